// Chain-evaluator rollout kernels: one model per lane, mixed precision.
#include "tdyn_rkm_chain.hpp"

namespace tdyn {

hipError_t launch_chain_mixed_lane(const RkmArgs& a, hipStream_t s, bool f0_only) {
    // the derivative alone (f0_only) is the general form's bit for bit, so
    // identity frames need no separate build there
    if (a.idf && !f0_only) return launch_chain_t<chain::IdF<chain::PrmL>, float>(a, s, false);
    return launch_chain_t<chain::PrmL, float>(a, s, f0_only);
}

} // namespace tdyn
