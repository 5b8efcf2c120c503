// Chain-evaluator rollout kernels: one model per batch, FP64 arithmetic with
// single-precision hand-over rows (integrator precision 2: G; precision 3: G
// and the trig cache).
#include "tdyn_rkm_chain.hpp"

namespace tdyn {

hipError_t launch_chain_split_batch(const RkmArgs& a, hipStream_t s, bool f0_only) {
    if (a.precision == 2) {
        if (a.idf && !f0_only) return launch_chain_t<chain::IdF<cdptr>, double, float, double>(a, s, false);
        return launch_chain_t<cdptr, double, float, double>(a, s, f0_only);
    }
    if (a.idf && !f0_only) return launch_chain_t<chain::IdF<cdptr>, double, float, float>(a, s, false);
    return launch_chain_t<cdptr, double, float, float>(a, s, f0_only);
}

} // namespace tdyn
