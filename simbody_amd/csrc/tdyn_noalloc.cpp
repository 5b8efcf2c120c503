// A rollout that allocates nothing: tdyn_rollout_noalloc.
//
// tdyn_rollout allocates the integrator state of a batch (state vectors,
// step sizes, status) at its first begin, and the batch's per-lane records
// at tdyn_batch_create.  Callers that cannot allocate device memory at the
// time of the call (code under graph capture, tight serving loops, harnesses
// that forbid it) use this variant instead: the records and the whole
// integrator state live in one file-scope device array, `scratch`, of
// TDYN_NOALLOC_SCRATCH_BYTES, one copy per GPU, held until the process ends.
// The model tables of the batch are read as they are (they were uploaded by
// tdyn_batch_create); q, u, fapp and stats are the caller's.  Every call on a
// GPU shares that GPU's array, so calls that overlap on one GPU (other host
// threads, other streams) race; the state of the previous call is not
// relied on (every row a call reads is written by that call first).
#include "tdyn_host.hpp"

#include <mutex>

using namespace tdyn;
using namespace tdyn::host;

namespace {

constexpr size_t kScratchBytes = TDYN_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The device address of this GPU's copy of `scratch` (looked up once per
// device, with that device current).
int scratch_ptr(unsigned char** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(TDYN_E_ARG, "tdyn_rollout_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hipfail(e, "tdyn_rollout_noalloc: hipGetSymbolAddress(scratch)");
        if (!p) return fail(TDYN_E_HIP, "tdyn_rollout_noalloc: no device address for scratch");
        g_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_ptr[dev]);
    return TDYN_OK;
}

__global__ void k_fill(double* p, double v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// counts[0] = lanes still running (status 0), counts[1] = lanes failed
// (status < 0).  One 256-thread block.
__global__ void __launch_bounds__(256) k_status_counts(const int* status, int n, int* counts) {
    __shared__ int run[256], bad[256];
    int r = 0, f = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) { r += status[i] == 0; f += status[i] < 0; }
    run[threadIdx.x] = r; bad[threadIdx.x] = f;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { run[threadIdx.x] += run[threadIdx.x + w]; bad[threadIdx.x] += bad[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { counts[0] = run[0]; counts[1] = bad[0]; }
}

size_t up256(size_t x) { return (x + 255) / 256 * 256; }

} // namespace

extern "C" int tdyn_rollout_noalloc(tdyn_batch* b, double* q, double* u, const double* fapp, double t0,
                                    double t_end, const tdyn_integrator_cfg* cfg, int32_t* stats, hipStream_t s) {
    if (!b || !q || !u || !cfg) return fail(TDYN_E_ARG, "tdyn_rollout_noalloc: batch, q, u and cfg are required");
    if (!(t_end > t0)) return fail(TDYN_E_ARG, "tdyn_rollout_noalloc: final time must exceed t0");
    RkmCfg c;
    int prec = 0;
    { const int rc = make_rkm_cfg(cfg, c, prec); if (rc) return rc; }
    if (prec != 0) return fail(TDYN_E_ARG, "tdyn_rollout_noalloc: FP64 only (precision 0)");
    if (b->evaluator == 3) return fail(TDYN_E_ARG, "tdyn_rollout_noalloc: runs the chain evaluator or the tree "
                                                   "sweeps (evaluator 0, 1 or 2), not the level-parallel evaluator");
    if (b->motion) return fail(TDYN_E_ARG, "tdyn_rollout_noalloc: prescribed motion runs on tdyn_rollout_begin only");
    if (b->any_cons) return fail(TDYN_E_ARG, "tdyn_rollout_noalloc: constraints run on tdyn_rollout_begin only");
    if (b->any_wit) return fail(TDYN_E_ARG, "tdyn_rollout_noalloc: event witnesses run on tdyn_rollout_begin only");
    const int L = b->n_inst;
    const int ny = b->max_nq + b->max_nu;
    const int rec = b->max_rec + 24;
    // layout of the array: records, state vectors, u scales, t, h, t_end,
    // status, statistics, two status counts
    size_t off = 0;
    const size_t o_rec = off;   off = up256(off + sizeof(double) * (size_t)rec * L);
    const size_t o_vec = off;   off = up256(off + sizeof(double) * ((size_t)11 * ny + 8) * L);
    const size_t o_us = off;    off = up256(off + sizeof(double) * (size_t)std::max(1, b->max_nu) * L);
    const size_t o_t = off;     off = up256(off + sizeof(double) * (size_t)L);
    const size_t o_h = off;     off = up256(off + sizeof(double) * (size_t)L);
    const size_t o_te = off;    off = up256(off + sizeof(double) * (size_t)L);
    const size_t o_st = off;    off = up256(off + sizeof(int) * (size_t)L);
    const size_t o_sx = off;    off = up256(off + sizeof(int) * 4 * (size_t)L);
    const size_t o_cnt = off;   off = up256(off + sizeof(int) * 2);
    if (off > kScratchBytes)
        return fail(TDYN_E_SIZE, "tdyn_rollout_noalloc: the batch needs " + std::to_string(off) +
                                 " bytes of scratch, more than the " + std::to_string(kScratchBytes) +
                                 " bytes of TDYN_NOALLOC_SCRATCH_BYTES");
    unsigned char* base = nullptr;
    { const int rc = scratch_ptr(&base); if (rc) return rc; }

    RkmArgs a = b->rk();
    a.cfg = c;
    a.precision = 0;
    a.scratch32 = nullptr;
    a.presc = b->known_p; a.udot_known = b->known_u;
    a.motion = nullptr;
    a.cons_ws = nullptr;
    a.chain = (!b->known_p && (b->evaluator == 2 || (b->evaluator == 0 && b->is_chain))) ? 1 : 0;
    a.idf = (a.chain ? b->idf : b->idf_all) ? 1 : 0;
    a.ext = (b->any_rev || b->any_euler || b->any_bendstretch || b->any_bushing || b->known_p) ? 1 : 0;
    a.vec_aos = 0; a.grp = 0;
    a.plan = nullptr; a.plan_idx = nullptr; a.plan_waves = 0; a.plan_ctr = nullptr;
    a.rec_stride = rec;
    a.scratch = reinterpret_cast<double*>(base + o_rec);
    a.vec = reinterpret_cast<double*>(base + o_vec);
    a.ny_max = ny;
    a.uscale = reinterpret_cast<double*>(base + o_us);
    a.t = reinterpret_cast<double*>(base + o_t);
    a.h = reinterpret_cast<double*>(base + o_h);
    double* te = reinterpret_cast<double*>(base + o_te);
    a.t_end = te;
    a.status = reinterpret_cast<int*>(base + o_st);
    a.stats = reinterpret_cast<int*>(base + o_sx);
    a.fapp = fapp;
    int* counts = reinterpret_cast<int*>(base + o_cnt);

    k_fill<<<dim3((L + 255) / 256), dim3(256), 0, s>>>(te, t_end, L);
    HIPCHK(hipGetLastError());
    HIPCHK(launch_pack_state(a, q, u, 1, nullptr, nullptr, s));
    HIPCHK(launch_rkm_init(a, t0, s));
    int h_counts[2] = {0, 0};
    for (int round = 0; round < 1000000; ++round) {
        RkmArgs x = a;
        x.cfg.max_steps = 1 << 20;
        HIPCHK(launch_rkm(x, s));
        k_status_counts<<<dim3(1), dim3(256), 0, s>>>(a.status, L, counts);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h_counts, counts, sizeof h_counts, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (h_counts[0] == 0) break;
    }
    if (h_counts[0] != 0) return fail(TDYN_E_STEP_FAILED, "tdyn_rollout_noalloc: rollout did not finish");
    HIPCHK(launch_pack_state(a, nullptr, nullptr, 0, q, u, s));
    if (stats) {
        if (b->permuted) HIPCHK(launch_lane_scatter(b->inst.p, L, nullptr, nullptr, nullptr, nullptr, a.stats, stats, s));
        else HIPCHK(hipMemcpyAsync(stats, a.stats, 4 * (size_t)L * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    return h_counts[1] ? fail(TDYN_E_STEP_FAILED, "Integrator::StepFailed for at least one instance") : TDYN_OK;
}
