// Serving slots over the batch machinery: one tdyn_batch of `slots`
// instances of the server's chain topology provides the tables, scratch and
// integrator state; each slot's own model constants live in per-lane rows
// (RkmArgs::prm_lane) that the chain evaluator reads instead of the shared
// table.  Admission and retirement move small pinned host records.
#include "tdyn_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

using namespace tdyn;
using namespace tdyn::host;

namespace {
template <class T> struct Pinned {
    T* p = nullptr; size_t n = 0;
    hipError_t reserve(size_t count) {
        if (count <= n) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr; n = 0;
        hipError_t e = hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault);
        if (e == hipSuccess) n = count;
        return e;
    }
    ~Pinned() { if (p) (void)hipHostFree(p); }
};
} // namespace

struct tdyn_server {
    tdyn_batch* B = nullptr;
    int nb = 0, nq = 0, nu = 0, slots = 0;
    size_t bytes_per_slot = 0;
    std::vector<int32_t> parent, mobilizer;
    DBuf<double> prm_lane, rec_dev, out_dev;
    DBuf<int> slot_dev, outi_dev;
    Pinned<double> rec_host, out_host;
    Pinned<int> slot_host, outi_host, status_host;
    std::vector<int> free_list;
    std::vector<char> occupied;
    bool idf = false;      // every model admitted so far has identity joint frames
    ~tdyn_server() { if (B) tdyn_batch_destroy(B); }
    int rec_stride() const { return nb * PRM_STRIDE + 3 + nq + nu + 1 + 6 * (nb + 1); }
    RkmArgs rk() const {
        RkmArgs a = B->rk();
        a.chain = 1;
        a.prm_lane = prm_lane.p;
        a.idf = idf ? 1 : 0;
        return a;
    }
};

extern "C" {

int tdyn_server_create(const tdyn_tree_desc* topology, const tdyn_integrator_cfg* cfg, size_t hbm_budget_bytes,
                       int max_slots, tdyn_server** out) {
    if (!out || !topology || !cfg) return fail(TDYN_E_ARG, "tdyn_server_create: null argument");
    *out = nullptr;
    { const int rc = validate_tree(*topology, 0); if (rc) return rc; }
    const int nb = topology->nb;
    for (int b = 1; b <= nb; ++b) {
        if (topology->parent[b-1] != b - 1)
            return fail(TDYN_E_ARG, "tdyn_server_create: the serving path runs serial-chain topologies (parent[b-1] == b-1)");
        if (mob_generic_only(topology->mobilizer[b-1]))
            return fail(TDYN_E_ARG, "tdyn_server_create: the serving path runs Pin, Slider, Universal, Ball and Free bodies");
        if (topology->mobilizer[b-1] == TDYN_FREE && b > 1)
            return fail(TDYN_E_ARG, "tdyn_server_create: the serving path runs a Free body as the first body only");
    }
    if (any_reversed(*topology))
        return fail(TDYN_E_ARG, "tdyn_server_create: the serving path runs forward mobilizers only");
    if (topology->use_euler_angles)
        return fail(TDYN_E_ARG, "tdyn_server_create: the serving path runs quaternion Ball and Free bodies only");
    RkmCfg c;
    int prec = 0;
    { const int rc = make_rkm_cfg(cfg, c, prec); if (rc) return rc; }
    if (prec > 1) return fail(TDYN_E_ARG, "tdyn_server_create: serving runs precision 0 (FP64) or 1 (mixed)");
    if (c.inf_norm || c.force_proj)
        return fail(TDYN_E_ARG, "tdyn_server_create: the chain slots run the RMS norm and the default projection "
                                "(the tree server takes the infinity norm and projection on every step)");
    int nq = 0, nu = 0, rec = 0;
    for (int b = 1; b <= nb; ++b) {
        const int ty = topology->mobilizer[b-1];
        nq += mob_nq(ty); nu += mob_dof(ty); rec += rec_size(mob_dof(ty));
    }
    const int ny = nq + nu;
    // device bytes per slot: state vectors, u scaling, scratch, model rows,
    // applied forces, t/h/t_end, status/stats, packed offsets, mixed-path rows
    const size_t per = 8 * ((size_t)11 * ny + 8 + std::max(1, nu) + (rec + 24) + nb * PRM_STRIDE + 3 + 6 * (nb + 1) + 3)
                       + 4 * (1 + 4 + 3) + (prec == 1 ? 4 * (size_t)(6 * nu + 4 * nb) : 0);
    size_t n = std::min<size_t>(max_slots > 0 ? (size_t)max_slots : (size_t)1 << 30, hbm_budget_bytes / per);
    n = n / 64 * 64;
    if (n == 0) return fail(TDYN_E_NOMEM, "tdyn_server_create: the HBM budget holds no 64-slot group (" +
                                          std::to_string(per) + " bytes per slot)");
    if (n > (size_t)INT32_MAX / 8) return fail(TDYN_E_SIZE, "tdyn_server_create: too many slots");
    auto* S = new tdyn_server;
    S->nb = nb; S->nq = nq; S->nu = nu; S->slots = (int)n; S->bytes_per_slot = per;
    S->parent.assign(topology->parent, topology->parent + nb);
    S->mobilizer.assign(topology->mobilizer, topology->mobilizer + nb);
    int rc = tdyn_batch_create(topology, 1, nullptr, (int)n, &S->B);
    if (rc) { delete S; return rc; }
    S->B->cfg = c;
    S->B->evaluator = 2;
    S->B->roll_evaluator = 2;
    S->B->precision = prec;
    if ((rc = ensure_rollout_buffers(S->B, prec))) { delete S; return rc; }
    S->B->rollout_ready = true;
    hipError_t e;
    if ((e = S->prm_lane.alloc((size_t)(nb * PRM_STRIDE + 3) * n)) != hipSuccess ||
        (e = S->B->fapp.alloc((size_t)S->B->total_b * 6)) != hipSuccess ||
        (e = hipMemset(S->B->fapp.p, 0, (size_t)S->B->total_b * 6 * sizeof(double))) != hipSuccess ||
        (e = hipMemset(S->B->stats.p, 0, 4 * n * sizeof(int))) != hipSuccess ||
        (e = hipMemsetD32((hipDeviceptr_t)S->B->status.p, 2, n)) != hipSuccess ||
        (e = S->status_host.reserve(n)) != hipSuccess) {
        delete S;
        return hipfail(e, "tdyn_server_create: device allocation");
    }
    S->idf = identity_frames(*topology);
    S->free_list.resize(n);
    for (size_t i = 0; i < n; ++i) S->free_list[i] = (int)(n - 1 - i);    // pop_back gives slot 0 first
    S->occupied.assign(n, 0);
    *out = S;
    return TDYN_OK;
}

int tdyn_server_destroy(tdyn_server* sv) { delete sv; return TDYN_OK; }

int tdyn_server_info(const tdyn_server* sv, int64_t* info) {
    if (!sv || !info) return fail(TDYN_E_ARG, "tdyn_server_info: null argument");
    info[0] = sv->slots; info[1] = (int64_t)sv->free_list.size();
    info[2] = sv->slots - (int64_t)sv->free_list.size(); info[3] = (int64_t)sv->bytes_per_slot;
    info[4] = sv->nq; info[5] = sv->nu;
    return TDYN_OK;
}

int tdyn_server_admit(tdyn_server* sv, int n, const tdyn_tree_desc* trees, const double* q, const double* u,
                      const double* t_end, const double* fapp, int32_t* slots, int* n_admitted, hipStream_t s) {
    if (!sv || n < 0 || (n > 0 && (!trees || !q || !u || !t_end || !slots)) || !n_admitted)
        return fail(TDYN_E_ARG, "tdyn_server_admit: null argument");
    *n_admitted = 0;
    const int nb = sv->nb, nq = sv->nq, nu = sv->nu;
    for (int i = 0; i < n; ++i) {
        const tdyn_tree_desc& d = trees[i];
        if (d.nb != nb || !d.parent || !d.mobilizer ||
            !std::equal(sv->parent.begin(), sv->parent.end(), d.parent) ||
            !std::equal(sv->mobilizer.begin(), sv->mobilizer.end(), d.mobilizer))
            return fail(TDYN_E_ARG, "tdyn_server_admit: request " + std::to_string(i) + " does not have the server's topology");
        // the slot kernels read neither reversal nor Euler-angle flags: refuse them, as tdyn_server_create does
        if (any_reversed(d))
            return fail(TDYN_E_ARG, "tdyn_server_admit: request " + std::to_string(i) +
                                    " has a reversed mobilizer; the serving path runs forward mobilizers only");
        if (d.use_euler_angles)
            return fail(TDYN_E_ARG, "tdyn_server_admit: request " + std::to_string(i) +
                                    " uses Euler angles; the serving path runs quaternion Ball and Free bodies only");
        const int rc = validate_tree(d, i);
        if (rc) return rc;
        if (!std::isfinite(t_end[i])) return fail(TDYN_E_ARG, "tdyn_server_admit: non-finite t_end");
    }
    // a model with general frames switches the server to the general kernel for good
    for (int i = 0; i < std::min<int>(n, (int)sv->free_list.size()); ++i)
        if (sv->idf && !identity_frames(trees[i])) sv->idf = false;
    const int k = std::min<int>(n, (int)sv->free_list.size());
    if (k == 0) return TDYN_OK;
    const int stride = sv->rec_stride();
    HIPCHK(sv->rec_host.reserve((size_t)k * stride));
    HIPCHK(sv->slot_host.reserve(k));
    if (sv->rec_dev.n < (size_t)k * stride) HIPCHK(sv->rec_dev.alloc((size_t)std::max(k, 64) * stride));
    if (sv->slot_dev.n < (size_t)k) HIPCHK(sv->slot_dev.alloc((size_t)std::max(k, 64)));
    for (int i = 0; i < k; ++i) {
        double* r = sv->rec_host.p + (size_t)i * stride;
        const tdyn_tree_desc& d = trees[i];
        for (int b = 1; b <= nb; ++b) pack_body_params(d, b, r + (size_t)(b-1) * PRM_STRIDE);
        double* x = r + (size_t)nb * PRM_STRIDE;
        x[0] = d.gravity[0]; x[1] = d.gravity[1]; x[2] = d.gravity[2];
        x += 3;
        for (int j = 0; j < nq; ++j) x[j] = q[(size_t)i * nq + j];
        x += nq;
        for (int j = 0; j < nu; ++j) x[j] = u[(size_t)i * nu + j];
        x += nu;
        x[0] = t_end[i];
        x += 1;
        for (int j = 0; j < 6 * (nb + 1); ++j) x[j] = fapp ? fapp[(size_t)i * 6 * (nb + 1) + j] : 0.0;
        const int slot = sv->free_list.back();
        sv->free_list.pop_back();
        sv->occupied[slot] = 1;
        sv->slot_host.p[i] = slot;
        slots[i] = slot;
    }
    HIPCHK(hipMemcpyAsync(sv->rec_dev.p, sv->rec_host.p, (size_t)k * stride * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(sv->slot_dev.p, sv->slot_host.p, (size_t)k * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(launch_srv_admit(sv->rk(), sv->rec_dev.p, sv->slot_dev.p, k, stride, 0.0, s));
    HIPCHK(hipStreamSynchronize(s));
    *n_admitted = k;
    return TDYN_OK;
}

int tdyn_server_advance(tdyn_server* sv, int max_steps, hipStream_t s) {
    if (!sv) return fail(TDYN_E_ARG, "tdyn_server_advance: null server");
    if (max_steps <= 0) return fail(TDYN_E_ARG, "max_steps must be positive");
    if (sv->free_list.size() == (size_t)sv->slots) return TDYN_OK;
    RkmArgs a = sv->rk();
    a.cfg.max_steps = max_steps;
    HIPCHK(launch_rkm(a, s));
    return TDYN_OK;
}

int tdyn_server_collect(tdyn_server* sv, int max_out, int32_t* slots, double* q, double* u, double* t,
                        int32_t* status, int32_t* stats, int* n_out, hipStream_t s) {
    if (!sv || !n_out || (max_out > 0 && !slots)) return fail(TDYN_E_ARG, "tdyn_server_collect: null argument");
    *n_out = 0;
    if (max_out <= 0 || sv->free_list.size() == (size_t)sv->slots) return TDYN_OK;
    const int N = sv->slots;
    HIPCHK(hipMemcpyAsync(sv->status_host.p, sv->B->status.p, (size_t)N * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<int> done;
    for (int i = 0; i < N && (int)done.size() < max_out; ++i)
        if (sv->occupied[i] && sv->status_host.p[i] != 0) done.push_back(i);
    const int k = (int)done.size();
    if (k == 0) return TDYN_OK;
    const int ny = sv->nq + sv->nu, ostride = ny + 1;
    HIPCHK(sv->slot_host.reserve(k));
    HIPCHK(sv->out_host.reserve((size_t)k * ostride));
    HIPCHK(sv->outi_host.reserve((size_t)k * 5));
    if (sv->slot_dev.n < (size_t)k) HIPCHK(sv->slot_dev.alloc((size_t)std::max(k, 64)));
    if (sv->out_dev.n < (size_t)k * ostride) HIPCHK(sv->out_dev.alloc((size_t)std::max(k, 64) * ostride));
    if (sv->outi_dev.n < (size_t)k * 5) HIPCHK(sv->outi_dev.alloc((size_t)std::max(k, 64) * 5));
    std::copy(done.begin(), done.end(), sv->slot_host.p);
    HIPCHK(hipMemcpyAsync(sv->slot_dev.p, sv->slot_host.p, (size_t)k * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(launch_srv_collect(sv->rk(), sv->slot_dev.p, k, sv->out_dev.p, ostride, sv->outi_dev.p, s));
    HIPCHK(hipMemcpyAsync(sv->out_host.p, sv->out_dev.p, (size_t)k * ostride * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(sv->outi_host.p, sv->outi_dev.p, (size_t)k * 5 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i < k; ++i) {
        const double* o = sv->out_host.p + (size_t)i * ostride;
        const int* oi = sv->outi_host.p + (size_t)i * 5;
        slots[i] = done[i];
        if (q) std::copy(o, o + sv->nq, q + (size_t)i * sv->nq);
        if (u) std::copy(o + sv->nq, o + ny, u + (size_t)i * sv->nu);
        if (t) t[i] = o[ny];
        if (status) status[i] = oi[0];
        if (stats) std::copy(oi + 1, oi + 5, stats + (size_t)i * 4);
        sv->occupied[done[i]] = 0;
        sv->free_list.push_back(done[i]);
    }
    *n_out = k;
    return TDYN_OK;
}

} // extern "C"
