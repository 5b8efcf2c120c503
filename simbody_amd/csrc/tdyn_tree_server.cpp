// Serving slots for requests of different trees: tdyn_tree_server_*.
//
// Each slot holds one request of any tree the tree sweeps run (any size,
// topology and mobilizer mix).  What a request needs is taken from two pools
// sized from the HBM budget when the server is created, by first fit, at its
// admission, and given back when it is collected:
//   - the body pool: nb + 1 consecutive body entries (the tree's body tables
//     and model constants; the first entry holds Ground's applied-force row);
//   - the state heap: the slot's records (the tree sweeps' hand-over, padded
//     3-dof bodies) followed by its 11 integrator vectors of nq + nu doubles.
// So a request's byte cost is its own tree's, and requests are admitted in
// arrival order while a slot is free and both pools have room for the next
// one.  The slots advance together in one launch of the tree sweeps'
// Runge-Kutta-Merson kernel (a lane per slot; empty and finished slots are
// skipped), each with its own step control.
#include "tdyn_host.hpp"

#include <algorithm>
#include <cmath>
#include <map>

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

// First-fit allocator of [0, cap) in units; free ranges coalesce.
struct Arena {
    std::map<long long, long long> free_;   // start -> length
    long long cap = 0, used = 0;
    void init(long long c) { cap = c; used = 0; free_.clear(); if (c > 0) free_[0] = c; }
    long long take(long long n) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const long long at = it->first, rest = it->second - n;
            free_.erase(it);
            if (rest) free_[at + n] = rest;
            used += n;
            return at;
        }
        return -1;
    }
    void give(long long at, long long n) {
        used -= n;
        auto nx = free_.lower_bound(at);
        if (nx != free_.end() && at + n == nx->first) { n += nx->second; nx = free_.erase(nx); }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == at) { pv->second += n; return; }
        }
        free_[at] = n;
    }
};

constexpr long long kAlign = 8;   // heap chunks start on 64-byte boundaries
constexpr size_t kBodyBytes = sizeof(int4) * 2 + sizeof(int) + sizeof(double) * (PRM_STRIDE + 6);
constexpr size_t kSlotBytes = sizeof(int4) * 2 + sizeof(double) * 4 + sizeof(int) * (1 + 4 + 4) +
                              sizeof(long long) * 2 + sizeof(double) * 3;

struct Held { long long body, nbody, heap, nheap; int nq, nu; };

// What one tree needs: its tables in the kernels' layout (local numbers).
struct TreeTables {
    int nb = 0, nq = 0, nu = 0, nquat = 0, rec = 0;
    std::vector<int> body_a, body_b, child;   // 4 nb, 4 nb (child begin local), nb
    std::vector<double> prm;                  // nb PRM_STRIDE
};
void tree_tables(const tdyn_tree_desc& d, TreeTables& t) {
    const int nb = d.nb;
    t.nb = nb;
    std::vector<std::vector<int>> ch(nb + 1);
    for (int b = 1; b <= nb; ++b) ch[d.parent[b-1]].push_back(b);
    int q = 0, u = 0, rec = 0;
    t.body_a.clear(); t.body_b.clear(); t.child.clear(); t.prm.assign((size_t)nb * PRM_STRIDE, 0.0);
    for (int b = 1; b <= nb; ++b) {
        const int ty = d.mobilizer[b-1];
        const int dty = ty == TDYN_BUSHING ? TDYN_FREE : ty;     // a Bushing runs as a flagged Free body
        const int4 ba = make_int4(d.parent[b-1], dty, q, u);
        const int4 bb = make_int4((int)t.child.size(), (int)ch[b].size(), rec, 0);
        t.body_a.insert(t.body_a.end(), {ba.x, ba.y, ba.z, ba.w});
        t.body_b.insert(t.body_b.end(), {bb.x, bb.y, bb.z, bb.w});
        for (int c : ch[b]) t.child.push_back(c);
        // the tree sweeps of different trees carry non-Free bodies in padded 3-dof records
        rec += rec_size(dty != TDYN_FREE ? 3 : 6);
        q += mob_nq(ty); u += mob_dof(ty);
        if ((ty == TDYN_BALL || ty == TDYN_FREE) && !d.use_euler_angles) ++t.nquat;
        pack_body_params(d, b, t.prm.data() + (size_t)(b-1) * PRM_STRIDE);
    }
    t.child.resize(nb, 0);
    t.nq = q; t.nu = u; t.rec = rec + 24;
}

} // namespace

struct tdyn_tree_server {
    int slots = 0;
    size_t budget = 0, fixed_bytes = 0;
    RkmCfg cfg{};
    // device: per slot, per body, heap
    DBuf<int4> tree_a, tree_b, body_a, body_b;
    DBuf<int> tree_of, child, status, stats, zero_off, b_off, slot_dev, outi_dev;
    DBuf<double> gravity, prm, fapp, heap, t, h, t_end, out_dev, tout_dev;
    DBuf<long long> pool, ioff_dev, doff_dev, qoff_dev;
    DBuf<int> ints_dev; DBuf<double> dbl_dev;
    Pinned<int> ints_h, slot_h, status_h, outi_h;
    Pinned<double> dbl_h, out_h, tout_h;
    Pinned<long long> ioff_h, doff_h, qoff_h;
    Arena bodies, heapA;
    std::vector<int> free_list;
    std::vector<Held> held;       // per slot; body < 0 when empty
    RkmArgs rk() const {
        RkmArgs a;
        std::memset(&a, 0, sizeof(a));
        a.T.tree_a = tree_a.p; a.T.tree_b = tree_b.p; a.T.body_a = body_a.p; a.T.body_b = body_b.p;
        a.T.child_idx = child.p; a.T.prm = prm.p; a.T.gravity = gravity.p;
        a.tree_of = tree_of.p; a.n_inst = slots; a.L = slots;
        a.scratch = heap.p; a.vec = heap.p; a.pool = pool.p;
        a.off = {zero_off.p, zero_off.p, b_off.p};
        a.fapp = fapp.p; a.t = t.p; a.h = h.p; a.t_end = t_end.p; a.status = status.p; a.stats = stats.p;
        a.cfg = cfg;
        return a;
    }
};

extern "C" {

int tdyn_tree_server_create(const tdyn_integrator_cfg* cfg, size_t hbm_budget_bytes, int max_slots,
                            tdyn_tree_server** out) {
    if (!out || !cfg) return fail(TDYN_E_ARG, "tdyn_tree_server_create: null argument");
    *out = nullptr;
    RkmCfg c;
    int prec = 0;
    { const int rc = make_rkm_cfg(cfg, c, prec); if (rc) return rc; }
    if (prec != 0) return fail(TDYN_E_ARG, "tdyn_tree_server_create: FP64 only (precision 0)");
    const int S = max_slots > 0 ? max_slots : 65536;
    const size_t fixed = (size_t)S * kSlotBytes;
    if (hbm_budget_bytes <= fixed + (1u << 20))
        return fail(TDYN_E_NOMEM, "tdyn_tree_server_create: the HBM budget does not cover the slots' own state (" +
                                  std::to_string(fixed) + " bytes) and a pool");
    // a quarter of the pools for the body tables, the rest for records and state
    const size_t rest = hbm_budget_bytes - fixed;
    const long long nbody = (long long)(rest / 4 / kBodyBytes);
    const long long nheap = (long long)((rest - (size_t)nbody * kBodyBytes) / sizeof(double)) / kAlign * kAlign;
    if (nbody > INT32_MAX / PRM_STRIDE) return fail(TDYN_E_SIZE, "tdyn_tree_server_create: budget too large");
    auto* sv = new tdyn_tree_server;
    sv->slots = S; sv->budget = hbm_budget_bytes; sv->fixed_bytes = fixed; sv->cfg = c;
    std::vector<int> iota(S);
    for (int i = 0; i < S; ++i) iota[i] = i;
    std::vector<int> two(S, 2);
    hipError_t e;
    if ((e = sv->tree_a.alloc(S)) != hipSuccess || (e = sv->tree_b.alloc(S)) != hipSuccess ||
        (e = sv->gravity.alloc(4 * (size_t)S)) != hipSuccess || (e = sv->tree_of.upload(iota)) != hipSuccess ||
        (e = sv->zero_off.alloc(S)) != hipSuccess || (e = sv->b_off.alloc(S)) != hipSuccess ||
        (e = sv->pool.alloc(2 * (size_t)S)) != hipSuccess || (e = sv->t.alloc(S)) != hipSuccess ||
        (e = sv->h.alloc(S)) != hipSuccess || (e = sv->t_end.alloc(S)) != hipSuccess ||
        (e = sv->status.upload(two)) != hipSuccess || (e = sv->stats.alloc(4 * (size_t)S)) != hipSuccess ||
        (e = sv->body_a.alloc((size_t)nbody)) != hipSuccess || (e = sv->body_b.alloc((size_t)nbody)) != hipSuccess ||
        (e = sv->child.alloc((size_t)nbody)) != hipSuccess ||
        (e = sv->prm.alloc((size_t)nbody * PRM_STRIDE)) != hipSuccess ||
        (e = sv->fapp.alloc((size_t)nbody * 6)) != hipSuccess || (e = sv->heap.alloc((size_t)nheap)) != hipSuccess ||
        (e = hipMemset(sv->zero_off.p, 0, (size_t)S * sizeof(int))) != hipSuccess) {
        delete sv;
        return hipfail(e, "tdyn_tree_server_create: device allocation");
    }
    sv->bodies.init(nbody);
    sv->heapA.init(nheap);
    sv->free_list.resize(S);
    for (int i = 0; i < S; ++i) sv->free_list[i] = S - 1 - i;     // pop_back gives slot 0 first
    sv->held.assign(S, Held{-1, 0, -1, 0, 0, 0});
    *out = sv;
    return TDYN_OK;
}

int tdyn_tree_server_destroy(tdyn_tree_server* sv) { delete sv; return TDYN_OK; }

int tdyn_tree_server_info(const tdyn_tree_server* sv, int64_t* info) {
    if (!sv || !info) return fail(TDYN_E_ARG, "tdyn_tree_server_info: null argument");
    info[0] = sv->slots;
    info[1] = (int64_t)sv->free_list.size();
    info[2] = sv->slots - (int64_t)sv->free_list.size();
    info[3] = (int64_t)sv->budget;
    info[4] = (int64_t)(sv->fixed_bytes + (size_t)sv->bodies.used * kBodyBytes + (size_t)sv->heapA.used * sizeof(double));
    info[5] = sv->bodies.cap;
    info[6] = sv->heapA.cap;
    return TDYN_OK;
}

int tdyn_tree_server_request_bytes(const tdyn_tree_desc* tree, int64_t* bytes) {
    if (!tree || !bytes) return fail(TDYN_E_ARG, "tdyn_tree_server_request_bytes: null argument");
    { const int rc = validate_tree(*tree, 0); if (rc) return rc; }
    TreeTables t;
    tree_tables(*tree, t);
    const long long heapn = ((long long)t.rec + 11LL * (t.nq + t.nu) + kAlign - 1) / kAlign * kAlign;
    *bytes = (int64_t)((size_t)(t.nb + 1) * kBodyBytes + (size_t)heapn * sizeof(double) + kSlotBytes);
    return TDYN_OK;
}

int tdyn_tree_server_admit(tdyn_tree_server* sv, int n, const tdyn_tree_desc* trees, const double* q,
                           const double* u, const double* t_end, const double* fapp, int32_t* slots,
                           int* n_admitted, hipStream_t s) {
    if (!sv || n < 0 || (n > 0 && (!trees || !q || !u || !t_end || !slots)) || !n_admitted)
        return fail(TDYN_E_ARG, "tdyn_tree_server_admit: null argument");
    *n_admitted = 0;
    for (int i = 0; i < n; ++i) {
        const int rc = validate_tree(trees[i], i);
        if (rc) return rc;
        if (!std::isfinite(t_end[i]) || !(t_end[i] > 0.0))
            return fail(TDYN_E_ARG, "tdyn_tree_server_admit: t_end must be positive and finite");
    }
    // place requests in order while a slot is free and both pools have room
    std::vector<TreeTables> tabs;
    std::vector<Held> place;
    size_t nints = 0, ndbl = 0;
    long long qo = 0, uo = 0, fo = 0;
    std::vector<long long> qoff, uoff, foff;
    for (int i = 0; i < n && (int)place.size() < (int)sv->free_list.size(); ++i) {
        TreeTables t;
        tree_tables(trees[i], t);
        const long long heapn = ((long long)t.rec + 11LL * (t.nq + t.nu) + kAlign - 1) / kAlign * kAlign;
        if (t.nb + 1 > sv->bodies.cap || heapn > sv->heapA.cap) {
            for (const Held& p : place) { sv->bodies.give(p.body, p.nbody); sv->heapA.give(p.heap, p.nheap); }
            return fail(TDYN_E_NOMEM, "tdyn_tree_server_admit: request " + std::to_string(i) +
                                      " needs more than the server's whole pools");
        }
        const long long b0 = sv->bodies.take(t.nb + 1);
        if (b0 < 0) break;
        const long long h0 = sv->heapA.take(heapn);
        if (h0 < 0) { sv->bodies.give(b0, t.nb + 1); break; }
        place.push_back(Held{b0, t.nb + 1LL, h0, heapn, t.nq, t.nu});
        nints += 8 + 9 * (size_t)t.nb;
        ndbl += (size_t)t.nb * PRM_STRIDE + 3 + t.nq + t.nu + 1 + 6 * (size_t)(t.nb + 1) + 2;
        qoff.push_back(qo); uoff.push_back(uo); foff.push_back(fo);
        qo += t.nq; uo += t.nu; fo += 6 * (t.nb + 1);
        tabs.push_back(std::move(t));
    }
    const int k = (int)place.size();
    if (k == 0) return TDYN_OK;
    auto undo = [&]() { for (const Held& p : place) { sv->bodies.give(p.body, p.nbody); sv->heapA.give(p.heap, p.nheap); } };
    // stage the records on the host (pinned), then one admission launch
    if (sv->ints_h.reserve(nints) != hipSuccess || sv->dbl_h.reserve(ndbl) != hipSuccess ||
        sv->ioff_h.reserve(k) != hipSuccess || sv->doff_h.reserve(k) != hipSuccess ||
        sv->slot_h.reserve(k) != hipSuccess) { undo(); return fail(TDYN_E_HIP, "tdyn_tree_server_admit: pinned staging"); }
    hipError_t e = hipSuccess;
    if (sv->ints_dev.n < nints) e = sv->ints_dev.alloc(std::max<size_t>(nints, 4096));
    if (e == hipSuccess && sv->dbl_dev.n < ndbl) e = sv->dbl_dev.alloc(std::max<size_t>(ndbl, 4096));
    if (e == hipSuccess && sv->ioff_dev.n < (size_t)k) e = sv->ioff_dev.alloc(std::max(k, 256));
    if (e == hipSuccess && sv->doff_dev.n < (size_t)k) e = sv->doff_dev.alloc(std::max(k, 256));
    if (e == hipSuccess && sv->slot_dev.n < (size_t)k) e = sv->slot_dev.alloc(std::max(k, 256));
    if (e != hipSuccess) { undo(); return hipfail(e, "tdyn_tree_server_admit: device staging"); }
    size_t io = 0, dof = 0;
    for (int i = 0; i < k; ++i) {
        const TreeTables& t = tabs[i];
        const Held& p = place[i];
        const int body_off = (int)p.body + 1;
        sv->ioff_h.p[i] = (long long)io; sv->doff_h.p[i] = (long long)dof;
        int* I = sv->ints_h.p + io;
        I[0] = t.nb; I[1] = t.nq; I[2] = t.nu; I[3] = body_off; I[4] = (int)p.body; I[5] = t.nquat; I[6] = t.rec; I[7] = 0;
        std::copy(t.body_a.begin(), t.body_a.end(), I + 8);
        int* bb = I + 8 + 4 * t.nb;
        for (int b = 0; b < t.nb; ++b) {
            bb[4*b] = t.body_b[4*b] + body_off; bb[4*b+1] = t.body_b[4*b+1];
            bb[4*b+2] = t.body_b[4*b+2]; bb[4*b+3] = t.body_b[4*b+3];
        }
        std::copy(t.child.begin(), t.child.end(), bb + 4 * t.nb);
        io += 8 + 9 * (size_t)t.nb;
        double* D = sv->dbl_h.p + dof;
        std::copy(t.prm.begin(), t.prm.end(), D);
        D += t.prm.size();
        D[0] = trees[i].gravity[0]; D[1] = trees[i].gravity[1]; D[2] = trees[i].gravity[2];
        D += 3;
        std::copy(q + qoff[i], q + qoff[i] + t.nq, D); D += t.nq;
        std::copy(u + uoff[i], u + uoff[i] + t.nu, D); D += t.nu;
        *D++ = t_end[i];
        for (int j = 0; j < 6 * (t.nb + 1); ++j) D[j] = fapp ? fapp[foff[i] + j] : 0.0;
        D += 6 * (t.nb + 1);
        D[0] = (double)(p.heap + t.rec); D[1] = (double)p.heap;      // vectors after the records
        dof += (size_t)t.nb * PRM_STRIDE + 3 + t.nq + t.nu + 1 + 6 * (size_t)(t.nb + 1) + 2;
        const int slot = sv->free_list.back();
        sv->free_list.pop_back();
        sv->held[slot] = p;
        sv->slot_h.p[i] = slot;
        slots[i] = slot;
    }
    HIPCHK(hipMemcpyAsync(sv->ints_dev.p, sv->ints_h.p, nints * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(sv->dbl_dev.p, sv->dbl_h.p, ndbl * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(sv->ioff_dev.p, sv->ioff_h.p, k * sizeof(long long), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(sv->doff_dev.p, sv->doff_h.p, k * sizeof(long long), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(sv->slot_dev.p, sv->slot_h.p, k * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(launch_tsrv_admit(sv->rk(), sv->slot_dev.p, k, sv->ioff_dev.p, sv->ints_dev.p, sv->doff_dev.p,
                             sv->dbl_dev.p, s));
    // the pinned staging is reused by the next call: wait for the copies
    HIPCHK(hipStreamSynchronize(s));
    *n_admitted = k;
    return TDYN_OK;
}

int tdyn_tree_server_advance(tdyn_tree_server* sv, int max_steps, hipStream_t s) {
    if (!sv) return fail(TDYN_E_ARG, "tdyn_tree_server_advance: null server");
    if (max_steps <= 0) return fail(TDYN_E_ARG, "max_steps must be positive");
    if (sv->free_list.size() == (size_t)sv->slots) return TDYN_OK;
    RkmArgs a = sv->rk();
    a.cfg.max_steps = max_steps;
    HIPCHK(launch_rkm(a, s));
    return TDYN_OK;
}

int tdyn_tree_server_collect(tdyn_tree_server* sv, int max_out, int64_t y_cap, int32_t* slots, double* y,
                             int64_t* y_off, double* t, int32_t* status, int32_t* stats, int* n_out,
                             hipStream_t s) {
    if (!sv || !n_out || (max_out > 0 && (!slots || !y || !y_off))) return fail(TDYN_E_ARG, "tdyn_tree_server_collect: null argument");
    *n_out = 0;
    if (max_out <= 0 || sv->free_list.size() == (size_t)sv->slots) return TDYN_OK;
    const int S = sv->slots;
    HIPCHK(sv->status_h.reserve(S));
    HIPCHK(hipMemcpyAsync(sv->status_h.p, sv->status.p, (size_t)S * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<int> done;
    long long yo = 0;
    for (int i = 0; i < S && (int)done.size() < max_out; ++i) {
        const Held& p = sv->held[i];
        if (p.body < 0 || sv->status_h.p[i] == 0) continue;
        if (yo + p.nq + p.nu > y_cap) break;
        y_off[done.size()] = yo;
        yo += p.nq + p.nu;
        done.push_back(i);
    }
    const int k = (int)done.size();
    if (k == 0) return TDYN_OK;
    HIPCHK(sv->slot_h.reserve(k)); HIPCHK(sv->qoff_h.reserve(k)); HIPCHK(sv->out_h.reserve((size_t)yo));
    HIPCHK(sv->tout_h.reserve(k)); HIPCHK(sv->outi_h.reserve(5 * (size_t)k));
    if (sv->slot_dev.n < (size_t)k) HIPCHK(sv->slot_dev.alloc(std::max(k, 256)));
    if (sv->qoff_dev.n < (size_t)k) HIPCHK(sv->qoff_dev.alloc(std::max(k, 256)));
    if (sv->out_dev.n < (size_t)yo) HIPCHK(sv->out_dev.alloc(std::max<size_t>(yo, 4096)));
    if (sv->tout_dev.n < (size_t)k) HIPCHK(sv->tout_dev.alloc(std::max(k, 256)));
    if (sv->outi_dev.n < 5 * (size_t)k) HIPCHK(sv->outi_dev.alloc(5 * (size_t)std::max(k, 256)));
    for (int i = 0; i < k; ++i) { sv->slot_h.p[i] = done[i]; sv->qoff_h.p[i] = y_off[i]; }
    HIPCHK(hipMemcpyAsync(sv->slot_dev.p, sv->slot_h.p, k * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(sv->qoff_dev.p, sv->qoff_h.p, k * sizeof(long long), hipMemcpyHostToDevice, s));
    HIPCHK(launch_tsrv_collect(sv->rk(), sv->slot_dev.p, k, sv->qoff_dev.p, sv->out_dev.p, sv->tout_dev.p,
                               sv->outi_dev.p, s));
    HIPCHK(hipMemcpyAsync(sv->out_h.p, sv->out_dev.p, (size_t)yo * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(sv->tout_h.p, sv->tout_dev.p, k * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(sv->outi_h.p, sv->outi_dev.p, 5 * (size_t)k * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::copy(sv->out_h.p, sv->out_h.p + yo, y);
    for (int i = 0; i < k; ++i) {
        slots[i] = done[i];
        if (t) t[i] = sv->tout_h.p[i];
        if (status) status[i] = sv->outi_h.p[5*i];
        if (stats) std::copy(sv->outi_h.p + 5*i + 1, sv->outi_h.p + 5*i + 5, stats + 4 * (size_t)i);
        Held& p = sv->held[done[i]];
        sv->bodies.give(p.body, p.nbody);
        sv->heapA.give(p.heap, p.nheap);
        p = Held{-1, 0, -1, 0, 0, 0};
        sv->free_list.push_back(done[i]);
    }
    *n_out = k;
    return TDYN_OK;
}

} // extern "C"
