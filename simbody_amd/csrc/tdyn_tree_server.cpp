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
#include <set>

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

// Best-fit allocator of [0, cap) in units: free ranges indexed by start (to
// coalesce neighbours on return) and by length (the smallest range that
// fits), so taking and giving back cost O(log ranges) however fragmented
// the pool is.  (A first-fit scan from the start measured 73 us per
// admission at 65,536 occupied slots and 17,000 free ranges.)
struct Arena {
    std::map<long long, long long> free_;          // start -> length
    std::set<std::pair<long long, long long>> by_len;   // (length, start)
    long long cap = 0, used = 0;
    void init(long long c) {
        cap = c; used = 0; free_.clear(); by_len.clear();
        if (c > 0) { free_[0] = c; by_len.insert({c, 0}); }
    }
    long long take(long long n) {
        auto f = by_len.lower_bound({n, -1});
        if (f == by_len.end()) return -1;
        const long long len = f->first, at = f->second, rest = len - n;
        by_len.erase(f);
        free_.erase(at);
        if (rest) { free_[at + n] = rest; by_len.insert({rest, at + n}); }
        used += n;
        return at;
    }
    void give(long long at, long long n) {
        used -= n;
        auto nx = free_.lower_bound(at);
        if (nx != free_.end() && at + n == nx->first) {
            by_len.erase({nx->second, nx->first});
            n += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == at) {
                by_len.erase({pv->second, pv->first});
                pv->second += n;
                by_len.insert({pv->second, pv->first});
                return;
            }
        }
        free_[at] = n;
        by_len.insert({n, at});
    }
};

constexpr long long kAlign = 8;   // heap chunks start on 64-byte boundaries
// per body-pool entry: body_a, body_b, bfs, bfs_b (int4), child (int), two
// level-list ints, prm and prm_bfs (PRM_STRIDE doubles each), applied force (6)
constexpr size_t kBodyBytes = sizeof(int4) * 4 + sizeof(int) * 3 + sizeof(double) * (2 * PRM_STRIDE + 6);
// per slot: tree_a, tree_b, gravity (4), t, h, t_end, tree_of, status, stats
// (4), b_off, pool offsets (2), plan_idx, the evaluator flag, and room for
// the plan of the level-parallel evaluator (one wavefront entry per slot)
constexpr size_t kSlotBytes = sizeof(int4) * 2 + sizeof(double) * 4 + sizeof(int) * (1 + 4 + 4) +
                              sizeof(long long) * 2 + sizeof(double) * 3 + sizeof(int) + 1 + sizeof(int4);
constexpr int kMaxQueues = 16;    // work queues of the level-parallel plan: 3 builds x 5 lane counts

struct Held {
    long long body, nbody, heap, nheap; int nq, nu;
    int grp, build, G;            // level-parallel: 1, its build (0 general, 1 identity frames, 2 extended), lanes
    long long cost;               // level steps of one derivative at G lanes
};

// What one tree needs: its tables in the kernels' layout (local numbers).
struct TreeTables {
    int nb = 0, nq = 0, nu = 0, nquat = 0, rec = 0;
    std::vector<int> body_a, body_b, child;   // 4 nb, 4 nb (child begin local), nb
    // the level-parallel evaluator's tables (when grp)
    int grp = 0, build = 0, G = 0;
    long long cost = 0;
    std::vector<int> bfs, bfs_b, starts;      // 4 nb, 4 nb, level starts (levels + 1)
    std::vector<int> cs, cl, at, order, pos, depth;   // working space, kept between requests
    long long body_entries() const {
        return std::max<long long>(nb + 1, (long long)tgroup::blocks(nb) * tgroup::BW);
    }
    long long heap_doubles() const {
        const long long r = grp ? (long long)tgroup::scratch_doubles(nb) : rec;
        return (r + 11LL * (nq + nu) + kAlign - 1) / kAlign * kAlign;
    }
};
// lanes: 1 the tree sweeps for every request; 0 the level-parallel
// evaluator where the tree allows it (no Free or Bushing body below body 1),
// its lanes from the tree's widest level (8 to 32); 4..64 that many lanes.
void tree_tables(const tdyn_tree_desc& d, TreeTables& t, int lanes) {
    const int nb = d.nb;
    t.nb = nb;
    // children of every body (Ground 0 included), in increasing order: the
    // range [cs[b], cs[b+1]) of cl
    std::vector<int>& cs = t.cs; std::vector<int>& cl = t.cl;
    cs.assign(nb + 2, 0); cl.resize(nb);
    for (int b = 1; b <= nb; ++b) ++cs[d.parent[b-1] + 1];
    for (int b = 0; b <= nb; ++b) cs[b + 1] += cs[b];
    t.at.assign(cs.begin(), cs.end() - 1);
    for (int b = 1; b <= nb; ++b) cl[t.at[d.parent[b-1]]++] = b;
    auto nch = [&](int b) { return cs[b + 1] - cs[b]; };
    int q = 0, u = 0, rec = 0;
    bool grp_ok = lanes != 1, ext = d.use_euler_angles != 0;
    t.body_a.clear(); t.body_b.clear(); t.child.clear();
    t.body_a.reserve(4 * (size_t)nb); t.body_b.reserve(4 * (size_t)nb); t.child.reserve(nb);
    for (int b = 1; b <= nb; ++b) {
        const int ty = d.mobilizer[b-1];
        const int dty = ty == TDYN_BUSHING ? TDYN_FREE : ty;     // a Bushing runs as a flagged Free body
        const int4 ba = make_int4(d.parent[b-1], dty, q, u);
        const int4 bb = make_int4((int)t.child.size(), nch(b), rec, 0);
        t.body_a.insert(t.body_a.end(), {ba.x, ba.y, ba.z, ba.w});
        t.body_b.insert(t.body_b.end(), {bb.x, bb.y, bb.z, bb.w});
        t.child.insert(t.child.end(), cl.begin() + cs[b], cl.begin() + cs[b + 1]);
        // the tree sweeps of different trees carry non-Free bodies in padded 3-dof records
        rec += rec_size(dty != TDYN_FREE ? 3 : 6);
        q += mob_nq(ty); u += mob_dof(ty);
        if ((ty == TDYN_BALL || ty == TDYN_FREE) && !d.use_euler_angles) ++t.nquat;
        // the level-parallel evaluator carries a 6-dof body as body 1 only
        if (b > 1 && (ty == TDYN_FREE || ty == TDYN_BUSHING)) grp_ok = false;
        ext = ext || ty == TDYN_BENDSTRETCH || ty == TDYN_BUSHING || (d.reversed && d.reversed[b-1]);
    }
    t.child.resize(nb, 0);
    t.nq = q; t.nu = u; t.rec = rec + 24;
    t.grp = grp_ok ? 1 : 0;
    t.bfs.clear(); t.bfs_b.clear(); t.starts.clear();
    if (!grp_ok) return;
    t.bfs.reserve(4 * (size_t)nb); t.bfs_b.reserve(4 * (size_t)nb);
    t.build = ext ? 2 : identity_frames(d) ? 1 : 0;
    // breadth-first order (roots, then each level's children in order), as
    // the batches' tables: body, parent's position, first child's position,
    // child count; mobilizer, qIndex, uIndex, body; the level starts
    std::vector<int>& order = t.order; std::vector<int>& pos = t.pos; std::vector<int>& depth = t.depth;
    order.reserve(nb);
    order.assign(cl.begin() + cs[0], cl.begin() + cs[1]);
    pos.assign(nb + 1, -1); depth.assign(nb + 1, 0);
    t.starts.push_back(0);
    for (size_t k = 0; k < order.size(); ++k) depth[order[k]] = 1;
    for (size_t k = 0; k < order.size(); ++k) {
        const int b = order[k];
        pos[b] = (int)k;
        if (k > 0 && depth[b] != depth[order[k - 1]]) t.starts.push_back((int)k);
        for (int j = cs[b]; j < cs[b + 1]; ++j) { depth[cl[j]] = depth[b] + 1; order.push_back(cl[j]); }
    }
    t.starts.push_back(nb);
    for (int k = 0; k < nb; ++k) {
        const int b = order[k], p = d.parent[b - 1];
        const int first = nch(b) == 0 ? 0 : pos[cl[cs[b]]];
        t.bfs.insert(t.bfs.end(), {b, p == 0 ? -1 : pos[p], first, nch(b)});
        t.bfs_b.insert(t.bfs_b.end(), {t.body_a[4*(b-1) + 1], t.body_a[4*(b-1) + 2], t.body_a[4*(b-1) + 3], b});
    }
    int maxw = 0;
    for (size_t k = 0; k + 1 < t.starts.size(); ++k) maxw = std::max(maxw, t.starts[k + 1] - t.starts[k]);
    int G = lanes;
    if (G <= 0) { G = 8; while (G < 32 && G < maxw) G *= 2; }
    t.G = G;
    t.cost = 0;
    for (size_t k = 0; k + 1 < t.starts.size(); ++k) t.cost += (t.starts[k + 1] - t.starts[k] + G - 1) / G;
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
    // level-parallel evaluator: per body entry bfs, bfs_b, two level-list
    // ints and the breadth-first constants; per slot the evaluator flag; the
    // plan (wavefronts, instances, queue counters)
    DBuf<int4> bfs, bfs_b, plan;
    DBuf<int> lvl, plan_idx, plan_ctr;
    DBuf<double> prm_bfs;
    DBuf<unsigned char> pool_grp;
    Pinned<int> ints_h, slot_h, status_h, outi_h, pidx_h;
    Pinned<double> dbl_h, out_h, tout_h;
    Pinned<long long> ioff_h, doff_h, qoff_h;
    Pinned<int4> plan_h;
    Arena bodies, heapA;
    std::vector<int> free_list;
    std::vector<Held> held;       // per slot; body < 0 when empty
    std::vector<TreeTables> tabs; // per admitted request of one admission call
    int lanes = 0;                // lanes setting for requests admitted from now on
    int resident = 0;             // wavefronts of the level-parallel evaluator the device keeps resident
    int n_grp = 0, n_sweep = 0;   // occupied slots on the level-parallel evaluator / on the tree sweeps
    bool plan_dirty = true;
    int plan_off[3] = {0, 0, 0}, plan_n[3] = {0, 0, 0}, plan_nctr = 0;   // per build: its wavefronts
    RkmArgs rk() const {
        RkmArgs a;
        std::memset(&a, 0, sizeof(a));
        a.T.tree_a = tree_a.p; a.T.tree_b = tree_b.p; a.T.body_a = body_a.p; a.T.body_b = body_b.p;
        a.T.child_idx = child.p; a.T.prm = prm.p; a.T.gravity = gravity.p;
        a.T.bfs = bfs.p; a.T.bfs_b = bfs_b.p; a.T.lvl = lvl.p; a.T.prm_bfs = prm_bfs.p;
        a.tree_of = tree_of.p; a.n_inst = slots; a.L = slots;
        a.scratch = heap.p; a.vec = heap.p; a.pool = pool.p; a.pool_grp = pool_grp.p;
        a.off = {zero_off.p, zero_off.p, b_off.p};
        a.fapp = fapp.p; a.t = t.p; a.h = h.p; a.t_end = t_end.p; a.status = status.p; a.stats = stats.p;
        a.plan_idx = plan_idx.p; a.plan_ctr = plan_ctr.p; a.plan_nctr = plan_nctr;
        a.cfg = cfg;
        return a;
    }
    // The level-parallel plan of the occupied slots: per build (one launch
    // each) one work queue per lane count, in decreasing order of cost, and
    // persistent wavefronts shared out over the queues by their lane-weighted
    // work (the device's resident wavefronts per launch).
    int build_plan(hipStream_t s);
};

int tdyn_tree_server::build_plan(hipStream_t s) {
    std::vector<int4> pl;
    std::vector<int> idx;
    int nctr = 0;
    for (int bd = 0; bd < 3; ++bd) {
        plan_off[bd] = (int)pl.size();
        std::map<int, std::vector<int>> byG;            // lanes -> slots
        for (int i = 0; i < slots; ++i)
            if (held[i].body >= 0 && held[i].grp && held[i].build == bd) byG[held[i].G].push_back(i);
        double work = 0;
        for (auto& kv : byG) for (int sl : kv.second) work += (double)kv.first * held[sl].cost;
        for (auto& kv : byG) {
            const int G = kv.first;
            std::vector<int>& v = kv.second;
            std::stable_sort(v.begin(), v.end(), [&](int x, int y) { return held[x].cost > held[y].cost; });
            double w = 0;
            for (int sl : v) w += (double)G * held[sl].cost;
            const long long need = ((long long)v.size() + 64 / G - 1) / (64 / G);
            const long long share = std::max<long long>(1, (long long)std::llround(resident * w / work));
            const long long waves = std::min(need, share);
            if (nctr >= kMaxQueues) return fail(TDYN_E_ARG, "tree server: too many work queues");
            for (long long k = 0; k < waves; ++k) pl.push_back(make_int4(G, (int)idx.size(), (int)v.size(), nctr + 1));
            idx.insert(idx.end(), v.begin(), v.end());
            ++nctr;
        }
        plan_n[bd] = (int)pl.size() - plan_off[bd];
    }
    plan_nctr = nctr;
    if (pl.empty()) { plan_dirty = false; return TDYN_OK; }
    // the pinned staging may still feed the previous upload
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(plan_h.reserve(pl.size())); HIPCHK(pidx_h.reserve(idx.size()));
    std::copy(pl.begin(), pl.end(), plan_h.p);
    std::copy(idx.begin(), idx.end(), pidx_h.p);
    HIPCHK(hipMemcpyAsync(plan.p, plan_h.p, pl.size() * sizeof(int4), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(plan_idx.p, pidx_h.p, idx.size() * sizeof(int), hipMemcpyHostToDevice, s));
    plan_dirty = false;
    return TDYN_OK;
}

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
        (e = sv->bfs.alloc((size_t)nbody)) != hipSuccess || (e = sv->bfs_b.alloc((size_t)nbody)) != hipSuccess ||
        (e = sv->lvl.alloc(2 * (size_t)nbody)) != hipSuccess ||
        (e = sv->prm_bfs.alloc((size_t)nbody * PRM_STRIDE)) != hipSuccess ||
        (e = sv->pool_grp.alloc(S)) != hipSuccess || (e = sv->plan.alloc(S)) != hipSuccess ||
        (e = sv->plan_idx.alloc(S)) != hipSuccess || (e = sv->plan_ctr.alloc(kMaxQueues)) != hipSuccess ||
        (e = hipMemset(sv->pool_grp.p, 0, (size_t)S)) != hipSuccess ||
        (e = hipMemset(sv->zero_off.p, 0, (size_t)S * sizeof(int))) != hipSuccess) {
        delete sv;
        return hipfail(e, "tdyn_tree_server_create: device allocation");
    }
    sv->bodies.init(nbody);
    sv->heapA.init(nheap);
    sv->free_list.resize(S);
    for (int i = 0; i < S; ++i) sv->free_list[i] = S - 1 - i;     // pop_back gives slot 0 first
    sv->held.assign(S, Held{-1, 0, -1, 0, 0, 0, 0, 0, 0, 0});
    sv->resident = std::max(1, tree_group_resident_waves());
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
    tree_tables(*tree, t, 0);
    *bytes = (int64_t)((size_t)t.body_entries() * kBodyBytes + (size_t)t.heap_doubles() * sizeof(double) + kSlotBytes);
    return TDYN_OK;
}

int tdyn_tree_server_set_lanes(tdyn_tree_server* sv, int lanes) {
    if (!sv) return fail(TDYN_E_ARG, "tdyn_tree_server_set_lanes: null server");
    if (!(lanes == 0 || lanes == 1 || (lanes >= 4 && lanes <= 64 && (lanes & (lanes - 1)) == 0)))
        return fail(TDYN_E_ARG, "tdyn_tree_server_set_lanes: lanes 0 (automatic), 1 (tree sweeps) or 4..64, a power of two");
    sv->lanes = lanes;
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
    std::vector<TreeTables>& tabs = sv->tabs;      // reused: their vectors keep their capacity
    std::vector<Held> place;
    size_t nints = 0, ndbl = 0;
    long long qo = 0, uo = 0, fo = 0;
    std::vector<long long> qoff, uoff, foff;
    for (int i = 0; i < n && (int)place.size() < (int)sv->free_list.size(); ++i) {
        if (tabs.size() <= place.size()) tabs.emplace_back();
        TreeTables& t = tabs[place.size()];
        tree_tables(trees[i], t, sv->lanes);
        const long long heapn = t.heap_doubles(), nbe = t.body_entries();
        if (nbe > sv->bodies.cap || heapn > sv->heapA.cap) {
            for (const Held& p : place) { sv->bodies.give(p.body, p.nbody); sv->heapA.give(p.heap, p.nheap); }
            return fail(TDYN_E_NOMEM, "tdyn_tree_server_admit: request " + std::to_string(i) +
                                      " needs more than the server's whole pools");
        }
        const long long b0 = sv->bodies.take(nbe);
        if (b0 < 0) break;
        const long long h0 = sv->heapA.take(heapn);
        if (h0 < 0) { sv->bodies.give(b0, nbe); break; }
        place.push_back(Held{b0, nbe, h0, heapn, t.nq, t.nu, t.grp, t.build, t.G, t.cost});
        nints += 12 + 9 * (size_t)t.nb + (t.grp ? 8 * (size_t)t.nb + t.starts.size() + 2 : 0);
        ndbl += (size_t)t.nb * PRM_STRIDE + 3 + t.nq + t.nu + 1 + 6 * (size_t)(t.nb + 1) + 2;
        qoff.push_back(qo); uoff.push_back(uo); foff.push_back(fo);
        qo += t.nq; uo += t.nu; fo += 6 * (t.nb + 1);
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
        // the level list and the breadth-first constants sit in the slot's
        // body-pool entries: two ints and PRM_STRIDE doubles per entry
        const int lvl_off = 2 * (int)p.body, pbfs_off = (int)p.body * PRM_STRIDE;
        const int nlev = t.grp ? (int)t.starts.size() - 1 : 0;
        I[0] = t.nb; I[1] = t.nq; I[2] = t.nu; I[3] = body_off; I[4] = (int)p.body; I[5] = t.nquat; I[6] = t.rec;
        I[7] = t.grp; I[8] = lvl_off; I[9] = pbfs_off; I[10] = nlev; I[11] = 0;
        std::copy(t.body_a.begin(), t.body_a.end(), I + 12);
        int* bb = I + 12 + 4 * t.nb;
        for (int b = 0; b < t.nb; ++b) {
            bb[4*b] = t.body_b[4*b] + body_off; bb[4*b+1] = t.body_b[4*b+1];
            bb[4*b+2] = t.body_b[4*b+2]; bb[4*b+3] = t.body_b[4*b+3];
        }
        std::copy(t.child.begin(), t.child.end(), bb + 4 * t.nb);
        io += 12 + 9 * (size_t)t.nb;
        if (t.grp) {
            int* g = I + 12 + 9 * t.nb;
            std::copy(t.bfs.begin(), t.bfs.end(), g);
            std::copy(t.bfs_b.begin(), t.bfs_b.end(), g + 4 * t.nb);
            int* lv = g + 8 * t.nb;
            lv[0] = nlev; lv[1] = pbfs_off;
            std::copy(t.starts.begin(), t.starts.end(), lv + 2);
            io += 8 * (size_t)t.nb + t.starts.size() + 2;
        }
        double* D = sv->dbl_h.p + dof;
        for (int b = 1; b <= t.nb; ++b) pack_body_params(trees[i], b, D + (size_t)(b-1) * PRM_STRIDE);
        D += (size_t)t.nb * PRM_STRIDE;
        D[0] = trees[i].gravity[0]; D[1] = trees[i].gravity[1]; D[2] = trees[i].gravity[2];
        D += 3;
        std::copy(q + qoff[i], q + qoff[i] + t.nq, D); D += t.nq;
        std::copy(u + uoff[i], u + uoff[i] + t.nu, D); D += t.nu;
        *D++ = t_end[i];
        for (int j = 0; j < 6 * (t.nb + 1); ++j) D[j] = fapp ? fapp[foff[i] + j] : 0.0;
        D += 6 * (t.nb + 1);
        // vectors after the records
        D[0] = (double)(p.heap + (t.grp ? tgroup::scratch_doubles(t.nb) : t.rec)); D[1] = (double)p.heap;
        dof += (size_t)t.nb * PRM_STRIDE + 3 + t.nq + t.nu + 1 + 6 * (size_t)(t.nb + 1) + 2;
        if (p.grp) ++sv->n_grp; else ++sv->n_sweep;
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
    sv->plan_dirty = true;
    *n_admitted = k;
    return TDYN_OK;
}

int tdyn_tree_server_advance(tdyn_tree_server* sv, int max_steps, hipStream_t s) {
    if (!sv) return fail(TDYN_E_ARG, "tdyn_tree_server_advance: null server");
    if (max_steps <= 0) return fail(TDYN_E_ARG, "max_steps must be positive");
    if (sv->free_list.size() == (size_t)sv->slots) return TDYN_OK;
    if (sv->n_grp > 0 && sv->plan_dirty) { const int rc = sv->build_plan(s); if (rc) return rc; }
    RkmArgs a = sv->rk();
    a.cfg.max_steps = max_steps;
    if (sv->n_grp > 0) {
        for (int bd = 0; bd < 3; ++bd) {
            if (!sv->plan_n[bd]) continue;
            RkmArgs g = a;
            g.grp = 1;
            g.idf = bd == 1 ? 1 : 0;
            g.ext = bd == 2 ? 1 : 0;
            g.plan = sv->plan.p + sv->plan_off[bd];
            g.plan_waves = sv->plan_n[bd];
            HIPCHK(launch_tree_group(g, s, false));
        }
    }
    if (sv->n_sweep > 0) HIPCHK(launch_rkm(a, s));
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
        if (p.grp) --sv->n_grp; else --sv->n_sweep;
        p = Held{-1, 0, -1, 0, 0, 0, 0, 0, 0, 0};
        sv->free_list.push_back(done[i]);
    }
    sv->plan_dirty = true;
    *n_out = k;
    return TDYN_OK;
}

} // extern "C"
