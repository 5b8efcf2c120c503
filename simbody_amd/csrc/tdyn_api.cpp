// Host side of the tdyn C ABI: validates and packs tree descriptions into
// device tables, owns the per-lane scratch and the device-resident integrator
// state, and launches the kernels of tdyn_kernels.hip.
#include "tdyn_host.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

using namespace tdyn;
using namespace tdyn::host;

namespace tdyn {
namespace host {
std::string& err_slot() { static thread_local std::string e; return e; }
} // namespace host
} // namespace tdyn

namespace {
bool finite_all(const double* p, int n) { for (int i = 0; i < n; ++i) if (!std::isfinite(p[i])) return false; return true; }
} // namespace

namespace tdyn {
namespace host {
int validate_tree(const tdyn_tree_desc& d, int t) {
    char buf[256];
    if (d.nb <= 0) { snprintf(buf, sizeof buf, "tree %d: nb must be positive (got %d)", t, d.nb); return fail(TDYN_E_ARG, buf); }
    if (!d.parent || !d.mobilizer || !d.mass || !d.com_B || !d.inertia_B || !d.X_PF || !d.X_BM)
        return fail(TDYN_E_ARG, "tree " + std::to_string(t) + ": null array in tdyn_tree_desc");
    for (int b = 1; b <= d.nb; ++b) {
        const int p = d.parent[b-1], ty = d.mobilizer[b-1];
        if (p < 0 || p >= b) { snprintf(buf, sizeof buf, "tree %d body %d: parent %d must be in [0,%d)", t, b, p, b); return fail(TDYN_E_ARG, buf); }
        if (ty < TDYN_PIN || ty > TDYN_BUSHING) { snprintf(buf, sizeof buf, "tree %d body %d: unknown mobilizer %d", t, b, ty); return fail(TDYN_E_ARG, buf); }
        if (ty == TDYN_SCREW && (!d.screw_pitch || !std::isfinite(d.screw_pitch[b-1]))) {
            snprintf(buf, sizeof buf, "tree %d body %d: Screw needs a finite screw_pitch", t, b); return fail(TDYN_E_ARG, buf);
        }
        if (ty == TDYN_BUSHING && p != 0) { snprintf(buf, sizeof buf, "tree %d body %d: Bushing mobilizer allowed on Ground only in this build", t, b); return fail(TDYN_E_ARG, buf); }
        if ((ty == TDYN_FREE || ty == TDYN_BENDSTRETCH || ty == TDYN_BUSHING) && d.reversed && d.reversed[b-1]) { snprintf(buf, sizeof buf, "tree %d body %d: a reversed %s mobilizer is not supported in this build", t, b, ty == TDYN_FREE ? "Free" : ty == TDYN_BUSHING ? "Bushing" : "BendStretch"); return fail(TDYN_E_ARG, buf); }
        if (!(d.mass[b-1] > 0) || !std::isfinite(d.mass[b-1])) { snprintf(buf, sizeof buf, "tree %d body %d: mass must be positive and finite", t, b); return fail(TDYN_E_ARG, buf); }
        if (!finite_all(d.com_B + 3*(b-1), 3) || !finite_all(d.inertia_B + 6*(b-1), 6) ||
            !finite_all(d.X_PF + 12*(b-1), 12) || !finite_all(d.X_BM + 12*(b-1), 12)) {
            snprintf(buf, sizeof buf, "tree %d body %d: non-finite mass properties or frames", t, b); return fail(TDYN_E_ARG, buf);
        }
    }
    if (!finite_all(d.gravity, 3)) return fail(TDYN_E_ARG, "non-finite gravity");
    return TDYN_OK;
}

bool identity_frames(const tdyn_tree_desc& d) {
    static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int b = 0; b < d.nb; ++b) {
        for (int k = 0; k < 9; ++k)
            if (d.X_PF[12*b + k] != I[k] || d.X_BM[12*b + k] != I[k]) return false;
        for (int k = 9; k < 12; ++k)
            if (d.X_PF[12*b + k] != 0.0) return false;
    }
    return true;
}

void pack_body_params(const tdyn_tree_desc& d, int b, double* P) {
    for (int k = 0; k < PRM_STRIDE; ++k) P[k] = 0.0;
    for (int k = 0; k < 9; ++k) P[PRM_R_PF + k] = d.X_PF[12*(b-1) + k];
    for (int k = 0; k < 3; ++k) P[PRM_P_PF + k] = d.X_PF[12*(b-1) + 9 + k];
    const double* Rbm = d.X_BM + 12*(b-1);
    const double* pbm = Rbm + 9;
    // X_MB = inverse(X_BM): R_MB = R_BM^T, p_MB = -R_BM^T p_BM
    for (int r = 0; r < 3; ++r) for (int c2 = 0; c2 < 3; ++c2) P[PRM_R_MB + 3*r + c2] = Rbm[3*c2 + r];
    for (int r = 0; r < 3; ++r)
        P[PRM_P_MB + r] = -(P[PRM_R_MB + 3*r]*pbm[0] + P[PRM_R_MB + 3*r + 1]*pbm[1] + P[PRM_R_MB + 3*r + 2]*pbm[2]);
    const double m = d.mass[b-1];
    P[PRM_MASS] = m;
    for (int k = 0; k < 3; ++k) P[PRM_COM + k] = d.com_B[3*(b-1) + k];
    for (int k = 0; k < 6; ++k) P[PRM_G + k] = d.inertia_B[6*(b-1) + k] / m;
    for (int k = 0; k < 3; ++k) P[PRM_P_BM + k] = pbm[k];
    if (d.mobilizer[b-1] == TDYN_SCREW) P[PRM_PITCH] = d.screw_pitch[b-1];
    if (d.reversed && d.reversed[b-1]) P[PRM_REV] = 1.0;
    if (d.use_euler_angles && (d.mobilizer[b-1] == TDYN_BALL || d.mobilizer[b-1] == TDYN_FREE)) P[PRM_EULER] = 1.0;
    if (d.mobilizer[b-1] == TDYN_BUSHING) P[PRM_EULER] = 2.0;   // runs as a Free body with x-y-z angles, u = qdot
}

bool any_reversed(const tdyn_tree_desc& d) {
    if (!d.reversed) return false;
    for (int b = 0; b < d.nb; ++b) if (d.reversed[b]) return true;
    return false;
}

int make_rkm_cfg(const tdyn_integrator_cfg* cfg, RkmCfg& c, int& precision) {
    if (cfg->precision < 0 || cfg->precision > 3)
        return fail(TDYN_E_ARG, "integrator precision must be 0 (FP64), 1 (mixed), 2 or 3 (FP64 with "
                                "single-precision hand-over rows)");
    precision = cfg->precision;
    const double acc = cfg->accuracy > 0 ? cfg->accuracy : 1e-3;
    if (!(acc < 1.0)) return fail(TDYN_E_ARG, "accuracy must be in (0,1)");
    c.accuracy = acc;
    c.cons_tol = cfg->constraint_tolerance > 0 ? cfg->constraint_tolerance : acc / 10;
    c.fixed_h = cfg->fixed_step > 0 ? cfg->fixed_step : 0.0;
    c.init_h = cfg->initial_step > 0 ? cfg->initial_step : -1.0;
    c.min_h = cfg->min_step > 0 ? cfg->min_step : -1.0;
    c.max_h = cfg->max_step > 0 ? cfg->max_step : -1.0;
    c.timescale = cfg->time_scale > 0 ? cfg->time_scale : 0.1;
    c.max_attempts = cfg->max_attempts > 0 ? cfg->max_attempts : 100000000;
    c.max_steps = 0;
    c.inf_norm = cfg->use_infinity_norm ? 1 : 0;
    c.force_proj = cfg->project_every_step ? 1 : 0;
    c.proj_interp = cfg->project_interpolated_states < 0 ? 0 : 1;
    if ((c.inf_norm || c.force_proj) && precision != 0)
        return fail(TDYN_E_ARG, "the infinity norm and projection on every step run in FP64 (precision 0)");
    return TDYN_OK;
}

// Wavefront plan of the level-parallel tree evaluator for the next rollout
// (returns 0 and leaves the plan empty for the one-lane streaming
// evaluator).  Lanes per instance: G, the tree_lanes setting, or automatic:
// lanes_min (0: the most lanes, up to 16, that keep the batch within about
// four groups per resident group slot).  The instances are dealt to groups
// of G lanes by a work queue: persistent wavefronts (as many as the device
// keeps resident) each run 64/G groups, and a group that finishes an
// instance takes the next from the queue, so a wavefront's lanes stay busy
// while its instances' step counts differ.  The queue is ordered by
// decreasing cost estimate, the level steps of one derivative (sum over the
// tree's levels of ceil(width / G)); a rollout lasts at least as long as
// its slowest instance, so with the automatic setting an instance whose
// estimate exceeds the mean load of a group slot runs first, in a wavefront
// of its own class with lanes_max lanes.  Measured on 16,384 trees of
// 8..128 bodies, before the queue (static wavefronts of instances ordered
// by size; body-steps/s): 4, 8, 16, 32, 64 lanes for every instance 0.56,
// 1.06, 1.32, 1.27, 0.95 x 1e8; per instance its widest level in [16, 32]
// 1.31e8.
int build_tree_plan(tdyn_batch* b, hipStream_t s) {
    if (b->tree_lanes == 1) return 0;
    const int L = b->n_inst;
    const int64_t resident = std::max(1, tree_group_resident_waves());
    int gq = b->tree_lanes > 1 ? b->tree_lanes : b->lanes_min;
    if (gq <= 0) {
        gq = 4;
        while (gq < 16 && (int64_t)L * 2 * gq <= 4 * 64 * resident) gq *= 2;
    }
    const int gbig = b->tree_lanes > 1 ? gq : std::max(gq, b->lanes_max);
    auto steps = [&](int l, int G) {
        int64_t c = 0;
        for (int w : b->h_tree_lw[b->h_lane_tree[l]]) c += (w + G - 1) / G;
        return c;
    };
    std::vector<int64_t> cost(L);
    int64_t total = 0;
    for (int l = 0; l < L; ++l) { cost[l] = steps(l, gq); total += cost[l]; }
    std::vector<int> ord(L);
    for (int l = 0; l < L; ++l) ord[l] = l;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    // the mean load of one group slot of the resident wavefronts
    const double mean_load = (double)total / (double)(resident * (64 / gq));
    std::vector<int> big, rest;
    for (int l : ord) ((gbig > gq && cost[l] > mean_load) ? big : rest).push_back(l);
    std::vector<int4> plan; std::vector<int> idx;
    for (size_t k = 0; k < big.size(); k += 64 / gbig) {
        const int cnt = (int)std::min(big.size() - k, (size_t)(64 / gbig));
        plan.push_back(make_int4(gbig, (int)idx.size(), cnt, 0));
        idx.insert(idx.end(), big.begin() + k, big.begin() + k + cnt);
    }
    if (!rest.empty()) {
        const int64_t waves = std::min<int64_t>(resident, ((int64_t)rest.size() + 64 / gq - 1) / (64 / gq));
        for (int64_t w = 0; w < waves; ++w) plan.push_back(make_int4(gq, (int)idx.size(), (int)rest.size(), 1));
        idx.insert(idx.end(), rest.begin(), rest.end());
    }
    hipError_t e;
    if ((e = b->plan.upload(plan)) != hipSuccess || (e = b->plan_idx.upload(idx)) != hipSuccess)
        return hipfail(e, "tree evaluator plan upload");
    if (!b->plan_ctr.p && (e = b->plan_ctr.alloc(1)) != hipSuccess) return hipfail(e, "tree evaluator queue counter");
    b->plan_waves = (int)plan.size();
    (void)s;
    return 0;
}

int ensure_constraints(tdyn_batch* b) {
    if (!b->cons_dirty) return TDYN_OK;
    std::vector<int4> ct, ca;
    std::vector<double> cp;
    for (size_t t = 0; t < b->h_cons.size(); ++t) {
        int rows = 0;
        for (const tdyn_constraint_desc& c : b->h_cons[t]) {
            ca.push_back(make_int4(c.type, c.body1, c.body2, 0));
            for (int k = 0; k < 3; ++k) cp.push_back(c.station1[k]);
            for (int k = 0; k < 3; ++k) cp.push_back(c.station2[k]);
            cp.push_back(c.length);
            rows += c.type == TDYN_CONS_BALL ? 3 : 1;
        }
        ct.push_back(make_int4((int)ca.size() - (int)b->h_cons[t].size(), (int)b->h_cons[t].size(), rows, 0));
    }
    if (ca.empty()) { ca.push_back(make_int4(0, 0, 0, 0)); cp.assign(7, 0.0); }
    hipError_t e;
    if ((e = b->cons_t.upload(ct)) != hipSuccess || (e = b->cons_a.upload(ca)) != hipSuccess ||
        (e = b->cons_p.upload(cp)) != hipSuccess)
        return hipfail(e, "constraint tables upload");
    const size_t ws = (size_t)tdyn::CWS_N * b->n_inst;
    if (b->cons_ws.n != ws && (e = b->cons_ws.alloc(ws)) != hipSuccess) return hipfail(e, "constraint workspace");
    b->cons_dirty = false;
    return TDYN_OK;
}

int ensure_witnesses(tdyn_batch* b) {
    if (!b->wit_dirty) return TDYN_OK;
    std::vector<int4> wt, wa;
    std::vector<double> wp;
    for (size_t t = 0; t < b->h_wit.size(); ++t) {
        int need = 0;
        for (const tdyn_witness_desc& w : b->h_wit[t]) {
            wa.push_back(make_int4(w.kind | (w.trigger << 4) | (w.action << 8), w.index, w.body, w.action_index));
            for (int k = 0; k < 3; ++k) wp.push_back(w.station[k]);
            for (int k = 0; k < 3; ++k) wp.push_back(w.direction[k]);
            wp.push_back(w.offset); wp.push_back(w.factor); wp.push_back(w.window);
            if (w.kind == TDYN_WIT_STATION_POS) need = std::max(need, 1);
            if (w.kind == TDYN_WIT_STATION_VEL) need = 2;
        }
        wt.push_back(make_int4((int)wa.size() - (int)b->h_wit[t].size(), (int)b->h_wit[t].size(), need, 0));
    }
    if (wa.empty()) { wa.push_back(make_int4(0, 0, 0, 0)); wp.assign(9, 0.0); }
    hipError_t e;
    if ((e = b->wit_t.upload(wt)) != hipSuccess || (e = b->wit_a.upload(wa)) != hipSuccess ||
        (e = b->wit_p.upload(wp)) != hipSuccess)
        return hipfail(e, "witness tables upload");
    b->wit_dirty = false;
    return TDYN_OK;
}

int ensure_rollout_buffers(tdyn_batch* b, int precision) {
    const int ny = b->max_nq + b->max_nu;
    const size_t L = b->n_inst;
    hipError_t e;
    if (precision != 0) {
        const size_t n32 = (size_t)(6 * b->max_nu + 4 * b->max_nb) * L;
        if (b->scratch32.n < n32 && (e = b->scratch32.alloc(n32)) != hipSuccess)
            return hipfail(e, "mixed-precision scratch allocation");
    }
    if (b->vec.p && b->ny_max == ny) return TDYN_OK;
    b->ny_max = 0;
    // 8 extra rows: the chain evaluator prefetches fixed-size q/u windows
    if ((e = b->vec.alloc((size_t)11 * ny * L + 8 * L)) != hipSuccess ||
        (e = b->uscale.alloc((size_t)std::max(1, b->max_nu) * L)) != hipSuccess ||
        (e = b->t.alloc(L)) != hipSuccess || (e = b->h.alloc(L)) != hipSuccess ||
        (e = b->t_end.alloc(L)) != hipSuccess || (e = b->status.alloc(L)) != hipSuccess ||
        (e = b->stats.alloc(4 * L)) != hipSuccess)
        return hipfail(e, "rollout state allocation");
    b->ny_max = ny;
    return TDYN_OK;
}
} // namespace host
} // namespace tdyn

extern "C" {

const char* tdyn_last_error(void) { return err_slot().c_str(); }

int tdyn_batch_create(const tdyn_tree_desc* trees, int n_trees, const int32_t* tree_of_instance,
                      int n_instances, tdyn_batch** out) {
    if (!out) return fail(TDYN_E_ARG, "out is null");
    *out = nullptr;
    if (!trees || n_trees <= 0) return fail(TDYN_E_ARG, "need at least one tree");
    if (n_instances <= 0) return fail(TDYN_E_ARG, "need at least one instance");
    for (int t = 0; t < n_trees; ++t) { int rc = validate_tree(trees[t], t); if (rc) return rc; }
    if (tree_of_instance)
        for (int i = 0; i < n_instances; ++i)
            if (tree_of_instance[i] < 0 || tree_of_instance[i] >= n_trees)
                return fail(TDYN_E_ARG, "tree_of_instance[" + std::to_string(i) + "] out of range");

    auto* B = new tdyn_batch;
    B->n_inst = n_instances; B->n_trees = n_trees;
    // reversed mobilizers and Euler-angle Balls run as padded generic
    // bodies, as in batches of different trees
    for (int t = 0; t < n_trees; ++t) {
        B->any_rev = B->any_rev || any_reversed(trees[t]);
        B->any_euler = B->any_euler || trees[t].use_euler_angles != 0;
        for (int b = 0; b < trees[t].nb; ++b) {
            B->any_bendstretch = B->any_bendstretch || trees[t].mobilizer[b] == TDYN_BENDSTRETCH;
            B->any_bushing = B->any_bushing || trees[t].mobilizer[b] == TDYN_BUSHING;
        }
    }
    B->uniform = ((tree_of_instance == nullptr) || n_trees == 1) && !B->any_rev && !B->any_euler && !B->any_bushing;
    std::vector<int4> body_b, pre, bfs, bfs_b; std::vector<int> child_idx, lvl; std::vector<double> prm, grav, prm_bfs;
    B->stream_ok = !B->uniform;
    std::vector<int> tq(n_trees), tu(n_trees), tnb(n_trees);
    int body_off = 0;
    for (int t = 0; t < n_trees; ++t) {
        const tdyn_tree_desc& d = trees[t];
        const int nb = d.nb;
        std::vector<std::vector<int>> ch(nb + 1);
        for (int b = 1; b <= nb; ++b) ch[d.parent[b-1]].push_back(b);
        int q = 0, u = 0, rec = 0, nquat = 0, gslot = 0;
        bool chain = true;
        for (int b = 1; b <= nb; ++b) {
            const int ty = d.mobilizer[b-1];
            // the kernels carry a Bushing as a Free body flagged PRM_EULER = 2
            const int dty = ty == TDYN_BUSHING ? TDYN_FREE : ty;
            // (the chain evaluator carries a Free body as the first body only)
            chain = chain && d.parent[b-1] == b - 1 && !mob_generic_only(ty) && !(ty == TDYN_FREE && b > 1);
            B->h_body_a.push_back(make_int4(d.parent[b-1], dty, q, u));
            body_b.push_back(make_int4((int)child_idx.size(), (int)ch[b].size(), rec, gslot));
            for (int c : ch[b]) child_idx.push_back(c);
            // batches of different trees carry non-Free bodies in padded 3-dof records
            const int rd = ((!B->uniform || mob_generic_only(dty)) && dty != TDYN_FREE) ? 3 : mob_dof(dty);
            q += mob_nq(ty); u += mob_dof(ty); rec += rec_size(rd); gslot += 7 * mob_dof(ty);
            if ((ty == TDYN_BALL || ty == TDYN_FREE) && !d.use_euler_angles) ++nquat;
            double P[PRM_STRIDE];
            pack_body_params(d, b, P);
            prm.insert(prm.end(), P, P + PRM_STRIDE);
        }
        // depth-first preorder (children in increasing order): body, depth, child count
        {
            std::vector<int> depth(nb + 1, 0), stack(ch[0].rbegin(), ch[0].rend());
            while (!stack.empty()) {
                const int b = stack.back(); stack.pop_back();
                depth[b] = depth[d.parent[b-1]] + 1;
                B->dmax = std::max(B->dmax, depth[b]);
                pre.push_back(make_int4(b, depth[b], (int)ch[b].size(), 0));
                stack.insert(stack.end(), ch[b].rbegin(), ch[b].rend());
            }
            for (int b = 2; b <= nb; ++b)
                if (d.mobilizer[b-1] == TDYN_FREE || d.mobilizer[b-1] == TDYN_BUSHING)
                    B->stream_ok = false;   // a 6-dof body must come first
        }
        for (int k = 0; k < 3; ++k) grav.push_back(d.gravity[k]);
        grav.push_back(0.0);
        B->h_tree_a.push_back(make_int4(nb, q, u, body_off));
        // breadth-first order (roots, then each level's children in order):
        // body, parent's position, first child's position, child count; and
        // the start of every level
        const int lvl_off = (int)lvl.size();
        size_t prm_base = 0;
        {
            std::vector<int> order(ch[0]), pos(nb + 1, -1), depth(nb + 1, 0);
            std::vector<int> starts{0};
            for (size_t k = 0; k < order.size(); ++k) depth[order[k]] = 1;
            for (size_t k = 0; k < order.size(); ++k) {
                const int b = order[k];
                pos[b] = (int)k;
                if (k > 0 && depth[b] != depth[order[k - 1]]) starts.push_back((int)k);
                for (int c : ch[b]) { depth[c] = depth[b] + 1; order.push_back(c); }
            }
            starts.push_back(nb);
            for (int k = 0; k < nb; ++k) {
                const int b = order[k], p = d.parent[b - 1];
                const int first = ch[b].empty() ? 0 : pos[ch[b][0]];
                bfs.push_back(make_int4(b, p == 0 ? -1 : pos[p], first, (int)ch[b].size()));
                const int4 ba = B->h_body_a[body_off + b - 1];
                bfs_b.push_back(make_int4(ba.y, ba.z, ba.w, b));
            }
            // model constants in breadth-first order, in blocks of BW positions
            // stored constant-major
            const size_t base = prm_bfs.size();
            const int BW = tdyn::tgroup::BW;
            prm_bfs.resize(base + (size_t)tdyn::tgroup::blocks(nb) * BW * PRM_STRIDE, 0.0);
            for (int k = 0; k < nb; ++k)
                for (int c = 0; c < PRM_STRIDE; ++c)
                    prm_bfs[base + (size_t)(k / BW) * PRM_STRIDE * BW + (size_t)c * BW + k % BW] =
                        prm[(size_t)(body_off + order[k] - 1) * PRM_STRIDE + c];
            prm_base = base;
            int maxw = 0;
            for (size_t k = 0; k + 1 < starts.size(); ++k) maxw = std::max(maxw, starts[k + 1] - starts[k]);
            B->h_tree_maxw.push_back(maxw);
            std::vector<int> lw;
            for (size_t k = 0; k + 1 < starts.size(); ++k) lw.push_back(starts[k + 1] - starts[k]);
            B->h_tree_lw.push_back(lw);
            lvl.push_back((int)starts.size() - 1);
            lvl.push_back((int)prm_base);
            lvl.insert(lvl.end(), starts.begin(), starts.end());
        }
        B->h_tree_b.push_back(make_int4(rec, nquat, gslot, lvl_off));
        tq[t] = q; tu[t] = u; tnb[t] = nb;
        if (t == 0) B->is_chain = chain;
        B->max_rec = std::max(B->max_rec, rec);
        body_off += nb;
    }
    B->h_tree_of.resize(n_instances);
    B->h_qoff.resize(n_instances); B->h_uoff.resize(n_instances); B->h_boff.resize(n_instances);
    int64_t q = 0, u = 0, bb = 0;
    for (int i = 0; i < n_instances; ++i) {
        const int t = tree_of_instance ? tree_of_instance[i] : 0;
        B->h_tree_of[i] = t;
        B->h_qoff[i] = (int)q; B->h_uoff[i] = (int)u; B->h_boff[i] = (int)bb;
        q += tq[t]; u += tu[t]; bb += tnb[t] + 1;
        B->max_nq = std::max(B->max_nq, tq[t]); B->max_nu = std::max(B->max_nu, tu[t]);
        B->max_nb = std::max(B->max_nb, tnb[t]);
    }
    if (q > INT32_MAX || u > INT32_MAX || bb * 21 > INT32_MAX || prm_bfs.size() > (size_t)INT32_MAX) { delete B; return fail(TDYN_E_SIZE, "batch too large for 32-bit offsets"); }
    B->total_q = q; B->total_u = u; B->total_b = bb;
    B->is_chain = B->is_chain && B->uniform;
    B->stream_stride = tdyn::tstream::scratch_doubles(B->dmax, B->max_nb);
    B->group_stride = tdyn::tgroup::scratch_doubles(B->max_nb);
    B->idf = B->is_chain && identity_frames(trees[0]);
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            B->n_simd = 4 * cus;
    }
    B->idf_all = true;
    for (int t = 0; t < n_trees && B->idf_all; ++t) B->idf_all = identity_frames(trees[t]);
    // Lane order.  A wave runs as long as its slowest lane, so instances of
    // different trees are assigned to lanes in order of decreasing work
    // (bodies weighted by their mobilizer's cost): the 64 lanes of a wave
    // then walk trees of similar size.  Per-lane tables are permuted;
    // per-instance results are mapped back through inst_of_lane.
    std::vector<int> lane_tree(n_instances), lane_q(n_instances), lane_u(n_instances), lane_b(n_instances);
    std::vector<int> inst_of_lane(n_instances);
    for (int i = 0; i < n_instances; ++i) inst_of_lane[i] = i;
    if (!B->uniform) {
        static const double cost[13] = {1.0, 0.9, 1.4, 1.8, 3.3, 1.5, 1.5, 1.2, 1.0, 1.9, 0.6, 1.3, 3.4};
        static_assert(sizeof cost / sizeof cost[0] == TDYN_BUSHING + 1, "one cost per mobilizer");
        std::vector<double> tw(n_trees, 0.0);
        for (int t = 0; t < n_trees; ++t)
            for (int bdy = 0; bdy < trees[t].nb; ++bdy) tw[t] += cost[trees[t].mobilizer[bdy]];
        std::stable_sort(inst_of_lane.begin(), inst_of_lane.end(), [&](int x, int y) {
            return tw[B->h_tree_of[x]] > tw[B->h_tree_of[y]];
        });
        B->permuted = true;
    }
    for (int l = 0; l < n_instances; ++l) {
        const int i = inst_of_lane[l];
        lane_tree[l] = B->h_tree_of[i]; lane_q[l] = B->h_qoff[i]; lane_u[l] = B->h_uoff[i]; lane_b[l] = B->h_boff[i];
    }
    B->h_lane_tree = lane_tree;
    {
    }
    hipError_t e;
    if ((e = B->body_a.upload(B->h_body_a)) != hipSuccess ||
        (e = B->body_b.upload(body_b)) != hipSuccess ||
        (e = B->child_idx.upload(child_idx.empty() ? std::vector<int>{0} : child_idx)) != hipSuccess ||
        (e = B->prm.upload(prm)) != hipSuccess ||
        (e = B->tree_a.upload(B->h_tree_a)) != hipSuccess ||
        (e = B->tree_b.upload(B->h_tree_b)) != hipSuccess ||
        (e = B->gravity.upload(grav)) != hipSuccess ||
        (e = B->tree_of.upload(lane_tree)) != hipSuccess ||
        (e = B->q_off.upload(lane_q)) != hipSuccess ||
        (e = B->u_off.upload(lane_u)) != hipSuccess ||
        (e = B->b_off.upload(lane_b)) != hipSuccess ||
        (e = B->inst.upload(inst_of_lane)) != hipSuccess ||
        (e = B->pre.upload(pre)) != hipSuccess ||
        (e = B->bfs.upload(bfs)) != hipSuccess ||

        (e = B->bfs_b.upload(bfs_b)) != hipSuccess ||
        (e = B->prm_bfs.upload(prm_bfs)) != hipSuccess ||
        (e = B->lvl.upload(lvl)) != hipSuccess ||
        (e = B->scratch.alloc((size_t)std::max({B->max_rec + 24, B->stream_ok ? B->stream_stride : 0,
                                                B->stream_ok ? B->group_stride : 0}) * n_instances)) != hipSuccess ||
        (e = B->op_status.alloc((size_t)n_instances)) != hipSuccess ||
        (e = hipMemset(B->op_status.p, 0, (size_t)n_instances * sizeof(int))) != hipSuccess) {
        delete B;
        return hipfail(e, "tdyn_batch_create: device allocation/upload");
    }
    *out = B;
    return TDYN_OK;
}

int tdyn_batch_destroy(tdyn_batch* b) { delete b; return TDYN_OK; }

int tdyn_batch_set_evaluator(tdyn_batch* b, int mode) {
    if (!b) return fail(TDYN_E_ARG, "null batch");
    if (mode < 0 || mode > 3) return fail(TDYN_E_ARG, "evaluator mode must be 0, 1, 2 or 3");
    if (mode == 2 && !b->is_chain) return fail(TDYN_E_ARG, "chain evaluator needs one serial-chain model for the whole batch");
    b->evaluator = mode;
    return TDYN_OK;
}

int tdyn_batch_set_known_udot(tdyn_batch* b, const int32_t* known, const double* udot_known) {
    if (!b) return fail(TDYN_E_ARG, "null batch");
    if ((known == nullptr) != (udot_known == nullptr))
        return fail(TDYN_E_ARG, "tdyn_batch_set_known_udot: known and udot_known go together");
    b->known_p = known; b->known_u = udot_known;
    return TDYN_OK;
}

int tdyn_batch_set_motion(tdyn_batch* b, const double* motion) {
    if (!b) return fail(TDYN_E_ARG, "null batch");
    if (!motion) { b->motion = nullptr; return TDYN_OK; }
    // the description is checked once, here: levels, finite values, and
    // position level only on mobilizers whose qdot is u
    const size_t nbody = (size_t)b->total_b;
    std::vector<double> m(4 * nbody);
    HIPCHK(hipMemcpy(m.data(), motion, m.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int> flags(nbody, 0);
    for (int i = 0; i < b->n_inst; ++i) {
        const int tr = b->h_tree_of.empty() ? 0 : b->h_tree_of[i];
        const int4 ta = b->h_tree_a[tr];
        for (int k = 1; k <= ta.x; ++k) {
            const double* r = m.data() + 4 * ((size_t)b->h_boff[i] + k);
            const int level = (int)r[0];
            if (r[0] != (double)level || level < 0 || level > 4 || !std::isfinite(r[1]) || !std::isfinite(r[2]) ||
                !std::isfinite(r[3]))
                return fail(TDYN_E_ARG, "tdyn_batch_set_motion: instance " + std::to_string(i) + " body " +
                                        std::to_string(k) + ": level must be 0-4 and a, w, phi finite");
            const int ty = b->h_body_a[ta.w + k - 1].y;
            if (level == 1 && (ty == TDYN_BALL || ty == TDYN_FREE))
                return fail(TDYN_E_ARG, "tdyn_batch_set_motion: position-level motion needs a mobilizer with "
                                        "qdot = u (not Ball or Free)");
            flags[b->h_boff[i] + k] = level != 0;
        }
    }
    HIPCHK(b->motion_flags.upload(flags));
    if (b->motion_udot.n != (size_t)std::max<int64_t>(1, b->total_u))
        HIPCHK(b->motion_udot.alloc((size_t)std::max<int64_t>(1, b->total_u)));
    HIPCHK(hipMemset(b->motion_udot.p, 0, b->motion_udot.n * sizeof(double)));
    b->motion = motion;
    return TDYN_OK;
}

int tdyn_batch_set_constraints(tdyn_batch* b, int tree, int n, const tdyn_constraint_desc* c) {
    if (!b) return fail(TDYN_E_ARG, "null batch");
    if (tree < 0 || tree >= b->n_trees) return fail(TDYN_E_ARG, "tdyn_batch_set_constraints: tree out of range");
    if (n < 0 || (n > 0 && !c)) return fail(TDYN_E_ARG, "tdyn_batch_set_constraints: n >= 0 descriptions required");
    if (n > 0 && b->any_euler)
        return fail(TDYN_E_ARG, "tdyn_batch_set_constraints: not with Euler-angle Ball or Free models");
    const int nb = b->h_tree_a[tree].x;
    int rows = 0;
    for (int i = 0; i < n; ++i) {
        const tdyn_constraint_desc& d = c[i];
        const std::string at = "tdyn_batch_set_constraints: constraint " + std::to_string(i) + ": ";
        if (d.type != TDYN_CONS_BALL && d.type != TDYN_CONS_ROD) return fail(TDYN_E_ARG, at + "unknown type");
        if (d.body1 < 0 || d.body1 > nb || d.body2 < 0 || d.body2 > nb || d.body1 == d.body2)
            return fail(TDYN_E_ARG, at + "bodies must be two different bodies of the tree (0 = Ground)");
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(d.station1[k]) || !std::isfinite(d.station2[k]))
                return fail(TDYN_E_ARG, at + "stations must be finite");
        if (d.type == TDYN_CONS_ROD && !(d.length > 0 && std::isfinite(d.length)))
            return fail(TDYN_E_ARG, at + "a rod's length must be positive");
        rows += d.type == TDYN_CONS_BALL ? 3 : 1;
    }
    if (rows > TDYN_MAX_CONSTRAINT_ROWS)
        return fail(TDYN_E_SIZE, "tdyn_batch_set_constraints: more than TDYN_MAX_CONSTRAINT_ROWS constraint equations");
    if (b->h_cons.size() != (size_t)b->n_trees) b->h_cons.resize(b->n_trees);
    b->h_cons[tree].assign(c, c + n);
    b->any_cons = false;
    for (const auto& v : b->h_cons) b->any_cons = b->any_cons || !v.empty();
    b->cons_dirty = true;
    b->rollout_ready = false;          // a rollout in progress was begun with the old model
    return ensure_constraints(b);
}

int tdyn_batch_set_witnesses(tdyn_batch* b, int tree, int n, const tdyn_witness_desc* w) {
    if (!b) return fail(TDYN_E_ARG, "null batch");
    if (tree < 0 || tree >= b->n_trees) return fail(TDYN_E_ARG, "tdyn_batch_set_witnesses: tree out of range");
    if (n < 0 || (n > 0 && !w)) return fail(TDYN_E_ARG, "tdyn_batch_set_witnesses: n >= 0 descriptions required");
    if (n > TDYN_MAX_WITNESSES) return fail(TDYN_E_SIZE, "tdyn_batch_set_witnesses: more than TDYN_MAX_WITNESSES");
    const int4 ta = b->h_tree_a[tree];
    const int nb = ta.x, nq = ta.y, nu = ta.z;
    for (int i = 0; i < n; ++i) {
        const tdyn_witness_desc& d = w[i];
        const std::string at = "tdyn_batch_set_witnesses: witness " + std::to_string(i) + ": ";
        if (d.kind < TDYN_WIT_Q || d.kind > TDYN_WIT_TIME) return fail(TDYN_E_ARG, at + "unknown kind");
        if (d.kind == TDYN_WIT_Q && (d.index < 0 || d.index >= nq)) return fail(TDYN_E_ARG, at + "no such q");
        if (d.kind == TDYN_WIT_U && (d.index < 0 || d.index >= nu)) return fail(TDYN_E_ARG, at + "no such u");
        if ((d.kind == TDYN_WIT_STATION_POS || d.kind == TDYN_WIT_STATION_VEL) && (d.body < 0 || d.body > nb))
            return fail(TDYN_E_ARG, at + "no such body (0 = Ground)");
        if (d.trigger < 1 || d.trigger > 3)
            return fail(TDYN_E_ARG, at + "trigger must be TDYN_TRIGGER_FALLING and/or TDYN_TRIGGER_RISING");
        if (d.action < TDYN_EVENT_REPORT || d.action > TDYN_EVENT_TERMINATE)
            return fail(TDYN_E_ARG, at + "unknown action");
        if (d.action == TDYN_EVENT_SCALE_U && d.action_index >= nu) return fail(TDYN_E_ARG, at + "no such u to scale");
        const double v[11] = {d.station[0], d.station[1], d.station[2], d.direction[0], d.direction[1],
                              d.direction[2], d.offset, d.factor, d.window, 0.0, 0.0};
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(v[k])) return fail(TDYN_E_ARG, at + "values must be finite");
        if (!(d.window > 0)) return fail(TDYN_E_ARG, at + "the localization window must be positive");
    }
    if (b->h_wit.size() != (size_t)b->n_trees) b->h_wit.resize(b->n_trees);
    b->h_wit[tree].assign(w, w + n);
    b->any_wit = false;
    for (const auto& v : b->h_wit) b->any_wit = b->any_wit || !v.empty();
    b->wit_dirty = true;
    b->rollout_ready = false;          // a rollout in progress was begun without them
    return ensure_witnesses(b);
}

int tdyn_calc_event_triggers(tdyn_batch* b, const double* q, const double* u, const double* t, double* values,
                             hipStream_t s) {
    if (!b || !q || !u || !values) return fail(TDYN_E_ARG, "getEventTriggers: batch, q, u and values are required");
    { const int rc = ensure_witnesses(b); if (rc) return rc; }
    const size_t L = b->n_inst;
    if (!b->any_wit) {
        HIPCHK(hipMemsetAsync(values, 0, L * TDYN_MAX_WITNESSES * sizeof(double), s));
        return TDYN_OK;
    }
    OpArgs a = b->op();
    a.q = q; a.u = u; a.times = t; a.out_wit = values;
    HIPCHK(launch_op(OP_WIT, a, s));
    return TDYN_OK;
}

int tdyn_batch_set_event_capacity(tdyn_batch* b, int capacity) {
    if (!b) return fail(TDYN_E_ARG, "null batch");
    if (capacity < 1 || capacity > (1 << 20)) return fail(TDYN_E_ARG, "tdyn_batch_set_event_capacity: 1 to 2^20");
    b->ev_cap = capacity;
    b->rollout_ready = false;
    return TDYN_OK;
}

int tdyn_rollout_get_events(tdyn_batch* b, double* windows, int32_t* ids, int32_t* counts, hipStream_t s) {
    if (!b || !b->rollout_ready) return fail(TDYN_E_STAGE, "tdyn_rollout_get_events: no rollout");
    if (!b->any_wit || !b->ev_count.p) return fail(TDYN_E_STAGE, "tdyn_rollout_get_events: the rollout has no witnesses");
    const size_t L = b->n_inst, E = L * (size_t)b->ev_cap * 2;
    if (b->permuted) {
        HIPCHK(launch_event_scatter(b->inst.p, (int)L, b->ev_cap, b->ev_win.p, b->ev_ids.p, b->ev_count.p, windows,
                                    ids, counts, s));
        return TDYN_OK;
    }
    if (windows) HIPCHK(hipMemcpyAsync(windows, b->ev_win.p, E * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (ids) HIPCHK(hipMemcpyAsync(ids, b->ev_ids.p, E * sizeof(int), hipMemcpyDeviceToDevice, s));
    if (counts) HIPCHK(hipMemcpyAsync(counts, b->ev_count.p, L * sizeof(int), hipMemcpyDeviceToDevice, s));
    return TDYN_OK;
}

int tdyn_calc_constrained_acceleration(tdyn_batch* b, const double* q, const double* u, const double* F,
                                       double* udot, double* lambda, hipStream_t s) {
    if (!b || !q || !u || !udot) return fail(TDYN_E_ARG, "realize(Acceleration): batch, q, u and udot are required");
    { const int rc = ensure_constraints(b); if (rc) return rc; }
    OpArgs a = b->op();
    a.q = q; a.u = u; a.fbody = F; a.out_u = udot; a.out_lam = lambda;
    HIPCHK(launch_op(OP_FWD_CONS, a, s));
    return TDYN_OK;
}

int tdyn_batch_set_tree_lanes(tdyn_batch* b, int lanes) {
    if (!b) return fail(TDYN_E_ARG, "null batch");
    if (!(lanes == 0 || lanes == 1 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64))
        return fail(TDYN_E_ARG, "lanes per instance must be 0 (automatic), 1, 4, 8, 16, 32 or 64");
    b->tree_lanes = lanes;
    return TDYN_OK;
}

int tdyn_batch_set_tree_lane_range(tdyn_batch* b, int lanes_min, int lanes_max) {
    if (!b) return fail(TDYN_E_ARG, "null batch");
    auto ok = [](int v) { return v == 4 || v == 8 || v == 16 || v == 32 || v == 64; };
    if (!(lanes_min == 0 || ok(lanes_min)) || !ok(lanes_max) || (lanes_min && lanes_min > lanes_max))
        return fail(TDYN_E_ARG, "lane range: lanes_min 0 or 4..64, lanes_max 4..64, powers of two, min <= max");
    b->lanes_min = lanes_min; b->lanes_max = lanes_max;
    return TDYN_OK;
}

int tdyn_batch_get_evaluator(const tdyn_batch* b, int* active, int* lanes) {
    if (!b || !active) return fail(TDYN_E_ARG, "null argument");
    int act = 0, ln = 0;
    if (b->rollout_ready) {
        const RkmArgs a = b->rk();
        act = a.chain ? 2 : a.grp ? 4 : a.vec_aos ? 3 : 1;
        ln = act >= 4 ? (b->tree_lanes > 1 ? b->tree_lanes : 0) : 1;
    }
    *active = act;
    if (lanes) *lanes = ln;
    return TDYN_OK;
}

int tdyn_batch_op_status(tdyn_batch* b, int32_t* status, int* n_singular, hipStream_t s) {
    if (!b) return fail(TDYN_E_ARG, "null batch");
    const size_t L = b->n_inst;
    if (status) HIPCHK(hipMemcpyAsync(status, b->op_status.p, L * sizeof(int), hipMemcpyDeviceToDevice, s));
    if (!n_singular) return TDYN_OK;
    std::vector<int> h(L);
    HIPCHK(hipMemcpyAsync(h.data(), b->op_status.p, L * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int k = 0, first = -1;
    for (size_t i = 0; i < L; ++i) if (h[i] != 0) { if (first < 0) first = (int)i; ++k; }
    *n_singular = k;
    if (k) return fail(TDYN_E_SINGULAR, "articulated inertia D is singular for " + std::to_string(k) +
                                        " instance(s), the first being instance " + std::to_string(first));
    return TDYN_OK;
}

int tdyn_batch_info(const tdyn_batch* b, int64_t* info) {
    if (!b || !info) return fail(TDYN_E_ARG, "null argument");
    info[0] = b->n_inst; info[1] = b->total_q; info[2] = b->total_u; info[3] = b->total_b;
    info[4] = b->max_nq; info[5] = b->max_nu; info[6] = b->max_nb; info[7] = b->n_trees;
    return TDYN_OK;
}

int tdyn_batch_offsets(const tdyn_batch* b, int32_t* q_off, int32_t* u_off, int32_t* body_off) {
    if (!b) return fail(TDYN_E_ARG, "null batch");
    for (int i = 0; i < b->n_inst; ++i) {
        if (q_off) q_off[i] = b->h_qoff[i];
        if (u_off) u_off[i] = b->h_uoff[i];
        if (body_off) body_off[i] = b->h_boff[i];
    }
    return TDYN_OK;
}

int tdyn_realize_kinematics(tdyn_batch* b, const double* q, const double* u, double* X_GB,
                            double* V_GB, double* qdot, double* energies, hipStream_t s) {
    if (!b || !q || !u) return fail(TDYN_E_ARG, "tdyn_realize_kinematics: batch, q and u are required");
    OpArgs a = b->op();
    a.q = q; a.u = u; a.out_body = X_GB; a.out_body2 = V_GB; a.out_q = qdot; a.out_scalar = energies;
    HIPCHK(launch_op(OP_KIN, a, s));
    return TDYN_OK;
}

int tdyn_realize_acceleration(tdyn_batch* b, const double* q, const double* u, const double* fapp,
                              double* udot, double* qdot, hipStream_t s) {
    if (!b || !q || !u || !udot) return fail(TDYN_E_ARG, "tdyn_realize_acceleration: batch, q, u and udot are required");
    if (b->any_cons) {
        // with constraints, realize(Acceleration) is the loop forward dynamics
        { const int rc = ensure_constraints(b); if (rc) return rc; }
        OpArgs a = b->op();
        a.q = q; a.u = u; a.fbody = fapp; a.out_u = udot; a.out_q = qdot;
        HIPCHK(launch_op(OP_FWD_CONS, a, s));
        return TDYN_OK;
    }
    OpArgs a = b->op();
    a.q = q; a.u = u; a.fbody = fapp; a.out_u = udot; a.out_q = qdot;
    HIPCHK(launch_op(OP_FWD_SYS, a, s));
    return TDYN_OK;
}

int tdyn_calc_acceleration_ignoring_constraints(tdyn_batch* b, const double* q, const double* u,
                                                const double* fmob, const double* F, double* udot,
                                                double* A_GB, double* qdotdot, hipStream_t s) {
    // The reference requires full-length fMob and F (SimbodyMatterSubsystem.cpp:205-216).
    if (!b || !q || !u || !udot) return fail(TDYN_E_ARG, "calcAccelerationIgnoringConstraints: batch, q, u and udot are required");
    if (!fmob || !F) return fail(TDYN_E_SIZE, "calcAccelerationIgnoringConstraints: appliedMobilityForces and appliedBodyForces must be full length");
    OpArgs a = b->op();
    a.q = q; a.u = u; a.fmob = fmob; a.fbody = F; a.out_u = udot; a.out_body = A_GB; a.out_q = qdotdot;
    HIPCHK(launch_op(OP_FWD_GIVEN, a, s));
    return TDYN_OK;
}

int tdyn_calc_acceleration_with_known_udot(tdyn_batch* b, const double* q, const double* u,
                                           const double* fmob, const double* F, const int32_t* known,
                                           const double* udot_known, double* udot, double* tau,
                                           double* A_GB, hipStream_t s) {
    if (!b || !q || !u || !udot || !tau) return fail(TDYN_E_ARG, "calcAccelerationWithKnownUDot: batch, q, u, udot and tau are required");
    if (!known || !udot_known) return fail(TDYN_E_ARG, "calcAccelerationWithKnownUDot: known and udot_known are required");
    OpArgs a = b->op();
    a.q = q; a.u = u; a.fmob = fmob; a.fbody = F; a.presc = known; a.udot_in = udot_known;
    a.out_u = udot; a.out_tau = tau; a.out_body = A_GB;
    HIPCHK(launch_op(OP_FWD_PRES, a, s));
    return TDYN_OK;
}

int tdyn_multiply_by_m_inv(tdyn_batch* b, const double* q, const double* v, double* MinvV, hipStream_t s) {
    if (!b || !q) return fail(TDYN_E_ARG, "multiplyByMInv: batch and q are required");
    if (!v || !MinvV) return fail(TDYN_E_SIZE, "multiplyByMInv: argument 'v' must have one entry per mobility");
    OpArgs a = b->op();
    a.q = q; a.fmob = v; a.out_u = MinvV;
    HIPCHK(launch_op(OP_MINV, a, s));
    return TDYN_OK;
}

int tdyn_calc_residual_force_ignoring_constraints(tdyn_batch* b, const double* q, const double* u,
                                                  const double* fmob, const double* F,
                                                  const double* udot, double* resid, hipStream_t s) {
    if (!b || !q || !u || !resid) return fail(TDYN_E_ARG, "calcResidualForceIgnoringConstraints: batch, q, u and resid are required");
    OpArgs a = b->op();
    a.q = q; a.u = u; a.fmob = fmob; a.fbody = F; a.udot_in = udot; a.out_u = resid;
    HIPCHK(launch_op(OP_RESID, a, s));
    return TDYN_OK;
}

int tdyn_calc_articulated_body_inertias(tdyn_batch* b, const double* q, double* P, double* PPlus, hipStream_t s) {
    if (!b || !q || !P) return fail(TDYN_E_ARG, "realizeArticulatedBodyInertias: batch, q and P are required");
    OpArgs a = b->op();
    a.q = q; a.out_abi = P; a.out_abi2 = PPlus;
    HIPCHK(launch_op(OP_ABI, a, s));
    return TDYN_OK;
}

int tdyn_multiply_by_m(tdyn_batch* b, const double* q, const double* v, double* Mv, hipStream_t s) {
    if (!b || !q || !v || !Mv) return fail(TDYN_E_ARG, "multiplyByM: batch, q, v and Mv are required");
    OpArgs a = b->op();
    a.q = q; a.udot_in = v; a.out_u = Mv;
    HIPCHK(launch_op(OP_MULM, a, s));
    return TDYN_OK;
}

int tdyn_multiply_by_system_jacobian(tdyn_batch* b, const double* q, const double* v, double* Jv, hipStream_t s) {
    if (!b || !q || !v || !Jv) return fail(TDYN_E_ARG, "multiplyBySystemJacobian: batch, q, v and Jv are required");
    OpArgs a = b->op();
    a.q = q; a.udot_in = v; a.out_body = Jv;
    HIPCHK(launch_op(OP_JAC, a, s));
    return TDYN_OK;
}

int tdyn_multiply_by_system_jacobian_transpose(tdyn_batch* b, const double* q, const double* F_G, double* f,
                                               hipStream_t s) {
    if (!b || !q || !F_G || !f)
        return fail(TDYN_E_ARG, "multiplyBySystemJacobianTranspose: batch, q, F_G and f are required");
    OpArgs a = b->op();
    a.q = q; a.fbody = F_G; a.out_u = f;
    HIPCHK(launch_op(OP_JACT, a, s));
    return TDYN_OK;
}

int tdyn_calc_body_acceleration_from_udot(tdyn_batch* b, const double* q, const double* u, const double* udot,
                                          double* A_GB, hipStream_t s) {
    if (!b || !q || !u || !A_GB)
        return fail(TDYN_E_ARG, "calcBodyAccelerationFromUDot: batch, q, u and A_GB are required");
    OpArgs a = b->op();
    a.q = q; a.u = u; a.udot_in = udot; a.out_body = A_GB;
    HIPCHK(launch_op(OP_BACC, a, s));
    return TDYN_OK;
}

int tdyn_calc_composite_body_inertias(tdyn_batch* b, const double* q, double* R, hipStream_t s) {
    if (!b || !q || !R) return fail(TDYN_E_ARG, "calcCompositeBodyInertias: batch, q and R are required");
    OpArgs a = b->op();
    a.q = q; a.out_body = R;
    HIPCHK(launch_op(OP_CBI, a, s));
    return TDYN_OK;
}

int tdyn_calc_mobilizer_reaction_forces(tdyn_batch* b, const double* q, const double* u,
                                        const double* fmob, const double* F, double* FM, hipStream_t s) {
    if (!b || !q || !u || !FM) return fail(TDYN_E_ARG, "calcMobilizerReactionForces: batch, q, u and FM are required");
    OpArgs a = b->op();
    a.q = q; a.u = u; a.fmob = fmob; a.fbody = F; a.out_body = FM;
    HIPCHK(launch_op(OP_REACT, a, s));
    return TDYN_OK;
}

// tdyn_rollout_begin; init false (tdyn_rollout_restore): the state is taken
// as it is, without initialize()'s projection and prescription (the caller
// then sets every lane's t, h, status and statistics).
static int rollout_begin_impl(tdyn_batch* b, const double* q, const double* u, const double* fapp,
                              double t0, const double* t_end, double t_end_all,
                              const tdyn_integrator_cfg* cfg, bool init, hipStream_t s);

int tdyn_rollout_begin(tdyn_batch* b, const double* q, const double* u, const double* fapp,
                       double t0, const double* t_end, double t_end_all,
                       const tdyn_integrator_cfg* cfg, hipStream_t s) {
    return rollout_begin_impl(b, q, u, fapp, t0, t_end, t_end_all, cfg, true, s);
}

static int rollout_begin_impl(tdyn_batch* b, const double* q, const double* u, const double* fapp,
                              double t0, const double* t_end, double t_end_all,
                              const tdyn_integrator_cfg* cfg, bool init, hipStream_t s) {
    if (!b || !q || !u || !cfg) return fail(TDYN_E_ARG, "tdyn_rollout_begin: batch, q, u and cfg are required");
    if (!t_end && !(t_end_all > t0)) return fail(TDYN_E_ARG, "tdyn_rollout_begin: final time must exceed t0");
    // Every check runs on locals first: a refused begin leaves the rollout in
    // progress (settings, layout, state) as it was.
    RkmCfg c;
    int prec = 0;
    { const int rc = make_rkm_cfg(cfg, c, prec); if (rc) return rc; }
    if (prec != 0 && !(b->is_chain && b->evaluator != 1))
        return fail(TDYN_E_ARG, "tdyn_rollout_begin: precision 1-3 runs on the chain evaluator (one serial-chain model per batch)");
    if ((c.inf_norm || c.force_proj) && b->evaluator == 2)
        return fail(TDYN_E_ARG, "tdyn_rollout_begin: the infinity norm and projection on every step run on the tree "
                                "evaluators (evaluator 0, 1 or 3), not the chain evaluator");
    if (b->known_p && (b->evaluator == 2 || (b->evaluator == 3 && b->tree_lanes == 1) || prec != 0))
        return fail(TDYN_E_ARG, "tdyn_rollout_begin: known udot runs on the tree sweeps or the level-parallel "
                                "tree evaluator (not one lane per instance), in FP64");
    if (b->motion && (b->known_p || b->evaluator == 2 || b->evaluator == 3 || prec != 0))
        return fail(TDYN_E_ARG, "tdyn_rollout_begin: prescribed motion runs on the tree sweeps (evaluator 0 or 1), "
                                "in FP64, without known udot");
    if (b->any_cons && (b->known_p || b->motion || b->evaluator == 2 || b->evaluator == 3 || prec != 0))
        return fail(TDYN_E_ARG, "tdyn_rollout_begin: constraints run on the tree sweeps (evaluator 0 or 1), in FP64, "
                                "without known udot or prescribed motion");
    // Rollouts with constraints of a batch of different trees (padded
    // records) are switched off: their one GPU run faulted, and the cause is
    // not found yet.  Batches of one tree run them.
    if (b->any_cons && !b->uniform)
        return fail(TDYN_E_ARG, "tdyn_rollout_begin: constraints in rollouts need a batch of one tree (a batch of "
                                "different trees takes tdyn_calc_constrained_acceleration only)");
    if (b->any_wit && (b->evaluator == 2 || b->evaluator == 3 || prec != 0))
        return fail(TDYN_E_ARG, "tdyn_rollout_begin: event witnesses run on the tree sweeps (evaluator 0 or 1), in "
                                "FP64");
    { const int rc = ensure_constraints(b); if (rc) return rc; }
    { const int rc = ensure_witnesses(b); if (rc) return rc; }
    const size_t L = b->n_inst;
    const bool aos = !b->motion && !b->any_cons && !b->any_wit && b->stream_ok && (b->evaluator == 0 || b->evaluator >= 3) &&
                     !(b->known_p && b->tree_lanes == 1);
    const int grp = aos && b->tree_lanes != 1 ? 1 : 0;
    // From here on the batch changes: a failure (allocation, upload) ends the
    // rollout in progress.
    b->rollout_ready = false;
    if (grp) {
        const int rc = build_tree_plan(b, s);
        if (rc) return rc;
    }
    { const int rc = ensure_rollout_buffers(b, prec); if (rc) return rc; }
    if (b->any_wit) {
        const size_t E = L * (size_t)b->ev_cap * 2;
        if (b->ev_ws.n != (size_t)tdyn::EW_N * L) HIPCHK(b->ev_ws.alloc((size_t)tdyn::EW_N * L));
        if (b->ev_win.n != E) HIPCHK(b->ev_win.alloc(E));
        if (b->ev_ids.n != E) HIPCHK(b->ev_ids.alloc(E));
        if (b->ev_count.n != L) HIPCHK(b->ev_count.alloc(L));
        HIPCHK(hipMemsetAsync(b->ev_count.p, 0, L * sizeof(int), s));
    }
    if (fapp) {
        if (b->fapp.n != (size_t)b->total_b * 6) HIPCHK(b->fapp.alloc((size_t)b->total_b * 6));
        HIPCHK(hipMemcpyAsync(b->fapp.p, fapp, (size_t)b->total_b * 6 * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        b->fapp.free();
    }
    if (t_end) {
        if (b->permuted) HIPCHK(launch_lane_gather(b->inst.p, (int)L, t_end, b->t_end.p, s));
        else HIPCHK(hipMemcpyAsync(b->t_end.p, t_end, L * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        std::vector<double> te(L, t_end_all);
        HIPCHK(hipMemcpyAsync(b->t_end.p, te.data(), L * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    // commit: the rollout keeps these settings until the next begin
    b->cfg = c;
    b->precision = prec;
    b->rollout_aos = aos;
    b->rollout_grp = grp;
    b->roll_known_p = b->known_p;
    b->roll_known_u = b->known_u;
    b->roll_motion = b->motion;
    b->roll_evaluator = b->evaluator;
    RkmArgs a = b->rk();
    HIPCHK(launch_pack_state(a, q, u, 1, nullptr, nullptr, s));
    if (init) {
        HIPCHK(launch_rkm_init(a, t0, s));
        HIPCHK(launch_cons_init(a, s));
    }
    b->rollout_ready = true;
    return TDYN_OK;
}

// ---- rollout snapshots -------------------------------------------------
extern "C++" {
namespace {
constexpr char kSnapMagic[8] = {'T', 'D', 'Y', 'N', 'S', 'N', 'P', '1'};
constexpr int kSnapHeader = 112;
// FNV-1a over the batch's model: tree and body tables, constants, gravity
// and the tree of every instance (a snapshot resumes only on the same model).
uint64_t model_hash(tdyn_batch* b, int& rc) {
    rc = 0;
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) { h ^= c[i]; h *= 1099511628211ull; }
    };
    mix(&b->n_inst, sizeof b->n_inst); mix(&b->n_trees, sizeof b->n_trees);
    mix(b->h_tree_of.data(), b->h_tree_of.size() * sizeof(int));
    mix(b->h_tree_a.data(), b->h_tree_a.size() * sizeof(int4));
    mix(b->h_body_a.data(), b->h_body_a.size() * sizeof(int4));
    std::vector<double> prm(b->prm.n), grav(b->gravity.n);
    if ((!prm.empty() && hipMemcpy(prm.data(), b->prm.p, prm.size() * sizeof(double), hipMemcpyDeviceToHost)) ||
        (!grav.empty() && hipMemcpy(grav.data(), b->gravity.p, grav.size() * sizeof(double), hipMemcpyDeviceToHost)))
        rc = TDYN_E_HIP;
    mix(prm.data(), prm.size() * sizeof(double)); mix(grav.data(), grav.size() * sizeof(double));
    return h;
}
int64_t snapshot_bytes(const tdyn_batch* b, int precision) {
    const int64_t n = b->n_inst, ny = b->total_q + b->total_u;
    return kSnapHeader + n * (3 * 8 + 5 * 4) + 8 * ny * (precision == 1 ? 2 : 1);
}
template <class T> void put(std::vector<unsigned char>& o, size_t at, const T& v) { std::memcpy(o.data() + at, &v, sizeof v); }
template <class T> T get(const unsigned char* p, size_t at) { T v; std::memcpy(&v, p + at, sizeof v); return v; }
} // namespace
} // extern "C++"

int tdyn_rollout_snapshot(tdyn_batch* b, void* out, int64_t capacity, int64_t* bytes, hipStream_t s) {
    if (!b || !b->rollout_ready) return fail(TDYN_E_STAGE, "tdyn_rollout_snapshot: no rollout");
    const int64_t need = snapshot_bytes(b, b->precision);
    if (bytes) *bytes = need;
    if (!out) return TDYN_OK;
    if (capacity < need) return fail(TDYN_E_SIZE, "tdyn_rollout_snapshot: the buffer is too small");
    const size_t L = b->n_inst, TQ = b->total_q, TU = b->total_u;
    int rc = 0;
    const uint64_t mh = model_hash(b, rc);
    if (rc) return fail(rc, "tdyn_rollout_snapshot: model tables copy failed");
    // per-lane arrays to host, then to instance order
    std::vector<double> t(L), h(L), te(L);
    std::vector<int> st(L), sx(4 * L), inst(L);
    HIPCHK(hipMemcpyAsync(t.data(), b->t.p, L * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(h.data(), b->h.p, L * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(te.data(), b->t_end.p, L * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(st.data(), b->status.p, L * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(sx.data(), b->stats.p, 4 * L * sizeof(int), hipMemcpyDeviceToHost, s));
    if (b->permuted) HIPCHK(hipMemcpyAsync(inst.data(), b->inst.p, L * sizeof(int), hipMemcpyDeviceToHost, s));
    else for (size_t l = 0; l < L; ++l) inst[l] = (int)l;
    // packed q, u (and the compensation rows of the mixed path), instance order
    const int nrow = b->precision == 1 ? 2 : 1;
    DBuf<double> dq, du;
    HIPCHK(dq.alloc(TQ * nrow)); HIPCHK(du.alloc(TU * nrow));
    RkmArgs a = b->rk();
    HIPCHK(launch_pack_state(a, nullptr, nullptr, 0, dq.p, du.p, s, VY));
    if (nrow == 2) HIPCHK(launch_pack_state(a, nullptr, nullptr, 0, dq.p + TQ, du.p + TU, s, VC0));
    std::vector<double> hq(TQ * nrow), hu(TU * nrow);
    HIPCHK(hipMemcpyAsync(hq.data(), dq.p, hq.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hu.data(), du.p, hu.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<unsigned char> o((size_t)need, 0);
    std::memcpy(o.data(), kSnapMagic, 8);
    put<uint32_t>(o, 8, 1u); put<int32_t>(o, 12, (int32_t)L);
    put<int64_t>(o, 16, (int64_t)TQ); put<int64_t>(o, 24, (int64_t)TU); put<uint64_t>(o, 32, mh);
    put<int32_t>(o, 40, b->precision); put<int32_t>(o, 44, b->roll_evaluator);
    const RkmCfg& c = b->cfg;
    const double cd[7] = {c.accuracy, c.cons_tol, c.fixed_h, c.init_h, c.min_h, c.max_h, c.timescale};
    for (int k = 0; k < 7; ++k) put<double>(o, 48 + 8*k, cd[k]);
    put<int32_t>(o, 104, c.max_attempts);
    size_t at = kSnapHeader;
    std::vector<double> ti(L), hi(L), tei(L);
    std::vector<int> sti(L), sxi(4 * L);
    for (size_t l = 0; l < L; ++l) {
        const int i = inst[l];
        ti[i] = t[l]; hi[i] = h[l]; tei[i] = te[l]; sti[i] = st[l];
        for (int k = 0; k < 4; ++k) sxi[4*i + k] = sx[4*l + k];
    }
    std::memcpy(o.data() + at, ti.data(), L * 8); at += L * 8;
    std::memcpy(o.data() + at, hi.data(), L * 8); at += L * 8;
    std::memcpy(o.data() + at, tei.data(), L * 8); at += L * 8;
    std::memcpy(o.data() + at, sti.data(), L * 4); at += L * 4;
    std::memcpy(o.data() + at, sxi.data(), 4 * L * 4); at += 4 * L * 4;
    std::memcpy(o.data() + at, hq.data(), TQ * 8); at += TQ * 8;
    std::memcpy(o.data() + at, hu.data(), TU * 8); at += TU * 8;
    if (nrow == 2) {
        std::memcpy(o.data() + at, hq.data() + TQ, TQ * 8); at += TQ * 8;
        std::memcpy(o.data() + at, hu.data() + TU, TU * 8); at += TU * 8;
    }
    std::memcpy(out, o.data(), o.size());
    return TDYN_OK;
}

int tdyn_rollout_restore(tdyn_batch* b, const void* in, int64_t bytes, const double* fapp, hipStream_t s) {
    if (!b || !in) return fail(TDYN_E_ARG, "tdyn_rollout_restore: batch and snapshot are required");
    const unsigned char* p = (const unsigned char*)in;
    if (bytes < kSnapHeader || std::memcmp(p, kSnapMagic, 8) != 0 || get<uint32_t>(p, 8) != 1u)
        return fail(TDYN_E_ARG, "tdyn_rollout_restore: not a snapshot of this format");
    const size_t L = b->n_inst, TQ = b->total_q, TU = b->total_u;
    const int prec = get<int32_t>(p, 40);
    if (get<int32_t>(p, 12) != (int32_t)L || get<int64_t>(p, 16) != (int64_t)TQ || get<int64_t>(p, 24) != (int64_t)TU ||
        prec < 0 || prec > 3 || bytes != snapshot_bytes(b, prec))
        return fail(TDYN_E_SIZE, "tdyn_rollout_restore: the snapshot's sizes are not this batch's");
    int rc = 0;
    const uint64_t mh = model_hash(b, rc);
    if (rc) return fail(rc, "tdyn_rollout_restore: model tables copy failed");
    if (mh != get<uint64_t>(p, 32)) return fail(TDYN_E_ARG, "tdyn_rollout_restore: the snapshot is of another model");
    tdyn_integrator_cfg cfg;
    std::memset(&cfg, 0, sizeof cfg);
    const double acc = get<double>(p, 48), ct = get<double>(p, 56), fh = get<double>(p, 64), ih = get<double>(p, 72),
                 mnh = get<double>(p, 80), mxh = get<double>(p, 88), ts = get<double>(p, 96);
    cfg.accuracy = acc; cfg.constraint_tolerance = ct; cfg.fixed_step = fh > 0 ? fh : 0.0;
    cfg.initial_step = ih > 0 ? ih : 0.0; cfg.min_step = mnh > 0 ? mnh : 0.0; cfg.max_step = mxh > 0 ? mxh : 0.0;
    cfg.time_scale = ts; cfg.max_attempts = get<int32_t>(p, 104); cfg.precision = prec;
    size_t at = kSnapHeader;
    std::vector<double> ti(L), hi(L), tei(L);
    std::vector<int> sti(L), sxi(4 * L);
    std::memcpy(ti.data(), p + at, L * 8); at += L * 8;
    std::memcpy(hi.data(), p + at, L * 8); at += L * 8;
    std::memcpy(tei.data(), p + at, L * 8); at += L * 8;
    std::memcpy(sti.data(), p + at, L * 4); at += L * 4;
    std::memcpy(sxi.data(), p + at, 4 * L * 4); at += 4 * L * 4;
    const size_t q_at = at;
    // the state rows go through the device (packed layout, instance order)
    const int nrow = prec == 1 ? 2 : 1;
    DBuf<double> dq, du, dte;
    HIPCHK(dq.alloc(TQ * nrow)); HIPCHK(du.alloc(TU * nrow)); HIPCHK(dte.alloc(L));
    HIPCHK(hipMemcpy(dq.p, p + q_at, TQ * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(du.p, p + q_at + TQ * 8, TU * 8, hipMemcpyHostToDevice));
    if (nrow == 2) {
        HIPCHK(hipMemcpy(dq.p + TQ, p + q_at + (TQ + TU) * 8, TQ * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(du.p + TU, p + q_at + (2*TQ + TU) * 8, TU * 8, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(dte.p, tei.data(), L * 8, hipMemcpyHostToDevice));
    if ((rc = rollout_begin_impl(b, dq.p, du.p, fapp, 0.0, dte.p, 0.0, &cfg, false, s))) return rc;
    b->rollout_ready = false;
    RkmArgs a = b->rk();
    if (nrow == 2) HIPCHK(launch_pack_state(a, dq.p + TQ, du.p + TU, 1, nullptr, nullptr, s, VC0));
    // per-instance t, h, status, statistics to lane order
    std::vector<int> inst(L);
    if (b->permuted) HIPCHK(hipMemcpy(inst.data(), b->inst.p, L * sizeof(int), hipMemcpyDeviceToHost));
    else for (size_t l = 0; l < L; ++l) inst[l] = (int)l;
    std::vector<double> t(L), h(L);
    std::vector<int> st(L), sx(4 * L);
    for (size_t l = 0; l < L; ++l) {
        const int i = inst[l];
        t[l] = ti[i]; h[l] = hi[i]; st[l] = sti[i];
        for (int k = 0; k < 4; ++k) sx[4*l + k] = sxi[4*i + k];
    }
    HIPCHK(hipMemcpyAsync(b->t.p, t.data(), L * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->h.p, h.data(), L * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->status.p, st.data(), L * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->stats.p, sx.data(), 4 * L * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    b->rollout_ready = true;
    return TDYN_OK;
}

int tdyn_rollout_advance(tdyn_batch* b, int max_steps, hipStream_t s) {
    if (!b || !b->rollout_ready) return fail(TDYN_E_STAGE, "tdyn_rollout_advance: call tdyn_rollout_begin first");
    if (max_steps <= 0) return fail(TDYN_E_ARG, "max_steps must be positive");
    RkmArgs a = b->rk();
    a.cfg.max_steps = max_steps;
    HIPCHK(launch_rkm(a, s));
    return TDYN_OK;
}

int tdyn_rollout_get(tdyn_batch* b, double* q, double* u, double* t, int32_t* status, int32_t* stats, hipStream_t s) {
    if (!b || !b->rollout_ready) return fail(TDYN_E_STAGE, "tdyn_rollout_get: no rollout");
    RkmArgs a = b->rk();
    if (q && u) HIPCHK(launch_pack_state(a, nullptr, nullptr, 0, q, u, s));
    else if (q || u) return fail(TDYN_E_ARG, "tdyn_rollout_get: q and u go together");
    const size_t L = b->n_inst;
    if (b->permuted) {
        HIPCHK(launch_lane_scatter(b->inst.p, (int)L, b->t.p, t, b->status.p, status, b->stats.p, stats, s));
    } else {
        if (t) HIPCHK(hipMemcpyAsync(t, b->t.p, L * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (status) HIPCHK(hipMemcpyAsync(status, b->status.p, L * sizeof(int), hipMemcpyDeviceToDevice, s));
        if (stats) HIPCHK(hipMemcpyAsync(stats, b->stats.p, 4 * L * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
    return TDYN_OK;
}

int tdyn_rollout_derivs(tdyn_batch* b, double* qdot, double* udot, hipStream_t s) {
    if (!b || !b->rollout_ready) return fail(TDYN_E_STAGE, "tdyn_rollout_derivs: no rollout");
    if (!qdot || !udot) return fail(TDYN_E_ARG, "tdyn_rollout_derivs: qdot and udot are required");
    RkmArgs a = b->rk();
    HIPCHK(launch_rollout_f0(a, s));
    HIPCHK(launch_pack_state(a, nullptr, nullptr, 0, qdot, udot, s, 2 /* VF0 */));
    return TDYN_OK;
}

int tdyn_rollout_extend(tdyn_batch* b, const double* t_end, double t_end_all, hipStream_t s) {
    if (!b || !b->rollout_ready) return fail(TDYN_E_STAGE, "tdyn_rollout_extend: call tdyn_rollout_begin first");
    if (!t_end && !std::isfinite(t_end_all)) return fail(TDYN_E_ARG, "tdyn_rollout_extend: final time must be finite");
    const int L = b->n_inst;
    const double* te = nullptr;
    if (t_end) {
        // the new times go through t_end's own buffer (lane order), then the
        // kernel reopens the lanes that finished before them
        if (b->permuted) HIPCHK(launch_lane_gather(b->inst.p, L, t_end, b->t_end.p, s));
        else HIPCHK(hipMemcpyAsync(b->t_end.p, t_end, (size_t)L * sizeof(double), hipMemcpyDeviceToDevice, s));
        te = b->t_end.p;
    }
    HIPCHK(launch_rollout_extend(L, te, t_end_all, b->t.p, b->t_end.p, b->status.p, s));
    return TDYN_OK;
}

int tdyn_rollout_reports(tdyn_batch* b, double* q, double* u, const double* fapp, double t0, double t_end,
                         const double* report_times, int n_reports, const tdyn_integrator_cfg* cfg,
                         double* q_reports, double* u_reports, int32_t* stats, hipStream_t s) {
    if (!b || !q || !u || !cfg || n_reports < 0 || (n_reports > 0 && (!report_times || !q_reports || !u_reports)))
        return fail(TDYN_E_ARG, "tdyn_rollout_reports: batch, q, u, cfg and (with reports) times and outputs are required");
    if (b->any_cons)
        return fail(TDYN_E_ARG, "tdyn_rollout_reports: not with constraints (interpolated states would need "
                                "their own projection)");
    if (b->any_wit)
        return fail(TDYN_E_ARG, "tdyn_rollout_reports: not with event witnesses (a report time inside an event's "
                                "window changes the localization)");
    std::vector<double> rep(report_times, report_times + n_reports);
    for (int k = 0; k < n_reports; ++k)
        if (!std::isfinite(rep[k]) || rep[k] < t0 || rep[k] > t_end || (k && !(rep[k] > rep[k-1])))
            return fail(TDYN_E_ARG, "tdyn_rollout_reports: report times must increase within [t0, t_end]");
    int rc = tdyn_rollout_begin(b, q, u, fapp, t0, nullptr, t_end, cfg, s);
    if (rc) return rc;
    const int L = b->n_inst;
    const size_t TQ = (size_t)b->total_q, TU = (size_t)b->total_u;
    // the two ends of the current step: packed q, u, qdot, udot and lane times
    DBuf<double> ends[2][4], tl[2], rt;
    DBuf<int> next_k;
    for (int e = 0; e < 2; ++e) {
        HIPCHK(ends[e][0].alloc(TQ)); HIPCHK(ends[e][1].alloc(TU));
        HIPCHK(ends[e][2].alloc(TQ)); HIPCHK(ends[e][3].alloc(TU)); HIPCHK(tl[e].alloc(L));
    }
    HIPCHK(rt.upload(rep.empty() ? std::vector<double>{0.0} : rep));
    int cur = 0;
    if ((rc = tdyn_rollout_get(b, ends[0][0].p, ends[0][1].p, nullptr, nullptr, nullptr, s)) ||
        (rc = tdyn_rollout_derivs(b, ends[0][2].p, ends[0][3].p, s)))
        return rc;
    HIPCHK(hipMemcpyAsync(tl[0].p, b->t.p, (size_t)L * sizeof(double), hipMemcpyDeviceToDevice, s));
    int k0 = 0;                                    // reports at t0: the initial state
    for (; k0 < n_reports && rep[k0] <= t0; ++k0) {
        HIPCHK(hipMemcpyAsync(q_reports + k0 * TQ, ends[0][0].p, TQ * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(u_reports + k0 * TU, ends[0][1].p, TU * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    HIPCHK(next_k.upload(std::vector<int>((size_t)L, k0)));
    const DevTables T = b->tables();
    std::vector<int> st((size_t)L);
    for (long step = 0;; ++step) {
        if (step > 100000000L) return fail(TDYN_E_STEP_FAILED, "rollout did not finish");
        if ((rc = tdyn_rollout_advance(b, 1, s))) return rc;
        const int nx = cur ^ 1;
        if ((rc = tdyn_rollout_get(b, ends[nx][0].p, ends[nx][1].p, nullptr, nullptr, nullptr, s)) ||
            (rc = tdyn_rollout_derivs(b, ends[nx][2].p, ends[nx][3].p, s)))
            return rc;
        if (n_reports > k0) {
            const double* prev[4] = {ends[cur][0].p, ends[cur][1].p, ends[cur][2].p, ends[cur][3].p};
            const double* now[4] = {ends[nx][0].p, ends[nx][1].p, ends[nx][2].p, ends[nx][3].p};
            HIPCHK(launch_report_interp(T, b->uniform ? nullptr : b->tree_of.p, b->q_off.p, b->u_off.p, L, rt.p,
                                        n_reports, next_k.p, tl[cur].p, b->t.p, prev, now, b->cfg.cons_tol,
                                        proj_opts(b->cfg) | (b->cfg.proj_interp ? 0 : PJ_NONE),
                                        q_reports, u_reports, (long)TQ, (long)TU, s));
        }
        HIPCHK(hipMemcpyAsync(tl[nx].p, b->t.p, (size_t)L * sizeof(double), hipMemcpyDeviceToDevice, s));
        cur = nx;
        HIPCHK(hipMemcpyAsync(st.data(), b->status.p, (size_t)L * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        bool running = false, failed = false;
        for (int x : st) { running |= x == 0; failed |= x < 0; }
        if (!running) {
            if ((rc = tdyn_rollout_get(b, q, u, nullptr, nullptr, stats, s))) return rc;
            HIPCHK(hipStreamSynchronize(s));
            return failed ? fail(TDYN_E_STEP_FAILED, "Integrator::StepFailed for at least one instance") : TDYN_OK;
        }
    }
}

int tdyn_rollout(tdyn_batch* b, double* q, double* u, const double* fapp, double t0, double t_end,
                 const tdyn_integrator_cfg* cfg, int32_t* stats, hipStream_t s) {
    int rc = tdyn_rollout_begin(b, q, u, fapp, t0, nullptr, t_end, cfg, s);
    if (rc) return rc;
    const size_t L = b->n_inst;
    std::vector<int> st(L);
    for (int round = 0; round < 1000000; ++round) {
        rc = tdyn_rollout_advance(b, 1 << 20, s);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(st.data(), b->status.p, L * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        bool running = false, failed = false;
        for (size_t i = 0; i < L; ++i) { running |= st[i] == 0; failed |= st[i] < 0; }
        if (!running) {
            rc = tdyn_rollout_get(b, q, u, nullptr, nullptr, stats, s);
            if (rc) return rc;
            HIPCHK(hipStreamSynchronize(s));
            return failed ? fail(TDYN_E_STEP_FAILED, "Integrator::StepFailed for at least one instance") : TDYN_OK;
        }
    }
    return fail(TDYN_E_STEP_FAILED, "rollout did not finish");
}

} // extern "C"
