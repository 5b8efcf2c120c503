// Host-side state shared by the C ABI implementation files: the batch
// handle, device buffers, error reporting and model packing.
#pragma once
#include "../../include/tdyn.h"
#include "tdyn_internal.hpp"

#include <cstring>
#include <string>
#include <vector>

namespace tdyn {
namespace host {
std::string& err_slot();     // this thread's last error message
inline int fail(int code, const std::string& msg) { err_slot() = msg; return code; }
inline int hipfail(hipError_t e, const char* what) {
    err_slot() = std::string(what) + ": " + hipGetErrorString(e);
    return TDYN_E_HIP;
}
#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return ::tdyn::host::hipfail(e_, #x); } while (0)

template <class T> struct DBuf {
    T* p = nullptr; size_t n = 0;
    hipError_t alloc(size_t count) {
        free();
        n = count;
        if (!count) return hipSuccess;
        return hipMalloc((void**)&p, count * sizeof(T));
    }
    hipError_t upload(const std::vector<T>& v) {
        hipError_t e = alloc(v.size());
        if (e != hipSuccess || v.empty()) return e;
        return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    }
    void free() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
    ~DBuf() { free(); }
};

// The PRM_STRIDE model constants of body b (1-based) of tree d: X_PF, X_MB =
// X_BM^-1, mass, mass center, unit inertia G = I/m about Bo, p_BM.
void pack_body_params(const tdyn_tree_desc& d, int b, double* P);
// True when every body's X_PF is the identity and X_BM has no rotation.
bool identity_frames(const tdyn_tree_desc& d);
// True when some body's mobilizer is reversed.
bool any_reversed(const tdyn_tree_desc& d);
// Validates one tree description (parents before children, known mobilizers,
// Free only on Ground, positive finite masses, finite frames).
int validate_tree(const tdyn_tree_desc& d, int t);
} // namespace host
} // namespace tdyn

struct tdyn_batch {
    using RkmCfg = tdyn::RkmCfg; using DevTables = tdyn::DevTables;
    using OpArgs = tdyn::OpArgs; using RkmArgs = tdyn::RkmArgs;
    int n_inst = 0, n_trees = 0;
    int64_t total_q = 0, total_u = 0, total_b = 0;
    int max_nq = 0, max_nu = 0, max_nb = 0, max_rec = 0;
    bool uniform = true;
    bool any_rev = false;        // some tree has a reversed mobilizer (the batch runs padded generic bodies)
    bool any_euler = false;      // some tree uses Euler angles for Ball and Free (likewise)
    bool any_bendstretch = false; // some tree has a BendStretch body
    bool any_bushing = false;     // some tree has a Bushing body
    const int* known_p = nullptr;     // tdyn_batch_set_known_udot: caller's device arrays (rollouts)
    const double* known_u = nullptr;
    // tdyn_batch_set_motion: the caller's device array, its per-body flags
    // and the per-instance prescribed udot the rollout kernels write
    const double* motion = nullptr;
    tdyn::host::DBuf<int> motion_flags;
    tdyn::host::DBuf<double> motion_udot;
    // what the current rollout was begun with: advance, get and derivs read
    // these, never the settings changed since tdyn_rollout_begin
    const int* roll_known_p = nullptr;
    const double* roll_known_u = nullptr;
    const double* roll_motion = nullptr;
    int roll_evaluator = 0;
    std::vector<int> h_qoff, h_uoff, h_boff, h_tree_of;
    std::vector<int4> h_tree_a, h_tree_b, h_body_a;
    // device tables
    tdyn::host::DBuf<int4> body_a, body_b, tree_a, tree_b, pre, bfs, bfs_b;
    tdyn::host::DBuf<int> lvl;
    tdyn::host::DBuf<int> child_idx, tree_of, q_off, u_off, b_off;   // per LANE
    tdyn::host::DBuf<int> inst;          // instance of each lane
    bool permuted = false;               // lanes are not in instance order
    tdyn::host::DBuf<double> prm, gravity, scratch, prm_bfs;
    tdyn::host::DBuf<int> op_status;     // per instance: status of the last operator that formed D
    // rollout state
    tdyn::host::DBuf<double> vec, uscale, t, h, t_end;
    tdyn::host::DBuf<int> status, stats;
    tdyn::host::DBuf<double> fapp;
    tdyn::host::DBuf<float> scratch32;   // mixed-precision hand-over rows
    int precision = 0;                   // integrator precision of the current rollout (0 FP64, 1 mixed)
    int ny_max = 0;
    bool rollout_ready = false;
    bool is_chain = false;       // uniform batch whose tree is a serial chain
    bool idf = false;            // tree 0's joint frames are identity rotations (R_PF = I, p_PF = 0, R_BM = I)
    bool idf_all = false;        // every tree's joint frames are
    int evaluator = 0;           // 0 auto, 1 tree sweeps, 2 chain (register-resident)
    bool stream_ok = false;      // different trees, Free only at body 1: the streaming tree evaluator applies
    int dmax = 0;                // deepest body of any tree
    int stream_stride = 0;       // scratch doubles per lane of the streaming tree evaluator
    bool rollout_aos = false;    // the current rollout keeps lane-contiguous vectors (streaming tree evaluator)
    int tree_lanes = 0;          // lanes per instance of the tree evaluators: 0 automatic, 1 streaming, 4-64 level-parallel
    int lanes_min = 0, lanes_max = 32;   // automatic lanes: range (lanes_min 0: from the batch size)
    int rollout_grp = 0;         // the current rollout runs the level-parallel evaluator (1) or not (0)
    int group_stride = 0;        // scratch doubles per instance of the level-parallel evaluator
    int n_simd = 1024;           // SIMDs of the device (4 per compute unit)
    std::vector<int> h_lane_tree, h_tree_maxw;   // tree of each lane; widest level of each tree
    std::vector<std::vector<int>> h_tree_lw;     // the width of every level of each tree
    tdyn::host::DBuf<int4> plan; tdyn::host::DBuf<int> plan_idx;   // wavefront plan of the level-parallel evaluator
    tdyn::host::DBuf<int> plan_ctr;              // its work queue's counter
    // holonomic constraints (tdyn_batch_set_constraints): per tree, and their
    // device tables and per-lane workspace (uploaded when changed)
    std::vector<std::vector<tdyn_constraint_desc>> h_cons;
    tdyn::host::DBuf<int4> cons_t, cons_a;
    tdyn::host::DBuf<double> cons_p, cons_ws;
    bool any_cons = false, cons_dirty = false;
    // event witnesses (tdyn_batch_set_witnesses): per tree, their device
    // tables (uploaded when changed), and the rollout's event workspace and
    // log (allocated by tdyn_rollout_begin)
    std::vector<std::vector<tdyn_witness_desc>> h_wit;
    tdyn::host::DBuf<int4> wit_t, wit_a;
    tdyn::host::DBuf<double> wit_p, ev_ws, ev_win;
    tdyn::host::DBuf<int> ev_ids, ev_count;
    int ev_cap = 64;
    bool any_wit = false, wit_dirty = false;
    int plan_waves = 0;
    RkmCfg cfg{};

    DevTables tables() const {
        DevTables T;
        T.body_a = body_a.p; T.body_b = body_b.p; T.child_idx = child_idx.p; T.prm = prm.p;
        T.tree_a = tree_a.p; T.tree_b = tree_b.p; T.gravity = gravity.p; T.pre = pre.p; T.bfs = bfs.p; T.lvl = lvl.p;
        T.bfs_b = bfs_b.p; T.prm_bfs = prm_bfs.p;
        T.cons_t = any_cons ? cons_t.p : nullptr; T.cons_a = cons_a.p; T.cons_p = cons_p.p;
        T.wit_t = any_wit ? wit_t.p : nullptr; T.wit_a = wit_a.p; T.wit_p = wit_p.p;
        return T;
    }
    OpArgs op() const {
        OpArgs a;
        std::memset(&a, 0, sizeof(a));
        a.T = tables(); a.tree_of = uniform ? nullptr : tree_of.p; a.n_inst = n_inst; a.L = n_inst;
        a.scratch = scratch.p; a.rec_stride = max_rec + 24; a.off = {q_off.p, u_off.p, b_off.p}; a.inst = permuted ? inst.p : nullptr;
        a.status = op_status.p;
        a.cons_ws = any_cons ? cons_ws.p : nullptr;
        return a;
    }
    RkmArgs rk() const {
        RkmArgs a;
        std::memset(&a, 0, sizeof(a));
        a.T = tables(); a.tree_of = uniform ? nullptr : tree_of.p; a.n_inst = n_inst; a.L = n_inst;
        a.scratch = scratch.p; a.rec_stride = max_rec + 24; a.off = {q_off.p, u_off.p, b_off.p};
        a.fapp = fapp.p; a.vec = vec.p; a.ny_max = ny_max; a.t = t.p; a.h = h.p; a.t_end = t_end.p;
        a.status = status.p; a.stats = stats.p; a.uscale = uscale.p; a.cfg = cfg; a.precision = precision;
        a.scratch32 = scratch32.p;
        a.chain = (!roll_known_p && !roll_motion && !any_cons && !any_wit && !cfg.inf_norm && !cfg.force_proj &&
                   (roll_evaluator == 2 || (roll_evaluator == 0 && is_chain))) ? 1 : 0;
        a.cons_ws = any_cons ? cons_ws.p : nullptr;
        if (any_wit) { a.ev_ws = ev_ws.p; a.ev_win = ev_win.p; a.ev_ids = ev_ids.p; a.ev_count = ev_count.p; }
        a.ev_cap = ev_cap;
        a.presc = roll_known_p; a.udot_known = roll_known_u;
        if (roll_motion) { a.motion = roll_motion; a.presc = motion_flags.p; a.udot_known = motion_udot.p; }
        a.idf = (a.chain ? idf : idf_all) ? 1 : 0;
        a.ext = (any_rev || any_euler || any_bendstretch || any_bushing || a.presc) ? 1 : 0;
        a.vec_aos = rollout_aos ? 1 : 0;
        a.stream_dmax = dmax;
        a.grp = rollout_aos ? rollout_grp : 0;
        a.plan = plan.p; a.plan_idx = plan_idx.p; a.plan_waves = plan_waves; a.plan_ctr = plan_ctr.p;
        if (rollout_aos) a.rec_stride = a.grp ? group_stride : stream_stride;
        return a;
    }
};


namespace tdyn {
namespace host {
// RkmCfg from the public integrator settings (defaults of
// IntegratorRep.h:737-740 and the System time scale).
int make_rkm_cfg(const tdyn_integrator_cfg* cfg, RkmCfg& c, int& precision);
// Allocates the device-resident integrator state of a batch (state vectors,
// per-lane time, step, status and statistics, and for precision 1 the
// mixed path's hand-over rows) for its current sizes.
int ensure_rollout_buffers(tdyn_batch* b, int precision);
// Uploads the constraint tables and allocates the per-lane workspace when
// the constraints changed.
int ensure_constraints(tdyn_batch* b);
// Uploads the witness tables when the witnesses changed.
int ensure_witnesses(tdyn_batch* b);
} // namespace host
} // namespace tdyn
