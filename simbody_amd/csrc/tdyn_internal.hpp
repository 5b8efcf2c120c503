// Internal layouts shared by the host packing code and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace tdyn {

// Model tables are read-only for the lifetime of a kernel: reading them
// through the constant address space lets wave-uniform loads become scalar
// (s_load) loads into SGPRs.
#define TD_CONST __attribute__((address_space(4)))
template <class T> __device__ __forceinline__ const TD_CONST T* cptr(const T* p) {
    return (const TD_CONST T*)p;
}
using cdptr = const TD_CONST double*;
// int4 table entry read through the constant address space.
__device__ __forceinline__ int4 ldi4(const int4* base, int idx) {
    const TD_CONST int* p = (const TD_CONST int*)(base + idx);
    return make_int4(p[0], p[1], p[2], p[3]);
}

enum { PIN = 0, SLIDER = 1, UNIVERSAL = 2, BALL = 3, FREE = 4, TRANSLATION = 5, PLANAR = 6, CYLINDER = 7, SCREW = 8,
       GIMBAL = 9, WELD = 10, BENDSTRETCH = 11, BUSHING = 12 };
// Kernel-side body kind: any non-Free mobilizer carried as a padded 3-dof body.
enum { GEN = 15 };
__host__ __device__ inline int mob_dof(int t) {
    switch (t) {
    case PIN: case SLIDER: case SCREW: return 1;
    case UNIVERSAL: case CYLINDER: case BENDSTRETCH: return 2;
    case BALL: case TRANSLATION: case PLANAR: case GIMBAL: return 3;
    case WELD: return 0;
    default: return 6;
    }
}
__host__ __device__ inline int mob_nq(int t) {
    switch (t) {
    case PIN: case SLIDER: case SCREW: return 1;
    case UNIVERSAL: case CYLINDER: case BENDSTRETCH: return 2;
    case TRANSLATION: case PLANAR: case GIMBAL: return 3;
    case BALL: return 4;
    case WELD: return 0;
    case BUSHING: return 6;
    default: return 7;
    }
}
// Translation, Planar, Cylinder, Screw, Gimbal, Weld and BendStretch run only
// as padded generic bodies; the typed code paths cover Pin, Slider,
// Universal, Ball and Free.  A Weld body (no mobilities) has all three padded
// columns zero.  BendStretch (a linear HDot_FM term) runs in the builds for
// extended models only (PrmTraits<>::ext, RkmArgs::ext).
__host__ __device__ inline bool mob_generic_only(int t) { return t >= TRANSLATION; }

// Per-body constants, PRM_STRIDE doubles per body (packed over all trees).
enum {
    PRM_R_PF = 0, PRM_P_PF = 9, PRM_R_MB = 12, PRM_P_MB = 21, PRM_MASS = 24,
    PRM_COM = 25, PRM_G = 28 /* unit inertia about Bo in B: xx yy zz xy xz yz */,
    PRM_P_BM = 34, PRM_PITCH = 37 /* Screw */, PRM_REV = 38 /* 1: reversed mobilizer */,
    PRM_EULER = 39 /* 1: Ball/Free q are x-y-z angles; 2: a Bushing (a Free body whose q are x-y-z
                       angles and p_FM, u = qdot, H_FM the Gimbal's angle columns) */, PRM_STRIDE = 40
};

// Per-body record in the per-lane scratch (doubles, relative to rec_off).
// Stored structure-of-arrays across lanes: scratch[(rec_off + k) * L + lane].
enum {
    REC_R = 0, REC_P = 9, REC_V = 12, REC_PHI = 18, REC_MKP = 21, REC_MKG = 24,
    REC_A = 30, REC_B = 36, REC_PPLUS = 42, REC_ZPLUS = 63, REC_AGB = 69,
    REC_DIEPS = 75 /* dof doubles, then H (6*dof) then G (6*dof) */
};
__host__ __device__ inline int rec_size(int dof) { return REC_DIEPS + 13 * dof; }

// Holonomic constraints: rows of one tree at most (the public limit
// TDYN_MAX_CONSTRAINT_ROWS), their kinds, and the per-lane workspace
// (doubles, lane-strided): W and a factor of it (CONS_MAXR^2 each), then the
// residual, solution and multiplier vectors (CONS_MAXR each).
constexpr int CONS_MAXR = 12;
enum { CONS_BALL = 0, CONS_ROD = 1 };
enum { CWS_W = 0, CWS_F = CONS_MAXR * CONS_MAXR, CWS_R = 2 * CONS_MAXR * CONS_MAXR, CWS_Y = CWS_R + CONS_MAXR,
       CWS_LAM = CWS_Y + CONS_MAXR, CWS_N = CWS_LAM + CONS_MAXR };
// Event witnesses (at most WIT_MAX per tree): kinds, trigger masks (bit 0
// falling, bit 1 rising) and handler actions as in include/tdyn.h.  The
// per-lane event workspace (doubles, lane-strided) holds the witness values
// at the step's start, at the two ends of the localization window and at
// the trial time.
constexpr int WIT_MAX = 8;
enum { WIT_Q = 0, WIT_U = 1, WIT_STATION_POS = 2, WIT_STATION_VEL = 3, WIT_TIME = 4 };
enum { EVA_REPORT = 0, EVA_SCALE_U = 1, EVA_TERMINATE = 2 };
enum { EW_E0 = 0, EW_LOW = WIT_MAX, EW_HIGH = 2 * WIT_MAX, EW_MID = 3 * WIT_MAX, EW_N = 4 * WIT_MAX };
// Rollout status of a lane whose event handler asked to terminate.
constexpr int ST_TERMINATED = 3;

struct DevTables {
    const int4* body_a;      // per global body: parent(local), type, qIndex, uIndex
    const int4* body_b;      // per global body: child_begin, child_count, rec_off, -
    const int* child_idx;    // local body numbers
    const double* prm;       // PRM_STRIDE per global body
    const int4* tree_a;      // per tree: nb, nq, nu, body_off
    const int4* tree_b;      // per tree: rec_size, nquat, hand-over base of the chain evaluator, lvl offset
    const double* gravity;   // per tree: 4 doubles
    const int4* pre;         // per global body: depth-first preorder of its tree: body, depth, child count, -
    const int4* bfs;         // per global body: breadth-first order of its tree: body, parent's position
                             // (-1 = Ground), first child's position, child count
    const int* lvl;          // per tree at tree_b.w: level count n, prm_bfs offset, then the n+1 level
                             // starts in bfs order
    const int4* bfs_b;       // per global body, breadth-first order: mobilizer type, qIndex, uIndex, -
    const double* prm_bfs;   // per tree at lvl[1]: breadth-first positions in blocks of 16, constant k of
                             // position j at [(j/16)*16*PRM_STRIDE + k*16 + j%16]
    const int4* cons_t;      // per tree (nullable: no constraints): first constraint, count, rows, -
    const int4* cons_a;      // per constraint: kind, body 1, body 2, -
    const double* cons_p;    // per constraint: station on body 1, station on body 2, rod length (7)
    const int4* wit_t;       // per tree (nullable: no witnesses): first witness, count, kinematics needed
                             // (0 none, 1 positions, 2 velocities), -
    const int4* wit_a;       // per witness: kind | trigger << 4 | action << 8, coordinate index, body,
                             // action index
    const double* wit_p;     // per witness: station (3), direction (3), offset, factor, window (9)
};

// Boundary layout offsets of each instance (instance-major packing).
struct InstOffsets {   // indexed by lane
    const int* q_off;   // start of the instance's q in packed q arrays
    const int* u_off;   // start of its u in packed u arrays
    const int* b_off;   // start of its (nb+1) bodies in packed per-body arrays
};

struct RkmCfg {
    double accuracy;    // Integrator::setAccuracy
    double cons_tol;    // accuracy/10 unless set
    double fixed_h;     // >0: setFixedStepSize
    double init_h, min_h, max_h;   // -1 = unset
    double timescale;   // System default time scale (0.1)
    int max_steps;      // accepted-step budget for this call (per lane)
    int max_attempts;   // attempt budget per lane for the whole rollout
    int inf_norm;       // Integrator::setUseInfinityNorm: every error and constraint norm the largest magnitude
    int force_proj;     // setProjectEveryStep: ForceProjection on every projection
    int proj_interp;    // setProjectInterpolatedStates (1, the default): project interpolated states
};
// Options of the device projections (cons_project_q/u): ForceProjection and
// ForceFullNewton at initialize; ForceProjection (setProjectEveryStep); the
// infinity norm (setUseInfinityNorm).
enum { PJ_INIT = 1, PJ_FORCE = 2, PJ_INF = 4, PJ_NONE = 8 /* reports: interpolated states not projected */ };
__host__ __device__ inline int proj_opts(const RkmCfg& c) {
    return (c.inf_norm ? PJ_INF : 0) | (c.force_proj ? PJ_FORCE : 0);
}
// Accumulates x into a norm: sum of squares, or (inf) the largest magnitude,
// a NaN kept.
__host__ __device__ inline void norm_acc(double& s, double x, bool inf) {
    if (inf) { const double a = fabs(x); s = (a > s || a != a) ? a : s; }
    else s += x*x;
}
__host__ __device__ inline double norm_end(double s, int n, bool inf) {
    return n ? (inf ? s : sqrt(s / n)) : 0.0;
}

// Launch entry points (tdyn_kernels.hip).
enum OpKind { OP_KIN = 0, OP_FWD_SYS = 1, OP_FWD_GIVEN = 2, OP_MINV = 3, OP_RESID = 4, OP_ABI = 5, OP_REACT = 6,
              OP_MULM = 7, OP_JAC = 8, OP_JACT = 9, OP_BACC = 10, OP_CBI = 11, OP_FWD_PRES = 12, OP_FWD_CONS = 13,
              OP_WIT = 14 };

struct OpArgs {
    DevTables T;
    const int* tree_of;      // null: every instance uses tree 0
    int n_inst;
    int L;                   // lane pitch of the scratch
    double* scratch;
    int rec_stride;          // doubles of scratch per lane (per-lane record layout of mixed-tree batches)
    InstOffsets off;
    const double* q;         // packed boundary layout
    const double* u;
    const double* fmob;      // nullable
    const double* fbody;     // nullable ((nb+1)*6 per instance)
    const double* udot_in;   // residual / body-acceleration input (nullable); the vector v of M v and J v
    double* out_u;           // udot / M^-1 f / residual   (packed u layout)
    double* out_q;           // qdot or qdotdot (packed q layout), nullable
    double* out_body;        // A_GB / X_GB / reactions / J v (packed (nb+1)*6 or *12), composite
                             // inertias ((nb+1)*10), nullable
    double* out_body2;       // V_GB (kinematics), nullable
    double* out_abi;         // P per body (21), nullable
    double* out_abi2;        // PPlus per body (21), nullable
    double* out_scalar;      // kinematics: ke, pe per instance (2), nullable
    int* status;             // per instance, nullable
    const int* inst;         // instance of each lane (nullable: lane = instance)
    const int* presc;        // OP_FWD_PRES: per body (nb+1 per instance) nonzero = udot known (udot_in)
    double* out_tau;         // OP_FWD_PRES: tau on the known mobilities, 0 elsewhere (packed u layout)
    double* cons_ws;         // OP_FWD_CONS: per-lane constraint workspace (CWS_N doubles, lane-strided)
    double* out_lam;         // OP_FWD_CONS: multipliers, CONS_MAXR per instance (nullable)
    const double* times;     // OP_WIT: time per instance (nullable: 0)
    double* out_wit;         // OP_WIT: witness values, WIT_MAX per instance
};
hipError_t launch_op(OpKind k, const OpArgs& a, hipStream_t s);

struct RkmArgs {
    DevTables T;
    const int* tree_of;
    int n_inst, L;
    double* scratch;
    int rec_stride;          // doubles of scratch per lane (per-lane record layout of mixed-tree batches)
    float* scratch32;        // mixed path: float hand-over rows (6*nu G rows, then 4*nb trig rows) per lane
    InstOffsets off;
    const double* fapp;      // constant applied body forces, packed, nullable
    double* vec;             // 11 SoA vectors of ny_max*L doubles: y y0 f0 fa fb ysave y1 ytmp yerr c0 c1
    int ny_max;
    double* t;               // per lane
    double* h;               // per lane: current step size
    const double* t_end;     // per lane
    int* status;             // per lane: 0 running, 1 done, 2 empty serving slot, 3 terminated by an event
                             // handler, <0 failed
    int* stats;              // per lane: attempts accepted rejected projections
    double* uscale;          // SoA nu_max*L
    RkmCfg cfg;
    int precision;           // 0 = FP64, 1 = mixed, 2/3 = FP64 with single-precision G (and trig) rows
    int chain;               // 1: uniform serial-chain model, use the register-resident chain evaluator
    int idf;                 // chain evaluator: every frame is an identity rotation at the joints
                             // (R_PF = I, p_PF = 0, R_MB = I), so the frame products are skipped
    int ext;                 // some body's mobilizer is reversed (PRM_REV), uses Euler angles (PRM_EULER)
                             // or is a BendStretch
    const double* prm_lane;  // chain evaluator with one model per lane (serving slots), nullable:
                             // row k of lane l at prm_lane[k*L + l], nb*PRM_STRIDE model rows then gravity (3)
    int vec_aos;             // 1: vectors lane-contiguous (entry k of lane l at ((v*L + l)*ny_max + k)) and
                             // the streaming tree evaluator (batches of different trees)
    int stream_dmax;         // streaming tree evaluator: deepest body of any tree
    int grp;                 // >1: level-parallel tree evaluator (vec_aos layout), run by the plan below
    const int4* plan;        // level-parallel evaluator, per wavefront: lanes per instance G, first entry of
                             // plan_idx, instance count, kind: 0 the wavefront runs those instances
                             // (count <= 64/G), k >= 1 its groups take instances from that range as a
                             // work queue through plan_ctr[k - 1]
    const int* plan_idx;     // instances (lane slots) in plan order
    int plan_waves;
    int* plan_ctr;           // the work queues' counters (zeroed before each launch)
    int plan_nctr;           // how many counters plan_ctr holds (0 counts as 1)
    const int* presc;        // known udot (tree sweeps only): per body (nb+1 per instance, packed) nonzero = known
    const double* udot_known; // the known udots (packed u layout)
    const double* motion;    // prescribed motion (tree sweeps, nullable): 4 per body (nb+1 per instance,
                             // packed body layout): level (1 position, 2 velocity, 3 acceleration sinusoid,
                             // 4 steady velocity), amplitude, rate, phase; presc holds its flags and
                             // udot_known receives the prescribed udot of each evaluation
    const long long* pool;   // tree-server slots (nullable): per lane, where its state vectors start in vec
                             // (11 lane-contiguous vectors of its own nq+nu) and its records in scratch,
                             // in doubles; the lane's tree is tree `lane` of the tables
    const unsigned char* pool_grp; // tree-server slots: 1 where the slot runs the level-parallel evaluator
                             // (its records and tables in that evaluator's layout; the tree sweeps skip it)
    double* cons_ws;         // holonomic constraints (nullable: none in the batch): per-lane workspace
                             // (CWS_N doubles, lane-strided); the tree sweeps only
    double* ev_ws;           // event witnesses (nullable: none): per-lane workspace (EW_N doubles,
                             // lane-strided); the tree sweeps of batches of one tree only
    double* ev_win;          // per lane, ev_cap events: localization window (tLow, tHigh)
    int* ev_ids;             // per lane, ev_cap events: witness, transition (1 falling, 2 rising)
    int* ev_count;           // per lane: events seen since the rollout began (may exceed ev_cap)
    int ev_cap;
};
// Per-lane scratch of the streaming tree evaluator (doubles): ST (24 per
// depth: R p V A of the last body with two or more children seen at that
// depth), ACC (27 per depth: P and z sums of a body's children after the
// first), then one hand-over record per preorder position: pose R p V (18),
// G (6 per dof), DI*eps (dof), trig cache (4).  Position 0 (body 1, 6 dof if
// Free) has 64, the others (3 dof, padded) 43.
namespace tstream {
enum { ST_W = 24, ACC_W = 27, POS0_W = 64, POS_W = 43 };
__host__ __device__ inline int scratch_doubles(int dmax, int nbmax) {
    return (ST_W + ACC_W) * (dmax + 1) + POS0_W + POS_W * (nbmax > 1 ? nbmax - 1 : 0);
}
// A lane-contiguous state vector and a body's hand-over values (G, DI*eps,
// trig cache) in a per-lane record.
struct LaneVec {
    double* p;
    __device__ __forceinline__ double& operator()(int k) const { return p[k]; }
};
struct LaneStore {
    double* p; int g, tc, de;
    __device__ __forceinline__ double& G(int i) const { return p[g + i]; }
    __device__ __forceinline__ double& T(int i) const { return p[tc + i]; }
    __device__ __forceinline__ double& DE(int i) const { return p[de + i]; }
};
} // namespace tstream
// Per-instance records of the level-parallel tree evaluator: NF fields per
// breadth-first position, in blocks of BW positions stored field-major
// (field f of position j at (j/BW)*NF*BW + f*BW + j%BW), so that the lanes of
// a group, which work on consecutive positions, touch consecutive doubles,
// while each lane still addresses a record's fields by immediate offsets.
// Fields: pose R p V (0..17), trig cache (18..21), P+ z+ phi for the parent
// (22..51), A_GB (52..57), G (58..93, 6 per dof) and DI*eps (94..99).
namespace tgroup {
enum { R_POSE = 0, R_TC = 18, R_PZ = 22, R_A = 52, R_G = 58, R_DE = 94, NF = 100, BW = 4 };
__host__ __device__ inline int blocks(int n) { return (n + BW - 1) / BW; }
__host__ __device__ inline int scratch_doubles(int nbmax) { return NF * BW * blocks(nbmax); }
__host__ __device__ inline int at(int j) { return (j / BW) * NF * BW + j % BW; }
} // namespace tgroup
// Entry k of state vector v of a lane, in either layout.
__device__ __forceinline__ double& vec_at(const RkmArgs& a, int v, int lane, int k) {
    return a.vec_aos ? a.vec[((size_t)v * a.L + lane) * a.ny_max + k] : a.vec[((size_t)v * a.ny_max + k) * a.L + lane];
}
// Rollout state vectors (rows of ny_max * L doubles in RkmArgs::vec): y, y0,
// f0, fa, fb, ysave, y1, ystage, yerr and the mixed path's Kahan rows.
enum { VY = 0, VY0 = 1, VF0 = 2, VFA = 3, VFB = 4, VYS = 5, VY1 = 6, VYT = 7, VYE = 8, VC0 = 9, VC1 = 10, NVEC = 11 };

hipError_t launch_rkm_init(const RkmArgs& a, double t0, hipStream_t s);
// The initial projection of the lanes whose trees carry constraints
// (initialize(): projectQ, projectU with ForceProjection), after
// launch_rkm_init; status -4 where it fails.
hipError_t launch_cons_init(const RkmArgs& a, hipStream_t s);
// Wavefronts of the level-parallel tree evaluator the device keeps resident.
int tree_group_resident_waves();
// Chain-evaluator rollouts (f0_only: only the derivative of the current state)
hipError_t launch_chain_fp64_batch(const RkmArgs& a, hipStream_t s, bool f0_only);
hipError_t launch_chain_fp64_lane(const RkmArgs& a, hipStream_t s, bool f0_only);
hipError_t launch_chain_mixed_batch(const RkmArgs& a, hipStream_t s, bool f0_only);
hipError_t launch_chain_mixed_lane(const RkmArgs& a, hipStream_t s, bool f0_only);
// precision 2, 3: FP64 arithmetic, single-precision G (and, 3, trig cache) rows
hipError_t launch_chain_split_batch(const RkmArgs& a, hipStream_t s, bool f0_only);
inline hipError_t launch_chain_variant(const RkmArgs& a, hipStream_t s, bool f0_only) {
    if (a.precision >= 2 && !a.prm_lane) return launch_chain_split_batch(a, s, f0_only);
    if (a.precision == 1)
        return a.prm_lane ? launch_chain_mixed_lane(a, s, f0_only) : launch_chain_mixed_batch(a, s, f0_only);
    return a.prm_lane ? launch_chain_fp64_lane(a, s, f0_only) : launch_chain_fp64_batch(a, s, f0_only);
}
hipError_t launch_rkm(const RkmArgs& a, hipStream_t s);
// Level-parallel tree evaluator (wavefront plan a.plan): rollout steps, or
// (f0_only) the derivative of the current state.
hipError_t launch_tree_group(const RkmArgs& a, hipStream_t s, bool f0_only);
// Tree-server slots (pooled tables and state, RkmArgs::pool): admission of n
// staged requests (per request i: ints at ioff[i], doubles at doff[i], laid
// out by tdyn_tree_server.cpp) into their slots, and collection of finished
// slots (q then u at qoff[i] of out, t, status and statistics), which marks
// the slots empty.
hipError_t launch_tsrv_admit(const RkmArgs& a, const int* slots, int n, const long long* ioff, const int* ints,
                             const long long* doff, const double* dbl, hipStream_t s);
hipError_t launch_tsrv_collect(const RkmArgs& a, const int* slots, int n, const long long* qoff, double* out,
                               double* t_out, int* outi, hipStream_t s);
hipError_t launch_srv_admit(const RkmArgs& a, const double* rec, const int* slots, int n, int rec_stride,
                            double t0, hipStream_t s);
hipError_t launch_srv_collect(const RkmArgs& a, const int* slots, int n, double* out, int out_stride, int* outi,
                              hipStream_t s);
// vec: which state vector to read when unpacking, or write when packing
// (0 = the state y, 2 = f0, VC0 = the mixed path's compensation)
hipError_t launch_pack_state(const RkmArgs& a, const double* q, const double* u, int to_soa,
                             double* q_out, double* u_out, hipStream_t s, int vec = 0);
hipError_t launch_rollout_f0(const RkmArgs& a, hipStream_t s);
// per-instance <-> per-lane copies for batches whose lanes are permuted
hipError_t launch_lane_gather(const int* inst, int n, const double* src_inst, double* dst_lane, hipStream_t s);
hipError_t launch_rollout_extend(int n, const double* te_lane, double te_all, const double* t, double* t_end,
                                 int* status, hipStream_t s);
// prev/now: packed q, u, qdot, udot (instance order) at a step's two ends
hipError_t launch_report_interp(const DevTables& T, const int* tree_of, const int* q_off, const int* u_off, int L,
                                const double* rt, int nrep, int* next_k, const double* t_prev, const double* t_now,
                                const double* const prev[4], const double* const now[4], double cons_tol,
                                int popt, double* q_out, double* u_out, long total_q, long total_u, hipStream_t s);
// the event log (per lane) to instance order
hipError_t launch_event_scatter(const int* inst, int n, int cap, const double* win, const int* ids, const int* cnt,
                                double* win_out, int* ids_out, int* cnt_out, hipStream_t s);
hipError_t launch_lane_scatter(const int* inst, int n, const double* t_lane, double* t_inst, const int* st_lane,
                               int* st_inst, const int* stats_lane, int* stats_inst, hipStream_t s);

} // namespace tdyn
