// Level-parallel tree evaluator: G lanes (G = 4, 8, 16, 32 or 64, a divisor
// of the 64-lane wavefront) cooperate on one instance of a batch of
// different trees.  A batch of a few thousand instances gives a one-lane-per-instance
// launch a fraction of the chip's 1,024 SIMDs, and a rollout lasts as long
// as its slowest instance (the largest tree with the most step attempts);
// splitting each instance over G lanes cuts that instance's time.
//
// Each instance's bodies are visited in breadth-first order, one tree level
// at a time (Simbody's level sweeps, SimbodyMatterSubsystemRep.cpp:
// realizeArticulatedBodyInertias / calcTreeForwardDynamicsOperator run the
// same recursions level by level).  The bodies of a level are dealt to the G
// lanes round-robin (position j to lane j mod G); the levels are separated
// by a fence, and the lanes of a group are lanes of one wavefront, so no
// barrier is needed.  The per-body arithmetic is the chain evaluator's pass
// bodies on padded 3-dof bodies (body 1 may be Free):
//   pass 1 (top-down): parent pose from the parent's record, this body's
//          pose, qdot and trig cache out;
//   pass 2 (bottom-up): the children's P+, z+ pulled from their records and
//          summed, then P, D, D^-1, G, z, P+, z+ of this body;
//   pass 3 (top-down): udot = DI eps - G^T A+ and A_GB from the parent's A.
// The Merson stage combinations, the error norm and the projection are split
// over the group's lanes by state entry or body, and the norms are summed
// across the group with butterfly shuffles, so every lane of a group holds
// the bitwise-same step size and takes the same branches.
#include <type_traits>
#include "tdyn_internal.hpp"
#include "tdyn_math.hpp"
#include "tdyn_chain.hpp"
#include "tdyn_rkm_chain.hpp"

namespace tdyn {
namespace {

using d3 = v3<double>;
using dm = m33<double>;
using dv = sv<double>;
using da = abi<double>;

// Orders the group's memory accesses: what one lane stored before is seen by
// every lane of the workgroup (here: the wavefront) after.
__device__ __forceinline__ void grp_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
// Sum over the G lanes of a group (aligned lane blocks of G).  The butterfly
// adds the same pairs in every lane, so all lanes get bitwise the same sum.
// G is wave-uniform (one G per wavefront).
__device__ __forceinline__ double grp_sum(int G, double x) {
    for (int m = 1; m < G; m <<= 1) x += __shfl_xor(x, m, 64);
    return x;
}
// The group's combined norm accumulator: the sum of squares, or (inf) the
// largest magnitude with a NaN kept (norm_acc's rule, lane against lane).
__device__ __forceinline__ double grp_norm(int G, double x, bool inf) {
    if (!inf) return grp_sum(G, x);
    for (int m = 1; m < G; m <<= 1) { const double y = __shfl_xor(x, m, 64); x = (y > x || y != y) ? y : x; }
    return x;
}

template <int V> using TyTag = std::integral_constant<int, V>;

// Blocked views (stride BW between fields): a position's record fields; the
// hand-over values of the pass bodies; a body's model constants.
struct RecS {
    double* p;
    __device__ __forceinline__ double& operator[](int f) const { return p[f * tgroup::BW]; }
};
struct StoreS {
    double* p;
    __device__ __forceinline__ double& G(int i) const { return p[(tgroup::R_G + i) * tgroup::BW]; }
    __device__ __forceinline__ double& T(int i) const { return p[(tgroup::R_TC + i) * tgroup::BW]; }
    __device__ __forceinline__ double& DE(int i) const { return p[(tgroup::R_DE + i) * tgroup::BW]; }
};
struct PrmS {
    const double* p;
    __device__ __forceinline__ double operator[](int k) const { return p[k * tgroup::BW]; }
    __device__ __forceinline__ PrmS operator+(int k) const { return {p + k * tgroup::BW}; }
};
// A body's P+, z+ and phi for its parent (30 doubles), field k at p[k * s]:
// in the LDS level frontier (s = G) or in the body's record (s = BW).
struct PZS {
    double* p; int s;
    __device__ __forceinline__ double& operator[](int k) const { return p[k * s]; }
};
// LDS level frontier of the pass-2 hand-over: per group two level buffers
// (the level being read and the level being written) of 30 fields x G
// positions, field-major so the lanes of a group touch consecutive doubles.
// 64/G groups x 2 x 30 x G doubles = 30 KB per wavefront for every G.  A
// level wider than G keeps its hand-over in the records.  (The pose and A
// hand-overs of passes 1 and 3 through the same frontier measured slower:
// the records those passes write are read back by later passes anyway.)
constexpr int kFrontierFields = 30;
constexpr size_t kFrontierBytes = 64 * 2 * kFrontierFields * sizeof(double);

template <int IDF>
struct GroupEval {
    // IDF: 0 general joint frames, 1 identity frames at the joints, 2 general
    // frames with reversed mobilizers or Euler-angle Balls and Frees
    using PA = std::conditional_t<IDF == 1, chain::IdF<PrmS>,
                                  std::conditional_t<IDF == 2, chain::ExtF<PrmS>, PrmS>>;
    using VR = tstream::LaneVec;
    int nb, nq, nlev, l, G;
    double* lf;        // this group's LDS frontier (2 x 30 x G doubles)
    const int4* bfs; const int4* bfs_b; const int* ls; const double* prm;
    double* S;
    d3 grav; const double* fapp;
    const int* pm; const double* uk;   // known udot: per-body flags (Ground first) and values, or null
    bool free0;
    static __device__ __forceinline__ VR vrow(const RkmArgs& a, int v, int inst) {
        if (a.pool) {
            // a tree-server slot: its 11 vectors of its own nq + nu, where
            // the slot's pool entry puts them (the slot's tree is tree inst)
            const int4 ta = ldi4(a.T.tree_a, inst);
            return VR{a.vec + a.pool[2*inst] + (size_t)v * (ta.y + ta.z)};
        }
        return VR{a.vec + ((size_t)v * a.L + inst) * a.ny_max};
    }
    __device__ GroupEval(const RkmArgs& a, int inst, int tr, int lane_in_group, int lanes, double* frontier)
        : l(lane_in_group), G(lanes), lf(frontier) {
        const int4 ta = ldi4(a.T.tree_a, tr);
        nb = ta.x; nq = ta.y;
        const int body_off = ta.w;
        bfs = a.T.bfs + body_off; bfs_b = a.T.bfs_b + body_off;
        const int* lv = a.T.lvl + ldi4(a.T.tree_b, tr).w;
        nlev = lv[0]; prm = a.T.prm_bfs + lv[1]; ls = lv + 2;
        S = a.scratch + (a.pool ? a.pool[2*inst + 1] : (long long)inst * a.rec_stride);
        const double* g = a.T.gravity + 4*tr;
        grav = V3(g[0], g[1], g[2]);
        fapp = a.fapp ? a.fapp + 6*(size_t)a.off.b_off[inst] : nullptr;
        pm = (IDF == 2 && a.presc) ? a.presc + a.off.b_off[inst] : nullptr;
        uk = pm ? a.udot_known + a.off.u_off[inst] : nullptr;
        free0 = ldi4(a.T.body_a, body_off).y == FREE;
    }
    __device__ __forceinline__ RecS rec(int j) const { return {S + tgroup::at(j)}; }
    // P+ z+ phi of position j of level lev: the LDS frontier when the level
    // fits it (at most G positions, the same answer in every lane of the
    // group), else the record
    __device__ __forceinline__ PZS pz(int j, int lev) const {
        const int s0 = ls[lev];
        if (ls[lev + 1] - s0 <= G) return {lf + (size_t)((lev & 1) * kFrontierFields) * G + (j - s0), G};
        return {S + tgroup::at(j) + tgroup::R_PZ * tgroup::BW, tgroup::BW};
    }
    __device__ __forceinline__ StoreS store(int j) const { return {S + tgroup::at(j)}; }
    __device__ __forceinline__ PA prm_of(int j) const {
        return PA(PrmS{prm + (j / tgroup::BW) * PRM_STRIDE * tgroup::BW + j % tgroup::BW});
    }
    __device__ __forceinline__ dv F(int b) const {
        if (!fapp) return svzero<double>();
        const double* f = fapp + 6*b;
        return {V3(f[0], f[1], f[2]), V3(f[3], f[4], f[5])};
    }
    template <int TYPE, bool G_ROWS>
    __device__ __forceinline__ void load(chain::Pre<double>& x, const VR& y, int4 bt, const StoreS& st) const {
        if constexpr (TYPE == FREE) {
            x.ty = FREE;
#pragma unroll
            for (int k = 0; k < 7; ++k) x.q[k] = y(bt.y + k);
#pragma unroll
            for (int k = 0; k < 6; ++k) x.u[k] = y(nq + bt.z + k);
            x.tc[0] = st.T(0);
        } else {
            const int ty = bt.x, d = mob_dof(ty), nqb = mob_nq(ty);
            x.ty = ty;
#pragma unroll
            for (int k = 0; k < 4; ++k) x.q[k] = k < nqb ? y(bt.y + k) : 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) x.u[k] = k < d ? y(nq + bt.z + k) : 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) x.tc[k] = st.T(k);
            if constexpr (G_ROWS) {
#pragma unroll
                for (int k = 0; k < 18; ++k) x.g[k] = st.G(k);
#pragma unroll
                for (int k = 0; k < 3; ++k) x.de[k] = st.DE(k);
            }
        }
    }
    __device__ __forceinline__ void ld_pose(const RecS& s, dm& R, d3& p, dv& V) const {
#pragma unroll
        for (int k = 0; k < 9; ++k) R.a[k] = s[k];
        p = V3(s[9], s[10], s[11]);
        V = {V3(s[12], s[13], s[14]), V3(s[15], s[16], s[17])};
    }
    __device__ __forceinline__ void st_pose(const RecS& s, const dm& R, const d3& p, const dv& V) const {
#pragma unroll
        for (int k = 0; k < 9; ++k) s[k] = R.a[k];
        s[9] = p.x; s[10] = p.y; s[11] = p.z;
        s[12] = V.w.x; s[13] = V.w.y; s[14] = V.w.z; s[15] = V.v.x; s[16] = V.v.y; s[17] = V.v.z;
    }
    // Runs f(TyTag<TYPE>, j) for this lane's positions of level lev:
    // j = ls[lev] + ... with j = l (mod G); position 0 (body 1) may be Free.
    template <class Fn> __device__ __forceinline__ void each(int lev, Fn&& f) const {
        const int s0 = ls[lev], e = ls[lev + 1];
        int j = s0 + ((l - s0) & (G - 1));
        if (j == 0) {
            if (free0) f(TyTag<FREE>{}, 0); else f(TyTag<GEN>{}, 0);
            j = G;
        }
        for (; j < e; j += G) f(TyTag<GEN>{}, j);
    }
    __device__ __forceinline__ void derivs(const VR& y, const VR& f) const {
        using namespace chain;
        const VR fu{f.p + nq};
        for (int lev = 0; lev < nlev; ++lev) {
            each(lev, [&](auto tag, int j) {
                constexpr int TY = decltype(tag)::value;
                const int4 e = ldi4(bfs, j), bt = ldi4(bfs_b, j);
                dm R; d3 p; dv V;
                if (e.y < 0) { R = ident<double>(); p = zero3<double>(); V = svzero<double>(); }
                else ld_pose(rec(e.y), R, p, V);
                const StoreS st = store(j);
                Pre<double> x;
                load<TY, false>(x, y, bt, st);
                pass1_body<TY, double>(prm_of(j), x, bt.y, f, st, R, p, V);
                st_pose(rec(j), R, p, V);
            });
            grp_sync();
        }
        for (int lev = nlev - 1; lev >= 0; --lev) {
            each(lev, [&](auto tag, int j) {
                constexpr int TY = decltype(tag)::value;
                const int4 e = ldi4(bfs, j), bt = ldi4(bfs_b, j);
                const RecS r = rec(j);
                dm R; d3 p; dv V;
                ld_pose(r, R, p, V);
                const StoreS st = store(j);
                Pre<double> x;
                load<TY, false>(x, y, bt, st);
                const double* ukn = nullptr;      // known udot (extended build only)
                if constexpr (IDF == 2) { if (pm && pm[bt.w]) ukn = uk + bt.z; }
                da PP; dv zp = svzero<double>(); d3 phic = zero3<double>();
                if (e.w > 0) {
                    // children's P+ and z+, shifted to this body, in child order
                    da aP;
                    aP.M = {0, 0, 0, 0, 0, 0}; aP.J = {0, 0, 0, 0, 0, 0};
#pragma unroll
                    for (int k = 0; k < 9; ++k) aP.F.a[k] = 0.0;
                    dv aZ = svzero<double>();
                    for (int c = e.z; c < e.z + e.w; ++c) {
                        const PZS q = pz(c, lev + 1);
                        da Q;
                        Q.M = {q[0], q[1], q[2], q[3], q[4], q[5]};
                        Q.J = {q[6], q[7], q[8], q[9], q[10], q[11]};
#pragma unroll
                        for (int k = 0; k < 9; ++k) Q.F.a[k] = q[12 + k];
                        const dv zc = {V3(q[21], q[22], q[23]), V3(q[24], q[25], q[26])};
                        const d3 ph = V3(q[27], q[28], q[29]);
                        abi_add_shifted(aP, Q, ph);
                        aZ = aZ + phi_mul(ph, zc);
                    }
                    pass2_body<TY, double>(prm_of(j), x, grav, F(bt.w), st, R, p, V, false, PP, zp, phic,
                                           &aP, &aZ, ukn);
                } else {
                    pass2_body<TY, double>(prm_of(j), x, grav, F(bt.w), st, R, p, V, false, PP, zp, phic,
                                           nullptr, nullptr, ukn);
                }
                if (e.y >= 0) {
                    const PZS q = pz(j, lev);
                    q[0] = PP.M.xx; q[1] = PP.M.yy; q[2] = PP.M.zz; q[3] = PP.M.xy; q[4] = PP.M.xz; q[5] = PP.M.yz;
                    q[6] = PP.J.xx; q[7] = PP.J.yy; q[8] = PP.J.zz; q[9] = PP.J.xy; q[10] = PP.J.xz; q[11] = PP.J.yz;
#pragma unroll
                    for (int k = 0; k < 9; ++k) q[12 + k] = PP.F.a[k];
                    q[21] = zp.w.x; q[22] = zp.w.y; q[23] = zp.w.z; q[24] = zp.v.x; q[25] = zp.v.y; q[26] = zp.v.z;
                    q[27] = phic.x; q[28] = phic.y; q[29] = phic.z;
                }
            });
            grp_sync();
        }
        for (int lev = 0; lev < nlev; ++lev) {
            each(lev, [&](auto tag, int j) {
                constexpr int TY = decltype(tag)::value;
                const int4 e = ldi4(bfs, j), bt = ldi4(bfs_b, j);
                dm R; d3 p; dv V, A;
                if (e.y < 0) { R = ident<double>(); p = zero3<double>(); V = svzero<double>(); A = svzero<double>(); }
                else {
                    const RecS s = rec(e.y);
                    ld_pose(s, R, p, V);
                    A = {V3(s[tgroup::R_A], s[tgroup::R_A + 1], s[tgroup::R_A + 2]),
                         V3(s[tgroup::R_A + 3], s[tgroup::R_A + 4], s[tgroup::R_A + 5])};
                }
                const StoreS st = store(j);
                Pre<double> x;
                load<TY, true>(x, y, bt, st);
                pass3_body<TY, double>(prm_of(j), x, bt.z, fu, st, R, p, V, A);
                const RecS s = rec(j);
                s[tgroup::R_A] = A.w.x; s[tgroup::R_A + 1] = A.w.y; s[tgroup::R_A + 2] = A.w.z;
                s[tgroup::R_A + 3] = A.v.x; s[tgroup::R_A + 4] = A.v.y; s[tgroup::R_A + 5] = A.v.z;
            });
            grp_sync();
        }
    }
};

// One group of G lanes runs one instance's TimeStepper loop (the one-lane
// loop of the other tree evaluators, RungeKuttaMersonIntegratorRep.cpp:
// 80-160 and AbstractIntegratorRep.cpp:448-502), its vector work split over
// the group.  l: the lane's index in the group.
template <int IDF>
__device__ __forceinline__ void grp_instance(const RkmArgs& a, int inst, int l, int G, double* frontier) {
    const int tr = a.tree_of[inst];
    const GroupEval<IDF> ev(a, inst, tr, l, G, frontier);
    const int4 ta = ldi4(a.T.tree_a, tr);
    const int nb = ta.x, nq = ta.y, nu = ta.z, ny = nq + nu;
    const int4* body_a = a.T.body_a + ta.w;
    const int nquat = ldi4(a.T.tree_b, tr).y;
    using VR = tstream::LaneVec;
    auto rows = [&](int v) { return GroupEval<IDF>::vrow(a, v, inst); };
    const VR Y = rows(VY), F0 = rows(VF0), FA = rows(VFA), FB = rows(VFB), YS = rows(VYS),
             Y1 = rows(VY1), YT = rows(VYT), YE = rows(VYE);

    const RkmCfg& cfg = a.cfg;
    const double acc = cfg.accuracy, consTol = cfg.cons_tol;
    double hmin = cfg.min_h, hmax = cfg.max_h;
    if (cfg.fixed_h > 0) { hmin = hmax = cfg.fixed_h; }
    const double projLimit = fmax(2*consTol, sqrt(consTol));
    double t = a.t[inst], hcur = a.h[inst];
    const double tEnd = a.t_end[inst];
    int attempts = a.stats[4*inst], accepted = a.stats[4*inst+1], rejected = a.stats[4*inst+2],
        projections = a.stats[4*inst+3];
    int status = 0, steps = 0;

    // error norm at y1 (IntegratorRep.h:454-523, calcRelativeScaling :644-655),
    // RMS or (setUseInfinityNorm) the largest magnitude
    const bool infn = cfg.inf_norm;
    auto errNorm = [&]() -> double {
        double sq = 0;
        for (int b = l; b < nb; b += G) {
            const int4 ba = ldi4(body_a, b);
            const int qi = ba.z;
            // (a Ball or Free in Euler-angle mode holds no quaternion; the
            // flag is read per body, not tested once per tree, which keeps
            // the compiler from duplicating the loop for the two modes)
            if ((ba.y == BALL || ba.y == FREE) && a.T.prm[(size_t)(ta.w + b) * PRM_STRIDE + PRM_EULER] == 0.0) {
                double qq[4], e[4], d4[4];
                for (int k = 0; k < 4; ++k) { qq[k] = Y1(qi + k); e[k] = YE(qi + k); }
                const d3 w = quat_dot_to_w(qq, e);
                quat_dot(qq, w, d4);
                if (infn) for (int k = 0; k < 4; ++k) norm_acc(sq, d4[k], true);
                else sq += d4[0]*d4[0] + d4[1]*d4[1] + d4[2]*d4[2] + d4[3]*d4[3];
                if (ba.y == FREE) for (int k = 4; k < 7; ++k) norm_acc(sq, YE(qi + k), infn);
            } else {
                for (int k = 0; k < mob_dof(ba.y); ++k) norm_acc(sq, YE(qi + k), infn);
            }
        }
        double su = 0;
        for (int k = l; k < nu; k += G) {
            const double vi = fabs(Y(nq + k));
            const double us = vi > 1.0 ? 1.0/vi : 1.0;
            norm_acc(su, us * YE(nq + k), infn);
        }
        sq = grp_norm(G, sq, infn); su = grp_norm(G, su, infn);
        const double qn = norm_end(sq, nq, infn);
        const double un = norm_end(su, nu, infn);
        return qn >= un ? qn : un;
    };

    while (steps < cfg.max_steps) {
        if (t >= tEnd) { status = 1; break; }
        bool ok = false;
        double t1 = t;
        int s0 = 0;          // f0 = f(y0) once per step
        do {
            bool limited = false;
            if (tEnd < t + 0.95*hcur) { limited = true; t1 = tEnd; }
            else if (tEnd > t + 1.001*hcur) t1 = t + hcur;
            else t1 = tEnd;
            if (!(t1 > t)) { status = -3; break; }
            if (++attempts > cfg.max_attempts) { status = -2; break; }
            const double h = t1 - t;
            for (int st = s0; st < 5; ++st) {
                VR in = Y, out = F0;
                if (st == 1) {
                    for (int k = l; k < ny; k += G) YT(k) = Y(k) + (h/3)*F0(k);
                    in = YT; out = FA;
                } else if (st == 2) {
                    for (int k = l; k < ny; k += G) YT(k) = Y(k) + (h/6)*(F0(k) + FA(k));
                    in = YT; out = FA;
                } else if (st == 3) {
                    for (int k = l; k < ny; k += G) YT(k) = Y(k) + (h/8)*(F0(k) + 3*FA(k));
                    in = YT; out = FB;
                } else if (st == 4) {
                    for (int k = l; k < ny; k += G) YS(k) = Y(k) + (h/2)*((F0(k) - 3*FA(k)) + 4*FB(k));
                    in = YS; out = FA;
                }
                grp_sync();
                ev.derivs(in, out);
            }
            s0 = 1;
            for (int k = l; k < ny; k += G) {
                const double y1 = Y(k) + (h/6)*((F0(k) + 4*FB(k)) + FA(k));
                Y1(k) = y1;
                YE(k) = 0.2 * fabs(y1 - YS(k));
            }
            grp_sync();
            bool converged = true;
            const double e1 = errNorm();
            if (!(e1 > 16.0*acc) && nquat > 0) {
                double s2 = 0;
                for (int b = l; b < nb; b += G) {
                    const int4 ba = ldi4(body_a, b);
                    // quaternions only (a Free body flagged PRM_EULER holds angles)
                    if (ba.y != BALL && (ba.y != FREE || a.T.prm[(size_t)(ta.w + b) * PRM_STRIDE + PRM_EULER] != 0.0)) continue;
                    double n2 = 0;
                    for (int k = 0; k < 4; ++k) { const double x = Y1(ba.z + k); n2 += x*x; }
                    norm_acc(s2, sqrt(n2) - 1.0, infn);
                }
                s2 = grp_norm(G, s2, infn);
                const double qe = norm_end(s2, nquat, infn);
                if (qe > projLimit) converged = false;
                else if (qe > consTol || cfg.force_proj) {
                    ++projections;
                    for (int b = l; b < nb; b += G) {
                        const int4 ba = ldi4(body_a, b);
                        // quaternions only (a Free body flagged PRM_EULER holds angles)
                        if (ba.y != BALL && (ba.y != FREE || a.T.prm[(size_t)(ta.w + b) * PRM_STRIDE + PRM_EULER] != 0.0)) continue;
                        double qq[4], n2 = 0;
                        for (int k = 0; k < 4; ++k) { qq[k] = Y1(ba.z + k); n2 += qq[k]*qq[k]; }
                        const double n = sqrt(n2);
                        for (int k = 0; k < 4; ++k) { qq[k] = qq[k] / n; Y1(ba.z + k) = qq[k]; }
                        double d = 0;
                        for (int k = 0; k < 4; ++k) d += YE(ba.z + k) * qq[k];
                        for (int k = 0; k < 4; ++k) YE(ba.z + k) = YE(ba.z + k) - d*qq[k];
                    }
                    grp_sync();
                }
            }
            const double err = converged ? errNorm() : __builtin_inf();
            ok = adjust_step(err, acc, limited, hmin, hmax, hcur);
            if (!ok) ++rejected;
        } while (!ok);
        if (status < 0) break;
        t = t1;
        for (int k = l; k < ny; k += G) Y(k) = Y1(k);
        grp_sync();
        ++accepted; ++steps;
    }
    if (status == 0 && t >= tEnd) status = 1;
    if (l == 0) {
        a.t[inst] = t; a.h[inst] = hcur; a.status[inst] = status;
        a.stats[4*inst] = attempts; a.stats[4*inst+1] = accepted; a.stats[4*inst+2] = rejected;
        a.stats[4*inst+3] = projections;
    }
}

// Work distribution: a plan of wavefronts.  Kind 0 runs up to 64/G given
// instances with G lanes each; kind k >= 1 (persistent wavefronts, as many
// as the device keeps resident) has each of its 64/G groups take instances
// from work queue k (a range of plan_idx in decreasing order of cost, its
// counter plan_ctr[k - 1]; the tree server keeps one queue per lane count)
// until the queue is empty.  The lanes of a wavefront reconverge after an
// instance, so its groups take their next instances together; the queue
// balances the wavefronts against each other (a wavefront whose instances
// needed few steps takes more of them) and starts the costliest first.
// (Switching instances per step instead, so that a group whose instance is
// done does not wait for the others of its wavefront, measured 3-4 % slower
// on 16,384 trees: the groups' steps fall out of phase and the wavefront
// runs the union of their paths.)  An instance's arithmetic does not depend
// on the group or wavefront that runs it.
template <int IDF, bool F0>
__device__ __forceinline__ void run_instance(const RkmArgs& a, int inst, int l, int G, double* frontier) {
    if constexpr (F0) {
        const GroupEval<IDF> ev(a, inst, a.tree_of[inst], l, G, frontier);
        ev.derivs(GroupEval<IDF>::vrow(a, VY, inst), GroupEval<IDF>::vrow(a, VF0, inst));
    } else {
        if (a.status[inst] != 0) return;
        grp_instance<IDF>(a, inst, l, G, frontier);
    }
}
template <int IDF, bool F0>
__device__ __forceinline__ void plan_wave(const RkmArgs& a, int G, int start, int count, int kind) {
    extern __shared__ double lds_frontier[];
    const int lg = __builtin_ctz((unsigned)G);
    const int grp = threadIdx.x >> lg, l = threadIdx.x & (G - 1);
    double* frontier = lds_frontier + (size_t)grp * 2 * kFrontierFields * G;
    const int leader = threadIdx.x & ~(G - 1);
    // (one call site of the evaluator for both kinds)
    int k = grp;
    for (;;) {
        if (kind != 0) {
            if (l == 0) k = atomicAdd(a.plan_ctr + (kind - 1), 1);
            k = __shfl(k, leader, 64);
        }
        if (k >= count) break;
        run_instance<IDF, F0>(a, a.plan_idx[start + k], l, G, frontier);
        if (kind == 0) break;
    }
}
// One build for every G (4 to 64, a power of two, the same in the whole
// wavefront): one copy of the evaluator's code serves all wavefronts.
template <int IDF, bool F0>
__global__ void __launch_bounds__(64) k_rkm_plan(RkmArgs a) {
    const int4 w = a.plan[blockIdx.x];
    const int G = __builtin_amdgcn_readfirstlane(w.x);
    plan_wave<IDF, F0>(a, G, __builtin_amdgcn_readfirstlane(w.y), __builtin_amdgcn_readfirstlane(w.z),
                       __builtin_amdgcn_readfirstlane(w.w));
}
__global__ void k_zero_int(int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

} // namespace

// a.idf: every tree's joint frames are identity rotations (Simbody's
// noR_PF/noX_MB node forms), the frame products are skipped; the results
// are bitwise those of the general form.  a.ext: some mobilizer is reversed
// or uses Euler angles, or some bodies have known udot; only that build
// carries the code for them.
hipError_t launch_tree_group(const RkmArgs& a, hipStream_t s, bool f0_only) {
    if (a.n_inst <= 0 || a.plan_waves <= 0) return hipSuccess;
    const dim3 grd((unsigned)a.plan_waves), blk(64);
    k_zero_int<<<1, 64, 0, s>>>(a.plan_ctr, a.plan_nctr > 0 ? a.plan_nctr : 1);
    if (a.ext) {
        if (f0_only) k_rkm_plan<2, true><<<grd, blk, kFrontierBytes, s>>>(a);
        else k_rkm_plan<2, false><<<grd, blk, kFrontierBytes, s>>>(a);
    } else if (a.idf) {
        if (f0_only) k_rkm_plan<1, true><<<grd, blk, kFrontierBytes, s>>>(a);
        else k_rkm_plan<1, false><<<grd, blk, kFrontierBytes, s>>>(a);
    } else {
        if (f0_only) k_rkm_plan<0, true><<<grd, blk, kFrontierBytes, s>>>(a);
        else k_rkm_plan<0, false><<<grd, blk, kFrontierBytes, s>>>(a);
    }
    return hipGetLastError();
}

int tree_group_resident_waves() {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rkm_plan<1, false>, 64, kFrontierBytes) != hipSuccess)
        return 1024;
    return std::max(1, cus * per_cu);
}

} // namespace tdyn
