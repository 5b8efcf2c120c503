// Serial-chain evaluator of the state derivative (realize(Acceleration)) for
// batches whose instances all share one chain model (parent(b) = b-1).
//
// Same algorithm as the tree sweeps (Simbody's ABA, RigidBodyNodeSpec.cpp:
// 249-446 and RigidBodyNode.cpp:54-213), organized so that a lane keeps the
// whole recursion state in registers:
//   pass 1 (base->tip): transforms and velocities only; qdot out.  Only the
//          tip's X_GB, V_GB survive the sweep.
//   pass 2 (tip->base): each body's parent transform and velocity are
//          recovered from the body's own (X_GP = X_GB X_PB^-1, V_GP = V_GB -
//          H u - ...), then H, Phi, coriolis a, gyroscopic b, M_k, gravity,
//          the articulated inertia P, D^-1, G, P+, z, eps, z+ -- the child's
//          P+, z+ and Phi arrive in registers from the previous iteration.
//          Only G and DI*eps (7 dof doubles) are written per body.
//   pass 3 (base->tip): kinematics recomputed forward, udot = DI eps - G^T A+,
//          A_GB = A+ + H udot + a.
// The model is uniform across the batch, so its constants are read with
// wave-uniform (scalar) loads and the mobilizer switch is a uniform branch.
#pragma once
#include "tdyn_internal.hpp"
#include "tdyn_math.hpp"

namespace tdyn {
namespace chain {

using d3 = v3<double>;
using dm = m33<double>;
using ds = s33<double>;
using dv = sv<double>;
using da = abi<double>;

template <int TYPE> struct MobC {
    static constexpr int D = TYPE <= SLIDER ? 1 : TYPE == UNIVERSAL ? 2 : TYPE == BALL ? 3 : 6;
    static constexpr int NQ = TYPE <= SLIDER ? 1 : TYPE == UNIVERSAL ? 2 : TYPE == BALL ? 4 : 7;
};

__device__ __forceinline__ dm ldm(cdptr p) { dm r; for (int i = 0; i < 9; ++i) r.a[i] = p[i]; return r; }
__device__ __forceinline__ d3 ld3(cdptr p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ dm mulABt(const dm& A, const dm& B) {   // A B^T
    dm C;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) C.a[r*3+c] = A.a[r*3]*B.a[c*3] + A.a[r*3+1]*B.a[c*3+1] + A.a[r*3+2]*B.a[c*3+2];
    return C;
}

// Mobilizer part: X_FM(q) and H_FM.
template <int TYPE> struct Joint {
    static constexpr int D = MobC<TYPE>::D;
    dm R_FM; d3 p_FM; dv H_FM[D];
    // tc: cached trig / normalization values of this q (written by the first
    // sweep with CACHED = false): Pin (cos,sin), Universal (c1,s1,c2,s2),
    // Ball/Free 1/|quaternion|.
    template <bool CACHED>
    __device__ __forceinline__ void set(const double* qv, double* tc) {
        p_FM = zero3<double>();
        if constexpr (TYPE == PIN) {
            double sn, cs;
            if constexpr (CACHED) { cs = tc[0]; sn = tc[1]; } else { sincos(qv[0], &sn, &cs); tc[0] = cs; tc[1] = sn; }
            R_FM.a[0]=cs; R_FM.a[1]=-sn; R_FM.a[2]=0; R_FM.a[3]=sn; R_FM.a[4]=cs; R_FM.a[5]=0;
            R_FM.a[6]=0;  R_FM.a[7]=0;   R_FM.a[8]=1;
            H_FM[0] = {V3(0.0, 0.0, 1.0), zero3<double>()};
        } else if constexpr (TYPE == SLIDER) {
            R_FM = ident<double>(); p_FM = V3(qv[0], 0.0, 0.0);
            H_FM[0] = {zero3<double>(), V3(1.0, 0.0, 0.0)};
        } else if constexpr (TYPE == UNIVERSAL) {
            double s1, c1, s2, c2;
            if constexpr (CACHED) { c1 = tc[0]; s1 = tc[1]; c2 = tc[2]; s2 = tc[3]; }
            else { sincos(qv[0], &s1, &c1); sincos(qv[1], &s2, &c2); tc[0] = c1; tc[1] = s1; tc[2] = c2; tc[3] = s2; }
            R_FM.a[0]= c2;    R_FM.a[1]= 0.0; R_FM.a[2]= s2;
            R_FM.a[3]= s2*s1; R_FM.a[4]= c1;  R_FM.a[5]= -s1*c2;
            R_FM.a[6]=-s2*c1; R_FM.a[7]= s1;  R_FM.a[8]= c1*c2;
            H_FM[0] = {V3(1.0, 0.0, 0.0), zero3<double>()};
            H_FM[1] = {col(R_FM, 1), zero3<double>()};
        } else {
            double oo;
            if constexpr (CACHED) oo = tc[0];
            else { oo = 1.0 / sqrt(qv[0]*qv[0] + qv[1]*qv[1] + qv[2]*qv[2] + qv[3]*qv[3]); tc[0] = oo; }
            R_FM = quat_rot(qv[0]*oo, qv[1]*oo, qv[2]*oo, qv[3]*oo);
            H_FM[0] = {V3(1.0, 0.0, 0.0), zero3<double>()};
            H_FM[1] = {V3(0.0, 1.0, 0.0), zero3<double>()};
            H_FM[2] = {V3(0.0, 0.0, 1.0), zero3<double>()};
            if constexpr (TYPE == FREE) {
                p_FM = V3(qv[4], qv[5], qv[6]);
                H_FM[3] = {zero3<double>(), V3(1.0, 0.0, 0.0)};
                H_FM[4] = {zero3<double>(), V3(0.0, 1.0, 0.0)};
                H_FM[5] = {zero3<double>(), V3(0.0, 0.0, 1.0)};
            }
        }
    }
};

// X_PB = X_PF X_FM X_MB (RigidBodyNodeSpec.h:554-569)
__device__ __forceinline__ void body_in_parent(cdptr P, const dm& R_FM, const d3& p_FM, dm& R_PB, d3& p_PB) {
    const dm R_PF = ldm(P + PRM_R_PF), R_MB = ldm(P + PRM_R_MB);
    const d3 p_PF = ld3(P + PRM_P_PF), p_MB = ld3(P + PRM_P_MB);
    const dm R_FB = mul(R_FM, R_MB);
    const d3 p_FB = p_FM + mul(R_FM, p_MB);
    R_PB = mul(R_PF, R_FB);
    p_PB = p_PF + mul(R_PF, p_FB);
}

// H_PB_G = R_GF (H_FM + H_MB_F)  (RigidBodyNodeSpec.cpp:44-74)
template <int TYPE>
__device__ __forceinline__ void hinge(cdptr P, const Joint<TYPE>& J, const dm& R_GP, dm& R_GF, d3& r_MB_F,
                                      dv* H) {
    R_GF = mul(R_GP, ldm(P + PRM_R_PF));
    r_MB_F = mul(J.R_FM, ld3(P + PRM_P_MB));
#pragma unroll
    for (int j = 0; j < MobC<TYPE>::D; ++j)
        H[j] = {mul(R_GF, J.H_FM[j].w), mul(R_GF, J.H_FM[j].v - cross(r_MB_F, J.H_FM[j].w))};
}

// Mobilizer coriolis acceleration a = (HDot u) + (0, w_GP x (v_GB - v_GP))
// (RigidBodyNodeSpec.cpp:82-129, RigidBodyNode.cpp:98-174).
template <int TYPE>
__device__ __forceinline__ dv coriolis(const Joint<TYPE>& J, const dm& R_GF, const d3& r_MB_F, const dv* H,
                                       const double* uv, const dv& V_GP, const dv& V_GB) {
    constexpr int D = MobC<TYPE>::D;
    dv V_FM = svzero<double>();
#pragma unroll
    for (int j = 0; j < D; ++j) V_FM = V_FM + uv[j]*J.H_FM[j];
    const d3 w_FM = V_FM.w;
    const d3 wxr = cross(w_FM, r_MB_F);
    dv VD = svzero<double>();
#pragma unroll
    for (int j = 0; j < D; ++j) {
        d3 hdw = zero3<double>();
        if constexpr (TYPE == UNIVERSAL) { if (j == 1) hdw = cross(w_FM, col(J.R_FM, 1)); }
        const d3 lin = (zero3<double>() - cross(r_MB_F, hdw)) - cross(wxr, J.H_FM[j].w);
        const dv HD = {mul(R_GF, hdw) + cross(V_GP.w, H[j].w), mul(R_GF, lin) + cross(V_GP.w, H[j].v)};
        VD = VD + uv[j]*HD;
    }
    return {VD.w, VD.v + cross(V_GP.w, V_GB.v - V_GP.v)};
}

// Per-lane SoA rows: row k of this lane at base[k * L] (base = array + lane).
struct Rows {
    double* p; int L;
    __device__ __forceinline__ double& operator()(int k) const { return p[(size_t)k * L]; }
};

// Everything one body's step needs from memory, loaded one body ahead so the
// loads are in flight while the previous body computes.
struct Pre {
    double q[7], u[6], tc[4], g[21];
};
// Raw Runge-Kutta sources of one body's stage input: y0, f0, fa, fb per
// q then u element.
struct Pre1 {
    double s[4][13];
};
// This lane's rows for one Merson attempt.
struct RkRows {
    Rows y0, f0, fa, fb;   // step-start state and stage derivatives
    Rows st;               // stage input (ytmp for stages 1-3, ysave for stage 4)
    Rows out;              // derivative output of this stage
    Rows y1, ye;           // stage 4: y1 and its error estimate
};
// Error-norm partial sums (IntegratorRep.h:454-523): weighted q part of the
// quaternions and of the other q, relative u part, quaternion norm errors.
struct NormSums { double sq_other, sq_quat, su, s2; };
// Type-exact prefetch: NQ q, D u, the trig cache entries the mobilizer uses
// and (pass 3) its 7*D G/DI*eps doubles.
template <int TYPE> struct TcN { static constexpr int n = TYPE == PIN ? 2 : TYPE == UNIVERSAL ? 4 : TYPE == SLIDER ? 0 : 1; };
template <int TYPE, bool G>
__device__ __forceinline__ void prefetch_t(Pre& x, const Rows& y, int nq, int qi, int ui, const Rows& sc, int tcoff,
                                           int gs, bool tc) {
    constexpr int NQ = MobC<TYPE>::NQ, D = MobC<TYPE>::D;
#pragma unroll
    for (int k = 0; k < NQ; ++k) x.q[k] = y(qi + k);
#pragma unroll
    for (int k = 0; k < D; ++k) x.u[k] = y(nq + ui + k);
    if (tc) {
#pragma unroll
        for (int k = 0; k < TcN<TYPE>::n; ++k) x.tc[k] = sc(tcoff + k);
    }
    if constexpr (G && D <= 3) {
#pragma unroll
        for (int k = 0; k < 7*D; ++k) x.g[k] = sc(gs + k);
    }
}
template <bool G>
__device__ __forceinline__ void prefetch(Pre& x, const Rows& y, int nq, int4 ba, const Rows& sc, int tcoff, int gs,
                                         bool tc) {
    switch (ba.y) {
    case PIN:       prefetch_t<PIN, G>(x, y, nq, ba.z, ba.w, sc, tcoff, gs, tc); break;
    case SLIDER:    prefetch_t<SLIDER, G>(x, y, nq, ba.z, ba.w, sc, tcoff, gs, tc); break;
    case UNIVERSAL: prefetch_t<UNIVERSAL, G>(x, y, nq, ba.z, ba.w, sc, tcoff, gs, tc); break;
    case BALL:      prefetch_t<BALL, G>(x, y, nq, ba.z, ba.w, sc, tcoff, gs, tc); break;
    default:        prefetch_t<FREE, G>(x, y, nq, ba.z, ba.w, sc, tcoff, gs, tc); break;
    }
}

// ---------------------------------------------------------------------------
// Pass 1 body: forward transforms and velocities, qdot out, trig cache out.
// ---------------------------------------------------------------------------
template <int TYPE>
__device__ __forceinline__ void pass1_body(cdptr P, const Pre& x, int qi, const Rows& f, const Rows& sc,
                                           int tcoff, dm& R, d3& p, dv& V) {
    constexpr int D = MobC<TYPE>::D;
    const double* qv = x.q; const double* uv = x.u;
    if constexpr (TYPE == BALL || TYPE == FREE) {
        double d4[4]; quat_dot(qv, V3(uv[0], uv[1], uv[2]), d4);
#pragma unroll
        for (int k = 0; k < 4; ++k) f(qi + k) = d4[k];
        if constexpr (TYPE == FREE) { f(qi+4) = uv[3]; f(qi+5) = uv[4]; f(qi+6) = uv[5]; }
    } else {
#pragma unroll
        for (int k = 0; k < D; ++k) f(qi + k) = uv[k];
    }
    Joint<TYPE> J;
    double tc[4];
    J.template set<false>(qv, tc);
    if constexpr (TYPE == PIN) { sc(tcoff) = tc[0]; sc(tcoff+1) = tc[1]; }
    else if constexpr (TYPE == UNIVERSAL) { sc(tcoff) = tc[0]; sc(tcoff+1) = tc[1]; sc(tcoff+2) = tc[2]; sc(tcoff+3) = tc[3]; }
    else if constexpr (TYPE == BALL || TYPE == FREE) { sc(tcoff) = tc[0]; }
    dm R_PB; d3 p_PB;
    body_in_parent(P, J.R_FM, J.p_FM, R_PB, p_PB);
    dm R_GF; d3 r_MB_F; dv H[D];
    hinge<TYPE>(P, J, R, R_GF, r_MB_F, H);
    const d3 phi = mul(R, p_PB);
    dv Vpb = svzero<double>();
#pragma unroll
    for (int j = 0; j < D; ++j) Vpb = Vpb + uv[j]*H[j];
    V = phiT_mul(phi, V) + Vpb;
    p = p + mul(R, p_PB);
    R = mul(R, R_PB);
}

// Small dense inverse of D: cofactor forms for dof<=3 (SmallMatrixMixed.h:
// 895-960), Gauss-Jordan without pivoting for dof=6 (D is SPD).
template <int D> __device__ __forceinline__ void dinv(const double* Dm, double* DI) {
    if constexpr (D == 1) DI[0] = 1.0 / Dm[0];
    else if constexpr (D == 2) {
        const double o = 1.0 / (Dm[0]*Dm[3] - Dm[1]*Dm[2]);
        DI[0] = o*Dm[3]; DI[1] = -(o*Dm[1]); DI[2] = -(o*Dm[2]); DI[3] = o*Dm[0];
    } else if constexpr (D == 3) {
        const double* A = Dm;
        const double d00 = A[4]*A[8]-A[5]*A[7], nd01 = A[5]*A[6]-A[3]*A[8], d02 = A[3]*A[7]-A[4]*A[6];
        const double o = 1.0 / (A[0]*d00 + A[1]*nd01 + A[2]*d02);
        const double nd10 = A[2]*A[7]-A[1]*A[8], d11 = A[0]*A[8]-A[2]*A[6], nd12 = A[1]*A[6]-A[0]*A[7],
                     d20 = A[1]*A[5]-A[2]*A[4], nd21 = A[2]*A[3]-A[0]*A[5], d22 = A[0]*A[4]-A[1]*A[3];
        DI[0]=o*d00;  DI[1]=o*nd10; DI[2]=o*d20;
        DI[3]=o*nd01; DI[4]=o*d11;  DI[5]=o*nd21;
        DI[6]=o*d02;  DI[7]=o*nd12; DI[8]=o*d22;
    } else {
        double M[D*D];
#pragma unroll
        for (int i = 0; i < D*D; ++i) { M[i] = Dm[i]; DI[i] = (i % (D+1) == 0) ? 1.0 : 0.0; }
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double ip = 1.0 / M[k*D+k];
#pragma unroll
            for (int j = 0; j < D; ++j) { M[k*D+j] *= ip; DI[k*D+j] *= ip; }
#pragma unroll
            for (int r = 0; r < D; ++r) {
                if (r == k) continue;
                const double fr = M[r*D+k];
#pragma unroll
                for (int j = 0; j < D; ++j) { M[r*D+j] -= fr*M[k*D+j]; DI[r*D+j] -= fr*DI[k*D+j]; }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pass 2 body: recover the parent's pose/velocity, then ABI + first ABA pass.
// In: R, p, V of this body; child's PP, zp, phic (hasChild).  Out: R, p, V
// of the parent; PP, zp, phic of this body for the parent.
// ---------------------------------------------------------------------------
template <int TYPE>
__device__ __forceinline__ void pass2_body(cdptr P, const Pre& x, const d3& grav, const dv& Fapp,
                                           const Rows& sc, int gs, dm& R, d3& p, dv& V, bool hasChild,
                                           da& PP, dv& zp, d3& phic) {
    constexpr int D = MobC<TYPE>::D;
    const double* uv = x.u;
    Joint<TYPE> J;
    double tc[4] = {x.tc[0], x.tc[1], x.tc[2], x.tc[3]};
    J.template set<true>(x.q, tc);
    dm R_PB; d3 p_PB;
    body_in_parent(P, J.R_FM, J.p_FM, R_PB, p_PB);
    const dm R_GB = R; const d3 p_GB = p; const dv V_GB = V;
    const dm R_GP = mulABt(R_GB, R_PB);
    const d3 phi = mul(R_GP, p_PB);
    const d3 p_GP = p_GB - phi;
    dm R_GF; d3 r_MB_F; dv H[D];
    hinge<TYPE>(P, J, R_GP, R_GF, r_MB_F, H);
    dv Vpb = svzero<double>();
#pragma unroll
    for (int j = 0; j < D; ++j) Vpb = Vpb + uv[j]*H[j];
    const d3 w_GP = V_GB.w - Vpb.w;
    const d3 v_GP = (V_GB.v - Vpb.v) - cross(w_GP, phi);
    const dv V_GP = {w_GP, v_GP};
    const dv a = coriolis<TYPE>(J, R_GF, r_MB_F, H, uv, V_GP, V_GB);
    const double m = P[PRM_MASS];
    const d3 Mk_p = mul(R_GB, ld3(P + PRM_COM));
    const ds Gu = {P[PRM_G], P[PRM_G+1], P[PRM_G+2], P[PRM_G+3], P[PRM_G+4], P[PRM_G+5]};
    const ds Mk_G = rotate_sym(R_GB, Gu);
    const d3 w = V_GB.w;
    const dv bg = {m*cross(w, mul(Mk_G, w)), m*cross(w, cross(w, Mk_p))};
    const d3 fg = m*grav;
    const dv F = dv{cross(Mk_p, fg), fg} + Fapp;
    da Pb = abi_from_rbi(m, Mk_p, Mk_G);
    if (hasChild) abi_add_shifted(Pb, PP, phic);
    dv PH[D];
#pragma unroll
    for (int j = 0; j < D; ++j) PH[j] = abi_mul(Pb, H[j]);
    double Dm[D*D], DI[D*D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int k = 0; k < D; ++k) Dm[r*D+k] = dot(H[r], PH[k]);
    dinv<D>(Dm, DI);
    dv G[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        dv acc = svzero<double>();
#pragma unroll
        for (int j = 0; j < D; ++j) acc = acc + DI[j*D+k]*PH[j];
        G[k] = acc;
    }
    dv z = (abi_mul(Pb, a) + bg) - F;
    if (hasChild) z = z + phi_mul(phic, zp);
    double eps[D];
#pragma unroll
    for (int j = 0; j < D; ++j) eps[j] = 0.0 - dot(H[j], z);
    dv zn = z;
#pragma unroll
    for (int j = 0; j < D; ++j) zn = zn + eps[j]*G[j];
    double mm[9], ms[9], in[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                const double gw = r == 0 ? G[j].w.x : r == 1 ? G[j].w.y : G[j].w.z;
                const double gv = r == 0 ? G[j].v.x : r == 1 ? G[j].v.y : G[j].v.z;
                const double pv = k == 0 ? PH[j].v.x : k == 1 ? PH[j].v.y : PH[j].v.z;
                const double pw = k == 0 ? PH[j].w.x : k == 1 ? PH[j].w.y : PH[j].w.z;
                a0 += gw*pv; a1 += gv*pv; a2 += gw*pw;
            }
            mm[r*3+k] = a0; ms[r*3+k] = a1; in[r*3+k] = a2;
        }
    PP.M = Pb.M - ds{ms[0], ms[4], ms[8], (ms[3]+ms[1])*0.5, (ms[6]+ms[2])*0.5, (ms[7]+ms[5])*0.5};
    PP.J = Pb.J - ds{in[0], in[4], in[8], (in[3]+in[1])*0.5, (in[6]+in[2])*0.5, (in[7]+in[5])*0.5};
#pragma unroll
    for (int i = 0; i < 9; ++i) PP.F.a[i] = Pb.F.a[i] - mm[i];
    zp = zn; phic = phi;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        sc(gs + 6*j + 0) = G[j].w.x; sc(gs + 6*j + 1) = G[j].w.y; sc(gs + 6*j + 2) = G[j].w.z;
        sc(gs + 6*j + 3) = G[j].v.x; sc(gs + 6*j + 4) = G[j].v.y; sc(gs + 6*j + 5) = G[j].v.z;
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
        double sacc = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) sacc += DI[r*D+k]*eps[k];
        sc(gs + 6*D + r) = sacc;
    }
    R = R_GP; p = p_GP; V = V_GP;
}

// ---------------------------------------------------------------------------
// Pass 3 body: forward kinematics again, then udot and A_GB.
// ---------------------------------------------------------------------------
template <int TYPE>
__device__ __forceinline__ void pass3_body(cdptr P, const Pre& x, int ui, const Rows& fu, const Rows& sc,
                                           int gs, dm& R, d3& p, dv& V, dv& A, double* ud_out = nullptr) {
    constexpr int D = MobC<TYPE>::D;
    const double* uv = x.u;
    Joint<TYPE> J;
    double tc[4] = {x.tc[0], x.tc[1], x.tc[2], x.tc[3]};
    J.template set<true>(x.q, tc);
    dm R_PB; d3 p_PB;
    body_in_parent(P, J.R_FM, J.p_FM, R_PB, p_PB);
    dm R_GF; d3 r_MB_F; dv H[D];
    hinge<TYPE>(P, J, R, R_GF, r_MB_F, H);
    const d3 phi = mul(R, p_PB);
    dv Vpb = svzero<double>();
#pragma unroll
    for (int j = 0; j < D; ++j) Vpb = Vpb + uv[j]*H[j];
    const dv V_GP = V;
    const dv V_GB = phiT_mul(phi, V_GP) + Vpb;
    const dv a = coriolis<TYPE>(J, R_GF, r_MB_F, H, uv, V_GP, V_GB);
    const dv Ap = phiT_mul(phi, A);
    double ud[D];
#pragma unroll
    for (int r = 0; r < D; ++r) {
        dv G;
        if constexpr (D <= 3) {
            G = {V3(x.g[6*r], x.g[6*r+1], x.g[6*r+2]), V3(x.g[6*r+3], x.g[6*r+4], x.g[6*r+5])};
            ud[r] = x.g[6*D + r] - dot(G, Ap);
        } else {
            G = {V3(sc(gs + 6*r), sc(gs + 6*r + 1), sc(gs + 6*r + 2)),
                 V3(sc(gs + 6*r + 3), sc(gs + 6*r + 4), sc(gs + 6*r + 5))};
            ud[r] = sc(gs + 6*D + r) - dot(G, Ap);
        }
    }
    dv Hu = svzero<double>();
#pragma unroll
    for (int j = 0; j < D; ++j) Hu = Hu + ud[j]*H[j];
    A = (Ap + Hu) + a;
#pragma unroll
    for (int j = 0; j < D; ++j) { fu(ui + j) = ud[j]; if (ud_out) ud_out[j] = ud[j]; }
    p = p + mul(R, p_PB);
    R = mul(R, R_PB);
    V = V_GB;
}

// ---------------------------------------------------------------------------
// The evaluator.  y, f: this lane's rows of the stage input and output
// vectors (q rows then u rows); sc: this lane's scratch rows (G records at
// body_b.w, trig cache at tree_b.z + 4(b-1)).
// ---------------------------------------------------------------------------
struct ChainEval {
    cdptr prm; const int4* body_a; const int4* body_b;
    int nb, nq, tcbase;
    Rows sc; d3 grav; const double* fapp;

    __device__ ChainEval(const RkmArgs& a, int lane, int tr) {
        const int4 ta = ldi4(a.T.tree_a, tr);
        nb = ta.x; nq = ta.y;
        prm = cptr(a.T.prm) + (size_t)ta.w * PRM_STRIDE;
        body_a = a.T.body_a + ta.w; body_b = a.T.body_b + ta.w;
        tcbase = ldi4(a.T.tree_b, tr).z;
        sc = {a.scratch + lane, a.L};
        const cdptr g = cptr(a.T.gravity) + 4*tr;
        grav = V3(g[0], g[1], g[2]);
        fapp = a.fapp ? a.fapp + 6*(size_t)a.off.b_off[lane] : nullptr;
    }
    __device__ __forceinline__ dv F(int b) const {
        if (!fapp) return svzero<double>();
        const double* fb = fapp + 6*b;
        return {V3(fb[0], fb[1], fb[2]), V3(fb[3], fb[4], fb[5])};
    }
    // A Free mobilizer can only connect body 1 to Ground (the host rejects it
    // elsewhere), so body 1 is peeled out of the loops when it is Free: the
    // 6-dof body's register demand then never overlaps the loop's.
    __device__ __forceinline__ void derivs(const Rows& y, const Rows& f) const {
        const Rows fu{f.p + (size_t)nq * f.L, f.L};
        Pre cur, nxt;
        const int4 a0 = ldi4(body_a, 0);
        const bool freeRoot = a0.y == FREE;
        // pass 1: base to tip
        dm R = ident<double>(); d3 p = zero3<double>(); dv V = svzero<double>();
        prefetch<false>(cur, y, nq, a0, sc, tcbase, 0, false);
        int bstart = 1;
        if (freeRoot) {
            pass1_body<FREE>(prm, cur, a0.z, f, sc, tcbase, R, p, V);
            bstart = 2;
            if (nb >= 2) { const int4 bn = ldi4(body_a, 1); prefetch<false>(cur, y, nq, bn, sc, tcbase + 4, 0, false); }
        }
        int4 ba = ldi4(body_a, bstart-1);
        for (int b = bstart; b <= nb; ++b) {
            int4 bn = ba;
            if (b < nb) { bn = ldi4(body_a, b); prefetch<false>(nxt, y, nq, bn, sc, tcbase + 4*b, 0, false); }
            const cdptr P = prm + (size_t)(b-1) * PRM_STRIDE;
            const int tco = tcbase + 4*(b-1);
            switch (ba.y) {
            case PIN:       pass1_body<PIN>(P, cur, ba.z, f, sc, tco, R, p, V); break;
            case SLIDER:    pass1_body<SLIDER>(P, cur, ba.z, f, sc, tco, R, p, V); break;
            case UNIVERSAL: pass1_body<UNIVERSAL>(P, cur, ba.z, f, sc, tco, R, p, V); break;
            default:        pass1_body<BALL>(P, cur, ba.z, f, sc, tco, R, p, V); break;
            }
            cur = nxt; ba = bn;
        }
        // pass 2: tip to base
        da PP; dv zp = svzero<double>(); d3 phic = zero3<double>();
        ba = ldi4(body_a, nb-1);
        prefetch<false>(cur, y, nq, ba, sc, tcbase + 4*(nb-1), 0, true);
        const int bstop = freeRoot ? 2 : 1;
        for (int b = nb; b >= bstop; --b) {
            int4 bn = ba;
            if (b > 1) { bn = ldi4(body_a, b-2); prefetch<false>(nxt, y, nq, bn, sc, tcbase + 4*(b-2), 0, true); }
            const int gs = ldi4(body_b, b-1).w;
            const cdptr P = prm + (size_t)(b-1) * PRM_STRIDE;
            const bool hc = b < nb;
            const dv Fb = F(b);
            switch (ba.y) {
            case PIN:       pass2_body<PIN>(P, cur, grav, Fb, sc, gs, R, p, V, hc, PP, zp, phic); break;
            case SLIDER:    pass2_body<SLIDER>(P, cur, grav, Fb, sc, gs, R, p, V, hc, PP, zp, phic); break;
            case UNIVERSAL: pass2_body<UNIVERSAL>(P, cur, grav, Fb, sc, gs, R, p, V, hc, PP, zp, phic); break;
            default:        pass2_body<BALL>(P, cur, grav, Fb, sc, gs, R, p, V, hc, PP, zp, phic); break;
            }
            cur = nxt; ba = bn;
        }
        if (freeRoot) pass2_body<FREE>(prm, cur, grav, F(1), sc, ldi4(body_b, 0).w, R, p, V, nb > 1, PP, zp, phic);
        // pass 3: base to tip
        R = ident<double>(); p = zero3<double>(); V = svzero<double>();
        dv A = svzero<double>();
        prefetch<true>(cur, y, nq, a0, sc, tcbase, ldi4(body_b, 0).w, true);
        if (freeRoot) {
            pass3_body<FREE>(prm, cur, a0.w, fu, sc, ldi4(body_b, 0).w, R, p, V, A);
            if (nb >= 2) { const int4 bn = ldi4(body_a, 1); prefetch<true>(cur, y, nq, bn, sc, tcbase + 4, ldi4(body_b, 1).w, true); }
        }
        ba = ldi4(body_a, bstart-1);
        for (int b = bstart; b <= nb; ++b) {
            int4 bn = ba;
            if (b < nb) { bn = ldi4(body_a, b); prefetch<true>(nxt, y, nq, bn, sc, tcbase + 4*b, ldi4(body_b, b).w, true); }
            const int gs = ldi4(body_b, b-1).w;
            const cdptr P = prm + (size_t)(b-1) * PRM_STRIDE;
            switch (ba.y) {
            case PIN:       pass3_body<PIN>(P, cur, ba.w, fu, sc, gs, R, p, V, A); break;
            case SLIDER:    pass3_body<SLIDER>(P, cur, ba.w, fu, sc, gs, R, p, V, A); break;
            case UNIVERSAL: pass3_body<UNIVERSAL>(P, cur, ba.w, fu, sc, gs, R, p, V, A); break;
            default:        pass3_body<BALL>(P, cur, ba.w, fu, sc, gs, R, p, V, A); break;
            }
            cur = nxt; ba = bn;
        }
    }

    // ---- Runge-Kutta-Merson stages fused into the sweeps -------------------
    // STAGE 0 evaluates f0 = f(y0).  STAGES 1-4 form the Merson stage input
    // (RungeKuttaMersonIntegrator.cpp:100-112) body by body while loading it
    // in pass 1 and store it to r.st for passes 2 and 3; STAGE 4 also forms
    // y1 = y0 + h/6 (f0 + 4 f3 + f4) and the error estimate 0.2|y1 - ysave|
    // (:114-137) in pass 3 and accumulates the error-norm sums.
    template <int TYPE, int STAGE>
    __device__ __forceinline__ void pre1(Pre1& x, const RkRows& r, int qi, int ui) const {
        constexpr int NQ = MobC<TYPE>::NQ, D = MobC<TYPE>::D;
#pragma unroll
        for (int k = 0; k < NQ + D; ++k) {
            const int idx = k < NQ ? qi + k : nq + ui + (k - NQ);
            x.s[0][k] = r.y0(idx);
            if constexpr (STAGE >= 1) x.s[1][k] = r.f0(idx);
            if constexpr (STAGE >= 2) x.s[2][k] = r.fa(idx);
            if constexpr (STAGE >= 4) x.s[3][k] = r.fb(idx);
        }
    }
    template <int STAGE>
    __device__ __forceinline__ void pre1_any(Pre1& x, const RkRows& r, int4 ba) const {
        switch (ba.y) {
        case PIN:       pre1<PIN, STAGE>(x, r, ba.z, ba.w); break;
        case SLIDER:    pre1<SLIDER, STAGE>(x, r, ba.z, ba.w); break;
        case UNIVERSAL: pre1<UNIVERSAL, STAGE>(x, r, ba.z, ba.w); break;
        case BALL:      pre1<BALL, STAGE>(x, r, ba.z, ba.w); break;
        default:        pre1<FREE, STAGE>(x, r, ba.z, ba.w); break;
        }
    }
    // stage input from the raw sources; stored to r.st (stages 1-4)
    template <int TYPE, int STAGE>
    __device__ __forceinline__ void form(const Pre1& x, Pre& o, const RkRows& r, int qi, int ui, double h) const {
        constexpr int NQ = MobC<TYPE>::NQ, D = MobC<TYPE>::D;
#pragma unroll
        for (int k = 0; k < NQ + D; ++k) {
            double v;
            if constexpr (STAGE == 0) v = x.s[0][k];
            else if constexpr (STAGE == 1) v = x.s[0][k] + (h/3)*x.s[1][k];
            else if constexpr (STAGE == 2) v = x.s[0][k] + (h/6)*(x.s[1][k] + x.s[2][k]);
            else if constexpr (STAGE == 3) v = x.s[0][k] + (h/8)*(x.s[1][k] + 3*x.s[2][k]);
            else v = x.s[0][k] + (h/2)*((x.s[1][k] - 3*x.s[2][k]) + 4*x.s[3][k]);
            if constexpr (STAGE >= 1) r.st(k < NQ ? qi + k : nq + ui + (k - NQ)) = v;
            if (k < NQ) o.q[k] = v; else o.u[k - NQ] = v;
        }
    }
    // STAGE 4, pass 3 epilogue of one body: y1, yerr and the norm sums.
    template <int TYPE>
    __device__ __forceinline__ void combine(const Pre& x, const double* y0, const double* f0, const double* fb,
                                            const double* ud, const RkRows& r, int qi, int ui, double h,
                                            NormSums& ns) const {
        constexpr int NQ = MobC<TYPE>::NQ, D = MobC<TYPE>::D;
        double f4[NQ + D];
        if constexpr (TYPE == BALL || TYPE == FREE) {
            quat_dot(x.q, V3(x.u[0], x.u[1], x.u[2]), f4);
            if constexpr (TYPE == FREE) { f4[4] = x.u[3]; f4[5] = x.u[4]; f4[6] = x.u[5]; }
        } else {
#pragma unroll
            for (int k = 0; k < D; ++k) f4[k] = x.u[k];
        }
#pragma unroll
        for (int k = 0; k < D; ++k) f4[NQ + k] = ud[k];
        double y1[NQ + D], e[NQ + D];
#pragma unroll
        for (int k = 0; k < NQ + D; ++k) {
            const double ys = k < NQ ? x.q[k] : x.u[k - NQ];
            y1[k] = y0[k] + (h/6)*((f0[k] + 4*fb[k]) + f4[k]);
            e[k] = 0.2 * fabs(y1[k] - ys);
            const int idx = k < NQ ? qi + k : nq + ui + (k - NQ);
            r.y1(idx) = y1[k]; r.ye(idx) = e[k];
        }
        if constexpr (TYPE == BALL || TYPE == FREE) {
            const d3 w = quat_dot_to_w(y1, e);
            double d4[4]; quat_dot(y1, w, d4);
            ns.sq_quat += d4[0]*d4[0] + d4[1]*d4[1] + d4[2]*d4[2] + d4[3]*d4[3];
            if constexpr (TYPE == FREE) ns.sq_other += e[4]*e[4] + e[5]*e[5] + e[6]*e[6];
            const double qe = sqrt(y1[0]*y1[0] + y1[1]*y1[1] + y1[2]*y1[2] + y1[3]*y1[3]) - 1.0;
            ns.s2 += qe*qe;
        } else {
#pragma unroll
            for (int k = 0; k < NQ; ++k) ns.sq_other += e[k]*e[k];
        }
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double vi = fabs(y0[NQ + k]);
            const double us = vi > 1.0 ? 1.0/vi : 1.0;
            const double xx = us * e[NQ + k];
            ns.su += xx*xx;
        }
    }

    template <int TYPE, int STAGE>
    __device__ __forceinline__ void p1(const Pre1& raw, const RkRows& r, const int4 ba, int b, double h,
                                       dm& R, d3& p, dv& V) const {
        Pre x;
        form<TYPE, STAGE>(raw, x, r, ba.z, ba.w, h);
        pass1_body<TYPE>(prm + (size_t)(b-1) * PRM_STRIDE, x, ba.z, r.out, sc, tcbase + 4*(b-1), R, p, V);
    }
    template <int TYPE, int STAGE>
    __device__ __forceinline__ void p3(const Pre& x, const RkRows& r, const Rows& fu, const int4 ba, int b, double h,
                                       dm& R, d3& p, dv& V, dv& A, NormSums& ns) const {
        constexpr int NQ = MobC<TYPE>::NQ, D = MobC<TYPE>::D;
        double y0[NQ + D], f0[NQ + D], fb[NQ + D];
        if constexpr (STAGE == 4) {
#pragma unroll
            for (int k = 0; k < NQ + D; ++k) {
                const int idx = k < NQ ? ba.z + k : nq + ba.w + (k - NQ);
                y0[k] = r.y0(idx); f0[k] = r.f0(idx); fb[k] = r.fb(idx);
            }
        }
        double ud[D];
        pass3_body<TYPE>(prm + (size_t)(b-1) * PRM_STRIDE, x, ba.w, fu, sc, ldi4(body_b, b-1).w, R, p, V, A, ud);
        if constexpr (STAGE == 4) combine<TYPE>(x, y0, f0, fb, ud, r, ba.z, ba.w, h, ns);
    }

    template <int STAGE>
    __device__ __forceinline__ void derivs_rk(const RkRows& r, double h, NormSums& ns) const {
        const Rows in = STAGE == 0 ? r.y0 : r.st;
        const Rows fu{r.out.p + (size_t)nq * r.out.L, r.out.L};
        const int4 a0 = ldi4(body_a, 0);
        const bool freeRoot = a0.y == FREE;
        // pass 1: base to tip (stage input formed here)
        dm R = ident<double>(); d3 p = zero3<double>(); dv V = svzero<double>();
        {
            Pre1 c1, n1;
            pre1_any<STAGE>(c1, r, a0);
            int bstart = 1;
            if (freeRoot) {
                p1<FREE, STAGE>(c1, r, a0, 1, h, R, p, V);
                bstart = 2;
                if (nb >= 2) pre1_any<STAGE>(c1, r, ldi4(body_a, 1));
            }
            int4 ba = ldi4(body_a, bstart-1);
            for (int b = bstart; b <= nb; ++b) {
                int4 bn = ba;
                if (b < nb) { bn = ldi4(body_a, b); pre1_any<STAGE>(n1, r, bn); }
                switch (ba.y) {
                case PIN:       p1<PIN, STAGE>(c1, r, ba, b, h, R, p, V); break;
                case SLIDER:    p1<SLIDER, STAGE>(c1, r, ba, b, h, R, p, V); break;
                case UNIVERSAL: p1<UNIVERSAL, STAGE>(c1, r, ba, b, h, R, p, V); break;
                default:        p1<BALL, STAGE>(c1, r, ba, b, h, R, p, V); break;
                }
                c1 = n1; ba = bn;
            }
        }
        // pass 2: tip to base
        Pre cur, nxt;
        da PP; dv zp = svzero<double>(); d3 phic = zero3<double>();
        int4 ba = ldi4(body_a, nb-1);
        prefetch<false>(cur, in, nq, ba, sc, tcbase + 4*(nb-1), 0, true);
        const int bstop = freeRoot ? 2 : 1;
        for (int b = nb; b >= bstop; --b) {
            int4 bn = ba;
            if (b > 1) { bn = ldi4(body_a, b-2); prefetch<false>(nxt, in, nq, bn, sc, tcbase + 4*(b-2), 0, true); }
            const int gs = ldi4(body_b, b-1).w;
            const cdptr P = prm + (size_t)(b-1) * PRM_STRIDE;
            const bool hc = b < nb;
            const dv Fb = F(b);
            switch (ba.y) {
            case PIN:       pass2_body<PIN>(P, cur, grav, Fb, sc, gs, R, p, V, hc, PP, zp, phic); break;
            case SLIDER:    pass2_body<SLIDER>(P, cur, grav, Fb, sc, gs, R, p, V, hc, PP, zp, phic); break;
            case UNIVERSAL: pass2_body<UNIVERSAL>(P, cur, grav, Fb, sc, gs, R, p, V, hc, PP, zp, phic); break;
            default:        pass2_body<BALL>(P, cur, grav, Fb, sc, gs, R, p, V, hc, PP, zp, phic); break;
            }
            cur = nxt; ba = bn;
        }
        if (freeRoot) pass2_body<FREE>(prm, cur, grav, F(1), sc, ldi4(body_b, 0).w, R, p, V, nb > 1, PP, zp, phic);
        // pass 3: base to tip
        R = ident<double>(); p = zero3<double>(); V = svzero<double>();
        dv A = svzero<double>();
        prefetch<true>(cur, in, nq, a0, sc, tcbase, ldi4(body_b, 0).w, true);
        int bstart = 1;
        if (freeRoot) {
            p3<FREE, STAGE>(cur, r, fu, a0, 1, h, R, p, V, A, ns);
            bstart = 2;
            if (nb >= 2) prefetch<true>(cur, in, nq, ldi4(body_a, 1), sc, tcbase + 4, ldi4(body_b, 1).w, true);
        }
        ba = ldi4(body_a, bstart-1);
        for (int b = bstart; b <= nb; ++b) {
            int4 bn = ba;
            if (b < nb) { bn = ldi4(body_a, b); prefetch<true>(nxt, in, nq, bn, sc, tcbase + 4*b, ldi4(body_b, b).w, true); }
            switch (ba.y) {
            case PIN:       p3<PIN, STAGE>(cur, r, fu, ba, b, h, R, p, V, A, ns); break;
            case SLIDER:    p3<SLIDER, STAGE>(cur, r, fu, ba, b, h, R, p, V, A, ns); break;
            case UNIVERSAL: p3<UNIVERSAL, STAGE>(cur, r, fu, ba, b, h, R, p, V, A, ns); break;
            default:        p3<BALL, STAGE>(cur, r, fu, ba, b, h, R, p, V, A, ns); break;
            }
            cur = nxt; ba = bn;
        }
    }
};

} // namespace chain
} // namespace tdyn
