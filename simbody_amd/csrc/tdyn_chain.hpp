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

__device__ __forceinline__ dm ldm(const double* p) { dm r; for (int i = 0; i < 9; ++i) r.a[i] = p[i]; return r; }
__device__ __forceinline__ d3 ld3(const double* p) { return {p[0], p[1], p[2]}; }
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
    __device__ __forceinline__ void set(const double* qv) {
        p_FM = zero3<double>();
        if constexpr (TYPE == PIN) {
            double sn, cs; sincos(qv[0], &sn, &cs);
            R_FM.a[0]=cs; R_FM.a[1]=-sn; R_FM.a[2]=0; R_FM.a[3]=sn; R_FM.a[4]=cs; R_FM.a[5]=0;
            R_FM.a[6]=0;  R_FM.a[7]=0;   R_FM.a[8]=1;
            H_FM[0] = {V3(0.0, 0.0, 1.0), zero3<double>()};
        } else if constexpr (TYPE == SLIDER) {
            R_FM = ident<double>(); p_FM = V3(qv[0], 0.0, 0.0);
            H_FM[0] = {zero3<double>(), V3(1.0, 0.0, 0.0)};
        } else if constexpr (TYPE == UNIVERSAL) {
            double s1, c1, s2, c2; sincos(qv[0], &s1, &c1); sincos(qv[1], &s2, &c2);
            R_FM.a[0]= c2;    R_FM.a[1]= 0.0; R_FM.a[2]= s2;
            R_FM.a[3]= s2*s1; R_FM.a[4]= c1;  R_FM.a[5]= -s1*c2;
            R_FM.a[6]=-s2*c1; R_FM.a[7]= s1;  R_FM.a[8]= c1*c2;
            H_FM[0] = {V3(1.0, 0.0, 0.0), zero3<double>()};
            H_FM[1] = {col(R_FM, 1), zero3<double>()};
        } else {
            const double len = sqrt(qv[0]*qv[0] + qv[1]*qv[1] + qv[2]*qv[2] + qv[3]*qv[3]);
            const double oo = 1.0 / len;
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
__device__ __forceinline__ void body_in_parent(const double* P, const dm& R_FM, const d3& p_FM, dm& R_PB, d3& p_PB) {
    const dm R_PF = ldm(P + PRM_R_PF), R_MB = ldm(P + PRM_R_MB);
    const d3 p_PF = ld3(P + PRM_P_PF), p_MB = ld3(P + PRM_P_MB);
    const dm R_FB = mul(R_FM, R_MB);
    const d3 p_FB = p_FM + mul(R_FM, p_MB);
    R_PB = mul(R_PF, R_FB);
    p_PB = p_PF + mul(R_PF, p_FB);
}

// H_PB_G = R_GF (H_FM + H_MB_F)  (RigidBodyNodeSpec.cpp:44-74)
template <int TYPE>
__device__ __forceinline__ void hinge(const double* P, const Joint<TYPE>& J, const dm& R_GP, dm& R_GF, d3& r_MB_F,
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

// G/DI*eps record of body b in the chain scratch: 7*dof doubles at slot gs.
struct GRec {
    double* p; int L, lane;
    __device__ __forceinline__ double& operator()(int k) const { return p[(size_t)k * L + lane]; }
};

// ---------------------------------------------------------------------------
// Pass 1 body: forward transforms and velocities, qdot out.
// ---------------------------------------------------------------------------
template <int TYPE, class QG, class UG, class QD>
__device__ __forceinline__ void pass1_body(const double* P, int qi, int ui, const QG& q, const UG& u, const QD& qd,
                                           dm& R, d3& p, dv& V) {
    constexpr int D = MobC<TYPE>::D, NQ = MobC<TYPE>::NQ;
    double qv[NQ], uv[D];
#pragma unroll
    for (int k = 0; k < NQ; ++k) qv[k] = q(qi + k);
#pragma unroll
    for (int k = 0; k < D; ++k) uv[k] = u(ui + k);
    if constexpr (TYPE == BALL || TYPE == FREE) {
        double d4[4]; quat_dot(qv, V3(uv[0], uv[1], uv[2]), d4);
#pragma unroll
        for (int k = 0; k < 4; ++k) qd(qi + k, d4[k]);
        if constexpr (TYPE == FREE) { qd(qi+4, uv[3]); qd(qi+5, uv[4]); qd(qi+6, uv[5]); }
    } else {
#pragma unroll
        for (int k = 0; k < D; ++k) qd(qi + k, uv[k]);
    }
    Joint<TYPE> J; J.set(qv);
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

// ---------------------------------------------------------------------------
// Pass 2 body: recover the parent's pose/velocity, then ABI + first ABA pass.
// In: R, p, V of this body; child's PP, zp, phic (hasChild).  Out: R, p, V
// of the parent; PP, zp, phic of this body for the parent.
// ---------------------------------------------------------------------------
template <int TYPE, class QG, class UG>
__device__ __forceinline__ void pass2_body(const double* P, int qi, int ui, const QG& q, const UG& u,
                                           const d3& grav, const dv& Fapp, const GRec& gr, int gs,
                                           dm& R, d3& p, dv& V, bool hasChild, da& PP, dv& zp, d3& phic) {
    constexpr int D = MobC<TYPE>::D, NQ = MobC<TYPE>::NQ;
    double qv[NQ], uv[D];
#pragma unroll
    for (int k = 0; k < NQ; ++k) qv[k] = q(qi + k);
#pragma unroll
    for (int k = 0; k < D; ++k) uv[k] = u(ui + k);
    Joint<TYPE> J; J.set(qv);
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
    // spatial inertia in G, gyroscopic force, gravity
    const double m = P[PRM_MASS];
    const d3 Mk_p = mul(R_GB, ld3(P + PRM_COM));
    const ds Gu = {P[PRM_G], P[PRM_G+1], P[PRM_G+2], P[PRM_G+3], P[PRM_G+4], P[PRM_G+5]};
    const ds Mk_G = rotate_sym(R_GB, Gu);
    const d3 w = V_GB.w;
    const dv bg = {m*cross(w, mul(Mk_G, w)), m*cross(w, cross(w, Mk_p))};
    const d3 fg = m*grav;
    const dv F = dv{cross(Mk_p, fg), fg} + Fapp;
    // articulated-body inertia
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
                const double f = M[r*D+k];
#pragma unroll
                for (int j = 0; j < D; ++j) { M[r*D+j] -= f*M[k*D+j]; DI[r*D+j] -= f*DI[k*D+j]; }
            }
        }
    }
    dv G[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        dv acc = svzero<double>();
#pragma unroll
        for (int j = 0; j < D; ++j) acc = acc + DI[j*D+k]*PH[j];
        G[k] = acc;
    }
    // z = P a + b - F + Phi_child z+_child ; eps = -H^T z ; z+ = z + G eps
    dv z = (abi_mul(Pb, a) + bg) - F;
    if (hasChild) z = z + phi_mul(phic, zp);
    double eps[D];
#pragma unroll
    for (int j = 0; j < D; ++j) eps[j] = 0.0 - dot(H[j], z);
    dv zn = z;
#pragma unroll
    for (int j = 0; j < D; ++j) zn = zn + eps[j]*G[j];
    // P+ = P - G PH^T (mass and inertia blocks symmetrized)
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
    // G and DI*eps for the outward pass
#pragma unroll
    for (int j = 0; j < D; ++j) {
        gr(gs + 6*j + 0) = G[j].w.x; gr(gs + 6*j + 1) = G[j].w.y; gr(gs + 6*j + 2) = G[j].w.z;
        gr(gs + 6*j + 3) = G[j].v.x; gr(gs + 6*j + 4) = G[j].v.y; gr(gs + 6*j + 5) = G[j].v.z;
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) s += DI[r*D+k]*eps[k];
        gr(gs + 6*D + r) = s;
    }
    R = R_GP; p = p_GP; V = V_GP;
}

// ---------------------------------------------------------------------------
// Pass 3 body: forward kinematics again, then udot and A_GB.
// ---------------------------------------------------------------------------
template <int TYPE, class QG, class UG, class UO>
__device__ __forceinline__ void pass3_body(const double* P, int qi, int ui, const QG& q, const UG& u, const UO& uo,
                                           const GRec& gr, int gs, dm& R, d3& p, dv& V, dv& A) {
    constexpr int D = MobC<TYPE>::D, NQ = MobC<TYPE>::NQ;
    double qv[NQ], uv[D];
#pragma unroll
    for (int k = 0; k < NQ; ++k) qv[k] = q(qi + k);
#pragma unroll
    for (int k = 0; k < D; ++k) uv[k] = u(ui + k);
    Joint<TYPE> J; J.set(qv);
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
        const dv G = {V3(gr(gs + 6*r), gr(gs + 6*r + 1), gr(gs + 6*r + 2)),
                      V3(gr(gs + 6*r + 3), gr(gs + 6*r + 4), gr(gs + 6*r + 5))};
        ud[r] = gr(gs + 6*D + r) - dot(G, Ap);
    }
    dv Hu = svzero<double>();
#pragma unroll
    for (int j = 0; j < D; ++j) Hu = Hu + ud[j]*H[j];
    A = (Ap + Hu) + a;
#pragma unroll
    for (int j = 0; j < D; ++j) uo(ui + j, ud[j]);
    p = p + mul(R, p_PB);
    R = mul(R, R_PB);
    V = V_GB;
}

} // namespace chain
} // namespace tdyn
