// Tree-dynamics kernels for gfx950 (MI355X), FP64.
//
// Mapping: one lane (work-item) per multibody instance.  Each lane walks its
// own tree in body-number order (a valid base-to-tip order, parents are
// numbered before children) and in reverse (tip-to-base).  Per-body
// intermediates that one sweep hands to the next live in a per-lane record
// in HBM, stored structure-of-arrays across lanes so that lanes running the
// same tree issue coalesced loads and stores.  The body loops are the
// level sweeps of Simbody's SimbodyMatterSubsystemRep (see the per-routine
// citations below); instances are independent, which is the parallelism the
// reference never exploits.
#include "tdyn_internal.hpp"
#include "tdyn_math.hpp"
#include "tdyn_chain.hpp"

namespace tdyn {

using d3 = v3<double>;
using dm = m33<double>;
using ds = s33<double>;
using dv = sv<double>;
using da = abi<double>;

// One lane's view of its instance.  Record storage: structure-of-arrays
// (row off of every lane contiguous: scr = scratch + lane, rs = L) when all
// lanes walk the same tree, so equal offsets coalesce; per-lane records
// (scr = scratch + lane * record stride, rs = 1) when lanes walk different
// trees, so that reading a parent's or a child's record, whose offset differs
// from lane to lane, touches a few cache lines per lane instead of one per
// value.
struct Lane {
    DevTables T;
    int lane, L, nb, nq, nu, body_off;
    double* scr; int rs;
    __device__ double& s(int off) const { return scr[(size_t)off * rs]; }
    __device__ int4 ba(int b) const { return ldi4(T.body_a, body_off + b - 1); }
    __device__ int4 bb(int b) const { return ldi4(T.body_b, body_off + b - 1); }
    __device__ cdptr prm(int b) const { return cptr(T.prm) + (size_t)(body_off + b - 1) * PRM_STRIDE; }
    __device__ int rec(int b) const { return ldi4(T.body_b, body_off + b - 1).z; }
};

__device__ __forceinline__ Lane make_lane_t(const DevTables& T, int tr, int lane, int L, double* scr,
                                           int per_lane_stride = 0) {
    Lane c;
    c.T = T; c.lane = lane; c.L = L;
    if (per_lane_stride > 0) { c.scr = scr + (size_t)lane * per_lane_stride; c.rs = 1; }
    else { c.scr = scr + lane; c.rs = L; }
    const int4 ta = ldi4(T.tree_a, tr);
    c.nb = ta.x; c.nq = ta.y; c.nu = ta.z; c.body_off = ta.w;
    return c;
}
__device__ __forceinline__ Lane make_lane(const DevTables& T, const int* tree_of, int lane, int L, double* scr) {
    Lane c;
    c.T = T; c.lane = lane; c.L = L; c.scr = scr + lane; c.rs = L;
    const int tr = tree_of ? tree_of[lane] : 0;
    const int4 ta = T.tree_a[tr];
    c.nb = ta.x; c.nq = ta.y; c.nu = ta.z; c.body_off = ta.w;
    return c;
}

// ---- record load/store --------------------------------------------------
__device__ __forceinline__ d3 ld3(const Lane& c, int o) { return {c.s(o), c.s(o+1), c.s(o+2)}; }
__device__ __forceinline__ void st3(const Lane& c, int o, d3 v) { c.s(o) = v.x; c.s(o+1) = v.y; c.s(o+2) = v.z; }
__device__ __forceinline__ dv ldsv(const Lane& c, int o) { return {ld3(c, o), ld3(c, o+3)}; }
__device__ __forceinline__ void stsv(const Lane& c, int o, const dv& v) { st3(c, o, v.w); st3(c, o+3, v.v); }
__device__ __forceinline__ dm ldm(const Lane& c, int o) { dm r; for (int i=0;i<9;++i) r.a[i] = c.s(o+i); return r; }
__device__ __forceinline__ void stm(const Lane& c, int o, const dm& r) { for (int i=0;i<9;++i) c.s(o+i) = r.a[i]; }
__device__ __forceinline__ ds lds(const Lane& c, int o) { return {c.s(o), c.s(o+1), c.s(o+2), c.s(o+3), c.s(o+4), c.s(o+5)}; }
__device__ __forceinline__ void sts(const Lane& c, int o, const ds& v) {
    c.s(o) = v.xx; c.s(o+1) = v.yy; c.s(o+2) = v.zz; c.s(o+3) = v.xy; c.s(o+4) = v.xz; c.s(o+5) = v.yz;
}
__device__ __forceinline__ da lda(const Lane& c, int o) { da P; P.M = lds(c, o); P.J = lds(c, o+6); P.F = ldm(c, o+12); return P; }
__device__ __forceinline__ void sta(const Lane& c, int o, const da& P) { sts(c, o, P.M); sts(c, o+6, P.J); stm(c, o+12, P.F); }
__device__ __forceinline__ void sta_global(double* o, const da& P) {
    o[0]=P.M.xx; o[1]=P.M.yy; o[2]=P.M.zz; o[3]=P.M.xy; o[4]=P.M.xz; o[5]=P.M.yz;
    o[6]=P.J.xx; o[7]=P.J.yy; o[8]=P.J.zz; o[9]=P.J.xy; o[10]=P.J.xz; o[11]=P.J.yz;
    for (int i=0;i<9;++i) o[12+i] = P.F.a[i];
}

__device__ __forceinline__ dm prm_m(cdptr p) { dm r; for (int i=0;i<9;++i) r.a[i] = p[i]; return r; }
__device__ __forceinline__ d3 prm_3(cdptr p) { return {p[0], p[1], p[2]}; }

template <int TYPE> struct Mob {
    static constexpr int D = TYPE <= SLIDER ? 1 : TYPE == UNIVERSAL ? 2 : TYPE == BALL ? 3 : 6;
    static constexpr int NQ = TYPE <= SLIDER ? 1 : TYPE == UNIVERSAL ? 2 : TYPE == BALL ? 4 : 7;
};

// ---------------------------------------------------------------------------
// Base-to-tip position (+ velocity) kinematics of one body.
// Restates RigidBodyNodeSpec<dof>::realizePosition/realizeVelocity
// (RigidBodyNodeSpec.h:229-333) with the mobilizer parts of
// RigidBodyNodeSpec_{Pin,Slider,Universal,Ball,Free}.h, calcBodyTransforms
// (RigidBodyNodeSpec.h:554-569), H_PB_G and its derivative
// (RigidBodyNodeSpec.cpp:44-129) and the joint-independent parts
// (RigidBodyNode.cpp:54-174).
// q(k), u(k): this instance's k-th q/u.  qd(k, v): store qdot.
// ---------------------------------------------------------------------------
template <int TYPE, bool VEL, class QG, class UG, class QD>
__device__ __forceinline__ void kin_body(const Lane& c, int b, const int4 ba, const QG& q, const UG& u, const QD& qd) {
    constexpr int D = Mob<TYPE>::D, NQ = Mob<TYPE>::NQ;
    const int p = ba.x, qi = ba.z, ui = ba.w;
    const cdptr P = c.prm(b);
    const int rb = c.rec(b);
    dm R_GP; d3 p_GP;
    int rp = 0;
    if (p == 0) { R_GP = ident<double>(); p_GP = zero3<double>(); }
    else { rp = c.rec(p); R_GP = ldm(c, rp + REC_R); p_GP = ld3(c, rp + REC_P); }

    // Ball/Free q as x-y-z angles (Euler-angle mode; PRM_EULER 2: a Bushing,
    // six q, u = qdot, RigidBodyNodeSpec_Bushing.h:152-220); ec: their
    // c0 s0 c1 s1
    const bool eul = (TYPE == BALL || TYPE == FREE) && P[PRM_EULER] != 0.0;
    const bool bush = TYPE == FREE && P[PRM_EULER] == 2.0;
    double qv[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) qv[k] = (TYPE == FREE && k == 6 && bush) ? 0.0 : q(qi + k);
    dm R_FM; d3 p_FM = zero3<double>();
    dv H_FM[D];
    double ec[4] = {1.0, 0.0, 1.0, 0.0};
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
    } else {   // BALL, FREE
        const double len = sqrt(qv[0]*qv[0] + qv[1]*qv[1] + qv[2]*qv[2] + qv[3]*qv[3]);
        const double oo = 1.0 / len;
        if (eul) {
            // x-y-z angles, Free translation in q3..q5 (RigidBodyNodeSpec_Free.h:139-188)
            double s2, c2;
            sincos(qv[0], &ec[1], &ec[0]); sincos(qv[1], &ec[3], &ec[2]); sincos(qv[2], &s2, &c2);
            R_FM = rot_xyz(ec[0], ec[1], ec[2], ec[3], c2, s2);
        } else
        R_FM = quat_rot(qv[0]*oo, qv[1]*oo, qv[2]*oo, qv[3]*oo);
        H_FM[0] = {V3(1.0, 0.0, 0.0), zero3<double>()};
        H_FM[1] = {V3(0.0, 1.0, 0.0), zero3<double>()};
        H_FM[2] = {V3(0.0, 0.0, 1.0), zero3<double>()};
        if constexpr (TYPE == FREE) {
            if (bush) {   // the Gimbal's angle columns (w_FM = H_FM qdot)
                H_FM[1] = {V3(0.0, ec[0], ec[1]), zero3<double>()};
                H_FM[2] = {V3(ec[3], -ec[1]*ec[2], ec[0]*ec[2]), zero3<double>()};
            }
            p_FM = eul ? V3(qv[3], qv[4], qv[5]) : V3(qv[4], qv[5], qv[6]);
            H_FM[3] = {zero3<double>(), V3(1.0, 0.0, 0.0)};
            H_FM[4] = {zero3<double>(), V3(0.0, 1.0, 0.0)};
            H_FM[5] = {zero3<double>(), V3(0.0, 0.0, 1.0)};
        }
    }
    const dm R_PF = prm_m(P + PRM_R_PF), R_MB = prm_m(P + PRM_R_MB);
    const d3 p_PF = prm_3(P + PRM_P_PF), p_MB = prm_3(P + PRM_P_MB);
    // X_PB = X_PF X_FM X_MB, X_GB = X_GP X_PB
    const dm R_FB = mul(R_FM, R_MB);
    const d3 p_FB = p_FM + mul(R_FM, p_MB);
    const dm R_PB = mul(R_PF, R_FB);
    const d3 p_PB = p_PF + mul(R_PF, p_FB);
    const dm R_GB = mul(R_GP, R_PB);
    const d3 p_GB = p_GP + mul(R_GP, p_PB);
    const dm R_GF = mul(R_GP, R_PF);
    const d3 r_MB_F = mul(R_FM, p_MB);
    dv H[D];
#pragma unroll
    for (int j = 0; j < D; ++j)
        H[j] = {mul(R_GF, H_FM[j].w), mul(R_GF, H_FM[j].v - cross(r_MB_F, H_FM[j].w))};
    const d3 phi = mul(R_GP, p_PB);
    const double m = P[PRM_MASS];
    const d3 Mk_p = mul(R_GB, prm_3(P + PRM_COM));
    const ds Gu = {P[PRM_G], P[PRM_G+1], P[PRM_G+2], P[PRM_G+3], P[PRM_G+4], P[PRM_G+5]};
    const ds Mk_G = rotate_sym(R_GB, Gu);
    stm(c, rb + REC_R, R_GB); st3(c, rb + REC_P, p_GB); st3(c, rb + REC_PHI, phi);
    st3(c, rb + REC_MKP, Mk_p); sts(c, rb + REC_MKG, Mk_G);
    const int rH = rb + REC_DIEPS + D;
#pragma unroll
    for (int j = 0; j < D; ++j) stsv(c, rH + 6*j, H[j]);

    if constexpr (VEL) {
        const dv V_GP = p == 0 ? svzero<double>() : ldsv(c, rp + REC_V);
        double uv[D];
#pragma unroll
        for (int k = 0; k < D; ++k) uv[k] = u(ui + k);
        if constexpr (TYPE == BALL || TYPE == FREE) {
            if (bush) {
#pragma unroll
                for (int k = 0; k < 6; ++k) qd(qi + k, uv[k]);
            } else if (eul) {
                const d3 ed = euler_dot(ec[0], ec[1], ec[2], ec[3], V3(uv[0], uv[1], uv[2]));
                qd(qi, ed.x); qd(qi + 1, ed.y); qd(qi + 2, ed.z);
                if constexpr (TYPE == FREE) { qd(qi+3, uv[3]); qd(qi+4, uv[4]); qd(qi+5, uv[5]); qd(qi+6, 0.0); }
                else qd(qi + 3, 0.0);
            } else {
                double d4[4]; quat_dot(qv, V3(uv[0], uv[1], uv[2]), d4);
#pragma unroll
                for (int k = 0; k < 4; ++k) qd(qi + k, d4[k]);
                if constexpr (TYPE == FREE) { qd(qi+4, uv[3]); qd(qi+5, uv[4]); qd(qi+6, uv[5]); }
            }
        } else {
#pragma unroll
            for (int k = 0; k < D; ++k) qd(qi + k, uv[k]);
        }
        dv V_FM = svzero<double>(), V_PB_G = svzero<double>();
#pragma unroll
        for (int j = 0; j < D; ++j) { V_FM = V_FM + uv[j]*H_FM[j]; V_PB_G = V_PB_G + uv[j]*H[j]; }
        const d3 w_GF = V_GP.w, w_FM = V_FM.w;
        const d3 wxr = cross(w_FM, r_MB_F);
        dv VD = svzero<double>();
#pragma unroll
        for (int j = 0; j < D; ++j) {
            d3 hdw = zero3<double>();
            if constexpr (TYPE == UNIVERSAL) { if (j == 1) hdw = cross(w_FM, col(R_FM, 1)); }
            if constexpr (TYPE == FREE) {
                // a Bushing's angle columns turn with the angles before them:
                // d/dt H_FM(j) = (sum_{k<j} qdot_k H_FM(k).w) x H_FM(j).w
                if (bush && (j == 1 || j == 2)) {
                    const d3 wj = j == 1 ? uv[0]*H_FM[0].w : uv[0]*H_FM[0].w + uv[1]*H_FM[1].w;
                    hdw = cross(wj, H_FM[j].w);
                }
            }
            const d3 lin = (zero3<double>() - cross(r_MB_F, hdw)) - cross(wxr, H_FM[j].w);
            const dv HD = {mul(R_GF, hdw) + cross(w_GF, H[j].w), mul(R_GF, lin) + cross(w_GF, H[j].v)};
            VD = VD + uv[j]*HD;
        }
        const dv V_GB = phiT_mul(phi, V_GP) + V_PB_G;
        const d3 w = V_GB.w;
        const dv bg = {m*cross(w, mul(Mk_G, w)), m*cross(w, cross(w, Mk_p))};
        const dv a = {VD.w, VD.v + cross(V_GP.w, V_GB.v - V_GP.v)};
        stsv(c, rb + REC_V, V_GB); stsv(c, rb + REC_A, a); stsv(c, rb + REC_B, bg);
    }
}

// Padded generic body (batches of different trees): any non-Free mobilizer
// is carried as a 3-dof body whose unused H columns, u entries and D rows are
// zero (D padded with ones on the diagonal, so D^-1 is the mobilizer's own
// inverse bordered by an identity block and every padded term adds an exact
// zero).  Only the joint's own X_FM, H_FM and qdot branch on the type, so
// lanes walking different trees stay on one instruction stream.
template <bool VEL, class QG, class UG, class QD>
__device__ __forceinline__ void kin_body_gen(const Lane& c, int b, const int4 ba, const QG& q, const UG& u,
                                             const QD& qd) {
    constexpr int D = 3;
    const int ty = ba.y, p = ba.x, qi = ba.z, ui = ba.w;
    const int d = mob_dof(ty);
    const cdptr P = c.prm(b);
    const int rb = c.rec(b);
    dm R_GP; d3 p_GP;
    int rp = 0;
    if (p == 0) { R_GP = ident<double>(); p_GP = zero3<double>(); }
    else { rp = c.rec(p); R_GP = ldm(c, rp + REC_R); p_GP = ld3(c, rp + REC_P); }
    double qv[4] = {0.0, 0.0, 0.0, 0.0};
    dm R_FM; d3 p_FM = zero3<double>();
    dv H_FM[D] = {svzero<double>(), svzero<double>(), svzero<double>()};
    const bool eul = ty == BALL && P[PRM_EULER] != 0.0;   // x-y-z angles, not a quaternion
    double ec[4] = {1.0, 0.0, 1.0, 0.0};                   // their c0 s0 c1 s1
    if (ty == PIN) {
        qv[0] = q(qi);
        double sn, cs; sincos(qv[0], &sn, &cs);
        R_FM.a[0]=cs; R_FM.a[1]=-sn; R_FM.a[2]=0; R_FM.a[3]=sn; R_FM.a[4]=cs; R_FM.a[5]=0;
        R_FM.a[6]=0;  R_FM.a[7]=0;   R_FM.a[8]=1;
        H_FM[0] = {V3(0.0, 0.0, 1.0), zero3<double>()};
    } else if (ty == SLIDER) {
        qv[0] = q(qi);
        R_FM = ident<double>(); p_FM = V3(qv[0], 0.0, 0.0);
        H_FM[0] = {zero3<double>(), V3(1.0, 0.0, 0.0)};
    } else if (ty == WELD) {
        R_FM = ident<double>();     // RigidBodyNode_Weld.cpp:389-423: no H
    } else if (ty == UNIVERSAL || ty == GIMBAL) {
        // Universal: body-fixed x then y; Gimbal: body-fixed x, y, z
        // (_Gimbal.h:141-176, Rotation.h:342-349).  With the third angle's
        // (sin, cos) = (0, 1) the 1-2-3 product is the Universal's rotation.
        qv[0] = q(qi); qv[1] = q(qi + 1);
        double s0, c0, s1, c1, s2 = 0.0, c2 = 1.0;
        sincos(qv[0], &s0, &c0); sincos(qv[1], &s1, &c1);
        if (ty == GIMBAL) { qv[2] = q(qi + 2); sincos(qv[2], &s2, &c2); }
        const double s0s1 = s0*s1, s2c0 = s2*c0, c0c2 = c0*c2, nc1 = -c1;
        R_FM.a[0]= c1*c2;             R_FM.a[1]= s2*nc1;            R_FM.a[2]= s1;
        R_FM.a[3]= s2c0 + s0s1*c2;    R_FM.a[4]= c0c2 - s0s1*s2;    R_FM.a[5]= s0*nc1;
        R_FM.a[6]= s0*s2 - s1*c0c2;   R_FM.a[7]= s0*c2 + s1*s2c0;   R_FM.a[8]= c0*c1;
        H_FM[0] = {V3(1.0, 0.0, 0.0), zero3<double>()};
        H_FM[1] = {V3(0.0, c0, s0), zero3<double>()};
        if (ty == GIMBAL) H_FM[2] = {V3(s1, s0*nc1, c0*c1), zero3<double>()};
    } else if (ty == TRANSLATION) {
        // RigidBodyNodeSpec_Translation.h: X_FM = (I, q), H_FM = (0, e_k)
        qv[0] = q(qi); qv[1] = q(qi + 1); qv[2] = q(qi + 2);
        R_FM = ident<double>(); p_FM = V3(qv[0], qv[1], qv[2]);
        H_FM[0] = {zero3<double>(), V3(1.0, 0.0, 0.0)};
        H_FM[1] = {zero3<double>(), V3(0.0, 1.0, 0.0)};
        H_FM[2] = {zero3<double>(), V3(0.0, 0.0, 1.0)};
    } else if (ty == PLANAR || ty == CYLINDER || ty == SCREW || ty == BENDSTRETCH) {
        // RigidBodyNodeSpec_Planar.h:115-156, _Cylinder.h:100-139,
        // _Screw.h:109-148: rotation about z by q0, H_FM constant in F
        qv[0] = q(qi);
        double sn, cs; sincos(qv[0], &sn, &cs);
        R_FM.a[0]=cs; R_FM.a[1]=-sn; R_FM.a[2]=0; R_FM.a[3]=sn; R_FM.a[4]=cs; R_FM.a[5]=0;
        R_FM.a[6]=0;  R_FM.a[7]=0;   R_FM.a[8]=1;
        H_FM[0] = {V3(0.0, 0.0, 1.0), zero3<double>()};
        if (ty == PLANAR) {
            qv[1] = q(qi + 1); qv[2] = q(qi + 2);
            p_FM = V3(qv[1], qv[2], 0.0);
            H_FM[1] = {zero3<double>(), V3(1.0, 0.0, 0.0)};
            H_FM[2] = {zero3<double>(), V3(0.0, 1.0, 0.0)};
        } else if (ty == CYLINDER) {
            qv[1] = q(qi + 1);
            p_FM = V3(0.0, 0.0, qv[1]);
            H_FM[1] = {zero3<double>(), V3(0.0, 0.0, 1.0)};
        } else if (ty == SCREW) {
            const double pt = P[PRM_PITCH];
            p_FM = V3(0.0, 0.0, qv[0]*pt);
            H_FM[0] = {V3(0.0, 0.0, 1.0), V3(0.0, 0.0, pt)};
        } else {
            // BendStretch (PolarCoords.h:128-172): slide q1 along the rotated x
            qv[1] = q(qi + 1);
            p_FM = mul(R_FM, V3(qv[1], 0.0, 0.0));
            H_FM[0] = {V3(0.0, 0.0, 1.0), cross(V3(0.0, 0.0, 1.0), p_FM)};
            H_FM[1] = {zero3<double>(), col(R_FM, 0)};
        }
    } else if (eul) {   // BALL with x-y-z angles (RigidBodyNodeSpec_Ball.h:113-158)
        qv[0] = q(qi); qv[1] = q(qi + 1); qv[2] = q(qi + 2);
        double s2, c2;
        sincos(qv[0], &ec[1], &ec[0]); sincos(qv[1], &ec[3], &ec[2]); sincos(qv[2], &s2, &c2);
        R_FM = rot_xyz(ec[0], ec[1], ec[2], ec[3], c2, s2);
        H_FM[0] = {V3(1.0, 0.0, 0.0), zero3<double>()};
        H_FM[1] = {V3(0.0, 1.0, 0.0), zero3<double>()};
        H_FM[2] = {V3(0.0, 0.0, 1.0), zero3<double>()};
    } else {   // BALL
#pragma unroll
        for (int k = 0; k < 4; ++k) qv[k] = q(qi + k);
        const double len = sqrt(qv[0]*qv[0] + qv[1]*qv[1] + qv[2]*qv[2] + qv[3]*qv[3]);
        const double oo = 1.0 / len;
        R_FM = quat_rot(qv[0]*oo, qv[1]*oo, qv[2]*oo, qv[3]*oo);
        H_FM[0] = {V3(1.0, 0.0, 0.0), zero3<double>()};
        H_FM[1] = {V3(0.0, 1.0, 0.0), zero3<double>()};
        H_FM[2] = {V3(0.0, 0.0, 1.0), zero3<double>()};
    }
    // Reversed mobilizer (RigidBodyNodeSpec.h:262-277): the above is X_MF and
    // H_MF; X_FM = ~X_MF and calcReverseMobilizerH_FM (RigidBodyNodeSpec.cpp:
    // 131-166): H_FM_w = -R_FM H_MF_w, H_FM_v = H_FM_w x p_FM - R_FM H_MF_v.
    // Hm keeps the mobilizer's own angular columns for HDot.
    const bool rv = P[PRM_REV] != 0.0;
    d3 Hm[D] = {H_FM[0].w, H_FM[1].w, H_FM[2].w};
    if (rv) {
        const dm R_MF = R_FM;
        p_FM = zero3<double>() - mulT(R_MF, p_FM);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c2 = 0; c2 < 3; ++c2) R_FM.a[3*r + c2] = R_MF.a[3*c2 + r];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const d3 w = zero3<double>() - mulT(R_MF, Hm[j]);
            H_FM[j] = {w, cross(w, p_FM) - mulT(R_MF, H_FM[j].v)};
        }
    }
    const dm R_PF = prm_m(P + PRM_R_PF), R_MB = prm_m(P + PRM_R_MB);
    const d3 p_PF = prm_3(P + PRM_P_PF), p_MB = prm_3(P + PRM_P_MB);
    const dm R_FB = mul(R_FM, R_MB);
    const d3 p_FB = p_FM + mul(R_FM, p_MB);
    const dm R_PB = mul(R_PF, R_FB);
    const d3 p_PB = p_PF + mul(R_PF, p_FB);
    const dm R_GB = mul(R_GP, R_PB);
    const d3 p_GB = p_GP + mul(R_GP, p_PB);
    const dm R_GF = mul(R_GP, R_PF);
    const d3 r_MB_F = mul(R_FM, p_MB);
    dv H[D];
#pragma unroll
    for (int j = 0; j < D; ++j)
        H[j] = {mul(R_GF, H_FM[j].w), mul(R_GF, H_FM[j].v - cross(r_MB_F, H_FM[j].w))};
    const d3 phi = mul(R_GP, p_PB);
    const double m = P[PRM_MASS];
    const d3 Mk_p = mul(R_GB, prm_3(P + PRM_COM));
    const ds Gu = {P[PRM_G], P[PRM_G+1], P[PRM_G+2], P[PRM_G+3], P[PRM_G+4], P[PRM_G+5]};
    const ds Mk_G = rotate_sym(R_GB, Gu);
    stm(c, rb + REC_R, R_GB); st3(c, rb + REC_P, p_GB); st3(c, rb + REC_PHI, phi);
    st3(c, rb + REC_MKP, Mk_p); sts(c, rb + REC_MKG, Mk_G);
    const int rH = rb + REC_DIEPS + D;
#pragma unroll
    for (int j = 0; j < D; ++j) stsv(c, rH + 6*j, H[j]);

    if constexpr (VEL) {
        const dv V_GP = p == 0 ? svzero<double>() : ldsv(c, rp + REC_V);
        double uv[D];
#pragma unroll
        for (int k = 0; k < D; ++k) uv[k] = k < d ? u(ui + k) : 0.0;
        if (ty == BALL && eul) {
            // angle rates N(q) w; the unused fourth slot's rate is zero
            const d3 ed = euler_dot(ec[0], ec[1], ec[2], ec[3], V3(uv[0], uv[1], uv[2]));
            qd(qi, ed.x); qd(qi + 1, ed.y); qd(qi + 2, ed.z); qd(qi + 3, 0.0);
        } else if (ty == BALL) {
            double d4[4]; quat_dot(qv, V3(uv[0], uv[1], uv[2]), d4);
#pragma unroll
            for (int k = 0; k < 4; ++k) qd(qi + k, d4[k]);
        } else {
            for (int k = 0; k < d; ++k) qd(qi + k, uv[k]);
        }
        dv V_FM = svzero<double>(), V_PB_G = svzero<double>();
#pragma unroll
        for (int j = 0; j < D; ++j) { V_FM = V_FM + uv[j]*H_FM[j]; V_PB_G = V_PB_G + uv[j]*H[j]; }
        const d3 w_GF = V_GP.w, w_FM = V_FM.w;
        const d3 wxr = cross(w_FM, r_MB_F);
        dv VD = svzero<double>();
#pragma unroll
        for (int j = 0; j < D; ++j) {
            d3 hdw = zero3<double>();
            // Universal and Gimbal (_Universal.h:173-189, _Gimbal.h:180-202):
            // d/dt H(j) = (sum_{k<j} u_k H(k).w) x H(j).w in the mobilizer's
            // own columns (HDot_MF for a reversed body)
            if (j >= 1 && (ty == UNIVERSAL || ty == GIMBAL))
                hdw = cross(j == 1 ? uv[0]*Hm[0] : uv[0]*Hm[0] + uv[1]*Hm[1], Hm[j]);
            d3 hdv = zero3<double>();
            // BendStretch (PolarCoords.h:174-200, never reversed)
            if (ty == BENDSTRETCH && j < 2)
                hdv = j == 0 ? cross(V3(0.0, 0.0, 1.0), V_FM.v) : cross(w_FM, col(R_FM, 0));
            if (rv) {
                // calcReverseMobilizerHDot_FM (RigidBodyNodeSpec.cpp:168-216)
                const d3 Hw = H_FM[j].w, Hv = H_FM[j].v;
                hdw = cross(w_FM, Hw) - mul(R_FM, hdw);
                hdv = cross(w_FM, Hv) - ((cross(p_FM, hdw) + cross(V_FM.v, Hw)) - cross(w_FM, cross(p_FM, Hw)));
            }
            const d3 lin = (hdv - cross(r_MB_F, hdw)) - cross(wxr, H_FM[j].w);
            const dv HD = {mul(R_GF, hdw) + cross(w_GF, H[j].w), mul(R_GF, lin) + cross(w_GF, H[j].v)};
            VD = VD + uv[j]*HD;
        }
        const dv V_GB = phiT_mul(phi, V_GP) + V_PB_G;
        const d3 w = V_GB.w;
        const dv bg = {m*cross(w, mul(Mk_G, w)), m*cross(w, cross(w, Mk_p))};
        const dv a = {VD.w, VD.v + cross(V_GP.w, V_GB.v - V_GP.v)};
        stsv(c, rb + REC_V, V_GB); stsv(c, rb + REC_A, a); stsv(c, rb + REC_B, bg);
    }
}

template <bool VEL, bool PAD = false, class QG, class UG, class QD>
__device__ void sweep_kin(const Lane& c, const QG& q, const UG& u, const QD& qd) {
    for (int b = 1; b <= c.nb; ++b) {
        const int4 ba = c.ba(b);
        if (PAD || mob_generic_only(ba.y)) {
            if (ba.y != FREE) { kin_body_gen<VEL>(c, b, ba, q, u, qd); continue; }
        }
        switch (ba.y) {
        case PIN:       kin_body<PIN, VEL>(c, b, ba, q, u, qd); break;
        case SLIDER:    kin_body<SLIDER, VEL>(c, b, ba, q, u, qd); break;
        case UNIVERSAL: kin_body<UNIVERSAL, VEL>(c, b, ba, q, u, qd); break;
        case BALL:      kin_body<BALL, VEL>(c, b, ba, q, u, qd); break;
        default:        kin_body<FREE, VEL>(c, b, ba, q, u, qd); break;
        }
    }
}

// Small dense inverse of D (row-major).  dof<=3: cofactor forms (the
// reference's SmallMatrixMixed.h:895-960); dof=6: Gauss-Jordan elimination
// without pivoting (D = H^T P H is symmetric positive definite for positive
// mass; the reference uses LAPACK LU with partial pivoting, agreement is at
// rounding level).  Returns true when D is not invertible as a symmetric
// positive definite matrix: a determinant (dof<=3) that is not positive, or
// (dof=6) an elimination pivot that is not above 1e-13 of its diagonal
// entry (for a positive definite D the ratio is in (0, 1]; below 1e-13, D is
// singular to working precision), NaN included.  The reference throws for
// the LAPACK case when a pivot is exactly zero (SmallMatrixMixed.h:878-892)
// and divides silently for dof<=3; both are reported here, and the inverse
// is formed anyway (inf/NaN or meaningless values).
template <int D> __device__ __forceinline__ bool inv_small(const double* A, double* Ai) {
    bool bad = false;
    if constexpr (D == 1) { bad = !(A[0] > 0.0); Ai[0] = 1.0 / A[0]; }
    else if constexpr (D == 2) {
        const double det = A[0]*A[3] - A[1]*A[2];
        bad = !(det > 0.0);
        const double o = 1.0 / det;
        Ai[0] = o*A[3]; Ai[1] = -(o*A[1]); Ai[2] = -(o*A[2]); Ai[3] = o*A[0];
    } else if constexpr (D == 3) {
        const double d00 = A[4]*A[8]-A[5]*A[7], nd01 = A[5]*A[6]-A[3]*A[8], d02 = A[3]*A[7]-A[4]*A[6];
        const double det = A[0]*d00 + A[1]*nd01 + A[2]*d02;
        bad = !(det > 0.0);
        const double o = 1.0 / det;
        const double nd10 = A[2]*A[7]-A[1]*A[8], d11 = A[0]*A[8]-A[2]*A[6], nd12 = A[1]*A[6]-A[0]*A[7],
                     d20 = A[1]*A[5]-A[2]*A[4], nd21 = A[2]*A[3]-A[0]*A[5], d22 = A[0]*A[4]-A[1]*A[3];
        Ai[0]=o*d00;  Ai[1]=o*nd10; Ai[2]=o*d20;
        Ai[3]=o*nd01; Ai[4]=o*d11;  Ai[5]=o*nd21;
        Ai[6]=o*d02;  Ai[7]=o*nd12; Ai[8]=o*d22;
    } else {
        double M[D*D];
#pragma unroll
        for (int i = 0; i < D*D; ++i) { M[i] = A[i]; Ai[i] = (i % (D+1) == 0) ? 1.0 : 0.0; }
#pragma unroll
        for (int k = 0; k < D; ++k) {
            bad = bad || !(M[k*D+k] > 1e-13 * A[k*D+k]);
            const double ip = 1.0 / M[k*D+k];
#pragma unroll
            for (int j = 0; j < D; ++j) { M[k*D+j] *= ip; Ai[k*D+j] *= ip; }
#pragma unroll
            for (int r = 0; r < D; ++r) {
                if (r == k) continue;
                const double f = M[r*D+k];
#pragma unroll
                for (int j = 0; j < D; ++j) { M[r*D+j] -= f*M[k*D+j]; Ai[r*D+j] -= f*Ai[k*D+j]; }
            }
        }
    }
    return bad;
}

// IN_BF (with IN_MINV): the body forces enter the first pass as applied
// forces, so the pass computes M^-1 J^T F (the multipliers' accelerations).
enum { IN_ABI_OUT = 1, IN_FWD = 2, IN_MINV = 4, IN_SYSF = 8, IN_PRES = 16, IN_BF = 32 };
// Known udot (prescribed motion at the acceleration level, locks): per-body
// flags m (nb+1 per instance, Ground first), the known udot up and the tau
// output (packed u layout), all offset to the instance.
struct Pres { const int* m; const double* up; double* tau; };

// ---------------------------------------------------------------------------
// Tip-to-base: articulated-body inertias (RigidBodyNodeSpec.cpp:249-325) fused
// with the first ABA pass (calcUDotPass1Inward :355-400) or the first pass of
// M^-1 f (multiplyByMInvPass1Inward :483-515).  F(b) -> body force, f(k) ->
// mobility force.
// ---------------------------------------------------------------------------
// TYPE = GEN: the padded generic body (see kin_body_gen); dof from ba.y.
template <int TYPE> struct MobD { static constexpr int D = TYPE == GEN ? 3 : Mob<TYPE>::D; };
template <int TYPE, int MODE, class FB, class FM>
__device__ __forceinline__ bool inward_body(const Lane& c, int b, const int4 ba, const int4 bb,
                                            const FB& fbody, const FM& fmob,
                                            double* outP, double* outPP, const double* grav, Pres pr = {}) {
    // (IN_PRES) a body with known udot: P+ = P, z+ = z and z += P H udot; its
    // G slot keeps P H and its DI*eps slot eps, so the outward pass forms
    // tau = eps - (P H)^T A+ where it would form udot (RigidBodyNodeSpec.cpp:
    // 292-296, 380-399, 434-443)
    bool pres = false;
    if constexpr ((MODE & IN_PRES) != 0) pres = pr.m[b] != 0;
    constexpr int D = MobD<TYPE>::D;
    const int d = TYPE == GEN ? mob_dof(ba.y) : D;
    const int rb = bb.z;
    const cdptr P0 = c.prm(b);
    const double m = P0[PRM_MASS];
    const d3 Mk_p = ld3(c, rb + REC_MKP);
    const ds Mk_G = lds(c, rb + REC_MKG);
    da P = abi_from_rbi(m, Mk_p, Mk_G);
    for (int k = 0; k < bb.y; ++k) {
        const int ch = cptr(c.T.child_idx)[bb.x + k];
        const int rc = c.rec(ch);
        abi_add_shifted(P, lda(c, rc + REC_PPLUS), ld3(c, rc + REC_PHI));
    }
    const int rH = rb + REC_DIEPS + D, rG = rb + REC_DIEPS + 7*D;
    dv H[D], PH[D];
#pragma unroll
    for (int j = 0; j < D; ++j) { H[j] = ldsv(c, rH + 6*j); PH[j] = abi_mul(P, H[j]); }
    double Dm[D*D], DI[D*D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int k = 0; k < D; ++k) Dm[r*D+k] = dot(H[r], PH[k]);
    if constexpr (TYPE == GEN) {
#pragma unroll
        for (int j = 0; j < D; ++j) if (j >= d) Dm[j*D+j] = 1.0;
    }
    const bool singular = inv_small<D>(Dm, DI);
    dv G[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        dv acc = svzero<double>();
#pragma unroll
        for (int j = 0; j < D; ++j) acc = acc + DI[j*D+k]*PH[j];
        G[k] = acc;
    }
    // P+ = P - G PH^T with the mass and inertia blocks symmetrized.
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
    da PP;
    PP.M = P.M - ds{ms[0], ms[4], ms[8], (ms[3]+ms[1])*0.5, (ms[6]+ms[2])*0.5, (ms[7]+ms[5])*0.5};
    PP.J = P.J - ds{in[0], in[4], in[8], (in[3]+in[1])*0.5, (in[6]+in[2])*0.5, (in[7]+in[5])*0.5};
#pragma unroll
    for (int i = 0; i < 9; ++i) PP.F.a[i] = P.F.a[i] - mm[i];
    if constexpr ((MODE & IN_PRES) != 0) {
        if (pres) {
            PP = P;
#pragma unroll
            for (int j = 0; j < D; ++j) G[j] = PH[j];
        }
    }
    sta(c, rb + REC_PPLUS, PP);
#pragma unroll
    for (int j = 0; j < D; ++j) stsv(c, rG + 6*j, G[j]);
    if constexpr (MODE & IN_ABI_OUT) {
        sta_global(outP + 21*b, P);
        if (outPP) sta_global(outPP + 21*b, PP);
    }
    if constexpr (MODE & (IN_FWD | IN_MINV)) {
        dv z;
        if constexpr (MODE & IN_FWD) {
            const dv a = ldsv(c, rb + REC_A), bg = ldsv(c, rb + REC_B);
            dv F = svzero<double>();
            if constexpr (MODE & IN_SYSF) {
                const d3 g = {grav[0], grav[1], grav[2]};
                const d3 f = m*g;
                F = F + dv{cross(Mk_p, f), f};
            }
            F = F + fbody(b);
            z = (abi_mul(P, a) + bg) - F;
            if constexpr ((MODE & IN_PRES) != 0) {
                if (pres) {
                    dv Hu = svzero<double>();
#pragma unroll
                    for (int j = 0; j < D; ++j) if (j < d) Hu = Hu + pr.up[ba.w + j]*H[j];
                    z = z + abi_mul(P, Hu);
                }
            }
        } else {
            z = svzero<double>();
            if constexpr ((MODE & IN_BF) != 0) z = z - fbody(b);
        }
        for (int k = 0; k < bb.y; ++k) {
            const int rc = c.rec(cptr(c.T.child_idx)[bb.x + k]);
            z = z + phi_mul(ld3(c, rc + REC_PHI), ldsv(c, rc + REC_ZPLUS));
        }
        double eps[D];
#pragma unroll
        for (int j = 0; j < D; ++j) eps[j] = (j < d ? fmob(ba.w + j) : 0.0) - dot(H[j], z);
        dv zp = z;
        if (!pres) {
#pragma unroll
            for (int j = 0; j < D; ++j) zp = zp + eps[j]*G[j];
        }
        stsv(c, rb + REC_ZPLUS, zp);
#pragma unroll
        for (int r = 0; r < D; ++r) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < D; ++k) s += DI[r*D+k]*eps[k];
            c.s(rb + REC_DIEPS + r) = pres ? eps[r] : s;
        }
    }
    return singular;
}

// Returns true when some body's D was not invertible (inv_small).
template <int MODE, bool PAD = false, class FB, class FM>
__device__ bool sweep_inward(const Lane& c, const FB& fbody, const FM& fmob, double* outP, double* outPP,
                             const double* grav, Pres pr = {}) {
    bool sing = false;
    for (int b = c.nb; b >= 1; --b) {
        const int4 ba = c.ba(b), bb = c.bb(b);
        if (PAD || mob_generic_only(ba.y)) {
            if (ba.y != FREE) { sing |= inward_body<GEN, MODE>(c, b, ba, bb, fbody, fmob, outP, outPP, grav, pr); continue; }
        }
        switch (ba.y) {
        case PIN: case SLIDER: sing |= inward_body<PIN, MODE>(c, b, ba, bb, fbody, fmob, outP, outPP, grav, pr); break;
        case UNIVERSAL:        sing |= inward_body<UNIVERSAL, MODE>(c, b, ba, bb, fbody, fmob, outP, outPP, grav, pr); break;
        case BALL:             sing |= inward_body<BALL, MODE>(c, b, ba, bb, fbody, fmob, outP, outPP, grav, pr); break;
        default:               sing |= inward_body<FREE, MODE>(c, b, ba, bb, fbody, fmob, outP, outPP, grav, pr); break;
        }
    }
    return sing;
}

// ---------------------------------------------------------------------------
// Base-to-tip second pass: calcUDotPass2Outward (RigidBodyNodeSpec.cpp:408-446,
// WITH_COR) or multiplyByMInvPass2Outward (:521-552), plus calcQDotDot for
// quaternions (Rotation.h:1006-1013).
// ---------------------------------------------------------------------------
template <int TYPE, bool WITH_COR, bool PRES = false, class QG, class UG, class UO, class QO>
__device__ __forceinline__ void outward_body(const Lane& c, int b, const int4 ba, const QG& q, const UG& u,
                                             const UO& uo, const QO& qdd, Pres pr = {}) {
    constexpr int D = MobD<TYPE>::D;
    const int d = TYPE == GEN ? mob_dof(ba.y) : D;
    const int rb = c.rec(b), p = ba.x;
    const dv A_GP = p == 0 ? svzero<double>() : ldsv(c, c.rec(p) + REC_AGB);
    const dv Ap = phiT_mul(ld3(c, rb + REC_PHI), A_GP);
    const int rH = rb + REC_DIEPS + D, rG = rb + REC_DIEPS + 7*D;
    double ud[D];
    dv A = Ap;
#pragma unroll
    for (int r = 0; r < D; ++r) ud[r] = c.s(rb + REC_DIEPS + r) - dot(ldsv(c, rG + 6*r), Ap);
    if constexpr (PRES) {
        // a body with known udot: what was formed is tau; udot is the known one
        const bool pres = pr.m[b] != 0;
#pragma unroll
        for (int r = 0; r < D; ++r) {
            if (r >= d) { ud[r] = 0.0; continue; }
            if (pr.tau) pr.tau[ba.w + r] = pres ? ud[r] : 0.0;
            if (pres) ud[r] = pr.up[ba.w + r];
        }
    }
    dv Hu = svzero<double>();
#pragma unroll
    for (int j = 0; j < D; ++j) Hu = Hu + ud[j]*ldsv(c, rH + 6*j);
    A = A + Hu;
    if constexpr (WITH_COR) A = A + ldsv(c, rb + REC_A);
    stsv(c, rb + REC_AGB, A);
#pragma unroll
    for (int j = 0; j < D; ++j) if (j < d) uo(ba.w + j, ud[j]);
    if constexpr (WITH_COR) {
        bool quat = TYPE == BALL || TYPE == FREE;
        if constexpr (TYPE == GEN) quat = ba.y == BALL;
        if constexpr (D >= 3) {
            if (TYPE == FREE && c.prm(b)[PRM_EULER] == 2.0) {      // Bushing: qdotdot = udot
#pragma unroll
                for (int k = 0; k < 6; ++k) qdd(ba.z + k, ud[k]);
                return;
            }
            if (quat && c.prm(b)[PRM_EULER] != 0.0) {
                // x-y-z angles: qdotdot = N b + NDot w (RigidBodyNodeSpec_Ball.h:355-392)
                double s0, c0, s1, c1;
                sincos(q(ba.z), &s0, &c0); sincos(q(ba.z + 1), &s1, &c1);
                const d3 w = {u(ba.w), u(ba.w+1), u(ba.w+2)};
                const d3 qd = euler_dot(c0, s0, c1, s1, w);
                const d3 ea = euler_dotdot(c0, s0, c1, s1, qd, V3(ud[0], ud[1], ud[2]));
                qdd(ba.z, ea.x); qdd(ba.z + 1, ea.y); qdd(ba.z + 2, ea.z);
                if constexpr (TYPE == FREE) { qdd(ba.z+3, ud[3]); qdd(ba.z+4, ud[4]); qdd(ba.z+5, ud[5]); qdd(ba.z+6, 0.0); }
                else qdd(ba.z + 3, 0.0);
                return;
            }
            if (quat) {
                double qq[4], nb4[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) qq[k] = q(ba.z + k);
                const d3 w = {u(ba.w), u(ba.w+1), u(ba.w+2)};
                quat_dot(qq, V3(ud[0], ud[1], ud[2]), nb4);
                const double sc = -0.25 * dot(w, w);
#pragma unroll
                for (int k = 0; k < 4; ++k) qdd(ba.z + k, nb4[k] + sc*qq[k]);
                if constexpr (TYPE == FREE) { qdd(ba.z+4, ud[3]); qdd(ba.z+5, ud[4]); qdd(ba.z+6, ud[5]); }
                return;
            }
        }
#pragma unroll
        for (int j = 0; j < D; ++j) if (j < d) qdd(ba.z + j, ud[j]);
    }
}

template <bool WITH_COR, bool PAD = false, bool PRES = false, class QG, class UG, class UO, class QO>
__device__ void sweep_outward2(const Lane& c, const QG& q, const UG& u, const UO& uo, const QO& qdd, Pres pr = {}) {
    for (int b = 1; b <= c.nb; ++b) {
        const int4 ba = c.ba(b);
        if (PAD || mob_generic_only(ba.y)) {
            if (ba.y != FREE) { outward_body<GEN, WITH_COR, PRES>(c, b, ba, q, u, uo, qdd, pr); continue; }
        }
        switch (ba.y) {
        case PIN: case SLIDER: outward_body<PIN, WITH_COR, PRES>(c, b, ba, q, u, uo, qdd, pr); break;
        case UNIVERSAL:        outward_body<UNIVERSAL, WITH_COR, PRES>(c, b, ba, q, u, uo, qdd, pr); break;
        case BALL:             outward_body<BALL, WITH_COR, PRES>(c, b, ba, q, u, uo, qdd, pr); break;
        default:               outward_body<FREE, WITH_COR, PRES>(c, b, ba, q, u, uo, qdd, pr); break;
        }
    }
}

// ---------------------------------------------------------------------------
// Inverse dynamics: calcBodyAccelerationsFromUdotOutward (RigidBodyNodeSpec.cpp:
// 566-582) and calcInverseDynamicsPass2Inward (:609-641).
// ---------------------------------------------------------------------------
template <int D, class UD>
__device__ __forceinline__ void resid_out_body(const Lane& c, int b, const int4 ba, const UD& udot) {
    const int rb = c.rec(b), p = ba.x;
    const dv A_GP = p == 0 ? svzero<double>() : ldsv(c, c.rec(p) + REC_AGB);
    dv A = phiT_mul(ld3(c, rb + REC_PHI), A_GP);
    const int rH = rb + REC_DIEPS + D;
    const int d = mob_dof(ba.y);     // < D only for a padded generic body
    dv Hu = svzero<double>();
#pragma unroll
    for (int j = 0; j < D; ++j) if (j < d) Hu = Hu + udot(ba.w + j)*ldsv(c, rH + 6*j);
    A = (A + Hu) + ldsv(c, rb + REC_A);
    stsv(c, rb + REC_AGB, A);
}
template <int D, class FB, class FM, class TO>
__device__ __forceinline__ void resid_in_body(const Lane& c, int b, const int4 ba, const int4 bb,
                                              const FB& fbody, const FM& fmob, const TO& tau) {
    const int rb = bb.z;
    const double m = c.prm(b)[PRM_MASS];
    dv F = rbi_mul(m, ld3(c, rb + REC_MKP), lds(c, rb + REC_MKG), ldsv(c, rb + REC_AGB)) + ldsv(c, rb + REC_B);
    F = F - fbody(b);
    for (int k = 0; k < bb.y; ++k) {
        const int rc = c.rec(cptr(c.T.child_idx)[bb.x + k]);
        F = F + phi_mul(ld3(c, rc + REC_PHI), ldsv(c, rc + REC_ZPLUS));
    }
    stsv(c, rb + REC_ZPLUS, F);
    const int rH = rb + REC_DIEPS + D;
    const int d = mob_dof(ba.y);
#pragma unroll
    for (int j = 0; j < D; ++j) if (j < d) tau(ba.w + j, dot(ldsv(c, rH + 6*j), F) - fmob(ba.w + j));
}

// ---------------------------------------------------------------------------
// O(n) operators on the position (and velocity) records.
// ---------------------------------------------------------------------------
// Outward: A_GB = Phi^T A_GP + H v -- the system Jacobian product J v
// (RigidBodyNodeSpec.cpp:760-785) and pass 1 of M v (:657-669).
template <int D, class VG>
__device__ __forceinline__ void jac_out_body(const Lane& c, int b, const int4 ba, const VG& v) {
    const int rb = c.rec(b), p = ba.x;
    const dv A_GP = p == 0 ? svzero<double>() : ldsv(c, c.rec(p) + REC_AGB);
    const dv A = phiT_mul(ld3(c, rb + REC_PHI), A_GP);
    const int rH = rb + REC_DIEPS + D;
    const int d = mob_dof(ba.y);
    dv Hv = svzero<double>();
#pragma unroll
    for (int j = 0; j < D; ++j) if (j < d) Hv = Hv + v(ba.w + j)*ldsv(c, rH + 6*j);
    stsv(c, rb + REC_AGB, A + Hv);
}
// Inward: z = F_own + sum_children Phi z_child, tau = H^T z.  MK: F_own =
// Mk_G A_GB (pass 2 of M v, RigidBodyNodeSpec.cpp:672-695); else the given
// body force (J^T F, :795-815).
template <int D, bool MK, class FB, class TO>
__device__ __forceinline__ void jact_in_body(const Lane& c, int b, const int4 ba, const int4 bb,
                                             const FB& fbody, const TO& tau) {
    const int rb = bb.z;
    dv z;
    if constexpr (MK) z = rbi_mul(c.prm(b)[PRM_MASS], ld3(c, rb + REC_MKP), lds(c, rb + REC_MKG), ldsv(c, rb + REC_AGB));
    else z = fbody(b);
    for (int k = 0; k < bb.y; ++k) {
        const int rc = c.rec(cptr(c.T.child_idx)[bb.x + k]);
        z = z + phi_mul(ld3(c, rc + REC_PHI), ldsv(c, rc + REC_ZPLUS));
    }
    stsv(c, rb + REC_ZPLUS, z);
    const int rH = rb + REC_DIEPS + D;
    const int d = mob_dof(ba.y);
#pragma unroll
    for (int j = 0; j < D; ++j) if (j < d) tau(ba.w + j, dot(ldsv(c, rH + 6*j), z));
}
// Unit inertia of a unit point mass at p about the origin, |p|^2 I - p p^T
// (UnitInertia::pointMassAt, MassProperties.h:866-867).
__device__ __forceinline__ ds point_mass(const d3& p) {
    const double xx = p.x*p.x, yy = p.y*p.y, zz = p.z*p.z;
    return {yy + zz, xx + zz, xx + yy, -(p.x*p.y), -(p.x*p.z), -(p.y*p.z)};
}
// Composite body inertia (RigidBodyNode.cpp:232-243): R = Mk + sum over the
// children of R_child shifted to this origin (SpatialInertia::shift,
// MassProperties.h:1103-1116) and added mass-weighted (operator+=, :1037-1045).
// Kept in the P+ slots of the record: m, mass center (3), unit inertia (6).
__device__ __forceinline__ void cbi_in_body(const Lane& c, int b, const int4 bb) {
    const int rb = bb.z;
    double m = c.prm(b)[PRM_MASS];
    d3 p = ld3(c, rb + REC_MKP);
    ds G = lds(c, rb + REC_MKG);
    for (int k = 0; k < bb.y; ++k) {
        const int rc = c.rec(cptr(c.T.child_idx)[bb.x + k]);
        const double mc = c.s(rc + REC_PPLUS);
        const d3 pc = ld3(c, rc + REC_PPLUS + 1);
        const ds Gc = lds(c, rc + REC_PPLUS + 4);
        const d3 pn = pc + ld3(c, rc + REC_PHI);          // shift by -phi: new origin is this body's
        const ds Gs = (Gc - point_mass(pc)) + point_mass(pn);
        const double mt = m + mc, oo = 1.0 / mt;
        p = oo*(m*p + mc*pn);
        G = {oo*(m*G.xx + mc*Gs.xx), oo*(m*G.yy + mc*Gs.yy), oo*(m*G.zz + mc*Gs.zz),
             oo*(m*G.xy + mc*Gs.xy), oo*(m*G.xz + mc*Gs.xz), oo*(m*G.yz + mc*Gs.yz)};
        m = mt;
    }
    c.s(rb + REC_PPLUS) = m;
    st3(c, rb + REC_PPLUS + 1, p);
    sts(c, rb + REC_PPLUS + 4, G);
}

// ---------------------------------------------------------------------------
// Accessors.
// ---------------------------------------------------------------------------
struct PackedGet {   // packed boundary layout, already offset to the instance
    const double* p;
    __device__ double operator()(int k) const { return p[k]; }
};
struct PackedGetOpt {
    const double* p;
    __device__ double operator()(int k) const { return p ? p[k] : 0.0; }
};
struct PackedPut {
    double* p;
    __device__ void operator()(int k, double v) const { if (p) p[k] = v; }
};
struct SoAGet {
    const double* p; int L; unsigned lane;
    __device__ double operator()(int k) const { return (p + (size_t)k * L)[lane]; }
};
struct SoAPut {
    double* p; int L; unsigned lane;
    __device__ void operator()(int k, double v) const { (p + (size_t)k * L)[lane] = v; }
};
struct NoPut { __device__ void operator()(int, double) const {} };
struct Zero { __device__ double operator()(int) const { return 0.0; } };
struct BodyForce {   // packed (nb+1)*6 per instance; nullptr = zero
    const double* p;
    __device__ dv operator()(int b) const {
        if (!p) return svzero<double>();
        const double* f = p + 6*b;
        return {V3(f[0], f[1], f[2]), V3(f[3], f[4], f[5])};
    }
};

// ---------------------------------------------------------------------------
// Holonomic constraints, Ground as the ancestor frame (the frame of the
// multipliers): Constraint::Ball (ConstraintImpl.h:1835-2000: station P on
// body 1, S on body 2; perr = p_S - p_P; velocity and acceleration rows at
// S and at C, the material point of body 1 at S; +lambda at S on body 2,
// -lambda at C on body 1) and Constraint::Rod (Constraint_RodImpl.h:193-300:
// perr = |p_Sb - p_Sf| - d; rows along Cz = (p_Sb - p_Sf)/r plus the Czd
// term; lambda Cz at Sb, -lambda Cz at Sf).  Everything is read from the
// records of the current kinematics; the workspace (CWS_*) holds W, its
// factor and the vectors.  Per-lane, on the tree sweeps.
// ---------------------------------------------------------------------------
// D of a record: padded generic bodies are 3-dof in batches of different trees.
#define TD_RDOF(ba) (((PAD || mob_generic_only((ba).y)) && (ba).y != FREE) ? 3 : mob_dof((ba).y))
struct Cons {
    const int4* ca; const double* cp; int n, m;   // the lane's tree's constraints and rows
    double* ws; int L;
    __device__ __forceinline__ double& w(int k) const { return ws[(size_t)k * L]; }
};
__device__ __forceinline__ Cons make_cons(const DevTables& T, int tr, double* ws, int lane, int L) {
    Cons k{nullptr, nullptr, 0, 0, nullptr, L};
    if (!T.cons_t || !ws) return k;
    const int4 t = ldi4(T.cons_t, tr);
    k.ca = T.cons_a + t.x; k.cp = T.cons_p + 7*(size_t)t.x; k.n = t.y; k.m = t.z; k.ws = ws + lane;
    return k;
}
struct CGeom { d3 p1, p2, r1, r2, Cz; double r; };
__device__ __forceinline__ CGeom cons_geom(const Lane& c, const int4 ca, const double* cp) {
    dm R1 = ident<double>(), R2 = ident<double>();
    d3 o1 = zero3<double>(), o2 = zero3<double>();
    if (ca.y > 0) { const int rb = c.rec(ca.y); R1 = ldm(c, rb + REC_R); o1 = ld3(c, rb + REC_P); }
    if (ca.z > 0) { const int rb = c.rec(ca.z); R2 = ldm(c, rb + REC_R); o2 = ld3(c, rb + REC_P); }
    const d3 l1 = mul(R1, V3(cp[0], cp[1], cp[2])), l2 = mul(R2, V3(cp[3], cp[4], cp[5]));
    CGeom g;
    g.p1 = o1 + l1; g.p2 = o2 + l2;
    g.r2 = l2;
    g.r1 = ca.x == CONS_BALL ? g.p2 - o1 : l1;
    g.r = 0.0; g.Cz = zero3<double>();
    if (ca.x == CONS_ROD) {
        const d3 d = g.p2 - g.p1;
        g.r = sqrt(dot(d, d));
        g.Cz = (1.0/g.r) * d;
    }
    return g;
}
__device__ __forceinline__ dv rec_sv(const Lane& c, int b, int field) {
    return b > 0 ? ldsv(c, c.rec(b) + field) : svzero<double>();
}
__device__ __forceinline__ d3 st_vel(const dv& V, const d3& r) { return V.v + cross(V.w, r); }
__device__ __forceinline__ d3 st_acc(const dv& A, const dv& V, const d3& r) {
    return (A.v + cross(A.w, r)) + cross(V.w, cross(V.w, r));
}
// Position errors into ws[out..].
__device__ void cons_perr(const Lane& c, const Cons& K, int out) {
    int row = 0;
    for (int i = 0; i < K.n; ++i) {
        const int4 ca = K.ca[i];
        const CGeom g = cons_geom(c, ca, K.cp + 7*i);
        if (ca.x == CONS_BALL) {
            const d3 d = g.p2 - g.p1;
            K.w(out + row) = d.x; K.w(out + row + 1) = d.y; K.w(out + row + 2) = d.z; row += 3;
        } else {
            K.w(out + row) = g.r - K.cp[7*i + 6]; row += 1;
        }
    }
}
// The constraint rows of the per-body spatial vectors in record field
// `field` into ws[out + row*stride]: velocity errors (REC_V), P v (the J v
// of the Jacobian sweeps), or with accel the acceleration errors of the
// body accelerations in REC_AGB (their velocity terms and the rod's Czd
// term from REC_V).
__device__ void cons_rows(const Lane& c, const Cons& K, int field, bool accel, int out, int stride) {
    int row = 0;
    for (int i = 0; i < K.n; ++i) {
        const int4 ca = K.ca[i];
        const CGeom g = cons_geom(c, ca, K.cp + 7*i);
        const dv X1 = rec_sv(c, ca.y, field), X2 = rec_sv(c, ca.z, field);
        d3 d, pd = zero3<double>();
        if (accel) {
            const dv V1 = rec_sv(c, ca.y, REC_V), V2 = rec_sv(c, ca.z, REC_V);
            d = st_acc(X2, V2, g.r2) - st_acc(X1, V1, g.r1);
            pd = st_vel(V2, g.r2) - st_vel(V1, g.r1);
        } else {
            d = st_vel(X2, g.r2) - st_vel(X1, g.r1);
        }
        if (ca.x == CONS_BALL) {
            K.w(out + row*stride) = d.x; K.w(out + (row + 1)*stride) = d.y; K.w(out + (row + 2)*stride) = d.z;
            row += 3;
        } else {
            double v = dot(d, g.Cz);
            if (accel) {
                const d3 Czd = (1.0/g.r) * (pd - dot(pd, g.Cz) * g.Cz);
                v = v + dot(pd, Czd);
            }
            K.w(out + row*stride) = v; row += 1;
        }
    }
}
// Body forces of the multipliers ws[lam..] with sign sgn, plus the applied
// forces fapp (nullable), for the inward sweeps.
struct ConsForce {
    const Lane* c; Cons K; int lam; double sgn; const double* fapp;
    __device__ dv operator()(int b) const {
        dv F = BodyForce{fapp}(b);
        int row = 0;
        for (int i = 0; i < K.n; ++i) {
            const int4 ca = K.ca[i];
            const int nr = ca.x == CONS_BALL ? 3 : 1;
            if (ca.y == b || ca.z == b) {
                const CGeom g = cons_geom(*c, ca, K.cp + 7*i);
                d3 f = ca.x == CONS_BALL ? V3(K.w(lam + row), K.w(lam + row + 1), K.w(lam + row + 2))
                                         : K.w(lam + row) * g.Cz;
                f = sgn * f;
                if (ca.z == b) F = F + dv{cross(g.r2, f), f};
                if (ca.y == b) F = F - dv{cross(g.r1, f), f};
            }
            row += nr;
        }
        return F;
    }
};
// Solves A y = r for the symmetric positive definite m x m A at ws[a]
// (row-major) by Cholesky in ws[CWS_F] (A is kept); r at ws[r], y at ws[y].
// False when a pivot is not positive.
__device__ bool cons_solve(const Cons& K, int a, int r, int y) {
    const int m = K.m;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j <= i; ++j) {
            double sacc = K.w(a + i*m + j);
            for (int k = 0; k < j; ++k) sacc -= K.w(CWS_F + i*m + k) * K.w(CWS_F + j*m + k);
            if (i == j) {
                if (!(sacc > 0.0)) return false;
                K.w(CWS_F + i*m + i) = sqrt(sacc);
            } else {
                K.w(CWS_F + i*m + j) = sacc / K.w(CWS_F + j*m + j);
            }
        }
    for (int i = 0; i < m; ++i) {
        double sacc = K.w(r + i);
        for (int k = 0; k < i; ++k) sacc -= K.w(CWS_F + i*m + k) * K.w(y + k);
        K.w(y + i) = sacc / K.w(CWS_F + i*m + i);
    }
    for (int i = m - 1; i >= 0; --i) {
        double sacc = K.w(y + i);
        for (int k = i + 1; k < m; ++k) sacc -= K.w(CWS_F + k*m + i) * K.w(y + k);
        K.w(y + i) = sacc / K.w(CWS_F + i*m + i);
    }
    return true;
}
__device__ __forceinline__ void cons_unit(const Cons& K, int off, int k) {
    for (int i = 0; i < K.m; ++i) K.w(off + i) = i == k ? 1.0 : 0.0;
}
// J^T F (tau, into the u functor) and J v (into REC_AGB) on the records.
template <bool PAD, class FB, class TO>
__device__ void cons_sweep_jt(const Lane& c, const FB& fb, const TO& tau) {
    for (int b = c.nb; b >= 1; --b) {
        const int4 ba = c.ba(b), bb = c.bb(b);
        switch (TD_RDOF(ba)) {
        case 1: jact_in_body<1, false>(c, b, ba, bb, fb, tau); break;
        case 2: jact_in_body<2, false>(c, b, ba, bb, fb, tau); break;
        case 3: jact_in_body<3, false>(c, b, ba, bb, fb, tau); break;
        default: jact_in_body<6, false>(c, b, ba, bb, fb, tau); break;
        }
    }
}
template <bool PAD, class VG>
__device__ void cons_sweep_j(const Lane& c, const VG& v) {
    for (int b = 1; b <= c.nb; ++b) {
        const int4 ba = c.ba(b);
        switch (TD_RDOF(ba)) {
        case 1: jac_out_body<1>(c, b, ba, v); break;
        case 2: jac_out_body<2>(c, b, ba, v); break;
        case 3: jac_out_body<3>(c, b, ba, v); break;
        default: jac_out_body<6>(c, b, ba, v); break;
        }
    }
}
// Loop forward dynamics (calcLoopForwardDynamicsOperator, SimbodyMatter-
// SubsystemRep.cpp:5076-5152): the tree's udot under gravity and fapp, the
// acceleration errors it leaves, W = G M^-1 G^T one column per multiplier
// (M^-1 J^T of the unit multiplier's body forces; the rows of the resulting
// body accelerations, which carry no velocity terms), W lambda = aerr
// (multipliers at ws[CWS_LAM]), and udot again under fapp minus the
// constraint forces.  qd gets qdot, ud udot.  False when W is not positive
// definite.
template <bool PAD, class QG, class UG, class QD, class UO>
__device__ bool cons_derivs(const Lane& c, const Cons& K, const double* fapp, const double* grav, const QG& q,
                            const UG& u, const QD& qd, const UO& ud) {
    sweep_kin<true, PAD>(c, q, u, qd);
    sweep_inward<IN_FWD | IN_SYSF, PAD>(c, BodyForce{fapp}, Zero{}, nullptr, nullptr, grav);
    sweep_outward2<true, PAD>(c, q, u, NoPut{}, NoPut{});
    cons_rows(c, K, REC_AGB, true, CWS_R, 1);
    for (int k = 0; k < K.m; ++k) {
        cons_unit(K, CWS_Y, k);
        sweep_inward<IN_MINV | IN_BF, PAD>(c, ConsForce{&c, K, CWS_Y, 1.0, nullptr}, Zero{}, nullptr, nullptr, grav);
        sweep_outward2<false, PAD>(c, q, u, NoPut{}, NoPut{});
        cons_rows(c, K, REC_AGB, false, CWS_W + k, K.m);
    }
    if (!cons_solve(K, CWS_W, CWS_R, CWS_LAM)) return false;
    sweep_inward<IN_FWD | IN_SYSF, PAD>(c, ConsForce{&c, K, CWS_LAM, -1.0, fapp}, Zero{}, nullptr, nullptr, grav);
    sweep_outward2<true, PAD>(c, q, u, ud, NoPut{});
    return true;
}

// ---- projections (unit weights: the reference's default constraint-error
// and u weights, StateImpl.h:508-566) --------------------------------------
// Per-body scale of J^T F: NInv NInv^T, 4|q|^2 on a quaternion's rotation
// rates (Rotation.h:742-748), 1 elsewhere; the batch has no Euler-angle
// models here.
__device__ __forceinline__ bool body_quat(const Lane& c, int4 ba, int b) {
    return (ba.y == BALL || ba.y == FREE) && c.prm(b)[PRM_EULER] == 0.0;
}
template <class QG, class UR>
__device__ void cons_scale_d(const Lane& c, const QG& q, const UR& t) {
    for (int b = 1; b <= c.nb; ++b) {
        const int4 ba = c.ba(b);
        if (!body_quat(c, ba, b)) continue;
        double n2 = 0;
        for (int k = 0; k < 4; ++k) { const double x = q(ba.z + k); n2 += x*x; }
        for (int k = 0; k < 3; ++k) t(ba.w + k) = (4.0*n2) * t(ba.w + k);
    }
}
// dq = N(q) du: the unnormalized quaternion form (Rotation.h:712-719) on Ball
// and Free rotations, identity elsewhere (qdot = u).
template <class QG, class UR, class QR>
__device__ void cons_mul_n(const Lane& c, const QG& q, const UR& du, const QR& dq) {
    for (int b = 1; b <= c.nb; ++b) {
        const int4 ba = c.ba(b);
        if (body_quat(c, ba, b)) {
            double qq[4], d4[4];
            for (int k = 0; k < 4; ++k) qq[k] = q(ba.z + k);
            quat_dot(qq, V3(du(ba.w), du(ba.w + 1), du(ba.w + 2)), d4);
            for (int k = 0; k < 4; ++k) dq(ba.z + k) = d4[k];
            if (ba.y == FREE) for (int k = 0; k < 3; ++k) dq(ba.z + 4 + k) = du(ba.w + 3 + k);
        } else {
            const int d = mob_dof(ba.y);
            for (int k = 0; k < d; ++k) dq(ba.z + k) = du(ba.w + k);
        }
    }
}
// Norm of the constraint errors at off: RMS, or the largest magnitude (inf).
__device__ __forceinline__ double cons_norm(const Cons& K, int off, bool inf) {
    double sacc = 0;
    for (int i = 0; i < K.m; ++i) norm_acc(sacc, K.w(off + i), inf);
    return norm_end(sacc, K.m, inf);
}
// Norm of |quaternion| - 1 over the quaternions of q.
template <class QG>
__device__ double cons_quat_norm(const Lane& c, const QG& q, int nquat, bool inf) {
    if (nquat == 0) return 0.0;
    double s2 = 0;
    for (int b = 1; b <= c.nb; ++b) {
        const int4 ba = c.ba(b);
        if (!body_quat(c, ba, b)) continue;
        double n2 = 0;
        for (int k = 0; k < 4; ++k) { const double x = q(ba.z + k); n2 += x*x; }
        norm_acc(s2, sqrt(n2) - 1.0, inf);
    }
    return norm_end(s2, nquat, inf);
}
template <class QR, class ER>
__device__ void cons_normalize(const Lane& c, const QR& q, const ER* e) {
    for (int b = 1; b <= c.nb; ++b) {
        const int4 ba = c.ba(b);
        if (!body_quat(c, ba, b)) continue;
        double qq[4], n2 = 0;
        for (int k = 0; k < 4; ++k) { qq[k] = q(ba.z + k); n2 += qq[k]*qq[k]; }
        const double n = sqrt(n2);
        for (int k = 0; k < 4; ++k) { qq[k] = qq[k] / n; q(ba.z + k) = qq[k]; }
        if (e) {
            double d = 0;
            for (int k = 0; k < 4; ++k) d += (*e)(ba.z + k) * qq[k];
            for (int k = 0; k < 4; ++k) (*e)(ba.z + k) = (*e)(ba.z + k) - d*qq[k];
        }
    }
}
// The least-norm Newton update dq = N NInv Pq^T (Pq Pq^T)^-1 r (Pq = P NInv,
// so Pq Pq^T = P D P^T and NInv Pq^T = D P^T) for the residual ws[r] at the
// current kinematics of q; du (u rows) and dq (q rows) are scratch rows.
template <bool PAD, class QG, class UR, class QR>
__device__ bool cons_q_update(const Lane& c, const Cons& K, const QG& q, int r, const UR& du, const QR& dq) {
    for (int k = 0; k < K.m; ++k) {
        cons_unit(K, CWS_Y, k);
        cons_sweep_jt<PAD>(c, ConsForce{&c, K, CWS_Y, 1.0, nullptr}, [&](int i, double v) { du(i) = v; });
        cons_scale_d(c, q, du);
        cons_sweep_j<PAD>(c, [&](int i) { return du(i); });
        cons_rows(c, K, REC_AGB, false, CWS_W + k, K.m);
    }
    if (!cons_solve(K, CWS_W, r, CWS_LAM)) return false;
    cons_sweep_jt<PAD>(c, ConsForce{&c, K, CWS_LAM, 1.0, nullptr}, [&](int i, double v) { du(i) = v; });
    cons_scale_d(c, q, du);
    cons_mul_n(c, q, du, dq);
    return true;
}
// projectQ (SimbodyMatterSubsystemRep.cpp:4096-4440) with the integrator's
// options (opt, PJ_* bits): LocalOnly unless PJ_INIT; ForceProjection and
// ForceFullNewton with PJ_INIT; ForceProjection with PJ_FORCE; every norm
// the infinity norm with PJ_INF.  q, u: the state rows (u only feeds the kinematics); e (nullable):
// the q error estimate; sv (q rows), du (u rows), dq (q rows): scratch.  The
// error-estimate projection uses Pq at the final q (the reference keeps the
// factorization of the last iteration's starting q).
template <bool PAD, class QR, class UR, class ER, class SR, class DUR, class DQR>
__device__ bool cons_project_q(const Lane& c, const Cons& K, const QR& q, const UR& u, const ER* e, const SR& sv,
                               const DUR& du, const DQR& dq, int nq, int nquat, double consTol, double projLimit,
                               int opt, bool* changed) {
    const bool init = opt & PJ_INIT, force = opt & (PJ_INIT | PJ_FORCE), inf = opt & PJ_INF;
    *changed = false;
    sweep_kin<false, PAD>(c, q, u, NoPut{});
    cons_perr(c, K, CWS_R);
    const double pn0 = cons_norm(K, CWS_R, inf), qn0 = cons_quat_norm(c, q, nquat, inf);
    if (fmax(pn0, qn0) > projLimit || pn0 != pn0) return false;
    if (pn0 == 0 || (pn0 <= consTol && !force)) {
        if (qn0 > consTol || force) {
            if (nquat) { cons_normalize(c, q, e); *changed = true; }
            if (cons_quat_norm(c, q, nquat, inf) > consTol) return false;
        }
        return true;
    }
    const double target = fmax(0.1*consTol, 1.9323057602939027e-14);   // SignificantReal = eps^(7/8)
    for (int j = 0; j < nq; ++j) sv(j) = q(j);
    double pn = pn0, prev = pn0;
    int its = 0;
    do {
        if (!cons_q_update<PAD>(c, K, q, CWS_R, du, dq)) return false;
        for (int j = 0; j < nq; ++j) q(j) = q(j) - dq(j);
        *changed = true; ++its;
        sweep_kin<false, PAD>(c, q, u, NoPut{});
        cons_perr(c, K, CWS_R);
        pn = cons_norm(K, CWS_R, inf);
        if (!init && its >= 2 && pn > prev) {
            for (int j = 0; j < nq; ++j) q(j) = q(j) + dq(j);
            sweep_kin<false, PAD>(c, q, u, NoPut{});
            cons_perr(c, K, CWS_R);
            pn = cons_norm(K, CWS_R, inf);
            break;
        }
        prev = pn;
    } while (pn > target && its < 20);
    if (!(pn <= consTol)) {
        if (!(pn < pn0)) {
            for (int j = 0; j < nq; ++j) q(j) = sv(j);
            sweep_kin<false, PAD>(c, q, u, NoPut{});
        }
        return false;
    }
    if (e) {
        // remove the error estimate's component off the constraint manifold:
        // r = Pq e = P NInv e, then e -= dq
        for (int b = 1; b <= c.nb; ++b) {
            const int4 ba = c.ba(b);
            if (body_quat(c, ba, b)) {
                double qq[4], ee[4];
                for (int k = 0; k < 4; ++k) { qq[k] = q(ba.z + k); ee[k] = (*e)(ba.z + k); }
                const d3 w = quat_dot_to_w(qq, ee);
                du(ba.w) = w.x; du(ba.w + 1) = w.y; du(ba.w + 2) = w.z;
                if (ba.y == FREE) for (int k = 0; k < 3; ++k) du(ba.w + 3 + k) = (*e)(ba.z + 4 + k);
            } else {
                const int d = mob_dof(ba.y);
                for (int k = 0; k < d; ++k) du(ba.w + k) = (*e)(ba.z + k);
            }
        }
        cons_sweep_j<PAD>(c, [&](int i) { return du(i); });
        cons_rows(c, K, REC_AGB, false, CWS_R, 1);
        if (!cons_q_update<PAD>(c, K, q, CWS_R, du, dq)) return false;
        for (int j = 0; j < nq; ++j) (*e)(j) = (*e)(j) - dq(j);
    }
    if (nquat) {
        cons_normalize(c, q, e);
        if (cons_quat_norm(c, q, nquat, inf) > consTol) return false;
    }
    return true;
}
// projectU (SimbodyMatterSubsystemRep.cpp:4666-4934) with the integrator's
// options (as projectQ): relative u scaling s = max(|u_i|, 1) from the u on
// entry, one factorization of P S^2 P^T for every iteration.  e (nullable):
// the u error estimate; sv, du (u rows): scratch.
template <bool PAD, class QR, class UR, class ER, class SR, class DUR>
__device__ bool cons_project_u(const Lane& c, const Cons& K, const QR& q, const UR& u, const ER* e, const SR& sv,
                               const DUR& du, int nu, double consTol, double projLimit, int opt, bool* changed) {
    const bool init = opt & PJ_INIT, force = opt & (PJ_INIT | PJ_FORCE), inf = opt & PJ_INF;
    *changed = false;
    sweep_kin<true, PAD>(c, q, u, NoPut{});
    cons_rows(c, K, REC_V, false, CWS_R, 1);
    const double vn0 = cons_norm(K, CWS_R, inf);
    if (vn0 > projLimit || vn0 != vn0) return false;
    if (vn0 == 0 || (vn0 <= consTol && !force)) return true;
    const double target = fmax(0.1*consTol, 1.9323057602939027e-14);
    for (int j = 0; j < nu; ++j) sv(j) = u(j);
    auto s2 = [&](int j) { const double a = fabs(sv(j)); const double sc = a > 1.0 ? a : 1.0; return sc*sc; };
    for (int k = 0; k < K.m; ++k) {
        cons_unit(K, CWS_Y, k);
        cons_sweep_jt<PAD>(c, ConsForce{&c, K, CWS_Y, 1.0, nullptr}, [&](int i, double v) { du(i) = v; });
        for (int j = 0; j < nu; ++j) du(j) = s2(j) * du(j);
        cons_sweep_j<PAD>(c, [&](int i) { return du(i); });
        cons_rows(c, K, REC_AGB, false, CWS_W + k, K.m);
    }
    auto solve_du = [&](int r) {
        if (!cons_solve(K, CWS_W, r, CWS_LAM)) return false;
        cons_sweep_jt<PAD>(c, ConsForce{&c, K, CWS_LAM, 1.0, nullptr}, [&](int i, double v) { du(i) = v; });
        for (int j = 0; j < nu; ++j) du(j) = s2(j) * du(j);
        return true;
    };
    double vn = vn0, prev = vn0;
    int its = 0;
    do {
        if (!solve_du(CWS_R)) return false;
        for (int j = 0; j < nu; ++j) u(j) = u(j) - du(j);
        *changed = true; ++its;
        sweep_kin<true, PAD>(c, q, u, NoPut{});
        cons_rows(c, K, REC_V, false, CWS_R, 1);
        vn = cons_norm(K, CWS_R, inf);
        if (!init && its >= 2 && vn > prev) {
            for (int j = 0; j < nu; ++j) u(j) = u(j) + du(j);
            sweep_kin<true, PAD>(c, q, u, NoPut{});
            cons_rows(c, K, REC_V, false, CWS_R, 1);
            vn = cons_norm(K, CWS_R, inf);
            break;
        }
        prev = vn;
    } while (vn > target && its < 7);
    if (!(vn <= consTol)) {
        if (!(vn < vn0)) for (int j = 0; j < nu; ++j) u(j) = sv(j);
        return false;
    }
    if (e) {
        cons_sweep_j<PAD>(c, [&](int i) { return (*e)(i); });
        cons_rows(c, K, REC_AGB, false, CWS_R, 1);
        if (!solve_du(CWS_R)) return false;
        for (int j = 0; j < nu; ++j) (*e)(j) = (*e)(j) - du(j);
    }
    return true;
}

// ---- event witnesses ----------------------------------------------------
// The trigger functions of TriggeredEventHandler / TriggeredEventReporter
// (EventHandler.h, EventReporter.h) over a fixed set of kinds, their
// EventTriggerInfo masks and windows (Event.h:196-258), and the handler
// actions (tdyn.h, TDYN_WIT_* and TDYN_EVENT_*).  Witness i's values live in
// the lane's event workspace at EW_E0 + i (the step's start) and in three
// rotating slots (window ends and trial).
struct Wit {
    const int4* wa; const double* wp; int n, need;
    double* ws; int L;
    __device__ __forceinline__ double& w(int k) const { return ws[(size_t)k * L]; }
};
// ws (nullable: the single-shot operator, which keeps no workspace).
__device__ __forceinline__ Wit make_wit(const DevTables& T, int tr, double* ws, int lane, int L) {
    Wit W{nullptr, nullptr, 0, 0, nullptr, L};
    if (!T.wit_t) return W;
    const int4 t = ldi4(T.wit_t, tr);
    W.wa = T.wit_a + t.x; W.wp = T.wit_p + 9*(size_t)t.x; W.n = t.y; W.need = t.z; W.ws = ws ? ws + lane : nullptr;
    return W;
}
// Witness values at (t, q, u) to out(i, value); the station kinds read the
// records of the kinematics of q, u (positions; velocities too when a
// velocity witness needs them).
template <bool PAD, class QG, class UG, class OUT>
__device__ void wit_eval(const Lane& c, const Wit& W, double t, const QG& q, const UG& u, const OUT& out) {
    if (W.need == 2) sweep_kin<true, PAD>(c, q, u, NoPut{});
    else if (W.need == 1) sweep_kin<false, PAD>(c, q, u, NoPut{});
    for (int i = 0; i < W.n; ++i) {
        const int4 wa = W.wa[i];
        const double* p = W.wp + 9*i;
        const int kind = wa.x & 15;
        double v;
        if (kind == WIT_Q) v = q(wa.y);
        else if (kind == WIT_U) v = u(wa.y);
        else if (kind == WIT_TIME) v = t;
        else {
            dm R = ident<double>();
            d3 o = zero3<double>();
            if (wa.z > 0) { const int rb = c.rec(wa.z); R = ldm(c, rb + REC_R); o = ld3(c, rb + REC_P); }
            const d3 r = mul(R, V3(p[0], p[1], p[2]));
            const d3 x = kind == WIT_STATION_POS ? o + r : st_vel(rec_sv(c, wa.z, REC_V), r);
            v = (p[3]*x.x + p[4]*x.y) + p[5]*x.z;
        }
        out(i, v - p[6]);
    }
}
// The same at the state rows yr, into the lane's workspace slots out + i.
template <bool PAD, class YR>
__device__ void wit_values(const Lane& c, const Wit& W, double t, const YR& yr, int out) {
    const SoAGet q{yr.base, yr.L, yr.lane}, u{yr.base + (size_t)c.nq * yr.L, yr.L, yr.lane};
    wit_eval<PAD>(c, W, t, q, u, [&](int i, double v) { W.w(out + i) = v; });
}
__device__ __forceinline__ int sign_of(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }
// estimateRootTime (IntegratorRep.h:351-376): secant with bias, bisection
// when an end value is exactly zero or the interval is at most minWindow,
// kept max(10% of the interval, minWindow/2) away from either end.
__device__ __forceinline__ double root_time(double tLow, double fLow, double tHigh, double fHigh, double bias,
                                            double minWindow) {
    const double h = tHigh - tLow;
    if (fLow == 0 || fHigh == 0 || h <= minWindow) return tLow + h/2;
    const double x = fHigh/(fHigh - bias*fLow);
    double tr = tHigh - x*h;
    const double buf = fmax(0.1*h, minWindow/2);
    tr = fmax(tr, tLow + buf);
    tr = fmin(tr, tHigh - buf);
    return tr;
}
// findEventCandidates (IntegratorRep.h:390-443) over the witnesses of the
// mask `viable`, window ends at workspace slots lo and hi: the candidate
// mask (Event::classifyTransition, Event.h:177-189, under each witness's
// trigger mask), the candidates reported rising, the earliest root estimate
// and the narrowest required window max(min(acc*timescale*window), minWindow).
__device__ unsigned find_cands(const Wit& W, unsigned viable, double tLow, int lo, double tHigh, int hi,
                               double bias, double minWindow, double accTs, unsigned& rising, double& earliest,
                               double& narrowest) {
    unsigned cand = 0;
    rising = 0;
    earliest = narrowest = __builtin_inf();
    for (int i = 0; i < W.n; ++i) {
        if (!((viable >> i) & 1u)) continue;
        const double el = W.w(lo + i), eh = W.w(hi + i);
        const int sl = sign_of(el), sh = sign_of(eh);
        const int seen = ((sl == sh || sl == 0) ? 0 : (sl == 1 ? 1 : 2)) & ((W.wa[i].x >> 4) & 3);
        if (!seen) continue;
        cand |= 1u << i;
        if (seen & 2) rising |= 1u << i;
        narrowest = fmax(fmin(narrowest, accTs * W.wp[9*i + 8]), minWindow);
        earliest = fmin(earliest, root_time(tLow, el, tHigh, eh, bias, minWindow));
    }
    return cand;
}

// ---------------------------------------------------------------------------
// Operator kernels (the batched SimbodyMatterSubsystem operators).
// ---------------------------------------------------------------------------
// Per-instance status of the operators that form articulated inertias:
// 0, or TDYN_E_SINGULAR (4) when some body's D was not invertible.
__device__ __forceinline__ void op_status(const OpArgs& a, int lane, bool singular) {
    if (a.status) a.status[a.inst ? a.inst[lane] : lane] = singular ? 4 : 0;
}
template <int OP, bool UNI>
__global__ void __launch_bounds__(64) k_op(OpArgs a) {
    constexpr bool PAD = !UNI;      // batches of different trees use padded records
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= a.n_inst) return;
    const int tr = UNI ? 0 : a.tree_of[lane];
    const Lane c = make_lane_t(a.T, tr, lane, a.L, a.scratch, PAD ? a.rec_stride : 0);
    const double* grav = a.T.gravity + 4*tr;
    const int qo = a.off.q_off[lane], uo = a.off.u_off[lane], bo = a.off.b_off[lane];
    const PackedGet q{a.q + qo};
    const PackedGet u{a.u ? a.u + uo : nullptr};
    if constexpr (OP == OP_KIN) {
        sweep_kin<true, PAD>(c, q, u, PackedPut{a.out_q ? a.out_q + qo : nullptr});
        double ke = 0, pe = 0;
        const d3 g = {grav[0], grav[1], grav[2]};
        for (int b = 1; b <= c.nb; ++b) {
            const int rb = c.rec(b);
            const double m = c.prm(b)[PRM_MASS];
            const dv V = ldsv(c, rb + REC_V);
            const d3 Mk_p = ld3(c, rb + REC_MKP);
            ke += dot(V, rbi_mul(m, Mk_p, lds(c, rb + REC_MKG), V)) * 0.5;
            pe -= m * dot(g, ld3(c, rb + REC_P) + Mk_p);
            if (a.out_body) {
                double* X = a.out_body + 12*(size_t)(bo + b);
                const dm R = ldm(c, rb + REC_R); const d3 pp = ld3(c, rb + REC_P);
                for (int i = 0; i < 9; ++i) X[i] = R.a[i];
                X[9] = pp.x; X[10] = pp.y; X[11] = pp.z;
            }
            if (a.out_body2) {
                double* Vo = a.out_body2 + 6*(size_t)(bo + b);
                Vo[0]=V.w.x; Vo[1]=V.w.y; Vo[2]=V.w.z; Vo[3]=V.v.x; Vo[4]=V.v.y; Vo[5]=V.v.z;
            }
        }
        if (a.out_body) { double* X = a.out_body + 12*(size_t)bo; for (int i = 0; i < 12; ++i) X[i] = (i == 0 || i == 4 || i == 8) ? 1.0 : 0.0; }
        if (a.out_body2) { double* Vo = a.out_body2 + 6*(size_t)bo; for (int i = 0; i < 6; ++i) Vo[i] = 0.0; }
        if (a.out_scalar) {
            const int inst = a.inst ? a.inst[lane] : lane;
            a.out_scalar[2*inst] = ke; a.out_scalar[2*inst+1] = pe;
        }
    } else if constexpr (OP == OP_FWD_PRES) {
        // forward dynamics with known udot on the flagged bodies: udot (known
        // and computed) to out_u, tau to out_tau, A_GB to out_body
        sweep_kin<true, PAD>(c, q, u, NoPut{});
        const BodyForce fb{a.fbody ? a.fbody + 6*(size_t)bo : nullptr};
        const PackedGetOpt fm{a.fmob ? a.fmob + uo : nullptr};
        const Pres pr{a.presc + bo, a.udot_in + uo, a.out_tau + uo};
        op_status(a, lane, sweep_inward<IN_FWD | IN_PRES, PAD>(c, fb, fm, nullptr, nullptr, grav, pr));
        sweep_outward2<true, PAD, true>(c, q, u, PackedPut{a.out_u + uo}, NoPut{}, pr);
        if (a.out_body) {
            for (int b = 0; b <= c.nb; ++b) {
                const dv A = b == 0 ? svzero<double>() : ldsv(c, c.rec(b) + REC_AGB);
                double* o = a.out_body + 6*(size_t)(bo + b);
                o[0]=A.w.x; o[1]=A.w.y; o[2]=A.w.z; o[3]=A.v.x; o[4]=A.v.y; o[5]=A.v.z;
            }
        }
    } else if constexpr (OP == OP_WIT) {
        // realize to the witnesses' stages and read the event triggers
        // (State::getEventTriggers); zero past the tree's witnesses
        const Wit W = make_wit(a.T, tr, nullptr, lane, a.L);
        double* o = a.out_wit + (size_t)WIT_MAX * (a.inst ? a.inst[lane] : lane);
        const double t = a.times ? a.times[a.inst ? a.inst[lane] : lane] : 0.0;
        for (int i = 0; i < WIT_MAX; ++i) o[i] = 0.0;
        wit_eval<PAD>(c, W, t, q, u, [&](int i, double v) { o[i] = v; });
    } else if constexpr (OP == OP_FWD_CONS) {
        // realize(Acceleration) with constraints: loop forward dynamics
        // (gravity and the applied body forces), udot and the multipliers
        const Cons K = make_cons(a.T, tr, a.cons_ws, lane, a.L);
        const BodyForce fb{a.fbody ? a.fbody + 6*(size_t)bo : nullptr};
        const PackedPut uout{a.out_u ? a.out_u + uo : nullptr};
        bool ok = true;
        const PackedPut qout{a.out_q ? a.out_q + qo : nullptr};
        if (K.m) {
            ok = cons_derivs<PAD>(c, K, a.fbody ? a.fbody + 6*(size_t)bo : nullptr, grav, q, u, qout, uout);
        } else {
            sweep_kin<true, PAD>(c, q, u, qout);
            ok = !sweep_inward<IN_FWD | IN_SYSF, PAD>(c, fb, Zero{}, nullptr, nullptr, grav);
            sweep_outward2<true, PAD>(c, q, u, uout, NoPut{});
        }
        if (a.out_lam) {
            double* lo = a.out_lam + (size_t)CONS_MAXR * (a.inst ? a.inst[lane] : lane);
            for (int i = 0; i < CONS_MAXR; ++i) lo[i] = (ok && i < K.m) ? K.w(CWS_LAM + i) : 0.0;
        }
        op_status(a, lane, !ok);
    } else if constexpr (OP == OP_FWD_SYS || OP == OP_FWD_GIVEN || OP == OP_REACT) {
        sweep_kin<true, PAD>(c, q, u, PackedPut{(OP == OP_FWD_SYS && a.out_q) ? a.out_q + qo : nullptr});
        const BodyForce fb{a.fbody ? a.fbody + 6*(size_t)bo : nullptr};
        const PackedGetOpt fm{(OP != OP_FWD_SYS && a.fmob) ? a.fmob + uo : nullptr};
        if constexpr (OP == OP_FWD_SYS) op_status(a, lane, sweep_inward<IN_FWD | IN_SYSF, PAD>(c, fb, fm, nullptr, nullptr, grav));
        else op_status(a, lane, sweep_inward<IN_FWD, PAD>(c, fb, fm, nullptr, nullptr, grav));
        const PackedPut uout{a.out_u ? a.out_u + uo : nullptr};
        if constexpr (OP == OP_FWD_GIVEN) {
            sweep_outward2<true, PAD>(c, q, u, uout, PackedPut{a.out_q ? a.out_q + qo : nullptr});
        } else {
            sweep_outward2<true, PAD>(c, q, u, uout, NoPut{});
        }
        if constexpr (OP == OP_REACT) {
            // Mobilizer reaction forces at M, in G (SimbodyMatterSubsystemRep.cpp:5786-5826).
            for (int b = 0; b <= c.nb; ++b) {
                double* o = a.out_body + 6*(size_t)(bo + b);
                if (b == 0) { for (int i = 0; i < 6; ++i) o[i] = 0.0; continue; }
                const int4 ba = c.ba(b);
                const int rb = c.rec(b);
                const d3 pB = ld3(c, rb + REC_P);
                dv A_GP = svzero<double>(); d3 pP = zero3<double>();
                if (ba.x != 0) { const int rp = c.rec(ba.x); A_GP = ldsv(c, rp + REC_AGB); pP = ld3(c, rp + REC_P); }
                const d3 l = pB - pP;
                const dv Ap = {A_GP.w, A_GP.v + cross(A_GP.w, l)};
                const dv FB = ldsv(c, rb + REC_ZPLUS) + abi_mul(lda(c, rb + REC_PPLUS), Ap);
                const d3 pBM = mul(ldm(c, rb + REC_R), prm_3(c.prm(b) + PRM_P_BM));
                const d3 mo = FB.w - cross(pBM, FB.v);
                o[0]=mo.x; o[1]=mo.y; o[2]=mo.z; o[3]=FB.v.x; o[4]=FB.v.y; o[5]=FB.v.z;
            }
        } else if (OP == OP_FWD_GIVEN && a.out_body) {
            for (int b = 0; b <= c.nb; ++b) {
                const dv A = b == 0 ? svzero<double>() : ldsv(c, c.rec(b) + REC_AGB);
                double* o = a.out_body + 6*(size_t)(bo + b);
                o[0]=A.w.x; o[1]=A.w.y; o[2]=A.w.z; o[3]=A.v.x; o[4]=A.v.y; o[5]=A.v.z;
            }
        }
    } else if constexpr (OP == OP_MINV) {
        sweep_kin<false, PAD>(c, q, u, NoPut{});
        const PackedGet f{a.fmob + uo};
        op_status(a, lane, sweep_inward<IN_MINV, PAD>(c, BodyForce{nullptr}, f, nullptr, nullptr, grav));
        sweep_outward2<false, PAD>(c, q, u, PackedPut{a.out_u + uo}, NoPut{});
    } else if constexpr (OP == OP_ABI) {
        sweep_kin<false, PAD>(c, q, u, NoPut{});
        op_status(a, lane, sweep_inward<IN_ABI_OUT, PAD>(c, BodyForce{nullptr}, Zero{}, a.out_abi + 21*(size_t)bo,
                                                         a.out_abi2 ? a.out_abi2 + 21*(size_t)bo : nullptr, grav));
        for (int i = 0; i < 21; ++i) {
            a.out_abi[21*(size_t)bo + i] = 0.0;
            if (a.out_abi2) a.out_abi2[21*(size_t)bo + i] = 0.0;
        }
    } else if constexpr (OP == OP_MULM || OP == OP_JAC) {
        sweep_kin<false, PAD>(c, q, u, NoPut{});
        const PackedGet v{a.udot_in + uo};
        for (int b = 1; b <= c.nb; ++b) {
            const int4 ba = c.ba(b);
            switch (TD_RDOF(ba)) {
            case 1: jac_out_body<1>(c, b, ba, v); break;
            case 2: jac_out_body<2>(c, b, ba, v); break;
            case 3: jac_out_body<3>(c, b, ba, v); break;
            default: jac_out_body<6>(c, b, ba, v); break;
            }
        }
        if constexpr (OP == OP_JAC) {
            for (int b = 0; b <= c.nb; ++b) {
                const dv A = b == 0 ? svzero<double>() : ldsv(c, c.rec(b) + REC_AGB);
                double* o = a.out_body + 6*(size_t)(bo + b);
                o[0]=A.w.x; o[1]=A.w.y; o[2]=A.w.z; o[3]=A.v.x; o[4]=A.v.y; o[5]=A.v.z;
            }
        } else {
            const PackedPut tau{a.out_u + uo};
            for (int b = c.nb; b >= 1; --b) {
                const int4 ba = c.ba(b), bb = c.bb(b);
                switch (TD_RDOF(ba)) {
                case 1: jact_in_body<1, true>(c, b, ba, bb, BodyForce{nullptr}, tau); break;
                case 2: jact_in_body<2, true>(c, b, ba, bb, BodyForce{nullptr}, tau); break;
                case 3: jact_in_body<3, true>(c, b, ba, bb, BodyForce{nullptr}, tau); break;
                default: jact_in_body<6, true>(c, b, ba, bb, BodyForce{nullptr}, tau); break;
                }
            }
        }
    } else if constexpr (OP == OP_JACT) {
        sweep_kin<false, PAD>(c, q, u, NoPut{});
        const BodyForce fb{a.fbody + 6*(size_t)bo};
        const PackedPut tau{a.out_u + uo};
        for (int b = c.nb; b >= 1; --b) {
            const int4 ba = c.ba(b), bb = c.bb(b);
            switch (TD_RDOF(ba)) {
            case 1: jact_in_body<1, false>(c, b, ba, bb, fb, tau); break;
            case 2: jact_in_body<2, false>(c, b, ba, bb, fb, tau); break;
            case 3: jact_in_body<3, false>(c, b, ba, bb, fb, tau); break;
            default: jact_in_body<6, false>(c, b, ba, bb, fb, tau); break;
            }
        }
    } else if constexpr (OP == OP_BACC) {
        // calcBodyAccelerationFromUDot (SimbodyMatterSubsystemRep.cpp:3592-3625)
        sweep_kin<true, PAD>(c, q, u, NoPut{});
        const PackedGetOpt ud{a.udot_in ? a.udot_in + uo : nullptr};
        for (int b = 1; b <= c.nb; ++b) {
            const int4 ba = c.ba(b);
            switch (TD_RDOF(ba)) {
            case 1: resid_out_body<1>(c, b, ba, ud); break;
            case 2: resid_out_body<2>(c, b, ba, ud); break;
            case 3: resid_out_body<3>(c, b, ba, ud); break;
            default: resid_out_body<6>(c, b, ba, ud); break;
            }
        }
        for (int b = 0; b <= c.nb; ++b) {
            const dv A = b == 0 ? svzero<double>() : ldsv(c, c.rec(b) + REC_AGB);
            double* o = a.out_body + 6*(size_t)(bo + b);
            o[0]=A.w.x; o[1]=A.w.y; o[2]=A.w.z; o[3]=A.v.x; o[4]=A.v.y; o[5]=A.v.z;
        }
    } else if constexpr (OP == OP_CBI) {
        // calcCompositeBodyInertias (SimbodyMatterSubsystemRep.cpp:5196-5205);
        // Ground's entry is its infinite mass (RigidBodyNode_Weld.cpp:168-171)
        sweep_kin<false, PAD>(c, q, u, NoPut{});
        for (int b = c.nb; b >= 1; --b) cbi_in_body(c, b, c.bb(b));
        for (int b = 0; b <= c.nb; ++b) {
            double* o = a.out_body + 10*(size_t)(bo + b);
            if (b == 0) {
                o[0] = __builtin_inf();
                for (int i = 1; i < 10; ++i) o[i] = (i >= 4 && i <= 6) ? 1.0 : 0.0;
                continue;
            }
            const int rb = c.rec(b);
            for (int i = 0; i < 10; ++i) o[i] = c.s(rb + REC_PPLUS + i);
        }
    } else if constexpr (OP == OP_RESID) {
        sweep_kin<true, PAD>(c, q, u, NoPut{});
        const PackedGetOpt ud{a.udot_in ? a.udot_in + uo : nullptr};
        const PackedGetOpt fm{a.fmob ? a.fmob + uo : nullptr};
        const BodyForce fb{a.fbody ? a.fbody + 6*(size_t)bo : nullptr};
        const PackedPut tau{a.out_u + uo};
        for (int b = 1; b <= c.nb; ++b) {
            const int4 ba = c.ba(b);
            switch (TD_RDOF(ba)) {
            case 1: resid_out_body<1>(c, b, ba, ud); break;
            case 2: resid_out_body<2>(c, b, ba, ud); break;
            case 3: resid_out_body<3>(c, b, ba, ud); break;
            default: resid_out_body<6>(c, b, ba, ud); break;
            }
        }
        for (int b = c.nb; b >= 1; --b) {
            const int4 ba = c.ba(b), bb = c.bb(b);
            switch (TD_RDOF(ba)) {
            case 1: resid_in_body<1>(c, b, ba, bb, fb, fm, tau); break;
            case 2: resid_in_body<2>(c, b, ba, bb, fb, fm, tau); break;
            case 3: resid_in_body<3>(c, b, ba, bb, fb, fm, tau); break;
            default: resid_in_body<6>(c, b, ba, bb, fb, fm, tau); break;
            }
        }
    }
}

#undef TD_RDOF

template <bool UNI>
static void launch_op_t(OpKind k, const OpArgs& a, hipStream_t s) {
    const dim3 blk(64), grd((a.n_inst + 63) / 64);
    switch (k) {
    case OP_KIN:       k_op<OP_KIN, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_FWD_SYS:   k_op<OP_FWD_SYS, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_FWD_GIVEN: k_op<OP_FWD_GIVEN, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_MINV:      k_op<OP_MINV, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_RESID:     k_op<OP_RESID, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_ABI:       k_op<OP_ABI, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_REACT:     k_op<OP_REACT, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_MULM:      k_op<OP_MULM, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_JAC:       k_op<OP_JAC, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_JACT:      k_op<OP_JACT, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_BACC:      k_op<OP_BACC, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_CBI:       k_op<OP_CBI, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_FWD_PRES:  k_op<OP_FWD_PRES, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_FWD_CONS:  k_op<OP_FWD_CONS, UNI><<<grd, blk, 0, s>>>(a); break;
    case OP_WIT:       k_op<OP_WIT, UNI><<<grd, blk, 0, s>>>(a); break;
    }
}
hipError_t launch_op(OpKind k, const OpArgs& a, hipStream_t s) {
    if (a.n_inst <= 0) return hipSuccess;
    if (a.tree_of) launch_op_t<false>(k, a, s); else launch_op_t<true>(k, a, s);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Runge-Kutta-Merson rollouts with per-instance step control.  One lane runs
// one instance's TimeStepper loop: RungeKuttaMersonIntegrator.cpp:86-140
// (stages), AbstractIntegratorRep.cpp:137-208 (projection gate), :448-502
// (adjustStepSize), :513-609 (takeOneStep), IntegratorRep.h:454-523, 644-655
// (error norm), SimbodyMatterSubsystemRep.cpp:4096-4184, 4448-4475 (quaternion
// projection).  Lanes stop after cfg.max_steps accepted steps so a rollout can
// be resumed by the next launch (the serving loop relies on this).
// ---------------------------------------------------------------------------

// Derivative evaluators: y (this lane's rows: q then u) -> f (qdot then udot).
template <bool PAD>
struct TreeEval {
    static constexpr bool kMotion = true;     // carries prescribed motion
    static constexpr bool kCons = true;       // carries holonomic constraints
    static constexpr bool kPad = PAD;
    using VR = chain::RowsT<double, false>;
    static __device__ __forceinline__ VR vrow(const RkmArgs& a, int v, int lane) {
        return VR{a.vec + (size_t)v * a.ny_max * a.L, a.L, (unsigned)lane};
    }
    Lane c; const double* grav; const double* fapp;
    Pres pr;   // known udot (locks, constant prescribed udot): pr.m null when none
    Cons K;    // the lane's constraints (K.m = 0: none)
    __device__ TreeEval(const RkmArgs& a, int lane, int tr)
        : c(make_lane_t(a.T, tr, lane, a.L, a.scratch, PAD ? a.rec_stride : 0)), grav(a.T.gravity + 4*tr),
          fapp(a.fapp ? a.fapp + 6*(size_t)a.off.b_off[lane] : nullptr),
          pr{a.presc ? a.presc + a.off.b_off[lane] : nullptr, a.presc ? a.udot_known + a.off.u_off[lane] : nullptr,
             nullptr},
          K(make_cons(a.T, tr, a.cons_ws, lane, a.L)) {}
    __device__ __forceinline__ void derivs(const chain::RowsT<double, false>& yr, const chain::RowsT<double, false>& fr) const {
        const SoAGet y{yr.base, yr.L, yr.lane}, u{yr.base + (size_t)c.nq * yr.L, yr.L, yr.lane};
        const SoAPut f{fr.base, fr.L, fr.lane}, fu{fr.base + (size_t)c.nq * fr.L, fr.L, fr.lane};
        if (K.m) {
            // loop forward dynamics; a singular G M^-1 G^T fails the step (NaN udot)
            if (!cons_derivs<PAD>(c, K, fapp, grav, y, u, f, fu))
                for (int k = 0; k < c.nu; ++k) fu(k, __builtin_nan(""));
            return;
        }
        sweep_kin<true, PAD>(c, y, u, f);
        if (pr.m) {
            sweep_inward<IN_FWD | IN_SYSF | IN_PRES, PAD>(c, BodyForce{fapp}, Zero{}, nullptr, nullptr, grav, pr);
            sweep_outward2<true, PAD, true>(c, y, u, fu, NoPut{}, pr);
        } else {
            sweep_inward<IN_FWD | IN_SYSF, PAD>(c, BodyForce{fapp}, Zero{}, nullptr, nullptr, grav);
            sweep_outward2<true, PAD>(c, y, u, fu, NoPut{});
        }
    }
};

// Tree-server slots: the tree sweeps on a slot's own pooled regions (its
// records, lane-contiguous, and its state vectors of its own size), so that
// slots of different trees take only the memory their trees need.
struct TreePoolEval {
    static constexpr bool kMotion = false;
    static constexpr bool kCons = false;
    static constexpr bool kPad = true;
    using VR = tstream::LaneVec;
    static __device__ __forceinline__ VR vrow(const RkmArgs& a, int v, int lane) {
        const int4 ta = ldi4(a.T.tree_a, lane);
        return VR{a.vec + a.pool[2*lane] + (size_t)v * (ta.y + ta.z)};
    }
    Lane c; const double* grav; const double* fapp;
    __device__ TreePoolEval(const RkmArgs& a, int lane, int tr)
        : c(make_lane_t(a.T, tr, 0, 1, a.scratch + a.pool[2*lane + 1], 1)), grav(a.T.gravity + 4*tr),
          fapp(a.fapp + 6*(size_t)a.off.b_off[lane]) {}
    __device__ __forceinline__ void derivs(const VR& yr, const VR& fr) const {
        const PackedGet y{yr.p}, u{yr.p + c.nq};
        const PackedPut f{fr.p}, fu{fr.p + c.nq};
        sweep_kin<true, true>(c, y, u, f);
        sweep_inward<IN_FWD | IN_SYSF, true>(c, BodyForce{fapp}, Zero{}, nullptr, nullptr, grav);
        sweep_outward2<true, true>(c, y, u, fu, NoPut{});
    }
};

// ---------------------------------------------------------------------------
// Streaming tree evaluator (rollouts of batches of different trees).  Each
// lane walks its tree in depth-first preorder (parents first, every subtree
// contiguous) with the chain evaluator's register-resident pass bodies on
// padded generic bodies:
//   pass 1 (preorder): the parent's pose comes from registers when the
//          parent is the previous body, else from a per-depth stack; the
//          body's pose is written to its position's hand-over row;
//   pass 2 (reverse preorder): the body's pose is read back; the children's
//          articulated inertias and bias forces arrive in registers from the
//          previous body (the first child) plus a per-depth accumulator of the
//          other children; the body's own P+, z+ go to the parent in
//          registers (when the parent is next) or into its accumulator;
//   pass 3 (preorder): as pass 1, with the parent's acceleration.
// All per-lane data is lane-contiguous (the lanes of a wave walk different
// trees, so no access pattern would coalesce across lanes).
// ---------------------------------------------------------------------------

struct TreeStreamEval {
    static constexpr bool kMotion = false;
    static constexpr bool kCons = false;
    static constexpr bool kPad = true;
    using VR = tstream::LaneVec;
    DevTables T;
    int nb, nq, body_off, acc0, pos0;
    const int4* pre;
    double* S;
    d3 grav; const double* fapp;
    static __device__ __forceinline__ VR vrow(const RkmArgs& a, int v, int lane) {
        return VR{a.vec + ((size_t)v * a.L + lane) * a.ny_max};
    }
    __device__ TreeStreamEval(const RkmArgs& a, int lane, int tr) : T(a.T) {
        const int4 ta = ldi4(a.T.tree_a, tr);
        nb = ta.x; nq = ta.y; body_off = ta.w;
        pre = a.T.pre + body_off;
        S = a.scratch + (size_t)lane * a.rec_stride;
        acc0 = tstream::ST_W * (a.stream_dmax + 1);
        pos0 = acc0 + tstream::ACC_W * (a.stream_dmax + 1);
        const double* g = a.T.gravity + 4*tr;
        grav = V3(g[0], g[1], g[2]);
        fapp = a.fapp ? a.fapp + 6*(size_t)a.off.b_off[lane] : nullptr;
        // accumulators start at zero; every reader clears what it read
        for (int k = 0; k < tstream::ACC_W * (a.stream_dmax + 1); ++k) S[acc0 + k] = 0.0;
    }
    __device__ __forceinline__ int pos(int i) const {
        return pos0 + (i == 0 ? 0 : tstream::POS0_W + (i - 1) * tstream::POS_W);
    }
    __device__ __forceinline__ tstream::LaneStore store(int i) const {
        const int o = pos(i);
        return i == 0 ? tstream::LaneStore{S, o + 18, o + 60, o + 54} : tstream::LaneStore{S, o + 18, o + 39, o + 36};
    }
    // (every generic body checks its reversed-mobilizer flag)
    __device__ __forceinline__ chain::ExtF<cdptr> prm(int b) const {
        return chain::ExtF<cdptr>(cptr(T.prm) + (size_t)(body_off + b - 1) * PRM_STRIDE);
    }
    __device__ __forceinline__ dv F(int b) const {
        if (!fapp) return svzero<double>();
        const double* f = fapp + 6*b;
        return {V3(f[0], f[1], f[2]), V3(f[3], f[4], f[5])};
    }
    // q, u (zero-padded), trig cache and (G_ROWS) G, DI*eps of position i
    template <int TYPE, bool G_ROWS>
    __device__ __forceinline__ void load(chain::Pre<double>& x, const VR& y, int4 ba, const tstream::LaneStore& st) const {
        if constexpr (TYPE == FREE) {
            x.ty = FREE;
#pragma unroll
            for (int k = 0; k < 7; ++k) x.q[k] = y(ba.z + k);
#pragma unroll
            for (int k = 0; k < 6; ++k) x.u[k] = y(nq + ba.w + k);
#pragma unroll
            for (int k = 0; k < 4; ++k) x.tc[k] = st.T(k);
        } else {
            const int ty = ba.y, d = mob_dof(ty), nqb = mob_nq(ty);
            x.ty = ty;
#pragma unroll
            for (int k = 0; k < 4; ++k) x.q[k] = k < nqb ? y(ba.z + k) : 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) x.u[k] = k < d ? y(nq + ba.w + k) : 0.0;
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
    __device__ __forceinline__ void ld_pose(const double* s, dm& R, d3& p, dv& V) const {
#pragma unroll
        for (int k = 0; k < 9; ++k) R.a[k] = s[k];
        p = V3(s[9], s[10], s[11]);
        V = {V3(s[12], s[13], s[14]), V3(s[15], s[16], s[17])};
    }
    __device__ __forceinline__ void st_pose(double* s, const dm& R, const d3& p, const dv& V) const {
#pragma unroll
        for (int k = 0; k < 9; ++k) s[k] = R.a[k];
        s[9] = p.x; s[10] = p.y; s[11] = p.z;
        s[12] = V.w.x; s[13] = V.w.y; s[14] = V.w.z; s[15] = V.v.x; s[16] = V.v.y; s[17] = V.v.z;
    }
    // Pass 1 at position i: the parent's pose (ground, registers or stack) to
    // this body's; qdot and trig cache out; pose to the position record and
    // to the depth stack.
    template <int TYPE>
    __device__ __forceinline__ void p1(int i, int& prev_d, const VR& y, const VR& f, dm& R, d3& p, dv& V) const {
        const int4 pe = ldi4(pre, i);
        const int b = pe.x, d = pe.y;
        const int4 ba = ldi4(T.body_a, body_off + b - 1);
        if (d == 1) { R = ident<double>(); p = zero3<double>(); V = svzero<double>(); }
        else if (prev_d != d - 1) ld_pose(S + tstream::ST_W * (d - 1), R, p, V);
        const tstream::LaneStore st = store(i);
        chain::Pre<double> x;
        load<TYPE, false>(x, y, ba, st);
        chain::pass1_body<TYPE, double>(prm(b), x, ba.z, f, st, R, p, V);
        st_pose(S + pos(i), R, p, V);
        if (pe.z > 1) st_pose(S + tstream::ST_W * d, R, p, V);   // later children need it
        prev_d = d;
    }
    // Pass 2 at position i (reverse order): P, D, G, z of this body from its
    // first child (registers, hc) and its other children (accumulator);
    // P+, z+ to the parent in registers or added into the parent's
    // accumulator.
    template <int TYPE>
    __device__ __forceinline__ void p2(int i, int& prev_d, const VR& y, chain::da& PP, chain::dv& zp, d3& phic) const {
        const int4 pe = ldi4(pre, i);
        const int b = pe.x, d = pe.y;
        const int4 ba = ldi4(T.body_a, body_off + b - 1);
        const bool hc = i + 1 < nb && prev_d == d + 1;
        dm R; d3 p; dv V;
        ld_pose(S + pos(i), R, p, V);
        const tstream::LaneStore st = store(i);
        chain::Pre<double> x;
        load<TYPE, false>(x, y, ba, st);
        if (pe.z > 1) {
            double* ac = S + acc0 + tstream::ACC_W * d;
            chain::da aP;
            aP.M = {ac[0], ac[1], ac[2], ac[3], ac[4], ac[5]};
            aP.J = {ac[6], ac[7], ac[8], ac[9], ac[10], ac[11]};
#pragma unroll
            for (int k = 0; k < 9; ++k) aP.F.a[k] = ac[12 + k];
            const dv aZ = {V3(ac[21], ac[22], ac[23]), V3(ac[24], ac[25], ac[26])};
#pragma unroll
            for (int k = 0; k < tstream::ACC_W; ++k) ac[k] = 0.0;
            chain::pass2_body<TYPE, double>(prm(b), x, grav, F(b), st, R, p, V, hc, PP, zp, phic, &aP, &aZ);
        } else {
            chain::pass2_body<TYPE, double>(prm(b), x, grav, F(b), st, R, p, V, hc, PP, zp, phic);
        }
        // PP, zp, phic: this body's P+, z+ and offset from its parent
        if (d > 1 && !(i >= 1 && ldi4(pre, i - 1).y == d - 1)) {
            double* ap = S + acc0 + tstream::ACC_W * (d - 1);
            chain::da Q;
            Q.M = {ap[0], ap[1], ap[2], ap[3], ap[4], ap[5]};
            Q.J = {ap[6], ap[7], ap[8], ap[9], ap[10], ap[11]};
#pragma unroll
            for (int k = 0; k < 9; ++k) Q.F.a[k] = ap[12 + k];
            abi_add_shifted(Q, PP, phic);
            const dv z = dv{V3(ap[21], ap[22], ap[23]), V3(ap[24], ap[25], ap[26])} + phi_mul(phic, zp);
            ap[0] = Q.M.xx; ap[1] = Q.M.yy; ap[2] = Q.M.zz; ap[3] = Q.M.xy; ap[4] = Q.M.xz; ap[5] = Q.M.yz;
            ap[6] = Q.J.xx; ap[7] = Q.J.yy; ap[8] = Q.J.zz; ap[9] = Q.J.xy; ap[10] = Q.J.xz; ap[11] = Q.J.yz;
#pragma unroll
            for (int k = 0; k < 9; ++k) ap[12 + k] = Q.F.a[k];
            ap[21] = z.w.x; ap[22] = z.w.y; ap[23] = z.w.z; ap[24] = z.v.x; ap[25] = z.v.y; ap[26] = z.v.z;
        }
        prev_d = d;
    }
    // Pass 3 at position i: udot = DI eps - G^T A+ and A_GB.
    template <int TYPE>
    __device__ __forceinline__ void p3(int i, int& prev_d, const VR& y, const VR& fu, dm& R, d3& p, dv& V, dv& A) const {
        const int4 pe = ldi4(pre, i);
        const int b = pe.x, d = pe.y;
        const int4 ba = ldi4(T.body_a, body_off + b - 1);
        if (d == 1) { R = ident<double>(); p = zero3<double>(); V = svzero<double>(); A = svzero<double>(); }
        else if (prev_d != d - 1) {
            const double* s = S + tstream::ST_W * (d - 1);
            ld_pose(s, R, p, V);
            A = {V3(s[18], s[19], s[20]), V3(s[21], s[22], s[23])};
        }
        const tstream::LaneStore st = store(i);
        chain::Pre<double> x;
        load<TYPE, true>(x, y, ba, st);
        chain::pass3_body<TYPE, double>(prm(b), x, ba.w, fu, st, R, p, V, A);
        if (pe.z > 1) {
            double* s = S + tstream::ST_W * d;
            st_pose(s, R, p, V);
            s[18] = A.w.x; s[19] = A.w.y; s[20] = A.w.z; s[21] = A.v.x; s[22] = A.v.y; s[23] = A.v.z;
        }
        prev_d = d;
    }
    __device__ __forceinline__ void derivs(const VR& y, const VR& f) const {
        const VR fu{f.p + nq};
        // position 0 is body 1, the only body that may be Free
        const bool free0 = ldi4(T.body_a, body_off).y == FREE;
        {
            dm R; d3 p; dv V; int pd = 0;
            if (free0) p1<FREE>(0, pd, y, f, R, p, V); else p1<GEN>(0, pd, y, f, R, p, V);
            for (int i = 1; i < nb; ++i) p1<GEN>(i, pd, y, f, R, p, V);
        }
        {
            chain::da PP; dv zp = svzero<double>(); d3 phic = zero3<double>(); int pd = 0;
            for (int i = nb - 1; i >= 1; --i) p2<GEN>(i, pd, y, PP, zp, phic);
            if (free0) p2<FREE>(0, pd, y, PP, zp, phic); else p2<GEN>(0, pd, y, PP, zp, phic);
        }
        {
            dm R; d3 p; dv V, A; int pd = 0;
            if (free0) p3<FREE>(0, pd, y, fu, R, p, V, A); else p3<GEN>(0, pd, y, fu, R, p, V, A);
            for (int i = 1; i < nb; ++i) p3<GEN>(i, pd, y, fu, R, p, V, A);
        }
    }
};

// One lane runs one instance's TimeStepper loop.  The state y0 lives in
// vector VY; a step evaluates f0 = f(y0) once and then, per attempt, the four
// Merson stages (RungeKuttaMersonIntegrator.cpp:86-140) through ONE inlined
// evaluator inside a stage loop.
// Prescribed motion at time t (prescribeQ/prescribeU, SimbodyMatterSubsystemRep.cpp:
// 3776-3830, and the prescribed udot; Motion::Sinusoid m(t) = a sin(w t + phi)
// and Motion::Steady, MotionImpl.h:182-300): q and u of the state rows y
// overwritten where the motion prescribes them (if qu), the prescribed udot
// into the lane's known-udot row.  Position level only on mobilizers with
// qdot = u (checked by the host).
template <class YR>
__device__ void prescribe_lane(const RkmArgs& a, int lane, int tr, double t, const YR& y, bool qu) {
    const int4 ta = ldi4(a.T.tree_a, tr);
    const int nb = ta.x, nq = ta.y;
    const double* m0 = a.motion + 4*(size_t)a.off.b_off[lane];
    double* ud = const_cast<double*>(a.udot_known) + a.off.u_off[lane];
    for (int b = 1; b <= nb; ++b) {
        const double* m = m0 + 4*b;
        const int level = (int)m[0];
        if (level == 0) continue;
        const int4 ba = ldi4(a.T.body_a, ta.w + b - 1);
        const int d = mob_dof(ba.y);
        double sn, cs;
        sincos(m[2]*t + m[3], &sn, &cs);
        const double f = m[1]*sn, fd = m[1]*m[2]*cs, fdd = -m[1]*m[2]*m[2]*sn;
        for (int i = 0; i < d; ++i) {
            if (level == 1) {
                if (qu) { y(ba.z + i) = f; y(nq + ba.w + i) = fd; }
                ud[ba.w + i] = fdd;
            } else if (level == 2) {
                if (qu) y(nq + ba.w + i) = f;
                ud[ba.w + i] = fd;
            } else if (level == 4) {
                if (qu) y(nq + ba.w + i) = m[1];
                ud[ba.w + i] = 0.0;
            } else {
                ud[ba.w + i] = f;
            }
        }
    }
}

// Events of the step just accepted from (t, Y, F0) to (t1, Y1)
// (AbstractIntegratorRep.cpp:583-756; handlers as TimeStepper::stepTo calls
// them, TimeStepper.cpp:178-205, Integrator.cpp:437-447).  The candidates
// over the whole step decide which events can trigger; the window (tLow,
// tHigh] shrinks around the earliest estimate by trial states interpolated
// inside the step (createInterpolatedState: the cubic Hermite interpolant,
// IntegratorRep.h:299-313, then the LocalOnly projection, quaternions
// normalized when their RMS norm error exceeds the constraint tolerance)
// until it is at most the narrowest required window; the advanced state
// moves back to tHigh by interpolation; every remaining candidate is
// recorded and its handler acts on that state.  A terminating handler that
// changes nothing leaves the state interpolated at tLow (the state
// ReachedEventTrigger reports), one that changes the state leaves it at
// tHigh.  The start-of-step values at EW_E0 then belong to the state the
// next step starts from.  F1, YT, FA, YS and YE are free rows.  Returns 0 (no
// event), 1 (handled), 2 (terminate) or -1 (no candidate in either half of a
// trial, no convergence within 400 trials, or a failed projection of an
// interpolated state).
template <class EV, class TRows>
__device__ int step_events(const RkmArgs& a, const EV& ev, const Wit& W, int lane, int tr, int nq, int nu, int nb,
                           int nquat, const int4* body_a, const double* prm_tree, double t, double& t1,
                           const TRows& Y, const TRows& F0, const TRows& Y1, const TRows& F1, const TRows& YT,
                           const TRows& FA, const TRows& YS, const TRows& YE) {
    const int ny = nq + nu;
    const double consTol = a.cfg.cons_tol, accTs = a.cfg.accuracy * a.cfg.timescale;
    int lo = EW_E0, hi = EW_HIGH, mid = EW_LOW;
    wit_values<EV::kPad>(ev.c, W, t1, Y1, hi);
    const double minWindow = 1.9323057602939027e-14 * fmax(1.0, t1);   // SignificantReal * max(1, t)
    unsigned rising;
    double earliest, narrowest;
    unsigned cand = find_cands(W, ~0u, t, lo, t1, hi, 1.0, minWindow, accTs, rising, earliest, narrowest);
    if (!cand) {
        for (int i = 0; i < W.n; ++i) W.w(EW_E0 + i) = W.w(hi + i);
        return 0;
    }
    // the interpolant at tm of (t, Y, F0) and (te, Y1, F1) into out, projected
    // (realizeAndProjectKinematicsWithThrow, IntegratorRep.h:858-897: with
    // constraints projectQ then projectU at the tolerance, no limit, no
    // error estimate, FA, YS and YE as scratch rows); false when that fails.
    // Under setProjectInterpolatedStates(false) only the state backed up to
    // the event (back_up) is projected (AbstractIntegratorRep.cpp:78-85,
    // 116-126).
    const bool inf = a.cfg.inf_norm, force = a.cfg.force_proj;
    auto interp = [&](double te, double tm, const TRows& out, bool back_up) -> bool {
        const double h = te - t, d = (tm - t) / (te - t);
        const double cy1 = d*d*(3.0 - 2.0*d), cy0 = 1.0 - cy1;
        const double hdd1 = h*d*(d - 1.0), cf1 = hdd1*d, cf0 = cf1 - hdd1;
        for (int k = 0; k < ny; ++k) out(k) = ((cy0*Y(k) + cy1*Y1(k)) + cf0*F0(k)) + cf1*F1(k);
        if constexpr (EV::kMotion) { if (a.motion) prescribe_lane(a, lane, tr, tm, out, true); }
        if (!back_up && !a.cfg.proj_interp) return true;
        if constexpr (EV::kCons) {
            if (ev.K.m > 0) {
                const TRows ou{out.base + (size_t)nq * out.L, out.L, out.lane};
                const TRows su{YS.base + (size_t)nq * YS.L, YS.L, YS.lane};
                const TRows du{FA.base + (size_t)nq * FA.L, FA.L, FA.lane};
                const double nolimit = __builtin_inf();
                bool ch;
                return cons_project_q<EV::kPad>(ev.c, ev.K, out, ou, (const TRows*)nullptr, YS, du, YE, nq, nquat,
                                                consTol, nolimit, proj_opts(a.cfg), &ch) &&
                       cons_project_u<EV::kPad>(ev.c, ev.K, out, ou, (const TRows*)nullptr, su, du, nu, consTol,
                                                nolimit, proj_opts(a.cfg), &ch);
            }
        }
        if (nquat == 0) return true;
        double s2 = 0;
        for (int b = 0; b < nb; ++b) {
            const int4 ba = ldi4(body_a, b);
            if (ba.y != BALL && (ba.y != FREE || prm_tree[(size_t)b * PRM_STRIDE + PRM_EULER] != 0.0)) continue;
            double n2 = 0;
            for (int k = 0; k < 4; ++k) { const double x = out(ba.z + k); n2 += x*x; }
            norm_acc(s2, sqrt(n2) - 1.0, inf);
        }
        if (!force && !(norm_end(s2, nquat, inf) > consTol)) return true;
        for (int b = 0; b < nb; ++b) {
            const int4 ba = ldi4(body_a, b);
            if (ba.y != BALL && (ba.y != FREE || prm_tree[(size_t)b * PRM_STRIDE + PRM_EULER] != 0.0)) continue;
            double n2 = 0;
            for (int k = 0; k < 4; ++k) { const double x = out(ba.z + k); n2 += x*x; }
            const double n = sqrt(n2);
            for (int k = 0; k < 4; ++k) out(ba.z + k) = out(ba.z + k) / n;
        }
        return true;
    };
    double tLow = t, tHigh = t1;
    const double t1o = t1;
    // f(y1): with prescribed motion the known udot at the time of y1
    auto end_derivs = [&](double te) {
        if constexpr (EV::kMotion) { if (a.motion) prescribe_lane(a, lane, tr, te, Y1, true); }
        ev.derivs(Y1, F1);
    };
    if (tHigh - tLow > narrowest) {
        end_derivs(t1o);
        double bias = 1.0;
        int side2 = 0, side1 = 0, it = 0;     // the half chosen two trials ago and at the last one
        do {
            if (++it > 400) return -1;
            if (side2 != 0 && side1 != 0) bias = side2 != side1 ? 1.0 : (side1 < 0 ? bias/2 : bias*2);
            const double tMid = earliest;
            if (!interp(t1o, tMid, YT, false)) return -1;
            wit_values<EV::kPad>(ev.c, W, tMid, YT, mid);
            unsigned r2;
            double e2, n2;
            unsigned nc = find_cands(W, cand, tLow, lo, tMid, mid, bias, minWindow, accTs, r2, e2, n2);
            if (nc) {
                side2 = side1; side1 = -1;
                tHigh = tMid;
                const int x = hi; hi = mid; mid = x;
            } else {
                nc = find_cands(W, cand, tMid, mid, tHigh, hi, bias, minWindow, accTs, r2, e2, n2);
                if (!nc) return -1;
                side2 = side1; side1 = 1;
                tLow = tMid;
                const int x = lo; lo = mid; mid = x;
            }
            cand = nc; rising = r2; earliest = e2; narrowest = n2;
        } while (tHigh - tLow > narrowest);
        if (tHigh < t1o) {
            if (!interp(t1o, tHigh, YT, true)) return -1;
            for (int k = 0; k < ny; ++k) Y1(k) = YT(k);
            t1 = tHigh;
        }
    }
    bool modified = false, term = false;
    int cnt = a.ev_count[lane];
    for (int i = 0; i < W.n; ++i) {
        if (!((cand >> i) & 1u)) continue;
        if (cnt < a.ev_cap) {
            const size_t e = (size_t)lane * a.ev_cap + cnt;
            a.ev_win[2*e] = tLow; a.ev_win[2*e + 1] = tHigh;
            a.ev_ids[2*e] = i; a.ev_ids[2*e + 1] = ((rising >> i) & 1u) ? 2 : 1;
        }
        ++cnt;
        const int act = (W.wa[i].x >> 8) & 15;
        if (act == EVA_SCALE_U) {
            const int ai = W.wa[i].w;
            const double f = W.wp[9*i + 7];
            for (int k = 0; k < nu; ++k)
                if (ai < 0 || ai == k) Y1(nq + k) = f * Y1(nq + k);
            modified = true;
        } else if (act == EVA_TERMINATE) {
            term = true;
        }
    }
    a.ev_count[lane] = cnt;
    if (term && !modified) {
        end_derivs(t1);
        if (!interp(t1, tLow, YT, false)) return -1;
        for (int k = 0; k < ny; ++k) Y1(k) = YT(k);
        t1 = tLow;
    }
    if (term) return 2;
    wit_values<EV::kPad>(ev.c, W, t1, Y1, EW_E0);
    return 1;
}

// UNI: every instance uses tree 0, so the tree tables are read with
// wave-uniform (scalar) loads.  EVT: the tree's event witnesses are watched
// (a batch of one tree on the tree sweeps).
template <class EV, bool UNI, bool EVT = false>
__global__ void __launch_bounds__(64) k_rkm(RkmArgs a) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= a.n_inst) return;
    if (a.status[lane] != 0) return;
    // a tree-server slot run by the level-parallel evaluator
    if constexpr (std::is_same_v<EV, TreePoolEval>) { if (a.pool_grp && a.pool_grp[lane]) return; }
    const int tr = UNI ? 0 : a.tree_of[lane];
    const EV ev(a, lane, tr);
    const int4 ta = ldi4(a.T.tree_a, tr);
    const int nb = ta.x, nq = ta.y, nu = ta.z, ny = nq + nu;
    const int4* body_a = a.T.body_a + ta.w;
    const int nquat = ldi4(a.T.tree_b, tr).y;
    const size_t NV = (size_t)a.ny_max * a.L;
    const int L = a.L;
    using TRows = typename EV::VR;
    auto rows = [&](int v) { return EV::vrow(a, v, lane); };
    const TRows Y = rows(VY), F0 = rows(VF0), FA = rows(VFA), FB = rows(VFB), YS = rows(VYS),
                      Y1 = rows(VY1), YT = rows(VYT), YE = rows(VYE);

    const RkmCfg& cfg = a.cfg;
    const double acc = cfg.accuracy, consTol = cfg.cons_tol;
    double hmin = cfg.min_h, hmax = cfg.max_h;
    if (cfg.fixed_h > 0) { hmin = hmax = cfg.fixed_h; }
    const double projLimit = fmax(2*consTol, sqrt(consTol));
    double t = a.t[lane], hcur = a.h[lane];
    const double tEnd = a.t_end[lane];
    int attempts = a.stats[4*lane], accepted = a.stats[4*lane+1], rejected = a.stats[4*lane+2],
        projections = a.stats[4*lane+3];
    int status = 0, steps = 0;
    Wit W{nullptr, nullptr, 0, 0, nullptr, L};
    if constexpr (EVT) {
        // the witnesses at the state the launch starts from (the values the
        // previous launch left are the same numbers)
        W = make_wit(a.T, tr, a.ev_ws, lane, L);
        if (W.n) wit_values<EV::kPad>(ev.c, W, t, Y, EW_E0);
    }

    // error norm at y1: q part weighted by N NInv for quaternions (IntegratorRep.h:
    // 454-523), u part relative to y0 (calcRelativeScaling, :644-655); RMS,
    // or the largest magnitude under setUseInfinityNorm
    const bool infn = cfg.inf_norm;
    auto errNorm = [&]() -> double {
        double sq = 0;
        for (int b = 0; b < nb; ++b) {
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
        const double qn = norm_end(sq, nq, infn);
        double su = 0;
#pragma unroll 8
        for (int k = 0; k < nu; ++k) {
            const double vi = fabs(Y(nq + k));
            const double us = vi > 1.0 ? 1.0/vi : 1.0;
            norm_acc(su, us * YE(nq + k), infn);
        }
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
                TRows in = Y, out = F0;
                double ts = t;             // the stage's time
                if (st == 1) {
#pragma unroll 8
                    for (int k = 0; k < ny; ++k) YT(k) = Y(k) + (h/3)*F0(k);
                    in = YT; out = FA; ts = t + h/3;
                } else if (st == 2) {
#pragma unroll 8
                    for (int k = 0; k < ny; ++k) YT(k) = Y(k) + (h/6)*(F0(k) + FA(k));
                    in = YT; out = FA; ts = t + h/3;
                } else if (st == 3) {
#pragma unroll 8
                    for (int k = 0; k < ny; ++k) YT(k) = Y(k) + (h/8)*(F0(k) + 3*FA(k));
                    in = YT; out = FB; ts = t + h/2;
                } else if (st == 4) {
#pragma unroll 8
                    for (int k = 0; k < ny; ++k) YS(k) = Y(k) + (h/2)*((F0(k) - 3*FA(k)) + 4*FB(k));
                    in = YS; out = FA; ts = t1;
                }
                if constexpr (EV::kMotion) {
                    if (a.motion) {
                        // the evaluated state gets the prescribed q and u; ysave
                        // keeps its own for the error estimate
                        if (st == 4) {
                            for (int k = 0; k < ny; ++k) YT(k) = YS(k);
                            in = YT;
                        }
                        prescribe_lane(a, lane, tr, ts, in, st != 0);
                    }
                }
                ev.derivs(in, out);
            }
            s0 = 1;
            bool motion = false;
            if constexpr (EV::kMotion) motion = a.motion != nullptr;
            if (motion) {
                // the step's end state gets the prescribed q and u at t1 before
                // the error estimate (setAdvancedStateAndRealizeKinematics)
                for (int k = 0; k < ny; ++k) Y1(k) = Y(k) + (h/6)*((F0(k) + 4*FB(k)) + FA(k));
                if constexpr (EV::kMotion) prescribe_lane(a, lane, tr, t1, Y1, true);
                for (int k = 0; k < ny; ++k) YE(k) = 0.2 * fabs(Y1(k) - YS(k));
            } else {
#pragma unroll 8
                for (int k = 0; k < ny; ++k) {
                    const double y1 = Y(k) + (h/6)*((F0(k) + 4*FB(k)) + FA(k));
                    Y1(k) = y1;
                    YE(k) = 0.2 * fabs(y1 - YS(k));
                }
            }
            bool converged = true;
            const double e1 = errNorm();
            bool cons = false;
            if constexpr (EV::kCons) cons = ev.K.m > 0;
            if (cons) {
                if constexpr (EV::kCons) {
                    // attemptDAEStep (AbstractIntegratorRep.cpp:165-207): projectQ
                    // then projectU of y1 and its error estimate, if the error
                    // is worth it
                    if (!(e1 > 16.0*acc)) {
                        const TRows Y1u{Y1.base + (size_t)nq * L, L, Y1.lane}, YEu{YE.base + (size_t)nq * L, L, YE.lane};
                        const TRows YSu{YS.base + (size_t)nq * L, L, YS.lane}, YTu{YT.base + (size_t)nq * L, L, YT.lane};
                        bool chq = false, chu = false;
                        if (!cons_project_q<EV::kPad>(ev.c, ev.K, Y1, Y1u, &YE, YS, YTu, YT, nq, nquat, consTol,
                                                      projLimit, proj_opts(cfg), &chq) ||
                            !cons_project_u<EV::kPad>(ev.c, ev.K, Y1, Y1u, &YEu, YSu, YTu, nu, consTol, projLimit,
                                                      proj_opts(cfg), &chu))
                            converged = false;
                        if (chq) ++projections;
                    }
                }
            } else if (!(e1 > 16.0*acc) && nquat > 0) {
                double s2 = 0;
                for (int b = 0; b < nb; ++b) {
                    const int4 ba = ldi4(body_a, b);
                    // quaternions only (a Free body flagged PRM_EULER holds angles)
                    if (ba.y != BALL && (ba.y != FREE || a.T.prm[(size_t)(ta.w + b) * PRM_STRIDE + PRM_EULER] != 0.0)) continue;
                    double n2 = 0;
                    for (int k = 0; k < 4; ++k) { const double x = Y1(ba.z + k); n2 += x*x; }
                    norm_acc(s2, sqrt(n2) - 1.0, infn);
                }
                const double qe = norm_end(s2, nquat, infn);
                if (qe > projLimit) converged = false;
                else if (qe > consTol || cfg.force_proj) {
                    ++projections;
                    for (int b = 0; b < nb; ++b) {
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
                }
            }
            const double err = converged ? errNorm() : __builtin_inf();
            // adjustStepSize, AbstractIntegratorRep.cpp:448-502
            double hnew;
            if (!isfinite(err)) hnew = 0.1*hcur;
            else if (err == 0) hnew = 5.0*hcur;
            else hnew = 0.9*hcur*pow(acc/err, 0.25);
            if (hnew > hcur) { if (limited || hnew < 1.2*hcur) hnew = hcur; }
            if (hnew < hcur) { if (err <= acc) hnew = hcur; else hnew = fmin(hnew, 0.9*hcur); }
            hnew = fmin(hnew, 5.0*hcur);
            hnew = fmax(hnew, 0.1*hcur);
            if (hmin != -1) hnew = fmax(hnew, hmin);
            if (hmax != -1) hnew = fmin(hnew, hmax);
            ok = hnew >= hcur;
            hcur = hnew;
            if (!ok) ++rejected;
        } while (!ok);
        if (status < 0) break;
        if constexpr (EVT) {
            if (W.n) {
                const int r = step_events(a, ev, W, lane, tr, nq, nu, nb, nquat, body_a,
                                          a.T.prm + (size_t)ta.w * PRM_STRIDE, t, t1, Y, F0, Y1, FB, YT, FA, YS, YE);
                if (r < 0) { status = -5; break; }
                if (r == 2) {
                    t = t1;
                    for (int k = 0; k < ny; ++k) Y(k) = Y1(k);
                    ++accepted; ++steps;
                    status = ST_TERMINATED;
                    break;
                }
            }
        }
        t = t1;
#pragma unroll 8
        for (int k = 0; k < ny; ++k) Y(k) = Y1(k);
        ++accepted; ++steps;
    }
    if (status == 0 && t >= tEnd) status = 1;
    a.t[lane] = t; a.h[lane] = hcur; a.status[lane] = status;
    a.stats[4*lane] = attempts; a.stats[4*lane+1] = accepted; a.stats[4*lane+2] = rejected; a.stats[4*lane+3] = projections;
}

template <class EV, bool UNI>
__global__ void __launch_bounds__(64) k_tree_f0(RkmArgs a) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= a.n_inst) return;
    const int tr = UNI ? 0 : a.tree_of[lane];
    const EV ev(a, lane, tr);
    if constexpr (EV::kMotion) { if (a.motion) prescribe_lane(a, lane, tr, a.t[lane], EV::vrow(a, VY, lane), true); }
    ev.derivs(EV::vrow(a, VY, lane), EV::vrow(a, VF0, lane));
}
hipError_t launch_rollout_f0(const RkmArgs& a, hipStream_t s) {
    if (a.n_inst <= 0) return hipSuccess;
    const dim3 grd((a.n_inst + 63)/64), blk(64);
    // the derivative is the general form's bit for bit, so identity frames
    // need no separate build here
    if (a.chain) return launch_chain_variant(a, s, true);
    if (a.grp) return launch_tree_group(a, s, true);
    if (a.vec_aos) k_tree_f0<TreeStreamEval, false><<<grd, blk, 0, s>>>(a);
    else if (a.tree_of) k_tree_f0<TreeEval<true>, false><<<grd, blk, 0, s>>>(a);
    else k_tree_f0<TreeEval<false>, true><<<grd, blk, 0, s>>>(a);
    return hipGetLastError();
}

// initialize(): forced quaternion normalization, t and h set.
__device__ void rkm_init_lane(const RkmArgs& a, int lane, double t0) {
    const Lane c = make_lane(a.T, a.tree_of, lane, a.L, a.scratch);
    if (a.motion && !a.vec_aos) {
        // prescribed q and u at t0 (initialize: prescribeQ, prescribeU)
        const chain::RowsT<double, false> y{a.vec + (size_t)VY * a.ny_max * a.L, a.L, (unsigned)lane};
        prescribe_lane(a, lane, a.tree_of ? a.tree_of[lane] : 0, t0, y, true);
    }
    // (a lane with constraints is projected by k_cons_init instead:
    // projectQ normalizes the quaternions as it projects)
    const bool constrained = a.cons_ws && a.T.cons_t && ldi4(a.T.cons_t, a.tree_of ? a.tree_of[lane] : 0).z > 0;
    for (int b = 1; b <= c.nb && !constrained; ++b) {
        const int4 ba = c.ba(b);
        if (ba.y != BALL && ba.y != FREE) continue;
        if (c.prm(b)[PRM_EULER] != 0.0) continue;      // x-y-z angles, not a quaternion
        double qq[4], n2 = 0;
        for (int k = 0; k < 4; ++k) { qq[k] = vec_at(a, VY, lane, ba.z + k); n2 += qq[k]*qq[k]; }
        const double n = sqrt(n2);
        for (int k = 0; k < 4; ++k) vec_at(a, VY, lane, ba.z + k) = qq[k] / n;
    }
    const RkmCfg& cfg = a.cfg;
    double hmin = cfg.min_h, hmax = cfg.max_h, hinit = cfg.init_h;
    if (cfg.fixed_h > 0) hmin = hmax = hinit = cfg.fixed_h;
    double h = hinit != -1 ? hinit : cfg.timescale / 10;
    if (hmin != -1) h = fmax(h, hmin);
    if (hmax != -1) h = fmin(h, hmax);
    a.t[lane] = t0; a.h[lane] = h; a.status[lane] = 0;
    for (int k = 0; k < 4; ++k) a.stats[4*lane + k] = 0;
    // Kahan compensation of the mixed path starts at zero
    for (int k = 0; k < c.nq + c.nu; ++k) vec_at(a, VC0, lane, k) = 0.0;
}
__global__ void __launch_bounds__(64) k_rkm_init(RkmArgs a, double t0) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= a.n_inst) return;
    rkm_init_lane(a, lane, t0);
}

// Serving slots.  Admission record of one request (doubles): nb*PRM_STRIDE
// model constants and 3 gravity, then q (nq), u (nu), t_end, and (nb+1)*6
// applied body forces.  One 64-thread block per admitted request scatters
// its record into the slot's structure-of-arrays rows, then initializes the
// slot's integrator state.
__global__ void __launch_bounds__(64) k_srv_admit(RkmArgs a, const double* rec, const int* slots, int n,
                                                  int rec_stride, double t0) {
    const int i = blockIdx.x;
    if (i >= n) return;
    const int slot = slots[i];
    const int4 ta = a.T.tree_a[0];
    const int nb = ta.x, nq = ta.y, nu = ta.z;
    const double* r = rec + (size_t)i * rec_stride;
    const int np = nb * PRM_STRIDE + 3;
    const size_t L = a.L;
    double* prm_lane = const_cast<double*>(a.prm_lane);
    for (int k = threadIdx.x; k < np; k += blockDim.x) prm_lane[(size_t)k * L + slot] = r[k];
    for (int k = threadIdx.x; k < nq + nu; k += blockDim.x) a.vec[(size_t)k * L + slot] = r[np + k];   // VY
    double* fa = const_cast<double*>(a.fapp) + 6 * (size_t)a.off.b_off[slot];
    for (int k = threadIdx.x; k < 6 * (nb + 1); k += blockDim.x) fa[k] = r[np + nq + nu + 1 + k];
    __syncthreads();
    if (threadIdx.x == 0) {
        const_cast<double*>(a.t_end)[slot] = r[np + nq + nu];
        rkm_init_lane(a, slot, t0);
    }
}
// Copies the finished slots' state out (q then u, t) and their status and
// statistics, and marks the slots empty (status 2).
__global__ void k_srv_collect(RkmArgs a, const int* slots, int n, double* out, int out_stride, int* outi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int slot = slots[i];
    const int4 ta = a.T.tree_a[0];
    const int ny = ta.y + ta.z;
    const size_t L = a.L;
    double* o = out + (size_t)i * out_stride;
    for (int k = 0; k < ny; ++k) o[k] = a.vec[(size_t)k * L + slot];
    o[ny] = a.t[slot];
    outi[5*i] = a.status[slot];
    for (int k = 0; k < 4; ++k) outi[5*i + 1 + k] = a.stats[4*slot + k];
    a.status[slot] = 2;
}
// Tree-server admission: one 64-thread block per request.  ints: nb, nq, nu,
// body_off, b_off, nquat, rec, grp (1: the level-parallel evaluator), lvl
// offset, prm_bfs offset, level count, 0, then body_a (4 nb), body_b (4 nb;
// child begin already offset into the pool), child_idx (nb), and for grp
// bfs (4 nb), bfs_b (4 nb) and the level list (level count + 3).  doubles:
// prm (nb PRM_STRIDE), gravity (3), q (nq), u (nu), t_end, applied forces
// (6 (nb+1)), the pool offsets of the slot's vectors and records (2); for
// grp the breadth-first constants are built here from prm and bfs.
__global__ void __launch_bounds__(64) k_tsrv_admit(RkmArgs a, const int* slots, int n, const long long* ioff,
                                                   const int* ints, const long long* doff, const double* dbl) {
    const int i = blockIdx.x;
    if (i >= n) return;
    const int slot = slots[i];
    const int* I = ints + ioff[i];
    const double* D = dbl + doff[i];
    const int nb = I[0], nq = I[1], nu = I[2], body_off = I[3], b_off = I[4];
    int4* body_a = const_cast<int4*>(a.T.body_a) + body_off;
    int4* body_b = const_cast<int4*>(a.T.body_b) + body_off;
    int* child = const_cast<int*>(a.T.child_idx) + body_off;
    double* prm = const_cast<double*>(a.T.prm) + (size_t)body_off * PRM_STRIDE;
    const int grp = I[7], lvl_off = I[8], pbfs_off = I[9], nlev = I[10];
    const int* ba = I + 12;
    const int* bb = ba + 4*nb;
    const int* ci = bb + 4*nb;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        body_a[k] = make_int4(ba[4*k], ba[4*k+1], ba[4*k+2], ba[4*k+3]);
        body_b[k] = make_int4(bb[4*k], bb[4*k+1], bb[4*k+2], bb[4*k+3]);
        child[k] = ci[k];
    }
    if (grp) {
        const int* bf = ci + nb;
        const int* bfb = bf + 4*nb;
        const int* lv = bfb + 4*nb;
        int4* bfs = const_cast<int4*>(a.T.bfs) + body_off;
        int4* bfs_b = const_cast<int4*>(a.T.bfs_b) + body_off;
        for (int k = threadIdx.x; k < nb; k += blockDim.x) {
            bfs[k] = make_int4(bf[4*k], bf[4*k+1], bf[4*k+2], bf[4*k+3]);
            bfs_b[k] = make_int4(bfb[4*k], bfb[4*k+1], bfb[4*k+2], bfb[4*k+3]);
        }
        int* lvl = const_cast<int*>(a.T.lvl) + lvl_off;
        for (int k = threadIdx.x; k < nlev + 3; k += blockDim.x) lvl[k] = lv[k];
    }
    for (int k = threadIdx.x; k < nb * PRM_STRIDE; k += blockDim.x) prm[k] = D[k];
    const double* g = D + nb * PRM_STRIDE;
    const double* q = g + 3;
    const double* u = q + nq;
    const double te = u[nu];
    const double* fa = u + nu + 1;
    const double* po = fa + 6 * (nb + 1);
    const long long vbase = (long long)po[0], rbase = (long long)po[1];
    if (grp) {
        // the constants in breadth-first order, blocks of BW positions stored
        // constant-major (padding positions zero), from the staged ones
        const int BW = tgroup::BW, nd = tgroup::blocks(nb) * BW * PRM_STRIDE;
        const int* bf = ci + nb;
        double* pb = const_cast<double*>(a.T.prm_bfs) + pbfs_off;
        for (int e = threadIdx.x; e < nd; e += blockDim.x) {
            const int blk = e / (PRM_STRIDE * BW), r = e % (PRM_STRIDE * BW), c = r / BW, k = blk * BW + r % BW;
            pb[e] = k < nb ? D[(size_t)(bf[4*k] - 1) * PRM_STRIDE + c] : 0.0;
        }
    }
    double* y = a.vec + vbase;
    for (int k = threadIdx.x; k < nq; k += blockDim.x) y[k] = q[k];
    for (int k = threadIdx.x; k < nu; k += blockDim.x) y[nq + k] = u[k];
    double* fapp = const_cast<double*>(a.fapp) + 6 * (size_t)b_off;
    for (int k = threadIdx.x; k < 6 * (nb + 1); k += blockDim.x) fapp[k] = fa[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        const_cast<int4*>(a.T.tree_a)[slot] = make_int4(nb, nq, nu, body_off);
        const_cast<int4*>(a.T.tree_b)[slot] = make_int4(I[6], I[5], 0, grp ? lvl_off : 0);
        const_cast<unsigned char*>(a.pool_grp)[slot] = (unsigned char)(grp ? 1 : 0);
        double* gr = const_cast<double*>(a.T.gravity) + 4 * slot;
        gr[0] = g[0]; gr[1] = g[1]; gr[2] = g[2]; gr[3] = 0.0;
        const_cast<int*>(a.off.b_off)[slot] = b_off;
        long long* pool = const_cast<long long*>(a.pool);
        pool[2*slot] = vbase; pool[2*slot + 1] = rbase;
        // initialize (as rkm_init_lane): quaternions normalized, t, h, stats
        for (int b = 0; b < nb; ++b) {
            const int4 x = body_a[b];
            if (x.y != BALL && x.y != FREE) continue;
            if (prm[(size_t)b * PRM_STRIDE + PRM_EULER] != 0.0) continue;
            double qq[4], n2 = 0;
            for (int k = 0; k < 4; ++k) { qq[k] = y[x.z + k]; n2 += qq[k]*qq[k]; }
            const double nn = sqrt(n2);
            for (int k = 0; k < 4; ++k) y[x.z + k] = qq[k] / nn;
        }
        const RkmCfg& cfg = a.cfg;
        double hmin = cfg.min_h, hmax = cfg.max_h, hinit = cfg.init_h;
        if (cfg.fixed_h > 0) hmin = hmax = hinit = cfg.fixed_h;
        double h = hinit != -1 ? hinit : cfg.timescale / 10;
        if (hmin != -1) h = fmax(h, hmin);
        if (hmax != -1) h = fmin(h, hmax);
        const_cast<double*>(a.t_end)[slot] = te;
        a.t[slot] = 0.0; a.h[slot] = h;
        for (int k = 0; k < 4; ++k) a.stats[4*slot + k] = 0;
        a.status[slot] = 0;
    }
}
__global__ void k_tsrv_collect(RkmArgs a, const int* slots, int n, const long long* qoff, double* out,
                               double* t_out, int* outi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int slot = slots[i];
    const int4 ta = a.T.tree_a[slot];
    const int ny = ta.y + ta.z;
    const double* y = a.vec + a.pool[2*slot];
    double* o = out + qoff[i];
    for (int k = 0; k < ny; ++k) o[k] = y[k];
    t_out[i] = a.t[slot];
    outi[5*i] = a.status[slot];
    for (int k = 0; k < 4; ++k) outi[5*i + 1 + k] = a.stats[4*slot + k];
    a.status[slot] = 2;
}
hipError_t launch_tsrv_admit(const RkmArgs& a, const int* slots, int n, const long long* ioff, const int* ints,
                             const long long* doff, const double* dbl, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_tsrv_admit<<<dim3(n), dim3(64), 0, s>>>(a, slots, n, ioff, ints, doff, dbl);
    return hipGetLastError();
}
hipError_t launch_tsrv_collect(const RkmArgs& a, const int* slots, int n, const long long* qoff, double* out,
                               double* t_out, int* outi, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_tsrv_collect<<<dim3((n + 63)/64), dim3(64), 0, s>>>(a, slots, n, qoff, out, t_out, outi);
    return hipGetLastError();
}
hipError_t launch_srv_admit(const RkmArgs& a, const double* rec, const int* slots, int n, int rec_stride,
                            double t0, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_srv_admit<<<dim3(n), dim3(64), 0, s>>>(a, rec, slots, n, rec_stride, t0);
    return hipGetLastError();
}
hipError_t launch_srv_collect(const RkmArgs& a, const int* slots, int n, double* out, int out_stride, int* outi,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_srv_collect<<<dim3((n + 63)/64), dim3(64), 0, s>>>(a, slots, n, out, out_stride, outi);
    return hipGetLastError();
}

hipError_t launch_rkm_init(const RkmArgs& a, double t0, hipStream_t s) {
    if (a.n_inst <= 0) return hipSuccess;
    k_rkm_init<<<dim3((a.n_inst + 63)/64), dim3(64), 0, s>>>(a, t0);
    return hipGetLastError();
}
// initialize() of a lane with constraints (Integrator.cpp:376-377): projectQ
// and projectU with ForceProjection and ForceFullNewton (no LocalOnly), with
// VYS and VYT as scratch rows; status -4 when either fails.
template <bool PAD>
__global__ void __launch_bounds__(64) k_cons_init(RkmArgs a) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= a.n_inst) return;
    const int tr = PAD ? a.tree_of[lane] : 0;
    const Cons K = make_cons(a.T, tr, a.cons_ws, lane, a.L);
    if (K.m == 0) return;
    const Lane c = make_lane_t(a.T, tr, lane, a.L, a.scratch, PAD ? a.rec_stride : 0);
    using TRows = typename TreeEval<PAD>::VR;
    const int4 ta = ldi4(a.T.tree_a, tr);
    const int nq = ta.y, nu = ta.z, nquat = ldi4(a.T.tree_b, tr).y;
    const TRows Y = TreeEval<PAD>::vrow(a, VY, lane), YS = TreeEval<PAD>::vrow(a, VYS, lane),
                YT = TreeEval<PAD>::vrow(a, VYT, lane);
    const TRows Yu{Y.base + (size_t)nq * a.L, a.L, Y.lane}, YSu{YS.base + (size_t)nq * a.L, a.L, YS.lane},
                YTu{YT.base + (size_t)nq * a.L, a.L, YT.lane};
    const double inf = __builtin_inf();
    bool ch = false;
    if (!cons_project_q<PAD>(c, K, Y, Yu, (const TRows*)nullptr, YS, YTu, YT, nq, nquat, a.cfg.cons_tol, inf,
                             PJ_INIT | proj_opts(a.cfg),
                             &ch) ||
        !cons_project_u<PAD>(c, K, Y, Yu, (const TRows*)nullptr, YSu, YTu, nu, a.cfg.cons_tol, inf,
                             PJ_INIT | proj_opts(a.cfg), &ch))
        a.status[lane] = -4;
}
hipError_t launch_cons_init(const RkmArgs& a, hipStream_t s) {
    if (a.n_inst <= 0 || !a.cons_ws) return hipSuccess;
    const dim3 grd((a.n_inst + 63)/64), blk(64);
    if (a.tree_of) k_cons_init<true><<<grd, blk, 0, s>>>(a);
    else k_cons_init<false><<<grd, blk, 0, s>>>(a);
    return hipGetLastError();
}
hipError_t launch_rkm(const RkmArgs& a, hipStream_t s) {
    if (a.n_inst <= 0) return hipSuccess;
    const dim3 grd((a.n_inst + 63)/64), blk(64);
    if (a.pool) { k_rkm<TreePoolEval, false><<<grd, blk, 0, s>>>(a); return hipGetLastError(); }
    if (a.chain) return launch_chain_variant(a, s, false);
    if (a.grp) return launch_tree_group(a, s, false);
    if (a.ev_ws) {
        // event witnesses: the tree sweeps
        if (a.vec_aos) return hipErrorInvalidValue;
        if (a.tree_of) k_rkm<TreeEval<true>, false, true><<<grd, blk, 0, s>>>(a);
        else k_rkm<TreeEval<false>, true, true><<<grd, blk, 0, s>>>(a);
        return hipGetLastError();
    }
    if (a.vec_aos) k_rkm<TreeStreamEval, false><<<grd, blk, 0, s>>>(a);
    else if (a.tree_of) k_rkm<TreeEval<true>, false><<<grd, blk, 0, s>>>(a);
    else k_rkm<TreeEval<false>, true><<<grd, blk, 0, s>>>(a);
    return hipGetLastError();
}

__global__ void k_lane_gather(const int* inst, int n, const double* src, double* dst) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) dst[l] = src[inst[l]];
}
__global__ void k_lane_scatter(const int* inst, int n, const double* tl, double* ti, const int* sl, int* si,
                               const int* xl, int* xi) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const int i = inst[l];
    if (ti) ti[i] = tl[l];
    if (si) si[i] = sl[l];
    if (xi) for (int k = 0; k < 4; ++k) xi[4*i + k] = xl[4*l + k];
}
// New final times (te, lane order, or te_all for every lane): a lane that had
// reached its old final time resumes if the new one lies ahead of it.
// The event log of each lane to its instance's rows (lanes permuted).
__global__ void k_event_scatter(const int* inst, int n, int cap, const double* win, const int* ids, const int* cnt,
                                double* win_out, int* ids_out, int* cnt_out) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const size_t i = inst[l];
    const size_t so = (size_t)l * cap * 2, d = i * cap * 2;
    if (win_out) for (int k = 0; k < 2*cap; ++k) win_out[d + k] = win[so + k];
    if (ids_out) for (int k = 0; k < 2*cap; ++k) ids_out[d + k] = ids[so + k];
    if (cnt_out) cnt_out[i] = cnt[l];
}
hipError_t launch_event_scatter(const int* inst, int n, int cap, const double* win, const int* ids, const int* cnt,
                                double* win_out, int* ids_out, int* cnt_out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_event_scatter<<<dim3((n + 63)/64), dim3(64), 0, s>>>(inst, n, cap, win, ids, cnt, win_out, ids_out, cnt_out);
    return hipGetLastError();
}
__global__ void k_rollout_extend(int n, const double* te, double te_all, const double* t, double* t_end,
                                 int* status) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const double e = te ? te[l] : te_all;
    t_end[l] = e;
    if (status[l] == 1 && t[l] < e) status[l] = 0;
}
hipError_t launch_rollout_extend(int n, const double* te_lane, double te_all, const double* t, double* t_end,
                                 int* status, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_rollout_extend<<<dim3((n + 255)/256), dim3(256), 0, s>>>(n, te_lane, te_all, t, t_end, status);
    return hipGetLastError();
}
// Reports by interpolation after one accepted step per lane (Simbody's
// default reporting, AbstractIntegratorRep.cpp:63-93, 310-330): every report
// time r in (t_prev, t_now] of the lane's instance gets the step's end state
// if r == t_now, else the cubic Hermite interpolant (interpolateOrder3,
// IntegratorRep.h:285-313) of the step's two ends, then projectQ of the
// interpolated state: its quaternions normalized when their RMS norm error
// exceeds cons_tol (SimbodyMatterSubsystemRep.cpp:4151-4160); popt: PJ_INF
// (the infinity norm), PJ_FORCE (always normalized), PJ_NONE (interpolated
// states not projected, setProjectInterpolatedStates(false)).  Packed
// instance-order states and derivatives; next_k: the lane's next report.
__global__ void k_report_interp(DevTables T, const int* tree_of, const int* q_off, const int* u_off, int L,
                                const double* rt, int nrep, int* next_k, const double* t_prev, const double* t_now,
                                const double* qp, const double* up, const double* fqp, const double* fup,
                                const double* qn, const double* un, const double* fqn, const double* fun,
                                double cons_tol, int popt, double* q_out, double* u_out, long total_q,
                                long total_u) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    const bool inf = popt & PJ_INF;
    const double t0 = t_prev[l], t1 = t_now[l];
    int k = next_k[l];
    if (!(t1 > t0) || k >= nrep || rt[k] > t1) return;
    const int tr = tree_of ? tree_of[l] : 0;
    const int4 ta = ldi4(T.tree_a, tr);
    const int nb = ta.x, nq = ta.y, nu = ta.z, nquat = ldi4(T.tree_b, tr).y;
    const int qo = q_off[l], uo = u_off[l];
    for (; k < nrep && rt[k] <= t1; ++k) {
        double* oq = q_out + (size_t)k * total_q + qo;
        double* ou = u_out + (size_t)k * total_u + uo;
        if (rt[k] == t1) {
            for (int j = 0; j < nq; ++j) oq[j] = qn[qo + j];
            for (int j = 0; j < nu; ++j) ou[j] = un[uo + j];
            continue;
        }
        const double h = t1 - t0, d = (rt[k] - t0) / h;
        const double cy1 = d*d*(3.0 - 2.0*d), cy0 = 1.0 - cy1;
        const double hdd1 = h*d*(d - 1.0), cf1 = hdd1*d, cf0 = cf1 - hdd1;
        for (int j = 0; j < nq; ++j)
            oq[j] = ((cy0*qp[qo + j] + cy1*qn[qo + j]) + cf0*fqp[qo + j]) + cf1*fqn[qo + j];
        for (int j = 0; j < nu; ++j)
            ou[j] = ((cy0*up[uo + j] + cy1*un[uo + j]) + cf0*fup[uo + j]) + cf1*fun[uo + j];
        if (nquat > 0 && !(popt & PJ_NONE)) {
            auto quat = [&](int b) {
                const int4 ba = ldi4(T.body_a, ta.w + b);
                return (ba.y == BALL || ba.y == FREE) && T.prm[(size_t)(ta.w + b) * PRM_STRIDE + PRM_EULER] == 0.0;
            };
            double s2 = 0;
            for (int b = 0; b < nb; ++b) {
                if (!quat(b)) continue;
                const double* qq = oq + ldi4(T.body_a, ta.w + b).z;
                norm_acc(s2, sqrt(qq[0]*qq[0] + qq[1]*qq[1] + qq[2]*qq[2] + qq[3]*qq[3]) - 1.0, inf);
            }
            if ((popt & PJ_FORCE) || norm_end(s2, nquat, inf) > cons_tol) {
                for (int b = 0; b < nb; ++b) {
                    if (!quat(b)) continue;
                    double* qq = oq + ldi4(T.body_a, ta.w + b).z;
                    const double n = sqrt(qq[0]*qq[0] + qq[1]*qq[1] + qq[2]*qq[2] + qq[3]*qq[3]);
                    for (int c = 0; c < 4; ++c) qq[c] = qq[c] / n;
                }
            }
        }
    }
    next_k[l] = k;
}
hipError_t launch_report_interp(const DevTables& T, const int* tree_of, const int* q_off, const int* u_off, int L,
                                const double* rt, int nrep, int* next_k, const double* t_prev, const double* t_now,
                                const double* const prev[4], const double* const now[4], double cons_tol,
                                int popt, double* q_out, double* u_out, long total_q, long total_u, hipStream_t s) {
    if (L <= 0) return hipSuccess;
    k_report_interp<<<dim3((L + 127)/128), dim3(128), 0, s>>>(T, tree_of, q_off, u_off, L, rt, nrep, next_k, t_prev,
        t_now, prev[0], prev[1], prev[2], prev[3], now[0], now[1], now[2], now[3], cons_tol, popt, q_out, u_out,
        total_q, total_u);
    return hipGetLastError();
}
hipError_t launch_lane_gather(const int* inst, int n, const double* src_inst, double* dst_lane, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_lane_gather<<<dim3((n + 255)/256), dim3(256), 0, s>>>(inst, n, src_inst, dst_lane);
    return hipGetLastError();
}
hipError_t launch_lane_scatter(const int* inst, int n, const double* t_lane, double* t_inst, const int* st_lane,
                               int* st_inst, const int* stats_lane, int* stats_inst, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_lane_scatter<<<dim3((n + 255)/256), dim3(256), 0, s>>>(inst, n, t_lane, t_inst, st_lane, st_inst,
                                                              stats_lane, stats_inst);
    return hipGetLastError();
}

// Packed boundary layout <-> SoA state vector VY (or, out only, vector v).
__global__ void k_pack(RkmArgs a, const double* q, const double* u, int to_soa, double* q_out, double* u_out, int v) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= a.n_inst) return;
    const int tr = a.tree_of ? a.tree_of[lane] : 0;
    const int4 ta = a.T.tree_a[tr];
    const int nq = ta.y, nu = ta.z;
    const int qo = a.off.q_off[lane], uo = a.off.u_off[lane];
    if (to_soa) {
        for (int k = 0; k < nq; ++k) vec_at(a, v, lane, k) = q[qo + k];
        for (int k = 0; k < nu; ++k) vec_at(a, v, lane, nq + k) = u[uo + k];
    } else {
        for (int k = 0; k < nq; ++k) q_out[qo + k] = vec_at(a, v, lane, k);
        for (int k = 0; k < nu; ++k) u_out[uo + k] = vec_at(a, v, lane, nq + k);
    }
}
hipError_t launch_pack_state(const RkmArgs& a, const double* q, const double* u, int to_soa,
                             double* q_out, double* u_out, hipStream_t s, int vec) {
    if (a.n_inst <= 0) return hipSuccess;
    k_pack<<<dim3((a.n_inst + 255)/256), dim3(256), 0, s>>>(a, q, u, to_soa, q_out, u_out, vec);
    return hipGetLastError();
}

} // namespace tdyn
