// Serial-chain evaluator of the state derivative (realize(Acceleration)) for
// batches whose instances all share one chain topology (parent(b) = b-1).
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
//          Only G and DI*eps (7 dof values) are written per body.
//   pass 3 (base->tip): kinematics recomputed forward, udot = DI eps - G^T A+,
//          A_GB = A+ + H udot + a.
//
// Precision.  K is the scalar of the kinematic recursion (X_GB, V_GB, H,
// HDot terms), the bias forces (coriolis a, gyroscopic b, gravity, z, z+)
// and the articulated inertias (P, P+, G).  K = double is the FP64 path.
// K = float is the mixed path; in it D = H^T P H is formed with FP64
// products and sums and inverted in FP64, G = P H D^-1 is formed in FP64
// and rounded once, DI*eps and udot are FP64, quaternion normalization is
// FP64, and the state and its Runge-Kutta combinations stay FP64 (the step
// update y1 = y0 + dy is Kahan-compensated across steps).
#pragma once
#include <type_traits>
#include "tdyn_internal.hpp"
#include "tdyn_math.hpp"

namespace tdyn {
namespace chain {

using d3 = v3<double>;
using dm = m33<double>;
using ds = s33<double>;
using dv = sv<double>;
using da = abi<double>;

// TYPE = GEN: a padded generic body (trees of different shapes in one wave):
// any non-Free mobilizer carried as 3 dof, the type chosen at run time; its
// unused H columns, u entries and D rows are zero and D is padded with ones
// on the diagonal, so every padded term adds an exact zero.
template <int TYPE> struct MobC {
    static constexpr int D = TYPE <= SLIDER ? 1 : TYPE == UNIVERSAL ? 2 : (TYPE == BALL || TYPE == GEN) ? 3 : 6;
    static constexpr int NQ = TYPE <= SLIDER ? 1 : TYPE == UNIVERSAL ? 2 : (TYPE == BALL || TYPE == GEN) ? 4 : 7;
};

#define TD_GLOBAL __attribute__((address_space(1)))
template <class T> __device__ __forceinline__ TD_GLOBAL T* uniform_ptr(T* p);

// Conversions between the two scalars (identity when K = double).
template <class K> __device__ __forceinline__ v3<K> tok(const d3& a) { return {K(a.x), K(a.y), K(a.z)}; }
template <class K> __device__ __forceinline__ sv<K> tok(const dv& a) { return {tok<K>(a.w), tok<K>(a.v)}; }
template <class K> __device__ __forceinline__ d3 tod(const v3<K>& a) { return {double(a.x), double(a.y), double(a.z)}; }
template <class K> __device__ __forceinline__ dv tod(const sv<K>& a) { return {tod(a.w), tod(a.v)}; }
// FP64 dot product of two K spatial vectors (the plain dot when K = double).
template <class K> __device__ __forceinline__ double dotd(const sv<K>& a, const sv<K>& b) {
    if constexpr (std::is_same<K, double>::value) return dot(a, b);
    else return dot(tod(a), tod(b));
}

// Model constants of the chain, PRM_STRIDE doubles per body.  Two access
// forms: one model for the whole batch (a constant-address-space pointer, so
// loads are wave-uniform scalar loads), or one model per lane (serving slots:
// structure-of-arrays rows, row k of this lane at p[k * L]).
struct PrmL {
    const double* p; int L; unsigned lane;     // p: row 0 (the same in every lane)
    __device__ __forceinline__ double operator[](int k) const {
        return *(const TD_GLOBAL double*)((const TD_GLOBAL char*)uniform_ptr(p + (size_t)k * L) + lane * 8u);
    }
    __device__ __forceinline__ PrmL operator+(size_t k) const { return {p + k * (size_t)L, L, lane}; }
};
// A model whose frames are all identity rotations at the joints (R_PF = I,
// p_PF = 0, R_MB = I; X_BM may translate), Simbody's noR_PF/noX_MB node
// specializations: the products with those frames are skipped.  They are
// exact in floating point (x*1 + y*0 + z*0 = x), so results are bitwise
// those of the general form.
template <class PA> struct IdF {
    PA p;
    __device__ __forceinline__ IdF() = default;
    __device__ __forceinline__ explicit IdF(PA q) : p(q) {}
    __device__ __forceinline__ double operator[](int k) const { return p[k]; }
    __device__ __forceinline__ IdF operator+(size_t k) const { return IdF(p + k); }
};
// A model that may hold reversed mobilizers (MobilizedBody::Reverse) or use
// Euler angles: each body's PRM_REV entry says whether its X_FM and H_FM are
// reversed (generic bodies), PRM_EULER whether its Ball or Free q are
// body-fixed x-y-z angles.
template <class PA> struct ExtF {
    PA p;
    __device__ __forceinline__ ExtF() = default;
    __device__ __forceinline__ explicit ExtF(PA q) : p(q) {}
    __device__ __forceinline__ double operator[](int k) const { return p[k]; }
    __device__ __forceinline__ ExtF operator+(size_t k) const { return ExtF(p + k); }
};
template <class PA> struct PrmTraits { static constexpr bool lane = false, idf = false, ext = false; };
template <> struct PrmTraits<PrmL> { static constexpr bool lane = true, idf = false, ext = false; };
template <class T> struct PrmTraits<IdF<T>> { static constexpr bool lane = PrmTraits<T>::lane, idf = true, ext = false; };
template <class T> struct PrmTraits<ExtF<T>> { static constexpr bool lane = PrmTraits<T>::lane, idf = false, ext = true; };

template <class K, class PA> __device__ __forceinline__ m33<K> ldm(const PA& p) {
    m33<K> r; for (int i = 0; i < 9; ++i) r.a[i] = K(p[i]); return r;
}
template <class K, class PA> __device__ __forceinline__ v3<K> ld3(const PA& p) { return {K(p[0]), K(p[1]), K(p[2])}; }
template <class K> __device__ __forceinline__ m33<K> mulABt(const m33<K>& A, const m33<K>& B) {   // A B^T
    m33<K> C;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) C.a[r*3+c] = A.a[r*3]*B.a[c*3] + A.a[r*3+1]*B.a[c*3+1] + A.a[r*3+2]*B.a[c*3+2];
    return C;
}

// Angle sin/cos: FP64 for the FP64 path; for K = float the FP64 angle is
// first reduced to [-pi, pi] in FP64 and the FP32 functions are applied.
template <class K> __device__ __forceinline__ void sincosk(double q, K* s, K* c) {
    if constexpr (std::is_same<K, double>::value) sincos(q, s, c);
    else {
        const double r = q - 6.283185307179586 * rint(q * 0.15915494309189535);
        sincosf(float(r), s, c);
    }
}

// Mobilizer part: X_FM(q) and H_FM.
template <int TYPE, class K> struct Joint {
    static constexpr int D = MobC<TYPE>::D;
    m33<K> R_FM; v3<K> p_FM; sv<K> H_FM[D];
    int ty = TYPE, d = D;       // run-time type and dof of a GEN body
    bool rv = false;            // reversed (GEN only); Hm: the mobilizer's own angular H columns
    v3<K> Hm[3];
    bool bu = false;            // a Bushing (FREE, extended builds): u = qdot, Gimbal angle columns
    // Reversed mobilizer (RigidBodyNodeSpec.h:262-277): what set() computed is
    // the mobilizer's X_MF and H_MF; X_FM = ~X_MF and, by
    // calcReverseMobilizerH_FM (RigidBodyNodeSpec.cpp:131-166),
    //   H_FM_w = -R_FM H_MF_w,  H_FM_v = H_FM_w x p_FM - R_FM H_MF_v.
    __device__ __forceinline__ void reverse() {
        static_assert(TYPE == GEN, "reversal is carried by generic bodies");
        rv = true;
        const m33<K> R_MF = R_FM;
        p_FM = -mulT(R_MF, p_FM);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) R_FM.a[3*r + c] = R_MF.a[3*c + r];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Hm[j] = H_FM[j].w;
            const v3<K> w = -mulT(R_MF, Hm[j]);
            H_FM[j] = {w, cross(w, p_FM) - mulT(R_MF, H_FM[j].v)};
        }
    }
    // tc: cached trig / normalization values of this q (written by the first
    // sweep with CACHED = false): Pin, Planar, Cylinder, Screw (cos,sin of
    // q0), Universal and Gimbal (cos,sin of q0 then of q1; a Gimbal's third
    // angle is recomputed, which keeps the cache at four entries),
    // Ball/Free 1/|quaternion| (computed in FP64).  pitch: the Screw pitch
    // (GEN only).
    // eul: Ball and Free hold body-fixed x-y-z angles instead of a quaternion
    // (Euler-angle mode, RigidBodyNodeSpec_Ball.h:113-158, _Free.h:139-188);
    // ec keeps (c0, s0, c1, s1) of those angles for their rates.
    K ec[4];
    // EXT: the build for extended models (BendStretch bodies possible).
    // bush: a Bushing carried as a Free body (RigidBodyNodeSpec_Bushing.h:
    // 152-191): the Euler-mode Free's X_FM, the Gimbal's angle columns.
    template <bool CACHED, bool EXT = false>
    __device__ __forceinline__ void set(const double* qv, K* tc, int rty = TYPE, K pitch = K(0), bool eul = false,
                                        bool bush = false) {
        const K z = K(0), o = K(1);
        p_FM = zero3<K>();
        if constexpr (TYPE == FREE) {
            if (eul) {
                // the root only: the three sin/cos are recomputed in every sweep
                K c0, s0, c1, s1, c2, s2;
                sincosk<K>(qv[0], &s0, &c0); sincosk<K>(qv[1], &s1, &c1); sincosk<K>(qv[2], &s2, &c2);
                ec[0] = c0; ec[1] = s0; ec[2] = c1; ec[3] = s1;
                R_FM = rot_xyz(c0, s0, c1, s1, c2, s2);
                p_FM = V3(K(qv[3]), K(qv[4]), K(qv[5]));
                H_FM[0] = {V3(o, z, z), zero3<K>()};
                H_FM[1] = {V3(z, o, z), zero3<K>()};
                H_FM[2] = {V3(z, z, o), zero3<K>()};
                H_FM[3] = {zero3<K>(), V3(o, z, z)};
                H_FM[4] = {zero3<K>(), V3(z, o, z)};
                H_FM[5] = {zero3<K>(), V3(z, z, o)};
                bu = bush;
                if (bush) {
                    H_FM[1] = {V3(z, c0, s0), zero3<K>()};
                    H_FM[2] = {V3(s1, -s0*c1, c0*c1), zero3<K>()};
                }
                return;
            }
        }
        if constexpr (TYPE == GEN) {
            ty = rty; d = mob_dof(rty);
            H_FM[0] = H_FM[1] = H_FM[2] = svzero<K>();
            // Joint angles: q0 about z (Pin, Planar, Cylinder, Screw), body-fixed
            // x, y (Universal) and x, y, z (Gimbal).  Their sin/cos come from one
            // sincos in a loop shared by every type; when CACHED the first two
            // come from the trig cache.
            const bool xyz = rty == UNIVERSAL || rty == GIMBAL || (rty == BALL && eul);
            const int na = rty == UNIVERSAL ? 2 : xyz ? 3
                         : (rty == SLIDER || rty == TRANSLATION || rty == BALL || rty == WELD) ? 0 : 1;
            K c0 = o, s0 = z, c1 = o, s1 = z, c2 = o, s2 = z;
            int k = 0;
            if constexpr (CACHED) { c0 = tc[0]; s0 = tc[1]; c1 = tc[2]; s1 = tc[3]; k = 2; }
            double qa = CACHED ? qv[2] : qv[0], qb = qv[1];      // angle k, angle k+1
#pragma unroll 1
            for (; k < na; ++k) {
                K sn, cs;
                sincosk<K>(qa, &sn, &cs);
                if (k == 0) { c0 = cs; s0 = sn; } else if (k == 1) { c1 = cs; s1 = sn; } else { c2 = cs; s2 = sn; }
                qa = qb; qb = qv[2];
            }
            if constexpr (!CACHED) { tc[0] = c0; tc[1] = s0; tc[2] = c1; tc[3] = s1; }
            ec[0] = c0; ec[1] = s0; ec[2] = c1; ec[3] = s1;
            if (rty == SLIDER) {
                R_FM = ident<K>(); p_FM = V3(K(qv[0]), z, z);
                H_FM[0] = {zero3<K>(), V3(o, z, z)};
            } else if (rty == WELD) {
                // RigidBodyNode_Weld.cpp:389-423: X_FM = identity, no H
                R_FM = ident<K>();
            } else if (rty == TRANSLATION) {
                // RigidBodyNodeSpec_Translation.h: X_FM = (I, q), H_FM = (0, e_k)
                R_FM = ident<K>(); p_FM = V3(K(qv[0]), K(qv[1]), K(qv[2]));
                H_FM[0] = {zero3<K>(), V3(o, z, z)};
                H_FM[1] = {zero3<K>(), V3(z, o, z)};
                H_FM[2] = {zero3<K>(), V3(z, z, o)};
            } else if (rty == BALL && !eul) {
                double oo;
                if constexpr (CACHED) oo = double(tc[0]);
                else { oo = 1.0 / sqrt(qv[0]*qv[0] + qv[1]*qv[1] + qv[2]*qv[2] + qv[3]*qv[3]); tc[0] = K(oo); }
                R_FM = quat_rot(K(qv[0]*oo), K(qv[1]*oo), K(qv[2]*oo), K(qv[3]*oo));
                H_FM[0] = {V3(o, z, z), zero3<K>()};
                H_FM[1] = {V3(z, o, z), zero3<K>()};
                H_FM[2] = {V3(z, z, o), zero3<K>()};
            } else if (xyz) {
                // Body-fixed 1-2-3 rotation (_Gimbal.h:141-176, Rotation.h:
                // 342-349); a Universal's third angle has (sin, cos) = (0, 1),
                // which gives its x-then-y rotation entry for entry.
                const K s0s1 = s0*s1, s2c0 = s2*c0, c0c2 = c0*c2, nc1 = -c1;
                R_FM.a[0]= c1*c2;             R_FM.a[1]= s2*nc1;            R_FM.a[2]= s1;
                R_FM.a[3]= s2c0 + s0s1*c2;    R_FM.a[4]= c0c2 - s0s1*s2;    R_FM.a[5]= s0*nc1;
                R_FM.a[6]= s0*s2 - s1*c0c2;   R_FM.a[7]= s0*c2 + s1*s2c0;   R_FM.a[8]= c0*c1;
                if (rty == BALL) {
                    // Euler-angle Ball: u is still w_FM in F
                    H_FM[0] = {V3(o, z, z), zero3<K>()};
                    H_FM[1] = {V3(z, o, z), zero3<K>()};
                    H_FM[2] = {V3(z, z, o), zero3<K>()};
                } else {
                    H_FM[0] = {V3(o, z, z), zero3<K>()};
                    H_FM[1] = {V3(z, c0, s0), zero3<K>()};
                    if (rty == GIMBAL) H_FM[2] = {V3(s1, s0*nc1, c0*c1), zero3<K>()};
                }
            } else {
                // Pin (_Pin.h), Planar (_Planar.h:115-156), Cylinder
                // (_Cylinder.h:100-139), Screw (_Screw.h:109-148): rotation
                // about z by q0, H_FM constant in F
                R_FM.a[0]=c0; R_FM.a[1]=-s0; R_FM.a[2]=z; R_FM.a[3]=s0; R_FM.a[4]=c0; R_FM.a[5]=z;
                R_FM.a[6]=z;  R_FM.a[7]=z;   R_FM.a[8]=o;
                H_FM[0] = {V3(z, z, o), zero3<K>()};
                if (rty == PLANAR) {
                    p_FM = V3(K(qv[1]), K(qv[2]), z);
                    H_FM[1] = {zero3<K>(), V3(o, z, z)};
                    H_FM[2] = {zero3<K>(), V3(z, o, z)};
                } else if (rty == CYLINDER) {
                    p_FM = V3(z, z, K(qv[1]));
                    H_FM[1] = {zero3<K>(), V3(z, z, o)};
                } else if (rty == SCREW) {
                    p_FM = V3(z, z, K(qv[0])*pitch);
                    H_FM[0] = {V3(z, z, o), V3(z, z, pitch)};
                } else if (EXT && rty == BENDSTRETCH) {
                    // RigidBodyNodeSpec_PolarCoords.h:128-172: slide q1 along
                    // the rotated x; H_FM(0) = (z, z x p_FM), H_FM(1) = (0, Mx)
                    p_FM = mul(R_FM, V3(K(qv[1]), z, z));
                    H_FM[0] = {V3(z, z, o), cross(V3(z, z, o), p_FM)};
                    H_FM[1] = {zero3<K>(), col(R_FM, 0)};
                }
            }
        } else if constexpr (TYPE == PIN) {
            K sn, cs;
            if constexpr (CACHED) { cs = tc[0]; sn = tc[1]; } else { sincosk<K>(qv[0], &sn, &cs); tc[0] = cs; tc[1] = sn; }
            R_FM.a[0]=cs; R_FM.a[1]=-sn; R_FM.a[2]=z; R_FM.a[3]=sn; R_FM.a[4]=cs; R_FM.a[5]=z;
            R_FM.a[6]=z;  R_FM.a[7]=z;   R_FM.a[8]=o;
            H_FM[0] = {V3(z, z, o), zero3<K>()};
        } else if constexpr (TYPE == SLIDER) {
            R_FM = ident<K>(); p_FM = V3(K(qv[0]), z, z);
            H_FM[0] = {zero3<K>(), V3(o, z, z)};
        } else if constexpr (TYPE == UNIVERSAL) {
            K s1, c1, s2, c2;
            if constexpr (CACHED) { c1 = tc[0]; s1 = tc[1]; c2 = tc[2]; s2 = tc[3]; }
            else { sincosk<K>(qv[0], &s1, &c1); sincosk<K>(qv[1], &s2, &c2); tc[0] = c1; tc[1] = s1; tc[2] = c2; tc[3] = s2; }
            R_FM.a[0]= c2;    R_FM.a[1]= z;  R_FM.a[2]= s2;
            R_FM.a[3]= s2*s1; R_FM.a[4]= c1; R_FM.a[5]= -s1*c2;
            R_FM.a[6]=-s2*c1; R_FM.a[7]= s1; R_FM.a[8]= c1*c2;
            H_FM[0] = {V3(o, z, z), zero3<K>()};
            H_FM[1] = {col(R_FM, 1), zero3<K>()};
        } else {
            double oo;
            if constexpr (CACHED) oo = double(tc[0]);
            else { oo = 1.0 / sqrt(qv[0]*qv[0] + qv[1]*qv[1] + qv[2]*qv[2] + qv[3]*qv[3]); tc[0] = K(oo); }
            R_FM = quat_rot(K(qv[0]*oo), K(qv[1]*oo), K(qv[2]*oo), K(qv[3]*oo));
            H_FM[0] = {V3(o, z, z), zero3<K>()};
            H_FM[1] = {V3(z, o, z), zero3<K>()};
            H_FM[2] = {V3(z, z, o), zero3<K>()};
            if constexpr (TYPE == FREE) {
                p_FM = V3(K(qv[4]), K(qv[5]), K(qv[6]));
                H_FM[3] = {zero3<K>(), V3(o, z, z)};
                H_FM[4] = {zero3<K>(), V3(z, o, z)};
                H_FM[5] = {zero3<K>(), V3(z, z, o)};
            }
        }
    }
};

// X_PB = X_PF X_FM X_MB (RigidBodyNodeSpec.h:554-569)
template <class K, class PA>
__device__ __forceinline__ void body_in_parent(const PA& P, const m33<K>& R_FM, const v3<K>& p_FM, m33<K>& R_PB,
                                               v3<K>& p_PB) {
    const v3<K> p_MB = ld3<K>(P + PRM_P_MB);
    if constexpr (PrmTraits<PA>::idf) {
        R_PB = R_FM;
        p_PB = p_FM + mul(R_FM, p_MB);
    } else {
        const m33<K> R_PF = ldm<K>(P + PRM_R_PF), R_MB = ldm<K>(P + PRM_R_MB);
        const v3<K> p_PF = ld3<K>(P + PRM_P_PF);
        const m33<K> R_FB = mul(R_FM, R_MB);
        const v3<K> p_FB = p_FM + mul(R_FM, p_MB);
        R_PB = mul(R_PF, R_FB);
        p_PB = p_PF + mul(R_PF, p_FB);
    }
}

// H_PB_G = R_GF (H_FM + H_MB_F)  (RigidBodyNodeSpec.cpp:44-74)
template <int TYPE, class K, class PA>
__device__ __forceinline__ void hinge(const PA& P, const Joint<TYPE, K>& J, const m33<K>& R_GP, m33<K>& R_GF,
                                      v3<K>& r_MB_F, sv<K>* H) {
    if constexpr (PrmTraits<PA>::idf) R_GF = R_GP;
    else R_GF = mul(R_GP, ldm<K>(P + PRM_R_PF));
    r_MB_F = mul(J.R_FM, ld3<K>(P + PRM_P_MB));
#pragma unroll
    for (int j = 0; j < MobC<TYPE>::D; ++j)
        H[j] = {mul(R_GF, J.H_FM[j].w), mul(R_GF, J.H_FM[j].v - cross(r_MB_F, J.H_FM[j].w))};
}

// Mobilizer coriolis acceleration a = (HDot u) + (0, w_GP x (v_GB - v_GP))
// (RigidBodyNodeSpec.cpp:82-129, RigidBodyNode.cpp:98-174).  HDot_FM is
// nonzero for Universal and Gimbal only.
// REV: the body may be reversed (J.rv); its HDot_FM then follows
// calcReverseMobilizerHDot_FM (RigidBodyNodeSpec.cpp:168-216):
//   HDot_FM_w = w_FM x H_FM_w - R_FM HDot_MF_w
//   HDot_FM_v = w_FM x H_FM_v - R_FM HDot_MF_v
//               - (p_FM x HDot_FM_w + v_FM x H_FM_w - w_FM x (p_FM x H_FM_w))
// with HDot_MF from the mobilizer's own H columns (HDot_MF_v is zero for
// every mobilizer carried here).
template <int TYPE, class K, bool REV = false>
__device__ __forceinline__ sv<K> coriolis(const Joint<TYPE, K>& J, const m33<K>& R_GF, const v3<K>& r_MB_F,
                                          const sv<K>* H, const double* uv, const sv<K>& V_GP, const sv<K>& V_GB) {
    constexpr int D = MobC<TYPE>::D;
    sv<K> V_FM = svzero<K>();
#pragma unroll
    for (int j = 0; j < D; ++j) V_FM = V_FM + K(uv[j])*J.H_FM[j];
    const v3<K> w_FM = V_FM.w;
    const v3<K> wxr = cross(w_FM, r_MB_F);
    sv<K> VD = svzero<K>();
#pragma unroll
    for (int j = 0; j < D; ++j) {
        v3<K> hdw = zero3<K>();
        if constexpr (TYPE == UNIVERSAL) { if (j == 1) hdw = cross(w_FM, col(J.R_FM, 1)); }
        if constexpr (TYPE == FREE && REV) {
            // a Bushing's angle columns turn with the angles before them
            // (_Bushing.h:195-220): d/dt H_FM(j) = (sum_{k<j} qdot_k H_FM(k).w) x H_FM(j).w
            if (J.bu && (j == 1 || j == 2)) {
                const v3<K> wj = j == 1 ? K(uv[0])*J.H_FM[0].w : K(uv[0])*J.H_FM[0].w + K(uv[1])*J.H_FM[1].w;
                hdw = cross(wj, J.H_FM[j].w);
            }
        }
        v3<K> hdv = zero3<K>();
        if constexpr (TYPE == GEN) {
            // Universal and Gimbal (_Universal.h:173-189, _Gimbal.h:180-202):
            // column j turns with the angles before it, d/dt H_FM(j) =
            // (sum_{k<j} u_k H_FM(k).w) x H_FM(j).w; a Universal's padded
            // third column is zero.  For a reversed body the same formula in
            // the mobilizer's own columns gives HDot_MF.
            v3<K> h0 = J.H_FM[0].w, h1 = J.H_FM[1].w, hj = J.H_FM[j].w;
            if constexpr (REV) { if (J.rv) { h0 = J.Hm[0]; h1 = J.Hm[1]; hj = J.Hm[j]; } }
            if (j >= 1 && (J.ty == UNIVERSAL || J.ty == GIMBAL)) {
                const v3<K> wj = j == 1 ? K(uv[0])*h0 : K(uv[0])*h0 + K(uv[1])*h1;
                hdw = cross(wj, hj);
            }
            if constexpr (REV) {
                // BendStretch (PolarCoords.h:174-200, never reversed):
                // HDot_FM(0) = (0, z x v_FM), HDot_FM(1) = (0, w_FM x Mx)
                if (J.ty == BENDSTRETCH && j < 2)
                    hdv = j == 0 ? cross(V3(K(0), K(0), K(1)), V_FM.v) : cross(w_FM, col(J.R_FM, 0));
                if (J.rv) {
                    const v3<K> Hw = J.H_FM[j].w, Hv = J.H_FM[j].v;
                    hdw = cross(w_FM, Hw) - mul(J.R_FM, hdw);
                    hdv = cross(w_FM, Hv) - ((cross(J.p_FM, hdw) + cross(V_FM.v, Hw)) - cross(w_FM, cross(J.p_FM, Hw)));
                }
            }
        }
        const v3<K> lin = (hdv - cross(r_MB_F, hdw)) - cross(wxr, J.H_FM[j].w);
        const sv<K> HD = {mul(R_GF, hdw) + cross(V_GP.w, H[j].w), mul(R_GF, lin) + cross(V_GP.w, H[j].v)};
        VD = VD + K(uv[j])*HD;
    }
    return {VD.w, VD.v + cross(V_GP.w, V_GB.v - V_GP.v)};
}

// A pointer that is the same in every active lane, marked as such: the
// compiler then keeps it in scalar registers and addresses per-lane
// elements with a 32-bit lane offset.
template <class T> __device__ __forceinline__ TD_GLOBAL T* uniform_ptr(T* p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (TD_GLOBAL T*)(((uint64_t)hi << 32) | lo);
}
// Per-lane SoA rows: row k of this lane at base[k * L + lane].  U: base and
// k are the same in every active lane (true for all rows of the chain
// evaluator, whose topology is shared by the batch), so the row address
// base + k*L is a scalar and the lane part a 32-bit offset.
template <class T, bool U = true> struct RowsT {
    T* base; int L; unsigned lane;
    __device__ __forceinline__ TD_GLOBAL T& operator()(int k) const {
        T* rb = base + (size_t)k * L;
        if constexpr (U) {
            // 32-bit byte offset: the access becomes scalar base + lane offset
            const uint32_t off = lane * (uint32_t)sizeof(T);
            return *(TD_GLOBAL T*)((TD_GLOBAL char*)uniform_ptr(rb) + off);
        } else {
            return ((TD_GLOBAL T*)rb)[lane];
        }
    }
    __device__ __forceinline__ RowsT shifted(int k) const { return {base + (size_t)k * L, L, lane}; }
};
using Rows = RowsT<double>;

// Where one body's sweep hand-over lives: G (6 dof values of type K), the
// trig cache (K) and DI*eps (dof doubles).  FP64 path: G and DI*eps at the
// body's G slot of the double scratch, trig cache at its own double rows.
// Mixed path: G and the trig cache in float rows (G at 6*uIndex, trig cache
// after all G rows), DI*eps at the double G slot.
// GS, TS: the storage types of G and of the trig cache (double, or float
// where those rows are single precision).
template <class GS, class TS = GS> struct KStore {
    RowsT<GS> kg; RowsT<TS> kt; Rows d; int g, tc, de;
    __device__ __forceinline__ TD_GLOBAL GS& G(int i) const { return kg(g + i); }
    __device__ __forceinline__ TD_GLOBAL TS& T(int i) const { return kt(tc + i); }
    __device__ __forceinline__ TD_GLOBAL double& DE(int i) const { return d(de + i); }
};

// Everything one body's step needs from memory, loaded one body ahead so the
// loads are in flight while the previous body computes.
template <class K> struct Pre {
    double q[7], u[6]; K tc[4]; K g[18]; double de[3];
    int ty;                 // GEN: the body's mobilizer
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
    Rows c0, c1;           // mixed path: Kahan compensation of y0 and of y1
};
// Error-norm partial sums (IntegratorRep.h:454-523): weighted q part of the
// quaternions and of the other q, relative u part, quaternion norm errors.
struct NormSums { double sq_other, sq_quat, su, s2; };
// Type-exact prefetch: NQ q, D u, the trig cache entries the mobilizer uses
// and (pass 3) its G and DI*eps values.
template <int TYPE> struct TcN { static constexpr int n = TYPE == PIN ? 2 : (TYPE == UNIVERSAL || TYPE == GEN) ? 4 : TYPE == SLIDER ? 0 : 1; };
template <int TYPE, bool G, class K, class YR, class ST>
__device__ __forceinline__ void prefetch_t(Pre<K>& x, const YR& y, int nq, int qi, int ui, const ST& st,
                                           bool tc) {
    constexpr int NQ = MobC<TYPE>::NQ, D = MobC<TYPE>::D;
#pragma unroll
    for (int k = 0; k < NQ; ++k) x.q[k] = y(qi + k);
#pragma unroll
    for (int k = 0; k < D; ++k) x.u[k] = y(nq + ui + k);
    if (tc) {
#pragma unroll
        for (int k = 0; k < TcN<TYPE>::n; ++k) x.tc[k] = st.T(k);
    }
    if constexpr (G && D <= 3) {
#pragma unroll
        for (int k = 0; k < 6*D; ++k) x.g[k] = st.G(k);
#pragma unroll
        for (int k = 0; k < D; ++k) x.de[k] = st.DE(k);
    }
}
// AF: the body may be a Free root.  Only body 1 can be Free (the host rejects
// Free elsewhere), and it is peeled out of every loop, so the loops prefetch
// with AF = false and never hold a 7-q/6-u body's registers.
template <bool G, bool AF, class K, class ST>
__device__ __forceinline__ void prefetch(Pre<K>& x, const Rows& y, int nq, int4 ba, const ST& st, bool tc) {
    switch (ba.y) {
    case PIN:       prefetch_t<PIN, G>(x, y, nq, ba.z, ba.w, st, tc); break;
    case SLIDER:    prefetch_t<SLIDER, G>(x, y, nq, ba.z, ba.w, st, tc); break;
    case UNIVERSAL: prefetch_t<UNIVERSAL, G>(x, y, nq, ba.z, ba.w, st, tc); break;
    default:
        if constexpr (AF) { if (ba.y == FREE) { prefetch_t<FREE, G>(x, y, nq, ba.z, ba.w, st, tc); break; } }
        prefetch_t<BALL, G>(x, y, nq, ba.z, ba.w, st, tc); break;
    }
}

// Euler-angle mode of a Ball (generic) or Free body, in models that may use it
// (a Bushing's q are such angles too).
template <int TYPE, class PA>
__device__ __forceinline__ bool euler_of(const PA& P) {
    if constexpr ((TYPE == GEN || TYPE == FREE) && PrmTraits<PA>::ext) return P[PRM_EULER] != 0.0;
    return false;
}
// A Bushing, carried as a Free body, in models that may hold one.
template <int TYPE, class PA>
__device__ __forceinline__ bool bushing_of(const PA& P) {
    if constexpr (TYPE == FREE && PrmTraits<PA>::ext) return P[PRM_EULER] == 2.0;
    return false;
}
// Reverses a generic body's joint when the model may hold reversed
// mobilizers and this body's is.
template <int TYPE, class K, class PA>
__device__ __forceinline__ void reverse_if(const PA& P, Joint<TYPE, K>& J) {
    if constexpr (TYPE == GEN && PrmTraits<PA>::ext) { if (P[PRM_REV] != 0.0) J.reverse(); }
}

// The pitch of a Screw body (padded generic bodies only), else zero.
template <int TYPE, class K, class PA>
__device__ __forceinline__ K screw_pitch(const PA& P, const Pre<K>& x) {
    if constexpr (TYPE == GEN) { if (x.ty == SCREW) return K(P[PRM_PITCH]); }
    return K(0);
}

// ---------------------------------------------------------------------------
// Pass 1 body: forward transforms and velocities, qdot out, trig cache out.
// ---------------------------------------------------------------------------
template <int TYPE, class K, class PA, class YR, class ST>
__device__ __forceinline__ void pass1_body(const PA& P, const Pre<K>& x, int qi, const YR& f, const ST& st,
                                           m33<K>& R, v3<K>& p, sv<K>& V) {
    constexpr int D = MobC<TYPE>::D;
    const double* qv = x.q; const double* uv = x.u;
    const bool eul = euler_of<TYPE>(P);
    if constexpr (TYPE == GEN) {
        if (x.ty == BALL) {
            if (!eul) {
                double d4[4]; quat_dot(qv, V3(uv[0], uv[1], uv[2]), d4);
#pragma unroll
                for (int k = 0; k < 4; ++k) f(qi + k) = d4[k];
            }
        } else {
            const int d = mob_dof(x.ty);
#pragma unroll
            for (int k = 0; k < D; ++k) if (k < d) f(qi + k) = uv[k];
        }
    } else if constexpr (TYPE == BALL || TYPE == FREE) {
        if (!eul) {
            double d4[4]; quat_dot(qv, V3(uv[0], uv[1], uv[2]), d4);
#pragma unroll
            for (int k = 0; k < 4; ++k) f(qi + k) = d4[k];
            if constexpr (TYPE == FREE) { f(qi+4) = uv[3]; f(qi+5) = uv[4]; f(qi+6) = uv[5]; }
        }
    } else {
#pragma unroll
        for (int k = 0; k < D; ++k) f(qi + k) = uv[k];
    }
    Joint<TYPE, K> J;
    K tc[4] = {K(0), K(0), K(0), K(0)};
    J.template set<false, PrmTraits<PA>::ext>(qv, tc, TYPE == GEN ? x.ty : TYPE, screw_pitch<TYPE, K>(P, x), eul,
                                           bushing_of<TYPE>(P));
    if constexpr (TYPE == GEN || TYPE == FREE) {
        if (eul && (TYPE == FREE || x.ty == BALL)) {
            if (TYPE == FREE && J.bu) {
                // Bushing: qdot = u
#pragma unroll
                for (int k = 0; k < 6; ++k) f(qi + k) = uv[k];
            } else {
                // x-y-z angle rates N(q) w (RigidBodyNodeSpec_Ball.h:183-214,
                // _Free.h:222-257); the fourth q slot is unused, its rate zero
                const v3<K> ed = euler_dot(J.ec[0], J.ec[1], J.ec[2], J.ec[3], V3(K(uv[0]), K(uv[1]), K(uv[2])));
                f(qi) = double(ed.x); f(qi + 1) = double(ed.y); f(qi + 2) = double(ed.z);
                if constexpr (TYPE == FREE) { f(qi+3) = uv[3]; f(qi+4) = uv[4]; f(qi+5) = uv[5]; f(qi+6) = 0.0; }
                else f(qi + 3) = 0.0;
            }
        }
    }
    reverse_if<TYPE>(P, J);
#pragma unroll
    for (int k = 0; k < TcN<TYPE>::n; ++k) st.T(k) = tc[k];
    m33<K> R_PB; v3<K> p_PB;
    body_in_parent<K>(P, J.R_FM, J.p_FM, R_PB, p_PB);
    m33<K> R_GF; v3<K> r_MB_F; sv<K> H[D];
    hinge<TYPE, K>(P, J, R, R_GF, r_MB_F, H);
    const v3<K> phi = mul(R, p_PB);
    sv<K> Vpb = svzero<K>();
#pragma unroll
    for (int j = 0; j < D; ++j) Vpb = Vpb + K(uv[j])*H[j];
    V = phiT_mul(phi, V) + Vpb;
    p = p + mul(R, p_PB);
    R = mul(R, R_PB);
}

// Small dense inverse of D (FP64 on both paths): cofactor forms for dof<=3
// (SmallMatrixMixed.h:895-960), Gauss-Jordan without pivoting for dof=6 (D is
// symmetric positive definite).
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
// accP, accZ (nullable): contributions of further children, already shifted
// to this body (trees); the child in PP/zp/phic is added by this routine.
template <int TYPE, class K, class PA, class ST>
__device__ __forceinline__ void pass2_body(const PA& P, const Pre<K>& x, const v3<K>& grav, const sv<K>& Fapp,
                                           const ST& st, m33<K>& R, v3<K>& p, sv<K>& V, bool hasChild,
                                           abi<K>& PP, sv<K>& zp, v3<K>& phic,
                                           const abi<K>* accP = nullptr, const sv<K>* accZ = nullptr,
                                           const double* ukn = nullptr) {
    // ukn: the body's known udot (locks, constant prescribed udot), or null.
    // Such a body hands its parent P+ = P and z+ = z with z += P H udot
    // (RigidBodyNodeSpec.cpp:292-296, 380-399) and keeps G = 0 and udot in
    // its DI*eps slot, so pass 3 forms udot = DI*eps - G^T A+ = udot exactly
    // (:434-443).
    constexpr int D = MobC<TYPE>::D;
    const double* uv = x.u;
    Joint<TYPE, K> J;
    K tc[4] = {x.tc[0], x.tc[1], x.tc[2], x.tc[3]};
    J.template set<true, PrmTraits<PA>::ext>(x.q, tc, TYPE == GEN ? x.ty : TYPE, screw_pitch<TYPE, K>(P, x),
                                          euler_of<TYPE>(P), bushing_of<TYPE>(P));
    reverse_if<TYPE>(P, J);
    m33<K> R_PB; v3<K> p_PB;
    body_in_parent<K>(P, J.R_FM, J.p_FM, R_PB, p_PB);
    const m33<K> R_GB = R; const v3<K> p_GB = p; const sv<K> V_GB = V;
    const m33<K> R_GP = mulABt(R_GB, R_PB);
    const v3<K> phi = mul(R_GP, p_PB);
    const v3<K> p_GP = p_GB - phi;
    m33<K> R_GF; v3<K> r_MB_F; sv<K> H[D];
    hinge<TYPE, K>(P, J, R_GP, R_GF, r_MB_F, H);
    sv<K> Vpb = svzero<K>();
#pragma unroll
    for (int j = 0; j < D; ++j) Vpb = Vpb + K(uv[j])*H[j];
    const v3<K> w_GP = V_GB.w - Vpb.w;
    const v3<K> v_GP = (V_GB.v - Vpb.v) - cross(w_GP, phi);
    const sv<K> V_GP = {w_GP, v_GP};
    const sv<K> a = coriolis<TYPE, K, PrmTraits<PA>::ext>(J, R_GF, r_MB_F, H, uv, V_GP, V_GB);
    const K m = K(P[PRM_MASS]);
    const v3<K> Mk_p = mul(R_GB, ld3<K>(P + PRM_COM));
    const s33<K> Gu = {K(P[PRM_G]), K(P[PRM_G+1]), K(P[PRM_G+2]), K(P[PRM_G+3]), K(P[PRM_G+4]), K(P[PRM_G+5])};
    const s33<K> Mk_G = rotate_sym(R_GB, Gu);
    const v3<K> w = V_GB.w;
    const sv<K> bg = {m*cross(w, mul(Mk_G, w)), m*cross(w, cross(w, Mk_p))};
    const v3<K> fg = m*grav;
    const sv<K> F = sv<K>{cross(Mk_p, fg), fg} + Fapp;
    abi<K> Pb = abi_from_rbi(m, Mk_p, Mk_G);
    if (accP) {
        Pb.M = Pb.M + accP->M; Pb.J = Pb.J + accP->J;
#pragma unroll
        for (int i = 0; i < 9; ++i) Pb.F.a[i] = Pb.F.a[i] + accP->F.a[i];
    }
    if (hasChild) abi_add_shifted(Pb, PP, phic);
    sv<K> PH[D];
#pragma unroll
    for (int j = 0; j < D; ++j) PH[j] = abi_mul(Pb, H[j]);
    double Dm[D*D], DI[D*D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int k = 0; k < D; ++k) Dm[r*D+k] = dotd(H[r], PH[k]);
    if constexpr (TYPE == GEN) {
#pragma unroll
        for (int j = 0; j < D; ++j) if (j >= J.d) Dm[j*D+j] = 1.0;
    }
    dinv<D>(Dm, DI);
    sv<K> G[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        dv acc = svzero<double>();
#pragma unroll
        for (int j = 0; j < D; ++j) acc = acc + DI[j*D+k]*tod(PH[j]);
        G[k] = tok<K>(acc);
    }
    sv<K> z = (abi_mul(Pb, a) + bg) - F;
    if (ukn) {
        sv<K> Hu = svzero<K>();
#pragma unroll
        for (int j = 0; j < D; ++j) if (TYPE != GEN || j < J.d) Hu = Hu + K(ukn[j])*H[j];
        z = z + abi_mul(Pb, Hu);
    }
    if (accZ) z = z + *accZ;
    if (hasChild) z = z + phi_mul(phic, zp);
    double eps[D];
#pragma unroll
    for (int j = 0; j < D; ++j) eps[j] = 0.0 - dotd(H[j], z);
    sv<K> zn = z;
#pragma unroll
    for (int j = 0; j < D; ++j) zn = zn + K(eps[j])*G[j];
    K mm[9], ms[9], in[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            K a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                const K gw = r == 0 ? G[j].w.x : r == 1 ? G[j].w.y : G[j].w.z;
                const K gv = r == 0 ? G[j].v.x : r == 1 ? G[j].v.y : G[j].v.z;
                const K pv = k == 0 ? PH[j].v.x : k == 1 ? PH[j].v.y : PH[j].v.z;
                const K pw = k == 0 ? PH[j].w.x : k == 1 ? PH[j].w.y : PH[j].w.z;
                a0 += gw*pv; a1 += gv*pv; a2 += gw*pw;
            }
            mm[r*3+k] = a0; ms[r*3+k] = a1; in[r*3+k] = a2;
        }
    const K hf = K(0.5);
    PP.M = Pb.M - s33<K>{ms[0], ms[4], ms[8], (ms[3]+ms[1])*hf, (ms[6]+ms[2])*hf, (ms[7]+ms[5])*hf};
    PP.J = Pb.J - s33<K>{in[0], in[4], in[8], (in[3]+in[1])*hf, (in[6]+in[2])*hf, (in[7]+in[5])*hf};
#pragma unroll
    for (int i = 0; i < 9; ++i) PP.F.a[i] = Pb.F.a[i] - mm[i];
    zp = zn; phic = phi;
    if (ukn) {
        PP = Pb; zp = z;
#pragma unroll
        for (int j = 0; j < D; ++j) G[j] = svzero<K>();
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
        st.G(6*j + 0) = G[j].w.x; st.G(6*j + 1) = G[j].w.y; st.G(6*j + 2) = G[j].w.z;
        st.G(6*j + 3) = G[j].v.x; st.G(6*j + 4) = G[j].v.y; st.G(6*j + 5) = G[j].v.z;
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
        double sacc = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) sacc += DI[r*D+k]*eps[k];
        if (ukn) sacc = (TYPE != GEN || r < J.d) ? ukn[r] : 0.0;
        st.DE(r) = sacc;
    }
    R = R_GP; p = p_GP; V = V_GP;
}

// ---------------------------------------------------------------------------
// Pass 3 body: forward kinematics again, then udot and A_GB.
// ---------------------------------------------------------------------------
template <int TYPE, class K, class PA, class YR, class ST>
__device__ __forceinline__ void pass3_body(const PA& P, const Pre<K>& x, int ui, const YR& fu, const ST& st,
                                           m33<K>& R, v3<K>& p, sv<K>& V, sv<K>& A, double* ud_out = nullptr) {
    constexpr int D = MobC<TYPE>::D;
    const double* uv = x.u;
    Joint<TYPE, K> J;
    K tc[4] = {x.tc[0], x.tc[1], x.tc[2], x.tc[3]};
    J.template set<true, PrmTraits<PA>::ext>(x.q, tc, TYPE == GEN ? x.ty : TYPE, screw_pitch<TYPE, K>(P, x),
                                          euler_of<TYPE>(P), bushing_of<TYPE>(P));
    reverse_if<TYPE>(P, J);
    m33<K> R_PB; v3<K> p_PB;
    body_in_parent<K>(P, J.R_FM, J.p_FM, R_PB, p_PB);
    m33<K> R_GF; v3<K> r_MB_F; sv<K> H[D];
    hinge<TYPE, K>(P, J, R, R_GF, r_MB_F, H);
    const v3<K> phi = mul(R, p_PB);
    sv<K> Vpb = svzero<K>();
#pragma unroll
    for (int j = 0; j < D; ++j) Vpb = Vpb + K(uv[j])*H[j];
    const sv<K> V_GP = V;
    const sv<K> V_GB = phiT_mul(phi, V_GP) + Vpb;
    const sv<K> a = coriolis<TYPE, K, PrmTraits<PA>::ext>(J, R_GF, r_MB_F, H, uv, V_GP, V_GB);
    const sv<K> Ap = phiT_mul(phi, A);
    double ud[D];
#pragma unroll
    for (int r = 0; r < D; ++r) {
        sv<K> G;
        if constexpr (D <= 3) {
            G = {V3(x.g[6*r], x.g[6*r+1], x.g[6*r+2]), V3(x.g[6*r+3], x.g[6*r+4], x.g[6*r+5])};
            ud[r] = x.de[r] - dotd(G, Ap);
        } else {
            G = {V3(K(st.G(6*r)), K(st.G(6*r + 1)), K(st.G(6*r + 2))),
                 V3(K(st.G(6*r + 3)), K(st.G(6*r + 4)), K(st.G(6*r + 5)))};
            ud[r] = st.DE(r) - dotd(G, Ap);
        }
    }
    sv<K> Hu = svzero<K>();
#pragma unroll
    for (int j = 0; j < D; ++j) Hu = Hu + K(ud[j])*H[j];
    A = (Ap + Hu) + a;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        if (TYPE == GEN && j >= J.d) continue;
        fu(ui + j) = ud[j]; if (ud_out) ud_out[j] = ud[j];
    }
    p = p + mul(R, p_PB);
    R = mul(R, R_PB);
    V = V_GB;
}

// ---------------------------------------------------------------------------
// The evaluator.  K: the arithmetic of kinematics, bias forces and
// articulated inertias (double, or float on the mixed path).  GS, TS: the
// storage of the hand-over rows G and trig cache.  sc: this lane's double
// scratch rows (G records at body_b.w, FP64 trig cache at tree_b.z +
// 4(b-1)); sf: the float scratch (G at 6 uIndex, then the trig cache at
// 6 nu + 4(b-1)), used for the rows stored in single precision.
// ---------------------------------------------------------------------------
template <class PA, class K, class GS = K, class TS = GS>
struct ChainEval {
    static constexpr bool MIXED = !std::is_same<K, double>::value;
    static constexpr bool G32 = std::is_same<GS, float>::value;
    static constexpr bool T32 = std::is_same<TS, float>::value;
    PA prm; const int4* body_a; const int4* body_b;
    int nb, nq, nu, tcbase;
    Rows sc; RowsT<float> sf; v3<K> grav; const double* fapp;

    __device__ ChainEval(const RkmArgs& a, int lane, int tr) {
        const int4 ta = ldi4(a.T.tree_a, tr);
        nb = ta.x; nq = ta.y; nu = ta.z;
        body_a = a.T.body_a + ta.w; body_b = a.T.body_b + ta.w;
        tcbase = ldi4(a.T.tree_b, tr).z;
        sc = {a.scratch, a.L, (unsigned)lane};
        sf = {a.scratch32, a.L, (unsigned)lane};
        if constexpr (PrmTraits<PA>::lane) {
            // per-lane model: nb*PRM_STRIDE rows, then 3 gravity rows
            const PrmL pl{a.prm_lane, a.L, (unsigned)lane};
            prm = PA(pl);
            const PrmL g = pl + (size_t)nb * PRM_STRIDE;
            grav = V3(K(g[0]), K(g[1]), K(g[2]));
        } else {
            prm = PA(cptr(a.T.prm) + (size_t)ta.w * PRM_STRIDE);
            const cdptr g = cptr(a.T.gravity) + 4*tr;
            grav = V3(K(g[0]), K(g[1]), K(g[2]));
        }
        fapp = a.fapp ? a.fapp + 6*(size_t)a.off.b_off[lane] : nullptr;
    }
    // Hand-over rows of body b (1-based, table entry ba, G slot gs).
    __device__ __forceinline__ KStore<GS, TS> store(int b, int4 ba, int gs) const {
        KStore<GS, TS> st;
        if constexpr (G32) { st.kg = sf; st.g = 6*ba.w; } else { st.kg = sc; st.g = gs; }
        if constexpr (T32) { st.kt = sf; st.tc = 6*nu + 4*(b-1); } else { st.kt = sc; st.tc = tcbase + 4*(b-1); }
        st.d = sc; st.de = gs + 6*mob_dof(ba.y);
        return st;
    }
    __device__ __forceinline__ sv<K> F(int b) const {
        if (!fapp) return svzero<K>();
        const double* fb = fapp + 6*b;
        return {V3(K(fb[0]), K(fb[1]), K(fb[2])), V3(K(fb[3]), K(fb[4]), K(fb[5]))};
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
    template <int STAGE, bool AF>
    __device__ __forceinline__ void pre1_any(Pre1& x, const RkRows& r, int4 ba) const {
        switch (ba.y) {
        case PIN:       pre1<PIN, STAGE>(x, r, ba.z, ba.w); break;
        case SLIDER:    pre1<SLIDER, STAGE>(x, r, ba.z, ba.w); break;
        case UNIVERSAL: pre1<UNIVERSAL, STAGE>(x, r, ba.z, ba.w); break;
        default:
            if constexpr (AF) { if (ba.y == FREE) { pre1<FREE, STAGE>(x, r, ba.z, ba.w); break; } }
            pre1<BALL, STAGE>(x, r, ba.z, ba.w); break;
        }
    }
    // stage input from the raw sources; stored to r.st (stages 1-4)
    template <int TYPE, int STAGE>
    __device__ __forceinline__ void form(const Pre1& x, Pre<K>& o, const RkRows& r, int qi, int ui, double h) const {
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
    __device__ __forceinline__ void combine(const Pre<K>& x, const double* y0, const double* f0, const double* fb,
                                            const double* c0, const double* ud, const RkRows& r, int qi, int ui,
                                            double h, NormSums& ns) const {
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
            const int idx = k < NQ ? qi + k : nq + ui + (k - NQ);
            if constexpr (MIXED) {
                // Kahan: the step's increment carries the rounding lost when
                // the previous increment was added to the state
                const double dy = (h/6)*((f0[k] + 4*fb[k]) + f4[k]) - c0[k];
                y1[k] = y0[k] + dy;
                r.c1(idx) = (y1[k] - y0[k]) - dy;
            } else {
                y1[k] = y0[k] + (h/6)*((f0[k] + 4*fb[k]) + f4[k]);
            }
            e[k] = 0.2 * fabs(y1[k] - ys);
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
                                       m33<K>& R, v3<K>& p, sv<K>& V) const {
        Pre<K> x;
        form<TYPE, STAGE>(raw, x, r, ba.z, ba.w, h);
        pass1_body<TYPE, K>(prm + (size_t)(b-1) * PRM_STRIDE, x, ba.z, r.out, store(b, ba, 0), R, p, V);
    }
    template <int TYPE, int STAGE>
    __device__ __forceinline__ void p3(const Pre<K>& x, const RkRows& r, const Rows& fu, const int4 ba, int b, double h,
                                       m33<K>& R, v3<K>& p, sv<K>& V, sv<K>& A, NormSums& ns) const {
        constexpr int NQ = MobC<TYPE>::NQ, D = MobC<TYPE>::D;
        double y0[NQ + D], f0[NQ + D], fb[NQ + D], c0[NQ + D];
        if constexpr (STAGE == 4) {
#pragma unroll
            for (int k = 0; k < NQ + D; ++k) {
                const int idx = k < NQ ? ba.z + k : nq + ba.w + (k - NQ);
                y0[k] = r.y0(idx); f0[k] = r.f0(idx); fb[k] = r.fb(idx);
                if constexpr (MIXED) c0[k] = r.c0(idx);
            }
        }
        double ud[D];
        pass3_body<TYPE, K>(prm + (size_t)(b-1) * PRM_STRIDE, x, ba.w, fu, store(b, ba, ldi4(body_b, b-1).w),
                            R, p, V, A, ud);
        if constexpr (STAGE == 4) combine<TYPE>(x, y0, f0, fb, c0, ud, r, ba.z, ba.w, h, ns);
    }

    template <int STAGE>
    __device__ __forceinline__ void derivs_rk(const RkRows& r, double h, NormSums& ns) const {
        const Rows in = STAGE == 0 ? r.y0 : r.st;
        const Rows fu = r.out.shifted(nq);
        const int4 a0 = ldi4(body_a, 0);
        const bool freeRoot = a0.y == FREE;
        // pass 1: base to tip (stage input formed here)
        m33<K> R = ident<K>(); v3<K> p = zero3<K>(); sv<K> V = svzero<K>();
        {
            // the raw sources two bodies ahead: a body's pass-1 work is
            // shorter than a memory round trip
            Pre1 c1, n1, n2;
            int bstart = 1;
            if (freeRoot) {
                Pre1 c0;
                pre1_any<STAGE, true>(c0, r, a0);
                p1<FREE, STAGE>(c0, r, a0, 1, h, R, p, V);
                bstart = 2;
            }
            int4 ba = ldi4(body_a, bstart-1), bn1 = ba;
            if (bstart <= nb) pre1_any<STAGE, false>(c1, r, ba);
            if (bstart < nb) { bn1 = ldi4(body_a, bstart); pre1_any<STAGE, false>(n1, r, bn1); }
            for (int b = bstart; b <= nb; ++b) {
                int4 bn2 = bn1;
                if (b + 1 < nb) { bn2 = ldi4(body_a, b + 1); pre1_any<STAGE, false>(n2, r, bn2); }
                switch (ba.y) {
                case PIN:       p1<PIN, STAGE>(c1, r, ba, b, h, R, p, V); break;
                case SLIDER:    p1<SLIDER, STAGE>(c1, r, ba, b, h, R, p, V); break;
                case UNIVERSAL: p1<UNIVERSAL, STAGE>(c1, r, ba, b, h, R, p, V); break;
                default:        p1<BALL, STAGE>(c1, r, ba, b, h, R, p, V); break;
                }
                c1 = n1; n1 = n2; ba = bn1; bn1 = bn2;
            }
        }
        // pass 2: tip to base
        Pre<K> cur, nxt;
        abi<K> PP; sv<K> zp = svzero<K>(); v3<K> phic = zero3<K>();
        int4 ba = ldi4(body_a, nb-1);
        const int bstop = freeRoot ? 2 : 1;
        if (nb >= bstop) prefetch<false, false>(cur, in, nq, ba, store(nb, ba, 0), true);
        for (int b = nb; b >= bstop; --b) {
            int4 bn = ba;
            if (b > bstop) { bn = ldi4(body_a, b-2); prefetch<false, false>(nxt, in, nq, bn, store(b-1, bn, 0), true); }
            const KStore<GS, TS> st = store(b, ba, ldi4(body_b, b-1).w);
            const PA P = prm + (size_t)(b-1) * PRM_STRIDE;
            const bool hc = b < nb;
            const sv<K> Fb = F(b);
            switch (ba.y) {
            case PIN:       pass2_body<PIN, K>(P, cur, grav, Fb, st, R, p, V, hc, PP, zp, phic); break;
            case SLIDER:    pass2_body<SLIDER, K>(P, cur, grav, Fb, st, R, p, V, hc, PP, zp, phic); break;
            case UNIVERSAL: pass2_body<UNIVERSAL, K>(P, cur, grav, Fb, st, R, p, V, hc, PP, zp, phic); break;
            default:        pass2_body<BALL, K>(P, cur, grav, Fb, st, R, p, V, hc, PP, zp, phic); break;
            }
            cur = nxt; ba = bn;
        }
        if (freeRoot) {
            Pre<K> c0;
            const KStore<GS, TS> st = store(1, a0, ldi4(body_b, 0).w);
            prefetch<false, true>(c0, in, nq, a0, st, true);
            pass2_body<FREE, K>(prm, c0, grav, F(1), st, R, p, V, nb > 1, PP, zp, phic);
        }
        // pass 3: base to tip
        R = ident<K>(); p = zero3<K>(); V = svzero<K>();
        sv<K> A = svzero<K>();
        int bstart = 1;
        if (freeRoot) {
            Pre<K> c0;
            prefetch<true, true>(c0, in, nq, a0, store(1, a0, ldi4(body_b, 0).w), true);
            p3<FREE, STAGE>(c0, r, fu, a0, 1, h, R, p, V, A, ns);
            bstart = 2;
        }
        if (bstart <= nb) {
            const int4 bs = ldi4(body_a, bstart-1);
            prefetch<true, false>(cur, in, nq, bs, store(bstart, bs, ldi4(body_b, bstart-1).w), true);
        }
        ba = ldi4(body_a, bstart-1);
        for (int b = bstart; b <= nb; ++b) {
            int4 bn = ba;
            if (b < nb) { bn = ldi4(body_a, b); prefetch<true, false>(nxt, in, nq, bn, store(b+1, bn, ldi4(body_b, b).w), true); }
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
