// Spatial algebra for the tree-dynamics kernels (gfx950, HIP).
//
// Small fixed-size types held in registers: 3-vectors, row-major 3x3,
// symmetric 3x3 stored as (xx yy zz xy xz yz), spatial vectors (angular,
// linear) and articulated-body inertias P = [J F; F^T M].  All operators are
// templated on the scalar S so the same code serves the FP64 path and the
// FP32 kinematics of the mixed-precision path.
#pragma once
#include <hip/hip_runtime.h>

namespace tdyn {

#define TD __device__ __forceinline__

template <class S> struct v3 { S x, y, z; };
template <class S> TD v3<S> V3(S a, S b, S c) { return v3<S>{a, b, c}; }
template <class S> TD v3<S> operator+(v3<S> a, v3<S> b) { return {a.x+b.x, a.y+b.y, a.z+b.z}; }
template <class S> TD v3<S> operator-(v3<S> a, v3<S> b) { return {a.x-b.x, a.y-b.y, a.z-b.z}; }
template <class S> TD v3<S> operator-(v3<S> a) { return {-a.x, -a.y, -a.z}; }
template <class S> TD v3<S> operator*(S s, v3<S> a) { return {s*a.x, s*a.y, s*a.z}; }
template <class S> TD S dot(v3<S> a, v3<S> b) { return a.x*b.x + a.y*b.y + a.z*b.z; }
template <class S> TD v3<S> cross(v3<S> a, v3<S> b) {
    return {a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x};
}
template <class S> TD v3<S> zero3() { return {S(0), S(0), S(0)}; }

template <class S> struct m33 { S a[9]; };   // row-major
template <class S> TD m33<S> ident() { m33<S> r; for (int i=0;i<9;++i) r.a[i]=S(0); r.a[0]=r.a[4]=r.a[8]=S(1); return r; }
template <class S> TD v3<S> mul(const m33<S>& A, v3<S> v) {
    return {A.a[0]*v.x + A.a[1]*v.y + A.a[2]*v.z,
            A.a[3]*v.x + A.a[4]*v.y + A.a[5]*v.z,
            A.a[6]*v.x + A.a[7]*v.y + A.a[8]*v.z};
}
template <class S> TD v3<S> mulT(const m33<S>& A, v3<S> v) {
    return {A.a[0]*v.x + A.a[3]*v.y + A.a[6]*v.z,
            A.a[1]*v.x + A.a[4]*v.y + A.a[7]*v.z,
            A.a[2]*v.x + A.a[5]*v.y + A.a[8]*v.z};
}
template <class S> TD m33<S> mul(const m33<S>& A, const m33<S>& B) {
    m33<S> C;
#pragma unroll
    for (int r=0;r<3;++r)
#pragma unroll
        for (int c=0;c<3;++c) C.a[r*3+c] = A.a[r*3]*B.a[c] + A.a[r*3+1]*B.a[3+c] + A.a[r*3+2]*B.a[6+c];
    return C;
}
template <class S> TD v3<S> col(const m33<S>& A, int c) { return {A.a[c], A.a[3+c], A.a[6+c]}; }

// symmetric 3x3: xx yy zz xy xz yz
template <class S> struct s33 { S xx, yy, zz, xy, xz, yz; };
template <class S> TD v3<S> mul(const s33<S>& A, v3<S> v) {
    return {A.xx*v.x + A.xy*v.y + A.xz*v.z,
            A.xy*v.x + A.yy*v.y + A.yz*v.z,
            A.xz*v.x + A.yz*v.y + A.zz*v.z};
}
template <class S> TD s33<S> operator+(const s33<S>& A, const s33<S>& B) {
    return {A.xx+B.xx, A.yy+B.yy, A.zz+B.zz, A.xy+B.xy, A.xz+B.xz, A.yz+B.yz};
}
template <class S> TD s33<S> operator-(const s33<S>& A, const s33<S>& B) {
    return {A.xx-B.xx, A.yy-B.yy, A.zz-B.zz, A.xy-B.xy, A.xz-B.xz, A.yz-B.yz};
}
// R S R^T with S symmetric (reexpress an inertia from B to G).
template <class S> TD s33<S> rotate_sym(const m33<S>& R, const s33<S>& I) {
    // T = R S (full), then lower triangle of T R^T.
    m33<S> T;
#pragma unroll
    for (int r=0;r<3;++r) {
        const S r0 = R.a[r*3], r1 = R.a[r*3+1], r2 = R.a[r*3+2];
        T.a[r*3+0] = r0*I.xx + r1*I.xy + r2*I.xz;
        T.a[r*3+1] = r0*I.xy + r1*I.yy + r2*I.yz;
        T.a[r*3+2] = r0*I.xz + r1*I.yz + r2*I.zz;
    }
    auto e = [&](int r, int c) { return T.a[r*3]*R.a[c*3] + T.a[r*3+1]*R.a[c*3+1] + T.a[r*3+2]*R.a[c*3+2]; };
    return {e(0,0), e(1,1), e(2,2), e(1,0), e(2,0), e(2,1)};
}

template <class S> struct sv { v3<S> w, v; };   // spatial vector (angular, linear)
template <class S> TD sv<S> SV(v3<S> a, v3<S> b) { return sv<S>{a, b}; }
template <class S> TD sv<S> svzero() { return {zero3<S>(), zero3<S>()}; }
template <class S> TD sv<S> operator+(const sv<S>& a, const sv<S>& b) { return {a.w+b.w, a.v+b.v}; }
template <class S> TD sv<S> operator-(const sv<S>& a, const sv<S>& b) { return {a.w-b.w, a.v-b.v}; }
template <class S> TD sv<S> operator*(S s, const sv<S>& a) { return {s*a.w, s*a.v}; }
template <class S> TD S dot(const sv<S>& a, const sv<S>& b) { return dot(a.w, b.w) + dot(a.v, b.v); }
// Phi * F (force shift child->parent), Phi^T * V (motion shift parent->child)
template <class S> TD sv<S> phi_mul(v3<S> l, const sv<S>& F) { return {F.w + cross(l, F.v), F.v}; }
template <class S> TD sv<S> phiT_mul(v3<S> l, const sv<S>& V) { return {V.w, V.v + cross(V.w, l)}; }

// Rigid spatial inertia about Bo in G: mass m, mass-center offset p, unit inertia G.
template <class S> TD sv<S> rbi_mul(S m, v3<S> p, const s33<S>& G, const sv<S>& x) {
    const v3<S> mom = mul(G, x.w) + cross(p, x.v);
    const v3<S> frc = x.v - cross(p, x.w);
    return {m*mom, m*frc};
}

// Articulated-body inertia.
template <class S> struct abi { s33<S> M, J; m33<S> F; };
template <class S> TD abi<S> abi_from_rbi(S m, v3<S> p, const s33<S>& G) {
    abi<S> P;
    P.M = {m, m, m, S(0), S(0), S(0)};
    P.J = {m*G.xx, m*G.yy, m*G.zz, m*G.xy, m*G.xz, m*G.yz};
    const v3<S> c = m*p;
    P.F.a[0]=S(0); P.F.a[1]=-c.z; P.F.a[2]= c.y;
    P.F.a[3]= c.z; P.F.a[4]=S(0); P.F.a[5]=-c.x;
    P.F.a[6]=-c.y; P.F.a[7]= c.x; P.F.a[8]=S(0);
    return P;
}
template <class S> TD sv<S> abi_mul(const abi<S>& P, const sv<S>& x) {
    return {mul(P.J, x.w) + mul(P.F, x.v), mulT(P.F, x.w) + mul(P.M, x.v)};
}
// P += Phi(s) Q Phi(s)^T :  F' = F + [s]x M ; J' = J + [s]x F^T - F' [s]x
template <class S> TD void abi_add_shifted(abi<S>& P, const abi<S>& Q, v3<S> s) {
    // [s]x M (M symmetric)
    const s33<S>& M = Q.M;
    const v3<S> M0 = {M.xx, M.xy, M.xz}, M1 = {M.xy, M.yy, M.yz}, M2 = {M.xz, M.yz, M.zz};  // columns
    m33<S> Fp;
    {
        const v3<S> c0 = cross(s, M0);
        const v3<S> c1 = cross(s, M1), c2 = cross(s, M2);
        Fp.a[0] = Q.F.a[0] + c0.x; Fp.a[1] = Q.F.a[1] + c1.x; Fp.a[2] = Q.F.a[2] + c2.x;
        Fp.a[3] = Q.F.a[3] + c0.y; Fp.a[4] = Q.F.a[4] + c1.y; Fp.a[5] = Q.F.a[5] + c2.y;
        Fp.a[6] = Q.F.a[6] + c0.z; Fp.a[7] = Q.F.a[7] + c1.z; Fp.a[8] = Q.F.a[8] + c2.z;
    }
    // X = [s]x F^T - Fp [s]x ; ([s]x F^T)(i,j) = (s x F_row_j)_i ; (Fp [s]x)(i,j) = (Fp_row_i x s)... computed per entry
    // ([s]x F^T)(i,j) = sum_k sx(i,k) F(j,k) = (s x Frow_j)_i
    const v3<S> F0 = {Q.F.a[0], Q.F.a[1], Q.F.a[2]}, F1 = {Q.F.a[3], Q.F.a[4], Q.F.a[5]}, F2 = {Q.F.a[6], Q.F.a[7], Q.F.a[8]};
    const v3<S> a0 = cross(s, F0), a1 = cross(s, F1), a2 = cross(s, F2);   // columns j of [s]x F^T
    // (Fp [s]x)(i,j) = sum_k Fp(i,k) sx(k,j) = (Fprow_i x s)... : row_i^T [s]x = -(s x row_i)^T = (row_i x s)^T
    const v3<S> P0 = {Fp.a[0], Fp.a[1], Fp.a[2]}, P1 = {Fp.a[3], Fp.a[4], Fp.a[5]}, P2 = {Fp.a[6], Fp.a[7], Fp.a[8]};
    const v3<S> b0 = cross(P0, s), b1 = cross(P1, s), b2 = cross(P2, s);   // rows i of Fp [s]x
    // lower triangle X(i,j), i>=j: a_j[i] - b_i[j]
    const s33<S> X = {a0.x - b0.x, a1.y - b1.y, a2.z - b2.z,
                      a0.y - b1.x, a0.z - b2.x, a1.z - b2.y};
    P.M = P.M + Q.M;
    P.J = P.J + (Q.J + X);
#pragma unroll
    for (int i=0;i<9;++i) P.F.a[i] = P.F.a[i] + Fp.a[i];
}

template <class S> struct xf { m33<S> R; v3<S> p; };
template <class S> TD xf<S> compose(const xf<S>& A, const xf<S>& B) {
    return {mul(A.R, B.R), A.p + mul(A.R, B.p)};
}

// Rotation from a normalized quaternion (w,x,y,z).
template <class S> TD m33<S> quat_rot(S q0, S q1, S q2, S q3) {
    const S q00=q0*q0, q11=q1*q1, q22=q2*q2, q33=q3*q3;
    const S q01=q0*q1, q02=q0*q2, q03=q0*q3, q12=q1*q2, q13=q1*q3, q23=q2*q3;
    const S a = q00-q11, c = q22-q33;
    m33<S> R;
    R.a[0]=q00+q11-q22-q33; R.a[1]=S(2)*(q12-q03); R.a[2]=S(2)*(q13+q02);
    R.a[3]=S(2)*(q12+q03);  R.a[4]=a+c;            R.a[5]=S(2)*(q23-q01);
    R.a[6]=S(2)*(q13-q02);  R.a[7]=S(2)*(q23+q01); R.a[8]=a-c;
    return R;
}
// qdot = N(q) w for an unnormalized quaternion.
template <class S> TD void quat_dot(const S q[4], v3<S> w, S d[4]) {
    const S e0=q[0]*S(0.5), e1=q[1]*S(0.5), e2=q[2]*S(0.5), e3=q[3]*S(0.5);
    d[0] = -e1*w.x - e2*w.y - e3*w.z;
    d[1] =  e0*w.x + e3*w.y - e2*w.z;
    d[2] = -e3*w.x + e0*w.y + e1*w.z;
    d[3] =  e2*w.x - e1*w.y + e0*w.z;
}
// w = NInv(q) qdot (unnormalized form).
template <class S> TD v3<S> quat_dot_to_w(const S q[4], const S d[4]) {
    const S e0=S(2)*q[0], e1=S(2)*q[1], e2=S(2)*q[2], e3=S(2)*q[3];
    return {-e1*d[0] + e0*d[1] - e3*d[2] + e2*d[3],
            -e2*d[0] + e3*d[1] + e0*d[2] - e1*d[3],
            -e3*d[0] - e2*d[1] + e1*d[2] + e0*d[3]};
}

// Body-fixed x-y-z (1-2-3) rotation from the cosines and sines of its three
// angles (Rotation.h:342-349).
template <class S> TD m33<S> rot_xyz(S c0, S s0, S c1, S s1, S c2, S s2) {
    const S s0s1 = s0*s1, s2c0 = s2*c0, c0c2 = c0*c2, nc1 = -c1;
    m33<S> R;
    R.a[0]= c1*c2;             R.a[1]= s2*nc1;            R.a[2]= s1;
    R.a[3]= s2c0 + s0s1*c2;    R.a[4]= c0c2 - s0s1*s2;    R.a[5]= s0*nc1;
    R.a[6]= s0*s2 - s1*c0c2;   R.a[7]= s0*c2 + s1*s2c0;   R.a[8]= c0*c1;
    return R;
}
// Rates of x-y-z angles from the angular velocity in the parent frame,
// qdot = N_P(q) w (Rotation.h:395-406).
template <class S> TD v3<S> euler_dot(S c0, S s0, S c1, S s1, v3<S> w) {
    const S oocosy = S(1)/c1;
    const S t = (s0*w.y - c0*w.z)*oocosy;
    return {w.x + t*s1, c0*w.y + s0*w.z, -t};
}
// Second derivatives of the angles from qdot and the angular acceleration
// b in the parent frame: N_P b + NDot_P w (Rotation.h:1041-1060).
template <class S> TD v3<S> euler_dotdot(S c0, S s0, S c1, S s1, v3<S> qd, v3<S> b) {
    const S oocosy = S(1)/c1;
    const v3<S> Nb = euler_dot(c0, s0, c1, s1, b);
    const S q1oc1 = qd.y*oocosy;
    return Nb + v3<S>{(qd.x*s1 - qd.z)*q1oc1, qd.x*qd.z*c1, (qd.z*s1 - qd.x)*q1oc1};
}

#undef TD
} // namespace tdyn
