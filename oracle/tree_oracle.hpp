// Oracle: scalar FP64 CPU restatement of Simbody's tree-dynamics path.
//
// THIS IS TEST INFRASTRUCTURE. It is the checker that the HIP path of
// simbody_amd is compared against, and the timed CPU baseline.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
// product path never calls into it.
//
// Every routine restates the algorithm of the Simbody 3.9 source file cited
// next to it (paths relative to the Simbody source tree).  It is written from
// the mathematics of those routines, in this file's own data structures; the
// operation order follows the reference where that is cheap to do so that the
// FP64 results agree to rounding.
//
// The code is templated on the scalar type T so that a counting scalar
// (flop_count.hpp) can report the algorithmic flop count of exactly this
// algorithm on exactly the benchmark inputs.
//
// Pinning (what checks this restatement; see tests/test_oracle_*.py):
//   * SD/FAST reaction-force golden vectors of
//     Simbody/tests/TestMobilizerReactionForces.cpp:405-435 (Pin + Slider,
//     gravity) pin position kinematics, spatial inertias, gravity, the
//     articulated-body inertias, the ABA passes and A_GB.
//   * Closed forms: single and double planar pendulum equations of motion,
//     a torque-free body, a Slider under gravity.
//   * Invariants of Simbody/tests/TestMassMatrix.cpp:658-887: M^-1(M v)=v,
//     residual(forward(f))=0, forward(u=0) = M^-1 f, symmetric positive M.
//   * Finite differences of the position kinematics pin V_GB and H; finite
//     differences of V_GB pin A_GB (coriolis and HDot terms).
//   * Energy conservation (TestMobilizedBody.cpp:615-662 style) and momentum
//     conservation of free-floating trees pin the velocity-dependent terms and
//     the 6x6 Free-joint D inverse.
//   * Reversed mobilizers: the same physical system built forward and with
//     the mobilizer reversed (Simbody/tests/TestReverseMobilizers.cpp
//     style) moves the same way.
//   * RK-Merson: the linear-test-equation stability polynomial and the
//     embedded error estimate (RungeKuttaMersonIntegrator.cpp:60-80 tableau).
//   Parity unpinned (no reference-held vector, reference not buildable here):
//   the exact adaptive step-size sequence of AbstractIntegratorRep
//   (:448-502) and the quaternion projection cadence are restated from the
//   source but only pinned by the accuracy/energy properties above.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>

namespace oracle {

using std::sqrt; using std::sin; using std::cos; using std::fabs; using std::pow;

// ---------------------------------------------------------------------------
// Small fixed-size algebra.
// ---------------------------------------------------------------------------
template <class T> struct V3 {
    T x[3];
    T& operator[](int i) { return x[i]; }
    const T& operator[](int i) const { return x[i]; }
};
template <class T> inline V3<T> mk3(T a, T b, T c) { V3<T> r; r[0]=a; r[1]=b; r[2]=c; return r; }
template <class T> inline V3<T> zero3() { return mk3<T>(T(0),T(0),T(0)); }
template <class T> inline V3<T> operator+(const V3<T>& a, const V3<T>& b) { return mk3<T>(a[0]+b[0], a[1]+b[1], a[2]+b[2]); }
template <class T> inline V3<T> operator-(const V3<T>& a, const V3<T>& b) { return mk3<T>(a[0]-b[0], a[1]-b[1], a[2]-b[2]); }
template <class T> inline V3<T> operator-(const V3<T>& a) { return mk3<T>(-a[0], -a[1], -a[2]); }
template <class T> inline V3<T> operator*(const T& s, const V3<T>& a) { return mk3<T>(s*a[0], s*a[1], s*a[2]); }
template <class T> inline T dot(const V3<T>& a, const V3<T>& b) { return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]; }
template <class T> inline V3<T> cross(const V3<T>& a, const V3<T>& b) {
    return mk3<T>(a[1]*b[2] - a[2]*b[1], a[2]*b[0] - a[0]*b[2], a[0]*b[1] - a[1]*b[0]);
}

// Row-major 3x3.
template <class T> struct M33 {
    T m[9];
    T& operator()(int r, int c) { return m[r*3+c]; }
    const T& operator()(int r, int c) const { return m[r*3+c]; }
};
template <class T> inline M33<T> ident33() { M33<T> r; for (int i=0;i<9;++i) r.m[i]=T(0); r.m[0]=r.m[4]=r.m[8]=T(1); return r; }
template <class T> inline M33<T> zero33() { M33<T> r; for (int i=0;i<9;++i) r.m[i]=T(0); return r; }
template <class T> inline V3<T> operator*(const M33<T>& A, const V3<T>& v) {
    return mk3<T>(A(0,0)*v[0] + A(0,1)*v[1] + A(0,2)*v[2],
                  A(1,0)*v[0] + A(1,1)*v[1] + A(1,2)*v[2],
                  A(2,0)*v[0] + A(2,1)*v[1] + A(2,2)*v[2]);
}
template <class T> inline V3<T> mulT(const M33<T>& A, const V3<T>& v) {   // A^T v
    return mk3<T>(A(0,0)*v[0] + A(1,0)*v[1] + A(2,0)*v[2],
                  A(0,1)*v[0] + A(1,1)*v[1] + A(2,1)*v[2],
                  A(0,2)*v[0] + A(1,2)*v[1] + A(2,2)*v[2]);
}
template <class T> inline M33<T> operator*(const M33<T>& A, const M33<T>& B) {
    M33<T> C;
    for (int r=0;r<3;++r) for (int c=0;c<3;++c)
        C(r,c) = A(r,0)*B(0,c) + A(r,1)*B(1,c) + A(r,2)*B(2,c);
    return C;
}
template <class T> inline M33<T> mulABt(const M33<T>& A, const M33<T>& B) {  // A B^T
    M33<T> C;
    for (int r=0;r<3;++r) for (int c=0;c<3;++c)
        C(r,c) = A(r,0)*B(c,0) + A(r,1)*B(c,1) + A(r,2)*B(c,2);
    return C;
}
template <class T> inline M33<T> mulAtB(const M33<T>& A, const M33<T>& B) {  // A^T B
    M33<T> C;
    for (int r=0;r<3;++r) for (int c=0;c<3;++c)
        C(r,c) = A(0,r)*B(0,c) + A(1,r)*B(1,c) + A(2,r)*B(2,c);
    return C;
}
template <class T> inline M33<T> operator+(const M33<T>& A, const M33<T>& B) { M33<T> C; for (int i=0;i<9;++i) C.m[i]=A.m[i]+B.m[i]; return C; }
template <class T> inline M33<T> operator-(const M33<T>& A, const M33<T>& B) { M33<T> C; for (int i=0;i<9;++i) C.m[i]=A.m[i]-B.m[i]; return C; }
template <class T> inline M33<T> transpose(const M33<T>& A) { M33<T> C; for (int r=0;r<3;++r) for (int c=0;c<3;++c) C(r,c)=A(c,r); return C; }
template <class T> inline M33<T> crossMat(const V3<T>& v) {
    M33<T> C;
    C(0,0)=T(0);  C(0,1)=-v[2]; C(0,2)= v[1];
    C(1,0)= v[2]; C(1,1)=T(0);  C(1,2)=-v[0];
    C(2,0)=-v[1]; C(2,1)= v[0]; C(2,2)=T(0);
    return C;
}
// Keep the lower triangle and mirror it: the reference stores SymMat33 as its
// lower triangle only (SimTKcommon SymMat<3>).
template <class T> inline M33<T> symFromLower(const M33<T>& A) {
    M33<T> C = A; C(0,1)=A(1,0); C(0,2)=A(2,0); C(1,2)=A(2,1); return C;
}

template <class T> struct SV { V3<T> a, b; };   // SpatialVec: (angular, linear)
template <class T> inline SV<T> sv(const V3<T>& a, const V3<T>& b) { SV<T> r; r.a=a; r.b=b; return r; }
template <class T> inline SV<T> svzero() { return sv(zero3<T>(), zero3<T>()); }
template <class T> inline SV<T> operator+(const SV<T>& x, const SV<T>& y) { return sv(x.a+y.a, x.b+y.b); }
template <class T> inline SV<T> operator-(const SV<T>& x, const SV<T>& y) { return sv(x.a-y.a, x.b-y.b); }
template <class T> inline T dot(const SV<T>& x, const SV<T>& y) { return dot(x.a,y.a) + dot(x.b,y.b); }

template <class T> struct Xform { M33<T> R; V3<T> p; };
template <class T> inline Xform<T> compose(const Xform<T>& A, const Xform<T>& B) {
    Xform<T> C; C.R = A.R*B.R; C.p = A.p + A.R*B.p; return C;
}
template <class T> inline Xform<T> invert(const Xform<T>& A) {
    Xform<T> C; C.R = transpose(A.R); C.p = -(C.R*A.p); return C;
}

// PhiMatrix action (SimTKcommon/Mechanics/.../SpatialAlgebra.h:685-788):
//   Phi  * F = (F.a + l x F.b, F.b)      (child-to-parent force shift)
//   Phi^T* V = (V.a, V.b + V.a x l)      (parent-to-child motion shift)
template <class T> inline SV<T> phiMul(const V3<T>& l, const SV<T>& F) { return sv(F.a + cross(l, F.b), F.b); }
template <class T> inline SV<T> phiTMul(const V3<T>& l, const SV<T>& V) { return sv(V.a, V.b + cross(V.a, l)); }

// Rigid-body spatial inertia about the body origin (MassProperties.h:992-1116):
// mass m, mass-center offset p, unit inertia G about the origin.
template <class T> struct SpatialInertia { T m; V3<T> p; M33<T> G; };
template <class T> inline SV<T> siMul(const SpatialInertia<T>& I, const SV<T>& v) {
    const V3<T> mom = I.G*v.a + cross(I.p, v.b);
    const V3<T> frc = v.b - cross(I.p, v.a);
    return sv(I.m*mom, I.m*frc);
}

// Articulated-body inertia P = [J F; F^T M] (MassProperties.h:1231-1331).
template <class T> struct ABI { M33<T> M, J, F; };
template <class T> inline ABI<T> abiFromRigid(const SpatialInertia<T>& I) {
    ABI<T> P;
    P.M = zero33<T>(); P.M(0,0)=P.M(1,1)=P.M(2,2)=I.m;
    for (int i=0;i<9;++i) P.J.m[i] = I.m*I.G.m[i];
    P.F = crossMat(I.m*I.p);
    return P;
}
template <class T> inline SV<T> abiMul(const ABI<T>& P, const SV<T>& v) {
    return sv(P.J*v.a + P.F*v.b, mulT(P.F, v.a) + P.M*v.b);
}
// Articulated shift across Phi(s): P' = Phi P Phi^T
// (SimTKcommon/Mechanics/src/MassProperties.cpp:100-127):
//   F' = F + [s]x M,  J' = J + [s]x F^T - F' [s]x  (symmetric), M' = M.
template <class T> inline ABI<T> abiShift(const ABI<T>& P, const V3<T>& s) {
    const M33<T> sx = crossMat(s);
    ABI<T> Q;
    Q.M = P.M;
    Q.F = P.F + sx*P.M;
    const M33<T> X = mulABt(sx, P.F) - Q.F*sx;   // [s]x F^T - F' [s]x
    Q.J = P.J + symFromLower(X);
    return Q;
}
template <class T> inline void abiAdd(ABI<T>& P, const ABI<T>& Q) {
    P.M = P.M + Q.M; P.J = P.J + Q.J; P.F = P.F + Q.F;
}

// ---------------------------------------------------------------------------
// Model.
// ---------------------------------------------------------------------------
enum Mobilizer { PIN=0, SLIDER=1, UNIVERSAL=2, BALL=3, FREE=4, TRANSLATION=5, PLANAR=6, CYLINDER=7, SCREW=8,
                 GIMBAL=9, WELD=10, BENDSTRETCH=11, BUSHING=12 };
inline int mobDof(int t) {
    static const int d[13] = {1, 1, 2, 3, 6, 3, 3, 2, 1, 3, 0, 2, 6};
    return d[t];
}
inline int mobNq(int t) {
    static const int n[13] = {1, 1, 2, 4, 7, 3, 3, 2, 1, 3, 0, 2, 6};
    return n[t];
}

// Plain-array tree description; identical to tdyn_tree_desc of include/tdyn.h.
struct TreeDesc {
    int32_t nb;
    const int32_t* parent;      // [nb], body numbers 1..nb, 0 = Ground, parent < body
    const int32_t* mobilizer;   // [nb]
    const double*  mass;        // [nb]
    const double*  com_B;       // [nb][3]
    const double*  inertia_B;   // [nb][6] about Bo in B: xx yy zz xy xz yz
    const double*  X_PF;        // [nb][12] row-major R then p
    const double*  X_BM;        // [nb][12]
    double gravity[3];
    const double*  pitch;       // [nb] Screw pitch (length per radian); nullable
    const int32_t* reversed;    // [nb] nonzero: reversed mobilizer; nullable
    int32_t use_euler_angles;   // Ball and Free use body-fixed x-y-z angles, not quaternions
};

struct Model {
    int nb = 0, nq = 0, nu = 0, nquat = 0;
    std::vector<int> parent, type, dof, qIndex, uIndex, level;
    std::vector<std::vector<int>> children;   // in body-number order
    std::vector<std::vector<int>> levels;     // rbNodeLevels, body numbers per level
    std::vector<double> mass;
    std::vector<V3<double>> com;
    std::vector<M33<double>> Gunit;           // unit inertia about Bo in B
    std::vector<Xform<double>> X_PF, X_MB;
    std::vector<V3<double>> p_BM;
    std::vector<double> pitch;
    std::vector<int> reversed;
    bool euler = false;                       // SimbodyMatterSubsystem::setUseEulerAngles
    V3<double> g;
    // Ball and Free carry a quaternion (four q slots either way; in Euler
    // mode the last slot of the four is unused: RigidBodyNodeSpec_Ball.h:
    // 183-214, _Free.h:222-257)
    bool isQuat(int b) const { return !euler && (type[b] == BALL || type[b] == FREE); }

    // Build from the plain description.  Body 0 is Ground.
    // Slot assignment follows SimbodyMatterSubsystemRep.cpp:256-293 (q and u
    // slots in body-number order, maxNQ per mobilizer).
    bool build(const TreeDesc& d) {
        nb = d.nb;
        if (nb <= 0) return false;
        const int n = nb + 1;
        parent.assign(n, -1); type.assign(n, -1); dof.assign(n, 0);
        qIndex.assign(n, 0); uIndex.assign(n, 0); level.assign(n, 0);
        children.assign(n, {}); mass.assign(n, 0.0);
        com.assign(n, zero3<double>()); Gunit.assign(n, zero33<double>());
        X_PF.assign(n, Xform<double>{ident33<double>(), zero3<double>()});
        X_MB = X_PF; p_BM.assign(n, zero3<double>()); pitch.assign(n, 0.0); reversed.assign(n, 0);
        nq = nu = nquat = 0;
        euler = d.use_euler_angles != 0;
        for (int b = 1; b <= nb; ++b) {
            const int i = b-1;
            const int p = d.parent[i];
            if (p < 0 || p >= b) return false;
            const int t = d.mobilizer[i];
            if (t < PIN || t > BUSHING) return false;
            if (t == BUSHING && p != 0) return false;
            parent[b] = p; type[b] = t; dof[b] = mobDof(t);
            level[b] = level[p] + 1;
            children[p].push_back(b);
            qIndex[b] = nq; uIndex[b] = nu;
            nq += mobNq(t); nu += dof[b];
            if ((t == BALL || t == FREE) && !euler) ++nquat;
            mass[b] = d.mass[i];
            if (!(mass[b] > 0)) return false;
            com[b] = mk3(d.com_B[3*i], d.com_B[3*i+1], d.com_B[3*i+2]);
            const double* I = d.inertia_B + 6*i;
            M33<double> Io;
            Io(0,0)=I[0]; Io(1,1)=I[1]; Io(2,2)=I[2];
            Io(0,1)=Io(1,0)=I[3]; Io(0,2)=Io(2,0)=I[4]; Io(1,2)=Io(2,1)=I[5];
            for (int k=0;k<9;++k) Gunit[b].m[k] = Io.m[k]/mass[b];
            for (int k=0;k<9;++k) X_PF[b].R.m[k] = d.X_PF[12*i+k];
            X_PF[b].p = mk3(d.X_PF[12*i+9], d.X_PF[12*i+10], d.X_PF[12*i+11]);
            Xform<double> X_BM;
            for (int k=0;k<9;++k) X_BM.R.m[k] = d.X_BM[12*i+k];
            X_BM.p = mk3(d.X_BM[12*i+9], d.X_BM[12*i+10], d.X_BM[12*i+11]);
            p_BM[b] = X_BM.p;
            pitch[b] = d.pitch ? d.pitch[i] : 0.0;
            reversed[b] = d.reversed ? (d.reversed[i] != 0) : 0;
            if (reversed[b] && (t == FREE || t == BENDSTRETCH || t == BUSHING)) return false;
            X_MB[b] = invert(X_BM);      // RigidBodyNode.h:1012 X_MB(~X_BM)
        }
        int maxLevel = 0;
        for (int b = 1; b <= nb; ++b) maxLevel = std::max(maxLevel, level[b]);
        levels.assign(maxLevel+1, {});
        for (int b = 0; b <= nb; ++b) levels[level[b]].push_back(b);
        g = mk3(d.gravity[0], d.gravity[1], d.gravity[2]);
        return true;
    }
};

// ---------------------------------------------------------------------------
// Per-body cache: the union of SBTreePositionCache, SBTreeVelocityCache,
// SBArticulatedBodyInertiaCache and SBTreeAccelerationCache entries
// (Simbody/src/SimbodyTreeState.h:648-1162) that the tree path needs.
// ---------------------------------------------------------------------------
template <class T> struct Body {
    // position
    T cq[3], sq[3];            // mobilizer q pool (sin/cos)
    T qerr;                    // |quaternion| - 1
    Xform<T> X_FM, X_PB, X_GB;
    V3<T> phi;                 // p_PB_G
    SV<T> H_FM[6], H[6];       // columns, H = H_PB_G
    SV<T> H_MF[6];             // reversed mobilizer: the mobilizer's own H
    SpatialInertia<T> Mk;      // about Bo, in G
    V3<T> com_G;
    // velocity
    SV<T> V_FM, V_PB_G, V_GB, VD_PB_G;
    SV<T> HDot_FM[6], HDot[6];
    SV<T> b, a, atot, MAb;     // gyroscopic, mobilizer coriolis, total coriolis, Mk*atot+b
    // articulated
    ABI<T> P, PPlus;
    T D[36], DI[36];
    SV<T> G[6], PAb;           // G = P H DI (columns), PAb = P a + b
    // acceleration / temporaries
    SV<T> z, zPlus, A_GB, F;
    T eps[6];
};

template <class T> struct Tree {
    const Model* M = nullptr;
    std::vector<Body<T>> B;
    explicit Tree(const Model& m) : M(&m), B(m.nb+1) { resetGround(); }
    void resetGround() {
        Body<T>& g = B[0];
        g.X_GB.R = ident33<T>(); g.X_GB.p = zero3<T>();
        g.V_GB = svzero<T>(); g.atot = svzero<T>(); g.A_GB = svzero<T>();
    }
};

template <class T> inline Xform<T> castX(const Xform<double>& X) {
    Xform<T> r; for (int i=0;i<9;++i) r.R.m[i]=T(X.R.m[i]); for (int i=0;i<3;++i) r.p[i]=T(X.p[i]); return r;
}
template <class T> inline V3<T> cast3(const V3<double>& v) { return mk3<T>(T(v[0]),T(v[1]),T(v[2])); }

// Rotation from a (normalized) quaternion, Rotation.cpp:600-610.
template <class T> inline M33<T> quatToRot(const T q[4]) {
    const T q00=q[0]*q[0], q11=q[1]*q[1], q22=q[2]*q[2], q33=q[3]*q[3];
    const T q01=q[0]*q[1], q02=q[0]*q[2], q03=q[0]*q[3];
    const T q12=q[1]*q[2], q13=q[1]*q[3], q23=q[2]*q[3];
    const T a = q00-q11, c = q22-q33;
    M33<T> R;
    R(0,0)=q00+q11-q22-q33; R(0,1)=T(2)*(q12-q03); R(0,2)=T(2)*(q13+q02);
    R(1,0)=T(2)*(q12+q03);  R(1,1)=a+c;            R(1,2)=T(2)*(q23-q01);
    R(2,0)=T(2)*(q13-q02);  R(2,1)=T(2)*(q23+q01); R(2,2)=a-c;
    return R;
}
// qdot = N(q) w for an unnormalized quaternion q (Rotation.h:712-719, 990-992).
template <class T> inline void quatDot(const T q[4], const V3<T>& w, T qd[4]) {
    const T e0=q[0]/T(2), e1=q[1]/T(2), e2=q[2]/T(2), e3=q[3]/T(2);
    qd[0] = -e1*w[0] - e2*w[1] - e3*w[2];
    qd[1] =  e0*w[0] + e3*w[1] - e2*w[2];
    qd[2] = -e3*w[0] + e0*w[1] + e1*w[2];
    qd[3] =  e2*w[0] - e1*w[1] + e0*w[2];
}
// u = NInv(q) qdot, unnormalized form (Rotation.h:742-748).
template <class T> inline V3<T> quatDotToW(const T q[4], const T qd[4]) {
    const T e0=T(2)*q[0], e1=T(2)*q[1], e2=T(2)*q[2], e3=T(2)*q[3];
    return mk3<T>(-e1*qd[0] + e0*qd[1] - e3*qd[2] + e2*qd[3],
                  -e2*qd[0] + e3*qd[1] + e0*qd[2] - e1*qd[3],
                  -e3*qd[0] - e2*qd[1] + e1*qd[2] + e0*qd[3]);
}
// Body-fixed 1-2-3 angle rates from the angular velocity in the parent,
// qdot = N_P(q) w (Rotation.h:395-406): c, s = cos, sin of the angles.
template <class T> inline V3<T> eulerDot(const T* c, const T* s, const V3<T>& w) {
    const T oocosy = T(1)/c[1];
    const T t = (s[0]*w[1] - c[0]*w[2])*oocosy;
    return mk3<T>(w[0] + t*s[1], c[0]*w[1] + s[0]*w[2], -t);
}
// qdotdot = N_P b + NDot_P w, written with qdot (Rotation.h:1041-1060).
template <class T> inline V3<T> eulerDotDot(const T* c, const T* s, const V3<T>& qd, const V3<T>& b) {
    const T oocosy = T(1)/c[1];
    const V3<T> Nb = eulerDot(c, s, b);
    const T q1oc1 = qd[1]*oocosy;
    return Nb + mk3<T>((qd[0]*s[1] - qd[2])*q1oc1, qd[0]*qd[2]*c[1], (qd[2]*s[1] - qd[0])*q1oc1);
}
// qdotdot = N b + NDot w = N b - |w|^2/4 q  (Rotation.h:1006-1013).
template <class T> inline void quatDotDot(const T q[4], const V3<T>& w, const V3<T>& bdot, T qdd[4]) {
    T Nb[4]; quatDot(q, bdot, Nb);
    const T s = T(-0.25)*dot(w,w);
    for (int i=0;i<4;++i) qdd[i] = Nb[i] + s*q[i];
}

// ---------------------------------------------------------------------------
// Position kinematics, base to tip.
// RigidBodyNodeSpec.h:229-288 (realizePosition) with the mobilizer-specific
// performQPrecalculations/calcX_FM/calcAcrossJointVelocityJacobian of
// RigidBodyNodeSpec_{Pin,Slider,Universal,Ball,Free,Gimbal,PolarCoords}.h, RigidBodyNode_Weld.cpp,
// calcBodyTransforms
// (RigidBodyNodeSpec.h:554-569), calcParentToChildVelocityJacobianInGround
// (RigidBodyNodeSpec.cpp:44-74) and calcJointIndependentKinematicsPos
// (RigidBodyNode.cpp:54-84).  Driver: SimbodyMatterSubsystemRep.cpp:1112-1144.
// ---------------------------------------------------------------------------
template <class T> void realizePosition(Tree<T>& tr, const T* q) {
    const Model& M = *tr.M;
    for (size_t L = 1; L < M.levels.size(); ++L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        const Body<T>& Pb = tr.B[M.parent[b]];
        const T* qb = q + M.qIndex[b];
        const int t = M.type[b], nd = M.dof[b];
        B.qerr = T(0);
        // q precalculations and X_FM
        B.X_FM.p = zero3<T>();
        switch (t) {
        case PIN: {
            B.cq[0] = cos(qb[0]); B.sq[0] = sin(qb[0]);
            M33<T> R = zero33<T>();
            R(0,0)=B.cq[0]; R(0,1)=-B.sq[0]; R(1,0)=B.sq[0]; R(1,1)=B.cq[0]; R(2,2)=T(1);
            B.X_FM.R = R;
            B.H_FM[0] = sv(mk3<T>(T(0),T(0),T(1)), zero3<T>());
        } break;
        case SLIDER: {
            B.X_FM.R = ident33<T>();
            B.X_FM.p = mk3<T>(qb[0], T(0), T(0));
            B.H_FM[0] = sv(zero3<T>(), mk3<T>(T(1),T(0),T(0)));
        } break;
        case UNIVERSAL: {
            // body-fixed X then Y sequence, Rotation.cpp:241-263
            B.cq[0] = cos(qb[0]); B.cq[1] = cos(qb[1]);
            B.sq[0] = sin(qb[0]); B.sq[1] = sin(qb[1]);
            const T c1=B.cq[0], s1=B.sq[0], c2=B.cq[1], s2=B.sq[1];
            M33<T> R;
            R(0,0)= c2;    R(0,1)= T(0); R(0,2)= s2;
            R(1,0)= s2*s1; R(1,1)= c1;   R(1,2)= -s1*c2;
            R(2,0)=-s2*c1; R(2,1)= s1;   R(2,2)= c1*c2;
            B.X_FM.R = R;
            B.H_FM[0] = sv(mk3<T>(T(1),T(0),T(0)), zero3<T>());
            B.H_FM[1] = sv(mk3<T>(R(0,1),R(1,1),R(2,1)), zero3<T>());   // R_FM.y()
        } break;
        // Translation, Planar, Cylinder, Screw: H_FM constant in F, so their
        // HDot_FM is zero (RigidBodyNodeSpec_Translation.h, _Planar.h:115-156,
        // _Cylinder.h:100-139, _Screw.h:109-148).
        case TRANSLATION: {
            B.X_FM.R = ident33<T>();
            B.X_FM.p = mk3<T>(qb[0], qb[1], qb[2]);
            for (int j=0;j<3;++j) { V3<T> e = zero3<T>(); e[j] = T(1); B.H_FM[j] = sv(zero3<T>(), e); }
        } break;
        case PLANAR: case CYLINDER: case SCREW: {
            B.cq[0] = cos(qb[0]); B.sq[0] = sin(qb[0]);
            M33<T> R = zero33<T>();
            R(0,0)=B.cq[0]; R(0,1)=-B.sq[0]; R(1,0)=B.sq[0]; R(1,1)=B.cq[0]; R(2,2)=T(1);
            B.X_FM.R = R;
            const V3<T> ez = mk3<T>(T(0),T(0),T(1));
            if (t == PLANAR) {
                B.X_FM.p = mk3<T>(qb[1], qb[2], T(0));
                B.H_FM[0] = sv(ez, zero3<T>());
                B.H_FM[1] = sv(zero3<T>(), mk3<T>(T(1),T(0),T(0)));
                B.H_FM[2] = sv(zero3<T>(), mk3<T>(T(0),T(1),T(0)));
            } else if (t == CYLINDER) {
                B.X_FM.p = mk3<T>(T(0), T(0), qb[1]);
                B.H_FM[0] = sv(ez, zero3<T>());
                B.H_FM[1] = sv(zero3<T>(), ez);
            } else {
                const T pt = T(M.pitch[b]);
                B.X_FM.p = mk3<T>(T(0), T(0), qb[0]*pt);
                B.H_FM[0] = sv(ez, mk3<T>(T(0), T(0), pt));
            }
        } break;
        case BENDSTRETCH: {
            // RigidBodyNodeSpec_PolarCoords.h:128-172: rotation q0 about z,
            // then a slide q1 along the rotated x; H_FM(0) = (z, z x p_FM),
            // H_FM(1) = (0, M's x axis in F)
            B.cq[0] = cos(qb[0]); B.sq[0] = sin(qb[0]);
            M33<T> R = zero33<T>();
            R(0,0)=B.cq[0]; R(0,1)=-B.sq[0]; R(1,0)=B.sq[0]; R(1,1)=B.cq[0]; R(2,2)=T(1);
            B.X_FM.R = R;
            B.X_FM.p = R*mk3<T>(qb[1], T(0), T(0));
            const V3<T> ez = mk3<T>(T(0),T(0),T(1));
            B.H_FM[0] = sv(ez, cross(ez, B.X_FM.p));
            B.H_FM[1] = sv(zero3<T>(), mk3<T>(R(0,0), R(1,0), R(2,0)));
        } break;
        case WELD:
            // no mobilities: X_FM = identity (RigidBodyNode_Weld.cpp:389-423);
            // with no H column the ABI passes give P+ = P and z+ = z
            B.X_FM.R = ident33<T>();
            break;
        case GIMBAL: case BUSHING: {
            // body-fixed 1-2-3 angles (RigidBodyNodeSpec_Gimbal.h:141-176,
            // Rotation.h:342-349); H_FM maps qdot to w_FM.
            for (int j=0;j<3;++j) { B.cq[j] = cos(qb[j]); B.sq[j] = sin(qb[j]); }
            const T c0=B.cq[0], c1=B.cq[1], c2=B.cq[2], s0=B.sq[0], s1=B.sq[1], s2=B.sq[2];
            const T s0s1 = s0*s1, s2c0 = s2*c0, c0c2 = c0*c2, nc1 = -c1;
            M33<T> R;
            R(0,0)= c1*c2;              R(0,1)= s2*nc1;             R(0,2)= s1;
            R(1,0)= s2c0 + s0s1*c2;     R(1,1)= c0c2 - s0s1*s2;     R(1,2)= s0*nc1;
            R(2,0)= s0*s2 - s1*c0c2;    R(2,1)= s0*c2 + s1*s2c0;    R(2,2)= c0*c1;
            B.X_FM.R = R;
            B.H_FM[0] = sv(mk3<T>(T(1), T(0), T(0)), zero3<T>());
            B.H_FM[1] = sv(mk3<T>(T(0), c0, s0), zero3<T>());
            B.H_FM[2] = sv(mk3<T>(s1, -s0*c1, c0*c1), zero3<T>());
            if (t == BUSHING) {
                // Bushing (RigidBodyNodeSpec_Bushing.h:152-191): the Gimbal's
                // rotation and angle columns, then p_FM = (q3, q4, q5) with
                // identity translation columns; u = qdot
                B.X_FM.p = mk3<T>(qb[3], qb[4], qb[5]);
                for (int j=0;j<3;++j) { V3<T> e = zero3<T>(); e[j] = T(1); B.H_FM[3+j] = sv(zero3<T>(), e); }
            }
        } break;
        case BALL: case FREE: if (M.euler) {
            // Euler-angle mode (RigidBodyNodeSpec_Ball.h:113-158, _Free.h:
            // 139-188): body-fixed 1-2-3 angles in q0..q2, Free translation
            // in q3..q5
            for (int j=0;j<3;++j) { B.cq[j] = cos(qb[j]); B.sq[j] = sin(qb[j]); }
            const T c0=B.cq[0], c1=B.cq[1], c2=B.cq[2], s0=B.sq[0], s1=B.sq[1], s2=B.sq[2];
            const T s0s1 = s0*s1, s2c0 = s2*c0, c0c2 = c0*c2, nc1 = -c1;
            M33<T> R;
            R(0,0)= c1*c2;              R(0,1)= s2*nc1;             R(0,2)= s1;
            R(1,0)= s2c0 + s0s1*c2;     R(1,1)= c0c2 - s0s1*s2;     R(1,2)= s0*nc1;
            R(2,0)= s0*s2 - s1*c0c2;    R(2,1)= s0*c2 + s1*s2c0;    R(2,2)= c0*c1;
            B.X_FM.R = R;
            for (int j=0;j<3;++j) { V3<T> e = zero3<T>(); e[j] = T(1); B.H_FM[j] = sv(e, zero3<T>()); }
            if (t == FREE) {
                B.X_FM.p = mk3<T>(qb[3], qb[4], qb[5]);
                for (int j=0;j<3;++j) { V3<T> e = zero3<T>(); e[j] = T(1); B.H_FM[3+j] = sv(zero3<T>(), e); }
            }
        } else {
            const T len = sqrt(qb[0]*qb[0] + qb[1]*qb[1] + qb[2]*qb[2] + qb[3]*qb[3]);
            B.qerr = len - T(1);
            const T oo = T(1)/len;
            T qn[4] = { qb[0]*oo, qb[1]*oo, qb[2]*oo, qb[3]*oo };
            B.X_FM.R = quatToRot(qn);
            for (int j=0;j<3;++j) {
                V3<T> e = zero3<T>(); e[j] = T(1);
                B.H_FM[j] = sv(e, zero3<T>());
            }
            if (t == FREE) {
                B.X_FM.p = mk3<T>(qb[4], qb[5], qb[6]);
                for (int j=0;j<3;++j) {
                    V3<T> e = zero3<T>(); e[j] = T(1);
                    B.H_FM[3+j] = sv(zero3<T>(), e);
                }
            }
        } break;
        }
        if (M.reversed[b]) {
            // Reversed mobilizer (RigidBodyNodeSpec.h:262-277): the mobilizer's
            // own transform and H are X_MF and H_MF; X_FM = ~X_MF and
            // calcReverseMobilizerH_FM (RigidBodyNodeSpec.cpp:131-166):
            //   H_FM_w = -R_FM H_MF_w,  H_FM_v = ~[p_FM]x H_FM_w - R_FM H_MF_v
            B.X_FM = invert(B.X_FM);
            const M33<T>& R = B.X_FM.R;
            const M33<T> np = transpose(crossMat(B.X_FM.p));
            for (int j = 0; j < nd; ++j) {
                B.H_MF[j] = B.H_FM[j];
                const V3<T> w = -(R*B.H_MF[j].a);
                B.H_FM[j] = sv(w, np*w - R*B.H_MF[j].b);
            }
        }
        // X_PB = X_PF X_FM X_MB ; X_GB = X_GP X_PB
        const Xform<T> X_PF = castX<T>(M.X_PF[b]);
        const Xform<T> X_MB = castX<T>(M.X_MB[b]);
        B.X_PB = compose(X_PF, compose(B.X_FM, X_MB));
        B.X_GB = compose(Pb.X_GB, B.X_PB);
        // H_PB_G = R_GF (H_FM + H_MB_F), H_MB_F = (0, -r_MB_F x H_FM.a)
        const M33<T> R_GF = Pb.X_GB.R * X_PF.R;
        const V3<T> r_MB_F = B.X_FM.R * X_MB.p;
        for (int j = 0; j < nd; ++j) {
            const V3<T> lin = B.H_FM[j].b - cross(r_MB_F, B.H_FM[j].a);
            B.H[j] = sv(R_GF*B.H_FM[j].a, R_GF*lin);
        }
        // joint-independent part
        B.phi = Pb.X_GB.R * B.X_PB.p;
        const M33<T>& R_GB = B.X_GB.R;
        M33<T> Gb; for (int k=0;k<9;++k) Gb.m[k] = T(M.Gunit[b].m[k]);
        B.Mk.m = T(M.mass[b]);
        B.Mk.G = symFromLower(mulABt(R_GB*Gb, R_GB));    // R G R^T
        B.Mk.p = R_GB*cast3<T>(M.com[b]);
        B.com_G = B.X_GB.p + B.Mk.p;
    }
}

// ---------------------------------------------------------------------------
// Velocity kinematics, base to tip.  RigidBodyNodeSpec.h:305-333,
// calcAcrossJointVelocityJacobianDot (_Universal.h:173-189, _Gimbal.h:180-202;
// zero otherwise),
// calcParentToChildVelocityJacobianInGroundDot (RigidBodyNodeSpec.cpp:82-129),
// calcJointIndependentKinematicsVel (RigidBodyNode.cpp:98-174), calcQDot
// (RigidBodyNodeSpec.h:129-134, _Ball.h:187-214, _Free.h:226-257).
// ---------------------------------------------------------------------------
template <class T> void realizeVelocity(Tree<T>& tr, const T* q, const T* u, T* qdot) {
    const Model& M = *tr.M;
    for (size_t L = 1; L < M.levels.size(); ++L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        const Body<T>& Pb = tr.B[M.parent[b]];
        const T* ub = u + M.uIndex[b];
        const T* qb = q + M.qIndex[b];
        const int t = M.type[b], nd = M.dof[b];
        if (qdot) {
            T* qd = qdot + M.qIndex[b];
            if (M.isQuat(b)) {
                quatDot(qb, mk3<T>(ub[0],ub[1],ub[2]), qd);
                if (t == FREE) { qd[4]=ub[3]; qd[5]=ub[4]; qd[6]=ub[5]; }
            } else if (t == BALL || t == FREE) {
                // Euler-angle rates qdot = N(q) w_FM (Rotation.h:395-406,
                // 1021-1028); the unused fourth slot's rate is zero
                const V3<T> ea = eulerDot(B.cq, B.sq, mk3<T>(ub[0],ub[1],ub[2]));
                qd[0]=ea[0]; qd[1]=ea[1]; qd[2]=ea[2];
                if (t == FREE) { qd[3]=ub[3]; qd[4]=ub[4]; qd[5]=ub[5]; qd[6]=T(0); }
                else qd[3] = T(0);
            } else for (int j=0;j<nd;++j) qd[j] = ub[j];
        }
        B.V_FM = svzero<T>(); B.V_PB_G = svzero<T>();
        for (int j=0;j<nd;++j) {
            B.V_FM.a = B.V_FM.a + ub[j]*B.H_FM[j].a; B.V_FM.b = B.V_FM.b + ub[j]*B.H_FM[j].b;
            B.V_PB_G.a = B.V_PB_G.a + ub[j]*B.H[j].a; B.V_PB_G.b = B.V_PB_G.b + ub[j]*B.H[j].b;
        }
        const bool rev = M.reversed[b] != 0;
        for (int j=0;j<nd;++j) B.HDot_FM[j] = svzero<T>();
        if (t == UNIVERSAL) {
            // the mobilizer's own frames and velocity (findX_F0M0, find_w_F0M0;
            // RigidBodyNode.h:738-848): ~X_FM and H_MF u when reversed
            const M33<T> R0 = rev ? transpose(B.X_FM.R) : B.X_FM.R;
            V3<T> w0 = B.V_FM.a;
            if (rev) { w0 = zero3<T>(); for (int j=0;j<nd;++j) w0 = w0 + ub[j]*B.H_MF[j].a; }
            const V3<T> y = mk3<T>(R0(0,1), R0(1,1), R0(2,1));
            B.HDot_FM[1] = sv(cross(w0, y), zero3<T>());
        } else if (t == BENDSTRETCH) {
            // RigidBodyNodeSpec_PolarCoords.h:174-200: HDot_FM(0) = (0, z x v_FM),
            // HDot_FM(1) = (0, w_FM x Mx)
            const V3<T> ez = mk3<T>(T(0),T(0),T(1));
            const V3<T> mx = mk3<T>(B.X_FM.R(0,0), B.X_FM.R(1,0), B.X_FM.R(2,0));
            B.HDot_FM[0] = sv(zero3<T>(), cross(ez, B.V_FM.b));
            B.HDot_FM[1] = sv(zero3<T>(), cross(B.V_FM.a, mx));
        } else if (t == GIMBAL || t == BUSHING) {
            // d/dt of H_FM with qdot = u (RigidBodyNodeSpec_Gimbal.h:180-202,
            // _Bushing.h:195-220; a Bushing's translation columns are constant)
            const T c0=B.cq[0], c1=B.cq[1], s0=B.sq[0], s1=B.sq[1];
            const T dc0 = -s0*ub[0], dc1 = -s1*ub[1], ds0 = c0*ub[0], ds1 = c1*ub[1];
            B.HDot_FM[1] = sv(mk3<T>(T(0), dc0, ds0), zero3<T>());
            B.HDot_FM[2] = sv(mk3<T>(ds1, -ds0*c1 - s0*dc1, dc0*c1 + c0*dc1), zero3<T>());
        }
        if (rev) {
            // calcReverseMobilizerHDot_FM (RigidBodyNodeSpec.cpp:168-216):
            //   HDot_FM = [w_FM]x H_FM - R_FM HDot_MF
            //   HDot_FM_v -= [p_FM]x HDot_FM_w + ([v_FM]x - [w_FM]x [p_FM]x) H_FM_w
            const M33<T>& R = B.X_FM.R;
            const M33<T> px = crossMat(B.X_FM.p), wx = crossMat(B.V_FM.a), vx = crossMat(B.V_FM.b);
            const M33<T> vwp = vx - wx*px;
            for (int j=0;j<nd;++j) {
                const SV<T> HD_MF = B.HDot_FM[j];
                SV<T> hd = sv(wx*B.H_FM[j].a - R*HD_MF.a, wx*B.H_FM[j].b - R*HD_MF.b);
                hd.b = hd.b - (px*hd.a + vwp*B.H_FM[j].a);
                B.HDot_FM[j] = hd;
            }
        }
        const Xform<T> X_PF = castX<T>(M.X_PF[b]);
        const Xform<T> X_MB = castX<T>(M.X_MB[b]);
        const M33<T> R_GF = Pb.X_GB.R * X_PF.R;
        const V3<T>& w_GF = Pb.V_GB.a;
        const V3<T> r_MB_F = B.X_FM.R * X_MB.p;
        const V3<T>& w_FM = B.V_FM.a;
        const V3<T> wxr = cross(w_FM, r_MB_F);
        B.VD_PB_G = svzero<T>();
        for (int j=0;j<nd;++j) {
            const V3<T> lin = B.HDot_FM[j].b - cross(r_MB_F, B.HDot_FM[j].a) - cross(wxr, B.H_FM[j].a);
            B.HDot[j] = sv(R_GF*B.HDot_FM[j].a + cross(w_GF, B.H[j].a),
                           R_GF*lin            + cross(w_GF, B.H[j].b));
            B.VD_PB_G.a = B.VD_PB_G.a + ub[j]*B.HDot[j].a;
            B.VD_PB_G.b = B.VD_PB_G.b + ub[j]*B.HDot[j].b;
        }
        // joint-independent
        const SV<T>& V_GP = Pb.V_GB;
        B.V_GB = phiTMul(B.phi, V_GP) + B.V_PB_G;
        const V3<T>& w = B.V_GB.a;
        B.b = sv(B.Mk.m*cross(w, B.Mk.G*w), B.Mk.m*cross(w, cross(w, B.Mk.p)));
        B.a = sv(B.VD_PB_G.a, B.VD_PB_G.b + cross(V_GP.a, B.V_GB.b - V_GP.b));
        B.atot = phiTMul(B.phi, Pb.atot) + B.a;
        B.MAb = siMul(B.Mk, B.atot) + B.b;
    }
}

// Gravity body forces (Force_Gravity.cpp:514-578): F = (p_CB_G x m g, m g).
template <class T> void addGravity(const Tree<T>& tr, SV<T>* F) {
    const Model& M = *tr.M;
    const V3<T> g = cast3<T>(M.g);
    for (int b = 1; b <= M.nb; ++b) {
        const Body<T>& B = tr.B[b];
        const V3<T> f = B.Mk.m*g;
        F[b] = F[b] + sv(cross(B.Mk.p, f), f);
    }
}
template <class T> T potentialEnergy(const Tree<T>& tr) {   // Force_Gravity.cpp:555
    const Model& M = *tr.M;
    const V3<T> g = cast3<T>(M.g);
    T pe = T(0);
    for (int b = 1; b <= M.nb; ++b) pe = pe - tr.B[b].Mk.m*dot(g, tr.B[b].com_G);
    return pe;
}
template <class T> T kineticEnergy(const Tree<T>& tr) {      // RigidBodyNode.cpp:182-188
    T ke = T(0);
    for (int b = 1; b <= tr.M->nb; ++b) {
        const Body<T>& B = tr.B[b];
        ke = ke + dot(B.V_GB, siMul(B.Mk, B.V_GB))/T(2);
    }
    return ke;
}

// Small dense inverse.  dof<=3 use the cofactor forms of
// SmallMatrixMixed.h:895-960; dof=6 (Free) uses LU with partial pivoting and
// column solves, the algorithm of LAPACK dgetrf/dgetri that the reference
// calls (SmallMatrixMixed.h:858-894); LAPACK itself is not vendored.
template <class T> bool invertSmall(int n, const T* A, T* Ai) {
    if (n == 1) { Ai[0] = T(1)/A[0]; return true; }
    if (n == 2) {
        const T d = A[0]*A[3] - A[1]*A[2];
        const T o = T(1)/d;
        Ai[0] = o*A[3]; Ai[1] = -(o*A[1]); Ai[2] = -(o*A[2]); Ai[3] = o*A[0];
        return true;
    }
    if (n == 3) {
        auto m = [&](int r, int c) -> const T& { return A[r*3+c]; };
        const T d00 = m(1,1)*m(2,2)-m(1,2)*m(2,1), nd01 = m(1,2)*m(2,0)-m(1,0)*m(2,2),
                d02 = m(1,0)*m(2,1)-m(1,1)*m(2,0);
        const T d = m(0,0)*d00 + m(0,1)*nd01 + m(0,2)*d02;
        const T o = T(1)/d;
        const T nd10 = m(0,2)*m(2,1)-m(0,1)*m(2,2), d11 = m(0,0)*m(2,2)-m(0,2)*m(2,0),
                nd12 = m(0,1)*m(2,0)-m(0,0)*m(2,1), d20 = m(0,1)*m(1,2)-m(0,2)*m(1,1),
                nd21 = m(0,2)*m(1,0)-m(0,0)*m(1,2), d22 = m(0,0)*m(1,1)-m(0,1)*m(1,0);
        Ai[0]=o*d00;  Ai[1]=o*nd10; Ai[2]=o*d20;
        Ai[3]=o*nd01; Ai[4]=o*d11;  Ai[5]=o*nd21;
        Ai[6]=o*d02;  Ai[7]=o*nd12; Ai[8]=o*d22;
        return true;
    }
    // n x n, LU with partial pivoting: P A = L U
    T LU[36]; int piv[6];
    for (int i=0;i<n*n;++i) LU[i]=A[i];
    for (int k=0;k<n;++k) {
        int p = k; double best = fabs(double(LU[k*n+k]));
        for (int r=k+1;r<n;++r) { const double v = fabs(double(LU[r*n+k])); if (v > best) { best=v; p=r; } }
        if (best == 0.0) return false;
        piv[k]=p;
        if (p != k) for (int c=0;c<n;++c) std::swap(LU[k*n+c], LU[p*n+c]);
        const T inv = T(1)/LU[k*n+k];
        for (int r=k+1;r<n;++r) {
            LU[r*n+k] = LU[r*n+k]*inv;
            const T l = LU[r*n+k];
            for (int c=k+1;c<n;++c) LU[r*n+c] = LU[r*n+c] - l*LU[k*n+c];
        }
    }
    for (int col=0; col<n; ++col) {
        T x[6];
        for (int i=0;i<n;++i) x[i] = T(0);
        // e_col permuted: apply the recorded row interchanges in order
        int target = col;
        T e[6]; for (int i=0;i<n;++i) e[i]=T(i==col?1:0);
        for (int k=0;k<n;++k) if (piv[k]!=k) std::swap(e[k], e[piv[k]]);
        (void)target;
        for (int i=0;i<n;++i) { T s = e[i]; for (int j=0;j<i;++j) s = s - LU[i*n+j]*x[j]; x[i]=s; }
        for (int i=n-1;i>=0;--i) { T s = x[i]; for (int j=i+1;j<n;++j) s = s - LU[i*n+j]*x[j]; x[i] = s/LU[i*n+i]; }
        for (int i=0;i<n;++i) Ai[i*n+col] = x[i];
    }
    return true;
}

// ---------------------------------------------------------------------------
// Articulated-body inertias, tip to base.  RigidBodyNodeSpec.cpp:249-325;
// driver SimbodyMatterSubsystemRep.cpp:1217-1246.  Returns false if a D is
// singular.
// ---------------------------------------------------------------------------
// presc (nullable, [nb+1]): bodies whose udot is known (prescribed motion or
// a lock); their P+ = P (RigidBodyNodeSpec.cpp:292-296).
template <class T> bool realizeABI(Tree<T>& tr, const int* presc = nullptr) {
    const Model& M = *tr.M;
    bool ok = true;
    for (int L = (int)M.levels.size()-1; L >= 1; --L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        const int nd = M.dof[b];
        B.P = abiFromRigid(B.Mk);
        for (int c : M.children[b]) abiAdd(B.P, abiShift(tr.B[c].PPlus, tr.B[c].phi));
        if (presc && presc[b]) { B.PPlus = B.P; continue; }
        SV<T> PH[6];
        for (int j=0;j<nd;++j) PH[j] = abiMul(B.P, B.H[j]);
        for (int r=0;r<nd;++r) for (int c=0;c<nd;++c) B.D[r*nd+c] = dot(B.H[r], PH[c]);
        if (!invertSmall(nd, B.D, B.DI)) ok = false;
        for (int c=0;c<nd;++c) {
            SV<T> acc = svzero<T>();
            for (int k=0;k<nd;++k) { acc.a = acc.a + B.DI[k*nd+c]*PH[k].a; acc.b = acc.b + B.DI[k*nd+c]*PH[k].b; }
            B.G[c] = acc;
        }
        M33<T> massMoment, mass, inertia;
        for (int r=0;r<3;++r) for (int c=0;c<3;++c) {
            T mm = T(0), ms = T(0), in = T(0);
            for (int j=0;j<nd;++j) {
                mm = mm + B.G[j].a[r]*PH[j].b[c];
                ms = ms + B.G[j].b[r]*PH[j].b[c];
                in = in + B.G[j].a[r]*PH[j].a[c];
            }
            massMoment(r,c)=mm; mass(r,c)=ms; inertia(r,c)=in;
        }
        M33<T> symMass, symIn;
        for (int r=0;r<3;++r) for (int c=0;c<3;++c) {
            if (r==c) { symMass(r,c)=mass(r,c); symIn(r,c)=inertia(r,c); }
            else { symMass(r,c)=(mass(r,c)+mass(c,r))/T(2); symIn(r,c)=(inertia(r,c)+inertia(c,r))/T(2); }
        }
        B.PPlus.M = B.P.M - symMass;
        B.PPlus.F = B.P.F - massMoment;
        B.PPlus.J = B.P.J - symIn;
    }
    return ok;
}
// P a + b, RigidBodyNode.cpp:202-213.
template <class T> void realizeABVelocity(Tree<T>& tr) {
    for (int b = 1; b <= tr.M->nb; ++b) tr.B[b].PAb = abiMul(tr.B[b].P, tr.B[b].a) + tr.B[b].b;
}

template <class T> inline SV<T> Hmul(const SV<T>* H, int nd, const T* x) {
    SV<T> r = svzero<T>();
    for (int j=0;j<nd;++j) { r.a = r.a + x[j]*H[j].a; r.b = r.b + x[j]*H[j].b; }
    return r;
}

// ---------------------------------------------------------------------------
// Forward dynamics: calcTreeAccelerations (SimbodyMatterSubsystemRep.cpp:
// 5258-5343) with calcUDotPass1Inward / calcUDotPass2Outward
// (RigidBodyNodeSpec.cpp:355-446) and calcQDotDot.  fmob may be null (zero);
// F is [nb+1] body forces in G (index 0 = Ground, ignored).
// ---------------------------------------------------------------------------
template <class T> void forwardDynamics(Tree<T>& tr, const T* q, const T* u, const T* fmob,
                                        const SV<T>* F, T* udot, T* qdotdot) {
    const Model& M = *tr.M;
    for (int L = (int)M.levels.size()-1; L >= 1; --L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        const int nd = M.dof[b];
        B.z = B.PAb - F[b];
        for (int c : M.children[b]) B.z = B.z + phiMul(tr.B[c].phi, tr.B[c].zPlus);
        for (int j=0;j<nd;++j) {
            const T f = fmob ? fmob[M.uIndex[b]+j] : T(0);
            B.eps[j] = f - dot(B.H[j], B.z);
        }
        B.zPlus = B.z + Hmul(B.G, nd, B.eps);
    }
    for (size_t L = 1; L < M.levels.size(); ++L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        const int nd = M.dof[b], t = M.type[b];
        const SV<T> Ap = phiTMul(B.phi, tr.B[M.parent[b]].A_GB);
        T* ud = udot + M.uIndex[b];
        for (int r=0;r<nd;++r) {
            T s = T(0);
            for (int c=0;c<nd;++c) s = s + B.DI[r*nd+c]*B.eps[c];
            ud[r] = s - dot(B.G[r], Ap);
        }
        B.A_GB = Ap + Hmul(B.H, nd, ud) + B.a;
        if (qdotdot) {
            T* qdd = qdotdot + M.qIndex[b];
            const T* qb = q + M.qIndex[b];
            const T* ub = u + M.uIndex[b];
            if (M.isQuat(b)) {
                quatDotDot(qb, mk3<T>(ub[0],ub[1],ub[2]), mk3<T>(ud[0],ud[1],ud[2]), qdd);
                if (t == FREE) { qdd[4]=ud[3]; qdd[5]=ud[4]; qdd[6]=ud[5]; }
            } else if (t == BALL || t == FREE) {
                // Euler-angle mode (RigidBodyNodeSpec_Ball.h:355-392)
                const V3<T> qd = eulerDot(B.cq, B.sq, mk3<T>(ub[0],ub[1],ub[2]));
                const V3<T> ea = eulerDotDot(B.cq, B.sq, qd, mk3<T>(ud[0],ud[1],ud[2]));
                qdd[0]=ea[0]; qdd[1]=ea[1]; qdd[2]=ea[2];
                if (t == FREE) { qdd[3]=ud[3]; qdd[4]=ud[4]; qdd[5]=ud[5]; qdd[6]=T(0); }
                else qdd[3] = T(0);
            } else for (int j=0;j<nd;++j) qdd[j] = ud[j];
        }
    }
}

// Forward dynamics with known udot on the bodies marked in presc ([nb+1]):
// calcUDotPass1Inward / calcUDotPass2Outward with their isUDotKnown branches
// (RigidBodyNodeSpec.cpp:355-446).  udot holds the known entries on input and
// receives the others; tau receives, on the prescribed mobilities,
// eps - H^T P A+ (zero elsewhere): with tau on the left-hand side,
// M udot + c + tau = f + J^T F, so the force the motion needs is -tau.
// Needs realizeABI(tr, presc) and realizeABVelocity first.
template <class T> void forwardDynamicsPrescribed(Tree<T>& tr, const int* presc, const T* fmob,
                                                  const SV<T>* F, T* udot, T* tau) {
    const Model& M = *tr.M;
    for (int L = (int)M.levels.size()-1; L >= 1; --L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        const int nd = M.dof[b];
        B.z = B.PAb - F[b];
        if (presc[b]) B.z = B.z + abiMul(B.P, Hmul(B.H, nd, udot + M.uIndex[b]));
        for (int c : M.children[b]) B.z = B.z + phiMul(tr.B[c].phi, tr.B[c].zPlus);
        for (int j=0;j<nd;++j) {
            const T f = fmob ? fmob[M.uIndex[b]+j] : T(0);
            B.eps[j] = f - dot(B.H[j], B.z);
        }
        B.zPlus = presc[b] ? B.z : B.z + Hmul(B.G, nd, B.eps);
    }
    for (size_t L = 1; L < M.levels.size(); ++L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        const int nd = M.dof[b];
        const SV<T> Ap = phiTMul(B.phi, tr.B[M.parent[b]].A_GB);
        T* ud = udot + M.uIndex[b];
        T* tb = tau + M.uIndex[b];
        if (presc[b]) {
            const SV<T> PA = abiMul(B.P, Ap);
            for (int r=0;r<nd;++r) tb[r] = B.eps[r] - dot(B.H[r], PA);
        } else {
            for (int r=0;r<nd;++r) {
                T s = T(0);
                for (int c=0;c<nd;++c) s = s + B.DI[r*nd+c]*B.eps[c];
                ud[r] = s - dot(B.G[r], Ap);
                tb[r] = T(0);
            }
        }
        B.A_GB = Ap + Hmul(B.H, nd, ud) + B.a;
    }
}

// udot = M^-1 f, RigidBodyNodeSpec.cpp:483-552, SimbodyMatterSubsystemRep.cpp:5357-5400.
template <class T> void multiplyByMInv(Tree<T>& tr, const T* f, T* out) {
    const Model& M = *tr.M;
    for (int L = (int)M.levels.size()-1; L >= 1; --L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        const int nd = M.dof[b];
        B.z = svzero<T>();
        for (int c : M.children[b]) B.z = B.z + phiMul(tr.B[c].phi, tr.B[c].zPlus);
        for (int j=0;j<nd;++j) B.eps[j] = f[M.uIndex[b]+j] - dot(B.H[j], B.z);
        B.zPlus = B.z + Hmul(B.G, nd, B.eps);
    }
    for (size_t L = 1; L < M.levels.size(); ++L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        const int nd = M.dof[b];
        const SV<T> Ap = phiTMul(B.phi, tr.B[M.parent[b]].A_GB);
        T* ud = out + M.uIndex[b];
        for (int r=0;r<nd;++r) {
            T s = T(0);
            for (int c=0;c<nd;++c) s = s + B.DI[r*nd+c]*B.eps[c];
            ud[r] = s - dot(B.G[r], Ap);
        }
        B.A_GB = Ap + Hmul(B.H, nd, ud);
    }
}

// Inverse dynamics: calcTreeResidualForces (SimbodyMatterSubsystemRep.cpp:
// 5532-5616) with calcBodyAccelerationsFromUdotOutward / calcInverseDynamicsPass2Inward
// (RigidBodyNodeSpec.cpp:566-641).  Null fmob/F/udot mean zeros.
template <class T> void residualForces(Tree<T>& tr, const T* fmob, const SV<T>* F, const T* udot, T* out) {
    const Model& M = *tr.M;
    for (size_t L = 1; L < M.levels.size(); ++L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        const int nd = M.dof[b];
        T ud[6];
        for (int j=0;j<nd;++j) ud[j] = udot ? udot[M.uIndex[b]+j] : T(0);
        B.A_GB = phiTMul(B.phi, tr.B[M.parent[b]].A_GB) + Hmul(B.H, nd, ud) + B.a;
    }
    for (int L = (int)M.levels.size()-1; L >= 1; --L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        const int nd = M.dof[b];
        B.F = siMul(B.Mk, B.A_GB) + B.b;
        if (F) B.F = B.F - F[b];
        for (int c : M.children[b]) B.F = B.F + phiMul(tr.B[c].phi, tr.B[c].F);
        for (int j=0;j<nd;++j) {
            const T f = fmob ? fmob[M.uIndex[b]+j] : T(0);
            out[M.uIndex[b]+j] = dot(B.H[j], B.F) - f;
        }
    }
}

// ---------------------------------------------------------------------------
// O(n) operators that reuse the position (and velocity) caches.
// ---------------------------------------------------------------------------
// Mass-matrix product f = M v: multiplyByMPass1Outward / multiplyByMPass2Inward
// (RigidBodyNodeSpec.cpp:657-695), driver SimbodyMatterSubsystemRep.cpp:
// 5411-5450.  Needs realizePosition().  Uses B.A_GB and B.F as temporaries.
template <class T> void multiplyByM(Tree<T>& tr, const T* v, T* out) {
    const Model& M = *tr.M;
    for (size_t L = 1; L < M.levels.size(); ++L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        B.A_GB = phiTMul(B.phi, tr.B[M.parent[b]].A_GB) + Hmul(B.H, M.dof[b], v + M.uIndex[b]);
    }
    for (int L = (int)M.levels.size()-1; L >= 1; --L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        B.F = siMul(B.Mk, B.A_GB);
        for (int c : M.children[b]) B.F = B.F + phiMul(tr.B[c].phi, tr.B[c].F);
        for (int j=0;j<M.dof[b];++j) out[M.uIndex[b]+j] = dot(B.H[j], B.F);
    }
}

// Jv = J v, one spatial vector per body, Ground's zero (RigidBodyNodeSpec.cpp:
// 760-785; driver SimbodyMatterSubsystemRep.cpp:6293).  Needs realizePosition().
template <class T> void systemJacobian(Tree<T>& tr, const T* v, SV<T>* Jv) {
    const Model& M = *tr.M;
    Jv[0] = svzero<T>();
    for (size_t L = 1; L < M.levels.size(); ++L)
    for (int b : M.levels[L]) {
        const Body<T>& B = tr.B[b];
        Jv[b] = phiTMul(B.phi, Jv[M.parent[b]]) + Hmul(B.H, M.dof[b], v + M.uIndex[b]);
    }
}

// f = J^T F for spatial forces F at the body origins, in G; F[0] (Ground) is
// ignored (RigidBodyNodeSpec.cpp:795-815).  Needs realizePosition().  Uses
// B.z as the temporary.
template <class T> void systemJacobianTranspose(Tree<T>& tr, const SV<T>* F, T* out) {
    const Model& M = *tr.M;
    for (int L = (int)M.levels.size()-1; L >= 1; --L)
    for (int b : M.levels[L]) {
        Body<T>& B = tr.B[b];
        B.z = F[b];
        for (int c : M.children[b]) B.z = B.z + phiMul(tr.B[c].phi, tr.B[c].z);
        for (int j=0;j<M.dof[b];++j) out[M.uIndex[b]+j] = dot(B.H[j], B.z);
    }
}

// A_GB = J udot + JDot u (calcBodyAccelerationsFromUdotOutward,
// RigidBodyNodeSpec.cpp:566-582; driver SimbodyMatterSubsystemRep.cpp:
// 3592-3625).  Needs realizeVelocity().  Null udot means zeros.
template <class T> void bodyAccelerationFromUDot(Tree<T>& tr, const T* udot, SV<T>* A) {
    const Model& M = *tr.M;
    A[0] = svzero<T>();
    for (size_t L = 1; L < M.levels.size(); ++L)
    for (int b : M.levels[L]) {
        const Body<T>& B = tr.B[b];
        const int nd = M.dof[b];
        T ud[6];
        for (int j=0;j<nd;++j) ud[j] = udot ? udot[M.uIndex[b]+j] : T(0);
        A[b] = phiTMul(B.phi, A[M.parent[b]]) + Hmul(B.H, nd, ud) + B.a;
    }
}

// ---------------------------------------------------------------------------
// Holonomic constraints, with Ground as the ancestor frame A in which the
// reference states them (its multipliers are forces in A; with A = Ground
// they are forces in G):
//   Constraint::Ball (ConstraintImpl.h:1835-2000): station P on body 1,
//     station S on body 2; perr = p_S - p_P; verr = v_S - v_C and aerr =
//     a_S - a_C with C the material point of body 1 at S (lever p_S - p_B1);
//     +lambda at S on body 2, -lambda at C on body 1;
//   Constraint::Rod (Constraint_RodImpl.h:193-300): stations Sf on body 1, Sb
//     on body 2, length d; perr = |p_Sb - p_Sf| - d; verr = (v_Sb - v_Sf).Cz;
//     aerr = (a_Sb - a_Sf).Cz + (v_Sb - v_Sf).Czd, Cz = (p_Sb - p_Sf)/r,
//     Czd = (pd - Cz (pd.Cz))/r; lambda Cz at Sb on body 2, -lambda Cz at Sf
//     on body 1.
// Station velocity v = V.lin + V.ang x r; station acceleration a = A.lin +
// A.ang x r + w x (w x r).
// ---------------------------------------------------------------------------
enum ConsType { CONS_BALL = 0, CONS_ROD = 1 };
struct ConsDef { int type, b1, b2; double s1[3], s2[3], len; };
inline int consRows(const ConsDef& c) { return c.type == CONS_BALL ? 3 : 1; }
inline int consRowsTotal(const std::vector<ConsDef>& cs) {
    int m = 0;
    for (const ConsDef& c : cs) m += consRows(c);
    return m;
}

template <class T> struct ConsGeom { V3<T> p1, p2, r1, r2, Cz; T r; };
template <class T> ConsGeom<T> consGeom(const Tree<T>& tr, const ConsDef& c) {
    const Xform<T>& X1 = tr.B[c.b1].X_GB;
    const Xform<T>& X2 = tr.B[c.b2].X_GB;
    ConsGeom<T> g;
    const V3<T> l1 = X1.R * mk3<T>(T(c.s1[0]), T(c.s1[1]), T(c.s1[2]));
    const V3<T> l2 = X2.R * mk3<T>(T(c.s2[0]), T(c.s2[1]), T(c.s2[2]));
    g.p1 = X1.p + l1; g.p2 = X2.p + l2;
    g.r2 = l2;
    g.r1 = c.type == CONS_BALL ? g.p2 - X1.p : l1;
    g.r = T(0); g.Cz = zero3<T>();
    if (c.type == CONS_ROD) {
        const V3<T> d = g.p2 - g.p1;
        g.r = sqrt(dot(d, d));
        g.Cz = (T(1)/g.r) * d;
    }
    return g;
}
template <class T> inline V3<T> stationVel(const SV<T>& V, const V3<T>& r) { return V.b + cross(V.a, r); }
template <class T> inline V3<T> stationAcc(const SV<T>& A, const SV<T>& V, const V3<T>& r) {
    return (A.b + cross(A.a, r)) + cross(V.a, cross(V.a, r));
}
// Position errors (needs realizePosition).
template <class T> void consPerr(const Tree<T>& tr, const std::vector<ConsDef>& cs, T* perr) {
    int k = 0;
    for (const ConsDef& c : cs) {
        const ConsGeom<T> g = consGeom(tr, c);
        if (c.type == CONS_BALL) { const V3<T> d = g.p2 - g.p1; perr[k] = d[0]; perr[k+1] = d[1]; perr[k+2] = d[2]; k += 3; }
        else perr[k++] = g.r - T(c.len);
    }
}
// Rows of the constraint Jacobian applied to per-body spatial velocities V
// (no centripetal terms): the velocity errors when V = V_GB, the rows of
// P v when V = J v.
template <class T> void consRowsOf(const Tree<T>& tr, const std::vector<ConsDef>& cs, const SV<T>* V, T* out) {
    int k = 0;
    for (const ConsDef& c : cs) {
        const ConsGeom<T> g = consGeom(tr, c);
        const V3<T> d = stationVel(V[c.b2], g.r2) - stationVel(V[c.b1], g.r1);
        if (c.type == CONS_BALL) { out[k] = d[0]; out[k+1] = d[1]; out[k+2] = d[2]; k += 3; }
        else out[k++] = dot(d, g.Cz);
    }
}
// Acceleration errors from body accelerations A (with their velocity terms)
// at the velocities of the tree.
template <class T> void consAerr(const Tree<T>& tr, const std::vector<ConsDef>& cs, const SV<T>* A, T* aerr) {
    int k = 0;
    for (const ConsDef& c : cs) {
        const ConsGeom<T> g = consGeom(tr, c);
        const SV<T>& V1 = tr.B[c.b1].V_GB; const SV<T>& V2 = tr.B[c.b2].V_GB;
        const V3<T> dd = stationAcc(A[c.b2], V2, g.r2) - stationAcc(A[c.b1], V1, g.r1);
        if (c.type == CONS_BALL) { aerr[k] = dd[0]; aerr[k+1] = dd[1]; aerr[k+2] = dd[2]; k += 3; }
        else {
            const V3<T> pd = stationVel(V2, g.r2) - stationVel(V1, g.r1);
            const V3<T> Czd = (T(1)/g.r) * (pd - dot(pd, g.Cz) * g.Cz);
            aerr[k++] = dot(dd, g.Cz) + dot(pd, Czd);
        }
    }
}
// Body forces (in G, about the body origins) of multipliers lambda, added
// to F with sign sgn.
template <class T> void consForces(const Tree<T>& tr, const std::vector<ConsDef>& cs, const T* lam, T sgn,
                                   SV<T>* F) {
    int k = 0;
    for (const ConsDef& c : cs) {
        const ConsGeom<T> g = consGeom(tr, c);
        V3<T> f;
        if (c.type == CONS_BALL) { f = mk3<T>(lam[k], lam[k+1], lam[k+2]); k += 3; }
        else f = lam[k++] * g.Cz;
        f = sgn * f;
        F[c.b2] = F[c.b2] + sv(cross(g.r2, f), f);
        F[c.b1] = F[c.b1] - sv(cross(g.r1, f), f);
    }
}
// Gaussian elimination with partial pivoting: solves A x = b in place (A n x
// n row-major, b -> x).  False when a pivot is zero.
template <class T> bool solveDense(int n, std::vector<T>& A, T* b) {
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int r = k + 1; r < n; ++r) if (fabs(double(A[r*n+k])) > fabs(double(A[piv*n+k]))) piv = r;
        if (A[piv*n+k] == T(0)) return false;
        if (piv != k) { for (int j = 0; j < n; ++j) std::swap(A[k*n+j], A[piv*n+j]); std::swap(b[k], b[piv]); }
        for (int r = k + 1; r < n; ++r) {
            const T f = A[r*n+k] / A[k*n+k];
            for (int j = k; j < n; ++j) A[r*n+j] = A[r*n+j] - f*A[k*n+j];
            b[r] = b[r] - f*b[k];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        T s = b[k];
        for (int j = k + 1; j < n; ++j) s = s - A[k*n+j]*b[j];
        b[k] = s / A[k*n+k];
    }
    return true;
}
// Loop forward dynamics (calcLoopForwardDynamicsOperator,
// SimbodyMatterSubsystemRep.cpp:5076-5152): udot of the tree under F, the
// acceleration errors aerr it leaves, W = G M^-1 G^T one column per
// multiplier (M^-1 of the generalized force of a unit multiplier; the rows
// of the resulting body accelerations, which carry no velocity terms), W
// lambda = aerr, and udot again under F minus the constraint forces.  Needs
// realizeABI and realizeABVelocity.  lam (nullable) receives the
// multipliers.  False when W is singular.
template <class T> bool loopForward(Tree<T>& tr, const std::vector<ConsDef>& cs, const T* q, const T* u,
                                    const SV<T>* F, T* udot, T* lam) {
    const Model& M = *tr.M;
    const int m = consRowsTotal(cs), nu = M.nu;
    forwardDynamics(tr, q, u, (const T*)nullptr, F, udot, (T*)nullptr);
    std::vector<SV<T>> A(M.nb + 1);
    for (int b = 0; b <= M.nb; ++b) A[b] = tr.B[b].A_GB;
    std::vector<T> aerr(m), W((size_t)m*m), e(m), f(nu), x(nu), col(m);
    consAerr(tr, cs, A.data(), aerr.data());
    std::vector<SV<T>> Fk(M.nb + 1);
    for (int k = 0; k < m; ++k) {
        for (int i = 0; i < m; ++i) e[i] = T(i == k ? 1 : 0);
        for (auto& v : Fk) v = svzero<T>();
        consForces(tr, cs, e.data(), T(1), Fk.data());
        systemJacobianTranspose(tr, Fk.data(), f.data());
        multiplyByMInv(tr, f.data(), x.data());
        for (int b = 0; b <= M.nb; ++b) A[b] = tr.B[b].A_GB;
        consRowsOf(tr, cs, A.data(), col.data());
        for (int i = 0; i < m; ++i) W[(size_t)i*m + k] = col[i];
    }
    if (!solveDense(m, W, aerr.data())) return false;
    std::vector<SV<T>> Fc(F, F + M.nb + 1);
    consForces(tr, cs, aerr.data(), T(-1), Fc.data());
    forwardDynamics(tr, q, u, (const T*)nullptr, Fc.data(), udot, (T*)nullptr);
    if (lam) for (int i = 0; i < m; ++i) lam[i] = aerr[i];
    return true;
}

// Unit inertia of a unit point mass at p, taken about the origin:
// |p|^2 I - p p^T (UnitInertia::pointMassAt, MassProperties.h:866-867).
template <class T> inline M33<T> pointMassUnitInertia(const V3<T>& p) {
    const T xx = p[0]*p[0], yy = p[1]*p[1], zz = p[2]*p[2];
    M33<T> G;
    G(0,0) = yy + zz; G(1,1) = xx + zz; G(2,2) = xx + yy;
    G(0,1) = G(1,0) = -(p[0]*p[1]);
    G(0,2) = G(2,0) = -(p[0]*p[2]);
    G(1,2) = G(2,1) = -(p[1]*p[2]);
    return G;
}
// SpatialInertia::shift(S): origin moved from O to O+S through the centroid
// (MassProperties.h:1103-1116, UnitInertia shifts :770-800).
template <class T> inline SpatialInertia<T> siShift(const SpatialInertia<T>& I, const V3<T>& S) {
    SpatialInertia<T> R;
    R.m = I.m;
    const V3<T> pNew = I.p - S;
    R.G = (I.G - pointMassUnitInertia(I.p)) + pointMassUnitInertia(pNew);
    R.p = pNew;
    return R;
}
// SpatialInertia += (MassProperties.h:1037-1045): mass-weighted mass center
// and unit inertia.
template <class T> inline void siAdd(SpatialInertia<T>& R, const SpatialInertia<T>& Q) {
    const T mtot = R.m + Q.m, oomtot = T(1)/mtot;
    R.p = oomtot*(R.m*R.p + Q.m*Q.p);
    M33<T> G;
    for (int i=0;i<9;++i) G.m[i] = oomtot*(R.m*R.G.m[i] + Q.m*Q.G.m[i]);
    R.G = G;
    R.m = mtot;
}
// Composite body inertias: each body with everything outboard of it welded,
// about its origin, in G (RigidBodyNode.cpp:232-243, driver
// SimbodyMatterSubsystemRep.cpp:5196-5205).  R[0] is Ground's infinite mass
// (RigidBodyNode_Weld.cpp:168-171).  Needs realizePosition().
template <class T> void compositeBodyInertias(const Tree<T>& tr, SpatialInertia<T>* R) {
    const Model& M = *tr.M;
    R[0].m = T(std::numeric_limits<double>::infinity()); R[0].p = zero3<T>(); R[0].G = ident33<T>();
    for (int L = (int)M.levels.size()-1; L >= 1; --L)
    for (int b : M.levels[L]) {
        R[b] = tr.B[b].Mk;
        for (int c : M.children[b]) siAdd(R[b], siShift(R[c], -tr.B[c].phi));
    }
}

// Mobilizer reaction forces at M, expressed in G (SimbodyMatterSubsystemRep.cpp:
// 5786-5826).  Requires forwardDynamics() to have run.  Checker only.
template <class T> void reactionForces(const Tree<T>& tr, SV<T>* FM) {
    const Model& M = *tr.M;
    FM[0] = svzero<T>();
    for (int b = 1; b <= M.nb; ++b) {
        const Body<T>& B = tr.B[b];
        const Body<T>& Pb = tr.B[M.parent[b]];
        const V3<T> p_PB_G = B.X_GB.p - Pb.X_GB.p;
        const SV<T> Ap = sv(Pb.A_GB.a, Pb.A_GB.b + cross(Pb.A_GB.a, p_PB_G));
        const SV<T> FB = B.zPlus + abiMul(B.PPlus, Ap);
        const V3<T> p_BM_G = B.X_GB.R * cast3<T>(M.p_BM[b]);
        FM[b] = sv(FB.a - cross(p_BM_G, FB.b), FB.b);
    }
}

// ---------------------------------------------------------------------------
// System derivative: the realize(Acceleration) of MultibodySystem with
// Force::Gravity plus constant body forces (Force::DiscreteForces).  ydot =
// [qdot; udot].  Fapp may be null; it has nb+1 entries.
// ---------------------------------------------------------------------------
template <class T> struct Workspace {
    std::vector<SV<T>> F;
    std::vector<T> tmp;
    // holonomic constraints (null: none) and the multipliers of the last
    // evaluation
    const std::vector<ConsDef>* cons = nullptr;
    std::vector<T> lam;
    // known udot (locks, constant prescribed udot): presc [nb+1], udotKnown
    // [nu]; null when none
    const int* presc = nullptr;
    const double* udotKnown = nullptr;
};

template <class T> bool derivs(Tree<T>& tr, Workspace<T>& ws, const T* y, const SV<T>* Fapp, T* ydot) {
    const Model& M = *tr.M;
    const T* q = y; const T* u = y + M.nq;
    realizePosition(tr, q);
    realizeVelocity(tr, q, u, ydot);
    ws.F.assign(M.nb+1, svzero<T>());
    addGravity(tr, ws.F.data());
    if (Fapp) for (int b=1;b<=M.nb;++b) ws.F[b] = ws.F[b] + Fapp[b];
    const bool ok = realizeABI(tr, ws.presc);
    realizeABVelocity(tr);
    if (ws.presc) {
        T* ud = ydot + M.nq;
        for (int i=0;i<M.nu;++i) ud[i] = T(ws.udotKnown[i]);
        ws.tmp.resize(M.nu);
        forwardDynamicsPrescribed(tr, ws.presc, (const T*)nullptr, ws.F.data(), ud, ws.tmp.data());
    } else if (ws.cons && !ws.cons->empty()) {
        ws.lam.resize(consRowsTotal(*ws.cons));
        if (!loopForward(tr, *ws.cons, q, u, ws.F.data(), ydot + M.nq, ws.lam.data())) return false;
    } else {
        forwardDynamics(tr, q, u, (const T*)nullptr, ws.F.data(), ydot + M.nq, (T*)nullptr);
    }
    return ok;
}

// ---------------------------------------------------------------------------
// Runge-Kutta-Merson with Simbody's step control.
//   attemptODEStep         RungeKuttaMersonIntegrator.cpp:86-140
//   attemptDAEStep         AbstractIntegratorRep.cpp:137-208
//   adjustStepSize         AbstractIntegratorRep.cpp:448-502
//   takeOneStep            AbstractIntegratorRep.cpp:513-609 (no events)
//   stepTo main loop       AbstractIntegratorRep.cpp:216-436 with finalTime set
//   methodInitialize       AbstractIntegratorRep.cpp:40-53 (h0 = timescale/10, timescale 0.1)
//   calcErrorNorm          IntegratorRep.h:454-488, scaleDQ :500-514, calcRelativeScaling :644-655
//   accuracy/consTol       IntegratorRep.h:737-740 (defaults 1e-3, acc/10)
//   projectQ               SimbodyMatterSubsystemRep.cpp:4096-4184 (no holonomic constraints)
//   normalizeQuaternions   SimbodyMatterSubsystemRep.cpp:4448-4475, _Ball.h:417-434, _Free.h:476-493
// ---------------------------------------------------------------------------
struct RkmConfig {
    double accuracy = 1e-3;     // Integrator::setAccuracy
    double fixed_h = 0.0;       // >0 : Integrator::setFixedStepSize(h)
    double init_h = -1.0, min_h = -1.0, max_h = -1.0;
    double timescale = 0.1;     // System default time scale
    long   max_attempts = 100000000;
};
// status: 0 success, 2 a derivative failed, 3 the step fell to zero, 4 too
// many attempts, 5 the initial projection failed, 6 event localization
// failed; terminated: an event handler ended the integration at time t.
struct RkmStats {
    long attempts=0, accepted=0, rejected=0, projections=0, evals=0; int status=0;
    int terminated = 0; double t = 0;
};

// One step attempt as the step controller saw it (test hook): the attempt's
// start and end times, the error norm before the projection gate (NaN when
// the stages did not converge), the quaternion error RMS(|q|-1) that the gate
// read (-1 when the gate was skipped), the error norm the controller used
// (+inf for a failed attempt), whether q was projected, whether the step was
// accepted, and the step size the controller chose next.
struct RkmAttempt { double t0, t1, err_pre, qerr, err, projected, accepted, h_next; };

// Prescribed motion of a body's mobilities (Motion.h:240-300, MotionImpl.h:
// 182-300): Motion::Sinusoid m(t) = a sin(w t + phi) applied to every q
// (position level: u = qdot = m', udot = m''; mobilizers with qdot = u
// only), to every u (velocity level: udot = m') or to every udot
// (acceleration level); Motion::Steady: every u = a, udot = 0.  Four doubles
// per body, Ground first: level, a, w, phi.
enum MotionLevel { MOTION_NONE = 0, MOTION_POSITION = 1, MOTION_VELOCITY = 2, MOTION_ACCELERATION = 3,
                   MOTION_STEADY = 4 };

// Event witnesses: the trigger functions of TriggeredEventHandler and
// TriggeredEventReporter (EventHandler.h, EventReporter.h) over a fixed set
// of kinds, with EventTriggerInfo's settings (Event.h:196-258: rising
// and/or falling sign transitions, the localization window in units of the
// time scale, default 0.1) and a fixed set of handler actions:
//   WIT_Q            q[index] - offset                       (Stage::Position)
//   WIT_U            u[index] - offset                       (Stage::Velocity)
//   WIT_STATION_POS  dir . p_G(station of body) - offset     (Stage::Position)
//   WIT_STATION_VEL  dir . v_G(station of body) - offset     (Stage::Velocity)
//   WIT_TIME         t - offset                              (Stage::Time)
//   EV_REPORT     a reporter: the event is recorded, the state is not changed
//   EV_SCALE_U    u[aindex] *= factor (every u when aindex < 0): impacts with
//                 restitution (factor -e), energy injection
//   EV_TERMINATE  the handler sets shouldTerminate
enum WitnessKind { WIT_Q = 0, WIT_U = 1, WIT_STATION_POS = 2, WIT_STATION_VEL = 3, WIT_TIME = 4 };
enum { TRIG_FALLING = 1, TRIG_RISING = 2 };    // Event::Trigger, Event.h:159-167
enum EventAction { EV_REPORT = 0, EV_SCALE_U = 1, EV_TERMINATE = 2 };
struct WitnessDef {
    int kind, index, body, trigger, action, aindex;
    double station[3], dir[3], offset, factor, window;
};
// One triggered event: the localization window (tLow, tHigh], the witness
// and the transition reported (TRIG_FALLING or TRIG_RISING,
// EventTriggerInfo::calcTransitionToReport).
struct EventRecord { double tLow, tHigh; int witness, transition; };
constexpr double kSignificantReal = 1.9323057602939027e-14;    // eps^(7/8)
constexpr int kMaxLocalizeIters = 400;

// SimTK::sign and Event::classifyTransition (Event.h:177-189): no event
// when the sign is unchanged or leaves zero; a transition to zero counts
// as a full sign change.
inline int signOf(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }
inline int classifyTransition(int before, int after) {
    if (before == after || before == 0) return 0;
    return before == 1 ? TRIG_FALLING : TRIG_RISING;
}
// estimateRootTime (IntegratorRep.h:351-376): the secant estimate with the
// caller's bias, bisection when an end value is exactly zero or the
// interval is at most minWindow, kept a buffer zone of max(10% of the
// interval, minWindow/2) away from either end.
inline double estimateRootTime(double tLow, double fLow, double tHigh, double fHigh, double bias,
                               double minWindow) {
    const double h = tHigh - tLow;
    if (fLow == 0 || fHigh == 0 || h <= minWindow) return tLow + h/2;
    const double x = fHigh/(fHigh - bias*fLow);
    double tRoot = tHigh - x*h;
    const double buf = std::max(0.1*h, minWindow/2);
    tRoot = std::max(tRoot, tLow + buf);
    tRoot = std::min(tRoot, tHigh - buf);
    return tRoot;
}
// findEventCandidates (IntegratorRep.h:390-443): the witnesses (all, or
// those of `viable`) whose sign transition over (tLow, tHigh] passes their
// mask; each one's root estimate and reported transition, the earliest
// estimate and the narrowest required window max(min(acc * timescale *
// window), minWindow).
struct EventCandidates {
    std::vector<int> idx, trans;
    std::vector<double> est;
    double earliest = 0, narrowest = 0;
};
inline void findEventCandidates(const std::vector<WitnessDef>& W, const std::vector<int>* viable, double tLow,
                                const double* eLow, double tHigh, const double* eHigh, double bias,
                                double minWindow, double acc, double timescale, EventCandidates& out) {
    out.idx.clear(); out.trans.clear(); out.est.clear();
    out.earliest = out.narrowest = std::numeric_limits<double>::infinity();
    const int n = viable ? (int)viable->size() : (int)W.size();
    for (int i = 0; i < n; ++i) {
        const int e = viable ? (*viable)[i] : i;
        const int seen = classifyTransition(signOf(eLow[e]), signOf(eHigh[e])) & W[e].trigger;
        if (!seen) continue;
        out.idx.push_back(e);
        out.trans.push_back((seen & TRIG_RISING) ? TRIG_RISING : TRIG_FALLING);
        out.narrowest = std::max(std::min(out.narrowest, acc*timescale*W[e].window), minWindow);
        out.est.push_back(estimateRootTime(tLow, eLow[e], tHigh, eHigh[e], bias, minWindow));
        out.earliest = std::min(out.earliest, out.est.back());
    }
}

template <class T> struct Rkm {
    const Model& M;
    Tree<T> tr;
    Workspace<T> ws;
    std::vector<RkmAttempt>* trace = nullptr;   // nullable: records every attempt
    const double* motion = nullptr;             // nullable: prescribed motion, 4 per body (Ground first)
    std::vector<int> mPresc;                    // bodies with prescribed udot (from motion)
    std::vector<double> mUdot;                  // their udot at the evaluation's time
    std::vector<T> yp;                          // the evaluated state with q, u prescribed
    std::vector<T> y0, f0, fa, fb, ysave, y1, ytmp, yerr, uscale;
    explicit Rkm(const Model& m) : M(m), tr(m) {
        const int ny = m.nq + m.nu;
        y0.resize(ny); f0.resize(ny); fa.resize(ny); fb.resize(ny);
        ysave.resize(ny); y1.resize(ny); ytmp.resize(ny); yerr.resize(ny); uscale.resize(m.nu);
    }

    // prescribeQ, prescribeU (SimbodyMatterSubsystemRep.cpp:3776-3830) and
    // the prescribed udot at time t: q and u of y overwritten where the
    // motion prescribes them, udot into mUdot.
    void prescribe(double t, T* y) {
        mPresc.assign(M.nb + 1, 0);
        mUdot.assign(M.nu, 0.0);
        for (int b = 1; b <= M.nb; ++b) {
            const double* m = motion + 4*b;
            const int level = (int)m[0];
            if (level == MOTION_NONE) continue;
            mPresc[b] = 1;
            const double a = m[1], w = m[2], ph = m[3];
            const double sn = std::sin(w*t + ph), cs = std::cos(w*t + ph);
            const double f = a*sn, fd = a*w*cs, fdd = -a*w*w*sn;
            const int qi = M.qIndex[b], ui = M.uIndex[b];
            for (int i = 0; i < M.dof[b]; ++i) {
                if (level == MOTION_POSITION) {
                    y[qi + i] = T(f); y[M.nq + ui + i] = T(fd); mUdot[ui + i] = fdd;
                } else if (level == MOTION_VELOCITY) {
                    y[M.nq + ui + i] = T(f); mUdot[ui + i] = fd;
                } else if (level == MOTION_STEADY) {
                    y[M.nq + ui + i] = T(a); mUdot[ui + i] = 0.0;
                } else {
                    mUdot[ui + i] = f;
                }
            }
        }
    }
    // The derivative at time t (setAdvancedStateAndRealizeDerivatives,
    // IntegratorRep.h:901-915): the prescribed q and u replace the state's,
    // the prescribed udot are known udot.  y itself is not changed.
    bool derivsAt(double t, const T* y, const SV<T>* Fapp, T* ydot) {
        if (!motion) return derivs(tr, ws, y, Fapp, ydot);
        const int ny = M.nq + M.nu;
        yp.assign(y, y + ny);
        prescribe(t, yp.data());
        ws.presc = mPresc.data(); ws.udotKnown = mUdot.data();
        const bool ok = derivs(tr, ws, yp.data(), Fapp, ydot);
        ws.presc = nullptr; ws.udotKnown = nullptr;
        return ok;
    }

    // ---- holonomic constraints (null: none) ----------------------------
    // Projection with the reference's default weights (every constraint
    // error and u weight 1, StateImpl.h:508-566): projectQ
    // (SimbodyMatterSubsystemRep.cpp:4096-4440) and projectU (:4666-4934).
    const std::vector<ConsDef>* cons = nullptr;
    // Integrator options (Integrator::setUseInfinityNorm, setProjectEveryStep,
    // setProjectInterpolatedStates; IntegratorRep.h:708-712, 780-790,
    // 858-875): every error and constraint norm the largest magnitude
    // instead of the RMS; ForceProjection on every projection (quaternions
    // normalized, one Newton iteration, even within tolerance); interpolated
    // report and localization states left unprojected.
    bool infNorm = false, forceProj = false, projInterp = true;
    std::vector<T> cP;
    // P (m x nu, row-major) at the current position kinematics: column j
    // holds the constraint rows of J e_j.
    void consP() {
        const int m = consRowsTotal(*cons), nu = M.nu;
        cP.assign((size_t)m*nu, T(0));
        std::vector<T> e(nu, T(0)), col(m);
        std::vector<SV<T>> Jv(M.nb + 1);
        for (int j = 0; j < nu; ++j) {
            e[j] = T(1);
            systemJacobian(tr, e.data(), Jv.data());
            e[j] = T(0);
            consRowsOf(tr, *cons, Jv.data(), col.data());
            for (int i = 0; i < m; ++i) cP[(size_t)i*nu + j] = col[i];
        }
    }
    // du = NInv(q) dq and dq = N(q) du: the unnormalized quaternion forms
    // (Rotation.h:712-748) on Ball and Free rotations, identity elsewhere
    // (every other mobilizer here has qdot = u).
    void mulNInv(const T* q, const T* dq, T* du) const {
        for (int b = 1; b <= M.nb; ++b) {
            const int qi = M.qIndex[b], ui = M.uIndex[b];
            if (M.isQuat(b)) {
                const V3<T> w = quatDotToW(q + qi, dq + qi);
                du[ui] = w[0]; du[ui+1] = w[1]; du[ui+2] = w[2];
                if (M.type[b] == FREE) for (int k = 0; k < 3; ++k) du[ui+3+k] = dq[qi+4+k];
            } else for (int k = 0; k < M.dof[b]; ++k) du[ui+k] = dq[qi+k];
        }
    }
    void mulN(const T* q, const T* du, T* dq) const {
        for (int b = 1; b <= M.nb; ++b) {
            const int qi = M.qIndex[b], ui = M.uIndex[b];
            if (M.isQuat(b)) {
                quatDot(q + qi, mk3<T>(du[ui], du[ui+1], du[ui+2]), dq + qi);
                if (M.type[b] == FREE) for (int k = 0; k < 3; ++k) dq[qi+4+k] = du[ui+3+k];
            } else for (int k = 0; k < M.dof[b]; ++k) dq[qi+k] = du[ui+k];
        }
    }
    // Norm of constraint errors: RMS, or the largest magnitude under
    // setUseInfinityNorm (ProjectOptions::UseInfinityNorm, SimbodyMatter-
    // SubsystemRep.cpp:4108, 4129, 4700).
    T consNorm(const std::vector<T>& v) const {
        if (v.empty()) return T(0);
        if (infNorm) {
            T m = T(0);
            for (const T& x : v) { const T a = T(fabs(double(x))); if (a > m) m = a; }
            return m;
        }
        T s = T(0);
        for (const T& x : v) s = s + x*x;
        return sqrt(s/T(v.size()));
    }
    // The least-norm Newton update in q (Pq = P NInv, dq = N NInv Pq^T
    // (Pq Pq^T)^-1 r) for the residual r at the current kinematics.
    bool qUpdate(const T* q, const std::vector<T>& r, std::vector<T>& dq) {
        const int m = (int)r.size(), nq = M.nq, nu = M.nu;
        consP();
        std::vector<T> Pq((size_t)m*nq, T(0)), rowu(nu), rowq(nq);
        for (int i = 0; i < m; ++i) {
            // row i of P NInv: (NInv^T P_i^T)^T, NInv^T from the transposed forms
            for (int j = 0; j < nu; ++j) rowu[j] = cP[(size_t)i*nu + j];
            for (int b = 1; b <= M.nb; ++b) {
                const int qi = M.qIndex[b], ui = M.uIndex[b];
                if (M.isQuat(b)) {
                    const T* qq = q + qi;
                    const T e0 = T(2)*qq[0], e1 = T(2)*qq[1], e2 = T(2)*qq[2], e3 = T(2)*qq[3];
                    const T a = rowu[ui], bb = rowu[ui+1], c = rowu[ui+2];
                    rowq[qi]   = -e1*a - e2*bb - e3*c;
                    rowq[qi+1] =  e0*a + e3*bb - e2*c;
                    rowq[qi+2] = -e3*a + e0*bb + e1*c;
                    rowq[qi+3] =  e2*a - e1*bb + e0*c;
                    if (M.type[b] == FREE) for (int k = 0; k < 3; ++k) rowq[qi+4+k] = rowu[ui+3+k];
                } else for (int k = 0; k < M.dof[b]; ++k) rowq[qi+k] = rowu[ui+k];
            }
            for (int j = 0; j < nq; ++j) Pq[(size_t)i*nq + j] = rowq[j];
        }
        std::vector<T> A((size_t)m*m), y(r);
        for (int i = 0; i < m; ++i)
            for (int k = 0; k < m; ++k) {
                T sacc = T(0);
                for (int j = 0; j < nq; ++j) sacc = sacc + Pq[(size_t)i*nq + j]*Pq[(size_t)k*nq + j];
                A[(size_t)i*m + k] = sacc;
            }
        if (!solveDense(m, A, y.data())) return false;
        std::vector<T> dqw(nq, T(0)), du(nu);
        for (int j = 0; j < nq; ++j) {
            T sacc = T(0);
            for (int i = 0; i < m; ++i) sacc = sacc + Pq[(size_t)i*nq + j]*y[i];
            dqw[j] = sacc;
        }
        mulNInv(q, dqw.data(), du.data());
        dq.assign(nq, T(0));
        mulN(q, du.data(), dq.data());
        return true;
    }
    // projectQ with the integrator's options: LocalOnly (stop when the
    // error grows from the second iteration on) unless initializing,
    // ForceProjection and ForceFullNewton when initializing.  q and its
    // error estimate qerr (nullable) are updated; false on failure.  The
    // error-estimate projection uses Pq at the final q (the reference solves
    // with the factorization of the last iteration's starting q; the two
    // differ by the last Newton step, below the constraint tolerance).
    bool projectQ(T* q, T* qerr, double consTol, double projLimit, bool init, bool* changed) {
        const int m = consRowsTotal(*cons), nq = M.nq;
        const double Significant = std::pow(2.220446049250313e-16, 0.875);
        *changed = false;
        std::vector<T> perr(m);
        realizePosition(tr, q);
        consPerr(tr, *cons, perr.data());
        const double pn0 = double(consNorm(perr)), qn0 = double(quatErrNorm(q));
        // ForceProjection: at initialize, or every step (setProjectEveryStep)
        const bool force = init || forceProj;
        if (std::max(pn0, qn0) > projLimit) return false;
        if (pn0 == 0 || (pn0 <= consTol && !force)) {
            if (qn0 > consTol || force) {
                if (M.nquat) { normalizeQuats(q, qerr); *changed = true; }
                if (double(quatErrNorm(q)) > consTol) return false;
            }
            return true;
        }
        const double target = std::max(0.1*consTol, Significant);
        std::vector<T> save(q, q + nq), dq;
        double pn = pn0, prev = pn0;
        int its = 0;
        do {
            if (!qUpdate(q, perr, dq)) return false;
            for (int j = 0; j < nq; ++j) q[j] = q[j] - dq[j];
            *changed = true; ++its;
            realizePosition(tr, q);
            consPerr(tr, *cons, perr.data());
            pn = double(consNorm(perr));
            if (!init && its >= 2 && pn > prev) {
                for (int j = 0; j < nq; ++j) q[j] = q[j] + dq[j];
                realizePosition(tr, q);
                consPerr(tr, *cons, perr.data());
                pn = double(consNorm(perr));
                break;
            }
            prev = pn;
        } while (pn > target && its < 20);
        if (pn > consTol) {
            if (pn >= pn0) { for (int j = 0; j < nq; ++j) q[j] = save[j]; realizePosition(tr, q); }
            return false;
        }
        if (qerr) {
            // remove the error estimate's component off the constraint manifold
            std::vector<T> du(M.nu), r(m);
            mulNInv(q, qerr, du.data());
            std::vector<SV<T>> Jv(M.nb + 1);
            systemJacobian(tr, du.data(), Jv.data());
            consRowsOf(tr, *cons, Jv.data(), r.data());
            if (!qUpdate(q, r, dq)) return false;
            for (int j = 0; j < nq; ++j) qerr[j] = qerr[j] - dq[j];
        }
        if (M.nquat) {
            normalizeQuats(q, qerr);
            if (double(quatErrNorm(q)) > consTol) return false;
        }
        return true;
    }
    // projectU with the integrator's options (as projectQ): u and its error
    // estimate uerr (nullable) updated at the kinematics of q; relative u
    // scaling max(|u_i|, 1) from the u on entry; one factorization of
    // P S^2 P^T for every iteration.
    bool projectU(const T* q, T* u, T* uerr, double consTol, double projLimit, bool init, bool* changed) {
        const int m = consRowsTotal(*cons), nu = M.nu;
        const double Significant = std::pow(2.220446049250313e-16, 0.875);
        *changed = false;
        std::vector<T> verr(m);
        std::vector<SV<T>> V(M.nb + 1);
        auto velErr = [&]() {
            realizeVelocity(tr, q, u, (T*)nullptr);
            for (int b = 0; b <= M.nb; ++b) V[b] = tr.B[b].V_GB;
            consRowsOf(tr, *cons, V.data(), verr.data());
            return double(consNorm(verr));
        };
        realizePosition(tr, q);
        const double vn0 = velErr();
        if (vn0 > projLimit) return false;
        if (vn0 == 0 || (vn0 <= consTol && !(init || forceProj))) return true;
        const double target = std::max(0.1*consTol, Significant);
        std::vector<T> sc(nu), save(u, u + nu), du(nu), y(m);
        for (int i = 0; i < nu; ++i) { const T a = T(fabs(double(u[i]))); sc[i] = a > T(1) ? a : T(1); }
        consP();
        std::vector<T> A0((size_t)m*m);
        for (int i = 0; i < m; ++i)
            for (int k = 0; k < m; ++k) {
                T sacc = T(0);
                for (int j = 0; j < nu; ++j) sacc = sacc + cP[(size_t)i*nu + j]*(sc[j]*sc[j])*cP[(size_t)k*nu + j];
                A0[(size_t)i*m + k] = sacc;
            }
        auto solveU = [&](const std::vector<T>& r) {
            std::vector<T> A(A0);
            y = r;
            if (!solveDense(m, A, y.data())) return false;
            for (int j = 0; j < nu; ++j) {
                T sacc = T(0);
                for (int i = 0; i < m; ++i) sacc = sacc + cP[(size_t)i*nu + j]*y[i];
                du[j] = (sc[j]*sc[j])*sacc;
            }
            return true;
        };
        double vn = vn0, prev = vn0;
        int its = 0;
        do {
            if (!solveU(verr)) return false;
            for (int j = 0; j < nu; ++j) u[j] = u[j] - du[j];
            *changed = true; ++its;
            vn = velErr();
            if (!init && its >= 2 && vn > prev) {
                for (int j = 0; j < nu; ++j) u[j] = u[j] + du[j];
                vn = velErr();
                break;
            }
            prev = vn;
        } while (vn > target && its < 7);
        if (vn > consTol) {
            if (vn >= vn0) for (int j = 0; j < nu; ++j) u[j] = save[j];
            return false;
        }
        if (uerr) {
            std::vector<SV<T>> Jv(M.nb + 1);
            std::vector<T> r(m);
            systemJacobian(tr, uerr, Jv.data());
            consRowsOf(tr, *cons, Jv.data(), r.data());
            if (!solveU(r)) return false;
            for (int j = 0; j < nu; ++j) uerr[j] = uerr[j] - du[j];
        }
        return true;
    }

    // Quaternion normalization of q (and, if given, removal of the radial
    // component of the q error estimate).
    void normalizeQuats(T* q, T* qerrest) {
        for (int b=1;b<=M.nb;++b) {
            if (!M.isQuat(b)) continue;
            T* qq = q + M.qIndex[b];
            const T n = sqrt(qq[0]*qq[0] + qq[1]*qq[1] + qq[2]*qq[2] + qq[3]*qq[3]);
            for (int i=0;i<4;++i) qq[i] = qq[i]/n;
            if (qerrest) {
                T* e = qerrest + M.qIndex[b];
                const T d = e[0]*qq[0] + e[1]*qq[1] + e[2]*qq[2] + e[3]*qq[3];
                for (int i=0;i<4;++i) e[i] = e[i] - d*qq[i];
            }
        }
    }
    // RMS of (|quat|-1) over the quaternions of q, or its largest magnitude
    // under the infinity norm; NaN-safe (a NaN error propagates).
    T quatErrNorm(const T* q) const {
        if (M.nquat == 0) return T(0);
        T s = T(0);
        for (int b=1;b<=M.nb;++b) {
            if (!M.isQuat(b)) continue;
            const T* qq = q + M.qIndex[b];
            const T e = sqrt(qq[0]*qq[0] + qq[1]*qq[1] + qq[2]*qq[2] + qq[3]*qq[3]) - T(1);
            if (infNorm) { const T a = T(fabs(double(e))); s = (a > s || std::isnan(double(a))) ? a : s; }
            else s = s + e*e;
        }
        return infNorm ? s : sqrt(s/T(M.nquat));
    }
    // max(RMS(N Wu NInv dq), RMS(uScale*du)), Wu = 1 (SimbodyMatterSubsystemRep.cpp:1630-1634);
    // under setUseInfinityNorm the largest magnitudes instead of the RMS
    // (calcWeightedInfNormQ, calcWeightedInfNorm, IntegratorRep.h:459-465).
    T errorNorm(const T* q, const T* err) const {
        T sq = T(0);
        auto acc = [&](T& s, T x) {
            if (infNorm) { const T a = T(fabs(double(x))); s = (a > s || std::isnan(double(a))) ? a : s; }
            else s = s + x*x;
        };
        for (int b=1;b<=M.nb;++b) {
            const int t = M.type[b];
            const T* e = err + M.qIndex[b];
            if (M.isQuat(b)) {
                const T* qq = q + M.qIndex[b];
                const V3<T> w = quatDotToW(qq, e);
                T d[4]; quatDot(qq, w, d);
                for (int i=0;i<4;++i) acc(sq, d[i]);
                if (t == FREE) for (int i=4;i<7;++i) acc(sq, e[i]);
            } else for (int i=0;i<M.dof[b];++i) acc(sq, e[i]);
        }
        const T qn = M.nq ? (infNorm ? sq : sqrt(sq/T(M.nq))) : T(0);
        T su = T(0);
        const T* eu = err + M.nq;
        for (int i=0;i<M.nu;++i) acc(su, uscale[i]*eu[i]);
        const T un = M.nu ? (infNorm ? su : sqrt(su/T(M.nu))) : T(0);
        return qn >= un ? qn : un;
    }

    // One RK-Merson attempt from (t0,y0,f0) to t1.  Fills y1, yerr, ysave.
    // Stage times t0 + h/3, t0 + h/3, t0 + h/2, t1; y1 gets the prescribed
    // q and u at t1 before the error estimate (setAdvancedStateAndRealize-
    // Kinematics, IntegratorRep.h:920-934, RungeKuttaMersonIntegrator.cpp:
    // 125-137).
    bool attemptODE(double t0, double t1, const SV<T>* Fapp) {
        const int ny = M.nq + M.nu;
        const T h = T(t1 - t0);
        const double hd = t1 - t0;
        bool ok = true;
        for (int i=0;i<ny;++i) ytmp[i] = y0[i] + (h/T(3))*f0[i];
        ok &= derivsAt(t0 + hd/3, ytmp.data(), Fapp, fa.data());
        for (int i=0;i<ny;++i) ytmp[i] = y0[i] + (h/T(6))*(f0[i] + fa[i]);
        ok &= derivsAt(t0 + hd/3, ytmp.data(), Fapp, fa.data());
        for (int i=0;i<ny;++i) ytmp[i] = y0[i] + (h/T(8))*(f0[i] + T(3)*fa[i]);
        ok &= derivsAt(t0 + hd/2, ytmp.data(), Fapp, fb.data());
        for (int i=0;i<ny;++i) ysave[i] = y0[i] + (h/T(2))*((f0[i] - T(3)*fa[i]) + T(4)*fb[i]);
        ok &= derivsAt(t1, ysave.data(), Fapp, fa.data());
        for (int i=0;i<ny;++i) y1[i] = y0[i] + (h/T(6))*((f0[i] + T(4)*fb[i]) + fa[i]);
        if (motion) prescribe(t1, y1.data());
        for (int i=0;i<ny;++i) yerr[i] = T(0.2)*T(fabs(double(y1[i]-ysave[i])));
        return ok;
    }

    // ---- event witnesses (null or empty: none) --------------------------
    const std::vector<WitnessDef>* wit = nullptr;
    std::vector<EventRecord>* events = nullptr;      // nullable: receives every triggered event
    std::vector<double> e0, eLow, eHigh, eMid;       // witness values: step start, window ends, trial
    std::vector<T> evF1, evY, wqd;
    // The witness values at (t, y); the station kinds realize the position
    // (and velocity) kinematics of y.
    void witnessValues(double t, const T* y, double* e) {
        bool pos = false, vel = false;
        for (const WitnessDef& w : *wit) {
            pos |= w.kind == WIT_STATION_POS || w.kind == WIT_STATION_VEL;
            vel |= w.kind == WIT_STATION_VEL;
        }
        if (pos) realizePosition(tr, y);
        if (vel) { wqd.resize(M.nq); realizeVelocity(tr, y, y + M.nq, wqd.data()); }
        for (size_t i = 0; i < wit->size(); ++i) {
            const WitnessDef& w = (*wit)[i];
            double v;
            if (w.kind == WIT_Q) v = double(y[w.index]);
            else if (w.kind == WIT_U) v = double(y[M.nq + w.index]);
            else if (w.kind == WIT_TIME) v = t;
            else {
                const Xform<T>& X = tr.B[w.body].X_GB;
                const V3<T> r = X.R * mk3<T>(T(w.station[0]), T(w.station[1]), T(w.station[2]));
                const V3<T> p = w.kind == WIT_STATION_POS ? X.p + r : stationVel(tr.B[w.body].V_GB, r);
                v = double((T(w.dir[0])*p[0] + T(w.dir[1])*p[1]) + T(w.dir[2])*p[2]);
            }
            e[i] = v - w.offset;
        }
    }
    // createInterpolatedState (AbstractIntegratorRep.cpp:63-90) between the
    // step's start (t0, y0, f0) and its end (t1, y1, evF1): the cubic Hermite
    // interpolant (interpolateOrder3, IntegratorRep.h:299-313) at t, then the
    // prescribed q and u at t and the LocalOnly projection at the constraint
    // tolerance (realizeAndProjectKinematicsWithThrow, IntegratorRep.h:
    // 858-897: prescribeQ, prescribeU; with
    // constraints projectQ then projectU, no projection limit, no error
    // estimate; without, the quaternions normalized when their RMS norm
    // error exceeds the tolerance).  False when the projection fails (the
    // reference throws).
    bool interpLocal(double t0, double t1, double t, double consTol, T* yt, bool backUp = false) {
        const int ny = M.nq + M.nu;
        const T h = T(t1 - t0), d = T((t - t0) / (t1 - t0));
        const T cy1 = d*d*(T(3) - T(2)*d), cy0 = T(1) - cy1;
        const T hdd1 = h*d*(d - T(1)), cf1 = hdd1*d, cf0 = cf1 - hdd1;
        for (int i=0;i<ny;++i) yt[i] = ((cy0*y0[i] + cy1*y1[i]) + cf0*f0[i]) + cf1*evF1[i];
        if (motion) prescribe(t, yt);       // prescribeQ, prescribeU of the interpolated state
        // setProjectInterpolatedStates(false) leaves interpolated states
        // unprojected, but not the advanced state backed up to an event
        // (AbstractIntegratorRep.cpp:78-85, 116-126)
        if (!backUp && !projInterp) return true;
        if (ws.cons) {
            bool ch;
            const double inf = std::numeric_limits<double>::infinity();
            return projectQ(yt, (T*)nullptr, consTol, inf, false, &ch) &&
                   projectU(yt, yt + M.nq, (T*)nullptr, consTol, inf, false, &ch);
        }
        if (M.nquat > 0 && (double(quatErrNorm(yt)) > consTol || forceProj)) normalizeQuats(yt, (T*)nullptr);
        return true;
    }
    // Events of the step just accepted from (t, y0, f0) to (t1, y1)
    // (takeOneStep's trigger search and localization, AbstractIntegratorRep.
    // cpp:583-756; the handlers as TimeStepper::stepTo calls them,
    // TimeStepper.cpp:178-205; Integrator.cpp:437-447).  e0 holds the
    // witnesses at the step's start.  The candidates over the whole step
    // decide which events can trigger; the window (tLow, tHigh] shrinks
    // around the earliest by trial states interpolated inside the step until
    // it is at most the narrowest required window; the advanced state is
    // then moved back to tHigh by interpolation, every remaining candidate is
    // recorded and its handler acts on that state.  A terminating handler
    // that changes nothing leaves the state interpolated at tLow (the state
    // ReachedEventTrigger reports); one that changes the state leaves it at
    // tHigh.  e0 then holds the witnesses of the state the next step starts
    // from.  Returns 0 (no event), 1 (events handled), 2 (terminate) or -1
    // (localization failed: no candidate in either half, no convergence in
    // kMaxLocalizeIters trials, or the projection of an interpolated state
    // failed).
    int stepEvents(double t, double& t1, const SV<T>* Fapp, const RkmConfig& cfg, double consTol) {
        const std::vector<WitnessDef>& W = *wit;
        const int nw = (int)W.size(), ny = M.nq + M.nu;
        const double acc = cfg.accuracy, ts = cfg.timescale;
        eHigh.resize(nw); eLow.resize(nw); eMid.resize(nw);
        witnessValues(t1, y1.data(), eHigh.data());
        const double minWindow = kSignificantReal * std::max(1.0, t1);
        EventCandidates C, N;
        findEventCandidates(W, nullptr, t, e0.data(), t1, eHigh.data(), 1.0, minWindow, acc, ts, C);
        if (C.idx.empty()) { e0 = eHigh; return 0; }
        double tLow = t, tHigh = t1;
        evF1.resize(ny); evY.resize(ny);
        if (tHigh - tLow > C.narrowest) {
            derivsAt(t1, y1.data(), Fapp, evF1.data());
            eLow = e0;
            double bias = 1;
            int side2 = 0, side1 = 0, it = 0;      // the half chosen two trials ago and at the last one
            do {
                if (++it > kMaxLocalizeIters) return -1;
                if (side2 != 0 && side1 != 0) bias = side2 != side1 ? 1.0 : (side1 < 0 ? bias/2 : bias*2);
                const double tMid = C.earliest;
                if (!interpLocal(t, t1, tMid, consTol, evY.data())) return -1;
                witnessValues(tMid, evY.data(), eMid.data());
                findEventCandidates(W, &C.idx, tLow, eLow.data(), tMid, eMid.data(), bias, minWindow, acc, ts, N);
                if (!N.idx.empty()) {
                    side2 = side1; side1 = -1;
                    tHigh = tMid; eHigh = eMid; C = N;
                    continue;
                }
                findEventCandidates(W, &C.idx, tMid, eMid.data(), tHigh, eHigh.data(), bias, minWindow, acc, ts, N);
                if (N.idx.empty()) return -1;
                side2 = side1; side1 = 1;
                tLow = tMid; eLow = eMid; C = N;
            } while (tHigh - tLow > C.narrowest);
            if (tHigh < t1) {
                if (!interpLocal(t, t1, tHigh, consTol, evY.data(), true)) return -1;
                for (int i=0;i<ny;++i) y1[i] = evY[i];
                t1 = tHigh;
            }
        }
        bool modified = false, term = false;
        for (size_t k = 0; k < C.idx.size(); ++k) {
            const WitnessDef& w = W[C.idx[k]];
            if (events) events->push_back({tLow, tHigh, C.idx[k], C.trans[k]});
            if (w.action == EV_SCALE_U) {
                for (int i = 0; i < M.nu; ++i)
                    if (w.aindex < 0 || w.aindex == i) y1[M.nq + i] = T(w.factor) * y1[M.nq + i];
                modified = true;
            } else if (w.action == EV_TERMINATE) term = true;
        }
        if (term && !modified) {
            derivsAt(t1, y1.data(), Fapp, evF1.data());
            if (!interpLocal(t, t1, tLow, consTol, evY.data())) return -1;
            for (int i=0;i<ny;++i) y1[i] = evY[i];
            t1 = tLow;
        }
        if (term) return 2;
        witnessValues(t1, y1.data(), e0.data());
        return 1;
    }

    static bool isFiniteD(double x) { return std::isfinite(x); }

    // Integrate y (q then u) from t0 to T.  Returns stats.
    RkmStats integrate(T* y, double t0, double Tend, const RkmConfig& cfg, const SV<T>* Fapp) {
        return integrateStops(y, t0, &Tend, 1, cfg, Fapp, nullptr);
    }

    // Integrate through the report times stops[0..n) (increasing; the last is
    // the final time), as TimeStepper::stepTo(stops[k]) in turn with
    // Integrator::setAllowInterpolation(false) (AbstractIntegratorRep.cpp:
    // 216-240: tMax = the report time; takeOneStep :512-530 stops each step
    // there): no step crosses a report time, and the step size carries from
    // one report interval to the next.  yout (nullable, n rows of nq+nu)
    // receives the state at each report time.
    //
    // reports (nullable, nrep increasing times in [t0, last stop]): Simbody's
    // default reporting with interpolation on (AbstractIntegratorRep.cpp:
    // 63-93, 310-330): steps ignore the report times; after each accepted
    // step every report time r in (t_prev, t] gets the step's end state if
    // r == t, else the cubic Hermite interpolant of (t_prev, y_prev, f_prev)
    // and (t, y, f(y)) (interpolateOrder3, IntegratorRep.h:285-313), whose
    // quaternions are then normalized when their RMS norm error exceeds the
    // constraint tolerance (projectQ on the interpolated state,
    // SimbodyMatterSubsystemRep.cpp:4151-4160).  rout: nrep rows.
    RkmStats integrateStops(T* y, double t0, const double* stops, int nstops, const RkmConfig& cfg,
                            const SV<T>* Fapp, T* yout, const double* reports = nullptr, int nrep = 0,
                            T* rout = nullptr) {
        RkmStats st;
        int k = 0, kr = 0;
        const int ny = M.nq + M.nu;
        const double acc = cfg.accuracy;
        const double consTol = acc/10;
        double hmin = cfg.min_h, hmax = cfg.max_h, hinit = cfg.init_h;
        if (cfg.fixed_h > 0) hmin = hmax = hinit = cfg.fixed_h;
        // initialize(): prescribed q and u at t0, forced projection of the
        // quaternions, then derivatives.
        if (motion) prescribe(t0, y);
        const bool hasCons = cons && !cons->empty();
        ws.cons = hasCons ? cons : nullptr;
        if (hasCons) {
            // initialize(): projectQ and projectU with ForceProjection and
            // ForceFullNewton (Integrator.cpp:376-377)
            bool ch;
            if (!projectQ(y, (T*)nullptr, consTol, std::numeric_limits<double>::infinity(), true, &ch) ||
                !projectU(y, y + M.nq, (T*)nullptr, consTol, std::numeric_limits<double>::infinity(), true, &ch)) {
                st.status = 5;
                return st;
            }
        } else {
            normalizeQuats(y, (T*)nullptr);
        }
        for (int i=0;i<ny;++i) y0[i] = y[i];
        double t = t0;
        st.t = t0;
        const bool hasWit = wit && !wit->empty();
        if (hasWit) { e0.resize(wit->size()); witnessValues(t0, y, e0.data()); }
        double hcur = (hinit != -1) ? hinit : cfg.timescale/10;
        if (hmin != -1) hcur = std::max(hcur, hmin);
        if (hmax != -1) hcur = std::min(hcur, hmax);
        const double projectionLimit = std::max(2*consTol, std::sqrt(consTol));
        for (; kr < nrep && reports[kr] <= t; ++kr)
            for (int i=0;i<ny;++i) rout[(size_t)kr*ny + i] = y[i];
        for (;;) {
            while (k < nstops && t >= stops[k]) {
                if (yout) for (int i=0;i<ny;++i) yout[(size_t)k*ny + i] = y[i];
                ++k;
            }
            if (k == nstops) break;
            // saveTimeAndStateAsPrevious + realizeStateDerivatives
            for (int i=0;i<ny;++i) y0[i] = y[i];
            for (int i=0;i<M.nu;++i) {
                const double vi = fabs(double(y0[M.nq+i]));
                uscale[i] = vi*1.0 > 1.0 ? T(1.0/vi) : T(1.0);
            }
            if (!derivsAt(t, y0.data(), Fapp, f0.data())) { st.status = 2; break; }
            ++st.evals;
            const double tMax = stops[k];
            bool succeeded = false;
            double t1 = t;
            do {
                bool limited = false;
                if (tMax < t + 0.95*hcur) { limited = true; t1 = tMax; }
                else if (tMax > t + 1.001*hcur) t1 = t + hcur;
                else t1 = tMax;
                if (!(t1 > t)) { st.status = 3; return st; }
                if (++st.attempts > cfg.max_attempts) { st.status = 4; return st; }
                bool converged = attemptODE(t, t1, Fapp);
                st.evals += 4;
                double errNorm;
                double errPre = std::numeric_limits<double>::quiet_NaN(), qErr = -1.0;
                bool projected = false;
                if (converged) {
                    const double e1 = double(errorNorm(y1.data(), yerr.data()));
                    errPre = e1;
                    if (!(e1 > 16.0*acc)) {   // NaN falls through to projection like the reference's comparison
                        const double qe = double(quatErrNorm(y1.data()));
                        qErr = qe;
                        if (hasCons) {
                            // attemptDAEStep (AbstractIntegratorRep.cpp:183-207)
                            bool chq = false, chu = false;
                            if (!projectQ(y1.data(), yerr.data(), consTol, projectionLimit, false, &chq) ||
                                !projectU(y1.data(), y1.data() + M.nq, yerr.data() + M.nq, consTol, projectionLimit,
                                          false, &chu))
                                converged = false;
                            if (chq) { ++st.projections; projected = true; }
                        } else if (qe > projectionLimit) converged = false;
                        else if (M.nquat > 0 && (qe > consTol || forceProj)) {
                            normalizeQuats(y1.data(), yerr.data()); ++st.projections; projected = true;
                        }
                    }
                }
                errNorm = converged ? double(errorNorm(y1.data(), yerr.data()))
                                    : std::numeric_limits<double>::infinity();
                // adjustStepSize
                const double Safety=0.9, MinShrink=0.1, MaxGrow=5, HysLow=0.9, HysHigh=1.2;
                double hnew;
                if (!isFiniteD(errNorm)) hnew = MinShrink*hcur;
                else if (errNorm == 0) hnew = MaxGrow*hcur;
                else hnew = Safety*hcur*std::pow(acc/errNorm, 1.0/4);
                if (hnew > hcur) { if (limited || hnew < HysHigh*hcur) hnew = hcur; }
                if (hnew < hcur) { if (errNorm <= acc) hnew = hcur; else hnew = std::min(hnew, HysLow*hcur); }
                hnew = std::min(hnew, MaxGrow*hcur);
                hnew = std::max(hnew, MinShrink*hcur);
                if (hmin != -1) hnew = std::max(hnew, hmin);
                if (hmax != -1) hnew = std::min(hnew, hmax);
                succeeded = (hnew >= hcur);
                hcur = hnew;
                if (!succeeded) ++st.rejected;
                if (trace) trace->push_back({t, t1, errPre, qErr, errNorm, projected ? 1.0 : 0.0,
                                             succeeded ? 1.0 : 0.0, hnew});
            } while (!succeeded);
            if (hasWit) {
                const int r = stepEvents(t, t1, Fapp, cfg, consTol);
                if (r < 0) { st.status = 6; return st; }
                if (r == 2) {
                    t = t1;
                    for (int i=0;i<ny;++i) y[i] = y1[i];
                    ++st.accepted;
                    st.terminated = 1; st.t = t;
                    return st;
                }
            }
            if (kr < nrep && reports[kr] <= t1) {
                std::vector<T> f1(ny);
                derivsAt(t1, y1.data(), Fapp, f1.data());
                for (; kr < nrep && reports[kr] <= t1; ++kr) {
                    T* yt = rout + (size_t)kr*ny;
                    if (reports[kr] == t1) { for (int i=0;i<ny;++i) yt[i] = y1[i]; continue; }
                    const T h = T(t1 - t), d = T((reports[kr] - t) / (t1 - t));
                    const T cy1 = d*d*(T(3) - T(2)*d), cy0 = T(1) - cy1;
                    const T hdd1 = h*d*(d - T(1)), cf1 = hdd1*d, cf0 = cf1 - hdd1;
                    for (int i=0;i<ny;++i) yt[i] = ((cy0*y0[i] + cy1*y1[i]) + cf0*f0[i]) + cf1*f1[i];
                    if (projInterp && M.nquat > 0 && (double(quatErrNorm(yt)) > consTol || forceProj))
                        normalizeQuats(yt, (T*)nullptr);
                }
            }
            t = t1;
            for (int i=0;i<ny;++i) y[i] = y1[i];
            ++st.accepted;
        }
        st.t = t;
        return st;
    }
};

} // namespace oracle
