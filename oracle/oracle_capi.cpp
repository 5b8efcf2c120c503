// C entry points of the oracle (test infrastructure; see tree_oracle.hpp).
// Loaded with ctypes by tests/ and by bench.py's cpu_baseline leg only.
#include "flop_count.hpp"
#include "tree_oracle.hpp"

#include <atomic>
#include <chrono>
#include <pthread.h>
#include <sched.h>
#include <time.h>
#include <memory>
#include <thread>
#include <vector>

using namespace oracle;

namespace {
struct Handle {
    Model m;
    std::unique_ptr<Tree<double>> tr;
    std::vector<ConsDef> cons;   // holonomic constraints of the model
    std::vector<WitnessDef> wit; // event witnesses of the model
    std::vector<EventRecord> events;   // the events of the last integration
    RkmStats last;                     // and its statistics (final time, termination)
    bool inf_norm = false, project_every_step = false, project_interpolated = true;   // integrator options
};
// Attaches the model's constraints, witnesses and integrator options to an integrator.
void attach(Handle* h, Rkm<double>& rk) {
    rk.infNorm = h->inf_norm; rk.forceProj = h->project_every_step; rk.projInterp = h->project_interpolated;
    if (!h->cons.empty()) rk.cons = &h->cons;
    h->events.clear();
    if (!h->wit.empty()) { rk.wit = &h->wit; rk.events = &h->events; }
}
inline Handle* H(void* p) { return static_cast<Handle*>(p); }

void loadForces(const Model& M, const double* F, std::vector<SV<double>>& out) {
    out.assign(M.nb+1, svzero<double>());
    if (!F) return;
    for (int b = 0; b <= M.nb; ++b)
        out[b] = sv(mk3(F[6*b], F[6*b+1], F[6*b+2]), mk3(F[6*b+3], F[6*b+4], F[6*b+5]));
}
void storeSV(const SV<double>& v, double* o) { for (int i=0;i<3;++i) { o[i]=v.a[i]; o[3+i]=v.b[i]; } }
void storeSym(const M33<double>& A, double* o) {   // xx yy zz xy xz yz
    o[0]=A(0,0); o[1]=A(1,1); o[2]=A(2,2); o[3]=A(1,0); o[4]=A(2,0); o[5]=A(2,1);
}
void storeABI(const ABI<double>& P, double* o) {
    storeSym(P.M, o); storeSym(P.J, o+6); for (int i=0;i<9;++i) o[12+i] = P.F.m[i];
}
void realizeAll(Handle* h, const double* q, const double* u) {
    realizePosition(*h->tr, q);
    if (u) realizeVelocity(*h->tr, q, u, (double*)nullptr);
}
} // namespace

extern "C" {

void* oracle_model_new(const TreeDesc* d) {
    auto* h = new Handle;
    if (!h->m.build(*d)) { delete h; return nullptr; }
    h->tr.reset(new Tree<double>(h->m));
    return h;
}
void oracle_model_free(void* p) { delete H(p); }

// Holonomic constraints of the model (Constraint::Ball, Constraint::Rod):
// per constraint ib[3] = type (0 Ball, 1 Rod), body 1, body 2 (0 = Ground)
// and dd[7] = station on body 1, station on body 2 (body frames), length.
// Returns -1 for an invalid description (bodies out of range or equal, an
// unknown type, a non-positive rod length, Euler-angle models).
int oracle_set_constraints(void* p, int n, const int* ib, const double* dd) {
    Handle* h = H(p); const Model& M = h->m;
    std::vector<ConsDef> cs;
    for (int i = 0; i < n; ++i) {
        ConsDef c;
        c.type = ib[3*i]; c.b1 = ib[3*i+1]; c.b2 = ib[3*i+2];
        for (int k = 0; k < 3; ++k) { c.s1[k] = dd[7*i+k]; c.s2[k] = dd[7*i+3+k]; }
        c.len = dd[7*i+6];
        if ((c.type != CONS_BALL && c.type != CONS_ROD) || c.b1 < 0 || c.b2 < 0 || c.b1 > M.nb || c.b2 > M.nb ||
            c.b1 == c.b2 || (c.type == CONS_ROD && !(c.len > 0)) || M.euler)
            return -1;
        cs.push_back(c);
    }
    h->cons = cs;
    return 0;
}
// Loop forward dynamics (gravity + Fapp, (nb+1)*6 nullable): udot [nu] and
// the multipliers lam [m] (forces in G).  Returns 1 when G M^-1 G^T is
// singular.
int oracle_loop_forward(void* p, const double* q, const double* u, const double* Fapp, double* udot, double* lam) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizePosition(tr, q);
    realizeVelocity(tr, q, u, (double*)nullptr);
    std::vector<SV<double>> F(M.nb+1, svzero<double>());
    addGravity(tr, F.data());
    if (Fapp) for (int b=1;b<=M.nb;++b) F[b] = F[b] + sv(mk3(Fapp[6*b], Fapp[6*b+1], Fapp[6*b+2]),
                                                        mk3(Fapp[6*b+3], Fapp[6*b+4], Fapp[6*b+5]));
    realizeABI(tr);
    realizeABVelocity(tr);
    return loopForward(tr, h->cons, q, u, F.data(), udot, lam) ? 0 : 1;
}
// Constraint errors at (q, u): perr [m], verr [m] (either nullable), and the
// acceleration errors aerr [m] (nullable) of the body accelerations of udot.
// Event witnesses of the model (WitnessDef): per witness ib[6] = kind,
// index, body, trigger mask, action, action index and dd[9] = station (3),
// direction (3), offset, factor, window.  Later integrations (not the one
// with reports by interpolation) localize and handle their events.  Returns
// -1 for an invalid description.
// Integrator options of the model's later integrations: infinity norm
// (Integrator::setUseInfinityNorm), forced projection on every step
// (setProjectEveryStep), projection of interpolated states
// (setProjectInterpolatedStates, default on).
int oracle_set_integrator_options(void* p, int inf_norm, int project_every_step, int project_interpolated) {
    Handle* h = H(p);
    h->inf_norm = inf_norm != 0; h->project_every_step = project_every_step != 0;
    h->project_interpolated = project_interpolated != 0;
    return 0;
}
int oracle_set_witnesses(void* p, int n, const int* ib, const double* dd) {
    Handle* h = H(p); const Model& M = h->m;
    std::vector<WitnessDef> ws;
    for (int i = 0; i < n; ++i) {
        WitnessDef w;
        w.kind = ib[6*i]; w.index = ib[6*i+1]; w.body = ib[6*i+2]; w.trigger = ib[6*i+3];
        w.action = ib[6*i+4]; w.aindex = ib[6*i+5];
        for (int k = 0; k < 3; ++k) { w.station[k] = dd[9*i+k]; w.dir[k] = dd[9*i+3+k]; }
        w.offset = dd[9*i+6]; w.factor = dd[9*i+7]; w.window = dd[9*i+8];
        for (int k = 0; k < 9; ++k) if (!std::isfinite(dd[9*i+k])) return -1;
        if (w.kind < WIT_Q || w.kind > WIT_TIME || w.trigger < 1 || w.trigger > 3 ||
            w.action < EV_REPORT || w.action > EV_TERMINATE || !(w.window > 0)) return -1;
        if (w.kind == WIT_Q && (w.index < 0 || w.index >= M.nq)) return -1;
        if (w.kind == WIT_U && (w.index < 0 || w.index >= M.nu)) return -1;
        if ((w.kind == WIT_STATION_POS || w.kind == WIT_STATION_VEL) && (w.body < 0 || w.body > M.nb)) return -1;
        if (w.action == EV_SCALE_U && w.aindex >= M.nu) return -1;
        ws.push_back(w);
    }
    h->wit = ws;
    return 0;
}
// The witness values at (t, y).
int oracle_witness_values(void* p, double t, const double* y, double* e) {
    Handle* h = H(p);
    Rkm<double> rk(h->m);
    rk.wit = &h->wit;
    rk.witnessValues(t, y, e);
    return 0;
}
// The last integration's events: up to cap rows of (tLow, tHigh) into t and
// (witness, transition) into ids; returns how many it had.  *t_final and
// *terminated (nullable) get its final time and whether a handler ended it.
int oracle_last_events(void* p, double* t, int* ids, int cap, double* t_final, int* terminated) {
    Handle* h = H(p);
    const int n = (int)h->events.size();
    for (int i = 0; i < n && i < cap; ++i) {
        t[2*i] = h->events[i].tLow; t[2*i+1] = h->events[i].tHigh;
        ids[2*i] = h->events[i].witness; ids[2*i+1] = h->events[i].transition;
    }
    if (t_final) *t_final = h->last.t;
    if (terminated) *terminated = h->last.terminated;
    return n;
}
// estimateRootTime of the localization (IntegratorRep.h:351-376).
double oracle_estimate_root_time(double tLow, double fLow, double tHigh, double fHigh, double bias,
                                 double minWindow) {
    return estimateRootTime(tLow, fLow, tHigh, fHigh, bias, minWindow);
}

int oracle_constraint_errors(void* p, const double* q, const double* u, const double* udot, double* perr,
                             double* verr, double* aerr) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizePosition(tr, q);
    realizeVelocity(tr, q, u, (double*)nullptr);
    if (perr) consPerr(tr, h->cons, perr);
    std::vector<SV<double>> V(M.nb+1);
    for (int b=0;b<=M.nb;++b) V[b] = tr.B[b].V_GB;
    if (verr) consRowsOf(tr, h->cons, V.data(), verr);
    if (aerr) {
        std::vector<SV<double>> A(M.nb+1);
        bodyAccelerationFromUDot(tr, udot, A.data());
        consAerr(tr, h->cons, A.data(), aerr);
    }
    return 0;
}

int oracle_model_sizes(void* p, int* out /*nb nq nu nquat nlevels*/) {
    const Model& M = H(p)->m;
    out[0]=M.nb; out[1]=M.nq; out[2]=M.nu; out[3]=M.nquat; out[4]=(int)M.levels.size();
    return 0;
}
int oracle_model_layout(void* p, int* qIndex, int* uIndex, int* level) {
    const Model& M = H(p)->m;
    for (int b=1;b<=M.nb;++b) { qIndex[b-1]=M.qIndex[b]; uIndex[b-1]=M.uIndex[b]; level[b-1]=M.level[b]; }
    return 0;
}

// Position+velocity kinematics outputs.  X_GB: (nb+1)*12 (R row-major, p);
// V_GB: (nb+1)*6; qdot: nq; energies: [ke, pe].  Any output may be null.
int oracle_kinematics(void* p, const double* q, const double* u, double* X_GB, double* V_GB,
                      double* qdot, double* energies, double* qerr) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizePosition(tr, q);
    realizeVelocity(tr, q, u, qdot);
    for (int b=0;b<=M.nb;++b) {
        if (X_GB) { for (int i=0;i<9;++i) X_GB[12*b+i]=tr.B[b].X_GB.R.m[i]; for (int i=0;i<3;++i) X_GB[12*b+9+i]=tr.B[b].X_GB.p[i]; }
        if (V_GB) storeSV(tr.B[b].V_GB, V_GB+6*b);
    }
    if (energies) { energies[0] = kineticEnergy(tr); energies[1] = potentialEnergy(tr); }
    if (qerr) { int k=0; for (size_t L=1; L<M.levels.size(); ++L) for (int b : M.levels[L]) if (M.isQuat(b)) qerr[k++] = tr.B[b].qerr; }
    return 0;
}

// Gravity + optional applied body forces, (nb+1)*6.
int oracle_system_forces(void* p, const double* q, const double* Fapp, double* Fout) {
    Handle* h = H(p); const Model& M = h->m;
    realizePosition(*h->tr, q);
    std::vector<SV<double>> F; loadForces(M, Fapp, F);
    addGravity(*h->tr, F.data());
    for (int b=0;b<=M.nb;++b) storeSV(F[b], Fout+6*b);
    return 0;
}

// calcAccelerationIgnoringConstraints(state, fmob, F, udot, A_GB) + qdotdot.
int oracle_forward(void* p, const double* q, const double* u, const double* fmob, const double* Fb,
                   double* udot, double* qdotdot, double* A_GB) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizeAll(h, q, u);
    const bool ok = realizeABI(tr);
    realizeABVelocity(tr);
    std::vector<SV<double>> F; loadForces(M, Fb, F);
    forwardDynamics(tr, q, u, fmob, F.data(), udot, qdotdot);
    if (A_GB) for (int b=0;b<=M.nb;++b) storeSV(tr.B[b].A_GB, A_GB+6*b);
    return ok ? 0 : 1;
}

// Forward dynamics with known udot on the bodies marked in presc ([nb+1]);
// udot in/out, tau out (zero on free mobilities).
int oracle_forward_prescribed(void* p, const double* q, const double* u, const double* fmob, const double* Fb,
                              const int* presc, double* udot, double* tau, double* A_GB) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizeAll(h, q, u);
    const bool ok = realizeABI(tr, presc);
    realizeABVelocity(tr);
    std::vector<SV<double>> F; loadForces(M, Fb, F);
    forwardDynamicsPrescribed(tr, presc, fmob, F.data(), udot, tau);
    if (A_GB) for (int b=0;b<=M.nb;++b) storeSV(tr.B[b].A_GB, A_GB+6*b);
    return ok ? 0 : 1;
}

// realize(Acceleration): gravity + optional constant applied forces.
int oracle_realize_acceleration(void* p, const double* q, const double* u, const double* Fapp,
                                double* udot, double* qdot, double* qdotdot) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizePosition(tr, q);
    realizeVelocity(tr, q, u, qdot);
    std::vector<SV<double>> F; loadForces(M, Fapp, F);
    addGravity(tr, F.data());
    const bool ok = realizeABI(tr);
    realizeABVelocity(tr);
    forwardDynamics(tr, q, u, (const double*)nullptr, F.data(), udot, qdotdot);
    return ok ? 0 : 1;
}

int oracle_minv(void* p, const double* q, const double* f, double* out) {
    Handle* h = H(p); Tree<double>& tr = *h->tr;
    realizePosition(tr, q);
    const bool ok = realizeABI(tr);
    multiplyByMInv(tr, f, out);
    return ok ? 0 : 1;
}

int oracle_residual(void* p, const double* q, const double* u, const double* fmob, const double* Fb,
                    const double* udot, double* out) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizeAll(h, q, u);
    std::vector<SV<double>> F;
    if (Fb) loadForces(M, Fb, F);
    residualForces(tr, fmob, Fb ? F.data() : (const SV<double>*)nullptr, udot, out);
    return 0;
}

// Articulated-body quantities per body: P, PPlus (21 each: M sym6, J sym6,
// F 9), D and DI packed by cumulative dof^2.  Any output may be null.
int oracle_abi(void* p, const double* q, double* P, double* PPlus, double* D, double* DI) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizePosition(tr, q);
    const bool ok = realizeABI(tr);
    int off = 0;
    for (int b=1;b<=M.nb;++b) {
        if (P) storeABI(tr.B[b].P, P + 21*(b-1));
        if (PPlus) storeABI(tr.B[b].PPlus, PPlus + 21*(b-1));
        const int n2 = M.dof[b]*M.dof[b];
        for (int i=0;i<n2;++i) { if (D) D[off+i]=tr.B[b].D[i]; if (DI) DI[off+i]=tr.B[b].DI[i]; }
        off += n2;
    }
    return ok ? 0 : 1;
}

// multiplyByM: out = M v (nu).
int oracle_multiply_by_m(void* p, const double* q, const double* v, double* out) {
    Handle* h = H(p); Tree<double>& tr = *h->tr;
    realizePosition(tr, q);
    multiplyByM(tr, v, out);
    return 0;
}

// multiplyBySystemJacobian: Jv = J v, (nb+1)*6.
int oracle_jacobian(void* p, const double* q, const double* v, double* Jv) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizePosition(tr, q);
    std::vector<SV<double>> J(M.nb+1);
    systemJacobian(tr, v, J.data());
    for (int b=0;b<=M.nb;++b) storeSV(J[b], Jv+6*b);
    return 0;
}

// multiplyBySystemJacobianTranspose: out = J^T F, F (nb+1)*6.
int oracle_jacobian_transpose(void* p, const double* q, const double* F, double* out) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizePosition(tr, q);
    std::vector<SV<double>> Fs; loadForces(M, F, Fs);
    systemJacobianTranspose(tr, Fs.data(), out);
    return 0;
}

// calcBodyAccelerationFromUDot: A_GB (nb+1)*6; null udot means zeros.
int oracle_body_acceleration(void* p, const double* q, const double* u, const double* udot, double* A) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizeAll(h, q, u);
    std::vector<SV<double>> As(M.nb+1);
    bodyAccelerationFromUDot(tr, udot, As.data());
    for (int b=0;b<=M.nb;++b) storeSV(As[b], A+6*b);
    return 0;
}

// calcCompositeBodyInertias: R (nb+1)*10 per body: mass, mass center (3),
// unit inertia about the origin (xx yy zz xy xz yz), all in G.
int oracle_composite_inertias(void* p, const double* q, double* R) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizePosition(tr, q);
    std::vector<SpatialInertia<double>> Rs(M.nb+1);
    compositeBodyInertias(tr, Rs.data());
    for (int b=0;b<=M.nb;++b) {
        double* o = R + 10*b;
        o[0] = Rs[b].m;
        for (int i=0;i<3;++i) o[1+i] = Rs[b].p[i];
        storeSym(Rs[b].G, o+4);
    }
    return 0;
}

int oracle_reactions(void* p, const double* q, const double* u, const double* fmob, const double* Fb,
                     double* FM) {
    Handle* h = H(p); const Model& M = h->m; Tree<double>& tr = *h->tr;
    realizeAll(h, q, u);
    const bool ok = realizeABI(tr);
    realizeABVelocity(tr);
    std::vector<SV<double>> F; loadForces(M, Fb, F);
    std::vector<double> ud(M.nu);
    forwardDynamics(tr, q, u, fmob, F.data(), ud.data(), (double*)nullptr);
    std::vector<SV<double>> R(M.nb+1);
    reactionForces(tr, R.data());
    for (int b=0;b<=M.nb;++b) storeSV(R[b], FM+6*b);
    return ok ? 0 : 1;
}

// TimeStepper(stepTo T) with RungeKuttaMersonIntegrator, finalTime = T.
// y: [q; u] in/out.  stats: attempts accepted rejected projections evals status
int oracle_integrate(void* p, double* y, const double* Fapp, double t0, double T,
                     double accuracy, double fixed_h, long max_attempts, long* stats) {
    Handle* h = H(p); const Model& M = h->m;
    Rkm<double> rk(M);
    attach(h, rk);
    std::vector<SV<double>> F; loadForces(M, Fapp, F);
    RkmConfig cfg; cfg.accuracy = accuracy; cfg.fixed_h = fixed_h;
    if (max_attempts > 0) cfg.max_attempts = max_attempts;
    RkmStats st = rk.integrate(y, t0, T, cfg, Fapp ? F.data() : (const SV<double>*)nullptr);
    h->last = st;
    if (stats) { stats[0]=st.attempts; stats[1]=st.accepted; stats[2]=st.rejected; stats[3]=st.projections; stats[4]=st.evals; stats[5]=st.status; }
    return st.status;
}

// Integration to T with the integrator settings c[6] = accuracy, fixed
// step, initial, minimum and maximum step (-1: unset) and time scale
// (Integrator::setInitialStepSize, setMinimumStepSize, setMaximumStepSize;
// System::setDefaultTimeScale).
int oracle_integrate_cfg(void* p, double* y, const double* Fapp, double t0, double T, const double* c,
                         long* stats) {
    Handle* h = H(p); const Model& M = h->m;
    Rkm<double> rk(M);
    attach(h, rk);
    std::vector<SV<double>> F; loadForces(M, Fapp, F);
    RkmConfig cfg; cfg.accuracy = c[0]; cfg.fixed_h = c[1]; cfg.init_h = c[2]; cfg.min_h = c[3]; cfg.max_h = c[4];
    cfg.timescale = c[5];
    RkmStats st = rk.integrate(y, t0, T, cfg, Fapp ? F.data() : (const SV<double>*)nullptr);
    h->last = st;
    if (stats) { stats[0]=st.attempts; stats[1]=st.accepted; stats[2]=st.rejected; stats[3]=st.projections; stats[4]=st.evals; stats[5]=st.status; }
    return st.status;
}

// The oracle's articulated-inertia shift (abiShift, the P' = Phi P Phi^T of
// MassProperties.cpp:100-127) on one 21-double ABI (storeABI layout; the
// symmetric blocks are read from their lower triangle).
int oracle_abi_shift(const double* P, const double* s, double* out) {
    ABI<double> A;
    auto sym = [](const double* o) {
        M33<double> X;
        X(0,0)=o[0]; X(1,1)=o[1]; X(2,2)=o[2];
        X(1,0)=X(0,1)=o[3]; X(2,0)=X(0,2)=o[4]; X(2,1)=X(1,2)=o[5];
        return X;
    };
    A.M = sym(P); A.J = sym(P+6);
    for (int i=0;i<9;++i) A.F.m[i] = P[12+i];
    storeABI(abiShift(A, mk3(s[0], s[1], s[2])), out);
    return 0;
}

// Integration to T recording every step attempt (Rkm::trace): rec gets up to
// max_rec rows of 8 doubles (RkmAttempt), *n_rec the number of attempts.
// init_h > 0 sets Integrator::setInitialStepSize.
int oracle_integrate_trace(void* p, double* y, const double* Fapp, double t0, double T, double accuracy,
                           double fixed_h, double init_h, double* rec, int max_rec, int* n_rec, long* stats) {
    Handle* h = H(p); const Model& M = h->m;
    Rkm<double> rk(M);
    attach(h, rk);
    std::vector<RkmAttempt> tr;
    rk.trace = &tr;
    std::vector<SV<double>> F; loadForces(M, Fapp, F);
    RkmConfig cfg; cfg.accuracy = accuracy; cfg.fixed_h = fixed_h;
    if (init_h > 0) cfg.init_h = init_h;
    RkmStats st = rk.integrate(y, t0, T, cfg, Fapp ? F.data() : (const SV<double>*)nullptr);
    const int n = (int)std::min<size_t>(tr.size(), (size_t)std::max(0, max_rec));
    for (int i=0;i<n;++i) {
        const RkmAttempt& a = tr[i];
        const double r[8] = {a.t0, a.t1, a.err_pre, a.qerr, a.err, a.projected, a.accepted, a.h_next};
        std::copy(r, r + 8, rec + 8*(size_t)i);
    }
    *n_rec = (int)tr.size();
    h->last = st;
    if (stats) { stats[0]=st.attempts; stats[1]=st.accepted; stats[2]=st.rejected; stats[3]=st.projections; stats[4]=st.evals; stats[5]=st.status; }
    return st.status;
}

// Integration to T with prescribed motion: motion [nb+1][4] (Ground row
// first: level, amplitude, rate, phase; levels as MotionLevel).
int oracle_integrate_motion(void* p, double* y, const double* Fapp, const double* motion, double t0, double T,
                            double accuracy, double fixed_h, long* stats) {
    Handle* h = H(p); const Model& M = h->m;
    Rkm<double> rk(M);
    attach(h, rk);
    rk.motion = motion;
    std::vector<SV<double>> F; loadForces(M, Fapp, F);
    RkmConfig cfg; cfg.accuracy = accuracy; cfg.fixed_h = fixed_h;
    RkmStats st = rk.integrate(y, t0, T, cfg, Fapp ? F.data() : (const SV<double>*)nullptr);
    h->last = st;
    if (stats) { stats[0]=st.attempts; stats[1]=st.accepted; stats[2]=st.rejected; stats[3]=st.projections; stats[4]=st.evals; stats[5]=st.status; }
    return st.status;
}

// Integration through report times stops[0..n) (the last is the final time)
// with interpolation off: every step stops at the next report time; yout
// [n][nq+nu] gets the state there (Rkm::integrateStops).
int oracle_integrate_stops(void* p, double* y, const double* Fapp, double t0, const double* stops, int n,
                           double accuracy, double fixed_h, double* yout, long* stats) {
    Handle* h = H(p); const Model& M = h->m;
    Rkm<double> rk(M);
    attach(h, rk);
    std::vector<SV<double>> F; loadForces(M, Fapp, F);
    RkmConfig cfg; cfg.accuracy = accuracy; cfg.fixed_h = fixed_h;
    RkmStats st = rk.integrateStops(y, t0, stops, n, cfg, Fapp ? F.data() : (const SV<double>*)nullptr, yout);
    h->last = st;
    if (stats) { stats[0]=st.attempts; stats[1]=st.accepted; stats[2]=st.rejected; stats[3]=st.projections; stats[4]=st.evals; stats[5]=st.status; }
    return st.status;
}

// Integration to T with reports by interpolation (Simbody's default
// reporting): rout [n][nq+nu] gets the state at each report time
// (Rkm::integrateStops with reports); y gets the state at T.
int oracle_integrate_reports(void* p, double* y, const double* Fapp, double t0, double T, const double* reports,
                             int n, double accuracy, double fixed_h, double* rout, long* stats) {
    Handle* h = H(p); const Model& M = h->m;
    // (a report time inside an event's window changes the localization
    // trials, AbstractIntegratorRep.cpp:640-680; not restated)
    if (!h->wit.empty()) return -1;
    Rkm<double> rk(M);
    attach(h, rk);
    std::vector<SV<double>> F; loadForces(M, Fapp, F);
    RkmConfig cfg; cfg.accuracy = accuracy; cfg.fixed_h = fixed_h;
    RkmStats st = rk.integrateStops(y, t0, &T, 1, cfg, Fapp ? F.data() : (const SV<double>*)nullptr, nullptr,
                                    reports, n, rout);
    h->last = st;
    if (stats) { stats[0]=st.attempts; stats[1]=st.accepted; stats[2]=st.rejected; stats[3]=st.projections; stats[4]=st.evals; stats[5]=st.status; }
    return st.status;
}

// Integration with known udot on the bodies flagged in presc ([nb+1]): the
// values udot_known ([nu]) on their mobilities at every derivative.
int oracle_integrate_known(void* p, double* y, const double* Fapp, const int* presc, const double* udot_known,
                           double t0, double T, double accuracy, double fixed_h, long* stats) {
    Handle* h = H(p); const Model& M = h->m;
    Rkm<double> rk(M);
    attach(h, rk);
    rk.ws.presc = presc; rk.ws.udotKnown = udot_known;
    std::vector<SV<double>> F; loadForces(M, Fapp, F);
    RkmConfig cfg; cfg.accuracy = accuracy; cfg.fixed_h = fixed_h;
    RkmStats st = rk.integrate(y, t0, T, cfg, Fapp ? F.data() : (const SV<double>*)nullptr);
    h->last = st;
    if (stats) { stats[0]=st.attempts; stats[1]=st.accepted; stats[2]=st.rejected; stats[3]=st.projections; stats[4]=st.evals; stats[5]=st.status; }
    return st.status;
}

// Same integration with the counting scalar.  tallies: addsub mul div sqrt trig.
int oracle_count_flops_integrate(void* p, double* y, const double* Fapp, double t0, double T,
                                 double accuracy, double fixed_h, unsigned long long* tallies, long* stats) {
    Handle* h = H(p); const Model& M = h->m;
    Rkm<CReal> rk(M);
    const int ny = M.nq + M.nu;
    std::vector<CReal> yc(ny); for (int i=0;i<ny;++i) yc[i] = CReal(y[i]);
    std::vector<SV<CReal>> F(M.nb+1, svzero<CReal>());
    if (Fapp) for (int b=0;b<=M.nb;++b)
        F[b] = sv(mk3<CReal>(Fapp[6*b],Fapp[6*b+1],Fapp[6*b+2]), mk3<CReal>(Fapp[6*b+3],Fapp[6*b+4],Fapp[6*b+5]));
    RkmConfig cfg; cfg.accuracy = accuracy; cfg.fixed_h = fixed_h;
    tally() = FlopTally();
    RkmStats st = rk.integrate(yc.data(), t0, T, cfg, Fapp ? F.data() : (const SV<CReal>*)nullptr);
    const FlopTally t = tally();
    tallies[0]=t.addsub; tallies[1]=t.mul; tallies[2]=t.div; tallies[3]=t.sqrt; tallies[4]=t.trig;
    for (int i=0;i<ny;++i) y[i] = yc[i].v;
    h->last = st;
    if (stats) { stats[0]=st.attempts; stats[1]=st.accepted; stats[2]=st.rejected; stats[3]=st.projections; stats[4]=st.evals; stats[5]=st.status; }
    return st.status;
}

// Flop count of one realize(Acceleration) (forward dynamics incl. gravity)
// and of one inverse-dynamics call on the given state.
int oracle_count_flops_ops(void* p, const double* q, const double* u, unsigned long long* fwd, unsigned long long* inv) {
    Handle* h = H(p); const Model& M = h->m;
    Tree<CReal> tr(M);
    Workspace<CReal> ws;
    const int ny = M.nq + M.nu;
    std::vector<CReal> y(ny), yd(ny), r(M.nu);
    for (int i=0;i<M.nq;++i) y[i] = q[i];
    for (int i=0;i<M.nu;++i) y[M.nq+i] = u[i];
    tally() = FlopTally();
    derivs(tr, ws, y.data(), (const SV<CReal>*)nullptr, yd.data());
    FlopTally t = tally();
    fwd[0]=t.addsub; fwd[1]=t.mul; fwd[2]=t.div; fwd[3]=t.sqrt; fwd[4]=t.trig;
    tally() = FlopTally();
    realizePosition(tr, y.data());
    realizeVelocity(tr, y.data(), y.data()+M.nq, (CReal*)nullptr);
    residualForces(tr, (const CReal*)nullptr, (const SV<CReal>*)nullptr, yd.data()+M.nq, r.data());
    t = tally();
    inv[0]=t.addsub; inv[1]=t.mul; inv[2]=t.div; inv[3]=t.sqrt; inv[4]=t.trig;
    return 0;
}

// CPU baseline: integrate n_inst instances (one instance per thread at a
// time, std::thread pool) and return wall seconds.  Instance i uses
// models[model_of_inst[i]]; Y is instance-major with stride ystride, and
// Fapp (nullable) is instance-major with stride fstride.  out_body_steps =
// sum over instances of accepted steps * nb.  pin != 0: worker k runs on the
// k-th CPU of the process's affinity mask (1:1 while n_threads <= its size).
// out_cpu_s (nullable): the workers' summed thread CPU time
// (CLOCK_THREAD_CPUTIME_ID), the measure of TestMultibodyPerformance.cpp:
// 168-198.
double oracle_time_integrate_batch(void** models, const int* model_of_inst, int n_inst,
                                   double* Y, int ystride, const double* Fapp, int fstride,
                                   double t0, double T, double accuracy, double fixed_h,
                                   int n_threads, long long* out_body_steps, long long* out_attempts,
                                   int pin, double* out_cpu_s) {
    std::atomic<int> next{0};
    std::atomic<long long> bodySteps{0}, attempts{0};
    std::vector<int> cpus;
    if (pin) {
        cpu_set_t set;
        CPU_ZERO(&set);
        if (sched_getaffinity(0, sizeof set, &set) == 0)
            for (int c = 0; c < CPU_SETSIZE; ++c) if (CPU_ISSET(c, &set)) cpus.push_back(c);
    }
    std::vector<double> cpuSec(n_threads, 0.0);
    auto worker = [&](int k) {
        if (!cpus.empty()) {
            cpu_set_t one;
            CPU_ZERO(&one);
            CPU_SET(cpus[k % cpus.size()], &one);
            (void)pthread_setaffinity_np(pthread_self(), sizeof one, &one);
        }
        timespec c0, c1;
        clock_gettime(CLOCK_THREAD_CPUTIME_ID, &c0);
        std::vector<std::unique_ptr<Rkm<double>>> cache;
        std::vector<void*> cacheKey;
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n_inst) break;
            Handle* h = H(models[model_of_inst[i]]);
            const Model& M = h->m;
            Rkm<double>* rk = nullptr;
            for (size_t k=0;k<cacheKey.size();++k) if (cacheKey[k]==(void*)h) rk = cache[k].get();
            if (!rk) { cache.emplace_back(new Rkm<double>(M)); cacheKey.push_back(h); rk = cache.back().get(); }
            std::vector<SV<double>> F;
            if (Fapp) loadForces(M, Fapp + (size_t)i*fstride, F);
            RkmConfig cfg; cfg.accuracy = accuracy; cfg.fixed_h = fixed_h;
            RkmStats st = rk->integrate(Y + (size_t)i*ystride, t0, T, cfg, Fapp ? F.data() : (const SV<double>*)nullptr);
            bodySteps += (long long)st.accepted * M.nb;
            attempts += st.attempts;
        }
        clock_gettime(CLOCK_THREAD_CPUTIME_ID, &c1);
        cpuSec[k] = (c1.tv_sec - c0.tv_sec) + 1e-9 * (c1.tv_nsec - c0.tv_nsec);
    };
    const auto t_start = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (int k=0;k<n_threads;++k) th.emplace_back(worker, k);
    for (auto& x : th) x.join();
    const auto t_end = std::chrono::steady_clock::now();
    if (out_body_steps) *out_body_steps = bodySteps.load();
    if (out_attempts) *out_attempts = attempts.load();
    if (out_cpu_s) { double c = 0; for (double x : cpuSec) c += x; *out_cpu_s = c; }
    return std::chrono::duration<double>(t_end - t_start).count();
}

} // extern "C"
