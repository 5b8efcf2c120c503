// tdyn_matter.hpp -- C++ drop-in over the tdyn C ABI (include/tdyn.h).
//
// `tdyn::MatterBatch` carries the operator names and argument meaning of
// Simbody's SimbodyMatterSubsystem (Simbody/include/simbody/internal/
// SimbodyMatterSubsystem.h) for a batch of independent tree instances; the
// integrator half mirrors RungeKuttaMersonIntegrator + TimeStepper.  Errors
// are thrown as tdyn::Error (the reference throws SimTK exceptions from the
// same places, SimbodyMatterSubsystem.cpp:247-261, 438-442).  All array
// arguments are device pointers in the packed layout of tdyn.h.
#pragma once
#include "tdyn.h"

#include <stdexcept>
#include <string>
#include <vector>

namespace tdyn {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }
private:
    int code_;
};

inline void check(int rc) {
    if (rc != TDYN_OK) throw Error(rc, tdyn_last_error());
}

// RungeKuttaMersonIntegrator settings (Integrator.cpp:243-270).
struct RungeKuttaMerson {
    tdyn_integrator_cfg cfg{1e-3, 0, 0, 0, 0, 0, 0.1, 0, 0};
    RungeKuttaMerson& setAccuracy(double a) { cfg.accuracy = a; return *this; }
    RungeKuttaMerson& setConstraintTolerance(double c) { cfg.constraint_tolerance = c; return *this; }
    RungeKuttaMerson& setFixedStepSize(double h) { cfg.fixed_step = h; return *this; }
    RungeKuttaMerson& setInitialStepSize(double h) { cfg.initial_step = h; return *this; }
    RungeKuttaMerson& setMinimumStepSize(double h) { cfg.min_step = h; return *this; }
    RungeKuttaMerson& setMaximumStepSize(double h) { cfg.max_step = h; return *this; }
    // Integrator.cpp:276-296 (FP64; the chain evaluator and chain server refuse the first two)
    RungeKuttaMerson& setUseInfinityNorm(bool on) { cfg.use_infinity_norm = on ? 1 : 0; return *this; }
    RungeKuttaMerson& setProjectEveryStep(bool on) { cfg.project_every_step = on ? 1 : 0; return *this; }
    RungeKuttaMerson& setProjectInterpolatedStates(bool on) { cfg.project_interpolated_states = on ? 0 : -1; return *this; }
};

class MatterBatch {
public:
    MatterBatch(const std::vector<tdyn_tree_desc>& trees, const std::vector<int32_t>& treeOfInstance) {
        check(tdyn_batch_create(trees.data(), (int)trees.size(),
                                treeOfInstance.empty() ? nullptr : treeOfInstance.data(),
                                (int)treeOfInstance.size(), &h_));
    }
    MatterBatch(const tdyn_tree_desc& tree, int nInstances) {
        check(tdyn_batch_create(&tree, 1, nullptr, nInstances, &h_));
    }
    ~MatterBatch() { if (h_) tdyn_batch_destroy(h_); }
    MatterBatch(const MatterBatch&) = delete;
    MatterBatch& operator=(const MatterBatch&) = delete;

    tdyn_batch* handle() const { return h_; }
    void setStream(hipStream_t s) { s_ = s; }
    // After each operator that forms articulated-body inertias, synchronize
    // the stream and throw tdyn::Error(TDYN_E_SINGULAR) when some instance's
    // D was singular, as the reference throws (SmallMatrixMixed.h:878-892).
    // On by default; off keeps the calls asynchronous (query
    // checkSingular() or tdyn_batch_op_status yourself).
    void setThrowOnSingular(bool on) { throwOnSingular_ = on; }
    void checkSingular() const { int n = 0; check(tdyn_batch_op_status(h_, nullptr, &n, s_)); }

    // sizes: instances, total q, total u, total bodies incl. Ground, max nq, max nu, max nb, trees
    std::vector<int64_t> info() const { std::vector<int64_t> v(8); check(tdyn_batch_info(h_, v.data())); return v; }

    // realizePositionKinematics + realizeVelocityKinematics, then X_GB, V_GB, qdot, [KE, PE]
    void realizeKinematics(const double* q, const double* u, double* X_GB, double* V_GB, double* qdot,
                           double* energies) const {
        check(tdyn_realize_kinematics(h_, q, u, X_GB, V_GB, qdot, energies, s_));
    }
    // System::realize(state, Stage::Acceleration) with gravity (+ constant applied forces)
    void realizeAcceleration(const double* q, const double* u, const double* appliedBodyForces, double* udot,
                             double* qdot = nullptr) const {
        check(tdyn_realize_acceleration(h_, q, u, appliedBodyForces, udot, qdot, s_));
        if (throwOnSingular_) checkSingular();
    }
    void calcAccelerationIgnoringConstraints(const double* q, const double* u, const double* appliedMobilityForces,
                                             const double* appliedBodyForces, double* udot, double* A_GB,
                                             double* qdotdot = nullptr) const {
        check(tdyn_calc_acceleration_ignoring_constraints(h_, q, u, appliedMobilityForces, appliedBodyForces,
                                                          udot, A_GB, qdotdot, s_));
        if (throwOnSingular_) checkSingular();
    }
    void multiplyByMInv(const double* q, const double* v, double* MinvV) const {
        check(tdyn_multiply_by_m_inv(h_, q, v, MinvV, s_));
        if (throwOnSingular_) checkSingular();
    }
    void calcResidualForceIgnoringConstraints(const double* q, const double* u, const double* appliedMobilityForces,
                                              const double* appliedBodyForces, const double* knownUdot,
                                              double* residualMobilityForces) const {
        check(tdyn_calc_residual_force_ignoring_constraints(h_, q, u, appliedMobilityForces, appliedBodyForces,
                                                            knownUdot, residualMobilityForces, s_));
    }
    // realizeArticulatedBodyInertias + getArticulatedBodyInertia for every body
    void realizeArticulatedBodyInertias(const double* q, double* P, double* PPlus = nullptr) const {
        check(tdyn_calc_articulated_body_inertias(h_, q, P, PPlus, s_));
        if (throwOnSingular_) checkSingular();
    }
    void calcMobilizerReactionForces(const double* q, const double* u, const double* appliedMobilityForces,
                                     const double* appliedBodyForces, double* FM_G) const {
        check(tdyn_calc_mobilizer_reaction_forces(h_, q, u, appliedMobilityForces, appliedBodyForces, FM_G, s_));
        if (throwOnSingular_) checkSingular();
    }
    // Forward dynamics with known udot on the bodies flagged in `known`
    // (prescribed motion at the acceleration level, locks; no single
    // SimbodyMatterSubsystem call: what realize(Acceleration) does with
    // Motion::Acceleration mobilizers).  See tdyn_calc_acceleration_with_known_udot.
    void calcAccelerationWithKnownUDot(const double* q, const double* u, const double* fMob, const double* F,
                                       const int32_t* known, const double* udotKnown, double* udot, double* tau,
                                       double* A_GB = nullptr) const {
        check(tdyn_calc_acceleration_with_known_udot(h_, q, u, fMob, F, known, udotKnown, udot, tau, A_GB, s_));
        if (throwOnSingular_) checkSingular();
    }
    // Constraint::Ball / Constraint::Rod on tree `tree` (tdyn_batch_set_constraints).
    void setConstraints(int tree, int n, const tdyn_constraint_desc* c) const {
        check(tdyn_batch_set_constraints(h_, tree, n, c));
    }
    // realize(Acceleration) with constraints and getConstraintMultipliers
    // (TDYN_MAX_CONSTRAINT_ROWS per instance).
    void calcConstrainedAcceleration(const double* q, const double* u, const double* appliedBodyForces,
                                     double* udot, double* multipliers) const {
        check(tdyn_calc_constrained_acceleration(h_, q, u, appliedBodyForces, udot, multipliers, s_));
        if (throwOnSingular_) checkSingular();
    }
    // TriggeredEventHandler / TriggeredEventReporter witnesses of tree `tree`
    // (tdyn_batch_set_witnesses) and the events recorded per instance.
    void setWitnesses(int tree, int n, const tdyn_witness_desc* w) const {
        check(tdyn_batch_set_witnesses(h_, tree, n, w));
    }
    void setEventCapacity(int capacity) const { check(tdyn_batch_set_event_capacity(h_, capacity)); }
    // State::getEventTriggers after realize: TDYN_MAX_WITNESSES per instance
    void calcEventTriggers(const double* q, const double* u, const double* t, double* values) const {
        check(tdyn_calc_event_triggers(h_, q, u, t, values, s_));
    }
    void multiplyByM(const double* q, const double* v, double* Mv) const {
        check(tdyn_multiply_by_m(h_, q, v, Mv, s_));
    }
    void multiplyBySystemJacobian(const double* q, const double* v, double* Jv) const {
        check(tdyn_multiply_by_system_jacobian(h_, q, v, Jv, s_));
    }
    void multiplyBySystemJacobianTranspose(const double* q, const double* F_G, double* f) const {
        check(tdyn_multiply_by_system_jacobian_transpose(h_, q, F_G, f, s_));
    }
    void calcBodyAccelerationFromUDot(const double* q, const double* u, const double* knownUDot,
                                      double* A_GB) const {
        check(tdyn_calc_body_acceleration_from_udot(h_, q, u, knownUDot, A_GB, s_));
    }
    void calcCompositeBodyInertias(const double* q, double* R) const {
        check(tdyn_calc_composite_body_inertias(h_, q, R, s_));
    }

    // TimeStepper(system, integ).initialize(state); stepTo(tEnd) for every instance.
    void stepTo(double* q, double* u, const double* appliedBodyForces, double t0, double tEnd,
                const RungeKuttaMerson& integ, int32_t* stats = nullptr) const {
        check(tdyn_rollout(h_, q, u, appliedBodyForces, t0, tEnd, &integ.cfg, stats, s_));
    }
    // stepTo(tEnd) with reports by interpolation at reportTimes (host array)
    void stepToWithReports(double* q, double* u, const double* appliedBodyForces, double t0, double tEnd,
                           const double* reportTimes, int nReports, const RungeKuttaMerson& integ,
                           double* qReports, double* uReports, int32_t* stats = nullptr) const {
        check(tdyn_rollout_reports(h_, q, u, appliedBodyForces, t0, tEnd, reportTimes, nReports, &integ.cfg,
                                   qReports, uReports, stats, s_));
    }
    // resumable form
    void initialize(const double* q, const double* u, const double* appliedBodyForces, double t0,
                    const double* tEnd, double tEndAll, const RungeKuttaMerson& integ) const {
        check(tdyn_rollout_begin(h_, q, u, appliedBodyForces, t0, tEnd, tEndAll, &integ.cfg, s_));
    }
    void advance(int maxSteps) const { check(tdyn_rollout_advance(h_, maxSteps, s_)); }
    // stepTo(next report time) with Integrator::setAllowInterpolation(false)
    void extendTo(const double* tEnd, double tEndAll) const { check(tdyn_rollout_extend(h_, tEnd, tEndAll, s_)); }
    void getState(double* q, double* u, double* t, int32_t* status, int32_t* stats) const {
        check(tdyn_rollout_get(h_, q, u, t, status, stats, s_));
    }
    // checkpoint / resume: the rollout's integrator state as bytes
    // (tdyn_rollout_snapshot layout), and a rollout begun from such bytes
    std::vector<unsigned char> snapshot() const {
        int64_t n = 0;
        check(tdyn_rollout_snapshot(h_, nullptr, 0, &n, s_));
        std::vector<unsigned char> out((size_t)n);
        check(tdyn_rollout_snapshot(h_, out.data(), n, &n, s_));
        return out;
    }
    void restore(const std::vector<unsigned char>& snap, const double* appliedBodyForces = nullptr) const {
        check(tdyn_rollout_restore(h_, snap.data(), (int64_t)snap.size(), appliedBodyForces, s_));
    }
    // the events of the resumable form's rollout: windows (tLow, tHigh), ids
    // (witness, transition) and counts per instance
    void getEvents(double* windows, int32_t* ids, int32_t* counts) const {
        check(tdyn_rollout_get_events(h_, windows, ids, counts, s_));
    }

private:
    tdyn_batch* h_ = nullptr;
    hipStream_t s_ = nullptr;
    bool throwOnSingular_ = true;
};

} // namespace tdyn
