/* tdyn: MI355X-native batched tree dynamics -- the C ABI.
 *
 * Drop-in boundary for Simbody's tree-dynamics operator surface
 * (SimbodyMatterSubsystem + RungeKuttaMersonIntegrator/TimeStepper) for
 * batches of independent, unconstrained multibody instances.  The reference
 * has no FFI: its surface is the C++ handle SimbodyMatterSubsystem used with a
 * State.  Each entry point below names the reference member it replaces
 * (paths relative to the Simbody source tree); the argument meaning, the
 * frames and the q/u orderings are the reference's:
 *   - bodies are numbered 1..nb in MobilizedBodyIndex order, Ground = 0;
 *   - q and u slots are assigned in body order (SimbodyMatterSubsystemRep.cpp:
 *     256-293); Ball/Free use quaternions (w,x,y,z) first, Free then the
 *     translation (RigidBodyNodeSpec_Free.h:178-186); Ball/Free u are the
 *     angular velocity in F then (Free) the linear velocity in F; Universal u
 *     are the Euler-angle rates (RigidBodyNodeSpec_Universal.h:51-53);
 *   - body forces are SpatialVec (torque, force), applied at the body origin,
 *     expressed in Ground, one per body including Ground
 *     (SimbodyMatterSubsystem.h:2210-2216);
 *   - per-instance arrays are packed instance after instance; use
 *     tdyn_batch_offsets() for where each instance starts.
 *
 * All data pointers are device pointers (hipMalloc'd, or a torch tensor's
 * data_ptr()); every call takes the HIP stream it enqueues on.  A handle is
 * not thread-safe: use one host thread per handle (the reference's State must
 * not be realized concurrently either, SimbodyMatterSubsystem.h:2749-2755).
 * All functions return TDYN_OK or an error code; tdyn_last_error() gives the
 * message of the last failure on the calling thread.
 */
#ifndef TDYN_H
#define TDYN_H
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes (the reference throws; SimbodyMatterSubsystem.cpp:247-261,
 * SmallMatrixMixed.h:878-892, AbstractIntegratorRep.cpp:426-431). */
#define TDYN_OK             0
#define TDYN_E_ARG          1   /* bad argument (SimTK_APIARGCHECK) */
#define TDYN_E_SIZE         2   /* size mismatch (SimTK_ERRCHK2) */
#define TDYN_E_STAGE        3   /* call order violation (SimTK_STAGECHECK) */
#define TDYN_E_SINGULAR     4   /* singular articulated inertia D */
#define TDYN_E_STEP_FAILED  5   /* Integrator::StepFailed for some instance */
#define TDYN_E_HIP          6   /* HIP runtime error */
#define TDYN_E_NOMEM        7

/* Mobilizer codes: MobilizedBody::Pin, Slider, Universal, Ball, Free,
 * Translation, Planar, Cylinder, Screw, Gimbal (RigidBodyNodeSpec_<Type>.h), Weld
 * (RigidBodyNode_Weld.cpp), BendStretch (RigidBodyNodeSpec_PolarCoords.h),
 * Bushing (RigidBodyNodeSpec_Bushing.h). */
#define TDYN_PIN         0
#define TDYN_SLIDER      1
#define TDYN_UNIVERSAL   2
#define TDYN_BALL        3
#define TDYN_FREE        4
#define TDYN_TRANSLATION 5   /* q = p_FM (3), u = v_FM in F (3) */
#define TDYN_PLANAR      6   /* q = (angle about z, x, y), u = their rates, translation in F */
#define TDYN_CYLINDER    7   /* q = (angle about z, z), u = their rates */
#define TDYN_SCREW       8   /* q = angle about z, translation pitch*q along z */
#define TDYN_GIMBAL      9   /* q = body-fixed x-y-z angles, u = qdot (singular at q1 = +-pi/2) */
#define TDYN_WELD       10   /* no q, no u: the body is fixed to its parent */
#define TDYN_BENDSTRETCH 11  /* q = (angle about z, slide along the rotated x), u = qdot; not reversible */
#define TDYN_BUSHING    12   /* q = body-fixed x-y-z angles then p_FM, u = qdot (6); on Ground only, not
                                reversible, not in the serving loop */

/* One tree (what MobilizedBody::<Type>(parent, X_PF, body, X_BM) calls build;
 * MobilizedBody.cpp:783-808).  Host pointers. */
typedef struct tdyn_tree_desc {
    int32_t nb;                /* mobilized bodies (Ground excluded) */
    const int32_t* parent;     /* [nb] parent body number, 0 = Ground, parent[b-1] < b */
    const int32_t* mobilizer;  /* [nb] TDYN_PIN .. TDYN_BUSHING (Bushing only on Ground) */
    const double* mass;        /* [nb] */
    const double* com_B;       /* [nb][3] mass center in B */
    const double* inertia_B;   /* [nb][6] inertia about Bo in B: xx yy zz xy xz yz */
    const double* X_PF;        /* [nb][12] inboard frame on the parent: R row-major, p */
    const double* X_BM;        /* [nb][12] outboard frame on the body */
    double gravity[3];         /* Force::Gravity direction*magnitude (Force_Gravity.cpp:210-229) */
    const double* screw_pitch; /* [nb] MobilizedBody::Screw pitch (length per radian), read for Screw
                                  bodies only; NULL: no Screw bodies */
    const int32_t* reversed;   /* [nb] nonzero: the mobilizer is reversed (MobilizedBody::Reverse,
                                  RigidBodyNodeSpec.h:262-277); not for Free; NULL: none reversed */
    int32_t use_euler_angles;  /* nonzero: SimbodyMatterSubsystem::setUseEulerAngles(true) -- Ball and
                                  Free q slots (4 and 7, unchanged) hold body-fixed x-y-z angles (Free:
                                  then the translation), the last slot unused; u keeps its meaning */
} tdyn_tree_desc;

typedef struct tdyn_batch tdyn_batch;

/* Build a batch: n_instances instances, instance i uses
 * trees[tree_of_instance[i]] (tree_of_instance NULL: all use trees[0]).
 * Replaces SimbodyMatterSubsystem construction + realizeTopology
 * (SimbodyMatterSubsystemRep.cpp:256-330). */
int tdyn_batch_create(const tdyn_tree_desc* trees, int n_trees, const int32_t* tree_of_instance,
                      int n_instances, tdyn_batch** out);
int tdyn_batch_destroy(tdyn_batch* b);

/* Derivative evaluator of the rollouts (no reference counterpart; a tuning
 * knob): 0 = automatic (the register-resident serial-chain evaluator when
 * every instance shares one chain model; the level-parallel tree evaluator
 * for a batch of different trees whose only Free body, if any, is body 1;
 * else the tree sweeps), 1 = tree sweeps, 2 = chain evaluator (TDYN_E_ARG
 * unless the batch is one chain model), 3 = the level-parallel tree
 * evaluator (falls back to the tree sweeps where the batch does not
 * qualify).  Takes effect at the next tdyn_rollout_begin. */
int tdyn_batch_set_evaluator(tdyn_batch* b, int mode);
/* Lanes per instance of the tree evaluators of a batch of different trees
 * (no reference counterpart; a tuning knob): 0 = automatic (the level-
 * parallel evaluator with, per instance, its tree's widest level rounded up
 * to a power of two, clamped to the lane range below), 1 = the streaming
 * evaluator (one lane walks the tree depth-first), 4, 8, 16, 32, 64 = the
 * level-parallel evaluator with that many lanes for every instance.  Takes
 * effect at the next tdyn_rollout_begin. */
int tdyn_batch_set_tree_lanes(tdyn_batch* b, int lanes);
/* Lane range of the automatic setting (no reference counterpart; a tuning
 * knob): lanes_min 0 (default: 4 to 16 from the batch size, about four
 * wavefronts per SIMD) or 4..64, lanes_max 4..64 (default 32); powers of two. */
int tdyn_batch_set_tree_lane_range(tdyn_batch* b, int lanes_min, int lanes_max);
/* The evaluator of the current rollout (no reference counterpart; for tests
 * and tuning): *active = 0 before the first tdyn_rollout_begin, else 1 tree
 * sweeps, 2 chain evaluator, 3 streaming tree evaluator, 4 level-parallel
 * tree evaluator; *lanes (nullable) = lanes per instance (0: chosen per
 * instance). */
int tdyn_batch_get_evaluator(const tdyn_batch* b, int* active, int* lanes);

/* Per-instance status of the last single-shot operator that formed the
 * articulated-body inertias (realize acceleration, calcAcceleration*,
 * calcAccelerationWithKnownUDot, multiplyByMInv, calcArticulatedBodyInertias,
 * calcMobilizerReactionForces): 0, or TDYN_E_SINGULAR where some body's
 * D = H^T P H was not invertible (a determinant, or a pivot of the 6x6
 * elimination, not positive).  The reference throws for a singular 6x6
 * (lapackInverse, SmallMatrixMixed.h:878-892); here the call itself returns
 * TDYN_OK (it is asynchronous), the instance's outputs hold inf/NaN, and this
 * query reports it.  status (device, nullable): n_instances int32 copied on
 * stream s.  n_singular (host, nullable): synchronizes s, counts the singular
 * instances and then returns TDYN_E_SINGULAR if there are any.  No reference
 * counterpart as a call (the reference's check is the throw). */
int tdyn_batch_op_status(tdyn_batch* b, int32_t* status, int* n_singular, hipStream_t s);

/* info[0..7] = n_instances, total q, total u, total bodies (incl. Ground),
 * max nq, max nu, max nb, n_trees. */
int tdyn_batch_info(const tdyn_batch* b, int64_t* info);
/* Host arrays of n_instances entries: where each instance starts in packed
 * q, u and per-body ((nb+1) per instance) arrays. */
int tdyn_batch_offsets(const tdyn_batch* b, int32_t* q_off, int32_t* u_off, int32_t* body_off);

/* realizePositionKinematics + realizeVelocityKinematics
 * (SimbodyMatterSubsystem.h:2691, 2705) followed by reads of the body
 * transforms X_GB [12 per body], spatial velocities V_GB [6 per body], qdot
 * (calcQDot), and energies [KE, PE] per instance
 * (SimbodyMatterSubsystem.h:438 calcKineticEnergy, Force_Gravity.cpp:555).
 * Any output may be NULL. */
int tdyn_realize_kinematics(tdyn_batch* b, const double* q, const double* u, double* X_GB,
                            double* V_GB, double* qdot, double* energies, hipStream_t s);

/* realize(state, Stage::Acceleration) of a MultibodySystem with
 * Force::Gravity and constant applied body forces fapp (NULL: none)
 * (SimbodyMatterSubsystemRep.cpp:1527-1555): udot, qdot (NULL ok).  With
 * constraints set (tdyn_batch_set_constraints), the loop forward dynamics,
 * as realizeLoopForwardDynamics (:5160-5200) makes it. */
int tdyn_realize_acceleration(tdyn_batch* b, const double* q, const double* u, const double* fapp,
                              double* udot, double* qdot, hipStream_t s);

/* calcAccelerationIgnoringConstraints(state, fMob, F, udot, A_GB)
 * (SimbodyMatterSubsystem.h:2171-2176, .cpp:199-235) plus qdotdot
 * (calcQDotDot).  fmob: packed u layout, F: per-body.  A_GB, qdotdot may be NULL. */
int tdyn_calc_acceleration_ignoring_constraints(tdyn_batch* b, const double* q, const double* u,
                                                const double* fmob, const double* F, double* udot,
                                                double* A_GB, double* qdotdot, hipStream_t s);

/* Forward dynamics with known udot on some bodies: prescribed motion at the
 * acceleration level and locks (the isUDotKnown branches of
 * RigidBodyNodeSpec.cpp:292-296, 355-446; what realize(Acceleration) does for
 * a MobilizedBody with Motion::Acceleration or lock(state, Motion::Acceleration)).
 * known: device int32, nb+1 per instance in the body layout of F (Ground
 * first, ignored), nonzero = that body's udot is known; udot_known (packed u
 * layout) holds those udots.  Out: udot (the known entries copied, the others
 * computed), tau (on the known mobilities the generalized force that Simbody
 * keeps on the left-hand side, M udot + c + tau = fmob + J^T F, i.e. minus
 * the force the motion needs; zero on free mobilities), A_GB ((nb+1)*6 per
 * instance, nullable).  fmob and F are nullable (zero). */
int tdyn_calc_acceleration_with_known_udot(tdyn_batch* b, const double* q, const double* u,
                                           const double* fmob, const double* F, const int32_t* known,
                                           const double* udot_known, double* udot, double* tau,
                                           double* A_GB, hipStream_t stream);

/* Known udot for rollouts (locks: known udot 0 with the locked u at its
 * value, usually 0; prescribed constant udot): the device arrays of
 * tdyn_calc_acceleration_with_known_udot, kept by pointer (the caller keeps
 * them alive while a rollout begun with them runs).  They take effect at the
 * next tdyn_rollout_begin: a rollout keeps the arrays (and the evaluator
 * setting) of its begin for every advance, get and derivs; NULL, NULL
 * clears.  No Simbody counterpart as a call: what
 * MobilizedBody::lock(state) and Motion prescriptions do to the State.
 * Such rollouts run in FP64 on the tree sweeps or, for a batch of different
 * trees, on the level-parallel tree evaluator (evaluator 0 or 3 with lanes
 * other than 1). */
int tdyn_batch_set_known_udot(tdyn_batch* b, const int32_t* known, const double* udot_known);

/* Prescribed motion for rollouts: Motion::Sinusoid and Motion::Steady
 * (Simbody/include/simbody/internal/Motion.h:240-300; MotionImpl.h:182-300)
 * on any bodies.  motion: device, 4 doubles per body (nb+1 per instance,
 * Ground first, the body layout of F): level, amplitude a, rate w, phase
 * phi, with level 0 none; 1 position, every q = a sin(w t + phi) and u =
 * qdot (mobilizers with qdot = u: not Ball or Free); 2
 * velocity, every u = a sin(w t + phi); 3 acceleration, every udot = a
 * sin(w t + phi); 4 steady velocity, every u = a.  As in the reference's
 * integrator (IntegratorRep.h:901-934), every stage is evaluated with the
 * prescribed q, u and udot at its own time and each step's end state gets
 * the prescribed q and u before its error estimate.  Checked and copied
 * once here (the array is read again by later rollouts: keep it alive);
 * takes effect at the next tdyn_rollout_begin; NULL clears.  Rollouts with
 * motion run on the tree sweeps in FP64 (evaluator 0 or 1), without known
 * udot.  The reference's counterpart: adding Motion objects to mobilized
 * bodies. */
int tdyn_batch_set_motion(tdyn_batch* b, const double* motion);

/* Holonomic constraints of a tree: Constraint::Ball (station1 on body1 and
 * station2 on body2 held together; ConstraintImpl.h:1835-2000) and
 * Constraint::Rod (the two stations held at distance `length`;
 * Constraint_RodImpl.h:193-300), stations in the body frames, body 0 =
 * Ground.  The multipliers are in Ground (the reference's ancestor frame A
 * is Ground for these) and, as the reference's, are subtracted from the
 * applied forces: a Ball's lambda is minus the force on body 2 at the
 * station; a Rod's, minus the tension along station1 -> station2. */
#define TDYN_CONS_BALL 0
#define TDYN_CONS_ROD  1
#define TDYN_MAX_CONSTRAINT_ROWS 12   /* constraint equations per tree (Ball 3, Rod 1) */
typedef struct tdyn_constraint_desc {
    int32_t type;              /* TDYN_CONS_BALL or TDYN_CONS_ROD */
    int32_t body1, body2;      /* two different bodies of the tree, 0 = Ground */
    double station1[3];        /* on body1, in its frame */
    double station2[3];        /* on body2, in its frame */
    double length;             /* Rod: the distance (> 0); Ball: unused */
} tdyn_constraint_desc;
/* Sets the constraints of tree `tree` of the batch (replacing earlier ones;
 * n = 0 clears), as adding Constraint::Ball / Constraint::Rod objects to the
 * matter subsystem does (Constraint.cpp).  Not with Euler-angle Ball or Free
 * models.  Rollouts with constraints need a batch of one tree (a batch of
 * different trees takes the single-shot operators only) and run the tree
 * sweeps in FP64 (evaluator 0 or 1), without known udot, prescribed motion,
 * reports or tdyn_rollout_noalloc: every derivative is the loop forward dynamics
 * (calcLoopForwardDynamicsOperator, SimbodyMatterSubsystemRep.cpp:
 * 5076-5152), the initial state is projected onto the constraints
 * (initialize(), Integrator.cpp:376-377; status -4 where that fails), and
 * each step's end state and error estimate are projected as attemptDAEStep
 * does (AbstractIntegratorRep.cpp:137-208: projectQ, projectU,
 * SimbodyMatterSubsystemRep.cpp:4096-4440, 4666-4934), with the reference's
 * default unit weights.  A rollout in progress ends (begin again). */
int tdyn_batch_set_constraints(tdyn_batch* b, int tree, int n, const tdyn_constraint_desc* c);
/* realize(Acceleration) of a constrained system (SimbodyMatterSubsystemRep.
 * cpp:5076-5152 via realizeLoopForwardDynamics): gravity and the applied
 * body forces F ((nb+1)*6 per instance, nullable), udot (packed u layout)
 * and the multipliers lambda (TDYN_MAX_CONSTRAINT_ROWS per instance, zero
 * past the tree's rows; nullable) -- the reference's
 * getConstraintMultipliers.  Instances of trees without constraints get the
 * tree's forward dynamics.  tdyn_batch_op_status reports TDYN_E_SINGULAR
 * where G M^-1 G^T is singular (redundant constraints are not reduced, as
 * the reference's FactorQTZ does). */
int tdyn_calc_constrained_acceleration(tdyn_batch* b, const double* q, const double* u, const double* F,
                                       double* udot, double* lambda, hipStream_t s);

/* Event witnesses of a tree: the trigger functions of TriggeredEventHandler
 * and TriggeredEventReporter (EventHandler.h, EventReporter.h; getValue)
 * over a fixed set of kinds, with EventTriggerInfo's settings (Event.h:
 * 196-258) and a fixed set of handler actions (handleEvent):
 *   TDYN_WIT_Q            q[index] - offset (Stage::Position)
 *   TDYN_WIT_U            u[index] - offset (Stage::Velocity)
 *   TDYN_WIT_STATION_POS  direction . (position in Ground of `station` on
 *                         `body`) - offset (Stage::Position)
 *   TDYN_WIT_STATION_VEL  direction . (its velocity in Ground) - offset
 *                         (Stage::Velocity)
 *   TDYN_WIT_TIME         t - offset (Stage::Time)
 *   TDYN_EVENT_REPORT     a reporter: the event is recorded, the state is
 *                         not changed
 *   TDYN_EVENT_SCALE_U    u[action_index] *= factor (every u of the
 *                         instance when action_index < 0): impacts with
 *                         restitution (factor -e), energy injection
 *   TDYN_EVENT_TERMINATE  the handler sets shouldTerminate: the instance
 *                         stops with status 3
 * index is the coordinate within the tree; bodies are 1..nb (0 = Ground). */
#define TDYN_WIT_Q            0
#define TDYN_WIT_U            1
#define TDYN_WIT_STATION_POS  2
#define TDYN_WIT_STATION_VEL  3
#define TDYN_WIT_TIME         4
#define TDYN_TRIGGER_FALLING  1   /* Event::Falling: positive to negative or zero */
#define TDYN_TRIGGER_RISING   2   /* Event::Rising: negative to positive or zero */
#define TDYN_EVENT_REPORT     0
#define TDYN_EVENT_SCALE_U    1
#define TDYN_EVENT_TERMINATE  2
#define TDYN_MAX_WITNESSES    8   /* per tree */
typedef struct tdyn_witness_desc {
    int32_t kind;              /* TDYN_WIT_* */
    int32_t index;             /* TDYN_WIT_Q, _U: the coordinate */
    int32_t body;              /* TDYN_WIT_STATION_*: the body */
    int32_t trigger;           /* TDYN_TRIGGER_FALLING | TDYN_TRIGGER_RISING (both: the reference's default) */
    int32_t action;            /* TDYN_EVENT_* */
    int32_t action_index;      /* TDYN_EVENT_SCALE_U: the u scaled, < 0 every u */
    double station[3];         /* in the body frame */
    double direction[3];       /* in Ground */
    double offset;             /* subtracted from the value */
    double factor;             /* TDYN_EVENT_SCALE_U */
    double window;             /* required localization window in units of accuracy * time scale
                                  (EventTriggerInfo::setRequiredLocalizationTimeWindow; the
                                  reference's default is 0.1) */
} tdyn_witness_desc;
/* Sets the witnesses of tree `tree` (replacing earlier ones; n = 0 clears),
 * as adding TriggeredEventHandler / TriggeredEventReporter objects to the
 * System does.  Rollouts then watch them after every accepted step as
 * takeOneStep does (AbstractIntegratorRep.cpp:583-756): the witnesses that
 * change sign across the step under their trigger masks are the candidates
 * (a sign change away from zero is not one, Event.h:177-189); the window
 * (tLow, tHigh] shrinks around the earliest by trial states interpolated
 * inside the step (cubic Hermite, then the quaternion projection) and the
 * secant / bisection estimates of estimateRootTime (IntegratorRep.h:
 * 351-376) until it is at most the narrowest required window (accuracy *
 * time scale * window, at least SignificantReal * max(1, t)); the state moves
 * back to tHigh by interpolation, every remaining candidate is recorded
 * (tdyn_rollout_get_events) and its handler acts on that state, in witness
 * order; integration continues from it with the step size it had
 * (TimeStepper.cpp:178-205, Integrator.cpp:437-447).  A terminating handler
 * that changes nothing leaves the instance at tLow, in the state
 * interpolated there (the state ReachedEventTrigger reports); one that
 * changes the state leaves it at tHigh.  As in the reference, a witness
 * that changes sign twice within one step is not seen (bound the step size
 * for such witnesses).  Rollouts with witnesses run the tree sweeps in
 * FP64 (evaluator 0 or 1), on batches of one tree or of different trees,
 * with or without constraints (whose rollouts need a batch of one tree;
 * trial and backed-up states are then projected onto them, projectQ and
 * projectU at the tolerance, as realizeAndProjectKinematicsWithThrow does,
 * IntegratorRep.h:858-897), with or without known udot and prescribed
 * motion (trial and backed-up states then carry the prescribed q and u at
 * their own times), not with reports or tdyn_rollout_noalloc; status -5
 * marks an instance whose localization or whose projection of an
 * interpolated state failed.  A rollout in progress ends (begin again). */
int tdyn_batch_set_witnesses(tdyn_batch* b, int tree, int n, const tdyn_witness_desc* w);
/* The witnesses' values at a state (realize(state) to their stages, then
 * State::getEventTriggers, State.h): q, u (packed), t (one time per
 * instance, nullable: 0) -> values [n_instances][TDYN_MAX_WITNESSES] (device;
 * zero past the tree's witnesses).  Any batch (different trees too). */
int tdyn_calc_event_triggers(tdyn_batch* b, const double* q, const double* u, const double* t, double* values,
                             hipStream_t s);
/* Events recorded per instance per rollout (default 64; later events are
 * counted, not recorded).  A rollout in progress ends (begin again). */
int tdyn_batch_set_event_capacity(tdyn_batch* b, int capacity);
/* The events of the current rollout (device outputs, each nullable):
 * windows [n_instances][capacity][2] (tLow, tHigh), ids
 * [n_instances][capacity][2] (witness, transition: TDYN_TRIGGER_FALLING or
 * TDYN_TRIGGER_RISING as EventTriggerInfo::calcTransitionToReport gives
 * it), counts [n_instances] (events seen; entries past the capacity are not
 * recorded), in the order they happened. */
int tdyn_rollout_get_events(tdyn_batch* b, double* windows, int32_t* ids, int32_t* counts, hipStream_t s);

/* multiplyByMInv(state, v, MinvV) (SimbodyMatterSubsystem.h:1343-1345,
 * .cpp:431-471). */
int tdyn_multiply_by_m_inv(tdyn_batch* b, const double* q, const double* v, double* MinvV,
                           hipStream_t s);

/* calcResidualForceIgnoringConstraints(state, fMob, F, udot, resid)
 * (SimbodyMatterSubsystem.h:2234-2239, .cpp:236-304).  NULL fmob, F or
 * udot mean all-zero, the reference's zero-length Vectors. */
int tdyn_calc_residual_force_ignoring_constraints(tdyn_batch* b, const double* q, const double* u,
                                                  const double* fmob, const double* F,
                                                  const double* udot, double* resid, hipStream_t s);

/* realizeArticulatedBodyInertias + getArticulatedBodyInertia (and the P+
 * "as felt through the mobilizer") for every body
 * (SimbodyMatterSubsystem.h:2730, 2867).  21 doubles per body: mass block
 * M (xx yy zz xy xz yz), inertia block J (same order), mass-moment block F
 * (row-major 3x3); P = [J F; F^T M] about Bo in G.  PPlus may be NULL. */
int tdyn_calc_articulated_body_inertias(tdyn_batch* b, const double* q, double* P, double* PPlus,
                                        hipStream_t s);

/* multiplyByM(state, v, Mv) (SimbodyMatterSubsystem.h:1262): Mv = M v in
 * O(n), per instance (nu doubles each, packed like u).  Needs q only. */
int tdyn_multiply_by_m(tdyn_batch* b, const double* q, const double* v, double* Mv, hipStream_t s);

/* multiplyBySystemJacobian(state, v, Jv) (SimbodyMatterSubsystem.h:554):
 * Jv = J v, the spatial velocity each body would have for speeds v, 6
 * doubles per body including Ground's zero ((nb+1)*6 per instance). */
int tdyn_multiply_by_system_jacobian(tdyn_batch* b, const double* q, const double* v, double* Jv,
                                     hipStream_t s);

/* multiplyBySystemJacobianTranspose(state, F_G, f) (SimbodyMatterSubsystem.h:
 * 646): f = J^T F_G for spatial forces (torque, force) at the body origins,
 * in G, (nb+1)*6 per instance with Ground's entry ignored; f packed like u. */
int tdyn_multiply_by_system_jacobian_transpose(tdyn_batch* b, const double* q, const double* F_G, double* f,
                                               hipStream_t s);

/* calcBodyAccelerationFromUDot(state, udot, A_GB) (SimbodyMatterSubsystem.h:
 * 2365): A_GB = J udot + JDot u, (nb+1)*6 per instance with Ground's zero;
 * udot NULL means zero (only the velocity-dependent part). */
int tdyn_calc_body_acceleration_from_udot(tdyn_batch* b, const double* q, const double* u, const double* udot,
                                          double* A_GB, hipStream_t s);

/* calcCompositeBodyInertias(state, R) (SimbodyMatterSubsystem.h:2323): the
 * spatial inertia of each body with everything outboard of it welded, about
 * its origin, in G.  10 doubles per body, Ground first ((nb+1)*10 per
 * instance): mass, mass center (3), unit inertia (xx yy zz xy xz yz);
 * Ground's entry is infinite mass, zero mass center, unit inertia 1. */
int tdyn_calc_composite_body_inertias(tdyn_batch* b, const double* q, double* R, hipStream_t s);

/* calcMobilizerReactionForces(state, FM_G) (SimbodyMatterSubsystemRep.cpp:
 * 5786-5826) for the given applied forces: reaction at each M frame origin,
 * expressed in G, 6 per body. */
int tdyn_calc_mobilizer_reaction_forces(tdyn_batch* b, const double* q, const double* u,
                                        const double* fmob, const double* F, double* FM,
                                        hipStream_t s);

/* ---- Integration: RungeKuttaMersonIntegrator + TimeStepper::stepTo ------- */
typedef struct tdyn_integrator_cfg {
    double accuracy;             /* Integrator::setAccuracy (<=0: 1e-3 default) */
    double constraint_tolerance; /* setConstraintTolerance (<=0: accuracy/10) */
    double fixed_step;           /* setFixedStepSize (<=0: adaptive) */
    double initial_step, min_step, max_step;   /* <=0: unset */
    double time_scale;           /* System default time scale (<=0: 0.1) */
    int32_t max_attempts;        /* per-instance attempt budget (<=0: 1e8) */
    int32_t precision;           /* 0 = FP64; 1 = mixed: FP32 kinematics, bias forces and
                                    articulated inertias, FP64 D = H^T P H and its inverse,
                                    quaternion normalization and Kahan-compensated state;
                                    2 = FP64 arithmetic with the G = P H D^-1 hand-over rows
                                    stored in FP32; 3 = as 2 with the joint trig cache in
                                    FP32 too.  1-3: chain evaluator of one model per batch
                                    only (TDYN_E_ARG otherwise; serving: 0 or 1) */
    int32_t use_infinity_norm;   /* Integrator::setUseInfinityNorm (Integrator.cpp:276-280): nonzero makes
                                    every error and constraint norm the largest magnitude instead of
                                    the RMS.  FP64 only; runs on the tree evaluators (a chain batch then
                                    takes the tree sweeps); the chain server refuses it */
    int32_t project_every_step;  /* setProjectEveryStep (Integrator.cpp:288-290): nonzero forces the
                                    projection (quaternion normalization, one Newton iteration) on every
                                    attempt that reaches it.  Same limits as use_infinity_norm */
    int32_t project_interpolated_states;   /* setProjectInterpolatedStates (Integrator.cpp:294-296):
                                    0 (default) projects interpolated report and event-localization
                                    states, -1 leaves them unprojected (the state backed up to an event
                                    is projected either way) */
} tdyn_integrator_cfg;

/* Autotuner of the tree evaluators' occupancy (no Simbody counterpart: it
 * picks how the batch maps onto the GPU).  For a batch of different trees it
 * times short probe rollouts (0 to t_probe, each candidate twice, the faster
 * kept) of the given state under fixed lanes per instance (8, 16, 32), ranges
 * of the automatic choice ([4,16], [8,32], [16,32], [16,64]), the one-lane
 * streaming evaluator and the tree sweeps, applies the fastest, and appends
 * it to cache_path (nullable) under a key of the batch's shape (instances,
 * mean bodies, depth, widest level, log2 buckets); a later call with a key
 * already in the file applies the cached choice without probing.  Batches of
 * one model have nothing to choose; with known udot the one-lane streaming
 * evaluator is not a candidate.  chosen (nullable,
 * 5 entries): evaluator mode, lanes, lane range min and max, 1 if it came
 * from the cache.  q, u are not modified; call tdyn_rollout_begin after. */
int tdyn_batch_autotune(tdyn_batch* b, const double* q, const double* u, const double* fapp,
                        const tdyn_integrator_cfg* cfg, double t_probe, const char* cache_path,
                        int32_t* chosen, hipStream_t stream);

/* TimeStepper::initialize (Integrator.cpp:333-435): load q,u (packed),
 * constant applied body forces fapp (NULL: none), start time t0 and
 * per-instance final times t_end (device, NULL: all t_end_all).  The state
 * stays resident on the device until tdyn_rollout_get.  A begin refused for
 * its arguments or settings (TDYN_E_ARG) leaves the rollout in progress as it
 * was; one that fails later (TDYN_E_HIP, TDYN_E_NOMEM) ends it. */
int tdyn_rollout_begin(tdyn_batch* b, const double* q, const double* u, const double* fapp,
                       double t0, const double* t_end, double t_end_all,
                       const tdyn_integrator_cfg* cfg, hipStream_t s);
/* Advance every unfinished instance by at most max_steps accepted steps
 * (TimeStepper::stepTo(t_end) with Integrator::setFinalTime(t_end),
 * AbstractIntegratorRep.cpp:216-436).  Resumable: calling again continues. */
int tdyn_rollout_advance(tdyn_batch* b, int max_steps, hipStream_t s);
/* TimeStepper::stepTo(report time) with Integrator::setAllowInterpolation(
 * false) (AbstractIntegratorRep.cpp:216-240, takeOneStep :512-530): moves
 * every instance's final time to t_end (device, per instance; NULL: all
 * t_end_all).  An instance that had reached its previous final time resumes
 * from its state, time, step size and statistics when the new time lies
 * ahead; tdyn_rollout_advance then stops every step at the new time, so
 * begin(t_1), advance, get, extend(t_2), advance, get, ... reports the
 * trajectory at t_1, t_2, ... exactly as Simbody's TimeStepper does with
 * interpolation off. */
int tdyn_rollout_extend(tdyn_batch* b, const double* t_end, double t_end_all, hipStream_t s);
/* Size of the device array that tdyn_rollout_noalloc works in (one copy per
 * GPU the library runs on, held until the process ends). */
#define TDYN_NOALLOC_SCRATCH_BYTES (32ull << 20)

/* tdyn_rollout without allocating: no reference counterpart (the reference
 * is a CPU library); for callers that cannot allocate device memory when
 * they call (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and results as tdyn_rollout (q, u
 * in and out, device; fapp, stats nullable; one final time t_end for every
 * instance), run by the chain evaluator or the tree sweeps as the batch's
 * evaluator setting selects (not the level-parallel evaluator: evaluator 3
 * is TDYN_E_ARG, and batches of different trees on the automatic setting run
 * the tree sweeps), FP64 only.  The per-lane records and the whole
 * integrator state live in a device array of TDYN_NOALLOC_SCRATCH_BYTES
 * instead of the batch's buffers (about 38 KB per instance of a 32-body
 * chain, so some 850 such instances); a batch that needs more is refused
 * with TDYN_E_SIZE.  Calls that overlap
 * on one GPU share the array and race; the batch's own rollout (begin /
 * advance / get) is not touched.  Synchronizes s before returning. */
int tdyn_rollout_noalloc(tdyn_batch* b, double* q, double* u, const double* fapp, double t0, double t_end,
                         const tdyn_integrator_cfg* cfg, int32_t* stats, hipStream_t s);

/* Copy out state (packed q, u), per-instance time t, status (0 running,
 * 1 reached t_end, 3 an event handler terminated it, <0 failed) and stats (attempts, accepted, rejected,
 * projections; 4 ints per instance).  Any output may be NULL. */
int tdyn_rollout_get(tdyn_batch* b, double* q, double* u, double* t, int32_t* status,
                     int32_t* stats, hipStream_t s);
/* Rollout snapshot (checkpoint / resume).  The reference's checkpoint is a
 * copy of the State (State.h:1206-1222; no serialization) and its resume is
 * Integrator::initialize(state) (Integrator.cpp:333-435), which projects and
 * restarts the step control.  Here a snapshot is the whole integrator state
 * of every instance in a documented byte layout (host memory), and a
 * restore continues each instance exactly where it was: same state, time,
 * step size, status and statistics, so the resumed rollout's results equal
 * the uninterrupted one's bit for bit.  Layout (little-endian):
 *   0   char[8] "TDYNSNP1"; 8 uint32 version (1); 12 int32 n_instances;
 *   16  int64 total_q; 24 int64 total_u; 32 uint64 model hash (FNV-1a of
 *       the batch's tree, body and constant tables, gravity and instance
 *       trees); 40 int32 precision; 44 int32 evaluator setting;
 *   48  double accuracy, constraint tolerance, fixed step, initial, minimum
 *       and maximum step (-1: unset), time scale; 104 int32 max attempts;
 *   112 per instance, instance order: double t[n], h[n], t_end[n]; int32
 *       status[n], stats[4n]; double q[total_q], u[total_u] (packed); with
 *       precision 1 also the Kahan compensation rows of q and u.
 * The event log is not part of a snapshot (a restored rollout's counts start
 * at zero).  tdyn_rollout_snapshot: bytes (nullable) gets the size; out
 * (host, nullable: size query only) of capacity bytes; synchronizes s. */
int tdyn_rollout_snapshot(tdyn_batch* b, void* out, int64_t capacity, int64_t* bytes, hipStream_t s);
/* Begins a rollout from a snapshot taken on a batch of the same model
 * (TDYN_E_ARG for another model or format, TDYN_E_SIZE for other sizes),
 * with the batch's own evaluator setting and constraints, witnesses, known
 * udot and motion as set now; fapp as in tdyn_rollout_begin.  No initial
 * projection.  Synchronizes s. */
int tdyn_rollout_restore(tdyn_batch* b, const void* snapshot, int64_t bytes, const double* fapp, hipStream_t s);
/* The state derivative (qdot, udot; packed) of the current rollout state,
 * evaluated by the rollout's evaluator at the rollout's precision: the f0
 * the integrator uses (realize(state, Stage::Acceleration) + getQDot/getUDot,
 * SimbodyMatterSubsystemRep.cpp:1527-1555).  Overwrites the integrator's
 * stage storage; call between advances. */
int tdyn_rollout_derivs(tdyn_batch* b, double* qdot, double* udot, hipStream_t s);
/* TimeStepper::stepTo(report times) ... to t_end with Simbody's default
 * reporting, interpolation on (AbstractIntegratorRep.cpp:63-93, 216-330):
 * the steps ignore the report times (the trajectory and stats are those of
 * tdyn_rollout to t_end); each report time r (report_times: n_reports HOST
 * doubles, increasing, within [t0, t_end]) gets the end state of the step
 * that reaches r when it lands on r, else the cubic Hermite interpolant of
 * that step's two ends (interpolateOrder3, IntegratorRep.h:285-313) with its
 * quaternions normalized when their RMS norm error exceeds the constraint
 * tolerance (projectQ of the interpolated state).  q_reports [n][total_q],
 * u_reports [n][total_u] (device); q, u in/out (state at t_end).  Runs one
 * accepted step per launch with a derivative evaluation at each step's end,
 * so it is slower than tdyn_rollout; synchronizes s. */
int tdyn_rollout_reports(tdyn_batch* b, double* q, double* u, const double* fapp, double t0, double t_end,
                         const double* report_times, int n_reports, const tdyn_integrator_cfg* cfg,
                         double* q_reports, double* u_reports, int32_t* stats, hipStream_t s);
/* begin + advance until every instance finished + get; synchronizes s.
 * Returns TDYN_E_STEP_FAILED if any instance failed. */
int tdyn_rollout(tdyn_batch* b, double* q, double* u, const double* fapp, double t0,
                 double t_end, const tdyn_integrator_cfg* cfg, int32_t* stats, hipStream_t s);

/* ---- Serving: continuous batching of rollout requests ------------------
 * No reference counterpart: Simbody runs one State per caller.  A server
 * owns device slots for trees of ONE topology (the parent and mobilizer
 * arrays of `topology`, a serial chain); every request brings its own
 * masses, frames, gravity, initial state, final time and (optionally)
 * applied body forces, and keeps its State on the device until collected.
 * The slot count follows from an HBM budget.  Requests admitted between
 * advances join the running ones; each slot runs the same arithmetic as a
 * one-instance tdyn_rollout of that request, so results do not depend on
 * what else is in the server.  Not thread-safe (one host thread). */
typedef struct tdyn_server tdyn_server;

/* slots = min(max_slots, hbm_budget_bytes / bytes per slot), rounded down to
 * a multiple of 64 (TDYN_E_NOMEM if that is 0). */
int tdyn_server_create(const tdyn_tree_desc* topology, const tdyn_integrator_cfg* cfg,
                       size_t hbm_budget_bytes, int max_slots, tdyn_server** out);
int tdyn_server_destroy(tdyn_server* sv);
/* info[0..5] = slots, free slots, occupied slots, bytes per slot, nq, nu. */
int tdyn_server_info(const tdyn_server* sv, int64_t* info);
/* Admit the first min(n, free slots) requests; host arrays: trees[n] (same
 * topology), q [n][nq], u [n][nu], t_end[n], fapp [n][(nb+1)*6] or NULL.
 * Writes the slot of each admitted request and their count.  All n requests
 * are validated before any is admitted.  Synchronizes s. */
int tdyn_server_admit(tdyn_server* sv, int n, const tdyn_tree_desc* trees, const double* q, const double* u,
                      const double* t_end, const double* fapp, int32_t* slots, int* n_admitted, hipStream_t s);
/* Advance every running slot by at most max_steps accepted steps (one launch, async). */
int tdyn_server_advance(tdyn_server* sv, int max_steps, hipStream_t s);
/* Retire up to max_out finished slots (reached t_end or failed): host
 * outputs slot ids, q [n][nq], u [n][nu], t, status (1 done, <0 failed) and
 * stats [n][4]; any output but slots may be NULL.  The slots become free.
 * Synchronizes s. */
int tdyn_server_collect(tdyn_server* sv, int max_out, int32_t* slots, double* q, double* u, double* t,
                        int32_t* status, int32_t* stats, int* n_out, hipStream_t s);

/* ---- Serving requests of different trees ---------------------------------
 * No reference counterpart.  A tree server's slots hold requests of any tree
 * of tdyn_batch_create (any size, topology and mobilizer mix), each with its
 * own constants, gravity, initial state, final time and applied body
 * forces.  A request's memory comes from two pools sized from the HBM budget
 * at creation -- body tables (nb + 1 entries, rounded up to whole blocks of
 * four) and records plus integrator state (its own nq + nu) -- so its byte
 * cost is its own tree's (tdyn_tree_server_request_bytes).  Requests are
 * admitted in order while a slot is free and both pools have room for the
 * next one; collection returns the memory.  All slots advance together, each
 * with its own step control.  A request runs on the level-parallel tree
 * evaluator (a group of lanes per request) when its tree allows it (no Free
 * or Bushing body below body 1), else on the one-lane tree sweeps; the
 * evaluator and its lane count are fixed at admission from the request's
 * own tree, so a slot's results do not depend on the other slots.  FP64;
 * not thread-safe (one host thread). */
typedef struct tdyn_tree_server tdyn_tree_server;

/* max_slots <= 0: 65,536 slots.  The budget pays the slots' own state (about
 * 160 bytes each), then a quarter of the rest goes to the body pool and the
 * remainder to the record and state pool. */
int tdyn_tree_server_create(const tdyn_integrator_cfg* cfg, size_t hbm_budget_bytes, int max_slots,
                            tdyn_tree_server** out);
int tdyn_tree_server_destroy(tdyn_tree_server* sv);
/* info[0..6] = slots, free slots, occupied slots, budget bytes, bytes in use
 * (slots' own state plus the pools' occupied parts), body pool capacity
 * (bodies), record and state pool capacity (doubles). */
int tdyn_tree_server_info(const tdyn_tree_server* sv, int64_t* info);
/* What one request of this tree takes from the budget (bytes), at the
 * automatic lane setting. */
int tdyn_tree_server_request_bytes(const tdyn_tree_desc* tree, int64_t* bytes);
/* Evaluator of the requests admitted from now on: 0 (default) automatic --
 * the level-parallel evaluator where the tree allows it, with lanes from the
 * tree's widest level (8, 16 or 32); 1 the one-lane tree sweeps for every
 * request; 4..64 (a power of two) the level-parallel evaluator with that
 * many lanes where the tree allows it. */
int tdyn_tree_server_set_lanes(tdyn_tree_server* sv, int lanes);
/* Admit requests in order while a slot is free and the pools have room; host
 * arrays: trees[n], q and u packed request after request (each its own nq,
 * nu), t_end[n] (> 0; every request starts at t = 0), fapp packed
 * (nb+1)*6 per request or NULL.  All n requests are validated before any is
 * admitted.  Writes the slots of the admitted (the first *n_admitted)
 * requests.  Synchronizes s. */
int tdyn_tree_server_admit(tdyn_tree_server* sv, int n, const tdyn_tree_desc* trees, const double* q,
                           const double* u, const double* t_end, const double* fapp, int32_t* slots,
                           int* n_admitted, hipStream_t s);
/* Advance every running slot by at most max_steps accepted steps (one launch, async). */
int tdyn_tree_server_advance(tdyn_tree_server* sv, int max_steps, hipStream_t s);
/* Retire up to max_out finished slots whose states fit y_cap doubles: host
 * outputs slot ids, y (each request's q then u, packed), y_off (where each
 * starts in y), t, status (1 done, <0 failed) and stats [n][4] (t, status,
 * stats nullable).  The slots and their memory become free.  Synchronizes s. */
int tdyn_tree_server_collect(tdyn_tree_server* sv, int max_out, int64_t y_cap, int32_t* slots, double* y,
                             int64_t* y_off, double* t, int32_t* status, int32_t* stats, int* n_out,
                             hipStream_t s);

const char* tdyn_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* TDYN_H */
