"""Host-side Python mirror of the tdyn C ABI (include/tdyn.h).

`MatterBatch` offers Simbody's SimbodyMatterSubsystem operator surface for a
batch of independent tree instances, under the reference's method names
(SimbodyMatterSubsystem.h): calcAccelerationIgnoringConstraints,
calcResidualForceIgnoringConstraints, multiplyByMInv,
realizeArticulatedBodyInertias/getArticulatedBodyInertia, calcKineticEnergy,
calcMobilizerReactionForces, plus realizeAcceleration of a MultibodySystem with
gravity, and `RungeKuttaMerson` / `TimeStepper.stepTo` for integration.

Arrays are torch float64 CUDA tensors in the packed instance-major layout of
the C ABI; the library's HIP kernels do all arithmetic.  There is no CPU
fallback: without the built libtdyn.so every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtdyn.so")

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)

TDYN_ERRORS = {1: "TDYN_E_ARG", 2: "TDYN_E_SIZE", 3: "TDYN_E_STAGE", 4: "TDYN_E_SINGULAR",
               5: "TDYN_E_STEP_FAILED", 6: "TDYN_E_HIP", 7: "TDYN_E_NOMEM"}

EXPORTS = ["tdyn_batch_create", "tdyn_batch_destroy", "tdyn_batch_set_evaluator", "tdyn_batch_get_evaluator",
           "tdyn_batch_set_tree_lanes", "tdyn_batch_set_tree_lane_range",
           "tdyn_batch_info", "tdyn_batch_op_status",
           "tdyn_batch_offsets",
           "tdyn_realize_kinematics", "tdyn_realize_acceleration",
           "tdyn_calc_acceleration_ignoring_constraints", "tdyn_calc_acceleration_with_known_udot",
           "tdyn_batch_set_known_udot", "tdyn_batch_set_motion", "tdyn_batch_autotune",
           "tdyn_batch_set_constraints", "tdyn_calc_constrained_acceleration",
           "tdyn_batch_set_witnesses", "tdyn_batch_set_event_capacity", "tdyn_rollout_get_events",
           "tdyn_calc_event_triggers", "tdyn_rollout_snapshot", "tdyn_rollout_restore",
           "tdyn_multiply_by_m_inv",
           "tdyn_calc_residual_force_ignoring_constraints", "tdyn_calc_articulated_body_inertias",
           "tdyn_calc_mobilizer_reaction_forces", "tdyn_multiply_by_m", "tdyn_multiply_by_system_jacobian",
           "tdyn_multiply_by_system_jacobian_transpose", "tdyn_calc_body_acceleration_from_udot",
           "tdyn_calc_composite_body_inertias", "tdyn_rollout_begin", "tdyn_rollout_advance",
           "tdyn_rollout_get", "tdyn_rollout_derivs", "tdyn_rollout", "tdyn_rollout_extend", "tdyn_rollout_reports",
           "tdyn_rollout_noalloc",
           "tdyn_last_error",
           "tdyn_server_create", "tdyn_server_destroy", "tdyn_server_info", "tdyn_server_admit",
           "tdyn_server_advance", "tdyn_server_collect",
           "tdyn_tree_server_create", "tdyn_tree_server_destroy", "tdyn_tree_server_info",
           "tdyn_tree_server_request_bytes", "tdyn_tree_server_admit", "tdyn_tree_server_advance",
           "tdyn_tree_server_collect", "tdyn_tree_server_set_lanes"]


class TdynError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{TDYN_ERRORS.get(code, code)}: {msg}")
        self.code = code


class TreeDescC(ctypes.Structure):
    _fields_ = [("nb", ctypes.c_int32), ("parent", _i32), ("mobilizer", _i32),
                ("mass", _d), ("com_B", _d), ("inertia_B", _d),
                ("X_PF", _d), ("X_BM", _d), ("gravity", ctypes.c_double * 3), ("screw_pitch", _d),
                ("reversed", _i32), ("use_euler_angles", ctypes.c_int32)]


class ConstraintDescC(ctypes.Structure):
    """tdyn_constraint_desc."""
    _fields_ = [("type", ctypes.c_int32), ("body1", ctypes.c_int32), ("body2", ctypes.c_int32),
                ("station1", ctypes.c_double * 3), ("station2", ctypes.c_double * 3), ("length", ctypes.c_double)]


CONS_BALL, CONS_ROD = 0, 1
MAX_CONSTRAINT_ROWS = 12


class WitnessDescC(ctypes.Structure):
    """tdyn_witness_desc."""
    _fields_ = [("kind", ctypes.c_int32), ("index", ctypes.c_int32), ("body", ctypes.c_int32),
                ("trigger", ctypes.c_int32), ("action", ctypes.c_int32), ("action_index", ctypes.c_int32),
                ("station", ctypes.c_double * 3), ("direction", ctypes.c_double * 3), ("offset", ctypes.c_double),
                ("factor", ctypes.c_double), ("window", ctypes.c_double)]


# Event witnesses: kinds, trigger masks, handler actions (tdyn.h TDYN_WIT_*,
# TDYN_TRIGGER_*, TDYN_EVENT_*), and the rollout status of an instance whose
# handler terminated it.
WIT_Q, WIT_U, WIT_STATION_POS, WIT_STATION_VEL, WIT_TIME = 0, 1, 2, 3, 4
TRIGGER_FALLING, TRIGGER_RISING, TRIGGER_BOTH = 1, 2, 3
EVENT_REPORT, EVENT_SCALE_U, EVENT_TERMINATE = 0, 1, 2
MAX_WITNESSES = 8
STATUS_TERMINATED = 3
WITNESS_DEFAULTS = dict(index=0, body=0, trigger=TRIGGER_BOTH, action=EVENT_REPORT, action_index=-1,
                        station=(0.0, 0.0, 0.0), direction=(0.0, 0.0, 0.0), offset=0.0, factor=1.0, window=0.1)


class IntegratorCfgC(ctypes.Structure):
    _fields_ = [("accuracy", ctypes.c_double), ("constraint_tolerance", ctypes.c_double),
                ("fixed_step", ctypes.c_double), ("initial_step", ctypes.c_double),
                ("min_step", ctypes.c_double), ("max_step", ctypes.c_double),
                ("time_scale", ctypes.c_double), ("max_attempts", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("use_infinity_norm", ctypes.c_int32),
                ("project_every_step", ctypes.c_int32), ("project_interpolated_states", ctypes.c_int32)]


_lib = None


def lib():
    """Load libtdyn.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} missing "
                               "(run __graft_entry__.build() or make -C simbody_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(L, name).restype = ctypes.c_int
        L.tdyn_last_error.restype = ctypes.c_char_p
        vp = ctypes.c_void_p
        L.tdyn_batch_create.argtypes = [ctypes.POINTER(TreeDescC), ctypes.c_int, _i32, ctypes.c_int,
                                        ctypes.POINTER(vp)]
        L.tdyn_batch_destroy.argtypes = [vp]
        L.tdyn_batch_set_evaluator.argtypes = [vp, ctypes.c_int]
        L.tdyn_batch_get_evaluator.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.tdyn_batch_set_tree_lanes.argtypes = [vp, ctypes.c_int]
        L.tdyn_batch_set_tree_lane_range.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.tdyn_batch_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.tdyn_batch_op_status.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int), vp]
        L.tdyn_batch_offsets.argtypes = [vp, _i32, _i32, _i32]
        L.tdyn_realize_kinematics.argtypes = [vp] + [vp] * 6 + [vp]
        L.tdyn_realize_acceleration.argtypes = [vp] + [vp] * 5 + [vp]
        L.tdyn_calc_acceleration_ignoring_constraints.argtypes = [vp] + [vp] * 7 + [vp]
        L.tdyn_calc_acceleration_with_known_udot.argtypes = [vp] + [vp] * 9 + [vp]
        L.tdyn_batch_set_known_udot.argtypes = [vp, vp, vp]
        L.tdyn_batch_set_motion.argtypes = [vp, vp]
        L.tdyn_batch_set_constraints.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ConstraintDescC)]
        L.tdyn_calc_constrained_acceleration.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.tdyn_batch_set_witnesses.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(WitnessDescC)]
        L.tdyn_batch_set_event_capacity.argtypes = [vp, ctypes.c_int]
        L.tdyn_rollout_get_events.argtypes = [vp, vp, vp, vp, vp]
        L.tdyn_calc_event_triggers.argtypes = [vp, vp, vp, vp, vp, vp]
        L.tdyn_rollout_snapshot.argtypes = [vp, vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), vp]
        L.tdyn_rollout_restore.argtypes = [vp, vp, ctypes.c_int64, vp, vp]
        L.tdyn_batch_autotune.argtypes = [vp, vp, vp, vp, ctypes.POINTER(IntegratorCfgC), ctypes.c_double,
                                          ctypes.c_char_p, vp, vp]
        L.tdyn_multiply_by_m_inv.argtypes = [vp] + [vp] * 3 + [vp]
        L.tdyn_calc_residual_force_ignoring_constraints.argtypes = [vp] + [vp] * 6 + [vp]
        L.tdyn_calc_articulated_body_inertias.argtypes = [vp] + [vp] * 3 + [vp]
        L.tdyn_calc_mobilizer_reaction_forces.argtypes = [vp] + [vp] * 5 + [vp]
        L.tdyn_multiply_by_m.argtypes = [vp] * 5
        L.tdyn_multiply_by_system_jacobian.argtypes = [vp] * 5
        L.tdyn_multiply_by_system_jacobian_transpose.argtypes = [vp] * 5
        L.tdyn_calc_body_acceleration_from_udot.argtypes = [vp] * 6
        L.tdyn_calc_composite_body_inertias.argtypes = [vp] * 4
        L.tdyn_rollout_begin.argtypes = [vp, vp, vp, vp, ctypes.c_double, vp, ctypes.c_double,
                                         ctypes.POINTER(IntegratorCfgC), vp]
        L.tdyn_rollout_advance.argtypes = [vp, ctypes.c_int, vp]
        L.tdyn_rollout_extend.argtypes = [vp, vp, ctypes.c_double, vp]
        L.tdyn_rollout_reports.argtypes = [vp, vp, vp, vp, ctypes.c_double, ctypes.c_double, vp, ctypes.c_int,
                                           ctypes.POINTER(IntegratorCfgC), vp, vp, vp, vp]
        L.tdyn_rollout_get.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.tdyn_rollout_noalloc.argtypes = [vp, vp, vp, vp, ctypes.c_double, ctypes.c_double,
                                           ctypes.POINTER(IntegratorCfgC), vp, vp]
        L.tdyn_rollout_derivs.argtypes = [vp, vp, vp, vp]
        L.tdyn_rollout.argtypes = [vp, vp, vp, vp, ctypes.c_double, ctypes.c_double,
                                   ctypes.POINTER(IntegratorCfgC), vp, vp]
        L.tdyn_server_create.argtypes = [ctypes.POINTER(TreeDescC), ctypes.POINTER(IntegratorCfgC), ctypes.c_size_t,
                                         ctypes.c_int, ctypes.POINTER(vp)]
        L.tdyn_server_destroy.argtypes = [vp]
        L.tdyn_server_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.tdyn_server_admit.argtypes = [vp, ctypes.c_int, ctypes.POINTER(TreeDescC), _d, _d, _d, _d, _i32,
                                        ctypes.POINTER(ctypes.c_int), vp]
        L.tdyn_server_advance.argtypes = [vp, ctypes.c_int, vp]
        L.tdyn_tree_server_create.argtypes = [ctypes.POINTER(IntegratorCfgC), ctypes.c_size_t, ctypes.c_int,
                                              ctypes.POINTER(vp)]
        L.tdyn_tree_server_destroy.argtypes = [vp]
        L.tdyn_tree_server_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.tdyn_tree_server_request_bytes.argtypes = [ctypes.POINTER(TreeDescC), ctypes.POINTER(ctypes.c_int64)]
        L.tdyn_tree_server_admit.argtypes = [vp, ctypes.c_int, ctypes.POINTER(TreeDescC), _d, _d, _d, _d, _i32,
                                             ctypes.POINTER(ctypes.c_int), vp]
        L.tdyn_tree_server_advance.argtypes = [vp, ctypes.c_int, vp]
        L.tdyn_tree_server_set_lanes.argtypes = [vp, ctypes.c_int]
        L.tdyn_tree_server_collect.argtypes = [vp, ctypes.c_int, ctypes.c_int64, _i32, _d,
                                               ctypes.POINTER(ctypes.c_int64), _d, _i32, _i32,
                                               ctypes.POINTER(ctypes.c_int), vp]
        L.tdyn_server_collect.argtypes = [vp, ctypes.c_int, _i32, _d, _d, _d, _i32, _i32,
                                          ctypes.POINTER(ctypes.c_int), vp]
        _lib = L
    return _lib


def fill_tree_desc(d, spec, keep):
    """Fill a TreeDescC from a TreeSpec-like object; the arrays it points to
    are appended to `keep`, which must outlive every use of d."""
    arrs = [np.ascontiguousarray(spec.parent, np.int32), np.ascontiguousarray(spec.mobilizer, np.int32)] + \
           [np.ascontiguousarray(x, np.float64) for x in (spec.mass, spec.com_B, spec.inertia_B, spec.X_PF, spec.X_BM)]
    keep.append(arrs)
    d.nb = len(arrs[0])
    d.parent = arrs[0].ctypes.data_as(_i32); d.mobilizer = arrs[1].ctypes.data_as(_i32)
    d.mass, d.com_B, d.inertia_B, d.X_PF, d.X_BM = (a.ctypes.data_as(_d) for a in arrs[2:])
    for j in range(3):
        d.gravity[j] = float(spec.gravity[j])
    pitch = getattr(spec, "pitch", None)
    if pitch is not None:
        pa = np.ascontiguousarray(pitch, np.float64)
        keep.append(pa)
        d.screw_pitch = pa.ctypes.data_as(_d)
    d.use_euler_angles = 1 if getattr(spec, "use_euler_angles", False) else 0
    rev = getattr(spec, "reversed", None)
    if rev is not None:
        ra = np.ascontiguousarray(rev, np.int32)
        keep.append(ra)
        d.reversed = ra.ctypes.data_as(_i32)
    return d


def _check(rc):
    if rc != 0:
        raise TdynError(rc, lib().tdyn_last_error().decode())


def _ptr(t):
    if t is None:
        return None
    import torch
    assert t.is_cuda and t.is_contiguous(), "device tensors must be contiguous CUDA tensors"
    assert t.dtype in (torch.float64, torch.int32)
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class RungeKuttaMerson:
    """RungeKuttaMersonIntegrator settings (Integrator.cpp:243-270)."""

    FP64, MIXED = 0, 1

    def __init__(self, accuracy=1e-3, fixed_step=0.0, constraint_tolerance=0.0, initial_step=0.0,
                 min_step=0.0, max_step=0.0, time_scale=0.1, max_attempts=0, precision=0,
                 use_infinity_norm=False, project_every_step=False, project_interpolated_states=True):
        self.accuracy, self.fixed_step = accuracy, fixed_step
        self.constraint_tolerance, self.initial_step = constraint_tolerance, initial_step
        self.min_step, self.max_step, self.time_scale = min_step, max_step, time_scale
        self.max_attempts = max_attempts
        self.precision = precision   # 0 FP64, 1 mixed (chain evaluator)
        # Integrator::setUseInfinityNorm, setProjectEveryStep, setProjectInterpolatedStates
        self.use_infinity_norm = use_infinity_norm
        self.project_every_step = project_every_step
        self.project_interpolated_states = project_interpolated_states

    def setUseInfinityNorm(self, on):
        self.use_infinity_norm = bool(on)

    def setProjectEveryStep(self, on):
        self.project_every_step = bool(on)

    def setProjectInterpolatedStates(self, on):
        self.project_interpolated_states = bool(on)

    def c(self):
        return IntegratorCfgC(self.accuracy, self.constraint_tolerance, self.fixed_step,
                              self.initial_step, self.min_step, self.max_step, self.time_scale,
                              self.max_attempts, self.precision, int(bool(self.use_infinity_norm)),
                              int(bool(self.project_every_step)), 0 if self.project_interpolated_states else -1)


class MatterBatch:
    """A batch of tree instances on one GPU (instance i uses trees[tree_of[i]])."""

    def __init__(self, trees, tree_of=None, n_instances=None, device="cuda:0"):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        if not isinstance(trees, (list, tuple)):
            trees = [trees]
        self.trees = list(trees)
        if tree_of is not None:
            tree_of = np.ascontiguousarray(tree_of, np.int32)
            n_instances = len(tree_of)
        elif n_instances is None:
            n_instances = 1
        self.tree_of = tree_of
        self._keep = []
        descs = (TreeDescC * len(trees))()
        for k, s in enumerate(trees):
            fill_tree_desc(descs[k], s, self._keep)
        with torch.cuda.device(self.device):
            h = ctypes.c_void_p()
            _check(lib().tdyn_batch_create(descs, len(trees),
                                           tree_of.ctypes.data_as(_i32) if tree_of is not None else None,
                                           n_instances, ctypes.byref(h)))
        self.h = h
        info = (ctypes.c_int64 * 8)()
        _check(lib().tdyn_batch_info(self.h, info))
        (self.n_instances, self.total_q, self.total_u, self.total_b,
         self.max_nq, self.max_nu, self.max_nb, self.n_trees) = list(info)
        qo = np.zeros(self.n_instances, np.int32); uo = np.zeros_like(qo); bo = np.zeros_like(qo)
        _check(lib().tdyn_batch_offsets(self.h, qo.ctypes.data_as(_i32), uo.ctypes.data_as(_i32),
                                        bo.ctypes.data_as(_i32)))
        self.q_off, self.u_off, self.b_off = qo, uo, bo

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.tdyn_batch_destroy(self.h)
            self.h = None

    def set_evaluator(self, mode):
        """0 automatic, 1 tree sweeps, 2 register-resident serial-chain
        evaluator, 3 level-parallel tree evaluator."""
        _check(lib().tdyn_batch_set_evaluator(self.h, int(mode)))

    def set_tree_lanes(self, lanes):
        """Lanes per instance of the tree evaluators: 0 automatic (per
        instance), 1 streaming (one lane per instance), 4/8/16/32/64
        level-parallel."""
        _check(lib().tdyn_batch_set_tree_lanes(self.h, int(lanes)))

    def set_tree_lane_range(self, lanes_min=0, lanes_max=32):
        """Range of the automatic lanes per instance (lanes_min 0: from the
        batch size)."""
        _check(lib().tdyn_batch_set_tree_lane_range(self.h, int(lanes_min), int(lanes_max)))

    def active_evaluator(self):
        """(evaluator, lanes per instance) of the current rollout: evaluator 0
        none yet, 1 tree sweeps, 2 chain, 3 streaming tree evaluator, 4
        level-parallel tree evaluator; lanes 0 when chosen per instance."""
        v = ctypes.c_int(0); n = ctypes.c_int(0)
        _check(lib().tdyn_batch_get_evaluator(self.h, ctypes.byref(v), ctypes.byref(n)))
        return v.value, n.value

    def uses_tree_stream(self):
        """True when the current rollout runs on one of the tree evaluators for
        batches of different trees (streaming or level-parallel)."""
        return self.active_evaluator()[0] in (3, 4)

    # ---- helpers ---------------------------------------------------------
    def _new(self, n):
        return self.torch.zeros(int(n), dtype=self.torch.float64, device=self.device)

    def to_device(self, a):
        return self.torch.as_tensor(np.ascontiguousarray(a, np.float64).reshape(-1), device=self.device)

    def pack(self, per_instance):
        """Concatenate per-instance arrays into the packed layout (device)."""
        return self.to_device(np.concatenate([np.asarray(x, np.float64).reshape(-1) for x in per_instance]))

    def split_q(self, packed):
        a = packed.detach().cpu().numpy() if hasattr(packed, "detach") else packed
        ends = list(self.q_off[1:]) + [self.total_q]
        return [a[s:e] for s, e in zip(self.q_off, ends)]

    def split_u(self, packed):
        a = packed.detach().cpu().numpy() if hasattr(packed, "detach") else packed
        ends = list(self.u_off[1:]) + [self.total_u]
        return [a[s:e] for s, e in zip(self.u_off, ends)]

    def split_b(self, packed, width):
        a = packed.detach().cpu().numpy() if hasattr(packed, "detach") else packed
        a = a.reshape(-1, width)
        ends = list(self.b_off[1:]) + [self.total_b]
        return [a[s:e] for s, e in zip(self.b_off, ends)]

    # ---- SimbodyMatterSubsystem surface -----------------------------------
    def realizeKinematics(self, q, u, want_X=True, want_V=True):
        X = self._new(12 * self.total_b) if want_X else None
        V = self._new(6 * self.total_b) if want_V else None
        qd = self._new(self.total_q); E = self._new(2 * self.n_instances)
        _check(lib().tdyn_realize_kinematics(self.h, _ptr(q), _ptr(u), _ptr(X), _ptr(V), _ptr(qd),
                                             _ptr(E), _stream()))
        return X, V, qd, E

    def op_status(self):
        """Per-instance status of the last operator that formed the
        articulated-body inertias (int32 device tensor): 0, or 4
        (TDYN_E_SINGULAR) where some body's D was not invertible."""
        import torch
        st = torch.empty(self.n_instances, dtype=torch.int32, device=self.device)
        _check(lib().tdyn_batch_op_status(self.h, _ptr(st), None, _stream()))
        return st

    def check_singular(self):
        """Raises TdynError(TDYN_E_SINGULAR) when the last operator that formed
        the articulated-body inertias met a singular D in some instance (what
        the reference throws, SmallMatrixMixed.h:878-892)."""
        n = ctypes.c_int(0)
        _check(lib().tdyn_batch_op_status(self.h, None, ctypes.byref(n), _stream()))

    def calcKineticEnergy(self, q, u):
        _, _, _, E = self.realizeKinematics(q, u, False, False)
        return E[0::2]

    def realizeAcceleration(self, q, u, fapp=None):
        ud = self._new(self.total_u); qd = self._new(self.total_q)
        _check(lib().tdyn_realize_acceleration(self.h, _ptr(q), _ptr(u), _ptr(fapp), _ptr(ud), _ptr(qd),
                                               _stream()))
        return ud, qd

    def calcAccelerationIgnoringConstraints(self, q, u, fmob, F, want_A=True, want_qdotdot=True):
        ud = self._new(self.total_u)
        A = self._new(6 * self.total_b) if want_A else None
        qdd = self._new(self.total_q) if want_qdotdot else None
        _check(lib().tdyn_calc_acceleration_ignoring_constraints(
            self.h, _ptr(q), _ptr(u), _ptr(fmob), _ptr(F), _ptr(ud), _ptr(A), _ptr(qdd), _stream()))
        return ud, A, qdd

    def calcAccelerationWithKnownUDot(self, q, u, fmob, F, known, udot_known, want_A=True):
        """Forward dynamics with known udot on the bodies flagged in `known`
        (int32 device tensor, nb+1 per instance in the body layout of F):
        prescribed motion at the acceleration level and locks.  Returns
        (udot, tau, A_GB); tau is nonzero on the known mobilities only, with
        M udot + c + tau = fmob + J^T F."""
        ud = self._new(self.total_u)
        tau = self._new(self.total_u)
        A = self._new(6 * self.total_b) if want_A else None
        _check(lib().tdyn_calc_acceleration_with_known_udot(
            self.h, _ptr(q), _ptr(u), _ptr(fmob), _ptr(F), _ptr(known), _ptr(udot_known), _ptr(ud), _ptr(tau),
            _ptr(A), _stream()))
        return ud, tau, A

    def autotune(self, q, u, integ, t_probe, fapp=None, cache=None):
        """Pick the tree evaluators' lanes per instance by timing short probe
        rollouts of this state (tdyn_batch_autotune).  Returns (evaluator
        mode, lanes, lanes_min, lanes_max, from_cache)."""
        out = (ctypes.c_int32 * 5)()
        cfg = integ.c()
        _check(lib().tdyn_batch_autotune(self.h, _ptr(q), _ptr(u), _ptr(fapp), ctypes.byref(cfg),
                                         ctypes.c_double(t_probe), cache.encode() if cache else None, out,
                                         _stream()))
        return tuple(out)

    def set_known_udot(self, known=None, udot_known=None):
        """Known udot for the following rollouts (locks, constant prescribed
        udot): device tensors as for calcAccelerationWithKnownUDot, kept alive
        by this object; None, None clears."""
        self._known = (known, udot_known)
        _check(lib().tdyn_batch_set_known_udot(self.h, _ptr(known), _ptr(udot_known)))

    def set_constraints(self, tree, constraints):
        """Constraint::Ball / Constraint::Rod on tree `tree`: a list of
        (type, body1, body2, station1, station2, length), type CONS_BALL or
        CONS_ROD, bodies 0 = Ground, stations in the body frames; [] clears."""
        n = len(constraints)
        arr = (ConstraintDescC * max(n, 1))()
        for i, (ty, b1, b2, s1, s2, ln) in enumerate(constraints):
            arr[i].type, arr[i].body1, arr[i].body2 = int(ty), int(b1), int(b2)
            for k in range(3):
                arr[i].station1[k] = float(s1[k]); arr[i].station2[k] = float(s2[k])
            arr[i].length = float(ln)
        _check(lib().tdyn_batch_set_constraints(self.h, int(tree), n, arr))

    def calcConstrainedAcceleration(self, q, u, F=None):
        """realize(Acceleration) with constraints: (udot, multipliers
        [n_instances, MAX_CONSTRAINT_ROWS])."""
        ud = self._new(self.total_u); lam = self._new(MAX_CONSTRAINT_ROWS * self.n_instances)
        _check(lib().tdyn_calc_constrained_acceleration(self.h, _ptr(q), _ptr(u), _ptr(F), _ptr(ud), _ptr(lam),
                                                       _stream()))
        return ud, lam.view(self.n_instances, MAX_CONSTRAINT_ROWS)

    def set_witnesses(self, tree, witnesses):
        """Event witnesses of tree `tree` (TriggeredEventHandler /
        TriggeredEventReporter): a list of dicts with `kind` (WIT_*) and any
        of index, body, trigger (TRIGGER_*), action (EVENT_*), action_index,
        station, direction, offset, factor, window (WITNESS_DEFAULTS); []
        clears."""
        n = len(witnesses)
        arr = (WitnessDescC * max(n, 1))()
        for i, w in enumerate(witnesses):
            w = dict(WITNESS_DEFAULTS, **w)
            d = arr[i]
            d.kind, d.index, d.body = int(w["kind"]), int(w["index"]), int(w["body"])
            d.trigger, d.action, d.action_index = int(w["trigger"]), int(w["action"]), int(w["action_index"])
            for k in range(3):
                d.station[k] = float(w["station"][k]); d.direction[k] = float(w["direction"][k])
            d.offset, d.factor, d.window = float(w["offset"]), float(w["factor"]), float(w["window"])
        _check(lib().tdyn_batch_set_witnesses(self.h, int(tree), n, arr))

    def calcEventTriggers(self, q, u, t=None):
        """The witnesses' values at (q, u, t) (State::getEventTriggers after
        realize): [n_instances, MAX_WITNESSES], zero past each tree's."""
        out = self._new(MAX_WITNESSES * self.n_instances)
        _check(lib().tdyn_calc_event_triggers(self.h, _ptr(q), _ptr(u), _ptr(t), _ptr(out), _stream()))
        return out.view(self.n_instances, MAX_WITNESSES)

    def set_event_capacity(self, capacity):
        """Events recorded per instance per rollout (later ones are only counted)."""
        self._ev_cap = int(capacity)
        _check(lib().tdyn_batch_set_event_capacity(self.h, self._ev_cap))

    def rollout_events(self):
        """The current rollout's events: windows [n, capacity, 2] (tLow,
        tHigh), ids [n, capacity, 2] (witness, transition: TRIGGER_FALLING
        or TRIGGER_RISING) and counts [n] (events seen)."""
        torch = self.torch
        cap = getattr(self, "_ev_cap", 64)
        n = self.n_instances
        w = self._new(n * cap * 2)
        ids = torch.zeros(n * cap * 2, dtype=torch.int32, device=self.device)
        cnt = torch.zeros(n, dtype=torch.int32, device=self.device)
        _check(lib().tdyn_rollout_get_events(self.h, _ptr(w), _ptr(ids), _ptr(cnt), _stream()))
        return w.view(n, cap, 2), ids.view(n, cap, 2), cnt

    def stepToWithEvents(self, q, u, integ, t_end, fapp=None, t0=0.0):
        """TimeStepper::stepTo(t_end) with the batch's event witnesses:
        returns a dict of q, u, t, status (1 reached t_end, STATUS_TERMINATED
        a handler ended it), stats [n, 4], windows, ids, counts
        (rollout_events)."""
        self.rollout_begin(q, u, integ, t_end, fapp=fapp, t0=t0)
        q1, u1, t1, st, stats = self._run_to_end()
        if bool((st < 0).any()):
            raise TdynError(5, "Integrator::StepFailed for at least one instance")
        w, ids, cnt = self.rollout_events()
        return dict(q=q1, u=u1, t=t1, status=st, stats=stats, windows=w, ids=ids, counts=cnt)

    def set_motion(self, motion=None):
        """Prescribed motion for the following rollouts (Motion::Sinusoid,
        Motion::Steady): a float64 device tensor of 4 per body (nb+1 per
        instance, Ground first): level (1 position, 2 velocity, 3
        acceleration sinusoid, 4 steady velocity), amplitude, rate, phase;
        None clears."""
        self._motion = motion
        _check(lib().tdyn_batch_set_motion(self.h, _ptr(motion)))

    def pack_body_flags(self, per_instance):
        """Concatenate per-instance body flags (nb+1 each, Ground first) into
        the packed int32 device layout of `known`."""
        return self.torch.as_tensor(np.concatenate([np.asarray(x, np.int32).reshape(-1) for x in per_instance]),
                                    dtype=self.torch.int32, device=self.device)

    def multiplyByMInv(self, q, v, out=None):
        out = self._new(self.total_u) if out is None else out
        _check(lib().tdyn_multiply_by_m_inv(self.h, _ptr(q), _ptr(v), _ptr(out), _stream()))
        return out

    def calcResidualForceIgnoringConstraints(self, q, u, fmob=None, F=None, udot=None):
        out = self._new(self.total_u)
        _check(lib().tdyn_calc_residual_force_ignoring_constraints(
            self.h, _ptr(q), _ptr(u), _ptr(fmob), _ptr(F), _ptr(udot), _ptr(out), _stream()))
        return out

    def realizeArticulatedBodyInertias(self, q, want_plus=True):
        P = self._new(21 * self.total_b)
        PP = self._new(21 * self.total_b) if want_plus else None
        _check(lib().tdyn_calc_articulated_body_inertias(self.h, _ptr(q), _ptr(P), _ptr(PP), _stream()))
        return P, PP

    def calcMobilizerReactionForces(self, q, u, fmob, F):
        out = self._new(6 * self.total_b)
        _check(lib().tdyn_calc_mobilizer_reaction_forces(self.h, _ptr(q), _ptr(u), _ptr(fmob), _ptr(F),
                                                         _ptr(out), _stream()))
        return out

    def multiplyByM(self, q, v, out=None):
        """M v per instance (packed like u)."""
        out = self._new(self.total_u) if out is None else out
        _check(lib().tdyn_multiply_by_m(self.h, _ptr(q), _ptr(v), _ptr(out), _stream()))
        return out

    def multiplyBySystemJacobian(self, q, v, out=None):
        """J v: (nb+1)*6 per instance, Ground's entry zero."""
        out = self._new(6 * self.total_b) if out is None else out
        _check(lib().tdyn_multiply_by_system_jacobian(self.h, _ptr(q), _ptr(v), _ptr(out), _stream()))
        return out

    def multiplyBySystemJacobianTranspose(self, q, F_G, out=None):
        """J^T F_G (packed like u); F_G (nb+1)*6 per instance, Ground ignored."""
        out = self._new(self.total_u) if out is None else out
        _check(lib().tdyn_multiply_by_system_jacobian_transpose(self.h, _ptr(q), _ptr(F_G), _ptr(out),
                                                                _stream()))
        return out

    def calcBodyAccelerationFromUDot(self, q, u, udot=None, out=None):
        """J udot + JDot u: (nb+1)*6 per instance; udot None means zero."""
        out = self._new(6 * self.total_b) if out is None else out
        _check(lib().tdyn_calc_body_acceleration_from_udot(self.h, _ptr(q), _ptr(u), _ptr(udot), _ptr(out),
                                                           _stream()))
        return out

    def calcCompositeBodyInertias(self, q, out=None):
        """(nb+1)*10 per instance: mass, mass center (3), unit inertia about
        the body origin (xx yy zz xy xz yz), in G; Ground first."""
        out = self._new(10 * self.total_b) if out is None else out
        _check(lib().tdyn_calc_composite_body_inertias(self.h, _ptr(q), _ptr(out), _stream()))
        return out

    # ---- RungeKuttaMersonIntegrator + TimeStepper ------------------------
    def rollout_begin(self, q, u, integ, t_end, fapp=None, t0=0.0):
        cfg = integ.c()
        if hasattr(t_end, "is_cuda"):
            _check(lib().tdyn_rollout_begin(self.h, _ptr(q), _ptr(u), _ptr(fapp), ctypes.c_double(t0),
                                            _ptr(t_end), ctypes.c_double(0.0), ctypes.byref(cfg), _stream()))
        else:
            _check(lib().tdyn_rollout_begin(self.h, _ptr(q), _ptr(u), _ptr(fapp), ctypes.c_double(t0),
                                            None, ctypes.c_double(float(t_end)), ctypes.byref(cfg), _stream()))

    def rollout_advance(self, max_steps):
        _check(lib().tdyn_rollout_advance(self.h, int(max_steps), _stream()))

    def rollout_extend(self, t_end):
        """Move every instance's final time to t_end (a float or a device
        tensor per instance); finished instances resume from their state and
        step size (stepTo(report time) with interpolation off)."""
        if hasattr(t_end, "is_cuda"):
            _check(lib().tdyn_rollout_extend(self.h, _ptr(t_end), ctypes.c_double(0.0), _stream()))
        else:
            _check(lib().tdyn_rollout_extend(self.h, None, ctypes.c_double(float(t_end)), _stream()))

    def _run_to_end(self):
        while True:
            self.rollout_advance(1 << 20)
            q, u, t, st, stats = self.rollout_get()
            if not bool((st == 0).any()):
                return q, u, t, st, stats

    def stepToReports(self, q, u, integ, report_times, fapp=None, t0=0.0):
        """TimeStepper::stepTo(t_k) for each report time in turn, with
        Integrator::setAllowInterpolation(false): no step crosses a report
        time and the step size carries across them; the last report time is
        the final time.  Returns (q [k], u [k] per report time, stats[n,4])."""
        times = [float(x) for x in report_times]
        if not times or any(b <= a for a, b in zip(times, times[1:])) or times[0] < t0:
            raise ValueError("report times must increase from t0")
        self.rollout_begin(q, u, integ, times[0], fapp=fapp, t0=t0)
        qs, us = [], []
        for k, tk in enumerate(times):
            if k:
                self.rollout_extend(tk)
            q1, u1, _, st, stats = self._run_to_end()
            if bool((st < 0).any()):
                raise TdynError(5, "Integrator::StepFailed for at least one instance")
            qs.append(q1); us.append(u1)
        return qs, us, stats

    def stepToWithReports(self, q, u, integ, t_end, report_times, fapp=None, t0=0.0):
        """stepTo(t_end) with reports by interpolation (Simbody's default):
        the trajectory of stepTo(t_end), and at each report time the Hermite
        interpolant of the step that reaches it.  Returns (q(t_end),
        u(t_end), [q at each report], [u at each report], stats[n,4])."""
        torch = self.torch
        q = q.clone(); u = u.clone()
        times = np.ascontiguousarray(report_times, dtype=np.float64)
        n = len(times)
        Qr = self._new(max(n, 1) * self.total_q); Ur = self._new(max(n, 1) * self.total_u)
        stats = torch.zeros(4 * self.n_instances, dtype=torch.int32, device=self.device)
        cfg = integ.c()
        _check(lib().tdyn_rollout_reports(self.h, _ptr(q), _ptr(u), _ptr(fapp), ctypes.c_double(t0),
                                          ctypes.c_double(t_end), times.ctypes.data_as(ctypes.c_void_p),
                                          ctypes.c_int(n), ctypes.byref(cfg), _ptr(Qr), _ptr(Ur), _ptr(stats),
                                          _stream()))
        Qr = Qr.view(-1, self.total_q); Ur = Ur.view(-1, self.total_u)
        return q, u, [Qr[k] for k in range(n)], [Ur[k] for k in range(n)], stats.view(-1, 4)

    def rollout_get(self):
        torch = self.torch
        q = self._new(self.total_q); u = self._new(self.total_u); t = self._new(self.n_instances)
        st = torch.zeros(self.n_instances, dtype=torch.int32, device=self.device)
        stats = torch.zeros(4 * self.n_instances, dtype=torch.int32, device=self.device)
        _check(lib().tdyn_rollout_get(self.h, _ptr(q), _ptr(u), _ptr(t), _ptr(st), _ptr(stats), _stream()))
        return q, u, t, st, stats.view(-1, 4)

    def rollout_snapshot(self):
        """The current rollout's integrator state as bytes (the layout of
        tdyn_rollout_snapshot): a checkpoint to store or send."""
        n = ctypes.c_int64(0)
        _check(lib().tdyn_rollout_snapshot(self.h, None, 0, ctypes.byref(n), _stream()))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().tdyn_rollout_snapshot(self.h, buf, n, ctypes.byref(n), _stream()))
        return buf.raw

    def rollout_restore(self, snapshot, fapp=None):
        """Resume a rollout from rollout_snapshot() bytes taken on a batch of
        the same model (advance / get / events continue from there)."""
        buf = ctypes.create_string_buffer(bytes(snapshot), len(snapshot))
        _check(lib().tdyn_rollout_restore(self.h, buf, len(snapshot), _ptr(fapp), _stream()))

    def rollout_derivs(self):
        """(qdot, udot) of the current rollout state by the rollout's
        evaluator and precision (the integrator's f0)."""
        qd = self._new(self.total_q); ud = self._new(self.total_u)
        _check(lib().tdyn_rollout_derivs(self.h, _ptr(qd), _ptr(ud), _stream()))
        return qd, ud

    def stepTo_noalloc(self, q, u, integ, t_end, fapp=None, t0=0.0, stats=None):
        """stepTo through tdyn_rollout_noalloc: no device allocation in the
        call (the integrator state lives in the library's fixed device
        array); q and u are advanced in place and stats (int32 device tensor
        of 4 per instance, or None) receives the statistics."""
        cfg = integ.c()
        _check(lib().tdyn_rollout_noalloc(self.h, _ptr(q), _ptr(u), _ptr(fapp), ctypes.c_double(t0),
                                          ctypes.c_double(t_end), ctypes.byref(cfg), _ptr(stats), _stream()))
        return q, u, stats

    def stepTo(self, q, u, integ, t_end, fapp=None, t0=0.0):
        """TimeStepper(system, integ).initialize(state); stepTo(t_end), with
        Integrator::setFinalTime(t_end).  Returns (q, u, stats[n,4])."""
        torch = self.torch
        q = q.clone(); u = u.clone()
        stats = torch.zeros(4 * self.n_instances, dtype=torch.int32, device=self.device)
        cfg = integ.c()
        _check(lib().tdyn_rollout(self.h, _ptr(q), _ptr(u), _ptr(fapp), ctypes.c_double(t0),
                                  ctypes.c_double(t_end), ctypes.byref(cfg), _ptr(stats), _stream()))
        return q, u, stats.view(-1, 4)
