"""GPU parity of holonomic constraints (Constraint::Ball, Constraint::Rod)
with the oracle, through the C ABI: the loop forward dynamics
(SimbodyMatterSubsystemRep.cpp:5076-5152) at 1e-10 relative in udot and the
multipliers, and rollouts whose initial and per-step projections (projectQ,
projectU, attemptDAEStep AbstractIntegratorRep.cpp:137-208) reproduce the
oracle's accepted and rejected step counts with states within 1e-8, on
batches of one tree (rollouts of batches of different constrained trees are
switched off); the single-shot operator on a batch of different trees; the
settings constraints cannot run with are refused."""
import numpy as np
import pytest

from common import rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel
from test_oracle_constraints import _ball_pair, _loop_state, _one_body, _spatial_loop, _tip

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

BALL_C, ROD_C = 0, 1
G = 9.8


def _four_bar():
    spec = M.TreeSpec(parent=np.array([0, 1], np.int32), mobilizer=np.array([M.PIN, M.PIN], np.int32),
                      mass=np.array([1.0, 1.0]), com_B=np.array([[0.5, 0, 0], [0.5, 0, 0]]),
                      inertia_B=np.tile([0.01, 0.34, 0.34, 0, 0, 0], (2, 1)),
                      X_PF=np.array([M._xform(), M._xform(p=(1.0, 0, 0))]),
                      X_BM=np.array([M._xform(), M._xform()]), gravity=np.array([0.0, -G, 0.0]))
    cons = [(ROD_C, 0, 2, (1.2, 0, 0), (1.0, 0, 0), 1.0)]
    t1 = np.deg2rad(80.0)
    Bp = np.array([np.cos(t1), np.sin(t1)]); D = np.array([1.2, 0.0])
    d = np.linalg.norm(D - Bp)
    t2 = np.arctan2(*(D - Bp)[::-1]) + np.arccos(d / 2) - t1
    return spec, cons, np.array([t1, t2])


def _cases(seed):
    """(spec, constraints, q, u, F) for several constrained models."""
    r = np.random.default_rng(seed)
    out = []
    # Free body on a Ball constraint
    A, B, qA, uA, qB, uB, s = _ball_pair(seed)
    out.append((_one_body(M.FREE, s), [(BALL_C, 0, 1, (0, 0, 0), s, 0.0)], qB, uB))
    # spatial loop: two Balls closed by a Ball constraint
    spec = _spatial_loop()
    q, u = _loop_state()
    P = _tip(OracleModel(spec), q)
    out.append((spec, [(BALL_C, 0, 2, tuple(P), (0, 1.0, 0), 0.0)], q, u + 0.2 * r.normal(size=6)))
    # planar four-bar closed by a Rod
    spec, cons, q = _four_bar()
    out.append((spec, cons, q, np.array([1.0 + r.normal() * 0.3, 0.0])))
    res = []
    for spec, cons, q, u in out:
        F = np.zeros((spec.nb + 1, 6)); F[1:] = 0.3 * r.normal(size=(spec.nb, 6))
        res.append((spec, cons, np.asarray(q, float), np.asarray(u, float), F))
    return res


def _batch(cases):
    from simbody_amd.matter import MatterBatch
    specs = [c[0] for c in cases]
    B = MatterBatch(specs, tree_of=np.arange(len(specs), dtype=np.int32))
    for i, c in enumerate(cases):
        B.set_constraints(i, c[1])
    return B


def _oracle(spec, cons):
    o = OracleModel(spec)
    assert o.set_constraints(cons) == 0
    return o


def test_loop_forward_dynamics_match_oracle():
    cases = _cases(1) + _cases(2)
    B = _batch(cases)
    q = B.pack([c[2] for c in cases]); u = B.pack([c[3] for c in cases]); F = B.pack([c[4] for c in cases])
    ud, lam = B.calcConstrainedAcceleration(q, u, F)
    ud2, _ = B.realizeAcceleration(q, u, F)
    uds, lam = B.split_u(ud), lam.cpu().numpy()
    assert np.array_equal(ud.cpu().numpy(), ud2.cpu().numpy())
    for i, (spec, cons, qi, ui, Fi) in enumerate(cases):
        o_ud, o_lam, rc = _oracle(spec, cons).loop_forward(qi, ui, Fi)
        assert rc == 0
        assert rel_inf(uds[i], o_ud) < 1e-10, i
        assert rel_inf(lam[i, :len(o_lam)], o_lam) < 1e-10, i
        assert np.all(lam[i, len(o_lam):] == 0)


@pytest.mark.parametrize("acc", [1e-6, 1e-8])
@pytest.mark.parametrize("model", [0, 1, 2])
def test_constrained_rollouts_match_oracle(acc, model):
    """Rollouts of one constrained tree (Free body on a Ball constraint,
    spatial loop, four-bar) over 8 instances with their own states and
    forces: the oracle's step counts and states, on both settings of the
    tree sweeps."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    cases = [_cases(10 + k)[model] for k in range(8)]
    spec, cons = cases[0][0], cases[0][1]
    if model == 0:                     # the station differs per seed: one seed's model for all
        cases = [(spec, cons) + _ball_state(np.asarray(cons[0][4]), k) for k in range(8)]
    B = MatterBatch([spec], n_instances=len(cases))
    B.set_constraints(0, cons)
    T = 0.5
    rkm = RungeKuttaMerson(accuracy=acc)
    o = _oracle(spec, cons)
    ref = [o.integrate(np.concatenate([c[2], c[3]]), T, acc, Fapp=c[4]) for c in cases]
    for mode in (0, 1):
        B.set_evaluator(mode)
        q1, u1, st = B.stepTo(B.pack([c[2] for c in cases]), B.pack([c[3] for c in cases]), rkm, T,
                              fapp=B.pack([c[4] for c in cases]))
        assert B.active_evaluator()[0] == 1
        q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
        for i in range(len(cases)):
            y, s = ref[i]
            assert s["status"] == 0
            assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), (mode, i, s, st[i])
            assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, (mode, i)


def _ball_state(s, k):
    """A state of the Free body held at station s by a Ball constraint at the
    Ground origin (on the constraint: p = -R s, v = w x p), and body forces."""
    r = np.random.default_rng(300 + k)
    q = r.normal(size=4); q /= np.linalg.norm(q)
    w = r.normal(size=3)
    p = -M.quat_to_rot(q) @ s
    F = np.zeros((2, 6)); F[1] = 0.3 * r.normal(size=6)
    return np.concatenate([q, p]), np.concatenate([w, np.cross(w, p)]), F


def test_shared_constrained_tree_and_initial_projection():
    """One constrained tree for every instance (typed records); instances
    start off the velocity constraint and, for some, off the position
    constraint by 1e-4: the initial projection puts them on it as the
    oracle's does."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    spec = _spatial_loop()
    q0, u0 = _loop_state()
    cons = [(BALL_C, 0, 2, tuple(_tip(OracleModel(spec), q0)), (0, 1.0, 0), 0.0)]
    n = 24
    r = np.random.default_rng(5)
    qs, us = [], []
    for i in range(n):
        q = q0.copy()
        if i % 3 == 0:
            q[4:8] += 1e-4 * r.normal(size=4)
        qs.append(q); us.append(u0 + 0.5 * r.normal(size=6))
    B = MatterBatch([spec], n_instances=n)
    B.set_constraints(0, cons)
    rkm = RungeKuttaMerson(accuracy=1e-6)
    q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), rkm, 0.3)
    q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
    o = _oracle(spec, cons)
    for i in range(n):
        y, s = o.integrate(np.concatenate([qs[i], us[i]]), 0.3, 1e-6)
        assert s["status"] == 0
        assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), (i, s, st[i])
        assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, i
        pe, ve, _ = o.constraint_errors(q1s[i], u1s[i])
        assert np.abs(pe).max() < 1e-6 and np.abs(ve).max() < 1e-6


def test_settings_constraints_cannot_run_with_are_refused():
    """Refused on the host, before any launch: the level-parallel evaluator,
    the mixed path, the rollout without allocation, reports, rollouts of a
    batch of different constrained trees (switched off: see the library's
    tdyn_batch_set_constraints notes), and invalid descriptions."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson, TdynError
    spec = _spatial_loop()
    q0, u0 = _loop_state()
    cons = [(BALL_C, 0, 2, tuple(_tip(OracleModel(spec), q0)), (0, 1.0, 0), 0.0)]
    B = MatterBatch([spec], n_instances=4)
    B.set_constraints(0, cons)
    q = B.pack([q0] * 4); u = B.pack([u0] * 4)
    rkm = RungeKuttaMerson(accuracy=1e-6)
    B.set_evaluator(2)
    with pytest.raises(TdynError):
        B.stepTo(q, u, rkm, 0.01)
    B.set_evaluator(0)
    with pytest.raises(TdynError):
        B.stepTo(q, u, RungeKuttaMerson(accuracy=1e-6, precision=1), 0.01)
    with pytest.raises(TdynError):
        B.stepTo_noalloc(q, u, rkm, 0.01)
    with pytest.raises(TdynError):
        B.stepToWithReports(q, u, rkm, 0.02, [0.01])
    B.stepTo(q, u, rkm, 0.01)
    cases = _cases(6)
    C = _batch(cases)
    with pytest.raises(TdynError):
        C.stepTo(C.pack([c[2] for c in cases]), C.pack([c[3] for c in cases]), rkm, 0.01)
    for bad in ([(BALL_C, 2, 2, (0, 0, 0), (0, 0, 0), 0.0)],          # the same body twice
                [(BALL_C, 0, 5, (0, 0, 0), (0, 0, 0), 0.0)],          # no body 5
                [(ROD_C, 0, 1, (0, 0, 0), (0, 0, 0), -1.0)],          # rod length
                [(BALL_C, 0, 1, (0, 0, 0), (0, 0, 0), 0.0)] * 5):     # 15 rows
        with pytest.raises(TdynError):
            B.set_constraints(0, bad)
