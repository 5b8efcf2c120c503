"""GPU parity of the integrator options setUseInfinityNorm,
setProjectEveryStep and setProjectInterpolatedStates (Integrator.cpp:276-296)
against the oracle: the oracle's accepted, rejected and projection counts
and its states within 1e-8 relative, on every evaluator that takes them --
the tree sweeps (a chain batch, which leaves the chain evaluator for them;
one-lane streaming over different trees; constrained trees), the
level-parallel evaluator over different trees, the tree server -- and the
report and event states with interpolated states unprojected.  The chain
evaluator, the chain server and the mixed path refuse the first two."""
import numpy as np
import pytest

from common import mixed_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel
import oracle_ctypes as O
from test_oracle_integrator_options import _ball_state, _free_ball

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

OPTS = [dict(use_infinity_norm=True), dict(project_every_step=True),
        dict(use_infinity_norm=True, project_every_step=True)]


def _oracle_opts(o):
    return dict(inf_norm=o.get("use_infinity_norm", False), project_every_step=o.get("project_every_step", False),
                project_interpolated=o.get("project_interpolated_states", True))


def _check(per, qs, us, q1s, u1s, st, acc, T, opts, Fs=None, cons=None):
    for i, t in enumerate(per):
        o = OracleModel(t)
        if cons is not None:
            assert o.set_constraints(cons) == 0
        o.set_integrator_options(**_oracle_opts(opts))
        y, s = o.integrate(np.concatenate([qs[i], us[i]]), T, acc, Fapp=None if Fs is None else Fs[i])
        assert (s["accepted"], s["rejected"], s["projections"]) == tuple(st[i, 1:4]), (i, s, st[i])
        assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, i


@pytest.mark.parametrize("opts", OPTS)
@pytest.mark.parametrize("kind", ["chain", "level-parallel", "streaming"])
def test_options_match_the_oracle(kind, opts):
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    acc, T = 1e-6, 0.05
    if kind == "chain":
        spec = M.spatial_chain(8, seed=3)
        per = [spec] * 16
        B = MatterBatch([spec], n_instances=16)
    else:
        per = mixed_trees(601, 12, nmin=4, nmax=30)
        B = MatterBatch(per, tree_of=np.arange(len(per), dtype=np.int32))
        if kind == "streaming":
            B.set_tree_lanes(1)
    qs, us = zip(*[M.random_state(t, 601, i) for i, t in enumerate(per)])
    Fs = [M.random_body_forces(t, 602, i) for i, t in enumerate(per)]
    q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=acc, **opts), T, fapp=B.pack(Fs))
    assert B.active_evaluator()[0] != 2
    _check(per, qs, us, B.split_q(q1), B.split_u(u1), st.cpu().numpy(), acc, T, opts, Fs)


@pytest.mark.parametrize("opts", OPTS)
def test_options_with_a_constraint(opts):
    """The point pendulum on a rod, states off the rod (the initial
    projection takes the options too)."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    spec = M.TreeSpec(parent=np.array([0], np.int32), mobilizer=np.array([M.TRANSLATION], np.int32),
                      mass=np.array([10.0]), com_B=np.zeros((1, 3)), inertia_B=np.array([[1e-3, 1e-3, 1e-3, 0, 0, 0]]),
                      X_PF=np.array([M._xform()]), X_BM=np.array([M._xform()]), gravity=np.array([0.0, -9.81, 0.0]))
    cons = [(1, 0, 1, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1.0)]
    n = 8
    rng = np.random.default_rng(5)
    qs = [np.array([np.cos(a), np.sin(a), 0.0]) * (1 + 1e-4 * rng.normal()) for a in rng.uniform(-3, 3, n)]
    us = [rng.normal(size=3) for _ in range(n)]
    B = MatterBatch([spec], n_instances=n)
    B.set_constraints(0, cons)
    q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-6, **opts), 0.2)
    _check([spec] * n, qs, us, B.split_q(q1), B.split_u(u1), st.cpu().numpy(), 1e-6, 0.2, opts, cons=cons)


@pytest.mark.parametrize("opts", OPTS)
def test_tree_server_options(opts):
    from simbody_amd.matter import RungeKuttaMerson
    from simbody_amd.serving import TreeServer
    per = mixed_trees(611, 10, nmin=4, nmax=30)
    qs, us = zip(*[M.random_state(t, 611, i) for i, t in enumerate(per)])
    T = 0.04
    sv = TreeServer(RungeKuttaMerson(accuracy=1e-6, **opts), 1 << 28, max_slots=32)
    slots = list(sv.admit(per, list(qs), list(us), np.full(len(per), T)))
    out = {}
    while len(out) < len(per):
        sv.advance(50)
        r = sv.collect()
        for j, s in enumerate(r["slots"]):
            out[slots.index(int(s))] = (r["y"][j], r["stats"][j])
    for i, t in enumerate(per):
        o = OracleModel(t)
        o.set_integrator_options(**_oracle_opts(opts))
        y, s = o.integrate(np.concatenate([qs[i], us[i]]), T, 1e-6)
        assert (s["accepted"], s["rejected"], s["projections"]) == tuple(out[i][1][1:4]), i
        assert rel_inf(out[i][0], y) < 1e-8, i


def test_unprojected_reports_and_events():
    """Torque-free Balls: reports with interpolated states unprojected equal
    the oracle's; a time witness that terminates with a velocity-scaling
    handler leaves the projected back-up state, as the oracle's."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    spec, y = _free_ball(), _ball_state()
    n = 4
    rng = np.random.default_rng(9)
    ys = [np.concatenate([y[:4], y[4:] * (1 + 0.2 * rng.normal())]) for _ in range(n)]
    qs, us = [v[:4] for v in ys], [v[4:] for v in ys]
    reports = np.linspace(0.013, 0.987, 41)
    integ = RungeKuttaMerson(accuracy=1e-3, project_interpolated_states=False)
    B = MatterBatch([spec], n_instances=n)
    _, _, Qr, Ur, st = B.stepToWithReports(B.pack(qs), B.pack(us), integ, 1.0, reports)
    for i in range(n):
        o = OracleModel(spec)
        o.set_integrator_options(project_interpolated=False)
        ro, _, _ = o.integrate_reports(ys[i], 1.0, reports, 1e-3)
        g = np.stack([np.concatenate([B.split_q(Qr[k])[i], B.split_u(Ur[k])[i]]) for k in range(len(reports))])
        assert np.max(np.abs(g - ro)) < 1e-8 * max(1.0, np.max(np.abs(ro))), i
    wit = [dict(kind=O.WIT_TIME, offset=0.43, action=O.EV_TERMINATE),
           dict(kind=O.WIT_TIME, offset=0.43, action=O.EV_SCALE_U, factor=0.5)]
    B2 = MatterBatch([spec], n_instances=n)
    B2.set_witnesses(0, wit)
    r = B2.stepToWithEvents(B2.pack(qs), B2.pack(us), integ, 1.0)
    q1, u1 = B2.split_q(r["q"]), B2.split_u(r["u"])
    for i in range(n):
        o = OracleModel(spec)
        o.set_integrator_options(project_interpolated=False)
        assert o.set_witnesses(wit) == 0
        yo, _ = o.integrate(ys[i], 1.0, 1e-3)
        assert rel_inf(np.concatenate([q1[i], u1[i]]), yo) < 1e-8, i
        assert abs(np.linalg.norm(q1[i]) - 1.0) <= 1e-4


def test_refusals():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson, TdynError
    from simbody_amd.serving import Server
    spec = M.spatial_chain(8, seed=3)
    q, u = M.random_state(spec, 1, 0)
    B = MatterBatch([spec], n_instances=1)
    B.set_evaluator(2)
    with pytest.raises(TdynError):
        B.stepTo(B.pack([q]), B.pack([u]), RungeKuttaMerson(accuracy=1e-6, use_infinity_norm=True), 0.01)
    B.set_evaluator(0)
    with pytest.raises(TdynError):
        B.stepTo(B.pack([q]), B.pack([u]), RungeKuttaMerson(accuracy=1e-6, precision=1, project_every_step=True), 0.01)
    with pytest.raises(TdynError):
        Server(spec, RungeKuttaMerson(accuracy=1e-6, use_infinity_norm=True), 1 << 28, max_slots=64)
