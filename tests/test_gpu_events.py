"""GPU parity of event witnesses with the oracle, through the C ABI: the
trigger search, localization and handlers of takeOneStep
(AbstractIntegratorRep.cpp:583-756) and TimeStepper::stepTo
(TimeStepper.cpp:178-205) in the tree-sweep rollouts reproduce the
oracle's events (count, witness, transition, window to 1e-9 in time), step
counts and final states (1e-9) on a bouncing ball (whose windows must also
hold the exact parabola crossings), a pendulum with position, velocity and
rate witnesses and energy injection, a spatial chain with quaternions, and
terminating handlers (status 3, time tLow); the event log's capacity, a
rollout resumed launch by launch, a batch of different trees, the
witnesses' values at a state (getEventTriggers), and the settings witnesses
cannot run with are refused on the host."""
import numpy as np
import pytest

from common import rel_inf
from simbody_amd import models as M
import oracle_ctypes as O
from oracle_ctypes import OracleModel
from test_oracle_events import G, _crossing, _parabola, pendulum, period, slider

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _gpu(spec, wits, qs, us, integ, T, fapp=None, cap=None, evaluator=0, chunk=None):
    from simbody_amd.matter import MatterBatch
    B = MatterBatch([spec], n_instances=len(qs))
    B.set_witnesses(0, wits)
    if cap:
        B.set_event_capacity(cap)
    B.set_evaluator(evaluator)
    q, u = B.pack(qs), B.pack(us)
    F = None if fapp is None else B.pack(fapp)
    if chunk is None:
        r = B.stepToWithEvents(q, u, integ, T, fapp=F)
    else:
        # resumed launch by launch, `chunk` accepted steps each
        B.rollout_begin(q, u, integ, T, fapp=F)
        while True:
            B.rollout_advance(chunk)
            q1, u1, t1, st, stats = B.rollout_get()
            if not bool((st == 0).any()):
                break
        w, ids, cnt = B.rollout_events()
        r = dict(q=q1, u=u1, t=t1, status=st, stats=stats, windows=w, ids=ids, counts=cnt)
    out = dict(q=B.split_q(r["q"]), u=B.split_u(r["u"]))
    for k in ("t", "status", "stats", "windows", "ids", "counts"):
        out[k] = r[k].cpu().numpy()
    assert B.active_evaluator()[0] == 1
    return out


def _oracle(spec, wits, qs, us, acc, T, fapp=None, max_h=None):
    o = OracleModel(spec)
    assert o.set_witnesses(wits) == 0
    res = []
    for i in range(len(qs)):
        y, st = o.integrate(np.concatenate([qs[i], us[i]]), T, acc, max_h=max_h,
                            Fapp=None if fapp is None else np.ascontiguousarray(fapp[i]))
        win, ids, tf, term = o.last_events()
        res.append(dict(y=y, st=st, win=win, ids=ids, t=tf, term=term))
    return res


def _compare(g, ref, cap=None, wtol=1e-9):
    """wtol: how far the windows may move.  A trial landing within rounding
    of a root can choose the other half, which moves the final window by up
    to its required width (accuracy * time scale * window); a handler that
    changes the state then acts at a time known to that width, and later
    roots move with it: the oracle and its FMA-contracted build, two valid
    roundings, differ that way."""
    for i, r in enumerate(ref):
        assert r["st"]["status"] == 0
        n = len(r["win"])
        assert g["counts"][i] == n, (i, g["counts"][i], n)
        m = n if cap is None else min(n, cap)
        assert np.array_equal(g["ids"][i, :m], r["ids"][:m]), i
        assert np.abs(g["windows"][i, :m] - r["win"][:m]).max(initial=0) < wtol, i
        assert (r["st"]["accepted"], r["st"]["rejected"]) == tuple(g["stats"][i, 1:3]), (i, r["st"], g["stats"][i])
        assert g["status"][i] == (3 if r["term"] else 1)
        assert abs(g["t"][i] - r["t"]) < 1e-12, i
        assert rel_inf(np.concatenate([g["q"][i], g["u"][i]]), r["y"]) < 1e-9, i


@pytest.mark.parametrize("acc", [1e-3, 1e-6])
def test_bouncing_balls_match_oracle_and_exact_crossings(acc):
    from simbody_amd.matter import RungeKuttaMerson
    r = np.random.default_rng(3)
    n, e, T = 16, 0.8, 2.0
    qs = [np.array([0.5 + r.uniform()]) for _ in range(n)]
    us = [np.array([r.uniform(-1, 2)]) for _ in range(n)]
    wits = [dict(kind=O.WIT_Q, index=0, trigger=O.TRIG_FALLING, action=O.EV_SCALE_U, action_index=0, factor=-e)]
    g = _gpu(slider(), wits, qs, us, RungeKuttaMerson(accuracy=acc, max_step=0.05), T)
    ref = _oracle(slider(), wits, qs, us, acc, T, max_h=0.05)
    _compare(g, ref)
    for i in range(n):
        x, v, t = qs[i][0], us[i][0], 0.0
        assert g["counts"][i] >= 2
        for lo, hi in g["windows"][i, :g["counts"][i]]:
            tc = t + _crossing(x, v)
            assert lo - 1e-12 < tc <= hi + 1e-12 and hi - lo <= acc * 0.01
            x, v = _parabola(x, v, hi - t)
            v, t = -e * v, hi


def _pendulum_case(n=8, seed=4):
    r = np.random.default_rng(seed)
    spec, I = pendulum()
    qs = [np.array([r.uniform(-0.3, 0.3)]) for _ in range(n)]
    us = [np.array([r.uniform(-1.0, 1.0)]) for _ in range(n)]
    wits = [dict(kind=O.WIT_STATION_POS, body=1, station=(1.0, 0, 0), direction=(1.0, 0, 0),
                 action=O.EV_SCALE_U, action_index=-1, factor=np.sqrt(1.1)),
            dict(kind=O.WIT_STATION_VEL, body=1, station=(1.0, 0, 0), direction=(0, 1.0, 0), trigger=O.TRIG_RISING),
            dict(kind=O.WIT_U, index=0, window=0.02),
            dict(kind=O.WIT_TIME, offset=0.77)]
    return spec, wits, qs, us, 1.7 * period(I)


@pytest.mark.parametrize("evaluator", [0, 1])
def test_pendulum_witnesses_match_oracle(evaluator):
    from simbody_amd.matter import RungeKuttaMerson
    spec, wits, qs, us, T = _pendulum_case()
    g = _gpu(spec, wits, qs, us, RungeKuttaMerson(accuracy=1e-7), T, evaluator=evaluator)
    _compare(g, _oracle(spec, wits, qs, us, 1e-7, T))


def test_resumed_rollout_and_event_capacity():
    from simbody_amd.matter import RungeKuttaMerson
    spec, wits, qs, us, T = _pendulum_case(seed=5)
    ref = _oracle(spec, wits, qs, us, 1e-7, T)
    g = _gpu(spec, wits, qs, us, RungeKuttaMerson(accuracy=1e-7), T, chunk=3)
    _compare(g, ref)
    g = _gpu(spec, wits, qs, us, RungeKuttaMerson(accuracy=1e-7), T, cap=2)
    assert np.all(g["counts"] > 2)
    _compare(g, ref, cap=2)


def test_spatial_chain_with_quaternions_matches_oracle():
    from simbody_amd.matter import RungeKuttaMerson
    spec = M.spatial_chain(nb=5, seed=7, types=(M.BALL, M.PIN, M.UNIVERSAL))
    spec.gravity = np.array([0.0, -G, 0.0])
    n = 8
    qs = [M.random_state(spec, 11, i, qscale=0.5, uscale=0.5)[0] for i in range(n)]
    us = [M.random_state(spec, 11, i, qscale=0.5, uscale=0.5)[1] for i in range(n)]
    fapp = [M.random_body_forces(spec, 11, i) for i in range(n)]
    wits = [dict(kind=O.WIT_STATION_POS, body=5, station=(0.1, 0.2, 0.0), direction=(0, 1.0, 0), offset=-1.0),
            dict(kind=O.WIT_STATION_VEL, body=3, station=(0.0, 0.3, 0.1), direction=(0.6, 0.0, 0.8),
                 trigger=O.TRIG_FALLING),
            dict(kind=O.WIT_U, index=1, action=O.EV_SCALE_U, action_index=2, factor=0.9)]
    g = _gpu(spec, wits, qs, us, RungeKuttaMerson(accuracy=1e-6), 1.5, fapp=fapp)
    ref = _oracle(spec, wits, qs, us, 1e-6, 1.5, fapp=fapp)
    assert sum(len(r["win"]) for r in ref) >= n
    _compare(g, ref, wtol=1e-6 * 0.1 * 0.1)


def test_terminating_handlers_stop_at_tlow():
    from simbody_amd.matter import RungeKuttaMerson, STATUS_TERMINATED
    n = 6
    qs = [np.array([1.0])] * n
    us = [np.array([v]) for v in np.linspace(-1.0, 4.0, n)]
    wits = [dict(kind=O.WIT_Q, index=0, trigger=O.TRIG_FALLING, action=O.EV_TERMINATE)]
    T = 0.9        # the fast throws have not landed by then
    g = _gpu(slider(), wits, qs, us, RungeKuttaMerson(accuracy=1e-6, max_step=0.05), T)
    ref = _oracle(slider(), wits, qs, us, 1e-6, T, max_h=0.05)
    _compare(g, ref)
    landed = [i for i in range(n) if _crossing(1.0, us[i][0]) < T]
    assert 0 < len(landed) < n
    for i in range(n):
        if i in landed:
            assert g["status"][i] == STATUS_TERMINATED
            lo, hi = g["windows"][i, 0]
            assert g["t"][i] == lo and lo - 1e-12 < _crossing(1.0, us[i][0]) <= hi + 1e-12
            assert g["q"][i][0] >= 0
        else:
            assert g["status"][i] == 1 and g["t"][i] == T and g["counts"][i] == 0


def test_settings_witnesses_cannot_run_with_are_refused():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson, TdynError
    spec, wits, qs, us, T = _pendulum_case(n=4)
    rkm = RungeKuttaMerson(accuracy=1e-6)
    B = MatterBatch([spec], n_instances=4)
    q, u = B.pack(qs), B.pack(us)
    with pytest.raises(TdynError):
        B.rollout_events()                                   # no rollout
    B.set_witnesses(0, wits)
    for ev in (2, 3):
        B.set_evaluator(ev)
        with pytest.raises(TdynError):
            B.stepTo(q, u, rkm, 0.1)
    B.set_evaluator(0)
    with pytest.raises(TdynError):
        B.stepTo_noalloc(q, u, rkm, 0.1)
    with pytest.raises(TdynError):
        B.stepToWithReports(q, u, rkm, 0.2, [0.1])
    B.stepTo(q, u, rkm, 0.1)
    for bad in ([dict(kind=O.WIT_Q, index=1)], [dict(kind=O.WIT_STATION_POS, body=2)],
                [dict(kind=O.WIT_Q, trigger=0)], [dict(kind=O.WIT_Q, window=0.0)],
                [dict(kind=9)], [dict(kind=O.WIT_Q, action=O.EV_SCALE_U, action_index=1)],
                [dict(kind=O.WIT_TIME)] * 9):
        with pytest.raises(TdynError):
            B.set_witnesses(0, bad)


def _scaled_diff(a, b):
    """SimTK-style relative difference: |a - b| / max(|a|, |b|, 1)."""
    return float(np.max(np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1.0)))


@pytest.mark.parametrize("shared", [False, True])
def test_event_triggers_match_oracle(shared):
    """State::getEventTriggers after realize (tdyn_calc_event_triggers) on a
    batch of different random trees with general joint frames (padded
    records), and on one shared tree: every witness kind at each instance's
    own time, 1e-12 against the oracle's witness values."""
    from simbody_amd.matter import MatterBatch
    n = 12
    specs = [M.random_tree(21, k % (1 if shared else 4), nmin=6, nmax=24, general_frames=True)
             for k in range(1 if shared else 4)]
    tree_of = np.array([k % len(specs) for k in range(n)], np.int32)
    r = np.random.default_rng(8)
    wits = []
    for s in specs:
        w = [dict(kind=O.WIT_Q, index=int(r.integers(s.nq)), offset=0.1),
             dict(kind=O.WIT_U, index=int(r.integers(s.nu)), offset=-0.2),
             dict(kind=O.WIT_TIME, offset=0.3)]
        for _ in range(3):
            w.append(dict(kind=O.WIT_STATION_POS, body=int(r.integers(s.nb + 1)), station=tuple(r.normal(size=3)),
                          direction=tuple(r.normal(size=3)), offset=float(r.normal())))
            w.append(dict(kind=O.WIT_STATION_VEL, body=int(r.integers(1, s.nb + 1)),
                          station=tuple(r.normal(size=3)), direction=tuple(r.normal(size=3))))
        wits.append(w[:8])
    B = MatterBatch(specs, tree_of=tree_of)
    for k, w in enumerate(wits):
        B.set_witnesses(k, w)
    qs, us = [], []
    for i in range(n):
        q, u = M.random_state(specs[tree_of[i]], 5, i)
        qs.append(q); us.append(u)
    ts = r.uniform(0, 1, n)
    vals = B.calcEventTriggers(B.pack(qs), B.pack(us), B.to_device(ts)).cpu().numpy()
    for i in range(n):
        o = OracleModel(specs[tree_of[i]])
        assert o.set_witnesses(wits[tree_of[i]]) == 0
        ref = o.witness_values(ts[i], np.concatenate([qs[i], us[i]]), len(wits[tree_of[i]]))
        assert _scaled_diff(vals[i, :len(ref)], ref) < 1e-12, i
        assert np.all(vals[i, len(ref):] == 0)


def test_batch_of_different_trees_matches_oracle():
    """Witnesses on a batch of different trees (padded records, lanes sorted
    by tree size): a bouncing ball, the pendulum and the spatial chain, each
    with its own witnesses; every instance's events, step counts and state
    as the oracle's."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    sl = slider()
    spec_p, wits_p, _, _, _ = _pendulum_case(n=3)
    ch = M.spatial_chain(nb=5, seed=7, types=(M.BALL, M.PIN, M.UNIVERSAL))
    ch.gravity = np.array([0.0, -G, 0.0])
    wits_c = [dict(kind=O.WIT_STATION_POS, body=5, station=(0.1, 0.2, 0.0), direction=(0, 1.0, 0), offset=-1.0),
              dict(kind=O.WIT_STATION_VEL, body=3, station=(0.0, 0.3, 0.1), direction=(0.6, 0.0, 0.8),
                   trigger=O.TRIG_FALLING),
              dict(kind=O.WIT_U, index=1, action=O.EV_SCALE_U, action_index=2, factor=0.9)]
    wits_s = [dict(kind=O.WIT_Q, index=0, trigger=O.TRIG_FALLING, action=O.EV_SCALE_U, action_index=0, factor=-0.8)]
    specs, wits = [sl, spec_p, ch], [wits_s, wits_p, wits_c]
    tree_of = np.array([0, 1, 2, 1, 0, 2, 2, 0, 1, 2, 0, 1], np.int32)
    r = np.random.default_rng(2)
    qs, us = [], []
    for i, t in enumerate(tree_of):
        if t == 0:
            qs.append(np.array([0.5 + r.uniform()])); us.append(np.array([r.uniform(-1, 2)]))
        elif t == 1:
            qs.append(np.array([r.uniform(-0.3, 0.3)])); us.append(np.array([r.uniform(-1, 1)]))
        else:
            q, u = M.random_state(ch, 11, i, qscale=0.5, uscale=0.5)
            qs.append(q); us.append(u)
    B = MatterBatch(specs, tree_of=tree_of)
    for k, w in enumerate(wits):
        B.set_witnesses(k, w)
    res = B.stepToWithEvents(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-6, max_step=0.05), 1.5)
    assert B.active_evaluator()[0] == 1
    g = dict(q=B.split_q(res["q"]), u=B.split_u(res["u"]))
    for k in ("t", "status", "stats", "windows", "ids", "counts"):
        g[k] = res[k].cpu().numpy()
    ref = []
    for i, t in enumerate(tree_of):
        ref.append(_oracle(specs[t], wits[t], [qs[i]], [us[i]], 1e-6, 1.5, max_h=0.05)[0])
    assert sum(len(x["win"]) for x in ref) >= 2 * len(tree_of)
    _compare(g, ref, wtol=1e-6 * 0.1 * 0.1)


def test_constrained_point_pendulum_matches_oracle():
    """Witnesses with constraints (the reference's integrator-test pendulum:
    a point mass held by a Rod, IntegratorTestFramework.h): trial and
    backed-up states are projected onto the constraints (projectQ then
    projectU at the tolerance), and the events, step counts and states
    follow the oracle's.  The projections' rounding differs between the
    two builds (sweeps against dense products), so windows agree to their
    required width and states to 1e-8."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    from test_oracle_events import point_pendulum
    spec, cons = point_pendulum()
    wits = [dict(kind=O.WIT_Q, index=0, action=O.EV_SCALE_U, action_index=-1, factor=1.1),
            dict(kind=O.WIT_TIME, offset=0.77),
            dict(kind=O.WIT_STATION_POS, body=1, direction=(0, 1.0, 0), offset=-0.9)]
    qs, us = [], []
    for k in range(8):
        a = -0.3 + 0.09 * k
        qs.append(np.array([np.cos(a), np.sin(a), 0.0]))
        us.append(np.array([-np.sin(a), np.cos(a), 0.0]) * (0.4 * k - 1.0))
    B = MatterBatch([spec], n_instances=8)
    B.set_constraints(0, cons)
    B.set_witnesses(0, wits)
    res = B.stepToWithEvents(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-7), 2.0)
    g = dict(q=B.split_q(res["q"]), u=B.split_u(res["u"]))
    for k in ("t", "status", "stats", "windows", "ids", "counts"):
        g[k] = res[k].cpu().numpy()
    o = OracleModel(spec)
    assert o.set_constraints(cons) == 0
    ref = []
    for i in range(8):
        assert o.set_witnesses(wits) == 0
        y, st = o.integrate(np.concatenate([qs[i], us[i]]), 2.0, 1e-7)
        win, ids, tf, term = o.last_events()
        ref.append(dict(y=y, st=st, win=win, ids=ids, t=tf, term=term))
    assert sum(len(r["win"]) for r in ref) >= 24
    for i, r in enumerate(ref):
        assert g["counts"][i] == len(r["win"]) and np.array_equal(g["ids"][i, :len(r["win"])], r["ids"]), i
        assert np.abs(g["windows"][i, :len(r["win"])] - r["win"]).max() <= 1e-7 * 0.1 * 0.1 * 2, i
        assert (r["st"]["accepted"], r["st"]["rejected"]) == tuple(g["stats"][i, 1:3]), (i, r["st"], g["stats"][i])
        assert rel_inf(np.concatenate([g["q"][i], g["u"][i]]), r["y"]) < 1e-8, i
        pe, ve, _ = o.constraint_errors(g["q"][i], g["u"][i])
        assert abs(pe[0]) < 1e-8 and abs(ve[0]) < 1e-8


def _station_wits(spec, seed, n=3):
    r = np.random.default_rng(seed)
    w = [dict(kind=O.WIT_TIME, offset=0.037)]
    for _ in range(n):
        w.append(dict(kind=O.WIT_STATION_POS, body=int(r.integers(1, spec.nb + 1)), station=tuple(r.normal(size=3)),
                      direction=tuple(r.normal(size=3)), offset=float(0.3 * r.normal())))
    return w


def _compare_runs(g, ref, n):
    for i in range(n):
        y, s, win, ids = ref[i]
        assert s["status"] == 0
        assert g["counts"][i] == len(win) and np.array_equal(g["ids"][i, :len(win)], ids), i
        assert np.abs(g["windows"][i, :len(win)] - win).max(initial=0) < 1e-8, i
        assert (s["accepted"], s["rejected"]) == tuple(g["stats"][i, 1:3]), (i, s, g["stats"][i])
        assert rel_inf(np.concatenate([g["q"][i], g["u"][i]]), y) < 1e-8, i


def _gpu_events(B, qs, us, fs, integ, T):
    res = B.stepToWithEvents(B.pack(qs), B.pack(us), integ, T, fapp=B.pack(fs))
    assert B.active_evaluator()[0] == 1
    g = dict(q=B.split_q(res["q"]), u=B.split_u(res["u"]))
    for k in ("stats", "windows", "ids", "counts"):
        g[k] = res[k].cpu().numpy()
    return g


def test_witnesses_with_prescribed_motion_match_oracle():
    """Witnesses on a chain whose root follows Motion::Sinusoid at position
    level: trial and backed-up states carry the prescribed q and u at their
    own times (createInterpolatedState's prescribeQ, prescribeU), the
    derivative at a step's end its known udot."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    from test_gpu_motion import _random_motion
    spec = M.spatial_chain(8, seed=13, types=(M.PIN, M.UNIVERSAL, M.BALL))
    n = 12
    wits = _station_wits(spec, 5)
    B = MatterBatch([spec], n_instances=n)
    B.set_witnesses(0, wits)
    qs, us, fs, ms = [], [], [], []
    for i in range(n):
        q, u = M.random_state(spec, 51, i)
        m = _random_motion(spec, 5100 + i)
        m[1] = (1, 0.4, 3.0, 0.1 * i)           # the root Pin at position level
        qs.append(q); us.append(u); fs.append(M.random_body_forces(spec, 52, i)); ms.append(m)
    B.set_motion(B.pack(ms))
    g = _gpu_events(B, qs, us, fs, RungeKuttaMerson(accuracy=1e-6), 0.6)
    o = OracleModel(spec)
    assert o.set_witnesses(wits) == 0
    ref = []
    for i in range(n):
        y, s = o.integrate_motion(np.concatenate([qs[i], us[i]]), 0.6, 1e-6, ms[i], Fapp=fs[i])
        ref.append((y, s) + o.last_events()[:2])
    assert sum(len(r[2]) for r in ref) >= 2 * n
    _compare_runs(g, ref, n)


def test_witnesses_with_known_udot_match_oracle():
    """Witnesses on rollouts with locks and constant known udot."""
    from simbody_amd.matter import RungeKuttaMerson
    from test_gpu_prescribed import _known_case
    per, B, qs, us, fs, masks, known = _known_case("shared")
    spec = per[0]
    wits = _station_wits(spec, 6)
    B.set_witnesses(0, wits)
    B.set_known_udot(B.pack_body_flags(masks), B.pack(known))
    g = _gpu_events(B, qs, us, fs, RungeKuttaMerson(accuracy=1e-6), 0.3)
    o = OracleModel(spec)
    assert o.set_witnesses(wits) == 0
    ref = []
    for i in range(len(per)):
        y, s = o.integrate_known(np.concatenate([qs[i], us[i]]), 0.3, 1e-6, masks[i], known[i], Fapp=fs[i])
        ref.append((y, s) + o.last_events()[:2])
    assert sum(len(r[2]) for r in ref) >= len(per)
    _compare_runs(g, ref, len(per))
