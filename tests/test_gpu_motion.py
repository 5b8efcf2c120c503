"""GPU parity of rollouts with prescribed motion (Motion::Sinusoid at
position, velocity and acceleration level, Motion::Steady; Motion.h:240-300,
MotionImpl.h:182-300) with the oracle, through the C ABI: every stage sees
the prescribed q, u and udot at its own time and each step's end state is
prescribed before its error estimate (IntegratorRep.h:901-934), so the
oracle's accepted and rejected step counts are matched exactly and the
states within 1e-8.  Batches of different trees mixing every mobilizer
(padded generic records) and one shared tree (typed records), each instance
with its own motions; a pendulum prescribed at position level follows its
sinusoid to rounding; the settings the motion cannot run with are refused."""
import numpy as np
import pytest

from common import ext_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

T_END = 0.1
ACC = 1e-6


def _random_motion(spec, seed, p=0.35):
    """(nb+1, 4) rows: level, a, w, phi; position level only where qdot = u."""
    r = np.random.default_rng(seed)
    m = np.zeros((spec.nb + 1, 4))
    for b in range(1, spec.nb + 1):
        if r.uniform() >= p:
            continue
        ty = int(spec.mobilizer[b - 1])
        levels = [2, 3, 4] if ty in (M.BALL, M.FREE) else [1, 2, 3, 4]
        lv = int(r.choice(levels))
        amp = {1: 0.5, 2: 1.0, 3: 2.0, 4: 0.6}[lv] * r.uniform(0.3, 1.0)
        m[b] = (lv, amp, r.uniform(1.0, 4.0), r.uniform(-np.pi, np.pi))
    return m


def _case(kind):
    from simbody_amd.matter import MatterBatch
    if kind == "ragged":
        per = ext_trees(311, 12, nmin=2, nmax=20)
        B = MatterBatch(per, tree_of=np.arange(len(per), dtype=np.int32))
    else:
        spec = M.spatial_chain(16, seed=13)
        per = [spec] * 32
        B = MatterBatch([spec], n_instances=32)
    qs, us, fs, ms = [], [], [], []
    for i, t in enumerate(per):
        q, u = M.random_state(t, 311, i)
        qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, 312, i))
        ms.append(_random_motion(t, 31100 + i))
    return per, B, qs, us, fs, ms


@pytest.mark.parametrize("kind", ["ragged", "shared"])
def test_rollouts_with_motion_match_oracle(kind):
    from simbody_amd.matter import RungeKuttaMerson
    per, B, qs, us, fs, ms = _case(kind)
    assert sum(int((m[:, 0] > 0).sum()) for m in ms) > 0
    mot = B.pack(ms)
    B.set_motion(mot)
    rkm = RungeKuttaMerson(accuracy=ACC)
    ref = [OracleModel(t).integrate_motion(np.concatenate([qs[i], us[i]]), T_END, ACC, ms[i], Fapp=fs[i])
           for i, t in enumerate(per)]
    for mode in (0, 1):
        B.set_evaluator(mode)
        q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), rkm, T_END, fapp=B.pack(fs))
        assert B.active_evaluator()[0] == 1, mode
        q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
        for i, t in enumerate(per):
            y, s = ref[i]
            assert s["status"] == 0
            assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), (mode, i, s, st[i])
            assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, (mode, i)


def test_motion_cleared_returns_to_free_dynamics():
    """set_motion(None): the next rollout is the unprescribed one."""
    from simbody_amd.matter import RungeKuttaMerson
    per, B, qs, us, fs, ms = _case("ragged")
    B.set_motion(B.pack(ms))
    rkm = RungeKuttaMerson(accuracy=ACC)
    B.stepTo(B.pack(qs), B.pack(us), rkm, 0.02, fapp=B.pack(fs))
    B.set_motion(None)
    B.set_evaluator(1)
    q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), rkm, T_END, fapp=B.pack(fs))
    q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
    for i, t in enumerate(per):
        y, s = OracleModel(t).integrate(np.concatenate([qs[i], us[i]]), T_END, ACC, Fapp=fs[i])
        assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), i
        assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, i


def test_position_sinusoid_is_followed_on_the_gpu():
    """A Pin prescribed at position level ends at a sin(w T + phi) with rate
    a w cos(w T + phi), whatever gravity and its initial state say; the
    unprescribed lower link of a double pendulum follows the oracle."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    a, w, ph, T = 0.7, 3.0, 0.4, 0.9
    spec = M.pendulum_chain(2)
    B = MatterBatch([spec], n_instances=3)
    B.set_evaluator(1)
    m = np.zeros((3, 4)); m[1] = (1, a, w, ph)
    B.set_motion(B.pack([m] * 3))
    qs = [np.array([0.1 * k, 0.6]) for k in range(3)]
    us = [np.array([-0.5, -0.4 + 0.1 * k]) for k in range(3)]
    rkm = RungeKuttaMerson(accuracy=1e-8)
    q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), rkm, T)
    q1s, u1s = B.split_q(q1), B.split_u(u1)
    for k in range(3):
        assert q1s[k][0] == pytest.approx(a * np.sin(w * T + ph), abs=1e-14)
        assert u1s[k][0] == pytest.approx(a * w * np.cos(w * T + ph), abs=1e-13)
        y, s = OracleModel(spec).integrate_motion(np.concatenate([qs[k], us[k]]), T, 1e-8, m)
        assert rel_inf(np.concatenate([q1s[k], u1s[k]]), y) < 1e-8, k


def test_settings_motion_cannot_run_with_are_refused():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson, TdynError
    rkm = RungeKuttaMerson(accuracy=ACC)
    spec = M.spatial_chain(8, seed=3)
    B = MatterBatch([spec], n_instances=4)
    q, u = M.random_state(spec, 5, 0)
    m = np.zeros((spec.nb + 1, 4)); m[2] = (2, 1.0, 2.0, 0.0)
    B.set_motion(B.pack([m] * 4))
    B.set_evaluator(2)                      # the chain evaluator carries no motion
    with pytest.raises(TdynError):
        B.stepTo(B.pack([q] * 4), B.pack([u] * 4), rkm, 0.01)
    B.set_evaluator(1)
    with pytest.raises(TdynError):          # FP64 only
        B.stepTo(B.pack([q] * 4), B.pack([u] * 4), RungeKuttaMerson(accuracy=ACC, precision=1), 0.01)
    with pytest.raises(TdynError):          # not beside known udot
        B.set_known_udot(B.pack_body_flags([np.zeros(spec.nb + 1)] * 4), B.pack([np.zeros(spec.nu)] * 4))
        B.stepTo(B.pack([q] * 4), B.pack([u] * 4), rkm, 0.01)
    B.set_known_udot(None, None)
    with pytest.raises(TdynError):          # nor in the rollout without allocation
        B.stepTo_noalloc(B.pack([q] * 4), B.pack([u] * 4), rkm, 0.01)
    B.stepTo(B.pack([q] * 4), B.pack([u] * 4), rkm, 0.01)
    # position level needs qdot = u: refused on a Ball, a level out of range
    # or a value that is not finite refused too
    trees = ext_trees(5, 3, nmin=4, nmax=6)
    trees[0].mobilizer[1] = M.BALL
    C = MatterBatch(trees, tree_of=np.arange(3, dtype=np.int32))
    for bad in ((2, (1, 0.5, 1.0, 0.0)), (1, (5, 0.5, 1.0, 0.0)), (1, (2, np.nan, 1.0, 0.0)), (1, (2.5, 1, 1, 0))):
        ms = [np.zeros((t.nb + 1, 4)) for t in trees]
        ms[0][bad[0]] = bad[1]
        with pytest.raises(TdynError):
            C.set_motion(C.pack(ms))
