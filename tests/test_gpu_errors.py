"""Error semantics through the C ABI (GPU).

- A singular articulated inertia D (SmallMatrixMixed.h:878-892, where the
  reference throws): a Free body with no rotational inertia, its mass center
  and joint point at its origin, gives D with a zero pivot.  The instance's outputs are
  inf/NaN (no fault), tdyn_batch_op_status reports TDYN_E_SINGULAR for it and
  0 for the others, whose results still match the oracle at 1e-12; the oracle
  flags the same instance.
- A refused tdyn_rollout_begin leaves the rollout in progress untouched, and
  settings changed after a begin (known udot, the evaluator) wait for the
  next begin.
- The serving loop's admit refuses reversed and Euler-angle requests before
  taking a slot; the autotuner's cache is keyed on known udot."""
import numpy as np
import pytest

from common import mixed_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _degenerate_free():
    """One Free body with no rotational inertia about its origin, which is
    also its mass center and its joint point: D = diag(0, 0, 0, m, m, m) in
    the mobilizer's frame, exactly singular (the reference's LAPACK LU stops
    at a zero pivot, and so does the oracle's)."""
    t = M.random_tree(7, 0, nmin=1, nmax=1, general_frames=True)
    t.mobilizer[0] = M.FREE
    t.inertia_B[0] = 0.0
    t.com_B[0] = 0.0
    t.X_BM[0, 9:] = 0.0
    return t


def test_singular_articulated_inertia_is_reported_per_instance():
    from simbody_amd.matter import MatterBatch, TdynError
    trees = mixed_trees(61, 2) + [_degenerate_free()] + mixed_trees(62, 2)
    n = len(trees)
    B = MatterBatch(trees, tree_of=np.arange(n))
    qs, us = zip(*[M.random_state(t, 3, i) for i, t in enumerate(trees)])
    q, u = B.pack(list(qs)), B.pack(list(us))
    for op in ("accel", "minv", "abi"):
        if op == "accel":
            out, _ = B.realizeAcceleration(q, u)
        elif op == "minv":
            out = B.multiplyByMInv(q, u)
        else:
            B.realizeArticulatedBodyInertias(q)
            out = None
        st = B.op_status().cpu().numpy()
        assert list(st) == [0, 0, 4, 0, 0], (op, st)
        with pytest.raises(TdynError) as e:
            B.check_singular()
        assert e.value.code == 4
        if out is not None:
            parts = B.split_u(out)
            assert not np.all(np.isfinite(parts[2]))
            for i in (0, 1, 3, 4):
                o = OracleModel(trees[i])
                ref = (o.realize_acceleration(qs[i], us[i])[0] if op == "accel" else o.minv(qs[i], us[i]))
                assert rel_inf(parts[i], ref) < 1e-12
    _, _, _, rc = OracleModel(trees[2]).realize_acceleration(qs[2], us[2])
    assert rc != 0
    # a batch without the degenerate tree clears the status
    B2 = MatterBatch(trees[:2], tree_of=np.arange(2))
    B2.realizeAcceleration(B2.pack(list(qs[:2])), B2.pack(list(us[:2])))
    assert list(B2.op_status().cpu().numpy()) == [0, 0]
    B2.check_singular()


def test_refused_begin_leaves_the_rollout_in_progress():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson, TdynError
    trees = mixed_trees(63, 6)
    n = len(trees)
    qs, us = zip(*[M.random_state(t, 5, i) for i, t in enumerate(trees)])
    integ = RungeKuttaMerson(accuracy=1e-6)
    finals = []
    for interrupt in (False, True):
        B = MatterBatch(trees, tree_of=np.arange(n))
        q, u = B.pack(list(qs)), B.pack(list(us))
        B.rollout_begin(q, u, integ, t_end=0.3)
        B.rollout_advance(3)
        if interrupt:
            with pytest.raises(TdynError):
                B.rollout_begin(q, u, RungeKuttaMerson(accuracy=1e-3, precision=1), t_end=0.1)
            known = torch.ones(B.total_b, dtype=torch.int32, device=q.device)
            B.set_known_udot(known, torch.zeros(B.total_u, dtype=torch.float64, device=q.device))
            B.set_evaluator(1)
        B.rollout_advance(1 << 20)
        qT, uT, _, st, stats = B.rollout_get()
        finals.append((qT.cpu().numpy(), uT.cpu().numpy(), stats.cpu().numpy()))
        del B
    assert np.array_equal(finals[0][0], finals[1][0]) and np.array_equal(finals[0][1], finals[1][1])
    assert np.array_equal(finals[0][2], finals[1][2])


def test_server_admit_refuses_reversed_and_euler_requests():
    from simbody_amd.matter import RungeKuttaMerson, TdynError
    from simbody_amd.serving import Server
    template = M.spatial_chain(6, seed=2)
    sv = Server(template, RungeKuttaMerson(accuracy=1e-6), 1 << 26, max_slots=64)
    free0 = sv.info()["free"]
    Y = M.random_states(template, 3, np.arange(2))
    for tweak in ("reversed", "euler"):
        t = M.spatial_chain(6, seed=2)
        if tweak == "reversed":
            t.reversed = np.zeros(t.nb, np.int32); t.reversed[1] = 1
        else:
            t.use_euler_angles = True
        with pytest.raises(TdynError):
            sv.admit([template, t], Y[:, :t.nq], Y[:, t.nq:], np.array([0.1, 0.1]))
        assert sv.info()["free"] == free0
    assert len(sv.admit([template], Y[:1, :template.nq], Y[:1, template.nq:], np.array([0.1]))) == 1
