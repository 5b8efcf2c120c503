"""Rollouts of batches of different trees run on the tree evaluators for
such batches: the streaming evaluator (one lane walks its tree
depth-first) and the level-parallel evaluator (4 to 64 lanes share one
instance, one tree level at a time); instance-contiguous state.  Held to
the parity bar of test_gpu_parity: the integrator's
derivative at the rollout's initial state within 1e-12 relative of the
oracle's udot and qdot (FP64), and whole rollouts with the oracle's
accept/reject counts.  Shapes: the largest trees of the workloads (128
bodies, depth 20), several roots on Ground, a Free root, wide and deep
branching (binary trees), one-body trees, and applied forces.  A batch whose
Free body is not body 1 falls back to the tree sweeps and is checked the same
way."""
import numpy as np
import pytest

from common import mixed_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


def _binary_tree(levels, seed):
    """Complete binary tree of 2**levels - 1 bodies, Ball and Pin joints."""
    nb = 2 ** levels - 1
    t = M.random_tree(seed, 0, nmin=nb, nmax=nb, general_frames=True)
    t.parent[:] = [b // 2 for b in range(1, nb + 1)]        # body b's parent is b // 2 (body 1 on Ground)
    t.mobilizer[:] = [M.BALL if b % 3 == 0 else M.PIN if b % 3 == 1 else M.UNIVERSAL for b in range(1, nb + 1)]
    return t


def _forest(seed, nb, roots):
    """Random tree whose bodies in `roots` are re-attached to Ground."""
    t = M.random_tree(seed, 1, nmin=nb, nmax=nb, general_frames=True)
    for b in roots:
        t.parent[b - 1] = 0
    return t


def _states(trees, seed, forces):
    qs, us, fs = [], [], []
    for i, t in enumerate(trees):
        q, u = M.random_state(t, seed, i)
        qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, seed, i) if forces else None)
    return qs, us, fs


# (evaluator mode, lanes per instance): 0 automatic (level-parallel, lanes
# per instance from the tree) or 3 level-parallel; lanes 1: streaming
LANES = [(0, 0), (0, 1), (3, 4), (3, 16), (3, 64)]
KIND = {0: 4, 3: 4}


def _check_derivs(trees, seed, forces=True, expect_stream=True, lanes=(0, 0)):
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    n = len(trees)
    mode, lanes = lanes
    B = MatterBatch(trees, tree_of=np.arange(n, dtype=np.int32))
    B.set_evaluator(mode)
    B.set_tree_lanes(lanes)
    qs, us, fs = _states(trees, seed, forces)
    B.rollout_begin(B.pack(qs), B.pack(us), RungeKuttaMerson(fixed_step=1e-3), t_end=1.0,
                    fapp=B.pack(fs) if forces else None)
    assert B.uses_tree_stream() == expect_stream
    if expect_stream:
        assert B.active_evaluator() == ((3, 1) if lanes == 1 else (KIND[mode], lanes))
    q0, u0 = B.rollout_get()[:2]             # quaternions normalized by the integrator's initialize()
    qd, ud = B.rollout_derivs()
    q0s, u0s, qds, uds = B.split_q(q0), B.split_u(u0), B.split_q(qd), B.split_u(ud)
    worst = 0.0
    for i, t in enumerate(trees):
        o_ud, o_qd, _, rc = OracleModel(t).realize_acceleration(q0s[i], u0s[i], fs[i])
        assert rc == 0
        e = rel_inf(uds[i], o_ud)
        worst = max(worst, e)
        assert e < TOL, (i, t.nb, e)
        assert rel_inf(qds[i], o_qd) < TOL, i
    return worst


@pytest.mark.parametrize("lanes", LANES)
def test_largest_trees_ragged_wave(lanes):
    """67 trees of 100..128 bodies (a wave plus 3 lanes), depth up to 20, Free
    roots, all five mobilizers, general joint frames."""
    trees = [M.random_tree(91, i, nmin=100, nmax=128, general_frames=True) for i in range(67)]
    trees[0] = M.random_tree(92, 0, nmin=128, nmax=128, general_frames=True)
    trees[1] = M.random_tree(92, 1, nmin=128, nmax=128, max_depth=20, general_frames=True)
    for k in range(0, 67, 5):
        trees[k].mobilizer[0] = M.FREE
    _check_derivs(trees, 3, lanes=lanes)


@pytest.mark.parametrize("lanes", LANES)
def test_mixed_small_trees_and_one_body_trees(lanes):
    trees = mixed_trees(17, 40, nmin=1, nmax=24)
    for k, m in enumerate([M.PIN, M.SLIDER, M.UNIVERSAL, M.BALL, M.FREE]):
        trees[k] = M.random_tree(18, k, nmin=1, nmax=1, general_frames=True)
        trees[k].mobilizer[0] = m
    _check_derivs(trees, 4, lanes=lanes)


@pytest.mark.parametrize("lanes", LANES)
def test_binary_trees_and_forests(lanes):
    """Bodies with two children at every depth (the accumulators of every
    level in use), several roots on Ground, a chain, without applied forces."""
    trees = [_binary_tree(7, 5), _binary_tree(5, 6), _binary_tree(1, 7),
             _forest(8, 40, [5, 9, 23]), _forest(9, 64, [2, 3, 4, 40]),
             M.spatial_chain(20, seed=3)]
    trees[3].mobilizer[0] = M.FREE
    _check_derivs(trees, 5, forces=False, lanes=lanes)
    _check_derivs(trees, 6, forces=True, lanes=lanes)


def test_free_body_not_first_falls_back():
    """A second Free root (body 5 on Ground): the batch runs on the tree sweeps."""
    t = _forest(10, 30, [5])
    t.mobilizer[0] = M.FREE
    t.mobilizer[4] = M.FREE
    trees = [t] + mixed_trees(11, 5)
    _check_derivs(trees, 7, expect_stream=False)


def test_tree_evaluators_rollouts_agree_and_match_oracle():
    """Adaptive rollouts (accuracy 1e-6, applied forces): the streaming and
    level-parallel evaluators (the same lanes for every instance, and lanes
    chosen per instance) and the tree sweeps take the same steps; all match
    the oracle's trajectory."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    trees = mixed_trees(41, 10, nmin=8, nmax=40) + [_binary_tree(4, 12), _forest(13, 20, [7, 11])]
    n = len(trees)
    qs, us, fs = _states(trees, 41, True)
    outs = []
    for mode, lanes, rng in ((1, 0, None), (0, 1, None), (3, 4, None), (3, 0, (8, 64)), (0, 0, (4, 8)),
                             (3, 64, None), (0, 0, None)):
        B = MatterBatch(trees, tree_of=np.arange(n, dtype=np.int32))
        B.set_evaluator(mode)
        B.set_tree_lanes(lanes)
        if rng:
            B.set_tree_lane_range(*rng)
        q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-6), 0.1, fapp=B.pack(fs))
        outs.append((B.split_q(q1), B.split_u(u1), st.cpu().numpy()))
    for o in outs[1:]:
        assert (o[2] == outs[0][2]).all()
    for i, t in enumerate(trees):
        y, s = OracleModel(t).integrate(np.concatenate([qs[i], us[i]]), 0.1, 1e-6, Fapp=fs[i])
        assert (s["accepted"], s["rejected"]) == tuple(outs[0][2][i, 1:3]), (i, s)
        y_w = np.concatenate([outs[0][0][i], outs[0][1][i]])
        for o in outs[1:]:
            y_e = np.concatenate([o[0][i], o[1][i]])
            assert rel_inf(y_e, y_w) < 1e-9, i
            assert rel_inf(y_e, y) < 1e-8, i


@pytest.mark.parametrize("lanes", [(0, 1), (3, 16), (0, 0)])
def test_identity_frame_trees(lanes):
    """Trees of the uneven-trees workload (identity joint frames, so the
    level-parallel evaluator skips the frame products): derivative and an
    adaptive rollout against the oracle."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    trees = [M.random_tree(95, i, nmin=8, nmax=128) for i in range(70)]
    for k in range(0, 70, 9):
        trees[k].mobilizer[0] = M.FREE
    _check_derivs(trees, 8, lanes=lanes)
    sub = trees[:8]
    qs, us, fs = _states(sub, 9, True)
    B = MatterBatch(sub, tree_of=np.arange(len(sub), dtype=np.int32))
    B.set_evaluator(lanes[0])
    B.set_tree_lanes(lanes[1])
    q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-6), 0.05, fapp=B.pack(fs))
    q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
    for i, t in enumerate(sub):
        y, s = OracleModel(t).integrate(np.concatenate([qs[i], us[i]]), 0.05, 1e-6, Fapp=fs[i])
        assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), (i, s, st[i])
        assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, i
