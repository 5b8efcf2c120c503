"""Edge cases of the HIP path against the oracle (FP64, 1e-12 relative on
udot, the bar of test_gpu_parity): one-body trees of every mobilizer, the
largest trees of the workloads (128 bodies, depth up to 20), instance counts
that do not fill a wave, a batch mixing chain-shaped and branched trees, and
rollouts of one-body instances through both evaluators."""
import numpy as np
import pytest

from common import rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


def _one_body(mob, seed):
    t = M.random_tree(seed, 0, nmin=1, nmax=1, general_frames=True)
    t.mobilizer[0] = mob
    return t


def _check_udot(trees, seed, forces=True):
    from simbody_amd.matter import MatterBatch
    n = len(trees)
    B = MatterBatch(trees, tree_of=np.arange(n)) if n > 1 else MatterBatch(trees, n_instances=1)
    qs, us, fs = [], [], []
    for i, t in enumerate(trees):
        q, u = M.random_state(t, seed, i)
        qs.append(q); us.append(u); fs.append(M.random_body_forces(t, seed, i))
    ud, _ = B.realizeAcceleration(B.pack(qs), B.pack(us), B.pack(fs) if forces else None)
    uds = B.split_u(ud)
    worst = 0.0
    for i, t in enumerate(trees):
        o, _, _, rc = OracleModel(t).realize_acceleration(qs[i], us[i], fs[i] if forces else None)
        e = rel_inf(uds[i], o)
        worst = max(worst, e)
        assert e < TOL, (i, t.nb, e)
    return worst


@pytest.mark.parametrize("mob", [M.PIN, M.SLIDER, M.UNIVERSAL, M.BALL, M.FREE])
def test_one_body_each_mobilizer_alone(mob):
    _check_udot([_one_body(mob, 50 + mob)], 3)


def test_one_body_trees_mixed_batch():
    trees = [_one_body(m, 60 + k) for k, m in enumerate([M.PIN, M.SLIDER, M.UNIVERSAL, M.BALL, M.FREE] * 3)]
    _check_udot(trees, 4)


def test_largest_trees_and_ragged_wave():
    """67 instances (a full wave plus 3 lanes) of 8..128-body trees, depth
    up to 20, including the 128-body extremes and Free roots."""
    trees = [M.random_tree(71, i, nmin=100, nmax=128) for i in range(67)]
    trees[0] = M.random_tree(72, 0, nmin=128, nmax=128)
    trees[1] = M.random_tree(72, 1, nmin=128, nmax=128, max_depth=20)
    for k in range(0, 67, 7):
        trees[k].mobilizer[0] = M.FREE
    assert max(t.nb for t in trees) == 128 and max(int(t.depth().max()) for t in trees) <= 20
    _check_udot(trees, 5)


def test_long_chain_128_bodies():
    spec = M.spatial_chain(128, seed=9)
    _check_udot([spec], 6)


def test_chain_and_branched_trees_in_one_batch():
    trees = [M.spatial_chain(20, seed=3)] + [M.random_tree(73, i, nmin=8, nmax=40) for i in range(9)]
    _check_udot(trees, 7)


@pytest.mark.parametrize("mob", [M.PIN, M.BALL, M.FREE])
def test_one_body_rollouts_match_oracle(mob):
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    t = _one_body(mob, 80 + mob)
    n = 5
    B = MatterBatch([t], n_instances=n)
    Y = np.stack([np.concatenate(M.random_state(t, 9, i)) for i in range(n)])
    q1, u1, st = B.stepTo(B.to_device(Y[:, :t.nq]), B.to_device(Y[:, t.nq:]), RungeKuttaMerson(accuracy=1e-7), 0.3)
    q1 = q1.cpu().numpy().reshape(n, -1); u1 = u1.cpu().numpy().reshape(n, -1); st = st.cpu().numpy()
    o = OracleModel(t)
    for i in range(n):
        y, s = o.integrate(Y[i], 0.3, 1e-7)
        assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), (i, s, st[i])
        assert rel_inf(np.concatenate([q1[i], u1[i]]), y) < 1e-8, i
