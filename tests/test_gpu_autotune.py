"""The tree evaluators' autotuner (tdyn_batch_autotune): it probes the
candidate lane settings on the batch's own state, applies one that the
batch accepts, leaves the caller's state untouched, records the choice in
its cache file and reuses it for a batch of the same shape; a rollout after
tuning still matches the oracle."""
import numpy as np
import pytest

from common import mixed_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_autotune_picks_caches_and_keeps_results(tmp_path):
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    trees = mixed_trees(211, 24, nmin=4, nmax=40)
    n = len(trees)
    qs, us = zip(*[M.random_state(t, 211, i) for i, t in enumerate(trees)])
    B = MatterBatch(list(trees), tree_of=np.arange(n, dtype=np.int32))
    q, u = B.pack(qs), B.pack(us)
    q0, u0 = q.clone(), u.clone()
    cache = str(tmp_path / "tune.txt")
    integ = RungeKuttaMerson(accuracy=1e-6)
    ch = B.autotune(q, u, integ, 0.01, cache=cache)
    assert ch[0] in (1, 3) and ch[4] == 0
    assert torch.equal(q, q0) and torch.equal(u, u0)
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and lines[0].startswith("tree-lanes ")
    B2 = MatterBatch(list(trees), tree_of=np.arange(n, dtype=np.int32))
    ch2 = B2.autotune(q, u, integ, 0.01, cache=cache)
    assert ch2[:4] == ch[:4] and ch2[4] == 1
    q1, u1, st = B.stepTo(q, u, integ, 0.05)
    q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
    for i in range(0, n, 4):
        y, s = OracleModel(trees[i]).integrate(np.concatenate([qs[i], us[i]]), 0.05, 1e-6)
        assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), i
        assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, i


def test_autotune_on_one_model_has_nothing_to_choose():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    spec = M.spatial_chain(8, seed=3)
    B = MatterBatch([spec], n_instances=64)
    Y = np.stack([np.concatenate(M.random_state(spec, 5, i)) for i in range(64)])
    ch = B.autotune(B.to_device(Y[:, :spec.nq]), B.to_device(Y[:, spec.nq:]), RungeKuttaMerson(accuracy=1e-6), 0.01)
    assert ch[0] == 0 and ch[4] == 0
