"""tdyn_rollout_noalloc against tdyn_rollout (GPU, through the C ABI).

The no-allocation variant runs the same kernels on the same state layout,
only with its records and integrator state in the library's fixed device
array, so its results must equal tdyn_rollout's bit for bit.  Two calls in a
row, the second with a smaller batch (the array keeps the first call's
contents, which must not leak into the second), for a chain batch (chain
evaluator, fixed step) and a batch of different trees (tree sweeps, step
control and applied forces); a batch larger than the array is refused with
TDYN_E_SIZE."""
import numpy as np
import pytest

from common import mixed_trees
from simbody_amd import models as M

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _pair(B, qs, us, integ, T, F=None):
    q0, u0 = B.pack(qs) if isinstance(qs, list) else qs, B.pack(us) if isinstance(us, list) else us
    fa = None if F is None else B.pack(F)
    q1, u1, st1 = B.stepTo(q0, u0, integ, T, fapp=fa)
    q2, u2 = q0.clone(), u0.clone()
    st2 = torch.zeros(4 * B.n_instances, dtype=torch.int32, device=q0.device)
    B.stepTo_noalloc(q2, u2, integ, T, fapp=fa, stats=st2)
    torch.cuda.synchronize()
    return (q1.cpu().numpy(), u1.cpu().numpy(), st1.cpu().numpy().reshape(-1)), \
           (q2.cpu().numpy(), u2.cpu().numpy(), st2.cpu().numpy())


@pytest.mark.parametrize("n", [300, 70])
def test_chain_batch_matches_rollout(n):
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    spec = M.spatial_chain(32, seed=2)
    Y = M.chain_batch_states(spec, n, seed=12 + n)
    B = MatterBatch([spec], n_instances=n)
    ref, got = _pair(B, B.to_device(Y[:, :spec.nq]), B.to_device(Y[:, spec.nq:]),
                     RungeKuttaMerson(fixed_step=1e-3), 0.02)
    assert B.active_evaluator()[0] == 2
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    assert (got[2].reshape(-1, 4)[:, 1] == 20).all()


@pytest.mark.parametrize("n", [40, 9])
def test_tree_batch_matches_rollout(n):
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    trees = mixed_trees(70 + n, n)
    B = MatterBatch(trees, tree_of=np.arange(n))
    B.set_evaluator(1)
    qs, us = zip(*[M.random_state(t, 4, i) for i, t in enumerate(trees)])
    F = [M.random_body_forces(t, 4, i) for i, t in enumerate(trees)]
    ref, got = _pair(B, list(qs), list(us), RungeKuttaMerson(accuracy=1e-6), 0.05, F)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    assert (got[2].reshape(-1, 4)[:, 1] > 0).all()


def test_too_large_batch_is_refused():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson, TdynError
    spec = M.spatial_chain(32, seed=2)
    n = 4096
    B = MatterBatch([spec], n_instances=n)
    Y = M.chain_batch_states(spec, n, seed=5)
    q, u = B.to_device(Y[:, :spec.nq]), B.to_device(Y[:, spec.nq:])
    with pytest.raises(TdynError) as e:
        B.stepTo_noalloc(q, u, RungeKuttaMerson(fixed_step=1e-3), 0.01)
    assert e.value.code == 2
