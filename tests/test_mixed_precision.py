"""Mixed-precision chain evaluator (integrator precision 1): FP32 kinematics,
bias forces and articulated inertias; FP64 D = H^T P H and its inverse,
quaternion normalization and Kahan-compensated state.

Stated bounds, against the FP64 oracle (relative, ||d||_inf / ||ref||_inf
per instance):
  udot, 32-body chains                                     <= 1e-5
  udot, stiff chains (neighbouring mass ratios up to 1e4)  <= 1e-3
  q, u after 50 fixed steps of h = 1e-3 vs the FP64 path   <= 2e-5
Precision 2 (FP64 arithmetic, FP32 G rows) and 3 (also the trig cache in
FP32): udot <= 1e-6 (2) and 5e-6 (3) on the chains, 1e-5 (2) and 1e-4 (3)
on the stiff chains.
The FP64 path through the same entry points stays at 1e-12 on the chains;
on the stiff chains its difference from the oracle (which does not fuse
a*b+c) grows with the conditioning of the mass matrix: it is held per
instance to max(1e-12, 0.02 cond(M) eps), the bound within which the oracle
compiled with and without fused multiply-adds agree
(tests/test_oracle_rounding.py).
"""
import numpy as np
import pytest

from common import rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

BOUND = {"chain": 1e-5, "stiff": 1e-3}
BOUND_SPLIT = {(2, "chain"): 1e-6, (3, "chain"): 5e-6, (2, "stiff"): 1e-5, (3, "stiff"): 1e-4}
BOUND64 = {"chain": 1e-12, "stiff": 1e-10}   # stiff: also per instance, cond-scaled (below)


def _spec(kind):
    return M.spatial_chain(32, seed=2) if kind == "chain" else M.stiff_chain(32, seed=2)


def _derivs(spec, Y, precision):
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    n = len(Y)
    B = MatterBatch([spec], n_instances=n)
    q = B.to_device(Y[:, :spec.nq]); u = B.to_device(Y[:, spec.nq:])
    B.rollout_begin(q, u, RungeKuttaMerson(fixed_step=1e-3, precision=precision), t_end=1.0)
    qd, ud = B.rollout_derivs()
    return qd.cpu().numpy().reshape(n, -1), ud.cpu().numpy().reshape(n, -1)


@pytest.mark.parametrize("kind", ["chain", "stiff"])
def test_udot_error_vs_oracle(kind):
    spec = _spec(kind)
    n = 64
    Y = M.chain_batch_states(spec, n, seed=31)
    qd64, ud64 = _derivs(spec, Y, 0)
    qd32, ud32 = _derivs(spec, Y, 1)
    o = OracleModel(spec)
    e64, e32 = [], []
    eps = np.finfo(np.float64).eps
    for i in range(n):
        ud, qd, _, rc = o.realize_acceleration(Y[i, :spec.nq], Y[i, spec.nq:])
        assert rc == 0
        e64.append(rel_inf(ud64[i], ud)); e32.append(rel_inf(ud32[i], ud))
        if kind == "stiff":
            q = Y[i, :spec.nq]
            Mm = np.stack([o.multiply_by_m(q, np.eye(spec.nu)[j]) for j in range(spec.nu)], 1)
            assert e64[-1] <= max(1e-12, 0.02 * np.linalg.cond(Mm) * eps), (i, e64[-1])
        assert rel_inf(qd64[i], qd) < 1e-12 and rel_inf(qd32[i], qd) < 1e-12   # qdot is FP64 on both paths
    print(f"{kind}: FP64 max {max(e64):.2e}; mixed max {max(e32):.2e} median {np.median(e32):.2e}")
    assert max(e64) < BOUND64[kind]
    assert max(e32) < BOUND[kind]


def test_mixed_rollout_tracks_fp64():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    spec = M.spatial_chain(32, seed=2)
    n = 128
    Y = M.chain_batch_states(spec, n, seed=32)
    outs = []
    for prec in (0, 1):
        B = MatterBatch([spec], n_instances=n)
        q = B.to_device(Y[:, :spec.nq]); u = B.to_device(Y[:, spec.nq:])
        B.rollout_begin(q, u, RungeKuttaMerson(fixed_step=1e-3, precision=prec), t_end=0.05)
        B.rollout_advance(1000)
        qf, uf, tf, st, stats = B.rollout_get()
        assert (st.cpu().numpy() == 1).all() and (stats.cpu().numpy()[:, 1] == 50).all()
        outs.append(np.concatenate([qf.cpu().numpy().reshape(n, -1), uf.cpu().numpy().reshape(n, -1)], 1))
    worst = max(rel_inf(outs[1][i], outs[0][i]) for i in range(n))
    print(f"mixed vs FP64 after 50 steps: {worst:.2e}")
    assert worst < 2e-5


def test_mixed_needs_the_chain_evaluator():
    from common import mixed_trees
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson, TdynError
    trees = mixed_trees(3, 4)
    B = MatterBatch(trees, tree_of=np.arange(4))
    qs, us = zip(*[M.random_state(t, 3, i) for i, t in enumerate(trees)])
    with pytest.raises(TdynError):
        B.rollout_begin(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-6, precision=1), t_end=0.1)


@pytest.mark.parametrize("kind", ["chain", "stiff"])
@pytest.mark.parametrize("prec", [2, 3])
def test_split_precision_udot(kind, prec):
    """FP64 arithmetic with single-precision hand-over rows: the udot error
    comes only from rounding G (and, precision 3, the joint trig cache) to
    FP32 between the sweeps."""
    spec = _spec(kind)
    n = 64
    Y = M.chain_batch_states(spec, n, seed=31)
    qd, ud = _derivs(spec, Y, prec)
    o = OracleModel(spec)
    e = []
    for i in range(n):
        ref, qref, _, rc = o.realize_acceleration(Y[i, :spec.nq], Y[i, spec.nq:])
        e.append(rel_inf(ud[i], ref))
        assert rel_inf(qd[i], qref) < 1e-12
    print(f"{kind} precision {prec}: udot max {max(e):.2e} median {np.median(e):.2e}")
    assert max(e) < BOUND_SPLIT[(prec, kind)]
