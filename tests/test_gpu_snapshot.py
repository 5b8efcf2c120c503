"""Rollout snapshots (checkpoint / resume) through the C ABI: a rollout
stopped after some steps, snapshotted to host bytes (the layout of
tdyn_rollout_snapshot), restored into a fresh batch of the same model and
run to the end gives the uninterrupted rollout's states, times, statuses and
statistics bit for bit -- on the chain evaluator (FP64 and mixed precision),
the level-parallel evaluator over different trees (lanes sorted by tree
size), the tree sweeps with constraints, and with event witnesses; a
snapshot of another model, of other sizes or of another format is
refused."""
import numpy as np
import pytest

from simbody_amd import models as M
import oracle_ctypes as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _run(make, qs, us, integ, T, split, fapp=None):
    """(uninterrupted result, result resumed from a snapshot after `split`
    accepted steps in a new batch)."""
    A = make()
    q, u = A.pack(qs), A.pack(us)
    F = None if fapp is None else A.pack(fapp)
    A.rollout_begin(q, u, integ, T, fapp=F)
    ref = A._run_to_end()
    B = make()
    B.rollout_begin(B.pack(qs), B.pack(us), integ, T, fapp=F)
    B.rollout_advance(split)
    snap = B.rollout_snapshot()
    del B
    C = make()
    C.rollout_restore(snap, fapp=F)
    out = C._run_to_end()
    return [x.cpu().numpy() for x in ref], [x.cpu().numpy() for x in out], snap, C


def _same(ref, out):
    for a, b in zip(ref, out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [0, 1])
def test_chain_resume_is_bitwise(precision):
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    spec = M.spatial_chain(nb=12, seed=3)
    n = 96
    qs = [M.random_state(spec, 4, i)[0] for i in range(n)]
    us = [M.random_state(spec, 4, i)[1] for i in range(n)]
    integ = RungeKuttaMerson(accuracy=1e-5, precision=precision)
    ref, out, snap, C = _run(lambda: MatterBatch([spec], n_instances=n), qs, us, integ, 0.3, 7)
    assert C.active_evaluator()[0] == 2
    assert np.all(ref[3] == 1) and np.all(ref[4][:, 1] > 7)
    _same(ref, out)
    assert snap[:8] == b"TDYNSNP1"
    ny = sum(len(x) for x in qs) + sum(len(x) for x in us)
    assert len(snap) == 112 + n * (3 * 8 + 5 * 4) + 8 * ny * (2 if precision == 1 else 1)


def test_different_trees_resume_is_bitwise():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    specs = [M.random_tree(31, k, nmin=6, nmax=30) for k in range(5)]
    tree_of = np.array([k % 5 for k in range(40)], np.int32)
    qs = [M.random_state(specs[t], 6, i)[0] for i, t in enumerate(tree_of)]
    us = [M.random_state(specs[t], 6, i)[1] for i, t in enumerate(tree_of)]
    fapp = [M.random_body_forces(specs[t], 6, i) for i, t in enumerate(tree_of)]
    ref, out, _, C = _run(lambda: MatterBatch(specs, tree_of=tree_of), qs, us, RungeKuttaMerson(accuracy=1e-5),
                          0.2, 5, fapp=fapp)
    assert C.active_evaluator()[0] != 2          # a tree evaluator (level-parallel here)
    _same(ref, out)


def test_constrained_and_witness_rollouts_resume_is_bitwise():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    from test_gpu_constraints import _spatial_loop, _loop_state, _tip
    from oracle_ctypes import OracleModel
    spec = _spatial_loop()
    q0, u0 = _loop_state()
    cons = [(0, 0, 2, tuple(_tip(OracleModel(spec), q0)), (0, 1.0, 0), 0.0)]
    r = np.random.default_rng(3)
    qs = [q0] * 16
    us = [u0 + 0.3 * r.normal(size=6) for _ in range(16)]

    def make():
        B = MatterBatch([spec], n_instances=16)
        B.set_constraints(0, cons)
        return B
    ref, out, _, _ = _run(make, qs, us, RungeKuttaMerson(accuracy=1e-6), 0.4, 6)
    _same(ref, out)

    from test_oracle_events import pendulum
    pspec, _ = pendulum()
    wits = [dict(kind=O.WIT_STATION_POS, body=1, station=(1.0, 0, 0), direction=(1.0, 0, 0),
                 action=O.EV_SCALE_U, action_index=-1, factor=1.05),
            dict(kind=O.WIT_U, index=0)]

    def makew():
        B = MatterBatch([pspec], n_instances=8)
        B.set_witnesses(0, wits)
        return B
    qs = [np.array([x]) for x in np.linspace(-0.3, 0.3, 8)]
    us = [np.array([x]) for x in np.linspace(-1.0, 1.0, 8)]
    ref, out, _, C = _run(makew, qs, us, RungeKuttaMerson(accuracy=1e-7), 3.0, 40)
    _same(ref, out)
    _, _, cnt = C.rollout_events()
    assert np.all(cnt.cpu().numpy() > 0)


def test_foreign_snapshots_are_refused():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson, TdynError
    spec = M.spatial_chain(nb=6, seed=3)
    B = MatterBatch([spec], n_instances=8)
    qs = [M.random_state(spec, 4, i)[0] for i in range(8)]
    us = [M.random_state(spec, 4, i)[1] for i in range(8)]
    B.rollout_begin(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-5), 0.1)
    B.rollout_advance(2)
    snap = B.rollout_snapshot()
    other = M.spatial_chain(nb=6, seed=4)                    # same sizes, other constants
    with pytest.raises(TdynError):
        MatterBatch([other], n_instances=8).rollout_restore(snap)
    with pytest.raises(TdynError):
        MatterBatch([spec], n_instances=9).rollout_restore(snap)
    with pytest.raises(TdynError):
        MatterBatch([spec], n_instances=8).rollout_restore(b"XXXXXXXX" + snap[8:])
    with pytest.raises(TdynError):
        MatterBatch([spec], n_instances=8).rollout_restore(snap[:-8])
    MatterBatch([spec], n_instances=8).rollout_restore(snap)
