"""GPU parity of rollouts through report times with interpolation off
(tdyn_rollout_extend; TimeStepper::stepTo(t_k) in turn with
Integrator::setAllowInterpolation(false)) against the oracle's
integrateStops: the state at every report time within 1e-8 and the final
accepted, rejected and projection counts identical, on every evaluator
(tree sweeps, streaming, level-parallel, chain), with one set of report
times for the batch and with per-instance final times; an earlier final
time leaves a finished instance where it is.  Reports by interpolation
(tdyn_rollout_reports, Simbody's default) against the oracle's
integrateStops with reports: the plain rollout's trajectory and counts, and
each Hermite report (quaternions projected) within 1e-8."""
import numpy as np
import pytest

from common import ext_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TIMES = [0.02, 0.05, 0.1]


def _ragged(seed, n=10):
    trees = ext_trees(seed, n, nmin=4, nmax=24)
    qs, us, fs = [], [], []
    for i, t in enumerate(trees):
        q, u = M.random_state(t, seed, i); qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, seed + 1, i))
    return trees, qs, us, fs


@pytest.mark.parametrize("mode,lanes,active", [(1, 0, 1), (3, 1, 3), (3, 8, 4), (0, 0, 4)])
def test_report_times_on_tree_evaluators(mode, lanes, active):
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    trees, qs, us, fs = _ragged(421)
    n = len(trees)
    B = MatterBatch(trees, tree_of=np.arange(n, dtype=np.int32))
    B.set_evaluator(mode); B.set_tree_lanes(lanes)
    Q, U, st = B.stepToReports(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-6), TIMES, fapp=B.pack(fs))
    assert B.active_evaluator()[0] == active
    st = st.cpu().numpy()
    for i, t in enumerate(trees):
        Y, s = OracleModel(t).integrate_stops(np.concatenate([qs[i], us[i]]), TIMES, 1e-6, Fapp=fs[i])
        assert (s["accepted"], s["rejected"], s["projections"]) == tuple(st[i, 1:4]), (i, s, st[i])
        for k in range(len(TIMES)):
            y = np.concatenate([B.split_q(Q[k])[i], B.split_u(U[k])[i]])
            assert rel_inf(y, Y[k]) < 1e-8, (i, k)


@pytest.mark.parametrize("fixed_h", [0.0, 0.004])
def test_report_times_on_chain_evaluator(fixed_h):
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    spec = M.spatial_chain(12, seed=9)
    n = 128
    B = MatterBatch([spec], n_instances=n)
    Y0 = np.stack([np.concatenate(M.random_state(spec, 423, i)) for i in range(n)])
    integ = RungeKuttaMerson(accuracy=1e-6, fixed_step=fixed_h) if fixed_h else RungeKuttaMerson(accuracy=1e-6)
    Q, U, st = B.stepToReports(B.to_device(Y0[:, :spec.nq]), B.to_device(Y0[:, spec.nq:]), integ, TIMES)
    assert B.active_evaluator()[0] == 2
    st = st.cpu().numpy()
    o = OracleModel(spec)
    for i in range(0, n, 16):
        Y, s = o.integrate_stops(Y0[i], TIMES, 1e-6, fixed_h=fixed_h)
        assert (s["accepted"], s["rejected"], s["projections"]) == tuple(st[i, 1:4]), (i, s, st[i])
        for k in range(len(TIMES)):
            y = np.concatenate([Q[k].cpu().numpy().reshape(n, -1)[i], U[k].cpu().numpy().reshape(n, -1)[i]])
            assert rel_inf(y, Y[k]) < 1e-8, (i, k)


def test_per_instance_final_times_and_backward_extend():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    trees, qs, us, fs = _ragged(425, n=8)
    n = len(trees)
    B = MatterBatch(trees, tree_of=np.arange(n, dtype=np.int32))
    rng = np.random.default_rng(425)
    t1 = rng.uniform(0.01, 0.05, size=n); t2 = t1 + rng.uniform(0.01, 0.05, size=n)
    integ = RungeKuttaMerson(accuracy=1e-6)
    B.rollout_begin(B.pack(qs), B.pack(us), integ, B.to_device(t1), fapp=B.pack(fs))
    qa, ua, ta, sa, _ = B._run_to_end()
    assert np.allclose(ta.cpu().numpy(), t1, rtol=0, atol=0) and bool((sa == 1).all())
    B.rollout_extend(B.to_device(t2))
    qb, ub, tb, sb, stats = B._run_to_end()
    assert np.array_equal(tb.cpu().numpy(), t2) and bool((sb == 1).all())
    stats = stats.cpu().numpy()
    for i, t in enumerate(trees):
        Y, s = OracleModel(t).integrate_stops(np.concatenate([qs[i], us[i]]), [t1[i], t2[i]], 1e-6, Fapp=fs[i])
        assert (s["accepted"], s["rejected"]) == tuple(stats[i, 1:3]), i
        assert rel_inf(np.concatenate([B.split_q(qa)[i], B.split_u(ua)[i]]), Y[0]) < 1e-8, i
        assert rel_inf(np.concatenate([B.split_q(qb)[i], B.split_u(ub)[i]]), Y[1]) < 1e-8, i
    # a final time behind every instance leaves them finished where they are
    B.rollout_extend(0.0)
    B.rollout_advance(1 << 20)
    qc, uc, tc, sc, _ = B.rollout_get()
    assert bool((sc == 1).all()) and torch.equal(qc, qb) and torch.equal(uc, ub) and torch.equal(tc, tb)


RTIMES = [0.0, 0.013, 0.05, 0.0777, 0.1]


@pytest.mark.parametrize("mode,lanes,active", [(1, 0, 1), (3, 1, 3), (3, 8, 4)])
def test_interpolated_reports_on_tree_evaluators(mode, lanes, active):
    """tdyn_rollout_reports (interpolation on): the trajectory and counts of
    the plain rollout, and at each report time the oracle's Hermite report
    (quaternions projected) within 1e-8."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    trees, qs, us, fs = _ragged(441)
    n = len(trees)
    B = MatterBatch(trees, tree_of=np.arange(n, dtype=np.int32))
    B.set_evaluator(mode); B.set_tree_lanes(lanes)
    q1, u1, Q, U, st = B.stepToWithReports(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-4), 0.1, RTIMES,
                                           fapp=B.pack(fs))
    assert B.active_evaluator()[0] == active
    st = st.cpu().numpy()
    for i, t in enumerate(trees):
        R, yT, s = OracleModel(t).integrate_reports(np.concatenate([qs[i], us[i]]), 0.1, RTIMES, 1e-4, Fapp=fs[i])
        assert (s["accepted"], s["rejected"], s["projections"]) == tuple(st[i, 1:4]), (i, s, st[i])
        assert rel_inf(np.concatenate([B.split_q(q1)[i], B.split_u(u1)[i]]), yT) < 1e-8, i
        for k in range(len(RTIMES)):
            y = np.concatenate([B.split_q(Q[k])[i], B.split_u(U[k])[i]])
            assert rel_inf(y, R[k]) < 1e-8, (i, k)


def test_interpolated_reports_on_chain_evaluator():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    spec = M.spatial_chain(12, seed=11)
    n = 128
    B = MatterBatch([spec], n_instances=n)
    Y0 = np.stack([np.concatenate(M.random_state(spec, 443, i)) for i in range(n)])
    q1, u1, Q, U, st = B.stepToWithReports(B.to_device(Y0[:, :spec.nq]), B.to_device(Y0[:, spec.nq:]),
                                           RungeKuttaMerson(accuracy=1e-4), 0.1, RTIMES)
    assert B.active_evaluator()[0] == 2
    st = st.cpu().numpy()
    o = OracleModel(spec)
    for i in range(0, n, 16):
        R, yT, s = o.integrate_reports(Y0[i], 0.1, RTIMES, 1e-4)
        assert (s["accepted"], s["rejected"], s["projections"]) == tuple(st[i, 1:4]), (i, s, st[i])
        for k in range(len(RTIMES)):
            y = np.concatenate([Q[k].cpu().numpy().reshape(n, -1)[i], U[k].cpu().numpy().reshape(n, -1)[i]])
            assert rel_inf(y, R[k]) < 1e-8, (i, k)


def test_report_api_errors():
    """Report times that do not increase within [t0, t_end] are refused
    before anything runs; extending a rollout that was never begun is a
    stage error."""
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson, TdynError
    spec = M.spatial_chain(4, seed=3)
    B = MatterBatch([spec], n_instances=4)
    with pytest.raises(TdynError) as e:
        B.rollout_extend(0.1)
    assert e.value.args[0].startswith("TDYN_E_STAGE")
    Y0 = np.stack([np.concatenate(M.random_state(spec, 451, i)) for i in range(4)])
    q, u = B.to_device(Y0[:, :spec.nq]), B.to_device(Y0[:, spec.nq:])
    for bad in ([0.05, 0.02], [0.02, 0.2], [-0.01, 0.05], [0.03, 0.03]):
        with pytest.raises(TdynError, match="report times must increase"):
            B.stepToWithReports(q, u, RungeKuttaMerson(accuracy=1e-4), 0.1, bad)
    with pytest.raises(ValueError):
        B.stepToReports(q, u, RungeKuttaMerson(accuracy=1e-4), [0.05, 0.02])
