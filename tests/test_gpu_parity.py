"""Parity of the HIP path (libtdyn.so via the C ABI) with the oracle.

FP64 tolerance: 1e-12 relative (||d||_inf / ||ref||_inf per
instance) on udot and residual forces; the other operator outputs are held to
the same bound.  Trajectories: 1e-9 relative after short horizons (the GPU
contracts a*b+c into FMA, the oracle does not, so single-step results differ
at rounding level and grow along a trajectory), identical accept/reject counts.
"""
import numpy as np
import pytest

from common import mixed_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


def _batch(trees, tree_of=None, n=None):
    from simbody_amd.matter import MatterBatch
    return MatterBatch(trees, tree_of=tree_of, n_instances=n)


def _ragged_case(seed=5, n=12):
    trees = mixed_trees(seed, n)
    tree_of = np.arange(n, dtype=np.int32)
    B = _batch(trees, tree_of)
    qs, us, fs, fm = [], [], [], []
    for i, t in enumerate(trees):
        q, u = M.random_state(t, seed, i)
        qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, seed, i))
        fm.append(np.random.default_rng(i).normal(size=t.nu))
    return trees, B, qs, us, fs, fm


def test_kinematics_energy():
    trees, B, qs, us, fs, fm = _ragged_case()
    X, V, qd, E = B.realizeKinematics(B.pack(qs), B.pack(us))
    Xs, Vs, qds = B.split_b(X, 12), B.split_b(V, 6), B.split_q(qd)
    E = E.cpu().numpy().reshape(-1, 2)
    for i, t in enumerate(trees):
        o = OracleModel(t).kinematics(qs[i], us[i])
        assert rel_inf(Xs[i], o["X_GB"]) < TOL, i
        assert rel_inf(Vs[i], o["V_GB"]) < TOL, i
        assert rel_inf(qds[i], o["qdot"]) < TOL, i
        assert abs(E[i, 0] - o["ke"]) <= TOL * max(abs(o["ke"]), 1.0)
        assert abs(E[i, 1] - o["pe"]) <= TOL * max(abs(o["pe"]), 1.0)


def test_realize_acceleration_with_applied_forces():
    trees, B, qs, us, fs, fm = _ragged_case(seed=6)
    ud, qd = B.realizeAcceleration(B.pack(qs), B.pack(us), B.pack(fs))
    uds = B.split_u(ud)
    worst = 0.0
    for i, t in enumerate(trees):
        o_ud, o_qd, _, rc = OracleModel(t).realize_acceleration(qs[i], us[i], fs[i])
        e = rel_inf(uds[i], o_ud)
        worst = max(worst, e)
        assert e < TOL, (i, e)
    print("worst udot rel err", worst)


def test_acceleration_ignoring_constraints():
    trees, B, qs, us, fs, fm = _ragged_case(seed=7)
    ud, A, qdd = B.calcAccelerationIgnoringConstraints(B.pack(qs), B.pack(us), B.pack(fm), B.pack(fs))
    uds, As, qdds = B.split_u(ud), B.split_b(A, 6), B.split_q(qdd)
    for i, t in enumerate(trees):
        o_ud, o_qdd, o_A, rc = OracleModel(t).forward(qs[i], us[i], fm[i], fs[i])
        assert rel_inf(uds[i], o_ud) < TOL, i
        assert rel_inf(As[i], o_A) < TOL, i
        assert rel_inf(qdds[i], o_qdd) < TOL, i


def test_multiply_by_m_inv():
    trees, B, qs, us, fs, fm = _ragged_case(seed=8)
    out = B.split_u(B.multiplyByMInv(B.pack(qs), B.pack(fm)))
    for i, t in enumerate(trees):
        assert rel_inf(out[i], OracleModel(t).minv(qs[i], fm[i])) < TOL, i


def test_residual_forces_and_zero_length_arguments():
    trees, B, qs, us, fs, fm = _ragged_case(seed=9)
    ud = [np.random.default_rng(100 + i).normal(size=t.nu) for i, t in enumerate(trees)]
    q, u = B.pack(qs), B.pack(us)
    full = B.split_u(B.calcResidualForceIgnoringConstraints(q, u, B.pack(fm), B.pack(fs), B.pack(ud)))
    none = B.split_u(B.calcResidualForceIgnoringConstraints(q, u))
    for i, t in enumerate(trees):
        o = OracleModel(t)
        assert rel_inf(full[i], o.residual(qs[i], us[i], fm[i], fs[i], ud[i])) < TOL, i
        assert rel_inf(none[i], o.residual(qs[i], us[i])) < TOL, i


def test_articulated_body_inertias():
    trees, B, qs, us, fs, fm = _ragged_case(seed=10)
    P, PP = B.realizeArticulatedBodyInertias(B.pack(qs))
    Ps, PPs = B.split_b(P, 21), B.split_b(PP, 21)
    for i, t in enumerate(trees):
        oP, oPP, _, _, rc = OracleModel(t).abi(qs[i], t.dofs())
        assert rel_inf(Ps[i][1:], oP) < TOL, i
        assert rel_inf(PPs[i][1:], oPP) < 1e-11, i     # P+ = P - G PH^T cancels


def test_reaction_forces():
    trees, B, qs, us, fs, fm = _ragged_case(seed=11)
    R = B.split_b(B.calcMobilizerReactionForces(B.pack(qs), B.pack(us), B.pack(fm), B.pack(fs)), 6)
    for i, t in enumerate(trees):
        assert rel_inf(R[i], OracleModel(t).reactions(qs[i], us[i], fm[i], fs[i])) < 1e-11, i


def test_shared_chain_batch_udot_and_residual():
    """One 32-body Pin/Universal/Ball chain model, many instances."""
    spec = M.spatial_chain(32, seed=2)
    n = 256
    B = _batch([spec], n=n)
    Y = M.chain_batch_states(spec, n, seed=21)
    q = B.to_device(Y[:, :spec.nq]); u = B.to_device(Y[:, spec.nq:])
    ud, _ = B.realizeAcceleration(q, u)
    ud = ud.cpu().numpy().reshape(n, -1)
    res = B.calcResidualForceIgnoringConstraints(q, u, None, None, B.to_device(ud)).cpu().numpy().reshape(n, -1)
    o = OracleModel(spec)
    worst = 0.0
    for i in range(0, n, 17):
        oud, _, _, _ = o.realize_acceleration(Y[i, :spec.nq], Y[i, spec.nq:])
        e = rel_inf(ud[i], oud)
        worst = max(worst, e)
        assert e < TOL, (i, e)
        F = o.system_forces(Y[i, :spec.nq])
        ores = o.residual(Y[i, :spec.nq], Y[i, spec.nq:], None, None, ud[i])
        assert rel_inf(res[i], ores) < TOL
        # residual of the (gravity-free) inverse dynamics at forward udot = -gravity generalized force
    print("worst chain udot rel err", worst)


def test_fixed_step_rollout_matches_oracle():
    from simbody_amd.matter import RungeKuttaMerson
    spec = M.spatial_chain(32, seed=2)
    n = 64
    B = _batch([spec], n=n)
    Y = M.chain_batch_states(spec, n, seed=22)
    q = B.to_device(Y[:, :spec.nq]); u = B.to_device(Y[:, spec.nq:])
    integ = RungeKuttaMerson(fixed_step=1e-3)
    q1, u1, st = B.stepTo(q, u, integ, 0.02)
    q1 = q1.cpu().numpy().reshape(n, -1); u1 = u1.cpu().numpy().reshape(n, -1); st = st.cpu().numpy()
    o = OracleModel(spec)
    for i in range(0, n, 9):
        y, s = o.integrate(Y[i], 0.02, 1e-3, fixed_h=1e-3)
        assert s["accepted"] == st[i, 1] == 20
        assert rel_inf(np.concatenate([q1[i], u1[i]]), y) < 1e-9, i


def test_adaptive_pendulum_rollout_matches_oracle():
    """10-link planar pendulum, RK-Merson accuracy 1e-8 to t = 1."""
    from simbody_amd.matter import RungeKuttaMerson
    spec = M.pendulum_chain(10)
    n = 8
    B = _batch([spec], n=n)
    Y = M.chain_batch_states(spec, n, seed=1)
    q = B.to_device(Y[:, :spec.nq]); u = B.to_device(Y[:, spec.nq:])
    q1, u1, st = B.stepTo(q, u, RungeKuttaMerson(accuracy=1e-8), 1.0)
    q1 = q1.cpu().numpy().reshape(n, -1); u1 = u1.cpu().numpy().reshape(n, -1); st = st.cpu().numpy()
    o = OracleModel(spec)
    for i in range(n):
        y, s = o.integrate(Y[i], 1.0, 1e-8)
        assert (s["attempts"], s["accepted"], s["rejected"]) == tuple(st[i, :3]), (i, s, st[i])
        assert rel_inf(np.concatenate([q1[i], u1[i]]), y) < 1e-8, i


def test_adaptive_tree_rollout_with_forces():
    """Uneven trees (Free roots included), applied forces, accuracy 1e-6."""
    from simbody_amd.matter import RungeKuttaMerson
    trees = mixed_trees(31, 6, nmin=8, nmax=24)
    B = _batch(trees, np.arange(6, dtype=np.int32))
    qs, us, fs = [], [], []
    for i, t in enumerate(trees):
        q, u = M.random_state(t, 31, i); qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, 31, i))
    q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-6), 0.1, fapp=B.pack(fs))
    q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
    for i, t in enumerate(trees):
        y, s = OracleModel(t).integrate(np.concatenate([qs[i], us[i]]), 0.1, 1e-6, Fapp=fs[i])
        assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), (i, s, st[i])
        assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, i


@pytest.mark.parametrize("free_root", [False, True])
def test_chain_and_tree_evaluators_agree(free_root):
    """The register-resident chain evaluator and the tree sweeps give the same
    rollout (same accept counts, states within rounding growth); with a Free
    root (a floating-base chain) as well."""
    from simbody_amd.matter import RungeKuttaMerson
    spec = M.spatial_chain(32, seed=5)
    if free_root:
        spec.mobilizer[0] = M.FREE
    n = 128
    Y = M.chain_batch_states(spec, n, seed=23)
    outs = []
    for mode in (1, 2):
        B = _batch([spec], n=n)
        B.set_evaluator(mode)
        q = B.to_device(Y[:, :spec.nq]); u = B.to_device(Y[:, spec.nq:])
        q1, u1, st = B.stepTo(q, u, RungeKuttaMerson(accuracy=1e-6), 0.05)
        outs.append((np.concatenate([q1.cpu().numpy().reshape(n, -1), u1.cpu().numpy().reshape(n, -1)], 1),
                     st.cpu().numpy()))
    assert (outs[0][1] == outs[1][1]).all()
    for i in range(n):
        assert rel_inf(outs[1][0][i], outs[0][0][i]) < 1e-9, i


def test_chain_evaluator_free_root_matches_oracle():
    from simbody_amd.matter import RungeKuttaMerson
    spec = M.spatial_chain(12, seed=8)
    spec.mobilizer[0] = M.FREE
    n = 16
    B = _batch([spec], n=n)
    Y = M.chain_batch_states(spec, n, seed=24)
    q = B.to_device(Y[:, :spec.nq]); u = B.to_device(Y[:, spec.nq:])
    q1, u1, st = B.stepTo(q, u, RungeKuttaMerson(accuracy=1e-6), 0.05)
    q1 = q1.cpu().numpy().reshape(n, -1); u1 = u1.cpu().numpy().reshape(n, -1); st = st.cpu().numpy()
    o = OracleModel(spec)
    for i in range(n):
        y, s = o.integrate(Y[i], 0.05, 1e-6)
        assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), (i, s, st[i])
        assert rel_inf(np.concatenate([q1[i], u1[i]]), y) < 1e-8, i


def test_pendulum10_fixture_on_gpu():
    """10-link pendulum fixture (tests/golden/pendulum10.json): udot at 1e-12, the
    adaptive RK-Merson run to t=1 with identical step counts and q(1) at 1e-9."""
    import json, os
    from simbody_amd.matter import RungeKuttaMerson
    g = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pendulum10.json")))
    spec = M.pendulum_chain(10)
    B = _batch([spec], n=1)
    q = B.to_device(np.array(g["q"])); u = B.to_device(np.array(g["u"]))
    ud, _ = B.realizeAcceleration(q, u)
    assert rel_inf(ud.cpu().numpy(), np.array(g["udot"])) < TOL
    q1, u1, st = B.stepTo(q, u, RungeKuttaMerson(accuracy=1e-8), 1.0)
    st = st.cpu().numpy()[0]
    assert tuple(st[:3]) == (g["attempts"], g["accepted"], g["rejected"])
    assert rel_inf(np.concatenate([q1.cpu().numpy(), u1.cpu().numpy()]), np.array(g["q1"] + g["u1"])) < 1e-9
