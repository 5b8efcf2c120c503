"""GPU parity for trees with Translation, Planar, Cylinder, Screw and Gimbal
mobilizers mixed with the other five (FP64, 1e-12 relative per instance
against the oracle, through the C ABI): every single-shot operator on
batches of different trees (padded generic records) and on one shared tree
(the same bodies through the typed batch kernels), and adaptive rollouts
through every tree evaluator with the oracle's step counts.  A serial chain
with these mobilizers runs on the tree evaluators (the register-resident
chain evaluator and the serving slots take Pin, Slider, Universal, Ball and
Free only)."""
import ctypes

import numpy as np
import pytest

from common import bs_trees, ext_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


def _case(kind, seed):
    from simbody_amd.matter import MatterBatch
    if kind == "ragged":
        per = ext_trees(seed, 12, nmin=1, nmax=24)
        B = MatterBatch(per, tree_of=np.arange(len(per), dtype=np.int32))
    else:
        spec = ext_trees(seed, 1, nmin=20, nmax=20)[0]
        per = [spec] * 40
        B = MatterBatch([spec], n_instances=40)
    qs, us, fs, fm, vs = [], [], [], [], []
    for i, t in enumerate(per):
        q, u = M.random_state(t, seed, i)
        qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, seed + 1, i))
        r = np.random.default_rng(seed * 100 + i)
        fm.append(r.normal(size=t.nu)); vs.append(r.normal(size=t.nu))
    return per, B, qs, us, fs, fm, vs


KINDS = ["ragged", "shared"]


@pytest.mark.parametrize("kind", KINDS)
def test_kinematics_and_dynamics_operators(kind):
    per, B, qs, us, fs, fm, vs = _case(kind, 61)
    q, u = B.pack(qs), B.pack(us)
    X, V, qd, E = B.realizeKinematics(q, u)
    ud, qd2 = B.realizeAcceleration(q, u, B.pack(fs))
    ud3, A, qdd = B.calcAccelerationIgnoringConstraints(q, u, B.pack(fm), B.pack(fs))
    mi = B.multiplyByMInv(q, B.pack(vs))
    res = B.calcResidualForceIgnoringConstraints(q, u, B.pack(fm), B.pack(fs), B.pack(vs))
    P, _ = B.realizeArticulatedBodyInertias(q)
    Xs, Vs, qds, uds, ud3s, As, qdds = (B.split_b(X, 12), B.split_b(V, 6), B.split_q(qd), B.split_u(ud),
                                        B.split_u(ud3), B.split_b(A, 6), B.split_q(qdd))
    mis, ress, Ps = B.split_u(mi), B.split_u(res), B.split_b(P, 21)
    E = E.cpu().numpy().reshape(-1, 2)
    for i, t in enumerate(per):
        o = OracleModel(t)
        K = o.kinematics(qs[i], us[i])
        assert rel_inf(Xs[i], K["X_GB"]) < TOL and rel_inf(Vs[i], K["V_GB"]) < TOL, i
        assert rel_inf(qds[i], K["qdot"]) < TOL, i
        assert abs(E[i, 0] - K["ke"]) <= TOL * max(abs(K["ke"]), 1.0), i
        o_ud, _, _, _ = o.realize_acceleration(qs[i], us[i], fs[i])
        assert rel_inf(uds[i], o_ud) < TOL, i
        f_ud, f_qdd, f_A, _ = o.forward(qs[i], us[i], fm[i], fs[i])
        assert rel_inf(ud3s[i], f_ud) < TOL and rel_inf(As[i], f_A) < TOL and rel_inf(qdds[i], f_qdd) < TOL, i
        assert rel_inf(mis[i], o.minv(qs[i], vs[i])) < TOL, i
        assert rel_inf(ress[i], o.residual(qs[i], us[i], fm[i], fs[i], vs[i])) < TOL, i
        oP, _, _, _, _ = o.abi(qs[i], t.dofs())
        assert rel_inf(Ps[i][1:], oP) < TOL, i


@pytest.mark.parametrize("kind", KINDS)
def test_next_row_operators(kind):
    per, B, qs, us, fs, fm, vs = _case(kind, 62)
    q, u = B.pack(qs), B.pack(us)
    Mv = B.split_u(B.multiplyByM(q, B.pack(vs)))
    J = B.split_b(B.multiplyBySystemJacobian(q, B.pack(vs)), 6)
    Jt = B.split_u(B.multiplyBySystemJacobianTranspose(q, B.pack(fs)))
    Ab = B.split_b(B.calcBodyAccelerationFromUDot(q, u, B.pack(vs)), 6)
    R = B.split_b(B.calcCompositeBodyInertias(q), 10)
    for i, t in enumerate(per):
        o = OracleModel(t)
        assert rel_inf(Mv[i], o.multiply_by_m(qs[i], vs[i])) < TOL, i
        assert rel_inf(J[i], o.jacobian(qs[i], vs[i])) < TOL, i
        assert rel_inf(Jt[i], o.jacobian_transpose(qs[i], fs[i])) < TOL, i
        assert rel_inf(Ab[i], o.body_acceleration(qs[i], us[i], vs[i])) < TOL, i
        assert rel_inf(np.asarray(R[i])[1:], o.composite_inertias(qs[i])[1:]) < TOL, i


def test_rollouts_through_every_tree_evaluator():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    trees = ext_trees(63, 10, nmin=4, nmax=30)
    n = len(trees)
    qs, us, fs = [], [], []
    for i, t in enumerate(trees):
        q, u = M.random_state(t, 63, i); qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, 64, i))
    ref = [OracleModel(t).integrate(np.concatenate([qs[i], us[i]]), 0.1, 1e-6, Fapp=fs[i])
           for i, t in enumerate(trees)]
    for mode, lanes in ((1, 0), (0, 1), (0, 0), (3, 8)):
        B = MatterBatch(trees, tree_of=np.arange(n, dtype=np.int32))
        B.set_evaluator(mode); B.set_tree_lanes(lanes)
        q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-6), 0.1, fapp=B.pack(fs))
        q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
        for i in range(n):
            y, s = ref[i]
            assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), (mode, lanes, i, s, st[i])
            assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, (mode, lanes, i)


def test_chain_with_new_mobilizers_uses_tree_evaluator():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson, TdynError
    spec = M.spatial_chain(12, seed=4)
    spec.mobilizer[:] = [M.SCREW, M.PLANAR, M.CYLINDER, M.TRANSLATION, M.GIMBAL, M.BALL] * 2
    spec.pitch = np.linspace(-0.3, 0.3, 12)
    n = 64
    B = MatterBatch([spec], n_instances=n)
    with pytest.raises(TdynError):
        B.set_evaluator(2)                     # the chain evaluator takes the classic five only
    Y = np.stack([np.concatenate(M.random_state(spec, 65, i)) for i in range(n)])
    q1, u1, st = B.stepTo(B.to_device(Y[:, :spec.nq]), B.to_device(Y[:, spec.nq:]),
                          RungeKuttaMerson(accuracy=1e-6), 0.05)
    assert B.active_evaluator()[0] == 1
    q1 = q1.cpu().numpy().reshape(n, -1); u1 = u1.cpu().numpy().reshape(n, -1); st = st.cpu().numpy()
    o = OracleModel(spec)
    for i in range(0, n, 8):
        y, s = o.integrate(Y[i], 0.05, 1e-6)
        assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), i
        assert rel_inf(np.concatenate([q1[i], u1[i]]), y) < 1e-8, i


def test_server_rejects_new_mobilizers_and_screw_needs_pitch():
    from simbody_amd.matter import IntegratorCfgC, MatterBatch, RungeKuttaMerson, TdynError, TreeDescC, \
        fill_tree_desc, lib
    spec = M.spatial_chain(4, seed=5)
    spec.mobilizer[:] = [M.PIN, M.CYLINDER, M.PIN, M.PIN]
    keep = []
    d = fill_tree_desc(TreeDescC(), spec, keep)
    out = ctypes.c_void_p()
    cfg = RungeKuttaMerson(accuracy=1e-6).c()
    rc = lib().tdyn_server_create(ctypes.byref(d), ctypes.byref(cfg), ctypes.c_size_t(1 << 26), 64,
                                  ctypes.byref(out))
    assert rc == 1 and b"Pin, Slider, Universal, Ball and Free" in lib().tdyn_last_error()
    spec.mobilizer[1] = M.SCREW
    spec.pitch = None
    with pytest.raises(TdynError, match="screw_pitch"):
        MatterBatch([spec], n_instances=2)


@pytest.mark.parametrize("kind", KINDS)
def test_bendstretch_operators(kind):
    """BendStretch bodies (linear HDot_FM) among the other mobilizers: every
    operator against the oracle at the FP64 bar."""
    from simbody_amd.matter import MatterBatch
    if kind == "ragged":
        per = bs_trees(71, 12, nmin=2, nmax=24)
        B = MatterBatch(per, tree_of=np.arange(len(per), dtype=np.int32))
    else:
        spec = bs_trees(71, 1, nmin=20, nmax=20)[0]
        per = [spec] * 40
        B = MatterBatch([spec], n_instances=40)
    qs, us, fs, fm, vs = [], [], [], [], []
    for i, t in enumerate(per):
        q, u = M.random_state(t, 71, i); qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, 72, i))
        r = np.random.default_rng(7100 + i)
        fm.append(r.normal(size=t.nu)); vs.append(r.normal(size=t.nu))
    q, u = B.pack(qs), B.pack(us)
    X, V, qd, E = B.realizeKinematics(q, u)
    ud, A, qdd = B.calcAccelerationIgnoringConstraints(q, u, B.pack(fm), B.pack(fs))
    res = B.calcResidualForceIgnoringConstraints(q, u, B.pack(fm), B.pack(fs), B.pack(vs))
    Ab = B.calcBodyAccelerationFromUDot(q, u, B.pack(vs))
    Xs, Vs, uds, As = B.split_b(X, 12), B.split_b(V, 6), B.split_u(ud), B.split_b(A, 6)
    ress, Abs = B.split_u(res), B.split_b(Ab, 6)
    for i, t in enumerate(per):
        o = OracleModel(t)
        K = o.kinematics(qs[i], us[i])
        assert rel_inf(Xs[i], K["X_GB"]) < TOL and rel_inf(Vs[i], K["V_GB"]) < TOL, i
        f_ud, _, f_A, _ = o.forward(qs[i], us[i], fm[i], fs[i])
        assert rel_inf(uds[i], f_ud) < TOL and rel_inf(As[i], f_A) < TOL, i
        assert rel_inf(ress[i], o.residual(qs[i], us[i], fm[i], fs[i], vs[i])) < TOL, i
        assert rel_inf(Abs[i], o.body_acceleration(qs[i], us[i], vs[i])) < TOL, i


def test_bendstretch_rollouts_through_every_tree_evaluator():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    trees = bs_trees(73, 10, nmin=4, nmax=30)
    n = len(trees)
    qs, us, fs = [], [], []
    for i, t in enumerate(trees):
        q, u = M.random_state(t, 73, i); qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, 74, i))
    ref = [OracleModel(t).integrate(np.concatenate([qs[i], us[i]]), 0.1, 1e-6, Fapp=fs[i])
           for i, t in enumerate(trees)]
    for mode, lanes in ((1, 0), (0, 1), (0, 0), (3, 8)):
        B = MatterBatch(trees, tree_of=np.arange(n, dtype=np.int32))
        B.set_evaluator(mode); B.set_tree_lanes(lanes)
        q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-6), 0.1, fapp=B.pack(fs))
        q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
        for i in range(n):
            y, s = ref[i]
            assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), (mode, lanes, i, s, st[i])
            assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, (mode, lanes, i)
