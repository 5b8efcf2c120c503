"""GPU parity for the Bushing mobilizer (FP64, 1e-12 relative per instance
against the oracle, through the C ABI): trees with a Bushing root under
which every other mobilizer hangs (quaternion Balls included, so the
quaternion projection must leave the Bushing's angles alone), every
single-shot operator on batches of different trees and on one shared tree,
and adaptive rollouts through every tree evaluator with the oracle's step
counts.  The chain evaluator and the serving loop refuse Bushing models."""
import ctypes

import numpy as np
import pytest

from common import rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel
from test_oracle_bushing import bushing_trees

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


def _case(kind, seed):
    from simbody_amd.matter import MatterBatch
    if kind == "ragged":
        per = bushing_trees(seed, 12, nmin=1, nmax=24)
        B = MatterBatch(per, tree_of=np.arange(len(per), dtype=np.int32))
    else:
        spec = bushing_trees(seed, 1, nmin=20, nmax=20)[0]
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


@pytest.mark.parametrize("kind", ["ragged", "shared"])
def test_operators_with_bushing(kind):
    per, B, qs, us, fs, fm, vs = _case(kind, 321)
    q, u = B.pack(qs), B.pack(us)
    X, V, qd, E = B.realizeKinematics(q, u)
    ud3, A, qdd = B.calcAccelerationIgnoringConstraints(q, u, B.pack(fm), B.pack(fs))
    mi = B.multiplyByMInv(q, B.pack(vs))
    res = B.calcResidualForceIgnoringConstraints(q, u, B.pack(fm), B.pack(fs), B.pack(vs))
    Mv = B.multiplyByM(q, B.pack(vs))
    J = B.multiplyBySystemJacobian(q, B.pack(vs))
    Jt = B.multiplyBySystemJacobianTranspose(q, B.pack(fs))
    Ab = B.calcBodyAccelerationFromUDot(q, u, B.pack(vs))
    Xs, Vs, qds, ud3s, As, qdds = (B.split_b(X, 12), B.split_b(V, 6), B.split_q(qd), B.split_u(ud3),
                                   B.split_b(A, 6), B.split_q(qdd))
    mis, ress, Mvs = B.split_u(mi), B.split_u(res), B.split_u(Mv)
    Js, Jts, Abs = B.split_b(J, 6), B.split_u(Jt), B.split_b(Ab, 6)
    for i, t in enumerate(per):
        o = OracleModel(t)
        K = o.kinematics(qs[i], us[i])
        assert rel_inf(Xs[i], K["X_GB"]) < TOL and rel_inf(Vs[i], K["V_GB"]) < TOL, i
        assert rel_inf(qds[i], K["qdot"]) < TOL, i
        f_ud, f_qdd, f_A, _ = o.forward(qs[i], us[i], fm[i], fs[i])
        assert rel_inf(ud3s[i], f_ud) < TOL and rel_inf(As[i], f_A) < TOL, i
        assert rel_inf(qdds[i], f_qdd) < TOL, i
        assert rel_inf(mis[i], o.minv(qs[i], vs[i])) < TOL, i
        assert rel_inf(ress[i], o.residual(qs[i], us[i], fm[i], fs[i], vs[i])) < TOL, i
        assert rel_inf(Mvs[i], o.multiply_by_m(qs[i], vs[i])) < TOL, i
        assert rel_inf(Js[i], o.jacobian(qs[i], vs[i])) < TOL, i
        assert rel_inf(Jts[i], o.jacobian_transpose(qs[i], fs[i])) < TOL, i
        assert rel_inf(Abs[i], o.body_acceleration(qs[i], us[i], vs[i])) < TOL, i


def test_rollouts_with_bushing_through_every_tree_evaluator():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    trees = bushing_trees(331, 10, nmin=4, nmax=30)
    n = len(trees)
    qs, us, fs = [], [], []
    for i, t in enumerate(trees):
        q, u = M.random_state(t, 331, i); qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, 332, i))
    ref = [OracleModel(t).integrate(np.concatenate([qs[i], us[i]]), 0.1, 1e-6, Fapp=fs[i])
           for i, t in enumerate(trees)]
    assert any(s["projections"] > 0 for _, s in ref) or any(OracleModel(t).nquat > 0 for t in trees)
    for mode, lanes, active in ((1, 0, 1), (3, 1, 3), (0, 0, 4), (3, 8, 4)):
        B = MatterBatch(trees, tree_of=np.arange(n, dtype=np.int32))
        B.set_evaluator(mode); B.set_tree_lanes(lanes)
        q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), RungeKuttaMerson(accuracy=1e-6), 0.1, fapp=B.pack(fs))
        assert B.active_evaluator()[0] == active, (mode, lanes)
        q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
        for i in range(n):
            y, s = ref[i]
            assert (s["accepted"], s["rejected"], s["projections"]) == tuple(st[i, 1:4]), (mode, lanes, i, s, st[i])
            assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, (mode, lanes, i)


def test_bushing_chain_runs_on_tree_evaluators_and_server_refuses_it():
    from simbody_amd.matter import IntegratorCfgC, MatterBatch, RungeKuttaMerson, TdynError, TreeDescC, \
        fill_tree_desc, lib
    spec = M.spatial_chain(10, seed=7)
    spec.mobilizer[0] = M.BUSHING
    n = 64
    B = MatterBatch([spec], n_instances=n)
    with pytest.raises(TdynError):
        B.set_evaluator(2)                     # the chain evaluator runs no Bushing
    Y = np.stack([np.concatenate(M.random_state(spec, 67, i)) for i in range(n)])
    q1, u1, st = B.stepTo(B.to_device(Y[:, :spec.nq]), B.to_device(Y[:, spec.nq:]),
                          RungeKuttaMerson(accuracy=1e-6), 0.05)
    assert B.active_evaluator()[0] != 2
    q1 = q1.cpu().numpy().reshape(n, -1); u1 = u1.cpu().numpy().reshape(n, -1); st = st.cpu().numpy()
    o = OracleModel(spec)
    for i in range(0, n, 8):
        y, s = o.integrate(Y[i], 0.05, 1e-6)
        assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), i
        assert rel_inf(np.concatenate([q1[i], u1[i]]), y) < 1e-8, i
    keep = []
    d = fill_tree_desc(TreeDescC(), spec, keep)
    out = ctypes.c_void_p()
    cfg = RungeKuttaMerson(accuracy=1e-6).c()
    rc = lib().tdyn_server_create(ctypes.byref(d), ctypes.byref(cfg), ctypes.c_size_t(1 << 26), 64,
                                  ctypes.byref(out))
    assert rc == 1 and b"serving path runs" in lib().tdyn_last_error()
