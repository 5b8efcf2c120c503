"""GPU parity for Free bodies off Ground (FP64, 1e-12 relative per instance
against the oracle, through the C ABI): every single-shot operator on
batches of different trees and on one shared tree with Free bodies
mid-tree, and adaptive rollouts with the oracle's step counts (such batches
run on the tree sweeps).  The chain evaluator and the serving loop, which
carry a Free body first only, refuse a chain with a Free body further on."""
import ctypes

import numpy as np
import pytest

from common import rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel
from test_oracle_free_offground import free_mid_trees

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("kind", ["ragged", "shared"])
def test_operators_and_rollouts_with_free_off_ground(kind):
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    if kind == "ragged":
        per = free_mid_trees(471, 10, nmin=4, nmax=24)
        B = MatterBatch(per, tree_of=np.arange(len(per), dtype=np.int32))
    else:
        spec = free_mid_trees(471, 1, nmin=20, nmax=20)[0]
        per = [spec] * 32
        B = MatterBatch([spec], n_instances=32)
    qs, us, fs, fm, vs = [], [], [], [], []
    for i, t in enumerate(per):
        q, u = M.random_state(t, 471, i); qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, 472, i))
        r = np.random.default_rng(47100 + i)
        fm.append(r.normal(size=t.nu)); vs.append(r.normal(size=t.nu))
    q, u = B.pack(qs), B.pack(us)
    X, V, qd, E = B.realizeKinematics(q, u)
    ud, A, qdd = B.calcAccelerationIgnoringConstraints(q, u, B.pack(fm), B.pack(fs))
    mi = B.multiplyByMInv(q, B.pack(vs))
    res = B.calcResidualForceIgnoringConstraints(q, u, B.pack(fm), B.pack(fs), B.pack(vs))
    Xs, Vs, uds, As, mis, ress = (B.split_b(X, 12), B.split_b(V, 6), B.split_u(ud), B.split_b(A, 6),
                                  B.split_u(mi), B.split_u(res))
    for i, t in enumerate(per):
        o = OracleModel(t)
        K = o.kinematics(qs[i], us[i])
        assert rel_inf(Xs[i], K["X_GB"]) < TOL and rel_inf(Vs[i], K["V_GB"]) < TOL, i
        f_ud, _, f_A, _ = o.forward(qs[i], us[i], fm[i], fs[i])
        assert rel_inf(uds[i], f_ud) < TOL and rel_inf(As[i], f_A) < TOL, i
        assert rel_inf(mis[i], o.minv(qs[i], vs[i])) < TOL, i
        assert rel_inf(ress[i], o.residual(qs[i], us[i], fm[i], fs[i], vs[i])) < TOL, i
    q1, u1, st = B.stepTo(q, u, RungeKuttaMerson(accuracy=1e-6), 0.05, fapp=B.pack(fs))
    assert B.active_evaluator()[0] == 1
    q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
    for i, t in enumerate(per[:12]):
        y, s = OracleModel(t).integrate(np.concatenate([qs[i], us[i]]), 0.05, 1e-6, Fapp=fs[i])
        assert (s["accepted"], s["rejected"], s["projections"]) == tuple(st[i, 1:4]), (i, s, st[i])
        assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, i


def test_chain_with_free_body_further_on_is_refused_by_chain_paths():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson, TdynError, TreeDescC, fill_tree_desc, lib
    spec = M.spatial_chain(8, seed=5)
    spec.mobilizer[3] = M.FREE
    B = MatterBatch([spec], n_instances=16)
    with pytest.raises(TdynError):
        B.set_evaluator(2)
    keep = []
    d = fill_tree_desc(TreeDescC(), spec, keep)
    out = ctypes.c_void_p()
    cfg = RungeKuttaMerson(accuracy=1e-6).c()
    rc = lib().tdyn_server_create(ctypes.byref(d), ctypes.byref(cfg), ctypes.c_size_t(1 << 26), 16,
                                  ctypes.byref(out))
    assert rc == 1 and b"Free body as the first body only" in lib().tdyn_last_error()
