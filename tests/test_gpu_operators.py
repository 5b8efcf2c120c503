"""GPU parity of the O(n) operators that reuse the tree caches (multiplyByM,
the system Jacobian and its transpose, body accelerations from udot,
composite body inertias) with the oracle, at the FP64 bar of
test_gpu_parity (1e-12 relative per instance), through the C ABI.  Batches:
different trees with general joint frames and Free roots (the padded
generic records), and one shared tree for many instances (the
structure-of-arrays records)."""
import numpy as np
import pytest

from common import mixed_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


def _case(kind, seed):
    from simbody_amd.matter import MatterBatch
    if kind == "ragged":
        trees = mixed_trees(seed, 13, nmin=1, nmax=30)
        B = MatterBatch(trees, tree_of=np.arange(len(trees), dtype=np.int32))
        per = trees
    else:
        spec = M.random_tree(seed, 0, nmin=40, nmax=40, general_frames=True)
        spec.mobilizer[0] = M.FREE
        B = MatterBatch([spec], n_instances=70)
        per = [spec] * 70
    qs, us, vs, Fs = [], [], [], []
    for i, t in enumerate(per):
        q, u = M.random_state(t, seed, i)
        qs.append(q); us.append(u)
        vs.append(np.random.default_rng(seed * 1000 + i).normal(size=t.nu))
        Fs.append(M.random_body_forces(t, seed + 1, i))
    return per, B, qs, us, vs, Fs


KINDS = ["ragged", "shared"]


@pytest.mark.parametrize("kind", KINDS)
def test_multiply_by_m(kind):
    per, B, qs, us, vs, Fs = _case(kind, 3)
    out = B.split_u(B.multiplyByM(B.pack(qs), B.pack(vs)))
    for i, t in enumerate(per):
        assert rel_inf(out[i], OracleModel(t).multiply_by_m(qs[i], vs[i])) < TOL, i


@pytest.mark.parametrize("kind", KINDS)
def test_system_jacobian_and_transpose(kind):
    per, B, qs, us, vs, Fs = _case(kind, 4)
    q = B.pack(qs)
    J = B.split_b(B.multiplyBySystemJacobian(q, B.pack(vs)), 6)
    Jt = B.split_u(B.multiplyBySystemJacobianTranspose(q, B.pack(Fs)))
    for i, t in enumerate(per):
        o = OracleModel(t)
        assert rel_inf(J[i], o.jacobian(qs[i], vs[i])) < TOL, i
        assert rel_inf(Jt[i], o.jacobian_transpose(qs[i], Fs[i])) < TOL, i


@pytest.mark.parametrize("kind", KINDS)
def test_body_acceleration_from_udot(kind):
    per, B, qs, us, vs, Fs = _case(kind, 5)
    q, u = B.pack(qs), B.pack(us)
    A = B.split_b(B.calcBodyAccelerationFromUDot(q, u, B.pack(vs)), 6)
    A0 = B.split_b(B.calcBodyAccelerationFromUDot(q, u, None), 6)
    for i, t in enumerate(per):
        o = OracleModel(t)
        assert rel_inf(A[i], o.body_acceleration(qs[i], us[i], vs[i])) < TOL, i
        assert rel_inf(A0[i], o.body_acceleration(qs[i], us[i], None)) < TOL, i


@pytest.mark.parametrize("kind", KINDS)
def test_composite_body_inertias(kind):
    per, B, qs, us, vs, Fs = _case(kind, 6)
    R = B.split_b(B.calcCompositeBodyInertias(B.pack(qs)), 10)
    for i, t in enumerate(per):
        ref = OracleModel(t).composite_inertias(qs[i])
        Ri = np.asarray(R[i]).reshape(-1, 10)
        assert np.isinf(Ri[0, 0]) and np.array_equal(Ri[0, 1:], ref[0, 1:]), i
        assert rel_inf(Ri[1:], ref[1:]) < TOL, i
