"""Pinning the oracle's Free mobilizer off Ground (a Free body whose parent
is another body, RigidBodyNodeSpec_Free.h; Simbody allows a Free anywhere)
(CPU): on trees mixing the other mobilizers with Free bodies mid-tree,
V_GB and A_GB against finite differences of X_GB and V_GB along (qdot,
udot), the mass-matrix invariants, and energy conservation."""
import numpy as np
import pytest

from common import ext_trees, rel_inf, simtk_eq
from simbody_amd import models as M
from oracle_ctypes import OracleModel
from test_oracle_mobilizers import _skew_inv


def free_mid_trees(seed, n, nmin=5, nmax=14):
    """ext_trees with bodies 3 and nb (when nb > 3) on Free mobilizers."""
    trees = ext_trees(seed, n, nmin, nmax)
    for t in trees:
        t.mobilizer[2] = M.FREE
        t.mobilizer[-1] = M.FREE
        assert t.parent[2] != 0 or t.parent[-1] != 0 or t.nb < 4
    return trees


TREES = free_mid_trees(461, 5)


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_free_off_ground_finite_differences(idx):
    spec = TREES[idx]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 461, idx)
    K = o.kinematics(q, u)
    qd = K["qdot"]
    ud, _, A, _ = o.forward(q, u, np.zeros(o.nu), o.system_forces(q))
    dt = 1e-6
    Kp = o.kinematics(q + dt * qd, u + dt * ud)
    Km = o.kinematics(q - dt * qd, u - dt * ud)
    for b in range(1, o.nb + 1):
        R = K["X_GB"][b, :9].reshape(3, 3)
        dR = (Kp["X_GB"][b, :9] - Km["X_GB"][b, :9]).reshape(3, 3) / (2 * dt)
        v_fd = (Kp["X_GB"][b, 9:] - Km["X_GB"][b, 9:]) / (2 * dt)
        scale = max(1, np.abs(K["V_GB"][b]).max())
        assert np.allclose(K["V_GB"][b], np.concatenate([_skew_inv(dR @ R.T), v_fd]), atol=1e-6 * scale), b
        A_fd = (Kp["V_GB"][b] - Km["V_GB"][b]) / (2 * dt)
        assert np.allclose(A[b], A_fd, atol=2e-5 * max(1, np.abs(A[b]).max())), b


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_free_off_ground_mass_matrix_invariants(idx):
    spec = TREES[idx]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 463, idx)
    nu = o.nu
    z = np.zeros(nu)
    Mm = np.column_stack([o.residual(q, z, None, None, np.eye(nu)[i]) for i in range(nu)])
    assert np.abs(Mm - Mm.T).max() <= 1e-13 * np.abs(Mm).max()
    assert np.linalg.eigvalsh(0.5 * (Mm + Mm.T)).min() > 0
    v = np.random.default_rng(idx).normal(size=nu)
    assert simtk_eq(o.minv(q, Mm @ v), v, n=nu)
    assert rel_inf(o.multiply_by_m(q, v), Mm @ v) < 1e-13


def test_free_off_ground_energy():
    spec = free_mid_trees(467, 1, nmin=8, nmax=10)[0]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 467, 0, uscale=0.3)
    K0 = o.kinematics(q, u)
    y, st = o.integrate(np.concatenate([q, u]), 0.5, 1e-10)
    assert st["status"] == 0 and st["accepted"] > 10
    K1 = o.kinematics(y[:o.nq], y[o.nq:])
    E0, E1 = K0["ke"] + K0["pe"], K1["ke"] + K1["pe"]
    assert abs(E1 - E0) < 1e-8 * max(abs(E0), K0["ke"])
