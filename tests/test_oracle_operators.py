"""Pinning the oracle's O(n) operators that reuse the tree caches (CPU): the
mass-matrix product, the system Jacobian and its transpose, body
accelerations from udot and composite body inertias.  Each is checked
against a route that does not share its recursion -- the dense mass matrix
built from residual columns, the kinetic energy, the velocity kinematics,
the forward dynamics, the adjoint identity and a direct parallel-axis sum
over each subtree -- so that a dropped term, a wrong sign or a transposed
operand fails at least one of them."""
import numpy as np
import pytest

from common import mixed_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

TREES = mixed_trees(23, 6, nmin=3, nmax=18, free_every=2)


def _rng(i):
    return np.random.default_rng(1000 + i)


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_multiply_by_m_matches_dense_residual_and_kinetic_energy(idx):
    spec = TREES[idx]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 23, idx)
    nu = o.nu
    z = np.zeros(nu)
    v = _rng(idx).normal(size=nu)
    Mv = o.multiply_by_m(q, v)
    # residual at u = 0 without applied forces is M udot, column by column
    Mm = np.column_stack([o.residual(q, z, None, None, np.eye(nu)[i]) for i in range(nu)])
    assert rel_inf(Mv, Mm @ v) < 1e-13
    assert rel_inf(o.minv(q, Mv), v) < 1e-11                   # M^-1 (M v) = v
    ke = o.kinematics(q, u)["ke"]
    assert abs(0.5 * u @ o.multiply_by_m(q, u) - ke) <= 1e-13 * abs(ke)   # KE = u^T M u / 2


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_system_jacobian_and_transpose(idx):
    spec = TREES[idx]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 23, idx)
    nu, nb = o.nu, spec.nb
    V = o.kinematics(q, u)["V_GB"]
    assert rel_inf(o.jacobian(q, u).reshape(-1), V.reshape(-1)) < 1e-13     # J u = V_GB
    assert np.all(o.jacobian(q, u)[0] == 0)
    F = _rng(idx).normal(size=(nb + 1, 6))
    v = _rng(idx + 7).normal(size=nu)
    JtF = o.jacobian_transpose(q, F)
    # adjoint: F . (J v) = (J^T F) . v, Ground's entry ignored
    lhs = np.sum(F[1:] * o.jacobian(q, v)[1:])
    assert abs(lhs - JtF @ v) <= 1e-12 * (np.abs(F[1:]).sum() * np.abs(o.jacobian(q, v)).max())
    # residual at u = 0, udot = 0 with body forces F is -J^T F
    F0 = F.copy(); F0[0] = 0
    assert rel_inf(JtF, -o.residual(q, np.zeros(nu), None, F0.reshape(-1), None)) < 1e-13


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_body_acceleration_from_udot(idx):
    spec = TREES[idx]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 23, idx)
    F = o.system_forces(q) + M.random_body_forces(spec, 5, idx)
    f = _rng(idx).normal(size=o.nu)
    ud, _, A, _ = o.forward(q, u, f, F.reshape(-1))
    assert rel_inf(o.body_acceleration(q, u, ud).reshape(-1), A.reshape(-1)) < 1e-13
    # udot = 0 leaves the velocity-dependent (coriolis) part: zero at u = 0
    assert np.all(o.body_acceleration(q, np.zeros(o.nu), None) == 0)


def _unit_inertia_B(spec, b):
    xx, yy, zz, xy, xz, yz = spec.inertia_B[b - 1]
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]]) / spec.mass[b - 1]


def _pm(p):
    return (p @ p) * np.eye(3) - np.outer(p, p)


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_composite_inertias_direct_subtree_sum(idx):
    """R_b from the parallel-axis theorem summed directly over the subtree of
    b (each body's central inertia moved to b's origin), not through the
    child-by-child shifts of the recursion."""
    spec = TREES[idx]
    o = OracleModel(spec)
    q, _ = M.random_state(spec, 23, idx)
    R = o.composite_inertias(q)
    X = o.kinematics(q, np.zeros(o.nu))["X_GB"]
    nb = spec.nb
    kids = [[] for _ in range(nb + 1)]
    for b in range(1, nb + 1):
        kids[spec.parent[b - 1]].append(b)

    def subtree(b):
        out = [b]
        for c in kids[b]:
            out += subtree(c)
        return out
    assert np.isinf(R[0, 0]) and np.all(R[0, 1:4] == 0) and np.all(R[0, 4:7] == 1)
    for b in range(1, nb + 1):
        Rb = X[b, :9].reshape(3, 3); ob = X[b, 9:]
        m = 0.0; mom = np.zeros(3); I = np.zeros((3, 3))
        for k in subtree(b):
            Rk = X[k, :9].reshape(3, 3); mk = spec.mass[k - 1]
            pk = Rk @ spec.com_B[k - 1]                  # mass center from Bk, in G
            Gk = Rk @ _unit_inertia_B(spec, k) @ Rk.T    # unit inertia about Bk, in G
            r = X[k, 9:] + pk - ob                        # mass center from Bb
            m += mk; mom += mk * r
            I += mk * (Gk - _pm(pk)) + mk * _pm(r)
        G = I / m
        ref = np.concatenate([[m], mom / m, [G[0, 0], G[1, 1], G[2, 2], G[0, 1], G[0, 2], G[1, 2]]])
        assert rel_inf(R[b], ref) < 1e-12, (b, R[b], ref)
        del Rb


def test_composite_inertia_of_locked_tree_gives_kinetic_energy():
    """With only the Free root moving (every other joint locked), the whole
    tree moves rigidly with body 1: KE = V1 . (R1 V1) / 2."""
    for idx in range(3):
        spec = M.random_tree(31, idx, nmin=10, nmax=30, general_frames=True)
        spec.mobilizer[0] = M.FREE
        o = OracleModel(spec)
        q, _ = M.random_state(spec, 31, idx)
        u = np.zeros(o.nu); u[:6] = _rng(idx).normal(size=6)
        kin = o.kinematics(q, u)
        V1 = kin["V_GB"][1]
        R1 = o.composite_inertias(q)[1]
        m, p, g = R1[0], R1[1:4], R1[4:]
        G = np.array([[g[0], g[3], g[4]], [g[3], g[1], g[5]], [g[4], g[5], g[2]]])
        w, v = V1[:3], V1[3:]
        RV = m * np.concatenate([G @ w + np.cross(p, v), v - np.cross(p, w)])
        assert abs(0.5 * V1 @ RV - kin["ke"]) <= 1e-12 * kin["ke"]
