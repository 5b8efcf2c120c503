"""Pinning the oracle's Translation, Planar, Cylinder and Screw mobilizers
(CPU): closed forms of single bodies on Ground, the mass-matrix invariants of
Simbody/tests/TestMassMatrix.cpp:658-887 and finite differences of the
kinematics on trees that mix all nine mobilizers, and energy conservation."""
import numpy as np
import pytest

from common import ext_trees, rel_inf, simtk_eq
from simbody_amd import models as M
from oracle_ctypes import OracleModel

G = 9.8


def _single(mob, com=(0.0, 0.0, 0.0), inertia=(0.3, 0.4, 0.5), pitch=0.0, gravity=(0.0, -G, 0.0)):
    eye = np.concatenate([np.eye(3).reshape(9), np.zeros(3)])
    return M.TreeSpec(parent=np.array([0], np.int32), mobilizer=np.array([mob], np.int32),
                      mass=np.array([2.0]), com_B=np.array([com], float),
                      inertia_B=np.array([[*inertia, 0.0, 0.0, 0.0]], float),
                      X_PF=eye[None].copy(), X_BM=eye[None].copy(),
                      gravity=np.array(gravity, float), pitch=np.array([pitch]))


def test_translation_free_fall():
    o = OracleModel(_single(M.TRANSLATION, com=(0.1, 0.2, -0.3)))
    ud, _, _, rc = o.realize_acceleration(np.array([0.3, -0.2, 1.0]), np.array([0.5, 1.0, -2.0]))
    assert rc == 0 and np.allclose(ud, [0.0, -G, 0.0], atol=1e-14)


def test_planar_free_fall_and_spin():
    """COM on the rotation axis: the angle rate is constant and the in-plane
    translation falls with g."""
    o = OracleModel(_single(M.PLANAR))
    ud, _, _, _ = o.realize_acceleration(np.array([0.7, 0.1, -0.4]), np.array([2.0, 0.3, 0.1]))
    assert np.allclose(ud, [0.0, 0.0, -G], atol=1e-14)


def test_planar_pendulum_about_moving_pivot():
    """Planar body at rest with its COM offset c along the body x axis,
    gravity -Y: a free planar body's COM falls with g and, with no torque
    about the COM, it does not spin, so theta'' = 0 and the origin's
    (x'', y'') = (0, -g)."""
    c = 0.4
    o = OracleModel(_single(M.PLANAR, com=(c, 0.0, 0.0)))
    th = 0.6
    ud, _, _, _ = o.realize_acceleration(np.array([th, 0.0, 0.0]), np.zeros(3))
    assert abs(ud[0]) < 1e-14
    assert abs(ud[1]) < 1e-14 and abs(ud[2] + G) < 1e-14


def test_cylinder_falls_along_its_axis():
    o = OracleModel(_single(M.CYLINDER, gravity=(0.0, 0.0, -G)))
    ud, _, _, _ = o.realize_acceleration(np.array([0.4, 0.2]), np.array([3.0, 0.5]))
    assert np.allclose(ud, [0.0, -G], atol=1e-14)


@pytest.mark.parametrize("pitch", [0.25, -0.1])
def test_screw_lagrangian(pitch):
    """Gravity along -z, COM on the axis: (Izz + m p^2) theta'' = -m g p."""
    m, Izz = 2.0, 0.5
    o = OracleModel(_single(M.SCREW, pitch=pitch, gravity=(0.0, 0.0, -G)))
    ud, _, _, _ = o.realize_acceleration(np.array([0.3]), np.array([1.5]))
    assert abs(ud[0] - (-m * G * pitch / (Izz + m * pitch ** 2))) < 1e-14
    K = o.kinematics(np.array([0.3]), np.array([1.5]))
    assert abs(K["X_GB"][1, 11] - 0.3 * pitch) < 1e-15          # z = p q


TREES = ext_trees(41, 5, nmin=5, nmax=14)


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_mass_matrix_invariants_all_mobilizers(idx):
    spec = TREES[idx]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 41, idx)
    nu = o.nu
    z = np.zeros(nu)
    Mm = np.column_stack([o.residual(q, z, None, None, np.eye(nu)[i]) for i in range(nu)])
    assert np.abs(Mm - Mm.T).max() <= 1e-13 * np.abs(Mm).max()
    assert np.linalg.eigvalsh(0.5 * (Mm + Mm.T)).min() > 0
    v = np.random.default_rng(idx).normal(size=nu)
    assert simtk_eq(o.minv(q, Mm @ v), v, n=nu)
    assert rel_inf(o.multiply_by_m(q, v), Mm @ v) < 1e-13
    f = np.random.default_rng(idx + 50).normal(size=nu)
    F = o.system_forces(q) + M.random_body_forces(spec, 3, idx)
    ud, _, _, _ = o.forward(q, u, f, F)
    assert np.abs(o.residual(q, u, f, F, ud)).max() < 1e-11 * max(1.0, np.abs(f).max())
    ke = o.kinematics(q, u)["ke"]
    assert abs(0.5 * u @ Mm @ u - ke) <= 1e-12 * ke


def _skew_inv(W):
    return np.array([W[2, 1] - W[1, 2], W[0, 2] - W[2, 0], W[1, 0] - W[0, 1]]) / 2


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_kinematics_finite_differences_all_mobilizers(idx):
    """V_GB against the time derivative of X_GB along qdot, and A_GB against
    that of V_GB along (qdot, udot): pins X_FM(q), H_FM, the q/u relation and
    the zero HDot_FM of the four constant-H mobilizers."""
    spec = TREES[idx]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 43, idx)
    K = o.kinematics(q, u)
    qd = K["qdot"]
    ud, _, A, _ = o.forward(q, u, np.zeros(o.nu), o.system_forces(q))
    dt = 1e-6
    Kp = o.kinematics(q + dt * qd, u + dt * ud)
    Km = o.kinematics(q - dt * qd, u - dt * ud)
    for b in range(1, o.nb + 1):
        R = K["X_GB"][b, :9].reshape(3, 3)
        dR = (Kp["X_GB"][b, :9] - Km["X_GB"][b, :9]).reshape(3, 3) / (2 * dt)
        w_fd = _skew_inv(dR @ R.T)
        v_fd = (Kp["X_GB"][b, 9:] - Km["X_GB"][b, 9:]) / (2 * dt)
        scale = max(1, np.abs(K["V_GB"][b]).max())
        assert np.allclose(K["V_GB"][b], np.concatenate([w_fd, v_fd]), atol=1e-6 * scale), b
        A_fd = (Kp["V_GB"][b] - Km["V_GB"][b]) / (2 * dt)
        assert np.allclose(A[b], A_fd, atol=2e-5 * max(1, np.abs(A[b]).max())), b


def test_energy_conservation_all_mobilizers():
    spec = ext_trees(47, 1, nmin=8, nmax=12)[0]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 47, 0)
    K0 = o.kinematics(q, u)
    y, st = o.integrate(np.concatenate([q, u]), 1.0, 1e-10)
    assert st["status"] == 0 and st["accepted"] > 10
    K1 = o.kinematics(y[:o.nq], y[o.nq:])
    E0, E1 = K0["ke"] + K0["pe"], K1["ke"] + K1["pe"]
    assert abs(E1 - E0) < 1e-8 * max(abs(E0), K0["ke"])
