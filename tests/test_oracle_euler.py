"""Pinning the oracle's Euler-angle mode (SimbodyMatterSubsystem::
setUseEulerAngles) for Ball and Free mobilizers (CPU).

In Euler mode Ball and Free keep their q slots (4 and 7) but hold body-fixed
x-y-z angles (and, for Free, the translation right after them), the last
slot unused; u keeps its meaning (RigidBodyNodeSpec_Ball.h:113-214,
_Free.h:139-257).  Checks:
- the same trees in quaternion and in Euler mode, at the same configuration
  (quaternion of the angles' rotation) and the same u, place, move and
  accelerate every body identically;
- qdot and qdotdot of the angles against the Gimbal's H(q), which maps the
  same angle rates to w_FM: H qdot = w, H qdd + HDot qdot = udot; the
  unused slot's rates are zero;
- rollouts in both modes from the same physical state end at the same
  configuration, and Euler-mode trees conserve energy and pass the
  mass-matrix invariants."""
import copy

import numpy as np
import pytest

from common import ext_trees, rel_inf, rot_to_quat, simtk_eq
from simbody_amd import models as M
from oracle_ctypes import OracleModel


def _rxyz(a):
    c, s = np.cos(a), np.sin(a)
    Rx = np.array([[1, 0, 0], [0, c[0], -s[0]], [0, s[0], c[0]]])
    Ry = np.array([[c[1], 0, s[1]], [0, 1, 0], [-s[1], 0, c[1]]])
    Rz = np.array([[c[2], -s[2], 0], [s[2], c[2], 0], [0, 0, 1]])
    return Rx @ Ry @ Rz


def _h(a):
    c0, s0, c1, s1 = np.cos(a[0]), np.sin(a[0]), np.cos(a[1]), np.sin(a[1])
    return np.array([[1, 0, s1], [0, c0, -s0 * c1], [0, s0, c0 * c1]])


def _hdot(a, ad):
    c0, s0, c1, s1 = np.cos(a[0]), np.sin(a[0]), np.cos(a[1]), np.sin(a[1])
    return np.array([[0, 0, c1 * ad[1]],
                     [0, -s0 * ad[0], -c0 * ad[0] * c1 + s0 * s1 * ad[1]],
                     [0, c0 * ad[0], -s0 * ad[0] * c1 - c0 * s1 * ad[1]]])


def _euler_trees(seed, n, nmin=4, nmax=14):
    """ext_trees in Euler mode with every third body a Ball and, in every
    other tree, a Free root."""
    trees = ext_trees(seed, n, nmin, nmax)
    for i, t in enumerate(trees):
        t.mobilizer[1::3] = M.BALL
        t.mobilizer[0] = M.FREE if i % 2 == 0 else M.BALL
        t.use_euler_angles = True
    return trees


def _to_quat_state(spec, q):
    """The quaternion-mode q of the same configuration."""
    out, qo = [], 0
    for t in spec.mobilizer:
        t = int(t)
        n = M.NQ[t]
        if t in (M.BALL, M.FREE):
            a = q[qo:qo + 3]
            out.append(rot_to_quat(_rxyz(a)))
            if t == M.FREE:
                out.append(q[qo + 3:qo + 6])
        else:
            out.append(q[qo:qo + n])
        qo += n
    return np.concatenate(out)


TREES = _euler_trees(131, 5)


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_euler_and_quaternion_modes_agree(idx):
    spec = TREES[idx]
    assert any(int(t) in (M.BALL, M.FREE) for t in spec.mobilizer)
    qspec = copy.deepcopy(spec); qspec.use_euler_angles = False
    oe, oq = OracleModel(spec), OracleModel(qspec)
    assert oe.nquat == 0 and oq.nquat > 0 and (oe.nq, oe.nu) == (oq.nq, oq.nu)
    q, u = M.random_state(spec, 131, idx)
    qq = _to_quat_state(spec, q)
    Ke, Kq = oe.kinematics(q, u), oq.kinematics(qq, u)
    assert np.abs(Ke["X_GB"] - Kq["X_GB"]).max() < 1e-12
    assert rel_inf(Ke["V_GB"], Kq["V_GB"]) < 1e-12
    F = M.random_body_forces(spec, 132, idx)
    f = np.random.default_rng(idx).normal(size=oe.nu)
    ude, qdde, Ae, _ = oe.forward(q, u, f, oe.system_forces(q) + F)
    udq, _, Aq, _ = oq.forward(qq, u, f, oq.system_forces(qq) + F)
    assert rel_inf(ude, udq) < 1e-11 and rel_inf(Ae, Aq) < 1e-11
    # angle rates through the Gimbal's H (same x-y-z convention)
    qo = uo = 0
    qd = Ke["qdot"]
    for t in spec.mobilizer:
        t = int(t)
        if t in (M.BALL, M.FREE):
            a, ad, add = q[qo:qo + 3], qd[qo:qo + 3], qdde[qo:qo + 3]
            w, wd = u[uo:uo + 3], ude[uo:uo + 3]
            assert np.abs(_h(a) @ ad - w).max() < 1e-12 * max(1, np.abs(w).max())
            assert np.abs(_h(a) @ add + _hdot(a, ad) @ ad - wd).max() < 1e-11 * max(1, np.abs(wd).max())
            last = qo + M.NQ[t] - 1
            assert qd[last] == 0.0 and qdde[last] == 0.0
            if t == M.FREE:
                assert np.array_equal(qd[qo + 3:qo + 6], u[uo + 3:uo + 6])
        qo += M.NQ[t]; uo += M.DOF[t]


def test_euler_rollout_matches_quaternion_rollout():
    spec = _euler_trees(137, 1, nmin=6, nmax=8)[0]
    qspec = copy.deepcopy(spec); qspec.use_euler_angles = False
    oe, oq = OracleModel(spec), OracleModel(qspec)
    q, u = M.random_state(spec, 137, 0)
    ye, se = oe.integrate(np.concatenate([q, u]), 0.3, 1e-10)
    yq, sq = oq.integrate(np.concatenate([_to_quat_state(spec, q), u]), 0.3, 1e-10)
    assert se["status"] == 0 and sq["status"] == 0 and se["projections"] == 0
    Ke = oe.kinematics(ye[:oe.nq], ye[oe.nq:])
    Kq = oq.kinematics(yq[:oq.nq], yq[oq.nq:])
    assert np.abs(Ke["X_GB"] - Kq["X_GB"]).max() < 1e-7
    assert np.abs(Ke["V_GB"] - Kq["V_GB"]).max() < 1e-7 * max(1, np.abs(Kq["V_GB"]).max())


def test_energy_conservation_euler():
    spec = _euler_trees(139, 1, nmin=8, nmax=10)[0]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 139, 0, uscale=0.3)
    K0 = o.kinematics(q, u)
    y, st = o.integrate(np.concatenate([q, u]), 0.5, 1e-10)
    assert st["status"] == 0 and st["accepted"] > 10
    K1 = o.kinematics(y[:o.nq], y[o.nq:])
    E0, E1 = K0["ke"] + K0["pe"], K1["ke"] + K1["pe"]
    assert abs(E1 - E0) < 1e-8 * max(abs(E0), K0["ke"])


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_mass_matrix_invariants_euler(idx):
    spec = TREES[idx]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 141, idx)
    nu = o.nu
    z = np.zeros(nu)
    Mm = np.column_stack([o.residual(q, z, None, None, np.eye(nu)[i]) for i in range(nu)])
    assert np.abs(Mm - Mm.T).max() <= 1e-13 * np.abs(Mm).max()
    assert np.linalg.eigvalsh(0.5 * (Mm + Mm.T)).min() > 0
    v = np.random.default_rng(idx).normal(size=nu)
    assert simtk_eq(o.minv(q, Mm @ v), v, n=nu)
    assert rel_inf(o.multiply_by_m(q, v), Mm @ v) < 1e-13
