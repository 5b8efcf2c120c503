"""Pinning the oracle's Bushing mobilizer (RigidBodyNodeSpec_Bushing.h) (CPU).

A Bushing holds body-fixed x-y-z angles and the translation p_FM (six q) and
uses u = qdot; its rotation and angle columns are the Gimbal's.  A Free body
in Euler-angle mode has the same six coordinates (and one unused slot) but
u = (w_FM, v_FM).  So a tree with a Bushing root and the same tree with an
Euler-mode Free root, at the same configuration and with u_free =
(H(q) qdot_angles, qdot_translation), must place, move and accelerate every
body identically, and their udots relate by w_dot = H qdd + HDot qdot.
Other bodies avoid Ball (Euler mode would turn them into angles too).
Checks also: mass-matrix invariants and energy conservation of Bushing
trees, and that a Bushing off Ground or reversed is refused."""
import copy

import numpy as np
import pytest

from common import ext_trees, rel_inf, simtk_eq
from simbody_amd import models as M
from oracle_ctypes import OracleModel


def _h(a):
    c0, s0, c1, s1 = np.cos(a[0]), np.sin(a[0]), np.cos(a[1]), np.sin(a[1])
    return np.array([[1, 0, s1], [0, c0, -s0 * c1], [0, s0, c0 * c1]])


def _hdot(a, ad):
    c0, s0, c1, s1 = np.cos(a[0]), np.sin(a[0]), np.cos(a[1]), np.sin(a[1])
    return np.array([[0, 0, c1 * ad[1]],
                     [0, -s0 * ad[0], -c0 * ad[0] * c1 + s0 * s1 * ad[1]],
                     [0, c0 * ad[0], -s0 * ad[0] * c1 - c0 * s1 * ad[1]]])


def bushing_trees(seed, n, nmin=4, nmax=14, balls=True):
    """ext_trees with a Bushing root (and, with balls=False, Gimbals in place
    of Balls)."""
    trees = ext_trees(seed, n, nmin, nmax)
    for t in trees:
        t.mobilizer[0] = M.BUSHING
        if not balls:
            t.mobilizer[t.mobilizer == M.BALL] = M.GIMBAL
    return trees


def _free_twin(spec):
    f = copy.deepcopy(spec)
    f.mobilizer[0] = M.FREE
    f.use_euler_angles = True
    return f


def _free_state(q, u):
    """The Euler-mode Free twin's q and u (root first in both layouts)."""
    a, ad = q[:3], u[:3]
    qf = np.concatenate([q[:6], [0.0], q[6:]])
    uf = np.concatenate([_h(a) @ ad, u[3:]])
    return qf, uf


TREES = bushing_trees(301, 5, balls=False)


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_bushing_matches_euler_free(idx):
    spec = TREES[idx]
    twin = _free_twin(spec)
    ob, of = OracleModel(spec), OracleModel(twin)
    assert (ob.nq, ob.nu, ob.nquat) == (of.nq - 1, of.nu, 0)
    q, u = M.random_state(spec, 301, idx)
    qf, uf = _free_state(q, u)
    Kb, Kf = ob.kinematics(q, u), of.kinematics(qf, uf)
    assert np.abs(Kb["X_GB"] - Kf["X_GB"]).max() < 1e-12
    assert rel_inf(Kb["V_GB"], Kf["V_GB"]) < 1e-12
    assert np.array_equal(Kb["qdot"][:6], u[:6])                  # u = qdot
    F = M.random_body_forces(spec, 302, idx)
    f = np.random.default_rng(idx).normal(size=ob.nu)
    fb = f.copy()
    fb[:3] = _h(q[:3]).T @ f[:3]        # same generalized work: f_b . qdot = f_w . w
    udb, qddb, Ab, _ = ob.forward(q, u, fb, ob.system_forces(q) + F)
    udf, _, Af, _ = of.forward(qf, uf, f, of.system_forces(qf) + F)
    assert rel_inf(Ab, Af) < 1e-11
    wdot = _h(q[:3]) @ udb[:3] + _hdot(q[:3], u[:3]) @ u[:3]
    assert rel_inf(np.concatenate([wdot, udb[3:]]), udf) < 1e-11
    assert np.array_equal(qddb[:6], udb[:6])                      # qdotdot = udot


def test_bushing_rollout_matches_euler_free():
    spec = bushing_trees(307, 1, nmin=6, nmax=8, balls=False)[0]
    twin = _free_twin(spec)
    ob, of = OracleModel(spec), OracleModel(twin)
    q, u = M.random_state(spec, 307, 0)
    qf, uf = _free_state(q, u)
    yb, sb = ob.integrate(np.concatenate([q, u]), 0.3, 1e-10)
    yf, sf = of.integrate(np.concatenate([qf, uf]), 0.3, 1e-10)
    assert sb["status"] == 0 and sf["status"] == 0 and sb["projections"] == 0
    Kb = ob.kinematics(yb[:ob.nq], yb[ob.nq:])
    Kf = of.kinematics(yf[:of.nq], yf[of.nq:])
    assert np.abs(Kb["X_GB"] - Kf["X_GB"]).max() < 1e-7
    assert np.abs(Kb["V_GB"] - Kf["V_GB"]).max() < 1e-7 * max(1, np.abs(Kf["V_GB"]).max())


def test_energy_conservation_bushing_with_quaternion_balls():
    """A Bushing root under which quaternion Balls hang: energy is kept and
    the Balls' quaternions stay normalized (projection skips the Bushing)."""
    spec = bushing_trees(311, 1, nmin=8, nmax=10)[0]
    spec.mobilizer[1::3] = M.BALL
    o = OracleModel(spec)
    assert o.nquat > 0
    q, u = M.random_state(spec, 311, 0, uscale=0.3)
    K0 = o.kinematics(q, u)
    y, st = o.integrate(np.concatenate([q, u]), 0.5, 1e-10)
    assert st["status"] == 0 and st["accepted"] > 10
    K1 = o.kinematics(y[:o.nq], y[o.nq:])
    E0, E1 = K0["ke"] + K0["pe"], K1["ke"] + K1["pe"]
    assert abs(E1 - E0) < 1e-8 * max(abs(E0), K0["ke"])


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_mass_matrix_invariants_bushing(idx):
    spec = bushing_trees(313, 5)[idx]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 313, idx)
    nu = o.nu
    z = np.zeros(nu)
    Mm = np.column_stack([o.residual(q, z, None, None, np.eye(nu)[i]) for i in range(nu)])
    assert np.abs(Mm - Mm.T).max() <= 1e-13 * np.abs(Mm).max()
    assert np.linalg.eigvalsh(0.5 * (Mm + Mm.T)).min() > 0
    v = np.random.default_rng(idx).normal(size=nu)
    assert simtk_eq(o.minv(q, Mm @ v), v, n=nu)
    assert rel_inf(o.multiply_by_m(q, v), Mm @ v) < 1e-13


def test_bushing_off_ground_or_reversed_is_refused():
    spec = bushing_trees(317, 1, nmin=4, nmax=4)[0]
    off = copy.deepcopy(spec)
    off.mobilizer[0] = M.PIN
    off.mobilizer[1] = M.BUSHING
    off.parent[1] = 1
    with pytest.raises(Exception):
        OracleModel(off)
    rev = copy.deepcopy(spec)
    rev.reversed = np.array([1, 0, 0, 0], np.int32)
    with pytest.raises(Exception):
        OracleModel(rev)
