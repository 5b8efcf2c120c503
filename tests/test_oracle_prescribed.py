"""Pinning the oracle's forward dynamics with known udot (prescribed motion
at the acceleration level and locks; the isUDotKnown branches of
RigidBodyNodeSpec.cpp:292-296 and 355-446) on the CPU, against the
independently pinned residual operator (calcResidualForceIgnoringConstraints)
and the unprescribed forward dynamics:
- every mobility prescribed: the returned tau is minus the inverse dynamics
  M udot + c - f - J^T F of the prescribed udot (Simbody keeps tau on the
  left-hand side, M udot + c + tau = f + J^T F; RigidBodyNodeSpec.cpp:436-443);
- the free dynamics' own udot prescribed: no force is needed (tau = 0);
- any subset prescribed (or locked, udot = 0): the full udot (known entries
  plus the computed ones) and tau satisfy M udot + c - f - J^T F = -tau, with
  tau zero on the free mobilities."""
import numpy as np
import pytest

from common import ext_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

TREES = ext_trees(171, 6, nmin=3, nmax=16)


def _setup(idx):
    spec = TREES[idx]
    o = OracleModel(spec)
    q, u = M.random_state(spec, 171, idx)
    rng = np.random.default_rng(idx)
    f = rng.normal(size=o.nu)
    F = o.system_forces(q) + M.random_body_forces(spec, 172, idx)
    return spec, o, q, u, f, F, rng


def _umask(spec, presc):
    m = []
    for b, t in enumerate(spec.mobilizer, start=1):
        m += [bool(presc[b])] * M.DOF[int(t)]
    return np.array(m, bool)


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_all_prescribed_is_inverse_dynamics(idx):
    spec, o, q, u, f, F, rng = _setup(idx)
    presc = np.ones(spec.nb + 1, np.int32)
    ud_p = rng.normal(size=o.nu)
    ud, tau, A, rc = o.forward_prescribed(q, u, f, F, presc, ud_p)
    assert rc == 0 and np.array_equal(ud, ud_p)
    assert rel_inf(-tau, o.residual(q, u, f, F, ud_p)) < 1e-12
    assert rel_inf(A, o.body_acceleration(q, u, ud_p)) < 1e-12


@pytest.mark.parametrize("idx", range(len(TREES)))
def test_prescribing_the_free_motion_needs_no_force(idx):
    spec, o, q, u, f, F, rng = _setup(idx)
    ud0, _, A0, _ = o.forward(q, u, f, F)
    presc = (rng.uniform(size=spec.nb + 1) < 0.5).astype(np.int32)
    known = np.where(_umask(spec, presc), ud0, 0.0)
    ud, tau, A, _ = o.forward_prescribed(q, u, f, F, presc, known)
    scale = max(1.0, np.abs(f).max(), np.abs(o.residual(q, u, f, F, np.zeros(o.nu))).max())
    assert np.abs(tau).max() < 1e-11 * scale
    assert rel_inf(ud, ud0) < 1e-11 and rel_inf(A, A0) < 1e-11


@pytest.mark.parametrize("idx", range(len(TREES)))
@pytest.mark.parametrize("lock", [False, True])
def test_subset_prescribed_satisfies_the_equations_of_motion(idx, lock):
    spec, o, q, u, f, F, rng = _setup(idx)
    presc = (rng.uniform(size=spec.nb + 1) < 0.4).astype(np.int32)
    presc[0] = 0
    if not presc[1:].any():
        presc[1] = 1
    mask = _umask(spec, presc)
    known = np.zeros(o.nu) if lock else np.where(mask, rng.normal(size=o.nu), 0.0)
    ud, tau, A, rc = o.forward_prescribed(q, u, f, F, presc, known)
    assert rc == 0
    assert np.array_equal(ud[mask], known[mask]) and np.all(tau[~mask] == 0.0)
    res = o.residual(q, u, f, F, ud)
    assert np.abs(res + tau).max() < 1e-11 * max(1.0, np.abs(f).max(), np.abs(tau).max())


def _lock_setup(seed, n=1, nmin=6, nmax=12):
    spec = ext_trees(seed, n, nmin, nmax)[0]
    o = OracleModel(spec)
    q, u = M.random_state(spec, seed, 0)
    rng = np.random.default_rng(seed)
    presc = np.zeros(spec.nb + 1, np.int32)
    presc[1 + rng.choice(spec.nb, size=max(1, spec.nb // 3), replace=False)] = 1
    mask = _umask(spec, presc)
    return spec, o, q, u, presc, mask


def test_locked_mobilities_hold_still_and_energy_is_conserved():
    """Locks (known udot 0 with u 0): the locked q and u do not move over a
    rollout, the others do, and the locks do no work."""
    spec, o, q, u, presc, mask = _lock_setup(191)
    u = np.where(mask, 0.0, u)
    qmask = np.zeros(o.nq, bool)
    qi, ui = spec.offsets()
    for b in range(1, spec.nb + 1):
        if presc[b]:
            qmask[qi[b - 1]:qi[b - 1] + M.NQ[int(spec.mobilizer[b - 1])]] = True
    y0 = np.concatenate([q, u])
    K0 = o.kinematics(q, u)
    y, st = o.integrate_known(y0, 0.5, 1e-10, presc, np.zeros(o.nu))
    assert st["status"] == 0 and st["accepted"] > 5
    q1, u1 = y[:o.nq], y[o.nq:]
    assert np.array_equal(u1[mask], u[mask])
    assert np.abs(q1[qmask] - q[qmask]).max() < 1e-14    # (quaternions only renormalised)
    assert np.abs(q1[~qmask] - q[~qmask]).max() > 1e-3
    K1 = o.kinematics(q1, u1)
    E0, E1 = K0["ke"] + K0["pe"], K1["ke"] + K1["pe"]
    assert abs(E1 - E0) < 1e-8 * max(abs(E0), K0["ke"])


def test_locked_pin_moves_like_a_weld():
    """A chain whose body 2 is a Pin locked at angle a moves exactly as the
    chain with a Weld there and the rotation folded into its X_PF."""
    spec = M.spatial_chain(6, seed=9)
    spec.mobilizer[:] = [M.PIN] * 6
    o = OracleModel(spec)
    q, u = M.random_state(spec, 193, 0)
    q[1] = 0.7; u[1] = 0.0
    presc = np.zeros(7, np.int32); presc[2] = 1
    y, st = o.integrate_known(np.concatenate([q, u]), 0.5, 1e-10, presc, np.zeros(6))
    weld = M.TreeSpec(**{**spec.__dict__})
    weld.mobilizer = np.array([M.PIN, M.WELD, M.PIN, M.PIN, M.PIN, M.PIN], np.int32)
    X = weld.X_PF.copy()
    R = X[1, :9].reshape(3, 3); c, s = np.cos(0.7), np.sin(0.7)
    X[1, :9] = (R @ np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])).ravel()
    weld.X_PF = X
    ow = OracleModel(weld)
    keep = np.array([0, 2, 3, 4, 5])
    yw, stw = ow.integrate(np.concatenate([q[keep], u[keep]]), 0.5, 1e-10)
    assert np.abs(y[:6][keep] - yw[:5]).max() < 1e-8 and np.abs(y[6:][keep] - yw[5:]).max() < 1e-7


def test_constant_known_udot_is_integrated_exactly():
    spec, o, q, u, presc, mask = _lock_setup(197)
    a = np.where(mask, np.random.default_rng(3).normal(size=o.nu), 0.0)
    y, st = o.integrate_known(np.concatenate([q, u]), 0.3, 1e-8, presc, a)
    assert st["status"] == 0
    assert np.abs(y[o.nq:][mask] - (u[mask] + 0.3 * a[mask])).max() < 1e-12
