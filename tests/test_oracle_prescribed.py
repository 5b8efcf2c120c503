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
