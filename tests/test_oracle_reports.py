"""Pinning the oracle's integration through report times with interpolation
off (TimeStepper::stepTo(t_k) in turn with
Integrator::setAllowInterpolation(false), AbstractIntegratorRep.cpp:216-240,
512-530) (CPU).

- one report time is the plain integration, bit for bit (state and counts);
- in fixed-step mode, report times on the step grid change nothing but the
  rounding of the step ends: the same accepted steps and states to 1e-12;
- the reported states are the trajectory: against a 1e-12-accuracy run to
  each report time, within the integrator's accuracy;
- energy is conserved at every report time of a conservative tree;
- report times only ever shorten steps: never fewer accepted steps.
Reports by interpolation (interpolation on, the reference's default,
AbstractIntegratorRep.cpp:63-93, IntegratorRep.h:285-313):
- the trajectory and counts are those of the plain integration, bit for bit;
  reports at t0 and at the final time are the end states;
- constant acceleration (a Slider under gravity): every report matches the
  closed form to rounding (RK-Merson and the cubic Hermite interpolant are
  exact for a quadratic);
- reports agree with a 1e-12-accuracy run; reported quaternions are within
  the constraint tolerance (projectQ of the interpolated state)."""
import numpy as np
import pytest

from common import rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel


def _tree(seed, nb=10):
    spec = M.random_tree(seed, 0, nmin=nb, nmax=nb, general_frames=True)
    spec.mobilizer[0] = M.FREE
    return spec


def test_one_report_time_is_the_plain_integration():
    spec = _tree(401)
    o = OracleModel(spec)
    q, u = M.random_state(spec, 401, 0)
    y0 = np.concatenate([q, u])
    F = M.random_body_forces(spec, 402, 0)
    y, s = o.integrate(y0, 0.2, 1e-6, Fapp=F)
    Y, s2 = o.integrate_stops(y0, [0.2], 1e-6, Fapp=F)
    assert np.array_equal(Y[0], y) and s == s2


def test_fixed_step_grid_report_times_change_nothing():
    spec = _tree(403)
    o = OracleModel(spec)
    q, u = M.random_state(spec, 403, 0)
    y0 = np.concatenate([q, u])
    h = 0.01
    Y, s = o.integrate_stops(y0, [0.05, 0.1, 0.2], 1e-3, fixed_h=h)
    assert s["accepted"] == 20 and s["status"] == 0
    for k, T in enumerate((0.05, 0.1, 0.2)):
        y, sk = o.integrate(y0, T, 1e-3, fixed_h=h)
        assert sk["accepted"] == round(T / h)
        assert rel_inf(Y[k], y) < 1e-12, k


def test_reported_states_follow_the_trajectory():
    spec = _tree(405)
    o = OracleModel(spec)
    q, u = M.random_state(spec, 405, 0, uscale=0.5)
    y0 = np.concatenate([q, u])
    times = [0.013, 0.05, 0.11, 0.3]
    Y, s = o.integrate_stops(y0, times, 1e-8)
    assert s["status"] == 0
    for k, T in enumerate(times):
        y, _ = o.integrate(y0, T, 1e-12)
        Ka = o.kinematics(Y[k][:o.nq], Y[k][o.nq:])
        Kb = o.kinematics(y[:o.nq], y[o.nq:])
        assert np.abs(Ka["X_GB"] - Kb["X_GB"]).max() < 1e-6, k
        assert np.abs(Ka["V_GB"] - Kb["V_GB"]).max() < 1e-6 * max(1, np.abs(Kb["V_GB"]).max()), k


def test_energy_at_every_report_time():
    spec = _tree(407)
    o = OracleModel(spec)
    q, u = M.random_state(spec, 407, 0, uscale=0.3)
    K0 = o.kinematics(q, u)
    E0 = K0["ke"] + K0["pe"]
    times = np.linspace(0.05, 0.5, 10)
    Y, s = o.integrate_stops(np.concatenate([q, u]), times, 1e-10)
    assert s["status"] == 0
    for k in range(len(times)):
        K = o.kinematics(Y[k][:o.nq], Y[k][o.nq:])
        assert abs(K["ke"] + K["pe"] - E0) < 1e-8 * max(abs(E0), K0["ke"]), k


@pytest.mark.parametrize("seed", [409, 411])
def test_report_times_never_save_steps(seed):
    spec = _tree(seed)
    o = OracleModel(spec)
    q, u = M.random_state(spec, seed, 0)
    y0 = np.concatenate([q, u])
    _, s1 = o.integrate(y0, 0.3, 1e-6)
    _, s2 = o.integrate_stops(y0, np.linspace(0.03, 0.3, 10), 1e-6)
    assert s2["accepted"] >= s1["accepted"]


# ---- reports by interpolation (Simbody's default) -------------------------

def test_interpolated_reports_leave_the_trajectory_alone():
    spec = _tree(431)
    o = OracleModel(spec)
    q, u = M.random_state(spec, 431, 0)
    y0 = np.concatenate([q, u])
    y, s = o.integrate(y0, 0.2, 1e-6)
    R, yT, s2 = o.integrate_reports(y0, 0.2, [0.0, 0.013, 0.1, 0.2], 1e-6)
    assert np.array_equal(yT, y) and s == s2
    assert np.array_equal(R[-1], y)                     # a report at the final time is the end state
    yi, _ = o.integrate(y0, 0.0, 1e-6)                  # initialize() only (quaternions normalized)
    assert np.array_equal(R[0], yi)


def test_interpolation_is_exact_for_constant_acceleration():
    """A Slider along x under gravity (-3, -9.8, 0): x = x0 + v0 t - 1.5 t^2.
    RK-Merson and the cubic Hermite interpolant are both exact for a
    quadratic, so every report matches to rounding; a swapped or mis-signed
    Hermite coefficient breaks it."""
    spec = M.random_tree(433, 0, nmin=1, nmax=1)
    spec.mobilizer[:] = M.SLIDER
    spec.X_PF[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]
    spec.gravity = np.array([-3.0, -9.8, 0.0])
    o = OracleModel(spec)
    x0, v0 = 0.3, 0.7
    times = np.array([0.0123, 0.05, 0.0777, 0.19, 0.31])
    R, yT, s = o.integrate_reports(np.array([x0, v0]), 0.4, times, 1e-4)
    assert s["accepted"] < len(times) + 5          # big steps: most reports fall inside one
    for k, t in enumerate(times):
        assert abs(R[k][0] - (x0 + v0 * t - 1.5 * t * t)) < 1e-13, k
        assert abs(R[k][1] - (v0 - 3.0 * t)) < 1e-13, k


def test_interpolated_reports_follow_the_trajectory():
    spec = _tree(435)
    o = OracleModel(spec)
    q, u = M.random_state(spec, 435, 0, uscale=0.5)
    y0 = np.concatenate([q, u])
    times = [0.013, 0.05, 0.11, 0.29]
    R, _, s = o.integrate_reports(y0, 0.3, times, 1e-8)
    assert s["status"] == 0
    for k, T in enumerate(times):
        y, _ = o.integrate(y0, T, 1e-12)
        Ka = o.kinematics(R[k][:o.nq], R[k][o.nq:])
        Kb = o.kinematics(y[:o.nq], y[o.nq:])
        assert np.abs(Ka["X_GB"] - Kb["X_GB"]).max() < 1e-6, k
        assert np.abs(Ka["V_GB"] - Kb["V_GB"]).max() < 1e-6 * max(1, np.abs(Kb["V_GB"]).max()), k


def test_interpolated_quaternions_are_projected():
    """At a loose accuracy the interpolant leaves the unit sphere by more
    than the constraint tolerance (accuracy/10); the reported quaternions
    are back within it."""
    spec = _tree(437, nb=12)
    spec.mobilizer[1::2] = M.BALL
    o = OracleModel(spec)
    q, u = M.random_state(spec, 437, 0, uscale=3.0)
    acc = 1e-2
    times = np.linspace(0.01, 0.5, 25)
    R, _, s = o.integrate_reports(np.concatenate([q, u]), 0.5, times, acc)
    assert s["status"] == 0
    qidx = o.layout()[0]
    qi = [int(qidx[b]) for b in range(spec.nb) if spec.mobilizer[b] in (M.BALL, M.FREE)]
    for k in range(len(times)):
        e = [np.linalg.norm(R[k][i:i + 4]) - 1.0 for i in qi]
        assert np.sqrt(np.mean(np.square(e))) <= acc / 10, k
