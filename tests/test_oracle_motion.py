"""Prescribed motion in the oracle (CPU): Motion::Sinusoid at position,
velocity and acceleration level and Motion::Steady (Motion.h:240-300,
MotionImpl.h:182-300), applied in the integrator as Simbody does
(prescribeQ/prescribeU and the prescribed udot at every stage time,
IntegratorRep.h:901-934; the step's end state prescribed before its error
estimate), pinned by closed forms:
- a Pin prescribed at position level follows a sin(w t + phi) and its rate
  exactly; a Steady Pin turns at its rate;
- a Slider under gravity prescribed at acceleration level integrates to
  the closed-form double integral of the sinusoid;
- the lower link of the double pendulum, driven by an upper link prescribed
  at position or velocity level, follows the closed-form Lagrange equation
  of the driven link (absolute angles t1 prescribed, t2'' = (s t1'^2 -
  g sin t2 - c t1'')/2 with c = cos(t1 - t2), s = sin(t1 - t2)) integrated by
  scipy at 1e-12."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

from simbody_amd import models as M
from oracle_ctypes import OracleModel

G = 9.8


def _motion(nb, rows):
    m = np.zeros((nb + 1, 4))
    for b, r in rows.items():
        m[b] = r
    return m


def test_position_sinusoid_is_followed_exactly():
    o = OracleModel(M.pendulum_chain(1))
    a, w, ph = 0.7, 3.0, 0.4
    for T in (0.3, 1.7):
        y, st = o.integrate_motion(np.array([0.1, -0.5]), T, 1e-8, _motion(1, {1: (1, a, w, ph)}))
        assert st["status"] == 0
        assert y[0] == pytest.approx(a * np.sin(w * T + ph), abs=1e-14)
        assert y[1] == pytest.approx(a * w * np.cos(w * T + ph), abs=1e-13)


def test_steady_velocity():
    o = OracleModel(M.pendulum_chain(1))
    y, st = o.integrate_motion(np.array([0.2, 5.0]), 1.3, 1e-8, _motion(1, {1: (4, 2.5, 0, 0)}))
    assert y[1] == 2.5 and y[0] == pytest.approx(0.2 + 2.5 * 1.3, rel=1e-13)


def test_acceleration_sinusoid_on_a_slider():
    spec = M.TreeSpec(parent=np.array([0], np.int32), mobilizer=np.array([M.SLIDER], np.int32),
                      mass=np.array([1.5]), com_B=np.zeros((1, 3)), inertia_B=np.array([[1.0, 1, 1, 0, 0, 0]]),
                      X_PF=np.array([M._xform()]), X_BM=np.array([M._xform()]), gravity=np.array([-G, 0, 0]))
    o = OracleModel(spec)
    a, w, ph, q0, u0, T = 2.0, 4.0, 0.3, 0.5, -1.0, 1.1
    y, st = o.integrate_motion(np.array([q0, u0]), T, 1e-10, _motion(1, {1: (3, a, w, ph)}))
    # udot = a sin(w t + ph) whatever gravity does: the motion is prescribed
    u = u0 + (a / w) * (np.cos(ph) - np.cos(w * T + ph))
    q = q0 + u0 * T + (a / w) * np.cos(ph) * T - (a / w ** 2) * (np.sin(w * T + ph) - np.sin(ph))
    assert y[1] == pytest.approx(u, abs=1e-8) and y[0] == pytest.approx(q, abs=1e-8)


def _driven_lower_link(t1f, T, q2, u2):
    """Lower link of the double pendulum with the upper link's absolute angle
    t1(t) given (t1f returns t1, t1', t1''); returns (q2, u2) at T."""
    def f(t, z):
        t1, d1, dd1 = t1f(t)
        t2, d2 = z
        c, s = np.cos(t1 - t2), np.sin(t1 - t2)
        return [d2, (s * d1 * d1 - G * np.sin(t2) - c * dd1) / 2]
    t1, d1, _ = t1f(0.0)
    sol = solve_ivp(f, (0, T), [t1 + q2, d1 + u2], rtol=1e-12, atol=1e-13, method="DOP853")
    t1T, d1T, _ = t1f(T)
    return sol.y[0, -1] - t1T, sol.y[1, -1] - d1T


def test_driven_double_pendulum_position_level():
    o = OracleModel(M.pendulum_chain(2))
    a, w, ph, T = 0.8, 2.5, 0.2, 1.5
    t1f = lambda t: (a * np.sin(w * t + ph), a * w * np.cos(w * t + ph), -a * w * w * np.sin(w * t + ph))
    y, st = o.integrate_motion(np.array([0.0, 0.6, 0.0, -0.4]), T, 1e-10, _motion(2, {1: (1, a, w, ph)}))
    q2, u2 = _driven_lower_link(t1f, T, 0.6, -0.4)
    assert y[0] == pytest.approx(t1f(T)[0], abs=1e-14)
    assert y[1] == pytest.approx(q2, abs=1e-7) and y[3] == pytest.approx(u2, abs=1e-7)


def test_driven_double_pendulum_velocity_level():
    o = OracleModel(M.pendulum_chain(2))
    a, w, ph, T, q10 = 1.2, 3.0, -0.5, 1.2, 0.3
    # u1 = a sin(w t + ph) prescribed, q1 integrated from it
    t1f = lambda t: (q10 + (a / w) * (np.cos(ph) - np.cos(w * t + ph)), a * np.sin(w * t + ph),
                     a * w * np.cos(w * t + ph))
    y, st = o.integrate_motion(np.array([q10, -0.2, 9.0, 0.5]), T, 1e-10, _motion(2, {1: (2, a, w, ph)}))
    q2, u2 = _driven_lower_link(t1f, T, -0.2, 0.5)
    assert y[0] == pytest.approx(t1f(T)[0], abs=1e-8) and y[2] == pytest.approx(t1f(T)[1], abs=1e-14)
    assert y[1] == pytest.approx(q2, abs=1e-7) and y[3] == pytest.approx(u2, abs=1e-7)


def test_fixed_step_convergence_is_fourth_order_with_prescribed_motion():
    """Merson is fourth order on the driven double pendulum only if every
    stage sees the prescribed motion at its own time (t0 + h/3, t0 + h/3,
    t0 + h/2, t1): halving a fixed step divides the error by about 16."""
    o = OracleModel(M.pendulum_chain(2))
    a, w, ph, T = 0.8, 2.5, 0.2, 1.0
    t1f = lambda t: (a * np.sin(w * t + ph), a * w * np.cos(w * t + ph), -a * w * w * np.sin(w * t + ph))
    q2, u2 = _driven_lower_link(t1f, T, 0.6, -0.4)
    err = []
    for h in (0.02, 0.01):
        y, st = o.integrate_motion(np.array([0.0, 0.6, 0.0, -0.4]), T, 1.0, _motion(2, {1: (1, a, w, ph)}),
                                   fixed_h=h)
        err.append(max(abs(y[1] - q2), abs(y[3] - u2)))
    assert 12 < err[0] / err[1] < 20, err
