"""Oracle pins for the integrator options Integrator::setUseInfinityNorm,
setProjectEveryStep and setProjectInterpolatedStates (IntegratorRep.h:
454-488, 708-712, 770-875; SimbodyMatterSubsystemRep.cpp:4108-4184,
4676-4720; AbstractIntegratorRep.cpp:63-126), by what the mathematics fixes:

- with one q and one u the infinity norm and the RMS norm are the same
  number, so the trajectories are identical;
- two decoupled pendulums, the second at rest at the bottom (no error) or an
  exact copy of the first: the infinity norm is the first pendulum's own
  error, so the step sequence is the first pendulum's alone; the RMS norm of
  (e, 0) is e/sqrt(2) and changes it;
- on one attempt (same state, same step) RMS <= inf <= sqrt(n) RMS;
- forced projection normalizes the quaternions (or takes a Newton step) on
  every attempt whose error passes the projection gate, and never when
  there are no quaternions and no constraints;
- unprojected report states are the projected ones before normalization;
- the state backed up to an event is projected whatever the option, the
  interpolated state a terminating handler leaves unchanged is not.

A planted mistake (a sum or an unsigned max for the infinity norm, a square
root of the max, the force flag ignored, the option applied to the back-up)
fails one of these."""
import numpy as np
import pytest

from simbody_amd import models as M
from oracle_ctypes import OracleModel
import oracle_ctypes as O

G = 9.81


def _pendulum(m=2.0, L=0.7, Ic=1e-2):
    I = m * L * L + Ic
    return dict(mobilizer=M.PIN, mass=m, com=[L, 0.0, 0.0], inertia=[Ic, I, I, 0, 0, 0])


def _spec(bodies, parents, gravity=(0.0, -G, 0.0)):
    n = len(bodies)
    return M.TreeSpec(parent=np.array(parents, np.int32), mobilizer=np.array([b["mobilizer"] for b in bodies], np.int32),
                      mass=np.array([b["mass"] for b in bodies]), com_B=np.array([b["com"] for b in bodies]),
                      inertia_B=np.array([b["inertia"] for b in bodies]),
                      X_PF=np.array([M._xform()] * n), X_BM=np.array([M._xform()] * n), gravity=np.array(gravity))


def _run(spec, y, T, acc, **opts):
    o = OracleModel(spec)
    o.set_integrator_options(**opts)
    return o.integrate(y, T, acc)


def test_infinity_norm_equals_rms_with_one_coordinate():
    spec = _spec([_pendulum()], [0])
    y = np.array([0.3, 1.7])
    ya, sa = _run(spec, y, 1.5, 1e-6)
    yb, sb = _run(spec, y, 1.5, 1e-6, inf_norm=True)
    assert sa == sb and np.array_equal(ya, yb)


@pytest.mark.parametrize("second", ["at_rest", "copy"])
def test_infinity_norm_is_the_largest_component(second):
    """{A, B}: B at rest at the bottom (zero error) or an exact copy of A;
    under the infinity norm the pair steps exactly as A alone; under the RMS
    norm the at-rest pair does not (its norm is A's over sqrt(2))."""
    A = _pendulum()
    one = _spec([A], [0])
    two = _spec([A, A], [0, 0])
    yA = np.array([0.3, 1.7])
    yB = np.array([-np.pi / 2, 0.0]) if second == "at_rest" else yA.copy()
    y2 = np.array([yA[0], yB[0], yA[1], yB[1]])
    ya, sa = _run(one, yA, 1.5, 1e-6, inf_norm=True)
    y2i, s2i = _run(two, y2, 1.5, 1e-6, inf_norm=True)
    assert (s2i["accepted"], s2i["rejected"]) == (sa["accepted"], sa["rejected"])
    assert np.max(np.abs(y2i[[0, 2]] - ya)) < 1e-12
    if second == "at_rest":
        _, s2r = _run(two, y2, 1.5, 1e-6)
        _, sar = _run(one, yA, 1.5, 1e-6)
        assert (s2r["accepted"], s2r["rejected"]) != (sar["accepted"], sar["rejected"])


def test_infinity_norm_bounds_on_one_attempt():
    """The first attempt (same state, same step) on the 10-link pendulum:
    RMS <= inf <= sqrt(n) RMS for the error norm, n = max(nq, nu)."""
    spec = M.pendulum_chain(10)
    q, u = M.random_state(spec, 5, 0)
    y = np.concatenate([q, u])
    tr = {}
    for inf in (False, True):
        o = OracleModel(spec)
        o.set_integrator_options(inf_norm=inf)
        _, _, t = o.integrate_trace(y, 0.05, 1e-6)
        tr[inf] = t
    assert tr[False].t1[0] == tr[True].t1[0]
    e_rms, e_inf = tr[False].err_pre[0], tr[True].err_pre[0]
    n = max(spec.nq, spec.nu)
    assert e_rms < e_inf <= np.sqrt(n) * e_rms * (1 + 1e-12)


def _free_ball():
    I = [1.0, 2.0, 3.0, 0.0, 0.0, 0.0]
    return _spec([dict(mobilizer=M.BALL, mass=1.0, com=[0.0, 0.0, 0.0], inertia=I)], [0], gravity=(0.0, 0.0, 0.0))


def _ball_state():
    q = np.array([0.9, 0.1, -0.3, 0.2]); q /= np.linalg.norm(q)
    return np.concatenate([q, [3.0, 4.0, 2.0]])


@pytest.mark.parametrize("acc", [1e-4, 1e-7])
def test_project_every_step(acc):
    """A torque-free Ball: forced, every attempt that passes the projection
    gate (error <= 16 acc) normalizes the quaternion; the final quaternion is
    unit to rounding.  Unforced, only those whose quaternion error exceeds
    the tolerance.  No quaternions, no constraints: nothing to project."""
    spec, y = _free_ball(), _ball_state()
    res = {}
    for force in (False, True):
        o = OracleModel(spec)
        o.set_integrator_options(project_every_step=force)
        yT, st, tr = o.integrate_trace(y, 1.0, acc)
        res[force] = (yT, st, tr)
    yT, st, tr = res[True]
    gate = np.sum(tr.err_pre <= 16 * acc)
    assert st["projections"] == gate and gate > 0
    assert abs(np.linalg.norm(yT[:4]) - 1.0) < 4e-16
    _, st0, tr0 = res[False]
    assert st0["projections"] == np.sum((tr0.err_pre <= 16 * acc) & (tr0.qerr > acc / 10))
    assert st0["projections"] < st["projections"]
    p = _spec([_pendulum()], [0])
    _, sp = _run(p, np.array([0.3, 1.7]), 1.0, acc, project_every_step=True)
    assert sp["projections"] == 0


def test_project_every_step_with_a_constraint():
    """The constrained point pendulum: forced, every attempt through the gate
    takes a Newton step (a projection) even within the tolerance."""
    spec = _spec([dict(mobilizer=M.TRANSLATION, mass=10.0, com=[0.0, 0.0, 0.0], inertia=[1e-3, 1e-3, 1e-3, 0, 0, 0])],
                 [0])
    cons = [(1, 0, 1, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1.0)]
    y = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    st = {}
    for force in (False, True):
        o = OracleModel(spec)
        assert o.set_constraints(cons) == 0
        o.set_integrator_options(project_every_step=force)
        yT, s, tr = o.integrate_trace(y, 1.0, 1e-6)
        st[force] = (s, tr, yT)
    s, tr, yT = st[True]
    assert s["projections"] == np.sum(tr.err_pre <= 16e-6)
    assert st[False][0]["projections"] < s["projections"]
    assert abs(np.linalg.norm(yT[:3]) - 1.0) < 1e-7


def test_unprojected_report_states():
    """Reports with setProjectInterpolatedStates(false) are the Hermite
    states before the quaternion normalization the default applies past the
    tolerance, so normalizing them where |q| - 1 exceeds it gives the default
    reports; the trajectory itself is unchanged."""
    spec, y = _free_ball(), _ball_state()
    reports = np.linspace(0.013, 0.987, 61)
    acc = 1e-3
    out = {}
    for proj in (True, False):
        o = OracleModel(spec)
        o.set_integrator_options(project_interpolated=proj)
        out[proj] = o.integrate_reports(y, 1.0, reports, acc)
    (rt, yt, st), (rf, yf, sf) = out[True], out[False]
    assert st == sf and np.array_equal(yt, yf)
    qn = np.linalg.norm(rf[:, :4], axis=1)
    past = np.abs(qn - 1.0) > acc / 10
    assert past.any() and (~past).any()
    expect = rf.copy()
    expect[past, :4] /= qn[past, None]
    assert np.max(np.abs(rt - expect)) < 1e-15


def test_event_back_up_is_projected_without_interpolated_projection():
    """A torque-free Ball and a time witness that terminates at a time where
    the interpolated quaternion is off unit length by more than twice the
    tolerance (the localization windows then do not depend on the option).
    The advanced state backed up to the window's end is projected whatever
    the option (a handler that scales the velocity leaves it as the final
    state: both runs agree to the bit, on unit length within the tolerance);
    a handler that changes nothing leaves the state interpolated at the
    window's start between the step's start and that projected end, the
    same in both runs."""
    spec, y = _free_ball(), _ball_state()
    acc = 1e-3
    tol = acc / 10
    grid = np.linspace(0.013, 0.987, 61)
    o = OracleModel(spec)
    o.set_integrator_options(project_interpolated=False)
    rf, _, _ = o.integrate_reports(y, 1.0, grid, acc)
    dev = np.abs(np.linalg.norm(rf[:, :4], axis=1) - 1.0)
    t_ev = float(grid[np.argmax(dev > 2 * tol)])
    assert dev.max() > 2 * tol

    def run(proj, scale):
        o = OracleModel(spec)
        o.set_integrator_options(project_interpolated=proj)
        ws = [dict(kind=O.WIT_TIME, offset=t_ev, action=O.EV_TERMINATE)]
        if scale:
            ws.append(dict(kind=O.WIT_TIME, offset=t_ev, action=O.EV_SCALE_U, factor=0.5))
        assert o.set_witnesses(ws) == 0
        yT, _ = o.integrate(y, 1.0, acc)
        win, _, tf, term = o.last_events(8)
        assert term and len(win) >= 1
        return yT, tf

    (a, ta), (b, tb) = run(False, False), run(True, False)
    assert ta == tb and np.array_equal(a, b)
    (c, tc), (d, td) = run(False, True), run(True, True)
    assert tc == td and np.array_equal(c, d)
    assert abs(np.linalg.norm(c[:4]) - 1.0) <= tol
