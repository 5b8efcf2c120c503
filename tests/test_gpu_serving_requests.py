"""Serving the chain workload's requests as they come (GPU): each request a
32-body chain with its own constants drawn from the chain workload's
distributions, a seeded initial state, horizon U(0.1, 2) s, RK-Merson
accuracy 1e-6, in both precisions; and mixed-precision adaptive rollouts of
the stiff chains.

Stated bounds:
- FP64 slots against the oracle: identical accepted and rejected counts and
  states within max(1e-8, 10 d) relative, d being how far the oracle built
  with fused multiply-adds ends from the plain oracle on the same request
  (chaotic chains amplify rounding over 2 s: d reaches 2e-7 on some
  requests, so no rounding of the algorithm can be held to 1e-8 there);
- mixed-precision slots (FP32 kinematics, bias forces and articulated
  inertias): every request finishes at its final time; accepted steps within
  5 % of the oracle's; states within 1e-5 relative of the FP64 oracle for
  horizons up to 0.35 s (longer chaotic horizons diverge under any
  perturbation); energy drift |E(T) - E(0)|/|E(0)| within 10 x the oracle's
  drift + 2e-5;
- in both precisions a request's result is the same alone and in the mix,
  to the bit;
- mixed adaptive rollouts of the stiff chains (accuracy 1e-6, T = 0.5 s):
  every instance finishes; median energy drift <= 1e-5 and max <= 2e-3;
  median state difference from the FP64 path <= 1e-2; at most 3 x the FP64
  path's accepted steps (FP32 rounding reaches the error target, so the
  mixed path takes more, smaller steps)."""
import numpy as np
import pytest

from common import rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ACC = 1e-6


def _requests(n=24, seed=51):
    template = M.spatial_chain(32, seed=2)
    params = M.chain_params(seed, np.arange(n), nb=32)
    specs = [M.chain_spec(params, i) for i in range(n)]
    Y = M.random_states(template, seed, np.arange(n))
    t_end = M.Rng(seed, 91).uniform(n, 0.1, 2.0)
    return template, specs, Y[:, :template.nq].copy(), Y[:, template.nq:].copy(), t_end


def _serve(template, specs, q, u, t_end, prec, slots=64):
    from simbody_amd.matter import RungeKuttaMerson
    from simbody_amd.serving import Server
    from test_serving import _run_all
    sv = Server(template, RungeKuttaMerson(accuracy=ACC, precision=prec), hbm_budget_bytes=1 << 34, max_slots=slots)
    return _run_all(sv, specs, q, u, t_end, steps=16)


def _energy(o, nq, y):
    k = o.kinematics(y[:nq], y[nq:])
    return k["ke"] + k["pe"]


@pytest.mark.parametrize("prec", [0, 1])
def test_chain_requests_against_the_oracle(prec):
    template, specs, q, u, t_end = _requests()
    out = _serve(template, specs, q, u, t_end, prec)
    nq = template.nq
    for i, t in enumerate(specs):
        qs, us, ts, st, stats = out[i]
        assert st == 1 and ts == t_end[i], (i, st, ts)
        y0 = np.concatenate([q[i], u[i]])
        o = OracleModel(t)
        yo, so = o.integrate(y0, float(t_end[i]), ACC)
        y = np.concatenate([qs, us])
        if prec == 0:
            yf, _ = OracleModel(t, variant="fma").integrate(y0, float(t_end[i]), ACC)
            assert (so["accepted"], so["rejected"]) == (stats[1], stats[2]), (i, so, stats)
            assert rel_inf(y, yo) <= max(1e-8, 10 * rel_inf(yf, yo)), (i, t_end[i], rel_inf(y, yo))
        else:
            assert abs(stats[1] - so["accepted"]) <= 0.05 * so["accepted"] + 1, (i, so, stats)
            if t_end[i] <= 0.35:
                assert rel_inf(y, yo) < 1e-5, (i, rel_inf(y, yo))
            E0 = _energy(o, nq, y0)
            d_mixed = abs(_energy(o, nq, y) - E0) / abs(E0)
            d_oracle = abs(_energy(o, nq, yo) - E0) / abs(E0)
            assert d_mixed <= 10 * d_oracle + 2e-5, (i, d_mixed, d_oracle)


@pytest.mark.parametrize("prec", [0, 1])
def test_chain_request_same_alone_and_in_the_mix(prec):
    template, specs, q, u, t_end = _requests(n=100, seed=53)
    mix = _serve(template, specs, q, u, t_end, prec, slots=64)
    for i in (2, 37, 71, 99):
        a = _serve(template, [specs[i]], q[[i]], u[[i]], t_end[[i]], prec)[0]
        m = mix[i]
        assert np.array_equal(a[0], m[0]) and np.array_equal(a[1], m[1]), (prec, i)
        assert a[2] == m[2] and a[3] == m[3] and tuple(a[4]) == tuple(m[4]), (prec, i)


def test_mixed_adaptive_rollouts_of_the_stiff_chains():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    spec = M.stiff_chain(32, seed=2)
    n = 64
    Y = M.chain_batch_states(spec, n, seed=33)
    outs = {}
    for prec in (0, 1):
        B = MatterBatch([spec], n_instances=n)
        B.rollout_begin(B.to_device(Y[:, :spec.nq]), B.to_device(Y[:, spec.nq:]),
                        RungeKuttaMerson(accuracy=ACC, precision=prec), t_end=0.5)
        B.rollout_advance(1 << 20)
        qf, uf, tf, st, stats = B.rollout_get()
        assert (st.cpu().numpy() == 1).all(), prec
        outs[prec] = (np.concatenate([qf.cpu().numpy().reshape(n, -1), uf.cpu().numpy().reshape(n, -1)], 1),
                      stats.cpu().numpy())
    o = OracleModel(spec)
    drift, diff = [], []
    for i in range(n):
        E0 = _energy(o, spec.nq, Y[i])
        drift.append(abs(_energy(o, spec.nq, outs[1][0][i]) - E0) / abs(E0))
        diff.append(rel_inf(outs[1][0][i], outs[0][0][i]))
    print(f"stiff mixed: drift median {np.median(drift):.2e} max {max(drift):.2e}; "
          f"state diff median {np.median(diff):.2e}")
    assert np.median(drift) <= 1e-5 and max(drift) <= 2e-3
    assert np.median(diff) <= 1e-2
    assert outs[1][1][:, 1].sum() <= 3 * outs[0][1][:, 1].sum()
