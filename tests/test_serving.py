"""Serving slots (tdyn_server_* through simbody_amd.serving).

Each request carries its own chain model.  Results must be the same alone
and in the mix: every slot runs the same kernel on its own lane, so the
tolerance against the request served alone is exact equality (0 ulp).
Against MatterBatch.stepTo of the same tree (a different kernel build: the
model constants are scalar there, per-lane here, so the compiler may fuse
a*b+c differently) the step counts are identical and states agree to 1e-10
relative; against the oracle the trajectory bar of the rollout parity tests
holds: identical accept/reject counts, states within 1e-8 relative.
"""
import numpy as np
import pytest

from common import rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NB = 12


def _requests(n, seed=41, horizon=(0.02, 0.08), forces=False):
    template = M.spatial_chain(NB, seed=1)
    params = M.chain_params(seed, np.arange(n), nb=NB)
    specs = [M.chain_spec(params, i) for i in range(n)]
    Y = M.random_states(template, seed, np.arange(n))
    rng = M.Rng(seed, 77)
    t_end = rng.uniform(n, *horizon)
    F = np.stack([M.random_body_forces(template, seed, i).reshape(-1) for i in range(n)]) if forces else None
    return template, specs, Y[:, :template.nq].copy(), Y[:, template.nq:].copy(), t_end, F


def _server(template, slots=64, acc=1e-6):
    from simbody_amd.matter import RungeKuttaMerson
    from simbody_amd.serving import Server
    return Server(template, RungeKuttaMerson(accuracy=acc), hbm_budget_bytes=1 << 34, max_slots=slots)


def _run_all(sv, specs, q, u, t_end, F=None, steps=3):
    """Admit everything that fits, advance until all collected."""
    n = len(specs)
    out = {}
    pending = list(range(n))
    slot_req = {}
    while len(out) < n:
        if pending:
            ids = pending[:sv.info()["free"]]
            slots = sv.admit([specs[i] for i in ids], q[ids], u[ids], t_end[ids], None if F is None else F[ids])
            for i, s in zip(ids, slots):
                slot_req[int(s)] = i
            pending = pending[len(slots):]
        sv.advance(steps)
        r = sv.collect()
        for j, s in enumerate(r["slots"]):
            out[slot_req.pop(int(s))] = (r["q"][j].copy(), r["u"][j].copy(), r["t"][j], r["status"][j],
                                         r["stats"][j].copy())
    return out


def test_server_capacity_from_budget_and_topology_check():
    from simbody_amd.matter import RungeKuttaMerson, TdynError
    from simbody_amd.serving import Server
    template, specs, q, u, t_end, _ = _requests(4)
    sv = Server(template, RungeKuttaMerson(accuracy=1e-6), hbm_budget_bytes=130 * 1 << 20, max_slots=0)
    inf = sv.info()
    assert inf["slots"] % 64 == 0 and inf["slots"] * inf["bytes_per_slot"] <= 130 * 1 << 20
    assert inf["slots"] + 64 > (130 * 1 << 20) // inf["bytes_per_slot"]
    with pytest.raises(TdynError):
        Server(template, RungeKuttaMerson(accuracy=1e-6), hbm_budget_bytes=1000)
    bad = M.chain_spec(M.chain_params(3, [0], nb=NB, types=(M.PIN,)), 0, types=(M.PIN,))
    with pytest.raises(TdynError):
        sv.admit([bad], q[:1], u[:1], t_end[:1])
    assert sv.info()["free"] == inf["slots"]
    small = _server(template, slots=64)
    ids = list(range(4)) * 20        # 80 requests, 64 slots
    slots = small.admit([specs[i] for i in ids], q[ids], u[ids], t_end[ids])
    assert len(slots) == 64 and len(set(slots.tolist())) == 64
    assert small.info()["free"] == 0


def test_server_matches_rollout_and_oracle():
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    template, specs, q, u, t_end, F = _requests(8, forces=True)
    sv = _server(template)
    out = _run_all(sv, specs, q, u, t_end, F)
    for i in range(8):
        qs, us, ts, st, stats = out[i]
        assert st == 1 and ts == t_end[i]
        B = MatterBatch([specs[i]], n_instances=1)
        q1, u1, st1 = B.stepTo(B.to_device(q[i]), B.to_device(u[i]), RungeKuttaMerson(accuracy=1e-6), float(t_end[i]),
                               fapp=B.to_device(F[i]))
        assert rel_inf(np.concatenate([qs, us]), np.concatenate([q1.cpu().numpy(), u1.cpu().numpy()])) < 1e-10, i
        assert tuple(stats) == tuple(st1.cpu().numpy()[0]), i
        y, s = OracleModel(specs[i]).integrate(np.concatenate([q[i], u[i]]), float(t_end[i]), 1e-6,
                                               Fapp=F[i].reshape(NB + 1, 6))
        assert (s["accepted"], s["rejected"]) == (stats[1], stats[2]), (i, s, stats)
        assert rel_inf(np.concatenate([qs, us]), y) < 1e-8, i


def test_result_same_alone_and_in_the_mix():
    """Six probe requests served alone, then inside a mix of 150 requests
    admitted in waves while others are running (more requests than slots,
    so probes wait for slots to free up): bitwise the same."""
    template, specs, q, u, t_end, _ = _requests(150, seed=43, horizon=(0.01, 0.1))
    probes = [3, 17, 64, 99, 120, 149]
    alone = {}
    for i in probes:
        sv = _server(template)
        alone[i] = _run_all(sv, [specs[i]], q[[i]], u[[i]], t_end[[i]])[0]
    sv = _server(template, slots=64)
    mix = _run_all(sv, specs, q, u, t_end, steps=2)
    for i in probes:
        a, m = alone[i], mix[i]
        assert np.array_equal(a[0], m[0]) and np.array_equal(a[1], m[1]), i
        assert a[2] == m[2] and a[3] == m[3] and tuple(a[4]) == tuple(m[4]), i


def test_serve_loop_poisson():
    """The request loop on Poisson arrivals: every request completes, and
    requests built from chain_params rows match the per-spec path exactly."""
    from simbody_amd.serving import Requests, serve
    n = 200
    template = M.spatial_chain(NB, seed=1)
    params = M.chain_params(44, np.arange(n), nb=NB)
    Y = M.random_states(template, 44, np.arange(n))
    t_end = M.Rng(44, 77).uniform(n, 0.01, 0.05)
    req = Requests.from_chain_params(template, params, Y[:, :template.nq], Y[:, template.nq:], t_end)
    sv = _server(template, slots=128)
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1 / 2000.0, size=n))
    r = serve(sv, req, arrivals, steps_per_advance=4, time_limit=60)
    assert r["completed"] == n
    assert (r["status"] == 1).all()
    assert np.isfinite(r["latency"]).all() and (r["latency"] > 0).all()
    assert sv.info()["free"] == sv.slots
    specs = [M.chain_spec(params, i) for i in range(n)]
    ref = _run_all(_server(template, slots=64), specs[:20], Y[:20, :template.nq], Y[:20, template.nq:], t_end[:20])
    for i in range(20):
        assert np.array_equal(ref[i][0], r["q"][i]) and np.array_equal(ref[i][1], r["u"][i]), i
