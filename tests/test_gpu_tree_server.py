"""Serving requests of different trees (tdyn_tree_server_* through
simbody_amd.serving.TreeServer).

Requests of different sizes, topologies and mobilizers (serial chains,
branched trees with Free roots, trees mixing the ten mobilizers, reversed
mobilizers, Free bodies below body 1) share one server.  Each slot runs on
its own pooled memory, on the level-parallel evaluator with lanes fixed by
its own tree or (a Free body below body 1, or lanes 1) on the one-lane tree
sweeps, so a request's result is the same alone and in the mix, to the bit;
against the oracle the trajectory bar of the rollout parity tests holds
(identical accepted and rejected counts, states within 1e-8 relative) under
every lane setting.  Memory is the request's own: a larger tree costs more
bytes, the pools admit requests until they are full, and give the memory
back when the requests are collected."""
import numpy as np
import pytest

from common import ext_trees, mixed_trees, rel_inf, rev_trees
from simbody_amd import models as M
from oracle_ctypes import OracleModel
from test_oracle_free_offground import free_mid_trees

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ACC = 1e-6


def _requests(n, seed, free_mid=False):
    """A mix of trees: 12-body chains, random trees with Free roots, trees of
    the ten mobilizers, trees with reversed mobilizers; sizes 4 to 40; with
    free_mid, also trees with Free bodies below body 1 (tree sweeps)."""
    specs = []
    gens = [lambda i: M.chain_spec(M.chain_params(seed, [i], nb=12), 0),
            lambda i: mixed_trees(seed * 3 + i, 1, nmin=4, nmax=40)[0],
            lambda i: ext_trees(seed * 5 + i, 1, nmin=4, nmax=24)[0],
            lambda i: rev_trees(seed * 7 + i, 1, nmin=4, nmax=24)[0]]
    if free_mid:
        gens.append(lambda i: free_mid_trees(seed * 11 + i, 1, nmin=4, nmax=24)[0])
    for i in range(n):
        specs.append(gens[i % len(gens)](i))
    ys = [M.random_state(t, seed, i) for i, t in enumerate(specs)]
    F = [M.random_body_forces(t, seed, i) for i, t in enumerate(specs)]
    t_end = M.Rng(seed, 5).uniform(n, 0.02, 0.08)
    return specs, [y[0] for y in ys], [y[1] for y in ys], t_end, F


def _server(slots=64, budget=1 << 28, lanes=0):
    from simbody_amd.matter import RungeKuttaMerson
    from simbody_amd.serving import TreeServer
    return TreeServer(RungeKuttaMerson(accuracy=ACC), budget, max_slots=slots, lanes=lanes)


def _run_all(sv, specs, q, u, t_end, F=None, steps=3):
    n = len(specs)
    out, pending, slot_req = {}, list(range(n)), {}
    guard = 0
    while len(out) < n:
        guard += 1
        assert guard < 100000
        if pending:
            ids = pending[:sv.info()["free"]]
            slots = sv.admit([specs[i] for i in ids], [q[i] for i in ids], [u[i] for i in ids], t_end[ids],
                             None if F is None else [F[i] for i in ids])
            for i, s in zip(ids, slots):
                slot_req[int(s)] = i
            pending = pending[len(slots):]
        sv.advance(steps)
        r = sv.collect()
        for j, s in enumerate(r["slots"]):
            out[slot_req.pop(int(s))] = (r["y"][j], r["t"][j], r["status"][j], r["stats"][j].copy())
    return out


@pytest.mark.parametrize("lanes", [0, 1, 4, 32])
def test_mixed_trees_match_the_oracle(lanes):
    """Every lane setting: automatic (level-parallel, lanes from each tree's
    widest level), 1 (every request on the tree sweeps), 4 and 32 lanes; the
    mix includes trees with Free bodies below body 1, which take the tree
    sweeps under every setting."""
    specs, q, u, t_end, F = _requests(20, 81, free_mid=True)
    out = _run_all(_server(lanes=lanes), specs, q, u, t_end, F)
    for i, t in enumerate(specs):
        y, ts, st, stats = out[i]
        assert st == 1 and ts == t_end[i], (i, st, ts)
        yo, s = OracleModel(t).integrate(np.concatenate([q[i], u[i]]), float(t_end[i]), ACC, Fapp=F[i])
        assert (s["accepted"], s["rejected"]) == (stats[1], stats[2]), (i, s, stats)
        assert rel_inf(y, yo) < 1e-8, (i, t.nb)


def test_result_same_alone_and_in_the_mix():
    """Probes served alone, then in a mix of 90 requests of different trees
    admitted in waves into 32 slots (probes wait for slots and pool memory):
    bitwise the same."""
    specs, q, u, t_end, F = _requests(90, 83, free_mid=True)
    probes = [1, 6, 12, 23, 46, 69, 70, 89]
    alone = {i: _run_all(_server(slots=64), [specs[i]], [q[i]], [u[i]], t_end[[i]], [F[i]])[0] for i in probes}
    mix = _run_all(_server(slots=32, budget=24 << 20), specs, q, u, t_end, F, steps=2)
    for i in probes:
        a, m = alone[i], mix[i]
        assert np.array_equal(a[0], m[0]), i
        assert a[1] == m[1] and a[2] == m[2] and tuple(a[3]) == tuple(m[3]), i


def test_memory_is_each_requests_own():
    from simbody_amd.matter import TdynError
    from simbody_amd.serving import TreeServer
    small = M.chain_spec(M.chain_params(3, [0], nb=8), 0)
    big = mixed_trees(5, 1, nmin=120, nmax=120)[0]
    cs, cb = TreeServer.request_bytes(small), TreeServer.request_bytes(big)
    assert cb > 10 * cs
    sv = _server(slots=256, budget=12 << 20)
    inf0 = sv.info()
    qs, us = M.random_state(small, 1, 0)
    qb, ub = M.random_state(big, 1, 0)
    # fill with big requests until the pools refuse, then small ones still fit
    nb_big = len(sv.admit([big] * 64, [qb] * 64, [ub] * 64, np.full(64, 0.01)))
    assert 0 < nb_big < 64
    n_small = len(sv.admit([small] * 64, [qs] * 64, [us] * 64, np.full(64, 0.01)))
    assert n_small > 0
    inf1 = sv.info()
    assert inf1["occupied"] == nb_big + n_small and inf1["bytes_in_use"] <= inf1["budget_bytes"]
    done = 0
    while done < nb_big + n_small:
        sv.advance(50)
        done += len(sv.collect()["slots"])
    inf2 = sv.info()
    assert inf2["occupied"] == 0 and inf2["bytes_in_use"] == inf0["bytes_in_use"]
    # a request larger than the whole pools is refused outright
    huge = M.spatial_chain(1000, seed=3)          # about 1 MB of records alone
    tiny = _server(slots=64, budget=(64 * 160) + (1 << 20) + (1 << 17))
    qh, uh = M.random_state(huge, 1, 0)
    with pytest.raises(TdynError):
        tiny.admit([huge], [qh], [uh], np.array([0.01]))
