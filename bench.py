"""Benchmark: body-steps/s of batched RK-Merson rollouts on one MI355X.

Workload (the headline configuration): 65,536 identical spatial serial chains
of 32 bodies (mobilizers cycling Pin/Universal/Ball), FP64, Merson fixed step
h = 1e-3 (Simbody's RungeKuttaMersonIntegrator with setFixedStepSize; Simbody
has no classic RK4).  One "step" = one accepted integrator step of every
instance; a body-step = one body advanced by one accepted step.  Inverse
dynamics (calcResidualForceIgnoringConstraints) on the final states is timed
as a separate figure.  Inputs are synthetic (seeded), resident in HBM before
the timed region.

python bench.py [--steps K] [--warmup W] [--instances N] [--no-cpu]
prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP64 flops per body-step of this workload (+,-,*,/ of the
# restated Simbody algorithm, counted by the oracle's counting scalar over
# fixed steps of the 32-body chain; tests/test_flop_count.py re-derives it).
FLOPS_PER_BODY_STEP = 9034.59375
FLOPS_RESID_PER_BODY = 1055.40625
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 (vector = matrix) peak
PEAK_HBM_GBS = 8000.0
# HBM bytes per body-step of k_rkm_chain on this workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE per dispatch (difference of a 10-step and a 1-step
# launch, over 9 steps x 2,097,152 bodies), each scaled by what the same
# counter reads for a plain 2 GiB device copy (FETCH x2.0, WRITE x1.0).
TRAFFIC_FETCH_PER_BODY_STEP = 1592.9
TRAFFIC_WRITE_PER_BODY_STEP = 1040.3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_threads():
    """Worker threads of the CPU baselines: every CPU this process may run on
    (its affinity mask), capped at 16, the host-CPU share of one GPU on the
    machines this runs on (more threads would oversubscribe that share)."""
    try:
        n_cpu = len(os.sched_getaffinity(0))
    except AttributeError:
        n_cpu = os.cpu_count() or 1
    return max(1, min(16, n_cpu))


def cpu_baseline(spec, Y, steps, h, budget_s):
    """The oracle (scalar FP64 C++ restatement) on host cores, one instance
    per pinned thread, on a bounded sample of the same workload: one run on
    one thread (the per-core rate) and one on every thread of the pool."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_ctypes import OracleModel, time_integrate_batch, host_cpu_info
    om = OracleModel(spec)
    threads = _cpu_threads()
    host = host_cpu_info()
    # calibrate with one short run per thread
    probe_steps = 20
    Yp = np.ascontiguousarray(Y[:threads].copy())
    sec, bs, _ = time_integrate_batch([om], np.zeros(threads, np.int32), Yp, probe_steps * h, 1e-3, h, threads)
    per_inst_full = sec * steps / probe_steps          # seconds per instance for `steps` steps, one thread
    # one thread: a quarter of the budget
    n1 = int(max(1, min(len(Y), int(0.25 * budget_s / max(per_inst_full, 1e-9)))))
    sec1, bs1, _, cpu1 = time_integrate_batch([om], np.zeros(n1, np.int32), np.ascontiguousarray(Y[:n1].copy()),
                                              steps * h, 1e-3, h, 1, cpu_time=True)
    n_inst = int(max(threads, min(len(Y), threads * max(1, int(0.75 * budget_s / max(per_inst_full, 1e-9))))))
    Ys = np.ascontiguousarray(Y[:n_inst].copy())
    sec, bs, _, cpus = time_integrate_batch([om], np.zeros(n_inst, np.int32), Ys, steps * h, 1e-3, h, threads,
                                            cpu_time=True)
    per_core = bs1 / sec1
    return {"value": bs / sec, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n_inst} chains x {steps} fixed steps (of the same workload), {threads} pinned threads, "
                      f"{sec:.2f} s wall, {cpus:.2f} s thread CPU",
            "per_core_value": per_core,
            "per_core_sample": f"{n1} chains x {steps} fixed steps, 1 pinned thread, {sec1:.2f} s wall, "
                               f"{cpu1:.2f} s thread CPU",
            "parallel_efficiency": bs / sec / (per_core * threads),
            "host": host,
            "all_online_cores_estimate": {
                "value": per_core * (host.get("nproc_online") or threads),
                "note": "per-core rate x online CPUs: an extrapolation, not a measurement; the pool is "
                        "capped at 16 threads, the host-CPU share of one GPU"}}


def row_pendulum(torch, M, MatterBatch, RungeKuttaMerson, cpu=True):
    """10-link planar Pin pendulum, one instance: realize(Acceleration), then
    RK-Merson accuracy 1e-8 to t = 1 s.  Latency-bound: wall times."""
    spec = M.pendulum_chain(10)
    rng = np.random.default_rng(20261202)
    q0 = rng.uniform(-np.pi, np.pi, 10); u0 = rng.normal(size=10)
    B = MatterBatch([spec], n_instances=1, device="cuda:0")
    q = B.to_device(q0); u = B.to_device(u0)
    integ = RungeKuttaMerson(accuracy=1e-8)
    B.realizeAcceleration(q, u); B.stepTo(q, u, integ, 1.0)       # warm (module load, allocation)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ud, _ = B.realizeAcceleration(q, u)
    torch.cuda.synchronize()
    t_acc = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, _, st = B.stepTo(q, u, integ, 1.0)
    torch.cuda.synchronize()
    t_roll = time.perf_counter() - t0
    st = st.cpu().numpy()[0]
    out = {"workload": "1 x 10-link planar Pin pendulum, FP64, RK-Merson accuracy 1e-8 to t=1 s",
           "gpu_realize_acceleration_ms": t_acc * 1e3, "gpu_rollout_ms": t_roll * 1e3,
           "accepted_steps": int(st[1]), "rejected_steps": int(st[2]),
           "gpu_body_steps_per_s": 10 * int(st[1]) / t_roll}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle_ctypes import OracleModel, time_integrate_batch
        om = OracleModel(spec)
        Y = np.concatenate([q0, u0])[None, :].copy()
        sec, bs, _ = time_integrate_batch([om], np.zeros(1, np.int32), Y, 1.0, 1e-8, 0.0, 1)
        out["cpu_rollout_ms"] = sec * 1e3
        out["cpu_body_steps_per_s"] = bs / sec
        out["cpu_cores"] = 1
    return out


def _evaluator_name(B):
    ev, lanes = B.active_evaluator()
    name = {1: "tree sweeps", 2: "chain", 3: "streaming tree evaluator",
            4: "level-parallel tree evaluator"}.get(ev, "?")
    return f"{name}, " + (f"{lanes} lane(s) per instance" if lanes else "lanes per instance from each tree")


def row_trees(torch, M, MatterBatch, RungeKuttaMerson, n=16384, cpu=True, cpu_budget=6.0, lanes=0, lane_range=(0, 32),
              evaluator=0, autotune=True):
    """Uneven branched trees (nb ~ U{8..128}, depth <= 20, Pin/Slider/
    Universal/Ball + Free roots), constant N(0,1) body forces and torques,
    per-instance RK-Merson step control at accuracy 1e-6 to t = 0.5 s, in
    one launch."""
    log(f"[bench] trees: generating {n}")
    trees = [M.random_tree(3, i) for i in range(n)]
    Ys = [M.random_states(t, 3, [i])[0] for i, t in enumerate(trees)]
    Fs = [M.random_body_forces(t, 3, i) for i, t in enumerate(trees)]
    nbs = np.array([t.nb for t in trees])
    B = MatterBatch(trees, tree_of=np.arange(n), device="cuda:0")
    B.set_evaluator(evaluator)
    B.set_tree_lanes(lanes)
    B.set_tree_lane_range(*lane_range)
    q = B.pack([y[:t.nq] for y, t in zip(Ys, trees)]); u = B.pack([y[t.nq:] for y, t in zip(Ys, trees)])
    F = B.pack(Fs)
    integ = RungeKuttaMerson(accuracy=1e-6)
    tuned = None
    if autotune:
        # probe rollouts of this state to t = 0.05 s pick the lanes per instance
        log("[bench] trees: autotune")
        tuned = B.autotune(q, u, integ, 0.05, fapp=F)
    torch.cuda.synchronize()
    log("[bench] trees: rollout")
    t0 = time.perf_counter()
    _, _, st = B.stepTo(q, u, integ, 0.5, fapp=F)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    st = st.cpu().numpy()
    body_steps = int((nbs * st[:, 1]).sum())
    out = {"workload": f"{n} uneven trees, nb~U{{8..128}} (mean {nbs.mean():.1f}), FP64, RK-Merson "
                       "accuracy 1e-6 to t=0.5 s, N(0,1) body forces",
           "value": body_steps / sec, "unit": "body-steps/s", "seconds": sec,
           "body_steps": body_steps, "accepted_mean": float(st[:, 1].mean()),
           "rejected_total": int(st[:, 2].sum()),
           "evaluator": _evaluator_name(B)}
    if tuned is not None:
        out["autotuned"] = {"mode": tuned[0], "lanes": tuned[1], "lanes_range": [tuned[2], tuned[3]],
                            "note": "tdyn_batch_autotune, probe rollouts to t=0.05 s outside the timed region"}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle_ctypes import OracleModel, time_integrate_batch
        threads = _cpu_threads()
        # per-instance cost scales with nb x steps: take instances in order until the budget estimate
        k = threads
        while True:
            oms = [OracleModel(t) for t in trees[:k]]
            ny = max(t.nq + t.nu for t in trees[:k]); nbm = max(t.nb for t in trees[:k])
            Y = np.zeros((k, ny)); Fa = np.zeros((k, 6 * (nbm + 1)))
            for i in range(k):
                Y[i, :len(Ys[i])] = Ys[i]; Fa[i, :Fs[i].size] = Fs[i].reshape(-1)
            sec_c, bs, _ = time_integrate_batch(oms, np.arange(k, dtype=np.int32), Y, 0.5, 1e-6, 0.0, threads, Fapp=Fa)
            if sec_c > cpu_budget / 4 or k >= n:
                break
            k = min(n, int(k * max(2.0, cpu_budget / max(sec_c, 1e-3) * 0.8)))
        out["cpu_baseline"] = {"value": bs / sec_c, "unit": "body-steps/s", "cores": threads, "kind": "port",
                               "sample": f"first {k} of the {n} trees, {threads} threads, {sec_c:.2f} s"}
    return out


def row_serving(torch, M, RungeKuttaMerson, rates=(1000, 8000, 32000), mixed_rates=(8000,), window=10.0,
                warmup=2.0, seed=5, budget_gib=24, steps_per_advance=16, n_models=8192):
    """Serving loop, steady state: Poisson arrivals for `window` seconds at
    each rate, the first `warmup` seconds discarded; every request is a
    32-body chain drawn from the chain workload's distributions (own masses,
    inertias and offsets; n_models distinct draws reused cyclically), a
    seeded initial state, horizon ~ U(0.1, 2) s, RK-Merson accuracy 1e-6.
    Slots come from an HBM budget; latency = arrival -> result on the host.
    Requests still running at window + 3 s count as not completed."""
    from simbody_amd.serving import Requests, Server, serve
    template = M.spatial_chain(32, seed=2)
    params = M.chain_params(seed, np.arange(n_models), nb=32)
    out = {"workload": f"Poisson arrivals for {window:g} s (first {warmup:g} s discarded); each request one "
                       "32-body chain with its own constants, horizon U(0.1,2) s, RK-Merson accuracy 1e-6; "
                       "FP64 and mixed precision",
           "hbm_budget_bytes": budget_gib << 30, "steps_per_advance": steps_per_advance, "sweep": []}
    for prec, rate in [(0, r) for r in rates] + [(1, r) for r in mixed_rates]:
        sv = Server(template, RungeKuttaMerson(accuracy=1e-6, precision=prec), budget_gib << 30, max_slots=65536)
        out["slots_fp64" if prec == 0 else "slots_mixed"] = sv.slots
        n = int(rate * window)
        Y = M.random_states(template, seed + rate, np.arange(n))
        horizon = M.Rng(seed + rate, 91).uniform(n, 0.1, 2.0)
        req = Requests.from_chain_params(template, params, Y[:, :template.nq], Y[:, template.nq:], horizon,
                                         model_of=np.arange(n) % n_models)
        arrivals = np.cumsum(np.random.default_rng(seed + rate).exponential(1.0 / rate, size=n))
        log(f"[bench] serving chains: precision {prec}, {rate} req/s, {n} requests")
        r = serve(sv, req, arrivals, steps_per_advance=steps_per_advance, time_limit=window + 3.0, warmup=warmup)
        st = r.get("steady", {})
        ok = np.isfinite(r["latency"])
        win = (r["done_at"] >= warmup) & (r["done_at"] <= window)
        body_steps = int((r["stats"][win, 1].astype(np.int64) * template.nb).sum())
        out["sweep"].append({"precision": "fp64" if prec == 0 else "mixed", "rate_req_s": rate, "requests": n,
                             "offered_req_s": st.get("offered_req_s"),
                             "completed_req_s": st.get("completed_in_window_req_s"),
                             "completed": int(r["completed"]), "failed": int((r["status"][ok] < 0).sum()),
                             "throughput_body_steps_s": body_steps / (window - warmup),
                             "latency_p50_s": st.get("latency_p50_s"), "latency_p99_s": st.get("latency_p99_s"),
                             "accepted_steps_mean": float(r["stats"][ok, 1].mean()) if ok.any() else None,
                             "max_occupied_slots": int(r["max_occupied"]), "launches": int(r["launches"]),
                             "wall_s": r["wall"]})
        log(f"[bench] serving: {out['sweep'][-1]}")
        del sv
    return out


def row_tree_serving(torch, M, RungeKuttaMerson, rates=(1000, 8000, 32000), window=10.0, warmup=2.0, seed=7,
                     budget_gib=24, steps_per_advance=16, n_models=2048, lanes=0):
    """Serving requests of different trees (TreeServer): Poisson arrivals for
    `window` seconds, the first `warmup` discarded; each request one uneven
    branched tree (nb ~ U{8..64}, depth <= 20, Pin/Slider/Universal/Ball and
    Free roots; n_models distinct trees reused cyclically), a seeded initial
    state, horizon ~ U(0.1, 1) s, RK-Merson accuracy 1e-6; each request's
    memory is its own tree's, from pools sized by the HBM budget.  lanes: the
    server's evaluator setting (0 level-parallel with lanes from each tree's
    widest level, 1 the one-lane tree sweeps)."""
    from simbody_amd.serving import TreeRequests, TreeServer, serve
    models = [M.random_tree(seed, i, nmin=8, nmax=64) for i in range(n_models)]
    out = {"workload": f"Poisson arrivals for {window:g} s (first {warmup:g} s discarded); each request one "
                       "uneven tree, nb~U{8..64}, horizon U(0.1,1) s, RK-Merson accuracy 1e-6, FP64",
           "hbm_budget_bytes": budget_gib << 30, "steps_per_advance": steps_per_advance,
           "evaluator": "one-lane tree sweeps" if lanes == 1 else
                        f"level-parallel, {'lanes from the widest level' if lanes == 0 else f'{lanes} lanes'}",
           "sweep": []}
    for rate in rates:
        sv = TreeServer(RungeKuttaMerson(accuracy=1e-6), budget_gib << 30, max_slots=65536, lanes=lanes)
        n = int(rate * window)
        specs = [models[i % n_models] for i in range(n)]
        qs, us = [None] * n, [None] * n
        for m in range(min(n, n_models)):          # states drawn per model, vectorized over its requests
            idx = np.arange(m, n, n_models)
            Ym = M.random_states(models[m], seed + rate, idx)
            for j, i in enumerate(idx):
                qs[i] = Ym[j, :models[m].nq]; us[i] = Ym[j, models[m].nq:]
        horizon = M.Rng(seed + rate, 93).uniform(n, 0.1, 1.0)
        req = TreeRequests(specs, qs, us, horizon)
        arrivals = np.cumsum(np.random.default_rng(seed + rate).exponential(1.0 / rate, size=n))
        log(f"[bench] serving trees: {rate} req/s, {n} requests")
        r = serve(sv, req, arrivals, steps_per_advance=steps_per_advance, time_limit=window + 3.0, warmup=warmup)
        st = r.get("steady", {})
        ok = np.isfinite(r["latency"])
        win = (r["done_at"] >= warmup) & (r["done_at"] <= window)
        nbs = np.array([t.nb for t in specs])
        body_steps = int((r["stats"][win, 1].astype(np.int64) * nbs[win]).sum())
        out["sweep"].append({"rate_req_s": rate, "requests": n, "offered_req_s": st.get("offered_req_s"),
                             "completed_req_s": st.get("completed_in_window_req_s"),
                             "completed": int(r["completed"]), "failed": int((r["status"][ok] < 0).sum()),
                             "throughput_body_steps_s": body_steps / (window - warmup),
                             "latency_p50_s": st.get("latency_p50_s"), "latency_p99_s": st.get("latency_p99_s"),
                             "mean_bodies": float(nbs.mean()), "max_occupied_slots": int(r["max_occupied"]),
                             "launches": int(r["launches"]), "wall_s": r["wall"],
                             "admit_s": r["admit_s"], "advance_collect_s": r["advance_collect_s"]})
        log(f"[bench] serving trees: {out['sweep'][-1]}")
        del sv
    return out


def row_operators(torch, M, MatterBatch, n=65536, reps=5, cpu=True, cpu_sample=2000):
    """Single-shot operators on the chain workload (65,536 x 32 bodies, one
    model for the batch): kernel time per call (HIP events on the launch
    stream, mean of `reps` calls after one warm call), bodies/s, and the
    boundary bytes (inputs and outputs of the call) per second."""
    spec = M.spatial_chain(32, seed=2)
    Y = M.chain_batch_states(spec, n, seed=2026)
    B = MatterBatch([spec], n_instances=n, device="cuda:0")
    q = B.to_device(Y[:, :spec.nq]); u = B.to_device(Y[:, spec.nq:])
    rng = np.random.default_rng(7)
    v = B.to_device(rng.normal(size=(n, spec.nu)))
    F = B.to_device(rng.normal(size=(n, (spec.nb + 1) * 6)))
    nq, nu, nbp = spec.nq, spec.nu, spec.nb + 1
    ou = torch.empty(n * nu, dtype=torch.float64, device="cuda:0")
    ob = torch.empty(n * nbp * 10, dtype=torch.float64, device="cuda:0")
    ob6 = ob[:n * nbp * 6]
    ops = {   # name: (call into preallocated outputs, boundary doubles per instance)
        "multiplyByM": (lambda: B.multiplyByM(q, v, out=ou), nq + 2 * nu),
        "multiplyByMInv": (lambda: B.multiplyByMInv(q, v, out=ou), nq + 2 * nu),
        "multiplyBySystemJacobian": (lambda: B.multiplyBySystemJacobian(q, v, out=ob6), nq + nu + 6 * nbp),
        "multiplyBySystemJacobianTranspose": (lambda: B.multiplyBySystemJacobianTranspose(q, F, out=ou),
                                              nq + 6 * nbp + nu),
        "calcBodyAccelerationFromUDot": (lambda: B.calcBodyAccelerationFromUDot(q, u, v, out=ob6),
                                         nq + 2 * nu + 6 * nbp),
        "calcCompositeBodyInertias": (lambda: B.calcCompositeBodyInertias(q, out=ob), nq + 10 * nbp),
    }
    out = {"workload": f"{n} chains x {spec.nb} bodies (the chain rollout's model and states), FP64, one call",
           "ops": {}}
    for name, (call, dbl) in ops.items():
        call(); torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            call()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out["ops"][name] = {"ms": ms, "bodies_per_s": n * spec.nb / (ms / 1e3),
                            "boundary_gbs": n * dbl * 8 / (ms / 1e3) / 1e9}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle_ctypes import OracleModel
        o = OracleModel(spec)
        Yc = Y[:cpu_sample]
        vc = rng.normal(size=(cpu_sample, nu))
        t0 = time.perf_counter()
        for i in range(cpu_sample):
            o.multiply_by_m(Yc[i, :nq], vc[i])
        sec = time.perf_counter() - t0
        out["cpu_baseline"] = {"op": "multiplyByM", "bodies_per_s": cpu_sample * spec.nb / sec, "cores": 1,
                               "kind": "port", "sample": f"{cpu_sample} chains, one thread, through ctypes"}
    return out


def row_mixed(torch, M, MatterBatch, RungeKuttaMerson, n=65536, T=0.5, h=1e-3):
    """Mixed precision vs FP64 on the 32-body chains (Merson fixed step h to
    T) and on a stiff variant (neighbouring-link mass ratios up to 1e4;
    fixed steps of 1e-3 are unstable there, so Merson step control at
    accuracy 1e-6 to T).  Errors of the mixed path are measured against the
    FP64 GPU path (which matches the FP64 oracle to 1e-12 on udot): udot at
    t=0, q and u at T, and the energy drift E(T)-E(0) of both paths."""
    out = {}
    steps = int(round(T / h))
    for variant, spec in (("chains", M.spatial_chain(32, seed=2)), ("stiff", M.stiff_chain(32, seed=2))):
        Y = M.chain_batch_states(spec, n, seed=2026)
        res = {}
        for prec in (0, 1):
            B = MatterBatch([spec], n_instances=n, device="cuda:0")
            q = B.to_device(Y[:, :spec.nq]); u = B.to_device(Y[:, spec.nq:])
            integ = (RungeKuttaMerson(fixed_step=h, precision=prec) if variant == "chains"
                     else RungeKuttaMerson(accuracy=1e-6, precision=prec))
            B.rollout_begin(q, u, integ, t_end=T)
            _, ud = B.rollout_derivs()
            _, _, _, E0 = B.realizeKinematics(q, u, False, False)
            B.rollout_advance(2)                       # warm-up steps (counted in the state, not timed)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
            e0.record(); B.rollout_advance(steps if variant == "chains" else 1 << 20); e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            qT, uT, tT, st, stats = B.rollout_get()
            _, _, _, E1 = B.realizeKinematics(qT, uT, False, False)
            torch.cuda.synchronize()
            assert (st.cpu().numpy() == 1).all()
            E0 = E0.cpu().numpy().reshape(n, 2).sum(1); E1 = E1.cpu().numpy().reshape(n, 2).sum(1)
            stats = stats.cpu().numpy()
            res[prec] = dict(ms=ms, ud=ud.cpu().numpy().reshape(n, -1),
                             y=np.concatenate([qT.cpu().numpy().reshape(n, -1), uT.cpu().numpy().reshape(n, -1)], 1),
                             drift=np.abs(E1 - E0) / np.maximum(np.abs(E0), 1e-300),
                             timed_body_steps=int((stats[:, 1].astype(np.int64) - 2).sum()) * spec.nb)
            del B
        def rel(a, b):
            return np.abs(a - b).max(1) / np.maximum(np.abs(b).max(1), 1e-300)
        eu = rel(res[1]["ud"], res[0]["ud"]); ey = rel(res[1]["y"], res[0]["y"])
        fp64_rate = res[0]["timed_body_steps"] / (res[0]["ms"] / 1e3)
        mixed_rate = res[1]["timed_body_steps"] / (res[1]["ms"] / 1e3)
        out[variant] = {
            "integrator": f"fixed step {h}" if variant == "chains" else "accuracy 1e-6",
            "fp64_body_steps_s": fp64_rate, "mixed_body_steps_s": mixed_rate,
            "speedup_mixed_over_fp64": mixed_rate / fp64_rate,
            "udot_rel_err_max": float(eu.max()), "udot_rel_err_median": float(np.median(eu)),
            "state_T_rel_err_max": float(ey.max()), "state_T_rel_err_median": float(np.median(ey)),
            "energy_drift_fp64_median": float(np.median(res[0]["drift"])),
            "energy_drift_mixed_median": float(np.median(res[1]["drift"])),
            "energy_drift_fp64_max": float(res[0]["drift"].max()),
            "energy_drift_mixed_max": float(res[1]["drift"].max())}
        log(f"[bench] mixed {variant}: {out[variant]}")
    out["workload"] = (f"{n} x 32-body chains and a stiff variant (neighbouring mass ratios up to 1e4), Merson "
                       f"fixed step h={h} to T={T}; errors of the mixed path relative to the FP64 path "
                       "(per instance ||d||_inf/||ref||_inf)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--instances", type=int, default=65536)
    ap.add_argument("--bodies", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the secondary workloads (pendulum latency, uneven trees)")
    ap.add_argument("--only", choices=["pendulum", "trees", "serving", "tree-serving", "mixed", "operators"],
                    default=None,
                    help="run only one secondary workload and print its JSON")
    ap.add_argument("--tree-lanes", type=int, default=0,
                    help="lanes per instance of the tree evaluator (0 automatic, 1, 2, 4, 8, 16)")
    ap.add_argument("--tree-evaluator", type=int, default=0,
                    help="evaluator of the uneven-trees row: 0 automatic, 1 tree sweeps, 3 level-parallel")
    ap.add_argument("--tree-lane-range", type=int, nargs=2, default=[0, 32],
                    help="automatic lanes per instance: min (0: from the batch size) and max")
    ap.add_argument("--no-tree-autotune", action="store_true",
                    help="uneven-trees row: keep the given lane settings instead of tdyn_batch_autotune's choice")
    ap.add_argument("--rates", type=int, nargs="+", default=None,
                    help="--only tree-serving: arrival rates (req/s) to sweep")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 2 GiB (counter calibration for HBM traffic)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 28                         # 2 GiB of doubles read + 2 GiB written
        x = torch.ones(n, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the instances are independent and "
                                   "would shard across GPUs by instance, but no multi-GPU run is provided"}))
        sys.exit(2)

    import torch
    from simbody_amd import models as M
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson
    if args.only == "pendulum":
        print(json.dumps(row_pendulum(torch, M, MatterBatch, RungeKuttaMerson, cpu=not args.no_cpu)), flush=True)
        return
    if args.only == "trees":
        print(json.dumps(row_trees(torch, M, MatterBatch, RungeKuttaMerson, cpu=not args.no_cpu,
                                   lanes=args.tree_lanes, lane_range=tuple(args.tree_lane_range),
                                   evaluator=args.tree_evaluator,
                                   autotune=not args.no_tree_autotune and args.tree_lanes == 0
                                   and args.tree_evaluator == 0 and tuple(args.tree_lane_range) == (0, 32))),
              flush=True)
        return
    if args.only == "operators":
        print(json.dumps(row_operators(torch, M, MatterBatch, cpu=not args.no_cpu)), flush=True)
        return
    if args.only == "mixed":
        print(json.dumps(row_mixed(torch, M, MatterBatch, RungeKuttaMerson)), flush=True)
        return
    if args.only == "serving":
        print(json.dumps(row_serving(torch, M, RungeKuttaMerson)), flush=True)
        return
    if args.only == "tree-serving":
        kw = {"rates": tuple(args.rates)} if args.rates else {}
        print(json.dumps(row_tree_serving(torch, M, RungeKuttaMerson, lanes=args.tree_lanes, **kw)), flush=True)
        return

    h = 1e-3
    spec = M.spatial_chain(args.bodies, seed=2)
    n = args.instances
    log(f"[bench] generating {n} chains x {spec.nb} bodies")
    Y = M.chain_batch_states(spec, n, seed=2026)
    B = MatterBatch([spec], n_instances=n, device="cuda:0")
    q = B.to_device(Y[:, :spec.nq]); u = B.to_device(Y[:, spec.nq:])
    integ = RungeKuttaMerson(fixed_step=h)
    B.rollout_begin(q, u, integ, t_end=1e30)
    torch.cuda.synchronize()
    log(f"[bench] warmup {args.warmup} steps")
    if args.warmup > 0:
        B.rollout_advance(args.warmup)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True); ev1 = torch.cuda.Event(enable_timing=True)
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    ev0.record()
    B.rollout_advance(args.steps)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1)
    qf, uf, tf, st, stats = B.rollout_get()
    torch.cuda.synchronize()
    st = st.cpu().numpy(); stats = stats.cpu().numpy()
    assert (st == 0).all(), "some instance failed"
    assert (stats[:, 1] == args.warmup + args.steps).all(), "unexpected accepted-step count"
    body_steps = n * spec.nb * args.steps
    value = body_steps / (ms / 1e3)
    # inverse dynamics on the same states (separate figure)
    ud, _ = B.realizeAcceleration(qf, uf)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    res = B.calcResidualForceIgnoringConstraints(qf, uf, None, None, ud)
    e1.record()
    torch.cuda.synchronize()
    id_ms = e0.elapsed_time(e1)
    achieved = FLOPS_PER_BODY_STEP * body_steps / (ms / 1e3) / 1e12
    traffic = (TRAFFIC_FETCH_PER_BODY_STEP + TRAFFIC_WRITE_PER_BODY_STEP) * body_steps
    out = {
        "metric": "body-steps/s on 1 MI355X, plain vs mixed/batched at stated err; p50/p99 latency",
        "value": value, "unit": "body-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded chains and states)",
        "config": {"workload": f"{n} spatial serial chains x {spec.nb} bodies (Pin/Universal/Ball), "
                               f"FP64, RK-Merson fixed step h={h}", "instances": n, "bodies": spec.nb,
                   "integrator": "RungeKuttaMerson fixed step", "precision": "fp64"},
        "roofline": {"bound": "fp64-valu", "compute": "fp64 vector ALU (no MFMA is used)",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "traffic_note": "HBM bytes of the timed launch (rocprofv3 FETCH_SIZE+WRITE_SIZE per "
                                     "body-step, copy-calibrated, x body-steps)",
                     "hbm_gbs": traffic / (ms / 1e3) / 1e9, "hbm_frac": traffic / (ms / 1e3) / 1e9 / PEAK_HBM_GBS,
                     "flops_per_body_step": FLOPS_PER_BODY_STEP, "kernel": "k_rkm_chain"},
        "inverse_dynamics": {"ms": id_ms, "bodies_per_s": n * spec.nb / (id_ms / 1e3),
                             "tflops": FLOPS_RESID_PER_BODY * n * spec.nb / (id_ms / 1e3) / 1e12},
        "wall_s_timed": wall,
    }
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cb = cpu_baseline(spec, Y, args.steps, h, args.cpu_budget)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = {"pool": value / cb["value"], "per_core": value / cb["per_core_value"],
                               "all_online_cores_estimate": value / cb["all_online_cores_estimate"]["value"],
                               "target": 50.0}
    if not args.no_rows:
        del B
        rows = {}
        log("[bench] pendulum latency")
        rows["pendulum10"] = row_pendulum(torch, M, MatterBatch, RungeKuttaMerson, cpu=not args.no_cpu)
        rows["uneven_trees"] = row_trees(torch, M, MatterBatch, RungeKuttaMerson, cpu=not args.no_cpu)
        rows["operators"] = row_operators(torch, M, MatterBatch, cpu=not args.no_cpu)
        rows["mixed_precision"] = row_mixed(torch, M, MatterBatch, RungeKuttaMerson)
        rows["serving"] = row_serving(torch, M, RungeKuttaMerson)
        rows["tree_serving"] = row_tree_serving(torch, M, RungeKuttaMerson)
        out["rows"] = rows
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
