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
TRAFFIC_FETCH_PER_BODY_STEP = 1615.9
TRAFFIC_WRITE_PER_BODY_STEP = 1078.7


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(spec, Y, steps, h, budget_s):
    """The oracle (scalar FP64 C++ restatement) on host cores, one instance
    per thread, on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_ctypes import OracleModel, time_integrate_batch
    om = OracleModel(spec)
    try:
        n_cpu = len(os.sched_getaffinity(0))
    except AttributeError:
        n_cpu = os.cpu_count() or 1
    threads = max(1, min(16, n_cpu))
    # calibrate with one short run per thread
    probe_steps = 20
    Yp = np.ascontiguousarray(Y[:threads].copy())
    sec, bs, _ = time_integrate_batch([om], np.zeros(threads, np.int32), Yp, probe_steps * h, 1e-3, h, threads)
    per_inst_full = sec * steps / probe_steps          # seconds per instance for `steps` steps, one thread
    n_inst = int(max(threads, min(len(Y), threads * max(1, int(budget_s / max(per_inst_full, 1e-9))))))
    Ys = np.ascontiguousarray(Y[:n_inst].copy())
    sec, bs, _ = time_integrate_batch([om], np.zeros(n_inst, np.int32), Ys, steps * h, 1e-3, h, threads)
    return {"value": bs / sec, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n_inst} chains x {steps} fixed steps (of the same workload), {threads} threads, {sec:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--instances", type=int, default=65536)
    ap.add_argument("--bodies", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
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
        print(json.dumps({"error": "this build runs on one GPU only; the instance batch would shard "
                                   "across GPUs by instance (see DESIGN.md), no multi-GPU run is provided"}))
        sys.exit(2)

    import torch
    from simbody_amd import models as M
    from simbody_amd.matter import MatterBatch, RungeKuttaMerson

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
        "roofline": {"bound": "mfma", "compute": "fp64 vector ALU (no MFMA is used)",
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
        out["cpu_baseline"] = cpu_baseline(spec, Y, args.steps, h, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
