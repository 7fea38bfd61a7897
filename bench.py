"""Benchmark: propagated state-steps/s of the orientation batch (D = 4096).

Workload: the anisotropic radical pair with 6 + 4 spin-1/2 nuclei (D = 4096),
field directions from the 1024-point Fibonacci sphere (|B| = 50 uT), symmetric
recombination k = 1e-3 /ns, dt = 2 ns, Lanczos order 16, complex FP32 states
with FP64 inner products / exponential / yield sums.  One bench step
propagates a batch of `--orientations` orientations x 1024 direct columns over
`--intervals` time steps (one msg_propagate call: upload, persistent kernel,
reduction, yields back on the host).  A state-step is one column advanced by
dt including its observable sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X, bytes/s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--orientations", type=int, default=32)
    ap.add_argument("--intervals", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--serve-intervals", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from molspin_amd import abi, backend, workloads

    sysobj, cfg = workloads.config2(n_orientations=1024)
    packed = sysobj.pack()
    ctx = backend.Context(0)
    sysd = ctx.upload(packed)
    D = sysd.dim
    m = cfg["krylov_size"]
    dt = cfg["time_step"]
    B = args.orientations
    T = args.intervals * dt
    prm = abi.run_params(total_time=T, time_step=dt, method=abi.MSG_METHOD_KRYLOV, krylov_size=m)
    ncols = sysd.direct_columns()
    state_steps = B * ncols * args.intervals

    def step(i):
        first = (i * B) % 1024
        fields = [cfg["fields"][(first + j) % 1024] for j in range(B)]
        y, _ = sysd.propagate(prm, fields=fields)
        return y

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize() if torch.cuda.is_available() else None
    kms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        y = step(args.warmup + i)
        kms.append(ctx.last_kernel_ms())
    torch.cuda.synchronize() if torch.cuda.is_available() else None
    wall = time.perf_counter() - t0
    value = state_steps * args.steps / wall
    kernel_s = float(np.mean(kms)) * 1e-3
    bytes_per_ss = 8.0 * D * (4 * m + 1)  # streamed-basis Lanczos, complex FP32
    achieved = bytes_per_ss * state_steps / kernel_s

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(packed, cfg, args.cpu_seconds)
    serving = None if args.no_serving else serve(ctx, args.requests, args.serve_intervals)

    line = {
        "metric": "propagated state-steps/s + p99 yield-request latency at |dPhiS|<=1e-5 vs FP64",
        "value": value,
        "unit": "state-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex-f32 states, f64 reductions",
        "data": "synthetic (seeded radical pair, Fibonacci field directions)",
        "config": {"workload": "orientation batch: D=4096 (6+4 spin-1/2 nuclei, anisotropic HF), "
                               "%d orientations x %d columns x %d steps of 2 ns per step, Lanczos m=%d"
                               % (B, ncols, args.intervals, m),
                   "orientations_per_step": B, "columns": ncols, "intervals": args.intervals,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel_ms": kernel_s * 1e3, "bytes_per_state_step": bytes_per_ss},
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


def serve(ctx, n_requests, intervals):
    """Serving loop: a seeded Poisson stream (mean 200 requests/s) of radical-pair
    field-sweep yield requests (n ~ U{2..12} nuclei, 16-128 fields each), each
    propagated over `intervals` steps of 2 ns; requests are admitted by HBM
    footprint into a continuous-batching server (one persistent launch per request
    on its own stream, up to 8 overlapping).
    Latency = completion - arrival.  Returns p50/p99 and state-steps/s."""
    import numpy as np
    from molspin_amd import abi, backend, workloads
    reqs, t_arr = [], 0.0
    for i in range(n_requests):
        s, p, gap = workloads.serving_request(i)
        t_arr += gap
        packed = s.pack()
        sysd = ctx.upload(packed)
        prm = abi.run_params(total_time=intervals * p["time_step"], time_step=p["time_step"])
        reqs.append(dict(arrival=t_arr, sys=sysd, packed=packed, prm=prm, fields=p["fields"],
                         steps=sysd.direct_columns() * len(p["fields"]) * intervals))
    srv = backend.Server(ctx, max_slots=8, hbm_budget_bytes=64 << 30)
    t0 = time.perf_counter()
    pending = list(range(n_requests))
    live = {}
    lat = []
    while pending or live:
        now = time.perf_counter() - t0
        while pending and reqs[pending[0]]["arrival"] <= now:
            i = pending.pop(0)
            r = reqs[i]
            live[srv.submit(r["sys"], r["prm"], r["fields"])] = i
        if not live:
            time.sleep(max(0.0, reqs[pending[0]]["arrival"] - now))
            continue
        if srv.pump(0) == len(live) and not (pending and reqs[pending[0]]["arrival"] <= now):
            time.sleep(0.0005)
        for rid in list(live):
            y = srv.collect(rid)
            if y is not None:
                i = live.pop(rid)
                lat.append(time.perf_counter() - t0 - reqs[i]["arrival"])
    elapsed = time.perf_counter() - t0
    srv.close()
    lat = np.array(lat)
    total = sum(r["steps"] for r in reqs)
    # plain one-system-at-a-time form: the first 6 requests, one field per call
    t1 = time.perf_counter()
    plain_steps = 0
    for r in reqs[:6]:
        for f in r["fields"]:
            r["sys"].propagate(r["prm"], fields=[f])
        plain_steps += r["steps"]
    plain = plain_steps / (time.perf_counter() - t1)
    return {"requests": n_requests, "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "state_steps_per_s": total / elapsed,
            "plain_one_at_a_time_state_steps_per_s": plain, "served_over_plain": total / elapsed / plain,
            "elapsed_s": elapsed, "intervals_per_request": intervals,
            "workload": "Poisson 200 req/s, config-4-style pairs (D 16..16384), 16-128 fields/request"}


def cpu_baseline(packed, cfg, seconds):
    """FP64 oracle (the reference algorithm: Arnoldi m=16 + Pade expm, CSR
    matvecs) on a bounded sample of the same workload, all host threads."""
    from molspin_amd import abi
    from oracle import binding as ref
    threads = min(16, os.cpu_count() or 1)
    prm = abi.run_params(total_time=cfg["time_step"] * 4000, time_step=cfg["time_step"], method=abi.MSG_METHOD_KRYLOV)
    cols = 2 * threads
    steps = 4
    # calibrate to about `seconds` of work
    t0 = time.perf_counter()
    ref.run_columns(packed, prm, field=cfg["fields"][0], col_range=(0, cols), max_steps=steps, threads=threads)
    t1 = time.perf_counter() - t0
    steps = max(4, min(4000, int(steps * seconds / max(t1, 1e-3))))
    t0 = time.perf_counter()
    ref.run_columns(packed, prm, field=cfg["fields"][0], col_range=(0, cols), max_steps=steps, threads=threads)
    el = time.perf_counter() - t0
    return {"value": cols * steps / el, "unit": "state-steps/s", "cores": threads, "kind": "port",
            "sample": "1 orientation x %d direct columns x %d steps (FP64 Arnoldi m=16, %d OpenMP threads)"
                      % (cols, steps, threads)}


if __name__ == "__main__":
    main()
