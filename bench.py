"""Benchmark of the Hilbert-space propagation path on one GPU.

Default workload (the headline line): the orientation batch -- the anisotropic
radical pair with 6 + 4 spin-1/2 nuclei (D = 4096), field directions from the
1024-point Fibonacci sphere (|B| = 50 uT), symmetric recombination k = 1e-3 /ns,
dt = 2 ns, Lanczos order 16, complex FP32 states with FP64 inner products /
exponential / yield sums.  One bench step propagates `--orientations`
orientations x 1024 direct columns over `--intervals` time steps (one
msg_propagate call: upload, persistent kernel, reduction, yields back on the
host).  A state-step is one column advanced by dt including its observable
sample.

Other workloads (`--workload`), same JSON line format:
  isotropic     the orientation batch with each hyperfine tensor replaced by its
                isotropic part (the kernels' short form of axial couplings);
  trajectories  8 + 6 nuclei (D = 65536), `--trajectories` spin-coherent columns,
                singlet-triplet dephasing 1e-4 /ns, dt = 1 ns (streaming Lanczos);
  dense         a 7-nucleus pair (D = 512) swept over 64 field strengths on the
                dense MFMA path (FP64 propagator built once per field);
  pair          a 12-nucleus pair (D = 16384) at `--fields` field points x 4096
                direct columns (Lanczos with the Krylov vector mirrored in LDS);
  uncoupled     the orientation batch through the exact uncoupled-radical method;
  rydmr         the orientation batch with a linearly polarized 5 MHz field on both
                electrons (dynamics=dynamic: per-interval electron fields);
  arnoldi       the orientation batch with singlet-only recombination (unequal rates:
                matrix-free Arnoldi on -iH - K);
  timeinf       time-infinity yields (method=timeinf) of a D = 512 pair with unequal
                sinks over 64 field strengths (propagator doubling, FP64);
  uneven        config 4's uneven batch: the first `--pairs` radical pairs (nuclear
                counts as drawn, D = 16 ... 16384) x `--pair-fields` log-spaced field
                strengths x all direct columns in one persistent launch
                (msg_propagate_uneven), against the same pairs one at a time.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12      # MI355X, bytes/s
MFMA_F32_PEAK = 157.3e12  # MI355X dense FP32 matrix, flop/s
VALU_F32_PEAK = 157.3e12  # MI355X FP32 vector (published; counts packed FP32)
VALU_F32_UNPACKED = 78.6e12  # FP32 vector FMA rate without packed FP32 (256 CUs x 64 lanes x 2 x 2.4 GHz)
METRIC = "propagated state-steps/s + p99 yield-request latency at |dPhiS|<=1e-5 vs FP64"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["orientations", "isotropic", "trajectories", "dense", "pair", "uncoupled",
                                           "rydmr", "arnoldi", "timeinf", "uneven"],
                    default="orientations")
    ap.add_argument("--orientations", type=int, default=32)
    ap.add_argument("--trajectories", type=int, default=512)
    ap.add_argument("--fields", type=int, default=2, help="field points of the pair workload")
    ap.add_argument("--pairs", type=int, default=24, help="radical pairs of the uneven workload")
    ap.add_argument("--pair-fields", type=int, default=8, help="field points per pair of the uneven workload")
    ap.add_argument("--intervals", type=int, default=None, help="time steps per bench step")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--serve-intervals", type=int, default=5)
    ap.add_argument("--copy-gib", type=float, default=None,
                    help="only time a device-to-device copy of this many GiB (three times) and exit: the kernel of "
                         "known size that rocprofv3's FETCH_SIZE and WRITE_SIZE are calibrated on")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_gib is not None:
        device_copy(args.copy_gib)
        return

    from molspin_amd import backend

    ctx = backend.Context(0)
    try:
        line = run_bench(ctx, args)
    finally:
        ctx.close()  # release the library's device state before interpreter teardown
    print(json.dumps(line), flush=True)


def device_copy(gib):
    """Three device-to-device copies of `gib` GiB (read + write of that many bytes
    each); prints the copy rate as one JSON line."""
    import torch
    n = int(gib * (1 << 30))
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    ms = []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    print(json.dumps({"copy_bytes": n, "ms": ms, "read_plus_write_GBps": [2 * n / (t * 1e6) for t in ms]}), flush=True)


def load_fixture(name):
    with open(os.path.join(ROOT, "tests", "golden", name + ".json")) as f:
        return json.load(f)


def run_bench(ctx, args):
    import numpy as np
    import torch

    if args.workload in ("orientations", "isotropic"):
        run = orientation_batch(ctx, args)
    elif args.workload == "trajectories":
        run = trajectory_batch(ctx, args)
    elif args.workload == "uncoupled":
        run = uncoupled_batch(ctx, args)
    elif args.workload == "pair":
        run = pair_batch(ctx, args)
    elif args.workload in ("rydmr", "arnoldi"):
        run = variant_batch(ctx, args)
    elif args.workload == "timeinf":
        run = timeinf_batch(ctx, args)
    elif args.workload == "uneven":
        run = uneven_batch(ctx, args)
    else:
        run = dense_sweep(ctx, args)

    def progress(what):
        print("[bench] %s" % what, file=sys.stderr, flush=True)

    for i in range(args.warmup):
        run["step"](i)
    torch.cuda.synchronize() if torch.cuda.is_available() else None
    progress("%d warmup steps done" % args.warmup)
    kms, outs = [], {}
    t0 = time.perf_counter()
    for i in range(args.steps):
        outs[args.warmup + i] = run["step"](args.warmup + i)
        kms.append(ctx.last_kernel_ms())
    torch.cuda.synchronize() if torch.cuda.is_available() else None
    wall = time.perf_counter() - t0
    value = run["state_steps"] * args.steps / wall
    kernel_s = float(np.mean(kms)) * 1e-3
    roof = run["roofline"](kernel_s)
    progress("%d timed steps done in %.2f s" % (args.steps, wall))
    # the error against the FP64 oracle: the timed launches' own results where a
    # committed fixture covers them, and a post-timing run of the fixture's set
    parity = run["parity"](outs) if run.get("parity") else None
    progress("parity check done")
    compare = run["compare"]() if run.get("compare") else None

    cpu = None
    if not args.no_cpu and run.get("packed") is not None:
        cpu = cpu_baseline(run["packed"], run["cpu_prm"], run["cpu_field"], run["cpu_columns"], args.cpu_seconds)
        progress("CPU baseline done")
    serving = None
    if not args.no_serving and args.workload == "orientations":
        serving = serve(ctx, args.requests, args.serve_intervals)
        progress("serving stream done")
        # the same request stream through the exact uncoupled-radical method
        serving["uncoupled_method"] = serve(ctx, args.requests, args.serve_intervals, method="uncoupled",
                                            compare_plain=False)
        progress("serving stream through the uncoupled method done")

    return {
        "metric": METRIC,
        "value": value,
        "unit": run.get("unit", "state-steps/s"),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex-f32 states, f64 reductions",
        "data": run["data"],
        "config": run["config"],
        "roofline": roof,
        "parity": parity,
        "compare": compare,
        "cpu_baseline": cpu,
        "serving": serving,
        "device": torch.cuda.get_device_name(0) if torch.cuda.is_available() else None,
    }


def lanczos_roofline(D, m, state_steps, kernel_s, traffic, traffic_source=None, issue_key=None):
    """Matrix-free Lanczos at D >= 16384: algorithmic HBM bytes per state-step
    8 D (4m + 1) (m iterations reading v_j, v_{j-1} and writing v_{j+1}, plus the
    m + 1 vectors of the final combination, complex FP32).  `issue_key` names the
    profiled VALU issue fraction of the kernel, reported beside the byte model."""
    bytes_per_ss = 8.0 * D * (4 * m + 1)
    achieved = bytes_per_ss * state_steps / kernel_s
    out = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
           "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_source": traffic_source,
           "kernel_ms": kernel_s * 1e3, "bytes_per_state_step": bytes_per_ss,
           "state_steps_per_launch": state_steps}
    if issue_key in PROFILED_VALU_ISSUE:
        v = PROFILED_VALU_ISSUE[issue_key]
        out["valu_issue_profiled"] = {"value": v[0], "source": v[1]}
    return out


def operator_flops_per_state_step(D, m, n_couplings, e_block=False, axial=False):
    """FP32 flops of one state-step of the matrix-free Lanczos kernels, counted per
    electron-pair group (4 amplitudes, D / 4 groups) from the kernel's arithmetic:
    per matvec 32 FMAs per electron-nucleus coupling (a 2x2 complex block on one
    electron of the flipped neighbour group; 8 for an axial coupling, whose block has
    one real entry per orientation), 8 for the scalar shift, 24 per electron field,
    64 for a general 4x4 electron block; per iteration 20 more for the three-term
    recurrence (w -= beta v_{j-1}, w -= alpha v_j, 1/beta scaling); and 16 m FMAs for
    the final combination psi = sum_j c_j v_j.  An FMA is 2 flops."""
    fma_matvec = (8 if axial else 32) * n_couplings + 8 + 48 + (64 if e_block else 0)
    fma_group = m * (fma_matvec + 20) + 16 * m
    return 2.0 * fma_group * (D // 4)


def valu_roofline(D, m, n_couplings, state_steps, kernel_s, hbm_traffic, traffic_source, valu_issue, axial=False):
    """D <= 8192: the Krylov basis of a column stays in LDS / L2 and the operator's
    FP32 VALU work binds.  `achieved` counts the kernel's FP32 flops
    (operator_flops_per_state_step) against MI355X's published FP32 vector peak
    (157.3 TFLOP/s, which assumes packed FP32); this build compiles with packed FP32
    disabled, so 78.6 TFLOP/s is its attainable ceiling (`frac_of_unpacked`).  The
    byte model is reported beside it (`hbm_model`)."""
    fl = operator_flops_per_state_step(D, m, n_couplings, axial=axial)
    achieved = fl * state_steps / kernel_s
    hbm = lanczos_roofline(D, m, state_steps, kernel_s, hbm_traffic, traffic_source)
    return {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_F32_PEAK / 1e12, "unit": "TFLOP/s",
            "frac": achieved / VALU_F32_PEAK, "frac_of_unpacked": achieved / VALU_F32_UNPACKED,
            "traffic": hbm_traffic, "traffic_source": traffic_source,
            "kernel_ms": kernel_s * 1e3, "flops_per_state_step": fl, "state_steps_per_launch": state_steps,
            "valu_issue_profiled": valu_issue,
            "hbm_model": {"achieved": hbm["achieved"], "unit": "GB/s", "frac": hbm["frac"],
                          "bytes_per_state_step": hbm["bytes_per_state_step"]}}


# HBM bytes per state-step of the dominant kernel, from rocprofv3 FETCH_SIZE /
# WRITE_SIZE runs of each workload at a smaller launch (calibrated against a
# device copy: the FETCH_SIZE count is doubled, WRITE_SIZE taken as is); the
# traffic of a launch scales with its state-steps at fixed D.  Each entry names the
# profile summary it comes from; a kernel change needs a new counter run.
MEASURED_TRAFFIC_PER_STATE_STEP = {
    "orientations": (1.049e+06, "profiles/k_lanczos_orientation_batch_D4096.md"),
    "isotropic": (1.049e+06, "profiles/k_lanczos_isotropic_orientation_batch_D4096.md"),
    "dense": (2.47e+04, "profiles/k_dense_field_sweep_D512.md"),
    "pair/-1": (6.622e+06, "profiles/k_lanczos_big_pair_D16384.md"),
    "arnoldi": (5.470e+06, "profiles/k_arnoldi_singlet_sink_D4096.md"),
    # D = 65536 by launch shape: tiled streaming (-2), split columns (-3)
    "trajectories/-2": (7.54e+07, "profiles/k_lanczos_big_trajectories_D65536.md"),
    "trajectories/-3": (2.527e+07, "profiles/k_lanczos_split_trajectories_D65536.md"),
    # config 4's first 24 pairs x 8 fields x 20 steps: mixed D, so only this launch size
    "uneven": (3.555e+06, "profiles/k_uneven_config4_pairs0_23.md"),
}
# VALU issue fraction (SQ_INSTS_VALU / (GRBM_GUI_ACTIVE / 8 XCDs x 256 CUs)) of the profiled kernels
PROFILED_VALU_ISSUE = {"orientations": (0.78, "profiles/k_lanczos_orientation_batch_D4096.md"),
                       "isotropic": (0.50, "profiles/k_lanczos_isotropic_orientation_batch_D4096.md"),
                       "pair/-1": (0.82, "profiles/k_lanczos_big_pair_D16384.md"),
                       "trajectories/-3": (0.31, "profiles/k_lanczos_split_trajectories_D65536.md"),
                       "arnoldi": (0.60, "profiles/k_arnoldi_singlet_sink_D4096.md"),
                       "uneven": (0.79, "profiles/k_uneven_config4_pairs0_23.md")}


def traffic_for(key, state_steps):
    t = MEASURED_TRAFFIC_PER_STATE_STEP.get(key)
    return (None, None) if t is None else (t[0] * state_steps, t[1] + " (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)")


def oracle_parity(sysd, packed, prm, fields, what, columns=None):
    """After timing: the device and the FP64 oracle on the same short run (`what`
    names it), max |dPhi| over every yield."""
    import numpy as np
    from oracle import binding as ref
    y, _ = sysd.propagate(prm, fields=fields, columns=columns)
    yo, _ = ref.run_hsgeneral(packed, prm, fields=fields, columns=columns, threads=16)
    return {"oracle_run": "FP64 oracle, " + what, "tolerance": {"yield_abs": 1e-5},
            "max_abs_dPhi": float(np.abs(y - yo).max())}


def orientation_batch(ctx, args):
    """The headline orientation batch; `--workload isotropic` runs it with each
    hyperfine tensor replaced by its isotropic part (the kernels' short form of axial
    couplings), checked against the oracle on a short run after timing."""
    from molspin_amd import abi, workloads
    iso = args.workload == "isotropic"
    sysobj, cfg = workloads.config2(n_orientations=1024, isotropic=iso)
    packed = sysobj.pack()
    sysd = ctx.upload(packed)
    D, m, dt = sysd.dim, cfg["krylov_size"], cfg["time_step"]
    B = args.orientations
    intervals = args.intervals or 50
    prm = abi.run_params(total_time=intervals * dt, time_step=dt, method=abi.MSG_METHOD_KRYLOV, krylov_size=m)
    ncols = sysd.direct_columns()
    state_steps = B * ncols * intervals
    shape = sysd.autotune(prm, fields=cfg["fields"][:B], sample_steps=5)

    def orients(i):
        return [(i * B + j) % 1024 for j in range(B)]

    def step(i):
        y, _ = sysd.propagate(prm, fields=[cfg["fields"][o] for o in orients(i)])
        return y

    def roof(kernel_s):
        key = "isotropic" if iso else "orientations"
        tr, src = traffic_for(key, state_steps)
        issue = PROFILED_VALU_ISSUE[key]
        return valu_roofline(D, m, 10, state_steps, kernel_s, tr, src,
                             {"value": issue[0], "source": issue[1]}, axial=iso)

    def parity_iso(outs):
        import numpy as np
        from oracle import binding as ref
        # 2 orientations, all columns, 10 steps: the device against the FP64 oracle
        prm_s = abi.run_params(total_time=10 * dt, time_step=dt, method=abi.MSG_METHOD_KRYLOV, krylov_size=m)
        fs = [cfg["fields"][o] for o in (0, 517)]
        y, _ = sysd.propagate(prm_s, fields=fs)
        yo, _ = ref.run_hsgeneral(packed, prm_s, fields=fs, threads=16)
        return {"oracle_run": "FP64 oracle, 2 orientations x %d columns x 10 steps" % ncols,
                "tolerance": {"yield_abs": 1e-5}, "max_abs_dPhi": float(np.abs(y - yo).max())}

    def parity(outs):
        import numpy as np
        if iso:
            return parity_iso(outs)
        fx = load_fixture("cfg2_bench")
        fy, ftr = np.array(fx["yields"]), np.array(fx["trace"])
        res = {"fixture": "tests/golden/cfg2_bench.json (FP64 oracle: Arnoldi m=16, CSR H)",
               "tolerance": {"yield_abs": 1e-5, "trace_max_rel": 1e-4}}
        # orientations of the timed launches that the fixture covers (same horizon)
        timed = []
        if intervals * dt == fx["total_time"]:
            for i, y in outs.items():
                for j, o in enumerate(orients(i)):
                    if o in fx["orientations"]:
                        timed.append(float(np.abs(y[j] - fy[fx["orientations"].index(o)]).max()))
        res["timed_launches"] = {"orientations_checked": len(timed),
                                 "max_abs_dPhi": max(timed) if timed else None}
        # the fixture's 8 orientations in one launch, yields and traces
        prm_f = abi.run_params(total_time=fx["total_time"], time_step=fx["time_step"], method=abi.MSG_METHOD_KRYLOV,
                               krylov_size=fx["krylov_size"])
        y, tr = sysd.propagate(prm_f, fields=[cfg["fields"][o] for o in fx["orientations"]], trace=True)
        terr = max(float(np.abs(tr[k] - ftr[k]).max() / np.abs(ftr[k]).max()) for k in range(len(ftr)))
        res["fixture_launch"] = {"orientations": len(fx["orientations"]), "max_abs_dPhi": float(np.abs(y - fy).max()),
                                 "trace_max_rel": terr}
        res["max_abs_dPhi"] = max([res["fixture_launch"]["max_abs_dPhi"]] + timed)
        return res

    return dict(step=step, state_steps=state_steps, roofline=roof, parity=parity, packed=packed,
                cpu_prm=abi.run_params(total_time=dt * 4000, time_step=dt, method=abi.MSG_METHOD_KRYLOV),
                cpu_field=cfg["fields"][0], cpu_columns=None,
                data="synthetic (seeded radical pair, Fibonacci field directions)",
                config={"workload": "orientation batch: D=4096 (6+4 spin-1/2 nuclei, %s HF), "
                                    "%d orientations x %d columns x %d steps of 2 ns per step, Lanczos m=%d"
                                    % ("isotropic" if iso else "anisotropic", B, ncols, intervals, m),
                        "orientations_per_step": B, "columns": ncols, "intervals": intervals,
                        "full_config": "1024 orientations x 1024 columns x 1500 steps (1.57e9 state-steps); "
                                       "each bench step is a 32-orientation, 50-step slice with the same "
                                       "per-state-step work",
                        "parallelism": "single GPU",
                        "launch_shape": {"groups_per_thread": shape.groups_per_thread, "threads": shape.threads,
                                         "slots": shape.slots, "chosen_by": "msg_autotune"}})


def trajectory_batch(ctx, args):
    from molspin_amd import abi, workloads
    M = args.trajectories
    sysobj, cfg = workloads.config3(n_traj=M)
    packed = sysobj.pack()
    sysd = ctx.upload(packed)
    D, m, dt = sysd.dim, cfg["krylov_size"], cfg["time_step"]
    intervals = args.intervals or 20
    cols = sysd.trace_samples(sysobj.state_index("S"), M, cfg["sample_seed"]) / math.sqrt(M)
    prm = abi.run_params(total_time=intervals * dt, time_step=dt, method=abi.MSG_METHOD_KRYLOV, krylov_size=m,
                         sampling=abi.MSG_SAMPLING_GIVEN, seed=2026)
    state_steps = M * intervals
    shape = sysd.autotune(prm, columns=cols[:, :min(M, 256)], sample_steps=2)

    def step(i):
        return sysd.propagate(prm, columns=cols)[0]

    def roof(kernel_s):
        key = "trajectories/%d" % shape.groups_per_thread
        tr, src = traffic_for(key, state_steps)
        return lanczos_roofline(D, m, state_steps, kernel_s, tr, src, key)

    def parity(outs):
        import numpy as np
        fx = load_fixture("cfg3_subset")
        Mf = fx["trajectories"]
        cf = sysd.trace_samples(sysobj.state_index("S"), Mf, fx["sample_seed"]) / math.sqrt(Mf)
        prm_f = abi.run_params(total_time=fx["total_time"], time_step=fx["time_step"], method=abi.MSG_METHOD_KRYLOV,
                               sampling=abi.MSG_SAMPLING_GIVEN, seed=fx["flip_seed"], krylov_size=fx["krylov_size"])
        y, _ = sysd.propagate(prm_f, columns=cf)
        err = float(np.abs(y[0] - np.array(fx["yields"])).max())
        return {"fixture": "tests/golden/cfg3_subset.json (FP64 oracle)", "tolerance": {"yield_abs": 1e-5},
                "fixture_launch": {"trajectories": Mf, "steps": int(fx["total_time"] / fx["time_step"]),
                                   "max_abs_dPhi": err},
                "max_abs_dPhi": err}

    cpu_prm = abi.run_params(total_time=dt * 4000, time_step=dt, method=abi.MSG_METHOD_KRYLOV,
                             sampling=abi.MSG_SAMPLING_GIVEN, seed=2026)
    return dict(step=step, state_steps=state_steps, roofline=roof, parity=parity, packed=packed, cpu_prm=cpu_prm,
                cpu_field=None, cpu_columns=cols,
                data="synthetic (seeded radical pair, host-drawn spin-coherent nuclear states)",
                config={"workload": "stochastic trajectories: D=65536 (8+6 spin-1/2 nuclei, isotropic HF), "
                                    "%d trajectories x %d steps of 1 ns, ST dephasing 1e-4/ns, Lanczos m=%d"
                                    % (M, intervals, m),
                        "trajectories": M, "intervals": intervals, "parallelism": "single GPU",
                        "launch_shape": {"groups_per_thread": shape.groups_per_thread, "threads": shape.threads,
                                         "slots": shape.slots, "chosen_by": "msg_autotune"}})


def pair_batch(ctx, args):
    """A D = 16384 radical pair (12 anisotropic spin-1/2 nuclei) of the serving mix
    at `--fields` field points, all 4096 direct columns per field."""
    from molspin_amd import abi, workloads
    idx = next(i for i in range(1000) if workloads.config4_pair(i)[1]["n"] == 12)
    sysobj, cfg = workloads.config4_pair(idx)
    packed = sysobj.pack()
    sysd = ctx.upload(packed)
    D, m, dt = sysd.dim, cfg["krylov_size"], cfg["time_step"]
    intervals = args.intervals or 10
    prm = abi.run_params(total_time=intervals * dt, time_step=dt, method=abi.MSG_METHOD_KRYLOV, krylov_size=m)
    fields = cfg["fields"][:args.fields]
    ncols = sysd.direct_columns()
    state_steps = len(fields) * ncols * intervals
    shape = sysd.autotune(prm, fields=fields, sample_steps=2)

    def step(i):
        return sysd.propagate(prm, fields=fields)[0]

    def roof(kernel_s):
        key = "pair/%d" % shape.groups_per_thread
        tr, src = traffic_for(key, state_steps)
        return lanczos_roofline(D, m, state_steps, kernel_s, tr, src, key)

    def parity(outs):
        from oracle import binding as ref
        cols = ref.coherent_samples(packed, sysobj.state_index("S"), 5, 4) / 2.0
        prm_s = abi.run_params(total_time=10 * dt, time_step=dt, method=abi.MSG_METHOD_KRYLOV, krylov_size=m,
                               sampling=abi.MSG_SAMPLING_GIVEN)
        return oracle_parity(sysd, packed, prm_s, fields, "%d fields x 4 spin-coherent columns x 10 steps"
                             % len(fields), columns=cols)

    return dict(step=step, state_steps=state_steps, roofline=roof, parity=parity, packed=packed,
                cpu_prm=abi.run_params(total_time=dt * 4000, time_step=dt, method=abi.MSG_METHOD_KRYLOV),
                cpu_field=fields[0], cpu_columns=None,
                data="synthetic (seeded radical pair, log-spaced field strengths)",
                config={"workload": "radical pair D=16384 (12 spin-1/2 nuclei, anisotropic HF, pair %d), "
                                    "%d fields x %d columns x %d steps of %g ns, Lanczos m=%d"
                                    % (idx, len(fields), ncols, intervals, dt, m),
                        "fields": len(fields), "columns": ncols, "intervals": intervals,
                        "parallelism": "single GPU",
                        "launch_shape": {"groups_per_thread": shape.groups_per_thread, "threads": shape.threads,
                                         "slots": shape.slots, "chosen_by": "msg_autotune"}})


def variant_batch(ctx, args):
    """The orientation batch (D = 4096, 6 + 4 anisotropic nuclei) with either a
    linearly polarized 5 MHz field of 50 uT on both electrons (rydmr) or
    singlet-only recombination at 2e-3 /ns (arnoldi)."""
    from molspin_amd import abi, workloads
    rydmr = args.workload == "rydmr"
    rng = workloads.SplitMix64(2, 0)
    hf = [workloads._aniso_hf(rng) for _ in range(10)]
    sysobj = workloads._radical_pair("cfg2", ["1/2"] * 6, ["1/2"] * 4, hf, 1e-3, swept=True)
    if rydmr:
        sysobj.add_zeeman(["E1", "E2"], (0.05, 0.0, 0.0), prefactor=0.001, time_dependence="linearpolarized",
                          frequency=2 * math.pi * 0.005)
    else:
        for t in sysobj.transitions:
            t["rate"] = 2e-3 if sysobj.states[t["source"]]["name"] == "S" else 0.0
    packed = sysobj.pack()
    sysd = ctx.upload(packed)
    D, m, dt = sysd.dim, 16, 2.0
    B = args.orientations if rydmr else min(args.orientations, 8)
    intervals = args.intervals or (50 if rydmr else 20)
    fields = [workloads.fibonacci_point(1024, i, 0.05) for i in range(B)]
    prm = abi.run_params(total_time=intervals * dt, time_step=dt, method=abi.MSG_METHOD_KRYLOV, krylov_size=m,
                         dynamics=rydmr)
    ncols = sysd.direct_columns()
    state_steps = B * ncols * intervals
    shape = sysd.autotune(prm, fields=fields, sample_steps=3)

    def step(i):
        return sysd.propagate(prm, fields=fields)[0]

    def roof(kernel_s):
        if rydmr:
            return lanczos_roofline(D, m, state_steps, kernel_s, None)
        # Arnoldi: per state-step m basis stores, one read of each earlier vector per
        # Gram-Schmidt sweep (v_j itself is in LDS; m (m - 1) / 2 reads) and m reads in
        # the final combination: 8 D (2 m + m (m - 1) / 2) bytes
        bss = 8.0 * D * (2 * m + m * (m - 1) / 2)
        tr, src = traffic_for("arnoldi", state_steps)
        return {"bound": "hbm", "achieved": bss * state_steps / kernel_s / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": bss * state_steps / kernel_s / HBM_PEAK, "traffic": tr, "traffic_source": src,
                "kernel_ms": kernel_s * 1e3, "bytes_per_state_step": bss, "state_steps_per_launch": state_steps}

    def parity(outs):
        from oracle import binding as ref
        cols = ref.coherent_samples(packed, sysobj.state_index("S"), 7, 16) / 4.0
        prm_s = abi.run_params(total_time=10 * dt, time_step=dt, method=abi.MSG_METHOD_KRYLOV, krylov_size=m,
                               dynamics=rydmr, sampling=abi.MSG_SAMPLING_GIVEN)
        return oracle_parity(sysd, packed, prm_s, fields[:2], "2 orientations x 16 spin-coherent columns x 10 steps",
                             columns=cols)

    what = ("linearly polarized 5 MHz, 50 uT field on both electrons (dynamics=dynamic)" if rydmr
            else "singlet-only recombination 2e-3/ns (Arnoldi on -iH - K)")
    return dict(step=step, state_steps=state_steps, roofline=roof, parity=parity, packed=packed,
                cpu_prm=abi.run_params(total_time=dt * 4000, time_step=dt, method=abi.MSG_METHOD_KRYLOV,
                                       dynamics=rydmr),
                cpu_field=fields[0], cpu_columns=None,
                data="synthetic (seeded radical pair, Fibonacci field directions)",
                config={"workload": "orientation batch D=4096 with %s: %d orientations x %d columns x %d steps of "
                                    "2 ns, Krylov m=%d" % (what, B, ncols, intervals, m),
                        "orientations_per_step": B, "columns": ncols, "intervals": intervals,
                        "parallelism": "single GPU",
                        "launch_shape": {"groups_per_thread": shape.groups_per_thread, "threads": shape.threads,
                                         "slots": shape.slots, "chosen_by": "msg_autotune"}})


def timeinf_batch(ctx, args):
    """Time-infinity yields (msg_time_infinity, doubling route) of a 7-nucleus pair
    (D = 512) with unequal sinks over the 64 field strengths of the dense sweep."""
    from molspin_amd import abi, workloads
    idx = next(i for i in range(1000) if workloads.config4_pair(i)[1]["n"] == 7)
    sysobj, cfg = workloads.config4_pair(idx)
    for t in sysobj.transitions:
        t["rate"] = 3e-3 if sysobj.states[t["source"]]["name"] == "S" else 1e-3
    packed = sysobj.pack()
    sysd = ctx.upload(packed)
    prm = abi.run_params()
    nf = len(cfg["fields"])

    def step(i):
        return sysd.time_infinity(prm, fields=cfg["fields"])

    def parity(outs):
        # the oracle's H, K and rho0 at 2 field strengths through scipy's Bartels-Stewart
        # solver (the oracle's own Kronecker-form solve is D^2 x D^2)
        import numpy as np
        import scipy.linalg as sl
        from oracle import binding as ref
        picks = (0, 40)
        y = sysd.time_infinity(prm, fields=[cfg["fields"][i] for i in picks])
        K, B = ref.reaction(packed), ref.direct_factors(packed)
        rho0 = B @ B.conj().T
        err = 0.0
        for k, i in enumerate(picks):
            X = sl.solve_continuous_lyapunov(-1j * ref.hamiltonian(packed, cfg["fields"][i]) - K, -rho0)
            yo = np.array([packed.desc.transitions[t].rate *
                           np.trace(ref.projector(packed, packed.desc.transitions[t].source_state) @ X).real
                           for t in ref.packed_transitions(packed)])
            err = max(err, float(np.abs(y[k] - yo).max()))
        return {"oracle_run": "scipy Bartels-Stewart on the oracle's H, K and rho0, 2 field strengths",
                "tolerance": {"yield_abs": 1e-9}, "max_abs_dPhi": err}

    return dict(step=step, state_steps=nf, roofline=lambda kernel_s: None, parity=parity, packed=None,
                data="synthetic (seeded radical pair, log-spaced field strengths)",
                unit="yield sets/s",
                config={"workload": "time-infinity yields: D=512 pair %d (7 spin-1/2 nuclei), unequal sinks "
                                    "k_S=3e-3, k_T=1e-3 /ns, %d field strengths per step (FP64 doubling)" % (idx, nf),
                        "fields": nf, "parallelism": "single GPU"})


def uneven_batch(ctx, args):
    """Config 4's uneven batch in one persistent launch: pairs 0 .. P-1 of the
    config (n ~ U{2..12} nuclei as drawn), F of the 64 log-spaced field strengths
    (every 64/F-th), all direct columns, `--intervals` steps of 2 ns."""
    import numpy as np
    from molspin_amd import abi, workloads
    P, F = args.pairs, args.pair_fields
    intervals = args.intervals or 20
    pairs = []
    for i in range(P):
        s, cfg = workloads.config4_pair(i)
        sysd = ctx.upload(s.pack())
        fields = cfg["fields"][::max(1, 64 // F)][:F]
        pairs.append(dict(sys=sysd, n=cfg["n"], fields=fields, cols=sysd.direct_columns()))
    prm = abi.run_params(total_time=intervals * 2.0, time_step=2.0, method=abi.MSG_METHOD_AUTO, krylov_size=16)
    ents = [dict(system=p["sys"], params=prm, fields=p["fields"]) for p in pairs]
    state_steps = sum(len(p["fields"]) * p["cols"] * intervals for p in pairs)
    mf_bytes = sum(len(p["fields"]) * p["cols"] * intervals * 8.0 * p["sys"].dim * 65 for p in pairs if p["sys"].dim > 512)

    def step(i):
        return ctx.propagate_uneven(ents)

    # the same work one pair at a time (each pair's whole field sweep per call), and
    # one pair and one field per call, timed once each outside the timed region
    def one_at_a_time():
        t0 = time.perf_counter()
        for p in pairs:
            p["sys"].propagate(prm, fields=p["fields"])
        t_pair = time.perf_counter() - t0
        t0 = time.perf_counter()
        for p in pairs:
            for f in p["fields"]:
                p["sys"].propagate(prm, fields=[f])
        t_field = time.perf_counter() - t0
        return {"pair_at_a_time_state_steps_per_s": state_steps / t_pair,
                "pair_and_field_at_a_time_state_steps_per_s": state_steps / t_field}

    def roof(kernel_s):
        # measured traffic applies to the profiled launch size only (mixed D)
        tr, src = traffic_for("uneven", state_steps) if (P, F, intervals) == (24, 8, 20) else (None, None)
        v = PROFILED_VALU_ISSUE["uneven"]
        return {"bound": "hbm", "achieved": mf_bytes / kernel_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": mf_bytes / kernel_s / HBM_PEAK, "traffic": tr, "traffic_source": src,
                "kernel_ms": kernel_s * 1e3, "state_steps_per_launch": state_steps,
                "valu_issue_profiled": {"value": v[0], "source": v[1]} if tr else None,
                "note": "byte model 8 D (4m + 1) per state-step summed over the matrix-free (D > 512) units; "
                        "dense units counted in state-steps only"}

    def parity(outs):
        fx = load_fixture("cfg4_uneven")
        prm_f = abi.run_params(total_time=fx["total_time"], time_step=fx["time_step"], method=abi.MSG_METHOD_AUTO,
                               krylov_size=fx["krylov_size"])
        ents_f = []
        for e in fx["pairs"]:
            s, _ = workloads.config4_pair(e["pair"])
            ents_f.append(dict(system=ctx.upload(s.pack()), params=prm_f, fields=e["fields"]))
        out = ctx.propagate_uneven(ents_f)
        errs = {str(e["n"]): float(np.abs(y - np.array(e["yields"])).max()) for e, (y, _) in zip(fx["pairs"], out)}
        return {"fixture": "tests/golden/cfg4_uneven.json (FP64 oracle; one pair per n = 2..12, 2 fields, 10 steps)",
                "tolerance": {"yield_abs": 1e-5}, "max_abs_dPhi_by_n": errs, "max_abs_dPhi": max(errs.values())}

    ns = sorted(p["n"] for p in pairs)
    return dict(step=step, state_steps=state_steps, roofline=roof, parity=parity, packed=None, compare=one_at_a_time,
                data="synthetic (seeded radical pairs, log-spaced field strengths)",
                config={"workload": "uneven batch (config 4 subset): pairs 0..%d (n = %s), %d fields x all direct "
                                    "columns x %d steps of 2 ns, one persistent launch" % (P - 1, ns, F, intervals),
                        "pairs": P, "fields_per_pair": F, "intervals": intervals,
                        "full_config": "256 pairs x 64 fields x 1500 steps (1.8e10 state-steps)",
                        "parallelism": "single GPU"})


def uncoupled_batch(ctx, args):
    """The orientation batch through the exact uncoupled-radical method (radicals of
    dimension 128 and 32): all 1024 orientations over the full 0-3000 ns grid per
    step; state-steps counted as for the Lanczos form (1024 columns per orientation)."""
    from molspin_amd import abi, workloads
    sysobj, cfg = workloads.config2(n_orientations=1024)
    packed = sysobj.pack()
    sysd = ctx.upload(packed)
    D, dt = sysd.dim, cfg["time_step"]
    intervals = args.intervals or int(round(cfg["total_time"] / dt))
    prm = abi.run_params(total_time=intervals * dt, time_step=dt, method=abi.MSG_METHOD_UNCOUPLED)
    B = 1024
    ncols = sysd.direct_columns()
    state_steps = B * ncols * intervals
    ns = intervals + 1
    # correlation sums (k_corr): per eigenvalue pair g <= h and sample, 15 FMAs for the 9
    # Hermitian products and the phase rotation (4 multiplies, 2 adds): 36 flops
    flops = B * (128 * 129 // 2 + 32 * 33 // 2) * ns * 36.0

    def step(i):
        return sysd.propagate(prm, fields=cfg["fields"])[0]

    def roof(kernel_s):
        return {"bound": "fp64-valu", "achieved": flops / kernel_s / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                "frac": flops / kernel_s / 78.6e12, "traffic": None, "kernel_ms": kernel_s * 1e3,
                "state_steps_per_launch": state_steps,
                "note": "k_corr's flops over the device time of the whole method (eigensolver, GEMMs, "
                        "correlation and combination kernels)"}

    def parity(outs):
        import numpy as np
        fx = load_fixture("cfg2_subset")
        prm_f = abi.run_params(total_time=fx["total_time"], time_step=fx["time_step"], method=abi.MSG_METHOD_UNCOUPLED)
        y, _ = sysd.propagate(prm_f, fields=[cfg["fields"][o] for o in fx["orientations"]])
        return {"fixture": "tests/golden/cfg2_subset.json (FP64 oracle)", "tolerance": {"yield_abs": 1e-5},
                "orientations": len(fx["orientations"]),
                "max_abs_dPhi": float(np.abs(y - np.array(fx["yields"])).max())}

    return dict(step=step, state_steps=state_steps, roofline=roof, parity=parity, packed=packed,
                cpu_prm=abi.run_params(total_time=dt * 4000, time_step=dt, method=abi.MSG_METHOD_KRYLOV),
                cpu_field=cfg["fields"][0], cpu_columns=None,
                data="synthetic (seeded radical pair, Fibonacci field directions)",
                config={"workload": "orientation batch, exact uncoupled-radical method: D=4096 as radicals of 128 and "
                                    "32, %d orientations x %d columns x %d steps of 2 ns per step" % (B, ncols, intervals),
                        "orientations_per_step": B, "columns": ncols, "intervals": intervals,
                        "parallelism": "single GPU"})


def dense_sweep(ctx, args):
    from molspin_amd import abi, workloads
    idx = next(i for i in range(1000) if workloads.config4_pair(i)[1]["n"] == 7)
    sysobj, cfg = workloads.config4_pair(idx)
    packed = sysobj.pack()
    sysd = ctx.upload(packed)
    D, dt = sysd.dim, cfg["time_step"]
    intervals = args.intervals or 500
    prm = abi.run_params(total_time=intervals * dt, time_step=dt, method=abi.MSG_METHOD_EXPONENTIAL)
    ncols = sysd.direct_columns()
    nf = len(cfg["fields"])
    state_steps = nf * ncols * intervals

    def step(i):
        return sysd.propagate(prm, fields=cfg["fields"])[0]

    def roof(kernel_s):
        flops = 8.0 * D * D * state_steps
        tr, src = traffic_for("dense", state_steps)
        return {"bound": "mfma", "achieved": flops / kernel_s / 1e12, "peak": MFMA_F32_PEAK / 1e12,
                "unit": "TFLOP/s", "frac": flops / kernel_s / MFMA_F32_PEAK,
                "traffic": tr, "traffic_source": src,
                "kernel_ms": kernel_s * 1e3, "flops_per_state_step": 8.0 * D * D,
                "state_steps_per_launch": state_steps}

    def parity(outs):
        prm_s = abi.run_params(total_time=10 * dt, time_step=dt, method=abi.MSG_METHOD_EXPONENTIAL)
        return oracle_parity(sysd, packed, prm_s, [cfg["fields"][i] for i in (0, 63)],
                             "2 field strengths x %d columns x 10 steps" % ncols)

    return dict(step=step, state_steps=state_steps, roofline=roof, parity=parity, packed=packed,
                cpu_prm=abi.run_params(total_time=dt * 4000, time_step=dt, method=abi.MSG_METHOD_EXPONENTIAL),
                cpu_field=cfg["fields"][0], cpu_columns=None,
                data="synthetic (seeded radical pair, log-spaced field strengths)",
                config={"workload": "dense field sweep: D=512 (7 spin-1/2 nuclei, anisotropic HF, pair %d), "
                                    "%d fields x %d columns x %d steps of 2 ns, FP32 MFMA propagation"
                                    % (idx, nf, ncols, intervals),
                        "fields": nf, "columns": ncols, "intervals": intervals, "parallelism": "single GPU"})


def serve(ctx, n_requests, intervals, method="auto", compare_plain=True):
    """Serving loop: a seeded Poisson stream (mean 200 requests/s) of radical-pair
    field-sweep yield requests (n ~ U{2..12} nuclei, 16-128 fields each), each
    propagated over `intervals` steps of 2 ns; requests are admitted by HBM
    footprint into a continuous-batching server (one persistent launch per request
    on its own stream, up to 8 overlapping).  Latency = completion - arrival.

    The plain form is the same requests one system and one field per call,
    sequentially; its time is estimated per request from two timed single-field
    calls (scaled by the request's field count) so the comparison covers the
    same mix without doubling the run time."""
    import numpy as np
    from molspin_amd import abi, backend, workloads
    reqs, t_arr = [], 0.0
    for i in range(n_requests):
        s, p, gap = workloads.serving_request(i)
        t_arr += gap
        packed = s.pack()
        sysd = ctx.upload(packed)
        prm = abi.run_params(total_time=intervals * p["time_step"], time_step=p["time_step"],
                             method=abi.MSG_METHOD_UNCOUPLED if method == "uncoupled" else abi.MSG_METHOD_AUTO)
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
                reqs[i]["result"] = y
    elapsed = time.perf_counter() - t0
    shared = srv.stats()
    srv.close()
    lat = np.array(lat)
    total = sum(r["steps"] for r in reqs)
    # every request against the FP64 oracle: first and last field of each sweep
    fx = load_fixture("serving40")
    errs = []
    if intervals == fx["intervals"]:
        for e in fx["requests"]:
            if e["request"] < n_requests:
                y = reqs[e["request"]]["result"]
                errs.append(float(np.abs(y[[0, -1]] - np.array(e["yields"])).max()))
    check = {"fixture": "tests/golden/serving40.json (FP64 oracle, first and last field of each request)",
             "requests_checked": len(errs), "max_abs_dPhi": max(errs) if errs else None, "tolerance": 1e-5}
    reduced = {"requests": n_requests, "intervals_per_request": intervals,
               "full_config": ">= 2000 requests of 1500-step sweeps (about 25 min of device time); this run is "
                              "the stream's first %d requests over %d steps of 2 ns" % (n_requests, intervals)}
    if not compare_plain:
        return {"requests": n_requests, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "state_steps_per_s": total / elapsed,
                "elapsed_s": elapsed, "method": method, "parity": check, "run_size": reduced,
                "shared_launches": shared[0], "requests_in_shared_launches": shared[1]}
    plain_s = 0.0
    for r in reqs:
        r["sys"].propagate(r["prm"], fields=[r["fields"][0]])  # first call per system uploads factors
        t1 = time.perf_counter()
        for f in r["fields"][1:3]:
            r["sys"].propagate(r["prm"], fields=[f])
        plain_s += (time.perf_counter() - t1) / 2 * len(r["fields"])
    plain = total / plain_s
    # small systems (D <= 512, dense path): the whole field sweep in one call against
    # one field per call, both timed in full on the same requests
    small = [r for r in reqs if r["sys"].dim <= 512]
    by_class = None
    if small:
        steps_small = sum(r["steps"] for r in small)
        t1 = time.perf_counter()
        for r in small:
            for f in r["fields"]:
                r["sys"].propagate(r["prm"], fields=[f])
        t_plain = time.perf_counter() - t1
        t1 = time.perf_counter()
        for r in small:
            r["sys"].propagate(r["prm"], fields=r["fields"])
        t_batched = time.perf_counter() - t1
        by_class = {"requests": len(small), "plain_state_steps_per_s": steps_small / t_plain,
                    "batched_state_steps_per_s": steps_small / t_batched, "batched_over_plain": t_plain / t_batched}
    return {"requests": n_requests, "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "state_steps_per_s": total / elapsed,
            "plain_one_at_a_time_state_steps_per_s": plain, "served_over_plain": total / elapsed / plain,
            "small_systems_D_le_512": by_class, "parity": check, "run_size": reduced,
            "shared_launches": shared[0], "requests_in_shared_launches": shared[1],
            "elapsed_s": elapsed, "intervals_per_request": intervals,
            "workload": "Poisson 200 req/s, seeded radical pairs (D 16..16384), 16-128 fields/request"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(packed, prm, field, columns, seconds):
    """FP64 oracle (the reference algorithm: Arnoldi m=16 + Pade expm on the
    Krylov matrix, CSR matvecs; dense Pade propagator on the dense workload) on
    a bounded sample of the same workload, in the two modes of the reference:
    one thread (HSGeneral's serial loop) and OpenMP over columns on the host's
    cores (at most 16, the box's CPU share).  `value` is the all-core figure."""
    from oracle import binding as ref
    nproc = os.cpu_count() or 1

    def timed(threads, cols, budget):
        if columns is not None:
            cols = min(cols, columns.shape[1])
        kw = dict(col_range=(0, cols), threads=threads)
        if field is not None:
            kw["field"] = field
        if columns is not None:
            kw["columns"] = columns
        steps = 2
        t0 = time.perf_counter()
        ref.run_columns(packed, prm, max_steps=steps, **kw)
        t1 = time.perf_counter() - t0
        steps = max(2, min(4000, int(steps * budget / max(t1, 1e-3))))
        t0 = time.perf_counter()
        ref.run_columns(packed, prm, max_steps=steps, **kw)
        el = time.perf_counter() - t0
        return cols * steps / el, cols, steps

    threads = min(16, nproc)
    v_all, c_all, s_all = timed(threads, 2 * threads, 0.6 * seconds)
    v_one, c_one, s_one = timed(1, 1, 0.4 * seconds)
    return {"value": v_all, "unit": "state-steps/s", "cores": threads, "kind": "port",
            "sample": "1 item x %d columns x %d steps (FP64 oracle, %d OpenMP threads over columns)"
                      % (c_all, s_all, threads),
            "single_thread": {"value": v_one, "unit": "state-steps/s", "cores": 1,
                              "sample": "1 item x %d column x %d steps" % (c_one, s_one)},
            "cpu_model": cpu_model(), "nproc": nproc}


if __name__ == "__main__":
    main()
