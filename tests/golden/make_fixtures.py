"""Writes the golden fixtures of tests/golden/ with the FP64 oracle.

The reference program cannot be built here (it needs Armadillo), so these
vectors come from the oracle, whose own pins (reference golden values, closed
forms, invariants) are checked in tests/test_oracle.py.  The fixtures let the
device tests compare against fixed numbers without re-running the FP64 oracle
at the larger sizes, and let the CPU suite detect any drift of the oracle.

    python tests/golden/make_fixtures.py [names...]

Fixtures:
  fib1024.json     the 1024-point Fibonacci sphere at |B| = 0.05 mT with the
                   reference's float truncation (RunSection/Utility.cpp:44-79)
  cfg1.json        config 1 (D = 48): yields and the 5001-point singlet trace
  cfg2_subset.json config 2 (D = 4096): 16 orientations, 0-50 ns
  cfg3_subset.json config 3 (D = 65536): 32 spin-coherent trajectories, 0-200 ns,
                   ST dephasing (flip stream seed 2026)
  symdecay_d512.json  infinite-time symmetric-decay yields of a 7-nucleus
                   anisotropic pair (D = 512) at two field strengths
  cfg2_bench.json  config 2 (D = 4096): 8 orientations spread over the sphere, all
                   1024 direct columns, 0-100 ns (the horizon of one bench step):
                   transition yields and their per-sample traces
  cfg3_full.json   config 3 (D = 65536): 8 spin-coherent trajectories over the full
                   0-2000 ns grid (2000 steps), ST dephasing (flip seed 2026)
  cfg4_uneven.json one config-4 pair per nuclear count n = 2 ... 12 (D = 16 ... 16384), two
                   field strengths, all direct columns, 0-20 ns: transition yields and traces
  serving40.json   the first 40 requests of the serving stream (D = 16 ... 16384): transition
                   yields at the first and last field of each sweep over 5 steps of 2 ns
"""
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from molspin_amd import abi, workloads  # noqa: E402
from oracle import binding as ref  # noqa: E402

THREADS = min(16, os.cpu_count() or 1)
CFG2_ORIENT = list(range(0, 1024, 64))
CFG2_TOTAL = 50.0
CFG3_M = 32
CFG3_TOTAL = 200.0
CFG3_SEED = 2026
CFG2_BENCH_ORIENT = list(range(0, 1024, 128))
CFG2_BENCH_TOTAL = 100.0
CFG3_FULL_M = 8
SERVING_N = 40
CFG4_FIELDS = (0, 40)
CFG4_TOTAL = 20.0
SERVING_INTERVALS = 5


def fib1024():
    _, cfg = workloads.config2(n_orientations=1024)
    return {"n": 1024, "magnitude_mT": 0.05, "points": [list(f) for f in cfg["fields"]]}


def cfg1():
    s, cfg = workloads.config1()
    p = s.pack()
    prm = abi.run_params(total_time=cfg["total_time"], time_step=cfg["time_step"], method=abi.MSG_METHOD_KRYLOV)
    y, _ = ref.run_hsgeneral(p, prm, threads=THREADS)
    prm_t = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                           total_time=cfg["total_time"], time_step=cfg["time_step"], method=abi.MSG_METHOD_KRYLOV)
    _, tr = ref.run_hsgeneral(p, prm_t, trace=True, threads=THREADS)
    return {"total_time": cfg["total_time"], "time_step": cfg["time_step"], "krylov_size": 16,
            "yields": y[0].tolist(), "state_names": [st["name"] for st in s.states],
            "trace": tr[0].T.tolist()}


def cfg2_subset():
    s, cfg = workloads.config2(n_orientations=1024)
    p = s.pack()
    fields = [cfg["fields"][i] for i in CFG2_ORIENT]
    prm = abi.run_params(total_time=CFG2_TOTAL, time_step=cfg["time_step"], method=abi.MSG_METHOD_KRYLOV)
    y, _ = ref.run_hsgeneral(p, prm, fields=fields, threads=THREADS)
    return {"orientations": CFG2_ORIENT, "total_time": CFG2_TOTAL, "time_step": cfg["time_step"],
            "krylov_size": 16, "yields": y.tolist()}


def cfg3_subset():
    s, cfg = workloads.config3(n_traj=CFG3_M)
    p = s.pack()
    cols = ref.coherent_samples(p, s.state_index("S"), cfg["sample_seed"], CFG3_M) / math.sqrt(CFG3_M)
    prm = abi.run_params(total_time=CFG3_TOTAL, time_step=cfg["time_step"], method=abi.MSG_METHOD_KRYLOV,
                         sampling=abi.MSG_SAMPLING_GIVEN, seed=CFG3_SEED)
    y, _ = ref.run_hsgeneral(p, prm, columns=cols, threads=THREADS)
    return {"trajectories": CFG3_M, "sample_seed": cfg["sample_seed"], "flip_seed": CFG3_SEED,
            "total_time": CFG3_TOTAL, "time_step": cfg["time_step"], "krylov_size": 16,
            "yields": y[0].tolist()}


def cfg2_bench():
    s, cfg = workloads.config2(n_orientations=1024)
    p = s.pack()
    fields = [cfg["fields"][i] for i in CFG2_BENCH_ORIENT]
    prm = abi.run_params(total_time=CFG2_BENCH_TOTAL, time_step=cfg["time_step"], method=abi.MSG_METHOD_KRYLOV)
    y, tr = ref.run_hsgeneral(p, prm, fields=fields, threads=THREADS, trace=True)
    return {"orientations": CFG2_BENCH_ORIENT, "total_time": CFG2_BENCH_TOTAL, "time_step": cfg["time_step"],
            "krylov_size": 16, "yields": y.tolist(), "trace": tr.tolist()}


def cfg3_full():
    s, cfg = workloads.config3(n_traj=CFG3_FULL_M)
    p = s.pack()
    cols = ref.coherent_samples(p, s.state_index("S"), cfg["sample_seed"], CFG3_FULL_M) / math.sqrt(CFG3_FULL_M)
    prm = abi.run_params(total_time=cfg["total_time"], time_step=cfg["time_step"], method=abi.MSG_METHOD_KRYLOV,
                         sampling=abi.MSG_SAMPLING_GIVEN, seed=CFG3_SEED)
    y, tr = ref.run_hsgeneral(p, prm, columns=cols, threads=THREADS, trace=True)
    return {"trajectories": CFG3_FULL_M, "sample_seed": cfg["sample_seed"], "flip_seed": CFG3_SEED,
            "total_time": cfg["total_time"], "time_step": cfg["time_step"], "krylov_size": 16,
            "yields": y[0].tolist(), "trace_every_10": tr[0, ::10].tolist()}


def cfg4_pairs():
    """The first config-4 pair index of each nuclear count 2 ... 12."""
    out = {}
    i = 0
    while len(out) < 11:
        n = workloads.config4_pair(i)[1]["n"]
        out.setdefault(n, i)
        i += 1
    return [out[n] for n in range(2, 13)]


def cfg4_uneven():
    pairs = []
    for idx in cfg4_pairs():
        s, cfg = workloads.config4_pair(idx)
        p = s.pack()
        fields = [cfg["fields"][i] for i in CFG4_FIELDS]
        prm = abi.run_params(total_time=CFG4_TOTAL, time_step=cfg["time_step"], method=abi.MSG_METHOD_KRYLOV)
        t0 = time.perf_counter()
        y, tr = ref.run_hsgeneral(p, prm, fields=fields, threads=THREADS, trace=True)
        print("  pair %d (n = %d): %.1f s" % (idx, cfg["n"], time.perf_counter() - t0), flush=True)
        pairs.append({"pair": idx, "n": cfg["n"], "fields": [list(f) for f in fields], "yields": y.tolist(),
                      "trace": tr.tolist()})
    return {"total_time": CFG4_TOTAL, "time_step": 2.0, "krylov_size": 16, "pairs": pairs}


def serving40():
    out = []
    for i in range(SERVING_N):
        s, cfg, _ = workloads.serving_request(i)
        p = s.pack()
        dt = cfg["time_step"]
        prm = abi.run_params(total_time=SERVING_INTERVALS * dt, time_step=dt, method=abi.MSG_METHOD_KRYLOV)
        fields = [cfg["fields"][0], cfg["fields"][-1]]
        y, _ = ref.run_hsgeneral(p, prm, fields=fields, threads=THREADS)
        out.append({"request": i, "n": cfg["n"], "fields": [list(f) for f in fields], "yields": y.tolist()})
    return {"intervals": SERVING_INTERVALS, "requests": out}


def symdecay_d512():
    idx = next(i for i in range(1000) if workloads.config4_pair(i)[1]["n"] == 7)
    s, cfg = workloads.config4_pair(idx)
    p = s.pack()
    fields = [cfg["fields"][0], cfg["fields"][40]]
    y = [ref.symmetric_decay(p, cfg["k"], f).tolist() for f in fields]
    return {"pair": idx, "rate": cfg["k"], "fields": [list(f) for f in fields],
            "states": [st["name"] for st in s.states], "yields": y}


MAKERS = {"fib1024": fib1024, "symdecay_d512": symdecay_d512, "cfg1": cfg1, "cfg2_subset": cfg2_subset,
          "cfg3_subset": cfg3_subset, "cfg2_bench": cfg2_bench, "cfg3_full": cfg3_full, "serving40": serving40,
          "cfg4_uneven": cfg4_uneven}


def main(names):
    for name in names or list(MAKERS):
        t0 = time.perf_counter()
        data = MAKERS[name]()
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print("%s written (%.1f s)" % (name, time.perf_counter() - t0), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
