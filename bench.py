"""Benchmark: median prover milliseconds for one Labinius proof at 2^31 evaluations on one MI355X.

One step = one whole proof (commit, transcript, partials, switch, row evaluation, challenges,
fold, proof bytes) over a table of 2^31 Boolean evaluations already resident in HBM; the table of
each step comes from its own 32-byte seed and is written on the device before the step's clock
starts (the reference's harness also writes the table outside the timer). The default modulus is
the reference's Q = 35,039,150,080.

Besides the JSON line it prints per-pass kernel times (HIP events on the library's stream), the
dominant kernel's roofline, and a CPU baseline: the scalar C oracle (same algorithm, port) timed on
a bounded sample (a full proof at 2^23 evaluations) and scaled by the work ratio to 2^31.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--log-vars 31] [--no-cpu-baseline]
"""
import argparse
import hashlib
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # int32 VALU: 256 CUs x 64 lanes x 2.4 GHz = 39.3 Tops/s
# Integer ops of the commit transform per (ring element, CRT prime): 3,242 Montgomery
# multiplications (2,594 twiddle/cube-root products of the 2,2,3,3,3,3 DIF + 648 MAC) at 4 ops and
# 7,344 modular additions/subtractions at 2 ops.
COMMIT_OPS_PER_ELEMENT_PRIME = 3242 * 4 + 7344 * 2


def bench_seed(i: int) -> bytes:
    return hashlib.sha256(b"labinius-mi355x/bench-seed/v1" + i.to_bytes(8, "little")).digest()


def cpu_baseline(log_vars_target: int):
    """The oracle's whole proof at k=23, scaled to the target shape (bounded sample, ~1-30 s)."""
    import oracle_ctypes as O
    O.build()
    k = 23
    seed = bench_seed(10_000)
    tab = O.table(seed, k)
    t0 = time.perf_counter()
    O.prove(seed, k, O.DEFAULT_Q, tab)
    dt = time.perf_counter() - t0
    scale = 2 ** (log_vars_target - k)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(dt * 1000 * scale, 1), "unit": "ms", "cores": cores, "kind": "port",
            "sample": f"CPU oracle whole proof at k={k} ({dt * 1000:.0f} ms, OpenMP), x{scale} work ratio to k={log_vars_target}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log-vars", type=int, default=31)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the prover's kernels run on one device per proof; "
              "a multi-GPU sharding is described in DESIGN.md but not built (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import amd
    m = amd.load()
    k = args.log_vars
    p = m.Prover(k)
    p.set_modulus(m.DEFAULT_Q)
    for w in range(args.warmup):      # with the pass timers on, so their first use is not timed
        p.generate_table(bench_seed(1000 + w))
        p.set_timing(True)
        p.prove(bench_seed(1000 + w))
        p.set_timing(False)

    # The proof lands in one caller-owned host buffer, allocated and touched once up front.
    out = np.zeros(m.proof_len(k, m.DEFAULT_Q), dtype=np.uint8)
    times, profs, timings = [], [], []
    for i in range(args.steps):
        seed = bench_seed(i)
        p.generate_table(seed)          # input generation, outside the clock
        p.set_timing(True)
        t0 = time.perf_counter()
        n, prof = p.prove_into(seed, out)  # synchronous: returns with the proof bytes in `out`
        times.append((time.perf_counter() - t0) * 1000)
        timings.append(p.timing())
        profs.append(prof)
        p.set_timing(False)
    med = statistics.median(times)
    mean_t = {f: statistics.mean(t[f] for t in timings) for f in timings[0]}
    n_words = 1 << (k - 7)
    re = (1 << (k - 17)) // 4
    table_bytes = n_words * 16
    primes = p.commit_primes()
    passes = {
        "commit_mac": {"ms": mean_t["commit_mac_ms"], "bound": "valu",
                       "achieved_tops": re * 1024 * primes * COMMIT_OPS_PER_ELEMENT_PRIME / (mean_t["commit_mac_ms"] / 1e3) / 1e12},
        "partials": {"ms": mean_t["eq_ms"] + mean_t["partials_ms"], "bound": "hbm",
                     "achieved_gbs": table_bytes / ((mean_t["eq_ms"] + mean_t["partials_ms"]) / 1e3) / 1e9},
        "switch": {"ms": mean_t["switch0_ms"] + mean_t["switch_rounds_ms"], "bound": "hbm",
                   "achieved_gbs": (table_bytes + 6_291_456 + 18_874_368) / ((mean_t["switch0_ms"] + mean_t["switch_rounds_ms"]) / 1e3) / 1e9},
        "row": {"ms": mean_t["row_ms"], "bound": "hbm", "achieved_gbs": table_bytes / (mean_t["row_ms"] / 1e3) / 1e9},
        "fold": {"ms": mean_t["fold_ms"], "bound": "hbm",
                 "achieved_gbs": (table_bytes + re * 648 * 4) / (mean_t["fold_ms"] / 1e3) / 1e9},
        "commit_finish": {"ms": mean_t["commit_finish_ms"], "bound": "valu"},
    }
    dom_name = max(passes, key=lambda n: passes[n]["ms"])
    dom = passes[dom_name]
    if dom["bound"] == "valu" and dom_name == "commit_mac":
        roofline = {"bound": "valu", "kernel": "k_commit_mac", "achieved": round(dom["achieved_tops"], 3),
                    "peak": round(VALU_PEAK_TOPS, 1), "unit": "Tops/s",
                    "frac": round(dom["achieved_tops"] / VALU_PEAK_TOPS, 4), "traffic": None}
    else:
        roofline = {"bound": dom["bound"], "kernel": dom_name, "achieved": round(dom.get("achieved_gbs", 0), 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom.get("achieved_gbs", 0) / HBM_PEAK_GBS, 4),
                    "traffic": None}
    line = {
        "metric": "prover ms (commit+open, 2^31 evals) on 1 MI355X",
        "value": round(med, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(statistics.mean(times), 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32/gf2", "data": "synthetic (BLAKE3 tables from fixed seeds)",
        "config": {"workload": f"one proof, 2^{k} Boolean evaluations (2^{k - 7} GHASH words), Q=35039150080, "
                               f"{primes} CRT primes", "log_vars": k, "modulus": m.DEFAULT_Q},
        "roofline": roofline,
        "passes_ms": {n: round(v["ms"], 3) for n, v in passes.items()},
        "host_phases_ms": {f.replace("_us", ""): round(statistics.mean(pr[f] for pr in profs) / 1000, 3)
                           for f in profs[0]},
        "step_ms": [round(t, 3) for t in times],
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(k)
    print(json.dumps(line), flush=True)
    p.close()


if __name__ == "__main__":
    main()
