"""Benchmark: median prover milliseconds for one Labinius proof at 2^31 evaluations on one MI355X.

One step = one whole proof (commit, transcript, partials, switch, row evaluation, challenges,
fold, proof bytes) over a table of 2^31 Boolean evaluations already resident in HBM; the table of
each step comes from its own 32-byte seed and is written on the device before the step's clock
starts (the reference's harness also writes the table outside the timer). The default modulus is
the reference's Q = 35,039,150,080.

Besides the JSON line it prints per-pass kernel times (HIP events on the library's stream), the
dominant kernel's roofline, and whole-process timings made before this process touches the GPU:
the CPU baseline (the oracle's command line, the same algorithm in scalar C with OpenMP, proving at
the same 2^31 shape from a table file on all host cores and on 8) and the labinius-fast executable,
both spawn to exit.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--log-vars 31] [--no-cpu-baseline] [--no-batched]
                    [--no-commit-micro]
    python bench.py --serving [--requests 10000] [--rate 400] [--workers 8]
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
# Proofs in flight each run on streams of their own; HIP maps a process's streams onto
# GPU_MAX_HW_QUEUES hardware queues (HIP's default is 4), and kernels of streams sharing a queue
# run one after another. The runtime reads the variable when it initialises, so main() sets it
# (--hw-queues) before the library loads.
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
VALU_SLOTS_T = 256 * 64 * 2.4e9 / 1e12  # int32 VALU issue: 256 CUs x 64 lanes x 2.4 GHz = 39.3 T/s
# Commit roofline in the algorithmic unit of the reference's transform: per ring element and CRT
# prime, 324 (Phi6 split) + 2*324 (radix 2) + 4*216*3 (radix 3 with the cube root) = 3,564 modular
# products plus 648 MAC = 4,212 "mulmod-eq". Peak: the int32 VALU issue rate at 3 slots per
# Montgomery product (two 32x32->64 multiply-adds and one 32-bit multiply, each one issue slot on
# gfx950: a throughput probe measured chains of them at 0.86-0.95 of that rate,
# profiles/valu_multiply_rates.md).
MULMOD_PER_ELEMENT_PRIME = 4212
MULMOD_SLOTS = 3
MULMOD_PEAK_T = VALU_SLOTS_T / MULMOD_SLOTS
# Moduli below 2^39 commit on the matrix cores: per (ring element, column) the product of the
# matrix's rotations (648 coefficients x 512 witness bits) in 5 int8 planes, 2 ops per term
# (multiply, add), against MI355X's dense int8 peak.
INT8_PEAK_T = 5000.0
TOEP_OPS_PER_ELEMENT_COLUMN = 2 * 648 * 512


# Commit micro-benchmark: 64 columns x 4096 binary ring elements (a 16 MiB witness from seed 0),
# Q = 3889 * 9721 * 17497, keys of 1 and 16 ring rows. The one-row commitment is one VALU kernel;
# its roofline counts int32 VALU issue slots per (ring element, column, CRT prime): the
# transform's 3,564 modular products at 3 slots each and the key row's 648 multiply-adds at 1 slot
# (one 32x32->64 multiply-add). With 16 rows the transform (the same 3,564 x 3 slots) is
# the larger kernel; the key-row products run on the matrix cores and are priced in bytes: the
# transformed witness (4 bytes per slot, column and element) and the key's digits (4 bytes per
# slot, row and element) read once per CRT prime.
Q_C3 = 661_473_542_593
C3_COLUMNS, C3_ELEMENTS = 64, 4096


def c3_slots(rows: int) -> int:
    return 3564 * MULMOD_SLOTS + rows * 648


def c3_product_bytes(rows: int, primes: int) -> int:
    return primes * C3_ELEMENTS * 648 * 4 * (C3_COLUMNS + 16)


def bench_seed(i: int) -> bytes:
    return hashlib.sha256(b"labinius-mi355x/bench-seed/v1" + i.to_bytes(8, "little")).digest()


REFERENCE_PUBLISHED_MS = 2972.27  # .yukon/README.md:162-178 (older baseline, Xeon 8488C, 8 CPUs, Q=55,144,219,393)
ORACLE_CLI = os.path.join(ROOT, "oracle", "build", "labinius-oracle")
GPU_CLI = os.path.join(ROOT, "labinius-fast-challenge_amd", "build", "labinius-fast")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed_process(cmd, env, cpus=None):
    """Wall-clock milliseconds from spawn to exit (the harness's timer, .yukon/benchmark.sh:142-162)
    and the process's stderr."""
    import subprocess
    pre = ["taskset", "-c", cpus] if cpus else []
    t0 = time.perf_counter()
    r = subprocess.run(pre + cmd, env=env, capture_output=True, text=True, timeout=900)
    dt = (time.perf_counter() - t0) * 1000
    if r.returncode != 0:
        raise RuntimeError(f"{os.path.basename(cmd[0])} exited {r.returncode}: {r.stderr[-400:]}")
    return dt, r.stderr


def process_runs(k, cpu_seeds, gpu_seeds, want_cpu):
    """Whole-process timings at 2^k evaluations, before this process touches the GPU: the table is
    written to a file first (outside the timer, as the harness does), then each prover runs as a
    child process reading it, spawn to exit. CPU baseline: the oracle's command line (the same
    algorithm in scalar C with OpenMP) on all host cores and on 8 (taskset -c 0-7); GPU: the
    labinius-fast executable (process start, device set-up, table read and upload, matrix, proof,
    proof write). Returns (cpu_baseline record or None, process record)."""
    import shutil
    import tempfile
    tmp = tempfile.mkdtemp(prefix="lbn-bench-")
    try:
        table = os.path.join(tmp, "witness-table.bin")
        with open(table, "wb") as f:
            f.write(np.random.default_rng(20261204).bytes((1 << k) // 8))
        base = dict(os.environ, LABINIUS_LOG_VARS=str(k))
        nproc = os.cpu_count() or 1
        cpu = None
        if want_cpu and os.path.exists(ORACLE_CLI):
            runs_all, runs_8 = [], []
            for i in range(cpu_seeds):
                env = dict(base, LABINIUS_SEED=bench_seed(i).hex(), OMP_NUM_THREADS=str(nproc))
                runs_all.append(timed_process([ORACLE_CLI, "--prove", os.path.join(tmp, "cpu.bin"), "--table", table],
                                              env)[0])
                env8 = dict(env, OMP_NUM_THREADS="8")
                if nproc >= 8 and shutil.which("taskset"):
                    runs_8.append(timed_process([ORACLE_CLI, "--prove", os.path.join(tmp, "cpu8.bin"), "--table", table],
                                                env8, cpus="0-7")[0])
            cpu = {"value": round(statistics.median(runs_all), 1), "unit": "ms", "cores": nproc, "kind": "port",
                   "sample": f"the oracle's command line (scalar C + OpenMP restatement), whole proof at k={k} "
                             f"from a table file, spawn to exit, {cpu_seeds} seed(s), median",
                   "cpu_model": cpu_model(), "nproc": nproc, "runs_ms": [round(t, 1) for t in runs_all],
                   "cores8_ms": round(statistics.median(runs_8), 1) if runs_8 else None,
                   "reference_published_ms": REFERENCE_PUBLISHED_MS}
        gpu = None
        if os.path.exists(GPU_CLI):
            runs, prof = [], None
            for i in range(gpu_seeds):
                env = dict(base, LABINIUS_SEED=bench_seed(i).hex())
                dt, err = timed_process([GPU_CLI, "--prove", os.path.join(tmp, "gpu.bin"), "--table", table], env)
                runs.append(dt)
                line = [x for x in err.splitlines() if x.startswith("labinius-profile")]
                if line:
                    prof = {kv.split("=")[0]: float(kv.split("=")[1]) for kv in line[-1].split()[1:]}
            gpu = {"median_ms": round(statistics.median(runs), 1), "runs_ms": [round(t, 1) for t in runs],
                   "what": "labinius-fast --prove P --table T, spawn to exit (process start, device set-up, table "
                           "read and upload, public matrix, proof, proof write)", "last_profile_us": prof}
        return cpu, gpu
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def pmc_traffic(pmc_dir, log_vars, kernel="k_commit_mac", copy_kernel="k_copy_u4"):
    """HBM bytes per launch of `kernel` from rocprofv3 counter CSVs: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch), each scaled so that the copy kernel (table bytes read once and written once) reads exactly
    its size. The launch is the kernel's largest dispatch: the whole-table commitment (the bench's
    deferred-upload proofs launch the MAC once per copied quarter of the ring elements)."""
    import csv
    import glob

    def per_kernel(sub, counter):
        vals = {}
        for f in glob.glob(os.path.join(pmc_dir, sub, "*counter_collection.csv")):
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] == counter:
                    vals.setdefault(r["Kernel_Name"].split("(")[0], []).append(float(r["Counter_Value"]) * 1024)
        return vals

    fetch, write = per_kernel("fetch", "FETCH_SIZE"), per_kernel("write", "WRITE_SIZE")
    if kernel not in fetch or kernel not in write or copy_kernel not in fetch or copy_kernel not in write:
        return None, None
    size = (1 << log_vars) // 8
    f_scale, w_scale = size / max(fetch[copy_kernel]), size / max(write[copy_kernel])
    f = max(fetch[kernel]) * f_scale
    w = max(write[kernel]) * w_scale
    return round(f + w), {"fetch_bytes": round(f), "write_bytes": round(w), "fetch_scale": round(f_scale, 4),
                          "write_scale": round(w_scale, 4)}


def committed_traffic(kernel):
    """The commit kernel's HBM bytes per launch at k = 31 as the last counter runs of this build measured
    them (the rocprofv3 FETCH_SIZE / WRITE_SIZE passes and copy-kernel scales written into
    profiles/<kernel>_traffic.txt, labelled "committed counters"; a run with --pmc-dir measures them
    afresh)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.txt")
    try:
        kv = dict(line.split("=", 1) for line in open(path).read().split() if "=" in line)
        f, w = int(kv["fetch_bytes"]), int(kv["write_bytes"])
        return f + w, {"fetch_bytes": f, "write_bytes": w, "fetch_scale": float(kv["fetch_scale"]),
                       "write_scale": float(kv["write_scale"]),
                       "source": f"committed counters: profiles/{kernel}_traffic.txt"}
    except (OSError, KeyError, ValueError):
        return None, None


def batched_instances(n=64, first=100):
    """The batched workload: instance i has seed bench_seed(first + i) and k = 20 + seed[0] mod 9."""
    return [(20 + bench_seed(first + i)[0] % 9, bench_seed(first + i)) for i in range(n)]


class ProofArena:
    """Host buffers for a list of proofs, allocated and written once before any clock so that no
    timed proof pays page faults on fresh memory: one PinnedBuffer (pinned=True: the proofs' copies
    write it directly) or one pageable array, cut into one 256-byte-aligned slice per proof. The
    paths compared on a workload get the same arena; proofs are taken out as bytes before close()."""

    def __init__(self, m, sizes, pinned):
        offs = [0]
        for n in sizes:
            offs.append(offs[-1] + (n + 255) // 256 * 256)
        self.pinned = m.PinnedBuffer(offs[-1]) if pinned else None
        arr = self.pinned.array if pinned else np.empty(offs[-1], dtype=np.uint8)
        arr[:] = 0
        self.views = [arr[offs[i]:offs[i] + n] for i, n in enumerate(sizes)]

    def close(self):
        self.views = None
        if self.pinned:
            self.pinned.close()


def warm(sch, ks, workers):
    """Context set-up outside the clock: `workers` contexts of each size (an int, or a count per size),
    then a round of proofs (first launches of every kernel)."""
    count = workers if isinstance(workers, dict) else {kk: workers for kk in ks}
    for kk in ks:
        sch.prewarm(kk, count[kk])
    hs = [sch.submit(kk, bench_seed(999_000 + 100 * kk + j)) for kk in ks for j in range(count[kk])]
    sch.drain()
    for h in hs:
        sch.result(h)


def run_jobs(m, jobs, workers, arena, reps=1, budget=200 << 30):
    """All jobs submitted at once to a scheduler with `workers` proofs in flight (contexts warmed
    first), timed `reps` times on the same scheduler, job i's proof into arena.views[i]; returns
    (proofs of every run, seconds per run)."""
    sch = m.Scheduler(m.DEFAULT_Q, hbm_budget=budget, workers=workers)
    try:
        warm(sch, sorted({kk for kk, _ in jobs}), workers)
        proofs, times = [], []
        for rep in range(reps):
            w0 = time.monotonic_ns()
            t0 = time.perf_counter()
            hs = [sch.submit(kk, seed, out=out) for (kk, seed), out in zip(jobs, arena.views)]
            sch.drain()
            times.append(time.perf_counter() - t0)
            if rep == 0:  # the first timed batch, on the kernel tracer's clock
                run_jobs.window_ns = (w0, time.monotonic_ns())
            proofs.append([sch.result(h)[0] for h in hs])
        return proofs, times
    finally:
        sch.close()


def fold_fill(k, proof):
    """||v||^2 of the proof's fold over the checker's cap at this shape, FOLD_CAP (108) per coefficient
    of the witness_len / 4 * N fold terms (verifier/proof_verify/src/main.rs:58-64); the fold is the
    proof's last 4 * 648 * 2^(k-19) bytes, i32 little-endian."""
    v = np.frombuffer(proof[len(proof) - 4 * 648 * (1 << (k - 19)):], dtype="<i4").astype(np.int64)
    return float((v * v).sum()) / float(np.ceil(108.0 * ((1 << (k - 7)) // 4 * 648)))


class Checker:
    """The library's checker (lbn_verify), one context per size; call(k, seed, proof) -> accepted.
    Rejections are kept in `rejected` as (k, reason, fold fill)."""

    def __init__(self, m):
        self.m, self.ctx, self.rejected = m, {}, []

    def __call__(self, k, seed, proof):
        if k not in self.ctx:
            self.ctx[k] = self.m.Prover(k)
        why = self.ctx[k].verify(seed, proof)
        if why is not None:
            self.rejected.append({"k": k, "seed": seed.hex(), "reason": why,
                                  "fold_fill": round(fold_fill(k, proof), 4)})
        return why is None

    def close(self):
        for p in self.ctx.values():
            p.close()
        self.ctx.clear()


def run_batch_api(m, jobs, threads, arena, reps):
    """The batch prover (lbn_prove_batch): one context per instance, created with its modulus outside
    the clock, then `reps` timed batches (each generates its tables on the device, as the scheduler's
    jobs do), proof i into arena.views[i]; returns (proofs of every run, seconds per run)."""
    provers = []
    try:
        for kk, _ in jobs:
            p = m.Prover(kk)
            p.set_modulus(m.DEFAULT_Q)
            provers.append(p)
        seeds = [seed for _, seed in jobs]
        outs = arena.views
        res = m.prove_batch(provers, seeds, outs, threads=threads)  # first launches of every kernel
        assert all(st == 0 for _, st in res), res
        proofs, times = [], []
        for rep in range(reps):
            w0 = time.monotonic_ns()
            t0 = time.perf_counter()
            res = m.prove_batch(provers, seeds, outs, threads=threads)
            times.append(time.perf_counter() - t0)
            if rep == 0:
                run_batch_api.window_ns = (w0, time.monotonic_ns())
            assert all(st == 0 for _, st in res), res
            proofs.append([o[:n].tobytes() for o, (n, _) in zip(outs, res)])
        return proofs, times
    finally:
        for p in provers:
            p.close()


def tune_shapes(m, ks, reps=3):
    """lbn_autotune for each table size in ks (a context of that size with the table of bench_seed(k)):
    the commitment's split count per shape, recorded for this process's later proofs; returns
    {k: {"best": splits, "ms": {splits: median ms}}}."""
    out = {}
    for kk in ks:
        p = m.Prover(kk)
        try:
            p.set_modulus(m.DEFAULT_Q)
            p.generate_table(bench_seed(kk))
            out[kk] = p.autotune(reps=reps)
        finally:
            p.close()
    return out


def batched(m, workers, reps=3, threads=16):
    """64 independent proofs of mixed size (k in [20, 28]): one at a time, through the scheduler with
    `workers` proofs in flight, and through the batch prover (lbn_prove_batch on `threads` host
    threads: the commitments per proof, then one device sequence for all switches and one launch each
    for all rows, challenges and folds, in two staggered groups); each way timed `reps` times (the
    batch is short), medians reported, proofs into the same pinned buffers. The headline is the
    batch prover's."""
    jobs = batched_instances()
    tuned = tune_shapes(m, sorted({kk for kk, _ in jobs}))  # launch shapes for every way below
    arena = ProofArena(m, [m.proof_len(kk, m.DEFAULT_Q) for kk, _ in jobs], pinned=True)
    try:  # the same pinned proof buffers for the three ways
        seq, t_seqs = run_jobs(m, jobs, 1, arena, reps)
        w_seq = run_jobs.window_ns
        par, t_pars = run_jobs(m, jobs, workers, arena, reps)
        w_par = run_jobs.window_ns
        bat, t_bats = run_batch_api(m, jobs, threads, arena, reps)
    finally:
        arena.close()
    batched.windows_ns = {"sequential": list(w_seq), "scheduler": list(w_par), "batch": list(run_batch_api.window_ns)}
    t_seq, t_par, t_bat = statistics.median(t_seqs), statistics.median(t_pars), statistics.median(t_bats)
    check = Checker(m)
    try:  # every proof of the first batch-prover run through the checker (the others are compared to it)
        verified = sum(check(kk, seed, proof) for (kk, seed), proof in zip(jobs, bat[0]))
    finally:
        check.close()
    n = len(jobs)
    return {"instances": n, "k": sorted({kk for kk, _ in jobs}), "path": "lbn_prove_batch", "threads": threads,
            "proofs_per_s": round(n / t_bat, 1), "sequential_proofs_per_s": round(n / t_seq, 1),
            "speedup": round(t_seq / t_bat, 2), "runs": reps,
            "proofs_per_s_runs": [round(n / t, 1) for t in t_bats],
            "sequential_proofs_per_s_runs": [round(n / t, 1) for t in t_seqs],
            "scheduler": {"workers": workers, "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                          "proofs_per_s": round(n / t_par, 1), "speedup": round(t_seq / t_par, 2),
                          "proofs_per_s_runs": [round(n / t, 1) for t in t_pars]},
            "identical_to_sequential": all(p == seq[0] for p in seq + par + bat),
            "verified": f"{verified}/{n}", "rejected": check.rejected,
            "proof_buffers": "one pinned arena (lbn_host_alloc), the same for the three ways",
            "mac_splits": {kk: v["best"] for kk, v in tuned.items()}}


def serving_requests(requests, rate, rng_seed=20261203):
    """The serving workload: (arrival s, k, seed) per request, Poisson arrivals at `rate` per second,
    k in {20..31} with P(k) ~ 2^-(k-20), request i's seed bench_seed(2_000_000 + i)."""
    import random
    rng = random.Random(rng_seed)
    ks = list(range(20, 32))
    wts = [2.0 ** -(kk - 20) for kk in ks]
    reqs, t = [], 0.0
    for i in range(requests):
        t += rng.expovariate(rate)
        reqs.append((t, rng.choices(ks, wts)[0], bench_seed(2_000_000 + i)))
    return reqs


SERVING_KS = list(range(20, 32))
BATCH_WORKERS, BATCH_MAX, BATCH_THREADS = 3, 64, 8  # continuous batching in the serving loop


def serving_arena(m, reqs):
    """Pageable proof buffers for the serving workload (written once up front; several GB, so not
    pinned), shared by the ways it is served."""
    return ProofArena(m, [m.proof_len(kk, m.DEFAULT_Q) for _, kk, _ in reqs], pinned=False)


def serve(m, reqs, w, budget, arena, check=None, batch=1):
    """The arrival sequence `reqs` served by a scheduler with `w` workers (contexts of every size
    warmed first; with batch > 1, continuous batching of up to `batch` admitted jobs per worker),
    request i's proof into arena.views[i]; latency = completion - arrival. With `check`, every proof
    goes through the checker after the clock. Returns (figures, the first 32 proofs)."""
    sch = m.Scheduler(m.DEFAULT_Q, hbm_budget=budget, workers=w)
    try:
        # with batching, as many contexts of the small sizes as a batch may hold (the large ones are rare)
        warm(sch, SERVING_KS, {kk: min(w * batch, 64) if kk <= 23 else w for kk in SERVING_KS})
        if batch > 1:
            sch.set_batch(batch, BATCH_THREADS)
        hs = []
        t0 = time.perf_counter()
        for (at, kk, seed), out in zip(reqs, arena.views):
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            lag = (time.perf_counter() - t0) - at  # submission behind the scheduled arrival
            hs.append((lag, sch.submit(kk, seed, out=out)))
        sch.drain()
        wall = time.perf_counter() - t0
        lat = sorted(lag * 1000 + (h.queue_us + h.run_us) / 1000 for lag, h in hs)
        st = sch.stats()
        proofs = [sch.result(h)[0] for _, h in hs[:32]]  # a sample, for the isolation check
        out = {"p50_ms": round(statistics.median(lat), 3), "p99_ms": round(lat[int(0.99 * (len(lat) - 1))], 3),
               "proofs_per_s": round(len(hs) / wall, 1), "wall_s": round(wall, 2),
               "contexts": st["contexts_created"], "evictions": st["evictions"],
               "peak_gib": round(st["peak_bytes"] / 2**30, 2)}
        if check:  # every proof through the checker after the clock, each released once checked
            def every_proof():
                yield from proofs
                for _, h in hs[32:]:
                    yield sch.result(h)[0]
            ok = sum(check(kk, seed, p) for (_, kk, seed), p in zip(reqs, every_proof()))
            out["verified"] = f"{ok}/{len(hs)}"
            # An honest fold can cross the cap (its sum of squares has a long tail): the proof is
            # still the reference algorithm's bytes, and the checker rejects it as the
            # reference's does. Each rejection is listed with its reason and fill.
            out["rejected"] = check.rejected
        return out, proofs
    finally:
        sch.close()


def serving(m, requests, rate, workers, budget):
    """Poisson arrivals of proving requests, k in {20..31} with P(k) ~ 2^-(k-20), fresh seed per
    request. The same arrival sequence is served by `workers` proofs in flight and by one at a time."""
    reqs = serving_requests(requests, rate)
    arena = serving_arena(m, reqs)
    checks = [Checker(m), Checker(m)]  # one per way, each with its own list of rejections
    try:
        mix, proofs_mix = serve(m, reqs, workers, budget, arena, checks[0])
        checks[0].close()
        cb, proofs_cb = serve(m, reqs, BATCH_WORKERS, budget, arena, checks[1], batch=BATCH_MAX)
        checks[1].close()
        seq, proofs_seq = serve(m, reqs, 1, budget, arena)
    finally:
        for c in checks:
            c.close()
        arena.close()
    return {"metric": "serving p50/p99 ms and proofs/s on 1 MI355X", "requests": requests, "rate_per_s": rate,
            "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
            "workers": workers, "hbm_budget_gib": budget / 2**30, "batched": mix,
            "continuous_batching": dict(cb, workers=BATCH_WORKERS, max_batch=BATCH_MAX, threads=BATCH_THREADS),
            "sequential": seq, "identical_sample": proofs_mix == proofs_seq == proofs_cb,
            "proof_buffers": "one pageable arena written before the clock, the same for every way"}


def serving_sweep(m, requests, rates, workers, budget):
    """The serving workload at rising arrival rates, `requests` requests per rate, served with
    `workers` proofs in flight and one at a time: where each saturates (throughput stops following the
    arrival rate and latency turns into queueing)."""
    rows = []
    for rate in rates:
        reqs = serving_requests(requests, rate)
        arena = serving_arena(m, reqs)
        try:
            mix, proofs_mix = serve(m, reqs, workers, budget, arena)
            cb, proofs_cb = serve(m, reqs, BATCH_WORKERS, budget, arena, batch=BATCH_MAX)
            seq, proofs_seq = serve(m, reqs, 1, budget, arena)
        finally:
            arena.close()
        rows.append({"rate_per_s": rate, "batched": mix, "continuous_batching": cb, "sequential": seq,
                     "identical_sample": proofs_mix == proofs_seq == proofs_cb})
        print(json.dumps({"sweep_row": rows[-1]}), file=sys.stderr, flush=True)
    return {"metric": "serving p50/p99 ms and proofs/s against the arrival rate on 1 MI355X",
            "requests_per_rate": requests, "workers": workers,
            "continuous_batching": {"workers": BATCH_WORKERS, "max_batch": BATCH_MAX, "threads": BATCH_THREADS},
            "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), "hbm_budget_gib": budget / 2**30,
            "rows": rows}


def commit_micro(m, steps, warmup):
    out = {"workload": f"commit of {C3_COLUMNS} columns x {C3_ELEMENTS} binary ring elements (16 MiB witness, "
                       f"seed 0), Q={Q_C3}", "q": Q_C3}
    for rows in (1, 16):
        c = m.Commitment(C3_COLUMNS, C3_ELEMENTS, rows, Q_C3)
        try:
            c.generate(bench_seed(0))
            for _ in range(max(1, warmup)):
                c.launch()
            c.read()
            c.set_timing(True)
            for _ in range(steps):
                c.launch()
            t = c.timing()
            primes = c.primes
        finally:
            c.close()
        ms = t["total_ms"] / t["runs"]
        kms = {k.replace("_ms", ""): t[k] / t["runs"] for k in ("transform_ms", "mac_ms", "sum_ms", "finish_ms")}
        rec = {"rows": rows, "primes": primes, "ms": round(ms, 4),
               "kernels_ms": {k: round(v, 4) for k, v in kms.items()}}
        if rows == 1:
            # one kernel: transform and MAC fused
            achieved = C3_COLUMNS * C3_ELEMENTS * primes * c3_slots(1) / (ms / 1e3) / 1e12
            rec["roofline"] = {"bound": "valu", "kernel": "k_commit_mac (+ slice sum, finish)",
                               "achieved": round(achieved, 3), "peak": round(VALU_SLOTS_T, 3),
                               "unit": "T int32 slots/s", "frac": round(achieved / VALU_SLOTS_T, 4)}
        else:
            achieved = C3_COLUMNS * C3_ELEMENTS * primes * 3564 * MULMOD_SLOTS / (kms["transform"] / 1e3) / 1e12
            rec["roofline"] = {"bound": "valu", "kernel": "k_commit_v (transform)", "achieved": round(achieved, 3),
                               "peak": round(VALU_SLOTS_T, 3), "unit": "T int32 slots/s",
                               "frac": round(achieved / VALU_SLOTS_T, 4)}
            gbs = c3_product_bytes(rows, primes) / (kms["mac"] / 1e3) / 1e9
            rec["products_roofline"] = {"bound": "hbm", "kernel": "k_commit_gemm_mfma", "achieved": round(gbs, 1),
                                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}
        out[f"rows{rows}"] = rec
    return out


def main():
    global BATCH_WORKERS, BATCH_MAX
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log-vars", type=int, default=31)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-instance batched measurement")
    ap.add_argument("--no-commit-micro", action="store_true", help="skip the commit micro-benchmark")
    ap.add_argument("--commit-micro-only", action="store_true",
                    help="run only the commit micro-benchmark (prints its own line)")
    ap.add_argument("--workers", type=int, default=None,
                    help="proofs in flight: default 12 for the batched run (64 proofs, highest throughput "
                         "measured), 8 for the serving loop (lower latency per proof)")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (hardware queues the proofs' streams map onto)")
    ap.add_argument("--serving", action="store_true", help="run the serving loop instead (prints its own line)")
    ap.add_argument("--tune-only", action="store_true",
                    help="run the launch-shape autotuner for k = 20..28 and 31 and print its table")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched measurement; its line also gives the monotonic-clock windows "
                         "(ns) of the sequential and the parallel batch, to cut a kernel trace to them")
    ap.add_argument("--requests", type=int, default=10_000)
    ap.add_argument("--rate", type=float, default=400.0, help="serving arrivals per second")
    ap.add_argument("--batch-workers", type=int, default=BATCH_WORKERS,
                    help="workers of the continuous-batching serving mode")
    ap.add_argument("--batch-max", type=int, default=BATCH_MAX, help="largest batch of that mode")
    ap.add_argument("--serving-sweep", default=None, metavar="RATES",
                    help="comma-separated arrival rates: the serving loop at each (--requests per rate, no "
                         "checker), batched and one at a time (prints its own line)")
    ap.add_argument("--pmc-dir", default=None,
                    help="directory holding FETCH_SIZE and WRITE_SIZE rocprofv3 --pmc runs of this bench with "
                         "--copy-calibration (subdirectories fetch/ and write/): fills roofline.traffic")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one plain device-to-device copy of the table size first (a known-size kernel "
                         "for scaling profiler byte counters)")
    ap.add_argument("--cpu-seeds", type=int, default=1, help="seeds of the CPU baseline (each all cores and 8)")
    ap.add_argument("--process-seeds", type=int, default=5, help="spawn-to-exit runs of the GPU executable")
    ap.add_argument("--no-process", action="store_true", help="skip the whole-process timings")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the prover's kernels run on one device per proof; "
              "a multi-GPU sharding is designed but not built (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if not 1 <= args.hw_queues <= 32:
        ap.error("--hw-queues must be between 1 and 32")
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    # child processes first: nothing in this process has touched the GPU yet
    proc_cpu, proc_gpu = None, None
    if not (args.serving or args.serving_sweep or args.tune_only or args.commit_micro_only or args.batched_only
            or args.no_process):
        proc_cpu, proc_gpu = process_runs(args.log_vars, args.cpu_seeds, args.process_seeds,
                                          not args.no_cpu_baseline)
    BATCH_WORKERS, BATCH_MAX = args.batch_workers, args.batch_max
    import amd
    m = amd.load()
    if args.serving:
        print(json.dumps(serving(m, args.requests, args.rate, args.workers or 8, 200 << 30)), flush=True)
        return
    if args.serving_sweep:
        rates = [float(x) for x in args.serving_sweep.split(",")]
        print(json.dumps(serving_sweep(m, args.requests, rates, args.workers or 8, 200 << 30)), flush=True)
        return
    if args.commit_micro_only:
        print(json.dumps(commit_micro(m, args.steps, args.warmup)), flush=True)
        return
    if args.tune_only:
        print(json.dumps({"autotune": tune_shapes(m, list(range(20, 29)) + [31], reps=5)}), flush=True)
        return
    if args.batched_only:
        b = batched(m, args.workers or 12)
        b["windows_ns"] = batched.windows_ns
        print(json.dumps(b), flush=True)
        return
    k = args.log_vars
    tuned = tune_shapes(m, [k])[k]  # the commitment's launch shape for this size, outside the clock
    p = m.Prover(k)
    p.set_modulus(m.DEFAULT_Q)
    if args.copy_calibration:  # the library's copy kernel over 2^k/8 bytes: read once, written once
        p.copy_probe()
    # The table handed over from host memory (pageable) as the reference's loader does: the
    # PCIe-inclusive cost, reported beside the step time, never inside it.
    host_table = np.frombuffer(np.random.default_rng(7).bytes((1 << k) // 8), dtype=np.uint8)
    t0 = time.perf_counter()
    p.load_table(host_table)
    table_h2d_ms = (time.perf_counter() - t0) * 1000
    del host_table
    for w in range(args.warmup):      # with the pass timers on, so their first use is not timed
        p.generate_table(bench_seed(1000 + w))
        p.set_timing(True)
        p.prove(bench_seed(1000 + w))
        p.set_timing(False)

    # The proof lands in one caller-owned host buffer, allocated and touched once up front: pinned
    # memory from the library (lbn_host_alloc), which the fold kernel and the commitment copy write
    # directly.
    pinned = m.PinnedBuffer(m.proof_len(k, m.DEFAULT_Q))
    out = pinned.array
    out[:] = 0
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
    # The PCIe-inclusive proof from a pinned host table (lbn_load_table_async: the upload goes in
    # ring-element ranges overlapping the commitment), reported beside the step, never as the value.
    host_pin = m.PinnedBuffer((1 << k) // 8)
    host_pin.array[:] = np.frombuffer(np.random.default_rng(8).bytes((1 << k) // 8), dtype=np.uint8)
    host_proof_ms = []
    for i in range(3):
        p.load_table_async(host_pin)
        t0 = time.perf_counter()
        p.prove_into(bench_seed(5000 + i), out)
        host_proof_ms.append((time.perf_counter() - t0) * 1000)
    host_pin.close()
    med = statistics.median(times)
    mean_t = {f: statistics.mean(t[f] for t in timings) for f in timings[0]}
    n_words = 1 << (k - 7)
    re = (1 << (k - 17)) // 4
    table_bytes = n_words * 16
    primes, planes = p.commit_primes(), p.commit_planes()
    if planes:
        per_launch_ms = mean_t["commit_mac_ms"]  # one k_toep_mac launch
        int8_t = re * 1024 * TOEP_OPS_PER_ELEMENT_COLUMN * planes / (per_launch_ms / 1e3) / 1e12
    else:
        per_launch_ms = mean_t["commit_mac_ms"] / primes  # one k_commit_mac launch per CRT prime
        mulmod_t = re * 1024 * MULMOD_PER_ELEMENT_PRIME / (per_launch_ms / 1e3) / 1e12
    passes = {
        "commit_mac": {"ms": mean_t["commit_mac_ms"], "bound": "mfma" if planes else "valu"},
        "partials": {"ms": mean_t["eq_ms"] + mean_t["partials_ms"], "bound": "hbm",
                     "achieved_gbs": table_bytes / ((mean_t["eq_ms"] + mean_t["partials_ms"]) / 1e3) / 1e9},
        "switch": {"ms": mean_t["switch0_ms"] + mean_t["switch_rounds_ms"], "bound": "hbm",
                   "achieved_gbs": (table_bytes + 6_291_456 + 18_874_368) / ((mean_t["switch0_ms"] + mean_t["switch_rounds_ms"]) / 1e3) / 1e9},
        "row": {"ms": mean_t["row_ms"], "bound": "hbm", "achieved_gbs": table_bytes / (mean_t["row_ms"] / 1e3) / 1e9},
        "fold": {"ms": mean_t["fold_ms"], "bound": "hbm",
                 "achieved_gbs": (table_bytes + re * 648 * 4) / (mean_t["fold_ms"] / 1e3) / 1e9},
        "commit_finish": {"ms": mean_t["commit_finish_ms"], "bound": "valu"},
    }
    if planes:
        roofline = {"bound": "mfma", "kernel": "k_toep_mac", "achieved": round(int8_t, 1), "peak": INT8_PEAK_T,
                    "unit": "TOP/s", "frac": round(int8_t / INT8_PEAK_T, 4), "traffic": None,
                    "launch_ms": round(per_launch_ms, 3)}
        kernel_name = "k_toep_mac"
    else:
        roofline = {"bound": "valu", "kernel": "k_commit_mac", "achieved": round(mulmod_t, 3),
                    "peak": round(MULMOD_PEAK_T, 3), "unit": "Tmulmod/s", "frac": round(mulmod_t / MULMOD_PEAK_T, 4),
                    "traffic": None, "launch_ms": round(per_launch_ms, 3)}
        kernel_name = "k_commit_mac"
    if args.pmc_dir:
        roofline["traffic"], roofline["traffic_detail"] = pmc_traffic(args.pmc_dir, k, kernel_name)
    elif k == 31:
        roofline["traffic"], roofline["traffic_detail"] = committed_traffic(kernel_name)
    line = {
        "metric": "prover ms (commit+open, 2^31 evals) on 1 MI355X",
        "value": round(med, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(statistics.mean(times), 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32/gf2", "data": "synthetic (BLAKE3 tables from fixed seeds)",
        "config": {"workload": f"one proof, 2^{k} Boolean evaluations (2^{k - 7} GHASH words), Q=35039150080, "
                               + (f"commitment on the matrix cores ({planes} int8 planes)" if planes
                                  else f"{primes} CRT primes"),
                   "log_vars": k, "modulus": m.DEFAULT_Q},
        "roofline": roofline,
        "autotune": tuned,
        "passes_ms": {n: round(v["ms"], 3) for n, v in passes.items()},
        "host_phases_ms": {f.replace("_us", ""): round(statistics.mean(pr[f] for pr in profs) / 1000, 3)
                           for f in profs[0]},
        "step_ms": [round(t, 3) for t in times],
        "table_h2d_ms": round(table_h2d_ms, 3),
        "host_table_proof_ms": {"median": round(statistics.median(host_proof_ms), 3),
                                "runs": [round(x, 3) for x in host_proof_ms],
                                "what": "lbn_load_table_async from pinned memory + lbn_prove: the 2^k/8-byte "
                                        "upload in four ring-element ranges overlapping the commitment"},
        "hbm_passes": {n: {"ms": round(v["ms"], 3), "gbs": round(v["achieved_gbs"], 1),
                           "frac": round(v["achieved_gbs"] / HBM_PEAK_GBS, 4)}
                       for n, v in passes.items() if v["bound"] == "hbm"},
    }
    p.close()
    out = None
    pinned.close()
    if not args.no_commit_micro:
        line["commit_micro"] = commit_micro(m, max(args.steps, 5), args.warmup)
    if not args.no_batched:
        line["batched"] = batched(m, args.workers or 12)
    if proc_cpu:
        line["cpu_baseline"] = proc_cpu
    if proc_gpu:
        line["process"] = proc_gpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
