"""The batched (64 instances of k in [20, 28]) and serving (Poisson requests of k in 20..31 under a
200 GB HBM budget) workloads the benchmark measures, run through the scheduler ABI (lbn_sched_*) on an
MI355X: every proof equals the same proof made alone on one context, both the library's checker
(lbn_verify) and the CPU oracle's accept every proof, and the oracle's own prover gives the same
bytes for one instance of each size. The serving requests run twice: 8 workers at 400 requests/s,
and continuous batching (lbn_sched_set_batch) at 3,000 requests/s.
"""
import sys
import time

import pytest

import amd
import oracle_ctypes as O
from bench import Checker, batched_instances, serving_requests

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def m():
    return amd.load()


@pytest.fixture(scope="module")
def alone(m):
    """Proof of (k, seed) from one context per size, one at a time, table generated on the device."""
    cache, provers = {}, {}

    def get(k, seed):
        if (k, seed) not in cache:
            if k not in provers:
                provers[k] = m.Prover(k)
                provers[k].set_modulus(m.DEFAULT_Q)
            provers[k].generate_table(seed)
            cache[(k, seed)] = provers[k].prove(seed)[0]
        return cache[(k, seed)]

    yield get
    for p in provers.values():
        p.close()


def product_checker_accepts(m, proofs):
    check = Checker(m)
    try:
        for i, (k, seed, proof) in enumerate(proofs):
            assert check(k, seed, proof), (i, k)
    finally:
        check.close()


def test_batched_64_instances(m, alone):
    jobs = batched_instances()
    ks = sorted({k for k, _ in jobs})
    say("batched: 64 instances, k", ks, "12 workers")
    assert ks == list(range(20, 29))
    sch = m.Scheduler(m.DEFAULT_Q, hbm_budget=200 << 30, workers=12)
    try:
        for k in ks:
            sch.prewarm(k, 2)
        hs = [(k, seed, sch.submit(k, seed)) for k, seed in jobs]
        sch.drain()
        proofs = [(k, seed, sch.result(h)[0]) for k, seed, h in hs]
        st = sch.stats()
        assert st["completed"] == 64 and st["failed"] == 0
    finally:
        sch.close()
    say("batched: solo proofs")
    for i, (k, seed, proof) in enumerate(proofs):
        assert proof == alone(k, seed), (i, k)
    say("batched: library checker on every proof")
    product_checker_accepts(m, proofs)
    say("batched: oracle checker on every proof")
    for i, (k, seed, proof) in enumerate(proofs):
        assert O.verify(seed, k, proof) is None, (i, k)
    say("batched: oracle prover, one instance per size")
    seen = set()
    for k, seed, proof in proofs:
        if k not in seen:
            seen.add(k)
            assert proof == O.prove(seed, k), k
    assert seen == set(ks)


@pytest.mark.parametrize("mode", ["workers", "batching"])
def test_serving_poisson_all_sizes(m, alone, mode):
    # workers: 8 proofs in flight at 400 requests/s; batching: continuous batching (2 workers, up to
    # 32 jobs per batch) at 3,000 requests/s, where batches form; the same requests either way
    rate = 400.0 if mode == "workers" else 3000.0
    reqs = serving_requests(520, rate)
    have = {k for _, k, _ in reqs}
    # one request of every size the distribution can draw, forced in at spread-out arrival times
    t_end = reqs[-1][0]
    extra = [(t_end * (k - 19) / 13, k, bytes([k] * 31 + [0x5C])) for k in range(20, 32)]
    reqs = sorted(reqs + extra)
    assert {k for _, k, _ in reqs} == set(range(20, 32))
    say("serving:", len(reqs), "requests over %.2f s, sizes drawn" % t_end, sorted(have))
    sch = m.Scheduler(m.DEFAULT_Q, hbm_budget=200 << 30, workers=8 if mode == "workers" else 2)
    try:
        sch.prewarm(20, 4)
        if mode == "batching":
            sch.set_batch(32, 8)
        hs = []
        t0 = time.perf_counter()
        for at, k, seed in reqs:
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            hs.append((k, seed, sch.submit(k, seed)))
        sch.drain()
        proofs = [(k, seed, sch.result(h)[0]) for k, seed, h in hs]
        st = sch.stats()
        say("serving: stats", st)
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["peak_bytes"] <= 200 << 30
    finally:
        sch.close()
    say("serving: solo proofs")
    for i, (k, seed, proof) in enumerate(proofs):
        assert proof == alone(k, seed), (i, k)
    say("serving: library checker on every proof")
    product_checker_accepts(m, proofs)
    if mode == "workers":  # (the batching run's proofs equal these, request for request)
        say("serving: oracle checker on every proof")
        for i, (k, seed, proof) in enumerate(proofs):
            assert O.verify(seed, k, proof) is None, (i, k)
