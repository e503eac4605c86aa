"""Batched / continuous proving through the scheduler ABI (lbn_sched_*), on an MI355X.

Every proof produced in a mix of concurrent jobs must be byte-identical to the proof of the same
job run alone on a single context, and to the CPU oracle's; memory pressure (a budget below the
working set) forces context eviction without changing any result.
"""
import sys

import numpy as np
import pytest

import amd
import oracle_ctypes as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def seed_of(i):
    return bytes((i * 37 + j * 11) & 255 for j in range(32))


@pytest.fixture(scope="module")
def m():
    mod = amd.load()
    O.build()
    return mod


@pytest.fixture(scope="module")
def alone(m):
    """Proof of (k, seed) from a single context, one at a time."""
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


MIX = [(20, 0), (22, 1), (21, 2), (20, 3), (23, 4), (22, 5), (20, 6), (21, 7), (23, 8), (20, 9)]


def test_mix_equals_alone_and_oracle(m, alone):
    say("scheduler: 10 jobs, k in 20..23, 4 workers")
    sch = m.Scheduler(m.DEFAULT_Q, hbm_budget=8 << 30, workers=4)
    try:
        hs = [(k, seed_of(i), sch.submit(k, seed_of(i))) for k, i in MIX]
        sch.drain()
        for k, seed, h in hs:
            proof, q_us, run_us = sch.result(h)
            assert run_us > 0
            assert proof == alone(k, seed), (k, seed[:2].hex())
        st = sch.stats()
        assert st["completed"] == len(MIX) and st["failed"] == 0
    finally:
        sch.close()
    for k, i in MIX[:3]:  # and the alone proofs are the oracle's
        say("oracle k", k)
        assert alone(k, seed_of(i)) == O.prove(seed_of(i), k)



def test_mix_into_caller_buffers(m, alone):
    say("scheduler: 10 jobs into slices of one pinned buffer and of one pageable array, 4 workers")
    sizes = [(m.proof_len(k, m.DEFAULT_Q) + 255) // 256 * 256 for k, _ in MIX]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    pin = m.PinnedBuffer(int(offs[-1]))
    page = np.zeros(int(offs[-1]), dtype=np.uint8)
    sch = m.Scheduler(m.DEFAULT_Q, hbm_budget=8 << 30, workers=4)
    try:
        for arr in (pin.array, page):
            hs = [(k, seed_of(i), sch.submit(k, seed_of(i), out=arr[offs[j]:offs[j] + sizes[j]]))
                  for j, (k, i) in enumerate(MIX)]
            sch.drain()
            for j, (k, seed, h) in enumerate(hs):
                proof = sch.result(h)[0]
                assert proof == alone(k, seed), (k, seed[:2].hex())
                assert arr[offs[j]:offs[j] + len(proof)].tobytes() == proof
    finally:
        sch.close()
        pin.close()

def test_prewarmed_contexts(m, alone):
    say("scheduler: prewarm 3 contexts of k=20 and 2 of k=21, 6 jobs on 3 workers")
    sch = m.Scheduler(m.DEFAULT_Q, hbm_budget=8 << 30, workers=3)
    try:
        sch.prewarm(20, 3)
        sch.prewarm(21, 2)
        sch.prewarm(20, 2)  # already held: creates nothing
        assert sch.stats()["contexts_created"] == 5
        jobs = [(20, 10), (21, 11), (20, 12), (20, 13), (21, 14), (20, 15)]
        hs = [(k, seed_of(i), sch.submit(k, seed_of(i))) for k, i in jobs]
        sch.drain()
        for k, seed, h in hs:
            assert sch.result(h)[0] == alone(k, seed), (k, seed[:2].hex())
        st = sch.stats()
        assert st["contexts_created"] == 5 and st["completed"] == len(jobs) and st["failed"] == 0
        with pytest.raises(m.LbnError):
            sch.prewarm(31, 64)  # beyond the budget: nothing created
        assert sch.stats()["contexts_created"] == 5
    finally:
        sch.close()


def test_eviction_under_a_tight_budget(m, alone):
    # room for one k=23 context plus a little: the mix of sizes must evict idle contexts
    budget = m.ctx_footprint(23) + m.ctx_footprint(20)
    say("scheduler: tight budget", budget)
    sch = m.Scheduler(m.DEFAULT_Q, hbm_budget=budget, workers=2)
    try:
        hs = [(k, seed_of(i), sch.submit(k, seed_of(i))) for k, i in MIX]
        sch.drain()
        for k, seed, h in hs:
            assert sch.result(h)[0] == alone(k, seed)
        st = sch.stats()
        assert st["evictions"] > 0
        assert st["peak_bytes"] <= budget
    finally:
        sch.close()


def test_host_table_and_errors(m, alone):
    sch = m.Scheduler(m.DEFAULT_Q, hbm_budget=512 << 20, workers=2)
    try:
        seed = seed_of(3)
        tab = O.table(seed, 20)
        h = sch.submit(20, seed, table=bytes(tab))
        assert sch.result(h)[0] == alone(20, seed)
        with pytest.raises(m.LbnError):
            sch.submit(31, seed)  # a 2^31 context (about 1 GiB) can never fit 512 MiB
        with pytest.raises(m.LbnError):
            sch.submit(18, seed)  # no such shape
    finally:
        sch.close()


def test_continuous_batching_equals_alone(m, alone):
    say("scheduler: continuous batching, 2 workers, batches of up to 16, 30 jobs k in 20..23")
    jobs = [(20 + (i * 7) % 4, 100 + i) for i in range(30)]
    sch = m.Scheduler(m.DEFAULT_Q, hbm_budget=16 << 30, workers=2)
    try:
        sch.set_batch(16, 8)
        hs = [(k, seed_of(i), sch.submit(k, seed_of(i))) for k, i in jobs]
        sch.drain()
        for k, seed, h in hs:
            assert sch.result(h)[0] == alone(k, seed), (k, seed[:2].hex())
        st = sch.stats()
        assert st["completed"] == len(jobs) and st["failed"] == 0
        with pytest.raises(m.LbnError):
            sch.set_batch(65, 8)
        sch.set_batch(1, 1)  # back to one proof per worker
        h = sch.submit(20, seed_of(5))
        assert sch.result(h)[0] == alone(20, seed_of(5))
    finally:
        sch.close()
