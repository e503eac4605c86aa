"""The batch prover (lbn_prove_batch: one device switch sequence for all instances) on an MI355X: the
benchmark's 64 instances of k in [20, 28] give, proof for proof, the bytes each context gives alone
(lbn_prove), the library's checker accepts every one, the oracle's prover agrees on one instance per
size; a batch with one instance, a batch of one size, the loaded table instead of a generated one, a
batch mixing matrix-core and CRT moduli, and item errors (a proof buffer too small, a context
without a modulus) behave like lbn_prove.
"""
import sys

import numpy as np
import pytest

import amd
import oracle_ctypes as O
from bench import Checker, batched_instances, bench_seed

pytestmark = pytest.mark.gpu
E_ARG, E_STATE = -1, -2  # LBN_E_ARG, LBN_E_STATE (include/labinius_gpu.h)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def m():
    return amd.load()


def solo(p, seed):
    p.generate_table(seed)
    return p.prove(seed)[0]


def run_batch(m, provers, seeds, gen_table=True, threads=12, caps=None):
    outs = [np.zeros(caps[i] if caps else m.proof_len(p.log_vars, m.DEFAULT_Q), dtype=np.uint8)
            for i, p in enumerate(provers)]
    res = m.prove_batch(provers, seeds, outs, gen_table=gen_table, threads=threads)
    return [(out[:n].tobytes() if st == 0 else None, st) for out, (n, st) in zip(outs, res)]


def test_batch_64_instances_equal_solo(m):
    jobs = batched_instances()
    say("batch prover: 64 instances, k", sorted({k for k, _ in jobs}))
    provers = []
    try:
        for k, _ in jobs:
            p = m.Prover(k)
            p.set_modulus(m.DEFAULT_Q)
            provers.append(p)
        got = run_batch(m, provers, [s for _, s in jobs])
        assert all(st == 0 for _, st in got), [st for _, st in got]
        say("batch prover: solo proofs on the same contexts")
        for i, ((k, seed), p) in enumerate(zip(jobs, provers)):
            assert got[i][0] == solo(p, seed), (i, k)
        say("batch prover: a second batch on the same contexts")
        again = run_batch(m, provers, [s for _, s in jobs], threads=4)
        assert [g[0] for g in again] == [g[0] for g in got]
    finally:
        for p in provers:
            p.close()
    say("batch prover: library checker on every proof")
    check = Checker(m)
    try:
        for i, ((k, seed), (proof, _)) in enumerate(zip(jobs, got)):
            assert check(k, seed, proof), (i, k)
    finally:
        check.close()
    say("batch prover: oracle prover, one instance per size up to 23")
    seen = set()
    for (k, seed), (proof, _) in zip(jobs, got):
        if k <= 23 and k not in seen:
            seen.add(k)
            assert proof == O.prove(seed, k), k


def test_batch_edge_cases(m):
    # one instance; one size; the loaded table
    p = m.Prover(20)
    p.set_modulus(m.DEFAULT_Q)
    seed = bench_seed(7)
    try:
        (proof, st), = run_batch(m, [p], [seed])
        assert st == 0 and proof == O.prove(seed, 20)
        p.generate_table(seed)
        (proof2, st2), = run_batch(m, [p], [seed], gen_table=False)
        assert st2 == 0 and proof2 == proof
    finally:
        p.close()
    ps = [m.Prover(21) for _ in range(5)]
    try:
        for q in ps:
            q.set_modulus(m.DEFAULT_Q)
        seeds = [bench_seed(200 + i) for i in range(5)]
        got = run_batch(m, ps, seeds, threads=2)
        for q, s, (pr, st) in zip(ps, seeds, got):
            assert st == 0 and pr == solo(q, s)
        # item errors: a buffer too small fails that item only
        caps = [m.proof_len(21, m.DEFAULT_Q)] * 5
        caps[2] = 100
        got2 = run_batch(m, ps, seeds, caps=caps)
        assert got2[2][1] == E_ARG and got2[2][0] is None
        for i in (0, 1, 3, 4):
            assert got2[i][1] == 0 and got2[i][0] == got[i][0]
    finally:
        for q in ps:
            q.close()
    # a context without a modulus fails with a state error, the others prove
    a, b = m.Prover(20), m.Prover(20)
    try:
        b.set_modulus(m.DEFAULT_Q)
        got = run_batch(m, [a, b], [bench_seed(1), bench_seed(2)])
        assert got[0][1] == E_STATE and got[1][1] == 0
        assert got[1][0] == O.prove(bench_seed(2), 20)
    finally:
        a.close()
        b.close()


def test_batch_mixed_moduli(m):
    # 18 proofs (two groups): the default Q (matrix-core commitment) and a 96-bit Q (CRT commitment,
    # 12-byte coefficients) in one batch; each equals its solo proof, one per modulus the oracle's
    big = (1 << 95) + 12345
    jobs = [(20 + i % 2, m.DEFAULT_Q if i % 3 else big, bench_seed(300 + i)) for i in range(18)]
    say("batch prover: 18 instances, k 20..21, moduli", sorted({q for _, q, _ in jobs}))
    provers = []
    try:
        for k, q, _ in jobs:
            p = m.Prover(k)
            p.set_modulus(q)
            provers.append(p)
        caps = [m.proof_len(k, q) for k, q, _ in jobs]
        outs = [np.zeros(c, dtype=np.uint8) for c in caps]
        res = m.prove_batch(provers, [s for _, _, s in jobs], outs, threads=8)
        assert all(st == 0 for _, st in res), res
        got = [o[:n].tobytes() for o, (n, _) in zip(outs, res)]
        for i, ((k, q, seed), p) in enumerate(zip(jobs, provers)):
            assert got[i] == solo(p, seed), (i, k, q)
    finally:
        for p in provers:
            p.close()
    for q in (m.DEFAULT_Q, big):
        i = next(i for i, (k, qq, _) in enumerate(jobs) if qq == q and k == 20)
        assert got[i] == O.prove(jobs[i][2], 20, q), q
