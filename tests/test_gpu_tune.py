"""The launch-shape autotuner (lbn_autotune, lbn_tune_set) on an MI355X: it times every element split
count of the matrix-core commitment up to the shape's limit and records the fastest; proofs made with
any recorded split count, and after tuning, are byte-identical to the default's and to the oracle's; a
modulus committed through the CRT transform has nothing to tune.
"""
import sys

import pytest

import amd
import oracle_ctypes as O
from bench import bench_seed

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def m():
    return amd.load()


@pytest.mark.parametrize("k", [20, 24])
def test_tuned_split_counts_keep_the_proof(m, k):
    seed = bench_seed(300 + k)
    p = m.Prover(k)
    try:
        p.set_modulus(m.DEFAULT_Q)
        p.generate_table(seed)
        base = p.prove(seed)[0]
        if k == 20:
            assert base == O.prove(seed, k)
        r = p.autotune(reps=2)
        say("k", k, "autotune", r)
        re = (1 << (k - 19))
        want = [c for c in (1, 2, 3, 4, 6, 8, 12, 24) if c <= min(re, 24)]
        assert sorted(r["ms"]) == want and r["best"] in want
        assert all(v > 0 for v in r["ms"].values())
        assert p.prove(seed)[0] == base
        for c in want:
            m.tune_set(k, c)
            assert p.prove(seed)[0] == base, c
    finally:
        m.tune_set(k, 0)
        p.close()


def test_crt_modulus_has_nothing_to_tune(m):
    p = m.Prover(20)
    try:
        p.set_modulus((1 << 95) + 12345)
        p.generate_table(bench_seed(1))
        with pytest.raises(m.LbnError):
            p.autotune(reps=1)
    finally:
        p.close()
    with pytest.raises(m.LbnError):
        m.tune_set(18, 1)  # no such shape
    with pytest.raises(m.LbnError):
        m.tune_set(20, 25)
