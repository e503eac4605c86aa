"""The product checker lbn_verify (verifier/proof_verify/src/main.rs:293-389) through the C ABI, on
an MI355X: it accepts the GPU prover's proofs at several shapes and moduli (the matrix-core
commitment below 2^39, the CRT one above, up to a 12-byte modulus whose products use all three
32-bit limbs), and on every tampered proof it gives the same verdict and the same reason as the
oracle's checker. Tampering the fold alone leaves the transcript unchanged, so those cases reach
the late checks (range, norm, the ring identity A v = sum_j c_j C_j).
"""
import sys
import time

import numpy as np
import pytest

import amd
import oracle_ctypes as O

pytestmark = pytest.mark.gpu

SEED = bytes([42] * 32)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def m():
    mod = amd.load()
    O.build()
    return mod


@pytest.fixture(scope="module")
def p20(m):
    p = m.Prover(20)
    yield p
    p.close()


def prove(p, seed, q):
    p.generate_table(seed)
    p.set_modulus(q)
    proof, _ = p.prove(seed)
    return proof


@pytest.fixture(scope="module")
def proof20(p20):
    say("verify: prove k=20 default modulus")
    return prove(p20, SEED, O.DEFAULT_Q)


def offsets(proof, k, q):
    """Byte offsets of the bincode fields (main.rs:98-124)."""
    lp = k - 7
    clen = 1024 * 648 * O.coefficient_bytes(q)
    o = {"version": 0, "modulus": 4, "clen": 20, "commitment": 28}
    o["partials"] = 28 + clen + 8
    o["rounds"] = o["partials"] + 2048 + 8
    o["opened"] = o["rounds"] + 48 * lp
    o["row_len"] = o["opened"] + 24
    o["row"] = o["row_len"] + 8
    o["fold_len"] = o["row"] + 20736
    o["fold"] = o["fold_len"] + 8
    o["ring_elements"] = 1 << (lp - 12)
    assert o["fold"] + 4 * 648 * o["ring_elements"] == len(proof)
    return o


def test_accepts_gpu_proofs_k20(m, p20, proof20):
    t = time.perf_counter()
    assert p20.verify(SEED, proof20) is None
    say(f"verify: k=20 accepted in {1e3 * (time.perf_counter() - t):.1f} ms")
    assert O.verify(SEED, 20, proof20) is None
    # a fresh context with no modulus set checks the same proof
    p = m.Prover(20)
    try:
        assert p.verify(SEED, proof20) is None
    finally:
        p.close()


@pytest.mark.parametrize("q", [(1 << 80) + 13, (1 << 96) - 17], ids=["q80", "q96"])
def test_accepts_crt_moduli(p20, q):
    seed = bytes(range(7, 39))
    say(f"verify: prove k=20 q={q}")
    proof = prove(p20, seed, q)
    assert p20.verify(seed, proof) is None
    assert O.verify(seed, 20, proof) is None
    bad = bytearray(proof)
    o = offsets(proof, 20, q)
    bad[o["fold"] + 4 * 777] ^= 1
    why = p20.verify(seed, bytes(bad))
    assert why == O.verify(seed, 20, bytes(bad)) == "the fold does not open the folded commitment"


def le32(x):
    return int(x).to_bytes(4, "little", signed=True)


def tampered(proof, o, case):
    b = bytearray(proof)
    if case == "version":
        b[0] = 2
    elif case == "modulus_low":
        b[4:20] = (34_816_965_048).to_bytes(16, "little")
    elif case == "modulus_high":
        b[4:20] = (1 << 96).to_bytes(16, "little")
    elif case == "modulus_other":
        b[4:20] = (O.DEFAULT_Q + 2).to_bytes(16, "little")
    elif case == "commitment":
        b[o["commitment"] + 1000] ^= 1
    elif case == "partial":
        b[o["partials"] + 16 * 5] ^= 1
    elif case == "round":
        b[o["rounds"] + 48 * 3 + 2] ^= 4
    elif case == "round_noncanonical":
        b[o["rounds"] + 48 * 2 + 16 + 5] |= 0x80
    elif case == "opened":
        b[o["opened"]] ^= 1
    elif case == "opened_noncanonical":
        b[o["opened"] + 23] = 0xFF
    elif case == "row":
        b[o["row"] + 7] ^= 1
    elif case == "fold_one":
        at = o["fold"] + 4 * 1000
        v = int.from_bytes(b[at:at + 4], "little", signed=True)
        b[at:at + 4] = le32(v + 1)
    elif case == "fold_range":
        at = o["fold"] + 4 * 99
        b[at:at + 4] = le32(40000)
    elif case == "fold_norm":
        at = o["fold"] + 4 * 99
        b[at:at + 4] = le32(32767)
    elif case == "trailing":
        b += b"\0"
    elif case == "truncated":
        del b[-1]
    elif case == "fold_short":
        b[o["fold_len"]:o["fold_len"] + 8] = (4 * 648 * o["ring_elements"] - 4).to_bytes(8, "little")
        del b[-4:]
    elif case == "row_short":
        b[o["row_len"]:o["row_len"] + 8] = (20735).to_bytes(8, "little")
        del b[o["row"] + 20735]
    elif case == "empty":
        b = bytearray()
    else:
        raise ValueError(case)
    return bytes(b)


CASES = ["version", "modulus_low", "modulus_high", "modulus_other", "commitment", "partial", "round",
         "round_noncanonical", "opened", "opened_noncanonical", "row", "fold_one", "fold_range", "fold_norm",
         "trailing", "truncated", "fold_short", "row_short", "empty"]


@pytest.mark.parametrize("case", CASES)
def test_rejections_match_the_oracle(p20, proof20, case):
    o = offsets(proof20, 20, O.DEFAULT_Q)
    bad = tampered(proof20, o, case)
    say(f"verify: tampered {case}")
    why = p20.verify(SEED, bad)
    ref = O.verify(SEED, 20, bad)
    assert why is not None
    assert why == ref


def test_wrong_seed_and_shape(m, p20, proof20):
    other = bytes([43] * 32)
    why = p20.verify(other, proof20)
    assert why is not None and why == O.verify(other, 20, proof20)
    p = m.Prover(21)
    try:
        why = p.verify(SEED, proof20)
        assert why == O.verify(SEED, 21, proof20) == "one switch round per packed coordinate"
    finally:
        p.close()
    # the context still proves after checking: the regenerated table is the seed's
    proof = prove(p20, SEED, O.DEFAULT_Q)
    assert proof == proof20


def test_accepts_gpu_proof_k31(m):
    p = m.Prover(31)
    try:
        say("verify: prove k=31")
        proof = prove(p, SEED, O.DEFAULT_Q)
        t = time.perf_counter()
        assert p.verify(SEED, proof) is None
        ms = 1e3 * (time.perf_counter() - t)
        t = time.perf_counter()
        assert p.verify(SEED, proof) is None
        say(f"verify: k=31 accepted in {ms:.1f} ms, again in {1e3 * (time.perf_counter() - t):.1f} ms")
        bad = bytearray(proof)
        o = offsets(proof, 31, O.DEFAULT_Q)
        at = o["fold"] + 4 * 1_000_003
        v = int.from_bytes(bad[at:at + 4], "little", signed=True)
        bad[at:at + 4] = le32(v - 1)
        assert p.verify(SEED, bytes(bad)) == "the fold does not open the folded commitment"
        say("verify: k=31 oracle checker")
        assert O.verify(SEED, 31, proof) is None
    finally:
        p.close()


def test_honest_fold_over_the_cap(m):
    """A table whose honest fold crosses the cap (the serving workload's request 5570, k = 23): the
    GPU proof is the oracle's, and both checkers reject it for the fold's length, as the
    reference's checker does (verifier/proof_verify/src/main.rs:476-480)."""
    from test_oracle import OVER_CAP_NORMSQ, OVER_CAP_SEED
    k = 23
    p = m.Prover(k)
    try:
        say("verify: k=23 fold over the cap")
        proof = prove(p, OVER_CAP_SEED, O.DEFAULT_Q)
        v = np.frombuffer(proof[len(proof) - 4 * 648 * (1 << (k - 19)):], dtype="<i4").astype(np.int64)
        assert int((v * v).sum()) == OVER_CAP_NORMSQ
        assert proof == O.prove(OVER_CAP_SEED, k)
        assert p.verify(OVER_CAP_SEED, proof) == "the fold is longer than the cap"
        assert O.verify(OVER_CAP_SEED, k, proof) == "the fold is longer than the cap"
    finally:
        p.close()
