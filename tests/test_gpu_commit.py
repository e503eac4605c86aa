"""The standalone commitment (lbn_cmt_*) against the CPU oracle, through the C ABI (needs an MI355X).

Keys of one row take the fused transform + MAC kernel, keys of several rows the transform, the
per-slot key-row products and the slice sums; both must give the oracle's schoolbook commitment
(rq.rs:264-287) byte for byte. At the commit micro-benchmark's shape (64 columns x 4096 ring
elements, Q = 3889 * 9721 * 17497, 16 key rows) the oracle checks sampled outputs, and two
size-independent properties cover every output: row 0 of the 16-row key equals the one-row
commitment, and the commitment is additive over a split of the witness into disjoint bits.
"""
import sys

import numpy as np
import pytest

import amd
import oracle_ctypes as O

pytestmark = pytest.mark.gpu

Q_C3 = 661_473_542_593


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def m():
    mod = amd.load()
    O.build()
    return mod


def commit(m, words, columns, elements, rows, q):
    c = m.Commitment(columns, elements, rows, q)
    try:
        c.load(words.tobytes())
        return c.commit()
    finally:
        c.close()


def oracle_rows(words, elements, rows, q, cols):
    return {(r, c): O.commit_key_row(words, elements, r, q, c, 1)[0] for r in rows for c in cols}


def as_ints(packed, width):
    b = packed.reshape(-1, width)
    return np.frombuffer(np.pad(b, ((0, 0), (0, 8 - width))).tobytes(), dtype="<u8")


def test_sixteen_rows_small_every_output(m):
    say("cmt: 64 columns x 32 elements, 16 rows, Q_C3")
    w = O.table(bytes([11] * 32), 20)  # 131072 bytes = 64 columns x 32 elements x 64
    got = commit(m, w, 64, 32, 16, Q_C3)
    assert got.shape == (16, 64, 648 * 5)
    for r in range(16):
        want = O.commit_key_row(w, 32, r, Q_C3, 0, 64)
        assert (got[r] == want).all(), r


def test_one_row_small_and_row_zero_of_many(m):
    say("cmt: 64 columns x 32 elements, 1 row vs 16 rows, default Q")
    w = O.table(bytes([8] * 32), 20)
    one = commit(m, w, 64, 32, 1, O.DEFAULT_Q)
    assert (one[0] == O.commit_key_row(w, 32, 0, O.DEFAULT_Q, 0, 64)).all()
    many = commit(m, w, 64, 32, 16, O.DEFAULT_Q)
    assert (many[0] == one[0]).all()
    # read as 1024 columns of 2 elements, the same bytes are the prover's 2^20 table: one key row
    # gives the prover's commitment
    p = m.Prover(20)
    try:
        p.load_table(w)
        p.set_modulus(O.DEFAULT_Q)
        assert commit(m, w, 1024, 2, 1, O.DEFAULT_Q).tobytes() == bytes(p.commit())
    finally:
        p.close()


def test_three_rows_two_chunks_default_q(m):
    """Rows below 16 (zero key rows in the product tiles) and a ragged second chunk (1100 = 1024 + 76
    elements)."""
    say("cmt: 64 columns x 1100 elements, 3 rows, default Q")
    rng = np.random.default_rng(3)
    w = rng.integers(0, 256, size=64 * 1100 * 64, dtype=np.uint8)
    got = commit(m, w, 64, 1100, 3, O.DEFAULT_Q)
    for (r, c), want in oracle_rows(w, 1100, (0, 2), O.DEFAULT_Q, (0, 63)).items():
        assert (got[r, c] == want).all(), (r, c)


def test_one_row_many_slices_five_primes(m):
    """A one-row key with the most CRT primes (Q near 2^96 over 2500 elements needs 5) and 125 MAC
    slices over a single column group."""
    say("cmt: 16 columns x 2500 elements, 1 row, Q = 2^96 - 1")
    q = (1 << 96) - 1
    rng = np.random.default_rng(5)
    w = rng.integers(0, 256, size=16 * 2500 * 64, dtype=np.uint8)
    c = m.Commitment(16, 2500, 1, q)
    try:
        assert c.primes == 5
        c.load(w.tobytes())
        got = c.commit()
    finally:
        c.close()
    for (r, col), want in oracle_rows(w, 2500, (0,), q, (0, 15)).items():
        assert (got[r, col] == want).all(), col


def test_commit_micro_shape(m):
    """64 columns x 4096 elements (a 16 MiB witness from the table writer's XOF), Q_C3, 16 rows."""
    seed = bytes(range(100, 132))
    say("cmt: C3 shape, generate on the device")
    c16 = m.Commitment(64, 4096, 16, Q_C3)
    try:
        assert c16.primes == 3
        c16.generate(seed)
        full = c16.commit()
        w = O.table(seed, 27)
        assert w.size == 64 * 4096 * 64
        c16.load(w.tobytes())  # the same bytes from the host give the same commitment
        assert (c16.commit() == full).all()
        say("cmt: C3 shape, additivity over a split")
        mask = np.random.default_rng(9).integers(0, 256, size=w.size, dtype=np.uint8)
        c16.load((w & mask).tobytes())
        a = c16.commit()
        c16.load((w & ~mask).tobytes())
        b = c16.commit()
    finally:
        c16.close()
    assert ((as_ints(a, 5) + as_ints(b, 5)) % Q_C3 == as_ints(full, 5)).all()
    say("cmt: C3 shape, one row")
    c1 = m.Commitment(64, 4096, 1, Q_C3)
    try:
        c1.load(w.tobytes())
        one = c1.commit()
    finally:
        c1.close()
    assert (one[0] == full[0]).all()
    say("cmt: C3 shape, oracle samples")
    for (r, col), want in oracle_rows(w, 4096, (0, 15), Q_C3, (0, 63)).items():
        assert (full[r, col] == want).all(), (r, col)


def test_errors(m):
    with pytest.raises(m.LbnError):
        m.Commitment(48, 64, 16, Q_C3)  # several rows need a multiple of 64 columns
    c = m.Commitment(16, 8, 1, Q_C3)
    try:
        with pytest.raises(m.LbnError):
            c.launch()  # no witness
        with pytest.raises(m.LbnError):
            c.load(bytes(100))
    finally:
        c.close()
