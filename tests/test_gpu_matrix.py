"""The public matrix expansion on the device (rq::ajtai_matrix, verifier/proof_verify/src/rq.rs:53-144)
through lbn_ajtai_elements, against the oracle's expansion (pinned by the reference's own stream
fixtures in test_oracle.py): every coefficient of the sampled elements, for moduli of 5 to 12
coefficient bytes, element indices at both ends of a matrix and past it (the key rows' indices),
and the set-up time of a k=31 context's modulus, whose matrix it expands.
"""
import sys
import time

import pytest

import amd
import oracle_ctypes as O

pytestmark = pytest.mark.gpu

MODULI = [34_816_965_049, O.DEFAULT_Q, (1 << 61) - 1, (1 << 80) + 13, (1 << 96) - 17]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("q", MODULI, ids=lambda q: f"bits{q.bit_length()}")
def test_device_expansion_equals_oracle(q):
    m = amd.load()
    for elements, first, count in [(2, 0, 2), (4096, 0, 3), (4096, 4094, 4), (4096, 16 * 4096 + 5, 2)]:
        say(f"matrix q={q} elements={elements} first={first} count={count}")
        got = m.ajtai_elements(q, elements, first, count)
        for i in range(count):
            assert got[i] == O.ajtai_element(q, elements, first + i), (q, elements, first + i)


def test_set_modulus_k31_time():
    m = amd.load()
    p = m.Prover(31)
    try:
        for q in (O.DEFAULT_Q, (1 << 61) - 1):
            t = time.perf_counter()
            p.set_modulus(q)
            say(f"set_modulus k=31 q bits {q.bit_length()}: {1e3 * (time.perf_counter() - t):.1f} ms")
    finally:
        p.close()
