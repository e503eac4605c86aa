"""The allocation-free entry point lbn_partials_static, through the C ABI on an MI355X: for a table
in device memory it gives the partials the context entry point lbn_partials gives for the same
table (and the oracle's), on two calls in a row, the second with a smaller table; sizes beyond its
static array are rejected with a message. Device memory for the tables comes from the HIP runtime
directly (the test's own plumbing)."""
import ctypes
import sys

import numpy as np
import pytest

import amd
import oracle_ctypes as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DeviceBytes:
    """A table copied into memory from hipMalloc (the library allocates nothing here)."""

    def __init__(self, data: np.ndarray):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(data.nbytes)) == 0
        assert self.hip.hipMemcpy(self.ptr, data.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(data.nbytes), 1) == 0
        self.n = data.nbytes

    def free(self):
        self.hip.hipFree(self.ptr)


def static_partials(m, dev, k, r_hi):
    out = np.zeros((128, 2), dtype=np.uint64)
    rc = m.lib().lbn_partials_static(0, dev.ptr, dev.n, k, r_hi.ctypes.data_as(ctypes.c_void_p),
                                     out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, m.lib().lbn_partials_static_error()
    return out


def test_static_partials_equal_the_context_entry_point():
    m = amd.load()
    assert m.lib().lbn_partials_static_scratch() <= 256 * 1024
    rng = np.random.default_rng(17)
    for k, seed in ((23, bytes([3] * 32)), (20, bytes([4] * 32))):  # the second call is smaller
        say("static partials k", k)
        tab = O.table(seed, k)
        r_hi = rng.integers(0, 2**63, size=(k - 7, 2), dtype=np.uint64)
        dev = DeviceBytes(tab)
        try:
            got = static_partials(m, dev, k, r_hi)
            got2 = static_partials(m, dev, k, r_hi)  # the array keeps the last call's sums: cleared
        finally:
            dev.free()
        p = m.Prover(k)
        try:
            p.load_table(tab)
            want = p.partials(r_hi)
        finally:
            p.close()
        assert (got == want).all() and (got2 == want).all()
        assert (want == O.partials(tab, k, r_hi)).all()


def test_static_partials_reject_bad_sizes():
    m = amd.load()
    L = m.lib()
    tab = O.table(bytes([5] * 32), 20)
    dev = DeviceBytes(tab)
    try:
        r_hi = np.zeros((24, 2), dtype=np.uint64)
        out = np.zeros((128, 2), dtype=np.uint64)
        P = ctypes.c_void_p
        args = (r_hi.ctypes.data_as(P), out.ctypes.data_as(P))
        assert L.lbn_partials_static(0, dev.ptr, dev.n - 16, 20, *args) == -1  # wrong length
        assert b"length" in L.lbn_partials_static_error()
        assert L.lbn_partials_static(0, dev.ptr, dev.n, 32, *args) == -1  # beyond the largest shape
        assert L.lbn_partials_static(0, ctypes.c_void_p(dev.ptr.value + 4), dev.n, 20, *args) == -1  # misaligned
    finally:
        dev.free()
