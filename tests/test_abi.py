"""The C-ABI library loads without a GPU and exports every symbol include/labinius_gpu.h declares;
host-only entry points answer without touching a device."""
import amd

import oracle_ctypes as O


def test_library_exports_header_symbols():
    m = amd.load()
    L = m.lib()
    names = m.exported_symbols()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n


def test_proof_len_matches_oracle_and_reference():
    m = amd.load()
    assert m.proof_len(31) == 13_958_612  # solution/proof_generate/src/main.rs:2335
    for k in (19, 20, 23, 31):
        for q in (O.DEFAULT_Q, 1 << 36, 55_144_219_393, (1 << 96) - 1):
            assert m.proof_len(k, q) == O.proof_len(k, q)


def test_bad_shape_is_rejected_without_a_device():
    m = amd.load()
    import ctypes
    h = ctypes.c_void_p()
    assert m.lib().lbn_create(0, 18, ctypes.byref(h)) == -1
    assert m.lib().lbn_create(0, 32, ctypes.byref(h)) == -1
    assert m.proof_len(18) == 0


def test_context_footprint_without_a_device():
    m = amd.load()
    sizes = [m.ctx_footprint(k) for k in range(19, 32)]
    assert all(a < b for a, b in zip(sizes, sizes[1:]))
    # the table, its eq expansion and the switch buffers dominate at k = 31: 2^24 words x 16 B x 2.5
    # + 2^23 x 24 B x 1.5
    assert sizes[-1] > (1 << 24) * 16 * 2.5 + (1 << 23) * 24 * 1.5
    assert m.ctx_footprint(31, (1 << 95) + 12345) > m.ctx_footprint(31, 1 << 40)  # more CRT primes
    # below 2^39 the commitment runs on the matrix cores: per ring element 3 x 4 x 272 x 5 dwords of
    # operand tables (4096 elements at k = 31), beside the table and the switch buffers
    assert m.ctx_footprint(31) > 4096 * 16320 * 4 + (1 << 24) * 16 * 2.5
    assert m.ctx_footprint(18) == 0 and m.ctx_footprint(31, 1 << 20) == 0


def test_scheduler_lifecycle_and_errors_without_a_device():
    m = amd.load()
    import ctypes
    h = ctypes.c_void_p()
    assert m.lib().lbn_sched_create(0, 1 << 20, 0, 1 << 30, 2, ctypes.byref(h)) == -4  # Q below the gate
    assert m.lib().lbn_sched_create(0, m.DEFAULT_Q, 0, 1 << 30, 0, ctypes.byref(h)) == -1
    sch = m.Scheduler(m.DEFAULT_Q, hbm_budget=512 << 20, workers=2)
    try:
        import pytest
        with pytest.raises(m.LbnError):
            sch.submit(31, bytes(32))  # never fits the budget
        with pytest.raises(m.LbnError):
            sch.submit(40, bytes(32))
        import numpy as np
        short = np.empty(m.proof_len(20, m.DEFAULT_Q) - 1, dtype=np.uint8)
        with pytest.raises(ValueError):
            sch.submit(20, bytes(32), out=short)  # a caller's buffer below the proof length
        with pytest.raises(ValueError):
            sch.submit(20, bytes(32), out=np.empty(m.proof_len(20, m.DEFAULT_Q), dtype=np.int32))
        assert sch.stats()["submitted"] == 0
        for bad in [(0, 8), (65, 8), (8, 0), (8, 65)]:  # continuous batching limits
            with pytest.raises(m.LbnError):
                sch.set_batch(*bad)
        sch.set_batch(32, 8)
        sch.set_batch(1, 1)
    finally:
        sch.close()


def test_commitment_arguments_without_a_device():
    m = amd.load()
    import ctypes
    L = m.lib()
    h = ctypes.c_void_p()
    q = 661_473_542_593
    assert L.lbn_cmt_len(64, 16, q, 0) == 16 * 64 * 648 * 5
    assert L.lbn_cmt_len(64, 1, (1 << 95) & (2**64 - 1), 1 << 31) == 64 * 648 * 12
    for cols, elems, rows in [(48, 64, 16), (8, 64, 1), (64, 0, 1), (64, 70000, 1), (64, 64, 17), (64, 64, 0)]:
        assert L.lbn_cmt_create(0, cols, elems, rows, q, 0, ctypes.byref(h)) == -1, (cols, elems, rows)
    assert L.lbn_cmt_create(0, 64, 64, 1, 1 << 30, 0, ctypes.byref(h)) == -4  # below the modulus gate
    assert not h.value


def test_pinned_allocation_arguments_without_a_device():
    m = amd.load()
    import ctypes
    L = m.lib()
    assert not L.lbn_host_alloc(0)
    L.lbn_host_free(None)
    buf = ctypes.create_string_buffer(64)
    L.lbn_host_free(ctypes.cast(buf, ctypes.c_void_p))  # not from lbn_host_alloc: ignored


def test_bench_entry_without_a_device():
    # the bench module compiles and its entry refuses more than one GPU before touching a device
    import subprocess
    import sys
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr, r.stderr[-2000:]
