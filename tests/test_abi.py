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
