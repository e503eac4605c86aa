"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes view of the CPU oracle (oracle/build/liblabinius_oracle.so). Imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker; the product never
imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liblabinius_oracle.so")
CLI_PATH = os.path.join(HERE, "build", "labinius-oracle")
DEFAULT_Q = 35_039_150_080


class Challenge(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_uint8 * 32), ("signs", ctypes.c_uint32), ("weight", ctypes.c_uint32)]


def build(force=False):
    """Compile the oracle with its own Makefile (gcc)."""
    if force or not (os.path.exists(LIB_PATH) and os.path.exists(CLI_PATH)):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        L = _lib
        P = ctypes.c_void_p
        L.orc_proof_len.restype = ctypes.c_size_t
        L.orc_proof_len.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_prove.argtypes = [P, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, P, P, ctypes.c_size_t,
                                ctypes.POINTER(ctypes.c_size_t)]
        L.orc_verify.argtypes = [P, ctypes.c_int, P, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
        L.orc_replay.argtypes = [P, ctypes.c_int, P, ctypes.c_size_t, P, P, P, P]
        L.orc_table_bytes.argtypes = [P, ctypes.c_uint64, P, ctypes.c_size_t]
        L.orc_blake3.argtypes = [ctypes.c_int, P, ctypes.c_size_t, P, ctypes.c_size_t, ctypes.c_uint64, P,
                                 ctypes.c_size_t]
        L.orc_partials.argtypes = [P, ctypes.c_int, P, P]
        L.orc_switch_run.argtypes = [P, ctypes.c_int, P, P, P, P, P]
        L.orc_row_eval.argtypes = [P, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P]
        L.orc_fold.argtypes = [P, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P]
        L.orc_fold_cap.restype = ctypes.c_uint64
        L.orc_commit_columns_packed.argtypes = [P, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                                ctypes.c_int, P]
        L.orc_coefficient_bytes.restype = ctypes.c_int
        L.orc_b128_mul.argtypes = [P, P, P]
        L.orc_f162_mul.argtypes = [P, P, P]
        L.orc_b128_mul_bitwise.argtypes = [P, P, P]
        L.orc_f162_mul_bitwise.argtypes = [P, P, P]
        L.orc_transcript_probe.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t, P, ctypes.c_size_t, ctypes.c_int,
                                           P, P]
        L.orc_challenges_probe.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t, ctypes.c_int, P]
        L.orc_attempt_from_bytes.restype = ctypes.c_long
        L.orc_attempt_from_bytes.argtypes = [P, ctypes.c_size_t, ctypes.c_int, P]
        L.orc_challenges_accounting.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t, ctypes.c_int, P, P]
        L.orc_canonical_inf_norm_sq.restype = ctypes.c_double
        L.orc_canonical_inf_norm_sq.argtypes = [P]
        L.orc_canonical_inf_norm_sq_naive.restype = ctypes.c_double
        L.orc_canonical_inf_norm_sq_naive.argtypes = [P]
        L.orc_sign_bits.restype = ctypes.c_uint32
        L.orc_sign_bits.argtypes = [P]
        L.orc_ajtai_element.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_size_t, P]
        L.orc_commit_key_row.argtypes = [P, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.c_int, ctypes.c_int, P]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if isinstance(a, np.ndarray) else ctypes.cast(a, ctypes.c_void_p)


def _buf(b):
    return ctypes.cast(ctypes.c_char_p(b), ctypes.c_void_p)


def table(seed: bytes, log_vars: int) -> np.ndarray:
    n = (1 << log_vars) // 8
    out = np.empty(n, dtype=np.uint8)
    lib().orc_table_bytes(_buf(seed), 0, _p(out), n)
    return out


def blake3(mode: int, key_or_ctx: bytes, msg: bytes, offset: int, n: int) -> bytes:
    out = ctypes.create_string_buffer(n)
    k = key_or_ctx or b""
    lib().orc_blake3(mode, _buf(k), len(k), _buf(msg), len(msg), offset, ctypes.cast(out, ctypes.c_void_p), n)
    return out.raw


def proof_len(log_vars: int, q: int = DEFAULT_Q) -> int:
    return lib().orc_proof_len(log_vars, q & (2**64 - 1), q >> 64)


def prove(seed: bytes, log_vars: int, q: int = DEFAULT_Q, table_bytes=None) -> bytes:
    cap = proof_len(log_vars, q)
    out = np.empty(cap, dtype=np.uint8)
    n = ctypes.c_size_t(0)
    t = _p(table_bytes) if table_bytes is not None else None
    rc = lib().orc_prove(_buf(seed), log_vars, q & (2**64 - 1), q >> 64, t, _p(out), cap, ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(f"oracle prove failed: {rc}")
    return out[: n.value].tobytes()


def verify(seed: bytes, log_vars: int, proof: bytes):
    """Returns None when accepted, else the rejection message."""
    err = ctypes.create_string_buffer(256)
    buf = np.frombuffer(proof, dtype=np.uint8)
    rc = lib().orc_verify(_buf(seed), log_vars, _p(buf), len(proof), err, 256)
    return None if rc == 0 else err.value.decode()


def replay(seed: bytes, log_vars: int, proof: bytes):
    lp = log_vars - 7
    r_hi = np.zeros((lp, 2), dtype=np.uint64)
    batch = np.zeros((128, 3), dtype=np.uint64)
    coords = np.zeros((lp, 3), dtype=np.uint64)
    ch = (Challenge * 1024)()
    buf = np.frombuffer(proof, dtype=np.uint8)
    rc = lib().orc_replay(_buf(seed), log_vars, _p(buf), len(proof), _p(r_hi), _p(batch), _p(coords),
                          ctypes.cast(ch, ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError(f"replay failed: {rc}")
    return r_hi, batch, coords, ch


def partials(tab: np.ndarray, log_vars: int, r_hi: np.ndarray) -> np.ndarray:
    out = np.zeros((128, 2), dtype=np.uint64)
    lib().orc_partials(_p(tab), log_vars, _p(np.ascontiguousarray(r_hi, dtype=np.uint64)), _p(out))
    return out


def switch_run(tab, log_vars, r_hi, batch, rhos):
    lp = log_vars - 7
    rounds = np.zeros((lp, 2, 3), dtype=np.uint64)
    opened = np.zeros(3, dtype=np.uint64)
    lib().orc_switch_run(_p(tab), log_vars, _p(np.ascontiguousarray(r_hi, dtype=np.uint64)),
                         _p(np.ascontiguousarray(batch, dtype=np.uint64)),
                         _p(np.ascontiguousarray(rhos, dtype=np.uint64)), _p(rounds), _p(opened))
    return rounds, opened


def row_eval(tab, log_vars, p0, first=0, count=1024):
    out = np.zeros((count, 3), dtype=np.uint64)
    lib().orc_row_eval(_p(tab), log_vars, _p(np.ascontiguousarray(p0, dtype=np.uint64)), first, count, _p(out))
    return out


def fold(tab, log_vars, challenges, first, count):
    out = np.zeros((count, 648), dtype=np.int32)
    bad = lib().orc_fold(_p(tab), log_vars, ctypes.cast(challenges, ctypes.c_void_p), first, count, _p(out))
    return out, bad


def commit_columns(tab, log_vars, q, first, count):
    w = coefficient_bytes(q)
    out = np.zeros(count * 648 * w, dtype=np.uint8)
    lib().orc_commit_columns_packed(_p(tab), log_vars, q & (2**64 - 1), q >> 64, first, count, _p(out))
    return out


def coefficient_bytes(q: int) -> int:
    return ((q - 1).bit_length() + 7) // 8


def commit_key_row(words, elements, row, q, first, count):
    """Packed commitment of `count` columns from `first` of a witness of `elements` ring elements per
    column (uint8 array, 64 bytes per element) under key row `row`: shape (count, 648 * width)."""
    w = coefficient_bytes(q)
    out = np.zeros(count * 648 * w, dtype=np.uint8)
    lib().orc_commit_key_row(_p(words), elements, row, q & (2**64 - 1), q >> 64, first, count, _p(out))
    return out.reshape(count, 648 * w)


def b128_mul(a, b, bitwise=False):
    x = np.array(a, dtype=np.uint64)
    y = np.array(b, dtype=np.uint64)
    o = np.zeros(2, dtype=np.uint64)
    (lib().orc_b128_mul_bitwise if bitwise else lib().orc_b128_mul)(_p(x), _p(y), _p(o))
    return o


def f162_mul(a, b, bitwise=False):
    x = np.array(a, dtype=np.uint64)
    y = np.array(b, dtype=np.uint64)
    o = np.zeros(3, dtype=np.uint64)
    (lib().orc_f162_mul_bitwise if bitwise else lib().orc_f162_mul)(_p(x), _p(y), _p(o))
    return o


def transcript_probe(domain: bytes, msg: bytes, label: bytes, n: int):
    b = np.zeros((n, 2), dtype=np.uint64)
    f = np.zeros((n, 3), dtype=np.uint64)
    lib().orc_transcript_probe(_buf(domain), len(domain), _buf(msg), len(msg), _buf(label), len(label), n,
                               _p(b), _p(f))
    return b, f


def challenges_probe(domain: bytes, msg: bytes, count: int):
    ch = (Challenge * count)()
    lib().orc_challenges_probe(_buf(domain), len(domain), _buf(msg), len(msg), count, ctypes.cast(ch, ctypes.c_void_p))
    return ch


def attempt_from_bytes(stream: bytes, weight: int = 28):
    """One unsigned Fisher-Yates attempt read from `stream` instead of the XOF: (sorted positions,
    bytes consumed), consumed = -1 when the stream ran out."""
    c = Challenge()
    n = lib().orc_attempt_from_bytes(_buf(stream), len(stream), weight, ctypes.byref(c))
    return list(c.positions[:weight]), n


def challenges_accounting(domain: bytes, msg: bytes, count: int):
    att = np.zeros(count, dtype=np.uint32)
    used = np.zeros(count, dtype=np.uint64)
    lib().orc_challenges_accounting(_buf(domain), len(domain), _buf(msg), len(msg), count, _p(att), _p(used))
    return att, used


def ajtai_element(q: int, elements: int, index: int):
    out = (ctypes.c_uint64 * (2 * 648))()
    lib().orc_ajtai_element(q & (2**64 - 1), q >> 64, elements, index, ctypes.cast(out, ctypes.c_void_p))
    return [out[2 * i] | (out[2 * i + 1] << 64) for i in range(648)]
