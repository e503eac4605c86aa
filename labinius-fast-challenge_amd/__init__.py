"""Python view of liblabinius_gpu, the MI355X Labinius prover (C ABI in include/labinius_gpu.h).

The methods mirror the reference prover's kernel entry points one to one (commit, partials,
switch rounds, row evaluation, fold; solution/proof_generate/src/main.rs:1692-2009) and the whole
prover. Every call goes to the HIP library; there is no CPU fallback: if the library is missing
the import-time loader raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liblabinius_gpu.so")
CLI_PATH = os.path.join(HERE, "build", "labinius-fast")
HEADER = os.path.join(os.path.dirname(HERE), "include", "labinius_gpu.h")
DEFAULT_Q = 35_039_150_080  # solution/proof_generate/src/parameters.rs:14

ERRORS = {-1: "argument", -2: "state", -3: "device", -4: "modulus", -5: "range", -6: "io", -7: "reject"}
E_REJECT = -7


class LbnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)} error: {msg}")
        self.code = code


class Challenge(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_uint8 * 32), ("signs", ctypes.c_uint32), ("weight", ctypes.c_uint32)]


class Profile(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("commit_us", "transcript_us", "partials_us", "switch_us", "row_us",
                                              "challenges_us", "fold_us", "encode_us", "total_us")]


class Job(ctypes.Structure):
    _fields_ = [("log_vars", ctypes.c_int), ("seed", ctypes.c_uint8 * 32), ("table", ctypes.c_void_p),
                ("proof", ctypes.c_void_p), ("cap", ctypes.c_size_t), ("len", ctypes.c_size_t),
                ("status", ctypes.c_int), ("queue_us", ctypes.c_double), ("run_us", ctypes.c_double),
                ("profile", Profile)]


class BatchItem(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("seed", ctypes.c_uint8 * 32), ("gen_table", ctypes.c_int),
                ("proof", ctypes.c_void_p), ("cap", ctypes.c_size_t), ("len", ctypes.c_size_t),
                ("status", ctypes.c_int), ("profile", Profile)]


class TuneResult(ctypes.Structure):
    _fields_ = [("log_vars", ctypes.c_int), ("n", ctypes.c_int), ("splits", ctypes.c_int * 8),
                ("ms", ctypes.c_double * 8), ("best", ctypes.c_int)]


class Timing(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("table_ms", "eq_ms", "partials_ms", "switch0_ms", "switch_rounds_ms",
                                              "row_ms", "fold_ms", "commit_mac_ms", "commit_finish_ms")] + \
               [("switch_rounds", ctypes.c_int)]




class CmtTiming(ctypes.Structure):
    _fields_ = [(f, ctypes.c_double) for f in ("total_ms", "transform_ms", "mac_ms", "sum_ms", "finish_ms")] + [
        ("runs", ctypes.c_int)]


class SchedStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "completed", "failed", "contexts_created", "evictions",
                                              "peak_bytes", "used_bytes")]


_lib = None


def lib():
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.lbn_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(P)]
        L.lbn_destroy.argtypes = [P]
        L.lbn_last_error.argtypes = [P]
        L.lbn_last_error.restype = ctypes.c_char_p
        L.lbn_load_table.argtypes = [P, P, ctypes.c_size_t]
        L.lbn_load_table_device.argtypes = [P, P, ctypes.c_size_t]
        L.lbn_load_table_async.argtypes = [P, P, ctypes.c_size_t]
        L.lbn_generate_table.argtypes = [P, P]
        L.lbn_set_modulus.argtypes = [P, ctypes.c_uint64, ctypes.c_uint64]
        L.lbn_commit.argtypes = [P, P, ctypes.c_size_t]
        L.lbn_partials.argtypes = [P, P, P]
        L.lbn_switch_begin.argtypes = [P, P]
        L.lbn_switch_msg.argtypes = [P, P]
        L.lbn_switch_fold.argtypes = [P, P]
        L.lbn_switch_final.argtypes = [P, P]
        L.lbn_row_eval.argtypes = [P, P, ctypes.c_int, P]
        L.lbn_fold.argtypes = [P, P, ctypes.c_int, P, ctypes.c_size_t]
        L.lbn_proof_len.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        L.lbn_proof_len.restype = ctypes.c_size_t
        L.lbn_prove.argtypes = [P, P, P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), P]
        L.lbn_verify.argtypes = [P, P, P, ctypes.c_size_t]
        L.lbn_prove_batch.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.lbn_autotune.argtypes = [P, ctypes.c_int, P]
        L.lbn_tune_set.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.lbn_ajtai_elements.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t,
                                         ctypes.c_size_t, ctypes.c_size_t, P]
        L.lbn_commit_primes.argtypes = [P]
        L.lbn_copy_probe.argtypes = [P]
        L.lbn_commit_planes.argtypes = [P]
        L.lbn_partials_static.argtypes = [ctypes.c_int, P, ctypes.c_size_t, ctypes.c_int, P, P]
        L.lbn_partials_static_scratch.restype = ctypes.c_size_t
        L.lbn_partials_static_error.restype = ctypes.c_char_p
        L.lbn_set_timing.argtypes = [P, ctypes.c_int]
        L.lbn_set_graphs.argtypes = [P, ctypes.c_int]
        L.lbn_reset_timing.argtypes = [P]
        L.lbn_get_timing.argtypes = [P, P]
        L.lbn_ctx_footprint.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        L.lbn_ctx_footprint.restype = ctypes.c_size_t
        L.lbn_sched_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int,
                                       ctypes.POINTER(P)]
        L.lbn_sched_submit.argtypes = [P, P]
        L.lbn_sched_wait.argtypes = [P, P]
        L.lbn_sched_drain.argtypes = [P]
        L.lbn_sched_get_stats.argtypes = [P, P]
        L.lbn_sched_destroy.argtypes = [P]
        L.lbn_sched_destroy.restype = None
        L.lbn_sched_prewarm.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.lbn_sched_set_batch.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.lbn_host_alloc.argtypes = [ctypes.c_size_t]
        L.lbn_host_alloc.restype = ctypes.c_void_p
        L.lbn_host_free.argtypes = [P]
        L.lbn_host_free.restype = None
        L.lbn_cmt_len.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        L.lbn_cmt_len.restype = ctypes.c_size_t
        L.lbn_cmt_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                     ctypes.c_uint64, ctypes.POINTER(P)]
        L.lbn_cmt_destroy.argtypes = [P]
        L.lbn_cmt_destroy.restype = None
        L.lbn_cmt_error.argtypes = [P]
        L.lbn_cmt_error.restype = ctypes.c_char_p
        L.lbn_cmt_primes.argtypes = [P]
        L.lbn_cmt_load.argtypes = [P, P, ctypes.c_size_t]
        L.lbn_cmt_load_device.argtypes = [P, P, ctypes.c_size_t]
        L.lbn_cmt_generate.argtypes = [P, P]
        L.lbn_cmt_launch.argtypes = [P]
        L.lbn_cmt_read.argtypes = [P, P, ctypes.c_size_t]
        L.lbn_cmt_set_timing.argtypes = [P, ctypes.c_int]
        L.lbn_cmt_get_timing.argtypes = [P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _b(b: bytes):
    return ctypes.cast(ctypes.c_char_p(b), ctypes.c_void_p)


def proof_len(log_vars: int, q: int = DEFAULT_Q) -> int:
    return lib().lbn_proof_len(log_vars, q & (2**64 - 1), q >> 64)


def prove_batch(provers, seeds, outs, gen_table=True, threads=12):
    """lbn_prove_batch: proof i of provers[i] (distinct contexts on one GPU, modulus set) for seeds[i]
    into outs[i] (writable contiguous uint8 arrays of at least proof_len bytes); with gen_table the
    table of seeds[i] is generated on the device first. Returns the list of (length, status)."""
    n = len(provers)
    items = (BatchItem * n)()
    for i, (p, seed, out) in enumerate(zip(provers, seeds, outs)):
        if out.dtype != np.uint8 or not out.flags.c_contiguous or not out.flags.writeable:
            raise ValueError("outs must be writable contiguous uint8 arrays")
        items[i].ctx = p.h
        ctypes.memmove(items[i].seed, seed, 32)
        items[i].gen_table = 1 if gen_table else 0
        items[i].proof = out.ctypes.data
        items[i].cap = out.size
    lib().lbn_prove_batch(ctypes.cast(items, ctypes.c_void_p), n, threads)
    return [(items[i].len, items[i].status) for i in range(n)]


def tune_set(log_vars: int, mac_splits: int, device: int = 0):
    """lbn_tune_set: record a commitment split count for (device, log_vars); 0 restores the default."""
    rc = lib().lbn_tune_set(device, log_vars, mac_splits)
    if rc != 0:
        raise LbnError(rc, "lbn_tune_set failed")


def coefficient_bytes(q: int) -> int:
    return ((q - 1).bit_length() + 7) // 8


class Prover:
    """One lbn_ctx: one table resident in HBM on one GPU."""

    def __init__(self, log_vars: int = 31, device: int = 0):
        self.log_vars = log_vars
        self.log_packed = log_vars - 7
        self.row_log = self.log_packed - 10
        self.ring_elements = (1 << self.row_log) // 4
        self.table_bytes = (1 << log_vars) // 8
        self.q = None
        h = ctypes.c_void_p()
        rc = lib().lbn_create(device, log_vars, ctypes.byref(h))
        if rc != 0:
            raise LbnError(rc, "lbn_create failed")
        self.h = h

    def close(self):
        if self.h:
            lib().lbn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise LbnError(rc, lib().lbn_last_error(self.h).decode())

    def load_table(self, table: np.ndarray):
        t = np.ascontiguousarray(table, dtype=np.uint8)
        self._check(lib().lbn_load_table(self.h, _p(t), t.size))

    def load_table_async(self, table):
        """lbn_load_table_async: the upload goes with the next proof (or other table reader). `table`
        is a uint8 numpy array or a PinnedBuffer; this object keeps it alive until then."""
        self._check(lib().lbn_load_table_async(self.h, _p(table) if isinstance(table, np.ndarray) else
                                               ctypes.c_void_p(table.ptr), table.size if isinstance(table, np.ndarray)
                                               else table.nbytes))
        self._pending = table

    def load_table_device(self, device_ptr: int, nbytes: int):
        """Copy the table from device memory (device_ptr: an address on this context's GPU)."""
        self._check(lib().lbn_load_table_device(self.h, ctypes.c_void_p(device_ptr), nbytes))

    def generate_table(self, seed: bytes):
        self._check(lib().lbn_generate_table(self.h, _b(seed)))

    def set_modulus(self, q: int = DEFAULT_Q):
        self._check(lib().lbn_set_modulus(self.h, q & (2**64 - 1), q >> 64))
        self.q = q

    def commit(self) -> bytes:
        n = 1024 * 648 * coefficient_bytes(self.q)
        out = np.zeros(n, dtype=np.uint8)
        self._check(lib().lbn_commit(self.h, _p(out), n))
        return out.tobytes()

    def partials(self, r_hi: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r_hi, dtype=np.uint64)
        out = np.zeros((128, 2), dtype=np.uint64)
        self._check(lib().lbn_partials(self.h, _p(r), _p(out)))
        return out

    def switch_begin(self, batch: np.ndarray):
        self._batch = np.ascontiguousarray(batch, dtype=np.uint64)
        self._check(lib().lbn_switch_begin(self.h, _p(self._batch)))

    def switch_msg(self) -> np.ndarray:
        out = np.zeros((2, 3), dtype=np.uint64)
        self._check(lib().lbn_switch_msg(self.h, _p(out)))
        return out

    def switch_fold(self, r: np.ndarray):
        rr = np.ascontiguousarray(r, dtype=np.uint64)
        self._check(lib().lbn_switch_fold(self.h, _p(rr)))

    def switch_final(self) -> np.ndarray:
        out = np.zeros(3, dtype=np.uint64)
        self._check(lib().lbn_switch_final(self.h, _p(out)))
        return out

    def row_eval(self, p0: np.ndarray) -> np.ndarray:
        p = np.ascontiguousarray(p0, dtype=np.uint64)
        out = np.zeros((1024, 3), dtype=np.uint64)
        self._check(lib().lbn_row_eval(self.h, _p(p), p.shape[0], _p(out)))
        return out

    def fold(self, challenges) -> np.ndarray:
        out = np.zeros((self.ring_elements, 648), dtype=np.int32)
        self._check(lib().lbn_fold(self.h, ctypes.cast(challenges, ctypes.c_void_p), 1024, _p(out), out.size))
        return out

    def autotune(self, reps: int = 3) -> dict:
        """lbn_autotune on the loaded table: {"best": splits, "ms": {splits: median ms of MAC + finish}};
        the choice is recorded for this device and shape."""
        r = TuneResult()
        self._check(lib().lbn_autotune(self.h, reps, ctypes.addressof(r)))
        return {"best": r.best, "ms": {int(r.splits[i]): round(r.ms[i], 4) for i in range(r.n)}}

    def prove_into(self, seed: bytes, out: np.ndarray):
        """lbn_prove into a caller-owned uint8 array (at least proof_len bytes); returns the proof
        length and the per-phase profile."""
        if out.dtype != np.uint8 or not out.flags.c_contiguous or not out.flags.writeable:
            raise ValueError("out must be a writable contiguous uint8 array")
        n = ctypes.c_size_t(0)
        prof = Profile()
        self._check(lib().lbn_prove(self.h, _b(seed), _p(out), out.size, ctypes.byref(n), ctypes.addressof(prof)))
        return n.value, {f: getattr(prof, f) for f, _ in Profile._fields_}

    def prove(self, seed: bytes):
        cap = proof_len(self.log_vars, self.q)
        buf = getattr(self, "_buf", None)
        if buf is None or buf.size < cap:
            buf = self._buf = np.zeros(cap, dtype=np.uint8)
        n, prof = self.prove_into(seed, buf)
        return buf[:n].tobytes(), prof

    def verify(self, seed: bytes, proof: bytes):
        """The checker (verifier/proof_verify/src/main.rs:381-389): None when `proof` verifies for
        the trace of `seed` at this shape, else the checker's reason (a str). Device errors raise.
        Regenerates the table of `seed` into this context."""
        buf = np.frombuffer(proof, dtype=np.uint8) if len(proof) else np.zeros(1, dtype=np.uint8)
        rc = lib().lbn_verify(self.h, _b(seed), _p(buf), len(proof))
        if rc == E_REJECT:
            return lib().lbn_last_error(self.h).decode()
        self._check(rc)
        return None

    def commit_primes(self) -> int:
        return lib().lbn_commit_primes(self.h)

    def commit_planes(self) -> int:
        """5 when the commitment runs on the matrix cores (moduli below 2^39), else 0."""
        return lib().lbn_commit_planes(self.h)

    def copy_probe(self):
        """One plain copy of the table's bytes on the device (a known-traffic kernel for counters)."""
        self._check(lib().lbn_copy_probe(self.h))

    def set_timing(self, enable=True):
        self._check(lib().lbn_set_timing(self.h, 1 if enable else 0))
        self._check(lib().lbn_reset_timing(self.h))

    def set_graphs(self, enable=True):
        """Replay the switch rounds as one captured HIP graph (the default) or launch them one by one."""
        self._check(lib().lbn_set_graphs(self.h, 1 if enable else 0))

    def timing(self) -> dict:
        t = Timing()
        self._check(lib().lbn_get_timing(self.h, ctypes.addressof(t)))
        return {f: getattr(t, f) for f, _ in Timing._fields_}


def ajtai_elements(q: int, elements: int, first: int, count: int, device: int = 0) -> np.ndarray:
    """Elements first .. first + count - 1 of the public matrix expansion for (q, elements), expanded
    on the GPU: a (count, 648) array of Python ints below q."""
    out = np.zeros((count, 648, 2), dtype=np.uint64)
    rc = lib().lbn_ajtai_elements(device, q & (2**64 - 1), q >> 64, elements, first, count, _p(out))
    if rc != 0:
        raise LbnError(rc, "lbn_ajtai_elements failed")
    return [[int(lo) | (int(hi) << 64) for lo, hi in row] for row in out]


def exported_symbols():
    """Names of the functions include/labinius_gpu.h declares."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(lbn_[a-z_0-9]+)\s*\(", text)))


def ctx_footprint(log_vars: int, q: int = DEFAULT_Q) -> int:
    """Device bytes of one context (the scheduler's admission unit)."""
    return lib().lbn_ctx_footprint(log_vars, q & (2**64 - 1), q >> 64)


class Scheduler:
    """Batched / continuous proving (lbn_sched_*): `workers` proofs in flight, contexts cached per
    table size under an HBM budget. submit() returns a handle; each finished proof equals the one a
    single Prover gives for the same seed and table."""

    def __init__(self, q: int = DEFAULT_Q, hbm_budget: int = 200 << 30, workers: int = 8, device: int = 0):
        self.q = q
        h = ctypes.c_void_p()
        rc = lib().lbn_sched_create(device, q & (2**64 - 1), q >> 64, hbm_budget, workers, ctypes.byref(h))
        if rc != 0:
            raise LbnError(rc, "lbn_sched_create failed")
        self.h = h
        self._live = {}

    def submit(self, log_vars: int, seed: bytes, table: bytes = None, out: np.ndarray = None):
        """Queues one proof; returns a handle for wait()/result(). The table (2^log_vars / 8 bytes)
        is generated on the device from the seed when not given. The proof goes into `out` (a
        writable contiguous uint8 array of at least proof_len bytes, for example a slice of a
        PinnedBuffer, which the proof's copies write directly) when given, else into a new array."""
        job = Job()
        job.log_vars = log_vars
        ctypes.memmove(job.seed, seed, 32)
        keep = [job]
        if table is not None:
            buf = ctypes.create_string_buffer(table, len(table))
            keep.append(buf)
            job.table = ctypes.cast(buf, ctypes.c_void_p)
        if out is None:
            out = np.empty(proof_len(log_vars, self.q), dtype=np.uint8)
        elif out.dtype != np.uint8 or not out.flags.c_contiguous or not out.flags.writeable:
            raise ValueError("out must be a writable contiguous uint8 array")
        elif out.size < proof_len(log_vars, self.q):
            raise ValueError("out is shorter than the proof")
        keep.append(out)
        job.proof = out.ctypes.data
        job.cap = out.size
        rc = lib().lbn_sched_submit(self.h, ctypes.addressof(job))
        if rc != 0:
            raise LbnError(rc, "lbn_sched_submit failed")
        self._live[id(job)] = keep
        return job

    def wait(self, job):
        return lib().lbn_sched_wait(self.h, ctypes.addressof(job))

    def drain(self):
        lib().lbn_sched_drain(self.h)

    def result(self, job):
        """(proof bytes, queue_us, run_us) of a finished job; raises on failure."""
        rc = self.wait(job)
        keep = self._live.pop(id(job))
        if rc != 0:
            raise LbnError(rc, "proof job failed")
        return keep[-1][: job.len].tobytes(), job.queue_us, job.run_us

    def set_batch(self, max_batch: int, threads: int = 8):
        """Continuous batching (lbn_sched_set_batch): a waking worker proves every job it can admit, up
        to max_batch, together (lbn_prove_batch on `threads` host threads); 1 turns it off."""
        rc = lib().lbn_sched_set_batch(self.h, max_batch, threads)
        if rc != 0:
            raise LbnError(rc, "lbn_sched_set_batch failed")

    def prewarm(self, log_vars: int, count: int):
        """Idle contexts of this size until the scheduler holds `count` of them (created now)."""
        rc = lib().lbn_sched_prewarm(self.h, log_vars, count)
        if rc != 0:
            raise LbnError(rc, "lbn_sched_prewarm failed")

    def stats(self) -> dict:
        st = SchedStats()
        lib().lbn_sched_get_stats(self.h, ctypes.addressof(st))
        return {f: getattr(st, f) for f, _ in SchedStats._fields_}

    def close(self):
        if self.h:
            lib().lbn_sched_destroy(self.h)
            self.h = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedBuffer:
    """Host memory from lbn_host_alloc (pinned, mapped for the device); `.array` is a uint8 view. A
    proof written into it skips the host-side staging copies."""

    def __init__(self, nbytes: int):
        ptr = lib().lbn_host_alloc(nbytes)
        if not ptr:
            raise LbnError(-3, "lbn_host_alloc failed")
        self.ptr, self.nbytes = ptr, nbytes
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * nbytes).from_address(ptr))

    def close(self):
        if self.ptr:
            self.array = None
            lib().lbn_host_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Commitment:
    """One lbn_cmt: the commitment of `columns` witness columns of `elements` binary ring elements
    under a key of `rows` ring rows (row r's element m is matrix element r * elements + m)."""

    def __init__(self, columns: int, elements: int, rows: int = 1, q: int = DEFAULT_Q, device: int = 0):
        self.columns, self.elements, self.rows, self.q = columns, elements, rows, q
        h = ctypes.c_void_p()
        rc = lib().lbn_cmt_create(device, columns, elements, rows, q & (2**64 - 1), q >> 64, ctypes.byref(h))
        if rc != 0:
            raise LbnError(rc, "lbn_cmt_create failed")
        self.h = h
        self.out_len = lib().lbn_cmt_len(columns, rows, q & (2**64 - 1), q >> 64)

    def _check(self, rc):
        if rc != 0:
            raise LbnError(rc, lib().lbn_cmt_error(self.h).decode())

    @property
    def primes(self) -> int:
        return lib().lbn_cmt_primes(self.h)

    def load(self, words: bytes):
        self._check(lib().lbn_cmt_load(self.h, _b(bytes(words)), len(words)))

    def load_device(self, ptr: int, nbytes: int):
        self._check(lib().lbn_cmt_load_device(self.h, ctypes.c_void_p(ptr), nbytes))

    def generate(self, seed: bytes):
        self._check(lib().lbn_cmt_generate(self.h, _b(seed)))

    def launch(self):
        self._check(lib().lbn_cmt_launch(self.h))

    def read(self) -> np.ndarray:
        """The last commitment: uint8 array of shape (rows, columns, 648 * width)."""
        out = np.empty(self.out_len, dtype=np.uint8)
        self._check(lib().lbn_cmt_read(self.h, _p(out), out.size))
        return out.reshape(self.rows, self.columns, -1)

    def commit(self) -> np.ndarray:
        self.launch()
        return self.read()

    def set_timing(self, enable: bool = True):
        self._check(lib().lbn_cmt_set_timing(self.h, 1 if enable else 0))

    def timing(self) -> dict:
        t = CmtTiming()
        self._check(lib().lbn_cmt_get_timing(self.h, ctypes.addressof(t)))
        return {f: getattr(t, f) for f, _ in CmtTiming._fields_}

    def close(self):
        if self.h:
            lib().lbn_cmt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
