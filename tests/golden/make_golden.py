"""Writes tests/golden/blake3_vectors.json from an independent BLAKE3 implementation.

The reference hashes through the `blake3` crate 1.8.7 (solution/proof_generate/Cargo.lock:21-23),
which is not vendored, and its own tests pin BLAKE3 against nothing external
(pcs/src/challenge.rs:864-886 cross-checks two of its own code paths). The independent
implementation used here is the one libLLVM-15 exports (`llvm_blake3_*`), called through ctypes.
It exists in the build container only; the fixtures it writes are committed, and nothing at test
time needs libLLVM.

Every vector is a plain input/output pair: BLAKE3 in its three modes with XOF reads at offsets,
the witness-table prefix for fixed seeds (verifier/proof_verify/src/main.rs:130-141), transcript
derivations as the byte stream the transcript hashes (pcs/src/challenge.rs:226-326, 724-749), the
matrix expansion stream (verifier/proof_verify/src/rq.rs:62-75), and challenge sign words
(pcs/src/challenge.rs:429-441, 875-879).

Run: python tests/golden/make_golden.py
"""
import ctypes
import json
import os
import struct

LIB = "/usr/lib/x86_64-linux-gnu/libLLVM-15.so"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "blake3_vectors.json")


class B3:
    def __init__(self):
        self.lib = ctypes.CDLL(LIB)
        for name in ("init", "init_keyed", "init_derive_key", "update", "finalize_seek"):
            setattr(self, name, getattr(self.lib, "llvm_blake3_hasher_" + name))
        self.finalize_seek.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_size_t]
        self.update.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]

    def new(self, mode=0, key=None):
        h = ctypes.create_string_buffer(8192)
        if mode == 0:
            self.init(h)
        elif mode == 1:
            self.init_keyed(h, ctypes.c_char_p(key))
        else:
            self.init_derive_key(h, ctypes.c_char_p(key))
        return h

    def upd(self, h, data):
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        self.update(h, buf, len(data))

    def out(self, h, offset, n):
        o = ctypes.create_string_buffer(n)
        self.finalize_seek(h, offset, o, n)
        return o.raw

    def clone(self, h):
        c = ctypes.create_string_buffer(8192)
        ctypes.memmove(c, h, 8192)
        return c


def u64(x):
    return struct.pack("<Q", x)


def main():
    b = B3()
    vec = {"hash": [], "table": [], "transcript": [], "matrix_stream": [], "signs": []}
    test_input = bytes(i % 251 for i in range(102400))
    key = b"whats the Elvish word for friend"
    ctx = b"BLAKE3 2019-12-27 16:29:52 test vectors context"
    for n in (0, 1, 63, 64, 65, 1023, 1024, 1025, 2048, 2049, 3072, 3073, 5000, 31744, 102400):
        for mode, k in ((0, None), (1, key), (2, ctx)):
            h = b.new(mode, k)
            b.upd(h, test_input[:n])
            vec["hash"].append({
                "mode": mode, "len": n,
                "out0": b.out(h, 0, 131).hex(),
                "out1000": b.out(h, 1000, 77).hex(),
            })
    seeds = [bytes([9] * 32), bytes([42] * 32), bytes(range(32))]
    for seed in seeds:
        h = b.new(1, seed)
        b.upd(h, b"labinius-fast/binary-table/v1")
        vec["table"].append({
            "seed": seed.hex(),
            "prefix": b.out(h, 0, 4096).hex(),
            "at_2pow28_minus_64": b.out(h, (1 << 28) - 64, 64).hex(),
            "at_131008": b.out(h, 131008, 64).hex(),
        })
    # Transcript::new(domain); absorb_bytes(msg); two sample_b128(label, n); one sample_f162.
    for domain, msg, label, n in ((b"labinius-fast/v1", b"", b"r-lo", 7),
                                  (b"labinius/test", bytes(range(200)), b"switch-round", 3),
                                  (b"labinius/test", bytes(3000), b"r-hi", 24)):
        h = b.new(0)
        b.upd(h, u64(len(domain)) + domain + u64(len(msg)) + msg)
        samples = []
        for counter, width in ((0, 16), (1, 16), (2, 24)):
            c = b.clone(h)
            b.upd(c, u64(counter) + u64(len(label)) + label)
            samples.append(b.out(c, 0, width * n).hex())
        vec["transcript"].append({"domain": domain.hex(), "msg": msg.hex(), "label": label.hex(), "n": n,
                                  "b128_first": samples[0], "b128_second": samples[1], "f162_raw": samples[2]})
    # Matrix expansion streams: keyed [0x5a;32], domain, q (16 LE), elements, index.
    for q, elements, index in ((35039150080, 4096, 0), (35039150080, 4096, 4095), (1 << 36, 2, 1),
                               (55144219393, 4096, 17), ((1 << 96) - 1, 64, 5)):
        h = b.new(1, bytes([0x5A] * 32))
        b.upd(h, b"labinius-fast/ajtai-matrix/v1" + q.to_bytes(16, "little") + u64(elements) + u64(index))
        vec["matrix_stream"].append({"q": str(q), "elements": elements, "index": index,
                                     "stream": b.out(h, 0, 2048).hex()})
    for positions in ([0, 5, 9, 17, 30, 31, 32, 33, 40, 44, 50, 51, 60, 63, 64, 70, 80, 81, 90, 99,
                       100, 101, 120, 130, 140, 150, 160, 161],
                      list(range(28)), list(range(134, 162))):
        h = b.new(2, b"labinius 2026 challenge signs v1")
        b.upd(h, u64(len(positions)) + bytes(positions))
        vec["signs"].append({"positions": positions, "word": struct.unpack("<I", b.out(h, 0, 4))[0]})
    with open(OUT, "w") as f:
        json.dump(vec, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
