// Throughput of the 32-bit integer operations that the commitment's modular products are made of,
// on the device this runs on: each thread runs 8 independent chains of 2048 steps of one operation
// (enough waves and chains that issue, not latency, bounds a chain), and the program prints the
// operations per second and the int32 VALU issue slots each costs against 256 CUs x 64 lanes x
// 2.4 GHz = 39.3 T slots/s. The roofline of the commitment kernel prices a Montgomery product from
// these figures. Build: hipcc --offload-arch=gfx950 -O3 (labinius-fast-challenge_amd/build.py).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int CHAINS = 8, STEPS = 2048, THREADS = 256, BLOCKS = 256 * 16;

enum Op { MAD64, MUL_LO, MUL_HI, MONTGOMERY, SHOUP, ADD_MIN, NOPS };
const char *const NAMES[NOPS] = {"mad_u64_u32 + xor", "mul_lo + add", "mul_hi + xor", "montgomery product",
                                 "shoup product", "add + lazy min"};
// VALU instructions per step of each chain (the step's own arithmetic)
const int INSTR[NOPS] = {2, 2, 2, 3, 4, 3};

template <int OP>
__global__ __launch_bounds__(THREADS) void k_rate(uint32_t *out, uint32_t seed, uint32_t w, uint32_t ws, uint32_t p,
                                                  uint32_t pinv_neg) {
    uint32_t a[CHAINS];
#pragma unroll
    for (int c = 0; c < CHAINS; c++) a[c] = seed * (threadIdx.x + 17 * c + 1) + blockIdx.x;
    for (int i = 0; i < STEPS; i++) {
#pragma unroll
        for (int c = 0; c < CHAINS; c++) {
            uint32_t x = a[c];
            if (OP == MAD64) {
                const uint64_t t = (uint64_t)x * w + p;
                x = (uint32_t)(t >> 32) ^ (uint32_t)t;
            } else if (OP == MUL_LO) {
                x = x * w + p;
            } else if (OP == MUL_HI) {
                x = __umulhi(x, w) ^ p;
            } else if (OP == MONTGOMERY) {  // x w R^-1 mod p in [0, 2p)
                const uint64_t t = (uint64_t)x * w;
                const uint32_t m = (uint32_t)t * pinv_neg;
                x = (uint32_t)((t + (uint64_t)m * p) >> 32);
            } else if (OP == SHOUP) {  // x w mod p in [0, 2p), ws = floor(w 2^32 / p)
                const uint32_t q = __umulhi(x, ws);
                x = x * w - q * p;
            } else {
                x = x + w;
                x = min(x, x - p);
            }
            a[c] = x;
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int c = 0; c < CHAINS; c++) r ^= a[c];
    out[(size_t)blockIdx.x * THREADS + threadIdx.x] = r;
}

template <int OP>
void launch(uint32_t *out, uint32_t w, uint32_t ws, uint32_t p, uint32_t pn) {
    k_rate<OP><<<BLOCKS, THREADS>>>(out, 3, w, ws, p, pn);
}

}  // namespace

int main() {
    uint32_t *out = nullptr;
    if (hipMalloc(&out, (size_t)BLOCKS * THREADS * sizeof(uint32_t)) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    const uint32_t p = 715809961u, w = 123456789u;  // the first CRT prime of the commitment
    const uint32_t ws = (uint32_t)(((uint64_t)w << 32) / p);
    uint32_t inv = 1;
    for (int i = 0; i < 5; i++) inv *= 2 - p * inv;  // p^-1 mod 2^32 (Newton)
    const uint32_t pn = 0u - inv;
    const double slots_per_s = 256.0 * 64 * 2.4e9;
    printf("%-20s %9s %12s %14s %12s\n", "operation", "ms", "T ops/s", "slots per op", "instructions");
    for (int op = 0; op < NOPS; op++) {
        float ms = 0;
        for (int rep = 0; rep < 2; rep++) {  // the first launch warms up
            if (hipEventRecord(e0) != hipSuccess) return 2;
            switch (op) {
                case MAD64: launch<MAD64>(out, w, ws, p, pn); break;
                case MUL_LO: launch<MUL_LO>(out, w, ws, p, pn); break;
                case MUL_HI: launch<MUL_HI>(out, w, ws, p, pn); break;
                case MONTGOMERY: launch<MONTGOMERY>(out, w, ws, p, pn); break;
                case SHOUP: launch<SHOUP>(out, w, ws, p, pn); break;
                default: launch<ADD_MIN>(out, w, ws, p, pn); break;
            }
            if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 2;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 2;
        }
        const double ops = (double)BLOCKS * THREADS * STEPS * CHAINS;
        const double rate = ops / (ms * 1e-3);
        printf("%-20s %9.3f %12.2f %14.2f %12d\n", NAMES[op], ms, rate / 1e12, slots_per_s / rate, INSTR[op]);
    }
    hipFree(out);
    return 0;
}
