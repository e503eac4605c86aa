// Device helpers shared by the kernels of liblabinius_gpu (gfx950, wave64).
//
// Binary fields: B128 = GF(2^128) in the GHASH basis (x^128 = x^7 + x^2 + x + 1), held as four
// little-endian u32 words; F162 = GF(2)[x]/(x^162 + x^81 + 1), held as six u32 words (bits
// 162..191 zero). CDNA4 has no carry-less multiply, so every product with a fixed operand is a
// GF(2)-linear map evaluated through windowed tables in LDS, and sums of selected values use the
// method of four Russians (M4RM) over 16-lane groups.
//
// Ring arithmetic: 31-bit primes p = 1 (mod 1944) in Montgomery form (R = 2^32).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbn {
namespace dev {

// ------------------------------------------------------------------------------------------
// GF(2^128) and F162 shifts
// ------------------------------------------------------------------------------------------

// y * x in GF(2^128), y as (lo, hi).
__device__ __forceinline__ void b128_mulx(uint64_t &lo, uint64_t &hi) {
    uint64_t carry = hi >> 63;
    hi = (hi << 1) | (lo >> 63);
    lo = (lo << 1) ^ (carry ? 0x87ull : 0ull);
}

// y * x in F162, y as three u64 limbs (limb 2 below 2^34).
__device__ __forceinline__ void f162_mulx(uint64_t &a0, uint64_t &a1, uint64_t &a2) {
    uint64_t top = (a2 >> 33) & 1;  // bit 161
    a2 = ((a2 << 1) | (a1 >> 63)) & ((1ull << 34) - 1);
    a1 = (a1 << 1) | (a0 >> 63);
    a0 = a0 << 1;
    if (top) {  // x^162 = x^81 + 1
        a0 ^= 1ull;
        a1 ^= 1ull << 17;  // bit 81 = limb 1 bit 17
    }
}

// Reduce a value of up to 320 bits (five u64) modulo x^162 + x^81 + 1.
__device__ __forceinline__ void f162_reduce5(uint64_t v[5], uint64_t out[3]) {
#pragma unroll
    for (int it = 0; it < 2; it++) {
        uint64_t h0 = (v[2] >> 34) | (v[3] << 30);
        uint64_t h1 = (v[3] >> 34) | (v[4] << 30);
        uint64_t h2 = v[4] >> 34;
        v[2] &= (1ull << 34) - 1;
        v[3] = 0;
        v[4] = 0;
        // v ^= H
        v[0] ^= h0;
        v[1] ^= h1;
        v[2] ^= h2;
        // v ^= H << 81: H << 81 = H << 64 << 17
        v[1] ^= h0 << 17;
        v[2] ^= (h0 >> 47) | (h1 << 17);
        v[3] ^= (h1 >> 47) | (h2 << 17);
        v[4] ^= h2 >> 47;
    }
    out[0] = v[0];
    out[1] = v[1];
    out[2] = v[2];
}

// Variable product a * b in F162, bit serial (setup kernels only).
__device__ inline void f162_mul_slow(const uint64_t a[3], const uint64_t b[3], uint64_t out[3]) {
    uint64_t r0 = 0, r1 = 0, r2 = 0;
    for (int i = 161; i >= 0; i--) {
        f162_mulx(r0, r1, r2);
        if ((a[i >> 6] >> (i & 63)) & 1) {
            r0 ^= b[0];
            r1 ^= b[1];
            r2 ^= b[2];
        }
    }
    out[0] = r0;
    out[1] = r1;
    out[2] = r2;
}

// ------------------------------------------------------------------------------------------
// M4RM: acc[k] ^= sum over a block step of bit_k(sel_j) * val_j, for k < 128.
//
// 256 threads, one element per thread per step. Within each group of 16 lanes the 16 x 128
// selection matrix is bit-transposed with four xor-shuffle stages, so that lane r holds, for each
// 16-bit chunk c, the bits of k = 16c + r over the group's 16 rows. The rows' values are folded
// into four 16-entry tables (rows 4t..4t+3), and each k costs four LDS lookups. Lane r keeps the
// accumulators of its eight k values: acc[c] <-> k = 16c + r.
// ------------------------------------------------------------------------------------------

constexpr int M4_THREADS = 256;

template <int V>
struct M4Lds {
    uint32_t val[V * M4_THREADS];            // staged values, value-major
    uint32_t tab[(M4_THREADS / 16) * 64 * V];  // per group: 4 tables x 16 entries x V
};

__device__ __forceinline__ void transpose16(uint32_t s[4], int r) {
    const uint32_t masks[4] = {0xFF00FF00u, 0xF0F0F0F0u, 0xCCCCCCCCu, 0xAAAAAAAAu};
    const int ds[4] = {8, 4, 2, 1};
#pragma unroll
    for (int st = 0; st < 4; st++) {
        const int d = ds[st];
        const uint32_t M = masks[st];
        const bool low = (r & d) == 0;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            uint32_t partner = (uint32_t)__shfl_xor((int)s[w], d);
            s[w] = low ? ((s[w] & ~M) | ((partner << d) & M)) : ((s[w] & M) | ((partner >> d) & ~M));
        }
    }
}

template <int V>
__device__ __forceinline__ void m4rm_step(uint32_t (&acc)[8][V], const uint32_t (&sel)[4],
                                          const uint32_t (&val)[V], M4Lds<V> &L) {
    const int tid = threadIdx.x, r = tid & 15, grp = tid >> 4;
#pragma unroll
    for (int v = 0; v < V; v++) L.val[v * M4_THREADS + tid] = val[v];
    uint32_t s[4] = {sel[0], sel[1], sel[2], sel[3]};
    transpose16(s, r);
    __syncthreads();
    {
        const int t = r >> 2, mh = r & 3, row0 = grp * 16 + 4 * t;
        uint32_t *T = L.tab + (grp * 64 + t * 16 + 4 * mh) * V;
#pragma unroll
        for (int v = 0; v < V; v++) {
            const uint32_t *col = L.val + v * M4_THREADS;
            uint32_t base = ((mh & 1) ? col[row0 + 2] : 0u) ^ ((mh & 2) ? col[row0 + 3] : 0u);
            uint32_t v0 = col[row0], v1 = col[row0 + 1];
            T[v] = base;
            T[V + v] = base ^ v0;
            T[2 * V + v] = base ^ v1;
            T[3 * V + v] = base ^ v0 ^ v1;
        }
    }
    __syncthreads();
    const uint32_t *G = L.tab + grp * 64 * V;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const uint32_t chunk = (s[c >> 1] >> (16 * (c & 1))) & 0xFFFFu;
        const uint32_t *a = G + (0 * 16 + (chunk & 15)) * V;
        const uint32_t *b = G + (1 * 16 + ((chunk >> 4) & 15)) * V;
        const uint32_t *cc = G + (2 * 16 + ((chunk >> 8) & 15)) * V;
        const uint32_t *d = G + (3 * 16 + (chunk >> 12)) * V;
#pragma unroll
        for (int v = 0; v < V; v++) acc[c][v] ^= a[v] ^ b[v] ^ cc[v] ^ d[v];
    }
    __syncthreads();
}

// Block-reduce the accumulators and XOR them into out[k * V + v] (global, atomic).
// `red` needs 4 * 128 * V words of LDS (it may alias M4Lds storage after the last step).
template <int V>
__device__ __forceinline__ void m4rm_flush(uint32_t (&acc)[8][V], uint32_t *red, uint32_t *out) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int c = 0; c < 8; c++)
#pragma unroll
        for (int v = 0; v < V; v++) {
            uint32_t x = acc[c][v];
            x ^= (uint32_t)__shfl_xor((int)x, 16);
            x ^= (uint32_t)__shfl_xor((int)x, 32);
            acc[c][v] = x;
        }
    __syncthreads();
    if (lane < 16)
#pragma unroll
        for (int c = 0; c < 8; c++)
#pragma unroll
            for (int v = 0; v < V; v++) red[(wave * 128 + 16 * c + lane) * V + v] = acc[c][v];
    __syncthreads();
    for (int i = tid; i < 128 * V; i += M4_THREADS) {
        uint32_t x = red[i] ^ red[128 * V + i] ^ red[256 * V + i] ^ red[384 * V + i];
        if (x) atomicXor(out + i, x);
    }
}

// Same reduction, but the block owns `out` exclusively (plain stores).
template <int V>
__device__ __forceinline__ void m4rm_flush_store(uint32_t (&acc)[8][V], uint32_t *red, uint32_t *out) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int c = 0; c < 8; c++)
#pragma unroll
        for (int v = 0; v < V; v++) {
            uint32_t x = acc[c][v];
            x ^= (uint32_t)__shfl_xor((int)x, 16);
            x ^= (uint32_t)__shfl_xor((int)x, 32);
            acc[c][v] = x;
        }
    __syncthreads();
    if (lane < 16)
#pragma unroll
        for (int c = 0; c < 8; c++)
#pragma unroll
            for (int v = 0; v < V; v++) red[(wave * 128 + 16 * c + lane) * V + v] = acc[c][v];
    __syncthreads();
    for (int i = tid; i < 128 * V; i += M4_THREADS)
        out[i] = red[i] ^ red[128 * V + i] ^ red[256 * V + i] ^ red[384 * V + i];
}

// ------------------------------------------------------------------------------------------
// Montgomery arithmetic modulo a 31-bit prime (R = 2^32).
// ------------------------------------------------------------------------------------------
struct PC {
    uint32_t p, pinv_neg;
};

// a * b * R^-1 mod p, in [0, 2p).
__device__ __forceinline__ uint32_t mont_lazy(uint32_t a, uint32_t b, PC c) {
    uint64_t t = (uint64_t)a * b;
    uint32_t m = (uint32_t)t * c.pinv_neg;
    uint64_t u = t + (uint64_t)m * c.p;
    return (uint32_t)(u >> 32);
}
// Corrections use unsigned min: for r in [0, 2p), min(r, r - p) is r mod p (r - p wraps when
// r < p). No compare-and-select, hence no scalar condition register between VALU ops.
__device__ __forceinline__ uint32_t mont(uint32_t a, uint32_t b, PC c) {
    const uint32_t r = mont_lazy(a, b, c);
    return min(r, r - c.p);
}
__device__ __forceinline__ uint32_t addm(uint32_t a, uint32_t b, uint32_t p) {
    const uint32_t r = a + b;
    return min(r, r - p);
}
__device__ __forceinline__ uint32_t subm(uint32_t a, uint32_t b, uint32_t p) {
    const uint32_t d = a - b;
    return min(d, d + p);
}

}  // namespace dev
}  // namespace lbn
