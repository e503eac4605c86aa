// Device helpers shared by the kernels of liblabinius_gpu (gfx950, wave64).
//
// Binary fields: B128 = GF(2^128) in the GHASH basis (x^128 = x^7 + x^2 + x + 1), held as four
// little-endian u32 words; F162 = GF(2)[x]/(x^162 + x^81 + 1), held as six u32 words (bits
// 162..191 zero). CDNA4 has no carry-less multiply, so every product with a fixed operand is a
// GF(2)-linear map evaluated through windowed tables in LDS, and sums of selected values use the
// method of four Russians (M4RM) over half-wave groups.
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
// M4RM: acc ^= sum over one step of bit_k(sel_j) * val_j, for k < 128, wave-local.
//
// A wave (64 lanes, one row each per step) is one group. Its 64 x 128 selection matrix is
// bit-transposed in registers, as two 64 x 64 blocks (words 2c, 2c+1 of every lane form block c):
// one exchange across the half-waves (ds_bpermute), then five stages within them (DPP for
// distances 1, 2 and 8, ds_swizzle for 4 and 16). Afterwards lane r holds in words 2c, 2c+1 the
// bits of k = 64c + r over rows 0..31 and 32..63. The rows' values are folded into sixteen
// 16-entry tables (rows 4t..4t+3; the four lanes of a quad build one table from their quad's
// values, broadcast by DPP), and each k costs sixteen LDS lookups. No block barrier: the wave
// orders its own LDS traffic (__syncwarp), so the waves of a block run independently.
//
// Table layout per wave (words): plane A = [16 tables][16 slots] x uint4 (value words 0..3), then
// for V = 6 plane B = [16][16] x uint2 (words 4..5). Entry e of table t sits in slot e ^ t, which
// spreads a quad's simultaneous writes over the LDS banks while the lookups into one table (all
// lanes read table n together) stay on distinct banks.
// acc[c][v] <-> k = 64c + lane.
// ------------------------------------------------------------------------------------------

template <int V>
struct M4W {
    static_assert(V == 4 || V == 6, "M4RM values are 4 or 6 words");
    static constexpr int WORDS = 256 * V;  // LDS words per wave (16-byte aligned: read as 16-byte rows)
};

template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
template <int PATTERN>
__device__ __forceinline__ uint32_t swz(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, PATTERN);
}
// Value of lane (lane ^ D), D < 32.
template <int D>
__device__ __forceinline__ uint32_t lane_xor(uint32_t x) {
    if constexpr (D == 1) return dpp<0xB1>(x);      // quad_perm [1,0,3,2]
    else if constexpr (D == 2) return dpp<0x4E>(x);  // quad_perm [2,3,0,1]
    else if constexpr (D == 8) return dpp<0x128>(x); // row_ror:8 within a row of 16
    else return swz<0x1F | (D << 10)>(x);            // bit mode: and 0x1F, xor D
}

// One exchange stage, branch-free: the lane with bit D of r clear keeps its bits outside M and
// takes the partner's word rotated left by D into M; its partner keeps M and takes the word
// rotated right by D into the complement. Both are one lane-dependent rotate and one bit select.
template <int D, uint32_t M>
__device__ __forceinline__ void transpose_stage(uint32_t (&s)[4], int r) {
    const bool low = (r & D) == 0;
    const uint32_t keep = low ? ~M : M;
    const uint32_t rot = low ? 32u - D : (uint32_t)D;  // rotate right by rot
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const uint32_t p = lane_xor<D>(s[w]);
        const uint32_t t = __builtin_amdgcn_alignbit(p, p, rot);
        s[w] = (s[w] & keep) | (t & ~keep);
    }
}

// 64 x 64 bit transpose of each block (s[2c], s[2c+1]) across the wave: afterwards bit i of the
// block in lane r (bit i of s[2c] for i < 32, bit i - 32 of s[2c+1] otherwise) is bit r of the
// block in lane i.
__device__ __forceinline__ void transpose64(uint32_t (&s)[4], int lane) {
    const bool lo = lane < 32;
#pragma unroll
    for (int c = 0; c < 2; c++) {  // distance 32: swap the off-diagonal 32 x 32 blocks
        const uint32_t send = lo ? s[2 * c + 1] : s[2 * c];
        const uint32_t got = (uint32_t)__shfl_xor((int)send, 32);
        s[2 * c + 1] = lo ? got : s[2 * c + 1];
        s[2 * c] = lo ? s[2 * c] : got;
    }
    transpose_stage<16, 0xFFFF0000u>(s, lane);
    transpose_stage<8, 0xFF00FF00u>(s, lane);
    transpose_stage<4, 0xF0F0F0F0u>(s, lane);
    transpose_stage<2, 0xCCCCCCCCu>(s, lane);
    transpose_stage<1, 0xAAAAAAAAu>(s, lane);
}

template <int V>
__device__ __forceinline__ void m4w_step(uint32_t (&acc)[2][V], const uint32_t (&sel)[4], const uint32_t (&val)[V],
                                         uint32_t *W) {
    const int lane = threadIdx.x & 63;
    uint32_t s[4] = {sel[0], sel[1], sel[2], sel[3]};
    transpose64(s, lane);
    uint4 *A = (uint4 *)W;
    uint2 *B = (uint2 *)(W + 1024);
    {
        const int t = lane >> 2, mh = lane & 3;
        uint32_t e[4][V];
#pragma unroll
        for (int v = 0; v < V; v++) {
            const uint32_t v0 = dpp<0x00>(val[v]), v1 = dpp<0x55>(val[v]);
            const uint32_t v2 = dpp<0xAA>(val[v]), v3 = dpp<0xFF>(val[v]);
            const uint32_t base = ((mh & 1) ? v2 : 0u) ^ ((mh & 2) ? v3 : 0u);
            e[0][v] = base;
            e[1][v] = base ^ v0;
            e[2][v] = base ^ v1;
            e[3][v] = base ^ v0 ^ v1;
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int slot = t * 16 + ((4 * mh + i) ^ t);
            A[slot] = make_uint4(e[i][0], e[i][1], e[i][2], e[i][3]);
            if constexpr (V == 6) B[slot] = make_uint2(e[i][4], e[i][5]);
        }
    }
    __syncwarp();
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int n = 0; n < 16; n++) {
            const uint32_t word = s[2 * c + (n >> 3)];
            const int slot = n * 16 + (((word >> (4 * (n & 7))) & 15) ^ n);
            const uint4 a = A[slot];
            acc[c][0] ^= a.x, acc[c][1] ^= a.y, acc[c][2] ^= a.z, acc[c][3] ^= a.w;
            if constexpr (V == 6) {
                const uint2 b = B[slot];
                acc[c][4] ^= b.x, acc[c][5] ^= b.y;
            }
        }
    __syncwarp();
}

// Block flush: every wave leaves its 128 x V sums (k-major) in its own LDS region
// `lds + wave * M4W<V>::WORDS`; after a block barrier they are XORed together and ATOMIC adds the
// result into out with atomicXor, otherwise the block owns out and stores it.
template <int V, bool ATOMIC>
__device__ __forceinline__ void m4w_flush(const uint32_t (&acc)[2][V], uint32_t *lds, uint32_t *out) {
    const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
    uint32_t *W = lds + (threadIdx.x >> 6) * M4W<V>::WORDS;
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int v = 0; v < V; v++) W[(64 * c + lane) * V + v] = acc[c][v];
    __syncthreads();
    for (int i = threadIdx.x; i < 128 * V; i += blockDim.x) {
        uint32_t x = 0;
        for (int w = 0; w < waves; w++) x ^= lds[w * M4W<V>::WORDS + i];
        if (ATOMIC) {
            if (x) atomicXor(out + i, x);
        } else {
            out[i] = x;
        }
    }
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
// x * R^-1 mod p, in [0, 2p), for any x < 2^32 p (so x + m p stays below 2^64).
__device__ __forceinline__ uint32_t redc64(uint64_t x, PC c) {
    const uint32_t m = (uint32_t)x * c.pinv_neg;
    return (uint32_t)((x + (uint64_t)m * c.p) >> 32);
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
