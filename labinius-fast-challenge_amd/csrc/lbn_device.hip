// liblabinius_gpu: the MI355X kernels of the Labinius prover and the device half of the C ABI.
//
// Passes over the 2^k-bit table (k = 31: 2^24 words, 256 MiB resident in HBM):
//   table  : BLAKE3 XOF of the seed, one root compression per 64-byte block (table writer)
//   eq     : eq_hi[j] = EH[j >> 12] * EL[j & 4095], a per-block "multiply by EH" byte table
//   partials: v[k] = sum_j bit_k(w_j) eq_hi[j]                       (M4RM, 4-word values)
//   switch : round 0 from the words, then fold+message fused rounds (M4RM, 6-word values)
//   row    : row[c] = sum_i eq(p0,i) lift(w_{c,i})                   (M4RM + x^b recombination)
//   fold   : v_m = sum_j c_j(-X^4) W_{j,m}, bit-sliced counters per challenge position
//   commit : per CRT prime, binary split/twist -> mixed-radix 2/3 DIF NTT -> MAC with the
//            transformed matrix -> inverse NTT -> Garner -> mod Q -> packed bytes
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "lbn_internal.h"
#include "lbn_kernels.h"

using namespace lbn;
using namespace lbn::dev;

namespace lbn {
bool prime_tables(const PrimeInfo &pi, std::vector<uint32_t> &tabs);
}  // namespace lbn


// The 18-th roots of unity of one prime: w18^j and w18^-j for j < 9, Montgomery form.
struct Roots18 {
    uint32_t f[9], i[9];
};

struct SwitchDev;
struct ToepOps;
struct CvJob;

struct lbn::DeviceState {
    hipStream_t stream = nullptr;
    uint4 *table = nullptr;
    uint4 *E[2] = {nullptr, nullptr};
    uint32_t *P[2] = {nullptr, nullptr};
    uint32_t *T = nullptr;  // [0,768): T0, [768,1536): Tinf, [2048,3584): the odd rounds' T0, Tinf
    uint32_t *pset = nullptr;  // the prefix sets and the first switch rounds' data (layout: namespace ps)
    uint4 *EL = nullptr, *EH = nullptr;  // EH = EL + 4096 (one allocation)
    uint32_t *eqrow = nullptr, *eqlo = nullptr;
    int *fold_off = nullptr;
    uint16_t *fold_cols = nullptr;  // inside fold_off's allocation
    uint32_t *ntt_tabs = nullptr;   // MAX_PRIMES * tab::SIZE
    uint32_t *ahat = nullptr;       // nprimes * RE * 648
    uint32_t *cacc = nullptr;       // nprimes * splits * 1024 * 648: the MAC's partial sums (< p)
    uint8_t *cbytes = nullptr;      // 1024 * 648 * width
    uint64_t *amat = nullptr;       // RE * 648 * 2 (staging for the matrix)
    size_t ahat_elems = 0;
    int alloc_primes = 0, alloc_width = 0;  // sizes of cacc / cbytes / cvs / hstage
    int splits = 0;                 // ring-element ranges of the MAC (blocks per column group)
    uint32_t *cvs = nullptr;        // chunk chaining values of the commitment bytes
    // Pinned host staging: the commitment bytes (copied on copy_stream once the commitment is
    // ready, while the opening runs) and the fold output, which the fold kernel writes directly.
    uint8_t *hstage = nullptr;
    bool staged_direct = false;     // the last commitment was copied straight into the proof buffer
    int32_t *hfold = nullptr;
    uint8_t *hbox = nullptr;        // pinned mailbox for the small transfers (regions in namespace mb)
    uint8_t *hbox_dev = nullptr;    // its device address: kernels write small results straight into it
    bool msg_clean = false;         // the switch message sums are zero (cleared with the partials)
    uint32_t *hcvs = nullptr;       // pinned: chunk chaining values of the commitment
    uint32_t *hcvs_dev = nullptr;   // its device address
    struct SwitchDev *sw = nullptr; // the switch rounds' device state (transcript, L-map, outputs)
    ChalArg *chal = nullptr;        // the challenge sampler's input (its output goes to the mailbox)
    int32_t *fold_host = nullptr;   // device address of hfold
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_commit = nullptr, ev_staged = nullptr;
    hipEvent_t ev[2 * 9] = {};           // start/stop per timed pass
    Roots18 roots[MAX_PRIMES] = {};
    // The switch rounds' launches (the same for every proof of this context: shapes, buffers
    // and the device state they read are fixed) captured once and replayed as one graph launch.
    hipGraphExec_t switch_graph = nullptr;
    bool graphs = true;
    hipStream_t own_stream = nullptr, own_copy = nullptr;  // created here; stream / copy_stream may be a caller's
    // The matrix-core commitment (moduli below 2^39): its operands, shared by the contexts of one
    // (device, Q, shape), and its partial sums.
    std::shared_ptr<::ToepOps> toep;
    int64_t *tpart = nullptr;
    int splits_try = 0;  // the MAC's split count while the autotuner times it (0: the chosen one)
    std::unique_ptr<::CvJob> cvjob;  // the staged commitment hashing of the batch protocol
    // A host table whose upload is deferred (lbn_load_table_async): the commitment copies it in
    // ring-element ranges, each range's MAC waiting for its copy (events on the copy stream).
    const uint8_t *pend = nullptr;
    hipEvent_t ev_chunk[4] = {};
};

static int capture_switch(lbn_ctx *ctx);

// Regions of a context's pinned mailbox. Every small host<->device transfer goes through it, so
// each is a plain DMA on the context's stream (pageable copies take the runtime's shared staging
// path, which serialises proofs running on other threads). Each entry point synchronises its stream
// before returning, so a region is never rewritten while a transfer from it is in flight.
namespace mb {
constexpr size_t EL = 0;                  // 4096 B128
constexpr size_t EH = EL + 65536;         // 2^(log_packed - 12) B128 <= 64 KiB (EL and EH: one copy)
constexpr size_t EQ = EH + 65536;         // row eq tables: 128 + 128 F162
constexpr size_t FOFF = EQ + 8192;        // fold list offsets (325 int) and position order (162 int)
constexpr size_t FCOLS = FOFF + 2048;     // fold column lists (<= 1024 x 32 u16); FOFF..FCOLS: one copy
constexpr size_t T = FCOLS + 65536;       // switch message sums (1536 u32)
constexpr size_t PART = T + 6144;         // partials (512 u32)
constexpr size_t ROW = PART + 2048;       // row (1024 F162)
constexpr size_t MISC = ROW + 24576;      // the fold's range flag, the last two switch values (256 B)
constexpr size_t HEAD = MISC + 256;       // commitment bytes before the hashed chunks (<= 1 KiB)
constexpr size_t TAIL = HEAD + 1024;      // and after them (<= 1 KiB)
constexpr size_t SW = TAIL + 1024;        // the device switch state (SwitchDev, < 8 KiB)
constexpr size_t CH = SW + 8192;          // the challenge sampler's input (ChalArg, < 8 KiB)
constexpr size_t CHOUT = CH + 8192;       // the sampled challenges (1024 lbn_challenge)
constexpr size_t HANDOFF_H2 = (size_t)1 << (SWITCH_HOST_STEPS - 1);  // items of the host's first fold
constexpr size_t HP = CHOUT + 1024 * 40;  // the switch handoff: that fold's input (4 h2 F162)
constexpr size_t HE = HP + 4 * HANDOFF_H2 * 24;  // and its eq values (2 h2 B128)
constexpr size_t LAM = HE + 2 * HANDOFF_H2 * 16;  // the prefix rounds' lambda values (128 B128)
constexpr size_t SIZE = LAM + 128 * 16;
}  // namespace mb

// The prefix-set layout of a context's pset buffer (words): the 64 sets M_u (128 B128 each), the
// partials' accumulator (128 B128), Q of the first switch rounds (64 x 128 F162) and their lambda
// values (128 B128).
namespace ps {
constexpr size_t M = 0, PART = 64 * 128 * 4, Q = PART + 512, LAM = Q + 64 * 128 * 6, WORDS = LAM + 512;
}


// Timed passes (indices into lbn_timing's fields, in order).
enum Pass { T_TABLE = 0, T_EQ, T_PARTIALS, T_SWITCH0, T_ROUNDS, T_ROW, T_FOLD, T_MAC, T_FINISH, T_COUNT };

static void tstart(lbn_ctx *ctx, int pass) {
    if (ctx->timing) hipEventRecord(ctx->dev->ev[2 * pass], ctx->dev->stream);
}
static void tstop(lbn_ctx *ctx, int pass) {
    if (ctx->timing) hipEventRecord(ctx->dev->ev[2 * pass + 1], ctx->dev->stream);
}
// After the stream has been synchronised: add the pass's elapsed time.
static void tcollect(lbn_ctx *ctx, int pass) {
    if (!ctx->timing) return;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->dev->ev[2 * pass], ctx->dev->ev[2 * pass + 1]) == hipSuccess)
        (&ctx->times.table_ms)[pass] += ms;
    else
        (void)hipGetLastError();  // a pass not run since the events were made: not an error of the proof
    if (pass == T_ROUNDS) ctx->times.switch_rounds++;
}

#define HIPCHK(ctx, x)                                                              \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess)                                                       \
            return (ctx)->fail(LBN_E_DEVICE, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

// =============================================================================================
// small kernels
// =============================================================================================

__global__ void k_copy_u4(const uint4 *src, uint4 *dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

static void fill_u32(hipStream_t s, void *p, size_t n, uint32_t v = 0) {
    if (!n) return;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    k_fill_u32<<<blocks, 256, 0, s>>>((uint32_t *)p, n, v);
}

// ---------------------------------------------------------------------------------------------
// The witness table on the device: BLAKE3 keyed by the seed, updated with the 29-byte context
// "labinius-fast/binary-table/v1", read as XOF (verifier/table_writer/src/main.rs:14-22). The
// input is one block of one chunk, so output block i is the root compression with counter i.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

__device__ __forceinline__ void b3_g(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d, uint32_t x,
                                     uint32_t y) {
    a = a + b + x;
    d = rotr32(d ^ a, 16);
    c = c + d;
    b = rotr32(b ^ c, 12);
    a = a + b + y;
    d = rotr32(d ^ a, 8);
    c = c + d;
    b = rotr32(b ^ c, 7);
}

struct B3Root {
    uint32_t cv[8];
    uint32_t m[16];
    uint32_t len, flags;
};

// The seven rounds of the BLAKE3 compression on state s with message m (m is permuted in place).
__device__ __forceinline__ void b3_rounds(uint32_t (&s)[16], uint32_t (&m)[16]) {
    const int perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
#pragma unroll
    for (int r = 0; r < 7; r++) {
        b3_g(s[0], s[4], s[8], s[12], m[0], m[1]);
        b3_g(s[1], s[5], s[9], s[13], m[2], m[3]);
        b3_g(s[2], s[6], s[10], s[14], m[4], m[5]);
        b3_g(s[3], s[7], s[11], s[15], m[6], m[7]);
        b3_g(s[0], s[5], s[10], s[15], m[8], m[9]);
        b3_g(s[1], s[6], s[11], s[12], m[10], m[11]);
        b3_g(s[2], s[7], s[8], s[13], m[12], m[13]);
        b3_g(s[3], s[4], s[9], s[14], m[14], m[15]);
        uint32_t t[16];
#pragma unroll
        for (int k = 0; k < 16; k++) t[k] = m[perm[k]];
#pragma unroll
        for (int k = 0; k < 16; k++) m[k] = t[k];
    }
}

__global__ void k_table_xof(B3Root root, uint32_t *out, size_t blocks) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= blocks) return;
    const uint32_t IV0 = 0x6A09E667u, IV1 = 0xBB67AE85u, IV2 = 0x3C6EF372u, IV3 = 0xA54FF53Au;
    uint32_t s[16] = {root.cv[0], root.cv[1], root.cv[2], root.cv[3], root.cv[4], root.cv[5], root.cv[6],
                      root.cv[7], IV0, IV1, IV2, IV3, (uint32_t)i, (uint32_t)(i >> 32), root.len, root.flags};
    uint32_t m[16];
#pragma unroll
    for (int k = 0; k < 16; k++) m[k] = root.m[k];
    b3_rounds(s, m);
    uint32_t *o = out + 16 * i;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        o[k] = s[k] ^ s[k + 8];
        o[k + 8] = s[k + 8] ^ root.cv[k];
    }
}

// rq::ajtai_matrix (verifier/proof_verify/src/rq.rs:53-144) on the device: element e (index
// first + e of the expansion for (q, elements)) is the keyed BLAKE3 XOF of domain || q || u64(elements)
// || u64(index), one 61-byte block whose root node is `root` with the index in bytes 53..60; its
// coefficients are the first 648 width-byte little-endian draws of the stream that, masked to
// bits(q - 1), are below q. One wave per element: each pass its 64 lanes make 64 output blocks (4 KiB
// of stream, after the previous pass's last block, kept in LDS), then the wave takes the draws that
// end in that stretch, 64 at a time in stream order, placing the accepted ones by ballot rank. The
// acceptance rate q / 2^bits is above 1/2, so a pass yields about 4096 / (2 width) coefficients;
// a wave that has not finished after AJ_PASSES passes sets *fail.
constexpr int AJ_WAVES = 4, AJ_PASSES = 64, AJ_WORDS = 16 * 65 + 4;
__global__ __launch_bounds__(64 * AJ_WAVES) void k_ajtai(B3Root root, uint64_t first, int count, int width,
                                                         uint64_t mask_lo, uint32_t mask_hi, uint64_t q_lo,
                                                         uint32_t q_hi, uint64_t *out, uint32_t *fail) {
    __shared__ uint32_t sbuf[AJ_WAVES][AJ_WORDS];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int e = blockIdx.x * AJ_WAVES + wv;
    if (e >= count) return;  // the whole wave: no block barrier below
    uint32_t *B = sbuf[wv];
    const uint64_t index = first + (uint64_t)e;
    uint32_t mt[16];
#pragma unroll
    for (int k = 0; k < 16; k++) mt[k] = root.m[k];
    mt[13] |= (uint32_t)(index << 8), mt[14] = (uint32_t)(index >> 24), mt[15] = (uint32_t)(index >> 56);
    const uint32_t IV0 = 0x6A09E667u, IV1 = 0xBB67AE85u, IV2 = 0x3C6EF372u, IV3 = 0xA54FF53Au;
    const uint64_t lanes_below = (1ull << lane) - 1;
    uint64_t *dst = out + (size_t)e * N * 2;
    int have = 0;
    uint64_t t = 0;  // the next draw
    for (int pass = 0; have < N && pass < AJ_PASSES; pass++) {
        {
            const uint64_t blk = 64ull * pass + lane;
            uint32_t st[16] = {root.cv[0], root.cv[1], root.cv[2], root.cv[3], root.cv[4], root.cv[5], root.cv[6],
                               root.cv[7], IV0, IV1, IV2, IV3, (uint32_t)blk, (uint32_t)(blk >> 32), root.len, root.flags};
            uint32_t m[16];
#pragma unroll
            for (int k = 0; k < 16; k++) m[k] = mt[k];
            b3_rounds(st, m);
            uint32_t *o = B + 16 * (1 + lane);
#pragma unroll
            for (int k = 0; k < 8; k++) o[k] = st[k] ^ st[k + 8], o[k + 8] = st[k + 8] ^ root.cv[k];
        }
        __syncwarp();
        // B holds stream bytes [4096 pass - 64, 4096 pass + 4096) (the first 64 only after pass 0)
        const int64_t base = 4096ll * pass - 64, end = 4096ll * (pass + 1);
        while (have < N) {
            const int64_t at = (int64_t)(t + lane) * width;
            const bool inside = at + width <= end;
            const uint64_t in = __ballot(inside);
            if (!in) break;
            bool acc = false;
            uint64_t lo = 0;
            uint32_t hi = 0;
            if (inside) {
                const int off = (int)(at - base), w0 = off >> 2;
                const uint32_t sh = 8u * (uint32_t)(off & 3);
                const uint32_t W0 = B[w0], W1 = B[w0 + 1], W2 = B[w0 + 2], W3 = B[w0 + 3];
                const uint32_t x0 = __builtin_amdgcn_alignbit(W1, W0, sh), x1 = __builtin_amdgcn_alignbit(W2, W1, sh);
                const uint32_t x2 = __builtin_amdgcn_alignbit(W3, W2, sh);
                lo = ((uint64_t)x0 | ((uint64_t)x1 << 32)) & mask_lo;
                hi = x2 & mask_hi;
                acc = hi < q_hi || (hi == q_hi && lo < q_lo);
            }
            const uint64_t bal = __ballot(acc);
            const int rank = have + __popcll(bal & lanes_below);
            if (acc && rank < N) dst[2 * rank] = lo, dst[2 * rank + 1] = hi;
            have += __popcll(bal);
            t += (uint64_t)__popcll(in);
            if (in != ~0ull) break;  // the rest end past this stretch
        }
        __syncwarp();
        if (lane < 16) B[lane] = B[16 * 64 + lane];  // this pass's last block opens the next stretch
        __syncwarp();
    }
    if (have < N && lane == 0) atomicOr(fail, 1u);
}

// Chaining values of whole 1 KiB chunks of a byte string (the commitment as the transcript
// absorbs it): thread c hashes bytes [off + 1024 c, +1024) as chunk index first + c, sixteen
// 64-byte blocks chained from the key. The start need not be 4-byte aligned: message words are
// funnel-shifted from aligned loads (the buffer has slack past the data).
struct B3Key {
    uint32_t k[8];
    uint32_t flags;
};

// One wave per window of 64 chunk indices (absolute indices 64 W .. 64 W + 63, lane = index mod 64):
// the lanes whose chunk lies in the run hash it, then the wave merges the aligned complete
// subtrees of up to 64 chunks that lie in the run (CV_WIN_LEVELS levels), leaving each subtree's
// value at its first chunk's slot, as k_cv_tree does for the levels above.
constexpr int CV_WIN_LEVELS = 6;
__device__ __forceinline__ void chunk_cvs_body(int bid, const uint8_t *bytes, size_t off, uint64_t first, size_t count,
                                               const B3Key &key, uint32_t *cvs, uint32_t *s_cv) {
    const int lane = threadIdx.x;
    const uint64_t a = (first / 64 + bid) * 64 + lane, end = first + count;
    const bool in = a >= first && a < end;
    const size_t c = (size_t)(a - first);
    const uint64_t index = a;
    uint32_t cv[8];
#pragma unroll
    for (int k = 0; k < 8; k++) cv[k] = key.k[k];
    const size_t base = off + (in ? c : 0) * 1024;
    const uint32_t *w = (const uint32_t *)(bytes + (base & ~(size_t)3));
    const uint32_t sh = (uint32_t)(base & 3) * 8;
    for (int b = 0; b < 16 && in; b++) {
        uint32_t m[16];
        uint32_t lo = w[16 * b];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t hi = w[16 * b + i + 1];
            m[i] = __builtin_amdgcn_alignbit(hi, lo, sh);
            lo = hi;
        }
        const uint32_t flags = key.flags | (b == 0 ? 1u : 0u) | (b == 15 ? 2u : 0u);  // CHUNK_START, CHUNK_END
        uint32_t st[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                           0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                           (uint32_t)index, (uint32_t)(index >> 32), 64u, flags};
        b3_rounds(st, m);
#pragma unroll
        for (int k = 0; k < 8; k++) cv[k] = st[k] ^ st[k + 8];
    }
    for (int l = 0; l < CV_WIN_LEVELS; l++) {
        const uint64_t span = 2ull << l;
#pragma unroll
        for (int k = 0; k < 8; k++) s_cv[8 * lane + k] = cv[k];
        __syncthreads();
        if (in && a % span == 0 && a + span <= end) {  // both halves in the run: a complete subtree
            uint32_t m[16];
#pragma unroll
            for (int k = 0; k < 8; k++) m[k] = cv[k], m[k + 8] = s_cv[8 * (lane + (1 << l)) + k];
            uint32_t st[16] = {key.k[0], key.k[1], key.k[2], key.k[3], key.k[4], key.k[5], key.k[6], key.k[7],
                               0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                               0u, 0u, 64u, key.flags | 4u};  // PARENT
            b3_rounds(st, m);
#pragma unroll
            for (int k = 0; k < 8; k++) cv[k] = st[k] ^ st[k + 8];
        }
        __syncthreads();
    }
    if (in)
#pragma unroll
        for (int k = 0; k < 8; k++) cvs[8 * c + k] = cv[k];
}
__global__ __launch_bounds__(64) void k_chunk_cvs(const uint8_t *bytes, size_t off, uint64_t first, size_t count,
                                                  B3Key key, uint32_t *cvs) {
    __shared__ uint32_t s_cv[64 * 8];
    chunk_cvs_body(blockIdx.x, bytes, off, first, count, key, cvs, s_cv);
}

// ---------------------------------------------------------------------------------------------
// eq[jh * 2^lo + jl] = EH[jh] * EL[jl] in GF(2^128): EH[jh] times the lo-coordinate expansion EL
// (lo <= 12), through a 16 x 256 byte-window table of y -> EH[jh] * y.
// ---------------------------------------------------------------------------------------------
// Block 0 also zeroes `clear` (the XOR accumulators the kernels after it fill), so that no
// separate fill is launched.
__device__ __forceinline__ void eq_expand_body(uint32_t jh, const uint4 *EH, const uint4 *EL, int lo, uint4 *E0,
                                               uint32_t *clear, int nclear, uint4 *basis, uint4 *tab) {
    const int t = threadIdx.x;
    if (jh == 0)
        for (int i = t; i < nclear; i += 256) clear[i] = 0;
    if (t < 128) {
        uint4 e = EH[jh];
        uint64_t lo = (uint64_t)e.x | ((uint64_t)e.y << 32), hi = (uint64_t)e.z | ((uint64_t)e.w << 32);
        for (int s = 0; s < t; s++) b128_mulx(lo, hi);
        basis[t] = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
    }
    __syncthreads();
    {
        const int w = t >> 4, bh = t & 15;
        uint4 base = make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++)
            if ((bh >> i) & 1) {
                uint4 b = basis[8 * w + 4 + i];
                base.x ^= b.x, base.y ^= b.y, base.z ^= b.z, base.w ^= b.w;
            }
        uint4 b0 = basis[8 * w], b1 = basis[8 * w + 1], b2 = basis[8 * w + 2], b3 = basis[8 * w + 3];
#pragma unroll
        for (int lo = 0; lo < 16; lo++) {
            uint4 v = base;
            if (lo & 1) v.x ^= b0.x, v.y ^= b0.y, v.z ^= b0.z, v.w ^= b0.w;
            if (lo & 2) v.x ^= b1.x, v.y ^= b1.y, v.z ^= b1.z, v.w ^= b1.w;
            if (lo & 4) v.x ^= b2.x, v.y ^= b2.y, v.z ^= b2.z, v.w ^= b2.w;
            if (lo & 8) v.x ^= b3.x, v.y ^= b3.y, v.z ^= b3.z, v.w ^= b3.w;
            tab[w * 256 + bh * 16 + lo] = v;
        }
    }
    __syncthreads();
    for (int jl = t; jl < (1 << lo); jl += 256) {
        uint4 y = EL[jl];
        uint32_t yw[4] = {y.x, y.y, y.z, y.w};
        uint4 acc = make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int w = 0; w < 16; w++) {
            uint4 v = tab[w * 256 + ((yw[w >> 2] >> (8 * (w & 3))) & 255)];
            acc.x ^= v.x, acc.y ^= v.y, acc.z ^= v.z, acc.w ^= v.w;
        }
        E0[((size_t)jh << lo) + jl] = acc;
    }
}
__global__ __launch_bounds__(256) void k_eq_expand(const uint4 *EH, const uint4 *EL, int lo, uint4 *E0,
                                                   uint32_t *clear, int nclear) {
    __shared__ uint4 basis[128];
    __shared__ uint4 tab[16 * 256];
    eq_expand_body(blockIdx.x, EH, EL, lo, E0, clear, nclear, basis, tab);
}

// ---------------------------------------------------------------------------------------------
// partials v[k] = sum_j bit_k(w_j) eq_hi[j] (crossfield.rs:152-165).
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_partials(const uint4 *table, const uint4 *E0, size_t n, uint32_t *out) {
    __shared__ __align__(16) uint32_t lds[4 * M4W<4>::WORDS];
    uint32_t *W = lds + (threadIdx.x >> 6) * M4W<4>::WORDS;
    uint32_t acc[2][4] = {};
    // two items per pass (n is a multiple of 256): both items' loads are in flight before the first
    // step, so the second step's operands arrive during the first
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t base = (size_t)blockIdx.x * 256; base < n; base += 2 * stride) {
        const size_t j = base + threadIdx.x;
        const bool two = base + stride < n;
        const uint4 w = table[j], e = E0[j];
        uint4 w2 = make_uint4(0, 0, 0, 0), e2 = make_uint4(0, 0, 0, 0);
        if (two) w2 = table[j + stride], e2 = E0[j + stride];
        {
            uint32_t sel[4] = {w.x, w.y, w.z, w.w};
            uint32_t val[4] = {e.x, e.y, e.z, e.w};
            m4w_step<4>(acc, sel, val, W);
        }
        if (two) {
            uint32_t sel[4] = {w2.x, w2.y, w2.z, w2.w};
            uint32_t val[4] = {e2.x, e2.y, e2.z, e2.w};
            m4w_step<4>(acc, sel, val, W);
        }
    }
    m4w_flush<4, true>(acc, lds, out);
}

// ---------------------------------------------------------------------------------------------
// The same partials without the expansion in memory (the allocation-free entry point):
// eq_hi[jh * 4096 + jl] = EH[jh] EL[jl], so v[k] = sum_jh EH[jh] R_jh[k] with
// R_jh[k] = sum_jl bit_k(w_{jh*4096+jl}) EL[jl]. One wave per row jh (waves stride over the rows):
// the wave-local M4RM sums R_jh over the row's 4096 words (EL read from L2), then 128 products by
// EH[jh] through a 2-bit window table of y -> EH[jh] y (64 windows x 4, 4 KiB: the M4RM region of
// the wave, free after the row's steps); each lane keeps the sums of bits l and l + 64; the block
// adds its waves' sums and atomicXors them into out (zeroed before).
// ---------------------------------------------------------------------------------------------
// y x^s in GF(2^128) (x^128 = x^7 + x^2 + x + 1), 0 <= s < 128.
__device__ __forceinline__ uint4 b128_mul_xs(uint4 y, int s) {
    typedef unsigned __int128 v128;
    const v128 v = ((v128)(((uint64_t)y.w << 32) | y.z) << 64) | (((uint64_t)y.y << 32) | y.x);
    v128 r = v;
    if (s) {
        const v128 h = v >> (128 - s);  // the s bits shifted out
        const v128 o = (h >> 127) ^ (h >> 126) ^ (h >> 121);  // the fold's own overflow
        r = (v << s) ^ h ^ (h << 1) ^ (h << 2) ^ (h << 7) ^ o ^ (o << 1) ^ (o << 2) ^ (o << 7);
    }
    const uint64_t lo = (uint64_t)r, hi = (uint64_t)(r >> 64);
    return make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}

constexpr int PF_WAVES = 4;
__global__ __launch_bounds__(64 * PF_WAVES) void k_partials_fact(const uint4 *table, const uint4 *EL, const uint4 *EH,
                                                                 int rows, uint32_t *out) {
    __shared__ __align__(16) uint32_t ws[PF_WAVES][M4W<4>::WORDS];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t *S = ws[wv];
    uint4 *win = (uint4 *)S;  // after a row's steps: win[w][v] = EH[jh] (v x^(2w)), 64 windows x 4
    uint4 prod[2] = {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};  // bits lane, lane + 64
    for (int jh = blockIdx.x * PF_WAVES + wv; jh < rows; jh += gridDim.x * PF_WAVES) {
        uint32_t acc[2][4] = {};
        const uint4 *tw = table + (size_t)jh * 4096;
        for (int base = 0; base < 4096; base += 128) {  // two items per pass, both loaded first
            const uint4 w0 = tw[base + lane], e0 = EL[base + lane];
            const uint4 w1 = tw[base + 64 + lane], e1 = EL[base + 64 + lane];
            {
                uint32_t sel[4] = {w0.x, w0.y, w0.z, w0.w};
                uint32_t val[4] = {e0.x, e0.y, e0.z, e0.w};
                m4w_step<4>(acc, sel, val, S);
            }
            {
                uint32_t sel[4] = {w1.x, w1.y, w1.z, w1.w};
                uint32_t val[4] = {e1.x, e1.y, e1.z, e1.w};
                m4w_step<4>(acc, sel, val, S);
            }
        }
        {  // lane l writes window l: 0, EH x^(2l), EH x^(2l+1) and their sum
            const uint4 e = EH[jh];
            const uint4 b0 = b128_mul_xs(e, 2 * lane), b1 = b128_mul_xs(e, 2 * lane + 1);
            win[4 * lane] = make_uint4(0, 0, 0, 0);
            win[4 * lane + 1] = b0;
            win[4 * lane + 2] = b1;
            win[4 * lane + 3] = make_uint4(b0.x ^ b1.x, b0.y ^ b1.y, b0.z ^ b1.z, b0.w ^ b1.w);
        }
        __syncwarp();
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int w = 0; w < 64; w++) {
                const uint4 y = win[4 * w + ((acc[c][w >> 4] >> (2 * (w & 15))) & 3)];
                prod[c].x ^= y.x, prod[c].y ^= y.y, prod[c].z ^= y.z, prod[c].w ^= y.w;
            }
        __syncwarp();  // the next row's steps rewrite the wave's region
    }
    // the block's sum: waves 1.. leave theirs in LDS, wave 0 adds them and the atomics follow
    __syncthreads();
    uint4 *red = (uint4 *)ws[wv];
    red[lane] = prod[0];
    red[lane + 64] = prod[1];
    __syncthreads();
    if (wv == 0) {
#pragma unroll
        for (int c = 0; c < 2; c++) {
            uint4 x = make_uint4(0, 0, 0, 0);
#pragma unroll
            for (int o = 0; o < PF_WAVES; o++) {
                const uint4 y = ((const uint4 *)ws[o])[lane + 64 * c];
                x.x ^= y.x, x.y ^= y.y, x.z ^= y.z, x.w ^= y.w;
            }
            uint32_t *dst = out + 4 * (lane + 64 * c);
            if (x.x) atomicXor(dst, x.x);
            if (x.y) atomicXor(dst + 1, x.y);
            if (x.z) atomicXor(dst + 2, x.z);
            if (x.w) atomicXor(dst + 3, x.w);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Prefix sets of the table: with j = u 2^S + m (u the top 6 index bits, S = log_packed - 6) the eq
// value of j factors as eq6(u) G_m, G the expansion of r_hi[6..] (crossfield.rs:87-97: the first
// coordinate is the most significant bit), so the partials (crossfield.rs:152-165) are
//   v[b] = sum_u eq6(u) M_u[b],   M_u[b] = sum_m bit_b(w_{u 2^S + m}) G_m,
// and the 64 sets M_u also give the switch's first six messages (k_prefix_msg). A block sums a
// chunk of 2^chunk_log words inside one prefix by wave-local M4RM, then XORs its sums into M_u and
// their products by eq6(u) (a 2-bit window table of y -> eq6(u) y) into v; both zeroed before.
// ---------------------------------------------------------------------------------------------
struct PrefixLds {
    __align__(16) uint32_t lds[4 * M4W<4>::WORDS];
    uint32_t bs[512];
};
__device__ __forceinline__ void prefix_sums_body(size_t bid, PrefixLds &L, const uint4 *table, const uint4 *G, int S,
                                                 int chunk_log, const uint4 *eq6, uint32_t *M, uint32_t *partials) {
    uint32_t *lds = L.lds, *bs = L.bs;
    uint32_t *W = lds + (threadIdx.x >> 6) * M4W<4>::WORDS;
    const size_t chunk = (size_t)1 << chunk_log, first = bid << chunk_log;
    const int u = (int)(first >> S);
    const uint4 *tw = table + first, *gw = G + (first & (((size_t)1 << S) - 1));
    uint32_t acc[2][4] = {};
    for (size_t base = 0; base < chunk; base += 512) {  // two items per pass, both loaded first
        const size_t i0 = base + threadIdx.x, i1 = i0 + 256;
        const uint4 z = make_uint4(0, 0, 0, 0);
        const uint4 w0 = i0 < chunk ? tw[i0] : z, g0 = i0 < chunk ? gw[i0] : z;
        const uint4 w1 = i1 < chunk ? tw[i1] : z, g1 = i1 < chunk ? gw[i1] : z;
        {
            uint32_t sel[4] = {w0.x, w0.y, w0.z, w0.w};
            uint32_t val[4] = {g0.x, g0.y, g0.z, g0.w};
            m4w_step<4>(acc, sel, val, W);
        }
        if (base + 256 < chunk) {
            uint32_t sel[4] = {w1.x, w1.y, w1.z, w1.w};
            uint32_t val[4] = {g1.x, g1.y, g1.z, g1.w};
            m4w_step<4>(acc, sel, val, W);
        }
    }
    m4w_flush<4, false>(acc, lds, bs);  // bs[4 b + v]: word v of the block's sum for bit b
    __syncthreads();
    const int t = threadIdx.x;
    for (int i = t; i < 512; i += 256)
        if (bs[i]) atomicXor(M + (size_t)u * 512 + i, bs[i]);
    uint4 *win = (uint4 *)lds;  // win[4 w + v] = eq6(u) (v x^(2w)), 64 windows of 2 bits
    if (t < 64) {
        const uint4 e = eq6[u];
        const uint4 b0 = b128_mul_xs(e, 2 * t), b1 = b128_mul_xs(e, 2 * t + 1);
        win[4 * t] = make_uint4(0, 0, 0, 0);
        win[4 * t + 1] = b0;
        win[4 * t + 2] = b1;
        win[4 * t + 3] = make_uint4(b0.x ^ b1.x, b0.y ^ b1.y, b0.z ^ b1.z, b0.w ^ b1.w);
    }
    __syncthreads();
    if (t < 128) {
        uint4 x = make_uint4(0, 0, 0, 0);
#pragma unroll 1
        for (int q = 0; q < 4; q++) {  // 16 windows per word of the sum
            const uint32_t y = bs[4 * t + q];
#pragma unroll
            for (int n = 0; n < 16; n++) {
                const uint4 e = win[4 * (16 * q + n) + ((y >> (2 * n)) & 3)];
                x.x ^= e.x, x.y ^= e.y, x.z ^= e.z, x.w ^= e.w;
            }
        }
        uint32_t *dst = partials + 4 * t;
        if (x.x) atomicXor(dst, x.x);
        if (x.y) atomicXor(dst + 1, x.y);
        if (x.z) atomicXor(dst + 2, x.z);
        if (x.w) atomicXor(dst + 3, x.w);
    }
}
__global__ __launch_bounds__(256) void k_prefix_sums(const uint4 *table, const uint4 *G, int S, int chunk_log,
                                                     const uint4 *eq6, uint32_t *M, uint32_t *partials) {
    __shared__ PrefixLds L;
    prefix_sums_body(blockIdx.x, L, table, G, S, chunk_log, eq6, M, partials);
}

// The view whose block range holds block b: the last i with starts(v[i]) <= b.
template <class V, class F>
__device__ __forceinline__ int view_of_block(const V *v, int n, int b, F starts) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts(v[mid]) <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// The partials of several proofs: the eq expansions (proof i's blocks from eq_start), the prefix
// sums (from ps_start), then each proof's 512 partial words copied to its mailbox.
struct PartView {
    const uint4 *EH, *EL;
    int lo;
    uint4 *E0;
    uint32_t *clear;
    int nclear, eq_start;
    const uint4 *table;
    int S, chunk_log, ps_start;
    const uint4 *eq6;
    uint32_t *M, *part, *part_out;
};
__global__ __launch_bounds__(256) void k_meq_expand(const PartView *v, int n) {
    __shared__ uint4 basis[128];
    __shared__ uint4 tab[16 * 256];
    const PartView &x = v[view_of_block(v, n, blockIdx.x, [](const PartView &y) { return y.eq_start; })];
    eq_expand_body(blockIdx.x - x.eq_start, x.EH, x.EL, x.lo, x.E0, x.clear, x.nclear, basis, tab);
}
__global__ __launch_bounds__(256) void k_mprefix_sums(const PartView *v, int n) {
    __shared__ PrefixLds L;
    const PartView &x = v[view_of_block(v, n, blockIdx.x, [](const PartView &y) { return y.ps_start; })];
    prefix_sums_body(blockIdx.x - x.ps_start, L, x.table, x.E0, x.S, x.chunk_log, x.eq6, x.M, x.part);
}
// Inputs staged in the proofs' pinned mailboxes, copied to their device buffers in one launch for
// all proofs (`per` blocks each) instead of one small copy per proof.
struct Gather {
    const uint4 *src;
    uint4 *dst;
    size_t n16;
};
__global__ __launch_bounds__(256) void k_mgather(const Gather *g, int per) {
    const Gather x = g[blockIdx.x / per];
    for (size_t i = (size_t)(blockIdx.x % per) * 256 + threadIdx.x; i < x.n16; i += (size_t)per * 256) x.dst[i] = x.src[i];
}
__global__ __launch_bounds__(256) void k_mpart_out(const PartView *v) {
    const PartView &x = v[blockIdx.x];
    for (int i = threadIdx.x; i < 512; i += 256) x.part_out[i] = x.part[i];
}

// ---------------------------------------------------------------------------------------------
// Switch round 0 (sumcheck.rs:89-117 with a_j = L(E0_j), p_j = lift(w_j)):
//   T0_k   = sum_{j<h} bit_k(E0_j) w_j
//   Tinf_k = sum_{j<h} bit_k(E1_j) (w_j + w_{j+h}),  E1_j = E0_j + E0_{j+h} (written)
// The host turns T into e0 = sum_k L[k] T0_k, einf = sum_k L[k] Tinf_k.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_switch_msg0(const uint4 *table, const uint4 *E0, size_t h, uint4 *E1,
                                                     uint32_t *T0, uint32_t *Tinf) {
    __shared__ __align__(16) uint32_t lds[4 * M4W<4>::WORDS];
    uint32_t *W = lds + (threadIdx.x >> 6) * M4W<4>::WORDS;
    uint32_t acc0[2][4] = {}, acci[2][4] = {};
    // two items per pass (h is a multiple of 256), both items' loads issued first
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t base = (size_t)blockIdx.x * 256; base < h; base += 2 * stride) {
        const bool two = base + stride < h;
        uint4 ea[2], eb[2], wa[2], wb[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const size_t j = base + u * stride + threadIdx.x;
            if (u == 0 || two) {
                ea[u] = E0[j], eb[u] = E0[j + h], wa[u] = table[j], wb[u] = table[j + h];
            } else {
                ea[u] = eb[u] = wa[u] = wb[u] = make_uint4(0, 0, 0, 0);
            }
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            if (u == 1 && !two) break;
            const size_t j = base + u * stride + threadIdx.x;
            const uint4 e1 = make_uint4(ea[u].x ^ eb[u].x, ea[u].y ^ eb[u].y, ea[u].z ^ eb[u].z, ea[u].w ^ eb[u].w);
            E1[j] = e1;
            {
                uint32_t sel[4] = {ea[u].x, ea[u].y, ea[u].z, ea[u].w};
                uint32_t val[4] = {wa[u].x, wa[u].y, wa[u].z, wa[u].w};
                m4w_step<4>(acc0, sel, val, W);
            }
            {
                uint32_t sel[4] = {e1.x, e1.y, e1.z, e1.w};
                uint32_t val[4] = {wa[u].x ^ wb[u].x, wa[u].y ^ wb[u].y, wa[u].z ^ wb[u].z, wa[u].w ^ wb[u].w};
                m4w_step<4>(acci, sel, val, W);
            }
        }
    }
    m4w_flush<4, true>(acc0, lds, T0);
    __syncthreads();
    m4w_flush<4, true>(acci, lds, Tinf);
}

// Per-block nibble-window table of y -> rho * y in F162: 41 windows x 16 entries x 6 words.
struct RhoArg {
    uint64_t r[3];
};
constexpr int RHO_WINDOWS = 41;

// rho * x^s for s < 128: the shifted product (below 2^290) reduced modulo x^162 + x^81 + 1.
__device__ __forceinline__ void f162_mulx_pow(const uint64_t (&a)[3], int s, uint64_t (&out)[3]) {
    const int ws = s >> 6, bs = s & 63;
    uint64_t sh[4];  // a << bs over four limbs
    sh[0] = a[0] << bs;
    sh[1] = (a[1] << bs) | (bs ? a[0] >> (64 - bs) : 0);
    sh[2] = (a[2] << bs) | (bs ? a[1] >> (64 - bs) : 0);
    sh[3] = bs ? a[2] >> (64 - bs) : 0;
    uint64_t v[5];
    v[0] = ws ? 0 : sh[0];
    v[1] = ws ? sh[0] : sh[1];
    v[2] = ws ? sh[1] : sh[2];
    v[3] = ws ? sh[2] : sh[3];
    v[4] = ws ? sh[3] : 0;
    f162_reduce5(v, out);
}

__device__ void build_rho_table(RhoArg rho, uint32_t *basis /*162*6*/, uint32_t *tab /*41*16*6*/) {
    const int t = threadIdx.x;
    if (t < 162) {
        // basis[t] = rho * x^t, through rho * x^128 for t >= 128
        const uint64_t r[3] = {rho.r[0], rho.r[1], rho.r[2]};
        uint64_t r128[3], x[3];
        f162_mulx_pow(r, 64, x);
        f162_mulx_pow(x, 64, r128);
        f162_mulx_pow(t < 128 ? r : r128, t & 127, x);
        const uint64_t a0 = x[0], a1 = x[1], a2 = x[2];
        uint32_t *b = basis + 6 * t;
        b[0] = (uint32_t)a0, b[1] = (uint32_t)(a0 >> 32), b[2] = (uint32_t)a1, b[3] = (uint32_t)(a1 >> 32);
        b[4] = (uint32_t)a2, b[5] = (uint32_t)(a2 >> 32);
    }
    __syncthreads();
    for (int e = t; e < RHO_WINDOWS * 16; e += blockDim.x) {
        const int w = e >> 4, nib = e & 15;
        uint32_t v[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int i = 4 * w + b;
            if (((nib >> b) & 1) && i < 162) {
#pragma unroll
                for (int k = 0; k < 6; k++) v[k] ^= basis[6 * i + k];
            }
        }
#pragma unroll
        for (int k = 0; k < 6; k++) tab[6 * e + k] = v[k];
    }
    __syncthreads();
}

// rho_mul with eight windows (one input word) per loop step, for kernels that run many products
// back to back: the table reads of one word are in flight at a time.
template <int NW>
__device__ __forceinline__ void rho_mul_lean(const uint32_t *tab, const uint32_t (&y)[6], uint32_t (&out)[6]) {
#pragma unroll
    for (int k = 0; k < 6; k++) out[k] = 0;
#pragma unroll 1
    for (int q = 0; q < (NW + 7) / 8; q++) {
        const uint32_t yq = q == 0 ? y[0] : q == 1 ? y[1] : q == 2 ? y[2] : q == 3 ? y[3] : q == 4 ? y[4] : y[5];
        const uint32_t *tq = tab + 6 * 16 * 8 * q;
#pragma unroll
        for (int n = 0; n < 8; n++) {
            if (8 * q + n < NW) {
                const uint32_t *e = tq + 6 * (n * 16 + ((yq >> (4 * n)) & 15));
#pragma unroll
                for (int k = 0; k < 6; k++) out[k] ^= e[k];
            }
        }
    }
}

// Two products by the same table, their lookups interleaved (two independent chains in flight).
template <int NW>
__device__ __forceinline__ void rho_mul_lean2(const uint32_t *tab, const uint32_t (&y0)[6], const uint32_t (&y1)[6],
                                              uint32_t (&o0)[6], uint32_t (&o1)[6]) {
#pragma unroll
    for (int k = 0; k < 6; k++) o0[k] = 0, o1[k] = 0;
#pragma unroll 1
    for (int q = 0; q < (NW + 7) / 8; q++) {
        const uint32_t a = q == 0 ? y0[0] : q == 1 ? y0[1] : q == 2 ? y0[2] : q == 3 ? y0[3] : q == 4 ? y0[4] : y0[5];
        const uint32_t b = q == 0 ? y1[0] : q == 1 ? y1[1] : q == 2 ? y1[2] : q == 3 ? y1[3] : q == 4 ? y1[4] : y1[5];
        const uint32_t *tq = tab + 6 * 16 * 8 * q;
#pragma unroll
        for (int n = 0; n < 8; n++) {
            if (8 * q + n < NW) {
                const uint32_t *e0 = tq + 6 * (n * 16 + ((a >> (4 * n)) & 15));
                const uint32_t *e1 = tq + 6 * (n * 16 + ((b >> (4 * n)) & 15));
#pragma unroll
                for (int k = 0; k < 6; k++) o0[k] ^= e0[k], o1[k] ^= e1[k];
            }
        }
    }
}

template <int NW>  // number of nibble windows of the input (32 for a lifted word, 41 for F162)
__device__ __forceinline__ void rho_mul(const uint32_t *tab, const uint32_t (&y)[6], uint32_t (&out)[6]) {
#pragma unroll
    for (int k = 0; k < 6; k++) out[k] = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        const uint32_t nib = (y[w >> 3] >> (4 * (w & 7))) & 15;
        const uint32_t *e = tab + 6 * (w * 16 + nib);
#pragma unroll
        for (int k = 0; k < 6; k++) out[k] ^= e[k];
    }
}

// ---------------------------------------------------------------------------------------------
// Fused switch round: fold round r with rho (sumcheck.rs:119-139) and the message of round r+1.
//   p'_j = p_j + rho (p_j + p_{j+h}) for j < h, with h = 2 h2 the current half;
//   T0_k   += bit_k(E_j) p'_j                 (E = E^(r+1), j < h2)
//   Tinf_k += bit_k(E_j + E_{j+h2}) (p'_j + p'_{j+h2}),   E^(r+2)_j = E_j + E_{j+h2} (written)
// WORDS: p is the table (lifted words) in round 0, F162 (6 words) afterwards.
// ---------------------------------------------------------------------------------------------
// Block 0 also clears `clear` (1536 words), the message buffer of the round after next.
// rho_dev, when not null, holds the challenge (written by the device transcript) instead of rho.
// rtab: RHO_WINDOWS * 16 * 6 words, lds: 4 * M4W<6>::WORDS words (16-byte aligned), both in LDS
template <bool WORDS>
__device__ __forceinline__ void switch_round_body(int bid, int nb, const uint32_t *__restrict__ pin, size_t h2,
                                                  RhoArg rho, const uint4 *__restrict__ Ein, uint4 *__restrict__ Eout,
                                                  uint32_t *__restrict__ pout, uint32_t *T0, uint32_t *Tinf,
                                                  uint32_t *clear, const uint64_t *rho_dev, uint32_t *rtab,
                                                  uint32_t *lds) {
    if (bid == 0)
        for (int i = threadIdx.x; i < 1536; i += blockDim.x) clear[i] = 0;
    const size_t h = 2 * h2;
    const size_t nthreads = (h2 + 255) / 256 * 256, stride = (size_t)nb * 256;
    // one item's operands: p_j, p_{j+h2}, p_{j+h}, p_{j+h+h2} and E_j, E_{j+h2} (zero past h2)
    uint32_t a[4][6];
    uint4 ea, eb;
    auto load_item = [&](size_t j) {
        const bool live = j < h2;
        const size_t jj = live ? j : 0;
        const size_t idx[4] = {jj, jj + h2, jj + h, jj + h + h2};
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (WORDS) {
                const uint4 w = ((const uint4 *)pin)[idx[u]];
                a[u][0] = w.x, a[u][1] = w.y, a[u][2] = w.z, a[u][3] = w.w, a[u][4] = 0, a[u][5] = 0;
            } else {
                const uint2 *src = (const uint2 *)(pin + 6 * idx[u]);
                const uint2 x0 = src[0], x1 = src[1], x2 = src[2];
                a[u][0] = x0.x, a[u][1] = x0.y, a[u][2] = x1.x, a[u][3] = x1.y, a[u][4] = x2.x, a[u][5] = x2.y;
            }
        }
        ea = Ein[jj], eb = Ein[jj + h2];
    };
    // the first item's loads are in flight while the challenge's tables are built
    size_t base = (size_t)bid * 256;
    if (base < nthreads) load_item(base + threadIdx.x);
    if (rho_dev) rho = RhoArg{{rho_dev[0], rho_dev[1], rho_dev[2]}};
    build_rho_table(rho, lds, rtab);
    uint32_t *W = lds + (threadIdx.x >> 6) * M4W<6>::WORDS;
    uint32_t acc0[2][6] = {}, acci[2][6] = {};
    for (; base < nthreads; base += stride) {
        const size_t j = base + threadIdx.x;
        const bool live = j < h2;
        uint32_t q0[6] = {0, 0, 0, 0, 0, 0}, q1[6] = {0, 0, 0, 0, 0, 0};
        uint32_t e[4] = {0, 0, 0, 0}, e2[4] = {0, 0, 0, 0};
        if (live) {
            uint32_t d0[6], d1[6], m0[6], m1[6];
#pragma unroll
            for (int k = 0; k < 6; k++) d0[k] = a[0][k] ^ a[2][k], d1[k] = a[1][k] ^ a[3][k];
            if (WORDS) {
                rho_mul<32>(rtab, d0, m0);
                rho_mul<32>(rtab, d1, m1);
            } else {
                rho_mul<RHO_WINDOWS>(rtab, d0, m0);
                rho_mul<RHO_WINDOWS>(rtab, d1, m1);
            }
#pragma unroll
            for (int k = 0; k < 6; k++) q0[k] = a[0][k] ^ m0[k], q1[k] = a[1][k] ^ m1[k];
            uint2 *o0 = (uint2 *)(pout + 6 * j), *o1 = (uint2 *)(pout + 6 * (j + h2));
            o0[0] = make_uint2(q0[0], q0[1]), o0[1] = make_uint2(q0[2], q0[3]), o0[2] = make_uint2(q0[4], q0[5]);
            o1[0] = make_uint2(q1[0], q1[1]), o1[1] = make_uint2(q1[2], q1[3]), o1[2] = make_uint2(q1[4], q1[5]);
            e[0] = ea.x, e[1] = ea.y, e[2] = ea.z, e[3] = ea.w;
            e2[0] = ea.x ^ eb.x, e2[1] = ea.y ^ eb.y, e2[2] = ea.z ^ eb.z, e2[3] = ea.w ^ eb.w;
            Eout[j] = make_uint4(e2[0], e2[1], e2[2], e2[3]);
        }
        if (base + stride < nthreads) load_item(base + stride + threadIdx.x);  // the next item's operands
        m4w_step<6>(acc0, e, q0, W);
        uint32_t dq[6];
#pragma unroll
        for (int k = 0; k < 6; k++) dq[k] = q0[k] ^ q1[k];
        m4w_step<6>(acci, e2, dq, W);
    }
    m4w_flush<6, true>(acc0, lds, T0);
    __syncthreads();
    m4w_flush<6, true>(acci, lds, Tinf);
}
template <bool WORDS>
__global__ __launch_bounds__(256) void k_switch_round(const uint32_t *__restrict__ pin, size_t h2, RhoArg rho,
                                                      const uint4 *__restrict__ Ein, uint4 *__restrict__ Eout,
                                                      uint32_t *__restrict__ pout, uint32_t *T0, uint32_t *Tinf,
                                                      uint32_t *clear, const uint64_t *rho_dev) {
    __shared__ uint32_t rtab[RHO_WINDOWS * 16 * 6];
    __shared__ __align__(16) uint32_t lds[4 * M4W<6>::WORDS];  // also holds the basis while rtab is built
    switch_round_body<WORDS>(blockIdx.x, gridDim.x, pin, h2, rho, Ein, Eout, pout, T0, Tinf, clear, rho_dev, rtab,
                             lds);
}

// ---------------------------------------------------------------------------------------------
// The switch rounds on the device (sumcheck.rs:89-139 driven as main.rs:975-992): after each
// round kernel, one block forms the round's message e0 = sum_k L[k] T0_k, einf = sum_k L[k]
// Tinf_k, absorbs it into the Fiat-Shamir transcript, samples the challenge rho
// (challenge.rs:226-326: absorb_f162 then sample_f162("switch-round")), and updates the linear
// map L_{r+1}[k] = L_r((1 + c) x^k) + rho L_r[k] (c = r_hi[r]); the next round kernel reads rho
// from device memory. The transcript is BLAKE3 over the absorbed stream, carried here in the
// same plain state the host hasher saves, so the host continues from where the device stops.
// ---------------------------------------------------------------------------------------------
static_assert(sizeof(B3State) % 4 == 0, "the transcript state is copied as words");
struct SwitchDev {
    B3State ts;
    uint64_t lmap[128][3];
    uint64_t r_hi[32][2];
    uint64_t rho[3];
    uint64_t rounds[32][6];
    uint64_t coords[32][3];
    uint64_t opened[3];
};

// BLAKE3 over a B3State in LDS, run by the four lanes of one quad (threads 0..3): the compression
// puts state column q in lane q (rows a = s[q], b = s[4+q], c = s[8+q], d = s[12+q]) and turns
// the diagonal step into quad rotations (DPP), so a compression is 7 x 2 G steps per lane. The
// bookkeeping (buffer, chunk, chaining-value stack) is the host Blake3 class statement by
// statement, done identically by the four lanes; every absorb here is a whole number of 32-bit
// words at a word-aligned buffer position, so the buffer is handled as words.
namespace b3d {
constexpr uint32_t CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8;

// x_q by bit selects on lane masks (no compare: its condition register would stall the DPP chain).
__device__ __forceinline__ uint32_t sel4(int q, uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3) {
    const uint32_t m1 = 0u - (uint32_t)(q & 1), m2 = 0u - (uint32_t)((q >> 1) & 1);
    const uint32_t lo = (x0 & ~m1) | (x1 & m1), hi = (x2 & ~m1) | (x3 & m1);
    return (lo & ~m2) | (hi & m2);
}
template <int J>
__device__ __forceinline__ uint32_t bcast(uint32_t x) {  // lane J of the quad
    return dpp<J * 0x55>(x);
}
__device__ __forceinline__ void g(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d, uint32_t x, uint32_t y) {
    b3_g(a, b, c, d, x, y);
}

// out (16 words, the same in the four lanes) = compress(cv, m, counter, len, flags).
__device__ __forceinline__ void compress(int q, const uint32_t (&cv)[8], const uint32_t (&m_in)[16], uint64_t counter, uint32_t len,
                         uint32_t flags, uint32_t (&out)[16]) {
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 16; i++) m[i] = m_in[i];
    uint32_t a = sel4(q, cv[0], cv[1], cv[2], cv[3]), b = sel4(q, cv[4], cv[5], cv[6], cv[7]);
    uint32_t c = sel4(q, 0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au);
    uint32_t d = sel4(q, (uint32_t)counter, (uint32_t)(counter >> 32), len, flags);
    const int perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
#pragma unroll
    for (int r = 0; r < 7; r++) {
        g(a, b, c, d, sel4(q, m[0], m[2], m[4], m[6]), sel4(q, m[1], m[3], m[5], m[7]));
        b = dpp<0x39>(b), c = dpp<0x4E>(c), d = dpp<0x93>(d);  // from lanes q+1, q+2, q+3
        g(a, b, c, d, sel4(q, m[8], m[10], m[12], m[14]), sel4(q, m[9], m[11], m[13], m[15]));
        b = dpp<0x93>(b), c = dpp<0x4E>(c), d = dpp<0x39>(d);  // back
        uint32_t t[16];
#pragma unroll
        for (int k = 0; k < 16; k++) t[k] = m[perm[k]];
#pragma unroll
        for (int k = 0; k < 16; k++) m[k] = t[k];
    }
    const uint32_t o0 = a ^ c, o1 = b ^ d, o2 = c ^ sel4(q, cv[0], cv[1], cv[2], cv[3]),
                   o3 = d ^ sel4(q, cv[4], cv[5], cv[6], cv[7]);
    out[0] = bcast<0>(o0), out[1] = bcast<1>(o0), out[2] = bcast<2>(o0), out[3] = bcast<3>(o0);
    out[4] = bcast<0>(o1), out[5] = bcast<1>(o1), out[6] = bcast<2>(o1), out[7] = bcast<3>(o1);
    out[8] = bcast<0>(o2), out[9] = bcast<1>(o2), out[10] = bcast<2>(o2), out[11] = bcast<3>(o2);
    out[12] = bcast<0>(o3), out[13] = bcast<1>(o3), out[14] = bcast<2>(o3), out[15] = bcast<3>(o3);
}
__device__ __forceinline__ uint32_t *bufw(B3State &s) { return (uint32_t *)s.buf; }
__device__ __forceinline__ void words16(B3State &s, uint32_t (&m)[16]) {  // the buffer, zero past buf_len
    const uint32_t n = s.buf_len;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t w = bufw(s)[i];
        m[i] = 4 * (uint32_t)i < n ? w : 0u;
    }
}
__device__ __forceinline__ void chunk_cv(int q, B3State &s, uint32_t (&out)[8]) {
    uint32_t m[16], cv[8], o[16];
    words16(s, m);
#pragma unroll
    for (int i = 0; i < 8; i++) cv[i] = s.cv[i];
    compress(q, cv, m, s.chunk, s.buf_len, s.flags | CHUNK_END | (s.blocks == 0 ? CHUNK_START : 0), o);
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = o[i];
}
__device__ __forceinline__ void add_chunk_cv(int q, B3State &s, uint32_t (&cv)[8], uint64_t total) {
    uint32_t key[8];
#pragma unroll
    for (int i = 0; i < 8; i++) key[i] = s.key[i];
    while ((total & 1) == 0) {
        uint32_t m[16], o[16];
        const int dpt = s.depth - 1;
#pragma unroll
        for (int i = 0; i < 8; i++) m[i] = s.stack[dpt][i], m[i + 8] = cv[i];
        compress(q, key, m, 0, 64, s.flags | PARENT, o);
#pragma unroll
        for (int i = 0; i < 8; i++) cv[i] = o[i];
        s.depth = dpt;
        total >>= 1;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) s.stack[s.depth][i] = cv[i];
    s.depth = s.depth + 1;
}
// Absorb one word (buf_len a multiple of 4).
__device__ __forceinline__ void update_word(int q, B3State &s, uint32_t w) {
    if ((uint32_t)s.blocks * 64 + s.buf_len == 1024) {  // full chunk and more input: close it
        uint32_t cv[8];
        chunk_cv(q, s, cv);
        add_chunk_cv(q, s, cv, s.chunk + 1);
        s.chunk = s.chunk + 1;
        s.blocks = 0;
        s.buf_len = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) s.cv[k] = s.key[k];
    }
    if (s.buf_len == 64) {
        uint32_t m[16], cv[8], o[16];
        words16(s, m);
#pragma unroll
        for (int k = 0; k < 8; k++) cv[k] = s.cv[k];
        compress(q, cv, m, s.chunk, 64, s.flags | (s.blocks == 0 ? CHUNK_START : 0), o);
#pragma unroll
        for (int k = 0; k < 8; k++) s.cv[k] = o[k];
        s.blocks = s.blocks + 1;
        s.buf_len = 0;
    }
    bufw(s)[s.buf_len / 4] = w;
    s.buf_len = s.buf_len + 4;
}
template <int N>
__device__ __forceinline__ void update_u64s(int q, B3State &s, const uint64_t (&x)[N]) {
#pragma unroll
    for (int i = 0; i < N; i++) {
        update_word(q, s, (uint32_t)x[i]);
        update_word(q, s, (uint32_t)(x[i] >> 32));
    }
}
// First output block of the root node.
__device__ __forceinline__ void root_block0(int q, B3State &s, uint32_t (&out)[16]) {
    uint32_t cv[8], m[16];
    uint32_t len, flags;
    if (s.depth == 0) {
#pragma unroll
        for (int i = 0; i < 8; i++) cv[i] = s.cv[i];
        words16(s, m);
        len = s.buf_len;
        flags = s.flags | CHUNK_END | (s.blocks == 0 ? CHUNK_START : 0);
    } else {
        uint32_t right[8], key[8];
        chunk_cv(q, s, right);
#pragma unroll
        for (int k = 0; k < 8; k++) key[k] = s.key[k];
        for (int i = s.depth - 1; i > 0; i--) {
            uint32_t mm[16], o[16];
#pragma unroll
            for (int k = 0; k < 8; k++) mm[k] = s.stack[i][k], mm[k + 8] = right[k];
            compress(q, key, mm, 0, 64, s.flags | PARENT, o);
#pragma unroll
            for (int k = 0; k < 8; k++) right[k] = o[k];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) cv[k] = key[k], m[k] = s.stack[0][k], m[k + 8] = right[k];
        len = 64;
        flags = s.flags | PARENT;
    }
    compress(q, cv, m, 0, len, flags | ROOT, out);
}
}  // namespace b3d

__device__ __forceinline__ void f162_mulx4(uint64_t (&a)[3]) {
    const uint64_t top = a[2] >> 30;  // bits 158..161
    a[2] = ((a[2] << 4) | (a[1] >> 60)) & ((1ull << 34) - 1);
    a[1] = (a[1] << 4) | (a[0] >> 60);
    a[0] <<= 4;
    a[0] ^= top;        // x^162 = x^81 + 1
    a[1] ^= top << 17;  // bit 81 is limb 1 bit 17
}

// v * x^k in GF(2^128), GHASH basis (x^128 = x^7 + x^2 + x + 1), k < 128.
__device__ __forceinline__ u128 b128_mulx_pow(u128 v, int k) {
    if (k == 0) return v;
    u128 lo = v << k;
    const u128 hi = v >> (128 - k);
    const u128 t = hi ^ (hi << 1) ^ (hi << 2) ^ (hi << 7);
    const u128 o = (hi >> 127) ^ (hi >> 126) ^ (hi >> 121);
    return lo ^ t ^ o ^ (o << 1) ^ (o << 2) ^ (o << 7);
}

// The control step's LDS: the transcript, the map, the challenge and its window tables, and the
// step's scratch tables.
struct CtlKeep {
    B3State st, cl;
    uint64_t L[128][3];
    uint64_t tabR41[41][16][3];  // y -> rho y over 4-bit windows
    uint64_t prod[2][3];
    uint64_t rho_s[3];
};
struct CtlTemp {
    uint64_t part[32][3];
    uint64_t tabL[32][16][3];
    uint64_t Tt[2][192][2];
    uint64_t rbas[162][3];
    uint64_t lpart[8][128][3];
};

// One control step for round r (1024 threads), on the transcript K.st and map K.L in
// LDS; the message comes from T, the round's records and the challenge go to sd.
__device__ __forceinline__ void switch_ctl_step(CtlKeep &K, CtlTemp &X, SwitchDev *sd, int r, int V,
                                                const uint32_t *T) {
    auto &st = K.st;
    auto &cl = K.cl;
    auto &L = K.L;
    auto &tabR41 = K.tabR41;
    auto &prod = K.prod;
    auto &rho_s = K.rho_s;
    auto &part = X.part;
    auto &tabL = X.tabL;
    auto &Tt = X.Tt;
    auto &rbas = X.rbas;
    auto &lpart = X.lpart;
    const int t = threadIdx.x, q = t & 3;
    // 4-bit tables of L: tabL[w][j] = sum of L[4w + i] over the set bits i of j (also used for the
    // map update below), and the message sums staged in LDS
    for (int e = t; e < 32 * 16; e += blockDim.x) {
        const int w = e >> 4, nib = e & 15;
        uint64_t v[3] = {0, 0, 0};
        for (int b = 0; b < 4; b++)
            if ((nib >> b) & 1)
                for (int k = 0; k < 3; k++) v[k] ^= L[4 * w + b][k];
        for (int k = 0; k < 3; k++) tabL[w][nib][k] = v[k];
    }
    // Bit-transpose of the message sums by wave ballots: Tt[kind][b][half] holds bit b of T_k for
    // k = 64 half + lane. Wave 2 word + half of kind 0 (waves 0..11), then kind 1.
    {
        const int lane = t & 63;
        for (int task = t >> 6; task < 2 * V; task += blockDim.x >> 6) {
            const int word = task >> 1, half = task & 1, k = 64 * half + lane;
            for (int kind = 0; kind < 2; kind++) {
                const uint32_t x = T[768 * kind + k * V + word];
                for (int bit = 0; bit < 32; bit++) {
                    const uint64_t mk = __ballot((x >> bit) & 1);
                    if (lane == 0) Tt[kind][32 * word + bit][half] = mk;
                }
            }
        }
    }
    __syncthreads();
    {
        // e = sum_k L[k] T_k = sum_b x^b U_b with U_b = sum_{k: bit b of T_k} L[k]: thread b (< 162)
        // for e0, 162 + b for einf; U_b through the tables of L, four values of k per lookup.
        const int kind = t < 162 ? 0 : 1, b = t - 162 * kind;
        uint64_t acc[3] = {0, 0, 0};
        if (t < 324 && b < 32 * V) {
            const uint64_t m0 = Tt[kind][b][0], m1 = Tt[kind][b][1];
#pragma unroll
            for (int w = 0; w < 32; w++) {
                const uint32_t nib = (uint32_t)(((w < 16 ? m0 : m1) >> (4 * (w & 15))) & 15);
                for (int k = 0; k < 3; k++) acc[k] ^= tabL[w][nib][k];
            }
            uint64_t y[3];  // U_b x^b
            if (b < 128) {
                f162_mulx_pow(acc, b, y);
            } else {
                uint64_t z[3];
                f162_mulx_pow(acc, 64, z);
                f162_mulx_pow(z, 64, y);
                f162_mulx_pow(y, b - 128, z);
                y[0] = z[0], y[1] = z[1], y[2] = z[2];
            }
            acc[0] = y[0], acc[1] = y[1], acc[2] = y[2];
        }
        // XOR over each half-wave (lane exchanges within 32 lanes), then over the half-waves of
        // each kind (the kind changes inside half-wave 5)
        uint32_t x[12];
#pragma unroll
        for (int w = 0; w < 3; w++) {
            x[4 * w + 0] = kind == 0 ? (uint32_t)acc[w] : 0u;
            x[4 * w + 1] = kind == 0 ? (uint32_t)(acc[w] >> 32) : 0u;
            x[4 * w + 2] = kind == 1 ? (uint32_t)acc[w] : 0u;
            x[4 * w + 3] = kind == 1 ? (uint32_t)(acc[w] >> 32) : 0u;
        }
#pragma unroll
        for (int i = 0; i < 12; i++) x[i] ^= lane_xor<1>(x[i]);
#pragma unroll
        for (int i = 0; i < 12; i++) x[i] ^= lane_xor<2>(x[i]);
#pragma unroll
        for (int i = 0; i < 12; i++) x[i] ^= lane_xor<4>(x[i]);
#pragma unroll
        for (int i = 0; i < 12; i++) x[i] ^= lane_xor<8>(x[i]);
#pragma unroll
        for (int i = 0; i < 12; i++) x[i] ^= lane_xor<16>(x[i]);
        if ((t & 31) == 0 && t < 352)
#pragma unroll
            for (int w = 0; w < 3; w++) {
                part[t >> 5][w] = (uint64_t)x[4 * w] | ((uint64_t)x[4 * w + 1] << 32);
                part[16 + (t >> 5)][w] = (uint64_t)x[4 * w + 2] | ((uint64_t)x[4 * w + 3] << 32);
            }
    }
    __syncthreads();
    if (t < 6) {  // half-waves 0..10 hold both kinds
        const int kind = t / 3, w = t % 3;
        uint64_t x = 0;
        for (int i = 0; i < 11; i++) x ^= part[16 * kind + i][w];
        prod[kind][w] = x;
    }
    __syncthreads();
    if (t < 4) {  // absorb_f162([e0, einf])
        const uint64_t msg[7] = {2, prod[0][0], prod[0][1], prod[0][2], prod[1][0], prod[1][1], prod[1][2]};
        if (t == 0)
            for (int w = 0; w < 6; w++) sd->rounds[r][w] = msg[w + 1];
        b3d::update_u64s(q, st, msg);
    }
    __syncthreads();
    {  // sample_f162("switch-round", 1): a reader at the counter, then counter += 1
        const uint32_t *src = (const uint32_t *)&st;
        uint32_t *dst = (uint32_t *)&cl;
        for (int i = t; i < (int)(sizeof(B3State) / 4); i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    if (t < 4) {
        const uint64_t head[2] = {st.counter, 12};
        b3d::update_u64s(q, cl, head);
        b3d::update_word(q, cl, 0x74697773u);  // "switch-round", little-endian words
        b3d::update_word(q, cl, 0x722d6863u);
        b3d::update_word(q, cl, 0x646e756fu);
        uint32_t o[16];
        b3d::root_block0(q, cl, o);
        if (t == 0) {
            st.counter = st.counter + 1;
            rho_s[0] = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
            rho_s[1] = (uint64_t)o[2] | ((uint64_t)o[3] << 32);
            rho_s[2] = ((uint64_t)o[4] | ((uint64_t)o[5] << 32)) & ((1ull << 34) - 1);
            for (int w = 0; w < 3; w++) sd->rho[w] = rho_s[w], sd->coords[r][w] = rho_s[w];
        }
    }
    __syncthreads();
    {
        // (the device runs the steps before the last; the host forms the opened value)
        // L_{r+1}[k] = L_r((1 + c) x^k) + rho L_r[k], through the tables of L built above and 4-bit
        // window tables of y -> rho y (from rho x^i, i < 162)
        if (t < 162) {
            const uint64_t rr[3] = {rho_s[0], rho_s[1], rho_s[2]};
            uint64_t z[3], y[3];
            if (t < 128) {
                f162_mulx_pow(rr, t, y);
            } else {
                f162_mulx_pow(rr, 64, z);
                f162_mulx_pow(z, 64, y);
                f162_mulx_pow(y, t - 128, z);
                y[0] = z[0], y[1] = z[1], y[2] = z[2];
            }
            rbas[t][0] = y[0], rbas[t][1] = y[1], rbas[t][2] = y[2];
        }
        __syncthreads();
        for (int e = t; e < 41 * 16; e += blockDim.x) {
            const int w = e >> 4, nib = e & 15;
            uint64_t v[3] = {0, 0, 0};
            for (int b = 0; b < 4; b++)
                if (((nib >> b) & 1) && 4 * w + b < 162)
                    for (int k = 0; k < 3; k++) v[k] ^= rbas[4 * w + b][k];
            for (int k = 0; k < 3; k++) tabR41[w][nib][k] = v[k];
        }
        __syncthreads();
        {  // k = t % 128; eight threads per k: quarters of each table walk
            const int k = t & 127, part = t >> 7;
            uint64_t acc[3] = {0, 0, 0};
            if (part < 4) {  // (1 + c) x^k through the tables of L, windows 8 part .. 8 part + 7
                const u128 one_c = (((u128)sd->r_hi[r][1] << 64) | sd->r_hi[r][0]) ^ (u128)1;
                const u128 z = b128_mulx_pow(one_c, k);
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int w = 8 * part + i;
                    const int nib = (int)((z >> (4 * w)) & 15);
#pragma unroll
                    for (int j = 0; j < 3; j++) acc[j] ^= tabL[w][nib][j];
                }
            } else {  // rho * L[k], windows 0..10, 11..20, 21..30 or 31..40
                const int w0 = part == 4 ? 0 : 10 * (part - 4) + 1, n = part == 4 ? 11 : 10;
#pragma unroll
                for (int i = 0; i < 11; i++)
                    if (i < n) {
                        const int w = w0 + i;
                        const int nib = (int)((L[k][w >> 4] >> (4 * (w & 15))) & 15);
#pragma unroll
                        for (int j = 0; j < 3; j++) acc[j] ^= tabR41[w][nib][j];
                    }
            }
#pragma unroll
            for (int j = 0; j < 3; j++) lpart[part][k][j] = acc[j];
        }
        __syncthreads();
        if (t < 384) {
            const int k = t / 3, j = t % 3;
            uint64_t x = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) x ^= lpart[i][k][j];
            L[k][j] = x;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void switch_state_load(CtlKeep &K, const SwitchDev *sd) {
    const int t = threadIdx.x;
    const uint32_t *src = (const uint32_t *)&sd->ts;
    uint32_t *dst = (uint32_t *)&K.st;
    for (int i = t; i < (int)(sizeof(B3State) / 4); i += blockDim.x) dst[i] = src[i];
    for (int i = t; i < 384; i += blockDim.x) K.L[i / 3][i % 3] = sd->lmap[i / 3][i % 3];
    __syncthreads();
}
__device__ __forceinline__ void switch_state_store(const CtlKeep &K, SwitchDev *sd) {
    const int t = threadIdx.x;
    const uint32_t *src = (const uint32_t *)&K.st;
    uint32_t *dst = (uint32_t *)&sd->ts;
    for (int i = t; i < (int)(sizeof(B3State) / 4); i += blockDim.x) dst[i] = src[i];
    for (int i = t; i < 384; i += blockDim.x) sd->lmap[i / 3][i % 3] = K.L[i / 3][i % 3];
}

// The first step takes the whole state from the host's pinned copy (`host`) and the last device step
// (r == last) returns it there, with the next fold's input (hp_n F162 words from hp_src, he_n B128
// from he_src) into hp_dst / he_dst, so the rounds need no copies around them.
__device__ __forceinline__ void switch_ctl_body(CtlKeep &K, CtlTemp &X, SwitchDev *sd, SwitchDev *host, int r,
                                                int last, int V, const uint32_t *T, const uint32_t *hp_src,
                                                uint32_t *hp_dst, int hp_n, const uint4 *he_src, uint4 *he_dst,
                                                int he_n) {
    static_assert(sizeof(SwitchDev) % 4 == 0, "state copied as words");
    constexpr int SW_WORDS = (int)(sizeof(SwitchDev) / 4);
    if (r == 0) {
        const volatile uint32_t *src = (const volatile uint32_t *)host;
        uint32_t *dst = (uint32_t *)sd;
        for (int i = threadIdx.x; i < SW_WORDS; i += blockDim.x) dst[i] = src[i];
        __threadfence_block();
        __syncthreads();
    }
    switch_state_load(K, sd);
    switch_ctl_step(K, X, sd, r, V, T);
    switch_state_store(K, sd);
    if (r == last) {
        __threadfence_block();
        __syncthreads();
        const uint32_t *src = (const uint32_t *)sd;
        volatile uint32_t *dst = (volatile uint32_t *)host;
        for (int i = threadIdx.x; i < SW_WORDS; i += blockDim.x) dst[i] = src[i];
        for (int i = threadIdx.x; i < 6 * hp_n; i += blockDim.x) hp_dst[i] = hp_src[i];
        for (int i = threadIdx.x; i < he_n; i += blockDim.x) he_dst[i] = he_src[i];
    }
}
__global__ __launch_bounds__(1024) void k_switch_ctl(SwitchDev *sd, SwitchDev *host, int r, int last, int V,
                                                     const uint32_t *T, const uint32_t *hp_src,
                                                     uint32_t *hp_dst, int hp_n, const uint4 *he_src, uint4 *he_dst,
                                                     int he_n) {
    __shared__ CtlKeep K;
    __shared__ CtlTemp X;
    switch_ctl_body(K, X, sd, host, r, last, V, T, hp_src, hp_dst, hp_n, he_src, he_dst, he_n);
}

// ---------------------------------------------------------------------------------------------
// The switch's first six rounds from the prefix sets. With S = log_packed - 6, round r < 6 holds
// p^(r)_(v', m) = sum_a eq(rho_0..r-1, a) lift(w_(a, v', m)) over the prefixes v' of 6 - r bits and
// E^(r)_(v', m) = eq(c_r..5, v') G_m (c = r_hi). Write
//   Q^(r)_v'[k] = sum_m p^(r)_(v', m) bit_k(G_m)   (F162, k < 128),
// so sum_m p_(v', m) L(lambda G_m) = sum_k Q_v'[k] L(lambda x^k) for any lambda, and with v' = (b, v)
// (b its top bit, nv = 2^(5 - r) values of v) the messages' per-bit sums are
//   T0_k   = sum_v sum_i bit_k(lambda0_v x^i) Q_(0,v)[i],                lambda0_v = (1 + c_r) eq(c_r+1..5, v)
//   Tinf_k = sum_v sum_i bit_k(lambda1_v x^i) (Q_(0,v)[i] + Q_(1,v)[i]),  lambda1_v = eq(c_r+1..5, v)
// (e0 = sum_k L[k] T0_k and einf likewise, formed by the control step as from the round kernels'
// sums). The fold is the round kernels' own, Q^(r+1)_v = Q_(0,v) + rho_r (Q_(0,v) + Q_(1,v)), and
// Q^(0)_u[k] = lift(N_u[k]) with N_u the transpose of M_u. Q is 64 x 128 F162 (6 words each).
// ---------------------------------------------------------------------------------------------
// Q^(0) from the sets, one wave per (u, half): lane l holds row b = 64 half + l of M_u; the ballot
// of bit k over the wave is bits 64 half .. 64 half + 63 of N_u[k], kept by lane k mod 64. Block 0
// also copies the lambda values from the host's mailbox (lam: lambda0 of round r at
// 64 - 2^(6-r) + v, lambda1 at 64 more) to lam_dev and zeroes round 0's message buffer `clear`.
__device__ __forceinline__ void prefix_q0_body(int bid, const uint32_t *M, uint32_t *Q, const uint4 *lam,
                                               uint4 *lam_dev, uint32_t *clear) {
    const int t = threadIdx.x, lane = t & 63;
    if (bid == 0) {
        if (t < 128) {
            const volatile uint32_t *l = (const volatile uint32_t *)(lam + t);
            lam_dev[t] = make_uint4(l[0], l[1], l[2], l[3]);
        }
        for (int i = t; i < 1536; i += blockDim.x) clear[i] = 0;
    }
    const int task = bid * 4 + (t >> 6), u = task >> 1, half = task & 1;
    const uint32_t *row = M + ((size_t)u * 128 + 64 * half + lane) * 4;
    const uint32_t r4[4] = {row[0], row[1], row[2], row[3]};
    uint64_t mine[2] = {0, 0};
#pragma unroll
    for (int wd = 0; wd < 4; wd++)
#pragma unroll 1
        for (int b = 0; b < 32; b++) {
            const int k = 32 * wd + b;
            const uint64_t m = __ballot((r4[wd] >> b) & 1);
            if (lane == (k & 63)) mine[wd >> 1] = m;
        }
#pragma unroll
    for (int c = 0; c < 2; c++) {
        uint32_t *q = Q + ((size_t)u * 128 + lane + 64 * c) * 6;
        q[2 * half] = (uint32_t)mine[c], q[2 * half + 1] = (uint32_t)(mine[c] >> 32);
        if (half == 0) q[4] = 0, q[5] = 0;
    }
}
__global__ __launch_bounds__(256) void k_prefix_q0(const uint32_t *M, uint32_t *Q, const uint4 *lam, uint4 *lam_dev,
                                                   uint32_t *clear) {
    prefix_q0_body(blockIdx.x, M, Q, lam, lam_dev, clear);
}

// Round r's sums (r < 6), for r > 0 after the fold of round r - 1 (rho from the device state,
// written by the control step before): item (v, i), v < nv, i < 128, folds Q[v], Q[2 nv + v] into
// Q[v] and Q[nv + v], Q[3 nv + v] into Q[nv + v] (in place: no item reads what another writes),
// then adds bit_k(lambda0_v x^i) Q[v][i] into T0 and bit_k(lambda1_v x^i) (Q[v][i] + Q[nv + v][i])
// into Tinf (zero before). Block 0 also zeroes `clear`, the next round's buffer.
// rtab: RHO_WINDOWS * 16 * 6 words, lds: 4 * M4W<6>::WORDS words (16-byte aligned), both in LDS
__device__ __forceinline__ void prefix_msg_body(int bid, uint32_t *Q, int r, const uint4 *lam, const SwitchDev *sd,
                                                uint32_t *T0, uint32_t *Tinf, uint32_t *clear, uint32_t *rtab,
                                                uint32_t *lds) {
    const int t = threadIdx.x, nv = 32 >> r, items = 128 * nv;
    if (bid == 0)
        for (int i = t; i < 1536; i += blockDim.x) clear[i] = 0;
    if (r > 0) build_rho_table(RhoArg{{sd->rho[0], sd->rho[1], sd->rho[2]}}, lds, rtab);
    const int it = bid * 256 + t;
    uint32_t a[6] = {0, 0, 0, 0, 0, 0}, d[6] = {0, 0, 0, 0, 0, 0};
    uint32_t s0[4] = {0, 0, 0, 0}, s1[4] = {0, 0, 0, 0};
    if (it < items) {
        const int v = it >> 7, i = it & 127;
        uint32_t *q0 = Q + ((size_t)v * 128 + i) * 6, *q1 = Q + ((size_t)(nv + v) * 128 + i) * 6;
        uint32_t b[6];
#pragma unroll
        for (int k = 0; k < 6; k++) a[k] = q0[k], b[k] = q1[k];
        if (r > 0) {  // the fold of round r - 1: Q[v] <- Q[v] + rho (Q[v] + Q[2nv + v]), likewise Q[nv + v]
            const uint32_t *q2 = Q + ((size_t)(2 * nv + v) * 128 + i) * 6, *q3 = Q + ((size_t)(3 * nv + v) * 128 + i) * 6;
            uint32_t x[6], y[6], mx[6], my[6];
#pragma unroll
            for (int k = 0; k < 6; k++) x[k] = a[k] ^ q2[k], y[k] = b[k] ^ q3[k];
            rho_mul<RHO_WINDOWS>(rtab, x, mx);
            rho_mul<RHO_WINDOWS>(rtab, y, my);
#pragma unroll
            for (int k = 0; k < 6; k++) a[k] ^= mx[k], b[k] ^= my[k], q0[k] = a[k], q1[k] = b[k];
        }
#pragma unroll
        for (int k = 0; k < 6; k++) d[k] = a[k] ^ b[k];
        const uint4 *l0 = lam + (64 - 2 * nv);
        const uint4 e0 = b128_mul_xs(l0[v], i), e1 = b128_mul_xs(l0[64 + v], i);
        s0[0] = e0.x, s0[1] = e0.y, s0[2] = e0.z, s0[3] = e0.w;
        s1[0] = e1.x, s1[1] = e1.y, s1[2] = e1.z, s1[3] = e1.w;
    }
    uint32_t *W = lds + (t >> 6) * M4W<6>::WORDS;
    uint32_t acc0[2][6] = {}, acci[2][6] = {};
    m4w_step<6>(acc0, s0, a, W);
    m4w_step<6>(acci, s1, d, W);
    m4w_flush<6, true>(acc0, lds, T0);
    __syncthreads();
    m4w_flush<6, true>(acci, lds, Tinf);
}
__global__ __launch_bounds__(256) void k_prefix_msg(uint32_t *Q, int r, const uint4 *lam, const SwitchDev *sd,
                                                    uint32_t *T0, uint32_t *Tinf, uint32_t *clear) {
    __shared__ uint32_t rtab[RHO_WINDOWS * 16 * 6];
    __shared__ __align__(16) uint32_t lds[4 * M4W<6>::WORDS];  // also the basis while rtab is built
    prefix_msg_body(blockIdx.x, Q, r, lam, sd, T0, Tinf, clear, rtab, lds);
}

// ---------------------------------------------------------------------------------------------
// The six folds at once (p^(6)_m = sum_u eq(rho_0..5, u) lift(w_(u 2^S + m)), the folds of rounds
// 0..5 as in k_switch_round) fused with round 6's message sums, as round 5's kernel would give
// them: for j < h2 = 2^(S-1) the item's outputs p^(6)_j and p^(6)_(j+h2), E^(6) = G (read) and
// E^(7)_j = G_j + G_(j+h2) (written), T0_k += bit_k(G_j) p_j and Tinf_k += bit_k(E^(7)_j)
// (p_j + p_(j+h2)). One output is a binary tree of folds over the six prefix bits, walked
// depth-first: leaf s (five bits u_1..u_5) folds the table words of u = s and u = 32 + s by rho_0,
// each further level the two halves by rho_1..rho_5; the challenges come from the device state
// (the records of the first six control steps). Block 0 also clears `clear` (1536 words).
// ---------------------------------------------------------------------------------------------
constexpr int BIND_THREADS = 512;

// Two outputs (p^(6)_j and p^(6)_(j+h2)) by a group of G = 2^GL lanes: lane g walks, for both, the
// subtree of the prefixes whose last GL bits are g (2^(5-GL) leaves, levels 1 .. 5-GL), the two
// walks' table lookups interleaved; then the group folds its G values across lanes (lanes g and
// g + G/2 by rho_(6-GL), ...); lane 0 of the group ends with the outputs.
__device__ __forceinline__ void f162_fold2(const uint32_t *tab, const uint32_t (&a0)[6], const uint32_t (&b0)[6],
                                           const uint32_t (&a1)[6], const uint32_t (&b1)[6], uint32_t (&o0)[6],
                                           uint32_t (&o1)[6]) {  // a + rho (a + b), twice
    uint32_t d0[6], d1[6], m0[6], m1[6];
#pragma unroll
    for (int k = 0; k < 6; k++) d0[k] = a0[k] ^ b0[k], d1[k] = a1[k] ^ b1[k];
    rho_mul_lean2<RHO_WINDOWS>(tab, d0, d1, m0, m1);
#pragma unroll
    for (int k = 0; k < 6; k++) o0[k] = a0[k] ^ m0[k], o1[k] = a1[k] ^ m1[k];
}

template <int GL>
__device__ __forceinline__ void bind_values(const uint4 *__restrict__ tw, size_t off, size_t stride,
                                            const uint32_t *rt, int g, uint32_t (&out0)[6], uint32_t (&out1)[6]) {
    constexpr int TW = RHO_WINDOWS * 16 * 6, LEAVES = 32 >> GL;
    uint32_t s1[2][6], s2[2][6], s3[2][6], s4[2][6], s5[2][6], v[2][6];
#pragma unroll 1
    for (int k = 0; k < LEAVES; k++) {
        const int sidx = g + (GL < 5 ? (int)((__builtin_bitreverse32((uint32_t)k) >> (27 + GL)) << GL) : 0);
        const uint4 a0 = tw[(size_t)sidx * stride], a1 = tw[(size_t)(32 + sidx) * stride];
        const uint4 b0 = tw[(size_t)sidx * stride + off], b1 = tw[(size_t)(32 + sidx) * stride + off];
        {
            uint32_t da[6] = {a0.x ^ a1.x, a0.y ^ a1.y, a0.z ^ a1.z, a0.w ^ a1.w, 0, 0};
            uint32_t db[6] = {b0.x ^ b1.x, b0.y ^ b1.y, b0.z ^ b1.z, b0.w ^ b1.w, 0, 0}, ma[6], mb[6];
            rho_mul_lean2<32>(rt, da, db, ma, mb);
            v[0][0] = a0.x ^ ma[0], v[0][1] = a0.y ^ ma[1], v[0][2] = a0.z ^ ma[2], v[0][3] = a0.w ^ ma[3];
            v[0][4] = ma[4], v[0][5] = ma[5];
            v[1][0] = b0.x ^ mb[0], v[1][1] = b0.y ^ mb[1], v[1][2] = b0.z ^ mb[2], v[1][3] = b0.w ^ mb[3];
            v[1][4] = mb[4], v[1][5] = mb[5];
        }
        // the depth-first combine: level L's pending left nodes wait in sL
#define BIND_LEVEL(L, SL)                                                        \
    if (GL < 6 - L) {                                                            \
        if (!(k & (1 << (L - 1)))) {                                             \
            _Pragma("unroll") for (int i = 0; i < 6; i++) SL[0][i] = v[0][i], SL[1][i] = v[1][i]; \
            continue;                                                            \
        }                                                                        \
        f162_fold2(rt + L * TW, SL[0], v[0], SL[1], v[1], v[0], v[1]);         \
    }
        BIND_LEVEL(1, s1)
        BIND_LEVEL(2, s2)
        BIND_LEVEL(3, s3)
        BIND_LEVEL(4, s4)
        BIND_LEVEL(5, s5)
#undef BIND_LEVEL
    }
#pragma unroll
    for (int lev = 0; lev < GL; lev++) {  // across the group: level 6 - GL + lev by rho_(6-GL+lev)
        const int half = (1 << GL) >> (lev + 1);
        uint32_t o[2][6];
#pragma unroll
        for (int i = 0; i < 6; i++)
            o[0][i] = (uint32_t)__shfl_xor((int)v[0][i], half), o[1][i] = (uint32_t)__shfl_xor((int)v[1][i], half);
        f162_fold2(rt + (6 - GL + lev) * TW, v[0], o[0], v[1], o[1], v[0], v[1]);  // lanes with bit `half` of g clear keep theirs
    }
#pragma unroll
    for (int i = 0; i < 6; i++) out0[i] = v[0][i], out1[i] = v[1][i];
}

// rt: 6 * RHO_WINDOWS * 16 * 6 words, lds: (BIND_THREADS / 64) * M4W<6>::WORDS words, both in LDS
template <int GL>
__device__ __forceinline__ void switch_bind_body(int bid, int nb, const uint4 *__restrict__ table, int S,
                                                 const SwitchDev *sd, const uint4 *__restrict__ G,
                                                 uint4 *__restrict__ Eout, uint32_t *__restrict__ pout, uint32_t *T0,
                                                 uint32_t *Tinf, uint32_t *clear, uint32_t *rt, uint32_t *lds) {
    if (bid == 0)
        for (int i = threadIdx.x; i < 1536; i += blockDim.x) clear[i] = 0;
    for (int i = 0; i < 6; i++)
        build_rho_table(RhoArg{{sd->coords[i][0], sd->coords[i][1], sd->coords[i][2]}}, lds,
                        rt + i * RHO_WINDOWS * 16 * 6);
    constexpr int GS = 1 << GL, GPB = BIND_THREADS / GS;  // lanes per output, groups per block
    const int g = threadIdx.x & (GS - 1);
    const size_t stride = (size_t)1 << S, h2 = stride / 2;
    const size_t rounds = (h2 + GPB - 1) / GPB;  // group passes over the items, grid-strided
    uint32_t *W = lds + (threadIdx.x >> 6) * M4W<6>::WORDS;
    uint32_t acc0[2][6] = {}, acci[2][6] = {};
    for (size_t pass = bid; pass < rounds; pass += nb) {
        const size_t j = pass * GPB + (threadIdx.x >> GL);
        uint32_t q0[6] = {0, 0, 0, 0, 0, 0}, q1[6] = {0, 0, 0, 0, 0, 0};
        uint32_t e[4] = {0, 0, 0, 0}, e2[4] = {0, 0, 0, 0};
        if (j < h2) {  // the whole group takes this branch
            bind_values<GL>(table + j, h2, stride, rt, g, q0, q1);
            if (g == 0) {
                uint2 *o0 = (uint2 *)(pout + 6 * j), *o1 = (uint2 *)(pout + 6 * (j + h2));
                o0[0] = make_uint2(q0[0], q0[1]), o0[1] = make_uint2(q0[2], q0[3]), o0[2] = make_uint2(q0[4], q0[5]);
                o1[0] = make_uint2(q1[0], q1[1]), o1[1] = make_uint2(q1[2], q1[3]), o1[2] = make_uint2(q1[4], q1[5]);
                const uint4 ea = G[j], eb = G[j + h2];
                e[0] = ea.x, e[1] = ea.y, e[2] = ea.z, e[3] = ea.w;
                e2[0] = ea.x ^ eb.x, e2[1] = ea.y ^ eb.y, e2[2] = ea.z ^ eb.z, e2[3] = ea.w ^ eb.w;
                Eout[j] = make_uint4(e2[0], e2[1], e2[2], e2[3]);
            } else {
#pragma unroll
                for (int k = 0; k < 6; k++) q0[k] = q1[k] = 0;
            }
        }
        m4w_step<6>(acc0, e, q0, W);
        uint32_t dq[6];
#pragma unroll
        for (int k = 0; k < 6; k++) dq[k] = q0[k] ^ q1[k];
        m4w_step<6>(acci, e2, dq, W);
    }
    m4w_flush<6, true>(acc0, lds, T0);
    __syncthreads();
    m4w_flush<6, true>(acci, lds, Tinf);
}
template <int GL>
__global__ __launch_bounds__(BIND_THREADS) void k_switch_bind(const uint4 *__restrict__ table, int S,
                                                              const SwitchDev *sd, const uint4 *__restrict__ G,
                                                              uint4 *__restrict__ Eout, uint32_t *__restrict__ pout,
                                                              uint32_t *T0, uint32_t *Tinf, uint32_t *clear) {
    __shared__ uint32_t rt[6 * RHO_WINDOWS * 16 * 6];
    __shared__ __align__(16) uint32_t lds[(BIND_THREADS / 64) * M4W<6>::WORDS];  // also the tables' basis
    switch_bind_body<GL>(blockIdx.x, gridDim.x, table, S, sd, G, Eout, pout, T0, Tinf, clear, rt, lds);
}

// ---------------------------------------------------------------------------------------------
// Row evaluation (eval.rs:173-184, 426-480): eq table of p0, then per column
//   U_b = sum_i bit_b(w_{c,i}) eq_i   (M4RM)   and   row[c] = sum_b x^b U_b  (mod x^162+x^81+1).
// ---------------------------------------------------------------------------------------------
// The eq table of p0 as the products lo[i mod 2^lo_bits] hi[i >> lo_bits]: block h builds the
// window table of y -> hi[h] y and its 2^lo_bits threads take one product each.
__device__ __forceinline__ void eq_row_body(int hb, const uint64_t *lo, const uint64_t *hi, int lo_bits,
                                            uint32_t *eqrow, uint32_t *tab, uint32_t *basis) {
    const uint64_t *h = hi + 3 * hb;
    build_rho_table(RhoArg{{h[0], h[1], h[2]}}, basis, tab);
    const int t = threadIdx.x;
    if (t >= (1 << lo_bits)) return;
    const uint64_t *a = lo + 3 * t;
    const uint32_t y[6] = {(uint32_t)a[0], (uint32_t)(a[0] >> 32), (uint32_t)a[1],
                           (uint32_t)(a[1] >> 32), (uint32_t)a[2], (uint32_t)(a[2] >> 32)};
    uint32_t r[6];
    rho_mul<RHO_WINDOWS>(tab, y, r);
    uint32_t *o = eqrow + 6 * (((size_t)hb << lo_bits) + t);
#pragma unroll
    for (int k = 0; k < 6; k++) o[k] = r[k];
}
__global__ __launch_bounds__(256) void k_eq_row(const uint64_t *lo, const uint64_t *hi, int lo_bits, uint32_t *eqrow) {
    __shared__ uint32_t tab[RHO_WINDOWS * 16 * 6], basis[162 * 6];
    eq_row_body(blockIdx.x, lo, hi, lo_bits, eqrow, tab, basis);
}

// Block c: U_b = sum_i bit_b(w_{c,i}) eq_i for b < 128 (M4RM over the column), then in the same block
// row[c] = sum_b x^b U_b (mod x^162 + x^81 + 1), one b per thread and an XOR reduction.
struct RowLds {
    __align__(16) uint32_t lds[4 * M4W<6>::WORDS];
    uint32_t s_u[128 * 6];
    uint64_t part[128 * 3];
};
__device__ __forceinline__ void row_body(RowLds &R, size_t c, const uint4 *table, const uint32_t *eqrow, size_t clen,
                                         uint64_t *row) {
    uint32_t *lds = R.lds, *s_u = R.s_u;
    uint64_t *part = R.part;
    uint32_t *W = lds + (threadIdx.x >> 6) * M4W<6>::WORDS;
    uint32_t acc[2][6] = {};
    // two items per pass, both items' loads issued first
    for (size_t base = 0; base < clen; base += 512) {
        uint32_t sel[2][4], val[2][6];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const size_t i = base + 256 * u + threadIdx.x;
#pragma unroll
            for (int k = 0; k < 4; k++) sel[u][k] = 0;
#pragma unroll
            for (int k = 0; k < 6; k++) val[u][k] = 0;
            if (i < clen) {
                const uint4 w = table[c * clen + i];
                sel[u][0] = w.x, sel[u][1] = w.y, sel[u][2] = w.z, sel[u][3] = w.w;
#pragma unroll
                for (int k = 0; k < 6; k++) val[u][k] = eqrow[6 * i + k];
            }
        }
        m4w_step<6>(acc, sel[0], val[0], W);
        if (base + 256 < clen) m4w_step<6>(acc, sel[1], val[1], W);
    }
    m4w_flush<6, false>(acc, lds, s_u);  // U_b of this column into LDS
    __syncthreads();
    const int b = threadIdx.x;
    if (b < 128) {
        const uint32_t *u = s_u + b * 6;
        const uint64_t x[3] = {(uint64_t)u[0] | ((uint64_t)u[1] << 32), (uint64_t)u[2] | ((uint64_t)u[3] << 32),
                               (uint64_t)u[4] | ((uint64_t)u[5] << 32)};
        // x << b into five words (b < 128)
        uint64_t v[5] = {0, 0, 0, 0, 0};
        const int wsh = b >> 6, bsh = b & 63;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v[k + wsh] ^= x[k] << bsh;
            if (bsh) v[k + wsh + 1] ^= x[k] >> (64 - bsh);
        }
        uint64_t r[3];
        f162_reduce5(v, r);
        part[3 * b] = r[0], part[3 * b + 1] = r[1], part[3 * b + 2] = r[2];
    }
    __syncthreads();
    for (int st = 64; st > 0; st >>= 1) {
        if (b < st)
            for (int k = 0; k < 3; k++) part[3 * b + k] ^= part[3 * (b + st) + k];
        __syncthreads();
    }
    if (b < 3) row[c * 3 + b] = part[b];
}
__global__ __launch_bounds__(256) void k_row(const uint4 *table, const uint32_t *eqrow, size_t clen, uint64_t *row) {
    __shared__ RowLds R;
    row_body(R, blockIdx.x, table, eqrow, clen, row);
}

// The row evaluations of several proofs in two launches: the eq tables (proof i's 2^hi_bits blocks
// from eq_start[i]) and the rows (1024 blocks per proof, one per column).
struct RowView {
    const uint4 *table;
    size_t clen;
    const uint64_t *lo, *hi;
    int lo_bits, eq_start;
    uint32_t *eqrow;
    uint64_t *row;
};
__global__ __launch_bounds__(256) void k_meq_row(const RowView *v, int n) {
    __shared__ uint32_t tab[RHO_WINDOWS * 16 * 6], basis[162 * 6];
    int i = 0;
    while (i + 1 < n && (int)blockIdx.x >= v[i + 1].eq_start) i++;
    const RowView &x = v[i];
    eq_row_body((int)blockIdx.x - x.eq_start, x.lo, x.hi, x.lo_bits, x.eqrow, tab, basis);
}
__global__ __launch_bounds__(256) void k_mrow(const RowView *v) {
    __shared__ RowLds R;
    const RowView &x = v[blockIdx.x / COLUMNS];
    row_body(R, blockIdx.x % COLUMNS, x.table, x.eqrow, x.clen, x.row);
}

// ---------------------------------------------------------------------------------------------
// The fold (reference.rs:78-115): for ring element m and residue class k = t mod 4,
//   raw_k[q] = sum_pos S_{pos,k}[q - pos],  S_{pos,k}[p] = sum_{j: pos in c_j} +-bit_p(w_{j,4m+k}),
// with the sign of (j, pos) = sign_t XOR (pos odd). The columns holding each position are listed
// once per fold (positives then negatives); each (pos, k, 32-bit word) task counts its 32 bit
// positions with NB-bit bit-sliced counters. Then X^648 = X^324 - 1 folds raw into 648 values:
//   v[t] = R[t] - R[t+648] - R[t+972] (t < 324),  v[t] = R[t] + R[t+324] (t >= 324).
// ---------------------------------------------------------------------------------------------
// One signed counter per task: a negative term adds the complement of its word, since
// -b = (1 - b) - 1, and the list length of the negatives is subtracted once at extraction.
// Words are added four at a time through carry-save full adders, then rippled.
__device__ __forceinline__ void full_add(uint32_t a, uint32_t b, uint32_t c, uint32_t &sum, uint32_t &carry) {
    const uint32_t t = a ^ b;
    sum = t ^ c;
    carry = (a & b) | (c & t);
}

// One stage of an 8 x 8 bit transpose in each byte lane: rows k and k + J exchange their J x J blocks.
template <int J>
__device__ __forceinline__ void delta_swaps(uint32_t (&x)[8], uint32_t m) {
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (!(k & J)) {
            const uint32_t t = ((x[k] >> J) ^ x[k | J]) & m;
            x[k] ^= t << J;
            x[k | J] ^= t;
        }
}

template <int NB>
__device__ __forceinline__ void ctr_ripple(uint32_t (&c)[NB], uint32_t x, int from) {
#pragma unroll
    for (int i = 0; i < NB; i++)
        if (i >= from) {
            const uint32_t t = c[i] & x;
            c[i] ^= x;
            x = t;
        }
}

template <int NB>
__device__ __forceinline__ void ctr_add4(uint32_t (&c)[NB], uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3) {
    uint32_t s1, k1, k2, k3;
    full_add(c[0], x0, x1, s1, k1);
    full_add(s1, x2, x3, c[0], k2);
    full_add(c[1], k1, k2, c[1], k3);
    ctr_ripple<NB>(c, k3, 2);
}

// Eight words into the counter, Harley-Seal style: planes 0..2 act as the ones / twos / fours of
// a carry-save tree (seven full adders), and only the eights ripple into the planes above.
template <int NB>
__device__ __forceinline__ void ctr_add8(uint32_t (&c)[NB], const uint32_t (&x)[8]) {
    uint32_t tA, tB, fA, fB, e8;
    full_add(c[0], x[0], x[1], c[0], tA);
    full_add(c[0], x[2], x[3], c[0], tB);
    full_add(c[1], tA, tB, c[1], fA);
    full_add(c[0], x[4], x[5], c[0], tA);
    full_add(c[0], x[6], x[7], c[0], tB);
    full_add(c[1], tA, tB, c[1], fB);
    full_add(c[2], fA, fB, c[2], e8);
    ctr_ripple<NB>(c, e8, 3);
}

// One block folds ring elements m = blockIdx.x, blockIdx.x + gridDim.x, ...: the column lists
// (28 x 1024 u16) are staged in LDS once, the 1024 columns' 64-byte element m is staged per m
// (column-major, 16 u32 per column), and task (pos, k) counts the four words of word k of every
// listed column with four NB-plane counters, so each list entry costs one u16 and one 16-byte LDS
// read. Tasks take the positions in the host's order (longest list first), so the 16 positions of
// a wave have lists of similar length.
constexpr int FOLD_THREADS = 704;  // 11 waves for the 648 (pos, k) tasks
constexpr int FOLD_MAX_ENTRIES = 1024 * 32;

struct FoldLds {
    uint4 w[1024 * 4];
    uint16_t cols[FOLD_MAX_ENTRIES];
    int off[2 * 162 + 1 + 162];  // list offsets, then the positions in task order
    int32_t raw[4 * 292];
};

// Ring elements m = m0, m0 + step, ... below m1 of one table; the lists are staged first. Every
// element ends with a block barrier, so a block may run the body again for another table.
template <int NB>
__device__ __forceinline__ void fold_body(FoldLds &L, const uint4 *table, size_t clen, size_t m0, size_t m1,
                                          size_t step, const int *off, const uint16_t *cols, int n_entries,
                                          int32_t *out, uint32_t *flag) {
    uint4 *s_w = L.w;
    uint16_t *s_cols = L.cols;
    int *s_off = L.off;
    int32_t *s_raw = L.raw;
    const int tid = threadIdx.x;
    for (int i = tid; i < n_entries; i += FOLD_THREADS) s_cols[i] = cols[i];
    for (int i = tid; i < 2 * 162 + 1 + 162; i += FOLD_THREADS) s_off[i] = off[i];
    for (size_t m = m0; m < m1; m += step) {
        for (int idx = tid; idx < 1024 * 4; idx += FOLD_THREADS) {
            const int j = idx >> 2, k = idx & 3;
            s_w[idx] = table[(size_t)j * clen + 4 * m + k];
        }
        for (int i = tid; i < 4 * 292; i += FOLD_THREADS) s_raw[i] = 0;
        __syncthreads();
        if (tid < 162 * 4) {
            const int pos = s_off[2 * 162 + 1 + (tid >> 2)], k = tid & 3;
            uint32_t c0[NB], c1[NB], c2[NB], c3[NB];
#pragma unroll
            for (int i = 0; i < NB; i++) c0[i] = c1[i] = c2[i] = c3[i] = 0;
            const int e0 = s_off[2 * pos], e1 = s_off[2 * pos + 1], e2 = s_off[2 * pos + 2];
            // positives then complemented negatives, eight list entries at a time, then four
            int e = e0;
            for (; e + 8 <= e2; e += 8) {
                int ci[8];
#pragma unroll
                for (int u = 0; u < 8; u++) ci[u] = s_cols[e + u];
                uint4 w[8];
#pragma unroll
                for (int u = 0; u < 8; u++) w[u] = s_w[ci[u] * 4 + k];
                uint32_t x[8], y[8], z[8], v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const uint32_t neg = e + u >= e1 ? ~0u : 0u;
                    x[u] = w[u].x ^ neg, y[u] = w[u].y ^ neg, z[u] = w[u].z ^ neg, v[u] = w[u].w ^ neg;
                }
                ctr_add8<NB>(c0, x);
                ctr_add8<NB>(c1, y);
                ctr_add8<NB>(c2, z);
                ctr_add8<NB>(c3, v);
            }
            for (; e + 4 <= e2; e += 4) {
                // the four list entries' indices, then their words, then the complement of the
                // negatives as a mask (no branch: lanes' sign boundaries differ)
                int ci[4];
#pragma unroll
                for (int u = 0; u < 4; u++) ci[u] = s_cols[e + u];
                uint4 w[4];
#pragma unroll
                for (int u = 0; u < 4; u++) w[u] = s_w[ci[u] * 4 + k];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const uint32_t neg = e + u >= e1 ? ~0u : 0u;
                    w[u] = make_uint4(w[u].x ^ neg, w[u].y ^ neg, w[u].z ^ neg, w[u].w ^ neg);
                }
                ctr_add4<NB>(c0, w[0].x, w[1].x, w[2].x, w[3].x);
                ctr_add4<NB>(c1, w[0].y, w[1].y, w[2].y, w[3].y);
                ctr_add4<NB>(c2, w[0].z, w[1].z, w[2].z, w[3].z);
                ctr_add4<NB>(c3, w[0].w, w[1].w, w[2].w, w[3].w);
            }
            for (; e < e2; e++) {
                uint4 w = s_w[s_cols[e] * 4 + k];
                const uint32_t neg = e >= e1 ? ~0u : 0u;
                w = make_uint4(w.x ^ neg, w.y ^ neg, w.z ^ neg, w.w ^ neg);
                ctr_ripple<NB>(c0, w.x, 0);
                ctr_ripple<NB>(c1, w.y, 0);
                ctr_ripple<NB>(c2, w.z, 0);
                ctr_ripple<NB>(c3, w.w, 0);
            }
            const int negs = e2 - e1;
            int32_t *raw = s_raw + k * 292 + pos;
            auto extract = [&](const uint32_t(&c)[NB], int base) {
                if (NB <= 8) {
                    // byte b of x[p] = the count of bit 8 b + p (bit i of it from plane i): an 8 x 8
                    // bit transpose in each byte lane, three delta swaps
                    uint32_t x[8];
#pragma unroll
                    for (int i = 0; i < 8; i++) x[i] = i < NB ? c[i] : 0u;
                    delta_swaps<4>(x, 0x0F0F0F0Fu);
                    delta_swaps<2>(x, 0x33333333u);
                    delta_swaps<1>(x, 0x55555555u);
#pragma unroll
                    for (int p = 0; p < 8; p++) {
                        const uint32_t f = x[p];
#pragma unroll
                        for (int b = 0; b < 4; b++) {
                            // unconditional: adding a zero count costs less than branching around it
                            atomicAdd(raw + base + p + 8 * b, (int)((f >> (8 * b)) & 0xFF) - negs);
                        }
                    }
                } else {
#pragma unroll
                    for (int p = 0; p < 16; p++) {
                        uint32_t f = 0;
#pragma unroll
                        for (int i = 0; i < NB; i++) f |= ((c[i] >> p) & 0x00010001u) << i;
#pragma unroll
                        for (int b = 0; b < 2; b++) {
                            atomicAdd(raw + base + p + 16 * b, (int)((f >> (16 * b)) & 0xFFFF) - negs);
                        }
                    }
                }
            };
            extract(c0, 0);
            extract(c1, 32);
            extract(c2, 64);
            extract(c3, 96);
        }
        __syncthreads();
        for (int t = tid; t < 648; t += FOLD_THREADS) {
            auto R = [&](int tp) -> int { return tp <= 1159 ? s_raw[(tp & 3) * 292 + (tp >> 2)] : 0; };
            const int v = t < 324 ? R(t) - R(t + 648) - R(t + 972) : R(t) + R(t + 324);
            out[m * 648 + t] = v;
            if (v < -32768 || v > 32767) *(volatile uint32_t *)flag = 1u;  // host memory: a plain store
        }
        __syncthreads();
    }
}
template <int NB>
__global__ __launch_bounds__(FOLD_THREADS) void k_fold(const uint4 *table, size_t clen, size_t re, const int *off,
                                                       const uint16_t *cols, int n_entries, int32_t *out,
                                                       uint32_t *flag) {
    __shared__ FoldLds L;
    fold_body<NB>(L, table, clen, blockIdx.x, re, gridDim.x, off, cols, n_entries, out, flag);
}

// The folds of several proofs in one launch: their ring elements in one sequence (proof i's
// elements from first[i]), each block a contiguous range of `per` of them, so the blocks' work is
// even whatever the proofs' sizes; a block that crosses into the next proof stages its lists.
struct FoldView {
    const uint4 *table;
    size_t clen, re, first;
    const int *off;
    const uint16_t *cols;
    int n_entries;
    int32_t *out;
    uint32_t *flag;
};
template <int NB>
__global__ __launch_bounds__(FOLD_THREADS) void k_mfold(const FoldView *v, int n, size_t per) {
    __shared__ FoldLds L;
    const size_t b0 = (size_t)blockIdx.x * per, b1 = b0 + per;
    for (int i = 0; i < n; i++) {
        const FoldView x = v[i];
        if (b1 <= x.first || b0 >= x.first + x.re) continue;
        const size_t lo = b0 > x.first ? b0 - x.first : 0, hi = b1 - x.first < x.re ? b1 - x.first : x.re;
        fold_body<NB>(L, x.table, x.clen, lo, hi, 1, x.off, x.cols, x.n_entries, x.out, x.flag);
    }
}

// ---------------------------------------------------------------------------------------------
// Commitment: exact A*W over Z via CRT over 31-bit primes p = 1 (mod 1944).
//
// Per prime, R_p = Z_p[X]/(X^648 - X^324 + 1) splits as X^324 = eta_h (eta_0 = w^324,
// eta_1 = w^1620, w a primitive 1944-th root); half h is twisted by beta_h^i (beta_0 = w,
// beta_1 = w^5) into a cyclic length-324 DFT with root rho = w^6, computed as an 18 x 18 four-step
// transform: pass A takes, for each n2 < 18, the 18-point DFT over n1 of x[18 n1 + n2] and twists
// output k1 by rho^(n2 k1); pass B takes, for each k1, the 18-point DFT over n2. Each 18-point DFT
// (radices 2, 3, 3) lives in one thread's registers; one LDS transpose sits between the passes.
// Thread (k1) of pass B holds output slots k1*18 + p, p < 18, which is where the MAC consumes
// them, so the slot order is this transform's own (the matrix goes through the same code).
// ---------------------------------------------------------------------------------------------
// 18-point DFT position p holds frequency KREV(p): p = 9 s1 + 3 s2 + s3 -> s1 + 2 s2 + 6 s3.
__host__ __device__ constexpr int krev18(int p) { return p / 9 + 2 * ((p / 3) % 3) + 6 * (p % 3); }

__device__ __forceinline__ void radix3(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t w3, PC pc) {
    const uint32_t t = mont(subm(b, c, pc.p), w3, pc);
    const uint32_t x0 = addm(addm(a, b, pc.p), c, pc.p);
    const uint32_t x1 = addm(subm(a, c, pc.p), t, pc.p);
    const uint32_t x2 = subm(subm(a, b, pc.p), t, pc.p);
    a = x0, b = x1, c = x2;
}

// Lazy corrections by unsigned min (x - 2p wraps when x < 2p).
__device__ __forceinline__ uint32_t red4(uint32_t x, uint32_t p2) { return min(x, x - p2); }  // [0,4p)->[0,2p)
__device__ __forceinline__ uint32_t red6(uint32_t x, uint32_t p2, uint32_t p4) {             // [0,6p)->[0,2p)
    x = min(x, x - p4);
    return min(x, x - p2);
}

// RED: outputs corrected into [0, 2p); otherwise x1 and x2 are left in (0, 6p), which suits a
// Montgomery product as the next step.
template <bool TW, bool RED = true>
__device__ __forceinline__ void radix3_lazy(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t w3, uint32_t w1,
                                            uint32_t w2, PC pc, uint32_t p2, uint32_t p4) {
    const uint32_t t = mont_lazy(b - c + p2, w3, pc);
    const uint32_t s = a + b + c;        // [0, 6p)
    const uint32_t x1 = a - c + t + p2;  // (0, 6p)
    const uint32_t x2 = a - b - t + p4;  // (0, 6p)
    a = RED ? red6(s, p2, p4) : s;
    if (TW) {
        b = mont_lazy(x1, w1, pc);
        c = mont_lazy(x2, w2, pc);
    } else if (RED) {
        b = red6(x1, p2, p4);
        c = red6(x2, p2, p4);
    } else {
        b = x1;
        c = x2;
    }
}

// First stage of the forward 18-point DFT: radix 2 over (j, j + 9) with twiddles w18^j.
__device__ __forceinline__ void dft18_stage1(uint32_t (&v)[18], const uint32_t (&w)[9], PC pc) {
    const uint32_t p2 = 2 * pc.p;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint32_t a = v[j], b = v[j + 9];
        v[j] = red4(a + b, p2);
        const uint32_t d = a - b + p2;  // (0, 4p)
        v[j + 9] = j ? mont_lazy(d, w[j], pc) : red4(d, p2);
    }
}

// Remaining stages (3, then 3). RED_OUT = false leaves the outputs in [0, 6p): every consumer
// that multiplies next (twiddle, MAC, Montgomery conversion) accepts any 32-bit operand.
template <bool RED_OUT>
__device__ __forceinline__ void dft18_stage23(uint32_t (&v)[18], const uint32_t (&w)[9], PC pc) {
    const uint32_t p2 = 2 * pc.p, p4 = 4 * pc.p;
#pragma unroll
    for (int B = 0; B < 18; B += 9) {
        radix3_lazy<false>(v[B], v[B + 3], v[B + 6], w[6], 0, 0, pc, p2, p4);
        radix3_lazy<true>(v[B + 1], v[B + 4], v[B + 7], w[6], w[2], w[4], pc, p2, p4);
        radix3_lazy<true>(v[B + 2], v[B + 5], v[B + 8], w[6], w[4], w[8], pc, p2, p4);
    }
#pragma unroll
    for (int B = 0; B < 18; B += 3) radix3_lazy<false, RED_OUT>(v[B], v[B + 1], v[B + 2], w[6], 0, 0, pc, p2, p4);
}

// Forward 18-point DFT (decimation in frequency: 2, then 3, 3), output in krev18 order.
// Lazy reduction: inputs lie in [0, 2p); the primes satisfy 6p < 2^32, so sums of up to three
// such values stay in 32 bits, and a Montgomery product with a constant below p accepts any
// 32-bit operand and returns [0, 2p) — a value that is multiplied next is never corrected.
template <bool RED_OUT>
__device__ __forceinline__ void dft18_fwd(uint32_t (&v)[18], const uint32_t (&w)[9], PC pc) {
    dft18_stage1(v, w, pc);
    dft18_stage23<RED_OUT>(v, w, pc);
}

// Exact inverse of dft18_fwd up to the factor 18 (decimation in time, inverse roots).
__device__ __forceinline__ void dft18_inv(uint32_t (&v)[18], const uint32_t (&w)[9], PC pc) {
#pragma unroll
    for (int B = 0; B < 18; B += 3) radix3(v[B], v[B + 1], v[B + 2], w[6], pc);
#pragma unroll
    for (int B = 0; B < 18; B += 9)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            if (j) {
                v[B + j + 3] = mont(v[B + j + 3], w[2 * j], pc);
                v[B + j + 6] = mont(v[B + j + 6], w[4 * j], pc);
            }
            radix3(v[B + j], v[B + j + 3], v[B + j + 6], w[6], pc);
        }
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint32_t a = v[j], b = j ? mont(v[j + 9], w[j], pc) : v[j + 9];
        v[j] = addm(a, b, pc.p);
        v[j + 9] = subm(a, b, pc.p);
    }
}

// Pass A for one half, after the DFT stages (outputs in [0, 6p), krev18 order): twist output k1 of
// thread n2 by rho^(n2 k1) and store Y[k1][n2]; the untwisted k1 = 0 is corrected into [0, 2p).
__device__ __forceinline__ void passA_twist_store(const uint32_t (&v)[18], int n2, const uint32_t *twpf,
                                                  uint32_t *Y, PC pc) {
    const uint32_t p2 = 2 * pc.p, p4 = 4 * pc.p;
#pragma unroll
    for (int p = 0; p < 18; p++) {
        const int k1 = krev18(p);
        const uint32_t y = k1 ? mont_lazy(v[p], twpf[n2 * 18 + k1], pc) : red6(v[p], p2, p4);
        Y[k1 * 18 + n2] = y;
    }
}

// Transformed matrix: ahat[m][slot] = T(A_m mod p)[slot], canonical residues. With row_stride
// > 0 the block's element is m = blockIdx.x of key row `row` and lands at ahat[m][slot][row] with
// row_stride rows per slot (the layout of the key-row products).
// One element per block; threads 0..35 run the two halves' transforms.
// The matrix is taken centred (a > Q/2 stands for a - Q, the same residue mod Q), which halves
// the range of the exact product and so the number of primes.
__device__ __forceinline__ uint32_t u128_mod(u128 x, uint32_t p) {
    const uint64_t two64 = ((uint64_t)0 - p) % p;  // 2^64 mod p
    return (uint32_t)((((uint64_t)(x >> 64) % p) * two64 + (uint64_t)x % p) % p);
}

__global__ __launch_bounds__(64) void k_ahat(const uint64_t *amat, const uint32_t *tabs, PC pc, Roots18 R,
                                             uint32_t r2, uint64_t q_lo, uint64_t q_hi, uint32_t *ahat,
                                             int row, int row_stride) {
    __shared__ uint32_t a[648];
    __shared__ uint32_t Y[2 * 324];
    const size_t m = blockIdx.x;
    const int t = threadIdx.x;
    const u128 Q = ((u128)q_hi << 64) | q_lo;
    for (int i = t; i < 648; i += 64) {
        const u128 x = ((u128)amat[(m * 648 + i) * 2 + 1] << 64) | amat[(m * 648 + i) * 2];
        if (x > Q / 2) {
            const uint32_t r = u128_mod(Q - x, pc.p);  // residue of x - Q
            a[i] = r ? pc.p - r : 0;
        } else {
            a[i] = u128_mod(x, pc.p);
        }
    }
    __syncthreads();
    const int h = t / 18, idx = t % 18;
    uint32_t w[9];
#pragma unroll
    for (int j = 0; j < 9; j++) w[j] = R.f[j];
    if (t < 36) {
        uint32_t v[18];
#pragma unroll
        for (int n1 = 0; n1 < 18; n1++) {
            const int i = 18 * n1 + idx;
            const uint32_t u = addm(a[i], mont(a[i + 324], tabs[tab::ETA + h], pc), pc.p);
            v[n1] = mont(u, tabs[tab::TW + h * 324 + i], pc);
        }
        dft18_fwd<false>(v, w, pc);
        passA_twist_store(v, idx, tabs + tab::TWPF, Y + h * 324, pc);
    }
    __syncthreads();
    if (t < 36) {
        uint32_t v[18];
#pragma unroll
        for (int n2 = 0; n2 < 18; n2++) v[n2] = Y[h * 324 + idx * 18 + n2];
        dft18_fwd<false>(v, w, pc);
#pragma unroll
        for (int p = 0; p < 18; p++) {
            const uint32_t x = red6(v[p], 2 * pc.p, 4 * pc.p);
            const size_t slot = m * 648 + h * 324 + idx * 18 + p;
            ahat[row_stride ? slot * row_stride + row : slot] = min(x, x - pc.p);
        }
    }
}

// Forward transform of the binary ring elements of TC columns, MAC against ahat, over the ring
// elements [m0, m1) of this block; its partial sums, reduced mod p, go to its own slice of cacc
// (slice blockIdx.y; the finish kernel adds the slices). 576 threads:
// thread t handles column t / 36, half (t / 18) % 2, and n2 (pass A) / k1 (pass B) = t % 18.
// The MAC is a plain 64-bit multiply-add of v < 6p by a < p (each product below 6p^2 < 2^32 p),
// folded every fourth element as acc = hi * (2^32 mod p) + lo (< 2^32 p): the largest value
// before a fold, 2^32 p + 4 * 6p^2 < 5 * 2^61.42, fits 64 bits. The block's sums are reduced to
// [0, p) and stored into slice blockIdx.y, laid out [slice][column][slot] over gridDim.x * TC
// columns.
constexpr int TC = 16;
constexpr int MAC_THREADS = TC * 36;

// 128-bit logical right shift by s in [0, 31] of four little-endian words.
__device__ __forceinline__ void shr128(const uint4 w, uint32_t s, uint32_t (&r)[4]) {
    r[0] = __builtin_amdgcn_alignbit(w.y, w.x, s);
    r[1] = __builtin_amdgcn_alignbit(w.z, w.y, s);
    r[2] = __builtin_amdgcn_alignbit(w.w, w.z, s);
    r[3] = w.w >> s;
}

template <int POS>
__device__ __forceinline__ uint32_t bit_at(const uint32_t (&r)[4]) {
    if (POS >= 128) return 0;
    return (r[(POS >> 5) & 3] >> (POS & 31)) & 1;
}

// The 4-bit index of the stage-1 table for pair (j, j + 9) of thread n2: the bits of coefficients
// 18 n1 + n2 (low half) and + 324 (high half) for n1 = j and j + 9. Coefficient c is bit c >> 2
// of word c & 3; for even n1 = 2u that is word n2 & 3, which shifted right by n2 >> 2 (ra) puts
// the bit at 9u (+81 for the high half), and for odd n1 = 2u + 1 word (n2 + 2) & 3 shifted by
// (n2 + 18) >> 2 (rb), again at 9u. Bits past 127 read as zero: coefficients >= 512 are zero.
template <int J>
__device__ __forceinline__ uint32_t stage1_sel(const uint32_t (&ra)[4], const uint32_t (&rb)[4]) {
    constexpr int U = J / 2;
    if constexpr ((J & 1) == 0)  // n1 = 2U (ra) and n1 = J + 9 = 2 (U + 4) + 1 (rb)
        return bit_at<9 * U>(ra) | (bit_at<9 * U + 81>(ra) << 1) | (bit_at<9 * U + 36>(rb) << 2) |
               (bit_at<9 * U + 117>(rb) << 3);
    else  // n1 = 2U + 1 (rb) and n1 = J + 9 = 2 (U + 5) (ra)
        return bit_at<9 * U>(rb) | (bit_at<9 * U + 81>(rb) << 1) | (bit_at<9 * U + 45>(ra) << 2) |
               (bit_at<9 * U + 126>(ra) << 3);
}

// st1 points at the half's table: entry (j, o, s) at ((j * 2 + o) * 2) * 16 + s.
template <int J>
__device__ __forceinline__ void stage1_lookup(const uint32_t (&ra)[4], const uint32_t (&rb)[4], const uint32_t *st1,
                                              uint32_t (&v)[18]) {
    const uint32_t s = stage1_sel<J>(ra, rb);
    v[J] = st1[(J * 2 + 0) * 32 + s];
    v[J + 9] = st1[(J * 2 + 1) * 32 + s];
}

// Pass A of one binary ring element for thread (half, n2 = idx): stage 1 by lookup, stages 2
// and 3, the twist by beta^n2 rho^(n2 k1), and the store of Y[k1][n2] of the half (bits: the
// element's four words; wa, wb, sa, sb: the thread's word indices and shifts).
__device__ __forceinline__ void commit_pass_a(const uint4 *bits, int wa, int wb, uint32_t sa, uint32_t sb,
                                              const uint32_t *st1, const uint32_t *twp, const uint32_t (&w)[9],
                                              PC pc, uint32_t *Yh, int idx) {
    uint32_t v[18], ra[4], rb[4];
    shr128(bits[wa], sa, ra);
    shr128(bits[wb], sb, rb);
    stage1_lookup<0>(ra, rb, st1, v);
    stage1_lookup<1>(ra, rb, st1, v);
    stage1_lookup<2>(ra, rb, st1, v);
    stage1_lookup<3>(ra, rb, st1, v);
    stage1_lookup<4>(ra, rb, st1, v);
    stage1_lookup<5>(ra, rb, st1, v);
    stage1_lookup<6>(ra, rb, st1, v);
    stage1_lookup<7>(ra, rb, st1, v);
    stage1_lookup<8>(ra, rb, st1, v);
    dft18_stage23<false>(v, w, pc);
#pragma unroll
    for (int p = 0; p < 18; p++) {
        const int k1 = krev18(p);
        Yh[k1 * 18 + idx] = mont_lazy(v[p], twp[k1], pc);
    }
}

// Pass B for thread (half, k1 = idx): the 18 outputs of slots h*324 + idx*18 + p, in [0, 6p).
__device__ __forceinline__ void commit_pass_b(const uint32_t *Yh, int idx, const uint32_t (&w)[9], PC pc,
                                              uint32_t (&v)[18]) {
#pragma unroll
    for (int n2 = 0; n2 < 18; n2++) v[n2] = Yh[idx * 18 + n2];
    dft18_fwd<false>(v, w, pc);
}

__global__ __launch_bounds__(MAC_THREADS, 5) void k_commit_mac(const uint32_t *table32, size_t clen, int m_per, int re,
                                                            const uint32_t *ahat, const uint32_t *tabs, PC pc,
                                                            uint32_t c32, uint32_t r2, Roots18 R,
                                                            uint32_t *cacc) {
    __shared__ uint32_t Y[TC * 2 * 324];
    __shared__ uint32_t s_st1[tab::ST1_SIZE];
    __shared__ uint32_t s_twp[2 * 324];
    __shared__ uint4 s_bits[TC * 4];
    const int t = threadIdx.x;
    const int col = t / 36, h = (t / 18) & 1, idx = t % 18;
    const int c0 = blockIdx.x * TC;
    const int m0 = blockIdx.y * m_per, m1 = min(re, m0 + m_per);
    for (int i = t; i < tab::ST1_SIZE; i += MAC_THREADS) s_st1[i] = tabs[tab::ST1 + i];
    for (int i = t; i < 2 * 324; i += MAC_THREADS) s_twp[i] = tabs[tab::TWPB + i];
    uint32_t w[9];
#pragma unroll
    for (int j = 0; j < 9; j++) w[j] = R.f[j];
    uint64_t acc[18];
#pragma unroll
    for (int p = 0; p < 18; p++) acc[p] = 0;
    auto fold = [&]() {
#pragma unroll
        for (int p = 0; p < 18; p++) acc[p] = (uint64_t)(uint32_t)(acc[p] >> 32) * c32 + (uint32_t)acc[p];
    };
    const uint32_t *st1 = s_st1 + h * 16;
    const uint32_t *twp = s_twp + (h * 18 + idx) * 18;
    uint32_t *Yh = Y + (col * 2 + h) * 324;
    const int wa = idx & 3, wb = (idx + 2) & 3;
    const uint32_t sa = idx >> 2, sb = (idx + 18) >> 2;
    const uint4 *tab4 = (const uint4 *)table32;
    for (int m = m0; m < m1; m++) {
        if (t < TC * 4) {
            const int bc = t >> 2, wd = t & 3;  // column bc, word wd of ring element m
            s_bits[t] = tab4[(size_t)(c0 + bc) * clen + 4 * (size_t)m + wd];
        }
        __syncthreads();
        commit_pass_a(s_bits + col * 4, wa, wb, sa, sb, st1, twp, w, pc, Yh, idx);
        __syncthreads();
        {
            // pass B and the MAC: slots h*324 + idx*18 + p
            uint32_t v[18];
            commit_pass_b(Yh, idx, w, pc, v);
            const uint2 *am = (const uint2 *)(ahat + (size_t)m * 648 + h * 324 + idx * 18);
#pragma unroll
            for (int q = 0; q < 9; q++) {
                const uint2 a2 = am[q];
                acc[2 * q] += (uint64_t)v[2 * q] * a2.x;
                acc[2 * q + 1] += (uint64_t)v[2 * q + 1] * a2.y;
            }
            if (((m - m0) & 3) == 3) fold();
        }
    }
    // Reduce to [0, p): one fold leaves acc < 2^32 p, so REDC(acc) = acc R^-1 mod p in [0, 2p), and a
    // Montgomery product with R^2 gives acc mod p.
    fold();
    uint32_t red[18];
#pragma unroll
    for (int p = 0; p < 18; p++) {
        const uint32_t m = (uint32_t)acc[p] * pc.pinv_neg;
        const uint32_t redc = (uint32_t)((acc[p] + (uint64_t)m * pc.p) >> 32);
        red[p] = mont(redc, r2, pc);
    }
    const size_t ncols = (size_t)gridDim.x * TC;
    uint2 *dst = (uint2 *)(cacc + ((size_t)blockIdx.y * ncols + c0 + col) * 648 + h * 324 + idx * 18);
#pragma unroll
    for (int q = 0; q < 9; q++) dst[q] = make_uint2(red[2 * q], red[2 * q + 1]);
}

// Inverse transform per column and prime, Garner over the primes, reduction mod Q, packing.
struct CrtArg {
    uint32_t p[MAX_PRIMES];
    uint32_t inv[MAX_PRIMES][MAX_PRIMES];  // inv[i][j] = p_j^-1 R mod p_i (j < i), Montgomery form
    uint64_t q_lo, q_hi;
    uint64_t half_m[3];                     // floor(M / 2), M = product of the primes (192 bits)
    uint64_t m[3];
    uint32_t r2[MAX_PRIMES];                // R^2 mod p
    int nprimes, width, splits;
    int cols;                               // outputs per slice (columns, or key rows x columns)
    PC pc[MAX_PRIMES];
    Roots18 roots[MAX_PRIMES];
};


// 192-bit helpers for the mixed-radix value (up to five 29.4-bit primes)
__device__ __forceinline__ void u192_muladd(uint64_t (&x)[3], uint32_t p, uint32_t v) {
    u128 t = (u128)x[0] * p + v;
    x[0] = (uint64_t)t;
    t = (u128)x[1] * p + (uint64_t)(t >> 64);
    x[1] = (uint64_t)t;
    t = (u128)x[2] * p + (uint64_t)(t >> 64);
    x[2] = (uint64_t)t;
}
__device__ __forceinline__ bool u192_gt(const uint64_t (&a)[3], const uint64_t (&b)[3]) {
    if (a[2] != b[2]) return a[2] > b[2];
    if (a[1] != b[1]) return a[1] > b[1];
    return a[0] > b[0];
}
__device__ __forceinline__ void u192_sub(const uint64_t (&a)[3], const uint64_t (&b)[3], uint64_t (&r)[3]) {
    const uint64_t r0 = a[0] - b[0];
    const uint64_t b0 = a[0] < b[0];
    const uint64_t r1 = a[1] - b[1] - b0;
    const uint64_t b1 = (a[1] < b[1]) || (a[1] == b[1] && b0);
    r[0] = r0, r[1] = r1, r[2] = a[2] - b[2] - b1;
}

// Four waves per column: the transforms use 36 threads, the slice sums and the CRT lift all.
constexpr int FIN_THREADS = 256;
__global__ __launch_bounds__(FIN_THREADS) void k_commit_finish(const uint32_t *cacc, const uint32_t *tabs_all,
                                                      CrtArg crt, uint8_t *outb) {
    __shared__ uint32_t Y[2 * 324];
    __shared__ uint32_t x[648];
    __shared__ uint32_t res[MAX_PRIMES][648];
    const int c = blockIdx.x, t = threadIdx.x;
    const int h = t / 18, idx = t % 18;
    for (int k = 0; k < crt.nprimes; k++) {
        const PC pc = crt.pc[k];
        const uint32_t *tabs = tabs_all + k * tab::SIZE;
        uint32_t w[9];
#pragma unroll
        for (int j = 0; j < 9; j++) w[j] = crt.roots[k].i[j];
        {  // the sum over the MAC's slices (at most 128 residues below p, so below 2^32 p),
           // coalesced over the slots, reduced as REDC then a Montgomery product with R^2; each
           // slice's eleven loads per thread are issued together, so the slices cost one latency
           // each rather than one per slot
            const uint32_t *src = cacc + (size_t)k * crt.splits * crt.cols * 648 + (size_t)c * 648;
            constexpr int PER = (648 + FIN_THREADS - 1) / FIN_THREADS;
            uint64_t sum[PER];
#pragma unroll
            for (int j = 0; j < PER; j++) sum[j] = 0;
            for (int sl = 0; sl < crt.splits; sl++) {
                const uint32_t *row = src + (size_t)sl * crt.cols * 648;
#pragma unroll
                for (int j = 0; j < PER; j++)
                    if (t + FIN_THREADS * j < 648) sum[j] += row[t + FIN_THREADS * j];
            }
#pragma unroll
            for (int j = 0; j < PER; j++)
                if (t + FIN_THREADS * j < 648) x[t + FIN_THREADS * j] = mont(redc64(sum[j], pc), crt.r2[k], pc);
        }
        __syncthreads();
        if (t < 36) {
            // inverse pass B: thread (h, k1 = idx)
            uint32_t v[18];
#pragma unroll
            for (int p = 0; p < 18; p++) v[p] = x[h * 324 + idx * 18 + p];
            dft18_inv(v, w, pc);
#pragma unroll
            for (int n2 = 0; n2 < 18; n2++) Y[h * 324 + idx * 18 + n2] = v[n2];
        }
        __syncthreads();
        if (t < 36) {
            // inverse pass A: thread (h, n2 = idx)
            uint32_t v[18];
#pragma unroll
            for (int p = 0; p < 18; p++) {
                const int k1 = krev18(p);
                const uint32_t y = Y[h * 324 + k1 * 18 + idx];
                v[p] = k1 ? mont(y, tabs[tab::TWPI + idx * 18 + k1], pc) : y;
            }
            dft18_inv(v, w, pc);
#pragma unroll
            for (int n1 = 0; n1 < 18; n1++) x[h * 324 + 18 * n1 + idx] = v[n1];
        }
        __syncthreads();
        for (int i = t; i < 324; i += FIN_THREADS) {
            const uint32_t u0 = mont(x[i], tabs[tab::UNTW + i], pc);
            const uint32_t u1 = mont(x[324 + i], tabs[tab::UNTW + 324 + i], pc);
            const uint32_t ahi = mont(subm(u0, u1, pc.p), tabs[tab::DINV], pc);
            res[k][i] = subm(u0, mont(ahi, tabs[tab::ETA0], pc), pc.p);
            res[k][i + 324] = ahi;
        }
        __syncthreads();
    }
    const u128 Q = ((u128)crt.q_hi << 64) | crt.q_lo;
    for (int tt = t; tt < 648; tt += FIN_THREADS) {
        // Garner: x = v0 + p0 (v1 + p1 (v2 + ...))
        uint32_t v[MAX_PRIMES];
        for (int i = 0; i < crt.nprimes; i++) {
            const uint32_t pi = crt.p[i];
            uint32_t xx = res[i][tt];
            for (int j = 0; j < i; j++) {
                const uint32_t vj = v[j] >= pi ? v[j] - pi : v[j];  // v[j] < p_j < 2 p_i
                xx = mont(subm(xx, vj, pi), crt.inv[i][j], crt.pc[i]);
            }
            v[i] = xx;
        }
        uint64_t X[3] = {v[crt.nprimes - 1], 0, 0};
        for (int i = crt.nprimes - 2; i >= 0; i--) u192_muladd(X, crt.p[i], v[i]);
        // centred lift: the exact coefficient is X or X - M
        const bool neg = u192_gt(X, crt.half_m);
        uint64_t mag[3];
        if (neg)
            u192_sub(crt.m, X, mag);
        else
            mag[0] = X[0], mag[1] = X[1], mag[2] = X[2];
        u128 r = 0;  // mag mod Q
        if (mag[1] == 0 && mag[2] == 0 && crt.q_hi == 0) {
            r = mag[0] % crt.q_lo;
        } else {  // shift-subtract from the top set bit (Q < 2^96)
            const int top = mag[2] ? 128 + 63 - __clzll(mag[2]) : mag[1] ? 64 + 63 - __clzll(mag[1]) : 63;
            for (int b = top; b >= 0; b--) {
                r = (r << 1) | ((mag[b >> 6] >> (b & 63)) & 1);
                if (r >= Q) r -= Q;
            }
        }
        if (neg && r) r = Q - r;
        uint8_t *o = outb + ((size_t)c * 648 + tt) * crt.width;
        for (int b = 0; b < crt.width; b++) o[b] = (uint8_t)(r >> (8 * b));
    }
}

// ---------------------------------------------------------------------------------------------
// Keys of several ring rows. The MAC above keeps 18 sums per thread for one key row; with R rows
// each transformed witness value meets R key values, a product of a (rows x elements) key by an
// (elements x columns) witness per slot, which needs R x columns sums per slot. So the transform
// and the products are split: k_commit_v stores the transformed witness of a chunk of ring
// elements (a buffer small enough to stay in the last-level cache), k_commit_gemm_mfma forms the
// products per slot on the matrix cores (int8 digits of the residues, exact int32 sums),
// k_commit_rowsum adds the slices and transposes them into the finish kernel's per-output layout.
// ---------------------------------------------------------------------------------------------

// Forward transform of the binary ring elements [m0, m1) of TC columns, no MAC: each value is
// reduced to [0, p) and stored at V[(m - mbase) * 648 + slot][column] (ncols columns per slot
// row), staged through LDS so that each slot row's TC columns leave as one 64-byte segment. The
// witness words are loaded one element ahead.
constexpr int VPAD = TC + 1;
static_assert(648 * VPAD >= TC * 2 * 324, "staging buffer smaller than the transpose buffer");
__global__ __launch_bounds__(MAC_THREADS, 2) void k_commit_v(const uint32_t *table32, size_t clen, int mbase,
                                                          int mend, int m_per, const uint32_t *tabs, PC pc,
                                                          Roots18 R, uint32_t *V, int ncols) {
    __shared__ uint32_t Y[648 * VPAD];  // the pass A -> B transpose, then the staged outputs
    __shared__ uint32_t s_st1[tab::ST1_SIZE];
    __shared__ uint32_t s_twp[2 * 324];
    __shared__ uint4 s_bits[TC * 4];
    const int t = threadIdx.x;
    const int col = t / 36, h = (t / 18) & 1, idx = t % 18;
    const int c0 = blockIdx.x * TC;
    const int m0 = mbase + blockIdx.y * m_per, m1 = min(mend, m0 + m_per);
    for (int i = t; i < tab::ST1_SIZE; i += MAC_THREADS) s_st1[i] = tabs[tab::ST1 + i];
    for (int i = t; i < 2 * 324; i += MAC_THREADS) s_twp[i] = tabs[tab::TWPB + i];
    uint32_t w[9];
#pragma unroll
    for (int j = 0; j < 9; j++) w[j] = R.f[j];
    const uint32_t *st1 = s_st1 + h * 16;
    const uint32_t *twp = s_twp + (h * 18 + idx) * 18;
    uint32_t *Yh = Y + (col * 2 + h) * 324;
    const int wa = idx & 3, wb = (idx + 2) & 3;
    const uint32_t sa = idx >> 2, sb = (idx + 18) >> 2;
    const uint4 *tab4 = (const uint4 *)table32;
    const uint32_t p2 = 2 * pc.p, p4 = 4 * pc.p;
    // threads below TC * 4 load word t & 3 of column c0 + t / 4, one element ahead
    const uint4 *mine = tab4 + (size_t)(c0 + (t >> 2)) * clen + (t & 3);
    uint4 next = make_uint4(0, 0, 0, 0);
    if (t < TC * 4 && m0 < m1) next = mine[4 * (size_t)m0];
    for (int m = m0; m < m1; m++) {
        if (t < TC * 4) {
            s_bits[t] = next;
            if (m + 1 < m1) next = mine[4 * (size_t)(m + 1)];
        }
        __syncthreads();  // the bits are in; the previous element's staged outputs are written out
        commit_pass_a(s_bits + col * 4, wa, wb, sa, sb, st1, twp, w, pc, Yh, idx);
        __syncthreads();
        uint32_t v[18];
        commit_pass_b(Yh, idx, w, pc, v);
        __syncthreads();  // every thread has its pass-B inputs: Y becomes the staging buffer
#pragma unroll
        for (int p = 0; p < 18; p++) {
            const uint32_t x = red6(v[p], p2, p4);
            Y[(h * 324 + idx * 18 + p) * VPAD + col] = min(x, x - pc.p);
        }
        __syncthreads();
        uint32_t *dst = V + (size_t)(m - mbase) * 648 * ncols + c0;
        for (int i = t; i < 648 * TC / 4; i += MAC_THREADS) {  // 16-byte stores of 4 columns
            const int slot = i / (TC / 4), c = 4 * (i % (TC / 4));
            const uint32_t *y = Y + slot * VPAD + c;
            *(uint4 *)(dst + (size_t)slot * ncols + c) = make_uint4(y[0], y[1], y[2], y[3]);
        }
    }
}

// Key-row products on the matrix cores (v_mfma_i32_16x16x64_i8). A residue x < p < 2^30 is
// four balanced signed bytes: y = x + 0x808080 and digit i = byte i of y ^ 0x00808080 (bytes 0..2
// less 128, byte 3 below 65), so x = sum_i digit_i 256^i. The product of two residues is then
// sum_{i,j} 256^(i+j) (digit_i a)(digit_j v): sixteen 8-bit products per pair, summed exactly in
// seven i32 classes s = i + j (|class| <= 4 * 128 * 128 * 1024 = 2^26 over at most 1024
// elements per block) and recombined mod p at the end. MFMA lane l holds 16 bytes of its row (A: key row
// l % 16) or column (B: witness column l % 16); byte e of lane l is ring element 16 (l / 16) + e of
// the 64-element step, the same for A and B. The result of tile n: lane l holds rows
// 4 (l / 16) .. +3 of column 16 n + l % 16.
__device__ __forceinline__ uint32_t bal_digits(uint32_t x) { return (x + 0x808080u) ^ 0x808080u; }

// Four words w0..w3 (four digits each) to four words d0..d3, d_i = digit i of w0..w3 (a 4 x 4 byte
// transpose).
__device__ __forceinline__ void byte_transpose4(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t &d0,
                                                uint32_t &d1, uint32_t &d2, uint32_t &d3) {
    const uint32_t lo01 = __builtin_amdgcn_perm(w1, w0, 0x05010400u), hi01 = __builtin_amdgcn_perm(w1, w0, 0x07030602u);
    const uint32_t lo23 = __builtin_amdgcn_perm(w3, w2, 0x05010400u), hi23 = __builtin_amdgcn_perm(w3, w2, 0x07030602u);
    d0 = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
    d1 = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
    d2 = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
    d3 = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
}

// The key's MFMA operands (set-up): amf[slot][64-element block][digit][lane] (16 bytes each) from
// ahat16[m][slot][16 rows]; elements past `elements` are zero.
__global__ __launch_bounds__(64) void k_ahat_mfma(const uint32_t *ahat16, int elements, int nmblk, uint4 *amf) {
    const int s = blockIdx.x, mb = blockIdx.y, l = threadIdx.x;
    const int r = l & 15, kb = l >> 4;
    uint32_t w[16];
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int m = mb * 64 + 16 * kb + e;
        w[e] = bal_digits(m < elements ? ahat16[((size_t)m * 648 + s) * 16 + r] : 0u);
    }
    uint32_t d[4][4];
#pragma unroll
    for (int g = 0; g < 4; g++) byte_transpose4(w[4 * g], w[4 * g + 1], w[4 * g + 2], w[4 * g + 3], d[0][g], d[1][g], d[2][g], d[3][g]);
    uint4 *dst = amf + ((size_t)s * nmblk + mb) * 4 * 64;
#pragma unroll
    for (int i = 0; i < 4; i++) dst[i * 64 + l] = make_uint4(d[i][0], d[i][1], d[i][2], d[i][3]);
}

// One wave per block: slot blockIdx.x, ring elements [y m_per, (y + 1) m_per) of the chunk (m_per a
// multiple of 64, at most 1024), columns 16 NT z .. 16 NT z + 16 NT - 1 (NT column tiles of 16);
// the block's sums mod p go to slice y ([slot][16][ncols]), stored or, with `accumulate`, added to
// what the slice holds (the chunks of one commitment share the slices). mblk0: the chunk's first
// 64-element block in the key's operand layout. A step's witness tile (64 elements x 16 NT
// columns of the slot) is loaded into registers one step ahead, written to LDS at the top of its
// step, and read back a column per lane (16 elements each) for the MFMA operands. A step may run
// past the block's range only at the end of the key (the chunk then ends with the key): those
// rows of V are allocated and hold stale values, and the key's digits there are zero, so they
// add nothing.
struct MfWeights {
    uint32_t w[7];  // 256^c R mod p: the digit classes' weights in Montgomery form
};
constexpr int MF_RSRC3 = 0x00020000;  // buffer resource word 3: 32-bit data format, raw access
typedef int mfma_v4i __attribute__((ext_vector_type(4)));
typedef unsigned int mf_v4u __attribute__((ext_vector_type(4)));
template <int NT>
__global__ __launch_bounds__(64) void k_commit_gemm_mfma(const uint4 *amf, int nmblk, int mblk0, const uint32_t *V,
                                                         int ncols, int mcount, int m_per, PC pc, MfWeights wt,
                                                         int accumulate, uint32_t *out) {
    constexpr int W = 16 * NT;           // columns per block
    constexpr int ROW = W + 1;           // LDS row of the tile (one word of padding spreads the column reads)
    constexpr int QL = W / 4;            // 16-byte column groups per row
    constexpr int RPL = 64 / QL;         // rows per load instruction
    constexpr int NQ = 64 / RPL;         // load instructions per step
    __shared__ uint32_t Vt[64 * ROW];
    const int l = threadIdx.x, s = blockIdx.x, cb = blockIdx.z * W;
    const int kb = l >> 4, cl = l & 15;
    const int m0 = blockIdx.y * m_per, m1 = min(mcount, m0 + m_per);
    const size_t rs = (size_t)648 * ncols;  // V's stride between ring elements
    // lane l loads row l / QL + RPL q (q < NQ) at columns 4 (l % QL) .. +3 of each step. The loads
    // are buffer loads over a per-step base (offsets below 64 * 648 * 4096 * 4 < 2^31): as
    // intrinsics they stay where they are written, one step ahead, instead of being merged with
    // the prologue's loads into one load at the top of the loop.
    const int lr = l / QL, lc = l % QL;
    const uint32_t lane_v = (uint32_t)((lr * rs + 4 * lc) * 4), lane_a = (uint32_t)l * 16;
    const char *vb = (const char *)(V + (size_t)s * ncols + cb);
    const char *ab = (const char *)(amf + ((size_t)s * nmblk + mblk0) * 256);
    auto load_v = [&](int mk, mf_v4u(&x)[NQ]) {
        const __amdgpu_buffer_rsrc_t r =
            __builtin_amdgcn_make_buffer_rsrc((void *)(vb + (size_t)mk * rs * 4), (short)0, 0x7fffffff, MF_RSRC3);
#pragma unroll
        for (int q = 0; q < NQ; q++)
            x[q] = __builtin_amdgcn_raw_buffer_load_b128(r, lane_v, (uint32_t)(RPL * q * rs * 4), 0);
    };
    auto load_a = [&](int mk, mf_v4u(&a)[4]) {
        const __amdgpu_buffer_rsrc_t r =
            __builtin_amdgcn_make_buffer_rsrc((void *)(ab + (size_t)(mk / 64) * 4096), (short)0, 4096, MF_RSRC3);
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = __builtin_amdgcn_raw_buffer_load_b128(r, lane_a + i * 1024, 0, 0);
    };
    mfma_v4i acc[7][NT];
#pragma unroll
    for (int c = 0; c < 7; c++)
#pragma unroll
        for (int n = 0; n < NT; n++) acc[c][n] = mfma_v4i{0, 0, 0, 0};
    // past the range the prefetch repeats the last step (cache hits, unused)
    const int last = m0 + (m1 - m0 - 1) / 64 * 64;
    mf_v4u vx[NQ], an[4];  // the next step's operands, in flight during a step
    load_v(m0, vx);
    load_a(m0, an);
    for (int mk = m0; mk < m1; mk += 64) {
        mfma_v4i av[4];
#pragma unroll
        for (int i = 0; i < 4; i++) av[i] = __builtin_bit_cast(mfma_v4i, an[i]);
        __syncthreads();  // the previous step's reads of the tile are done
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            uint32_t *d = Vt + (lr + RPL * q) * ROW + 4 * lc;
            d[0] = vx[q][0], d[1] = vx[q][1], d[2] = vx[q][2], d[3] = vx[q][3];
        }
        __syncthreads();
        const int mn = min(mk + 64, last);
        load_v(mn, vx);
        load_a(mn, an);
        __builtin_amdgcn_sched_barrier(0);  // all of the next step's loads issue before the products
#pragma unroll
        for (int n = 0; n < NT; n++) {
            uint32_t w[16];
#pragma unroll
            for (int e = 0; e < 16; e++) w[e] = bal_digits(Vt[(16 * kb + e) * ROW + 16 * n + cl]);
            uint32_t d[4][4];
#pragma unroll
            for (int g = 0; g < 4; g++)
                byte_transpose4(w[4 * g], w[4 * g + 1], w[4 * g + 2], w[4 * g + 3], d[0][g], d[1][g], d[2][g], d[3][g]);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const mfma_v4i bv = mfma_v4i{(int)d[j][0], (int)d[j][1], (int)d[j][2], (int)d[j][3]};
#pragma unroll
                for (int i = 0; i < 4; i++)
                    acc[i + j][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[i], bv, acc[i + j][n], 0, 0, 0);
            }
        }
    }
    // recombine: sum_c 256^c class_c mod p, as two Montgomery reductions of four and three
    // products below p^2 each (4 p^2 < 2^32 p)
    uint32_t *dst = out + (size_t)blockIdx.y * 648 * 16 * ncols + ((size_t)s * 16 + 4 * kb) * ncols + cb + cl;
#pragma unroll
    for (int n = 0; n < NT; n++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint64_t lo = 0, hi = 0;
#pragma unroll
            for (int c = 0; c < 7; c++) {
                const int32_t v = acc[c][n][i];  // |v| <= 2^26 < p
                const uint64_t t = (uint64_t)(v < 0 ? (uint32_t)(v + (int32_t)pc.p) : (uint32_t)v) * wt.w[c];
                if (c < 4)
                    lo += t;
                else
                    hi += t;
            }
            uint32_t r = redc64(lo, pc) + redc64(hi, pc);  // < 4p
            r = min(r, r - 2 * pc.p);
            r = min(r, r - pc.p);
            uint32_t *o = dst + (size_t)i * ncols + 16 * n;
            if (accumulate) {
                r += *o;
                r = min(r, r - pc.p);
            }
            *o = r;
        }
}

// The key-row products' slices added and transposed for the finish kernel:
// crow[(r * ncols + c) * 648 + slot] = sum_sl slices[sl][slot][r][c] mod p. Block (x, y): slots
// 16x .. 16x+15 (the last block 8), outputs rc = 64y .. 64y+63 of the row-major (r, c) order;
// each thread keeps four independent sums in flight over the slices.
__global__ __launch_bounds__(256) void k_commit_rowsum(const uint32_t *slices, int nslices, int ncols, PC pc,
                                                       uint32_t r2, uint32_t *crow) {
    __shared__ uint32_t tile[16][65];
    const int t = threadIdx.x;
    const int sb = blockIdx.x * 16, rc0 = blockIdx.y * 64;
    const int r = rc0 / ncols, c0 = rc0 % ncols;
    const size_t sw = (size_t)648 * 16 * ncols;
    const int ci = t & 63;
    for (int k = t >> 6; k < 16; k += 4) {
        if (sb + k >= 648) break;
        const uint32_t *src = slices + ((size_t)(sb + k) * 16 + r) * ncols + c0 + ci;
        uint64_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;  // at most 128 residues below p in all
        int sl = 0;
        for (; sl + 4 <= nslices; sl += 4) {
            a0 += src[(size_t)sl * sw];
            a1 += src[(size_t)(sl + 1) * sw];
            a2 += src[(size_t)(sl + 2) * sw];
            a3 += src[(size_t)(sl + 3) * sw];
        }
        for (; sl < nslices; sl++) a0 += src[(size_t)sl * sw];
        tile[k][ci] = mont(redc64(a0 + a1 + a2 + a3, pc), r2, pc);
    }
    __syncthreads();
    const int rc = t >> 2, q = t & 3;  // four slots of one output per thread, 16-byte stores
    if (sb + 4 * q < 648)
        *(uint4 *)(crow + (size_t)(rc0 + rc) * 648 + sb + 4 * q) =
            make_uint4(tile[4 * q][rc], tile[4 * q + 1][rc], tile[4 * q + 2][rc], tile[4 * q + 3][rc]);
}

// Slices of the one-row MAC added: out[i] = sum_sl in[sl * n + i] mod p for i < n (n = columns x
// 648 per prime, blockIdx.y the prime), so that the finish kernel reads one slice.
__global__ __launch_bounds__(256) void k_commit_slicesum(const uint32_t *in, int nslices, size_t n, CrtArg crt,
                                                         uint32_t *out) {
    const int k = blockIdx.y;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const PC pc = crt.pc[k];
    const uint32_t *src = in + (size_t)k * nslices * n + i;
    uint64_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int sl = 0;
    for (; sl + 4 <= nslices; sl += 4) {
        a0 += src[(size_t)sl * n];
        a1 += src[(size_t)(sl + 1) * n];
        a2 += src[(size_t)(sl + 2) * n];
        a3 += src[(size_t)(sl + 3) * n];
    }
    for (; sl < nslices; sl++) a0 += src[(size_t)sl * n];
    out[(size_t)k * n + i] = mont(redc64(a0 + a1 + a2 + a3, pc), crt.r2[k], pc);
}

// The chaining values of aligned complete subtrees of the commitment's chunks: tree b covers chunk
// CVs cvs[start[b] ..] (2^level[b] of them, relative to the run's first chunk). One workgroup
// merges level by level (parent = compress(key, left || right, counter 0, 64 bytes,
// flags | PARENT)), in LDS when the run fits (IN_LDS: count * 32 bytes of dynamic LDS), else in
// place in global memory with a device-scope fence between levels. Wide levels take one lane per
// parent, narrow ones (at most a quarter of the lanes busy) one quad. Tree b's value is left at
// cvs[8 b].
constexpr int MAX_CV_TREES = 64;
constexpr size_t CV_TREE_LDS_MAX = 120 * 1024;
// ---------------------------------------------------------------------------------------------
// Commitment on the matrix cores (moduli below 2^39): the product in the coefficient domain as one
// int8 GEMM. Column c's commitment is C_c = sum_m A_m W_{c,m} in Z[X]/(X^648 - X^324 + 1); with
// Rot(A)[t][i] the coefficient t of X^i A, C_c[t] = sum_m sum_{i<512} Rot(A_m)[t][i] w_{c,m,i}
// (coefficients >= 512 of W are zero). Each entry, taken centred mod Q (|e| <= Q/2 < 2^38), is
// split into five balanced signed bytes e = sum_d 256^d e_d, so C_c[t] = sum_d 256^d S_d[t][c] with
// S_d = Rot_d W an exact int32 product of int8 digits and 0/1 bits (|S_d| <= 2^21 * 128): the
// witness stays binary and the matrix cores take 64 (element, bit) terms per instruction. The
// reduction X^648 = X^324 - 1 gives, for t < 648 and i < 512,
//   Rot(A)[t][i] = [t >= i] a_{t-i} + [t >= 324, i > t-324] a_{t+324-i}
//                - [t < 324, i > t] a_{t+648-i} - [t < 188, i > t+324] a_{t+972-i}.
// K order of one ring element: 8 steps of 64 terms; in step s = 2q + h, lane group g (lanes
// 16g..16g+15) takes word g of the element and operand byte 4e + j holds bit
// p = 32q + 8j + 4h + e of that word, i.e. coefficient i = 4p + g. With the word's dword q in a
// register, the four operand registers of a step are (dword >> (4h + e)) & 0x01010101.
// ---------------------------------------------------------------------------------------------
constexpr int TP_PLANES = 5;
constexpr int TP_TILES = 41;                 // 16-row tiles over the 648 coefficients (656 rows)
constexpr int TP_ROWS = 16 * TP_TILES;

// Rot(A)[t][i] depends on t and x = t - i only through one of three sequences (the regime of the
// row): for t >= 324, s0(x) = [x >= 0] a_x + [x < 324] a_(x+324); for 188 <= t < 324,
// s1(x) = [x >= 0] a_x - [x < 0] a_(x+648); for t < 188, s2(x) = s1(x) - [x < -324] a_(x+972)
// (a_i = 0 outside 0 <= i < 648; rows 648..671 of the padded tiles read values nobody keeps).
// The K order above puts, in operand register e' of step s = 2q + h' of lane (row t, group g), the
// bytes j' of coefficients i = 128 q + 16 h' + 4 e' + 32 j' + g: the sequence at
// k - 32 j' with k = t - g - 128 q - 16 h' - 4 e'. So one dword per (regime, plane, k),
//   U[k] = bytes (s(k), s(k - 32), s(k - 64), s(k - 96)) of the plane's digits,
// is one operand register, and a lane's four registers of a step are U at k, k - 4, k - 8, k - 12.
// Stored per element as [regime][k mod 4][272][plane] dwords for k in [-415, 672], each phase in
// descending k (j = 0 holds the largest k), so a lane reads upward from its base, the five planes of
// one k side by side.
constexpr int TW_K0 = -415;                                    // smallest k
constexpr int TW_PHASE = 272;                                  // dwords per phase array
constexpr int TW_X0 = -511, TW_XN = 1184;                      // sequence domain [-511, 672]
constexpr size_t TW_ELEM_U32 = 3 * TP_PLANES * 4 * TW_PHASE;   // 16,320 dwords per ring element
constexpr int TW_WIN = 112;                                    // a 32-row group's window per phase
__host__ __device__ constexpr int toep_regime(int t) { return t >= 324 ? 0 : t >= 188 ? 1 : 2; }

// The operand tables (set-up): block m builds element m's 3 x 4 x 272 x 5 dwords.
__global__ __launch_bounds__(256) void k_toep_build(const uint64_t *amat, int re, uint64_t q, uint32_t *out) {
    __shared__ int64_t a[648];
    __shared__ int8_t dig[3][TP_PLANES][TW_XN];
    const int m = blockIdx.x, tid = threadIdx.x;
    const uint64_t half = q / 2;
    const int64_t hq = (int64_t)half, qq = (int64_t)q;
    for (int i = tid; i < 648; i += 256) {
        const uint64_t x = amat[((size_t)m * 648 + i) * 2];  // < q < 2^39: the high word is zero
        a[i] = x > half ? (int64_t)x - qq : (int64_t)x;
    }
    __syncthreads();
    auto A = [&](int i) -> int64_t { return i >= 0 && i < 648 ? a[i] : 0; };
    for (int idx = tid; idx < 3 * TW_XN; idx += 256) {
        const int r = idx / TW_XN, x = idx % TW_XN + TW_X0;
        int64_t v = r == 0 ? (x >= 0 ? A(x) : 0) + (x < 324 ? A(x + 324) : 0) : (x >= 0 ? A(x) : -A(x + 648));
        if (r == 2 && x < -324) v -= A(x + 972);
        while (v > hq) v -= qq;  // |v| <= q before, <= q / 2 after
        while (v < -hq) v += qq;
#pragma unroll
        for (int d = 0; d < TP_PLANES; d++) {
            const int64_t dd = ((v + 128) & 255) - 128;  // balanced digit in [-128, 127]
            v = (v - dd) >> 8;
            dig[r][d][x - TW_X0] = (int8_t)dd;
        }
    }
    __syncthreads();
    for (int idx = tid; idx < (int)TW_ELEM_U32; idx += 256) {
        const int d = idx % TP_PLANES, j = (idx / TP_PLANES) % TW_PHASE, rest = idx / (TP_PLANES * TW_PHASE);
        const int ph = rest % 4, r = rest / 4;
        const int k = TW_K0 + ph + 4 * (TW_PHASE - 1 - j);
        uint32_t w = 0;
#pragma unroll
        for (int jj = 0; jj < 4; jj++) w |= (uint32_t)(uint8_t)dig[r][d][k - 32 * jj - TW_X0] << (8 * jj);
        out[(size_t)m * TW_ELEM_U32 + idx] = w;
    }
}

typedef int tp_v4i __attribute__((ext_vector_type(4)));

// The GEMM: block (row group rg, column half ch, split sp) sums ring elements [m0, m1) of the split
// for rows 32 rg .. 32 rg + 31 (tiles 2 rg and 2 rg + 1; group 20 has tile 40 only) and columns
// 512 ch .. 512 ch + 511; wave w takes 64 of those columns (4 tiles of 16). Each witness operand
// feeds both row tiles (a 32 x 64 wave tile: 160 accumulator registers, two waves per SIMD), so
// the bit-to-byte expansion costs 0.75 VALU instructions per product. Per ring element the block
// copies the window of the operand tables its rows read — k in [32 rg - 415, 32 rg + 32): 112 x 5
// dwords per (regime slot, phase), 8.75 KiB per regime — into LDS one element ahead by
// direct-to-LDS loads; each operand register is then one LDS dword at a constant offset from one of
// the lane's two bases (its row's regime slot, phase (t - g + 415) mod 4 and position; steps 0-3
// and 4-7), two ds_read2_b32 per operand, the next plane's registers read while the current
// plane's 8 products run.
constexpr int T2_GROUPS = (TP_TILES + 1) / 2;   // 21 row groups
constexpr int T2_WAVES = 8;
constexpr int T2_NT = 4;                        // column tiles of 16 per wave
constexpr int T2_RUN = TW_WIN * TP_PLANES;      // dwords of one (slot, phase) window: 560
constexpr int T2_PS = 568;                      // LDS stride of a phase: phases 24 banks apart
constexpr int T2_SLOT = 4 * T2_PS;              // dwords of one regime slot
constexpr int T2_BUF = 2 * T2_SLOT;             // dwords of one element's window (two regime slots)

__device__ __forceinline__ tp_v4i toep_bexp(uint32_t x, int h) {
    return tp_v4i{(int)((x >> h) & 0x01010101u), (int)((x >> (h + 1)) & 0x01010101u),
                  (int)((x >> (h + 2)) & 0x01010101u), (int)((x >> (h + 3)) & 0x01010101u)};
}

template <bool TWO>
__device__ __forceinline__ void toep2_body(const uint32_t *__restrict__ ops, const uint4 *__restrict__ table,
                                           size_t clen, int rg, int ch, int sp, int m0, int m1, int64_t *part,
                                           uint32_t *sU) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = l >> 4, n = l & 15;
    const int r0 = 2 * rg;
    const int c0 = 512 * ch + 16 * T2_NT * w;
    // regime slots: 0 = the regime of row 32 rg, 1 = that of row 32 rg + 31 (the same for 19 groups)
    const int reg0 = toep_regime(32 * rg), reg1 = toep_regime(32 * rg + 31);
    // lane base (dwords within a buffer) of row tile j: row t = 32 rg + 16 j + n; the operand of step s,
    // plane d is at base + 5 (32 (s >> 1) + 4 (s & 1) + e') + d, e' = 0..3
    auto lane_base = [&](int j) {
        const int t = 32 * rg + 16 * j + n, tau = 16 * j + n - g + 415;
        return (toep_regime(t) == reg0 ? 0 : T2_SLOT) + (tau & 3) * T2_PS + TP_PLANES * (TW_WIN - 1 - (tau >> 2));
    };
    const int lb0 = lane_base(0), lb1 = lane_base(1);
    tp_v4i acc0[TP_PLANES][T2_NT], acc1[TP_PLANES][T2_NT];
#pragma unroll
    for (int d = 0; d < TP_PLANES; d++)
#pragma unroll
        for (int t = 0; t < T2_NT; t++) acc0[d][t] = acc1[d][t] = tp_v4i{0, 0, 0, 0};
    // dword q of word g of element mm of column c0 + 16 t + n (steps 2q, 2q + 1), addressed by a 32-bit
    // byte offset from the table (256 MiB at most)
    const uint32_t wbase = (uint32_t)(((size_t)(c0 + n) * clen + g) * 16);
    const uint32_t wtile = (uint32_t)(16 * clen * 16);
    auto wload = [&](uint32_t *dst, int mm, int q) {
#pragma unroll
        for (int t = 0; t < T2_NT; t++)
            dst[t] = *(const uint32_t *)((const char *)table + (wbase + (uint32_t)t * wtile + (uint32_t)mm * 64u + 4u * q));
    };
    // element mm's window into buffer bb: run (slot, phase) = 560 dwords from the table of the slot's
    // regime at j = 160 - 8 rg (k descending), 140 x 16 bytes in three loads of 64 lanes; wave w
    // copies pieces w + 8 i of the (runs x 3) pieces
    const int npieces = (reg0 == reg1 ? 4 : 8) * 3;
    auto dma = [&](int mm, int bb) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int pi = w + 8 * i;
            const int run = pi / 3, piece = pi % 3;
            const int lane16 = 64 * piece + l;  // 16-byte index within the run
            if (pi < npieces && lane16 < T2_RUN / 4) {
                const int sl = run >> 2, ph = run & 3;
                const uint32_t *src = ops + (size_t)mm * TW_ELEM_U32 +
                                      ((size_t)((sl ? reg1 : reg0) * 4 + ph) * TW_PHASE + (160 - 8 * rg)) * TP_PLANES +
                                      4 * lane16;
                const uint32_t lds = __builtin_amdgcn_readfirstlane(
                    (uint32_t)(uintptr_t)(sU + bb * T2_BUF + sl * T2_SLOT + ph * T2_PS + 256 * piece));
                asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(lds), "v"(src) : "memory", "m0");
            }
        }
    };
    // operand registers of (step s, plane d) from the lane base `p` of a row tile
    auto aop = [&](const uint32_t *p, int s, int d) {
        const int o = TP_PLANES * (32 * (s >> 1) + 4 * (s & 1)) + d;
        return tp_v4i{(int)p[o], (int)p[o + TP_PLANES], (int)p[o + 2 * TP_PLANES], (int)p[o + 3 * TP_PLANES]};
    };
    uint32_t wc[T2_NT], wn[T2_NT];
    tp_v4i bv[T2_NT];
    wload(wc, m0, 0);
#pragma unroll
    for (int t = 0; t < T2_NT; t++) bv[t] = toep_bexp(wc[t], 0);
    dma(m0, 0);
    for (int m = m0; m < m1; m++) {
        const int bb = (m - m0) & 1;
        __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of element m and its witness words have landed
        __syncthreads();                // every wave's copies; every wave is done with the other buffer
        if (m + 1 < m1) dma(m + 1, bb ^ 1);
        __builtin_amdgcn_sched_barrier(0);
        // steps 0-3 read from the first bases, steps 4-7 from bases 5 x 64 dwords further
        const uint32_t *pa0 = sU + bb * T2_BUF + lb0, *pa1 = sU + bb * T2_BUF + lb1;
        tp_v4i a0 = aop(pa0, 0, 0), a1;
        if (TWO) a1 = aop(pa1, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            // the witness dword of the next step pair: (m, q + 1), or (m + 1, 0) (after the split's last
            // element a word of that element again, loaded and never used: no predicate in the loop)
            wload(wn, q < 3 ? m : min(m + 1, m1 - 1), (q + 1) & 3);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int s = 2 * q + h;
#pragma unroll
                for (int d = 0; d < TP_PLANES; d++) {
                    __builtin_amdgcn_sched_barrier(0);
                    const int jn = s * TP_PLANES + d + 1;  // the next (step, plane) of the element
                    tp_v4i p0, p1;
                    if (jn < 8 * TP_PLANES) {
                        const int sn = jn / TP_PLANES, dn = jn % TP_PLANES;
                        const int hoff = sn >= 4 ? TP_PLANES * 64 : 0;
                        p0 = aop(pa0 + hoff, sn & 3, dn);
                        if (TWO) p1 = aop(pa1 + hoff, sn & 3, dn);
                    }
#pragma unroll
                    for (int t = 0; t < T2_NT; t++) {
                        acc0[d][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, bv[t], acc0[d][t], 0, 0, 0);
                        if (TWO) acc1[d][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, bv[t], acc1[d][t], 0, 0, 0);
                        // after the last plane's products of column tile t: its witness operands of the
                        // next step, bits from 4 of this pair's dword, or bits from 0 of the next pair's
                        if (d == TP_PLANES - 1) {
                            bv[t] = h == 0 ? toep_bexp(wc[t], 4) : toep_bexp(wn[t], 0);
                        }
                    }
                    if (jn < 8 * TP_PLANES) {
                        a0 = p0;
                        if (TWO) a1 = p1;
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < T2_NT; t++) wc[t] = wn[t];
        }
    }
    // D layout: lane l holds rows 4 g + i (i < 4) of the tile, column n of each column tile
#pragma unroll
    for (int t = 0; t < T2_NT; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int64_t v0 = 0, v1 = 0;
#pragma unroll
            for (int d = 0; d < TP_PLANES; d++) {
                v0 += (int64_t)acc0[d][t][i] * ((int64_t)1 << (8 * d));
                v1 += (int64_t)acc1[d][t][i] * ((int64_t)1 << (8 * d));
            }
            const size_t col = (size_t)c0 + 16 * t + n;
            part[((size_t)sp * TP_ROWS + 16 * r0 + 4 * g + i) * COLUMNS + col] = v0;
            if (TWO) part[((size_t)sp * TP_ROWS + 16 * (r0 + 1) + 4 * g + i) * COLUMNS + col] = v1;
        }
}

// Blocks: 42 per split (21 row groups x 2 column halves), consecutive logical blocks on one XCD so
// the two halves of a row group read the operand window through the same L2.
__global__ __launch_bounds__(64 * T2_WAVES) void k_toep_mac(const uint32_t *__restrict__ ops,
                                                            const uint4 *__restrict__ table, size_t clen, int re,
                                                            int splits, int sp0, int nsp, int64_t *part) {
    const int nb = 2 * T2_GROUPS * nsp, b = blockIdx.x;
    const int full = nb / 8 * 8;
    const int lb = b < full ? (b % 8) * (nb / 8) + b / 8 : b;
    const int ch = lb & 1, rg = (lb >> 1) % T2_GROUPS, sp = sp0 + (lb >> 1) / T2_GROUPS;
    const int m0 = (int)((int64_t)sp * re / splits), m1 = (int)((int64_t)(sp + 1) * re / splits);
    __shared__ __align__(16) uint32_t sU[2 * T2_BUF];
    if (2 * rg + 1 < TP_TILES) toep2_body<true>(ops, table, clen, rg, ch, sp, m0, m1, part, sU);
    else toep2_body<false>(ops, table, clen, rg, ch, sp, m0, m1, part, sU);
}

// Sums the splits (|C| <= 2^20 Q < 2^59), reduces mod Q and packs `width` bytes per coefficient at
// out[(column * 648 + t) * width]. Block (x, y): columns 64 x .. 64 x + 63, rows 16 y .. 16 y + 15;
// 256 threads.
__global__ __launch_bounds__(256) void k_toep_finish(const int64_t *part, int splits, uint64_t q, int width,
                                                     uint8_t *out) {
    __shared__ uint64_t c[64][17];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int col = 64 * blockIdx.x + tx;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int row = 16 * blockIdx.y + 4 * k + ty;
        int64_t v = 0;
        for (int sp = 0; sp < splits; sp++) v += part[((size_t)sp * TP_ROWS + row) * COLUMNS + col];
        int64_t rm = v % (int64_t)q;
        if (rm < 0) rm += (int64_t)q;
        c[tx][4 * k + ty] = (uint64_t)rm;
    }
    __syncthreads();
    const int rows = min(16, 648 - 16 * (int)blockIdx.y);
    const int span = 16 * width;  // bytes of one column's 16 rows
    for (int pos = threadIdx.x; pos < 64 * span; pos += 256) {
        const int cc = pos / span, within = pos % span;
        const int row = within / width, byte = within % width;
        if (row < rows)
            out[((size_t)(64 * blockIdx.x + cc) * 648 + 16 * blockIdx.y + row) * width + byte] =
                (uint8_t)(c[cc][row] >> (8 * byte));
    }
}

struct CvTrees {
    int n;
    uint32_t start[MAX_CV_TREES];
    int level[MAX_CV_TREES];
};
// A commitment hashing staged for a batched launch: the chunk run, its subtrees, the key and the
// head and tail bytes around it.
struct CvJob {
    size_t off, count, head, tail;
    uint64_t first;
    CvTrees tr;
    B3Key key;
};

// The subtree roots go to `roots` (the host's pinned buffer), and the commitment bytes before and
// after the hashed chunks (cb[0, head), cb[tail, tail + tail_len)) to head_out / tail_out (the
// host's mailbox), so the host needs no copies after the kernel.
template <bool IN_LDS>
__device__ __forceinline__ void cv_tree_body(uint32_t *s_cv, uint32_t *cvs, int count, const CvTrees &tr,
                                             const B3Key &key, int l0, uint32_t *roots, const uint8_t *cb, int head,
                                             size_t tail, int tail_len, uint8_t *head_out, uint8_t *tail_out) {
    uint32_t *base = IN_LDS ? s_cv : cvs;
    const int t = threadIdx.x, q = t & 3, quad = t >> 2;
    for (int i = t; i < head; i += blockDim.x) head_out[i] = cb[i];
    for (int i = t; i < tail_len; i += blockDim.x) tail_out[i] = cb[tail + i];
    if (IN_LDS) {  // 16-byte copies, several in flight per thread
        const uint4 *src = (const uint4 *)cvs;
        uint4 *dst = (uint4 *)s_cv;
#pragma unroll 8
        for (int k = t; k < count * 2; k += blockDim.x) dst[k] = src[k];
        __syncthreads();
    }
    uint32_t kw[8];
#pragma unroll
    for (int k = 0; k < 8; k++) kw[k] = key.k[k];
    int maxl = 0;
    for (int b = 0; b < tr.n; b++) maxl = max(maxl, tr.level[b]);
    for (int l = l0; l < maxl; l++) {  // the levels below l0 are merged already
        int total = 0;
        for (int b = 0; b < tr.n; b++)
            if (tr.level[b] > l) total += 1 << (tr.level[b] - l - 1);
        const bool quads = total <= (int)blockDim.x / 4;
        const int first = quads ? quad : t, step = quads ? blockDim.x / 4 : blockDim.x;
        for (int w = first; w < total; w += step) {  // with quads, the same w in the quad's lanes
            int b = 0, rem = w;
            for (;; b++) {
                if (tr.level[b] <= l) continue;
                const int cnt = 1 << (tr.level[b] - l - 1);
                if (rem < cnt) break;
                rem -= cnt;
            }
            uint32_t *left = base + 8 * ((size_t)tr.start[b] + ((size_t)rem << (l + 1)));
            const uint32_t *right = left + 8 * ((size_t)1 << l);
            uint32_t m[16];
#pragma unroll
            for (int k = 0; k < 8; k++) m[k] = left[k], m[k + 8] = right[k];
            if (quads) {
                uint32_t o[16];
                b3d::compress(q, kw, m, 0, 64, key.flags | b3d::PARENT, o);
                if (q == 0)
#pragma unroll
                    for (int k = 0; k < 8; k++) left[k] = o[k];
            } else {
                uint32_t st[16] = {kw[0], kw[1], kw[2], kw[3], kw[4], kw[5], kw[6], kw[7],
                                   0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                   0u, 0u, 64u, key.flags | b3d::PARENT};
                b3_rounds(st, m);
#pragma unroll
                for (int k = 0; k < 8; k++) left[k] = st[k] ^ st[k + 8];
            }
        }
        if (!IN_LDS) __threadfence();
        __syncthreads();
    }
    uint32_t v[8];
    if (t < tr.n)
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = base[8 * (size_t)tr.start[t] + k];
    if (t < tr.n)
#pragma unroll
        for (int k = 0; k < 8; k++) roots[8 * t + k] = v[k];
}
template <bool IN_LDS>
__global__ __launch_bounds__(1024) void k_cv_tree(uint32_t *cvs, int count, CvTrees tr, B3Key key, int l0,
                                                  uint32_t *roots, const uint8_t *cb, int head, size_t tail,
                                                  int tail_len, uint8_t *head_out, uint8_t *tail_out) {
    extern __shared__ __align__(16) uint32_t s_cv[];
    cv_tree_body<IN_LDS>(s_cv, cvs, count, tr, key, l0, roots, cb, head, tail, tail_len, head_out, tail_out);
}

// The commitment hashing and the partials of several proofs: one launch per kernel, proof i's
// blocks from its view's start to the next view's, each proof's arguments in its view.
struct CvView {
    const uint8_t *bytes;
    size_t off;
    uint64_t first;
    size_t count;
    B3Key key;
    uint32_t *cvs;
    int start;  // first window block
    CvTrees tr;
    uint32_t *roots;
    int head, tail_len;
    size_t tail;
    uint8_t *head_out, *tail_out;
};
__global__ __launch_bounds__(64) void k_mchunk_cvs(const CvView *v, int n) {
    __shared__ uint32_t s_cv[64 * 8];
    const CvView &x = v[view_of_block(v, n, blockIdx.x, [](const CvView &y) { return y.start; })];
    chunk_cvs_body((int)blockIdx.x - x.start, x.bytes, x.off, x.first, x.count, x.key, x.cvs, s_cv);
}
template <bool IN_LDS>
__global__ __launch_bounds__(1024) void k_mcv_tree(const CvView *v) {
    extern __shared__ __align__(16) uint32_t s_cv[];
    const CvView &x = v[blockIdx.x];
    cv_tree_body<IN_LDS>(s_cv, x.cvs, (int)x.count, x.tr, x.key, CV_WIN_LEVELS, x.roots, x.bytes, x.head, x.tail,
                         x.tail_len, x.head_out, x.tail_out);
}

// ---------------------------------------------------------------------------------------------
// The folding challenges (pcs/src/challenge.rs:396-715; sampled as main.rs:1877-1893): challenge
// i comes from the transcript reader at counter + i under "short-challenge" (the transcript
// absorbs u64(counter + i), u64(15) and the label into a copy of its state, and reads the XOF of
// that copy's root). One wave per challenge: its first quad runs the copy's chunk state and the
// output stream (the transcript's chaining-value stack is shared and read-only in LDS; entries a completed
// chunk merges with are popped from a private view of it) and draws the positions and signs; the
// 81 sums of the norm check are split over the lanes and the wave votes. A rejected attempt
// continues the same stream, as the host sampler does.
// ---------------------------------------------------------------------------------------------
namespace chal {
constexpr uint32_t START = 1, END = 2, PARENT = 4, ROOT = 8;
constexpr int WAVES = 4;  // challenges per block

// The compression by the wave's first quad (lanes 0..3 run every step of the stream with the same
// values; their LDS stores agree), out written by lane 0.
__device__ __forceinline__ void compress(const uint32_t *cv, const uint32_t *m_in, uint64_t counter, uint32_t len,
                                         uint32_t flags, uint32_t *out, bool full) {
    const int q = threadIdx.x & 3;
    uint32_t c[8], m[16], o[16];
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = cv[k];
#pragma unroll
    for (int k = 0; k < 16; k++) m[k] = m_in[k];
    b3d::compress(q, c, m, counter, len, flags, o);
#pragma unroll
    for (int k = 0; k < 8; k++) out[k] = o[k];
    if (full)
#pragma unroll
        for (int k = 0; k < 8; k++) out[k + 8] = o[k + 8];
}

// The first quad's state, all arrays in the wave's LDS slice.
struct Wave {
    uint32_t cv[8], bw[16];  // chunk chaining value and the open block
    uint32_t priv[2][8];     // stack entries above the shared ones
    uint32_t rcv[8], rm[16]; // root node
    uint32_t rlen, rflags;   // its length and flags
    uint8_t pos[32];         // the chosen attempt's positions and signs
    uint32_t signs;
};

// One speculative attempt per quad: attempt g read from byte 56 g of the stream, which is where
// it starts when every earlier attempt drew its 28 two-byte values without a rejected draw.
constexpr int SPEC = 16;
struct Spec {
    uint32_t out[16];  // current output block of the stream
    uint8_t perm[164];
    uint8_t pos[32];
    uint32_t signs, used;  // sign bits; bytes the attempt consumed
};

struct Reader {
    Wave *w;
    const uint32_t (*stack)[8];  // shared, in LDS
    const uint32_t *key;
    uint32_t flags, buf_len, rlen, rflags;
    int blocks, sd, np;
    uint64_t chunk;

    __device__ void chunk_cv(uint32_t *o) const {
        compress(w->cv, w->bw, chunk, buf_len, flags | END | (blocks == 0 ? START : 0), o, false);
    }
    __device__ const uint32_t *entry(int i) const { return i < sd ? stack[i] : w->priv[i - sd]; }
    __device__ void parent(const uint32_t *l, uint32_t *r) const {
        uint32_t m[16];
#pragma unroll
        for (int k = 0; k < 8; k++) m[k] = l[k], m[k + 8] = r[k];
        compress(key, m, 0, 64, flags | PARENT, r, false);
    }
    __device__ void byte(uint32_t b) {
        if ((uint32_t)blocks * 64 + buf_len == 1024) {  // a full chunk and more input: close it
            uint32_t c[8];
            chunk_cv(c);
            uint64_t total = chunk + 1;
            while ((total & 1) == 0) {
                const uint32_t *top = entry(sd + np - 1);
                uint32_t l[8];
#pragma unroll
                for (int k = 0; k < 8; k++) l[k] = top[k];
                if (np) np--;
                else sd--;
                parent(l, c);
                total >>= 1;
            }
#pragma unroll
            for (int k = 0; k < 8; k++) w->priv[np][k] = c[k];
            np++;
            chunk++;
            blocks = 0;
            buf_len = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) w->cv[k] = key[k];
#pragma unroll
            for (int k = 0; k < 16; k++) w->bw[k] = 0;
        }
        if (buf_len == 64) {
            uint32_t o[8];
            compress(w->cv, w->bw, chunk, 64, flags | (blocks == 0 ? START : 0), o, false);
#pragma unroll
            for (int k = 0; k < 8; k++) w->cv[k] = o[k];
#pragma unroll
            for (int k = 0; k < 16; k++) w->bw[k] = 0;
            blocks++;
            buf_len = 0;
        }
        w->bw[buf_len >> 2] |= (b & 0xFFu) << (8 * (buf_len & 3));
        buf_len++;
    }
    __device__ void finish() {
        const int depth = sd + np;
        if (depth == 0) {
#pragma unroll
            for (int k = 0; k < 8; k++) w->rcv[k] = w->cv[k];
#pragma unroll
            for (int k = 0; k < 16; k++) w->rm[k] = w->bw[k];
            rlen = buf_len;
            rflags = flags | END | (blocks == 0 ? START : 0);
        } else {
            uint32_t right[8];
            chunk_cv(right);
            for (int i = depth - 1; i > 0; i--) parent(entry(i), right);
            const uint32_t *l = entry(0);
#pragma unroll
            for (int k = 0; k < 8; k++) w->rcv[k] = key[k], w->rm[k] = l[k], w->rm[k + 8] = right[k];
            rlen = 64;
            rflags = flags | PARENT;
        }
    }
};
}  // namespace chal

__device__ __forceinline__ void challenges_body(int bid, const ChalArg *arg, lbn_challenge *out) {
    __shared__ chal::Wave waves[chal::WAVES];
    __shared__ chal::Spec spec[chal::WAVES][chal::SPEC];
    __shared__ uint32_t s_stack[54][8];
    __shared__ uint32_t s_key[8];
    __shared__ double s_cos[243], s_sin[243];
    __shared__ int s_units[81];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const B3State *ts = &arg->ts;
    const int depth = ts->depth;
    for (int k = t; k < depth * 8; k += blockDim.x) s_stack[k >> 3][k & 7] = ts->stack[k >> 3][k & 7];
    for (int k = t; k < 243; k += blockDim.x) s_cos[k] = arg->cosv[k], s_sin[k] = arg->sinv[k];
    if (t < 81) s_units[t] = arg->units[t];
    if (t < 8) s_key[t] = ts->key[t];
    __syncthreads();
    const int i = bid * chal::WAVES + wv;  // this wave's challenge
    chal::Wave *w = &waves[wv];
    chal::Reader r;
    if (lane < 4) {
        r.w = w;
        r.stack = s_stack;
        r.key = s_key;
        r.flags = ts->flags;
        for (int k = 0; k < 8; k++) w->cv[k] = ts->cv[k];
        for (int k = 0; k < 16; k++) w->bw[k] = 0;
        r.buf_len = ts->buf_len;
        for (uint32_t b = 0; b < r.buf_len; b++) w->bw[b >> 2] |= (uint32_t)ts->buf[b] << (8 * (b & 3));
        r.blocks = ts->blocks;
        r.chunk = ts->chunk;
        r.sd = depth;
        r.np = 0;
        // reader_at("short-challenge", i): u64(counter + i), u64(15), the label
        const uint64_t ctr = ts->counter + (uint64_t)i;
        for (int k = 0; k < 8; k++) r.byte((uint32_t)(ctr >> (8 * k)));
        for (int k = 0; k < 8; k++) r.byte(k == 0 ? 15u : 0u);
        const char label[16] = "short-challenge";
        for (int k = 0; k < 15; k++) r.byte((uint32_t)(uint8_t)label[k]);
        r.finish();
        w->rlen = r.rlen, w->rflags = r.rflags;
    }
    const double bound_sq = arg->bound_sq;
    // canonical embedding norm of (pos, signs): unit k of Z/243 per lane (k = lane, lane + 64), the
    // same additions in the same order as the host; true (on every lane) if a unit exceeds the bound
    auto too_long = [&](const uint8_t *pos, uint32_t signs) {
        bool bad = false;
        for (int k = lane; k < 81; k += 64) {
            const int u = s_units[k];
            double re = 0.0, im = 0.0;
            for (int tt = 0; tt < WEIGHT; tt++) {
                const int a = ((int)pos[tt] * u) % 243;
                const bool neg = (signs >> tt) & 1;
                re = re + (neg ? -s_cos[a] : s_cos[a]);
                im = im + (neg ? -s_sin[a] : s_sin[a]);
            }
            const double a2 = re * re, b2 = im * im;
            if (a2 + b2 > bound_sq) bad = true;
        }
        return __ballot(bad) != 0;
    };
    // positions and signs of one attempt from a byte source: partial Fisher-Yates over 162
    // positions, two-byte draws below a multiple of the range; then the sign hasher over
    // u64(WEIGHT) || positions, first 4 output bytes. Run by the four lanes of one quad.
    auto attempt = [&](auto &&next, uint8_t *perm, uint8_t *pos, uint32_t &signs) {
        for (int k = 0; k < 162; k++) perm[k] = (uint8_t)k;
        uint64_t seen[3] = {0, 0, 0};
        for (int k = 0; k < WEIGHT; k++) {
            const uint32_t n = 162 - k, limit = (65536u / n) * n;
            uint32_t v;
            for (;;) {
                const uint32_t lo = next(), hi = next();
                v = lo | (hi << 8);
                if (v < limit) break;
            }
            const int j = k + (int)(v % n);
            const uint8_t pj = perm[j];
            perm[j] = perm[k];
            seen[pj >> 6] |= 1ull << (pj & 63);
        }
        int n = 0;
        for (int ww = 0; ww < 3; ww++)
            for (uint64_t b = seen[ww]; b; b &= b - 1) pos[n++] = (uint8_t)(64 * ww + __builtin_ctzll(b));
        uint32_t m[16], o[8];
        for (int k = 0; k < 16; k++) m[k] = 0;
        m[0] = WEIGHT;
        for (int k = 0; k < WEIGHT; k++) m[2 + (k >> 2)] |= (uint32_t)pos[k] << (8 * (k & 3));
        chal::compress(arg->sign_key, m, 0, 8 + WEIGHT, arg->sign_flags | chal::START | chal::END | chal::ROOT, o, false);
        signs = o[0] & ((1u << WEIGHT) - 1);
    };
    // The attempts in parallel, one per quad, then the first that passes: a rejection (about a
    // third of the attempts) no longer costs a whole attempt's latency. Quad h runs the attempt
    // that starts at byte base + 56 h, its start when the attempts before it consumed 56 bytes
    // each (no rejected draw); scanning in order, an attempt that consumed more ends the round and
    // the next round starts right after it.
    __syncwarp();  // the root node is in LDS
    uint64_t base = 0;
    for (;;) {
        {
            chal::Spec *sq = &spec[wv][lane >> 2];
            const uint64_t start = base + 56u * (uint32_t)(lane >> 2);
            uint64_t blk = start >> 6;
            uint32_t pp = (uint32_t)(start & 63), used = 0;
            chal::compress(w->rcv, w->rm, blk, w->rlen, w->rflags | chal::ROOT, sq->out, true);
            auto next = [&]() -> uint32_t {
                if (pp == 64) {
                    chal::compress(w->rcv, w->rm, ++blk, w->rlen, w->rflags | chal::ROOT, sq->out, true);
                    pp = 0;
                }
                const uint32_t b = (sq->out[pp >> 2] >> (8 * (pp & 3))) & 0xFF;
                pp++;
                used++;
                return b;
            };
            uint32_t sg;
            attempt(next, sq->perm, sq->pos, sg);
            sq->signs = sg;
            sq->used = used;
        }
        __syncwarp();
        int chosen = -1;
        uint64_t nbase = base + 56u * chal::SPEC;
        for (int g = 0; g < chal::SPEC; g++) {
            const chal::Spec *sq = &spec[wv][g];
            if (!too_long(sq->pos, sq->signs)) {
                chosen = g;
                break;
            }
            if (sq->used != 56) {
                nbase = base + 56u * g + sq->used;
                break;
            }
        }
        if (chosen >= 0) {
            if (lane < 4) {
                const chal::Spec *sq = &spec[wv][chosen];
                for (int k = 0; k < 32; k++) w->pos[k] = sq->pos[k];
                w->signs = sq->signs;
            }
            __syncwarp();
            break;
        }
        __syncwarp();  // every lane has read the round's attempts before they are overwritten
        base = nbase;
    }
    if (lane < 8) {  // positions (32 bytes, zero past WEIGHT), signs, weight
        uint32_t v;
        if (lane < 7)
            v = (uint32_t)w->pos[4 * lane] | ((uint32_t)w->pos[4 * lane + 1] << 8) | ((uint32_t)w->pos[4 * lane + 2] << 16) |
                ((uint32_t)w->pos[4 * lane + 3] << 24);
        else
            v = 0;
        ((uint32_t *)&out[i])[lane] = v;
    }
    if (lane == 8) out[i].signs = w->signs;
    if (lane == 9) out[i].weight = WEIGHT;
}
__global__ __launch_bounds__(64 * chal::WAVES) void k_challenges(const ChalArg *arg, lbn_challenge *out) {
    challenges_body(blockIdx.x, arg, out);
}
// The challenges of many proofs in one launch: instance i's input and output at arg[i], out[i], its
// challenges in blocks [i COLUMNS / WAVES, (i + 1) COLUMNS / WAVES).
struct ChalPtrs {
    const ChalArg *arg;
    lbn_challenge *out;
};
__global__ __launch_bounds__(64 * chal::WAVES) void k_mchallenges(const ChalPtrs *p) {
    constexpr int PER = COLUMNS / chal::WAVES;
    const ChalPtrs x = p[blockIdx.x / PER];
    challenges_body(blockIdx.x % PER, x.arg, x.out);
}

// =============================================================================================
// Host side of the device passes
// =============================================================================================

int lbn::dev_init(lbn_ctx *ctx) {
    auto *d = new DeviceState();
    ctx->dev = d;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamCreateWithFlags(&d->own_stream, hipStreamNonBlocking));
    d->stream = d->own_stream;
    const Shape &s = ctx->shape;
    HIPCHK(ctx, hipMalloc(&d->table, s.n_words * 16));
    HIPCHK(ctx, hipMalloc(&d->E[0], s.n_words * 16));
    HIPCHK(ctx, hipMalloc(&d->E[1], s.n_words / 2 * 16));
    HIPCHK(ctx, hipMalloc(&d->P[0], s.n_words / 2 * 24));
    HIPCHK(ctx, hipMalloc(&d->P[1], s.n_words / 4 * 24));
    HIPCHK(ctx, hipMalloc(&d->T, 3584 * 4));
    HIPCHK(ctx, hipMalloc(&d->pset, ps::WORDS * 4));
    // EL then EH in one allocation, laid out like their mailbox regions (one copy per proof)
    HIPCHK(ctx, hipMalloc(&d->EL, (4096 + std::max<size_t>((size_t)1 << (s.log_packed - 12), 128)) * 16));
    d->EH = d->EL + 4096;
    HIPCHK(ctx, hipMalloc(&d->eqrow, s.column_len * 24));
    HIPCHK(ctx, hipMalloc(&d->eqlo, 256 * 24));
    // the fold's offsets then its column lists, at the mailbox's relative offset (one copy per proof)
    static_assert((2 * 162 + 1 + 162) * sizeof(int) <= mb::FCOLS - mb::FOFF, "fold offsets region");
    HIPCHK(ctx, hipMalloc(&d->fold_off, (mb::FCOLS - mb::FOFF) + (size_t)COLUMNS * 32 * sizeof(uint16_t)));
    d->fold_cols = (uint16_t *)((uint8_t *)d->fold_off + (mb::FCOLS - mb::FOFF));
    HIPCHK(ctx, hipMalloc(&d->ntt_tabs, (size_t)MAX_PRIMES * tab::SIZE * 4));
    HIPCHK(ctx, hipHostMalloc((void **)&d->hfold, s.ring_elements * N * sizeof(int32_t)));
    HIPCHK(ctx, hipHostMalloc((void **)&d->hbox, mb::SIZE));
    HIPCHK(ctx, hipMalloc(&d->sw, sizeof(SwitchDev)));
    HIPCHK(ctx, hipMalloc(&d->chal, (sizeof(ChalArg) + 15) / 16 * 16));  // whole 16-byte words
    HIPCHK(ctx, hipHostGetDevicePointer((void **)&d->fold_host, d->hfold, 0));
    HIPCHK(ctx, hipHostGetDevicePointer((void **)&d->hbox_dev, d->hbox, 0));
    HIPCHK(ctx, hipStreamCreateWithFlags(&d->own_copy, hipStreamNonBlocking));
    d->copy_stream = d->own_copy;
    HIPCHK(ctx, hipEventCreateWithFlags(&d->ev_commit, hipEventDisableTiming));
    HIPCHK(ctx, hipEventCreateWithFlags(&d->ev_staged, hipEventDisableTiming));
    for (auto &e : d->ev) HIPCHK(ctx, hipEventCreate(&e));
    for (auto &e : d->ev_chunk) HIPCHK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return capture_switch(ctx);
}

void lbn::dev_use_streams(lbn_ctx *ctx, void *main, void *copy) {
    DeviceState *d = ctx->dev;
    d->stream = main ? (hipStream_t)main : d->own_stream;
    d->copy_stream = copy ? (hipStream_t)copy : d->own_copy;
}

void lbn::dev_free(lbn_ctx *ctx) {
    DeviceState *d = ctx->dev;
    if (!d) return;
    hipSetDevice(ctx->device);
    void *ptrs[] = {d->table, d->E[0], d->E[1], d->P[0], d->P[1], d->T, d->pset, d->EL, d->eqrow, d->eqlo,
                    d->fold_off, d->ntt_tabs, d->ahat, d->cacc, d->cbytes, d->amat, d->cvs, d->sw,
                    d->chal, d->tpart};
    if (d->stream) hipStreamSynchronize(d->stream);
    if (d->copy_stream) hipStreamSynchronize(d->copy_stream);
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (d->hstage) hipHostFree(d->hstage);
    if (d->hfold) hipHostFree(d->hfold);
    if (d->hbox) hipHostFree(d->hbox);
    if (d->hcvs) hipHostFree(d->hcvs);
    for (auto &e : d->ev)
        if (e) hipEventDestroy(e);
    for (auto &e : d->ev_chunk)
        if (e) hipEventDestroy(e);
    if (d->ev_commit) hipEventDestroy(d->ev_commit);
    if (d->ev_staged) hipEventDestroy(d->ev_staged);
    if (d->switch_graph) hipGraphExecDestroy(d->switch_graph);
    if (d->own_copy) hipStreamDestroy(d->own_copy);
    if (d->own_stream) hipStreamDestroy(d->own_stream);
    d->toep.reset();
    delete d;
    ctx->dev = nullptr;
}

// Ring elements [m0, m1) of every column of the pending host table (one 2D copy on the copy stream);
// the main stream waits for it.
static int upload_range(lbn_ctx *ctx, int g, size_t m0, size_t m1) {
    DeviceState *d = ctx->dev;
    const size_t pitch = ctx->shape.column_len * 16;
    HIPCHK(ctx, hipMemcpy2DAsync((uint8_t *)d->table + 64 * m0, pitch, d->pend + 64 * m0, pitch, 64 * (m1 - m0),
                                 COLUMNS, hipMemcpyHostToDevice, d->copy_stream));
    HIPCHK(ctx, hipEventRecord(d->ev_chunk[g], d->copy_stream));
    HIPCHK(ctx, hipStreamWaitEvent(d->stream, d->ev_chunk[g], 0));
    return LBN_OK;
}

// The whole pending table, before a pass other than the commitment reads it.
static int table_ready(lbn_ctx *ctx) {
    if (!ctx->dev->pend) return LBN_OK;
    const int rc = upload_range(ctx, 0, 0, ctx->shape.ring_elements);
    ctx->dev->pend = nullptr;
    return rc;
}

extern "C" int lbn_load_table(lbn_ctx *ctx, const void *host_bytes, size_t bytes) {
    if (!ctx || !host_bytes) return LBN_E_ARG;
    if (bytes != ctx->shape.table_bytes)
        return ctx->fail(LBN_E_ARG, "witness table has the wrong length");
    ctx->dev->pend = nullptr;  // replaces a deferred upload
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemcpyAsync(ctx->dev->table, host_bytes, bytes, hipMemcpyHostToDevice, ctx->dev->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->dev->stream));
    ctx->table_loaded = true;
    ctx->partials_done = false;
    ctx->round = -1;
    return LBN_OK;
}

extern "C" int lbn_load_table_async(lbn_ctx *ctx, const void *host_bytes, size_t bytes) {
    if (!ctx || !host_bytes) return LBN_E_ARG;
    if (bytes != ctx->shape.table_bytes)
        return ctx->fail(LBN_E_ARG, "witness table has the wrong length");
    ctx->dev->pend = (const uint8_t *)host_bytes;
    ctx->table_loaded = true;
    ctx->partials_done = false;
    ctx->round = -1;
    return LBN_OK;
}

extern "C" int lbn_load_table_device(lbn_ctx *ctx, const void *device_bytes, size_t bytes) {
    if (!ctx || !device_bytes) return LBN_E_ARG;
    if (bytes != ctx->shape.table_bytes)
        return ctx->fail(LBN_E_ARG, "witness table has the wrong length");
    ctx->dev->pend = nullptr;  // replaces a deferred upload
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if ((uintptr_t)device_bytes % 16) return ctx->fail(LBN_E_ARG, "device table must be 16-byte aligned");
    // a plain 16-byte-per-lane copy kernel (reads the table once, writes it once)
    const size_t words = bytes / 16;
    k_copy_u4<<<(unsigned)std::min<size_t>((words + 255) / 256, 65536), 256, 0, ctx->dev->stream>>>(
        (const uint4 *)device_bytes, ctx->dev->table, words);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->dev->stream));
    ctx->table_loaded = true;
    ctx->partials_done = false;
    ctx->round = -1;
    return LBN_OK;
}

extern "C" int lbn_copy_probe(lbn_ctx *ctx) {
    if (!ctx) return LBN_E_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    {
        const int rc_ = table_ready(ctx);  // a deferred upload completes first
        if (rc_ != LBN_OK) return rc_;
    }
    // the table's bytes into the eq buffer (the same size) with the plain copy kernel; the eq
    // expansion is rebuilt by the next lbn_partials
    const size_t words = ctx->shape.table_bytes / 16;
    k_copy_u4<<<(unsigned)std::min<size_t>((words + 255) / 256, 65536), 256, 0, ctx->dev->stream>>>(
        ctx->dev->table, ctx->dev->E[0], words);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->dev->stream));
    ctx->partials_done = false;
    ctx->round = -1;
    return LBN_OK;
}

// The root node of the table writer's XOF for a seed (verifier/table_writer/src/main.rs:14-22):
// keyed by the seed, one block holding the domain string.
static B3Root table_root(const uint8_t seed[32]) {
    static const char CTXT[] = "labinius-fast/binary-table/v1";
    B3Root root{};
    for (int i = 0; i < 8; i++)
        root.cv[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) | ((uint32_t)seed[4 * i + 2] << 16) |
                     ((uint32_t)seed[4 * i + 3] << 24);
    uint8_t block[64] = {0};
    memcpy(block, CTXT, sizeof CTXT - 1);
    for (int i = 0; i < 16; i++)
        root.m[i] = (uint32_t)block[4 * i] | ((uint32_t)block[4 * i + 1] << 8) | ((uint32_t)block[4 * i + 2] << 16) |
                    ((uint32_t)block[4 * i + 3] << 24);
    root.len = sizeof CTXT - 1;
    root.flags = 16u /*KEYED_HASH*/ | 1u /*CHUNK_START*/ | 2u /*CHUNK_END*/ | 8u /*ROOT*/;
    return root;
}

extern "C" int lbn_generate_table(lbn_ctx *ctx, const uint8_t seed[32]) {
    if (!ctx || !seed) return LBN_E_ARG;
    ctx->dev->pend = nullptr;  // replaces a deferred upload
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const B3Root root = table_root(seed);
    const size_t blocks = ctx->shape.table_bytes / 64;
    tstart(ctx, T_TABLE);
    k_table_xof<<<(unsigned)((blocks + 255) / 256), 256, 0, ctx->dev->stream>>>(root, (uint32_t *)ctx->dev->table,
                                                                              blocks);
    tstop(ctx, T_TABLE);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->dev->stream));
    tcollect(ctx, T_TABLE);
    ctx->table_loaded = true;
    ctx->partials_done = false;
    ctx->round = -1;
    return LBN_OK;
}

// ----------------------------------------------------------------------------- pinned buffers
// Host buffers from lbn_host_alloc are pinned and mapped into the device's address space; the
// prover writes the fold and the commitment bytes straight into such a proof buffer instead of
// staging them. Only ranges inside a live lbn_host_alloc allocation qualify.
namespace {
std::mutex g_pin_m;
std::map<uintptr_t, size_t> g_pins;  // base -> bytes
}  // namespace

extern "C" void *lbn_host_alloc(size_t bytes) {
    if (!bytes) return nullptr;
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocPortable | hipHostMallocMapped) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_pin_m);
    g_pins[(uintptr_t)p] = bytes;
    return p;
}

extern "C" void lbn_host_free(void *p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_pin_m);
        if (!g_pins.erase((uintptr_t)p)) return;  // not ours
    }
    hipHostFree(p);
}

// The device address of [p, p + n) if it lies in one lbn_host_alloc allocation, else nullptr.
static void *pinned_alias(const void *p, size_t n) {
    const uintptr_t a = (uintptr_t)p;
    uintptr_t base;
    {
        std::lock_guard<std::mutex> lk(g_pin_m);
        auto it = g_pins.upper_bound(a);
        if (it == g_pins.begin()) return nullptr;
        --it;
        if (a + n > it->first + it->second) return nullptr;
        base = it->first;
    }
    void *dev = nullptr;
    if (hipHostGetDevicePointer(&dev, (void *)base, 0) != hipSuccess || !dev) {
        (void)hipGetLastError();
        return nullptr;
    }
    return (uint8_t *)dev + (a - base);
}

// ----------------------------------------------------------------------------- modulus, commit

static u128 mul_mod_small(u128 a, uint32_t b, u128 m) { return (a * b) % m; }

// Primes: the largest p = 1 (mod 1944) with 6p < 2^32, each with a primitive 1944-th root.
static const PrimeInfo CRT_PRIMES[MAX_PRIMES] = {{715809961u, 0, 0, 181294596u},
                                                 {715808017u, 0, 0, 501708134u},
                                                 {715806073u, 0, 0, 148596466u},
                                                 {715757473u, 0, 0, 282162646u},
                                                 {715722481u, 0, 0, 150155317u}};

// CRT primes for the exact commitment of `re` ring elements per column under modulus q (0 if
// the range cannot be covered). With the matrix centred (|a| <= Q/2) and at most 512 set bits
// per element, a coefficient of one product is at most 3 * 512 * Q/2 = 768 Q, so |coefficient|
// <= re * 768 * Q; the CRT range must exceed twice that. Counted in doubles (M can pass 2^128).
static int crt_prime_count(u128 q, size_t re) {
    const double need_bits = std::log2(2.0 * (double)re * 768.0) + std::log2((double)q) + 0.01;
    int np = 0;
    double have = 0;
    while (np < MAX_PRIMES && have <= need_bits) have += std::log2((double)CRT_PRIMES[np++].p);
    return have > need_bits ? np : 0;
}

static int modulus_width(u128 q) {
    int bits = 0;
    for (u128 x = q - 1; x; x >>= 1) bits++;
    return (bits + 7) / 8;
}

// Ring-element ranges of the MAC over `groups` groups of TC columns: about 1024 blocks per prime
// (groups x splits), at most 128 slices (the finish kernel's slice sum stays below 2^32 p).
// Device bytes of the buffers that scale with the CRT prime count and the coefficient width.
constexpr int MAX_SLICES = 128;
static int mac_m_per(size_t re, int groups = COLUMNS / TC) {
    const size_t want = std::max<size_t>(1, 1024 / (size_t)groups);
    const int msplit = (int)std::max<size_t>(1, std::min<size_t>({re, want, (size_t)MAX_SLICES}));
    return (int)((re + msplit - 1) / msplit);
}
static int mac_splits(size_t re, int groups = COLUMNS / TC) {
    const int m_per = mac_m_per(re, groups);
    return (int)((re + m_per - 1) / m_per);
}

static size_t commit_buffer_bytes(int np, int width, size_t re) {
    const size_t clen = (size_t)COLUMNS * N * width;
    return (size_t)np * mac_splits(re) * COLUMNS * N * 4 + (clen + 64) + (clen / 1024 * 32 + 64);
}

// The matrix-core commitment serves moduli below 2^39: a centred entry (|e| <= Q/2 < 2^38) is five
// balanced signed bytes and |C| <= 2^20 Q fits an int64. Larger moduli use the CRT transform.
static bool toep_modulus(u128 q) { return q < ((u128)1 << 39); }
// Ring-element splits of the GEMM (42 blocks each): 6 with the table resident (252 blocks, one
// round over 256 CUs), 24 with a deferred upload (four launches of 6 splits, one per copied range).
static int toep_splits(size_t re, bool deferred) { return (int)std::min<size_t>(re, deferred ? 24 : 6); }
// Split counts chosen by the autotuner (lbn_autotune / lbn_tune_set) per (device, log_vars); a shape
// without an entry uses toep_splits.
static std::mutex tune_mu;
static std::map<std::pair<int, int>, int> tune_splits;
static size_t toep_ops_bytes(size_t re) { return re * TW_ELEM_U32 * 4; }
static size_t toep_part_bytes(size_t re) { return (size_t)toep_splits(re, true) * TP_ROWS * COLUMNS * 8; }

// The GEMM's matrix operands depend only on (device, Q, ring elements): built once, shared by the
// contexts that use them (weak entries: freed with the last context holding them).
struct ToepOps {
    int device = 0;
    uint32_t *ptr = nullptr;
    ~ToepOps() {
        if (ptr) {
            int cur = 0;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(device);
            (void)hipFree(ptr);
            (void)hipSetDevice(cur);
        }
    }
};
static std::mutex toep_mu;
static std::map<std::tuple<int, u128, size_t>, std::weak_ptr<ToepOps>> toep_cache;

extern "C" size_t lbn_ctx_footprint(int log_vars, uint64_t q_lo, uint64_t q_hi) {
    Shape s;
    if (!Shape::of(log_vars, s)) return 0;
    const u128 q = ((u128)q_hi << 64) | q_lo;
    if (q < (u128)34816965049ull || q >= ((u128)1 << 96)) return 0;
    const bool toep = toep_modulus(q);
    const int np = toep ? 0 : crt_prime_count(q, s.ring_elements);
    if (!toep && !np) return 0;
    const size_t n = s.n_words;
    size_t b = n * 16 + n * 16 + n / 2 * 16 + n / 2 * 24 + n / 4 * 24;  // table, E0, E1, P0, P1
    b += 3584 * 4 + ps::WORDS * 4;  // T, pset
    b += (4096 + std::max<size_t>((size_t)1 << (s.log_packed - 12), 128)) * 16;  // EL, EH
    b += s.column_len * 24 + 256 * 24 + (size_t)COLUMNS * 128 * 24 + (size_t)COLUMNS * 24;
    b += 4 + (2 * 162 + 1) * sizeof(int) + (size_t)COLUMNS * 32 * sizeof(uint16_t);
    b += (size_t)MAX_PRIMES * tab::SIZE * 4;
    b += (size_t)np * s.ring_elements * N * 4;  // ahat
    b += s.ring_elements * N * 16;              // matrix staging while the modulus is set
    b += commit_buffer_bytes(np, modulus_width(q), s.ring_elements);
    if (toep) b += toep_ops_bytes(s.ring_elements) + toep_part_bytes(s.ring_elements);  // counted per context
    return b;
}

// Elements first .. first + count - 1 of the matrix expansion for (q, elements) into out (device,
// 2 x u64 per coefficient), on `stream`; waits for it. Indices past `elements` continue the same
// keyed stream family (a key of several ring rows takes row r's element m at r * elements + m).
static hipError_t ajtai_device(hipStream_t stream, u128 q, size_t elements, size_t first, size_t count,
                               uint64_t *out, bool *finished) {
    int bits = 0;
    for (u128 x = q - 1; x; x >>= 1) bits++;
    const int width = (bits + 7) / 8;
    const u128 mask = ((u128)1 << bits) - 1;
    uint8_t key[32];
    memset(key, 0x5a, 32);
    Blake3 h = Blake3::keyed(key);
    static const char DOMAIN[] = "labinius-fast/ajtai-matrix/v1";
    h.update(DOMAIN, sizeof DOMAIN - 1);
    uint8_t qb[16];
    for (int i = 0; i < 16; i++) qb[i] = (uint8_t)(q >> (8 * i));
    h.update(qb, 16);
    h.update_u64(elements);
    h.update_u64(0);  // the index bytes (53..60) are filled in per element
    const Blake3::Root r = h.root();
    static_assert(sizeof DOMAIN - 1 + 16 + 8 == 53, "the element index sits at bytes 53..60 of the block");
    if (r.len != 61) return hipErrorInvalidValue;
    B3Root root{};
    memcpy(root.cv, r.cv, 32);
    memcpy(root.m, r.m, 64);
    root.len = r.len;
    root.flags = r.flags | Blake3::ROOT;
    uint32_t *fail = nullptr;
    hipError_t e = hipMalloc(&fail, 4);
    if (e == hipSuccess) e = hipMemsetAsync(fail, 0, 4, stream);
    if (e == hipSuccess && count) {
        k_ajtai<<<(unsigned)((count + AJ_WAVES - 1) / AJ_WAVES), 64 * AJ_WAVES, 0, stream>>>(
            root, (uint64_t)first, (int)count, width, (uint64_t)mask, (uint32_t)(mask >> 64), (uint64_t)q,
            (uint32_t)(q >> 64), out, fail);
        e = hipGetLastError();
    }
    uint32_t flag = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, fail, 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (fail) (void)hipFree(fail);
    *finished = flag == 0;
    return e;
}

extern "C" int lbn_ajtai_elements(int device, uint64_t q_lo, uint64_t q_hi, size_t elements, size_t first,
                                  size_t count, uint64_t *out) {
    const u128 q = ((u128)q_hi << 64) | q_lo;
    if (!out || q < 2 || q >= ((u128)1 << 96) || count > ((size_t)1 << 20)) return LBN_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return LBN_E_DEVICE;
    uint64_t *d = nullptr;
    hipStream_t st = nullptr;
    bool done = false;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&d, count * N * 16 + 16);
    if (e == hipSuccess) e = ajtai_device(st, q, elements, first, count, d, &done);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, count * N * 16, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d) (void)hipFree(d);
    if (st) (void)hipStreamDestroy(st);
    return e != hipSuccess ? LBN_E_DEVICE : done ? LBN_OK : LBN_E_DEVICE;
}

// The first np CRT primes with their Montgomery constants, their transform tables (all,
// MAX_PRIMES * tab::SIZE words) and their 18-th roots.
static bool setup_primes(int np, PrimeInfo *primes, Roots18 *roots, std::vector<uint32_t> &all) {
    all.assign((size_t)MAX_PRIMES * tab::SIZE, 0);
    for (int k = 0; k < np; k++) {
        PrimeInfo pi = CRT_PRIMES[k];
        uint32_t inv = 1;  // p^-1 mod 2^32 by Newton
        for (int it = 0; it < 5; it++) inv *= 2 - pi.p * inv;
        pi.pinv_neg = (uint32_t)(0u - inv);
        pi.r2 = (uint32_t)(((u128)1 << 64) % pi.p);
        primes[k] = pi;
        std::vector<uint32_t> t;
        if (!prime_tables(pi, t)) return false;
        std::copy(t.begin(), t.end(), all.begin() + (size_t)k * tab::SIZE);
        for (int j = 0; j < 9; j++) {
            roots[k].f[j] = all[(size_t)k * tab::SIZE + tab::R18F + j];
            roots[k].i[j] = all[(size_t)k * tab::SIZE + tab::R18I + j];
        }
    }
    return true;
}

static int alloc_commit_buffers(lbn_ctx *ctx, int np, int width) {
    DeviceState *d = ctx->dev;
    if (d->alloc_primes >= np && d->alloc_width >= width) return LBN_OK;
    HIPCHK(ctx, hipStreamSynchronize(d->copy_stream));
    for (void *p : {(void *)d->cacc, (void *)d->cbytes, (void *)d->cvs})
        if (p) HIPCHK(ctx, hipFree(p));
    if (d->hstage) HIPCHK(ctx, hipHostFree(d->hstage));
    if (d->hcvs) HIPCHK(ctx, hipHostFree(d->hcvs));
    d->cacc = nullptr, d->cbytes = nullptr, d->cvs = nullptr, d->hstage = nullptr, d->hcvs = nullptr;
    d->hcvs_dev = nullptr;
    d->alloc_primes = d->alloc_width = 0;
    const size_t clen = (size_t)COLUMNS * N * width;
    d->splits = mac_splits(ctx->shape.ring_elements);
    if (np) HIPCHK(ctx, hipMalloc(&d->cacc, (size_t)np * d->splits * COLUMNS * N * 4));
    HIPCHK(ctx, hipMalloc(&d->cbytes, clen + 64));  // slack: the chunk hashing reads one word past
    HIPCHK(ctx, hipMalloc(&d->cvs, clen / 1024 * 32 + 64));
    HIPCHK(ctx, hipHostMalloc((void **)&d->hstage, clen));
    HIPCHK(ctx, hipHostMalloc((void **)&d->hcvs, clen / 1024 * 32 + 64));
    HIPCHK(ctx, hipHostGetDevicePointer((void **)&d->hcvs_dev, d->hcvs, 0));
    d->alloc_primes = np;
    d->alloc_width = width;
    return LBN_OK;
}

// The matrix-core commitment's operands for Q (from the cache, or built here from the public
// matrix) and the context's partial-sum slices.
static int setup_toep(lbn_ctx *ctx, u128 q) {
    DeviceState *d = ctx->dev;
    const size_t re = ctx->shape.ring_elements;
    if (!d->tpart) {
        HIPCHK(ctx, hipMalloc(&d->tpart, toep_part_bytes(re)));
    }
    const auto key = std::make_tuple(ctx->device, q, re);
    std::lock_guard<std::mutex> lk(toep_mu);
    if (auto sp = toep_cache[key].lock()) {
        d->toep = sp;
        return LBN_OK;
    }
    d->toep.reset();
    auto ops = std::make_shared<ToepOps>();
    ops->device = ctx->device;
    HIPCHK(ctx, hipMalloc(&ops->ptr, toep_ops_bytes(re)));
    HIPCHK(ctx, hipMalloc(&d->amat, re * N * 16));
    {
        bool done = false;
        HIPCHK(ctx, ajtai_device(d->stream, q, re, 0, re, d->amat, &done));
        if (!done) return ctx->fail(LBN_E_DEVICE, "matrix expansion did not finish");
    }
    k_toep_build<<<(unsigned)re, 256, 0, d->stream>>>(d->amat, (int)re, (uint64_t)q, ops->ptr);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(d->stream));
    HIPCHK(ctx, hipFree(d->amat));
    d->amat = nullptr;
    for (auto it = toep_cache.begin(); it != toep_cache.end();)
        it = it->second.expired() ? toep_cache.erase(it) : std::next(it);
    toep_cache[key] = ops;
    d->toep = ops;
    return LBN_OK;
}

extern "C" int lbn_set_modulus(lbn_ctx *ctx, uint64_t q_lo, uint64_t q_hi) {
    if (!ctx) return LBN_E_ARG;
    const u128 q = ((u128)q_hi << 64) | q_lo;
    if (q < (u128)34816965049ull || q >= ((u128)1 << 96))
        return ctx->fail(LBN_E_MODULUS, "commitment modulus outside [34816965049, 2^96)");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    DeviceState *d = ctx->dev;
    const Shape &s = ctx->shape;
    const bool toep = toep_modulus(q);
    const int np = toep ? 0 : crt_prime_count(q, s.ring_elements);
    if (!toep && !np) return ctx->fail(LBN_E_MODULUS, "modulus too large for the CRT range");
    ctx->modulus_set = false;
    {
        const int rc = alloc_commit_buffers(ctx, np, modulus_width(q));
        if (rc != LBN_OK) return rc;
    }
    ctx->nprimes = np;
    if (toep) {
        const int rc = setup_toep(ctx, q);
        if (rc != LBN_OK) return rc;
        ctx->q = q;
        ctx->width = modulus_width(q);
        ctx->modulus_set = true;
        return LBN_OK;
    }
    d->toep.reset();
    std::vector<uint32_t> all;
    if (!setup_primes(np, ctx->primes, d->roots, all)) return ctx->fail(LBN_E_ARG, "NTT table construction failed");
    HIPCHK(ctx, hipMemcpyAsync(d->ntt_tabs, all.data(), all.size() * 4, hipMemcpyHostToDevice, d->stream));
    // The public matrix (rq.rs:114-144), expanded, reduced and transformed on the device.
    HIPCHK(ctx, hipMalloc(&d->amat, s.ring_elements * N * 16));
    {
        bool done = false;
        HIPCHK(ctx, ajtai_device(d->stream, q, s.ring_elements, 0, s.ring_elements, d->amat, &done));
        if (!done) return ctx->fail(LBN_E_DEVICE, "matrix expansion did not finish");
    }
    const size_t need = (size_t)np * s.ring_elements * N;
    if (d->ahat_elems < need) {
        if (d->ahat) HIPCHK(ctx, hipFree(d->ahat));
        HIPCHK(ctx, hipMalloc(&d->ahat, need * 4));
        d->ahat_elems = need;
    }
    for (int k = 0; k < np; k++) {
        PC pc{ctx->primes[k].p, ctx->primes[k].pinv_neg};
        const Roots18 R = d->roots[k];
        k_ahat<<<(unsigned)s.ring_elements, 64, 0, d->stream>>>(d->amat, d->ntt_tabs + (size_t)k * tab::SIZE, pc, R,
                                                                ctx->primes[k].r2, q_lo, q_hi,
                                                                d->ahat + (size_t)k * s.ring_elements * N, 0, 0);
        HIPCHK(ctx, hipGetLastError());
    }
    HIPCHK(ctx, hipStreamSynchronize(d->stream));
    HIPCHK(ctx, hipFree(d->amat));  // only the transform needs the matrix itself
    d->amat = nullptr;
    ctx->q = q;
    ctx->width = modulus_width(q);
    ctx->modulus_set = true;
    (void)mul_mod_small;
    return LBN_OK;
}

extern "C" int lbn_commit_primes(const lbn_ctx *ctx) { return ctx ? ctx->nprimes : 0; }
extern "C" int lbn_commit_planes(const lbn_ctx *ctx) {
    return ctx && ctx->modulus_set && ctx->dev && ctx->dev->toep ? TP_PLANES : 0;
}

static uint32_t powmod32(uint64_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1;
    a %= p;
    while (e) {
        if (e & 1) r = r * a % p;
        a = a * a % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

// Host memcpy split over the worker pool (1 MiB pieces).
static void copy_parallel(uint8_t *dst, const uint8_t *src, size_t n) {
    const size_t piece = (size_t)1 << 20, pieces = (n + piece - 1) / piece;
    parallel_for(pieces, [&](size_t i) {
        const size_t lo = i * piece;
        memcpy(dst + lo, src + lo, std::min(piece, n - lo));
    });
}

extern "C" int lbn_commit(lbn_ctx *ctx, uint8_t *out, size_t out_len) {
    if (!ctx || !out) return LBN_E_ARG;
    if (!ctx->table_loaded || !ctx->modulus_set) return ctx->fail(LBN_E_STATE, "table and modulus must be loaded");
    const size_t clen_bytes = (size_t)COLUMNS * N * ctx->width;
    if (out_len != clen_bytes) return ctx->fail(LBN_E_ARG, "commitment buffer has the wrong length");
    const int rc = dev_commit_launch(ctx);
    if (rc != LBN_OK) return rc;
    DeviceState *d = ctx->dev;
    HIPCHK(ctx, hipMemcpyAsync(out, d->cbytes, clen_bytes, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(ctx, hipStreamSynchronize(d->stream));
    tcollect(ctx, T_MAC);
    tcollect(ctx, T_FINISH);
    return LBN_OK;
}

// Launch-shape autotuner: the matrix-core commitment (MAC + finish) timed on the context's loaded table
// for each split count, the fastest recorded for (device, log_vars). The result is a launch shape
// only: every split count gives the same commitment bytes.
extern "C" int lbn_autotune(lbn_ctx *ctx, int reps, lbn_tune_result *out) {
    if (!ctx || !out || reps < 1 || reps > 100) return LBN_E_ARG;
    if (!ctx->table_loaded || !ctx->modulus_set) return ctx->fail(LBN_E_STATE, "table and modulus must be loaded");
    DeviceState *d = ctx->dev;
    if (!d->toep) return ctx->fail(LBN_E_STATE, "the modulus commits through the CRT transform (nothing to tune)");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    {
        const int rc = table_ready(ctx);  // a deferred upload completes first
        if (rc != LBN_OK) return rc;
    }
    *out = lbn_tune_result{};
    out->log_vars = ctx->shape.log_vars;
    const int cap = toep_splits(ctx->shape.ring_elements, true);
    static const int cand[8] = {1, 2, 3, 4, 6, 8, 12, 24};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(ctx, hipEventCreate(&e0));
    HIPCHK(ctx, hipEventCreate(&e1));
    int rc = LBN_OK;
    double best_ms = 1e30;
    for (int c = 0; c < 8 && rc == LBN_OK; c++) {
        if (cand[c] > cap) break;
        d->splits_try = cand[c];
        std::vector<float> ts;
        for (int r = 0; r <= reps && rc == LBN_OK; r++) {  // the first launch is not timed
            if (hipEventRecord(e0, d->stream) != hipSuccess) rc = ctx->fail(LBN_E_DEVICE, "event record failed");
            if (rc == LBN_OK) rc = dev_commit_launch(ctx);
            if (rc == LBN_OK && hipEventRecord(e1, d->stream) != hipSuccess) rc = ctx->fail(LBN_E_DEVICE, "event record failed");
            float ms = 0;
            if (rc == LBN_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = ctx->fail(LBN_E_DEVICE, "event timing failed");
            if (rc == LBN_OK && r > 0) ts.push_back(ms);
        }
        if (rc != LBN_OK) break;
        std::sort(ts.begin(), ts.end());
        const double med = ts[ts.size() / 2];
        out->splits[out->n] = cand[c];
        out->ms[out->n] = med;
        out->n++;
        if (med < best_ms) best_ms = med, out->best = cand[c];
    }
    d->splits_try = 0;
    (void)hipStreamSynchronize(d->stream);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != LBN_OK) return rc;
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_splits[{ctx->device, ctx->shape.log_vars}] = out->best;
    return LBN_OK;
}

extern "C" int lbn_tune_set(int device, int log_vars, int mac_splits) {
    Shape s;
    if (!Shape::of(log_vars, s) || mac_splits < 0 || mac_splits > 24) return LBN_E_ARG;
    std::lock_guard<std::mutex> lk(tune_mu);
    if (mac_splits == 0) tune_splits.erase({device, log_vars});
    else tune_splits[{device, log_vars}] = mac_splits;
    return LBN_OK;
}

// The commitment's hashing in three steps: the arguments checked and the commitment's copy to the
// host queued (stage), the chunk and subtree kernels (launch, or one launch for several proofs),
// the subtree roots and the bytes around the hashed chunks taken from pinned memory (collect).

static int cvs_stage(lbn_ctx *ctx, size_t off, uint64_t first, size_t count, const Blake3::Subtrees &trees,
                     const uint32_t key[8], uint32_t flags, uint8_t *dst, size_t head, size_t tail, CvJob &job) {
    DeviceState *d = ctx->dev;
    const size_t clen = (size_t)COLUMNS * N * ctx->width;
    if (off + count * 1024 > clen || head > clen || tail > clen)
        return ctx->fail(LBN_E_ARG, "commitment chunk range out of bounds");
    if (trees.size() > MAX_CV_TREES) return ctx->fail(LBN_E_ARG, "too many commitment subtrees");
    CvTrees &tr = job.tr;
    tr = CvTrees{};
    tr.n = (int)trees.size();
    {
        uint64_t at = first;  // the subtrees tile [first, first + count) in order
        for (int b = 0; b < tr.n; b++) {
            if (trees[b].first != at || trees[b].second < 0 || trees[b].second > 30)
                return ctx->fail(LBN_E_ARG, "commitment subtrees do not tile the chunks");
            tr.start[b] = (uint32_t)(at - first);
            tr.level[b] = trees[b].second;
            at += 1ull << trees[b].second;
        }
        if (at != first + count) return ctx->fail(LBN_E_ARG, "commitment subtrees do not tile the chunks");
    }
    if (head > 1024 || clen - tail > 1024) return ctx->fail(LBN_E_ARG, "commitment head or tail above one chunk");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // the staging copy runs on its own stream once the commitment is final, into the caller's
    // buffer when it is one of ours (pinned); every argument is checked before it is queued
    d->staged_direct = pinned_alias(dst, clen) != nullptr;
    HIPCHK(ctx, hipStreamWaitEvent(d->copy_stream, d->ev_commit, 0));
    HIPCHK(ctx, hipMemcpyAsync(d->staged_direct ? dst : d->hstage, d->cbytes, clen, hipMemcpyDeviceToHost,
                               d->copy_stream));
    HIPCHK(ctx, hipEventRecord(d->ev_staged, d->copy_stream));
    job.off = off, job.first = first, job.count = count, job.head = head, job.tail = tail;
    for (int i = 0; i < 8; i++) job.key.k[i] = key[i];
    job.key.flags = flags;
    return LBN_OK;
}

static unsigned cvs_windows(const CvJob &job) {
    return job.count ? (unsigned)((job.first + job.count - 1) / 64 - job.first / 64 + 1) : 0u;
}

static int cvs_launch(lbn_ctx *ctx, const CvJob &job) {
    DeviceState *d = ctx->dev;
    const size_t clen = (size_t)COLUMNS * N * ctx->width;
    if (job.count) {
        k_chunk_cvs<<<cvs_windows(job), 64, 0, d->stream>>>(d->cbytes, job.off, job.first, job.count, job.key, d->cvs);
        HIPCHK(ctx, hipGetLastError());
        // the roots, head and tail go straight to pinned host memory
        const int tail_len = (int)(clen - job.tail);
        uint8_t *ho = d->hbox_dev + mb::HEAD, *to = d->hbox_dev + mb::TAIL;
        if (job.count * 32 <= CV_TREE_LDS_MAX)
            k_cv_tree<true><<<1, 1024, job.count * 32, d->stream>>>(d->cvs, (int)job.count, job.tr, job.key,
                                                                    CV_WIN_LEVELS, d->hcvs_dev, d->cbytes,
                                                                    (int)job.head, job.tail, tail_len, ho, to);
        else
            k_cv_tree<false><<<1, 1024, 0, d->stream>>>(d->cvs, (int)job.count, job.tr, job.key, CV_WIN_LEVELS,
                                                        d->hcvs_dev, d->cbytes, (int)job.head, job.tail, tail_len,
                                                        ho, to);
        HIPCHK(ctx, hipGetLastError());
    } else {
        if (job.head)
            HIPCHK(ctx, hipMemcpyAsync(d->hbox + mb::HEAD, d->cbytes, job.head, hipMemcpyDeviceToHost, d->stream));
        if (job.tail < clen)
            HIPCHK(ctx, hipMemcpyAsync(d->hbox + mb::TAIL, d->cbytes + job.tail, clen - job.tail,
                                       hipMemcpyDeviceToHost, d->stream));
    }
    return LBN_OK;
}

static void cvs_collect(lbn_ctx *ctx, const CvJob &job, uint32_t *cvs, uint8_t *dst) {
    DeviceState *d = ctx->dev;
    const size_t clen = (size_t)COLUMNS * N * ctx->width;
    memcpy(cvs, d->hcvs, (size_t)job.tr.n * 32);
    memcpy(dst, d->hbox + mb::HEAD, job.head);
    memcpy(dst + job.tail, d->hbox + mb::TAIL, clen - job.tail);
    tcollect(ctx, T_MAC);
    tcollect(ctx, T_FINISH);
}

int lbn::dev_commit_cvs(lbn_ctx *ctx, size_t off, uint64_t first, size_t count, const Blake3::Subtrees &trees,
                        const uint32_t key[8], uint32_t flags, uint32_t *cvs, uint8_t *dst, size_t head, size_t tail) {
    CvJob job;
    {
        const int rc = cvs_stage(ctx, off, first, count, trees, key, flags, dst, head, tail, job);
        if (rc != LBN_OK) return rc;
    }
    {
        const int rc = cvs_launch(ctx, job);
        if (rc != LBN_OK) return rc;
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->dev->stream));
    cvs_collect(ctx, job, cvs, dst);
    return LBN_OK;
}

int lbn::dev_cvs_stage(lbn_ctx *ctx, const Blake3::Subtrees &trees, size_t off, uint64_t first, size_t count,
                       const uint32_t key[8], uint32_t flags, uint8_t *dst, size_t head, size_t tail) {
    DeviceState *d = ctx->dev;
    if (!d->cvjob) d->cvjob.reset(new CvJob());
    return cvs_stage(ctx, off, first, count, trees, key, flags, dst, head, tail, *d->cvjob);
}

void lbn::dev_cvs_collect(lbn_ctx *ctx, uint32_t *cvs, uint8_t *dst) { cvs_collect(ctx, *ctx->dev->cvjob, cvs, dst); }

void lbn::dev_quiesce(lbn_ctx *ctx) {
    DeviceState *d = ctx->dev;
    if (!d) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(d->copy_stream);
    (void)hipStreamSynchronize(d->stream);
}

int lbn::dev_commit_bytes(lbn_ctx *ctx, uint8_t *dst) {
    DeviceState *d = ctx->dev;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipEventSynchronize(d->ev_staged));
    if (!d->staged_direct) copy_parallel(dst, d->hstage, (size_t)COLUMNS * N * ctx->width);
    return LBN_OK;
}

// Arguments of the finish kernel: the primes, their Garner constants and the range M.
static CrtArg make_crt(const PrimeInfo *primes, const Roots18 *roots, int np, u128 q, int width, int splits,
                       int cols) {
    CrtArg crt{};
    crt.nprimes = np;
    crt.width = width;
    crt.splits = splits;
    crt.cols = cols;
    crt.q_lo = (uint64_t)q;
    crt.q_hi = (uint64_t)(q >> 64);
    uint64_t M[3] = {1, 0, 0};  // product of the primes, 192 bits
    for (int i = 0; i < np; i++) {
        crt.p[i] = primes[i].p;
        crt.pc[i] = PC{primes[i].p, primes[i].pinv_neg};
        crt.r2[i] = primes[i].r2;
        crt.roots[i] = roots[i];
        u128 t = (u128)M[0] * crt.p[i];
        M[0] = (uint64_t)t;
        t = (u128)M[1] * crt.p[i] + (uint64_t)(t >> 64);
        M[1] = (uint64_t)t;
        M[2] = M[2] * crt.p[i] + (uint64_t)(t >> 64);
        for (int j = 0; j < i; j++)
            crt.inv[i][j] =
                (uint32_t)(((uint64_t)powmod32(primes[j].p % primes[i].p, primes[i].p - 2, primes[i].p) << 32) %
                           primes[i].p);
    }
    for (int w = 0; w < 3; w++) crt.m[w] = M[w];
    crt.half_m[0] = (M[0] >> 1) | (M[1] << 63);
    crt.half_m[1] = (M[1] >> 1) | (M[2] << 63);
    crt.half_m[2] = M[2] >> 1;
    return crt;
}

// The MAC's element splits for this commitment: 24 with a deferred upload (four launches of 6), else
// the autotuner's choice for the shape (or the one being timed), else toep_splits.
static int commit_splits(lbn_ctx *ctx) {
    const size_t re = ctx->shape.ring_elements;
    const int cap = toep_splits(re, true);  // the partial sums hold this many
    if (ctx->dev->pend) return cap;
    if (ctx->dev->splits_try > 0) return std::min(ctx->dev->splits_try, cap);
    std::lock_guard<std::mutex> lk(tune_mu);
    const auto it = tune_splits.find({ctx->device, ctx->shape.log_vars});
    return it != tune_splits.end() ? std::max(1, std::min(it->second, cap)) : toep_splits(re, false);
}

int lbn::dev_commit_launch(lbn_ctx *ctx) {
    if (!ctx->table_loaded || !ctx->modulus_set) return ctx->fail(LBN_E_STATE, "table and modulus must be loaded");
    const Shape &s = ctx->shape;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // a staging copy of an earlier commitment still reading cbytes finishes first
    HIPCHK(ctx, hipStreamWaitEvent(ctx->dev->stream, ctx->dev->ev_staged, 0));
    DeviceState *d = ctx->dev;
    const int re = (int)s.ring_elements;
    if (d->toep) {
        const int sp = commit_splits(ctx);
        // a deferred upload goes in (at most) four ring-element ranges, each range's MAC after its copy
        const int groups = d->pend ? std::min(4, sp) : 1;
        tstart(ctx, T_MAC);
        for (int g = 0; g < groups; g++) {
            const int s0 = sp * g / groups, s1 = sp * (g + 1) / groups;
            if (d->pend) {
                const int rc = upload_range(ctx, g, (size_t)((int64_t)s0 * re / sp), (size_t)((int64_t)s1 * re / sp));
                if (rc != LBN_OK) return rc;
            }
            k_toep_mac<<<2 * T2_GROUPS * (s1 - s0), 64 * T2_WAVES, 0, d->stream>>>(
                d->toep->ptr, d->table, s.column_len, re, sp, s0, s1 - s0, d->tpart);
            HIPCHK(ctx, hipGetLastError());
        }
        d->pend = nullptr;
        tstop(ctx, T_MAC);
        tstart(ctx, T_FINISH);
        k_toep_finish<<<dim3(COLUMNS / 64, TP_TILES), 256, 0, d->stream>>>(d->tpart, sp, (uint64_t)ctx->q, ctx->width,
                                                                          d->cbytes);
        tstop(ctx, T_FINISH);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(d->ev_commit, d->stream));
        return LBN_OK;
    }
    const int np = ctx->nprimes;
    const int m_per = mac_m_per(re), msplit = mac_splits(re);
    if (msplit != d->splits) return ctx->fail(LBN_E_STATE, "MAC slices not allocated for this shape");
    {
        const int rc = table_ready(ctx);
        if (rc != LBN_OK) return rc;
    }
    tstart(ctx, T_MAC);
    for (int k = 0; k < np; k++) {
        PC pc{ctx->primes[k].p, ctx->primes[k].pinv_neg};
        dim3 grid(COLUMNS / TC, msplit);
        k_commit_mac<<<grid, MAC_THREADS, 0, d->stream>>>((const uint32_t *)d->table, s.column_len, m_per, re,
                                                           d->ahat + (size_t)k * re * N,
                                                           d->ntt_tabs + (size_t)k * tab::SIZE, pc,
                                                           (uint32_t)((1ull << 32) % pc.p), ctx->primes[k].r2,
                                                           d->roots[k],
                                                           d->cacc + (size_t)k * msplit * COLUMNS * N);
        HIPCHK(ctx, hipGetLastError());
    }
    tstop(ctx, T_MAC);
    const CrtArg crt = make_crt(ctx->primes, d->roots, np, ctx->q, ctx->width, msplit, COLUMNS);
    tstart(ctx, T_FINISH);
    k_commit_finish<<<COLUMNS, FIN_THREADS, 0, d->stream>>>(d->cacc, d->ntt_tabs, crt, d->cbytes);
    tstop(ctx, T_FINISH);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(d->ev_commit, d->stream));
    return LBN_OK;
}

// ----------------------------------------------------------------------------- partials, switch

// partials (crossfield.rs:152-165) through the prefix sets: G = eq(r_hi[6..]) expanded into E[0]
// (its high coordinates, at most 6, times the low ones, at most 12), then the sets M_u and the
// partials in one pass over the table (k_prefix_sums). M_u and G stay for the switch.
// The partials in three steps: the eq tables of r_hi on the host and their upload (prepare), the eq
// expansion and prefix sums (launch, or one launch each for several proofs), the 128 partials read
// from the mailbox (collect).
struct PartShape {
    int lp, S, lo, nh, chunk_log;
};
static PartShape part_shape(const Shape &s) {
    PartShape p;
    p.lp = s.log_packed, p.S = p.lp - 6, p.lo = std::min(p.S, 12), p.nh = p.S - p.lo;
    p.chunk_log = std::min(p.S, 14);
    return p;
}

int lbn::dev_partials_prepare(lbn_ctx *ctx, const uint64_t *r_hi, bool upload) {
    if (!ctx->table_loaded) return ctx->fail(LBN_E_STATE, "no table loaded");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    {
        const int rc_ = table_ready(ctx);  // a deferred upload completes first
        if (rc_ != LBN_OK) return rc_;
    }
    DeviceState *d = ctx->dev;
    const PartShape p = part_shape(ctx->shape);
    ctx->r_hi.assign((const B128 *)r_hi, (const B128 *)r_hi + p.lp);
    B128 *EL = (B128 *)(d->hbox + mb::EL), *EH = (B128 *)(d->hbox + mb::EH);
    eq_expand_b128(ctx->r_hi.data() + p.lp - p.lo, p.lo, EL);
    eq_expand_b128(ctx->r_hi.data() + 6, p.nh, EH);
    eq_expand_b128(ctx->r_hi.data(), 6, EH + 64);  // eq6, after the at most 64 high values
    static_assert(mb::EH == mb::EL + 4096 * 16, "EL and EH adjacent, as on the device");
    if (upload)
        HIPCHK(ctx, hipMemcpyAsync(d->EL, d->hbox + mb::EL, (4096 + 128) * 16, hipMemcpyHostToDevice, d->stream));
    d->msg_clean = false;
    return LBN_OK;
}

void lbn::dev_partials_collect(lbn_ctx *ctx, uint64_t *partials) {
    const uint32_t *out = (const uint32_t *)(ctx->dev->hbox + mb::PART);
    for (int k = 0; k < 128; k++) {
        partials[2 * k] = (uint64_t)out[4 * k] | ((uint64_t)out[4 * k + 1] << 32);
        partials[2 * k + 1] = (uint64_t)out[4 * k + 2] | ((uint64_t)out[4 * k + 3] << 32);
    }
    ctx->partials_done = true;
    ctx->round = -1;
}

extern "C" int lbn_partials(lbn_ctx *ctx, const uint64_t *r_hi, uint64_t *partials) {
    if (!ctx || !r_hi || !partials) return LBN_E_ARG;
    {
        const int rc_ = dev_partials_prepare(ctx, r_hi, true);
        if (rc_ != LBN_OK) return rc_;
    }
    DeviceState *d = ctx->dev;
    const PartShape p = part_shape(ctx->shape);
    tstart(ctx, T_EQ);
    // also zeroes the sets and the partials' accumulator
    k_eq_expand<<<1u << p.nh, 256, 0, d->stream>>>(d->EH, d->EL, p.lo, d->E[0], d->pset, (int)ps::Q);
    tstop(ctx, T_EQ);
    HIPCHK(ctx, hipGetLastError());
    tstart(ctx, T_PARTIALS);
    k_prefix_sums<<<1u << (p.lp - p.chunk_log), 256, 0, d->stream>>>(d->table, d->E[0], p.S, p.chunk_log, d->EH + 64,
                                                                     d->pset + ps::M, d->pset + ps::PART);
    tstop(ctx, T_PARTIALS);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(d->hbox + mb::PART, d->pset + ps::PART, 512 * 4, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(ctx, hipStreamSynchronize(d->stream));
    tcollect(ctx, T_EQ);
    tcollect(ctx, T_PARTIALS);
    dev_partials_collect(ctx, partials);
    return LBN_OK;
}

// ----------------------------------------------------------------------------- checker arithmetic
// The two sides of the checker's identity A v = sum_j c_j(-X^4) C_j (verifier/proof_verify/src/
// rq.rs:188-256), as exact integer coefficients of the products before the reduction into R_Q
// (indices e < 1295): each side is returned as three limb sums raw[e][k], k < 3, with
// raw[e] = sum_k 2^(32 k) raw[e][k]; the host reduces them modulo Q and folds X^648 = X^324 - 1.
constexpr int VT_RAW = 2 * N - 1;  // 1295
constexpr int VT_M_PER = 8;        // ring elements per block of the A v kernel (512 blocks at 4096)

// A v: block b takes ring elements [8 b, 8 b + 8); matrix entries in [0, Q), Q < 2^96, as three
// 32-bit limbs, v as int16. Thread t keeps the sums of e = t + 256 q (q < 6); a sum has at most
// 8 x 648 terms below 2^32 x 2^15 in size, well inside int64.
__global__ __launch_bounds__(256) void k_vt_atv(const uint64_t *amat, const int16_t *v, int re, int64_t *out) {
    __shared__ uint32_t a[3][648];
    __shared__ int16_t vv[648];
    const int t = threadIdx.x;
    const int m0 = blockIdx.x * VT_M_PER, m1 = min(re, m0 + VT_M_PER);
    int64_t acc[6][3] = {};
    for (int m = m0; m < m1; m++) {
        for (int i = t; i < 648; i += 256) {
            const uint64_t lo = amat[((size_t)m * 648 + i) * 2], hi = amat[((size_t)m * 648 + i) * 2 + 1];
            a[0][i] = (uint32_t)lo, a[1][i] = (uint32_t)(lo >> 32), a[2][i] = (uint32_t)hi;
            vv[i] = v[(size_t)m * 648 + i];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int e = t + 256 * q;
            if (e >= VT_RAW) break;
            const int i0 = max(0, e - 647), i1 = min(e, 647);
            for (int i = i0; i <= i1; i++) {
                const int64_t x = vv[i];
                acc[q][0] += (int64_t)a[0][e - i] * x;
                acc[q][1] += (int64_t)a[1][e - i] * x;
                acc[q][2] += (int64_t)a[2][e - i] * x;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int e = t + 256 * q;
        if (e >= VT_RAW) break;
#pragma unroll
        for (int k = 0; k < 3; k++) out[((size_t)blockIdx.x * VT_RAW + e) * 3 + k] = acc[q][k];
    }
}

// sum_j c_j(-X^4) C_j: block (x, y) takes e = 256 x + t and columns 64 y .. 64 y + 63; the
// coefficient X^(4 pos) of c_j(-X^4) is +-1 with sign (1 - 2 s_t) (-1)^pos. C_j's coefficients
// are read from the packed commitment (`width` bytes LE each, < Q < 2^96).
__global__ __launch_bounds__(256) void k_vt_comb(const uint8_t *cbytes, int width, const lbn_challenge *ch,
                                                 int64_t *out) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= VT_RAW) return;
    int64_t acc[3] = {0, 0, 0};
    for (int j = blockIdx.y * 64; j < blockIdx.y * 64 + 64; j++) {
        const lbn_challenge c = ch[j];
        for (uint32_t s = 0; s < c.weight; s++) {
            const int pos = c.positions[s];
            const int idx = e - 4 * pos;
            if (idx < 0 || idx >= 648) continue;
            const bool neg = (((c.signs >> s) & 1) != 0) != ((pos & 1) != 0);
            const uint8_t *b = cbytes + ((size_t)j * 648 + idx) * width;
            uint32_t limb[3] = {0, 0, 0};
            for (int i = 0; i < width; i++) limb[i >> 2] |= (uint32_t)b[i] << (8 * (i & 3));
#pragma unroll
            for (int k = 0; k < 3; k++) acc[k] += neg ? -(int64_t)limb[k] : (int64_t)limb[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) out[((size_t)blockIdx.y * VT_RAW + e) * 3 + k] = acc[k];
}

int lbn::dev_verify_products(lbn_ctx *ctx, u128 q, const int16_t *v, const uint8_t *commitment, int width,
                             const lbn_challenge *ch, __int128 *lhs, __int128 *rhs) {
    const size_t re = ctx->shape.ring_elements;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int nb = (int)((re + VT_M_PER - 1) / VT_M_PER), ny = COLUMNS / 64;
    const size_t clen = (size_t)COLUMNS * N * width;
    uint64_t *damat = nullptr;
    int16_t *dv = nullptr;
    uint8_t *dc = nullptr;
    lbn_challenge *dch = nullptr;
    int64_t *dout = nullptr;
    std::vector<int64_t> out((size_t)(nb + ny) * VT_RAW * 3);
    auto release = [&]() {
        for (void *p : {(void *)damat, (void *)dv, (void *)dc, (void *)dch, (void *)dout})
            if (p) (void)hipFree(p);
    };
    hipError_t e = hipSuccess;
    auto ok = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
        return e == hipSuccess;
    };
    bool expanded = false;
    if (ok(hipMalloc(&damat, re * N * 16)) && ok(hipMalloc(&dv, re * N * 2)) && ok(hipMalloc(&dc, clen + 16)) &&
        ok(hipMalloc(&dch, COLUMNS * sizeof(lbn_challenge))) && ok(hipMalloc(&dout, out.size() * 8)) &&
        ok(ajtai_device(ctx->dev->stream, q, re, 0, re, damat, &expanded)) && expanded &&
        ok(hipMemcpyAsync(dv, v, re * N * 2, hipMemcpyHostToDevice, ctx->dev->stream)) &&
        ok(hipMemcpyAsync(dc, commitment, clen, hipMemcpyHostToDevice, ctx->dev->stream)) &&
        ok(hipMemcpyAsync(dch, ch, COLUMNS * sizeof(lbn_challenge), hipMemcpyHostToDevice, ctx->dev->stream))) {
        k_vt_atv<<<nb, 256, 0, ctx->dev->stream>>>(damat, dv, (int)re, dout);
        k_vt_comb<<<dim3((VT_RAW + 255) / 256, ny), 256, 0, ctx->dev->stream>>>(dc, width, dch,
                                                                               dout + (size_t)nb * VT_RAW * 3);
        if (ok(hipGetLastError()))
            if (ok(hipMemcpyAsync(out.data(), dout, out.size() * 8, hipMemcpyDeviceToHost, ctx->dev->stream)))
                ok(hipStreamSynchronize(ctx->dev->stream));
    }
    release();
    if (e != hipSuccess) return ctx->fail(LBN_E_DEVICE, std::string("checker products: ") + hipGetErrorString(e));
    if (!expanded) return ctx->fail(LBN_E_DEVICE, "matrix expansion did not finish");
    for (int i = 0; i < VT_RAW * 3; i++) {
        __int128 l = 0, r = 0;
        for (int b = 0; b < nb; b++) l += out[(size_t)b * VT_RAW * 3 + i];
        for (int y = 0; y < ny; y++) r += out[(size_t)(nb + y) * VT_RAW * 3 + i];
        lhs[i] = l, rhs[i] = r;
    }
    return LBN_OK;
}

// ----------------------------------------------------------------------------- allocation-free partials
// lbn_partials_static: the partials of a table in the caller's device memory, with every buffer the
// kernels need besides the table (EL, EH, the 128 sums) in this file-scope array. Host code never
// passes `scratch` itself: its name is a host-side stand-in, so each device's address comes from
// hipGetSymbolAddress on the first call made on that device, and is used from then on.
constexpr size_t STATIC_EL = 0;                         // 4096 B128
constexpr size_t STATIC_EH = STATIC_EL + 4096 * 16;     // up to 4096 B128 (2^(k-19) rows, k <= 31)
constexpr size_t STATIC_OUT = STATIC_EH + 4096 * 16;    // 128 B128 accumulators
constexpr size_t STATIC_SCRATCH_BYTES = STATIC_OUT + 128 * 16;
static __device__ __align__(256) uint8_t scratch[STATIC_SCRATCH_BYTES];
static std::mutex static_mu;
static uint8_t *static_ptr[64];  // per device, from hipGetSymbolAddress
static thread_local std::string static_err;

static int static_fail(int code, const char *msg) {
    static_err = msg;
    return code;
}

extern "C" size_t lbn_partials_static_scratch(void) { return STATIC_SCRATCH_BYTES; }
extern "C" const char *lbn_partials_static_error(void) { return static_err.c_str(); }

extern "C" int lbn_partials_static(int device, const void *table_device, size_t bytes, int log_vars,
                                   const uint64_t *r_hi, uint64_t *partials) {
    Shape s;
    if (!Shape::of(log_vars, s)) return static_fail(LBN_E_ARG, "log_vars outside [19, 31]");
    if (!table_device || !r_hi || !partials) return static_fail(LBN_E_ARG, "null argument");
    if (bytes != s.table_bytes) return static_fail(LBN_E_ARG, "table length is not 2^log_vars / 8 bytes");
    if ((uintptr_t)table_device % 16) return static_fail(LBN_E_ARG, "device table must be 16-byte aligned");
    const int lp = s.log_packed, nh = lp - 12;
    const size_t rows = (size_t)1 << nh;
    if (STATIC_EH + rows * 16 > STATIC_OUT)
        return static_fail(LBN_E_ARG, "more eq rows than the static scratch holds (at most 4096: log_vars <= 31)");
    if (device < 0 || device >= 64) return static_fail(LBN_E_ARG, "device index outside [0, 64)");
    std::lock_guard<std::mutex> lk(static_mu);  // one call at a time: the calls share the array
    if (hipSetDevice(device) != hipSuccess) return static_fail(LBN_E_DEVICE, "hipSetDevice failed");
    if (!static_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return static_fail(LBN_E_DEVICE, "hipGetSymbolAddress of the scratch array failed");
        }
        static_ptr[device] = (uint8_t *)p;
    }
    uint8_t *sc = static_ptr[device];
    B128 EH[4096], EL[4096];
    eq_expand_b128((const B128 *)r_hi, nh, EH);
    eq_expand_b128((const B128 *)r_hi + nh, 12, EL);
    if (hipMemcpyAsync(sc + STATIC_EL, EL, sizeof EL, hipMemcpyHostToDevice, 0) != hipSuccess ||
        hipMemcpyAsync(sc + STATIC_EH, EH, rows * 16, hipMemcpyHostToDevice, 0) != hipSuccess)
        return static_fail(LBN_E_DEVICE, "copy of the eq factors failed");
    // the sums start from zero: a fill kernel (the array keeps the last call's contents)
    k_fill_u32<<<1, 256, 0, 0>>>((uint32_t *)(sc + STATIC_OUT), 512, 0u);
    k_partials_fact<<<(unsigned)std::max<size_t>(1, std::min<size_t>(rows / PF_WAVES, 1024)), 64 * PF_WAVES, 0, 0>>>(
        (const uint4 *)table_device, (const uint4 *)(sc + STATIC_EL), (const uint4 *)(sc + STATIC_EH), (int)rows,
        (uint32_t *)(sc + STATIC_OUT));
    if (hipGetLastError() != hipSuccess) return static_fail(LBN_E_DEVICE, "kernel launch failed");
    uint32_t out[512];
    if (hipMemcpy(out, sc + STATIC_OUT, sizeof out, hipMemcpyDeviceToHost) != hipSuccess)
        return static_fail(LBN_E_DEVICE, "copy of the partials failed");
    for (int k = 0; k < 128; k++) {
        partials[2 * k] = (uint64_t)out[4 * k] | ((uint64_t)out[4 * k + 1] << 32);
        partials[2 * k + 1] = (uint64_t)out[4 * k + 2] | ((uint64_t)out[4 * k + 3] << 32);
    }
    return LBN_OK;
}

// T holds 128 values of V words: e = sum_k L[k] * T_k (F162 products on the host).
static F162 combine(const std::vector<F162> &L, const uint32_t *T, int V) {
    F162 acc{{0, 0, 0}};
    for (int k = 0; k < 128; k++) {
        const uint32_t *t = T + V * k;
        F162 x{{(uint64_t)t[0] | ((uint64_t)t[1] << 32), (uint64_t)t[2] | ((uint64_t)t[3] << 32),
                V == 6 ? ((uint64_t)t[4] | ((uint64_t)t[5] << 32)) : 0}};
        if (x.w[0] | x.w[1] | x.w[2]) acc = f162_add(acc, f162_mul(L[k], x));
    }
    return acc;
}

// Message sums of switch round r (T0 then Tinf, 768 words each) alternate between two buffers;
// the kernel that produces round r + 1's sums clears round r's buffer for round r + 2.
static uint32_t *msg_buf(DeviceState *d, int r) { return d->T + ((r & 1) ? 2048 : 0); }

static int launch_blocks(size_t items) { return (int)std::max<size_t>(1, std::min<size_t>((items + 255) / 256, 1024)); }

// The start of the switch for lbn_prove (SwitchVerifier::start's prover side, crossfield.rs:645-660):
// the map L_0 from the batch; the rounds (dev_switch_rounds) start from the prefix sets.
int lbn::dev_switch_prepare(lbn_ctx *ctx, const uint64_t *batch) {
    if (!ctx->partials_done) return ctx->fail(LBN_E_STATE, "lbn_partials must run first");
    ctx->lmap.assign((const F162 *)batch, (const F162 *)batch + 128);
    ctx->round = 0;
    return LBN_OK;
}

// The round-by-round switch (lbn_switch_msg / lbn_switch_fold) works on the whole eq expansion:
// it is expanded here (EH over the first log_packed - 12 coordinates, EL over the last 12), then
// round 0's message sums.
extern "C" int lbn_switch_begin(lbn_ctx *ctx, const uint64_t *batch) {
    if (!ctx || !batch) return LBN_E_ARG;
    if (!ctx->partials_done) return ctx->fail(LBN_E_STATE, "lbn_partials must run first");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    {
        const int rc_ = table_ready(ctx);  // a deferred upload completes first
        if (rc_ != LBN_OK) return rc_;
    }
    DeviceState *d = ctx->dev;
    ctx->lmap.assign((const F162 *)batch, (const F162 *)batch + 128);
    ctx->round = 0;
    const int lp = ctx->shape.log_packed, nh = lp - 12;
    const size_t h = ctx->shape.n_words / 2;
    eq_expand_b128(ctx->r_hi.data(), nh, (B128 *)(d->hbox + mb::EH));
    eq_expand_b128(ctx->r_hi.data() + nh, 12, (B128 *)(d->hbox + mb::EL));
    HIPCHK(ctx, hipMemcpyAsync(d->EL, d->hbox + mb::EL, (4096 + ((size_t)1 << nh)) * 16, hipMemcpyHostToDevice,
                               d->stream));
    // also zeroes both message buffers
    k_eq_expand<<<1u << nh, 256, 0, d->stream>>>(d->EH, d->EL, 12, d->E[0], d->T, 3584);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(d->stream));  // the mailbox regions are rewritten by the next call
    tstart(ctx, T_SWITCH0);
    k_switch_msg0<<<launch_blocks(h), 256, 0, d->stream>>>(d->table, d->E[0], h, d->E[1], d->T, d->T + 768);
    tstop(ctx, T_SWITCH0);
    HIPCHK(ctx, hipGetLastError());
    return LBN_OK;
}

extern "C" int lbn_switch_msg(lbn_ctx *ctx, uint64_t *msg) {
    if (!ctx || !msg) return LBN_E_ARG;
    if (ctx->round < 0 || ctx->round >= ctx->shape.log_packed) return ctx->fail(LBN_E_STATE, "no switch round pending");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    DeviceState *d = ctx->dev;
    const uint32_t *T = (const uint32_t *)(d->hbox + mb::T);
    HIPCHK(ctx, hipMemcpyAsync(d->hbox + mb::T, msg_buf(d, ctx->round), 1536 * 4, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(ctx, hipStreamSynchronize(d->stream));
    tcollect(ctx, ctx->round == 0 ? T_SWITCH0 : T_ROUNDS);
    const int V = ctx->round == 0 ? 4 : 6;
    F162 e0 = combine(ctx->lmap, T, V), einf = combine(ctx->lmap, T + 768, V);
    memcpy(msg, e0.w, 24);
    memcpy(msg + 3, einf.w, 24);
    return LBN_OK;
}

// Fold of round r with rho (by value, or from device memory when rho_dev is set) fused with the
// message of round r + 1.
static void launch_round(lbn_ctx *ctx, int r, RhoArg ra, const uint64_t *rho_dev) {
    DeviceState *d = ctx->dev;
    const size_t h2 = (ctx->shape.n_words >> (r + 1)) / 2;
    const uint4 *Ein = d->E[(r + 1) & 1];
    uint4 *Eout = d->E[r & 1];
    uint32_t *pout = d->P[r & 1];
    uint32_t *T = msg_buf(d, r + 1);
    if (r == 0)
        k_switch_round<true><<<launch_blocks(h2), 256, 0, d->stream>>>((const uint32_t *)d->table, h2, ra, Ein, Eout,
                                                                        pout, T, T + 768, msg_buf(d, r), rho_dev);
    else
        k_switch_round<false><<<launch_blocks(h2), 256, 0, d->stream>>>(d->P[(r - 1) & 1], h2, ra, Ein, Eout, pout,
                                                                         T, T + 768, msg_buf(d, r), rho_dev);
}

// The switch rounds' device launches. Rounds 0..5 run on the prefix sets (k_prefix_msg, then control
// step r) and the bind kernel does their folds with round 6's message, as round 5's kernel would
// leave them (p^(6) in P[1], E^(7) in E[1], the sums in round 6's buffer); then control step r and
// round r's fold with
// round r + 1's message for 6 <= r < r0 = max(7, lp - SWITCH_HOST_STEPS) (control steps) and
// r < r0 - 1 (folds); control step r0 - 1 hands the state and round r0 - 1's fold input (written by
// the kernel of round r0 - 2) to the host. Everything they read (shape, buffers, the device state
// d->sw, the mailbox) is fixed per context.
static int switch_host_from(int lp) { return std::max(7, lp - SWITCH_HOST_STEPS); }
static void launch_switch(lbn_ctx *ctx) {
    DeviceState *d = ctx->dev;
    const int lp = ctx->shape.log_packed, r0 = switch_host_from(lp), S = lp - 6;
    const int h2 = 1 << (lp - r0 - 1);  // the handoff fold's items (at most mb::HANDOFF_H2)
    const uint4 *lam = (const uint4 *)(d->pset + ps::LAM);
    k_prefix_q0<<<32, 256, 0, d->stream>>>(d->pset + ps::M, d->pset + ps::Q, (const uint4 *)(d->hbox_dev + mb::LAM),
                                           (uint4 *)lam, msg_buf(d, 0));
    for (int r = 0; r < 6; r++) {
        uint32_t *T = msg_buf(d, r);
        k_prefix_msg<<<(unsigned)((128 << (5 - r)) + 255) / 256, 256, 0, d->stream>>>(d->pset + ps::Q, r, lam, d->sw, T,
                                                                                    T + 768, msg_buf(d, r + 1));
        k_switch_ctl<<<1, 1024, 0, d->stream>>>(d->sw, (SwitchDev *)(d->hbox_dev + mb::SW), r, -1, 6, T, nullptr,
                                                nullptr, 0, nullptr, nullptr, 0);
    }
    {
        // lanes per output: enough lanes in flight for the device (2^17) when the items are few
        const size_t items = (size_t)1 << (S - 1);
        int gl = 0;
        while (gl < 5 && (items << gl) < ((size_t)1 << 17)) gl++;
        const unsigned blocks = (unsigned)std::min<size_t>(((items << gl) + BIND_THREADS - 1) / BIND_THREADS, 512);
        uint32_t *T = msg_buf(d, 6), *clr = msg_buf(d, 7);
        uint4 *G = d->E[0], *E7 = d->E[1];
        uint32_t *P6 = d->P[1];
        switch (gl) {
            case 0: k_switch_bind<0><<<blocks, BIND_THREADS, 0, d->stream>>>(d->table, S, d->sw, G, E7, P6, T, T + 768, clr); break;
            case 1: k_switch_bind<1><<<blocks, BIND_THREADS, 0, d->stream>>>(d->table, S, d->sw, G, E7, P6, T, T + 768, clr); break;
            case 2: k_switch_bind<2><<<blocks, BIND_THREADS, 0, d->stream>>>(d->table, S, d->sw, G, E7, P6, T, T + 768, clr); break;
            case 3: k_switch_bind<3><<<blocks, BIND_THREADS, 0, d->stream>>>(d->table, S, d->sw, G, E7, P6, T, T + 768, clr); break;
            case 4: k_switch_bind<4><<<blocks, BIND_THREADS, 0, d->stream>>>(d->table, S, d->sw, G, E7, P6, T, T + 768, clr); break;
            default: k_switch_bind<5><<<blocks, BIND_THREADS, 0, d->stream>>>(d->table, S, d->sw, G, E7, P6, T, T + 768, clr); break;
        }
    }
    for (int r = 6; r < r0; r++) {
        const bool hand = r == r0 - 1;  // the handoff fold's input: P[(r0 - 2) & 1] (P[1] from the bind at r0 = 7)
        k_switch_ctl<<<1, 1024, 0, d->stream>>>(
            d->sw, (SwitchDev *)(d->hbox_dev + mb::SW), r, r0 - 1, 6, msg_buf(d, r),
            hand ? d->P[(r0 - 2) & 1] : nullptr, (uint32_t *)(d->hbox_dev + mb::HP), hand ? 4 * h2 : 0,
            hand ? d->E[r0 & 1] : nullptr, (uint4 *)(d->hbox_dev + mb::HE), hand ? 2 * h2 : 0);
        if (r < r0 - 1) launch_round(ctx, r, RhoArg{{0, 0, 0}}, d->sw->rho);
    }
}

// Captured once per context (at creation) and replayed by every proof.
static int capture_switch(lbn_ctx *ctx) {
    DeviceState *d = ctx->dev;
    HIPCHK(ctx, hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
    launch_switch(ctx);
    hipGraph_t g = nullptr;
    const hipError_t cap = hipGetLastError(), end = hipStreamEndCapture(d->stream, &g);
    hipError_t inst = hipSuccess;
    if (cap == hipSuccess && end == hipSuccess) inst = hipGraphInstantiate(&d->switch_graph, g, nullptr, nullptr, 0);
    if (g) hipGraphDestroy(g);
    HIPCHK(ctx, cap);
    HIPCHK(ctx, end);
    HIPCHK(ctx, inst);
    return LBN_OK;
}

// What the device rounds start from, written to the context's mailbox: the transcript state, the
// map L_0, r_hi and the prefix rounds' eq weights.
static void switch_mailbox(lbn_ctx *ctx, const B3State &ts) {
    const int lp = ctx->shape.log_packed;
    DeviceState *d = ctx->dev;
    SwitchDev *h = (SwitchDev *)(d->hbox + mb::SW);
    static_assert(sizeof(SwitchDev) <= 8192, "mailbox region");
    h->ts = ts;
    memcpy(h->lmap, ctx->lmap.data(), 128 * sizeof(F162));
    for (int r = 0; r < lp; r++) h->r_hi[r][0] = ctx->r_hi[r].w[0], h->r_hi[r][1] = ctx->r_hi[r].w[1];
    {  // the prefix rounds' eq weights: lambda1 of round r = eq(c_r+1..5, .), lambda0 = (1 + c_r) lambda1
        B128 *lam = (B128 *)(d->hbox + mb::LAM);
        for (int r = 0; r < 6; r++) {
            const int nv = 32 >> r, off = 64 - 2 * nv;
            eq_expand_b128(ctx->r_hi.data() + r + 1, 5 - r, lam + 64 + off);
            const B128 one_c = b128_add(ctx->r_hi[r], B128{{1, 0}});
            for (int v = 0; v < nv; v++) lam[off + v] = b128_mul(one_c, lam[64 + off + v]);
        }
    }
}

// The device rounds' results from the mailbox (after the stream has finished them).
static void switch_readback(lbn_ctx *ctx, B3State &ts, uint64_t *rounds, uint64_t *coords, SwitchHandoff &ho) {
    const int lp = ctx->shape.log_packed, r0 = switch_host_from(lp);
    DeviceState *d = ctx->dev;
    const SwitchDev *h = (const SwitchDev *)(d->hbox + mb::SW);
    ts = h->ts;
    memcpy(rounds, h->rounds, (size_t)r0 * 48);
    memcpy(coords, h->coords, (size_t)r0 * 24);
    const size_t h2 = (size_t)1 << (lp - r0 - 1);
    ho.r0 = r0;
    memcpy(ho.rho.w, h->rho, 24);
    ho.L.assign((const F162 *)h->lmap, (const F162 *)h->lmap + 128);
    ho.p.resize(4 * h2);
    memcpy(ho.p.data(), d->hbox + mb::HP, 4 * h2 * sizeof(F162));
    ho.e.resize(2 * h2);
    memcpy(ho.e.data(), d->hbox + mb::HE, 2 * h2 * sizeof(B128));
    ctx->round = lp;  // the host finishes the rounds (and sets ctx->opened)
}

int lbn::dev_switch_rounds(lbn_ctx *ctx, B3State &ts, uint64_t *rounds, uint64_t *coords, SwitchHandoff &ho) {
    if (ctx->round != 0) return ctx->fail(LBN_E_STATE, "lbn_switch_begin must run first");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    DeviceState *d = ctx->dev;
    switch_mailbox(ctx, ts);
    // the prefix kernel reads the state from the mailbox, the last control step writes it back there
    tstart(ctx, T_ROUNDS);
    if (d->graphs)
        HIPCHK(ctx, hipGraphLaunch(d->switch_graph, d->stream));
    else
        launch_switch(ctx);
    tstop(ctx, T_ROUNDS);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(d->stream));
    tcollect(ctx, T_ROUNDS);
    switch_readback(ctx, ts, rounds, coords, ho);
    return LBN_OK;
}

// ---------------------------------------------------------------------------------------------
// The switch of many proofs at once: each launch of the device round sequence (prefix rounds,
// bind, control steps and folds) covers every instance, so a batch of n proofs makes one launch
// per step instead of n. Instance i's blocks are [start[i], start[i + 1]) of the grid; an instance
// whose device rounds are over takes no blocks. Every instance's buffers and state are its own
// context's, so each proof is the one its context would make alone.
// ---------------------------------------------------------------------------------------------
struct SwView {
    const uint4 *table;
    uint32_t *pset, *T;
    SwitchDev *sw, *mbox_sw;
    const uint4 *mbox_lam;
    uint32_t *hp;
    uint4 *he;
    uint4 *E[2];
    uint32_t *P[2];
    size_t n_words;
    int lp, S, r0, gl;
};
struct SwMap {
    int n;
    int start[SWB_MAX + 1];
};
__device__ __forceinline__ uint32_t *swb_msg(const SwView &x, int r) { return x.T + ((r & 1) ? 2048 : 0); }
__device__ __forceinline__ int swb_find(const SwMap &mp, int b) {  // the last i with start[i] <= b
    int lo = 0, hi = mp.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (mp.start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
__global__ __launch_bounds__(256) void k_mprefix_q0(const SwView *v) {
    const SwView &x = v[blockIdx.x / 32];
    prefix_q0_body(blockIdx.x % 32, x.pset + ps::M, x.pset + ps::Q, x.mbox_lam, (uint4 *)(x.pset + ps::LAM),
                   swb_msg(x, 0));
}
__global__ __launch_bounds__(256) void k_mprefix_msg(const SwView *v, int r, int per) {
    __shared__ uint32_t rtab[RHO_WINDOWS * 16 * 6];
    __shared__ __align__(16) uint32_t lds[4 * M4W<6>::WORDS];
    const SwView &x = v[blockIdx.x / per];
    uint32_t *T = swb_msg(x, r);
    prefix_msg_body(blockIdx.x % per, x.pset + ps::Q, r, (const uint4 *)(x.pset + ps::LAM), x.sw, T, T + 768,
                    swb_msg(x, r + 1), rtab, lds);
}
// control step r of every instance (block i: instance i): a prefix round, or a later one while
// r < r0 of the instance (its last device step hands over to the host)
__global__ __launch_bounds__(1024) void k_mswitch_ctl(const SwView *v, int r, int prefix) {
    __shared__ CtlKeep K;
    __shared__ CtlTemp X;
    const SwView &x = v[blockIdx.x];
    if (prefix) {
        switch_ctl_body(K, X, x.sw, x.mbox_sw, r, -1, 6, swb_msg(x, r), nullptr, nullptr, 0, nullptr, nullptr, 0);
        return;
    }
    if (r >= x.r0) return;  // the whole block
    const bool hand = r == x.r0 - 1;
    const int h2 = 1 << (x.lp - x.r0 - 1);
    switch_ctl_body(K, X, x.sw, x.mbox_sw, r, x.r0 - 1, 6, swb_msg(x, r), hand ? x.P[(x.r0 - 2) & 1] : nullptr, x.hp,
                    hand ? 4 * h2 : 0, hand ? x.E[x.r0 & 1] : nullptr, x.he, hand ? 2 * h2 : 0);
}
__global__ __launch_bounds__(BIND_THREADS) void k_mswitch_bind(const SwView *v, SwMap mp) {
    __shared__ uint32_t rt[6 * RHO_WINDOWS * 16 * 6];
    __shared__ __align__(16) uint32_t lds[(BIND_THREADS / 64) * M4W<6>::WORDS];
    const int i = swb_find(mp, blockIdx.x), b = blockIdx.x - mp.start[i], nb = mp.start[i + 1] - mp.start[i];
    const SwView &x = v[i];
    uint32_t *T = swb_msg(x, 6), *clr = swb_msg(x, 7);
#define MBIND(G) \
    switch_bind_body<G>(b, nb, x.table, x.S, x.sw, x.E[0], x.E[1], x.P[1], T, T + 768, clr, rt, lds)
    switch (x.gl) {
        case 0: MBIND(0); break;
        case 1: MBIND(1); break;
        case 2: MBIND(2); break;
        case 3: MBIND(3); break;
        case 4: MBIND(4); break;
        default: MBIND(5); break;
    }
#undef MBIND
}
// the fold of round r (6 <= r < r0 - 1) with round r + 1's message sums
__global__ __launch_bounds__(256) void k_mswitch_round(const SwView *v, SwMap mp, int r) {
    __shared__ uint32_t rtab[RHO_WINDOWS * 16 * 6];
    __shared__ __align__(16) uint32_t lds[4 * M4W<6>::WORDS];
    const int i = swb_find(mp, blockIdx.x), b = blockIdx.x - mp.start[i], nb = mp.start[i + 1] - mp.start[i];
    const SwView &x = v[i];
    const size_t h2 = (x.n_words >> (r + 1)) / 2;
    uint32_t *T = swb_msg(x, r + 1);
    switch_round_body<false>(b, nb, x.P[(r - 1) & 1], h2, RhoArg{{0, 0, 0}}, x.E[(r + 1) & 1], x.E[r & 1],
                             x.P[r & 1], T, T + 768, swb_msg(x, r), x.sw->rho, rtab, lds);
}

// One per device: the stream and the view array of batched switches (calls on one device are
// serialised).
struct SwBatchDev {
    hipStream_t stream = nullptr;
    SwView *views = nullptr;
    struct ChalPtrs *chal = nullptr;
    RowView *rows = nullptr;
    FoldView *folds = nullptr;
    CvView *cvs = nullptr;
    PartView *parts = nullptr;
    Gather *gather = nullptr;
};
// Proof i's `bytes` staged at its mailbox's `region` into its device buffer dst[i], for every
// proof of the batch in one launch on the batch stream.
static int gather_launch(lbn_ctx *const *ctxs, int n, SwBatchDev &bd, size_t region, uint4 *const *dst,
                         const size_t *bytes) {
    std::vector<Gather> g(n);
    size_t most = 0;
    for (int i = 0; i < n; i++) {
        g[i].src = (const uint4 *)(ctxs[i]->dev->hbox_dev + region);
        g[i].dst = dst[i];
        g[i].n16 = (bytes[i] + 15) / 16;
        most = std::max(most, g[i].n16);
    }
    const int per = (int)std::min<size_t>(8, (most + 255) / 256);
    HIPCHK(ctxs[0], hipMemcpyAsync(bd.gather, g.data(), sizeof(Gather) * n, hipMemcpyHostToDevice, bd.stream));
    if (per > 0) k_mgather<<<n * per, 256, 0, bd.stream>>>(bd.gather, per);
    HIPCHK(ctxs[0], hipGetLastError());
    return LBN_OK;
}
static int swb_ready(lbn_ctx *ctx, SwBatchDev &bd) {
    if (bd.stream) return LBN_OK;
    // a set left half made by a failed allocation is completed by the next call, not allocated again
    if (!bd.views) HIPCHK(ctx, hipMalloc(&bd.views, sizeof(SwView) * SWB_MAX));
    if (!bd.chal) HIPCHK(ctx, hipMalloc(&bd.chal, sizeof(ChalPtrs) * SWB_MAX));
    if (!bd.rows) HIPCHK(ctx, hipMalloc(&bd.rows, sizeof(RowView) * SWB_MAX));
    if (!bd.folds) HIPCHK(ctx, hipMalloc(&bd.folds, sizeof(FoldView) * SWB_MAX));
    if (!bd.cvs) HIPCHK(ctx, hipMalloc(&bd.cvs, sizeof(CvView) * SWB_MAX));
    if (!bd.parts) HIPCHK(ctx, hipMalloc(&bd.parts, sizeof(PartView) * SWB_MAX));
    if (!bd.gather) HIPCHK(ctx, hipMalloc(&bd.gather, sizeof(Gather) * SWB_MAX));
    HIPCHK(ctx, hipStreamCreateWithFlags(&bd.stream, hipStreamNonBlocking));  // last: it marks the set ready
    return LBN_OK;
}
// Batch stream sets per device: a batched step leases one for its launches and its final wait, so
// the steps of batches proved at once (scheduler workers, staggered groups) overlap on the device.
constexpr int SWB_SETS = 4;
struct SwBatchSets {
    std::mutex m;
    std::condition_variable cv;
    SwBatchDev set[SWB_SETS];
    bool busy[SWB_SETS] = {};
};
static std::mutex swb_map_mu;
static std::map<int, std::unique_ptr<SwBatchSets>> swb_sets;
static SwBatchSets &swb_sets_of(int device) {
    std::lock_guard<std::mutex> lk(swb_map_mu);
    auto &p = swb_sets[device];
    if (!p) p.reset(new SwBatchSets());
    return *p;
}
class BatchLease {
  public:
    explicit BatchLease(int device) : s_(swb_sets_of(device)) {
        std::unique_lock<std::mutex> lk(s_.m);
        s_.cv.wait(lk, [&] {
            for (int i = 0; i < SWB_SETS; i++)
                if (!s_.busy[i]) return (idx_ = i), true;
            return false;
        });
        s_.busy[idx_] = true;
    }
    ~BatchLease() {
        {
            std::lock_guard<std::mutex> lk(s_.m);
            s_.busy[idx_] = false;
        }
        s_.cv.notify_one();
    }
    SwBatchDev &bd() { return s_.set[idx_]; }

  private:
    SwBatchSets &s_;
    int idx_ = 0;
};

// The contexts of a batch: on one device, each once.
static int batch_check(lbn_ctx *const *ctxs, int n) {
    if (n < 1 || n > SWB_MAX) return LBN_E_ARG;
    for (int i = 0; i < n; i++) {
        if (ctxs[i]->device != ctxs[0]->device) return ctxs[i]->fail(LBN_E_ARG, "a batch runs on one device");
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return ctxs[i]->fail(LBN_E_ARG, "a context appears twice in the batch");
    }
    return LBN_OK;
}

void lbn::dev_switch_stage(lbn_ctx *ctx, const B3State &ts) { switch_mailbox(ctx, ts); }
void lbn::dev_switch_collect(lbn_ctx *ctx, B3State &ts, uint64_t *rounds, uint64_t *coords, SwitchHandoff &ho) {
    switch_readback(ctx, ts, rounds, coords, ho);
}
int lbn::dev_sync(lbn_ctx *ctx) {
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->dev->stream));
    return LBN_OK;
}

int lbn::dev_switch_batch(lbn_ctx *const *ctxs, int n) {
    {
        const int rc = batch_check(ctxs, n);
        if (rc != LBN_OK) return rc;
    }
    const int dev = ctxs[0]->device;
    for (int i = 0; i < n; i++)
        if (ctxs[i]->round != 0) return ctxs[i]->fail(LBN_E_STATE, "lbn_switch_begin must run first");
    BatchLease lease(dev);
    HIPCHK(ctxs[0], hipSetDevice(dev));
    SwBatchDev &bd = lease.bd();
    {
        const int rc = swb_ready(ctxs[0], bd);
        if (rc != LBN_OK) return rc;
    }
    std::vector<SwView> hv(n);
    int max_r0 = 0;
    for (int i = 0; i < n; i++) {
        lbn_ctx *ctx = ctxs[i];
        DeviceState *d = ctx->dev;
        SwView &x = hv[i];
        x.table = d->table;
        x.pset = d->pset;
        x.T = d->T;
        x.sw = d->sw;
        x.mbox_sw = (SwitchDev *)(d->hbox_dev + mb::SW);
        x.mbox_lam = (const uint4 *)(d->hbox_dev + mb::LAM);
        x.hp = (uint32_t *)(d->hbox_dev + mb::HP);
        x.he = (uint4 *)(d->hbox_dev + mb::HE);
        x.E[0] = d->E[0], x.E[1] = d->E[1];
        x.P[0] = d->P[0], x.P[1] = d->P[1];
        x.n_words = ctx->shape.n_words;
        x.lp = ctx->shape.log_packed;
        x.S = x.lp - 6;
        x.r0 = switch_host_from(x.lp);
        const size_t items = (size_t)1 << (x.S - 1);
        int gl = 0;
        while (gl < 5 && (items << gl) < ((size_t)1 << 17)) gl++;
        x.gl = gl;
        max_r0 = std::max(max_r0, x.r0);
    }
    hipStream_t st = bd.stream;
    HIPCHK(ctxs[0], hipMemcpyAsync(bd.views, hv.data(), sizeof(SwView) * n, hipMemcpyHostToDevice, st));
    k_mprefix_q0<<<32 * n, 256, 0, st>>>(bd.views);
    for (int r = 0; r < 6; r++) {
        const int per = ((128 << (5 - r)) + 255) / 256;
        k_mprefix_msg<<<per * n, 256, 0, st>>>(bd.views, r, per);
        k_mswitch_ctl<<<n, 1024, 0, st>>>(bd.views, r, 1);
    }
    SwMap mp{};
    mp.n = n;
    for (int i = 0; i < n; i++) {
        const size_t items = (size_t)1 << (hv[i].S - 1);
        const int blocks = (int)std::min<size_t>(((items << hv[i].gl) + BIND_THREADS - 1) / BIND_THREADS, 512);
        mp.start[i + 1] = mp.start[i] + blocks;
    }
    k_mswitch_bind<<<mp.start[n], BIND_THREADS, 0, st>>>(bd.views, mp);
    for (int r = 6; r < max_r0; r++) {
        k_mswitch_ctl<<<n, 1024, 0, st>>>(bd.views, r, 0);
        for (int i = 0; i < n; i++) {
            const int blocks = r < hv[i].r0 - 1 ? launch_blocks((hv[i].n_words >> (r + 1)) / 2) : 0;
            mp.start[i + 1] = mp.start[i] + blocks;
        }
        if (mp.start[n] > 0) k_mswitch_round<<<mp.start[n], 256, 0, st>>>(bd.views, mp, r);
    }
    HIPCHK(ctxs[0], hipGetLastError());
    HIPCHK(ctxs[0], hipStreamSynchronize(st));
    return LBN_OK;
}

// L_{r+1}(y) = L_r((1 + c) y) + rho L_r(y), c = r_hi[r]: the images of x^k, k < 128, through
// 16 tables of 256 sums of L_r over 8-bit windows of y (method of four Russians).
static void update_lmap(lbn_ctx *ctx, int r, const F162 &rho) { lmap_next(ctx->lmap, ctx->r_hi[r], rho); }

extern "C" int lbn_switch_fold(lbn_ctx *ctx, const uint64_t *rr) {
    if (!ctx || !rr) return LBN_E_ARG;
    const int lp = ctx->shape.log_packed;
    if (ctx->round < 0 || ctx->round >= lp) return ctx->fail(LBN_E_STATE, "no switch round pending");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    {
        const int rc_ = table_ready(ctx);  // a deferred upload completes first
        if (rc_ != LBN_OK) return rc_;
    }
    DeviceState *d = ctx->dev;
    const int r = ctx->round;
    F162 rho{{rr[0], rr[1], rr[2]}};
    const size_t n = ctx->shape.n_words;
    const size_t h = n >> (r + 1);  // current half
    if (r == lp - 1) {
        // p^(lp-1) has two elements: opened = p0 + rho (p0 + p1).
        uint32_t two[12];
        const uint32_t *src = (r == 0) ? (const uint32_t *)d->table : d->P[(r - 1) & 1];
        const uint32_t *box = (const uint32_t *)(d->hbox + mb::MISC);
        HIPCHK(ctx, hipMemcpyAsync(d->hbox + mb::MISC, src, r == 0 ? 32 : 48, hipMemcpyDeviceToHost, d->stream));
        HIPCHK(ctx, hipStreamSynchronize(d->stream));
        if (r == 0) {
            for (int u = 0; u < 2; u++)
                for (int k = 0; k < 6; k++) two[6 * u + k] = k < 4 ? box[4 * u + k] : 0;
        } else {
            memcpy(two, box, 48);
        }
        F162 p0{{(uint64_t)two[0] | ((uint64_t)two[1] << 32), (uint64_t)two[2] | ((uint64_t)two[3] << 32),
                 (uint64_t)two[4] | ((uint64_t)two[5] << 32)}};
        F162 p1{{(uint64_t)two[6] | ((uint64_t)two[7] << 32), (uint64_t)two[8] | ((uint64_t)two[9] << 32),
                 (uint64_t)two[10] | ((uint64_t)two[11] << 32)}};
        ctx->opened = f162_add(p0, f162_mul(rho, f162_add(p0, p1)));
        ctx->round = lp;
        return LBN_OK;
    }
    tstart(ctx, T_ROUNDS);
    launch_round(ctx, r, RhoArg{{rr[0], rr[1], rr[2]}}, nullptr);
    tstop(ctx, T_ROUNDS);
    HIPCHK(ctx, hipGetLastError());
    update_lmap(ctx, r, rho);  // on the host while the kernel runs; only the next message needs it
    ctx->round = r + 1;
    return LBN_OK;
}

extern "C" int lbn_switch_final(lbn_ctx *ctx, uint64_t *opened) {
    if (!ctx || !opened) return LBN_E_ARG;
    if (ctx->round != ctx->shape.log_packed) return ctx->fail(LBN_E_STATE, "switch rounds not finished");
    memcpy(opened, ctx->opened.w, 24);
    return LBN_OK;
}

// ----------------------------------------------------------------------------- row evaluation

static int row_bits_lo(const Shape &s) { return std::min(s.row_log, 7); }

int lbn::dev_row_prepare(lbn_ctx *ctx, const uint64_t *p0, bool upload) {
    const Shape &s = ctx->shape;
    if (!ctx->table_loaded) return ctx->fail(LBN_E_STATE, "no table loaded");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    {
        const int rc_ = table_ready(ctx);  // a deferred upload completes first
        if (rc_ != LBN_OK) return rc_;
    }
    DeviceState *d = ctx->dev;
    const int lo_bits = row_bits_lo(s), hi_bits = s.row_log - lo_bits;
    std::vector<F162> lo((size_t)1 << lo_bits), hi((size_t)1 << hi_bits);
    eq_table_f162((const F162 *)p0, lo_bits, lo.data());
    eq_table_f162((const F162 *)p0 + lo_bits, hi_bits, hi.data());
    memcpy(d->hbox + mb::EQ, lo.data(), lo.size() * 24);
    memcpy(d->hbox + mb::EQ + 128 * 24, hi.data(), hi.size() * 24);
    if (upload)
        HIPCHK(ctx, hipMemcpyAsync(d->eqlo, d->hbox + mb::EQ, (128 + hi.size()) * 24, hipMemcpyHostToDevice,
                                   d->stream));
    return LBN_OK;
}

extern "C" int lbn_row_eval(lbn_ctx *ctx, const uint64_t *p0, int n_p0, uint64_t *row) {
    if (!ctx || !p0 || !row) return LBN_E_ARG;
    const Shape &s = ctx->shape;
    if (n_p0 != s.row_log) return ctx->fail(LBN_E_ARG, "p0 must have log_vars - 17 coordinates");
    {
        const int rc_ = dev_row_prepare(ctx, p0, true);
        if (rc_ != LBN_OK) return rc_;
    }
    DeviceState *d = ctx->dev;
    const int lo_bits = row_bits_lo(s), hi_bits = s.row_log - lo_bits;
    tstart(ctx, T_ROW);
    k_eq_row<<<1u << hi_bits, 256, 0, d->stream>>>((const uint64_t *)d->eqlo,
                                                   (const uint64_t *)((uint8_t *)d->eqlo + 128 * 24), lo_bits, d->eqrow);
    HIPCHK(ctx, hipGetLastError());
    k_row<<<COLUMNS, 256, 0, d->stream>>>(d->table, d->eqrow, s.column_len,
                                          (uint64_t *)(d->hbox_dev + mb::ROW));  // into the mailbox
    tstop(ctx, T_ROW);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(d->stream));
    memcpy(row, d->hbox + mb::ROW, (size_t)COLUMNS * 24);
    tcollect(ctx, T_ROW);
    return LBN_OK;
}

// ----------------------------------------------------------------------------- fold

int lbn::dev_fold_prepare(lbn_ctx *ctx, const lbn_challenge *ch, int n_challenges, int32_t *out, size_t n_out,
                          FoldPlan &plan, bool upload) {
    const Shape &s = ctx->shape;
    if (n_challenges != COLUMNS) return ctx->fail(LBN_E_ARG, "one challenge per column (1024)");
    if (n_out != s.ring_elements * N) return ctx->fail(LBN_E_ARG, "fold output has the wrong length");
    if (!ctx->table_loaded) return ctx->fail(LBN_E_STATE, "no table loaded");
    // Columns per challenge position, positives then negatives (sign = sign_t xor (pos odd)).
    std::vector<int> cnt(2 * 162, 0);
    for (int j = 0; j < COLUMNS; j++) {
        if (ch[j].weight > 32) return ctx->fail(LBN_E_ARG, "challenge weight above 32");
        for (uint32_t t = 0; t < ch[j].weight; t++) {
            const int pos = ch[j].positions[t];
            if (pos >= 162) return ctx->fail(LBN_E_ARG, "challenge position above 161");
            const int neg = (int)((ch[j].signs >> t) & 1) ^ (pos & 1);
            cnt[2 * pos + neg]++;
        }
    }
    std::vector<int> off(2 * 162 + 1, 0);
    for (int i = 0; i < 2 * 162; i++) off[i + 1] = off[i] + cnt[i];
    std::vector<uint16_t> cols(std::max(1, off[2 * 162]));
    std::vector<int> fillp(off.begin(), off.end() - 1);
    for (int j = 0; j < COLUMNS; j++)
        for (uint32_t t = 0; t < ch[j].weight; t++) {
            const int pos = ch[j].positions[t];
            const int neg = (int)((ch[j].signs >> t) & 1) ^ (pos & 1);
            cols[fillp[2 * pos + neg]++] = (uint16_t)j;
        }
    int maxlen = 0;  // one counter per position holds positives and complemented negatives
    for (int pos = 0; pos < 162; pos++) maxlen = std::max(maxlen, cnt[2 * pos] + cnt[2 * pos + 1]);
    // positions by list length, longest first: a wave's 16 positions then walk lists of similar
    // length (a wave runs as long as its longest list)
    for (int pos = 0; pos < 162; pos++) off.push_back(pos);
    std::stable_sort(off.begin() + 2 * 162 + 1, off.end(), [&](int a, int b) {
        return cnt[2 * a] + cnt[2 * a + 1] > cnt[2 * b] + cnt[2 * b + 1];
    });
    int nb = 1;
    while ((1 << nb) <= maxlen) nb++;
    nb = std::max(nb, 6);
    const int n_entries = off[2 * 162];
    static_assert((2 * 162 + 1 + 162) * sizeof(int) <= mb::FCOLS - mb::FOFF, "fold offsets region");
    if (n_entries > FOLD_MAX_ENTRIES || nb > 12) return ctx->fail(LBN_E_ARG, "challenge lists too long");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    {
        const int rc_ = table_ready(ctx);  // a deferred upload completes first
        if (rc_ != LBN_OK) return rc_;
    }
    DeviceState *d = ctx->dev;
    memcpy(d->hbox + mb::FOFF, off.data(), off.size() * sizeof(int));
    memcpy(d->hbox + mb::FCOLS, cols.data(), cols.size() * 2);
    // offsets and lists in one copy (the device buffers keep the mailbox's relative layout)
    plan.upload_bytes = (mb::FCOLS - mb::FOFF) + cols.size() * 2;
    if (upload)
        HIPCHK(ctx, hipMemcpyAsync(d->fold_off, d->hbox + mb::FOFF, plan.upload_bytes, hipMemcpyHostToDevice,
                                   d->stream));
    // the range flag lives in the mailbox: cleared here, set by the kernel, read after the sync
    *(volatile uint32_t *)(d->hbox + mb::MISC) = 0;
    // into the caller's buffer when it is one of ours (pinned), else into the context's staging
    int32_t *direct = ((uintptr_t)out & 3) ? nullptr : (int32_t *)pinned_alias(out, n_out * 4);
    plan.nb = nb;
    plan.n_entries = n_entries;
    plan.out = out;
    plan.n_out = n_out;
    plan.dst = direct ? direct : d->fold_host;
    return LBN_OK;
}

int lbn::dev_fold_collect(lbn_ctx *ctx, const FoldPlan &plan) {
    DeviceState *d = ctx->dev;
    // The kernel wrote the fold straight into pinned host memory over the host link, overlapped
    // with its own work; once its stream is idle the values are visible here.
    if (*(volatile uint32_t *)(d->hbox + mb::MISC))
        return ctx->fail(LBN_E_RANGE, "a fold coefficient does not fit the wire codec");
    if (plan.dst == d->fold_host) copy_parallel((uint8_t *)plan.out, (const uint8_t *)d->hfold, plan.n_out * 4);
    return LBN_OK;
}

extern "C" int lbn_fold(lbn_ctx *ctx, const lbn_challenge *ch, int n_challenges, int32_t *out, size_t n_out) {
    if (!ctx || !ch || !out) return LBN_E_ARG;
    FoldPlan plan;
    {
        const int rc_ = dev_fold_prepare(ctx, ch, n_challenges, out, n_out, plan, true);
        if (rc_ != LBN_OK) return rc_;
    }
    const Shape &s = ctx->shape;
    DeviceState *d = ctx->dev;
    int cus = 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) cus = 256;
    const unsigned grid = (unsigned)std::min<size_t>(s.ring_elements, (size_t)cus);
    uint32_t *flag = (uint32_t *)(d->hbox_dev + mb::MISC);
    tstart(ctx, T_FOLD);
#define FOLD_CASE(B)                                                                                          \
    case B:                                                                                                   \
        k_fold<B><<<grid, FOLD_THREADS, 0, d->stream>>>(d->table, s.column_len, s.ring_elements, d->fold_off,   \
                                                        d->fold_cols, plan.n_entries, plan.dst, flag);        \
        break;
    switch (plan.nb) {
        FOLD_CASE(6)
        FOLD_CASE(7)
        FOLD_CASE(8)
        FOLD_CASE(9)
        FOLD_CASE(10)
        FOLD_CASE(11)
        FOLD_CASE(12)
        default:
            return ctx->fail(LBN_E_ARG, "challenge lists too long");
    }
#undef FOLD_CASE
    tstop(ctx, T_FOLD);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(d->stream));
    tcollect(ctx, T_FOLD);
    return dev_fold_collect(ctx, plan);
}

void lbn::dev_challenges_stage(lbn_ctx *ctx, const ChalArg &arg) { memcpy(ctx->dev->hbox + mb::CH, &arg, sizeof(ChalArg)); }
void lbn::dev_challenges_collect(lbn_ctx *ctx, lbn_challenge *out) {
    memcpy(out, ctx->dev->hbox + mb::CHOUT, COLUMNS * sizeof(lbn_challenge));
}

int lbn::dev_challenges_batch(lbn_ctx *const *ctxs, int n) {
    {
        const int rc = batch_check(ctxs, n);
        if (rc != LBN_OK) return rc;
    }
    const int dev = ctxs[0]->device;
    BatchLease lease(dev);
    HIPCHK(ctxs[0], hipSetDevice(dev));
    SwBatchDev &bd = lease.bd();
    {
        const int rc = swb_ready(ctxs[0], bd);
        if (rc != LBN_OK) return rc;
    }
    static_assert((sizeof(ChalArg) + 15) / 16 * 16 <= 8192, "the sampler's input in its mailbox region");
    std::vector<ChalPtrs> hp(n);
    std::vector<uint4 *> dst(n);
    std::vector<size_t> bytes(n, sizeof(ChalArg));
    for (int i = 0; i < n; i++) {
        DeviceState *d = ctxs[i]->dev;
        hp[i] = ChalPtrs{d->chal, (lbn_challenge *)(d->hbox_dev + mb::CHOUT)};
        dst[i] = (uint4 *)d->chal;
    }
    {
        const int rc = gather_launch(ctxs, n, bd, mb::CH, dst.data(), bytes.data());
        if (rc != LBN_OK) return rc;
    }
    HIPCHK(ctxs[0], hipMemcpyAsync(bd.chal, hp.data(), sizeof(ChalPtrs) * n, hipMemcpyHostToDevice, bd.stream));
    k_mchallenges<<<n * (COLUMNS / chal::WAVES), 64 * chal::WAVES, 0, bd.stream>>>(bd.chal);
    HIPCHK(ctxs[0], hipGetLastError());
    HIPCHK(ctxs[0], hipStreamSynchronize(bd.stream));
    return LBN_OK;
}

void lbn::dev_row_collect(lbn_ctx *ctx, uint64_t *row) {
    memcpy(row, ctx->dev->hbox + mb::ROW, (size_t)COLUMNS * 24);
}

int lbn::dev_row_batch(lbn_ctx *const *ctxs, int n) {
    {
        const int rc = batch_check(ctxs, n);
        if (rc != LBN_OK) return rc;
    }
    BatchLease lease(ctxs[0]->device);
    HIPCHK(ctxs[0], hipSetDevice(ctxs[0]->device));
    SwBatchDev &bd = lease.bd();
    {
        const int rc = swb_ready(ctxs[0], bd);
        if (rc != LBN_OK) return rc;
    }
    std::vector<RowView> hv(n);
    int eq_blocks = 0;
    for (int i = 0; i < n; i++) {
        lbn_ctx *ctx = ctxs[i];
        DeviceState *d = ctx->dev;
        const Shape &s = ctx->shape;
        RowView &x = hv[i];
        x.table = d->table;
        x.clen = s.column_len;
        x.lo = (const uint64_t *)(d->hbox_dev + mb::EQ);  // read once per thread straight from the mailbox
        x.hi = (const uint64_t *)(d->hbox_dev + mb::EQ + 128 * 24);
        x.lo_bits = row_bits_lo(s);
        x.eq_start = eq_blocks;
        x.eqrow = d->eqrow;
        x.row = (uint64_t *)(d->hbox_dev + mb::ROW);
        eq_blocks += 1 << (s.row_log - x.lo_bits);
    }
    hipStream_t st = bd.stream;
    HIPCHK(ctxs[0], hipMemcpyAsync(bd.rows, hv.data(), sizeof(RowView) * n, hipMemcpyHostToDevice, st));
    k_meq_row<<<eq_blocks, 256, 0, st>>>(bd.rows, n);
    HIPCHK(ctxs[0], hipGetLastError());
    k_mrow<<<n * COLUMNS, 256, 0, st>>>(bd.rows);
    HIPCHK(ctxs[0], hipGetLastError());
    HIPCHK(ctxs[0], hipStreamSynchronize(st));
    return LBN_OK;
}

int lbn::dev_fold_batch(lbn_ctx *const *ctxs, int n, const FoldPlan *plans) {
    {
        const int rc = batch_check(ctxs, n);
        if (rc != LBN_OK) return rc;
    }
    const int dev = ctxs[0]->device;
    BatchLease lease(dev);
    HIPCHK(ctxs[0], hipSetDevice(dev));
    SwBatchDev &bd = lease.bd();
    {
        const int rc = swb_ready(ctxs[0], bd);
        if (rc != LBN_OK) return rc;
    }
    std::vector<FoldView> hv(n);
    size_t total = 0;
    int nb = 6;
    for (int i = 0; i < n; i++) {
        lbn_ctx *ctx = ctxs[i];
        DeviceState *d = ctx->dev;
        FoldView &x = hv[i];
        x.table = d->table;
        x.clen = ctx->shape.column_len;
        x.re = ctx->shape.ring_elements;
        x.first = total;
        x.off = d->fold_off;
        x.cols = d->fold_cols;
        x.n_entries = plans[i].n_entries;
        x.out = plans[i].dst;
        x.flag = (uint32_t *)(d->hbox_dev + mb::MISC);
        total += x.re;
        nb = std::max(nb, plans[i].nb);  // wider counters give the same counts
    }
    int cus = 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    const size_t per = (total + cus - 1) / cus;
    const unsigned grid = (unsigned)((total + per - 1) / per);
    hipStream_t st = bd.stream;
    {  // the column lists staged in the mailboxes
        std::vector<uint4 *> dst(n);
        std::vector<size_t> bytes(n);
        for (int i = 0; i < n; i++) dst[i] = (uint4 *)ctxs[i]->dev->fold_off, bytes[i] = plans[i].upload_bytes;
        const int rc = gather_launch(ctxs, n, bd, mb::FOFF, dst.data(), bytes.data());
        if (rc != LBN_OK) return rc;
    }
    HIPCHK(ctxs[0], hipMemcpyAsync(bd.folds, hv.data(), sizeof(FoldView) * n, hipMemcpyHostToDevice, st));
    switch (nb) {
        case 6: k_mfold<6><<<grid, FOLD_THREADS, 0, st>>>(bd.folds, n, per); break;
        case 7: k_mfold<7><<<grid, FOLD_THREADS, 0, st>>>(bd.folds, n, per); break;
        case 8: k_mfold<8><<<grid, FOLD_THREADS, 0, st>>>(bd.folds, n, per); break;
        case 9: k_mfold<9><<<grid, FOLD_THREADS, 0, st>>>(bd.folds, n, per); break;
        case 10: k_mfold<10><<<grid, FOLD_THREADS, 0, st>>>(bd.folds, n, per); break;
        case 11: k_mfold<11><<<grid, FOLD_THREADS, 0, st>>>(bd.folds, n, per); break;
        case 12: k_mfold<12><<<grid, FOLD_THREADS, 0, st>>>(bd.folds, n, per); break;
        default: return ctxs[0]->fail(LBN_E_ARG, "challenge lists too long");
    }
    HIPCHK(ctxs[0], hipGetLastError());
    HIPCHK(ctxs[0], hipStreamSynchronize(st));
    return LBN_OK;
}

void lbn::dev_batch_quiesce(int device) {
    SwBatchSets &sets = swb_sets_of(device);
    std::vector<hipStream_t> st;
    {
        std::lock_guard<std::mutex> lk(sets.m);
        for (auto &b : sets.set)
            if (b.stream) st.push_back(b.stream);
    }
    (void)hipSetDevice(device);
    for (hipStream_t x : st) (void)hipStreamSynchronize(x);
}

// The batched steps' common start: the contexts checked, the device current and its batch buffers.
static int batch_begin(lbn_ctx *const *ctxs, int n, BatchLease &lease, SwBatchDev *&bd) {
    HIPCHK(ctxs[0], hipSetDevice(ctxs[0]->device));
    bd = &lease.bd();
    return swb_ready(ctxs[0], *bd);
}

int lbn::dev_cvs_batch(lbn_ctx *const *ctxs, int n) {
    {
        const int rc = batch_check(ctxs, n);
        if (rc != LBN_OK) return rc;
    }
    BatchLease lease(ctxs[0]->device);
    SwBatchDev *bd = nullptr;
    {
        const int rc = batch_begin(ctxs, n, lease, bd);
        if (rc != LBN_OK) return rc;
    }
    std::vector<CvView> hv(n);
    int windows = 0;
    size_t max_count = 0;
    for (int i = 0; i < n; i++) {
        lbn_ctx *ctx = ctxs[i];
        DeviceState *d = ctx->dev;
        if (!d->cvjob) return ctx->fail(LBN_E_STATE, "no commitment hashing staged");
        const CvJob &job = *d->cvjob;
        const size_t clen = (size_t)COLUMNS * N * ctx->width;
        CvView &x = hv[i];
        x.bytes = d->cbytes;
        x.off = job.off;
        x.first = job.first;
        x.count = job.count;
        x.key = job.key;
        x.cvs = d->cvs;
        x.start = windows;
        x.tr = job.tr;
        x.roots = d->hcvs_dev;
        x.head = (int)job.head;
        x.tail = job.tail;
        x.tail_len = (int)(clen - job.tail);
        x.head_out = d->hbox_dev + mb::HEAD;
        x.tail_out = d->hbox_dev + mb::TAIL;
        windows += (int)cvs_windows(job);
        max_count = std::max(max_count, job.count);
    }
    hipStream_t st = bd->stream;
    HIPCHK(ctxs[0], hipMemcpyAsync(bd->cvs, hv.data(), sizeof(CvView) * n, hipMemcpyHostToDevice, st));
    if (windows) {
        k_mchunk_cvs<<<windows, 64, 0, st>>>(bd->cvs, n);
        HIPCHK(ctxs[0], hipGetLastError());
    }
    // every proof's subtrees, head and tail (a proof without whole chunks copies only those)
    if (max_count * 32 <= CV_TREE_LDS_MAX)
        k_mcv_tree<true><<<n, 1024, max_count * 32, st>>>(bd->cvs);
    else
        k_mcv_tree<false><<<n, 1024, 0, st>>>(bd->cvs);
    HIPCHK(ctxs[0], hipGetLastError());
    HIPCHK(ctxs[0], hipStreamSynchronize(st));
    return LBN_OK;
}

int lbn::dev_partials_batch(lbn_ctx *const *ctxs, int n) {
    {
        const int rc = batch_check(ctxs, n);
        if (rc != LBN_OK) return rc;
    }
    BatchLease lease(ctxs[0]->device);
    SwBatchDev *bd = nullptr;
    {
        const int rc = batch_begin(ctxs, n, lease, bd);
        if (rc != LBN_OK) return rc;
    }
    std::vector<PartView> hv(n);
    int eq_blocks = 0, ps_blocks = 0;
    for (int i = 0; i < n; i++) {
        lbn_ctx *ctx = ctxs[i];
        DeviceState *d = ctx->dev;
        const PartShape p = part_shape(ctx->shape);
        PartView &x = hv[i];
        x.EH = d->EH;
        x.EL = d->EL;
        x.lo = p.lo;
        x.E0 = d->E[0];
        x.clear = d->pset;
        x.nclear = (int)ps::Q;
        x.eq_start = eq_blocks;
        x.table = d->table;
        x.S = p.S;
        x.chunk_log = p.chunk_log;
        x.ps_start = ps_blocks;
        x.eq6 = d->EH + 64;
        x.M = d->pset + ps::M;
        x.part = d->pset + ps::PART;
        x.part_out = (uint32_t *)(d->hbox_dev + mb::PART);
        eq_blocks += 1 << p.nh;
        ps_blocks += 1 << (p.lp - p.chunk_log);
    }
    hipStream_t st = bd->stream;
    {  // the eq tables staged in the mailboxes
        std::vector<uint4 *> dst(n);
        std::vector<size_t> bytes(n, (4096 + 128) * 16);
        for (int i = 0; i < n; i++) dst[i] = ctxs[i]->dev->EL;
        const int rc = gather_launch(ctxs, n, *bd, mb::EL, dst.data(), bytes.data());
        if (rc != LBN_OK) return rc;
    }
    HIPCHK(ctxs[0], hipMemcpyAsync(bd->parts, hv.data(), sizeof(PartView) * n, hipMemcpyHostToDevice, st));
    k_meq_expand<<<eq_blocks, 256, 0, st>>>(bd->parts, n);  // also zeroes each proof's sets and partials
    HIPCHK(ctxs[0], hipGetLastError());
    k_mprefix_sums<<<ps_blocks, 256, 0, st>>>(bd->parts, n);
    HIPCHK(ctxs[0], hipGetLastError());
    k_mpart_out<<<n, 256, 0, st>>>(bd->parts);
    HIPCHK(ctxs[0], hipGetLastError());
    HIPCHK(ctxs[0], hipStreamSynchronize(st));
    return LBN_OK;
}

int lbn::dev_challenges(lbn_ctx *ctx, const ChalArg &arg, lbn_challenge *out) {
    HIPCHK(ctx, hipSetDevice(ctx->device));
    DeviceState *d = ctx->dev;
    static_assert(sizeof(ChalArg) <= 8192, "mailbox region");
    memcpy(d->hbox + mb::CH, &arg, sizeof(ChalArg));
    HIPCHK(ctx, hipMemcpyAsync(d->chal, d->hbox + mb::CH, sizeof(ChalArg), hipMemcpyHostToDevice, d->stream));
    // the challenges go straight into the mailbox
    k_challenges<<<COLUMNS / chal::WAVES, 64 * chal::WAVES, 0, d->stream>>>(
        d->chal, (lbn_challenge *)(d->hbox_dev + mb::CHOUT));
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(d->stream));
    memcpy(out, d->hbox + mb::CHOUT, COLUMNS * sizeof(lbn_challenge));
    return LBN_OK;
}

// ----------------------------------------------------------------------------- timing

extern "C" int lbn_set_graphs(lbn_ctx *ctx, int enable) {
    if (!ctx) return LBN_E_ARG;
    ctx->dev->graphs = enable != 0;
    return LBN_OK;
}

extern "C" int lbn_set_timing(lbn_ctx *ctx, int enable) {
    if (!ctx) return LBN_E_ARG;
    ctx->timing = enable != 0;
    return LBN_OK;
}
extern "C" int lbn_reset_timing(lbn_ctx *ctx) {
    if (!ctx) return LBN_E_ARG;
    ctx->times = lbn_timing{};
    return LBN_OK;
}
extern "C" int lbn_get_timing(const lbn_ctx *ctx, lbn_timing *out) {
    if (!ctx || !out) return LBN_E_ARG;
    *out = ctx->times;
    return LBN_OK;
}

// ============================================================================= standalone commitment
// lbn_cmt: the commitment of `columns` witness columns of `elements` binary ring elements (four
// words each, column-major like the prover's table) under a key of `rows` ring rows, row r's
// element m being element r * elements + m of the matrix expansion for (Q, elements).

namespace {
struct CmtEv {
    int kind;
    hipEvent_t a, b;
};
}  // namespace

struct lbn_cmt {
    int device = 0, columns = 0, elements = 0, rows = 0, nprimes = 0, width = 0;
    u128 q = 0;
    std::string err;
    hipStream_t stream = nullptr;
    uint4 *table = nullptr;
    uint32_t *tabs = nullptr;  // MAX_PRIMES * tab::SIZE
    uint32_t *ahat = nullptr;  // rows == 1: [k][m][648] (rows > 1: [k][m][648][16] during set-up)
    uint4 *amf = nullptr;      // rows > 1: the key's MFMA operands [k][slot][64-element block][digit][lane]
    uint32_t *V = nullptr;     // rows > 1: transformed witness of one chunk, [m][648][columns]
    uint32_t *cacc = nullptr;  // MAC slices [k][slice][column][648], or product slices [k][slice][648][16][columns]
    uint32_t *crow = nullptr;  // summed slices [k][row * columns + column][648]
    uint8_t *out = nullptr;    // rows * columns * 648 * width
    PrimeInfo primes[MAX_PRIMES] = {};
    Roots18 roots[MAX_PRIMES] = {};
    int splits = 0, m_per = 0;                 // rows == 1: MAC slices
    int chunk = 0, nmblk = 0, nslices = 0;     // rows > 1: elements per chunk, 64-element blocks, slices
    bool loaded = false, timing = false;
    std::vector<CmtEv> evs;
    size_t ev_used = 0;
    lbn_cmt_timing times{};
    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
};

namespace {
constexpr int CMT_CHUNK = 1024;  // ring elements per transformed chunk (columns x 2.6 KB each)
constexpr int CMT_VM_PER = 4;    // ring elements per block of k_commit_v
constexpr int CMT_MF_PER = 1024;  // ring elements of a chunk per block of k_commit_gemm_mfma (a whole chunk)
constexpr int CMT_MF_NT = 2;      // its column tiles of 16 (columns are a multiple of 64)
static_assert(CMT_CHUNK % 64 == 0 && CMT_MF_PER % 64 == 0 && CMT_MF_PER <= 1024,
              "the product kernel steps over whole 64-element blocks, at most 1024 per block");
enum CmtKind { CK_TOTAL = 0, CK_TRANSFORM, CK_MAC, CK_SUM, CK_FINISH };

int cmt_begin(lbn_cmt *c, int kind) {
    if (!c->timing) return -1;
    if (c->ev_used == c->evs.size()) {
        CmtEv e{};
        if (hipEventCreate(&e.a) != hipSuccess || hipEventCreate(&e.b) != hipSuccess) return -1;
        c->evs.push_back(e);
    }
    CmtEv &e = c->evs[c->ev_used];
    e.kind = kind;
    hipEventRecord(e.a, c->stream);
    return (int)c->ev_used++;
}
void cmt_end(lbn_cmt *c, int i) {
    if (i >= 0) hipEventRecord(c->evs[(size_t)i].b, c->stream);
}
void cmt_free(lbn_cmt *c) {
    for (void *p : {(void *)c->table, (void *)c->tabs, (void *)c->ahat, (void *)c->amf, (void *)c->V,
                    (void *)c->cacc, (void *)c->crow, (void *)c->out})
        if (p) hipFree(p);
    for (CmtEv &e : c->evs) {
        hipEventDestroy(e.a);
        hipEventDestroy(e.b);
    }
    if (c->stream) hipStreamDestroy(c->stream);
}
}  // namespace

extern "C" size_t lbn_cmt_len(int columns, int rows, uint64_t q_lo, uint64_t q_hi) {
    const u128 q = ((u128)q_hi << 64) | q_lo;
    if (columns <= 0 || rows <= 0 || q < 2) return 0;
    return (size_t)rows * columns * N * modulus_width(q);
}

extern "C" int lbn_cmt_create(int device, int columns, int elements, int rows, uint64_t q_lo, uint64_t q_hi,
                              lbn_cmt **out) {
    if (!out) return LBN_E_ARG;
    *out = nullptr;
    if (rows < 1 || rows > 16 || elements < 1 || elements > 65536 || columns < TC || columns > 4096 ||
        columns % (rows == 1 ? TC : 64))
        return LBN_E_ARG;
    const u128 q = ((u128)q_hi << 64) | q_lo;
    if (q < (u128)34816965049ull || q >= ((u128)1 << 96)) return LBN_E_MODULUS;
    const int np = crt_prime_count(q, (size_t)elements);
    if (!np) return LBN_E_MODULUS;
    auto *c = new lbn_cmt();
    c->device = device, c->columns = columns, c->elements = elements, c->rows = rows;
    c->q = q, c->nprimes = np, c->width = modulus_width(q);
    const int rc = [&]() -> int {
        HIPCHK(c, hipSetDevice(device));
        HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        const size_t words = (size_t)columns * elements * 4;
        HIPCHK(c, hipMalloc(&c->table, words * 16));
        std::vector<uint32_t> all;
        if (!setup_primes(np, c->primes, c->roots, all)) return c->fail(LBN_E_ARG, "NTT table construction failed");
        HIPCHK(c, hipMalloc(&c->tabs, all.size() * 4));
        HIPCHK(c, hipMemcpyAsync(c->tabs, all.data(), all.size() * 4, hipMemcpyHostToDevice, c->stream));
        const size_t re = (size_t)elements;
        const int rstride = rows == 1 ? 0 : 16;
        HIPCHK(c, hipMalloc(&c->ahat, (size_t)np * re * N * (rows == 1 ? 1 : 16) * 4));
        if (rows > 1) HIPCHK(c, hipMemsetAsync(c->ahat, 0, (size_t)np * re * N * 16 * 4, c->stream));
        // the key, one row at a time: expanded, reduced and transformed on the device
        uint64_t *amat = nullptr;
        HIPCHK(c, hipMalloc(&amat, re * N * 16));
        for (int r = 0; r < rows; r++) {
            bool done = false;
            const hipError_t ex = ajtai_device(c->stream, q, re, (size_t)r * re, re, amat, &done);
            if (ex != hipSuccess || !done) {
                (void)hipFree(amat);
                return c->fail(LBN_E_DEVICE, ex != hipSuccess ? hipGetErrorString(ex) : "matrix expansion did not finish");
            }
            for (int k = 0; k < np; k++) {
                PC pc{c->primes[k].p, c->primes[k].pinv_neg};
                k_ahat<<<(unsigned)re, 64, 0, c->stream>>>(amat, c->tabs + (size_t)k * tab::SIZE, pc, c->roots[k],
                                                           c->primes[k].r2, q_lo, q_hi,
                                                           c->ahat + (size_t)k * re * N * (rows == 1 ? 1 : 16), r,
                                                           rstride);
                HIPCHK(c, hipGetLastError());
            }
            HIPCHK(c, hipStreamSynchronize(c->stream));  // amat and A are reused by the next row
        }
        HIPCHK(c, hipFree(amat));
        if (rows == 1) {
            c->m_per = mac_m_per(re, columns / TC);
            c->splits = mac_splits(re, columns / TC);
            HIPCHK(c, hipMalloc(&c->cacc, (size_t)np * c->splits * columns * N * 4));
            HIPCHK(c, hipMalloc(&c->crow, (size_t)np * columns * N * 4));
        } else {
            // the key rows' digits in the matrix cores' operand order; the chunks start on 64-element
            // blocks (CMT_CHUNK is a multiple of 64)
            c->nmblk = (elements + 63) / 64;
            const size_t per = (size_t)N * c->nmblk * 4 * 64;
            HIPCHK(c, hipMalloc(&c->amf, (size_t)np * per * 16));
            for (int k = 0; k < np; k++) {
                k_ahat_mfma<<<dim3(N, c->nmblk), 64, 0, c->stream>>>(c->ahat + (size_t)k * re * N * 16, elements,
                                                                      c->nmblk, c->amf + (size_t)k * per);
                HIPCHK(c, hipGetLastError());
            }
            HIPCHK(c, hipStreamSynchronize(c->stream));
            HIPCHK(c, hipFree(c->ahat));
            c->ahat = nullptr;
            c->chunk = std::min(elements, CMT_CHUNK);
            c->nslices = (c->chunk + CMT_MF_PER - 1) / CMT_MF_PER;
            // whole 64-element steps of the product kernel stay inside V
            HIPCHK(c, hipMalloc(&c->V, (size_t)((c->chunk + 63) / 64 * 64) * N * columns * 4));
            HIPCHK(c, hipMemsetAsync(c->V, 0, (size_t)((c->chunk + 63) / 64 * 64) * N * columns * 4, c->stream));
            HIPCHK(c, hipMalloc(&c->cacc, (size_t)np * c->nslices * N * 16 * columns * 4));
            HIPCHK(c, hipMalloc(&c->crow, (size_t)np * rows * columns * N * 4));
        }
        HIPCHK(c, hipMalloc(&c->out, (size_t)rows * columns * N * c->width));
        return LBN_OK;
    }();
    if (rc != LBN_OK) {
        cmt_free(c);
        delete c;
        return rc;
    }
    *out = c;
    return LBN_OK;
}

extern "C" void lbn_cmt_destroy(lbn_cmt *c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    cmt_free(c);
    delete c;
}

extern "C" const char *lbn_cmt_error(const lbn_cmt *c) { return c ? c->err.c_str() : "null commitment context"; }
extern "C" int lbn_cmt_primes(const lbn_cmt *c) { return c ? c->nprimes : 0; }

extern "C" int lbn_cmt_load(lbn_cmt *c, const void *host_words, size_t bytes) {
    if (!c || !host_words) return LBN_E_ARG;
    if (bytes != (size_t)c->columns * c->elements * 64) return c->fail(LBN_E_ARG, "witness has the wrong length");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->table, host_words, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->loaded = true;
    return LBN_OK;
}

extern "C" int lbn_cmt_load_device(lbn_cmt *c, const void *device_words, size_t bytes) {
    if (!c || !device_words) return LBN_E_ARG;
    if (bytes != (size_t)c->columns * c->elements * 64) return c->fail(LBN_E_ARG, "witness has the wrong length");
    if ((uintptr_t)device_words % 16) return c->fail(LBN_E_ARG, "device witness must be 16-byte aligned");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t words = bytes / 16;
    k_copy_u4<<<(unsigned)std::min<size_t>((words + 255) / 256, 65536), 256, 0, c->stream>>>(
        (const uint4 *)device_words, c->table, words);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->loaded = true;
    return LBN_OK;
}

extern "C" int lbn_cmt_generate(lbn_cmt *c, const uint8_t seed[32]) {
    if (!c || !seed) return LBN_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const B3Root root = table_root(seed);
    const size_t blocks = (size_t)c->columns * c->elements;  // 64 bytes per ring element
    k_table_xof<<<(unsigned)((blocks + 255) / 256), 256, 0, c->stream>>>(root, (uint32_t *)c->table, blocks);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->loaded = true;
    return LBN_OK;
}

extern "C" int lbn_cmt_launch(lbn_cmt *c) {
    if (!c) return LBN_E_ARG;
    if (!c->loaded) return c->fail(LBN_E_STATE, "no witness loaded");
    HIPCHK(c, hipSetDevice(c->device));
    const int np = c->nprimes, cols = c->columns, re = c->elements;
    const size_t clen = (size_t)re * 4;
    const int tot = cmt_begin(c, CK_TOTAL);
    CrtArg crt;
    const uint32_t *fin = c->crow;  // the finish kernel's input: one slice per prime
    if (c->rows == 1) {
        const int e = cmt_begin(c, CK_MAC);
        for (int k = 0; k < np; k++) {
            const PC pc{c->primes[k].p, c->primes[k].pinv_neg};
            k_commit_mac<<<dim3(cols / TC, c->splits), MAC_THREADS, 0, c->stream>>>(
                (const uint32_t *)c->table, clen, c->m_per, re, c->ahat + (size_t)k * re * N,
                c->tabs + (size_t)k * tab::SIZE, pc, (uint32_t)((1ull << 32) % pc.p), c->primes[k].r2, c->roots[k],
                c->cacc + (size_t)k * c->splits * cols * N);
            HIPCHK(c, hipGetLastError());
        }
        cmt_end(c, e);
        crt = make_crt(c->primes, c->roots, np, c->q, c->width, 1, cols);
        const size_t n = (size_t)cols * N;
        const int es = cmt_begin(c, CK_SUM);
        k_commit_slicesum<<<dim3((unsigned)((n + 255) / 256), np), 256, 0, c->stream>>>(c->cacc, c->splits, n, crt,
                                                                                        c->crow);
        cmt_end(c, es);
        HIPCHK(c, hipGetLastError());
    } else {
        const size_t slice_words = (size_t)N * 16 * cols;
        const size_t per = (size_t)N * c->nmblk * 4 * 64;
        for (int k = 0; k < np; k++) {
            const PC pc{c->primes[k].p, c->primes[k].pinv_neg};
            MfWeights wt;
            for (uint64_t c8 = 0, x = 1; c8 < 7; c8++, x = (x << 8) % pc.p) wt.w[c8] = (uint32_t)((x << 32) % pc.p);
            uint32_t *slices = c->cacc + (size_t)k * c->nslices * slice_words;
            for (int mb = 0; mb < re; mb += c->chunk) {
                const int mc = std::min(c->chunk, re - mb);
                int e = cmt_begin(c, CK_TRANSFORM);
                k_commit_v<<<dim3(cols / TC, (mc + CMT_VM_PER - 1) / CMT_VM_PER), MAC_THREADS, 0, c->stream>>>(
                    (const uint32_t *)c->table, clen, mb, mb + mc, CMT_VM_PER, c->tabs + (size_t)k * tab::SIZE, pc,
                    c->roots[k], c->V, cols);
                cmt_end(c, e);
                HIPCHK(c, hipGetLastError());
                const int gy = (mc + CMT_MF_PER - 1) / CMT_MF_PER;
                e = cmt_begin(c, CK_MAC);
                k_commit_gemm_mfma<CMT_MF_NT><<<dim3(N, gy, cols / (16 * CMT_MF_NT)), 64, 0, c->stream>>>(
                    c->amf + (size_t)k * per, c->nmblk, mb / 64, c->V, cols, mc, CMT_MF_PER, pc, wt, mb > 0, slices);
                cmt_end(c, e);
                HIPCHK(c, hipGetLastError());
            }
            const int e = cmt_begin(c, CK_SUM);
            k_commit_rowsum<<<dim3((N + 15) / 16, c->rows * cols / 64), 256, 0, c->stream>>>(
                slices, c->nslices, cols, pc, c->primes[k].r2, c->crow + (size_t)k * c->rows * cols * N);
            cmt_end(c, e);
            HIPCHK(c, hipGetLastError());
        }
        crt = make_crt(c->primes, c->roots, np, c->q, c->width, 1, c->rows * cols);
    }
    const int e = cmt_begin(c, CK_FINISH);
    k_commit_finish<<<(unsigned)(c->rows * cols), FIN_THREADS, 0, c->stream>>>(fin, c->tabs, crt, c->out);
    cmt_end(c, e);
    HIPCHK(c, hipGetLastError());
    cmt_end(c, tot);
    return LBN_OK;
}

extern "C" int lbn_cmt_read(lbn_cmt *c, uint8_t *out, size_t len) {
    if (!c || !out) return LBN_E_ARG;
    const size_t need = (size_t)c->rows * c->columns * N * c->width;
    if (len < need) return c->fail(LBN_E_ARG, "output buffer too small");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(out, c->out, need, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBN_OK;
}

extern "C" int lbn_cmt_set_timing(lbn_cmt *c, int enable) {
    if (!c) return LBN_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->timing = enable != 0;
    c->ev_used = 0;
    c->times = lbn_cmt_timing{};
    return LBN_OK;
}

extern "C" int lbn_cmt_get_timing(lbn_cmt *c, lbn_cmt_timing *t) {
    if (!c || !t) return LBN_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < c->ev_used; i++) {
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, c->evs[i].a, c->evs[i].b));
        switch (c->evs[i].kind) {
            case CK_TOTAL: c->times.total_ms += ms, c->times.runs++; break;
            case CK_TRANSFORM: c->times.transform_ms += ms; break;
            case CK_MAC: c->times.mac_ms += ms; break;
            case CK_SUM: c->times.sum_ms += ms; break;
            default: c->times.finish_ms += ms; break;
        }
    }
    c->ev_used = 0;
    *t = c->times;
    return LBN_OK;
}
