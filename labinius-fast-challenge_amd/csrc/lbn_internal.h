// Internal declarations shared by the device (.hip) and host (.cpp) halves of liblabinius_gpu.
#pragma once
#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/labinius_gpu.h"
#include "lbn_blake3.h"

typedef unsigned __int128 u128;

namespace lbn {

constexpr int N = 648;           // ring degree of R_Q = Z_Q[X]/(X^648 - X^324 + 1)
constexpr int HALF_N = 324;      // the middle exponent X^324
constexpr int PACK = 128;        // bits per GHASH word
constexpr int COLUMN_LOG = 10;   // columns = 1024 (verifier/proof_verify/src/main.rs:37)
constexpr int COLUMNS = 1 << COLUMN_LOG;
constexpr int WEIGHT = 28;       // challenge weight (pcs/src/challenge.rs:56)
constexpr int MAX_PRIMES = 5;    // CRT primes (below 2^32/6) for the exact commitment

struct Shape {
    int log_vars = 0, log_packed = 0, row_log = 0;
    size_t n_words = 0, column_len = 0, ring_elements = 0, table_bytes = 0;
    static bool of(int log_vars, Shape &s);
};

// GF(2^128) (GHASH basis) and F162 = GF(2)[x]/(x^162 + x^81 + 1) on the host, PCLMUL based.
struct B128 {
    uint64_t w[2];
};
struct F162 {
    uint64_t w[3];
};
B128 b128_mul(B128 a, B128 b);
F162 f162_mul(F162 a, F162 b);
inline B128 b128_add(B128 a, B128 b) { return B128{{a.w[0] ^ b.w[0], a.w[1] ^ b.w[1]}}; }
inline F162 f162_add(F162 a, F162 b) { return F162{{a.w[0] ^ b.w[0], a.w[1] ^ b.w[1], a.w[2] ^ b.w[2]}}; }
inline F162 f162_lift(B128 a) { return F162{{a.w[0], a.w[1], 0}}; }
inline bool f162_eq(F162 a, F162 b) { return a.w[0] == b.w[0] && a.w[1] == b.w[1] && a.w[2] == b.w[2]; }
// eq expansions (crossfield.rs:66-97: first coordinate = most significant index bit) and the
// LSB-first eq table (eval.rs:51-65).
void eq_expand_b128(const B128 *r, int n, B128 *out);
void eq_expand_f162(const F162 *r, int n, F162 *out);
void eq_table_f162(const F162 *p, int n, F162 *out);

// One CRT prime of the commitment transform.
struct PrimeInfo {
    uint32_t p, pinv_neg, r2;  // -p^-1 mod 2^32, R^2 mod p (R = 2^32)
    uint32_t omega;            // primitive 1944-th root of unity
};

// A BLAKE3 hasher's whole state plus a transcript's sample counter, as plain data: the transcript
// moves to the device for the switch rounds and back (Blake3::save_state / load_state).
struct B3State {
    uint32_t key[8];
    uint32_t flags;
    int32_t depth;
    uint32_t stack[54][8];
    uint32_t cv[8];
    uint64_t chunk;
    int32_t blocks;
    uint32_t buf_len;
    uint8_t buf[64];
    uint64_t counter;
};

// Everything the device needs to sample the folding challenges exactly as the host sampler does:
// the transcript state after the row evaluation, the sign hasher's key and flags, the norm bound
// and the host's cos/sin tables of the 243rd roots (so both sides add the same doubles).
struct ChalArg {
    B3State ts;
    uint32_t sign_key[8];
    uint32_t sign_flags;
    uint32_t pad;
    double bound_sq;
    double cosv[243], sinv[243];
    int32_t units[81];
};

// Layout of one prime's commitment-transform constants (u32 offsets), built by prime_tables().
namespace tab {
constexpr int TWB = 0;                 // [2][324][4] binary split+twist values: 0, b, eta b, (1+eta) b
constexpr int TW = TWB + 2 * 324 * 4;  // [2][324] beta_h^i (Montgomery)
constexpr int ETA = TW + 2 * 324;      // [2] eta_h (Montgomery)
constexpr int UNTW = ETA + 2;          // [2][324] beta_h^-i / 324 (Montgomery)
constexpr int ETA0 = UNTW + 2 * 324;   // eta_0 (Montgomery)
constexpr int DINV = ETA0 + 1;         // 1 / (eta_0 - eta_1) (Montgomery)
constexpr int TWPF = DINV + 1;         // [18][18] rho^(n2 k1) (Montgomery), rho = w^6
constexpr int TWPI = TWPF + 324;       // [18][18] rho^-(n2 k1)
constexpr int R18F = TWPI + 324;       // [9] w18^j, w18 = rho^18 (Montgomery)
constexpr int R18I = R18F + 9;         // [9] w18^-j
// The commitment's pass A over a binary ring element: x[n1] = TWB[h][18 n1 + n2][2 bits]
// = c(bits) beta_h^(18 n1) * beta_h^n2, so the DFT over n1 is that of x'[n1] = c(bits)
// beta_h^(18 n1), the same for every n2, scaled by beta_h^n2.
// ST1 [9 j][2 o][2 h][16 s]: its first radix-2 stage, o = 0 giving x'[j] + x'[j+9] and o = 1
// giving (x'[j] - x'[j+9]) w18^j, for the 4 bits s = (bits of n1 = j) | (bits of n1 = j+9) << 2.
// Plain residues in [0, p).
// TWPB [2 h][18 n2][18 k1]: beta_h^n2 rho^(n2 k1) (Montgomery), the pass-A twist with that scale.
constexpr int ST1 = R18I + 9;
constexpr int ST1_SIZE = 9 * 2 * 2 * 16;
constexpr int TWPB = ST1 + ST1_SIZE;
constexpr int SIZE = TWPB + 2 * 324;
}  // namespace tab

// Device-side state (owned by lbn_ctx; defined in lbn_device.hip).
struct DeviceState;

}  // namespace lbn

struct lbn_ctx {
    int device = 0;
    lbn::Shape shape;
    std::string err;
    bool table_loaded = false;
    bool modulus_set = false;
    u128 q = 0;
    int width = 0;
    int nprimes = 0;
    lbn::PrimeInfo primes[lbn::MAX_PRIMES];
    // switch state kept on the host: r_hi, the linear map a = L(E) and the round index.
    std::vector<lbn::B128> r_hi;
    std::vector<lbn::F162> lmap;  // 128 entries
    int round = -1;
    bool partials_done = false;
    lbn::F162 opened{};
    lbn::DeviceState *dev = nullptr;
    bool timing = false;
    lbn_timing times{};
    int fail(int code, const std::string &msg) {
        err = msg;
        return code;
    }
};

namespace lbn {
// Device entry points used by the host prover (lbn_device.hip).
int dev_init(lbn_ctx *ctx);
void dev_free(lbn_ctx *ctx);
// The commitment kernels, leaving the packed commitment in device memory.
int dev_commit_launch(lbn_ctx *ctx);
// Chaining values of `count` whole BLAKE3 chunks of the device commitment from byte `off`, chunk
// indices first.., under key/flags, into cvs (8 words each); also copies the commitment bytes
// [0, head) and [tail, clen) to the same offsets of dst, and starts copying all of it to pinned
// host memory (dev_commit_bytes finishes that copy into dst).
int dev_commit_cvs(lbn_ctx *ctx, size_t off, uint64_t first, size_t count, const Blake3::Subtrees &trees,
                   const uint32_t key[8], uint32_t flags, uint32_t *cvs, uint8_t *dst, size_t head, size_t tail);
int dev_commit_bytes(lbn_ctx *ctx, uint8_t *dst);
// The checker's ring identity A v = sum_j c_j(-X^4) C_j (verifier/proof_verify/src/rq.rs:188-256):
// both sides' coefficients before the reduction into R_Q (1295 of them), each as three limb sums
// lhs/rhs[3 e + k] with value sum_k 2^(32 k) lhs[3 e + k], computed on the device from the public
// matrix of Q, the fold v (ring_elements x 648), the packed commitment and the 1024 challenges.
int dev_verify_products(lbn_ctx *ctx, u128 q, const int16_t *v, const uint8_t *commitment, int width,
                        const lbn_challenge *ch, __int128 *lhs, __int128 *rhs);
// Waits for everything queued on the context's streams (errors ignored): a failing lbn_prove calls
// it before returning, so no copy or kernel still writes the caller's proof buffer.
void dev_quiesce(lbn_ctx *ctx);
// The switch rounds after lbn_switch_begin, all but the last few on the device: control steps
// 0 .. r0 - 1 and the folds of rounds 0 .. r0 - 2, with the transcript (state ts in, advanced state
// out) absorbing each message and sampling each challenge there. The last SWITCH_HOST_STEPS control
// steps and the folds before them (at most 2^(SWITCH_HOST_STEPS - 1) items each) cost less on the
// host than their launches; the handoff carries what they start from. rounds: r0 x 2 F162, coords:
// r0 F162 (3 u64 each).
constexpr int SWITCH_HOST_STEPS = 7;
struct SwitchHandoff {
    int r0 = 0;              // first control step left to the host
    F162 rho{};              // the challenge of round r0 - 1
    std::vector<F162> L;     // the linear map L_{r0} (128 entries)
    std::vector<F162> p;     // the input of round r0 - 1's fold: 4 h2 values, h2 = 2^(lp - r0 - 1)
    std::vector<B128> e;     // and its eq values (2 h2)
};
int dev_switch_rounds(lbn_ctx *ctx, B3State &ts, uint64_t *rounds, uint64_t *coords, SwitchHandoff &ho);
// The batch protocol (lbn_prove_batch): up to SWB_MAX proofs on contexts of one device advance
// together. Between the batched launches, host threads run each proof's host steps and its
// *_stage / *_prepare call, then dev_sync (its stream idle); a *_batch call then launches for
// every instance at once on a stream of its own, each instance reading and writing only its own
// context's buffers, and returns when the device is done; each proof's *_collect takes its
// results from its mailbox. Instance i's results are those of the one-proof entry point.
constexpr int SWB_MAX = 64;
int dev_sync(lbn_ctx *ctx);
// After a batched call failed part-way: waits for whatever it launched on the device's batch stream.
void dev_batch_quiesce(int device);
// The commitment's hashing: dev_cvs_stage takes dev_commit_cvs's arguments (the commitment's
// copy to the host is queued there), dev_cvs_batch hashes every staged proof's chunks and subtrees,
// dev_cvs_collect gives what dev_commit_cvs gives (the subtree roots and dst's head and tail bytes).
int dev_cvs_stage(lbn_ctx *ctx, const Blake3::Subtrees &trees, size_t off, uint64_t first, size_t count,
                  const uint32_t key[8], uint32_t flags, uint8_t *dst, size_t head, size_t tail);
int dev_cvs_batch(lbn_ctx *const *ctxs, int n);
void dev_cvs_collect(lbn_ctx *ctx, uint32_t *cvs, uint8_t *dst);
// The partials: dev_partials_prepare computes the eq tables of r_hi (log_packed values) into the
// mailbox (and queues their upload with `upload`; the batch copies them itself),
// dev_partials_batch runs every proof's eq expansion and prefix sums, and dev_partials_collect
// gives the 128 partials (2 u64 each) that lbn_partials gives.
int dev_partials_prepare(lbn_ctx *ctx, const uint64_t *r_hi, bool upload);
int dev_partials_batch(lbn_ctx *const *ctxs, int n);
void dev_partials_collect(lbn_ctx *ctx, uint64_t *partials);
// The switch: dev_switch_stage (after dev_switch_prepare) writes the start state from ts to the
// mailbox; dev_switch_batch runs the device rounds; dev_switch_collect returns what
// dev_switch_rounds returns.
void dev_switch_stage(lbn_ctx *ctx, const B3State &ts);
int dev_switch_batch(lbn_ctx *const *ctxs, int n);
void dev_switch_collect(lbn_ctx *ctx, B3State &ts, uint64_t *rounds, uint64_t *coords, SwitchHandoff &ho);
// The switch's start for dev_switch_rounds: the map L_0 from the batch (no kernel; the rounds start
// from the prefix sets lbn_partials left).
int dev_switch_prepare(lbn_ctx *ctx, const uint64_t *batch);
// The map update of switch round r: L_{r+1}[k] = L_r((1 + c) x^k) + rho L_r[k], c = r_hi[r].
void lmap_next(std::vector<F162> &L, B128 c, const F162 &rho);
// The 1024 folding challenges sampled on the device from the transcript state in arg (the
// counter of arg->ts is the first challenge's); out receives them in column order.
int dev_challenges(lbn_ctx *ctx, const ChalArg &arg, lbn_challenge *out);
// The same in the batch protocol: the sampler's input staged in the mailbox, one launch (which
// copies every staged input first), the 1024 challenges collected in column order.
void dev_challenges_stage(lbn_ctx *ctx, const ChalArg &arg);
int dev_challenges_batch(lbn_ctx *const *ctxs, int n);
void dev_challenges_collect(lbn_ctx *ctx, lbn_challenge *out);
// The row evaluation in the batch protocol: dev_row_prepare computes the eq tables of p0 (row_log
// coordinates) into the mailbox (and queues their upload with `upload`; the batch kernels read
// them there); dev_row_batch evaluates every row in one launch per kernel; dev_row_collect gives
// the 1024 F162 values (3 u64 each) that lbn_row_eval gives.
int dev_row_prepare(lbn_ctx *ctx, const uint64_t *p0, bool upload);
int dev_row_batch(lbn_ctx *const *ctxs, int n);
void dev_row_collect(lbn_ctx *ctx, uint64_t *row);
// The fold in the batch protocol: dev_fold_prepare checks the challenges, builds the column lists
// per position in the mailbox (and queues their upload with `upload`; the batch copies them
// itself); dev_fold_batch folds every
// instance in one launch; dev_fold_collect checks the context's range flag and puts its fold in
// `out` (n_out int32) if the kernel did not write it there directly. lbn_fold runs the same parts
// on the context's own stream.
struct FoldPlan {
    int nb = 0;          // counter planes the longest list needs
    int n_entries = 0;   // list entries
    int32_t *out = nullptr;
    size_t n_out = 0;
    int32_t *dst = nullptr;  // where the kernel writes: `out` itself when pinned, else the context's staging
    size_t upload_bytes = 0; // the lists' bytes staged in the mailbox
};
int dev_fold_prepare(lbn_ctx *ctx, const lbn_challenge *ch, int n_challenges, int32_t *out, size_t n_out,
                     FoldPlan &plan, bool upload);
int dev_fold_batch(lbn_ctx *const *ctxs, int n, const FoldPlan *plans);
int dev_fold_collect(lbn_ctx *ctx, const FoldPlan &plan);
// The streams the context's work goes to (hipStream_t, as void *): a caller driving many contexts
// from a few threads gives each thread its own pair; null restores the context's own streams.
// Only between entry points (every entry point leaves its streams idle).
void dev_use_streams(lbn_ctx *ctx, void *main, void *copy);

// Host worker pool: f(i) for every i < n, on up to host_threads() threads (the caller included).
int host_threads();
void parallel_for(size_t n, const std::function<void(size_t)> &f);
}  // namespace lbn
