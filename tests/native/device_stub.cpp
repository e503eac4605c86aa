// A stand-in for the device half of liblabinius_gpu (lbn_device.hip) for host-only builds under the
// sanitizers: every device entry point succeeds without a GPU and fills its outputs with
// deterministic values derived from its inputs, so that the host code (the worker pool, the
// scheduler, the transcript, the switch's host rounds, the proof encoding and the checker's
// decoding) runs its real paths. It models nothing of the GPU's results: proofs made through it
// are not valid proofs. Test infrastructure only.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../labinius-fast-challenge_amd/csrc/lbn_internal.h"

struct lbn::DeviceState {
    std::vector<uint8_t> commitment;
    // the batch protocol's staged inputs and results
    lbn::B3State ts{};
    std::vector<uint64_t> rounds = std::vector<uint64_t>(6 * 64), coords = std::vector<uint64_t>(3 * 64);
    lbn::SwitchHandoff ho;
    std::vector<uint64_t> p0, row;
    lbn::ChalArg arg{};
    std::vector<lbn_challenge> ch;
    const lbn_challenge *fold_ch = nullptr;
    struct {
        size_t off = 0, count = 0, head = 0, tail = 0;
        uint64_t first = 0;
        lbn::Blake3::Subtrees trees;
        uint32_t key[8] = {}, flags = 0;
        uint8_t *dst = nullptr;
    } cv;
    std::vector<uint64_t> r_hi, partials;
};

namespace {
uint64_t mix(uint64_t x) {  // splitmix64 finaliser
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
}  // namespace

namespace lbn {
int dev_init(lbn_ctx *ctx) {
    ctx->dev = new DeviceState();
    return LBN_OK;
}
void dev_free(lbn_ctx *ctx) {
    delete ctx->dev;
    ctx->dev = nullptr;
}
int dev_commit_launch(lbn_ctx *ctx) {
    if (!ctx->table_loaded || !ctx->modulus_set) return ctx->fail(LBN_E_STATE, "table and modulus must be loaded");
    auto &c = ctx->dev->commitment;
    c.resize((size_t)COLUMNS * N * ctx->width);
    for (size_t i = 0; i < c.size(); i++) c[i] = (uint8_t)mix(i + 7 * (uint64_t)ctx->shape.log_vars);
    return LBN_OK;
}
int dev_commit_cvs(lbn_ctx *ctx, size_t off, uint64_t first, size_t count, const Blake3::Subtrees &trees,
                   const uint32_t key[8], uint32_t flags, uint32_t *cvs, uint8_t *dst, size_t head, size_t tail) {
    (void)off, (void)count;
    for (size_t t = 0; t < trees.size(); t++)
        for (int k = 0; k < 8; k++) cvs[8 * t + k] = (uint32_t)mix(trees[t].first * 64 + trees[t].second + k) ^ key[k] ^ flags;
    const auto &c = ctx->dev->commitment;
    memcpy(dst, c.data(), head);
    memcpy(dst + tail, c.data() + tail, c.size() - tail);
    (void)first;
    return LBN_OK;
}
int dev_commit_bytes(lbn_ctx *ctx, uint8_t *dst) {
    memcpy(dst, ctx->dev->commitment.data(), ctx->dev->commitment.size());
    return LBN_OK;
}
int dev_verify_products(lbn_ctx *ctx, u128 q, const int16_t *v, const uint8_t *commitment, int width,
                        const lbn_challenge *ch, __int128 *lhs, __int128 *rhs) {
    (void)q, (void)width, (void)ch;
    for (int i = 0; i < 3 * (2 * N - 1); i++) lhs[i] = rhs[i] = (__int128)(v[i % (N * ctx->shape.ring_elements)] + commitment[i]);
    return LBN_OK;
}
void dev_quiesce(lbn_ctx *) {}
int dev_switch_prepare(lbn_ctx *ctx, const uint64_t *batch) {
    if (!ctx->partials_done) return ctx->fail(LBN_E_STATE, "lbn_partials must run first");
    ctx->lmap.assign((const F162 *)batch, (const F162 *)batch + 128);
    ctx->round = 0;
    return LBN_OK;
}
int dev_switch_rounds(lbn_ctx *ctx, B3State &ts, uint64_t *rounds, uint64_t *coords, SwitchHandoff &ho) {
    if (ctx->round != 0) return ctx->fail(LBN_E_STATE, "lbn_switch_begin must run first");
    const int lp = ctx->shape.log_packed, r0 = std::max(7, lp - SWITCH_HOST_STEPS);
    const uint64_t top = (1ull << 34) - 1;
    for (int r = 0; r < r0; r++) {
        for (int k = 0; k < 6; k++) rounds[6 * r + k] = mix(ts.counter * 97 + 6 * r + k) & (k % 3 == 2 ? top : ~0ull);
        for (int k = 0; k < 3; k++) coords[3 * r + k] = mix(1000 + 3 * r + k) & (k == 2 ? top : ~0ull);
    }
    const size_t h2 = (size_t)1 << (lp - r0 - 1);
    ho.r0 = r0;
    ho.rho = F162{{mix(1), mix(2), mix(3) & top}};
    ho.L.assign(ctx->lmap.begin(), ctx->lmap.end());
    ho.p.resize(4 * h2);
    for (size_t i = 0; i < ho.p.size(); i++) ho.p[i] = F162{{mix(5 * i), mix(5 * i + 1), mix(5 * i + 2) & top}};
    ho.e.resize(2 * h2);
    for (size_t i = 0; i < ho.e.size(); i++) ho.e[i] = B128{{mix(7 * i), mix(7 * i + 3)}};
    ctx->round = lp;
    return LBN_OK;
}
int dev_sync(lbn_ctx *) { return LBN_OK; }
void dev_batch_quiesce(int) {}
int dev_cvs_stage(lbn_ctx *ctx, const Blake3::Subtrees &trees, size_t off, uint64_t first, size_t count,
                  const uint32_t key[8], uint32_t flags, uint8_t *dst, size_t head, size_t tail) {
    auto &c = ctx->dev->cv;
    c.off = off, c.first = first, c.count = count, c.head = head, c.tail = tail, c.trees = trees, c.flags = flags;
    c.dst = dst;
    memcpy(c.key, key, 32);
    return LBN_OK;
}
int dev_cvs_batch(lbn_ctx *const *, int) { return LBN_OK; }
void dev_cvs_collect(lbn_ctx *ctx, uint32_t *cvs, uint8_t *dst) {
    auto &c = ctx->dev->cv;
    dev_commit_cvs(ctx, c.off, c.first, c.count, c.trees, c.key, c.flags, cvs, dst, c.head, c.tail);
}
int dev_partials_prepare(lbn_ctx *ctx, const uint64_t *r_hi, bool) {
    ctx->dev->r_hi.assign(r_hi, r_hi + 2 * ctx->shape.log_packed);
    return LBN_OK;
}
int dev_partials_batch(lbn_ctx *const *ctxs, int n) {
    for (int i = 0; i < n; i++) {
        DeviceState *d = ctxs[i]->dev;
        d->partials.resize(256);
        const int rc = lbn_partials(ctxs[i], d->r_hi.data(), d->partials.data());
        if (rc != LBN_OK) return rc;
    }
    return LBN_OK;
}
void dev_partials_collect(lbn_ctx *ctx, uint64_t *partials) {
    memcpy(partials, ctx->dev->partials.data(), 256 * 8);
}
void dev_switch_stage(lbn_ctx *ctx, const B3State &ts) { ctx->dev->ts = ts; }
int dev_switch_batch(lbn_ctx *const *ctxs, int n) {
    for (int i = 0; i < n; i++) {
        DeviceState *d = ctxs[i]->dev;
        const int rc = dev_switch_rounds(ctxs[i], d->ts, d->rounds.data(), d->coords.data(), d->ho);
        if (rc != LBN_OK) return rc;
    }
    return LBN_OK;
}
void dev_switch_collect(lbn_ctx *ctx, B3State &ts, uint64_t *rounds, uint64_t *coords, SwitchHandoff &ho) {
    DeviceState *d = ctx->dev;
    const int r0 = std::max(7, ctx->shape.log_packed - SWITCH_HOST_STEPS);
    ts = d->ts;
    memcpy(rounds, d->rounds.data(), (size_t)r0 * 48);
    memcpy(coords, d->coords.data(), (size_t)r0 * 24);
    ho = d->ho;
}
int dev_challenges(lbn_ctx *, const ChalArg &arg, lbn_challenge *out) {
    for (int j = 0; j < COLUMNS; j++) {
        lbn_challenge &c = out[j];
        memset(&c, 0, sizeof c);
        c.weight = WEIGHT;
        for (int i = 0; i < WEIGHT; i++) c.positions[i] = (uint8_t)(5 * i + (j + arg.ts.counter) % 5);
        c.signs = (uint32_t)mix(j) & ((1u << WEIGHT) - 1);
    }
    return LBN_OK;
}
void dev_challenges_stage(lbn_ctx *ctx, const ChalArg &arg) { ctx->dev->arg = arg; }
int dev_challenges_batch(lbn_ctx *const *ctxs, int n) {
    for (int i = 0; i < n; i++) {
        DeviceState *d = ctxs[i]->dev;
        d->ch.resize(COLUMNS);
        dev_challenges(ctxs[i], d->arg, d->ch.data());
    }
    return LBN_OK;
}
void dev_challenges_collect(lbn_ctx *ctx, lbn_challenge *out) {
    memcpy(out, ctx->dev->ch.data(), COLUMNS * sizeof(lbn_challenge));
}
void dev_use_streams(lbn_ctx *, void *, void *) {}
int dev_row_prepare(lbn_ctx *ctx, const uint64_t *p0, bool) {
    ctx->dev->p0.assign(p0, p0 + 3 * ctx->shape.row_log);
    return LBN_OK;
}
int dev_row_batch(lbn_ctx *const *ctxs, int n) {
    for (int i = 0; i < n; i++) {
        DeviceState *d = ctxs[i]->dev;
        d->row.resize(3 * COLUMNS);
        const int rc = lbn_row_eval(ctxs[i], d->p0.data(), ctxs[i]->shape.row_log, d->row.data());
        if (rc != LBN_OK) return rc;
    }
    return LBN_OK;
}
void dev_row_collect(lbn_ctx *ctx, uint64_t *row) { memcpy(row, ctx->dev->row.data(), 24 * COLUMNS); }
int dev_fold_prepare(lbn_ctx *ctx, const lbn_challenge *ch, int n_challenges, int32_t *out, size_t n_out,
                     FoldPlan &plan, bool) {
    if (n_challenges != COLUMNS || n_out != (size_t)N * ctx->shape.ring_elements) return LBN_E_ARG;
    ctx->dev->fold_ch = ch;
    plan.out = out;
    plan.n_out = n_out;
    return LBN_OK;
}
int dev_fold_batch(lbn_ctx *const *ctxs, int n, const FoldPlan *plans) {
    for (int i = 0; i < n; i++) {
        const int rc = lbn_fold(ctxs[i], ctxs[i]->dev->fold_ch, COLUMNS, plans[i].out, plans[i].n_out);
        if (rc != LBN_OK) return rc;
    }
    return LBN_OK;
}
int dev_fold_collect(lbn_ctx *, const FoldPlan &) { return LBN_OK; }
}  // namespace lbn

extern "C" int lbn_load_table(lbn_ctx *ctx, const void *host_bytes, size_t bytes) {
    if (!ctx || !host_bytes) return LBN_E_ARG;
    if (bytes != ctx->shape.table_bytes) return ctx->fail(LBN_E_ARG, "witness table has the wrong length");
    ctx->table_loaded = true;
    ctx->partials_done = false;
    return LBN_OK;
}
extern "C" int lbn_generate_table(lbn_ctx *ctx, const uint8_t seed[32]) {
    if (!ctx || !seed) return LBN_E_ARG;
    ctx->table_loaded = true;
    ctx->partials_done = false;
    return LBN_OK;
}
extern "C" int lbn_set_modulus(lbn_ctx *ctx, uint64_t q_lo, uint64_t q_hi) {
    if (!ctx) return LBN_E_ARG;
    const u128 q = ((u128)q_hi << 64) | q_lo;
    if (q < (u128)34816965049ull || q >= ((u128)1 << 96)) return ctx->fail(LBN_E_MODULUS, "modulus out of range");
    int bits = 0;
    for (u128 x = q - 1; x; x >>= 1) bits++;
    ctx->q = q;
    ctx->width = (bits + 7) / 8;
    ctx->modulus_set = true;
    return LBN_OK;
}
extern "C" size_t lbn_ctx_footprint(int log_vars, uint64_t q_lo, uint64_t q_hi) {
    lbn::Shape s;
    const u128 q = ((u128)q_hi << 64) | q_lo;
    if (!lbn::Shape::of(log_vars, s) || q < (u128)34816965049ull || q >= ((u128)1 << 96)) return 0;
    return 2 * s.table_bytes + ((size_t)64 << 20);
}
extern "C" int lbn_partials(lbn_ctx *ctx, const uint64_t *r_hi, uint64_t *partials) {
    if (!ctx || !r_hi || !partials) return LBN_E_ARG;
    if (!ctx->table_loaded) return ctx->fail(LBN_E_STATE, "no table loaded");
    const int lp = ctx->shape.log_packed;
    ctx->r_hi.assign((const lbn::B128 *)r_hi, (const lbn::B128 *)r_hi + lp);
    for (int i = 0; i < 256; i++) partials[i] = mix(r_hi[i % (2 * lp)] + i);
    ctx->partials_done = true;
    ctx->round = -1;
    return LBN_OK;
}
extern "C" int lbn_row_eval(lbn_ctx *ctx, const uint64_t *p0, int n_p0, uint64_t *row) {
    if (!ctx || !p0 || !row || n_p0 != ctx->shape.row_log) return LBN_E_ARG;
    for (int c = 0; c < 3 * lbn::COLUMNS; c++) row[c] = mix(p0[c % (3 * n_p0)] + c) & (c % 3 == 2 ? (1ull << 34) - 1 : ~0ull);
    return LBN_OK;
}
extern "C" int lbn_fold(lbn_ctx *ctx, const lbn_challenge *ch, int n_challenges, int32_t *out, size_t n_out) {
    if (!ctx || !ch || !out || n_challenges != lbn::COLUMNS) return LBN_E_ARG;
    if (n_out != (size_t)lbn::N * ctx->shape.ring_elements) return ctx->fail(LBN_E_ARG, "fold buffer has the wrong length");
    for (size_t i = 0; i < n_out; i++) out[i] = (int32_t)(mix(i + ch[i % 1024].signs) % 201) - 100;
    return LBN_OK;
}

// The scheduler's worker streams.
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *stream, unsigned int) {
    *stream = (hipStream_t) new int(0);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t stream) {
    delete (int *)stream;
    return hipSuccess;
}
