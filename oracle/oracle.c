/* ORACLE — TEST INFRASTRUCTURE ONLY. Never linked into, loaded by, or called from the product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
 *
 * A plain scalar restatement of the reference prover (verifier/proof_verify/src/reference.rs)
 * and checker (verifier/proof_verify/src/main.rs, rq.rs) plus the library semantics they call.
 * Each function cites the reference lines it follows. It favours the obvious formulation over
 * speed: i128 schoolbook ring products, bit-loop carry-less products, arrays for the sumcheck.
 *
 * Pins (see tests/): BLAKE3 against libLLVM's implementation (golden fixtures); field products
 * against their bit-by-bit definitions and the order-243 / Frobenius identities; the switch and
 * opening against the checker's own identities (switch verifier, `A v = sum c_j C_j`, parity,
 * claim) and brute force on tiny shapes; the proof layout against the 13,958,612-byte length at
 * the challenge shape (solution/proof_generate/src/main.rs:2335) and fold_cap 293,534,171,136 /
 * bound 541,787 (verifier/proof_verify/src/main.rs:476-480).
 */
#include "oracle.h"
#include "b3.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define N 648
#define SPLIT 324
#define PACK 128
#define COLUMN_LOG 10
#define WEIGHT 28
#define N162 162
#define CONDUCTOR 243

/* verifier/proof_verify/src/main.rs:54 */
static const u128 MIN_MODULUS = (u128)34816965049ull;
/* pcs/src/recursion/mod.rs:81 */
static const double FOLD_CAP = 108.0;

int orc_shape_of(int log_vars, orc_shape *s) {
    if (log_vars < 19 || log_vars > 31) return -1;
    s->log_vars = log_vars;
    s->log_packed = log_vars - 7;
    s->column_log = COLUMN_LOG;
    s->columns = 1 << COLUMN_LOG;
    s->row_log = s->log_packed - COLUMN_LOG;
    s->column_len = (size_t)1 << s->row_log;
    s->ring_elements = s->column_len / 4;
    s->n_words = (size_t)1 << s->log_packed;
    s->table_bytes = ((size_t)1 << log_vars) / 8;
    return 0;
}

/* =============================== fields ================================================== */

/* Carry-less 64 x 64 -> 128 product: PCLMULQDQ where the build enables it, else the shift-and-xor
 * definition. Both are checked against the bit-by-bit field products in tests/test_oracle.py. */
#if defined(__PCLMUL__) && defined(__SSE4_1__)
#include <immintrin.h>
static u128 clmul64(uint64_t a, uint64_t b) {
    __m128i r = _mm_clmulepi64_si128(_mm_cvtsi64_si128((long long)a), _mm_cvtsi64_si128((long long)b), 0);
    return ((u128)(uint64_t)_mm_extract_epi64(r, 1) << 64) | (uint64_t)_mm_cvtsi128_si64(r);
}
#else
static u128 clmul64(uint64_t a, uint64_t b) {
    u128 r = 0;
    for (int i = 0; i < 64; i++)
        if ((a >> i) & 1) r ^= (u128)b << i;
    return r;
}
#endif

/* GF(2^128) in the GHASH polynomial basis, modulo x^128 + x^7 + x^2 + x + 1
 * (pcs/src/fields/scalar.rs:4, 46-68). */
void orc_b128_mul(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]) {
    u128 ll = clmul64(a[0], b[0]), hh = clmul64(a[1], b[1]);
    u128 mid = clmul64(a[0], b[1]) ^ clmul64(a[1], b[0]);
    u128 lo = ll ^ (mid << 64);
    u128 hi = hh ^ (mid >> 64);
    /* hi * x^128 = hi * (x^7 + x^2 + x + 1); the bits shifted past 128 are folded once more. */
    u128 t = hi ^ (hi << 1) ^ (hi << 2) ^ (hi << 7);
    u128 o = (hi >> 127) ^ (hi >> 126) ^ (hi >> 121);
    t ^= o ^ (o << 1) ^ (o << 2) ^ (o << 7);
    lo ^= t;
    out[0] = (uint64_t)lo;
    out[1] = (uint64_t)(lo >> 64);
}

static int bit_of(const uint64_t *v, int i) { return (int)((v[i >> 6] >> (i & 63)) & 1); }
static void flip(uint64_t *v, int i) { v[i >> 6] ^= 1ull << (i & 63); }

void orc_b128_mul_bitwise(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]) {
    uint64_t p[4] = {0, 0, 0, 0};
    for (int i = 0; i < 128; i++)
        if (bit_of(a, i))
            for (int j = 0; j < 128; j++)
                if (bit_of(b, j)) flip(p, i + j);
    for (int j = 255; j >= 128; j--)
        if (bit_of(p, j)) {
            flip(p, j);
            flip(p, j - 128);
            flip(p, j - 121);
            flip(p, j - 126);
            flip(p, j - 127);
        }
    out[0] = p[0];
    out[1] = p[1];
}

/* F162 = GF(2)[x]/(x^162 + x^81 + 1), three u64 limbs, limb 2 below 2^34
 * (pcs/src/fields/scalar.rs:15, 97-121). */
void orc_f162_mul(const uint64_t a[3], const uint64_t b[3], uint64_t out[3]) {
    uint64_t p[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            u128 t = clmul64(a[i], b[j]);
            p[i + j] ^= (uint64_t)t;
            p[i + j + 1] ^= (uint64_t)(t >> 64);
        }
    /* x^(162+k) = x^(81+k) + x^k. The product has degree <= 322; with H = p >> 162 (161 bits) and
     * L = p mod x^162, p = L + H + H x^81, and H x^81 = (H mod x^81) x^81 + G x^162 with
     * G = H >> 81 (80 bits), which folds once more into G x^81 + G, both below x^162. */
    uint64_t H[3], G[2], t[3];
    for (int i = 0; i < 3; i++) H[i] = (p[i + 2] >> 34) | (i + 3 < 6 ? p[i + 3] << 30 : 0);
    H[2] &= (1ull << 33) - 1; /* bits 162..322 */
    G[0] = (H[1] >> 17) | (H[2] << 47);
    G[1] = H[2] >> 17;
    /* t = (H mod x^81) + G, placed at x^81 */
    uint64_t hm0 = H[0] ^ G[0], hm1 = (H[1] & ((1ull << 17) - 1)) ^ G[1];
    t[0] = 0; /* x^81 is bit 17 of limb 1 */
    t[1] = hm0 << 17;
    t[2] = (hm0 >> 47) | (hm1 << 17);
    out[0] = p[0] ^ H[0] ^ G[0] ^ t[0];
    out[1] = p[1] ^ H[1] ^ G[1] ^ t[1];
    out[2] = (p[2] & ((1ull << 34) - 1)) ^ H[2] ^ t[2];
}

void orc_f162_mul_bitwise(const uint64_t a[3], const uint64_t b[3], uint64_t out[3]) {
    uint64_t p[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 162; i++)
        if (bit_of(a, i))
            for (int j = 0; j < 162; j++)
                if (bit_of(b, j)) flip(p, i + j);
    for (int j = 323; j >= 162; j--)
        if (bit_of(p, j)) {
            flip(p, j);
            flip(p, j - 162);
            flip(p, j - 81);
        }
    out[0] = p[0];
    out[1] = p[1];
    out[2] = p[2];
}

static orc_b128 bmul(orc_b128 a, orc_b128 b) {
    orc_b128 r;
    orc_b128_mul(a.w, b.w, r.w);
    return r;
}
static orc_b128 badd(orc_b128 a, orc_b128 b) {
    orc_b128 r = {{a.w[0] ^ b.w[0], a.w[1] ^ b.w[1]}};
    return r;
}
static orc_f162 fmul(orc_f162 a, orc_f162 b) {
    orc_f162 r;
    orc_f162_mul(a.w, b.w, r.w);
    return r;
}
static orc_f162 fadd(orc_f162 a, orc_f162 b) {
    orc_f162 r = {{a.w[0] ^ b.w[0], a.w[1] ^ b.w[1], a.w[2] ^ b.w[2]}};
    return r;
}
static int feq(orc_f162 a, orc_f162 b) {
    return a.w[0] == b.w[0] && a.w[1] == b.w[1] && a.w[2] == b.w[2];
}
static const orc_f162 F_ZERO = {{0, 0, 0}};
static const orc_f162 F_ONE = {{1, 0, 0}};
/* F162::from_b128 (scalar.rs:77): zero extension. */
static orc_f162 lift(orc_b128 x) {
    orc_f162 r = {{x.w[0], x.w[1], 0}};
    return r;
}

/* =============================== blake3 test access ====================================== */

void orc_blake3(int mode, const uint8_t *key_or_ctx, size_t ctx_len, const uint8_t *msg,
                size_t msg_len, uint64_t offset, uint8_t *out, size_t out_len) {
    b3_hasher h;
    if (mode == 0)
        b3_init(&h);
    else if (mode == 1)
        b3_init_keyed(&h, key_or_ctx);
    else
        b3_init_derive_key(&h, (const char *)key_or_ctx, ctx_len);
    b3_update(&h, msg, msg_len);
    b3_finalize_seek(&h, offset, out, out_len);
}

/* =============================== table ================================================== */

static const char TABLE_CONTEXT[] = "labinius-fast/binary-table/v1";

/* verifier/proof_verify/src/main.rs:130-141: BLAKE3 keyed by the seed, updated with the
 * context, read as XOF. */
void orc_table_bytes(const uint8_t seed[32], uint64_t offset, uint8_t *out, size_t len) {
    b3_hasher h;
    b3_init_keyed(&h, seed);
    b3_update(&h, TABLE_CONTEXT, sizeof TABLE_CONTEXT - 1);
    const size_t step = (size_t)1 << 20;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)((len + step - 1) / step); i++) {
        size_t at = (size_t)i * step;
        size_t n = len - at < step ? len - at : step;
        b3_finalize_seek(&h, offset + at, out + at, n);
    }
}

static orc_b128 word_at(const uint8_t *table, size_t j) {
    orc_b128 w;
    memcpy(&w.w[0], table + 16 * j, 8);
    memcpy(&w.w[1], table + 16 * j + 8, 8);
    return w;
}

/* =============================== transcript ============================================= */
/* pcs/src/challenge.rs:226-326 and reader_at 724-749: one sequential BLAKE3 over a byte
 * stream; every derivation hashes stream || u64le(counter) || u64le(|label|) || label. */

typedef struct {
    b3_hasher h;
    uint64_t counter;
} tr_t;

static void put_u64(b3_hasher *h, uint64_t x) {
    uint8_t b[8];
    for (int i = 0; i < 8; i++) b[i] = (uint8_t)(x >> (8 * i));
    b3_update(h, b, 8);
}

static void tr_new(tr_t *t, const void *domain, size_t n) {
    b3_init(&t->h);
    put_u64(&t->h, n);
    b3_update(&t->h, domain, n);
    t->counter = 0;
}
static void tr_absorb_bytes(tr_t *t, const void *b, size_t n) {
    put_u64(&t->h, n);
    b3_update(&t->h, b, n);
}
static void tr_absorb_u64(tr_t *t, uint64_t x) { put_u64(&t->h, x); }
static void tr_absorb_b128(tr_t *t, const orc_b128 *xs, size_t n) {
    put_u64(&t->h, n);
    for (size_t i = 0; i < n; i++) {
        put_u64(&t->h, xs[i].w[0]);
        put_u64(&t->h, xs[i].w[1]);
    }
}
static void tr_absorb_f162(tr_t *t, const orc_f162 *xs, size_t n) {
    put_u64(&t->h, n);
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) put_u64(&t->h, xs[i].w[k]);
}
/* Transcript::reader: the derivation, then counter += 1. */
static void tr_reader(tr_t *t, const char *label, b3_hasher *out) {
    *out = t->h;
    size_t n = strlen(label);
    put_u64(out, t->counter);
    put_u64(out, n);
    b3_update(out, label, n);
    t->counter++;
}
static uint64_t le64(const uint8_t *b) {
    uint64_t x = 0;
    for (int i = 7; i >= 0; i--) x = (x << 8) | b[i];
    return x;
}
static void tr_sample_b128(tr_t *t, const char *label, int n, orc_b128 *out) {
    b3_hasher x;
    tr_reader(t, label, &x);
    uint8_t *buf = (uint8_t *)malloc((size_t)16 * n);
    b3_finalize_seek(&x, 0, buf, (size_t)16 * n);
    for (int i = 0; i < n; i++) {
        out[i].w[0] = le64(buf + 16 * i);
        out[i].w[1] = le64(buf + 16 * i + 8);
    }
    free(buf);
}
static void tr_sample_f162(tr_t *t, const char *label, int n, orc_f162 *out) {
    b3_hasher x;
    tr_reader(t, label, &x);
    uint8_t *buf = (uint8_t *)malloc((size_t)24 * n);
    b3_finalize_seek(&x, 0, buf, (size_t)24 * n);
    for (int i = 0; i < n; i++) {
        out[i].w[0] = le64(buf + 24 * i);
        out[i].w[1] = le64(buf + 24 * i + 8);
        out[i].w[2] = le64(buf + 24 * i + 16) & ((1ull << 34) - 1);
    }
    free(buf);
}

void orc_transcript_probe(const uint8_t *domain, size_t dlen, const uint8_t *msg, size_t mlen,
                          const uint8_t *label, size_t llen, int n, uint64_t *b128_out,
                          uint64_t *f162_out) {
    tr_t t;
    tr_new(&t, domain, dlen);
    tr_absorb_bytes(&t, msg, mlen);
    char lab[64];
    memcpy(lab, label, llen);
    lab[llen] = 0;
    orc_b128 *b = (orc_b128 *)malloc(sizeof(orc_b128) * n);
    tr_sample_b128(&t, lab, n, b);
    tr_sample_b128(&t, lab, n, b); /* the second sample: the counter advanced */
    memcpy(b128_out, b, sizeof(orc_b128) * n);
    orc_f162 *f = (orc_f162 *)malloc(sizeof(orc_f162) * n);
    tr_sample_f162(&t, lab, n, f);
    memcpy(f162_out, f, sizeof(orc_f162) * n);
    free(b);
    free(f);
}

/* =============================== short challenges ======================================= */
/* pcs/src/challenge.rs: Xof 329-372, attempt 662-688, signed 429-441, within/norm 547-623. */

typedef struct {
    b3_hasher h;
    uint64_t off;
    uint8_t buf[256];
    int pos;
    /* Test hook: when raw is set, bytes come from raw[0..raw_len) instead of the XOF (reads past
     * the end return 0 and set overrun); consumed counts every byte read either way. */
    const uint8_t *raw;
    size_t raw_len, consumed;
    int overrun;
} xof_t;

static void xof_init(xof_t *x, const b3_hasher *h) {
    x->h = *h;
    x->off = 0;
    x->pos = 256;
    x->raw = NULL;
    x->raw_len = x->consumed = 0;
    x->overrun = 0;
}
static uint8_t xof_byte(xof_t *x) {
    if (x->raw) {
        if (x->consumed >= x->raw_len) {
            x->overrun = 1;
            x->consumed++;
            return 0;
        }
        return x->raw[x->consumed++];
    }
    x->consumed++;
    if (x->pos == 256) {
        b3_finalize_seek(&x->h, x->off, x->buf, 256);
        x->off += 256;
        x->pos = 0;
    }
    return x->buf[x->pos++];
}
static uint32_t xof_u16(xof_t *x) {
    uint32_t lo = xof_byte(x);
    uint32_t hi = xof_byte(x);
    return lo | (hi << 8);
}
/* Xof::below: accept r < floor(65536 / n) * n, return r mod n. */
static uint32_t xof_below(xof_t *x, uint32_t n) {
    uint32_t limit = (65536u / n) * n;
    for (;;) {
        uint32_t r = xof_u16(x);
        if (r < limit) return r % n;
    }
}

static void attempt(xof_t *x, int weight, orc_challenge *out) {
    uint8_t perm[N162];
    uint64_t seen[3] = {0, 0, 0};
    for (int i = 0; i < N162; i++) perm[i] = (uint8_t)i;
    for (int i = 0; i < weight; i++) {
        int j = i + (int)xof_below(x, (uint32_t)(N162 - i));
        uint8_t p = perm[j];
        perm[j] = perm[i];
        seen[p >> 6] |= 1ull << (p & 63);
    }
    memset(out, 0, sizeof *out);
    out->weight = (uint32_t)weight;
    int n = 0;
    for (int p = 0; p < N162; p++)
        if ((seen[p >> 6] >> (p & 63)) & 1) out->positions[n++] = (uint8_t)p;
}

/* The sign word: BLAKE3 derive_key("labinius 2026 challenge signs v1") of
 * u64le(weight) || positions, first four bytes little-endian (challenge.rs:429-441, 875-879). */
uint32_t orc_sign_bits(const orc_challenge *c) {
    static const char CTX[] = "labinius 2026 challenge signs v1";
    b3_hasher h;
    b3_init_derive_key(&h, CTX, sizeof CTX - 1);
    uint8_t m[40];
    for (int i = 0; i < 8; i++) m[i] = (uint8_t)((uint64_t)c->weight >> (8 * i));
    memcpy(m + 8, c->positions, c->weight);
    b3_update(&h, m, 8 + c->weight);
    uint8_t o[4];
    b3_finalize_seek(&h, 0, o, 4);
    uint32_t w = (uint32_t)o[0] | ((uint32_t)o[1] << 8) | ((uint32_t)o[2] << 16) | ((uint32_t)o[3] << 24);
    uint32_t mask = (uint32_t)((1ull << c->weight) - 1);
    return w & mask;
}

static double COS[CONDUCTOR], SIN[CONDUCTOR];
static int UNITS[N162];
static int trig_ready = 0;
static void trig_init(void) {
    if (trig_ready) return;
    /* challenge.rs:504-533: angle = 2 pi m / 243 in f64, libm cos/sin. */
    for (int m = 0; m < CONDUCTOR; m++) {
        volatile double angle = 2.0 * M_PI * (double)m / (double)CONDUCTOR;
        COS[m] = cos(angle);
        SIN[m] = sin(angle);
    }
    int n = 0;
    for (int x = 1; x < CONDUCTOR; x++)
        if (x % 3 != 0) UNITS[n++] = x;
    trig_ready = 1;
}

/* max over the first 81 units u of |sum_t (+-) zeta^(pos_t u)|^2: per root, terms added in
 * position order starting from 0.0, then re*re + im*im with two roundings and one add
 * (challenge.rs:565-623; the accumulation order fixes the f64 result bit for bit). */
double orc_canonical_inf_norm_sq(const orc_challenge *c) {
    trig_init();
    double best = 0.0;
    for (int k = 0; k < N162 / 2; k++) {
        double re = 0.0, im = 0.0;
        for (uint32_t t = 0; t < c->weight; t++) {
            int m = (c->positions[t] * UNITS[k]) % CONDUCTOR;
            int neg = (c->signs >> t) & 1;
            re = re + (neg ? -COS[m] : COS[m]);
            im = im + (neg ? -SIN[m] : SIN[m]);
        }
        double a = re * re;
        double b = im * im;
        double x = a + b;
        if (x > best) best = x;
    }
    return best;
}

static int within(const orc_challenge *c, double bound_sq) {
    trig_init();
    for (int k = 0; k < N162 / 2; k++) {
        double re = 0.0, im = 0.0;
        for (uint32_t t = 0; t < c->weight; t++) {
            int m = (c->positions[t] * UNITS[k]) % CONDUCTOR;
            int neg = (c->signs >> t) & 1;
            re = re + (neg ? -COS[m] : COS[m]);
            im = im + (neg ? -SIN[m] : SIN[m]);
        }
        double a = re * re;
        double b = im * im;
        double x = a + b;
        if (x > bound_sq) return 0;
    }
    return 1;
}

/* Horner over all 162 roots (challenge.rs:627-645), the independent cross-check. */
double orc_canonical_inf_norm_sq_naive(const orc_challenge *c) {
    trig_init();
    int coeffs[N162];
    memset(coeffs, 0, sizeof coeffs);
    for (uint32_t i = 0; i < c->weight; i++)
        coeffs[c->positions[i]] = 1 - 2 * (int)((c->signs >> i) & 1);
    double best = 0.0;
    for (int k = 0; k < N162; k++) {
        double angle = 2.0 * M_PI * (double)UNITS[k] / (double)CONDUCTOR;
        double zr = cos(angle), zi = sin(angle);
        double ar = 0.0, ai = 0.0;
        for (int p = N162 - 1; p >= 0; p--) {
            double nr = ar * zr - ai * zi, ni = ar * zi + ai * zr;
            ar = nr + (double)coeffs[p];
            ai = ni;
        }
        double m = ar * ar + ai * ai;
        if (m > best) best = m;
    }
    return best;
}

/* sample_short_challenge (challenge.rs:662-688): one derivation "short-challenge", attempts
 * read one continuous XOF stream until the signed candidate is within the bound. */
static void sample_short_challenge(tr_t *t, orc_challenge *out) {
    const double bound = 12.0;
    const double bound_sq = bound * bound + 1e-12;
    b3_hasher h;
    tr_reader(t, "short-challenge", &h);
    xof_t x;
    xof_init(&x, &h);
    for (;;) {
        attempt(&x, WEIGHT, out);
        out->signs = orc_sign_bits(out);
        if (within(out, bound_sq)) return;
    }
}

/* One attempt (Fisher-Yates draw of `weight` positions, unsigned) read from a given byte stream in
 * place of the XOF, for known-answer and byte-accounting tests of the draw. Returns the bytes the
 * attempt consumed, or -1 when it needed more than len bytes. */
long orc_attempt_from_bytes(const uint8_t *stream, size_t len, int weight, orc_challenge *out) {
    xof_t x;
    memset(&x, 0, sizeof x);
    x.raw = stream;
    x.raw_len = len;
    attempt(&x, weight, out);
    return x.overrun ? -1 : (long)x.consumed;
}

/* Bytes of the "short-challenge" stream that sample_short_challenge consumes per attempt, for
 * `count` challenges drawn after absorbing msg into a transcript started with domain:
 * attempts[i] = attempts challenge i took, consumed[i] = its bytes. */
void orc_challenges_accounting(const uint8_t *domain, size_t dlen, const uint8_t *msg, size_t mlen, int count,
                               uint32_t *attempts, uint64_t *consumed) {
    tr_t t;
    tr_new(&t, domain, dlen);
    tr_absorb_bytes(&t, msg, mlen);
    const double bound_sq = 12.0 * 12.0 + 1e-12;
    for (int i = 0; i < count; i++) {
        b3_hasher h;
        tr_reader(&t, "short-challenge", &h);
        xof_t x;
        xof_init(&x, &h);
        orc_challenge c;
        uint32_t n = 0;
        for (;;) {
            n++;
            attempt(&x, WEIGHT, &c);
            c.signs = orc_sign_bits(&c);
            if (within(&c, bound_sq)) break;
        }
        attempts[i] = n;
        consumed[i] = x.consumed;
    }
}

void orc_challenges_probe(const uint8_t *domain, size_t dlen, const uint8_t *msg, size_t mlen,
                          int count, orc_challenge *out) {
    tr_t t;
    tr_new(&t, domain, dlen);
    tr_absorb_bytes(&t, msg, mlen);
    for (int i = 0; i < count; i++) sample_short_challenge(&t, &out[i]);
}

/* =============================== ring R_Q =============================================== */

int orc_modulus_bits(u128 q) {
    u128 x = q - 1;
    int b = 0;
    while (x) {
        b++;
        x >>= 1;
    }
    return b;
}
int orc_coefficient_bytes(u128 q) { return (orc_modulus_bits(q) + 7) / 8; }

/* rq::ajtai_matrix element `index` (rq.rs:53-144). */
void orc_ajtai_element(u128 q, size_t elements, size_t index, u128 *out) {
    static const uint8_t SEED[32] = {0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a,
                                     0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a,
                                     0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a,
                                     0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a, 0x5a};
    static const char DOMAIN[] = "labinius-fast/ajtai-matrix/v1";
    b3_hasher h;
    b3_init_keyed(&h, SEED);
    b3_update(&h, DOMAIN, sizeof DOMAIN - 1);
    uint8_t qb[16];
    for (int i = 0; i < 16; i++) qb[i] = (uint8_t)(q >> (8 * i));
    b3_update(&h, qb, 16);
    put_u64(&h, (uint64_t)elements);
    put_u64(&h, (uint64_t)index);
    xof_t x;
    xof_init(&x, &h);
    int width = orc_coefficient_bytes(q), bits = orc_modulus_bits(q);
    u128 mask = bits == 128 ? ~(u128)0 : (((u128)1 << bits) - 1);
    for (int t = 0; t < N; t++) {
        for (;;) {
            u128 v = 0;
            for (int b = 0; b < width; b++) v |= (u128)xof_byte(&x) << (8 * b);
            v &= mask;
            if (v < q) {
                out[t] = v;
                break;
            }
        }
    }
}

static u128 *ajtai_matrix(u128 q, size_t elements) {
    u128 *m = (u128 *)malloc(sizeof(u128) * N * elements);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)elements; i++) orc_ajtai_element(q, elements, (size_t)i, m + N * i);
    return m;
}

static i128 rem_euclid(i128 x, i128 m) {
    i128 r = x % m;
    return r < 0 ? r + m : r;
}

/* rq::fold_into_ring (rq.rs:152-165). */
static void fold_into_ring(i128 *raw, u128 q, u128 *out) {
    i128 m = (i128)q;
    for (int k = 2 * N - 2; k >= N; k--) {
        i128 v = rem_euclid(raw[k], m);
        raw[k] = 0;
        raw[k - SPLIT] += v;
        raw[k - N] -= v;
    }
    for (int t = 0; t < N; t++) out[t] = (u128)rem_euclid(raw[t], m);
}

/* rq::commit_binary_column (rq.rs:264-287) for column `column` of the table. */
/* Schoolbook commitment of one column of `elements` ring elements (column c at word
 * c * 4 * elements): bit p of word k of element m is coefficient 4p + k, and each set bit adds
 * X^(4p+k) A_m to the raw product (rq.rs:264-287), folded into R_Q at the end. */
static void commit_column_words(const uint8_t *table, size_t elements, u128 q, const u128 *matrix, int column,
                                u128 *out) {
    if (q * (u128)(elements * 512) < ((u128)1 << 63)) {
        /* The same sums in u64: every raw coefficient is a sum of at most elements * 512 matrix
         * values below q, so it stays below 2^63 and converts to i128 exactly. */
        uint64_t *raw64 = (uint64_t *)calloc(2 * N - 1, sizeof(uint64_t));
        uint64_t a64[N];
        for (size_t m = 0; m < elements; m++) {
            for (int t = 0; t < N; t++) a64[t] = (uint64_t)matrix[N * m + t];
            for (int k = 0; k < 4; k++) {
                orc_b128 w = word_at(table, (size_t)column * 4 * elements + 4 * m + k);
                for (int h = 0; h < 2; h++)
                    for (uint64_t b = w.w[h]; b; b &= b - 1) {
                        uint64_t *win = raw64 + 4 * (64 * h + __builtin_ctzll(b)) + k;
                        for (int t = 0; t < N; t++) win[t] += a64[t];
                    }
            }
        }
        i128 raw[2 * N - 1];
        for (int t = 0; t < 2 * N - 1; t++) raw[t] = (i128)raw64[t];
        free(raw64);
        fold_into_ring(raw, q, out);
        return;
    }
    i128 *raw = (i128 *)calloc(2 * N - 1, sizeof(i128));
    for (size_t m = 0; m < elements; m++) {
        const u128 *a = matrix + N * m;
        for (int k = 0; k < 4; k++) {
            orc_b128 w = word_at(table, (size_t)column * 4 * elements + 4 * m + k);
            for (int p = 0; p < 128; p++) {
                if (!bit_of(w.w, p)) continue;
                i128 *win = raw + 4 * p + k;
                for (int t = 0; t < N; t++) win[t] += (i128)a[t];
            }
        }
    }
    fold_into_ring(raw, q, out);
    free(raw);
}

void orc_commit_column(const uint8_t *table, int log_vars, u128 q, const u128 *matrix, int column,
                       u128 *out) {
    orc_shape s;
    orc_shape_of(log_vars, &s);
    commit_column_words(table, s.ring_elements, q, matrix, column, out);
}

static void pack_coeff(u128 x, int width, uint8_t *dst) {
    for (int b = 0; b < width; b++) dst[b] = (uint8_t)(x >> (8 * b));
}

void orc_commit_columns_packed(const uint8_t *table, int log_vars, uint64_t q_lo, uint64_t q_hi,
                               int first, int count, uint8_t *out) {
    orc_shape s;
    orc_shape_of(log_vars, &s);
    u128 q = ((u128)q_hi << 64) | q_lo;
    u128 *matrix = ajtai_matrix(q, s.ring_elements);
    int width = orc_coefficient_bytes(q);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < count; i++) {
        u128 col[N];
        orc_commit_column(table, log_vars, q, matrix, first + i, col);
        for (int t = 0; t < N; t++) pack_coeff(col[t], width, out + ((size_t)i * N + t) * width);
    }
    free(matrix);
}

/* rq::a_times_v (rq.rs:188-217). */
/* Each thread accumulates its own range of elements as rq.rs does (reducing every 64 elements);
 * the per-thread accumulators are then summed modulo q, which is the same residue. */
static void a_times_v_range(const u128 *a, const int16_t *v, size_t e0, size_t e1, u128 q, i128 *raw) {
    i128 m = (i128)q;
    for (size_t e = e0; e < e1; e++) {
        const u128 *el = a + N * e;
        const int16_t *fo = v + N * e;
        for (int j = 0; j < N; j++) {
            int c = fo[j];
            if (c == 0) continue;
            u128 w = (u128)(c < 0 ? -c : c);
            i128 *win = raw + j;
            if (c > 0)
                for (int t = 0; t < N; t++) win[t] += (i128)(el[t] * w);
            else
                for (int t = 0; t < N; t++) win[t] -= (i128)(el[t] * w);
        }
        if ((e + 1) % 64 == 0)
            for (int t = 0; t < 2 * N - 1; t++) raw[t] = rem_euclid(raw[t], m);
    }
    for (int t = 0; t < 2 * N - 1; t++) raw[t] = rem_euclid(raw[t], m);
}

static void a_times_v(const u128 *a, const int16_t *v, size_t elements, u128 q, u128 *out) {
    const int parts = 64;
    i128 *raws = (i128 *)calloc((size_t)parts * (2 * N - 1), sizeof(i128));
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < parts; p++) {
        size_t e0 = elements * p / parts, e1 = elements * (p + 1) / parts;
        a_times_v_range(a, v, e0, e1, q, raws + (size_t)p * (2 * N - 1));
    }
    i128 *raw = (i128 *)calloc(2 * N - 1, sizeof(i128));
    for (int p = 0; p < parts; p++)
        for (int t = 0; t < 2 * N - 1; t++) raw[t] = rem_euclid(raw[t] + raws[(size_t)p * (2 * N - 1) + t], (i128)q);
    fold_into_ring(raw, q, out);
    free(raw);
    free(raws);
}

/* rq::challenge_combination (rq.rs:226-256): c_j sits in R_648 as c(-X^4). */
static void challenge_combination(const u128 *columns, const orc_challenge *ch, int n, u128 q,
                                  u128 *out) {
    i128 *raw = (i128 *)calloc(2 * N - 1, sizeof(i128));
    i128 mq = (i128)q;
    for (int j = 0; j < n; j++) {
        const u128 *col = columns + (size_t)N * j;
        for (uint32_t t = 0; t < ch[j].weight; t++) {
            int m = ch[j].positions[t];
            int sign = (1 - 2 * (int)((ch[j].signs >> t) & 1)) * ((m % 2 == 0) ? 1 : -1);
            i128 *win = raw + 4 * m;
            if (sign > 0)
                for (int i = 0; i < N; i++) win[i] += (i128)col[i];
            else
                for (int i = 0; i < N; i++) win[i] -= (i128)col[i];
        }
        if ((j + 1) % 64 == 0)
            for (int t = 0; t < 2 * N - 1; t++) raw[t] = rem_euclid(raw[t], mq);
    }
    fold_into_ring(raw, q, out);
    free(raw);
}

/* =============================== switch ================================================= */

/* eq_expand_b128_into (crossfield.rs:87-97). */
void orc_eq_expand_b128(const orc_b128 *r, int n, orc_b128 *t) {
    t[0].w[0] = 1;
    t[0].w[1] = 0;
    orc_b128 *tmp = NULL;
    for (int round = 0; round < n; round++) {
        long long len = 1ll << round;
        if (len < 4096) {
            for (long long i = len - 1; i >= 0; i--) {
                orc_b128 hi = bmul(t[i], r[round]);
                t[2 * i + 1] = hi;
                t[2 * i] = badd(t[i], hi);
            }
            continue;
        }
        /* large rounds: the same doubling from a copy of the previous round, in parallel */
        if (!tmp) tmp = (orc_b128 *)malloc(sizeof(orc_b128) << (n - 1));
        memcpy(tmp, t, sizeof(orc_b128) * len);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < len; i++) {
            orc_b128 hi = bmul(tmp[i], r[round]);
            t[2 * i + 1] = hi;
            t[2 * i] = badd(tmp[i], hi);
        }
    }
    free(tmp);
}

/* eq_expand_f162 (crossfield.rs:66-78): the same doubling over F162. */
void orc_eq_expand_f162(const orc_f162 *r, int n, orc_f162 *t) {
    t[0] = F_ONE;
    for (int round = 0; round < n; round++)
        for (long long i = (1ll << round) - 1; i >= 0; i--) {
            orc_f162 hi = fmul(t[i], r[round]);
            t[2 * i + 1] = hi;
            t[2 * i] = fadd(t[i], hi);
        }
}

/* eq_table (eval.rs:51-65): variable k is bit k of the index. */
void orc_eq_table(const orc_f162 *p, int n, orc_f162 *t) {
    t[0] = F_ONE;
    size_t len = 1;
    for (int k = 0; k < n; k++) {
        for (size_t b = 0; b < len; b++) {
            orc_f162 y = fmul(t[b], p[k]);
            t[b + len] = y;
            t[b] = fadd(t[b], y);
        }
        len *= 2;
    }
}

static void partials_from_eq(const uint8_t *table, size_t n, const orc_b128 *eq, orc_b128 *v) {
    memset(v, 0, sizeof(orc_b128) * PACK);
#pragma omp parallel
    {
        orc_b128 loc[PACK];
        memset(loc, 0, sizeof loc);
#pragma omp for schedule(static)
        for (long long j = 0; j < (long long)n; j++) {
            orc_b128 w = word_at(table, (size_t)j);
            for (int h = 0; h < 2; h++)
                for (uint64_t b = w.w[h]; b; b &= b - 1) {
                    const int k = 64 * h + __builtin_ctzll(b);
                    loc[k] = badd(loc[k], eq[j]);
                }
        }
#pragma omp critical
        for (int k = 0; k < PACK; k++) v[k] = badd(v[k], loc[k]);
    }
}

void orc_partials(const uint8_t *table, int log_vars, const uint64_t *r_hi, uint64_t *out) {
    orc_shape s;
    orc_shape_of(log_vars, &s);
    orc_b128 *eq = (orc_b128 *)malloc(sizeof(orc_b128) * s.n_words);
    orc_eq_expand_b128((const orc_b128 *)r_hi, s.log_packed, eq);
    partials_from_eq(table, s.n_words, eq, (orc_b128 *)out);
    free(eq);
}


typedef struct {
    orc_f162 *a, *p;
    size_t half;
} switch_t;

/* SwitchProver::from_workspace (crossfield.rs:575-600). */
static void switch_init(switch_t *sw, const uint8_t *table, size_t n, const orc_b128 *eq,
                        const orc_f162 *batch) {
    sw->a = (orc_f162 *)malloc(sizeof(orc_f162) * n);
    sw->p = (orc_f162 *)malloc(sizeof(orc_f162) * n);
    /* a_j = psi(eq_j) = sum_k bit_k(eq_j) batch[k] (crossfield.rs:116-129 via poly_from_eq 403-445),
     * summed per byte of eq_j: byte_sums[i][v] = sum_{b: bit b of v} batch[8 i + b] */
    orc_f162 (*byte_sums)[256] = (orc_f162 (*)[256])malloc(sizeof(orc_f162) * 16 * 256);
    for (int i = 0; i < 16; i++) {
        byte_sums[i][0] = F_ZERO;
        for (int v = 1; v < 256; v++) {
            int b = __builtin_ctz((unsigned)v);
            byte_sums[i][v] = fadd(byte_sums[i][v & (v - 1)], batch[8 * i + b]);
        }
    }
#pragma omp parallel for schedule(static)
    for (long long j = 0; j < (long long)n; j++) {
        orc_f162 acc = F_ZERO;
        for (int i = 0; i < 16; i++) acc = fadd(acc, byte_sums[i][(eq[j].w[i >> 3] >> (8 * (i & 7))) & 255]);
        sw->a[j] = acc;
        sw->p[j] = lift(word_at(table, (size_t)j));
    }
    free(byte_sums);
    sw->half = n;
}
/* sumcheck::msg (sumcheck.rs:89-117). */
static void switch_msg(switch_t *sw, orc_f162 msg[2]) {
    sw->half /= 2;
    size_t h = sw->half;
    orc_f162 e0 = F_ZERO, einf = F_ZERO;
    /* field sums: each thread adds its range, the partial sums are added after (order-free) */
#pragma omp parallel if (h >= 4096)
    {
        orc_f162 l0 = F_ZERO, linf = F_ZERO;
#pragma omp for schedule(static)
        for (long long j = 0; j < (long long)h; j++) {
            l0 = fadd(l0, fmul(sw->a[j], sw->p[j]));
            linf = fadd(linf, fmul(fadd(sw->a[j], sw->a[j + h]), fadd(sw->p[j], sw->p[j + h])));
        }
#pragma omp critical
        {
            e0 = fadd(e0, l0);
            einf = fadd(einf, linf);
        }
    }
    msg[0] = e0;
    msg[1] = einf;
}
/* sumcheck::fold (sumcheck.rs:119-139). */
static void switch_fold(switch_t *sw, orc_f162 r) {
    size_t h = sw->half;
#pragma omp parallel for schedule(static)
    for (long long j = 0; j < (long long)h; j++) {
        sw->a[j] = fadd(sw->a[j], fmul(r, fadd(sw->a[j], sw->a[j + h])));
        sw->p[j] = fadd(sw->p[j], fmul(r, fadd(sw->p[j], sw->p[j + h])));
    }
}
static void switch_free(switch_t *sw) {
    free(sw->a);
    free(sw->p);
}

void orc_switch_run(const uint8_t *table, int log_vars, const uint64_t *r_hi, const uint64_t *batch,
                    const uint64_t *rhos, uint64_t *rounds, uint64_t *opened) {
    orc_shape s;
    orc_shape_of(log_vars, &s);
    orc_b128 *eq = (orc_b128 *)malloc(sizeof(orc_b128) * s.n_words);
    orc_eq_expand_b128((const orc_b128 *)r_hi, s.log_packed, eq);
    switch_t sw;
    switch_init(&sw, table, s.n_words, eq, (const orc_f162 *)batch);
    free(eq);
    for (int r = 0; r < s.log_packed; r++) {
        orc_f162 msg[2];
        switch_msg(&sw, msg);
        memcpy(rounds + 6 * r, msg, sizeof msg);
        orc_f162 rho;
        memcpy(&rho, rhos + 3 * r, sizeof rho);
        switch_fold(&sw, rho);
    }
    memcpy(opened, &sw.p[0], sizeof(orc_f162));
    switch_free(&sw);
}

/* slice_sum (crossfield.rs:625-637). */
static orc_f162 slice_sum(const orc_b128 *v, const orc_f162 *batch) {
    orc_f162 s = F_ZERO;
    for (int k = 0; k < PACK; k++) {
        orc_b128 u = {{0, 0}};
        for (int i = 0; i < PACK; i++)
            if (bit_of(v[i].w, k)) u.w[i >> 6] |= 1ull << (i & 63);
        s = fadd(s, fmul(lift(u), batch[k]));
    }
    return s;
}

/* transparent_coeff (crossfield.rs:99-108) with the Tensor of crossfield.rs:10-63. */
static orc_f162 transparent_coeff(const orc_b128 *r_hi, const orc_f162 *r_pp, int l,
                                  const orc_f162 *batch) {
    orc_f162 t[PACK], vs[PACK], hs[PACK], self_c[PACK];
    memset(t, 0, sizeof t);
    t[0] = F_ONE;
    for (int i = 0; i < l; i++) {
        for (int c = 0; c < PACK; c++) vs[c] = fmul(t[c], r_pp[i]);
        memcpy(self_c, t, sizeof t);
        memset(hs, 0, sizeof hs);
        for (int k = PACK - 1; k >= 0; k--) {
            /* shift_x */
            orc_f162 top = hs[PACK - 1];
            for (int c = PACK - 1; c >= 1; c--) hs[c] = hs[c - 1];
            hs[0] = F_ZERO;
            hs[0] = fadd(hs[0], top);
            hs[1] = fadd(hs[1], top);
            hs[2] = fadd(hs[2], top);
            hs[7] = fadd(hs[7], top);
            if (bit_of(r_hi[i].w, k))
                for (int c = 0; c < PACK; c++) hs[c] = fadd(hs[c], self_c[c]);
        }
        for (int c = 0; c < PACK; c++) t[c] = fadd(fadd(t[c], vs[c]), hs[c]);
    }
    orc_f162 acc = F_ZERO;
    for (int c = 0; c < PACK; c++) acc = fadd(acc, fmul(t[c], batch[c]));
    return acc;
}

/* =============================== opening ================================================ */

void orc_row_eval(const uint8_t *table, int log_vars, const uint64_t *p0, int first, int count,
                  uint64_t *out) {
    orc_shape s;
    orc_shape_of(log_vars, &s);
    orc_f162 *eq = (orc_f162 *)malloc(sizeof(orc_f162) * s.column_len);
    orc_eq_table((const orc_f162 *)p0, s.row_log, eq);
#pragma omp parallel for schedule(dynamic, 4)
    for (int c = 0; c < count; c++) {
        orc_f162 acc = F_ZERO;
        size_t base = (size_t)(first + c) * s.column_len;
        for (size_t i = 0; i < s.column_len; i++) acc = fadd(acc, fmul(eq[i], lift(word_at(table, base + i))));
        memcpy(out + 3 * c, &acc, sizeof acc);
    }
    free(eq);
}

/* fold_element (reference.rs:78-115). Returns nonzero when a coefficient leaves i16. */
static int fold_element(const uint8_t *table, const orc_shape *s, const orc_challenge *ch, size_t m,
                        int32_t *out) {
    int32_t raw[2 * N - 1];
    int32_t bits[N];
    memset(raw, 0, sizeof raw);
    for (int j = 0; j < s->columns; j++) {
        memset(bits, 0, sizeof bits);
        for (int k = 0; k < 4; k++) {
            orc_b128 w = word_at(table, (size_t)j * s->column_len + 4 * m + k);
            for (int p = 0; p < 128; p++) bits[4 * p + k] = bit_of(w.w, p);
        }
        for (uint32_t t = 0; t < ch[j].weight; t++) {
            int position = ch[j].positions[t];
            int negative = (int)((ch[j].signs >> t) & 1) ^ (position & 1);
            int32_t *win = raw + 4 * position;
            if (negative)
                for (int i = 0; i < N; i++) win[i] -= bits[i];
            else
                for (int i = 0; i < N; i++) win[i] += bits[i];
        }
    }
    for (int k = 2 * N - 2; k >= N; k--) {
        int32_t v = raw[k];
        raw[k] = 0;
        raw[k - SPLIT] += v;
        raw[k - N] -= v;
    }
    int bad = 0;
    for (int t = 0; t < N; t++) {
        if (raw[t] < -32768 || raw[t] > 32767) bad = 1;
        out[t] = raw[t];
    }
    return bad;
}

int orc_fold(const uint8_t *table, int log_vars, const orc_challenge *ch, int first, int count,
             int32_t *out) {
    orc_shape s;
    orc_shape_of(log_vars, &s);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int i = 0; i < count; i++) bad |= fold_element(table, &s, ch, (size_t)(first + i), out + (size_t)N * i);
    return bad;
}

/* verifier/proof_verify/src/main.rs:62-64: FOLD_CAP per coefficient over the fold's
 * 2^(k-7)/4 * 648 coefficients, as f64, ceiling. */
uint64_t orc_fold_cap(int log_vars) {
    double n = (double)(((uint64_t)1 << (log_vars - 7)) / 4 * N);
    return (uint64_t)ceil(FOLD_CAP * n);
}
static uint64_t isqrt64(uint64_t x) {
    uint64_t r = (uint64_t)sqrt((double)x);
    while (r * r > x) r--;
    while ((r + 1) * (r + 1) <= x) r++;
    return r;
}

/* =============================== proof codec ============================================ */
/* bincode fixint little-endian, field order of `Proof` (verifier/proof_verify/src/main.rs:98-124;
 * solution/proof_generate/src/main.rs:49-74). */

static size_t f162_bytes(size_t count) { return (count * 162 + 7) / 8; }

size_t orc_proof_len(int log_vars, uint64_t q_lo, uint64_t q_hi) {
    orc_shape s;
    if (orc_shape_of(log_vars, &s)) return 0;
    u128 q = ((u128)q_hi << 64) | q_lo;
    size_t w = (size_t)orc_coefficient_bytes(q);
    return 4 + 16 + 8 + (size_t)s.columns * N * w + 8 + 16 * PACK + 8 + 48 * (size_t)s.log_packed + 24 + 8 +
           f162_bytes((size_t)s.columns) + 8 + 4 * (size_t)N * s.ring_elements;
}

typedef struct {
    uint8_t *p;
    size_t at, cap;
    int bad;
} wr_t;
static void wr(wr_t *w, const void *b, size_t n) {
    if (w->at + n > w->cap) {
        w->bad = 1;
        return;
    }
    memcpy(w->p + w->at, b, n);
    w->at += n;
}
static void wr_u64(wr_t *w, uint64_t x) {
    uint8_t b[8];
    for (int i = 0; i < 8; i++) b[i] = (uint8_t)(x >> (8 * i));
    wr(w, b, 8);
}

/* wire::pack_f162 (wire/mod.rs:445-453, BitWriter 86-138): 64 + 64 + 34 bits per element,
 * least significant first, the last partial byte flushed. */
static void pack_f162(const orc_f162 *v, size_t n, uint8_t *out) {
    size_t nb = f162_bytes(n);
    memset(out, 0, nb);
    size_t bit = 0;
    for (size_t i = 0; i < n; i++)
        for (int b = 0; b < 162; b++, bit++)
            if ((v[i].w[b >> 6] >> (b & 63)) & 1) out[bit >> 3] |= (uint8_t)(1u << (bit & 7));
}
static void unpack_f162(const uint8_t *in, size_t n, orc_f162 *v) {
    size_t bit = 0;
    for (size_t i = 0; i < n; i++) {
        memset(&v[i], 0, sizeof v[i]);
        for (int b = 0; b < 162; b++, bit++)
            if ((in[bit >> 3] >> (bit & 7)) & 1) v[i].w[b >> 6] |= 1ull << (b & 63);
    }
}

/* =============================== the prover ============================================= */

static void instance_transcript(tr_t *t, const uint8_t seed[32], int log_vars, u128 q,
                                const uint8_t *commitment, size_t clen) {
    /* verifier/proof_verify/src/main.rs:166-174 */
    static const char D[] = "labinius-fast/v1";
    tr_new(t, D, sizeof D - 1);
    tr_absorb_bytes(t, seed, 32);
    tr_absorb_u64(t, (uint64_t)log_vars);
    tr_absorb_u64(t, COLUMN_LOG);
    uint8_t qb[16];
    for (int i = 0; i < 16; i++) qb[i] = (uint8_t)(q >> (8 * i));
    tr_absorb_bytes(t, qb, 16);
    tr_absorb_bytes(t, commitment, clen);
}

static orc_b128 claim_of(const orc_b128 *partials, const orc_b128 *r_lo) {
    orc_b128 eq_lo[PACK];
    orc_eq_expand_b128(r_lo, 7, eq_lo);
    orc_b128 acc = {{0, 0}};
    for (int b = 0; b < PACK; b++) acc = badd(acc, bmul(partials[b], eq_lo[b]));
    return acc;
}

/* split_point (verifier main.rs:257-264): p0 = reverse(coords[10..]), p1 = reverse(coords[..10]). */
static void split_point(const orc_f162 *coords, int lp, orc_f162 *p0, orc_f162 *p1) {
    int rl = lp - COLUMN_LOG;
    for (int i = 0; i < rl; i++) p0[i] = coords[lp - 1 - i];
    for (int i = 0; i < COLUMN_LOG; i++) p1[i] = coords[COLUMN_LOG - 1 - i];
}

static void absorb_row(tr_t *t, const orc_f162 *row, int n) {
    /* RowEvaluation::absorb (scheme/commitment.rs:320-331) */
    static const char L[] = "labinius/row-evaluation";
    tr_absorb_bytes(t, L, sizeof L - 1);
    uint8_t *b = (uint8_t *)malloc((size_t)24 * n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++)
            for (int y = 0; y < 8; y++) b[24 * i + 8 * k + y] = (uint8_t)(row[i].w[k] >> (8 * y));
    tr_absorb_bytes(t, b, (size_t)24 * n);
    free(b);
}

static int accept_modulus(u128 q, int log_vars) {
    if (q < MIN_MODULUS) return -1;
    if (q <= 2 * (u128)isqrt64(orc_fold_cap(log_vars))) return -2;
    if (q >= ((u128)1 << 96)) return -3;
    return 0;
}

/* reference::prove (verifier/proof_verify/src/reference.rs:137-177). */
int orc_prove(const uint8_t seed[32], int log_vars, uint64_t q_lo, uint64_t q_hi,
              const uint8_t *table_in, uint8_t *proof, size_t cap, size_t *len) {
    orc_shape s;
    if (orc_shape_of(log_vars, &s)) return -1;
    u128 q = ((u128)q_hi << 64) | q_lo;
    if (accept_modulus(q, log_vars)) return -2;
    size_t plen = orc_proof_len(log_vars, q_lo, q_hi);
    if (cap < plen) return -3;
    uint8_t *table = (uint8_t *)table_in;
    if (!table) {
        table = (uint8_t *)malloc(s.table_bytes);
        orc_table_bytes(seed, 0, table, s.table_bytes);
    }
    int width = orc_coefficient_bytes(q);
    size_t clen = (size_t)s.columns * N * width;
    uint8_t *commitment = (uint8_t *)malloc(clen);
    orc_commit_columns_packed(table, log_vars, q_lo, q_hi, 0, s.columns, commitment);

    tr_t t;
    instance_transcript(&t, seed, log_vars, q, commitment, clen);
    orc_b128 r_lo[7];
    orc_b128 *r_hi = (orc_b128 *)malloc(sizeof(orc_b128) * s.log_packed);
    tr_sample_b128(&t, "r-lo", 7, r_lo);
    tr_sample_b128(&t, "r-hi", s.log_packed, r_hi);
    orc_b128 *eq = (orc_b128 *)malloc(sizeof(orc_b128) * s.n_words);
    orc_eq_expand_b128(r_hi, s.log_packed, eq);
    orc_b128 partials[PACK];
    partials_from_eq(table, s.n_words, eq, partials);
    orc_b128 value = claim_of(partials, r_lo);
    tr_absorb_b128(&t, &value, 1);
    tr_absorb_b128(&t, partials, PACK);
    orc_f162 rb[7], batch[PACK];
    tr_sample_f162(&t, "switch-batch", 7, rb);
    orc_eq_expand_f162(rb, 7, batch);
    switch_t sw;
    switch_init(&sw, table, s.n_words, eq, batch);
    free(eq);
    orc_f162 *rounds = (orc_f162 *)malloc(sizeof(orc_f162) * 2 * s.log_packed);
    orc_f162 *coords = (orc_f162 *)malloc(sizeof(orc_f162) * s.log_packed);
    for (int r = 0; r < s.log_packed; r++) {
        switch_msg(&sw, rounds + 2 * r);
        tr_absorb_f162(&t, rounds + 2 * r, 2);
        tr_sample_f162(&t, "switch-round", 1, &coords[r]);
        switch_fold(&sw, coords[r]);
    }
    orc_f162 opened = sw.p[0];
    switch_free(&sw);
    tr_absorb_f162(&t, &opened, 1);
    orc_f162 *p0 = (orc_f162 *)malloc(sizeof(orc_f162) * s.row_log), p1[COLUMN_LOG];
    split_point(coords, s.log_packed, p0, p1);
    orc_f162 *row = (orc_f162 *)malloc(sizeof(orc_f162) * s.columns);
    orc_row_eval(table, log_vars, (const uint64_t *)p0, 0, s.columns, (uint64_t *)row);
    absorb_row(&t, row, s.columns);
    orc_challenge *ch = (orc_challenge *)malloc(sizeof(orc_challenge) * s.columns);
    for (int j = 0; j < s.columns; j++) sample_short_challenge(&t, &ch[j]);
    int32_t *folded = (int32_t *)malloc(sizeof(int32_t) * N * s.ring_elements);
    int bad = orc_fold(table, log_vars, ch, 0, (int)s.ring_elements, folded);

    wr_t w = {proof, 0, cap, 0};
    uint8_t b4[4] = {1, 0, 0, 0};
    wr(&w, b4, 4);
    wr_u64(&w, (uint64_t)q);
    wr_u64(&w, (uint64_t)(q >> 64));
    wr_u64(&w, clen);
    wr(&w, commitment, clen);
    wr_u64(&w, PACK);
    for (int b = 0; b < PACK; b++) {
        wr_u64(&w, partials[b].w[0]);
        wr_u64(&w, partials[b].w[1]);
    }
    wr_u64(&w, (uint64_t)s.log_packed);
    for (int r = 0; r < s.log_packed; r++)
        for (int m = 0; m < 2; m++)
            for (int k = 0; k < 3; k++) wr_u64(&w, rounds[2 * r + m].w[k]);
    for (int k = 0; k < 3; k++) wr_u64(&w, opened.w[k]);
    size_t rb_len = f162_bytes((size_t)s.columns);
    uint8_t *rowb = (uint8_t *)malloc(rb_len);
    pack_f162(row, (size_t)s.columns, rowb);
    wr_u64(&w, rb_len);
    wr(&w, rowb, rb_len);
    size_t fl = 4 * (size_t)N * s.ring_elements;
    wr_u64(&w, fl);
    for (size_t i = 0; i < (size_t)N * s.ring_elements; i++) {
        uint8_t b[4];
        uint32_t x = (uint32_t)folded[i];
        for (int y = 0; y < 4; y++) b[y] = (uint8_t)(x >> (8 * y));
        wr(&w, b, 4);
    }
    *len = w.at;
    free(rowb);
    free(folded);
    free(ch);
    free(row);
    free(p0);
    free(coords);
    free(rounds);
    free(r_hi);
    free(commitment);
    if (!table_in) free(table);
    if (w.bad || w.at != plen) return -4;
    return bad ? -5 : 0;
}

/* =============================== the checker ============================================ */

typedef struct {
    u128 q;
    const uint8_t *commitment;
    size_t clen;
    orc_b128 partials[PACK];
    orc_f162 *rounds; /* 2 per round */
    orc_f162 opened;
    const uint8_t *row;
    size_t row_len;
    const uint8_t *folded;
    size_t folded_len;
} proof_t;

typedef struct {
    const uint8_t *p;
    size_t at, len;
    int bad;
} rd_t;
static const uint8_t *rd(rd_t *r, size_t n) {
    if (r->at + n > r->len) {
        r->bad = 1;
        return NULL;
    }
    const uint8_t *x = r->p + r->at;
    r->at += n;
    return x;
}
static uint64_t rd_u64(rd_t *r) {
    const uint8_t *b = rd(r, 8);
    return b ? le64(b) : 0;
}

/* codec().deserialize with limit 64 MiB and reject_trailing_bytes (verifier main.rs:119-124). */
static int decode(const uint8_t *buf, size_t len, int lp, proof_t *P, char *err, size_t el) {
    rd_t r = {buf, 0, len, 0};
    if (len > 64u * 1024 * 1024) {
        snprintf(err, el, "proof too large");
        return -1;
    }
    const uint8_t *v = rd(&r, 4);
    if (!v || v[0] != 1 || v[1] || v[2] || v[3]) {
        snprintf(err, el, "bad version");
        return -1;
    }
    uint64_t lo = rd_u64(&r), hi = rd_u64(&r);
    P->q = ((u128)hi << 64) | lo;
    P->clen = rd_u64(&r);
    P->commitment = rd(&r, P->clen);
    uint64_t np = rd_u64(&r);
    if (r.bad || np != PACK) {
        snprintf(err, el, "128 partial evaluations");
        return -1;
    }
    for (int b = 0; b < PACK; b++) {
        P->partials[b].w[0] = rd_u64(&r);
        P->partials[b].w[1] = rd_u64(&r);
    }
    uint64_t nr = rd_u64(&r);
    if (r.bad || nr != (uint64_t)lp) {
        snprintf(err, el, "one switch round per packed coordinate");
        return -1;
    }
    P->rounds = (orc_f162 *)malloc(sizeof(orc_f162) * 2 * lp);
    for (int i = 0; i < 2 * lp; i++)
        for (int k = 0; k < 3; k++) P->rounds[i].w[k] = rd_u64(&r);
    for (int k = 0; k < 3; k++) P->opened.w[k] = rd_u64(&r);
    P->row_len = rd_u64(&r);
    P->row = rd(&r, P->row_len);
    P->folded_len = rd_u64(&r);
    P->folded = rd(&r, P->folded_len);
    if (r.bad || r.at != len) {
        snprintf(err, el, "malformed proof bytes");
        free(P->rounds);
        P->rounds = NULL;
        return -1;
    }
    return 0;
}

static int checked_f162(orc_f162 x) { return (x.w[2] >> 34) == 0; }

/* The checker's trace, transcript and both checks (verifier/proof_verify/src/main.rs:293-409). */
static int verify_impl(const uint8_t seed[32], int log_vars, const uint8_t *buf, size_t len,
                       char *err, size_t el, uint64_t *r_hi_out, uint64_t *batch_out,
                       uint64_t *coords_out, orc_challenge *ch_out, int full) {
    orc_shape s;
    if (orc_shape_of(log_vars, &s)) {
        snprintf(err, el, "unsupported shape");
        return 1;
    }
    proof_t P;
    memset(&P, 0, sizeof P);
    if (decode(buf, len, s.log_packed, &P, err, el)) return 2;
    int rc = 0;
    uint8_t *table = NULL;
    orc_b128 *r_hi = NULL, *eq = NULL;
    orc_f162 *coords = NULL, *row = NULL, *p0 = NULL, *comp = NULL, *eqt = NULL;
    orc_challenge *ch = NULL;
    u128 *columns = NULL, *matrix = NULL;
    int16_t *v = NULL;
#define FAIL(code, msg)                       \
    do {                                      \
        snprintf(err, el, "%s", msg);         \
        rc = code;                            \
        goto done;                            \
    } while (0)
    int ar = accept_modulus(P.q, log_vars);
    if (ar == -1) FAIL(3, "commitment modulus is below the security bound");
    if (ar == -2) FAIL(3, "commitment modulus does not clear the no-wrap bound");
    if (ar == -3) FAIL(3, "commitment modulus is above what the checker's arithmetic covers");
    table = (uint8_t *)malloc(s.table_bytes);
    orc_table_bytes(seed, 0, table, s.table_bytes);
    tr_t t;
    instance_transcript(&t, seed, log_vars, P.q, P.commitment, P.clen);
    /* check_field_switch */
    orc_b128 r_lo[7];
    r_hi = (orc_b128 *)malloc(sizeof(orc_b128) * s.log_packed);
    tr_sample_b128(&t, "r-lo", 7, r_lo);
    tr_sample_b128(&t, "r-hi", s.log_packed, r_hi);
    orc_b128 value;
    if (full) {
        eq = (orc_b128 *)malloc(sizeof(orc_b128) * s.n_words);
        orc_eq_expand_b128(r_hi, s.log_packed, eq);
        orc_b128 trusted[PACK];
        partials_from_eq(table, s.n_words, eq, trusted);
        value = claim_of(trusted, r_lo);
    } else {
        value = claim_of(P.partials, r_lo);
    }
    tr_absorb_b128(&t, &value, 1);
    tr_absorb_b128(&t, P.partials, PACK);
    orc_f162 rb[7], batch[PACK];
    tr_sample_f162(&t, "switch-batch", 7, rb);
    orc_eq_expand_f162(rb, 7, batch);
    /* SwitchVerifier::start (crossfield.rs:645-660) */
    orc_b128 re = claim_of(P.partials, r_lo);
    if (re.w[0] != value.w[0] || re.w[1] != value.w[1]) FAIL(4, "partial evaluation mismatch");
    orc_f162 sum = slice_sum(P.partials, batch);
    coords = (orc_f162 *)malloc(sizeof(orc_f162) * s.log_packed);
    for (int r = 0; r < s.log_packed; r++) {
        orc_f162 e0 = P.rounds[2 * r], einf = P.rounds[2 * r + 1];
        if (!checked_f162(e0) || !checked_f162(einf)) FAIL(5, "noncanonical F162 element");
        tr_absorb_f162(&t, P.rounds + 2 * r, 2);
        tr_sample_f162(&t, "switch-round", 1, &coords[r]);
        orc_f162 rr = coords[r];
        orc_f162 e1 = fadd(sum, e0);
        /* s = e0 + r (e0 + e1 + einf) + r^2 einf (crossfield.rs:666-671) */
        sum = fadd(fadd(e0, fmul(rr, fadd(fadd(e0, e1), einf))), fmul(fmul(rr, rr), einf));
    }
    if (!checked_f162(P.opened)) FAIL(5, "noncanonical F162 element");
    if (full) {
        orc_f162 d = transparent_coeff(r_hi, coords, s.log_packed, batch);
        if (feq(d, F_ZERO)) FAIL(6, "the switch left the opened value unbound");
        if (!feq(sum, fmul(d, P.opened))) FAIL(6, "sumcheck final check failed");
    }
    tr_absorb_f162(&t, &P.opened, 1);
    /* check_opening */
    p0 = (orc_f162 *)malloc(sizeof(orc_f162) * s.row_log);
    orc_f162 p1[COLUMN_LOG];
    split_point(coords, s.log_packed, p0, p1);
    if (P.row_len != f162_bytes((size_t)s.columns)) FAIL(7, "malformed row evaluation");
    row = (orc_f162 *)malloc(sizeof(orc_f162) * s.columns);
    unpack_f162(P.row, (size_t)s.columns, row);
    if (P.folded_len != 4 * (size_t)N * s.ring_elements) FAIL(7, "folded witness has the wrong length");
    size_t nf = (size_t)N * s.ring_elements;
    v = (int16_t *)malloc(sizeof(int16_t) * nf);
    for (size_t i = 0; i < nf; i++) {
        int32_t x = (int32_t)(uint32_t)(P.folded[4 * i] | (P.folded[4 * i + 1] << 8) |
                                        (P.folded[4 * i + 2] << 16) | ((uint32_t)P.folded[4 * i + 3] << 24));
        if (x < -32768 || x > 32767) FAIL(7, "coefficient out of range");
        v[i] = (int16_t)x;
    }
    absorb_row(&t, row, s.columns);
    ch = (orc_challenge *)malloc(sizeof(orc_challenge) * s.columns);
    for (int j = 0; j < s.columns; j++) sample_short_challenge(&t, &ch[j]);
    if (r_hi_out) memcpy(r_hi_out, r_hi, sizeof(orc_b128) * s.log_packed);
    if (batch_out) memcpy(batch_out, batch, sizeof batch);
    if (coords_out) memcpy(coords_out, coords, sizeof(orc_f162) * s.log_packed);
    if (ch_out) memcpy(ch_out, ch, sizeof(orc_challenge) * s.columns);
    if (!full) goto done;
    {
        orc_f162 eqp1[1 << COLUMN_LOG];
        orc_eq_table(p1, COLUMN_LOG, eqp1);
        orc_f162 acc = F_ZERO;
        for (int c = 0; c < s.columns; c++) acc = fadd(acc, fmul(eqp1[c], row[c]));
        if (!feq(acc, P.opened)) FAIL(8, "the row evaluation does not reduce to the opened value");
    }
    {
        uint64_t normsq = 0, worst = 0;
        for (size_t i = 0; i < nf; i++) {
            uint64_t mgn = (uint64_t)(v[i] < 0 ? -(int)v[i] : v[i]);
            normsq += mgn * mgn;
            if (mgn > worst) worst = mgn;
        }
        if (normsq > orc_fold_cap(log_vars)) FAIL(9, "the fold is longer than the cap");
        if ((u128)worst > (P.q - 1) / 2) FAIL(9, "the fold is not centered");
    }
    {
        int width = orc_coefficient_bytes(P.q);
        if (P.clen != (size_t)s.columns * N * width) FAIL(10, "the commitment has the wrong length");
        columns = (u128 *)malloc(sizeof(u128) * N * s.columns);
        for (size_t i = 0; i < (size_t)N * s.columns; i++) {
            u128 x = 0;
            for (int b = 0; b < width; b++) x |= (u128)P.commitment[i * width + b] << (8 * b);
            if (x >= P.q) FAIL(10, "a commitment coefficient is not reduced");
            columns[i] = x;
        }
        matrix = ajtai_matrix(P.q, s.ring_elements);
        u128 lhs[N], rhs[N];
        a_times_v(matrix, v, s.ring_elements, P.q, lhs);
        challenge_combination(columns, ch, s.columns, P.q, rhs);
        if (memcmp(lhs, rhs, sizeof lhs)) FAIL(11, "the fold does not open the folded commitment");
    }
    {
        /* components_mod_2 (verifier main.rs:268-280) against the folded row. */
        comp = (orc_f162 *)calloc(4 * s.ring_elements, sizeof(orc_f162));
        for (size_t m = 0; m < s.ring_elements; m++)
            for (int p = 0; p < N / 4; p++)
                for (int k = 0; k < 4; k++)
                    if (v[(size_t)N * m + 4 * p + k] & 1) comp[4 * m + k].w[p >> 6] |= 1ull << (p & 63);
        eqt = (orc_f162 *)malloc(sizeof(orc_f162) * s.column_len);
        orc_eq_table(p0, s.row_log, eqt);
        orc_f162 lhs = F_ZERO, rhs = F_ZERO;
        for (size_t i = 0; i < s.column_len; i++) lhs = fadd(lhs, fmul(eqt[i], comp[i]));
        for (int j = 0; j < s.columns; j++) {
            orc_f162 par = F_ZERO;
            for (uint32_t i = 0; i < ch[j].weight; i++) par.w[ch[j].positions[i] >> 6] |= 1ull << (ch[j].positions[i] & 63);
            rhs = fadd(rhs, fmul(par, row[j]));
        }
        if (!feq(lhs, rhs)) FAIL(12, "the fold does not match the folded row evaluation");
    }
done:
    free(P.rounds);
    free(table);
    free(r_hi);
    free(eq);
    free(coords);
    free(row);
    free(p0);
    free(comp);
    free(eqt);
    free(ch);
    free(columns);
    free(matrix);
    free(v);
    return rc;
#undef FAIL
}

int orc_verify(const uint8_t seed[32], int log_vars, const uint8_t *proof, size_t len, char *err,
               size_t errlen) {
    if (errlen) err[0] = 0;
    return verify_impl(seed, log_vars, proof, len, err, errlen, NULL, NULL, NULL, NULL, 1);
}

int orc_replay(const uint8_t seed[32], int log_vars, const uint8_t *proof, size_t len,
               uint64_t *r_hi, uint64_t *batch, uint64_t *coords, orc_challenge *challenges) {
    char err[128];
    return verify_impl(seed, log_vars, proof, len, err, sizeof err, r_hi, batch, coords, challenges, 0);
}

/* Commitment of columns first .. first + count - 1 of a witness of `elements` ring elements per
 * column under key row `row`: the row's element m is element row * elements + m of the matrix
 * expansion for (q, elements) (rq.rs:53-144 with the index continued past `elements`). Packed
 * coefficient bytes, [column][t]. */
void orc_commit_key_row(const uint8_t *words, size_t elements, int row, uint64_t q_lo, uint64_t q_hi, int first,
                        int count, uint8_t *out) {
    u128 q = ((u128)q_hi << 64) | q_lo;
    u128 *key = (u128 *)malloc(sizeof(u128) * N * elements);
#pragma omp parallel for schedule(static)
    for (long long m = 0; m < (long long)elements; m++)
        orc_ajtai_element(q, elements, (size_t)row * elements + (size_t)m, key + N * m);
    int width = orc_coefficient_bytes(q);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < count; i++) {
        u128 col[N];
        commit_column_words(words, elements, q, key, first + i, col);
        for (int t = 0; t < N; t++) pack_coeff(col[t], width, out + ((size_t)i * N + t) * width);
    }
    free(key);
}
