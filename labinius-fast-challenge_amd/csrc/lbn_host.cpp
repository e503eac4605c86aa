// Host half of liblabinius_gpu: binary-field products (PCLMULQDQ), the Fiat-Shamir transcript,
// the short-challenge sampler's constants, the NTT constant tables, the whole-proof orchestration
// lbn_prove (the order of solution/proof_generate/src/main.rs:1692-2009 and
// verifier/proof_verify/src/reference.rs:137-177) and the checker lbn_verify. The table-sized
// passes and the public matrix expansion run on the GPU.
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <functional>
#include <memory>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "lbn_blake3.h"
#include "lbn_internal.h"

namespace lbn {

// ============================================================================= fields

static inline __m128i clmul(uint64_t a, uint64_t b) {
    return _mm_clmulepi64_si128(_mm_set_epi64x(0, (long long)a), _mm_set_epi64x(0, (long long)b), 0x00);
}
static inline u128 to_u128(__m128i x) {
    return (u128)(uint64_t)_mm_cvtsi128_si64(x) | ((u128)(uint64_t)_mm_extract_epi64(x, 1) << 64);
}

// GHASH basis: x^128 = x^7 + x^2 + x + 1 (pcs/src/fields/scalar.rs:4, 46-68).
B128 b128_mul(B128 a, B128 b) {
    u128 lo = to_u128(clmul(a.w[0], b.w[0]));
    u128 hi = to_u128(clmul(a.w[1], b.w[1]));
    u128 mid = to_u128(clmul(a.w[0], b.w[1])) ^ to_u128(clmul(a.w[1], b.w[0]));
    lo ^= mid << 64;
    hi ^= mid >> 64;
    u128 t = hi ^ (hi << 1) ^ (hi << 2) ^ (hi << 7);
    u128 o = (hi >> 127) ^ (hi >> 126) ^ (hi >> 121);
    lo ^= t ^ o ^ (o << 1) ^ (o << 2) ^ (o << 7);
    return B128{{(uint64_t)lo, (uint64_t)(lo >> 64)}};
}

// x^162 = x^81 + 1 (pcs/src/fields/scalar.rs:97-121).
F162 f162_mul(F162 a, F162 b) {
    uint64_t v[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            u128 t = to_u128(clmul(a.w[i], b.w[j]));
            v[i + j] ^= (uint64_t)t;
            v[i + j + 1] ^= (uint64_t)(t >> 64);
        }
    for (int it = 0; it < 2; it++) {
        uint64_t h0 = (v[2] >> 34) | (v[3] << 30), h1 = (v[3] >> 34) | (v[4] << 30), h2 = (v[4] >> 34) | (v[5] << 30);
        v[2] &= (1ull << 34) - 1;
        v[3] = v[4] = v[5] = 0;
        v[0] ^= h0, v[1] ^= h1, v[2] ^= h2;
        v[1] ^= h0 << 17;
        v[2] ^= (h0 >> 47) | (h1 << 17);
        v[3] ^= (h1 >> 47) | (h2 << 17);
        v[4] ^= h2 >> 47;
    }
    return F162{{v[0], v[1], v[2]}};
}

void eq_expand_b128(const B128 *r, int n, B128 *t) {
    t[0] = B128{{1, 0}};
    for (int round = 0; round < n; round++)
        for (long long i = (1ll << round) - 1; i >= 0; i--) {
            B128 hi = b128_mul(t[i], r[round]);
            t[2 * i + 1] = hi;
            t[2 * i] = b128_add(t[i], hi);
        }
}

void eq_expand_f162(const F162 *r, int n, F162 *t) {
    t[0] = F162{{1, 0, 0}};
    for (int round = 0; round < n; round++)
        for (long long i = (1ll << round) - 1; i >= 0; i--) {
            F162 hi = f162_mul(t[i], r[round]);
            t[2 * i + 1] = hi;
            t[2 * i] = f162_add(t[i], hi);
        }
}

void eq_table_f162(const F162 *p, int n, F162 *t) {
    t[0] = F162{{1, 0, 0}};
    size_t len = 1;
    for (int k = 0; k < n; k++, len *= 2)
        for (size_t b = 0; b < len; b++) {
            F162 y = f162_mul(t[b], p[k]);
            t[b + len] = y;
            t[b] = f162_add(t[b], y);
        }
}

int host_threads() {
    unsigned n = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(n ? n : 1u, 16u));
}

// A persistent pool shared by every caller: workers sleep on a condition variable until a job is
// posted; each job hands out indices through an atomic counter. Jobs of concurrent callers (proofs
// on other threads) are served side by side: a worker takes the job with the fewest helpers, and
// every caller works on its own job too. A job lives on its caller's stack, so the caller returns
// only when all indices are done and no worker still holds the job.
namespace {
struct PoolJob {
    const std::function<void(size_t)> *f;
    size_t n;
    std::atomic<size_t> next{0};
    size_t done = 0;  // under the pool mutex
    int users = 0;    // workers inside the job (under the pool mutex)
};

class Pool {
   public:
    static Pool &get() {
        static Pool pool;
        return pool;
    }
    void run(size_t n, const std::function<void(size_t)> &f) {
        PoolJob job;
        job.f = &f;
        job.n = n;
        {
            std::lock_guard<std::mutex> lk(m_);
            jobs_.push_back(&job);
        }
        cv_.notify_all();
        size_t mine = work(job);
        std::unique_lock<std::mutex> lk(m_);
        job.done += mine;
        done_.wait(lk, [&] { return job.done == n && job.users == 0; });
        jobs_.erase(std::find(jobs_.begin(), jobs_.end(), &job));
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
    }

   private:
    Pool() {
        for (int i = 1; i < host_threads(); i++) workers_.emplace_back([this] { loop(); });
    }
    static size_t work(PoolJob &job) {
        size_t count = 0;
        for (size_t i; (i = job.next.fetch_add(1)) < job.n; count++) (*job.f)(i);
        return count;
    }
    PoolJob *pick() {  // under m_: an unfinished job with the fewest helpers
        PoolJob *best = nullptr;
        for (PoolJob *j : jobs_)
            if (j->next.load() < j->n && (!best || j->users < best->users)) best = j;
        return best;
    }
    void loop() {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            PoolJob *job = nullptr;
            cv_.wait(lk, [&] { return stop_ || (job = pick()) != nullptr; });
            if (stop_) return;
            job->users++;
            lk.unlock();
            const size_t count = work(*job);
            lk.lock();
            job->done += count;
            job->users--;
            if (job->done == job->n && job->users == 0) done_.notify_all();
        }
    }
    std::mutex m_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> workers_;
    std::vector<PoolJob *> jobs_;
    bool stop_ = false;
};
}  // namespace

void parallel_for(size_t n, const std::function<void(size_t)> &f) {
    if (n <= 1 || host_threads() <= 1) {
        for (size_t i = 0; i < n; i++) f(i);
        return;
    }
    Pool::get().run(n, f);
}

// ============================================================================= transcript
// pcs/src/challenge.rs:226-326 and 724-749: one BLAKE3 over the absorbed stream; a derivation
// hashes stream || u64le(counter) || u64le(|label|) || label and advances the counter.

class Transcript {
   public:
    explicit Transcript(const char *domain) {
        size_t n = strlen(domain);
        h_.update_u64(n);
        h_.update(domain, n);
    }
    void absorb_bytes(const void *b, size_t n) {
        h_.update_u64(n);
        h_.update(b, n);
    }
    // absorb_bytes with the whole chunks hashed by cvs_of (see Blake3::update_with).
    void absorb_bytes_with(const uint8_t *b, size_t n,
                           const std::function<void(size_t, uint64_t, size_t, const Blake3::Subtrees &, uint32_t *)> &cvs_of) {
        h_.update_u64(n);
        h_.update_with(b, n, cvs_of);
    }
    const Blake3 &hasher() const { return h_; }
    // What absorb_bytes_with(b, n, cvs_of) will pass to cvs_of, and the key words and flags it uses.
    Blake3::WithPlan plan_bytes_with(size_t n) const {
        Blake3 h = h_;
        h.update_u64(n);
        return h.plan_with(n);
    }
    void save(B3State &s) const {
        h_.save_state(s);
        s.counter = counter_;
    }
    void load(const B3State &s) {
        h_.load_state(s);
        counter_ = s.counter;
    }
    void absorb_u64(uint64_t x) { h_.update_u64(x); }
    void absorb_b128(const B128 *xs, size_t n) {
        h_.update_u64(n);
        for (size_t i = 0; i < n; i++) h_.update_u64(xs[i].w[0]), h_.update_u64(xs[i].w[1]);
    }
    void absorb_f162(const F162 *xs, size_t n) {
        h_.update_u64(n);
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) h_.update_u64(xs[i].w[k]);
    }
    Blake3 reader_at(const char *label, uint64_t offset) const {
        Blake3 r = h_;
        size_t n = strlen(label);
        r.update_u64(counter_ + offset);
        r.update_u64(n);
        r.update(label, n);
        return r;
    }
    Blake3 reader(const char *label) {
        Blake3 r = reader_at(label, 0);
        counter_++;
        return r;
    }
    void skip(uint64_t n) { counter_ += n; }
    std::vector<B128> sample_b128(const char *label, int n) {
        std::vector<uint8_t> b((size_t)16 * n);
        reader(label).xof(0, b.data(), b.size());
        std::vector<B128> out(n);
        memcpy(out.data(), b.data(), b.size());
        return out;
    }
    std::vector<F162> sample_f162(const char *label, int n) {
        std::vector<uint8_t> b((size_t)24 * n);
        reader(label).xof(0, b.data(), b.size());
        std::vector<F162> out(n);
        memcpy(out.data(), b.data(), b.size());
        for (auto &x : out) x.w[2] &= (1ull << 34) - 1;
        return out;
    }

   private:
    Blake3 h_;
    uint64_t counter_ = 0;
};

// The linear map L: GF(2^128) -> F162, y -> sum of L[k] over the set bits k of y, through 32 tables of
// the 16 sums over a 4-bit window (12 KB, built with one add per entry).
struct LmapTables {
    F162 t[32][16];
    explicit LmapTables(const std::vector<F162> &L) {
        for (int w = 0; w < 32; w++) {
            t[w][0] = F162{{0, 0, 0}};
            for (int b = 1; b < 16; b++) t[w][b] = f162_add(t[w][b & (b - 1)], L[4 * w + __builtin_ctz(b)]);
        }
    }
    F162 operator()(B128 y) const {
        F162 acc{{0, 0, 0}};
        for (int w = 0; w < 32; w++) acc = f162_add(acc, t[w][(y.w[w >> 4] >> (4 * (w & 15))) & 15]);
        return acc;
    }
};

// L_{r+1}[k] = L_r((1 + c) x^k) + rho L_r[k] for k < 128, from the tables of L_r.
static void lmap_next(std::vector<F162> &L, const LmapTables &tab, B128 c, const F162 &rho) {
    std::vector<F162> next(128);
    B128 z = b128_add(c, B128{{1, 0}});  // (1 + c) x^k
    for (int k = 0; k < 128; k++) {
        next[k] = f162_add(tab(z), f162_mul(rho, L[k]));
        const uint64_t carry = z.w[1] >> 63;  // z *= x
        z.w[1] = (z.w[1] << 1) | (z.w[0] >> 63);
        z.w[0] = (z.w[0] << 1) ^ (carry ? 0x87ull : 0ull);
    }
    L.swap(next);
}

void lmap_next(std::vector<F162> &L, B128 c, const F162 &rho) {
    const LmapTables tab(L);
    lmap_next(L, tab, c, rho);
}

// The last switch rounds on the host (sumcheck.rs:89-139 driven as main.rs:975-992), from where the
// device stopped: round r's fold of p by rho (p'_j = p_j + rho (p_j + p_{j+h}), j < h = |p| / 2), then
// control step r + 1: the message, absorb, sample rho, update L, and after the last one the opened
// value p0 + rho (p0 + p1), absorbed. The message of round r + 1 is e0 = sum_k L[k] T0_k with
// T0_k = sum of p'_j over the j < h2 whose eq value E_j has bit k, that is e0 = sum_j p'_j L(E_j),
// and einf = sum_j (p'_j + p'_{j+h2}) L(E_j + E_{j+h2}); the device sums per bit first.
static void switch_host_tail(lbn_ctx *ctx, Transcript &t, SwitchHandoff &ho, F162 *rounds, F162 *coords) {
    const int lp = ctx->shape.log_packed;
    std::vector<F162> p = std::move(ho.p), q;
    std::vector<B128> e = std::move(ho.e), e2;
    std::vector<F162> &L = ho.L;
    F162 rho = ho.rho;
    for (int r = ho.r0 - 1; r < lp - 1; r++) {
        const int s = r + 1;  // the control step after this fold
        const LmapTables tab(L);  // of L_s
        const size_t h2 = p.size() / 4, h = 2 * h2;
        q.assign(h, F162{{0, 0, 0}});
        e2.assign(h2, B128{{0, 0}});
        F162 msg[2] = {F162{{0, 0, 0}}, F162{{0, 0, 0}}};
        for (size_t j = 0; j < h2; j++) {
            const F162 q0 = f162_add(p[j], f162_mul(rho, f162_add(p[j], p[j + h])));
            const F162 q1 = f162_add(p[j + h2], f162_mul(rho, f162_add(p[j + h2], p[j + h + h2])));
            q[j] = q0, q[j + h2] = q1;
            e2[j] = b128_add(e[j], e[j + h2]);
            msg[0] = f162_add(msg[0], f162_mul(q0, tab(e[j])));
            msg[1] = f162_add(msg[1], f162_mul(f162_add(q0, q1), tab(e2[j])));
        }
        p.swap(q);
        e.swap(e2);
        rounds[2 * s] = msg[0], rounds[2 * s + 1] = msg[1];
        t.absorb_f162(msg, 2);
        rho = t.sample_f162("switch-round", 1)[0];
        coords[s] = rho;
        if (s == lp - 1) {
            const F162 opened = f162_add(p[0], f162_mul(rho, f162_add(p[0], p[1])));
            ctx->opened = opened;
            t.absorb_f162(&opened, 1);
        } else {
            lmap_next(L, tab, ctx->r_hi[s], rho);
        }
    }
}

// ============================================================================= short challenges
// pcs/src/challenge.rs: Fisher-Yates attempt 662-688 over one XOF stream, keyed sign word
// 429-441, rejection by max_u |c(zeta^u)|^2 <= 144 + 1e-12 in f64 with the reference's order of
// operations (565-623): per root, terms added in position order from 0.0, then re*re + im*im.

// The sampler's fixed data for the device (challenge.rs:396-715): the cos/sin tables of the 243rd
// roots of unity as the host computes them, the units of Z/243 in increasing order, the sign
// hasher (derive-key context "labinius 2026 challenge signs v1") and the norm bound 12^2 + 1e-12.
struct ChallengeSampler {
    ChalArg arg{};
    ChallengeSampler() {
        for (int m = 0; m < 243; m++) {
            volatile double angle = 2.0 * M_PI * (double)m / 243.0;
            arg.cosv[m] = std::cos(angle);
            arg.sinv[m] = std::sin(angle);
        }
        int n = 0;
        for (int x = 1; x < 243 && n < 81; x++)
            if (x % 3) arg.units[n++] = x;
        const Blake3 sign_key = Blake3::derive_key("labinius 2026 challenge signs v1");
        memcpy(arg.sign_key, sign_key.key_words(), 32);
        arg.sign_flags = sign_key.flags();
        const double bound = 12.0;
        arg.bound_sq = bound * bound + 1e-12;
    }
};

// ============================================================================= NTT tables

static uint32_t pw(uint64_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1;
    a %= p;
    while (e) {
        if (e & 1) r = r * a % p;
        a = a * a % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

bool prime_tables(const PrimeInfo &pi, std::vector<uint32_t> &t) {
    const uint32_t p = pi.p;
    auto mf = [&](uint32_t x) { return (uint32_t)(((uint64_t)x << 32) % p); };  // Montgomery form
    auto inv = [&](uint32_t x) { return pw(x, p - 2, p); };
    const uint32_t w = pi.omega;
    if (pw(w, 1944, p) != 1 || pw(w, 972, p) == 1 || pw(w, 648, p) == 1) return false;
    const uint32_t eta[2] = {pw(w, 324, p), pw(w, 1620, p)};
    const uint32_t beta[2] = {w, pw(w, 5, p)};
    const uint32_t rho = pw(w, 6, p), w18 = pw(rho, 18, p);
    using namespace tab;
    t.assign(SIZE, 0);
    const uint32_t i324 = inv(324);
    for (int h = 0; h < 2; h++)
        for (int i = 0; i < 324; i++) {
            const uint32_t b = pw(beta[h], i, p);
            t[TWB + (h * 324 + i) * 4 + 0] = 0;
            t[TWB + (h * 324 + i) * 4 + 1] = b;
            t[TWB + (h * 324 + i) * 4 + 2] = (uint32_t)((uint64_t)eta[h] * b % p);
            t[TWB + (h * 324 + i) * 4 + 3] = (uint32_t)((uint64_t)(1 + eta[h]) % p * b % p);
            t[TW + h * 324 + i] = mf(b);
            t[UNTW + h * 324 + i] = mf((uint32_t)((uint64_t)pw(inv(beta[h]), i, p) * i324 % p));
        }
    t[ETA] = mf(eta[0]);
    t[ETA + 1] = mf(eta[1]);
    t[ETA0] = mf(eta[0]);
    t[DINV] = mf(inv((eta[0] + p - eta[1]) % p));
    const uint32_t rhoi = inv(rho);
    for (int n2 = 0; n2 < 18; n2++)
        for (int k1 = 0; k1 < 18; k1++) {
            t[TWPF + n2 * 18 + k1] = mf(pw(rho, (uint64_t)n2 * k1, p));
            t[TWPI + n2 * 18 + k1] = mf(pw(rhoi, (uint64_t)n2 * k1, p));
        }
    const uint32_t w18i = inv(w18);
    for (int j = 0; j < 9; j++) {
        t[R18F + j] = mf(pw(w18, j, p));
        t[R18I + j] = mf(pw(w18i, j, p));
    }
    for (int h = 0; h < 2; h++) {
        const uint64_t c[4] = {0, 1, eta[h], (1 + (uint64_t)eta[h]) % p};
        for (int j = 0; j < 9; j++) {
            const uint64_t wj = pw(w18, j, p);
            const uint64_t b0 = pw(beta[h], 18 * j, p), b1 = pw(beta[h], 18 * (j + 9), p);
            for (int sel = 0; sel < 16; sel++) {
                const uint64_t x0 = c[sel & 3] * b0 % p, x1 = c[sel >> 2] * b1 % p;
                t[ST1 + ((j * 2 + 0) * 2 + h) * 16 + sel] = (uint32_t)((x0 + x1) % p);
                t[ST1 + ((j * 2 + 1) * 2 + h) * 16 + sel] = (uint32_t)((x0 + p - x1) % p * wj % p);
            }
        }
        for (int n2 = 0; n2 < 18; n2++)
            for (int k1 = 0; k1 < 18; k1++)
                t[TWPB + (h * 18 + n2) * 18 + k1] =
                    mf((uint32_t)((uint64_t)pw(beta[h], n2, p) * pw(rho, (uint64_t)n2 * k1, p) % p));
    }
    return true;
}

}  // namespace lbn

using namespace lbn;

// ============================================================================= C ABI (host)

// The shape of a 2^log_vars-evaluation instance (verifier/proof_verify/src/main.rs:30-48 at
// log_vars = 31): 2^(k-7) GHASH words, 1024 columns of 2^(k-17) words, 4 words per ring element.
bool lbn::Shape::of(int log_vars, Shape &s) {
    if (log_vars < 19 || log_vars > 31) return false;
    s.log_vars = log_vars;
    s.log_packed = log_vars - 7;
    s.row_log = s.log_packed - COLUMN_LOG;
    s.n_words = (size_t)1 << s.log_packed;
    s.column_len = (size_t)1 << s.row_log;
    s.ring_elements = s.column_len / 4;
    s.table_bytes = ((size_t)1 << log_vars) / 8;
    return true;
}

extern "C" int lbn_create(int device, int log_vars, lbn_ctx **out) {
    if (!out) return LBN_E_ARG;
    *out = nullptr;
    auto *ctx = new lbn_ctx();
    ctx->device = device;
    if (!Shape::of(log_vars, ctx->shape)) {
        delete ctx;
        return LBN_E_ARG;
    }
    int rc = dev_init(ctx);
    if (rc != LBN_OK) {
        dev_free(ctx);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return LBN_OK;
}

extern "C" void lbn_destroy(lbn_ctx *ctx) {
    if (!ctx) return;
    dev_free(ctx);
    delete ctx;
}

extern "C" const char *lbn_last_error(const lbn_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static size_t f162_bytes(size_t n) { return (n * 162 + 7) / 8; }

extern "C" size_t lbn_proof_len(int log_vars, uint64_t q_lo, uint64_t q_hi) {
    Shape s;
    if (!Shape::of(log_vars, s)) return 0;
    u128 q = ((u128)q_hi << 64) | q_lo;
    if (q < 2) return 0;
    int bits = 0;
    for (u128 x = q - 1; x; x >>= 1) bits++;
    size_t w = (size_t)(bits + 7) / 8;
    return 4 + 16 + 8 + (size_t)COLUMNS * N * w + 8 + 16 * 128 + 8 + 48 * (size_t)s.log_packed + 24 + 8 +
           f162_bytes(COLUMNS) + 8 + 4 * (size_t)N * s.ring_elements;
}

namespace {
struct Writer {
    uint8_t *p;
    size_t at = 0;
    void bytes(const void *b, size_t n) {
        memcpy(p + at, b, n);
        at += n;
    }
    void u64(uint64_t x) {
        for (int i = 0; i < 8; i++) p[at++] = (uint8_t)(x >> (8 * i));
    }
};
double now_us() {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
// wire::pack_f162 (pcs/src/wire/mod.rs:445-453): 162 bits per element, least significant first.
void pack_f162(const F162 *v, size_t n, uint8_t *out) {
    std::vector<uint64_t> w((n * 162 + 63) / 64 + 2, 0);
    auto put = [&](size_t bit, uint64_t x) {  // x < 2^64 at bit offset `bit`
        const size_t i = bit >> 6, b = bit & 63;
        w[i] |= x << b;
        if (b) w[i + 1] |= x >> (64 - b);
    };
    for (size_t i = 0; i < n; i++) {
        put(162 * i, v[i].w[0]);
        put(162 * i + 64, v[i].w[1]);
        put(162 * i + 128, v[i].w[2] & ((1ull << 34) - 1));
    }
    for (size_t k = 0; k < f162_bytes(n); k++) out[k] = (uint8_t)(w[k >> 3] >> (8 * (k & 7)));
}
}  // namespace

#define TRY(x)                        \
    do {                              \
        int rc_ = (x);                \
        if (rc_ != LBN_OK) return rc_; \
    } while (0)

namespace {
// One proof between its phases: the front (commitment, transcript, partials, the switch's start),
// the device switch rounds (alone or batched with other proofs), the back (the switch's host steps,
// row, folding challenges, fold, bincode). Field offsets of the bincode layout:
// verifier/proof_verify/src/main.rs:98-124.
struct ProveJob {
    lbn_ctx *ctx;
    const uint8_t *seed;
    uint8_t *proof;
    size_t plen = 0, clen = 0, row_bytes = 0, fold_bytes = 0;
    size_t off_commit = 0, off_partials = 0, off_fold = 0;
    Transcript t{"labinius-fast/v1"};
    std::vector<B128> partials;
    std::vector<F162> rounds, coords;
    B3State ts{};
    SwitchHandoff ho;
    lbn_profile pf{};
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0, t6 = 0;
    std::vector<B128> r_lo, r_hi;
    std::vector<F162> p0, row;
    FoldPlan fp;
    ChalArg ca{};
    std::vector<lbn_challenge> ch;
    bool launched = false;  // device work queued: an error return waits for the device first
    ProveJob(lbn_ctx *c, const uint8_t *sd, uint8_t *pr) : ctx(c), seed(sd), proof(pr) {}
    ~ProveJob() {
        if (launched) dev_quiesce(ctx);
    }
};

// The proof's layout and the commitment's launch.
int prove_front_launch(ProveJob &j, size_t cap) {
    lbn_ctx *ctx = j.ctx;
    if (!ctx->table_loaded || !ctx->modulus_set) return ctx->fail(LBN_E_STATE, "table and modulus must be loaded");
    const Shape &s = ctx->shape;
    const u128 q = ctx->q;
    j.plen = lbn_proof_len(s.log_vars, (uint64_t)q, (uint64_t)(q >> 64));
    if (cap < j.plen) return ctx->fail(LBN_E_ARG, "proof buffer too small");
    const int lp = s.log_packed;
    j.clen = (size_t)COLUMNS * N * ctx->width;
    j.row_bytes = f162_bytes(COLUMNS);
    j.fold_bytes = 4 * (size_t)N * s.ring_elements;
    j.off_commit = 4 + 16 + 8;
    j.off_partials = j.off_commit + j.clen;  // u64 len + 128 x u128
    const size_t off_rounds = j.off_partials + 8 + 16 * 128;
    const size_t off_opened = off_rounds + 8 + 48 * (size_t)lp;
    const size_t off_row = off_opened + 24;
    j.off_fold = off_row + 8 + j.row_bytes;
    j.t0 = now_us();
    // From the first launch on, an error return first waits for the device: the commitment's copy
    // and the fold kernel may write into `proof` (directly when it is pinned).
    j.launched = true;
    return dev_commit_launch(ctx);
}

// verifier/proof_verify/src/main.rs:166-182 (absorb k and 10, the shape of this instance)
void absorb_instance(ProveJob &j) {
    Transcript &t = j.t;
    const u128 q = j.ctx->q;
    t.absorb_bytes(j.seed, 32);
    t.absorb_u64((uint64_t)j.ctx->shape.log_vars);
    t.absorb_u64(COLUMN_LOG);
    uint8_t qb[16];
    for (int i = 0; i < 16; i++) qb[i] = (uint8_t)(q >> (8 * i));
    t.absorb_bytes(qb, 16);
}

// After the commitment's absorb: the point's r_lo and r_hi.
void sample_point(ProveJob &j) {
    j.t1 = now_us();
    j.r_lo = j.t.sample_b128("r-lo", 7);
    j.r_hi = j.t.sample_b128("r-hi", j.ctx->shape.log_packed);
    j.t2 = now_us();
}

// After the partials (in j.partials): the claim and the partials into the transcript, the switch's
// batch and its start.
int absorb_claim(ProveJob &j) {
    lbn_ctx *ctx = j.ctx;
    const int lp = ctx->shape.log_packed;
    Transcript &t = j.t;
    {
        B128 eq_lo[128];
        eq_expand_b128(j.r_lo.data(), 7, eq_lo);
        B128 claim{{0, 0}};
        for (int b = 0; b < 128; b++) claim = b128_add(claim, b128_mul(j.partials[b], eq_lo[b]));
        t.absorb_b128(&claim, 1);
        t.absorb_b128(j.partials.data(), 128);
    }
    std::vector<F162> rb = t.sample_f162("switch-batch", 7);
    std::vector<F162> batch(128);
    eq_expand_f162(rb.data(), 7, batch.data());
    j.t3 = now_us();
    j.pf.commit_us = j.t1 - j.t0;
    j.pf.transcript_us = j.t2 - j.t1;
    j.pf.partials_us = j.t3 - j.t2;

    // The rounds run on the device with the transcript (main.rs:975-992: per round absorb [e0, einf],
    // sample "switch-round", fold; then absorb [opened]); the host continues from its final state.
    j.rounds.resize(2 * (size_t)lp);
    j.coords.resize(lp);
    TRY(dev_switch_prepare(ctx, (const uint64_t *)batch.data()));
    t.save(j.ts);
    return LBN_OK;
}

// After the commitment's launch: the transcript, the partials and the switch's start.
int prove_front_rest(ProveJob &j) {
    lbn_ctx *ctx = j.ctx;
    Transcript &t = j.t;
    absorb_instance(j);
    // The commitment's whole BLAKE3 chunks are hashed where the bytes are (device); the host reads
    // back their chaining values and the bytes around them. All bytes reach the proof at the end.
    int cvs_rc = LBN_OK;
    uint8_t *cdst = j.proof + j.off_commit;
    t.absorb_bytes_with(cdst, j.clen,
                        [&](size_t off, uint64_t first, size_t count, const Blake3::Subtrees &trees, uint32_t *cvs) {
                            cvs_rc = dev_commit_cvs(ctx, off, first, count, trees, t.hasher().key_words(),
                                                    t.hasher().flags(), cvs, cdst, off, off + count * Blake3::CHUNK);
                        });
    TRY(cvs_rc);
    sample_point(j);
    j.partials.resize(128);
    TRY(lbn_partials(ctx, (const uint64_t *)j.r_hi.data(), (uint64_t *)j.partials.data()));
    return absorb_claim(j);
}

// The same in the batch protocol: the instance absorbed and the commitment's hashing staged ...
int front_stage_hash(ProveJob &j) {
    absorb_instance(j);
    const Blake3::WithPlan w = j.t.plan_bytes_with(j.clen);
    return dev_cvs_stage(j.ctx, w.trees, w.take, w.first, w.whole, j.t.hasher().key_words(), j.t.hasher().flags(),
                         j.proof + j.off_commit, w.take, w.take + w.whole * Blake3::CHUNK);
}
// ... after dev_cvs_batch: the commitment absorbed with the staged hashing's results, the point
// sampled and the partials prepared ...
int front_absorb_commitment(ProveJob &j) {
    uint8_t *cdst = j.proof + j.off_commit;
    j.t.absorb_bytes_with(cdst, j.clen, [&](size_t, uint64_t, size_t, const Blake3::Subtrees &, uint32_t *cvs) {
        dev_cvs_collect(j.ctx, cvs, cdst);  // the plan staged from this same state
    });
    sample_point(j);
    return dev_partials_prepare(j.ctx, (const uint64_t *)j.r_hi.data(), false);
}
// ... after dev_partials_batch: the claim and the switch's start.
int front_claim(ProveJob &j) {
    j.partials.resize(128);
    dev_partials_collect(j.ctx, (uint64_t *)j.partials.data());
    return absorb_claim(j);
}

int prove_front(ProveJob &j, size_t cap) {
    TRY(prove_front_launch(j, cap));
    return prove_front_rest(j);
}

// The switch's host steps, the row and its absorb; leaves the folding challenges' input in j.ca.
// After the batched switch rounds: the switch's host steps and the row's point p0.
int prove_back_point(ProveJob &j) {
    lbn_ctx *ctx = j.ctx;
    const Shape &s = ctx->shape;
    const int lp = s.log_packed;
    j.t.load(j.ts);
    switch_host_tail(ctx, j.t, j.ho, j.rounds.data(), j.coords.data());
    j.t4 = now_us();
    // EvaluationPoint::msb_first (scheme/commitment.rs:163-174): p0 = reverse(coords[10..]).
    j.p0.resize(s.row_log);
    j.row.resize(COLUMNS);
    for (int i = 0; i < s.row_log; i++) j.p0[i] = j.coords[lp - 1 - i];
    return LBN_OK;
}

// After the row evaluation (in j.row): its absorb and the challenges' transcript state.
int prove_back_row(ProveJob &j) {
    Transcript &t = j.t;
    j.t5 = now_us();
    // RowEvaluation::absorb (scheme/commitment.rs:320-331), then 1024 challenges in parallel,
    // each one derivation at its own counter (challenge.rs:778-802).
    {
        static const char L[] = "labinius/row-evaluation";
        t.absorb_bytes(L, sizeof L - 1);
        std::vector<uint8_t> rb2((size_t)24 * COLUMNS);
        memcpy(rb2.data(), j.row.data(), rb2.size());
        t.absorb_bytes(rb2.data(), rb2.size());
    }
    static const ChallengeSampler sampler;
    j.ca = sampler.arg;
    t.save(j.ca.ts);  // challenge i reads the transcript at its counter + i
    j.ch.resize(COLUMNS);
    return LBN_OK;
}

int prove_back(ProveJob &j) {
    TRY(prove_back_point(j));
    TRY(lbn_row_eval(j.ctx, (const uint64_t *)j.p0.data(), j.ctx->shape.row_log, (uint64_t *)j.row.data()));
    return prove_back_row(j);
}

// Before the fold: the challenges leave the transcript.
void prove_tail_start(ProveJob &j) {
    j.t.skip(COLUMNS);
    j.t6 = now_us();
}
int32_t *fold_out(ProveJob &j) { return (int32_t *)(j.proof + j.off_fold + 8); }

// After the fold (in the proof): the commitment bytes and the bincode.
int prove_tail_encode(ProveJob &j, size_t *len) {
    lbn_ctx *ctx = j.ctx;
    const Shape &s = ctx->shape;
    const u128 q = ctx->q;
    const int lp = s.log_packed;
    const F162 opened = ctx->opened;
    auto &row = j.row;
    const double t4 = j.t4, t5 = j.t5, t6 = j.t6;
    TRY(dev_commit_bytes(ctx, j.proof + j.off_commit));
    const double t7 = now_us();

    Writer w{j.proof};
    const uint32_t version = 1;
    w.bytes(&version, 4);
    w.u64((uint64_t)q);
    w.u64((uint64_t)(q >> 64));
    w.u64(j.clen);
    w.at = j.off_partials;
    w.u64(128);
    for (auto &x : j.partials) w.u64(x.w[0]), w.u64(x.w[1]);
    w.u64((uint64_t)lp);
    for (auto &x : j.rounds)
        for (int k = 0; k < 3; k++) w.u64(x.w[k]);
    for (int k = 0; k < 3; k++) w.u64(opened.w[k]);
    w.u64(j.row_bytes);
    pack_f162(row.data(), COLUMNS, j.proof + w.at);
    w.at += j.row_bytes;
    w.u64(j.fold_bytes);
    *len = j.plen;
    j.launched = false;
    const double t8 = now_us();
    j.pf.switch_us = t4 - j.t3;
    j.pf.row_us = t5 - t4;
    j.pf.challenges_us = t6 - t5;
    j.pf.fold_us = t7 - t6;
    j.pf.encode_us = t8 - t7;
    j.pf.total_us = t8 - j.t0;
    return LBN_OK;
}

// After the folding challenges (in j.ch): the fold, the commitment bytes and the bincode.
int prove_tail(ProveJob &j, size_t *len) {
    prove_tail_start(j);
    TRY(lbn_fold(j.ctx, j.ch.data(), COLUMNS, fold_out(j), (size_t)N * j.ctx->shape.ring_elements));
    return prove_tail_encode(j, len);
}
}  // namespace

extern "C" int lbn_prove(lbn_ctx *ctx, const uint8_t seed[32], uint8_t *proof, size_t cap, size_t *len,
                         lbn_profile *prof) {
    if (!ctx || !seed || !proof || !len) return LBN_E_ARG;
    ProveJob j(ctx, seed, proof);
    TRY(prove_front(j, cap));
    TRY(dev_switch_rounds(ctx, j.ts, (uint64_t *)j.rounds.data(), (uint64_t *)j.coords.data(), j.ho));
    TRY(prove_back(j));
    TRY(dev_challenges(ctx, j.ca, j.ch.data()));
    TRY(prove_tail(j, len));
    if (prof) *prof = j.pf;
    return LBN_OK;
}

namespace {
constexpr int BATCH_GROUPS = 3;     // a batch runs as up to this many staggered groups
constexpr int BATCH_GROUP_MIN = 8;  // of at least this many proofs each
// The batch's host threads, started once per batch: run(f) calls f(order[k]) for every k, on the
// threads and the caller, and returns when every call has returned.
class PhasePool {
  public:
    PhasePool(int helpers, const std::vector<int> &order) : order_(order) {
        for (int k = 0; k < helpers; k++) th_.emplace_back([this] { loop(); });
    }
    ~PhasePool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        go_.notify_all();
        for (auto &t : th_) t.join();
    }
    void run(const std::function<void(int)> &f) {
        {
            std::lock_guard<std::mutex> lk(m_);
            f_ = &f;
            next_ = 0;
            busy_ = (int)th_.size();
            gen_++;
        }
        go_.notify_all();
        work(f);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return busy_ == 0; });
        f_ = nullptr;
    }

  private:
    void work(const std::function<void(int)> &f) {
        const int n = (int)order_.size();
        for (int k; (k = next_.fetch_add(1)) < n;) f(order_[k]);
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)> *f;
            {
                std::unique_lock<std::mutex> lk(m_);
                go_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                f = f_;
            }
            work(*f);
            std::lock_guard<std::mutex> lk(m_);
            if (--busy_ == 0) done_.notify_all();
        }
    }
    const std::vector<int> &order_;
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable go_, done_;
    const std::function<void(int)> *f_ = nullptr;
    std::atomic<int> next_{0};
    uint64_t gen_ = 0;
    int busy_ = 0;
    bool stop_ = false;
};
}  // namespace

// One group of a batch, phase by phase: the commitments on `threads` host threads (each proof on
// its own context and streams), then the transcripts and partials, then one device sequence for
// the switch rounds of all of them (one launch per device step), one launch pair for all rows, one
// launch for all folding challenges and one for all folds, with each proof's host steps between
// them on the threads; `committed` is called once the commitments are done. An item that fails
// keeps its error and leaves the others going.
static void batch_group(lbn_batch_item *items, std::vector<std::unique_ptr<ProveJob>> &jobs,
                        const std::vector<int> &order, int threads, const std::function<void()> &committed) {
    PhasePool pool(std::min(threads, (int)order.size()) - 1, order);
    auto run = [&](const std::function<void(int)> &f) { pool.run(f); };
    std::vector<int> li;
    std::vector<lbn_ctx *> lc;
    // the proofs still going, and a batched call over them (a failure fails each of them)
    auto batch = [&](const std::function<int(lbn_ctx *const *, int)> &call) {
        li.clear(), lc.clear();
        for (int i : order)
            if (items[i].status == 1) li.push_back(i), lc.push_back(items[i].ctx);
        if (li.empty()) return;
        const int rc = call(lc.data(), (int)lc.size());
        if (rc != LBN_OK) {
            dev_batch_quiesce(lc[0]->device);  // nothing it launched still writes the contexts' buffers
            for (int i : li) items[i].status = rc;
        }
    };
    // each phase on the threads: proof i's step, then its stream idle
    auto phase = [&](const std::function<int(int)> &step) {
        run([&](int i) {
            lbn_batch_item &it = items[i];
            if (it.status != 1) return;
            int rc = step(i);
            if (rc == LBN_OK) rc = dev_sync(it.ctx);
            if (rc != LBN_OK) it.status = rc;
        });
    };
    // The commitments first, on their own: their matrix-core blocks take whole CUs, which the
    // short kernels of the other proofs' later steps would hold up.
    phase([&](int i) {  // table, commitment
        lbn_batch_item &it = items[i];
        int rc = it.gen_table ? lbn_generate_table(it.ctx, it.seed) : LBN_OK;
        if (rc == LBN_OK) rc = prove_front_launch(*jobs[i], it.cap);
        return rc;
    });
    committed();
    phase([&](int i) { return front_stage_hash(*jobs[i]); });  // the instance, the commitment's hashing
    batch(dev_cvs_batch);
    phase([&](int i) { return front_absorb_commitment(*jobs[i]); });  // the commitment, the point
    batch(dev_partials_batch);
    phase([&](int i) {  // the claim, the switch's start state
        const int rc = front_claim(*jobs[i]);
        if (rc == LBN_OK) dev_switch_stage(jobs[i]->ctx, jobs[i]->ts);
        return rc;
    });
    batch(dev_switch_batch);
    phase([&](int i) {  // the switch's host steps, the row's eq tables
        ProveJob &j = *jobs[i];
        dev_switch_collect(j.ctx, j.ts, (uint64_t *)j.rounds.data(), (uint64_t *)j.coords.data(), j.ho);
        int rc = prove_back_point(j);
        if (rc == LBN_OK) rc = dev_row_prepare(j.ctx, (const uint64_t *)j.p0.data(), false);
        return rc;
    });
    batch(dev_row_batch);
    phase([&](int i) {  // the row into the transcript, the challenges' transcript state
        ProveJob &j = *jobs[i];
        dev_row_collect(j.ctx, (uint64_t *)j.row.data());
        const int rc = prove_back_row(j);
        if (rc == LBN_OK) dev_challenges_stage(j.ctx, j.ca);
        return rc;
    });
    batch(dev_challenges_batch);
    phase([&](int i) {  // the challenges, the fold's column lists
        ProveJob &j = *jobs[i];
        dev_challenges_collect(j.ctx, j.ch.data());
        prove_tail_start(j);
        return dev_fold_prepare(j.ctx, j.ch.data(), COLUMNS, fold_out(j), (size_t)N * j.ctx->shape.ring_elements,
                                j.fp, false);
    });
    {
        std::vector<FoldPlan> fps;
        batch([&](lbn_ctx *const *cx, int m) {
            for (int i : li) fps.push_back(jobs[i]->fp);
            return dev_fold_batch(cx, m, fps.data());
        });
    }
    run([&](int i) {  // the fold's range check, the commitment bytes, the bincode
        lbn_batch_item &it = items[i];
        if (it.status != 1) return;
        ProveJob &j = *jobs[i];
        int rc = dev_fold_collect(it.ctx, j.fp);
        if (rc == LBN_OK) rc = prove_tail_encode(j, &it.len);
        if (rc == LBN_OK) it.profile = j.pf;
        it.status = rc;
    });
}

// Many proofs together: up to BATCH_GROUPS staggered groups of at least BATCH_GROUP_MIN proofs.
extern "C" int lbn_prove_batch(lbn_batch_item *items, int n, int threads) {
    if (!items || n < 1 || n > SWB_MAX || threads < 1) return LBN_E_ARG;
    for (int i = 0; i < n; i++) {
        if (!items[i].ctx || !items[i].proof) return LBN_E_ARG;
        for (int k = 0; k < i; k++)
            if (items[k].ctx == items[i].ctx) return LBN_E_ARG;
        if (items[i].ctx->device != items[0].ctx->device) return LBN_E_ARG;
        items[i].status = 1;
        items[i].len = 0;
    }
    std::vector<std::unique_ptr<ProveJob>> jobs(n);
    for (int i = 0; i < n; i++) jobs[i].reset(new ProveJob(items[i].ctx, items[i].seed, items[i].proof));
    // the host threads take the items largest table first, so that no large proof starts last
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return items[a].ctx->shape.log_vars > items[b].ctx->shape.log_vars;
    });
    // Groups of alternate items, each starting once the one before it has its commitments: a
    // group's commitments then run beside the earlier groups' hashing, switch, rows, challenges and
    // folds, which leave most CUs idle, instead of beside its own.
    const int groups = std::max(1, std::min({BATCH_GROUPS, n / BATCH_GROUP_MIN, threads}));
    std::vector<std::vector<int>> g(groups);
    for (int k = 0; k < n; k++) g[k % groups].push_back(order[k]);
    std::mutex m;
    std::condition_variable cv;
    int started = 1;  // groups released
    auto release = [&](int next) {
        {
            std::lock_guard<std::mutex> lk(m);
            started = std::max(started, next + 1);
        }
        cv.notify_all();
    };
    std::vector<std::thread> later;
    for (int q = 1; q < groups; q++)
        later.emplace_back([&, q] {
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return started > q; });
            }
            batch_group(items, jobs, g[q], threads / groups, [&, q] { release(q + 1); });
        });
    batch_group(items, jobs, g[0], threads - (groups - 1) * (threads / groups), [&] { release(1); });
    for (auto &t : later) t.join();
    int first = LBN_OK;
    for (int i = 0; i < n; i++) {
        if (first == LBN_OK && items[i].status != LBN_OK) first = items[i].status;
        jobs[i].reset();  // a failed proof waits for its device work here
    }
    return first;
}

// ============================================================================= checker
// verify (verifier/proof_verify/src/main.rs:381-389) with check_proof_shape, check_field_switch and
// check_opening (main.rs:293-379) over the bincode bytes of `Proof` (main.rs:98-124). The GHASH
// trace is regenerated on the device from the seed (main.rs:135-160) and the device computes the
// trusted partials, the folding challenges and the integer sums of the ring identity; the
// transcript, the switch verifier, the F162 checks and the residues of the sums stay on the host.

namespace {
constexpr double FOLD_CAP = 108.0;                // pcs/src/recursion/mod.rs:81
constexpr u128 MIN_MODULUS = 34816965049ull;       // verifier/proof_verify/src/main.rs:54
constexpr u128 MAX_MODULUS = (u128)1 << 96;        // verifier/proof_verify/src/rq.rs:25
constexpr size_t MAX_PROOF_BYTES = 64ull << 20;    // the codec's byte limit (main.rs:119-124)

// bincode's fixed-width little-endian reads; a read past the end marks the reader bad.
struct Reader {
    const uint8_t *p;
    size_t len, at = 0;
    bool bad = false;
    const uint8_t *take(uint64_t n) {
        if (bad || n > len - at) {
            bad = true;
            return nullptr;
        }
        const uint8_t *r = p + at;
        at += (size_t)n;
        return r;
    }
    uint64_t u64() {
        const uint8_t *b = take(8);
        uint64_t x = 0;
        if (b)
            for (int i = 0; i < 8; i++) x |= (uint64_t)b[i] << (8 * i);
        return x;
    }
};

struct ProofFields {
    u128 q = 0;
    const uint8_t *commitment = nullptr, *row = nullptr, *folded = nullptr;
    uint64_t clen = 0, row_len = 0, folded_len = 0;
    std::vector<B128> partials;
    std::vector<F162> rounds;  // [e0, einf] per round
    F162 opened{};
};

// The proof's fields; a message when the bytes are not one `Proof` of this shape.
const char *decode_proof(const uint8_t *buf, size_t len, int lp, ProofFields &P) {
    if (len > MAX_PROOF_BYTES) return "proof too large";
    Reader r{buf, len};
    const uint8_t *v = r.take(4);
    if (!v || v[0] != 1 || v[1] || v[2] || v[3]) return "bad version";
    const uint64_t qlo = r.u64(), qhi = r.u64();
    P.q = ((u128)qhi << 64) | qlo;
    P.clen = r.u64();
    P.commitment = r.take(P.clen);
    if (r.u64() != 128 || r.bad) return "128 partial evaluations";
    P.partials.resize(128);
    for (auto &x : P.partials) x.w[0] = r.u64(), x.w[1] = r.u64();
    if (r.u64() != (uint64_t)lp || r.bad) return "one switch round per packed coordinate";
    P.rounds.resize(2 * (size_t)lp);
    for (auto &x : P.rounds)
        for (int k = 0; k < 3; k++) x.w[k] = r.u64();
    for (int k = 0; k < 3; k++) P.opened.w[k] = r.u64();
    P.row_len = r.u64();
    P.row = r.take(P.row_len);
    P.folded_len = r.u64();
    P.folded = r.take(P.folded_len);
    if (r.bad || r.at != len) return "malformed proof bytes";
    return nullptr;
}

uint64_t fold_cap(const Shape &s) {  // main.rs:62-64 at this shape
    return (uint64_t)std::ceil(FOLD_CAP * (double)(((uint64_t)1 << s.log_packed) / 4 * N));
}

uint64_t isqrt(uint64_t x) {
    uint64_t r = (uint64_t)std::sqrt((double)x);
    while (r * r > x) r--;
    while ((r + 1) * (r + 1) <= x) r++;
    return r;
}

bool canonical(const F162 &x) { return (x.w[2] >> 34) == 0; }

// claim (main.rs:184-190): the partials weighted by eq(r_lo, .).
B128 claim_of(const B128 *partials, const B128 *r_lo) {
    B128 eq[128];
    eq_expand_b128(r_lo, 7, eq);
    B128 acc{{0, 0}};
    for (int b = 0; b < 128; b++) acc = b128_add(acc, b128_mul(partials[b], eq[b]));
    return acc;
}

// slice_sum (crossfield.rs:625-637): sum over bit positions k of the batch weight times the
// GHASH element whose bit i is bit k of partial i, lifted to F162.
F162 slice_sum(const B128 *v, const F162 *batch) {
    F162 s{{0, 0, 0}};
    for (int k = 0; k < 128; k++) {
        B128 u{{0, 0}};
        for (int i = 0; i < 128; i++) u.w[i >> 6] |= ((v[i].w[k >> 6] >> (k & 63)) & 1) << (i & 63);
        s = f162_add(s, f162_mul(f162_lift(u), batch[k]));
    }
    return s;
}

// transparent_coeff (crossfield.rs:99-108 over the Tensor of crossfield.rs:10-63): the product
// over i of (1 + r_pp[i] (x) 1 + 1 (x) r_hi[i]) in F162 (x) GF(2^128), held as the F162
// coefficients of x^c, c < 128, on the GHASH side (x^128 = x^7 + x^2 + x + 1), then paired with
// the batch weights.
F162 transparent_coeff(const B128 *r_hi, const F162 *r_pp, int l, const F162 *batch) {
    const F162 zero{{0, 0, 0}};
    std::vector<F162> t(128, zero), h(128);
    t[0] = F162{{1, 0, 0}};
    for (int i = 0; i < l; i++) {
        // h = (1 (x) r_hi[i]) t by Horner over the bits of r_hi[i], most significant first
        std::fill(h.begin(), h.end(), zero);
        for (int k = 127; k >= 0; k--) {
            const F162 top = h[127];
            std::rotate(h.begin(), h.end() - 1, h.end());
            h[0] = top;
            h[1] = f162_add(h[1], top), h[2] = f162_add(h[2], top), h[7] = f162_add(h[7], top);
            if ((r_hi[i].w[k >> 6] >> (k & 63)) & 1)
                for (int c = 0; c < 128; c++) h[c] = f162_add(h[c], t[c]);
        }
        for (int c = 0; c < 128; c++) t[c] = f162_add(f162_add(t[c], f162_mul(t[c], r_pp[i])), h[c]);
    }
    F162 acc = zero;
    for (int c = 0; c < 128; c++) acc = f162_add(acc, f162_mul(t[c], batch[c]));
    return acc;
}

// wire::unpack_f162 (pcs/src/wire/mod.rs:456-470): 162 bits per element, least significant first.
void unpack_f162(const uint8_t *in, size_t n, F162 *out) {
    auto bits = [&](size_t at, int cnt) {  // cnt <= 64 bits from bit offset `at`
        uint64_t x = 0;
        for (int b = 0; b < cnt; b++) x |= (uint64_t)((in[(at + b) >> 3] >> ((at + b) & 7)) & 1) << b;
        return x;
    };
    for (size_t i = 0; i < n; i++) out[i] = F162{{bits(162 * i, 64), bits(162 * i + 64, 64), bits(162 * i + 128, 34)}};
}

// BLAKE3 chaining values of whole-chunk subtrees of host bytes: every chunk on the pool, then the
// parents of each subtree (the absorb of a proof's commitment, 3-8 MB).
void host_subtree_cvs(const uint8_t *data, uint64_t first, size_t count, const Blake3::Subtrees &trees,
                      const uint32_t key[8], uint32_t flags, uint32_t *cvs) {
    std::vector<uint32_t> leaf(count * 8);
    parallel_for(count, [&](size_t i) {
        uint32_t cv[8], m[16], out[16];
        memcpy(cv, key, 32);
        for (int b = 0; b < 16; b++) {
            memcpy(m, data + i * Blake3::CHUNK + 64 * b, 64);
            const uint32_t f = flags | (b == 0 ? Blake3::CHUNK_START : 0) | (b == 15 ? Blake3::CHUNK_END : 0);
            Blake3::compress(cv, m, first + i, 64, f, out);
            memcpy(cv, out, 32);
        }
        memcpy(&leaf[8 * i], cv, 32);
    });
    size_t at = 0;
    for (size_t t = 0; t < trees.size(); t++) {
        size_t n = (size_t)1 << trees[t].second;
        uint32_t *lv = &leaf[8 * at];
        at += n;
        for (; n > 1; n /= 2)
            for (size_t i = 0; i < n / 2; i++) {
                uint32_t out[16];
                Blake3::compress(key, lv + 16 * i, 0, 64, flags | Blake3::PARENT, out);
                memcpy(lv + 8 * i, out, 32);
            }
        memcpy(cvs + 8 * t, lv, 32);
    }
}

// Residues modulo q < 2^96.
u128 add_mod(u128 a, u128 b, u128 q) {
    const u128 s = a + b;
    return s >= q ? s - q : s;
}
u128 mul_mod(u128 a, u128 b, u128 q) {  // b in 32-bit digits, most significant first
    u128 r = 0;
    for (int i = 2; i >= 0; i--) r = add_mod((r << 32) % q, (a * (uint32_t)(b >> (32 * i))) % q, q);
    return r;
}
u128 signed_mod(__int128 x, u128 q) {
    const __int128 m = x % (__int128)q;
    return (u128)(m < 0 ? m + (__int128)q : m);
}

// The ring element sum_e raw_e X^e (raw_e = sum_k 2^(32 k) s[3 e + k]) of R_Q, with
// X^648 = X^324 - 1 applied from the top (rq.rs:152-165).
std::vector<u128> ring_of_limbs(const __int128 *s, u128 q) {
    const u128 p32 = ((u128)1 << 32) % q, p64 = ((u128)1 << 64) % q;
    std::vector<u128> raw(2 * N - 1);
    for (int e = 0; e < 2 * N - 1; e++)
        raw[e] = add_mod(add_mod(signed_mod(s[3 * e], q), mul_mod(signed_mod(s[3 * e + 1], q), p32, q), q),
                         mul_mod(signed_mod(s[3 * e + 2], q), p64, q), q);
    for (int k = 2 * N - 2; k >= N; k--) {
        raw[k - N / 2] = add_mod(raw[k - N / 2], raw[k], q);
        raw[k - N] = add_mod(raw[k - N], q - raw[k], q);
    }
    raw.resize(N);
    return raw;
}
}  // namespace

extern "C" int lbn_verify(lbn_ctx *ctx, const uint8_t seed[32], const uint8_t *proof, size_t len) {
    if (!ctx || !seed || (!proof && len)) return LBN_E_ARG;
    const Shape &s = ctx->shape;
    const int lp = s.log_packed;
    auto reject = [&](const char *why) { return ctx->fail(LBN_E_REJECT, why); };
    ProofFields P;
    if (const char *why = decode_proof(proof, len, lp, P)) return reject(why);
    // accept_modulus (main.rs:85-97)
    const u128 q = P.q;
    if (q < MIN_MODULUS) return reject("commitment modulus is below the security bound");
    if (q <= 2 * (u128)isqrt(fold_cap(s))) return reject("commitment modulus does not clear the no-wrap bound");
    if (q >= MAX_MODULUS) return reject("commitment modulus is above what the checker's arithmetic covers");
    TRY(lbn_generate_table(ctx, seed));

    // transcript (main.rs:166-174)
    Transcript t("labinius-fast/v1");
    t.absorb_bytes(seed, 32);
    t.absorb_u64((uint64_t)s.log_vars);
    t.absorb_u64(COLUMN_LOG);
    uint8_t qb[16];
    for (int i = 0; i < 16; i++) qb[i] = (uint8_t)(q >> (8 * i));
    t.absorb_bytes(qb, 16);
    t.absorb_bytes_with(P.commitment, (size_t)P.clen,
                        [&](size_t off, uint64_t first, size_t count, const Blake3::Subtrees &trees, uint32_t *cvs) {
                            host_subtree_cvs(P.commitment + off, first, count, trees, t.hasher().key_words(),
                                             t.hasher().flags(), cvs);
                        });

    // check_field_switch (main.rs:306-334)
    const std::vector<B128> r_lo = t.sample_b128("r-lo", 7), r_hi = t.sample_b128("r-hi", lp);
    std::vector<B128> trusted(128);
    TRY(lbn_partials(ctx, (const uint64_t *)r_hi.data(), (uint64_t *)trusted.data()));
    const B128 value = claim_of(trusted.data(), r_lo.data());
    t.absorb_b128(&value, 1);
    t.absorb_b128(P.partials.data(), 128);
    const std::vector<F162> rb = t.sample_f162("switch-batch", 7);
    std::vector<F162> batch(128);
    eq_expand_f162(rb.data(), 7, batch.data());
    // SwitchVerifier::start, round, finish (crossfield.rs:645-690)
    const B128 re = claim_of(P.partials.data(), r_lo.data());
    if (re.w[0] != value.w[0] || re.w[1] != value.w[1]) return reject("partial evaluation mismatch");
    F162 sum = slice_sum(P.partials.data(), batch.data());
    std::vector<F162> coords(lp);
    for (int r = 0; r < lp; r++) {
        const F162 e0 = P.rounds[2 * r], einf = P.rounds[2 * r + 1];
        if (!canonical(e0) || !canonical(einf)) return reject("noncanonical F162 element");
        t.absorb_f162(&P.rounds[2 * r], 2);
        const F162 rho = t.sample_f162("switch-round", 1)[0];
        coords[r] = rho;
        const F162 e1 = f162_add(sum, e0);
        // s = e0 + rho (e0 + e1 + einf) + rho^2 einf (crossfield.rs:666-671)
        sum = f162_add(f162_add(e0, f162_mul(rho, f162_add(f162_add(e0, e1), einf))),
                       f162_mul(f162_mul(rho, rho), einf));
    }
    if (!canonical(P.opened)) return reject("noncanonical F162 element");
    {
        const F162 d = transparent_coeff(r_hi.data(), coords.data(), lp, batch.data());
        if (f162_eq(d, F162{{0, 0, 0}})) return reject("the switch left the opened value unbound");
        if (!f162_eq(sum, f162_mul(d, P.opened))) return reject("sumcheck final check failed");
    }
    t.absorb_f162(&P.opened, 1);

    // check_opening (main.rs:341-379); split_point (main.rs:257-264)
    std::vector<F162> p0(s.row_log), p1(COLUMN_LOG);
    for (int i = 0; i < s.row_log; i++) p0[i] = coords[lp - 1 - i];
    for (int i = 0; i < COLUMN_LOG; i++) p1[i] = coords[COLUMN_LOG - 1 - i];
    if (P.row_len != f162_bytes(COLUMNS)) return reject("malformed row evaluation");
    std::vector<F162> row(COLUMNS);
    unpack_f162(P.row, COLUMNS, row.data());
    const size_t nf = (size_t)N * s.ring_elements;
    if (P.folded_len != 4 * nf) return reject("folded witness has the wrong length");
    // unpack_folded (main.rs:229-248): i32 words that must fit i16
    std::vector<int16_t> v(nf);
    for (size_t i = 0; i < nf; i++) {
        const uint8_t *b = P.folded + 4 * i;
        const int32_t x = (int32_t)((uint32_t)b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 | (uint32_t)b[3] << 24);
        if (x < -32768 || x > 32767) return reject("coefficient out of range");
        v[i] = (int16_t)x;
    }
    {  // RowEvaluation::absorb (scheme/commitment.rs:320-331)
        static const char L[] = "labinius/row-evaluation";
        t.absorb_bytes(L, sizeof L - 1);
        t.absorb_bytes(row.data(), (size_t)24 * COLUMNS);
    }
    static const ChallengeSampler sampler;
    std::vector<lbn_challenge> ch(COLUMNS);
    {
        ChalArg ca = sampler.arg;
        t.save(ca.ts);
        TRY(dev_challenges(ctx, ca, ch.data()));
    }
    t.skip(COLUMNS);
    {
        std::vector<F162> eq1((size_t)1 << COLUMN_LOG);
        eq_table_f162(p1.data(), COLUMN_LOG, eq1.data());
        F162 acc{{0, 0, 0}};
        for (int c = 0; c < COLUMNS; c++) acc = f162_add(acc, f162_mul(eq1[c], row[c]));
        if (!f162_eq(acc, P.opened)) return reject("the row evaluation does not reduce to the opened value");
    }
    {
        uint64_t normsq = 0, worst = 0;
        for (size_t i = 0; i < nf; i++) {
            const uint64_t m = (uint64_t)std::abs((int)v[i]);
            normsq += m * m;
            worst = std::max(worst, m);
        }
        if (normsq > fold_cap(s)) return reject("the fold is longer than the cap");
        if ((u128)worst > (q - 1) / 2) return reject("the fold is not centered");
    }
    int width = 0;
    for (u128 x = q - 1; x; x >>= 8) width++;
    // rq::unpack_ring_elements (rq.rs:308-324)
    if (P.clen != (uint64_t)COLUMNS * N * width) return reject("the commitment has the wrong length");
    {
        std::atomic<bool> unreduced{false};
        const size_t per = 1 << 16, total = (size_t)COLUMNS * N;
        parallel_for((total + per - 1) / per, [&](size_t b) {
            for (size_t i = b * per; i < std::min(total, (b + 1) * per); i++) {
                u128 x = 0;
                for (int k = 0; k < width; k++) x |= (u128)P.commitment[i * width + k] << (8 * k);
                if (x >= q) unreduced = true;
            }
        });
        if (unreduced) return reject("a commitment coefficient is not reduced");
    }
    {
        std::vector<__int128> ls(3 * (2 * N - 1)), rs(3 * (2 * N - 1));
        TRY(dev_verify_products(ctx, q, v.data(), P.commitment, width, ch.data(), ls.data(), rs.data()));
        if (ring_of_limbs(ls.data(), q) != ring_of_limbs(rs.data(), q))
            return reject("the fold does not open the folded commitment");
    }
    {
        // components_mod_2 (main.rs:268-280) under eq(p0) against the challenges' parities on the row
        const size_t rows = 4 * s.ring_elements;
        std::vector<F162> eq0(rows);
        eq_table_f162(p0.data(), s.row_log, eq0.data());
        F162 lhs{{0, 0, 0}}, rhs{{0, 0, 0}};
        for (size_t m = 0; m < s.ring_elements; m++)
            for (int k = 0; k < 4; k++) {
                F162 c{{0, 0, 0}};
                for (int p = 0; p < N / 4; p++) c.w[p >> 6] |= (uint64_t)(v[N * m + 4 * p + k] & 1) << (p & 63);
                lhs = f162_add(lhs, f162_mul(eq0[4 * m + k], c));
            }
        for (int j = 0; j < COLUMNS; j++) {
            F162 par{{0, 0, 0}};
            for (uint32_t i = 0; i < ch[j].weight; i++) par.w[ch[j].positions[i] >> 6] |= 1ull << (ch[j].positions[i] & 63);
            rhs = f162_add(rhs, f162_mul(par, row[j]));
        }
        if (!f162_eq(lhs, rhs)) return reject("the fold does not match the folded row evaluation");
    }
    return LBN_OK;
}
