// Host-code stress under ThreadSanitizer or AddressSanitizer + UBSan (built by tests/native/Makefile
// against device_stub.cpp instead of the device half): the shared worker pool from many callers at
// once, the scheduler with twelve workers proving jobs of mixed sizes (same seed, same bytes), the
// checker's decoder on truncated, mutated and random proof bytes, and the transcript's absorb with
// chunks hashed elsewhere against the plain absorb. Exits non-zero on a failed check; the
// sanitizers report their own findings. Test infrastructure only.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "../../labinius-fast-challenge_amd/csrc/lbn_internal.h"

using namespace lbn;

static int failures = 0;
static int scale = 4;  // iterations in quarters: 4 by default, fewer under ThreadSanitizer
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            fprintf(stderr, "check failed %s:%d: %s\n", __FILE__, __LINE__, #c); \
            failures++;                                                   \
        }                                                                 \
    } while (0)

static void pool_stress() {
    std::vector<std::thread> callers;
    std::atomic<int> bad{0};
    for (int c = 0; c < 8; c++)
        callers.emplace_back([c, &bad] {
            std::mt19937 rng(c);
            for (int it = 0; it < 40 * scale; it++) {
                const size_t n = rng() % 3000 + 1;
                std::vector<std::atomic<int>> hit(n);
                for (auto &h : hit) h = 0;
                std::atomic<uint64_t> sum{0};
                parallel_for(n, [&](size_t i) {
                    hit[i]++;
                    sum += i;
                });
                if (sum != n * (n - 1) / 2) bad++;
                for (auto &h : hit)
                    if (h != 1) bad++;
            }
        });
    for (auto &t : callers) t.join();
    CHECK(bad == 0);
}

// max_batch > 1: continuous batching (several admitted jobs proved by lbn_prove_batch).
static void scheduler_stress(int max_batch) {
    lbn_sched *s = nullptr;
    CHECK(lbn_sched_create(0, 35039150080ull, 0, (size_t)64 << 30, max_batch > 1 ? 3 : 12, &s) == LBN_OK);
    if (!s) return;
    if (max_batch > 1) CHECK(lbn_sched_set_batch(s, max_batch, 4) == LBN_OK);
    std::vector<lbn_job> jobs(12 + 15 * scale);
    std::vector<std::vector<uint8_t>> out(jobs.size());
    for (size_t i = 0; i < jobs.size(); i++) {
        lbn_job &j = jobs[i];
        memset(&j, 0, sizeof j);
        j.log_vars = 19 + (int)(i % 5);
        for (int b = 0; b < 32; b++) j.seed[b] = (uint8_t)(i % 12 + b);  // every seed twice at least
        out[i].resize(lbn_proof_len(j.log_vars, 35039150080ull, 0));
        j.proof = out[i].data();
        j.cap = out[i].size();
        CHECK(lbn_sched_submit(s, &j) == LBN_OK);
        if (i == jobs.size() / 2) CHECK(lbn_sched_prewarm(s, 21, 3) == LBN_OK);
    }
    CHECK(lbn_sched_drain(s) == LBN_OK);
    for (size_t i = 0; i < jobs.size(); i++) {
        CHECK(lbn_sched_wait(s, &jobs[i]) == LBN_OK);
        for (size_t k = 0; k < i; k++)
            if (jobs[k].log_vars == jobs[i].log_vars && !memcmp(jobs[k].seed, jobs[i].seed, 32))
                CHECK(jobs[k].len == jobs[i].len && !memcmp(out[k].data(), out[i].data(), jobs[i].len));
    }
    lbn_sched_stats st{};
    CHECK(lbn_sched_get_stats(s, &st) == LBN_OK);
    CHECK(st.completed == jobs.size() && st.failed == 0);
    lbn_sched_destroy(s);
}

// The batch prover's phases on its host threads: every item's proof equals the same context's solo
// proof (the stub's outputs are deterministic), over batches of several sizes and thread counts.
static void batch_stress() {
    const int n = 24;  // the full batch runs as up to three groups, the half batch as one
    std::vector<lbn_ctx *> ctx(n, nullptr);
    for (int i = 0; i < n; i++) {
        CHECK(lbn_create(0, 19 + i % 4, &ctx[i]) == LBN_OK);
        if (!ctx[i]) return;
        CHECK(lbn_set_modulus(ctx[i], 35039150080ull, 0) == LBN_OK);
    }
    std::vector<std::vector<uint8_t>> solo(n), out(n);
    for (int i = 0; i < n; i++) {
        uint8_t seed[32] = {(uint8_t)i, 7};
        CHECK(lbn_generate_table(ctx[i], seed) == LBN_OK);
        solo[i].resize(lbn_proof_len(19 + i % 4, 35039150080ull, 0));
        size_t len = 0;
        CHECK(lbn_prove(ctx[i], seed, solo[i].data(), solo[i].size(), &len, nullptr) == LBN_OK);
        out[i].resize(solo[i].size());
    }
    for (int rep = 0; rep < 2 * scale; rep++) {
        const int m = rep % 2 ? n : n / 2 + 1, threads = 1 + rep % 5;
        std::vector<lbn_batch_item> items(m);
        for (int i = 0; i < m; i++) {
            memset(&items[i], 0, sizeof items[i]);
            items[i].ctx = ctx[i];
            items[i].seed[0] = (uint8_t)i, items[i].seed[1] = 7;
            items[i].gen_table = 1;
            items[i].proof = out[i].data();
            items[i].cap = out[i].size();
        }
        CHECK(lbn_prove_batch(items.data(), m, threads) == LBN_OK);
        for (int i = 0; i < m; i++) CHECK(items[i].len == solo[i].size() && out[i] == solo[i]);
    }
    for (lbn_ctx *c : ctx) lbn_destroy(c);
}

static void checker_decode_fuzz() {
    lbn_ctx *ctx = nullptr;
    CHECK(lbn_create(0, 20, &ctx) == LBN_OK);
    if (!ctx) return;
    CHECK(lbn_set_modulus(ctx, 35039150080ull, 0) == LBN_OK);
    uint8_t seed[32] = {1, 2, 3};
    CHECK(lbn_generate_table(ctx, seed) == LBN_OK);
    std::vector<uint8_t> proof(lbn_proof_len(20, 35039150080ull, 0));
    size_t len = 0;
    CHECK(lbn_prove(ctx, seed, proof.data(), proof.size(), &len, nullptr) == LBN_OK);
    CHECK(len == proof.size());
    const int whole = lbn_verify(ctx, seed, proof.data(), len);
    CHECK(whole == LBN_OK || whole == LBN_E_REJECT);
    std::mt19937_64 rng(99);
    for (int it = 0; it < 75 * scale; it++) {
        std::vector<uint8_t> bad(proof.begin(), proof.begin() + (rng() % (len + 1)));  // truncated copies
        CHECK(lbn_verify(ctx, seed, bad.data(), bad.size()) == LBN_E_REJECT);
    }
    for (int it = 0; it < 75 * scale; it++) {  // mutated length fields and bytes anywhere
        std::vector<uint8_t> bad(proof);
        const size_t at = rng() % len;
        bad[at] ^= (uint8_t)(1 + rng() % 255);
        if (it % 3 == 0) {
            const size_t fields[] = {20, 28 + (size_t)COLUMNS * N * 5, len - 4 * (size_t)N * 2 - 8};
            const uint64_t v = rng() % 4 ? rng() : ~0ull;
            memcpy(bad.data() + fields[it % 3 == 0 ? rng() % 3 : 0], &v, 8);
        }
        const int rc = lbn_verify(ctx, seed, bad.data(), bad.size());
        CHECK(rc == LBN_OK || rc == LBN_E_REJECT);
    }
    for (int it = 0; it < 25 * scale; it++) {
        std::vector<uint8_t> junk(rng() % 4096);
        for (auto &b : junk) b = (uint8_t)rng();
        CHECK(lbn_verify(ctx, seed, junk.data(), junk.size()) == LBN_E_REJECT);
    }
    lbn_destroy(ctx);
}

// Chaining value of a whole-chunk subtree of data (chunks first .. first + 2^level - 1).
static void subtree_cv(const uint8_t *data, uint64_t first, int level, const uint32_t key[8], uint32_t flags,
                       uint32_t out[8]) {
    std::vector<uint32_t> cvs((size_t)8 << level);
    for (size_t c = 0; c < ((size_t)1 << level); c++) {
        uint32_t cv[8], m[16], o[16];
        memcpy(cv, key, 32);
        for (int b = 0; b < 16; b++) {
            memcpy(m, data + c * Blake3::CHUNK + 64 * b, 64);
            Blake3::compress(cv, m, first + c, 64,
                             flags | (b == 0 ? Blake3::CHUNK_START : 0) | (b == 15 ? Blake3::CHUNK_END : 0), o);
            memcpy(cv, o, 32);
        }
        memcpy(&cvs[8 * c], cv, 32);
    }
    for (size_t n = (size_t)1 << level; n > 1; n /= 2)
        for (size_t i = 0; i < n / 2; i++) {
            uint32_t o[16];
            Blake3::compress(key, &cvs[16 * i], 0, 64, flags | Blake3::PARENT, o);
            memcpy(&cvs[8 * i], o, 32);
        }
    memcpy(out, cvs.data(), 32);
}

static void transcript_absorb() {
    std::mt19937 rng(5);
    for (int it = 0; it < 40; it++) {
        const size_t pre = rng() % 3000, n = rng() % 40000;
        std::vector<uint8_t> a(pre), b(n + 64);
        for (auto &x : a) x = (uint8_t)rng();
        for (auto &x : b) x = (uint8_t)rng();
        Blake3 h1, h2;
        h1.update(a.data(), a.size());
        h2.update(a.data(), a.size());
        h1.update(b.data(), n);
        h2.update_with(b.data(), n, [&](size_t off, uint64_t first, size_t count, const Blake3::Subtrees &trees, uint32_t *cvs) {
            size_t at = off;
            for (size_t t = 0; t < trees.size(); t++) {
                subtree_cv(b.data() + at, trees[t].first, trees[t].second, h2.key_words(), h2.flags(), cvs + 8 * t);
                at += Blake3::CHUNK << trees[t].second;
            }
            (void)first, (void)count;
        });
        uint8_t d1[64], d2[64];
        h1.xof(0, d1, 64);
        h2.xof(0, d2, 64);
        CHECK(!memcmp(d1, d2, 64));
    }
}

int main(int argc, char **argv) {
    if (argc > 1) scale = atoi(argv[1]);
    pool_stress();
    fprintf(stderr, "pool done\n");
    transcript_absorb();
    fprintf(stderr, "transcript done\n");
    checker_decode_fuzz();
    fprintf(stderr, "checker decode done\n");
    scheduler_stress(1);
    scheduler_stress(8);
    fprintf(stderr, "scheduler done\n");
    batch_stress();
    fprintf(stderr, "batch done, failures %d\n", failures);
    return failures ? 1 : 0;
}
