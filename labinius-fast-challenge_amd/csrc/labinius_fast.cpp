// labinius-fast: the prover executable, a drop-in for the reference's
//   labinius-fast --prove PROOF --table TABLE [--modulus Q]        (main.rs:2050-2090)
// with the seed in LABINIUS_SEED (64 hex characters, main.rs:76-80). It reads the witness table
// (exactly 2^k/8 bytes), proves on the GPU through liblabinius_gpu and writes the bincode proof.
// The profile line's load_us is spawn to the end of the table read, which overlaps the GPU set-up
// (alloc_us: runtime start and the pinned buffer, create_us: the context, modulus_us: the public
// matrix); the upload runs inside the proof.
// Extensions: LABINIUS_LOG_VARS (default 31) selects a smaller shape; LABINIUS_DEVICE the GPU.
// Any failure exits non-zero with a message, as the reference's aborts do.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/labinius_gpu.h"

typedef unsigned __int128 u128;

static void die(const char *msg) {
    fprintf(stderr, "labinius-fast: %s\n", msg);
    exit(101);
}

static bool parse_seed(const char *s, uint8_t out[32]) {
    if (!s || strlen(s) != 64) return false;
    for (int i = 0; i < 32; i++) {
        unsigned v;
        if (sscanf(s + 2 * i, "%2x", &v) != 1) return false;
        out[i] = (uint8_t)v;
    }
    return true;
}

static bool parse_u128(const char *s, u128 *out) {
    while (*s == ' ' || *s == '\t' || *s == '\n') s++;
    u128 v = 0;
    int digits = 0;
    for (; *s >= '0' && *s <= '9'; s++, digits++) {
        u128 n = v * 10 + (u128)(*s - '0');
        if (n / 10 != v) return false;
        v = n;
    }
    while (*s == ' ' || *s == '\t' || *s == '\n') s++;
    if (!digits || *s) return false;
    *out = v;
    return true;
}

int main(int argc, char **argv) {
    const auto start = std::chrono::steady_clock::now();
    if (!(argc == 5 || argc == 7)) die("usage: labinius-fast --prove FILE --table FILE [--modulus Q]");
    if (strcmp(argv[1], "--prove")) die("unknown command");
    if (strcmp(argv[3], "--table")) die("witness table path is required");
    uint8_t seed[32];
    if (!parse_seed(getenv("LABINIUS_SEED"), seed)) die("LABINIUS_SEED must be 64 hex characters");
    u128 q = (u128)35039150080ull;  // solution/proof_generate/src/parameters.rs:14
    if (argc == 7) {
        if (strcmp(argv[5], "--modulus")) die("unknown prover option");
        if (!parse_u128(argv[6], &q)) die("invalid commitment modulus");
    }
    const int k = getenv("LABINIUS_LOG_VARS") ? atoi(getenv("LABINIUS_LOG_VARS")) : 31;
    const int device = getenv("LABINIUS_DEVICE") ? atoi(getenv("LABINIUS_DEVICE")) : 0;
    if (k < 19 || k > 31) die("LABINIUS_LOG_VARS must be in [19, 31]");
    const size_t table_bytes = ((size_t)1 << k) / 8;

    // The table is read by eight threads into pinned memory while the GPU context and the public
    // matrix are set up, then uploaded with the proof in ring-element ranges that overlap the
    // commitment. (The pinned allocation is the first HIP call: it starts the runtime.)
    const int fd = open(argv[4], O_RDONLY);
    if (fd < 0) die("cannot open the witness table");
    struct stat st;
    if (fstat(fd, &st) != 0 || (size_t)st.st_size != table_bytes) die("witness table has the wrong length");
    std::vector<uint8_t> fallback;
    uint8_t *table = (uint8_t *)lbn_host_alloc(table_bytes);
    const bool pinned = table != nullptr;
    if (!pinned) {
        fallback.resize(table_bytes);
        table = fallback.data();
    }
    const auto allocated = std::chrono::steady_clock::now();
    std::atomic<bool> read_ok{true};
    std::atomic<long long> read_end_us{0};
    std::vector<std::thread> readers;
    const int nr = 8;
    for (int i = 0; i < nr; i++)
        readers.emplace_back([&, i] {
            size_t at = table_bytes / nr * i;
            const size_t end = i == nr - 1 ? table_bytes : table_bytes / nr * (i + 1);
            while (at < end) {
                const ssize_t r = pread(fd, table + at, end - at, (off_t)at);
                if (r <= 0) {
                    read_ok = false;
                    return;
                }
                at += (size_t)r;
            }
            const long long t = std::chrono::duration_cast<std::chrono::microseconds>(
                                    std::chrono::steady_clock::now() - start).count();
            long long prev = read_end_us.load();
            while (t > prev && !read_end_us.compare_exchange_weak(prev, t)) {
            }
        });
    lbn_ctx *ctx = nullptr;
    if (lbn_create(device, k, &ctx) != LBN_OK) die("cannot create the GPU context");
    const auto created = std::chrono::steady_clock::now();
    if (lbn_set_modulus(ctx, (uint64_t)q, (uint64_t)(q >> 64)) != LBN_OK) die(lbn_last_error(ctx));
    const auto modulus = std::chrono::steady_clock::now();
    for (auto &t : readers) t.join();
    close(fd);
    if (!read_ok) die("cannot read the witness table");
    const auto loaded = std::chrono::steady_clock::now();
    if (lbn_load_table_async(ctx, table, table_bytes) != LBN_OK) die(lbn_last_error(ctx));
    const auto setup = std::chrono::steady_clock::now();
    const size_t cap = lbn_proof_len(k, (uint64_t)q, (uint64_t)(q >> 64));
    std::vector<uint8_t> proof(cap);
    size_t len = 0;
    lbn_profile prof{};
    if (lbn_prove(ctx, seed, proof.data(), cap, &len, &prof) != LBN_OK) die(lbn_last_error(ctx));
    const auto proved = std::chrono::steady_clock::now();
    FILE *o = fopen(argv[2], "wb");
    if (!o || fwrite(proof.data(), 1, len, o) != len || fclose(o) != 0) die("cannot write the proof");
    const auto written = std::chrono::steady_clock::now();
    auto us = [](auto a, auto b) { return (long long)std::chrono::duration_cast<std::chrono::microseconds>(b - a).count(); };
    fprintf(stderr,
            "labinius-profile load_us=%lld setup_us=%lld commit_us=%.0f transcript_us=%.0f partials_us=%.0f "
            "switch_us=%.0f row_us=%.0f challenges_us=%.0f fold_us=%.0f encode_us=%.0f prove_us=%.0f write_us=%lld "
            "total_us=%lld commit_primes=%d alloc_us=%lld create_us=%lld modulus_us=%lld read_end_us=%lld\n",
            us(start, loaded), us(loaded, setup), prof.commit_us, prof.transcript_us, prof.partials_us, prof.switch_us,
            prof.row_us, prof.challenges_us, prof.fold_us, prof.encode_us, prof.total_us, us(proved, written),
            us(start, written), lbn_commit_primes(ctx), us(start, allocated), us(allocated, created),
            us(created, modulus), read_end_us.load());
    lbn_destroy(ctx);
    if (pinned) lbn_host_free(table);
    return 0;
}
