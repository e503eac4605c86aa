/* ORACLE — TEST INFRASTRUCTURE ONLY. Command line of the CPU oracle:
 *   labinius-oracle --write-table FILE        (verifier/table_writer/src/main.rs:14-22)
 *   labinius-oracle --prove FILE [--table FILE] [--modulus Q]
 *                                             (verifier/proof_verify/src/main.rs:423-430; the
 *                                              prover's form, solution/proof_generate/src/main.rs:2056-2069)
 *   labinius-oracle --verify FILE             (verifier/proof_verify/src/main.rs:431-455)
 * Seed from LABINIUS_SEED (64 hex), shape from LABINIUS_LOG_VARS (default 31). */
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int seed_from_env(uint8_t seed[32]) {
    const char *s = getenv("LABINIUS_SEED");
    if (!s || strlen(s) != 64) return -1;
    for (int i = 0; i < 32; i++) {
        unsigned v;
        if (sscanf(s + 2 * i, "%2x", &v) != 1) return -1;
        seed[i] = (uint8_t)v;
    }
    return 0;
}

static int parse_u128(const char *s, u128 *out) {
    u128 v = 0;
    if (!*s) return -1;
    for (; *s; s++) {
        if (*s < '0' || *s > '9') return -1;
        u128 n = v * 10 + (u128)(*s - '0');
        if (n / 10 != v) return -1;
        v = n;
    }
    *out = v;
    return 0;
}

int main(int argc, char **argv) {
    uint8_t seed[32];
    if (seed_from_env(seed)) { fprintf(stderr, "LABINIUS_SEED is required (64 hex)\n"); return 2; }
    int k = getenv("LABINIUS_LOG_VARS") ? atoi(getenv("LABINIUS_LOG_VARS")) : 31;
    orc_shape s;
    if (orc_shape_of(k, &s)) { fprintf(stderr, "unsupported LABINIUS_LOG_VARS\n"); return 2; }
    if (argc == 3 && !strcmp(argv[1], "--write-table")) {
        uint8_t *t = malloc(s.table_bytes);
        orc_table_bytes(seed, 0, t, s.table_bytes);
        FILE *f = fopen(argv[2], "wb");
        if (!f || fwrite(t, 1, s.table_bytes, f) != s.table_bytes) { fprintf(stderr, "write failed\n"); return 1; }
        fclose(f);
        return 0;
    }
    if (argc >= 3 && !strcmp(argv[1], "--prove")) {
        /* --prove FILE [--table FILE] [--modulus Q], as the prover's command line
         * (solution/proof_generate/src/main.rs:2056-2069); without --table the table is the seed's */
        const char *table_path = NULL;
        u128 q = 35039150080ull; /* solution/proof_generate/src/parameters.rs:14 */
        for (int i = 3; i < argc; i += 2) {
            if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", argv[i]); return 2; }
            if (!strcmp(argv[i], "--table")) table_path = argv[i + 1];
            else if (!strcmp(argv[i], "--modulus")) {
                if (parse_u128(argv[i + 1], &q)) { fprintf(stderr, "invalid modulus\n"); return 2; }
            } else { fprintf(stderr, "unknown argument %s\n", argv[i]); return 2; }
        }
        uint8_t *table = NULL;
        if (table_path) {
            FILE *tf = fopen(table_path, "rb");
            if (!tf) { fprintf(stderr, "cannot open table\n"); return 1; }
            table = malloc(s.table_bytes + 1);
            size_t got = fread(table, 1, s.table_bytes + 1, tf);
            fclose(tf);
            if (got != s.table_bytes) { fprintf(stderr, "table has the wrong length\n"); return 1; }
        }
        size_t cap = orc_proof_len(k, (uint64_t)q, (uint64_t)(q >> 64)), len = 0;
        if (!cap) { fprintf(stderr, "modulus outside the accepted range\n"); return 2; }
        uint8_t *p = malloc(cap);
        int rc = orc_prove(seed, k, (uint64_t)q, (uint64_t)(q >> 64), table, p, cap, &len);
        if (rc) { fprintf(stderr, "prove failed: %d\n", rc); return 1; }
        FILE *f = fopen(argv[2], "wb");
        if (!f || fwrite(p, 1, len, f) != len) { fprintf(stderr, "write failed\n"); return 1; }
        fclose(f);
        return 0;
    }
    if (argc == 3 && !strcmp(argv[1], "--verify")) {
        FILE *f = fopen(argv[2], "rb");
        if (!f) { fprintf(stderr, "cannot open proof\n"); return 1; }
        fseek(f, 0, SEEK_END);
        long n = ftell(f);
        fseek(f, 0, SEEK_SET);
        if (n < 0 || n > 64l * 1024 * 1024) { fprintf(stderr, "proof too large\n"); return 1; }
        uint8_t *p = malloc((size_t)n + 1);
        if (fread(p, 1, (size_t)n, f) != (size_t)n) { fprintf(stderr, "read failed\n"); return 1; }
        fclose(f);
        char err[256];
        int rc = orc_verify(seed, k, p, (size_t)n, err, sizeof err);
        if (rc) { fprintf(stderr, "rejected: %s\n", err); return 1; }
        u128 q = 0;
        for (int i = 0; i < 16; i++) q |= (u128)p[4 + i] << (8 * i);
        char qs[48]; int at = 47; qs[at] = 0; u128 x = q;
        do { qs[--at] = (char)('0' + (int)(x % 10)); x /= 10; } while (x);
        printf("{\"modulus\":\"%s\",\"modulusBits\":%d,\"polynomialSize\":%llu,\"packedElements\":%llu,\"field\":\"GHASH\"}\n",
               qs + at, orc_modulus_bits(q), 1ull << k, 1ull << (k - 7));
        return 0;
    }
    fprintf(stderr, "usage: labinius-oracle --write-table FILE | --prove FILE [--table FILE] [--modulus Q] | --verify FILE\n");
    return 2;
}
