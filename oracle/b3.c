/* ORACLE — TEST INFRASTRUCTURE ONLY. Never linked into the product library.
 *
 * BLAKE3, restated from the published specification (Aumasson, Neves, O'Connor, Wilcox-O'Hearn,
 * "BLAKE3: one function, fast everywhere", 2020): compression function, chunk chaining, binary
 * tree with a chaining-value stack, keyed and derive-key modes, and the extendable output read
 * from any byte offset.
 *
 * The reference reaches BLAKE3 through the `blake3` crate 1.8.7
 * (solution/proof_generate/Cargo.lock:21-23), which is not vendored under the reference tree.
 * Call sites on the proving path: verifier/proof_verify/src/main.rs:135-141 (table),
 * verifier/proof_verify/src/rq.rs:62-75 (matrix expansion),
 * solution/labinius/crates/pcs/src/challenge.rs:226-326,402-441,721-749 (transcript, sign key).
 * Pinned against libLLVM's independent BLAKE3 by tests/golden (hash, keyed, derive-key, XOF).
 */
#include "b3.h"
#include <string.h>

static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                               0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};

static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void g(uint32_t *s, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
    s[a] = s[a] + s[b] + mx;
    s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + my;
    s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 7);
}

void b3_compress(const uint32_t cv[8], const uint8_t block[64], uint64_t counter,
                 uint32_t block_len, uint32_t flags, uint32_t out[16]) {
    uint32_t m[16], s[16], t[16];
    for (int i = 0; i < 16; i++)
        m[i] = (uint32_t)block[4 * i] | ((uint32_t)block[4 * i + 1] << 8) |
               ((uint32_t)block[4 * i + 2] << 16) | ((uint32_t)block[4 * i + 3] << 24);
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    for (int i = 0; i < 4; i++) s[8 + i] = IV[i];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
    for (int r = 0; r < 7; r++) {
        g(s, 0, 4, 8, 12, m[0], m[1]);
        g(s, 1, 5, 9, 13, m[2], m[3]);
        g(s, 2, 6, 10, 14, m[4], m[5]);
        g(s, 3, 7, 11, 15, m[6], m[7]);
        g(s, 0, 5, 10, 15, m[8], m[9]);
        g(s, 1, 6, 11, 12, m[10], m[11]);
        g(s, 2, 7, 8, 13, m[12], m[13]);
        g(s, 3, 4, 9, 14, m[14], m[15]);
        for (int i = 0; i < 16; i++) t[i] = m[PERM[i]];
        memcpy(m, t, sizeof m);
    }
    for (int i = 0; i < 8; i++) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

static void words_from_bytes(const uint8_t *b, uint32_t *w, int n) {
    for (int i = 0; i < n; i++)
        w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
               ((uint32_t)b[4 * i + 3] << 24);
}

static void init_base(b3_hasher *h, const uint32_t key[8], uint32_t flags) {
    memset(h, 0, sizeof *h);
    memcpy(h->key, key, 32);
    memcpy(h->cv, key, 32);
    h->flags = flags;
}

void b3_init(b3_hasher *h) { init_base(h, IV, 0); }

void b3_init_keyed(b3_hasher *h, const uint8_t key[32]) {
    uint32_t k[8];
    words_from_bytes(key, k, 8);
    init_base(h, k, B3_KEYED_HASH);
}

void b3_init_derive_key(b3_hasher *h, const char *context, size_t context_len) {
    b3_hasher c;
    init_base(&c, IV, B3_DERIVE_KEY_CONTEXT);
    b3_update(&c, context, context_len);
    uint8_t ck[32];
    b3_finalize_seek(&c, 0, ck, 32);
    uint32_t k[8];
    words_from_bytes(ck, k, 8);
    init_base(h, k, B3_DERIVE_KEY_MATERIAL);
}

/* The chaining value of the current (complete) chunk, which is not the last chunk. */
static void chunk_end_cv(const b3_hasher *h, uint32_t cv[8]) {
    uint32_t out[16];
    uint8_t block[64];
    memset(block, 0, 64); /* a partial last block is zero padded */
    memcpy(block, h->buf, h->buf_len);
    uint32_t f = h->flags | B3_CHUNK_END | (h->blocks_done == 0 ? B3_CHUNK_START : 0);
    b3_compress(h->cv, block, h->chunk_counter, (uint32_t)h->buf_len, f, out);
    memcpy(cv, out, 32);
}

static void parent_cv(const uint32_t key[8], uint32_t flags, const uint32_t l[8], const uint32_t r[8],
                      uint32_t cv[8]) {
    uint8_t block[64];
    for (int i = 0; i < 8; i++) {
        for (int b = 0; b < 4; b++) {
            block[4 * i + b] = (uint8_t)(l[i] >> (8 * b));
            block[32 + 4 * i + b] = (uint8_t)(r[i] >> (8 * b));
        }
    }
    uint32_t out[16];
    b3_compress(key, block, 0, 64, flags | B3_PARENT, out);
    memcpy(cv, out, 32);
}

/* Push a finished chunk's chaining value; merge complete subtrees (the spec's lazy merge). */
static void push_chunk_cv(b3_hasher *h, uint32_t cv[8], uint64_t total_chunks) {
    while ((total_chunks & 1) == 0) {
        h->stack_len--;
        parent_cv(h->key, h->flags, h->stack[h->stack_len], cv, cv);
        total_chunks >>= 1;
    }
    memcpy(h->stack[h->stack_len], cv, 32);
    h->stack_len++;
}

void b3_update(b3_hasher *h, const void *data, size_t len) {
    const uint8_t *p = (const uint8_t *)data;
    while (len > 0) {
        /* A full chunk followed by more input: finalize it into the tree. */
        if (h->blocks_done * 64 + h->buf_len == B3_CHUNK_LEN) {
            uint32_t cv[8];
            chunk_end_cv(h, cv);
            push_chunk_cv(h, cv, h->chunk_counter + 1);
            h->chunk_counter++;
            h->blocks_done = 0;
            h->buf_len = 0;
            memcpy(h->cv, h->key, 32);
        }
        /* A full block followed by more input: compress it into the chunk state. */
        if (h->buf_len == 64) {
            uint32_t out[16];
            uint32_t f = h->flags | (h->blocks_done == 0 ? B3_CHUNK_START : 0);
            b3_compress(h->cv, h->buf, h->chunk_counter, 64, f, out);
            memcpy(h->cv, out, 32);
            h->blocks_done++;
            h->buf_len = 0;
        }
        size_t take = 64 - h->buf_len;
        if (take > len) take = len;
        memcpy(h->buf + h->buf_len, p, take);
        h->buf_len += take;
        p += take;
        len -= take;
    }
}

/* The root node: its input chaining value, block, length and flags. XOF block i is the root
 * compression with counter i. */
static void root_node(const b3_hasher *h, uint32_t cv[8], uint8_t block[64], uint32_t *blen,
                      uint32_t *flags) {
    if (h->stack_len == 0) {
        memcpy(cv, h->cv, 32);
        memset(block, 0, 64);
        memcpy(block, h->buf, h->buf_len);
        *blen = (uint32_t)h->buf_len;
        *flags = h->flags | B3_CHUNK_END | (h->blocks_done == 0 ? B3_CHUNK_START : 0);
        return;
    }
    uint32_t right[8];
    chunk_end_cv(h, right);
    int i = h->stack_len - 1;
    while (i > 0) {
        parent_cv(h->key, h->flags, h->stack[i], right, right);
        i--;
    }
    memcpy(cv, h->key, 32);
    for (int k = 0; k < 8; k++) {
        for (int b = 0; b < 4; b++) {
            block[4 * k + b] = (uint8_t)(h->stack[0][k] >> (8 * b));
            block[32 + 4 * k + b] = (uint8_t)(right[k] >> (8 * b));
        }
    }
    *blen = 64;
    *flags = h->flags | B3_PARENT;
}

void b3_finalize_seek(const b3_hasher *h, uint64_t offset, uint8_t *out, size_t len) {
    uint32_t cv[8], blen, flags;
    uint8_t block[64];
    root_node(h, cv, block, &blen, &flags);
    uint64_t counter = offset / 64;
    size_t skip = (size_t)(offset % 64);
    while (len > 0) {
        uint32_t w[16];
        uint8_t bytes[64];
        b3_compress(cv, block, counter, blen, flags | B3_ROOT, w);
        for (int i = 0; i < 16; i++)
            for (int b = 0; b < 4; b++) bytes[4 * i + b] = (uint8_t)(w[i] >> (8 * b));
        size_t take = 64 - skip;
        if (take > len) take = len;
        memcpy(out, bytes + skip, take);
        out += take;
        len -= take;
        skip = 0;
        counter++;
    }
}

void b3_hash(const void *data, size_t len, uint8_t out[32]) {
    b3_hasher h;
    b3_init(&h);
    b3_update(&h, data, len);
    b3_finalize_seek(&h, 0, out, 32);
}
