/* ORACLE — TEST INFRASTRUCTURE ONLY. BLAKE3 from the published specification. */
#ifndef ORACLE_B3_H
#define ORACLE_B3_H
#include <stddef.h>
#include <stdint.h>

#define B3_CHUNK_LEN 1024
#define B3_CHUNK_START 1u
#define B3_CHUNK_END 2u
#define B3_PARENT 4u
#define B3_ROOT 8u
#define B3_KEYED_HASH 16u
#define B3_DERIVE_KEY_CONTEXT 32u
#define B3_DERIVE_KEY_MATERIAL 64u

typedef struct {
    uint32_t key[8];
    uint32_t flags;
    uint32_t stack[54][8];
    int stack_len;
    uint32_t cv[8];       /* chaining value inside the current chunk */
    uint64_t chunk_counter;
    int blocks_done;      /* blocks of the current chunk already compressed */
    uint8_t buf[64];
    size_t buf_len;
} b3_hasher;

void b3_compress(const uint32_t cv[8], const uint8_t block[64], uint64_t counter,
                 uint32_t block_len, uint32_t flags, uint32_t out[16]);
void b3_init(b3_hasher *h);
void b3_init_keyed(b3_hasher *h, const uint8_t key[32]);
void b3_init_derive_key(b3_hasher *h, const char *context, size_t context_len);
void b3_update(b3_hasher *h, const void *data, size_t len);
/* Extendable output of everything absorbed so far, from byte `offset`; `h` is unchanged. */
void b3_finalize_seek(const b3_hasher *h, uint64_t offset, uint8_t *out, size_t len);
void b3_hash(const void *data, size_t len, uint8_t out[32]);
#endif
