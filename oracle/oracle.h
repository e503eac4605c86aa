/* ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A plain, scalar CPU restatement of the reference Labinius prover and checker, used only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to check (never to produce)
 * the product's results. Nothing in the product links or calls it.
 *
 * Shape generalisation: `log_vars` = k in [19, 31]; at k = 31 every constant equals the
 * reference's (verifier/proof_verify/src/main.rs:32-69). For k < 31 the columns stay 1024,
 * the transcript absorbs k, the matrix binds RING_ELEMENTS = 2^(k-19), and fold_cap scales with
 * the witness length.
 */
#ifndef ORACLE_H
#define ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned __int128 u128;
typedef __int128 i128;

typedef struct { uint64_t w[2]; } orc_b128;
typedef struct { uint64_t w[3]; } orc_f162;

typedef struct {
    uint8_t positions[32];
    uint32_t signs;
    uint32_t weight;
} orc_challenge;

/* ---- shape ---- */
typedef struct {
    int log_vars, log_packed, column_log, columns, row_log;
    size_t column_len, ring_elements, n_words, table_bytes;
} orc_shape;
int orc_shape_of(int log_vars, orc_shape *s);

/* ---- fields (pcs/src/fields/scalar.rs:46-121) ---- */
void orc_b128_mul(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]);
void orc_f162_mul(const uint64_t a[3], const uint64_t b[3], uint64_t out[3]);
/* Bit-by-bit definitions (pcs/tests/fields.rs:33-72), for cross-checking the above. */
void orc_b128_mul_bitwise(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]);
void orc_f162_mul_bitwise(const uint64_t a[3], const uint64_t b[3], uint64_t out[3]);

/* ---- blake3 (test access) ---- */
void orc_blake3(int mode, const uint8_t *key_or_ctx, size_t ctx_len, const uint8_t *msg,
                size_t msg_len, uint64_t offset, uint8_t *out, size_t out_len);

/* ---- table (verifier/proof_verify/src/main.rs:130-153) ---- */
void orc_table_bytes(const uint8_t seed[32], uint64_t offset, uint8_t *out, size_t len);

/* ---- transcript probes for golden tests ---- */
/* Transcript::new(domain); absorb_bytes(msg); then sample_b128(label, n) twice, f162 once. */
void orc_transcript_probe(const uint8_t *domain, size_t dlen, const uint8_t *msg, size_t mlen,
                          const uint8_t *label, size_t llen, int n, uint64_t *b128_out /*2n*/,
                          uint64_t *f162_out /*3n*/);

/* ---- challenges (pcs/src/challenge.rs:396-715) ---- */
/* `count` consecutive sample_short_challenge(w=28, bound=12) after absorbing `msg` into a
 * transcript started with `domain`. */
void orc_challenges_probe(const uint8_t *domain, size_t dlen, const uint8_t *msg, size_t mlen,
                          int count, orc_challenge *out);
/* One unsigned attempt (attempt, challenge.rs:662-688) drawn from `stream` in place of the XOF;
 * returns the bytes consumed, or -1 when the stream was too short. */
long orc_attempt_from_bytes(const uint8_t *stream, size_t len, int weight, orc_challenge *out);
/* Attempts and XOF bytes consumed per challenge for `count` challenges (as orc_challenges_probe). */
void orc_challenges_accounting(const uint8_t *domain, size_t dlen, const uint8_t *msg, size_t mlen, int count,
                               uint32_t *attempts, uint64_t *consumed);
double orc_canonical_inf_norm_sq(const orc_challenge *c);
double orc_canonical_inf_norm_sq_naive(const orc_challenge *c);
uint32_t orc_sign_bits(const orc_challenge *c);

/* ---- ring R_Q (verifier/proof_verify/src/rq.rs) ---- */
int orc_modulus_bits(u128 q);
int orc_coefficient_bytes(u128 q);
void orc_ajtai_element(u128 q, size_t elements, size_t index, u128 *out648);
/* Commitment of one column (schoolbook, rq.rs:264-287), reduced to [0, q). */
void orc_commit_column(const uint8_t *table, int log_vars, u128 q, const u128 *matrix,
                       int column, u128 *out648);
void orc_commit_columns_packed(const uint8_t *table, int log_vars, uint64_t q_lo, uint64_t q_hi,
                               int first, int count, uint8_t *out);
/* Columns first .. first+count-1 of a witness of `elements` ring elements per column (4 words
 * each, column-major) committed under key row `row` (matrix elements row * elements + m), packed
 * [column][t]. */
void orc_commit_key_row(const uint8_t *words, size_t elements, int row, uint64_t q_lo, uint64_t q_hi, int first,
                        int count, uint8_t *out);

/* ---- switch and opening components ---- */
/* eq_expand_b128 (crossfield.rs:87-97): r[0] becomes the MSB of the index. */
void orc_eq_expand_b128(const orc_b128 *r, int n, orc_b128 *out);
void orc_eq_expand_f162(const orc_f162 *r, int n, orc_f162 *out);
/* eq_table (eval.rs:51-65): variable k is bit k of the index. */
void orc_eq_table(const orc_f162 *p, int n, orc_f162 *out);
/* partial_evals (crossfield.rs:152-165): v[k] = sum_j bit_k(w_j) eq_hi[j]. */
void orc_partials(const uint8_t *table, int log_vars, const uint64_t *r_hi, uint64_t *out);
/* One whole switch (crossfield.rs:563-623 + sumcheck.rs:89-139) driven by given challenges
 * instead of a transcript: rounds[24][2][3] and opened[3]. */
void orc_switch_run(const uint8_t *table, int log_vars, const uint64_t *r_hi, const uint64_t *batch,
                    const uint64_t *rhos, uint64_t *rounds, uint64_t *opened);
/* row_evaluate (reference.rs:52-70) for columns [first, first+count). p0 has k-17 entries. */
void orc_row_eval(const uint8_t *table, int log_vars, const uint64_t *p0, int first, int count,
                  uint64_t *out);
/* fold_element (reference.rs:78-115) for ring elements [first, first+count). 0 on success. */
int orc_fold(const uint8_t *table, int log_vars, const orc_challenge *ch, int first, int count,
             int32_t *out);
uint64_t orc_fold_cap(int log_vars);

/* ---- whole prover / checker ---- */
size_t orc_proof_len(int log_vars, uint64_t q_lo, uint64_t q_hi);
/* table may be NULL (then regenerated from the seed). Returns 0 on success. */
int orc_prove(const uint8_t seed[32], int log_vars, uint64_t q_lo, uint64_t q_hi,
              const uint8_t *table, uint8_t *proof, size_t cap, size_t *len);
/* Returns 0 when the proof is accepted; otherwise a nonzero code and a message. */
int orc_verify(const uint8_t seed[32], int log_vars, const uint8_t *proof, size_t len, char *err,
               size_t errlen);
/* The transcript-derived values a prover needs, replayed from a finished proof (for per-pass
 * parity tests): r_hi[k-7][2], batch[128][3], coords[k-7][3], challenges[1024]. */
int orc_replay(const uint8_t seed[32], int log_vars, const uint8_t *proof, size_t len,
               uint64_t *r_hi, uint64_t *batch, uint64_t *coords, orc_challenge *challenges);

#ifdef __cplusplus
}
#endif
#endif
