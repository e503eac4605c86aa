/* labinius_gpu.h — C ABI of the MI355X Labinius prover (liblabinius_gpu.so).
 *
 * The entry points replace the kernel calls the reference prover's `prove()` makes
 * (solution/proof_generate/src/main.rs:1692-2009) and the whole-proof executable contract
 * (`labinius-fast --prove P --table T [--modulus Q]`, main.rs:2050-2090). Plain pointers and
 * sizes only; every buffer is caller-owned host memory unless a name says "device". Every call
 * returns 0 on success or a negative LBN_E_* code; lbn_last_error() describes the last failure
 * on that context. No call aborts the process.
 *
 * Threading: one context per host thread; calls on one context are serial (the Fiat-Shamir
 * order forces it). Contexts are independent and may run concurrently.
 *
 * Shape: log_vars = k in [19, 31]; k = 31 is the challenge shape (2^24 GHASH words, 1024
 * columns of 16,384 words, 4096 ring elements per column). Smaller k keep 1024 columns and
 * shrink the column length; they are this library's generalisation for tests and batching.
 */
#ifndef LABINIUS_GPU_H
#define LABINIUS_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBN_OK 0
#define LBN_E_ARG (-1)     /* invalid argument or shape */
#define LBN_E_STATE (-2)   /* call out of order (no table, no modulus, no partials, ...) */
#define LBN_E_DEVICE (-3)  /* HIP runtime error */
#define LBN_E_MODULUS (-4) /* modulus outside [34,816,965,049, 2^96) */
#define LBN_E_RANGE (-5)   /* a fold coefficient left i16 (cannot happen for honest input) */
#define LBN_E_IO (-6)      /* file I/O (CLI helpers) */
#define LBN_E_REJECT (-7)  /* lbn_verify: the proof does not verify (reason in lbn_last_error) */

typedef struct lbn_ctx lbn_ctx;

/* A weight-w signed challenge of R_162: coefficient positions[i] is -1 when bit i of signs is
 * set, +1 otherwise; positions sorted and distinct (pcs/src/challenge.rs:396-406). */
typedef struct {
    uint8_t positions[32];
    uint32_t signs;
    uint32_t weight;
} lbn_challenge;

/* Per-phase wall times of the last lbn_prove, microseconds (the reference prints the same split
 * as its `labinius-profile` line, main.rs:1999-2008). */
typedef struct {
    double commit_us, transcript_us, partials_us, switch_us, row_us, challenges_us, fold_us,
        encode_us, total_us;
} lbn_profile;

/* Context on HIP device `device` for shape `log_vars`: allocates the instance's HBM. */
int lbn_create(int device, int log_vars, lbn_ctx **out);
void lbn_destroy(lbn_ctx *ctx);
const char *lbn_last_error(const lbn_ctx *ctx);

/* The witness table: 2^log_vars / 8 bytes, word j = u128 LE of bytes 16j..16j+16
 * (Table::load, main.rs:387-429; decode_table, verifier/proof_verify/src/main.rs:143-153).
 * lbn_load_table copies host memory to HBM; lbn_load_table_device copies from device memory. */
int lbn_load_table(lbn_ctx *ctx, const void *host_bytes, size_t bytes);
int lbn_load_table_device(lbn_ctx *ctx, const void *device_bytes, size_t bytes);
/* The same table with its upload deferred into the proof, as the reference folds the table's page
 * faults into the commitment's first read (main.rs:387-429, 1736-1740): lbn_prove / lbn_commit copy
 * it in four ring-element ranges, the commitment MAC of each range starting once its copy is in;
 * any other entry point that reads the table copies it whole first. host_bytes must stay valid and
 * unchanged until that call returns. Pinned memory (lbn_host_alloc) lets the ranges go by DMA
 * while the MAC runs; other memory works, more slowly. */
int lbn_load_table_async(lbn_ctx *ctx, const void *host_bytes, size_t bytes);
/* Fill the table on the device from a 32-byte seed: the BLAKE3 XOF the trusted table writer
 * produces (verifier/table_writer/src/main.rs:14-22). */
int lbn_generate_table(lbn_ctx *ctx, const uint8_t seed[32]);

/* Commitment modulus Q (any value in [34,816,965,049, 2^96)); builds the public matrix
 * rq::ajtai_matrix(Q, ring_elements) (verifier/proof_verify/src/rq.rs:114-144) and its
 * transform into HBM. Public data: independent of the seed. */
int lbn_set_modulus(lbn_ctx *ctx, uint64_t q_lo, uint64_t q_hi);

/* Commitment A*W mod Q of the loaded table, packed ceil(bits(Q-1)/8) bytes LE per coefficient,
 * column-major (column c coefficient t at c*648+t). Replaces commit_columns_rows
 * (pcs/src/simd/sms_commit.rs:1473) + commitment_bytes_sms (main.rs:1080-1127). */
int lbn_commit(lbn_ctx *ctx, uint8_t *out, size_t out_len);

/* The 128 GHASH partial evaluations at r_hi (k-7 B128, as u64 pairs), and the eq expansion kept
 * on the device for the switch. Replaces eq_expand_and_partials_factored_blocks +
 * partials_from_sets (pcs/src/fields/crossfield.rs:1136, 1451). */
int lbn_partials(lbn_ctx *ctx, const uint64_t *r_hi, uint64_t *partials);

/* The GHASH->F162 switch sumcheck over the loaded table (crossfield.rs LeanSwitch 2168-2549,
 * semantics SwitchProver 563-623; driven as main.rs:948-992): begin with the 128 batch weights
 * (F162 as 3 u64), then per round msg -> [e0, einf] and fold(r); final -> opened. */
int lbn_switch_begin(lbn_ctx *ctx, const uint64_t *batch);
int lbn_switch_msg(lbn_ctx *ctx, uint64_t *msg);
int lbn_switch_fold(lbn_ctx *ctx, const uint64_t *r);
int lbn_switch_final(lbn_ctx *ctx, uint64_t *opened);

/* Row evaluation row[c] = sum_i eq(p0,i) lift(w[c*len+i]) for the 1024 columns, p0 with
 * k-17 F162 coordinates (eval.rs row_evaluate_words 426-480). row: 1024 x 3 u64. */
int lbn_row_eval(lbn_ctx *ctx, const uint64_t *p0, int n_p0, uint64_t *row);

/* The exact integer fold v_m = sum_j c_j(-X^4) W_{j,m} over the ring elements, as i32 LE
 * (fold_witness_bitsliced_into, pcs/src/fold.rs:1370; reference.rs:78-115). n_challenges = 1024,
 * out holds ring_elements*648 values. */
int lbn_fold(lbn_ctx *ctx, const lbn_challenge *challenges, int n_challenges, int32_t *out,
             size_t n_out);

/* The whole prover (main.rs prove 1692-2009; reference.rs:137-177): commit, transcript,
 * partials, switch, row, folding challenges, fold, bincode proof. The table and modulus must be
 * loaded. proof_len is the exact proof size; prof may be NULL. */
size_t lbn_proof_len(int log_vars, uint64_t q_lo, uint64_t q_hi);
int lbn_prove(lbn_ctx *ctx, const uint8_t seed[32], uint8_t *proof, size_t cap, size_t *len,
              lbn_profile *prof);

/* Many proofs at once, no counterpart in the reference (which proves one table per process,
 * solution/proof_generate/src/main.rs:2050-2090): each item is lbn_prove on its own context (all on
 * one device, distinct, modulus set; the table of `seed` generated on the device first when
 * gen_table is nonzero, else the loaded one), with the switch rounds of all items run as one
 * device sequence — one launch per prefix round, bind, control step and fold covering every
 * instance — instead of one sequence per proof; the commitments, partials, rows, challenges and
 * folds run on the items' own streams from `threads` host threads. Each proof equals the one
 * lbn_prove gives alone. 1 <= n <= 64. Returns LBN_OK or the first item's error; every item has
 * its own status (0 or a negative LBN_E_* code). */
typedef struct {
    lbn_ctx *ctx;
    uint8_t seed[32];
    int gen_table;
    uint8_t *proof;
    size_t cap;
    size_t len;          /* out */
    int status;          /* out */
    lbn_profile profile; /* out: the proof's host phases (the switch phase spans the batched rounds) */
} lbn_batch_item;
int lbn_prove_batch(lbn_batch_item *items, int n, int threads);

/* Launch-shape autotuner, no counterpart in the reference: times the matrix-core commitment (MAC and
 * finish) on the context's loaded table for each element split count (1, 2, 3, 4, 6, 8, 12, 24, up to
 * the ring elements), `reps` timed launches each after one untimed, and records the fastest for this
 * device and log_vars: later commitments of that shape on the device launch with it (a deferred
 * upload keeps 24). Launch shape only: the commitment bytes are the same for every count. The
 * modulus must commit on the matrix cores (below 2^39). */
typedef struct {
    int log_vars;
    int n;           /* split counts timed */
    int splits[8];
    double ms[8];    /* median MAC + finish time of each, milliseconds */
    int best;        /* the split count recorded */
} lbn_tune_result;
int lbn_autotune(lbn_ctx *ctx, int reps, lbn_tune_result *out);
/* Records a split count for (device, log_vars), for example one an earlier lbn_autotune chose;
 * 0 restores the default. */
int lbn_tune_set(int device, int log_vars, int mac_splits);

/* The checker (verifier/proof_verify/src/main.rs:381-389 verify, with check_proof_shape,
 * check_field_switch and check_opening, main.rs:293-379): accepts `proof` (bincode bytes) for the
 * trace of `seed` at the context's shape under the modulus the proof carries. Returns LBN_OK,
 * LBN_E_REJECT with the checker's reason in lbn_last_error, or a device error. The table of `seed`
 * is regenerated into the context (replacing a loaded table); no modulus needs to be set. */
int lbn_verify(lbn_ctx *ctx, const uint8_t seed[32], const uint8_t *proof, size_t len);

/* rq::ajtai_matrix (verifier/proof_verify/src/rq.rs:53-144), elements first .. first + count - 1 of
 * the expansion for (Q, elements), computed on device `device` (the expansion every context and the
 * checker use): out receives count x 648 coefficients below Q, each as two uint64 (low, high).
 * Indices past `elements` continue the same keyed stream family. */
int lbn_ajtai_elements(int device, uint64_t q_lo, uint64_t q_hi, size_t elements, size_t first, size_t count,
                       uint64_t *out);

/* Pinned host memory for proof buffers (no counterpart in the reference, which writes the proof
 * from an ordinary Vec): when lbn_prove's proof buffer lies in an lbn_host_alloc allocation, the
 * fold kernel writes its output into it directly and the commitment bytes are copied into it by DMA,
 * instead of both being staged and copied by the host. Any other buffer works as before. */
void *lbn_host_alloc(size_t bytes);
void lbn_host_free(void *p);

/* The 128 partials of lbn_partials (crossfield.rs:152-165; the same values) for a table already in
 * device memory (16-byte aligned, 2^log_vars / 8 bytes), without a context and without allocating:
 * no counterpart in the reference; for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc). Everything besides the
 * table lives in one file-scope device array of lbn_partials_static_scratch() bytes (136 KiB, enough
 * for every shape up to log_vars 31), held by each device the library runs on until the process
 * ends. r_hi: log_vars - 7 B128 (u64 pairs); partials: 128 B128 out (host memory). Calls on one
 * device share the array; the library serialises them (one at a time per process), so they are
 * safe to make from several threads but never run concurrently. Returns LBN_E_ARG for a shape or
 * size beyond the array's limit, LBN_E_DEVICE for a HIP failure; lbn_partials_static_error() gives
 * the calling thread's last message. */
int lbn_partials_static(int device, const void *table_device, size_t bytes, int log_vars, const uint64_t *r_hi,
                        uint64_t *partials);
size_t lbn_partials_static_scratch(void);
const char *lbn_partials_static_error(void);

/* Diagnostics (no counterpart in the reference): one plain copy of the table's bytes into a device
 * buffer of the same size, 16 bytes per lane, read once and written once: a kernel of known traffic
 * that scales a profiler's byte counters. Clears the state lbn_partials left. */
int lbn_copy_probe(lbn_ctx *ctx);

/* How the commitment is computed for the current modulus (diagnostics): moduli below 2^39 take the
 * matrix cores (an int8 product of the binary witness and the matrix's rotations in five byte
 * planes; lbn_commit_planes returns 5 and lbn_commit_primes 0), larger ones a transform over
 * lbn_commit_primes CRT primes (lbn_commit_planes returns 0). */
int lbn_commit_primes(const lbn_ctx *ctx);
int lbn_commit_planes(const lbn_ctx *ctx);

/* Kernel timing (diagnostics): when enabled, every pass records HIP events on the context's own
 * stream around its kernels; lbn_get_timing returns the milliseconds accumulated since the last
 * lbn_reset_timing. No counterpart in the reference (it prints wall-clock phases only). */
typedef struct {
    double table_ms, eq_ms, partials_ms, switch0_ms, switch_rounds_ms, row_ms, fold_ms, commit_mac_ms,
        commit_finish_ms;
    int switch_rounds;
} lbn_timing;
int lbn_set_timing(lbn_ctx *ctx, int enable);

/* Graph replay of the device-driven switch (default on): the 2 log_packed - 1 kernel launches of
 * the rounds are captured as a HIP graph on a context's first proof and replayed as one graph
 * launch by every later proof; disabled, they are launched one by one. The proof bytes are the
 * same either way. No counterpart in the reference. */
int lbn_set_graphs(lbn_ctx *ctx, int enable);
int lbn_reset_timing(lbn_ctx *ctx);
int lbn_get_timing(const lbn_ctx *ctx, lbn_timing *out);

/* ----------------------------------------------------------------------------------------------
 * Batched and continuous proving on one GPU. No counterpart in the reference, which proves one
 * table per process (solution/proof_generate/src/main.rs:2050-2090); these serve many proofs of
 * mixed table sizes at once: worker threads each run lbn_prove on a context of their own (own HIP
 * streams), contexts are cached per table size, and jobs are admitted against an HBM budget
 * (idle contexts of other sizes are evicted, least recently used first). Each proof equals the
 * one lbn_prove gives for the same seed, table and Q alone.
 * ---------------------------------------------------------------------------------------------- */

/* Device bytes one context of 2^log_vars evaluations holds under modulus Q (0 if the shape or Q
 * is invalid): the admission unit of the scheduler. */
size_t lbn_ctx_footprint(int log_vars, uint64_t q_lo, uint64_t q_hi);

typedef struct lbn_sched lbn_sched;

typedef struct {
    int log_vars;        /* 19..31 */
    uint8_t seed[32];    /* proof seed; also the table seed when table is NULL */
    const void *table;   /* optional host witness table (2^log_vars / 8 bytes); NULL: generated on
                            the device from seed (verifier/table_writer/src/main.rs:14-22) */
    uint8_t *proof;      /* caller-owned output, at least lbn_proof_len bytes */
    size_t cap;
    size_t len;          /* out: proof length */
    int status;          /* out: 1 while queued or running, then 0 or a negative LBN_E_* code */
    double queue_us;     /* out: submission to admission */
    double run_us;       /* out: admission to completion (context set-up, table, proof) */
    lbn_profile profile; /* out: the proof's host phases (lbn_prove) */
} lbn_job;

typedef struct {
    uint64_t submitted, completed, failed, contexts_created, evictions;
    uint64_t peak_bytes; /* most device bytes reserved at once */
    uint64_t used_bytes; /* reserved now */
} lbn_sched_stats;

/* workers: proofs in flight at once (1..64); hbm_budget: device bytes the contexts may hold. */
int lbn_sched_create(int device, uint64_t q_lo, uint64_t q_hi, size_t hbm_budget, int workers,
                     lbn_sched **out);
/* Queues a job (it must stay valid until finished); LBN_E_RANGE if it can never fit the budget. */
int lbn_sched_submit(lbn_sched *s, lbn_job *job);
/* Blocks until the job has finished; returns its status. */
int lbn_sched_wait(lbn_sched *s, const lbn_job *job);
/* Blocks until every submitted job has finished. */
int lbn_sched_drain(lbn_sched *s);
int lbn_sched_get_stats(lbn_sched *s, lbn_sched_stats *out);
/* Drains, stops the workers and frees every context. */
void lbn_sched_destroy(lbn_sched *s);
/* Creates idle contexts of 2^log_vars evaluations (modulus set) until the scheduler holds count
 * of that size, on the calling thread, so that later jobs of that size find a context ready (a
 * context created for a job costs milliseconds and synchronises the device). LBN_E_RANGE, with
 * none created, when the budget cannot hold them. No counterpart in the reference. */
int lbn_sched_prewarm(lbn_sched *s, int log_vars, int count);
/* Continuous batching: from now on a worker that wakes takes every queued job it can admit, up to
 * max_batch (1..64), and proves them together with lbn_prove_batch on `threads` host threads (a
 * single job runs as lbn_prove). max_batch 1 restores one proof per worker. Each proof still
 * equals the one lbn_prove gives alone. No counterpart in the reference. */
int lbn_sched_set_batch(lbn_sched *s, int max_batch, int threads);

/* ----------------------------------------------------------------------------------------------
 * Standalone commitment (the commit kernels on their own, for keys of one or several ring rows).
 * Replaces the library's column commitment sms::commit_columns_rows
 * (pcs/src/simd/sms_commit.rs:1473) and its NTT-domain form simd::commit::commit_limbs
 * (pcs/tests/commit.rs:29-42), with the result in the coefficient domain mod Q as rq.rs defines it
 * (verifier/proof_verify/src/rq.rs:264-287). A witness is `columns` columns of `elements` binary
 * ring elements, 4 u128 words (64 bytes) per element, column-major like the prover's table
 * (element m of column c at bytes (c * elements + m) * 64). Ring element m of key row r is element
 * r * elements + m of rq::ajtai_matrix's expansion for (Q, elements) (rq.rs:53-144), so row 0 is
 * ajtai_matrix(Q, elements) itself. Output: rows x columns x 648 coefficients in [0, Q), each
 * ceil(bits(Q-1)/8) bytes LE, row-major ([row][column][t]).
 * ---------------------------------------------------------------------------------------------- */
typedef struct lbn_cmt lbn_cmt;

typedef struct {
    double total_ms;     /* whole commitments (first kernel start to last kernel end) */
    double transform_ms; /* forward transforms of the witness (key of several rows only) */
    double mac_ms;       /* products with the key (rows == 1: transform and MAC fused) */
    double sum_ms;       /* sums of the partial-sum slices */
    double finish_ms;    /* inverse transforms, CRT, reduction mod Q, packing */
    int runs;
} lbn_cmt_timing;

/* Output bytes for a key of `rows` rows over `columns` columns under Q (0 if invalid). */
size_t lbn_cmt_len(int columns, int rows, uint64_t q_lo, uint64_t q_hi);
/* rows in [1, 16]; elements in [1, 65536]; columns in [16, 4096], a multiple of 16 (rows == 1)
 * or of 64 (rows > 1); Q in [34,816,965,049, 2^96). Expands and transforms the key (setup). */
int lbn_cmt_create(int device, int columns, int elements, int rows, uint64_t q_lo, uint64_t q_hi,
                   lbn_cmt **out);
void lbn_cmt_destroy(lbn_cmt *c);
const char *lbn_cmt_error(const lbn_cmt *c);
int lbn_cmt_primes(const lbn_cmt *c);
/* The witness: columns * elements * 64 bytes from host or (16-byte aligned) device memory, or the
 * table writer's XOF of a seed (verifier/table_writer/src/main.rs:14-22), its first bytes. */
int lbn_cmt_load(lbn_cmt *c, const void *host_words, size_t bytes);
int lbn_cmt_load_device(lbn_cmt *c, const void *device_words, size_t bytes);
int lbn_cmt_generate(lbn_cmt *c, const uint8_t seed[32]);
/* Queues one commitment on the context's stream (returns without waiting). */
int lbn_cmt_launch(lbn_cmt *c);
/* Waits for the queued commitments and copies the last one's packed coefficients. */
int lbn_cmt_read(lbn_cmt *c, uint8_t *out, size_t len);
/* Kernel timing with HIP events on the context's stream (diagnostics, no counterpart in the
 * reference): set_timing resets the sums; get_timing waits and returns them. */
int lbn_cmt_set_timing(lbn_cmt *c, int enable);
int lbn_cmt_get_timing(lbn_cmt *c, lbn_cmt_timing *t);

#ifdef __cplusplus
}
#endif
#endif
