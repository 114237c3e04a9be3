/*
 * prothcore.h -- C ABI of the MI355X (gfx950) Proth squaring engine.
 *
 * Drop-in boundary for ProthLogic 4.2's squaring path.  Plain C: no C++
 * types, no exceptions across the boundary, plain pointers and sizes.  Every
 * function returns PL_OK (0) or a negative PL_ERR_* code; pl_last_error()
 * gives a human-readable reason for the calling thread's last failure.
 *
 * The residue of a context stays on the device between calls; the host
 * sees canonical residues only through pl_get_residue / pl_res64 /
 * checkpoints.  A context owns one HIP stream (default priority) and is not
 * safe to use from two threads at once; distinct contexts are independent
 * and may run concurrently from several threads on one device (one test per
 * worker thread, as P/dcore.cpp:125-192 dispatches): at most one of them at
 * a time runs the one-launch outer pass (whose workgroups wait on each
 * other), the others take the two-launch pass meanwhile, and a chunk whose
 * carry hand-off still gives up (kernels of another process holding the
 * CUs) is recomputed from its start on the two-launch pass.
 *
 * Each declaration names the reference interface it replaces
 * (P/ = prothlogic4.2/ in the reference tree).
 */
#ifndef PROTHCORE_H
#define PROTHCORE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PL_OK 0
#define PL_ERR_ARG (-1)      /* bad argument (sizes, k even, n out of range) */
#define PL_ERR_HIP (-2)      /* HIP runtime failure or no usable GPU        */
#define PL_ERR_ROUNDOFF (-3) /* FFT round-off above the guard: result void  */
#define PL_ERR_STATE (-4)    /* device carry did not settle (internal bug)  */
#define PL_ERR_IO (-5)       /* checkpoint file could not be read/written   */

#define PL_ABI_VERSION 6

typedef struct pl_ctx pl_ctx;

typedef struct pl_info {
    uint64_t k;
    uint32_t n;
    uint32_t fft_len;      /* complex points L (= digits of the residue)   */
    uint32_t rows;         /* R: strided DFT length                         */
    uint32_t cols;         /* C: contiguous DFT length                      */
    double bits_per_digit; /* n / L                                         */
    uint64_t device_bytes; /* HBM held by the context                       */
    uint32_t passes;       /* kernels per steady-state squaring: 0 when the
                              steady state runs as one resident launch per
                              chunk of squarings (PL_RESIDENT=0 turns it
                              off), 2 with the one-launch outer pass, 3 with
                              it in two launches (PL_CARRY_TURN=0), 5
                              unfused                                        */
    uint32_t handoff_fallbacks; /* chunks recomputed on the two-launch pass
                              after a carry hand-off wait gave up            */
    uint32_t row_radix;    /* largest radix of the steady-state row pass    */
    uint32_t col_radix;    /* largest radix of the steady-state outer pass  */
    uint32_t run;          /* digits per carry run of the outer pass (8 or
                              16; 0 unfused)                                 */
} pl_info;

/* Thread-local description of the last error (never NULL). */
const char* pl_last_error(void);
int pl_abi_version(void);
/* Number of visible HIP devices (0 when none; not an error). */
int pl_device_count(int* count);

/* FFT-length rule (replaces auto_tune_fft_len, P/dcore.cpp:20-81): the
 * smallest power-of-two L whose n/L keeps FP64 round-off within the guard
 * and whose limb layout always holds the residue (no "no-fit" band). */
int pl_plan_fft_len(uint32_t n, uint64_t k, uint32_t* fft_len);

/* Measured FFT length (replaces auto_tune_fft_len, P/dcore.cpp:20-81, and
 * keeps its probe: limbs all ones with limbs[0] = k|1, 100 squarings, max
 * round-off against `err_guard`; <= 0 selects the reference's 0.30).  Tries
 * the rule's length, the next shorter power of two the round-off limits
 * allow, and the next longer one; times 256 squarings at each length that
 * passes and picks the fastest.  *fft_len gets the choice, *max_err (nullable)
 * its probe round-off, *us_per_squaring (nullable) its measured time. */
int pl_autotune_fft_len(uint32_t n, uint64_t k, int device, double err_guard, uint32_t* fft_len, double* max_err,
                        double* us_per_squaring);

/* Launch-plan autotuner (no counterpart in the reference, whose one length
 * test is P/dcore.cpp:20-81).  For a length the engine can run several
 * shapes L = R*C, row and column radices 8 or 16, carry runs of 8 or 16
 * digits and the outer pass as one launch or two.  A tuning run times them
 * on the device (coordinate search from the built-in rule: shape, row radix,
 * column radix, runs, outer-pass form), each from the same random residues
 * for the same number of whole 256-squaring chunks, and keeps the fastest
 * whose residue after those squarings equals the rule plan's and whose
 * round-off stays within 0.30.  The choice goes into a process-wide table
 * that contexts (pl_ctx_create) and pool groups created later take, keyed by
 * device, length, carry layout, slot-count class (1, 2-8, 9-64, more) and
 * outer-pass constraint.  pl_autotune_fft_len also tunes the length it
 * picks.  Each trial is reported (trials may be NULL; at most max_trials are
 * written, *n_trials gets the count); *best (nullable) is the choice. */
typedef struct pl_tune_trial {
    uint32_t rows, cols;     /* R x C                                          */
    uint32_t row_radix;      /* 8 or 16                                        */
    uint32_t col_radix;      /* 8 or 16                                        */
    uint32_t run;            /* digits per carry run (8, 16; 0 unfused)        */
    uint32_t one_launch;     /* outer pass as one launch                       */
    double us_per_squaring;  /* device time per squaring of every slot         */
    double max_err;          /* round-off of the timed squarings               */
    int32_t exact;           /* residue equal to the rule plan's               */
    int32_t chosen;          /* the trial the table now holds                  */
} pl_tune_trial;
/* Tunes the plan of a single context for (k, n) (pl_ctx_create's, slots 1). */
int pl_autotune_plan(int device, uint64_t k, uint32_t n, pl_tune_trial* trials, uint32_t max_trials,
                     uint32_t* n_trials, pl_tune_trial* best);
/* Puts `choice` (rows, radices, run, one_launch) into the table for the
 * plan a single context for (k, n) would take; PL_ERR_ARG when the kernels
 * have no such plan for that length or carry layout. */
int pl_tune_pin(int device, uint64_t k, uint32_t n, const pl_tune_trial* choice);
/* Empties the table (every later plan takes the built-in rule). */
int pl_tune_clear(void);
/* Writes / merges the table as text (one entry a line). */
int pl_tune_save(const char* path);
int pl_tune_load(const char* path);

/* Context create/destroy (replaces fft_init / fft_destroy,
 * P/fft_engine.h:10-11).  fft_len = 0 picks pl_plan_fft_len's length.
 * The residue starts at 0.  Environment (read at creation): PL_CARRY_TURN=0
 * runs the outer pass as two launches (3 kernels per squaring instead of 2);
 * PL_FUSED=0 runs it unfused (5 kernels).  pl_info.passes reports which
 * form the context prefers.  PL_NO_GRAPHS=1 (contexts and pool groups)
 * launches whole 256-squaring chunks directly instead of replaying their
 * captured graphs: the same kernels, for dispatch-tracing profilers. */
int pl_ctx_create(uint64_t k, uint32_t n, uint32_t fft_len, int device, pl_ctx** out);
void pl_ctx_destroy(pl_ctx* ctx);
int pl_ctx_info(const pl_ctx* ctx, pl_info* out);

/* Residue in/out as little-endian u64 limbs of the canonical value in
 * [0, N) (the limb layout of P/proth_logic.cpp:442-446).  set reduces any
 * value mod N; get zero-fills limbs above the value and fails with
 * PL_ERR_ARG when n_limbs cannot hold it. */
int pl_set_residue(pl_ctx* ctx, const uint64_t* limbs, size_t n_limbs);
int pl_get_residue(pl_ctx* ctx, uint64_t* limbs, size_t n_limbs);

/* x <- x^2 mod N, `iters` times, on the device (the loop body of
 * P/proth_logic.cpp:485-549).  *max_err (nullable) gets the largest
 * |s - rint(s)| over all iterations; err_trace (nullable, `iters` entries)
 * gets each iteration's maximum.  Fails with PL_ERR_ROUNDOFF when an
 * iteration exceeds `err_guard` (<= 0 selects the default 0.40); the
 * residue is then unreliable. */
int pl_square(pl_ctx* ctx, uint64_t iters, double err_guard, double* max_err, double* err_trace);

/* x <- x*a mod N for a small a (< 2^32): the seed's powering step. */
int pl_mul_small(pl_ctx* ctx, uint64_t a);

/* Low 64 bits of the canonical residue (the build's Res64), and the
 * verdict test x == N-1 (replaces is_nm1_limbs, P/proth_logic.cpp:336-355). */
int pl_res64(pl_ctx* ctx, uint64_t* res64);
int pl_is_n_minus_1(pl_ctx* ctx, int* is_nm1);

/* Per-call compatibility shim with the exact semantics of
 * fft_square_karatsuba(ctx, v, n_limbs, n_bits, k) (P/fft_engine.h:22,
 * P/fft_engine.cpp:986-1073): v[0..n_limbs) <- v^2 mod (k*2^n_bits+1), and
 * *last_max_diff (nullable) <- that squaring's max round-off
 * (FFTContext::last_max_diff).  n_bits and k must match the context. */
int pl_fft_square_karatsuba(pl_ctx* ctx, uint64_t* v, size_t n_limbs, unsigned n_bits, uint64_t k,
                            double* last_max_diff);

/* No counterpart in the reference: pl_fft_square_karatsuba for callers
 * that cannot allocate device memory (code under graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  v[0..n_limbs) (host LE
 * u64 limbs) <- v^2 mod (k*2^n_bits+1), *last_max_diff (nullable) <- the
 * squaring's max round-off, with no context: every device buffer the call
 * uses lies in one static array per GPU (6 MiB), so the call makes no
 * device allocation (hipMalloc, hipFree, stream or graph creation; host
 * conversion buffers are ordinary heap memory) and runs on the legacy
 * stream.  n_bits must need a transform length of at most 2^18 (about
 * 4.5 million bits): longer ones fail with PL_ERR_ARG.  Calls on one device
 * are serialised within a process and share its array, which keeps the
 * last call's contents (each call clears what it reads).  Errors are
 * described by pl_last_error(). */
int pl_fft_square_karatsuba_noalloc(int device, uint64_t* v, size_t n_limbs, unsigned n_bits, uint64_t k,
                                    double* last_max_diff);

/* Diagnostics (no counterpart in the reference): average duration in ms of
 * each kernel of one squaring, timed with HIP events on the context's own
 * stream between consecutive kernels of `reps` real squarings run in their
 * steady-state order, less the events' own cost (the excess of their sum
 * over the same loop timed end to end, in equal parts per steady kernel),
 * into ms[0..7]: ms[0] forward column pass,
 * ms[1] row pass (forward DFT, pointwise square, inverse DFT), ms[2] inverse
 * column pass, ms[3] flat carry/fold, ms[4] block fix-up, ms[5] inverse
 * column DFT + carry/fold, ms[6] carry-in + forward column DFT, ms[7] the
 * one-launch outer pass (inverse column DFT, carry/fold, carry hand-off,
 * forward column DFT).  Kernels the context does not run read 0.  The
 * residue is preserved.
 * *traffic_bytes (nullable) gets the algorithmic HBM bytes one launch of
 * each full-array kernel moves (read + write of the L complex points). */
int pl_kernel_times(pl_ctx* ctx, int reps, double* ms, uint64_t* traffic_bytes);

/* Diagnostics (no counterpart in the reference): device time of the resident
 * launches (pl_info.passes == 0) that the context's pl_square calls ran, as
 * they ran: HIP events around each launch on the context's stream, summed
 * over completed chunks.  *ms (nullable) gets the total, *launches and
 * *squarings (nullable) how many launches and how many squarings they
 * carried; reset != 0 clears the sums after reading them. */
int pl_loop_time(pl_ctx* ctx, int reset, double* ms, uint64_t* launches, uint64_t* squarings);

/* Diagnostics (no counterpart in the reference): *ms <- average duration of
 * a plain device-to-device copy of `bytes` bytes (reads and writes `bytes`
 * each) over `reps` launches on `device`.  The measured streaming rate of
 * the card, and the known-size kernel that calibrates byte counters. */
int pl_copy_probe(int device, uint64_t bytes, int reps, double* ms);

/* Small-prime gate (P/proth_logic.cpp:41-50, 405-408): 1 when some odd
 * prime p <= 67 divides k*2^n+1. */
int pl_gate_rejects(uint64_t k, uint32_t n);
/* Base rule (P/proth_logic.cpp:423-435): first b of {2,3,5,...,37}, then
 * 2..99, with Jacobi(b, N) = -1; 0 when none qualifies. */
uint64_t pl_choose_base(uint64_t k, uint32_t n);

/* Progress report of pl_proth_test: `done` squarings of `total` (n-1)
 * applied, max round-off so far. */
typedef void (*pl_progress_fn)(void* user, uint64_t done, uint64_t total, double max_err);

typedef struct pl_test_args {
    uint64_t k;
    uint32_t n;
    uint32_t fft_len;           /* 0: pl_plan_fft_len                           */
    int device;
    const char* checkpoint_path;/* NULL: no checkpoints; else this file         */
    uint64_t checkpoint_every;  /* squarings between checkpoints (0: 10000)     */
    uint64_t stop_after;        /* nonzero: stop after that many squarings      */
    double err_guard;           /* <= 0: 0.40                                   */
    int report_step;            /* > 0: call `progress` each time the percentage
                                   (i+1)*100/(n-1) reaches a new multiple of it
                                   (the report rule of P/proth_logic.cpp:541-548) */
    pl_progress_fn progress;    /* nullable                                     */
    void* progress_user;
    uint64_t* final_limbs;      /* nullable: the final canonical residue as LE
                                   u64 limbs, zero-filled (the reference leaves
                                   it in `limbs`, P/proth_logic.cpp:550-577)    */
    size_t final_n_limbs;
    uint32_t gerbicz_block;     /* > 0: Gerbicz check with blocks of this many
                                   squarings (planned fault tolerance of the
                                   reference, README.md:19; no counterpart in
                                   its code).  0: off                          */
    uint32_t gerbicz_blocks;    /* blocks per verified epoch (0: 64)            */
    uint64_t inject_error_at;   /* tests only: add 1 to the residue once, after
                                   the first block that passes this squaring
                                   count (0: never); the check must catch it   */
} pl_test_args;

typedef struct pl_test_result {
    int verdict;        /* 1 prime, 0 composite, -1 stopped early (stop_after) */
    int gated;          /* 1 when the small-prime gate rejected the candidate  */
    uint64_t base;      /* a                                                   */
    uint64_t res64;     /* low 64 bits of a^(k*2^iters_done) mod N            */
    uint64_t iters_done;/* squarings applied after the seed (n-1 when done)    */
    uint64_t resumed_at;/* first squaring index run in this call               */
    double max_err;     /* max round-off over the run                          */
    uint32_t fft_len;
    uint32_t gerbicz_checks;    /* epochs verified or rejected                 */
    uint32_t gerbicz_rollbacks; /* epochs rejected and recomputed              */
} pl_test_result;

/* Full Proth test a^((N-1)/2) mod N with the residue resident on the device
 * (replaces execute_proth_test_fft, P/proth_logic.h:12 /
 * P/proth_logic.cpp:384-580, and its checkpoint call sites :449-451,
 * :536-538, :569, :576).  The checkpoint stores the next squaring index, so
 * a resumed run applies exactly n-1 squarings in total.
 *
 * With gerbicz_block = B > 0 the squarings run in epochs of M blocks: after
 * each block the running product d of the block-end residues u_0..u_j is
 * updated on the device (d <- d*u_j), and at the epoch's end
 * d_(M-1)^(2^B) * u_0 must equal d_M (u_j^(2^B) = u_(j+1) for an error-free
 * chain).  A verified epoch end is the only state checkpointed and the
 * rollback point: a failed check recomputes the epoch from there (three
 * failures in a row end the test with PL_ERR_STATE).  Cost: B squarings and
 * M+1 products per epoch of M*B squarings. */
int pl_proth_test(const pl_test_args* args, pl_test_result* out);

/* Checkpoint files.  Format v2 (written): 64-byte header
 *   {char magic[8]="PLCKPT02"; u64 k; u32 n; u32 base; u64 next_iter;
 *    u64 n_limbs; u64 checksum(FNV-1a over limbs); u64 reserved}
 * followed by n_limbs LE u64 limbs of the canonical residue.
 * pl_load_checkpoint also reads the reference's format
 * ({u64 i; size_t n_limbs} + limbs, P/checkpoint_manager.h:10-13) and maps
 * its "last applied index" i to next_iter = i+1.  Replaces
 * save_checkpoint / load_checkpoint (P/checkpoint_manager.h:15-16). */
int pl_save_checkpoint(const char* path, uint64_t k, uint32_t n, uint64_t base, uint64_t next_iter,
                       const uint64_t* limbs, size_t n_limbs);
int pl_load_checkpoint(const char* path, uint64_t k, uint32_t n, uint64_t* base, uint64_t* next_iter,
                       uint64_t* limbs, size_t n_limbs);

/* ------------------------------------------------------------------------
 * Candidate sieve (the step before the test).  A k is removed when some
 * prime 3 <= p <= pmax divides k*2^n+1 (k == -2^(-n) mod p).  Errors are
 * described by pl_sieve_last_error().
 * ---------------------------------------------------------------------- */

/* Replaces execute_full_streaming_sieve (P/presieve_core.h:8-12,
 * P/presieve_core.cpp:15-88, over init_gpu_sieve .. finalize_gpu_sieve,
 * P/gpu_engine.h:13-30): the odd k in [k_start, k_end] that survive, in
 * ascending order.  Writes min(cap, survivors) of them to out_ks and sets
 * *n_out to the number of survivors (call with cap = 0 to count).  The range
 * is limited to 2^36 values and pmax to 2^62. */
int pl_sieve_range(uint64_t k_start, uint64_t k_end, uint32_t n, uint64_t pmax, int device, uint64_t* out_ks,
                   size_t cap, size_t* n_out);

/* Replaces execute_sparse_streaming_sieve (P/presieve_core.h:14-18,
 * P/presieve_core.cpp:93-167): keep[i] <- 1 when ks[i] survives every prime
 * in [pstart, pmax], else 0.  The reference compares k itself, not k mod p,
 * with the target (P/gpu_engine.cu:397), so it removes almost nothing; this
 * is the test the reference intends. */
int pl_sieve_list(const uint64_t* ks, size_t count, uint32_t n, uint64_t pstart, uint64_t pmax, int device,
                  uint8_t* keep);
const char* pl_sieve_last_error(void);

/* ------------------------------------------------------------------------
 * Continuous batching.  Replaces the reference's per-thread dispatch of one
 * test per worker (dispatch, P/dcore.cpp:84-193, calling run_task_with_ctx,
 * P/worker.cpp:27-70, for each k of a PROTH_FF_BATCH task :51-67) with one
 * GPU-resident pool: candidates of uneven n are grouped by transform
 * length, every group advances all its tests with one launch per kernel kind
 * per squaring step (graph-captured), and finished tests free their slots
 * for queued ones.  Tests are admitted while a slot of their length is free
 * and a new length group fits the HBM budget.  A pool is not thread-safe.
 * ---------------------------------------------------------------------- */
typedef struct pl_pool pl_pool;

typedef struct pl_pool_result {
    uint64_t tag;       /* the caller's id from pl_pool_submit                 */
    uint64_t k;
    uint32_t n;
    uint32_t fft_len;
    int verdict;        /* 1 prime, 0 composite, -1 stopped early (stop_after),
                           -2 void (see error)                                 */
    int gated;          /* 1: rejected by the small-prime gate, no squarings   */
    uint64_t base;      /* a                                                   */
    uint64_t res64;     /* low 64 bits of a^(k*2^s) mod N, s = squarings after
                           the seed (n-1 for a full test)                      */
    uint64_t squarings; /* device squarings incl. the seed powering            */
    double max_err;     /* max round-off over the test                         */
    double t_submit;    /* seconds since the pool was created                  */
    double t_admit;
    double t_done;
    int error;          /* PL_OK; else the result is void (verdict -2, no
                           Res64): PL_ERR_ROUNDOFF when the round-off passed
                           0.40, PL_ERR_STATE when the device carry did not
                           settle in two restarts of the test from its seed  */
} pl_pool_result;

/* hbm_budget_bytes = 0: 64 GiB; max_slots_per_length = 0: 256;
 * k_max = 0: 2^32-1.  Pool groups overlap on their own streams and always
 * take the two-launch outer pass (no workgroup waits on another).  Each
 * length keeps up to two groups: carry runs of 8 digits for tests whose own
 * digit size lets 8 digits absorb the carries of any k up to k_max, and runs
 * of 16 for the rest (sized for the narrowest digits the length admits).  A
 * group takes the tuned plan for its key when pl_pool_autotune (or a table
 * loaded with pl_tune_load) recorded one before the group was created. */
int pl_pool_create(int device, uint64_t hbm_budget_bytes, uint32_t max_slots_per_length, uint64_t k_max,
                   pl_pool** out);
void pl_pool_destroy(pl_pool* pool);
/* Queue one candidate.  stop_after = 0 runs the full Proth test; otherwise
 * the test stops after that many squarings past the seed (verdict -1). */
int pl_pool_submit(pl_pool* pool, uint64_t k, uint32_t n, uint64_t tag, uint64_t stop_after);
/* Admit queued candidates, advance every active test by `steps` squarings
 * (fewer for tests that finish), retire finished tests and copy up to `cap`
 * results into out (*n_out of them; the rest stay for the next call).  A
 * length group whose device carry did not settle has every test it holds
 * restarted from its seed (or reported void, pl_pool_result.error). */
int pl_pool_run(pl_pool* pool, uint32_t steps, pl_pool_result* out, size_t cap, size_t* n_out);
int pl_pool_status(pl_pool* pool, size_t* queued, size_t* active, size_t* finished_unread, uint64_t* hbm_bytes);
/* Tunes the plan of the length group a candidate (k, n) would join, with
 * the slot count that group would get from this pool's budget (see
 * pl_autotune_plan).  A key the table already holds is not timed again
 * (*n_trials = 0, *best = the recorded choice), so it can be called for
 * every candidate of a batch.  A group that already exists keeps its plan. */
int pl_pool_autotune(pl_pool* pool, uint64_t k, uint32_t n, pl_tune_trial* trials, uint32_t max_trials,
                     uint32_t* n_trials, pl_tune_trial* best);
const char* pl_pool_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* PROTHCORE_H */
