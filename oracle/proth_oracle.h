/*
 * proth_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, scalar) of ProthLogic 4.2's Proth squaring path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline; the
 * product (prothlogic_amd) never links or calls it.
 *
 * Reference files restated (paths relative to the reference tree,
 * P/ = prothlogic4.2/):
 *   P/fft_engine.cpp:18-136, 234-494, 499-820, 825-1073  (engine)
 *   P/llr_fft_pack.cpp:12-41                             (pack)
 *   P/proth_logic.cpp:27-50, 336-355, 384-580            (gate, verdict, driver)
 *   P/dcore.cpp:20-81                                    (FFT-length autotune)
 *   P/checkpoint_manager.cpp:8-45                        (checkpoint format)
 *
 * Pinning of each function is listed in the header of proth_oracle.c.
 */
#ifndef PROTH_ORACLE_H
#define PROTH_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

/* P/fft_engine.cpp:18-73 (fft_init) / :75-80 (fft_destroy). */
or_ctx* or_fft_init(size_t fft_len);
void or_fft_destroy(or_ctx* c);
size_t or_fft_len(const or_ctx* c);
double or_last_max_diff(const or_ctx* c);

/* Natural-order complex DFT of length L in place on (re, im) -- the
 * dispatch of P/fft_engine.cpp:994-1014: Bailey four-step for even log2 L,
 * mixed radix-8 Stockham for odd log2 L.  Forward uses e^{-2 pi i/L};
 * inverse is unnormalised. */
void or_fft_inplace(or_ctx* c, double* re, double* im, int inverse);

/* P/llr_fft_pack.cpp:12-41.  `in` must be readable for
 * round_up(4*n_limbs, 8) 16-bit chunks (the reference's vector loads). */
void or_pack(or_ctx* c, const uint64_t* in, size_t n_limbs);

/* P/fft_engine.cpp:825-984: round, max-err, base-2^16 carry, fold mod
 * k*2^n+1.  Writes out[0..n_limbs). */
void or_unpack(or_ctx* c, uint64_t* out, size_t n_limbs, unsigned n_bits, uint64_t k);

/* P/fft_engine.cpp:986-1073: v <- v^2 mod (k*2^n+1) in place. */
void or_square(or_ctx* c, uint64_t* v, size_t n_limbs, unsigned n_bits, uint64_t k);

/* P/proth_logic.cpp:41-50 + 405-408: 1 if some p<=67 divides k*2^n+1. */
int or_gate_rejects(uint64_t k, unsigned n);
/* P/proth_logic.cpp:423-435 (mpz_jacobi replaced by a small-modulus
 * Jacobi symbol of b over N = k*2^n+1). Returns 0 when no base qualifies. */
uint64_t or_choose_base(uint64_t k, unsigned n);
int or_jacobi_small(uint64_t b, uint64_t k, unsigned n);

/* P/dcore.cpp:20-81.  fit_fix=0 restates the reference rule as written;
 * fit_fix=1 additionally doubles L until 16*L >= n + 2*bitlen(k) + 16, the
 * condition under which the reference's limb layout holds x^2 mod N. */
size_t or_auto_tune_fft_len(unsigned n, uint64_t k, int fit_fix, double* err_out);
int or_fits(size_t fft_len, unsigned n, uint64_t k);

/* x <- a^k mod N by left-to-right binary powering with or_square and an
 * exact multiply-by-small (stands in for mpz_powm at
 * P/proth_logic.cpp:438-440).  limbs has fft_len/4 entries. */
void or_seed(or_ctx* c, uint64_t* limbs, uint64_t a, uint64_t k, unsigned n);

/* P/proth_logic.cpp:336-355. */
int or_is_nm1(const uint64_t* limbs, size_t n_limbs, unsigned n, uint64_t k);

/* P/checkpoint_manager.cpp:8-45 (format {u64 i; size_t n_limbs} + limbs). */
int or_save_checkpoint(uint64_t i, size_t n_limbs, const uint64_t* data, const char* filename);
int or_load_checkpoint(uint64_t* out_i, size_t n_limbs, uint64_t* data, const char* filename);

/* P/proth_logic.cpp:384-580 without console/primes.txt output.
 * limbs: caller buffer of >= max(fft_len, 2n/64+64) u64 (P/dcore.cpp:143).
 * checkpoint_dir: NULL disables checkpoint I/O, else files are
 *   <dir>/check_k{k}_n{n}.bin exactly as the reference names them.
 * stop_after: if nonzero, stop after that many loop iterations (simulated
 *   kill; the checkpoint is left in place).  Returns 1 prime, 0 composite,
 *   -1 stopped.  *base_out gets a (0 if gated), *max_err_out the max
 *   round-off seen, *iters_out the number of squarings executed. */
int or_execute_proth_test(uint64_t k, unsigned n, uint64_t* limbs, or_ctx* c,
                          const char* checkpoint_dir, uint64_t stop_after,
                          uint64_t* base_out, double* max_err_out, uint64_t* iters_out);

/* Exact helpers on little-endian u64 limbs used by the tests. */
void or_mod_N(uint64_t* v, size_t n_limbs, unsigned n, uint64_t k); /* v < 2^(64*n_limbs) -> v mod N */

#ifdef __cplusplus
}
#endif
#endif
