// proth_logic.h -- the Proth test entry point of P/proth_logic.h:12 over the
// MI355X engine.
#pragma once

#include <cstdint>

#include "fft_engine.h"

// P/proth_logic.cpp:384-580.  Gate, base, seed a^k, n-1 squarings and the
// verdict x == N-1, with the residue resident on the device throughout.
// Returns true for a prime (false also for candidates the small-prime gate
// rejects, as the reference does).  Keeps the reference's call-site
// behaviour:
//   * checkpoints every 10000 squarings in check_k<k>_n<n>.bin (current
//     directory), resumed when present (reference-layout files too) and
//     removed at the end;
//   * "[STEP] k=.. -> p% | Max Err: .." each time the percentage reaches a
//     new multiple of report_step (report_step <= 0: no progress lines);
//   * the final canonical residue left in limbs (limbs[0] is the Res64),
//     zero-filled up to max(ctx.fft_len/4, (n+127)/64) limbs, which the
//     reference's callers allocate (P/dcore.cpp:143 sizes the buffer
//     max(L, 2n/64 + 64));
//   * a prime is appended to primes.txt as "k * 2^n + 1 (a=A)".
// ctx.last_max_diff gets the largest round-off of the whole test.  Throws
// std::runtime_error when the device fails or the round-off passes the
// guard (the reference has no error path).
bool execute_proth_test_fft(uint64_t k, unsigned n, uint64_t* limbs, FFTContext& ctx, int report_step);
