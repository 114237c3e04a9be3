// fft_engine.h -- the squaring API of P/fft_engine.h:10-22 over the MI355X
// engine.  fft_init / fft_destroy / fft_square_karatsuba keep their
// signatures and semantics; the CPU engine's internals declared there
// (Stockham/Bailey passes, pack/unpack) have no callers outside the CPU
// engine and are not provided.
#pragma once

#include <cstddef>
#include <cstdint>

#include "fft_types.h"

// P/fft_engine.h:10.  Records fft_len and selects the device (environment
// PL_DEVICE, default 0).  Returns false when no HIP device is visible or
// fft_len is 0 (the reference returns false when its allocation fails).
bool fft_init(FFTContext& ctx, FFTBackend backend, size_t fft_len);

// P/fft_engine.h:11.  Releases the device context.
void fft_destroy(FFTContext& ctx);

// P/fft_engine.h:22 (semantics of P/fft_engine.cpp:986-1073): v[0..n_limbs)
// <- v^2 mod (k_val*2^n_bits + 1), canonical; ctx.last_max_diff <- that
// squaring's round-off.  v is read as a value (any size, reduced mod N);
// limbs of the result above n_limbs are dropped, as the reference's unpack
// writes n_limbs limbs.  Throws std::runtime_error when the device fails
// (the reference has no error path; a silent wrong residue is worse).
void fft_square_karatsuba(FFTContext& ctx, uint64_t* v, size_t n_limbs, unsigned n_bits, uint64_t k_val);
