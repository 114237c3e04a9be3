// autotune.h -- auto_tune_fft_len of P/dcore.cpp:20-81 over the device
// autotuner (pl_autotune_fft_len).
#pragma once

#include <cstddef>
#include <cstdint>

// Same probe as the reference (limbs all ones, limbs[0] = k|1, 100
// squarings, accepted while max round-off <= 0.30), run on the device at
// the rule's length and its two power-of-two neighbours; returns the fastest
// length that passes.  The value is the device transform length: the
// number of residue digits, usable as fft_len for fft_init and for sizing
// limb buffers the way the reference's dispatch does (max(L, 2n/64 + 64)
// limbs).  Returns 0 when no device is visible or no length passes.
size_t auto_tune_fft_len(unsigned actual_n, uint64_t sample_k);
