// fft_types.h -- FFTContext and FFTBackend as ProthLogic 4.2's callers use
// them (P/fft_types.h:8-60), bound to the MI355X squaring engine of
// include/prothcore.h.  The fields callers read keep their names and
// meaning: fft_len (limb buffers are sized from it, fft_len/4 limbs at
// P/proth_logic.cpp:442), last_max_diff (round-off of the last squaring
// call), num_workers, current_i and the prof_* counters.  The CPU engine's
// work buffers and tables do not exist here: the residue lives in HBM.
#pragma once

#include <cstddef>
#include <cstdint>

struct pl_ctx;

// INTERNAL is kept so existing calls (fft_init(ctx, FFTBackend::INTERNAL,
// L), P/dcore.cpp:136) compile; in this build it selects the one engine
// there is, the HIP device engine, exactly like HIP.
enum class FFTBackend { INTERNAL, HIP };

struct FFTContext {
    FFTBackend backend = FFTBackend::HIP;
    size_t fft_len = 0;
    int num_workers = 1;
    double last_max_diff = 0.0;
    uint64_t current_i = 0;

    double prof_pack = 0.0;
    double prof_fwd = 0.0;
    double prof_square = 0.0;
    double prof_inv = 0.0;
    double prof_unpack = 0.0;
    uint64_t prof_calls = 0;

    // device binding: created on the first squaring for a (k, n) pair
    int device = 0;
    pl_ctx* dev = nullptr;
    uint64_t dev_k = 0;
    unsigned dev_n = 0;
};
