// worker.h -- the worker entry points of P/worker.h over the MI355X engine.
#pragma once

#include <cstdint>
#include <string>

#include "fft_engine.h"
#include "task.h"

// P/worker.cpp:27-70.  Payload keys as in the reference: "n", "k",
// "report_step" (default 25), "k_list" (comma separated).
//   label "Proth":          one test through execute_proth_test_fft;
//                           returns "prime" or "PRP" (the reference's words).
//   label "PROTH_FF_BATCH": every k of k_list with the same n.  Instead of
//                           one test after another, the candidates go to one
//                           device pool (pl_pool_*) and advance together;
//                           each prime prints "[BATCH-RESULT] PRIME FOUND:
//                           k=.." and is appended to primes.txt.  Returns
//                           "BATCH".
//   anything else:          "fail".
// limbs is used as in the reference (the "Proth" test's final residue).
std::string run_task_with_ctx(const Task& t, FFTContext& ctx, uint64_t* limbs);
