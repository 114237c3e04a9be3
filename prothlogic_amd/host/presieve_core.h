// presieve_core.h -- the candidate sieve of P/presieve_core.h:8-18 over the
// device sieve (pl_sieve_range / pl_sieve_list).
#pragma once

#include <cstdint>
#include <vector>

// Odd k in [k_start, k_end_inclusive] such that no prime 3 <= p <= pmax
// divides k*2^n_pow2 + 1, ascending.  Throws std::runtime_error when the
// device fails.
std::vector<uint64_t> execute_full_streaming_sieve(uint64_t k_start, uint64_t k_end_inclusive, unsigned n_pow2,
                                                   uint64_t pmax);

// The k of input_ks (order kept) that no prime in [pstart, pmax] removes.
std::vector<uint64_t> execute_sparse_streaming_sieve(const std::vector<uint64_t>& input_ks, unsigned n_pow2,
                                                     uint64_t pstart, uint64_t pmax);
