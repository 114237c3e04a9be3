// presieve_core.cpp -- execute_full_streaming_sieve / execute_sparse_streaming_sieve
// over the C ABI (see presieve_core.h).
#include "presieve_core.h"

#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>

#include "prothcore.h"

namespace {

int device_from_env()
{
    const char* d = std::getenv("PL_DEVICE");
    return d ? std::atoi(d) : 0;
}

}  // namespace

std::vector<uint64_t> execute_full_streaming_sieve(uint64_t k_start, uint64_t k_end_inclusive, unsigned n_pow2,
                                                   uint64_t pmax)
{
    if (k_end_inclusive < k_start) return {};
    const size_t cap = (size_t)((k_end_inclusive - k_start) / 2 + 1);
    std::vector<uint64_t> ks(cap);
    size_t got = 0;
    if (pl_sieve_range(k_start, k_end_inclusive, n_pow2, pmax, device_from_env(), ks.data(), cap, &got) != PL_OK)
        throw std::runtime_error(std::string("execute_full_streaming_sieve: ") + pl_sieve_last_error());
    ks.resize(got);
    std::cout << "[DONE] Candidates after sieve " << got << " of " << (k_end_inclusive - k_start + 1) << std::endl;
    return ks;
}

std::vector<uint64_t> execute_sparse_streaming_sieve(const std::vector<uint64_t>& input_ks, unsigned n_pow2,
                                                     uint64_t pstart, uint64_t pmax)
{
    if (input_ks.empty()) return {};
    std::vector<uint8_t> keep(input_ks.size(), 1);
    if (pl_sieve_list(input_ks.data(), input_ks.size(), n_pow2, pstart, pmax, device_from_env(), keep.data()) != PL_OK)
        throw std::runtime_error(std::string("execute_sparse_streaming_sieve: ") + pl_sieve_last_error());
    std::vector<uint64_t> out;
    for (size_t i = 0; i < input_ks.size(); ++i)
        if (keep[i]) out.push_back(input_ks[i]);
    std::cout << "[SPARSE DONE] Candidates after sieve: " << out.size() << " of " << input_ks.size() << std::endl;
    return out;
}
