// worker.cpp -- run_task_with_ctx over execute_proth_test_fft and the device
// pool (see worker.h).
#include "worker.h"

#include <fstream>
#include <iostream>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "proth_logic.h"
#include "prothcore.h"

namespace {

std::mutex g_batch_mutex;

std::vector<uint64_t> k_list(const std::string& s)
{
    std::vector<uint64_t> out;
    std::istringstream in(s);
    std::string tok;
    while (std::getline(in, tok, ','))
        if (!tok.empty()) out.push_back(std::stoull(tok));
    return out;
}

std::string batch(const std::vector<uint64_t>& ks, unsigned n, int device)
{
    pl_pool* pool = nullptr;
    if (pl_pool_create(device, 0, 0, 0, &pool) != PL_OK)
        throw std::runtime_error(std::string("PROTH_FF_BATCH: ") + pl_pool_last_error());
    std::vector<pl_pool_result> res(256);
    size_t pending = 0;
    try {
        for (size_t i = 0; i < ks.size(); ++i) {
            if (pl_pool_submit(pool, ks[i], n, i, 0) != PL_OK)
                throw std::runtime_error(std::string("PROTH_FF_BATCH: ") + pl_pool_last_error());
            ++pending;
        }
        while (pending > 0) {
            size_t got = 0;
            if (pl_pool_run(pool, 64, res.data(), res.size(), &got) != PL_OK)
                throw std::runtime_error(std::string("PROTH_FF_BATCH: ") + pl_pool_last_error());
            for (size_t i = 0; i < got; ++i) {
                const pl_pool_result& r = res[i];
                --pending;
                if (r.error != PL_OK) {
                    std::lock_guard<std::mutex> lock(g_batch_mutex);
                    std::cout << "[BATCH-RESULT] k=" << r.k << ": no verdict (error " << r.error << ")\n";
                    continue;
                }
                if (r.verdict != 1) continue;
                std::lock_guard<std::mutex> lock(g_batch_mutex);
                std::cout << "[BATCH-RESULT] PRIME FOUND: k=" << r.k << "\n";
                std::ofstream f("primes.txt", std::ios::app);
                if (f) f << r.k << " * 2^" << r.n << " + 1 (a=" << r.base << ")\n";
            }
        }
    } catch (...) {
        pl_pool_destroy(pool);
        throw;
    }
    pl_pool_destroy(pool);
    return "BATCH";
}

}  // namespace

std::string run_task_with_ctx(const Task& t, FFTContext& ctx, uint64_t* limbs)
{
    unsigned n = 0;
    int report_step = 25;
    if (t.payload.count("n")) n = (unsigned)std::stoul(t.payload.at("n"));
    if (t.payload.count("report_step")) report_step = std::stoi(t.payload.at("report_step"));
    if (t.label == "Proth") {
        const uint64_t k = (uint64_t)std::stoull(t.payload.at("k"));
        return execute_proth_test_fft(k, n, limbs, ctx, report_step) ? "prime" : "PRP";
    }
    if (t.label == "PROTH_FF_BATCH") return batch(k_list(t.payload.at("k_list")), n, ctx.device);
    return "fail";
}
