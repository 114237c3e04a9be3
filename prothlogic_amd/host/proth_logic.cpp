// proth_logic.cpp -- execute_proth_test_fft over pl_proth_test (see
// proth_logic.h).
#include "proth_logic.h"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <string>

#include "prothcore.h"

namespace {

std::mutex g_out_mutex;

struct Report {
    uint64_t k;
};

void on_progress(void* user, uint64_t done, uint64_t total, double max_err)
{
    const Report* r = static_cast<const Report*>(user);
    const int pct = (int)(done * 100 / total);
    char err[32];
    std::snprintf(err, sizeof err, "%.2e", max_err);
    std::lock_guard<std::mutex> lock(g_out_mutex);
    std::cout << "    [STEP] k=" << r->k << " -> " << pct << "% | Max Err: " << err << std::endl;
}

}  // namespace

bool execute_proth_test_fft(uint64_t k, unsigned n, uint64_t* limbs, FFTContext& ctx, int report_step)
{
    if (pl_gate_rejects(k, n)) return false;
    const std::string ckpt = "check_k" + std::to_string(k) + "_n" + std::to_string(n) + ".bin";
    const size_t nl = std::max<size_t>(ctx.fft_len / 4, ((size_t)n + 127) / 64);
    Report rep{ k };
    pl_test_args a{};
    a.k = k;
    a.n = n;
    a.fft_len = 0;
    a.device = ctx.device;
    a.checkpoint_path = ckpt.c_str();
    a.checkpoint_every = 10000;
    a.report_step = report_step;
    a.progress = report_step > 0 ? on_progress : nullptr;
    a.progress_user = &rep;
    a.final_limbs = limbs;
    a.final_n_limbs = nl;
    pl_test_result r{};
    if (pl_proth_test(&a, &r) != PL_OK) throw std::runtime_error(std::string("execute_proth_test_fft: ") + pl_last_error());
    ctx.last_max_diff = r.max_err;
    ctx.current_i = r.iters_done;
    const bool prime = r.verdict == 1;
    {
        std::lock_guard<std::mutex> lock(g_out_mutex);
        if (prime)
            std::cout << "\n[!] FOUND PRIME: " << k << " * 2^" << n << " + 1 (a=" << r.base << ")" << std::endl;
        else
            std::cout << "[+] " << k << " * 2^" << n << " + 1 is COMPOSITE" << std::endl;
    }
    if (prime) {
        std::lock_guard<std::mutex> lock(g_out_mutex);
        std::ofstream f("primes.txt", std::ios::app);
        if (f) f << k << " * 2^" << n << " + 1 (a=" << r.base << ")\n";
    }
    return prime;
}
