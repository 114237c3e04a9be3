// cli.cpp -- command-line front end with the flags of P/cli.cpp:39-68 and its
// flow (P/cli.cpp:70-168): sieve (full range, or the list of --in-file),
// optional survivor file, then every survivor tested on the device.  Where
// the reference builds worker tasks (P/pipeline_proth.cpp:20-75) and
// dispatches one test per CPU thread, the survivors go to one device pool
// that advances them together (continuous batching); primes are appended to
// --out-primes as "k * 2^n + 1 (a=A)".
//
//   prothlogic --n <n> --k-start <a> --k-end <b> [--pmax P] [--pstart P0]
//              [--report R] [--out-primes F] [--out-sieve F] [--in-file F]
//              [--sieve-only] [--workers W]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>
#include <algorithm>

#include "presieve_core.h"
#include "proth_logic.h"
#include "prothcore.h"

namespace {

struct Options {
    unsigned n = 0;
    uint64_t k_start = 0, k_end = 0, pstart = 0, pmax = 50000000000ull;
    int workers = 0;  // accepted for compatibility; the device pool sizes itself
    int report_step = 25;
    std::string out_primes = "primes.txt", out_sieve, in_file;
    bool sieve_only = false;
};

bool parse(int argc, char** argv, Options& o)
{
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char* what) -> const char* {
            if (i + 1 >= argc) {
                std::cerr << "[ERROR] " << what << " needs a value\n";
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "--n") o.n = (unsigned)std::stoul(next("--n"));
        else if (a == "--k-start") o.k_start = std::stoull(next("--k-start"));
        else if (a == "--k-end") o.k_end = std::stoull(next("--k-end"));
        else if (a == "--pmax") o.pmax = std::stoull(next("--pmax"));
        else if (a == "--pstart") o.pstart = std::stoull(next("--pstart"));
        else if (a == "--workers") o.workers = std::stoi(next("--workers"));
        else if (a == "--report") o.report_step = std::stoi(next("--report"));
        else if (a == "--out-primes") o.out_primes = next("--out-primes");
        else if (a == "--out-sieve") o.out_sieve = next("--out-sieve");
        else if (a == "--in-file") o.in_file = next("--in-file");
        else if (a == "--sieve-only") o.sieve_only = true;
        else {
            std::cerr << "[ERROR] unknown argument " << a << "\n";
            return false;
        }
    }
    return true;
}

std::vector<uint64_t> load_ks(const std::string& path)
{
    std::vector<uint64_t> ks;
    std::ifstream f(path);
    if (!f) {
        std::cerr << "[ERROR] Cannot open file : " << path << "\n";
        return ks;
    }
    uint64_t k;
    while (f >> k) ks.push_back(k);
    return ks;
}

// Every candidate through one device pool; returns the number of primes.
size_t test_all(const std::vector<uint64_t>& ks, unsigned n, const std::string& out_primes)
{
    pl_pool* pool = nullptr;
    if (pl_pool_create(0, 0, 0, 0, &pool) != PL_OK) throw std::runtime_error(pl_pool_last_error());
    size_t pending = 0, primes = 0, done = 0;
    for (size_t i = 0; i < ks.size(); ++i) {
        if (pl_pool_submit(pool, ks[i], n, i, 0) != PL_OK) {
            pl_pool_destroy(pool);
            throw std::runtime_error(pl_pool_last_error());
        }
        ++pending;
    }
    std::vector<pl_pool_result> res(256);
    std::ofstream out(out_primes, std::ios::app);
    while (pending) {
        size_t got = 0;
        if (pl_pool_run(pool, 64, res.data(), res.size(), &got) != PL_OK) {
            pl_pool_destroy(pool);
            throw std::runtime_error(pl_pool_last_error());
        }
        for (size_t i = 0; i < got; ++i) {
            const pl_pool_result& r = res[i];
            --pending;
            ++done;
            if (r.error != PL_OK) {
                // void result (round-off above the guard, or carries that never settled)
                std::cout << "\n[ERROR] " << r.k << " * 2^" << r.n << " + 1: no verdict (error " << r.error
                          << ", max err " << r.max_err << ")" << std::endl;
                continue;
            }
            if (r.verdict == 1) {
                ++primes;
                std::cout << "\n[!] FOUND PRIME: " << r.k << " * 2^" << r.n << " + 1 (a=" << r.base << ")" << std::endl;
                out << r.k << " * 2^" << r.n << " + 1 (a=" << r.base << ")\n";
                out.flush();
            }
        }
        if (got) std::cout << "\r[TEST] " << done << " / " << ks.size() << " done" << std::flush;
    }
    std::cout << "\n";
    pl_pool_destroy(pool);
    return primes;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc <= 1) {
        std::cout << "Usage: prothlogic --n <n> --k-start <a> --k-end <b> [--pmax P] [--pstart P0] [--report R]\n"
                     "                  [--out-primes F] [--out-sieve F] [--in-file F] [--sieve-only]\n";
        return 0;
    }
    Options o;
    if (!parse(argc, argv, o)) return 1;
    if (o.n == 0 || (o.k_start == 0 && o.in_file.empty())) {
        std::cerr << "[ERROR] Missing required parameters --n and --k-start" << std::endl;
        return 1;
    }
    if (o.k_end < o.k_start) o.k_end = o.k_start;
    using clk = std::chrono::steady_clock;
    try {
        std::vector<uint64_t> survivors;
        if (!o.in_file.empty()) {
            const std::vector<uint64_t> ks = load_ks(o.in_file);
            std::cout << "[SYSTEM] Loaded " << ks.size() << " candidates from " << o.in_file << "\n";
            const auto t0 = clk::now();
            if (!ks.empty()) survivors = execute_sparse_streaming_sieve(ks, o.n, o.pstart, o.pmax);
            std::cout << "[DONE] Sparse sieve time: " << std::chrono::duration<double>(clk::now() - t0).count() << " s\n";
        } else if (o.k_start == o.k_end) {
            // one candidate: the full test with progress lines (P/cli.cpp:92-101)
            FFTContext ctx;
            if (!fft_init(ctx, FFTBackend::HIP, 4096)) {
                std::cerr << "[ERROR] no HIP device\n";
                return 1;
            }
            // the caller-side buffer of P/dcore.cpp:143: max(fft_len, 2n/64 + 64) limbs
            std::vector<uint64_t> limbs(std::max<size_t>(4096, (size_t)o.n / 32 + 64), 0);
            const bool prime = execute_proth_test_fft(o.k_start, o.n, limbs.data(), ctx, o.report_step);
            char r64[32];
            std::snprintf(r64, sizeof r64, "%016llX", (unsigned long long)limbs[0]);
            std::cout << "Res64: " << r64 << "  max round-off " << ctx.last_max_diff << "\n";
            if (prime && o.out_primes != "primes.txt") {
                std::ofstream out(o.out_primes, std::ios::app);
                out << o.k_start << " * 2^" << o.n << " + 1 (a=" << pl_choose_base(o.k_start, o.n) << ")\n";
            }
            fft_destroy(ctx);
            return 0;
        } else {
            const auto t0 = clk::now();
            survivors = execute_full_streaming_sieve(o.k_start, o.k_end, o.n, o.pmax);
            std::cout << "[DONE] Sieve time: " << std::chrono::duration<double>(clk::now() - t0).count() << " s\n";
        }
        if (!o.out_sieve.empty() && !survivors.empty()) {
            std::ofstream f(o.out_sieve);
            for (uint64_t k : survivors) f << k << "\n";
            std::cout << "[SIEVE] Saving " << survivors.size() << " surviving candidates to a file: " << o.out_sieve
                      << "\n";
        }
        if (o.sieve_only) {
            std::cout << "[INFO] The --sieve-only flag is active. Stopping before the tests.\n";
            return 0;
        }
        if (survivors.empty()) {
            std::cout << "[INFO] No candidates survived the sieve. Finishing." << std::endl;
            return 0;
        }
        std::cout << "\nStarting Proth Test (" << survivors.size() << " candidates to test)" << std::endl;
        const auto t0 = clk::now();
        const size_t primes = test_all(survivors, o.n, o.out_primes);
        const double s = std::chrono::duration<double>(clk::now() - t0).count();
        std::cout << "Calculations finished: " << primes << " prime(s) of " << survivors.size() << " in " << s
                  << " s (" << 3600.0 * survivors.size() / s << " tests/hour)" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "[ERROR] " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
