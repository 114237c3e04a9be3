// selftest.cpp -- drives the C++ mirror (fft_engine.h, proth_logic.h,
// checkpoint_manager.h, autotune.h, worker.h) the way the reference's
// callers do, and prints one JSON line per case for an external checker.
// Run it in an empty directory: it writes checkpoints and primes.txt there.
//
//   prothhost_selftest                    the mirror's call sites, one by one
//   prothhost_selftest threads k:n,k:n    one full test per thread, at once
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "autotune.h"
#include "checkpoint_manager.h"
#include "fft_engine.h"
#include "proth_logic.h"
#include "prothcore.h"
#include "worker.h"

namespace {

std::string hex_limbs(const uint64_t* v, size_t n)
{
    std::string s;
    char b[17];
    for (size_t i = n; i-- > 0;) {
        std::snprintf(b, sizeof b, "%016llx", (unsigned long long)v[i]);
        s += b;
    }
    return s;
}

// The caller-side buffer size of P/dcore.cpp:143.
size_t limbs_for(size_t L, unsigned n) { return std::max<size_t>(L, (size_t)(2 * n) / 64 + 64); }

void full_test(uint64_t k, unsigned n)
{
    const size_t L = auto_tune_fft_len(n, k);
    FFTContext ctx;
    if (!fft_init(ctx, FFTBackend::INTERNAL, L)) throw std::runtime_error("fft_init failed");
    std::vector<uint64_t> limbs(limbs_for(L, n), 0);
    const bool prime = execute_proth_test_fft(k, n, limbs.data(), ctx, 50);
    std::printf("{\"case\": \"full\", \"k\": %llu, \"n\": %u, \"fft_len\": %zu, \"prime\": %d, \"res64\": \"%016llx\", "
                "\"max_err\": %.6g}\n",
                (unsigned long long)k, n, L, prime ? 1 : 0, (unsigned long long)limbs[0], ctx.last_max_diff);
    std::fflush(stdout);
    fft_destroy(ctx);
}

void square_chain(uint64_t k, unsigned n, int steps)
{
    FFTContext ctx;
    if (!fft_init(ctx, FFTBackend::INTERNAL, 1024)) throw std::runtime_error("fft_init failed");
    const size_t nl = (size_t)n / 64 + 2;
    std::vector<uint64_t> v(nl, 0);
    uint64_t s = 0x9E3779B97F4A7C15ull;
    for (size_t i = 0; i + 1 < nl; ++i) {
        s ^= s << 13, s ^= s >> 7, s ^= s << 17;
        v[i] = s;
    }
    const std::string x0 = hex_limbs(v.data(), nl);
    for (int i = 0; i < steps; ++i) fft_square_karatsuba(ctx, v.data(), nl, n, k);
    std::printf("{\"case\": \"square\", \"k\": %llu, \"n\": %u, \"steps\": %d, \"x0\": \"%s\", \"x\": \"%s\", "
                "\"last_max_diff\": %.6g}\n",
                (unsigned long long)k, n, steps, x0.c_str(), hex_limbs(v.data(), nl).c_str(), ctx.last_max_diff);
    std::fflush(stdout);
    fft_destroy(ctx);
}

// A checkpoint in the reference's layout saved after squaring index i (so
// i+1 squarings applied past the seed), then the test resumed from it.
void resume_from_reference_checkpoint(uint64_t k, unsigned n, uint64_t i)
{
    const uint64_t a = pl_choose_base(k, n);
    pl_ctx* c = nullptr;
    if (pl_ctx_create(k, n, 0, 0, &c) != PL_OK) throw std::runtime_error(pl_last_error());
    // seed a^k by left-to-right powering, then i+1 squarings
    uint64_t one[1] = { a };
    pl_set_residue(c, one, 1);
    int top = 63;
    while (top > 0 && !((k >> top) & 1)) --top;
    for (int b = top - 1; b >= 0; --b) {
        if (pl_square(c, 1, 0.0, nullptr, nullptr) != PL_OK) throw std::runtime_error(pl_last_error());
        if ((k >> b) & 1) pl_mul_small(c, a);
    }
    if (pl_square(c, i + 1, 0.0, nullptr, nullptr) != PL_OK) throw std::runtime_error(pl_last_error());
    const size_t nl = (size_t)n / 64 + 8;  // more limbs than N needs, as the reference's L/4 buffer has
    std::vector<uint64_t> x(nl, 0);
    if (pl_get_residue(c, x.data(), nl) != PL_OK) throw std::runtime_error(pl_last_error());
    pl_ctx_destroy(c);
    const std::string name = "check_k" + std::to_string(k) + "_n" + std::to_string(n) + ".bin";
    if (!save_checkpoint(i, nl, x.data(), name)) throw std::runtime_error("save_checkpoint failed");
    uint64_t back_i = 0;
    std::vector<uint64_t> y(nl, 0);
    const bool loaded = load_checkpoint(back_i, nl, y.data(), name) && back_i == i && y == x;
    FFTContext ctx;
    if (!fft_init(ctx, FFTBackend::INTERNAL, 4 * nl)) throw std::runtime_error("fft_init failed");
    std::vector<uint64_t> limbs(limbs_for(4 * nl, n), 0);
    const bool prime = execute_proth_test_fft(k, n, limbs.data(), ctx, 0);
    FILE* left = std::fopen(name.c_str(), "rb");
    std::printf("{\"case\": \"resume\", \"k\": %llu, \"n\": %u, \"saved_i\": %llu, \"roundtrip\": %d, \"prime\": %d, "
                "\"res64\": \"%016llx\", \"squarings_run\": %llu, \"checkpoint_removed\": %d}\n",
                (unsigned long long)k, n, (unsigned long long)i, loaded ? 1 : 0, prime ? 1 : 0,
                (unsigned long long)limbs[0], (unsigned long long)ctx.current_i, left ? 0 : 1);
    if (left) std::fclose(left);
    std::fflush(stdout);
    fft_destroy(ctx);
}

void worker_tasks(unsigned n, const std::string& ks, uint64_t single_k)
{
    FFTContext ctx;
    const size_t L = 4096;
    if (!fft_init(ctx, FFTBackend::INTERNAL, L)) throw std::runtime_error("fft_init failed");
    std::vector<uint64_t> limbs(limbs_for(L, n), 0);
    Task one{ n, single_k, "Proth", { { "n", std::to_string(n) }, { "k", std::to_string(single_k) }, { "report_step", "0" } } };
    const std::string r1 = run_task_with_ctx(one, ctx, limbs.data());
    Task many{ n, 0, "PROTH_FF_BATCH", { { "n", std::to_string(n) }, { "k_list", ks } } };
    const std::string r2 = run_task_with_ctx(many, ctx, limbs.data());
    Task bad{ n, 0, "Unknown", {} };
    const std::string r3 = run_task_with_ctx(bad, ctx, limbs.data());
    std::printf("{\"case\": \"worker\", \"n\": %u, \"single_k\": %llu, \"single\": \"%s\", \"batch\": \"%s\", "
                "\"other\": \"%s\", \"k_list\": \"%s\"}\n",
                n, (unsigned long long)single_k, r1.c_str(), r2.c_str(), r3.c_str(), ks.c_str());
    std::fflush(stdout);
    fft_destroy(ctx);
}

// One execute_proth_test_fft per std::thread, each with its own FFTContext
// and limbs, as the reference's dispatch runs its workers (P/dcore.cpp:134-175).
// spec: "k:n,k:n,...".
int worker_threads(const std::string& spec)
{
    std::vector<std::pair<uint64_t, unsigned>> cands;
    size_t pos = 0;
    while (pos < spec.size()) {
        size_t comma = spec.find(',', pos);
        if (comma == std::string::npos) comma = spec.size();
        const std::string item = spec.substr(pos, comma - pos);
        const size_t colon = item.find(':');
        if (colon == std::string::npos) return 2;
        cands.emplace_back(std::stoull(item.substr(0, colon)), (unsigned)std::stoul(item.substr(colon + 1)));
        pos = comma + 1;
    }
    std::vector<std::string> lines(cands.size());
    std::vector<std::thread> ts;
    for (size_t i = 0; i < cands.size(); ++i)
        ts.emplace_back([&, i] {
            const uint64_t k = cands[i].first;
            const unsigned n = cands[i].second;
            char buf[256];
            try {
                uint32_t L = 0;
                if (pl_plan_fft_len(n, k, &L) != PL_OK) throw std::runtime_error(pl_last_error());
                FFTContext ctx;
                if (!fft_init(ctx, FFTBackend::INTERNAL, L)) throw std::runtime_error("fft_init failed");
                std::vector<uint64_t> limbs(limbs_for(L, n), 0);
                const bool prime = execute_proth_test_fft(k, n, limbs.data(), ctx, 0);
                std::snprintf(buf, sizeof buf,
                              "{\"case\": \"thread\", \"k\": %llu, \"n\": %u, \"prime\": %d, \"res64\": \"%016llx\"}",
                              (unsigned long long)k, n, prime ? 1 : 0, (unsigned long long)limbs[0]);
                fft_destroy(ctx);
            } catch (const std::exception& e) {
                std::snprintf(buf, sizeof buf, "{\"case\": \"error\", \"k\": %llu, \"what\": \"%s\"}",
                              (unsigned long long)k, e.what());
            }
            lines[i] = buf;
        });
    for (auto& t : ts) t.join();
    int rc = 0;
    for (const auto& l : lines) {
        std::printf("%s\n", l.c_str());
        if (l.find("\"error\"") != std::string::npos) rc = 1;
    }
    std::fflush(stdout);
    return rc;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc >= 3 && std::string(argv[1]) == "threads") return worker_threads(argv[2]);
    try {
        std::printf("{\"case\": \"start\", \"abi\": %d}\n", pl_abi_version());
        std::fflush(stdout);
        full_test(3, 20909);     // prime
        full_test(1023, 32768);  // the n=32768 workload's candidate
        square_chain(843, 12000, 3);
        square_chain(65535, 60000, 2);
        resume_from_reference_checkpoint(3, 20909, 10000);
        worker_tasks(20909, "3,5,7,9,11,13,15,17,19,21", 5);
    } catch (const std::exception& e) {
        std::printf("{\"case\": \"error\", \"what\": \"%s\"}\n", e.what());
        return 1;
    }
    std::printf("{\"case\": \"done\"}\n");
    return 0;
}
