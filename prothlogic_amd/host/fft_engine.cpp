// fft_engine.cpp -- fft_init / fft_destroy / fft_square_karatsuba over the
// C ABI (include/prothcore.h).  See fft_engine.h.
#include "fft_engine.h"

#include <algorithm>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "prothcore.h"

namespace {

void bind(FFTContext& ctx, uint64_t k, unsigned n)
{
    if (ctx.dev && ctx.dev_k == k && ctx.dev_n == n) return;
    if (ctx.dev) pl_ctx_destroy(ctx.dev);
    ctx.dev = nullptr;
    pl_ctx* c = nullptr;
    if (pl_ctx_create(k, n, 0, ctx.device, &c) != PL_OK)
        throw std::runtime_error(std::string("fft_square_karatsuba: ") + pl_last_error());
    ctx.dev = c;
    ctx.dev_k = k;
    ctx.dev_n = n;
}

}  // namespace

bool fft_init(FFTContext& ctx, FFTBackend backend, size_t fft_len)
{
    ctx = FFTContext{};
    (void)backend;  // INTERNAL and HIP both select the device engine
    ctx.backend = FFTBackend::HIP;
    ctx.fft_len = fft_len;
    if (const char* d = std::getenv("PL_DEVICE")) ctx.device = std::atoi(d);
    int count = 0;
    if (pl_device_count(&count) != PL_OK || count <= ctx.device || ctx.device < 0) return false;
    return fft_len > 0;
}

void fft_destroy(FFTContext& ctx)
{
    if (ctx.dev) pl_ctx_destroy(ctx.dev);
    ctx.dev = nullptr;
    ctx.dev_k = 0;
    ctx.dev_n = 0;
}

void fft_square_karatsuba(FFTContext& ctx, uint64_t* v, size_t n_limbs, unsigned n_bits, uint64_t k_val)
{
    bind(ctx, k_val, n_bits);
    // a buffer that always holds the canonical result (< N has n_bits +
    // bitlen(k) + 1 bits); its low n_limbs limbs go back to the caller
    const size_t need = (size_t)n_bits / 64 + 3;
    std::vector<uint64_t> w(std::max(n_limbs, need), 0);
    for (size_t i = 0; i < n_limbs; ++i) w[i] = v[i];
    double e = 0.0;
    if (pl_fft_square_karatsuba(ctx.dev, w.data(), w.size(), n_bits, k_val, &e) != PL_OK)
        throw std::runtime_error(std::string("fft_square_karatsuba: ") + pl_last_error());
    for (size_t i = 0; i < n_limbs; ++i) v[i] = w[i];
    ctx.last_max_diff = e;
    ++ctx.prof_calls;
}
