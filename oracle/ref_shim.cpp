// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C-ABI shim, written for this repository, over the reference's own C++
// functions so that tests and the CPU-baseline leg of bench.py can call the
// reference engine through ctypes.  It is compiled together with the
// reference's unmodified sources by oracle/Makefile into oracle/_ref/ (never
// committed; built only where the reference tree is present).
//
// Bound reference entry points:
//   fft_init / fft_destroy / fft_square_karatsuba   P/fft_engine.h:10,11,22
//   auto_tune_fft_len                               P/dcore.cpp:20 (no header)
//   execute_proth_test_fft                          P/proth_logic.h:12
#include "fft_engine.h"
#include "proth_logic.h"

#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <new>

size_t auto_tune_fft_len(unsigned actual_n, uint64_t sample_k);  // P/dcore.cpp:20

extern "C" {

void* ref_fft_init(size_t fft_len)
{
    FFTContext* c = new (std::nothrow) FFTContext();
    if (!c) return nullptr;
    if (!fft_init(*c, FFTBackend::INTERNAL, fft_len)) {
        delete c;
        return nullptr;
    }
    return c;
}

void ref_fft_destroy(void* c)
{
    if (!c) return;
    fft_destroy(*static_cast<FFTContext*>(c));
    delete static_cast<FFTContext*>(c);
}

void ref_fft_square(void* c, uint64_t* v, size_t n_limbs, unsigned n_bits, uint64_t k)
{
    fft_square_karatsuba(*static_cast<FFTContext*>(c), v, n_limbs, n_bits, k);
}

double ref_last_max_diff(void* c) { return static_cast<FFTContext*>(c)->last_max_diff; }

size_t ref_auto_tune_fft_len(unsigned n, uint64_t k) { return auto_tune_fft_len(n, k); }

// Runs the reference driver for one candidate at a caller-chosen length,
// exactly as a dispatch worker would (P/dcore.cpp:134-175), in the current
// working directory (the driver reads/writes its checkpoint and primes.txt
// there).  The final limbs are copied into out[0..out_cap).
int ref_execute_proth_test(uint64_t k, unsigned n, size_t fft_len, uint64_t* out, size_t out_cap)
{
    FFTContext ctx;
    if (!fft_init(ctx, FFTBackend::INTERNAL, fft_len)) return -2;
    size_t safe = std::max(fft_len, (size_t)((2 * (size_t)n) / 64 + 64));
    size_t bytes = safe * sizeof(uint64_t);
    uint64_t* limbs = (uint64_t*)mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (limbs == MAP_FAILED) {
        fft_destroy(ctx);
        return -3;
    }
    std::memset(limbs, 0, bytes);
    bool prime = execute_proth_test_fft(k, n, limbs, ctx, 0);
    std::memcpy(out, limbs, std::min(out_cap, safe) * sizeof(uint64_t));
    munmap(limbs, bytes);
    fft_destroy(ctx);
    return prime ? 1 : 0;
}

}  // extern "C"
