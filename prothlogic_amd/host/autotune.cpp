// autotune.cpp -- auto_tune_fft_len over pl_autotune_fft_len (see
// autotune.h).
#include "autotune.h"

#include <cstdlib>
#include <iostream>

#include "prothcore.h"

size_t auto_tune_fft_len(unsigned actual_n, uint64_t sample_k)
{
    int device = 0;
    if (const char* d = std::getenv("PL_DEVICE")) device = std::atoi(d);
    std::cout << "\n[AUTOTUNE] starting engine calibration for N=" << actual_n << "...\n";
    uint32_t L = 0;
    double err = 0.0, us = 0.0;
    if (pl_autotune_fft_len(actual_n, sample_k | 1, device, 0.30, &L, &err, &us) != PL_OK) {
        std::cout << "[AUTOTUNE] failed: " << pl_last_error() << "\n\n";
        return 0;
    }
    std::cout << "  Max Err after 100 iterations = " << err << ", " << us << " us/squaring\n";
    std::cout << "[AUTOTUNE] Optimal FFT_Len: " << L << "\n\n";
    return L;
}
