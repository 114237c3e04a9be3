// noalloc.hip -- pl_fft_square_karatsuba_noalloc (include/prothcore.h): the
// per-call squaring shim for callers that cannot allocate device memory.
//
// Every device buffer it uses -- the weighted residue, the digit staging,
// the flat carries, the round-off partials, the status word and the
// transform tables -- lies in `scratch`, one static array per GPU, laid out
// as a one-slot Group (engine.h) for the call's transform length.  The host
// only ever passes `ptr`, the array's device address from
// hipGetSymbolAddress, never the array's name.  The array keeps the last
// call's contents: a call that changes n or k lays the Group out again
// (zeroing its carries, status and round-off partials and uploading its
// tables), and every call rewrites the residue and clears the status and
// partials it reads.  Calls are serialised per device within the process;
// calls from several processes on one GPU each have their own array.
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "bignum.h"
#include "engine.h"
#include "prothcore.h"

// 6 MiB: a one-slot group at L = 2^18 needs 5.7 MiB (16 B per point of
// residue, 4 B per point of digit staging, the round-off partials and the
// tables).
constexpr size_t kNoAllocBytes = 6u << 20;
constexpr uint32_t kNoAllocMaxLogL = 18;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;

struct DevScratch {
    std::mutex mu;
    void* ptr = nullptr;  // device address of this device's `scratch`
    std::unique_ptr<pleng::Group> g;
    std::unique_ptr<plbn::Modulus> M;
    uint32_t n = 0;
    uint64_t k = 0;
};
DevScratch g_scratch[kMaxDevices];

}  // namespace

namespace plerr {
int fail(int code, const std::string& msg);  // sets pl_last_error() (capi.cpp)
}

namespace {
int na_fail(int code, const std::string& m) { return plerr::fail(code, m); }
}  // namespace

extern "C" {

int pl_fft_square_karatsuba_noalloc(int device, uint64_t* v, size_t n_limbs, unsigned n_bits, uint64_t k,
                                    double* last_max_diff)
{
    if (!v || n_limbs == 0) return na_fail(PL_ERR_ARG, "NULL or empty limb buffer");
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return na_fail(PL_ERR_HIP, "no HIP device visible");
    if (device < 0 || device >= nd || device >= kMaxDevices) return na_fail(PL_ERR_ARG, "device index out of range");
    pleng::Plan plan;
    std::string why;
    if (!pleng::choose_plan(n_bits, k, 0, plan, why)) return na_fail(PL_ERR_ARG, why);
    if (plan.logL > kNoAllocMaxLogL)
        return na_fail(PL_ERR_ARG, "n needs a transform length above 2^" + std::to_string(kNoAllocMaxLogL) +
                                       ", the most the static scratch of the no-allocation entry point holds");
    if (!pleng::choose_layout(plan, (double)n_bits / plan.L, k, 1, false, why)) return na_fail(PL_ERR_ARG, why);
    DevScratch& d = g_scratch[device];
    std::lock_guard<std::mutex> lock(d.mu);
    try {
        if (hipSetDevice(device) != hipSuccess) return na_fail(PL_ERR_HIP, "hipSetDevice failed");
        if (!d.ptr) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return na_fail(PL_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            d.ptr = p;
        }
        if (pleng::Group::layout_bytes(plan, 1) > kNoAllocBytes)
            return na_fail(PL_ERR_ARG, "transform too large for the static scratch");
        if (!d.g || d.n != n_bits || d.k != k) {
            d.g.reset();
            d.M.reset();
            // the one-slot layout for this candidate, in the static array, on
            // the legacy stream (no stream is created)
            d.g = std::make_unique<pleng::Group>(plan, 1, device, d.ptr, kNoAllocBytes, nullptr);
            d.M = std::make_unique<plbn::Modulus>(n_bits, k);
            d.n = n_bits;
            d.k = k;
        }
        pleng::Group& G = *d.g;
        const plbn::Modulus& M = *d.M;
        // The slot holds x itself: one squaring leaves x^2 * 2^(-2n) =
        // x^2 * k^2, and two multiplications by k^-1 on the host give x^2.
        const plbn::Limbs x = plbn::mod_N(M, plbn::Limbs(v, v + n_limbs));
        std::vector<int32_t> digits;
        int64_t xhi = 0;
        plbn::to_digits(M, plan.L, x, digits, xhi);
        G.clear_status();
        G.upload(0, digits, xhi);
        const double err = G.square_slots(0, 1, 1, nullptr);
        const int st = G.status();
        if (st) {
            G.clear_status();
            return na_fail(PL_ERR_STATE, "device carry did not settle (status " + std::to_string(st) + ")");
        }
        G.download(0, digits, xhi);
        const plbn::Limbs y = plbn::mul_kinv(M, plbn::mul_kinv(M, plbn::from_digits(M, plan.L, digits, xhi)));
        size_t used = y.size();
        while (used > 0 && y[used - 1] == 0) --used;
        if (used > n_limbs) return na_fail(PL_ERR_ARG, "n_limbs too small for the result");
        for (size_t i = 0; i < n_limbs; ++i) v[i] = i < y.size() ? y[i] : 0;
        if (last_max_diff) *last_max_diff = err;
        return PL_OK;
    } catch (const std::invalid_argument& e) {
        return na_fail(PL_ERR_ARG, e.what());
    } catch (const std::exception& e) {
        d.g.reset();  // the layout may be half set up: lay it out again next call
        return na_fail(PL_ERR_HIP, e.what());
    }
}

}  // extern "C"
