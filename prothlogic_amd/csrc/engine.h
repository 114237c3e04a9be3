// engine.h -- device-resident modular squaring x <- x^2 * 2^(-2n) mod N,
// N = k*2^n + 1, on one gfx950 GPU.  See engine.hip for the algorithm.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace pleng {

// Complex length L = R*C digits; R = strided (column) DFT length, C =
// contiguous (row) DFT length.
struct Plan {
    uint32_t n = 0;
    uint64_t k = 0;
    uint32_t L = 0, logL = 0, R = 0, C = 0;
    uint32_t s = 0;      // log2 of the low half of the two-level tables
    bool fused = false;  // steady state uses the split outer pass (3 kernels per squaring)
};

// Rules for L: smallest power of two in [2^10, 2^22] whose bits per digit
// n/L stays within max_bits_per_digit(L) and above 8.  Returns false when
// no length qualifies.
bool choose_plan(uint32_t n, uint64_t k, uint32_t force_L, Plan& out, std::string& why);
double max_bits_per_digit(uint32_t logL);

struct DeviceTables;
struct Params;

class Engine {
  public:
    Engine(const Plan& p, int device, bool allow_fused = true);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    const Plan& plan() const { return plan_; }
    hipStream_t stream() const { return stream_; }
    size_t device_bytes() const { return bytes_; }

    // Digits (variable base, balanced) + top scalar in/out of the device.
    void upload(const std::vector<int32_t>& digits, int64_t xhi);
    void download(std::vector<int32_t>& digits, int64_t& xhi);

    // iters squarings on the device; per-iteration max round-off into
    // err_trace (iters entries) when non-null.  Returns the max round-off.
    double square(uint64_t iters, double* err_trace);

    // device status word (non-zero: a carry did not settle within a run)
    int status();

    // Average duration (ms) of each kernel, timed with HIP events on this
    // engine's stream over `reps` launches each: [0] forward column pass,
    // [1] row pass (DFT, square, inverse DFT), [2] inverse column pass,
    // [3] flat carry, [4] block fix-up, [5] inverse column DFT + carry
    // (split outer pass), [6] carry-in + forward column DFT (split outer
    // pass); [5] and [6] are 0 when the plan does not use them.  The residue
    // is saved before and restored after.
    void kernel_times(int reps, double* ms);
    static constexpr int kNumKernels = 7;

    static constexpr int kGraphChunk = 64;

  private:
    Params params(uint64_t slot) const;
    void carry(const Params& p, bool do_carry, bool do_fixup);
    void enqueue_chunk(int g);
    void build_graph(int iters);

    Plan plan_;
    int device_;
    hipStream_t stream_ = nullptr;
    size_t bytes_ = 0;
    void* mem_ = nullptr;
    double2* z_ = nullptr;
    long long* carries_ = nullptr;
    long long* runc_ = nullptr;
    int* rdig_ = nullptr;
    long long* xhi_ = nullptr;
    unsigned long long* err_ = nullptr;  // kGraphChunk slots (bit patterns of non-negative doubles)
    int* status_ = nullptr;
    int32_t* digits_dev_ = nullptr;
    DeviceTables* tabs_ = nullptr;
    uint32_t carry_blocks_ = 0;
    int run_ = 16;         // digits per thread in the flat carry kernel
    int carry_tpb_ = 64;   // threads per flat carry block
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    int graph_iters_ = 0;
    std::vector<unsigned long long> err_host_;
};

void check(hipError_t e, const char* what);

}  // namespace pleng
