// sieve.hip -- Proth candidate sieve on one gfx950 GPU: removes every k for
// which some prime p <= pmax divides k*2^n + 1, i.e. k == -2^(-n) (mod p).
// The step before the squaring path (P/presieve_core.cpp:15-167 over
// P/gpu_engine.cu:82-462; targets as P/gpu_engine.cu:304-308), rebuilt as:
//   1. primes of a window [lo, hi) by a segmented sieve of Eratosthenes in
//      LDS (one block per 2^16 odd numbers, sieving primes <= sqrt(hi) from
//      the host), compacted per block into a small-prime list (p <= 2^18)
//      and a large-prime list;
//   2. t_p = -2^(-n) mod p per prime (64-bit Montgomery powering);
//   3. strikes into a bitmap of the k range: small primes (many strikes
//      each) block-cooperatively into a 2^18-bit LDS tile that is ANDed into
//      the bitmap once, large primes one thread per prime with global
//      atomic ANDs (at most range/p + 1 strikes each).
// The list form (sparse sieve) checks k mod p == t_p for every k of a list
// cached in LDS, with a Barrett reduction per k.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "prothcore.h"

namespace plsieve {

constexpr int kSegOdd = 1 << 16;        // odd numbers per prime-sieve block
constexpr int kSegWords = kSegOdd / 32;
constexpr int kTileBits = 1 << 18;      // k values per strike tile
constexpr int kTileWords = kTileBits / 32;
constexpr uint64_t kSmallMax = (uint64_t)kTileBits;  // small primes: p <= 2^18
constexpr uint64_t kSmallCap = 24000;                // pi(2^18) = 23000

void hcheck(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ math

__device__ __forceinline__ uint64_t mont_ninv(uint64_t p)  // -p^-1 mod 2^64 (p odd)
{
    uint64_t x = p;                      // p*p == 1 mod 8
    for (int i = 0; i < 5; ++i) x *= 2 - p * x;
    return 0 - x;
}

// a*b*2^-64 mod p for a, b < p < 2^63
__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b, uint64_t p, uint64_t ninv)
{
    const uint64_t lo = a * b, hi = __umul64hi(a, b);
    const uint64_t m = lo * ninv;
    const uint64_t mlo = m * p, mhi = __umul64hi(m, p);
    const uint64_t carry = (lo + mlo) < lo ? 1 : 0;  // lo + mlo == 0 mod 2^64
    uint64_t t = hi + mhi + carry;
    return t >= p ? t - p : t;
}

// -2^(-n) mod p for an odd prime p < 2^63: ((p+1)/2)^n, negated.
__device__ uint64_t proth_target(uint64_t p, uint32_t n)
{
    const uint64_t ninv = mont_ninv(p);
    const uint64_t r1 = (0 - p) % p;  // 2^64 mod p
    uint64_t r2 = r1;                  // 2^128 mod p by 64 doublings
    for (int i = 0; i < 64; ++i) r2 = r2 >= p - r2 ? r2 - (p - r2) : r2 + r2;
    uint64_t b = mont_mul((p + 1) / 2, r2, p, ninv);                    // Montgomery form
    uint64_t x = r1;                                                    // 1 in Montgomery form
    for (uint32_t e = n; e; e >>= 1) {
        if (e & 1) x = mont_mul(x, b, p, ninv);
        b = mont_mul(b, b, p, ninv);
    }
    const uint64_t v = mont_mul(x, 1, p, ninv);  // leave the Montgomery form
    return v == 0 ? 0 : p - v;
}

// ------------------------------------------------------------------ kernels

// Primes of the odd numbers lo + 2i, i < count (lo odd), block b taking
// i in [b*kSegOdd, (b+1)*kSegOdd).  Sieving primes q (3 <= q, q*q < hi)
// strike their odd multiples >= q*q in an LDS bitmap; survivors x with
// 3 <= x <= pmax go to the small list (x <= kSmallMax) or the large list.
// One atomic per block and list reserves the output range.
__global__ __launch_bounds__(256) void k_prime_segment(uint64_t lo, uint64_t count, uint64_t pmax,
                                                       const uint32_t* __restrict__ q, uint32_t nq,
                                                       uint64_t* __restrict__ small, unsigned long long* n_small,
                                                       uint64_t* __restrict__ large, unsigned long long* n_large,
                                                       uint64_t large_cap)
{
    __shared__ uint32_t bits[kSegWords];
    __shared__ uint32_t cnt_s[256], cnt_l[256];
    __shared__ unsigned long long base_s, base_l;
    const uint64_t i0 = (uint64_t)blockIdx.x * kSegOdd;
    const uint64_t seg_n = min((uint64_t)kSegOdd, count - i0);
    const uint64_t x0 = lo + 2 * i0;               // first odd number of the segment
    const uint64_t x_end = x0 + 2 * seg_n;         // exclusive
    for (int w = threadIdx.x; w < kSegWords; w += blockDim.x) bits[w] = 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < nq; j += blockDim.x) {
        const uint64_t p = q[j];
        if (p * p >= x_end) continue;
        uint64_t m = ((x0 + p - 1) / p) * p;
        if (m < p * p) m = p * p;
        if ((m & 1) == 0) m += p;
        for (uint64_t i = (m - x0) / 2; i < seg_n; i += p) atomicAnd(&bits[i >> 5], ~(1u << (i & 31)));
    }
    __syncthreads();
    // count this thread's primes (words w = tid, tid + 256, ...)
    uint32_t cs = 0, cl = 0;
    for (int w = threadIdx.x; w < kSegWords; w += blockDim.x) {
        uint32_t v = bits[w];
        while (v) {
            const int b = __ffs(v) - 1;
            v &= v - 1;
            const uint64_t i = (uint64_t)w * 32 + b;
            if (i >= seg_n) break;
            const uint64_t x = x0 + 2 * i;
            if (x < 3 || x > pmax) continue;
            if (x <= kSmallMax) ++cs;
            else ++cl;
        }
    }
    cnt_s[threadIdx.x] = cs;
    cnt_l[threadIdx.x] = cl;
    __syncthreads();
    if (threadIdx.x == 0) {  // exclusive scan over the block's 256 threads
        uint32_t as = 0, al = 0;
        for (int t = 0; t < 256; ++t) {
            const uint32_t s = cnt_s[t], l = cnt_l[t];
            cnt_s[t] = as;
            cnt_l[t] = al;
            as += s;
            al += l;
        }
        base_s = as ? atomicAdd(n_small, (unsigned long long)as) : 0;
        base_l = al ? atomicAdd(n_large, (unsigned long long)al) : 0;
    }
    __syncthreads();
    uint64_t os = base_s + cnt_s[threadIdx.x], ol = base_l + cnt_l[threadIdx.x];
    for (int w = threadIdx.x; w < kSegWords; w += blockDim.x) {
        uint32_t v = bits[w];
        while (v) {
            const int b = __ffs(v) - 1;
            v &= v - 1;
            const uint64_t i = (uint64_t)w * 32 + b;
            if (i >= seg_n) break;
            const uint64_t x = x0 + 2 * i;
            if (x < 3 || x > pmax) continue;
            if (x <= kSmallMax) {
                if (os < kSmallCap) small[os++] = x;
            } else if (ol < large_cap) {
                large[ol++] = x;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_targets(const uint64_t* __restrict__ p, uint64_t* __restrict__ t,
                                                 uint64_t count, uint32_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) t[i] = proth_target(p[i], n);
}

__device__ __forceinline__ uint64_t first_offset(uint64_t k0, uint64_t p, uint64_t t)
{
    const uint64_t r = k0 % p;
    return t >= r ? t - r : t + p - r;
}

// Small primes: block b owns k0 + [b*kTileBits, (b+1)*kTileBits); for each
// prime in turn every thread strikes every TPB-th multiple in the LDS tile.
__global__ __launch_bounds__(256) void k_strike_small(uint32_t* __restrict__ bitmap, uint64_t k0, uint64_t nbits,
                                                      const uint64_t* __restrict__ p, const uint64_t* __restrict__ t,
                                                      uint64_t count)
{
    __shared__ uint32_t tile[kTileWords];
    const uint64_t j0 = (uint64_t)blockIdx.x * kTileBits;
    const uint64_t tile_n = min((uint64_t)kTileBits, nbits - j0);
    for (int w = threadIdx.x; w < kTileWords; w += blockDim.x) tile[w] = 0xFFFFFFFFu;
    __syncthreads();
    for (uint64_t e = 0; e < count; ++e) {
        const uint64_t pe = p[e];
        const uint64_t off = first_offset(k0 + j0, pe, t[e]);
        for (uint64_t j = off + threadIdx.x * pe; j < tile_n; j += (uint64_t)blockDim.x * pe)
            atomicAnd(&tile[j >> 5], ~(1u << (j & 31)));
    }
    __syncthreads();
    const uint64_t w0 = j0 / 32;
    const uint64_t nw = (tile_n + 31) / 32;
    for (uint64_t w = threadIdx.x; w < nw; w += blockDim.x) {
        const uint32_t m = tile[w];
        if (m != 0xFFFFFFFFu) atomicAnd(&bitmap[w0 + w], m);
    }
}

// Large primes: one thread per prime, global atomic ANDs.
__global__ __launch_bounds__(256) void k_strike_large(uint32_t* __restrict__ bitmap, uint64_t k0, uint64_t nbits,
                                                      const uint64_t* __restrict__ p, const uint64_t* __restrict__ t,
                                                      uint64_t count)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t pe = p[i];
    for (uint64_t j = first_offset(k0, pe, t[i]); j < nbits; j += pe) atomicAnd(&bitmap[j >> 5], ~(1u << (j & 31)));
}

__global__ void k_fill_u32(uint32_t* a, uint64_t count, uint32_t v)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = v;
}

// List form: one thread per prime, every k of the list (cached in LDS,
// `nk` <= kMaxList) tested as k mod p == t_p with a Barrett reduction;
// a hit clears keep[k's index].
constexpr int kMaxList = 4096;

// k mod p with mu = floor((2^64 - 1)/p): the quotient estimate is short by
// at most 2.
__device__ __forceinline__ uint64_t mod_barrett(uint64_t k, uint64_t p, uint64_t mu)
{
    const uint64_t q = __umul64hi(k, mu);
    uint64_t r = k - q * p;
    while (r >= p) r -= p;
    return r;
}

__global__ __launch_bounds__(256) void k_strike_list(const uint64_t* __restrict__ ks, uint32_t nk, uint8_t* keep,
                                                     const uint64_t* __restrict__ p, const uint64_t* __restrict__ t,
                                                     uint64_t count)
{
    __shared__ uint64_t lk[kMaxList];
    for (uint32_t i = threadIdx.x; i < nk; i += blockDim.x) lk[i] = ks[i];
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t pe = p[i], te = t[i];
    const uint64_t mu = ~0ull / pe;
    for (uint32_t j = 0; j < nk; ++j)
        if (mod_barrett(lk[j], pe, mu) == te) keep[j] = 0;
}

// ------------------------------------------------------------------ host

// Odd primes q with q*q <= limit (host sieve; limit < 2^64 so q < 2^32).
std::vector<uint32_t> sieving_primes(uint64_t limit)
{
    uint64_t r = (uint64_t)std::sqrt((long double)limit);
    while ((r + 1) * (r + 1) <= limit) ++r;
    while (r * r > limit) --r;
    std::vector<uint8_t> comp(r + 1, 0);
    std::vector<uint32_t> q;
    for (uint64_t i = 2; i <= r; ++i) {
        if (comp[i]) continue;
        if (i > 2) q.push_back((uint32_t)i);
        for (uint64_t j = i * i; j <= r; j += i) comp[j] = 1;
    }
    return q;
}

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { hcheck(hipMalloc(&p, std::max<size_t>(bytes, 16)), "hipMalloc"); }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

// Calls `visit(p, t, count, small)` on device arrays of primes and their
// targets for every window of odd numbers in [pstart, pmax].
template <class Visit>
void for_each_prime_batch(uint64_t pstart, uint64_t pmax, uint32_t n, hipStream_t st, Visit visit)
{
    // odd numbers per window (PL_SIEVE_WINDOW_ODD, a power of two >= 2^16,
    // tunes it: more windows use less HBM for the prime lists)
    uint64_t window_odd = 1ull << 26;
    if (const char* e = std::getenv("PL_SIEVE_WINDOW_ODD")) {
        const uint64_t v = std::strtoull(e, nullptr, 10);
        if (v >= (uint64_t)kSegOdd && (v & (v - 1)) == 0) window_odd = v;
    }
    const std::vector<uint32_t> q = sieving_primes(pmax);
    DevBuf dq(q.size() * sizeof(uint32_t));
    if (!q.empty())
        hcheck(hipMemcpyAsync(dq.p, q.data(), q.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st), "H2D q");
    const uint64_t small_cap = kSmallCap;
    const uint64_t large_cap = window_odd;  // a window cannot hold more primes than odd numbers
    DevBuf ps(small_cap * 8), ts(small_cap * 8), pl(large_cap * 8), tl(large_cap * 8), cnt(16);
    uint64_t lo = std::max<uint64_t>(pstart, 3) | 1;
    while (lo <= pmax) {
        const uint64_t count = std::min<uint64_t>(window_odd, (pmax - lo) / 2 + 1);
        hcheck(hipMemsetAsync(cnt.p, 0, 16, st), "reset counts");
        const uint32_t blocks = (uint32_t)((count + kSegOdd - 1) / kSegOdd);
        hipLaunchKernelGGL(k_prime_segment, dim3(blocks), dim3(256), 0, st, lo, count, pmax, dq.as<uint32_t>(),
                           (uint32_t)q.size(), ps.as<uint64_t>(), cnt.as<unsigned long long>(), pl.as<uint64_t>(),
                           cnt.as<unsigned long long>() + 1, large_cap);
        hcheck(hipGetLastError(), "k_prime_segment");
        unsigned long long c[2] = { 0, 0 };
        hcheck(hipMemcpyAsync(c, cnt.p, 16, hipMemcpyDeviceToHost, st), "D2H counts");
        hcheck(hipStreamSynchronize(st), "prime window");
        if (c[0] > small_cap || c[1] > large_cap) throw std::runtime_error("prime buffer overflow");
        if (c[0]) {
            hipLaunchKernelGGL(k_targets, dim3((unsigned)((c[0] + 255) / 256)), dim3(256), 0, st, ps.as<uint64_t>(),
                               ts.as<uint64_t>(), (uint64_t)c[0], n);
            visit(ps.as<uint64_t>(), ts.as<uint64_t>(), (uint64_t)c[0], true);
        }
        if (c[1]) {
            hipLaunchKernelGGL(k_targets, dim3((unsigned)((c[1] + 255) / 256)), dim3(256), 0, st, pl.as<uint64_t>(),
                               tl.as<uint64_t>(), (uint64_t)c[1], n);
            visit(pl.as<uint64_t>(), tl.as<uint64_t>(), (uint64_t)c[1], false);
        }
        hcheck(hipStreamSynchronize(st), "strike window");
        if (pmax - lo < 2 * count) break;
        lo += 2 * count;
    }
}

thread_local std::string g_err;

int sfail(int code, const std::string& m)
{
    g_err = m;
    return code;
}

int select_device(int device)
{
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return sfail(PL_ERR_HIP, "no HIP device visible");
    if (device < 0 || device >= nd) return sfail(PL_ERR_ARG, "device index out of range");
    if (hipSetDevice(device) != hipSuccess) return sfail(PL_ERR_HIP, "hipSetDevice failed");
    return PL_OK;
}

}  // namespace plsieve

using namespace plsieve;

extern "C" {

const char* pl_sieve_last_error(void) { return g_err.c_str(); }

int pl_sieve_range(uint64_t k_start, uint64_t k_end, uint32_t n, uint64_t pmax, int device, uint64_t* out_ks,
                   size_t cap, size_t* n_out)
{
    if (!n_out || (!out_ks && cap)) return sfail(PL_ERR_ARG, "NULL argument");
    if (k_end < k_start) {
        *n_out = 0;
        return PL_OK;
    }
    if (k_end - k_start >= (1ull << 36)) return sfail(PL_ERR_ARG, "k range above 2^36 values");
    if (pmax >= (1ull << 62)) return sfail(PL_ERR_ARG, "pmax must be below 2^62");
    int rc = select_device(device);
    if (rc) return rc;
    hipStream_t st = nullptr;
    try {
        hcheck(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        const uint64_t nbits = k_end - k_start + 1;
        const uint64_t nwords = (nbits + 31) / 32;
        DevBuf bm(nwords * 4);
        hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, st, bm.as<uint32_t>(),
                           nwords, 0xFFFFFFFFu);
        if (pmax >= 3) {
            for_each_prime_batch(3, pmax, n, st, [&](const uint64_t* p, const uint64_t* t, uint64_t c, bool small) {
                if (small)
                    hipLaunchKernelGGL(k_strike_small, dim3((unsigned)((nbits + kTileBits - 1) / kTileBits)), dim3(256),
                                       0, st, bm.as<uint32_t>(), k_start, nbits, p, t, c);
                else
                    hipLaunchKernelGGL(k_strike_large, dim3((unsigned)((c + 255) / 256)), dim3(256), 0, st,
                                       bm.as<uint32_t>(), k_start, nbits, p, t, c);
                hcheck(hipGetLastError(), "strike");
            });
        }
        std::vector<uint32_t> h(nwords);
        hcheck(hipMemcpyAsync(h.data(), bm.p, nwords * 4, hipMemcpyDeviceToHost, st), "D2H bitmap");
        hcheck(hipStreamSynchronize(st), "sieve sync");
        size_t total = 0;
        for (uint64_t w = 0; w < nwords; ++w) {
            uint32_t v = h[w];
            while (v) {
                const int b = __builtin_ctz(v);
                v &= v - 1;
                const uint64_t j = w * 32 + (uint64_t)b;
                if (j >= nbits) break;
                const uint64_t k = k_start + j;
                if ((k & 1) == 0) continue;  // the reference keeps odd k only (P/presieve_core.cpp:78-83)
                if (total < cap) out_ks[total] = k;
                ++total;
            }
        }
        *n_out = total;
    } catch (const std::exception& e) {
        if (st) (void)hipStreamDestroy(st);
        return sfail(PL_ERR_HIP, e.what());
    }
    (void)hipStreamDestroy(st);
    return PL_OK;
}

int pl_sieve_list(const uint64_t* ks, size_t count, uint32_t n, uint64_t pstart, uint64_t pmax, int device,
                  uint8_t* keep)
{
    if ((!ks || !keep) && count) return sfail(PL_ERR_ARG, "NULL argument");
    if (pmax >= (1ull << 62)) return sfail(PL_ERR_ARG, "pmax must be below 2^62");
    for (size_t i = 0; i < count; ++i) keep[i] = 1;
    if (count == 0 || pmax < 3 || pstart > pmax) return PL_OK;
    int rc = select_device(device);
    if (rc) return rc;
    hipStream_t st = nullptr;
    try {
        hcheck(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        DevBuf dk(count * 8), dkeep(count);
        hcheck(hipMemcpyAsync(dk.p, ks, count * 8, hipMemcpyHostToDevice, st), "H2D ks");
        hcheck(hipMemcpyAsync(dkeep.p, keep, count, hipMemcpyHostToDevice, st), "H2D keep");
        for_each_prime_batch(pstart, pmax, n, st, [&](const uint64_t* p, const uint64_t* t, uint64_t c, bool) {
            for (size_t c0 = 0; c0 < count; c0 += kMaxList) {
                const uint32_t nk = (uint32_t)std::min<size_t>(kMaxList, count - c0);
                hipLaunchKernelGGL(k_strike_list, dim3((unsigned)((c + 255) / 256)), dim3(256), 0, st,
                                   dk.as<uint64_t>() + c0, nk, dkeep.as<uint8_t>() + c0, p, t, c);
                hcheck(hipGetLastError(), "k_strike_list");
            }
        });
        hcheck(hipMemcpyAsync(keep, dkeep.p, count, hipMemcpyDeviceToHost, st), "D2H keep");
        hcheck(hipStreamSynchronize(st), "sieve sync");
    } catch (const std::exception& e) {
        if (st) (void)hipStreamDestroy(st);
        return sfail(PL_ERR_HIP, e.what());
    }
    (void)hipStreamDestroy(st);
    return PL_OK;
}

}  // extern "C"
