// engine.hip -- CDNA4 (gfx950) kernels for the modular squaring of a Proth
// residue, N = k*2^n + 1, kept on the device for the whole test.
//
// Representation.  The residue is x = X_lo + X_hi*2^n where X_lo has L
// balanced digits x_m at bit positions e_m = ceil(n*m/L) (variable base,
// widths b_m in {floor(n/L), ceil(n/L)}) and X_hi is a small signed scalar
// (|X_hi| <~ k/2).  Digits live in the device array z as weighted complex
// values z_m = x_m * a_m * w^m with the IBDWT weight a_m = 2^(e_m - n*m/L)
// and the right-angle weight w = exp(i*pi/(2L)).
//
// Squaring.  With u^L == i, the complex cyclic convolution of z is the
// negacyclic convolution of length 2L of the real digits, whose low half
// (coefficient m) comes back in Re(z_m) and high half (coefficient m+L, bit
// position n+e_m) in Im(z_m).  X_lo has only L digits, so nothing wraps and
// the product is the exact linear one.  X_hi*2^n maps to the constant i*X_hi,
// so the pointwise step Z <- Z*(Z + 2i*X_hi) adds 2*X_hi*X_lo*2^n; the last
// term X_hi^2*2^(2n) is added exactly in the carry step.
//
// Reduction.  k*2^n == -1 (mod N) gives 2^-n == -k, so for T = T_lo + 2^n*T_hi
// one fold T*2^-n == T_hi - k*T_lo (mod N) needs no division.  Two folds per
// squaring keep |x| below ~k*2^(n-1); the residue is therefore held in the
// Montgomery form y = x*2^(2n) mod N (conversions happen on the host, at the
// start and end of a test only).  The three carry passes (T_lo, the first fold
// S, the second fold R) run per run of consecutive digits in int64; each run
// publishes its own three carry-outs, computed from its own digits only, and
// the next run folds them into its first digits (one hop, no chain).
//
// FFT.  L = R*C, four-step on the R x C row-major array (digit m at r*C + c):
// R-point DFTs down the columns with the W_L^(c*rho) twiddle, then C-point
// DFTs along rows; the spectrum stays in that transposed order.  One squaring
// is two kernels in steady state:
//   k_outer : inverse column DFT -> unweight/round/carry/fold -> weight ->
//             forward column DFT (block of NT columns, runs of NT digits)
//   k_row   : forward row DFT -> Z*(Z + 2i*X_hi) -> inverse row DFT
// k_fwd_col / k_inv_col / k_carry / k_fixup are the unfused pieces used to
// enter and leave the steady state (and for lengths the fused path does not
// cover).
#include "engine.h"
#include "fft_lds.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace pleng {

void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct Params {
    double2* z;
    const double2* tw;      // W_4096^i, i < 4096
    const double2* twl_lo;  // W_L^j, j < 2^s
    const double2* twl_hi;  // W_L^(j*2^s)
    const double* pw_lo;    // 2^(j/L)
    const double* pw_hi;    // 2^(j*2^s/L)
    const double* ipw_lo;   // 2^(-j/L)
    const double* ipw_hi;   // 2^(-j*2^s/L) / L
    const double2* om_lo;   // exp(i*pi*j/(2L))
    const double2* om_hi;   // exp(i*pi*j*2^s/(2L))
    long long* carries;     // flat carry kernel: 3 per block (own carry-outs of its last run)
    long long* runc;        // outer kernel: 3 per run (block-major, then row)
    unsigned* flags;        // outer kernel: per-block publish counters
    unsigned* ticket;       // outer kernel: logical block ticket counter
    long long* xhi;
    unsigned long long* err;
    int* status;
    unsigned long long n;
    long long k;
    unsigned L, logL, R, C, s, smask;
};

struct DeviceTables {
    double2* tw;
    double2 *twl_lo, *twl_hi, *om_lo, *om_hi;
    double *pw_lo, *pw_hi, *ipw_lo, *ipw_hi;
};

// ------------------------------------------------------------- helpers

__device__ __forceinline__ double2 twL(const Params& p, unsigned j)  // W_L^j, j < L
{
    return cmul(p.twl_hi[j >> p.s], p.twl_lo[j & p.smask]);
}

__device__ __forceinline__ unsigned long long epos(const Params& p, unsigned m)
{
    return (p.n * m + p.L - 1) >> p.logL;
}
__device__ __forceinline__ int width(const Params& p, unsigned m) { return (int)(epos(p, m + 1) - epos(p, m)); }
__device__ __forceinline__ unsigned fidx(const Params& p, unsigned m)
{
    return (unsigned)((0ull - p.n * (unsigned long long)m) & (p.L - 1));
}
__device__ __forceinline__ double2 omega(const Params& p, unsigned m)
{
    return cmul(p.om_hi[m >> p.s], p.om_lo[m & p.smask]);
}
// a_m * w^m
__device__ __forceinline__ double2 fwd_weight(const Params& p, unsigned m)
{
    unsigned f = fidx(p, m);
    double a = p.pw_hi[f >> p.s] * p.pw_lo[f & p.smask];
    double2 w = omega(p, m);
    return make_double2(w.x * a, w.y * a);
}
// conj(w^m) / (L * a_m)
__device__ __forceinline__ double2 inv_weight(const Params& p, unsigned m)
{
    unsigned f = fidx(p, m);
    double ia = p.ipw_hi[f >> p.s] * p.ipw_lo[f & p.smask];
    double2 w = omega(p, m);
    return make_double2(w.x * ia, -w.y * ia);
}

// v = d + q*2^b with d in [-2^(b-1), 2^(b-1)); returns d, sets carry = q.
__device__ __forceinline__ long long bal(long long v, int b, long long& carry)
{
    long long q = (v + (1ll << (b - 1))) >> b;
    carry = q;
    return v - q * (1ll << b);
}

// One digit through the three carry passes: T_lo, S = T_hi - k*T_lo,
// R = S_hi - k*S_lo.  Returns the digit of R; advances the carries.
__device__ __forceinline__ long long carry3(double tr, double hr, long long k, int b, long long& q1, long long& q2,
                                           long long& q3)
{
    const long long d1 = bal((long long)tr + q1, b, q1);
    const long long d2 = bal((long long)hr - k * d1 + q2, b, q2);
    return bal(-k * d2 + q3, b, q3);
}

// Adjustment a run receives from its predecessor: c2in is the first fold's
// carry arriving at the run's first bit (it already includes -k times the
// predecessor's T_lo carry); it is spread over balanced digits delta_j, which
// enter the second fold as -k*delta_j; extra is the second fold's carry-in at
// the same bit.  The sum is normalised into W balanced digits nu_j added to
// the run's digits.  Sets status when W digits do not absorb it.
template <int W>
__device__ __forceinline__ void ripple(const Params& p, unsigned m0, long long c2in, long long extra, long long (&nu)[W])
{
    long long rem = c2in;
    long long a[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        long long c;
        const long long d = bal(rem, width(p, m0 + j), c);
        a[j] = -p.k * d;
        rem = c;
    }
    a[0] += extra;
    long long c = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        long long cc;
        nu[j] = bal(a[j] + c, width(p, m0 + j), cc);
        c = cc;
    }
    if (rem != 0 || c != 0) atomicOr(p.status, 1);
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// ------------------------------------------------------------- FFT kernels

template <int R, int NT, int TPB>
__global__ __launch_bounds__(TPB) void k_fwd_col(Params p)
{
    __shared__ double2 lds[LdsLayout<R, NT>::kSlots];
    const unsigned c0 = blockIdx.x * NT;
    const double2* __restrict__ z = p.z;
    dft<R, NT, TPB, false>(
        lds, p.tw, [&](int tr, int idx) { return z[(size_t)idx * p.C + c0 + tr]; },
        [&](int tr, int idx, double2 v) {
            const unsigned c = c0 + tr;
            p.z[(size_t)idx * p.C + c] = cmul(v, twL(p, c * (unsigned)idx));
        });
}

template <int R, int NT, int TPB>
__global__ __launch_bounds__(TPB) void k_inv_col(Params p)
{
    __shared__ double2 lds[LdsLayout<R, NT>::kSlots];
    const unsigned c0 = blockIdx.x * NT;
    const double2* __restrict__ z = p.z;
    dft<R, NT, TPB, true>(
        lds, p.tw,
        [&](int tr, int idx) {
            const unsigned c = c0 + tr;
            return cmulc(z[(size_t)idx * p.C + c], twL(p, c * (unsigned)idx));
        },
        [&](int tr, int idx, double2 v) { p.z[(size_t)idx * p.C + c0 + tr] = v; });
}

// Row pass: NT rows; forward C-point DFT, Z <- Z*(Z + 2i*X_hi) folded into
// the last forward stage, inverse C-point DFT.
template <int C, int NT, int TPB>
__global__ __launch_bounds__(TPB) void k_row(Params p)
{
    using Lay = LdsLayout<C, NT>;
    __shared__ double2 lds[Lay::kSlots];
    const size_t base = (size_t)blockIdx.x * NT * C;
    const double2* __restrict__ z = p.z;
    const double x2 = 2.0 * (double)(*p.xhi);
    dft<C, NT, TPB, false, false, true>(
        lds, p.tw, [&](int tr, int idx) { return z[base + (size_t)tr * C + idx]; },
        [&](int tr, int idx, double2 X) { lds[Lay::at(tr, idx)] = cmul(X, make_double2(X.x, X.y + x2)); });
    __syncthreads();
    dft<C, NT, TPB, true, true>(
        lds, p.tw, [&](int tr, int idx) { return lds[Lay::at(tr, idx)]; },
        [&](int tr, int idx, double2 v) { p.z[base + (size_t)tr * C + idx] = v; });
}

// ------------------------------------------------------------- flat carry

// Unweight, round (with max round-off), three carry passes per run of RUN
// digits, in-block run hand-off, re-weight.  The first run of each block is
// finished by k_fixup.
template <int RUN, int TPB>
__global__ __launch_bounds__(TPB) void k_carry(Params p)
{
    __shared__ long long sq[3][TPB];
    const unsigned tid = threadIdx.x;
    const unsigned m0 = (blockIdx.x * TPB + tid) * RUN;
    const long long k = p.k;
    long long x[RUN];
    long long q1 = 0, q2 = 0, q3 = 0;
    double emax = 0.0;
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
        const unsigned m = m0 + j;
        const double2 v = cmul(p.z[m], inv_weight(p, m));
        const double tr = rint(v.x), hr = rint(v.y);
        emax = fmax(emax, fmax(fabs(v.x - tr), fabs(v.y - hr)));
        x[j] = carry3(tr, hr, k, width(p, m), q1, q2, q3);
    }
    sq[0][tid] = q1;
    sq[1][tid] = q2;
    sq[2][tid] = q3;
    __syncthreads();
    if (tid > 0) {
        long long nu[RUN];
        ripple<RUN>(p, m0, sq[1][tid - 1] - k * sq[0][tid - 1], sq[2][tid - 1], nu);
#pragma unroll
        for (int j = 0; j < RUN; ++j) x[j] += nu[j];
    }
    if (tid == TPB - 1) {
        p.carries[3 * blockIdx.x + 0] = q1;
        p.carries[3 * blockIdx.x + 1] = q2;
        p.carries[3 * blockIdx.x + 2] = q3;
    }
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
        const unsigned m = m0 + j;
        const double2 w = fwd_weight(p, m);
        const double xd = (double)x[j];
        p.z[m] = make_double2(xd * w.x, xd * w.y);
    }
    emax = wave_max(emax);
    if ((tid & 63) == 0) atomicMax(p.err, (unsigned long long)__double_as_longlong(emax));
}

// Block boundaries: the first run of block w takes the carries of block
// w-1's last run.  Run 0 takes the top: +c1_top (T_lo's carry at bit n is
// part of T_hi) and c2_top + X_hi^2 (the second fold's top scalar); the new
// X_hi is the top run's third carry.
template <int RUN, int CTPB>
__global__ void k_fixup(Params p, unsigned nblocks)
{
    const unsigned w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nblocks) return;
    const unsigned m0 = w * CTPB * RUN;
    const long long k = p.k;
    long long c2in, extra;
    if (w > 0) {
        const long long* c = p.carries + 3 * (w - 1);
        c2in = c[1] - k * c[0];
        extra = c[2];
    } else {
        const long long* c = p.carries + 3 * (nblocks - 1);
        const long long xh = *p.xhi;
        c2in = c[0];
        extra = c[1] + xh * xh;
    }
    long long nu[RUN];
    ripple<RUN>(p, m0, c2in, extra, nu);
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
        if (nu[j] == 0) continue;
        const unsigned m = m0 + j;
        const double2 v = cmul(p.z[m], inv_weight(p, m));
        const double xd = rint(v.x * (double)p.L) + (double)nu[j];
        const double2 wt = fwd_weight(p, m);
        p.z[m] = make_double2(xd * wt.x, xd * wt.y);
    }
    if (w == 0) *p.xhi = p.carries[3 * (nblocks - 1) + 2];
}

// ------------------------------------------------------------- fused outer pass

__device__ __forceinline__ unsigned ld_acquire(const unsigned* f)
{
    return __hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

// Block b owns columns [b*NT, b*NT+NT): for each row r its digits
// r*C + b*NT .. +NT-1 form one run.  The predecessor of run (b, r) is
// (b-1, r), or (nb-1, r-1) for b = 0; run (0, 0) takes the top rule.
template <int R, int NT, int TPB>
__global__ __launch_bounds__(TPB) void k_outer(Params p)
{
    using Lay = LdsLayout<R, NT>;
    constexpr int RPT = R / TPB;  // runs per thread
    static_assert(R % TPB == 0, "runs must split evenly over threads");
    __shared__ double2 lds[Lay::kSlots];
    __shared__ unsigned s_blk;
    __shared__ double s_err[TPB / 64];

    const unsigned nb = p.C / NT;
    if (threadIdx.x == 0) s_blk = atomicAdd(p.ticket, 1u) % nb;
    __syncthreads();
    const unsigned b = s_blk;
    const unsigned c0 = b * NT;
    const long long k = p.k;

    // 1. inverse column DFT into LDS (natural row order)
    const double2* __restrict__ z = p.z;
    dft<R, NT, TPB, true, false, true>(
        lds, p.tw,
        [&](int tr, int idx) {
            const unsigned c = c0 + tr;
            return cmulc(z[(size_t)idx * p.C + c], twL(p, c * (unsigned)idx));
        },
        [&](int tr, int idx, double2 v) { lds[Lay::at(tr, idx)] = v; });
    __syncthreads();

    // 2. own carries of each run
    long long x[RPT][NT];
    double emax = 0.0;
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned r = threadIdx.x + u * TPB;
        const unsigned m0 = r * p.C + c0;
        long long q1 = 0, q2 = 0, q3 = 0;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const unsigned m = m0 + j;
            const double2 v = cmul(lds[Lay::at(j, r)], inv_weight(p, m));
            const double tr = rint(v.x), hr = rint(v.y);
            emax = fmax(emax, fmax(fabs(v.x - tr), fabs(v.y - hr)));
            x[u][j] = carry3(tr, hr, k, width(p, m), q1, q2, q3);
        }
        long long* rc = p.runc + 3 * ((size_t)b * R + r);
        rc[0] = q1;
        rc[1] = q2;
        rc[2] = q3;
    }
    // 3. publish, then wait for the predecessor block
    __threadfence();
    __syncthreads();
    const unsigned pred = b == 0 ? nb - 1 : b - 1;
    if (threadIdx.x == 0) {
        const unsigned mine = __hip_atomic_fetch_add(p.flags + b, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) + 1;
        // bounded wait: every block publishes before it waits, so the bound
        // is never reached in a healthy launch; if it is, the status word
        // reports it and the kernel still exits
        unsigned spins = 0;
        while (ld_acquire(p.flags + pred) < mine) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > (1u << 22)) {
                atomicOr(p.status, 4);
                break;
            }
        }
    }
    __syncthreads();
    __threadfence();

    // 4. fold the predecessor's carries into each run's first digits (only
    //    block 0 reads X_hi: it is also the block that replaces it)
    const long long xh = b == 0 ? *p.xhi : 0;
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned r = threadIdx.x + u * TPB;
        const unsigned m0 = r * p.C + c0;
        long long c2in, extra;
        if (b == 0 && r == 0) {
            const long long* rc = p.runc + 3 * ((size_t)(nb - 1) * R + (R - 1));
            c2in = __hip_atomic_load(rc + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            extra = __hip_atomic_load(rc + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + xh * xh;
        } else {
            const size_t src = b == 0 ? (size_t)(nb - 1) * R + (r - 1) : (size_t)(b - 1) * R + r;
            const long long* rc = p.runc + 3 * src;
            const long long c1 = __hip_atomic_load(rc + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const long long c2 = __hip_atomic_load(rc + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            c2in = c2 - k * c1;
            extra = __hip_atomic_load(rc + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        long long nu[NT];
        ripple<NT>(p, m0, c2in, extra, nu);
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const unsigned m = m0 + j;
            const double2 w = fwd_weight(p, m);
            const double xd = (double)(x[u][j] + nu[j]);
            lds[Lay::at(j, r)] = make_double2(xd * w.x, xd * w.y);
        }
    }
    if (b == 0 && threadIdx.x == 0) {
        const long long* rc = p.runc + 3 * ((size_t)(nb - 1) * R + (R - 1));
        *p.xhi = __hip_atomic_load(rc + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    emax = wave_max(emax);
    if ((threadIdx.x & 63) == 0) s_err[threadIdx.x / 64] = emax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = s_err[0];
#pragma unroll
        for (int i = 1; i < TPB / 64; ++i) e = fmax(e, s_err[i]);
        atomicMax(p.err, (unsigned long long)__double_as_longlong(e));
    }

    // 5. forward column DFT from LDS, twiddle, store
    dft<R, NT, TPB, false, true, false>(
        lds, p.tw, [&](int tr, int idx) { return lds[Lay::at(tr, idx)]; },
        [&](int tr, int idx, double2 v) {
            const unsigned c = c0 + tr;
            p.z[(size_t)idx * p.C + c] = cmul(v, twL(p, c * (unsigned)idx));
        });
}

__global__ void k_set_digits(Params p, const int* __restrict__ d)
{
    const unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= p.L) return;
    const double2 w = fwd_weight(p, m);
    const double x = (double)d[m];
    p.z[m] = make_double2(x * w.x, x * w.y);
}

__global__ void k_get_digits(Params p, int* __restrict__ d)
{
    const unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= p.L) return;
    const double2 v = cmul(p.z[m], inv_weight(p, m));
    d[m] = (int)llrint(v.x * (double)p.L);
}

__global__ void k_fill_u64(unsigned long long* a, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = 0ull;
}

// ------------------------------------------------------------- host side

static constexpr int kRun = 16;
static constexpr int kCarryTpb = 64;
static constexpr int kOuterNT = 8;

double max_bits_per_digit(uint32_t logL)
{
    if (logL <= 14) return 18.5;
    if (logL <= 17) return 18.0;
    if (logL <= 20) return 17.5;
    return 17.0;
}

static unsigned bitlen64(unsigned long long x)
{
    unsigned b = 0;
    while (x) {
        ++b;
        x >>= 1;
    }
    return b;
}

// Digits a carry-in can reach (see ripple): bounds the carries from the
// worst-case coefficient magnitude L*2^(2b+1).
static int ripple_digits(const Plan& pl)
{
    const unsigned bmin = pl.n >> pl.logL;
    const unsigned bmax = bmin + 1;
    const unsigned kb = bitlen64(pl.k);
    const unsigned c2bits = bmax + 4 + pl.logL + kb + 1;
    const unsigned a3bits = c2bits + kb + 1;
    return (int)((a3bits + 1 + bmin - 1) / bmin) + 1;
}

bool choose_plan(uint32_t n, uint64_t k, uint32_t force_L, Plan& out, std::string& why)
{
    if (k == 0 || (k & 1) == 0) {
        why = "k must be odd";
        return false;
    }
    if (k >= (1ull << 32)) {
        why = "k must be below 2^32";
        return false;
    }
    uint32_t logL = 0;
    if (force_L) {
        if (force_L & (force_L - 1)) {
            why = "fft_len must be a power of two";
            return false;
        }
        while ((1u << logL) < force_L) ++logL;
        if (logL < 10 || logL > 22) {
            why = "fft_len must be in [2^10, 2^22]";
            return false;
        }
        if ((double)n / (double)force_L > max_bits_per_digit(logL) + 1.0) {
            why = "fft_len too short for n (round-off)";
            return false;
        }
    } else {
        for (logL = 10; logL <= 22; ++logL)
            if ((double)n / (double)(1u << logL) <= max_bits_per_digit(logL)) break;
        if (logL > 22) {
            why = "n too large (needs fft_len > 2^22)";
            return false;
        }
    }
    const uint32_t L = 1u << logL;
    if (n < 8u * L) {
        why = "n too small: needs at least 8 bits per digit (n >= 8*fft_len)";
        return false;
    }
    out.n = n;
    out.k = k;
    out.L = L;
    out.logL = logL;
    const uint32_t logR = std::max<uint32_t>(logL > 12 ? logL - 12 : 0, std::min<uint32_t>(9, logL / 2));
    out.R = 1u << logR;
    out.C = L / out.R;
    out.s = (logL + 1) / 2;
    out.fused = false;
    return true;
}

static Params make_params(const Plan& pl, double2* z, const DeviceTables* t, long long* carries, long long* runc,
                          unsigned* flags, unsigned* ticket, long long* xhi, unsigned long long* err, int* status)
{
    Params p;
    p.z = z;
    p.tw = t->tw;
    p.twl_lo = t->twl_lo;
    p.twl_hi = t->twl_hi;
    p.pw_lo = t->pw_lo;
    p.pw_hi = t->pw_hi;
    p.ipw_lo = t->ipw_lo;
    p.ipw_hi = t->ipw_hi;
    p.om_lo = t->om_lo;
    p.om_hi = t->om_hi;
    p.carries = carries;
    p.runc = runc;
    p.flags = flags;
    p.ticket = ticket;
    p.xhi = xhi;
    p.err = err;
    p.status = status;
    p.n = pl.n;
    p.k = (long long)pl.k;
    p.L = pl.L;
    p.logL = pl.logL;
    p.R = pl.R;
    p.C = pl.C;
    p.s = pl.s;
    p.smask = (1u << pl.s) - 1;
    return p;
}

template <int R>
struct ColCfg {
    static constexpr int NT = R <= 1024 ? 8 : 4;
    static constexpr int TPB = (NT * R / 16) < 64 ? 64 : (NT * R / 16);
};

template <int C>
struct RowCfg {
    static constexpr int NT = (1024 / C) > 1 ? (1024 / C) : 1;
    static constexpr int TPB = (NT * C / 16) < 64 ? 64 : (NT * C / 16);
};

enum class Kind { FwdCol, InvCol, Outer };

template <int R>
static void launch_col(Kind kind, const Params& p, hipStream_t st)
{
    constexpr int NT = ColCfg<R>::NT, TPB = ColCfg<R>::TPB;
    dim3 grid(p.C / NT);
    switch (kind) {
    case Kind::FwdCol: hipLaunchKernelGGL((k_fwd_col<R, NT, TPB>), grid, dim3(TPB), 0, st, p); break;
    case Kind::InvCol: hipLaunchKernelGGL((k_inv_col<R, NT, TPB>), grid, dim3(TPB), 0, st, p); break;
    case Kind::Outer:
        if constexpr (R >= 64 && R <= 1024 && NT == kOuterNT && R % TPB == 0)
            hipLaunchKernelGGL((k_outer<R, NT, TPB>), grid, dim3(TPB), 0, st, p);
        else
            throw std::runtime_error("no fused outer kernel for this column length");
        break;
    }
}

template <int C>
static void launch_row(const Params& p, hipStream_t st)
{
    constexpr int NT = RowCfg<C>::NT, TPB = RowCfg<C>::TPB;
    dim3 grid(p.R / NT);
    hipLaunchKernelGGL((k_row<C, NT, TPB>), grid, dim3(TPB), 0, st, p);
}

static void dispatch_col(Kind kind, const Params& p, hipStream_t st)
{
    switch (p.R) {
    case 32: launch_col<32>(kind, p, st); break;
    case 64: launch_col<64>(kind, p, st); break;
    case 128: launch_col<128>(kind, p, st); break;
    case 256: launch_col<256>(kind, p, st); break;
    case 512: launch_col<512>(kind, p, st); break;
    case 1024: launch_col<1024>(kind, p, st); break;
    default: throw std::runtime_error("unsupported column length");
    }
}

static void dispatch_row(const Params& p, hipStream_t st)
{
    switch (p.C) {
    case 32: launch_row<32>(p, st); break;
    case 64: launch_row<64>(p, st); break;
    case 128: launch_row<128>(p, st); break;
    case 256: launch_row<256>(p, st); break;
    case 512: launch_row<512>(p, st); break;
    case 1024: launch_row<1024>(p, st); break;
    case 2048: launch_row<2048>(p, st); break;
    case 4096: launch_row<4096>(p, st); break;
    default: throw std::runtime_error("unsupported row length");
    }
}

template <int R>
static int outer_max_blocks_per_cu()
{
    constexpr int NT = ColCfg<R>::NT, TPB = ColCfg<R>::TPB;
    if constexpr (R >= 64 && R <= 1024 && NT == kOuterNT && R % TPB == 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_outer<R, NT, TPB>, TPB, 0) != hipSuccess) return 0;
        return nb;
    }
    return 0;
}

static int outer_blocks_per_cu(uint32_t R)
{
    switch (R) {
    case 64: return outer_max_blocks_per_cu<64>();
    case 128: return outer_max_blocks_per_cu<128>();
    case 256: return outer_max_blocks_per_cu<256>();
    case 512: return outer_max_blocks_per_cu<512>();
    case 1024: return outer_max_blocks_per_cu<1024>();
    default: return 0;
    }
}

// Correctly rounded tables from long double (64-bit mantissa) arithmetic.
static void host_tables(const Plan& pl, std::vector<double2>& tw, std::vector<double2>& twl_lo,
                        std::vector<double2>& twl_hi, std::vector<double2>& om_lo, std::vector<double2>& om_hi,
                        std::vector<double>& pw_lo, std::vector<double>& pw_hi, std::vector<double>& ipw_lo,
                        std::vector<double>& ipw_hi)
{
    const long double PI = 3.141592653589793238462643383279502884L;
    const uint32_t L = pl.L, S = 1u << pl.s, H = L / S;
    tw.resize(4096);
    for (uint32_t i = 0; i < 4096; ++i) {
        long double a = -2.0L * PI * (long double)i / 4096.0L;
        tw[i] = make_double2((double)cosl(a), (double)sinl(a));
    }
    twl_lo.resize(S);
    om_lo.resize(S);
    pw_lo.resize(S);
    ipw_lo.resize(S);
    for (uint32_t j = 0; j < S; ++j) {
        long double a = -2.0L * PI * (long double)j / (long double)L;
        twl_lo[j] = make_double2((double)cosl(a), (double)sinl(a));
        long double o = PI * (long double)j / (2.0L * (long double)L);
        om_lo[j] = make_double2((double)cosl(o), (double)sinl(o));
        pw_lo[j] = (double)exp2l((long double)j / (long double)L);
        ipw_lo[j] = (double)exp2l(-(long double)j / (long double)L);
    }
    twl_hi.resize(H);
    om_hi.resize(H);
    pw_hi.resize(H);
    ipw_hi.resize(H);
    for (uint32_t j = 0; j < H; ++j) {
        long double jj = (long double)j * (long double)S;
        long double a = -2.0L * PI * jj / (long double)L;
        twl_hi[j] = make_double2((double)cosl(a), (double)sinl(a));
        long double o = PI * jj / (2.0L * (long double)L);
        om_hi[j] = make_double2((double)cosl(o), (double)sinl(o));
        pw_hi[j] = (double)exp2l(jj / (long double)L);
        ipw_hi[j] = (double)(exp2l(-jj / (long double)L) / (long double)L);
    }
}

Engine::Engine(const Plan& p, int device, bool allow_fused) : plan_(p), device_(device)
{
    check(hipSetDevice(device), "hipSetDevice");
    check(hipStreamCreate(&stream_), "hipStreamCreate");
    std::vector<double2> tw, twl_lo, twl_hi, om_lo, om_hi;
    std::vector<double> pw_lo, pw_hi, ipw_lo, ipw_hi;
    host_tables(p, tw, twl_lo, twl_hi, om_lo, om_hi, pw_lo, pw_hi, ipw_lo, ipw_hi);
    carry_blocks_ = p.L / (kRun * kCarryTpb);

    // The fused pass needs: a kernel for this column length, runs (NT digits)
    // long enough for any carry-in to settle, and every block resident at
    // once (blocks spin on their predecessor's flag).
    if (allow_fused && ripple_digits(p) <= kOuterNT) {
        int bpc = outer_blocks_per_cu(p.R);
        int cus = 0;
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "CU count");
        const unsigned nb = p.C / kOuterNT;
        if (bpc > 0 && (long long)bpc * cus >= (long long)nb) plan_.fused = true;
    }
    const unsigned outer_blocks = p.C / kOuterNT;

    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t off = 0;
    const size_t o_z = off; off += al(sizeof(double2) * p.L);
    const size_t o_car = off; off += al(sizeof(long long) * 3 * carry_blocks_);
    const size_t o_runc = off; off += al(plan_.fused ? sizeof(long long) * 3 * (size_t)p.R * outer_blocks : 0);
    const size_t o_flags = off; off += al(sizeof(unsigned) * (outer_blocks + 1));
    const size_t o_xhi = off; off += al(sizeof(long long));
    const size_t o_err = off; off += al(sizeof(unsigned long long) * kGraphChunk);
    const size_t o_st = off; off += al(sizeof(int));
    const size_t o_dig = off; off += al(sizeof(int32_t) * p.L);
    const size_t zero_end = off;
    const size_t o_tw = off; off += al(sizeof(double2) * tw.size());
    const size_t o_tlo = off; off += al(sizeof(double2) * twl_lo.size());
    const size_t o_thi = off; off += al(sizeof(double2) * twl_hi.size());
    const size_t o_olo = off; off += al(sizeof(double2) * om_lo.size());
    const size_t o_ohi = off; off += al(sizeof(double2) * om_hi.size());
    const size_t o_plo = off; off += al(sizeof(double) * pw_lo.size());
    const size_t o_phi = off; off += al(sizeof(double) * pw_hi.size());
    const size_t o_ilo = off; off += al(sizeof(double) * ipw_lo.size());
    const size_t o_ihi = off; off += al(sizeof(double) * ipw_hi.size());
    bytes_ = off;
    check(hipMalloc(&mem_, bytes_), "hipMalloc");
    char* b = static_cast<char*>(mem_);
    z_ = reinterpret_cast<double2*>(b + o_z);
    carries_ = reinterpret_cast<long long*>(b + o_car);
    runc_ = reinterpret_cast<long long*>(b + o_runc);
    flags_ = reinterpret_cast<unsigned*>(b + o_flags);
    ticket_ = flags_ + outer_blocks;
    xhi_ = reinterpret_cast<long long*>(b + o_xhi);
    err_ = reinterpret_cast<unsigned long long*>(b + o_err);
    status_ = reinterpret_cast<int*>(b + o_st);
    digits_dev_ = reinterpret_cast<int32_t*>(b + o_dig);
    tabs_ = new DeviceTables;
    tabs_->tw = reinterpret_cast<double2*>(b + o_tw);
    tabs_->twl_lo = reinterpret_cast<double2*>(b + o_tlo);
    tabs_->twl_hi = reinterpret_cast<double2*>(b + o_thi);
    tabs_->om_lo = reinterpret_cast<double2*>(b + o_olo);
    tabs_->om_hi = reinterpret_cast<double2*>(b + o_ohi);
    tabs_->pw_lo = reinterpret_cast<double*>(b + o_plo);
    tabs_->pw_hi = reinterpret_cast<double*>(b + o_phi);
    tabs_->ipw_lo = reinterpret_cast<double*>(b + o_ilo);
    tabs_->ipw_hi = reinterpret_cast<double*>(b + o_ihi);
    auto up = [&](void* dst, const void* src, size_t n) {
        check(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync H2D");
    };
    up(tabs_->tw, tw.data(), sizeof(double2) * tw.size());
    up(tabs_->twl_lo, twl_lo.data(), sizeof(double2) * twl_lo.size());
    up(tabs_->twl_hi, twl_hi.data(), sizeof(double2) * twl_hi.size());
    up(tabs_->om_lo, om_lo.data(), sizeof(double2) * om_lo.size());
    up(tabs_->om_hi, om_hi.data(), sizeof(double2) * om_hi.size());
    up(tabs_->pw_lo, pw_lo.data(), sizeof(double) * pw_lo.size());
    up(tabs_->pw_hi, pw_hi.data(), sizeof(double) * pw_hi.size());
    up(tabs_->ipw_lo, ipw_lo.data(), sizeof(double) * ipw_lo.size());
    up(tabs_->ipw_hi, ipw_hi.data(), sizeof(double) * ipw_hi.size());
    check(hipMemsetAsync(mem_, 0, zero_end, stream_), "hipMemsetAsync");
    check(hipStreamSynchronize(stream_), "init sync");
    err_host_.resize(kGraphChunk);
}

Engine::~Engine()
{
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (mem_) (void)hipFree(mem_);
    if (stream_) (void)hipStreamDestroy(stream_);
    delete tabs_;
}

Params Engine::params(uint64_t slot) const
{
    return make_params(plan_, z_, tabs_, carries_, runc_, flags_, ticket_, xhi_, err_ + slot, status_);
}

void Engine::upload(const std::vector<int32_t>& digits, int64_t xhi)
{
    check(hipSetDevice(device_), "hipSetDevice");
    if (digits.size() != plan_.L) throw std::runtime_error("digit count mismatch");
    long long xh = xhi;
    check(hipMemcpyAsync(digits_dev_, digits.data(), sizeof(int32_t) * plan_.L, hipMemcpyHostToDevice, stream_),
          "upload digits");
    check(hipMemcpyAsync(xhi_, &xh, sizeof(long long), hipMemcpyHostToDevice, stream_), "upload xhi");
    Params p = params(0);
    hipLaunchKernelGGL(k_set_digits, dim3((plan_.L + 255) / 256), dim3(256), 0, stream_, p, digits_dev_);
    check(hipGetLastError(), "k_set_digits");
    check(hipStreamSynchronize(stream_), "upload sync");
}

void Engine::download(std::vector<int32_t>& digits, int64_t& xhi)
{
    check(hipSetDevice(device_), "hipSetDevice");
    Params p = params(0);
    hipLaunchKernelGGL(k_get_digits, dim3((plan_.L + 255) / 256), dim3(256), 0, stream_, p, digits_dev_);
    check(hipGetLastError(), "k_get_digits");
    digits.resize(plan_.L);
    long long xh = 0;
    check(hipMemcpyAsync(digits.data(), digits_dev_, sizeof(int32_t) * plan_.L, hipMemcpyDeviceToHost, stream_),
          "download digits");
    check(hipMemcpyAsync(&xh, xhi_, sizeof(long long), hipMemcpyDeviceToHost, stream_), "download xhi");
    check(hipStreamSynchronize(stream_), "download sync");
    xhi = xh;
}

// g squarings from the digit state back to the digit state; iteration i's
// round-off lands in err slot i.
void Engine::enqueue_chunk(int g)
{
    const Params p0 = params(0);
    dispatch_col(Kind::FwdCol, p0, stream_);
    dispatch_row(p0, stream_);
    for (int i = 0; i + 1 < g; ++i) {
        const Params p = params((uint64_t)i);
        if (plan_.fused) {
            dispatch_col(Kind::Outer, p, stream_);
        } else {
            dispatch_col(Kind::InvCol, p, stream_);
            hipLaunchKernelGGL((k_carry<kRun, kCarryTpb>), dim3(carry_blocks_), dim3(kCarryTpb), 0, stream_, p);
            hipLaunchKernelGGL((k_fixup<kRun, kCarryTpb>), dim3((carry_blocks_ + 63) / 64), dim3(64), 0, stream_, p,
                               carry_blocks_);
            dispatch_col(Kind::FwdCol, p, stream_);
        }
        dispatch_row(p, stream_);
    }
    const Params pl = params((uint64_t)(g - 1));
    dispatch_col(Kind::InvCol, pl, stream_);
    hipLaunchKernelGGL((k_carry<kRun, kCarryTpb>), dim3(carry_blocks_), dim3(kCarryTpb), 0, stream_, pl);
    hipLaunchKernelGGL((k_fixup<kRun, kCarryTpb>), dim3((carry_blocks_ + 63) / 64), dim3(64), 0, stream_, pl,
                       carry_blocks_);
}

void Engine::build_graph(int iters)
{
    if (graph_exec_) {
        (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
    }
    if (graph_) {
        (void)hipGraphDestroy(graph_);
        graph_ = nullptr;
    }
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kGraphChunk), 0, stream_, err_, kGraphChunk);
    enqueue_chunk(iters);
    check(hipStreamEndCapture(stream_, &graph_), "hipStreamEndCapture");
    check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    graph_iters_ = iters;
}

double Engine::square(uint64_t iters, double* err_trace)
{
    check(hipSetDevice(device_), "hipSetDevice");
    double emax = 0.0;
    uint64_t done = 0;
    while (done < iters) {
        const int g = (int)std::min<uint64_t>(kGraphChunk, iters - done);
        if (g == kGraphChunk) {
            if (!graph_exec_) build_graph(kGraphChunk);
            check(hipGraphLaunch(graph_exec_, stream_), "hipGraphLaunch");
        } else {
            hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kGraphChunk), 0, stream_, err_, kGraphChunk);
            enqueue_chunk(g);
            check(hipGetLastError(), "square launch");
        }
        check(hipMemcpyAsync(err_host_.data(), err_, sizeof(unsigned long long) * g, hipMemcpyDeviceToHost, stream_),
              "err D2H");
        check(hipStreamSynchronize(stream_), "square sync");
        for (int i = 0; i < g; ++i) {
            double e;
            std::memcpy(&e, &err_host_[i], sizeof e);
            if (err_trace) err_trace[done + i] = e;
            if (e > emax) emax = e;
        }
        done += g;
    }
    return emax;
}

void Engine::kernel_times(int reps, double* ms)
{
    check(hipSetDevice(device_), "hipSetDevice");
    std::vector<int32_t> saved;
    int64_t saved_xhi = 0;
    download(saved, saved_xhi);
    const Params p = params(0);
    hipEvent_t a, b;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    for (int kind = 0; kind < kNumKernels; ++kind) {
        if (kind == 5 && !plan_.fused) {
            ms[kind] = 0.0;
            continue;
        }
        for (int pass = 0; pass < 2; ++pass) {
            const int cnt = pass == 0 ? 1 : reps;
            if (pass == 1) check(hipEventRecord(a, stream_), "hipEventRecord");
            for (int i = 0; i < cnt; ++i) {
                switch (kind) {
                case 0: dispatch_col(Kind::FwdCol, p, stream_); break;
                case 1: dispatch_row(p, stream_); break;
                case 2: dispatch_col(Kind::InvCol, p, stream_); break;
                case 3:
                    hipLaunchKernelGGL((k_carry<kRun, kCarryTpb>), dim3(carry_blocks_), dim3(kCarryTpb), 0, stream_, p);
                    break;
                case 4:
                    hipLaunchKernelGGL((k_fixup<kRun, kCarryTpb>), dim3((carry_blocks_ + 63) / 64), dim3(64), 0,
                                       stream_, p, carry_blocks_);
                    break;
                default: dispatch_col(Kind::Outer, p, stream_); break;
                }
            }
            if (pass == 1) check(hipEventRecord(b, stream_), "hipEventRecord");
        }
        check(hipEventSynchronize(b), "hipEventSynchronize");
        float t = 0;
        check(hipEventElapsedTime(&t, a, b), "hipEventElapsedTime");
        ms[kind] = (double)t / reps;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    // the timing launches leave arbitrary data and status behind
    check(hipMemsetAsync(status_, 0, sizeof(int), stream_), "status reset");
    upload(saved, saved_xhi);
}

int Engine::status()
{
    int s = 0;
    check(hipMemcpy(&s, status_, sizeof(int), hipMemcpyDeviceToHost), "status D2H");
    return s;
}

}  // namespace pleng
