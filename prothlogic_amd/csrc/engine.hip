// engine.hip -- CDNA4 (gfx950) kernels for one modular squaring of a Proth
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
// S, the second fold R) run per run of RUN consecutive digits in int64; each
// run publishes its own three carry-outs, computed from its own digits only,
// and the next run folds them into its first digits (one hop, no chain).
//
// FFT.  L = R*C, four-step: R-point DFTs down the columns (stride C) with the
// W_L^(c*rho) twiddle, then C-point DFTs along rows; the frequency data stays
// in that transposed order, the pointwise square runs in the row kernel
// between the forward and inverse row DFTs, and the inverse column kernel
// undoes the first step.  Each DFT is a Stockham radix-8/4/2 autosort staged
// in LDS.
#include "engine.h"

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
    long long* carries;     // 3 per carry block: own carry-outs of its last run
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

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

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

// ------------------------------------------------------------- butterflies

template <bool INV>
__device__ __forceinline__ double2 mul_mi(double2 a)  // forward: -i*a, inverse: +i*a
{
    return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft2(double2* v)
{
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(double2* v)
{
    double2 s0 = cadd(v[0], v[2]), s1 = csub(v[0], v[2]);
    double2 s2 = cadd(v[1], v[3]), s3 = csub(v[1], v[3]);
    double2 d = mul_mi<INV>(s3);
    v[0] = cadd(s0, s2);
    v[1] = cadd(s1, d);
    v[2] = csub(s0, s2);
    v[3] = csub(s1, d);
}

template <bool INV>
__device__ __forceinline__ void dft8(double2* v)
{
    constexpr double h = 0.70710678118654752440;
    constexpr double sg = INV ? -1.0 : 1.0;
    double2 e[4] = { v[0], v[2], v[4], v[6] };
    double2 o[4] = { v[1], v[3], v[5], v[7] };
    dft4<INV>(e);
    dft4<INV>(o);
    double2 t1 = make_double2(h * (o[1].x + sg * o[1].y), h * (o[1].y - sg * o[1].x));
    double2 t2 = mul_mi<INV>(o[2]);
    double2 t3 = make_double2(h * (-o[3].x + sg * o[3].y), h * (-o[3].y - sg * o[3].x));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], t1);
    v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2);
    v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3);
    v[7] = csub(e[3], t3);
}

template <int P, bool INV>
__device__ __forceinline__ void dftp(double2* v)
{
    if constexpr (P == 8) dft8<INV>(v);
    else if constexpr (P == 4) dft4<INV>(v);
    else dft2<INV>(v);
}

// One Stockham autosort stage of radix P over NT transforms of length N in
// LDS (transform t at lds + t*N).  Stage with span NS: butterfly j reads
// j + r*N/P, twiddles by W_(NS*P)^(r*(j mod NS)), writes
// (j/NS)*NS*P + (j mod NS) + r*NS.
template <int N, int NT, int TPB, int P, int NS, bool INV>
__device__ __forceinline__ void lds_stage(double2* lds, const double2* __restrict__ tw)
{
    constexpr int NB = NT * (N / P);
    constexpr int PER = (NB + TPB - 1) / TPB;
    constexpr int SCALE = 4096 / N;
    double2 v[PER][P];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int t = threadIdx.x + u * TPB;
        if ((NB % TPB == 0) || t < NB) {
            const int tr = t / (N / P), j = t % (N / P);
            const double2* base = lds + tr * N;
#pragma unroll
            for (int r = 0; r < P; ++r) v[u][r] = base[j + r * (N / P)];
            if constexpr (NS > 1) {
                const int jj = j % NS;
#pragma unroll
                for (int r = 1; r < P; ++r) {
                    double2 w = tw[(r * jj * (N / (NS * P))) * SCALE];
                    if (INV) w.y = -w.y;
                    v[u][r] = cmul(v[u][r], w);
                }
            }
            dftp<P, INV>(v[u]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int t = threadIdx.x + u * TPB;
        if ((NB % TPB == 0) || t < NB) {
            const int tr = t / (N / P), j = t % (N / P);
            double2* base = lds + tr * N;
            const int o = (j / NS) * NS * P + (j % NS);
#pragma unroll
            for (int r = 0; r < P; ++r) base[o + r * NS] = v[u][r];
        }
    }
    __syncthreads();
}

template <int N, int NT, int TPB, bool INV, int NS = 1>
__device__ __forceinline__ void lds_fft(double2* lds, const double2* __restrict__ tw)
{
    if constexpr (NS < N) {
        constexpr int REM = N / NS;
        constexpr int P = REM >= 8 ? 8 : (REM >= 4 ? 4 : 2);
        lds_stage<N, NT, TPB, P, NS, INV>(lds, tw);
        lds_fft<N, NT, TPB, INV, NS * P>(lds, tw);
    }
}

// ------------------------------------------------------------- FFT kernels

// Forward column pass: NT consecutive columns c0.. of the R x C array,
// R-point DFT down each column, then the W_L^(c*rho) twiddle.
template <int R, int NT, int TPB>
__global__ __launch_bounds__(TPB) void k_fwd_col(Params p)
{
    __shared__ double2 lds[NT * R];
    const unsigned c0 = blockIdx.x * NT;
    for (int t = threadIdx.x; t < NT * R; t += TPB) {
        const int r = t / NT, b = t % NT;
        lds[b * R + r] = p.z[(size_t)r * p.C + c0 + b];
    }
    __syncthreads();
    lds_fft<R, NT, TPB, false>(lds, p.tw);
    for (int t = threadIdx.x; t < NT * R; t += TPB) {
        const int rho = t / NT, b = t % NT;
        const unsigned c = c0 + b;
        p.z[(size_t)rho * p.C + c] = cmul(lds[b * R + rho], twL(p, c * (unsigned)rho));
    }
}

// Row pass: NT rows; forward C-point DFT, Z <- Z*(Z + 2i*X_hi), inverse.
template <int C, int NT, int TPB>
__global__ __launch_bounds__(TPB) void k_row(Params p)
{
    __shared__ double2 lds[NT * C];
    const size_t base = (size_t)blockIdx.x * NT * C;
    for (int t = threadIdx.x; t < NT * C; t += TPB) lds[t] = p.z[base + t];
    __syncthreads();
    lds_fft<C, NT, TPB, false>(lds, p.tw);
    const double x2 = 2.0 * (double)(*p.xhi);
    for (int t = threadIdx.x; t < NT * C; t += TPB) {
        double2 X = lds[t];
        lds[t] = cmul(X, make_double2(X.x, X.y + x2));
    }
    __syncthreads();
    lds_fft<C, NT, TPB, true>(lds, p.tw);
    for (int t = threadIdx.x; t < NT * C; t += TPB) p.z[base + t] = lds[t];
}

// Inverse column pass: conjugate twiddle, inverse R-point DFT down columns.
template <int R, int NT, int TPB>
__global__ __launch_bounds__(TPB) void k_inv_col(Params p)
{
    __shared__ double2 lds[NT * R];
    const unsigned c0 = blockIdx.x * NT;
    for (int t = threadIdx.x; t < NT * R; t += TPB) {
        const int rho = t / NT, b = t % NT;
        const unsigned c = c0 + b;
        double2 w = twL(p, c * (unsigned)rho);
        w.y = -w.y;
        lds[b * R + rho] = cmul(p.z[(size_t)rho * p.C + c], w);
    }
    __syncthreads();
    lds_fft<R, NT, TPB, true>(lds, p.tw);
    for (int t = threadIdx.x; t < NT * R; t += TPB) {
        const int r = t / NT, b = t % NT;
        p.z[(size_t)r * p.C + c0 + b] = lds[b * R + r];
    }
}

// ------------------------------------------------------------- carry

// Folds one run's carry-in into its first digits.  c2in is the first fold's
// carry arriving at the run's first bit (it already includes -k times the
// previous run's T_lo carry); it is spread over balanced digits delta_j,
// which enter the second fold as -k*delta_j; extra is the second fold's
// carry-in at the same bit.  Both are normalised into this run's digits.
template <int RUN>
__device__ __forceinline__ void adjust_run(const Params& p, unsigned m0, long long* x, long long c2in, long long extra)
{
    long long a[RUN];
    long long rem = c2in;
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
        long long c;
        long long d = bal(rem, width(p, m0 + j), c);
        a[j] = -p.k * d;
        rem = c;
    }
    if (rem != 0) atomicOr(p.status, 1);
    a[0] += extra;
    long long c = 0;
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
        long long cc;
        x[j] += bal(a[j] + c, width(p, m0 + j), cc);
        c = cc;
    }
    if (c != 0) atomicOr(p.status, 2);
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

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
        const int b = width(p, m);
        const long long d1 = bal((long long)tr + q1, b, q1);       // T_lo
        const long long d2 = bal((long long)hr - k * d1 + q2, b, q2);  // S = T_hi - k*T_lo
        x[j] = bal(-k * d2 + q3, b, q3);                            // R = S_hi - k*S_lo
    }
    sq[0][tid] = q1;
    sq[1][tid] = q2;
    sq[2][tid] = q3;
    __syncthreads();
    if (tid > 0) {
        const long long c1 = sq[0][tid - 1], c2 = sq[1][tid - 1], c3 = sq[2][tid - 1];
        adjust_run<RUN>(p, m0, x, c2 - k * c1, c3);
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
    long long x[RUN];
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
        const unsigned m = m0 + j;
        const double2 v = cmul(p.z[m], inv_weight(p, m));
        x[j] = llrint(v.x * (double)p.L);
    }
    adjust_run<RUN>(p, m0, x, c2in, extra);
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
        const unsigned m = m0 + j;
        const double2 wt = fwd_weight(p, m);
        const double xd = (double)x[j];
        p.z[m] = make_double2(xd * wt.x, xd * wt.y);
    }
    if (w == 0) *p.xhi = p.carries[3 * (nblocks - 1) + 2];
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

double max_bits_per_digit(uint32_t logL)
{
    if (logL <= 14) return 18.5;
    if (logL <= 17) return 18.0;
    if (logL <= 20) return 17.5;
    return 17.0;
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
    out.R = 1u << (logL / 2);
    out.C = L / out.R;
    out.s = (logL + 1) / 2;
    return true;
}

static Params make_params(const Plan& pl, double2* z, const DeviceTables* t, long long* carries, long long* xhi,
                          unsigned long long* err, int* status)
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
static void launch_col(bool inverse, const Params& p, hipStream_t st)
{
    constexpr int NT = (4096 / R) < 8 ? (4096 / R) : 8;
    constexpr int TPB = NT * R >= 2048 ? 256 : (NT * R >= 1024 ? 128 : 64);
    dim3 grid(p.C / NT);
    if (inverse) hipLaunchKernelGGL((k_inv_col<R, NT, TPB>), grid, dim3(TPB), 0, st, p);
    else hipLaunchKernelGGL((k_fwd_col<R, NT, TPB>), grid, dim3(TPB), 0, st, p);
}

template <int C>
static void launch_row(const Params& p, hipStream_t st)
{
    constexpr int NT = (1024 / C) > 1 ? (1024 / C) : 1;
    constexpr int TPB = NT * C >= 2048 ? 256 : 128;
    dim3 grid(p.R / NT);
    hipLaunchKernelGGL((k_row<C, NT, TPB>), grid, dim3(TPB), 0, st, p);
}

static void dispatch_col(bool inverse, const Params& p, hipStream_t st)
{
    switch (p.R) {
    case 32: launch_col<32>(inverse, p, st); break;
    case 64: launch_col<64>(inverse, p, st); break;
    case 128: launch_col<128>(inverse, p, st); break;
    case 256: launch_col<256>(inverse, p, st); break;
    case 512: launch_col<512>(inverse, p, st); break;
    case 1024: launch_col<1024>(inverse, p, st); break;
    case 2048: launch_col<2048>(inverse, p, st); break;
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

Engine::Engine(const Plan& p, int device) : plan_(p), device_(device)
{
    check(hipSetDevice(device), "hipSetDevice");
    check(hipStreamCreate(&stream_), "hipStreamCreate");
    std::vector<double2> tw, twl_lo, twl_hi, om_lo, om_hi;
    std::vector<double> pw_lo, pw_hi, ipw_lo, ipw_hi;
    host_tables(p, tw, twl_lo, twl_hi, om_lo, om_hi, pw_lo, pw_hi, ipw_lo, ipw_hi);
    carry_blocks_ = p.L / (kRun * kCarryTpb);

    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t off = 0;
    const size_t o_z = off; off += al(sizeof(double2) * p.L);
    const size_t o_car = off; off += al(sizeof(long long) * 3 * carry_blocks_);
    const size_t o_xhi = off; off += al(sizeof(long long));
    const size_t o_err = off; off += al(sizeof(unsigned long long) * kGraphChunk);
    const size_t o_st = off; off += al(sizeof(int));
    const size_t o_dig = off; off += al(sizeof(int32_t) * p.L);
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
    check(hipMemsetAsync(mem_, 0, o_tw, stream_), "hipMemsetAsync");
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

void Engine::upload(const std::vector<int32_t>& digits, int64_t xhi)
{
    check(hipSetDevice(device_), "hipSetDevice");
    if (digits.size() != plan_.L) throw std::runtime_error("digit count mismatch");
    long long xh = xhi;
    check(hipMemcpyAsync(digits_dev_, digits.data(), sizeof(int32_t) * plan_.L, hipMemcpyHostToDevice, stream_),
          "upload digits");
    check(hipMemcpyAsync(xhi_, &xh, sizeof(long long), hipMemcpyHostToDevice, stream_), "upload xhi");
    Params p = make_params(plan_, z_, tabs_, carries_, xhi_, err_, status_);
    hipLaunchKernelGGL(k_set_digits, dim3((plan_.L + 255) / 256), dim3(256), 0, stream_, p, digits_dev_);
    check(hipGetLastError(), "k_set_digits");
    check(hipStreamSynchronize(stream_), "upload sync");
}

void Engine::download(std::vector<int32_t>& digits, int64_t& xhi)
{
    check(hipSetDevice(device_), "hipSetDevice");
    Params p = make_params(plan_, z_, tabs_, carries_, xhi_, err_, status_);
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

void Engine::enqueue_one(uint64_t slot)
{
    Params p = make_params(plan_, z_, tabs_, carries_, xhi_, err_ + slot, status_);
    dispatch_col(false, p, stream_);
    dispatch_row(p, stream_);
    dispatch_col(true, p, stream_);
    hipLaunchKernelGGL((k_carry<kRun, kCarryTpb>), dim3(carry_blocks_), dim3(kCarryTpb), 0, stream_, p);
    hipLaunchKernelGGL((k_fixup<kRun, kCarryTpb>), dim3((carry_blocks_ + 63) / 64), dim3(64), 0, stream_, p,
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
    for (int i = 0; i < iters; ++i) enqueue_one((uint64_t)i);
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
            for (int i = 0; i < g; ++i) enqueue_one((uint64_t)i);
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

void Engine::kernel_times(int reps, double* ms5)
{
    check(hipSetDevice(device_), "hipSetDevice");
    std::vector<int32_t> saved;
    int64_t saved_xhi = 0;
    download(saved, saved_xhi);
    Params p = make_params(plan_, z_, tabs_, carries_, xhi_, err_, status_);
    hipEvent_t a, b;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    for (int kind = 0; kind < kNumKernels; ++kind) {
        // one untimed launch, then reps timed launches
        for (int pass = 0; pass < 2; ++pass) {
            const int cnt = pass == 0 ? 1 : reps;
            if (pass == 1) check(hipEventRecord(a, stream_), "hipEventRecord");
            for (int i = 0; i < cnt; ++i) {
                switch (kind) {
                case 0: dispatch_col(false, p, stream_); break;
                case 1: dispatch_row(p, stream_); break;
                case 2: dispatch_col(true, p, stream_); break;
                case 3:
                    hipLaunchKernelGGL((k_carry<kRun, kCarryTpb>), dim3(carry_blocks_), dim3(kCarryTpb), 0, stream_, p);
                    break;
                default:
                    hipLaunchKernelGGL((k_fixup<kRun, kCarryTpb>), dim3((carry_blocks_ + 63) / 64), dim3(64), 0,
                                       stream_, p, carry_blocks_);
                }
            }
            if (pass == 1) check(hipEventRecord(b, stream_), "hipEventRecord");
        }
        check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0;
        check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        ms5[kind] = (double)ms / reps;
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
