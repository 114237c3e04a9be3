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
// S, the second fold R) run per run of consecutive digits, exactly in FP64;
// each run publishes its own three carry-outs, computed from its own digits
// only, and the next run folds them into its first digits (one hop, no
// chain).
//
// FFT.  L = R*C, four-step on the R x C row-major array (digit m at r*C + c):
// R-point DFTs down the columns with the W_L^(c*rho) twiddle, then C-point
// DFTs along rows; the spectrum stays in that transposed order.  One squaring
// is two kernels in steady state:
//   k_carry_turn : inverse column DFT -> unweight/round -> own carries of
//                  each run of 8 digits -> carry hand-off to the next column
//                  block -> predecessor's carries folded in -> weight ->
//                  forward column DFT
//   k_row        : forward row DFT -> Z*(Z + 2i*X_hi) -> inverse row DFT
// k_inv_carry + k_adjust_fwd are the same outer pass in two launches (runs
// of 16 digits, or PL_CARRY_TURN=0); k_fwd_col / k_inv_col / k_carry /
// k_fixup are the unfused pieces used to enter and leave the steady state
// (and for lengths the split pass does not cover).
//
// Slots.  Every launch covers the slots of a Group (blockIdx.y = slot); a
// slot binds its own n, k, residue, carry staging, X_hi and round-off slot.
// Scheduled stepping multiplies the second fold by a per-slot small factor
// (x <- x^2 * a when the current bit of k is set), which runs the seed
// exponentiation a^k on the device with the same kernels, and retires a
// slot after its last squaring (inactive slots' blocks exit at once).
#include "engine.h"
#include "fft_lds.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
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
    // Factored weights of digit m = r*C + c (per slot, they depend on n):
    // fwd a_m*w^m = wrf[r]*wcf[c] * (1/2 when F_r + G_c >= L), inverse
    // conj(w^m)/(L*a_m) = wri[r]*wci[c] * (2 when F_r + G_c >= L), with
    // F_r = (-n*r*C) mod L, G_c = (-n*c) mod L.
    const double2* wrf;
    const double2* wri;
    const double2* wcf;
    const double2* wci;
    const double2* tcq;     // W_L^(c*q*R/16), c < C, q < 16 (index c*16 + q)
    long long* carries;     // flat carry kernel: 3 per block (own carry-outs of its last run)
    long long* runc;        // split outer pass: 3 carry-outs per run (block-major, then row);
                            // the one-launch pass stores them as tagged words
    unsigned* epoch;        // per slot: [0] the next hand-off epoch, [1] the last one used,
                            // [2] the round-off column of the squaring the next outer pass
                            // carries (the row pass advances it), [3] unused
    int* dig;               // split outer pass: NT int32 digits per run (same order)
    long long* xhi;
    const long long* xhi_mul;  // product pass: the other factor's X_hi (nullptr: squaring)
    unsigned long long* err;  // per-block round-off partials: block b, column c at err[b*kChunk + c]
    int* status;
    const SlotDesc* desc;
    SlotState* state;
    size_t carr_stride;     // per-slot strides of carries / runc / err
    size_t runc_stride;
    unsigned err_stride;
    unsigned slot0;         // first slot of this launch
    int sched;              // apply per-slot multipliers and the active flag
    unsigned spin_limit;    // polls before a carry hand-off wait gives up (status bit 2)
    unsigned tag_skew;      // 0; a test of the give-up path expects tags that never come
    int per_slot;           // slots hold different candidates: read n, k from desc
                            // (else n, k below are every slot's, and no kernel
                            // starts with a dependent descriptor load)
    int ztag;               // row pass: store points with ready parity 1 (the pass
                            // entering a k_loop launch)
    unsigned poll_sleep;    // k_loop: pause between reloads of points not yet ready
                            // (units of s_sleep 1, about 64 clocks)
    unsigned step_off;      // scheduled stepping: steps of this chunk before this launch's
                            // (the slot states advance once per chunk, k_advance)
    // bound per slot by bind()
    unsigned long long n;
    long long k;
    long long mult;
    int active;
    unsigned L, logL, R, C, s, smask;
};

struct DeviceTables {
    double2* tw;
    double2 *twl_lo, *twl_hi, *om_lo, *om_hi;
    double *pw_lo, *pw_hi, *ipw_lo, *ipw_hi;
    double2* tcq;
    double2 *wrf, *wri, *wcf, *wci;  // per slot: R, R, C, C entries
};

// The slot's view of the launch parameters.  SLOTS = false (plain squaring
// of slots that all hold the plan's candidate: per_slot = sched = 0) skips
// the descriptor test, whose kernel-argument load and branch would
// otherwise sit in front of every first memory request of the kernel
// (0.2-0.45 us per squaring at n = 2^21..2^24).
template <bool SLOTS = true>
__device__ __forceinline__ Params bind(const Params& base, unsigned y)
{
    Params p = base;
    const unsigned slot = base.slot0 + y;
    p.z += (size_t)slot * base.L;
    p.dig += (size_t)slot * base.L;
    p.runc += (size_t)slot * base.runc_stride;
    p.epoch += 4 * (size_t)slot;
    p.carries += (size_t)slot * base.carr_stride;
    p.xhi += slot;
    p.err += (size_t)slot * base.err_stride;  // err_stride = blocks * kChunk
    p.wrf += (size_t)slot * base.R;
    p.wri += (size_t)slot * base.R;
    p.wcf += (size_t)slot * base.C;
    p.wci += (size_t)slot * base.C;
    p.mult = 1;
    p.active = 1;
    if (SLOTS && (base.per_slot || base.sched)) {
        const SlotDesc d = base.desc[slot];
        p.n = d.n;
        p.k = d.k;
        if (base.sched) {
            // the state as of the chunk's start, moved on by the chunk's
            // steps before this one (a slot stops once it has carried
            // total - 1 squarings; at step_off 0 the state is taken as it is,
            // which lets the leave sequence run a finished slot)
            const SlotState st = base.state[slot];
            const unsigned long long c = st.carried + base.step_off;
            p.active = st.active && (base.step_off == 0 || c + 1 < d.total);
            const unsigned long long i = c + 1;  // squaring being carried
            if (d.a && i <= d.topbit && ((d.k >> (d.topbit - i)) & 1)) p.mult = d.a;
        }
    }
    return p;
}

// ------------------------------------------------------------- helpers

// v with the lowest mantissa bit of both parts set: the ready parity of a
// row pass's points (resident.h).
__device__ __forceinline__ double2 with_parity1(double2 v)
{
    return make_double2(__longlong_as_double(__double_as_longlong(v.x) | 1ll),
                        __longlong_as_double(__double_as_longlong(v.y) | 1ll));
}

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
__device__ __forceinline__ double2 fwd_weight_f(const Params& p, unsigned m, unsigned f)
{
    double a = p.pw_hi[f >> p.s] * p.pw_lo[f & p.smask];
    double2 w = omega(p, m);
    return make_double2(w.x * a, w.y * a);
}
__device__ __forceinline__ double2 inv_weight_f(const Params& p, unsigned m, unsigned f)
{
    double ia = p.ipw_hi[f >> p.s] * p.ipw_lo[f & p.smask];
    double2 w = omega(p, m);
    return make_double2(w.x * ia, -w.y * ia);
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

// Exact balanced split in FP64 for integers below 2^53: v = d + q*2^b.
// (v + 2^(b-1)) is exact, the scaling by 2^-b is exact, floor is exact and
// the fma residual is an integer of at most b bits, so nothing rounds.
__device__ __forceinline__ double balf(double v, double two_b, double inv_two_b, double& carry)
{
    const double q = floor(fma(v, inv_two_b, 0.5));
    carry = q;
    return fma(-q, two_b, v);
}

// Digit widths and weight indices walk a run incrementally: with
// f_m = (-n*m) mod L, b_m = floor(n/L) + (f_m < n mod L) and
// f_(m+1) = f_m - (n mod L) (+L when it would go negative).
struct DigitWalk {
    unsigned f, nr, L, bq;
    __device__ __forceinline__ DigitWalk(const Params& p, unsigned m0)
        : f(fidx(p, m0)), nr((unsigned)(p.n & (p.L - 1))), L(p.L), bq((unsigned)(p.n >> p.logL))
    {
    }
    __device__ __forceinline__ bool wide() const { return f < nr; }  // b_m = bq + 1
    __device__ __forceinline__ void next() { f = f >= nr ? f - nr : f + L - nr; }
};

// One digit through the three carry passes: T_lo, S = T_hi - k*T_lo,
// R = mult*(S_hi - k*S_lo), all exact in FP64 (|T|, |k*d|, |k*mult*d| stay
// below 2^52 while the FFT round-off is below the guard).  Returns the digit
// of R; advances the carries (exact integers held in doubles).
__device__ __forceinline__ double carry3f(double tr, double hr, double kd, double kmd, double two_b, double inv_two_b,
                                          double& q1, double& q2, double& q3)
{
    const double d1 = balf(tr + q1, two_b, inv_two_b, q1);
    const double d2 = balf(fma(-kd, d1, hr) + q2, two_b, inv_two_b, q2);
    return balf(fma(-kmd, d2, q3), two_b, inv_two_b, q3);
}

// Adjustment a run receives from its predecessor: c2in is the first fold's
// carry arriving at the run's first bit (it already includes -k times the
// predecessor's T_lo carry); it is spread over balanced digits delta_j, which
// enter the second fold as -k*mult*delta_j; extra is the second fold's
// carry-in at the same bit.  The sum is normalised into W balanced digits
// nu_j added to the run's digits.  128-bit intermediates: c2in and extra can
// exceed 2^63 for k near 2^32.  Sets status when W digits do not absorb it.
__device__ __forceinline__ __int128 bal128(__int128 v, int b, __int128& carry)
{
    const __int128 q = (v + ((__int128)1 << (b - 1))) >> b;
    carry = q;
    return v - q * ((__int128)1 << b);
}

__device__ __forceinline__ long long bal64(long long v, int b, long long& carry)
{
    const long long q = (v + (1ll << (b - 1))) >> b;
    carry = q;
    return v - q * (1ll << b);
}

__device__ __forceinline__ bool fits62(__int128 v)
{
    return v < ((__int128)1 << 62) && v > -((__int128)1 << 62);
}

// Balanced split of a 64-bit v by a width b <= 31: v = d + q*2^b with d in
// [-2^(b-1), 2^(b-1)).  With t = v + 2^(b-1), q = floor(t / 2^b) (an
// arithmetic shift) and d = (t mod 2^b) - 2^(b-1), which the low 32 bits
// of t decide: one 64-bit add, one 64-bit shift and 32-bit masking, where a
// multiply by 2^b would cost a 64-bit product.
__device__ __forceinline__ int bal_lo(long long v, int b, long long& q)
{
    const long long t = v + (long long)(1u << (b - 1));
    q = t >> b;
    return (int)((unsigned)t & ((1u << b) - 1u)) - (int)(1u << (b - 1));
}

template <int W>
__device__ __forceinline__ void ripple(const Params& p, unsigned m0, __int128 c2in, __int128 extra, long long (&nu)[W])
{
    int bw[W];
    DigitWalk walk(p, m0);
#pragma unroll
    for (int j = 0; j < W; ++j) {
        bw[j] = (int)walk.bq + (walk.wide() ? 1 : 0);
        walk.next();
    }
    const long long km = p.k * p.mult;
    if (fits62(c2in) && fits62(extra)) {
        // 64-bit path: |c2in|, |extra| < 2^62 and |km*d| < 2^(32+7+18) keep
        // every sum below 2^63.
        long long rem = (long long)c2in;
        int d[W];
#pragma unroll
        for (int j = 0; j < W; ++j) d[j] = bal_lo(rem, bw[j], rem);
        long long c = (long long)extra;
        if (km < (1ll << 31)) {
            // k*mult fits 32 bits (every pool step, k < 2^31): 32x32 -> 64 products
            const int nk = -(int)km;
#pragma unroll
            for (int j = 0; j < W; ++j) nu[j] = bal_lo((long long)nk * d[j] + c, bw[j], c);
        } else {
#pragma unroll
            for (int j = 0; j < W; ++j) nu[j] = bal_lo(-km * d[j] + c, bw[j], c);
        }
        if (rem != 0 || c != 0) atomicOr(p.status, 1);
        return;
    }
    __int128 rem = c2in;
    long long a[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        __int128 c;
        const long long d = (long long)bal128(rem, bw[j], c);
        a[j] = -km * d;
        rem = c;
    }
    __int128 c = extra;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        __int128 cc;
        nu[j] = (long long)bal128((__int128)a[j] + c, bw[j], cc);
        c = cc;
    }
    if (rem != 0 || c != 0) atomicOr(p.status, 1);
}

// Second-fold carry-in of run (0, 0): mult*(c2_top + X_hi^2) -- for a
// product X_hi*Y_hi, the term the pointwise step leaves out -- and the first
// fold's carry-in c1_top (T_lo's carry at bit n is part of T_hi).
__device__ __forceinline__ void top_rule(const Params& p, const long long* top, __int128& c2in, __int128& extra)
{
    const long long xh = *p.xhi;
    const long long yh = p.xhi_mul ? *p.xhi_mul : xh;
    c2in = top[0];
    extra = ((__int128)top[1] + (__int128)xh * yh) * p.mult;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Four-step twiddles W_L^(c*idx) of the points one thread moves in the
// first (or last) stage of a column DFT, where idx = j + r*R/P:
// W_L^(c*j) (two-level table) times W_L^(c*r*R/P) = q^r with
// q = tcq[c*16 + 16/P], powered by multiplication (one table load per
// column instead of P-1; 0.5 us per squaring at n = 2^21).
template <int R, int NT, int TPB, int P>
struct ColTw {
    using S = StageShape<R, NT, TPB, P>;
    double2 w[S::PER][P];
    // SCALED: every factor also carries f, a factor that is constant down
    // column c (k_carry_turn folds the column weight in here: the column DFT
    // is linear, so scaling its input or output by f scales the whole column)
    template <bool SCALED = false>
    __device__ __forceinline__ void load(const Params& p, unsigned c, double2 f = make_double2(1.0, 0.0),
                                         unsigned tid = threadIdx.x)
    {
#pragma unroll
        for (int u = 0; u < S::PER; ++u) {
            if (!S::live(u, tid)) continue;
            double2 a = twL(p, c * (unsigned)S::j(u, tid));
            if constexpr (SCALED) a = cmul(a, f);
            const double2 q = p.tcq[(size_t)c * 16 + 16 / P];
            w[u][0] = a;
            double2 qr = q;
#pragma unroll
            for (int r = 1; r < P; ++r) {
                w[u][r] = cmul(a, qr);
                if (r + 1 < P) qr = cmul(qr, q);
            }
        }
    }
};

// ------------------------------------------------------------- FFT kernels

// Column passes alone (entering the steady state from digits, and the
// unfused outer pass).  The four-step factors of the points a thread moves
// come from ColTw: one table load per thread instead of two per point.
template <int R, int NT, int TPB>
__global__ __launch_bounds__(TPB) void k_fwd_col(Params pb)
{
    const Params p = bind(pb, blockIdx.y);
    if (!p.active) return;
    __shared__ double2 lds[LdsLayout<R, NT>::kSlots];
    StageTw<R, NT, TPB> t;
    load_stage_tw(t, p.tw);
    const unsigned c0 = blockIdx.x * NT;
    ColTw<R, NT, TPB, Radix<R>::last()> ct;
    ct.load(p, c0 + threadIdx.x % NT);
    const double2* __restrict__ z = p.z;
    dft<R, NT, TPB, false>(
        lds, t, [&](int tr, int idx, int, int) { return z[(size_t)idx * p.C + c0 + tr]; },
        [&](int tr, int idx, double2 v, int u, int r) { p.z[(size_t)idx * p.C + c0 + tr] = cmul(v, ct.w[u][r]); });
}

template <int R, int NT, int TPB>
__global__ __launch_bounds__(TPB) void k_inv_col(Params pb)
{
    const Params p = bind(pb, blockIdx.y);
    if (!p.active) return;
    __shared__ double2 lds[LdsLayout<R, NT>::kSlots];
    StageTw<R, NT, TPB> t;
    load_stage_tw(t, p.tw);
    const unsigned c0 = blockIdx.x * NT;
    ColTw<R, NT, TPB, Radix<R>::first()> ct;
    ct.load(p, c0 + threadIdx.x % NT);
    const double2* __restrict__ z = p.z;
    dft<R, NT, TPB, true>(
        lds, t, [&](int tr, int idx, int u, int r) { return cmulc(z[(size_t)idx * p.C + c0 + tr], ct.w[u][r]); },
        [&](int tr, int idx, double2 v, int, int) { p.z[(size_t)idx * p.C + c0 + tr] = v; });
}

// Row pass: NT rows; forward C-point DFT, Z <- Z*(Z + 2i*X_hi), inverse
// C-point DFT.  Where the first and last radix agree (C = 64, 512, 4096 at
// radix 8) the forward last stage, the pointwise step and the inverse first
// stage run in registers.  The stage twiddles are read once and serve both
// directions.
template <int C, int NT, int TPB, bool SLOTS, int RADIX>
__global__ __launch_bounds__(TPB) void k_row(Params pb)
{
    const Params p = bind<SLOTS>(pb, blockIdx.y);
    if (!p.active) return;
    // two LDS buffers (ping-pong stages, half the barriers); at C = 4096
    // that is 128 KB, which one workgroup per CU has to itself anyway
    __shared__ double2 lds[2][LdsLayout<C, NT>::kSlots];
    double2* const buf[2] = { lds[0], lds[1] };
    StageTw<C, NT, TPB, RADIX> t;
    load_stage_tw(t, p.tw);
    const size_t base = (size_t)blockIdx.x * NT * C;
    const double2* __restrict__ z = p.z;
    const double x2 = 2.0 * (double)(*p.xhi);
    // The next carry hand-off's epoch (k_carry_turn), set by block 0.  Every
    // lane reads the last one with a vector load that nothing waits for until
    // the row is stored (a load in the writing lane's branch alone would be
    // waited for there, in front of block 0's row loads).
    const unsigned last = __hip_atomic_load(p.epoch + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned col = __hip_atomic_load(p.epoch + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    dft_square_pass_pp<C, NT, TPB, RADIX>(
        buf, t, [&](int tr, int idx, int, int) { return z[base + (size_t)tr * C + idx]; },
        [&](double2 X) { return cmul(X, make_double2(X.x, X.y + x2)); },
        [&](int tr, int idx, double2 v, int, int) {
            store_stream(&p.z[base + (size_t)tr * C + idx], p.ztag ? with_parity1(v) : v);
        });
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        p.epoch[0] = last + 1u;
        p.epoch[2] = col + 1u;  // the next outer pass carries the next squaring
    }
}

// Row pass of a product: this slot's row times the same row of another
// residue in the same column-transformed state (zo, X_hi,o).  Forward
// C-point DFTs of both, Z <- Zd*(Zo + i*X_hi,o) + i*X_hi,d*Zo (the product
// of Zd + i*X_hi,d and Zo + i*X_hi,o without its constant -X_hi,d*X_hi,o,
// which the carry's top rule adds back as X_hi,d*X_hi,o*2^(2n)), inverse
// DFT into this slot.  For zo == z it is the squaring pass.
template <int C, int NT, int TPB>
__global__ __launch_bounds__(TPB) void k_row_mul(Params pb, const double2* __restrict__ zo, const long long* __restrict__ xo_p)
{
    const Params p = bind(pb, blockIdx.y);
    if (!p.active) return;
    using Lay = LdsLayout<C, NT>;
    __shared__ double2 lds[Lay::kSlots];
    __shared__ double2 ldo[Lay::kSlots];
    StageTw<C, NT, TPB, PL_ROW_RADIX> t;
    load_stage_tw(t, p.tw);
    const size_t base = (size_t)blockIdx.x * NT * C;
    const double2* __restrict__ z = p.z;
    const double xd = (double)(*p.xhi), xo = (double)(*xo_p);
    dft<C, NT, TPB, false, false, true>(
        ldo, t, [&](int tr, int idx, int, int) { return zo[base + (size_t)tr * C + idx]; },
        [&](int tr, int idx, double2 v, int, int) { ldo[Lay::at(tr, idx)] = v; });
    __syncthreads();
    dft<C, NT, TPB, false, false, true>(
        lds, t, [&](int tr, int idx, int, int) { return z[base + (size_t)tr * C + idx]; },
        [&](int tr, int idx, double2 X, int, int) {
            const double2 O = ldo[Lay::at(tr, idx)];
            lds[Lay::at(tr, idx)] = cadd(cmul(X, make_double2(O.x, O.y + xo)), make_double2(-xd * O.y, xd * O.x));
        });
    __syncthreads();
    dft<C, NT, TPB, true, true>(
        lds, t, [&](int tr, int idx, int, int) { return lds[Lay::at(tr, idx)]; },
        [&](int tr, int idx, double2 v, int, int) { store_stream(&p.z[base + (size_t)tr * C + idx], v); });
}

// ------------------------------------------------------------- flat carry

// Unweight, round (with max round-off), three carry passes per run of RUN
// digits, in-block run hand-off, re-weight.  The block's TPB*RUN digits are
// staged through LDS so global loads and stores are coalesced; in LDS run t
// occupies slots t*(RUN+1) .. +RUN-1 (one pad slot per run keeps the
// per-thread accesses conflict-free).  The first run of each block is
// finished by k_fixup.
template <int RUN, int TPB>
__global__ __launch_bounds__(TPB) void k_carry(Params pb)
{
    const Params p = bind(pb, blockIdx.y);
    if (!p.active) return;
    // round-off column, requested first (its use is the kernel's last store)
    const unsigned ecol = p.epoch[2] & (Group::kChunk - 1);
    constexpr int SPAN = TPB * RUN;
    __shared__ double2 buf[SPAN + TPB];
    __shared__ long long sq[3][TPB];
    __shared__ double s_err[TPB / 64];
    const unsigned tid = threadIdx.x;
    const unsigned base = blockIdx.x * SPAN;
    for (int i = tid; i < SPAN; i += TPB) buf[i + i / RUN] = p.z[base + i];
    __syncthreads();

    const unsigned m0 = base + tid * RUN;
    double2* run = buf + tid * (RUN + 1);
    const long long k = p.k;
    const double kd = (double)p.k, kmd = (double)(p.k * p.mult);
    DigitWalk walk(p, m0);
    const double tb0 = ldexp(1.0, (int)walk.bq), tb1 = 2.0 * tb0;
    const double itb0 = 1.0 / tb0, itb1 = 0.5 * itb0;
    double q1 = 0, q2 = 0, q3 = 0;
    double emax = 0.0;
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
        const unsigned m = m0 + j;
        const double2 v = cmul(run[j], inv_weight_f(p, m, walk.f));
        const double tr = rint(v.x), hr = rint(v.y);
        emax = fmax(emax, fmax(fabs(v.x - tr), fabs(v.y - hr)));
        const bool wd = walk.wide();
        run[j].x = carry3f(tr, hr, kd, kmd, wd ? tb1 : tb0, wd ? itb1 : itb0, q1, q2, q3);
        walk.next();
    }
    sq[0][tid] = (long long)q1;
    sq[1][tid] = (long long)q2;
    sq[2][tid] = (long long)q3;
    __syncthreads();
    long long nu[RUN];
    if (tid > 0) {
        ripple<RUN>(p, m0, (__int128)sq[1][tid - 1] - (__int128)k * sq[0][tid - 1], sq[2][tid - 1], nu);
    } else {
#pragma unroll
        for (int j = 0; j < RUN; ++j) nu[j] = 0;
    }
    if (tid == TPB - 1) {
        p.carries[3 * blockIdx.x + 0] = sq[0][tid];
        p.carries[3 * blockIdx.x + 1] = sq[1][tid];
        p.carries[3 * blockIdx.x + 2] = sq[2][tid];
    }
    DigitWalk walk2(p, m0);
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
        const double2 w = fwd_weight_f(p, m0 + j, walk2.f);
        walk2.next();
        const double xd = run[j].x + (double)nu[j];
        run[j] = make_double2(xd * w.x, xd * w.y);
    }
    emax = wave_max(emax);
    if ((tid & 63) == 0) s_err[tid / 64] = emax;
    __syncthreads();
    for (int i = tid; i < SPAN; i += TPB) p.z[base + i] = buf[i + i / RUN];
    if (tid == 0) {
        double e = s_err[0];
#pragma unroll
        for (int i = 1; i < TPB / 64; ++i) e = fmax(e, s_err[i]);
        p.err[(size_t)blockIdx.x * Group::kChunk + ecol] =
            (unsigned long long)__double_as_longlong(e);
    }
}

// Block boundaries: the first run of block w takes the carries of block
// w-1's last run.  Run 0 takes the top: +c1_top (T_lo's carry at bit n is
// part of T_hi) and c2_top + X_hi^2 (the second fold's top scalar); the new
// X_hi is the top run's third carry.
template <int RUN, int CTPB>
__global__ __launch_bounds__(64) void k_fixup(Params pb, unsigned nblocks)
{
    const Params p = bind(pb, blockIdx.y);
    if (!p.active) return;
    const unsigned w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nblocks) return;
    const unsigned m0 = w * CTPB * RUN;
    const long long k = p.k;
    __int128 c2in, extra;
    if (w > 0) {
        const long long* c = p.carries + 3 * (w - 1);
        c2in = (__int128)c[1] - (__int128)k * c[0];
        extra = c[2];
    } else {
        top_rule(p, p.carries + 3 * (nblocks - 1), c2in, extra);
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

// ------------------------------------------------------------- split outer pass

// Column block b owns columns [b*NT, b*NT+NT): for each row r its digits
// r*C + b*NT .. +NT-1 form one run.  The predecessor of run (b, r) is
// (b-1, r), or (nb-1, r-1) for b = 0; run (0, 0) takes the top rule.  Run
// data between the two kernels is block-major: run (b, r) has its NT int32
// digits at dig + (b*R + r)*NT and its three carry-outs at runc + (b*R + r)*3.

// Inverse column DFT, unweight, round (max round-off), own carries of every
// run; writes the runs' digits (before their carry-in) and carry-outs.  All
// table values the block needs (stage twiddles, four-step twiddles, row and
// column weights) are requested before the transform, beside the data.
template <int R, int NT, int TPB, int MAXR>
__global__ __launch_bounds__(TPB) void k_inv_carry(Params pb)
{
    const Params p = bind(pb, blockIdx.y);
    if (!p.active) return;
    const unsigned ecol = p.epoch[2] & (Group::kChunk - 1);
    using Lay = LdsLayout<R, NT>;
    constexpr int RPT = (R + TPB - 1) / TPB;
    constexpr int P0 = Radix<R, MAXR>::first();
    __shared__ double2 lds[Lay::kSlots];
    __shared__ double s_err[TPB / 64];
    const unsigned b = blockIdx.x;
    const unsigned c0 = b * NT;
    const double2* __restrict__ z = p.z;
    StageTw<R, NT, TPB, MAXR> t;
    load_stage_tw(t, p.tw);
    ColTw<R, NT, TPB, P0> ct;
    ct.load(p, c0 + threadIdx.x % NT);
    double2 wr[RPT];
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned r = threadIdx.x + u * TPB;
        wr[u] = r < (unsigned)R ? p.wri[r] : make_double2(0.0, 0.0);
    }
    double2 wc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) wc[j] = p.wci[c0 + j];
    dft<R, NT, TPB, true, false, true>(
        lds, t, [&](int tr, int idx, int u, int r) { return cmulc(z[(size_t)idx * p.C + c0 + tr], ct.w[u][r]); },
        [&](int tr, int idx, double2 v, int, int) { lds[Lay::at(tr, idx)] = v; });
    __syncthreads();

    const unsigned long long Lm = p.L - 1;
    const unsigned nr = (unsigned)(p.n & Lm);
    unsigned G[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) G[j] = (unsigned)((0ull - p.n * (unsigned long long)(c0 + j)) & Lm);
    double emax = 0.0;
    const double kd = (double)p.k, kmd = (double)(p.k * p.mult);
    const double tb0 = ldexp(1.0, (int)(p.n >> p.logL)), tb1 = 2.0 * tb0;
    const double itb0 = 1.0 / tb0, itb1 = 0.5 * itb0;
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned r = threadIdx.x + u * TPB;
        if (r >= (unsigned)R) break;
        const unsigned F = (unsigned)((0ull - p.n * (unsigned long long)r * p.C) & Lm);
        double q1 = 0, q2 = 0, q3 = 0;
        int dg[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            unsigned f = F + G[j];
            const bool wrap = f >= p.L;
            if (wrap) f -= p.L;
            double2 w = cmul(wr[u], wc[j]);
            if (wrap) w = make_double2(2.0 * w.x, 2.0 * w.y);
            const double2 v = cmul(lds[Lay::at(j, r)], w);
            const double tr = rint(v.x), hr = rint(v.y);
            emax = fmax(emax, fmax(fabs(v.x - tr), fabs(v.y - hr)));
            const bool wd = f < nr;
            dg[j] = (int)carry3f(tr, hr, kd, kmd, wd ? tb1 : tb0, wd ? itb1 : itb0, q1, q2, q3);
        }
        const size_t run = (size_t)b * R + r;
        int4* dst = reinterpret_cast<int4*>(p.dig + run * NT);
#pragma unroll
        for (int j = 0; j < NT / 4; ++j) dst[j] = make_int4(dg[4 * j], dg[4 * j + 1], dg[4 * j + 2], dg[4 * j + 3]);
        long long* rc = p.runc + 3 * run;
        rc[0] = (long long)q1;
        rc[1] = (long long)q2;
        rc[2] = (long long)q3;
    }
    emax = wave_max(emax);
    if ((threadIdx.x & 63) == 0) s_err[threadIdx.x / 64] = emax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = s_err[0];
#pragma unroll
        for (int i = 1; i < TPB / 64; ++i) e = fmax(e, s_err[i]);
        p.err[(size_t)blockIdx.x * Group::kChunk + ecol] =
            (unsigned long long)__double_as_longlong(e);
    }
}

// Each run takes its predecessor's carries into its first digits, the
// digits are weighted, and the block runs its forward column DFT.  Block 0
// also replaces X_hi with the top run's third carry (it is the only reader
// of the old value: X_hi^2 enters run (0, 0)).
template <int R, int NT, int TPB, int MAXR>
__global__ __launch_bounds__(TPB) void k_adjust_fwd(Params pb)
{
    const Params p = bind(pb, blockIdx.y);
    if (!p.active) return;
    using Lay = LdsLayout<R, NT>;
    constexpr int RPT = (R + TPB - 1) / TPB;
    constexpr int PL = Radix<R, MAXR>::last();
    __shared__ double2 lds[Lay::kSlots];
    const unsigned nb = p.C / NT;
    const unsigned b = blockIdx.x;
    const unsigned c0 = b * NT;
    const long long k = p.k;
    const long long* top = p.runc + 3 * ((size_t)(nb - 1) * R + (R - 1));
    StageTw<R, NT, TPB, MAXR> t;
    load_stage_tw(t, p.tw);
    ColTw<R, NT, TPB, PL> ct;
    ct.load(p, c0 + threadIdx.x % NT);
    double2 wc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) wc[j] = p.wcf[c0 + j];
    const unsigned long long Lm = p.L - 1;
    unsigned G[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) G[j] = (unsigned)((0ull - p.n * (unsigned long long)(c0 + j)) & Lm);
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned r = threadIdx.x + u * TPB;
        if (r >= (unsigned)R) break;
        const unsigned m0 = r * p.C + c0;
        const double2 wr = p.wrf[r];
        __int128 c2in, extra;
        if (b == 0 && r == 0) {
            top_rule(p, top, c2in, extra);
        } else {
            const long long* rc = p.runc + 3 * (b == 0 ? (size_t)(nb - 1) * R + (r - 1) : (size_t)(b - 1) * R + r);
            c2in = (__int128)rc[1] - (__int128)k * rc[0];
            extra = rc[2];
        }
        long long nu[NT];
        ripple<NT>(p, m0, c2in, extra, nu);
        const int4* src = reinterpret_cast<const int4*>(p.dig + ((size_t)b * R + r) * NT);
        int dg[NT];
#pragma unroll
        for (int j = 0; j < NT / 4; ++j) {
            const int4 q = src[j];
            dg[4 * j] = q.x;
            dg[4 * j + 1] = q.y;
            dg[4 * j + 2] = q.z;
            dg[4 * j + 3] = q.w;
        }
        const unsigned F = (unsigned)((0ull - p.n * (unsigned long long)r * p.C) & Lm);
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            double2 w = cmul(wr, wc[j]);
            if (F + G[j] >= p.L) w = make_double2(0.5 * w.x, 0.5 * w.y);
            const double xd = (double)dg[j] + (double)nu[j];
            lds[Lay::at(j, r)] = make_double2(xd * w.x, xd * w.y);
        }
    }
    // run (0, 0) -- thread 0, u = 0 -- read the old X_hi above, in this thread
    if (b == 0 && threadIdx.x == 0) *p.xhi = top[2];
    __syncthreads();
    dft<R, NT, TPB, false, true, false>(
        lds, t, [&](int tr, int idx, int, int) { return lds[Lay::at(tr, idx)]; },
        [&](int tr, int idx, double2 v, int u, int r) {
            store_stream(&p.z[(size_t)idx * p.C + c0 + tr], cmul(v, ct.w[u][r]));
        });
}

// Leaving the steady state after k_inv_carry: each run takes its
// predecessor's carries into its first digits and the weighted digits are
// stored in their natural places (the digit state k_get_digits and
// k_fwd_col read), with no transform.  Block 0 replaces X_hi with the top
// run's third carry, as k_adjust_fwd does.
template <int R, int NT, int TPB>
__global__ __launch_bounds__(TPB) void k_adjust_out(Params pb)
{
    const Params p = bind(pb, blockIdx.y);
    if (!p.active) return;
    constexpr int RPT = (R + TPB - 1) / TPB;
    const unsigned nb = p.C / NT;
    const unsigned b = blockIdx.x;
    const unsigned c0 = b * NT;
    const long long k = p.k;
    const long long* top = p.runc + 3 * ((size_t)(nb - 1) * R + (R - 1));
    double2 wc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) wc[j] = p.wcf[c0 + j];
    const unsigned long long Lm = p.L - 1;
    unsigned G[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) G[j] = (unsigned)((0ull - p.n * (unsigned long long)(c0 + j)) & Lm);
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned r = threadIdx.x + u * TPB;
        if (r >= (unsigned)R) break;
        const unsigned m0 = r * p.C + c0;
        const double2 wr = p.wrf[r];
        __int128 c2in, extra;
        if (b == 0 && r == 0) {
            top_rule(p, top, c2in, extra);
        } else {
            const long long* rc = p.runc + 3 * (b == 0 ? (size_t)(nb - 1) * R + (r - 1) : (size_t)(b - 1) * R + r);
            c2in = (__int128)rc[1] - (__int128)k * rc[0];
            extra = rc[2];
        }
        long long nu[NT];
        ripple<NT>(p, m0, c2in, extra, nu);
        const int4* src = reinterpret_cast<const int4*>(p.dig + ((size_t)b * R + r) * NT);
        const unsigned F = (unsigned)((0ull - p.n * (unsigned long long)r * p.C) & Lm);
        double2* out = p.z + m0;
#pragma unroll
        for (int j4 = 0; j4 < NT / 4; ++j4) {
            const int4 q = src[j4];
            const int dg[4] = { q.x, q.y, q.z, q.w };
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j = 4 * j4 + i;
                double2 w = cmul(wr, wc[j]);
                if (F + G[j] >= p.L) w = make_double2(0.5 * w.x, 0.5 * w.y);
                const double xd = (double)(dg[i] + (int)nu[j]);
                out[j] = make_double2(xd * w.x, xd * w.y);
            }
        }
    }
    if (b == 0 && threadIdx.x == 0) *p.xhi = top[2];
}

// The outer pass in one launch: inverse column DFT, unweight, round, own
// carries of every run (as k_inv_carry), then each run's carry-outs are
// handed straight to its successor run in the next column block, and the
// block goes on with the carry-in, weighting and forward column DFT (as
// k_adjust_fwd) without its digits leaving registers.
//
// Hand-off.  A run's three carry-outs are exact integers below 2^53 in
// magnitude (they come out of FP64 arithmetic that is exact), so each goes
// out as one 64-bit word q*2^10 + tag, where tag is the launch's epoch mod
// 2^10, with an agent-scope (device-coherent) atomic store.  The successor's
// thread polls the three words with agent-scope loads until all carry its
// epoch: the data is its own ready flag, each word is written whole, and
// nothing but these atomics crosses workgroups, so there is no L2 write-back
// or invalidate (an agent-scope release/acquire pair made the launch four
// times slower at L = 2^20, flushing the z lines other blocks had just
// written) and no second round trip for a separate flag.  The epoch is one
// word per slot: k_row sets it to the last hand-off's epoch plus one (after
// its row is stored), and k_carry_turn reads it with a scalar load and
// records the epoch it used.  Between two hand-offs of a slot
// there is always a row pass, so consecutive hand-offs never share a tag,
// however many row passes run without one.
//
// Every block publishes before it waits and waits only on blockIdx.x - 1,
// except block 0, which waits on the slot's last block; workgroups start in
// increasing order on each XCD.  In a launch that has the device to itself,
// every block but block 0 waits on words published before it started, so
// the blocks ahead of a slot's last block finish and make room for it, and
// block 0's wait ends.  Kernels of other streams that spin the same way can
// take that room instead (Group::no_turn is for callers that overlap them).
// A wait therefore still gives up after kSpinLimit polls (status bit 2, well
// over 0.1 s where a normal wait is microseconds), so that no launch can
// hang.  PL_HANDOFF_GIVE_UP=1 (tests of that path) makes every wait expect a
// tag no run writes and give up after 64 polls.
//
// The words lie in three planes (word i of run q at runc[i*runs + q]), so the
// 64 lanes of a wave store and poll 512 contiguous bytes per instruction
// instead of touching each line three times.
constexpr unsigned kSpinLimit = 1u << 18;
constexpr int kTagBits = 10;
constexpr unsigned kTagMask = (1u << kTagBits) - 1;

__device__ __forceinline__ unsigned long long tag_word(double q, unsigned ep)
{
    return ((unsigned long long)(long long)q << kTagBits) | ep;
}
__device__ __forceinline__ long long untag(unsigned long long w) { return (long long)w >> kTagBits; }
__device__ __forceinline__ unsigned tag_of(unsigned long long w) { return (unsigned)w & kTagMask; }

template <int R, int NT, int TPB, bool SLOTS, int MAXR>
__global__ __launch_bounds__(TPB) void k_carry_turn(Params pb)
{
    const Params p = bind<SLOTS>(pb, blockIdx.y);
    if (!p.active) return;
    using Lay = LdsLayout<R, NT>;
    constexpr int RPT = (R + TPB - 1) / TPB;
    constexpr int P0 = Radix<R, MAXR>::first();
    constexpr int PL = Radix<R, MAXR>::last();
    // Two LDS buffers (ping-pong stages, half the barriers) while two
    // workgroups still fit a CU; one buffer for R*NT > 2048 points.
    constexpr bool PP = Lay::kSlots <= 2048;
    constexpr int IO = pp_out<R, MAXR, 0>();  // buffer the inverse DFT leaves the columns in
    __shared__ double2 lds[PP ? 2 : 1][Lay::kSlots];
    double2* const buf[2] = { lds[0], lds[PP ? 1 : 0] };
    double2* const io = buf[PP ? IO : 0];
    __shared__ double s_err[TPB / 64];
    const unsigned nb = p.C / NT;
    const unsigned b = blockIdx.x;
    const unsigned c0 = b * NT;
    const double2* __restrict__ z = p.z;
    unsigned long long* const rw = reinterpret_cast<unsigned long long*>(p.runc);
    const size_t runs = (size_t)nb * R;
    StageTw<R, NT, TPB, MAXR> t;
    load_stage_tw(t, p.tw);
    const unsigned e = p.epoch[0] & kTagMask;
    const unsigned ecol = p.epoch[2] & (Group::kChunk - 1);
    if (b == 0 && threadIdx.x == 0) p.epoch[1] = e;
    unsigned ep[RPT];
#pragma unroll
    for (int u = 0; u < RPT; ++u) ep[u] = e;
    double2 wri[RPT];
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned r = threadIdx.x + u * TPB;
        wri[u] = r < (unsigned)R ? p.wri[r] : make_double2(0.0, 0.0);
    }
    {
        // the inverse column weight rides on the four-step factors: the data
        // is multiplied by conj(a*conj(wci)*q^r) = conj(a*q^r)*wci
        ColTw<R, NT, TPB, P0> ct;
        const double2 wc = p.wci[c0 + threadIdx.x % NT];
        ct.template load<true>(p, c0 + threadIdx.x % NT, make_double2(wc.x, -wc.y));
        auto ld = [&](int tr, int idx, int u, int r) { return cmulc(z[(size_t)idx * p.C + c0 + tr], ct.w[u][r]); };
        auto st = [&](int tr, int idx, double2 v, int, int) { io[Lay::at(tr, idx)] = v; };
        if constexpr (PP) dft_pp<R, NT, TPB, true, 0>(buf, t, ld, st);
        else dft<R, NT, TPB, true, false, true>(io, t, ld, st);
    }
    __syncthreads();

    const unsigned long long Lm = p.L - 1;
    const unsigned nr = (unsigned)(p.n & Lm);
    unsigned G[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) G[j] = (unsigned)((0ull - p.n * (unsigned long long)(c0 + j)) & Lm);
    double emax = 0.0;
    const long long k = p.k;
    const double kd = (double)p.k, kmd = (double)(p.k * p.mult);
    const double tb0 = ldexp(1.0, (int)(p.n >> p.logL)), tb1 = 2.0 * tb0;
    const double itb0 = 1.0 / tb0, itb1 = 0.5 * itb0;
    int dg[RPT][NT];
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned r = threadIdx.x + u * TPB;
        if (r >= (unsigned)R) break;
        const unsigned F = (unsigned)((0ull - p.n * (unsigned long long)r * p.C) & Lm);
        double q1 = 0, q2 = 0, q3 = 0;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            double2 w = wri[u];
            unsigned f = F + G[j];
            const bool wrap = f >= p.L;
            if (wrap) {
                f -= p.L;
                w = make_double2(2.0 * w.x, 2.0 * w.y);
            }
            const double2 v = cmul(io[Lay::at(j, r)], w);
            const double tr = rint(v.x), hr = rint(v.y);
            emax = fmax(emax, fmax(fabs(v.x - tr), fabs(v.y - hr)));
            const bool wd = f < nr;
            dg[u][j] = (int)carry3f(tr, hr, kd, kmd, wd ? tb1 : tb0, wd ? itb1 : itb0, q1, q2, q3);
        }
        unsigned long long* w = rw + (size_t)b * R + r;
        __hip_atomic_store(w, tag_word(q1, ep[u]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(w + runs, tag_word(q2, ep[u]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(w + 2 * runs, tag_word(q3, ep[u]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // Predecessor of run (b, r): (b-1, r); (nb-1, r-1) for b = 0; run (0, 0)
    // takes the top run (nb-1, R-1) through the top rule.  The first poll of
    // its words goes out now, so that the round-off reduction's lane
    // exchanges and the forward-side tables overlap its round trip.
    const unsigned long long* pw[RPT];
    unsigned long long w0[RPT], w1[RPT], w2[RPT];
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned r = threadIdx.x + u * TPB;
        if (r >= (unsigned)R) break;
        pw[u] = rw + (b == 0 ? (r == 0 ? (size_t)(nb - 1) * R + (R - 1) : (size_t)(nb - 1) * R + (r - 1))
                             : (size_t)(b - 1) * R + r);
        w0[u] = __hip_atomic_load(pw[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w1[u] = __hip_atomic_load(pw[u] + runs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w2[u] = __hip_atomic_load(pw[u] + 2 * runs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    emax = wave_max(emax);
    if ((threadIdx.x & 63) == 0) s_err[threadIdx.x / 64] = emax;

    // forward-side tables (the forward column weight folded into the
    // four-step factors), requested while the predecessors finish
    ColTw<R, NT, TPB, PL> ct;
    ct.template load<true>(p, c0 + threadIdx.x % NT, p.wcf[c0 + threadIdx.x % NT]);
    double2 wrf[RPT];
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned r = threadIdx.x + u * TPB;
        wrf[u] = r < (unsigned)R ? p.wrf[r] : make_double2(0.0, 0.0);
    }

    // Poll until the three words carry this launch's tag.  Each thread
    // rewrites only the LDS slots of its own runs, which only it read in the
    // carry phase: no barrier before the weighted digits go in.
    bool late = false;
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned r = threadIdx.x + u * TPB;
        if (r >= (unsigned)R) break;
        const unsigned want = (ep[u] + p.tag_skew) & kTagMask;
        unsigned spins = 0;
        while (!(tag_of(w0[u]) == want && tag_of(w1[u]) == want && tag_of(w2[u]) == want)) {
            if (++spins >= p.spin_limit) {
                late = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            w0[u] = __hip_atomic_load(pw[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            w1[u] = __hip_atomic_load(pw[u] + runs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            w2[u] = __hip_atomic_load(pw[u] + 2 * runs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const long long cin[3] = { untag(w0[u]), untag(w1[u]), untag(w2[u]) };
        __int128 c2in, extra;
        if (b == 0 && r == 0) {
            top_rule(p, cin, c2in, extra);
            *p.xhi = cin[2];  // the old X_hi was read just above, in this thread
        } else {
            c2in = (__int128)cin[1] - (__int128)k * cin[0];
            extra = cin[2];
        }
        long long nu[NT];
        ripple<NT>(p, r * p.C + c0, c2in, extra, nu);
        const unsigned F = (unsigned)((0ull - p.n * (unsigned long long)r * p.C) & Lm);
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            double2 w = wrf[u];
            if (F + G[j] >= p.L) w = make_double2(0.5 * w.x, 0.5 * w.y);
            // digit and adjustment are balanced digits (|.| <= 2^b/2, b < 32):
            // their sum is exact in 32-bit integers
            const double xd = (double)(dg[u][j] + (int)nu[j]);
            io[Lay::at(j, r)] = make_double2(xd * w.x, xd * w.y);
        }
    }
    if (late) atomicOr(p.status, 2);
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = s_err[0];
#pragma unroll
        for (int i = 1; i < TPB / 64; ++i) e = fmax(e, s_err[i]);
        p.err[(size_t)blockIdx.x * Group::kChunk + ecol] =
            (unsigned long long)__double_as_longlong(e);
    }
    auto ld = [&](int tr, int idx, int, int) { return io[Lay::at(tr, idx)]; };
    auto st = [&](int tr, int idx, double2 v, int u, int r) {
        store_stream(&p.z[(size_t)idx * p.C + c0 + tr], cmul(v, ct.w[u][r]));
    };
    if constexpr (PP) dft_pp<R, NT, TPB, false, 1 - IO>(buf, t, ld, st);
    else dft<R, NT, TPB, false, true, false>(io, t, ld, st);
}

}  // namespace pleng
#include "resident.h"
namespace pleng {

__global__ void k_set_digits(Params pb, const int* __restrict__ d)
{
    const Params p = bind(pb, blockIdx.y);
    const unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= p.L) return;
    const double2 w = fwd_weight(p, m);
    const double x = (double)d[m];
    p.z[m] = make_double2(x * w.x, x * w.y);
}

__global__ void k_get_digits(Params pb, int* __restrict__ d)
{
    const Params p = bind(pb, blockIdx.y);
    const unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= p.L) return;
    const double2 v = cmul(p.z[m], inv_weight(p, m));
    d[m] = (int)llrint(v.x * (double)p.L);
}

// Start of a one-launch chunk: slots' residues (count double2) and top
// scalars (nx) copied to the backup the chunk is recomputed from if a carry
// hand-off gives up.
__global__ __launch_bounds__(256) void k_backup(const double2* __restrict__ z, double2* __restrict__ zb, size_t count,
                                                const long long* __restrict__ x, long long* __restrict__ xb, unsigned nx)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = i0; i < count; i += stride) zb[i] = z[i];
    if (i0 < nx) xb[i0] = x[i0];
}

// Round-off column of the next carried squaring, for slots [slot0, slot0 + n).
__global__ void k_set_col(unsigned* epoch, unsigned slot0, unsigned n, unsigned col)
{
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) epoch[4 * (size_t)(slot0 + i) + 2] = col;
}

__global__ void k_fill_u64(unsigned long long* a, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = 0ull;
}

// Round-off of every squaring of a chunk: column i of slot s is the max of
// the blocks' partials (non-negative doubles compare as their bit patterns).
// One block of 1024 threads per slot: thread t takes column t % 64 of the
// blocks t/64, t/64 + 16, ... (partials [slot][block][column], so a wave
// reads 512 contiguous bytes), then the 16 partial maxima of each column
// meet in LDS.
__global__ __launch_bounds__(1024) void k_err_reduce(const unsigned long long* __restrict__ part,
                                                     unsigned long long* __restrict__ out, unsigned blocks,
                                                     unsigned slot0)
{
    constexpr int K = Group::kChunk, GROUPS = 1024 / K;
    __shared__ unsigned long long red[GROUPS][K];
    const unsigned s = slot0 + blockIdx.x, i = threadIdx.x % K, g0 = threadIdx.x / K;
    const unsigned long long* src = part + (size_t)s * blocks * K + i;
    unsigned long long m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    unsigned b = g0;
    for (; b + 3 * GROUPS < blocks; b += 4 * GROUPS) {
        m0 = max(m0, src[(size_t)b * K]);
        m1 = max(m1, src[(size_t)(b + GROUPS) * K]);
        m2 = max(m2, src[(size_t)(b + 2 * GROUPS) * K]);
        m3 = max(m3, src[(size_t)(b + 3 * GROUPS) * K]);
    }
    for (; b < blocks; b += GROUPS) m0 = max(m0, src[(size_t)b * K]);
    red[g0][i] = max(max(m0, m1), max(m2, m3));
    __syncthreads();
    if (g0 == 0) {
        unsigned long long m = red[0][i];
#pragma unroll
        for (int g = 1; g < GROUPS; ++g) m = max(m, red[g][i]);
        out[(size_t)s * K + i] = m;
    }
}

// Plain streaming copy of n 16-byte elements (grid-stride): reads and writes
// exactly 16*n bytes, the reference point for bandwidth and byte counters.
__global__ __launch_bounds__(256) void k_copy_probe(const double2* __restrict__ a, double2* __restrict__ b, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

// End of a chunk of scheduled steps: a slot that has carried total-1
// squarings stops (its last squaring is carried by the leave sequence).
__global__ void k_advance(SlotState* st, const SlotDesc* desc, unsigned slots, unsigned g)
{
    // g steps at once: the slot stops at the first step after which it has
    // carried total - 1 squarings
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots || !st[i].active) return;
    const unsigned long long c0 = st[i].carried, total = desc[i].total;
    if (c0 + 1 + g >= total) {
        st[i].carried = total - 1;
        st[i].active = 0;
    } else {
        st[i].carried = c0 + g;
    }
}

// ------------------------------------------------------------- host side

// One largest-radix butterfly per thread per stage, within [64, 512] (at
// 1024 threads the 128-register limit would spill the register twiddles).
constexpr int tpb_for(int points, int radix = PL_MAX_RADIX)
{
    return points / radix < 64 ? 64 : (points / radix > 512 ? 512 : points / radix);
}

// The residue's top scalar X_hi (|X_hi| <~ k/2) enters the spectrum as
// 2i*X_hi*Z, which scales the transform's relative round-off of Z by k:
// about k * 2^(b-1) * log2(L) * 2^-53 in every coefficient.  Keeping that
// under ~0.05 bounds b by 44.5 - log2(k); it binds only for k above 2^26.
double max_bits_per_digit(uint32_t logL, uint64_t k)
{
    const double table = logL <= 14 ? 18.5 : logL <= 17 ? 18.0 : logL <= 20 ? 17.5 : 17.0;
    const double by_k = 44.5 - std::log2((double)(k ? k : 1));
    return std::min(table, by_k);
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

// Digits a run's carry-in can reach (see ripple), bounded from the
// worst-case coefficient magnitude L*2^(2b+1): the first fold's carry-in is
// below 2^(b + 4 + log2 L + log2 k + 1), and the normalised second-fold
// adjustment is k*mult times that.
static int ripple_digits(uint32_t logL, double bmin, uint64_t kmax, uint32_t mult_max)
{
    const unsigned bq = (unsigned)bmin;
    const unsigned kb = bitlen64(kmax);
    const unsigned c2bits = bq + 1 + 4 + logL + kb + 1;
    const unsigned a3bits = c2bits + kb + bitlen64(mult_max) + 1;
    return (int)((a3bits + 1 + bq - 1) / bq) + 1;
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
        if ((double)n / (double)force_L > max_bits_per_digit(logL, k) + 1.0) {
            why = "fft_len too short for n (round-off)";
            return false;
        }
    } else {
        for (logL = 10; logL <= 22; ++logL)
            if ((double)n / (double)(1u << logL) <= max_bits_per_digit(logL, k)) break;
        if (logL > 22) {
            why = "n too large (needs fft_len > 2^22)";
            return false;
        }
    }
    const uint32_t L = 1u << logL;
    if (n < 7u * L) {
        why = "n too small: needs at least 7 bits per digit (n >= 7*fft_len)";
        return false;
    }
    out.n = n;
    out.k = k;
    out.L = L;
    out.logL = logL;
    // Column DFTs of at most 256 points (more, smaller column blocks: C/8
    // workgroups) unless the row DFT would pass 4096 points, and of R =
    // 2^(logL/2 - 1) for L = 2^14..2^16.  Measured on MI355X (us per squaring,
    // one-launch outer pass): L = 2^14 64x256 10.9 vs 128x128 12.2; 2^15
    // 64x512 10.7 vs 128x256 12.1; 2^16 128x512 12.3 vs 256x256 13.7; 2^17
    // 256x512 14.5 vs 128x1024 14.7; 2^19 256x2048 24.5 vs 128x4096 27.0 and
    // 512x1024 28.0; 2^20 256x4096 30.9 vs 512x2048 37.5.
    const uint32_t logR = logL >= 14 && logL <= 16
                              ? logL / 2 - 1
                              : std::max<uint32_t>(logL > 12 ? logL - 12 : 0, std::min<uint32_t>(8, logL / 2));
    out.R = 1u << logR;
    out.C = L / out.R;
    out.s = (logL + 1) / 2;
    out.nt = 0;
    out.run = 16;
    return true;
}

std::vector<uint32_t> shape_options(uint32_t logL, bool split)
{
    std::vector<uint32_t> out;
    for (uint32_t lr = 5; lr <= 10 && lr < logL; ++lr) {
        const uint32_t lc = logL - lr;
        if (lc < 5 || lc > 12) continue;
        if (split && (lr < 6 || lr > 9)) continue;
        out.push_back(lr);
    }
    return out;
}

void set_shape(Plan& p, uint32_t logR)
{
    p.R = 1u << logR;
    p.C = p.L / p.R;
}

bool choose_layout(Plan& p, double bmin, uint64_t kmax, uint32_t mult_max, bool allow_split, std::string& why)
{
    const int w = ripple_digits(p.logL, bmin, kmax, mult_max);
    if (w > 16) {
        why = "carries of this candidate would not settle within 16 digits";
        return false;
    }
    p.run = w <= 8 ? 8 : 16;
    p.nt = 0;
    if (allow_split && p.R >= 64 && p.R <= 512) p.nt = w <= 8 ? 8 : 16;
    return true;
}

template <int R>
struct ColCfg {
    static constexpr int NT = R <= 1024 ? 8 : 4;
    static constexpr int TPB = tpb_for(NT * R);
};

template <int R, int NT, int MAXR = 8>
struct SplitCfg {
    static constexpr int TPB = tpb_for(NT * R, MAXR);
};

// Row blocks of 512 points (NT = 512 / C rows for short rows): more, smaller
// workgroups than 1024-point blocks, 0.6-0.9 us per squaring faster at
// C <= 512 on MI355X and the same for longer rows.
template <int C, int RADIX = 8>
struct RowCfg {
    static constexpr int NT = (512 / C) > 1 ? (512 / C) : 1;
    static constexpr int TPB = tpb_for(NT * C, RADIX);
};

enum Kind {
    kFwdCol = 0,
    kRow = 1,
    kInvCol = 2,
    kCarry = 3,
    kFixup = 4,
    kInvCarry = 5,
    kAdjFwd = 6,
    kCarryTurn = 7,
    kAdvance = 8,
    kAdjOut = 9
};

// The steady-state outer kernels (one-launch pass, split pass) at the
// column radix MAXR; k_adjust_out transforms nothing and has one form.
template <int R, int NT, int MAXR>
static void launch_outer(int kind, const Params& p, uint32_t slots, hipStream_t st)
{
    constexpr int TPB = SplitCfg<R, NT, MAXR>::TPB;
    const dim3 grid(p.C / NT, slots);
    if (kind == kInvCarry) {
        hipLaunchKernelGGL((k_inv_carry<R, NT, TPB, MAXR>), grid, dim3(TPB), 0, st, p);
    } else if (kind == kAdjFwd) {
        hipLaunchKernelGGL((k_adjust_fwd<R, NT, TPB, MAXR>), grid, dim3(TPB), 0, st, p);
    } else if (kind == kAdjOut) {
        constexpr int TPB8 = SplitCfg<R, NT>::TPB;
        hipLaunchKernelGGL((k_adjust_out<R, NT, TPB8>), grid, dim3(TPB8), 0, st, p);
    } else if constexpr (NT == 8) {
        if (p.per_slot || p.sched) hipLaunchKernelGGL((k_carry_turn<R, 8, TPB, true, MAXR>), grid, dim3(TPB), 0, st, p);
        else hipLaunchKernelGGL((k_carry_turn<R, 8, TPB, false, MAXR>), grid, dim3(TPB), 0, st, p);
    } else {
        throw std::runtime_error("the one-launch outer pass needs runs of 8 digits");
    }
}

template <int R>
static void launch_col(int kind, const Params& p, uint32_t slots, uint32_t nt, uint32_t radix, hipStream_t st)
{
    if (kind == kFwdCol || kind == kInvCol) {
        constexpr int NT = ColCfg<R>::NT, TPB = ColCfg<R>::TPB;
        dim3 grid(p.C / NT, slots);
        if (kind == kFwdCol) hipLaunchKernelGGL((k_fwd_col<R, NT, TPB>), grid, dim3(TPB), 0, st, p);
        else hipLaunchKernelGGL((k_inv_col<R, NT, TPB>), grid, dim3(TPB), 0, st, p);
        return;
    }
    if constexpr (R >= 64 && R <= 512) {
        if (nt == 8 && radix == 16) return launch_outer<R, 8, 16>(kind, p, slots, st);
        if (nt == 8) return launch_outer<R, 8, 8>(kind, p, slots, st);
        if (nt == 16 && radix == 16) return launch_outer<R, 16, 16>(kind, p, slots, st);
        if (nt == 16) return launch_outer<R, 16, 8>(kind, p, slots, st);
    }
    throw std::runtime_error("no split outer kernels for this column length");
}

template <int C>
static void launch_row(const Params& p, uint32_t slots, uint32_t radix, hipStream_t st)
{
    const bool multi = p.per_slot || p.sched;
    if (radix == 16) {
        constexpr int NT = RowCfg<C, 16>::NT, TPB = RowCfg<C, 16>::TPB;
        const dim3 grid(p.R / NT, slots);
        if (multi) hipLaunchKernelGGL((k_row<C, NT, TPB, true, 16>), grid, dim3(TPB), 0, st, p);
        else hipLaunchKernelGGL((k_row<C, NT, TPB, false, 16>), grid, dim3(TPB), 0, st, p);
    } else {
        constexpr int NT = RowCfg<C>::NT, TPB = RowCfg<C>::TPB;
        const dim3 grid(p.R / NT, slots);
        if (multi) hipLaunchKernelGGL((k_row<C, NT, TPB, true, 8>), grid, dim3(TPB), 0, st, p);
        else hipLaunchKernelGGL((k_row<C, NT, TPB, false, 8>), grid, dim3(TPB), 0, st, p);
    }
}

static void dispatch_col(int kind, const Params& p, uint32_t slots, uint32_t nt, uint32_t radix, hipStream_t st)
{
    switch (p.R) {
    case 32: launch_col<32>(kind, p, slots, nt, radix, st); break;
    case 64: launch_col<64>(kind, p, slots, nt, radix, st); break;
    case 128: launch_col<128>(kind, p, slots, nt, radix, st); break;
    case 256: launch_col<256>(kind, p, slots, nt, radix, st); break;
    case 512: launch_col<512>(kind, p, slots, nt, radix, st); break;
    case 1024: launch_col<1024>(kind, p, slots, nt, radix, st); break;
    default: throw std::runtime_error("unsupported column length");
    }
}

static void dispatch_row(const Params& p, uint32_t slots, uint32_t radix, hipStream_t st)
{
    switch (p.C) {
    case 32: launch_row<32>(p, slots, radix, st); break;
    case 64: launch_row<64>(p, slots, radix, st); break;
    case 128: launch_row<128>(p, slots, radix, st); break;
    case 256: launch_row<256>(p, slots, radix, st); break;
    case 512: launch_row<512>(p, slots, radix, st); break;
    case 1024: launch_row<1024>(p, slots, radix, st); break;
    case 2048: launch_row<2048>(p, slots, radix, st); break;
    case 4096: launch_row<4096>(p, slots, radix, st); break;
    default: throw std::runtime_error("unsupported row length");
    }
}

template <int C>
static void launch_row_mul(const Params& p, const double2* zo, const long long* xo, hipStream_t st)
{
    constexpr int NT = RowCfg<C>::NT, TPB = RowCfg<C>::TPB;
    hipLaunchKernelGGL((k_row_mul<C, NT, TPB>), dim3(p.R / NT, 1), dim3(TPB), 0, st, p, zo, xo);
}

static void dispatch_row_mul(const Params& p, const double2* zo, const long long* xo, hipStream_t st)
{
    switch (p.C) {
    case 32: launch_row_mul<32>(p, zo, xo, st); break;
    case 64: launch_row_mul<64>(p, zo, xo, st); break;
    case 128: launch_row_mul<128>(p, zo, xo, st); break;
    case 256: launch_row_mul<256>(p, zo, xo, st); break;
    case 512: launch_row_mul<512>(p, zo, xo, st); break;
    case 1024: launch_row_mul<1024>(p, zo, xo, st); break;
    case 2048: launch_row_mul<2048>(p, zo, xo, st); break;
    case 4096: launch_row_mul<4096>(p, zo, xo, st); break;
    default: throw std::runtime_error("unsupported row length");
    }
}

// Resident-loop variants: shape R x C, tile widths, threads, column and row
// radix.  One column tile of NTD columns and one row tile of NTR rows per
// workgroup (C/NTD = R/NTR workgroups); the carry phase takes one run of 8
// digits per thread (threads past the tile's runs wait).
struct LoopShape {
    uint32_t R, C, ntd, ntr, tpb, cr, rr;
    const void* kern;
};
#define PL_LOOP(R_, C_, NTD_, NTR_, T_, CR_, RR_) \
    { R_, C_, NTD_, NTR_, T_, CR_, RR_, reinterpret_cast<const void*>(&k_loop<R_, C_, NTD_, NTR_, T_, CR_, RR_>) }
static const LoopShape kLoopShapes[] = {
    // preferred variant of each shape first, and for each length the shape an
    // untuned context takes first (measured on MI355X)
    PL_LOOP(64, 512, 8, 1, 128, 4, 4),      // L = 2^15: 64 workgroups
    PL_LOOP(128, 512, 8, 2, 256, 4, 4),     // L = 2^16: 64 workgroups
    PL_LOOP(128, 1024, 8, 1, 256, 4, 4),    // L = 2^17: 128 workgroups
    PL_LOOP(128, 1024, 8, 1, 128, 8, 8),
    PL_LOOP(256, 512, 8, 4, 256, 8, 8),     // L = 2^17: 64 workgroups
    PL_LOOP(256, 1024, 8, 2, 256, 8, 8),    // L = 2^18: 128 workgroups (13.5 us per squaring)
    PL_LOOP(128, 2048, 16, 1, 512, 4, 4),   // L = 2^18: 128 workgroups (14.9 us)
    PL_LOOP(256, 2048, 16, 2, 512, 8, 8),   // L = 2^19: 128 workgroups
    PL_LOOP(256, 4096, 16, 1, 512, 8, 8),   // L = 2^20: 256 workgroups
};
#undef PL_LOOP
constexpr int kNumLoopShapes = (int)(sizeof(kLoopShapes) / sizeof(kLoopShapes[0]));

// Whether variant i's workgroups all fit on the device at once.
static bool loop_fits(int device, int i)
{
    const LoopShape& sh = kLoopShapes[i];
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sh.kern, (int)sh.tpb, 0) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return false;
    return (long long)per_cu * cus >= (long long)(sh.C / sh.ntd);
}

// The first variant for this shape that fits (PL_LOOP_VAR=i picks variant i
// when it has the shape); -1 when none does.
static int loop_variant(int device, uint32_t R, uint32_t C)
{
    if (const char* e = std::getenv("PL_LOOP_VAR")) {
        const int i = std::atoi(e);
        if (i >= 0 && i < kNumLoopShapes && kLoopShapes[i].R == R && kLoopShapes[i].C == C && loop_fits(device, i))
            return i;
    }
    for (int i = 0; i < kNumLoopShapes; ++i)
        if (kLoopShapes[i].R == R && kLoopShapes[i].C == C && loop_fits(device, i)) return i;
    return -1;
}

uint32_t resident_logR(int device, uint32_t L)
{
    if (const char* e = std::getenv("PL_RESIDENT"))
        if (e[0] == '0') return 0;
    check(hipSetDevice(device), "hipSetDevice");
    for (int i = 0; i < kNumLoopShapes; ++i)
        if (kLoopShapes[i].R * kLoopShapes[i].C == L && loop_fits(device, i)) {
            uint32_t l = 0;
            while ((1u << l) < kLoopShapes[i].R) ++l;
            return l;
        }
    return 0;
}

static void launch_loop(int var, const Params& p, unsigned nsq, unsigned long long* xw, hipStream_t st)
{
    if (var < 0 || var >= kNumLoopShapes) throw std::runtime_error("no resident loop for this shape");
    const LoopShape& sh = kLoopShapes[var];
    Params pa = p;
    void* args[] = { &pa, &nsq, &xw };
    check(hipLaunchKernel(sh.kern, dim3(sh.C / sh.ntd), dim3(sh.tpb), args, 0, st), "k_loop launch");
}

template <int RUN, int TPB>
static void launch_carry_fixup(const Params& p, unsigned blocks, uint32_t slots, bool carry, bool fixup, hipStream_t st)
{
    if (carry) hipLaunchKernelGGL((k_carry<RUN, TPB>), dim3(blocks, slots), dim3(TPB), 0, st, p);
    if (fixup)
        hipLaunchKernelGGL((k_fixup<RUN, TPB>), dim3((blocks + 63) / 64, slots), dim3(64), 0, st, p, blocks);
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

// W_L^(c*q*R/16) = exp(-2*pi*i*c*q/(16*C)) for c < C, q < 16.
static void host_tcq(const Plan& pl, std::vector<double2>& tcq)
{
    const long double PI = 3.141592653589793238462643383279502884L;
    tcq.resize((size_t)pl.C * 16);
    for (uint32_t c = 0; c < pl.C; ++c)
        for (uint32_t q = 0; q < 16; ++q) {
            const long double a = -2.0L * PI * (long double)((uint64_t)c * q) / (16.0L * (long double)pl.C);
            tcq[(size_t)c * 16 + q] = make_double2((double)cosl(a), (double)sinl(a));
        }
}

// Factored weights of one slot (see Params): row r and column c factors of
// a_m*w^m and conj(w^m)/(L*a_m) for digit m = r*C + c.
static void host_slot_weights(const Plan& pl, uint64_t n, double2* wrf, double2* wri, double2* wcf, double2* wci)
{
    const long double PI = 3.141592653589793238462643383279502884L;
    const long double L = (long double)pl.L;
    const uint64_t Lm = pl.L - 1;
    for (uint32_t r = 0; r < pl.R; ++r) {
        const uint64_t F = (0ull - n * (uint64_t)r * pl.C) & Lm;
        const long double a = exp2l((long double)F / L);
        const long double o = PI * (long double)((uint64_t)r * pl.C) / (2.0L * L);
        wrf[r] = make_double2((double)(a * cosl(o)), (double)(a * sinl(o)));
        wri[r] = make_double2((double)(cosl(o) / a), (double)(-sinl(o) / a));
    }
    for (uint32_t c = 0; c < pl.C; ++c) {
        const uint64_t G = (0ull - n * (uint64_t)c) & Lm;
        const long double a = exp2l((long double)G / L);
        const long double o = PI * (long double)c / (2.0L * L);
        wcf[c] = make_double2((double)(a * cosl(o)), (double)(a * sinl(o)));
        wci[c] = make_double2((double)(cosl(o) / (a * L)), (double)(-sinl(o) / (a * L)));
    }
}

static uint32_t carry_tpb_for(const Plan& p) { return p.L >= p.run * 256u * 16u ? 256 : 64; }

// The turn token of a device: held by at most one plain-squaring chunk at a
// time in this process, for as long as its one-launch outer passes are in
// flight (until the chunk has been read back).  Two one-launch passes in
// flight at once could each hold the CUs the other's last blocks need; a
// launch that has no spinning neighbour always finishes, because every other
// kernel's blocks complete without waiting and free their CUs.  Kernels of
// other processes are outside this token: their effect is bounded by the
// waits' give-up and the chunk's recomputation on the split pass.
namespace {
constexpr int kMaxTokenDevices = 64;
std::mutex g_turn_token[kMaxTokenDevices];

struct TurnToken {
    int dev = -1;
    bool held = false;
    TurnToken(int device, bool want)
    {
        if (want && device >= 0 && device < kMaxTokenDevices && g_turn_token[device].try_lock()) {
            dev = device;
            held = true;
        }
    }
    ~TurnToken()
    {
        if (held) g_turn_token[dev].unlock();
    }
    TurnToken(const TurnToken&) = delete;
    TurnToken& operator=(const TurnToken&) = delete;
};
}  // namespace

size_t Group::slot_bytes(const Plan& p)
{
    // residue (16 B/pt) + split-pass digits (4 B/pt) + run carries (24 B per
    // run) + flat-carry block carries + factored weights + scalars
    const size_t runs = p.nt ? p.L / p.nt : 0;
    const uint32_t tpb = carry_tpb_for(p);
    return (size_t)p.L * 16 + (size_t)p.L * 4 + runs * (p.nt == 8 ? 48 : 24) + (size_t)(p.L / (p.run * tpb)) * 24 +
           (size_t)(p.R + p.C) * 32 + 64 + kChunk * 8;
}

// Byte offsets of a group's buffers in its one device allocation.
struct GroupLayout {
    size_t z, car, runc, runt, epoch, rdig, xhi, xw, err, errp, st, dig, desc, state, zero_end;
    size_t tw, tlo, thi, olo, ohi, plo, phi, ilo, ihi, tcq, wrf, wri, wcf, wci, bytes;
    uint32_t carry_tpb, carry_blocks, eb;
};

static GroupLayout group_layout(const Plan& p, uint32_t cap)
{
    GroupLayout g{};
    g.carry_tpb = carry_tpb_for(p);
    g.carry_blocks = p.L / (p.run * g.carry_tpb);
    g.eb = std::max<uint32_t>(g.carry_blocks, p.nt ? p.C / p.nt : 1);
    const size_t runs = p.nt ? (size_t)p.L / p.nt : 0;
    const size_t S = (size_t)1 << p.s, H = p.L / S;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t off = 0;
    auto put = [&](size_t& at, size_t bytes) {
        at = off;
        off += al(bytes);
    };
    put(g.z, sizeof(double2) * p.L * cap);
    put(g.car, sizeof(long long) * 3 * g.carry_blocks * cap);
    put(g.runc, sizeof(long long) * 3 * runs * cap);
    put(g.runt, p.nt == 8 ? sizeof(long long) * 3 * runs * cap : 0);
    put(g.epoch, sizeof(unsigned) * 4 * cap);
    put(g.rdig, p.nt ? sizeof(int) * (size_t)p.L * cap : 0);
    put(g.xhi, sizeof(long long) * cap);
    put(g.xw, sizeof(unsigned long long) * cap);
    put(g.err, sizeof(unsigned long long) * Group::kChunk * cap);
    put(g.errp, sizeof(unsigned long long) * Group::kChunk * g.eb * cap);
    put(g.st, sizeof(int));
    put(g.dig, sizeof(int32_t) * p.L);
    put(g.desc, sizeof(SlotDesc) * cap);
    put(g.state, sizeof(SlotState) * cap);
    g.zero_end = off;
    put(g.tw, sizeof(double2) * 4096);
    put(g.tlo, sizeof(double2) * S);
    put(g.thi, sizeof(double2) * H);
    put(g.olo, sizeof(double2) * S);
    put(g.ohi, sizeof(double2) * H);
    put(g.plo, sizeof(double) * S);
    put(g.phi, sizeof(double) * H);
    put(g.ilo, sizeof(double) * S);
    put(g.ihi, sizeof(double) * H);
    put(g.tcq, sizeof(double2) * (size_t)p.C * 16);
    put(g.wrf, sizeof(double2) * p.R * cap);
    put(g.wri, sizeof(double2) * p.R * cap);
    put(g.wcf, sizeof(double2) * p.C * cap);
    put(g.wci, sizeof(double2) * p.C * cap);
    g.bytes = off;
    return g;
}

size_t Group::layout_bytes(const Plan& p, uint32_t capacity) { return group_layout(p, capacity).bytes; }

Group::Group(const Plan& p, uint32_t capacity, int device) : plan_(p), cap_(capacity), device_(device)
{
    init(nullptr, 0, nullptr);
}

Group::Group(const Plan& p, uint32_t capacity, int device, void* mem, size_t mem_bytes, hipStream_t stream)
    : plan_(p), cap_(capacity), device_(device)
{
    if (!mem) throw std::invalid_argument("external group memory is NULL");
    init(mem, mem_bytes, stream);
}

void Group::init(void* ext_mem, size_t ext_bytes, hipStream_t ext_stream)
{
    const Plan& p = plan_;
    if (cap_ == 0) throw std::invalid_argument("group capacity must be positive");
    check(hipSetDevice(device_), "hipSetDevice");
    const GroupLayout lay = group_layout(p, cap_);
    if (ext_mem) {
        if (ext_bytes < lay.bytes) throw std::invalid_argument("external group memory too small for this length");
        owns_ = false;
        stream_ = ext_stream;
        mem_ = ext_mem;
        turn_ = false;  // the one-launch pass's fallback would allocate its backup
    } else {
        // non-blocking: no implicit ordering with the legacy stream, so host
        // threads driving other groups (and capturing their graphs) never
        // depend on this one
        check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    }
    std::vector<double2> tw, twl_lo, twl_hi, om_lo, om_hi;
    std::vector<double> pw_lo, pw_hi, ipw_lo, ipw_hi;
    host_tables(p, tw, twl_lo, twl_hi, om_lo, om_hi, pw_lo, pw_hi, ipw_lo, ipw_hi);
    carry_tpb_ = lay.carry_tpb;
    carry_blocks_ = lay.carry_blocks;
    if (owns_) {
        if (const char* e = std::getenv("PL_CARRY_TURN")) turn_ = e[0] != '0';
    }
    if (const char* e = std::getenv("PL_HANDOFF_GIVE_UP")) give_up_ = e[0] == '1';
    if (const char* e = std::getenv("PL_NO_GRAPHS")) no_graphs_ = e[0] == '1';
    loop_var_ = owns_ ? loop_variant(device_, p.R, p.C) : -1;
    resident_ = loop_var_ >= 0;
    if (const char* e = std::getenv("PL_POLL_SLEEP")) poll_sleep_ = (unsigned)std::max(1, std::atoi(e));
    if (const char* e = std::getenv("PL_RESIDENT")) resident_ = resident_ && e[0] != '0';
    eb_ = lay.eb;
    const size_t o_z = lay.z, o_car = lay.car, o_runc = lay.runc, o_runt = lay.runt, o_epoch = lay.epoch;
    const size_t o_rdig = lay.rdig, o_xhi = lay.xhi, o_err = lay.err, o_errp = lay.errp, o_st = lay.st;
    const size_t o_dig = lay.dig, o_desc = lay.desc, o_state = lay.state, zero_end = lay.zero_end;
    const size_t o_tw = lay.tw, o_tlo = lay.tlo, o_thi = lay.thi, o_olo = lay.olo, o_ohi = lay.ohi;
    const size_t o_plo = lay.plo, o_phi = lay.phi, o_ilo = lay.ilo, o_ihi = lay.ihi, o_tcq = lay.tcq;
    const size_t o_wrf = lay.wrf, o_wri = lay.wri, o_wcf = lay.wcf, o_wci = lay.wci;
    std::vector<double2> tcq;
    host_tcq(p, tcq);
    bytes_ = lay.bytes;
    if (owns_) check(hipMalloc(&mem_, bytes_), "hipMalloc");
    char* b = static_cast<char*>(mem_);
    z_ = reinterpret_cast<double2*>(b + o_z);
    carries_ = reinterpret_cast<long long*>(b + o_car);
    runc_ = reinterpret_cast<long long*>(b + o_runc);
    runt_ = p.nt == 8 ? reinterpret_cast<long long*>(b + o_runt) : nullptr;
    epoch_ = reinterpret_cast<unsigned*>(b + o_epoch);
    rdig_ = reinterpret_cast<int*>(b + o_rdig);
    xhi_ = reinterpret_cast<long long*>(b + o_xhi);
    xw_ = reinterpret_cast<unsigned long long*>(b + lay.xw);
    err_ = reinterpret_cast<unsigned long long*>(b + o_err);
    errp_ = reinterpret_cast<unsigned long long*>(b + o_errp);
    status_ = reinterpret_cast<int*>(b + o_st);
    digits_dev_ = reinterpret_cast<int32_t*>(b + o_dig);
    desc_ = reinterpret_cast<SlotDesc*>(b + o_desc);
    state_ = reinterpret_cast<SlotState*>(b + o_state);
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
    tabs_->tcq = reinterpret_cast<double2*>(b + o_tcq);
    tabs_->wrf = reinterpret_cast<double2*>(b + o_wrf);
    tabs_->wri = reinterpret_cast<double2*>(b + o_wri);
    tabs_->wcf = reinterpret_cast<double2*>(b + o_wcf);
    tabs_->wci = reinterpret_cast<double2*>(b + o_wci);
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
    up(tabs_->tcq, tcq.data(), sizeof(double2) * tcq.size());
    {
        std::vector<double2> wrf((size_t)p.R * cap_), wri((size_t)p.R * cap_), wcf((size_t)p.C * cap_),
            wci((size_t)p.C * cap_);
        host_slot_weights(p, p.n, wrf.data(), wri.data(), wcf.data(), wci.data());
        for (uint32_t i = 1; i < cap_; ++i) {
            std::copy(wrf.begin(), wrf.begin() + p.R, wrf.begin() + (size_t)i * p.R);
            std::copy(wri.begin(), wri.begin() + p.R, wri.begin() + (size_t)i * p.R);
            std::copy(wcf.begin(), wcf.begin() + p.C, wcf.begin() + (size_t)i * p.C);
            std::copy(wci.begin(), wci.begin() + p.C, wci.begin() + (size_t)i * p.C);
        }
        up(tabs_->wrf, wrf.data(), sizeof(double2) * wrf.size());
        up(tabs_->wri, wri.data(), sizeof(double2) * wri.size());
        up(tabs_->wcf, wcf.data(), sizeof(double2) * wcf.size());
        up(tabs_->wci, wci.data(), sizeof(double2) * wci.size());
        check(hipStreamSynchronize(stream_), "table upload");
    }
    check(hipMemsetAsync(mem_, 0, zero_end, stream_), "hipMemsetAsync");
    check(hipStreamSynchronize(stream_), "init sync");
    desc_host_.assign(cap_, SlotDesc{ p.n, (long long)p.k, 0, 0, 0 });
    check(hipMemcpyAsync(desc_, desc_host_.data(), sizeof(SlotDesc) * cap_, hipMemcpyHostToDevice, stream_),
          "desc upload");
    check(hipStreamSynchronize(stream_), "desc sync");
    err_host_.resize((size_t)kChunk * cap_);
}

Group::~Group()
{
    // nothing launched on the stream may still use the graphs or the memory
    (void)hipStreamSynchronize(stream_);
    for (auto& pr : lev_)
        for (hipEvent_t& e : pr)
            if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; ++i) {
        if (pipe_done_[i]) (void)hipEventDestroy(pipe_done_[i]);
        if (pipe_err_[i]) (void)hipHostFree(pipe_err_[i]);
    }
    if (pipe_mem_) (void)hipFree(pipe_mem_);
    for (auto& kv : plain_graphs_) {
        (void)hipGraphExecDestroy(kv.second.first);
        (void)hipGraphDestroy(kv.second.second);
    }
    for (auto& kv : step_graphs_) {
        (void)hipGraphExecDestroy(kv.second.first);
        (void)hipGraphDestroy(kv.second.second);
    }
    if (owns_) {
        if (mem_) (void)hipFree(mem_);
        if (zbak_) (void)hipFree(zbak_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    delete tabs_;
}

Params Group::params(uint32_t slot0, bool check_active) const
{
    Params p;
    p.z = z_;
    p.tw = tabs_->tw;
    p.twl_lo = tabs_->twl_lo;
    p.twl_hi = tabs_->twl_hi;
    p.pw_lo = tabs_->pw_lo;
    p.pw_hi = tabs_->pw_hi;
    p.ipw_lo = tabs_->ipw_lo;
    p.ipw_hi = tabs_->ipw_hi;
    p.om_lo = tabs_->om_lo;
    p.om_hi = tabs_->om_hi;
    p.tcq = tabs_->tcq;
    p.wrf = tabs_->wrf;
    p.wri = tabs_->wri;
    p.wcf = tabs_->wcf;
    p.wci = tabs_->wci;
    p.carries = carries_;
    p.runc = runc_;
    p.epoch = epoch_;
    p.dig = rdig_;
    p.xhi = xhi_;
    p.xhi_mul = nullptr;
    p.err = errp_;
    p.status = status_;
    p.desc = desc_;
    p.state = state_;
    p.carr_stride = 3 * (size_t)carry_blocks_;
    p.runc_stride = plan_.nt ? 3 * (size_t)(plan_.L / plan_.nt) : 0;
    p.err_stride = kChunk * eb_;
    p.slot0 = slot0;
    p.sched = check_active ? 1 : 0;
    p.spin_limit = give_up_ ? 64u : kSpinLimit;
    p.tag_skew = give_up_ ? 1u : 0u;
    p.per_slot = uniform_ ? 0 : 1;
    p.ztag = 0;
    p.poll_sleep = poll_sleep_;
    p.step_off = 0;
    p.n = plan_.n;
    p.k = (long long)plan_.k;
    p.mult = 1;
    p.active = 1;
    p.L = plan_.L;
    p.logL = plan_.logL;
    p.R = plan_.R;
    p.C = plan_.C;
    p.s = plan_.s;
    p.smask = (1u << plan_.s) - 1;
    return p;
}

// One kernel kind over `slots` slots starting at p.slot0.
void Group::launch(int kind, const Params& p, uint32_t slots)
{
    launched_ = true;
    switch (kind) {
    case kRow: dispatch_row(p, slots, plan_.row_radix, stream_); break;
    case kCarry:
    case kFixup: {
        const bool c = kind == kCarry, f = kind == kFixup;
        if (plan_.run == 8 && carry_tpb_ == 256) launch_carry_fixup<8, 256>(p, carry_blocks_, slots, c, f, stream_);
        else if (plan_.run == 8) launch_carry_fixup<8, 64>(p, carry_blocks_, slots, c, f, stream_);
        else if (carry_tpb_ == 256) launch_carry_fixup<16, 256>(p, carry_blocks_, slots, c, f, stream_);
        else launch_carry_fixup<16, 64>(p, carry_blocks_, slots, c, f, stream_);
        break;
    }
    case kAdvance:
        hipLaunchKernelGGL(k_advance, dim3((slots + 63) / 64), dim3(64), 0, stream_, state_, desc_, slots,
                           p.step_off);
        break;
    default: dispatch_col(kind, p, slots, plan_.nt, plan_.col_radix, stream_);
    }
}

void Group::set_slot(uint32_t slot, uint32_t n, uint64_t k, uint32_t a, uint64_t total)
{
    if (slot >= cap_) throw std::invalid_argument("slot out of range");
    if (n >> plan_.logL < 7 || (double)n / plan_.L > max_bits_per_digit(plan_.logL, k) + 1.0)
        throw std::invalid_argument("n does not suit this group's transform length");
    check(hipSetDevice(device_), "hipSetDevice");
    unsigned top = 0;
    while (top < 63 && (k >> (top + 1))) ++top;
    desc_host_[slot] = SlotDesc{ n, (long long)k, a, top, total };
    if (uniform_ && (n != plan_.n || k != plan_.k)) {
        // plain-squaring graphs were captured reading n, k from the launch
        // parameters; from now on every launch reads each slot's descriptor
        uniform_ = false;
        for (auto& kv : plain_graphs_) {
            (void)hipGraphExecDestroy(kv.second.first);
            (void)hipGraphDestroy(kv.second.second);
        }
        plain_graphs_.clear();
    }
    check(hipMemcpyAsync(desc_ + slot, &desc_host_[slot], sizeof(SlotDesc), hipMemcpyHostToDevice, stream_),
          "desc upload");
    std::vector<double2> wrf(plan_.R), wri(plan_.R), wcf(plan_.C), wci(plan_.C);
    host_slot_weights(plan_, n, wrf.data(), wri.data(), wcf.data(), wci.data());
    auto up = [&](double2* dst, const std::vector<double2>& v) {
        check(hipMemcpyAsync(dst, v.data(), sizeof(double2) * v.size(), hipMemcpyHostToDevice, stream_),
              "weight upload");
    };
    up(tabs_->wrf + (size_t)slot * plan_.R, wrf);
    up(tabs_->wri + (size_t)slot * plan_.R, wri);
    up(tabs_->wcf + (size_t)slot * plan_.C, wcf);
    up(tabs_->wci + (size_t)slot * plan_.C, wci);
    check(hipStreamSynchronize(stream_), "desc sync");
}

void Group::set_state(uint32_t slot, uint64_t carried, bool active)
{
    SlotState s{ carried, active ? 1 : 0, 0 };
    check(hipMemcpyAsync(state_ + slot, &s, sizeof s, hipMemcpyHostToDevice, stream_), "state upload");
    check(hipStreamSynchronize(stream_), "state sync");
}

void Group::upload(uint32_t slot, const std::vector<int32_t>& digits, int64_t xhi)
{
    check(hipSetDevice(device_), "hipSetDevice");
    if (digits.size() != plan_.L) throw std::runtime_error("digit count mismatch");
    long long xh = xhi;
    check(hipMemcpyAsync(digits_dev_, digits.data(), sizeof(int32_t) * plan_.L, hipMemcpyHostToDevice, stream_),
          "upload digits");
    check(hipMemcpyAsync(xhi_ + slot, &xh, sizeof(long long), hipMemcpyHostToDevice, stream_), "upload xhi");
    Params p = params(slot, false);
    hipLaunchKernelGGL(k_set_digits, dim3((plan_.L + 255) / 256, 1), dim3(256), 0, stream_, p, digits_dev_);
    check(hipGetLastError(), "k_set_digits");
    check(hipStreamSynchronize(stream_), "upload sync");
}

void Group::download(uint32_t slot, std::vector<int32_t>& digits, int64_t& xhi)
{
    check(hipSetDevice(device_), "hipSetDevice");
    Params p = params(slot, false);
    hipLaunchKernelGGL(k_get_digits, dim3((plan_.L + 255) / 256, 1), dim3(256), 0, stream_, p, digits_dev_);
    check(hipGetLastError(), "k_get_digits");
    digits.resize(plan_.L);
    long long xh = 0;
    check(hipMemcpyAsync(digits.data(), digits_dev_, sizeof(int32_t) * plan_.L, hipMemcpyDeviceToHost, stream_),
          "download digits");
    check(hipMemcpyAsync(&xh, xhi_ + slot, sizeof(long long), hipMemcpyDeviceToHost, stream_), "download xhi");
    check(hipStreamSynchronize(stream_), "download sync");
    xhi = xh;
}

// g squarings of every slot from the digit state back to the digit state;
// iteration i's round-off lands in err column i.
// g plain squarings of slots [sq0_, sq0_ + sqn_).
void Group::enqueue_plain_chunk(int g, bool use_turn, int parity)
{
    const uint32_t s0 = sq0_, ns = sqn_;
    if (use_turn) {
        const size_t count = (size_t)plan_.L * ns;
        const unsigned blocks = (unsigned)std::min<size_t>((count + 255) / 256, 2048);
        double2* bz = parity < 0 ? zbak_ : pipe_z_[parity];
        long long* bx = parity < 0 ? xbak_ : pipe_x_[parity];
        hipLaunchKernelGGL(k_backup, dim3(blocks), dim3(256), 0, stream_, z_ + (size_t)s0 * plan_.L, bz, count,
                           xhi_ + s0, bx, ns);
    }
    launch(kFwdCol, params(s0, false), ns);
    if (use_turn && g >= 2 && resident_on()) {
        // the steady state as one resident launch: the entry row pass stores
        // ready-tagged points, k_loop runs the g-1 steady squarings
        Params pe = params(s0, false);
        pe.ztag = 1;
        launch(kRow, pe, ns);
        Params pt = params(s0, false);
        pt.runc = runt_;
        launched_ = true;
        ensure_loop_events();
        const int ev = parity < 0 ? 0 : parity;
        check(hipEventRecord(lev_[ev][0], stream_), "hipEventRecord");
        launch_loop(loop_var_, pt, (unsigned)(g - 1), xw_ + s0, stream_);
        check(hipEventRecord(lev_[ev][1], stream_), "hipEventRecord");
        enqueue_exit(params(s0, false), ns);
        return;
    }
    launch(kRow, params(s0, false), ns);
    for (int i = 0; i + 1 < g; ++i) {
        const Params p = params(s0, false);
        if (use_turn) {
            Params pt = p;
            pt.runc = runt_;
            launch(kCarryTurn, pt, ns);
        } else if (plan_.nt) {
            launch(kInvCarry, p, ns);
            launch(kAdjFwd, p, ns);
        } else {
            launch(kInvCol, p, ns);
            launch(kCarry, p, ns);
            launch(kFixup, p, ns);
            launch(kFwdCol, p, ns);
        }
        launch(kRow, p, ns);
    }
    enqueue_exit(params(s0, false), ns);
}

// Back to the digit state after a squaring's row pass: the split outer
// pass's first kernel and k_adjust_out where the split pass exists (2
// kernels), else inverse column pass, flat carry and fix-up (3).
void Group::enqueue_exit(const Params& p, uint32_t slots)
{
    if (plan_.nt) {
        launch(kInvCarry, p, slots);
        launch(kAdjOut, p, slots);
    } else {
        launch(kInvCol, p, slots);
        launch(kCarry, p, slots);
        launch(kFixup, p, slots);
    }
}

// g scheduled steps of every active slot; step i's round-off in err column i.
// Step chunks take the split outer pass: the pool overlaps them on several
// streams, and a single test's seed powering is a few dozen steps.
void Group::enqueue_step_chunk(int g)
{
    for (int i = 0; i < g; ++i) {
        Params p = params(0, true);
        p.step_off = (unsigned)i;
        if (plan_.nt) {
            launch(kInvCarry, p, step_n_);
            launch(kAdjFwd, p, step_n_);
        } else {
            launch(kInvCol, p, step_n_);
            launch(kCarry, p, step_n_);
            launch(kFixup, p, step_n_);
            launch(kFwdCol, p, step_n_);
        }
        launch(kRow, p, step_n_);
    }
    // the slots' states move on once for the whole chunk
    Params pa = params(0, true);
    pa.step_off = (unsigned)g;
    launch(kAdvance, pa, step_n_);
}

void Group::build_graph(hipGraphExec_t& exec, hipGraph_t& graph, bool stepping, int g, bool use_turn)
{
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    const uint32_t ns = stepping ? step_n_ : cap_;
    enqueue_err_clear(ns, stepping ? 0u : ~0u);
    if (stepping) enqueue_step_chunk(g);
    else enqueue_plain_chunk(g, use_turn);
    enqueue_err_reduce(0, ns);
    check(hipStreamEndCapture(stream_, &graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

// Zero the round-off partials of slots [0, slots) and set their round-off
// column to col0 (fill kernels, so the chunk can be graph-captured).  A plain
// chunk starts at column -1: its entry row pass advances it to 0.
void Group::enqueue_err_clear(uint32_t slots, unsigned col0)
{
    const int cnt = (int)((size_t)kChunk * eb_ * slots);
    hipLaunchKernelGGL(k_fill_u64, dim3((cnt + 255) / 256), dim3(256), 0, stream_, errp_, cnt);
    hipLaunchKernelGGL(k_set_col, dim3((slots + 63) / 64), dim3(64), 0, stream_, epoch_, 0u, slots, col0);
}

void Group::enqueue_set_col(uint32_t slot, unsigned col)
{
    hipLaunchKernelGGL(k_set_col, dim3(1), dim3(64), 0, stream_, epoch_, slot, 1u, col);
}

void Group::enqueue_err_reduce(uint32_t slot0, uint32_t slots)
{
    hipLaunchKernelGGL(k_err_reduce, dim3(slots), dim3(1024), 0, stream_, errp_, err_, eb_, slot0);
}

// Reads the chunk's err columns: trace gets slot 0's per-column values,
// slot_err each slot's max; returns the max over everything.
double Group::read_err(int g, double* trace, double* slot_err, uint32_t trace_slot, uint32_t nslots)
{
    const uint32_t live = nslots ? nslots : cap_;
    check(hipMemcpyAsync(err_host_.data(), err_, sizeof(unsigned long long) * kChunk * cap_, hipMemcpyDeviceToHost,
                         stream_),
          "err D2H");
    sync_status("chunk sync");
    double emax = 0.0;
    for (uint32_t s = 0; s < cap_; ++s) {
        double smax = 0.0;
        for (int i = 0; i < g && s < live; ++i) {
            double e;
            std::memcpy(&e, &err_host_[(size_t)s * kChunk + i], sizeof e);
            if (s == trace_slot && trace) trace[i] = e;
            smax = std::max(smax, e);
        }
        if (slot_err) slot_err[s] = smax;
        emax = std::max(emax, smax);
    }
    return emax;
}

double Group::square(uint64_t iters, double* err_trace) { return square_slots(0, cap_, iters, err_trace); }

// One plain chunk of g squarings of slots [sq0_, sq0_ + sqn_): a captured
// graph for whole chunks, direct launches for a partial one.  (Partial
// chunks as a few captured pieces -- entry, blocks of 128, 64, ..., 1
// steady squarings, exit -- measured slower on MI355X than the launches.)
void Group::run_plain_chunk(int g, bool use_turn)
{
    // A resident chunk is a handful of launches around one long one, so it
    // gains nothing from a graph, and the events around its resident launch
    // must be recorded on the stream itself to be timed.
    const bool resident = use_turn && g >= 2 && resident_on();
    if (g == kChunk && !no_graphs_ && !resident) {
        const uint64_t key = ((uint64_t)sq0_ << 33) | ((uint64_t)sqn_ << 1) | (use_turn ? 1u : 0u);
        auto it = plain_graphs_.find(key);
        if (it == plain_graphs_.end()) {
            hipGraphExec_t ex = nullptr;
            hipGraph_t gr = nullptr;
            build_graph(ex, gr, false, kChunk, use_turn);
            it = plain_graphs_.emplace(key, std::make_pair(ex, gr)).first;
        }
        check(hipGraphLaunch(it->second.first, stream_), "hipGraphLaunch");
        launched_ = true;
    } else {
        enqueue_err_clear(cap_, ~0u);
        enqueue_plain_chunk(g, use_turn);
        enqueue_err_reduce(0, cap_);
        check(hipGetLastError(), "square launch");
    }
}

// The backup of a one-launch chunk's start (allocated on first use, before
// any graph that copies into it is captured; the copy itself is the chunk's
// first kernel).
void Group::ensure_backup()
{
    if (zbak_) return;
    void* m = nullptr;
    check(hipMalloc(&m, sizeof(double2) * plan_.L * cap_ + 256 + sizeof(long long) * cap_), "hipMalloc backup");
    zbak_ = static_cast<double2*>(m);
    xbak_ = reinterpret_cast<long long*>(static_cast<char*>(m) + sizeof(double2) * plan_.L * cap_ + 256);
}

// Events around each resident launch (created before any graph that records
// them is captured; no system-scope fence, which would add a cache
// write-back the chunk itself does not do).
void Group::ensure_loop_events()
{
    for (auto& pr : lev_)
        for (hipEvent_t& e : pr)
            if (!e) check(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "hipEventCreate");
}

// Buffers of the pipelined resident chunks: per parity a backup of the
// chunk's start (one slot), a pinned host copy of its round-off columns and
// status word, and an event marking its end.
void Group::ensure_pipeline()
{
    ensure_loop_events();
    if (pipe_mem_) return;
    const size_t zb = sizeof(double2) * plan_.L, per = zb + 256;
    check(hipMalloc(&pipe_mem_, 2 * per), "hipMalloc pipeline backup");
    for (int i = 0; i < 2; ++i) {
        char* b = static_cast<char*>(pipe_mem_) + (size_t)i * per;
        pipe_z_[i] = reinterpret_cast<double2*>(b);
        pipe_x_[i] = reinterpret_cast<long long*>(b + zb);
        check(hipHostMalloc(reinterpret_cast<void**>(&pipe_err_[i]), sizeof(unsigned long long) * kChunk * cap_ + 64),
              "hipHostMalloc");
        pipe_st_[i] = reinterpret_cast<int*>(pipe_err_[i] + (size_t)kChunk * cap_);
        check(hipEventCreateWithFlags(&pipe_done_[i], hipEventDisableTiming), "hipEventCreate");
    }
}

void Group::loop_time(bool reset, double* ms, uint64_t* launches, uint64_t* squarings)
{
    if (ms) *ms = loop_ms_;
    if (launches) *launches = loop_launches_;
    if (squarings) *squarings = loop_squarings_;
    if (reset) {
        loop_ms_ = 0;
        loop_launches_ = loop_squarings_ = 0;
    }
}

// After a launch whose hand-off gave up: the tagged words may hold tags of
// its unfinished range, which the next launch would reuse (a resident launch
// advances the slot's epoch only when it completes).  Moving the epoch 256
// on puts the next launches' tags clear of every tag a word can hold (the
// last completed chunk's and the failed range, each under 256 wide).
void Group::skip_epochs(uint32_t slot)
{
    unsigned ep[4];
    check(hipMemcpyAsync(ep, epoch_ + 4 * (size_t)slot, sizeof ep, hipMemcpyDeviceToHost, stream_), "epoch D2H");
    check(hipStreamSynchronize(stream_), "epoch sync");
    const unsigned e = (ep[0] + (unsigned)kChunk) & kTagMask;
    ep[1] = (e + kTagMask) & kTagMask;
    ep[0] = e;
    check(hipMemcpyAsync(epoch_ + 4 * (size_t)slot, ep, sizeof ep, hipMemcpyHostToDevice, stream_), "epoch H2D");
    check(hipStreamSynchronize(stream_), "epoch sync");
}

void Group::restore_start(uint32_t slot0, uint32_t slots)
{
    check(hipMemcpyAsync(z_ + (size_t)slot0 * plan_.L, zbak_, sizeof(double2) * plan_.L * slots,
                         hipMemcpyDeviceToDevice, stream_),
          "restore z");
    check(hipMemcpyAsync(xhi_ + slot0, xbak_, sizeof(long long) * slots, hipMemcpyDeviceToDevice, stream_),
          "restore xhi");
}

// Whole resident chunks of slot sq0_, pipelined: chunk i+1 is enqueued
// before chunk i is waited for, so the device never waits for the host
// between chunks.  Each chunk saves its start in its parity's backup and
// copies its round-off columns and status word to its parity's pinned
// buffers.  Runs chunks of at least 2 squarings while `iters - done` allows
// and the token is held by the caller; returns the squarings done (the
// caller finishes any rest).  A chunk whose hand-off gave up is recomputed
// from its start on the split pass, and the chunk after it is discarded
// and left to the caller.
uint64_t Group::square_pipelined(uint64_t iters, double* err_trace, double& emax)
{
    ensure_pipeline();
    const uint32_t s0 = sq0_;
    uint64_t done = 0, queued = 0;
    int inflight[2] = { 0, 0 };  // squarings of the chunk in flight per parity (0: none)
    uint64_t start[2] = { 0, 0 };
    int next = 0;
    auto enqueue = [&](int par, int g) {
        enqueue_err_clear(cap_, ~0u);
        enqueue_plain_chunk(g, true, par);
        enqueue_err_reduce(0, cap_);
        check(hipGetLastError(), "square launch");
        check(hipMemcpyAsync(pipe_err_[par], err_, sizeof(unsigned long long) * kChunk * cap_, hipMemcpyDeviceToHost,
                             stream_),
              "err D2H");
        check(hipMemcpyAsync(pipe_st_[par], status_, sizeof(int), hipMemcpyDeviceToHost, stream_), "status D2H");
        check(hipEventRecord(pipe_done_[par], stream_), "hipEventRecord");
        launched_ = true;
        inflight[par] = g;
        start[par] = queued;
        queued += (uint64_t)g;
    };
    auto room = [&]() { return iters - queued >= 2 ? (int)std::min<uint64_t>(kChunk, iters - queued) : 0; };
    if (const int g = room()) enqueue(next, g);
    while (inflight[next]) {
        const int par = next, other = 1 - next;
        if (!inflight[other])
            if (const int g = room()) enqueue(other, g);
        check(hipEventSynchronize(pipe_done_[par]), "chunk sync");
        const int g = inflight[par];
        inflight[par] = 0;
        const int st = *pipe_st_[par];
        if (st & 2) {
            // a hand-off wait gave up: drain, recompute this chunk from its
            // start on the split pass; the chunk after it is void
            check(hipStreamSynchronize(stream_), "chunk sync");
            clear_status();
            check(hipMemcpyAsync(z_ + (size_t)s0 * plan_.L, pipe_z_[par], sizeof(double2) * plan_.L,
                                 hipMemcpyDeviceToDevice, stream_),
                  "restore z");
            check(hipMemcpyAsync(xhi_ + s0, pipe_x_[par], sizeof(long long), hipMemcpyDeviceToDevice, stream_),
                  "restore xhi");
            skip_epochs(s0);
            run_plain_chunk(g, false);
            double* tr = err_trace ? err_trace + start[par] : nullptr;
            emax = std::max(emax, read_err(g, tr, nullptr, s0));
            ++fallbacks_;
            return start[par] + (uint64_t)g;
        }
        seen_status_ |= st;
        float ms = 0;
        check(hipEventElapsedTime(&ms, lev_[par][0], lev_[par][1]), "hipEventElapsedTime");
        loop_ms_ += ms;
        ++loop_launches_;
        loop_squarings_ += (uint64_t)(g - 1);
        for (int i = 0; i < g; ++i) {
            double e;
            std::memcpy(&e, &pipe_err_[par][(size_t)s0 * kChunk + i], sizeof e);
            if (err_trace) err_trace[start[par] + (uint64_t)i] = e;
            emax = std::max(emax, e);
        }
        done = start[par] + (uint64_t)g;
        next = other;
    }
    launched_ = false;
    return done;
}

double Group::square_slots(uint32_t slot0, uint32_t slots, uint64_t iters, double* err_trace)
{
    if (slots == 0 || slot0 + slots > cap_) throw std::invalid_argument("square_slots: slot range out of bounds");
    check(hipSetDevice(device_), "hipSetDevice");
    sq0_ = slot0;
    sqn_ = slots;
    double emax = 0.0;
    uint64_t done = 0;
    if (iters >= 2 && resident_on()) {
        TurnToken token(device_, true);
        if (token.held) done = square_pipelined(iters, err_trace, emax);
    }
    while (done < iters) {
        const int g = (int)std::min<uint64_t>(kChunk, iters - done);
        double* tr = err_trace ? err_trace + done : nullptr;
        // one-launch passes need this device's turn token and g >= 2 (a
        // chunk of one squaring has no steady-state squaring)
        TurnToken token(device_, turn() && g >= 2);
        if (token.held) {
            ensure_backup();
            run_plain_chunk(g, true);
            double e = read_err(g, tr, nullptr, slot0);
            if (g >= 2 && resident_on() && !(seen_status_ & 2)) {
                // the resident launch's own duration, from the events around it
                float ms = 0;
                check(hipEventElapsedTime(&ms, lev_[0][0], lev_[0][1]), "hipEventElapsedTime");
                loop_ms_ += ms;
                ++loop_launches_;
                loop_squarings_ += (uint64_t)(g - 1);
            }
            if (seen_status_ & 2) {
                // a hand-off wait gave up (another process's kernels held the
                // CUs its predecessor needed, or the test hook): the chunk's
                // result is void; recompute it from its start on the split pass
                clear_status();
                restore_start(slot0, slots);
                for (uint32_t i = 0; i < slots; ++i) skip_epochs(slot0 + i);
                run_plain_chunk(g, false);
                e = read_err(g, tr, nullptr, slot0);
                ++fallbacks_;
            }
            emax = std::max(emax, e);
        } else {
            run_plain_chunk(g, false);
            emax = std::max(emax, read_err(g, tr, nullptr, slot0));
        }
        done += g;
    }
    return emax;
}

void Group::copy_slot(uint32_t dst, uint32_t src)
{
    if (dst >= cap_ || src >= cap_) throw std::invalid_argument("copy_slot: slot out of range");
    if (dst == src) return;
    check(hipSetDevice(device_), "hipSetDevice");
    check(hipMemcpyAsync(z_ + (size_t)dst * plan_.L, z_ + (size_t)src * plan_.L, sizeof(double2) * plan_.L,
                         hipMemcpyDeviceToDevice, stream_),
          "copy_slot z");
    check(hipMemcpyAsync(xhi_ + dst, xhi_ + src, sizeof(long long), hipMemcpyDeviceToDevice, stream_), "copy_slot xhi");
    check(hipStreamSynchronize(stream_), "copy_slot sync");
}

// dst <- dst*src mod N: forward column pass of both, the two-operand row
// pass, and the inverse column pass + carry with the product's top rule.
double Group::mul_slot(uint32_t dst, uint32_t src)
{
    if (dst >= cap_ || src >= cap_ || dst == src) throw std::invalid_argument("mul_slot: bad slots");
    check(hipSetDevice(device_), "hipSetDevice");
    {
        const int cnt = (int)((size_t)kChunk * eb_);
        hipLaunchKernelGGL(k_fill_u64, dim3((cnt + 255) / 256), dim3(256), 0, stream_, errp_ + (size_t)dst * cnt, cnt);
        enqueue_set_col(dst, 0u);
    }
    launch(kFwdCol, params(dst, false), 1);
    launch(kFwdCol, params(src, false), 1);
    Params p = params(dst, false);
    dispatch_row_mul(p, z_ + (size_t)src * plan_.L, xhi_ + src, stream_);
    p.xhi_mul = xhi_ + src;
    enqueue_exit(p, 1);
    enqueue_err_reduce(dst, 1);
    check(hipGetLastError(), "mul_slot");
    unsigned long long e = 0;
    check(hipMemcpyAsync(&e, err_ + (size_t)dst * kChunk, sizeof e, hipMemcpyDeviceToHost, stream_), "err D2H");
    sync_status("mul_slot sync");
    double d;
    std::memcpy(&d, &e, sizeof d);
    return d;
}

void Group::enter(uint32_t slot)
{
    check(hipSetDevice(device_), "hipSetDevice");
    const Params p = params(slot, false);
    launch(kFwdCol, p, 1);
    launch(kRow, p, 1);
    check(hipGetLastError(), "enter");
}

void Group::step_begin(uint32_t g, uint32_t slots)
{
    if (g == 0 || g > (uint32_t)kStepChunk) throw std::invalid_argument("step_begin: 1..kStepChunk steps");
    if (pending_steps_) throw std::logic_error("step_begin: the previous chunk was not read back");
    if (slots > cap_) throw std::invalid_argument("step_begin: more slots than the group holds");
    check(hipSetDevice(device_), "hipSetDevice");
    uint32_t ns = cap_;
    if (slots) {
        ns = 8;
        while (ns < slots) ns *= 2;
        ns = std::min(ns, cap_);
    }
    step_n_ = ns;
    if (g == (uint32_t)kStepChunk && !no_graphs_) {
        auto it = step_graphs_.find(ns);
        if (it == step_graphs_.end()) {
            hipGraphExec_t ex = nullptr;
            hipGraph_t gr = nullptr;
            build_graph(ex, gr, true, kStepChunk, false);
            it = step_graphs_.emplace(ns, std::make_pair(ex, gr)).first;
        }
        check(hipGraphLaunch(it->second.first, stream_), "hipGraphLaunch");
        launched_ = true;
    } else {
        enqueue_err_clear(ns, 0u);
        enqueue_step_chunk((int)g);
        enqueue_err_reduce(0, ns);
        check(hipGetLastError(), "step launch");
    }
    pending_steps_ = g;
}

bool Group::step_ready()
{
    if (!pending_steps_) return true;
    const hipError_t e = hipStreamQuery(stream_);
    if (e == hipErrorNotReady) return false;
    check(e, "hipStreamQuery");
    return true;
}

double Group::step_end(double* slot_err)
{
    if (!pending_steps_) throw std::logic_error("step_end: no chunk in flight");
    check(hipSetDevice(device_), "hipSetDevice");
    const int g = (int)pending_steps_;
    pending_steps_ = 0;
    return read_err(g, nullptr, slot_err, 0, step_n_);
}

double Group::step(uint32_t steps, double* slot_err)
{
    double emax = 0.0;
    std::vector<double> se(cap_, 0.0), tmp(cap_, 0.0);
    uint32_t done = 0;
    while (done < steps) {
        const uint32_t g = std::min<uint32_t>(kStepChunk, steps - done);
        step_begin(g);
        emax = std::max(emax, step_end(tmp.data()));
        for (uint32_t s = 0; s < cap_; ++s) se[s] = std::max(se[s], tmp[s]);
        done += g;
    }
    if (slot_err) std::copy(se.begin(), se.end(), slot_err);
    return emax;
}

double Group::leave(uint32_t slot)
{
    check(hipSetDevice(device_), "hipSetDevice");
    {
        const int cnt = (int)((size_t)kChunk * eb_);
        hipLaunchKernelGGL(k_fill_u64, dim3((cnt + 255) / 256), dim3(256), 0, stream_, errp_ + (size_t)slot * cnt, cnt);
        enqueue_set_col(slot, 0u);
    }
    // the schedule multiplier of the last squaring comes from the slot's state
    Params p = params(slot, true);
    p.sched = 1;
    SlotState st{};
    check(hipMemcpyAsync(&st, state_ + slot, sizeof st, hipMemcpyDeviceToHost, stream_), "state D2H");
    check(hipStreamSynchronize(stream_), "state sync");
    // the kernels skip inactive slots; the leave sequence runs regardless
    SlotState on{ st.carried, 1, 0 };
    check(hipMemcpyAsync(state_ + slot, &on, sizeof on, hipMemcpyHostToDevice, stream_), "state H2D");
    enqueue_exit(p, 1);
    enqueue_err_reduce(slot, 1);
    SlotState fin{ st.carried + 1, 0, 0 };
    check(hipMemcpyAsync(state_ + slot, &fin, sizeof fin, hipMemcpyHostToDevice, stream_), "state H2D");
    unsigned long long e = 0;
    check(hipMemcpyAsync(&e, err_ + (size_t)slot * kChunk, sizeof e, hipMemcpyDeviceToHost, stream_), "err D2H");
    sync_status("leave sync");
    double d;
    std::memcpy(&d, &e, sizeof d);
    return d;
}

std::vector<SlotState> Group::states()
{
    std::vector<SlotState> v(cap_);
    check(hipMemcpyAsync(v.data(), state_, sizeof(SlotState) * cap_, hipMemcpyDeviceToHost, stream_), "state D2H");
    check(hipStreamSynchronize(stream_), "state sync");
    return v;
}

void Group::kernel_times(int reps, double* ms, uint32_t nslots)
{
    if (nslots == 0 || nslots > cap_) throw std::invalid_argument("kernel_times: slot count out of range");
    check(hipSetDevice(device_), "hipSetDevice");
    std::vector<int32_t> saved;
    int64_t saved_xhi = 0;
    download(0, saved, saved_xhi);
    const Params p = params(0, false);
    // Real squarings in their steady-state order, an event between every two
    // kernels: each kernel is timed on the data its predecessor just wrote,
    // as it runs inside a squaring chunk.  The first and last kernels of a
    // chunk (forward column pass from digits, inverse column + carry + fix-up
    // back to digits) are timed on the same pass around the loop.
    TurnToken token(device_, turn());
    Params pt = p;
    pt.runc = runt_;
    std::vector<int> seq;
    if (token.held)
        seq = { kCarryTurn, kRow };
    else if (plan_.nt)
        seq = { kInvCarry, kAdjFwd, kRow };
    else
        seq = { kInvCol, kCarry, kFixup, kFwdCol, kRow };
    const int m = (int)seq.size();
    std::vector<hipEvent_t> ev((size_t)(m + 1) * (size_t)reps + 8);
    // no system-scope fence per event: between kernels of one stream it only
    // adds a cache write-back that the squaring itself does not pay
    for (auto& e : ev) check(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "hipEventCreate");
    for (int kind = 0; kind < kNumKernels; ++kind) ms[kind] = 0.0;
    // enter: the first squaring's forward transform (timed once, then the
    // forward column pass as the unfused steady state times it)
    check(hipEventRecord(ev[0], stream_), "hipEventRecord");
    launch(kFwdCol, p, nslots);
    check(hipEventRecord(ev[1], stream_), "hipEventRecord");
    launch(kRow, p, nslots);
    size_t e = 2;
    check(hipEventRecord(ev[e], stream_), "hipEventRecord");
    const size_t loop0 = e;
    for (int r = 0; r < reps; ++r)
        for (int j = 0; j < m; ++j) {
            launch(seq[(size_t)j], seq[(size_t)j] == kCarryTurn ? pt : p, nslots);
            check(hipEventRecord(ev[++e], stream_), "hipEventRecord");
        }
    // leave: back to digits
    const size_t tail0 = e;
    launch(kInvCol, p, nslots);
    check(hipEventRecord(ev[++e], stream_), "hipEventRecord");
    launch(kCarry, p, nslots);
    check(hipEventRecord(ev[++e], stream_), "hipEventRecord");
    launch(kFixup, p, nslots);
    check(hipEventRecord(ev[++e], stream_), "hipEventRecord");
    check(hipEventSynchronize(ev[e]), "hipEventSynchronize");
    auto span = [&](size_t i, size_t j) {
        float t = 0;
        check(hipEventElapsedTime(&t, ev[i], ev[j]), "hipEventElapsedTime");
        return (double)t;
    };
    for (int r = 0; r < reps; ++r)
        for (int j = 0; j < m; ++j) {
            const size_t i = loop0 + (size_t)r * (size_t)m + (size_t)j;
            ms[seq[(size_t)j]] += span(i, i + 1) / reps;
        }
    // The same steady loop again with events at its two ends only: the
    // excess of the per-kernel sum over this per-squaring time is what the
    // events between kernels cost (about 1 us each); it is taken off each
    // steady kernel in equal parts.
    {
        hipEvent_t a, b;
        check(hipEventCreateWithFlags(&a, hipEventDisableSystemFence), "hipEventCreate");
        check(hipEventCreateWithFlags(&b, hipEventDisableSystemFence), "hipEventCreate");
        launch(kFwdCol, p, nslots);
        launch(kRow, p, nslots);
        check(hipEventRecord(a, stream_), "hipEventRecord");
        for (int r = 0; r < reps; ++r)
            for (int j = 0; j < m; ++j) launch(seq[(size_t)j], seq[(size_t)j] == kCarryTurn ? pt : p, nslots);
        check(hipEventRecord(b, stream_), "hipEventRecord");
        launch(kInvCol, p, nslots);
        launch(kCarry, p, nslots);
        launch(kFixup, p, nslots);
        check(hipEventSynchronize(b), "hipEventSynchronize");
        float t = 0;
        check(hipEventElapsedTime(&t, a, b), "hipEventElapsedTime");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        double sum = 0;
        for (int j = 0; j < m; ++j) sum += ms[seq[(size_t)j]];
        const double per_event = std::max(0.0, (sum - (double)t / reps) / m);
        for (int j = 0; j < m; ++j) ms[seq[(size_t)j]] = std::max(0.0, ms[seq[(size_t)j]] - per_event);
    }
    if (plan_.nt) {
        ms[kFwdCol] = span(0, 1);
        ms[kInvCol] = span(tail0, tail0 + 1);
        ms[kCarry] = span(tail0 + 1, tail0 + 2);
        ms[kFixup] = span(tail0 + 2, tail0 + 3);
    }
    for (auto& ee : ev) (void)hipEventDestroy(ee);
    if (status()) clear_status();
    upload(0, saved, saved_xhi);
}

// Enqueues the status word's read-back behind whatever the stream holds,
// waits for the stream and records the word (one wait for both).
void Group::sync_status(const char* what)
{
    check(hipMemcpyAsync(status_host_, status_, sizeof(int), hipMemcpyDeviceToHost, stream_), "status D2H");
    check(hipStreamSynchronize(stream_), what);
    seen_status_ = *status_host_;
    launched_ = false;
}

int Group::status()
{
    // the word read at the last wait holds unless kernels ran since
    if (launched_) sync_status("status sync");
    return seen_status_;
}

void Group::clear_status()
{
    check(hipMemsetAsync(status_, 0, sizeof(int), stream_), "status reset");
    check(hipStreamSynchronize(stream_), "status sync");
    seen_status_ = 0;
}

double copy_probe(int device, size_t bytes, int reps)
{
    check(hipSetDevice(device), "hipSetDevice");
    const size_t n = bytes / sizeof(double2);
    if (n == 0 || reps <= 0) throw std::runtime_error("copy_probe: empty copy");
    double2 *a = nullptr, *b = nullptr;
    check(hipMalloc(&a, n * sizeof(double2)), "hipMalloc");
    hipError_t e = hipMalloc(&b, n * sizeof(double2));
    if (e != hipSuccess) {
        (void)hipFree(a);
        check(e, "hipMalloc");
    }
    hipStream_t st = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float ms = 0;
    try {
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&t0), "hipEventCreate");
        check(hipEventCreate(&t1), "hipEventCreate");
        check(hipMemsetAsync(a, 0, n * sizeof(double2), st), "hipMemsetAsync");
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
        hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, st, a, b, n);
        check(hipEventRecord(t0, st), "hipEventRecord");
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, st, a, b, n);
        check(hipEventRecord(t1, st), "hipEventRecord");
        check(hipEventSynchronize(t1), "hipEventSynchronize");
        check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
    } catch (...) {
        if (st) (void)hipStreamSynchronize(st);
        (void)hipFree(a);
        (void)hipFree(b);
        throw;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    return (double)ms / reps;
}

}  // namespace pleng
