// fft_lds.h -- device building blocks: radix-16/8/4/2 Stockham DFT stages
// over NT interleaved transforms of length N staged in LDS.
//
// Thread t of a TPB-thread workgroup owns transform tr = t % NT and the
// butterflies j = t/NT + u*(TPB/NT) of every stage, so a wavefront touches
// NT adjacent transforms at each index: global accesses of a column tile are
// NT*16 contiguous bytes, and LDS accesses hit neighbouring 16-byte slots.
//
// LDS layout: element idx of transform tr is slot s = NT*idx + tr, XOR
// swizzled within aligned groups of 2^PL_LDS_SWZ slots (s ^ ((s >> SWZ) &
// (2^SWZ - 1))).  The swizzle keeps the three access patterns free of bank
// conflicts: lanes stepping tr (every column tile), lanes stepping idx by the
// stage radix (first-stage stores), and lanes stepping the transform index
// while each reads NT consecutive slots (the row-per-thread carry phase).
//
// A transform is run as: first stage reading from a caller functor (global
// memory or LDS), middle stages LDS -> LDS, last stage writing to a caller
// functor (global memory or LDS).  Functors get (tr, idx, u, r): u and r
// are the (unrolled, compile-time) butterfly and leg numbers, so callers can
// keep per-point values in registers indexed by them.  Stage with span NS
// and radix P: butterfly j reads j + r*N/P, twiddles by W_(NS*P)^(r*(j mod
// NS)), applies the P-point DFT, writes (j/NS)*NS*P + (j mod NS) + r*NS
// (natural order out).  The first stage therefore reads idx = j + r*N/P0 and
// the last writes idx = j + r*N/P_last.
//
// Stage twiddles are formed once per kernel in registers (StageTw) before
// the data arrive: per stage and butterfly W^e comes from the 4096-point
// table, W^2e, W^4e, W^8e are its squares and the other powers their
// products, formed at use; the inverse transform uses the same registers
// conjugated.
#pragma once
#include <hip/hip_runtime.h>

namespace pleng {

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b)  // a * conj(b)
{
    return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
// Streaming store of one point (global_store_dwordx4 ... nt): a pass's
// output is next read by the other pass's kernel, never by this one, and
// kept out of L2 it leaves the end-of-kernel write-back less to do (0.9 us
// per squaring at L = 2^20 on MI355X, neutral at shorter lengths).
typedef double pl_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_stream(double2* dst, double2 v)
{
    const pl_d2v w = { v.x, v.y };
    __builtin_nontemporal_store(w, reinterpret_cast<pl_d2v*>(dst));
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

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
__device__ __forceinline__ void dft4(double2& a0, double2& a1, double2& a2, double2& a3)
{
    double2 s0 = cadd(a0, a2), s1 = csub(a0, a2);
    double2 s2 = cadd(a1, a3), s3 = csub(a1, a3);
    double2 d = mul_mi<INV>(s3);
    a0 = cadd(s0, s2);
    a1 = cadd(s1, d);
    a2 = csub(s0, s2);
    a3 = csub(s1, d);
}

template <bool INV>
__device__ __forceinline__ void dft4(double2* v)
{
    dft4<INV>(v[0], v[1], v[2], v[3]);
}

// multiply by W8^q = exp(-+ i*pi*q/4), q = 1, 2, 3
template <bool INV, int Q>
__device__ __forceinline__ double2 w8(double2 a)
{
    constexpr double h = 0.70710678118654752440;
    constexpr double sg = INV ? -1.0 : 1.0;
    if constexpr (Q == 1) return make_double2(h * (a.x + sg * a.y), h * (a.y - sg * a.x));
    else if constexpr (Q == 2) return mul_mi<INV>(a);
    else return make_double2(h * (-a.x + sg * a.y), h * (-a.y - sg * a.x));
}

template <bool INV>
__device__ __forceinline__ void dft8(double2* v)
{
    double2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    double2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    o1 = w8<INV, 1>(o1);
    o2 = w8<INV, 2>(o2);
    o3 = w8<INV, 3>(o3);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

// 16-point DFT as 4 x 4: inner DFT-4 over n1 (stride 4), twiddle
// W16^(n2*k1), outer DFT-4 over n2; X[k1 + 4*k2] lands in v[k1 + 4*k2].
template <bool INV>
__device__ __forceinline__ void dft16(double2* v)
{
    constexpr double c1 = 0.92387953251128675613;  // cos(pi/8)
    constexpr double s1 = 0.38268343236508977173;  // sin(pi/8)
    constexpr double sg = INV ? -1.0 : 1.0;
    double2 y[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        double2 a0 = v[n2], a1 = v[4 + n2], a2 = v[8 + n2], a3 = v[12 + n2];
        dft4<INV>(a0, a1, a2, a3);
        y[4 * n2 + 0] = a0;
        y[4 * n2 + 1] = a1;
        y[4 * n2 + 2] = a2;
        y[4 * n2 + 3] = a3;
    }
    // twiddles W16^(n2*k1) for n2, k1 in 1..3: exponents 1,2,3 / 2,4,6 / 3,6,9
    const double2 w1 = make_double2(c1, -sg * s1);
    const double2 w3 = make_double2(s1, -sg * c1);
    const double2 w9 = make_double2(-c1, sg * s1);  // W16^9 = -W16^1
    y[5] = cmul(y[5], w1);
    y[6] = w8<INV, 1>(y[6]);
    y[7] = cmul(y[7], w3);
    y[9] = w8<INV, 1>(y[9]);
    y[10] = w8<INV, 2>(y[10]);
    y[11] = w8<INV, 3>(y[11]);
    y[13] = cmul(y[13], w3);
    y[14] = w8<INV, 3>(y[14]);
    y[15] = cmul(y[15], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        double2 a0 = y[k1], a1 = y[4 + k1], a2 = y[8 + k1], a3 = y[12 + k1];
        dft4<INV>(a0, a1, a2, a3);
        v[k1] = a0;
        v[k1 + 4] = a1;
        v[k1 + 8] = a2;
        v[k1 + 12] = a3;
    }
}

template <int P, bool INV>
__device__ __forceinline__ void dftp(double2* v)
{
    if constexpr (P == 16) dft16<INV>(v);
    else if constexpr (P == 8) dft8<INV>(v);
    else if constexpr (P == 4) dft4<INV>(v);
    else dft2<INV>(v);
}

#ifndef PL_MAX_RADIX
#define PL_MAX_RADIX 8
#endif
#ifndef PL_LDS_SWZ
#define PL_LDS_SWZ 3
#endif
#ifndef PL_ROW_RADIX
#define PL_ROW_RADIX PL_MAX_RADIX
#endif

template <int N, int MAXR = PL_MAX_RADIX>
struct Radix {
    // radix of the stage whose span is NS (largest power of two up to MAXR
    // that still fits)
    static constexpr int of(int NS)
    {
        return (N / NS) >= 16 && MAXR >= 16 ? 16
             : (N / NS) >= 8 && MAXR >= 8   ? 8
             : (N / NS) >= 4                ? 4
                                            : 2;
    }
    static constexpr int first() { return of(1); }
    static constexpr int last()
    {
        int ns = 1, p = of(1);
        while (ns * p < N) {
            ns *= p;
            p = of(ns);
        }
        return p;
    }
};

template <int N, int NT>
struct LdsLayout {
    static constexpr int kSlots = NT * N;
    static_assert(kSlots % (1 << PL_LDS_SWZ) == 0, "swizzle group must divide the buffer");
    __device__ static __forceinline__ int at(int tr, int idx)
    {
        const int s = NT * idx + tr;
        return s ^ ((s >> PL_LDS_SWZ) & ((1 << PL_LDS_SWZ) - 1));
    }
};

// Threads, butterflies per thread and the butterfly index of a stage.
template <int N, int NT, int TPB, int P>
struct StageShape {
    static constexpr int G = TPB / NT;    // butterfly groups per transform
    static constexpr int NBT = N / P;     // butterflies per transform
    static constexpr int PER = NBT >= G ? NBT / G : 1;
    static_assert(TPB % NT == 0, "TPB must be a multiple of NT");
    static_assert(NBT < G || NBT % G == 0, "butterflies must split evenly");
    // tid: the thread's index in the workgroup (a caller that loops over
    // whole transforms passes an opaque copy, so that the address arithmetic
    // derived from it is not hoisted out of the loop and kept live across it)
    __device__ static __forceinline__ int j(int u, unsigned tid = threadIdx.x) { return (int)(tid / NT) + u * G; }
    __device__ static __forceinline__ bool live(int u, unsigned tid = threadIdx.x) { return NBT >= G || j(u, tid) < NBT; }
};

// Register-resident stage twiddles of one thread for every stage of an
// N-point transform: for the stage with span NS and butterfly u, w[stage][u]
// holds W^e, W^2e, W^4e, W^8e with W^e = W_(NS*P)^(j mod NS); the other
// powers are formed where they are used (twiddles()).
template <int N, int NT, int TPB, int MAXR = PL_MAX_RADIX>
struct StageTw {
    static constexpr int G = TPB / NT;
    static constexpr int PMAX = (N / 2) >= G ? (N / 2) / G : 1;
    double2 w[8][PMAX][4];
};

// All P-1 twiddles of a butterfly from its base powers b = {W^e, W^2e, W^4e, W^8e}.
template <int P>
__device__ __forceinline__ void twiddles(const double2 (&b)[4], double2* w)
{
    w[1] = b[0];
    if constexpr (P >= 4) {
        w[2] = b[1];
        w[3] = cmul(b[0], b[1]);
    }
    if constexpr (P >= 8) {
        w[4] = b[2];
        w[5] = cmul(b[0], b[2]);
        w[6] = cmul(b[1], b[2]);
        w[7] = cmul(w[3], b[2]);
    }
    if constexpr (P >= 16) {
        w[8] = b[3];
#pragma unroll
        for (int r = 9; r < 16; ++r) w[r] = cmul(w[r - 8], b[3]);
    }
}

template <int N, int NT, int TPB, int MAXR, int NS, int SI>
__device__ __forceinline__ void tw_fill(StageTw<N, NT, TPB, MAXR>& t, const double2* __restrict__ tw, unsigned tid)
{
    constexpr int P = Radix<N, MAXR>::of(NS);
    if constexpr (NS > 1) {
        using S = StageShape<N, NT, TPB, P>;
        constexpr int SCALE = 4096 / N;       // tw is the 4096-point table
        constexpr int STEP = N / (NS * P);
#pragma unroll
        for (int u = 0; u < S::PER; ++u) {
            if (!S::live(u, tid)) continue;
            // e < (N/P) * SCALE = 4096/P: the index stays in the table.
            // W^e from the table, W^2e, W^4e, W^8e by squaring: one load per
            // stage and butterfly instead of up to four (each load a thread
            // issues before its data costs ~0.2 us at L = 2^20; the squarings
            // add a few ulp to a twiddle, far inside the round-off guard)
            const int e = (S::j(u, tid) % NS) * STEP * SCALE;
            t.w[SI][u][0] = tw[e];
            if constexpr (P >= 4) t.w[SI][u][1] = cmul(t.w[SI][u][0], t.w[SI][u][0]);
            if constexpr (P >= 8) t.w[SI][u][2] = cmul(t.w[SI][u][1], t.w[SI][u][1]);
            if constexpr (P >= 16) t.w[SI][u][3] = cmul(t.w[SI][u][2], t.w[SI][u][2]);
        }
    }
    if constexpr (NS * P < N) tw_fill<N, NT, TPB, MAXR, NS * P, SI + 1>(t, tw, tid);
}

template <int N, int NT, int TPB, int MAXR>
__device__ __forceinline__ void load_stage_tw(StageTw<N, NT, TPB, MAXR>& t, const double2* __restrict__ tw,
                                              unsigned tid = threadIdx.x)
{
    tw_fill<N, NT, TPB, MAXR, 1, 0>(t, tw, tid);
}

// One stage.  ld(tr, idx, u, r) -> double2, st(tr, idx, v, u, r).  When
// both go to the same LDS buffer the caller passes SYNC_BETWEEN = true so
// every read of the stage completes before any write.
template <int N, int NT, int TPB, int MAXR, int P, int NS, int SI, bool INV, bool SYNC_BETWEEN, class Ld, class St>
__device__ __forceinline__ void stage(const StageTw<N, NT, TPB, MAXR>& t, Ld ld, St st, unsigned tid)
{
    using S = StageShape<N, NT, TPB, P>;
    const int tr = tid % NT;
    double2 v[S::PER][P];
#pragma unroll
    for (int u = 0; u < S::PER; ++u) {
        if (S::live(u, tid)) {
            const int j = S::j(u, tid);
#pragma unroll
            for (int r = 0; r < P; ++r) v[u][r] = ld(tr, j + r * S::NBT, u, r);
            if constexpr (NS > 1) {
                double2 w[P];
                twiddles<P>(t.w[SI][u], w);
#pragma unroll
                for (int r = 1; r < P; ++r) v[u][r] = INV ? cmulc(v[u][r], w[r]) : cmul(v[u][r], w[r]);
            }
            dftp<P, INV>(v[u]);
        }
    }
    if constexpr (SYNC_BETWEEN) __syncthreads();
#pragma unroll
    for (int u = 0; u < S::PER; ++u) {
        if (S::live(u, tid)) {
            const int j = S::j(u, tid);
            const int o = (j / NS) * NS * P + (j % NS);
#pragma unroll
            for (int r = 0; r < P; ++r) st(tr, o + r * NS, v[u][r], u, r);
        }
    }
}

// Full N-point DFT of NT transforms: first stage ld_in -> LDS, middle
// stages in LDS, last stage LDS -> st_out.  lds must hold
// LdsLayout<N,NT>::kSlots double2.  IN_LDS: ld_in reads this same LDS
// buffer; OUT_LDS: st_out writes it (each then needs a barrier between the
// stage's reads and writes).  The caller provides the barrier before the
// call when LDS was in use; the call ends without a trailing barrier.
template <int N, int NT, int TPB, int MAXR, bool INV, bool IN_LDS, bool OUT_LDS, int NS, int SI, class LdIn, class StOut>
__device__ __forceinline__ void dft_rest(double2* lds, const StageTw<N, NT, TPB, MAXR>& t, LdIn ld_in, StOut st_out, unsigned tid)
{
    using Lay = LdsLayout<N, NT>;
    constexpr int P = Radix<N, MAXR>::of(NS);
    auto lds_ld = [&](int tr, int idx, int, int) { return lds[Lay::at(tr, idx)]; };
    auto lds_st = [&](int tr, int idx, double2 v, int, int) { lds[Lay::at(tr, idx)] = v; };
    if constexpr (NS * P >= N) {
        if constexpr (NS == 1) stage<N, NT, TPB, MAXR, P, NS, SI, INV, IN_LDS && OUT_LDS>(t, ld_in, st_out, tid);
        else stage<N, NT, TPB, MAXR, P, NS, SI, INV, OUT_LDS>(t, lds_ld, st_out, tid);
    } else {
        if constexpr (NS == 1) stage<N, NT, TPB, MAXR, P, NS, SI, INV, IN_LDS>(t, ld_in, lds_st, tid);
        else stage<N, NT, TPB, MAXR, P, NS, SI, INV, true>(t, lds_ld, lds_st, tid);
        __syncthreads();
        dft_rest<N, NT, TPB, MAXR, INV, IN_LDS, OUT_LDS, NS * P, SI + 1>(lds, t, ld_in, st_out, tid);
    }
}

template <int N, int NT, int TPB, bool INV, bool IN_LDS = false, bool OUT_LDS = false, int MAXR, class LdIn,
          class StOut>
__device__ __forceinline__ void dft(double2* lds, const StageTw<N, NT, TPB, MAXR>& t, LdIn ld_in, StOut st_out,
                                    unsigned tid = threadIdx.x)
{
    dft_rest<N, NT, TPB, MAXR, INV, IN_LDS, OUT_LDS, 1, 0>(lds, t, ld_in, st_out, tid);
}

// Span and stage number of the last stage of an N-point transform.
template <int N, int MAXR>
struct LastStage {
    static constexpr int ns()
    {
        int ns = 1, p = Radix<N, MAXR>::of(1);
        while (ns * p < N) {
            ns *= p;
            p = Radix<N, MAXR>::of(ns);
        }
        return ns;
    }
    static constexpr int si()
    {
        int ns = 1, s = 0, p = Radix<N, MAXR>::of(1);
        while (ns * p < N) {
            ns *= p;
            p = Radix<N, MAXR>::of(ns);
            ++s;
        }
        return s;
    }
};

// ------------------------------------------------------------ ping-pong
//
// The same transforms over two LDS buffers: a stage reads one buffer and
// writes the other, so it needs no barrier between its reads and its writes,
// only the one after its writes (the stage before last read the buffer this
// stage writes, and that barrier orders them).  Half the barriers of the
// one-buffer form, for twice the LDS.  A stage with span NS writes buf[DST]
// (or st_out when it is the last) and reads buf[1 - DST] (ld_in when NS = 1).

template <int N, int MAXR = PL_MAX_RADIX>
struct Stages {
    static constexpr int count()
    {
        int ns = 1, s = 1, p = Radix<N, MAXR>::of(1);
        while (ns * p < N) {
            ns *= p;
            p = Radix<N, MAXR>::of(ns);
            ++s;
        }
        return s;
    }
};

template <int N, int NT, int TPB, int MAXR, bool INV, int NS, int SI, int DST, class LdIn, class StOut>
__device__ __forceinline__ void dft_pp_rest(double2* const (&buf)[2], const StageTw<N, NT, TPB, MAXR>& t, LdIn ld_in,
                                            StOut st_out, unsigned tid)
{
    using Lay = LdsLayout<N, NT>;
    constexpr int P = Radix<N, MAXR>::of(NS);
    auto rd = [&](int tr, int idx, int, int) { return buf[1 - DST][Lay::at(tr, idx)]; };
    auto wr = [&](int tr, int idx, double2 v, int, int) { buf[DST][Lay::at(tr, idx)] = v; };
    if constexpr (NS * P >= N) {
        if constexpr (NS == 1) stage<N, NT, TPB, MAXR, P, NS, SI, INV, false>(t, ld_in, st_out, tid);
        else stage<N, NT, TPB, MAXR, P, NS, SI, INV, false>(t, rd, st_out, tid);
    } else {
        if constexpr (NS == 1) stage<N, NT, TPB, MAXR, P, NS, SI, INV, false>(t, ld_in, wr, tid);
        else stage<N, NT, TPB, MAXR, P, NS, SI, INV, false>(t, rd, wr, tid);
        __syncthreads();
        dft_pp_rest<N, NT, TPB, MAXR, INV, NS * P, SI + 1, 1 - DST>(buf, t, ld_in, st_out, tid);
    }
}

// Full N-point DFT over two buffers; the first stage writes buf[DST0].  When
// st_out writes LDS it must write buf[pp_out<N, MAXR, DST0>()], the buffer
// the last stage does not read.  Ends without a trailing barrier.
template <int N, int MAXR, int DST0>
__device__ constexpr int pp_out()
{
    return (DST0 + Stages<N, MAXR>::count() - 1) & 1;
}

template <int N, int NT, int TPB, bool INV, int DST0, int MAXR, class LdIn, class StOut>
__device__ __forceinline__ void dft_pp(double2* const (&buf)[2], const StageTw<N, NT, TPB, MAXR>& t, LdIn ld_in,
                                       StOut st_out, unsigned tid = threadIdx.x)
{
    dft_pp_rest<N, NT, TPB, MAXR, INV, 1, 0, DST0>(buf, t, ld_in, st_out, tid);
}

// Every stage but the last, ping-pong; returns after a barrier with the data
// in buf[pp_head_out<N, MAXR, DST0>()].
template <int N, int MAXR, int DST0>
__device__ constexpr int pp_head_out()
{
    return (DST0 + Stages<N, MAXR>::count() - 2) & 1;
}

template <int N, int NT, int TPB, int MAXR, bool INV, int NS, int SI, int DST, class LdIn>
__device__ __forceinline__ void dft_pp_head(double2* const (&buf)[2], const StageTw<N, NT, TPB, MAXR>& t, LdIn ld_in, unsigned tid)
{
    using Lay = LdsLayout<N, NT>;
    constexpr int P = Radix<N, MAXR>::of(NS);
    if constexpr (NS * P < N) {
        auto rd = [&](int tr, int idx, int, int) { return buf[1 - DST][Lay::at(tr, idx)]; };
        auto wr = [&](int tr, int idx, double2 v, int, int) { buf[DST][Lay::at(tr, idx)] = v; };
        if constexpr (NS == 1) stage<N, NT, TPB, MAXR, P, NS, SI, INV, false>(t, ld_in, wr, tid);
        else stage<N, NT, TPB, MAXR, P, NS, SI, INV, false>(t, rd, wr, tid);
        __syncthreads();
        dft_pp_head<N, NT, TPB, MAXR, INV, NS * P, SI + 1, 1 - DST>(buf, t, ld_in, tid);
    }
}

// The squaring pass over two buffers: forward DFT from ld_in, pointwise map,
// inverse DFT to st_out.  Where the first and last radix agree, the turn runs
// in registers: forward last-stage butterfly j produces the values at
// j + r*N/P, exactly the inputs of inverse first-stage butterfly j, so the
// forward last stage, the pointwise map and the inverse first stage need one
// LDS round trip instead of two.
template <int N, int NT, int TPB, int MAXR, class LdIn, class Pw, class StOut>
__device__ __forceinline__ void dft_square_pass_pp(double2* const (&buf)[2], const StageTw<N, NT, TPB, MAXR>& t,
                                                   LdIn ld_in, Pw pw, StOut st_out, unsigned tid = threadIdx.x)
{
    using Lay = LdsLayout<N, NT>;
    constexpr int P0 = Radix<N, MAXR>::first();
    if constexpr (P0 == Radix<N, MAXR>::last() && LastStage<N, MAXR>::ns() > 1) {
        constexpr int H = pp_head_out<N, MAXR, 0>();  // the forward head's output
        dft_pp_head<N, NT, TPB, MAXR, false, 1, 0, 0>(buf, t, ld_in, tid);
        // the turn: forward last stage from buf[H], inverse first stage into buf[1 - H]
        constexpr int SI = LastStage<N, MAXR>::si();
        using S = StageShape<N, NT, TPB, P0>;
        const int tr = tid % NT;
#pragma unroll
        for (int u = 0; u < S::PER; ++u) {
            if (S::live(u, tid)) {
                const int j = S::j(u, tid);
                double2 v[P0];
#pragma unroll
                for (int r = 0; r < P0; ++r) v[r] = buf[H][Lay::at(tr, j + r * S::NBT)];
                double2 w[P0];
                twiddles<P0>(t.w[SI][u], w);
#pragma unroll
                for (int r = 1; r < P0; ++r) v[r] = cmul(v[r], w[r]);
                dftp<P0, false>(v);
#pragma unroll
                for (int r = 0; r < P0; ++r) v[r] = pw(v[r]);
                dftp<P0, true>(v);
#pragma unroll
                for (int r = 0; r < P0; ++r) buf[1 - H][Lay::at(tr, j * P0 + r)] = v[r];
            }
        }
        __syncthreads();
        dft_pp_rest<N, NT, TPB, MAXR, true, P0, 1, H>(buf, t, ld_in, st_out, tid);
    } else {
        constexpr int F = pp_out<N, MAXR, 0>();  // where the forward pass leaves the spectrum
        dft_pp<N, NT, TPB, false, 0>(
            buf, t, ld_in, [&](int tr, int idx, double2 X, int, int) { buf[F][Lay::at(tr, idx)] = pw(X); }, tid);
        __syncthreads();
        dft_pp<N, NT, TPB, true, 1 - F>(
            buf, t, [&](int tr, int idx, int, int) { return buf[F][Lay::at(tr, idx)]; }, st_out, tid);
    }
}

}  // namespace pleng
