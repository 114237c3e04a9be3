// resident.h -- the steady state of plain squaring as one resident launch
// (k_loop), included by engine.hip after the kernels it shares code with.
//
// Every workgroup stays on its CU for a whole chunk of squarings.  Per
// squaring it runs the outer pass of its column tile (inverse column DFT,
// round, carries with the hand-off to the next tile, forward column DFT)
// and then the row pass of its row tile (forward row DFT, Z*(Z + 2i*X_hi),
// inverse row DFT), the work k_carry_turn and k_row do as two launches.
// The launch boundary between the passes is replaced by a handshake on the
// points themselves, so a squaring costs two exchanges of the residue
// between CUs and no kernel start or end.
//
// Point traffic.  A tile written by one workgroup is read by all the
// others (the four-step transposes), and workgroups sit on different XCDs,
// whose L2 caches are not coherent with each other.  Points are therefore
// stored with device-scope (sc1) vector stores, which write through the
// XCD's L2 to the device-coherent level, and loaded with device-scope
// loads, which do not return a stale L2 copy.  Each point is its own ready
// flag: the lowest mantissa bit of its real and of its imaginary part holds
// the parity of the pass that wrote it (0: outer pass, 1: row pass).  Every
// location is written once per pass, by the workgroup that reads it in that
// pass (each pass transforms its tile in place), so before a pass writes a
// location it holds the other pass's parity, and a reader simply reloads
// its points until all carry the parity it waits for.  Both halves are
// tagged, so a 16-byte point read half old and half new is seen as not
// ready.  Forcing the bit moves a value by at most one unit in the last
// place, an error far inside the round-off guard.  The row pass that enters
// the steady state before a k_loop launch stores its points with parity 1
// (Params::ztag).
//
// The carries use the tagged words of k_carry_turn.  X_hi, which the outer
// pass's top run replaces and every row tile needs, goes out as one more
// tagged word (xw).  All waits are bounded: a wait gives up after
// spin_limit polls, or as soon as it sees another workgroup's give-up in the
// status word (bit 2), and the whole launch then ends early; the chunk is
// recomputed from its saved start on the split pass, as for k_carry_turn.
// The launch needs every workgroup resident at once (the host checks the
// occupancy and launches at most one such chunk per device at a time).
#pragma once

namespace pleng {

typedef unsigned pl_u4v __attribute__((ext_vector_type(4)));
constexpr int kDevScope = 16;                 // buffer cache policy bits: sc1 (device scope)
constexpr int kBufWord3 = 0x00020000;         // raw buffer descriptor word 3 (gfx9: 32-bit data format)

// The residue's points as a raw buffer: device-scope loads and tagged
// device-scope stores of one 16-byte point.  Offsets past the slot's L
// points read as zero and drop stores (the descriptor's range).
struct PtBuf {
    __amdgpu_buffer_rsrc_t rs;
    __device__ __forceinline__ PtBuf(double2* z, unsigned L)
        : rs(__builtin_amdgcn_make_buffer_rsrc(z, 0, (int)(L * 16u), kBufWord3))
    {
    }
    __device__ __forceinline__ pl_u4v load(unsigned i) const
    {
        return __builtin_amdgcn_raw_buffer_load_b128(rs, i * 16u, 0, kDevScope);
    }
    __device__ __forceinline__ void store(unsigned i, double2 v, unsigned tag) const
    {
        pl_u4v b;
        b.x = ((unsigned)__double2loint(v.x) & ~1u) | tag;
        b.y = (unsigned)__double2hiint(v.x);
        b.z = ((unsigned)__double2loint(v.y) & ~1u) | tag;
        b.w = (unsigned)__double2hiint(v.y);
        __builtin_amdgcn_raw_buffer_store_b128(b, rs, i * 16u, 0, kDevScope);
    }
};

__device__ __forceinline__ bool pt_ready(pl_u4v v, unsigned tag)
{
    return (((v.x ^ tag) | (v.z ^ tag)) & 1u) == 0;
}
__device__ __forceinline__ double2 pt_value(pl_u4v v)
{
    return make_double2(__hiloint2double((int)v.y, (int)v.x), __hiloint2double((int)v.w, (int)v.z));
}

// A pointer the compiler cannot see through, so that table loads are not
// hoisted out of the squaring loop: kept live across both passes, the two
// passes' twiddles together would not fit the register file.
template <class T>
__device__ __forceinline__ T* opaque(T* ptr)
{
    asm volatile("" : "+s"(ptr));
    return ptr;
}

// The thread's index as a value the compiler cannot see through, read anew
// for every pass: the address arithmetic of a pass's stages derives from it
// and is then recomputed in each pass rather than hoisted out of the
// squaring loop and held in registers across both passes.
__device__ __forceinline__ unsigned pass_tid()
{
    unsigned t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Another workgroup gave up a wait of this launch.
__device__ __forceinline__ bool launch_aborted(const Params& p)
{
    return (__hip_atomic_load(p.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 2) != 0;
}

// Loads the points one thread reads in the first stage of a transform
// (ix(u, r): point index of leg r of butterfly u) until each carries `tag`.
// False when the wait gave up (status bit 2 is then set).
template <int PER, int P, class Live, class Ix>
__device__ __forceinline__ bool poll_points(const PtBuf& zb, Live live, Ix ix, unsigned tag, const Params& p,
                                            double2 (&g)[PER][P])
{
    static_assert(PER * P <= 32, "one pending bit per point");
    pl_u4v v[PER][P];
#pragma unroll
    for (int u = 0; u < PER; ++u)
#pragma unroll
        for (int r = 0; r < P; ++r)
            if (live(u)) v[u][r] = zb.load(ix(u, r));
    unsigned spins = 0;
    bool ok = true;
    for (;;) {
        unsigned pend = 0;
#pragma unroll
        for (int u = 0; u < PER; ++u)
#pragma unroll
            for (int r = 0; r < P; ++r)
                if (live(u) && !pt_ready(v[u][r], tag)) pend |= 1u << (u * P + r);
        if (!__any(pend != 0)) break;
        if (++spins >= p.spin_limit || ((spins & 63) == 0 && launch_aborted(p))) {
            ok = false;
            break;
        }
        for (unsigned z = 0; z < p.poll_sleep; ++z) __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int u = 0; u < PER; ++u)
#pragma unroll
            for (int r = 0; r < P; ++r)
                if (pend & (1u << (u * P + r))) v[u][r] = zb.load(ix(u, r));
    }
#pragma unroll
    for (int u = 0; u < PER; ++u)
#pragma unroll
        for (int r = 0; r < P; ++r) g[u][r] = pt_value(v[u][r]);
    if (!ok) atomicOr(p.status, 2);
    return ok;
}

// Waits for a tagged word to carry `tag`; false when the wait gave up.
__device__ __forceinline__ bool poll_word(const unsigned long long* w, unsigned tag, const Params& p,
                                          unsigned long long& out)
{
    unsigned spins = 0;
    out = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (tag_of(out) != tag) {
        if (++spins >= p.spin_limit || ((spins & 63) == 0 && launch_aborted(p))) {
            atomicOr(p.status, 2);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
        out = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
}

// Outer pass of one squaring over the column tile of workgroup w: columns
// [w*NTD, w*NTD + NTD), NTD/8 runs of 8 digits per row.  Thread q handles
// runs q + u*T (u < RPT) of the tile: row r = run % R, run column block
// w*(NTD/8) + run / R.  Otherwise as k_carry_turn, with its point loads and
// stores replaced by tagged device-scope ones and X_hi published as a
// tagged word.
template <int R, int NTD, int T, int MAXR>
__device__ __forceinline__ bool loop_outer(const Params& p, double2* const (&buf)[2], const PtBuf& zb, unsigned w,
                                           unsigned e, unsigned ecol, double* s_err, unsigned long long* xw)
{
    using Lay = LdsLayout<R, NTD>;
    constexpr int BPT = NTD / 8;           // run column blocks per tile
    constexpr int RUNS = R * BPT;
    static_assert(NTD % 8 == 0 && (RUNS % T == 0 || T % RUNS == 0), "runs must split evenly over the threads");
    constexpr int RPT = RUNS >= T ? RUNS / T : 1;  // threads q >= RUNS have no run
    constexpr int P0 = Radix<R, MAXR>::first();
    constexpr int PL = Radix<R, MAXR>::last();
    constexpr int IO = pp_out<R, MAXR, 0>();
    double2* const io = buf[IO];
    const unsigned nb = p.C / 8;
    const unsigned c0 = w * NTD;
    const unsigned tid = pass_tid();
    unsigned long long* const rw = reinterpret_cast<unsigned long long*>(p.runc);
    const size_t runs = (size_t)nb * R;
    StageTw<R, NTD, T, MAXR> t;
    load_stage_tw(t, p.tw, tid);
    double2 wri[RPT];
#pragma unroll
    for (int u = 0; u < RPT; ++u) wri[u] = p.wri[(tid + u * T) % R];
    {
        using S = StageShape<R, NTD, T, P0>;
        ColTw<R, NTD, T, P0> ct;
        const unsigned cc = c0 + tid % NTD;
        const double2 wc = p.wci[cc];
        ct.template load<true>(p, cc, make_double2(wc.x, -wc.y), tid);
        double2 g[S::PER][P0];
        const bool ok = poll_points<S::PER, P0>(
            zb, [&](int u) { return S::live(u, tid); },
            [&](int u, int r) { return (unsigned)(S::j(u, tid) + r * S::NBT) * p.C + cc; }, 1u, p, g);
        if (__syncthreads_or(!ok)) return false;
        dft_pp<R, NTD, T, true, 0>(
            buf, t, [&](int, int, int u, int r) { return cmulc(g[u][r], ct.w[u][r]); },
            [&](int tr, int idx, double2 v, int, int) { io[Lay::at(tr, idx)] = v; }, tid);
    }
    __syncthreads();

    const unsigned long long Lm = p.L - 1;
    const unsigned nr = (unsigned)(p.n & Lm);
    double emax = 0.0;
    const long long k = p.k;
    const double kd = (double)p.k, kmd = (double)(p.k * p.mult);
    const double tb0 = ldexp(1.0, (int)(p.n >> p.logL)), tb1 = 2.0 * tb0;
    const double itb0 = 1.0 / tb0, itb1 = 0.5 * itb0;
    int dg[RPT][8];
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned q = tid + u * T, r = q % R, sub = q / R;
        if (q >= (unsigned)RUNS) break;
        const unsigned b = w * BPT + sub;
        const unsigned F = (unsigned)((0ull - p.n * (unsigned long long)r * p.C) & Lm);
        // G_j = (-n*(8b + j)) mod L, stepped by -(n mod L) (DigitWalk)
        unsigned Gj = (unsigned)((0ull - p.n * (unsigned long long)(b * 8)) & Lm);
        double q1 = 0, q2 = 0, q3 = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            double2 wv = wri[u];
            unsigned f = F + Gj;
            Gj = Gj >= nr ? Gj - nr : Gj + p.L - nr;
            if (f >= p.L) {
                f -= p.L;
                wv = make_double2(2.0 * wv.x, 2.0 * wv.y);
            }
            const double2 v = cmul(io[Lay::at(sub * 8 + j, r)], wv);
            const double tr = rint(v.x), hr = rint(v.y);
            emax = fmax(emax, fmax(fabs(v.x - tr), fabs(v.y - hr)));
            const bool wd = f < nr;
            dg[u][j] = (int)carry3f(tr, hr, kd, kmd, wd ? tb1 : tb0, wd ? itb1 : itb0, q1, q2, q3);
        }
        unsigned long long* wp = rw + (size_t)b * R + r;
        __hip_atomic_store(wp, tag_word(q1, e), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(wp + runs, tag_word(q2, e), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(wp + 2 * runs, tag_word(q3, e), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // predecessors' words: first poll now, so the reduction and the
    // forward-side tables overlap its round trip
    const unsigned long long* pw[RPT];
    unsigned long long w0[RPT], w1[RPT], w2[RPT];
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned q = tid + u * T, r = q % R, b = w * BPT + q / R;
        if (q >= (unsigned)RUNS) break;
        pw[u] = rw + (b == 0 ? (r == 0 ? (size_t)(nb - 1) * R + (R - 1) : (size_t)(nb - 1) * R + (r - 1))
                             : (size_t)(b - 1) * R + r);
        w0[u] = __hip_atomic_load(pw[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w1[u] = __hip_atomic_load(pw[u] + runs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w2[u] = __hip_atomic_load(pw[u] + 2 * runs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    emax = wave_max(emax);
    if ((tid & 63) == 0) s_err[tid / 64] = emax;

    ColTw<R, NTD, T, PL> ct;
    {
        const unsigned cc = c0 + tid % NTD;
        ct.template load<true>(p, cc, p.wcf[cc], tid);
    }
    double2 wrf[RPT];
#pragma unroll
    for (int u = 0; u < RPT; ++u) wrf[u] = p.wrf[(tid + u * T) % R];

    bool ok = true;
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const unsigned q = tid + u * T, r = q % R, sub = q / R;
        if (q >= (unsigned)RUNS) break;
        const unsigned b = w * BPT + sub;
        const unsigned want = (e + p.tag_skew) & kTagMask;  // tag_skew: tests of the give-up path
        unsigned spins = 0;
        while (ok && !(tag_of(w0[u]) == want && tag_of(w1[u]) == want && tag_of(w2[u]) == want)) {
            if (++spins >= p.spin_limit || ((spins & 63) == 0 && launch_aborted(p))) {
                ok = false;
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
            // the old X_hi, last stored by this same thread (or before the launch)
            const long long xh = __hip_atomic_load(p.xhi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            c2in = cin[0];
            extra = ((__int128)cin[1] + (__int128)xh * xh) * p.mult;
            __hip_atomic_store(p.xhi, cin[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(xw, ((unsigned long long)cin[2] << kTagBits) | e, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            c2in = (__int128)cin[1] - (__int128)k * cin[0];
            extra = cin[2];
        }
        long long nu[8];
        ripple<8>(p, r * p.C + b * 8, c2in, extra, nu);
        const unsigned F = (unsigned)((0ull - p.n * (unsigned long long)r * p.C) & Lm);
        unsigned Gj = (unsigned)((0ull - p.n * (unsigned long long)(b * 8)) & Lm);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            double2 wv = wrf[u];
            if (F + Gj >= p.L) wv = make_double2(0.5 * wv.x, 0.5 * wv.y);
            Gj = Gj >= nr ? Gj - nr : Gj + p.L - nr;
            const double xd = (double)(dg[u][j] + (int)nu[j]);
            io[Lay::at(sub * 8 + j, r)] = make_double2(xd * wv.x, xd * wv.y);
        }
    }
    if (!ok) atomicOr(p.status, 2);
    if (__syncthreads_or(!ok)) return false;
    if (tid == 0) {
        double em = s_err[0];
#pragma unroll
        for (int i = 1; i < T / 64; ++i) em = fmax(em, s_err[i]);
        p.err[(size_t)w * Group::kChunk + ecol] = (unsigned long long)__double_as_longlong(em);
    }
    dft_pp<R, NTD, T, false, 1 - IO>(
        buf, t, [&](int tr, int idx, int, int) { return io[Lay::at(tr, idx)]; },
        [&](int tr, int idx, double2 v, int u, int r) {
            zb.store((unsigned)idx * p.C + c0 + tr, cmul(v, ct.w[u][r]), 0u);
        },
        tid);
    return true;
}

// Row pass of one squaring over rows [w*NTR, w*NTR + NTR), as k_row, with
// tagged device-scope point loads and stores and X_hi from its tagged word.
template <int C, int NTR, int T, int RADIX>
__device__ __forceinline__ bool loop_row(const Params& p, double2* const (&buf)[2], const PtBuf& zb, unsigned w,
                                         unsigned e, const unsigned long long* xw)
{
    constexpr int P0 = Radix<C, RADIX>::first();
    using S = StageShape<C, NTR, T, P0>;
    const unsigned tid = pass_tid();
    StageTw<C, NTR, T, RADIX> t;
    load_stage_tw(t, p.tw, tid);
    const unsigned base = w * NTR * C;
    const unsigned tr0 = tid % NTR;
    double2 g[S::PER][P0];
    bool ok = poll_points<S::PER, P0>(
        zb, [&](int u) { return S::live(u, tid); },
        [&](int u, int r) { return base + tr0 * C + (unsigned)(S::j(u, tid) + r * S::NBT); }, 0u, p, g);
    unsigned long long xv = 0;
    ok = ok && poll_word(xw, e, p, xv);
    if (__syncthreads_or(!ok)) return false;
    const double x2 = 2.0 * (double)untag(xv);
    dft_square_pass_pp<C, NTR, T, RADIX>(
        buf, t, [&](int, int, int u, int r) { return g[u][r]; },
        [&](double2 X) { return cmul(X, make_double2(X.x, X.y + x2)); },
        [&](int tr, int idx, double2 v, int, int) { zb.store(base + (unsigned)tr * C + (unsigned)idx, v, 1u); },
        tid);
    return true;
}

// nsq squarings (outer pass then row pass) of slot 0, in one launch of
// C/NTD = R/NTR workgroups of T threads, all resident at once.  Entered
// after a row pass that stored parity 1; left after a row pass (parity 1),
// so the exit kernels follow as after k_row.  Epoch and round-off column
// advance as nsq k_carry_turn + k_row pairs would advance them.
template <int R, int C, int NTD, int NTR, int T, int MAXR, int RRADIX>
__global__ __launch_bounds__(T) void k_loop(Params pb, unsigned nsq, unsigned long long* xw)
{
    static_assert(C / NTD == R / NTR, "one column tile and one row tile per workgroup");
    constexpr int SC = LdsLayout<R, NTD>::kSlots, SR = LdsLayout<C, NTR>::kSlots;
    constexpr int SL = SC > SR ? SC : SR;
    __shared__ double2 lds[2][SL];
    __shared__ double s_err[T / 64];
    double2* const buf[2] = { lds[0], lds[1] };
    const Params p = bind<false>(pb, 0);
    const PtBuf zb(p.z, p.L);
    const unsigned w = blockIdx.x;
    const unsigned e0 = p.epoch[0], col0 = p.epoch[2];
    unsigned s = 0;
    for (; s < nsq; ++s) {
        const unsigned e = (e0 + s) & kTagMask;
        Params q = p;
        q.tw = opaque(q.tw);
        q.twl_lo = opaque(q.twl_lo);
        q.twl_hi = opaque(q.twl_hi);
        q.tcq = opaque(q.tcq);
        q.wri = opaque(q.wri);
        q.wrf = opaque(q.wrf);
        q.wci = opaque(q.wci);
        q.wcf = opaque(q.wcf);
        if (!loop_outer<R, NTD, T, MAXR>(q, buf, zb, w, e, (col0 + s) & (Group::kChunk - 1), s_err, xw)) break;
        q.tw = opaque(q.tw);
        if (!loop_row<C, NTR, T, RRADIX>(q, buf, zb, w, e, xw)) break;
    }
    // every workgroup read the epochs before block 0 can get here: block 0's
    // last row pass waited for every tile's last outer pass
    if (w == 0 && threadIdx.x == 0 && s == nsq) {
        const unsigned last = (e0 + nsq - 1) & kTagMask;
        p.epoch[1] = last;
        p.epoch[0] = last + 1u;
        p.epoch[2] = col0 + nsq;
    }
}

}  // namespace pleng
