// capi.cpp -- the C ABI of include/prothcore.h over pleng::Engine and the
// host bignum helpers.  No exception crosses this boundary.
#include "prothcore.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fcntl.h>
#include <unistd.h>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "bignum.h"
#include "engine.h"
#include "prewrap.h"
#include "tune.h"

namespace {
thread_local std::string g_err = "no error";
}  // namespace

namespace plerr {
int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}
}  // namespace plerr

namespace {

using plerr::fail;

template <class F>
int guarded(F&& f)
{
    try {
        return f();
    } catch (const std::invalid_argument& e) {
        return fail(PL_ERR_ARG, e.what());
    } catch (const std::exception& e) {
        return fail(PL_ERR_HIP, e.what());
    }
}

using u128 = unsigned __int128;

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)(((u128)a * b) % m); }

uint64_t powmod(uint64_t a, uint64_t e, uint64_t m)
{
    uint64_t r = 1 % m;
    a %= m;
    while (e) {
        if (e & 1) r = mulmod(r, a, m);
        a = mulmod(a, a, m);
        e >>= 1;
    }
    return r;
}

int jacobi_u64(uint64_t a, uint64_t m)  // m odd
{
    int t = 1;
    a %= m;
    while (a) {
        while ((a & 1) == 0) {
            a >>= 1;
            uint64_t r = m & 7;
            if (r == 3 || r == 5) t = -t;
        }
        std::swap(a, m);
        if ((a & 3) == 3 && (m & 3) == 3) t = -t;
        a %= m;
    }
    return m == 1 ? t : 0;
}

// Jacobi(b, k*2^n+1) for small b by reciprocity (N mod 8 for the factors of
// two, N mod b' for the odd part; (N-1)/2 = k*2^(n-1) is even for n >= 2).
int jacobi_small(uint64_t b, uint64_t k, uint32_t n)
{
    const uint64_t n8 = ((k % 8) * powmod(2, n, 8) + 1) % 8;
    int t = 1;
    while (b && (b & 1) == 0) {
        b >>= 1;
        if (n8 == 3 || n8 == 5) t = -t;
    }
    if (b == 0) return 0;
    if (b == 1) return t;
    const uint64_t nmod = ((k % b) * powmod(2, n, b) + 1) % b;
    if (n == 1 && (k & 1) && ((b - 1) / 2) % 2 == 1) t = -t;
    if (nmod == 0) return 0;
    return t * jacobi_u64(nmod, b);
}


}  // namespace

struct pl_ctx {
    pleng::Plan plan;
    // slot 0: the residue; slots 1 .. extra: the caller's (Gerbicz check);
    // then the normal-form scratch (a temporary and two constants)
    std::unique_ptr<pleng::Group> eng;
    std::unique_ptr<plbn::Modulus> M;
    int device = 0;
    plprep::Scratch scr;
    uint64_t consts_a = ~0ull;  // the a whose constants the scratch holds (~0: none loaded)
    // While the exact value is small, slot 0 holds x itself (normal form).
    bool normal = false;
    plprep::SmallValue small;

    plbn::Limbs raw_of(uint32_t slot) { return plprep::raw(*eng, *M, slot); }
    // Montgomery form of slot 0 (only outside the normal form)
    plbn::Limbs mont_value() { return raw_of(0); }
    void set_mont_value(const plbn::Limbs& y) { plprep::set_raw(*eng, *M, 0, y); }
    plbn::Limbs mont_of(uint32_t slot) { return raw_of(slot); }
    plbn::Limbs value()
    {
        plbn::Limbs r = raw_of(0);
        return normal ? r : plbn::from_mont(*M, r);
    }
    void set_value(const plbn::Limbs& x0)
    {
        const plbn::Limbs x = plbn::mod_N(*M, x0);
        normal = small.set(*M, x);
        if (normal) plprep::set_raw(*eng, *M, 0, x);
        else set_mont_value(plbn::to_mont(*M, x));
    }
    void ensure_consts(uint64_t a)
    {
        if (consts_a == ~0ull || (a && consts_a != a)) {
            plprep::load_constants(*eng, *M, a, scr);
            consts_a = a;
        }
    }
    // One squaring in normal form: x <- x^2 * (times_a ? a : 1).
    double normal_step(bool times_a, uint64_t a)
    {
        ensure_consts(times_a ? a : 0);
        const double e = plprep::step(*eng, 0, scr, times_a);
        small.grow(true, times_a ? a : 0);
        return e;
    }
    // Normal form -> Montgomery form (once the exact value has wrapped).
    double leave_normal()
    {
        ensure_consts(0);
        const double e = plprep::to_mont(*eng, 0, scr);
        normal = false;
        return e;
    }
    bool wrapped() const { return small.wrapped(*M); }
    // `iters` squarings of slot 0 (in normal form first while the value is small).
    double square(uint64_t iters, double* trace)
    {
        double e = 0.0;
        uint64_t done = 0;
        while (normal && done < iters) {
            if (wrapped()) break;
            const double s = normal_step(false, 0);
            if (trace) trace[done] = s;
            e = std::max(e, s);
            ++done;
        }
        if (normal && wrapped()) e = std::max(e, leave_normal());
        if (done < iters) e = std::max(e, eng->square_slots(0, 1, iters - done, trace ? trace + done : nullptr));
        return e;
    }
    // The normal-form squarings until the value has wrapped (at most `limit`),
    // then the move to the Montgomery form; returns the squarings applied.
    uint64_t finish_normal(uint64_t limit, double& emax)
    {
        uint64_t done = 0;
        while (normal && !wrapped() && done < limit) {
            emax = std::max(emax, normal_step(false, 0));
            ++done;
        }
        if (normal && wrapped()) emax = std::max(emax, leave_normal());
        return done;
    }
};

namespace {

// The built-in plan of a single context: length rule (or fft_len), then the
// carry layout for the test's own digits and k.
bool rule_plan(uint32_t n, uint64_t k, uint32_t fft_len, pleng::Plan& p, std::string& why)
{
    if (!pleng::choose_plan(n, k, fft_len, p, why)) return false;
    const char* fz = std::getenv("PL_FUSED");
    const bool allow_split = !(fz && fz[0] == '0');
    return pleng::choose_layout(p, (double)n / p.L, k, 1, allow_split, why);
}

int check_device(int device)
{
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return fail(PL_ERR_HIP, "no HIP device visible");
    if (device < 0 || device >= nd) return fail(PL_ERR_ARG, "device index out of range");
    return PL_OK;
}

uint32_t log2_of(uint32_t x)
{
    uint32_t l = 0;
    while ((1u << l) < x) ++l;
    return l;
}

int create_ctx(uint64_t k, uint32_t n, uint32_t fft_len, int device, uint32_t capacity, pl_ctx** out)
{
    if (!out) return fail(PL_ERR_ARG, "out is NULL");
    *out = nullptr;
    return guarded([&] {
        auto c = std::make_unique<pl_ctx>();
        std::string why;
        if (!pleng::choose_plan(n, k, fft_len, c->plan, why)) return fail(PL_ERR_ARG, why);
        if (int rc = check_device(device)) return rc;
        c->device = device;
        c->M = std::make_unique<plbn::Modulus>(n, k);
        if (!rule_plan(n, k, fft_len, c->plan, why)) return fail(PL_ERR_ARG, why);
        bool turn = true;
        const bool tuned = pltune::apply(device, c->plan, 1, false, &turn);
        if (!tuned && c->plan.nt == 8) {
            // untuned: the shape whose steady state runs as one resident launch
            if (const uint32_t lr = pleng::resident_logR(device, c->plan.L)) pleng::set_shape(c->plan, lr);
        }
        c->eng = std::make_unique<pleng::Group>(c->plan, capacity + 3, device);
        if (tuned && !turn) c->eng->no_turn();
        c->scr = plprep::Scratch{ capacity, capacity + 1, capacity + 2 };
        *out = c.release();
        return PL_OK;
    });
}

}  // namespace

extern "C" {

const char* pl_last_error(void) { return g_err.c_str(); }

int pl_abi_version(void) { return PL_ABI_VERSION; }

int pl_device_count(int* count)
{
    if (!count) return fail(PL_ERR_ARG, "count is NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return PL_OK;
}

int pl_plan_fft_len(uint32_t n, uint64_t k, uint32_t* fft_len)
{
    if (!fft_len) return fail(PL_ERR_ARG, "fft_len is NULL");
    pleng::Plan p;
    std::string why;
    if (!pleng::choose_plan(n, k, 0, p, why)) return fail(PL_ERR_ARG, why);
    *fft_len = p.L;
    return PL_OK;
}

int pl_ctx_create(uint64_t k, uint32_t n, uint32_t fft_len, int device, pl_ctx** out)
{
    return create_ctx(k, n, fft_len, device, 1, out);
}

void pl_ctx_destroy(pl_ctx* ctx) { delete ctx; }

int pl_ctx_info(const pl_ctx* ctx, pl_info* out)
{
    if (!ctx || !out) return fail(PL_ERR_ARG, "NULL argument");
    out->k = ctx->plan.k;
    out->n = ctx->plan.n;
    out->fft_len = ctx->plan.L;
    out->rows = ctx->plan.R;
    out->cols = ctx->plan.C;
    out->bits_per_digit = (double)ctx->plan.n / (double)ctx->plan.L;
    out->device_bytes = ctx->eng->device_bytes();
    out->passes = (uint32_t)ctx->eng->launches_per_squaring();
    out->handoff_fallbacks = ctx->eng->handoff_fallbacks();
    out->row_radix = ctx->plan.row_radix;
    out->col_radix = ctx->plan.col_radix;
    out->run = ctx->plan.nt;
    return PL_OK;
}

int pl_set_residue(pl_ctx* ctx, const uint64_t* limbs, size_t n_limbs)
{
    if (!ctx || (!limbs && n_limbs)) return fail(PL_ERR_ARG, "NULL argument");
    return guarded([&] {
        plbn::Limbs x(limbs, limbs + n_limbs);
        if (x.empty()) x.push_back(0);
        ctx->set_value(plbn::mod_N(*ctx->M, x));
        return PL_OK;
    });
}

int pl_get_residue(pl_ctx* ctx, uint64_t* limbs, size_t n_limbs)
{
    if (!ctx || !limbs) return fail(PL_ERR_ARG, "NULL argument");
    return guarded([&] {
        plbn::Limbs x = ctx->value();
        size_t used = x.size();
        while (used > 0 && x[used - 1] == 0) --used;
        if (used > n_limbs) return fail(PL_ERR_ARG, "n_limbs too small for the residue");
        for (size_t i = 0; i < n_limbs; ++i) limbs[i] = i < x.size() ? x[i] : 0;
        return PL_OK;
    });
}

int pl_square(pl_ctx* ctx, uint64_t iters, double err_guard, double* max_err, double* err_trace)
{
    if (!ctx) return fail(PL_ERR_ARG, "NULL context");
    if (err_guard <= 0) err_guard = 0.40;
    return guarded([&] {
        double e = ctx->square(iters, err_trace);
        if (max_err) *max_err = e;
        int st = ctx->eng->status();
        if (st) ctx->eng->clear_status();
        if (st & 2) return fail(PL_ERR_STATE, "carry hand-off between workgroups timed out");
        if (st) return fail(PL_ERR_STATE, "device carry did not settle (status " + std::to_string(st) + ")");
        if (e > err_guard)
            return fail(PL_ERR_ROUNDOFF, "round-off " + std::to_string(e) + " above guard " + std::to_string(err_guard));
        return PL_OK;
    });
}

int pl_mul_small(pl_ctx* ctx, uint64_t a)
{
    if (!ctx) return fail(PL_ERR_ARG, "NULL context");
    return guarded([&] {
        // linear in the raw state, so the same in either form
        plprep::set_raw(*ctx->eng, *ctx->M, 0, plbn::mul_small_mod(*ctx->M, ctx->raw_of(0), a));
        if (ctx->normal) ctx->small.grow(false, a);
        return PL_OK;
    });
}

int pl_res64(pl_ctx* ctx, uint64_t* res64)
{
    if (!ctx || !res64) return fail(PL_ERR_ARG, "NULL argument");
    return guarded([&] {
        *res64 = plbn::low64(ctx->value());
        return PL_OK;
    });
}

int pl_is_n_minus_1(pl_ctx* ctx, int* is_nm1)
{
    if (!ctx || !is_nm1) return fail(PL_ERR_ARG, "NULL argument");
    return guarded([&] {
        *is_nm1 = plbn::is_n_minus_1(*ctx->M, ctx->value()) ? 1 : 0;
        return PL_OK;
    });
}

int pl_fft_square_karatsuba(pl_ctx* ctx, uint64_t* v, size_t n_limbs, unsigned n_bits, uint64_t k,
                            double* last_max_diff)
{
    if (!ctx || !v) return fail(PL_ERR_ARG, "NULL argument");
    if (n_bits != ctx->plan.n || k != ctx->plan.k) return fail(PL_ERR_ARG, "n_bits/k differ from the context");
    int rc = pl_set_residue(ctx, v, n_limbs);
    if (rc) return rc;
    double e = 0;
    rc = pl_square(ctx, 1, 1.0, &e, nullptr);
    if (rc) return rc;
    if (last_max_diff) *last_max_diff = e;
    return pl_get_residue(ctx, v, n_limbs);
}

int pl_kernel_times(pl_ctx* ctx, int reps, double* ms, uint64_t* traffic_bytes)
{
    if (!ctx || !ms || reps <= 0) return fail(PL_ERR_ARG, "bad argument");
    return guarded([&] {
        ctx->eng->kernel_times(reps, ms, 1);
        if (traffic_bytes) *traffic_bytes = 32ull * ctx->plan.L;
        return PL_OK;
    });
}

int pl_loop_time(pl_ctx* ctx, int reset, double* ms, uint64_t* launches, uint64_t* squarings)
{
    if (!ctx) return fail(PL_ERR_ARG, "NULL context");
    return guarded([&] {
        ctx->eng->loop_time(reset != 0, ms, launches, squarings);
        return PL_OK;
    });
}

int pl_copy_probe(int device, uint64_t bytes, int reps, double* ms)
{
    if (!ms || bytes < 16 || reps <= 0) return fail(PL_ERR_ARG, "bad argument");
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return fail(PL_ERR_HIP, "no HIP device visible");
    if (device < 0 || device >= nd) return fail(PL_ERR_ARG, "device index out of range");
    return guarded([&] {
        *ms = pleng::copy_probe(device, (size_t)bytes, reps);
        return PL_OK;
    });
}

int pl_autotune_fft_len(uint32_t n, uint64_t k, int device, double err_guard, uint32_t* fft_len, double* max_err,
                        double* us_per_squaring)
{
    if (!fft_len) return fail(PL_ERR_ARG, "fft_len is NULL");
    uint32_t L0 = 0;
    int rc = pl_plan_fft_len(n, k, &L0);
    if (rc) return rc;
    const double guard = err_guard > 0 ? err_guard : 0.30;
    uint32_t best = 0;
    double best_us = 0, best_err = 0;
    std::string last_why = "no length passed the round-off probe";
    for (uint32_t L : { L0 / 2, L0, L0 * 2 }) {
        if (L < 1024) continue;
        pl_ctx* raw = nullptr;
        if (pl_ctx_create(k, n, L, device, &raw) != PL_OK) {
            last_why = g_err;
            continue;  // this length is not allowed for n (too few or too many bits per digit)
        }
        std::unique_ptr<pl_ctx> c(raw);
        // the reference probe: all-ones limbs with limbs[0] = k|1 (P/dcore.cpp:52-56)
        std::vector<uint64_t> v((n + 63) / 64, ~0ull);
        v[0] = k | 1;
        double e = 0;
        if (pl_set_residue(c.get(), v.data(), v.size()) != PL_OK) return PL_ERR_HIP;
        rc = pl_square(c.get(), 100, 1.0, &e, nullptr);
        if (rc == PL_ERR_STATE || e > guard) continue;
        if (rc) return rc;
        const int reps = 256;
        hipEvent_t t0, t1;
        float ms = 0;
        rc = guarded([&] {
            pleng::check(hipEventCreate(&t0), "hipEventCreate");
            pleng::check(hipEventCreate(&t1), "hipEventCreate");
            pleng::check(hipEventRecord(t0, c->eng->stream()), "hipEventRecord");
            c->eng->square_slots(0, 1, reps, nullptr);
            pleng::check(hipEventRecord(t1, c->eng->stream()), "hipEventRecord");
            pleng::check(hipEventSynchronize(t1), "hipEventSynchronize");
            pleng::check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
            (void)hipEventDestroy(t0);
            (void)hipEventDestroy(t1);
            return PL_OK;
        });
        if (rc) return rc;
        const double us = 1000.0 * ms / reps;
        if (!best || us < best_us) {
            best = L;
            best_us = us;
            best_err = e;
        }
    }
    if (!best) return fail(PL_ERR_ROUNDOFF, last_why);
    // then shape, radices and outer pass at that length
    rc = guarded([&] {
        pleng::Plan p;
        std::string why;
        if (!rule_plan(n, k, best, p, why)) return fail(PL_ERR_ARG, why);
        pltune::Trial bt;
        pltune::tune(device, p, 1, false, bt);
        best_us = std::min(best_us, bt.us);
        return PL_OK;
    });
    if (rc) return rc;
    *fft_len = best;
    if (max_err) *max_err = best_err;
    if (us_per_squaring) *us_per_squaring = best_us;
    return PL_OK;
}

int pl_autotune_plan(int device, uint64_t k, uint32_t n, pl_tune_trial* trials, uint32_t max_trials,
                     uint32_t* n_trials, pl_tune_trial* best)
{
    if (int rc = check_device(device)) return rc;
    return guarded([&] {
        pleng::Plan p;
        std::string why;
        if (!rule_plan(n, k, 0, p, why)) return fail(PL_ERR_ARG, why);
        pltune::Trial b;
        const auto ts = pltune::tune(device, p, 1, false, b);
        pltune::report(ts, b, trials, max_trials, n_trials, best);
        return PL_OK;
    });
}

int pl_tune_pin(int device, uint64_t k, uint32_t n, const pl_tune_trial* choice)
{
    if (!choice) return fail(PL_ERR_ARG, "choice is NULL");
    if (device < 0) return fail(PL_ERR_ARG, "device index out of range");
    pleng::Plan p;
    std::string why;
    if (!rule_plan(n, k, 0, p, why)) return fail(PL_ERR_ARG, why);
    pltune::Choice c;
    if (choice->rows == 0 || (choice->rows & (choice->rows - 1)) || choice->rows >= p.L)
        return fail(PL_ERR_ARG, "rows must be a power of two below fft_len");
    c.logR = log2_of(choice->rows);
    if (choice->cols && choice->cols * choice->rows != p.L) return fail(PL_ERR_ARG, "rows * cols must equal fft_len");
    c.row_radix = choice->row_radix;
    c.col_radix = choice->col_radix;
    c.nt = choice->run;
    c.turn = choice->one_launch != 0;
    if ((c.row_radix != 8 && c.row_radix != 16) || (c.col_radix != 8 && c.col_radix != 16))
        return fail(PL_ERR_ARG, "radices must be 8 or 16");
    if (!(c.nt == p.nt || (p.nt == 8 && c.nt == 16)))
        return fail(PL_ERR_ARG, "run must be the carry layout's (" + std::to_string(p.nt) + ") or 16 in place of 8");
    if (c.turn && c.nt != 8) return fail(PL_ERR_ARG, "the one-launch outer pass needs runs of 8 digits");
    const auto opts = pleng::shape_options(p.logL, c.nt != 0);
    if (std::find(opts.begin(), opts.end(), c.logR) == opts.end())
        return fail(PL_ERR_ARG, "no kernels for this shape with this outer pass");
    pltune::set(pltune::key_for(device, p, 1, false), c);
    return PL_OK;
}

int pl_tune_clear(void)
{
    pltune::clear();
    return PL_OK;
}

int pl_tune_save(const char* path)
{
    if (!path) return fail(PL_ERR_ARG, "path is NULL");
    std::string why;
    if (!pltune::save(path, why)) return fail(PL_ERR_IO, why);
    return PL_OK;
}

int pl_tune_load(const char* path)
{
    if (!path) return fail(PL_ERR_ARG, "path is NULL");
    std::string why;
    if (!pltune::load(path, why)) return fail(PL_ERR_IO, why);
    return PL_OK;
}

int pl_gate_rejects(uint64_t k, uint32_t n)
{
    static const uint64_t ps[] = { 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67 };
    for (uint64_t p : ps) {
        uint64_t km = k % p;
        if (km == 0) continue;
        if ((km * powmod(2, n, p)) % p == p - 1) return 1;
    }
    return 0;
}

uint64_t pl_choose_base(uint64_t k, uint32_t n)
{
    static const uint64_t pref[] = { 2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37 };
    for (uint64_t b : pref)
        if (jacobi_small(b, k, n) == -1) return b;
    for (uint64_t b = 2; b < 100; ++b)
        if (jacobi_small(b, k, n) == -1) return b;
    return 0;
}

// ------------------------------------------------------------ checkpoints

// FNV-1a over a byte range, continuing from h.
static uint64_t fnv1a(const void* p, size_t bytes, uint64_t h = 1469598103934665603ull)
{
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < bytes; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

struct CkptV2 {
    char magic[8];
    uint64_t k;
    uint32_t n;
    uint32_t base;
    uint64_t next_iter;
    uint64_t n_limbs;
    uint64_t checksum;
    uint64_t reserved;
};
static_assert(sizeof(CkptV2) == 56 || sizeof(CkptV2) == 64, "header layout");

// The checksum covers the whole header (with the checksum field taken as 0)
// and the limbs, so a flipped k, n, base or next_iter is caught like a
// flipped limb.
static uint64_t ckpt_checksum(CkptV2 h, const uint64_t* limbs, size_t n_limbs)
{
    h.checksum = 0;
    unsigned char hdr[64] = { 0 };
    std::memcpy(hdr, &h, sizeof h);
    return fnv1a(limbs, 8 * n_limbs, fnv1a(hdr, 64));
}

int pl_save_checkpoint(const char* path, uint64_t k, uint32_t n, uint64_t base, uint64_t next_iter,
                       const uint64_t* limbs, size_t n_limbs)
{
    if (!path || !limbs) return fail(PL_ERR_ARG, "NULL argument");
    unsigned char hdr[64] = { 0 };
    CkptV2 h;
    std::memcpy(h.magic, "PLCKPT02", 8);
    h.k = k;
    h.n = n;
    h.base = (uint32_t)base;
    h.next_iter = next_iter;
    h.n_limbs = n_limbs;
    h.reserved = 0;
    h.checksum = 0;
    h.checksum = ckpt_checksum(h, limbs, n_limbs);
    std::memcpy(hdr, &h, sizeof h);
    std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return fail(PL_ERR_IO, "cannot open " + tmp);
    // written, flushed and synced to the disk before it replaces the last
    // good checkpoint (the order of P/checkpoint_manager.cpp:13-19)
    bool ok = std::fwrite(hdr, 1, 64, f) == 64 && std::fwrite(limbs, 8, n_limbs, f) == n_limbs;
    ok = ok && std::fflush(f) == 0;
    ok = ok && ::fsync(fileno(f)) == 0;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        std::remove(tmp.c_str());
        return fail(PL_ERR_IO, "short write or sync failure on " + tmp);
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return fail(PL_ERR_IO, "rename failed for " + tmp);
    }
    // the rename itself reaches the disk with the directory's sync
    std::string dir = path;
    const size_t slash = dir.find_last_of('/');
    dir = slash == std::string::npos ? "." : (slash == 0 ? "/" : dir.substr(0, slash));
    const int dfd = ::open(dir.c_str(), O_RDONLY | O_DIRECTORY);
    if (dfd >= 0) {
        (void)::fsync(dfd);
        ::close(dfd);
    }
    return PL_OK;
}

int pl_load_checkpoint(const char* path, uint64_t k, uint32_t n, uint64_t* base, uint64_t* next_iter,
                       uint64_t* limbs, size_t n_limbs)
{
    if (!path || !limbs || !next_iter) return fail(PL_ERR_ARG, "NULL argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(PL_ERR_IO, std::string("cannot open ") + path);
    std::fseek(f, 0, SEEK_END);
    long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    unsigned char hdr[64] = { 0 };
    size_t got = std::fread(hdr, 1, 64, f);
    int rc = PL_OK;
    if (got >= 16 && std::memcmp(hdr, "PLCKPT02", 8) == 0 && got == 64) {
        CkptV2 h;
        std::memcpy(&h, hdr, sizeof h);
        if (h.k != k || h.n != n) rc = fail(PL_ERR_IO, "checkpoint is for another candidate");
        else if ((long)(64 + 8 * h.n_limbs) != size) rc = fail(PL_ERR_IO, "checkpoint truncated");
        else {
            std::vector<uint64_t> buf(h.n_limbs);
            std::fseek(f, 64, SEEK_SET);
            if (std::fread(buf.data(), 8, h.n_limbs, f) != h.n_limbs) rc = fail(PL_ERR_IO, "short read");
            else if (ckpt_checksum(h, buf.data(), buf.size()) != h.checksum)
                rc = fail(PL_ERR_IO, "checkpoint checksum mismatch");
            else {
                for (size_t i = 0; i < n_limbs; ++i) limbs[i] = i < buf.size() ? buf[i] : 0;
                for (size_t i = n_limbs; i < buf.size(); ++i)
                    if (buf[i]) rc = fail(PL_ERR_ARG, "n_limbs too small for the checkpoint");
                if (rc == PL_OK) {
                    *next_iter = h.next_iter;
                    if (base) *base = h.base;
                }
            }
        }
    } else if (got >= 16) {
        // Reference layout {u64 current_i; size_t n_limbs} + limbs; current_i
        // is the index of the last squaring applied (P/proth_logic.cpp:536-537).
        uint64_t cur_i, nl;
        std::memcpy(&cur_i, hdr, 8);
        std::memcpy(&nl, hdr + 8, 8);
        if ((long)(16 + 8 * nl) != size) rc = fail(PL_ERR_IO, "reference checkpoint truncated");
        else {
            std::vector<uint64_t> buf(nl);
            std::fseek(f, 16, SEEK_SET);
            if (std::fread(buf.data(), 8, nl, f) != nl) rc = fail(PL_ERR_IO, "short read");
            else {
                for (size_t i = 0; i < n_limbs; ++i) limbs[i] = i < buf.size() ? buf[i] : 0;
                for (size_t i = n_limbs; i < buf.size(); ++i)
                    if (buf[i]) rc = fail(PL_ERR_ARG, "n_limbs too small for the checkpoint");
                if (rc == PL_OK) {
                    *next_iter = cur_i + 1;
                    if (base) *base = 0;
                }
            }
        }
    } else {
        rc = fail(PL_ERR_IO, "checkpoint header truncated");
    }
    std::fclose(f);
    return rc;
}

// ------------------------------------------------------------ full test

int pl_proth_test(const pl_test_args* args, pl_test_result* out)
{
    if (!args || !out) return fail(PL_ERR_ARG, "NULL argument");
    std::memset(out, 0, sizeof *out);
    const uint64_t k = args->k;
    const uint32_t n = args->n;
    if (pl_gate_rejects(k, n)) {
        out->gated = 1;
        out->verdict = 0;
        return PL_OK;
    }
    const uint64_t a = pl_choose_base(k, n);
    out->base = a;
    // slots: the residue, plus for the Gerbicz check the running product,
    // the epoch's first residue, the check value and a scratch copy
    enum { kX = 0, kD = 1, kU0 = 2, kCk = 3, kT = 4 };
    const bool gerbicz = args->gerbicz_block > 0;
    pl_ctx* raw = nullptr;
    int rc = create_ctx(k, n, args->fft_len, args->device, gerbicz ? 5 : 1, &raw);
    if (rc) return rc;
    std::unique_ptr<pl_ctx> ctx(raw);
    out->fft_len = ctx->plan.L;
    const double guard = args->err_guard > 0 ? args->err_guard : 0.40;
    const uint64_t every = args->checkpoint_every ? args->checkpoint_every : 10000;
    const uint64_t total = n - 1;
    const size_t nl = ctx->M->nl;
    return guarded([&] {
        uint64_t next = 0;
        bool resumed = false;
        if (args->checkpoint_path) {
            std::vector<uint64_t> buf(nl, 0);
            uint64_t b = 0, nx = 0;
            FILE* probe = std::fopen(args->checkpoint_path, "rb");
            if (probe) {
                std::fclose(probe);
                int lr = pl_load_checkpoint(args->checkpoint_path, k, n, &b, &nx, buf.data(), nl);
                if (lr != PL_OK) return lr;  // a corrupt or foreign file is an error, not a silent restart
                if (nx > total) return fail(PL_ERR_IO, "checkpoint iteration beyond n-1");
                ctx->set_value(plbn::mod_N(*ctx->M, plbn::Limbs(buf.begin(), buf.end())));
                next = nx;
                resumed = true;
            }
        }
        if (!resumed) {
            // seed a^k mod N on the device, left to right from x = a: one
            // squaring per lower bit of k, times a where the bit is set, in
            // normal form (the values are small)
            if (a == 0) {
                ctx->set_value(plbn::Limbs(1, 0));
            } else {
                ctx->set_value(plbn::Limbs(1, a));
                unsigned top = 0;
                while (top < 63 && (k >> (top + 1))) ++top;
                for (int b = (int)top - 1; b >= 0; --b)
                    out->max_err = std::max(out->max_err, ctx->normal_step(((k >> b) & 1) != 0, a));
                if (ctx->eng->status()) return fail(PL_ERR_STATE, "device carry did not settle (seed)");
            }
        }
        out->resumed_at = next;
        // progress points: squaring count t at which floor(t*100/total) first
        // reaches the next multiple of report_step (P/proth_logic.cpp:541-548)
        const int step_pct = args->progress && args->report_step > 0 ? args->report_step : 0;
        uint64_t next_pct = 0, report_at = total + 1;
        auto arm = [&](uint64_t done) {
            if (!step_pct) return;
            const uint64_t pct = done * 100 / total;
            next_pct = (pct / (uint64_t)step_pct + 1) * (uint64_t)step_pct;
            report_at = next_pct > 100 ? total + 1 : (next_pct * total + 99) / 100;
        };
        arm(next);
        // test hook: a one-unit error in the residue, once
        bool injected = false;
        auto maybe_inject = [&] {
            if (!args->inject_error_at || injected || next <= args->inject_error_at) return;
            plbn::Limbs y = ctx->raw_of(0);
            plbn::Limbs one(y.size(), 0);
            one[0] = 1;
            plbn::add_in(y, one);
            plprep::set_raw(*ctx->eng, *ctx->M, 0, plbn::mod_N(*ctx->M, y));
            injected = true;
        };
        if (gerbicz) {
            pleng::Group& G = *ctx->eng;
            const uint64_t B = args->gerbicz_block;
            const uint64_t Mb = args->gerbicz_blocks ? args->gerbicz_blocks : 64;
            const uint64_t end = args->stop_after ? std::min<uint64_t>(total, next + args->stop_after) : total;
            // the epochs run on Montgomery forms: first the squarings that
            // still have a small exact value (a few dozen at most)
            {
                double e = 0.0;
                const uint64_t d = ctx->finish_normal(end - next, e);
                out->max_err = std::max(out->max_err, e);
                if (G.status()) return fail(PL_ERR_STATE, "device carry did not settle (normal form)");
                next += d;
                out->iters_done += d;
                if (ctx->normal && next < end) {
                    // the value stays small (0 or 1): the squarings go on in normal form
                    e = ctx->square(end - next, nullptr);
                    out->max_err = std::max(out->max_err, e);
                    out->iters_done += end - next;
                    next = end;
                }
            }
            uint64_t verified = next, last_ckpt = next;
            int failures = 0;
            auto start_epoch = [&] {
                G.copy_slot(kU0, kX);
                G.copy_slot(kD, kX);
            };
            start_epoch();
            while (next < end) {
                const uint64_t rem = end - next;
                const uint64_t b = std::min<uint64_t>(B, rem);
                const uint64_t m = std::max<uint64_t>(1, std::min<uint64_t>(Mb, rem / b));
                double e = 0.0;
                for (uint64_t j = 1; j <= m; ++j) {
                    e = std::max(e, G.square_slots(kX, 1, b, nullptr));
                    next += b;
                    maybe_inject();
                    if (j == m) G.copy_slot(kCk, kD);  // d_(m-1)
                    G.copy_slot(kT, kX);
                    e = std::max(e, G.mul_slot(kD, kT));  // d_j = d_(j-1) * u_j
                }
                e = std::max(e, G.square_slots(kCk, 1, b, nullptr));
                G.copy_slot(kT, kU0);
                e = std::max(e, G.mul_slot(kCk, kT));  // d_(m-1)^(2^b) * u_0
                const bool settled = G.status() == 0;
                if (!settled) G.clear_status();
                const bool ok = settled && e <= guard && ctx->mont_of(kCk) == ctx->mont_of(kD);
                ++out->gerbicz_checks;
                if (!ok) {
                    ++out->gerbicz_rollbacks;
                    if (++failures >= 3)
                        return fail(e > guard ? PL_ERR_ROUNDOFF : PL_ERR_STATE,
                                    "Gerbicz check failed three times in a row at squaring " + std::to_string(verified));
                    G.copy_slot(kX, kU0);
                    next = verified;
                    start_epoch();
                    continue;
                }
                failures = 0;
                out->max_err = std::max(out->max_err, e);
                out->iters_done += next - verified;
                verified = next;
                while (step_pct && next >= report_at) {
                    args->progress(args->progress_user, next, total, out->max_err);
                    arm(next);
                }
                if (args->checkpoint_path && next < total && next - last_ckpt >= every) {
                    plbn::Limbs x = ctx->value();
                    int s = pl_save_checkpoint(args->checkpoint_path, k, n, a, next, x.data(), x.size());
                    if (s) return s;
                    last_ckpt = next;
                }
                if (next < end) start_epoch();
            }
            if (args->stop_after && next < total && args->checkpoint_path) {
                plbn::Limbs x = ctx->value();
                int s = pl_save_checkpoint(args->checkpoint_path, k, n, a, next, x.data(), x.size());
                if (s) return s;
            }
        }
        while (next < total && !gerbicz) {
            uint64_t chunk = std::min<uint64_t>(every - (next % every), total - next);
            if (report_at > next && report_at - next < chunk) chunk = report_at - next;
            if (args->stop_after && out->iters_done + chunk > args->stop_after)
                chunk = args->stop_after - out->iters_done;
            if (chunk == 0) break;
            double e = ctx->square(chunk, nullptr);
            if (e > out->max_err) out->max_err = e;
            int st = ctx->eng->status();
            if (st) return fail(PL_ERR_STATE, "device carry did not settle");
            if (e > guard) return fail(PL_ERR_ROUNDOFF, "round-off " + std::to_string(e) + " above guard");
            next += chunk;
            out->iters_done += chunk;
            maybe_inject();
            if (step_pct && next >= report_at) {
                args->progress(args->progress_user, next, total, out->max_err);
                arm(next);
            }
            if (args->checkpoint_path && next < total && (next % every == 0 || args->stop_after)) {
                plbn::Limbs x = ctx->value();
                int s = pl_save_checkpoint(args->checkpoint_path, k, n, a, next, x.data(), x.size());
                if (s) return s;
            }
            if (args->stop_after && out->iters_done >= args->stop_after) break;
        }
        plbn::Limbs x = ctx->value();
        out->res64 = plbn::low64(x);
        if (args->final_limbs) {
            for (size_t i = 0; i < args->final_n_limbs; ++i) args->final_limbs[i] = i < x.size() ? x[i] : 0;
            for (size_t i = args->final_n_limbs; i < x.size(); ++i)
                if (x[i]) return fail(PL_ERR_ARG, "final_n_limbs cannot hold the residue");
        }
        if (next < total) {
            out->verdict = -1;
            return PL_OK;
        }
        out->verdict = plbn::is_n_minus_1(*ctx->M, x) ? 1 : 0;
        if (args->checkpoint_path) std::remove(args->checkpoint_path);
        return PL_OK;
    });
}

}  // extern "C"
