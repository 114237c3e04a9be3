// pool.cpp -- continuous batching of Proth tests on one GPU (pl_pool_* in
// include/prothcore.h).
//
// Candidates wait in a FIFO.  Admission places a candidate in a free slot of
// the group for its transform length, creating that group when its whole
// allocation fits in the remaining HBM budget.  Every run() call advances
// all active slots of every group by scheduled steps (seed powering a^k,
// then n-1 squarings, all on the device), then retires the slots whose tests
// have finished and reports their verdicts.  Freed slots are refilled at the
// next admission, so requests that arrive while others are on the GPU join
// the running batch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "bignum.h"
#include "engine.h"
#include "prewrap.h"
#include "prothcore.h"
#include "tune.h"

namespace plpool {

using clk = std::chrono::steady_clock;

struct Request {
    uint64_t k;
    uint32_t n;
    uint64_t tag;
    uint64_t stop_after;
    double t_submit;
    uint32_t attempts;  // runs restarted after the device carry did not settle
};

// A test whose round-off passes this (as pl_proth_test's default guard) has
// no reliable residue; its result is reported void.
constexpr double kErrGuard = 0.40;
// Restarts from the seed after a group's carry did not settle, before the
// test is reported void.
constexpr uint32_t kMaxAttempts = 2;

struct Occupant {
    bool used = false;
    Request req{};
    uint64_t base = 0;
    uint64_t total = 0;
    double t_admit = 0;
    double max_err = 0;
    std::unique_ptr<plbn::Modulus> M;
};

struct LenGroup {
    pleng::Plan plan;
    std::unique_ptr<pleng::Group> g;
    std::vector<Occupant> slots;  // the group's test slots; three scratch slots follow them
    plprep::Scratch scr;
};

}  // namespace plpool

using namespace plpool;

struct pl_pool {
    int device = 0;
    uint64_t budget = 0;
    uint64_t used_bytes = 0;
    uint32_t max_slots = 0;
    uint64_t k_max = 0;
    // by (log2 L, digits per carry run): a test joins the group of its
    // length whose runs absorb the carries at its own digit size
    std::map<std::pair<uint32_t, uint32_t>, std::unique_ptr<LenGroup>> groups;
    std::deque<Request> queue;
    std::vector<pl_pool_result> done;
    clk::time_point t0 = clk::now();
    double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }
};

namespace {
thread_local std::string g_pool_err;
int pfail(int code, const std::string& m)
{
    g_pool_err = m;
    return code;
}

unsigned topbit64(uint64_t k)
{
    unsigned t = 0;
    while (t < 63 && (k >> (t + 1))) ++t;
    return t;
}

// The group for a test: its transform length, and carry runs of 8 digits
// when those absorb the carries at the test's own digit size (for any k up
// to k_max), else 16.  A group holds only tests whose digit size is at
// least the one its layout was chosen for: runs of 8 at floor(n / L) bits,
// runs of 16 at the fewest bits the length admits (n just above the
// previous length's limit).  Steps carry no multiplier (the seed is powered
// in normal form before a slot starts stepping), so the bound takes 1.
// The built-in plan and the map key of the group a test joins.
bool group_plan(const pl_pool* P, uint32_t n, uint64_t k, pleng::Plan& pl, std::pair<uint32_t, uint32_t>& key,
                std::string& why)
{
    if (!pleng::choose_plan(n, k, 0, pl, why)) return false;
    pleng::Plan own = pl;
    if (!pleng::choose_layout(own, std::floor((double)n / pl.L), P->k_max, 1, true, why)) return false;
    const uint32_t nt = own.nt == 8 ? 8u : 16u;
    key = std::make_pair(pl.logL, nt);
    if (nt == 8) {
        pl = own;
    } else {
        const double bmin = pl.logL > 10 ? pleng::max_bits_per_digit(pl.logL - 1, P->k_max) / 2.0 : 7.0;
        if (!pleng::choose_layout(pl, std::max(7.0, bmin), P->k_max, 1, true, why)) return false;
    }
    return true;
}

// Slots a new group with this plan gets from what is left of the budget.
uint32_t group_capacity(const pl_pool* P, const pleng::Plan& pl)
{
    const size_t per_slot = pleng::Group::slot_bytes(pl);
    const uint64_t left = P->budget > P->used_bytes ? P->budget - P->used_bytes : 0;
    return (uint32_t)std::min<uint64_t>(P->max_slots, left / (per_slot + 1));
}

LenGroup* group_for(pl_pool* P, uint32_t n, uint64_t k, std::string& why)
{
    pleng::Plan pl;
    std::pair<uint32_t, uint32_t> key;
    if (!group_plan(P, n, k, pl, key, why)) return nullptr;
    auto it = P->groups.find(key);
    if (it != P->groups.end()) return it->second.get();
    const uint32_t cap = group_capacity(P, pl);
    if (cap == 0) {
        why = "HBM budget exhausted";
        return nullptr;
    }
    pltune::apply(P->device, pl, cap, true, nullptr);  // a measured plan for this key, if one was recorded
    auto lg = std::make_unique<LenGroup>();
    lg->plan = pl;
    lg->g = std::make_unique<pleng::Group>(pl, cap + 3, P->device);
    lg->g->no_turn();  // groups step concurrently on their own streams
    lg->slots.resize(cap);
    lg->scr = plprep::Scratch{ cap, cap + 1, cap + 2 };
    P->used_bytes += lg->g->device_bytes();
    LenGroup* raw = lg.get();
    P->groups[key] = std::move(lg);
    return raw;
}

pl_pool_result void_result(pl_pool* P, const Request& q, uint32_t L, int error, double max_err);
void evict(pl_pool* P, LenGroup& lg, uint32_t s);

pl_pool_result finished_result(pl_pool* P, const Occupant& o, uint32_t L, const plbn::Limbs& x)
{
    pl_pool_result r{};
    r.tag = o.req.tag;
    r.k = o.req.k;
    r.n = o.req.n;
    r.fft_len = L;
    r.base = o.base;
    r.res64 = plbn::low64(x);
    r.max_err = o.max_err;
    r.squarings = o.total;
    const bool full = !o.req.stop_after || o.req.stop_after >= (uint64_t)o.req.n - 1;
    r.verdict = full ? (plbn::is_n_minus_1(*o.M, x) ? 1 : 0) : -1;
    r.gated = 0;
    r.error = PL_OK;
    r.t_submit = o.req.t_submit;
    r.t_admit = o.t_admit;
    r.t_done = P->now();
    return r;
}

// The first squarings of an admitted test, while its exact value is small:
// the seed a^k and then squarings until the value has wrapped modulo N, in
// normal form on the slot (prewrap.h), then the move to the Montgomery form
// the scheduled steps use.  Returns the squarings applied; all of them
// (o.total) when the test ends before its value wraps, the slot then still
// holding the value itself.
uint64_t prewrap(LenGroup& lg, uint32_t s, Occupant& o, bool& normal)
{
    pleng::Group& G = *lg.g;
    const plbn::Modulus& M = *o.M;
    for (uint32_t t : { lg.scr.t, lg.scr.m1, lg.scr.ma }) G.set_slot(t, o.req.n, o.req.k, 0, 0);
    plprep::load_constants(G, M, o.base, lg.scr);
    plprep::SmallValue v;
    normal = v.set(M, plbn::Limbs(1, o.base));
    plprep::set_raw(G, M, s, plbn::Limbs(1, o.base));
    const unsigned top = topbit64(o.req.k);
    uint64_t done = 0;
    for (int b = (int)top - 1; b >= 0 && done < o.total; --b, ++done) {
        const bool bit = ((o.req.k >> b) & 1) != 0;
        o.max_err = std::max(o.max_err, plprep::step(G, s, lg.scr, bit));
        v.grow(true, bit ? o.base : 0);
    }
    while (!v.wrapped(M) && done < o.total) {
        o.max_err = std::max(o.max_err, plprep::step(G, s, lg.scr, false));
        v.grow(true, 0);
        ++done;
    }
    if (done < o.total) {
        o.max_err = std::max(o.max_err, plprep::to_mont(G, s, lg.scr));
        normal = false;
    }
    return done;
}

void admit(pl_pool* P)
{
    // slots whose carry did not settle during their first squarings; evicted
    // (and requeued) after the walk over the queue
    std::vector<std::pair<LenGroup*, uint32_t>> failed;
    for (auto it = P->queue.begin(); it != P->queue.end();) {
        const Request r = *it;
        std::string why;
        LenGroup* lg = group_for(P, r.n, r.k, why);
        if (!lg) {
            ++it;  // no room for this length yet; later candidates may still fit
            continue;
        }
        int free_slot = -1;
        for (size_t s = 0; s < lg->slots.size(); ++s)
            if (!lg->slots[s].used) {
                free_slot = (int)s;
                break;
            }
        if (free_slot < 0) {
            ++it;
            continue;
        }
        Occupant& o = lg->slots[free_slot];
        o.used = true;
        o.req = r;
        o.base = pl_choose_base(r.k, r.n);
        o.M = std::make_unique<plbn::Modulus>(r.n, r.k);
        o.t_admit = P->now();
        o.max_err = 0;
        const unsigned top = topbit64(r.k);
        const uint64_t sq = r.stop_after ? std::min<uint64_t>(r.stop_after, r.n - 1) : r.n - 1;
        o.total = top + sq;
        const uint32_t s = (uint32_t)free_slot;
        lg->g->set_slot(s, r.n, r.k, (uint32_t)o.base, o.total);
        bool normal = false;
        const uint64_t done = prewrap(*lg, s, o, normal);
        it = P->queue.erase(it);
        if (lg->g->status()) {
            lg->g->clear_status();
            failed.emplace_back(lg, s);
            continue;
        }
        if (done >= o.total) {
            // the whole test ran before its value wrapped: the slot holds x
            const plbn::Limbs x = plprep::raw(*lg->g, *o.M, s);
            if (!(o.max_err <= kErrGuard)) P->done.push_back(void_result(P, o.req, lg->plan.L, PL_ERR_ROUNDOFF, o.max_err));
            else P->done.push_back(finished_result(P, o, lg->plan.L, normal ? x : plbn::from_mont(*o.M, x)));
            o.used = false;
            o.M.reset();
            continue;
        }
        // the scheduled steps carry the rest (the seed's multipliers are done)
        lg->g->set_state(s, done, done + 1 < o.total);
        lg->g->enter(s);
    }
    for (auto& f : failed) evict(P, *f.first, f.second);
}

pl_pool_result void_result(pl_pool* P, const Request& q, uint32_t L, int error, double max_err)
{
    pl_pool_result r{};
    r.tag = q.tag;
    r.k = q.k;
    r.n = q.n;
    r.fft_len = L;
    r.verdict = -2;
    r.error = error;
    r.max_err = max_err;
    r.base = pl_choose_base(q.k, q.n);
    r.t_submit = q.t_submit;
    r.t_done = P->now();
    return r;
}

// Takes a slot's test off the device after its group's carry did not settle:
// the test starts again from its seed (ahead of later arrivals), or is
// reported void once it has been restarted kMaxAttempts times.
void evict(pl_pool* P, LenGroup& lg, uint32_t s)
{
    Occupant& o = lg.slots[s];
    Request q = o.req;
    lg.g->set_state(s, 0, false);
    if (++q.attempts > kMaxAttempts) P->done.push_back(void_result(P, q, lg.plan.L, PL_ERR_STATE, o.max_err));
    else P->queue.push_front(q);
    o.used = false;
    o.M.reset();
}

void evict_group(pl_pool* P, LenGroup& lg)
{
    lg.g->clear_status();
    for (uint32_t s = 0; s < lg.slots.size(); ++s)
        if (lg.slots[s].used) evict(P, lg, s);
}

void retire(pl_pool* P)
{
    for (auto& kv : P->groups) {
        LenGroup& lg = *kv.second;
        const auto st = lg.g->states();
        for (size_t s = 0; s < lg.slots.size(); ++s) {
            Occupant& o = lg.slots[s];
            if (!o.used || st[s].active) continue;
            o.max_err = std::max(o.max_err, lg.g->leave((uint32_t)s));
            if (lg.g->status()) {
                // the last squaring's carry did not settle: no reliable residue
                lg.g->clear_status();
                evict(P, lg, (uint32_t)s);
                continue;
            }
            if (!(o.max_err <= kErrGuard)) {
                P->done.push_back(void_result(P, o.req, lg.plan.L, PL_ERR_ROUNDOFF, o.max_err));
                P->done.back().t_admit = o.t_admit;
                P->done.back().squarings = o.total;
                o.used = false;
                o.M.reset();
                continue;
            }
            const plbn::Limbs x = plbn::from_mont(*o.M, plprep::raw(*lg.g, *o.M, (uint32_t)s));
            const pl_pool_result r = finished_result(P, o, lg.plan.L, x);
            P->done.push_back(r);
            o.used = false;
            o.M.reset();
        }
    }
}

size_t active_count(pl_pool* P)
{
    size_t a = 0;
    for (auto& kv : P->groups)
        for (auto& o : kv.second->slots) a += o.used ? 1 : 0;
    return a;
}

}  // namespace

extern "C" {

const char* pl_pool_last_error(void) { return g_pool_err.c_str(); }

int pl_pool_create(int device, uint64_t hbm_budget_bytes, uint32_t max_slots_per_length, uint64_t k_max,
                   pl_pool** out)
{
    if (!out) return pfail(PL_ERR_ARG, "out is NULL");
    *out = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return pfail(PL_ERR_HIP, "no HIP device visible");
    if (device < 0 || device >= nd) return pfail(PL_ERR_ARG, "device index out of range");
    auto P = std::make_unique<pl_pool>();
    P->device = device;
    P->budget = hbm_budget_bytes ? hbm_budget_bytes : (64ull << 30);
    P->max_slots = max_slots_per_length ? max_slots_per_length : 256;
    P->k_max = k_max ? k_max : (1ull << 32) - 1;
    *out = P.release();
    return PL_OK;
}

void pl_pool_destroy(pl_pool* P) { delete P; }

int pl_pool_submit(pl_pool* P, uint64_t k, uint32_t n, uint64_t tag, uint64_t stop_after)
{
    if (!P) return pfail(PL_ERR_ARG, "NULL pool");
    if (k == 0 || (k & 1) == 0 || k > P->k_max) return pfail(PL_ERR_ARG, "k must be odd and at most k_max");
    pleng::Plan pl;
    std::string why;
    if (!pleng::choose_plan(n, k, 0, pl, why)) return pfail(PL_ERR_ARG, why);
    const double t = P->now();
    if (pl_gate_rejects(k, n)) {
        pl_pool_result r{};
        r.tag = tag;
        r.k = k;
        r.n = n;
        r.gated = 1;
        r.verdict = 0;
        r.error = PL_OK;
        r.t_submit = r.t_admit = r.t_done = t;
        P->done.push_back(r);
        return PL_OK;
    }
    P->queue.push_back(Request{ k, n, tag, stop_after, t, 0 });
    return PL_OK;
}

int pl_pool_run(pl_pool* P, uint32_t steps, pl_pool_result* out, size_t cap, size_t* n_out)
{
    if (!P || (!out && cap) || !n_out) return pfail(PL_ERR_ARG, "NULL argument");
    try {
        admit(P);
        // Every group with work enqueues a chunk on its own stream before any
        // is read back, so the groups' kernels overlap on the device.  A
        // chunk launches only the group's slots up to its highest occupied
        // one (admission fills the lowest free slot first), so a sparsely
        // filled group does not dispatch blocks for all its capacity.
        std::vector<LenGroup*> live;
        std::vector<uint32_t> span;
        for (auto& kv : P->groups) {
            uint32_t hi = 0;
            for (size_t s = 0; s < kv.second->slots.size(); ++s)
                if (kv.second->slots[s].used) hi = (uint32_t)s + 1;
            if (hi) {
                live.push_back(kv.second.get());
                span.push_back(hi);
            }
        }
        // Groups step independently: each enqueues its next chunk as soon as
        // its last one is read back, so a group of short transforms is not
        // held to the pace of a group of long ones.  The call returns once
        // every group has stepped at least `steps` times; groups that get
        // there first keep stepping meanwhile (their tests keep advancing).
        std::vector<uint32_t> done(live.size(), 0);
        std::vector<uint32_t> inflight(live.size(), 0);
        auto behind = [&]() {
            for (size_t i = 0; i < live.size(); ++i)
                if (done[i] < steps) return true;
            return false;
        };
        auto launch = [&](size_t i) {
            const uint32_t want = done[i] < steps ? steps - done[i] : pleng::Group::kStepChunk;
            const uint32_t g = std::min<uint32_t>(pleng::Group::kStepChunk, want);
            live[i]->g->step_begin(g, span[i]);
            inflight[i] = g;
        };
        for (size_t i = 0; i < live.size(); ++i) launch(i);
        for (;;) {
            bool waiting = false;
            for (size_t i = 0; i < live.size(); ++i) {
                if (!inflight[i]) continue;
                LenGroup* lg = live[i];
                if (!lg->g->step_ready()) {
                    waiting = true;
                    continue;
                }
                // step_end reports every slot of the group, scratch slots included
                std::vector<double> se(lg->g->capacity(), 0.0);
                lg->g->step_end(se.data());
                for (size_t s = 0; s < lg->slots.size(); ++s)
                    if (lg->slots[s].used) lg->slots[s].max_err = std::max(lg->slots[s].max_err, se[s]);
                done[i] += inflight[i];
                inflight[i] = 0;
                if (lg->g->status()) {
                    done[i] = std::max(done[i], steps);  // evicted below; no more steps on a void state
                    continue;
                }
                if (behind()) launch(i);
            }
            bool any = false;
            for (uint32_t f : inflight) any = any || f != 0;
            if (!any) break;
            if (waiting) std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        // A group whose carry did not settle in these steps holds no reliable
        // residue in any slot: its tests restart from their seeds.
        for (LenGroup* lg : live)
            if (lg->g->status()) evict_group(P, *lg);
        retire(P);
        admit(P);
    } catch (const std::exception& e) {
        return pfail(PL_ERR_HIP, e.what());
    }
    const size_t m = std::min(cap, P->done.size());
    for (size_t i = 0; i < m; ++i) out[i] = P->done[i];
    P->done.erase(P->done.begin(), P->done.begin() + (long)m);
    *n_out = m;
    return PL_OK;
}

int pl_pool_autotune(pl_pool* P, uint64_t k, uint32_t n, pl_tune_trial* trials, uint32_t max_trials,
                     uint32_t* n_trials, pl_tune_trial* best)
{
    if (!P) return pfail(PL_ERR_ARG, "NULL pool");
    try {
        pleng::Plan pl;
        std::pair<uint32_t, uint32_t> key;
        std::string why;
        if (!group_plan(P, n, k, pl, key, why)) return pfail(PL_ERR_ARG, why);
        const uint32_t cap = group_capacity(P, pl);
        if (cap == 0) return pfail(PL_ERR_ARG, "HBM budget exhausted");
        pltune::Trial b;
        pltune::Choice have;
        if (pltune::get(pltune::key_for(P->device, pl, cap, true), have)) {
            // tuned already (by this pool or another in the process): report it
            b.c = have;
            pleng::Plan tp = pl;
            pltune::apply_choice(tp, have);
            b.R = tp.R;
            b.C = tp.C;
            b.exact = true;
            pltune::report({}, b, trials, max_trials, n_trials, best);
            return PL_OK;
        }
        const auto ts = pltune::tune(P->device, pl, cap, true, b);
        pltune::report(ts, b, trials, max_trials, n_trials, best);
        return PL_OK;
    } catch (const std::exception& e) {
        return pfail(PL_ERR_HIP, e.what());
    }
}

int pl_pool_status(pl_pool* P, size_t* queued, size_t* active, size_t* finished_unread, uint64_t* hbm_bytes)
{
    if (!P) return pfail(PL_ERR_ARG, "NULL pool");
    if (queued) *queued = P->queue.size();
    if (active) *active = active_count(P);
    if (finished_unread) *finished_unread = P->done.size();
    if (hbm_bytes) *hbm_bytes = P->used_bytes;
    return PL_OK;
}

}  // extern "C"
