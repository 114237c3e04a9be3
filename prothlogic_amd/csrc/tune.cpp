// tune.cpp -- the measured plan table (see tune.h).
#include "tune.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <random>
#include <sstream>
#include <stdexcept>
#include <tuple>

#include "bignum.h"
#include "prewrap.h"

namespace pltune {

namespace {

std::mutex g_mu;
std::map<Key, Choice> g_table;

// The reference's autotune guard (P/dcore.cpp:63): a candidate whose
// round-off passes it is not taken.
constexpr double kGuard = 0.30;

uint32_t log2u(uint32_t x)
{
    uint32_t l = 0;
    while ((1u << l) < x) ++l;
    return l;
}

bool same(const Choice& a, const Choice& b)
{
    return a.logR == b.logR && a.row_radix == b.row_radix && a.col_radix == b.col_radix && a.nt == b.nt &&
           a.turn == b.turn;
}

// One alternative, timed: the same random start in every slot set, one
// untimed chunk (it captures the chunk graph), then `reps` timed chunks on
// the group's stream.  The result of slot 0 is compared with the first
// trial's (the rule plan's).
struct Bench {
    int device;
    pleng::Plan rule;
    uint32_t slots;
    bool split_only;
    plbn::Modulus M;
    std::vector<plbn::Limbs> start;
    uint32_t reps;
    plbn::Limbs ref;
    bool have_ref = false;

    Bench(int dev, const pleng::Plan& p, uint32_t s, bool so) : device(dev), rule(p), slots(s), split_only(so), M(p.n, p.k)
    {
        std::mt19937_64 rng(0x9e3779b97f4a7c15ull ^ ((uint64_t)p.n << 20) ^ p.k);
        for (uint32_t i = 0; i < slots; ++i) {
            plbn::Limbs x(M.nl, 0);
            for (size_t j = 0; j < (p.n + 63) / 64 + 1 && j < x.size(); ++j) x[j] = rng();
            start.push_back(plbn::mod_N(M, x));
        }
        // at least ~4 ms of timed work: a squaring step costs about the
        // larger of the launch floor (~8 us) and ~70 bytes per point of HBM
        // traffic at ~4 TB/s
        const double est_us = std::max(8.0, 70.0 * (double)p.L * slots / 4.0e6);
        const double chunk_ms = est_us * pleng::Group::kChunk / 1000.0;
        reps = chunk_ms >= 4.0 ? 1u : (uint32_t)std::min(8.0, 4.0 / chunk_ms) + 1u;
    }

    Trial run(const Choice& c)
    {
        Trial t;
        t.c = c;
        pleng::Plan p = rule;
        apply_choice(p, c);
        t.R = p.R;
        t.C = p.C;
        pleng::Group g(p, slots, device);
        if (!c.turn || split_only) g.no_turn();
        for (uint32_t s = 0; s < slots; ++s) plprep::set_raw(g, M, s, start[s]);
        const uint64_t chunk = pleng::Group::kChunk;
        g.square_slots(0, slots, chunk, nullptr);
        hipEvent_t e0, e1;
        pleng::check(hipEventCreate(&e0), "hipEventCreate");
        pleng::check(hipEventCreate(&e1), "hipEventCreate");
        float ms = 0;
        try {
            pleng::check(hipEventRecord(e0, g.stream()), "hipEventRecord");
            t.max_err = g.square_slots(0, slots, chunk * reps, nullptr);
            pleng::check(hipEventRecord(e1, g.stream()), "hipEventRecord");
            pleng::check(hipEventSynchronize(e1), "hipEventSynchronize");
            pleng::check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (g.status() != 0) throw std::runtime_error("tuning trial left a device status error");
        t.us = 1000.0 * ms / (double)(chunk * reps);
        const plbn::Limbs v = plprep::raw(g, M, 0);
        if (!have_ref) {
            ref = v;
            have_ref = true;
        }
        t.exact = v == ref;
        return t;
    }
};

}  // namespace

bool Key::operator<(const Key& o) const
{
    return std::tie(device, logL, nt, slots_class, split_only) <
           std::tie(o.device, o.logL, o.nt, o.slots_class, o.split_only);
}

uint32_t slots_class(uint32_t slots) { return slots <= 1 ? 0 : slots <= 8 ? 1 : slots <= 64 ? 2 : 3; }

Key key_for(int device, const pleng::Plan& p, uint32_t slots, bool split_only)
{
    Key k;
    k.device = device;
    k.logL = p.logL;
    k.nt = p.nt;
    k.slots_class = slots_class(slots);
    k.split_only = split_only;
    return k;
}

void apply_choice(pleng::Plan& p, const Choice& c)
{
    if (c.logR) pleng::set_shape(p, c.logR);
    p.row_radix = c.row_radix;
    p.col_radix = c.col_radix;
    p.nt = c.nt;
}

bool apply(int device, pleng::Plan& p, uint32_t slots, bool split_only, bool* turn)
{
    Choice c;
    if (!get(key_for(device, p, slots, split_only), c)) return false;
    apply_choice(p, c);
    if (turn) *turn = c.turn;
    return true;
}

std::vector<Trial> tune(int device, const pleng::Plan& rule, uint32_t slots, bool split_only, Trial& best)
{
    Bench b(device, rule, std::max(1u, slots), split_only);
    std::vector<Trial> trials;
    Choice base;
    base.logR = log2u(rule.R);
    base.row_radix = rule.row_radix;
    base.col_radix = rule.col_radix;
    base.nt = rule.nt;
    base.turn = !split_only && rule.nt == 8;
    trials.push_back(b.run(base));
    best = trials[0];
    auto consider = [&](const Choice& c) {
        for (const Trial& t : trials)
            if (same(t.c, c)) return;
        trials.push_back(b.run(c));
        const Trial& t = trials.back();
        if (t.exact && t.max_err <= kGuard && t.us < best.us) best = t;
    };
    // shape: every R the kernels allow for this outer pass
    for (uint32_t lr : pleng::shape_options(rule.logL, rule.nt != 0)) {
        Choice c = best.c;
        c.logR = lr;
        consider(c);
    }
    // radices of the steady-state row and column passes
    {
        Choice c = best.c;
        c.row_radix = 16;
        consider(c);
    }
    if (rule.nt) {
        Choice c = best.c;
        c.col_radix = 16;
        consider(c);
    }
    // runs of 16 digits (split pass only) where the rule took 8
    if (rule.nt == 8) {
        Choice c = best.c;
        c.nt = 16;
        c.turn = false;
        consider(c);
    }
    // outer pass in two launches instead of one
    if (!split_only && best.c.nt == 8 && best.c.turn) {
        Choice c = best.c;
        c.turn = false;
        consider(c);
    }
    set(key_for(device, rule, slots, split_only), best.c);
    return trials;
}

static void to_trial(const Trial& t, const Choice& chosen, pl_tune_trial& o)
{
    o.rows = t.R;
    o.cols = t.C;
    o.row_radix = t.c.row_radix;
    o.col_radix = t.c.col_radix;
    o.run = t.c.nt;
    o.one_launch = t.c.turn && t.c.nt == 8 ? 1u : 0u;
    o.us_per_squaring = t.us;
    o.max_err = t.max_err;
    o.exact = t.exact ? 1 : 0;
    o.chosen = same(t.c, chosen) ? 1 : 0;
}

void report(const std::vector<Trial>& ts, const Trial& best, pl_tune_trial* trials, uint32_t max_trials,
            uint32_t* n_trials, pl_tune_trial* out_best)
{
    if (trials)
        for (size_t i = 0; i < ts.size() && i < max_trials; ++i) to_trial(ts[i], best.c, trials[i]);
    if (n_trials) *n_trials = (uint32_t)ts.size();
    if (out_best) to_trial(best, best.c, *out_best);
}

void set(const Key& k, const Choice& c)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[k] = c;
}

bool get(const Key& k, Choice& c)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(k);
    if (it == g_table.end()) return false;
    c = it->second;
    return true;
}

void clear()
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
}

// Text format, one entry a line after the header:
//   device logL nt slots_class split_only : logR row_radix col_radix nt turn
static const char* kHeader = "# prothlogic tuned plans v1";

bool save(const std::string& path, std::string& why)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const std::string tmp = path + ".tmp";
    {
        std::ofstream f(tmp, std::ios::trunc);
        if (!f) {
            why = "cannot open " + tmp;
            return false;
        }
        f << kHeader << "\n";
        for (const auto& [k, c] : g_table)
            f << k.device << ' ' << k.logL << ' ' << k.nt << ' ' << k.slots_class << ' ' << (k.split_only ? 1 : 0)
              << " : " << c.logR << ' ' << c.row_radix << ' ' << c.col_radix << ' ' << c.nt << ' ' << (c.turn ? 1 : 0)
              << "\n";
        f.flush();
        if (!f) {
            why = "write failed: " + tmp;
            return false;
        }
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        why = "rename failed: " + path;
        return false;
    }
    return true;
}

bool load(const std::string& path, std::string& why)
{
    std::ifstream f(path);
    if (!f) {
        why = "cannot open " + path;
        return false;
    }
    std::string line;
    if (!std::getline(f, line) || line != kHeader) {
        why = "not a tuned-plan file: " + path;
        return false;
    }
    std::map<Key, Choice> in;
    int ln = 1;
    while (std::getline(f, line)) {
        ++ln;
        if (line.empty()) continue;
        std::istringstream is(line);
        Key k;
        Choice c;
        int so = 0, turn = 0;
        std::string colon;
        if (!(is >> k.device >> k.logL >> k.nt >> k.slots_class >> so >> colon >> c.logR >> c.row_radix >> c.col_radix >>
              c.nt >> turn) ||
            colon != ":") {
            why = path + ":" + std::to_string(ln) + ": malformed entry";
            return false;
        }
        k.split_only = so != 0;
        c.turn = turn != 0;
        // every value must name a plan the kernels have
        const auto opts = pleng::shape_options(k.logL, c.nt != 0);
        const bool shape_ok = std::find(opts.begin(), opts.end(), c.logR) != opts.end();
        const bool radix_ok = (c.row_radix == 8 || c.row_radix == 16) && (c.col_radix == 8 || c.col_radix == 16);
        const bool nt_ok = c.nt == k.nt || (k.nt == 8 && c.nt == 16);
        if (!shape_ok || !radix_ok || !nt_ok || k.logL < 10 || k.logL > 22) {
            why = path + ":" + std::to_string(ln) + ": entry names no available plan";
            return false;
        }
        in[k] = c;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& [k, c] : in) g_table[k] = c;
    return true;
}

}  // namespace pltune
