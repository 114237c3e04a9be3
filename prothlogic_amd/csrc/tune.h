// tune.h -- per-length launch plans measured on the device.
//
// For one transform length the engine can run several shapes L = R*C, row
// and column radices 8 or 16, runs of 8 or 16 digits in the split outer
// pass, and the outer pass as one launch or two.  choose_plan's rule picks
// one of them from a sweep; the tuner times the alternatives on the device
// the plan will run on, with the slot count it will run with, and keeps the
// fastest whose result equals the rule plan's.  Every later context or pool
// group with the same key then takes the measured choice.
//
// The table is process-wide and thread-safe; it can be written to and read
// from a text file so that one measurement serves later processes on the
// same machine.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "engine.h"
#include "prothcore.h"

namespace pltune {

// What a tuned entry overrides.
struct Choice {
    uint32_t logR = 0;
    uint32_t row_radix = 8, col_radix = 8;
    uint32_t nt = 0;   // runs of the split outer pass (0: unfused)
    bool turn = true;  // one-launch outer pass where nt == 8
};

// Which plans an entry applies to: device, length, the carry layout the
// rule chose (nt), slot-count class and whether only the split pass may run
// (pool groups).
struct Key {
    int device = 0;
    uint32_t logL = 0;
    uint32_t nt = 0;
    uint32_t slots_class = 0;
    bool split_only = false;
    bool operator<(const Key& o) const;
};

uint32_t slots_class(uint32_t slots);  // 1, 2-8, 9-64, more: 0..3
Key key_for(int device, const pleng::Plan& rule_plan, uint32_t slots, bool split_only);

struct Trial {
    Choice c;
    uint32_t R = 0, C = 0;
    double us = 0;       // microseconds per squaring of every slot
    double max_err = 0;  // round-off of the timed squarings
    bool exact = false;  // result equal to the rule plan's
};

// Applies the table's entry for this plan, if any: shape, radices, nt.
// Returns true when an entry was found; *turn gets its outer-pass form.
bool apply(int device, pleng::Plan& p, uint32_t slots, bool split_only, bool* turn);
// Applies a choice to a plan (no table involved).
void apply_choice(pleng::Plan& p, const Choice& c);

// Times the alternatives for rule_plan (choose_plan + choose_layout's
// result) with `slots` slots and records the fastest exact one.
// Coordinate search: shape, then row radix, then column radix, then runs and
// outer-pass form.  Throws on HIP errors.
std::vector<Trial> tune(int device, const pleng::Plan& rule_plan, uint32_t slots, bool split_only, Trial& best);

// Copies trials into the C ABI's records (at most max_trials; trials and
// out_best nullable) and marks the one equal to best.
void report(const std::vector<Trial>& ts, const Trial& best, pl_tune_trial* trials, uint32_t max_trials,
            uint32_t* n_trials, pl_tune_trial* out_best);

void set(const Key& k, const Choice& c);
bool get(const Key& k, Choice& c);
void clear();
bool save(const std::string& path, std::string& why);
bool load(const std::string& path, std::string& why);

}  // namespace pltune
