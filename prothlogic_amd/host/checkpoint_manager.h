// checkpoint_manager.h -- the checkpoint API of P/checkpoint_manager.h:10-16.
//
// save_checkpoint / load_checkpoint keep the reference's file layout
// ({u64 current_i; size_t n_limbs} + n_limbs LE u64 limbs), so files move
// between the two builds.  The device test driver (pl_proth_test) writes the
// self-checking format v2 of include/prothcore.h instead and reads both.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

struct CheckpointHeader {
    uint64_t current_i;
    size_t n_limbs;
};

// Writes filename + ".tmp", flushes it to disk, renames it over filename.
bool save_checkpoint(uint64_t i, size_t n_limbs, uint64_t* data, const std::string& filename);
// False when the file is missing, short, or holds a different n_limbs.
bool load_checkpoint(uint64_t& out_i, size_t n_limbs, uint64_t* data, const std::string& filename);
