// checkpoint_manager.cpp -- reference-layout checkpoints (see
// checkpoint_manager.h).
#include "checkpoint_manager.h"

#include <cstdio>
#include <unistd.h>

bool save_checkpoint(uint64_t i, size_t n_limbs, uint64_t* data, const std::string& filename)
{
    const std::string tmp = filename + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return false;
    const CheckpointHeader h{ i, n_limbs };
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(data, sizeof(uint64_t), n_limbs, f) == n_limbs;
    ok = ok && std::fflush(f) == 0 && fsync(fileno(f)) == 0;
    ok = std::fclose(f) == 0 && ok;
    if (!ok) {
        std::remove(tmp.c_str());
        return false;
    }
    return std::rename(tmp.c_str(), filename.c_str()) == 0;
}

bool load_checkpoint(uint64_t& out_i, size_t n_limbs, uint64_t* data, const std::string& filename)
{
    FILE* f = std::fopen(filename.c_str(), "rb");
    if (!f) return false;
    CheckpointHeader h{};
    bool ok = std::fread(&h, sizeof h, 1, f) == 1 && h.n_limbs == n_limbs &&
              std::fread(data, sizeof(uint64_t), n_limbs, f) == n_limbs;
    std::fclose(f);
    if (ok) out_i = h.current_i;
    return ok;
}
