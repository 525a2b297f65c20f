// Snapshot files of a trained NeRF (reference src/services/file-manager.cuh:
// 17-198): a 300-byte header, the fp32 parameters (the density network --
// hash table, then its MLP -- followed by the colour network and, with an
// appearance embedding, the embeddings [n_appearances][16]), then the
// occupancy bitfield. Same names, layout and errors as the reference. The
// colour network is stored as tiny-cuda-nn's fully fused MLP counts it: its
// first layer [64][32 + 16 embedding columns], and its 3-row output layer
// padded with zero rows to 16 (n_color_params = 7168, or 8192 with an
// embedding), so the header's counts match a file of the reference's.
//
// Optionally (save with include_optimizer_state) an optimiser section follows
// the bitfield, for an exact resume of training (the reference serialises the
// same state through NGPAdamOptimizer::serialize, src/core/adam-optimizer.cuh:
// 203-224, but does not write it into these files): the 8 bytes "TNADAM01",
// uint32 training step, uint32 n (parameters, in this build's order: hash
// table, density MLP, colour MLP unpadded), float m1[n], float m2[n],
// uint32 steps[n]. A reader that stops after the bitfield (the reference)
// ignores it; a file without it loads with a fresh optimiser state.
#pragma once

#include <cstdint>
#include <string>

#include "controllers.hpp"

namespace turbo {

struct FileManager {
    static constexpr uint32_t kFileVersion = 0;

    struct NeRFFileData {
        char project_name[16];
        char dataset_path[256];
        uint32_t version;
        uint32_t n_color_params;
        uint32_t n_density_params;
        uint32_t n_appearance_embedding_params;
        uint32_t n_occ_grid_bits;
        uint32_t aabb_scale;
        uint32_t n_appearances;
        size_t n_params() const {
            return (size_t)n_color_params + n_density_params + n_appearance_embedding_params;
        }
    };
    static_assert(sizeof(NeRFFileData) == 300, "the snapshot header is 300 bytes");

    // colour-network parameters as stored: W3 [64][32], W4 [64][64], W5 [16][64]
    static constexpr uint32_t kStoredColorParams = 64 * 32 + 64 * 64 + 16 * 64;
    static constexpr char kOptimizerTag[8] = {'T', 'N', 'A', 'D', 'A', 'M', '0', '1'};

    static void save(const NeRFProxy* proxy, const std::string& file_path, bool include_optimizer_state = false);
    // Returns true when the file held optimiser state (then restored, with the
    // training step).
    static bool load(NeRFProxy* proxy, const std::string& file_path);
};

}  // namespace turbo
