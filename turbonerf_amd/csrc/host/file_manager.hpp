// Snapshot files of a trained NeRF (reference src/services/file-manager.cuh:
// 17-198): a 300-byte header, the fp32 parameters (the density network --
// hash table, then its MLP -- followed by the colour network), then the
// occupancy bitfield. Same names, layout and errors as the reference; the
// optimiser state is not stored, as there.
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

    static void save(const NeRFProxy* proxy, const std::string& file_path);
    static void load(NeRFProxy* proxy, const std::string& file_path);
};

}  // namespace turbo
