// FileManager (see file_manager.hpp).
#include "file_manager.hpp"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <vector>

namespace turbo {

void FileManager::save(const NeRFProxy* proxy, const std::string& file_path) {
    if (!proxy || proxy->nerfs.empty()) throw std::runtime_error("FileManager::save: no NeRF");
    const NeRF& nerf = proxy->nerfs[0];
    const NetworkParamsWorkspace& p = nerf.params;
    if (p.n_total == 0 || p.params_fp.size() < p.n_total || nerf.occupancy_grid.bitfield.size() == 0)
        throw std::runtime_error("FileManager::save: the NeRF has no trained state yet");
    std::ofstream file(file_path, std::ios::binary);
    if (!file) throw std::runtime_error("Could not open file for writing");

    NeRFFileData data;
    std::memset(&data, 0, sizeof(data));
    std::strncpy(data.project_name, "TurboNeRF", sizeof(data.project_name));
    data.version = kFileVersion;
    if (proxy->dataset && proxy->dataset->file_path) {
        const std::string& dp = *proxy->dataset->file_path;
        const size_t n = std::min(dp.size(), sizeof(data.dataset_path) - 1);
        std::memcpy(data.dataset_path, dp.data(), n);
    }
    data.n_density_params = p.n_grid + TN_DENSITY_MLP_PARAMS;
    data.n_color_params = TN_COLOR_MLP_PARAMS;
    data.n_appearance_embedding_params = 0;
    data.n_occ_grid_bits = (uint32_t)nerf.occupancy_grid.bitfield.size();
    data.aabb_scale = (uint32_t)proxy->training_bbox.get().size();
    data.n_appearances = 0;

    std::vector<float> params(data.n_params());
    TN_HIP(hipMemcpy(params.data(), p.params_fp.data(), params.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<uint8_t> bits(data.n_occ_grid_bits);
    TN_HIP(hipMemcpy(bits.data(), nerf.occupancy_grid.bitfield.data(), bits.size(), hipMemcpyDeviceToHost));

    file.write(reinterpret_cast<const char*>(&data), sizeof(data));
    file.write(reinterpret_cast<const char*>(params.data()), (std::streamsize)(params.size() * sizeof(float)));
    file.write(reinterpret_cast<const char*>(bits.data()), (std::streamsize)bits.size());
    if (!file) throw std::runtime_error("Could not write file");
}

void FileManager::load(NeRFProxy* proxy, const std::string& file_path) {
    if (!proxy) throw std::runtime_error("FileManager::load: no NeRF");
    std::ifstream file(file_path, std::ios::binary);
    if (!file) throw std::runtime_error("Could not open file for reading");
    NeRFFileData data;
    file.read(reinterpret_cast<char*>(&data), sizeof(data));
    if (!file || std::strncmp(data.project_name, "TurboNeRF", sizeof(data.project_name)) != 0)
        throw std::runtime_error("File is not a TurboNeRF file");
    if (data.version != kFileVersion)
        throw std::runtime_error("File version " + std::to_string(data.version) +
                                 " is not supported. Supported version is " + std::to_string(kFileVersion));
    if (data.n_appearance_embedding_params != 0 || data.n_appearances != 0)
        throw std::runtime_error("FileManager::load: appearance embeddings are not part of this build");
    if (data.n_color_params != TN_COLOR_MLP_PARAMS)
        throw std::runtime_error("FileManager::load: a colour network of " + std::to_string(data.n_color_params) +
                                 " parameters, this build's has " + std::to_string(TN_COLOR_MLP_PARAMS));

    std::vector<float> params(data.n_params());
    file.read(reinterpret_cast<char*>(params.data()), (std::streamsize)(params.size() * sizeof(float)));
    std::vector<uint8_t> bits(data.n_occ_grid_bits);
    file.read(reinterpret_cast<char*>(bits.data()), (std::streamsize)bits.size());
    if (!file) throw std::runtime_error("FileManager::load: the file is truncated");

    if (proxy->nerfs.empty()) proxy->nerfs.resize(1);
    NeRF& nerf = proxy->nerfs[0];
    NerfNetwork& net = nerf.network;
    const int aabb = std::max(1, (int)data.aabb_scale);
    // the hash-table size is not stored: take the layout whose parameter count matches
    const int configured = net.config.log2_hashmap_size;
    bool found = false;
    for (int k = -1; k <= 24 && !found; ++k) {
        const int log2_T = k < 0 ? configured : k;
        if (k >= 0 && (k < 8 || k == configured)) continue;
        net.config.log2_hashmap_size = log2_T;
        net.aabb_scale = 0;  // rebuild the layout
        net.update_aabb_scale_if_needed(aabb);
        found = net.grid.n_params + (uint32_t)TN_DENSITY_MLP_PARAMS == data.n_density_params;
    }
    if (!found) {
        net.config.log2_hashmap_size = configured;
        net.aabb_scale = 0;
        net.update_aabb_scale_if_needed(aabb);
        throw std::runtime_error("FileManager::load: a density network of " + std::to_string(data.n_density_params) +
                                 " parameters matches no hash-grid layout of this build");
    }

    const hipStream_t s = DeviceManager::stream();
    NetworkParamsWorkspace& p = nerf.params;
    net.update_params_if_needed(s, p);
    TN_HIP(hipMemcpyAsync(p.params_fp.data(), params.data(), params.size() * sizeof(float), hipMemcpyHostToDevice, s));
    if (net.config.precision == 1) TN_TRY(tn_cast_params(p.n_total, p.params_fp.data(), 1, p.params_lp.data(), as_tn(s)));
    // a fresh optimiser state
    p.gradients.zero(s);
    p.m1.zero(s);
    p.m2.zero(s);
    p.steps.zero(s);

    OccupancyGrid& og = nerf.occupancy_grid;
    og.set_aabb_scale((float)aabb);
    og.initialize_if_needed(s);
    if (bits.size() != og.bitfield.size())
        throw std::runtime_error("FileManager::load: an occupancy bitfield of " + std::to_string(bits.size()) +
                                 " bytes, this build's has " + std::to_string(og.bitfield.size()));
    TN_HIP(hipMemcpyAsync(og.bitfield.data(), bits.data(), bits.size(), hipMemcpyHostToDevice, s));
    TN_HIP(hipStreamSynchronize(s));

    const BoundingBox bbox((float)aabb);
    proxy->training_bbox.set(bbox);
    proxy->render_bbox.set(bbox);
    proxy->is_valid = true;
    proxy->can_render = true;
}

}  // namespace turbo
