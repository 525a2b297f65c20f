// FileManager (see file_manager.hpp).
#include "file_manager.hpp"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <vector>

namespace turbo {

namespace {

constexpr uint32_t kW5Rows = 3, kW5PaddedRows = 16, kHidden = 64, kColorIn = 32, kEmb = TN_APPEARANCE_DIMS;
constexpr uint32_t kW3 = kHidden * kColorIn, kW4 = kHidden * kHidden;

// colour parameters of this build (W3 [64][32], W4, W5 [3][64], and with an
// embedding W3e [64][16] kept apart) -> tiny-cuda-nn's form: W3 [64][32 + 16]
// with the embedding columns last, W4, W5 padded to 16 rows
uint32_t stored_color_params(bool emb, bool padded) {
    return kHidden * (kColorIn + (emb ? kEmb : 0)) + kW4 + (padded ? kW5PaddedRows : kW5Rows) * kHidden;
}
void pack_color(const float* in, const float* w3e, float* out) {
    const uint32_t win = kColorIn + (w3e ? kEmb : 0);
    for (uint32_t r = 0; r < kHidden; ++r) {
        std::memcpy(out + r * win, in + r * kColorIn, kColorIn * sizeof(float));
        if (w3e) std::memcpy(out + r * win + kColorIn, w3e + r * kEmb, kEmb * sizeof(float));
    }
    out += kHidden * win;
    std::memcpy(out, in + kW3, (kW4 + kW5Rows * kHidden) * sizeof(float));
    std::fill(out + kW4 + kW5Rows * kHidden, out + kW4 + kW5PaddedRows * kHidden, 0.0f);
}
// the stored form (padded or not) -> this build's
void unpack_color(const float* in, bool emb, float* out, float* w3e) {
    const uint32_t win = kColorIn + (emb ? kEmb : 0);
    for (uint32_t r = 0; r < kHidden; ++r) {
        std::memcpy(out + r * kColorIn, in + r * win, kColorIn * sizeof(float));
        if (emb) std::memcpy(w3e + r * kEmb, in + r * win + kColorIn, kEmb * sizeof(float));
    }
    // W4 and the first 3 rows of W5 (a padded form's extra rows follow them)
    std::memcpy(out + kW3, in + kHidden * win, (kW4 + kW5Rows * kHidden) * sizeof(float));
}

template <typename T>
void write_vec(std::ofstream& f, const std::vector<T>& v) {
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

template <typename T>
void read_vec(std::ifstream& f, std::vector<T>& v) {
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

}  // namespace

void FileManager::save(const NeRFProxy* proxy, const std::string& file_path, bool include_optimizer_state) {
    if (!proxy || proxy->nerfs.empty()) throw std::runtime_error("FileManager::save: no NeRF");
    const NeRF& nerf = proxy->nerfs[0];
    const NetworkParamsWorkspace& p = nerf.params;
    if (p.n_total == 0 || p.params_fp.size() < p.n_total || nerf.occupancy_grid.bitfield.size() == 0)
        throw std::runtime_error("FileManager::save: the NeRF has no trained state yet");
    if (include_optimizer_state && (p.m1.size() < p.n_total || p.m2.size() < p.n_total || p.steps.size() < p.n_total))
        throw std::runtime_error("FileManager::save: the NeRF has no optimiser state");
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
    const uint32_t n_density = p.n_grid + TN_DENSITY_MLP_PARAMS;
    const uint32_t n_app = p.n_appearances;
    data.n_density_params = n_density;
    data.n_color_params = stored_color_params(n_app > 0, true);
    data.n_appearance_embedding_params = n_app * kEmb;
    data.n_occ_grid_bits = (uint32_t)nerf.occupancy_grid.bitfield.size();
    data.aabb_scale = (uint32_t)proxy->training_bbox.get().size();
    data.n_appearances = n_app;

    std::vector<float> params(p.n_total);
    TN_HIP(hipMemcpy(params.data(), p.params_fp.data(), params.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<float> stored(data.n_params());
    std::memcpy(stored.data(), params.data(), n_density * sizeof(float));
    pack_color(params.data() + n_density, n_app ? params.data() + p.w3e_offset() : nullptr,
               stored.data() + n_density);
    // the embeddings follow the colour network, as the reference stores them
    if (n_app)
        std::memcpy(stored.data() + n_density + data.n_color_params, params.data() + p.emb_offset(),
                    (size_t)n_app * kEmb * sizeof(float));
    std::vector<uint8_t> bits(data.n_occ_grid_bits);
    TN_HIP(hipMemcpy(bits.data(), nerf.occupancy_grid.bitfield.data(), bits.size(), hipMemcpyDeviceToHost));

    file.write(reinterpret_cast<const char*>(&data), sizeof(data));
    write_vec(file, stored);
    write_vec(file, bits);
    if (include_optimizer_state) {
        std::vector<float> m1(p.n_total), m2(p.n_total);
        std::vector<uint32_t> steps(p.n_total);
        TN_HIP(hipMemcpy(m1.data(), p.m1.data(), m1.size() * sizeof(float), hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(m2.data(), p.m2.data(), m2.size() * sizeof(float), hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(steps.data(), p.steps.data(), steps.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        const uint32_t head[2] = {proxy->training_step, p.n_total};
        file.write(kOptimizerTag, sizeof(kOptimizerTag));
        file.write(reinterpret_cast<const char*>(head), sizeof(head));
        write_vec(file, m1);
        write_vec(file, m2);
        write_vec(file, steps);
    }
    if (!file) throw std::runtime_error("Could not write file");
}

bool FileManager::load(NeRFProxy* proxy, const std::string& file_path) {
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
    const uint32_t n_app = data.n_appearances;
    if (data.n_appearance_embedding_params != n_app * kEmb)
        throw std::runtime_error("FileManager::load: " + std::to_string(data.n_appearance_embedding_params) +
                                 " embedding parameters for " + std::to_string(n_app) +
                                 " appearances; this build's embeddings have " + std::to_string(kEmb) + " each");
    // the padded form tiny-cuda-nn counts, or the unpadded one
    const bool emb = n_app > 0;
    if (data.n_color_params != stored_color_params(emb, true) && data.n_color_params != stored_color_params(emb, false))
        throw std::runtime_error("FileManager::load: a colour network of " + std::to_string(data.n_color_params) +
                                 " parameters, this build's has " + std::to_string(stored_color_params(emb, true)));
    // every cascade shares one byte per cell of the 128^3 grid
    const uint32_t grid_bytes = 128u * 128u * 128u;
    if (data.n_occ_grid_bits != grid_bytes)
        throw std::runtime_error("FileManager::load: an occupancy bitfield of " +
                                 std::to_string(data.n_occ_grid_bits) + " bytes, this build's has " +
                                 std::to_string(grid_bytes));
    const int aabb = std::max(1, (int)data.aabb_scale);

    std::vector<float> stored(data.n_params());
    read_vec(file, stored);
    std::vector<uint8_t> bits(data.n_occ_grid_bits);
    read_vec(file, bits);
    if (!file) throw std::runtime_error("FileManager::load: the file is truncated");

    // the hash-table size is not stored: find the layout whose parameter
    // count matches, before any state of the NeRF changes
    if (proxy->nerfs.empty()) proxy->nerfs.resize(1);
    NeRF& nerf = proxy->nerfs[0];
    NerfNetwork& net = nerf.network;
    int log2_T = -1;
    for (int k = -1; k <= 24 && log2_T < 0; ++k) {
        const int cand = k < 0 ? net.config.log2_hashmap_size : k;
        if (k >= 0 && (k < 8 || k == net.config.log2_hashmap_size)) continue;
        NerfNetwork probe;
        probe.config = net.config;
        probe.config.log2_hashmap_size = cand;
        probe.update_aabb_scale_if_needed(aabb);
        if (probe.grid.n_params + (uint32_t)TN_DENSITY_MLP_PARAMS == data.n_density_params) log2_T = cand;
    }
    if (log2_T < 0)
        throw std::runtime_error("FileManager::load: a density network of " + std::to_string(data.n_density_params) +
                                 " parameters matches no hash-grid layout of this build");
    const uint32_t n_total =
        data.n_density_params + TN_COLOR_MLP_PARAMS + (emb ? TN_APPEARANCE_WEIGHTS + n_app * kEmb : 0);

    // the optional optimiser section, read in full before anything changes
    bool has_opt = false;
    uint32_t opt_step = 0;
    std::vector<float> m1, m2;
    std::vector<uint32_t> steps;
    char tag[sizeof(kOptimizerTag)];
    file.read(tag, sizeof(tag));
    if (file && std::memcmp(tag, kOptimizerTag, sizeof(tag)) == 0) {
        uint32_t head[2];
        file.read(reinterpret_cast<char*>(head), sizeof(head));
        if (!file || head[1] != n_total)
            throw std::runtime_error("FileManager::load: the optimiser section does not match the parameters");
        opt_step = head[0];
        m1.resize(n_total);
        m2.resize(n_total);
        steps.resize(n_total);
        read_vec(file, m1);
        read_vec(file, m2);
        read_vec(file, steps);
        if (!file) throw std::runtime_error("FileManager::load: the file is truncated");
        has_opt = true;
    }

    std::vector<float> params(n_total);
    std::memcpy(params.data(), stored.data(), data.n_density_params * sizeof(float));
    const uint32_t w3e_at = data.n_density_params + TN_COLOR_MLP_PARAMS;
    unpack_color(stored.data() + data.n_density_params, emb, params.data() + data.n_density_params,
                 params.data() + w3e_at);
    if (emb)
        std::memcpy(params.data() + w3e_at + TN_APPEARANCE_WEIGHTS,
                    stored.data() + data.n_density_params + data.n_color_params, (size_t)n_app * kEmb * sizeof(float));

    net.config.log2_hashmap_size = log2_T;
    net.config.appearance_dims = emb ? (int)kEmb : 0;
    net.n_appearances = n_app;
    net.aabb_scale = 0;  // rebuild the layout
    net.update_aabb_scale_if_needed(aabb);
    const hipStream_t s = DeviceManager::stream();
    NetworkParamsWorkspace& p = nerf.params;
    net.update_params_if_needed(s, p);
    TN_HIP(hipMemcpyAsync(p.params_fp.data(), params.data(), params.size() * sizeof(float), hipMemcpyHostToDevice, s));
    if (net.config.precision == 1) TN_TRY(tn_cast_params(p.n_total, p.params_fp.data(), 1, p.params_lp.data(), as_tn(s)));
    p.gradients.zero(s);
    if (has_opt) {
        TN_HIP(hipMemcpyAsync(p.m1.data(), m1.data(), m1.size() * sizeof(float), hipMemcpyHostToDevice, s));
        TN_HIP(hipMemcpyAsync(p.m2.data(), m2.data(), m2.size() * sizeof(float), hipMemcpyHostToDevice, s));
        TN_HIP(hipMemcpyAsync(p.steps.data(), steps.data(), steps.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    } else {
        // a fresh optimiser state, as the reference loads
        p.m1.zero(s);
        p.m2.zero(s);
        p.steps.zero(s);
    }

    OccupancyGrid& og = nerf.occupancy_grid;
    og.set_aabb_scale((float)aabb);
    og.initialize_if_needed(s);
    TN_HIP(hipMemcpyAsync(og.bitfield.data(), bits.data(), bits.size(), hipMemcpyHostToDevice, s));
    TN_HIP(hipStreamSynchronize(s));

    const BoundingBox bbox((float)aabb);
    proxy->training_bbox.set(bbox);
    proxy->render_bbox.set(bbox);
    proxy->is_valid = true;
    proxy->can_render = true;
    if (has_opt) proxy->training_step = opt_step;
    return has_opt;
}

}  // namespace turbo
