// Autotuner (see autotune.hpp).
#include "autotune.hpp"

#include <algorithm>
#include <chrono>
#include <stdexcept>

#include "common.hpp"

namespace turbo {

namespace {

// order of the coordinate search: the knobs of the longest kernels first.
// The hash-grid backward's workgroups per level are not searched: the
// trainer forms each batch's backward plan ahead with the batch, at a fixed
// 256 per level (the knob shapes the one-call tn_hashgrid_backward_binned).
constexpr int kOrder[] = {TN_TUNE_STEP_HALVES,           TN_TUNE_HASH_BWD_LEVEL_GROUPS,
                          TN_TUNE_MLP_BWD_BLOCKS_PER_CU, TN_TUNE_MLP_FWD_BLOCKS_PER_CU,
                          TN_TUNE_HASH_FWD_BLOCK,        TN_TUNE_MARCH_MAX_BLOCKS,
                          TN_TUNE_ADAM_BLOCK,            TN_TUNE_HASH_FWD_LDS_LEVELS,
                          TN_TUNE_HASH_ADAM_FUSED};

int get(int knob) {
    int v = 0;
    TN_TRY(tn_tune_get(knob, &v));
    return v;
}

}  // namespace

const char* Autotuner::knob_name(int knob) {
    switch (knob) {
    case TN_TUNE_HASH_FWD_BLOCK: return "hash_fwd_block";
    case TN_TUNE_HASH_BWD_LEVEL_BLOCKS: return "hash_bwd_level_blocks";
    case TN_TUNE_MLP_FWD_BLOCKS_PER_CU: return "mlp_fwd_blocks_per_cu";
    case TN_TUNE_MLP_BWD_BLOCKS_PER_CU: return "mlp_bwd_blocks_per_cu";
    case TN_TUNE_MARCH_MAX_BLOCKS: return "march_max_blocks";
    case TN_TUNE_ADAM_BLOCK: return "adam_block";
    case TN_TUNE_HASH_BWD_LEVEL_GROUPS: return "hash_bwd_level_groups";
    case TN_TUNE_HASH_FWD_LDS_LEVELS: return "hash_fwd_lds_levels";
    case TN_TUNE_STEP_HALVES: return "step_halves";
    case TN_TUNE_HASH_ADAM_FUSED: return "hash_adam_fused";
    default: throw std::runtime_error("autotune: unknown knob");
    }
}

std::vector<int> Autotuner::candidates(int knob) {
    switch (knob) {
    case TN_TUNE_HASH_FWD_BLOCK: return {64, 128, 256};
    case TN_TUNE_HASH_BWD_LEVEL_BLOCKS: return {64, 128, 256};
    case TN_TUNE_MLP_FWD_BLOCKS_PER_CU: return {1, 2, 3, 4};
    case TN_TUNE_MLP_BWD_BLOCKS_PER_CU: return {1, 2};
    case TN_TUNE_MARCH_MAX_BLOCKS: return {512, 1024, 2048, 4096};
    case TN_TUNE_ADAM_BLOCK: return {128, 256, 512, 1024};
    case TN_TUNE_HASH_BWD_LEVEL_GROUPS: return {1, 2, 4, 8};
    case TN_TUNE_HASH_FWD_LDS_LEVELS: return {0, 1, 2};
    case TN_TUNE_STEP_HALVES: return {0, 1};
    case TN_TUNE_HASH_ADAM_FUSED: return {0, 1};
    default: throw std::runtime_error("autotune: unknown knob");
    }
}

std::map<std::string, int> Autotuner::current() {
    std::map<std::string, int> m;
    for (int k = 0; k < TN_TUNE_COUNT; ++k) m[knob_name(k)] = get(k);
    return m;
}

void Autotuner::set(const std::string& name, int value) {
    for (int k = 0; k < TN_TUNE_COUNT; ++k)
        if (name == knob_name(k)) {
            if (tn_tune_set(k, value) != TN_OK)
                throw std::runtime_error("set_tuning: " + name + " = " + std::to_string(value) + " is out of range");
            return;
        }
    throw std::runtime_error("set_tuning: no knob named " + name);
}

AutotuneResult Autotuner::tune(const std::function<void(uint32_t)>& run, const std::function<uint64_t()>& samples,
                               uint32_t rounds, uint32_t repetitions) {
    if (rounds == 0 || repetitions == 0) throw std::runtime_error("autotune: rounds and repetitions must be > 0");
    using clock = std::chrono::steady_clock;
    auto trial = [&]() {
        run(2);  // captures the graphs of a new launch shape, untimed
        const uint64_t s0 = samples();
        const auto t0 = clock::now();
        run(rounds);
        const uint64_t s1 = samples();
        const double ns = std::chrono::duration<double, std::nano>(clock::now() - t0).count();
        return s1 > s0 ? ns / (double)(s1 - s0) : 1e30;
    };
    auto median = [](std::vector<double> v) {
        std::sort(v.begin(), v.end());
        return v[v.size() / 2];
    };
    AutotuneResult res;
    {
        std::vector<double> t;
        for (uint32_t r = 0; r < repetitions; ++r) t.push_back(trial());
        res.ns_per_sample_before = median(t);
    }
    for (const int knob : kOrder) {
        const std::vector<int> cand = candidates(knob);
        std::vector<std::vector<double>> t(cand.size());
        for (uint32_t r = 0; r < repetitions; ++r)
            for (size_t c = 0; c < cand.size(); ++c) {
                TN_TRY(tn_tune_set(knob, cand[c]));
                t[c].push_back(trial());
            }
        size_t best = 0;
        std::vector<std::pair<int, double>> rows;
        for (size_t c = 0; c < cand.size(); ++c) {
            const double m = median(t[c]);
            rows.emplace_back(cand[c], m);
            if (m < rows[best].second) best = c;
        }
        TN_TRY(tn_tune_set(knob, cand[best]));
        res.trials[knob_name(knob)] = rows;
    }
    res.picks = current();
    {
        std::vector<double> t;
        for (uint32_t r = 0; r < repetitions; ++r) t.push_back(trial());
        res.ns_per_sample_after = median(t);
    }
    return res;
}

}  // namespace turbo
