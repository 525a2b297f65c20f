// Launch-shape knobs: one process-wide table that the kernel launchers read
// on the host when they enqueue work, set by tn_tune_set (an autotuner, or a
// caller that knows its workload). Ranges are those every launcher accepts:
// the hash-forward kernel is bounded to 256 threads per workgroup, the binned
// backward's row reduction to 256 count/scatter workgroups per level, and the
// MLP backward's partial-sum workspace is sized for 2 workgroups per CU.
#include <atomic>

#include "common.hpp"

namespace tn {

namespace {

std::atomic<int> g_tune[TN_TUNE_COUNT] = {256, 128, 3, 1, 2048, 256, 1, 0, 0, 0};
std::atomic<uint32_t> g_generation{0};

bool tune_valid(int knob, int v) {
    switch (knob) {
    case TN_TUNE_HASH_FWD_BLOCK: return v == 64 || v == 128 || v == 256;
    case TN_TUNE_HASH_BWD_LEVEL_BLOCKS: return v >= 8 && v <= 256;
    case TN_TUNE_MLP_FWD_BLOCKS_PER_CU: return v >= 1 && v <= 4;
    case TN_TUNE_MLP_BWD_BLOCKS_PER_CU: return v >= 1 && v <= 2;
    case TN_TUNE_MARCH_MAX_BLOCKS: return v >= 256 && v <= 8192;
    case TN_TUNE_ADAM_BLOCK: return v >= 64 && v <= 1024 && v % 64 == 0;
    case TN_TUNE_HASH_BWD_LEVEL_GROUPS: return v >= 1 && v <= 16;
    case TN_TUNE_HASH_FWD_LDS_LEVELS: return v >= 0 && v <= 4;
    case TN_TUNE_STEP_HALVES: return v == 0 || v == 1;
    case TN_TUNE_HASH_ADAM_FUSED: return v == 0 || v == 1;
    default: return false;
    }
}

}  // namespace

int tune_value(int knob) { return g_tune[knob].load(std::memory_order_relaxed); }

}  // namespace tn

extern "C" int tn_tune_set(int knob, int value) {
    if (!tn::tune_valid(knob, value)) return TN_ERR_INVALID_ARGUMENT;
    if (tn::g_tune[knob].exchange(value, std::memory_order_relaxed) != value)
        tn::g_generation.fetch_add(1, std::memory_order_relaxed);
    return TN_OK;
}

extern "C" int tn_tune_get(int knob, int* value) {
    if (knob < 0 || knob >= TN_TUNE_COUNT || !value) return TN_ERR_INVALID_ARGUMENT;
    *value = tn::tune_value(knob);
    return TN_OK;
}

extern "C" uint32_t tn_tune_generation(void) { return tn::g_generation.load(std::memory_order_relaxed); }
