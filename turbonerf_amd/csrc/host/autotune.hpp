// Launch-shape autotuner. The kernels' launch shapes (workgroup sizes, grid
// caps, workgroups per CU; enum tn_tune_knob) are process-wide knobs; the
// tuner picks each knob's value by timing real training rounds of the
// caller's workload (one coordinate at a time, candidates interleaved over
// the repetitions so drift affects them alike) and keeps the fastest. The
// rounds it times are ordinary training steps: the NeRFs go on training.
// No reference counterpart (the reference's launch shapes are compile-time
// constants).
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace turbo {

struct AutotuneResult {
    std::map<std::string, int> picks;
    // per knob: (value, median ns per trained sample) of each candidate
    std::map<std::string, std::vector<std::pair<int, double>>> trials;
    double ns_per_sample_before = 0.0, ns_per_sample_after = 0.0;
};

struct Autotuner {
    // run(n): enqueue n training rounds; samples(): samples trained so far
    // (synchronises). Each trial runs 2 untimed rounds (graph capture) and
    // then `rounds` timed ones.
    static AutotuneResult tune(const std::function<void(uint32_t)>& run, const std::function<uint64_t()>& samples,
                               uint32_t rounds, uint32_t repetitions);
    static const char* knob_name(int knob);
    static std::vector<int> candidates(int knob);
    static std::map<std::string, int> current();
    static void set(const std::string& name, int value);
};

}  // namespace turbo
