// HIP-event timing of solver stages on the stream each stage runs on.
#pragma once

#include "hip_common.hpp"

#include <utility>
#include <vector>

namespace cupf {

// Pairs of HIP events recorded around each stage on the stream it runs on.
struct StageTimer {
    static constexpr int kStages = 7;
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    std::vector<int> stage_of;
    size_t used = 0;
    ~StageTimer() {
        for (auto& e : pool) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }
    int begin(int stage, hipStream_t s) {
        if (!on || stage < 0) return -1;
        if (used == pool.size()) {
            std::pair<hipEvent_t, hipEvent_t> e;
            CUPF_HIP(hipEventCreate(&e.first));
            CUPF_HIP(hipEventCreate(&e.second));
            pool.push_back(e);
            stage_of.push_back(stage);
        }
        stage_of[used] = stage;
        CUPF_HIP(hipEventRecord(pool[used].first, s));
        return int(used++);
    }
    void end(int idx, hipStream_t s) {
        if (idx >= 0) CUPF_HIP(hipEventRecord(pool[idx].second, s));
    }
    void collect(double* ms, int64_t* count) {
        for (int k = 0; k < kStages; ++k) { ms[k] = 0.0; count[k] = 0; }
        for (size_t i = 0; i < used; ++i) {
            CUPF_HIP(hipEventSynchronize(pool[i].second));
            float t = 0.f;
            CUPF_HIP(hipEventElapsedTime(&t, pool[i].first, pool[i].second));
            ms[stage_of[i]] += t;
            count[stage_of[i]] += 1;
        }
        used = 0;
    }
};

struct StageScope {
    StageTimer* t;
    hipStream_t s;
    int idx;
    StageScope(StageTimer* timer, int stage, hipStream_t stream)
        : t(timer), s(stream), idx(timer ? timer->begin(stage, stream) : -1) {}
    StageScope(StageTimer& timer, int stage, hipStream_t stream)
        : StageScope(&timer, stage, stream) {}
    ~StageScope() {
        if (t) t->end(idx, s);
    }
};

}  // namespace cupf
