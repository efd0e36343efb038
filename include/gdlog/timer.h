// gdlog/timer.h -- KernelTimer of the reference (include/timer.cuh:5-30),
// measuring on the engine context's stream with hipEvents; seconds.
#pragma once

#include <hip/hip_runtime.h>

#include "gdlog/relation.h"

struct KernelTimer {
    hipEvent_t start;
    hipEvent_t stop;

    KernelTimer() {
        (void)hipEventCreate(&start);
        (void)hipEventCreate(&stop);
    }
    ~KernelTimer() {
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(stop);
    }
    void start_timer() { (void)hipEventRecord(start, (hipStream_t)gdl_stream(gdlog::context())); }
    void stop_timer() { (void)hipEventRecord(stop, (hipStream_t)gdl_stream(gdlog::context())); }
    float get_spent_time() {
        float elapsed = 0.f;
        (void)hipEventSynchronize(stop);
        (void)hipEventElapsedTime(&elapsed, start, stop);
        return elapsed / 1000.0f;
    }
};
