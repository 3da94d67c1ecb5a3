// Per-device state of the plugin: HIP stream, the scene resident in HBM (rebuilt when the
// host's scene changes) and reusable scratch buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "prender_gpu.h"
#include "wbvh.h"

namespace pgpu {

#define PG_CHECK(x)                                                                                 \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Grows to at least n bytes (contents are not preserved).
    void ensure(size_t n) {
        if (n <= bytes && p) return;
        release();
        PG_CHECK(hipMalloc(&p, n ? n : 16));
        bytes = n ? n : 16;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
    template <class T> void upload(const T *src, size_t count, hipStream_t s) {
        ensure(count * sizeof(T));
        if (count) PG_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

// Launch and batching parameters chosen at run time: these defaults, then the tuning file that
// the autotuner writes (read once, when the plugin first enumerates devices), then
// prgpu_tune_set per device.
struct Tuning {
    int traceBlocksPerCU = 8;                // persistent traversal grid per CU (1..8: LDS-bound)
    int shadeBlocksPerCU = 16;               // grid of the surface, emission and material kernels
    int64_t pathBudget = int64_t(32) << 20;  // paths in flight per batch: sizes every queue (~15 GB)
    int guideBatch = 1 << 18;                // training records per Adam step
    int guideSteps = 4;                      // Adam steps per batch of paths, at most
    int guideMinSteps = 1;                   // ... and at least, cycling over the records
    float guideLr = 1e-3f;                   // Adam learning rate
    int guideBounces = 2;                    // bounces whose diffuse vertices use the guide (0: all)
    float guideBlend = 0.5f;                 // weight of the MIS-proposal term in the training loss
    int guideReplay = 1 << 20;               // training records kept across batches (0: the batch's only)
    int graphRounds = 0;                     // bounce rounds per captured HIP graph (even; 0: none)
    int materialSort = 0;                    // sort the material queues by material and medium each bounce
    int raySort = 0;                         // sort the next bounce's rays by direction octant and origin cell
};

struct PTScene;     // path-tracing data (pt.hip)
struct GuideState;  // learned path-guiding model (guide_host.h)

struct Device {
    int id = 0;
    int cus = 0;
    Tuning tune;
    hipStream_t stream = nullptr;
    // Geometry
    uint64_t geomKey = 0;
    DevBuf nodes, recs;
    int depth = 0;
    int nodeCount = 0, recCount = 0;
    float sceneLo[3] = {0, 0, 0}, sceneHi[3] = {0, 0, 0};
    // Traversal scratch
    DevBuf spill;
    int spillCap = 0, spillStride = 0;
    DevBuf flags;  // [0] traversal stack overflow
    uint32_t *scratch = nullptr;  // this GPU's copy of the allocation-free query's scratch array
    // Path tracing
    uint64_t ptKey = 0;
    std::unique_ptr<PTScene> pt;
    std::unique_ptr<GuideState> guide;
    DevBuf rays, hits;
    // Device-resident FP64 film (prgpu_film_reset / prgpu_film_read)
    DevBuf filmSum, filmSq;
    int filmW = 0, filmH = 0;
    // Profiling (prgpu_profile)
    int profileMode = 0;
    prgpu::ProfileStatsEx prof{};
    DevBuf travCounters;  // 6 x uint64: closest rays/nodes/recs, shadow rays/nodes/recs
    std::vector<hipEvent_t> eventPool;
    hipEvent_t pipeEvents[2] = {nullptr, nullptr};  // read-back events of the pipelined bounce loop
    Device();
    ~Device();
};

}  // namespace pgpu
