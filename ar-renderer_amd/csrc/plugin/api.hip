// C-ABI entry points of the MI355X plugin (include/prender_gpu.h), device management and the
// ray-batch query kernel. Every entry point takes the one global lock, never throws across the
// ABI and reports failures as a nonzero return with a message (reference
// src/gpu/prender_gpu.cpp:1576-1730 for the contract).
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "device.h"
#include "guide_host.h"
#include "pt_host.h"
#include "sheen_table.h"
#include "trace_dev.h"

namespace pgpu {

namespace {

// Most resident traversal blocks per CU (the persistent grid; the spill slabs are sized for it): LDS holds kStackLds * 256 * 4 B per
// block, 160 KiB per CU.
#ifndef PG_TRACE_BLOCKS_PER_CU
#define PG_TRACE_BLOCKS_PER_CU 8
#endif
constexpr int kTraceBlocksPerCU = PG_TRACE_BLOCKS_PER_CU;
static_assert(kTraceBlocksPerCU * kStackLds * kTraceBlock * 4 <= 160 * 1024, "traversal blocks exceed the LDS of a CU");

std::mutex gMutex;
std::vector<std::unique_ptr<Device>> gDevices;
bool gEnumerated = false;

// Scratch of the allocation-free ray query (prgpu_trace_rays_noalloc), one copy per GPU, held
// until the process ends. Words [0, kScratchHeadWords) hold the traversal stack-overflow flag
// (word 0); the rest holds the per-lane spill slabs of the traversal stacks. Host code never
// passes `scratch` itself to a kernel: it uses the device address that hipGetSymbolAddress
// returns for the current device (Device::scratch).
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchWords = kScratchBytes / sizeof(uint32_t);
constexpr size_t kScratchHeadWords = 64;  // 256 B, so the spill slabs start 256-B aligned
static __device__ __align__(256) uint32_t scratch[kScratchWords];

void SetErr(char *err, int len, const std::string &m) {
    if (err && len > 0) std::snprintf(err, size_t(len), "%s", m.c_str());
}

// ---- run-time tuning ----------------------------------------------------------------------------
// Keys of the tuning file and of prgpu_tune_set / prgpu_tune_get, with their admissible ranges.
struct TuneKey {
    const char *name;
    double lo, hi;
    bool integer;
};
const TuneKey kTuneKeys[] = {
    {"trace_blocks_per_cu", 1, kTraceBlocksPerCU, true}, {"shade_blocks_per_cu", 1, 64, true},
    {"path_budget", 1 << 16, 1 << 26, true},           {"guide_batch", 256, 1 << 20, true},
    {"guide_steps", 1, 64, true},                      {"guide_min_steps", 0, 64, true},
    {"guide_lr", 1e-6, 1.0, false},                    {"guide_bounces", 0, 64, true},
    {"guide_blend", 0.0, 1.0, false},                  {"graph_rounds", 0, 16, true},
    {"guide_replay", 0, 1 << 24, true},                {"material_sort", 0, 1, true},
    {"ray_sort", 0, 1, true},
};

bool TuneSet(Tuning &t, const std::string &key, double v) {
    for (const TuneKey &k : kTuneKeys) {
        if (key != k.name) continue;
        if (!(v >= k.lo && v <= k.hi) || (k.integer && v != std::floor(v))) return false;
        if (key == "trace_blocks_per_cu") t.traceBlocksPerCU = int(v);
        else if (key == "shade_blocks_per_cu") t.shadeBlocksPerCU = int(v);
        else if (key == "path_budget") t.pathBudget = int64_t(v);
        else if (key == "guide_batch") t.guideBatch = int(v);
        else if (key == "guide_steps") t.guideSteps = int(v);
        else if (key == "guide_min_steps") t.guideMinSteps = int(v);
        else if (key == "guide_lr") t.guideLr = float(v);
        else if (key == "guide_bounces") t.guideBounces = int(v);
        else if (key == "guide_blend") t.guideBlend = float(v);
        else if (key == "guide_replay") t.guideReplay = int(v);
        else if (key == "material_sort") t.materialSort = int(v);
        else if (key == "ray_sort") t.raySort = int(v);
        else if (key == "graph_rounds") {
            if (int(v) % 2) return false;  // a replay must end on the queue it started from
            t.graphRounds = int(v);
        }
        return true;
    }
    return false;
}

bool TuneGet(const Tuning &t, const std::string &key, double *v) {
    if (key == "trace_blocks_per_cu") *v = t.traceBlocksPerCU;
    else if (key == "shade_blocks_per_cu") *v = t.shadeBlocksPerCU;
    else if (key == "path_budget") *v = double(t.pathBudget);
    else if (key == "guide_batch") *v = t.guideBatch;
    else if (key == "guide_steps") *v = t.guideSteps;
    else if (key == "guide_min_steps") *v = t.guideMinSteps;
    else if (key == "guide_lr") *v = t.guideLr;
    else if (key == "guide_bounces") *v = t.guideBounces;
    else if (key == "guide_blend") *v = t.guideBlend;
    else if (key == "graph_rounds") *v = t.graphRounds;
    else if (key == "guide_replay") *v = t.guideReplay;
    else if (key == "material_sort") *v = t.materialSort;
    else if (key == "ray_sort") *v = t.raySort;
    else return false;
    return true;
}

Tuning gTuning;           // defaults for every device, after the tuning file
std::string gTuningFile;  // the file read ("" when none)
std::string gTuningError; // the first problem found in it ("" when none)

// The tuning file: $PRGPU_TUNING_FILE, else prgpu_tuning.txt beside this library. Lines of
// "key value"; '#' starts a comment. An unreadable line or value is reported and skipped.
void LoadTuningFile() {
    std::string path;
    if (const char *e = std::getenv("PRGPU_TUNING_FILE")) {
        path = e;
    } else {
        Dl_info info;
        if (dladdr(reinterpret_cast<const void *>(&LoadTuningFile), &info) && info.dli_fname) {
            std::string lib = info.dli_fname;
            const size_t slash = lib.find_last_of('/');
            path = (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/prgpu_tuning.txt";
        }
    }
    if (path.empty()) return;
    std::FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return;
    gTuningFile = path;
    char line[512];
    int no = 0;
    while (std::fgets(line, sizeof(line), f)) {
        ++no;
        std::string s(line);
        const size_t hash = s.find('#');
        if (hash != std::string::npos) s.resize(hash);
        char key[128];
        double v = 0;
        char extra[8];
        const int got = std::sscanf(s.c_str(), " %127s %lf %7s", key, &v, extra);
        if (got <= 0) continue;  // blank or comment
        if (got != 2 || !TuneSet(gTuning, key, v)) {
            if (gTuningError.empty()) gTuningError = path + ":" + std::to_string(no) + ": invalid tuning line";
        }
    }
    std::fclose(f);
}

void Enumerate() {
    if (gEnumerated) return;
    gEnumerated = true;
    LoadTuningFile();
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    for (int i = 0; i < n; ++i) {
        auto d = std::make_unique<Device>();
        d->id = i;
        d->tune = gTuning;
        gDevices.push_back(std::move(d));
    }
}

Device &Get(int index) {
    if (index < 0 || index >= int(gDevices.size())) throw std::runtime_error("no such GPU device");
    Device &d = *gDevices[size_t(index)];
    PG_CHECK(hipSetDevice(d.id));
    if (!d.stream) {
        PG_CHECK(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        PG_CHECK(hipGetDeviceProperties(&prop, d.id));
        d.cus = prop.multiProcessorCount;
        d.flags.ensure(64);
        PG_CHECK(hipMemset(d.flags.p, 0, 64));
        // Device address of this GPU's copy of the allocation-free query's scratch array.
        void *sp = nullptr;
        PG_CHECK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
        d.scratch = static_cast<uint32_t *>(sp);
    }
    return d;
}

__global__ void k_scratch_clear(uint32_t *words, int n) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) words[i] = 0u;
}


// With kStats (profile mode 2) every lane adds its rays, nodes fetched and leaf records tested
// to counters[0..2] (closest hit) or counters[3..5] (any hit).
template <bool kStats>
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays(TraceCtx c, const prgpu::RayEx *rays, prgpu::HitEx *hits,
                                                            int64_t n, int anyHit, unsigned long long *counters) {
    __shared__ uint32_t lds[kStackLds * kTraceBlock];
    const int tid = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    TraceStats st;
    unsigned long long traced = 0;
    for (int64_t i = tid; i < n; i += stride) {
        const prgpu::RayEx rr = rays[i];
        DRay r{rr.o.x, rr.o.y, rr.o.z, rr.d.x, rr.d.y, rr.d.z};
        DHit h;
        bool hit = anyHit ? dev_trace<true, kStats>(c, r, rr.tMax, &h, lds, tid, &st)
                          : dev_trace<false, kStats>(c, r, rr.tMax, &h, lds, tid, &st);
        prgpu::HitEx o;
        o.t = hit ? h.t : 0.f;
        o.prim = hit ? h.prim : -1;
        o.b1 = hit ? h.b1 : 0.f;
        o.b2 = hit ? h.b2 : 0.f;
        hits[i] = o;
        ++traced;
    }
    if (kStats && traced) {
        unsigned long long *ctr = counters + (anyHit ? 3 : 0);
        atomicAdd(&ctr[0], traced);
        atomicAdd(&ctr[1], (unsigned long long)st.nodes);
        atomicAdd(&ctr[2], (unsigned long long)st.recs);
    }
}

}  // namespace

// Content fingerprint of the geometry: counts, pointers and a strided sample of the primitives.
// The reference keys its upload cache on the host pointer and counts only; a strided content
// sample additionally catches buffers that are reused with new contents.
uint64_t GeometryKey(const prgpu::Primitive *prims, int count) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t n) {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    };
    mix(&count, sizeof(count));
    uintptr_t ptr = uintptr_t(prims);
    mix(&ptr, sizeof(ptr));
    if (count > 0) {
        int step = std::max(1, count / 4096);
        for (int i = 0; i < count; i += step) mix(&prims[i], sizeof(prgpu::Primitive));
        mix(&prims[count - 1], sizeof(prgpu::Primitive));
    }
    return h;
}

void EnsureGeometry(Device &d, const prgpu::Primitive *prims, int count) {
    uint64_t key = GeometryKey(prims, count);
    if (key == d.geomKey && d.nodes.p) return;
    WideBVH bvh;
    BuildWideBVH(prims, count, &bvh);
    if (bvh.nodes.empty()) {  // empty scene: a root with no children
        bvh.nodes.emplace_back();
        std::memset(&bvh.nodes[0], 0, sizeof(WNode));
        bvh.depth = 1;
    }
    d.nodes.upload(bvh.nodes.data(), bvh.nodes.size(), d.stream);
    d.recs.upload(bvh.recs.data(), bvh.recs.size(), d.stream);  // an empty scene gets a small unread buffer
    d.depth = bvh.depth;
    d.nodeCount = int(bvh.nodes.size());
    d.recCount = int(bvh.recs.size());
    for (int a = 0; a < 3; ++a) {
        d.sceneLo[a] = bvh.lo[a];
        d.sceneHi[a] = bvh.hi[a];
    }
    // At most 7 entries per level are pending on a traversal stack.
    int need = 7 * bvh.depth + 8;
    d.spillCap = std::max(0, need - kStackLds);
    d.spillStride = d.cus * kTraceBlocksPerCU * kTraceBlock;
    if (d.spillCap > 0) d.spill.ensure(size_t(d.spillCap) * size_t(d.spillStride) * sizeof(uint32_t));
    PG_CHECK(hipStreamSynchronize(d.stream));
    d.geomKey = key;
}

TraceCtx MakeTraceCtx(Device &d) {
    TraceCtx c;
    c.nodes = d.nodes.as<WNode>();
    c.recs = d.recs.as<TriRec>();
    c.spill = d.spill.as<uint32_t>();
    c.spillStride = d.spillStride;
    c.spillCap = d.spillCap;
    c.errorFlag = d.flags.as<int>();
    return c;
}

int TraceGrid(const Device &d, int64_t n) {
    int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    const int perCU = std::max(1, std::min(d.tune.traceBlocksPerCU, kTraceBlocksPerCU));
    return int(std::max<int64_t>(1, std::min<int64_t>(blocks, int64_t(d.cus) * perCU)));
}

void CheckTraceFlags(Device &d) {
    int f = 0;
    PG_CHECK(hipMemcpyAsync(&f, d.flags.p, sizeof(int), hipMemcpyDeviceToHost, d.stream));
    PG_CHECK(hipStreamSynchronize(d.stream));
    if (f) {
        PG_CHECK(hipMemsetAsync(d.flags.p, 0, sizeof(int), d.stream));
        throw std::runtime_error("BVH traversal stack overflow");
    }
}

}  // namespace pgpu

using namespace pgpu;

PRGPU_API int prgpu_tune_set(int deviceIndex, const char *key, double value) {
    std::lock_guard<std::mutex> lock(gMutex);
    Enumerate();
    if (!key || deviceIndex < 0 || deviceIndex >= int(gDevices.size())) return 1;
    return TuneSet(gDevices[size_t(deviceIndex)]->tune, key, value) ? 0 : 1;
}

PRGPU_API int prgpu_tune_get(int deviceIndex, const char *key, double *value) {
    std::lock_guard<std::mutex> lock(gMutex);
    Enumerate();
    if (!key || !value || deviceIndex < 0 || deviceIndex >= int(gDevices.size())) return 1;
    return TuneGet(gDevices[size_t(deviceIndex)]->tune, key, value) ? 0 : 1;
}

PRGPU_API int prgpu_tune_source(char *path, int pathLen, char *error, int errorLen) {
    std::lock_guard<std::mutex> lock(gMutex);
    Enumerate();
    SetErr(path, pathLen, gTuningFile);
    SetErr(error, errorLen, gTuningError);
    return gTuningFile.empty() ? 1 : 0;
}

PRGPU_API int prgpu_api_version() { return prgpu::kApiVersion; }
PRGPU_API int prgpu_api_ext_version() { return prgpu::kExtVersion; }

PRGPU_API int prgpu_list_devices(prgpu::DeviceInfo *out, int max) {
    std::lock_guard<std::mutex> lock(gMutex);
    Enumerate();
    int n = 0;
    for (const auto &d : gDevices) {
        if (n >= max) break;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d->id) != hipSuccess) continue;
        prgpu::DeviceInfo &di = out[n++];
        std::memset(&di, 0, sizeof(di));
        std::snprintf(di.name, sizeof(di.name), "%s (%s)", prop.name, prop.gcnArchName);
        di.globalMemory = uint64_t(prop.totalGlobalMem);
        di.computeUnits = prop.multiProcessorCount;
        di.isGpu = 1;
        std::snprintf(di.backend, sizeof(di.backend), "hip");
    }
    return n;
}

PRGPU_API int prgpu_trace_rays(int deviceIndex, const prgpu::SceneDescEx *sd, const prgpu::RayEx *rays, int64_t n,
                               int anyHit, prgpu::HitEx *hits, char *err, int errLen) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        if (!sd || sd->apiVersion != prgpu::kApiVersion || sd->extVersion != prgpu::kExtVersion)
            throw std::runtime_error("API version mismatch");
        if (n < 0 || (n > 0 && (!rays || !hits))) throw std::runtime_error("invalid ray batch");
        Device &d = Get(deviceIndex);
        EnsureGeometry(d, sd->prims, sd->primCount);
        if (n == 0) return 0;
        d.rays.upload(rays, size_t(n), d.stream);
        d.hits.ensure(size_t(n) * sizeof(prgpu::HitEx));
        if (d.profileMode >= 2 && d.travCounters.p)
            hipLaunchKernelGGL(k_trace_rays<true>, dim3(TraceGrid(d, n)), dim3(kTraceBlock), 0, d.stream,
                               MakeTraceCtx(d), d.rays.as<prgpu::RayEx>(), d.hits.as<prgpu::HitEx>(), n, anyHit,
                               d.travCounters.as<unsigned long long>());
        else
            hipLaunchKernelGGL(k_trace_rays<false>, dim3(TraceGrid(d, n)), dim3(kTraceBlock), 0, d.stream,
                               MakeTraceCtx(d), d.rays.as<prgpu::RayEx>(), d.hits.as<prgpu::HitEx>(), n, anyHit,
                               static_cast<unsigned long long *>(nullptr));
        PG_CHECK(hipGetLastError());
        PG_CHECK(hipMemcpyAsync(hits, d.hits.p, size_t(n) * sizeof(prgpu::HitEx), hipMemcpyDeviceToHost, d.stream));
        CheckTraceFlags(d);
        return 0;
    } catch (const std::exception &e) {
        SetErr(err, errLen, e.what());
        return 1;
    }
}

namespace {
// True when [p, p + bytes) lies inside one device allocation of device `id`.
bool DeviceRange(const void *p, size_t bytes, int id) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (a.type != hipMemoryTypeDevice || a.device != id) return false;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, hipDeviceptr_t(const_cast<void *>(p))) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    const uintptr_t b = uintptr_t(base), q = uintptr_t(p);
    return q >= b && q - b <= size && bytes <= size - (q - b);
}
}  // namespace

PRGPU_API int prgpu_trace_rays_noalloc(int deviceIndex, const prgpu::SceneDescEx *sd, const prgpu::RayEx *rays,
                                       int64_t n, int anyHit, prgpu::HitEx *hits, char *err, int errLen) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        if (!sd || sd->apiVersion != prgpu::kApiVersion || sd->extVersion != prgpu::kExtVersion)
            throw std::runtime_error("API version mismatch");
        if (deviceIndex < 0 || deviceIndex >= int(gDevices.size()) || !gDevices[size_t(deviceIndex)]->stream ||
            !gDevices[size_t(deviceIndex)]->scratch)
            throw std::runtime_error("device not initialised: make the scene resident with prgpu_trace_rays or "
                                     "prgpu_render_ex on this device first");
        Device &d = *gDevices[size_t(deviceIndex)];
        PG_CHECK(hipSetDevice(d.id));
        if (!d.nodes.p || GeometryKey(sd->prims, sd->primCount) != d.geomKey)
            throw std::runtime_error("scene geometry is not resident on this device: make it resident with "
                                     "prgpu_trace_rays or prgpu_render_ex first");
        if (n < 0) throw std::runtime_error("invalid ray batch");
        if (n == 0) return 0;
        if (!DeviceRange(rays, size_t(n) * sizeof(prgpu::RayEx), d.id) ||
            !DeviceRange(hits, size_t(n) * sizeof(prgpu::HitEx), d.id))
            throw std::runtime_error("rays and hits must be device allocations of this GPU holding n records");
        // The spill slabs of all lanes must fit the scratch: cap the grid to what it holds.
        const int64_t slabWords = int64_t(kScratchWords - kScratchHeadWords);
        int grid = TraceGrid(d, n);
        if (d.spillCap > 0) {
            const int64_t maxBlocks = slabWords / (int64_t(d.spillCap) * kTraceBlock);
            if (maxBlocks < 1)
                throw std::runtime_error("BVH too deep for the allocation-free query: its traversal stacks need " +
                                         std::to_string(int64_t(d.spillCap) * kTraceBlock * 4) + " B of spill per block, "
                                         "the scratch holds " + std::to_string(slabWords * 4) + " B");
            grid = int(std::min<int64_t>(grid, maxBlocks));
        }
        TraceCtx c = MakeTraceCtx(d);
        c.spill = d.scratch + kScratchHeadWords;
        c.spillStride = grid * kTraceBlock;
        c.errorFlag = reinterpret_cast<int *>(d.scratch);
        // The scratch keeps the previous call's contents: the overflow flag must start at zero.
        hipLaunchKernelGGL(k_scratch_clear, dim3(1), dim3(64), 0, d.stream, d.scratch, 1);
        hipLaunchKernelGGL(k_trace_rays<false>, dim3(grid), dim3(kTraceBlock), 0, d.stream, c, rays, hits, n, anyHit,
                           static_cast<unsigned long long *>(nullptr));
        PG_CHECK(hipGetLastError());
        int f = 0;
        PG_CHECK(hipMemcpyAsync(&f, d.scratch, sizeof(int), hipMemcpyDeviceToHost, d.stream));
        PG_CHECK(hipStreamSynchronize(d.stream));
        if (f) throw std::runtime_error("BVH traversal stack overflow");
        return 0;
    } catch (const std::exception &e) {
        SetErr(err, errLen, e.what());
        return 1;
    }
}

PRGPU_API int prgpu_render_ex(int deviceIndex, const prgpu::SceneDescEx *sd, const prgpu::RenderParamsEx *rp,
                              double *sum, double *sumsq, char *err, int errLen) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        if (!sd || sd->apiVersion != prgpu::kApiVersion || sd->extVersion != prgpu::kExtVersion)
            throw std::runtime_error("API version mismatch");
        if (!rp || (!sum && sumsq)) throw std::runtime_error("invalid arguments");
        Device &d = Get(deviceIndex);
        EnsureGeometry(d, sd->prims, sd->primCount);
        RenderExImpl(d, *sd, *rp, sum, sumsq);
        return 0;
    } catch (const std::exception &e) {
        SetErr(err, errLen, e.what());
        return 1;
    }
}

PRGPU_API int prgpu_sample_radiance(int deviceIndex, const prgpu::SceneDescEx *sd, const prgpu::RenderParamsEx *rp,
                                    prgpu::SampleRecordEx *out, char *err, int errLen) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        if (!sd || sd->apiVersion != prgpu::kApiVersion || sd->extVersion != prgpu::kExtVersion)
            throw std::runtime_error("API version mismatch");
        if (!rp || !out) throw std::runtime_error("invalid arguments");
        Device &d = Get(deviceIndex);
        EnsureGeometry(d, sd->prims, sd->primCount);
        SampleRadianceImpl(d, *sd, *rp, out);
        return 0;
    } catch (const std::exception &e) {
        SetErr(err, errLen, e.what());
        return 1;
    }
}

PRGPU_API int prgpu_render(int deviceIndex, const prgpu::SceneDesc *sd, const prgpu::RenderParams *rp, float *outXYZ,
                           char *err, int errLen) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        if (deviceIndex < 0 || deviceIndex >= int(gDevices.size())) throw std::runtime_error("no such GPU device");
        if (!sd || sd->apiVersion != prgpu::kApiVersion) throw std::runtime_error("API version mismatch");
        if (!rp || !outXYZ) throw std::runtime_error("invalid arguments");
        Device &d = Get(deviceIndex);
        RenderV4Impl(d, *sd, *rp, outXYZ);
        return 0;
    } catch (const std::exception &e) {
        SetErr(err, errLen, e.what());
        return 1;
    }
}

PRGPU_API int prgpu_film_reset(int deviceIndex, int width, int height) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        if (width <= 0 || height <= 0) throw std::runtime_error("invalid film size");
        Device &d = Get(deviceIndex);
        size_t n = size_t(width) * size_t(height) * 3 * sizeof(double);
        d.filmSum.ensure(n);
        d.filmSq.ensure(n);
        PG_CHECK(hipMemsetAsync(d.filmSum.p, 0, n, d.stream));
        PG_CHECK(hipMemsetAsync(d.filmSq.p, 0, n, d.stream));
        PG_CHECK(hipStreamSynchronize(d.stream));
        d.filmW = width;
        d.filmH = height;
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

PRGPU_API int prgpu_film_read(int deviceIndex, double *sum, double *sumsq) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        Device &d = Get(deviceIndex);
        if (!d.filmSum.p || !sum) return 1;
        size_t n = size_t(d.filmW) * size_t(d.filmH) * 3 * sizeof(double);
        PG_CHECK(hipMemcpyAsync(sum, d.filmSum.p, n, hipMemcpyDeviceToHost, d.stream));
        if (sumsq) PG_CHECK(hipMemcpyAsync(sumsq, d.filmSq.p, n, hipMemcpyDeviceToHost, d.stream));
        PG_CHECK(hipStreamSynchronize(d.stream));
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

PRGPU_API int prgpu_profile(int deviceIndex, int mode) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        Device &d = Get(deviceIndex);
        d.profileMode = mode < 0 ? 0 : (mode > 2 ? 2 : mode);
        if (d.profileMode >= 2 && !d.travCounters.p) {
            d.travCounters.ensure(8 * sizeof(unsigned long long));
            PG_CHECK(hipMemsetAsync(d.travCounters.p, 0, 8 * sizeof(unsigned long long), d.stream));
            PG_CHECK(hipStreamSynchronize(d.stream));
        }
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

PRGPU_API int prgpu_profile_read(int deviceIndex, prgpu::ProfileStatsEx *out, int reset) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        Device &d = Get(deviceIndex);
        if (!out) return 1;
        unsigned long long c[8] = {0};
        if (d.travCounters.p) {
            PG_CHECK(hipMemcpyAsync(c, d.travCounters.p, sizeof(c), hipMemcpyDeviceToHost, d.stream));
            PG_CHECK(hipStreamSynchronize(d.stream));
        }
        *out = d.prof;
        out->closestRays = int64_t(c[0]);
        out->closestNodes = int64_t(c[1]);
        out->closestRecs = int64_t(c[2]);
        out->shadowRays = int64_t(c[3]);
        out->shadowNodes = int64_t(c[4]);
        out->shadowRecs = int64_t(c[5]);
        out->nodeBytes = int64_t(sizeof(WNode));
        out->recBytes = int64_t(sizeof(TriRec));
        if (reset) {
            d.prof = prgpu::ProfileStatsEx{};
            if (d.travCounters.p) {
                PG_CHECK(hipMemsetAsync(d.travCounters.p, 0, 8 * sizeof(unsigned long long), d.stream));
                PG_CHECK(hipStreamSynchronize(d.stream));
            }
        }
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

PRGPU_API int prgpu_release(int deviceIndex) {
    std::lock_guard<std::mutex> lock(gMutex);
    if (deviceIndex < 0 || deviceIndex >= int(gDevices.size())) return 1;
    Device &d = *gDevices[size_t(deviceIndex)];
    if (hipSetDevice(d.id) != hipSuccess) return 1;
    if (d.stream) (void)hipStreamSynchronize(d.stream);
    ReleasePT(d);
    d.nodes.release();
    d.recs.release();
    d.spill.release();
    d.rays.release();
    d.hits.release();
    d.filmSum.release();
    d.filmSq.release();
    d.filmW = d.filmH = 0;
    d.travCounters.release();
    for (hipEvent_t e : d.eventPool) (void)hipEventDestroy(e);
    d.eventPool.clear();
    d.geomKey = d.ptKey = 0;
    return 0;
}

// ---- path-guiding model probes -----------------------------------------------------------------

namespace {
pgpu::GuideState &EnsureGuide(pgpu::Device &d) {
    if (!d.guide) {
        d.guide = std::make_unique<pgpu::GuideState>();
        d.guide->Init(d.stream);
    }
    return *d.guide;
}
}  // namespace

PRGPU_API int prgpu_guide_param_count() { return pgpu::kGParams; }

PRGPU_API int prgpu_sheen_table(float *out) {
    if (out) pgpu::SheenAlbedoTable(out);
    return prgpu::kSheenAlpha * prgpu::kSheenMu;
}

PRGPU_API int prgpu_guide_params(int deviceIndex, float *params, int count, int write) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        if (!params || count != pgpu::kGParams) throw std::runtime_error("invalid arguments");
        Device &d = Get(deviceIndex);
        pgpu::GuideState &g = EnsureGuide(d);
        if (write) g.SetParams(params, d.stream);
        else PG_CHECK(hipMemcpyAsync(params, g.params.p, size_t(count) * sizeof(float), hipMemcpyDeviceToHost, d.stream));
        PG_CHECK(hipStreamSynchronize(d.stream));
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

PRGPU_API int prgpu_guide_eval(int deviceIndex, const float *records, int n, float *out41) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        if (n < 0 || (n > 0 && (!records || !out41))) throw std::runtime_error("invalid arguments");
        Device &d = Get(deviceIndex);
        pgpu::GuideState &g = EnsureGuide(d);
        if (n == 0) return 0;
        DevBuf in, out;
        in.upload(records, size_t(n) * 9, d.stream);
        out.ensure(size_t(n) * pgpu::kGOut * sizeof(float));
        g.Eval(in.as<float>(), n, out.as<float>(), d.stream);
        PG_CHECK(hipMemcpyAsync(out41, out.p, size_t(n) * pgpu::kGOut * sizeof(float), hipMemcpyDeviceToHost, d.stream));
        PG_CHECK(hipStreamSynchronize(d.stream));
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

PRGPU_API int prgpu_guide_train(int deviceIndex, const float *records, int n, int steps, float lr, float *grad,
                                float *stats) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        if (n <= 0 || !records || steps < 0 || (steps == 0 && !grad)) throw std::runtime_error("invalid arguments");
        Device &d = Get(deviceIndex);
        pgpu::GuideState &g = EnsureGuide(d);
        g.blend = d.tune.guideBlend;
        DevBuf in;
        in.upload(records, size_t(n) * 16, d.stream);
        if (steps == 0) {
            g.Gradient(in.as<float>(), nullptr, 0, n, d.stream);
            PG_CHECK(hipMemcpyAsync(grad, g.grad.p, size_t(pgpu::kGParams) * sizeof(float), hipMemcpyDeviceToHost, d.stream));
        } else {
            for (int k = 0; k < steps; ++k) g.Step(in.as<float>(), nullptr, 0, n, lr, d.stream);
        }
        if (stats) PG_CHECK(hipMemcpyAsync(stats, g.stats.p, 2 * sizeof(float), hipMemcpyDeviceToHost, d.stream));
        PG_CHECK(hipStreamSynchronize(d.stream));
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

PRGPU_API int prgpu_guide_records(int deviceIndex, float *records, int max, int *produced) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        Device &d = Get(deviceIndex);
        if (!d.guide || max < 0 || (max > 0 && !records)) return -1;
        int counts[2] = {0, 0};
        PG_CHECK(hipMemcpyAsync(counts, d.guide->trainCount.p, 2 * sizeof(int), hipMemcpyDeviceToHost, d.stream));
        PG_CHECK(hipStreamSynchronize(d.stream));
        const int n = std::max(0, std::min(int(d.guide->ringValid), max));
        if (n > 0)
            PG_CHECK(hipMemcpyAsync(records, d.guide->train.p, size_t(n) * 16 * sizeof(float), hipMemcpyDeviceToHost,
                                    d.stream));
        PG_CHECK(hipStreamSynchronize(d.stream));
        if (produced) *produced = counts[1];
        return n;
    } catch (const std::exception &) {
        return -1;
    }
}

PRGPU_API int prgpu_guide_stats(int deviceIndex, float *stats, int *steps) {
    std::lock_guard<std::mutex> lock(gMutex);
    try {
        Enumerate();
        Device &d = Get(deviceIndex);
        if (!d.guide) return 1;
        if (stats) PG_CHECK(hipMemcpyAsync(stats, d.guide->stats.p, 2 * sizeof(float), hipMemcpyDeviceToHost, d.stream));
        if (steps) *steps = d.guide->steps;
        PG_CHECK(hipStreamSynchronize(d.stream));
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}
