// A stand-in for libprender_gpu.so that needs no GPU, for CPU tests of the host library's loader,
// device listing and render session. It exports the v4 entry points and the render extension with
// the real signatures (include/prender_gpu.h). prgpu_render_ex adds, for pixel i and sample index s,
// f = (i + 1) * (s + 1) into every channel of the film sum and f^2 into the sum of squares, so the
// film after samples [0, n) is a closed form of n that batch splits and resumes cannot change.
// Environment knobs: FAKE_PRGPU_SLEEP_MS (per call), FAKE_PRGPU_FAIL_AT (sample index that fails).
#include "prender_gpu.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace {
int EnvInt(const char *name, int dflt) {
    const char *v = std::getenv(name);
    return v && *v ? std::atoi(v) : dflt;
}
}  // namespace

PRGPU_API int prgpu_api_version() { return prgpu::kApiVersion; }
PRGPU_API int prgpu_api_ext_version() { return prgpu::kExtVersion; }

PRGPU_API int prgpu_list_devices(prgpu::DeviceInfo *out, int max) {
    if (max < 1) return 0;
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->name, sizeof(out->name), "Fake \"GPU\" 0");
    std::snprintf(out->backend, sizeof(out->backend), "fake");
    out->globalMemory = uint64_t(3) << 29;  // 1536 MiB
    out->computeUnits = 7;
    out->isGpu = 1;
    return 1;
}

PRGPU_API int prgpu_render(int, const prgpu::SceneDesc *, const prgpu::RenderParams *, float *, char *err, int errLen) {
    std::snprintf(err, size_t(errLen), "the fake plugin has no v4 renderer");
    return 1;
}

PRGPU_API int prgpu_render_ex(int device, const prgpu::SceneDescEx *scene, const prgpu::RenderParamsEx *params,
                              double *sum, double *sumsq, char *err, int errLen) {
    if (device != 0) {
        std::snprintf(err, size_t(errLen), "device index %d out of range", device);
        return 1;
    }
    const int failAt = EnvInt("FAKE_PRGPU_FAIL_AT", -1);
    if (failAt >= params->firstSample && failAt < params->firstSample + params->spp) {
        std::snprintf(err, size_t(errLen), "injected failure at sample %d", failAt);
        return 1;
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(EnvInt("FAKE_PRGPU_SLEEP_MS", 0)));
    const long long n = (long long)scene->camera.width * scene->camera.height;
    for (long long i = 0; i < n; ++i)
        for (int s = params->firstSample; s < params->firstSample + params->spp; ++s) {
            const double f = double(i + 1) * double(s + 1);
            for (int c = 0; c < 3; ++c) {
                sum[3 * i + c] += f;
                if (sumsq) sumsq[3 * i + c] += f * f;
            }
        }
    return 0;
}
