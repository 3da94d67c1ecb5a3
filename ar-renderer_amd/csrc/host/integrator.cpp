// Host side of the GPU path tracer: plugin loading and the progressive batch loop, restating
// the reference GpuPathIntegrator (src/integrators/gpu_path.cpp:38-101, 458-513) for Linux.
#include "prender_host.h"
#include "scene_build.h"

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

namespace {

void SetErr(char *err, int len, const std::string &m) {
    if (err && len > 0) std::snprintf(err, size_t(len), "%s", m.c_str());
}

struct Plugin {
    void *handle = nullptr;
    int (*version)() = nullptr;
    int (*listDevices)(prgpu::DeviceInfo *, int) = nullptr;
    int (*render)(int, const prgpu::SceneDesc *, const prgpu::RenderParams *, float *, char *, int) = nullptr;
    int (*extVersion)() = nullptr;
    int (*renderEx)(int, const prgpu::SceneDescEx *, const prgpu::RenderParamsEx *, double *, double *, char *, int) =
        nullptr;
};

struct GeneratedScene {
    phost::SceneBuilder sb;
    phost::SceneFileSettings settings;
    std::vector<std::string> warnings;
};

}  // namespace

PH_API void *ph_generate(int config, int width, int height, float detail, char *err, int errLen) {
    auto gs = std::make_unique<GeneratedScene>();
    phost::GenOptions o;
    o.width = width;
    o.height = height;
    o.detail = detail > 0 ? detail : 1.f;
    std::string e;
    if (!phost::GenerateConfig(config, o, &gs->sb, &e)) {
        SetErr(err, errLen, e);
        return nullptr;
    }
    gs->sb.Desc();
    return gs.release();
}

PH_API void *ph_load_scene(const char *path, const char *rigs, int width, int height, char *err, int errLen) {
    if (!path) {
        SetErr(err, errLen, "invalid arguments");
        return nullptr;
    }
    auto gs = std::make_unique<GeneratedScene>();
    std::vector<std::string> rigList;
    if (rigs) {
        std::string r(rigs), cur;
        for (char c : r + ",") {
            if (c == ',') {
                if (!cur.empty()) rigList.push_back(cur);
                cur.clear();
            } else if (c != ' ') {
                cur += c;
            }
        }
    }
    std::string e;
    if (!phost::LoadSceneFile(path, rigs ? &rigList : nullptr, width, height, &gs->sb, &gs->settings, &gs->warnings,
                              &e)) {
        SetErr(err, errLen, e);
        return nullptr;
    }
    gs->sb.Desc();
    return gs.release();
}

PH_API int ph_scene_warnings(void *s, char *out, int outLen) {
    std::string all;
    for (const std::string &w : static_cast<GeneratedScene *>(s)->warnings) all += w + "\n";
    if (out && outLen > 0) std::snprintf(out, size_t(outLen), "%s", all.c_str());
    return int(all.size());
}

PH_API void ph_scene_file_settings(void *s, float *clampLuminance, double *timeLimit, int32_t *flags) {
    const phost::SceneFileSettings &f = static_cast<GeneratedScene *>(s)->settings;
    if (clampLuminance) *clampLuminance = f.clampLuminance;
    if (timeLimit) *timeLimit = f.timeLimit;
    if (flags) *flags = (f.russianRoulette ? 1 : 0) | (f.lightBVH ? 2 : 0);
}

PH_API void ph_scene_free(void *s) { delete static_cast<GeneratedScene *>(s); }

PH_API const prgpu::SceneDescEx *ph_scene_desc(void *s) { return &static_cast<GeneratedScene *>(s)->sb.Desc(); }

PH_API void ph_scene_defaults(void *s, int32_t *out, uint64_t *seed) {
    const phost::SceneBuilder &sb = static_cast<GeneratedScene *>(s)->sb;
    out[0] = sb.spp;
    out[1] = sb.maxDepth;
    out[2] = sb.rrStartDepth;
    *seed = sb.seed;
}

PH_API void *ph_plugin_open(const char *path, char *err, int errLen) {
    auto p = std::make_unique<Plugin>();
    p->handle = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!p->handle) {
        const char *e = dlerror();
        SetErr(err, errLen, std::string("cannot load GPU plugin: ") + (e ? e : path));
        return nullptr;
    }
    p->version = reinterpret_cast<int (*)()>(dlsym(p->handle, "prgpu_api_version"));
    p->listDevices = reinterpret_cast<decltype(p->listDevices)>(dlsym(p->handle, "prgpu_list_devices"));
    p->render = reinterpret_cast<decltype(p->render)>(dlsym(p->handle, "prgpu_render"));
    if (!p->version || !p->listDevices || !p->render || p->version() != prgpu::kApiVersion) {
        SetErr(err, errLen, "GPU plugin has an incompatible interface");
        dlclose(p->handle);
        return nullptr;
    }
    p->extVersion = reinterpret_cast<int (*)()>(dlsym(p->handle, "prgpu_api_ext_version"));
    if (p->extVersion && p->extVersion() == prgpu::kExtVersion)
        p->renderEx = reinterpret_cast<decltype(p->renderEx)>(dlsym(p->handle, "prgpu_render_ex"));
    return p.release();
}

PH_API void ph_plugin_close(void *pl) {
    auto *p = static_cast<Plugin *>(pl);
    if (!p) return;
    if (p->handle) dlclose(p->handle);
    delete p;
}

PH_API int ph_plugin_has_ext(void *pl) { return static_cast<Plugin *>(pl)->renderEx ? 1 : 0; }

PH_API int ph_plugin_list_devices(void *pl, prgpu::DeviceInfo *out, int max) {
    return static_cast<Plugin *>(pl)->listDevices(out, max);
}

PH_API int ph_render(void *pl, int device, const prgpu::SceneDescEx *scene, const prgpu::RenderParamsEx *base,
                     int targetSpp, double timeLimit, double *sum, double *sumsq, int32_t *sppDone, double *seconds,
                     char *err, int errLen) {
    auto *p = static_cast<Plugin *>(pl);
    if (!p || !p->renderEx) {
        SetErr(err, errLen, "GPU plugin lacks the extension interface");
        return 1;
    }
    const int target = timeLimit > 0 ? std::numeric_limits<int>::max() : std::max(1, targetSpp);
    int done = 0, batch = 1;
    char msg[1024] = {0};
    auto start = std::chrono::steady_clock::now();
    while (done < target) {
        prgpu::RenderParamsEx rp = *base;
        rp.spp = std::min(batch, target - done);
        rp.firstSample = base->firstSample + done;
        auto t0 = std::chrono::steady_clock::now();
        if (p->renderEx(device, scene, &rp, sum, sumsq, msg, sizeof(msg)) != 0) {
            SetErr(err, errLen, std::string("GPU render failed: ") + msg);
            return 1;
        }
        double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        done += rp.spp;
        if (dt < 0.15) batch = std::min(batch * 2, 256);
        else if (dt > 0.5 && batch > 1) batch /= 2;
        double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        if (timeLimit > 0 && elapsed >= timeLimit) break;
    }
    if (sppDone) *sppDone = done;
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    return 0;
}

PH_API int ph_render_session(void *pl, int device, const prgpu::SceneDescEx *scene, const prgpu::RenderParamsEx *base,
                             int targetSpp, double timeLimit, const PhRenderControl *ctl, double *sum, double *sumsq,
                             int32_t *sppDone, double *scale, char *err, int errLen) {
    using Clock = std::chrono::steady_clock;
    auto *p = static_cast<Plugin *>(pl);
    if (!p || !p->renderEx) {
        SetErr(err, errLen, "GPU plugin lacks the extension interface");
        return 1;
    }
    if (!scene || !base || !sum) {
        SetErr(err, errLen, "invalid arguments");
        return 1;
    }
    const PhRenderControl none{};
    const PhRenderControl &c = ctl ? *ctl : none;
    const std::string ckptPath = c.checkpointPath ? c.checkpointPath : "";
    const std::string resumePath = c.resumePath ? c.resumePath : "";
    const auto start = Clock::now();
    auto lastProgress = start, lastCheckpoint = start;
    auto seconds = [](Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); };

    const int target = timeLimit > 0 ? std::numeric_limits<int>::max() : std::max(1, targetSpp);
    int done = 0, batch = 1;
    char msg[1024] = {0};
    if (!resumePath.empty()) {
        int32_t passes = 0;
        if (ph_checkpoint_read(resumePath.c_str(), scene, base->maxDepth, base->seed, sum, sumsq, &passes, msg,
                               sizeof(msg)) != 0) {
            SetErr(err, errLen, msg);
            return 1;
        }
        done = passes;
    }
    auto checkpoint = [&] {
        if (ckptPath.empty()) return;
        char e[1024] = {0};
        if (ph_checkpoint_write(ckptPath.c_str(), scene, base->maxDepth, base->seed, sum, sumsq, done, e,
                                sizeof(e)) != 0)
            std::fprintf(stderr, "warning: %s\n", e);
    };
    bool cancel = c.cancel && *c.cancel;
    while (done < target && !cancel) {
        prgpu::RenderParamsEx rp = *base;
        rp.spp = std::min(batch, target - done);
        rp.firstSample = done;
        const auto t0 = Clock::now();
        if (p->renderEx(device, scene, &rp, sum, sumsq, msg, sizeof(msg)) != 0) {
            SetErr(err, errLen, std::string("GPU render failed: ") + msg);
            return 1;
        }
        const auto t1 = Clock::now();
        const double dt = seconds(t0, t1);
        done += rp.spp;
        if (dt < 0.15) batch = std::min(batch * 2, 256);
        else if (dt > 0.5 && batch > 1) batch /= 2;
        const double elapsed = seconds(start, t1);
        if (c.onProgress && (done == target || seconds(lastProgress, t1) >= c.progressInterval)) {
            lastProgress = t1;
            const double fraction = timeLimit > 0 ? std::min(1.0, elapsed / timeLimit) : double(done) / target;
            if (c.onProgress(c.user, "render", fraction, elapsed, double(done)) != 0) cancel = true;
        }
        if (!ckptPath.empty() && seconds(lastCheckpoint, Clock::now()) >= c.checkpointInterval) {
            lastCheckpoint = Clock::now();
            checkpoint();
        }
        if (timeLimit > 0 && elapsed >= timeLimit) break;
        if (c.cancel && *c.cancel) cancel = true;
    }
    checkpoint();
    if (sppDone) *sppDone = done;
    if (scale) *scale = done > 0 ? 1.0 / done : 0.0;
    return 0;
}

PH_API void ph_resolve_srgb(const double *xyz, int64_t pixels, double scale, float *rgb) {
    for (int64_t i = 0; i < pixels; ++i) phost::XYZToLinearSRGB(xyz + 3 * i, scale, rgb + 3 * i);
}

PH_API void ph_fit_sigmoid(const float *rgb, float *coeffs) { phost::RgbTableCoeffs(rgb, coeffs); }
PH_API int ph_rgb_fit_tables(double *rgbw, double *toXyz, double *white) {
    return phost::RgbFitTables(rgbw, toXyz, white);
}
PH_API int ph_rgb_table_record(float *out, int n) {
    const std::vector<float> &t = phost::RgbToSpectrumTableRecord();
    if (out && n >= int(t.size())) std::copy(t.begin(), t.end(), out);
    return int(t.size());
}

PH_API int ph_envmap_sample(const float *rgb, int width, int height, const float *u, int n, float *out, float *avgY) {
    if (!rgb || width <= 0 || height <= 0 || n < 0 || (n > 0 && (!u || !out))) return 1;
    phost::EnvMapDistribution d;
    d.Build(rgb, width, height);
    for (int i = 0; i < n; ++i) {
        float *o = out + 4 * size_t(i);
        d.Sample(u[2 * size_t(i)], u[2 * size_t(i) + 1], &o[0], &o[1], &o[2]);
        o[3] = d.Pdf(o[0], o[1]);
    }
    if (avgY) *avgY = d.avgY;
    return 0;
}

PH_API int ph_preetham_sky(const float *sunDir, float turbidity, float groundAlbedo, int width, int height, float *rgb) {
    if (!sunDir || !rgb || width <= 0 || height <= 0) return 1;
    phost::PreethamSky(phost::f3(sunDir[0], sunDir[1], sunDir[2]), turbidity, groundAlbedo, width, height, rgb);
    return 0;
}

PH_API float ph_spectrum_value(const prgpu::SpectrumEx *s, const float *knots, float lambda) {
    return phost::SpectrumValue(*s, knots, lambda);
}
