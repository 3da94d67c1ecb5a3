#pragma once
// C API of the prender host library (libprender_host.so): scene generation, the plugin loader and
// the progressive render loop that drives libprender_gpu.so -- the Linux counterpart of the
// reference's GpuPathIntegrator (src/integrators/gpu_path.cpp), which is Windows-only there.

#include "prender_gpu.h"

#define PH_API extern "C" __attribute__((visibility("default")))

// Scene generation ------------------------------------------------------------------------------
// Builds benchmark configuration `config` (1..5) at width x height (0 = the configuration's own),
// with geometry density `detail` (1 = full). Returns a handle or null with a message in err.
PH_API void *ph_generate(int config, int width, int height, float detail, char *err, int errLen);
PH_API void ph_scene_free(void *scene);
PH_API const prgpu::SceneDescEx *ph_scene_desc(void *scene);
// Render defaults of a generated scene: out = {spp, maxDepth, rrStartDepth}; *seed.
PH_API void ph_scene_defaults(void *scene, int32_t *out, uint64_t *seed);

// Plugin loading (reference gpu_path.cpp:38-101) ---------------------------------------------------
// dlopen()s the plugin and resolves the v4 symbols (and the extension ones when present). Returns a
// handle, or null with the reason in err (missing file, missing symbol, version mismatch).
PH_API void *ph_plugin_open(const char *path, char *err, int errLen);
PH_API void ph_plugin_close(void *plugin);
PH_API int ph_plugin_has_ext(void *plugin);
PH_API int ph_plugin_list_devices(void *plugin, prgpu::DeviceInfo *out, int max);

// Progressive render (reference gpu_path.cpp:458-513) -------------------------------------------
// Renders targetSpp samples per pixel (or until timeLimit seconds when timeLimit > 0) in batches
// that double while a batch takes < 0.15 s and halve above 0.5 s (1..256 spp), starting at sample
// index base->firstSample. Adds into the FP64 film sum / sumsq (width*height*3). Returns 0 on success.
PH_API int ph_render(void *plugin, int device, const prgpu::SceneDescEx *scene, const prgpu::RenderParamsEx *base,
                     int targetSpp, double timeLimit, double *sum, double *sumsq, int32_t *sppDone, double *seconds,
                     char *err, int errLen);

// Checkpoints (reference src/integrators/integrator.cpp:16-90) ---------------------------------------
// Scene fingerprint stored in checkpoint headers (the quantities of src/scene/scene.cpp:61-70).
PH_API uint64_t ph_scene_fingerprint(const prgpu::SceneDescEx *scene);
// Writes header + FP64 film sums + body {i32 passes done, optional FP64 sum-of-squares film}
// atomically (via path.tmp). sumsq may be null. 0 on success.
PH_API int ph_checkpoint_write(const char *path, const prgpu::SceneDescEx *scene, int32_t maxDepth, uint64_t seed,
                               const double *sum, const double *sumsq, int32_t passes, char *err, int errLen);
// Validates the header against (scene, maxDepth, seed) and reads the films and the pass count;
// errors as the reference: "is not a prender checkpoint", "was made for a different scene,
// integrator, device, resolution, seed or depth", "is truncated or corrupt". sumsq may be null.
PH_API int ph_checkpoint_read(const char *path, const prgpu::SceneDescEx *scene, int32_t maxDepth, uint64_t seed,
                              double *sum, double *sumsq, int32_t *passes, char *err, int errLen);

// Film helpers -------------------------------------------------------------------------------------
// Linear sRGB (the film's output colour space) of an FP64 XYZ film scaled by `scale` (src/core/film.cpp:170-175).
PH_API void ph_resolve_srgb(const double *xyz, int64_t pixels, double scale, float *rgb);
// Sigmoid-polynomial coefficients of an sRGB reflectance (RGB uplift used by the generators).
PH_API void ph_fit_sigmoid(const float *rgb, float *coeffs);
PH_API float ph_spectrum_value(const prgpu::SpectrumEx *s, const float *knots, float lambda);
