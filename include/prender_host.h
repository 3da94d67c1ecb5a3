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
// Loads a scene file (reference src/scene/scene_loader.cpp:1264-1386): a .prscene.json document with
// its includes, OBJ / PLY meshes and .vol grids, plus light rigs -- rigs == null: the file's
// active_rigs; otherwise a comma-separated list replacing them ("" = none). width / height > 0
// override the film size. Returns a scene handle (ph_scene_desc, ph_scene_defaults and
// ph_scene_free apply) or null with the reason in err. What the flat scene cannot express is
// either dropped with a warning, as the reference's GPU exporter drops it, or a load error.
PH_API void *ph_load_scene(const char *path, const char *rigs, int width, int height, char *err, int errLen);
// Warnings of a loaded scene, one per line; returns the full length (out may be short or null).
PH_API int ph_scene_warnings(void *scene, char *out, int outLen);
// Render settings of a loaded scene: film clamp, time limit (s), flags bit 0 = Russian roulette,
// bit 1 = light-BVH light selection (else by power).
PH_API void ph_scene_file_settings(void *scene, float *clampLuminance, double *timeLimit, int32_t *flags);
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

// Render session (reference GpuPathIntegrator::Render, src/integrators/gpu_path.cpp:440-513, with the
// RenderControl of src/integrators/integrator.h:49-95) ------------------------------------------------
// Called after every batch, at most once per progressInterval seconds, and always after the batch
// that reaches targetSpp. stage is "render"; fraction = spp done / targetSpp, or elapsed / timeLimit
// (at most 1) for a time-limited job. A nonzero return cancels the job after that batch.
typedef int (*ph_progress_fn)(void *user, const char *stage, double fraction, double elapsed, double samplesPerPixel);

struct PhRenderControl {
    const char *checkpointPath;  // null or empty: no checkpoints
    const char *resumePath;      // null or empty: start at sample 0
    double checkpointInterval;   // seconds between periodic checkpoints
    double progressInterval;     // seconds between progress calls (the reference uses 0.5)
    ph_progress_fn onProgress;   // may be null
    void *user;                  // passed to onProgress
    volatile int32_t *cancel;    // may be null; nonzero stops the job before its next batch
};

// Renders until targetSpp samples per pixel (timeLimit <= 0) or until timeLimit seconds have passed,
// in the adaptive batches of ph_render, adding into the caller's FP64 film sum / sumsq (w*h*3; sumsq
// may be null). With a resume path the films and the sample count come from that checkpoint first
// (its header must match scene, base->maxDepth and base->seed); sample s of pixel p is the same
// function of (p, s, seed) either way, so a resumed job ends with the film of an uninterrupted one.
// base->spp and base->firstSample are ignored: passes run from the samples already done. A checkpoint
// is written every checkpointInterval seconds and once at the end (also after a cancel or a time
// limit); a failed periodic or final write is reported on stderr and does not fail the job, as in the
// reference. *sppDone = samples per pixel in the film, *scale = 1 / *sppDone (0 when none).
// Returns 0 on success, 1 with a message in err (checkpoint rejected, render failed).
PH_API int ph_render_session(void *plugin, int device, const prgpu::SceneDescEx *scene,
                             const prgpu::RenderParamsEx *base, int targetSpp, double timeLimit,
                             const PhRenderControl *control, double *sum, double *sumsq, int32_t *sppDone,
                             double *scale, char *err, int errLen);

// CLI events (reference src/cli/main.cpp:299-305, 453-468, 576-587) ----------------------------------
// Each writes one JSON line, as the reference's nlohmann::json dump() prints it (sorted keys, no
// spaces), into out (NUL-terminated, truncated to outLen - 1) and returns its full length.
// The devices event: the CPU ("cpu", cpuThreads or the hardware thread count when <= 0), then the
// devices of the plugin at pluginPath ("gpu:<i>"), and gpu_status "ok" or the loader's message.
PH_API int ph_list_devices_json(const char *pluginPath, int cpuThreads, char *out, int outLen);
PH_API int ph_progress_json(const char *stage, double fraction, double elapsed, double samplesPerPixel,
                            double acceptRate, char *out, int outLen);  // accept_rate written only when >= 0
PH_API int ph_stage_json(const char *name, char *out, int outLen);
PH_API int ph_error_json(const char *message, char *out, int outLen);

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
// Sigmoid-polynomial coefficients of an sRGB reflectance: the RGB-to-spectrum table lookup every
// RGB spectrum of the generators and the loader goes through (RGBToSpectrumTable::operator(),
// src/core/rgb2spec.cpp:278-304).
PH_API void ph_fit_sigmoid(const float *rgb, float *coeffs);
// The table fit's inputs (the reference's fit tables, src/core/rgb2spec.cpp:39-74): 3 x n
// spectrum-to-linear-sRGB weights over n fine wavelengths, the sRGB-to-XYZ matrix (9) and its white
// (3). Returns n (189); null outputs are skipped.
PH_API int ph_rgb_fit_tables(double *rgbw, double *toXyz, double *white);
// The RGB-to-spectrum table as its knots record (resolution, z nodes, coefficients); returns its
// length and copies it into out when n is at least that.
PH_API int ph_rgb_table_record(float *out, int n);
// Preetham daylight map of a sun_sky light: width x height linear-sRGB texels (kcd/m^2),
// equirectangular with v = 0 at the zenith, ground below the horizon (replaces MakePreethamSky,
// src/lights/light.h:265-268, src/lights/sky.cpp:146-208). 0 on success, 1 on bad arguments.
PH_API int ph_preetham_sky(const float *sunDir, float turbidity, float groundAlbedo, int width, int height, float *rgb);
// Sampling of an equirectangular image environment (replaces the Distribution2D of
// EnvironmentLight, src/lights/light.cpp:221-239, 279-296; src/core/sampling.cpp:5-67): for each of
// the n pairs in u, writes (u, v, pdf of the sample, density at (u, v)) to out; avgY gets the map's
// sin-weighted mean luminance. 0 on success, 1 on bad arguments.
PH_API int ph_envmap_sample(const float *rgb, int width, int height, const float *u, int n, float *out, float *avgY);
PH_API float ph_spectrum_value(const prgpu::SpectrumEx *s, const float *knots, float lambda);
