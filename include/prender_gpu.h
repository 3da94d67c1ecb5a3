#pragma once
// C ABI of the prender GPU plugin (libprender_gpu.so), MI355X / gfx950 implementation.
//
// Part 1 is the version-4 interface that PRender's host loads for `--device gpu`
// (reference: src/gpu/gpu_api.h). Every struct keeps the reference's exact LP64 layout so
// an unmodified v4 host can pass its buffers straight through; the layout is pinned by the
// static_asserts at the end of this file.
//
// Part 2 is an extension, discovered with dlsym("prgpu_api_ext_version"), that carries what
// v4 cannot express: analytic spectra (exact per-wavelength evaluation instead of 5 nm
// tables), distant lights, grid media, the light-BVH sampler, the Russian-roulette start
// depth, FP64 sum and sum-of-squares film output, a ray-batch query used by the closest-hit
// parity tests, and a per-sample radiance query used by the same-RNG-stream parity tests.
// The v4 entry points are unchanged by it.

#include <stdint.h>

#define PRGPU_API extern "C" __attribute__((visibility("default")))

namespace prgpu {

// ---------------------------------------------------------------------------------------------
// Part 1: v4 interface (reference src/gpu/gpu_api.h:15-173).

constexpr int kApiVersion = 4;             // gpu_api.h:17
constexpr int kSpectrumSamples = 95;       // gpu_api.h:20 (360..830 nm every 5 nm)
constexpr float kSpectrumMin = 360.f, kSpectrumStep = 5.f;
constexpr int kFilterBins = 64;            // gpu_api.h:23
constexpr int kSheenMu = 32, kSheenAlpha = 16;  // gpu_api.h:25

struct Float3 {
    float x, y, z;
};

struct Spectrum {  // gpu_api.h:31-33: samples at kSpectrumMin + kSpectrumStep * i, linear interpolation
    float v[kSpectrumSamples];
};

struct BVHNode {  // gpu_api.h:35-40: the host's binary BVH, depth-first order
    Float3 bmin, bmax;
    int32_t offset;   // leaf: first slot in primOrder; interior: index of the second child
    uint16_t nPrims;  // 0 for interior nodes
    uint8_t axis, pad;
};

enum PrimType : uint32_t { PrimTriangle = 0, PrimSphere = 1 };

struct Primitive {  // gpu_api.h:44-56
    uint32_t type;
    int32_t material;                     // index into materials
    int32_t light;                        // index into lights for area emitters, else -1
    int32_t mediumInside, mediumOutside;  // -1 = vacuum
    uint32_t pad;
    Float3 p0, p1, p2;                    // triangle vertices; sphere: p0 = centre
    Float3 n0, n1, n2;                    // per-vertex normals, all zero = flat
    float u0, v0, u1, v1, u2, v2;         // texture coordinates
    float r;                              // sphere radius
    uint32_t flip;                        // sphere: normals point inwards
};

enum MaterialType : uint32_t {
    MatDiffuse = 0, MatConductor = 1, MatDielectric = 2, MatBlack = 3, MatInterface = 4,
    MatCoatedDiffuse = 5, MatCoatedConductor = 6
};

struct Material {  // gpu_api.h:61-81
    uint32_t type;
    int32_t albedo;        // spectrum: diffuse reflectance / conductor artist reflectance, -1 none
    int32_t albedoB;       // second checker spectrum, -1 none
    float checkerScale;    // checker frequency in uv, 0 = no checker
    int32_t eta, k;        // spectra: conductor n and k, dielectric IOR
    float alphaX, alphaY;  // GGX alpha = roughness^2; max < 1e-3 is smooth
    uint32_t dispersive;   // IOR varies with wavelength (secondary wavelengths terminate)
    uint32_t thin;
    int32_t coatEta;
    float coatAlpha, thickness, coatG;
    int32_t coatAlbedo;
    uint32_t coatDispersive;
    int32_t coatMaxDepth, coatSamples;
    int32_t sheenColor;
    float sheenRoughness, sheenWeight;
    uint32_t shading;  // extension scenes: knots offset + 1 of the shading-perturbation record (normal / bump
                       // map, kShadeHeader), 0 = none; v4 padding (the v4 entry point clears it)
};

struct Medium {  // gpu_api.h:84-87: homogeneous, sigma = spectrum * scale, Henyey-Greenstein g
    int32_t sigmaA, sigmaS;
    float scale, g;
};

enum FilterType : uint32_t { FilterBox = 0, FilterGaussian = 1, FilterBlackmanHarris = 2 };

struct Filter {  // gpu_api.h:91-95
    uint32_t type;
    float radius, sigma, norm;
    float cdf[kFilterBins + 1];
};

enum LightType : uint32_t { LightArea = 0, LightPoint = 1 };

struct Light {  // gpu_api.h:99-110
    uint32_t type;
    int32_t primitive;  // area light: emitting primitive
    int32_t spectrum;   // radiance (area) or intensity (point)
    float scale;
    Float3 position;    // point light
    uint32_t twoSided;
    float cosPower;
    float area;
    float pmf;          // power-based selection probability
    uint32_t pad;
};

struct Camera {  // gpu_api.h:112-118
    float cameraToWorld[12];  // row-major 3x4
    float tanHalfFov, aspect;
    float lensRadius, focalDistance;
    int32_t width, height, margin;
    uint32_t pad;
};

struct SceneDesc {  // gpu_api.h:120-146
    int32_t apiVersion;
    const BVHNode *nodes;
    int32_t nodeCount;
    const int32_t *primOrder;
    const Primitive *prims;
    int32_t primCount;
    const Material *materials;
    int32_t materialCount;
    const Light *lights;
    int32_t lightCount;
    const Spectrum *spectra;
    int32_t spectrumCount;
    const Medium *media;
    int32_t mediumCount;
    int32_t cameraMedium;
    int32_t envSpectrum;
    float envScale;
    float envPmf;
    Float3 worldCenter;
    float worldRadius;
    Camera camera;
    Filter filter;
    const float *sheenAlbedo;
    const float *cie;
    float cieYIntegral;
};

struct RenderParams {  // gpu_api.h:148-154
    int32_t spp;
    int32_t firstSample;
    uint64_t seed;
    int32_t maxDepth;      // <= 0: unlimited
    float clampLuminance;  // 0 = off
};

struct DeviceInfo {  // gpu_api.h:156-162
    char name[256];
    uint64_t globalMemory;
    int32_t computeUnits;
    int32_t isGpu;
    char backend[32];
};

// ---------------------------------------------------------------------------------------------
// Part 2: extension (no counterpart in the reference ABI; see the file comment).

constexpr int kExtVersion = 4;

// Analytic spectrum: value(lambda) = scale * base(lambda). Mirrors the reference Spectrum
// classes (src/core/spectrum.h:151-231, src/core/rgb2spec.h:13-74) so the device evaluates
// exactly what the CPU path tracer evaluates at each hero wavelength.
enum SpectrumKind : uint32_t {
    SpecConstant = 0,   // base = p[0]
    SpecSigmoid = 1,    // base = S(p[0] l^2 + p[1] l + p[2]), S(x) = 1/2 + x / (2 sqrt(1 + x^2))
    SpecBlackbody = 2,  // base = Planck(l, T = p[0]) * p[1]   (p[1] = 1 / Planck(peak))
    SpecPiecewise = 3,  // base = piecewise-linear over knots[offset .. offset+2*count): lambdas then values
    SpecSellmeier = 4,  // base = sqrt(max(1, 1 + sum_i p[i] l2 / (l2 - p[3+i]))), l in um, count terms (<= 3)
    SpecCauchy = 5,     // base = p[0] + p[1] / l^2
    SpecTable95 = 6,    // base = linear interpolation of knots[offset .. offset+95) at 5 nm from 360 nm (v4 tables)
    SpecImage = 7,      // an image texture: the record at knots[offset] (kTexHeader), looked up at the hit's uv and
                        // uplifted as an albedo (RGBAlbedoSample); only in a material's albedo slot
};

// Record of an image texture in SceneDescEx::knots (ImageSpectrumTexture / ImageFloatTexture,
// reference src/materials/texture.h:67-124, texture.cpp:8-52): [0] width w, [1] height h, [2] wrap
// (0 repeat, 1 clamp, 2 mirror), [3] flip v (1: v = 1 - v before the lookup), [4] [5] uv scale,
// [6] [7] uv offset, [8] [9] [10] RGB multiplier, [11] the knots offset of the RGB-to-spectrum table
// record; then w x h linear RGB texels, top row first.
constexpr int kTexHeader = 12;  // (record offsets held in floats are below 2^24)
// Record of a material's shading perturbation (Material::PerturbShading, src/materials/
// material.cpp:16-58), at knots[Material::shading - 1]: [0] 1 with a normal map, [1] its texture
// record's knots offset, [2] strength, [3] flip green (1 / 0); [4] 1 with a bump map, [5] its height
// texture record's knots offset, [6] the channel read (-1: the mean of r, g, b), [7] the bump scale.
constexpr int kShadeHeader = 8;

struct SpectrumEx {  // 48 bytes
    uint32_t kind;
    float scale;
    int32_t offset, count;
    float p[8];
};

enum LightKindEx : uint32_t {
    LightExArea = 0,         // DiffuseAreaLight on primitive `primitive` (reference src/lights/light.cpp:10-103)
    LightExPoint = 1,        // PointLight at `position`
    LightExDistant = 2,      // DistantLight arriving from `direction` (src/lights/sky.cpp:52-60)
    LightExEnvironment = 3,  // constant EnvironmentLight (src/lights/light.cpp:215-297)
    LightExSpot = 4,         // SpotLight at `position` pointing along `direction` (src/lights/light.cpp:117-212)
    LightExGoniometric = 5,  // GoniometricLight (IES type C profile) at `position` (src/lights/sky.cpp:10-50)
    LightExSun = 6,          // SunLight: a disk around `direction` (unit, towards the sun) of angular radius
                             // acos(cosFalloffEnd), uniform radiance (src/lights/sky.cpp:82-127)
};

// Record of a goniometric light's profile in SceneDescEx::knots, at LightEx::profileOffset:
// [0..8] the luminaire frame's x, y and z axes in world space (z = the IES nadir), [9] nV, [10] nH,
// [11] the integral of the relative intensity over the sphere (steradians), then nV vertical
// angles, nH horizontal angles (degrees, ascending) and nH x nV relative candela values in [0, 1]
// (row h holds the vertical profile at horizontal angle h). profileSize = 12 + nV + nH + nV nH.
constexpr int kIesHeader = 12;

// Record of an image environment light in SceneDescEx::knots, at LightEx::profileOffset
// (profileSize floats; a LightExEnvironment whose profileSize is 0 is a constant environment with
// its spectrum): [0] width w, [1] height h, [2] cos r and [3] sin r of its rotation r about +y
// (a world direction d is looked up at (cos r dx - sin r dz, dy, sin r dx + cos r dz)), [4] the
// integral of the sampling function, [5] the map's sin-weighted mean luminance, [6] the index of
// the D65 spectrum (unit luminance) in spectra, [7] the knots offset of the RGB-to-spectrum table
// record; then w x h linear sRGB texels (rows from the zenith down), the sampling function
// luminance x sin(theta) + 1e-6 (w x h), per row its conditional CDF (w + 1), the row integrals
// (h) and the marginal CDF (h + 1).
constexpr int kEnvHeader = 8;
// RGB-to-spectrum table record: [0] the resolution n = kRgbTableRes, n z nodes, then the sigmoid
// polynomial coefficients (3 per node) of 3 x n x n x n nodes, [max component][z][y][x].
constexpr int kRgbTableRes = 64;

struct LightEx {  // 64 bytes
    uint32_t kind;
    int32_t primitive;
    int32_t spectrum;
    float scale;
    Float3 position;   // point and spot lights
    Float3 direction;  // distant light: unit vector towards the light; spot light: unit axis
    uint32_t twoSided;
    float cosPower;
    // spot light: cosines of the falloff start and of the total width (cos(radians(deg)) in FP32);
    // intensity is scaled by smoothstep(cos theta, cosFalloffEnd, cosFalloffStart)
    float cosFalloffStart, cosFalloffEnd;
    int32_t profileOffset, profileSize;  // goniometric light: its profile record in knots (kIesHeader)
};

enum MediumKindEx : uint32_t { MediumHomogeneous = 0, MediumGrid = 1 };

struct MediumEx {  // 120 bytes
    uint32_t kind;
    int32_t sigmaA, sigmaS;  // spectra
    float scale, g;
    int32_t nx, ny, nz;      // grid resolution (cell-centred samples, x fastest)
    int64_t gridOffset;      // first density value in SceneDescEx::gridData
    float maxDensity;
    Float3 boundsMin, boundsMax;    // density bounds in medium space
    float worldToMedium[12];        // row-major 3x4 affine
};

struct CameraEx {  // thin-lens perspective camera (reference src/cameras/camera.cpp:7-48)
    float cameraToWorld[12];  // row-major 3x4
    float tanHalfFov;
    float lensRadius, focalDistance;
    int32_t width, height;    // image resolution
    int32_t filterType;       // FilterType
    float filterRadius;       // pixels; the film margin is ceil(radius - 0.5)
    int32_t pad;
};

struct SceneDescEx {
    int32_t apiVersion;  // kApiVersion
    int32_t extVersion;  // kExtVersion
    const Primitive *prims;
    int32_t primCount;
    const Material *materials;  // spectrum indices refer to `spectra`
    int32_t materialCount;
    const SpectrumEx *spectra;
    int32_t spectrumCount;
    const float *knots;
    int32_t knotCount;
    const LightEx *lights;
    int32_t lightCount;
    const MediumEx *media;
    int32_t mediumCount;
    const float *gridData;
    int64_t gridCount;
    int32_t cameraMedium;
    int32_t pad;
    CameraEx camera;
};

enum LightSamplerKind : int32_t { LightSamplerBVH = 0, LightSamplerPower = 1 };

// RenderParamsEx::flags. Learned path guiding (no counterpart in the reference): diffuse
// vertices sample a one-sample MIS mixture of a learned 8-lobe vMF distribution and the BSDF;
// with RenderGuideTrain the model is trained after every batch of the call (Adam on the KL
// divergence from the incident radiance); RenderGuideReset starts from the initial model. The
// model persists on the device across calls for the same scene.
enum RenderFlagsEx : int32_t { RenderGuide = 1, RenderGuideTrain = 2, RenderGuideReset = 4 };

struct RenderParamsEx {
    int32_t spp;               // passes rendered by this call
    int32_t firstSample;       // sample index of the first pass
    uint64_t seed;
    int32_t maxDepth;          // <= 0: unlimited
    float clampLuminance;      // firefly clamp on sample luminance, 0 = off
    int32_t rrStartDepth;      // Russian roulette once depth > rrStartDepth (reference: 1)
    int32_t russianRoulette;   // 0 disables
    int32_t lightSampler;      // LightSamplerKind
    int32_t flags;             // reserved, 0
};

struct RayEx {  // one query ray for prgpu_trace_rays
    Float3 o, d;
    float tMax;
};

struct HitEx {
    float t;
    int32_t prim;  // original primitive index, -1 = miss
    float b1, b2;  // triangle barycentrics
};

// One camera sample's result for prgpu_sample_radiance: the radiance that the CPU path
// tracer's Li returns for the same (pixel, sample index, seed), before the film.
struct SampleRecordEx {
    float L[4];
    float lambda[4];
    float pdf[4];
    float pFilmX, pFilmY;
    float weight;        // camera ray weight; 0 = no ray
    uint32_t flags;      // bit 0: L holds NaN/Inf (the film drops the sample)
};

// Profiling counters of one kind of kernel launch (prgpu_profile_read). Intersect: the closest hits
// (and, without media or interfaces, the previous bounce's shadow rays); surface: the surface and
// emission stages and the sort of the material queues (tuning key material_sort); material: the
// BSDF-family and medium-scattering kernels; guide: guiding inference and training.
enum KernelKindEx : int {
    KernCamera = 0, KernIntersect = 1, KernSurface = 2, KernShadow = 3, KernFinish = 4, KernGather = 5, KernGuide = 6,
    KernMaterial = 7, KernCount = 8
};
struct ProfileStatsEx {
    double ms[8];          // summed HIP-event durations per kernel kind
    int64_t launches[8];
    int64_t items[8];      // rays / paths / pixels processed
    // Traversal counters (profile mode 2): rays traced, wide nodes fetched, leaf records tested.
    int64_t closestRays, closestNodes, closestRecs;
    int64_t shadowRays, shadowNodes, shadowRecs;
    int64_t nodeBytes, recBytes;  // record sizes of the layout in HBM
};

} // namespace prgpu

// ---- v4 entry points (reference gpu_api.h:166-173) ------------------------------------------
// Number of devices written to out (at most max).
PRGPU_API int prgpu_list_devices(prgpu::DeviceInfo *out, int max);
// Renders params->spp samples per pixel and ADDS, per pixel, the sum over those samples of the
// filter-weighted XYZ (normalised filter, norm 1 / (4 * CIE_Y_Integral)) into
// outXYZ[3 * (y * width + x)]; the caller divides by the total spp. 0 on success, else a
// message in err. Coated and sheen materials use scene->sheenAlbedo when it is not null.
PRGPU_API int prgpu_render(int deviceIndex, const prgpu::SceneDesc *scene, const prgpu::RenderParams *params,
                           float *outXYZ, char *err, int errLen);
PRGPU_API int prgpu_api_version();

// ---- extension entry points (no reference counterpart) --------------------------------------
PRGPU_API int prgpu_api_ext_version();
// Renders params->spp passes over the extended film. Each pass splats every sample with the
// film's reconstruction filter (as the CPU film does) and ADDS, per image pixel and channel,
// the pass value to sumXYZ and its square to sumSqXYZ (both width*height*3, may be null
// for sumSqXYZ). 0 on success.
PRGPU_API int prgpu_render_ex(int deviceIndex, const prgpu::SceneDescEx *scene, const prgpu::RenderParamsEx *params,
                              double *sumXYZ, double *sumSqXYZ, char *err, int errLen);
// Closest-hit (anyHit = 0) or any-hit (anyHit = 1) queries for n rays; rays and hits are
// host arrays. For any-hit queries only HitEx::prim (>= 0: occluded) is meaningful.
PRGPU_API int prgpu_trace_rays(int deviceIndex, const prgpu::SceneDescEx *scene, const prgpu::RayEx *rays, int64_t n,
                               int anyHit, prgpu::HitEx *hits, char *err, int errLen);
// Allocation-free closest-hit / any-hit query, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
// reference counterpart. rays and hits are DEVICE arrays of this GPU holding n records each;
// the scene's geometry must already be resident (a previous prgpu_trace_rays or prgpu_render_ex
// with the same scene on this device), else the call fails. It calls no hipMalloc / hipFree:
// its traversal-stack spill and overflow flag live in a fixed 64 MiB scratch array per GPU (the
// grid is capped to the lanes whose spill slabs fit; a tree too deep for one block's slabs is
// rejected). Calls that overlap on one GPU share that scratch and race; the call is synchronous
// (the hits are written when it returns). Results equal prgpu_trace_rays' on the same rays.
PRGPU_API int prgpu_trace_rays_noalloc(int deviceIndex, const prgpu::SceneDescEx *scene, const prgpu::RayEx *rays,
                                       int64_t n, int anyHit, prgpu::HitEx *hits, char *err, int errLen);
// Radiance of one camera sample per extended-film pixel for sample index params->firstSample
// (params->spp is ignored): out has (width + 2m) * (height + 2m) records in raster order.
PRGPU_API int prgpu_sample_radiance(int deviceIndex, const prgpu::SceneDescEx *scene, const prgpu::RenderParamsEx *params,
                                    prgpu::SampleRecordEx *out, char *err, int errLen);
// Device-resident film: prgpu_render_ex with sumXYZ == null adds into an FP64 sum / sum-of-squares
// film held in HBM (no per-call PCIe transfer); reset zeroes it for a width x height image and
// read copies it out (sumSqXYZ may be null).
PRGPU_API int prgpu_film_reset(int deviceIndex, int width, int height);
PRGPU_API int prgpu_film_read(int deviceIndex, double *sumXYZ, double *sumSqXYZ);
// Profiling: mode 0 off, 1 HIP-event timing of every kernel launch on the plugin's stream, 2 timing
// plus traversal counters (a counting variant of the traversal kernels). read() returns the
// totals since the last reset.
PRGPU_API int prgpu_profile(int deviceIndex, int mode);
PRGPU_API int prgpu_profile_read(int deviceIndex, prgpu::ProfileStatsEx *out, int reset);
// Run-time tuning (no reference counterpart): launch and batching parameters, set per device.
// Keys: trace_blocks_per_cu (persistent traversal blocks per CU, 1..8), shade_blocks_per_cu
// (blocks per CU of the surface / material kernels), path_budget (paths in flight per batch: the
// queue capacity), guide_batch (records per Adam step), guide_steps and guide_min_steps (Adam
// steps per batch of paths), guide_lr, guide_bounces (bounces that sample the guide, 0 = all),
// guide_blend (weight of the MIS-proposal term in the training loss, the rest fits the mixture),
// graph_rounds (bounce rounds replayed per captured HIP graph, even, 0 = eager launches),
// guide_replay (training records kept across batches in a ring, 0 = train on each batch's own),
// material_sort (1: each bounce sorts the material queues by material, and the medium scattering
// vertices by medium, before the material stage; 0: family order only), ray_sort (1: each bounce
// sorts the next bounce's rays by direction octant and origin cell; 0: the order they leave the
// material stage in).
// Defaults come from the tuning file the autotuner writes ($PRGPU_TUNING_FILE, else
// prgpu_tuning.txt beside the library: "key value" lines), read when devices are first
// enumerated. set / get return 0, or 1 for an unknown key, a value out of range or no device.
PRGPU_API int prgpu_tune_set(int deviceIndex, const char *key, double value);
PRGPU_API int prgpu_tune_get(int deviceIndex, const char *key, double *value);
// The tuning file read (0) or none (1), and the first invalid line found in it (empty if none).
PRGPU_API int prgpu_tune_source(char *path, int pathLen, char *error, int errorLen);
// Frees every device buffer the plugin holds for the device.
PRGPU_API int prgpu_release(int deviceIndex);

// ---- path-guiding model probes (no reference counterpart) ---------------------------------------
// The model's parameters are one FP32 vector of prgpu_guide_param_count() values: per layer the
// weights Wt[out][in] (rows 64, 64, 64, 64, 48; 64 columns) then the biases.
PRGPU_API int prgpu_guide_param_count();
// Reads (write = 0) or replaces (write = 1; also resets the optimiser) the device's model.
PRGPU_API int prgpu_guide_params(int deviceIndex, float *params, int count, int write);
// The network's 41 raw outputs for n records of 9 floats (position normalised to the scene box,
// normal, outgoing direction; the network's inputs are the position's and the normal's encodings,
// the direction is kept in the record but not read), evaluated with the BF16 matrix-core kernel.
PRGPU_API int prgpu_guide_eval(int deviceIndex, const float *records, int n, float *out41);
// Gradient of the normalised training loss over n records of 16 floats (position, normal, wo,
// sampled direction, p_bsdf, weight, 2 unused) into grad (param_count floats); stats = {summed
// weight, summed weighted loss}. With steps > 0, instead runs that many Adam steps (learning rate
// lr) on the records and returns the stats of the last step.
PRGPU_API int prgpu_guide_train(int deviceIndex, const float *records, int n, int steps, float lr, float *grad,
                                float *stats);
// The training records (16 floats each, as prgpu_guide_train takes them) that the device's
// record ring holds (the records its training renders kept, at most guide_replay of them or, with
// no replay, the last batch's): copies at most max into records, returns the count copied (-1
// when the device has no model) and sets *produced (may be null) to the records the last batch
// produced before its uniform subsample.
PRGPU_API int prgpu_guide_records(int deviceIndex, float *records, int max, int *produced);
// State of the device's model: stats = {summed weight, summed weighted loss} of its last training
// step, *steps = Adam steps taken since it was created or reset. 1 when no model exists.
PRGPU_API int prgpu_guide_stats(int deviceIndex, float *stats, int *steps);
// Writes the sheen directional-albedo table (kSheenAlpha x kSheenMu, alpha-major) that coated and
// sheen materials use when the scene description carries none. Host only; returns the count.
PRGPU_API int prgpu_sheen_table(float *out);

// ---- layout pins (measured on the reference's LP64 build) -----------------------------------
static_assert(sizeof(prgpu::Float3) == 12, "Float3");
static_assert(sizeof(prgpu::Spectrum) == 380, "Spectrum");
static_assert(sizeof(prgpu::BVHNode) == 32, "BVHNode");
static_assert(sizeof(prgpu::Primitive) == 128, "Primitive");
static_assert(__builtin_offsetof(prgpu::Primitive, p0) == 24, "Primitive.p0");
static_assert(__builtin_offsetof(prgpu::Primitive, n0) == 60, "Primitive.n0");
static_assert(__builtin_offsetof(prgpu::Primitive, u0) == 96, "Primitive.u0");
static_assert(__builtin_offsetof(prgpu::Primitive, r) == 120, "Primitive.r");
static_assert(sizeof(prgpu::Material) == 88, "Material");
static_assert(sizeof(prgpu::Medium) == 16, "Medium");
static_assert(sizeof(prgpu::Filter) == 276, "Filter");
static_assert(sizeof(prgpu::Light) == 48, "Light");
static_assert(sizeof(prgpu::Camera) == 80, "Camera");
static_assert(sizeof(prgpu::SceneDesc) == 520, "SceneDesc");
static_assert(__builtin_offsetof(prgpu::SceneDesc, camera) == 140, "SceneDesc.camera");
static_assert(__builtin_offsetof(prgpu::SceneDesc, filter) == 220, "SceneDesc.filter");
static_assert(__builtin_offsetof(prgpu::SceneDesc, sheenAlbedo) == 496, "SceneDesc.sheenAlbedo");
static_assert(__builtin_offsetof(prgpu::SceneDesc, cie) == 504, "SceneDesc.cie");
static_assert(sizeof(prgpu::RenderParams) == 24, "RenderParams");
static_assert(__builtin_offsetof(prgpu::RenderParams, seed) == 8, "RenderParams.seed");
static_assert(sizeof(prgpu::DeviceInfo) == 304, "DeviceInfo");
static_assert(sizeof(prgpu::SpectrumEx) == 48, "SpectrumEx");
static_assert(sizeof(prgpu::LightEx) == 64, "LightEx");
static_assert(sizeof(prgpu::MediumEx) == 120, "MediumEx");
static_assert(sizeof(prgpu::SampleRecordEx) == 64, "SampleRecordEx");
static_assert(sizeof(prgpu::ProfileStatsEx) == 256, "ProfileStatsEx");
