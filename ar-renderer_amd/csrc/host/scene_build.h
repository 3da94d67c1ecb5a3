// Host-side scene assembly for the prender GPU plugin: a flat-array builder that produces the
// plugin's SceneDescEx, the spectral helpers a scene needs (blackbodies normalised to unit
// luminance, named metals and glasses, RGB reflectance uplift), procedural meshes, and the
// seeded generators of the benchmark scenes.
//
// This is the part of the reference that builds pr::Scene before GpuPathIntegrator exports it
// (src/scene/scene_loader.cpp, src/geometry/shape.cpp:302-353, src/core/noise.cpp); it feeds the
// plugin through the C ABI in include/prender_gpu.h.
#pragma once

#include "prender_gpu.h"

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace phost {

struct F3 {
    float x = 0, y = 0, z = 0;
};
inline F3 f3(float x, float y, float z) { return {x, y, z}; }
inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 norm3(F3 v) {
    float l = std::sqrt(dot3(v, v));
    float inv = 1.f / l;
    return {v.x * inv, v.y * inv, v.z * inv};
}
inline F3 crossd(F3 a, F3 b) {
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return {float(ay * bz - az * by), float(az * bx - ax * bz), float(ax * by - ay * bx)};
}
inline prgpu::Float3 pf(F3 v) { return {v.x, v.y, v.z}; }

// PCG32 seeded like the reference's RNG(seqIndex) (src/core/rng.h:19): scene generators draw
// their random parameters from RNG(config number).
struct SceneRng {
    uint64_t state = 0, inc = 1;
    explicit SceneRng(uint64_t seq);
    uint32_t next();
    float uniform();
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
};

struct Mesh {
    std::vector<F3> p, n;
    std::vector<float> uv;  // 2 per vertex
    std::vector<int> idx;
};

Mesh MakeUVSphere(F3 c, float r, int segments, int rings);
Mesh MakeIcosphere(F3 c, float r, int subdivisions);
Mesh MakeQuad(F3 corner, F3 edgeU, F3 edgeV);  // normal = normalize(edgeU x edgeV) ... see .cpp
Mesh MakeBox(F3 center, F3 size);
// Convex hull of a point set as an outward-wound triangle mesh (flat facets, no normals).
Mesh ConvexHull(const std::vector<F3> &points);
// Vertices of a 57-facet round brilliant with girdle radius 1, table up (+y).
std::vector<F3> RoundBrilliantPoints(int ringSegments = 64);

// Improved Perlin noise and fBm (src/core/noise.cpp:36-60).
float PerlinNoise(F3 p);
float FBm(F3 p, int octaves, float lacunarity = 2.f, float gain = 0.5f);

// Camera-to-world for a camera at pos looking at target, +z forward (src/core/math.cpp:87-105).
void LookAt(F3 pos, F3 target, F3 up, float out12[12]);

struct IesProfile;

class SceneBuilder {
  public:
    // ---- spectra ----
    int Constant(float v);
    int Sigmoid(float c0, float c1, float c2, float scale = 1.f);
    int RGBAlbedo(float r, float g, float b);     // smooth reflectance spectrum for an sRGB albedo
    int RGBUnbounded(float r, float g, float b);  // unbounded positive spectrum (scattering coefficients)
    int Scaled(int spectrum, float s);            // a copy of `spectrum` multiplied by s
    int BlackbodyUnitY(float T);                  // Planck at T normalised to luminance Y = 1
    int Metal(const std::string &name, bool k);   // Au, Ag, Cu, Al: n (k = false) or k
    int Glass(const std::string &name);           // Sellmeier IOR: BK7, SF11, fused_silica, diamond, sapphire
    int Sellmeier(const float *B, const float *C, int n);
    int Piecewise(const std::vector<float> &lambdas, const std::vector<float> &values, float scale = 1.f);
    int Table95(const std::vector<float> &values);  // 95 values, 360..830 nm every 5 nm
    int D65Illuminant();                            // CIE D65 normalised to luminance Y = 1 (one per scene)
    int Cauchy(float A, float Bnm2);                // IOR A + B / lambda^2 (lambda in nm)

    // ---- materials ----
    int Diffuse(int reflectance);
    int Conductor(int eta, int k, float alphaX, float alphaY);
    int ConductorReflectance(int reflectance, float alphaX, float alphaY);
    int Dielectric(int eta, float alphaX, float alphaY, bool thin = false);
    int Interface();
    // Coated materials (src/materials/material.h:124-170): a rough dielectric coat of IOR `eta`
    // over a diffuse or conductor base, with a slab of unit extinction between them whose
    // scattering albedo is `albedo` (-1: purely absorbing) and phase asymmetry g.
    struct Coat {
        int eta = -1;
        float roughness = 0;     // artist roughness; alpha = roughness^2
        float thickness = 0.01f;
        int albedo = -1;
        float g = 0;
        int maxDepth = 64, nSamples = 1;
    };
    int CoatedDiffuse(const Coat &coat, int reflectance);
    int CoatedConductor(const Coat &coat, int eta, int k, float alphaX, float alphaY);
    int CoatedConductorReflectance(const Coat &coat, int reflectance, float alphaX, float alphaY);
    // Charlie sheen over the material's closure (src/materials/material.cpp:8-12).
    void SetSheen(int material, int color, float roughness, float weight);

    // ---- geometry ----
    // wrapU: a triangle whose u values straddle the seam of a wrapped parameterisation (the
    // generators' UV spheres) gets a continuous u; meshes from files keep their uvs verbatim.
    int AddMesh(const Mesh &m, int material, int mediumInside = -1, int mediumOutside = -1, bool normals = true,
                bool wrapU = true);
    int AddSphere(F3 c, float r, int material, int mediumInside = -1, int mediumOutside = -1);
    // Every triangle of the mesh becomes its own area light (src/scene/scene_loader.cpp:785-801).
    void AddEmitter(const Mesh &m, int spectrum, float scale, bool twoSided = false);

    // ---- lights, media, camera ----
    int AddPointLight(F3 p, int spectrum, float scale);
    int AddDistantLight(F3 towardsLight, int spectrum, float scale);
    // Spot light at p along dir (normalised here); the cone is given in degrees, as the reference's
    // SpotLight takes it (src/lights/light.cpp:161-166), and stored as FP32 cosines.
    int AddSpotLight(F3 p, F3 dir, int spectrum, float scale, float totalWidthDeg, float falloffStartDeg);
    // Goniometric light at pos with the luminaire frame {x, y, z} (z = the IES nadir); the profile is
    // appended to knots as the record prender_gpu.h describes (kIesHeader).
    int AddGoniometricLight(F3 pos, const F3 frame[3], const IesProfile &p, int spectrum, float scale);
    // Sun disk towards wSun of the given angular radius (degrees, clamped to [0.01, 45]), radiance
    // spectrum * scale (src/lights/sky.cpp:84-86).
    int AddSunLight(F3 wSun, float angularRadiusDeg, int spectrum, float scale);
    // Solar radiance spectrum for a sun at sunDir through a clear atmosphere of turbidity T: a 5778 K
    // blackbody at 2e6 cd/m^2 for the nominal disk, attenuated by Rayleigh and aerosol optical depth
    // over the Kasten-Young air mass, every 5 nm (src/lights/sky.cpp:207-233).
    int SunRadiance(F3 sunDir, float T, float angularRadiusDeg);
    int AddEnvironment(int spectrum, float scale);
    // Image environment (an equirectangular linear-RGB map, rows from the zenith down, w x h):
    // its record (prgpu::kEnvHeader) and, once per scene, the RGB-to-spectrum table go to knots.
    int AddEnvironmentMap(const float *rgb, int w, int h, float scale, float rotationDeg);
    // Image texture record (prgpu::kTexHeader: linear RGB texels, top row first, w x h; wrap 0
    // repeat, 1 clamp, 2 mirror; uv scale and offset; RGB multiplier) in knots: its offset.
    int AddImageTexture(const float *rgb, int w, int h, int wrap, bool flipV, const float uvScale[2],
                        const float uvOffset[2], const float mult[3]);
    // An albedo spectrum that looks the texture record up at the hit (prgpu::SpecImage).
    int ImageSpectrum(int textureRecord);
    // Normal map (texture record or -1, strength, flip green) and bump map (height texture record or
    // -1, channel -1 / 0 / 1 / 2, scale) of a material (prgpu::kShadeHeader).
    void SetShading(int material, int normalRecord, float strength, bool flipGreen, int bumpRecord, int channel,
                    float bumpScale);
    int AddHomogeneous(int sigmaA, int sigmaS, float scale, float g);
    int AddGrid(int sigmaA, int sigmaS, float scale, float g, int nx, int ny, int nz, const std::vector<float> &data,
                F3 boundsMin, F3 boundsMax);
    void SetCamera(F3 pos, F3 target, float fovYDeg, int width, int height, float lensRadius = 0, float focalDistance = 0);
    void SetFilter(prgpu::FilterType t, float radius);

    const prgpu::SceneDescEx &Desc();
    float SpectrumY(int id) const;

    std::vector<prgpu::Primitive> prims;
    std::vector<prgpu::Material> materials;
    std::vector<prgpu::SpectrumEx> spectra;
    std::vector<float> knots;
    std::vector<prgpu::LightEx> lights;
    std::vector<prgpu::MediumEx> media;
    std::vector<float> grid;
    prgpu::CameraEx camera{};
    int cameraMedium = -1;

    // Render defaults that come with a generated scene.
    int spp = 16, maxDepth = -1, rrStartDepth = 1;
    uint64_t seed = 1;

  private:
    int AddSpectrum(const prgpu::SpectrumEx &s);
    int RgbTable();
    int d65_ = -1;
    int rgbTable_ = -1;  // knots offset of the RGB-to-spectrum table record, once added
    prgpu::SceneDescEx desc_{};
};

float SpectrumValue(const prgpu::SpectrumEx &s, const float *knots, float lambda);
// Sigmoid-polynomial coefficients (c0 l^2 + c1 l + c2 inside S, l in nm) of an sRGB colour by
// the RGB-to-spectrum table, as the reference's RGBToSpectrumTable::operator() looks them up
// (src/core/rgb2spec.cpp:278-304): every RGB reflectance, unbounded and illuminant spectrum.
void RgbTableCoeffs(const float rgb[3], float coeffs[3]);
// The table fit's inputs (src/core/rgb2spec.cpp:39-74): per fine wavelength (2.5 nm over
// [360, 830]) the spectrum-to-linear-sRGB weights (D65, white adaptation, trapezoid weight, 1 /
// CIE Y integral), 3 x n; the sRGB-to-XYZ matrix (9) and its white (3). Returns n; null outputs
// are skipped.
int RgbFitTables(double *rgbw, double *toXyz, double *white);
// Linear sRGB of a film XYZ value, with the renderer's spectral white adaptation (src/core/film.cpp:170-175).
void XYZToLinearSRGB(const double xyz[3], double scale, float rgb[3]);
// Linear sRGB of an XYZ value by the sRGB matrix alone (IEC 61966-2-1, D65 white).
void XYZToSRGB(float X, float Y, float Z, float rgb[3]);
// Preetham et al. (1999) daylight as an equirectangular linear-sRGB map of width x height texels
// (kcd/m^2), below-horizon texels holding the Lambertian ground lit by sun and sky
// (src/lights/sky.cpp:146-208). Writes 3 * width * height floats, rows top to bottom.
void PreethamSky(F3 sunDir, float turbidity, float groundAlbedo, int width, int height, float *rgb);

// Sampling tables of an equirectangular image environment (src/lights/light.cpp:221-239,
// src/core/sampling.cpp:5-67): weights luminance x sin(theta) + 1e-6 per texel, per-row
// conditional CDFs and integrals, the marginal CDF; continuous sampling and density on [0, 1]^2.
// The RGB-to-spectrum table (prgpu::kRgbTableRes^3 nodes per maximum component, the layout of
// the reference's RGBToSpectrumTable, src/core/rgb2spec.cpp:243-276), as its knots record:
// resolution, z nodes, coefficients. Built once per process.
const std::vector<float> &RgbToSpectrumTableRecord();

struct EnvMapDistribution {
    int nu = 0, nv = 0;
    std::vector<float> func, condCdf, rowInt, margCdf;
    float margInt = 0, avgY = 0;
    void Build(const float *rgb, int w, int h);
    void Sample(float u0, float u1, float *u, float *v, float *pdf) const;
    float Pdf(float u, float v) const;
};

// IESNA LM-63 type C profile (src/lights/ies.cpp): angles in degrees, candela normalised by the peak.
struct IesProfile {
    std::vector<float> v, h, cd;  // cd[h * nV + v]
    float peak = 0, integral = 0;
};
bool ParseIES(const std::string &text, IesProfile *p, std::string *err);
float IesEvaluateAngles(const IesProfile &p, float thetaDeg, float phiDeg);
float IesEvaluate(const IesProfile &p, F3 wLocal);  // wLocal in the luminaire frame (+z = nadir)
void IesIntegral(IesProfile *p);

// Settings of a scene file that the flat scene does not hold (src/scene/scene.h:29-42).
struct SceneFileSettings {
    float clampLuminance = 0;  // firefly clamp (film.clamp)
    double timeLimit = 0;      // seconds; 0 = none
    bool russianRoulette = true;
    bool lightBVH = true;      // false: light selection by power
};
// Loads a .prscene.json file (with its includes, meshes and the given or active light rigs) into
// sb, at width x height (0: the file's film size). Returns false with the reason in err.
bool LoadSceneFile(const std::string &path, const std::vector<std::string> *rigs, int width, int height,
                   SceneBuilder *sb, SceneFileSettings *settings, std::vector<std::string> *warnings, std::string *err);

struct GenOptions {
    int width = 0, height = 0;  // 0: the configuration's resolution
    float detail = 1.f;         // geometry density factor (1 = full configuration)
};
// Builds benchmark configuration 1..5. Returns false for unknown configurations.
bool GenerateConfig(int config, const GenOptions &opt, SceneBuilder *sb, std::string *err);

}  // namespace phost
