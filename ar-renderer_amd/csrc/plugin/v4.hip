// The reference's v4 entry point, prgpu_render (src/gpu/gpu_api.h:168-172), on top of the
// wavefront path tracer: the v4 scene (tabulated spectra, power-pmf lights, homogeneous media,
// importance-sampled pixel filter) is restated as an extension scene and rendered with the CPU
// path tracer's sample streams, then added into the caller's float XYZ buffer.
//
// The reference GPU module estimates each pixel from the samples it draws inside that pixel,
// weighted by filter / pdf (src/gpu/prender_gpu.cpp:1498-1560); this plugin splats every sample
// with the film's filter as the CPU film does (src/core/film.cpp:141-157). Both estimate the
// same filtered image: per pixel, the sum over spp samples of the normalised filter convolution
// of the XYZ radiance (norm 1 / (4 * CIE_Y_Integral)).
#include "pt_host.h"

#include <cstring>
#include <stdexcept>
#include <vector>

namespace pgpu {

namespace {

struct V4Scene {
    std::vector<prgpu::Material> materials;  // the caller's, with the padding word cleared
    std::vector<prgpu::SpectrumEx> spectra;
    std::vector<float> knots;
    std::vector<prgpu::LightEx> lights;
    std::vector<prgpu::MediumEx> media;
    prgpu::SceneDescEx desc{};
};

// Tabulated 95-sample spectra (gpu_path.cpp:148-153) as Table95 rows. A table whose samples are
// all equal is a constant (a plain IOR), so dispersion is decided as the CPU decides it
// (src/materials/material.cpp:109: secondary wavelengths terminate only for a varying IOR).
void ExportSpectra(const prgpu::SceneDesc &sd, V4Scene &out) {
    out.spectra.resize(size_t(sd.spectrumCount));
    out.knots.reserve(size_t(sd.spectrumCount) * prgpu::kSpectrumSamples);
    for (int i = 0; i < sd.spectrumCount; ++i) {
        const float *v = sd.spectra[i].v;
        prgpu::SpectrumEx &s = out.spectra[size_t(i)];
        s = prgpu::SpectrumEx{};
        s.scale = 1;
        bool constant = true;
        for (int k = 1; k < prgpu::kSpectrumSamples; ++k) constant = constant && v[k] == v[0];
        if (constant) {
            s.kind = prgpu::SpecConstant;
            s.p[0] = v[0];
            continue;
        }
        s.kind = prgpu::SpecTable95;
        s.offset = int32_t(out.knots.size());
        s.count = prgpu::kSpectrumSamples;
        out.knots.insert(out.knots.end(), v, v + prgpu::kSpectrumSamples);
    }
}

void ExportLights(const prgpu::SceneDesc &sd, V4Scene &out) {
    auto spec = [&](int s) {
        if (s < 0 || s >= sd.spectrumCount) throw std::runtime_error("invalid scene: light spectrum");
        return s;
    };
    for (int i = 0; i < sd.lightCount; ++i) {
        const prgpu::Light &l = sd.lights[i];
        prgpu::LightEx e{};
        e.spectrum = spec(l.spectrum);
        e.scale = l.scale;
        if (l.type == prgpu::LightArea) {
            e.kind = prgpu::LightExArea;
            e.primitive = l.primitive;
            e.twoSided = l.twoSided;
            e.cosPower = l.cosPower;
        } else if (l.type == prgpu::LightPoint) {
            e.kind = prgpu::LightExPoint;
            e.primitive = -1;
            e.position = l.position;
        } else {
            throw std::runtime_error("invalid scene: unknown light type");
        }
        out.lights.push_back(e);
    }
    // The constant environment (SceneDesc::envSpectrum) comes after the listed lights, so the
    // primitives' light indices stay valid.
    if (sd.envSpectrum >= 0) {
        prgpu::LightEx e{};
        e.kind = prgpu::LightExEnvironment;
        e.primitive = -1;
        e.spectrum = spec(sd.envSpectrum);
        e.scale = sd.envScale;
        out.lights.push_back(e);
    }
}

void ExportMedia(const prgpu::SceneDesc &sd, V4Scene &out) {
    for (int i = 0; i < sd.mediumCount; ++i) {
        const prgpu::Medium &m = sd.media[i];
        prgpu::MediumEx e{};
        e.kind = prgpu::MediumHomogeneous;
        e.sigmaA = m.sigmaA;
        e.sigmaS = m.sigmaS;
        e.scale = m.scale;
        e.g = m.g;
        out.media.push_back(e);
    }
}

void Export(const prgpu::SceneDesc &sd, V4Scene &out) {
    if (sd.primCount < 0 || (sd.primCount > 0 && !sd.prims)) throw std::runtime_error("invalid scene: primitives");
    if (sd.spectrumCount > 0 && !sd.spectra) throw std::runtime_error("invalid scene: spectra");
    if (sd.lightCount > 0 && !sd.lights) throw std::runtime_error("invalid scene: lights");
    if (sd.mediumCount > 0 && !sd.media) throw std::runtime_error("invalid scene: media");
    if (sd.materialCount > 0 && !sd.materials) throw std::runtime_error("invalid scene: materials");
    ExportSpectra(sd, out);
    ExportLights(sd, out);
    ExportMedia(sd, out);
    prgpu::SceneDescEx &d = out.desc;
    d.apiVersion = prgpu::kApiVersion;
    d.extVersion = prgpu::kExtVersion;
    d.prims = sd.prims;  // same 128-byte layout
    d.primCount = sd.primCount;
    // Same 88-byte layout and spectrum indices; the v4 padding word is the extension's shading
    // record index, which a v4 scene never has.
    out.materials.assign(sd.materials, sd.materials + std::max(0, sd.materialCount));
    for (prgpu::Material &m : out.materials) m.shading = 0;
    d.materials = out.materials.data();
    d.materialCount = sd.materialCount;
    d.spectra = out.spectra.data();
    d.spectrumCount = int32_t(out.spectra.size());
    d.knots = out.knots.data();
    d.knotCount = int32_t(out.knots.size());
    d.lights = out.lights.data();
    d.lightCount = int32_t(out.lights.size());
    d.media = out.media.data();
    d.mediumCount = int32_t(out.media.size());
    d.gridData = nullptr;
    d.gridCount = 0;
    d.cameraMedium = sd.cameraMedium;
    const prgpu::Camera &c = sd.camera;
    std::memcpy(d.camera.cameraToWorld, c.cameraToWorld, sizeof(c.cameraToWorld));
    d.camera.tanHalfFov = c.tanHalfFov;
    d.camera.lensRadius = c.lensRadius;
    d.camera.focalDistance = c.focalDistance;
    d.camera.width = c.width;
    d.camera.height = c.height;
    d.camera.filterType = int32_t(sd.filter.type);
    d.camera.filterRadius = sd.filter.radius;
}

}  // namespace

void RenderV4Impl(Device &dev, const prgpu::SceneDesc &sd, const prgpu::RenderParams &rp, float *outXYZ) {
    if (rp.spp <= 0) return;
    V4Scene s;
    Export(sd, s);
    // The scene's own geometry: the traversal reads the device-resident BVH of these primitives.
    EnsureGeometry(dev, sd.prims, sd.primCount);
    prgpu::RenderParamsEx p{};
    p.spp = rp.spp;
    p.firstSample = rp.firstSample;
    p.seed = rp.seed;
    p.maxDepth = rp.maxDepth;
    p.clampLuminance = rp.clampLuminance;
    p.rrStartDepth = 1;  // the reference's fixed Russian-roulette start (src/integrators/path.cpp:181-188)
    p.russianRoulette = 1;
    p.lightSampler = prgpu::LightSamplerBVH;
    const size_t n = size_t(sd.camera.width) * size_t(sd.camera.height) * 3;
    std::vector<double> sum(n, 0.0);
    // Coated and sheen materials: the caller's sheen albedo table when it passes one (the
    // reference host fills it from SheenBxDF::Albedo), else the same table built here.
    RenderExImpl(dev, s.desc, p, sum.data(), nullptr, sd.sheenAlbedo);
    for (size_t i = 0; i < n; ++i) outXYZ[i] += float(sum[i]);
}

}  // namespace pgpu
