// Driver over the compiled reference components (TEST INFRASTRUCTURE ONLY).
//
// Built by Makefile.ref from the reference's own sources under /root/reference/src (those that
// compile unmodified with the image's g++) into oracle/_ref/libprref.so. It exposes C entry
// points mirroring the oracle's component probes (rf_* vs or_*), so the golden-vector script in
// tests/golden/ can record the reference's outputs and the tests can check the oracle against
// them. Nothing here is shipped or called by the product.
#include "cameras/camera.h"
#include "core/color.h"
#include "core/film.h"
#include "core/math.h"
#include "core/rgb2spec.h"
#include "core/sampling.h"
#include "core/spectrum.h"
#include "geometry/bvh.h"
#include "geometry/shape.h"
#include "integrators/integrator.h"
#include "lights/ies.h"
#include "lights/light.h"
#include "lights/light_bvh.h"
#include "materials/bxdf.h"
#include "materials/material.h"
#include "materials/texture.h"
#include "scene/scene.h"

#include "../include/prender_gpu.h"

// The reference's image reader decodes PNG with stb_image asking for 3 channels, Radiance HDR with
// stb_image's float loader (src/core/image.cpp:79-101) and OpenEXR with tinyexr's LoadEXR
// (src/core/exr_read.cpp:19-37); the same vendored decoders, built here for golden vectors.
#define STB_IMAGE_IMPLEMENTATION
#define STBI_ONLY_PNG
#define STBI_ONLY_HDR
#include "stb/stb_image.h"
#define TINYEXR_IMPLEMENTATION
#define TINYEXR_USE_MINIZ 1
#define TINYEXR_USE_THREAD 0
#include "tinyexr/tinyexr.h"

#include <cstring>
#include <memory>
#include <unordered_map>
#include <sstream>
#include <vector>

using namespace pr;

namespace {

// Reference spectrum objects for the analytic descriptors the generators emit.
struct SigmoidSpec : Spectrum {
    RGBSigmoidPolynomial p;
    explicit SigmoidSpec(RGBSigmoidPolynomial q) : p(q) {}
    float operator()(float l) const override { return p(l); }
    float MaxValue() const override { return std::numeric_limits<float>::quiet_NaN(); }  // not exercised
};

SpectrumPtr MakeSpectrum(const prgpu::SpectrumEx &s, const float *knots) {
    SpectrumPtr base;
    switch (s.kind) {
    case prgpu::SpecConstant: base = std::make_shared<ConstantSpectrum>(s.p[0]); break;
    case prgpu::SpecSigmoid: base = std::make_shared<SigmoidSpec>(RGBSigmoidPolynomial(s.p[0], s.p[1], s.p[2])); break;
    case prgpu::SpecBlackbody: base = std::make_shared<BlackbodySpectrum>(s.p[0]); break;
    case prgpu::SpecPiecewise: {
        std::vector<float> l(knots + s.offset, knots + s.offset + s.count);
        std::vector<float> v(knots + s.offset + s.count, knots + s.offset + 2 * s.count);
        base = std::make_shared<PiecewiseLinearSpectrum>(l, v);
        break;
    }
    case prgpu::SpecSellmeier: {
        std::vector<float> B(s.p, s.p + s.count), C(s.p + 3, s.p + 3 + s.count);
        base = std::make_shared<SellmeierIOR>(B, C);
        break;
    }
    case prgpu::SpecCauchy: base = std::make_shared<CauchyIOR>(s.p[0], s.p[1]); break;
    case prgpu::SpecTable95: {  // 95 nodes every 5 nm from 360 nm, linear between them
        std::vector<float> l(95), v(knots + s.offset, knots + s.offset + 95);
        for (int i = 0; i < 95; ++i) l[size_t(i)] = 360.f + 5.f * float(i);
        base = std::make_shared<PiecewiseLinearSpectrum>(l, v);
        break;
    }
    default: return nullptr;
    }
    if (s.scale != 1.f) return std::make_shared<ScaledSpectrum>(s.scale, base);
    return base;
}

Vec3f V(const prgpu::Float3 &f) { return {f.x, f.y, f.z}; }

// The degree value nearest acos(c) whose cos(Radians(deg)) reproduces c exactly (within 64 ulps):
// the flat scene keeps cone angles as cosines, the reference's lights take degrees.
float ExactDegrees(float c) {
    float d = Degrees(std::acos(c));
    for (int k = 0; k < 64; ++k) {
        float lo = d, hi = d;
        for (int j = 0; j < k; ++j) lo = std::nextafter(lo, -1e30f), hi = std::nextafter(hi, 1e30f);
        if (std::cos(Radians(lo)) == c) return lo;
        if (std::cos(Radians(hi)) == c) return hi;
    }
    return d;
}

struct RefScene {
    Scene scene;
    std::vector<const Light *> lightOrder;
    std::unordered_map<const Light *, int> lightIndex;
};

SampledWavelengths Lambdas(float u, int terminate) {
    SampledWavelengths w = SampledWavelengths::SampleVisible(u);
    if (terminate) w.TerminateSecondary();
    return w;
}

}  // namespace

extern "C" {

void rf_rng_stream(uint64_t pixel, uint64_t sampleIndex, uint64_t seed, int n, float *out) {
    IndependentSampler s;
    s.StartPixelSample(pixel, sampleIndex, seed);
    for (int i = 0; i < n; ++i) out[i] = s.Get1D();
}

void rf_sample_visible(float u, float *l, float *pdf) {
    SampledWavelengths w = SampledWavelengths::SampleVisible(u);
    SampledSpectrum p = w.PDF();
    for (int i = 0; i < 4; ++i) {
        l[i] = w[i];
        pdf[i] = p[i];
    }
}

void rf_cie(float l, float *xyz) {
    xyz[0] = CIE_X(l);
    xyz[1] = CIE_Y(l);
    xyz[2] = CIE_Z(l);
}
float rf_cie_y_integral() { return CIE_Y_Integral(); }

// stb_image's float RGB decode of a Radiance HDR image (stbi_loadf_from_memory, 3 channels);
// returns w * h * 3 and writes the texels when they fit `cap`, -1 on failure.
int rf_stbi_rgbf(const unsigned char *data, int len, int *w, int *h, float *out, int cap) {
    int n = 0;
    float *d = stbi_loadf_from_memory(data, len, w, h, &n, 3);
    if (!d) return -1;
    const int size = *w * *h * 3;
    if (out && size <= cap) std::memcpy(out, d, sizeof(float) * size_t(size));
    stbi_image_free(d);
    return size;
}

// tinyexr's LoadEXR decode as the reference keeps it (R, G, B of its RGBA output); returns
// w * h * 3 and writes the texels when they fit `cap`, -1 on failure.
int rf_exr_rgb(const unsigned char *data, int len, int *w, int *h, float *out, int cap) {
    float *rgba = nullptr;
    const char *msg = nullptr;
    if (LoadEXRFromMemory(&rgba, w, h, data, size_t(len), &msg) != TINYEXR_SUCCESS) {
        if (msg) FreeEXRErrorMessage(msg);
        return -1;
    }
    const int size = *w * *h * 3;
    if (out && size <= cap)
        for (int i = 0; i < *w * *h; ++i)
            for (int c = 0; c < 3; ++c) out[3 * i + c] = rgba[4 * i + c];
    free(rgba);
    return size;
}

// stb_image's 8-bit RGB decode of an encoded image (stbi_load_from_memory with 3 channels, as
// ReadImage asks for); returns w * h * 3 and writes the texels when they fit `cap`, -1 on failure.
int rf_stbi_rgb8(const unsigned char *data, int len, int *w, int *h, unsigned char *out, int cap) {
    int n = 0;
    unsigned char *d = stbi_load_from_memory(data, len, w, h, &n, 3);
    if (!d) return -1;
    const int size = *w * *h * 3;
    if (out && size <= cap) std::memcpy(out, d, size_t(size));
    stbi_image_free(d);
    return size;
}

// The RGB-to-spectrum fit's inputs as the reference defines them (src/core/rgb2spec.cpp:39-74),
// from its own colour matrices, white adaptation, D65 and CIE functions: per fine wavelength (189
// over [360, 830], trapezoid weights) the spectrum-to-linear-sRGB weights, then the sRGB-to-XYZ
// matrix and its white.
void rf_rgb_fit_tables(double *rgbw, double *toXyz, double *white) {
    const int n = 189;
    const Mat3 toRGB = XYZToRGBMatrix(ColorSpaceId::sRGB) * SpectralWhiteAdaptation();
    const Mat3 &xyz = RGBToXYZMatrix(ColorSpaceId::sRGB);
    const Spectrum &d65 = StdIlluminantD65();
    const double h = (LambdaMax - LambdaMin) / (n - 1), norm = 1.0 / CIE_Y_Integral();
    for (int i = 0; i < n; ++i) {
        const double l = LambdaMin + i * h;
        const double w = (i == 0 || i == n - 1) ? 0.5 * h : h;
        const double I = d65(float(l));
        const double c[3] = {CIE_X(float(l)), CIE_Y(float(l)), CIE_Z(float(l))};
        for (int k = 0; k < 3; ++k) {
            double s = 0;
            for (int j = 0; j < 3; ++j) s += toRGB.m[k][j] * c[j];
            rgbw[k * n + i] = s * I * w * norm;
        }
    }
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) toXyz[3 * i + j] = xyz.m[i][j];
        white[i] = double(xyz.m[i][0]) + xyz.m[i][1] + xyz.m[i][2];
    }
}

// XYZ of radiance L carried by SampleVisible(u) (secondary wavelengths terminated when asked).
void rf_to_xyz(const float *L, float u, int terminate, float *xyz) {
    SampledWavelengths w = Lambdas(u, terminate);
    SampledSpectrum s;
    for (int i = 0; i < 4; ++i) s[i] = L[i];
    XYZ x = ToXYZ(s, w);
    xyz[0] = x.x;
    xyz[1] = x.y;
    xyz[2] = x.z;
}

float rf_spectrum_eval(const prgpu::SpectrumEx *s, const float *knots, float l) { return (*MakeSpectrum(*s, knots))(l); }
float rf_spectrum_y(const prgpu::SpectrumEx *s, const float *knots) { return SpectrumToY(*MakeSpectrum(*s, knots)); }
float rf_spectrum_max(const prgpu::SpectrumEx *s, const float *knots) { return MakeSpectrum(*s, knots)->MaxValue(); }
// Named data as the reference loader builds it: metal n/k and glass Sellmeier IOR.
float rf_named_metal(const char *name, int k, float l) {
    SpectrumPtr eta, kk;
    if (!GetNamedMetal(name, &eta, &kk)) return -1;
    return k ? (*kk)(l) : (*eta)(l);
}
float rf_named_ior(const char *name, float l) {
    SpectrumPtr s = GetNamedIOR(name);
    return s ? (*s)(l) : -1;
}
// The reference's solar radiance spectrum (sky.cpp:207-233) at 95 wavelengths from 360 nm, 5 nm apart.
// The reference's image-environment sampling on an RGB map: its Distribution2D over the weights
// the EnvironmentLight constructor forms (restated here: luminance x sin(theta) + 1e-6), sampled at
// n pairs u; out per pair: (u, v, pdf, Distribution2D::PDF(u, v), EnvironmentLight::PDF_Li of the
// sampled direction), the last from the reference light itself, which builds its own table.
void rf_envmap_sample(const float *rgb, int w, int h, const float *u, int n, float *out) {
    auto img = std::make_shared<Image>(w, h);
    std::vector<float> f(size_t(w) * h);
    for (int y = 0; y < h; ++y) {
        float sinTheta = std::sin(Pi * (y + 0.5f) / h);
        for (int x = 0; x < w; ++x) {
            const float *c = rgb + 3 * (size_t(y) * w + x);
            img->Set(x, y, RGB(c[0], c[1], c[2]));
            float lum = 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2];
            f[size_t(y) * w + x] = std::max(lum, 0.f) * sinTheta + 1e-6f;
        }
    }
    Distribution2D d(f.data(), w, h);
    EnvironmentLight light(img, 1.f, 0.f);
    for (int i = 0; i < n; ++i) {
        float pdf;
        Vec2f s = d.SampleContinuous(Vec2f(u[2 * i], u[2 * i + 1]), &pdf);
        float *o = out + 5 * size_t(i);
        o[0] = s.x, o[1] = s.y, o[2] = pdf, o[3] = d.PDF(s);
        o[4] = light.PDF_Li(LightSampleContext(), EquirectToDir(s));
    }
}

// The reference's image and mapping from a texture record in this build's layout (kTexHeader:
// w, h, wrap, flip v, uv scale, uv offset, multiplier, table offset; texels top row first).
static std::shared_ptr<const ImageData> RecordImage(const float *rec, UVMapping *map) {
    const int w = int(rec[0]), h = int(rec[1]);
    Image img(w, h);
    const float *px = rec + prgpu::kTexHeader;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float *c = px + 3 * (size_t(y) * w + x);
            img.Set(x, y, RGB(c[0], c[1], c[2]));
        }
    map->scale = Vec2f(rec[4], rec[5]);
    map->offset = Vec2f(rec[6], rec[7]);
    return std::make_shared<const ImageData>(std::move(img), WrapMode(int(rec[2])), rec[3] != 0);
}

// ImageData::Lookup(UVMapping::Map(uv)) (texture.cpp:22-36) at n points: out = n x RGB.
void rf_tex_lookup(const float *rec, const float *uv, int n, float *out) {
    UVMapping map;
    auto img = RecordImage(rec, &map);
    for (int i = 0; i < n; ++i) {
        RGB c = img->Lookup(map.Map(Vec2f(uv[2 * i], uv[2 * i + 1])));
        out[3 * i] = c.r, out[3 * i + 1] = c.g, out[3 * i + 2] = c.b;
    }
}

// Material::PerturbShading (material.cpp:16-58) of a diffuse material carrying the normal and bump
// maps of a shading record in this build's layout (kShadeHeader, texture records in knots). ctx
// per case: p, uv, n, ns, dpdus, dpdu, dpdv (20 floats); out per case: ns, dpdus, changed (7).
void rf_perturb_shading(const float *knots, const float *shade, const float *ctx, int n, float *out) {
    // (a constant reflectance: the RGB uplift lives in rgb2spec.cpp, which is not built here)
    DiffuseMaterial m(std::make_shared<ConstantSpectrumTexture>(std::make_shared<ConstantSpectrum>(0.5f), RGB(0.5f, 0.5f, 0.5f)));
    if (shade[0] != 0) {
        NormalMapParams np;
        np.image = RecordImage(knots + int(shade[1]), &np.mapping);
        np.strength = shade[2];
        np.flipGreen = shade[3] != 0;
        m.SetNormalMap(np);
    }
    if (shade[4] != 0) {
        const float *hrec = knots + int(shade[5]);
        UVMapping map;
        auto img = RecordImage(hrec, &map);
        BumpMapParams bp;
        bp.height = std::make_shared<ImageFloatTexture>(img, map, int(shade[6]), hrec[8]);
        bp.scale = shade[7];
        m.SetBumpMap(bp);
    }
    for (int i = 0; i < n; ++i) {
        const float *c = ctx + 20 * size_t(i);
        MaterialEvalContext e;
        e.p = Vec3f(c[0], c[1], c[2]);
        e.uv = Vec2f(c[3], c[4]);
        e.n = Vec3f(c[5], c[6], c[7]);
        e.ns = Vec3f(c[8], c[9], c[10]);
        e.dpdus = Vec3f(c[11], c[12], c[13]);
        e.dpdu = Vec3f(c[14], c[15], c[16]);
        e.dpdv = Vec3f(c[17], c[18], c[19]);
        const bool changed = m.PerturbShading(e);
        float *o = out + 7 * size_t(i);
        o[0] = e.ns.x, o[1] = e.ns.y, o[2] = e.ns.z, o[3] = e.dpdus.x, o[4] = e.dpdus.y, o[5] = e.dpdus.z;
        o[6] = changed ? 1.f : 0.f;
    }
}

// The reference's Preetham daylight map (MakePreethamSky), width x height x RGB, zenith row first.
void rf_preetham_sky(const float *dir, float T, float albedo, int w, int h, float *out) {
    std::shared_ptr<Image> img = MakePreethamSky(Vec3f(dir[0], dir[1], dir[2]), T, albedo, w, h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            RGB c = img->Get(x, y);
            float *o = out + 3 * (size_t(y) * w + x);
            o[0] = c.r, o[1] = c.g, o[2] = c.b;
        }
}

void rf_sun_radiance(const float *dir, float T, float radiusDeg, float *out95) {
    SpectrumPtr s = MakeSunRadiance(Vec3f(dir[0], dir[1], dir[2]), T, radiusDeg);
    for (int i = 0; i < 95; ++i) out95[i] = (*s)(360.f + 5.f * float(i));
}
float rf_blackbody_unit_y(float T, float l) {
    auto bb = std::make_shared<BlackbodySpectrum>(T);
    float y = SpectrumToY(*bb);
    return ScaledSpectrum(1 / y, bb)(l);
}

// BxDF probe with the same descriptor as or_bxdf: [type, alphaX, alphaY, etaD, R0..3, eta0..3, k0..3].
void rf_bxdf(const float *desc, int op, const float *wo, const float *wi, float uc, const float *u, float *out) {
    int type = int(desc[0]);
    TrowbridgeReitzDistribution d(desc[1], desc[2]);
    SampledSpectrum R, eta, k;
    for (int i = 0; i < 4; ++i) {
        R[i] = desc[4 + i];
        eta[i] = desc[8 + i];
        k[i] = desc[12 + i];
    }
    std::unique_ptr<BxDF> b;
    if (type == 1) b = std::make_unique<DiffuseBxDF>(R);
    else if (type == 2) b = std::make_unique<ConductorBxDF>(d, eta, k);
    else if (type == 3) b = std::make_unique<DielectricBxDF>(desc[3], d);
    else b = std::make_unique<ThinDielectricBxDF>(desc[3]);
    Vec3f o(wo[0], wo[1], wo[2]);
    if (op == 0) {
        SampledSpectrum f = b->f(o, Vec3f(wi[0], wi[1], wi[2]), TransportMode::Radiance);
        for (int i = 0; i < 4; ++i) out[i] = f[i];
    } else if (op == 1) {
        out[0] = b->PDF(o, Vec3f(wi[0], wi[1], wi[2]), TransportMode::Radiance, BxDFReflTransFlags::All);
    } else {
        auto s = b->Sample_f(o, uc, Vec2f(u[0], u[1]), TransportMode::Radiance, BxDFReflTransFlags::All);
        for (int i = 0; i < 11; ++i) out[i] = 0;
        if (s) {
            for (int i = 0; i < 4; ++i) out[i] = s->f[i];
            out[4] = s->wi.x;
            out[5] = s->wi.y;
            out[6] = s->wi.z;
            out[7] = s->pdf;
            out[8] = float(uint32_t(s->flags));
            out[9] = s->eta;
            out[10] = 1;
        }
    }
}

// Closure probe with the descriptor of or_closure (oracle.h): plain, coated (LayeredBxDF over a
// DielectricBxDF coat) and sheen-wrapped BxDFs, with transport mode and sample flags.
void rf_closure(const float *desc, int op, const float *wo, const float *wi, float uc, const float *u, int mode,
                int rt, float *out) {
    const int type = int(desc[0]);
    SampledSpectrum R, eta, k, albedo, sheen;
    for (int i = 0; i < 4; ++i) {
        R[i] = desc[4 + i];
        eta[i] = desc[8 + i];
        k[i] = desc[12 + i];
        albedo[i] = desc[19 + i];
        sheen[i] = desc[28 + i];
    }
    TrowbridgeReitzDistribution d(desc[1], desc[2]);
    std::vector<std::unique_ptr<BxDF>> keep;
    auto plain = [&](int t, float etaD) -> const BxDF * {
        if (t == 1) keep.push_back(std::make_unique<DiffuseBxDF>(R));
        else if (t == 2) keep.push_back(std::make_unique<ConductorBxDF>(d, eta, k));
        else if (t == 3) keep.push_back(std::make_unique<DielectricBxDF>(etaD, d));
        else keep.push_back(std::make_unique<ThinDielectricBxDF>(etaD));
        return keep.back().get();
    };
    const BxDF *b;
    if (type == 5 || type == 6) {
        keep.push_back(std::make_unique<DielectricBxDF>(desc[3], TrowbridgeReitzDistribution(desc[16], desc[16])));
        const BxDF *top = keep.back().get();
        const BxDF *bottom = plain(type == 5 ? 1 : 2, 1.f);
        keep.push_back(std::make_unique<LayeredBxDF>(top, bottom, desc[17], albedo, desc[18], desc[25] != 0,
                                                     int(desc[23]), int(desc[24])));
        b = keep.back().get();
    } else {
        b = plain(type, desc[3]);
    }
    if (desc[26] > 0) {
        keep.push_back(std::make_unique<SheenBxDF>(b, sheen, desc[27], desc[26]));
        b = keep.back().get();
    }
    const TransportMode m = mode ? TransportMode::Importance : TransportMode::Radiance;
    const BxDFReflTransFlags f = BxDFReflTransFlags(rt);
    Vec3f o(wo[0], wo[1], wo[2]);
    if (op == 0) {
        SampledSpectrum v = b->f(o, Vec3f(wi[0], wi[1], wi[2]), m);
        for (int i = 0; i < 4; ++i) out[i] = v[i];
    } else if (op == 1) {
        out[0] = b->PDF(o, Vec3f(wi[0], wi[1], wi[2]), m, f);
    } else if (op == 2) {
        auto s = b->Sample_f(o, uc, Vec2f(u[0], u[1]), m, f);
        for (int i = 0; i < 12; ++i) out[i] = 0;
        if (s) {
            for (int i = 0; i < 4; ++i) out[i] = s->f[i];
            out[4] = s->wi.x;
            out[5] = s->wi.y;
            out[6] = s->wi.z;
            out[7] = s->pdf;
            out[8] = float(uint32_t(s->flags));
            out[9] = s->eta;
            out[10] = 1;
            out[11] = s->pdfIsProportional ? 1.f : 0.f;
        }
    } else {
        out[0] = float(uint32_t(b->Flags()));
        out[1] = b->PDFIsExact() ? 1.f : 0.f;
    }
}

float rf_sheen_albedo(float cosTheta, float alpha) { return SheenBxDF::Albedo(cosTheta, alpha); }

// Order in which the compiler evaluates the two arguments of Vec2f(r(), r()) (the reference's
// layered random walk draws its 2D samples that way): 0 = left to right, 1 = right to left.
int rf_pair_order() {
    int n = 0;
    auto r = [&n]() { return float(n++); };
    Vec2f v(r(), r());
    return v.x == 0 ? 0 : 1;
}

float rf_fr_dielectric(float c, float eta) { return FrDielectric(c, eta); }
float rf_fr_complex(float c, float eta, float k) {
    return FrComplex(c, SampledSpectrum(eta), SampledSpectrum(k))[0];
}
float rf_hg(float c, float g) { return HenyeyGreenstein(c, g); }
void rf_sample_hg(const float *wo, float g, const float *u, float *wi, float *pdf) {
    Vec3f r = SampleHenyeyGreenstein(Vec3f(wo[0], wo[1], wo[2]), g, Vec2f(u[0], u[1]), pdf);
    wi[0] = r.x;
    wi[1] = r.y;
    wi[2] = r.z;
}
void rf_filter(int type, float radius, const float *x, int n, float *out, float *norm) {
    FilterSettings fs;
    fs.type = type == 0 ? FilterType::Box : (type == 1 ? FilterType::Gaussian : FilterType::BlackmanHarris);
    fs.radius = radius;
    Filter f(fs);
    for (int i = 0; i < n; ++i) out[i] = f.Eval1D(x[i]);
    *norm = f.Norm();
}

// ---- scene-level probes: the reference BVH, light BVH, lights and camera on a flat scene ----

void *rf_scene_create(const prgpu::SceneDescEx *d, float fovY) {
    auto rs = std::make_unique<RefScene>();
    Scene &sc = rs->scene;
    sc.primitives.resize(size_t(d->primCount));
    for (int i = 0; i < d->primCount; ++i) {
        const prgpu::Primitive &p = d->prims[i];
        std::unique_ptr<Shape> shape;
        if (p.type == prgpu::PrimSphere) {
            shape = std::make_unique<Sphere>(V(p.p0), p.r, p.flip != 0);
        } else {
            auto m = std::make_shared<TriangleMesh>();
            m->p = {V(p.p0), V(p.p1), V(p.p2)};
            bool normals = !(p.n0.x == 0 && p.n0.y == 0 && p.n0.z == 0 && p.n1.x == 0 && p.n1.y == 0 && p.n1.z == 0 &&
                             p.n2.x == 0 && p.n2.y == 0 && p.n2.z == 0);
            if (normals) m->n = {V(p.n0), V(p.n1), V(p.n2)};
            m->uv = {Vec2f(p.u0, p.v0), Vec2f(p.u1, p.v1), Vec2f(p.u2, p.v2)};
            m->indices = {0, 1, 2};
            sc.meshes.push_back(m);
            shape = std::make_unique<Triangle>(m, 0);
        }
        sc.primitives[size_t(i)].shape = shape.get();
        sc.shapes.push_back(std::move(shape));
    }
    for (int i = 0; i < d->lightCount; ++i) {
        const prgpu::LightEx &l = d->lights[i];
        SpectrumPtr sp = MakeSpectrum(d->spectra[l.spectrum], d->knots);
        std::shared_ptr<Light> light;
        if (l.kind == prgpu::LightExArea) {
            Primitive &pr = sc.primitives[size_t(l.primitive)];
            auto al = std::make_shared<DiffuseAreaLight>(pr.shape, sp, l.scale, l.twoSided != 0, l.cosPower,
                                                         &pr.mediumInterface);
            pr.areaLight = al.get();
            light = al;
        } else if (l.kind == prgpu::LightExPoint) {
            light = std::make_shared<PointLight>(V(l.position), sp, l.scale);
        } else if (l.kind == prgpu::LightExSpot) {
            // the flat scene keeps the cone as cosines; the reference takes degrees: the degree value
            // nearest acos(c) whose cos(Radians(deg)) reproduces c exactly (within 64 ulps)
            float width = ExactDegrees(l.cosFalloffEnd);
            float start = ExactDegrees(l.cosFalloffStart);
            light = std::make_shared<SpotLight>(V(l.position), V(l.direction), sp, l.scale, width, start);
        } else if (l.kind == prgpu::LightExGoniometric) {
            // the reference's IESProfile is built by its own LM-63 parser from the record's angles
            // and relative candela, written back out as a type C file
            const float *r = d->knots + l.profileOffset;
            int nV = int(r[9]), nH = int(r[10]);
            std::ostringstream t;
            t.precision(9);
            t << "IESNA:LM-63-2002\nTILT=NONE\n1 -1 1 " << nV << " " << nH << " 1 2 0 0 0\n1 1 100\n";
            for (int i = 0; i < nV + nH + nV * nH; ++i) t << r[prgpu::kIesHeader + i] << (i % 10 == 9 ? "\n" : " ");
            std::string err;
            auto prof = IESProfile::Parse(t.str(), &err);
            Frame f;
            f.x = Vec3f(r[0], r[1], r[2]);
            f.y = Vec3f(r[3], r[4], r[5]);
            f.z = Vec3f(r[6], r[7], r[8]);
            if (prof) light = std::make_shared<GoniometricLight>(V(l.position), f, prof, sp, l.scale);
            else light = std::make_shared<PointLight>(V(l.position), sp, 0.f);
        } else if (l.kind == prgpu::LightExSun) {
            light = std::make_shared<SunLight>(V(l.direction), ExactDegrees(l.cosFalloffEnd), sp, l.scale);
        } else if (l.kind == prgpu::LightExDistant) {
            light = std::make_shared<DistantLight>(V(l.direction), sp, l.scale);
        } else {
            light = std::make_shared<EnvironmentLight>(sp, l.scale);
        }
        rs->lightIndex[light.get()] = i;
        rs->lightOrder.push_back(light.get());
        sc.lights.push_back(light);
    }
    const prgpu::CameraEx &c = d->camera;
    Mat4 m;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 4; ++k) m.m[r][k] = c.cameraToWorld[r * 4 + k];
    CameraParams cp;
    cp.cameraToWorld = Transform(m);
    cp.fovY = fovY;
    cp.lensRadius = c.lensRadius;
    cp.focalDistance = c.focalDistance;
    FilterSettings fs;
    fs.type = c.filterType == 0 ? FilterType::Box : (c.filterType == 1 ? FilterType::Gaussian : FilterType::BlackmanHarris);
    fs.radius = c.filterRadius;
    Filter filt(fs);
    CameraFilmInfo info{c.width, c.height, int(std::ceil(filt.Radius() - 0.5f))};
    sc.camera = std::make_unique<PerspectiveCamera>(cp, info);
    sc.Build();
    return rs.release();
}

void rf_scene_destroy(void *h) { delete static_cast<RefScene *>(h); }

void rf_trace(void *h, const prgpu::RayEx *rays, int64_t n, int anyHit, prgpu::HitEx *hits) {
    const Scene &sc = static_cast<RefScene *>(h)->scene;
    for (int64_t i = 0; i < n; ++i) {
        Ray r(V(rays[i].o), V(rays[i].d));
        prgpu::HitEx &o = hits[i];
        o.t = 0;
        o.b1 = o.b2 = 0;
        if (anyHit) {
            o.prim = sc.GetBVH().IntersectP(r, rays[i].tMax) ? 0 : -1;
            continue;
        }
        BVH::Hit bh;
        if (sc.GetBVH().Intersect(r, rays[i].tMax, &bh)) {
            o.t = bh.t;
            o.prim = bh.prim;
            o.b1 = bh.info.b1;
            o.b2 = bh.info.b2;
        } else {
            o.prim = -1;
        }
    }
}

// Closest-hit queries on the reference's binary BVH with its traversal statistics: a replay of
// BVH::Intersect (geometry/bvh.cpp:137-191: slab test with the far distances widened by
// 1 + 6 * 1.2e-7, near child first by the split axis' direction sign, 64-entry stack) that also
// counts, per ray, the nodes whose box is tested and the primitives tested. counts: 2 per ray.
// The hits are those of BVH::Intersect; the caller checks them against rf_trace.
void rf_trace_counts(void *h, const prgpu::RayEx *rays, int64_t n, prgpu::HitEx *hits, int32_t *counts) {
    const Scene &sc = static_cast<RefScene *>(h)->scene;
    const BVH &bvh = sc.GetBVH();
    const auto &nodes = bvh.Nodes();
    const auto &order = bvh.Order();
    const float widen = 1 + 2 * 1.2e-7f * 3;
    for (int64_t i = 0; i < n; ++i) {
        Ray r(V(rays[i].o), V(rays[i].d));
        float tMax = rays[i].tMax;
        Vec3f inv(1 / r.d.x, 1 / r.d.y, 1 / r.d.z);
        const int neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
        int stack[64], sp = 0, cur = 0, nodeCount = 0, primCount = 0;
        prgpu::HitEx &o = hits[i];
        o = prgpu::HitEx{0.f, -1, 0.f, 0.f};
        while (!nodes.empty()) {
            const BVH::Node &nd = nodes[size_t(cur)];
            ++nodeCount;
            const Bounds3f &b = nd.bounds;
            float t0 = (b[neg[0]].x - r.o.x) * inv.x, t1 = (b[1 - neg[0]].x - r.o.x) * inv.x * widen;
            float y0 = (b[neg[1]].y - r.o.y) * inv.y, y1 = (b[1 - neg[1]].y - r.o.y) * inv.y * widen;
            bool inside = !(t0 > y1 || y0 > t1);
            if (inside) {
                t0 = y0 > t0 ? y0 : t0;
                t1 = y1 < t1 ? y1 : t1;
                float z0 = (b[neg[2]].z - r.o.z) * inv.z, z1 = (b[1 - neg[2]].z - r.o.z) * inv.z * widen;
                inside = !(t0 > z1 || z0 > t1);
                if (inside) {
                    t0 = z0 > t0 ? z0 : t0;
                    t1 = z1 < t1 ? z1 : t1;
                    inside = t0 < tMax && t1 > 0;
                }
            }
            if (inside && nd.nPrims == 0) {
                if (neg[nd.axis]) {
                    stack[sp++] = cur + 1;
                    cur = nd.offset;
                } else {
                    stack[sp++] = nd.offset;
                    cur = cur + 1;
                }
                continue;
            }
            if (inside) {
                for (int k = 0; k < nd.nPrims; ++k) {
                    int pi = nd.offset + k;
                    ++primCount;
                    float t;
                    HitInfo info;
                    if (sc.primitives[size_t(order[size_t(pi)])].shape->Intersect(r, tMax, &t, &info)) {
                        tMax = t;
                        o = prgpu::HitEx{t, order[size_t(pi)], info.b1, info.b2};
                    }
                }
            }
            if (sp == 0) break;
            cur = stack[--sp];
        }
        counts[2 * i] = nodeCount;
        counts[2 * i + 1] = primCount;
    }
}

int rf_light_sample(void *h, const float *p, const float *n, float u, float *pmf) {
    RefScene &rs = *static_cast<RefScene *>(h);
    *pmf = 0;
    const Light *l = rs.scene.lightBVH.Sample(Vec3f(p[0], p[1], p[2]), Vec3f(n[0], n[1], n[2]), u, pmf);
    return l ? rs.lightIndex[l] : -1;
}

float rf_light_pmf(void *h, const float *p, const float *n, int light) {
    RefScene &rs = *static_cast<RefScene *>(h);
    return rs.scene.lightBVH.PMF(Vec3f(p[0], p[1], p[2]), Vec3f(n[0], n[1], n[2]), rs.lightOrder[size_t(light)]);
}

float rf_light_power_pmf(void *h, int light) {
    RefScene &rs = *static_cast<RefScene *>(h);
    return rs.scene.lightSampler.PMF(rs.lightOrder[size_t(light)]);
}

void rf_camera_ray(void *h, float fx, float fy, float lx, float ly, float *out) {
    const Scene &sc = static_cast<RefScene *>(h)->scene;
    CameraSample cs{Vec2f(fx, fy), Vec2f(lx, ly)};
    auto r = sc.camera->GenerateRay(cs);
    out[0] = r->o.x;
    out[1] = r->o.y;
    out[2] = r->o.z;
    out[3] = r->d.x;
    out[4] = r->d.y;
    out[5] = r->d.z;
}

// Li sample from a light for hero wavelengths SampleVisible(ul): out = L0..3, wi, pdf, ok.
void rf_light_sample_li(void *h, int light, const float *p, const float *n, const float *u, float ul, float *out) {
    RefScene &rs = *static_cast<RefScene *>(h);
    SampledWavelengths w = SampledWavelengths::SampleVisible(ul);
    LightSampleContext ctx;
    ctx.p = Vec3f(p[0], p[1], p[2]);
    ctx.n = ctx.ns = Vec3f(n[0], n[1], n[2]);
    for (int i = 0; i < 9; ++i) out[i] = 0;
    auto ls = rs.lightOrder[size_t(light)]->SampleLi(ctx, Vec2f(u[0], u[1]), w);
    if (!ls) return;
    for (int i = 0; i < 4; ++i) out[i] = ls->L[i];
    out[4] = ls->wi.x;
    out[5] = ls->wi.y;
    out[6] = ls->wi.z;
    out[7] = ls->pdf;
    out[8] = 1;
}

float rf_light_pdf_li(void *h, int light, const float *p, const float *n, const float *wi) {
    RefScene &rs = *static_cast<RefScene *>(h);
    LightSampleContext ctx;
    ctx.p = Vec3f(p[0], p[1], p[2]);
    ctx.n = ctx.ns = Vec3f(n[0], n[1], n[2]);
    return rs.lightOrder[size_t(light)]->PDF_Li(ctx, Vec3f(wi[0], wi[1], wi[2]));
}

void rf_world_sphere(void *h, float *out) {
    const Scene &sc = static_cast<RefScene *>(h)->scene;
    out[0] = sc.worldCenter.x;
    out[1] = sc.worldCenter.y;
    out[2] = sc.worldCenter.z;
    out[3] = sc.worldRadius;
}

}  // extern "C"
