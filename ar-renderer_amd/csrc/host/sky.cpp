// The Preetham, Shirley and Smits (1999) analytic daylight model as an equirectangular radiance
// map, following the reference's MakePreethamSky (src/lights/sky.cpp:131-208). The Perez
// distribution coefficients and the zenith luminance and chromaticity polynomials are the
// published ones (Preetham et al. 1999, appendix A.2), so they read as in the paper.
#include "scene_build.h"

#include <algorithm>
#include <cmath>

namespace phost {

namespace {

constexpr float kPiF = 3.14159265358979323846f;

// Perez et al. (1993) sky distribution F(theta, gamma); the zenith angle's cosine is kept at 0.01
// or above so the horizon stays finite.
struct PerezModel {
    float a, b, c, d, e;
    float operator()(float cosTheta, float gamma, float cosGamma) const {
        const float ct = std::max(cosTheta, 0.01f);
        return (1 + a * std::exp(b / ct)) * (1 + c * std::exp(d * gamma) + e * cosGamma * cosGamma);
    }
};

// Direction of equirectangular coordinates (u, v) in [0, 1]^2: v = 0 at +y (zenith), u around y
// from +x towards +z (src/lights/light.h:287-291).
F3 EquirectDirection(float u, float v) {
    const float theta = v * kPiF, phi = u * 2 * kPiF, st = std::sin(theta);
    return f3(st * std::cos(phi), std::cos(theta), st * std::sin(phi));
}

float Smooth(float x, float lo, float hi) {
    if (lo == hi) return x < lo ? 0.f : 1.f;
    const float t = std::clamp((x - lo) / (hi - lo), 0.f, 1.f);
    return t * t * (3 - 2 * t);
}

}  // namespace

void PreethamSky(F3 sunDirIn, float turbidity, float groundAlbedo, int width, int height, float *rgb) {
    const F3 sun = norm3(sunDirIn);
    const float T = std::clamp(turbidity, 1.7f, 10.f);
    // A sun below the horizon is evaluated at the horizon; the daylight then fades out.
    const float thetaS = std::acos(std::clamp(std::max(sun.y, 0.f), -1.f, 1.f));
    const PerezModel lumY{0.1787f * T - 1.4630f, -0.3554f * T + 0.4275f, -0.0227f * T + 5.3251f,
                          0.1206f * T - 2.5771f, -0.0670f * T + 0.3703f};
    const PerezModel chromX{-0.0193f * T - 0.2592f, -0.0665f * T + 0.0008f, -0.0004f * T + 0.2125f,
                            -0.0641f * T - 0.8989f, -0.0033f * T + 0.0452f};
    const PerezModel chromY{-0.0167f * T - 0.2608f, -0.0950f * T + 0.0092f, -0.0079f * T + 0.2102f,
                            -0.0441f * T - 1.6537f, -0.0109f * T + 0.0529f};
    // Zenith luminance (kcd/m^2) and chromaticity.
    const float chi = (4.f / 9.f - T / 120.f) * (kPiF - 2 * thetaS);
    const float zenithY = std::max(0.f, (4.0453f * T - 4.9710f) * std::tan(chi) - 0.2155f * T + 2.4192f);
    const float s1 = thetaS, s2 = s1 * s1, s3 = s2 * s1;
    const float zenithX = T * T * (0.00166f * s3 - 0.00375f * s2 + 0.00209f * s1) +
                          T * (-0.02903f * s3 + 0.06377f * s2 - 0.03202f * s1 + 0.00394f) +
                          (0.11693f * s3 - 0.21196f * s2 + 0.06052f * s1 + 0.25886f);
    const float zenithYc = T * T * (0.00275f * s3 - 0.00610f * s2 + 0.00317f * s1) +
                           T * (-0.04214f * s3 + 0.08970f * s2 - 0.04153f * s1 + 0.00516f) +
                           (0.15346f * s3 - 0.26756f * s2 + 0.06670f * s1 + 0.26688f);
    const float cosS = std::cos(thetaS);
    const float normY = lumY(1, thetaS, cosS), normX = chromX(1, thetaS, cosS), normYc = chromY(1, thetaS, cosS);
    const float daylight = Smooth(sun.y, -0.1f, 0.02f);
    const F3 sunAbove = norm3(f3(sun.x, std::max(sun.y, 0.f), sun.z));
    double skyIrradiance = 0;  // on a horizontal plane, kcd/m^2 * sr
    for (int j = 0; j < height; ++j)
        for (int i = 0; i < width; ++i) {
            float *px = rgb + 3 * (size_t(j) * size_t(width) + size_t(i));
            px[0] = px[1] = px[2] = 0;
            const F3 w = EquirectDirection((float(i) + 0.5f) / float(width), (float(j) + 0.5f) / float(height));
            if (!(w.y > 0)) continue;
            const float cosG = std::clamp(w.x * sunAbove.x + w.y * sunAbove.y + w.z * sunAbove.z, -1.f, 1.f);
            const float gamma = std::acos(cosG);
            const float Y = zenithY * lumY(w.y, gamma, cosG) / normY * daylight;
            const float x = zenithX * chromX(w.y, gamma, cosG) / normX;
            const float y = zenithYc * chromY(w.y, gamma, cosG) / normYc;
            if (!(y > 0 && Y > 0)) continue;
            XYZToSRGB(x / y * Y, Y, (1 - x - y) / y * Y, px);
            for (int k = 0; k < 3; ++k) px[k] = std::max(0.f, px[k]);
            const float solidAngle = (2 * kPiF / float(width)) * (kPiF / float(height)) *
                                     std::sqrt(std::max(0.f, 1 - w.y * w.y));
            skyIrradiance += Y * w.y * solidAngle;
        }
    // The ground below the horizon: a Lambertian of the given albedo under the sky and the sun
    // (whose luminance the sun-radiance spectrum gives, at the nominal 0.2665 degree radius).
    float sunIrradiance = 0;
    if (sun.y > 0) {
        SceneBuilder b;
        const float r = (kPiF / 180.f) * 0.2665f;
        sunIrradiance = b.SpectrumY(b.SunRadiance(sun, T, 0.2665f)) * kPiF * r * r * sun.y;
    }
    const float ground = groundAlbedo * (1 / kPiF) * (float(skyIrradiance) + sunIrradiance);
    for (int j = 0; j < height; ++j)
        for (int i = 0; i < width; ++i)
            if (EquirectDirection((float(i) + 0.5f) / float(width), (float(j) + 0.5f) / float(height)).y <= 0) {
                float *px = rgb + 3 * (size_t(j) * size_t(width) + size_t(i));
                px[0] = px[1] = px[2] = ground;
            }
}

}  // namespace phost
