// Sheen directional albedo table (host side of the plugin, compiled without FMA contraction).
//
// Reference src/materials/bxdf.cpp:535-563 (SheenBxDF::Albedo): for 16 roughnesses
// alpha_i = lerp(i / 15, 0.05, 1) and 32 cosines mu_j = max(1e-3, j / 31), the albedo of the
// Charlie lobe for wo = (sqrt(1 - mu^2), 0, mu) by a 32 x 32 stratified grid of cosine-weighted
// directions (concentric disk map), each term f * pi, summed in double. The kernels interpolate
// it bilinearly. The CPU path tracer builds the same table; prgpu_render takes the caller's.
#include <algorithm>
#include <cmath>

#include "sheen_table.h"

namespace pgpu {
namespace {

struct Vec {
    float x, y, z;
};

float Charlie(Vec wo, Vec wi, float alpha) {  // D(wh) / (4 (ci + co - ci co)), same hemisphere
    if (!(wo.z * wi.z > 0)) return 0;
    float co = std::abs(wo.z), ci = std::abs(wi.z);
    if (co == 0 || ci == 0) return 0;
    Vec h{wo.x + wi.x, wo.y + wi.y, wo.z + wi.z};
    float inv = 1.f / std::sqrt(h.x * h.x + h.y * h.y + h.z * h.z);
    float c = std::abs(h.z * inv), ia = 1 / alpha;
    float sin2 = std::max(0.f, 1 - c * c);
    float D = (2 + ia) * std::pow(sin2, 0.5f * ia) * 0.15915494309189533577f;
    return D / (4 * (ci + co - ci * co));
}

Vec CosineDirection(float ux, float uy) {
    float ox = 2 * ux - 1, oy = 2 * uy - 1, dx = 0, dy = 0;
    if (ox != 0 || oy != 0) {
        float r, theta;
        if (std::abs(ox) > std::abs(oy)) {
            r = ox;
            theta = 0.78539816339744830961f * (oy / ox);
        } else {
            r = oy;
            theta = 1.57079632679489661923f - 0.78539816339744830961f * (ox / oy);
        }
        dx = r * std::cos(theta);
        dy = r * std::sin(theta);
    }
    return {dx, dy, std::sqrt(std::max(0.f, 1 - dx * dx - dy * dy))};
}

}  // namespace

void SheenAlbedoTable(float *out) {
    constexpr int kMu = 32, kAlpha = 16, kN = 32;
    for (int ia = 0; ia < kAlpha; ++ia) {
        float t = float(ia) / (kAlpha - 1);
        float a = (1 - t) * 0.05f + t * 1.f;
        for (int im = 0; im < kMu; ++im) {
            float mu = std::max(1e-3f, float(im) / (kMu - 1));
            Vec wo{std::sqrt(std::max(0.f, 1 - mu * mu)), 0, mu};
            double sum = 0;
            for (int i = 0; i < kN; ++i)
                for (int j = 0; j < kN; ++j)
                    sum += Charlie(wo, CosineDirection((i + 0.5f) / kN, (j + 0.5f) / kN), a) * 3.14159265358979323846f;
            out[ia * kMu + im] = float(sum / (kN * kN));
        }
    }
}

}  // namespace pgpu
