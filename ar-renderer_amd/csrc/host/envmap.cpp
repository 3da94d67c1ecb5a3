// Sampling tables of an image environment light: the piecewise-constant 2D distribution over the
// equirectangular map weighted by luminance x sin(theta) + 1e-6 (src/lights/light.cpp:221-239),
// its continuous sampling and density (src/core/sampling.cpp:5-67). The tables are laid out as the
// device will read them from the scene's knots: per row the conditional CDF (nu + 1 floats), then
// the row integrals, then the marginal CDF (nv + 1 floats).
#include "scene_build.h"

#include <algorithm>
#include <cmath>

namespace phost {

namespace {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kBelowOne = 0x1.fffffep-1f;  // the largest float below 1

// Normalised CDF of a non-negative step function over n cells, accumulated in double; returns its
// integral over [0, 1]. A zero function gets the uniform CDF.
float BuildCdf(const float *f, int n, float *cdf) {
    double acc = 0;
    cdf[0] = 0;
    for (int i = 0; i < n; ++i) {
        acc += double(std::fabs(f[i])) / n;
        cdf[i + 1] = float(acc);
    }
    const float integral = float(acc);
    for (int i = 1; i <= n; ++i) cdf[i] = integral == 0 ? float(i) / float(n) : float(cdf[i] / acc);
    cdf[n] = 1;
    return integral;
}

// Largest i in [0, n - 1] with cdf[i] <= u (bisection over the n + 1 entries).
int FindCell(const float *cdf, int n, float u) {
    int first = 1, len = n - 1;
    while (len > 0) {
        const int half = len >> 1, mid = first + half;
        if (cdf[mid] <= u) {
            first = mid + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return std::clamp(first - 1, 0, n - 1);
}

float SampleCells(const float *f, const float *cdf, int n, float integral, float u, float *pdf, int *cell) {
    const int o = FindCell(cdf, n, u);
    float du = u - cdf[o];
    if (cdf[o + 1] - cdf[o] > 0) du /= cdf[o + 1] - cdf[o];
    *pdf = integral > 0 ? std::fabs(f[o]) / integral : 1.f;
    *cell = o;
    return std::min((float(o) + du) / float(n), kBelowOne);
}

}  // namespace

void EnvMapDistribution::Build(const float *rgb, int w, int h) {
    nu = w;
    nv = h;
    func.assign(size_t(w) * size_t(h), 0.f);
    double sumY = 0, sumW = 0;
    for (int y = 0; y < h; ++y) {
        const float sinTheta = std::sin(kPiF * (float(y) + 0.5f) / float(h));
        for (int x = 0; x < w; ++x) {
            const float *c = rgb + 3 * (size_t(y) * size_t(w) + size_t(x));
            const float lum = 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2];
            func[size_t(y) * size_t(w) + size_t(x)] = std::max(lum, 0.f) * sinTheta + 1e-6f;
            sumY += lum * sinTheta;
            sumW += sinTheta;
        }
    }
    avgY = float(sumY / std::max(1e-9, sumW));
    condCdf.assign(size_t(h) * size_t(w + 1), 0.f);
    rowInt.assign(size_t(h), 0.f);
    for (int y = 0; y < h; ++y)
        rowInt[size_t(y)] = BuildCdf(&func[size_t(y) * size_t(w)], w, &condCdf[size_t(y) * size_t(w + 1)]);
    margCdf.assign(size_t(h) + 1, 0.f);
    margInt = BuildCdf(rowInt.data(), h, margCdf.data());
}

void EnvMapDistribution::Sample(float u0, float u1, float *u, float *v, float *pdf) const {
    float pv, pu;
    int row, col;
    *v = SampleCells(rowInt.data(), margCdf.data(), nv, margInt, u1, &pv, &row);
    *u = SampleCells(&func[size_t(row) * size_t(nu)], &condCdf[size_t(row) * size_t(nu + 1)], nu,
                     rowInt[size_t(row)], u0, &pu, &col);
    *pdf = pu * pv;
}

float EnvMapDistribution::Pdf(float u, float v) const {
    const int iu = std::clamp(int(u * float(nu)), 0, nu - 1), iv = std::clamp(int(v * float(nv)), 0, nv - 1);
    if (margInt == 0) return 0;
    return func[size_t(iv) * size_t(nu) + size_t(iu)] / margInt;
}

}  // namespace phost
