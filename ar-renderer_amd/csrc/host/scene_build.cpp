#include "scene_build.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <atomic>
#include <map>
#include <stdexcept>
#include <thread>

namespace phost {

// ---------------------------------------------------------------------------------------------
// Random numbers for scene generation (PCG32, src/core/rng.h:19-45)

static uint64_t MixBits64(uint64_t v) {
    v ^= (v >> 31);
    v *= 0x7fb5d329728ea185ULL;
    v ^= (v >> 27);
    v *= 0x81dadef4bc2dd44dULL;
    v ^= (v >> 33);
    return v;
}

SceneRng::SceneRng(uint64_t seq) {
    state = 0;
    inc = (seq << 1u) | 1u;
    next();
    state += MixBits64(seq);
    next();
}
uint32_t SceneRng::next() {
    uint64_t old = state;
    state = old * 0x5851f42d4c957f2dULL + inc;
    uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31));
}
float SceneRng::uniform() { return std::min(0x1.fffffep-1f, float(next()) * 0x1p-32f); }

// ---------------------------------------------------------------------------------------------
// Spectra

static const float kPi = 3.14159265358979323846f;

static float Planck(float lambdaNm, float T) {  // src/core/spectrum.cpp:116-122
    if (T <= 0) return 0;
    const double c = 299792458., h = 6.62606957e-34, kb = 1.3806488e-23;
    double l = lambdaNm * 1e-9;
    return float((2 * h * c * c) / (std::pow(l, 5) * (std::exp((h * c) / (l * kb * T)) - 1)));
}

static float CieG(float x, float mu, float s1, float s2) {
    float t = (x - mu) / (x < mu ? s1 : s2);
    return std::exp(-0.5f * t * t);
}
static float CieX(float l) {
    return 1.056f * CieG(l, 599.8f, 37.9f, 31.0f) + 0.362f * CieG(l, 442.0f, 16.0f, 26.7f) -
           0.065f * CieG(l, 501.1f, 20.4f, 26.2f);
}
static float CieY(float l) { return 0.821f * CieG(l, 568.8f, 46.9f, 40.5f) + 0.286f * CieG(l, 530.9f, 16.3f, 31.1f); }
static float CieZ(float l) { return 1.217f * CieG(l, 437.0f, 11.8f, 36.0f) + 0.681f * CieG(l, 459.0f, 26.0f, 13.8f); }
static float CieYIntegral() {
    static const float v = [] {
        double s = 0;
        for (int l = 360; l <= 830; ++l) s += CieY(float(l));
        return float(s);
    }();
    return v;
}

template <class P> static int FindIntervalH(int size, const P &pred) {
    int first = 1, len = size - 2;
    while (len > 0) {
        int half = len >> 1, mid = first + half;
        bool r = pred(mid);
        first = r ? mid + 1 : first;
        len = r ? len - (half + 1) : half;
    }
    return std::clamp(first - 1, 0, size - 2);
}

float SpectrumValue(const prgpu::SpectrumEx &s, const float *knots, float l) {
    float b = 0;
    switch (s.kind) {
    case prgpu::SpecConstant: b = s.p[0]; break;
    case prgpu::SpecSigmoid: {
        float x = std::fma(l, std::fma(l, s.p[0], s.p[1]), s.p[2]);
        b = std::isinf(x) ? (x > 0 ? 1.f : 0.f) : 0.5f + x / (2 * std::sqrt(1 + x * x));
        break;
    }
    case prgpu::SpecBlackbody: b = Planck(l, s.p[0]) * s.p[1]; break;
    case prgpu::SpecPiecewise: {
        const float *L = knots + s.offset, *V = L + s.count;
        int n = s.count;
        if (l <= L[0]) b = V[0];
        else if (l >= L[n - 1]) b = V[n - 1];
        else {
            int o = FindIntervalH(n, [&](int i) { return L[i] <= l; });
            float t = (l - L[o]) / (L[o + 1] - L[o]);
            b = (1 - t) * V[o] + t * V[o + 1];
        }
        break;
    }
    case prgpu::SpecSellmeier: {
        double l2 = (double(l) * 1e-3) * (double(l) * 1e-3);
        double n2 = 1;
        for (int i = 0; i < s.count; ++i) n2 += s.p[i] * l2 / (l2 - s.p[3 + i]);
        b = float(std::sqrt(std::max(n2, 1.0)));
        break;
    }
    case prgpu::SpecCauchy: b = s.p[0] + s.p[1] / (l * l); break;
    case prgpu::SpecTable95: {
        const float *v = knots + s.offset;
        float x = std::clamp((l - 360.f) / 5.f, 0.f, 94.f - 1e-3f);
        int i = int(x);
        float t = x - float(i);
        b = v[i] * (1 - t) + v[i + 1] * t;
        break;
    }
    }
    return s.scale == 1.f ? b : s.scale * b;
}

static void SpectrumXYZ(const prgpu::SpectrumEx &s, const float *knots, float xyz[3]) {  // spectrum.cpp:82-94
    double X = 0, Y = 0, Z = 0;
    for (int l = 360; l <= 830; ++l) {
        float v = SpectrumValue(s, knots, float(l));
        X += CieX(float(l)) * v;
        Y += CieY(float(l)) * v;
        Z += CieZ(float(l)) * v;
    }
    float inv = 1.f / CieYIntegral();
    xyz[0] = float(X) * inv;
    xyz[1] = float(Y) * inv;
    xyz[2] = float(Z) * inv;
}

// CIE D65 relative power, 300..830 nm every 10 nm (the standard illuminant table, as in
// src/core/spectrum.cpp:140-147), normalised below to unit luminance.
static const float kD65[] = {
    0.0341f,  3.2945f,  20.236f,  37.0535f, 39.9488f, 44.9117f, 46.6383f, 52.0891f, 49.9755f,
    54.6482f, 82.7549f, 91.486f,  93.4318f, 86.6823f, 104.865f, 117.008f, 117.812f, 114.861f,
    115.923f, 108.811f, 109.354f, 107.802f, 104.79f,  107.689f, 104.405f, 104.046f, 100.0f,
    96.3342f, 95.788f,  88.6856f, 90.0062f, 89.5991f, 87.6987f, 83.2886f, 83.6992f, 80.0268f,
    80.2146f, 82.2778f, 78.2842f, 69.7213f, 71.6091f, 74.349f,  61.604f,  69.8856f, 75.087f,
    63.5927f, 46.4182f, 66.8054f, 63.3828f, 64.304f,  59.4519f, 51.959f,  57.4406f, 60.3125f};

struct D65Table {
    std::vector<float> l, v;
    D65Table() {
        for (size_t i = 0; i < sizeof(kD65) / sizeof(float); ++i) {
            l.push_back(300.f + 10.f * float(i));
            v.push_back(kD65[i]);
        }
        prgpu::SpectrumEx s{};
        s.kind = prgpu::SpecPiecewise;
        s.scale = 1;
        s.offset = 0;
        s.count = int(l.size());
        std::vector<float> k = l;
        k.insert(k.end(), v.begin(), v.end());
        float xyz[3];
        SpectrumXYZ(s, k.data(), xyz);
        for (float &x : v) x /= xyz[1];
    }
    float operator()(float lambda) const {
        if (lambda <= l.front()) return v.front();
        if (lambda >= l.back()) return v.back();
        int o = FindIntervalH(int(l.size()), [&](int i) { return l[size_t(i)] <= lambda; });
        float t = (lambda - l[size_t(o)]) / (l[size_t(o + 1)] - l[size_t(o)]);
        return (1 - t) * v[size_t(o)] + t * v[size_t(o + 1)];
    }
};
static const D65Table &D65() {
    static const D65Table t;
    return t;
}

// 3x3 float matrices as the reference colour code uses them (src/core/color.h:20-30).
struct M3 {
    float m[3][3];
    void apply(float x, float y, float z, float out[3]) const {
        out[0] = m[0][0] * x + m[0][1] * y + m[0][2] * z;
        out[1] = m[1][0] * x + m[1][1] * y + m[1][2] * z;
        out[2] = m[2][0] * x + m[2][1] * y + m[2][2] * z;
    }
    M3 operator*(const M3 &b) const {
        M3 r{};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += double(m[i][k]) * b.m[k][j];
                r.m[i][j] = float(s);
            }
        return r;
    }
    M3 inverse() const {
        double a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0], h = m[2][1],
               i = m[2][2];
        double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
        double inv = 1.0 / (a * A + b * B + c * C);
        M3 r{};
        r.m[0][0] = float(A * inv);
        r.m[0][1] = float(-(b * i - c * h) * inv);
        r.m[0][2] = float((b * f - c * e) * inv);
        r.m[1][0] = float(B * inv);
        r.m[1][1] = float((a * i - c * g) * inv);
        r.m[1][2] = float(-(a * f - c * d) * inv);
        r.m[2][0] = float(C * inv);
        r.m[2][1] = float(-(a * h - b * g) * inv);
        r.m[2][2] = float((a * e - b * d) * inv);
        return r;
    }
};
// sRGB primaries, D65 (IEC 61966-2-1).
static const M3 kXYZToSRGB = {{{3.2404542f, -1.5371385f, -0.4985314f},
                               {-0.9692660f, 1.8760108f, 0.0415560f},
                               {0.0556434f, -0.2040259f, 1.0572252f}}};
static const M3 kSRGBToXYZ = {{{0.4124564f, 0.3575761f, 0.1804375f},
                               {0.2126729f, 0.7151522f, 0.0721750f},
                               {0.0193339f, 0.1191920f, 0.9503041f}}};

// Bradford adaptation from the spectral D65 white seen through the analytic CMFs to the exact
// CIE D65 white (src/core/color.cpp:85-101).
static const M3 &WhiteAdaptation() {
    static const M3 w = [] {
        const D65Table &d = D65();
        double X = 0, Y = 0, Z = 0;
        for (int l = 360; l <= 830; ++l) {
            float v = d(float(l));
            X += CieX(float(l)) * v;
            Y += CieY(float(l)) * v;
            Z += CieZ(float(l)) * v;
        }
        float inv = 1.f / CieYIntegral();
        float src[3] = {float(X) * inv, float(Y) * inv, float(Z) * inv};
        float dst[3] = {0.95047f, 1.f, 1.08883f};
        static const M3 B = {{{0.8951f, 0.2664f, -0.1614f}, {-0.7502f, 1.7135f, 0.0367f}, {0.0389f, -0.0685f, 1.0296f}}};
        M3 Binv = B.inverse();
        float s[3], t[3];
        B.apply(src[0], src[1], src[2], s);
        B.apply(dst[0], dst[1], dst[2], t);
        M3 D = {{{t[0] / s[0], 0, 0}, {0, t[1] / s[1], 0}, {0, 0, t[2] / s[2]}}};
        return Binv * (D * B);
    }();
    return w;
}

void XYZToSRGB(float X, float Y, float Z, float rgb[3]) { kXYZToSRGB.apply(X, Y, Z, rgb); }

void XYZToLinearSRGB(const double xyz[3], double scale, float rgb[3]) {
    float X = float(xyz[0] * scale), Y = float(xyz[1] * scale), Z = float(xyz[2] * scale);
    float a[3];
    WhiteAdaptation().apply(X, Y, Z, a);
    kXYZToSRGB.apply(a[0], a[1], a[2], rgb);
}

// ---- RGB uplift: Gauss-Newton fit of a sigmoid polynomial in CIELAB ---------------------------
namespace {
constexpr int kFine = 189;  // 2.5 nm over [360, 830]
struct FitData {
    double lambda[kFine];
    double rgbw[3][kFine];
    double toXyz[3][3];
    double white[3];
};
const FitData &Fit() {
    static const FitData f = [] {
        FitData t{};
        M3 toRGB = kXYZToSRGB * WhiteAdaptation();
        double h = (830.0 - 360.0) / (kFine - 1);
        double norm = 1.0 / CieYIntegral();
        for (int i = 0; i < kFine; ++i) {
            double l = 360.0 + i * h;
            double w = (i == 0 || i == kFine - 1) ? 0.5 * h : h;
            double I = D65()(float(l));
            double xyz[3] = {CieX(float(l)), CieY(float(l)), CieZ(float(l))};
            t.lambda[i] = l;
            for (int k = 0; k < 3; ++k) {
                double s = 0;
                for (int j = 0; j < 3; ++j) s += toRGB.m[k][j] * xyz[j];
                t.rgbw[k][i] = s * I * w * norm;
            }
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) t.toXyz[i][j] = kSRGBToXYZ.m[i][j];
        for (int i = 0; i < 3; ++i) t.white[i] = t.toXyz[i][0] + t.toXyz[i][1] + t.toXyz[i][2];
        return t;
    }();
    return f;
}
void Lab(const FitData &t, double *p) {
    double X = 0, Y = 0, Z = 0;
    for (int j = 0; j < 3; ++j) {
        X += p[j] * t.toXyz[0][j];
        Y += p[j] * t.toXyz[1][j];
        Z += p[j] * t.toXyz[2][j];
    }
    auto f = [](double v) {
        const double d = 6.0 / 29.0;
        return v > d * d * d ? std::cbrt(v) : v / (d * d * 3.0) + (4.0 / 29.0);
    };
    double fx = f(X / t.white[0]), fy = f(Y / t.white[1]), fz = f(Z / t.white[2]);
    p[0] = 116.0 * fy - 16.0;
    p[1] = 500.0 * (fx - fy);
    p[2] = 200.0 * (fy - fz);
}
void Residual(const FitData &t, const double *c, const double *rgb, double *res) {
    double out[3] = {0, 0, 0};
    for (int i = 0; i < kFine; ++i) {
        double x = (t.lambda[i] - 360.0) / (830.0 - 360.0);
        double v = (c[0] * x + c[1]) * x + c[2];
        double s = 0.5 * v / std::sqrt(1.0 + v * v) + 0.5;
        for (int j = 0; j < 3; ++j) out[j] += t.rgbw[j][i] * s;
    }
    Lab(t, out);
    double target[3] = {rgb[0], rgb[1], rgb[2]};
    Lab(t, target);
    for (int j = 0; j < 3; ++j) res[j] = target[j] - out[j];
}
bool Solve3(double A[3][3], const double b[3], double x[3]) {
    double M[3][4];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) M[i][j] = A[i][j];
        M[i][3] = b[i];
    }
    for (int c = 0; c < 3; ++c) {
        int piv = c;
        for (int r = c + 1; r < 3; ++r)
            if (std::abs(M[r][c]) > std::abs(M[piv][c])) piv = r;
        if (std::abs(M[piv][c]) < 1e-15) return false;
        for (int k = 0; k < 4; ++k) std::swap(M[c][k], M[piv][k]);
        for (int r = c + 1; r < 3; ++r) {
            double f = M[r][c] / M[c][c];
            for (int k = c; k < 4; ++k) M[r][k] -= f * M[c][k];
        }
    }
    for (int i = 2; i >= 0; --i) {
        double s = M[i][3];
        for (int k = i + 1; k < 3; ++k) s -= M[i][k] * x[k];
        x[i] = s / M[i][i];
    }
    return true;
}
void GaussNewton(const FitData &t, const double rgb[3], double c[3], int iters) {
    for (int it = 0; it < iters; ++it) {
        double r[3], J[3][3], dx[3];
        Residual(t, c, rgb, r);
        for (int i = 0; i < 3; ++i) {
            double a[3] = {c[0], c[1], c[2]}, r0[3], r1[3];
            a[i] = c[i] - 1e-5;
            Residual(t, a, rgb, r0);
            a[i] = c[i] + 1e-5;
            Residual(t, a, rgb, r1);
            for (int j = 0; j < 3; ++j) J[j][i] = (r1[j] - r0[j]) / 2e-5;
        }
        if (!Solve3(J, r, dx)) break;
        double e = 0;
        for (int j = 0; j < 3; ++j) {
            c[j] -= dx[j];
            e += r[j] * r[j];
        }
        double mx = std::max(std::max(c[0], c[1]), c[2]);
        if (mx > 200)
            for (int j = 0; j < 3; ++j) c[j] *= 200 / mx;
        if (e < 1e-6) break;
    }
}
}  // namespace

// The table's nodes, as the reference's generator places and fits them (rgb2spec.cpp:243-276):
// z nodes smoothstep(smoothstep(k / (res - 1))); for maximum component l and the other two at
// x, y in [0, 1] of it, each (l, y, x) column of z is fitted from z index res / 5 upwards and then
// downwards, every fit starting from the previous one's coefficients.
const std::vector<float> &RgbToSpectrumTableRecord() {
    static const std::vector<float> rec = [] {
        const int R = prgpu::kRgbTableRes;
        std::vector<float> out(1 + size_t(R) + size_t(3) * R * R * R * 3, 0.f);
        out[0] = float(R);
        float *z = out.data() + 1, *co = z + R;
        auto smoothstep = [](double x) { return x * x * (3.0 - 2.0 * x); };
        for (int k = 0; k < R; ++k) z[k] = float(smoothstep(smoothstep(double(k) / (R - 1))));
        const FitData &t = Fit();
        const double c0 = 360.0, c1 = 1.0 / (830.0 - 360.0);
        auto column = [&](int l, int j, int i) {
            const double y = double(j) / (R - 1), x = double(i) / (R - 1);
            auto fitRange = [&](int kStart, int kEnd, int step) {
                double c[3] = {0, 0, 0};
                for (int k = kStart; k != kEnd; k += step) {
                    const double b = z[k];
                    double rgb[3];
                    rgb[l] = b;
                    rgb[(l + 1) % 3] = x * b;
                    rgb[(l + 2) % 3] = y * b;
                    GaussNewton(t, rgb, c, 15);
                    const size_t idx = ((size_t(l) * R + size_t(k)) * R + size_t(j)) * R + size_t(i);
                    const double A = c[0], B = c[1], C = c[2];
                    co[3 * idx + 0] = float(A * c1 * c1);
                    co[3 * idx + 1] = float(B * c1 - 2 * A * c0 * c1 * c1);
                    co[3 * idx + 2] = float(C - B * c0 * c1 + A * c0 * c0 * c1 * c1);
                }
            };
            fitRange(R / 5, R, 1);
            fitRange(R / 5, -1, -1);
        };
        const int jobs = 3 * R;
        std::atomic<int> next{0};
        auto worker = [&] {
            for (int lj = next++; lj < jobs; lj = next++)
                for (int i = 0; i < R; ++i) column(lj / R, lj % R, i);
        };
        const unsigned n = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned k = 1; k < n; ++k) pool.emplace_back(worker);
        worker();
        for (std::thread &th : pool) th.join();
        return out;
    }();
    return rec;
}

int RgbFitTables(double *rgbw, double *toXyz, double *white) {
    const FitData &t = Fit();
    for (int k = 0; k < 3; ++k) {
        if (rgbw)
            for (int i = 0; i < kFine; ++i) rgbw[k * kFine + i] = t.rgbw[k][i];
        if (white) white[k] = t.white[k];
        if (toXyz)
            for (int j = 0; j < 3; ++j) toXyz[3 * k + j] = t.toXyz[k][j];
    }
    return kFine;
}

// RGBToSpectrumTable::operator() (src/core/rgb2spec.cpp:278-304): the colour clamped to [0, 1];
// grey in closed form; otherwise the maximum component picks the table, the other two give its
// (x, y) cell at (res - 1) / max, FindInterval over the z nodes gives the z cell, and the node
// coefficients are interpolated trilinearly.
void RgbTableCoeffs(const float rgbIn[3], float c[3]) {
    float rgb[3];
    for (int k = 0; k < 3; ++k) rgb[k] = std::clamp(rgbIn[k], 0.f, 1.f);
    if (rgb[0] == rgb[1] && rgb[1] == rgb[2]) {
        const float v = rgb[0];
        c[0] = c[1] = 0;
        c[2] = v <= 0 ? -INFINITY : (v >= 1 ? INFINITY : (v - .5f) / std::sqrt(v * (1 - v)));
        return;
    }
    const std::vector<float> &t = RgbToSpectrumTableRecord();
    const int n = int(t[0]);
    const float *zn = t.data() + 1, *co = zn + n;
    const int maxc = (rgb[0] > rgb[1]) ? ((rgb[0] > rgb[2]) ? 0 : 2) : ((rgb[1] > rgb[2]) ? 1 : 2);
    const float z = rgb[maxc];
    const float x = rgb[(maxc + 1) % 3] * float(n - 1) / z;
    const float y = rgb[(maxc + 2) % 3] * float(n - 1) / z;
    const int xi = std::min(int(x), n - 2), yi = std::min(int(y), n - 2);
    int first = 1, len = n - 2;  // FindInterval(n, z nodes[i] < z)
    while (len > 0) {
        const int half = len >> 1, middle = first + half;
        const bool below = zn[middle] < z;
        first = below ? middle + 1 : first;
        len = below ? len - (half + 1) : half;
    }
    const int zi = std::clamp(first - 1, 0, n - 2);
    const float dx = x - float(xi), dy = y - float(yi), dz = (z - zn[zi]) / (zn[zi + 1] - zn[zi]);
    auto at = [&](int zz, int yy, int xx, int i) {
        return co[3 * (((size_t(maxc) * size_t(n) + size_t(zz)) * size_t(n) + size_t(yy)) * size_t(n) + size_t(xx)) + size_t(i)];
    };
    auto lerp = [](float a, float u, float v) { return (1 - a) * u + a * v; };
    for (int i = 0; i < 3; ++i)
        c[i] = lerp(dz, lerp(dy, lerp(dx, at(zi, yi, xi, i), at(zi, yi, xi + 1, i)),
                             lerp(dx, at(zi, yi + 1, xi, i), at(zi, yi + 1, xi + 1, i))),
                    lerp(dy, lerp(dx, at(zi + 1, yi, xi, i), at(zi + 1, yi, xi + 1, i)),
                         lerp(dx, at(zi + 1, yi + 1, xi, i), at(zi + 1, yi + 1, xi + 1, i))));
}

// ---------------------------------------------------------------------------------------------
// SceneBuilder

int SceneBuilder::AddSpectrum(const prgpu::SpectrumEx &s) {
    spectra.push_back(s);
    return int(spectra.size() - 1);
}

int SceneBuilder::Constant(float v) {
    prgpu::SpectrumEx s{};
    s.kind = prgpu::SpecConstant;
    s.scale = 1;
    s.p[0] = v;
    return AddSpectrum(s);
}

int SceneBuilder::Sigmoid(float c0, float c1, float c2, float scale) {
    prgpu::SpectrumEx s{};
    s.kind = prgpu::SpecSigmoid;
    s.scale = scale;
    s.p[0] = c0;
    s.p[1] = c1;
    s.p[2] = c2;
    return AddSpectrum(s);
}

int SceneBuilder::RGBAlbedo(float r, float g, float b) {  // RGBAlbedoSpectrum (rgb2spec.cpp:306)
    const float rgb[3] = {r, g, b};
    float c[3];
    RgbTableCoeffs(rgb, c);
    return Sigmoid(c[0], c[1], c[2]);
}

int SceneBuilder::RGBUnbounded(float r, float g, float b) {  // RGBUnboundedSpectrum (rgb2spec.cpp:309-313)
    const float m = std::max(0.f, std::max(r, std::max(g, b)));
    const float scale = 2 * m;
    const float inv = scale > 0 ? 1 / scale : 0.f;
    const float rgb[3] = {r * inv, g * inv, b * inv};
    float c[3];
    RgbTableCoeffs(rgb, c);
    return Sigmoid(c[0], c[1], c[2], scale);
}

int SceneBuilder::Scaled(int id, float f) {
    prgpu::SpectrumEx s = spectra[size_t(id)];
    s.scale *= f;
    return AddSpectrum(s);
}

int SceneBuilder::BlackbodyUnitY(float T) {
    // BlackbodySpectrum(T) normalised to peak 1, then NormalizedToY (scene_loader.cpp:197-201).
    prgpu::SpectrumEx s{};
    s.kind = prgpu::SpecBlackbody;
    s.scale = 1;
    s.p[0] = T;
    float lambdaMax = 2.8977721e-3f / T * 1e9f;
    s.p[1] = 1 / Planck(lambdaMax, T);
    float xyz[3];
    SpectrumXYZ(s, knots.data(), xyz);
    if (xyz[1] > 0) s.scale = 1 / xyz[1];
    return AddSpectrum(s);
}

int SceneBuilder::Piecewise(const std::vector<float> &l, const std::vector<float> &v, float scale) {
    prgpu::SpectrumEx s{};
    s.kind = prgpu::SpecPiecewise;
    s.scale = scale;
    s.offset = int(knots.size());
    s.count = int(l.size());
    knots.insert(knots.end(), l.begin(), l.end());
    knots.insert(knots.end(), v.begin(), v.end());
    return AddSpectrum(s);
}

int SceneBuilder::Table95(const std::vector<float> &values) {
    prgpu::SpectrumEx s{};
    s.kind = prgpu::SpecTable95;
    s.scale = 1;
    s.offset = int(knots.size());
    s.count = 95;
    knots.insert(knots.end(), values.begin(), values.begin() + 95);
    return AddSpectrum(s);
}

int SceneBuilder::D65Illuminant() {
    if (d65_ < 0) d65_ = Piecewise(D65().l, D65().v);
    return d65_;
}

int SceneBuilder::Cauchy(float A, float B) {
    prgpu::SpectrumEx s{};
    s.kind = prgpu::SpecCauchy;
    s.scale = 1;
    s.p[0] = A;
    s.p[1] = B;
    return AddSpectrum(s);
}

// Measured optical constants sampled across the visible range (the reference's named metals,
// src/core/spectrum.cpp:164-189).
int SceneBuilder::Metal(const std::string &name, bool k) {
    static const std::vector<float> L = {360, 400, 450, 500, 550, 600, 650, 700, 750, 830};
    static const std::map<std::string, std::pair<std::vector<float>, std::vector<float>>> tab = {
        {"Au", {{1.72f, 1.66f, 1.40f, 0.97f, 0.43f, 0.25f, 0.17f, 0.16f, 0.16f, 0.17f},
                {1.92f, 1.96f, 1.88f, 1.87f, 2.46f, 2.98f, 3.50f, 3.95f, 4.40f, 5.10f}}},
        {"Cu", {{1.26f, 1.18f, 1.17f, 1.13f, 1.03f, 0.27f, 0.21f, 0.21f, 0.23f, 0.27f},
                {2.30f, 2.40f, 2.53f, 2.56f, 2.58f, 3.24f, 3.67f, 4.18f, 4.60f, 5.20f}}},
        {"Ag", {{0.07f, 0.05f, 0.04f, 0.05f, 0.06f, 0.06f, 0.07f, 0.08f, 0.09f, 0.10f},
                {1.60f, 2.10f, 2.66f, 3.13f, 3.59f, 4.00f, 4.44f, 4.80f, 5.20f, 5.80f}}},
        {"Al", {{0.40f, 0.49f, 0.62f, 0.77f, 0.96f, 1.20f, 1.47f, 1.83f, 2.40f, 2.70f},
                {4.35f, 4.86f, 5.47f, 6.08f, 6.69f, 7.26f, 7.79f, 8.31f, 8.62f, 8.30f}}},
    };
    auto it = tab.find(name);
    if (it == tab.end()) return -1;
    return Piecewise(L, k ? it->second.second : it->second.first);
}

int SceneBuilder::Sellmeier(const float *B, const float *C, int n) {
    prgpu::SpectrumEx s{};
    s.kind = prgpu::SpecSellmeier;
    s.scale = 1;
    s.count = n;
    for (int i = 0; i < n; ++i) {
        s.p[i] = B[i];
        s.p[3 + i] = C[i];
    }
    return AddSpectrum(s);
}

// Sellmeier coefficients of the named glasses (src/core/spectrum.cpp:191-201; standard catalogue values).
int SceneBuilder::Glass(const std::string &name) {
    struct G {
        const char *n;
        float B[3], C[3];
        int terms;
    };
    static const G glasses[] = {
        {"BK7", {1.03961212f, 0.231792344f, 1.01046945f}, {0.00600069867f, 0.0200179144f, 103.560653f}, 3},
        {"SF11", {1.73759695f, 0.313747346f, 1.89878101f}, {0.013188707f, 0.0623068142f, 155.23629f}, 3},
        {"fused_silica", {0.6961663f, 0.4079426f, 0.8974794f}, {0.00467914826f, 0.0135120631f, 97.9340025f}, 3},
        {"diamond", {0.3306f, 4.3356f, 0}, {0.030625f, 0.011236f, 0}, 2},
        {"sapphire", {1.4313493f, 0.65054713f, 5.3414021f}, {0.00527993f, 0.0142383f, 325.017834f}, 3},
    };
    for (const G &g : glasses)
        if (name == g.n) return Sellmeier(g.B, g.C, g.terms);
    return -1;
}

int SceneBuilder::Diffuse(int reflectance) {
    prgpu::Material m{};
    m.type = prgpu::MatDiffuse;
    m.albedo = reflectance;
    m.albedoB = m.eta = m.k = m.coatEta = m.coatAlbedo = m.sheenColor = -1;
    materials.push_back(m);
    return int(materials.size() - 1);
}

int SceneBuilder::Conductor(int eta, int k, float ax, float ay) {
    prgpu::Material m{};
    m.type = prgpu::MatConductor;
    m.albedo = m.albedoB = m.coatEta = m.coatAlbedo = m.sheenColor = -1;
    m.eta = eta;
    m.k = k;
    m.alphaX = ax;
    m.alphaY = ay;
    materials.push_back(m);
    return int(materials.size() - 1);
}

int SceneBuilder::ConductorReflectance(int refl, float ax, float ay) {
    int id = Conductor(-1, -1, ax, ay);
    materials[size_t(id)].albedo = refl;
    return id;
}

static bool ConstantSpectrum(const prgpu::SpectrumEx &s) {
    return s.kind == prgpu::SpecConstant || (s.kind == prgpu::SpecCauchy && s.p[1] == 0);
}

int SceneBuilder::Dielectric(int eta, float ax, float ay, bool thin) {
    prgpu::Material m{};
    m.type = prgpu::MatDielectric;
    m.albedo = m.albedoB = m.k = m.coatEta = m.coatAlbedo = m.sheenColor = -1;
    m.eta = eta;
    m.alphaX = ax;
    m.alphaY = ay;
    m.thin = thin ? 1 : 0;
    const prgpu::SpectrumEx &s = spectra[size_t(eta)];
    m.dispersive = ConstantSpectrum(s) ? 0 : 1;
    materials.push_back(m);
    return int(materials.size() - 1);
}


static void SetCoat(prgpu::Material &m, const SceneBuilder::Coat &c, const std::vector<prgpu::SpectrumEx> &spectra) {
    m.coatEta = c.eta;
    float r = std::min(std::max(c.roughness, 0.f), 1.f);
    m.coatAlpha = r * r;
    m.thickness = c.thickness;
    m.coatAlbedo = c.albedo;
    m.coatG = c.g;
    m.coatMaxDepth = c.maxDepth;
    m.coatSamples = c.nSamples;
    m.coatDispersive = ConstantSpectrum(spectra[size_t(c.eta)]) ? 0 : 1;
}

int SceneBuilder::CoatedDiffuse(const Coat &coat, int reflectance) {
    int id = Diffuse(reflectance);
    prgpu::Material &m = materials[size_t(id)];
    m.type = prgpu::MatCoatedDiffuse;
    SetCoat(m, coat, spectra);
    return id;
}

int SceneBuilder::CoatedConductor(const Coat &coat, int eta, int k, float ax, float ay) {
    int id = Conductor(eta, k, ax, ay);
    prgpu::Material &m = materials[size_t(id)];
    m.type = prgpu::MatCoatedConductor;
    SetCoat(m, coat, spectra);
    return id;
}

int SceneBuilder::CoatedConductorReflectance(const Coat &coat, int refl, float ax, float ay) {
    int id = CoatedConductor(coat, -1, -1, ax, ay);
    materials[size_t(id)].albedo = refl;
    return id;
}

void SceneBuilder::SetSheen(int material, int color, float roughness, float weight) {
    prgpu::Material &m = materials[size_t(material)];
    m.sheenColor = color;
    m.sheenRoughness = roughness;
    m.sheenWeight = weight;
}

int SceneBuilder::Interface() {
    prgpu::Material m{};
    m.type = prgpu::MatInterface;
    m.albedo = m.albedoB = m.eta = m.k = m.coatEta = m.coatAlbedo = m.sheenColor = -1;
    materials.push_back(m);
    return int(materials.size() - 1);
}

int SceneBuilder::AddMesh(const Mesh &mesh, int material, int mIn, int mOut, bool normals, bool wrapU) {
    int first = int(prims.size());
    for (size_t t = 0; t + 2 < mesh.idx.size(); t += 3) {
        prgpu::Primitive p{};
        p.type = prgpu::PrimTriangle;
        p.material = material;
        p.light = -1;
        p.mediumInside = mIn;
        p.mediumOutside = mOut;
        int a = mesh.idx[t], b = mesh.idx[t + 1], c = mesh.idx[t + 2];
        p.p0 = pf(mesh.p[size_t(a)]);
        p.p1 = pf(mesh.p[size_t(b)]);
        p.p2 = pf(mesh.p[size_t(c)]);
        if (normals && !mesh.n.empty()) {
            p.n0 = pf(mesh.n[size_t(a)]);
            p.n1 = pf(mesh.n[size_t(b)]);
            p.n2 = pf(mesh.n[size_t(c)]);
        }
        if (!mesh.uv.empty()) {
            p.u0 = mesh.uv[2 * size_t(a)];
            p.v0 = mesh.uv[2 * size_t(a) + 1];
            p.u1 = mesh.uv[2 * size_t(b)];
            p.v1 = mesh.uv[2 * size_t(b) + 1];
            p.u2 = mesh.uv[2 * size_t(c)];
            p.v2 = mesh.uv[2 * size_t(c) + 1];
            // A triangle that straddles the u seam of a wrapped parameterisation keeps a
            // continuous u (its tangent frame stays smooth).
            float lo = std::min(p.u0, std::min(p.u1, p.u2)), hi = std::max(p.u0, std::max(p.u1, p.u2));
            if (wrapU && hi - lo > 0.5f) {
                if (p.u0 < 0.5f) p.u0 += 1;
                if (p.u1 < 0.5f) p.u1 += 1;
                if (p.u2 < 0.5f) p.u2 += 1;
            }
        } else {  // the reference's default parameterisation (shape.cpp:231-236)
            p.u0 = 0; p.v0 = 0; p.u1 = 1; p.v1 = 0; p.u2 = 1; p.v2 = 1;
        }
        prims.push_back(p);
    }
    return first;
}

int SceneBuilder::AddSphere(F3 c, float r, int material, int mIn, int mOut) {
    prgpu::Primitive p{};
    p.type = prgpu::PrimSphere;
    p.material = material;
    p.light = -1;
    p.mediumInside = mIn;
    p.mediumOutside = mOut;
    p.p0 = pf(c);
    p.r = r;
    prims.push_back(p);
    return int(prims.size() - 1);
}

void SceneBuilder::AddEmitter(const Mesh &m, int spectrum, float scale, bool twoSided) {
    // Emitters carry the reference loader's zero-reflectance diffuse material (scene_loader.cpp:1030-1036).
    int black = Diffuse(Constant(0.f));
    int first = AddMesh(m, black, -1, -1, true);
    for (int i = first; i < int(prims.size()); ++i) {
        prgpu::LightEx l{};
        l.kind = prgpu::LightExArea;
        l.primitive = i;
        l.spectrum = spectrum;
        l.scale = scale;
        l.twoSided = twoSided ? 1 : 0;
        l.cosPower = 0;
        prims[size_t(i)].light = int(lights.size());
        lights.push_back(l);
    }
}

int SceneBuilder::AddPointLight(F3 p, int spectrum, float scale) {
    prgpu::LightEx l{};
    l.kind = prgpu::LightExPoint;
    l.primitive = -1;
    l.spectrum = spectrum;
    l.scale = scale;
    l.position = pf(p);
    lights.push_back(l);
    return int(lights.size() - 1);
}

int SceneBuilder::AddSpotLight(F3 p, F3 dir, int spectrum, float scale, float width, float start) {
    prgpu::LightEx l{};
    l.kind = prgpu::LightExSpot;
    l.primitive = -1;
    l.spectrum = spectrum;
    l.scale = scale;
    l.position = pf(p);
    l.direction = pf(norm3(dir));  // Normalize: v * (1 / |v|) (src/core/math.h:79-82)
    const float rad = kPi / 180.f;
    l.cosFalloffEnd = std::cos(rad * width);
    l.cosFalloffStart = std::cos(rad * std::min(start, width));
    lights.push_back(l);
    return int(lights.size() - 1);
}

int SceneBuilder::AddSunLight(F3 w, float radiusDeg, int spectrum, float scale) {
    prgpu::LightEx l{};
    l.kind = prgpu::LightExSun;
    l.primitive = -1;
    l.spectrum = spectrum;
    l.scale = scale;
    l.direction = pf(norm3(w));
    l.cosFalloffEnd = std::cos((kPi / 180.f) * std::clamp(radiusDeg, 0.01f, 45.f));
    lights.push_back(l);
    return int(lights.size() - 1);
}

int SceneBuilder::SunRadiance(F3 sunDir, float T, float radiusDeg) {
    F3 d = norm3(sunDir);
    float elevDeg = (180.f / kPi) * std::asin(std::clamp(d.y, -1.f, 1.f));
    float zenithDeg = 90 - elevDeg;
    float m = zenithDeg < 95 ? 1 / (std::cos((kPi / 180.f) * std::min(zenithDeg, 90.f)) +
                                    0.50572f * std::pow(96.07995f - zenithDeg, -1.6364f))
                             : 1e3f;
    float beta = std::max(0.f, 0.04608f * T - 0.04586f);
    const prgpu::SpectrumEx bb = spectra[size_t(BlackbodyUnitY(5778.f))];  // bb / Y(bb)
    spectra.pop_back();
    float r = 0.2665f / std::max(radiusDeg, 0.01f);
    float sizeScale = r * r;
    std::vector<float> l, v;
    for (int lambda = 360; lambda <= 830; lambda += 5) {
        float um = float(lambda) * 1e-3f;
        float tauR = 0.008735f * std::pow(um, -4.08f);
        float tauA = beta * std::pow(um, -1.3f);
        float trans = elevDeg < -1 ? 0.f : std::exp(-m * (tauR + tauA));
        l.push_back(float(lambda));
        v.push_back(2.0e6f * sizeScale * SpectrumValue(bb, knots.data(), float(lambda)) * trans);
    }
    return Piecewise(l, v);
}

int SceneBuilder::AddDistantLight(F3 w, int spectrum, float scale) {
    prgpu::LightEx l{};
    l.kind = prgpu::LightExDistant;
    l.primitive = -1;
    l.spectrum = spectrum;
    l.scale = scale;
    l.direction = pf(norm3(w));
    lights.push_back(l);
    return int(lights.size() - 1);
}

int SceneBuilder::AddEnvironment(int spectrum, float scale) {
    prgpu::LightEx l{};
    l.kind = prgpu::LightExEnvironment;
    l.primitive = -1;
    l.spectrum = spectrum;
    l.scale = scale;
    lights.push_back(l);
    return int(lights.size() - 1);
}

int SceneBuilder::AddEnvironmentMap(const float *rgb, int w, int h, float scale, float rotationDeg) {
    // EnvironmentLight(map, scale, rotation) (src/lights/light.cpp:221-239): world directions go
    // into the map's frame by the inverse of Rotate(rotation, +y).
    EnvMapDistribution dist;
    dist.Build(rgb, w, h);
    const int table = RgbTable();
    const int d65 = D65Illuminant();
    const float rad = (3.14159265358979323846f / 180.f) * rotationDeg;
    const int off = int(knots.size());
    const float header[prgpu::kEnvHeader] = {float(w), float(h), std::cos(rad), std::sin(rad), dist.margInt,
                                             dist.avgY, float(d65), float(table)};
    knots.insert(knots.end(), header, header + prgpu::kEnvHeader);
    knots.insert(knots.end(), rgb, rgb + size_t(w) * size_t(h) * 3);
    knots.insert(knots.end(), dist.func.begin(), dist.func.end());
    knots.insert(knots.end(), dist.condCdf.begin(), dist.condCdf.end());
    knots.insert(knots.end(), dist.rowInt.begin(), dist.rowInt.end());
    knots.insert(knots.end(), dist.margCdf.begin(), dist.margCdf.end());
    prgpu::LightEx l{};
    l.kind = prgpu::LightExEnvironment;
    l.primitive = -1;
    l.spectrum = d65;
    l.scale = scale;
    l.profileOffset = off;
    l.profileSize = int(knots.size()) - off;
    lights.push_back(l);
    return int(lights.size() - 1);
}

// Knots offsets that records hold as floats must be exact in a float (below 2^24).
static int FloatOffset(size_t off) {
    if (off >= (size_t(1) << 24)) throw std::runtime_error("scene too large: a texture record would start beyond 2^24 knots");
    return int(off);
}

// The RGB-to-spectrum table record, added to knots once per scene.
int SceneBuilder::RgbTable() {
    if (rgbTable_ < 0) {
        const std::vector<float> &t = RgbToSpectrumTableRecord();
        rgbTable_ = FloatOffset(knots.size());
        knots.insert(knots.end(), t.begin(), t.end());
    }
    return rgbTable_;
}

int SceneBuilder::AddImageTexture(const float *rgb, int w, int h, int wrap, bool flipV, const float uvScale[2],
                                  const float uvOffset[2], const float mult[3]) {
    const int table = RgbTable();
    const int off = FloatOffset(knots.size());
    const float header[prgpu::kTexHeader] = {float(w),    float(h),    float(wrap), flipV ? 1.f : 0.f,
                                             uvScale[0],  uvScale[1],  uvOffset[0], uvOffset[1],
                                             mult[0],     mult[1],     mult[2],     float(table)};
    knots.insert(knots.end(), header, header + prgpu::kTexHeader);
    knots.insert(knots.end(), rgb, rgb + size_t(w) * size_t(h) * 3);
    return off;
}

int SceneBuilder::ImageSpectrum(int textureRecord) {
    prgpu::SpectrumEx s{};
    s.kind = prgpu::SpecImage;
    s.scale = 1;
    s.offset = textureRecord;
    return AddSpectrum(s);
}

void SceneBuilder::SetShading(int material, int normalRecord, float strength, bool flipGreen, int bumpRecord,
                              int channel, float bumpScale) {
    const int off = int(knots.size());
    const float r[prgpu::kShadeHeader] = {normalRecord >= 0 ? 1.f : 0.f, float(std::max(0, normalRecord)), strength,
                                          flipGreen ? 1.f : 0.f,          bumpRecord >= 0 ? 1.f : 0.f,
                                          float(std::max(0, bumpRecord)), float(channel), bumpScale};
    knots.insert(knots.end(), r, r + prgpu::kShadeHeader);
    materials[size_t(material)].shading = uint32_t(off + 1);
}

int SceneBuilder::AddHomogeneous(int sa, int ss, float scale, float g) {
    prgpu::MediumEx m{};
    m.kind = prgpu::MediumHomogeneous;
    m.sigmaA = sa;
    m.sigmaS = ss;
    m.scale = scale;
    m.g = g;
    media.push_back(m);
    return int(media.size() - 1);
}

int SceneBuilder::AddGrid(int sa, int ss, float scale, float g, int nx, int ny, int nz, const std::vector<float> &data,
                          F3 bmin, F3 bmax) {
    prgpu::MediumEx m{};
    m.kind = prgpu::MediumGrid;
    m.sigmaA = sa;
    m.sigmaS = ss;
    m.scale = scale;
    m.g = g;
    m.nx = nx;
    m.ny = ny;
    m.nz = nz;
    m.gridOffset = int64_t(grid.size());
    float mx = 0;
    for (float v : data) {
        float c = (std::isfinite(v) && v > 0) ? v : 0.f;  // medium.cpp:108-115
        grid.push_back(c);
        mx = std::max(mx, c);
    }
    m.maxDensity = mx;
    m.boundsMin = pf(bmin);
    m.boundsMax = pf(bmax);
    float id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    std::memcpy(m.worldToMedium, id, sizeof(id));
    media.push_back(m);
    return int(media.size() - 1);
}

void LookAt(F3 pos, F3 target, F3 up, float o[12]) {
    F3 dir = norm3(target - pos);
    F3 right = crossd(dir, norm3(up));
    if (dot3(right, right) < 1e-12f) {
        float sign = std::copysign(1.f, dir.z);
        float a = -1 / (sign + dir.z);
        float b = dir.x * dir.y * a;
        right = f3(1 + sign * dir.x * dir.x * a, sign * b, -sign * dir.x);
    }
    right = norm3(right);
    F3 newUp = crossd(right, dir);
    float m[12] = {right.x, newUp.x, dir.x, pos.x, right.y, newUp.y, dir.y, pos.y, right.z, newUp.z, dir.z, pos.z};
    std::memcpy(o, m, sizeof(m));
}

void SceneBuilder::SetCamera(F3 pos, F3 target, float fov, int w, int h, float lensRadius, float focal) {
    LookAt(pos, target, f3(0, 1, 0), camera.cameraToWorld);
    camera.tanHalfFov = std::tan(((kPi / 180.f) * fov) / 2);
    camera.lensRadius = lensRadius;
    F3 d = target - pos;
    camera.focalDistance = focal > 0 ? focal : std::sqrt(dot3(d, d));
    camera.width = w;
    camera.height = h;
    if (camera.filterRadius == 0) {
        camera.filterType = prgpu::FilterBlackmanHarris;
        camera.filterRadius = 1.5f;
    }
}

void SceneBuilder::SetFilter(prgpu::FilterType t, float r) {
    camera.filterType = int(t);
    camera.filterRadius = t == prgpu::FilterBox ? 0.5f : r;
}

float SceneBuilder::SpectrumY(int id) const {
    float xyz[3];
    SpectrumXYZ(spectra[size_t(id)], knots.data(), xyz);
    return xyz[1];
}

const prgpu::SceneDescEx &SceneBuilder::Desc() {
    prgpu::SceneDescEx &d = desc_;
    d = prgpu::SceneDescEx{};
    d.apiVersion = prgpu::kApiVersion;
    d.extVersion = prgpu::kExtVersion;
    d.prims = prims.data();
    d.primCount = int(prims.size());
    d.materials = materials.data();
    d.materialCount = int(materials.size());
    d.spectra = spectra.data();
    d.spectrumCount = int(spectra.size());
    d.knots = knots.data();
    d.knotCount = int(knots.size());
    d.lights = lights.data();
    d.lightCount = int(lights.size());
    d.media = media.data();
    d.mediumCount = int(media.size());
    d.gridData = grid.data();
    d.gridCount = int64_t(grid.size());
    d.cameraMedium = cameraMedium;
    d.camera = camera;
    return d;
}

// ---------------------------------------------------------------------------------------------
// Meshes

Mesh MakeUVSphere(F3 c, float r, int seg, int rings) {
    Mesh m;
    for (int i = 0; i <= rings; ++i) {
        float th = kPi * float(i) / float(rings);
        for (int j = 0; j <= seg; ++j) {
            float ph = 2 * kPi * float(j) / float(seg);
            F3 n = f3(std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph));
            if (i == 0) n = f3(0, 1, 0);
            if (i == rings) n = f3(0, -1, 0);
            m.p.push_back(c + n * r);
            m.n.push_back(n);
            m.uv.push_back(float(j) / float(seg));
            m.uv.push_back(float(i) / float(rings));
        }
    }
    auto v = [&](int i, int j) { return i * (seg + 1) + j; };
    for (int i = 0; i < rings; ++i)
        for (int j = 0; j < seg; ++j) {
            int a = v(i, j), b = v(i, j + 1), cc = v(i + 1, j + 1), d = v(i + 1, j);
            if (i == 0) {
                m.idx.insert(m.idx.end(), {a, cc, d});
            } else if (i == rings - 1) {
                m.idx.insert(m.idx.end(), {a, b, d});
            } else {
                m.idx.insert(m.idx.end(), {a, b, cc});
                m.idx.insert(m.idx.end(), {a, cc, d});
            }
        }
    return m;
}

Mesh MakeIcosphere(F3 c, float r, int subdiv) {
    const float t = (1.f + std::sqrt(5.f)) / 2.f;
    std::vector<F3> v = {f3(-1, t, 0), f3(1, t, 0),  f3(-1, -t, 0), f3(1, -t, 0), f3(0, -1, t), f3(0, 1, t),
                         f3(0, -1, -t), f3(0, 1, -t), f3(t, 0, -1),  f3(t, 0, 1),  f3(-t, 0, -1), f3(-t, 0, 1)};
    for (F3 &p : v) p = norm3(p);
    std::vector<int> f = {0, 11, 5, 0, 5,  1, 0, 1, 7, 0, 7,  10, 0, 10, 11, 1, 5, 9, 5, 11, 4,  11, 10, 2,  10, 7,
                          6, 7,  1, 8, 3, 9, 4, 3, 4, 2, 3, 2,  6,  3, 6,  8,  3, 8, 9, 4, 9, 5, 2, 4,  11, 6,  2,  10,
                          8, 6,  7, 9, 8, 1};
    for (int s = 0; s < subdiv; ++s) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            v.push_back(norm3((v[size_t(a)] + v[size_t(b)]) * 0.5f));
            int id = int(v.size() - 1);
            mid[key] = id;
            return id;
        };
        std::vector<int> nf;
        for (size_t i = 0; i < f.size(); i += 3) {
            int a = f[i], b = f[i + 1], cc = f[i + 2];
            int ab = midpoint(a, b), bc = midpoint(b, cc), ca = midpoint(cc, a);
            nf.insert(nf.end(), {a, ab, ca, b, bc, ab, cc, ca, bc, ab, bc, ca});
        }
        f.swap(nf);
    }
    Mesh m;
    for (F3 n : v) {
        m.p.push_back(c + n * r);
        m.n.push_back(n);
        float u = std::atan2(n.z, n.x) / (2 * kPi) + 0.5f;
        float w = std::acos(std::clamp(n.y, -1.f, 1.f)) / kPi;
        m.uv.push_back(u);
        m.uv.push_back(w);
    }
    m.idx = f;
    return m;
}

Mesh MakeQuad(F3 corner, F3 U, F3 V) {
    Mesh m;
    F3 n = norm3(crossd(U, V));
    m.p = {corner, corner + U, corner + U + V, corner + V};
    m.n = {n, n, n, n};
    m.uv = {0, 0, 1, 0, 1, 1, 0, 1};
    m.idx = {0, 1, 2, 0, 2, 3};
    return m;
}

Mesh MakeBox(F3 center, F3 size) {  // outward normals (shape.cpp:331-353)
    Mesh m;
    F3 h = size * 0.5f;
    struct Face {
        F3 n, u, v;
    };
    const Face faces[6] = {{f3(1, 0, 0), f3(0, 0, -1), f3(0, 1, 0)},  {f3(-1, 0, 0), f3(0, 0, 1), f3(0, 1, 0)},
                           {f3(0, 1, 0), f3(1, 0, 0), f3(0, 0, -1)},  {f3(0, -1, 0), f3(1, 0, 0), f3(0, 0, 1)},
                           {f3(0, 0, 1), f3(1, 0, 0), f3(0, 1, 0)},   {f3(0, 0, -1), f3(-1, 0, 0), f3(0, 1, 0)}};
    for (const Face &fc : faces) {
        int base = int(m.p.size());
        F3 cc = f3(fc.n.x * h.x, fc.n.y * h.y, fc.n.z * h.z);
        F3 U = f3(fc.u.x * h.x, fc.u.y * h.y, fc.u.z * h.z), V = f3(fc.v.x * h.x, fc.v.y * h.y, fc.v.z * h.z);
        m.p.push_back(center + cc - U - V);
        m.p.push_back(center + cc + U - V);
        m.p.push_back(center + cc + U + V);
        m.p.push_back(center + cc - U + V);
        for (int i = 0; i < 4; ++i) m.n.push_back(fc.n);
        m.uv.insert(m.uv.end(), {0, 0, 1, 0, 1, 1, 0, 1});
        m.idx.insert(m.idx.end(), {base, base + 1, base + 2, base, base + 2, base + 3});
    }
    return m;
}

// ---------------------------------------------------------------------------------------------
// Convex hull (incremental, in double precision; the point sets here have ~150 points)

Mesh ConvexHull(const std::vector<F3> &pts) {
    struct D3 {
        double x, y, z;
    };
    auto sub = [](D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; };
    auto cross = [](D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; };
    auto dot = [](D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    std::vector<D3> P(pts.size());
    double scale = 0;
    for (size_t i = 0; i < pts.size(); ++i) {
        P[i] = D3{pts[i].x, pts[i].y, pts[i].z};
        scale = std::max(scale, std::max(std::abs(P[i].x), std::max(std::abs(P[i].y), std::abs(P[i].z))));
    }
    const double eps = 1e-9 * std::max(1.0, scale);
    struct Face {
        int a, b, c;
        D3 n;
        double d;
        bool alive;
    };
    std::vector<Face> faces;
    auto make = [&](int a, int b, int c) {
        D3 n = cross(sub(P[size_t(b)], P[size_t(a)]), sub(P[size_t(c)], P[size_t(a)]));
        double l = std::sqrt(dot(n, n));
        n = D3{n.x / l, n.y / l, n.z / l};
        faces.push_back(Face{a, b, c, n, dot(n, P[size_t(a)]), true});
    };
    Mesh m;
    const int N = int(P.size());
    if (N < 4) return m;
    // Initial tetrahedron: two far-apart points, the point farthest from their line, and the
    // point farthest from that plane.
    int i0 = 0, i1 = 0;
    for (int i = 1; i < N; ++i)
        if (P[size_t(i)].x < P[size_t(i0)].x) i0 = i;
    double best = -1;
    for (int i = 0; i < N; ++i) {
        D3 d = sub(P[size_t(i)], P[size_t(i0)]);
        if (dot(d, d) > best) best = dot(d, d), i1 = i;
    }
    int i2 = -1;
    best = -1;
    for (int i = 0; i < N; ++i) {
        D3 c = cross(sub(P[size_t(i1)], P[size_t(i0)]), sub(P[size_t(i)], P[size_t(i0)]));
        if (dot(c, c) > best) best = dot(c, c), i2 = i;
    }
    int i3 = -1;
    best = -1;
    D3 n012 = cross(sub(P[size_t(i1)], P[size_t(i0)]), sub(P[size_t(i2)], P[size_t(i0)]));
    for (int i = 0; i < N; ++i) {
        double dd = std::abs(dot(n012, sub(P[size_t(i)], P[size_t(i0)])));
        if (dd > best) best = dd, i3 = i;
    }
    if (best <= eps) return m;  // degenerate (planar) set
    if (dot(n012, sub(P[size_t(i3)], P[size_t(i0)])) > 0) std::swap(i1, i2);
    make(i0, i1, i2);
    make(i0, i3, i1);
    make(i1, i3, i2);
    make(i2, i3, i0);
    for (int p = 0; p < N; ++p) {
        if (p == i0 || p == i1 || p == i2 || p == i3) continue;
        std::vector<int> visible;
        for (int f = 0; f < int(faces.size()); ++f)
            if (faces[size_t(f)].alive && dot(faces[size_t(f)].n, P[size_t(p)]) - faces[size_t(f)].d > eps) visible.push_back(f);
        if (visible.empty()) continue;
        // Horizon: directed edges of visible faces whose twin belongs to no visible face.
        std::vector<std::pair<int, int>> edges;
        for (int f : visible) {
            const Face &F = faces[size_t(f)];
            int v[3] = {F.a, F.b, F.c};
            for (int e = 0; e < 3; ++e) edges.push_back({v[e], v[(e + 1) % 3]});
        }
        for (int f : visible) faces[size_t(f)].alive = false;
        for (const auto &e : edges) {
            bool twin = false;
            for (const auto &o : edges)
                if (o.first == e.second && o.second == e.first) {
                    twin = true;
                    break;
                }
            if (!twin) make(e.first, e.second, p);
        }
    }
    std::vector<int> remap(size_t(N), -1);
    for (const Face &F : faces) {
        if (!F.alive) continue;
        for (int v : {F.a, F.b, F.c}) {
            if (remap[size_t(v)] < 0) {
                remap[size_t(v)] = int(m.p.size());
                m.p.push_back(pts[size_t(v)]);
            }
            m.idx.push_back(remap[size_t(v)]);
        }
    }
    m.uv.assign(m.p.size() * 2, 0.f);
    return m;
}

// A round brilliant with ideal proportions (table 57%, crown 34.5 deg, pavilion 40.75 deg, girdle
// 3%, star 50%, lower-girdle 77%): the corners of the table, the star points and the lower-girdle
// points sit where neighbouring bezel / pavilion-main planes meet, so each kite is planar, and the
// girdle is a ring of segments (the cut of the reference's diamonds generator,
// scenes/diamonds/generate.py:30-58).
std::vector<F3> RoundBrilliantPoints(int ring) {
    const double kDeg = kPi / 180.0;
    const double table = 0.57, star = 0.5, lower = 0.77, g = 0.03 * 2;
    const double tc = std::tan(34.5 * kDeg), tp = std::tan(40.75 * kDeg), c22 = std::cos(22.5 * kDeg);
    std::vector<F3> pts;
    for (int k = 0; k < 8; ++k) {
        double a = 45.0 * k * kDeg, b = a + 22.5 * kDeg;
        pts.push_back(f3(float(table * std::cos(a)), float(g / 2 + (1 - table) * tc), float(table * std::sin(a))));
        double rs = table * c22 + star * (1 - table * c22);
        pts.push_back(f3(float(rs * std::cos(b)), float(g / 2 + (1 - rs * c22) * tc), float(rs * std::sin(b))));
        double rl = 1 - lower;
        pts.push_back(f3(float(rl * std::cos(b)), float(-g / 2 - (1 - rl * c22) * tp), float(rl * std::sin(b))));
    }
    for (int i = 0; i < ring; ++i) {
        double a = 2 * kPi * i / ring;
        pts.push_back(f3(float(std::cos(a)), float(g / 2), float(std::sin(a))));
        pts.push_back(f3(float(std::cos(a)), float(-g / 2), float(std::sin(a))));
    }
    pts.push_back(f3(0, float(-g / 2 - tp), 0));  // culet
    return pts;
}

// ---------------------------------------------------------------------------------------------
// Noise: Perlin's improved noise with his reference permutation (src/core/noise.cpp:7-60).

static const int kPermBase[256] = {
    151, 160, 137, 91,  90,  15,  131, 13,  201, 95,  96,  53,  194, 233, 7,   225, 140, 36,  103, 30,  69,  142,
    8,   99,  37,  240, 21,  10,  23,  190, 6,   148, 247, 120, 234, 75,  0,   26,  197, 62,  94,  252, 219, 203,
    117, 35,  11,  32,  57,  177, 33,  88,  237, 149, 56,  87,  174, 20,  125, 136, 171, 168, 68,  175, 74,  165,
    71,  134, 139, 48,  27,  166, 77,  146, 158, 231, 83,  111, 229, 122, 60,  211, 133, 230, 220, 105, 92,  41,
    55,  46,  245, 40,  244, 102, 143, 54,  65,  25,  63,  161, 1,   216, 80,  73,  209, 76,  132, 187, 208, 89,
    18,  169, 200, 196, 135, 130, 116, 188, 159, 86,  164, 100, 109, 198, 173, 186, 3,   64,  52,  217, 226, 250,
    124, 123, 5,   202, 38,  147, 118, 126, 255, 82,  85,  212, 207, 206, 59,  227, 47,  16,  58,  17,  182, 189,
    28,  42,  223, 183, 170, 213, 119, 248, 152, 2,   44,  154, 163, 70,  221, 153, 101, 155, 167, 43,  172, 9,
    129, 22,  39,  253, 19,  98,  108, 110, 79,  113, 224, 232, 178, 185, 112, 104, 218, 246, 97,  228, 251, 34,
    242, 193, 238, 210, 144, 12,  191, 179, 162, 241, 81,  51,  145, 235, 249, 14,  239, 107, 49,  192, 214, 31,
    181, 199, 106, 157, 184, 84,  204, 176, 115, 121, 50,  45,  127, 4,   150, 254, 138, 236, 205, 93,  222, 114,
    67,  29,  24,  72,  243, 141, 128, 195, 78,  66,  215, 61,  156, 180};

static int Perm(int i) { return kPermBase[i & 255]; }
static float Fade(float t) { return t * t * t * (t * (t * 6 - 15) + 10); }
static float Grad(int hash, float x, float y, float z) {
    int h = hash & 15;
    float u = h < 8 ? x : y;
    float v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
}
static float LerpH(float t, float a, float b) { return (1 - t) * a + t * b; }

float PerlinNoise(F3 p) {
    float fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
    int X = int(fx) & 255, Y = int(fy) & 255, Z = int(fz) & 255;
    float x = p.x - fx, y = p.y - fy, z = p.z - fz;
    float u = Fade(x), v = Fade(y), w = Fade(z);
    int A = Perm(X) + Y, AA = Perm(A) + Z, AB = Perm(A + 1) + Z;
    int B = Perm(X + 1) + Y, BA = Perm(B) + Z, BB = Perm(B + 1) + Z;
    return LerpH(w,
                 LerpH(v, LerpH(u, Grad(Perm(AA), x, y, z), Grad(Perm(BA), x - 1, y, z)),
                       LerpH(u, Grad(Perm(AB), x, y - 1, z), Grad(Perm(BB), x - 1, y - 1, z))),
                 LerpH(v, LerpH(u, Grad(Perm(AA + 1), x, y, z - 1), Grad(Perm(BA + 1), x - 1, y, z - 1)),
                       LerpH(u, Grad(Perm(AB + 1), x, y - 1, z - 1), Grad(Perm(BB + 1), x - 1, y - 1, z - 1))));
}

float FBm(F3 p, int octaves, float lacunarity, float gain) {
    float sum = 0, amp = 0.5f, freq = 1, norm = 0;
    for (int i = 0; i < std::max(1, octaves); ++i) {
        sum += amp * PerlinNoise(p * freq);
        norm += amp;
        amp *= gain;
        freq *= lacunarity;
    }
    return sum / norm;
}

// ---------------------------------------------------------------------------------------------
// Benchmark configurations

static void GenPyramid(const GenOptions &o, SceneBuilder *sb) {
    // 30 unit spheres in a 4-level pyramid (16 + 9 + 4 + 1) on a 40 x 40 ground, one 2 x 2
    // 6500 K quad light at y = 12; pinhole camera at (0, 6, -18), 40 deg.
    SceneRng rng(1);
    int W = o.width > 0 ? o.width : 256, H = o.height > 0 ? o.height : 256;
    sb->SetCamera(f3(0, 6, -18), f3(0, 0, 0), 40.f, W, H);
    sb->AddMesh(MakeQuad(f3(-20, 0, -20), f3(0, 0, 40), f3(40, 0, 0)), sb->Diffuse(sb->Constant(0.5f)));
    const float ys[4] = {1.f, 2.41f, 3.83f, 5.24f};
    int segs = std::max(4, int(std::lround(32 * o.detail))), rings = std::max(3, int(std::lround(16 * o.detail)));
    for (int level = 0; level < 4; ++level) {
        int n = 4 - level;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                float x = 2.f * float(i) - float(n - 1), z = 2.f * float(j) - float(n - 1);
                float r = rng.uniform(0.1f, 0.9f), g = rng.uniform(0.1f, 0.9f), b = rng.uniform(0.1f, 0.9f);
                int mat = sb->Diffuse(sb->RGBAlbedo(r, g, b));
                sb->AddMesh(MakeUVSphere(f3(x, ys[level], z), 1.f, segs, rings), mat);
            }
    }
    // Rect light facing -y (its emitting side), as a loader "rect" light with direction (0,-1,0).
    sb->AddEmitter(MakeQuad(f3(-1, 12, -1), f3(2, 0, 0), f3(0, 0, 2)), sb->BlackbodyUnitY(6500.f), 60.f);
    sb->spp = 64;
    sb->maxDepth = 8;
    sb->rrStartDepth = 3;
    sb->seed = 1;
}

static void GenTerrain(const GenOptions &o, SceneBuilder *sb) {
    // 1000 x 500 fBm heightfield (amplitude 3) + 200 level-4 icospheres with 40/30/30
    // Lambert / rough anisotropic conductor / rough BK7 materials, lit by 64 blackbody quads.
    SceneRng rng(2);
    int W = o.width > 0 ? o.width : 1920, H = o.height > 0 ? o.height : 1080;
    const int NX = std::max(8, int(std::lround(1000 * o.detail))), NZ = std::max(4, int(std::lround(500 * o.detail)));
    const float X0 = -50, X1 = 50, Z0 = -25, Z1 = 25, amp = 3.f;
    F3 noiseOff = f3(rng.uniform(0, 100), rng.uniform(0, 100), rng.uniform(0, 100));
    auto height = [&](float x, float z) { return amp * FBm(f3(x * 0.06f, 0.f, z * 0.06f) + noiseOff, 6); };
    Mesh t;
    t.p.resize(size_t(NX) * NZ);
    t.uv.resize(size_t(NX) * NZ * 2);
    for (int j = 0; j < NZ; ++j)
        for (int i = 0; i < NX; ++i) {
            float x = X0 + (X1 - X0) * float(i) / float(NX - 1), z = Z0 + (Z1 - Z0) * float(j) / float(NZ - 1);
            t.p[size_t(j) * NX + i] = f3(x, height(x, z), z);
            t.uv[2 * (size_t(j) * NX + i)] = float(i) / float(NX - 1) * 10;
            t.uv[2 * (size_t(j) * NX + i) + 1] = float(j) / float(NZ - 1) * 5;
        }
    for (int j = 0; j + 1 < NZ; ++j)
        for (int i = 0; i + 1 < NX; ++i) {
            int a = j * NX + i, b = a + 1, c = a + NX + 1, d = a + NX;
            t.idx.insert(t.idx.end(), {a, c, b, a, d, c});  // normals face +y
        }
    sb->AddMesh(t, sb->Diffuse(sb->RGBAlbedo(0.42f, 0.36f, 0.28f)), -1, -1, false);

    const char *metals[4] = {"Au", "Ag", "Cu", "Al"};
    int bk7 = sb->Glass("BK7");
    int subdiv = o.detail >= 1.f ? 4 : (o.detail >= 0.25f ? 3 : 2);
    const int nSpheres = 200;
    for (int s = 0; s < nSpheres; ++s) {
        float r = rng.uniform(0.3f, 1.5f);
        float x = rng.uniform(X0 + 5, X1 - 5), z = rng.uniform(Z0 + 3, Z1 - 3);
        float y = height(x, z) + 0.7f * r;
        float kind = rng.uniform();
        int mat;
        if (kind < 0.4f) {
            mat = sb->Diffuse(sb->RGBAlbedo(rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f)));
        } else if (kind < 0.7f) {
            int m = std::min(3, int(rng.uniform() * 4));
            float alpha = rng.uniform(0.05f, 0.5f), aniso = rng.uniform(0.f, 0.8f);
            // Anisotropy a stretches the lobe: alpha_x = alpha / sqrt(1 - 0.9 a), alpha_y = alpha sqrt(1 - 0.9 a).
            float asp = std::sqrt(1 - 0.9f * aniso);
            mat = sb->Conductor(sb->Metal(metals[m], false), sb->Metal(metals[m], true), alpha / asp, alpha * asp);
        } else {
            float alpha = rng.uniform(0.05f, 0.5f);
            mat = sb->Dielectric(bk7, alpha, alpha);
        }
        sb->AddMesh(MakeIcosphere(f3(x, y, z), r, subdiv), mat);
    }
    for (int l = 0; l < 64; ++l) {
        float x = rng.uniform(X0 + 2, X1 - 2), z = rng.uniform(Z0 + 2, Z1 - 2);
        float y = height(x, z) + rng.uniform(6.f, 10.f);
        float T = rng.uniform(2700.f, 9000.f);
        sb->AddEmitter(MakeQuad(f3(x - 0.75f, y, z - 0.75f), f3(1.5f, 0, 0), f3(0, 0, 1.5f)), sb->BlackbodyUnitY(T), 25.f);
    }
    sb->SetCamera(f3(0, 16, -42), f3(0, 0, 2), 40.f, W, H);
    sb->spp = 1024;
    sb->maxDepth = 16;
    sb->rrStartDepth = 1;
    sb->seed = 1;
}

static void GenMedia(const GenOptions &o, SceneBuilder *sb) {
    // A 6 x 4 x 6 box of heterogeneous medium (a 256^3 grid of max(0, fBm) densities; sigma_t =
    // d * RGBUnbounded(1, 0.8, 0.5) * 4, 90% scattering, Henyey-Greenstein g = 0.6) beside a
    // subsurface sphere (smooth IOR 1.33 interface around a homogeneous medium with
    // sigma_s ~ U[0.5, 2]^3), on a grey ground, lit by one 5800 K distant light.
    SceneRng rng(3);
    int W = o.width > 0 ? o.width : 1280, H = o.height > 0 ? o.height : 720;
    sb->SetCamera(f3(0, 4, -14), f3(0.5f, 2.f, 0), 40.f, W, H);
    sb->AddMesh(MakeQuad(f3(-20, 0, -20), f3(0, 0, 40), f3(40, 0, 0)), sb->Diffuse(sb->Constant(0.5f)));

    const int N = std::max(8, int(std::lround(256 * o.detail)));
    const F3 bmin = f3(-5, 0.5f, -3), bmax = f3(1, 4.5f, 3);
    F3 off = f3(rng.uniform(0, 100), rng.uniform(0, 100), rng.uniform(0, 100));
    std::vector<float> density(size_t(N) * N * N);
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                F3 p = f3((x + 0.5f) / N, (y + 0.5f) / N, (z + 0.5f) / N);
                density[(size_t(z) * N + y) * N + x] = std::max(0.f, FBm(p * 4.f + off, 6));
            }
    int st = sb->RGBUnbounded(1.0f, 0.8f, 0.5f);
    int grid = sb->AddGrid(sb->Scaled(st, 0.1f), sb->Scaled(st, 0.9f), 4.f, 0.6f, N, N, N, density, bmin, bmax);
    sb->AddMesh(MakeBox((bmin + bmax) * 0.5f, bmax - bmin), sb->Interface(), grid, -1);

    float sr = rng.uniform(0.5f, 2.f), sg = rng.uniform(0.5f, 2.f), sbl = rng.uniform(0.5f, 2.f);
    int sss = sb->AddHomogeneous(sb->Constant(0.05f), sb->RGBUnbounded(sr, sg, sbl), 1.f, 0.f);
    sb->AddSphere(f3(4, 1.5f, 0), 1.5f, sb->Dielectric(sb->Constant(1.33f), 0, 0), sss, -1);

    sb->AddDistantLight(f3(-0.5f, 1.f, -0.6f), sb->BlackbodyUnitY(5800.f), 3.f);
    sb->spp = 512;
    sb->maxDepth = 256;
    sb->rrStartDepth = 1;
    sb->seed = 1;
}

// A box with each face split into n x n quads (two triangles each), rotated about +y.
static Mesh MakeTessellatedBox(F3 c, F3 size, int n, float rotY) {
    Mesh m;
    F3 h = size * 0.5f;
    const float cs = std::cos(rotY), sn = std::sin(rotY);
    auto place = [&](F3 p) { return c + f3(cs * p.x + sn * p.z, p.y, -sn * p.x + cs * p.z); };
    struct Face {
        F3 n, u, v;
    };
    const Face faces[6] = {{f3(1, 0, 0), f3(0, 0, -1), f3(0, 1, 0)}, {f3(-1, 0, 0), f3(0, 0, 1), f3(0, 1, 0)},
                           {f3(0, 1, 0), f3(1, 0, 0), f3(0, 0, -1)}, {f3(0, -1, 0), f3(1, 0, 0), f3(0, 0, 1)},
                           {f3(0, 0, 1), f3(1, 0, 0), f3(0, 1, 0)},  {f3(0, 0, -1), f3(-1, 0, 0), f3(0, 1, 0)}};
    for (const Face &fc : faces) {
        F3 cc = f3(fc.n.x * h.x, fc.n.y * h.y, fc.n.z * h.z);
        F3 U = f3(fc.u.x * h.x, fc.u.y * h.y, fc.u.z * h.z), V = f3(fc.v.x * h.x, fc.v.y * h.y, fc.v.z * h.z);
        int base = int(m.p.size());
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                float a = 2.f * float(i) / float(n) - 1.f, b = 2.f * float(j) / float(n) - 1.f;
                m.p.push_back(place(cc + U * a + V * b));
                m.uv.insert(m.uv.end(), {float(i) / float(n), float(j) / float(n)});
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int a = base + j * (n + 1) + i, b = a + 1, c2 = a + n + 2, d = a + n + 1;
                m.idx.insert(m.idx.end(), {a, b, c2, a, c2, d});
            }
    }
    return m;
}

static void GenRoom(const GenOptions &o, SceneBuilder *sb) {
    // A closed 10 x 10 x 4 room lit only by a 5800 K sun through a 1 x 1 window, holding 120
    // finely tessellated boxes (19 x 19 quads per face, albedo U[0.2, 0.8]^3): almost all light
    // arrives indirectly, the case learned path guiding is meant for.
    SceneRng rng(4);
    int W = o.width > 0 ? o.width : 1920, H = o.height > 0 ? o.height : 1080;
    sb->SetCamera(f3(-4.2f, 2.2f, -4.6f), f3(1.5f, 1.0f, 3.f), 60.f, W, H);
    int wall = sb->Diffuse(sb->Constant(0.7f)), floorM = sb->Diffuse(sb->Constant(0.5f));
    const float X = 5, Y = 4, Z = 5;
    sb->AddMesh(MakeQuad(f3(-X, 0, -Z), f3(0, 0, 2 * Z), f3(2 * X, 0, 0)), floorM);       // floor, +y
    sb->AddMesh(MakeQuad(f3(-X, Y, -Z), f3(2 * X, 0, 0), f3(0, 0, 2 * Z)), wall);         // ceiling, -y
    sb->AddMesh(MakeQuad(f3(-X, 0, -Z), f3(0, Y, 0), f3(0, 0, 2 * Z)), wall);             // x = -5, +x
    sb->AddMesh(MakeQuad(f3(X, 0, -Z), f3(0, 0, 2 * Z), f3(0, Y, 0)), wall);              // x = +5, -x
    sb->AddMesh(MakeQuad(f3(-X, 0, -Z), f3(2 * X, 0, 0), f3(0, Y, 0)), wall);             // z = -5, +z
    // z = +5 wall with a window over x in [1, 2], y in [2, 3], facing -z.
    const float wx0 = 1, wx1 = 2, wy0 = 2, wy1 = 3;
    sb->AddMesh(MakeQuad(f3(-X, 0, Z), f3(0, Y, 0), f3(wx0 + X, 0, 0)), wall);
    sb->AddMesh(MakeQuad(f3(wx1, 0, Z), f3(0, Y, 0), f3(X - wx1, 0, 0)), wall);
    sb->AddMesh(MakeQuad(f3(wx0, 0, Z), f3(0, wy0, 0), f3(wx1 - wx0, 0, 0)), wall);
    sb->AddMesh(MakeQuad(f3(wx0, wy1, Z), f3(0, Y - wy1, 0), f3(wx1 - wx0, 0, 0)), wall);
    const int n = std::max(1, int(std::lround(19 * o.detail)));
    const int boxes = std::max(4, int(std::lround(120 * std::min(1.f, o.detail * 4))));
    for (int b = 0; b < boxes; ++b) {
        F3 size = f3(rng.uniform(0.3f, 1.0f), rng.uniform(0.3f, 1.0f), rng.uniform(0.3f, 1.0f));
        F3 c = f3(rng.uniform(-4.4f, 4.4f), 0.5f * size.y, rng.uniform(-4.4f, 4.4f));
        float rot = rng.uniform(0.f, kPi);
        int mat = sb->Diffuse(sb->RGBAlbedo(rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f)));
        sb->AddMesh(MakeTessellatedBox(c, size, n, rot), mat, -1, -1, false);
    }
    sb->AddDistantLight(f3(-0.35f, 0.55f, 0.76f), sb->BlackbodyUnitY(5800.f), 600.f);
    sb->spp = 256;
    sb->maxDepth = 32;
    sb->rrStartDepth = 1;
    sb->seed = 1;
}

static void GenGems(const GenOptions &o, SceneBuilder *sb) {
    // 150 round-brilliant diamonds (Sellmeier dispersion) at random scales U[0.5, 2] and uniformly
    // random orientations above a rough silver plane (GGX alpha 0.1), lit by a 0.2 x 0.2 quad at
    // 5000 K and a dim constant environment; thin-lens camera at f/2.8.
    SceneRng rng(5);
    int W = o.width > 0 ? o.width : 3840, H = o.height > 0 ? o.height : 2160;
    const float fov = 30.f;
    const F3 eye = f3(0, 14, -32), at = f3(0, 0, 0);
    // f-number to lens radius as the reference loader converts it (src/scene/scene_loader.cpp:1156-1159):
    // focal length for a 24 mm sensor height, 100 scene units per metre.
    const float focalMm = 24.f / (2 * std::tan((kPi / 180.f) * fov / 2)), fStop = 2.8f, unitsPerMetre = 100.f;
    const float lensRadius = (focalMm / 1000.f) / (2 * fStop) * unitsPerMetre;
    sb->SetCamera(eye, at, fov, W, H, lensRadius, 0.f);

    int ag = sb->Conductor(sb->Metal("Ag", false), sb->Metal("Ag", true), 0.1f, 0.1f);
    sb->AddMesh(MakeQuad(f3(-40, 0, -25), f3(0, 0, 50), f3(80, 0, 0)), ag);

    const float B[3] = {0.3306f, 4.3356f, 0.f}, C[3] = {0.030625f, 0.011236f, 0.f};
    int diamond = sb->Dielectric(sb->Sellmeier(B, C, 2), 0, 0);
    const Mesh cut = ConvexHull(RoundBrilliantPoints(std::max(16, int(std::lround(64 * std::min(1.f, o.detail * 4))))));
    const int nGems = std::max(4, int(std::lround(150 * std::min(1.f, o.detail * 4))));
    std::vector<F3> centres;
    std::vector<float> radii;
    for (int g = 0; g < nGems; ++g) {
        float sc = rng.uniform(0.5f, 2.f), bound = 1.15f * sc;
        F3 c;
        bool ok = false;
        for (int attempt = 0; attempt < 1000 && !ok; ++attempt) {
            c = f3(rng.uniform(-32, 32), bound, rng.uniform(-20, 20));
            ok = true;
            for (size_t k = 0; k < centres.size() && ok; ++k) {
                F3 d = c - centres[k];
                ok = dot3(d, d) > (bound + radii[k]) * (bound + radii[k]);
            }
        }
        if (!ok) continue;
        centres.push_back(c);
        radii.push_back(bound);
        // Uniform random rotation (Shoemake's quaternion from three uniforms).
        float u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
        float qa = std::sqrt(1 - u1) * std::sin(2 * kPi * u2), qb = std::sqrt(1 - u1) * std::cos(2 * kPi * u2);
        float qc = std::sqrt(u1) * std::sin(2 * kPi * u3), qd = std::sqrt(u1) * std::cos(2 * kPi * u3);
        float x = qa, y = qb, z = qc, w = qd;
        float R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                      2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                      2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)};
        Mesh gem = cut;
        for (F3 &p : gem.p) {
            F3 q = f3(R[0] * p.x + R[1] * p.y + R[2] * p.z, R[3] * p.x + R[4] * p.y + R[5] * p.z,
                      R[6] * p.x + R[7] * p.y + R[8] * p.z);
            p = c + q * sc;
        }
        sb->AddMesh(gem, diamond, -1, -1, false);
    }
    sb->AddEmitter(MakeQuad(f3(-0.1f, 15, -0.1f), f3(0.2f, 0, 0), f3(0, 0, 0.2f)), sb->BlackbodyUnitY(5000.f), 5000.f);
    sb->AddEnvironment(sb->Constant(1.f), 0.05f);
    sb->spp = 4096;
    sb->maxDepth = 32;
    sb->rrStartDepth = 1;
    sb->seed = 1;
}

static void GenMaterials(const GenOptions &o, SceneBuilder *sb) {
    // The pyramid of spheres with coated and sheen materials (no reference configuration uses
    // them; they exercise the coated / sheen closures): clear, rough, scattering and dispersive
    // coats over diffuse bases, coated gold and coated artist conductors, sheen over diffuse,
    // coated and rough dielectric bases. Quad light plus a dim constant environment.
    SceneRng rng(6);
    int W = o.width > 0 ? o.width : 256, H = o.height > 0 ? o.height : 256;
    sb->SetCamera(f3(0, 6, -18), f3(0, 0, 0), 40.f, W, H);
    sb->AddMesh(MakeQuad(f3(-20, 0, -20), f3(0, 0, 40), f3(40, 0, 0)), sb->Diffuse(sb->Constant(0.5f)));
    const float ys[4] = {1.f, 2.41f, 3.83f, 5.24f};
    int segs = std::max(4, int(std::lround(32 * o.detail))), rings = std::max(3, int(std::lround(16 * o.detail)));
    const int glass = sb->Constant(1.5f), water = sb->Constant(1.33f), bk7 = sb->Glass("BK7");
    const int auN = sb->Metal("Au", false), auK = sb->Metal("Au", true);
    int k = 0;
    for (int level = 0; level < 4; ++level) {
        int n = 4 - level;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j, ++k) {
                float x = 2.f * float(i) - float(n - 1), z = 2.f * float(j) - float(n - 1);
                float r = rng.uniform(0.1f, 0.9f), g = rng.uniform(0.1f, 0.9f), b = rng.uniform(0.1f, 0.9f);
                int albedo = sb->RGBAlbedo(r, g, b);
                SceneBuilder::Coat c;
                c.eta = glass;
                int mat;
                switch (k % 8) {
                case 0: mat = sb->CoatedDiffuse(c, albedo); break;
                case 1:
                    c.roughness = 0.3f;
                    c.eta = water;
                    mat = sb->CoatedDiffuse(c, albedo);
                    break;
                case 2:
                    c.thickness = 0.1f;
                    c.albedo = sb->RGBAlbedo(0.8f, 0.6f, 0.3f);
                    c.g = 0.3f;
                    mat = sb->CoatedDiffuse(c, albedo);
                    break;
                case 3:
                    c.eta = bk7;
                    mat = sb->CoatedDiffuse(c, albedo);
                    break;
                case 4: mat = sb->CoatedConductor(c, auN, auK, 0.04f, 0.04f); break;
                case 5:
                    c.roughness = 0.2f;
                    mat = sb->CoatedConductorReflectance(c, albedo, 0.1f, 0.1f);
                    break;
                case 6:
                    mat = sb->Diffuse(albedo);
                    sb->SetSheen(mat, sb->RGBAlbedo(0.9f, 0.9f, 0.9f), 0.3f, 1.f);
                    break;
                default:
                    c.roughness = 0.1f;
                    mat = sb->CoatedDiffuse(c, albedo);
                    sb->SetSheen(mat, sb->RGBAlbedo(1.f, 0.8f, 0.6f), 0.5f, 0.7f);
                    break;
                }
                sb->AddMesh(MakeUVSphere(f3(x, ys[level], z), 1.f, segs, rings), mat);
            }
    }
    sb->AddEmitter(MakeQuad(f3(-1, 12, -1), f3(2, 0, 0), f3(0, 0, 2)), sb->BlackbodyUnitY(6500.f), 60.f);
    sb->AddEnvironment(sb->Constant(1.f), 0.05f);
    sb->spp = 64;
    sb->maxDepth = 8;
    sb->rrStartDepth = 3;
    sb->seed = 1;
}

bool GenerateConfig(int config, const GenOptions &o, SceneBuilder *sb, std::string *err) {
    switch (config) {
    case 1: GenPyramid(o, sb); return true;
    case 2: GenTerrain(o, sb); return true;
    case 3: GenMedia(o, sb); return true;
    case 4: GenRoom(o, sb); return true;
    case 5: GenGems(o, sb); return true;
    case 6: GenMaterials(o, sb); return true;
    default:
        if (err) *err = "configuration " + std::to_string(config) + " is not available";
        return false;
    }
}

}  // namespace phost
