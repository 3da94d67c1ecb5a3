// Oracle: CPU restatement of the reference CPU volumetric path tracer, used ONLY as test
// infrastructure (the checker in tests/, the cpu_baseline leg of bench.py and smoke()). No
// product code links or calls it.
//
// This header restates the reference's arithmetic helpers. Every function cites the reference
// file:line it follows; operation order is kept where the result is compared bit for bit
// (triangle intersection, BVH build and traversal) or to 1e-4 (everything on the radiance path).
// Build with -ffp-contract=off: the reference compares against an x86 build without FMA
// contraction (its own /arch:AVX2 MSVC build does not contract either).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace orc {

// ---- constants (src/core/math.h:11-19) -------------------------------------------------------
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv2Pi = 0.15915494309189533577f;
constexpr float kInv4Pi = 0.07957747154594766788f;
constexpr float kPiOver2 = 1.57079632679489661923f;
constexpr float kPiOver4 = 0.78539816339744830961f;
constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr float kOneMinusEps = 0x1.fffffep-1f;
constexpr float kLambdaMin = 360.f, kLambdaMax = 830.f;

template <class T> inline T sq(T v) { return v * v; }
template <class T, class A, class B> inline T clampv(T v, A lo, B hi) {  // math.h:21-24
    return v < T(lo) ? T(lo) : (v > T(hi) ? T(hi) : v);
}
inline float lerpf(float t, float a, float b) { return (1 - t) * a + t * b; }  // math.h:25
inline float safe_sqrt(float x) { return std::sqrt(std::max(0.f, x)); }        // math.h:27
inline float safe_asin(float x) { return std::asin(clampv(x, -1.f, 1.f)); }
inline float safe_acos(float x) { return std::acos(clampv(x, -1.f, 1.f)); }
// Hermite smoothstep between a and b; a step at a when a == b (src/core/math.h:32-36).
inline float smooth_step(float x, float a, float b) {
    if (a == b) return x < a ? 0.f : 1.f;
    float t = clampv((x - a) / (b - a), 0.f, 1.f);
    return t * t * (3 - 2 * t);
}
inline float deg2rad(float d) { return (kPi / 180.f) * d; }
inline float rad2deg(float r) { return (180.f / kPi) * r; }

// Largest i in [0, size-2] with pred(i) true (math.h:45-54).
template <class P> inline int find_interval(int size, const P &pred) {
    int first = 1, len = size - 2;
    while (len > 0) {
        int half = len >> 1, mid = first + half;
        bool r = pred(mid);
        first = r ? mid + 1 : first;
        len = r ? len - (half + 1) : half;
    }
    return clampv(first - 1, 0, size - 2);
}

// ---- vectors (math.h:57-150) -----------------------------------------------------------------
struct V2 {
    float x = 0, y = 0;
};

struct V3 {
    float x = 0, y = 0, z = 0;
    V3() = default;
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    bool zero() const { return x == 0 && y == 0 && z == 0; }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return a * s; }
inline V3 operator/(V3 a, float s) {  // math.h:79-82: multiply by the reciprocal
    float inv = 1.f / s;
    return {a.x * inv, a.y * inv, a.z * inv};
}
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float absdot(V3 a, V3 b) { return std::abs(dot(a, b)); }
inline V3 cross_d(V3 a, V3 b) {  // math.h:95-99: double internally
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return {float(ay * bz - az * by), float(az * bx - ax * bz), float(ax * by - ay * bx)};
}
inline V3 cross_f(V3 a, V3 b) {  // src/geometry/shape.cpp:7-9: float cross for ray/triangle
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float len2(V3 v) { return dot(v, v); }
inline float len(V3 v) { return std::sqrt(len2(v)); }
inline V3 normalize(V3 v) { return v / len(v); }
inline float dist(V3 a, V3 b) { return len(a - b); }
inline float dist2(V3 a, V3 b) { return len2(a - b); }
inline V3 vabs(V3 v) { return {std::abs(v.x), std::abs(v.y), std::abs(v.z)}; }
inline V3 vmin(V3 a, V3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
inline float max_comp(V3 v) { return std::max(v.x, std::max(v.y, v.z)); }
inline int max_dim(V3 v) { return (v.x > v.y) ? (v.x > v.z ? 0 : 2) : (v.y > v.z ? 1 : 2); }
inline V3 face_forward(V3 n, V3 v) { return dot(n, v) < 0.f ? -n : n; }

inline void coordinate_system(V3 v1, V3 *v2, V3 *v3) {  // math.h:120-126 (Duff et al. 2017)
    float sign = std::copysign(1.f, v1.z);
    float a = -1 / (sign + v1.z);
    float b = v1.x * v1.y * a;
    *v2 = V3(1 + sign * v1.x * v1.x * a, sign * b, -sign * v1.x);
    *v3 = V3(b, sign + v1.y * v1.y * a, -v1.y);
}

struct Frame {  // math.h:128-150
    V3 x{1, 0, 0}, y{0, 1, 0}, z{0, 0, 1};
    static Frame from_z(V3 zz) {
        Frame f;
        f.z = zz;
        coordinate_system(zz, &f.x, &f.y);
        return f;
    }
    static Frame from_xz(V3 xin, V3 zz) {
        V3 xx = xin - zz * dot(xin, zz);
        float l2 = len2(xx);
        if (l2 < 1e-12f || !std::isfinite(l2)) return from_z(zz);
        xx = xx / std::sqrt(l2);
        Frame f;
        f.z = zz;
        f.x = xx;
        f.y = cross_d(zz, xx);
        return f;
    }
    V3 to_local(V3 v) const { return {dot(v, x), dot(v, y), dot(v, z)}; }
    V3 from_local(V3 v) const { return x * v.x + y * v.y + z * v.z; }
};

// Local shading-frame trigonometry (math.h:153-177).
inline float cos2t(V3 w) { return w.z * w.z; }
inline float abscos(V3 w) { return std::abs(w.z); }
inline float sin2t(V3 w) { return std::max(0.f, 1 - cos2t(w)); }
inline float sint(V3 w) { return std::sqrt(sin2t(w)); }
inline float tan2t(V3 w) { return sin2t(w) / cos2t(w); }
inline float cosphi(V3 w) {
    float s = sint(w);
    return s == 0 ? 1 : clampv(w.x / s, -1.f, 1.f);
}
inline float sinphi(V3 w) {
    float s = sint(w);
    return s == 0 ? 0 : clampv(w.y / s, -1.f, 1.f);
}
inline bool same_hemi(V3 a, V3 b) { return a.z * b.z > 0; }
inline V3 spherical_dir(float sinT, float cosT, float phi) {
    return {clampv(sinT, -1.f, 1.f) * std::cos(phi), clampv(sinT, -1.f, 1.f) * std::sin(phi), clampv(cosT, -1.f, 1.f)};
}
inline V3 reflect(V3 wo, V3 n) { return -wo + n * (2 * dot(wo, n)); }  // math.h:179

inline bool refract(V3 wi, V3 n, float eta, float *etap, V3 *wt) {  // math.h:183-197
    float ci = dot(n, wi);
    if (ci < 0) {
        eta = 1 / eta;
        ci = -ci;
        n = -n;
    }
    float s2i = std::max(0.f, 1 - sq(ci));
    float s2t = s2i / sq(eta);
    if (s2t >= 1) return false;
    float ct = safe_sqrt(1 - s2t);
    *wt = -wi / eta + n * (ci / eta - ct);
    if (etap) *etap = eta;
    return true;
}

struct Ray {
    V3 o, d;
    V3 at(float t) const { return o + d * t; }
};

struct Box {  // math.h:207-237
    V3 lo{kInf, kInf, kInf}, hi{-kInf, -kInf, -kInf};
    bool empty() const { return lo.x > hi.x || lo.y > hi.y || lo.z > hi.z; }
    V3 diag() const { return hi - lo; }
    V3 centroid() const { return (lo + hi) * 0.5f; }
    float area() const {
        if (empty()) return 0;
        V3 d = diag();
        return 2 * (d.x * d.y + d.x * d.z + d.y * d.z);
    }
    int max_extent() const { return max_dim(diag()); }
    V3 offset(V3 p) const {
        V3 o = p - lo;
        if (hi.x > lo.x) o.x /= hi.x - lo.x;
        if (hi.y > lo.y) o.y /= hi.y - lo.y;
        if (hi.z > lo.z) o.z /= hi.z - lo.z;
        return o;
    }
    bool inside(V3 p) const {
        return p.x >= lo.x && p.x <= hi.x && p.y >= lo.y && p.y <= hi.y && p.z >= lo.z && p.z <= hi.z;
    }
    void bounding_sphere(V3 *c, float *r) const {
        *c = centroid();
        *r = empty() ? 0.f : dist(*c, hi);
    }
    const V3 &corner(int i) const { return i == 0 ? lo : hi; }
};
inline Box join(const Box &b, V3 p) { return {vmin(b.lo, p), vmax(b.hi, p)}; }
inline Box join(const Box &a, const Box &b) { return {vmin(a.lo, b.lo), vmax(a.hi, b.hi)}; }

// Row-major 3x4 affine transform applied as the reference Transform::Point/Vector do
// (math.h:291-304; the implicit last row is 0 0 0 1, so w == 1).
struct Affine {
    float m[12];
    V3 point(V3 p) const {
        float x = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
        float y = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
        float z = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
        float w = 0 * p.x + 0 * p.y + 0 * p.z + 1;
        return w == 1 ? V3(x, y, z) : V3(x, y, z) / w;
    }
    V3 vector(V3 v) const {
        return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
                m[8] * v.x + m[9] * v.y + m[10] * v.z};
    }
};

// Rotation about an axis (math.cpp:71-86), as used by the light-BVH cone union.
inline V3 rotate_vector(float degrees, V3 axis, V3 v) {
    V3 a = normalize(axis);
    float s = std::sin(deg2rad(degrees)), c = std::cos(deg2rad(degrees));
    float m00 = a.x * a.x + (1 - a.x * a.x) * c;
    float m01 = a.x * a.y * (1 - c) - a.z * s;
    float m02 = a.x * a.z * (1 - c) + a.y * s;
    float m10 = a.x * a.y * (1 - c) + a.z * s;
    float m11 = a.y * a.y + (1 - a.y * a.y) * c;
    float m12 = a.y * a.z * (1 - c) - a.x * s;
    float m20 = a.x * a.z * (1 - c) - a.y * s;
    float m21 = a.y * a.z * (1 - c) + a.x * s;
    float m22 = a.z * a.z + (1 - a.z * a.z) * c;
    return {m00 * v.x + m01 * v.y + m02 * v.z, m10 * v.x + m11 * v.y + m12 * v.z, m20 * v.x + m21 * v.y + m22 * v.z};
}

// ---- random numbers (src/core/rng.h:11-105) ----------------------------------------------------
inline uint64_t mix_bits(uint64_t v) {  // rng.h:21-28
    v ^= (v >> 31);
    v *= 0x7fb5d329728ea185ULL;
    v ^= (v >> 27);
    v *= 0x81dadef4bc2dd44dULL;
    v ^= (v >> 33);
    return v;
}

struct Pcg32 {  // rng.h:11-64 (O'Neill PCG32 XSH-RR)
    uint64_t state = 0x853c49e6748fea9bULL, inc = 0xda3e39cb94b95bdbULL;
    void set_sequence(uint64_t seq, uint64_t offset) {
        state = 0u;
        inc = (seq << 1u) | 1u;
        next();
        state += offset;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    float uniform() { return std::min(kOneMinusEps, float(next()) * 0x1p-32f); }
};

inline uint64_t murmur64a(const unsigned char *key, size_t n, uint64_t seed) {  // rng.h:67-97
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    const int r = 47;
    uint64_t h = seed ^ (n * m);
    const unsigned char *end = key + 8 * (n / 8);
    for (; key != end; key += 8) {
        uint64_t k;
        std::memcpy(&k, key, 8);
        k *= m;
        k ^= k >> r;
        k *= m;
        h ^= k;
        h *= m;
    }
    int rem = int(n & 7);
    for (int i = rem - 1; i >= 0; --i) h ^= uint64_t(key[i]) << (8 * i);
    if (rem) h *= m;
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}

// Hash of the raw bytes of the arguments, packed back to back (rng.h:99-105).
template <class... A> inline uint64_t hash_args(A... args) {
    unsigned char buf[(sizeof(A) + ... + 0)];
    size_t off = 0;
    ((std::memcpy(buf + off, &args, sizeof(A)), off += sizeof(A)), ...);
    return murmur64a(buf, sizeof(buf), 0);
}

// The per-sample stream of IndependentSampler::StartPixelSample (src/integrators/integrator.h:29-38).
struct Sampler {
    Pcg32 rng;
    void start(uint64_t pixel, uint64_t sampleIndex, uint64_t seed) {
        rng.set_sequence(hash_args(pixel, seed), mix_bits(sampleIndex * 0x9E3779B97F4A7C15ull + 1));
    }
    float get1() { return rng.uniform(); }
    V2 get2() {
        float a = get1();
        float b = get1();
        return {a, b};
    }
};

// ---- sampled spectra and hero wavelengths (src/core/spectrum.h:12-125) ---------------------------
struct SS {
    float v[4] = {0, 0, 0, 0};
    SS() = default;
    explicit SS(float c) { v[0] = v[1] = v[2] = v[3] = c; }
    float operator[](int i) const { return v[i]; }
    float &operator[](int i) { return v[i]; }
    bool nonzero() const { return v[0] != 0 || v[1] != 0 || v[2] != 0 || v[3] != 0; }
    float maxc() const { return std::max(std::max(v[0], v[1]), std::max(v[2], v[3])); }
    float avg() const { return (v[0] + v[1] + v[2] + v[3]) / 4; }
    bool has_nan() const {
        for (float x : v)
            if (std::isnan(x) || std::isinf(x)) return true;
        return false;
    }
};
inline SS operator+(const SS &a, const SS &b) {
    SS r;
    for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] + b.v[i];
    return r;
}
inline SS operator-(const SS &a, const SS &b) {
    SS r;
    for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] - b.v[i];
    return r;
}
inline SS operator*(const SS &a, const SS &b) {
    SS r;
    for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] * b.v[i];
    return r;
}
inline SS operator*(const SS &a, float s) {
    SS r;
    for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] * s;
    return r;
}
inline SS operator/(const SS &a, const SS &b) {  // spectrum.h:50-54: x/0 = 0
    SS r;
    for (int i = 0; i < 4; ++i) r.v[i] = b.v[i] != 0 ? a.v[i] / b.v[i] : 0.f;
    return r;
}
inline SS operator/(const SS &a, float s) { return a * (1.f / s); }  // spectrum.h:56
inline SS ss_exp(const SS &s) {
    SS r;
    for (int i = 0; i < 4; ++i) r.v[i] = std::exp(s.v[i]);
    return r;
}
inline SS ss_sqrt(const SS &s) {
    SS r;
    for (int i = 0; i < 4; ++i) r.v[i] = safe_sqrt(s.v[i]);
    return r;
}
inline SS ss_clamp0(const SS &s) {
    SS r;
    for (int i = 0; i < 4; ++i) r.v[i] = std::max(0.f, s.v[i]);
    return r;
}

struct Lambda {
    float l[4] = {0, 0, 0, 0}, pdf[4] = {0, 0, 0, 0};
    SS pdfs() const {
        SS r;
        for (int i = 0; i < 4; ++i) r.v[i] = pdf[i];
        return r;
    }
    bool secondary_terminated() const { return pdf[1] == 0 && pdf[2] == 0 && pdf[3] == 0; }
    void terminate_secondary() {  // spectrum.h:111-115
        if (secondary_terminated()) return;
        pdf[1] = pdf[2] = pdf[3] = 0;
        pdf[0] /= 4;
    }
};

inline float sample_visible_wavelength(float u) {  // spectrum.cpp:11-13
    return 538 - 138.888889f * std::atanh(0.85691062f - 1.82750197f * u);
}
inline float visible_wavelength_pdf(float l) {  // spectrum.cpp:15-18
    if (l < kLambdaMin || l > kLambdaMax) return 0;
    return 0.0039398042f / sq(std::cosh(0.0072f * (l - 538)));
}
inline Lambda sample_visible(float u) {  // spectrum.cpp:32-41
    Lambda w;
    for (int i = 0; i < 4; ++i) {
        float up = u + float(i) / 4;
        if (up > 1) up -= 1;
        w.l[i] = clampv(sample_visible_wavelength(up), kLambdaMin, kLambdaMax);
        w.pdf[i] = visible_wavelength_pdf(w.l[i]);
    }
    return w;
}

// CIE 1931 multi-lobe fit (Wyman, Sloan & Shirley 2013; spectrum.cpp:47-57).
inline float cie_g(float x, float mu, float s1, float s2) {
    float t = (x - mu) / (x < mu ? s1 : s2);
    return std::exp(-0.5f * t * t);
}
inline float cie_x(float l) {
    return 1.056f * cie_g(l, 599.8f, 37.9f, 31.0f) + 0.362f * cie_g(l, 442.0f, 16.0f, 26.7f) -
           0.065f * cie_g(l, 501.1f, 20.4f, 26.2f);
}
inline float cie_y(float l) { return 0.821f * cie_g(l, 568.8f, 46.9f, 40.5f) + 0.286f * cie_g(l, 530.9f, 16.3f, 31.1f); }
inline float cie_z(float l) { return 1.217f * cie_g(l, 437.0f, 11.8f, 36.0f) + 0.681f * cie_g(l, 459.0f, 26.0f, 13.8f); }
inline float cie_y_integral() {  // spectrum.cpp:59-66
    static const float v = [] {
        double s = 0;
        for (int l = int(kLambdaMin); l <= int(kLambdaMax); ++l) s += cie_y(float(l));
        return float(s);
    }();
    return v;
}

struct XYZ {
    float x = 0, y = 0, z = 0;
};
inline XYZ to_xyz(const SS &s, const Lambda &w) {  // spectrum.cpp:68-80
    double X = 0, Y = 0, Z = 0;
    for (int i = 0; i < 4; ++i) {
        if (w.pdf[i] == 0) continue;
        float v = s.v[i] / w.pdf[i];
        X += cie_x(w.l[i]) * v;
        Y += cie_y(w.l[i]) * v;
        Z += cie_z(w.l[i]) * v;
    }
    float norm = 1.f / (4 * cie_y_integral());
    return {float(X) * norm, float(Y) * norm, float(Z) * norm};
}

// ---- sampling warps (src/core/sampling.h:10-86) ------------------------------------------------
inline V2 sample_disk_concentric(V2 u) {
    float ox = 2 * u.x - 1, oy = 2 * u.y - 1;
    if (ox == 0 && oy == 0) return {0, 0};
    float theta, r;
    if (std::abs(ox) > std::abs(oy)) {
        r = ox;
        theta = kPiOver4 * (oy / ox);
    } else {
        r = oy;
        theta = kPiOver2 - kPiOver4 * (ox / oy);
    }
    return {r * std::cos(theta), r * std::sin(theta)};
}
inline V3 sample_cosine_hemisphere(V2 u) {
    V2 d = sample_disk_concentric(u);
    float z = safe_sqrt(1 - d.x * d.x - d.y * d.y);
    return {d.x, d.y, z};
}
inline V3 sample_uniform_sphere(V2 u) {
    float z = 1 - 2 * u.x;
    float r = safe_sqrt(1 - z * z);
    float phi = 2 * kPi * u.y;
    return {r * std::cos(phi), r * std::sin(phi), z};
}
inline void sample_uniform_triangle(V2 u, float b[3]) {
    float b0, b1;
    if (u.x < u.y) {
        b0 = u.x / 2;
        b1 = u.y - b0;
    } else {
        b1 = u.y / 2;
        b0 = u.x - b1;
    }
    b[0] = b0;
    b[1] = b1;
    b[2] = 1 - b0 - b1;
}
inline float power_heuristic(float fPdf, float gPdf) {  // sampling.h:81-86 with nf = ng = 1
    float f = fPdf, g = gPdf;
    if (std::isinf(sq(f))) return 1;
    if (f == 0 && g == 0) return 0;
    return sq(f) / (sq(f) + sq(g));
}

}  // namespace orc
