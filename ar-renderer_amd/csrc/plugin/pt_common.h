// Shading arithmetic of the MI355X path tracer, compiled for both the plugin's host code (light
// BVH construction) and the gfx950 kernels. Functions follow the reference CPU path tracer's
// definitions (cited per function) with its operation order, so the per-sample radiance agrees
// with the CPU to float round-off; the file is built without FMA contraction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "prender_gpu.h"

#define PG_HD __host__ __device__ __forceinline__

namespace pgpu {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv2Pi = 0.15915494309189533577f;
constexpr float kInv4Pi = 0.07957747154594766788f;
constexpr float kPiOver2 = 1.57079632679489661923f;
constexpr float kPiOver4 = 0.78539816339744830961f;
constexpr float kOneMinusEps = 0x1.fffffep-1f;
constexpr float kInfF = __builtin_huge_valf();

template <class T> PG_HD T sq(T v) { return v * v; }
PG_HD float minf(float a, float b) { return (b < a) ? b : a; }  // std::min semantics
PG_HD float maxf(float a, float b) { return (a < b) ? b : a; }  // std::max semantics
PG_HD float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
PG_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
PG_HD float lerp(float t, float a, float b) { return (1 - t) * a + t * b; }
PG_HD float safe_sqrt(float x) { return sqrtf(maxf(0.f, x)); }
PG_HD float safe_asin(float x) { return asinf(clampf(x, -1.f, 1.f)); }
PG_HD float safe_acos(float x) { return acosf(clampf(x, -1.f, 1.f)); }
// Relative intensity of a goniometric light's IES record (layout: prender_gpu.h, kIesHeader) in the
// direction w given in world space (src/lights/ies.cpp:80-105, src/lights/sky.cpp:12-27).
PG_HD int ies_interval(const float *a, int n, float x) {  // FindInterval (src/core/math.h:45-54)
    int first = 1, len = n - 2;
    while (len > 0) {
        int half = len >> 1, middle = first + half;
        bool r = a[middle] <= x;
        first = r ? middle + 1 : first;
        len = r ? len - (half + 1) : half;
    }
    return clampi(first - 1, 0, n - 2);
}
PG_HD float ies_eval(const float *rec, float wx, float wy, float wz) {
    // to the luminaire frame
    float lx = wx * rec[0] + wy * rec[1] + wz * rec[2];
    float ly = wx * rec[3] + wy * rec[4] + wz * rec[5];
    float lz = wx * rec[6] + wy * rec[7] + wz * rec[8];
    const int nV = int(rec[9]), nH = int(rec[10]);
    const float *V = rec + prgpu::kIesHeader, *H = V + nV, *C = H + nH;
    float theta = (180.f / kPi) * acosf(clampf(lz, -1.f, 1.f));
    float phi = (180.f / kPi) * atan2f(ly, lx);
    if (theta < V[0] || theta > V[nV - 1]) return 0;
    float hLast = H[nH - 1];
    phi = fmodf(phi, 360.f);
    if (phi < 0) phi += 360;
    if (nH == 1 || hLast == 0) phi = 0;
    else if (hLast <= 90) {
        phi = fmodf(phi, 180.f);
        if (phi > 90) phi = 180 - phi;
    } else if (hLast <= 180) {
        if (phi > 180) phi = 360 - phi;
    }
    int vi = 0;
    float tv = 0;
    if (nV > 1) {
        vi = ies_interval(V, nV, theta);
        tv = clampf((theta - V[vi]) / maxf(1e-6f, V[vi + 1] - V[vi]), 0.f, 1.f);
    }
    auto vertical = [&](int hi) {
        const float *row = C + size_t(hi) * size_t(nV);
        return nV == 1 ? row[0] : lerp(tv, row[vi], row[vi + 1]);
    };
    if (nH == 1) return vertical(0);
    int hi = ies_interval(H, nH, phi);
    float th = clampf((phi - H[hi]) / maxf(1e-6f, H[hi + 1] - H[hi]), 0.f, 1.f);
    return lerp(th, vertical(hi), vertical(hi + 1));
}

// Hermite smoothstep of x between a and b, a step at a when a == b (src/core/math.h:32-36).
PG_HD float smooth_step(float x, float a, float b) {
    if (a == b) return x < a ? 0.f : 1.f;
    float t = clampf((x - a) / (b - a), 0.f, 1.f);
    return t * t * (3 - 2 * t);
}
PG_HD bool is_finite(float x) { return __builtin_isfinite(x); }
PG_HD bool is_inf(float x) { return __builtin_isinf(x); }
PG_HD bool is_nan(float x) { return __builtin_isnan(x); }

// ---- vectors -----------------------------------------------------------------------------------
struct V3 {
    float x, y, z;
    PG_HD V3() : x(0), y(0), z(0) {}
    PG_HD V3(float a, float b, float c) : x(a), y(b), z(c) {}
    PG_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    PG_HD bool zero() const { return x == 0 && y == 0 && z == 0; }
};
PG_HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PG_HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PG_HD V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
PG_HD V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
PG_HD V3 operator/(V3 a, float s) {
    float inv = 1.f / s;
    return {a.x * inv, a.y * inv, a.z * inv};
}
PG_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PG_HD float absdot(V3 a, V3 b) { return fabsf(dot(a, b)); }
PG_HD V3 cross_d(V3 a, V3 b) {
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return {float(ay * bz - az * by), float(az * bx - ax * bz), float(ax * by - ay * bx)};
}
PG_HD V3 cross_f(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
PG_HD float len2(V3 v) { return dot(v, v); }
PG_HD float len(V3 v) { return sqrtf(len2(v)); }
PG_HD V3 normalize(V3 v) { return v / len(v); }
PG_HD float dist2(V3 a, V3 b) { return len2(a - b); }
PG_HD float dist(V3 a, V3 b) { return len(a - b); }
PG_HD V3 vmin(V3 a, V3 b) { return {minf(a.x, b.x), minf(a.y, b.y), minf(a.z, b.z)}; }
PG_HD V3 vmax(V3 a, V3 b) { return {maxf(a.x, b.x), maxf(a.y, b.y), maxf(a.z, b.z)}; }
PG_HD float max_comp(V3 v) { return maxf(v.x, maxf(v.y, v.z)); }
PG_HD V3 vabs(V3 v) { return {fabsf(v.x), fabsf(v.y), fabsf(v.z)}; }
PG_HD V3 face_forward(V3 n, V3 v) { return dot(n, v) < 0.f ? -n : n; }
PG_HD V3 f3(const prgpu::Float3 &f) { return {f.x, f.y, f.z}; }

PG_HD void coordinate_system(V3 v1, V3 *v2, V3 *v3) {  // src/core/math.h:120-126
    float sign = copysignf(1.f, v1.z);
    float a = -1 / (sign + v1.z);
    float b = v1.x * v1.y * a;
    *v2 = V3(1 + sign * v1.x * v1.x * a, sign * b, -sign * v1.x);
    *v3 = V3(b, sign + v1.y * v1.y * a, -v1.y);
}

struct Frame {  // src/core/math.h:128-150
    V3 x, y, z;
    PG_HD static Frame from_z(V3 zz) {
        Frame f;
        f.z = zz;
        coordinate_system(zz, &f.x, &f.y);
        return f;
    }
    PG_HD static Frame from_xz(V3 xin, V3 zz) {
        V3 xx = xin - zz * dot(xin, zz);
        float l2 = len2(xx);
        if (l2 < 1e-12f || !is_finite(l2)) return from_z(zz);
        xx = xx / sqrtf(l2);
        Frame f;
        f.z = zz;
        f.x = xx;
        f.y = cross_d(zz, xx);
        return f;
    }
    PG_HD V3 to_local(V3 v) const { return {dot(v, x), dot(v, y), dot(v, z)}; }
    PG_HD V3 from_local(V3 v) const { return x * v.x + y * v.y + z * v.z; }
};

PG_HD float cos2t(V3 w) { return w.z * w.z; }
PG_HD float abscos(V3 w) { return fabsf(w.z); }
PG_HD float sin2t(V3 w) { return maxf(0.f, 1 - cos2t(w)); }
PG_HD float sint(V3 w) { return sqrtf(sin2t(w)); }
PG_HD float tan2t(V3 w) { return sin2t(w) / cos2t(w); }
PG_HD float cosphi(V3 w) {
    float s = sint(w);
    return s == 0 ? 1 : clampf(w.x / s, -1.f, 1.f);
}
PG_HD float sinphi(V3 w) {
    float s = sint(w);
    return s == 0 ? 0 : clampf(w.y / s, -1.f, 1.f);
}
PG_HD bool same_hemi(V3 a, V3 b) { return a.z * b.z > 0; }
PG_HD V3 spherical_dir(float sinT, float cosT, float phi) {
    return {clampf(sinT, -1.f, 1.f) * cosf(phi), clampf(sinT, -1.f, 1.f) * sinf(phi), clampf(cosT, -1.f, 1.f)};
}
PG_HD V3 reflect(V3 wo, V3 n) { return -wo + n * (2 * dot(wo, n)); }
PG_HD bool refract(V3 wi, V3 n, float eta, float *etap, V3 *wt) {  // src/core/math.h:183-197
    float ci = dot(n, wi);
    if (ci < 0) {
        eta = 1 / eta;
        ci = -ci;
        n = -n;
    }
    float s2i = maxf(0.f, 1 - sq(ci));
    float s2t = s2i / sq(eta);
    if (s2t >= 1) return false;
    float ct = safe_sqrt(1 - s2t);
    *wt = -wi / eta + n * (ci / eta - ct);
    *etap = eta;
    return true;
}

struct Box3 {
    V3 lo, hi;
    PG_HD Box3() : lo(kInfF, kInfF, kInfF), hi(-kInfF, -kInfF, -kInfF) {}
    PG_HD Box3(V3 a, V3 b) : lo(a), hi(b) {}
    PG_HD bool empty() const { return lo.x > hi.x || lo.y > hi.y || lo.z > hi.z; }
    PG_HD V3 diag() const { return hi - lo; }
    PG_HD V3 centroid() const { return (lo + hi) * 0.5f; }
    PG_HD float area() const {
        if (empty()) return 0;
        V3 d = diag();
        return 2 * (d.x * d.y + d.x * d.z + d.y * d.z);
    }
    PG_HD V3 offset(V3 p) const {
        V3 o = p - lo;
        if (hi.x > lo.x) o.x /= hi.x - lo.x;
        if (hi.y > lo.y) o.y /= hi.y - lo.y;
        if (hi.z > lo.z) o.z /= hi.z - lo.z;
        return o;
    }
    PG_HD bool inside(V3 p) const {
        return p.x >= lo.x && p.x <= hi.x && p.y >= lo.y && p.y <= hi.y && p.z >= lo.z && p.z <= hi.z;
    }
};
PG_HD Box3 join(const Box3 &a, const Box3 &b) { return {vmin(a.lo, b.lo), vmax(a.hi, b.hi)}; }
PG_HD Box3 join(const Box3 &a, V3 p) { return {vmin(a.lo, p), vmax(a.hi, p)}; }

// ---- random numbers ----------------------------------------------------------------------------
PG_HD uint64_t mix_bits(uint64_t v) {  // src/core/rng.h:21-28
    v ^= (v >> 31);
    v *= 0x7fb5d329728ea185ULL;
    v ^= (v >> 27);
    v *= 0x81dadef4bc2dd44dULL;
    v ^= (v >> 33);
    return v;
}

struct Pcg {  // PCG32 XSH-RR (src/core/rng.h:11-64)
    uint64_t state, inc;
    PG_HD void set_sequence(uint64_t seq, uint64_t offset) {
        state = 0u;
        inc = (seq << 1u) | 1u;
        next();
        state += offset;
        next();
    }
    PG_HD uint32_t next() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    PG_HD float uniform() { return minf(kOneMinusEps, float(next()) * 0x1p-32f); }
};

// MurmurHash64A over n 64-bit words plus a 4-byte tail (src/core/rng.h:67-105).
PG_HD uint64_t murmur_words(const uint64_t *w, int nWords, bool tail4, uint32_t tail) {
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    const int r = 47;
    uint64_t n = uint64_t(nWords) * 8 + (tail4 ? 4 : 0);
    uint64_t h = 0 ^ (n * m);
    for (int i = 0; i < nWords; ++i) {
        uint64_t k = w[i];
        k *= m;
        k ^= k >> r;
        k *= m;
        h ^= k;
        h *= m;
    }
    if (tail4) {
        h ^= uint64_t(tail);
        h *= m;
    }
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}
PG_HD uint64_t fbits(float a, float b) { return uint64_t(__float_as_uint(a)) | (uint64_t(__float_as_uint(b)) << 32); }

// ---- spectra (src/core/spectrum.cpp, src/core/rgb2spec.h) --------------------------------------------
// Planck's law in double precision as the CPU evaluates it (spectrum.cpp:116-122), with l^5 as four
// multiplications instead of pow: within 2 ulp of the double power, so the float result is the
// CPU's, and the light's spectrum, evaluated at every next-event estimate, no longer pays a
// double-precision pow.
PG_HD float planck(float lambdaNm, float T) {
    if (T <= 0) return 0;
    const double c = 299792458., h = 6.62606957e-34, kb = 1.3806488e-23;
    const double l = lambdaNm * 1e-9, l2 = l * l;
    return float((2 * h * c * c) / (l2 * l2 * l * (exp((h * c) / (l * kb * T)) - 1)));
}

// ---- image environment light (record layout: prgpu::kEnvHeader) ----------------------------------
// The map's lookup (src/lights/light.cpp:249-261, light.h:287-297), the RGB-to-spectrum table
// lookup (src/core/rgb2spec.cpp:278-304) and the piecewise-constant 2D distribution over the map
// (src/core/sampling.cpp:24-67), as float operations in the reference's order.
struct EnvMap {
    const float *hdr;
    int w, h;
    const float *rgb, *func, *cond, *rowInt, *marg;
};
PG_HD EnvMap env_map(const float *rec) {
    EnvMap e;
    e.hdr = rec;
    e.w = int(rec[0]);
    e.h = int(rec[1]);
    e.rgb = rec + prgpu::kEnvHeader;
    e.func = e.rgb + 3 * size_t(e.w) * size_t(e.h);
    e.cond = e.func + size_t(e.w) * size_t(e.h);
    e.rowInt = e.cond + size_t(e.h) * size_t(e.w + 1);
    e.marg = e.rowInt + e.h;
    return e;
}
// World direction to the map's frame (Rotate(r, +y) inverted) and back.
PG_HD V3 env_to_light(const EnvMap &e, V3 w) { return V3(e.hdr[2] * w.x - e.hdr[3] * w.z, w.y, e.hdr[3] * w.x + e.hdr[2] * w.z); }
PG_HD V3 env_to_world(const EnvMap &e, V3 w) { return V3(e.hdr[2] * w.x + e.hdr[3] * w.z, w.y, -e.hdr[3] * w.x + e.hdr[2] * w.z); }
// DirToEquirect of a unit direction: (phi / 2 pi, theta / pi), theta from +y.
PG_HD void dir_to_equirect(V3 d, float *u, float *v) {
    float theta = safe_acos(d.y);
    float phi = atan2f(d.z, d.x);
    if (phi < 0) phi += 2 * kPi;
    *u = phi * kInv2Pi;
    *v = theta * kInvPi;
}
PG_HD V3 equirect_to_dir(float u, float v) {
    float theta = v * kPi, phi = u * 2 * kPi;
    float s = sinf(theta);
    return V3(s * cosf(phi), cosf(theta), s * sinf(phi));
}
// Bilinear RGB with horizontal wrap and vertical clamp at map coordinates (u, v).
PG_HD void env_rgb(const EnvMap &e, float u, float v, float out[3]) {
    float x = u * float(e.w) - 0.5f, y = v * float(e.h) - 0.5f;
    int x0 = int(floorf(x)), y0 = int(floorf(y));
    float dx = x - float(x0), dy = y - float(y0);
    auto tex = [&](int xi, int yi) {
        xi = ((xi % e.w) + e.w) % e.w;
        yi = clampi(yi, 0, e.h - 1);
        return e.rgb + 3 * (size_t(yi) * size_t(e.w) + size_t(xi));
    };
    const float *a = tex(x0, y0), *b = tex(x0 + 1, y0), *c = tex(x0, y0 + 1), *d = tex(x0 + 1, y0 + 1);
    for (int i = 0; i < 3; ++i)
        out[i] = (1 - dx) * (1 - dy) * a[i] + dx * (1 - dy) * b[i] + (1 - dx) * dy * c[i] + dx * dy * d[i];
}
// Sigmoid-polynomial coefficients of an RGB in [0, 1]^3 from the table record (resolution n, z
// nodes, coefficients [max component][z][y][x][3]).
PG_HD void rgb_sigmoid_coeffs(const float *table, const float rgbIn[3], float c[3]) {
    float rgb[3] = {clampf(rgbIn[0], 0.f, 1.f), clampf(rgbIn[1], 0.f, 1.f), clampf(rgbIn[2], 0.f, 1.f)};
    if (rgb[0] == rgb[1] && rgb[1] == rgb[2]) {
        float v = rgb[0];
        c[0] = c[1] = 0;
        c[2] = v <= 0 ? -__builtin_huge_valf() : (v >= 1 ? __builtin_huge_valf() : (v - .5f) / sqrtf(v * (1 - v)));
        return;
    }
    const int n = int(table[0]);
    const float *zn = table + 1, *co = zn + n;
    int maxc = (rgb[0] > rgb[1]) ? ((rgb[0] > rgb[2]) ? 0 : 2) : ((rgb[1] > rgb[2]) ? 1 : 2);
    float z = rgb[maxc];
    float x = rgb[(maxc + 1) % 3] * float(n - 1) / z;
    float y = rgb[(maxc + 2) % 3] * float(n - 1) / z;
    int xi = int(x) < n - 2 ? int(x) : n - 2, yi = int(y) < n - 2 ? int(y) : n - 2;
    int first = 1, len = n - 2;  // FindInterval(n, zNodes[i] < z)
    while (len > 0) {
        int half = len >> 1, middle = first + half;
        bool r = zn[middle] < z;
        first = r ? middle + 1 : first;
        len = r ? len - (half + 1) : half;
    }
    int zi = clampi(first - 1, 0, n - 2);
    float dx = x - float(xi), dy = y - float(yi), dz = (z - zn[zi]) / (zn[zi + 1] - zn[zi]);
    auto at = [&](int zz, int yy, int xx, int i) {
        return co[3 * (((size_t(maxc) * size_t(n) + size_t(zz)) * size_t(n) + size_t(yy)) * size_t(n) + size_t(xx)) + i];
    };
    for (int i = 0; i < 3; ++i)
        c[i] = lerp(dz, lerp(dy, lerp(dx, at(zi, yi, xi, i), at(zi, yi, xi + 1, i)),
                             lerp(dx, at(zi, yi + 1, xi, i), at(zi, yi + 1, xi + 1, i))),
                    lerp(dy, lerp(dx, at(zi + 1, yi, xi, i), at(zi + 1, yi, xi + 1, i)),
                         lerp(dx, at(zi + 1, yi + 1, xi, i), at(zi + 1, yi + 1, xi + 1, i))));
}
// ---- image textures (record layout: prgpu::kTexHeader) ------------------------------------------
// ImageData::Lookup with the texture's UVMapping (src/materials/texture.cpp:8-36, texture.h:61-64):
// bilinear RGB at the mapped (u, v), texel centres at half-integers, the wrap mode on both axes.
PG_HD int tex_wrap(int v, int n, int mode) {
    if (mode == 0) {
        v %= n;
        return v < 0 ? v + n : v;
    }
    if (mode == 1) return clampi(v, 0, n - 1);
    const int period = 2 * n;
    v %= period;
    if (v < 0) v += period;
    return v < n ? v : period - 1 - v;
}
PG_HD void tex_rgb(const float *rec, float u, float v, float out[3]) {
    const int w = int(rec[0]), h = int(rec[1]), mode = int(rec[2]);
    const float su = u * rec[4] + rec[6], sv = v * rec[5] + rec[7];
    const float vv = rec[3] != 0 ? 1 - sv : sv;
    // (coordinates beyond +-2^24 texels are clamped before the integer conversion)
    const float x = clampf(su * float(w) - 0.5f, -16777216.f, 16777216.f);
    const float y = clampf(vv * float(h) - 0.5f, -16777216.f, 16777216.f);
    const int x0 = int(floorf(x)), y0 = int(floorf(y));
    const float dx = x - float(x0), dy = y - float(y0);
    const float *px = rec + prgpu::kTexHeader;
    auto texel = [&](int xi, int yi) {
        return px + 3 * (size_t(tex_wrap(yi, h, mode)) * size_t(w) + size_t(tex_wrap(xi, w, mode)));
    };
    const float *a = texel(x0, y0), *b = texel(x0 + 1, y0), *c = texel(x0, y0 + 1), *d = texel(x0 + 1, y0 + 1);
    for (int i = 0; i < 3; ++i)
        out[i] = (1 - dx) * (1 - dy) * a[i] + dx * (1 - dy) * b[i] + (1 - dx) * dy * c[i] + dx * dy * d[i];
}
// ImageFloatTexture::Evaluate (texture.h:111-116): one channel, or the mean of the three, times the
// texture's multiplier.
PG_HD float tex_float(const float *rec, int channel, float u, float v) {
    float c[3];
    tex_rgb(rec, u, v, c);
    // a select rather than c[channel]: a run-time index would put c in scratch memory
    const float x = channel < 0 ? (c[0] + c[1] + c[2]) / 3 : (channel == 0 ? c[0] : (channel == 1 ? c[1] : c[2]));
    return x * rec[8];
}

// The sigmoid polynomial at wavelength l (RGBSigmoidPolynomial, src/core/rgb2spec.h:13-27).
PG_HD float sigmoid_poly(const float c[3], float l) {
    float x = fmaf(l, fmaf(l, c[0], c[1]), c[2]);
    if (is_inf(x)) return x > 0 ? 1.f : 0.f;
    return 0.5f + x / (2 * sqrtf(1 + x * x));
}
// Distribution1D::SampleContinuous over n cells with its CDF (n + 1 entries) and integral.
PG_HD float env_sample_cells(const float *f, const float *cdf, int n, float integral, float u, float *pdf, int *cell) {
    int o = ies_interval(cdf, n + 1, u);
    float du = u - cdf[o];
    if (cdf[o + 1] - cdf[o] > 0) du /= (cdf[o + 1] - cdf[o]);
    *pdf = integral > 0 ? f[o] / integral : 1.f;
    *cell = o;
    return minf((float(o) + du) / float(n), kOneMinusEps);
}
// Distribution2D::SampleContinuous: the row by the marginal (u1), then the column (u0).
PG_HD void env_sample_uv(const EnvMap &e, float u0, float u1, float *u, float *v, float *pdf) {
    float pv, pu;
    int row, col;
    *v = env_sample_cells(e.rowInt, e.marg, e.h, e.hdr[4], u1, &pv, &row);
    *u = env_sample_cells(e.func + size_t(row) * size_t(e.w), e.cond + size_t(row) * size_t(e.w + 1), e.w, e.rowInt[row],
                          u0, &pu, &col);
    *pdf = pu * pv;
}
// Distribution2D::PDF at map coordinates (u, v).
PG_HD float env_pdf_uv(const EnvMap &e, float u, float v) {
    int iu = clampi(int(u * float(e.w)), 0, e.w - 1), iv = clampi(int(v * float(e.h)), 0, e.h - 1);
    if (e.hdr[4] == 0) return 0;
    return e.func[size_t(iv) * size_t(e.w) + size_t(iu)] / e.hdr[4];
}

struct SpecTable {
    const prgpu::SpectrumEx *s;
    const float *knots;
};

PG_HD float spec_base(const prgpu::SpectrumEx &s, const float *knots, float l) {
    switch (s.kind) {
    case prgpu::SpecConstant: return s.p[0];
    case prgpu::SpecSigmoid: {
        float x = fmaf(l, fmaf(l, s.p[0], s.p[1]), s.p[2]);
        if (is_inf(x)) return x > 0 ? 1.f : 0.f;
        return 0.5f + x / (2 * sqrtf(1 + x * x));
    }
    case prgpu::SpecBlackbody: return planck(l, s.p[0]) * s.p[1];
    case prgpu::SpecPiecewise: {
        const float *L = knots + s.offset, *V = L + s.count;
        int n = s.count;
        if (n == 0) return 0;
        if (l <= L[0]) return V[0];
        if (l >= L[n - 1]) return V[n - 1];
        int first = 1, len = n - 2;  // FindInterval (src/core/math.h:45-54)
        while (len > 0) {
            int half = len >> 1, mid = first + half;
            bool r = L[mid] <= l;
            first = r ? mid + 1 : first;
            len = r ? len - (half + 1) : half;
        }
        int o = clampi(first - 1, 0, n - 2);
        float t = (l - L[o]) / (L[o + 1] - L[o]);
        return lerp(t, V[o], V[o + 1]);
    }
    case prgpu::SpecSellmeier: {
        double l2 = double(l) * 1e-3;
        l2 = l2 * l2;
        double n2 = 1;
        for (int i = 0; i < s.count; ++i) n2 += s.p[i] * l2 / (l2 - s.p[3 + i]);
        return float(sqrt(n2 > 1.0 ? n2 : 1.0));
    }
    case prgpu::SpecCauchy: return s.p[0] + s.p[1] / (l * l);
    case prgpu::SpecTable95: {
        const float *v = knots + s.offset;
        float x = (l - prgpu::kSpectrumMin) / prgpu::kSpectrumStep;
        x = clampf(x, 0.f, float(prgpu::kSpectrumSamples - 1) - 1e-3f);
        int i = int(x);
        float t = x - float(i);
        return v[i] * (1 - t) + v[i + 1] * t;
    }
    }
    return 0;
}
PG_HD float spec_eval(const SpecTable &t, int id, float l) {
    if (id < 0) return 0;
    const prgpu::SpectrumEx &s = t.s[id];
    float b = spec_base(s, t.knots, l);
    return s.scale == 1.f ? b : s.scale * b;
}
PG_HD bool spec_is_constant(const SpecTable &t, int id) {
    const prgpu::SpectrumEx &s = t.s[id];
    return s.kind == prgpu::SpecConstant || (s.kind == prgpu::SpecCauchy && s.p[1] == 0);
}
PG_HD float spec_max(const SpecTable &t, int id) {
    const prgpu::SpectrumEx &s = t.s[id];
    float m = 0;
    switch (s.kind) {
    case prgpu::SpecConstant: m = s.p[0]; break;
    case prgpu::SpecSigmoid: {  // src/core/rgb2spec.cpp:24-31
        m = maxf(spec_base(s, t.knots, 360.f), spec_base(s, t.knots, 830.f));
        if (s.p[0] != 0) {
            float l = -s.p[1] / (2 * s.p[0]);
            if (l >= 360.f && l <= 830.f) m = maxf(m, spec_base(s, t.knots, l));
        }
        break;
    }
    case prgpu::SpecBlackbody: m = 1.f; break;
    case prgpu::SpecPiecewise:
        for (int i = 0; i < s.count; ++i) m = maxf(m, t.knots[s.offset + s.count + i]);
        break;
    case prgpu::SpecTable95:
        for (int i = 0; i < prgpu::kSpectrumSamples; ++i) m = maxf(m, t.knots[s.offset + i]);
        break;
    default: m = spec_base(s, t.knots, 360.f); break;
    }
    return s.scale == 1.f ? m : s.scale * m;
}

// CIE 1931 fit (src/core/spectrum.cpp:47-57).
PG_HD float cie_g(float x, float mu, float s1, float s2) {
    float t = (x - mu) / (x < mu ? s1 : s2);
    return expf(-0.5f * t * t);
}
PG_HD float cie_x(float l) {
    return 1.056f * cie_g(l, 599.8f, 37.9f, 31.0f) + 0.362f * cie_g(l, 442.0f, 16.0f, 26.7f) -
           0.065f * cie_g(l, 501.1f, 20.4f, 26.2f);
}
PG_HD float cie_y(float l) { return 0.821f * cie_g(l, 568.8f, 46.9f, 40.5f) + 0.286f * cie_g(l, 530.9f, 16.3f, 31.1f); }
PG_HD float cie_z(float l) { return 1.217f * cie_g(l, 437.0f, 11.8f, 36.0f) + 0.681f * cie_g(l, 459.0f, 26.0f, 13.8f); }

inline float host_spectrum_y(const SpecTable &t, int id, float cieYIntegral) {  // SpectrumToXYZ(...).y
    double Y = 0;
    for (int l = 360; l <= 830; ++l) Y += cie_y(float(l)) * spec_eval(t, id, float(l));
    return float(Y) * (1.f / cieYIntegral);
}
inline float host_cie_y_integral() {
    double s = 0;
    for (int l = 360; l <= 830; ++l) s += cie_y(float(l));
    return float(s);
}

// ---- light BVH node and importance (src/lights/light_bvh.cpp:45-83) -------------------------------
struct alignas(16) LNode {
    float lo[3], hi[3];
    float w[3];
    float phi, cosTheta_o, cosTheta_e;
    int32_t child1;  // interior: second child (the first follows the node); -1 for leaves
    int32_t light;   // leaf: light index; -1 for interior nodes
    int32_t parent;
    uint32_t twoSided;
};
static_assert(sizeof(LNode) == 64, "LNode");

PG_HD float lb_importance(const LNode &lb, V3 p, V3 n) {
    V3 lo(lb.lo[0], lb.lo[1], lb.lo[2]), hi(lb.hi[0], lb.hi[1], lb.hi[2]);
    V3 pc = (lo + hi) * 0.5f;
    float d2 = dist2(p, pc);
    d2 = maxf(d2, len(hi - lo) / 2);
    V3 wi = p - pc;
    float l = len(wi);
    wi = l > 0 ? wi / l : V3(0, 0, 1);
    float cos_w = dot(V3(lb.w[0], lb.w[1], lb.w[2]), wi);
    if (lb.twoSided) cos_w = fabsf(cos_w);
    float sin_w = safe_sqrt(1 - sq(cos_w));
    // Bounding sphere of the bounds: centre = centroid, radius = distance to hi (empty: 0).
    bool emptyB = lo.x > hi.x || lo.y > hi.y || lo.z > hi.z;
    V3 c = (lo + hi) * 0.5f;
    float r = emptyB ? 0.f : dist(c, hi);
    float cos_b = -1;
    float dc2 = dist2(p, c);
    if (dc2 >= r * r) cos_b = safe_sqrt(1 - r * r / dc2);
    float sin_b = safe_sqrt(1 - sq(cos_b));
    float sin_o = safe_sqrt(1 - sq(lb.cosTheta_o));
    float cos_x = (cos_w > lb.cosTheta_o) ? 1.f : cos_w * lb.cosTheta_o + sin_w * sin_o;
    float sin_x = (cos_w > lb.cosTheta_o) ? 0.f : sin_w * lb.cosTheta_o - cos_w * sin_o;
    float cos_p = (cos_x > cos_b) ? 1.f : cos_x * cos_b + sin_x * sin_b;
    if (cos_p <= lb.cosTheta_e) return 0;
    float imp = lb.phi * cos_p / d2;
    if (!n.zero()) {
        float cos_i = absdot(wi, n);
        float sin_i = safe_sqrt(1 - sq(cos_i));
        imp *= (cos_i > cos_b) ? 1.f : cos_i * cos_b + sin_i * sin_b;
    }
    return maxf(imp, 0.f);
}

}  // namespace pgpu
