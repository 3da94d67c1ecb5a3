// Oracle: CPU restatement of the reference CPU path tracer (pr::PathIntegrator and everything
// it calls), for tests only -- see pt_base.h. It consumes the same flat scene description
// (prgpu::SceneDescEx) that the GPU plugin consumes, builds its own binary SAH BVH and light
// BVH the way the reference does, and exposes a C API (oracle.h) used by tests/ and by the
// CPU baseline of bench.py.
#include "pt_base.h"
#include "oracle.h"

#include <atomic>
#include <chrono>
#include <complex>
#include <future>
#include <memory>
#include <optional>
#include <thread>
#include <unordered_map>
#include <vector>

namespace orc {

using prgpu::LightEx;
using prgpu::Material;
using prgpu::MediumEx;
using prgpu::Primitive;
using prgpu::SpectrumEx;

inline V3 v3(const prgpu::Float3 &f) { return {f.x, f.y, f.z}; }

// =============================================================================================
// Spectra (src/core/spectrum.cpp:116-136, src/core/rgb2spec.h:13-27, rgb2spec.cpp:24-31)

static float planck(float lambdaNm, float T) {  // spectrum.cpp:116-122
    if (T <= 0) return 0;
    const double c = 299792458., h = 6.62606957e-34, kb = 1.3806488e-23;
    double l = lambdaNm * 1e-9;
    double Le = (2 * h * c * c) / (std::pow(l, 5) * (std::exp((h * c) / (l * kb * T)) - 1));
    return float(Le);
}

static float sigmoid_s(float x) {  // rgb2spec.h:21-24
    if (std::isinf(x)) return x > 0 ? 1.f : 0.f;
    return 0.5f + x / (2 * std::sqrt(1 + x * x));
}

struct SpectrumTable {
    std::vector<SpectrumEx> specs;
    std::vector<float> knots;

    float base(const SpectrumEx &s, float l) const {
        switch (s.kind) {
        case prgpu::SpecConstant: return s.p[0];
        case prgpu::SpecSigmoid:  // EvaluatePolynomial(l, c2, c1, c0) = fma(l, fma(l, c0, c1), c2)
            return sigmoid_s(std::fma(l, std::fma(l, s.p[0], s.p[1]), s.p[2]));
        case prgpu::SpecBlackbody: return planck(l, s.p[0]) * s.p[1];
        case prgpu::SpecPiecewise: {  // spectrum.cpp:101-108
            const float *L = knots.data() + s.offset, *V = L + s.count;
            int n = s.count;
            if (n == 0) return 0;
            if (l <= L[0]) return V[0];
            if (l >= L[n - 1]) return V[n - 1];
            int o = find_interval(n, [&](int i) { return L[i] <= l; });
            float t = (l - L[o]) / (L[o + 1] - L[o]);
            return lerpf(t, V[o], V[o + 1]);
        }
        case prgpu::SpecSellmeier: {  // spectrum.cpp:131-136 (double precision)
            double l2 = sq(double(l) * 1e-3);
            double n2 = 1;
            for (int i = 0; i < s.count; ++i) n2 += s.p[i] * l2 / (l2 - s.p[3 + i]);
            return float(std::sqrt(std::max(n2, 1.0)));
        }
        case prgpu::SpecCauchy: return s.p[0] + s.p[1] / (l * l);
        case prgpu::SpecTable95: {
            const float *v = knots.data() + s.offset;
            float x = (l - prgpu::kSpectrumMin) / prgpu::kSpectrumStep;
            x = clampv(x, 0.f, float(prgpu::kSpectrumSamples - 1) - 1e-3f);
            int i = int(x);
            float t = x - float(i);
            return v[i] * (1 - t) + v[i + 1] * t;
        }
        }
        return 0;
    }
    // ScaledSpectrum: scale * base (spectrum.h:198-208); unscaled kinds carry scale 1.
    float eval(int id, float l) const {
        if (id < 0) return 0;
        const SpectrumEx &s = specs[size_t(id)];
        float b = base(s, l);
        return s.scale == 1.f ? b : s.scale * b;
    }
    SS sample(int id, const Lambda &w) const {
        SS r;
        for (int i = 0; i < 4; ++i) r.v[i] = eval(id, w.l[i]);
        return r;
    }
    bool is_constant(int id) const {
        const SpectrumEx &s = specs[size_t(id)];
        return s.kind == prgpu::SpecConstant || (s.kind == prgpu::SpecCauchy && s.p[1] == 0);
    }
    float max_value(int id) const {
        const SpectrumEx &s = specs[size_t(id)];
        float m = 0;
        switch (s.kind) {
        case prgpu::SpecConstant: m = s.p[0]; break;
        case prgpu::SpecSigmoid: {  // rgb2spec.cpp:24-31
            m = std::max(base(s, kLambdaMin), base(s, kLambdaMax));
            if (s.p[0] != 0) {
                float l = -s.p[1] / (2 * s.p[0]);
                if (l >= kLambdaMin && l <= kLambdaMax) m = std::max(m, base(s, l));
            }
            break;
        }
        case prgpu::SpecBlackbody: m = 1.f; break;
        case prgpu::SpecPiecewise:
            for (int i = 0; i < s.count; ++i) m = std::max(m, knots[size_t(s.offset + s.count + i)]);
            break;
        case prgpu::SpecTable95:
            for (int i = 0; i < prgpu::kSpectrumSamples; ++i) m = std::max(m, knots[size_t(s.offset + i)]);
            break;
        default: m = base(s, kLambdaMin); break;  // Sellmeier / Cauchy: value at LambdaMin
        }
        return s.scale == 1.f ? m : s.scale * m;
    }
    float to_y(int id) const {  // SpectrumToXYZ(...).y (spectrum.cpp:82-94)
        double Y = 0;
        for (int l = int(kLambdaMin); l <= int(kLambdaMax); ++l) Y += cie_y(float(l)) * eval(id, float(l));
        float inv = 1.f / cie_y_integral();
        return float(Y) * inv;
    }
};

// =============================================================================================
// Interactions (src/scene/interaction.h:20-77)

inline float ray_epsilon(V3 p) { return 5e-5f * std::max(1.f, max_comp(vabs(p))); }
constexpr float kShadowEps = 1e-4f;

struct Intr {
    V3 p, n, wo;
    V2 uv;
    bool surface_mi = false;  // mediumInterface set (surface hits)
    int mIn = -1, mOut = -1;  // medium interface
    int medium = -1;          // used when !surface_mi
    int get_medium(V3 w) const {
        if (surface_mi) return dot(w, n) > 0 ? mOut : mIn;
        return medium;
    }
    V3 offset_origin(V3 w) const {
        if (n.zero()) return p;
        float eps = ray_epsilon(p);
        V3 off = n * eps;
        if (dot(w, n) < 0) off = -off;
        return p + off;
    }
    Ray spawn(V3 d) const { return {offset_origin(d), d}; }
    Ray spawn_to(const Intr &it) const {
        V3 o = offset_origin(it.p - p);
        V3 target = it.offset_origin(o - it.p);
        return {o, target - o};
    }
};

struct SurfHit : Intr {
    V3 ns, dpdus, dpdu, dpdv;
    int prim = -1;
};

// =============================================================================================
// Shapes (src/geometry/shape.cpp)

struct HitInfo {
    float b1 = 0, b2 = 0;
};

struct Shape {
    const Primitive *p = nullptr;
    bool sphere() const { return p->type == prgpu::PrimSphere; }
    bool has_normals() const { return !(v3(p->n0).zero() && v3(p->n1).zero() && v3(p->n2).zero()); }

    Box bounds() const {
        if (sphere()) {
            V3 c = v3(p->p0);
            float r = p->r;
            return {c - V3(r, r, r), c + V3(r, r, r)};
        }
        V3 a = v3(p->p0), b = v3(p->p1), c = v3(p->p2);
        return join(Box{vmin(a, b), vmax(a, b)}, c);  // shape.cpp:186-190
    }
    float area() const {
        if (sphere()) return 4 * kPi * p->r * p->r;
        V3 a = v3(p->p0), b = v3(p->p1), c = v3(p->p2);
        return 0.5f * len(cross_d(b - a, c - a));  // shape.cpp:192-196
    }

    bool intersect(const Ray &ray, float tMax, float *tHit, HitInfo *hit) const {
        if (sphere()) {  // shape.cpp:42-61, double precision
            V3 c = v3(p->p0);
            float r = p->r;
            double ox = ray.o.x - c.x, oy = ray.o.y - c.y, oz = ray.o.z - c.z;
            double dx = ray.d.x, dy = ray.d.y, dz = ray.d.z;
            double a = dx * dx + dy * dy + dz * dz;
            double b = ox * dx + oy * dy + oz * dz;
            double cc = ox * ox + oy * oy + oz * oz - double(r) * r;
            double k = b / a;
            double lx = ox - k * dx, ly = oy - k * dy, lz = oz - k * dz;
            double disc = double(r) * r - (lx * lx + ly * ly + lz * lz);
            if (disc < 0) return false;
            double q = -b - std::copysign(std::sqrt(a * disc), b);
            double t0 = cc / q, t1 = q / a;
            if (t0 > t1) std::swap(t0, t1);
            double t = t0;
            if (t <= 0) t = t1;
            if (t <= 0 || t >= tMax) return false;
            *tHit = float(t);
            return true;
        }
        // Moller-Trumbore in float without contraction (shape.cpp:202-222).
        V3 p0 = v3(p->p0), p1 = v3(p->p1), p2 = v3(p->p2);
        V3 e1 = p1 - p0, e2 = p2 - p0;
        V3 pvec = cross_f(ray.d, e2);
        float det = dot(e1, pvec);
        if (det == 0 || !std::isfinite(det)) return false;
        float inv = 1 / det;
        V3 tvec = ray.o - p0;
        float u = dot(tvec, pvec) * inv;
        if (u < 0 || u > 1) return false;
        V3 qvec = cross_f(tvec, e1);
        float v = dot(ray.d, qvec) * inv;
        if (v < 0 || u + v > 1) return false;
        float t = dot(e2, qvec) * inv;
        if (t <= 0 || t >= tMax) return false;
        *tHit = t;
        hit->b1 = u;
        hit->b2 = v;
        return true;
    }

    void interaction(const Ray &ray, float tHit, const HitInfo &hit, SurfHit *si) const {
        if (sphere()) {  // shape.cpp:63-85
            V3 c = v3(p->p0);
            float r = p->r;
            V3 pHit = ray.at(tHit);
            V3 rel = pHit - c;
            float l = len(rel);
            V3 n = l > 0 ? rel / l : V3(0, 1, 0);
            pHit = c + n * r;
            float phi = std::atan2(n.z, n.x);
            if (phi < 0) phi += 2 * kPi;
            float theta = safe_acos(n.y);
            si->p = pHit;
            si->n = p->flip ? -n : n;
            si->uv = {phi * kInv2Pi, theta * kInvPi};
            si->wo = -normalize(ray.d);
            V3 dpdu = V3(-rel.z, 0, rel.x) * (2 * kPi);
            float cosT = n.y, sinT = safe_sqrt(1 - cosT * cosT);
            V3 dpdv = sinT > 1e-6f ? V3(cosT * rel.x / sinT, -r * sinT, cosT * rel.z / sinT) * kPi : V3();
            if (len2(dpdu) < 1e-12f || len2(dpdv) < 1e-12f) coordinate_system(n, &dpdu, &dpdv);
            si->dpdu = dpdu;
            si->dpdv = dpdv;
            si->ns = si->n;
            si->dpdus = dpdu;
            return;
        }
        // shape.cpp:224-264
        V3 p0 = v3(p->p0), p1 = v3(p->p1), p2 = v3(p->p2);
        float b1 = hit.b1, b2 = hit.b2, b0 = 1 - b1 - b2;
        si->p = p0 * b0 + p1 * b1 + p2 * b2;
        V3 n = normalize(cross_f(p1 - p0, p2 - p0));
        V2 uv0{p->u0, p->v0}, uv1{p->u1, p->v1}, uv2{p->u2, p->v2};
        si->uv = {uv0.x * b0 + uv1.x * b1 + uv2.x * b2, uv0.y * b0 + uv1.y * b1 + uv2.y * b2};
        V2 duv02{uv0.x - uv2.x, uv0.y - uv2.y}, duv12{uv1.x - uv2.x, uv1.y - uv2.y};
        V3 dp02 = p0 - p2, dp12 = p1 - p2;
        float det = duv02.x * duv12.y - duv02.y * duv12.x;
        V3 dpdu, dpdv;
        bool degenerate = std::abs(det) < 1e-9f;
        if (!degenerate) {
            float inv = 1 / det;
            dpdu = (dp02 * duv12.y - dp12 * duv02.y) * inv;
            dpdv = (dp12 * duv02.x - dp02 * duv12.x) * inv;
            degenerate = len2(cross_d(dpdu, n)) == 0 || !std::isfinite(dpdu.x) || !std::isfinite(dpdv.x);
        }
        if (degenerate) coordinate_system(n, &dpdu, &dpdv);
        si->dpdu = dpdu;
        si->dpdv = dpdv;
        V3 ns = n;
        if (has_normals()) {
            ns = v3(p->n0) * b0 + v3(p->n1) * b1 + v3(p->n2) * b2;
            float l = len(ns);
            ns = l > 0 ? ns / l : n;
            n = face_forward(n, ns);
        }
        si->n = n;
        si->ns = ns;
        si->dpdus = dpdu;
        si->wo = -normalize(ray.d);
    }

    void normal_bounds(V3 *w, float *cosTheta) const {  // shape.cpp:266-273
        if (sphere()) {
            *w = V3(0, 0, 1);
            *cosTheta = -1;
            return;
        }
        V3 p0 = v3(p->p0), p1 = v3(p->p1), p2 = v3(p->p2);
        V3 n = normalize(cross_f(p1 - p0, p2 - p0));
        if (has_normals()) n = face_forward(n, v3(p->n0) + v3(p->n1) + v3(p->n2));
        *w = n;
        *cosTheta = 1;
    }

    // Area sampling (shape.cpp:87-97 sphere, 275-297 triangle).
    bool sample_area(V2 u, Intr *it, float *pdf) const {
        if (sphere()) {
            V3 n = sample_uniform_sphere(u);
            V3 c = v3(p->p0);
            it->p = c + n * p->r;
            it->n = p->flip ? -n : n;
            float phi = std::atan2(n.z, n.x);
            if (phi < 0) phi += 2 * kPi;
            it->uv = {phi * kInv2Pi, safe_acos(n.y) * kInvPi};
            *pdf = 1 / area();
            return true;
        }
        V3 p0 = v3(p->p0), p1 = v3(p->p1), p2 = v3(p->p2);
        float b[3];
        sample_uniform_triangle(u, b);
        it->p = p0 * b[0] + p1 * b[1] + p2 * b[2];
        V3 n = normalize(cross_f(p1 - p0, p2 - p0));
        if (has_normals()) {
            V3 ns = v3(p->n0) * b[0] + v3(p->n1) * b[1] + v3(p->n2) * b[2];
            n = face_forward(n, ns);
        }
        it->n = n;
        it->uv = {p->u0 * b[0] + p->u1 * b[1] + p->u2 * b[2], p->v0 * b[0] + p->v1 * b[1] + p->v2 * b[2]};
        *pdf = 1 / area();
        return true;
    }

    // Solid-angle sampling from a reference point (shape.cpp:12-24 generic, 99-128 sphere cone).
    bool sample_from(V3 ref, V2 u, Intr *it, float *pdf) const {
        if (sphere()) {
            V3 c = v3(p->p0);
            float r = p->r;
            float dc2 = dist2(ref, c);
            if (!(dc2 <= r * r * 1.0001f)) {
                float sinThetaMax = r / std::sqrt(dc2);
                float sin2ThetaMax = sq(sinThetaMax);
                float cosThetaMax = safe_sqrt(1 - sin2ThetaMax);
                float omc = 1 - cosThetaMax;
                float cosTheta = (cosThetaMax - 1) * u.x + 1;
                float sin2Theta = 1 - sq(cosTheta);
                if (sin2ThetaMax < 0.00068523f) {
                    sin2Theta = sin2ThetaMax * u.x;
                    cosTheta = std::sqrt(1 - sin2Theta);
                    omc = sin2ThetaMax / 2;
                }
                float cosAlpha = sin2Theta / sinThetaMax + cosTheta * safe_sqrt(1 - sin2Theta / sq(sinThetaMax));
                float sinAlpha = safe_sqrt(1 - sq(cosAlpha));
                float phi = u.y * 2 * kPi;
                V3 w = spherical_dir(sinAlpha, cosAlpha, phi);
                Frame fr = Frame::from_z(normalize(c - ref));
                V3 n = fr.from_local(-w);
                it->p = c + n * r;
                it->n = p->flip ? -n : n;
                float uphi = std::atan2(n.z, n.x);
                if (uphi < 0) uphi += 2 * kPi;
                it->uv = {uphi * kInv2Pi, safe_acos(n.y) * kInvPi};
                *pdf = 1 / (2 * kPi * omc);
                return true;
            }
        }
        float apdf;
        if (!sample_area(u, it, &apdf)) return false;
        V3 wi = it->p - ref;
        float d2 = len2(wi);
        if (d2 == 0) return false;
        wi = wi / std::sqrt(d2);
        float cosTheta = absdot(it->n, -wi);
        if (cosTheta == 0) return false;
        apdf *= d2 / cosTheta;
        if (!std::isfinite(apdf)) return false;
        *pdf = apdf;
        return true;
    }

    float pdf_from(V3 ref, V3 wi) const {  // shape.cpp:26-37 generic, 130-141 sphere
        if (sphere()) {
            V3 c = v3(p->p0);
            float r = p->r;
            float dc2 = dist2(ref, c);
            if (!(dc2 <= r * r * 1.0001f)) {
                float sin2ThetaMax = r * r / dc2;
                float cosThetaMax = safe_sqrt(1 - sin2ThetaMax);
                float omc = 1 - cosThetaMax;
                if (sin2ThetaMax < 0.00068523f) omc = sin2ThetaMax / 2;
                V3 toC = (c - ref) / std::sqrt(dc2);
                if (dot(toC, wi) < cosThetaMax - 1e-6f) return 0;
                return 1 / (2 * kPi * omc);
            }
        }
        Ray ray{ref, wi};
        float tHit;
        HitInfo hit;
        if (!intersect(ray, kInf, &tHit, &hit)) return 0;
        SurfHit si;
        interaction(ray, tHit, hit, &si);
        float cosTheta = absdot(si.n, -wi);
        if (cosTheta == 0) return 0;
        float pdf = (1 / area()) * dist2(ref, si.p) / cosTheta;
        return std::isfinite(pdf) ? pdf : 0.f;
    }
};

// =============================================================================================
// Binary SAH BVH (src/geometry/bvh.cpp)

struct BNode {
    Box b;
    int offset = 0;
    uint16_t nPrims = 0;
    uint8_t axis = 0;
};

class Bvh {
  public:
    struct Hit {
        float t = kInf;
        int prim = -1;
        HitInfo info;
    };

    void build(const std::vector<Shape> &shapes) {  // bvh.cpp:8-28
        shapes_ = &shapes;
        nodes_.clear();
        order_.clear();
        if (shapes.empty()) return;
        std::vector<BuildPrim> prims(shapes.size());
        for (size_t i = 0; i < shapes.size(); ++i) {
            prims[i].b = shapes[i].bounds();
            prims[i].c = prims[i].b.centroid();
            prims[i].index = int(i);
        }
        nodes_.reserve(2 * prims.size());
        recurse(prims, 0, int(prims.size()), nodes_, 0);
        order_.resize(prims.size());
        for (size_t i = 0; i < prims.size(); ++i) order_[i] = prims[i].index;
    }

    bool intersect(const Ray &ray, float tMax, Hit *hit) const {  // bvh.cpp:158-198
        if (nodes_.empty()) return false;
        V3 invDir(1 / ray.d.x, 1 / ray.d.y, 1 / ray.d.z);
        int neg[3] = {invDir.x < 0, invDir.y < 0, invDir.z < 0};
        int stack[64];
        int sp = 0, cur = 0;
        bool found = false;
        while (true) {
            const BNode &n = nodes_[size_t(cur)];
            if (box_hit(n.b, ray.o, invDir, neg, tMax)) {
                if (n.nPrims > 0) {
                    for (int i = 0; i < n.nPrims; ++i) {
                        int slot = n.offset + i;
                        float t;
                        HitInfo info;
                        if ((*shapes_)[size_t(order_[size_t(slot)])].intersect(ray, tMax, &t, &info)) {
                            tMax = t;
                            hit->t = t;
                            hit->prim = order_[size_t(slot)];
                            hit->info = info;
                            found = true;
                        }
                    }
                    if (sp == 0) break;
                    cur = stack[--sp];
                } else if (neg[n.axis]) {
                    stack[sp++] = cur + 1;
                    cur = n.offset;
                } else {
                    stack[sp++] = n.offset;
                    cur = cur + 1;
                }
            } else {
                if (sp == 0) break;
                cur = stack[--sp];
            }
        }
        return found;
    }

    bool intersect_any(const Ray &ray, float tMax) const {  // bvh.cpp:200-232
        if (nodes_.empty()) return false;
        V3 invDir(1 / ray.d.x, 1 / ray.d.y, 1 / ray.d.z);
        int neg[3] = {invDir.x < 0, invDir.y < 0, invDir.z < 0};
        int stack[64];
        int sp = 0, cur = 0;
        while (true) {
            const BNode &n = nodes_[size_t(cur)];
            if (box_hit(n.b, ray.o, invDir, neg, tMax)) {
                if (n.nPrims > 0) {
                    for (int i = 0; i < n.nPrims; ++i) {
                        float t;
                        HitInfo info;
                        if ((*shapes_)[size_t(order_[size_t(n.offset + i)])].intersect(ray, tMax, &t, &info)) return true;
                    }
                    if (sp == 0) break;
                    cur = stack[--sp];
                } else if (neg[n.axis]) {
                    stack[sp++] = cur + 1;
                    cur = n.offset;
                } else {
                    stack[sp++] = n.offset;
                    cur = cur + 1;
                }
            } else {
                if (sp == 0) break;
                cur = stack[--sp];
            }
        }
        return false;
    }

    Box bounds() const { return nodes_.empty() ? Box() : nodes_[0].b; }
    const std::vector<BNode> &nodes() const { return nodes_; }
    const std::vector<int> &order() const { return order_; }

  private:
    struct BuildPrim {
        Box b;
        V3 c;
        int index;
    };

    // Conservative slab test (bvh.cpp:137-156): far distances widened by 1 + 2 * gamma(3).
    static bool box_hit(const Box &b, V3 o, V3 invDir, const int neg[3], float tMax) {
        float tMin = (b.corner(neg[0]).x - o.x) * invDir.x;
        float tMaxX = (b.corner(1 - neg[0]).x - o.x) * invDir.x;
        float tyMin = (b.corner(neg[1]).y - o.y) * invDir.y;
        float tyMax = (b.corner(1 - neg[1]).y - o.y) * invDir.y;
        tMaxX *= 1 + 2 * 1.2e-7f * 3;
        tyMax *= 1 + 2 * 1.2e-7f * 3;
        if (tMin > tyMax || tyMin > tMaxX) return false;
        if (tyMin > tMin) tMin = tyMin;
        if (tyMax < tMaxX) tMaxX = tyMax;
        float tzMin = (b.corner(neg[2]).z - o.z) * invDir.z;
        float tzMax = (b.corner(1 - neg[2]).z - o.z) * invDir.z;
        tzMax *= 1 + 2 * 1.2e-7f * 3;
        if (tMin > tzMax || tzMin > tMaxX) return false;
        if (tzMin > tMin) tMin = tzMin;
        if (tzMax < tMaxX) tMaxX = tzMax;
        return (tMin < tMax) && (tMaxX > 0);
    }

    int recurse(std::vector<BuildPrim> &prims, int start, int end, std::vector<BNode> &out, int depth) {  // bvh.cpp:30-111
        int self = int(out.size());
        out.emplace_back();
        Box bounds, cb;
        for (int i = start; i < end; ++i) {
            bounds = join(bounds, prims[size_t(i)].b);
            cb = join(cb, prims[size_t(i)].c);
        }
        out[size_t(self)].b = bounds;
        int n = end - start;
        auto leaf = [&] {
            out[size_t(self)].offset = start;
            out[size_t(self)].nPrims = uint16_t(n);
            return self;
        };
        if (n <= 2) return leaf();
        int dim = cb.max_extent();
        float cmin = cb.lo[dim], cmax = cb.hi[dim];
        if (cmax == cmin) {
            if (n <= 255) return leaf();
            int mid = (start + end) / 2;
            out[size_t(self)].axis = uint8_t(dim);
            children(prims, start, mid, end, out, self, depth);
            return self;
        }
        constexpr int kBuckets = 16;
        struct Bucket {
            int count = 0;
            Box b;
        } buckets[kBuckets];
        auto bucket_of = [&](const BuildPrim &p) {
            int b = int(kBuckets * ((p.c[dim] - cmin) / (cmax - cmin)));
            return std::min(b, kBuckets - 1);
        };
        for (int i = start; i < end; ++i) {
            int b = bucket_of(prims[size_t(i)]);
            buckets[b].count++;
            buckets[b].b = join(buckets[b].b, prims[size_t(i)].b);
        }
        float costs[kBuckets - 1];
        {
            int below = 0;
            Box bb;
            for (int i = 0; i < kBuckets - 1; ++i) {
                bb = join(bb, buckets[i].b);
                below += buckets[i].count;
                costs[i] = below * bb.area();
            }
            int above = 0;
            Box ba;
            for (int i = kBuckets - 1; i >= 1; --i) {
                ba = join(ba, buckets[i].b);
                above += buckets[i].count;
                costs[i - 1] += above * ba.area();
            }
        }
        int minBucket = 0;
        float minCost = kInf;
        for (int i = 0; i < kBuckets - 1; ++i)
            if (costs[i] < minCost) {
                minCost = costs[i];
                minBucket = i;
            }
        const float traversalCost = 0.5f;
        float leafCost = float(n);
        minCost = traversalCost + minCost / bounds.area();
        if (n <= 4 && leafCost <= minCost) return leaf();
        // std::partition / std::nth_element as in the reference: their element order decides
        // the leaf order and therefore which primitive wins an exact tie.
        auto it = std::partition(prims.begin() + start, prims.begin() + end,
                                 [&](const BuildPrim &p) { return bucket_of(p) <= minBucket; });
        int mid = int(it - prims.begin());
        if (mid == start || mid == end) {
            mid = (start + end) / 2;
            std::nth_element(prims.begin() + start, prims.begin() + mid, prims.begin() + end,
                             [&](const BuildPrim &a, const BuildPrim &b) { return a.c[dim] < b.c[dim]; });
        }
        out[size_t(self)].axis = uint8_t(dim);
        children(prims, start, mid, end, out, self, depth);
        return self;
    }

    void children(std::vector<BuildPrim> &prims, int start, int mid, int end, std::vector<BNode> &nodes, int parent,
                  int depth) {  // bvh.cpp:113-135: large subtrees near the root build concurrently
        if (end - start >= 65536 && depth < 5) {
            std::vector<BNode> right;
            right.reserve(size_t(2 * (end - mid)));
            auto fut = std::async(std::launch::async, [&] { recurse(prims, mid, end, right, depth + 1); });
            recurse(prims, start, mid, nodes, depth + 1);
            fut.get();
            int base = int(nodes.size());
            for (BNode n : right) {
                if (n.nPrims == 0) n.offset += base;
                nodes.push_back(n);
            }
            nodes[size_t(parent)].offset = base;
            return;
        }
        recurse(prims, start, mid, nodes, depth + 1);
        int second = recurse(prims, mid, end, nodes, depth + 1);
        nodes[size_t(parent)].offset = second;
    }

    const std::vector<Shape> *shapes_ = nullptr;
    std::vector<BNode> nodes_;
    std::vector<int> order_;
};

// =============================================================================================
// BxDFs (src/materials/bxdf.h, bxdf.cpp:13-256)

enum BxFlags : uint32_t {
    kRefl = 1, kTrans = 2, kDiffuse = 4, kGlossy = 8, kSpecular = 16,
};

struct TrDist {  // bxdf.h:76-115
    float ax = 0, ay = 0;
    TrDist() = default;
    TrDist(float x, float y) : ax(x), ay(y) {
        if (!smooth()) {
            ax = std::max(ax, 1e-4f);
            ay = std::max(ay, 1e-4f);
        }
    }
    bool smooth() const { return std::max(ax, ay) < 1e-3f; }
    float D(V3 wm) const {
        float t2 = tan2t(wm);
        if (std::isinf(t2)) return 0;
        float c4 = sq(cos2t(wm));
        if (c4 < 1e-16f) return 0;
        float e = t2 * (sq(cosphi(wm) / ax) + sq(sinphi(wm) / ay));
        return 1 / (kPi * ax * ay * c4 * sq(1 + e));
    }
    float lambda(V3 w) const {
        float t2 = tan2t(w);
        if (std::isinf(t2)) return 0;
        float a2 = sq(cosphi(w) * ax) + sq(sinphi(w) * ay);
        return (std::sqrt(1 + a2 * t2) - 1) / 2;
    }
    float G1(V3 w) const { return 1 / (1 + lambda(w)); }
    float G(V3 wo, V3 wi) const { return 1 / (1 + lambda(wo) + lambda(wi)); }
    float Dv(V3 w, V3 wm) const { return G1(w) / abscos(w) * D(wm) * absdot(w, wm); }
    float pdf(V3 w, V3 wm) const { return Dv(w, wm); }
    V3 sample_wm(V3 w, V2 u) const {  // bxdf.cpp:19-30 (Heitz 2018 visible normals)
        V3 wh = normalize(V3(ax * w.x, ay * w.y, w.z));
        if (wh.z < 0) wh = -wh;
        V3 T1 = (wh.z < 0.99999f) ? normalize(cross_d(V3(0, 0, 1), wh)) : V3(1, 0, 0);
        V3 T2 = cross_d(wh, T1);
        float r = std::sqrt(u.x);
        float theta = 2 * kPi * u.y;
        V2 pd{r * std::cos(theta), r * std::sin(theta)};
        float h = std::sqrt(1 - sq(pd.x));
        pd.y = lerpf((1 + wh.z) / 2, h, pd.y);
        float pz = std::sqrt(std::max(0.f, 1 - pd.x * pd.x - pd.y * pd.y));
        V3 nh = T1 * pd.x + T2 * pd.y + wh * pz;
        return normalize(V3(ax * nh.x, ay * nh.y, std::max(1e-6f, nh.z)));
    }
};

inline float fr_dielectric(float ci, float eta) {  // bxdf.cpp:32-45
    ci = clampv(ci, -1.f, 1.f);
    if (ci < 0) {
        eta = 1 / eta;
        ci = -ci;
    }
    float s2i = 1 - sq(ci);
    float s2t = s2i / sq(eta);
    if (s2t >= 1) return 1.f;
    float ct = safe_sqrt(1 - s2t);
    float rpa = (eta * ci - ct) / (eta * ci + ct);
    float rpe = (ci - eta * ct) / (ci + eta * ct);
    return (sq(rpa) + sq(rpe)) / 2;
}

// Conductor Fresnel with complex IOR eta + i k in complex<float> (bxdf.cpp:47-62).
inline float fr_complex1(float ci, std::complex<float> eta) {
    using C = std::complex<float>;
    ci = clampv(ci, 0.f, 1.f);
    float s2i = 1 - sq(ci);
    C s2t = s2i / (eta * eta);
    C ct = std::sqrt(C(1.f) - s2t);
    C rpa = (eta * ci - ct) / (eta * ci + ct);
    C rpe = (C(ci) - eta * ct) / (C(ci) + eta * ct);
    return (std::norm(rpa) + std::norm(rpe)) / 2;
}
inline SS fr_complex(float ci, const SS &eta, const SS &k) {
    SS r;
    for (int i = 0; i < 4; ++i) r.v[i] = fr_complex1(ci, std::complex<float>(eta.v[i], k.v[i]));
    return r;
}

inline float hg_phase(float cosTheta, float g) {  // bxdf.cpp:64-67
    float denom = 1 + sq(g) + 2 * g * cosTheta;
    return kInv4Pi * (1 - sq(g)) / (denom * safe_sqrt(denom));
}
inline V3 sample_hg(V3 wo, float g, V2 u, float *pdf) {  // bxdf.cpp:69-80
    float cosTheta;
    if (std::abs(g) < 1e-3f) cosTheta = 1 - 2 * u.x;
    else cosTheta = -1 / (2 * g) * (1 + sq(g) - sq((1 - sq(g)) / (1 + g - 2 * g * u.x)));
    cosTheta = clampv(cosTheta, -1.f, 1.f);
    float sinTheta = safe_sqrt(1 - sq(cosTheta));
    float phi = 2 * kPi * u.y;
    Frame f = Frame::from_z(wo);
    V3 wi = f.from_local(spherical_dir(sinTheta, cosTheta, phi));
    if (pdf) *pdf = hg_phase(cosTheta, g);
    return wi;
}

// Sample flags (BxDFReflTransFlags, bxdf.h:39-40) and transport mode (bxdf.h:35-38). The mode only
// changes the dielectric's transmission (the 1/eta^2 radiance scaling); the layered walk samples
// the exit interface in the adjoint mode.
enum : uint32_t { kRtRefl = 1, kRtTrans = 2, kRtAll = 3 };
enum class Mode { Radiance, Importance };
inline Mode adjoint(Mode m) { return m == Mode::Radiance ? Mode::Importance : Mode::Radiance; }

struct BSample {
    SS f;
    V3 wi;
    float pdf = 0;
    uint32_t flags = 0;
    float eta = 1;
    bool proportional = false;  // pdf only proportional to the sampling density (bxdf.h:44-47)
};

enum BxType { BxNone = 0, BxDiffuse, BxConductor, BxDielectric, BxThin };

struct BxDF {  // the plain closures (bxdf.cpp:82-256)
    BxType type = BxNone;
    SS R;          // diffuse reflectance
    SS eta, k;     // conductor
    float etaD = 1;  // dielectric
    TrDist dist;

    uint32_t flags() const {
        switch (type) {
        case BxDiffuse: return R.nonzero() ? (kDiffuse | kRefl) : 0;
        case BxConductor: return dist.smooth() ? (kSpecular | kRefl) : (kGlossy | kRefl);
        case BxDielectric: {
            uint32_t f = (etaD == 1) ? kTrans : (kRefl | kTrans);
            return f | (dist.smooth() ? kSpecular : kGlossy);
        }
        case BxThin: return kRefl | kTrans | kSpecular;
        default: return 0;
        }
    }

    SS f(V3 wo, V3 wi, Mode mode = Mode::Radiance) const {
        switch (type) {
        case BxDiffuse:  // bxdf.cpp:85-88
            if (!same_hemi(wo, wi)) return SS(0.f);
            return R * kInvPi;
        case BxConductor: {  // bxdf.cpp:107-116
            if (!same_hemi(wo, wi) || dist.smooth()) return SS(0.f);
            float co = abscos(wo), ci = abscos(wi);
            if (ci == 0 || co == 0) return SS(0.f);
            V3 wm = wi + wo;
            if (len2(wm) == 0) return SS(0.f);
            wm = normalize(wm);
            SS F = fr_complex(absdot(wo, wm), eta, k);
            return F * (dist.D(wm) * dist.G(wo, wi) / (4 * ci * co));
        }
        case BxDielectric: {  // bxdf.cpp:196-213
            if (etaD == 1 || dist.smooth()) return SS(0.f);
            float co = wo.z, ci = wi.z;
            bool refl = ci * co > 0;
            float etap = 1;
            if (!refl) etap = co > 0 ? etaD : (1 / etaD);
            V3 wm = wi * etap + wo;
            if (ci == 0 || co == 0 || len2(wm) == 0) return SS(0.f);
            wm = face_forward(normalize(wm), V3(0, 0, 1));
            if (dot(wm, wi) * ci < 0 || dot(wm, wo) * co < 0) return SS(0.f);
            float F = fr_dielectric(dot(wo, wm), etaD);
            if (refl) return SS(dist.D(wm) * dist.G(wo, wi) * F / std::abs(4 * ci * co));
            float denom = sq(dot(wi, wm) + dot(wo, wm) / etap) * ci * co;
            float ft = dist.D(wm) * (1 - F) * dist.G(wo, wi) * std::abs(dot(wi, wm) * dot(wo, wm) / denom);
            if (mode == Mode::Radiance) ft /= sq(etap);
            return SS(ft);
        }
        default: return SS(0.f);
        }
    }

    float pdf(V3 wo, V3 wi, Mode = Mode::Radiance, uint32_t rt = kRtAll) const {
        switch (type) {
        case BxDiffuse:  // bxdf.cpp:99-102
            if (!(rt & kRtRefl) || !same_hemi(wo, wi)) return 0;
            return abscos(wi) * kInvPi;
        case BxConductor: {  // bxdf.cpp:138-145
            if (!(rt & kRtRefl)) return 0;
            if (!same_hemi(wo, wi) || dist.smooth()) return 0;
            V3 wm = wo + wi;
            if (len2(wm) == 0) return 0;
            wm = face_forward(normalize(wm), V3(0, 0, 1));
            return dist.pdf(wo, wm) / (4 * absdot(wo, wm));
        }
        case BxDielectric: {  // bxdf.cpp:215-234
            if (etaD == 1 || dist.smooth()) return 0;
            float co = wo.z, ci = wi.z;
            bool refl = ci * co > 0;
            float etap = 1;
            if (!refl) etap = co > 0 ? etaD : (1 / etaD);
            V3 wm = wi * etap + wo;
            if (ci == 0 || co == 0 || len2(wm) == 0) return 0;
            wm = face_forward(normalize(wm), V3(0, 0, 1));
            if (dot(wm, wi) * ci < 0 || dot(wm, wo) * co < 0) return 0;
            float R = fr_dielectric(dot(wo, wm), etaD), T = 1 - R;
            float pr = (rt & kRtRefl) ? R : 0.f, pt = (rt & kRtTrans) ? T : 0.f;
            if (pr == 0 && pt == 0) return 0;
            if (refl) return dist.pdf(wo, wm) / (4 * absdot(wo, wm)) * pr / (pr + pt);
            float denom = sq(dot(wi, wm) + dot(wo, wm) / etap);
            float dwm_dwi = absdot(wi, wm) / denom;
            return dist.pdf(wo, wm) * dwm_dwi * pt / (pr + pt);
        }
        default: return 0;
        }
    }

    std::optional<BSample> sample(V3 wo, float uc, V2 u, Mode mode = Mode::Radiance, uint32_t rt = kRtAll) const {
        switch (type) {
        case BxDiffuse: {  // bxdf.cpp:90-97
            if (!(rt & kRtRefl)) return std::nullopt;
            V3 wi = sample_cosine_hemisphere(u);
            if (wo.z < 0) wi.z *= -1;
            float p = abscos(wi) * kInvPi;
            return BSample{R * kInvPi, wi, p, kDiffuse | kRefl, 1};
        }
        case BxConductor: {  // bxdf.cpp:118-136
            if (!(rt & kRtRefl)) return std::nullopt;
            if (dist.smooth()) {
                V3 wi(-wo.x, -wo.y, wo.z);
                SS f = fr_complex(abscos(wi), eta, k) / abscos(wi);
                return BSample{f, wi, 1, kSpecular | kRefl, 1};
            }
            if (wo.z == 0) return std::nullopt;
            V3 wm = dist.sample_wm(wo, u);
            V3 wi = reflect(wo, wm);
            if (!same_hemi(wo, wi)) return std::nullopt;
            float p = dist.pdf(wo, wm) / (4 * absdot(wo, wm));
            float co = abscos(wo), ci = abscos(wi);
            if (ci == 0 || co == 0) return std::nullopt;
            SS F = fr_complex(absdot(wo, wm), eta, k);
            SS f = F * (dist.D(wm) * dist.G(wo, wi) / (4 * ci * co));
            return BSample{f, wi, p, kGlossy | kRefl, 1};
        }
        case BxDielectric: {  // bxdf.cpp:150-194
            if (etaD == 1 || dist.smooth()) {
                float R = fr_dielectric(wo.z, etaD), T = 1 - R;
                float pr = (rt & kRtRefl) ? R : 0.f, pt = (rt & kRtTrans) ? T : 0.f;
                if (pr == 0 && pt == 0) return std::nullopt;
                if (uc < pr / (pr + pt)) {
                    V3 wi(-wo.x, -wo.y, wo.z);
                    SS fr(R / abscos(wi));
                    return BSample{fr, wi, pr / (pr + pt), kSpecular | kRefl, 1};
                }
                V3 wi;
                float etap;
                if (!refract(wo, V3(0, 0, 1), etaD, &etap, &wi)) return std::nullopt;
                SS ft(T / abscos(wi));
                if (mode == Mode::Radiance) ft = ft / sq(etap);
                return BSample{ft, wi, pt / (pr + pt), kSpecular | kTrans, etap};
            }
            V3 wm = dist.sample_wm(wo, u);
            float R = fr_dielectric(dot(wo, wm), etaD), T = 1 - R;
            float pr = (rt & kRtRefl) ? R : 0.f, pt = (rt & kRtTrans) ? T : 0.f;
            if (pr == 0 && pt == 0) return std::nullopt;
            if (uc < pr / (pr + pt)) {
                V3 wi = reflect(wo, wm);
                if (!same_hemi(wo, wi)) return std::nullopt;
                float p = dist.pdf(wo, wm) / (4 * absdot(wo, wm)) * pr / (pr + pt);
                SS f(dist.D(wm) * dist.G(wo, wi) * R / (4 * wi.z * wo.z));
                return BSample{f, wi, p, kGlossy | kRefl, 1};
            }
            float etap;
            V3 wi;
            bool tir = !refract(wo, wm, etaD, &etap, &wi);
            if (tir || same_hemi(wo, wi) || wi.z == 0) return std::nullopt;
            float denom = sq(dot(wi, wm) + dot(wo, wm) / etap);
            float dwm_dwi = absdot(wi, wm) / denom;
            float p = dist.pdf(wo, wm) * dwm_dwi * pt / (pr + pt);
            SS ft(T * dist.D(wm) * dist.G(wo, wi) * std::abs(dot(wi, wm) * dot(wo, wm) / (wi.z * wo.z * denom)));
            if (mode == Mode::Radiance) ft = ft / sq(etap);
            return BSample{ft, wi, p, kGlossy | kTrans, etap};
        }
        case BxThin: {  // bxdf.cpp:239-256
            float R = fr_dielectric(abscos(wo), etaD), T = 1 - R;
            if (R < 1) {
                R += sq(T) * R / (1 - sq(R));
                T = 1 - R;
            }
            float pr = (rt & kRtRefl) ? R : 0.f, pt = (rt & kRtTrans) ? T : 0.f;
            if (pr == 0 && pt == 0) return std::nullopt;
            if (uc < pr / (pr + pt)) {
                V3 wi(-wo.x, -wo.y, wo.z);
                return BSample{SS(R / abscos(wi)), wi, pr / (pr + pt), kSpecular | kRefl, 1};
            }
            V3 wi = -wo;
            return BSample{SS(T / abscos(wi)), wi, pt / (pr + pt), kSpecular | kTrans, 1};
        }
        default: return std::nullopt;
        }
    }
};

inline bool usable(const std::optional<BSample> &s) { return s && s->f.nonzero() && s->pdf != 0 && s->wi.z != 0; }

// Stochastic layered closure (reference bxdf.cpp:260-517, the position-free random walk of
// pbrt-v4's LayeredBxDF): a dielectric coat over a base, optionally with a homogeneous slab of
// unit extinction and the given albedo between them. The walk's random stream is hashed from
// the directions (f: Hash(0, wo), Hash(wi); pdf: Hash(0, wi), Hash(wo); sample: Hash(0, wo),
// Hash(uc, u)), so all three are deterministic functions of their arguments. Each 2D draw takes
// its second coordinate first: the reference's Vec2f(r(), r()) evaluates its arguments right to
// left (MSVC and g++ alike; pinned against the compiled reference).
struct Layered {
    BxDF top, bottom;
    float thick = 0.01f, g = 0;
    SS albedo;
    bool twoSided = true;
    int maxDepth = 64, nSamples = 1;

    struct Walk {
        Pcg32 rng;
        Walk(uint64_t seq, uint64_t offset) { rng.set_sequence(seq, offset); }
        float u() { return std::min(rng.uniform(), kOneMinusEps); }
        V2 u2() {
            float y = u();
            float x = u();
            return V2{x, y};
        }
    };
    static float tr(float dz, V3 w) {
        if (std::abs(dz) <= std::numeric_limits<float>::min()) return 1;
        return std::exp(-std::abs(dz / w.z));
    }
    static float free_flight(float u, V3 w) { return -std::log(1 - u) / (1.f / std::abs(w.z)); }  // sampling.h:88

    uint32_t flags() const {  // bxdf.cpp:260-268
        uint32_t t = top.flags(), b = bottom.flags(), f = kRefl;
        if (t & kSpecular) f |= kSpecular;
        if ((t & kDiffuse) || (b & kDiffuse) || albedo.nonzero()) f |= kDiffuse;
        else if ((t & kGlossy) || (b & kGlossy)) f |= kGlossy;
        if ((t & kTrans) && (b & kTrans)) f |= kTrans;
        return f;
    }

    SS f(V3 wo, V3 wi, Mode mode) const {  // bxdf.cpp:270-373
        if (twoSided && wo.z < 0) {
            wo = -wo;
            wi = -wi;
        }
        const bool fromTop = twoSided || wo.z > 0;
        const bool same = same_hemi(wo, wi);
        const bool exitBottom = same != fromTop;  // the walk leaves through the interface on wi's side
        const BxDF &enter = fromTop ? top : bottom;
        const BxDF &exitI = exitBottom ? bottom : top, &inner = exitBottom ? top : bottom;
        const float exitZ = exitBottom ? 0.f : thick;
        const bool exitSpec = (exitI.flags() & kSpecular) != 0, innerSpec = (inner.flags() & kSpecular) != 0;
        SS f(0.f);
        if (same) f = enter.f(wo, wi, mode) * float(nSamples);
        Walk w(hash_args(uint64_t(0), wo), hash_args(wi));
        for (int s = 0; s < nSamples; ++s) {
            float uc = w.u();
            V2 u0 = w.u2();
            auto wos = enter.sample(wo, uc, u0, mode, kRtTrans);
            if (!usable(wos)) continue;
            uc = w.u();
            V2 u1 = w.u2();
            auto wis = exitI.sample(wi, uc, u1, adjoint(mode), kRtTrans);
            if (!usable(wis)) continue;
            SS beta = wos->f * abscos(wos->wi) / wos->pdf;
            float z = fromTop ? thick : 0;
            V3 d = wos->wi;
            for (int depth = 0; depth < maxDepth; ++depth) {
                if (depth > 3 && beta.maxc() < 0.25f) {
                    float q = std::max(0.f, 1 - beta.maxc());
                    if (w.u() < q) break;
                    beta = beta / (1 - q);
                }
                if (!albedo.nonzero()) {
                    z = (z == thick) ? 0 : thick;
                    beta = beta * tr(thick, d);
                } else {
                    float zp = d.z > 0 ? (z + free_flight(w.u(), d)) : (z - free_flight(w.u(), d));
                    if (0 < zp && zp < thick) {
                        // Scattering in the slab: next-event estimate through the exit interface's
                        // sampled direction, then continue with a phase-function sample.
                        float wt = 1;
                        float phaseExit = hg_phase(dot(-d, -wis->wi), g);
                        if (!exitSpec) wt = power_heuristic(wis->pdf, phaseExit);
                        f = f + beta * albedo * phaseExit * wt * tr(zp - exitZ, wis->wi) * wis->f / wis->pdf;
                        float phasePdf;
                        V3 wp = sample_hg(-d, g, w.u2(), &phasePdf);
                        if (phasePdf == 0 || wp.z == 0) continue;
                        beta = beta * albedo;
                        d = wp;
                        z = zp;
                        if (((z < exitZ && d.z > 0) || (z > exitZ && d.z < 0)) && !exitSpec) {
                            SS fExit = exitI.f(-d, wi, mode);
                            if (fExit.nonzero()) {
                                float exitPdf = exitI.pdf(-d, wi, mode, kRtTrans);
                                float wt2 = power_heuristic(phasePdf, exitPdf);
                                f = f + beta * tr(zp - exitZ, wp) * fExit * wt2;
                            }
                        }
                        continue;
                    }
                    z = clampv(zp, 0.f, thick);
                }
                if (z == exitZ) {
                    float uc2 = w.u();
                    V2 u2 = w.u2();
                    auto bs = exitI.sample(-d, uc2, u2, mode, kRtRefl);
                    if (!usable(bs)) break;
                    beta = beta * (bs->f * abscos(bs->wi) / bs->pdf);
                    d = bs->wi;
                } else {
                    if (!innerSpec) {
                        float wt = 1;
                        if (!exitSpec) wt = power_heuristic(wis->pdf, inner.pdf(-d, -wis->wi, mode));
                        f = f + beta * inner.f(-d, -wis->wi, mode) * abscos(wis->wi) * wt * tr(thick, wis->wi) * wis->f /
                                    wis->pdf;
                    }
                    float uc2 = w.u();
                    V2 u2 = w.u2();
                    auto bs = inner.sample(-d, uc2, u2, mode, kRtRefl);
                    if (!usable(bs)) break;
                    beta = beta * (bs->f * abscos(bs->wi) / bs->pdf);
                    d = bs->wi;
                    if (!exitSpec) {
                        SS fExit = exitI.f(-d, wi, mode);
                        if (fExit.nonzero()) {
                            float wt = 1;
                            if (!innerSpec) wt = power_heuristic(bs->pdf, exitI.pdf(-d, wi, mode, kRtTrans));
                            f = f + beta * tr(thick, bs->wi) * fExit * wt;
                        }
                    }
                }
            }
        }
        return f / float(nSamples);
    }

    std::optional<BSample> sample(V3 wo, float uc, V2 u, Mode mode) const {  // bxdf.cpp:375-454
        bool flip = false;
        if (twoSided && wo.z < 0) {
            wo = -wo;
            flip = true;
        }
        const bool fromTop = twoSided || wo.z > 0;
        auto bs = (fromTop ? top : bottom).sample(wo, uc, u, mode);
        if (!usable(bs)) return std::nullopt;
        if (bs->flags & kRefl) {
            if (flip) bs->wi = -bs->wi;
            bs->proportional = true;
            return bs;
        }
        V3 d = bs->wi;
        bool specPath = (bs->flags & kSpecular) != 0;
        Walk w(hash_args(uint64_t(0), wo), hash_args(uc, u));
        SS f = bs->f * abscos(bs->wi);
        float pdf = bs->pdf;
        float z = fromTop ? thick : 0;
        for (int depth = 0; depth < maxDepth; ++depth) {
            float rrBeta = f.maxc() / pdf;
            if (depth > 3 && rrBeta < 0.25f) {
                float q = std::max(0.f, 1 - rrBeta);
                if (w.u() < q) return std::nullopt;
                pdf *= 1 - q;
            }
            if (d.z == 0) return std::nullopt;
            if (albedo.nonzero()) {
                float zp = d.z > 0 ? (z + free_flight(w.u(), d)) : (z - free_flight(w.u(), d));
                if (zp == z) return std::nullopt;
                if (0 < zp && zp < thick) {
                    float phasePdf;
                    V3 wp = sample_hg(-d, g, w.u2(), &phasePdf);
                    if (phasePdf == 0 || wp.z == 0) return std::nullopt;
                    f = f * (albedo * phasePdf);
                    pdf *= phasePdf;
                    specPath = false;
                    d = wp;
                    z = zp;
                    continue;
                }
                z = clampv(zp, 0.f, thick);
            } else {
                z = (z == thick) ? 0 : thick;
                f = f * tr(thick, d);
            }
            const BxDF &iface = (z == 0) ? bottom : top;
            float uc2 = w.u();
            V2 u2 = w.u2();
            auto bs2 = iface.sample(-d, uc2, u2, mode);
            if (!usable(bs2)) return std::nullopt;
            f = f * bs2->f;
            pdf *= bs2->pdf;
            specPath = specPath && (bs2->flags & kSpecular);
            d = bs2->wi;
            if (bs2->flags & kTrans) {
                uint32_t fl = (same_hemi(wo, d) ? kRefl : kTrans) | (specPath ? kSpecular : kGlossy);
                if (flip) d = -d;
                return BSample{f, d, pdf, fl, 1.f, true};
            }
            f = f * abscos(bs2->wi);
        }
        return std::nullopt;
    }

    float pdf(V3 wo, V3 wi, Mode mode) const {  // bxdf.cpp:456-517
        if (twoSided && wo.z < 0) {
            wo = -wo;
            wi = -wi;
        }
        Walk w(hash_args(uint64_t(0), wi), hash_args(wo));
        const bool fromTop = twoSided || wo.z > 0;
        const bool same = same_hemi(wo, wi);
        float sum = 0;
        if (same) sum += nSamples * (fromTop ? top.pdf(wo, wi, mode, kRtRefl) : bottom.pdf(wo, wi, mode, kRtRefl));
        for (int s = 0; s < nSamples; ++s) {
            if (same) {
                // Transmission through the entry interface, one bounce off the other.
                const BxDF &rI = fromTop ? bottom : top, &tI = fromTop ? top : bottom;
                float a = w.u();
                V2 ua = w.u2();
                auto wos = tI.sample(wo, a, ua, mode, kRtTrans);
                float b = w.u();
                V2 ub = w.u2();
                auto wis = tI.sample(wi, b, ub, adjoint(mode), kRtTrans);
                if (wos && wos->f.nonzero() && wos->pdf > 0 && wis && wis->f.nonzero() && wis->pdf > 0) {
                    if (!(tI.flags() & (kDiffuse | kGlossy))) {
                        sum += rI.pdf(-wos->wi, -wis->wi, mode);
                    } else {
                        float c = w.u();
                        V2 uc2 = w.u2();
                        auto rs = rI.sample(-wos->wi, c, uc2, mode);
                        if (rs && rs->f.nonzero() && rs->pdf > 0) {
                            if (!(rI.flags() & (kDiffuse | kGlossy))) {
                                sum += tI.pdf(-rs->wi, wi, mode);
                            } else {
                                float rPdf = rI.pdf(-wos->wi, -wis->wi, mode);
                                sum += power_heuristic(wis->pdf, rPdf) * rPdf;
                                float tPdf = tI.pdf(-rs->wi, wi, mode);
                                sum += power_heuristic(rs->pdf, tPdf) * tPdf;
                            }
                        }
                    }
                }
            } else {
                // Transmission through both interfaces.
                const BxDF &toI = fromTop ? top : bottom, &tiI = fromTop ? bottom : top;
                float a = w.u();
                V2 ua = w.u2();
                auto wos = toI.sample(wo, a, ua, mode);
                if (!usable(wos) || (wos->flags & kRefl)) continue;
                float b = w.u();
                V2 ub = w.u2();
                auto wis = tiI.sample(wi, b, ub, adjoint(mode));
                if (!usable(wis) || (wis->flags & kRefl)) continue;
                if (toI.flags() & kSpecular) sum += tiI.pdf(-wos->wi, wi, mode);
                else if (tiI.flags() & kSpecular) sum += toI.pdf(wo, -wis->wi, mode);
                else sum += (toI.pdf(wo, -wis->wi, mode) + tiI.pdf(-wos->wi, wi, mode)) / 2;
            }
        }
        return lerpf(0.9f, 1 / (4 * kPi), sum / nSamples);
    }
};

// "Charlie" sheen (reference bxdf.cpp:519-631): the microfibre lobe D / (4 (ci + co - ci co))
// and its directional albedo, tabulated over 16 alphas x 32 cosines by 32 x 32 stratified
// cosine samples and interpolated bilinearly.
constexpr int kSheenMuN = 32, kSheenAlphaN = 16;
inline float sheen_value(V3 wo, V3 wi, float alpha) {
    if (!same_hemi(wo, wi)) return 0;
    float co = abscos(wo), ci = abscos(wi);
    if (co == 0 || ci == 0) return 0;
    V3 wh = normalize(wo + wi);
    float c = abscos(wh), inv = 1 / alpha;
    float sin2 = std::max(0.f, 1 - c * c);
    float D = (2 + inv) * std::pow(sin2, 0.5f * inv) * kInv2Pi;
    return D / (4 * (ci + co - ci * co));
}
inline const float *sheen_table() {
    static const std::vector<float> table = [] {
        std::vector<float> t(size_t(kSheenMuN) * kSheenAlphaN);
        for (int ia = 0; ia < kSheenAlphaN; ++ia) {
            float a = lerpf(float(ia) / (kSheenAlphaN - 1), 0.05f, 1.f);
            for (int im = 0; im < kSheenMuN; ++im) {
                float mu = std::max(1e-3f, float(im) / (kSheenMuN - 1));
                V3 wo(safe_sqrt(1 - mu * mu), 0, mu);
                double sum = 0;
                for (int i = 0; i < 32; ++i)
                    for (int j = 0; j < 32; ++j) {
                        V3 wi = sample_cosine_hemisphere(V2{(i + 0.5f) / 32, (j + 0.5f) / 32});
                        sum += sheen_value(wo, wi, a) * kPi;
                    }
                t[size_t(ia) * kSheenMuN + im] = float(sum / (32 * 32));
            }
        }
        return t;
    }();
    return table.data();
}
inline float sheen_albedo(const float *t, float cosTheta, float alpha) {
    float fa = (clampv(alpha, 0.05f, 1.f) - 0.05f) / 0.95f * (kSheenAlphaN - 1);
    float fm = clampv(std::abs(cosTheta), 0.f, 1.f) * (kSheenMuN - 1);
    int ia = std::min(int(fa), kSheenAlphaN - 2), im = std::min(int(fm), kSheenMuN - 2);
    float ta = fa - ia, tm = fm - im;
    auto at = [&](int a, int m) { return t[a * kSheenMuN + m]; };
    return lerpf(ta, lerpf(tm, at(ia, im), at(ia, im + 1)), lerpf(tm, at(ia + 1, im), at(ia + 1, im + 1)));
}

struct SheenLobe {
    SS color;
    float alpha = 0.3f, weight = 0, colorMax = 0;
    const float *table = nullptr;
    SS f(V3 wo, V3 wi) const { return color * (weight * sheen_value(wo, wi, alpha)); }
    float base_scale(V3 wo, V3 wi) const {
        float e = std::max(sheen_albedo(table, wo.z, alpha), sheen_albedo(table, wi.z, alpha));
        return std::max(0.f, 1 - weight * colorMax * e);
    }
    float prob(V3 wo, uint32_t baseFlags) const {
        float e = weight * colorMax * sheen_albedo(table, wo.z, alpha);
        if (!baseFlags) return 1;
        return clampv(e, 0.f, 0.9f);
    }
};

// The closure of one surface hit (Material::GetBSDF, material.cpp:5-14): a plain or layered
// base, optionally under a sheen lobe. Local-frame directions.
struct Closure {
    BxDF bx;
    bool layered = false, sheen = false;
    Layered lay;
    SheenLobe sh;

    uint32_t base_flags() const { return layered ? lay.flags() : bx.flags(); }
    uint32_t flags() const { return sheen ? (base_flags() | kGlossy | kRefl) : base_flags(); }
    bool pdf_exact() const { return !layered; }
    SS base_f(V3 wo, V3 wi, Mode m) const { return layered ? lay.f(wo, wi, m) : bx.f(wo, wi, m); }
    float base_pdf(V3 wo, V3 wi, Mode m, uint32_t rt) const { return layered ? lay.pdf(wo, wi, m) : bx.pdf(wo, wi, m, rt); }
    std::optional<BSample> base_sample(V3 wo, float uc, V2 u, Mode m, uint32_t rt) const {
        return layered ? lay.sample(wo, uc, u, m) : bx.sample(wo, uc, u, m, rt);
    }

    SS f(V3 wo, V3 wi, Mode m = Mode::Radiance) const {  // bxdf.cpp:589-591
        if (!sheen) return base_f(wo, wi, m);
        return sh.f(wo, wi) + base_f(wo, wi, m) * sh.base_scale(wo, wi);
    }
    float pdf(V3 wo, V3 wi, Mode m = Mode::Radiance, uint32_t rt = kRtAll) const {  // bxdf.cpp:593-599
        if (!sheen) return base_pdf(wo, wi, m, rt);
        float ps = sh.prob(wo, base_flags());
        float pSheen = (rt & kRtRefl) && same_hemi(wo, wi) ? abscos(wi) * kInvPi : 0.f;
        return ps * pSheen + (1 - ps) * base_pdf(wo, wi, m, rt);
    }
    std::optional<BSample> sample(V3 wo, float uc, V2 u, Mode m = Mode::Radiance, uint32_t rt = kRtAll) const {
        if (!sheen) return base_sample(wo, uc, u, m, rt);
        // bxdf.cpp:601-629: pick the sheen lobe with probability ps, else the base.
        float ps = sh.prob(wo, base_flags());
        if (!(rt & kRtRefl)) ps = 0;
        const bool exact = pdf_exact();
        if (uc < ps) {
            V3 wi = sample_cosine_hemisphere(u);
            if (wo.z < 0) wi.z = -wi.z;
            if (exact) {
                float p = pdf(wo, wi, m, rt);
                if (p == 0) return std::nullopt;
                return BSample{f(wo, wi, m), wi, p, kGlossy | kRefl, 1};
            }
            float p = ps * (abscos(wi) * kInvPi);
            return BSample{sh.f(wo, wi), wi, p, kGlossy | kRefl, 1, true};
        }
        float uc2 = ps < 1 ? std::min((uc - ps) / (1 - ps), kOneMinusEps) : 0.f;
        auto bs = base_sample(wo, uc2, u, m, rt);
        if (!bs || bs->pdf == 0) return std::nullopt;
        if ((bs->flags & kSpecular) || !exact || bs->proportional) {
            bs->f = bs->f * sh.base_scale(wo, bs->wi);
            bs->pdf *= (1 - ps);
            if (!(bs->flags & kSpecular)) bs->proportional = true;
            return bs;
        }
        float p = pdf(wo, bs->wi, m, rt);
        if (p == 0) return std::nullopt;
        bs->f = f(wo, bs->wi, m);
        bs->pdf = p;
        return bs;
    }
};

struct BSDF {  // src/materials/bsdf.h:8-51
    Frame frame;
    Closure c;
    bool valid = false;
    SS f(V3 woW, V3 wiW) const {
        V3 wi = frame.to_local(wiW), wo = frame.to_local(woW);
        if (wo.z == 0) return SS(0.f);
        return c.f(wo, wi);
    }
    float pdf(V3 woW, V3 wiW) const {
        V3 wo = frame.to_local(woW), wi = frame.to_local(wiW);
        if (wo.z == 0) return 0;
        return c.pdf(wo, wi);
    }
    std::optional<BSample> sample(V3 woW, float u, V2 u2) const {
        V3 wo = frame.to_local(woW);
        if (wo.z == 0 || !(c.flags() & (kRefl | kTrans))) return std::nullopt;
        auto bs = c.sample(wo, u, u2);
        if (!bs || !bs->f.nonzero() || bs->pdf == 0 || bs->wi.z == 0 || bs->f.has_nan() || !std::isfinite(bs->pdf))
            return std::nullopt;
        bs->wi = frame.from_local(bs->wi);
        return bs;
    }
    bool has_non_specular() const { return (c.flags() & (kDiffuse | kGlossy)) != 0; }
};

// =============================================================================================
// Lights and light sampling (src/lights/light.cpp, light_bvh.cpp, sky.cpp:52-60)

struct LightB {  // LightBounds (light.h:30-38)
    Box bounds;
    V3 w{0, 0, 1};
    float phi = 0;
    float cosTheta_o = 1, cosTheta_e = 0;
    bool twoSided = false;
};

struct Cone {
    V3 w{0, 0, 1};
    float cosTheta = kInf;
    Cone() = default;
    Cone(V3 ww, float c) : w(normalize(ww)), cosTheta(c) {}
    bool empty() const { return cosTheta == kInf; }
};

inline float angle_between(V3 a, V3 b) {  // light_bvh.cpp:10-13
    if (dot(a, b) < 0) return kPi - 2 * safe_asin(len(a + b) / 2);
    return 2 * safe_asin(len(b - a) / 2);
}

inline Cone cone_union(const Cone &a, const Cone &b) {  // light_bvh.cpp:15-29
    if (a.empty()) return b;
    if (b.empty()) return a;
    float ta = safe_acos(a.cosTheta), tb = safe_acos(b.cosTheta);
    float td = angle_between(a.w, b.w);
    if (std::min(td + tb, kPi) <= ta) return a;
    if (std::min(td + ta, kPi) <= tb) return b;
    float to = (ta + td + tb) / 2;
    if (to >= kPi) return Cone(V3(0, 0, 1), -1);
    float tr = to - ta;
    V3 wr = cross_d(a.w, b.w);
    if (len2(wr) == 0) return Cone(V3(0, 0, 1), -1);
    V3 w = rotate_vector(rad2deg(tr), wr, a.w);
    return Cone(w, std::cos(to));
}

inline LightB lb_union(const LightB &a, const LightB &b) {  // light_bvh.cpp:31-43
    if (a.phi == 0) return b;
    if (b.phi == 0) return a;
    Cone c = cone_union(Cone(a.w, a.cosTheta_o), Cone(b.w, b.cosTheta_o));
    LightB r;
    r.bounds = join(a.bounds, b.bounds);
    r.w = c.w;
    r.phi = a.phi + b.phi;
    r.cosTheta_o = c.cosTheta;
    r.cosTheta_e = std::min(a.cosTheta_e, b.cosTheta_e);
    r.twoSided = a.twoSided || b.twoSided;
    return r;
}

inline float lb_importance(const LightB &lb, V3 p, V3 n) {  // light_bvh.cpp:45-83
    V3 pc = (lb.bounds.lo + lb.bounds.hi) * 0.5f;
    float d2 = dist2(p, pc);
    d2 = std::max(d2, len(lb.bounds.diag()) / 2);
    auto cos_sub = [](float sA, float cA, float sB, float cB) {
        if (cA > cB) return 1.f;
        return cA * cB + sA * sB;
    };
    auto sin_sub = [](float sA, float cA, float sB, float cB) {
        if (cA > cB) return 0.f;
        return sA * cB - cA * sB;
    };
    V3 wi = p - pc;
    float l = len(wi);
    wi = l > 0 ? wi / l : V3(0, 0, 1);
    float cos_w = dot(lb.w, wi);
    if (lb.twoSided) cos_w = std::abs(cos_w);
    float sin_w = safe_sqrt(1 - sq(cos_w));
    V3 c;
    float r;
    lb.bounds.bounding_sphere(&c, &r);
    float cos_b = -1;
    float dc2 = dist2(p, c);
    if (dc2 >= r * r) cos_b = safe_sqrt(1 - r * r / dc2);
    float sin_b = safe_sqrt(1 - sq(cos_b));
    float sin_o = safe_sqrt(1 - sq(lb.cosTheta_o));
    float cos_x = cos_sub(sin_w, cos_w, sin_o, lb.cosTheta_o);
    float sin_x = sin_sub(sin_w, cos_w, sin_o, lb.cosTheta_o);
    float cos_p = cos_sub(sin_x, cos_x, sin_b, cos_b);
    if (cos_p <= lb.cosTheta_e) return 0;
    float imp = lb.phi * cos_p / d2;
    if (!n.zero()) {
        float cos_i = absdot(wi, n);
        float sin_i = safe_sqrt(1 - sq(cos_i));
        imp *= cos_sub(sin_i, cos_i, sin_b, cos_b);
    }
    return std::max(imp, 0.f);
}

inline float lb_cost(const LightB &b, const Box &bounds, int dim) {  // light_bvh.cpp:90-99
    float t_o = safe_acos(b.cosTheta_o), t_e = safe_acos(b.cosTheta_e);
    float t_w = std::min(t_o + t_e, kPi);
    float s_o = safe_sqrt(1 - sq(b.cosTheta_o));
    float M = 2 * kPi * (1 - b.cosTheta_o) +
              kPi / 2 * (2 * t_w * s_o - std::cos(t_o - 2 * t_w) - 2 * t_o * s_o + b.cosTheta_o);
    V3 d = bounds.diag();
    float Kr = d[dim] > 0 ? max_comp(d) / d[dim] : 1.f;
    return b.phi * M * Kr * b.bounds.area();
}

struct LightSample {
    SS L;
    V3 wi;
    float pdf = 0;
    Intr pLight;
};

struct Scene;

// RGBToSpectrumTable::operator() (src/core/rgb2spec.cpp:278-304) over a table record (resolution n,
// n z nodes, coefficients [max component][z][y][x][3]).
static void rgb_table_coeffs(const float *table, const float rgb[3], float co[3]) {
    float rgbc[3] = {clampv(rgb[0], 0.f, 1.f), clampv(rgb[1], 0.f, 1.f), clampv(rgb[2], 0.f, 1.f)};
    if (rgbc[0] == rgbc[1] && rgbc[1] == rgbc[2]) {
        float v = rgbc[0];
        co[0] = co[1] = 0;
        co[2] = v <= 0 ? -std::numeric_limits<float>::infinity()
                       : (v >= 1 ? std::numeric_limits<float>::infinity() : (v - .5f) / std::sqrt(v * (1 - v)));
        return;
    }
    const int n = int(table[0]);
    const float *zNodes = table + 1, *cf = zNodes + n;
    int maxc = (rgbc[0] > rgbc[1]) ? ((rgbc[0] > rgbc[2]) ? 0 : 2) : ((rgbc[1] > rgbc[2]) ? 1 : 2);
    float z = rgbc[maxc];
    float x = rgbc[(maxc + 1) % 3] * (n - 1) / z;
    float y = rgbc[(maxc + 2) % 3] * (n - 1) / z;
    int xi = std::min(int(x), n - 2), yi = std::min(int(y), n - 2);
    int zi = find_interval(n, [&](int i) { return zNodes[i] < z; });
    float dx = x - xi, dy = y - yi, dz = (z - zNodes[zi]) / (zNodes[zi + 1] - zNodes[zi]);
    auto at = [&](int zz, int yy, int xx, int i) {
        return cf[3 * (((size_t(maxc) * n + zz) * n + yy) * n + xx) + i];
    };
    for (int i = 0; i < 3; ++i)
        co[i] = lerpf(dz,
                      lerpf(dy, lerpf(dx, at(zi, yi, xi, i), at(zi, yi, xi + 1, i)),
                            lerpf(dx, at(zi, yi + 1, xi, i), at(zi, yi + 1, xi + 1, i))),
                      lerpf(dy, lerpf(dx, at(zi + 1, yi, xi, i), at(zi + 1, yi, xi + 1, i)),
                            lerpf(dx, at(zi + 1, yi + 1, xi, i), at(zi + 1, yi + 1, xi + 1, i))));
}

// ---- image textures (src/materials/texture.cpp:8-52, texture.h:61-124) ----------------------------
static int tex_wrap(int v, int n, int mode) {  // ImageData::Texel's wrap (texture.cpp:10-20)
    switch (mode) {
    case 0: v %= n; return v < 0 ? v + n : v;
    case 1: return clampv(v, 0, n - 1);
    default: {
        int period = 2 * n;
        v %= period;
        if (v < 0) v += period;
        return v < n ? v : period - 1 - v;
    }
    }
}
// ImageData::Lookup(UVMapping::Map(uv)) on a texture record (kTexHeader).
static void tex_rgb(const float *rec, V2 uv, float out[3]) {
    const int w = int(rec[0]), h = int(rec[1]), mode = int(rec[2]);
    V2 st{uv.x * rec[4] + rec[6], uv.y * rec[5] + rec[7]};
    float v = rec[3] != 0 ? 1 - st.y : st.y;
    float x = st.x * w - 0.5f, y = v * h - 0.5f;
    int x0 = int(std::floor(x)), y0 = int(std::floor(y));
    float dx = x - x0, dy = y - y0;
    auto texel = [&](int xi, int yi) {
        return rec + prgpu::kTexHeader + 3 * (size_t(tex_wrap(yi, h, mode)) * size_t(w) + size_t(tex_wrap(xi, w, mode)));
    };
    const float *a = texel(x0, y0), *b = texel(x0 + 1, y0), *c = texel(x0, y0 + 1), *d = texel(x0 + 1, y0 + 1);
    for (int i = 0; i < 3; ++i)
        out[i] = (1 - dx) * (1 - dy) * a[i] + dx * (1 - dy) * b[i] + (1 - dx) * dy * c[i] + dx * dy * d[i];
}
static float tex_float(const float *rec, int channel, V2 uv) {  // ImageFloatTexture::Evaluate
    float c[3];
    tex_rgb(rec, uv, c);
    float v = channel < 0 ? (c[0] + c[1] + c[2]) / 3 : c[channel];
    return v * rec[8];
}
// Material::PerturbShading (src/materials/material.cpp:16-58) with a shading record (kShadeHeader,
// texture records in knots): a bump map's finite-difference height gradient in uv tilts the
// shading normal, a normal map's tangent-space normal replaces it in the shading frame; the result
// is kept in the geometric normal's hemisphere and the shading tangent re-orthogonalised. Returns
// whether the frame changed.
static bool perturb_record(const float *knots, const float *r, V2 uv, V3 n, V3 dpdu, V3 dpdv, V3 *nsIo, V3 *dpdusIo) {
    V3 ns = *nsIo;
    bool changed = false;
    if (r[4] != 0) {
        const float *hrec = knots + int(r[5]);
        const int ch = int(r[6]);
        const float du = 1.f / 2048, dv = 1.f / 2048;
        V2 cu = uv, cv = uv;
        cu.x += du;
        cv.y += dv;
        float h0 = tex_float(hrec, ch, uv);
        float dhdu = (tex_float(hrec, ch, cu) - h0) / du * r[7];
        float dhdv = (tex_float(hrec, ch, cv) - h0) / dv * r[7];
        V3 pu = dpdu + ns * dhdu, pv = dpdv + ns * dhdv;
        V3 nb = cross_d(pu, pv);
        float l = len(nb);
        if (l > 0 && std::isfinite(l)) {
            nb = nb / l;
            if (dot(nb, ns) < 0) nb = -nb;
            ns = nb;
            *dpdusIo = pu;
            changed = true;
        }
    }
    if (r[0] != 0) {
        float c[3];
        tex_rgb(knots + int(r[1]), uv, c);
        V3 t(2 * c[0] - 1, 2 * c[1] - 1, 2 * c[2] - 1);
        t.x *= r[2];
        t.y *= r[3] != 0 ? -r[2] : r[2];
        Frame f = Frame::from_xz(*dpdusIo, ns);
        if (dot(f.y, dpdv) < 0) f.y = -f.y;  // the bitangent follows the parameterisation's v
        V3 nm = f.x * t.x + f.y * t.y + f.z * std::max(t.z, 1e-3f);
        float l = len(nm);
        if (l > 0 && std::isfinite(l)) {
            ns = nm / l;
            changed = true;
        }
    }
    if (!changed) return false;
    if (dot(ns, n) < 0) ns = -ns;
    *nsIo = ns;
    *dpdusIo = *dpdusIo - ns * dot(*dpdusIo, ns);
    return true;
}

struct EnvImage;

struct LightObj {
    LightEx d;
    float area = 0;
    std::shared_ptr<EnvImage> env;  // image environment (d.profileSize > 0)
    bool delta() const {
        return d.kind == prgpu::LightExPoint || d.kind == prgpu::LightExDistant || d.kind == prgpu::LightExSpot ||
               d.kind == prgpu::LightExGoniometric;
    }
    bool infinite_type() const { return d.kind == prgpu::LightExEnvironment || d.kind == prgpu::LightExSun; }
};

class LightBvh {
  public:
    struct Node {
        LightB lb;
        int child1 = -1;
        int light = -1;
    };
    std::vector<int> bounded, infinite;  // light indices
    std::vector<Node> nodes;
    std::vector<int> parent;
    std::unordered_map<int, int> leafOf;

    void build(const std::vector<std::optional<LightB>> &bounds) {  // light_bvh.cpp:103-115
        std::vector<std::pair<int, LightB>> items;
        for (size_t i = 0; i < bounds.size(); ++i) {
            if (!bounds[i]) {
                infinite.push_back(int(i));
            } else if (bounds[i]->phi > 0) {
                items.push_back({int(bounded.size()), *bounds[i]});
                bounded.push_back(int(i));
            }
        }
        if (!items.empty()) rec(items, 0, int(items.size()), -1);
    }
    float p_infinite() const {
        return infinite.empty() ? 0.f : float(infinite.size()) / float(infinite.size() + (nodes.empty() ? 0 : 1));
    }
    int sample(V3 p, V3 n, float u, float *pmf) const {  // light_bvh.cpp:179-214
        float pInf = p_infinite();
        if (u < pInf) {
            u /= pInf;
            int index = std::min(int(u * infinite.size()), int(infinite.size()) - 1);
            *pmf = pInf / infinite.size();
            return infinite[size_t(index)];
        }
        if (nodes.empty()) return -1;
        u = std::min((u - pInf) / (1 - pInf), kOneMinusEps);
        int ni = 0;
        float prob = 1 - pInf;
        while (true) {
            const Node &node = nodes[size_t(ni)];
            if (node.light >= 0) {
                if (ni > 0 || lb_importance(node.lb, p, n) > 0) {
                    *pmf = prob;
                    return bounded[size_t(node.light)];
                }
                return -1;
            }
            const Node &c0 = nodes[size_t(ni + 1)], &c1 = nodes[size_t(node.child1)];
            float ci0 = lb_importance(c0.lb, p, n), ci1 = lb_importance(c1.lb, p, n);
            if (ci0 == 0 && ci1 == 0) return -1;
            float p0 = ci0 / (ci0 + ci1);
            if (u < p0) {
                prob *= p0;
                u = std::min(u / p0, kOneMinusEps);
                ni = ni + 1;
            } else {
                prob *= 1 - p0;
                u = std::min((u - p0) / (1 - p0), kOneMinusEps);
                ni = node.child1;
            }
        }
    }
    float pmf(V3 p, V3 n, int light) const {  // light_bvh.cpp:216-234
        auto it = leafOf.find(light);
        if (it == leafOf.end()) {
            for (int l : infinite)
                if (l == light) return p_infinite() / infinite.size();
            return 0;
        }
        int leaf = it->second;
        if (leaf == 0) return lb_importance(nodes[0].lb, p, n) > 0 ? 1 - p_infinite() : 0.f;
        float pm = 1 - p_infinite();
        for (int child = leaf, par = parent[size_t(leaf)]; par >= 0; child = par, par = parent[size_t(par)]) {
            const Node &node = nodes[size_t(par)];
            float ci0 = lb_importance(nodes[size_t(par + 1)].lb, p, n), ci1 = lb_importance(nodes[size_t(node.child1)].lb, p, n);
            if (ci0 == 0 && ci1 == 0) return 0;
            pm *= (child == node.child1 ? ci1 : ci0) / (ci0 + ci1);
        }
        return pm;
    }

  private:
    int rec(std::vector<std::pair<int, LightB>> &items, int start, int end, int par) {  // light_bvh.cpp:117-177
        if (end - start == 1) {
            int ni = int(nodes.size());
            Node node;
            node.lb = items[size_t(start)].second;
            node.light = items[size_t(start)].first;
            nodes.push_back(node);
            parent.push_back(par);
            leafOf[bounded[size_t(node.light)]] = ni;
            return ni;
        }
        Box bounds, cb;
        for (int i = start; i < end; ++i) {
            const LightB &lb = items[size_t(i)].second;
            bounds = join(bounds, lb.bounds);
            cb = join(cb, lb.bounds.centroid());
        }
        float minCost = kInf;
        int minBucket = -1, minDim = -1;
        constexpr int kB = 12;
        for (int dim = 0; dim < 3; ++dim) {
            if (cb.hi[dim] == cb.lo[dim]) continue;
            LightB bucket[kB];
            for (int i = start; i < end; ++i) {
                V3 pc = items[size_t(i)].second.bounds.centroid();
                int b = int(kB * cb.offset(pc)[dim]);
                b = clampv(b, 0, kB - 1);
                bucket[b] = lb_union(bucket[b], items[size_t(i)].second);
            }
            for (int i = 0; i < kB - 1; ++i) {
                LightB b0, b1;
                for (int j = 0; j <= i; ++j) b0 = lb_union(b0, bucket[j]);
                for (int j = i + 1; j < kB; ++j) b1 = lb_union(b1, bucket[j]);
                float cost = lb_cost(b0, bounds, dim) + lb_cost(b1, bounds, dim);
                if (cost > 0 && cost < minCost) {
                    minCost = cost;
                    minBucket = i;
                    minDim = dim;
                }
            }
        }
        int mid;
        if (minDim == -1) {
            mid = (start + end) / 2;
        } else {
            auto it = std::partition(items.begin() + start, items.begin() + end, [&](const std::pair<int, LightB> &l) {
                int b = int(kB * cb.offset(l.second.bounds.centroid())[minDim]);
                return clampv(b, 0, kB - 1) <= minBucket;
            });
            mid = int(it - items.begin());
            if (mid == start || mid == end) mid = (start + end) / 2;
        }
        int ni = int(nodes.size());
        nodes.push_back(Node());
        parent.push_back(par);
        rec(items, start, mid, ni);
        int c1 = rec(items, mid, end, ni);
        nodes[size_t(ni)].child1 = c1;
        nodes[size_t(ni)].lb = lb_union(nodes[size_t(ni + 1)].lb, nodes[size_t(c1)].lb);
        return ni;
    }
};

// Piecewise-constant discrete distribution (src/core/sampling.cpp:5-42), for power sampling.
struct Distrib1D {
    std::vector<float> func, cdf;
    float funcInt = 0;
    void init(const std::vector<float> &f) {
        func = f;
        int n = int(func.size());
        for (float &v : func) v = std::abs(v);
        cdf.assign(size_t(n + 1), 0.f);
        double acc = 0;
        for (int i = 1; i < n + 1; ++i) {
            acc += double(func[size_t(i - 1)]) / n;
            cdf[size_t(i)] = float(acc);
        }
        funcInt = float(acc);
        if (funcInt == 0) {
            for (int i = 1; i < n + 1; ++i) cdf[size_t(i)] = float(i) / float(n);
        } else {
            for (int i = 1; i < n + 1; ++i) cdf[size_t(i)] = float(cdf[size_t(i)] / acc);
        }
        cdf[size_t(n)] = 1;
    }
    float pmf(int i) const { return funcInt > 0 ? func[size_t(i)] / (funcInt * int(func.size())) : 1.f / int(func.size()); }
    int sample(float u, float *p) const {
        int off = find_interval(int(cdf.size()), [&](int i) { return cdf[size_t(i)] <= u; });
        *p = pmf(off);
        return off;
    }
    float sample_continuous(float u, float *pdf, int *offOut) const {  // sampling.cpp:24-32
        int n = int(func.size());
        int off = find_interval(int(cdf.size()), [&](int i) { return cdf[size_t(i)] <= u; });
        *offOut = off;
        float du = u - cdf[size_t(off)];
        if (cdf[size_t(off) + 1] - cdf[size_t(off)] > 0) du /= (cdf[size_t(off) + 1] - cdf[size_t(off)]);
        *pdf = funcInt > 0 ? func[size_t(off)] / funcInt : 1.f;
        return std::min((off + du) / n, kOneMinusEps);
    }
};

// Image environment light (src/lights/light.cpp:221-297): the map and its rotation from the
// record, the sampling distribution rebuilt from the texels (not read from the record), the
// RGB-to-spectrum table from its record (src/core/rgb2spec.cpp:278-304, 316-343).
struct EnvImage {
    int w = 0, h = 0, d65 = -1;
    float c = 1, s = 0, avgY = 0;
    std::vector<float> rgb, table;
    std::vector<Distrib1D> cond;
    Distrib1D marg;
    void init(const float *rec, const float *knots) {
        w = int(rec[0]);
        h = int(rec[1]);
        c = rec[2];
        s = rec[3];
        d65 = int(rec[6]);
        rgb.assign(rec + prgpu::kEnvHeader, rec + prgpu::kEnvHeader + size_t(3) * w * h);
        const float *t = knots + int(rec[7]);
        const int n = int(t[0]);
        table.assign(t, t + 1 + n + size_t(3) * n * n * n * 3);
        std::vector<float> f(size_t(w) * h);
        double sumY = 0, sumW = 0;
        for (int y = 0; y < h; ++y) {  // light.cpp:226-237
            float sinTheta = std::sin(kPi * (y + 0.5f) / h);
            for (int x = 0; x < w; ++x) {
                const float *px = &rgb[3 * (size_t(y) * w + x)];
                float lum = 0.2126f * px[0] + 0.7152f * px[1] + 0.0722f * px[2];
                f[size_t(y) * w + x] = std::max(lum, 0.f) * sinTheta + 1e-6f;
                sumY += lum * sinTheta;
                sumW += sinTheta;
            }
        }
        avgY = float(sumY / std::max(1e-9, sumW));
        cond.resize(size_t(h));
        std::vector<float> m(static_cast<size_t>(h), 0.f);
        for (int y = 0; y < h; ++y) {  // Distribution2D (sampling.cpp:44-50)
            cond[size_t(y)].init(std::vector<float>(f.begin() + size_t(y) * w, f.begin() + size_t(y + 1) * w));
            m[size_t(y)] = cond[size_t(y)].funcInt;
        }
        marg.init(m);
    }
    V3 to_light(V3 v) const { return V3(c * v.x + 0.f * v.y + (-s) * v.z, v.y, s * v.x + 0.f * v.y + c * v.z); }
    V3 to_world(V3 v) const { return V3(c * v.x + 0.f * v.y + s * v.z, v.y, (-s) * v.x + 0.f * v.y + c * v.z); }
    static void dir_to_equirect(V3 d, float *u, float *v) {  // light.h:292-297
        float theta = safe_acos(d.y);
        float phi = std::atan2(d.z, d.x);
        if (phi < 0) phi += 2 * kPi;
        *u = phi * kInv2Pi;
        *v = theta * kInvPi;
    }
    static V3 equirect_to_dir(float u, float v) {  // light.h:287-291
        float theta = v * kPi, phi = u * 2 * kPi;
        float st = std::sin(theta);
        return V3(st * std::cos(phi), std::cos(theta), st * std::sin(phi));
    }
    void coeffs(float r, float g, float b, float co[3]) const {
        const float rgb[3] = {r, g, b};
        rgb_table_coeffs(table.data(), rgb, co);
    }
    SS lookup(const SpectrumTable &spec, float scale, V3 wWorld, const Lambda &lam) const {  // light.cpp:249-261
        float u, v;
        dir_to_equirect(normalize(to_light(wWorld)), &u, &v);
        float x = u * w - 0.5f, y = v * h - 0.5f;
        int x0 = int(std::floor(x)), y0 = int(std::floor(y));
        float dx = x - x0, dy = y - y0;
        auto tex = [&](int xi, int yi) {
            xi = ((xi % w) + w) % w;
            yi = clampv(yi, 0, h - 1);
            return &rgb[3 * (size_t(yi) * w + xi)];
        };
        const float *a = tex(x0, y0), *b = tex(x0 + 1, y0), *cc = tex(x0, y0 + 1), *d = tex(x0 + 1, y0 + 1);
        float px[3];
        for (int i = 0; i < 3; ++i)
            px[i] = (1 - dx) * (1 - dy) * a[i] + dx * (1 - dy) * b[i] + (1 - dx) * dy * cc[i] + dx * dy * d[i];
        for (int i = 0; i < 3; ++i) px[i] = px[i] * scale;
        // RGBIlluminantSample = RGBUnboundedSample x D65 (rgb2spec.cpp:326-343)
        float m = std::max(0.f, std::max(px[0], std::max(px[1], px[2])));
        if (m == 0) return SS(0.f);
        float sc = 2 * m, inv = 1 / sc, co[3];
        coeffs(px[0] * inv, px[1] * inv, px[2] * inv, co);
        SS out;
        for (int i = 0; i < 4; ++i)
            out.v[i] = sc * sigmoid_s(std::fma(lam.l[i], std::fma(lam.l[i], co[0], co[1]), co[2])) * spec.eval(d65, lam.l[i]);
        return out;
    }
    bool sample(V2 u, V3 *wi, float *pdf) const {  // light.cpp:272-280, Distribution2D::SampleContinuous
        float pdfs[2];
        int row, col;
        float d1 = marg.sample_continuous(u.y, &pdfs[1], &row);
        float d0 = cond[size_t(row)].sample_continuous(u.x, &pdfs[0], &col);
        float mapPdf = pdfs[0] * pdfs[1];
        if (mapPdf == 0) return false;
        float sinTheta = std::sin(d1 * kPi);
        if (sinTheta == 0) return false;
        *wi = to_world(equirect_to_dir(d0, d1));
        *pdf = mapPdf / (2 * kPi * kPi * sinTheta);
        return true;
    }
    float pdf(V3 wi) const {  // light.cpp:291-297, Distribution2D::PDF
        float u, v;
        dir_to_equirect(normalize(to_light(wi)), &u, &v);
        float sinTheta = std::sin(v * kPi);
        if (sinTheta == 0) return 0;
        int iu = clampv(int(u * w), 0, w - 1), iv = clampv(int(v * h), 0, h - 1);
        if (marg.funcInt == 0) return 0;
        return cond[size_t(iv)].func[size_t(iu)] / marg.funcInt / (2 * kPi * kPi * sinTheta);
    }
};

// =============================================================================================
// Media (src/media/medium.cpp)

struct DistSample {
    bool scattered = false;
    float t = 0;
    SS weight;
    SS pdfRatio{1.f};
};

// =============================================================================================
// Scene

struct Scene {
    SpectrumTable spec;
    std::vector<Primitive> prims;
    std::vector<Material> mats;
    std::vector<Shape> shapes;
    std::vector<LightObj> lights;
    std::vector<MediumEx> media;
    std::vector<float> grid;
    Bvh bvh;
    LightBvh lbvh;
    Distrib1D power;
    int cameraMedium = -1;
    prgpu::CameraEx cam;
    Affine c2w;
    V3 worldCenter;
    float worldRadius = 1;
    // film
    int filterType = 0;
    float filterRadius = 0.5f, filterSigma = 0.5f, filterNorm = 1;
    int margin = 0;

    bool is_interface(int pi) const {
        const Primitive &p = prims[size_t(pi)];
        bool noMat = p.material < 0 || mats[size_t(p.material)].type == prgpu::MatInterface ||
                     mats[size_t(p.material)].type == prgpu::MatBlack;
        return noMat && p.light < 0;
    }
    bool has_bsdf(int pi) const {
        const Primitive &p = prims[size_t(pi)];
        if (p.material < 0) return false;
        uint32_t t = mats[size_t(p.material)].type;
        return t == prgpu::MatDiffuse || t == prgpu::MatConductor || t == prgpu::MatDielectric;
    }

    // ---- film filter (src/core/film.cpp:9-39) ----
    float filter_1d(float x) const {
        x = std::abs(x);
        if (x > filterRadius) return 0;
        switch (filterType) {
        case prgpu::FilterBox: return 1;
        case prgpu::FilterGaussian:
            return std::max(0.f, std::exp(-x * x / (2 * filterSigma * filterSigma)) -
                                     std::exp(-filterRadius * filterRadius / (2 * filterSigma * filterSigma)));
        default: {
            float t = 0.5f + x / (2 * filterRadius);
            const float a0 = 0.35875f, a1 = 0.48829f, a2 = 0.14128f, a3 = 0.01168f;
            return a0 - a1 * std::cos(2 * kPi * t) + a2 * std::cos(4 * kPi * t) - a3 * std::cos(6 * kPi * t);
        }
        }
    }
    void init_filter(int type, float radius) {
        filterType = type;
        filterRadius = radius;
        if (type == prgpu::FilterBox) filterRadius = std::max(0.5f, std::min(filterRadius, 0.5f));
        if (type == prgpu::FilterGaussian) filterSigma = filterRadius / 3.f;
        const int n = 4096;
        double s1 = 0;
        for (int i = 0; i < n; ++i) {
            float x = -filterRadius + (i + 0.5f) * (2 * filterRadius / n);
            s1 += filter_1d(x);
        }
        s1 *= 2 * filterRadius / n;
        filterNorm = float(1.0 / (s1 * s1));
        margin = int(std::ceil(filterRadius - 0.5f));
    }

    // ---- light interface ----
    std::optional<LightB> light_bounds(int li) const {
        const LightObj &L = lights[size_t(li)];
        switch (L.d.kind) {
        case prgpu::LightExArea: {  // light.cpp:89-103
            const Shape &s = shapes[size_t(L.d.primitive)];
            LightB lb;
            lb.bounds = s.bounds();
            V3 w;
            float ct;
            s.normal_bounds(&w, &ct);
            lb.w = w;
            lb.cosTheta_o = ct;
            lb.cosTheta_e = 0;
            lb.twoSided = L.d.twoSided != 0;
            lb.phi = L.d.scale * spec.max_value(L.d.spectrum) * s.area() * kPi * (L.d.twoSided ? 2 : 1);
            return lb;
        }
        case prgpu::LightExPoint: {  // light.cpp:108-115
            LightB lb;
            lb.bounds = Box{v3(L.d.position), v3(L.d.position)};
            lb.phi = 4 * kPi * L.d.scale * spec.max_value(L.d.spectrum);
            lb.cosTheta_o = -1;
            lb.cosTheta_e = 0;
            return lb;
        }
        case prgpu::LightExSpot: {  // light.cpp:117-125
            LightB lb;
            lb.bounds = Box{v3(L.d.position), v3(L.d.position)};
            lb.w = v3(L.d.direction);
            lb.phi = 4 * kPi * L.d.scale * spec.max_value(L.d.spectrum);
            lb.cosTheta_o = L.d.cosFalloffStart;
            lb.cosTheta_e = std::cos(safe_acos(L.d.cosFalloffEnd) - safe_acos(L.d.cosFalloffStart));
            return lb;
        }
        case prgpu::LightExGoniometric: {  // light.h:199-206
            LightB lb;
            lb.bounds = Box{v3(L.d.position), v3(L.d.position)};
            lb.phi = 4 * kPi * L.d.scale * spec.max_value(L.d.spectrum);
            lb.cosTheta_o = -1;
            lb.cosTheta_e = 0;
            return lb;
        }
        default: return std::nullopt;  // distant and environment lights
        }
    }
    // Relative intensity of a goniometric light towards world direction w: the profile's type C
    // lookup in the luminaire frame, with phi folded by the profile's symmetry (ies.cpp:80-105).
    float ies_profile(const LightObj &L, V3 w) const {
        const float *r = spec.knots.data() + L.d.profileOffset;
        V3 x(r[0], r[1], r[2]), y(r[3], r[4], r[5]), z(r[6], r[7], r[8]);
        V3 wl(dot(w, x), dot(w, y), dot(w, z));
        const int nV = int(r[9]), nH = int(r[10]);
        const float *V = r + prgpu::kIesHeader, *H = V + nV, *C = H + nH;
        float theta = (180.f / kPi) * safe_acos(wl.z);
        float phi = (180.f / kPi) * std::atan2(wl.y, wl.x);
        if (theta < V[0] || theta > V[nV - 1]) return 0;
        float hLast = H[nH - 1];
        phi = std::fmod(phi, 360.f);
        if (phi < 0) phi += 360;
        if (nH == 1 || hLast == 0) phi = 0;
        else if (hLast <= 90) {
            phi = std::fmod(phi, 180.f);
            if (phi > 90) phi = 180 - phi;
        } else if (hLast <= 180) {
            if (phi > 180) phi = 360 - phi;
        }
        auto vertical = [&](int hi) {
            const float *row = C + size_t(hi) * size_t(nV);
            if (nV == 1) return row[0];
            int vi = find_interval(nV, [&](int i) { return V[i] <= theta; });
            float t = (theta - V[vi]) / std::max(1e-6f, V[vi + 1] - V[vi]);
            return lerpf(clampv(t, 0.f, 1.f), row[vi], row[vi + 1]);
        };
        if (nH == 1) return vertical(0);
        int hi = find_interval(nH, [&](int i) { return H[i] <= phi; });
        float t = (phi - H[hi]) / std::max(1e-6f, H[hi + 1] - H[hi]);
        return lerpf(clampv(t, 0.f, 1.f), vertical(hi), vertical(hi + 1));
    }
    float light_power(int li) const {
        const LightObj &L = lights[size_t(li)];
        switch (L.d.kind) {
        case prgpu::LightExArea: {  // light.cpp:15-18
            float lobe = 2 * kPi / (L.d.cosPower + 2);
            return L.d.scale * spec.to_y(L.d.spectrum) * L.area * lobe * (L.d.twoSided ? 2 : 1);
        }
        case prgpu::LightExPoint: return 4 * kPi * L.d.scale * spec.to_y(L.d.spectrum);
        case prgpu::LightExGoniometric:  // light.h:198: the profile's integral is in its record
            return L.d.scale * spec.to_y(L.d.spectrum) * spec.knots[size_t(L.d.profileOffset) + 11];
        case prgpu::LightExSpot:  // light.cpp:173-175
            return L.d.scale * spec.to_y(L.d.spectrum) * 2 * kPi *
                   ((1 - L.d.cosFalloffStart) + (L.d.cosFalloffStart - L.d.cosFalloffEnd) / 2);
        case prgpu::LightExDistant: return L.d.scale * spec.to_y(L.d.spectrum) * kPi * worldRadius * worldRadius;
        case prgpu::LightExSun:  // sky.cpp:88-90
            return L.d.scale * spec.to_y(L.d.spectrum) * 2 * kPi * (1 - L.d.cosFalloffEnd) * kPi * worldRadius *
                   worldRadius;
        default:  // environment (light.cpp:241): the map's mean luminance, or the constant's
            return 4 * kPi * kPi * worldRadius * worldRadius * L.d.scale *
                   (L.env ? L.env->avgY : spec.to_y(L.d.spectrum));
        }
    }
    SS area_L(const LightObj &L, V3 n, V3 w, const Lambda &lam) const {  // light.cpp:20-25
        float c = dot(n, w);
        if (!L.d.twoSided && c <= 0) return SS(0.f);
        float prof = L.d.cosPower > 0 ? std::pow(std::abs(c), L.d.cosPower) : 1.f;
        return spec.sample(L.d.spectrum, lam) * (L.d.scale * prof);
    }
    std::optional<LightSample> sample_li(int li, V3 p, V3 n, V2 u, const Lambda &lam) const {
        const LightObj &L = lights[size_t(li)];
        (void)n;
        switch (L.d.kind) {
        case prgpu::LightExArea: {  // light.cpp:27-42
            const Shape &s = shapes[size_t(L.d.primitive)];
            Intr it;
            float pdf;
            if (!s.sample_from(p, u, &it, &pdf) || pdf == 0 || len2(it.p - p) == 0) return std::nullopt;
            V3 wi = normalize(it.p - p);
            SS Le = area_L(L, it.n, -wi, lam);
            if (!Le.nonzero()) return std::nullopt;
            const Primitive &pr = prims[size_t(L.d.primitive)];
            it.surface_mi = true;
            it.mIn = pr.mediumInside;
            it.mOut = pr.mediumOutside;
            return LightSample{Le, wi, pdf, it};
        }
        case prgpu::LightExPoint: {  // light.cpp:126-137
            V3 d = v3(L.d.position) - p;
            float d2 = len2(d);
            if (d2 == 0) return std::nullopt;
            LightSample ls;
            ls.wi = d / std::sqrt(d2);
            ls.L = spec.sample(L.d.spectrum, lam) * (L.d.scale / d2);
            ls.pdf = 1;
            ls.pLight.p = v3(L.d.position);
            return ls;
        }
        case prgpu::LightExSpot: {  // light.cpp:168-192
            V3 d = v3(L.d.position) - p;
            float d2 = len2(d);
            if (d2 == 0) return std::nullopt;
            V3 wi = d / std::sqrt(d2);
            float f = smooth_step(dot(-wi, v3(L.d.direction)), L.d.cosFalloffEnd, L.d.cosFalloffStart);
            if (f == 0) return std::nullopt;
            LightSample ls;
            ls.wi = wi;
            ls.L = spec.sample(L.d.spectrum, lam) * (L.d.scale * f / d2);
            ls.pdf = 1;
            ls.pLight.p = v3(L.d.position);
            return ls;
        }
        case prgpu::LightExGoniometric: {  // sky.cpp:12-27
            V3 d = v3(L.d.position) - p;
            float d2 = len2(d);
            if (d2 == 0) return std::nullopt;
            V3 wi = d / std::sqrt(d2);
            float prof = ies_profile(L, -wi);
            if (prof <= 0) return std::nullopt;
            LightSample ls;
            ls.wi = wi;
            ls.L = spec.sample(L.d.spectrum, lam) * (L.d.scale * prof / d2);
            ls.pdf = 1;
            ls.pLight.p = v3(L.d.position);
            return ls;
        }
        case prgpu::LightExSun: {  // sky.cpp:97-105
            float cosMax = L.d.cosFalloffEnd;
            float cosT = (1 - u.x) + u.x * cosMax;
            V3 wl = spherical_dir(safe_sqrt(1 - cosT * cosT), cosT, u.y * 2 * kPi);
            LightSample ls;
            ls.wi = Frame::from_z(v3(L.d.direction)).from_local(wl);
            ls.pdf = 1 / (2 * kPi * (1 - cosMax));
            ls.L = spec.sample(L.d.spectrum, lam) * L.d.scale;
            ls.pLight.p = p + ls.wi * (2 * worldRadius);
            return ls;
        }
        case prgpu::LightExDistant: {  // sky.cpp:52-60
            LightSample ls;
            V3 w = v3(L.d.direction);
            ls.wi = w;
            ls.L = spec.sample(L.d.spectrum, lam) * L.d.scale;
            ls.pdf = 1;
            ls.pLight.p = p + w * (2 * worldRadius);
            return ls;
        }
        default: {  // environment (light.cpp:258-289)
            if (L.env) {
                LightSample ls;
                if (!L.env->sample(u, &ls.wi, &ls.pdf)) return std::nullopt;
                ls.L = L.env->lookup(spec, L.d.scale, ls.wi, lam);
                ls.pLight.p = p + ls.wi * (2 * worldRadius);
                return ls;
            }
            V3 wi = sample_uniform_sphere(u);
            LightSample ls;
            ls.wi = wi;
            ls.pdf = kInv4Pi;
            ls.L = spec.sample(L.d.spectrum, lam) * L.d.scale;
            ls.pLight.p = p + wi * (2 * worldRadius);
            return ls;
        }
        }
    }
    float pdf_li(int li, V3 p, V3 wi) const {
        const LightObj &L = lights[size_t(li)];
        if (L.d.kind == prgpu::LightExArea) return shapes[size_t(L.d.primitive)].pdf_from(p, wi);
        if (L.d.kind == prgpu::LightExEnvironment) return L.env ? L.env->pdf(wi) : kInv4Pi;
        if (L.d.kind == prgpu::LightExSun)  // sky.cpp:107-109
            return dot(normalize(wi), v3(L.d.direction)) >= L.d.cosFalloffEnd ? 1 / (2 * kPi * (1 - L.d.cosFalloffEnd))
                                                                               : 0.f;
        return 0;
    }
    int light_sample(bool useBvh, V3 p, V3 n, float u, float *pmf) const {
        if (useBvh) return lbvh.sample(p, n, u, pmf);
        if (lights.empty()) return -1;
        return power.sample(u, pmf);
    }
    float light_pmf(bool useBvh, V3 p, V3 n, int li) const {
        if (useBvh) return lbvh.pmf(p, n, li);
        return power.pmf(li);
    }

    // ---- media ----
    SS sigma(int sid, float scale, const Lambda &lam) const { return spec.sample(sid, lam) * scale; }

    SS medium_tr(int mi, const Ray &ray, float tMax, const Lambda &lam) const {
        const MediumEx &m = media[size_t(mi)];
        if (m.kind == prgpu::MediumHomogeneous) {  // medium.cpp:15-23
            SS st = sigma(m.sigmaA, m.scale, lam) + sigma(m.sigmaS, m.scale, lam);
            if (std::isinf(tMax)) {
                SS r;
                for (int i = 0; i < 4; ++i) r.v[i] = st.v[i] > 0 ? 0.f : 1.f;
                return r;
            }
            SS e;
            for (int i = 0; i < 4; ++i) e.v[i] = -st.v[i] * tMax;
            return ss_exp(e);
        }
        // Grid: ratio tracking with a hashed RNG (medium.cpp:205-230).
        Ray mray;
        float t0, t1;
        if (!grid_clip(m, ray, tMax, &mray, &t0, &t1)) return SS(1.f);
        SS st = (spec.sample(m.sigmaA, lam) + spec.sample(m.sigmaS, lam)) * m.scale;
        float mu = m.maxDensity * st.maxc();
        if (mu <= 0) return SS(1.f);
        Pcg32 rng;
        rng.set_sequence(hash_args(ray.o, ray.d), hash_args(tMax));
        SS T(1.f);
        float t = t0;
        for (int iter = 0; iter < 100000; ++iter) {
            t -= std::log(1 - rng.uniform()) / mu;
            if (t >= t1) break;
            float d = grid_density(m, mray.at(t));
            T = T * ss_clamp0(SS(1.f) - st * (d / mu));
            float mx = T.maxc();
            if (mx <= 0) return SS(0.f);
            if (mx < 0.1f) {
                if (rng.uniform() < 0.5f) return SS(0.f);
                T = T * 2.f;
            }
        }
        return T;
    }

    DistSample medium_sample(int mi, const Ray &ray, float tMax, float u, const Lambda &lam) const {
        const MediumEx &m = media[size_t(mi)];
        DistSample ds;
        if (m.kind == prgpu::MediumHomogeneous) {  // SampleDistanceHero, medium.cpp:71-103
            SS ss = sigma(m.sigmaS, m.scale, lam);
            SS st = sigma(m.sigmaA, m.scale, lam) + ss;
            float t = st.v[0] > 0 ? -std::log(1 - std::min(u, kOneMinusEps)) / st.v[0] : kInf;
            if (t < tMax) {
                SS e;
                for (int i = 0; i < 4; ++i) e.v[i] = -st.v[i] * t;
                SS Tr = ss_exp(e);
                float pdf0 = st.v[0] * Tr.v[0];
                ds.scattered = true;
                ds.t = t;
                if (!(pdf0 > 0)) {
                    ds.weight = SS(0.f);
                    return ds;
                }
                ds.weight = Tr * ss / pdf0;
                ds.pdfRatio = st * Tr / pdf0;
                return ds;
            }
            SS Tr = medium_tr(mi, ray, tMax, lam);
            ds.scattered = false;
            ds.t = tMax;
            if (!(Tr.v[0] > 0)) {
                ds.weight = SS(0.f);
                return ds;
            }
            ds.weight = Tr / Tr.v[0];
            ds.pdfRatio = Tr / Tr.v[0];
            return ds;
        }
        // Grid: spectral tracking (medium.cpp:232-285).
        ds.weight = SS(1.f);
        ds.t = tMax;
        Ray mray;
        float t0, t1;
        if (!grid_clip(m, ray, tMax, &mray, &t0, &t1)) return ds;
        SS sa = spec.sample(m.sigmaA, lam) * m.scale, ssc = spec.sample(m.sigmaS, lam) * m.scale;
        SS st = sa + ssc;
        float mu = m.maxDensity * st.maxc();
        if (mu <= 0) return ds;
        SS lpdf = lam.pdfs();
        auto avg = [&](const SS &s) {
            float sum = 0;
            int n = 0;
            for (int i = 0; i < 4; ++i)
                if (lpdf.v[i] != 0) {
                    sum += std::abs(s.v[i]);
                    ++n;
                }
            return n ? sum / n : 0.f;
        };
        Pcg32 rng;
        rng.set_sequence(hash_args(u, ray.o), hash_args(ray.d, tMax));
        float t = t0;
        for (int iter = 0; iter < 100000; ++iter) {
            t -= std::log(1 - rng.uniform()) / mu;
            if (t >= t1) break;
            float d = grid_density(m, mray.at(t));
            SS sA = sa * d, sS = ssc * d;
            SS sN = SS(mu) - sA - sS;
            float pa = avg(sA), ps = avg(sS), pn = avg(sN);
            float norm = pa + ps + pn;
            if (norm <= 0) continue;
            float xi = rng.uniform() * norm;
            if (xi < ps) {
                ds.scattered = true;
                ds.t = t;
                ds.weight = ds.weight * (sS / (mu * ps / norm));
                return ds;
            }
            if (xi < ps + pa) {
                ds.weight = SS(0.f);
                ds.t = t;
                return ds;
            }
            ds.weight = ds.weight * (sN / (mu * pn / norm));
            if (!ds.weight.nonzero()) return ds;
        }
        ds.t = tMax;
        return ds;
    }

    bool grid_clip(const MediumEx &m, const Ray &ray, float tMax, Ray *mray, float *t0, float *t1) const {  // medium.cpp:182-203
        Affine w2m;
        std::memcpy(w2m.m, m.worldToMedium, sizeof(w2m.m));
        *mray = Ray{w2m.point(ray.o), w2m.vector(ray.d)};
        *t0 = 0;
        *t1 = tMax;
        Box b{v3(m.boundsMin), v3(m.boundsMax)};
        if (b.empty()) return std::isfinite(tMax);
        for (int a = 0; a < 3; ++a) {
            float inv = 1 / mray->d[a];
            float tn = (b.lo[a] - mray->o[a]) * inv, tf = (b.hi[a] - mray->o[a]) * inv;
            if (tn > tf) std::swap(tn, tf);
            if (std::isnan(tn) || std::isnan(tf)) {
                if (mray->o[a] < b.lo[a] || mray->o[a] > b.hi[a]) return false;
                continue;
            }
            *t0 = std::max(*t0, tn);
            *t1 = std::min(*t1, tf);
            if (*t0 > *t1) return false;
        }
        return true;
    }

    float grid_at(const MediumEx &m, int x, int y, int z) const {
        x = clampv(x, 0, m.nx - 1);
        y = clampv(y, 0, m.ny - 1);
        z = clampv(z, 0, m.nz - 1);
        return grid[size_t(m.gridOffset) + (size_t(z) * m.ny + y) * m.nx + x];
    }
    float grid_density(const MediumEx &m, V3 p) const {  // medium.cpp:117-128
        Box b{v3(m.boundsMin), v3(m.boundsMax)};
        if (!b.inside(p)) return 0;
        V3 o = b.offset(p);
        float x = o.x * m.nx - 0.5f, y = o.y * m.ny - 0.5f, z = o.z * m.nz - 0.5f;
        int x0 = int(std::floor(x)), y0 = int(std::floor(y)), z0 = int(std::floor(z));
        float dx = x - x0, dy = y - y0, dz = z - z0;
        float c00 = lerpf(dx, grid_at(m, x0, y0, z0), grid_at(m, x0 + 1, y0, z0));
        float c10 = lerpf(dx, grid_at(m, x0, y0 + 1, z0), grid_at(m, x0 + 1, y0 + 1, z0));
        float c01 = lerpf(dx, grid_at(m, x0, y0, z0 + 1), grid_at(m, x0 + 1, y0, z0 + 1));
        float c11 = lerpf(dx, grid_at(m, x0, y0 + 1, z0 + 1), grid_at(m, x0 + 1, y0 + 1, z0 + 1));
        return lerpf(dz, lerpf(dy, c00, c10), lerpf(dy, c01, c11));
    }
    float medium_g(int mi) const { return clampv(media[size_t(mi)].g, -0.99f, 0.99f); }

    // ---- geometry queries (src/scene/scene.cpp:29-59) ----
    bool intersect(const Ray &ray, float tMax, SurfHit *si, float *tHit) const {
        Bvh::Hit h;
        if (!bvh.intersect(ray, tMax, &h)) return false;
        shapes[size_t(h.prim)].interaction(ray, h.t, h.info, si);
        si->prim = h.prim;
        si->surface_mi = true;
        si->mIn = prims[size_t(h.prim)].mediumInside;
        si->mOut = prims[size_t(h.prim)].mediumOutside;
        if (tHit) *tHit = h.t;
        return true;
    }

    SS tr(const Intr &p0, const Intr &p1, const Lambda &lam) const {
        Ray ray = p0.spawn_to(p1);
        int medium = p0.get_medium(ray.d);
        SS T(1.f);
        for (int iter = 0; iter < 256; ++iter) {
            float tHit = 1 - kShadowEps;
            SurfHit si;
            bool hit = intersect(ray, 1 - kShadowEps, &si, &tHit);
            if (hit && !is_interface(si.prim)) return SS(0.f);
            if (medium >= 0) {
                float l = len(ray.d);
                T = T * medium_tr(medium, Ray{ray.o, ray.d / l}, (hit ? tHit : 1.f) * l, lam);
            }
            if (!T.nonzero()) return T;
            if (!hit) return T;
            medium = si.get_medium(ray.d);
            ray = si.spawn_to(p1);
        }
        return SS(0.f);
    }

    // A material's albedo at the hit: a spectrum, or an image texture as an albedo
    // (ImageSpectrumTexture::Evaluate with RGBAlbedoSample, texture.cpp:38-52).
    SS albedo(int id, const SurfHit &si, const Lambda &lam) const {
        const SpectrumEx &s = spec.specs[size_t(id)];
        if (s.kind != prgpu::SpecImage) return spec.sample(id, lam);
        const float *rec = spec.knots.data() + s.offset;
        float c[3], co[3];
        tex_rgb(rec, si.uv, c);
        for (int i = 0; i < 3; ++i) c[i] = c[i] * rec[8 + i];
        rgb_table_coeffs(spec.knots.data() + int(rec[11]), c, co);
        SS r;
        for (int i = 0; i < 4; ++i) r.v[i] = sigmoid_s(std::fma(lam.l[i], std::fma(lam.l[i], co[0], co[1]), co[2]));
        return r;
    }
    // Material::PerturbShading with the material's record (knots[m.shading - 1]); EvaluateSurface
    // (material.cpp:60-66) then keeps the perturbed shading normal and tangent in the interaction.
    void perturb_shading(SurfHit &si) const {
        const Material &m = mats[size_t(prims[size_t(si.prim)].material)];
        if (m.shading == 0) return;
        perturb_record(spec.knots.data(), spec.knots.data() + (m.shading - 1), si.uv, si.n, si.dpdu, si.dpdv, &si.ns,
                       &si.dpdus);
    }

    // ---- materials (src/materials/material.cpp:5-14, 64-166) ----
    int checker(const prgpu::Material &m, const SurfHit &si, int sid) const {  // CheckerSpectrumTexture (texture.h:126-134)
        if (m.albedoB >= 0 && m.checkerScale != 0) {
            float sx = si.uv.x * m.checkerScale + 0.f, sy = si.uv.y * m.checkerScale + 0.f;
            bool odd = (int(std::floor(sx)) + int(std::floor(sy))) & 1;
            if (odd) return m.albedoB;
        }
        return sid;
    }
    void diffuse_bx(const prgpu::Material &m, const SurfHit &si, const Lambda &lam, BxDF &bx) const {
        SS r = albedo(checker(m, si, m.albedo), si, lam);
        for (int i = 0; i < 4; ++i) r.v[i] = clampv(r.v[i], 0.f, 1.f);
        bx.type = BxDiffuse;
        bx.R = r;
    }
    // Conductor closure; under a coat its complex IOR is relative to the coat (eta, k / coatEta).
    void conductor_bx(const prgpu::Material &m, const SurfHit &si, const Lambda &lam, float coatEta, BxDF &bx) const {
        bx.type = BxConductor;
        bx.dist = TrDist(m.alphaX, m.alphaY);
        if (m.eta >= 0 && m.k >= 0) {
            bx.eta = spec.sample(m.eta, lam);
            bx.k = spec.sample(m.k, lam);
            if (coatEta != 0) {
                bx.eta = bx.eta / coatEta;
                bx.k = bx.k / coatEta;
            }
        } else {
            SS r = albedo(checker(m, si, m.albedo), si, lam);
            for (int i = 0; i < 4; ++i) r.v[i] = clampv(r.v[i], 0.f, 0.9999f);
            bx.eta = SS(1.f);
            bx.k = ss_sqrt(r) * 2.f / ss_sqrt(ss_clamp0(SS(1.f) - r));
        }
    }
    BSDF eval_surface(const SurfHit &si, Lambda &lam) const {
        BSDF b;
        const Primitive &pr = prims[size_t(si.prim)];
        const Material &m = mats[size_t(pr.material)];
        Closure &c = b.c;
        BxDF &bx = c.bx;
        if (m.type == prgpu::MatDiffuse) {
            diffuse_bx(m, si, lam, bx);
        } else if (m.type == prgpu::MatConductor) {
            conductor_bx(m, si, lam, 0.f, bx);
        } else if (m.type == prgpu::MatCoatedDiffuse || m.type == prgpu::MatCoatedConductor) {
            // MakeCoatTop (material.cpp:120-131): a rough dielectric coat at the hero wavelength's IOR.
            float ce = spec.eval(m.coatEta, lam.l[0]);
            if (!spec.is_constant(m.coatEta)) lam.terminate_secondary();
            if (ce == 0) ce = 1;
            Layered &L = c.lay;
            c.layered = true;
            L.top.type = BxDielectric;
            L.top.etaD = ce;
            L.top.dist = TrDist(m.coatAlpha, m.coatAlpha);
            L.thick = std::max(m.thickness, std::numeric_limits<float>::min());
            L.g = m.coatG;
            L.albedo = m.coatAlbedo >= 0 ? ss_clamp0(spec.sample(m.coatAlbedo, lam)) : SS(0.f);
            L.twoSided = true;
            L.maxDepth = m.coatMaxDepth;
            L.nSamples = std::max(1, m.coatSamples);
            if (m.type == prgpu::MatCoatedDiffuse) diffuse_bx(m, si, lam, L.bottom);
            else conductor_bx(m, si, lam, ce, L.bottom);
        } else {  // dielectric
            float e = spec.eval(m.eta, lam.l[0]);
            if (!spec.is_constant(m.eta)) lam.terminate_secondary();
            if (e == 0) e = 1;
            bx.etaD = e;
            if (m.thin) {
                bx.type = BxThin;
            } else {
                bx.type = BxDielectric;
                bx.dist = TrDist(m.alphaX, m.alphaY);
            }
        }
        if (m.sheenWeight > 0 && m.sheenColor >= 0) {  // material.cpp:8-12
            c.sheen = true;
            SS col = ss_clamp0(spec.sample(m.sheenColor, lam));
            c.sh.color = col;
            c.sh.alpha = clampv(m.sheenRoughness, 0.05f, 1.f);
            c.sh.weight = clampv(m.sheenWeight, 0.f, 1.f);
            c.sh.colorMax = clampv(col.maxc(), 0.f, 1.f);
            c.sh.table = sheen_table();
        }
        b.frame = Frame::from_xz(si.dpdus, si.ns);
        b.valid = true;
        return b;
    }
};

// =============================================================================================
// Path tracer (src/integrators/path.cpp:82-238)

struct Settings {
    int maxDepth = -1;
    int rrStart = 1;
    bool rr = true;
    bool lightBvh = true;
};

static SS sample_ld(const Scene &sc, const Settings &st, const Intr &intr, const BSDF *bsdf, float g, V3 ns, Lambda &lam,
                    Sampler &smp) {
    float uLight = smp.get1();
    V2 uL = smp.get2();
    float pmf = 0;
    V3 cp = intr.p, cn = intr.n;
    int li = sc.light_sample(st.lightBvh, cp, cn, uLight, &pmf);
    if (li < 0 || pmf == 0) return SS(0.f);
    auto ls = sc.sample_li(li, cp, cn, uL, lam);
    if (!ls || !ls->L.nonzero() || ls->pdf == 0) return SS(0.f);
    V3 wo = intr.wo, wi = ls->wi;
    SS f;
    float scatterPdf;
    if (bsdf) {
        f = bsdf->f(wo, wi) * absdot(wi, ns);
        if (!f.nonzero()) return SS(0.f);
        scatterPdf = bsdf->pdf(wo, wi);
    } else {
        float p = hg_phase(dot(wo, wi), g);
        f = SS(p);
        scatterPdf = p;
    }
    SS T = sc.tr(intr, ls->pLight, lam);
    if (!T.nonzero()) return SS(0.f);
    float lightPdf = pmf * ls->pdf;
    if (sc.lights[size_t(li)].delta()) return f * T * ls->L / lightPdf;
    float w = power_heuristic(lightPdf, scatterPdf);
    return f * T * ls->L * (w / lightPdf);
}

static SS li(const Scene &sc, const Settings &st, Ray ray, Lambda &lam, Sampler &smp) {
    SS L(0.f), beta(1.f);
    bool specularBounce = false;
    int depth = 0;
    float etaScale = 1;
    float prevPdf = 1;
    V3 prevP, prevN;
    int medium = sc.cameraMedium;
    const int maxDepth = st.maxDepth > 0 ? st.maxDepth : std::numeric_limits<int>::max();
    int crossings = 0;
    SS rPath(1.f), Lhero(0.f);
    auto add = [&](const SS &c) {
        Lhero = Lhero + c;
        float a = rPath.avg();
        if (a > 0 && std::isfinite(a)) L = L + c / a;
    };
    auto eff_beta = [&] {
        if (lam.secondary_terminated()) return beta;
        float a = rPath.avg();
        return a > 0 && std::isfinite(a) ? beta / a : SS(0.f);
    };
    auto add_emitted = [&](int li, const SS &Le, V3 dir) {
        if (!Le.nonzero()) return;
        if (depth == 0 || specularBounce) {
            add(beta * Le);
        } else {
            float pmf = sc.light_pmf(st.lightBvh, prevP, prevN, li);
            float lightPdf = pmf * sc.pdf_li(li, prevP, dir);
            float w = power_heuristic(prevPdf, lightPdf);
            add(beta * Le * w);
        }
    };
    auto roulette = [&]() -> bool {  // returns true when the path is terminated
        if (st.rr && depth > st.rrStart) {
            float m = eff_beta().maxc() * etaScale;
            if (m < 1) {
                float q = std::max(0.f, 1 - m);
                if (smp.get1() < q) return true;
                beta = beta / (1 - q);
            }
        }
        return false;
    };

    while (true) {
        float tHit = kInf;
        SurfHit si;
        bool hit = sc.intersect(ray, kInf, &si, &tHit);

        if (medium >= 0) {
            DistSample ds = sc.medium_sample(medium, ray, hit ? tHit : kInf, smp.get1(), lam);
            beta = beta * ds.weight;
            rPath = rPath * ds.pdfRatio;
            if (!beta.nonzero()) break;
            if (ds.scattered) {
                if (depth++ >= maxDepth) break;
                Intr mi;
                mi.p = ray.at(ds.t);
                mi.wo = -ray.d;
                mi.medium = medium;
                float g = sc.medium_g(medium);
                add(beta * sample_ld(sc, st, mi, nullptr, g, V3(), lam, smp));
                float phasePdf;
                V3 wi = sample_hg(mi.wo, g, smp.get2(), &phasePdf);
                if (phasePdf == 0) break;
                prevPdf = phasePdf;
                prevP = mi.p;
                prevN = mi.n;
                specularBounce = false;
                ray = Ray{mi.p, wi};
                if (roulette()) break;
                continue;
            }
        }

        if (!hit) {
            for (size_t i = 0; i < sc.lights.size(); ++i) {
                const LightObj &Li = sc.lights[i];
                if (Li.d.kind == prgpu::LightExEnvironment)  // EnvironmentLight::Le (light.cpp:263-265)
                    add_emitted(int(i), Li.env ? Li.env->lookup(sc.spec, Li.d.scale, ray.d, lam)
                                               : sc.spec.sample(Li.d.spectrum, lam) * Li.d.scale,
                                ray.d);
                else if (Li.d.kind == prgpu::LightExSun &&  // SunLight::Le (sky.cpp:92-95)
                         dot(normalize(ray.d), v3(Li.d.direction)) >= Li.d.cosFalloffEnd)
                    add_emitted(int(i), sc.spec.sample(Li.d.spectrum, lam) * Li.d.scale, ray.d);
            }
            break;
        }

        const Primitive &pr = sc.prims[size_t(si.prim)];
        if (pr.light >= 0)
            add_emitted(pr.light, sc.area_L(sc.lights[size_t(pr.light)], si.n, -ray.d, lam), ray.d);

        if (sc.is_interface(si.prim)) {
            if (++crossings > 1000) break;
            medium = si.get_medium(ray.d);
            ray = si.spawn(ray.d);
            continue;
        }

        if (!sc.has_bsdf(si.prim)) break;
        sc.perturb_shading(si);
        BSDF bsdf = sc.eval_surface(si, lam);
        if (depth++ >= maxDepth) break;

        if (bsdf.has_non_specular()) add(beta * sample_ld(sc, st, si, &bsdf, 0, si.ns, lam, smp));

        V3 wo = si.wo;
        float uc = smp.get1();
        V2 u2 = smp.get2();
        auto bs = bsdf.sample(wo, uc, u2);
        if (!bs) break;
        beta = beta * (bs->f * (absdot(bs->wi, si.ns) / bs->pdf));
        prevPdf = bs->proportional ? bsdf.pdf(wo, bs->wi) : bs->pdf;  // path.cpp:220
        specularBounce = (bs->flags & kSpecular) != 0;
        if (bs->flags & kTrans) etaScale *= sq(bs->eta);
        prevP = si.p;
        prevN = si.n;
        ray = si.spawn(bs->wi);
        medium = si.get_medium(bs->wi);

        if (!beta.nonzero() || beta.has_nan()) break;
        if (roulette()) break;
    }
    return lam.secondary_terminated() ? Lhero : L;
}

// Camera ray (src/cameras/camera.cpp:36-48).
static bool camera_ray(const Scene &sc, V2 pFilm, V2 pLens, Ray *out) {
    const prgpu::CameraEx &c = sc.cam;
    float aspect = float(c.width) / float(c.height);
    float x = (2 * pFilm.x / c.width - 1) * c.tanHalfFov * aspect;
    float y = (1 - 2 * pFilm.y / c.height) * c.tanHalfFov;
    V3 dir = normalize(V3(x, y, 1));
    V3 o(0, 0, 0);
    float lensRadius = std::max(0.f, c.lensRadius);
    float focal = std::max(1e-4f, c.focalDistance);
    if (lensRadius > 0) {
        V2 pl = sample_disk_concentric(pLens);
        pl.x *= lensRadius;
        pl.y *= lensRadius;
        float ft = focal / dir.z;
        V3 pFocus = dir * ft;
        o = V3(pl.x, pl.y, 0);
        dir = normalize(pFocus - o);
    }
    *out = Ray{sc.c2w.point(o), normalize(sc.c2w.vector(dir))};
    return true;
}

struct SampleOut {
    SS L;
    Lambda lam;
    V2 pFilm;
    float weight = 0;
};

static SampleOut render_sample(const Scene &sc, const Settings &st, int px, int py, int si, uint64_t seed) {
    // path.cpp:61-76
    const int m = sc.margin;
    const int sw = sc.cam.width + 2 * m;
    Sampler smp;
    smp.start(uint64_t(py) * sw + px, uint64_t(si), seed);
    SampleOut o;
    o.lam = sample_visible(smp.get1());
    V2 u = smp.get2();
    o.pFilm = {px - m + u.x, py - m + u.y};
    V2 lens = smp.get2();
    float weight = 1;
    Ray ray;
    if (camera_ray(sc, o.pFilm, lens, &ray) && weight > 0) {
        o.L = li(sc, st, ray, o.lam, smp) * weight;
        o.weight = weight;
    }
    return o;
}

// =============================================================================================
// Scene construction from the flat description

static Scene *build_scene(const prgpu::SceneDescEx *d) {
    auto sc = std::make_unique<Scene>();
    sc->spec.specs.assign(d->spectra, d->spectra + d->spectrumCount);
    if (d->knotCount > 0) sc->spec.knots.assign(d->knots, d->knots + d->knotCount);
    sc->prims.assign(d->prims, d->prims + d->primCount);
    sc->mats.assign(d->materials, d->materials + d->materialCount);
    if (d->mediumCount > 0) sc->media.assign(d->media, d->media + d->mediumCount);
    if (d->gridCount > 0) sc->grid.assign(d->gridData, d->gridData + d->gridCount);
    sc->cameraMedium = d->cameraMedium;
    sc->cam = d->camera;
    std::memcpy(sc->c2w.m, d->camera.cameraToWorld, sizeof(sc->c2w.m));
    sc->shapes.resize(sc->prims.size());
    for (size_t i = 0; i < sc->prims.size(); ++i) sc->shapes[i].p = &sc->prims[i];
    sc->bvh.build(sc->shapes);
    sc->init_filter(d->camera.filterType, d->camera.filterRadius);
    // Scene::Build (scene.cpp:9-27): world sphere from the BVH bounds and the camera position.
    Box b = sc->bvh.bounds();
    b = join(b, sc->c2w.point(V3(0, 0, 0)));
    b.bounding_sphere(&sc->worldCenter, &sc->worldRadius);
    sc->worldRadius = std::max(sc->worldRadius, 1e-3f);
    for (int i = 0; i < d->lightCount; ++i) {
        LightObj L;
        L.d = d->lights[i];
        if (L.d.kind == prgpu::LightExArea) L.area = sc->shapes[size_t(L.d.primitive)].area();
        if (L.d.kind == prgpu::LightExEnvironment && L.d.profileSize > 0) {
            L.env = std::make_shared<EnvImage>();
            L.env->init(d->knots + L.d.profileOffset, d->knots);
        }
        sc->lights.push_back(L);
    }
    // Preprocess of distant/environment lights uses the same bounds: radius = max(r, 1e-3).
    std::vector<std::optional<LightB>> lbs;
    std::vector<float> power;
    for (size_t i = 0; i < sc->lights.size(); ++i) {
        lbs.push_back(sc->light_bounds(int(i)));
        float p = sc->light_power(int(i));
        power.push_back(std::isfinite(p) && p > 0 ? p : 0.f);
    }
    // LightSampler (light.cpp:326-340): every light keeps a tiny share.
    float maxP = 0;
    for (float p : power) maxP = std::max(maxP, p);
    for (float &p : power) p = std::max(p, maxP > 0 ? maxP * 1e-4f : 1.f);
    if (!power.empty()) sc->power.init(power);
    sc->lbvh.build(lbs);
    return sc.release();
}

}  // namespace orc

// =============================================================================================
// C API

using namespace orc;

namespace {

Settings settings_of(const prgpu::RenderParamsEx *p) {
    Settings s;
    s.maxDepth = p->maxDepth;
    s.rrStart = p->rrStartDepth;
    s.rr = p->russianRoulette != 0;
    s.lightBvh = p->lightSampler == prgpu::LightSamplerBVH;
    return s;
}

// Persistent-worker parallel loop over [0, n) with an atomic counter (src/core/parallel.cpp:23-131).
template <class F> void parallel_for(int64_t n, int threads, const F &fn) {
    if (threads <= 1 || n <= 1) {
        for (int64_t i = 0; i < n; ++i) fn(i, 0);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (int64_t i; (i = next.fetch_add(1)) < n;) fn(i, t);
        });
    for (auto &th : pool) th.join();
}

inline void atomic_add(std::atomic<double> &a, double v) {
    double old = a.load(std::memory_order_relaxed);
    while (!a.compare_exchange_weak(old, old + v, std::memory_order_relaxed)) {
    }
}

}  // namespace

extern "C" {

void *or_scene_create(const prgpu::SceneDescEx *d) { return build_scene(d); }
void or_scene_destroy(void *h) { delete static_cast<Scene *>(h); }

int or_scene_info(void *h, int64_t *out) {
    const Scene &sc = *static_cast<Scene *>(h);
    out[0] = int64_t(sc.bvh.nodes().size());
    out[1] = int64_t(sc.lbvh.nodes.size());
    out[2] = int64_t(sc.lbvh.infinite.size());
    out[3] = sc.margin;
    return 0;
}

void or_world_sphere(void *h, float *out) {
    const Scene &sc = *static_cast<Scene *>(h);
    out[0] = sc.worldCenter.x;
    out[1] = sc.worldCenter.y;
    out[2] = sc.worldCenter.z;
    out[3] = sc.worldRadius;
}

// Binary BVH node dump: 8 floats per node (lo xyz, hi xyz, offset, nPrims) + axis in order[].
int or_bvh_dump(void *h, float *nodes, int32_t *order) {
    const Scene &sc = *static_cast<Scene *>(h);
    const auto &ns = sc.bvh.nodes();
    for (size_t i = 0; i < ns.size(); ++i) {
        float *o = nodes + 8 * i;
        o[0] = ns[i].b.lo.x; o[1] = ns[i].b.lo.y; o[2] = ns[i].b.lo.z;
        o[3] = ns[i].b.hi.x; o[4] = ns[i].b.hi.y; o[5] = ns[i].b.hi.z;
        o[6] = float(ns[i].offset); o[7] = float(ns[i].nPrims);
    }
    for (size_t i = 0; i < sc.bvh.order().size(); ++i) order[i] = sc.bvh.order()[i];
    return int(ns.size());
}

void or_trace(void *h, const prgpu::RayEx *rays, int64_t n, int anyHit, prgpu::HitEx *hits, int threads) {
    const Scene &sc = *static_cast<Scene *>(h);
    parallel_for((n + 1023) / 1024, threads, [&](int64_t blk, int) {
        for (int64_t i = blk * 1024; i < std::min(n, (blk + 1) * 1024); ++i) {
            Ray r{v3(rays[i].o), v3(rays[i].d)};
            prgpu::HitEx &o = hits[i];
            if (anyHit) {
                o.prim = sc.bvh.intersect_any(r, rays[i].tMax) ? 0 : -1;
                o.t = 0; o.b1 = o.b2 = 0;
                continue;
            }
            Bvh::Hit bh;
            if (sc.bvh.intersect(r, rays[i].tMax, &bh)) {
                o.t = bh.t; o.prim = bh.prim; o.b1 = bh.info.b1; o.b2 = bh.info.b2;
            } else {
                o.t = 0; o.prim = -1; o.b1 = o.b2 = 0;
            }
        }
    });
}

// Closest hit by testing every primitive; ties[i] = number of primitives whose hit distance
// equals the minimum exactly (1 = unique).
void or_trace_brute(void *h, const prgpu::RayEx *rays, int64_t n, prgpu::HitEx *hits, int32_t *ties, int threads) {
    const Scene &sc = *static_cast<Scene *>(h);
    parallel_for(n, threads, [&](int64_t i, int) {
        Ray r{v3(rays[i].o), v3(rays[i].d)};
        float best = rays[i].tMax;
        int bp = -1, cnt = 0;
        HitInfo bi;
        for (size_t k = 0; k < sc.shapes.size(); ++k) {
            float t;
            HitInfo info;
            if (!sc.shapes[k].intersect(r, rays[i].tMax, &t, &info)) continue;
            if (t < best) {
                best = t; bp = int(k); bi = info; cnt = 1;
            } else if (t == best && bp >= 0) {
                ++cnt;
            }
        }
        hits[i].t = bp >= 0 ? best : 0;
        hits[i].prim = bp;
        hits[i].b1 = bi.b1;
        hits[i].b2 = bi.b2;
        ties[i] = cnt;
    });
}

// Light-BVH dump: per node 12 floats: bounds lo/hi, w, phi, cosTheta_o, cosTheta_e ; ints: child1, light, twoSided.
int or_light_bvh_dump(void *h, float *f, int32_t *iv) {
    const Scene &sc = *static_cast<Scene *>(h);
    const auto &ns = sc.lbvh.nodes;
    for (size_t i = 0; i < ns.size(); ++i) {
        const LightB &b = ns[i].lb;
        float *o = f + 12 * i;
        o[0] = b.bounds.lo.x; o[1] = b.bounds.lo.y; o[2] = b.bounds.lo.z;
        o[3] = b.bounds.hi.x; o[4] = b.bounds.hi.y; o[5] = b.bounds.hi.z;
        o[6] = b.w.x; o[7] = b.w.y; o[8] = b.w.z;
        o[9] = b.phi; o[10] = b.cosTheta_o; o[11] = b.cosTheta_e;
        iv[3 * i + 0] = ns[i].child1;
        iv[3 * i + 1] = ns[i].light >= 0 ? sc.lbvh.bounded[size_t(ns[i].light)] : -1;
        iv[3 * i + 2] = b.twoSided ? 1 : 0;
    }
    return int(ns.size());
}

int or_light_sample(void *h, int useBvh, const float *p, const float *n, float u, float *pmf) {
    const Scene &sc = *static_cast<Scene *>(h);
    *pmf = 0;
    return sc.light_sample(useBvh != 0, V3(p[0], p[1], p[2]), V3(n[0], n[1], n[2]), u, pmf);
}

float or_light_pmf(void *h, int useBvh, const float *p, const float *n, int light) {
    const Scene &sc = *static_cast<Scene *>(h);
    return sc.light_pmf(useBvh != 0, V3(p[0], p[1], p[2]), V3(n[0], n[1], n[2]), light);
}

void or_camera_ray(void *h, float fx, float fy, float lx, float ly, float *out) {
    const Scene &sc = *static_cast<Scene *>(h);
    Ray r;
    camera_ray(sc, V2{fx, fy}, V2{lx, ly}, &r);
    out[0] = r.o.x; out[1] = r.o.y; out[2] = r.o.z;
    out[3] = r.d.x; out[4] = r.d.y; out[5] = r.d.z;
}

// One camera sample per extended-film pixel at sample index params->firstSample.
void or_sample_radiance(void *h, const prgpu::RenderParamsEx *params, prgpu::SampleRecordEx *out, int threads) {
    const Scene &sc = *static_cast<Scene *>(h);
    Settings st = settings_of(params);
    const int m = sc.margin, sw = sc.cam.width + 2 * m, sh = sc.cam.height + 2 * m;
    parallel_for(int64_t(sh), threads, [&](int64_t py, int) {
        for (int px = 0; px < sw; ++px) {
            SampleOut s = render_sample(sc, st, px, int(py), params->firstSample, params->seed);
            prgpu::SampleRecordEx &r = out[size_t(py) * sw + px];
            for (int i = 0; i < 4; ++i) {
                r.L[i] = s.L.v[i];
                r.lambda[i] = s.lam.l[i];
                r.pdf[i] = s.lam.pdf[i];
            }
            r.pFilmX = s.pFilm.x;
            r.pFilmY = s.pFilm.y;
            r.weight = s.weight;
            r.flags = s.L.has_nan() ? 1u : 0u;
        }
    });
}

// Progressive render (path.cpp:8-80): one pass per sample index over 16x16 tiles of the
// extended film, filter splatting with atomic FP64 adds (film.cpp:133-157). When sumsq is
// non-null each pass is accumulated separately and its square added per pixel. Returns the
// render wall time in seconds (scene build excluded).
double or_render(void *h, const prgpu::RenderParamsEx *params, double *sum, double *sumsq, int threads) {
    const Scene &sc = *static_cast<Scene *>(h);
    Settings st = settings_of(params);
    const int W = sc.cam.width, H = sc.cam.height, m = sc.margin;
    const int sw = W + 2 * m, sh = H + 2 * m;
    const int tile = 16, tx = (sw + tile - 1) / tile, ty = (sh + tile - 1) / tile;
    const size_t npix = size_t(W) * H * 3;
    std::unique_ptr<std::atomic<double>[]> acc(new std::atomic<double>[npix]);
    auto clear = [&] {
        for (size_t i = 0; i < npix; ++i) acc[i].store(0.0, std::memory_order_relaxed);
    };
    clear();
    if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
    const float clampY = params->clampLuminance;
    auto t0 = std::chrono::steady_clock::now();
    for (int pass = 0; pass < params->spp; ++pass) {
        const int si = params->firstSample + pass;
        parallel_for(int64_t(tx) * ty, threads, [&](int64_t t, int) {
            int tX = int(t % tx), tY = int(t / tx);
            for (int py = tY * tile; py < std::min(sh, (tY + 1) * tile); ++py)
                for (int px = tX * tile; px < std::min(sw, (tX + 1) * tile); ++px) {
                    SampleOut s = render_sample(sc, st, px, py, si, params->seed);
                    if (s.weight <= 0 || s.L.has_nan() || !s.L.nonzero()) continue;
                    XYZ xyz = to_xyz(s.L, s.lam);
                    if (clampY > 0 && xyz.y > clampY) {
                        float k = clampY / xyz.y;
                        xyz = {xyz.x * k, xyz.y * k, xyz.z * k};
                    }
                    if (!std::isfinite(xyz.x) || !std::isfinite(xyz.y) || !std::isfinite(xyz.z)) continue;
                    float r = sc.filterRadius;
                    V2 p = s.pFilm;
                    int x0 = std::max(0, int(std::ceil(p.x - r - 0.5f)));
                    int x1 = std::min(W - 1, int(std::floor(p.x + r - 0.5f)));
                    int y0 = std::max(0, int(std::ceil(p.y - r - 0.5f)));
                    int y1 = std::min(H - 1, int(std::floor(p.y + r - 0.5f)));
                    for (int y = y0; y <= y1; ++y)
                        for (int x = x0; x <= x1; ++x) {
                            float w = sc.filter_1d(p.x - (x + 0.5f)) * sc.filter_1d(p.y - (y + 0.5f)) * sc.filterNorm;
                            if (w == 0) continue;
                            size_t idx = 3 * (size_t(y) * W + x);
                            atomic_add(acc[idx + 0], double(xyz.x) * w);
                            atomic_add(acc[idx + 1], double(xyz.y) * w);
                            atomic_add(acc[idx + 2], double(xyz.z) * w);
                        }
                }
        });
        if (sumsq) {
            for (size_t i = 0; i < npix; ++i) {
                double v = acc[i].load(std::memory_order_relaxed);
                sum[i] += v;
                sumsq[i] += v * v;
            }
            clear();
        }
    }
    if (!sumsq)
        for (size_t i = 0; i < npix; ++i) sum[i] += acc[i].load(std::memory_order_relaxed);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---- component probes (pinned against the compiled reference components) ----

void or_rng_stream(uint64_t pixel, uint64_t sampleIndex, uint64_t seed, int n, float *out) {
    Sampler s;
    s.start(pixel, sampleIndex, seed);
    for (int i = 0; i < n; ++i) out[i] = s.get1();
}
void or_sample_visible(float u, float *l, float *pdf) {
    Lambda w = sample_visible(u);
    for (int i = 0; i < 4; ++i) {
        l[i] = w.l[i];
        pdf[i] = w.pdf[i];
    }
}
void or_cie(float l, float *xyz) {
    xyz[0] = cie_x(l);
    xyz[1] = cie_y(l);
    xyz[2] = cie_z(l);
}
float or_cie_y_integral() { return cie_y_integral(); }
void or_to_xyz(const float *L, const float *l, const float *pdf, float *xyz) {
    SS s;
    Lambda w;
    for (int i = 0; i < 4; ++i) {
        s.v[i] = L[i];
        w.l[i] = l[i];
        w.pdf[i] = pdf[i];
    }
    XYZ x = to_xyz(s, w);
    xyz[0] = x.x; xyz[1] = x.y; xyz[2] = x.z;
}
float or_spectrum_eval(const prgpu::SpectrumEx *s, const float *knots, int nknots, float l) {
    SpectrumTable t;
    t.specs.push_back(*s);
    if (nknots > 0) t.knots.assign(knots, knots + nknots);
    return t.eval(0, l);
}
float or_spectrum_max(const prgpu::SpectrumEx *s, const float *knots, int nknots) {
    SpectrumTable t;
    t.specs.push_back(*s);
    if (nknots > 0) t.knots.assign(knots, knots + nknots);
    return t.max_value(0);
}
float or_spectrum_y(const prgpu::SpectrumEx *s, const float *knots, int nknots) {
    SpectrumTable t;
    t.specs.push_back(*s);
    if (nknots > 0) t.knots.assign(knots, knots + nknots);
    return t.to_y(0);
}

// BxDF probe. desc: [type, alphaX, alphaY, etaD, R0..3, eta0..3, k0..3].
// op 0: f(wo, wi) -> out[0..3]; op 1: pdf -> out[0]; op 2: sample(wo, uc, u) ->
// out = f0..3, wi xyz, pdf, flags, eta, ok.
void or_bxdf(const float *desc, int op, const float *wo, const float *wi, float uc, const float *u, float *out) {
    BxDF b;
    b.type = BxType(int(desc[0]));
    b.dist = TrDist(desc[1], desc[2]);
    b.etaD = desc[3];
    for (int i = 0; i < 4; ++i) {
        b.R.v[i] = desc[4 + i];
        b.eta.v[i] = desc[8 + i];
        b.k.v[i] = desc[12 + i];
    }
    V3 o(wo[0], wo[1], wo[2]);
    if (op == 0) {
        SS f = b.f(o, V3(wi[0], wi[1], wi[2]));
        for (int i = 0; i < 4; ++i) out[i] = f.v[i];
    } else if (op == 1) {
        out[0] = b.pdf(o, V3(wi[0], wi[1], wi[2]));
    } else {
        auto s = b.sample(o, uc, V2{u[0], u[1]});
        for (int i = 0; i < 11; ++i) out[i] = 0;
        if (s) {
            for (int i = 0; i < 4; ++i) out[i] = s->f.v[i];
            out[4] = s->wi.x; out[5] = s->wi.y; out[6] = s->wi.z;
            out[7] = s->pdf;
            out[8] = float(s->flags);
            out[9] = s->eta;
            out[10] = 1;
        }
    }
}

// Closure probe: desc = [type (1 diffuse, 2 conductor, 3 dielectric, 4 thin, 5 coated diffuse,
// 6 coated conductor), alphaX, alphaY, etaD | coatEta, R[4], eta[4], k[4], coatAlpha, thickness,
// g, albedo[4], maxDepth, nSamples, twoSided, sheenWeight, sheenAlpha, sheenColor[4]].
// op 0: f (4), 1: pdf (1), 2: sample (f[4], wi[3], pdf, flags, eta, ok, proportional), 3: flags, exact.
void or_closure(const float *desc, int op, const float *wo, const float *wi, float uc, const float *u, int mode, int rt,
                float *out) {
    const int type = int(desc[0]);
    auto plain = [&](int t, float etaD, BxDF &b) {
        b.type = BxType(t);
        b.dist = TrDist(desc[1], desc[2]);
        b.etaD = etaD;
        for (int i = 0; i < 4; ++i) {
            b.R.v[i] = desc[4 + i];
            b.eta.v[i] = desc[8 + i];
            b.k.v[i] = desc[12 + i];
        }
    };
    Closure c;
    if (type == 5 || type == 6) {
        c.layered = true;
        Layered &L = c.lay;
        L.top.type = BxDielectric;
        L.top.etaD = desc[3];
        L.top.dist = TrDist(desc[16], desc[16]);
        plain(type == 5 ? BxDiffuse : BxConductor, 1.f, L.bottom);
        L.thick = std::max(desc[17], std::numeric_limits<float>::min());
        L.g = desc[18];
        for (int i = 0; i < 4; ++i) L.albedo.v[i] = desc[19 + i];
        L.maxDepth = int(desc[23]);
        L.nSamples = int(desc[24]);
        L.twoSided = desc[25] != 0;
    } else {
        plain(type, desc[3], c.bx);
    }
    if (desc[26] > 0) {
        c.sheen = true;
        for (int i = 0; i < 4; ++i) c.sh.color.v[i] = desc[28 + i];
        c.sh.alpha = clampv(desc[27], 0.05f, 1.f);
        c.sh.weight = clampv(desc[26], 0.f, 1.f);
        c.sh.colorMax = clampv(c.sh.color.maxc(), 0.f, 1.f);
        c.sh.table = sheen_table();
    }
    const Mode m = mode ? Mode::Importance : Mode::Radiance;
    V3 o(wo[0], wo[1], wo[2]);
    if (op == 0) {
        SS f = c.f(o, V3(wi[0], wi[1], wi[2]), m);
        for (int i = 0; i < 4; ++i) out[i] = f.v[i];
    } else if (op == 1) {
        out[0] = c.pdf(o, V3(wi[0], wi[1], wi[2]), m, uint32_t(rt));
    } else if (op == 2) {
        auto s = c.sample(o, uc, V2{u[0], u[1]}, m, uint32_t(rt));
        for (int i = 0; i < 12; ++i) out[i] = 0;
        if (s) {
            for (int i = 0; i < 4; ++i) out[i] = s->f.v[i];
            out[4] = s->wi.x; out[5] = s->wi.y; out[6] = s->wi.z;
            out[7] = s->pdf;
            out[8] = float(s->flags);
            out[9] = s->eta;
            out[10] = 1;
            out[11] = s->proportional ? 1.f : 0.f;
        }
    } else {
        out[0] = float(c.flags());
        out[1] = c.pdf_exact() ? 1.f : 0.f;
    }
}
float or_sheen_albedo(float cosTheta, float alpha) { return sheen_albedo(sheen_table(), cosTheta, alpha); }
void or_sheen_table(float *out) { std::memcpy(out, sheen_table(), sizeof(float) * kSheenMuN * kSheenAlphaN); }

float or_fr_dielectric(float c, float eta) { return fr_dielectric(c, eta); }
float or_fr_complex(float c, float eta, float k) { return fr_complex1(c, std::complex<float>(eta, k)); }
float or_hg(float c, float g) { return hg_phase(c, g); }
void or_sample_hg(const float *wo, float g, const float *u, float *wi, float *pdf) {
    V3 r = sample_hg(V3(wo[0], wo[1], wo[2]), g, V2{u[0], u[1]}, pdf);
    wi[0] = r.x; wi[1] = r.y; wi[2] = r.z;
}
void or_filter(int type, float radius, const float *x, int n, float *out, float *norm) {
    Scene s;
    s.init_filter(type, radius);
    for (int i = 0; i < n; ++i) out[i] = s.filter_1d(x[i]);
    *norm = s.filterNorm;
}

// Area-light sample from a reference point for a light index: out = L0..3, wi, pdf, p, n, ok.
void or_light_sample_li(void *h, int light, const float *p, const float *n, const float *u, const float *lam,
                        const float *lpdf, float *out) {
    const Scene &sc = *static_cast<Scene *>(h);
    Lambda w;
    for (int i = 0; i < 4; ++i) {
        w.l[i] = lam[i];
        w.pdf[i] = lpdf[i];
    }
    for (int i = 0; i < 15; ++i) out[i] = 0;
    auto ls = sc.sample_li(light, V3(p[0], p[1], p[2]), V3(n[0], n[1], n[2]), V2{u[0], u[1]}, w);
    if (!ls) return;
    for (int i = 0; i < 4; ++i) out[i] = ls->L.v[i];
    out[4] = ls->wi.x; out[5] = ls->wi.y; out[6] = ls->wi.z;
    out[7] = ls->pdf;
    out[8] = ls->pLight.p.x; out[9] = ls->pLight.p.y; out[10] = ls->pLight.p.z;
    out[11] = ls->pLight.n.x; out[12] = ls->pLight.n.y; out[13] = ls->pLight.n.z;
    out[14] = 1;
}
float or_light_pdf_li(void *h, int light, const float *p, const float *wi) {
    const Scene &sc = *static_cast<Scene *>(h);
    return sc.pdf_li(light, V3(p[0], p[1], p[2]), V3(wi[0], wi[1], wi[2]));
}
float or_light_power(void *h, int light) { return static_cast<Scene *>(h)->light_power(light); }

// Texture probes: the lookup of a texture record at n uv points (out n x RGB), and the shading
// perturbation of a shading record for n contexts (p, uv, n, ns, dpdus, dpdu, dpdv: 20 floats;
// out: ns, dpdus, changed: 7 floats).
void or_tex_lookup(const float *rec, const float *uv, int n, float *out) {
    for (int i = 0; i < n; ++i) tex_rgb(rec, V2{uv[2 * i], uv[2 * i + 1]}, out + 3 * size_t(i));
}
void or_perturb_shading(const float *knots, const float *shade, const float *ctx, int n, float *out) {
    for (int i = 0; i < n; ++i) {
        const float *c = ctx + 20 * size_t(i);
        V3 ns(c[8], c[9], c[10]), dpdus(c[11], c[12], c[13]);
        bool changed = perturb_record(knots, shade, V2{c[3], c[4]}, V3(c[5], c[6], c[7]), V3(c[14], c[15], c[16]),
                                      V3(c[17], c[18], c[19]), &ns, &dpdus);
        float *o = out + 7 * size_t(i);
        o[0] = ns.x, o[1] = ns.y, o[2] = ns.z, o[3] = dpdus.x, o[4] = dpdus.y, o[5] = dpdus.z;
        o[6] = changed ? 1.f : 0.f;
    }
}

// Medium probes: transmittance and distance sampling along a ray.
void or_medium_tr(void *h, int medium, const float *o, const float *d, float tMax, const float *lam, const float *lpdf,
                  float *out) {
    const Scene &sc = *static_cast<Scene *>(h);
    Lambda w;
    for (int i = 0; i < 4; ++i) {
        w.l[i] = lam[i];
        w.pdf[i] = lpdf[i];
    }
    SS t = sc.medium_tr(medium, Ray{V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2])}, tMax, w);
    for (int i = 0; i < 4; ++i) out[i] = t.v[i];
}
void or_medium_sample(void *h, int medium, const float *o, const float *d, float tMax, float u, const float *lam,
                      const float *lpdf, float *out) {
    const Scene &sc = *static_cast<Scene *>(h);
    Lambda w;
    for (int i = 0; i < 4; ++i) {
        w.l[i] = lam[i];
        w.pdf[i] = lpdf[i];
    }
    DistSample ds = sc.medium_sample(medium, Ray{V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2])}, tMax, u, w);
    out[0] = ds.scattered ? 1.f : 0.f;
    out[1] = ds.t;
    for (int i = 0; i < 4; ++i) {
        out[2 + i] = ds.weight.v[i];
        out[6 + i] = ds.pdfRatio.v[i];
    }
}

}  // extern "C"
