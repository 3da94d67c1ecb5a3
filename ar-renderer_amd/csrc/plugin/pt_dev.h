// Device-side shading of the wavefront path tracer: cameras, surface interactions, BSDFs,
// lights and the light BVH, media. Each function follows the reference CPU path tracer (cited
// per function) in operation order; see pt_common.h for the shared arithmetic.
#pragma once

#include "pt_common.h"
#include "trace_dev.h"

namespace pgpu {

// ---- sampled spectra -------------------------------------------------------------------------
struct SS {
    float v[4];
    PG_HD SS() : v{0, 0, 0, 0} {}
    PG_HD explicit SS(float c) : v{c, c, c, c} {}
    PG_HD SS(float a, float b, float c, float d) : v{a, b, c, d} {}
    PG_HD bool nonzero() const { return v[0] != 0 || v[1] != 0 || v[2] != 0 || v[3] != 0; }
    PG_HD float maxc() const { return maxf(maxf(v[0], v[1]), maxf(v[2], v[3])); }
    PG_HD float avg() const { return (v[0] + v[1] + v[2] + v[3]) / 4; }
    PG_HD bool has_nan() const {
        for (int i = 0; i < 4; ++i)
            if (is_nan(v[i]) || is_inf(v[i])) return true;
        return false;
    }
};
PG_HD SS operator+(const SS &a, const SS &b) { return {a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]}; }
PG_HD SS operator-(const SS &a, const SS &b) { return {a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2], a.v[3] - b.v[3]}; }
PG_HD SS operator*(const SS &a, const SS &b) { return {a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]}; }
PG_HD SS operator*(const SS &a, float s) { return {a.v[0] * s, a.v[1] * s, a.v[2] * s, a.v[3] * s}; }
PG_HD SS operator/(const SS &a, const SS &b) {
    SS r;
    for (int i = 0; i < 4; ++i) r.v[i] = b.v[i] != 0 ? a.v[i] / b.v[i] : 0.f;
    return r;
}
PG_HD SS operator/(const SS &a, float s) { return a * (1.f / s); }
PG_HD SS ss_sqrt(const SS &s) { return {safe_sqrt(s.v[0]), safe_sqrt(s.v[1]), safe_sqrt(s.v[2]), safe_sqrt(s.v[3])}; }
PG_HD SS ss_clamp0(const SS &s) { return {maxf(0.f, s.v[0]), maxf(0.f, s.v[1]), maxf(0.f, s.v[2]), maxf(0.f, s.v[3])}; }
PG_HD SS ss_exp_neg(const SS &st, float t) {  // Exp(-st * t)
    return {expf(-st.v[0] * t), expf(-st.v[1] * t), expf(-st.v[2] * t), expf(-st.v[3] * t)};
}
__device__ __forceinline__ SS ld4(const float4 *p, int i) {
    float4 f = p[i];
    return {f.x, f.y, f.z, f.w};
}
__device__ __forceinline__ void st4(float4 *p, int i, const SS &s) { p[i] = make_float4(s.v[0], s.v[1], s.v[2], s.v[3]); }

struct Lambda {
    float l[4], pdf[4];
    PG_HD bool terminated() const { return pdf[1] == 0 && pdf[2] == 0 && pdf[3] == 0; }
    PG_HD void terminate_secondary() {
        if (terminated()) return;
        pdf[1] = pdf[2] = pdf[3] = 0;
        pdf[0] /= 4;
    }
};

PG_HD Lambda sample_visible(float u) {  // src/core/spectrum.cpp:11-41
    Lambda w;
    for (int i = 0; i < 4; ++i) {
        float up = u + float(i) / 4;
        if (up > 1) up -= 1;
        float l = 538 - 138.888889f * atanhf(0.85691062f - 1.82750197f * up);
        l = clampf(l, 360.f, 830.f);
        w.l[i] = l;
        w.pdf[i] = (l < 360.f || l > 830.f) ? 0.f : 0.0039398042f / sq(coshf(0.0072f * (l - 538)));
    }
    return w;
}

PG_HD SS spec_sample(const SpecTable &t, int id, const Lambda &w) {
    return {spec_eval(t, id, w.l[0]), spec_eval(t, id, w.l[1]), spec_eval(t, id, w.l[2]), spec_eval(t, id, w.l[3])};
}

// ---- sampling warps (src/core/sampling.h) ------------------------------------------------------
struct V2 {
    float x, y;
};
PG_HD V2 disk_concentric(V2 u) {
    float ox = 2 * u.x - 1, oy = 2 * u.y - 1;
    if (ox == 0 && oy == 0) return {0, 0};
    float theta, r;
    if (fabsf(ox) > fabsf(oy)) {
        r = ox;
        theta = kPiOver4 * (oy / ox);
    } else {
        r = oy;
        theta = kPiOver2 - kPiOver4 * (ox / oy);
    }
    return {r * cosf(theta), r * sinf(theta)};
}
PG_HD V3 cosine_hemisphere(V2 u) {
    V2 d = disk_concentric(u);
    return {d.x, d.y, safe_sqrt(1 - d.x * d.x - d.y * d.y)};
}
PG_HD V3 uniform_sphere(V2 u) {
    float z = 1 - 2 * u.x;
    float r = safe_sqrt(1 - z * z);
    float phi = 2 * kPi * u.y;
    return {r * cosf(phi), r * sinf(phi), z};
}
PG_HD float power_heuristic(float f, float g) {
    if (is_inf(sq(f))) return 1;
    if (f == 0 && g == 0) return 0;
    return sq(f) / (sq(f) + sq(g));
}

// ---- device scene ------------------------------------------------------------------------------
struct DevScene {
    const prgpu::Primitive *prims;
    const prgpu::Material *mats;
    SpecTable spec;
    const prgpu::LightEx *lights;
    const float *lightArea;
    int lightCount;
    const LNode *lnodes;
    int lnodeCount;
    const int *leafOf;     // light -> leaf node, -1 when not in the tree
    const int *infinite;   // infinite (unbounded) lights
    int infCount;
    int envCount;          // environment lights (emission seen by escaping paths)
    int emitters;          // any area or environment light: the emission stage runs
    const float *powerCdf;  // power sampler: cdf[count + 1], func[count]
    const float *powerFunc;
    float powerInt;
    int powerCount;
    int lightSampler;
    const prgpu::MediumEx *media;
    const float *grid;
    int cameraMedium;
    float worldR;
    int hasInterfaces;
    const float *sheenTable;  // sheen directional albedo, 16 alphas x 32 cosines
    // camera and film
    float c2w[12];
    float tanHalf, aspect, lensR, focal;
    int width, height, margin;
    int filterType;
    float filterRadius, filterSigma, filterNorm;
    float xyzNorm;  // 1 / (4 * CIE_Y integral)
    TraceCtx trace;
};

PG_HD bool is_interface(const DevScene &s, const prgpu::Primitive &p) {
    bool noMat = p.material < 0;
    if (!noMat) {
        uint32_t t = s.mats[p.material].type;
        noMat = t == prgpu::MatInterface || t == prgpu::MatBlack;
    }
    return noMat && p.light < 0;
}
PG_HD bool has_bsdf(const DevScene &s, const prgpu::Primitive &p) {
    if (p.material < 0) return false;
    uint32_t t = s.mats[p.material].type;
    return t == prgpu::MatDiffuse || t == prgpu::MatConductor || t == prgpu::MatDielectric;
}

// ---- surface interactions (src/geometry/shape.cpp:63-85, 224-264) --------------------------------
struct Surf {
    V3 p, n, ns, dpdus, wo;
    V3 dpdv;  // the parameterisation's v derivative (bump mapping)
    float u, v;
};

PG_HD bool prim_has_normals(const prgpu::Primitive &p) {
    return !(p.n0.x == 0 && p.n0.y == 0 && p.n0.z == 0 && p.n1.x == 0 && p.n1.y == 0 && p.n1.z == 0 && p.n2.x == 0 &&
             p.n2.y == 0 && p.n2.z == 0);
}

// Both branches leave their results in locals and the record is written once at the end: stores
// of different fields from the two branches, merged at the join, would make the compiler keep the
// record in scratch memory.
PG_HD void interaction(const prgpu::Primitive &pr, V3 ro, V3 rd, float tHit, float b1, float b2, Surf *si) {
    V3 p, n, ns, dpdu, dpdv;
    float u, v;
    if (pr.type == prgpu::PrimSphere) {
        V3 c = f3(pr.p0);
        float r = pr.r;
        V3 pHit = ro + rd * tHit;
        V3 rel = pHit - c;
        float l = len(rel);
        V3 nn = l > 0 ? rel / l : V3(0, 1, 0);
        p = c + nn * r;
        float phi = atan2f(nn.z, nn.x);
        if (phi < 0) phi += 2 * kPi;
        float theta = safe_acos(nn.y);
        n = pr.flip ? -nn : nn;
        u = phi * kInv2Pi;
        v = theta * kInvPi;
        dpdu = V3(-rel.z, 0, rel.x) * (2 * kPi);
        float cosT = nn.y, sinT = safe_sqrt(1 - cosT * cosT);
        dpdv = sinT > 1e-6f ? V3(cosT * rel.x / sinT, -r * sinT, cosT * rel.z / sinT) * kPi : V3();
        if (len2(dpdu) < 1e-12f || len2(dpdv) < 1e-12f) coordinate_system(nn, &dpdu, &dpdv);
        ns = n;
    } else {
        V3 p0 = f3(pr.p0), p1 = f3(pr.p1), p2 = f3(pr.p2);
        float b0 = 1 - b1 - b2;
        p = p0 * b0 + p1 * b1 + p2 * b2;
        n = normalize(cross_f(p1 - p0, p2 - p0));
        u = pr.u0 * b0 + pr.u1 * b1 + pr.u2 * b2;
        v = pr.v0 * b0 + pr.v1 * b1 + pr.v2 * b2;
        float du02 = pr.u0 - pr.u2, dv02 = pr.v0 - pr.v2, du12 = pr.u1 - pr.u2, dv12 = pr.v1 - pr.v2;
        V3 dp02 = p0 - p2, dp12 = p1 - p2;
        float det = du02 * dv12 - dv02 * du12;
        bool degenerate = fabsf(det) < 1e-9f;
        if (!degenerate) {
            float inv = 1 / det;
            dpdu = (dp02 * dv12 - dp12 * dv02) * inv;
            dpdv = (dp12 * du02 - dp02 * du12) * inv;
            degenerate = len2(cross_d(dpdu, n)) == 0 || !is_finite(dpdu.x) || !is_finite(dpdv.x);
        }
        if (degenerate) coordinate_system(n, &dpdu, &dpdv);
        ns = n;
        if (prim_has_normals(pr)) {
            ns = f3(pr.n0) * b0 + f3(pr.n1) * b1 + f3(pr.n2) * b2;
            float l = len(ns);
            ns = l > 0 ? ns / l : n;
            n = face_forward(n, ns);
        }
    }
    si->p = p;
    si->n = n;
    si->ns = ns;
    si->u = u;
    si->v = v;
    si->dpdus = dpdu;
    si->dpdv = dpdv;
    si->wo = -normalize(rd);
}

PG_HD float ray_eps(V3 p) { return 5e-5f * maxf(1.f, max_comp(vabs(p))); }  // src/scene/interaction.h:20
PG_HD V3 offset_origin(V3 p, V3 n, V3 w) {  // interaction.h:33-39
    if (n.zero()) return p;
    float eps = ray_eps(p);
    V3 off = n * eps;
    if (dot(w, n) < 0) off = -off;
    return p + off;
}

// ---- BSDFs (src/materials/bxdf.cpp:13-256) ------------------------------------------------------
constexpr uint32_t kRefl = 1, kTrans = 2, kDiffuse = 4, kGlossy = 8, kSpecular = 16;
enum BxType : int { BxNone = 0, BxDiffuse = 1, BxConductor = 2, BxDielectric = 3, BxThin = 4 };

struct TrDist {
    float ax, ay;
    PG_HD void init(float x, float y) {
        ax = x;
        ay = y;
        if (!smooth()) {
            ax = maxf(ax, 1e-4f);
            ay = maxf(ay, 1e-4f);
        }
    }
    PG_HD bool smooth() const { return maxf(ax, ay) < 1e-3f; }
    PG_HD float D(V3 wm) const {
        float t2 = tan2t(wm);
        if (is_inf(t2)) return 0;
        float c4 = sq(cos2t(wm));
        if (c4 < 1e-16f) return 0;
        float e = t2 * (sq(cosphi(wm) / ax) + sq(sinphi(wm) / ay));
        return 1 / (kPi * ax * ay * c4 * sq(1 + e));
    }
    PG_HD float lambda(V3 w) const {
        float t2 = tan2t(w);
        if (is_inf(t2)) return 0;
        float a2 = sq(cosphi(w) * ax) + sq(sinphi(w) * ay);
        return (sqrtf(1 + a2 * t2) - 1) / 2;
    }
    PG_HD float G1(V3 w) const { return 1 / (1 + lambda(w)); }
    PG_HD float G(V3 wo, V3 wi) const { return 1 / (1 + lambda(wo) + lambda(wi)); }
    PG_HD float pdf(V3 w, V3 wm) const { return G1(w) / abscos(w) * D(wm) * absdot(w, wm); }
    PG_HD V3 sample_wm(V3 w, V2 u) const {
        V3 wh = normalize(V3(ax * w.x, ay * w.y, w.z));
        if (wh.z < 0) wh = -wh;
        V3 T1 = (wh.z < 0.99999f) ? normalize(cross_d(V3(0, 0, 1), wh)) : V3(1, 0, 0);
        V3 T2 = cross_d(wh, T1);
        float r = sqrtf(u.x);
        float theta = 2 * kPi * u.y;
        float px = r * cosf(theta), py = r * sinf(theta);
        float h = sqrtf(1 - sq(px));
        py = lerp((1 + wh.z) / 2, h, py);
        float pz = sqrtf(maxf(0.f, 1 - px * px - py * py));
        V3 nh = T1 * px + T2 * py + wh * pz;
        return normalize(V3(ax * nh.x, ay * nh.y, maxf(1e-6f, nh.z)));
    }
};

PG_HD float fr_dielectric(float ci, float eta) {
    ci = clampf(ci, -1.f, 1.f);
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

// Complex arithmetic for the conductor Fresnel term (bxdf.cpp:47-56).
struct Cx {
    float re, im;
};
PG_HD Cx cmul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
PG_HD Cx cadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
PG_HD Cx csub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
PG_HD Cx cdiv(Cx a, Cx b) {  // Smith's algorithm
    if (fabsf(b.re) >= fabsf(b.im)) {
        float r = b.im / b.re, den = b.re + b.im * r;
        return {(a.re + a.im * r) / den, (a.im - a.re * r) / den};
    }
    float r = b.re / b.im, den = b.re * r + b.im;
    return {(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}
PG_HD Cx csqrt(Cx z) {  // principal square root
    if (z.re == 0 && z.im == 0) return {0, z.im};
    float m = hypotf(z.re, z.im);
    if (z.re >= 0) {
        float t = sqrtf(0.5f * (m + z.re));
        return {t, z.im / (2 * t)};
    }
    float t = sqrtf(0.5f * (m - z.re));
    return {fabsf(z.im) / (2 * t), copysignf(t, z.im)};
}
PG_HD float fr_complex1(float ci, float eta, float k) {
    ci = clampf(ci, 0.f, 1.f);
    float s2i = 1 - sq(ci);
    Cx e{eta, k};
    Cx e2 = cmul(e, e);
    Cx s2t = cdiv(Cx{s2i, 0}, e2);
    Cx ct = csqrt(csub(Cx{1, 0}, s2t));
    Cx eci{e.re * ci, e.im * ci};
    Cx rpa = cdiv(csub(eci, ct), cadd(eci, ct));
    Cx ect = cmul(e, ct);
    Cx rpe = cdiv(csub(Cx{ci, 0}, ect), cadd(Cx{ci, 0}, ect));
    float n1 = rpa.re * rpa.re + rpa.im * rpa.im, n2 = rpe.re * rpe.re + rpe.im * rpe.im;
    return (n1 + n2) / 2;
}
PG_HD SS fr_complex(float ci, const SS &eta, const SS &k) {
    return {fr_complex1(ci, eta.v[0], k.v[0]), fr_complex1(ci, eta.v[1], k.v[1]), fr_complex1(ci, eta.v[2], k.v[2]),
            fr_complex1(ci, eta.v[3], k.v[3])};
}

PG_HD float hg_phase(float cosTheta, float g) {
    float denom = 1 + sq(g) + 2 * g * cosTheta;
    return kInv4Pi * (1 - sq(g)) / (denom * safe_sqrt(denom));
}
PG_HD V3 sample_hg(V3 wo, float g, V2 u, float *pdf) {  // bxdf.cpp:69-80
    float cosTheta;
    if (fabsf(g) < 1e-3f) cosTheta = 1 - 2 * u.x;
    else cosTheta = -1 / (2 * g) * (1 + sq(g) - sq((1 - sq(g)) / (1 + g - 2 * g * u.x)));
    cosTheta = clampf(cosTheta, -1.f, 1.f);
    float sinTheta = safe_sqrt(1 - sq(cosTheta));
    float phi = 2 * kPi * u.y;
    Frame f = Frame::from_z(wo);
    V3 wi = f.from_local(spherical_dir(sinTheta, cosTheta, phi));
    *pdf = hg_phase(cosTheta, g);
    return wi;
}

// Sample flags (reflection / transmission) and transport mode (src/materials/bxdf.h:35-40). The
// mode only changes dielectric transmission (the 1/eta'^2 radiance scaling); the layered walk
// samples its exit interface in the adjoint (importance) mode.
constexpr uint32_t kRtRefl = 1, kRtTrans = 2, kRtAll = 3;

struct BSample {
    SS f;
    V3 wi;
    float pdf;
    uint32_t flags;
    float eta;
    bool proportional = false;  // pdf only proportional to the sampling density (layered, sheen)
};

struct BxDF {
    int type;
    SS a, b;  // diffuse: a = R; conductor: a = eta, b = k
    float etaD;
    TrDist dist;

    PG_HD uint32_t flags() const {
        switch (type) {
        case BxDiffuse: return a.nonzero() ? (kDiffuse | kRefl) : 0u;
        case BxConductor: return dist.smooth() ? (kSpecular | kRefl) : (kGlossy | kRefl);
        case BxDielectric: return ((etaD == 1) ? kTrans : (kRefl | kTrans)) | (dist.smooth() ? kSpecular : kGlossy);
        case BxThin: return kRefl | kTrans | kSpecular;
        default: return 0u;
        }
    }
    PG_HD SS f(V3 wo, V3 wi, bool radiance = true) const {
        if (type == BxDiffuse) {
            if (!same_hemi(wo, wi)) return SS(0.f);
            return a * kInvPi;
        }
        if (type == BxConductor) {
            if (!same_hemi(wo, wi) || dist.smooth()) return SS(0.f);
            float co = abscos(wo), ci = abscos(wi);
            if (ci == 0 || co == 0) return SS(0.f);
            V3 wm = wi + wo;
            if (len2(wm) == 0) return SS(0.f);
            wm = normalize(wm);
            SS F = fr_complex(absdot(wo, wm), a, b);
            return F * (dist.D(wm) * dist.G(wo, wi) / (4 * ci * co));
        }
        if (type == BxDielectric) {
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
            if (refl) return SS(dist.D(wm) * dist.G(wo, wi) * F / fabsf(4 * ci * co));
            float denom = sq(dot(wi, wm) + dot(wo, wm) / etap) * ci * co;
            float ft = dist.D(wm) * (1 - F) * dist.G(wo, wi) * fabsf(dot(wi, wm) * dot(wo, wm) / denom);
            if (radiance) ft /= sq(etap);
            return SS(ft);
        }
        return SS(0.f);
    }
    PG_HD float pdf(V3 wo, V3 wi, uint32_t rt = kRtAll) const {
        if (type == BxDiffuse) {
            if (!(rt & kRtRefl) || !same_hemi(wo, wi)) return 0;
            return abscos(wi) * kInvPi;
        }
        if (type == BxConductor) {
            if (!(rt & kRtRefl) || !same_hemi(wo, wi) || dist.smooth()) return 0;
            V3 wm = wo + wi;
            if (len2(wm) == 0) return 0;
            wm = face_forward(normalize(wm), V3(0, 0, 1));
            return dist.pdf(wo, wm) / (4 * absdot(wo, wm));
        }
        if (type == BxDielectric) {
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
        return 0;
    }
    PG_HD bool sample(V3 wo, float uc, V2 u, BSample *s, bool radiance = true, uint32_t rt = kRtAll) const {
        if (type == BxDiffuse) {
            if (!(rt & kRtRefl)) return false;
            V3 wi = cosine_hemisphere(u);
            if (wo.z < 0) wi.z *= -1;
            *s = {a * kInvPi, wi, abscos(wi) * kInvPi, kDiffuse | kRefl, 1};
            return true;
        }
        if (type == BxConductor) {
            if (!(rt & kRtRefl)) return false;
            if (dist.smooth()) {
                V3 wi(-wo.x, -wo.y, wo.z);
                *s = {fr_complex(abscos(wi), a, b) / abscos(wi), wi, 1, kSpecular | kRefl, 1};
                return true;
            }
            if (wo.z == 0) return false;
            V3 wm = dist.sample_wm(wo, u);
            V3 wi = reflect(wo, wm);
            if (!same_hemi(wo, wi)) return false;
            float p = dist.pdf(wo, wm) / (4 * absdot(wo, wm));
            float co = abscos(wo), ci = abscos(wi);
            if (ci == 0 || co == 0) return false;
            SS F = fr_complex(absdot(wo, wm), a, b);
            *s = {F * (dist.D(wm) * dist.G(wo, wi) / (4 * ci * co)), wi, p, kGlossy | kRefl, 1};
            return true;
        }
        if (type == BxDielectric) {
            if (etaD == 1 || dist.smooth()) {
                float R = fr_dielectric(wo.z, etaD), T = 1 - R;
                float pr = (rt & kRtRefl) ? R : 0.f, pt = (rt & kRtTrans) ? T : 0.f;
                if (pr == 0 && pt == 0) return false;
                if (uc < pr / (pr + pt)) {
                    V3 wi(-wo.x, -wo.y, wo.z);
                    *s = {SS(R / abscos(wi)), wi, pr / (pr + pt), kSpecular | kRefl, 1};
                    return true;
                }
                V3 wi;
                float etap;
                if (!refract(wo, V3(0, 0, 1), etaD, &etap, &wi)) return false;
                SS ft(T / abscos(wi));
                *s = {radiance ? ft / sq(etap) : ft, wi, pt / (pr + pt), kSpecular | kTrans, etap};
                return true;
            }
            V3 wm = dist.sample_wm(wo, u);
            float R = fr_dielectric(dot(wo, wm), etaD), T = 1 - R;
            float pr = (rt & kRtRefl) ? R : 0.f, pt = (rt & kRtTrans) ? T : 0.f;
            if (pr == 0 && pt == 0) return false;
            if (uc < pr / (pr + pt)) {
                V3 wi = reflect(wo, wm);
                if (!same_hemi(wo, wi)) return false;
                float p = dist.pdf(wo, wm) / (4 * absdot(wo, wm)) * pr / (pr + pt);
                *s = {SS(dist.D(wm) * dist.G(wo, wi) * R / (4 * wi.z * wo.z)), wi, p, kGlossy | kRefl, 1};
                return true;
            }
            float etap;
            V3 wi;
            bool tir = !refract(wo, wm, etaD, &etap, &wi);
            if (tir || same_hemi(wo, wi) || wi.z == 0) return false;
            float denom = sq(dot(wi, wm) + dot(wo, wm) / etap);
            float dwm_dwi = absdot(wi, wm) / denom;
            float p = dist.pdf(wo, wm) * dwm_dwi * pt / (pr + pt);
            SS ft(T * dist.D(wm) * dist.G(wo, wi) * fabsf(dot(wi, wm) * dot(wo, wm) / (wi.z * wo.z * denom)));
            *s = {radiance ? ft / sq(etap) : ft, wi, p, kGlossy | kTrans, etap};
            return true;
        }
        if (type == BxThin) {
            float R = fr_dielectric(abscos(wo), etaD), T = 1 - R;
            if (R < 1) {
                R += sq(T) * R / (1 - sq(R));
                T = 1 - R;
            }
            float pr = (rt & kRtRefl) ? R : 0.f, pt = (rt & kRtTrans) ? T : 0.f;
            if (pr == 0 && pt == 0) return false;
            if (uc < pr / (pr + pt)) {
                V3 wi(-wo.x, -wo.y, wo.z);
                *s = {SS(R / abscos(wi)), wi, pr / (pr + pt), kSpecular | kRefl, 1};
                return true;
            }
            V3 wi = -wo;
            *s = {SS(T / abscos(wi)), wi, pt / (pr + pt), kSpecular | kTrans, 1};
            return true;
        }
        return false;
    }
};

struct BSDF {  // src/materials/bsdf.h:8-51
    Frame frame;
    BxDF bx;
    PG_HD SS f(V3 woW, V3 wiW) const {
        V3 wi = frame.to_local(wiW), wo = frame.to_local(woW);
        if (wo.z == 0) return SS(0.f);
        return bx.f(wo, wi);
    }
    PG_HD float pdf(V3 woW, V3 wiW) const {
        V3 wo = frame.to_local(woW), wi = frame.to_local(wiW);
        if (wo.z == 0) return 0;
        return bx.pdf(wo, wi);
    }
    PG_HD bool sample(V3 woW, float u, V2 u2, BSample *bs) const {
        V3 wo = frame.to_local(woW);
        if (wo.z == 0 || !(bx.flags() & (kRefl | kTrans))) return false;
        if (!bx.sample(wo, u, u2, bs)) return false;
        if (!bs->f.nonzero() || bs->pdf == 0 || bs->wi.z == 0 || bs->f.has_nan() || !is_finite(bs->pdf)) return false;
        bs->wi = frame.from_local(bs->wi);
        return true;
    }
    PG_HD bool non_specular() const { return (bx.flags() & (kDiffuse | kGlossy)) != 0; }
};

// ---- coated and sheen closures -------------------------------------------------------------------
// FP32 results of exp and log rounded from double precision: these agree with the CPU's
// correctly rounded expf / logf where the single-precision device routines can differ in the
// last place, and the layered walk's random stream is hashed from the directions it produces.
PG_HD float exp_f32(float x) { return float(exp(double(x))); }
PG_HD float log_f32(float x) { return float(log(double(x))); }

PG_HD bool usable(const BSample &s) { return s.f.nonzero() && s.pdf != 0 && s.wi.z != 0; }

// Stochastic layered closure (reference src/materials/bxdf.cpp:260-517, pbrt-v4's position-free
// random walk): a dielectric coat over a diffuse or conductor base, with a slab of unit
// extinction and albedo `albedo` between them. Its random stream is hashed from the arguments
// (f: (0, wo), wi; pdf: (0, wi), wo; sample: (0, wo), (uc, u)), so f, pdf and sample are
// deterministic; every 2D draw takes its second coordinate first, the order the reference's
// compiled Vec2f(r(), r()) evaluates in.
struct Layered {
    BxDF top, bottom;
    SS albedo;
    float thick, g;
    int maxDepth, nSamples;

    struct Walk {
        Pcg rng;
        PG_HD float u() { return minf(rng.uniform(), kOneMinusEps); }
        PG_HD V2 u2() {
            float y = u();
            float x = u();
            return {x, y};
        }
    };
    PG_HD static Walk walk(V3 a, uint64_t offset) {  // RNG(Hash(0ull, a), offset)
        const uint64_t w[2] = {0ull, fbits(a.x, a.y)};
        Walk r;
        r.rng.set_sequence(murmur_words(w, 2, true, __float_as_uint(a.z)), offset);
        return r;
    }
    PG_HD static uint64_t hash_dir(V3 a) {
        const uint64_t w[1] = {fbits(a.x, a.y)};
        return murmur_words(w, 1, true, __float_as_uint(a.z));
    }
    PG_HD static float tr(float dz, V3 w) {
        if (fabsf(dz) <= 1.17549435e-38f) return 1;
        return exp_f32(-fabsf(dz / w.z));
    }
    PG_HD static float free_flight(float u, V3 w) { return -log_f32(1 - u) / (1.f / fabsf(w.z)); }

    PG_HD uint32_t flags() const {
        uint32_t t = top.flags(), b = bottom.flags(), f = kRefl;
        if (t & kSpecular) f |= kSpecular;
        if ((t & kDiffuse) || (b & kDiffuse) || albedo.nonzero()) f |= kDiffuse;
        else if ((t & kGlossy) || (b & kGlossy)) f |= kGlossy;
        if ((t & kTrans) && (b & kTrans)) f |= kTrans;
        return f;
    }

    PG_HD SS f(V3 wo, V3 wi) const {  // radiance transport
        if (wo.z < 0) {  // two-sided
            wo = -wo;
            wi = -wi;
        }
        const bool same = same_hemi(wo, wi);
        // Entered through the coat; the walk leaves through the interface on wi's side.
        const BxDF &exitI = same ? top : bottom, &inner = same ? bottom : top;
        const float exitZ = same ? thick : 0.f;
        const bool exitSpec = (exitI.flags() & kSpecular) != 0, innerSpec = (inner.flags() & kSpecular) != 0;
        SS f(0.f);
        if (same) f = top.f(wo, wi) * float(nSamples);
        Walk w = walk(wo, hash_dir(wi));
        for (int s = 0; s < nSamples; ++s) {
            BSample wos, wis;
            float uc = w.u();
            V2 u0 = w.u2();
            if (!top.sample(wo, uc, u0, &wos, true, kRtTrans) || !usable(wos)) continue;
            uc = w.u();
            V2 u1 = w.u2();
            if (!exitI.sample(wi, uc, u1, &wis, false, kRtTrans) || !usable(wis)) continue;
            SS beta = wos.f * abscos(wos.wi) / wos.pdf;
            float z = thick;
            V3 d = wos.wi;
            for (int depth = 0; depth < maxDepth; ++depth) {
                if (depth > 3 && beta.maxc() < 0.25f) {
                    float q = maxf(0.f, 1 - beta.maxc());
                    if (w.u() < q) break;
                    beta = beta / (1 - q);
                }
                if (!albedo.nonzero()) {
                    z = (z == thick) ? 0 : thick;
                    beta = beta * tr(thick, d);
                } else {
                    const float dz = free_flight(w.u(), d);
                    const float zp = d.z > 0 ? (z + dz) : (z - dz);
                    if (0 < zp && zp < thick) {
                        float wt = 1;
                        float phaseExit = hg_phase(dot(-d, -wis.wi), g);
                        if (!exitSpec) wt = power_heuristic(wis.pdf, phaseExit);
                        f = f + beta * albedo * phaseExit * wt * tr(zp - exitZ, wis.wi) * wis.f / wis.pdf;
                        float phasePdf;
                        V3 wp = sample_hg(-d, g, w.u2(), &phasePdf);
                        if (phasePdf == 0 || wp.z == 0) continue;
                        beta = beta * albedo;
                        d = wp;
                        z = zp;
                        if (((z < exitZ && d.z > 0) || (z > exitZ && d.z < 0)) && !exitSpec) {
                            SS fExit = exitI.f(-d, wi);
                            if (fExit.nonzero()) {
                                float wt2 = power_heuristic(phasePdf, exitI.pdf(-d, wi, kRtTrans));
                                f = f + beta * tr(zp - exitZ, wp) * fExit * wt2;
                            }
                        }
                        continue;
                    }
                    z = clampf(zp, 0.f, thick);
                }
                BSample bs;
                if (z == exitZ) {
                    float uc2 = w.u();
                    V2 u2 = w.u2();
                    if (!exitI.sample(-d, uc2, u2, &bs, true, kRtRefl) || !usable(bs)) break;
                    beta = beta * (bs.f * abscos(bs.wi) / bs.pdf);
                    d = bs.wi;
                } else {
                    if (!innerSpec) {
                        float wt = 1;
                        if (!exitSpec) wt = power_heuristic(wis.pdf, inner.pdf(-d, -wis.wi));
                        f = f + beta * inner.f(-d, -wis.wi) * abscos(wis.wi) * wt * tr(thick, wis.wi) * wis.f / wis.pdf;
                    }
                    float uc2 = w.u();
                    V2 u2 = w.u2();
                    if (!inner.sample(-d, uc2, u2, &bs, true, kRtRefl) || !usable(bs)) break;
                    beta = beta * (bs.f * abscos(bs.wi) / bs.pdf);
                    d = bs.wi;
                    if (!exitSpec) {
                        SS fExit = exitI.f(-d, wi);
                        if (fExit.nonzero()) {
                            float wt = 1;
                            if (!innerSpec) wt = power_heuristic(bs.pdf, exitI.pdf(-d, wi, kRtTrans));
                            f = f + beta * tr(thick, bs.wi) * fExit * wt;
                        }
                    }
                }
            }
        }
        return f / float(nSamples);
    }

    PG_HD bool sample(V3 wo, float uc, V2 u, BSample *out) const {
        const bool flip = wo.z < 0;
        if (flip) wo = -wo;
        BSample bs;
        if (!top.sample(wo, uc, u, &bs) || !usable(bs)) return false;
        if (bs.flags & kRefl) {
            if (flip) bs.wi = -bs.wi;
            bs.proportional = true;
            *out = bs;
            return true;
        }
        V3 d = bs.wi;
        bool specPath = (bs.flags & kSpecular) != 0;
        const uint64_t uw[1] = {fbits(uc, u.x)};
        Walk w = walk(wo, murmur_words(uw, 1, true, __float_as_uint(u.y)));
        SS f = bs.f * abscos(bs.wi);
        float pdf = bs.pdf;
        float z = thick;
        for (int depth = 0; depth < maxDepth; ++depth) {
            float rrBeta = f.maxc() / pdf;
            if (depth > 3 && rrBeta < 0.25f) {
                float q = maxf(0.f, 1 - rrBeta);
                if (w.u() < q) return false;
                pdf *= 1 - q;
            }
            if (d.z == 0) return false;
            if (albedo.nonzero()) {
                const float dz = free_flight(w.u(), d);
                const float zp = d.z > 0 ? (z + dz) : (z - dz);
                if (zp == z) return false;
                if (0 < zp && zp < thick) {
                    float phasePdf;
                    V3 wp = sample_hg(-d, g, w.u2(), &phasePdf);
                    if (phasePdf == 0 || wp.z == 0) return false;
                    f = f * (albedo * phasePdf);
                    pdf *= phasePdf;
                    specPath = false;
                    d = wp;
                    z = zp;
                    continue;
                }
                z = clampf(zp, 0.f, thick);
            } else {
                z = (z == thick) ? 0 : thick;
                f = f * tr(thick, d);
            }
            const BxDF &iface = (z == 0) ? bottom : top;
            float uc2 = w.u();
            V2 u2 = w.u2();
            BSample b2;
            if (!iface.sample(-d, uc2, u2, &b2) || !usable(b2)) return false;
            f = f * b2.f;
            pdf *= b2.pdf;
            specPath = specPath && (b2.flags & kSpecular);
            d = b2.wi;
            if (b2.flags & kTrans) {
                uint32_t fl = (same_hemi(wo, d) ? kRefl : kTrans) | (specPath ? kSpecular : kGlossy);
                if (flip) d = -d;
                *out = {f, d, pdf, fl, 1.f, true};
                return true;
            }
            f = f * abscos(b2.wi);
        }
        return false;
    }

    PG_HD float pdf(V3 wo, V3 wi) const {
        if (wo.z < 0) {
            wo = -wo;
            wi = -wi;
        }
        Walk w = walk(wi, hash_dir(wo));
        const bool same = same_hemi(wo, wi);
        float sum = 0;
        if (same) sum += nSamples * top.pdf(wo, wi, kRtRefl);
        for (int s = 0; s < nSamples; ++s) {
            BSample wos, wis;
            if (same) {
                // Transmission through the coat, one bounce off the base.
                float a = w.u();
                V2 ua = w.u2();
                const bool okO = top.sample(wo, a, ua, &wos, true, kRtTrans);
                float b = w.u();
                V2 ub = w.u2();
                const bool okI = top.sample(wi, b, ub, &wis, false, kRtTrans);
                if (okO && wos.f.nonzero() && wos.pdf > 0 && okI && wis.f.nonzero() && wis.pdf > 0) {
                    if (!(top.flags() & (kDiffuse | kGlossy))) {
                        sum += bottom.pdf(-wos.wi, -wis.wi);
                    } else {
                        float c = w.u();
                        V2 uc2 = w.u2();
                        BSample rs;
                        if (bottom.sample(-wos.wi, c, uc2, &rs) && rs.f.nonzero() && rs.pdf > 0) {
                            if (!(bottom.flags() & (kDiffuse | kGlossy))) {
                                sum += top.pdf(-rs.wi, wi);
                            } else {
                                float rPdf = bottom.pdf(-wos.wi, -wis.wi);
                                sum += power_heuristic(wis.pdf, rPdf) * rPdf;
                                float tPdf = top.pdf(-rs.wi, wi);
                                sum += power_heuristic(rs.pdf, tPdf) * tPdf;
                            }
                        }
                    }
                }
            } else {
                // Transmission through both interfaces.
                float a = w.u();
                V2 ua = w.u2();
                if (!top.sample(wo, a, ua, &wos) || !usable(wos) || (wos.flags & kRefl)) continue;
                float b = w.u();
                V2 ub = w.u2();
                if (!bottom.sample(wi, b, ub, &wis, false) || !usable(wis) || (wis.flags & kRefl)) continue;
                if (top.flags() & kSpecular) sum += bottom.pdf(-wos.wi, wi);
                else if (bottom.flags() & kSpecular) sum += top.pdf(wo, -wis.wi);
                else sum += (top.pdf(wo, -wis.wi) + bottom.pdf(-wos.wi, wi)) / 2;
            }
        }
        return lerp(0.9f, 1 / (4 * kPi), sum / nSamples);
    }
};

// "Charlie" sheen (reference src/materials/bxdf.cpp:519-631) and its directional albedo table
// (16 alphas x 32 cosines, bilinear).
constexpr int kSheenMuN = 32, kSheenAlphaN = 16;
PG_HD float sheen_value(V3 wo, V3 wi, float alpha) {
    if (!same_hemi(wo, wi)) return 0;
    float co = abscos(wo), ci = abscos(wi);
    if (co == 0 || ci == 0) return 0;
    V3 wh = normalize(wo + wi);
    float c = abscos(wh), inv = 1 / alpha;
    float sin2 = maxf(0.f, 1 - c * c);
    float D = (2 + inv) * float(pow(double(sin2), double(0.5f * inv))) * kInv2Pi;
    return D / (4 * (ci + co - ci * co));
}
PG_HD float sheen_albedo(const float *t, float cosTheta, float alpha) {
    float fa = (clampf(alpha, 0.05f, 1.f) - 0.05f) / 0.95f * (kSheenAlphaN - 1);
    float fm = clampf(fabsf(cosTheta), 0.f, 1.f) * (kSheenMuN - 1);
    int ia = int(fa), im = int(fm);
    ia = ia < kSheenAlphaN - 2 ? ia : kSheenAlphaN - 2;
    im = im < kSheenMuN - 2 ? im : kSheenMuN - 2;
    float ta = fa - ia, tm = fm - im;
    const float *r0 = t + ia * kSheenMuN + im, *r1 = r0 + kSheenMuN;
    return lerp(ta, lerp(tm, r0[0], r0[1]), lerp(tm, r1[0], r1[1]));
}

struct SheenLobe {
    SS color;
    float alpha, weight, colorMax;
    const float *table;
    PG_HD SS f(V3 wo, V3 wi) const { return color * (weight * sheen_value(wo, wi, alpha)); }
    PG_HD float base_scale(V3 wo, V3 wi) const {
        float e = maxf(sheen_albedo(table, wo.z, alpha), sheen_albedo(table, wi.z, alpha));
        return maxf(0.f, 1 - weight * colorMax * e);
    }
    PG_HD float prob(V3 wo, uint32_t baseFlags) const {
        float e = weight * colorMax * sheen_albedo(table, wo.z, alpha);
        if (!baseFlags) return 1;
        return clampf(e, 0.f, 0.9f);
    }
};

// The closure of the coated / sheen family (src/materials/material.cpp:5-14): a plain or
// layered base, optionally under a sheen lobe, in the shading frame. Radiance transport and all
// sample flags, as the path tracer calls it.
struct LBSDF {
    Frame frame;
    BxDF bx;
    Layered lay;
    SheenLobe sh;
    bool layered, sheen;

    PG_HD uint32_t base_flags() const { return layered ? lay.flags() : bx.flags(); }
    PG_HD uint32_t flags() const { return sheen ? (base_flags() | kGlossy | kRefl) : base_flags(); }
    PG_HD SS base_f(V3 wo, V3 wi) const { return layered ? lay.f(wo, wi) : bx.f(wo, wi); }
    PG_HD float base_pdf(V3 wo, V3 wi) const { return layered ? lay.pdf(wo, wi) : bx.pdf(wo, wi); }
    PG_HD bool base_sample(V3 wo, float uc, V2 u, BSample *s) const {
        return layered ? lay.sample(wo, uc, u, s) : bx.sample(wo, uc, u, s);
    }
    PG_HD SS local_f(V3 wo, V3 wi) const {
        if (!sheen) return base_f(wo, wi);
        return sh.f(wo, wi) + base_f(wo, wi) * sh.base_scale(wo, wi);
    }
    PG_HD float local_pdf(V3 wo, V3 wi) const {
        if (!sheen) return base_pdf(wo, wi);
        float ps = sh.prob(wo, base_flags());
        float pSheen = same_hemi(wo, wi) ? abscos(wi) * kInvPi : 0.f;
        return ps * pSheen + (1 - ps) * base_pdf(wo, wi);
    }
    PG_HD bool local_sample(V3 wo, float uc, V2 u, BSample *bs) const {
        if (!sheen) return base_sample(wo, uc, u, bs);
        float ps = sh.prob(wo, base_flags());
        const bool exact = !layered;
        if (uc < ps) {
            V3 wi = cosine_hemisphere(u);
            if (wo.z < 0) wi.z = -wi.z;
            if (exact) {
                float p = local_pdf(wo, wi);
                if (p == 0) return false;
                *bs = {local_f(wo, wi), wi, p, kGlossy | kRefl, 1};
                return true;
            }
            *bs = {sh.f(wo, wi), wi, ps * (abscos(wi) * kInvPi), kGlossy | kRefl, 1, true};
            return true;
        }
        float uc2 = ps < 1 ? minf((uc - ps) / (1 - ps), kOneMinusEps) : 0.f;
        if (!base_sample(wo, uc2, u, bs) || bs->pdf == 0) return false;
        if ((bs->flags & kSpecular) || !exact || bs->proportional) {
            bs->f = bs->f * sh.base_scale(wo, bs->wi);
            bs->pdf *= (1 - ps);
            if (!(bs->flags & kSpecular)) bs->proportional = true;
            return true;
        }
        float p = local_pdf(wo, bs->wi);
        if (p == 0) return false;
        bs->f = local_f(wo, bs->wi);
        bs->pdf = p;
        return true;
    }

    PG_HD SS f(V3 woW, V3 wiW) const {
        V3 wi = frame.to_local(wiW), wo = frame.to_local(woW);
        if (wo.z == 0) return SS(0.f);
        return local_f(wo, wi);
    }
    PG_HD float pdf(V3 woW, V3 wiW) const {
        V3 wo = frame.to_local(woW), wi = frame.to_local(wiW);
        if (wo.z == 0) return 0;
        return local_pdf(wo, wi);
    }
    PG_HD bool sample(V3 woW, float u, V2 u2, BSample *bs) const {
        V3 wo = frame.to_local(woW);
        if (wo.z == 0 || !(flags() & (kRefl | kTrans))) return false;
        if (!local_sample(wo, u, u2, bs)) return false;
        if (!bs->f.nonzero() || bs->pdf == 0 || bs->wi.z == 0 || bs->f.has_nan() || !is_finite(bs->pdf)) return false;
        bs->wi = frame.from_local(bs->wi);
        return true;
    }
    PG_HD bool non_specular() const { return (flags() & (kDiffuse | kGlossy)) != 0; }
};

// A material's albedo spectrum at the hit (src/materials/material.cpp:64-117): a spectrum, or an
// image texture (ImageSpectrumTexture as an albedo, texture.cpp:38-52: the multiplied bilinear
// lookup, clamped to [0, 1], uplifted through the RGB-to-spectrum table, RGBAlbedoSample).
PG_HD SS albedo_sample(const DevScene &s, int id, const Surf &si, const Lambda &lam) {
    const prgpu::SpectrumEx &sp = s.spec.s[id];
    if (sp.kind != prgpu::SpecImage) return spec_sample(s.spec, id, lam);
    const float *rec = s.spec.knots + sp.offset;
    float rgb[3], c[3];
    tex_rgb(rec, si.u, si.v, rgb);
    for (int i = 0; i < 3; ++i) rgb[i] = rgb[i] * rec[8 + i];
    rgb_sigmoid_coeffs(s.spec.knots + int(rec[11]), rgb, c);
    SS out;
    for (int i = 0; i < 4; ++i) out.v[i] = sigmoid_poly(c, lam.l[i]);
    return out;
}

// Material::PerturbShading (src/materials/material.cpp:16-58): a bump map's finite-difference
// height gradient in uv tilts the shading normal, then a normal map's tangent-space normal
// replaces it in the shading frame; the result is kept in the geometric normal's hemisphere and
// the shading tangent re-orthogonalised. The record is at knots[m.shading - 1] (kShadeHeader).
PG_HD void perturb_shading(const DevScene &s, const prgpu::Material &m, Surf *si) {
    if (m.shading == 0) return;
    const float *knots = s.spec.knots;
    const float *r = knots + (m.shading - 1);
    V3 ns = si->ns;
    bool changed = false;
    if (r[4] != 0) {
        const float *hrec = knots + int(r[5]);
        const int ch = int(r[6]);
        const float du = 1.f / 2048, dv = 1.f / 2048;
        const float h0 = tex_float(hrec, ch, si->u, si->v);
        const float dhdu = (tex_float(hrec, ch, si->u + du, si->v) - h0) / du * r[7];
        const float dhdv = (tex_float(hrec, ch, si->u, si->v + dv) - h0) / dv * r[7];
        const V3 dpdu = si->dpdus + ns * dhdu, dpdv = si->dpdv + ns * dhdv;
        V3 n = cross_d(dpdu, dpdv);
        const float l = len(n);
        if (l > 0 && is_finite(l)) {
            n = n / l;
            if (dot(n, ns) < 0) n = -n;
            ns = n;
            si->dpdus = dpdu;
            changed = true;
        }
    }
    if (r[0] != 0) {
        float c[3];
        tex_rgb(knots + int(r[1]), si->u, si->v, c);
        V3 t(2 * c[0] - 1, 2 * c[1] - 1, 2 * c[2] - 1);
        t.x *= r[2];
        t.y *= r[3] != 0 ? -r[2] : r[2];
        Frame f = Frame::from_xz(si->dpdus, ns);
        if (dot(f.y, si->dpdv) < 0) f.y = -f.y;  // the bitangent follows the parameterisation's v
        const V3 n = f.x * t.x + f.y * t.y + f.z * maxf(t.z, 1e-3f);
        const float l = len(n);
        if (l > 0 && is_finite(l)) {
            ns = n / l;
            changed = true;
        }
    }
    if (!changed) return;
    if (dot(ns, si->n) < 0) ns = -ns;
    si->ns = ns;
    si->dpdus = si->dpdus - ns * dot(si->dpdus, ns);
}

// src/materials/material.cpp:64-117 (textures: constant spectra and uv checkers).
PG_HD int checker_pick(const prgpu::Material &m, float u, float v, int a) {
    if (m.albedoB >= 0 && m.checkerScale != 0) {
        float sx = u * m.checkerScale + 0.f, sy = v * m.checkerScale + 0.f;
        bool odd = (int(floorf(sx)) + int(floorf(sy))) & 1;
        if (odd) return m.albedoB;
    }
    return a;
}

PG_HD void diffuse_bx(const DevScene &s, const prgpu::Material &m, const Surf &si, const Lambda &lam, BxDF &bx) {
    SS r = albedo_sample(s, checker_pick(m, si.u, si.v, m.albedo), si, lam);
    for (int i = 0; i < 4; ++i) r.v[i] = clampf(r.v[i], 0.f, 1.f);
    bx.type = BxDiffuse;
    bx.a = r;
}
// Conductor closure; under a coat (coatEta != 0) its complex IOR is relative to the coat
// (src/materials/material.cpp:140-166).
PG_HD void conductor_bx(const DevScene &s, const prgpu::Material &m, const Surf &si, const Lambda &lam, float coatEta,
                        BxDF &bx) {
    bx.type = BxConductor;
    bx.dist.init(m.alphaX, m.alphaY);
    if (m.eta >= 0 && m.k >= 0) {
        bx.a = spec_sample(s.spec, m.eta, lam);
        bx.b = spec_sample(s.spec, m.k, lam);
        if (coatEta != 0) {
            bx.a = bx.a / coatEta;
            bx.b = bx.b / coatEta;
        }
    } else {
        SS r = albedo_sample(s, checker_pick(m, si.u, si.v, m.albedo), si, lam);
        for (int i = 0; i < 4; ++i) r.v[i] = clampf(r.v[i], 0.f, 0.9999f);
        bx.a = SS(1.f);
        bx.b = ss_sqrt(r) * 2.f / ss_sqrt(ss_clamp0(SS(1.f) - r));
    }
}
PG_HD void dielectric_bx(const DevScene &s, const prgpu::Material &m, Lambda &lam, BxDF &bx) {
    float e = spec_eval(s.spec, m.eta, lam.l[0]);
    if (!spec_is_constant(s.spec, m.eta)) lam.terminate_secondary();
    if (e == 0) e = 1;
    bx.etaD = e;
    if (m.thin) {
        bx.type = BxThin;
    } else {
        bx.type = BxDielectric;
        bx.dist.init(m.alphaX, m.alphaY);
    }
}

// kFamily: -1 = any plain material, else the material family the caller guarantees (0 diffuse,
// 1 conductor, 2 dielectric), which removes the other families' code from the caller.
template <int kFamily = -1>
PG_HD BSDF eval_surface(const DevScene &s, const prgpu::Material &m, const Surf &si, Lambda &lam) {
    BSDF b;
    BxDF &bx = b.bx;
    bx.etaD = 1;
    bx.dist.ax = bx.dist.ay = 0;
    if (kFamily == 0 || (kFamily < 0 && m.type == prgpu::MatDiffuse)) diffuse_bx(s, m, si, lam, bx);
    else if (kFamily == 1 || (kFamily < 0 && m.type == prgpu::MatConductor)) conductor_bx(s, m, si, lam, 0.f, bx);
    else dielectric_bx(s, m, lam, bx);
    b.frame = Frame::from_xz(si.dpdus, si.ns);
    return b;
}

PG_HD bool is_coated(const prgpu::Material &m) {
    return m.type == prgpu::MatCoatedDiffuse || m.type == prgpu::MatCoatedConductor;
}
PG_HD bool has_sheen(const prgpu::Material &m) { return m.sheenWeight > 0 && m.sheenColor >= 0; }

// The coated / sheen family (Material::GetBSDF, src/materials/material.cpp:5-14, 120-166): a
// layered closure under a rough dielectric coat whose IOR is taken at the hero wavelength, or
// a plain closure, optionally wrapped in the sheen lobe.
PG_HD LBSDF eval_closure(const DevScene &s, const prgpu::Material &m, const Surf &si, Lambda &lam) {
    LBSDF b;
    b.layered = b.sheen = false;
    b.bx.etaD = 1;
    b.bx.dist.ax = b.bx.dist.ay = 0;
    if (is_coated(m)) {
        float ce = spec_eval(s.spec, m.coatEta, lam.l[0]);
        if (!spec_is_constant(s.spec, m.coatEta)) lam.terminate_secondary();
        if (ce == 0) ce = 1;
        Layered &L = b.lay;
        b.layered = true;
        L.top.type = BxDielectric;
        L.top.etaD = ce;
        L.top.dist.init(m.coatAlpha, m.coatAlpha);
        L.bottom.etaD = 1;
        L.thick = maxf(m.thickness, 1.17549435e-38f);
        L.g = m.coatG;
        L.albedo = m.coatAlbedo >= 0 ? ss_clamp0(spec_sample(s.spec, m.coatAlbedo, lam)) : SS(0.f);
        L.maxDepth = m.coatMaxDepth;
        L.nSamples = m.coatSamples > 1 ? m.coatSamples : 1;
        if (m.type == prgpu::MatCoatedDiffuse) diffuse_bx(s, m, si, lam, L.bottom);
        else conductor_bx(s, m, si, lam, ce, L.bottom);
    } else if (m.type == prgpu::MatDiffuse) {
        diffuse_bx(s, m, si, lam, b.bx);
    } else if (m.type == prgpu::MatConductor) {
        conductor_bx(s, m, si, lam, 0.f, b.bx);
    } else {
        dielectric_bx(s, m, lam, b.bx);
    }
    if (has_sheen(m)) {
        b.sheen = true;
        SS c = ss_clamp0(spec_sample(s.spec, m.sheenColor, lam));
        b.sh.color = c;
        b.sh.alpha = clampf(m.sheenRoughness, 0.05f, 1.f);
        b.sh.weight = clampf(m.sheenWeight, 0.f, 1.f);
        b.sh.colorMax = clampf(c.maxc(), 0.f, 1.f);
        b.sh.table = s.sheenTable;
    }
    b.frame = Frame::from_xz(si.dpdus, si.ns);
    return b;
}

// ---- shapes as light emitters (src/geometry/shape.cpp:12-37, 87-141, 275-297) ---------------------
struct ShapeSample {
    V3 p, n;
    float u, v;
    float pdf;
};

PG_HD bool shape_sample_area(const prgpu::Primitive &pr, float area, V2 u, ShapeSample *ss) {
    if (pr.type == prgpu::PrimSphere) {
        V3 n = uniform_sphere(u);
        ss->p = f3(pr.p0) + n * pr.r;
        ss->n = pr.flip ? -n : n;
        float phi = atan2f(n.z, n.x);
        if (phi < 0) phi += 2 * kPi;
        ss->u = phi * kInv2Pi;
        ss->v = safe_acos(n.y) * kInvPi;
        ss->pdf = 1 / area;
        return true;
    }
    V3 p0 = f3(pr.p0), p1 = f3(pr.p1), p2 = f3(pr.p2);
    float b0, b1;
    if (u.x < u.y) {
        b0 = u.x / 2;
        b1 = u.y - b0;
    } else {
        b1 = u.y / 2;
        b0 = u.x - b1;
    }
    float b2 = 1 - b0 - b1;
    ss->p = p0 * b0 + p1 * b1 + p2 * b2;
    V3 n = normalize(cross_f(p1 - p0, p2 - p0));
    if (prim_has_normals(pr)) n = face_forward(n, f3(pr.n0) * b0 + f3(pr.n1) * b1 + f3(pr.n2) * b2);
    ss->n = n;
    ss->u = pr.u0 * b0 + pr.u1 * b1 + pr.u2 * b2;
    ss->v = pr.v0 * b0 + pr.v1 * b1 + pr.v2 * b2;
    ss->pdf = 1 / area;
    return true;
}

PG_HD bool shape_sample_from(const prgpu::Primitive &pr, float area, V3 ref, V2 u, ShapeSample *ss) {
    if (pr.type == prgpu::PrimSphere) {
        V3 c = f3(pr.p0);
        float r = pr.r;
        float dc2 = dist2(ref, c);
        if (!(dc2 <= r * r * 1.0001f)) {
            float sinThetaMax = r / sqrtf(dc2);
            float sin2ThetaMax = sq(sinThetaMax);
            float cosThetaMax = safe_sqrt(1 - sin2ThetaMax);
            float omc = 1 - cosThetaMax;
            float cosTheta = (cosThetaMax - 1) * u.x + 1;
            float sin2Theta = 1 - sq(cosTheta);
            if (sin2ThetaMax < 0.00068523f) {
                sin2Theta = sin2ThetaMax * u.x;
                cosTheta = sqrtf(1 - sin2Theta);
                omc = sin2ThetaMax / 2;
            }
            float cosAlpha = sin2Theta / sinThetaMax + cosTheta * safe_sqrt(1 - sin2Theta / sq(sinThetaMax));
            float sinAlpha = safe_sqrt(1 - sq(cosAlpha));
            float phi = u.y * 2 * kPi;
            V3 w = spherical_dir(sinAlpha, cosAlpha, phi);
            Frame fr = Frame::from_z(normalize(c - ref));
            V3 n = fr.from_local(-w);
            ss->p = c + n * r;
            ss->n = pr.flip ? -n : n;
            float uphi = atan2f(n.z, n.x);
            if (uphi < 0) uphi += 2 * kPi;
            ss->u = uphi * kInv2Pi;
            ss->v = safe_acos(n.y) * kInvPi;
            ss->pdf = 1 / (2 * kPi * omc);
            return true;
        }
    }
    if (!shape_sample_area(pr, area, u, ss)) return false;
    V3 wi = ss->p - ref;
    float d2 = len2(wi);
    if (d2 == 0) return false;
    wi = wi / sqrtf(d2);
    float cosTheta = absdot(ss->n, -wi);
    if (cosTheta == 0) return false;
    float pdf = ss->pdf * (d2 / cosTheta);
    if (!is_finite(pdf)) return false;
    ss->pdf = pdf;
    return true;
}

PG_HD float shape_pdf_from(const prgpu::Primitive &pr, float area, V3 ref, V3 wi) {
    if (pr.type == prgpu::PrimSphere) {
        V3 c = f3(pr.p0);
        float r = pr.r;
        float dc2 = dist2(ref, c);
        if (!(dc2 <= r * r * 1.0001f)) {
            float sin2ThetaMax = r * r / dc2;
            float cosThetaMax = safe_sqrt(1 - sin2ThetaMax);
            float omc = 1 - cosThetaMax;
            if (sin2ThetaMax < 0.00068523f) omc = sin2ThetaMax / 2;
            V3 toC = (c - ref) / sqrtf(dc2);
            if (dot(toC, wi) < cosThetaMax - 1e-6f) return 0;
            return 1 / (2 * kPi * omc);
        }
    }
    DRay r{ref.x, ref.y, ref.z, wi.x, wi.y, wi.z};
    float t, b1 = 0, b2 = 0;
    bool hit;
    if (pr.type == prgpu::PrimSphere) {
        hit = dev_hit_sphere(r, pr.p0.x, pr.p0.y, pr.p0.z, pr.r, kInfF, &t);
    } else {
        V3 p0 = f3(pr.p0), e1 = f3(pr.p1) - p0, e2 = f3(pr.p2) - p0;
        hit = dev_hit_triangle(r, p0.x, p0.y, p0.z, e1.x, e1.y, e1.z, e2.x, e2.y, e2.z, kInfF, &t, &b1, &b2);
    }
    if (!hit) return 0;
    Surf si;
    interaction(pr, ref, wi, t, b1, b2, &si);
    float cosTheta = absdot(si.n, -wi);
    if (cosTheta == 0) return 0;
    float pdf = (1 / area) * dist2(ref, si.p) / cosTheta;
    return is_finite(pdf) ? pdf : 0.f;
}

// ---- lights (src/lights/light.cpp, sky.cpp:52-60, light_bvh.cpp:179-234) ---------------------------
PG_HD SS area_L(const DevScene &s, const prgpu::LightEx &L, V3 n, V3 w, const Lambda &lam) {
    float c = dot(n, w);
    if (!L.twoSided && c <= 0) return SS(0.f);
    float prof = L.cosPower > 0 ? powf(fabsf(c), L.cosPower) : 1.f;
    return spec_sample(s.spec, L.spectrum, lam) * (L.scale * prof);
}

struct LightSample {
    SS L;
    V3 wi;
    float pdf;
    V3 pl, nl;  // point on the light (and its normal, zero for non-surfaces)
};

// Radiance of an image environment in world direction w (EnvironmentLight::Lookup,
// src/lights/light.cpp:249-261): the bilinear map value times the light's scale, as an RGB
// illuminant (RGBIlluminantSample, src/core/rgb2spec.cpp:322-343): 2 m sigmoid(rgb / (2 m)) D65.
PG_HD SS env_image_L(const DevScene &s, const prgpu::LightEx &L, V3 w, const Lambda &lam) {
    const EnvMap e = env_map(s.spec.knots + L.profileOffset);
    float u, v, rgb[3];
    dir_to_equirect(normalize(env_to_light(e, w)), &u, &v);
    env_rgb(e, u, v, rgb);
    for (int i = 0; i < 3; ++i) rgb[i] = rgb[i] * L.scale;
    const float m = maxf(0.f, maxf(rgb[0], maxf(rgb[1], rgb[2])));
    if (m == 0) return SS(0.f);
    const float scale = 2 * m, inv = 1 / scale;
    const float q[3] = {rgb[0] * inv, rgb[1] * inv, rgb[2] * inv};
    float c[3];
    rgb_sigmoid_coeffs(s.spec.knots + int(e.hdr[7]), q, c);
    const SS d65 = spec_sample(s.spec, int(e.hdr[6]), lam);
    SS out;
    for (int i = 0; i < 4; ++i) out.v[i] = scale * sigmoid_poly(c, lam.l[i]) * d65.v[i];
    return out;
}
PG_HD bool env_is_image(const prgpu::LightEx &L) { return L.kind == prgpu::LightExEnvironment && L.profileSize > 0; }
// Radiance an escaping ray in direction w receives from environment light L.
PG_HD SS env_L(const DevScene &s, const prgpu::LightEx &L, V3 w, const Lambda &lam) {
    return env_is_image(L) ? env_image_L(s, L, w, lam) : spec_sample(s.spec, L.spectrum, lam) * L.scale;
}

PG_HD bool sample_li(const DevScene &s, int li, V3 p, V2 u, const Lambda &lam, LightSample *ls) {
    const prgpu::LightEx &L = s.lights[li];
    if (L.kind == prgpu::LightExArea) {
        const prgpu::Primitive &pr = s.prims[L.primitive];
        ShapeSample ss;
        if (!shape_sample_from(pr, s.lightArea[li], p, u, &ss) || ss.pdf == 0 || len2(ss.p - p) == 0) return false;
        V3 wi = normalize(ss.p - p);
        SS Le = area_L(s, L, ss.n, -wi, lam);
        if (!Le.nonzero()) return false;
        *ls = {Le, wi, ss.pdf, ss.p, ss.n};
        return true;
    }
    if (L.kind == prgpu::LightExPoint) {
        V3 d = f3(L.position) - p;
        float d2 = len2(d);
        if (d2 == 0) return false;
        *ls = {spec_sample(s.spec, L.spectrum, lam) * (L.scale / d2), d / sqrtf(d2), 1, f3(L.position), V3()};
        return true;
    }
    if (L.kind == prgpu::LightExSpot) {  // light.cpp:168-192
        V3 d = f3(L.position) - p;
        float d2 = len2(d);
        if (d2 == 0) return false;
        V3 wi = d / sqrtf(d2);
        float f = smooth_step(dot(-wi, f3(L.direction)), L.cosFalloffEnd, L.cosFalloffStart);
        if (f == 0) return false;
        *ls = {spec_sample(s.spec, L.spectrum, lam) * (L.scale * f / d2), wi, 1, f3(L.position), V3()};
        return true;
    }
    if (L.kind == prgpu::LightExGoniometric) {  // sky.cpp:12-27
        V3 d = f3(L.position) - p;
        float d2 = len2(d);
        if (d2 == 0) return false;
        V3 wi = d / sqrtf(d2);
        float prof = ies_eval(s.spec.knots + L.profileOffset, -wi.x, -wi.y, -wi.z);
        if (prof <= 0) return false;
        *ls = {spec_sample(s.spec, L.spectrum, lam) * (L.scale * prof / d2), wi, 1, f3(L.position), V3()};
        return true;
    }
    if (L.kind == prgpu::LightExSun) {  // sky.cpp:97-105
        float cosMax = L.cosFalloffEnd;
        float cosT = (1 - u.x) + u.x * cosMax;  // SampleUniformCone (sampling.h:47-53)
        V3 wl = spherical_dir(safe_sqrt(1 - cosT * cosT), cosT, u.y * 2 * kPi);
        V3 wi = Frame::from_z(f3(L.direction)).from_local(wl);
        *ls = {spec_sample(s.spec, L.spectrum, lam) * L.scale, wi, 1 / (2 * kPi * (1 - cosMax)),
               p + wi * (2 * s.worldR), V3()};
        return true;
    }
    if (L.kind == prgpu::LightExDistant) {
        V3 w = f3(L.direction);
        *ls = {spec_sample(s.spec, L.spectrum, lam) * L.scale, w, 1, p + w * (2 * s.worldR), V3()};
        return true;
    }
    if (env_is_image(L)) {  // EnvironmentLight::SampleLi (src/lights/light.cpp:267-289)
        const EnvMap e = env_map(s.spec.knots + L.profileOffset);
        float mu, mv, mapPdf;
        env_sample_uv(e, u.x, u.y, &mu, &mv, &mapPdf);
        if (mapPdf == 0) return false;
        float sinTheta = sinf(mv * kPi);
        if (sinTheta == 0) return false;
        V3 wi = env_to_world(e, equirect_to_dir(mu, mv));
        *ls = {env_image_L(s, L, wi, lam), wi, mapPdf / (2 * kPi * kPi * sinTheta), p + wi * (2 * s.worldR), V3()};
        return true;
    }
    V3 wi = uniform_sphere(u);
    *ls = {spec_sample(s.spec, L.spectrum, lam) * L.scale, wi, kInv4Pi, p + wi * (2 * s.worldR), V3()};
    return true;
}

PG_HD float pdf_li(const DevScene &s, int li, V3 p, V3 wi) {
    const prgpu::LightEx &L = s.lights[li];
    if (L.kind == prgpu::LightExArea) return shape_pdf_from(s.prims[L.primitive], s.lightArea[li], p, wi);
    if (env_is_image(L)) {  // EnvironmentLight::PDF_Li (src/lights/light.cpp:291-297)
        const EnvMap e = env_map(s.spec.knots + L.profileOffset);
        float mu, mv;
        dir_to_equirect(normalize(env_to_light(e, wi)), &mu, &mv);
        float sinTheta = sinf(mv * kPi);
        if (sinTheta == 0) return 0;
        return env_pdf_uv(e, mu, mv) / (2 * kPi * kPi * sinTheta);
    }
    if (L.kind == prgpu::LightExEnvironment) return kInv4Pi;
    if (L.kind == prgpu::LightExSun)  // sky.cpp:107-109
        return dot(normalize(wi), f3(L.direction)) >= L.cosFalloffEnd ? 1 / (2 * kPi * (1 - L.cosFalloffEnd)) : 0.f;
    return 0;
}

PG_HD bool light_is_delta(const DevScene &s, int li) {
    uint32_t k = s.lights[li].kind;
    return k == prgpu::LightExPoint || k == prgpu::LightExDistant || k == prgpu::LightExSpot ||
           k == prgpu::LightExGoniometric;
}

PG_HD float p_infinite(const DevScene &s) {
    return s.infCount == 0 ? 0.f : float(s.infCount) / float(s.infCount + (s.lnodeCount == 0 ? 0 : 1));
}

PG_HD int light_sample(const DevScene &s, V3 p, V3 n, float u, float *pmf) {
    *pmf = 0;
    if (s.lightSampler == prgpu::LightSamplerPower) {  // Distribution1D::SampleDiscrete
        if (s.powerCount == 0) return -1;
        int size = s.powerCount + 1, first = 1, len = size - 2;
        while (len > 0) {
            int half = len >> 1, mid = first + half;
            bool r = s.powerCdf[mid] <= u;
            first = r ? mid + 1 : first;
            len = r ? len - (half + 1) : half;
        }
        int off = clampi(first - 1, 0, size - 2);
        *pmf = s.powerInt > 0 ? s.powerFunc[off] / (s.powerInt * s.powerCount) : 1.f / s.powerCount;
        return off;
    }
    float pInf = p_infinite(s);
    if (u < pInf) {
        u /= pInf;
        int index = min(int(u * s.infCount), s.infCount - 1);
        *pmf = pInf / s.infCount;
        return s.infinite[index];
    }
    if (s.lnodeCount == 0) return -1;
    u = minf((u - pInf) / (1 - pInf), kOneMinusEps);
    int ni = 0;
    float prob = 1 - pInf;
    for (int guard = 0; guard < 64; ++guard) {
        const LNode &node = s.lnodes[ni];
        if (node.light >= 0) {
            if (ni > 0 || lb_importance(node, p, n) > 0) {
                *pmf = prob;
                return node.light;
            }
            return -1;
        }
        float ci0 = lb_importance(s.lnodes[ni + 1], p, n), ci1 = lb_importance(s.lnodes[node.child1], p, n);
        if (ci0 == 0 && ci1 == 0) return -1;
        float p0 = ci0 / (ci0 + ci1);
        if (u < p0) {
            prob *= p0;
            u = minf(u / p0, kOneMinusEps);
            ni = ni + 1;
        } else {
            prob *= 1 - p0;
            u = minf((u - p0) / (1 - p0), kOneMinusEps);
            ni = node.child1;
        }
    }
    return -1;
}

PG_HD float light_pmf(const DevScene &s, V3 p, V3 n, int li) {
    if (s.lightSampler == prgpu::LightSamplerPower)
        return s.powerInt > 0 ? s.powerFunc[li] / (s.powerInt * s.powerCount) : 1.f / s.powerCount;
    int leaf = s.leafOf[li];
    if (leaf < 0) {
        for (int i = 0; i < s.infCount; ++i)
            if (s.infinite[i] == li) return p_infinite(s) / s.infCount;
        return 0;
    }
    if (leaf == 0) return lb_importance(s.lnodes[0], p, n) > 0 ? 1 - p_infinite(s) : 0.f;
    float pm = 1 - p_infinite(s);
    for (int child = leaf, par = s.lnodes[leaf].parent, guard = 0; par >= 0 && guard < 64;
         child = par, par = s.lnodes[par].parent, ++guard) {
        const LNode &node = s.lnodes[par];
        float ci0 = lb_importance(s.lnodes[par + 1], p, n), ci1 = lb_importance(s.lnodes[node.child1], p, n);
        if (ci0 == 0 && ci1 == 0) return 0;
        pm *= (child == node.child1 ? ci1 : ci0) / (ci0 + ci1);
    }
    return pm;
}

// ---- media (src/media/medium.cpp) -----------------------------------------------------------------
// Natural log of a float, rounded from double precision: the correctly rounded value except in
// rare double-rounding cases. The CPU's logf is within 0.52 ulp and returns the same bits for
// more than 99% of arguments, so free-flight distances, and with them the ray origins that the
// grid medium hashes into its random stream, match the CPU's far more often than with a 1-ulp
// float log.

struct DistSample {
    bool scattered;
    float t;
    SS weight, pdfRatio;
};

PG_HD bool grid_clip(const prgpu::MediumEx &m, V3 ro, V3 rd, float tMax, V3 *mo, V3 *md, float *t0, float *t1) {
    const float *w = m.worldToMedium;
    *mo = V3(w[0] * ro.x + w[1] * ro.y + w[2] * ro.z + w[3], w[4] * ro.x + w[5] * ro.y + w[6] * ro.z + w[7],
             w[8] * ro.x + w[9] * ro.y + w[10] * ro.z + w[11]);
    *md = V3(w[0] * rd.x + w[1] * rd.y + w[2] * rd.z, w[4] * rd.x + w[5] * rd.y + w[6] * rd.z,
             w[8] * rd.x + w[9] * rd.y + w[10] * rd.z);
    *t0 = 0;
    *t1 = tMax;
    Box3 b(f3(m.boundsMin), f3(m.boundsMax));
    if (b.empty()) return is_finite(tMax);
    for (int a = 0; a < 3; ++a) {
        float inv = 1 / (*md)[a];
        float tn = (b.lo[a] - (*mo)[a]) * inv, tf = (b.hi[a] - (*mo)[a]) * inv;
        if (tn > tf) {
            float t = tn;
            tn = tf;
            tf = t;
        }
        if (is_nan(tn) || is_nan(tf)) {
            if ((*mo)[a] < b.lo[a] || (*mo)[a] > b.hi[a]) return false;
            continue;
        }
        *t0 = maxf(*t0, tn);
        *t1 = minf(*t1, tf);
        if (*t0 > *t1) return false;
    }
    return true;
}

PG_HD float grid_at(const DevScene &s, const prgpu::MediumEx &m, int x, int y, int z) {
    x = clampi(x, 0, m.nx - 1);
    y = clampi(y, 0, m.ny - 1);
    z = clampi(z, 0, m.nz - 1);
    return s.grid[m.gridOffset + (int64_t(z) * m.ny + y) * m.nx + x];
}
PG_HD float grid_density(const DevScene &s, const prgpu::MediumEx &m, V3 p) {
    Box3 b(f3(m.boundsMin), f3(m.boundsMax));
    if (!b.inside(p)) return 0;
    V3 o = b.offset(p);
    float x = o.x * m.nx - 0.5f, y = o.y * m.ny - 0.5f, z = o.z * m.nz - 0.5f;
    int x0 = int(floorf(x)), y0 = int(floorf(y)), z0 = int(floorf(z));
    float dx = x - x0, dy = y - y0, dz = z - z0;
    float c00 = lerp(dx, grid_at(s, m, x0, y0, z0), grid_at(s, m, x0 + 1, y0, z0));
    float c10 = lerp(dx, grid_at(s, m, x0, y0 + 1, z0), grid_at(s, m, x0 + 1, y0 + 1, z0));
    float c01 = lerp(dx, grid_at(s, m, x0, y0, z0 + 1), grid_at(s, m, x0 + 1, y0, z0 + 1));
    float c11 = lerp(dx, grid_at(s, m, x0, y0 + 1, z0 + 1), grid_at(s, m, x0 + 1, y0 + 1, z0 + 1));
    return lerp(dz, lerp(dy, c00, c10), lerp(dy, c01, c11));
}

PG_HD SS medium_tr(const DevScene &s, int mi, V3 ro, V3 rd, float tMax, const Lambda &lam) {
    const prgpu::MediumEx &m = s.media[mi];
    if (m.kind == prgpu::MediumHomogeneous) {
        SS st = spec_sample(s.spec, m.sigmaA, lam) * m.scale + spec_sample(s.spec, m.sigmaS, lam) * m.scale;
        if (is_inf(tMax)) return {st.v[0] > 0 ? 0.f : 1.f, st.v[1] > 0 ? 0.f : 1.f, st.v[2] > 0 ? 0.f : 1.f,
                                  st.v[3] > 0 ? 0.f : 1.f};
        return ss_exp_neg(st, tMax);
    }
    V3 mo, md;
    float t0, t1;
    if (!grid_clip(m, ro, rd, tMax, &mo, &md, &t0, &t1)) return SS(1.f);
    SS st = (spec_sample(s.spec, m.sigmaA, lam) + spec_sample(s.spec, m.sigmaS, lam)) * m.scale;
    float mu = m.maxDensity * st.maxc();
    if (mu <= 0) return SS(1.f);
    uint64_t w1[3] = {fbits(ro.x, ro.y), fbits(ro.z, rd.x), fbits(rd.y, rd.z)};
    Pcg rng;
    rng.set_sequence(murmur_words(w1, 3, false, 0), murmur_words(nullptr, 0, true, __float_as_uint(tMax)));
    SS T(1.f);
    float t = t0;
    for (int iter = 0; iter < 100000; ++iter) {
        t -= log_f32(1 - rng.uniform()) / mu;
        if (t >= t1) break;
        float d = grid_density(s, m, mo + md * t);
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

PG_HD DistSample medium_sample(const DevScene &s, int mi, V3 ro, V3 rd, float tMax, float u, const Lambda &lam) {
    const prgpu::MediumEx &m = s.media[mi];
    DistSample ds;
    ds.scattered = false;
    ds.pdfRatio = SS(1.f);
    if (m.kind == prgpu::MediumHomogeneous) {  // SampleDistanceHero (medium.cpp:71-103)
        SS ss = spec_sample(s.spec, m.sigmaS, lam) * m.scale;
        SS st = spec_sample(s.spec, m.sigmaA, lam) * m.scale + ss;
        float t = st.v[0] > 0 ? -log_f32(1 - minf(u, kOneMinusEps)) / st.v[0] : kInfF;
        if (t < tMax) {
            SS Tr = ss_exp_neg(st, t);
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
        SS Tr = medium_tr(s, mi, ro, rd, tMax, lam);
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
    V3 mo, md;
    float t0, t1;
    if (!grid_clip(m, ro, rd, tMax, &mo, &md, &t0, &t1)) return ds;
    SS sa = spec_sample(s.spec, m.sigmaA, lam) * m.scale, ssc = spec_sample(s.spec, m.sigmaS, lam) * m.scale;
    SS st = sa + ssc;
    float mu = m.maxDensity * st.maxc();
    if (mu <= 0) return ds;
    auto avg = [&](const SS &x) {
        float sum = 0;
        int n = 0;
        for (int i = 0; i < 4; ++i)
            if (lam.pdf[i] != 0) {
                sum += fabsf(x.v[i]);
                ++n;
            }
        return n ? sum / n : 0.f;
    };
    uint64_t wa[2] = {fbits(u, ro.x), fbits(ro.y, ro.z)}, wb[2] = {fbits(rd.x, rd.y), fbits(rd.z, tMax)};
    Pcg rng;
    rng.set_sequence(murmur_words(wa, 2, false, 0), murmur_words(wb, 2, false, 0));
    float t = t0;
    for (int iter = 0; iter < 100000; ++iter) {
        t -= log_f32(1 - rng.uniform()) / mu;
        if (t >= t1) break;
        float d = grid_density(s, m, mo + md * t);
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

}  // namespace pgpu
