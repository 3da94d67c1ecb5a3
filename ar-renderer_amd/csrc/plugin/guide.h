// Learned path guiding: the model's layout and its device-side mixture (decode, pdf, sampling).
//
// The model maps a surface vertex (position in the scene box, normal, outgoing direction) to a
// mixture of 8 von Mises-Fisher lobes over world directions plus the probability s of sampling
// that mixture instead of the BSDF. A guided vertex samples the one-sample MIS proposal
//     q(w) = s * sum_k pi_k vMF(w; mu_k, kappa_k) + (1 - s) * p_bsdf(w)
// and uses q as its scattering pdf (in the throughput and in both MIS weights), which keeps the
// estimator unbiased whatever the network predicts. Training minimises a blend of the KL
// divergences of q and of the mixture from the incident radiance (Mueller et al. 2019, "Neural
// Importance Sampling", where the selection probability is learned through the MIS proposal).
//
// Network: 64 features -> 4 x 64 ReLU -> 41 outputs (row-padded to 48), BF16 weights stored
// transposed (Wt[out][in]) with FP32 biases and FP32 accumulation. Parameters live in one FP32
// vector: per layer the weights Wt (rows x 64) then the biases (rows), rows = 64, 64, 64, 64, 48.
#pragma once

#include <hip/hip_runtime.h>

#include "pt_common.h"

namespace pgpu {

constexpr int kGIn = 64, kGHidden = 64, kGOutRows = 48, kGOut = 41, kGLobes = 8, kGLayers = 5;
#ifndef PG_GUIDE_SEL_MIN
#define PG_GUIDE_SEL_MIN 0.1f
#endif
constexpr float kGSelMin = PG_GUIDE_SEL_MIN, kGSelMax = 0.9f, kGKappaMin = 0.01f, kGKappaMax = 1000.f;
// Training loss per record: -w [kGBlend log q + (1 - kGBlend) log p_guide]. The first term fits the
// one-sample MIS proposal q (and is the only one that moves s); the second fits the mixture itself,
// so its lobes keep learning the incident radiance while s is small (with q alone, the lobes'
// gradient scales with s, and a poor early mixture drives s to its floor and stalls there).
constexpr float kGBlend = 0.5f;
// Training records a path keeps during a training render: one per guided vertex, for its first
// kGuideRecords guided vertices.
#ifndef PG_GUIDE_RECORDS
#define PG_GUIDE_RECORDS 2
#endif
constexpr int kGuideRecords = PG_GUIDE_RECORDS;

PG_HD constexpr int guide_rows(int l) { return l == kGLayers - 1 ? kGOutRows : kGHidden; }
PG_HD constexpr int guide_woff(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += guide_rows(i) * kGIn + guide_rows(i);
    return o;
}
PG_HD constexpr int guide_boff(int l) { return guide_woff(l) + guide_rows(l) * kGIn; }
constexpr int kGParams = guide_woff(kGLayers);       // 19760 (19305 used)
constexpr int kGWeights = (4 * kGHidden + kGOutRows) * kGIn;  // bf16 weights, all layers

// Feature f (0..63) of a vertex: sin / cos of 2^j pi x on each normalised axis (48), then the
// normal's encoding (16): the normal, sin / cos of pi n on each axis, sin / cos of 2 pi n on each
// axis, one zero. The incident radiance a diffuse vertex samples depends on where it is and how it
// faces, not on the direction the path arrived from.
__device__ __forceinline__ float guide_feature(int f, const float *pos, const float *nrm, const float *wo) {
    (void)wo;
    if (f < 48) {
        const int axis = f >> 4, j = (f & 15) >> 1;
        const float arg = pos[axis] * float(1 << j);
        return (f & 1) ? cospif(arg) : sinpif(arg);
    }
    if (f < 51) return nrm[f - 48];
    if (f < 63) {
        const int g = f - 51, axis = (g % 6) >> 1;
        const float arg = nrm[axis] * (g < 6 ? 1.f : 2.f);
        return (g & 1) ? cospif(arg) : sinpif(arg);
    }
    return 0.f;
}

__device__ __forceinline__ float softplusf(float x) { return x > 20.f ? x : log1pf(expf(x)); }

// vMF density with the stable normaliser kappa / (2 pi (1 - e^{-2 kappa})).
__device__ __forceinline__ float vmf_pdf(V3 mu, float kappa, V3 w) {
    return kappa / (2.f * kPi * -expm1f(-2.f * kappa)) * expf(kappa * (dot(mu, w) - 1.f));
}

// Lobe k of the raw outputs o[48]: mean direction (normalised; +z when degenerate), concentration.
__device__ __forceinline__ void guide_lobe(const float *o, int k, V3 *mu, float *kappa) {
    V3 v(o[5 * k], o[5 * k + 1], o[5 * k + 2]);
    float l = sqrtf(dot(v, v));
    *mu = l > 1e-6f ? v / l : V3(0.f, 0.f, 1.f);
    *kappa = fminf(kGKappaMin + softplusf(o[5 * k + 3]), kGKappaMax);
}

// A decoded mixture, as the inference kernel stores it per vertex (12 float4): lobes
// (mu.x, mu.y, mu.z, kappa) in [0, 8), the pdf coefficients c_k = pi_k kappa_k / (2 pi (1 -
// e^{-2 kappa_k})) in [8, 10), the lobe CDF pi_0 .. pi_0 + .. + pi_6 and s in [10, 12).
struct GuideMix {
    const float4 *d;
    float sel;
};

__device__ __forceinline__ GuideMix guide_mix(const float4 *d) {
    GuideMix m;
    m.d = d;
    m.sel = d[11].w;
    return m;
}

// Decodes the 48 raw outputs o into the stored form (used by the inference kernel).
__device__ __forceinline__ void guide_decode(const float *o, float4 *d) {
    float amax = o[4];
    for (int k = 1; k < kGLobes; ++k) amax = fmaxf(amax, o[5 * k + 4]);
    float w[kGLobes], z = 0.f;
    for (int k = 0; k < kGLobes; ++k) {
        w[k] = expf(o[5 * k + 4] - amax);
        z += w[k];
    }
    float c[kGLobes], cdf[kGLobes];
    float acc = 0.f;
    for (int k = 0; k < kGLobes; ++k) {
        V3 mu;
        float kappa;
        guide_lobe(o, k, &mu, &kappa);
        const float pi = w[k] / z;
        d[k] = make_float4(mu.x, mu.y, mu.z, kappa);
        c[k] = pi * kappa / (2.f * kPi * -expm1f(-2.f * kappa));
        acc += pi;
        cdf[k] = acc;
    }
    d[8] = make_float4(c[0], c[1], c[2], c[3]);
    d[9] = make_float4(c[4], c[5], c[6], c[7]);
    d[10] = make_float4(cdf[0], cdf[1], cdf[2], cdf[3]);
    d[11] = make_float4(cdf[4], cdf[5], cdf[6], kGSelMin + (kGSelMax - kGSelMin) / (1.f + expf(-o[40])));
}

__device__ __forceinline__ float guide_pdf(const GuideMix &m, V3 w) {
    const float4 c0 = m.d[8], c1 = m.d[9];
    const float c[kGLobes] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    float p = 0.f;
#pragma unroll
    for (int k = 0; k < kGLobes; ++k) {
        const float4 l = m.d[k];
        p += c[k] * expf(l.w * (l.x * w.x + l.y * w.y + l.z * w.z - 1.f));
    }
    return p;
}

// Lobe by the CDF (ul), then the vMF lobe (Jakob 2012) in the frame around its mean.
__device__ __forceinline__ V3 guide_sample(const GuideMix &m, float ul, float u1, float u2) {
    const float4 f0 = m.d[10], f1 = m.d[11];
    const float cdf[kGLobes - 1] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z};
    int k = 0;
    while (k < kGLobes - 1 && !(ul < cdf[k])) ++k;
    const float4 l = m.d[k];
    const V3 mu(l.x, l.y, l.z);
    const float kappa = l.w;
    float cosT = 1.f + log1pf(u1 * expm1f(-2.f * kappa)) / kappa;
    cosT = fminf(1.f, fmaxf(-1.f, cosT));
    float sinT = sqrtf(fmaxf(0.f, 1.f - cosT * cosT)), phi = 2.f * kPi * u2;
    float sign = copysignf(1.f, mu.z), a = -1.f / (sign + mu.z), b = mu.x * mu.y * a;
    V3 tt(1.f + sign * mu.x * mu.x * a, sign * b, -sign * mu.x), ss(b, sign + mu.y * mu.y * a, -mu.y);
    return tt * (sinT * cosf(phi)) + ss * (sinT * sinf(phi)) + mu * cosT;
}

// Raw-output helpers for the training kernel (the loss is differentiated through the raw outputs).
struct GuideRaw {
    const float *o;
    float amax, z, sel;
};

__device__ __forceinline__ GuideRaw guide_raw(const float *o) {
    GuideRaw m;
    m.o = o;
    m.amax = o[4];
    for (int k = 1; k < kGLobes; ++k) m.amax = fmaxf(m.amax, o[5 * k + 4]);
    m.z = 0.f;
    for (int k = 0; k < kGLobes; ++k) m.z += expf(o[5 * k + 4] - m.amax);
    m.sel = kGSelMin + (kGSelMax - kGSelMin) / (1.f + expf(-o[40]));
    return m;
}

}  // namespace pgpu
