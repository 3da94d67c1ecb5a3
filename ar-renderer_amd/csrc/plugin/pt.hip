// Wavefront spectral path tracer for gfx950.
//
// A batch holds P passes x (sample width x sample height) camera samples, one path slot each.
// Per bounce: k_intersect (closest hit of every queued path), k_shade (one iteration of the
// reference's Li loop: media, emission, interfaces, BSDF, next-event estimation, Russian
// roulette; surviving paths are appended to the next queue, shadow rays to the shadow queue),
// k_shadow (transmittance of the shadow rays; visible contributions are added to their path).
// When the queue drains, k_finish turns each path's radiance into a film sample and k_gather
// splats the samples of every pass onto the image with the film's reconstruction filter,
// adding the pass value and its square to the FP64 sum and sum-of-squares accumulators.
//
// Path state is structure-of-arrays in HBM (float4 / int4 per field) so that the lanes of a
// wave read and write consecutive 16-byte words.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <optional>
#include <unordered_map>
#include <vector>

#include "guide_host.h"
#include "pt_dev.h"
#include "sheen_table.h"
#include "pt_host.h"

namespace pgpu {

// =============================================================================================
// Host: light BVH (restates src/lights/light_bvh.cpp:10-177 so its PMF matches the CPU's)

namespace {

struct HostLB {
    Box3 bounds;
    V3 w{0, 0, 1};
    float phi = 0, cosTheta_o = 1, cosTheta_e = 0;
    bool twoSided = false;
};

struct HostCone {
    V3 w{0, 0, 1};
    float cosTheta = kInfF;
    HostCone() = default;
    HostCone(V3 ww, float c) : w(normalize(ww)), cosTheta(c) {}
    bool empty() const { return cosTheta == kInfF; }
};

V3 RotateVector(float degrees, V3 axis, V3 v) {  // src/core/math.cpp:71-86
    V3 a = normalize(axis);
    float rad = (kPi / 180.f) * degrees;
    float s = std::sin(rad), c = std::cos(rad);
    float m00 = a.x * a.x + (1 - a.x * a.x) * c, m01 = a.x * a.y * (1 - c) - a.z * s, m02 = a.x * a.z * (1 - c) + a.y * s;
    float m10 = a.x * a.y * (1 - c) + a.z * s, m11 = a.y * a.y + (1 - a.y * a.y) * c, m12 = a.y * a.z * (1 - c) - a.x * s;
    float m20 = a.x * a.z * (1 - c) - a.y * s, m21 = a.y * a.z * (1 - c) + a.x * s, m22 = a.z * a.z + (1 - a.z * a.z) * c;
    return {m00 * v.x + m01 * v.y + m02 * v.z, m10 * v.x + m11 * v.y + m12 * v.z, m20 * v.x + m21 * v.y + m22 * v.z};
}

float AngleBetween(V3 a, V3 b) {
    if (dot(a, b) < 0) return kPi - 2 * safe_asin(len(a + b) / 2);
    return 2 * safe_asin(len(b - a) / 2);
}

HostCone ConeUnion(const HostCone &a, const HostCone &b) {
    if (a.empty()) return b;
    if (b.empty()) return a;
    float ta = safe_acos(a.cosTheta), tb = safe_acos(b.cosTheta);
    float td = AngleBetween(a.w, b.w);
    if (std::min(td + tb, kPi) <= ta) return a;
    if (std::min(td + ta, kPi) <= tb) return b;
    float to = (ta + td + tb) / 2;
    if (to >= kPi) return HostCone(V3(0, 0, 1), -1);
    float tr = to - ta;
    V3 wr = cross_d(a.w, b.w);
    if (len2(wr) == 0) return HostCone(V3(0, 0, 1), -1);
    V3 w = RotateVector((180.f / kPi) * tr, wr, a.w);
    return HostCone(w, std::cos(to));
}

HostLB LBUnion(const HostLB &a, const HostLB &b) {
    if (a.phi == 0) return b;
    if (b.phi == 0) return a;
    HostCone c = ConeUnion(HostCone(a.w, a.cosTheta_o), HostCone(b.w, b.cosTheta_o));
    HostLB r;
    r.bounds = join(a.bounds, b.bounds);
    r.w = c.w;
    r.phi = a.phi + b.phi;
    r.cosTheta_o = c.cosTheta;
    r.cosTheta_e = std::min(a.cosTheta_e, b.cosTheta_e);
    r.twoSided = a.twoSided || b.twoSided;
    return r;
}

float LBCost(const HostLB &b, const Box3 &bounds, int dim) {
    float t_o = safe_acos(b.cosTheta_o), t_e = safe_acos(b.cosTheta_e);
    float t_w = std::min(t_o + t_e, kPi);
    float s_o = safe_sqrt(1 - sq(b.cosTheta_o));
    float M = 2 * kPi * (1 - b.cosTheta_o) +
              kPi / 2 * (2 * t_w * s_o - std::cos(t_o - 2 * t_w) - 2 * t_o * s_o + b.cosTheta_o);
    V3 d = bounds.diag();
    float Kr = d[dim] > 0 ? max_comp(d) / d[dim] : 1.f;
    return b.phi * M * Kr * b.bounds.area();
}

struct LightTreeBuilder {
    std::vector<int> bounded, infinite;
    std::vector<HostLB> lbs;
    std::vector<int> child1, light, parent;

    int Rec(std::vector<std::pair<int, HostLB>> &items, int start, int end, int par) {
        if (end - start == 1) {
            int ni = int(lbs.size());
            lbs.push_back(items[size_t(start)].second);
            child1.push_back(-1);
            light.push_back(bounded[size_t(items[size_t(start)].first)]);
            parent.push_back(par);
            return ni;
        }
        Box3 bounds, cb;
        for (int i = start; i < end; ++i) {
            bounds = join(bounds, items[size_t(i)].second.bounds);
            cb = join(cb, items[size_t(i)].second.bounds.centroid());
        }
        float minCost = kInfF;
        int minBucket = -1, minDim = -1;
        constexpr int kB = 12;
        for (int dim = 0; dim < 3; ++dim) {
            if (cb.hi[dim] == cb.lo[dim]) continue;
            HostLB bucket[kB];
            for (int i = start; i < end; ++i) {
                V3 pc = items[size_t(i)].second.bounds.centroid();
                int b = int(kB * cb.offset(pc)[dim]);
                b = clampi(b, 0, kB - 1);
                bucket[b] = LBUnion(bucket[b], items[size_t(i)].second);
            }
            for (int i = 0; i < kB - 1; ++i) {
                HostLB b0, b1;
                for (int j = 0; j <= i; ++j) b0 = LBUnion(b0, bucket[j]);
                for (int j = i + 1; j < kB; ++j) b1 = LBUnion(b1, bucket[j]);
                float cost = LBCost(b0, bounds, dim) + LBCost(b1, bounds, dim);
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
            auto it = std::partition(items.begin() + start, items.begin() + end, [&](const std::pair<int, HostLB> &l) {
                int b = int(kB * cb.offset(l.second.bounds.centroid())[minDim]);
                return clampi(b, 0, kB - 1) <= minBucket;
            });
            mid = int(it - items.begin());
            if (mid == start || mid == end) mid = (start + end) / 2;
        }
        int ni = int(lbs.size());
        lbs.emplace_back();
        child1.push_back(-1);
        light.push_back(-1);
        parent.push_back(par);
        Rec(items, start, mid, ni);
        int c1 = Rec(items, mid, end, ni);
        child1[size_t(ni)] = c1;
        lbs[size_t(ni)] = LBUnion(lbs[size_t(ni + 1)], lbs[size_t(c1)]);
        return ni;
    }
};

float PrimArea(const prgpu::Primitive &p) {
    if (p.type == prgpu::PrimSphere) return 4 * kPi * p.r * p.r;
    V3 a = f3(p.p0), b = f3(p.p1), c = f3(p.p2);
    return 0.5f * len(cross_d(b - a, c - a));
}

Box3 PrimBounds(const prgpu::Primitive &p) {
    if (p.type == prgpu::PrimSphere) {
        V3 c = f3(p.p0);
        return {c - V3(p.r, p.r, p.r), c + V3(p.r, p.r, p.r)};
    }
    V3 a = f3(p.p0), b = f3(p.p1), c = f3(p.p2);
    return join(Box3(vmin(a, b), vmax(a, b)), c);
}

}  // namespace

// =============================================================================================
// Device-resident path-tracing scene and batch buffers

// Minimum waves per SIMD a kernel is compiled for (its register budget: 8 waves = 64 VGPRs,
// 6 = 80, 5 = 96, 4 = 128, 3 = 168); 0 leaves the choice to the compiler. The traversal is
// latency-bound and its LDS stacks allow 8 waves per SIMD, which 64 VGPRs reach without spilling.
// The plain material families take 175-185 VGPRs unconstrained (2 waves); measured on
// configuration 2 at 1920x1080, 4 spp: 5.13 ms per step at 2 waves, 4.69 at 3, 4.80 at 4 (the
// spills of a 128-VGPR budget cost more than the occupancy gains). The surface stage is fastest
// unconstrained (92 VGPRs: 1.61 ms; 1.71 at 6 waves, 2.23 at 8).
#define PG_WAVES(n) __attribute__((amdgpu_waves_per_eu((n) > 0 ? (n) : 1, 8)))
#ifndef PG_SURFACE_WAVES
#define PG_SURFACE_WAVES 0
#endif
#ifndef PG_MATERIAL_WAVES
#define PG_MATERIAL_WAVES 3
#endif
constexpr int kSurfaceWaves = PG_SURFACE_WAVES, kMaterialWaves = PG_MATERIAL_WAVES;

// Bufs::counts: queue lengths and the ray fetch counters of the persistent traversals.
constexpr int kFetchIntersect = 16, kFetchShadow = 32, kCountsInts = 48, kCountsBytes = kCountsInts * 4;
static_assert(kFetchParts <= 16, "fetch counters overlap");

struct PTScene {
    DevBuf prims, mats, specs, knots, lights, lightArea, lnodes, leafOf, infinite, powerCdf, powerFunc, media, grid, sheen;
    int lightCount = 0, lnodeCount = 0, infCount = 0, powerCount = 0;
    float powerInt = 0;
    float worldR = 1;
    int hasInterfaces = 0, hasMedia = 0, hasLayered = 0;  // hasLayered: coated or sheen materials
    int sortable = 0;  // more than one material or medium: the material sort can group anything
    float cieYInt = 1;
    // host copies for inspection
    std::vector<LNode> hostNodes;
    // batch buffers
    size_t cap = 0;
    DevBuf o, d, rng, lam, lpdf, beta, rpath, L, Lh, prevP, prevN, st, hit, q0, q1, qm, qk, qms, sortBins, pfilm, xyz;
    DevBuf nk, qr, raySortBins;
    DevBuf h0, h1, h2, h3, qe, qeLpdf, qes;
    DevBuf sceneBuf;  // the DevScene the kernels read through a pointer
    DevBuf so, sd, sp, sn, sf, sL, sb, sq;  // shadow records and the shadow queue
    int *countsHost = nullptr;
    DevBuf counts;
    // The captured bounce rounds (see TraceBatch) and the key of what they were captured with.
    hipGraphExec_t bounceExec = nullptr;
    uint64_t bounceKey = 0;
    ~PTScene() {
        if (bounceExec) (void)hipGraphExecDestroy(bounceExec);
        if (countsHost) (void)hipHostFree(countsHost);
    }
};

Device::Device() = default;
Device::~Device() {
    pt.reset();
    for (hipEvent_t &e : pipeEvents)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}
void ReleasePT(Device &d) {
    d.pt.reset();
    d.guide.reset();
}

namespace {

uint64_t PTKey(const prgpu::SceneDescEx &sd, const float *sheenTable) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t n) {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    };
    mix(&sd.primCount, sizeof(int));
    if (sd.materialCount) mix(sd.materials, sizeof(prgpu::Material) * size_t(sd.materialCount));
    if (sd.spectrumCount) mix(sd.spectra, sizeof(prgpu::SpectrumEx) * size_t(sd.spectrumCount));
    if (sd.knotCount) mix(sd.knots, sizeof(float) * size_t(sd.knotCount));
    if (sd.lightCount) mix(sd.lights, sizeof(prgpu::LightEx) * size_t(sd.lightCount));
    if (sd.mediumCount) mix(sd.media, sizeof(prgpu::MediumEx) * size_t(sd.mediumCount));
    mix(&sd.gridCount, sizeof(sd.gridCount));
    if (sd.gridCount > 0) {
        int64_t step = std::max<int64_t>(1, sd.gridCount / 4096);
        for (int64_t i = 0; i < sd.gridCount; i += step) mix(&sd.gridData[i], sizeof(float));
    }
    mix(&sd.cameraMedium, sizeof(int));
    mix(&sd.camera, sizeof(sd.camera));
    if (sheenTable) mix(sheenTable, sizeof(float) * prgpu::kSheenAlpha * prgpu::kSheenMu);
    return h ^ GeometryKey(sd.prims, sd.primCount);
}

void Validate(const prgpu::SceneDescEx &sd) {
    auto bad = [](const char *m) { throw std::runtime_error(std::string("invalid scene: ") + m); };
    if (sd.primCount < 0 || (sd.primCount > 0 && !sd.prims)) bad("primitives");
    if (sd.camera.width <= 0 || sd.camera.height <= 0) bad("film size");
    for (int i = 0; i < sd.primCount; ++i) {
        const prgpu::Primitive &p = sd.prims[i];
        if (p.material >= sd.materialCount || p.light >= sd.lightCount || p.mediumInside >= sd.mediumCount ||
            p.mediumOutside >= sd.mediumCount)
            bad("primitive references out of range");
    }
    auto specOK = [&](int s, bool optional) { return (optional && s < 0) || (s >= 0 && s < sd.spectrumCount); };
    // An image texture record (kTexHeader) and its RGB-to-spectrum table in bounds.
    auto texOK = [&](int64_t o) {
        if (o < 0 || o + prgpu::kTexHeader > sd.knotCount) return false;
        const float *r = sd.knots + o;
        const int64_t w = int64_t(r[0]), h = int64_t(r[1]), tab = int64_t(r[11]);
        if (!(w >= 1 && h >= 1 && float(w) == r[0] && float(h) == r[1] && w * h <= (int64_t(1) << 26) &&
              (r[2] == 0 || r[2] == 1 || r[2] == 2) && (r[3] == 0 || r[3] == 1) &&
              o + prgpu::kTexHeader + 3 * w * h <= sd.knotCount && float(tab) == r[11] && tab >= 0 && tab < sd.knotCount))
            return false;
        const int64_t n = int64_t(sd.knots[tab]);
        return n >= 2 && n <= 256 && float(n) == sd.knots[tab] && tab + 1 + n + 3 * n * n * n * 3 <= sd.knotCount;
    };
    auto isImage = [&](int s) { return s >= 0 && s < sd.spectrumCount && sd.spectra[s].kind == prgpu::SpecImage; };
    for (int i = 0; i < sd.materialCount; ++i) {
        const prgpu::Material &m = sd.materials[i];
        // Image textures only in the albedo slot, and not inside a checker.
        if (isImage(m.albedoB) || isImage(m.eta) || isImage(m.k) || isImage(m.coatEta) || isImage(m.coatAlbedo) ||
            isImage(m.sheenColor) || (isImage(m.albedo) && m.albedoB >= 0 && m.checkerScale != 0))
            bad("image texture outside a material's albedo");
        if (m.shading != 0) {
            const int64_t o = int64_t(m.shading) - 1;
            bool ok = o + prgpu::kShadeHeader <= sd.knotCount;
            if (ok) {
                const float *r = sd.knots + o;
                ok = (r[0] == 0 || r[0] == 1) && (r[4] == 0 || r[4] == 1) && (r[3] == 0 || r[3] == 1) &&
                     (r[0] == 0 || (float(int64_t(r[1])) == r[1] && texOK(int64_t(r[1])))) &&
                     (r[4] == 0 || (float(int64_t(r[5])) == r[5] && texOK(int64_t(r[5])) &&
                                    (r[6] == -1 || r[6] == 0 || r[6] == 1 || r[6] == 2)));
            }
            if (!ok) bad("shading perturbation record");
        }
        if (m.type == prgpu::MatDiffuse && !specOK(m.albedo, false)) bad("diffuse reflectance");
        if (m.type == prgpu::MatDiffuse && m.albedoB >= sd.spectrumCount) bad("checker spectrum");
        if (m.type == prgpu::MatConductor && !((specOK(m.eta, false) && specOK(m.k, false)) || specOK(m.albedo, false)))
            bad("conductor spectra");
        if (m.type == prgpu::MatDielectric && !specOK(m.eta, false)) bad("dielectric IOR");
        if (m.type == prgpu::MatCoatedDiffuse && !(specOK(m.albedo, false) && m.albedoB < sd.spectrumCount))
            bad("coated diffuse reflectance");
        if (m.type == prgpu::MatCoatedConductor && !((specOK(m.eta, false) && specOK(m.k, false)) || specOK(m.albedo, false)))
            bad("coated conductor spectra");
        if (is_coated(m) && !(specOK(m.coatEta, false) && specOK(m.coatAlbedo, true))) bad("coat spectra");
        if (m.sheenWeight > 0 && !specOK(m.sheenColor, true)) bad("sheen colour");
        if (m.type > prgpu::MatCoatedConductor) bad("material type");
    }
    for (int i = 0; i < sd.spectrumCount; ++i) {
        const prgpu::SpectrumEx &s = sd.spectra[i];
        if (s.kind == prgpu::SpecPiecewise && (s.count < 1 || s.offset < 0 || s.offset + 2 * s.count > sd.knotCount))
            bad("piecewise spectrum knots");
        if (s.kind == prgpu::SpecTable95 && (s.offset < 0 || s.offset + prgpu::kSpectrumSamples > sd.knotCount))
            bad("tabulated spectrum");
        if (s.kind == prgpu::SpecSellmeier && (s.count < 0 || s.count > 3)) bad("Sellmeier terms");
        if (s.kind == prgpu::SpecImage && !texOK(s.offset)) bad("image texture record");
        if (s.kind > prgpu::SpecImage) bad("spectrum kind");
    }
    for (int i = 0; i < sd.lightCount; ++i) {
        const prgpu::LightEx &l = sd.lights[i];
        if (!specOK(l.spectrum, false) || isImage(l.spectrum)) bad("light spectrum");
        if (l.kind == prgpu::LightExArea && (l.primitive < 0 || l.primitive >= sd.primCount)) bad("area light primitive");
        if (l.kind > prgpu::LightExSun) bad("light kind");
        if (l.kind == prgpu::LightExSun && !(l.cosFalloffEnd > -1 && l.cosFalloffEnd < 1)) bad("sun disk");
        if (l.kind == prgpu::LightExGoniometric) {
            bool ok = l.profileOffset >= 0 && l.profileSize > prgpu::kIesHeader &&
                      int64_t(l.profileOffset) + l.profileSize <= sd.knotCount;
            if (ok) {
                const float *r = sd.knots + l.profileOffset;
                int nV = int(r[9]), nH = int(r[10]);
                ok = nV >= 1 && nH >= 1 && float(nV) == r[9] && float(nH) == r[10] &&
                     int64_t(prgpu::kIesHeader) + nV + nH + int64_t(nV) * nH == l.profileSize;
            }
            if (!ok) bad("goniometric light profile");
        }
        if (l.kind == prgpu::LightExSpot &&
            !(l.cosFalloffEnd >= -1 && l.cosFalloffEnd <= l.cosFalloffStart && l.cosFalloffStart <= 1))
            bad("spot light cone");
        if (l.kind == prgpu::LightExEnvironment && l.profileSize != 0) {  // image environment record
            bool ok = l.profileOffset >= 0 && l.profileSize > prgpu::kEnvHeader &&
                      int64_t(l.profileOffset) + l.profileSize <= sd.knotCount;
            if (ok) {
                const float *r = sd.knots + l.profileOffset;
                const int64_t w = int64_t(r[0]), h = int64_t(r[1]), d65 = int64_t(r[6]), tab = int64_t(r[7]);
                ok = w >= 1 && h >= 1 && float(w) == r[0] && float(h) == r[1] && w * h <= (int64_t(1) << 26) &&
                     int64_t(prgpu::kEnvHeader) + 3 * w * h + w * h + h * (w + 1) + h + (h + 1) == l.profileSize &&
                     float(d65) == r[6] && specOK(int(d65), false) && float(tab) == r[7] && tab >= 0 &&
                     tab < sd.knotCount;
                if (ok) {
                    const int64_t n = int64_t(sd.knots[tab]);
                    ok = n >= 2 && n <= 256 && float(n) == sd.knots[tab] &&
                         tab + 1 + n + 3 * n * n * n * 3 <= sd.knotCount;
                }
            }
            if (!ok) bad("environment map record");
        }
    }
    for (int i = 0; i < sd.mediumCount; ++i) {
        const prgpu::MediumEx &m = sd.media[i];
        if (!specOK(m.sigmaA, false) || !specOK(m.sigmaS, false) || isImage(m.sigmaA) || isImage(m.sigmaS))
            bad("medium spectra");
        if (m.kind == prgpu::MediumGrid &&
            (m.nx <= 0 || m.ny <= 0 || m.nz <= 0 || m.gridOffset < 0 ||
             m.gridOffset + int64_t(m.nx) * m.ny * m.nz > sd.gridCount))
            bad("grid medium data");
    }
    if (sd.cameraMedium >= sd.mediumCount) bad("camera medium");
}

// Uploads count elements, or one zeroed element when the array is empty (kernels never index it then).
template <class T> void UploadOrDummy(DevBuf &b, const T *src, int count, hipStream_t s) {
    if (count > 0 && src) {
        b.upload(src, size_t(count), s);
    } else {
        T zero{};
        std::memset(&zero, 0, sizeof(T));
        b.upload(&zero, 1, s);
        PG_CHECK(hipStreamSynchronize(s));
    }
}

// The sheen albedo table of the CPU path tracer, built once per process.
const std::vector<float> &HostSheenTable() {
    static const std::vector<float> t = [] {
        std::vector<float> v(size_t(prgpu::kSheenAlpha) * prgpu::kSheenMu);
        SheenAlbedoTable(v.data());
        return v;
    }();
    return t;
}

void EnsurePTScene(Device &dev, const prgpu::SceneDescEx &sd, const float *sheenTable) {
    // The traversal reads the device-resident BVH: it must be this scene's.
    if (!dev.nodes.p || dev.geomKey != GeometryKey(sd.prims, sd.primCount))
        throw std::runtime_error("internal error: the scene's geometry is not on the device");
    uint64_t key = PTKey(sd, sheenTable);
    if (dev.pt && dev.ptKey == key) return;
    Validate(sd);
    dev.guide.reset();  // a model learned for another scene
    auto pt = std::make_unique<PTScene>();
    hipStream_t s = dev.stream;
    UploadOrDummy(pt->prims, sd.prims, sd.primCount, s);
    UploadOrDummy(pt->mats, sd.materials, sd.materialCount, s);
    UploadOrDummy(pt->specs, sd.spectra, sd.spectrumCount, s);
    std::vector<float> knots(sd.knots, sd.knots + sd.knotCount);
    knots.push_back(0.f);
    pt->knots.upload(knots.data(), knots.size(), s);
    UploadOrDummy(pt->media, sd.media, sd.mediumCount, s);
    std::vector<float> grid(sd.gridData, sd.gridData + sd.gridCount);
    grid.push_back(0.f);
    pt->grid.upload(grid.data(), grid.size(), s);
    pt->cieYInt = host_cie_y_integral();
    pt->sheen.upload(sheenTable ? sheenTable : HostSheenTable().data(), size_t(prgpu::kSheenAlpha) * prgpu::kSheenMu, s);
    pt->hasLayered = 0;
    pt->sortable = (sd.materialCount > 1 || sd.mediumCount > 1) ? 1 : 0;
    for (int i = 0; i < sd.materialCount; ++i)
        pt->hasLayered |= (is_coated(sd.materials[i]) || has_sheen(sd.materials[i])) ? 1 : 0;

    // World sphere: the primitives' bounds joined with the camera position (src/scene/scene.cpp:9-27).
    Box3 b(V3(dev.sceneLo[0], dev.sceneLo[1], dev.sceneLo[2]), V3(dev.sceneHi[0], dev.sceneHi[1], dev.sceneHi[2]));
    if (sd.primCount == 0) b = Box3();
    const float *m = sd.camera.cameraToWorld;
    V3 camPos(m[0] * 0.f + m[1] * 0.f + m[2] * 0.f + m[3], m[4] * 0.f + m[5] * 0.f + m[6] * 0.f + m[7],
              m[8] * 0.f + m[9] * 0.f + m[10] * 0.f + m[11]);
    b = join(b, camPos);
    V3 wc = b.centroid();
    float wr = b.empty() ? 0.f : dist(wc, b.hi);
    pt->worldR = std::max(wr, 1e-3f);

    // Lights: areas, bounds, powers; light BVH and power distribution.
    SpecTable st{sd.spectra, sd.knots};
    std::vector<float> areas(size_t(std::max(1, sd.lightCount)), 0.f);
    std::vector<std::optional<HostLB>> lbs;
    std::vector<float> power;
    for (int i = 0; i < sd.lightCount; ++i) {
        const prgpu::LightEx &l = sd.lights[i];
        std::optional<HostLB> lb;
        float pw = 0;
        if (l.kind == prgpu::LightExArea) {
            const prgpu::Primitive &p = sd.prims[l.primitive];
            areas[size_t(i)] = PrimArea(p);
            HostLB x;
            x.bounds = PrimBounds(p);
            if (p.type == prgpu::PrimSphere) {
                x.w = V3(0, 0, 1);
                x.cosTheta_o = -1;
            } else {
                V3 n = normalize(cross_f(f3(p.p1) - f3(p.p0), f3(p.p2) - f3(p.p0)));
                if (prim_has_normals(p)) n = face_forward(n, f3(p.n0) + f3(p.n1) + f3(p.n2));
                x.w = n;
                x.cosTheta_o = 1;
            }
            x.cosTheta_e = 0;
            x.twoSided = l.twoSided != 0;
            x.phi = l.scale * spec_max(st, l.spectrum) * areas[size_t(i)] * kPi * (l.twoSided ? 2 : 1);
            lb = x;
            float lobe = 2 * kPi / (l.cosPower + 2);
            pw = l.scale * host_spectrum_y(st, l.spectrum, pt->cieYInt) * areas[size_t(i)] * lobe * (l.twoSided ? 2 : 1);
        } else if (l.kind == prgpu::LightExPoint) {
            HostLB x;
            x.bounds = Box3(f3(l.position), f3(l.position));
            x.phi = 4 * kPi * l.scale * spec_max(st, l.spectrum);
            x.cosTheta_o = -1;
            x.cosTheta_e = 0;
            lb = x;
            pw = 4 * kPi * l.scale * host_spectrum_y(st, l.spectrum, pt->cieYInt);
        } else if (l.kind == prgpu::LightExSpot) {  // light.cpp:117-125, 173-175
            HostLB x;
            x.bounds = Box3(f3(l.position), f3(l.position));
            x.w = f3(l.direction);
            x.phi = 4 * kPi * l.scale * spec_max(st, l.spectrum);
            x.cosTheta_o = l.cosFalloffStart;
            x.cosTheta_e = std::cos(safe_acos(l.cosFalloffEnd) - safe_acos(l.cosFalloffStart));
            lb = x;
            pw = l.scale * host_spectrum_y(st, l.spectrum, pt->cieYInt) * 2 * kPi *
                 ((1 - l.cosFalloffStart) + (l.cosFalloffStart - l.cosFalloffEnd) / 2);
        } else if (l.kind == prgpu::LightExGoniometric) {  // light.h:198-206
            HostLB x;
            x.bounds = Box3(f3(l.position), f3(l.position));
            x.phi = 4 * kPi * l.scale * spec_max(st, l.spectrum);
            x.cosTheta_o = -1;
            x.cosTheta_e = 0;
            lb = x;
            pw = l.scale * host_spectrum_y(st, l.spectrum, pt->cieYInt) * sd.knots[l.profileOffset + 11];
        } else if (l.kind == prgpu::LightExSun) {  // sky.cpp:88-90
            pw = l.scale * host_spectrum_y(st, l.spectrum, pt->cieYInt) * 2 * kPi * (1 - l.cosFalloffEnd) * kPi *
                 pt->worldR * pt->worldR;
        } else if (l.kind == prgpu::LightExDistant) {
            pw = l.scale * host_spectrum_y(st, l.spectrum, pt->cieYInt) * kPi * pt->worldR * pt->worldR;
        } else if (l.profileSize > 0) {  // image environment: its mean luminance (light.cpp:241)
            pw = 4 * kPi * kPi * pt->worldR * pt->worldR * l.scale * sd.knots[l.profileOffset + 5];
        } else {
            pw = 4 * kPi * kPi * pt->worldR * pt->worldR * l.scale * host_spectrum_y(st, l.spectrum, pt->cieYInt);
        }
        lbs.push_back(lb);
        power.push_back(std::isfinite(pw) && pw > 0 ? pw : 0.f);
    }
    LightTreeBuilder tb;
    std::vector<std::pair<int, HostLB>> items;
    for (size_t i = 0; i < lbs.size(); ++i) {
        if (!lbs[i]) {
            tb.infinite.push_back(int(i));
        } else if (lbs[i]->phi > 0) {
            items.push_back({int(tb.bounded.size()), *lbs[i]});
            tb.bounded.push_back(int(i));
        }
    }
    if (!items.empty()) tb.Rec(items, 0, int(items.size()), -1);
    std::vector<LNode> nodes(tb.lbs.size());
    std::vector<int> leafOf(size_t(std::max(1, sd.lightCount)), -1);
    for (size_t i = 0; i < nodes.size(); ++i) {
        LNode &n = nodes[i];
        const HostLB &x = tb.lbs[i];
        n.lo[0] = x.bounds.lo.x; n.lo[1] = x.bounds.lo.y; n.lo[2] = x.bounds.lo.z;
        n.hi[0] = x.bounds.hi.x; n.hi[1] = x.bounds.hi.y; n.hi[2] = x.bounds.hi.z;
        n.w[0] = x.w.x; n.w[1] = x.w.y; n.w[2] = x.w.z;
        n.phi = x.phi;
        n.cosTheta_o = x.cosTheta_o;
        n.cosTheta_e = x.cosTheta_e;
        n.child1 = tb.child1[i];
        n.light = tb.light[i];
        n.parent = tb.parent[i];
        n.twoSided = x.twoSided ? 1u : 0u;
        if (n.light >= 0) leafOf[size_t(n.light)] = int(i);
    }
    pt->hostNodes = nodes;
    pt->lnodeCount = int(nodes.size());
    if (nodes.empty()) nodes.push_back(LNode{});  // no bounded lights: a placeholder the kernels never read
    pt->lnodes.upload(nodes.data(), nodes.size(), s);
    pt->leafOf.upload(leafOf.data(), leafOf.size(), s);
    std::vector<int> inf = tb.infinite;
    pt->infCount = int(inf.size());
    inf.push_back(-1);
    pt->infinite.upload(inf.data(), inf.size(), s);
    pt->lightCount = sd.lightCount;
    UploadOrDummy(pt->lights, sd.lights, sd.lightCount, s);
    pt->lightArea.upload(areas.data(), areas.size(), s);
    // Power sampler: Distribution1D over max(power, 1e-4 max) (src/lights/light.cpp:326-340, sampling.cpp:5-27).
    {
        float maxP = 0;
        for (float p : power) maxP = std::max(maxP, p);
        for (float &p : power) p = std::max(p, maxP > 0 ? maxP * 1e-4f : 1.f);
        int n = int(power.size());
        std::vector<float> cdf(size_t(n + 1), 0.f);
        double acc = 0;
        for (int i = 1; i < n + 1; ++i) {
            acc += double(power[size_t(i - 1)]) / n;
            cdf[size_t(i)] = float(acc);
        }
        float funcInt = float(acc);
        if (funcInt == 0) {
            for (int i = 1; i < n + 1; ++i) cdf[size_t(i)] = float(i) / float(n);
        } else {
            for (int i = 1; i < n + 1; ++i) cdf[size_t(i)] = float(cdf[size_t(i)] / acc);
        }
        if (n > 0) cdf[size_t(n)] = 1;
        power.push_back(0.f);
        pt->powerCdf.upload(cdf.data(), cdf.size(), s);
        pt->powerFunc.upload(power.data(), power.size(), s);
        pt->powerInt = funcInt;
        pt->powerCount = n;
    }
    pt->hasInterfaces = 0;
    for (int i = 0; i < sd.primCount && !pt->hasInterfaces; ++i) {
        const prgpu::Primitive &p = sd.prims[i];
        bool noMat = p.material < 0 || sd.materials[p.material].type == prgpu::MatInterface ||
                     sd.materials[p.material].type == prgpu::MatBlack;
        if (noMat && p.light < 0) pt->hasInterfaces = 1;
    }
    pt->hasMedia = (sd.cameraMedium >= 0) ? 1 : 0;
    for (int i = 0; i < sd.primCount && !pt->hasMedia; ++i)
        if (sd.prims[i].mediumInside >= 0 || sd.prims[i].mediumOutside >= 0) pt->hasMedia = 1;
    PG_CHECK(hipHostMalloc(reinterpret_cast<void **>(&pt->countsHost), kCountsBytes, hipHostMallocDefault));
    pt->counts.ensure(kCountsBytes);
    PG_CHECK(hipMemsetAsync(pt->counts.p, 0, kCountsBytes, s));
    PG_CHECK(hipStreamSynchronize(s));
    dev.pt = std::move(pt);
    dev.ptKey = key;
}

float FilterEval1D(int type, float radius, float sigma, float x) {  // src/core/film.cpp:24-39
    x = std::abs(x);
    if (x > radius) return 0;
    if (type == prgpu::FilterBox) return 1;
    if (type == prgpu::FilterGaussian)
        return std::max(0.f, std::exp(-x * x / (2 * sigma * sigma)) - std::exp(-radius * radius / (2 * sigma * sigma)));
    float t = 0.5f + x / (2 * radius);
    const float a0 = 0.35875f, a1 = 0.48829f, a2 = 0.14128f, a3 = 0.01168f;
    return a0 - a1 * std::cos(2 * kPi * t) + a2 * std::cos(4 * kPi * t) - a3 * std::cos(6 * kPi * t);
}

DevScene MakeDevScene(Device &dev, const prgpu::SceneDescEx &sd, const prgpu::RenderParamsEx &rp) {
    PTScene &pt = *dev.pt;
    DevScene s{};
    s.prims = pt.prims.as<prgpu::Primitive>();
    s.mats = pt.mats.as<prgpu::Material>();
    s.spec = SpecTable{pt.specs.as<prgpu::SpectrumEx>(), pt.knots.as<float>()};
    s.lights = pt.lights.as<prgpu::LightEx>();
    s.lightArea = pt.lightArea.as<float>();
    s.lightCount = pt.lightCount;
    s.lnodes = pt.lnodes.as<LNode>();
    s.lnodeCount = pt.lnodeCount;
    s.leafOf = pt.leafOf.as<int>();
    s.infinite = pt.infinite.as<int>();
    s.infCount = pt.infCount;
    s.envCount = 0;
    s.emitters = 0;
    for (int i = 0; i < sd.lightCount; ++i) {
        const uint32_t k = sd.lights[i].kind;
        s.envCount += (k == prgpu::LightExEnvironment || k == prgpu::LightExSun) ? 1 : 0;
        s.emitters |= (k == prgpu::LightExArea || k == prgpu::LightExEnvironment || k == prgpu::LightExSun) ? 1 : 0;
    }
    s.powerCdf = pt.powerCdf.as<float>();
    s.powerFunc = pt.powerFunc.as<float>();
    s.powerInt = pt.powerInt;
    s.powerCount = pt.powerCount;
    s.lightSampler = rp.lightSampler;
    s.media = pt.media.as<prgpu::MediumEx>();
    s.grid = pt.grid.as<float>();
    s.cameraMedium = sd.cameraMedium;
    s.worldR = pt.worldR;
    s.hasInterfaces = pt.hasInterfaces;
    s.sheenTable = pt.sheen.as<float>();
    std::memcpy(s.c2w, sd.camera.cameraToWorld, sizeof(s.c2w));
    s.tanHalf = sd.camera.tanHalfFov;
    s.aspect = float(sd.camera.width) / float(sd.camera.height);
    s.lensR = std::max(0.f, sd.camera.lensRadius);
    s.focal = std::max(1e-4f, sd.camera.focalDistance);
    s.width = sd.camera.width;
    s.height = sd.camera.height;
    // Film filter and margin (src/core/film.cpp:9-22, 42).
    int ft = sd.camera.filterType;
    float r = sd.camera.filterRadius;
    float sigma = 0.5f;
    if (ft == prgpu::FilterBox) r = std::max(0.5f, std::min(r, 0.5f));
    if (ft == prgpu::FilterGaussian) sigma = r / 3.f;
    const int n = 4096;
    double s1 = 0;
    for (int i = 0; i < n; ++i) s1 += FilterEval1D(ft, r, sigma, -r + (i + 0.5f) * (2 * r / n));
    s1 *= 2 * r / n;
    s.filterType = ft;
    s.filterRadius = r;
    s.filterSigma = sigma;
    s.filterNorm = float(1.0 / (s1 * s1));
    s.margin = int(std::ceil(r - 0.5f));
    s.xyzNorm = 1.f / (4 * pt.cieYInt);
    s.trace = MakeTraceCtx(dev);
    return s;
}

// The kernels read the scene through a pointer into HBM: passing the struct by value and then
// referencing its fields would force a private copy of the kernel argument.
const DevScene *UploadDevScene(Device &dev, const DevScene &s) {
    dev.pt->sceneBuf.upload(&s, 1, dev.stream);
    return dev.pt->sceneBuf.as<DevScene>();
}

// =============================================================================================
// Kernels

struct Params {
    int maxDepth;  // already mapped: <= 0 -> INT_MAX
    int rrStart;
    int rr;
    float clampY;
    uint64_t seed;
    int firstSample;
    int guide, train;  // learned path guiding: sample with the model / collect training records
    float trainKeep;   // probability that a training record is kept (a uniform subsample over the image)
    GuideBox box;      // the box vertex positions are normalised to for the model
};

struct Bufs {
    float4 *o, *d;
    ulonglong2 *rng;
    float4 *lam, *lpdf, *beta, *rpath, *L, *Lh, *prevP, *prevN;
    int4 *st;
    float4 *hit;
    float4 *h0, *h1, *h2, *h3;  // surface hit records for the material stage
    int *q0, *q1;
    int *qm;     // material queues, QCount x qcap
    int *qk;     // with the material sort: each material-queue entry's sort key (its material, or for a
                 // medium scattering vertex its medium), written beside it; null without the sort
    int *qms;    // the material queues sorted by key, which the material stage then reads
    int *sortTotals, *sortCursor;  // per family and key bin: entries, then each bin's next position
    int *nk;     // with the ray sort: each next-queue entry's key (direction octant, origin cell); null without
    int *qr;     // the next queue sorted by that key, which the next bounce traces
    int *raySortTotals, *raySortCursor;
    int qcap;
    int *counts;  // [0] queue A, [1] queue B, [2] and [11] shadow queues (alternating bounces when the
                  // shadow rays are traced with the next bounce's closest hits), [3..5] plain
                  // material queues, [6] medium scattering vertices, [9] emission queue, [10]
                  // coated / sheen material queue,
                  // [16..16+P) closest-hit fetch counters, [32..32+P) shadow fetch counters
                  // (P = kFetchParts)
    float2 *pfilm;
    float4 *xyz;
    // shadow records, one per path slot (a path makes at most one shadow ray per bounce)
    float4 *so, *sd, *sp, *sn, *sf, *sL, *sb;
    // guiding: decoded mixture per diffuse-queue entry (12 float4), kGuideRecords training
    // records per path (record j: 6 float4 planes of qcap at 6 j qcap: {position, state},
    // {normal, p_bsdf}, wo, w, throughput, radiance snapshot), compacted training records (16
    // floats) and their counter
    int *qes;       // emission queue (path slots)
    int4 *qe;       // per slot: slot, depth and flags on arrival at the emitter, primitive (-1: escaped)
    float4 *qeLpdf; // and the wavelength pdfs on arrival
    int *sq;        // shadow queue (path slots; the shadow records are indexed by slot)
    int shadowIdx;  // index in counts of the shadow queue's length this bounce appends to
    const float4 *gout;
    float4 *trec;
    float *train;
    int *trainCount;
    int trainCap;    // records the batch may add
    int trainBase;   // ring position of the batch's first record
    int trainRing;   // ring capacity (records)
};

constexpr int kBlock = 256;
constexpr uint32_t kSpecularFlag = 1;

// Queue append aggregated per wave: the active lanes take consecutive entries with one atomic.
__device__ __forceinline__ int wave_append(int *counter) {
    const unsigned long long mask = __ballot(1);
    const int lane = __lane_id();
    const int leader = __ffsll(mask) - 1;
    const int prefix = __popcll(mask & ((1ull << lane) - 1ull));
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    return base + prefix;
}

// ---- queue appends aggregated per tile --------------------------------------------------------
// A device-scope atomic on one cache line completes about every 12 ns on MI355X, one after the
// other, so one atomic per wave and queue (a launch of 880k paths: ~14k waves, up to 6 queues)
// serializes into hundreds of microseconds (measured on configuration 2: the surface stage took
// 5.8 ms per step with per-wave appends, 1.6 ms with per-tile ones). The shading kernels therefore
// take their items in tiles of kAppendK items per thread: each item's queue entries (at most two)
// get a position within the tile from LDS counters, one atomic instruction (a lane per queue; the
// counters share a cache line) reserves every queue's range for the whole tile, and the entries
// are then scattered. Queue order does not affect any path's result. One item per thread is the
// fastest: larger tiles shrink the grid (material kernels 5.2, 6.6 and 9.3 ms per step at 1, 2
// and 4 items per thread).
#ifndef PG_APPEND_K
#define PG_APPEND_K 1
#endif
constexpr int kAppendK = PG_APPEND_K;
constexpr int kTile = kAppendK * kBlock;
static_assert(kTile <= 2048, "tile positions are 11 bits");
// Queue ids: the material families 0..4 (the last: medium scattering vertices), then the next
// bounce, emission and shadow queues.
enum QueueId : int { QNext = 5, QEmit = 6, QShadow = 7, QIds = 8 };
constexpr uint32_t kNoQueue = 15;

struct TileStage {
    int cnt[8];             // entries of the current tile per queue
    int base[8];            // each queue's reserved range for the tile
    int slot[kTile];        // the item's path slot
    uint32_t code[kTile];   // two entries of 15 bits: queue id << 11 | position in the tile's range
    int key[kTile];         // the material-sort key of the first entry (material queues only)
};

__device__ __forceinline__ void tile_begin(TileStage &t) {
    if (threadIdx.x < 8) t.cnt[threadIdx.x] = 0;
    __syncthreads();
}

// Position of this lane's entry among the tile's entries of queue q (q < 0: no entry). Every lane
// of the wave calls it; the loop runs once per distinct queue in the wave.
__device__ __forceinline__ int lds_append(int *cnt, int q) {
    int pos = 0;
    unsigned long long pending = __ballot(q >= 0);
    const int lane = __lane_id();
    while (pending) {
        const int leader = __ffsll(pending) - 1;
        const int qq = __shfl(q, leader);
        const unsigned long long m = __ballot(q == qq);
        int base = 0;
        if (lane == leader) base = atomicAdd(&cnt[qq], __popcll(m));
        base = __shfl(base, leader);
        if (q == qq) pos = base + __popcll(m & ((1ull << lane) - 1ull));
        pending &= ~m;
    }
    return pos;
}

__device__ __forceinline__ void tile_stage(TileStage &t, int j, int slot, int q1, int q2, int key = 0) {
    const int p1 = lds_append(t.cnt, q1), p2 = lds_append(t.cnt, q2);
    const uint32_t c1 = q1 < 0 ? (kNoQueue << 11) : (uint32_t(q1) << 11) | uint32_t(p1);
    const uint32_t c2 = q2 < 0 ? (kNoQueue << 11) : (uint32_t(q2) << 11) | uint32_t(p2);
    t.slot[j] = slot;
    t.code[j] = c1 | (c2 << 16);
    t.key[j] = key;
}

__device__ __forceinline__ int *queue_ptr(const Bufs &b, int *next, int q) {
    return q < QNext ? b.qm + q * b.qcap : (q == QNext ? next : (q == QEmit ? b.qes : b.sq));
}
// Index in Bufs::counts of material family q's queue length (the layered family's is 10, the
// medium scattering vertices' 6).
__device__ __host__ constexpr int mat_count_index(int q) { return q == 3 ? 10 : (q == 4 ? 6 : 3 + q); }
__device__ __forceinline__ int *count_ptr(const Bufs &b, int *nextCount, int q) {
    return q < QNext ? b.counts + mat_count_index(q) : (q == QNext ? nextCount : (q == QEmit ? b.counts + 9 : b.counts + b.shadowIdx));
}

// Reserves the tile's ranges (one atomic instruction) and scatters the staged entries.
__device__ __forceinline__ void tile_flush(TileStage &t, const Bufs &b, int *next, int *nextCount) {
    __syncthreads();
    if (threadIdx.x < QIds) {
        const int q = int(threadIdx.x), c = t.cnt[q];
        t.base[q] = c > 0 ? atomicAdd(count_ptr(b, nextCount, q), c) : 0;
        t.cnt[q] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kAppendK; ++k) {
        const int j = k * kBlock + int(threadIdx.x);
        const uint32_t code = t.code[j];
        const int slot = t.slot[j];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const uint32_t ce = (code >> (16 * e)) & 0x7fffu;
            const int q = int(ce >> 11);
            if (q != int(kNoQueue)) {
                const int pos = t.base[q] + int(ce & 0x7ffu);
                queue_ptr(b, next, q)[pos] = slot;
                // a material or next-queue entry (only an item's first) keeps its sort key beside it
                if (e == 0 && q < int(QNext) && b.qk) b.qk[size_t(q) * b.qcap + pos] = t.key[j];
                if (e == 0 && q == int(QNext) && b.nk) b.nk[pos] = t.key[j];
            }
        }
    }
}

__device__ __forceinline__ Pcg load_rng(const Bufs &b, int slot) {
    ulonglong2 r = b.rng[slot];
    Pcg p;
    p.state = r.x;
    p.inc = r.y;
    return p;
}
__device__ __forceinline__ V2 get2(Pcg &r) {
    float a = r.uniform();
    float c = r.uniform();
    return {a, c};
}

// Start of a bounce: every queue length and fetch counter but the current path queue's is zeroed
// (a kernel rather than a memset, so that the bounce can be captured in a graph).
__global__ __launch_bounds__(64) void k_reset_counts(int *counts, int keep, int keep2) {
    for (int i = int(threadIdx.x); i < kCountsInts; i += 64)
        if (i != keep && i != keep2) counts[i] = 0;
}

// Camera samples (src/integrators/path.cpp:61-70, src/cameras/camera.cpp:36-48).
__global__ __launch_bounds__(kBlock) void k_camera(const DevScene *__restrict__ sp, Params p, Bufs b, int n, int npix, int sw) {
    const DevScene &s = *sp;
    int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= n) return;
    int pass = slot / npix, pix = slot - pass * npix;
    int py = pix / sw, px = pix - py * sw;
    Pcg rng;
    uint64_t words[2] = {uint64_t(py) * uint64_t(sw) + uint64_t(px), p.seed};
    uint64_t si = uint64_t(p.firstSample + pass);
    rng.set_sequence(murmur_words(words, 2, false, 0), mix_bits(si * 0x9E3779B97F4A7C15ull + 1));
    Lambda lam = sample_visible(rng.uniform());
    V2 u = get2(rng);
    float fx = px - s.margin + u.x, fy = py - s.margin + u.y;
    V2 lens = get2(rng);
    float x = (2 * fx / s.width - 1) * s.tanHalf * s.aspect;
    float y = (1 - 2 * fy / s.height) * s.tanHalf;
    V3 dir = normalize(V3(x, y, 1));
    V3 o(0, 0, 0);
    if (s.lensR > 0) {
        V2 pl = disk_concentric(lens);
        pl.x = pl.x * s.lensR;
        pl.y = pl.y * s.lensR;
        float ft = s.focal / dir.z;
        V3 pFocus = dir * ft;
        o = V3(pl.x, pl.y, 0);
        dir = normalize(pFocus - o);
    }
    const float *m = s.c2w;
    V3 wo(m[0] * o.x + m[1] * o.y + m[2] * o.z + m[3], m[4] * o.x + m[5] * o.y + m[6] * o.z + m[7],
          m[8] * o.x + m[9] * o.y + m[10] * o.z + m[11]);
    V3 wd = normalize(V3(m[0] * dir.x + m[1] * dir.y + m[2] * dir.z, m[4] * dir.x + m[5] * dir.y + m[6] * dir.z,
                         m[8] * dir.x + m[9] * dir.y + m[10] * dir.z));
    b.o[slot] = make_float4(wo.x, wo.y, wo.z, 1.f);           // w: prevPdf
    b.d[slot] = make_float4(wd.x, wd.y, wd.z, 1.f);           // w: etaScale
    b.rng[slot] = make_ulonglong2(rng.state, rng.inc);
    b.lam[slot] = make_float4(lam.l[0], lam.l[1], lam.l[2], lam.l[3]);
    b.lpdf[slot] = make_float4(lam.pdf[0], lam.pdf[1], lam.pdf[2], lam.pdf[3]);
    b.beta[slot] = make_float4(1, 1, 1, 1);
    b.rpath[slot] = make_float4(1, 1, 1, 1);
    b.L[slot] = make_float4(0, 0, 0, 0);
    b.Lh[slot] = make_float4(0, 0, 0, 0);
    b.prevP[slot] = make_float4(0, 0, 0, 0);
    b.prevN[slot] = make_float4(0, 0, 0, 0);
    b.st[slot] = make_int4(0, 0, s.cameraMedium, 0);
    b.pfilm[slot] = make_float2(fx, fy);
    b.q0[slot] = slot;
    if (p.train)  // no guided vertex yet
        for (int j = 0; j < kGuideRecords; ++j) b.trec[6 * j * size_t(b.qcap) + slot] = make_float4(0, 0, 0, 0);
}

// Closest hits of the queued paths. Persistent: each lane takes a new ray from the queue front as
// soon as its previous one completes (fetch_rays_parts), so a wave's lanes stay busy while ray
// lengths differ and the rays in flight stay a narrow, coherent part of the queue.
template <bool kStats>
__global__ __launch_bounds__(kTraceBlock) PG_WAVES(8) void k_intersect(const DevScene *__restrict__ sp, Bufs b, const int *queue,
                                                           const int *count, int *fetch, unsigned long long *counters) {
    const DevScene &s = *sp;
    __shared__ uint32_t lds[kStackLds * kTraceBlock];
    const int n = *count;
    const int tid = int(blockIdx.x * blockDim.x + threadIdx.x);
    TraceStats st;
    uint32_t rays = 0;
    int idx = -1, slot = 0;
    bool started = false;
    DRay r{};
    TravState t{};
    FetchState fs = fetch_begin();
    while (fetch_rays_parts(fetch, n, idx, fs)) {
        if (idx >= 0 && !started) {
            slot = queue[idx];
            float4 o = b.o[slot], d = b.d[slot];
            r = DRay{o.x, o.y, o.z, d.x, d.y, d.z};
            trace_begin(t, r, kInfF);
            started = true;
        }
        if (!trace_turn(idx >= 0, t.cur)) continue;
        if (trace_step<false, kStats>(s.trace, r, t, lds, tid, &st)) {
            b.hit[slot] = t.found ? make_float4(t.hit.t, __int_as_float(t.hit.prim), t.hit.b1, t.hit.b2)
                                  : make_float4(0, __int_as_float(-1), 0, 0);
            ++rays;
            idx = -1;
            started = false;
        }
    }
    if (kStats && rays) {
        atomicAdd(&counters[0], (unsigned long long)rays);
        atomicAdd(&counters[1], (unsigned long long)st.nodes);
        atomicAdd(&counters[2], (unsigned long long)st.recs);
    }
}

struct PathRegs {
    V3 ro, rd;
    float prevPdf, etaScale;
    Pcg rng;
    Lambda lam;
    SS beta, rpath, L, Lh;
    V3 prevP, prevN;
    int depth, flags, medium, crossings;
};

__device__ __forceinline__ void add_contrib(PathRegs &r, const SS &c) {  // path.cpp:131-135
    r.Lh = r.Lh + c;
    float a = r.rpath.avg();
    if (a > 0 && is_finite(a)) r.L = r.L + c / a;
}

// Russian roulette (path.cpp:181-188, 228-235); true when the path is terminated.
__device__ __forceinline__ bool roulette(int rr, int rrStart, PathRegs &r) {
    if (rr && r.depth > rrStart) {
        SS eff;
        if (r.lam.terminated()) {
            eff = r.beta;
        } else {
            float a = r.rpath.avg();
            eff = a > 0 && is_finite(a) ? r.beta / a : SS(0.f);
        }
        float m = eff.maxc() * r.etaScale;
        if (m < 1) {
            float q = maxf(0.f, 1 - m);
            if (r.rng.uniform() < q) return true;
            r.beta = r.beta / (1 - q);
        }
    }
    return false;
}

// Next-event estimation up to the transmittance (path.cpp:82-111): samples a light and, when
// the unoccluded contribution is nonzero, writes the path's shadow record, which k_shadow
// completes; returns whether it did (the caller then appends the slot to the shadow queue).
template <bool kMedia, class B = BSDF>
__device__ __forceinline__ bool sample_ld(const DevScene &s, const Bufs &b, int slot, PathRegs &r, V3 p, V3 n, V3 wo, const B *bsdf,
                          float g, V3 ns, bool surface, int mIn, int mOut, const GuideMix *gm = nullptr) {
    float uLight = r.rng.uniform();
    V2 uL = get2(r.rng);
    float pmf;
    int li = light_sample(s, p, n, uLight, &pmf);
    if (li < 0 || pmf == 0) return false;
    LightSample ls;
    if (!sample_li(s, li, p, uL, r.lam, &ls) || !ls.L.nonzero() || ls.pdf == 0) return false;
    V3 wi = ls.wi;
    SS f;
    // A delta light takes no MIS weight: the scattering pdf is needed only for the others.
    const bool delta = light_is_delta(s, li);
    float scatterPdf = 0.f;
    if (bsdf) {
        f = bsdf->f(wo, wi) * absdot(wi, ns);
        if (!f.nonzero()) return false;
        if (!delta) {
            scatterPdf = bsdf->pdf(wo, wi);
            // A guided vertex scatters with the one-sample MIS proposal: that is the pdf to weigh against.
            if (gm) scatterPdf = gm->sel * guide_pdf(*gm, wi) + (1.f - gm->sel) * scatterPdf;
        }
    } else {
        float ph = hg_phase(dot(wo, wi), g);
        f = SS(ph);
        scatterPdf = ph;
    }
    float lightPdf = pmf * ls.pdf;
    float k = delta ? 1.f / lightPdf : power_heuristic(lightPdf, scatterPdf) / lightPdf;
    // Shadow ray p0.SpawnRayTo(pLight) (src/scene/interaction.h:43-47) and its starting medium.
    V3 o = surface ? offset_origin(p, n, ls.pl - p) : p;
    V3 target = offset_origin(ls.pl, ls.nl, o - ls.pl);
    V3 d = target - o;
    int medium = kMedia ? (surface ? (dot(d, n) > 0 ? mOut : mIn) : r.medium) : -1;
    const int rec = slot;
    b.so[rec] = make_float4(o.x, o.y, o.z, __int_as_float(medium));
    b.sd[rec] = make_float4(d.x, d.y, d.z, __int_as_float(slot));
    b.sp[rec] = make_float4(ls.pl.x, ls.pl.y, ls.pl.z, k);
    b.sn[rec] = make_float4(ls.nl.x, ls.nl.y, ls.nl.z, r.rpath.avg());
    st4(b.sf, rec, f);
    st4(b.sL, rec, ls.L);
    st4(b.sb, rec, r.beta);
    return true;
}

__device__ __forceinline__ void add_emitted(const DevScene &s, PathRegs &r, int li, const SS &Le, V3 dir) {  // path.cpp:143-153
    if (!Le.nonzero()) return;
    if (r.depth == 0 || (r.flags & kSpecularFlag)) {
        add_contrib(r, r.beta * Le);
    } else {
        float pmf = light_pmf(s, r.prevP, r.prevN, li);
        float lightPdf = pmf * pdf_li(s, li, r.prevP, dir);
        float w = power_heuristic(r.prevPdf, lightPdf);
        add_contrib(r, r.beta * Le * w);
    }
}

// Sort key of a ray for the traversal (tuning key ray_sort): its direction octant, then the Morton
// code of its origin's cell in an 8 x 8 x 8 grid over the scene's bounds (3 + 9 bits).
constexpr int kRaySortBins = 1 << 12;
__device__ __forceinline__ int ray_key(const Params &p, V3 o, V3 d) {
    auto cell = [&](float v, int a) {
        const float c = (v - p.box.lo[a]) * p.box.inv[a] * 8.f;
        return c >= 0.f ? (c < 8.f ? int(c) : 7) : 0;
    };
    const int cx = cell(o.x, 0), cy = cell(o.y, 1), cz = cell(o.z, 2);
    int m = 0;
    for (int k = 0; k < 3; ++k)
        m |= (((cx >> k) & 1) << (3 * k)) | (((cy >> k) & 1) << (3 * k + 1)) | (((cz >> k) & 1) << (3 * k + 2));
    return ((d.x < 0 ? 1 : 0) | (d.y < 0 ? 2 : 0) | (d.z < 0 ? 4 : 0)) << 9 | m;
}

// ---- stage 1 of a bounce: everything up to the BSDF (path.cpp:155-210) ----------------------------
// Medium distance sampling, escape and area emission, interfaces, the depth limit. Paths that
// continue at a surface are compacted into one queue per BSDF family together with their hit
// record, and paths that scatter in a medium into the phase-function family with the scattering
// point: the material stage runs both kinds of vertex (next-event estimation, then the direction
// sample), so this stage carries no light-sampling code and keeps its registers for the hit
// records' gathers. Paths that meet emission (an emitting primitive, or escape with an
// environment light) are appended to the emission queue, whose MIS-weighted contributions k_emit
// adds before the BSDF stage changes the throughput.
enum MatQueue : int { QDiffuse = 0, QConductor = 1, QDielectric = 2, QLayered = 3, QPhase = 4, QCount = 5 };
static_assert(int(QCount) == int(QNext), "the next-bounce queue id follows the material families");

// One path's surface stage; *q1 is the queue the path continues in (a material family, QNext:
// straight to the next bounce, -1: ended) and *q2 the emission or shadow queue (or -1).
template <bool kMedia, bool kInterfaces>
__device__ __forceinline__ void surface_vertex(const DevScene &s, const Params &p, const Bufs &b, int slot, int *q1, int *q2,
                                               int *key) {
    PathRegs r;
    float4 o = b.o[slot], d = b.d[slot];
    r.ro = V3(o.x, o.y, o.z);
    r.prevPdf = o.w;
    r.rd = V3(d.x, d.y, d.z);
    r.etaScale = d.w;
    float4 lp = b.lpdf[slot];
    const float4 lpIn = lp;
    r.lam.pdf[0] = lp.x, r.lam.pdf[1] = lp.y, r.lam.pdf[2] = lp.z, r.lam.pdf[3] = lp.w;
    int4 st = b.st[slot];
    r.depth = st.x;
    r.flags = st.y;
    r.medium = st.z;
    r.crossings = st.w;
    const bool lam0 = r.lam.terminated();
    // Training: the radiance gathered before this vertex closes the record of the guided
    // vertex the path just left (its own next-event estimate included).
    if (p.train)
        for (int j = 0; j < kGuideRecords; ++j) {
            const size_t base = 6 * j * size_t(b.qcap);
            float4 t0 = b.trec[base + slot];
            if (t0.w == 1.f) {
                b.trec[base + 5 * size_t(b.qcap) + slot] = b.Lh[slot];
                t0.w = 2.f;
                b.trec[base + slot] = t0;
            }
        }
    const float4 hv = b.hit[slot];
    const int prim = __float_as_int(hv.y);
    const bool hit = prim >= 0;
    const float tHit = hv.x;
    bool alive = true, toMaterial = false, done = false, emits = false;
    int mq = 0;

    if (kMedia && r.medium >= 0) {
        // Distance sampling (path.cpp:159-191): the throughput and the spectral-MIS ratios take
        // the sample's weights; a scattering event moves the path to the phase-function queue.
        float4 l = b.lam[slot];
        r.lam.l[0] = l.x, r.lam.l[1] = l.y, r.lam.l[2] = l.z, r.lam.l[3] = l.w;
        r.beta = ld4(b.beta, slot);
        r.rpath = ld4(b.rpath, slot);
        r.rng = load_rng(b, slot);
        DistSample ds = medium_sample(s, r.medium, r.ro, r.rd, hit ? tHit : kInfF, r.rng.uniform(), r.lam);
        r.beta = r.beta * ds.weight;
        r.rpath = r.rpath * ds.pdfRatio;
        if (!r.beta.nonzero()) {
            alive = false;
            done = true;
        } else if (ds.scattered) {
            done = true;
            if (r.depth++ >= p.maxDepth) {
                alive = false;
            } else {
                const V3 mp = r.ro + r.rd * ds.t;
                b.h0[slot] = make_float4(mp.x, mp.y, mp.z, 0.f);
                toMaterial = true;
                mq = QPhase;
                *key = r.medium;
            }
        }
        b.rng[slot] = make_ulonglong2(r.rng.state, r.rng.inc);
        st4(b.beta, slot, r.beta);
        st4(b.rpath, slot, r.rpath);
    }

    const int depthIn = r.depth, flagsIn = r.flags;
    if (!done) {
        if (!hit) {
            emits = s.envCount > 0;
            alive = false;
        } else {
            const prgpu::Primitive &pr = s.prims[prim];
            emits = pr.light >= 0;
            if (kInterfaces && is_interface(s, pr)) {
                if (++r.crossings > 1000) {
                    alive = false;
                } else {
                    Surf si;
                    interaction(pr, r.ro, r.rd, tHit, hv.z, hv.w, &si);
                    r.medium = dot(r.rd, si.n) > 0 ? pr.mediumOutside : pr.mediumInside;
                    r.ro = offset_origin(si.p, si.n, r.rd);
                }
            } else if (!has_bsdf(s, pr)) {
                alive = false;
            } else {
                const prgpu::Material &m = s.mats[pr.material];
                // A dispersive dielectric or coat keeps only the hero wavelength
                // (material.cpp:106-110, 120-124), also when the depth limit ends the path here.
                if (m.type == prgpu::MatDielectric && !spec_is_constant(s.spec, m.eta)) r.lam.terminate_secondary();
                if (is_coated(m) && !spec_is_constant(s.spec, m.coatEta)) r.lam.terminate_secondary();
                if (r.depth++ >= p.maxDepth) {
                    alive = false;
                } else {
                    Surf si;
                    interaction(pr, r.ro, r.rd, tHit, hv.z, hv.w, &si);
                    perturb_shading(s, m, &si);
                    toMaterial = true;
                    *key = pr.material;
                    mq = (is_coated(m) || has_sheen(m)) ? QLayered
                         : m.type == prgpu::MatDiffuse ? QDiffuse
                         : (m.type == prgpu::MatConductor ? QConductor : QDielectric);
                    uint32_t media = uint32_t(pr.mediumInside & 0xffff) | (uint32_t(pr.mediumOutside & 0xffff) << 16);
                    b.h0[slot] = make_float4(si.p.x, si.p.y, si.p.z, __int_as_float(pr.material));
                    b.h1[slot] = make_float4(si.n.x, si.n.y, si.n.z, si.u);
                    b.h2[slot] = make_float4(si.ns.x, si.ns.y, si.ns.z, si.v);
                    b.h3[slot] = make_float4(si.dpdus.x, si.dpdus.y, si.dpdus.z, __uint_as_float(media));
                }
            }
        }
    }
    if (emits) {
        // The emission is evaluated with the wavelength pdfs and depth the path arrived with.
        b.qe[slot] = make_int4(slot, depthIn, flagsIn, prim);
        b.qeLpdf[slot] = lpIn;
        *q2 = QEmit;
    }

    if (r.lam.terminated() != lam0)
        b.lpdf[slot] = make_float4(r.lam.pdf[0], r.lam.pdf[1], r.lam.pdf[2], r.lam.pdf[3]);
    b.st[slot] = make_int4(r.depth, r.flags, r.medium, r.crossings);
    if (kMedia || kInterfaces) {
        b.o[slot] = make_float4(r.ro.x, r.ro.y, r.ro.z, r.prevPdf);
        b.d[slot] = make_float4(r.rd.x, r.rd.y, r.rd.z, r.etaScale);
    }
    *q1 = toMaterial ? mq : (alive ? int(QNext) : -1);
    if (!toMaterial && alive && b.nk) *key = ray_key(p, r.ro, r.rd);
}

template <bool kMedia, bool kInterfaces>
__global__ __launch_bounds__(kBlock) PG_WAVES(kSurfaceWaves) void k_surface(const DevScene *__restrict__ sp, Params p, Bufs b, const int *queue,
                                                    const int *count, int *nextQueue, int *nextCount) {
    __shared__ TileStage ts;
    const DevScene &s = *sp;
    const int n = *count;
    tile_begin(ts);
    for (int t0 = int(blockIdx.x) * kTile; t0 < n; t0 += int(gridDim.x) * kTile) {
#pragma unroll 1
        for (int k = 0; k < kAppendK; ++k) {
            const int j = k * kBlock + int(threadIdx.x), i = t0 + j;
            int slot = 0, q1 = -1, q2 = -1, key = 0;
            if (i < n) {
                slot = queue[i];
                surface_vertex<kMedia, kInterfaces>(s, p, b, slot, &q1, &q2, &key);
            }
            tile_stage(ts, j, slot, q1, q2, key);
        }
        tile_flush(ts, b, nextQueue, nextCount);
    }
}

// Emission met by the paths of the emission queue, with its MIS weight (path.cpp:143-153,
// 159-167): area light of the hit primitive, or the environment for an escaped path.
// Emission met by one path of the emission queue, with its MIS weight.
__device__ __forceinline__ void emit_vertex(const DevScene &s, const Bufs &b, int slot) {
    const int4 e = b.qe[slot];
    const int prim = e.w;
    PathRegs r;
    float4 o = b.o[slot], d = b.d[slot], l = b.lam[slot], lp = b.qeLpdf[slot];
    r.ro = V3(o.x, o.y, o.z);
    r.prevPdf = o.w;
    r.rd = V3(d.x, d.y, d.z);
    r.lam = {{l.x, l.y, l.z, l.w}, {lp.x, lp.y, lp.z, lp.w}};
    r.depth = e.y;
    r.flags = e.z;
    r.beta = ld4(b.beta, slot);
    r.rpath = ld4(b.rpath, slot);
    r.L = ld4(b.L, slot);
    r.Lh = ld4(b.Lh, slot);
    float4 pp = b.prevP[slot], pn = b.prevN[slot];
    r.prevP = V3(pp.x, pp.y, pp.z);
    r.prevN = V3(pn.x, pn.y, pn.z);
    if (prim < 0) {
        for (int li = 0; li < s.infCount; ++li) {
            int L = s.infinite[li];
            if (s.lights[L].kind == prgpu::LightExEnvironment)
                add_emitted(s, r, L, env_L(s, s.lights[L], r.rd, r.lam), r.rd);
            else if (s.lights[L].kind == prgpu::LightExSun &&  // SunLight::Le (sky.cpp:92-95)
                     dot(normalize(r.rd), f3(s.lights[L].direction)) >= s.lights[L].cosFalloffEnd)
                add_emitted(s, r, L, spec_sample(s.spec, s.lights[L].spectrum, r.lam) * s.lights[L].scale, r.rd);
        }
    } else {
        const prgpu::Primitive &pr = s.prims[prim];
        const float4 hv = b.hit[slot];
        Surf si;
        interaction(pr, r.ro, r.rd, hv.x, hv.z, hv.w, &si);
        add_emitted(s, r, pr.light, area_L(s, s.lights[pr.light], si.n, -r.rd, r.lam), r.rd);
    }
    st4(b.L, slot, r.L);
    st4(b.Lh, slot, r.Lh);
}

__global__ __launch_bounds__(kBlock) void k_emit(const DevScene *__restrict__ sp, Bufs b) {
    const DevScene &s = *sp;
    const int n = b.counts[9];
    const int stride = int(gridDim.x * blockDim.x);
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += stride) emit_vertex(s, b, b.qes[i]);
}

// ---- stage 2: one BSDF family (path.cpp:208-235) ----------------------------------------------------
// Builds the BSDF, next-event estimation when it has a non-specular lobe, BSDF sampling,
// throughput update, spawn, Russian roulette.
// One path's BSDF stage (i: its position in the family's queue); *q1 is QNext when the path
// continues, *q2 QShadow when it made a shadow ray.
template <int kType, bool kMedia, bool kGuided>
__device__ __forceinline__ void material_vertex(const DevScene &s, const Params &p, const Bufs &b, int i, int slot, int *q1,
                                                int *q2, int *key) {
    PathRegs r;
    float4 d = b.d[slot];
    r.rd = V3(d.x, d.y, d.z);
    r.etaScale = d.w;
    r.rng = load_rng(b, slot);
    float4 l = b.lam[slot], lp = b.lpdf[slot];
    r.lam = {{l.x, l.y, l.z, l.w}, {lp.x, lp.y, lp.z, lp.w}};
    r.beta = ld4(b.beta, slot);
    r.rpath = ld4(b.rpath, slot);
    int4 st = b.st[slot];
    r.depth = st.x;
    r.flags = st.y;
    r.medium = st.z;
    float4 h0 = b.h0[slot], h1 = b.h1[slot], h2 = b.h2[slot], h3 = b.h3[slot];
    Surf si;
    si.p = V3(h0.x, h0.y, h0.z);
    si.n = V3(h1.x, h1.y, h1.z);
    si.ns = V3(h2.x, h2.y, h2.z);
    si.dpdus = V3(h3.x, h3.y, h3.z);
    si.u = h1.w;
    si.v = h2.w;
    si.wo = -normalize(r.rd);
    const prgpu::Material &m = s.mats[__float_as_int(h0.w)];
    uint32_t media = __float_as_uint(h3.w);
    int mIn = int(int16_t(media & 0xffff)), mOut = int(int16_t(media >> 16));
    auto bsdf = [&] {
        if constexpr (kType == QLayered) return eval_closure(s, m, si, r.lam);
        else return eval_surface<kType>(s, m, si, r.lam);
    }();
    bool alive = true;
    if constexpr (kGuided) {
        // One-sample MIS between the learned mixture and the BSDF (q = s p_guide + (1 - s) p_bsdf).
        const GuideMix gm = guide_mix(b.gout + size_t(i) * 12);
        const SS beta0 = r.beta;
        if (sample_ld<kMedia>(s, b, slot, r, si.p, si.n, si.wo, &bsdf, 0, si.ns, true, mIn, mOut, &gm)) *q2 = QShadow;
        const float usel = r.rng.uniform(), uc = r.rng.uniform();
        const V2 u2 = get2(r.rng);
        V3 wi;
        SS f;
        float pb;
        if (usel < gm.sel) {
            wi = guide_sample(gm, uc, u2.x, u2.y);
            f = bsdf.f(si.wo, wi);
            pb = bsdf.pdf(si.wo, wi);
        } else {
            BSample bs;
            if (bsdf.sample(si.wo, uc, u2, &bs)) {
                wi = bs.wi;
                f = bs.f;
                pb = bs.pdf;
            } else {
                f = SS(0.f);
                pb = 0.f;
            }
        }
        const float q = gm.sel * guide_pdf(gm, wi) + (1.f - gm.sel) * pb;
        if (!(q > 0.f) || !f.nonzero()) {
            alive = false;
        } else {
            r.beta = r.beta * (f * (absdot(wi, si.ns) / q));
            r.prevPdf = q;
            r.flags &= ~kSpecularFlag;
            r.ro = offset_origin(si.p, si.n, wi);
            r.rd = wi;
            if (kMedia) r.medium = dot(wi, si.n) > 0 ? mOut : mIn;
            if (p.train) {  // the path's first free record slot takes this guided vertex
                const size_t c = size_t(b.qcap);
                int j = 0;
                while (j < kGuideRecords && b.trec[6 * j * c + slot].w != 0.f) ++j;
                if (j < kGuideRecords) {
                    float4 *t = b.trec + 6 * j * c;
                    t[slot] = make_float4((si.p.x - p.box.lo[0]) * p.box.inv[0], (si.p.y - p.box.lo[1]) * p.box.inv[1],
                                          (si.p.z - p.box.lo[2]) * p.box.inv[2], 1.f);
                    t[c + slot] = make_float4(si.n.x, si.n.y, si.n.z, pb);
                    t[2 * c + slot] = make_float4(si.wo.x, si.wo.y, si.wo.z, 0.f);
                    t[3 * c + slot] = make_float4(wi.x, wi.y, wi.z, 0.f);
                    st4(t, 4 * c + slot, beta0);
                }
            }
            if (!r.beta.nonzero() || r.beta.has_nan()) alive = false;
            else if (roulette(p.rr, p.rrStart, r)) alive = false;
        }
    } else {
        if (bsdf.non_specular() && sample_ld<kMedia>(s, b, slot, r, si.p, si.n, si.wo, &bsdf, 0, si.ns, true, mIn, mOut))
            *q2 = QShadow;
        float uc = r.rng.uniform();
        V2 u2 = get2(r.rng);
        BSample bs;
        if (!bsdf.sample(si.wo, uc, u2, &bs)) {
            alive = false;
        } else {
            r.beta = r.beta * (bs.f * (absdot(bs.wi, si.ns) / bs.pdf));
            r.prevPdf = bs.pdf;
            // A stochastic closure's sampling pdf is only proportional: MIS uses its pdf (path.cpp:220).
            if constexpr (kType == QLayered)
                if (bs.proportional) r.prevPdf = bsdf.pdf(si.wo, bs.wi);
            r.flags = (bs.flags & kSpecular) ? (r.flags | kSpecularFlag) : (r.flags & ~kSpecularFlag);
            if (bs.flags & kTrans) r.etaScale *= sq(bs.eta);
            r.ro = offset_origin(si.p, si.n, bs.wi);
            r.rd = bs.wi;
            if (kMedia) r.medium = dot(bs.wi, si.n) > 0 ? mOut : mIn;
            if (!r.beta.nonzero() || r.beta.has_nan()) alive = false;
            else if (roulette(p.rr, p.rrStart, r)) alive = false;
        }
    }
    b.rng[slot] = make_ulonglong2(r.rng.state, r.rng.inc);
    if (alive) {
        b.o[slot] = make_float4(r.ro.x, r.ro.y, r.ro.z, r.prevPdf);
        b.d[slot] = make_float4(r.rd.x, r.rd.y, r.rd.z, r.etaScale);
        st4(b.beta, slot, r.beta);
        b.prevP[slot] = make_float4(si.p.x, si.p.y, si.p.z, 0.f);
        b.prevN[slot] = make_float4(si.n.x, si.n.y, si.n.z, 0.f);
        b.st[slot] = make_int4(r.depth, r.flags, r.medium, st.w);
        *q1 = QNext;
        if (b.nk) *key = ray_key(p, r.ro, r.rd);
    }
}

// A medium scattering vertex (path.cpp:166-190): next-event estimation with the Henyey-Greenstein
// phase function as the scattering density, then the phase-function direction sample and Russian
// roulette, drawing from the path's stream in that order. k_surface has applied the distance
// sample's weights, counted the vertex against the depth limit and left the scattering point in
// the hit record.
__device__ __forceinline__ void phase_vertex(const DevScene &s, const Params &p, const Bufs &b, int slot, int *q1, int *q2,
                                             int *key) {
    PathRegs r;
    const float4 d = b.d[slot];
    r.rd = V3(d.x, d.y, d.z);
    r.etaScale = d.w;
    r.rng = load_rng(b, slot);
    const float4 l = b.lam[slot], lp = b.lpdf[slot];
    r.lam = {{l.x, l.y, l.z, l.w}, {lp.x, lp.y, lp.z, lp.w}};
    r.beta = ld4(b.beta, slot);
    r.rpath = ld4(b.rpath, slot);
    const int4 st = b.st[slot];
    r.depth = st.x;
    r.flags = st.y;
    r.medium = st.z;
    const float4 h0 = b.h0[slot];
    const V3 mp(h0.x, h0.y, h0.z), mwo = -r.rd;
    const float g = clampf(s.media[r.medium].g, -0.99f, 0.99f);
    if (sample_ld<true>(s, b, slot, r, mp, V3(), mwo, static_cast<const BSDF *>(nullptr), g, V3(), false, -1, -1))
        *q2 = QShadow;
    float phasePdf;
    const V3 wi = sample_hg(mwo, g, get2(r.rng), &phasePdf);
    bool alive = phasePdf != 0;
    if (alive) {
        r.prevPdf = phasePdf;
        r.flags &= ~kSpecularFlag;
        r.ro = mp;
        r.rd = wi;
        if (roulette(p.rr, p.rrStart, r)) alive = false;
    }
    b.rng[slot] = make_ulonglong2(r.rng.state, r.rng.inc);
    if (alive) {
        b.o[slot] = make_float4(r.ro.x, r.ro.y, r.ro.z, r.prevPdf);
        b.d[slot] = make_float4(r.rd.x, r.rd.y, r.rd.z, r.etaScale);
        st4(b.beta, slot, r.beta);
        b.prevP[slot] = make_float4(mp.x, mp.y, mp.z, 0.f);
        b.prevN[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        b.st[slot] = make_int4(r.depth, r.flags, r.medium, st.w);
        *q1 = QNext;
        if (b.nk) *key = ray_key(p, r.ro, r.rd);
    }
}

template <int kType, bool kMedia, bool kGuided = false>
__global__ __launch_bounds__(kBlock) PG_WAVES(kType == QLayered ? 0 : kMaterialWaves) void k_material(const DevScene *__restrict__ sp, Params p, Bufs b, int *nextQueue,
                                                     int *nextCount) {
    __shared__ TileStage ts;
    const DevScene &s = *sp;
    const int n = b.counts[mat_count_index(kType)];
    const int *queue = b.qm + kType * b.qcap;
    tile_begin(ts);
    for (int t0 = int(blockIdx.x) * kTile; t0 < n; t0 += int(gridDim.x) * kTile) {
#pragma unroll 1
        for (int k = 0; k < kAppendK; ++k) {
            const int j = k * kBlock + int(threadIdx.x), i = t0 + j;
            int slot = 0, q1 = -1, q2 = -1, key = 0;
            if (i < n) {
                slot = queue[i];
                if constexpr (kType == QPhase) phase_vertex(s, p, b, slot, &q1, &q2, &key);
                else material_vertex<kType, kMedia, kGuided>(s, p, b, i, slot, &q1, &q2, &key);
            }
            tile_stage(ts, j, slot, q1, q2, key);
        }
        tile_flush(ts, b, nextQueue, nextCount);
    }
}

// The plain families (diffuse from family f0 on, conductor, dielectric and, in scenes with media,
// the medium scattering vertices) in one launch over their queues laid end to end: in late
// bounces, where each family holds a few thousand paths, separate launches cost a latency tail
// each. A wave meets at most the boundaries between the queues.
#ifndef PG_FUSED_MATERIAL
#define PG_FUSED_MATERIAL 1
#endif
constexpr bool kFusedMaterial = PG_FUSED_MATERIAL != 0;
template <bool kMedia>
__global__ __launch_bounds__(kBlock) PG_WAVES(kMaterialWaves) void k_material_plain(const DevScene *__restrict__ sp, Params p, Bufs b,
                                                                                 int *nextQueue, int *nextCount, int f0) {
    __shared__ TileStage ts;
    const DevScene &s = *sp;
    const int nD = f0 <= QDiffuse ? b.counts[mat_count_index(QDiffuse)] : 0;
    const int nC = b.counts[mat_count_index(QConductor)];
    const int nE = b.counts[mat_count_index(QDielectric)];
    const int nP = kMedia ? b.counts[mat_count_index(QPhase)] : 0;
    const int nDCE = nD + nC + nE, n = nDCE + nP;
    tile_begin(ts);
    for (int t0 = int(blockIdx.x) * kTile; t0 < n; t0 += int(gridDim.x) * kTile) {
#pragma unroll 1
        for (int k = 0; k < kAppendK; ++k) {
            const int j = k * kBlock + int(threadIdx.x), i = t0 + j;
            int slot = 0, q1 = -1, q2 = -1, key = 0;
            if (i < nD) {
                slot = b.qm[QDiffuse * b.qcap + i];
                material_vertex<QDiffuse, kMedia, false>(s, p, b, i, slot, &q1, &q2, &key);
            } else if (i < nD + nC) {
                const int ii = i - nD;
                slot = b.qm[QConductor * b.qcap + ii];
                material_vertex<QConductor, kMedia, false>(s, p, b, ii, slot, &q1, &q2, &key);
            } else if (i < nDCE) {
                const int ii = i - nD - nC;
                slot = b.qm[QDielectric * b.qcap + ii];
                material_vertex<QDielectric, kMedia, false>(s, p, b, ii, slot, &q1, &q2, &key);
            } else if (kMedia && i < n) {
                slot = b.qm[QPhase * b.qcap + (i - nDCE)];
                phase_vertex(s, p, b, slot, &q1, &q2, &key);
            }
            tile_stage(ts, j, slot, q1, q2, key);
        }
        tile_flush(ts, b, nextQueue, nextCount);
    }
}

// ---- the material queues sorted by material and medium ---------------------------------------
// Between the surface and material stages each material queue is sorted by its entries' keys (the
// material, or for the medium scattering vertices the medium), so that a wave of the material
// stage holds one material's paths: a material's branches (smooth or rough lobes, IOR models,
// spectrum kinds, the phase function of one medium) then run once per wave instead of once per
// material present in it. A counting sort: k_sort_count adds each block's key counts (per family
// and key bin; a key is folded into kSortBins bins) into totals, k_sort_scan turns the totals into
// each bin's first position within its family's queue, and k_sort_scatter, over the same slices as
// k_sort_count, reserves each block's range per bin with one atomic and places the entries. Every
// count is aggregated over a wave (one LDS atomic per distinct key in the wave) and over a block
// (one device atomic per bin the block holds). Order within a bin follows the atomics and does not
// affect any path's result.
constexpr int kSortBins = kBlock;
constexpr int kSortSlots = int(QCount) * kSortBins;

// The slice [*lo, *hi) of family f's queue that block `blk` of `nblk` sorts.
__device__ __forceinline__ void sort_slice(const Bufs &b, int f, int blk, int nblk, int *n, int *lo, int *hi) {
    *n = b.counts[mat_count_index(f)];
    const int per = (*n + nblk - 1) / nblk;
    *lo = min(*n, blk * per);
    *hi = min(*n, *lo + per);
}

__device__ __forceinline__ void sort_block_counts(const Bufs &b, int *hist) {
    for (int i = int(threadIdx.x); i < kSortSlots; i += kBlock) hist[i] = 0;
    __syncthreads();
    for (int f = 0; f < int(QCount); ++f) {
        int n, lo, hi;
        sort_slice(b, f, int(blockIdx.x), int(gridDim.x), &n, &lo, &hi);
        const int *keys = b.qk + size_t(f) * b.qcap;
        for (int i0 = lo; i0 < hi; i0 += kBlock) {
            const int i = i0 + int(threadIdx.x);
            const int bin = i < hi ? f * kSortBins + (keys[i] & (kSortBins - 1)) : -1;
            (void)lds_append(hist, bin);
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_sort_count(Bufs b) {
    __shared__ int hist[kSortSlots];
    sort_block_counts(b, hist);
    for (int i = int(threadIdx.x); i < kSortSlots; i += kBlock)
        if (hist[i]) atomicAdd(&b.sortTotals[i], hist[i]);
}

// One block: per family, the exclusive scan of its bins' totals gives each bin's first position;
// the totals are cleared for the next bounce.
__global__ __launch_bounds__(kBlock) void k_sort_scan(Bufs b) {
    static_assert(kSortBins == kBlock, "one thread per bin");
    __shared__ int sc[kSortBins];
    const int t = int(threadIdx.x);
    for (int f = 0; f < int(QCount); ++f) {
        const int i = f * kSortBins + t;
        const int v = b.sortTotals[i];
        sc[t] = v;
        __syncthreads();
        for (int off = 1; off < kSortBins; off <<= 1) {
            const int add = t >= off ? sc[t - off] : 0;
            __syncthreads();
            sc[t] += add;
            __syncthreads();
        }
        b.sortCursor[i] = sc[t] - v;
        b.sortTotals[i] = 0;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_sort_scatter(Bufs b) {
    __shared__ int hist[kSortSlots];
    sort_block_counts(b, hist);
    for (int i = int(threadIdx.x); i < kSortSlots; i += kBlock)
        if (hist[i]) hist[i] = atomicAdd(&b.sortCursor[i], hist[i]);
    __syncthreads();
    for (int f = 0; f < int(QCount); ++f) {
        int n, lo, hi;
        sort_slice(b, f, int(blockIdx.x), int(gridDim.x), &n, &lo, &hi);
        const size_t off = size_t(f) * b.qcap;
        for (int i0 = lo; i0 < hi; i0 += kBlock) {
            const int i = i0 + int(threadIdx.x);
            const int bin = i < hi ? f * kSortBins + (b.qk[off + i] & (kSortBins - 1)) : -1;
            const int pos = lds_append(hist, bin);
            if (bin >= 0 && pos < n) b.qms[off + pos] = b.qm[off + i];  // positions stay in the family's queue
        }
    }
}

// The same counting sort for the next bounce's path queue by ray key (tuning key ray_sort), so
// that a traversal wave takes rays of one direction octant from one region of the scene.
__device__ __forceinline__ void ray_sort_counts(const Bufs &b, int n, int *hist) {
    for (int i = int(threadIdx.x); i < kRaySortBins; i += kBlock) hist[i] = 0;
    __syncthreads();
    const int per = (n + int(gridDim.x) - 1) / int(gridDim.x);
    const int lo = min(n, int(blockIdx.x) * per), hi = min(n, lo + per);
    for (int i0 = lo; i0 < hi; i0 += kBlock) {
        const int i = i0 + int(threadIdx.x);
        (void)lds_append(hist, i < hi ? (b.nk[i] & (kRaySortBins - 1)) : -1);
    }
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_ray_sort_count(Bufs b, const int *count) {
    __shared__ int hist[kRaySortBins];
    ray_sort_counts(b, *count, hist);
    for (int i = int(threadIdx.x); i < kRaySortBins; i += kBlock)
        if (hist[i]) atomicAdd(&b.raySortTotals[i], hist[i]);
}

// One block: the exclusive scan of the bin totals (each thread sums a run of consecutive bins, the
// runs are scanned in LDS); the totals are cleared for the next bounce.
__global__ __launch_bounds__(kBlock) void k_ray_sort_scan(Bufs b) {
    constexpr int kRun = kRaySortBins / kBlock;
    __shared__ int sc[kBlock];
    const int t = int(threadIdx.x);
    int v[kRun], run = 0;
    for (int k = 0; k < kRun; ++k) {
        v[k] = b.raySortTotals[t * kRun + k];
        run += v[k];
    }
    sc[t] = run;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const int add = t >= off ? sc[t - off] : 0;
        __syncthreads();
        sc[t] += add;
        __syncthreads();
    }
    int pos = sc[t] - run;
    for (int k = 0; k < kRun; ++k) {
        b.raySortCursor[t * kRun + k] = pos;
        pos += v[k];
        b.raySortTotals[t * kRun + k] = 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_ray_sort_scatter(Bufs b, const int *count, const int *src) {
    __shared__ int hist[kRaySortBins];
    const int n = *count;
    ray_sort_counts(b, n, hist);
    for (int i = int(threadIdx.x); i < kRaySortBins; i += kBlock)
        if (hist[i]) hist[i] = atomicAdd(&b.raySortCursor[i], hist[i]);
    __syncthreads();
    const int per = (n + int(gridDim.x) - 1) / int(gridDim.x);
    const int lo = min(n, int(blockIdx.x) * per), hi = min(n, lo + per);
    for (int i0 = lo; i0 < hi; i0 += kBlock) {
        const int i = i0 + int(threadIdx.x);
        const int bin = i < hi ? (b.nk[i] & (kRaySortBins - 1)) : -1;
        const int pos = lds_append(hist, bin);
        if (bin >= 0 && pos < n) b.qr[pos] = src[i];  // positions stay in the queue
    }
}

// Shadow rays run from the offset origin to the offset light point, d = target - origin: the
// query's range is (0, 1 - 1e-4) in units of d (src/scene/interaction.h:43-47, scene.cpp:41-59).
constexpr float kShadowTMax = 1 - 1e-4f;

// The deferred add() of a visible shadow record: beta * (f * Tr * L * k), and the rPath
// average at sampling time for the spectral-MIS estimate.
__device__ __forceinline__ void shadow_add(const Bufs &b, int slot, const SS &T, float k, float a) {
    SS c = ld4(b.sb, slot) * (ld4(b.sf, slot) * T * ld4(b.sL, slot) * k);
    SS Lh = ld4(b.Lh, slot) + c;
    st4(b.Lh, slot, Lh);
    if (a > 0 && is_finite(a)) st4(b.L, slot, ld4(b.L, slot) + c / a);
}

// Closest hits of the current queue and, in the same persistent launch, the shadow rays of the
// previous bounce (scenes without media or interfaces, where a shadow ray is one any-hit query).
// The two are independent: a shadow ray only adds to its path's radiance, which this bounce's
// surface and emission stages read after the launch, so contributions are added in the same order
// as with a launch of their own and every sample's result is unchanged. A wave takes closest rays
// until that queue is drained, then shadow rays, so a bounce has one traversal tail instead of
// two; late bounces, where a few thousand rays take the whole launch latency, cost half as much.
#ifndef PG_FUSED_SHADOW
#define PG_FUSED_SHADOW 1
#endif
constexpr bool kFusedShadow = PG_FUSED_SHADOW != 0;
constexpr int kShadowCountB = 11;  // the second shadow queue length (see Bufs::counts)
__host__ __device__ constexpr int shadow_count_index(int curIdx) { return curIdx ? kShadowCountB : 2; }

template <bool kStats>
__global__ __launch_bounds__(kTraceBlock) PG_WAVES(8) void k_trace_bounce(const DevScene *__restrict__ sp, Bufs b,
                                                                        const int *queue, const int *count,
                                                                        const int *shadowCount, int *fetch,
                                                                        int *shadowFetch, unsigned long long *counters) {
    const DevScene &s = *sp;
    __shared__ uint32_t lds[kStackLds * kTraceBlock];
    const int n = *count, ns = *shadowCount;
    const int tid = int(blockIdx.x * blockDim.x + threadIdx.x);
    TraceStats stc, sts;
    uint32_t raysC = 0, raysS = 0;
    int idx = -1, slot = 0;
    bool shadow = false, started = false;
    DRay r{};
    TravState t{};
    FetchState fc = fetch_begin(), fs = fetch_begin();
    while (true) {
        bool more;
        if (!fc.done) {
            const bool idle = idx < 0;
            more = fetch_rays_parts(fetch, n, idx, fc);
            if (idle && idx >= 0) shadow = false;
        }
        if (fc.done) {
            const bool idle = idx < 0;
            more = fetch_rays_parts(shadowFetch, ns, idx, fs);
            if (idle && idx >= 0) shadow = true;
        }
        if (!more) break;
        if (idx >= 0 && !started) {
            if (shadow) {
                slot = b.sq[idx];
                const float4 so = b.so[slot], sd = b.sd[slot];
                r = DRay{so.x, so.y, so.z, sd.x, sd.y, sd.z};
                trace_begin(t, r, kShadowTMax);
                t.any = true;
            } else {
                slot = queue[idx];
                const float4 o = b.o[slot], d = b.d[slot];
                r = DRay{o.x, o.y, o.z, d.x, d.y, d.z};
                trace_begin(t, r, kInfF);
            }
            started = true;
        }
        if (!trace_turn(idx >= 0, t.cur)) continue;
        TraceStats one;
        const bool fin = trace_step<false, kStats, true>(s.trace, r, t, lds, tid, &one);
        if (kStats) {
            sts.nodes += shadow ? one.nodes : 0u;
            sts.recs += shadow ? one.recs : 0u;
            stc.nodes += shadow ? 0u : one.nodes;
            stc.recs += shadow ? 0u : one.recs;
        }
        if (fin) {
            if (shadow) {
                ++raysS;
                if (!t.found) {
                    const float4 spk = b.sp[slot], snk = b.sn[slot];
                    shadow_add(b, slot, SS(1.f), spk.w, snk.w);
                }
            } else {
                ++raysC;
                b.hit[slot] = t.found ? make_float4(t.hit.t, __int_as_float(t.hit.prim), t.hit.b1, t.hit.b2)
                                      : make_float4(0, __int_as_float(-1), 0, 0);
            }
            idx = -1;
            started = false;
        }
    }
    if (kStats && (raysC || raysS)) {
        atomicAdd(&counters[0], (unsigned long long)raysC);
        atomicAdd(&counters[1], (unsigned long long)stc.nodes);
        atomicAdd(&counters[2], (unsigned long long)stc.recs);
        atomicAdd(&counters[3], (unsigned long long)raysS);
        atomicAdd(&counters[4], (unsigned long long)sts.nodes);
        atomicAdd(&counters[5], (unsigned long long)sts.recs);
    }
}

// Transmittance of the shadow records (src/scene/scene.cpp:41-59) and the deferred add().
// Without media or interfaces a shadow ray is one any-hit query, traversed persistently as in
// k_intersect; otherwise each ray walks its interface crossings to completion.
// Transmittance of one shadow record (src/scene/scene.cpp:41-59): one any-hit query, or with
// interfaces a walk through their crossings, with the media's transmittance along the way; then
// the deferred add().
template <bool kMedia, bool kInterfaces, bool kStats>
__device__ __forceinline__ void shadow_walk(const DevScene &s, const Bufs &b, int slot, uint32_t *lds, int tid,
                                            TraceStats *ts) {
    float4 so = b.so[slot], sd = b.sd[slot], spt = b.sp[slot], sn = b.sn[slot];
    V3 ro(so.x, so.y, so.z), rd(sd.x, sd.y, sd.z), pl(spt.x, spt.y, spt.z), nl(sn.x, sn.y, sn.z);
    int medium = __float_as_int(so.w);
    Lambda lam;
    {
        float4 l = b.lam[slot], lp = b.lpdf[slot];
        lam = {{l.x, l.y, l.z, l.w}, {lp.x, lp.y, lp.z, lp.w}};
    }
    SS T(1.f);
    for (int iter = 0;; ++iter) {
        if (iter == 256) {
            T = SS(0.f);
            break;
        }
        DRay r{ro.x, ro.y, ro.z, rd.x, rd.y, rd.z};
        DHit h;
        if (!kInterfaces) {
            if (dev_trace<true, kStats>(s.trace, r, kShadowTMax, &h, lds, tid, ts)) {
                T = SS(0.f);
            } else if (kMedia && medium >= 0) {
                float l = len(rd);
                T = T * medium_tr(s, medium, ro, rd / l, 1.f * l, lam);
            }
            break;
        }
        bool hit = dev_trace<false, kStats>(s.trace, r, kShadowTMax, &h, lds, tid, ts);
        const prgpu::Primitive *pr = hit ? &s.prims[h.prim] : nullptr;
        if (hit && !is_interface(s, *pr)) {
            T = SS(0.f);
            break;
        }
        if (kMedia && medium >= 0) {
            float l = len(rd);
            T = T * medium_tr(s, medium, ro, rd / l, (hit ? h.t : 1.f) * l, lam);
        }
        if (!T.nonzero() || !hit) break;
        Surf si;
        interaction(*pr, ro, rd, h.t, h.b1, h.b2, &si);
        medium = dot(rd, si.n) > 0 ? pr->mediumOutside : pr->mediumInside;
        ro = offset_origin(si.p, si.n, pl - si.p);
        V3 target = offset_origin(pl, nl, ro - pl);
        rd = target - ro;
    }
    if (!T.nonzero()) return;
    shadow_add(b, slot, T, spt.w, sn.w);
}

template <bool kStats, bool kMedia, bool kInterfaces>
__global__ __launch_bounds__(kTraceBlock) void k_shadow(const DevScene *__restrict__ sp, Bufs b, const int *count,
                                                        int *fetch, unsigned long long *counters) {
    const DevScene &s = *sp;
    __shared__ uint32_t lds[kStackLds * kTraceBlock];
    const int n = *count;
    const int tid = int(blockIdx.x * blockDim.x + threadIdx.x);
    TraceStats ts;
    uint32_t rays = 0;
    if constexpr (!kMedia && !kInterfaces) {
        int idx = -1, slot = 0;
        bool started = false;
        DRay r{};
        TravState t{};
        FetchState fs = fetch_begin();
        while (fetch_rays_parts(fetch, n, idx, fs)) {
            if (idx >= 0 && !started) {
                slot = b.sq[idx];
                float4 so = b.so[slot], sd = b.sd[slot];
                r = DRay{so.x, so.y, so.z, sd.x, sd.y, sd.z};
                trace_begin(t, r, kShadowTMax);
                started = true;
            }
            if (!trace_turn(idx >= 0, t.cur)) continue;
            if (trace_step<true, kStats>(s.trace, r, t, lds, tid, &ts)) {
                ++rays;
                if (!t.found) {
                    float4 spk = b.sp[slot], snk = b.sn[slot];
                    shadow_add(b, slot, SS(1.f), spk.w, snk.w);
                }
                idx = -1;
                started = false;
            }
        }
    } else {
        const int stride = int(gridDim.x * blockDim.x);
        for (int i = tid; i < n; i += stride) {
            ++rays;
            shadow_walk<kMedia, kInterfaces, kStats>(s, b, b.sq[i], lds, tid, &ts);
        }
    }
    if (kStats && rays) {
        atomicAdd(&counters[3], (unsigned long long)rays);
        atomicAdd(&counters[4], (unsigned long long)ts.nodes);
        atomicAdd(&counters[5], (unsigned long long)ts.recs);
    }
}

// Final radiance per sample: film XYZ (path.cpp:71-73, film.cpp:133-139) or a parity record.
__global__ __launch_bounds__(kBlock) void k_finish(const DevScene *__restrict__ sp, Params p, Bufs b, int n,
                                                   prgpu::SampleRecordEx *rec) {
    const DevScene &s = *sp;
    int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= n) return;
    float4 lp = b.lpdf[slot], l = b.lam[slot];
    Lambda lam = {{l.x, l.y, l.z, l.w}, {lp.x, lp.y, lp.z, lp.w}};
    SS L = lam.terminated() ? ld4(b.Lh, slot) : ld4(b.L, slot);
    L = L * 1.f;  // camera ray weight
    if (p.train)
        for (int r = 0; r < kGuideRecords; ++r) {
            // Training weight of a guided vertex of the path: the radiance gathered beyond it divided
            // by the throughput that reached it, f L_i / q in the KL gradient (averaged over the hero
            // wavelengths).
            const size_t c = size_t(b.qcap);
            const float4 *t = b.trec + 6 * r * c;
            if (t[slot].w != 2.f) continue;
            const SS Lh = ld4(b.Lh, slot), snap = ld4(t, 5 * c + slot), beta0 = ld4(t, 4 * c + slot);
            float w = 0.f;
            int k = 0;
            for (int j = 0; j < 4; ++j)
                if (beta0.v[j] > 0.f && lam.pdf[j] > 0.f) {
                    w += (Lh.v[j] - snap.v[j]) / beta0.v[j];
                    ++k;
                }
            w = k ? w / float(k) : 0.f;
            if (w > 0.f && is_finite(w)) {
                // Every record is counted (trainCount[1]); a uniform subsample of probability
                // trainKeep, decided by a hash of the sample, is written, so that the records a
                // batch trains on cover the whole image rather than its first rows.
                wave_append(b.trainCount + 1);
                const uint64_t h = mix_bits((uint64_t(uint32_t(p.firstSample)) << 36) ^ (uint64_t(slot) << 2) ^ uint64_t(r));
                const float u = float(uint32_t(h >> 40)) * 0x1p-24f;
                if (!(u < p.trainKeep)) continue;
                const int at = wave_append(b.trainCount);
                if (at < b.trainCap) {
                    float4 a = t[slot], n1 = t[c + slot], wo = t[2 * c + slot], wi = t[3 * c + slot];
                    const size_t ring = size_t((int64_t(b.trainBase) + at) % b.trainRing);
                    float4 *dst = reinterpret_cast<float4 *>(b.train + ring * 16);
                    dst[0] = make_float4(a.x, a.y, a.z, n1.x);
                    dst[1] = make_float4(n1.y, n1.z, wo.x, wo.y);
                    dst[2] = make_float4(wo.z, wi.x, wi.y, wi.z);
                    dst[3] = make_float4(n1.w, w, 0.f, 0.f);
                }
            }
        }
    float2 pf = b.pfilm[slot];
    if (rec) {
        prgpu::SampleRecordEx &o = rec[slot];
        for (int k = 0; k < 4; ++k) {
            o.L[k] = L.v[k];
            o.lambda[k] = lam.l[k];
            o.pdf[k] = lam.pdf[k];
        }
        o.pFilmX = pf.x;
        o.pFilmY = pf.y;
        o.weight = 1.f;
        o.flags = L.has_nan() ? 1u : 0u;
        return;
    }
    float4 out = make_float4(0, 0, 0, 0);
    if (!L.has_nan() && L.nonzero()) {
        double X = 0, Y = 0, Z = 0;  // ToXYZ (src/core/spectrum.cpp:68-80)
        for (int k = 0; k < 4; ++k) {
            if (lam.pdf[k] == 0) continue;
            float v = L.v[k] / lam.pdf[k];
            X += cie_x(lam.l[k]) * v;
            Y += cie_y(lam.l[k]) * v;
            Z += cie_z(lam.l[k]) * v;
        }
        float x = float(X) * s.xyzNorm, y = float(Y) * s.xyzNorm, z = float(Z) * s.xyzNorm;
        if (p.clampY > 0 && y > p.clampY) {
            float kk = p.clampY / y;
            x *= kk;
            y *= kk;
            z *= kk;
        }
        if (is_finite(x) && is_finite(y) && is_finite(z)) out = make_float4(x, y, z, 1.f);
    }
    b.xyz[slot] = out;
}

// The reconstruction filter's 1D profile with its per-launch constants hoisted.
struct FilterK {
    int type;
    float r, invDiam, inv2s2, tail;  // radius, 1 / (2 r), 1 / (2 sigma^2), the Gaussian's value at r
};
__device__ __forceinline__ FilterK filter_k(const DevScene &s) {
    FilterK k;
    k.type = s.filterType;
    k.r = s.filterRadius;
    k.invDiam = 1 / (2 * s.filterRadius);
    k.inv2s2 = 1 / (2 * s.filterSigma * s.filterSigma);
    k.tail = expf(-s.filterRadius * s.filterRadius * k.inv2s2);
    return k;
}
__device__ __forceinline__ float filter_1d(const FilterK &k, float x) {
    x = fabsf(x);
    if (x > k.r) return 0;
    if (k.type == prgpu::FilterBox) return 1;
    if (k.type == prgpu::FilterGaussian) return maxf(0.f, expf(-x * x * k.inv2s2) - k.tail);
    // Blackman-Harris, a0 - a1 cos(2 pi t) + a2 cos(4 pi t) - a3 cos(6 pi t) (film.cpp:9-39), with
    // one cosine: v_cos_f32 takes its argument in revolutions, and cos 2x = 2c^2 - 1, cos 3x =
    // 4c^3 - 3c. Every pixel evaluates the profile for each of its neighbours' samples in every
    // pass, so three library cosines made the gather VALU-bound.
    const float t = 0.5f + x * k.invDiam;
    const float a0 = 0.35875f, a1 = 0.48829f, a2 = 0.14128f, a3 = 0.01168f;
    const float c = __builtin_amdgcn_cosf(t), c2 = c * c;
    return a0 - a1 * c + a2 * (2 * c2 - 1) - a3 * (4 * c2 - 3) * c;
}

// Film: each output pixel gathers the samples whose filter footprint covers it (the CPU film's
// AddXYZ splat, film.cpp:141-157, evaluated per pixel so no atomics are needed), one pass at a
// time, and adds the pass value and its square to the FP64 accumulators.
__global__ __launch_bounds__(kBlock) void k_gather(const DevScene *__restrict__ sp, Bufs b, int passes, int npix, int sw,
                                                   int reach, double *sum, double *sumsq) {
    const DevScene &s = *sp;
    int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= s.width * s.height) return;
    int y = pix / s.width, x = pix - y * s.width;
    const int sh = npix / sw;
    const float r = s.filterRadius;
    const FilterK fk = filter_k(s);
    double accX = 0, accY = 0, accZ = 0, sqX = 0, sqY = 0, sqZ = 0;
    for (int pass = 0; pass < passes; ++pass) {
        double eX = 0, eY = 0, eZ = 0;
        for (int sy = y + s.margin - reach; sy <= y + s.margin + reach; ++sy) {
            if (sy < 0 || sy >= sh) continue;
            for (int sx = x + s.margin - reach; sx <= x + s.margin + reach; ++sx) {
                if (sx < 0 || sx >= sw) continue;
                int slot = pass * npix + sy * sw + sx;
                float4 v = b.xyz[slot];
                if (v.w == 0.f) continue;
                float2 pf = b.pfilm[slot];
                int x0 = max(0, int(ceilf(pf.x - r - 0.5f))), x1 = min(s.width - 1, int(floorf(pf.x + r - 0.5f)));
                int y0 = max(0, int(ceilf(pf.y - r - 0.5f))), y1 = min(s.height - 1, int(floorf(pf.y + r - 0.5f)));
                if (x < x0 || x > x1 || y < y0 || y > y1) continue;
                float w = filter_1d(fk, pf.x - (x + 0.5f)) * filter_1d(fk, pf.y - (y + 0.5f)) * s.filterNorm;
                if (w == 0) continue;
                eX += double(v.x) * w;
                eY += double(v.y) * w;
                eZ += double(v.z) * w;
            }
        }
        accX += eX;
        accY += eY;
        accZ += eZ;
        sqX += eX * eX;
        sqY += eY * eY;
        sqZ += eZ * eZ;
    }
    size_t idx = 3 * size_t(pix);
    sum[idx] += accX;
    sum[idx + 1] += accY;
    sum[idx + 2] += accZ;
    if (sumsq) {
        sumsq[idx] += sqX;
        sumsq[idx + 1] += sqY;
        sumsq[idx + 2] += sqZ;
    }
}

// The same gather over 16 x 16 pixel tiles: per pass the tile's samples and their apron (reach
// samples on each side) are staged in LDS once, each with its value and its separable weights
// for the 2 reach + 1 pixel columns and rows it can reach (zero outside its footprint and for an
// empty sample), and every pixel then sums its neighbours in the order and with the products of
// k_gather, so the film is the same bit for bit. A sample is loaded and its filter evaluated once
// per pass instead of once per pixel that reaches it: (2 reach + 1)^2 times fewer loads and
// 2 reach + 1 times fewer filter evaluations.
constexpr int kGatherTile = 16, kGatherMaxReach = 3;
constexpr int kGatherApron = kGatherTile + 2 * kGatherMaxReach, kGatherTaps = 2 * kGatherMaxReach + 1;
static_assert(kGatherTile * kGatherTile == kBlock, "one thread per tile pixel");

__global__ __launch_bounds__(kBlock) void k_gather_tiled(const DevScene *__restrict__ sp, Bufs b, int passes, int npix,
                                                         int sw, int reach, double *sum, double *sumsq) {
    const DevScene &s = *sp;
    __shared__ float sv[3][kGatherApron * kGatherApron];
    __shared__ float swx[kGatherTaps][kGatherApron * kGatherApron], swy[kGatherTaps][kGatherApron * kGatherApron];
    const int tx = int(threadIdx.x) % kGatherTile, ty = int(threadIdx.x) / kGatherTile;
    const int X0 = int(blockIdx.x) * kGatherTile, Y0 = int(blockIdx.y) * kGatherTile;
    const int x = X0 + tx, y = Y0 + ty;
    const bool inside = x < s.width && y < s.height;
    const int sh = npix / sw, side = kGatherTile + 2 * reach, taps = 2 * reach + 1;
    const float r = s.filterRadius;
    const FilterK fk = filter_k(s);
    double accX = 0, accY = 0, accZ = 0, sqX = 0, sqY = 0, sqZ = 0;
    for (int pass = 0; pass < passes; ++pass) {
        __syncthreads();  // the previous pass's sums are done with the tile
        for (int k = int(threadIdx.x); k < side * side; k += kBlock) {
            const int ay = k / side, ax = k - ay * side;
            const int sx = X0 + s.margin - reach + ax, sy = Y0 + s.margin - reach + ay;
            float v[3] = {0.f, 0.f, 0.f};
            bool live = false;
            float2 pf = make_float2(0.f, 0.f);
            int x0 = 0, x1 = -1, y0 = 0, y1 = -1;
            if (sx >= 0 && sx < sw && sy >= 0 && sy < sh) {
                const int slot = pass * npix + sy * sw + sx;
                const float4 vv = b.xyz[slot];
                if (vv.w != 0.f) {
                    live = true;
                    v[0] = vv.x, v[1] = vv.y, v[2] = vv.z;
                    pf = b.pfilm[slot];
                    x0 = max(0, int(ceilf(pf.x - r - 0.5f)));
                    x1 = min(s.width - 1, int(floorf(pf.x + r - 0.5f)));
                    y0 = max(0, int(ceilf(pf.y - r - 0.5f)));
                    y1 = min(s.height - 1, int(floorf(pf.y + r - 0.5f)));
                }
            }
            for (int c = 0; c < 3; ++c) sv[c][k] = v[c];
            // tap e: the pixel e - reach columns (rows) from this sample's own
            for (int e = 0; e < taps; ++e) {
                const int px = sx - s.margin + e - reach, py = sy - s.margin + e - reach;
                swx[e][k] = live && px >= x0 && px <= x1 ? filter_1d(fk, pf.x - (px + 0.5f)) : 0.f;
                swy[e][k] = live && py >= y0 && py <= y1 ? filter_1d(fk, pf.y - (py + 0.5f)) : 0.f;
            }
        }
        __syncthreads();
        if (!inside) continue;
        double eX = 0, eY = 0, eZ = 0;
        for (int oy = -reach; oy <= reach; ++oy) {
            for (int ox = -reach; ox <= reach; ++ox) {
                const int k = (ty + reach + oy) * side + (tx + reach + ox);
                const float w = swx[reach - ox][k] * swy[reach - oy][k] * s.filterNorm;
                if (w == 0) continue;
                eX += double(sv[0][k]) * w;
                eY += double(sv[1][k]) * w;
                eZ += double(sv[2][k]) * w;
            }
        }
        accX += eX;
        accY += eY;
        accZ += eZ;
        sqX += eX * eX;
        sqY += eY * eY;
        sqZ += eZ * eZ;
    }
    if (!inside) return;
    const size_t idx = 3 * (size_t(y) * size_t(s.width) + size_t(x));
    sum[idx] += accX;
    sum[idx + 1] += accY;
    sum[idx + 2] += accZ;
    if (sumsq) {
        sumsq[idx] += sqX;
        sumsq[idx + 1] += sqY;
        sumsq[idx + 2] += sqZ;
    }
}

// =============================================================================================
// Host driver

void EnsureBatch(PTScene &pt, size_t n) {
    if (pt.cap >= n) return;
    auto f4 = [&](DevBuf &b) { b.ensure(n * sizeof(float4)); };
    f4(pt.o); f4(pt.d); f4(pt.lam); f4(pt.lpdf); f4(pt.beta); f4(pt.rpath); f4(pt.L); f4(pt.Lh);
    f4(pt.prevP); f4(pt.prevN); f4(pt.hit); f4(pt.xyz);
    f4(pt.h0); f4(pt.h1); f4(pt.h2); f4(pt.h3); f4(pt.qe); f4(pt.qeLpdf);
    pt.qes.ensure(n * sizeof(int));
    pt.sq.ensure(n * sizeof(int));
    pt.qm.ensure(n * QCount * sizeof(int));
    pt.qk.ensure(n * QCount * sizeof(int));
    pt.qms.ensure(n * QCount * sizeof(int));
    if (!pt.sortBins.p) {  // the totals start at zero; k_sort_scan clears them after each use
        pt.sortBins.ensure(2 * kSortSlots * sizeof(int));
        PG_CHECK(hipMemset(pt.sortBins.p, 0, pt.sortBins.bytes));
    }
    pt.nk.ensure(n * sizeof(int));
    pt.qr.ensure(n * sizeof(int));
    if (!pt.raySortBins.p) {  // likewise for k_ray_sort_scan
        pt.raySortBins.ensure(2 * kRaySortBins * sizeof(int));
        PG_CHECK(hipMemset(pt.raySortBins.p, 0, pt.raySortBins.bytes));
    }
    f4(pt.so); f4(pt.sd); f4(pt.sp); f4(pt.sn); f4(pt.sf); f4(pt.sL); f4(pt.sb);
    pt.rng.ensure(n * sizeof(ulonglong2));
    pt.st.ensure(n * sizeof(int4));
    pt.q0.ensure(n * sizeof(int));
    pt.q1.ensure(n * sizeof(int));
    pt.pfilm.ensure(n * sizeof(float2));
    pt.cap = n;
}

Bufs MakeBufs(PTScene &pt) {
    Bufs b{};  // zeroed, padding included: the bounce-graph key hashes its bytes
    b.o = pt.o.as<float4>(); b.d = pt.d.as<float4>(); b.rng = pt.rng.as<ulonglong2>();
    b.lam = pt.lam.as<float4>(); b.lpdf = pt.lpdf.as<float4>(); b.beta = pt.beta.as<float4>();
    b.rpath = pt.rpath.as<float4>(); b.L = pt.L.as<float4>(); b.Lh = pt.Lh.as<float4>();
    b.prevP = pt.prevP.as<float4>(); b.prevN = pt.prevN.as<float4>(); b.st = pt.st.as<int4>();
    b.hit = pt.hit.as<float4>(); b.q0 = pt.q0.as<int>(); b.q1 = pt.q1.as<int>(); b.counts = pt.counts.as<int>();
    b.h0 = pt.h0.as<float4>(); b.h1 = pt.h1.as<float4>(); b.h2 = pt.h2.as<float4>(); b.h3 = pt.h3.as<float4>();
    b.qm = pt.qm.as<int>();
    b.qk = nullptr;  // set by a batch that sorts the material queues
    b.qms = pt.qms.as<int>();
    b.sortTotals = pt.sortBins.as<int>();
    b.sortCursor = b.sortTotals + kSortSlots;
    b.nk = nullptr;  // set by a batch that sorts its rays
    b.qr = pt.qr.as<int>();
    b.raySortTotals = pt.raySortBins.as<int>();
    b.raySortCursor = b.raySortTotals + kRaySortBins;
    b.qes = pt.qes.as<int>();
    b.sq = pt.sq.as<int>();
    b.shadowIdx = 2;
    b.qe = pt.qe.as<int4>();
    b.qeLpdf = pt.qeLpdf.as<float4>();
    b.qcap = int(pt.cap);
    b.pfilm = pt.pfilm.as<float2>(); b.xyz = pt.xyz.as<float4>();
    b.so = pt.so.as<float4>(); b.sd = pt.sd.as<float4>(); b.sp = pt.sp.as<float4>(); b.sn = pt.sn.as<float4>();
    b.sf = pt.sf.as<float4>(); b.sL = pt.sL.as<float4>(); b.sb = pt.sb.as<float4>();
    b.gout = nullptr;
    b.trec = nullptr;
    b.train = nullptr;
    b.trainCount = nullptr;
    b.trainCap = 0;
    b.trainBase = 0;
    b.trainRing = 1;
    return b;
}

// Guiding buffers of the device's model (sized for the batch), when a render uses them.
void AttachGuide(Device &dev, Bufs &b) {
    if (!dev.guide) return;
    GuideState &g = *dev.guide;
    b.gout = g.out.as<float4>();
    b.trec = g.rec.as<float4>();
    b.train = g.train.as<float>();
    b.trainCount = g.trainCount.as<int>();
    b.trainCap = int(std::min<size_t>(g.trainCap, size_t(g.batchAdd)));
    b.trainBase = int(g.ringHead);
    b.trainRing = int(std::max<size_t>(1, g.trainCap));
}

Params MakeParams(const prgpu::RenderParamsEx &rp) {
    Params p;
    p.maxDepth = rp.maxDepth > 0 ? rp.maxDepth : 0x7fffffff;
    p.rrStart = rp.rrStartDepth;
    p.rr = rp.russianRoulette;
    p.clampY = rp.clampLuminance;
    p.seed = rp.seed;
    p.firstSample = rp.firstSample;
    p.guide = (rp.flags & prgpu::RenderGuide) ? 1 : 0;
    p.train = p.guide && (rp.flags & prgpu::RenderGuideTrain) ? 1 : 0;
    p.trainKeep = 1.f;
    p.box = GuideBox{};
    return p;
}

// Event-timed launches when profiling is on: each launch is bracketed by two events on the
// plugin's stream; durations are collected after the batch's synchronisation.
struct Timer {
    Device &dev;
    struct Pending {
        int kind;
        hipEvent_t a, b;
        int64_t items;
        const int *itemsHost;  // items read after the sync (queue counts), or null
    };
    std::vector<Pending> pending;
    explicit Timer(Device &d) : dev(d) {}
    hipEvent_t Get() {
        if (dev.eventPool.empty()) {
            hipEvent_t e;
            PG_CHECK(hipEventCreate(&e));
            return e;
        }
        hipEvent_t e = dev.eventPool.back();
        dev.eventPool.pop_back();
        return e;
    }
    template <class F> void Run(int kind, int64_t items, const int *itemsHost, F launch) {
        if (!dev.profileMode) {
            launch();
            return;
        }
        Pending p{kind, Get(), Get(), items, itemsHost};
        PG_CHECK(hipEventRecord(p.a, dev.stream));
        launch();
        PG_CHECK(hipEventRecord(p.b, dev.stream));
        pending.push_back(p);
    }
    // Call after a stream synchronisation; itemsHost values are read at this point.
    void Collect() {
        for (Pending &p : pending) {
            float ms = 0;
            PG_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
            dev.prof.ms[p.kind] += ms;
            dev.prof.launches[p.kind] += 1;
            dev.prof.items[p.kind] += p.itemsHost ? *p.itemsHost : p.items;
            dev.eventPool.push_back(p.a);
            dev.eventPool.push_back(p.b);
        }
        pending.clear();
    }
};

// Traces one batch of `passes` passes (samples firstSample .. firstSample + passes - 1) to
// completion; the per-sample results are left in the xyz / record buffers.


// Bounce rounds per captured graph (tuning key graph_rounds; even: a replay ends on the queue it
// started from); 0 = no graphs, the default. On configuration 2 at 1080p, replaying 4 captured
// rounds took 27.9 ms per step against 25.4 ms for the pipelined eager loop (late, nearly empty
// bounces still launch the whole-batch grids a graph bakes in, and a replay costs more than it
// saves in synchronisation).
// What a captured bounce graph bakes in: the kernel variants, grids, buffer addresses and the
// parameters the bounce kernels read (not the sample range, seed or clamp, which only the camera
// and finish kernels use).
uint64_t BounceKey(const Params &p, const Bufs &b, const DevScene *ds, int tgrid, int grid, bool media, bool interfaces,
                   int emitters, int layered) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *v, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(v);
        for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
    };
    const int ints[10] = {p.maxDepth, p.rrStart, p.rr, p.guide, p.train, tgrid, grid, media ? 1 : 0, interfaces ? 1 : 0,
                          emitters * 2 + layered};
    mix(ints, sizeof(ints));
    mix(&p.box, sizeof(p.box));
    mix(&b, sizeof(b));
    mix(&ds, sizeof(ds));
    return h;
}

void TraceBatch(Device &dev, const DevScene &s, const DevScene *ds, Params p, int passes, int npix, int sw) {
    PTScene &pt = *dev.pt;
    hipStream_t st = dev.stream;
    const int n = passes * npix;
    Bufs b = MakeBufs(pt);
    if (p.guide) {
        if (!dev.guide || dev.guide->cap < pt.cap) throw std::runtime_error("guiding state not prepared");
        AttachGuide(dev, b);
    }
    Timer timer(dev);
    const bool stats = dev.profileMode >= 2;
    const bool media = pt.hasMedia != 0, interfaces = pt.hasInterfaces != 0;
    unsigned long long *ctr = dev.travCounters.as<unsigned long long>();
    timer.Run(prgpu::KernCamera, n, nullptr, [&] {
        hipLaunchKernelGGL(k_camera, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, ds, p, b, n, npix, sw);
    });
    PG_CHECK(hipGetLastError());
    int active = n;
    hipLaunchKernelGGL(k_reset_counts, dim3(1), dim3(64), 0, st, b.counts, -1, -1);
    pt.countsHost[0] = n;
    PG_CHECK(hipMemcpyAsync(b.counts, pt.countsHost, sizeof(int), hipMemcpyHostToDevice, st));
    const int shadeGrid = std::max(1, std::min((active + kTile - 1) / kTile, dev.cus * dev.tune.shadeBlocksPerCU));
    // One bounce: closest hits of the current queue, the surface and emission stages, the
    // material families (which fill the next queue), the shadow rays. Every kernel reads its
    // queue length from device memory, so rounds can be queued without reading them back.
    // Without media or interfaces a bounce's shadow rays are traced in the next bounce's traversal
    // launch (k_trace_bounce): a bounce appends them to shadow queue shadow_count_index(curIdx) and
    // the next one, whose curIdx is the other, consumes them. After the last bounce a round over the
    // empty path queue traces the last shadow rays.
    const bool fused = kFusedShadow && !media && !interfaces;
    const bool sorted = dev.tune.materialSort != 0 && pt.sortable != 0;
    if (sorted) b.qk = pt.qk.as<int>();
    // The ray sort reorders the next queue into its own buffer, which the next bounce then traces
    // (not with captured bounce graphs, whose rounds fix their queues).
    const bool raySorted = dev.tune.raySort != 0 && dev.tune.graphRounds == 0;
    if (raySorted) b.nk = pt.nk.as<int>();
    const int sortGrid = std::max(1, std::min(shadeGrid, dev.cus * 4));
    int pendingShadow = 0;  // shadow rays the next round traces (profiled rounds: read back)
    auto round = [&](int curIdx, int bounce, int tgrid, int grid) {
        const int nxtIdx = 1 - curIdx;
        int *cur = raySorted && bounce > 0 ? b.qr : (curIdx ? b.q1 : b.q0), *nxt = curIdx ? b.q0 : b.q1;
        Bufs rb = b;
        const int sIn = shadow_count_index(1 - curIdx);
        rb.shadowIdx = fused ? shadow_count_index(curIdx) : 2;
        hipLaunchKernelGGL(k_reset_counts, dim3(1), dim3(64), 0, st, b.counts, curIdx, fused ? sIn : -1);
        timer.Run(prgpu::KernIntersect, int64_t(active) + (fused ? pendingShadow : 0), nullptr, [&] {
            if (fused) {
                auto *k = stats ? k_trace_bounce<true> : k_trace_bounce<false>;
                hipLaunchKernelGGL(k, dim3(tgrid), dim3(kTraceBlock), 0, st, ds, rb, cur, b.counts + curIdx, b.counts + sIn,
                                   b.counts + kFetchIntersect, b.counts + kFetchShadow, ctr);
            } else if (stats) {
                hipLaunchKernelGGL(k_intersect<true>, dim3(tgrid), dim3(kTraceBlock), 0, st, ds, rb, cur, b.counts + curIdx,
                                   b.counts + kFetchIntersect, ctr);
            } else {
                hipLaunchKernelGGL(k_intersect<false>, dim3(tgrid), dim3(kTraceBlock), 0, st, ds, rb, cur, b.counts + curIdx,
                                   b.counts + kFetchIntersect, ctr);
            }
        });
        // Guiding applies to the vertices of the first guideBounces bounces (0: all).
        const bool guided = p.guide && (dev.tune.guideBounces == 0 || bounce < dev.tune.guideBounces);
        timer.Run(prgpu::KernSurface, active, nullptr, [&] {
            auto *k = media ? (interfaces ? k_surface<true, true> : k_surface<true, false>)
                            : (interfaces ? k_surface<false, true> : k_surface<false, false>);
            hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, st, ds, p, rb, cur, b.counts + curIdx, nxt,
                               b.counts + nxtIdx);
            if (s.emitters) hipLaunchKernelGGL(k_emit, dim3(grid), dim3(kBlock), 0, st, ds, rb);
            if (sorted) {
                hipLaunchKernelGGL(k_sort_count, dim3(sortGrid), dim3(kBlock), 0, st, rb);
                hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(kBlock), 0, st, rb);
                hipLaunchKernelGGL(k_sort_scatter, dim3(sortGrid), dim3(kBlock), 0, st, rb);
            }
        });
        // The material stage reads the queues sorted by material and medium when the batch sorts.
        Bufs mb = rb;
        if (sorted) mb.qm = b.qms;
        if (guided)
            timer.Run(prgpu::KernGuide, 0, nullptr, [&] {
                dev.guide->Infer(mb.qm + size_t(QDiffuse) * b.qcap, b.counts + mat_count_index(QDiffuse), active, b.h0,
                                 b.h1, b.d, p.box, dev.cus, st);
            });
        timer.Run(prgpu::KernMaterial, active, nullptr, [&] {
            if (guided) {
                auto *kd = media ? k_material<QDiffuse, true, true> : k_material<QDiffuse, false, true>;
                hipLaunchKernelGGL(kd, dim3(grid), dim3(kBlock), 0, st, ds, p, mb, nxt, b.counts + nxtIdx);
            }
            if (kFusedMaterial) {
                auto *kp = media ? k_material_plain<true> : k_material_plain<false>;
                hipLaunchKernelGGL(kp, dim3(grid), dim3(kBlock), 0, st, ds, p, mb, nxt, b.counts + nxtIdx,
                                   guided ? int(QConductor) : int(QDiffuse));
            } else {
                auto *kd = media ? k_material<QDiffuse, true> : k_material<QDiffuse, false>;
                auto *kc = media ? k_material<QConductor, true> : k_material<QConductor, false>;
                auto *ke = media ? k_material<QDielectric, true> : k_material<QDielectric, false>;
                if (!guided) hipLaunchKernelGGL(kd, dim3(grid), dim3(kBlock), 0, st, ds, p, mb, nxt, b.counts + nxtIdx);
                hipLaunchKernelGGL(kc, dim3(grid), dim3(kBlock), 0, st, ds, p, mb, nxt, b.counts + nxtIdx);
                hipLaunchKernelGGL(ke, dim3(grid), dim3(kBlock), 0, st, ds, p, mb, nxt, b.counts + nxtIdx);
                auto *kph = k_material<QPhase, true>;
                if (media) hipLaunchKernelGGL(kph, dim3(grid), dim3(kBlock), 0, st, ds, p, mb, nxt, b.counts + nxtIdx);
            }
            if (pt.hasLayered) {
                auto *kl = media ? k_material<QLayered, true> : k_material<QLayered, false>;
                hipLaunchKernelGGL(kl, dim3(grid), dim3(kBlock), 0, st, ds, p, mb, nxt, b.counts + nxtIdx);
            }
            if (raySorted) {  // the next bounce traces the queue sorted by ray key
                hipLaunchKernelGGL(k_ray_sort_count, dim3(sortGrid), dim3(kBlock), 0, st, rb, b.counts + nxtIdx);
                hipLaunchKernelGGL(k_ray_sort_scan, dim3(1), dim3(kBlock), 0, st, rb);
                hipLaunchKernelGGL(k_ray_sort_scatter, dim3(sortGrid), dim3(kBlock), 0, st, rb, b.counts + nxtIdx, nxt);
            }
        });
        if (!fused)
            timer.Run(prgpu::KernShadow, 0, &pt.countsHost[2], [&] {
                auto *k = stats ? (media ? (interfaces ? k_shadow<true, true, true> : k_shadow<true, true, false>)
                                         : (interfaces ? k_shadow<true, false, true> : k_shadow<true, false, false>))
                                : (media ? (interfaces ? k_shadow<false, true, true> : k_shadow<false, true, false>)
                                         : (interfaces ? k_shadow<false, false, true> : k_shadow<false, false, false>));
                hipLaunchKernelGGL(k, dim3(tgrid), dim3(kTraceBlock), 0, st, ds, rb, b.counts + 2, b.counts + kFetchShadow,
                                   ctr);
            });
    };

    const int graphRounds = dev.tune.graphRounds;
    if (!dev.profileMode && !p.guide && graphRounds > 0) {
        // Unprofiled, unguided renders replay graphRounds captured bounces per launch (an even
        // count, so every replay starts from queue A) with grids sized for the whole batch, and
        // read the path count back once per replay instead of once per bounce.
        const int tgrid = TraceGrid(dev, n);
        const uint64_t key = BounceKey(p, b, ds, tgrid, shadeGrid, media, interfaces, s.emitters, pt.hasLayered) ^
                             (uint64_t(graphRounds) << 56);
        if (!pt.bounceExec || pt.bounceKey != key) {
            if (pt.bounceExec) PG_CHECK(hipGraphExecDestroy(pt.bounceExec));
            pt.bounceExec = nullptr;
            hipGraph_t g = nullptr;
            PG_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                for (int k = 0; k < graphRounds; ++k) round(k & 1, k, tgrid, shadeGrid);
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            PG_CHECK(hipStreamEndCapture(st, &g));
            const hipError_t e = hipGraphInstantiate(&pt.bounceExec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            PG_CHECK(e);
            pt.bounceKey = key;
        }
        while (active > 0) {
            PG_CHECK(hipGraphLaunch(pt.bounceExec, st));
            PG_CHECK(hipMemcpyAsync(pt.countsHost, b.counts, sizeof(int), hipMemcpyDeviceToHost, st));
            PG_CHECK(hipStreamSynchronize(st));
            active = pt.countsHost[0];
        }
        if (fused) round(0, graphRounds, tgrid, 1);  // the last replay's shadow rays
    } else if (!dev.profileMode) {
        // Pipelined: round k + 1 is queued before the host waits for the path count of round k,
        // so the device never idles on a read-back. A round's grids are sized by the count the host
        // last saw, an upper bound (paths only ever leave the queues); the kernels read the exact
        // count on the device. The count of the paths entering round k + 1 is copied right after
        // round k, before round k + 2 could clear it.
        int *slot = pt.countsHost + 8;  // two read-back slots, alternating
        for (hipEvent_t &e : dev.pipeEvents)
            if (!e) PG_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        hipEvent_t ready[2] = {dev.pipeEvents[0], dev.pipeEvents[1]};
        int bound = n, curIdx = 0, k = 0;
        // With fused shadow rays a round's traversal also takes the previous round's shadow rays,
        // at most one per path that entered it (<= bnd as well).
        auto launch = [&](int r, int curI, int bnd) {
            round(curI, r, TraceGrid(dev, int64_t(bnd) * (fused && r > 0 ? 2 : 1)),
                  std::max(1, std::min((bnd + kTile - 1) / kTile, shadeGrid)));
            PG_CHECK(hipMemcpyAsync(slot + (r & 1), b.counts + (1 - curI), sizeof(int), hipMemcpyDeviceToHost, st));
            PG_CHECK(hipEventRecord(ready[r & 1], st));
        };
        launch(0, 0, bound);
        while (true) {
            // Round k is queued and bounds round k + 1 (bound >= paths entering round k >= those entering k + 1).
            launch(k + 1, 1 - curIdx, bound);
            PG_CHECK(hipGetLastError());
            PG_CHECK(hipEventSynchronize(ready[k & 1]));
            bound = slot[k & 1];  // paths entering round k + 1
            curIdx = 1 - curIdx;
            ++k;
            if (bound == 0) break;  // round k (queued, empty) is the last
        }
        PG_CHECK(hipStreamSynchronize(st));
    } else {
        int curIdx = 0, bounce = 0;
        // A round runs while paths remain or, with fused shadow rays, while the last round left
        // shadow rays (the final round's path queue is empty).
        while (active > 0 || pendingShadow > 0) {
            const int tgrid = TraceGrid(dev, int64_t(active) + pendingShadow);
            const int grid = std::max(1, std::min((active + kTile - 1) / kTile, shadeGrid));
            round(curIdx, bounce, tgrid, grid);
            PG_CHECK(hipGetLastError());
            PG_CHECK(hipMemcpyAsync(pt.countsHost, b.counts, (kShadowCountB + 1) * sizeof(int), hipMemcpyDeviceToHost, st));
            PG_CHECK(hipStreamSynchronize(st));
            timer.Collect();
            pendingShadow = fused ? pt.countsHost[shadow_count_index(curIdx)] : 0;
            curIdx = 1 - curIdx;
            active = pt.countsHost[curIdx];
            ++bounce;
        }
    }
    CheckTraceFlags(dev);
}

// Guided rendering trains the model after every batch: up to guideSteps Adam steps of guideBatch
// records each (the records of the batch's paths, in the order they were written).
// Passes per batch: bounded by the path budget (the queue capacity), so that the state fits
// comfortably in HBM.
int PassesPerBatch(const Device &dev, int npix, int spp) {
    return int(std::max<int64_t>(1, std::min<int64_t>(spp, dev.tune.pathBudget / std::max(1, npix))));
}

}  // namespace

void RenderExImpl(Device &dev, const prgpu::SceneDescEx &sd, const prgpu::RenderParamsEx &rp, double *sum,
                  double *sumsq, const float *sheenTable) {
    if (rp.spp <= 0) return;
    EnsurePTScene(dev, sd, sheenTable);
    DevScene s = MakeDevScene(dev, sd, rp);
    const DevScene *ds = UploadDevScene(dev, s);
    const int sw = s.width + 2 * s.margin, sh = s.height + 2 * s.margin, npix = sw * sh;
    const int per = PassesPerBatch(dev, npix, rp.spp);
    PTScene &pt = *dev.pt;
    EnsureBatch(pt, size_t(per) * size_t(npix));
    const size_t filmN = size_t(s.width) * s.height * 3;
    // Host film: stage it in HBM for the call. Resident film: accumulate in place.
    DevBuf hsum, hsq;
    double *dsum, *dsq;
    if (sum) {
        hsum.upload(sum, filmN, dev.stream);
        if (sumsq) hsq.upload(sumsq, filmN, dev.stream);
        dsum = hsum.as<double>();
        dsq = sumsq ? hsq.as<double>() : nullptr;
    } else {
        if (!dev.filmSum.p || dev.filmW != s.width || dev.filmH != s.height)
            throw std::runtime_error("resident film not reset for this resolution");
        dsum = dev.filmSum.as<double>();
        dsq = dev.filmSq.as<double>();
    }
    const int reach = int(std::ceil(s.filterRadius + 0.5f));
    Params p = MakeParams(rp);
    if (p.guide) {
        if (!dev.guide || (rp.flags & prgpu::RenderGuideReset)) {
            dev.guide = std::make_unique<GuideState>();
            dev.guide->Init(dev.stream);
        }
        dev.guide->EnsureBatch(pt.cap);  // per-path planes are indexed with the batch capacity
        dev.guide->blend = dev.tune.guideBlend;
        if (p.train)
            dev.guide->EnsureTrain(std::max(size_t(dev.tune.guideBatch) * size_t(dev.tune.guideSteps),
                                            size_t(dev.tune.guideReplay)));
    }
    for (int a = 0; a < 3; ++a) {  // the scene's bounds: guiding's position encoding, the ray sort's cells
        p.box.lo[a] = dev.sceneLo[a];
        const float ext = dev.sceneHi[a] - dev.sceneLo[a];
        p.box.inv[a] = ext > 0 ? 1.f / ext : 0.f;
    }
    for (int done = 0; done < rp.spp; done += per) {
        int passes = std::min(per, rp.spp - done);
        p.firstSample = rp.firstSample + done;
        if (p.train) {
            GuideState &g = *dev.guide;
            PG_CHECK(hipMemsetAsync(g.trainCount.p, 0, 2 * sizeof(int), dev.stream));
            // The records live in a ring of trainCap records that keeps the last batches' records
            // (guide_replay): a batch adds up to a quarter of the ring, or the whole of what the
            // training steps use when there is no replay. A uniform subsample keeps about that
            // many, from the records the last batch produced per path (for the first, an estimate
            // of one per path).
            const size_t use = size_t(dev.tune.guideBatch) * size_t(dev.tune.guideSteps);
            g.batchAdd = dev.tune.guideReplay > 0 ? std::max<size_t>(1, g.trainCap / 4) : std::min(use, g.trainCap);
            if (dev.tune.guideReplay == 0) g.ringHead = 0, g.ringValid = 0;
            const double expected = g.recordsPerPath > 0 ? g.recordsPerPath * double(passes) * npix : double(passes) * npix;
            p.trainKeep = float(std::min(1.0, double(g.batchAdd) / std::max(1.0, expected)));
        }
        TraceBatch(dev, s, ds, p, passes, npix, sw);
        Bufs b = MakeBufs(pt);
        if (p.guide) AttachGuide(dev, b);
        int n = passes * npix;
        Timer timer(dev);
        timer.Run(prgpu::KernFinish, n, nullptr, [&] {
            hipLaunchKernelGGL(k_finish, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, dev.stream, ds, p, b, n,
                               static_cast<prgpu::SampleRecordEx *>(nullptr));
        });
        int pixels = s.width * s.height;
        timer.Run(prgpu::KernGather, int64_t(pixels) * passes, nullptr, [&] {
            if (reach <= kGatherMaxReach)
                hipLaunchKernelGGL(k_gather_tiled,
                                   dim3((s.width + kGatherTile - 1) / kGatherTile, (s.height + kGatherTile - 1) / kGatherTile),
                                   dim3(kBlock), 0, dev.stream, ds, b, passes, npix, sw, reach, dsum, dsq);
            else
                hipLaunchKernelGGL(k_gather, dim3((pixels + kBlock - 1) / kBlock), dim3(kBlock), 0, dev.stream, ds, b,
                                   passes, npix, sw, reach, dsum, dsq);
        });
        PG_CHECK(hipGetLastError());
        if (p.train) {
            // One Adam step per slice of guideBatch records of the batch (at most guideSteps), and at
            // least guideMinSteps steps, cycling over the slices, when the batch has records.
            GuideState &g = *dev.guide;
            int counts[2] = {0, 0};  // records written, records produced
            PG_CHECK(hipMemcpyAsync(counts, g.trainCount.p, 2 * sizeof(int), hipMemcpyDeviceToHost, dev.stream));
            PG_CHECK(hipStreamSynchronize(dev.stream));
            const size_t added = std::min<size_t>(size_t(std::max(0, counts[0])), g.batchAdd);
            g.recordsPerPath = double(counts[1]) / double(std::max(1, passes * npix));
            g.ringHead = (g.ringHead + added) % std::max<size_t>(1, g.trainCap);
            g.ringValid = std::min(g.trainCap, g.ringValid + added);
            // Adam steps on slices of guideBatch records of the ring: one per slice of the batch's
            // records (at most guideSteps) and at least guideMinSteps, the slices rotating over the
            // whole ring from call to call.
            const Tuning &t = dev.tune;
            const int valid = int(g.ringValid);
            const int slices = std::min(t.guideSteps, (int(std::min<size_t>(added, g.trainCap)) + t.guideBatch - 1) / t.guideBatch);
            const int steps = slices > 0 ? std::max(slices, t.guideMinSteps) : 0;
            const int ringSlices = std::max(1, valid / std::max(1, t.guideBatch));
            if (steps > 0)
                timer.Run(prgpu::KernGuide, steps, nullptr, [&] {
                    for (int k = 0; k < steps; ++k) {
                        const int first = int((g.rotor++ % size_t(ringSlices)) * size_t(t.guideBatch));
                        g.Step(g.train.as<float>(), nullptr, first, std::min(t.guideBatch, valid - first), t.guideLr,
                               dev.stream);
                    }
                });
        }
        if (dev.profileMode) {
            PG_CHECK(hipStreamSynchronize(dev.stream));
            timer.Collect();
        }
    }
    if (sum) {
        PG_CHECK(hipMemcpyAsync(sum, dsum, filmN * sizeof(double), hipMemcpyDeviceToHost, dev.stream));
        if (sumsq) PG_CHECK(hipMemcpyAsync(sumsq, dsq, filmN * sizeof(double), hipMemcpyDeviceToHost, dev.stream));
    }
    PG_CHECK(hipStreamSynchronize(dev.stream));
}

void SampleRadianceImpl(Device &dev, const prgpu::SceneDescEx &sd, const prgpu::RenderParamsEx &rp,
                        prgpu::SampleRecordEx *out) {
    EnsurePTScene(dev, sd, nullptr);
    DevScene s = MakeDevScene(dev, sd, rp);
    const DevScene *ds = UploadDevScene(dev, s);
    const int sw = s.width + 2 * s.margin, sh = s.height + 2 * s.margin, npix = sw * sh;
    PTScene &pt = *dev.pt;
    EnsureBatch(pt, size_t(npix));
    Params p = MakeParams(rp);
    p.guide = p.train = 0;  // per-sample records are the plain path tracer's
    TraceBatch(dev, s, ds, p, 1, npix, sw);
    DevBuf rec;
    rec.ensure(size_t(npix) * sizeof(prgpu::SampleRecordEx));
    Bufs b = MakeBufs(pt);
    hipLaunchKernelGGL(k_finish, dim3((npix + kBlock - 1) / kBlock), dim3(kBlock), 0, dev.stream, ds, p, b, npix,
                       rec.as<prgpu::SampleRecordEx>());
    PG_CHECK(hipGetLastError());
    PG_CHECK(hipMemcpyAsync(out, rec.p, size_t(npix) * sizeof(prgpu::SampleRecordEx), hipMemcpyDeviceToHost, dev.stream));
    PG_CHECK(hipStreamSynchronize(dev.stream));
}

}  // namespace pgpu
