// Checkpoint files for progressive GPU renders, in the reference's format
// (src/integrators/integrator.cpp:16-90, src/core/film.cpp:61-83):
//
//   "PRCKPT02" | u32 integrator type | 0x100 (GPU) | i32 maxDepth | i32 width | i32 height |
//   i32 chains | u64 scene fingerprint | u64 seed | u64 n = width*height*3 | n x f64 film sums |
//   u64 body size | body
//
// The body of a GPU path-tracer checkpoint is the i32 number of samples per pixel done
// (gpu_path.cpp:466-474); this host appends the FP64 sum-of-squares film after it when it has
// one (a reader that stops after the pass count ignores it). A resume continues at sample index
// `passes`, so a resumed render equals an uninterrupted one up to FP64 summation order.
//
// The fingerprint hashes what the reference's Scene::Fingerprint hashes (src/scene/scene.cpp:61-70)
// as far as the flat scene carries it: primitive and light counts, each primitive's bounds and
// whether it has a material and an area light, and the camera position; each light contributes
// its kind, spectrum row and scale in place of its luminous power.
#include "prender_gpu.h"
#include "prender_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

void SetErr(char *err, int len, const std::string &m) {
    if (err && len > 0) std::snprintf(err, size_t(len), "%s", m.c_str());
}

// MurmurHash64A (Austin Appleby, public domain), seed 0, over a byte string: the hash the
// reference applies to the bytes of its arguments (src/core/rng.h:67-105).
uint64_t Murmur64(const uint8_t *data, size_t len) {
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    const int r = 47;
    uint64_t h = 0 ^ (len * m);
    size_t nblocks = len / 8;
    for (size_t i = 0; i < nblocks; ++i) {
        uint64_t k;
        std::memcpy(&k, data + 8 * i, 8);
        k *= m;
        k ^= k >> r;
        k *= m;
        h ^= k;
        h *= m;
    }
    const uint8_t *tail = data + 8 * nblocks;
    switch (len & 7) {
    case 7: h ^= uint64_t(tail[6]) << 48; [[fallthrough]];
    case 6: h ^= uint64_t(tail[5]) << 40; [[fallthrough]];
    case 5: h ^= uint64_t(tail[4]) << 32; [[fallthrough]];
    case 4: h ^= uint64_t(tail[3]) << 24; [[fallthrough]];
    case 3: h ^= uint64_t(tail[2]) << 16; [[fallthrough]];
    case 2: h ^= uint64_t(tail[1]) << 8; [[fallthrough]];
    case 1:
        h ^= uint64_t(tail[0]);
        h *= m;
    }
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}

struct Bytes {
    std::vector<uint8_t> b;
    template <class T> Bytes &put(const T &v) {
        const uint8_t *p = reinterpret_cast<const uint8_t *>(&v);
        b.insert(b.end(), p, p + sizeof(T));
        return *this;
    }
    uint64_t hash() const { return Murmur64(b.data(), b.size()); }
};

constexpr char kMagic[8] = {'P', 'R', 'C', 'K', 'P', 'T', '0', '2'};
constexpr uint32_t kPathIntegrator = 0, kGpuDevice = 0x100;
constexpr int32_t kChains = 1024;  // IntegratorSettings::chains default (src/scene/scene.h:36)

std::vector<uint8_t> Header(const prgpu::SceneDescEx &sd, int32_t maxDepth, uint64_t seed) {
    Bytes h;
    h.b.insert(h.b.end(), kMagic, kMagic + 8);
    h.put(uint32_t(kPathIntegrator | kGpuDevice)).put(maxDepth).put(int32_t(sd.camera.width));
    h.put(int32_t(sd.camera.height)).put(kChains).put(ph_scene_fingerprint(&sd)).put(seed);
    return h.b;
}

}  // namespace

PH_API uint64_t ph_scene_fingerprint(const prgpu::SceneDescEx *sd) {
    Bytes c;
    c.put(uint64_t(sd->primCount)).put(uint64_t(sd->lightCount));
    uint64_t h = c.hash();
    for (int i = 0; i < sd->primCount; ++i) {
        const prgpu::Primitive &p = sd->prims[i];
        float lo[3], hi[3];
        if (p.type == prgpu::PrimSphere) {
            const float c3[3] = {p.p0.x, p.p0.y, p.p0.z};
            for (int a = 0; a < 3; ++a) lo[a] = c3[a] - p.r, hi[a] = c3[a] + p.r;
        } else {
            const prgpu::Float3 v[3] = {p.p0, p.p1, p.p2};
            lo[0] = std::min({v[0].x, v[1].x, v[2].x}), hi[0] = std::max({v[0].x, v[1].x, v[2].x});
            lo[1] = std::min({v[0].y, v[1].y, v[2].y}), hi[1] = std::max({v[0].y, v[1].y, v[2].y});
            lo[2] = std::min({v[0].z, v[1].z, v[2].z}), hi[2] = std::max({v[0].z, v[1].z, v[2].z});
        }
        bool hasMaterial = p.material >= 0 && p.material < sd->materialCount &&
                           sd->materials[p.material].type != prgpu::MatInterface;
        Bytes e;
        e.put(h).put(lo).put(hi).put(hasMaterial).put(p.light >= 0);
        h = e.hash();
    }
    for (int i = 0; i < sd->lightCount; ++i) {
        const prgpu::LightEx &l = sd->lights[i];
        Bytes e;
        e.put(h).put(l.kind).put(l.scale);
        if (l.spectrum >= 0 && l.spectrum < sd->spectrumCount) e.put(sd->spectra[l.spectrum]);
        h = e.hash();
    }
    const float *m = sd->camera.cameraToWorld;
    Bytes e;
    e.put(h).put(m[3]).put(m[7]).put(m[11]);
    return e.hash();
}

PH_API int ph_checkpoint_write(const char *path, const prgpu::SceneDescEx *sd, int32_t maxDepth, uint64_t seed,
                               const double *sum, const double *sumsq, int32_t passes, char *err, int errLen) {
    if (!path || !sd || !sum) {
        SetErr(err, errLen, "invalid arguments");
        return 1;
    }
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        SetErr(err, errLen, "cannot write checkpoint " + tmp);
        return 1;
    }
    std::vector<uint8_t> hdr = Header(*sd, maxDepth, seed);
    const uint64_t n = uint64_t(sd->camera.width) * uint64_t(sd->camera.height) * 3;
    const uint64_t bodySize = sizeof(int32_t) + (sumsq ? n * sizeof(double) : 0);
    bool ok = std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() && std::fwrite(&n, sizeof(n), 1, f) == 1 &&
              std::fwrite(sum, sizeof(double), size_t(n), f) == size_t(n) &&
              std::fwrite(&bodySize, sizeof(bodySize), 1, f) == 1 && std::fwrite(&passes, sizeof(passes), 1, f) == 1;
    if (ok && sumsq) ok = std::fwrite(sumsq, sizeof(double), size_t(n), f) == size_t(n);
    ok = std::fclose(f) == 0 && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        SetErr(err, errLen, std::string("failed to write checkpoint ") + path);
        return 1;
    }
    return 0;
}

PH_API int ph_checkpoint_read(const char *path, const prgpu::SceneDescEx *sd, int32_t maxDepth, uint64_t seed,
                              double *sum, double *sumsq, int32_t *passes, char *err, int errLen) {
    if (!path || !sd || !sum || !passes) {
        SetErr(err, errLen, "invalid arguments");
        return 1;
    }
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        SetErr(err, errLen, std::string("cannot open checkpoint ") + path);
        return 1;
    }
    std::vector<uint8_t> want = Header(*sd, maxDepth, seed), got(want.size());
    bool readHeader = std::fread(got.data(), 1, got.size(), f) == got.size();
    if (!readHeader || std::memcmp(got.data(), want.data(), 8) != 0) {
        std::fclose(f);
        SetErr(err, errLen, std::string(path) + " is not a prender checkpoint");
        return 1;
    }
    if (got != want) {
        std::fclose(f);
        SetErr(err, errLen,
               std::string(path) + " was made for a different scene, integrator, device, resolution, seed or depth");
        return 1;
    }
    const uint64_t expect = uint64_t(sd->camera.width) * uint64_t(sd->camera.height) * 3;
    uint64_t n = 0, bodySize = 0;
    bool ok = std::fread(&n, sizeof(n), 1, f) == 1 && n == expect &&
              std::fread(sum, sizeof(double), size_t(n), f) == size_t(n) &&
              std::fread(&bodySize, sizeof(bodySize), 1, f) == 1 && bodySize >= sizeof(int32_t) &&
              bodySize < (uint64_t(1) << 40) && std::fread(passes, sizeof(int32_t), 1, f) == 1;
    if (ok && sumsq) {
        if (bodySize >= sizeof(int32_t) + n * sizeof(double))
            ok = std::fread(sumsq, sizeof(double), size_t(n), f) == size_t(n);
        else
            std::fill(sumsq, sumsq + n, 0.0);  // a checkpoint without the variance film
    }
    std::fclose(f);
    if (!ok || *passes < 0) {
        SetErr(err, errLen, std::string(path) + " is truncated or corrupt");
        return 1;
    }
    return 0;
}
