// Device traversal of the 8-wide quantised BVH (gfx950, wave64).
//
// One ray per lane. A lane's traversal stack lives in LDS, interleaved by lane so that the 64
// lanes of a wave touch 64 consecutive banks; entries beyond the LDS part spill to a per-thread
// HBM slab (sized by the host from the tree depth: at most 7 entries per level are pending).
// Children that a ray's slab test hits are ordered near-to-far with an 8-key sorting network
// (key = distance bits | slot), the nearest is taken next and the rest are pushed far-to-near.
//
// The triangle test is the reference's Moller-Trumbore in float (src/geometry/shape.cpp:202-222)
// operation for operation; the file is compiled without FMA contraction so that the hit
// distances, and therefore the closest-hit primitive ids, are bit-identical to the CPU's.
#pragma once

#include <hip/hip_runtime.h>

#include "wbvh.h"

namespace pgpu {

// Tuning knobs (compile-time): LDS stack entries per lane and the idle-lane threshold of the
// per-lane ray fetch.
#ifndef PG_STACK_LDS
#define PG_STACK_LDS 16
#endif
#ifndef PG_REFILL
#define PG_REFILL 48
#endif
#ifndef PG_LEAF_DEFER
#define PG_LEAF_DEFER 0
#endif
#ifndef PG_FETCH_PARTS
#define PG_FETCH_PARTS 1
#endif
#ifndef PG_ORDERED_SLAB
#define PG_ORDERED_SLAB 1
#endif
#ifndef PG_CHILD_RELOAD
#define PG_CHILD_RELOAD 0
#endif
constexpr int kTraceBlock = 256;
constexpr int kStackLds = PG_STACK_LDS;  // LDS entries per lane

struct DRay {
    float ox, oy, oz, dx, dy, dz;
};

struct DHit {
    float t;
    int prim;  // original primitive index, -1 = miss
    float b1, b2;
};

__device__ __forceinline__ float dev_inv(float d) {
    // 1/d, with +-FLT_MAX instead of +-inf so that 0 * inv never produces a NaN in the slab test.
    float r = 1.f / d;
    return __builtin_isinf(r) ? __builtin_copysignf(3.402823466e38f, d) : r;
}

// Reference Moller-Trumbore, no contraction. Returns true and updates t/b1/b2 on a closer hit.
__device__ __forceinline__ bool dev_hit_triangle(const DRay &r, float p0x, float p0y, float p0z, float e1x, float e1y,
                                                 float e1z, float e2x, float e2y, float e2z, float tMax, float *tOut,
                                                 float *uOut, float *vOut) {
    // pvec = cross(d, e2)
    float px = r.dy * e2z - r.dz * e2y;
    float py = r.dz * e2x - r.dx * e2z;
    float pz = r.dx * e2y - r.dy * e2x;
    float det = e1x * px + e1y * py + e1z * pz;
    if (det == 0.f || !__builtin_isfinite(det)) return false;
    float inv = 1.f / det;
    float tx = r.ox - p0x, ty = r.oy - p0y, tz = r.oz - p0z;
    float u = (tx * px + ty * py + tz * pz) * inv;
    if (u < 0.f || u > 1.f) return false;
    // qvec = cross(tvec, e1)
    float qx = ty * e1z - tz * e1y;
    float qy = tz * e1x - tx * e1z;
    float qz = tx * e1y - ty * e1x;
    float v = (r.dx * qx + r.dy * qy + r.dz * qz) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    float t = (e2x * qx + e2y * qy + e2z * qz) * inv;
    if (t <= 0.f || t >= tMax) return false;
    *tOut = t;
    *uOut = u;
    *vOut = v;
    return true;
}

// Robust quadratic in double precision (src/geometry/shape.cpp:42-61).
__device__ __forceinline__ bool dev_hit_sphere(const DRay &r, float cx, float cy, float cz, float rad, float tMax,
                                               float *tOut) {
    double ox = r.ox - cx, oy = r.oy - cy, oz = r.oz - cz;
    double dx = r.dx, dy = r.dy, dz = r.dz;
    double a = dx * dx + dy * dy + dz * dz;
    double b = ox * dx + oy * dy + oz * dz;
    double c = ox * ox + oy * oy + oz * oz - double(rad) * rad;
    double k = b / a;
    double lx = ox - k * dx, ly = oy - k * dy, lz = oz - k * dz;
    double disc = double(rad) * rad - (lx * lx + ly * ly + lz * lz);
    if (disc < 0) return false;
    double q = -b - __builtin_copysign(__builtin_sqrt(a * disc), b);
    double t0 = c / q, t1 = q / a;
    if (t0 > t1) {
        double s = t0;
        t0 = t1;
        t1 = s;
    }
    double t = t0;
    if (t <= 0) t = t1;
    if (t <= 0 || t >= tMax) return false;
    *tOut = float(t);
    return true;
}

__device__ __forceinline__ void cswap(uint32_t &a, uint32_t &b) {
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo;
    b = hi;
}

// Optimal 19-comparator network for 8 keys.
__device__ __forceinline__ void sort8(uint32_t k[8]) {
    cswap(k[0], k[2]); cswap(k[1], k[3]); cswap(k[4], k[6]); cswap(k[5], k[7]);
    cswap(k[0], k[4]); cswap(k[1], k[5]); cswap(k[2], k[6]); cswap(k[3], k[7]);
    cswap(k[0], k[1]); cswap(k[2], k[3]); cswap(k[4], k[5]); cswap(k[6], k[7]);
    cswap(k[2], k[4]); cswap(k[3], k[5]);
    cswap(k[1], k[4]); cswap(k[3], k[6]);
    cswap(k[1], k[2]); cswap(k[3], k[4]); cswap(k[5], k[6]);
}

__device__ __forceinline__ float pow2f(int e) { return __int_as_float((e + 127) << 23); }

struct TraceCtx {
    const WNode *nodes;
    const TriRec *recs;
    uint32_t *spill;       // per-thread spill slab: entry k at spill[k * spillStride + tid]
    int spillStride;       // number of threads sharing the slab layout
    int spillCap;          // entries available per thread in the slab
    int *errorFlag;        // set to 1 if a traversal stack would overflow its slab
};

// Counters for the algorithmic-traffic figures (nodes fetched, leaf records tested).
struct TraceStats {
    uint32_t nodes = 0, recs = 0;
};

// Traversal state of one ray: the inverse direction, the current node or leaf reference, the
// stack depth and the best hit so far. The stack itself lives in LDS (+ the HBM spill slab).
struct TravState {
    float ix, iy, iz;
    float tMax;
    uint32_t cur;
    int sp;
    bool found;
    bool any;  // any-hit query (read by trace_step<..., kRuntimeAny = true>)
    DHit hit;
};

// Inverse direction for the ordered slab test: |1/d| clamped to 2^64, so that the per-node
// products with the node scale (a power of two at most 2^60) and the origin offset stay finite
// and an axis the ray is parallel to yields signed, very large but never infinite distances.
__device__ __forceinline__ float slab_inv(float d) {
    float r = 1.f / d;
    return __builtin_copysignf(fminf(fabsf(r), 0x1p64f), d);
}

__device__ __forceinline__ void trace_begin(TravState &t, const DRay &r, float tMax) {
#if PG_ORDERED_SLAB
    t.ix = slab_inv(r.dx);
    t.iy = slab_inv(r.dy);
    t.iz = slab_inv(r.dz);
#else
    t.ix = dev_inv(r.dx);
    t.iy = dev_inv(r.dy);
    t.iz = dev_inv(r.dz);
#endif
    t.tMax = tMax;
    t.cur = 0;  // root node
    t.sp = 0;
    t.found = false;
    t.any = false;
    t.hit.prim = -1;
}

__device__ __forceinline__ uint32_t trace_pop(const TraceCtx &c, TravState &t, const uint32_t *lds, uint32_t lane,
                                              uint32_t bd, int tid) {
    --t.sp;
    return t.sp < kStackLds ? lds[t.sp * bd + lane] : c.spill[size_t(t.sp - kStackLds) * c.spillStride + tid];
}

// One traversal step: one wide node (slab-test its children, push the hit ones far to near and
// continue with the nearest) or one leaf (test its records). Returns true when the traversal of
// this ray is complete. `lds` is the block's stack area (kStackLds * blockDim.x entries), `tid`
// the global thread index for the spill slab. An any-hit query (kAny, or t.any with kRuntimeAny)
// completes at its first hit.
template <bool kAny, bool kStats, bool kRuntimeAny = false>
__device__ __forceinline__ bool trace_step(const TraceCtx &c, const DRay &r, TravState &t, uint32_t *lds, int tid,
                                           TraceStats *st) {
    const uint32_t lane = threadIdx.x;
    const uint32_t bd = blockDim.x;
    if (!(t.cur & kLeafBit)) {
        const uint4 *np = reinterpret_cast<const uint4 *>(c.nodes + t.cur);
        uint4 h0 = np[0], q0 = np[3], q1 = np[4], q2 = np[5];
        if (kStats) st->nodes++;
        float ox = __uint_as_float(h0.x), oy = __uint_as_float(h0.y), oz = __uint_as_float(h0.z);
        int ex = int(int8_t(h0.w & 0xff)), ey = int(int8_t((h0.w >> 8) & 0xff)), ez = int(int8_t((h0.w >> 16) & 0xff));
        int nch = int(h0.w >> 24);
        float sx = pow2f(ex), sy = pow2f(ey), sz = pow2f(ez);
        // Per-axis origin terms shared by all children.
        float bx = (ox - r.ox), by = (oy - r.oy), bz = (oz - r.oz);
#if !PG_CHILD_RELOAD
        const uint4 c0 = np[1], c1 = np[2];
        uint32_t child[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#endif
        uint32_t qlx[2] = {q0.x, q0.y}, qly[2] = {q0.z, q0.w}, qlz[2] = {q1.x, q1.y};
        uint32_t qhx[2] = {q1.z, q1.w}, qhy[2] = {q2.x, q2.y}, qhz[2] = {q2.z, q2.w};
        uint32_t key[8];
#if PG_ORDERED_SLAB
        // Ordered slab test: the ray's direction signs pick each axis's entry (near) and exit (far)
        // planes once per node, and a plane distance is one FMA, q (scale / d) + (origin - o) / d.
        // The quantised child boxes are rounded outwards by a grid step, far more than these
        // roundings, and the widening below is kept.
        // exact (power-of-two scales); clamped so that q * a stays finite for any node exponent
        const float ax = __builtin_copysignf(fminf(fabsf(sx * t.ix), 0x1p118f), t.ix);
        const float ay = __builtin_copysignf(fminf(fabsf(sy * t.iy), 0x1p118f), t.iy);
        const float az = __builtin_copysignf(fminf(fabsf(sz * t.iz), 0x1p118f), t.iz);
        const float cx = bx * t.ix, cy = by * t.iy, cz = bz * t.iz;
        const bool px = t.ix >= 0.f, py = t.iy >= 0.f, pz = t.iz >= 0.f;
        const uint32_t nx[2] = {px ? qlx[0] : qhx[0], px ? qlx[1] : qhx[1]};
        const uint32_t fx[2] = {px ? qhx[0] : qlx[0], px ? qhx[1] : qlx[1]};
        const uint32_t ny[2] = {py ? qly[0] : qhy[0], py ? qly[1] : qhy[1]};
        const uint32_t fy[2] = {py ? qhy[0] : qly[0], py ? qhy[1] : qly[1]};
        const uint32_t nz[2] = {pz ? qlz[0] : qhz[0], pz ? qlz[1] : qhz[1]};
        const uint32_t fz[2] = {pz ? qhz[0] : qlz[0], pz ? qhz[1] : qlz[1]};
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int w = i >> 2, sh = (i & 3) * 8;
            float tn = fmaxf(fmaxf(__builtin_fmaf(float((nx[w] >> sh) & 0xff), ax, cx),
                                   __builtin_fmaf(float((ny[w] >> sh) & 0xff), ay, cy)),
                             __builtin_fmaf(float((nz[w] >> sh) & 0xff), az, cz));
            float tf = fminf(fminf(__builtin_fmaf(float((fx[w] >> sh) & 0xff), ax, cx),
                                   __builtin_fmaf(float((fy[w] >> sh) & 0xff), ay, cy)),
                             __builtin_fmaf(float((fz[w] >> sh) & 0xff), az, cz));
            tf = tf * (1.f + 2.f * 1.2e-7f * 3.f);
            tn = tn - fabsf(tn) * (2.f * 1.2e-7f * 3.f);
            bool h = (i < nch) && tn <= tf && tn < t.tMax && tf >= 0.f;
            key[i] = h ? ((__float_as_uint(fmaxf(tn, 0.f)) & ~7u) | uint32_t(i)) : 0xffffffffu;
        }
#else
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int w = i >> 2, sh = (i & 3) * 8;
            float lx = float((qlx[w] >> sh) & 0xff) * sx + bx, hx = float((qhx[w] >> sh) & 0xff) * sx + bx;
            float ly = float((qly[w] >> sh) & 0xff) * sy + by, hy = float((qhy[w] >> sh) & 0xff) * sy + by;
            float lz = float((qlz[w] >> sh) & 0xff) * sz + bz, hz = float((qhz[w] >> sh) & 0xff) * sz + bz;
            float t0x = lx * t.ix, t1x = hx * t.ix, t0y = ly * t.iy, t1y = hy * t.iy, t0z = lz * t.iz, t1z = hz * t.iz;
            float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
            float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
            // Conservative widening as the reference slab test (bvh.cpp:137-156), on both ends.
            tf = tf * (1.f + 2.f * 1.2e-7f * 3.f);
            tn = tn - fabsf(tn) * (2.f * 1.2e-7f * 3.f);
            bool h = (i < nch) && tn <= tf && tn < t.tMax && tf >= 0.f;
            key[i] = h ? ((__float_as_uint(fmaxf(tn, 0.f)) & ~7u) | uint32_t(i)) : 0xffffffffu;
        }
#endif
        sort8(key);
        if (key[0] == 0xffffffffu) {
            if (t.sp == 0) return true;
            t.cur = trace_pop(c, t, lds, lane, bd, tid);
            return false;
        }
        // Push hit children far to near; continue with the nearest. A child's reference is read by
        // its slot index (key & 7): from the node, just loaded and still in the vector L1, rather
        // than from the register copy, whose per-lane indexing costs a compare-select chain.
#if PG_CHILD_RELOAD
        const uint32_t *cref = reinterpret_cast<const uint32_t *>(np) + 4;
#define PG_CHILD(k) cref[(k) & 7u]
#else
#define PG_CHILD(k) child[(k) & 7u]
#endif
#pragma unroll
        for (int i = 7; i >= 1; --i) {
            if (key[i] != 0xffffffffu) {
                uint32_t ref = PG_CHILD(key[i]);
                if (t.sp < kStackLds) {
                    lds[t.sp * bd + lane] = ref;
                } else if (t.sp - kStackLds < c.spillCap) {
                    c.spill[size_t(t.sp - kStackLds) * c.spillStride + tid] = ref;
                } else {
                    *c.errorFlag = 1;
                    --t.sp;  // keep the stack in bounds; the error flag reports the lost subtree
                }
                ++t.sp;
            }
        }
        t.cur = PG_CHILD(key[0]);
#undef PG_CHILD
        return false;
    }
    // Leaf: test its records.
    const uint32_t first = t.cur & 0x07ffffffu, cnt = ((t.cur >> 27) & 7u) + 1u;
    for (uint32_t k = 0; k < cnt; ++k) {
        const float4 *rp = reinterpret_cast<const float4 *>(c.recs + first + k);
        float4 a = rp[0], b = rp[1], d = rp[2];
        if (kStats) st->recs++;
        uint32_t prim = __float_as_uint(d.y);
        float th, u = 0.f, v = 0.f;
        bool h;
        if (prim & kSphereBit) h = dev_hit_sphere(r, a.x, a.y, a.z, a.w, t.tMax, &th);
        else h = dev_hit_triangle(r, a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, d.x, t.tMax, &th, &u, &v);
        if (h) {
            t.found = true;
            t.tMax = th;
            t.hit.t = th;
            t.hit.prim = int(prim & ~kSphereBit);
            t.hit.b1 = u;
            t.hit.b2 = v;
            if (kAny || (kRuntimeAny && t.any)) return true;
        }
    }
    if (t.sp == 0) return true;
    t.cur = trace_pop(c, t, lds, lane, bd, tid);
    return false;
}

// Closest (kAny = false) or any (kAny = true) hit along r in (0, tMax), traversed to completion.
template <bool kAny, bool kStats>
__device__ __forceinline__ bool dev_trace(const TraceCtx &c, const DRay &r, float tMax, DHit *hit, uint32_t *lds,
                                          int tid, TraceStats *st) {
    TravState t;
    trace_begin(t, r, tMax);
    while (!trace_step<kAny, kStats>(c, r, t, lds, tid, st)) {
    }
    *hit = t.hit;
    return t.found;
}

// Per-lane ray fetch for persistent traversal kernels: lanes without a ray take consecutive
// indices from a global counter, one atomic per wave, once at least `kRefill` lanes of the wave
// are idle (or no lane has work left). Sets `idx` (< n) or marks the lane exhausted.
constexpr int kRefill = PG_REFILL;

// Leaf deferral for persistent traversal: when > 0, a wave's lanes that sit at a leaf wait while
// other lanes still descend inner nodes, until at least kLeafDefer of them have gathered (or no
// lane is at an inner node); then only the leaf lanes test their records. Returns whether this
// lane advances in the current iteration.
constexpr int kLeafDefer = PG_LEAF_DEFER;
__device__ __forceinline__ bool trace_turn(bool busy, uint32_t cur) {
    if (kLeafDefer <= 0) return busy;
    const bool atLeaf = busy && (cur & kLeafBit);
    const unsigned long long leafMask = __ballot(atLeaf), innerMask = __ballot(busy && !atLeaf);
    const bool leafTurn = innerMask == 0 || __popcll(leafMask) >= kLeafDefer;
    return busy && (atLeaf == leafTurn);
}
// Ray fetch of the persistent traversals. With kFetchParts > 1 the queue is split into that many
// contiguous parts, one per XCD (workgroups are dispatched to the 8 XCDs round-robin, so block b
// runs on XCD b % 8): a wave drains its own XCD's part first, then the others in turn. Measured on
// configuration 2 this is slower (8 parts: closest hit 8.9 ms per step against 6.7 ms with one
// part): one queue front keeps every XCD on the same rays, whose nodes then stay in the shared
// MALL. The default is one part. Per-wave reservoirs of larger chunks (a static share per wave,
// then guided chunks of 64 to 500 rays) were slower too (8.9 to 13.4 ms against 6.4 ms): they
// widen the part of the queue in flight, and the fetch atomics (about two per wave and launch) do
// not limit the traversal. `part` and `left` are wave-uniform; the loop ends when every part is
// drained and no lane holds a ray.
constexpr int kFetchParts = PG_FETCH_PARTS;
struct FetchState {
    int part, left;
    bool done;
};
__device__ __forceinline__ FetchState fetch_begin() {
    FetchState f;
    f.part = int(blockIdx.x % kFetchParts);
    f.left = kFetchParts;
    f.done = false;
    return f;
}
__device__ __forceinline__ bool fetch_rays_parts(int *counters, int n, int &idx, FetchState &f) {
    const bool idle = idx < 0 && !f.done;
    const unsigned long long idleMask = __ballot(idle);
    const unsigned long long busyMask = __ballot(idx >= 0);
    if (idleMask != 0 && (__popcll(idleMask) >= kRefill || busyMask == 0)) {
        const int lane = __lane_id();
        const int leader = __ffsll(idleMask) - 1;
        const int want = __popcll(idleMask);
        const int lo = int(int64_t(f.part) * n / kFetchParts), hi = int(int64_t(f.part + 1) * n / kFetchParts);
        int base = 0;
        if (lane == leader) base = atomicAdd(&counters[f.part], want);
        base = __shfl(base, leader);
        if (idle) {
            const int k = lo + base + __popcll(idleMask & ((1ull << lane) - 1ull));
            if (k < hi) idx = k;
        }
        if (lo + base + want >= hi) {  // this part is drained: the next one
            f.part = f.part + 1 == kFetchParts ? 0 : f.part + 1;
            if (--f.left == 0) f.done = true;
        }
    }
    return __ballot(idx >= 0) != 0 || !f.done;
}

}  // namespace pgpu
