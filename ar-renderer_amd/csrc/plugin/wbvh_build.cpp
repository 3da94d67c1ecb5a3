// Host build of the 8-wide quantised BVH: a binned-SAH binary build (subtrees above 64k
// primitives build on their own threads) collapsed into 8-wide nodes by repeatedly opening the
// child with the largest surface area, then quantised outwards to 8 bits per axis.
#include "wbvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>

namespace pgpu {

namespace {

struct Box {
    float lo[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                   std::numeric_limits<float>::infinity()};
    float hi[3] = {-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                   -std::numeric_limits<float>::infinity()};
    void grow(const Box &b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    void grow(const float *p) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    float area() const {
        float d[3];
        for (int a = 0; a < 3; ++a) d[a] = std::max(0.f, hi[a] - lo[a]);
        return 2 * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2]);
    }
};

struct BPrim {
    Box b;
    float c[3];
    int id;
};

struct BNode {
    Box b;
    int left = -1, right = -1;  // children (internal)
    int first = 0, count = 0;   // prims (leaf, count > 0)
};

class BinaryBuilder {
  public:
    std::vector<BPrim> prims;
    std::vector<BNode> nodes;
    std::atomic<int> used{0};

    int Build() {
        nodes.resize(std::max<size_t>(1, 2 * prims.size()));
        return Recurse(0, int(prims.size()), 0);
    }

  private:
    int Alloc() { return used.fetch_add(1); }

#ifndef PG_BVH_LEAF_CANDIDATE
#define PG_BVH_LEAF_CANDIDATE 4
#endif
#ifndef PG_BVH_TRAVERSAL_COST
#define PG_BVH_TRAVERSAL_COST 1.0f
#endif
    static constexpr int kLeafCandidate = PG_BVH_LEAF_CANDIDATE;
    static constexpr float kTraversalCost = PG_BVH_TRAVERSAL_COST;

    int Recurse(int b, int e, int depth) {
        int self = Alloc();
        BNode &node = nodes[size_t(self)];
        Box bounds, cb;
        for (int i = b; i < e; ++i) {
            bounds.grow(prims[size_t(i)].b);
            cb.grow(prims[size_t(i)].c);
        }
        node.b = bounds;
        int n = e - b;
        auto leaf = [&] {
            node.first = b;
            node.count = n;
            return self;
        };
        if (n == 1) return leaf();
        constexpr int kBins = 32;
        float bestCost = std::numeric_limits<float>::infinity();
        int bestAxis = -1, bestBin = 0;
        for (int a = 0; a < 3; ++a) {
            float cmin = cb.lo[a], cmax = cb.hi[a];
            if (!(cmax > cmin)) continue;
            float k = kBins / (cmax - cmin);
            Box bin[kBins];
            int cnt[kBins] = {0};
            for (int i = b; i < e; ++i) {
                int j = std::min(kBins - 1, int((prims[size_t(i)].c[a] - cmin) * k));
                cnt[j]++;
                bin[j].grow(prims[size_t(i)].b);
            }
            float rightArea[kBins];
            int rightCnt[kBins];
            Box acc;
            int c = 0;
            for (int j = kBins - 1; j > 0; --j) {
                acc.grow(bin[j]);
                c += cnt[j];
                rightArea[j] = acc.area();
                rightCnt[j] = c;
            }
            Box left;
            int lc = 0;
            for (int j = 0; j < kBins - 1; ++j) {
                left.grow(bin[j]);
                lc += cnt[j];
                if (lc == 0 || rightCnt[j + 1] == 0) continue;
                float cost = left.area() * float(lc) + rightArea[j + 1] * float(rightCnt[j + 1]);
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = a;
                    bestBin = j;
                }
            }
        }
        // A range becomes a leaf when it is small and a split would not pay: the SAH cost of the best
        // split, relative to one primitive test, plus the cost of one more traversal step.
        float area = bounds.area();
        float splitCost = kTraversalCost + (area > 0 ? bestCost / area : 0.f);
        if (n <= kLeafCandidate && (bestAxis < 0 || float(n) <= splitCost)) return leaf();
        int mid;
        if (bestAxis < 0) {
            if (n <= kMaxLeafPrims) return leaf();
            mid = (b + e) / 2;  // coincident centroids: split by index
        } else {
            float cmin = cb.lo[bestAxis], k = kBins / (cb.hi[bestAxis] - cmin);
            auto it = std::partition(prims.begin() + b, prims.begin() + e, [&](const BPrim &p) {
                return std::min(kBins - 1, int((p.c[bestAxis] - cmin) * k)) <= bestBin;
            });
            mid = int(it - prims.begin());
            if (mid == b || mid == e) mid = (b + e) / 2;
        }
        int l, r;
        if (n >= 65536 && depth < 6) {
            auto fut = std::async(std::launch::async, [&] { return Recurse(mid, e, depth + 1); });
            l = Recurse(b, mid, depth + 1);
            r = fut.get();
        } else {
            l = Recurse(b, mid, depth + 1);
            r = Recurse(mid, e, depth + 1);
        }
        nodes[size_t(self)].left = l;
        nodes[size_t(self)].right = r;
        return self;
    }
};

// Rounds a child interval [lo, hi] outwards onto the node grid origin + q * 2^e (exact product,
// one rounding in the add, as the kernel decodes it).
inline float Decode(float o, int q, float s) { return o + float(q) * s; }

class Collapser {
  public:
    const BinaryBuilder &bb;
    WideBVH &out;
    explicit Collapser(const BinaryBuilder &b, WideBVH &o) : bb(b), out(o) {}

    uint32_t Emit(int bn, int depth) {
        out.depth = std::max(out.depth, depth);
        int self = int(out.nodes.size());
        out.nodes.emplace_back();
        std::vector<int> kids;
        const BNode &root = bb.nodes[size_t(bn)];
        if (root.count > 0) {
            kids.push_back(bn);
        } else {
            kids.push_back(root.left);
            kids.push_back(root.right);
            while (kids.size() < 8) {
                int best = -1;
                float bestA = -1;
                for (size_t i = 0; i < kids.size(); ++i) {
                    const BNode &k = bb.nodes[size_t(kids[i])];
                    if (k.count > 0) continue;
                    float a = k.b.area();
                    if (a > bestA) {
                        bestA = a;
                        best = int(i);
                    }
                }
                if (best < 0) break;
                int open = kids[size_t(best)];
                kids[size_t(best)] = bb.nodes[size_t(open)].left;
                kids.insert(kids.begin() + best + 1, bb.nodes[size_t(open)].right);
            }
        }
        uint32_t refs[8] = {0};
        for (size_t i = 0; i < kids.size(); ++i) {
            const BNode &k = bb.nodes[size_t(kids[i])];
            if (k.count > 0) {
                uint32_t first = uint32_t(out.recs.size());
                for (int j = 0; j < k.count; ++j) out.recs.push_back(Record(bb.prims[size_t(k.first + j)].id));
                refs[i] = kLeafBit | (uint32_t(k.count - 1) << 27) | first;
            } else {
                refs[i] = Emit(kids[i], depth + 1);
            }
        }
        WNode &w = out.nodes[size_t(self)];
        std::memset(&w, 0, sizeof(w));
        w.nChildren = uint8_t(kids.size());
        for (size_t i = 0; i < kids.size(); ++i) w.child[i] = refs[i];
        Quantise(kids, &w);
        return uint32_t(self);
    }

    const prgpu::Primitive *prims = nullptr;

  private:
    TriRec Record(int id) const {
        const prgpu::Primitive &p = prims[size_t(id)];
        TriRec r{};
        if (p.type == prgpu::PrimSphere) {
            r.p0x = p.p0.x; r.p0y = p.p0.y; r.p0z = p.p0.z;
            r.e1x = p.r;
            r.prim = uint32_t(id) | kSphereBit;
        } else {
            r.p0x = p.p0.x; r.p0y = p.p0.y; r.p0z = p.p0.z;
            r.e1x = p.p1.x - p.p0.x; r.e1y = p.p1.y - p.p0.y; r.e1z = p.p1.z - p.p0.z;
            r.e2x = p.p2.x - p.p0.x; r.e2y = p.p2.y - p.p0.y; r.e2z = p.p2.z - p.p0.z;
            r.prim = uint32_t(id);
        }
        return r;
    }

    void Quantise(const std::vector<int> &kids, WNode *w) const {
        Box all;
        for (int k : kids) all.grow(bb.nodes[size_t(k)].b);
        float org[3] = {all.lo[0], all.lo[1], all.lo[2]};
        w->ox = org[0];
        w->oy = org[1];
        w->oz = org[2];
        for (int a = 0; a < 3; ++a) {
            float ext = all.hi[a] - org[a];
            int e = ext > 0 ? int(std::ceil(std::log2(double(ext) / 255.0))) : -100;
            e = std::clamp(e, -120, 127);
            for (;; ++e) {
                float s = std::ldexp(1.f, e);
                bool ok = true;
                uint8_t lo8[8] = {0}, hi8[8] = {0};
                for (size_t i = 0; i < kids.size() && ok; ++i) {
                    const Box &cb = bb.nodes[size_t(kids[i])].b;
                    int qlo = std::clamp(int(std::floor((cb.lo[a] - org[a]) / s)), 0, 255);
                    while (qlo > 0 && Decode(org[a], qlo, s) > cb.lo[a]) --qlo;
                    int qhi = std::clamp(int(std::ceil((cb.hi[a] - org[a]) / s)), 0, 255);
                    while (qhi < 255 && Decode(org[a], qhi, s) < cb.hi[a]) ++qhi;
                    if (Decode(org[a], qlo, s) > cb.lo[a] || Decode(org[a], qhi, s) < cb.hi[a]) ok = false;
                    lo8[i] = uint8_t(qlo);
                    hi8[i] = uint8_t(qhi);
                }
                if (ok || e >= 127) {
                    int8_t ee = int8_t(e);
                    if (a == 0) w->ex = ee;
                    if (a == 1) w->ey = ee;
                    if (a == 2) w->ez = ee;
                    for (size_t i = 0; i < kids.size(); ++i) {
                        w->qlo[a][i] = lo8[i];
                        w->qhi[a][i] = hi8[i];
                    }
                    break;
                }
            }
        }
    }
};

}  // namespace

void BuildWideBVH(const prgpu::Primitive *prims, int count, WideBVH *out) {
    out->nodes.clear();
    out->recs.clear();
    out->depth = 0;
    BinaryBuilder bb;
    bb.prims.resize(size_t(count));
    Box scene;
    for (int i = 0; i < count; ++i) {
        const prgpu::Primitive &p = prims[i];
        BPrim &bp = bb.prims[size_t(i)];
        if (p.type == prgpu::PrimSphere) {
            float c[3] = {p.p0.x, p.p0.y, p.p0.z};
            for (int a = 0; a < 3; ++a) {
                bp.b.lo[a] = c[a] - p.r;
                bp.b.hi[a] = c[a] + p.r;
            }
        } else {
            const prgpu::Float3 *v[3] = {&p.p0, &p.p1, &p.p2};
            for (auto *q : v) {
                float x[3] = {q->x, q->y, q->z};
                bp.b.grow(x);
            }
        }
        for (int a = 0; a < 3; ++a) bp.c[a] = (bp.b.lo[a] + bp.b.hi[a]) * 0.5f;
        bp.id = i;
        scene.grow(bp.b);
    }
    for (int a = 0; a < 3; ++a) {
        out->lo[a] = scene.lo[a];
        out->hi[a] = scene.hi[a];
    }
    if (count == 0) return;
    int root = bb.Build();
    Collapser c(bb, *out);
    c.prims = prims;
    out->nodes.reserve(size_t(bb.used.load()) / 4 + 1);
    out->recs.reserve(size_t(count));
    c.Emit(root, 1);
}

}  // namespace pgpu
