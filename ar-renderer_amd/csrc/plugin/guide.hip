// Learned path guiding on gfx950 matrix cores: BF16 MFMA inference for every guided vertex of a
// bounce, and a fused BF16 forward / backward training kernel with FP32 accumulation, followed
// by an Adam step. The model and the mixture are described in guide.h.
//
// Matrix layout. Every layer is computed transposed, Y^T = W^T X^T, with gfx950's 16x16x32 BF16
// MFMAs: the A operand is a 16-row block of Wt[out][in] (lane l holds row l % 16 and 8 of the 32
// K elements), the B operand a 32-feature x 16-vertex block of the layer input (lane l holds vertex
// l % 16 and 8 features), and the accumulator holds 16 outputs x 16 vertices with lane l owning
// vertex l % 16, outputs 4 (l / 16) .. +3. With the K order chosen below, the accumulators of two
// output blocks are, lane for lane, the B operand of the next layer: activations stay in registers
// from layer to layer; weights sit in LDS. A wave evaluates 16 vertices at a time.
// LDS rows of weights and staging tiles are padded by 16 bytes, so a row starts 16-byte aligned and
// the 16 rows an A-operand read takes do not all fall in the same LDS banks.
#include "guide_host.h"

#include <cmath>
#include <vector>

namespace pgpu {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kGBlock = 256;
constexpr int kInferBlock = 512;       // inference: 8 waves share one LDS copy of the model (2 blocks per CU)
constexpr int kGTileRow = kGOutRows + 1;  // floats per row of the raw-output tiles (padded)
constexpr int kGBiases = 4 * kGHidden + kGOutRows;  // 304
constexpr int kGRec = 16;                            // floats per training record

__host__ __device__ constexpr int wl_off(int l) { return l * kGHidden * kGIn; }  // bf16 weight array
__host__ __device__ constexpr int bl_off(int l) { return l * kGHidden; }         // bias array

#ifndef PG_GUIDE_NATIVE_BF16
#define PG_GUIDE_NATIVE_BF16 1
#endif
__device__ __forceinline__ short f2bf(float f) {  // round to nearest even
#if PG_GUIDE_NATIVE_BF16
    return __builtin_bit_cast(short, static_cast<__bf16>(f));  // gfx950: v_cvt_pk_bf16_f32
#else
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return short(u >> 16);
#endif
}
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float(uint32_t(uint16_t(h)) << 16); }

constexpr int kGRows = kGWeights / kGIn;  // 304 weight rows, all layers

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ---- forward pass on the gfx950 16x16x32 BF16 matrix cores ---------------------------------------
// The K dimension of one 16x16x32 MFMA is 32 wide; lane l supplies 8 of its elements (k = 8 (l / 16)
// .. +7). The accumulator of a layer holds, in lane l, rows 4 (l / 16) .. +3 of each 16-row output
// block j. Taking the K order of block m as {32 m + 4 q .. +3, 32 m + 16 + 4 q .. +3} for lane group
// q = l / 16 (any K order works when A and B agree), the accumulators of output blocks 2 m and
// 2 m + 1 are, lane for lane, the B operand of K-block m of the next layer. The weights are stored
// in LDS with every row's 64 columns in that order, so one 16-byte LDS read fetches a lane's A
// operand.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

constexpr int kG32Stride = kGIn + 8;  // bf16 per LDS weight row (16-byte padding: rows start 16-byte aligned)

struct GuideLds32 {
    alignas(16) short w[kGRows * kG32Stride];
    float b[kGBiases];
};

__host__ __device__ constexpr int row32(int l, int r) { return (l * kGHidden + r) * kG32Stride; }
// Stored column of input column c in the K order above.
__host__ __device__ constexpr int perm32(int c) {
    return (c & ~31) + 8 * ((c & 15) >> 2) + (c & 3) + ((c & 16) ? 4 : 0);
}

__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0,
                                                   0);
}

// Loads the BF16 weights (Wt[out][in], 4 consecutive columns per 8-byte word, which stay consecutive
// in the stored order) and the FP32 biases.
__device__ __forceinline__ void load_model32(GuideLds32 &L, const short *wt, const float *bias) {
    for (int i = threadIdx.x; i < kGWeights / 4; i += blockDim.x) {
        const int row = i >> 4, c = 4 * (i & 15);
        *reinterpret_cast<uint2 *>(&L.w[row * kG32Stride + perm32(c)]) = reinterpret_cast<const uint2 *>(wt)[i];
    }
    for (int i = threadIdx.x; i < kGBiases; i += blockDim.x) L.b[i] = bias[i];
}

__device__ __forceinline__ s16x8 wblock32(const GuideLds32 &L, int l, int j, int m, int lane) {
    return *reinterpret_cast<const s16x8 *>(&L.w[row32(l, j * 16 + (lane & 15)) + 32 * m + 8 * (lane >> 4)]);
}

__device__ __forceinline__ s16x8 pack8(f32x4 a, f32x4 b) {
    s16x8 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        r[i] = f2bf(a[i]);
        r[4 + i] = f2bf(b[i]);
    }
    return r;
}

// Features of this lane's vertex in the K order: B[0] = axis 0 then axis 1 frequencies 2q, 2q + 1
// (sin, cos each), B[1] = axis 2 frequencies, then features 48 + 4q .. 51 + 4q of the normal's
// encoding (guide.h: n, sin / cos of pi n per axis, sin / cos of 2 pi n per axis, a zero).
// Feature f of the model is element (f & 3) + 4 ((f >> 4) & 1) of block f >> 5.
__device__ __forceinline__ void features32(const float *pos, const float *nrm, const float *wo, int lane, s16x8 x[2]) {
    (void)wo;
    const int q = lane >> 4;
    const float f0 = float(1 << (2 * q)), f1 = 2.f * f0;
    float v[3][4];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sincospif(pos[a] * f0, &v[a][0], &v[a][1]);
        sincospif(pos[a] * f1, &v[a][2], &v[a][3]);
    }
    float s1[3], c1[3], s2[3], c2[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sincospif(nrm[a], &s1[a], &c1[a]);
        sincospif(2.f * nrm[a], &s2[a], &c2[a]);
    }
    // features 48..63: n0 n1 n2 s1[0] | c1[0] s1[1] c1[1] s1[2] | c1[2] s2[0] c2[0] s2[1] | c2[1] s2[2] c2[2] 0
    const float e[4] = {q == 0 ? nrm[0] : (q == 1 ? c1[0] : (q == 2 ? c1[2] : c2[1])),
                        q == 0 ? nrm[1] : (q == 1 ? s1[1] : (q == 2 ? s2[0] : s2[2])),
                        q == 0 ? nrm[2] : (q == 1 ? c1[1] : (q == 2 ? c2[0] : c2[2])),
                        q == 0 ? s1[0] : (q == 1 ? s1[2] : (q == 2 ? s2[1] : 0.f))};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        x[0][i] = f2bf(v[0][i]);
        x[0][4 + i] = f2bf(v[1][i]);
        x[1][i] = f2bf(v[2][i]);
        x[1][4 + i] = f2bf(e[i]);
    }
}

// One hidden layer: relu(W x + b), B layout (K order above) in and out. The accumulators start at
// the bias.
__device__ __forceinline__ void hidden32(const GuideLds32 &L, int l, const s16x8 xin[2], s16x8 xout[2], int lane) {
    f32x4 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float4 bj = *reinterpret_cast<const float4 *>(&L.b[bl_off(l) + j * 16 + 4 * (lane >> 4)]);
        acc[j] = f32x4{bj.x, bj.y, bj.z, bj.w};
    }
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = mfma32(wblock32(L, l, j, m, lane), xin[m], acc[j]);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[j][i] = fmaxf(acc[j][i], 0.f);
    xout[0] = pack8(acc[0], acc[1]);
    xout[1] = pack8(acc[2], acc[3]);
}

// The output layer: 48 raw outputs (rows 41..47 zero), accumulator layout.
__device__ __forceinline__ void output32(const GuideLds32 &L, const s16x8 xin[2], f32x4 out[3], int lane) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float4 bj = *reinterpret_cast<const float4 *>(&L.b[bl_off(4) + j * 16 + 4 * (lane >> 4)]);
        out[j] = f32x4{bj.x, bj.y, bj.z, bj.w};
    }
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int j = 0; j < 3; ++j) out[j] = mfma32(wblock32(L, 4, j, m, lane), xin[m], out[j]);
}

__device__ __forceinline__ void forward32(const GuideLds32 &L, const float *pos, const float *nrm, const float *wo,
                                          int lane, f32x4 o[3]) {
    s16x8 xa[2], xb[2];
    features32(pos, nrm, wo, lane, xa);
    hidden32(L, 0, xa, xb, lane);
    hidden32(L, 1, xb, xa, lane);
    hidden32(L, 2, xa, xb, lane);
    hidden32(L, 3, xb, xa, lane);
    output32(L, xa, o, lane);
}

// ---- inference -------------------------------------------------------------------------------
// The decoded mixture (guide.h) for the queued vertices: queue[i] is a path slot whose hit record
// (position h0, geometric normal h1) and incoming direction d describe the vertex; the mixture
// goes to out[12 i ..]. Vertex positions are normalised to the box (lo, inv extent).
__global__ __launch_bounds__(kInferBlock) void k_guide_infer(const short *wt, const float *bias, const int *queue,
                                                             const int *count, const float4 *h0, const float4 *h1,
                                                             const float4 *dir, GuideBox box, float4 *out) {
    __shared__ GuideLds32 L;
    __shared__ float tile[kInferBlock / 64][16][kGTileRow];
    load_model32(L, wt, bias);
    __syncthreads();
    const int n = *count, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int waves = int(gridDim.x) * (kInferBlock / 64);
    for (int tile_i = int(blockIdx.x) * (kInferBlock / 64) + wave; tile_i * 16 < n; tile_i += waves) {
        const int idx = tile_i * 16 + (lane & 15);
        const bool valid = idx < n;
        float pos[3] = {0.f, 0.f, 0.f}, nrm[3] = {0.f, 0.f, 0.f}, wo[3] = {0.f, 0.f, 0.f};
        if (valid) {
            const int slot = queue[idx];
            float4 p = h0[slot], nn = h1[slot], d = dir[slot];
            pos[0] = (p.x - box.lo[0]) * box.inv[0];
            pos[1] = (p.y - box.lo[1]) * box.inv[1];
            pos[2] = (p.z - box.lo[2]) * box.inv[2];
            nrm[0] = nn.x, nrm[1] = nn.y, nrm[2] = nn.z;
            float il = -1.f / sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
            wo[0] = d.x * il, wo[1] = d.y * il, wo[2] = d.z * il;
        }
        f32x4 o[3];
        forward32(L, pos, nrm, wo, lane, o);
        // Decode the mixture with all 64 lanes: the four lanes of a vertex (lane % 16) take two lobes
        // each and share the softmax normaliser and the lobe CDF through shuffles.
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) tile[wave][lane & 15][j * 16 + 4 * (lane >> 4) + i] = o[j][i];
        wave_sync();
        {
            const int sidx = lane & 15, q = lane >> 4, k0 = 2 * q;
            const float *t = tile[wave][sidx];
            const float a0 = t[5 * k0 + 4], a1 = t[5 * k0 + 9];
            float amax = fmaxf(a0, a1);
            amax = fmaxf(amax, __shfl_xor(amax, 16));
            amax = fmaxf(amax, __shfl_xor(amax, 32));
            const float w0 = expf(a0 - amax), w1 = expf(a1 - amax), pair = w0 + w1;
            const float p0 = __shfl(pair, sidx), p1 = __shfl(pair, sidx + 16), p2 = __shfl(pair, sidx + 32),
                        p3 = __shfl(pair, sidx + 48);
            const float z = p0 + p1 + p2 + p3;
            const float excl = (q > 0 ? p0 : 0.f) + (q > 1 ? p1 : 0.f) + (q > 2 ? p2 : 0.f);
            const int idx = tile_i * 16 + sidx;
            if (idx < n) {
                float *of = reinterpret_cast<float *>(out + size_t(idx) * 12);
                const float w[2] = {w0, w1};
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    V3 mu;
                    float kappa;
                    guide_lobe(t, k0 + h, &mu, &kappa);
                    const float pi = w[h] / z;
                    reinterpret_cast<float4 *>(of)[k0 + h] = make_float4(mu.x, mu.y, mu.z, kappa);
                    of[32 + k0 + h] = pi * kappa / (2.f * kPi * -expm1f(-2.f * kappa));
                    if (k0 + h < kGLobes - 1) of[40 + k0 + h] = (excl + w0 + (h ? w1 : 0.f)) / z;
                }
                if (q == 3) of[47] = kGSelMin + (kGSelMax - kGSelMin) / (1.f + expf(-t[40]));
            }
        }
        wave_sync();
    }
}

// Raw outputs for explicit records (9 floats: normalised position, normal, outgoing direction).
__global__ __launch_bounds__(kGBlock) void k_guide_eval(const short *wt, const float *bias, const float *recs, int n,
                                                        float *out41) {
    __shared__ GuideLds32 L;
    load_model32(L, wt, bias);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int waves = int(gridDim.x) * (kGBlock / 64);
    for (int tile = int(blockIdx.x) * (kGBlock / 64) + int(threadIdx.x >> 6); tile * 16 < n; tile += waves) {
        const int idx = tile * 16 + (lane & 15);
        const bool valid = idx < n;
        float r[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (valid)
            for (int k = 0; k < 9; ++k) r[k] = recs[size_t(idx) * 9 + k];
        f32x4 o[3];
        forward32(L, r, r + 3, r + 6, lane, o);
        if (valid)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = j * 16 + 4 * (lane >> 4) + i;
                    if (row < kGOut) out41[size_t(idx) * kGOut + row] = o[j][i];
                }
    }
}

// ---- training ----------------------------------------------------------------------------------
// d(-wt (b log q + (1 - b) log p_guide)) / d(raw outputs) for one record, b = blend, as the FP64
// model defines it (the weight is not yet normalised by the batch's summed weight).
__device__ void output_grad(const float *o, const float *rec, float blend, float *dout, float *lossOut) {
    for (int i = 0; i < kGOutRows; ++i) dout[i] = 0.f;
    const float wt = rec[13], pb = rec[12];
    *lossOut = 0.f;
    if (!(wt > 0.f)) return;
    const V3 w(rec[9], rec[10], rec[11]);
    GuideRaw m = guide_raw(o);
    float pk[kGLobes], pm = 0.f, pi[kGLobes];
    for (int k = 0; k < kGLobes; ++k) {
        V3 mu;
        float kappa;
        guide_lobe(o, k, &mu, &kappa);
        pk[k] = vmf_pdf(mu, kappa, w);
        pi[k] = expf(o[5 * k + 4] - m.amax) / m.z;
        pm += pi[k] * pk[k];
    }
    const float q = m.sel * pm + (1.f - m.sel) * pb;
    if (!(q > 0.f) || !(pm > 0.f)) return;
    *lossOut = -wt * (blend * logf(q) + (1.f - blend) * logf(pm));
    const float dq = -wt * blend / q;
    const float sig = (m.sel - kGSelMin) / (kGSelMax - kGSelMin);
    dout[40] = dq * (pm - pb) * (kGSelMax - kGSelMin) * sig * (1.f - sig);
    const float dpm = dq * m.sel - wt * (1.f - blend) / pm;
    for (int k = 0; k < kGLobes; ++k) {
        const float *ok = o + 5 * k;
        dout[5 * k + 4] = dpm * pi[k] * (pk[k] - pm);
        V3 v(ok[0], ok[1], ok[2]);
        const float vl = sqrtf(dot(v, v));
        V3 mu = vl > 1e-6f ? v / vl : V3(0.f, 0.f, 1.f);
        const float sp = kGKappaMin + softplusf(ok[3]), kappa = fminf(sp, kGKappaMax);
        const float c = dot(mu, w), dpk = dpm * pi[k];
        const float dk = pk[k] * (1.f / kappa - 2.f * expf(-2.f * kappa) / -expm1f(-2.f * kappa) + (c - 1.f));
        if (sp < kGKappaMax) dout[5 * k + 3] = dpk * dk / (1.f + expf(-ok[3]));
        if (vl > 1e-6f) {
            const float s = dpk * pk[k] * kappa / vl;
            dout[5 * k + 0] = s * (w.x - c * mu.x);
            dout[5 * k + 1] = s * (w.y - c * mu.y);
            dout[5 * k + 2] = s * (w.z - c * mu.z);
        }
    }
}

// Training, one block per CU, 4 waves, 64 records per pass (16 per wave):
//  - forward with the 16x16x32 layers above, every layer's input kept in registers;
//  - the loss gradient of the 48 raw outputs per record (output_grad);
//  - per layer, from the output: the layer input X and output gradient dY of the block's 64
//    records are staged in LDS as [feature][record]; the weight gradient dW = dY X^T (K = the 64
//    records) is an MFMA product whose 16x16 tiles are split among the 4 waves and accumulated in
//    their registers over all passes (no atomics); the input gradient dX = W^T dY (a second,
//    transposed copy of the weights in LDS, columns in the K order) masked by the ReLU, in registers.
// At the end every wave writes its tiles, and 64 threads their bias sums, to the block's partial.
constexpr int kTWaves = 4;
constexpr int kTRecs = 16 * kTWaves;   // records per pass of a block
constexpr int kTStage = kTRecs + 8;    // bf16 per staging row (64 records, 16-byte padding)
constexpr int kTTiles[kGLayers] = {16, 16, 16, 16, 12};  // 16x16 weight-gradient tiles per layer

struct TrainLds {
    GuideLds32 m;                                   // Wt[out][in], columns in the K order
    alignas(16) short wT[4][kGIn * kG32Stride];     // layers 1..4: W^T[in][out], out in the K order
    alignas(16) short sx[kGIn * kTStage];           // X_l[in][record]
    alignas(16) short sd[kGIn * kTStage];           // dY_l[out][record]
    float tile[kTWaves][16][kGTileRow];             // raw outputs, then their gradients
    float wsum, loss;
};

static_assert(sizeof(TrainLds) <= 160 * 1024, "training LDS exceeds a CU");

__device__ __forceinline__ s16x8 lds8(const short *p) { return *reinterpret_cast<const s16x8 *>(p); }

// Stages this wave's X (B layout: K order of the 64 inputs) and dY (accumulator layout, nb blocks)
// as [feature][record] rows; record column = 16 wave + s.
__device__ __forceinline__ void stage(TrainLds &T, int wave, int lane, const s16x8 x[2], const f32x4 *dy, int nb) {
    const int s = 16 * wave + (lane & 15), q = lane >> 4;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int i = 0; i < 8; ++i) T.sx[(32 * m + 4 * q + (i & 3) + (i & 4 ? 16 : 0)) * kTStage + s] = x[m][i];
    for (int j = 0; j < nb; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) T.sd[(16 * j + 4 * q + i) * kTStage + s] = f2bf(dy[j][i]);
    for (int j = nb; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) T.sd[(16 * j + 4 * q + i) * kTStage + s] = 0;
}

// This wave's share of dW_l += dY X^T over the staged records: tiles t = wave, wave + 4, .. of the
// layer's (rows / 16) x 4 tiles, tile t = (row block t / 4, column block t % 4).
__device__ __forceinline__ void weight_grad32(const TrainLds &T, int l, int wave, int lane, f32x4 *acc) {
    const int r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int t = wave + 4 * k;
        if (t >= kTTiles[l]) break;
        const int jb = t >> 2, cb = t & 3;
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            const s16x8 a = lds8(&T.sd[(16 * jb + r) * kTStage + 32 * kb + 8 * q]);
            const s16x8 b = lds8(&T.sx[(16 * cb + r) * kTStage + 32 * kb + 8 * q]);
            acc[k] = mfma32(a, b, acc[k]);
        }
    }
}

// dX = W_l^T dY, accumulator layout in (nb blocks of dY rows) and out (4 blocks of 16 inputs),
// masked by the ReLU of the layer input x (B layout).
__device__ __forceinline__ void input_grad32(const TrainLds &T, int l, const f32x4 *dy, int nb, const s16x8 x[2],
                                             f32x4 dx[4], int lane) {
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    const s16x8 b0 = pack8(dy[0], dy[1]), b1 = pack8(dy[2], nb > 3 ? dy[3] : zero);
    const short *w = T.wT[l - 1];
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) {
        const short *row = &w[(16 * jb + (lane & 15)) * kG32Stride + 8 * (lane >> 4)];
        f32x4 acc = mfma32(lds8(row), b0, zero);
        acc = mfma32(lds8(row + 32), b1, acc);
        const int m = jb >> 1, h = 4 * (jb & 1);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = bf2f(x[m][h + i]) > 0.f ? acc[i] : 0.f;
        dx[jb] = acc;
    }
}

// One pass over records [first, first + n) (16 floats each, within the *count written when count
// is given, at most `limit`): per block, the summed gradient of the unnormalised loss (see
// output_grad), the summed weight and loss, written to part[block * (kGParams + 2) ..].
__global__ __launch_bounds__(kGBlock) void k_guide_train(const short *wt, const float *bias, const float *recs,
                                                         const int *count, int limit, int first, int n, float blend,
                                                         float *part) {
    extern __shared__ __align__(16) unsigned char smem[];
    TrainLds &T = *reinterpret_cast<TrainLds *>(smem);
    load_model32(T.m, wt, bias);
    // The transposed weights of layers 1..4 (rows = inputs, columns = outputs in the K order; the
    // output layer's outputs 48..63 are zero).
    for (int i = threadIdx.x; i < 4 * kGIn * kGIn; i += blockDim.x) {
        const int l = 1 + i / (kGIn * kGIn), e = i % (kGIn * kGIn), out = e / kGIn, in = e % kGIn;
        const short v = out < guide_rows(l) ? wt[wl_off(l) + out * kGIn + in] : short(0);
        T.wT[l - 1][in * kG32Stride + perm32(out)] = v;
    }
    if (threadIdx.x == 0) T.wsum = T.loss = 0.f;
    __syncthreads();
    const int avail = count ? min(n, max(0, min(*count, limit) - first)) : n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, s = lane & 15;
    f32x4 gacc[kGLayers][4];
#pragma unroll
    for (int l = 0; l < kGLayers; ++l)
#pragma unroll
        for (int k = 0; k < 4; ++k) gacc[l][k] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bacc[kGLayers] = {0.f, 0.f, 0.f, 0.f, 0.f};  // thread t < 64: bias gradient of output t of each layer
    for (int pass = int(blockIdx.x); pass * kTRecs < avail; pass += int(gridDim.x)) {
        const int idx = pass * kTRecs + 16 * wave + s;
        const bool valid = idx < avail;
        float rec[kGRec];
        for (int k = 0; k < kGRec; ++k) rec[k] = valid ? recs[size_t(first + idx) * kGRec + k] : 0.f;
        s16x8 x[kGLayers][2];
        features32(rec, rec + 3, rec + 6, lane, x[0]);
        hidden32(T.m, 0, x[0], x[1], lane);
        hidden32(T.m, 1, x[1], x[2], lane);
        hidden32(T.m, 2, x[2], x[3], lane);
        hidden32(T.m, 3, x[3], x[4], lane);
        f32x4 o[3];
        output32(T.m, x[4], o, lane);
        float(*tl)[kGTileRow] = T.tile[wave];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) tl[s][j * 16 + 4 * (lane >> 4) + i] = o[j][i];
        wave_sync();
        if (lane < 16) {
            float dout[kGOutRows], loss;
            output_grad(tl[s], rec, blend, dout, &loss);
            for (int i = 0; i < kGOutRows; ++i) tl[s][i] = valid ? dout[i] : 0.f;
            if (valid && rec[13] > 0.f) {
                atomicAdd(&T.wsum, rec[13]);
                atomicAdd(&T.loss, loss);
            }
        }
        wave_sync();
        f32x4 dy[4];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) dy[j][i] = tl[s][j * 16 + 4 * (lane >> 4) + i];
        dy[3] = f32x4{0.f, 0.f, 0.f, 0.f};
        // Layers 4 .. 0: stage, weight and bias gradients, input gradient.
#pragma unroll
        for (int l = kGLayers - 1; l >= 0; --l) {
            const int nb = l == kGLayers - 1 ? 3 : 4;
            __syncthreads();  // the previous layer's staging has been read
            stage(T, wave, lane, x[l], dy, nb);
            __syncthreads();
            weight_grad32(T, l, wave, lane, gacc[l]);
            if (threadIdx.x < guide_rows(l)) {
                float sum = 0.f;
                const short *row = &T.sd[threadIdx.x * kTStage];
                for (int k = 0; k < kTRecs; k += 8) {
                    const s16x8 v = lds8(row + k);
#pragma unroll
                    for (int i = 0; i < 8; ++i) sum += bf2f(v[i]);
                }
                bacc[l] += sum;
            }
            if (l > 0) {
                f32x4 dx[4];
                input_grad32(T, l, dy, nb, x[l], dx, lane);
#pragma unroll
                for (int j = 0; j < 4; ++j) dy[j] = dx[j];
            }
        }
    }
    // Tiles: lane l holds column 16 cb + l % 16 (input) and rows 16 jb + 4 (l / 16) .. +3 (outputs).
    float *out = part + size_t(blockIdx.x) * (kGParams + 2);
    const int q = lane >> 4;
#pragma unroll
    for (int l = 0; l < kGLayers; ++l)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int t = wave + 4 * k;
            if (t >= kTTiles[l]) break;
            const int jb = t >> 2, cb = t & 3;
#pragma unroll
            for (int i = 0; i < 4; ++i) out[guide_woff(l) + (16 * jb + 4 * q + i) * kGIn + 16 * cb + s] = gacc[l][k][i];
        }
    if (threadIdx.x < kGIn)
#pragma unroll
        for (int l = 0; l < kGLayers; ++l)
            if (int(threadIdx.x) < guide_rows(l)) out[guide_boff(l) + threadIdx.x] = bacc[l];
    __syncthreads();
    if (threadIdx.x == 0) {
        out[kGParams] = T.wsum;
        out[kGParams + 1] = T.loss;
    }
}

// Sums the block partials: gradient / summed weight (grad), and summed weight and loss. Each block
// first reduces the summed weight over the partials, then one thread per parameter sums its column
// of the partials with eight independent accumulators.
__global__ __launch_bounds__(256) void k_guide_reduce(const float *part, int blocks, float *grad, float *stats) {
    __shared__ float ws[256];
    float w = 0.f;
    for (int b = int(threadIdx.x); b < blocks; b += 256) w += part[size_t(b) * (kGParams + 2) + kGParams];
    ws[threadIdx.x] = w;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) ws[threadIdx.x] += ws[threadIdx.x + h];
        __syncthreads();
    }
    w = ws[0];
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i > kGParams + 1) return;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int b = 0;
    for (; b + 8 <= blocks; b += 8)
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] += part[size_t(b + k) * (kGParams + 2) + i];
    for (; b < blocks; ++b) acc[0] += part[size_t(b) * (kGParams + 2) + i];
    const float s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    if (i < kGParams) grad[i] = w > 0.f ? s / w : 0.f;
    else stats[i - kGParams] = s;
}

// Adam (Kingma and Ba 2015) on the FP32 master parameters; refreshes the BF16 weights and FP32 biases.
__global__ void k_guide_adam(const float *grad, float *P, float *m, float *v, short *wt, float *bias, float lr,
                             float c1, float c2) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= kGParams) return;
    const float g = grad[i];
    const float mi = 0.9f * m[i] + 0.1f * g, vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi;
    v[i] = vi;
    const float p = P[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    P[i] = p;
    // Map the parameter index to its bf16 weight or bias slot.
    int l = 0;
    while (l < kGLayers - 1 && i >= guide_woff(l + 1)) ++l;
    const int r = i - guide_woff(l);
    if (r < guide_rows(l) * kGIn) wt[wl_off(l) + r] = f2bf(p);
    else bias[bl_off(l) + (r - guide_rows(l) * kGIn)] = p;
}

__global__ void k_guide_split(const float *P, short *wt, float *bias) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= kGParams) return;
    int l = 0;
    while (l < kGLayers - 1 && i >= guide_woff(l + 1)) ++l;
    const int r = i - guide_woff(l);
    if (r < guide_rows(l) * kGIn) wt[wl_off(l) + r] = f2bf(P[i]);
    else bias[bl_off(l) + (r - guide_rows(l) * kGIn)] = P[i];
}

constexpr int kTrainBlocks = 256;

}  // namespace

// ---- host ---------------------------------------------------------------------------------------

size_t GuideTrainLdsBytes() { return sizeof(TrainLds); }

// Initial parameters: He-uniform hidden layers, a small output layer, zero biases except the
// selection output, so the first proposal is a set of broad lobes sampled with s = 0.2 (most
// samples still follow the BSDF until training sharpens the mixture). Deterministic (PCG32
// stream 54).
std::vector<float> GuideInitialParams() {
    std::vector<float> P(size_t(kGParams), 0.f);
    uint64_t state = 0x853c49e6748fea9bull, inc = (54ull << 1u) | 1u;
    auto next = [&]() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u), rot = uint32_t(old >> 59u);
        uint32_t r = (xs >> rot) | (xs << ((32u - rot) & 31u));
        return float(r) * 0x1p-32f;
    };
    for (int l = 0; l < kGLayers; ++l) {
        const float a = l < kGLayers - 1 ? std::sqrt(6.f / kGIn) : 0.1f * std::sqrt(6.f / kGIn);
        const int rows = l < kGLayers - 1 ? kGHidden : kGOut;
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < kGIn; ++c) P[size_t(guide_woff(l) + r * kGIn + c)] = (2.f * next() - 1.f) * a;
    }
    // s = kGSelMin + (kGSelMax - kGSelMin) sigmoid(b) = 0.2
    // Initial selection probability: an eighth of the way from its floor to its ceiling.
    P[size_t(guide_boff(kGLayers - 1) + 40)] = std::log(0.125f / 0.875f);
    return P;
}

void GuideState::SetParams(const float *P, hipStream_t s) {
    params.upload(P, size_t(kGParams), s);
    m.ensure(size_t(kGParams) * sizeof(float));
    v.ensure(size_t(kGParams) * sizeof(float));
    PG_CHECK(hipMemsetAsync(m.p, 0, size_t(kGParams) * sizeof(float), s));
    PG_CHECK(hipMemsetAsync(v.p, 0, size_t(kGParams) * sizeof(float), s));
    wt.ensure(size_t(kGWeights) * sizeof(short));
    bias.ensure(size_t(kGBiases) * sizeof(float));
    hipLaunchKernelGGL(k_guide_split, dim3((kGParams + 255) / 256), dim3(256), 0, s, params.as<float>(),
                       wt.as<short>(), bias.as<float>());
    PG_CHECK(hipGetLastError());
    steps = 0;
}

void GuideState::Init(hipStream_t s) {
    std::vector<float> P = GuideInitialParams();
    SetParams(P.data(), s);
    grad.ensure(size_t(kGParams + 2) * sizeof(float));
    stats.ensure(4 * sizeof(float));
    part.ensure(size_t(kTrainBlocks) * (kGParams + 2) * sizeof(float));
    trainCount.ensure(2 * sizeof(int));
    PG_CHECK(hipMemsetAsync(trainCount.p, 0, 2 * sizeof(int), s));
    static bool attr = false;
    if (!attr) {
        PG_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(k_guide_train),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(TrainLds))));
        attr = true;
    }
}

void GuideState::EnsureBatch(size_t n) {
    if (cap >= n) return;
    out.ensure(n * 48 * sizeof(float));
    rec.ensure(n * 6 * size_t(kGuideRecords) * sizeof(float4));
    cap = n;
}

void GuideState::EnsureTrain(size_t n) {
    if (trainCap >= n) return;
    train.ensure(n * kGRec * sizeof(float));
    trainCap = n;
    ringHead = ringValid = 0;  // the ring starts over in the new buffer
}

void GuideState::Infer(const int *queue, const int *count, int maxCount, const float4 *h0, const float4 *h1,
                       const float4 *dir, const GuideBox &box, int cus, hipStream_t s) {
    if (maxCount <= 0) return;
    const int tiles = (maxCount + 15) / 16;
    const int blocks = std::max(1, std::min((tiles + kInferBlock / 64 - 1) / (kInferBlock / 64), cus * 8));
    hipLaunchKernelGGL(k_guide_infer, dim3(blocks), dim3(kInferBlock), 0, s, wt.as<short>(), bias.as<float>(), queue,
                       count, h0, h1, dir, box, out.as<float4>());
    PG_CHECK(hipGetLastError());
}

void GuideState::Gradient(const float *recs, const int *count, int first, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_train, dim3(kTrainBlocks), dim3(kGBlock), sizeof(TrainLds), s, wt.as<short>(),
                       bias.as<float>(), recs, count, count ? int(trainCap) : n, first, n, blend, part.as<float>());
    PG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_guide_reduce, dim3((kGParams + 2 + 255) / 256), dim3(256), 0, s, part.as<float>(),
                       kTrainBlocks, grad.as<float>(), stats.as<float>());
    PG_CHECK(hipGetLastError());
}

void GuideState::Step(const float *recs, const int *count, int first, int n, float lr, hipStream_t s) {
    Gradient(recs, count, first, n, s);
    ++steps;
    const float c1 = 1.f - std::pow(0.9f, float(steps)), c2 = 1.f - std::pow(0.999f, float(steps));
    hipLaunchKernelGGL(k_guide_adam, dim3((kGParams + 255) / 256), dim3(256), 0, s, grad.as<float>(),
                       params.as<float>(), m.as<float>(), v.as<float>(), wt.as<short>(), bias.as<float>(), lr, c1, c2);
    PG_CHECK(hipGetLastError());
}

void GuideState::Eval(const float *recs9, int n, float *out41, hipStream_t s) {
    if (n <= 0) return;
    const int blocks = std::max(1, std::min((n + 63) / 64, 1024));
    hipLaunchKernelGGL(k_guide_eval, dim3(blocks), dim3(kGBlock), 0, s, wt.as<short>(), bias.as<float>(), recs9, n,
                       out41);
    PG_CHECK(hipGetLastError());
}

}  // namespace pgpu
