// Learned path guiding on gfx950 matrix cores: BF16 MFMA inference for every guided vertex of a
// bounce, and a fused BF16 forward / backward training kernel with FP32 accumulation, followed
// by an Adam step. The model and the mixture are described in guide.h.
//
// Matrix layout. Every layer is computed transposed, Y^T = W^T X^T, with 16x16x16 BF16 MFMAs:
// the A operand is a 16x16 block of Wt[out][in] (lane l holds row l % 16, columns
// 4 (l / 16) .. +3: one 8-byte LDS read), the B operand a 16-feature x 16-sample block of the
// layer input (lane l holds sample l % 16, features 4 (l / 16) .. +3), and the accumulator holds
// 16 outputs x 16 samples with lane l owning sample l % 16, outputs 4 (l / 16) .. +3: exactly
// the B-operand layout of the next layer, so activations stay in registers from layer to layer.
// A wave evaluates 16 vertices at a time; the 19 456 BF16 weights and 304 biases sit in LDS.
// LDS rows of weights (and of the training kernel's staging tiles) are padded by 8 bytes: the
// 16 lanes of an A-operand read take 16 consecutive rows, which at a 128-byte stride would all
// fall in one LDS bank (a 16-way conflict); at 136 bytes they cover the 32 banks once.
#include "guide_host.h"

#include <cmath>
#include <vector>

namespace pgpu {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kGBlock = 256;
constexpr int kInferBlock = 256;       // inference: 4 waves share one LDS copy of the model (3 blocks per CU)
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

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

constexpr int kGLdsStride = kGIn + 4;                 // bf16 per LDS weight row
constexpr int kGRows = kGWeights / kGIn;               // 304 weight rows, all layers
__host__ __device__ constexpr int lds_row(int l, int r) { return (l * kGHidden + r) * kGLdsStride; }

struct GuideLds {
    short w[kGRows * kGLdsStride];
    float b[kGBiases];
};

__device__ __forceinline__ void load_model(GuideLds &L, const short *wt, const float *bias) {
    for (int i = threadIdx.x; i < kGWeights / 4; i += blockDim.x)
        *reinterpret_cast<uint2 *>(&L.w[(i >> 4) * kGLdsStride + 4 * (i & 15)]) = reinterpret_cast<const uint2 *>(wt)[i];
    for (int i = threadIdx.x; i < kGBiases; i += blockDim.x) L.b[i] = bias[i];
}

__device__ __forceinline__ bf16x4 wblock(const GuideLds &L, int l, int j, int kb, int lane) {
    return *reinterpret_cast<const bf16x4 *>(&L.w[lds_row(l, j * 16 + (lane & 15)) + kb * 16 + 4 * (lane >> 4)]);
}

// Features of this lane's vertex (sample lane % 16) in B layout: lane group q = lane / 16 holds
// features 16 kb + 4 q .. +3, i.e. frequencies 2q and 2q + 1 (sin, cos) of axis kb for kb < 3,
// and for kb = 3 the normal and outgoing direction (q = 0, 1) or zeros.
__device__ __forceinline__ void features(const float *pos, const float *nrm, const float *wo, int lane, bf16x4 x[4]) {
    const int q = lane >> 4;
    const float f0 = float(1 << (2 * q)), f1 = 2.f * f0;
#pragma unroll
    for (int kb = 0; kb < 3; ++kb) {
        float s0, c0, s1, c1;
        sincospif(pos[kb] * f0, &s0, &c0);
        sincospif(pos[kb] * f1, &s1, &c1);
        x[kb][0] = f2bf(s0);
        x[kb][1] = f2bf(c0);
        x[kb][2] = f2bf(s1);
        x[kb][3] = f2bf(c1);
    }
    const bool g0 = q == 0, g1 = q == 1;
    x[3][0] = f2bf(g0 ? nrm[0] : (g1 ? wo[1] : 0.f));
    x[3][1] = f2bf(g0 ? nrm[1] : (g1 ? wo[2] : 0.f));
    x[3][2] = f2bf(g0 ? nrm[2] : 0.f);
    x[3][3] = f2bf(g0 ? wo[0] : 0.f);
}

// One hidden layer: x_out = relu(W x_in + b), B layout in and out.
__device__ __forceinline__ void hidden(const GuideLds &L, int l, const bf16x4 xin[4], bf16x4 xout[4], int lane) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma(wblock(L, l, j, kb, lane), xin[kb], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) xout[j][i] = f2bf(fmaxf(acc[i] + L.b[bl_off(l) + j * 16 + 4 * (lane >> 4) + i], 0.f));
    }
}

// The output layer: 48 raw outputs (rows 41..47 zero), accumulator layout.
__device__ __forceinline__ void output(const GuideLds &L, const bf16x4 xin[4], f32x4 out[3], int lane) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma(wblock(L, 4, j, kb, lane), xin[kb], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] += L.b[bl_off(4) + j * 16 + 4 * (lane >> 4) + i];
        out[j] = acc;
    }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ---- inference -------------------------------------------------------------------------------
// The decoded mixture (guide.h) for the queued vertices: queue[i] is a path slot whose hit record
// (position h0, geometric normal h1) and incoming direction d describe the vertex; the mixture
// goes to out[12 i ..]. Vertex positions are normalised to the box (lo, inv extent).
__global__ __launch_bounds__(kInferBlock) void k_guide_infer(const short *wt, const float *bias, const int *queue,
                                                             const int *count, const float4 *h0, const float4 *h1,
                                                             const float4 *dir, GuideBox box, float4 *out) {
    __shared__ GuideLds L;
    __shared__ float tile[kInferBlock / 64][16][kGTileRow];
    load_model(L, wt, bias);
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
        bf16x4 xa[4], xb[4];
        features(pos, nrm, wo, lane, xa);
        hidden(L, 0, xa, xb, lane);
        hidden(L, 1, xb, xa, lane);
        hidden(L, 2, xa, xb, lane);
        hidden(L, 3, xb, xa, lane);
        f32x4 o[3];
        output(L, xa, o, lane);
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
    __shared__ GuideLds L;
    load_model(L, wt, bias);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int waves = int(gridDim.x) * (kGBlock / 64);
    for (int tile = int(blockIdx.x) * (kGBlock / 64) + int(threadIdx.x >> 6); tile * 16 < n; tile += waves) {
        const int idx = tile * 16 + (lane & 15);
        const bool valid = idx < n;
        float r[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (valid)
            for (int k = 0; k < 9; ++k) r[k] = recs[size_t(idx) * 9 + k];
        bf16x4 xa[4], xb[4];
        features(r, r + 3, r + 6, lane, xa);
        hidden(L, 0, xa, xb, lane);
        hidden(L, 1, xb, xa, lane);
        hidden(L, 2, xa, xb, lane);
        hidden(L, 3, xb, xa, lane);
        f32x4 o[3];
        output(L, xa, o, lane);
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

constexpr int kGStage = 16 + 4;  // bf16 per staging row (16 samples, padded)

struct TrainLds {
    GuideLds m;
    float acc[kGParams];            // this block's gradient, in the parameter layout
    short stage[4][2][kGIn * kGStage];  // per wave: layer input X[in][s] and output gradient dY[out][s]
    float tile[4][16][kGTileRow];  // per wave: raw outputs, then their gradients
    float wsum, loss;
};

static_assert(sizeof(TrainLds) <= 160 * 1024, "training LDS exceeds a CU");

// dW_l += dY_l X_l^T over this wave's 16 samples and db_l += sum_s dY_l (LDS atomics), with the
// layer's input X (B layout, bf16) and output gradient dY (accumulator layout, fp32, nb blocks).
__device__ __forceinline__ void weight_grad(TrainLds &T, int wave, int lane, int l, const bf16x4 x[4], const f32x4 *dy,
                                            int nb) {
    short *sx = T.stage[wave][0], *sd = T.stage[wave][1];
    const int s = lane & 15, q = lane >> 4;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
        for (int i = 0; i < 4; ++i) sx[(kb * 16 + 4 * q + i) * kGStage + s] = x[kb][i];
    for (int j = 0; j < nb; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) sd[(j * 16 + 4 * q + i) * kGStage + s] = f2bf(dy[j][i]);
    wave_sync();
    float *gw = T.acc + guide_woff(l), *gb = T.acc + guide_boff(l);
    for (int j = 0; j < nb; ++j) {
        const bf16x4 a = *reinterpret_cast<const bf16x4 *>(&sd[(j * 16 + s) * kGStage + 4 * q]);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bf16x4 b = *reinterpret_cast<const bf16x4 *>(&sx[(c * 16 + s) * kGStage + 4 * q]);
            f32x4 d = mfma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int i = 0; i < 4; ++i) atomicAdd(&gw[(j * 16 + 4 * q + i) * kGIn + c * 16 + s], d[i]);
        }
    }
    if (lane < nb * 16) {
        float sum = 0.f;
        for (int k = 0; k < 16; ++k) sum += bf2f(sd[lane * kGStage + k]);
        atomicAdd(&gb[lane], sum);
    }
    wave_sync();
}

// dX = W^T dY (accumulator layout in and out), masked by the ReLU of the layer input x.
__device__ __forceinline__ void input_grad(const GuideLds &L, int l, const f32x4 *dy, int nb, const bf16x4 x[4],
                                           f32x4 dx[4], int lane) {
    bf16x4 b[4];
    for (int kb = 0; kb < nb; ++kb)
#pragma unroll
        for (int i = 0; i < 4; ++i) b[kb][i] = f2bf(dy[kb][i]);
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < nb; ++kb) {
            bf16x4 a;
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = L.w[lds_row(l, kb * 16 + 4 * (lane >> 4) + i) + jb * 16 + (lane & 15)];
            acc = mfma(a, b[kb], acc);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = bf2f(x[jb][i]) > 0.f ? acc[i] : 0.f;
        dx[jb] = acc;
    }
}

// One pass over n records (16 floats each): per block, the summed gradient of sum_r -wt_r log q_r
// (unnormalised), the summed weight and loss, written to part[block * (kGParams + 2) ..].
__global__ __launch_bounds__(kGBlock) void k_guide_train(const short *wt, const float *bias, const float *recs,
                                                         const int *count, int limit, int first, int n, float blend,
                                                         float *part) {
    extern __shared__ __align__(16) unsigned char smem[];
    TrainLds &T = *reinterpret_cast<TrainLds *>(smem);
    load_model(T.m, wt, bias);
    for (int i = threadIdx.x; i < kGParams; i += blockDim.x) T.acc[i] = 0.f;
    if (threadIdx.x == 0) T.wsum = T.loss = 0.f;
    __syncthreads();
    // Records [first, first + n), within the *count written (at most `limit`, the buffer's size).
    const int avail = count ? min(n, max(0, min(*count, limit) - first)) : n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, s = lane & 15;
    const int waves = int(gridDim.x) * (kGBlock / 64);
    for (int tile = int(blockIdx.x) * (kGBlock / 64) + wave; tile * 16 < avail; tile += waves) {
        const int idx = tile * 16 + s;
        const bool valid = idx < avail;
        float rec[kGRec];
        for (int k = 0; k < kGRec; ++k) rec[k] = valid ? recs[size_t(first + idx) * kGRec + k] : 0.f;
        bf16x4 x[5][4];
        features(rec, rec + 3, rec + 6, lane, x[0]);
        hidden(T.m, 0, x[0], x[1], lane);
        hidden(T.m, 1, x[1], x[2], lane);
        hidden(T.m, 2, x[2], x[3], lane);
        hidden(T.m, 3, x[3], x[4], lane);
        f32x4 o[3];
        output(T.m, x[4], o, lane);
        float(*tl)[kGTileRow] = T.tile[wave];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) tl[s][j * 16 + 4 * (lane >> 4) + i] = o[j][i];
        wave_sync();
        if (lane < 16) {
            float dout[kGOutRows], loss;
            output_grad(tl[s], rec, blend, dout, &loss);
            for (int i = 0; i < kGOutRows; ++i) tl[s][i] = dout[i];
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
        wave_sync();
        // Layer 4 (output, 48 rows), then the hidden layers 3..0.
        weight_grad(T, wave, lane, 4, x[4], dy, 3);
        f32x4 dx[4];
        input_grad(T.m, 4, dy, 3, x[4], dx, lane);
        for (int l = 3; l >= 0; --l) {
            weight_grad(T, wave, lane, l, x[l], dx, 4);
            if (l == 0) break;
            f32x4 nd[4];
            input_grad(T.m, l, dx, 4, x[l], nd, lane);
#pragma unroll
            for (int j = 0; j < 4; ++j) dx[j] = nd[j];
        }
    }
    __syncthreads();
    float *out = part + size_t(blockIdx.x) * (kGParams + 2);
    for (int i = threadIdx.x; i < kGParams; i += blockDim.x) out[i] = T.acc[i];
    if (threadIdx.x == 0) {
        out[kGParams] = T.wsum;
        out[kGParams + 1] = T.loss;
    }
}

// Sums the block partials: gradient / summed weight (grad), and summed weight and loss.
__global__ void k_guide_reduce(const float *part, int blocks, float *grad, float *stats) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i > kGParams + 1) return;
    float w = 0.f;
    for (int b = 0; b < blocks; ++b) w += part[size_t(b) * (kGParams + 2) + kGParams];
    float s = 0.f;
    for (int b = 0; b < blocks; ++b) s += part[size_t(b) * (kGParams + 2) + i];
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
