// hulk_mlp.hip - learned hit-distance bound: BF16 MFMA forward / backward
// kernels with FP32 accumulation, Adam, and the fused per-pixel bound.
//
// Network (no reference counterpart):
//   x  = ((o - c) * (1/h), d * (1/|d|))        c, h: scene box centre / half size
//   z  = F x (F: 32 x 6, fixed), e = (sin z, cos z) (64)
//   h1 = relu(W1 e + b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3)   (128 each)
//   y  = w4 . h3 + b4 = predicted log t_hit; bound = exp(y + margin)
// Weights, e and activations are BF16; every product accumulates in FP32
// (v_mfma_f32_16x16x32_bf16).  Master weights and Adam state are FP32.
//
// MFMA fragment convention (16x16x32, wave64): lane l, g = l / 16,
// r = l % 16.  A fragment: A[row r][k 8g .. 8g+7]; B fragment:
// B[k 8g .. 8g+7][col r]; accumulator: C[row 4g + i][col r], i = 0..3.
// A and B use the same lane -> k map, so the k order inside the
// instruction does not change the sum.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "hulk_abi.h"
#include "hulk_device.hpp"

namespace hulk {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kIn = 64, kHid = 128;
// flat parameter offsets: W1, b1, W2, b2, W3, b3, w4, b4
constexpr uint32_t oW1 = 0, ob1 = oW1 + kHid * kIn, oW2 = ob1 + kHid, ob2 = oW2 + kHid * kHid,
                   oW3 = ob2 + kHid, ob3 = oW3 + kHid * kHid, ow4 = ob3 + kHid, ob4 = ow4 + kHid,
                   kParams = ob4 + 1;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);  // NaN
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 pack8(const uint16_t* h) {
    uint4 v;
    v.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
    v.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
    v.z = (uint32_t)h[4] | ((uint32_t)h[5] << 16);
    v.w = (uint32_t)h[6] | ((uint32_t)h[7] << 16);
    return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct FwdParams {
    uint32_t camera;
    Camera cam;
    const Camera* cams;    // bound mode over several frames: camera of frame f (null: cam)
    uint32_t per_frame;    // evaluations per frame (bound mode)
    uint32_t pf_m, wq_m;   // division multipliers of per_frame and the evaluations per row (udiv)
    uint32_t W, H;
    const hulk_ray* rays;
    uint32_t n;
    float c[3], ih[3];
    const float* fourier;  // 32 x 6
    const uint4* Wf;       // W1, W2, W3 as 80 MFMA operand fragments (k_refresh)
    const uint16_t* w4;
    const float* p;        // master params (biases)
    float* pred;
    float* tbound;
    float margin, max_log;
    uint32_t stride;                   // bound mode: one evaluation per stride x stride pixels
    uint16_t *encT, *h1T, *h2T, *h3T;  // [feature][ld] (training)
    uint32_t ld;
};

// Ray tiles of 16 per wave and group: 32 rays share each weight fragment
// read from LDS; at 4 tiles the activations of two layers (8 x 4 x 2
// VGPRs each) no longer fit in 256 VGPRs beside the accumulators.
constexpr uint32_t kTiles = 2;
constexpr uint32_t kTrainTiles = 1;  // training also writes every activation: one tile keeps it out of scratch
constexpr uint32_t kFwdWaves = 8;  // waves per forward block (two per SIMD)
// Weight fragments in MFMA operand order: fragment f of a layer is the A
// operand of output tile t = f / chunks and K chunk kk = f % chunks; lane l
// holds 8 contiguous bf16 (16 B), so a wave reads a fragment as 1 KiB.
constexpr uint32_t kFrag1 = 8 * 2, kFrag2 = 8 * 4, kFragAll = kFrag1 + 2 * kFrag2;  // 80 fragments, 80 KiB

// Input feature of K position p (0..7) of lane group g in K chunk kk of
// layers 2 and 3.  The accumulator of the layer before holds neurons
// 16 t + 4 g + i (i = 0..3) of output tile t in lane group g, so a B
// fragment made of tiles 2 kk and 2 kk + 1 carries neurons 32 kk + 4 g + i
// and 32 kk + 16 + 4 g + i; the weights are stored with the same K order,
// which keeps the activations in registers from layer to layer.
__host__ __device__ __forceinline__ uint32_t hidden_k(uint32_t kk, uint32_t g, uint32_t p) {
    return 32u * kk + (p < 4u ? 4u * g + p : 16u + 4u * g + (p - 4u));
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// a / d for a divisor d >= 1 that is the same for every lane, with
// m = floor((2^32 - 1) / d) from the host (udiv_multiplier): umulhi(a, m)
// is at most two below the quotient.
__host__ __device__ inline uint32_t udiv_multiplier(uint32_t d) { return 0xFFFFFFFFu / d; }
__device__ __forceinline__ uint32_t udiv(uint32_t a, uint32_t d, uint32_t m) {
    uint32_t q = __umulhi(a, m);
    while (a - q * d >= d) q++;
    return q;
}

// Two floats to packed BF16, round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t cvt2(float lo, float hi) {
    const bf16x2 v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
}
// acc + 2 products of packed BF16 pairs in FP32 (v_dot2_f32_bf16).
__device__ __forceinline__ float dot2(uint32_t a, uint32_t b, float acc) {
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, a), __builtin_bit_cast(bf16x2, b), acc, false);
}
__device__ __forceinline__ bf16x8 join(uint2 lo, uint2 hi) {
    return __builtin_bit_cast(bf16x8, make_uint4(lo.x, lo.y, hi.x, hi.y));
}
__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// ReLU + BF16 of accumulator rows 16 t + 4 g + (0..3) of one ray tile (the
// accumulator starts at the bias); the feature-major copy goes to hT when
// training.
__device__ __forceinline__ uint2 activate(const f32x4& acc, uint16_t* hT, uint32_t ld, uint32_t row0, uint32_t ray,
                                          bool valid) {
    const uint2 h = make_uint2(cvt2(fmaxf(acc[0], 0.0f), fmaxf(acc[1], 0.0f)),
                               cvt2(fmaxf(acc[2], 0.0f), fmaxf(acc[3], 0.0f)));
    if (hT && valid) {
        hT[(size_t)(row0 + 0) * ld + ray] = (uint16_t)h.x;
        hT[(size_t)(row0 + 1) * ld + ray] = (uint16_t)(h.x >> 16);
        hT[(size_t)(row0 + 2) * ld + ray] = (uint16_t)h.y;
        hT[(size_t)(row0 + 3) * ld + ray] = (uint16_t)(h.y >> 16);
    }
    return h;
}

// Fused forward.  A block of kFwdWaves waves stages the 80 weight
// fragments (W1, W2, W3 in MFMA operand order, 80 KiB) in LDS once; each
// wave then takes groups of 64 rays (4 tiles of 16) until none is left.
// Per group: the Fourier encoding in registers (B operand of layer 1),
// then three layers whose activations never leave the registers (see
// hidden_k); layer 3 folds the output dot with w4.  Every weight fragment
// read from LDS feeds 4 MFMAs.  Bound mode with stride s evaluates the ray
// of pixel (s*qx, s*qy) and writes its bound to the s x s pixel block.
template <bool TRAIN, uint32_t TILES>
__global__ void __launch_bounds__(64 * kFwdWaves) k_mlp_fwd(FwdParams P) {
    __shared__ uint4 wl[kFragAll * 64];
    const uint32_t tid = threadIdx.x;
    const uint32_t l = tid & 63u, g = l >> 4, r = l & 15u;
    for (uint32_t i = tid; i < kFragAll * 64; i += 64 * kFwdWaves) wl[i] = P.Wf[i];
    __syncthreads();
    const uint32_t groups = (P.n + 16u * TILES - 1u) / (16u * TILES);
    const uint32_t s = P.stride ? P.stride : 1u;
    const uint32_t Wq = (P.W + s - 1) / s;
    const uint32_t frame_px = P.W * P.H;
    for (uint32_t gi = blockIdx.x * kFwdWaves + (tid >> 6); gi < groups; gi += gridDim.x * kFwdWaves) {
        uint32_t ray[TILES];
        bool valid[TILES];
        bf16x8 b1[TILES][2];
        float x[TILES][6];
        uint32_t fbase[TILES], qx[TILES], qy[TILES];  // bound mode: frame's first pixel, block coordinates
#pragma unroll
        for (uint32_t j = 0; j < TILES; j++) {
            ray[j] = gi * 16u * TILES + 16u * j + r;
            valid[j] = ray[j] < P.n;
            float o[3] = {0, 0, 0}, d[3] = {0, 0, 1};
            if (valid[j]) {
                if (P.camera) {
                    Ray rr;
                    const uint32_t f = P.cams ? udiv(ray[j], P.per_frame, P.pf_m) : 0u;
                    const uint32_t loc = ray[j] - f * P.per_frame;
                    qy[j] = udiv(loc, Wq, P.wq_m);
                    qx[j] = loc - qy[j] * Wq;
                    fbase[j] = f * frame_px;
                    cast_ray(P.cams ? P.cams[f] : P.cam, qx[j] * s, qy[j] * s, P.W, P.H, rr);
                    o[0] = rr.o.x; o[1] = rr.o.y; o[2] = rr.o.z;
                    d[0] = rr.d.x; d[1] = rr.d.y; d[2] = rr.d.z;
                } else {
                    const hulk_ray q = P.rays[ray[j]];
                    o[0] = q.start.x; o[1] = q.start.y; o[2] = q.start.z;
                    d[0] = q.dir.x; d[1] = q.dir.y; d[2] = q.dir.z;
                }
            }
            for (int i = 0; i < 3; i++) x[j][i] = (o[i] - P.c[i]) * P.ih[i];
            const float dd = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
            const float inv = 1.0f / __builtin_sqrtf(dd);
            for (int i = 0; i < 3; i++) x[j][3 + i] = d[i] * inv;
        }
        // Fourier features: row 8 g + k of F against each tile's ray, sin and cos as BF16 pairs
        uint32_t es[TILES][4], ec[TILES][4];
#pragma unroll
        for (uint32_t k = 0; k < 8; k += 2) {
            float zs[TILES][2];
#pragma unroll
            for (uint32_t kq = 0; kq < 2; kq++) {
                const float* F = P.fourier + (8 * g + k + kq) * 6;
                const float f0 = F[0], f1 = F[1], f2 = F[2], f3 = F[3], f4 = F[4], f5 = F[5];
#pragma unroll
                for (uint32_t j = 0; j < TILES; j++) {
                    float z = f0 * x[j][0];
                    z = __builtin_fmaf(f1, x[j][1], z);
                    z = __builtin_fmaf(f2, x[j][2], z);
                    z = __builtin_fmaf(f3, x[j][3], z);
                    z = __builtin_fmaf(f4, x[j][4], z);
                    z = __builtin_fmaf(f5, x[j][5], z);
                    zs[j][kq] = z;
                }
            }
#pragma unroll
            for (uint32_t j = 0; j < TILES; j++) {
                es[j][k / 2] = cvt2(__sinf(zs[j][0]), __sinf(zs[j][1]));
                ec[j][k / 2] = cvt2(__cosf(zs[j][0]), __cosf(zs[j][1]));
                if (TRAIN && valid[j]) {
                    for (uint32_t kq = 0; kq < 2; kq++) {
                        P.encT[(size_t)(8 * g + k + kq) * P.ld + ray[j]] = (uint16_t)(es[j][k / 2] >> (16 * kq));
                        P.encT[(size_t)(32 + 8 * g + k + kq) * P.ld + ray[j]] = (uint16_t)(ec[j][k / 2] >> (16 * kq));
                    }
                }
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < TILES; j++) {
            b1[j][0] = as_bf16x8(make_uint4(es[j][0], es[j][1], es[j][2], es[j][3]));
            b1[j][1] = as_bf16x8(make_uint4(ec[j][0], ec[j][1], ec[j][2], ec[j][3]));
        }
        // layer 1 (K = 64: sin, then cos features)
        uint2 h[8][TILES];
#pragma unroll
        for (uint32_t t = 0; t < 8; t++) {
            const bf16x8 a0 = as_bf16x8(wl[(2 * t) * 64 + l]);
            const bf16x8 a1 = as_bf16x8(wl[(2 * t + 1) * 64 + l]);
            const float4 bias = *reinterpret_cast<const float4*>(P.p + ob1 + 16 * t + 4 * g);
#pragma unroll
            for (uint32_t j = 0; j < TILES; j++) {
                f32x4 acc = {bias.x, bias.y, bias.z, bias.w};
                acc = mfma(a0, b1[j][0], acc);
                acc = mfma(a1, b1[j][1], acc);
                h[t][j] = activate(acc, TRAIN ? P.h1T : nullptr, P.ld, 16 * t + 4 * g, ray[j], valid[j]);
            }
        }
        // layers 2 and 3 (K = 128) from registers; layer 3 folds the output dot
        float part[TILES];
#pragma unroll
        for (uint32_t j = 0; j < TILES; j++) part[j] = 0.0f;
#pragma unroll
        for (uint32_t layer = 0; layer < 2; layer++) {
            const uint32_t f0 = kFrag1 + layer * kFrag2;
            const float* bias_l = P.p + (layer == 0 ? ob2 : ob3);
            uint16_t* hT = TRAIN ? (layer == 0 ? P.h2T : P.h3T) : nullptr;
            uint2 hn[8][TILES];
#pragma unroll
            for (uint32_t t = 0; t < 8; t++) {
                const float4 bias = *reinterpret_cast<const float4*>(bias_l + 16 * t + 4 * g);
                f32x4 acc[TILES];
#pragma unroll
                for (uint32_t j = 0; j < TILES; j++) acc[j] = f32x4{bias.x, bias.y, bias.z, bias.w};
#pragma unroll
                for (uint32_t kk = 0; kk < 4; kk++) {
                    const bf16x8 a = as_bf16x8(wl[(f0 + 4 * t + kk) * 64 + l]);
#pragma unroll
                    for (uint32_t j = 0; j < TILES; j++)
                        acc[j] = mfma(a, join(h[2 * kk][j], h[2 * kk + 1][j]), acc[j]);
                }
                const uint2 w4p = *reinterpret_cast<const uint2*>(P.w4 + 16 * t + 4 * g);
#pragma unroll
                for (uint32_t j = 0; j < TILES; j++) {
                    hn[t][j] = activate(acc[j], hT, P.ld, 16 * t + 4 * g, ray[j], valid[j]);
                    if (layer == 1) part[j] = dot2(hn[t][j].y, w4p.y, dot2(hn[t][j].x, w4p.x, part[j]));
                }
            }
            if (layer == 0) {
#pragma unroll
                for (uint32_t t = 0; t < 8; t++)
#pragma unroll
                    for (uint32_t j = 0; j < TILES; j++) h[t][j] = hn[t][j];
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < TILES; j++) {
            float pj = part[j];
            pj += __shfl_xor(pj, 16, 64);
            pj += __shfl_xor(pj, 32, 64);
            const float y = pj + P.p[ob4];
            if (g == 0 && valid[j]) {
                if (P.pred) P.pred[ray[j]] = y;
                if (P.tbound) {
                    const float e = y + P.margin;
                    const float tb = (e < P.max_log) ? __expf(e) : kFltMax;
                    if (P.camera) {
                        for (uint32_t yy = qy[j] * s; yy < min(P.H, qy[j] * s + s); yy++)
                            for (uint32_t xx = qx[j] * s; xx < min(P.W, qx[j] * s + s); xx++)
                                P.tbound[(size_t)fbase[j] + (size_t)yy * P.W + xx] = tb;
                    } else {
                        P.tbound[ray[j]] = tb;
                    }
                }
            }
        }
    }
}

// C[M][N] = sum_k A[m][k] * Bt[n][k]; A, Bt row-major with k contiguous.
// One wave per 32 x 32 tile, split-K over blockIdx.z.
struct GemmParams {
    uint32_t M, N, K, k_chunk;
    const uint16_t* A;
    uint32_t lda;
    const uint16_t* Bt;
    uint32_t ldb;
    uint32_t mode;  // 0: atomicAdd into C; 1: relu-mask and write bf16 twice
    float* C;
    uint32_t ldc;
    const uint16_t* maskT;
    uint32_t ldm;
    uint16_t* outT;
    uint32_t ldoT;
    uint16_t* out;
    uint32_t ldo;
};

__global__ void __launch_bounds__(64) k_gemm_nt(GemmParams G) {
    const uint32_t l = threadIdx.x, g = l >> 4, r = l & 15u;
    const uint32_t m0 = blockIdx.x * 32u, n0 = blockIdx.y * 32u;
    const uint32_t k_begin = blockIdx.z * G.k_chunk;
    const uint32_t k_end = min(G.K, k_begin + G.k_chunk);
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t k0 = k_begin; k0 < k_end; k0 += 32) {
        bf16x8 fa[2], fb[2];
#pragma unroll
        for (int a = 0; a < 2; a++) fa[a] = ld8(G.A + (size_t)(m0 + 16 * a + r) * G.lda + k0 + 8 * g);
#pragma unroll
        for (int b = 0; b < 2; b++) fb[b] = ld8(G.Bt + (size_t)(n0 + 16 * b + r) * G.ldb + k0 + 8 * g);
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = mfma(fa[a], fb[b], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (uint32_t i = 0; i < 4; i++) {
                const uint32_t m = m0 + 16 * a + 4 * g + i, n = n0 + 16 * b + r;
                float v = acc[a][b][i];
                if (G.mode == 0) {
                    atomicAdd(G.C + (size_t)m * G.ldc + n, v);
                } else {
                    if (!(bf2f(G.maskT[(size_t)m * G.ldm + n]) > 0.0f)) v = 0.0f;
                    const uint16_t h = f2bf(v);
                    G.outT[(size_t)m * G.ldoT + n] = h;
                    G.out[(size_t)n * G.ldo + m] = h;
                }
            }
}

// dy = 2 (y^ - y) / n; dz3[n][r] = bf16(dy[r] * w4[n] * [h3[n][r] > 0]) in
// both layouts; loss accumulates sum (y^ - y)^2 / n.
__global__ void k_dz3(const float* pred, const float* target, uint32_t n, uint32_t ld, const uint16_t* w4,
                      const uint16_t* h3T, float* dy, uint16_t* dz3T, uint16_t* dz3, float* loss) {
    const uint32_t ray = blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= n) return;
    const float e = pred[ray] - target[ray];
    const float d = 2.0f * e / (float)n;
    dy[ray] = d;
    atomicAdd(loss, e * e / (float)n);
    for (uint32_t k = 0; k < kHid; k++) {
        float v = (bf2f(h3T[(size_t)k * ld + ray]) > 0.0f) ? d * bf2f(w4[k]) : 0.0f;
        const uint16_t h = f2bf(v);
        dz3T[(size_t)k * ld + ray] = h;
        dz3[(size_t)ray * kHid + k] = h;
    }
}

// out[f] += scale * sum_r w[r] * T[f][r]   (w = null: 1); one block per row.
__global__ void k_rowsum(const uint16_t* T, uint32_t ld, uint32_t n, const float* w, float* out) {
    __shared__ float red[256];
    const uint32_t f = blockIdx.x;
    float s = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += bf2f(T[(size_t)f * ld + i]) * (w ? w[i] : 1.0f);
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[f] += red[0];
}

__global__ void k_sum(const float* v, uint32_t n, float* out) {
    __shared__ float red[256];
    float s = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] += red[0];
}

// BF16 copies of the master weights: W1, W2, W3 and w4 row-major, W2 and
// W3 transposed (backward), and W1, W2, W3 as the forward kernel's MFMA
// operand fragments (Wf: fragment f, lane l = 16 g + r, 8 bf16; layer 1
// rows 16 t + r, K 32 kk + 8 g + p; layers 2 and 3 K order hidden_k).
__global__ void k_refresh(const float* p, uint16_t* W1, uint16_t* W2, uint16_t* W3, uint16_t* w4, uint16_t* W2T,
                          uint16_t* W3T, uint16_t* Wf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kFragAll * 64 * 8) {
        const uint32_t pp = i & 7u, lane = (i >> 3) & 63u, f = i >> 9;
        const uint32_t g = lane >> 4, r = lane & 15u;
        float w;
        if (f < kFrag1) {
            const uint32_t t = f >> 1, kk = f & 1u;
            w = p[oW1 + (16 * t + r) * kIn + 32 * kk + 8 * g + pp];
        } else {
            const uint32_t f2 = (f - kFrag1) % kFrag2, t = f2 >> 2, kk = f2 & 3u;
            const uint32_t base = f < kFrag1 + kFrag2 ? oW2 : oW3;
            w = p[base + (16 * t + r) * kHid + hidden_k(kk, g, pp)];
        }
        Wf[i] = f2bf(w);
    }
    if (i < kHid * kIn) W1[i] = f2bf(p[oW1 + i]);
    if (i < kHid * kHid) {
        const uint32_t a = i / kHid, b = i % kHid;
        W2[i] = f2bf(p[oW2 + i]);
        W3[i] = f2bf(p[oW3 + i]);
        W2T[b * kHid + a] = f2bf(p[oW2 + i]);
        W3T[b * kHid + a] = f2bf(p[oW3 + i]);
    }
    if (i < kHid) w4[i] = f2bf(p[ow4 + i]);
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, uint32_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_zero(float* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace mlp
}  // namespace hulk

using namespace hulk;
using namespace hulk::mlp;

// ------------------------------------------------------------------ host --

struct hulk_mlp {
    hulk_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    uint32_t max_batch = 0;
    float c[3], ih[3];
    float* p = nullptr;       // master params
    float* grad = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* fourier = nullptr;
    uint16_t *W1 = nullptr, *W2 = nullptr, *W3 = nullptr, *w4 = nullptr, *W2T = nullptr, *W3T = nullptr;
    uint16_t* Wf = nullptr;  // forward operand fragments (k_refresh)
    int num_cus = 256;
    Camera* d_cams = nullptr;  // cameras of a multi-frame bound call (device, pinned staging)
    Camera* h_cams = nullptr;
    uint32_t cam_cap = 0;
    hipEvent_t cams_ev = nullptr;
    uint16_t *encT = nullptr, *h1T = nullptr, *h2T = nullptr, *h3T = nullptr;
    uint16_t *dz3T = nullptr, *dz3 = nullptr, *dz2T = nullptr, *dz2 = nullptr, *dz1T = nullptr, *dz1 = nullptr;
    float *pred = nullptr, *dy = nullptr, *loss = nullptr, *host_loss = nullptr;
    uint32_t step = 0;
};

extern "C" int hulk_ctx_scene_box(const hulk_ctx* c, float out[6]);  // hulk_api.hip
extern "C" void* hulk_stream(hulk_ctx* c);

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

#define MLP_TRY(m, expr)                                                           \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            (m)->err = std::string(#expr) + ": " + hipGetErrorString(_e);          \
            return HULK_E_HIP;                                                     \
        }                                                                          \
    } while (0)

int mlp_refresh(hulk_mlp* m) {
    hipLaunchKernelGGL(k_refresh, dim3((kFragAll * 64 * 8 + 255) / 256), dim3(256), 0, m->stream, m->p, m->W1, m->W2,
                       m->W3, m->w4, m->W2T, m->W3T, m->Wf);
    MLP_TRY(m, hipGetLastError());
    return HULK_OK;
}

FwdParams fwd_params(hulk_mlp* m) {
    FwdParams P;
    std::memset(&P, 0, sizeof(P));
    for (int i = 0; i < 3; i++) {
        P.c[i] = m->c[i];
        P.ih[i] = m->ih[i];
    }
    P.fourier = m->fourier;
    P.Wf = reinterpret_cast<const uint4*>(m->Wf);
    P.w4 = m->w4;
    P.p = m->p;
    return P;
}

int launch_fwd(hulk_mlp* m, const FwdParams& P) {
    // one block per CU (the 80 KiB of weight fragments in LDS), fewer when
    // there are fewer groups of 64 rays than waves
    const bool train = P.encT != nullptr;
    const uint32_t per_group = 16u * (train ? kTrainTiles : kTiles);
    const uint32_t groups = (P.n + per_group - 1) / per_group;
    const uint32_t grid = std::max<uint32_t>(
        1u, std::min<uint32_t>((groups + kFwdWaves - 1) / kFwdWaves, static_cast<uint32_t>(m->num_cus)));
    if (train)
        hipLaunchKernelGGL((k_mlp_fwd<true, kTrainTiles>), dim3(grid), dim3(64 * kFwdWaves), 0, m->stream, P);
    else
        hipLaunchKernelGGL((k_mlp_fwd<false, kTiles>), dim3(grid), dim3(64 * kFwdWaves), 0, m->stream, P);
    MLP_TRY(m, hipGetLastError());
    return HULK_OK;
}

// Forward with stored activations + backward into m->grad (zeroed first).
int mlp_backward(hulk_mlp* m, const hulk_ray* rays, const float* y, uint32_t n) {
    const uint32_t ld = m->max_batch;
    FwdParams P = fwd_params(m);
    P.rays = rays;
    P.n = n;
    P.pred = m->pred;
    P.encT = m->encT;
    P.h1T = m->h1T;
    P.h2T = m->h2T;
    P.h3T = m->h3T;
    P.ld = ld;
    int r;
    hipLaunchKernelGGL(k_zero, dim3((kParams + 255) / 256), dim3(256), 0, m->stream, m->grad, kParams);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, m->stream, m->loss, 1u);
    if ((r = launch_fwd(m, P))) return r;
    hipLaunchKernelGGL(k_dz3, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->pred, y, n, ld, m->w4, m->h3T, m->dy,
                       m->dz3T, m->dz3, m->loss);
    // w4, b4
    hipLaunchKernelGGL(k_rowsum, dim3(kHid), dim3(256), 0, m->stream, m->h3T, ld, n, m->dy, m->grad + ow4);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, m->stream, m->dy, n, m->grad + ob4);
    const uint32_t splits = std::max<uint32_t>(1u, std::min<uint32_t>(64u, n / 256u));
    uint32_t chunk = ((n + splits - 1) / splits + 31u) & ~31u;
    auto gemm_w = [&](const uint16_t* dzT, const uint16_t* xT, uint32_t K_out, float* out) {
        GemmParams G;
        std::memset(&G, 0, sizeof(G));
        G.M = kHid;
        G.N = K_out;
        G.K = n;
        G.k_chunk = chunk;
        G.A = dzT;
        G.lda = ld;
        G.Bt = xT;
        G.ldb = ld;
        G.mode = 0;
        G.C = out;
        G.ldc = K_out;
        hipLaunchKernelGGL(k_gemm_nt, dim3(kHid / 32, K_out / 32, (n + chunk - 1) / chunk), dim3(64), 0, m->stream, G);
    };
    auto gemm_h = [&](const uint16_t* WT, const uint16_t* dz, const uint16_t* hT, uint16_t* dzT_out,
                      uint16_t* dz_out) {
        GemmParams G;
        std::memset(&G, 0, sizeof(G));
        G.M = kHid;
        G.N = n;
        G.K = kHid;
        G.k_chunk = kHid;
        G.A = WT;
        G.lda = kHid;
        G.Bt = dz;
        G.ldb = kHid;
        G.mode = 1;
        G.maskT = hT;
        G.ldm = ld;
        G.outT = dzT_out;
        G.ldoT = ld;
        G.out = dz_out;
        G.ldo = kHid;
        hipLaunchKernelGGL(k_gemm_nt, dim3(kHid / 32, n / 32, 1), dim3(64), 0, m->stream, G);
    };
    // layer 3
    gemm_w(m->dz3T, m->h2T, kHid, m->grad + oW3);
    hipLaunchKernelGGL(k_rowsum, dim3(kHid), dim3(256), 0, m->stream, m->dz3T, ld, n, nullptr, m->grad + ob3);
    gemm_h(m->W3T, m->dz3, m->h2T, m->dz2T, m->dz2);
    // layer 2
    gemm_w(m->dz2T, m->h1T, kHid, m->grad + oW2);
    hipLaunchKernelGGL(k_rowsum, dim3(kHid), dim3(256), 0, m->stream, m->dz2T, ld, n, nullptr, m->grad + ob2);
    gemm_h(m->W2T, m->dz2, m->h1T, m->dz1T, m->dz1);
    // layer 1
    gemm_w(m->dz1T, m->encT, kIn, m->grad + oW1);
    hipLaunchKernelGGL(k_rowsum, dim3(kHid), dim3(256), 0, m->stream, m->dz1T, ld, n, nullptr, m->grad + ob1);
    MLP_TRY(m, hipGetLastError());
    return HULK_OK;
}

}  // namespace

extern "C" uint32_t hulk_mlp_param_count(void) { return kParams; }

extern "C" void hulk_mlp_destroy(hulk_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void* bufs[] = {m->p, m->grad, m->m, m->v, m->fourier, m->W1, m->W2, m->W3, m->w4, m->W2T, m->W3T, m->Wf,
                    m->encT, m->h1T, m->h2T, m->h3T, m->dz3T, m->dz3, m->dz2T, m->dz2, m->dz1T, m->dz1,
                    m->pred, m->dy, m->loss};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (m->host_loss) (void)hipHostFree(m->host_loss);
    if (m->d_cams) (void)hipFree(m->d_cams);
    if (m->h_cams) (void)hipHostFree(m->h_cams);
    if (m->cams_ev) (void)hipEventDestroy(m->cams_ev);
    delete m;
}

extern "C" const char* hulk_mlp_last_error(const hulk_mlp* m) { return m ? m->err.c_str() : "null mlp"; }

extern "C" int hulk_mlp_create(hulk_ctx* ctx, uint64_t seed, uint32_t max_batch, hulk_mlp** out) {
    if (!ctx || !out || max_batch == 0 || (max_batch % 64) != 0) return HULK_E_ILLEGAL_INPUT;
    *out = nullptr;
    float box[6];
    if (hulk_ctx_scene_box(ctx, box) != HULK_OK) return HULK_E_NO_SCENE;
    hulk_mlp* m = new (std::nothrow) hulk_mlp;
    if (!m) return HULK_E_ALLOC;
    m->ctx = ctx;
    m->stream = static_cast<hipStream_t>(hulk_stream(ctx));
    m->max_batch = max_batch;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        m->num_cus = prop.multiProcessorCount;
    for (int i = 0; i < 3; i++) {
        m->c[i] = (box[i] + box[3 + i]) * 0.5f;
        m->ih[i] = 1.0f / ((box[3 + i] - box[i]) * 0.5f);
    }
    const size_t nb = max_batch;
    struct A {
        void** p;
        size_t bytes;
    } allocs[] = {
        {(void**)&m->p, kParams * 4}, {(void**)&m->grad, kParams * 4}, {(void**)&m->m, kParams * 4},
        {(void**)&m->v, kParams * 4}, {(void**)&m->fourier, 32 * 6 * 4}, {(void**)&m->W1, kHid * kIn * 2},
        {(void**)&m->W2, kHid * kHid * 2}, {(void**)&m->W3, kHid * kHid * 2}, {(void**)&m->w4, kHid * 2},
        {(void**)&m->W2T, kHid * kHid * 2}, {(void**)&m->W3T, kHid * kHid * 2}, {(void**)&m->Wf, kFragAll * 64 * 16},
        {(void**)&m->encT, kIn * nb * 2},
        {(void**)&m->h1T, kHid * nb * 2}, {(void**)&m->h2T, kHid * nb * 2}, {(void**)&m->h3T, kHid * nb * 2},
        {(void**)&m->dz3T, kHid * nb * 2}, {(void**)&m->dz3, kHid * nb * 2}, {(void**)&m->dz2T, kHid * nb * 2},
        {(void**)&m->dz2, kHid * nb * 2}, {(void**)&m->dz1T, kHid * nb * 2}, {(void**)&m->dz1, kHid * nb * 2},
        {(void**)&m->pred, nb * 4}, {(void**)&m->dy, nb * 4}, {(void**)&m->loss, 64}};
    for (auto& a : allocs) {
        if (hipMalloc(a.p, a.bytes) != hipSuccess) {
            *a.p = nullptr;
            hulk_mlp_destroy(m);
            return HULK_E_ALLOC;
        }
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&m->host_loss), 64, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&m->cams_ev, hipEventDisableTiming) != hipSuccess) {
        hulk_mlp_destroy(m);
        return HULK_E_ALLOC;
    }
    // He-normal weights, zero biases, Fourier matrix N(0, 4^2).
    Rng rng{seed};
    std::vector<float> p(kParams, 0.0f), F(32 * 6);
    for (uint32_t i = 0; i < kHid * kIn; i++) p[oW1 + i] = (float)(rng.normal() * std::sqrt(2.0 / kIn));
    for (uint32_t i = 0; i < kHid * kHid; i++) p[oW2 + i] = (float)(rng.normal() * std::sqrt(2.0 / kHid));
    for (uint32_t i = 0; i < kHid * kHid; i++) p[oW3 + i] = (float)(rng.normal() * std::sqrt(2.0 / kHid));
    for (uint32_t i = 0; i < kHid; i++) p[ow4 + i] = (float)(rng.normal() * std::sqrt(1.0 / kHid));
    for (auto& f : F) f = (float)(rng.normal() * 4.0);
    std::vector<float> zeros(kParams, 0.0f);
    if (hipMemcpyAsync(m->p, p.data(), kParams * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->m, zeros.data(), kParams * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->v, zeros.data(), kParams * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->fourier, F.data(), F.size() * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        mlp_refresh(m) != HULK_OK || hipStreamSynchronize(m->stream) != hipSuccess) {
        hulk_mlp_destroy(m);
        return HULK_E_HIP;
    }
    *out = m;
    return HULK_OK;
}

extern "C" int hulk_mlp_get_params(hulk_mlp* m, float* params, float* fourier) {
    if (!m) return HULK_E_ILLEGAL_INPUT;
    if (params) MLP_TRY(m, hipMemcpyAsync(params, m->p, kParams * 4, hipMemcpyDeviceToHost, m->stream));
    if (fourier) MLP_TRY(m, hipMemcpyAsync(fourier, m->fourier, 32 * 6 * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    return HULK_OK;
}

extern "C" int hulk_mlp_set_params(hulk_mlp* m, const float* params, const float* fourier) {
    if (!m) return HULK_E_ILLEGAL_INPUT;
    if (params) MLP_TRY(m, hipMemcpyAsync(m->p, params, kParams * 4, hipMemcpyHostToDevice, m->stream));
    if (fourier) MLP_TRY(m, hipMemcpyAsync(m->fourier, fourier, 32 * 6 * 4, hipMemcpyHostToDevice, m->stream));
    int r = mlp_refresh(m);
    if (r) return r;
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    return HULK_OK;
}

extern "C" int hulk_mlp_forward(hulk_mlp* m, const hulk_ray* rays_dev, uint32_t n, float* pred_dev) {
    if (!m || (n && (!rays_dev || !pred_dev))) return HULK_E_ILLEGAL_INPUT;
    if (!n) return HULK_OK;
    FwdParams P = fwd_params(m);
    P.rays = rays_dev;
    P.n = n;
    P.pred = pred_dev;
    int r = launch_fwd(m, P);
    if (r) return r;
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    return HULK_OK;
}

extern "C" int hulk_mlp_loss_grad(hulk_mlp* m, const hulk_ray* rays_dev, const float* y_dev, uint32_t n,
                                  float* loss, float* grad_host) {
    if (!m || !rays_dev || !y_dev || n == 0 || n > m->max_batch || (n % 64) != 0) return HULK_E_ILLEGAL_INPUT;
    int r = mlp_backward(m, rays_dev, y_dev, n);
    if (r) return r;
    MLP_TRY(m, hipMemcpyAsync(m->host_loss, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
    if (grad_host) MLP_TRY(m, hipMemcpyAsync(grad_host, m->grad, kParams * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    if (loss) *loss = m->host_loss[0];
    return HULK_OK;
}

extern "C" int hulk_mlp_train_step(hulk_mlp* m, const hulk_ray* rays_dev, const float* y_dev, uint32_t n, float lr,
                                   float* loss) {
    if (!m || !rays_dev || !y_dev || n == 0 || n > m->max_batch || (n % 64) != 0) return HULK_E_ILLEGAL_INPUT;
    int r = mlp_backward(m, rays_dev, y_dev, n);
    if (r) return r;
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - (float)std::pow(b1, m->step), c2 = 1.0f - (float)std::pow(b2, m->step);
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, m->stream, m->p, m->grad, m->m, m->v,
                       kParams, lr, b1, b2, 1e-8f, c1, c2);
    MLP_TRY(m, hipGetLastError());
    if ((r = mlp_refresh(m))) return r;
    if (loss) {
        MLP_TRY(m, hipMemcpyAsync(m->host_loss, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
        MLP_TRY(m, hipStreamSynchronize(m->stream));
        *loss = m->host_loss[0];
    }
    return HULK_OK;
}

namespace {

Camera to_camera(const hulk_camera& c) {
    Camera k;
    k.from[0] = c.from.x; k.from[1] = c.from.y; k.from[2] = c.from.z;
    k.vf[0] = c.vertical_factor.x; k.vf[1] = c.vertical_factor.y; k.vf[2] = c.vertical_factor.z;
    k.hf[0] = c.horizontal_factor.x; k.hf[1] = c.horizontal_factor.y; k.hf[2] = c.horizontal_factor.z;
    k.ll[0] = c.lower_left.x; k.ll[1] = c.lower_left.y; k.ll[2] = c.lower_left.z;
    return k;
}

// Bounds of n_cams frames of W x H pixels in one launch: frame f's bounds
// start at tbound + f * W * H.
int bound_frames(hulk_mlp* m, const hulk_camera* cams, uint32_t n_cams, uint32_t W, uint32_t H, float margin,
                 float max_log, uint32_t stride, float* tbound_dev, uint32_t flags) {
    if (!m || !cams || !tbound_dev || n_cams == 0) return HULK_E_ILLEGAL_INPUT;
    if (stride == 0 || stride > 8 || W == 0 || H == 0) {
        m->err = "stride must be in [1, 8] and the frame non-empty";
        return HULK_E_ILLEGAL_INPUT;
    }
    const uint64_t per_frame = (uint64_t)((W + stride - 1) / stride) * ((H + stride - 1) / stride);
    if (per_frame * n_cams > 0xFFFFFFC0ull || (uint64_t)W * H * n_cams > 0xFFFFFFFFull) {
        m->err = "too many pixels in one call";
        return HULK_E_ILLEGAL_INPUT;
    }
    FwdParams P = fwd_params(m);
    P.camera = 1;
    P.cam = to_camera(cams[0]);
    if (n_cams > 1) {
        if (n_cams > m->cam_cap) {
            if (m->d_cams) (void)hipFree(m->d_cams);
            if (m->h_cams) (void)hipHostFree(m->h_cams);
            m->d_cams = nullptr;
            m->h_cams = nullptr;
            m->cam_cap = 0;
            MLP_TRY(m, hipMalloc(reinterpret_cast<void**>(&m->d_cams), n_cams * sizeof(Camera)));
            MLP_TRY(m, hipHostMalloc(reinterpret_cast<void**>(&m->h_cams), n_cams * sizeof(Camera), hipHostMallocDefault));
            m->cam_cap = n_cams;
        }
        // the pinned staging copy may still feed the previous call's upload
        MLP_TRY(m, hipEventSynchronize(m->cams_ev));
        for (uint32_t f = 0; f < n_cams; f++) m->h_cams[f] = to_camera(cams[f]);
        MLP_TRY(m, hipMemcpyAsync(m->d_cams, m->h_cams, n_cams * sizeof(Camera), hipMemcpyHostToDevice, m->stream));
        MLP_TRY(m, hipEventRecord(m->cams_ev, m->stream));
        P.cams = m->d_cams;
    }
    P.W = W;
    P.H = H;
    P.stride = stride;
    P.per_frame = static_cast<uint32_t>(per_frame);
    P.pf_m = udiv_multiplier(P.per_frame);
    P.wq_m = udiv_multiplier((W + stride - 1) / stride);
    P.n = static_cast<uint32_t>(per_frame * n_cams);
    P.tbound = tbound_dev;
    P.margin = margin;
    P.max_log = max_log;
    int r = launch_fwd(m, P);
    if (r) return r;
    if (!(flags & HULK_FLAG_NO_SYNC)) MLP_TRY(m, hipStreamSynchronize(m->stream));
    return HULK_OK;
}

}  // namespace

extern "C" int hulk_mlp_bound_camera(hulk_mlp* m, const hulk_camera* cam, uint32_t W, uint32_t H, float margin,
                                     float max_log, uint32_t stride, float* tbound_dev, uint32_t flags) {
    return bound_frames(m, cam, 1, W, H, margin, max_log, stride, tbound_dev, flags);
}

extern "C" int hulk_mlp_bound_cameras(hulk_mlp* m, const hulk_camera* cams, uint32_t n_cams, uint32_t W, uint32_t H,
                                      float margin, float max_log, uint32_t stride, float* tbound_dev,
                                      uint32_t flags) {
    return bound_frames(m, cams, n_cams, W, H, margin, max_log, stride, tbound_dev, flags);
}
