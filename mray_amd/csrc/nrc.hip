// Neural radiance cache on gfx950: multiresolution hash encoding + 64-wide
// fully fused MLP on bf16 MFMA (fp32 accumulation), trained online with Adam.
//
// No reference counterpart (the reference only keeps an irradiance hash grid,
// Tracer/HashGrid.h:17-476); the methods are the published ones: hash
// encoding of Mueller et al. 2022 (Instant NGP), one-blob encoding of Mueller
// et al. 2019, network and relative L2 loss of Mueller et al. 2021 (Real-time
// Neural Radiance Caching).
//
// Query record (11 floats): position in [0,1]^3 (scene box), direction and
// normal as (theta, phi) in [0,1]^2, roughness in [0,1], albedo RGB.
// Network input (64): 16 levels x 2 hash features, one-blob(4) of the five
// scalars, albedo, a constant 1, zero padding.  Six 64x64 layers (the output
// layer's columns 3..63 stay zero), ReLU after the first five.
//
// Kernels (one wave = 32 samples as one 32x64 activation tile in LDS; the
// MFMA is gfx950's v_mfma_f32_32x32x16_bf16, 16 k per instruction):
//   KNrcForward<TRAIN>  encode + 6 layers; inference writes RGB, training also
//                       stores every layer input (bf16) and dL/dy
//   KNrcBackward        dZ of every layer (bf16) and dL/d(grid features)
//   KNrcWgrad           dW = X^T dZ per layer, split over sample chunks
//   KNrcGridBwd         scatter of the grid-feature gradient (fp32 atomics)
//   KNrcAdam            Adam on fp32 masters, refreshes the bf16 / fp16 copies
//                       and clears the gradients for the next step
#include "common.h"
#include "scene.h"

#include <hip/hip_fp16.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace mray
{

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr uint32_t kNrcLayers = 6;
constexpr uint32_t kW = 64;                  // width of every layer
constexpr uint32_t kLdsRow = 72;             // padded LDS row: 16-B aligned, conflict-free b128 reads
constexpr uint32_t kNrcFeat = 2;
constexpr uint32_t kNrcMaxLevels = 16;
constexpr uint32_t kFwdWaves = 4;            // 128 samples per block
constexpr uint32_t kLayerParams = kW * kW;
constexpr uint32_t kMlpParams = kNrcLayers * kLayerParams;
// kernel kinds timed by mray_nrc_kernel_times: inference forward, training
// forward, backward (data gradients), weight gradients, grid scatter, Adam
constexpr uint32_t kNrcKernelKinds = 6;

__device__ __forceinline__ uint16_t F2Bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);         // round to nearest even
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float Bf2F(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

struct NrcView
{
    const uint16_t* wT;       // [6][n][k] bf16 (forward operand)
    const uint16_t* wK;       // [6][k][n] bf16 (backward operand)
    const __half* grid;       // [L][T][2]
    float scale[kNrcMaxLevels];
    uint32_t res[kNrcMaxLevels];
    uint32_t dense[kNrcMaxLevels];
    uint32_t levels;
    uint32_t log2T;
};

// ---- encoding --------------------------------------------------------------
__device__ __forceinline__ uint32_t GridIndex(const NrcView& v, uint32_t l, uint32_t x, uint32_t y,
                                              uint32_t z)
{
    if(v.dense[l])
    {
        uint32_t r = v.res[l];
        return x + y * r + z * r * r;
    }
    uint32_t h = (x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u);
    return h & ((1u << v.log2T) - 1u);
}

struct Corners
{
    uint32_t idx[8];
    float w[8];
};

__device__ __forceinline__ void LevelCorners(const NrcView& v, uint32_t l, const float* pos, Corners& c)
{
    float p[3], fr[3];
    uint32_t i0[3];
    for(int d = 0; d < 3; d++)
    {
        p[d] = pos[d] * v.scale[l] + 0.5f;
        float fl = floorf(p[d]);
        i0[d] = uint32_t(fl);
        fr[d] = p[d] - fl;
    }
    for(uint32_t k = 0; k < 8; k++)
    {
        uint32_t ox = k & 1u, oy = (k >> 1) & 1u, oz = (k >> 2) & 1u;
        float w = (ox ? fr[0] : 1.0f - fr[0]);
        w = w * (oy ? fr[1] : 1.0f - fr[1]);
        w = w * (oz ? fr[2] : 1.0f - fr[2]);
        c.w[k] = w;
        c.idx[k] = GridIndex(v, l, i0[0] + ox, i0[1] + oy, i0[2] + oz);
    }
}

__device__ __forceinline__ float OneBlob(float s, uint32_t j)
{
    float c = (float(j) + 0.5f) * 0.25f;
    float d = s - c;
    return expf(-d * d * 8.0f);              // sigma = 1/4
}

// Input row of one sample, written by two lanes: `half` 0 takes levels 0..7
// and columns 32..47, `half` 1 levels 8..15 and columns 48..63.
__device__ __forceinline__ void EncodeHalf(const NrcView& v, const float* rec, uint32_t half,
                                           uint16_t* row)
{
    float pos[3] = {ClampF(rec[0], 0.0f, 1.0f), ClampF(rec[1], 0.0f, 1.0f), ClampF(rec[2], 0.0f, 1.0f)};
    const uint32_t l0 = half * 8u;
    for(uint32_t l = l0; l < l0 + 8u; l++)
    {
        float f0 = 0.0f, f1 = 0.0f;
        if(l < v.levels)
        {
            Corners c;
            LevelCorners(v, l, pos, c);
            const __half2* g = reinterpret_cast<const __half2*>(v.grid) + (size_t(l) << v.log2T);
            for(uint32_t k = 0; k < 8; k++)
            {
                float2 e = __half22float2(g[c.idx[k]]);
                f0 += c.w[k] * e.x;
                f1 += c.w[k] * e.y;
            }
        }
        row[2 * l] = F2Bf(f0);
        row[2 * l + 1] = F2Bf(f1);
    }
    if(half == 0)
    {
        // one-blob of dir.theta, dir.phi, normal.theta, normal.phi (cols 32..47)
        for(uint32_t q = 0; q < 4; q++)
            for(uint32_t j = 0; j < 4; j++) row[32 + 4 * q + j] = F2Bf(OneBlob(rec[3 + q], j));
    }
    else
    {
        for(uint32_t j = 0; j < 4; j++) row[48 + j] = F2Bf(OneBlob(rec[7], j));
        row[52] = F2Bf(rec[8]);
        row[53] = F2Bf(rec[9]);
        row[54] = F2Bf(rec[10]);
        row[55] = F2Bf(1.0f);
        for(uint32_t j = 56; j < 64; j++) row[j] = 0;
    }
}

// ---- MFMA tile products ------------------------------------------------------
// acc[nt] = X(32x64, LDS rows of kLdsRow) * B, where the operand for output
// column n and reduction index k is bRow(n)[k] (4 consecutive k per lane).
__device__ __forceinline__ void TileProduct(const uint16_t* sA, const uint16_t* sB, v16f acc[2])
{
    // v_mfma_f32_32x32x16_bf16: lane l supplies row (l mod 32) of A and
    // column (l mod 32) of B over k = 8 (l / 32) .. +7 of the 16-k step
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = lane & 31u, kh = (lane >> 5) * 8u;
    acc[0] = v16f{0};
    acc[1] = v16f{0};
    #pragma unroll
    for(uint32_t kt = 0; kt < kW / 16u; kt++)
    {
        const uint32_t k = kt * 16u + kh;
        v8bf a = *reinterpret_cast<const v8bf*>(sA + r * kLdsRow + k);
        v8bf b0 = *reinterpret_cast<const v8bf*>(sB + r * kLdsRow + k);
        v8bf b1 = *reinterpret_cast<const v8bf*>(sB + (r + 32u) * kLdsRow + k);
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b0, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b1, acc[1], 0, 0, 0);
    }
}

// Row of accumulator element e for this lane (32x32 output layout).
__device__ __forceinline__ uint32_t AccRow(uint32_t e)
{
    const uint32_t lane = threadIdx.x & 63u;
    return 8u * (e >> 2) + 4u * (lane >> 5) + (e & 3u);
}

__device__ __forceinline__ void StageWeights(const uint16_t* __restrict__ g, uint16_t* s)
{
    // [6][64][64] -> padded LDS rows
    for(uint32_t i = threadIdx.x; i < kNrcLayers * kW * kW / 4u; i += blockDim.x)
    {
        uint32_t e = i * 4u;
        uint32_t row = e / kW, col = e % kW;
        *reinterpret_cast<uint2*>(s + row * kLdsRow + col) = reinterpret_cast<const uint2*>(g)[i];
    }
}

// ---- forward ---------------------------------------------------------------
struct NrcTrainBuf
{
    uint16_t* x;       // [6][B][64] bf16 layer inputs
    uint16_t* dz;      // [6][B][64] bf16 gradients of the layer outputs
    float* dy;         // [B][3]
    float* dx0;        // [B][32]
    float* loss;       // [1]
};

template<bool TRAIN>
__global__ void __launch_bounds__(kFwdWaves * 64)
KNrcForward(NrcView v, const float* __restrict__ rec, const uint32_t* __restrict__ dCount,
            uint32_t count, float* __restrict__ out, const float* __restrict__ target,
            NrcTrainBuf tb, uint32_t batch)
{
    __shared__ __attribute__((aligned(16))) uint16_t sW[kNrcLayers * kW * kLdsRow];
    __shared__ __attribute__((aligned(16))) uint16_t sX[kFwdWaves][32 * kLdsRow];
    const uint32_t n = dCount ? *dCount : count;
    // a block without a tile (few or no queries this bounce) skips the
    // weight staging: the test is uniform over the block
    if(blockIdx.x * kFwdWaves * 32u >= n) return;
    StageWeights(v.wT, sW);
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint16_t* X = sX[wave];
    float lossAcc = 0.0f;
    for(uint32_t tile = blockIdx.x * kFwdWaves + wave; tile * 32u < n; tile += gridDim.x * kFwdWaves)
    {
        const uint32_t base = tile * 32u;
        // encode: lane -> (sample lane%32, half lane/32)
        {
            const uint32_t s = base + (lane & 31u);
            uint16_t* row = X + (lane & 31u) * kLdsRow;
            if(s < n) EncodeHalf(v, rec + size_t(s) * 11u, lane >> 5, row);
            else
                for(uint32_t j = (lane >> 5) * 32u; j < (lane >> 5) * 32u + 32u; j++) row[j] = 0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for(uint32_t l = 0; l < kNrcLayers; l++)
        {
            if(TRAIN)
            {
                // store this layer's input rows (32 x 64 bf16)
                for(uint32_t e = lane; e < 32u * 16u; e += 64u)
                {
                    uint32_t rr = e >> 4, c4 = (e & 15u) * 4u;
                    uint32_t s = base + rr;
                    if(s < n)
                        *reinterpret_cast<uint2*>(tb.x + (size_t(l) * batch + s) * kW + c4) =
                            *reinterpret_cast<const uint2*>(X + rr * kLdsRow + c4);
                }
            }
            v16f acc[2];
            TileProduct(X, sW + l * kW * kLdsRow, acc);
            __builtin_amdgcn_wave_barrier();
            if(l + 1u < kNrcLayers)
            {
                #pragma unroll
                for(uint32_t t = 0; t < 2; t++)
                    #pragma unroll
                    for(uint32_t e = 0; e < 16; e++)
                        X[AccRow(e) * kLdsRow + t * 32u + (lane & 31u)] = F2Bf(MaxF(acc[t][e], 0.0f));
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            else
            {
                // output columns 0..2 live in lanes 0..2 and 32..34 of tile 0
                const uint32_t c = lane & 31u;
                if(c < 3u)
                {
                    #pragma unroll
                    for(uint32_t e = 0; e < 16; e++)
                    {
                        uint32_t s = base + AccRow(e);
                        if(s >= n) continue;
                        float y = acc[0][e];
                        if(!TRAIN) out[size_t(s) * 3u + c] = y;
                        else
                        {
                            float t = target[size_t(s) * 3u + c];
                            float den = y * y + 0.01f;
                            float d = y - t;
                            lossAcc += d * d / den;
                            tb.dy[size_t(s) * 3u + c] = 2.0f * d / den / float(batch * 3u);
                        }
                    }
                }
            }
        }
    }
    if(TRAIN)
    {
        for(uint32_t o = 32; o > 0; o >>= 1) lossAcc += __shfl_down(lossAcc, o);
        if(lane == 0 && lossAcc != 0.0f) atomicAdd(tb.loss, lossAcc / float(batch * 3u));
    }
}

// ---- backward (data gradients) ----------------------------------------------
__global__ void __launch_bounds__(kFwdWaves * 64)
KNrcBackward(NrcView v, NrcTrainBuf tb, uint32_t batch)
{
    __shared__ __attribute__((aligned(16))) uint16_t sW[kNrcLayers * kW * kLdsRow];
    __shared__ __attribute__((aligned(16))) uint16_t sG[kFwdWaves][32 * kLdsRow];
    StageWeights(v.wK, sW);
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint16_t* G = sG[wave];
    for(uint32_t tile = blockIdx.x * kFwdWaves + wave; tile * 32u < batch; tile += gridDim.x * kFwdWaves)
    {
        const uint32_t base = tile * 32u;
        // dZ of the output layer: dL/dy in columns 0..2, zero elsewhere
        for(uint32_t e = lane; e < 32u * kW; e += 64u)
        {
            uint32_t rr = e / kW, c = e % kW;
            uint32_t s = base + rr;
            float g = (c < 3u && s < batch) ? tb.dy[size_t(s) * 3u + c] : 0.0f;
            G[rr * kLdsRow + c] = F2Bf(g);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for(int l = int(kNrcLayers) - 1; l >= 0; l--)
        {
            // keep dZ_{l+1} for the weight gradient of layer l
            for(uint32_t e = lane; e < 32u * 16u; e += 64u)
            {
                uint32_t rr = e >> 4, c4 = (e & 15u) * 4u;
                uint32_t s = base + rr;
                if(s < batch)
                    *reinterpret_cast<uint2*>(tb.dz + (size_t(l) * batch + s) * kW + c4) =
                        *reinterpret_cast<const uint2*>(G + rr * kLdsRow + c4);
            }
            v16f acc[2];
            // dX = dZ * W^T: output column k, reduction over n -> W[k][n] rows
            TileProduct(G, sW + uint32_t(l) * kW * kLdsRow, acc);
            __builtin_amdgcn_wave_barrier();
            if(l > 0)
            {
                // ReLU of layer l-1's output = this layer's input X_l
                #pragma unroll
                for(uint32_t t = 0; t < 2; t++)
                    #pragma unroll
                    for(uint32_t e = 0; e < 16; e++)
                    {
                        uint32_t rr = AccRow(e), c = t * 32u + (lane & 31u);
                        uint32_t s = base + rr;
                        float xin = (s < batch) ? Bf2F(tb.x[(size_t(l) * batch + s) * kW + c]) : 0.0f;
                        G[rr * kLdsRow + c] = F2Bf(xin > 0.0f ? acc[t][e] : 0.0f);
                    }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            else
            {
                // gradient of the 32 grid features
                #pragma unroll
                for(uint32_t e = 0; e < 16; e++)
                {
                    uint32_t s = base + AccRow(e);
                    if(s < batch) tb.dx0[size_t(s) * 32u + (lane & 31u)] = acc[0][e];
                }
            }
        }
    }
}

// ---- weight gradients -------------------------------------------------------
// blockIdx.y = layer; each block reduces a chunk of samples; wave w owns the
// 32x32 tile (w/2, w%2) of dW[k][n] = sum_s X[s][k] dZ[s][n].
__global__ void __launch_bounds__(256)
KNrcWgrad(NrcTrainBuf tb, uint32_t batch, uint32_t chunk, float* __restrict__ wgrad)
{
    __shared__ __attribute__((aligned(16))) uint16_t sXc[64 * kLdsRow];
    __shared__ __attribute__((aligned(16))) uint16_t sGc[64 * kLdsRow];
    const uint32_t l = blockIdx.y;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t mt = wave >> 1, nt = wave & 1u;
    const uint32_t s0 = blockIdx.x * chunk;
    const uint32_t s1 = min(batch, s0 + chunk);
    v16f acc = v16f{0};
    for(uint32_t sb = s0; sb < s1; sb += 64u)
    {
        for(uint32_t e = threadIdx.x; e < 64u * 16u; e += 256u)
        {
            uint32_t rr = e >> 4, c4 = (e & 15u) * 4u;
            uint32_t s = sb + rr;
            uint2 xv = make_uint2(0, 0), gv = make_uint2(0, 0);
            if(s < s1)
            {
                xv = *reinterpret_cast<const uint2*>(tb.x + (size_t(l) * batch + s) * kW + c4);
                gv = *reinterpret_cast<const uint2*>(tb.dz + (size_t(l) * batch + s) * kW + c4);
            }
            *reinterpret_cast<uint2*>(sXc + rr * kLdsRow + c4) = xv;
            *reinterpret_cast<uint2*>(sGc + rr * kLdsRow + c4) = gv;
        }
        __syncthreads();
        const uint32_t kRow = mt * 32u + (lane & 31u);
        const uint32_t nCol = nt * 32u + (lane & 31u);
        #pragma unroll
        for(uint32_t kt = 0; kt < 4; kt++)
        {
            // reduction over 16 samples per instruction: lane half h takes
            // samples 8h .. 8h+7 of the step
            const uint32_t sr = kt * 16u + (lane >> 5) * 8u;
            v8bf a, b;
            #pragma unroll
            for(uint32_t j = 0; j < 8; j++)
            {
                a[j] = __builtin_bit_cast(__bf16, sXc[(sr + j) * kLdsRow + kRow]);
                b[j] = __builtin_bit_cast(__bf16, sGc[(sr + j) * kLdsRow + nCol]);
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    float* g = wgrad + size_t(l) * kLayerParams;
    #pragma unroll
    for(uint32_t e = 0; e < 16; e++)
    {
        uint32_t k = mt * 32u + AccRow(e);
        uint32_t nn = nt * 32u + (lane & 31u);
        if(acc[e] != 0.0f) atomicAdd(g + k * kW + nn, acc[e]);
    }
}

// ---- grid gradients -----------------------------------------------------------
// One thread per (sample, level): the 16 atomics of a level's 8 corners.
__global__ void __launch_bounds__(256)
KNrcGridBwd(NrcView v, const float* __restrict__ rec, const float* __restrict__ dx0,
            uint32_t batch, float* __restrict__ ggrad)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = t / v.levels, l = t % v.levels;
    if(s >= batch) return;
    const float g0 = dx0[size_t(s) * 32u + 2u * l], g1 = dx0[size_t(s) * 32u + 2u * l + 1u];
    if(g0 == 0.0f && g1 == 0.0f) return;
    const float* r = rec + size_t(s) * 11u;
    float pos[3] = {ClampF(r[0], 0.0f, 1.0f), ClampF(r[1], 0.0f, 1.0f), ClampF(r[2], 0.0f, 1.0f)};
    Corners c;
    LevelCorners(v, l, pos, c);
    float* gl = ggrad + (size_t(l) << v.log2T) * 2u;
    for(uint32_t k = 0; k < 8; k++)
    {
        atomicAdd(gl + size_t(c.idx[k]) * 2u, c.w[k] * g0);
        atomicAdd(gl + size_t(c.idx[k]) * 2u + 1u, c.w[k] * g1);
    }
}

// ---- Adam ---------------------------------------------------------------------
struct AdamArgs
{
    float lr, b1, b2, eps, c1, c2;   // c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)
};

__device__ __forceinline__ float AdamOne(float& p, float& m, float& vv, float& g, const AdamArgs& a)
{
    float gg = g;
    g = 0.0f;
    m = a.b1 * m + (1.0f - a.b1) * gg;
    vv = a.b2 * vv + (1.0f - a.b2) * gg * gg;
    float mh = m * a.c1, vh = vv * a.c2;
    p = p - a.lr * mh / (sqrtf(vh) + a.eps);
    return p;
}

__global__ void __launch_bounds__(256)
KNrcAdamMlp(float* __restrict__ w, float* __restrict__ m, float* __restrict__ vv,
            float* __restrict__ g, uint16_t* __restrict__ wT, uint16_t* __restrict__ wK, AdamArgs a)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= kMlpParams) return;
    uint32_t l = i / kLayerParams, e = i % kLayerParams, k = e / kW, n = e % kW;
    float p = w[i], mm = m[i], v2 = vv[i], gg = g[i];
    if(l == kNrcLayers - 1u && n >= 3u) { p = 0.0f; gg = 0.0f; }
    AdamOne(p, mm, v2, gg, a);
    if(l == kNrcLayers - 1u && n >= 3u) p = 0.0f;
    w[i] = p; m[i] = mm; vv[i] = v2; g[i] = 0.0f;
    uint16_t b = F2Bf(p);
    wK[i] = b;
    wT[l * kLayerParams + n * kW + k] = b;
}

// Four parameters per thread (count is a multiple of 4).  LAZY: entries
// with a zero gradient keep parameter and moments, and are not written.
template<bool LAZY>
__global__ void __launch_bounds__(256)
KNrcAdamGrid(float* __restrict__ p, float* __restrict__ m, float* __restrict__ vv,
             float* __restrict__ g, __half* __restrict__ h, size_t count, AdamArgs a)
{
    const size_t n4 = count / 4u;
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for(; i < n4; i += size_t(gridDim.x) * blockDim.x)
    {
        float4 g4 = reinterpret_cast<const float4*>(g)[i];
        if(LAZY && g4.x == 0.0f && g4.y == 0.0f && g4.z == 0.0f && g4.w == 0.0f) continue;
        float4 p4 = reinterpret_cast<const float4*>(p)[i];
        float4 m4 = reinterpret_cast<const float4*>(m)[i];
        float4 v4 = reinterpret_cast<const float4*>(vv)[i];
        float* pp = &p4.x;
        float* mm = &m4.x;
        float* v2 = &v4.x;
        float* gg = &g4.x;
        for(int k = 0; k < 4; k++)
            if(!LAZY || gg[k] != 0.0f) AdamOne(pp[k], mm[k], v2[k], gg[k], a);
        reinterpret_cast<float4*>(p)[i] = p4;
        reinterpret_cast<float4*>(m)[i] = m4;
        reinterpret_cast<float4*>(vv)[i] = v4;
        reinterpret_cast<float4*>(g)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        __half2 h01 = __floats2half2_rn(p4.x, p4.y), h23 = __floats2half2_rn(p4.z, p4.w);
        reinterpret_cast<__half2*>(h)[2 * i] = h01;
        reinterpret_cast<__half2*>(h)[2 * i + 1] = h23;
    }
}

__global__ void KNrcZero(float* __restrict__ p, size_t n)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for(; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0.0f;
}

__global__ void KNrcRefresh(const float* __restrict__ w, uint16_t* __restrict__ wT,
                            uint16_t* __restrict__ wK, const float* __restrict__ grid,
                            __half* __restrict__ h, size_t gridCount)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for(size_t j = i; j < kMlpParams; j += size_t(gridDim.x) * blockDim.x)
    {
        uint32_t l = uint32_t(j / kLayerParams), e = uint32_t(j % kLayerParams);
        uint32_t k = e / kW, n = e % kW;
        uint16_t b = F2Bf(w[j]);
        wK[j] = b;
        wT[l * kLayerParams + n * kW + k] = b;
    }
    for(size_t j = i; j < gridCount; j += size_t(gridDim.x) * blockDim.x) h[j] = __float2half(grid[j]);
}

} // namespace mray

// ===========================================================================
// Object + C-ABI
// ===========================================================================
using namespace mray;

struct MRayNrc
{
    MRayNrcConfig cfg{};
    DeviceArena arena;
    NrcView view{};
    uint16_t* wT = nullptr;
    uint16_t* wK = nullptr;
    float* w = nullptr;
    float* wm = nullptr;
    float* wv = nullptr;
    float* wg = nullptr;
    __half* gridH = nullptr;
    float* grid = nullptr;
    float* gm = nullptr;
    float* gv = nullptr;
    float* gg = nullptr;
    size_t gridCount = 0;
    NrcTrainBuf tb{};
    uint32_t maxBatch = 0;
    uint64_t step = 0;
    float* hLoss = nullptr;       // pinned
    uint32_t fwdBlocks = 0;
    // per-kernel timing (mray_nrc_set_profiling): event pairs around each
    // launch, harvested by mray_nrc_kernel_times
    bool profiling = false;
    std::vector<hipEvent_t> evPool;
    std::vector<std::pair<uint32_t, std::pair<hipEvent_t, hipEvent_t>>> evPending;
    double kMs[kNrcKernelKinds] = {};
    uint64_t kLaunches[kNrcKernelKinds] = {};
};

// Event pair around one launch of kernel kind k (profiling only).
struct NrcTimer
{
    MRayNrc* n;
    uint32_t k;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    NrcTimer(MRayNrc* n_, uint32_t k_, hipStream_t s_) : n(n_), k(k_), s(s_)
    {
        if(!n->profiling) return;
        a = Take();
        b = Take();
        if(a && b) (void)hipEventRecord(a, s);
    }
    ~NrcTimer()
    {
        if(!a || !b) return;
        (void)hipEventRecord(b, s);
        n->evPending.push_back({k, {a, b}});
    }
    hipEvent_t Take()
    {
        if(!n->evPool.empty()) { hipEvent_t e = n->evPool.back(); n->evPool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        return (hipEventCreate(&e) == hipSuccess) ? e : nullptr;
    }
};

static int NrcLaunchCheck() { MRAY_HIP_CHECK(hipGetLastError()); return MRAY_OK; }


extern "C" {

int mray_nrc_create(const MRayNrcConfig* hCfg, MRayNrc** out)
{
    if(!hCfg || !out) { SetError("nrc_create: NULL argument"); return MRAY_ERR_INVALID; }
    *out = nullptr;
    const MRayNrcConfig& c = *hCfg;
    if(c.levels == 0 || c.levels > kNrcMaxLevels || c.log2TableSize < 10 || c.log2TableSize > 24 ||
       c.minResolution < 2 || c.maxResolution < c.minResolution || c.maxBatch == 0 ||
       (c.maxBatch % 32u) != 0)
    {
        SetError("nrc_create: levels 1..16, log2TableSize 10..24, 2 <= minRes <= maxRes, "
                 "maxBatch a non-zero multiple of 32");
        return MRAY_ERR_INVALID;
    }
    MRayNrc* n = new MRayNrc();
    n->cfg = c;
    NrcView& v = n->view;
    v.levels = c.levels;
    v.log2T = c.log2TableSize;
    const double b = (c.levels > 1)
        ? std::exp((std::log(double(c.maxResolution)) - std::log(double(c.minResolution))) / (c.levels - 1))
        : 1.0;
    for(uint32_t l = 0; l < kNrcMaxLevels; l++)
    {
        float s = (l < c.levels) ? float(double(c.minResolution) * std::pow(b, double(l)) - 1.0) : 1.0f;
        v.scale[l] = s;
        v.res[l] = uint32_t(std::ceil(s)) + 2u;
        uint64_t cells = uint64_t(v.res[l]) * v.res[l] * v.res[l];
        v.dense[l] = (cells <= (1ull << c.log2TableSize)) ? 1u : 0u;
    }
    n->gridCount = size_t(c.levels) << c.log2TableSize;
    n->gridCount *= kNrcFeat;
    n->maxBatch = c.maxBatch;
    auto A = [&](void** p, size_t bytes) { return DevAlloc(n->arena, p, bytes); };
    int rv = MRAY_OK;
    auto fail = [&](int r) { n->arena.Release(); if(n->hLoss) (void)hipHostFree(n->hLoss); delete n; return r; };
    if((rv = A((void**)&n->wT, kMlpParams * 2)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->wK, kMlpParams * 2)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->w, kMlpParams * 4)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->wm, kMlpParams * 4)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->wv, kMlpParams * 4)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->wg, kMlpParams * 4)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->gridH, n->gridCount * 2)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->grid, n->gridCount * 4)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->gm, n->gridCount * 4)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->gv, n->gridCount * 4)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->gg, n->gridCount * 4)) != MRAY_OK) return fail(rv);
    const size_t B = c.maxBatch;
    if((rv = A((void**)&n->tb.x, kNrcLayers * B * kW * 2)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->tb.dz, kNrcLayers * B * kW * 2)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->tb.dy, B * 3 * 4)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->tb.dx0, B * 32 * 4)) != MRAY_OK) return fail(rv);
    if((rv = A((void**)&n->tb.loss, 4)) != MRAY_OK) return fail(rv);
    if(hipHostMalloc((void**)&n->hLoss, sizeof(float), hipHostMallocDefault) != hipSuccess)
    {
        n->hLoss = nullptr;
        SetError("nrc_create: pinned allocation failed");
        return fail(MRAY_ERR_HIP);
    }
    v.wT = n->wT;
    v.wK = n->wK;
    v.grid = n->gridH;
    // initial parameters: grid U(-1e-4, 1e-4); layers U(-a, a), a = sqrt(6/(in+out))
    // (Glorot); the output layer keeps only columns 0..2
    {
        std::mt19937 mt(uint32_t(c.seed ^ (c.seed >> 32)));
        std::uniform_real_distribution<float> ug(-1e-4f, 1e-4f);
        std::vector<float> hg(n->gridCount);
        for(float& x : hg) x = ug(mt);
        std::vector<float> hw(kMlpParams, 0.0f);
        for(uint32_t l = 0; l < kNrcLayers; l++)
        {
            uint32_t fanOut = (l + 1 == kNrcLayers) ? 3u : kW;
            float a = std::sqrt(6.0f / float(kW + fanOut));
            std::uniform_real_distribution<float> uw(-a, a);
            for(uint32_t k = 0; k < kW; k++)
                for(uint32_t o = 0; o < fanOut; o++) hw[l * kLayerParams + k * kW + o] = uw(mt);
        }
        if(hipMemcpy(n->grid, hg.data(), n->gridCount * 4, hipMemcpyHostToDevice) != hipSuccess ||
           hipMemcpy(n->w, hw.data(), kMlpParams * 4, hipMemcpyHostToDevice) != hipSuccess)
        {
            SetError("nrc_create: upload failed");
            return fail(MRAY_ERR_HIP);
        }
    }
    KNrcZero<<<256, 256>>>(n->wm, kMlpParams);
    KNrcZero<<<256, 256>>>(n->wv, kMlpParams);
    KNrcZero<<<256, 256>>>(n->wg, kMlpParams);
    KNrcZero<<<2048, 256>>>(n->gm, n->gridCount);
    KNrcZero<<<2048, 256>>>(n->gv, n->gridCount);
    KNrcZero<<<2048, 256>>>(n->gg, n->gridCount);
    KNrcRefresh<<<2048, 256>>>(n->w, n->wT, n->wK, n->grid, n->gridH, n->gridCount);
    if(hipDeviceSynchronize() != hipSuccess || hipGetLastError() != hipSuccess)
    {
        SetError("nrc_create: initialisation kernels failed");
        return fail(MRAY_ERR_HIP);
    }
    int dev = 0;
    hipDeviceProp_t prop;
    if(hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return fail(MRAY_ERR_HIP);
    n->fwdBlocks = uint32_t(std::max(1, prop.multiProcessorCount)) * 2u;
    *out = n;
    return MRAY_OK;
}

int mray_nrc_destroy(MRayNrc* n)
{
    if(!n) return MRAY_OK;
    (void)hipDeviceSynchronize();
    if(n->hLoss) (void)hipHostFree(n->hLoss);
    for(auto& p : n->evPending)
    {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    for(hipEvent_t e : n->evPool) (void)hipEventDestroy(e);
    n->arena.Release();
    delete n;
    return MRAY_OK;
}

int mray_nrc_infer(MRayNrc* n, const float* dRecords, uint32_t count, const uint32_t* dCount,
                   float* dOut, void* stream)
{
    if(!n || !dRecords || !dOut) { SetError("nrc_infer: NULL argument"); return MRAY_ERR_INVALID; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t blocks = dCount ? n->fwdBlocks
                             : std::min(n->fwdBlocks, DivUp(std::max(count, 1u), 32u * kFwdWaves));
    {
        NrcTimer t(n, 0, s);
        KNrcForward<false><<<blocks, kFwdWaves * 64, 0, s>>>(n->view, dRecords, dCount, count, dOut,
                                                             nullptr, NrcTrainBuf{}, 0);
    }
    return NrcLaunchCheck();
}

int mray_nrc_backward(MRayNrc* n, const float* dRecords, const float* dTargets, uint32_t count,
                      void* stream)
{
    if(!n || !dRecords || !dTargets) { SetError("nrc_backward: NULL argument"); return MRAY_ERR_INVALID; }
    if(count == 0 || count > n->maxBatch || (count % 32u) != 0)
    {
        SetError("nrc_backward: count must be a non-zero multiple of 32, at most maxBatch");
        return MRAY_ERR_INVALID;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    KNrcZero<<<1, 64, 0, s>>>(n->tb.loss, 1);
    uint32_t tiles = count / 32u;
    uint32_t blocks = std::min(n->fwdBlocks, DivUp(tiles, kFwdWaves));
    {
        NrcTimer t(n, 1, s);
        KNrcForward<true><<<blocks, kFwdWaves * 64, 0, s>>>(n->view, dRecords, nullptr, count, nullptr,
                                                            dTargets, n->tb, count);
    }
    {
        NrcTimer t(n, 2, s);
        KNrcBackward<<<blocks, kFwdWaves * 64, 0, s>>>(n->view, n->tb, count);
    }
    uint32_t chunk = 512;
    dim3 wg(DivUp(count, chunk), kNrcLayers);
    {
        NrcTimer t(n, 3, s);
        KNrcWgrad<<<wg, 256, 0, s>>>(n->tb, count, chunk, n->wg);
    }
    {
        NrcTimer t(n, 4, s);
        KNrcGridBwd<<<DivUp(size_t(count) * n->view.levels, 256u), 256, 0, s>>>(n->view, dRecords,
                                                                             n->tb.dx0, count, n->gg);
    }
    return NrcLaunchCheck();
}

int mray_nrc_adam(MRayNrc* n, void* stream)
{
    if(!n) { SetError("nrc_adam: NULL argument"); return MRAY_ERR_INVALID; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    n->step++;
    AdamArgs a;
    a.lr = n->cfg.learningRate;
    a.b1 = n->cfg.beta1;
    a.b2 = n->cfg.beta2;
    a.eps = n->cfg.epsilon;
    a.c1 = float(1.0 / (1.0 - std::pow(double(a.b1), double(n->step))));
    a.c2 = float(1.0 / (1.0 - std::pow(double(a.b2), double(n->step))));
    NrcTimer t(n, 5, s);
    KNrcAdamMlp<<<DivUp(kMlpParams, 256u), 256, 0, s>>>(n->w, n->wm, n->wv, n->wg, n->wT, n->wK, a);
    if(n->cfg.flags & MRAY_NRC_LAZY_ADAM)
        KNrcAdamGrid<true><<<4096, 256, 0, s>>>(n->grid, n->gm, n->gv, n->gg, n->gridH, n->gridCount, a);
    else
        KNrcAdamGrid<false><<<4096, 256, 0, s>>>(n->grid, n->gm, n->gv, n->gg, n->gridH, n->gridCount, a);
    return NrcLaunchCheck();
}

int mray_nrc_train(MRayNrc* n, const float* dRecords, const float* dTargets, uint32_t count,
                   void* stream)
{
    MRAY_TRY(mray_nrc_backward(n, dRecords, dTargets, count, stream));
    return mray_nrc_adam(n, stream);
}

int mray_nrc_last_loss(MRayNrc* n, float* hLoss)
{
    if(!n || !hLoss) { SetError("nrc_last_loss: NULL argument"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    MRAY_HIP_CHECK(hipMemcpy(hLoss, n->tb.loss, sizeof(float), hipMemcpyDeviceToHost));
    return MRAY_OK;
}

int mray_nrc_param_counts(const MRayNrc* n, uint64_t* hMlp, uint64_t* hGrid)
{
    if(!n || !hMlp || !hGrid) { SetError("nrc_param_counts: NULL argument"); return MRAY_ERR_INVALID; }
    *hMlp = kMlpParams;
    *hGrid = n->gridCount;
    return MRAY_OK;
}

int mray_nrc_read(MRayNrc* n, uint32_t which, float* hOut)
{
    if(!n || !hOut) { SetError("nrc_read: NULL argument"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    const float* src = nullptr;
    size_t cnt = 0;
    switch(which)
    {
        case MRAY_NRC_WEIGHTS: src = n->w; cnt = kMlpParams; break;
        case MRAY_NRC_WEIGHT_GRADS: src = n->wg; cnt = kMlpParams; break;
        case MRAY_NRC_GRID: src = n->grid; cnt = n->gridCount; break;
        case MRAY_NRC_GRID_GRADS: src = n->gg; cnt = n->gridCount; break;
        default: SetError("nrc_read: unknown buffer"); return MRAY_ERR_INVALID;
    }
    MRAY_HIP_CHECK(hipMemcpy(hOut, src, cnt * 4, hipMemcpyDeviceToHost));
    return MRAY_OK;
}

int mray_nrc_write(MRayNrc* n, uint32_t which, const float* hIn)
{
    if(!n || !hIn) { SetError("nrc_write: NULL argument"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    if(which == MRAY_NRC_WEIGHTS)
        MRAY_HIP_CHECK(hipMemcpy(n->w, hIn, kMlpParams * 4, hipMemcpyHostToDevice));
    else if(which == MRAY_NRC_GRID)
        MRAY_HIP_CHECK(hipMemcpy(n->grid, hIn, n->gridCount * 4, hipMemcpyHostToDevice));
    else { SetError("nrc_write: only weights and grid are writable"); return MRAY_ERR_INVALID; }
    KNrcRefresh<<<2048, 256>>>(n->w, n->wT, n->wK, n->grid, n->gridH, n->gridCount);
    MRAY_HIP_CHECK(hipGetLastError());
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    return MRAY_OK;
}

int mray_nrc_set_profiling(MRayNrc* n, uint32_t on)
{
    if(!n) { SetError("nrc_set_profiling: NULL argument"); return MRAY_ERR_INVALID; }
    n->profiling = on != 0;
    return MRAY_OK;
}

int mray_nrc_kernel_times(MRayNrc* n, double* hMs, uint64_t* hLaunches)
{
    if(!n || !hMs || !hLaunches) { SetError("nrc_kernel_times: NULL argument"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    for(auto& p : n->evPending)
    {
        float ms = 0.0f;
        MRAY_HIP_CHECK(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        n->kMs[p.first] += ms;
        n->kLaunches[p.first] += 1;
        n->evPool.push_back(p.second.first);
        n->evPool.push_back(p.second.second);
    }
    n->evPending.clear();
    for(uint32_t k = 0; k < kNrcKernelKinds; k++)
    {
        hMs[k] = n->kMs[k];
        hLaunches[k] = n->kLaunches[k];
    }
    return MRAY_OK;
}

int mray_nrc_zero_grads(MRayNrc* n, void* stream)
{
    if(!n) { SetError("nrc_zero_grads: NULL argument"); return MRAY_ERR_INVALID; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    KNrcZero<<<256, 256, 0, s>>>(n->wg, kMlpParams);
    KNrcZero<<<2048, 256, 0, s>>>(n->gg, n->gridCount);
    return NrcLaunchCheck();
}

} // extern "C"
