// mt_mlp.hip -- radiance-cache MLP kernels for gfx950: inference, training forward/backward,
// weight gradients and Adam, all on v_mfma_f32_32x32x16_bf16.
//
// Products are computed transposed, Z^T = W^T H^T: M = output features (32-row tiles),
// N = 32 samples (one per lane column), K = input features in chunks of 16.  A 32x32 result
// leaves lane (g = lane/32, n = lane%32) with rows q*8 + g*4 + j (q, j < 4) of sample n; the
// next layer's B operand wants 8 k-slots per lane group, so the input features of every layer
// are numbered in that order (mlp::perm) and the packed weights follow it: activations go from
// one product to the next without leaving registers.  The training pass keeps activations and
// gradients in [feature][32 samples] slabs per group of 32 samples, so that the weight gradient
// (samples as K) loads its operand fragments straight from them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "mt_internal.hpp"
#include "mt_mlp.hpp"

namespace mt {

using namespace mlp;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kWave = 64;
constexpr float kPiF = 3.14159265358979323846f;
constexpr float kLossEps = 0.01f;  // relative error r = (y - t) / sqrt(y^2 + eps), y held
constexpr float kHuber = 10.0f;    // loss r^2 for |r| <= 10, linear beyond (rare huge targets)

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ uint16_t bf16_bits(float x) {
    const __bf16 h = static_cast<__bf16>(x);
    uint16_t u;
    __builtin_memcpy(&u, &h, 2);
    return u;
}
__device__ __forceinline__ float bf16_float(uint16_t u) {
    return __uint_as_float(static_cast<uint32_t>(u) << 16);
}

// encoding dimension e (0..31) of a normalised position p (in [0,1]^3) and a direction d:
// e < 24: sin/cos(2^f pi p_a) for axis a = e/8, f = (e%8)/2; e >= 24: sin/cos(2^f pi u) for
// u = theta/pi (e < 28) or phi/pi (e >= 28), f = (e%4)/2
__device__ __forceinline__ float encode(int e, const float p[3], float th, float ph) {
    float x;
    int r;
    if (e < 24) {
        x = p[e >> 3];
        r = e & 7;
    } else {
        x = (e < 28) ? th : ph;
        r = (e - 24) & 3;
    }
    // |a| <= 8 pi: the hardware sine and cosine are accurate far below the BF16 rounding the
    // encoding then takes
    const float a = kPiF * x * static_cast<float>(1 << (r >> 1));
    return (r & 1) ? __cosf(a) : __sinf(a);
}

struct EncCtx {
    float lo[3], inv[3];
};

// the two K-chunks of the encoding held by this lane (its lane group's 16 dimensions)
__device__ __forceinline__ void encode_lane(const EncCtx& E, float4 pos, float4 dir, int g,
                                            bf16x8 out[2], float* h_out) {
    float p[3] = {(pos.x - E.lo[0]) * E.inv[0], (pos.y - E.lo[1]) * E.inv[1],
                  (pos.z - E.lo[2]) * E.inv[2]};
    const float dl = sqrtf(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z);
    const float cy = dl > 0.0f ? fminf(fmaxf(dir.y / dl, -1.0f), 1.0f) : 1.0f;
    const float th = acosf(cy) / kPiF;
    const float ph = atan2f(dir.z, dir.x) / kPiF;
    for (int c = 0; c < 2; ++c)
        for (int t = 0; t < 8; ++t) {
            const int e = perm(c, g, t);
            const float v = encode(e, p, th, ph);
            out[c][t] = static_cast<__bf16>(v);
            if (h_out) h_out[c * 8 + t] = static_cast<float>(out[c][t]);
        }
}

// one dense layer: acc[mt] = W^T x + b for the lane's sample, from K-chunks in[]
template <int kTiles, int kChunks>
__device__ __forceinline__ void dense(const bf16x8* __restrict__ W, const float* __restrict__ bias,
                                      const bf16x8 in[kChunks], f32x16 acc[kTiles], int lane) {
    const int g = lane >> 5;
#pragma unroll
    for (int mt = 0; mt < kTiles; ++mt) {
        f32x16 a = {};
#pragma unroll
        for (int c = 0; c < kChunks; ++c) a = mfma(W[(mt * kChunks + c) * kWave + lane], in[c], a);
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int o = mt * 32 + q * 8 + g * 4 + j;
                a[q * 4 + j] += bias ? bias[o] : 0.0f;
            }
        acc[mt] = a;
    }
}

// ReLU of two 32-row tiles into the four K-chunks of the next layer; mask bit (mt*16 + r) set
// where the pre-activation is positive
__device__ __forceinline__ void relu_chunks(const f32x16 acc[2], bf16x8 out[4], uint32_t& mask) {
    mask = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int mt = c >> 1, q0 = (c & 1) * 2;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int r = (q0 + (t >> 2)) * 4 + (t & 3);
            const float z = acc[mt][r];
            if (z > 0.0f) mask |= 1u << (mt * 16 + r);
            out[c][t] = static_cast<__bf16>(z > 0.0f ? z : 0.0f);
        }
    }
}

// masked gradient chunks (dZ = dH * relu'(Z)) from two dH tiles
__device__ __forceinline__ void grad_chunks(const f32x16 acc[2], uint32_t mask, bf16x8 out[4]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int mt = c >> 1, q0 = (c & 1) * 2;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int r = (q0 + (t >> 2)) * 4 + (t & 3);
            out[c][t] = static_cast<__bf16>(((mask >> (mt * 16 + r)) & 1u) ? acc[mt][r] : 0.0f);
        }
    }
}

// stores the chunks of a feature vector of sample s (0..31 in its group) into the group's
// [feature][32 samples] slab; the 32 lanes of a lane group write 64 contiguous bytes per feature
__device__ __forceinline__ void store_slab(uint16_t* slab, const bf16x8 v[], int nchunks, int g, int s) {
    for (int c = 0; c < nchunks; ++c)
        for (int t = 0; t < 8; ++t) {
            const __bf16 h = v[c][t];
            uint16_t u;
            __builtin_memcpy(&u, &h, 2);
            slab[perm(c, g, t) * 32 + s] = u;
        }
}

struct Net {
    const bf16x8* fwd;  // packed forward operands
    const bf16x8* bwd;  // packed backward operands
    const float* master;
    const float* bias[kLayers];  // per layer (global master weights, or an LDS copy)
};

__host__ __device__ inline Net make_net(const uint16_t* fwd, const uint16_t* bwd, const float* master) {
    Net N{reinterpret_cast<const bf16x8*>(fwd), reinterpret_cast<const bf16x8*>(bwd), master, {}};
    for (int l = 0; l < kLayers; ++l) N.bias[l] = master + kBOff[l];
    return N;
}

// full forward of one group of 32 samples; returns the 3 outputs in lanes g == 0 (acc5[0..2])
__device__ __forceinline__ void forward(const Net& N, const bf16x8 enc[2], f32x16& out5,
                                        uint32_t masks[4], bf16x8 h[4][4], int lane) {
    f32x16 acc[2];
    dense<2, 2>(N.fwd + kFwdOff[0] / 8, N.bias[0], enc, acc, lane);
    relu_chunks(acc, h[0], masks[0]);
#pragma unroll
    for (int l = 1; l < 4; ++l) {
        dense<2, 4>(N.fwd + kFwdOff[l] / 8, N.bias[l], h[l - 1], acc, lane);
        relu_chunks(acc, h[l], masks[l]);
    }
    f32x16 o[1];
    dense<1, 4>(N.fwd + kFwdOff[4] / 8, nullptr, h[3], o, lane);
    out5 = o[0];
}

// a step's sample count: n_host, or (when n_dev is set) min(n_host, *n_dev - offset) clamped at 0
__device__ __forceinline__ uint32_t step_count(const uint32_t* n_dev, uint32_t offset, uint32_t n_host) {
    if (!n_dev) return n_host;
    const uint32_t avail = *n_dev;
    return avail > offset ? min(n_host, avail - offset) : 0u;
}

__global__ __launch_bounds__(256) void k_mlp_infer(Net N, EncCtx E, uint32_t n_host, const uint32_t* n_dev,
                                                   const float4* pos, const float4* dir, float4* out) {
    // the packed forward weights (32 KB) and the biases, staged once per block in LDS: every
    // group of 32 queries reads its 32 operand fragments from there, not from the caches
    __shared__ bf16x8 sW[kFwdSize / 8];
    __shared__ float sB[4 * 64 + 3];
    for (int i = threadIdx.x; i < kFwdSize / 8; i += blockDim.x) sW[i] = N.fwd[i];
    for (int i = threadIdx.x; i < 4 * 64 + 3; i += blockDim.x) {
        const int l = i / 64 < 4 ? i / 64 : 4;
        sB[i] = N.bias[l][i - 64 * l];
    }
    __syncthreads();
    Net L = N;
    L.fwd = sW;
    for (int l = 0; l < kLayers; ++l) L.bias[l] = sB + 64 * l;
    const uint32_t n = step_count(n_dev, 0, n_host);
    const int lane = threadIdx.x & 63, g = lane >> 5, s = lane & 31;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t base = wave * 32; base < n; base += waves * 32) {
        const uint32_t i = base + s;
        const bool ok = i < n;
        const float4 p = ok ? pos[i] : make_float4(0, 0, 0, 0);
        const float4 d = ok ? dir[i] : make_float4(0, 0, 1, 0);
        bf16x8 enc[2];
        encode_lane(E, p, d, g, enc, nullptr);
        f32x16 o;
        uint32_t masks[4];
        bf16x8 h[4][4];
        forward(L, enc, o, masks, h, lane);
        if (ok && g == 0) {
            const float* b = L.bias[4];
            out[i] = make_float4(o[0] + b[0], o[1] + b[1], o[2] + b[2], 0.0f);
        }
    }
}

// training pass for 32 samples per wave: forward keeping H0..H4, loss gradient (relative L2 or L2),
// backward through the layers keeping dZ1..dZ5 (bf16, natural feature order)
__global__ __launch_bounds__(256) void k_mlp_train(Net N, EncCtx E, uint32_t n_host, const uint32_t* n_dev,
                                                   uint32_t offset, const float4* pos,
                                                   const float4* dir, const float4* target,
                                                   CacheTrainBuffers T) {
    const uint32_t n = step_count(n_dev, offset, n_host);
    if (n == 0) return;
    pos += offset;
    dir += offset;
    target += offset;
    const float inv_count = 1.0f / static_cast<float>(n);
    const int lane = threadIdx.x & 63, g = lane >> 5, s = lane & 31;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    float loss = 0.0f;
    for (uint32_t base = wave * 32; base < n; base += waves * 32) {
        const uint32_t i = base + s;
        const bool ok = i < n;
        const float4 p = ok ? pos[i] : make_float4(0, 0, 0, 0);
        const float4 d = ok ? dir[i] : make_float4(0, 0, 1, 0);
        bf16x8 enc[2];
        encode_lane(E, p, d, g, enc, nullptr);
        f32x16 o;
        uint32_t masks[4];
        bf16x8 h[4][4];
        forward(N, enc, o, masks, h, lane);
        // the group's slabs: every lane stores, lanes past n with dZ = 0 (finite activations
        // of a zero input), so that they add nothing to the weight gradients
        uint16_t* Hs = T.H + static_cast<size_t>(base / 32) * kHWidth * 32;
        uint16_t* Ds = T.DZ + static_cast<size_t>(base / 32) * kDZWidth * 32;
        store_slab(Hs + kHOff[0] * 32, enc, 2, g, s);
        for (int l = 1; l < 5; ++l) store_slab(Hs + kHOff[l] * 32, h[l - 1], 4, g, s);
        // output gradient (lanes g == 0 hold outputs 0..2 in o[0..2])
        f32x16 dz5 = {};
        if (g == 0 && ok) {
            const float* b = N.master + kBOff[4];
            const float4 t = target[i];
            const float tv[3] = {t.x, t.y, t.z};
            for (int c = 0; c < 3; ++c) {
                const float y = o[c] + b[c];
                if (T.l2) {  // L2: mean over samples and channels of (y - t)^2
                    const float e = y - tv[c];
                    loss += e * e;
                    dz5[c] = 2.0f * e * inv_count / 3.0f;
                    continue;
                }
                const float sd = sqrtf(y * y + kLossEps);
                const float r = (y - tv[c]) / sd;
                const float ar = fabsf(r);
                loss += ar <= kHuber ? r * r : kHuber * (2.0f * ar - kHuber);
                // d/dy of the Huber loss of r with the denominator held, averaged over samples
                // and channels
                dz5[c] = 2.0f * fminf(fmaxf(r, -kHuber), kHuber) / sd * inv_count / 3.0f;
            }
        }
        // dZ5 chunk 0 of the output layer: k-slots t<4 of group 0 are outputs 0..3
        bf16x8 dzc[4];
        for (int t = 0; t < 8; ++t) dzc[0][t] = static_cast<__bf16>(t < 4 ? dz5[t] : 0.0f);
        if (g == 0)
            for (int c = 0; c < 4; ++c) Ds[(kDZOff[4] + c) * 32 + s] = c < 3 ? bf16_bits(dz5[c]) : 0;
        // layer 5 backward: dH4 = W5 dZ5 (K = one chunk of the 3 outputs)
        f32x16 acc[2];
        for (int mt = 0; mt < 2; ++mt) {
            f32x16 a = {};
            a = mfma(N.bwd[(kBwdOff[4] / 8) + (mt * 1 + 0) * kWave + lane], dzc[0], a);
            acc[mt] = a;
        }
        for (int l = 3; l >= 0; --l) {
            bf16x8 dz[4];
            grad_chunks(acc, masks[l], dz);
            store_slab(Ds + kDZOff[l] * 32, dz, 4, g, s);
            if (l == 0) break;
            for (int mt = 0; mt < 2; ++mt) {
                f32x16 a = {};
                for (int c = 0; c < 4; ++c)
                    a = mfma(N.bwd[(kBwdOff[l] / 8) + (mt * 4 + c) * kWave + lane], dz[c], a);
                acc[mt] = a;
            }
        }
    }
    // per-block loss sum
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    __shared__ float red[4];
    if (lane == 0) red[threadIdx.x >> 6] = loss;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(T.loss, (red[0] + red[1] + red[2] + red[3]) * inv_count / 3.0f);
}

// weight gradients: every block sums dW = dZ^T H over its range of 32-sample groups with the
// samples as K.  The slabs are [feature][32 samples], so the A fragment of lane (g, r) for an
// output tile (layer l, rows o) is dZ feature o over samples c*16 + g*8 .. +7 -- 16 contiguous
// bytes -- and the B fragment is H feature i over the same samples: both load straight from
// memory.  The bias gradient of output o is the sum of the A fragments of the tile with i
// tile 0.  Each block writes its partial sums.
__global__ __launch_bounds__(256) void k_mlp_wgrad(CacheTrainBuffers T, uint32_t n_host, const uint32_t* n_dev,
                                                   uint32_t offset) {
    const uint32_t n = step_count(n_dev, offset, n_host);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 5, r = lane & 31;
    // 16 output tiles: layer, o tile, i tile -- 4 per wave
    int tl[4], to[4], ti[4];
    {
        int k = 0, idx = 0;
        for (int l = 0; l < kLayers; ++l)
            for (int a = 0; a < (kOut[l] + 31) / 32; ++a)
                for (int b = 0; b < kIn[l] / 32; ++b, ++k)
                    if ((k & 3) == w) {
                        tl[idx] = l;
                        to[idx] = a;
                        ti[idx] = b;
                        ++idx;
                    }
    }
    f32x16 acc[4] = {};
    float bias_acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t groups = (n + 31) / 32;
    const uint32_t per = (groups + gridDim.x - 1) / gridDim.x;
    const uint32_t g0 = min(groups, blockIdx.x * per), g1 = min(groups, g0 + per);
    for (uint32_t grp = g0; grp < g1; ++grp) {
        const uint16_t* Hs = T.H + static_cast<size_t>(grp) * kHWidth * 32;
        const uint16_t* Ds = T.DZ + static_cast<size_t>(grp) * kDZWidth * 32;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int l = tl[k];
            const int o = to[k] * 32 + r;  // A row: output feature
            const int i = ti[k] * 32 + r;  // B column: input feature
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int sm = c * 16 + g * 8;
                bf16x8 a;
                if (o < kOut[l]) {
                    a = *reinterpret_cast<const bf16x8*>(Ds + (kDZOff[l] + o) * 32 + sm);
                    if (ti[k] == 0)
                        for (int t = 0; t < 8; ++t) bias_acc[k] += static_cast<float>(a[t]);
                } else {
                    for (int t = 0; t < 8; ++t) a[t] = static_cast<__bf16>(0.0f);
                }
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(Hs + (kHOff[l] + i) * 32 + sm);
                acc[k] = mfma(a, b, acc[k]);
            }
        }
    }
    float* out = T.partial + static_cast<size_t>(blockIdx.x) * kParams;
    for (int k = 0; k < 4; ++k) {
        const int l = tl[k];
        for (int q = 0; q < 4; ++q)
            for (int j = 0; j < 4; ++j) {
                const int o = to[k] * 32 + q * 8 + g * 4 + j;
                const int i = ti[k] * 32 + r;
                if (o < kOut[l]) out[kWOff[l] + o * kIn[l] + i] = acc[k][q * 4 + j];
            }
        // lane groups 0 and 1 hold the two halves of every group's samples
        const float bsum = bias_acc[k] + __shfl_xor(bias_acc[k], 32);
        const int o = to[k] * 32 + r;
        if (ti[k] == 0 && g == 0 && o < kOut[l]) out[kBOff[l] + o] = bsum;
    }
}

// Adam on the fp32 master weights from the sum of the blocks' partial gradients
// (64 parameters per block; its four waves sum a quarter of the partials each)
__global__ __launch_bounds__(256) void k_mlp_adam(float* master, float* m, float* v, const float* partial,
                                                  uint32_t blocks, float lr, float b1, float b2, float eps,
                                                  float c1, float c2, float* grad_out, uint32_t n_host,
                                                  const uint32_t* n_dev, uint32_t offset,
                                                  unsigned long long* trained) {
    const uint32_t n = step_count(n_dev, offset, n_host);
    if (n == 0) return;  // an empty step changes nothing (uniform over the grid)
    const uint32_t p = blockIdx.x * 64 + (threadIdx.x & 63), quarter = threadIdx.x >> 6;
    __shared__ float red[4][64];
    float part = 0.0f;
    if (p < static_cast<uint32_t>(kParams))
        for (uint32_t b = quarter; b < blocks; b += 4) part += partial[static_cast<size_t>(b) * kParams + p];
    red[quarter][threadIdx.x & 63] = part;
    __syncthreads();
    if (quarter != 0 || p >= static_cast<uint32_t>(kParams)) return;
    if (p == 0 && trained) atomicAdd(trained, static_cast<unsigned long long>(n));
    const uint32_t t = threadIdx.x & 63;
    const float gsum = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
    if (grad_out) grad_out[p] = gsum;
    const float mm = b1 * m[p] + (1.0f - b1) * gsum;
    const float vv = b2 * v[p] + (1.0f - b2) * gsum * gsum;
    m[p] = mm;
    v[p] = vv;
    master[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// bf16 operand packs from the master weights
__global__ void k_mlp_pack(const float* master, uint16_t* fwd, uint16_t* bwd) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kFwdSize) {
        int l = 0;
        while (l + 1 < kLayers && e >= kFwdOff[l + 1]) ++l;
        const int local = e - kFwdOff[l];
        const int t = local & 7, lane = (local >> 3) & 63, tc = local >> 9;
        const int c = tc % kFwdChunks[l], mt = tc / kFwdChunks[l];
        const int o = mt * 32 + (lane & 31);
        const int i = perm(c, lane >> 5, t);
        const float w = (o < kOut[l]) ? master[kWOff[l] + o * kIn[l] + i] : 0.0f;
        fwd[e] = bf16_bits(w);
    }
    if (e < kBwdSize) {
        int l = 1;
        while (l + 1 < kLayers && e >= kBwdOff[l + 1]) ++l;
        const int local = e - kBwdOff[l];
        const int t = local & 7, lane = (local >> 3) & 63, tc = local >> 9;
        const int c = tc % kBwdChunks[l], mt = tc / kBwdChunks[l];
        const int i = mt * 32 + (lane & 31);   // row: input feature of layer l
        const int o = perm(c, lane >> 5, t);   // k-slot: output feature of layer l
        const float w = (o < kOut[l]) ? master[kWOff[l] + o * kIn[l] + i] : 0.0f;
        bwd[e] = bf16_bits(w);
    }
}

__global__ void k_zero(float* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

template <typename T>
int dalloc(std::vector<void*>& owned, T** p, size_t n) {
    void* q = nullptr;
    if (device_malloc(&q, std::max<size_t>(n, 1) * sizeof(T)))
        return fail(MT_E_NOMEM, "device allocation of cache buffers failed");
    owned.push_back(q);
    *p = static_cast<T*>(q);
    return MT_OK;
}

}  // namespace

Mlp::~Mlp() {
    for (void* p : owned_) (void)hipFree(p);
}

int Mlp::init(uint32_t seed, uint32_t train_capacity, float lr) {
    lr_ = lr;
    train_capacity = (train_capacity + 31) & ~31u;
    tb_.capacity = train_capacity;
    tb_.blocks = 256;  // weight-gradient blocks: one per CU, each summing its slabs into a partial
    int rc;
    if ((rc = dalloc(owned_, &master_, kParams)) || (rc = dalloc(owned_, &adam_m_, kParams)) ||
        (rc = dalloc(owned_, &adam_v_, kParams)) || (rc = dalloc(owned_, &fwd_, kFwdSize)) ||
        (rc = dalloc(owned_, &bwd_, kBwdSize)) ||
        (rc = dalloc(owned_, &tb_.H, static_cast<size_t>(train_capacity) * kHWidth)) ||
        (rc = dalloc(owned_, &tb_.DZ, static_cast<size_t>(train_capacity) * kDZWidth)) ||
        (rc = dalloc(owned_, &tb_.partial, static_cast<size_t>(tb_.blocks) * kParams)) ||
        (rc = dalloc(owned_, &tb_.loss, 1)) || (rc = dalloc(owned_, &trained_, 1)) ||
        (rc = dalloc(owned_, &scale_, 2)))
        return rc;
    // the counter of trained samples starts at zero and the output scale at {1, not set}
    // (device memory is not cleared)
    const float sc0[2] = {1.0f, 0.0f};
    if (hipMemset(trained_, 0, sizeof(*trained_)) != hipSuccess ||
        hipMemcpy(scale_, sc0, sizeof(sc0), hipMemcpyHostToDevice) != hipSuccess)
        return fail(MT_E_HIP, "cache counter reset failed");
    // He-uniform initial weights, zero biases
    std::vector<float> w(kParams, 0.0f);
    std::mt19937 rng(seed);
    for (int l = 0; l < kLayers; ++l) {
        const float a = std::sqrt(6.0f / static_cast<float>(kIn[l]));
        std::uniform_real_distribution<float> U(-a, a);
        for (int k = 0; k < kOut[l] * kIn[l]; ++k) w[kWOff[l] + k] = (l == kLayers - 1) ? 0.1f * U(rng) : U(rng);
    }
    return set_master(w.data());
}

int Mlp::set_master(const float* in) {
    if (hipMemcpy(master_, in, sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess)
        return fail(MT_E_HIP, "copy of cache weights failed");
    std::vector<float> z(kParams, 0.0f);
    if (hipMemcpy(adam_m_, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(adam_v_, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess)
        return fail(MT_E_HIP, "reset of the optimiser state failed");
    step_ = 0;
    const int rc = pack(nullptr);
    if (rc) return rc;
    return hipDeviceSynchronize() == hipSuccess ? MT_OK : fail(MT_E_HIP, "cache pack failed");
}

uint64_t Mlp::trained_samples() {
    unsigned long long v = 0;
    if (trained_) (void)hipMemcpy(&v, trained_, sizeof(v), hipMemcpyDeviceToHost);
    return v;
}

int Mlp::get_master(float* out) {
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out, master_, sizeof(float) * kParams, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(MT_E_HIP, "copy of cache weights failed");
    return MT_OK;
}

void Mlp::set_bounds(const float lo[3], const float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo_[a] = lo[a];
        const float ext = hi[a] - lo[a];
        inv_ext_[a] = ext > 0.0f ? 1.0f / ext : 1.0f;
    }
}

int Mlp::pack(hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((kFwdSize + 255) / 256), dim3(256), 0, s, master_, fwd_, bwd_);
    return hipGetLastError() == hipSuccess ? MT_OK : fail(MT_E_HIP, "k_mlp_pack launch failed");
}

int Mlp::infer(uint32_t n, const float4* pos, const float4* dir, float4* out, hipStream_t s,
                const uint32_t* n_dev) {
    if (n == 0) return MT_OK;
    const Net N = make_net(fwd_, bwd_, master_);
    EncCtx E{{lo_[0], lo_[1], lo_[2]}, {inv_ext_[0], inv_ext_[1], inv_ext_[2]}};
    const uint32_t groups = (n + 31) / 32;
    const uint32_t blocks = std::min<uint32_t>((groups + 3) / 4, 4096);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, s, N, E, n, n_dev, pos, dir, out);
    return hipGetLastError() == hipSuccess ? MT_OK : fail(MT_E_HIP, "k_mlp_infer launch failed");
}

int Mlp::train(uint32_t n, const float4* pos, const float4* dir, const float4* target, hipStream_t s,
               float* loss_host, float* grad_host, const uint32_t* n_dev, uint32_t offset) {
    if (n == 0) return MT_OK;
    if (n > tb_.capacity) return fail(MT_E_INVALID, "training batch exceeds the cache's capacity");
    const Net N = make_net(fwd_, bwd_, master_);
    EncCtx E{{lo_[0], lo_[1], lo_[2]}, {inv_ext_[0], inv_ext_[1], inv_ext_[2]}};
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, tb_.loss, 1u);
    const uint32_t groups = (n + 31) / 32;
    const uint32_t blocks = std::min<uint32_t>((groups + 3) / 4, 2048);
    hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(256), 0, s, N, E, n, n_dev, offset, pos, dir,
                       target, tb_);
    const uint32_t wb = std::min<uint32_t>(tb_.blocks, (n + 31) / 32);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(wb), dim3(256), 0, s, tb_, n, n_dev, offset);
    ++step_;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, static_cast<float>(step_));
    const float c2 = 1.0f - std::pow(b2, static_cast<float>(step_));
    float* gdev = nullptr;
    if (grad_host && dalloc(owned_, &gdev, kParams)) return MT_E_NOMEM;
    hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + 63) / 64), dim3(256), 0, s, master_, adam_m_,
                       adam_v_, tb_.partial, wb, lr_, b1, b2, 1e-8f, c1, c2, gdev, n, n_dev, offset,
                       trained_);
    int rc = pack(s);
    if (rc) return rc;
    if (hipGetLastError() != hipSuccess) return fail(MT_E_HIP, "cache training launch failed");
    if (loss_host || grad_host) {
        if (hipStreamSynchronize(s) != hipSuccess) return fail(MT_E_HIP, "cache training failed");
        if (loss_host && hipMemcpy(loss_host, tb_.loss, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            return fail(MT_E_HIP, "loss readback failed");
        if (grad_host) {
            const hipError_t e = hipMemcpy(grad_host, gdev, sizeof(float) * kParams, hipMemcpyDeviceToHost);
            owned_.pop_back();
            (void)hipFree(gdev);
            if (e != hipSuccess) return fail(MT_E_HIP, "gradient readback failed");
        }
    }
    return MT_OK;
}

}  // namespace mt
