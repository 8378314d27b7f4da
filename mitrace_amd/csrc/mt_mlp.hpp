// mt_mlp.hpp -- the radiance-cache MLP (internal): 32 encoded inputs -> 4 hidden layers of 64
// (ReLU) -> 3 outputs, BF16 operands on MFMA with FP32 accumulation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace mt {

namespace mlp {

constexpr int kLayers = 5;
constexpr int kIn[kLayers] = {32, 64, 64, 64, 64};
constexpr int kOut[kLayers] = {64, 64, 64, 64, 3};
// fp32 master parameters, natural layout: W_l[out][in] then b_l[out] per layer
constexpr int kWOff[kLayers] = {0, 2112, 6272, 10432, 14592};
constexpr int kBOff[kLayers] = {2048, 6208, 10368, 14528, 14784};
constexpr int kParams = 14787;
// bf16 operand packs: forward A = W^T fragments [layer][out tile][in chunk][lane][8],
// backward A = W fragments [layer][in tile][out chunk][lane][8] (layers 2..5)
constexpr int kFwdTiles[kLayers] = {2, 2, 2, 2, 1};
constexpr int kFwdChunks[kLayers] = {2, 4, 4, 4, 4};
constexpr int kFwdOff[kLayers] = {0, 2048, 6144, 10240, 14336};  // bf16 elements
constexpr int kFwdSize = 16384;
constexpr int kBwdTiles = 2;
constexpr int kBwdChunks[kLayers] = {0, 4, 4, 4, 1};
constexpr int kBwdOff[kLayers] = {0, 0, 4096, 8192, 12288};
constexpr int kBwdSize = 13312;
// per-sample activations kept by the training pass (bf16, [group][feature][32 samples]):
// H0 (32, the encoding) H1..H4 (64 each), and dZ1..dZ4 (64 each) dZ5 (3, stored as 4)
constexpr int kHWidth = 32 + 4 * 64;      // 288
constexpr int kDZWidth = 4 * 64 + 4;      // 260
constexpr int kHOff[kLayers] = {0, 32, 96, 160, 224};      // input of layer l
constexpr int kDZOff[kLayers] = {0, 64, 128, 192, 256};    // output gradient of layer l

// feature held in k-slot t (0..7) of lane group g (0..1) of K-chunk c: the order in which a
// 32x32 MFMA result (lane = sample, rows q*8 + g*4 + j) feeds the next product directly
inline __host__ __device__ int perm(int c, int g, int t) {
    return (c >> 1) * 32 + ((c & 1) * 2 + (t >> 2)) * 8 + g * 4 + (t & 3);
}

}  // namespace mlp

struct CacheTrainBuffers {
    uint16_t* H;          // [n / 32][kHWidth][32] bf16: per group of 32 samples, feature-major
    uint16_t* DZ;         // [n / 32][kDZWidth][32] bf16
    float* partial;       // [blocks][kParams]
    float* loss;          // [blocks]
    uint32_t capacity;    // samples
    uint32_t blocks;      // weight-gradient blocks
    int32_t l2;           // loss: 0 relative L2 (Huber-clipped), 1 L2
};

class Mlp {
   public:
    ~Mlp();
    int init(uint32_t seed, uint32_t train_capacity, float lr);
    // device pointers: pos/dir float4 (xyz used), out float4; bounds = scene box for the encoding.
    // n_dev (optional): the count is min(n, *n_dev), read by the kernels (no host sync)
    int infer(uint32_t n, const float4* pos, const float4* dir, float4* out, hipStream_t s,
              const uint32_t* n_dev = nullptr);
    // one Adam step on n samples with targets float4 (xyz); *loss_host (optional) receives the
    // mean loss (synchronises the stream).  n_dev (optional): the step takes samples
    // [offset, offset + min(n, *n_dev - offset)) and is skipped when that is empty
    int train(uint32_t n, const float4* pos, const float4* dir, const float4* target, hipStream_t s,
              float* loss_host, float* grad_host, const uint32_t* n_dev = nullptr,
              uint32_t offset = 0);
    uint64_t trained_samples();
    int get_master(float* out);
    int set_master(const float* in);
    void set_bounds(const float lo[3], const float hi[3]);
    void set_lr(float lr) { lr_ = lr; }
    void set_loss(uint32_t l2) { tb_.l2 = l2 ? 1 : 0; }
    uint64_t steps() const { return step_; }
    // the render passes' output scale: {scale, set}; the cache learns targets divided by the
    // scale and its predictions are multiplied by it; {1, 0} until the first training chunk of
    // a cached render sets it to that chunk's mean target
    float* scale() const { return scale_; }

   private:
    int pack(hipStream_t s);
    std::vector<void*> owned_;
    float* master_ = nullptr;
    float* adam_m_ = nullptr;
    float* adam_v_ = nullptr;
    uint16_t* fwd_ = nullptr;
    uint16_t* bwd_ = nullptr;
    CacheTrainBuffers tb_{};
    float lo_[3] = {0, 0, 0}, inv_ext_[3] = {1, 1, 1};
    float lr_ = 1e-2f;
    uint64_t step_ = 0;
    unsigned long long* trained_ = nullptr;  // samples of all non-empty steps
    float* scale_ = nullptr;
};

}  // namespace mt
