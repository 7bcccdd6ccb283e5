// mpm_surrogate_net.hpp — the grid U-Net of the learned surrogate (host side).
//
// Four levels, 32-64-128-128 channels, 3x3 convolutions with GELU, 2x2 average
// pooling down, nearest 2x upsampling and channel concatenation up:
//
//   e0: x(32)            -> a0 (32)   H x W
//   e1: pool(a0) (32)    -> a1 (64)   H/2
//   e2: pool(a1) (64)    -> a2 (128)  H/4
//   e3: pool(a2) (128)   -> a3 (128)  H/8
//   d2: [up(a3) | a2]    -> b2 (128)  H/4
//   d1: [up(b2) | a1]    -> b1 (64)   H/2
//   d0: [up(b1) | a0]    -> b0 (32)   H
//   head: 1x1, 32 -> 2 (FP32), then the per-image momentum projection.
//
// Parameters (FP32, one buffer): per conv layer W[9][cout][cin] then b[cout], in
// the order e0 e1 e2 e3 d2 d1 d0, then the head wh[2][32], bh[2].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "mpm_unet.hpp"

namespace mpm {

struct UNetLayer {
    int c1, c2, cout, level;
    size_t w_off, b_off;
};

class UNet {
public:
    static constexpr int IN_CH = 32;
    static constexpr int NLAYERS = 7;

    // h, w: image size (multiples of 8); max_batch images per call
    int init(int h, int w, int max_batch, uint64_t seed, hipStream_t st, std::string* err);
    void release();
    ~UNet() { release(); }

    int64_t num_params() const { return (int64_t)n_params_; }
    float* params() { return params_; }
    float* grads() { return grads_; }
    int h() const { return h_; }
    int w() const { return w_; }
    int max_batch() const { return max_batch_; }

    // Refresh the BF16 weight copies after the FP32 parameters changed.
    hipError_t sync_weights();
    // Forward pass: x [nb][h][w][32] BF16, wm [nb][h*w] projection weights; the
    // projected output (2 channels, FP32) is left in out().  On `stream` when given,
    // else on the network's own stream.
    hipError_t forward(int nb, const unet::bf16* x, const float* wm, hipStream_t stream = nullptr);
    // Loss of the projected output against target [nb][h*w][2] into *loss (device
    // scalar, accumulated); with backward, the parameter gradient into grads() (zeroed first).
    hipError_t loss_and_backward(int nb, const unet::bf16* x, const float* wm, const float* target, float* loss,
                                 bool backward);
    hipError_t adam(float lr, float beta1 = 0.9f, float beta2 = 0.999f, float eps = 1e-8f);
    float* out() { return out_; }
    // forward convolution kernel: 0 by layer shape, 1 register-streaming, 2 LDS-tiled
    void set_conv_kernel(int k) { conv_kernel_ = k; }

private:
    int h_ = 0, w_ = 0, max_batch_ = 0;
    int conv_kernel_ = 0;
    hipStream_t st_ = nullptr;
    UNetLayer layers_[NLAYERS] = {};
    size_t n_params_ = 0, head_w_off_ = 0, head_b_off_ = 0;
    float* params_ = nullptr;
    float* grads_ = nullptr;
    float* adam_m_ = nullptr;
    float* adam_v_ = nullptr;
    int adam_t_ = 0;
    unet::bf16* wf_[NLAYERS] = {};
    unet::bf16* wb_[NLAYERS] = {};
    // activations (forward) and their gradients (backward), BF16 NHWC
    unet::bf16 *z0_, *a0_, *z1_, *a1_, *z2_, *a2_, *z3_, *a3_;
    unet::bf16 *zd2_, *b2_, *zd1_, *b1_, *zd0_, *b0_;
    unet::bf16 *g_b0_, *g_u0_, *g_a0_, *g_b1_, *g_u1_, *g_a1_, *g_b2_, *g_u2_, *g_a2_, *g_a3_, *g_p2_, *g_p1_,
        *g_p0_;
    float* out_ = nullptr;
    float* dproj_ = nullptr;
    float* part_ = nullptr;    // [max_batch][64][2] partial projection sums
    float* part2_ = nullptr;   // [max_batch][64][2] partial gradient sums
    float* hpart_ = nullptr;   // [max_batch][64][66] head gradient partials
    std::vector<void*> allocs_;
    void* alloc(size_t bytes, hipError_t* e);
};

}  // namespace mpm
