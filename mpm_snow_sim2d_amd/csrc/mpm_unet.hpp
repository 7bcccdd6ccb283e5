// mpm_unet.hpp — kernels of the grid U-Net surrogate (mpm_unet.hip).
//
// Tensors are NHWC: [image][row][column][channel], channels innermost, BF16 for
// activations and their gradients, FP32 for parameters, their gradients and the
// optimiser state.  3x3 convolutions are "same"-padded implicit GEMMs on the
// gfx950 BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation); every
// channel count is a multiple of 32.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mpm {
namespace unet {

typedef __bf16 bf16;

// Forward 3x3 convolution over the channel concatenation [x1 | x2] (x2 may be
// empty, c2 = 0).  wt: [9][cout][c1+c2] BF16 (tap = 3*(dy+1) + (dx+1)).
// Forward epilogue: z = acc + bias, a = GELU(z) (either pointer may be null).
// Split epilogue (backward data): output channels [0, split) go to o1 (row stride
// split), [split, cout) to o2 (row stride cout - split), each stored or added.
struct ConvArgs {
    const bf16* x1;
    const bf16* x2;
    int c1, c2;
    int mode1;          // x1 stored at this resolution (0), half (1: upsampled), double (2: pooled)
    int nb, h, w;
    const bf16* wt;
    const float* bias;
    int cout;
    bf16* z;
    bf16* a;
    bf16* o1;
    bf16* o2;
    int split;
    int add1, add2;
};
// kernel: 0 chooses by layer shape, 1 the register-streaming kernel, 2 the LDS-tiled one
hipError_t conv3x3_forward(const ConvArgs& a, hipStream_t st, int kernel = 0);
hipError_t conv3x3_split(const ConvArgs& a, hipStream_t st);

// dW[tap][co][ci] += sum over pixels of dz[p][co] * x[p shifted by tap][ci]  (FP32, atomics);
// x = [x1 | x2] as the forward convolution saw it (mode1 as in ConvArgs)
hipError_t conv3x3_wgrad(const bf16* x1, const bf16* x2, int c1, int c2, int mode1, const bf16* dz, int cout,
                         int nb, int h, int w, float* dw, hipStream_t st);
// db[co] += sum over pixels of dz[p][co]
hipError_t bias_grad(const bf16* dz, int npix, int cout, float* db, hipStream_t st);

// BF16 copies of one layer's FP32 weights [9][cout][cin]: forward layout and the
// tap-flipped, transposed layout [9][cin][cout] of the backward-data convolution.
hipError_t cast_weights(const float* w, int cout, int cin, bf16* wf, bf16* wb, hipStream_t st);

// 2x2 average pooling and nearest 2x upsampling, with their adjoints.
hipError_t avgpool2(const bf16* x, int nb, int h, int w, int c, bf16* y, hipStream_t st);
hipError_t avgpool2_bwd(const bf16* dy, int nb, int h, int w, int c, bf16* dx, int add, hipStream_t st);
hipError_t upsample2(const bf16* x, int nb, int h, int w, int c, bf16* y, hipStream_t st);
hipError_t upsample2_bwd(const bf16* dy, int nb, int h, int w, int c, bf16* dx, hipStream_t st);

// dz = da * GELU'(z), elementwise over n values.
hipError_t gelu_bwd(const bf16* da, const bf16* z, int64_t n, bf16* dz, hipStream_t st);

// Head: out[p][o] = bh[o] + sum_c a[p][c] wh[o][c] (o < 2, c < 32, FP32 out) and the
// per-block partial sums part [nb][64][2] of wm_i out_i (wm: per-node weights m_i/M
// summing to one per image).
hipError_t head_forward(const bf16* a, int nb, int hw, const float* wh, const float* bh, const float* wm,
                        float* out, float* part, hipStream_t st);
// Momentum projection in place, out' = out - sum_i wm_i out_i per image (the sum of
// the image's partials); with a target, the loss
// L = (1/nb) sum_images sum_i wm_i |out'_i - target_i|^2 added into *loss and
// dL/dout' written to dproj with its per-block partial sums in part2 [nb][64][2].
hipError_t head_project_loss(float* out, const float* part, const float* target, const float* wm, int nb,
                             int hw, float* loss, float* dproj, float* part2, hipStream_t st);
// Backward of projection and head: da (BF16, [p][32]), dwh/dbh accumulated (FP32).
// hpart: scratch of nb * 64 rows of 66 floats (per-workgroup gradient partials).
hipError_t head_backward(const bf16* a, const float* dproj, const float* part2, const float* wm, int nb,
                         int hw, const float* wh, bf16* da, float* dwh, float* dbh, float* hpart, hipStream_t st);

// Adam (FP32): m, v moments; lr_t includes the bias corrections.
hipError_t adam_step(float* w, const float* g, float* m, float* v, int64_t n, float lr_t, float beta1,
                     float beta2, float eps_t, hipStream_t st);
hipError_t fill_f32(float* p, int64_t n, float value, hipStream_t st);
hipError_t fill_bf16(bf16* p, int64_t n, hipStream_t st);

}  // namespace unet
}  // namespace mpm
