// mpm_surrogate_net.hip — host orchestration of the grid U-Net (layer table,
// buffers, forward, backward, Adam) over the kernels of mpm_unet.hip.
#include "mpm_surrogate_net.hpp"

#include <cmath>
#include <random>

namespace mpm {

using unet::bf16;

#define NET_TRY(expr)                          \
    do {                                       \
        hipError_t e_ = (expr);                \
        if (e_ != hipSuccess) return e_;       \
    } while (0)

void* UNet::alloc(size_t bytes, hipError_t* e)
{
    void* p = nullptr;
    if (*e != hipSuccess) return nullptr;
    *e = hipMalloc(&p, bytes < 256 ? 256 : bytes);
    if (*e == hipSuccess) allocs_.push_back(p);
    return p;
}

void UNet::release()
{
    for (void* p : allocs_) (void)hipFree(p);
    allocs_.clear();
    params_ = grads_ = nullptr;
}

int UNet::init(int h, int w, int max_batch, uint64_t seed, hipStream_t st, std::string* err)
{
    if (h < 8 || w < 8 || h % 8 || w % 8) {
        *err = "image size must be a positive multiple of 8";
        return -1;
    }
    if (max_batch < 1) {
        *err = "max_batch must be >= 1";
        return -1;
    }
    h_ = h;
    w_ = w;
    max_batch_ = max_batch;
    st_ = st;
    const int spec[NLAYERS][4] = {
        {IN_CH, 0, 32, 0}, {32, 0, 64, 1}, {64, 0, 128, 2}, {128, 0, 128, 3},
        {128, 128, 128, 2}, {128, 64, 64, 1}, {64, 32, 32, 0}};
    size_t off = 0;
    for (int l = 0; l < NLAYERS; ++l) {
        UNetLayer& L = layers_[l];
        L.c1 = spec[l][0];
        L.c2 = spec[l][1];
        L.cout = spec[l][2];
        L.level = spec[l][3];
        L.w_off = off;
        off += (size_t)9 * L.cout * (L.c1 + L.c2);
        L.b_off = off;
        off += L.cout;
    }
    head_w_off_ = off;
    off += 2 * 32;
    head_b_off_ = off;
    off += 2;
    n_params_ = off;

    hipError_t e = hipSuccess;
    params_ = (float*)alloc(n_params_ * 4, &e);
    grads_ = (float*)alloc(n_params_ * 4, &e);
    adam_m_ = (float*)alloc(n_params_ * 4, &e);
    adam_v_ = (float*)alloc(n_params_ * 4, &e);
    for (int l = 0; l < NLAYERS; ++l) {
        const size_t nw = (size_t)9 * layers_[l].cout * (layers_[l].c1 + layers_[l].c2);
        wf_[l] = (bf16*)alloc(nw * 2, &e);
        wb_[l] = (bf16*)alloc(nw * 2, &e);
    }
    const size_t B = max_batch, P0 = B * h * w, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    auto act = [&](size_t pix, int c) { return (bf16*)alloc(pix * c * 2, &e); };
    z0_ = act(P0, 32); a0_ = act(P0, 32);
    z1_ = act(P1, 64); a1_ = act(P1, 64);
    z2_ = act(P2, 128); a2_ = act(P2, 128);
    z3_ = act(P3, 128); a3_ = act(P3, 128);
    zd2_ = act(P2, 128); b2_ = act(P2, 128);
    zd1_ = act(P1, 64); b1_ = act(P1, 64);
    zd0_ = act(P0, 32); b0_ = act(P0, 32);
    g_b0_ = act(P0, 32); g_u0_ = act(P0, 64); g_a0_ = act(P0, 32);
    g_b1_ = act(P1, 64); g_u1_ = act(P1, 128); g_a1_ = act(P1, 64);
    g_b2_ = act(P2, 128); g_u2_ = act(P2, 128); g_a2_ = act(P2, 128);
    g_a3_ = act(P3, 128); g_p2_ = act(P3, 128); g_p1_ = act(P2, 64); g_p0_ = act(P1, 32);
    out_ = (float*)alloc(P0 * 2 * 4, &e);
    dproj_ = (float*)alloc(P0 * 2 * 4, &e);
    part_ = (float*)alloc(B * 64 * 2 * 4, &e);
    part2_ = (float*)alloc(B * 64 * 2 * 4, &e);
    hpart_ = (float*)alloc(B * 64 * 66 * 4, &e);
    if (e != hipSuccess) {
        *err = std::string("U-Net allocation: ") + hipGetErrorString(e);
        release();
        return -2;
    }

    // He-normal convolution weights, zero biases, a small random head
    std::vector<float> host(n_params_, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < NLAYERS; ++l) {
        const UNetLayer& L = layers_[l];
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * (L.c1 + L.c2))));
        const size_t nw = (size_t)9 * L.cout * (L.c1 + L.c2);
        for (size_t i = 0; i < nw; ++i) host[L.w_off + i] = nd(rng);
    }
    std::normal_distribution<float> nh(0.f, 1e-2f);
    for (int i = 0; i < 64; ++i) host[head_w_off_ + i] = nh(rng);
    e = hipMemcpyAsync(params_, host.data(), n_params_ * 4, hipMemcpyHostToDevice, st_);
    if (e == hipSuccess) e = unet::fill_f32(adam_m_, (int64_t)n_params_, 0.f, st_);
    if (e == hipSuccess) e = unet::fill_f32(adam_v_, (int64_t)n_params_, 0.f, st_);
    if (e == hipSuccess) e = sync_weights();
    if (e == hipSuccess) e = hipStreamSynchronize(st_);
    if (e != hipSuccess) {
        *err = std::string("U-Net initialisation: ") + hipGetErrorString(e);
        release();
        return -2;
    }
    adam_t_ = 0;
    return 0;
}

hipError_t UNet::sync_weights()
{
    for (int l = 0; l < NLAYERS; ++l) {
        const UNetLayer& L = layers_[l];
        NET_TRY(unet::cast_weights(params_ + L.w_off, L.cout, L.c1 + L.c2, wf_[l], wb_[l], st_));
    }
    return hipSuccess;
}

hipError_t UNet::forward(int nb, const bf16* x, const float* wm, hipStream_t stream)
{
    const int H = h_, W = w_;
    hipStream_t st_ = stream ? stream : this->st_;
    // x1 of e1..e3 is the pooled output of the level above, x1 of d2..d0 the
    // upsampled output of the level below: the convolutions read them in place
    auto conv = [&](int l, const bf16* x1, int mode1, const bf16* x2, int hh, int ww, bf16* z, bf16* a) {
        const UNetLayer& L = layers_[l];
        unet::ConvArgs c{};
        c.x1 = x1; c.x2 = x2; c.c1 = L.c1; c.c2 = L.c2; c.mode1 = mode1;
        c.nb = nb; c.h = hh; c.w = ww;
        c.wt = wf_[l]; c.bias = params_ + L.b_off; c.cout = L.cout;
        c.z = z; c.a = a;
        return unet::conv3x3_forward(c, st_, conv_kernel_);
    };
    NET_TRY(conv(0, x, 0, nullptr, H, W, z0_, a0_));
    NET_TRY(conv(1, a0_, 2, nullptr, H / 2, W / 2, z1_, a1_));
    NET_TRY(conv(2, a1_, 2, nullptr, H / 4, W / 4, z2_, a2_));
    NET_TRY(conv(3, a2_, 2, nullptr, H / 8, W / 8, z3_, a3_));
    NET_TRY(conv(4, a3_, 1, a2_, H / 4, W / 4, zd2_, b2_));
    NET_TRY(conv(5, b2_, 1, a1_, H / 2, W / 2, zd1_, b1_));
    NET_TRY(conv(6, b1_, 1, a0_, H, W, zd0_, b0_));
    NET_TRY(unet::head_forward(b0_, nb, H * W, params_ + head_w_off_, params_ + head_b_off_, wm, out_, part_, st_));
    return unet::head_project_loss(out_, part_, nullptr, wm, nb, H * W, nullptr, nullptr, nullptr, st_);
}

hipError_t UNet::loss_and_backward(int nb, const bf16* x, const float* wm, const float* target, float* loss,
                                   bool backward)
{
    const int H = h_, W = w_;
    NET_TRY(forward(nb, x, wm));
    // forward() left the projected output: the loss pass subtracts nothing more
    NET_TRY(unet::head_project_loss(out_, nullptr, target, wm, nb, H * W, loss, backward ? dproj_ : nullptr,
                                    part2_, st_));
    if (!backward) return hipSuccess;

    NET_TRY(unet::fill_f32(grads_, (int64_t)n_params_, 0.f, st_));
    NET_TRY(unet::head_backward(b0_, dproj_, part2_, wm, nb, H * W, params_ + head_w_off_, g_b0_,
                                grads_ + head_w_off_, grads_ + head_b_off_, hpart_, st_));

    auto wgrad = [&](int l, const bf16* x1, int mode1, const bf16* x2, const bf16* dz, int hh,
                     int ww) -> hipError_t {
        const UNetLayer& L = layers_[l];
        NET_TRY(unet::conv3x3_wgrad(x1, x2, L.c1, L.c2, mode1, dz, L.cout, nb, hh, ww, grads_ + L.w_off, st_));
        return unet::bias_grad(dz, nb * hh * ww, L.cout, grads_ + L.b_off, st_);
    };
    auto dgrad = [&](int l, const bf16* dz, int hh, int ww, bf16* o1, bf16* o2, int add1, int add2) {
        const UNetLayer& L = layers_[l];
        unet::ConvArgs c{};
        c.x1 = dz; c.x2 = nullptr; c.c1 = L.cout; c.c2 = 0;
        c.nb = nb; c.h = hh; c.w = ww;
        c.wt = wb_[l]; c.bias = nullptr; c.cout = L.c1 + L.c2;
        c.o1 = o1; c.o2 = o2; c.split = L.c1; c.add1 = add1; c.add2 = add2;
        return unet::conv3x3_split(c, st_);
    };
    const int64_t P0 = (int64_t)nb * H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;

    // d0: [u0 | a0] -> b0
    NET_TRY(unet::gelu_bwd(g_b0_, zd0_, P0 * 32, g_b0_, st_));
    NET_TRY(wgrad(6, b1_, 1, a0_, g_b0_, H, W));
    NET_TRY(dgrad(6, g_b0_, H, W, g_u0_, g_a0_, 0, 0));
    NET_TRY(unet::upsample2_bwd(g_u0_, nb, H / 2, W / 2, 64, g_b1_, st_));
    // d1: [u1 | a1] -> b1
    NET_TRY(unet::gelu_bwd(g_b1_, zd1_, P1 * 64, g_b1_, st_));
    NET_TRY(wgrad(5, b2_, 1, a1_, g_b1_, H / 2, W / 2));
    NET_TRY(dgrad(5, g_b1_, H / 2, W / 2, g_u1_, g_a1_, 0, 0));
    NET_TRY(unet::upsample2_bwd(g_u1_, nb, H / 4, W / 4, 128, g_b2_, st_));
    // d2: [u2 | a2] -> b2
    NET_TRY(unet::gelu_bwd(g_b2_, zd2_, P2 * 128, g_b2_, st_));
    NET_TRY(wgrad(4, a3_, 1, a2_, g_b2_, H / 4, W / 4));
    NET_TRY(dgrad(4, g_b2_, H / 4, W / 4, g_u2_, g_a2_, 0, 0));
    NET_TRY(unet::upsample2_bwd(g_u2_, nb, H / 8, W / 8, 128, g_a3_, st_));
    // e3: p2 -> a3
    NET_TRY(unet::gelu_bwd(g_a3_, z3_, P3 * 128, g_a3_, st_));
    NET_TRY(wgrad(3, a2_, 2, nullptr, g_a3_, H / 8, W / 8));
    NET_TRY(dgrad(3, g_a3_, H / 8, W / 8, g_p2_, nullptr, 0, 0));
    NET_TRY(unet::avgpool2_bwd(g_p2_, nb, H / 4, W / 4, 128, g_a2_, 1, st_));
    // e2: p1 -> a2
    NET_TRY(unet::gelu_bwd(g_a2_, z2_, P2 * 128, g_a2_, st_));
    NET_TRY(wgrad(2, a1_, 2, nullptr, g_a2_, H / 4, W / 4));
    NET_TRY(dgrad(2, g_a2_, H / 4, W / 4, g_p1_, nullptr, 0, 0));
    NET_TRY(unet::avgpool2_bwd(g_p1_, nb, H / 2, W / 2, 64, g_a1_, 1, st_));
    // e1: p0 -> a1
    NET_TRY(unet::gelu_bwd(g_a1_, z1_, P1 * 64, g_a1_, st_));
    NET_TRY(wgrad(1, a0_, 2, nullptr, g_a1_, H / 2, W / 2));
    NET_TRY(dgrad(1, g_a1_, H / 2, W / 2, g_p0_, nullptr, 0, 0));
    NET_TRY(unet::avgpool2_bwd(g_p0_, nb, H, W, 32, g_a0_, 1, st_));
    // e0: x -> a0 (no input gradient)
    NET_TRY(unet::gelu_bwd(g_a0_, z0_, P0 * 32, g_a0_, st_));
    NET_TRY(wgrad(0, x, 0, nullptr, g_a0_, H, W));
    return hipSuccess;
}

hipError_t UNet::adam(float lr, float beta1, float beta2, float eps)
{
    ++adam_t_;
    const double b1t = 1.0 - std::pow((double)beta1, adam_t_), b2t = 1.0 - std::pow((double)beta2, adam_t_);
    const float lr_t = (float)(lr * std::sqrt(b2t) / b1t);
    const float eps_t = (float)(eps * std::sqrt(b2t));
    NET_TRY(unet::adam_step(params_, grads_, adam_m_, adam_v_, (int64_t)n_params_, lr_t, beta1, beta2, eps_t, st_));
    return sync_weights();
}

}  // namespace mpm
