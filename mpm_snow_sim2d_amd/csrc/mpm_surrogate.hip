// mpm_surrogate.hip — the learned grid surrogate of the 2D solver: grid features
// captured from a solver's particles, a device-resident training set, training
// steps of the U-Net (mpm_surrogate_net.hip), and the macro step that advances a
// scene by k substeps with the predicted grid velocity and displacement, with the
// plain solver as fallback.
//
// Image: grid nodes (i, j) with 0 <= i < w, 0 <= j < h of a scene whose grid has
// w = WIDTH cells (nodes 0..WIDTH); the last node row and column lie inside the
// wall band and are left to the plain solver.
//
// Per-node raw features of a sample (FP32 [9][h*w]): mass m_i = sum_p w_ip m_p,
// velocity v_i = sum_p w_ip m_p v_p / m_i, mean J_P, mean log singular value of F_E
// (= log|det F_E| / 2), and two Lagrangian target fields, velocity and
// displacement k substeps later, y_p = v_p(t + k dt) and y_p = x_p(t + k dt) - x_p(t),
// at the weights w_ip(t) of the positions at capture time.  Each target field is
// the mass-weighted least-squares fit of the particle values,
//   min_t sum_p m_p |y_p - sum_i w_ip t_i|^2,
// approximated by refine_iters Jacobi-preconditioned Richardson steps from the
// transfer t_i = sum_p w_ip m_p y_p / m_i (the preconditioner, the diagonal of the
// normal equations' row sums, is m_i):
//   t_i <- t_i + sum_p w_ip m_p (y_p - sum_j w_jp t_j) / m_i.
// Gathered back to the particles, the fitted fields reproduce the particles'
// values instead of smoothing them twice, and each step keeps sum_i m_i t_i
// (refine_iters = 0, the default, keeps the transfers themselves).
// Network input (BF16, 32 channels, normalised): [m/m_ref, v/v_ref, (J-1)*10,
// logsig*10, occupancy, mode, 0...]; the network predicts the velocity field and,
// when the macro step uses the displacement (macro flags), also the displacement
// field, selected by the constant mode plane (0 velocity, 1 displacement) of a
// second image in the same batch, as the increment over ballistic motion,
//   v_pred,i = v_i + g k dt e_y + d_ref * out'_i                        (mode 0)
//   u_pred,i = k dt v_i + g dt^2 k (k + 1) / 2 e_y + d_ref k dt * out'_i (mode 1)
// (the ballistic terms are the solver's own free-fall recurrence), where out' is
// projected so that sum_i m_i out'_i = 0 (the prediction changes the momentum of
// the scene by exactly the gravity impulse).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpm/mpm.h"
#include "mpm_device.hpp"
#include "mpm_solver_impl.hpp"
#include "mpm_surrogate_net.hpp"
#include "mpm_unet.hpp"

using namespace mpm;
using unet::bf16;

namespace {

constexpr int NCH = 9;           // raw sample channels
constexpr int MACRO_X = MPM_MACRO_DISPLACEMENT_X;   // macro step: positions from the displacement
constexpr int MACRO_F = MPM_MACRO_DISPLACEMENT_F;   // macro step: F_E from the displacement gradient
constexpr float V_REF = 3.0f;    // velocity scale of the input (m/s)

struct SlotInfo {
    float particle_mass;   // sum_p m_p of the scene
    float grid_mass;       // sum_i m_i over the image
    float band_mass;       // particle mass within the wall band
    float T;               // k * dt
    float gravity;
    float m_ref;           // RHO * dx^2
    float ug;              // g dt^2 k (k + 1) / 2: free-fall displacement in k substeps
    float pad;
};

int sfail(int code, const std::string& m) { return mpm::detail::set_error(code, m); }

#define S_TRY(expr)                                                                           \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return sfail(MPM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Sum of v over the workgroup (256 threads), returned to thread 0.
__device__ inline float block_sum(float v, float* sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
}

// Scatter of the features (mode 0) or of the target momentum at the saved
// positions (mode 1) of the particles in slots [first, first + count).  The
// particles of a workgroup are spatially sorted (the solver bins them), so their
// stencils cover a small box of nodes: the workgroup accumulates that box in LDS
// and adds it to the image once per node; a workgroup whose box does not fit
// (after particles have mixed) adds its contributions directly.
//
// The LDS box accumulates in exact fixed point (the LDS integer add retires tens of
// times faster than the LDS float add on gfx950): channel c of the workgroup's
// particles is scaled by a power of two 2^e_c with max_p |val_c| 2^e_c <= 2^45, and
// each FP32 contribution x = rint(w * val * 2^e_c) becomes an int64 through the exact
// split hi = floor(x / 2^22), lo = x - hi 2^22 (two single-instruction FP32->int32
// conversions; there is no FP32->int64 instruction) and is added to an int64 word.
// A node's weights from one particle are <= 9/16 and at most 256 particles add to
// it, so a sum stays below 2^53: exact in FP64 when converted back, and independent
// of the order of the additions; a
// contribution loses nothing above 2^-22 of the workgroup's largest, and at most
// half a unit of 2^-45 of it below.
constexpr int CAP_NODES = 1024;   // LDS box capacity in nodes (5 channels of int64, 40 KB)

__device__ inline float capture_scale(float bound)
{
    if (!(bound > 0.f) || !(bound <= 3.0e38f)) return 1.f;   // empty, or not finite: no scaling
    int k;
    frexpf(bound, &k);                 // bound < 2^k
    return ldexpf(1.f, min(max(45 - k, -126), 126));
}

__device__ inline float wave_max_f(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__device__ inline int wave_min_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

__device__ inline int wave_max_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

template <int MODE>
__global__ void __launch_bounds__(256) k_capture(DevParticles P, int first, int count, Phys ph, int h, int w,
                                                float band, float* __restrict__ sample, float* __restrict__ xs,
                                                float2* __restrict__ part)
{
    constexpr int NCAP = MODE == 0 ? 5 : 4;   // channels scattered
    __shared__ float sh[4];
    __shared__ float s_bound[NCAP][4];
    __shared__ int s_box[4][4];
    __shared__ unsigned long long tile[NCAP * CAP_NODES];
    // lane l of wave w takes the workgroup's particle 4l + w: the particles are sorted
    // by cell, so the 64 lanes of a wave fall in different cells and their LDS adds
    // rarely meet on one address (same-address adds serialise)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x * 256 + ((lane << 2) | wave);
    const int n = P.n;
    const size_t hw = (size_t)h * w;
    float pm = 0.f, inband = 0.f;
    float val[NCAP];
#pragma unroll
    for (int c = 0; c < NCAP; ++c) val[c] = 0.f;
    Stencil<2, 2> st;
    const bool live = k < count;
    if (live) {
        const int p = first + k;
        const uint32_t id = P.pid[p];
        float x0 = P.x[p], x1 = P.x[n + p];
        const float c0 = x0, c1 = x1;   // positions now
        if constexpr (MODE == 0) {
            xs[2 * (size_t)id] = x0;
            xs[2 * (size_t)id + 1] = x1;
        } else {
            x0 = xs[2 * (size_t)id];
            x1 = xs[2 * (size_t)id + 1];
        }
        const float m = P.mass[p];
        pm = m;
        const float xg[2] = {x0 * ph.inv_dx, x1 * ph.inv_dx};
        if (xg[0] < band || xg[0] > (float)ph.W - band || xg[1] < band || xg[1] > (float)ph.H - band) inband = m;
        st.init(xg);
        const float v0 = P.v[p], v1 = P.v[n + p];
        if constexpr (MODE == 0) {
            const float F[4] = {P.F[p], P.F[n + p], P.F[2 * n + p], P.F[3 * n + p]};
            val[0] = m;
            val[1] = m * v0;
            val[2] = m * v1;
            val[3] = m * P.Jp[p];
            val[4] = m * 0.5f * logf(fmaxf(fabsf(det<2>(F)), 1e-30f));
        } else {
            val[0] = m * v0;
            val[1] = m * v1;
            val[2] = m * (c0 - x0);
            val[3] = m * (c1 - x1);
        }
    }
    // the workgroup's largest |value| per channel (-> fixed-point scales) and the box
    // of its stencils, reduced per wave, then across the four waves
#pragma unroll
    for (int c = 0; c < NCAP; ++c) {
        float b = fabsf(val[c]);
        if (!(b <= 3.0e38f)) b = INFINITY;   // NaN counts as unbounded (fmaxf would drop it)
        b = wave_max_f(b);
        if (lane == 0) s_bound[c][wave] = b;
    }
    {
        const int b0 = wave_min_i(live ? st.base[0] : INT_MAX), b1 = wave_min_i(live ? st.base[1] : INT_MAX);
        const int b2 = wave_max_i(live ? st.base[0] : INT_MIN), b3 = wave_max_i(live ? st.base[1] : INT_MIN);
        if (lane == 0) {
            s_box[0][wave] = b0;
            s_box[1][wave] = b1;
            s_box[2][wave] = b2;
            s_box[3][wave] = b3;
        }
    }
    __syncthreads();
    // a workgroup holding a value that is not finite adds in floating point
    // (directly), so the non-finite value reaches the image
    float scale[NCAP];
    double inv_scale[NCAP];
    bool finite = true;
#pragma unroll
    for (int c = 0; c < NCAP; ++c) {
        const float bm = fmaxf(fmaxf(s_bound[c][0], s_bound[c][1]), fmaxf(s_bound[c][2], s_bound[c][3]));
        finite &= bm <= 3.0e38f;
        scale[c] = capture_scale(bm);
        inv_scale[c] = 1.0 / (double)scale[c];
    }
    const int bx0 = min(min(s_box[0][0], s_box[0][1]), min(s_box[0][2], s_box[0][3]));
    const int by0 = min(min(s_box[1][0], s_box[1][1]), min(s_box[1][2], s_box[1][3]));
    const int bx1 = max(max(s_box[2][0], s_box[2][1]), max(s_box[2][2], s_box[2][3]));
    const int by1 = max(max(s_box[3][0], s_box[3][1]), max(s_box[3][2], s_box[3][3]));
    const bool any_live = bx1 >= bx0;
    // box extents in 64 bits: bases of non-finite positions saturate at the int range
    const long long bwl = any_live ? (long long)bx1 - bx0 + 3 : 0, bhl = any_live ? (long long)by1 - by0 + 3 : 0;
    const size_t ch0 = MODE == 0 ? 0 : 5;
    if (any_live && finite && bwl <= CAP_NODES && bhl <= CAP_NODES && bwl * bhl <= CAP_NODES) {
        const int bw = (int)bwl, nn = (int)(bwl * bhl);
        for (int q = threadIdx.x; q < nn * NCAP; q += 256) tile[q] = 0ull;
        __syncthreads();
        if (live) {
            float vs[NCAP];
#pragma unroll
            for (int c = 0; c < NCAP; ++c) vs[c] = val[c] * scale[c];
            // each lane walks its 3x3 stencil from a start rotated by its lane index,
            // so lanes whose particles share a cell (dense or compressed regions) add
            // to different nodes at each step
            float wx[3], wy[3];
            int ox[3], oy[3];
            {
                const int r0 = lane % 3, r1 = (lane / 3) % 3;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int ax = a + r0 >= 3 ? a + r0 - 3 : a + r0;
                    const int ay = a + r1 >= 3 ? a + r1 - 3 : a + r1;
                    wx[a] = ax == 0 ? st.w[0][0] : ax == 1 ? st.w[0][1] : st.w[0][2];
                    wy[a] = ay == 0 ? st.w[1][0] : ay == 1 ? st.w[1][1] : st.w[1][2];
                    ox[a] = st.base[0] + ax - bx0;
                    oy[a] = (st.base[1] + ay - by0) * bw;
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const float wt = wx[a] * wy[b];
                    const int q = oy[b] + ox[a];
#pragma unroll
                    for (int c = 0; c < NCAP; ++c) {
                        const float x = rintf(wt * vs[c]);
                        const float hi = floorf(x * 2.384185791015625e-7f);   // x / 2^22
                        const float lo = fmaf(-hi, 4194304.0f, x);            // exact, in [0, 2^22)
                        const long long v = (long long)(int)hi * 4194304ll + (long long)(int)lo;
                        if (v != 0) atomicAdd(&tile[c * nn + q], (unsigned long long)v);
                    }
                }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nn; q += 256) {
            const int i = bx0 + q % bw, j = by0 + q / bw;
            if (i < 0 || i >= w || j < 0 || j >= h) continue;
            const size_t g = (size_t)j * w + i;
#pragma unroll
            for (int c = 0; c < NCAP; ++c) {
                const long long t = (long long)tile[c * nn + q];
                if (t == 0) continue;
                const double v = (double)t * inv_scale[c];
                atomicAdd(sample + (ch0 + c) * hw + g, (float)v);
            }
        }
    } else if (live) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int i = st.base[0] + a, j = st.base[1] + b;
                if (i < 0 || i >= w || j < 0 || j >= h) continue;
                const float wt = st.w[0][a] * st.w[1][b];
                const size_t g = (size_t)j * w + i;
#pragma unroll
                for (int c = 0; c < NCAP; ++c) atomicAdd(sample + (ch0 + c) * hw + g, wt * val[c]);
            }
    }
    // per-workgroup sums of particle mass and wall-band mass (reduced by
    // k_info_reduce: no same-address atomics across workgroups)
    if constexpr (MODE == 0) {
        const float a = block_sum(pm, sh);
        const float b = block_sum(inband, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = make_float2(a, b);
    }
}

// Momentum sums to per-node means (channels 1-4 in mode 0, 5-8 in mode 1), and
// per-workgroup sums of the image's grid mass.
__global__ void __launch_bounds__(256) k_capture_finish(float* __restrict__ sample, int hw, float* __restrict__ part,
                                                       int mode)
{
    __shared__ float sh[4];
    const int q = blockIdx.x * 256 + threadIdx.x;
    float m = 0.f;
    if (q < hw) {
        m = sample[q];
        const float inv = m > 0.f ? 1.f / m : 0.f;
        if (mode == 0) {
            for (int c = 1; c <= 4; ++c) sample[(size_t)c * hw + q] *= inv;
        } else {
            for (int c = 5; c <= 8; ++c) sample[(size_t)c * hw + q] *= inv;
        }
    }
    if (mode == 0) {
        const float a = block_sum(m, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = a;
    }
}

// One Richardson step of the target fields' least-squares fit, first half: per
// particle of [first, first + count), the residual r_p = y_p - sum_i w_ip t_i of the
// four target channels (velocity now, displacement since the saved position) at the
// saved position, scattered as w_ip m_p r_p into acc [4][h*w].  Nodes outside the
// image take no part, in the gather as in the scatter (as in the capture).
__global__ void __launch_bounds__(256) k_jacobi_residual(DevParticles P, int first, int count, Phys ph, int h,
                                                        int w, const float* __restrict__ sample,
                                                        const float* __restrict__ xs, float* __restrict__ acc)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const int p = first + k, n = P.n;
    const size_t hw = (size_t)h * w;
    const uint32_t id = P.pid[p];
    const float s0 = xs[2 * (size_t)id], s1 = xs[2 * (size_t)id + 1];
    float r[4] = {P.v[p], P.v[n + p], P.x[p] - s0, P.x[n + p] - s1};
    const float m = P.mass[p];
    const float xg[2] = {s0 * ph.inv_dx, s1 * ph.inv_dx};
    Stencil<2, 2> st;
    st.init(xg);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int i = st.base[0] + a, j = st.base[1] + b;
            if (i < 0 || i >= w || j < 0 || j >= h) continue;
            const float wt = st.w[0][a] * st.w[1][b];
            const size_t q = (size_t)j * w + i;
#pragma unroll
            for (int c = 0; c < 4; ++c) r[c] -= wt * sample[(5 + c) * hw + q];
        }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int i = st.base[0] + a, j = st.base[1] + b;
            if (i < 0 || i >= w || j < 0 || j >= h) continue;
            const float wm = st.w[0][a] * st.w[1][b] * m;
            const size_t q = (size_t)j * w + i;
#pragma unroll
            for (int c = 0; c < 4; ++c) atomicAdd(acc + c * hw + q, wm * r[c]);
        }
}

// Second half: t_i += acc_i / m_i on occupied nodes; acc is left zero for the next step.
__global__ void __launch_bounds__(256) k_jacobi_apply(float* __restrict__ sample, float* __restrict__ acc, int hw)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= hw) return;
    const float m = sample[q];
    const float inv = m > 0.f ? 1.f / m : 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        sample[(size_t)(5 + c) * hw + q] += acc[(size_t)c * hw + q] * inv;
        acc[(size_t)c * hw + q] = 0.f;
    }
}

// The slot's particle, wall-band and grid mass from the workgroup sums.
__global__ void __launch_bounds__(256) k_info_reduce(const float2* __restrict__ pp, int npp,
                                                    const float* __restrict__ pg, int npg, SlotInfo* info)
{
    __shared__ float sh[4];
    float a = 0.f, b = 0.f, c = 0.f;
    for (int i = threadIdx.x; i < npp; i += 256) {
        const float2 v = pp[i];
        a += v.x;
        b += v.y;
    }
    for (int i = threadIdx.x; i < npg; i += 256) c += pg[i];
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    c = block_sum(c, sh);
    if (threadIdx.x == 0) {
        info->particle_mass = a;
        info->band_mass = b;
        info->grid_mass = c;
    }
}

// Batch of network inputs, projection weights and targets from dataset slots:
// `images` (1 or 2) images per slot, image images * b + mode for slot b (mode 0
// velocity, 1 displacement).
__global__ void __launch_bounds__(256) k_pack(const float* __restrict__ data, const SlotInfo* __restrict__ info,
                                             const int* __restrict__ slots, int nb, int images, int hw, float d_ref,
                                             bf16* __restrict__ in, float* __restrict__ wm,
                                             float* __restrict__ target)
{
    const int64_t n = (int64_t)images * nb * hw;
    for (int64_t t = blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int img = (int)(t / hw), q = (int)(t - (int64_t)img * hw);
        const int mode = images == 2 ? (img & 1) : 0;
        const int slot = slots[images == 2 ? img >> 1 : img];
        const float* s = data + (size_t)slot * NCH * hw;
        const SlotInfo I = info[slot];
        const float m = s[q];
        const bool occ = m > 0.f;
        float f[8] = {m / I.m_ref, s[hw + q] / V_REF, s[2 * hw + q] / V_REF,
                      occ ? (s[3 * (size_t)hw + q] - 1.f) * 10.f : 0.f, s[4 * (size_t)hw + q] * 10.f,
                      occ ? 1.f : 0.f, (float)mode, 0.f};
        typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
        bf16x8 v0, z;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            v0[j] = (bf16)f[j];
            z[j] = (bf16)0.f;
        }
        bf16x8* dst = reinterpret_cast<bf16x8*>(in + (size_t)t * 32);
        dst[0] = v0;
        dst[1] = z;
        dst[2] = z;
        dst[3] = z;
        wm[t] = I.grid_mass > 0.f ? m / I.grid_mass : 0.f;
        if (target) {
            const float vx = s[hw + q], vy = s[2 * (size_t)hw + q];
            if (mode == 0) {
                target[2 * t] = occ ? (s[5 * (size_t)hw + q] - vx) / d_ref : 0.f;
                target[2 * t + 1] = occ ? (s[6 * (size_t)hw + q] - vy - I.gravity * I.T) / d_ref : 0.f;
            } else {
                const float su = d_ref * I.T;
                target[2 * t] = occ ? (s[7 * (size_t)hw + q] - I.T * vx) / su : 0.f;
                target[2 * t + 1] = occ ? (s[8 * (size_t)hw + q] - I.T * vy - I.ug) / su : 0.f;
            }
        }
    }
}

// Macro step: particles of [first, first+count) advanced by T = k dt with the
// predicted grid velocity v_pred,i = v_i + g T e_y + d_ref * outv_i and
// displacement u_pred,i = T v_i + ug e_y + d_ref T * outu_i (outv == NULL: ballistic,
// zero increments):
//   x' = x + sum_i w_ip u_pred,i,   F_E' = clamp((I + D^-1 sum_i w_ip u_pred,i (x_i - x_p)^T) F_E),
//   v' = sum_i w_ip v_pred,i,       C' = D^-1 sum_i w_ip v_pred,i (x_i - x_p)^T,
// or, where flags leaves out MACRO_X / MACRO_F, the velocity-only forms
// x' = x + T (v + v') / 2 and F_E' = clamp((I + T C') F_E);
// with the singular values of F_E' clamped as in G2P (g2p.glsl:80-93) and J_P
// scaled by the clamped ratio.  Nodes outside the image take the particle's own
// ballistic motion.
__global__ void __launch_bounds__(256) k_macro_g2p(DevParticles P, int first, int count, Phys ph, int h, int w,
                                                  const float* __restrict__ sample, const float* __restrict__ outv,
                                                  const float* __restrict__ outu, float d_ref, float T, float ug,
                                                  int flags, const int* __restrict__ gate)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    // a rollout whose residual test has failed leaves the particles as they are
    if (gate && gate[0] != 0) return;
    const int p = first + k, n = P.n;
    const size_t hw = (size_t)h * w;
    const float xp[2] = {P.x[p], P.x[n + p]};
    const float vp[2] = {P.v[p], P.v[n + p]};
    const float xg[2] = {xp[0] * ph.inv_dx, xp[1] * ph.inv_dx};
    Stencil<2, 2> st;
    st.init(xg);
    float nv[2] = {0.f, 0.f}, nB[4] = {0.f, 0.f, 0.f, 0.f}, nu[2] = {0.f, 0.f}, nG[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int i = st.base[0] + a, j = st.base[1] + b;
            const float wgt = st.w[0][a] * st.w[1][b];
            float vt[2] = {vp[0], vp[1]}, dv[2] = {0.f, 0.f}, du[2] = {0.f, 0.f};
            if (i >= 0 && i < w && j >= 0 && j < h) {
                const size_t q = (size_t)j * w + i;
                vt[0] = sample[hw + q];
                vt[1] = sample[2 * hw + q];
                if (outv) {
                    dv[0] = d_ref * outv[2 * q];
                    dv[1] = d_ref * outv[2 * q + 1];
                }
                if (outu) {
                    du[0] = d_ref * T * outu[2 * q];
                    du[1] = d_ref * T * outu[2 * q + 1];
                }
            }
            const float vi[2] = {vt[0] + dv[0], vt[1] + ph.gravity * T + dv[1]};
            const float ui[2] = {T * vt[0] + du[0], T * vt[1] + ug + du[1]};
            const float dp[2] = {st.d[0][a] * ph.dx, st.d[1][b] * ph.dx};
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                nv[r] += wgt * vi[r];
                nu[r] += wgt * ui[r];
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    nB[r * 2 + c] += wgt * (vi[r] * (-dp[c]));
                    nG[r * 2 + c] += wgt * (ui[r] * (-dp[c]));
                }
            }
        }
    float Cn[4], M[4], Fo[4], Fh[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        Cn[e] = ph.Dinv * nB[e];
        M[e] = ((e % 3) == 0 ? 1.f : 0.f) + ((flags & MACRO_F) ? ph.Dinv * nG[e] : T * Cn[e]);
        Fo[e] = P.F[e * n + p];
    }
    matmul<2>(M, Fo, Fh);
    float U[4], sig[2], V[4], sc[2];
    svd<2>(Fh, U, sig, V);
    float ratio = 1.f;
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        float sv = fminf(fmaxf(sig[d], ph.theta0), ph.theta1);
        sv = fmaxf(sv, 1e-6f);
        sc[d] = sv;
        ratio *= sig[d] / sv;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
            P.F[(r * 2 + c) * n + p] = (U[r * 2] * sc[0]) * V[c * 2] + (U[r * 2 + 1] * sc[1]) * V[c * 2 + 1];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        P.x[d * n + p] = (flags & MACRO_X) ? xp[d] + nu[d] : xp[d] + 0.5f * T * (vp[d] + nv[d]);
        P.v[d * n + p] = nv[d];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) P.C[e * n + p] = Cn[e];
    P.Jp[p] = P.Jp[p] * ratio;
}

__global__ void k_init_info(SlotInfo* info, float T, float g, float m_ref, float ug)
{
    info->particle_mass = 0.f;
    info->grid_mass = 0.f;
    info->band_mass = 0.f;
    info->T = T;
    info->gravity = g;
    info->m_ref = m_ref;
    info->ug = ug;
    info->pad = 0.f;
}

__global__ void __launch_bounds__(256) k_zero(float* p, int64_t n)
{
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0.f;
}

__global__ void k_copy_info(SlotInfo* dst, const SlotInfo* src) { *dst = *src; }

}  // namespace

// The residual test of a macro step on the device (the host form is in
// mpm_surrogate_advance): the live slot's share of the scene's mass in the wall band
// or outside the image against tol.  gate[2] counts the macro steps taken through
// the gate; the first one that fails is recorded in gate[0] (1-based) and from then on
// the gated kernels leave the particles alone; gate[1] holds the largest residual seen
// before it (as the bits of a non-negative float, so integer max orders them).
__global__ void k_residual_gate(const SlotInfo* __restrict__ info, float tol, int* __restrict__ gate)
{
    if (gate[0] != 0) return;
    const SlotInfo I = *info;
    const float M = I.particle_mass > 0.f ? I.particle_mass : 1.f;
    const float res = fmaxf(I.band_mass / M, fabsf(M - I.grid_mass) / M);
    const int step = gate[2] + 1;
    gate[2] = step;
    if (!(res <= tol)) {
        gate[0] = step;
        return;
    }
    gate[1] = max(gate[1], __float_as_int(res));
}

__global__ void k_gate_reset(int* gate)
{
    gate[0] = 0;
    gate[1] = 0;
    gate[2] = 0;
}

struct mpm_surrogate {
    hipStream_t st = nullptr;
    UNet net;
    int h = 0, w = 0, max_batch = 0, capacity = 0;
    float d_ref = 0.1f;
    int macro_flags = 0;           // MPM_MACRO_DISPLACEMENT_* of the macro step
    int refine_iters = 0;          // least-squares refinement steps of the target fields
    int images() const { return macro_flags ? 2 : 1; }   // network images per slot
    float* data = nullptr;         // (capacity + 1) slots x NCH x h*w; the last is the live slot
    SlotInfo* info = nullptr;      // capacity + 1
    float* xs = nullptr;           // saved positions by original particle index
    int64_t xs_cap = 0;
    float2* part_p = nullptr;      // per-workgroup particle / wall-band mass of a capture (xs_cap / 256 + 1)
    float* acc = nullptr;          // [4][h*w] scatter of the target fit's residuals (zero between uses)
    float* part_g = nullptr;       // per-workgroup grid mass of a capture (h * w / 256 + 1)
    bf16* in = nullptr;            // network batch: 2 images (velocity, displacement) per slot
    float* wm = nullptr;
    float* target = nullptr;
    int* slots_d = nullptr;
    int* live_slot_d = nullptr;     // the live slot's index, for k_pack
    float* loss_d = nullptr;
    SlotInfo* info_h = nullptr;     // pinned copy of the live slot's info
    hipEvent_t ev_solver = nullptr; // orders a capture after the solver's work
    hipEvent_t ev_capture = nullptr;// orders the solver's next work after a capture
    hipEvent_t ev_info = nullptr;   // the live slot's info has reached info_h
    int live() const { return capacity; }
    // mpm_surrogate_rollout: device residual gate (failed step, largest residual,
    // steps taken) and the graph of one gated macro step with what it was recorded for
    int* gate = nullptr;
    int* gate_h = nullptr;
    hipGraphExec_t macro_graph = nullptr;
    const void* graph_solver = nullptr;
    const void* graph_particles = nullptr;
    int graph_k = 0, graph_mode = 0, graph_flags = 0, graph_conv = 0;
    float graph_tol = 0.f;
    void drop_graph()
    {
        if (macro_graph) (void)hipGraphExecDestroy(macro_graph);
        macro_graph = nullptr;
        graph_solver = graph_particles = nullptr;
    }
};

namespace {

int check_scene(mpm_surrogate* g, mpm_solver* s, int scene)
{
    if (!g || !s) return sfail(MPM_E_INVALID, "NULL argument");
    if (s->dim != 2) return sfail(MPM_E_INVALID, "the surrogate handles 2D scenes");
    if (scene < 0 || scene >= s->nscenes) return sfail(MPM_E_INVALID, "scene out of range");
    const mpm_params& P = s->scene_params[scene];
    if (P.spline_order != 2) return sfail(MPM_E_INVALID, "the surrogate uses the quadratic B-spline");
    if (P.WIDTH != g->w || P.HEIGHT != g->h)
        return sfail(MPM_E_INVALID, "scene grid does not match the surrogate image (WIDTH x HEIGHT cells)");
    return MPM_OK;
}

int ensure_xs(mpm_surrogate* g, mpm_solver* s)
{
    if (g->xs_cap >= s->n) return MPM_OK;
    if (g->xs) S_TRY(hipFree(g->xs));
    if (g->part_p) S_TRY(hipFree(g->part_p));
    g->xs = nullptr;
    g->part_p = nullptr;
    g->xs_cap = 0;
    S_TRY(hipMalloc(&g->xs, (size_t)std::max(s->n, 1) * 2 * sizeof(float)));
    S_TRY(hipMalloc(&g->part_p, ((size_t)s->n / 256 + 1) * sizeof(float2)));
    g->xs_cap = s->n;
    return MPM_OK;
}

// Capture into `slot`: mode 0 the features (and saved positions), mode 1 the target.
// On the solver's own stream (on_solver_stream: a macro step's capture, ordered with
// the solver's work by the stream itself), or on the surrogate's stream with events
// ordering it after the solver's work and the solver's next work after it (no host
// synchronisation either way).
int capture(mpm_surrogate* g, mpm_solver* s, int scene, int slot, int mode, int k_substeps,
            bool on_solver_stream = false)
{
    int r = check_scene(g, s, scene);
    if (r) return r;
    if ((r = ensure_xs(g, s))) return r;
    hipStream_t st = on_solver_stream ? s->stream : g->st;
    if (!on_solver_stream) {
        S_TRY(hipEventRecord(g->ev_solver, s->stream));
        S_TRY(hipStreamWaitEvent(g->st, g->ev_solver, 0));
    }
    const mpm_params& P = s->scene_params[scene];
    const Phys ph = s->scenes_h[scene].ph;
    const int hw = g->h * g->w;
    float* sample = g->data + (size_t)slot * NCH * hw;
    SlotInfo* info = g->info + slot;
    const int first = (int)s->p_off[scene], count = (int)(s->p_off[scene + 1] - s->p_off[scene]);
    if (mode == 0) {
        k_zero<<<std::min((NCH * hw + 255) / 256, 4096), 256, 0, st>>>(sample, (int64_t)NCH * hw);
        const float dt = P.DELTA_TIME;
        k_init_info<<<1, 1, 0, st>>>(info, (float)k_substeps * dt, P.GRAVITY, P.RHO * P.dx * P.dx,
                                     P.GRAVITY * dt * dt * (0.5f * (float)k_substeps * (float)(k_substeps + 1)));
    }
    const float band = P.BORDER_MARGIN + 2.0f;
    const int nbp = (count + 255) / 256, nbg = (hw + 255) / 256;
    if (count > 0) {
        if (mode == 0)
            k_capture<0><<<nbp, 256, 0, st>>>(s->dev_particles(), first, count, ph, g->h, g->w, band, sample,
                                                 g->xs, g->part_p);
        else
            k_capture<1><<<nbp, 256, 0, st>>>(s->dev_particles(), first, count, ph, g->h, g->w, band, sample,
                                                 g->xs, g->part_p);
    }
    k_capture_finish<<<nbg, 256, 0, st>>>(sample, hw, g->part_g, mode);
    if (mode == 0) k_info_reduce<<<1, 256, 0, st>>>(g->part_p, nbp, g->part_g, nbg, info);
    if (mode == 1 && count > 0) {
        // least-squares refinement of the target fields (channels 5-8)
        for (int it = 0; it < g->refine_iters; ++it) {
            k_jacobi_residual<<<nbp, 256, 0, st>>>(s->dev_particles(), first, count, ph, g->h, g->w, sample, g->xs,
                                                   g->acc);
            k_jacobi_apply<<<nbg, 256, 0, st>>>(sample, g->acc, hw);
        }
    }
    S_TRY(hipGetLastError());
    if (!on_solver_stream) {
        S_TRY(hipEventRecord(g->ev_capture, g->st));
        S_TRY(hipStreamWaitEvent(s->stream, g->ev_capture, 0));
    }
    return MPM_OK;
}

int upload_slots(mpm_surrogate* g, const int* slots, int nb)
{
    if (nb < 1 || nb > g->max_batch) return sfail(MPM_E_INVALID, "batch size out of range");
    for (int b = 0; b < nb; ++b)
        if (slots[b] < 0 || slots[b] >= g->capacity) return sfail(MPM_E_INVALID, "slot out of range");
    S_TRY(hipMemcpyAsync(g->slots_d, slots, sizeof(int) * nb, hipMemcpyHostToDevice, g->st));
    return MPM_OK;
}

// One macro step with the residual test on the device, enqueued on the solver's
// stream (mpm_surrogate_rollout records it into a graph): capture of the live slot,
// the gate, the network (MPM_ADVANCE_NETWORK) and the gated macro G2P.  The caller
// has sized the capture's buffers (ensure_xs) and ordered the solver's stream after
// the surrogate stream's work.
int enqueue_gated_macro(mpm_surrogate* g, mpm_solver* s, int k_substeps, int mode, float tol)
{
    hipStream_t st = s->stream;
    int r = capture(g, s, 0, g->live(), 0, k_substeps, true);
    if (r) return r;
    k_residual_gate<<<1, 1, 0, st>>>(g->info + g->live(), tol, g->gate);
    const int hw = g->h * g->w;
    const float* out = nullptr;
    if (mode == MPM_ADVANCE_NETWORK) {
        const int ni = g->images();
        k_pack<<<std::min((ni * hw + 255) / 256, 8192), 256, 0, st>>>(g->data, g->info, g->live_slot_d, 1, ni, hw,
                                                                      g->d_ref, g->in, g->wm, nullptr);
        S_TRY(hipGetLastError());
        S_TRY(g->net.forward(ni, g->in, g->wm, st));
        out = g->net.out();
    }
    const Phys ph = s->scenes_h[0].ph;
    const float T = (float)k_substeps * ph.dt;
    const float ug = ph.gravity * ph.dt * ph.dt * (0.5f * (float)k_substeps * (float)(k_substeps + 1));
    const float* sample = g->data + (size_t)g->live() * NCH * hw;
    const float* outu = out && g->images() == 2 ? out + (size_t)2 * hw : nullptr;
    k_macro_g2p<<<(s->n + 255) / 256, 256, 0, st>>>(s->dev_particles(), 0, s->n, ph, g->h, g->w, sample, out, outu,
                                                    g->d_ref, T, ug, g->macro_flags, g->gate);
    S_TRY(hipGetLastError());
    return MPM_OK;
}

}  // namespace

extern "C" {

int mpm_surrogate_create(int height, int width, int max_batch, int capacity, uint64_t seed, mpm_surrogate** out)
{
    if (!out) return sfail(MPM_E_INVALID, "out is NULL");
    *out = nullptr;
    if (capacity < 0 || max_batch < 1) return sfail(MPM_E_INVALID, "capacity/max_batch out of range");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1) return sfail(MPM_E_HIP, "no HIP device");
    mpm_surrogate* g = new mpm_surrogate();
    g->h = height;
    g->w = width;
    g->max_batch = max_batch;
    g->capacity = capacity;
    auto bail = [&](int code) {
        mpm_surrogate_destroy(g);
        return code;
    };
    if (hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) != hipSuccess)
        return bail(sfail(MPM_E_HIP, "hipStreamCreate failed"));
    std::string err;
    int r = g->net.init(height, width, 2 * max_batch, seed, g->st, &err);
    if (r) return bail(sfail(r == -1 ? MPM_E_INVALID : MPM_E_HIP, err));
    const size_t hw = (size_t)height * width;
    if (hipMalloc(&g->data, (size_t)(capacity + 1) * NCH * hw * sizeof(float)) != hipSuccess ||
        hipMalloc(&g->info, (size_t)(capacity + 1) * sizeof(SlotInfo)) != hipSuccess ||
        hipMalloc(&g->in, (size_t)2 * max_batch * hw * 32 * sizeof(bf16)) != hipSuccess ||
        hipMalloc(&g->wm, (size_t)2 * max_batch * hw * sizeof(float)) != hipSuccess ||
        hipMalloc(&g->target, (size_t)2 * max_batch * hw * 2 * sizeof(float)) != hipSuccess ||
        hipMalloc(&g->acc, (size_t)4 * hw * sizeof(float)) != hipSuccess ||
        hipMalloc(&g->slots_d, (size_t)max_batch * sizeof(int)) != hipSuccess ||
        hipMalloc(&g->loss_d, 256) != hipSuccess ||
        hipMalloc(&g->part_g, (hw / 256 + 1) * sizeof(float)) != hipSuccess ||
        hipMalloc(&g->live_slot_d, sizeof(int)) != hipSuccess ||
        hipMalloc(&g->gate, 4 * sizeof(int)) != hipSuccess ||
        hipHostMalloc(&g->gate_h, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&g->info_h, sizeof(SlotInfo), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_solver, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_capture, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_info, hipEventDisableTiming) != hipSuccess)
        return bail(sfail(MPM_E_HIP, "surrogate allocation failed"));
    const int live_slot = capacity;
    if (hipMemcpyAsync(g->live_slot_d, &live_slot, sizeof(int), hipMemcpyHostToDevice, g->st) != hipSuccess ||
        hipMemsetAsync(g->info, 0, (size_t)(capacity + 1) * sizeof(SlotInfo), g->st) != hipSuccess ||
        hipMemsetAsync(g->in, 0, (size_t)2 * max_batch * hw * 32 * sizeof(bf16), g->st) != hipSuccess ||
        hipMemsetAsync(g->wm, 0, (size_t)2 * max_batch * hw * sizeof(float), g->st) != hipSuccess ||
        hipMemsetAsync(g->target, 0, (size_t)2 * max_batch * hw * 2 * sizeof(float), g->st) != hipSuccess ||
        hipMemsetAsync(g->acc, 0, (size_t)4 * hw * sizeof(float), g->st) != hipSuccess ||
        hipStreamSynchronize(g->st) != hipSuccess)
        return bail(sfail(MPM_E_HIP, "surrogate initialisation failed"));
    *out = g;
    return MPM_OK;
}

int mpm_surrogate_destroy(mpm_surrogate* g)
{
    if (!g) return MPM_OK;
    if (g->st) (void)hipStreamSynchronize(g->st);
    g->drop_graph();
    (void)hipFree(g->gate);
    if (g->gate_h) (void)hipHostFree(g->gate_h);
    g->net.release();
    (void)hipFree(g->data);
    (void)hipFree(g->info);
    (void)hipFree(g->xs);
    (void)hipFree(g->part_p);
    (void)hipFree(g->part_g);
    (void)hipFree(g->in);
    (void)hipFree(g->wm);
    (void)hipFree(g->target);
    (void)hipFree(g->acc);
    (void)hipFree(g->slots_d);
    (void)hipFree(g->live_slot_d);
    (void)hipFree(g->loss_d);
    if (g->info_h) (void)hipHostFree(g->info_h);
    if (g->ev_solver) (void)hipEventDestroy(g->ev_solver);
    if (g->ev_capture) (void)hipEventDestroy(g->ev_capture);
    if (g->ev_info) (void)hipEventDestroy(g->ev_info);
    if (g->st) (void)hipStreamDestroy(g->st);
    delete g;
    return MPM_OK;
}

int mpm_surrogate_num_params(const mpm_surrogate* g, int64_t* n)
{
    if (!g || !n) return sfail(MPM_E_INVALID, "NULL argument");
    *n = const_cast<mpm_surrogate*>(g)->net.num_params();
    return MPM_OK;
}

int mpm_surrogate_get_params(mpm_surrogate* g, float* host)
{
    if (!g || !host) return sfail(MPM_E_INVALID, "NULL argument");
    S_TRY(hipMemcpyAsync(host, g->net.params(), g->net.num_params() * 4, hipMemcpyDeviceToHost, g->st));
    S_TRY(hipStreamSynchronize(g->st));
    return MPM_OK;
}

int mpm_surrogate_set_params(mpm_surrogate* g, const float* host)
{
    if (!g || !host) return sfail(MPM_E_INVALID, "NULL argument");
    S_TRY(hipMemcpyAsync(g->net.params(), host, g->net.num_params() * 4, hipMemcpyHostToDevice, g->st));
    S_TRY(g->net.sync_weights());
    S_TRY(hipStreamSynchronize(g->st));
    return MPM_OK;
}

int mpm_surrogate_set_macro(mpm_surrogate* g, int flags)
{
    if (!g) return sfail(MPM_E_INVALID, "NULL argument");
    g->drop_graph();   // the recorded macro step holds the old setting
    if (flags & ~(MPM_MACRO_DISPLACEMENT_X | MPM_MACRO_DISPLACEMENT_F)) return sfail(MPM_E_INVALID, "unknown macro flags");
    g->macro_flags = flags;
    return MPM_OK;
}

int mpm_surrogate_set_targets(mpm_surrogate* g, int refine_iters)
{
    if (!g) return sfail(MPM_E_INVALID, "NULL argument");
    if (refine_iters < 0 || refine_iters > 64) return sfail(MPM_E_INVALID, "refine_iters must be in [0, 64]");
    g->refine_iters = refine_iters;
    return MPM_OK;
}


int mpm_surrogate_set_conv(mpm_surrogate* g, int kernel)
{
    if (!g) return sfail(MPM_E_INVALID, "NULL argument");
    g->drop_graph();   // the recorded macro step holds the old setting
    if (kernel < MPM_CONV_AUTO || kernel > MPM_CONV_LDS) return sfail(MPM_E_INVALID, "unknown convolution kernel");
    g->net.set_conv_kernel(kernel);
    return MPM_OK;
}

int mpm_surrogate_set_scale(mpm_surrogate* g, float d_ref)
{
    if (!g) return sfail(MPM_E_INVALID, "NULL argument");
    g->drop_graph();   // the recorded macro step holds the old setting
    if (!(d_ref > 0.f)) return sfail(MPM_E_INVALID, "d_ref must be > 0");
    g->d_ref = d_ref;
    return MPM_OK;
}

int mpm_surrogate_forward(mpm_surrogate* g, int nb, const uint16_t* input, const float* wm, float* out)
{
    if (!g || !input || !wm || !out) return sfail(MPM_E_INVALID, "NULL argument");
    if (nb < 1 || nb > g->max_batch) return sfail(MPM_E_INVALID, "batch size out of range");
    const size_t hw = (size_t)g->h * g->w;
    S_TRY(hipMemcpyAsync(g->in, input, nb * hw * 32 * 2, hipMemcpyHostToDevice, g->st));
    S_TRY(hipMemcpyAsync(g->wm, wm, nb * hw * 4, hipMemcpyHostToDevice, g->st));
    S_TRY(g->net.forward(nb, g->in, g->wm));
    S_TRY(hipMemcpyAsync(out, g->net.out(), nb * hw * 2 * 4, hipMemcpyDeviceToHost, g->st));
    S_TRY(hipStreamSynchronize(g->st));
    return MPM_OK;
}

int mpm_surrogate_gradient(mpm_surrogate* g, int nb, const uint16_t* input, const float* wm, const float* target,
                           float* loss, float* grad)
{
    if (!g || !input || !wm || !target) return sfail(MPM_E_INVALID, "NULL argument");
    if (nb < 1 || nb > g->max_batch) return sfail(MPM_E_INVALID, "batch size out of range");
    const size_t hw = (size_t)g->h * g->w;
    S_TRY(hipMemcpyAsync(g->in, input, nb * hw * 32 * 2, hipMemcpyHostToDevice, g->st));
    S_TRY(hipMemcpyAsync(g->wm, wm, nb * hw * 4, hipMemcpyHostToDevice, g->st));
    S_TRY(hipMemcpyAsync(g->target, target, nb * hw * 2 * 4, hipMemcpyHostToDevice, g->st));
    S_TRY(unet::fill_f32(g->loss_d, 1, 0.f, g->st));
    S_TRY(g->net.loss_and_backward(nb, g->in, g->wm, g->target, g->loss_d, grad != nullptr));
    if (loss) S_TRY(hipMemcpyAsync(loss, g->loss_d, 4, hipMemcpyDeviceToHost, g->st));
    if (grad) S_TRY(hipMemcpyAsync(grad, g->net.grads(), g->net.num_params() * 4, hipMemcpyDeviceToHost, g->st));
    S_TRY(hipStreamSynchronize(g->st));
    return MPM_OK;
}

int mpm_surrogate_record(mpm_surrogate* g, mpm_solver* s, int scene, int slot, int what, int k_substeps)
{
    if (!g) return sfail(MPM_E_INVALID, "surrogate is NULL");
    if (slot < 0 || slot >= g->capacity) return sfail(MPM_E_INVALID, "slot out of range");
    if (what != 0 && what != 1) return sfail(MPM_E_INVALID, "what must be 0 (input) or 1 (target)");
    if (k_substeps < 1) return sfail(MPM_E_INVALID, "k_substeps must be >= 1");
    return capture(g, s, scene, slot, what, k_substeps);
}

int mpm_surrogate_get_sample(mpm_surrogate* g, int slot, float* host, float* scalars)
{
    if (!g) return sfail(MPM_E_INVALID, "surrogate is NULL");
    if (slot < 0 || slot > g->capacity) return sfail(MPM_E_INVALID, "slot out of range");
    const size_t hw = (size_t)g->h * g->w;
    if (host) S_TRY(hipMemcpyAsync(host, g->data + (size_t)slot * NCH * hw, NCH * hw * 4, hipMemcpyDeviceToHost, g->st));
    SlotInfo I{};
    S_TRY(hipMemcpyAsync(&I, g->info + slot, sizeof(I), hipMemcpyDeviceToHost, g->st));
    S_TRY(hipStreamSynchronize(g->st));
    if (scalars) {
        scalars[0] = I.particle_mass;
        scalars[1] = I.grid_mass;
        scalars[2] = I.band_mass;
        scalars[3] = I.T;
        scalars[4] = I.gravity;
        scalars[5] = I.m_ref;
    }
    return MPM_OK;
}

int mpm_surrogate_train(mpm_surrogate* g, const int* slots, int nb, float lr, float* loss)
{
    if (!g || !slots) return sfail(MPM_E_INVALID, "NULL argument");
    int r = upload_slots(g, slots, nb);
    if (r) return r;
    const int hw = g->h * g->w;
    const int ni = g->images();
    k_pack<<<std::min((ni * nb * hw + 255) / 256, 8192), 256, 0, g->st>>>(g->data, g->info, g->slots_d, nb, ni, hw,
                                                                          g->d_ref, g->in, g->wm, g->target);
    S_TRY(hipGetLastError());
    S_TRY(unet::fill_f32(g->loss_d, 1, 0.f, g->st));
    S_TRY(g->net.loss_and_backward(ni * nb, g->in, g->wm, g->target, g->loss_d, lr > 0.f));
    if (lr > 0.f) S_TRY(g->net.adam(lr));
    if (loss) {
        S_TRY(hipMemcpyAsync(loss, g->loss_d, 4, hipMemcpyDeviceToHost, g->st));
        S_TRY(hipStreamSynchronize(g->st));
    }
    return MPM_OK;
}

int mpm_surrogate_baseline_loss(mpm_surrogate* g, const int* slots, int nb, float* loss)
{
    // the loss of the zero increment (ballistic grid velocity), on the same samples
    if (!g || !slots || !loss) return sfail(MPM_E_INVALID, "NULL argument");
    int r = upload_slots(g, slots, nb);
    if (r) return r;
    const int hw = g->h * g->w;
    const int ni = g->images();
    k_pack<<<std::min((ni * nb * hw + 255) / 256, 8192), 256, 0, g->st>>>(g->data, g->info, g->slots_d, nb, ni, hw,
                                                                          g->d_ref, g->in, g->wm, g->target);
    S_TRY(hipGetLastError());
    std::vector<float> t((size_t)ni * nb * hw * 2), wm((size_t)ni * nb * hw);
    S_TRY(hipMemcpyAsync(t.data(), g->target, t.size() * 4, hipMemcpyDeviceToHost, g->st));
    S_TRY(hipMemcpyAsync(wm.data(), g->wm, wm.size() * 4, hipMemcpyDeviceToHost, g->st));
    S_TRY(hipStreamSynchronize(g->st));
    double l = 0.0;
    for (size_t i = 0; i < wm.size(); ++i) l += (double)wm[i] * ((double)t[2 * i] * t[2 * i] + (double)t[2 * i + 1] * t[2 * i + 1]);
    *loss = (float)(l / (ni * nb));
    return MPM_OK;
}

int mpm_surrogate_time(mpm_surrogate* g, int nb, int iters, int train, double* ms)
{
    if (!g || !ms) return sfail(MPM_E_INVALID, "NULL argument");
    if (nb < 1 || nb > g->net.max_batch() || iters < 1) return sfail(MPM_E_INVALID, "nb/iters out of range");
    hipEvent_t a = nullptr, b = nullptr;
    S_TRY(hipEventCreate(&a));
    S_TRY(hipEventCreate(&b));
    // whatever the batch buffers hold: the kernels' cost does not depend on values
    int r = MPM_OK;
    float t = 0.f;
    if (hipEventRecord(a, g->st) != hipSuccess) r = sfail(MPM_E_HIP, "hipEventRecord failed");
    for (int i = 0; i < iters && r == MPM_OK; ++i) {
        hipError_t e = train ? g->net.loss_and_backward(nb, g->in, g->wm, g->target, g->loss_d, true)
                             : g->net.forward(nb, g->in, g->wm);
        if (e != hipSuccess) r = sfail(MPM_E_HIP, std::string("network: ") + hipGetErrorString(e));
    }
    if (r == MPM_OK && (hipEventRecord(b, g->st) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                        hipEventElapsedTime(&t, a, b) != hipSuccess))
        r = sfail(MPM_E_HIP, "event timing failed");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (r == MPM_OK) *ms = (double)t / iters;
    return r;
}

int mpm_surrogate_advance(mpm_surrogate* g, mpm_solver* s, int k_substeps, int mode, float tol, int* used,
                          float* residual)
{
    if (!g || !s) return sfail(MPM_E_INVALID, "NULL argument");
    if (s->nscenes != 1) return sfail(MPM_E_INVALID, "advance runs on a single-scene solver");
    if (k_substeps < 1) return sfail(MPM_E_INVALID, "k_substeps must be >= 1");
    if (mode < MPM_ADVANCE_PLAIN || mode > MPM_ADVANCE_BALLISTIC) return sfail(MPM_E_INVALID, "unknown mode");
    if (used) *used = MPM_ADVANCE_PLAIN;
    if (residual) *residual = 0.f;
    if (mode == MPM_ADVANCE_PLAIN || s->n == 0) return mpm_step(s, k_substeps);
    // the whole macro step runs on the solver's stream, in order with its work; the
    // solver's stream first waits for the surrogate stream's pending work (training,
    // recording), which shares the network's and the capture's buffers
    hipStream_t st = s->stream;
    S_TRY(hipEventRecord(g->ev_solver, g->st));
    S_TRY(hipStreamWaitEvent(st, g->ev_solver, 0));
    int r = capture(g, s, 0, g->live(), 0, k_substeps, true);
    if (r) return r;
    const int hw = g->h * g->w;
    S_TRY(hipMemcpyAsync(g->info_h, g->info + g->live(), sizeof(SlotInfo), hipMemcpyDeviceToHost, st));
    S_TRY(hipEventRecord(g->ev_info, st));
    // the network runs while the host waits for the residual test's inputs (it reads
    // only the live slot; if the test fails its output is not used)
    const float* out = nullptr;
    if (mode == MPM_ADVANCE_NETWORK) {
        // one batch: velocity (image 0) and, when used, displacement (image 1)
        const int ni = g->images();
        k_pack<<<std::min((ni * hw + 255) / 256, 8192), 256, 0, st>>>(g->data, g->info, g->live_slot_d, 1, ni, hw,
                                                                      g->d_ref, g->in, g->wm, nullptr);
        S_TRY(hipGetLastError());
        S_TRY(g->net.forward(ni, g->in, g->wm, st));
        out = g->net.out();
    }
    S_TRY(hipEventSynchronize(g->ev_info));
    const SlotInfo I = *g->info_h;
    // share of mass that can exchange momentum with the walls in the next k
    // substeps, and mass outside the image: the prediction conserves momentum up to
    // the gravity impulse, so these bound its momentum and mass residuals
    const float M = I.particle_mass > 0.f ? I.particle_mass : 1.f;
    const float res = std::max(I.band_mass / M, std::fabs(M - I.grid_mass) / M);
    if (residual) *residual = res;
    if (!(res <= tol)) {
        // later work on the surrogate's stream waits for this call's use of the
        // shared buffers (capture, network)
        S_TRY(hipEventRecord(g->ev_capture, st));
        S_TRY(hipStreamWaitEvent(g->st, g->ev_capture, 0));
        // plain substeps after an earlier macro step are refused by mpm_step
        if (s->after_macro)
            return sfail(MPM_E_STATE, "the residual test failed after an earlier macro step, and plain "
                                      "substeps after a macro step are not enabled");
        return mpm_step(s, k_substeps);
    }

    const Phys ph = s->scenes_h[0].ph;
    const float T = (float)k_substeps * ph.dt;
    const float ug = ph.gravity * ph.dt * ph.dt * (0.5f * (float)k_substeps * (float)(k_substeps + 1));
    const float* sample = g->data + (size_t)g->live() * NCH * hw;
    const float* outu = out && g->images() == 2 ? out + (size_t)2 * hw : nullptr;
    k_macro_g2p<<<(s->n + 255) / 256, 256, 0, st>>>(s->dev_particles(), 0, s->n, ph, g->h, g->w, sample, out, outu,
                                                    g->d_ref, T, ug, g->macro_flags, nullptr);
    S_TRY(hipGetLastError());
    // later work on the surrogate's stream (training, recording) waits for this
    // macro step's use of the shared buffers: the live slot and the network output
    S_TRY(hipEventRecord(g->ev_capture, st));
    S_TRY(hipStreamWaitEvent(g->st, g->ev_capture, 0));
    // the solver's grid and binning no longer match its particles
    s->grid_state = mpm::detail::GRID_DIRTY;
    s->items_valid = false;
    s->after_macro = true;
    s->substeps_done += k_substeps;
    if (used) *used = mode;
    return MPM_OK;
}

int mpm_surrogate_rollout(mpm_surrogate* g, mpm_solver* s, int k_substeps, int n_macro, int mode, float tol,
                          int* n_network, int* n_plain, float* max_residual)
{
    if (!g || !s) return sfail(MPM_E_INVALID, "NULL argument");
    if (s->nscenes != 1) return sfail(MPM_E_INVALID, "rollout runs on a single-scene solver");
    if (k_substeps < 1 || n_macro < 0) return sfail(MPM_E_INVALID, "k_substeps must be >= 1 and n_macro >= 0");
    if (mode < MPM_ADVANCE_PLAIN || mode > MPM_ADVANCE_BALLISTIC) return sfail(MPM_E_INVALID, "unknown mode");
    int nn = 0, np = 0;
    float worst = 0.f;
    auto report = [&] {
        if (n_network) *n_network = nn;
        if (n_plain) *n_plain = np;
        if (max_residual) *max_residual = worst;
    };
    report();
    if (n_macro == 0) return MPM_OK;
    if (mode == MPM_ADVANCE_PLAIN || s->n == 0) {
        const int r = mpm_step(s, k_substeps * n_macro);
        if (r == MPM_OK) np = n_macro;
        report();
        return r;
    }
    // macro steps through the host test until the solver has taken one (a failed
    // test runs plain substeps only before that)
    int done = 0;
    while (done < n_macro && !s->after_macro) {
        int used = MPM_ADVANCE_PLAIN;
        float res = 0.f;
        const int r = mpm_surrogate_advance(g, s, k_substeps, mode, tol, &used, &res);
        if (r) { report(); return r; }
        if (used == MPM_ADVANCE_PLAIN) ++np;
        else { ++nn; worst = std::max(worst, res); }
        ++done;
    }
    if (done == n_macro) { report(); return MPM_OK; }

    // the rest: replays of one recorded macro step with the test on the device
    int r = check_scene(g, s, 0);
    if (r) { report(); return r; }
    if ((r = ensure_xs(g, s))) { report(); return r; }
    const void* particles = s->dev_particles().x;
    if (g->macro_graph && (g->graph_solver != s || g->graph_particles != particles || g->graph_k != k_substeps ||
                           g->graph_mode != mode || g->graph_tol != tol))
        g->drop_graph();
    // the solver's stream waits for pending surrogate work (training, recording)
    S_TRY(hipEventRecord(g->ev_solver, g->st));
    S_TRY(hipStreamWaitEvent(s->stream, g->ev_solver, 0));
    if (!g->macro_graph) {
        S_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        r = enqueue_gated_macro(g, s, k_substeps, mode, tol);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(s->stream, &graph);
        if (r) { if (graph) (void)hipGraphDestroy(graph); report(); return r; }
        if (e != hipSuccess) { report(); return sfail(MPM_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e)); }
        const hipError_t ei = hipGraphInstantiate(&g->macro_graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) { g->macro_graph = nullptr; report(); return sfail(MPM_E_HIP, "hipGraphInstantiate failed"); }
        g->graph_solver = s;
        g->graph_particles = particles;
        g->graph_k = k_substeps;
        g->graph_mode = mode;
        g->graph_tol = tol;
    }
    k_gate_reset<<<1, 1, 0, s->stream>>>(g->gate);
    S_TRY(hipGetLastError());
    for (int i = done; i < n_macro; ++i) S_TRY(hipGraphLaunch(g->macro_graph, s->stream));
    S_TRY(hipMemcpyAsync(g->gate_h, g->gate, 3 * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    // later work on the surrogate's stream waits for the rollout's use of the shared
    // buffers (live slot, network)
    S_TRY(hipEventRecord(g->ev_capture, s->stream));
    S_TRY(hipStreamWaitEvent(g->st, g->ev_capture, 0));
    S_TRY(hipStreamSynchronize(s->stream));
    const int failed = g->gate_h[0];
    const int taken = failed ? failed - 1 : n_macro - done;
    nn += taken;
    worst = std::max(worst, __builtin_bit_cast(float, g->gate_h[1]));
    s->grid_state = mpm::detail::GRID_DIRTY;
    s->items_valid = false;
    s->after_macro = true;
    s->substeps_done += (int64_t)k_substeps * taken;
    report();
    if (failed)
        return sfail(MPM_E_STATE, "the residual test failed at macro step " + std::to_string(done + failed) +
                                      " of the rollout, after an earlier macro step: the particles hold the state "
                                      "before it, and plain substeps after a macro step are not enabled");
    return MPM_OK;
}

}  // extern "C"
