// mpm_unet.hip — kernels of the grid U-Net surrogate: BF16 matrix-core 3x3
// convolutions (forward, backward data, backward weights), pooling, upsampling,
// GELU, the projected velocity head with its loss, and Adam.
//
// Matrix-core fragment layout (v_mfma_f32_16x16x32_bf16, wave64), with r = lane % 16
// and g = lane / 16:
//   A (16 x 32, rows i, reduction k):  lane holds A[r][8g .. 8g+7]
//   B (32 x 16, reduction k, cols j):  lane holds B[8g .. 8g+7][r]
//   D (16 x 16, FP32):                 lane holds D[4g .. 4g+3][r]
// The convolutions put output channels on the rows and pixels on the columns, so
// that a lane's four results are consecutive channels of one NHWC pixel.
// A and B place the same reduction index in the same (lane group, element) slot,
// so the product does not depend on how the hardware orders the reduction.
#include <hip/hip_runtime.h>

#include "mpm_unet.hpp"

namespace mpm {
namespace unet {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ float gelu(float z)
{
    return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float z)
{
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * __expf(-0.5f * z * z);
    return cdf + z * pdf;
}

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (bf16)0.0f;
    return v;
}

// Eight channels [c, c+8) of pixel (b, y, x) of an input as a layer at resolution
// h x w sees it: stored at that resolution (mode 0), at half resolution and
// upsampled by nearest copy (mode 1), or at double resolution and 2x2-average
// pooled (mode 2, summed in FP32 in raster order and rounded once).
__device__ __forceinline__ bf16x8 load_in(const bf16* x, int cs, int c, int b, int y, int xx, int h, int w,
                                          int mode)
{
    if (mode == 0) return *reinterpret_cast<const bf16x8*>(x + ((size_t)(b * h + y) * w + xx) * cs + c);
    if (mode == 1) {
        const int hh = h >> 1, ww = w >> 1;
        return *reinterpret_cast<const bf16x8*>(x + ((size_t)(b * hh + (y >> 1)) * ww + (xx >> 1)) * cs + c);
    }
    const int h2 = 2 * h, w2 = 2 * w;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(
                x + ((size_t)(b * h2 + 2 * y + a) * w2 + 2 * xx + bb) * cs + c);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += (float)v[j];
        }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = (bf16)(0.25f * s[j]);
    return o;
}

// ---------------------------------------------------------------------------
// 3x3 convolution as implicit GEMM: rows = pixels, columns = output channels,
// reduction = (tap, 32-channel chunk).  Each wave owns a tile of 16*MT pixels x
// 16*NT output channels; a workgroup of 4 waves holds 4/KS tiles and splits each
// tile's reduction KS ways (waves s = 0..KS-1 take chunks s, s+KS, ...), adding the
// partial sums through LDS before the epilogue.  Small (low-resolution) layers use
// KS > 1 so that they still fill the chip.
template <int MT, int NT, int KS, bool SPLIT>
__global__ void __launch_bounds__(256) k_conv3x3(ConvArgs a, int ntiles_co)
{
    constexpr int TPW = 4 / KS;                     // tiles per workgroup
    constexpr int NACC = MT * NT * 4;               // accumulator floats per lane
    __shared__ float red[(KS > 1 ? (KS - 1) * TPW : 1) * NACC * 64];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int ks = wave % KS, tloc = wave / KS;
    const int tile = blockIdx.x * TPW + tloc;
    const int co_t = tile % ntiles_co, px_t = tile / ntiles_co;
    const int cin = a.c1 + a.c2;
    const int hw = a.h * a.w;
    const int npix = a.nb * hw;
    const int co0 = co_t * 16 * NT;
    const int pbase = px_t * 16 * MT;

    int pb[MT], py[MT], px[MT];
    bool pv[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int p = pbase + mt * 16 + r;
        pv[mt] = p < npix;
        const int pp = pv[mt] ? p : 0;
        pb[mt] = pp / hw;
        const int rem = pp - pb[mt] * hw;
        py[mt] = rem / a.w;
        px[mt] = rem - py[mt] * a.w;
    }

    f32x4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[mt][t] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int nchunk = cin >> 5;
    const int nk = 9 * nchunk;
    const bool tile_live = pbase < npix;
    // fragments of reduction step kk: A (MT pixel tiles) and B (NT channel tiles)
    auto load = [&](int kk, bf16x8* af, bf16x8* bfr) {
        const int tap = kk / nchunk;
        const int ci0 = (kk - tap * nchunk) << 5;
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const bool first = ci0 < a.c1;
        const bf16* xs = first ? a.x1 : a.x2;
        const int cs = first ? a.c1 : a.c2;
        const int cc = (first ? ci0 : ci0 - a.c1) + 8 * g;
        const int md = first ? a.mode1 : 0;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int yy = py[mt] + dy, xx = px[mt] + dx;
            const bool ok = pv[mt] && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
            af[mt] = ok ? load_in(xs, cs, cc, pb[mt], yy, xx, a.h, a.w, md) : zero8();
        }
        const bf16* wrow = a.wt + ((size_t)tap * a.cout + co0 + r) * cin + ci0 + 8 * g;
#pragma unroll
        for (int t = 0; t < NT; ++t) bfr[t] = *reinterpret_cast<const bf16x8*>(wrow + (size_t)16 * t * cin);
    };
    if (tile_live && ks < nk) {
        // a ring of DEPTH reduction steps in registers: the loads of step i + DEPTH
        // are issued as soon as step i has gone to the matrix cores, so DEPTH load
        // round trips overlap
        constexpr int DEPTH = 4;
        const int nsteps = (nk - ks + KS - 1) / KS;
        bf16x8 fa[DEPTH][MT], fb[DEPTH][NT];
#pragma unroll
        for (int j = 0; j < DEPTH; ++j)
            if (j < nsteps) load(ks + j * KS, fa[j], fb[j]);
        for (int i = 0; i < nsteps; i += DEPTH) {
#pragma unroll
            for (int j = 0; j < DEPTH; ++j) {
                if (i + j < nsteps) {
                    // weights as the A operand (rows = output channels), pixels as B:
                    // each lane ends up with 4 consecutive channels of one pixel
#pragma unroll
                    for (int t = 0; t < NT; ++t)
#pragma unroll
                        for (int mt = 0; mt < MT; ++mt)
                            acc[mt][t] =
                                __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[j][t], fa[j][mt], acc[mt][t], 0, 0, 0);
                    if (i + j + DEPTH < nsteps) load(ks + (i + j + DEPTH) * KS, fa[j], fb[j]);
                }
            }
        }
    }

    if constexpr (KS > 1) {
        if (ks > 0) {
            float* dst = red + ((tloc * (KS - 1) + ks - 1) * NACC) * 64 + lane;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int e = 0; e < 4; ++e) dst[((mt * NT + t) * 4 + e) * 64] = acc[mt][t][e];
        }
        __syncthreads();
        if (ks > 0) return;
#pragma unroll
        for (int j = 0; j < KS - 1; ++j) {
            const float* src = red + ((tloc * (KS - 1) + j) * NACC) * 64 + lane;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc[mt][t][e] += src[((mt * NT + t) * 4 + e) * 64];
        }
    }
    if (!tile_live) return;

    // D[4g+e][r] of tile (mt, t): output channel co0 + 16t + 4g + e of pixel
    // pbase + 16mt + r; four channels go out in one 8-byte store
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int p = pbase + mt * 16 + r;
        if (p >= npix) continue;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int cb = co0 + 16 * t + 4 * g;
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = acc[mt][t][e] + ((!SPLIT && a.bias) ? a.bias[cb + e] : 0.f);
            if constexpr (!SPLIT) {
                bf16x4 zo, ao;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    zo[e] = (bf16)v[e];
                    ao[e] = (bf16)gelu(v[e]);
                }
                if (a.z) *reinterpret_cast<bf16x4*>(a.z + (size_t)p * a.cout + cb) = zo;
                if (a.a) *reinterpret_cast<bf16x4*>(a.a + (size_t)p * a.cout + cb) = ao;
            } else {
                // split is a multiple of 32: the four channels lie on one side
                const bool lo = cb < a.split;
                const int stride = lo ? a.split : a.cout - a.split;
                bf16* d = lo ? a.o1 + (size_t)p * stride + cb : a.o2 + (size_t)p * stride + (cb - a.split);
                const int add = lo ? a.add1 : a.add2;
                bf16x4 o;
                if (add) {
                    const bf16x4 old = *reinterpret_cast<const bf16x4*>(d);
#pragma unroll
                    for (int e = 0; e < 4; ++e) o[e] = (bf16)((float)old[e] + v[e]);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) o[e] = (bf16)v[e];
                }
                *reinterpret_cast<bf16x4*>(d) = o;
            }
        }
    }
}

template <int MT, int NT, int KS, bool SPLIT>
hipError_t launch_conv_cfg(const ConvArgs& a, int npix, hipStream_t st)
{
    const int ntiles_co = a.cout / (16 * NT);
    const int ntiles = ((npix + 16 * MT - 1) / (16 * MT)) * ntiles_co;
    constexpr int TPW = 4 / KS;
    k_conv3x3<MT, NT, KS, SPLIT><<<(ntiles + TPW - 1) / TPW, 256, 0, st>>>(a, ntiles_co);
    return hipGetLastError();
}

// Tile shape and reduction split by layer shape: 16*NT output channels per wave
// (NT = 4 when the channel count allows, else 2), the largest of 64/32/16 pixels
// per wave that still gives >= 1024 waves, and the reduction split until about
// 2048 waves run.
template <bool SPLIT>
hipError_t launch_conv(const ConvArgs& a, hipStream_t st)
{
    const int npix = a.nb * a.h * a.w;
    if (npix == 0) return hipSuccess;
    const int nt = (a.cout % 64 == 0) ? 4 : 2;
    const int co_tiles = a.cout / (16 * nt);
    auto waves_at = [&](int m) { return ((npix + 16 * m - 1) / (16 * m)) * co_tiles; };
    const int mt = waves_at(4) >= 1024 ? 4 : waves_at(2) >= 1024 ? 2 : 1;
    const int tiles = waves_at(mt);
    const int chunks = 9 * ((a.c1 + a.c2) / 32);
    int ks = 1;
    while (ks < 4 && tiles * ks < 2048 && chunks >= 4 * ks) ks *= 2;
#define CONV_CASE(M, N, K)                                        \
    if (mt == M && nt == N && ks == K) return launch_conv_cfg<M, N, K, SPLIT>(a, npix, st);
    CONV_CASE(1, 2, 1) CONV_CASE(1, 2, 2) CONV_CASE(1, 2, 4)
    CONV_CASE(1, 4, 1) CONV_CASE(1, 4, 2) CONV_CASE(1, 4, 4)
    CONV_CASE(2, 2, 1) CONV_CASE(2, 2, 2) CONV_CASE(2, 2, 4)
    CONV_CASE(2, 4, 1) CONV_CASE(2, 4, 2) CONV_CASE(2, 4, 4)
    CONV_CASE(4, 2, 1) CONV_CASE(4, 2, 2) CONV_CASE(4, 2, 4)
    CONV_CASE(4, 4, 1) CONV_CASE(4, 4, 2) CONV_CASE(4, 4, 4)
#undef CONV_CASE
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// LDS-tiled forward 3x3 convolution.  A workgroup computes TH x 16 pixels (rows x
// columns) of one image for 16*NT output channels.  Per 32-channel input chunk it
// stages the (TH+2) x 18 pixel patch (tile and one-pixel halo, zero outside the
// image) and the 9 x 16*NT x 32 weights in LDS once, then each wave runs all nine
// taps over its TH/4 rows from LDS: every staged value is used by up to nine taps
// (input) or all TH/4 x 16 pixels of the wave (weights) instead of being fetched
// again per tap.  Operands: weights as A (rows = output channels), pixels as B, as
// in k_conv3x3; a pixel's or weight row's 32 channels sit at a pitch of 40 BF16
// (80 bytes) so that the 16-byte fragment reads of 16 neighbouring rows spread
// over the LDS banks.
template <int TH, int NT>
__global__ void __launch_bounds__(256) k_conv3x3_lds(ConvArgs a, int tiles_x, int tiles_y, int co_blocks)
{
    constexpr int TW = 16, PW = TW + 2, PH = TH + 2;
    constexpr int MT = TH / 4;           // rows per wave
    constexpr int CP = 40;               // LDS pitch (bf16) of a pixel's / weight row's 32 channels
    constexpr int COW = 16 * NT;         // output channels per workgroup
    __shared__ __attribute__((aligned(16))) bf16 s_in[PH * PW * CP];
    __shared__ __attribute__((aligned(16))) bf16 s_w[9 * COW * CP];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    int bid = blockIdx.x;
    const int cblk = bid % co_blocks;
    bid /= co_blocks;
    const int tx = bid % tiles_x;
    bid /= tiles_x;
    const int ty = bid % tiles_y;
    const int b = bid / tiles_y;
    const int y0 = ty * TH, x0 = tx * TW, co0 = cblk * COW;
    const int cin = a.c1 + a.c2;
    const int nchunk = cin >> 5;

    f32x4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[mt][t] = f32x4{0.f, 0.f, 0.f, 0.f};

    // chunk staging through registers: the global loads of chunk ch + 1 are in
    // flight while the waves run chunk ch from LDS
    constexpr int NV_IN = (PH * PW * 4 + 255) / 256, NV_W = (9 * COW * 4 + 255) / 256;
    bf16x8 rin[NV_IN], rw[NV_W];
    auto fetch = [&](int ch) {
        const int ci0 = ch << 5;
        const bool first = ci0 < a.c1;
        const bf16* xs = first ? a.x1 : a.x2;
        const int cs = first ? a.c1 : a.c2;
        const int cc = first ? ci0 : ci0 - a.c1;
        const int md = first ? a.mode1 : 0;
#pragma unroll
        for (int i = 0; i < NV_IN; ++i) {
            const int v = threadIdx.x + 256 * i;
            const int pix = v >> 2, q = v & 3;
            const int py = pix / PW, px = pix - py * PW;
            const int yy = y0 + py - 1, xx = x0 + px - 1;
            rin[i] = zero8();
            if (v < PH * PW * 4 && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w)
                rin[i] = load_in(xs, cs, cc + 8 * q, b, yy, xx, a.h, a.w, md);
        }
#pragma unroll
        for (int i = 0; i < NV_W; ++i) {
            const int v = threadIdx.x + 256 * i;
            const int row = v >> 2, q = v & 3;   // row = tap * COW + channel
            const int tap = row / COW, co = row - tap * COW;
            rw[i] = zero8();
            if (v < 9 * COW * 4)
                rw[i] = *reinterpret_cast<const bf16x8*>(a.wt + ((size_t)tap * a.cout + co0 + co) * cin + ci0 + 8 * q);
        }
    };
    fetch(0);
    for (int ch = 0; ch < nchunk; ++ch) {
#pragma unroll
        for (int i = 0; i < NV_IN; ++i) {
            const int v = threadIdx.x + 256 * i;
            if (v < PH * PW * 4) *reinterpret_cast<bf16x8*>(&s_in[(v >> 2) * CP + 8 * (v & 3)]) = rin[i];
        }
#pragma unroll
        for (int i = 0; i < NV_W; ++i) {
            const int v = threadIdx.x + 256 * i;
            if (v < 9 * COW * 4) *reinterpret_cast<bf16x8*>(&s_w[(v >> 2) * CP + 8 * (v & 3)]) = rw[i];
        }
        __syncthreads();
        if (ch + 1 < nchunk) fetch(ch + 1);
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;   // patch offsets of the tap (0..2)
            bf16x8 fw[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t)
                fw[t] = *reinterpret_cast<const bf16x8*>(&s_w[(tap * COW + 16 * t + r) * CP + 8 * g]);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int prow = wave * MT + mt + dy;
                const bf16x8 fx = *reinterpret_cast<const bf16x8*>(&s_in[(prow * PW + r + dx) * CP + 8 * g]);
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw[t], fx, acc[mt][t], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // D[4g+e][r] of (mt, t): output channel co0 + 16t + 4g + e of pixel
    // (y0 + wave*MT + mt, x0 + r); four channels go out in one 8-byte store
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int y = y0 + wave * MT + mt, x = x0 + r;
        if (y >= a.h || x >= a.w) continue;
        const size_t p = ((size_t)b * a.h + y) * a.w + x;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int cb = co0 + 16 * t + 4 * g;
            bf16x4 zo, ao;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float v = acc[mt][t][e] + (a.bias ? a.bias[cb + e] : 0.f);
                zo[e] = (bf16)v;
                ao[e] = (bf16)gelu(v);
            }
            if (a.z) *reinterpret_cast<bf16x4*>(a.z + p * a.cout + cb) = zo;
            if (a.a) *reinterpret_cast<bf16x4*>(a.a + p * a.cout + cb) = ao;
        }
    }
}

template <int TH>
hipError_t launch_conv_lds(const ConvArgs& a, hipStream_t st)
{
    const int tiles_x = (a.w + 15) / 16, tiles_y = (a.h + TH - 1) / TH, co_blocks = a.cout / 32;
    const long long nwg = (long long)a.nb * tiles_y * tiles_x * co_blocks;
    k_conv3x3_lds<TH, 2><<<(unsigned)nwg, 256, 0, st>>>(a, tiles_x, tiles_y, co_blocks);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Weight gradient: per tap a GEMM dW[co][ci] = sum_p dz[p][co] x[p+tap][ci].
// Workgroup: one tap, a 32x32 (co, ci) tile, a range of pixels; each wave takes
// 32-pixel slabs, transposes them through its own LDS region ([channel][pixel])
// and accumulates the 2x2 16x16 output tiles; the four waves reduce through LDS
// and the workgroup adds its tile into dW with FP32 atomics.
constexpr int WG_PIX = 2048;   // pixels per workgroup (split of the reduction)
constexpr int WG_LDS = 40;     // LDS row stride (bf16) of a transposed slab

__global__ void __launch_bounds__(256) k_conv3x3_wgrad(const bf16* __restrict__ x1, const bf16* __restrict__ x2,
                                                      int c1, int c2, int mode1, const bf16* __restrict__ dz,
                                                      int cout, int nb, int h, int w, float* __restrict__ dw)
{
    __shared__ __attribute__((aligned(16))) bf16 tz[4][32 * WG_LDS];
    __shared__ __attribute__((aligned(16))) bf16 tx[4][32 * WG_LDS];
    __shared__ float red[4][32 * 33];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int cin = c1 + c2;
    const int ntile_co = cout / 32;
    const int co_t = blockIdx.x % ntile_co, ci_t = blockIdx.x / ntile_co;
    const int tap = blockIdx.y;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int hw = h * w, npix = nb * hw;
    const int p_begin = blockIdx.z * WG_PIX;

    const int ci_abs = ci_t * 32;
    const bool first = ci_abs < c1;
    const bf16* xs = first ? x1 : x2;
    const int cs = first ? c1 : c2;
    const int cc = first ? ci_abs : ci_abs - c1;

    // staging: lane loads pixel (lane >> 1), channels 16 * (lane & 1) .. +15
    const int sp = lane >> 1, sc = (lane & 1) * 16;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int s = p_begin + wave * 32; s < p_begin + WG_PIX; s += 128) {
        const int p = s + sp;
        bf16x8 z0 = zero8(), z1 = zero8(), v0 = zero8(), v1 = zero8();
        if (p < npix) {
            const bf16* zr = dz + (size_t)p * cout + co_t * 32 + sc;
            z0 = *reinterpret_cast<const bf16x8*>(zr);
            z1 = *reinterpret_cast<const bf16x8*>(zr + 8);
            const int b = p / hw, rem = p - b * hw;
            const int yy = rem / w + dy, xx = rem % w + dx;
            if (yy >= 0 && yy < h && xx >= 0 && xx < w) {
                const int md = first ? mode1 : 0;
                v0 = load_in(xs, cs, cc + sc, b, yy, xx, h, w, md);
                v1 = load_in(xs, cs, cc + sc + 8, b, yy, xx, h, w, md);
            }
        }
        __syncthreads();
        bf16* mz = tz[wave];
        bf16* mx = tx[wave];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            mz[(sc + j) * WG_LDS + sp] = z0[j];
            mz[(sc + 8 + j) * WG_LDS + sp] = z1[j];
            mx[(sc + j) * WG_LDS + sp] = v0[j];
            mx[(sc + 8 + j) * WG_LDS + sp] = v1[j];
        }
        __syncthreads();
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            af[i] = *reinterpret_cast<const bf16x8*>(mz + (16 * i + r) * WG_LDS + 8 * g);
            bfr[i] = *reinterpret_cast<const bf16x8*>(mx + (16 * i + r) * WG_LDS + 8 * g);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }

    // D[4g+e][r] of tile (i, j): co = 16i + 4g + e, ci = 16j + r
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) red[wave][(16 * i + 4 * g + e) * 33 + 16 * j + r] = acc[i][j][e];
    __syncthreads();
    for (int o = threadIdx.x; o < 1024; o += 256) {
        const int co = o >> 5, ci = o & 31;
        const float v = red[0][co * 33 + ci] + red[1][co * 33 + ci] + red[2][co * 33 + ci] + red[3][co * 33 + ci];
        atomicAdd(dw + ((size_t)tap * cout + co_t * 32 + co) * cin + ci_abs + ci, v);
    }
}

__global__ void __launch_bounds__(256) k_bias_grad(const bf16* __restrict__ dz, int npix, int cout,
                                                  float* __restrict__ db)
{
    // thread: 8 consecutive channels (one 16-byte load per pixel) of every
    // (256 / (cout/8))-th pixel of the workgroup's chunks; partial sums through LDS
    __shared__ float part[256 * 8];
    const int ngrp = cout >> 3, lanes = 256 / ngrp;
    const int cg = threadIdx.x % ngrp, k = threadIdx.x / ngrp;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (k < lanes)
        for (int p0 = blockIdx.x * 1024; p0 < npix; p0 += gridDim.x * 1024)
            for (int p = p0 + k; p < min(p0 + 1024, npix); p += lanes) {
                const bf16x8 v = *reinterpret_cast<const bf16x8*>(dz + (size_t)p * cout + 8 * cg);
#pragma unroll
                for (int j = 0; j < 8; ++j) s[j] += (float)v[j];
            }
#pragma unroll
    for (int j = 0; j < 8; ++j) part[threadIdx.x * 8 + j] = s[j];
    __syncthreads();
    if (threadIdx.x < cout) {
        const int g = threadIdx.x >> 3, j = threadIdx.x & 7;
        float t = 0.f;
        for (int kk = 0; kk < lanes; ++kk) t += part[(kk * ngrp + g) * 8 + j];
        atomicAdd(db + threadIdx.x, t);
    }
}

__global__ void __launch_bounds__(256) k_cast_weights(const float* __restrict__ w, int cout, int cin,
                                                     bf16* __restrict__ wf, bf16* __restrict__ wb)
{
    const int64_t n = (int64_t)9 * cout * cin;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int tap = (int)(i / ((int64_t)cout * cin));
        const int rem = (int)(i - (int64_t)tap * cout * cin);
        const int co = rem / cin, ci = rem - co * cin;
        const bf16 v = (bf16)w[i];
        wf[i] = v;
        wb[((size_t)(8 - tap) * cin + ci) * cout + co] = v;
    }
}

// ---------------------------------------------------------------------------
// Pooling / upsampling (8 channels per thread; c is a multiple of 8)
__global__ void __launch_bounds__(256) k_avgpool2(const bf16* __restrict__ x, int nb, int h, int w, int c,
                                                 bf16* __restrict__ y)
{
    const int ho = h / 2, wo = w / 2, c8 = c / 8;
    const int64_t n = (int64_t)nb * ho * wo * c8;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int k = (int)(i % c8);
        int64_t t = i / c8;
        const int xo = (int)(t % wo);
        t /= wo;
        const int yo = (int)(t % ho);
        const int b = (int)(t / ho);
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int a = 0; a < 2; ++a)
            for (int bb = 0; bb < 2; ++bb) {
                const bf16x8 v = *reinterpret_cast<const bf16x8*>(
                    x + ((size_t)(b * h + 2 * yo + a) * w + 2 * xo + bb) * c + 8 * k);
#pragma unroll
                for (int j = 0; j < 8; ++j) s[j] += (float)v[j];
            }
        bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = (bf16)(0.25f * s[j]);
        *reinterpret_cast<bf16x8*>(y + ((size_t)(b * ho + yo) * wo + xo) * c + 8 * k) = o;
    }
}

// dx (h x w) from dy (h/2 x w/2): dx = dy / 4 (stored or added)
__global__ void __launch_bounds__(256) k_avgpool2_bwd(const bf16* __restrict__ dy, int nb, int h, int w, int c,
                                                     bf16* __restrict__ dx, int add)
{
    const int ho = h / 2, wo = w / 2, c8 = c / 8;
    const int64_t n = (int64_t)nb * h * w * c8;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int k = (int)(i % c8);
        int64_t t = i / c8;
        const int x = (int)(t % w);
        t /= w;
        const int y = (int)(t % h);
        const int b = (int)(t / h);
        const bf16x8 g = *reinterpret_cast<const bf16x8*>(dy + ((size_t)(b * ho + y / 2) * wo + x / 2) * c + 8 * k);
        bf16x8* d = reinterpret_cast<bf16x8*>(dx + ((size_t)(b * h + y) * w + x) * c + 8 * k);
        bf16x8 o;
        if (add) {
            const bf16x8 old = *d;
#pragma unroll
            for (int j = 0; j < 8; ++j) o[j] = (bf16)((float)old[j] + 0.25f * (float)g[j]);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) o[j] = (bf16)(0.25f * (float)g[j]);
        }
        *d = o;
    }
}

// y (2h x 2w) = nearest upsample of x (h x w)
__global__ void __launch_bounds__(256) k_upsample2(const bf16* __restrict__ x, int nb, int h, int w, int c,
                                                  bf16* __restrict__ y)
{
    const int h2 = 2 * h, w2 = 2 * w, c8 = c / 8;
    const int64_t n = (int64_t)nb * h2 * w2 * c8;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int k = (int)(i % c8);
        int64_t t = i / c8;
        const int xx = (int)(t % w2);
        t /= w2;
        const int yy = (int)(t % h2);
        const int b = (int)(t / h2);
        *reinterpret_cast<bf16x8*>(y + ((size_t)(b * h2 + yy) * w2 + xx) * c + 8 * k) =
            *reinterpret_cast<const bf16x8*>(x + ((size_t)(b * h + yy / 2) * w + xx / 2) * c + 8 * k);
    }
}

// dx (h x w) = sum of the four dy (2h x 2w) it was copied to
__global__ void __launch_bounds__(256) k_upsample2_bwd(const bf16* __restrict__ dy, int nb, int h, int w, int c,
                                                      bf16* __restrict__ dx)
{
    const int h2 = 2 * h, w2 = 2 * w, c8 = c / 8;
    const int64_t n = (int64_t)nb * h * w * c8;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int k = (int)(i % c8);
        int64_t t = i / c8;
        const int x = (int)(t % w);
        t /= w;
        const int y = (int)(t % h);
        const int b = (int)(t / h);
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int a = 0; a < 2; ++a)
            for (int bb = 0; bb < 2; ++bb) {
                const bf16x8 v = *reinterpret_cast<const bf16x8*>(
                    dy + ((size_t)(b * h2 + 2 * y + a) * w2 + 2 * x + bb) * c + 8 * k);
#pragma unroll
                for (int j = 0; j < 8; ++j) s[j] += (float)v[j];
            }
        bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = (bf16)s[j];
        *reinterpret_cast<bf16x8*>(dx + ((size_t)(b * h + y) * w + x) * c + 8 * k) = o;
    }
}

__global__ void __launch_bounds__(256) k_gelu_bwd(const bf16* __restrict__ da, const bf16* __restrict__ z,
                                                 int64_t n, bf16* __restrict__ dz)
{
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        dz[i] = (bf16)((float)da[i] * gelu_grad((float)z[i]));
}

// ---------------------------------------------------------------------------
// Head (32 -> 2 channels, FP32) and the per-image momentum projection.
constexpr int HEAD_C = 32;
constexpr int HEAD_PART = 2 * HEAD_C + 2;   // head gradient values per partial row

__device__ __forceinline__ float block_sum(float v, float* sh)
{
    // 256 threads: wave reduction then across the four waves
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
}

// Per-block partial sums part[image][block][2] of wm_i out_i (no atomics: the
// projection kernel adds the partials of its image itself).
constexpr int HEAD_BLOCKS = 64;

__global__ void __launch_bounds__(256) k_head_forward(const bf16* __restrict__ a, int nb, int hw,
                                                     const float* __restrict__ wh, const float* __restrict__ bh,
                                                     const float* __restrict__ wm, float* __restrict__ out,
                                                     float* __restrict__ part)
{
    __shared__ float sh[4];
    // blockIdx.y = image; pixels of the image grid-strided over blockIdx.x
    const int b = blockIdx.y;
    float s0 = 0.f, s1 = 0.f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < hw; i += gridDim.x * 256) {
        const size_t p = (size_t)b * hw + i;
        const bf16* ar = a + p * HEAD_C;
        float o0 = bh[0], o1 = bh[1];
#pragma unroll
        for (int k = 0; k < HEAD_C / 8; ++k) {
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(ar + 8 * k);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                o0 += (float)v[j] * wh[8 * k + j];
                o1 += (float)v[j] * wh[HEAD_C + 8 * k + j];
            }
        }
        out[2 * p] = o0;
        out[2 * p + 1] = o1;
        s0 += wm[p] * o0;
        s1 += wm[p] * o1;
    }
    const float t0 = block_sum(s0, sh);
    const float t1 = block_sum(s1, sh);
    if (threadIdx.x == 0) {
        part[((size_t)b * HEAD_BLOCKS + blockIdx.x) * 2] = t0;
        part[((size_t)b * HEAD_BLOCKS + blockIdx.x) * 2 + 1] = t1;
    }
}

// out' = out - (sum of the image's partials) in place; with a target, the loss and
// dL/dout' with its per-block partial sums part2 [image][block][2].
__global__ void __launch_bounds__(256) k_head_loss(float* __restrict__ out, const float* __restrict__ target,
                                                  const float* __restrict__ wm, int nb, int hw,
                                                  const float* __restrict__ part, int nparts,
                                                  float* __restrict__ loss, float* __restrict__ dproj,
                                                  float* __restrict__ part2)
{
    __shared__ float sh[4];
    const int b = blockIdx.y;
    float m0 = 0.f, m1 = 0.f;
    for (int j = 0; j < nparts; ++j) {
        m0 += part[((size_t)b * HEAD_BLOCKS + j) * 2];
        m1 += part[((size_t)b * HEAD_BLOCKS + j) * 2 + 1];
    }
    const float inv_nb = 1.0f / (float)nb;
    float l = 0.f, g0 = 0.f, g1 = 0.f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < hw; i += gridDim.x * 256) {
        const size_t p = (size_t)b * hw + i;
        const float o0 = out[2 * p] - m0, o1 = out[2 * p + 1] - m1;
        out[2 * p] = o0;
        out[2 * p + 1] = o1;
        if (target) {
            const float d0 = o0 - target[2 * p], d1 = o1 - target[2 * p + 1];
            const float wmi = wm[p];
            l += wmi * (d0 * d0 + d1 * d1);
            if (dproj) {
                const float q0 = 2.f * wmi * d0 * inv_nb, q1 = 2.f * wmi * d1 * inv_nb;
                dproj[2 * p] = q0;
                dproj[2 * p + 1] = q1;
                g0 += q0;
                g1 += q1;
            }
        }
    }
    if (!target) return;
    const float tl = block_sum(l, sh);
    const float t0 = block_sum(g0, sh);
    const float t1 = block_sum(g1, sh);
    if (threadIdx.x == 0) {
        atomicAdd(loss, tl * inv_nb);
        if (dproj) {
            part2[((size_t)b * HEAD_BLOCKS + blockIdx.x) * 2] = t0;
            part2[((size_t)b * HEAD_BLOCKS + blockIdx.x) * 2 + 1] = t1;
        }
    }
}

__global__ void __launch_bounds__(256) k_head_backward(const bf16* __restrict__ a, const float* __restrict__ dproj,
                                                      const float* __restrict__ part2, int nparts,
                                                      const float* __restrict__ wm, int nb, int hw,
                                                      const float* __restrict__ wh, bf16* __restrict__ da,
                                                      float* __restrict__ hpart)
{
    __shared__ float red[4][HEAD_PART];
    const int b = blockIdx.y;
    float t0 = 0.f, t1 = 0.f;
    for (int j = 0; j < nparts; ++j) {
        t0 += part2[((size_t)b * HEAD_BLOCKS + j) * 2];
        t1 += part2[((size_t)b * HEAD_BLOCKS + j) * 2 + 1];
    }
    float gw[2 * HEAD_C];
#pragma unroll
    for (int k = 0; k < 2 * HEAD_C; ++k) gw[k] = 0.f;
    float gb0 = 0.f, gb1 = 0.f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < hw; i += gridDim.x * 256) {
        const size_t p = (size_t)b * hw + i;
        // projection adjoint: dout_j = dproj_j - wm_j * sum_i dproj_i
        const float d0 = dproj[2 * p] - wm[p] * t0, d1 = dproj[2 * p + 1] - wm[p] * t1;
        gb0 += d0;
        gb1 += d1;
        const bf16* ar = a + p * HEAD_C;
        bf16* dr = da + p * HEAD_C;
#pragma unroll
        for (int k = 0; k < HEAD_C / 8; ++k) {
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(ar + 8 * k);
            bf16x8 o;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int c = 8 * k + j;
                o[j] = (bf16)(d0 * wh[c] + d1 * wh[HEAD_C + c]);
                gw[c] += d0 * (float)v[j];
                gw[HEAD_C + c] += d1 * (float)v[j];
            }
            *reinterpret_cast<bf16x8*>(dr + 8 * k) = o;
        }
    }
    // the workgroup's sums of the 2*HEAD_C weight and 2 bias gradients: per wave by
    // shuffles, across the four waves through LDS, then to this workgroup's row of
    // hpart (reduced by k_head_reduce: no same-address global atomics)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < HEAD_PART; ++k) {
        float v = k < 2 * HEAD_C ? gw[k] : (k == 2 * HEAD_C ? gb0 : gb1);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < HEAD_PART)
        hpart[((size_t)b * gridDim.x + blockIdx.x) * HEAD_PART + threadIdx.x] =
            red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// dwh (2*HEAD_C values) and dbh (2) accumulate the sum of the nrows partial rows:
// workgroup k sums column k.
__global__ void __launch_bounds__(256) k_head_reduce(const float* __restrict__ hpart, int nrows,
                                                    float* __restrict__ dwh, float* __restrict__ dbh)
{
    __shared__ float sh[4];
    const int k = blockIdx.x;
    float v = 0.f;
    for (int r = threadIdx.x; r < nrows; r += 256) v += hpart[(size_t)r * HEAD_PART + k];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) {
        if (k < 2 * HEAD_C) dwh[k] += v;
        else dbh[k - 2 * HEAD_C] += v;
    }
}

__global__ void __launch_bounds__(256) k_adam(float* __restrict__ w, const float* __restrict__ g,
                                             float* __restrict__ m, float* __restrict__ v, int64_t n, float lr_t,
                                             float beta1, float beta2, float eps_t)
{
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float gi = g[i];
        const float mi = beta1 * m[i] + (1.f - beta1) * gi;
        const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr_t * mi / (sqrtf(vi) + eps_t);
    }
}

__global__ void __launch_bounds__(256) k_fill_f32(float* p, int64_t n, float value)
{
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = value;
}

__global__ void __launch_bounds__(256) k_fill_bf16(bf16* p, int64_t n)
{
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = (bf16)0.0f;
}

inline unsigned grid_for(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

inline unsigned head_blocks(int hw)
{
    const int b = (hw + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > HEAD_BLOCKS ? HEAD_BLOCKS : b));
}

}  // namespace

// Forward convolutions: the LDS-tiled kernel with the tallest tile (most reuse of
// the staged weights) that still gives >= 1024 workgroups (four per CU, so that
// their staging and matrix work overlap), 4-row tiles for smaller layers down to
// 64 workgroups (measured faster than the register-streaming kernel there), and
// the register-streaming kernel with its split reduction below that.
hipError_t conv3x3_forward(const ConvArgs& a, hipStream_t st, int kernel)
{
    if (a.nb * a.h * a.w == 0) return hipSuccess;
    auto wgs = [&](int th) { return (long long)a.nb * ((a.h + th - 1) / th) * ((a.w + 15) / 16) * (a.cout / 32); };
    if (kernel != 1 && a.cout % 32 == 0 && (a.c1 + a.c2) % 32 == 0) {
        if (wgs(16) >= 1024) return launch_conv_lds<16>(a, st);
        if (wgs(8) >= 1024) return launch_conv_lds<8>(a, st);
        if (wgs(4) >= 64 || kernel == 2) return launch_conv_lds<4>(a, st);
    }
    return launch_conv<false>(a, st);
}
hipError_t conv3x3_split(const ConvArgs& a, hipStream_t st) { return launch_conv<true>(a, st); }

hipError_t conv3x3_wgrad(const bf16* x1, const bf16* x2, int c1, int c2, int mode1, const bf16* dz, int cout,
                         int nb, int h, int w, float* dw, hipStream_t st)
{
    const int npix = nb * h * w;
    if (npix == 0) return hipSuccess;
    dim3 grid((cout / 32) * ((c1 + c2) / 32), 9, (npix + WG_PIX - 1) / WG_PIX), blk(256);
    k_conv3x3_wgrad<<<grid, blk, 0, st>>>(x1, x2, c1, c2, mode1, dz, cout, nb, h, w, dw);
    return hipGetLastError();
}

hipError_t bias_grad(const bf16* dz, int npix, int cout, float* db, hipStream_t st)
{
    if (npix == 0) return hipSuccess;
    k_bias_grad<<<std::min((npix + 1023) / 1024, 256), 256, 0, st>>>(dz, npix, cout, db);
    return hipGetLastError();
}

hipError_t cast_weights(const float* w, int cout, int cin, bf16* wf, bf16* wb, hipStream_t st)
{
    k_cast_weights<<<grid_for((int64_t)9 * cout * cin), 256, 0, st>>>(w, cout, cin, wf, wb);
    return hipGetLastError();
}

hipError_t avgpool2(const bf16* x, int nb, int h, int w, int c, bf16* y, hipStream_t st)
{
    k_avgpool2<<<grid_for((int64_t)nb * (h / 2) * (w / 2) * (c / 8)), 256, 0, st>>>(x, nb, h, w, c, y);
    return hipGetLastError();
}

hipError_t avgpool2_bwd(const bf16* dy, int nb, int h, int w, int c, bf16* dx, int add, hipStream_t st)
{
    k_avgpool2_bwd<<<grid_for((int64_t)nb * h * w * (c / 8)), 256, 0, st>>>(dy, nb, h, w, c, dx, add);
    return hipGetLastError();
}

hipError_t upsample2(const bf16* x, int nb, int h, int w, int c, bf16* y, hipStream_t st)
{
    k_upsample2<<<grid_for((int64_t)nb * 4 * h * w * (c / 8)), 256, 0, st>>>(x, nb, h, w, c, y);
    return hipGetLastError();
}

hipError_t upsample2_bwd(const bf16* dy, int nb, int h, int w, int c, bf16* dx, hipStream_t st)
{
    k_upsample2_bwd<<<grid_for((int64_t)nb * h * w * (c / 8)), 256, 0, st>>>(dy, nb, h, w, c, dx);
    return hipGetLastError();
}

hipError_t gelu_bwd(const bf16* da, const bf16* z, int64_t n, bf16* dz, hipStream_t st)
{
    k_gelu_bwd<<<grid_for(n), 256, 0, st>>>(da, z, n, dz);
    return hipGetLastError();
}

hipError_t head_forward(const bf16* a, int nb, int hw, const float* wh, const float* bh, const float* wm,
                        float* out, float* part, hipStream_t st)
{
    dim3 grid(head_blocks(hw), nb);
    k_head_forward<<<grid, 256, 0, st>>>(a, nb, hw, wh, bh, wm, out, part);
    return hipGetLastError();
}

hipError_t head_project_loss(float* out, const float* part, const float* target, const float* wm, int nb,
                             int hw, float* loss, float* dproj, float* part2, hipStream_t st)
{
    dim3 grid(head_blocks(hw), nb);
    k_head_loss<<<grid, 256, 0, st>>>(out, target, wm, nb, hw, part, part ? (int)head_blocks(hw) : 0, loss, dproj,
                                      part2);
    return hipGetLastError();
}

hipError_t head_backward(const bf16* a, const float* dproj, const float* part2, const float* wm, int nb,
                         int hw, const float* wh, bf16* da, float* dwh, float* dbh, float* hpart, hipStream_t st)
{
    dim3 grid(head_blocks(hw), nb);
    k_head_backward<<<grid, 256, 0, st>>>(a, dproj, part2, (int)head_blocks(hw), wm, nb, hw, wh, da, hpart);
    k_head_reduce<<<HEAD_PART, 256, 0, st>>>(hpart, (int)(head_blocks(hw) * nb), dwh, dbh);
    return hipGetLastError();
}

hipError_t adam_step(float* w, const float* g, float* m, float* v, int64_t n, float lr_t, float beta1,
                     float beta2, float eps_t, hipStream_t st)
{
    k_adam<<<grid_for(n), 256, 0, st>>>(w, g, m, v, n, lr_t, beta1, beta2, eps_t);
    return hipGetLastError();
}

hipError_t fill_f32(float* p, int64_t n, float value, hipStream_t st)
{
    k_fill_f32<<<grid_for(n), 256, 0, st>>>(p, n, value);
    return hipGetLastError();
}

hipError_t fill_bf16(bf16* p, int64_t n, hipStream_t st)
{
    k_fill_bf16<<<grid_for(n), 256, 0, st>>>(p, n);
    return hipGetLastError();
}

}  // namespace unet
}  // namespace mpm
