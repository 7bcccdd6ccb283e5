// mpm_fused_kernel.hpp — the tiled, LDS-staged particle kernel k_tiled (template),
// shared by mpm_fused.hip and mpm_fused_cubic.hip (the 3D cubic Coulomb-wall
// instantiation, compiled with a higher full-unroll limit).
#pragma once
// The tiled, LDS-staged particle kernels.
//
// k_tiled<..., G2P=true>  is one substep's particle work in a single pass:
//   G2P from grid `src` (shaders/{2d,3d}/g2p.glsl) — gather, MLS C, position, F_E
//   update with the clamped SVD, plastic determinant — immediately followed by the
//   P2G of the updated particle into grid `dst` (shaders/{2d,3d}/p2g.glsl of the
//   next substep).  The stress uses the SVD G2P just computed: with
//   F_E = U diag(s) V^T and R = U V^T,  2 mu (F - R) F^T = U diag(2 mu (s - 1) s) U^T,
//   so each particle's state is read once, written once, and decomposed once per
//   substep.
// k_tiled<..., G2P=false> is P2G alone from the stored state (first substep).
//
// Work decomposition: one work item = the particles binned into one 64-node grid
// block (at most 512 of them, sorted by cell; at most two per thread, in as few
// wave passes as the count needs).  The workgroup stages the block's node tile
// plus a halo (stencil reach + a drift margin of MG cells for particles that moved
// since the last binning) in LDS: velocities for the gather and a zeroed
// accumulator for the scatter.  The tile is flushed to HBM with
// one global float atomic per touched node component.  A particle whose stencil
// leaves the tile (or whose state is not finite) takes the direct global path, so
// the result never depends on how stale the binning is.
//
// Tile accumulation is exact fixed point.  On gfx950 the LDS float atomic retires
// about a third of a lane per clock per CU, the integer LDS add about twelve
// (measured), so each contribution is scaled by a per-item power of two 2^e with
// |value| * 2^e <= 2^21 (e from the item's largest per-particle bound), rounded to
// an integer and summed exactly in LDS, two components per 64-bit add (pack_fixed;
// at most 512 contributions per node, so no half overflows): the tile sum is exact
// and independent of lane order, and the only rounding is the per-contribution
// quantisation (<= 2^-22 of the item's largest contribution bound) and the one
// conversion of the sum back to FP32.
#include <type_traits>

#include "mpm_kernels.hpp"

namespace mpm {

template <int DIM> struct TileShape {
    static constexpr int BS = BlockShape<DIM>::BS;
    static constexpr int MG = DIM == 3 ? 1 : 2;          // drift margin (cells)
    static constexpr int T = BS + 3 + 2 * MG;            // tile edge (nodes)
    static constexpr int TN = DIM == 3 ? T * T * T : T * T;
    static constexpr int NB = DIM == 3 ? 27 : 9;         // grid blocks a tile overlaps
};

// Fixed-point contributions travel as the bits of fma(a, b, 1.5 * 2^23): the
// fused multiply-add rounds once, to a float whose unit in the last place is 1, so
// for |a * b| < 2^22 its bits are 0x4B400000 + a * b rounded to the nearest integer
// (ties to even).  Two of them pack into one 64-bit word hi * 2^32 + lo (two's
// complement), whose sums are exact: the 64-bit adds carry the low half's sign into
// the high half, and both halves of a node's sum stay within 2^30.  Packing the
// biased bits and removing both biases with one 64-bit subtraction gives that word
// directly.
constexpr unsigned FIXED_BIAS = 0x4B400000u;

__device__ inline int fixed_bits(float a, float b) { return __float_as_int(fmaf(a, b, 12582912.0f)); }

// Workgroups reach the 8 XCDs in turn (workgroup b runs on XCD b mod 8), each XCD
// with its own L2, while consecutive work items share grid nodes (tile halos, the
// items of one super-block).  The item kernels therefore take items in rounds of
// gridDim.x (a multiple of 8 XCD_RUN) and give each XCD runs of XCD_RUN consecutive
// items, the runs dealt to the XCDs in turn so that dense and sparse regions still
// spread over all of them: this is the workgroup's item within a round (used by
// k_super).
constexpr int XCD_RUN = 4;
__device__ inline int xcd_item_slot()
{
    const int k8 = (int)(blockIdx.x >> 3);
    return ((k8 / XCD_RUN) * 8 + (int)(blockIdx.x & 7)) * XCD_RUN + k8 % XCD_RUN;
}

// v / W for 0 <= v < 128 without a division: for W = 3 a multiply and shift (86/256
// exceeds 1/3 by less than 1/(3 * 128)), for W = 4 a shift.
template <int W>
__device__ inline unsigned small_div(unsigned v)
{
    static_assert(W == 3 || W == 4, "stencil width");
    if constexpr (W == 4) return v >> 2;
    else return __umul24(v, 86u) >> 8;
}

// One axis of a lane's rotated stencil walk.  At step s the lane takes the node at
// offset (s + rot) mod W from its base, rot = (phase - rb) mod W, whose residue mod W
// is then phase + s whatever the base.  wa, xa: the weights and distance terms in
// node order; wr, xr: the same in walk order; er: the walk's node offsets along the
// axis (tile base rb >= 0 included) times the axis stride.  Every output is one
// select per nonzero rotation from the inputs, on W - 1 compares of rot.
template <int W>
__device__ inline void walk_axis(int phase, int rb, int strd, const float* wa, const float* xa, float* wr,
                                 float* xr, int* er)
{
    // phase + 96 - rb lies in (0, 128) for rb in [0, 96), and 96 is a multiple of W
    const unsigned v = (unsigned)(phase + 96 - rb);
    const int rot = (int)(v - small_div<W>(v) * W);
    const int rbs = __umul24(rb, strd);
    // the inputs as opaque values, so that the selects stay selects between
    // registers (not a load from a private array at a computed index)
    float a[W], b[W];
#pragma unroll
    for (int s = 0; s < W; ++s) {
        a[s] = wa[s];
        b[s] = xa[s];
        asm volatile("" : "+v"(a[s]), "+v"(b[s]));
    }
#pragma unroll
    for (int s = 0; s < W; ++s) {
        float wv = a[s], xv = b[s];
        int ov = s * strd;
#pragma unroll
        for (int k = 1; k < W; ++k)
            if (rot == k) {
                wv = a[(s + k) % W];
                xv = b[(s + k) % W];
                ov = ((s + k) % W) * strd;
            }
        wr[s] = wv;
        xr[s] = xv;
        er[s] = rbs + ov;
    }
}

__device__ inline unsigned long long pack_fixed(int lo_bits, int hi_bits)
{
    constexpr unsigned long long B = ((unsigned long long)FIXED_BIAS << 32) | FIXED_BIAS;
    return ((((unsigned long long)(unsigned)hi_bits) << 32) | (unsigned)lo_bits) - B;
}

// The two halves of a packed sum: lo is the low word as a signed int; hi follows
// once lo's sign extension is taken back out.
__device__ inline int unpack_lo(unsigned long long s) { return (int)(unsigned)s; }
__device__ inline int unpack_hi(unsigned long long s)
{
    return (int)((long long)(s - (unsigned long long)(long long)unpack_lo(s)) >> 32);
}

template <int DIM>
__device__ inline void tile_coords(int e, int* t)
{
    constexpr int T = TileShape<DIM>::T;
    t[0] = e % T;
    if constexpr (DIM == 2) { t[1] = e / T; t[2] = 0; }
    else { t[1] = (e / T) % T; t[2] = e / (T * T); }
}

template <int DIM>
__device__ inline bool in_grid(const Phys& ph, int i, int j, int k)
{
    bool ok = i >= 0 && i <= ph.W && j >= 0 && j <= ph.H;
    if constexpr (DIM == 3) ok = ok && k >= 0 && k <= ph.D;
    return ok;
}

// Power-of-two scale 2^e with bound * 2^e <= 2^21: an item has at most 512
// particles, each adding at most once to a node, so a node's sum stays within
// 2^30 and the int32 accumulator cannot overflow.
__device__ inline float fixed_scale(float bound)
{
    if (!(bound > 0.f)) return 1.f;
    int k;
    frexpf(bound, &k);                 // bound < 2^k
    int e = min(max(21 - k, -126), 126);
    return ldexpf(1.f, e);
}

// Workgroup barrier ordering LDS accesses only: each wave waits for its own LDS
// operations, not for its global stores and atomics.  The waves of a workgroup
// exchange data only through LDS here (particle state, grid accumulators and block
// flags are written for later kernels), so the particle stores of G2P and the
// global atomics of a tile flush drain while the workgroup moves on.
__device__ inline void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ inline float wave_max(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

__device__ inline float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// What the scatter needs of one particle, kept in registers between the phases.
template <int DIM>
struct P2GIn {
    float xg[DIM];        // position in grid units
    float mvp[DIM];       // m v
    float A[DIM * DIM];   // m C - (V D^-1 dt) tau
    float pm;
    int mode;             // 0 none, 1 through the tile, 2 direct global atomics
};

// P2G contribution of one particle to one node: (w m, w (m v + A (x_i - x_p))).
template <int DIM>
__device__ inline void node_contrib(float w, float pm, const float* mvp, const float* A, const float* xi,
                                    float* vals)
{
    vals[0] = w * pm;
#pragma unroll
    for (int r = 0; r < DIM; ++r) {
        float s = mvp[r];
#pragma unroll
        for (int c = 0; c < DIM; ++c) s += A[r * DIM + c] * xi[c];
        vals[1 + r] = w * s;
    }
}

template <int DIM, int ORDER, bool COULOMB, bool G2P, bool BATCH>
__global__ void __launch_bounds__(256, 3) k_tiled(DevParticles P, DevGrid src_all, DevGrid dst_all,
                                                  TileWork work, Phys ph_one, const SceneDesc* scenes)
{
    using TS = TileShape<DIM>;
    constexpr int BS = TS::BS, MG = TS::MG, T = TS::T, TN = TS::TN;
    static_assert(T <= 96, "walk_axis takes tile bases below 96");
    constexpr int W = Spline<ORDER>::W;
    constexpr int TV = G2P ? TN : 1;
    constexpr int RMAX = 2;                      // particles per thread per item
    constexpr float REACH = ORDER == 3 ? 2.f : 1.5f;   // max |x_i - x_p| / dx per axis
    // Scatter accumulators.  3D: per tile node two packed fixed-point pairs
    // (pack_fixed), [0] (mass, p_x) and [1] (p_y, p_z): one 64-bit LDS add carries two
    // of the four components.  2D: one 32-bit word per component, [c][node]: three
    // components pack into 64-bit pairs only with a wasted half, which measured
    // slower there (81.5 against 88.9 us per launch on the 64-scene batch).
    constexpr bool PACKED = DIM == 3;
    __shared__ unsigned long long s_acc_raw[PACKED ? 2 * TN : ((DIM + 1) * TN + 1) / 2];
    unsigned long long (*s_acc)[TN] = reinterpret_cast<unsigned long long (*)[TN]>(s_acc_raw);
    int (*s_acc32)[TN] = reinterpret_cast<int (*)[TN]>(s_acc_raw);
    // gather velocities.  3D: component-major, friction velocity in rows 0..2 and,
    // with Coulomb walls, the collision velocity in rows 3..5: one float per row and
    // node, rather than a float4 per node, whose 16-byte stride maps a node's bank to
    // its tile index mod 16, spreads a wave's reads over the LDS banks (falling cube:
    // ~198 -> ~195 us per launch, mean of three runs).  2D: a float4 per node,
    // friction velocity in .xy, collision velocity in a second float4 with Coulomb
    // walls (the component-major rows measured no faster there).
    constexpr bool VSOA = DIM == 3;
    constexpr int NVR = VSOA ? (COULOMB ? 2 * DIM : DIM) : 1;
    __shared__ float s_vel[G2P && VSOA ? NVR : 1][G2P && VSOA ? TV : 1];
    __shared__ float4 s_vel4[COULOMB && G2P && !VSOA ? 2 : 1][G2P && !VSOA ? TV : 1];
    __shared__ int s_touch[TS::NB];
    __shared__ float s_red[2][4];

    // friction (vF) and collision (vC) velocity of tile node e0 + off from the staged
    // tile (off a constant where the caller has one: it then folds into the reads'
    // offsets)
    auto load_node_velocity = [&](int e0, int off, float* vF, float* vC) {
        if constexpr (VSOA) {
            const float* sv = &s_vel[0][0] + e0;
            vF[2] = vC[2] = 0.f;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                vF[d] = sv[d * TV + off];
                vC[d] = COULOMB ? sv[((DIM + d) % NVR) * TV + off] : vF[d];
            }
        } else {
            const float4* s4 = &s_vel4[0][0] + e0;
            const float4 f4 = s4[off];
            vF[0] = f4.x; vF[1] = f4.y; vF[2] = f4.z;
            if constexpr (COULOMB) {
                const float4 c4 = s4[(COULOMB ? TV : 0) + off];
                vC[0] = c4.x; vC[1] = c4.y; vC[2] = c4.z;
            } else {
                vC[0] = vF[0]; vC[1] = vF[1]; vC[2] = vF[2];
            }
        }
    };

    const int n = P.n;
    const int tid = threadIdx.x;
    const int nit = *work.n_items;

    // Items are walked grid-stride in key (spatial) order.  Neither an XCD-contiguous
    // assignment (workgroup b runs on XCD b % 8) nor a block-parity colour order of
    // the items measured faster on the falling-cube workload, so neither is made; the
    // super-block kernel's runs of XCD_RUN items per XCD (xcd_item_slot) measured
    // 61.3 -> 60.8 us on the 64-scene batch but 8.9 -> 9.1 us on the 2D snowballs.
    for (int it = blockIdx.x; it < nit; it += gridDim.x) {
        int4 item = work.items[it];
        // a batch of independent scenes: the item's scene supplies parameters and grid
        Phys ph = ph_one;
        DevGrid src = src_all, dst = dst_all;
        int blk = item.x;
        if constexpr (BATCH) {
            const SceneDesc sd = scenes[item.w];
            ph = sd.ph;
            src = scene_grid(src_all, sd);
            dst = scene_grid(dst_all, sd);
            blk -= sd.block_off;
        }
        int first = item.y, cnt = min(item.z, 256 * RMAX);
        int bc[3];
        bc[0] = blk % dst.nbx;
        int rest = blk / dst.nbx;
        if constexpr (DIM == 2) { bc[1] = rest; bc[2] = 0; }
        else { bc[1] = rest % dst.nby; bc[2] = rest / dst.nby; }
        int o[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) o[a] = (a < DIM) ? bc[a] * BS - 1 - MG : 0;

        // ---- stage the tile: velocities to gather from, zeroed accumulators (all
        // loads issued before the first LDS store, so their latencies overlap).
        // The lane's tile entries are derived from an opaque copy of its index, so
        // they are recomputed per item (a few integer operations) instead of being
        // hoisted out of the item loop and kept live across the G2P phase.
        int ltid = tid;
        asm volatile("" : "+v"(ltid));
        {
            constexpr int NR = (TN + 255) / 256;
            float4 va[NR], vb[NR];
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                va[k] = make_float4(0.f, 0.f, 0.f, 0.f);
                vb[k] = va[k];
                const int e = ltid + 256 * k;
                if (G2P && e < TN) {
                    int t[3];
                    tile_coords<DIM>(e, t);
                    int i = o[0] + t[0], j = o[1] + t[1], kk = o[2] + t[2];
                    if (in_grid<DIM>(ph, i, j, kk)) {
                        int addr = node_addr<DIM>(src, i, j, kk);
                        va[k] = src.a[addr];
                        if constexpr (COULOMB) vb[k] = src.vc[addr];
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const int e = ltid + 256 * k;
                if (e >= TN) continue;
                if constexpr (PACKED) {
                    s_acc[0][e] = 0ull;
                    s_acc[1][e] = 0ull;
                } else {
#pragma unroll
                    for (int c = 0; c <= DIM; ++c) s_acc32[c][e] = 0;
                }
                if constexpr (G2P && VSOA) {
                    const float f[3] = {va[k].y, va[k].z, va[k].w}, c[3] = {vb[k].x, vb[k].y, vb[k].z};
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        s_vel[d][e] = f[d];
                        if constexpr (COULOMB) s_vel[(DIM + d) % NVR][e] = c[d];
                    }
                } else if constexpr (G2P) {
                    s_vel4[0][e] = make_float4(va[k].y, va[k].z, va[k].w, 0.f);
                    if constexpr (COULOMB) s_vel4[COULOMB ? 1 : 0][e] = vb[k];
                }
            }
        }
        if (tid < TS::NB) s_touch[tid] = 0;
        lds_barrier();

        // ---- phase 1: per particle update (G2P) and what its scatter needs
        P2GIn<DIM> in[RMAX];
        float bound_m = 0.f, bound_p = 0.f;
        // The item's particles form npass = ceil(cnt / 64) wave passes, so only the
        // passes the item needs run.  3D: a pass holds 64 consecutive particles
        // (coalesced loads and stores of the particle state, which measured faster
        // than spreading lanes over cells).  2D: pass pi holds particles pi,
        // pi + npass, pi + 2 npass, ..., so the 64 lanes spread over the block's cells
        // and the scatter's LDS adds of one pass rarely meet on one address (particles
        // are sorted by cell), which measured faster there.  Wave w runs passes
        // ((w + it) & 3) + 4r; the rotation moves an item's odd pass between waves.
        const int npass = (cnt + 63) >> 6;
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            in[r].mode = 0;
            const int pass = 4 * r + (((tid >> 6) + it) & 3);
            if (pass >= npass) continue;
            const int q = DIM == 3 ? pass * 64 + (tid & 63) : pass + npass * (tid & 63);
            if (q >= cnt) continue;
            const int p = first + q;
            float x[DIM], v[DIM], C[DIM * DIM], tau[DIM * DIM], Jp;
#pragma unroll
            for (int d = 0; d < DIM; ++d) x[d] = P.x[d * n + p];
            Jp = P.Jp[p];
            const float vol = P.vol[p];
            const float pm = P.mass[p];

            if constexpr (G2P) {
                // g2p.glsl: gather, B (C = D^-1 B), position, F_E, plasticity
                float xg[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) xg[d] = x[d] * ph.inv_dx;
                Stencil<DIM, ORDER> st;
                st.init(xg);
                bool tile_ok = true;
                int rb[3] = {0, 0, 0};
#pragma unroll
                for (int a = 0; a < DIM; ++a) {
                    rb[a] = st.base[a] - o[a];
                    tile_ok &= rb[a] >= 0 && rb[a] + W <= T;
                }
                float nv[DIM], nB[DIM * DIM], nx[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) { nv[d] = 0.f; nx[d] = 0.f; }
#pragma unroll
                for (int e = 0; e < DIM * DIM; ++e) nB[e] = 0.f;
                // whole stencil inside the grid and the tile (nearly every particle):
                // no per-node bounds tests, and the stencil sums are contracted one
                // axis at a time (below)
                bool inside = tile_ok;
                {
                    const int lim[3] = {ph.W, ph.H, ph.D};
#pragma unroll
                    for (int a = 0; a < DIM; ++a) inside &= st.base[a] >= 0 && st.base[a] + W - 1 <= lim[a];
                }
                auto gather = [&](auto fast) {
#pragma unroll
                    for (int a0 = 0; a0 < W; ++a0)
#pragma unroll
                        for (int a1 = 0; a1 < W; ++a1)
#pragma unroll
                            for (int a2 = 0; a2 < (DIM == 3 ? W : 1); ++a2) {
                                int idx[3] = {st.base[0] + a0, st.base[1] + a1,
                                              DIM == 3 ? st.base[DIM - 1] + a2 : 0};
                                if constexpr (!decltype(fast)::value)
                                    if (!in_grid<DIM>(ph, idx[0], idx[1], idx[2])) continue;
                                float dp[DIM];
                                dp[0] = st.d[0][a0] * ph.dx;
                                dp[1] = st.d[1][a1] * ph.dx;
                                float w = st.w[0][a0] * st.w[1][a1];
                                if constexpr (DIM == 3) {
                                    dp[DIM - 1] = st.d[DIM - 1][a2] * ph.dx;
                                    w = w * st.w[DIM - 1][a2];
                                }
                                float wdp[DIM];
#pragma unroll
                                for (int c = 0; c < DIM; ++c) wdp[c] = w * dp[c];
                                float vF[3], vC[3];
                                if (decltype(fast)::value || tile_ok) {
                                    int e = DIM == 3 ? ((rb[2] + a2) * T + (rb[1] + a1)) * T + (rb[0] + a0)
                                                     : (rb[1] + a1) * T + (rb[0] + a0);
                                    load_node_velocity(e, 0, vF, vC);
                                } else {
                                    int addr = node_addr<DIM>(src, idx[0], idx[1], idx[2]);
                                    float4 a4 = src.a[addr];
                                    vF[0] = a4.y; vF[1] = a4.z; vF[2] = a4.w;
                                    if constexpr (COULOMB) {
                                        float4 c4 = src.vc[addr];
                                        vC[0] = c4.x; vC[1] = c4.y; vC[2] = c4.z;
                                    } else {
                                        vC[0] = vF[0]; vC[1] = vF[1]; vC[2] = vF[2];
                                    }
                                }
#pragma unroll
                                for (int rr = 0; rr < DIM; ++rr) {
                                    nv[rr] += vF[rr] * w;
#pragma unroll
                                    for (int c = 0; c < DIM; ++c) nB[rr * DIM + c] -= vF[rr] * wdp[c];
                                    nx[rr] += (ph.dt * vC[rr] + (float)idx[rr] * ph.dx) * w;
                                }
                            }
                };
                // Interior particles: the node weight is a product of per-axis
                // weights, so the sums over the W^DIM nodes factor axis by axis:
                //   nv   = sum_i w_i v_i
                //   nB_c = sum_i w_i v_i (x_i - x_p)_c   (one factor w_c -> w_c (x_i - x_p)_c)
                //   nx   = sum_i w_i (x_i + dt vC_i) = dx sum_i w_i idx_i + dt sum_i w_i vC_i
                // with sum_i w_i idx_i,c = (sum_s w_c[s] (base_c + s)) prod_{a != c} sum_s w_a[s].
                // Same sums as the node loop up to the order of the additions
                // (about 14 FLOPs per node instead of 30).
                auto gather_separable = [&]() {
                    float wd[DIM][W];
#pragma unroll
                    for (int a = 0; a < DIM; ++a)
#pragma unroll
                        for (int s2 = 0; s2 < W; ++s2) wd[a][s2] = -st.w[a][s2] * (st.d[a][s2] * ph.dx);
                    float nc[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) nc[d] = 0.f;
                    // the stencil's first tile node; the others at constant offsets
                    const int e0 = DIM == 3 ? (rb[2] * T + rb[1]) * T + rb[0] : rb[1] * T + rb[0];
#pragma unroll
                    for (int a0 = 0; a0 < W; ++a0) {
                        // sums over a1 (and a2): S v, D1 = sum wd1 v, D2 = sum wd2 v, Cs vC
                        float S1[DIM], D1[DIM], D2[DIM], C1[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; ++d) { S1[d] = 0.f; D1[d] = 0.f; D2[d] = 0.f; C1[d] = 0.f; }
#pragma unroll
                        for (int a1 = 0; a1 < W; ++a1) {
                            float S2[DIM], E2[DIM], C2[DIM];
#pragma unroll
                            for (int d = 0; d < DIM; ++d) { S2[d] = 0.f; E2[d] = 0.f; C2[d] = 0.f; }
#pragma unroll
                            for (int a2 = 0; a2 < (DIM == 3 ? W : 1); ++a2) {
                                const int off = DIM == 3 ? (a2 * T + a1) * T + a0 : a1 * T + a0;
                                float vF[3], vC[3];
                                load_node_velocity(e0, off, vF, vC);
                                if constexpr (DIM == 3) {
                                    const float w2 = st.w[DIM - 1][a2], wd2 = wd[DIM - 1][a2];
#pragma unroll
                                    for (int d = 0; d < DIM; ++d) {
                                        S2[d] += w2 * vF[d];
                                        E2[d] += wd2 * vF[d];
                                        if constexpr (COULOMB) C2[d] += w2 * vC[d];
                                    }
                                } else {
#pragma unroll
                                    for (int d = 0; d < DIM; ++d) { S2[d] = vF[d]; C2[d] = vC[d]; }
                                }
                            }
                            const float w1 = st.w[1][a1], wd1 = wd[1][a1];
#pragma unroll
                            for (int d = 0; d < DIM; ++d) {
                                S1[d] += w1 * S2[d];
                                D1[d] += wd1 * S2[d];
                                if constexpr (DIM == 3) D2[d] += w1 * E2[d];
                                if constexpr (COULOMB) C1[d] += w1 * C2[d];
                            }
                        }
                        const float w0 = st.w[0][a0], wd0 = wd[0][a0];
#pragma unroll
                        for (int d = 0; d < DIM; ++d) {
                            nv[d] += w0 * S1[d];
                            nB[d * DIM + 0] += wd0 * S1[d];
                            nB[d * DIM + 1] += w0 * D1[d];
                            if constexpr (DIM == 3) nB[d * DIM + 2] += w0 * D2[d];
                            if constexpr (COULOMB) nc[d] += w0 * C1[d];
                        }
                    }
                    float sw[DIM], mi[DIM];
#pragma unroll
                    for (int a = 0; a < DIM; ++a) {
                        sw[a] = 0.f;
                        mi[a] = 0.f;
#pragma unroll
                        for (int s2 = 0; s2 < W; ++s2) {
                            sw[a] += st.w[a][s2];
                            mi[a] += st.w[a][s2] * (float)(st.base[a] + s2);
                        }
                    }
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        float m = mi[d];
#pragma unroll
                        for (int a = 0; a < DIM; ++a)
                            if (a != d) m *= sw[a];
                        nx[d] = m * ph.dx + ph.dt * (COULOMB ? nc[d] : nv[d]);
                    }
                };
                if (inside) gather_separable();
                else gather(std::false_type{});
                float M[DIM * DIM], Fo[DIM * DIM], Fh[DIM * DIM], F[DIM * DIM];
#pragma unroll
                for (int e = 0; e < DIM * DIM; ++e) {
                    C[e] = ph.Dinv * nB[e];
                    M[e] = ((e % (DIM + 1)) == 0 ? 1.f : 0.f) + ph.dt * C[e];
                    Fo[e] = P.F[e * n + p];
                }
                matmul<DIM>(M, Fo, Fh);
                float U[DIM * DIM], sig[DIM], V[DIM * DIM], sc[DIM];
                svd<DIM>(Fh, U, sig, V);
                float ratio = 1.f, J = 1.f;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    float sv = fminf(fmaxf(sig[d], ph.theta0), ph.theta1);
                    sv = fmaxf(sv, 1e-6f);
                    sc[d] = sv;
                    ratio *= sig[d] / sv;
                    J *= sv;
                }
#pragma unroll
                for (int rr = 0; rr < DIM; ++rr)
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        float sacc = (U[rr * DIM] * sc[0]) * V[c * DIM];
#pragma unroll
                        for (int qq = 1; qq < DIM; ++qq) sacc += (U[rr * DIM + qq] * sc[qq]) * V[c * DIM + qq];
                        F[rr * DIM + c] = sacc;
                    }
                Jp = Jp * ratio;
#pragma unroll
                for (int d = 0; d < DIM; ++d) { x[d] = nx[d]; v[d] = nv[d]; }
                // v and C only when a reader follows (the next substep's gather
                // recomputes both from the grid)
#pragma unroll
                for (int d = 0; d < DIM; ++d) P.x[d * n + p] = x[d];
#pragma unroll
                for (int e = 0; e < DIM * DIM; ++e) P.F[e * n + p] = F[e];
                P.Jp[p] = Jp;
                if (work.store_vc) {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) P.v[d * n + p] = v[d];
#pragma unroll
                    for (int e = 0; e < DIM * DIM; ++e) P.C[e * n + p] = C[e];
                }
                // stress of the new state: U diag(2 mu (s-1) s + la (J-1) J) U^T
                float harden = expf(ph.xi * (1.f - Jp));
                float tmu = 2.f * (ph.mu * harden);
                float sI = (ph.lambda * harden) * (J - 1.f) * J;
                float tdiag[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) tdiag[d] = tmu * (sc[d] - 1.f) * sc[d] + sI;
#pragma unroll
                for (int rr = 0; rr < DIM; ++rr)
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        float sacc = U[rr * DIM] * tdiag[0] * U[c * DIM];
#pragma unroll
                        for (int qq = 1; qq < DIM; ++qq) sacc += U[rr * DIM + qq] * tdiag[qq] * U[c * DIM + qq];
                        tau[rr * DIM + c] = sacc;
                    }
            } else {
                // stress from the stored state (p2g.glsl:44-59)
                float F[DIM * DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) v[d] = P.v[d * n + p];
#pragma unroll
                for (int e = 0; e < DIM * DIM; ++e) { C[e] = P.C[e * n + p]; F[e] = P.F[e * n + p]; }
                float U[DIM * DIM], sig[DIM], V[DIM * DIM];
                svd<DIM>(F, U, sig, V);
                kirchhoff<DIM>(ph, F, U, V, det<DIM>(F), Jp, tau);
            }

            if (vol == 0.f) continue;                                 // p2g.glsl:42
            // momentum of node i: w (m v + A (x_i - x_p)), A = m C - (V D^-1 dt) tau
            // D^-1 dt formed here from the uniform parameters (an opaque copy keeps
            // it from being hoisted out of the item loop into a long-lived register)
            float dinv = ph.Dinv;
            asm volatile("" : "+v"(dinv));
            const float vfd = vol * (dinv * ph.dt);
            float bp = 0.f;
#pragma unroll
            for (int rr = 0; rr < DIM; ++rr) {
                in[r].mvp[rr] = pm * v[rr];
                float rowsum = 0.f;
#pragma unroll
                for (int c = 0; c < DIM; ++c) {
                    float a = pm * C[rr * DIM + c] - vfd * tau[rr * DIM + c];
                    in[r].A[rr * DIM + c] = a;
                    rowsum += fabsf(a);
                }
                bp = fmaxf(bp, fabsf(in[r].mvp[rr]) + REACH * ph.dx * rowsum);
            }
            in[r].pm = pm;
            bool tile_ok = true;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                in[r].xg[d] = x[d] * ph.inv_dx;
                int rb = Spline<ORDER>::base(in[r].xg[d]) - o[d];
                tile_ok &= rb >= 0 && rb + W <= T;
            }
            // a bound of 0.5 * FLT_MAX keeps bound * count finite; beyond it (or NaN)
            // the particle takes the float path so non-finite values propagate
            bool finite = bp < 1.0e37f && fabsf(pm) < 1.0e37f;
            if (tile_ok && finite) {
                in[r].mode = 1;
                bound_m = fmaxf(bound_m, fabsf(pm));
                bound_p = fmaxf(bound_p, bp);
            } else {
                in[r].mode = 2;
            }
        }

        // ---- the item's fixed-point scales
        bound_m = wave_max(bound_m);
        bound_p = wave_max(bound_p);
        if ((tid & 63) == 0) { s_red[0][tid >> 6] = bound_m; s_red[1][tid >> 6] = bound_p; }
        lds_barrier();
        const float scale_m = fixed_scale(fmaxf(fmaxf(s_red[0][0], s_red[0][1]), fmaxf(s_red[0][2], s_red[0][3])));
        const float scale_p = fixed_scale(fmaxf(fmaxf(s_red[1][0], s_red[1][1]), fmaxf(s_red[1][2], s_red[1][3])));

        // ---- phase 2: scatter (p2g.glsl) into the tile, or directly to HBM
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            if (in[r].mode == 0) continue;
            int base[3] = {0, 0, 0}, rb[3] = {0, 0, 0};
            float wa[DIM][W], xa[DIM][W];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                base[d] = Spline<ORDER>::base(in[r].xg[d]);
                rb[d] = base[d] - o[d];
                float dd[W];
                Spline<ORDER>::weights(in[r].xg[d], base[d], wa[d], dd);
#pragma unroll
                for (int s2 = 0; s2 < W; ++s2) xa[d][s2] = -dd[s2] * ph.dx;
            }
            const bool via_tile = in[r].mode == 1;
            bool inside = via_tile;
            {
                const int lim[3] = {ph.W, ph.H, ph.D};
#pragma unroll
                for (int d = 0; d < DIM; ++d) inside &= base[d] >= 0 && base[d] + W - 1 <= lim[d];
            }
            if (inside) {
                // through the tile, whole stencil inside: the inputs pre-multiplied by
                // the item's power-of-two scales (exact), one rounding per component
                // (fixed_bits) and LDS adds (3D: two of packed component pairs per
                // node).  Each lane walks its stencil from a start rotated by its
                // phase, so that lanes whose particles share stencil nodes (3D: sorted
                // neighbours, a pass holds consecutive particles; 2D: a pass's lanes
                // spread over the block's cells) add to different words at each step
                // instead of serialising on one (2D batch: 75.9 against 69.1 us; a
                // rotation by the lane's digits alone measured slower there, 88.3
                // against 81.5 us).
                const float pms = in[r].pm * scale_m;
                float mvs[DIM], As[DIM * DIM];
#pragma unroll
                for (int rr = 0; rr < DIM; ++rr) {
                    mvs[rr] = in[r].mvp[rr] * scale_p;
#pragma unroll
                    for (int c = 0; c < DIM; ++c) As[rr * DIM + c] = in[r].A[rr * DIM + c] * scale_p;
                }
                float wr[DIM][W], xr[DIM][W];
                int er[DIM][W];
                {
                    const int strd[3] = {1, T, T * T};
                    // the lane's rotation, from an opaque copy of its index: derived per
                    // particle (a few integer operations) instead of being hoisted out
                    // of the item loop as compare masks kept live (or spilled) across G2P
                    unsigned lr = tid & 63;
                    asm volatile("" : "+v"(lr));
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        // the lane's phase (its index's base-W digits) fixes
                        // (base + rot) mod W, so that lanes of different phases never
                        // add to one word at the same step (as in k_super)
                        const unsigned q = small_div<W>(lr);
                        walk_axis<W>((int)(lr - q * W), rb[d], strd[d], wa[d], xa[d], wr[d], xr[d], er[d]);
                        lr = q;
                    }
                }
#pragma unroll
                for (int a0 = 0; a0 < W; ++a0)
#pragma unroll
                    for (int a1 = 0; a1 < W; ++a1) {
                        const float w01 = wr[0][a0] * wr[1][a1];
                        // momentum rows up to the last axis' term (the same chain of
                        // additions as the full sum, which continues below)
                        float sb[DIM];
#pragma unroll
                        for (int rr = 0; rr < DIM; ++rr)
                            sb[rr] = (mvs[rr] + As[rr * DIM] * xr[0][a0]) + As[rr * DIM + 1] * xr[1][a1];
#pragma unroll
                        for (int a2 = 0; a2 < (DIM == 3 ? W : 1); ++a2) {
                            float w = w01;
                            float sm[DIM];
#pragma unroll
                            for (int rr = 0; rr < DIM; ++rr) sm[rr] = sb[rr];
                            if constexpr (DIM == 3) {
                                w = w * wr[2][a2];
#pragma unroll
                                for (int rr = 0; rr < DIM; ++rr) sm[rr] += As[rr * DIM + 2] * xr[2][a2];
                            }
                            int qb[4];
                            qb[0] = fixed_bits(w, pms);
#pragma unroll
                            for (int rr = 0; rr < DIM; ++rr) qb[1 + rr] = fixed_bits(w, sm[rr]);
                            const int e = DIM == 3 ? er[0][a0] + er[1][a1] + er[DIM - 1][a2] : er[0][a0] + er[1][a1];
                            if constexpr (PACKED) {
                                atomicAdd(&s_acc[0][e], pack_fixed(qb[0], qb[1]));
                                atomicAdd(&s_acc[1][e], pack_fixed(qb[2], qb[3]));
                            } else {
#pragma unroll
                                for (int c = 0; c <= DIM; ++c) atomicAdd(&s_acc32[c][e], qb[c] - (int)FIXED_BIAS);
                            }
                        }
                    }
                continue;
            }
            auto scatter = [&](auto fast) {
#pragma unroll
                for (int a0 = 0; a0 < W; ++a0)
#pragma unroll
                    for (int a1 = 0; a1 < W; ++a1)
#pragma unroll
                        for (int a2 = 0; a2 < (DIM == 3 ? W : 1); ++a2) {
                            const int idx[3] = {base[0] + a0, base[1] + a1, DIM == 3 ? base[DIM - 1] + a2 : 0};
                            if constexpr (!decltype(fast)::value)
                                if (!in_grid<DIM>(ph, idx[0], idx[1], idx[2])) continue;
                            float w = wa[0][a0] * wa[1][a1];
                            float xi[DIM];
                            xi[0] = xa[0][a0];
                            xi[1] = xa[1][a1];
                            if constexpr (DIM == 3) {
                                w = w * wa[2][a2];
                                xi[2] = xa[2][a2];
                            }
                            float vals[DIM + 1];
                            node_contrib<DIM>(w, in[r].pm, in[r].mvp, in[r].A, xi, vals);
                            if (decltype(fast)::value || via_tile) {
                                const int e = DIM == 3 ? ((rb[2] + a2) * T + (rb[1] + a1)) * T + (rb[0] + a0)
                                                       : (rb[1] + a1) * T + (rb[0] + a0);
                                int qb[4];
                                qb[0] = fixed_bits(vals[0], scale_m);
#pragma unroll
                                for (int c = 1; c <= DIM; ++c) qb[c] = fixed_bits(vals[c], scale_p);
                                if constexpr (PACKED) {
                                    atomicAdd(&s_acc[0][e], pack_fixed(qb[0], qb[1]));
                                    atomicAdd(&s_acc[1][e], pack_fixed(qb[2], qb[3]));
                                } else {
#pragma unroll
                                    for (int c = 0; c <= DIM; ++c) {
                                        const int q = qb[c] - (int)FIXED_BIAS;
                                        if (q != 0) atomicAdd(&s_acc32[c][e], q);
                                    }
                                }
                            } else {
                                float* dstp =
                                    reinterpret_cast<float*>(&dst.a[node_addr<DIM>(dst, idx[0], idx[1], idx[2])]);
#pragma unroll
                                for (int c = 0; c <= DIM; ++c) unsafeAtomicAdd(dstp + c, vals[c]);
                                dst.flags[block_of<DIM>(dst, idx[0], idx[1], idx[2])] = 1u;
                            }
                        }
            };
            scatter(std::false_type{});
        }
        lds_barrier();

        // ---- flush the tile: fixed point back to FP32, one global atomic per component.
        // Lanes take (node, component) pairs, component fastest, so one atomic
        // instruction covers 16 (3D) or 21 (2D) whole nodes: the grid stores a node's
        // DIM + 1 accumulators contiguously and a block row of nodes contiguously, and
        // the memory system serves the lanes of one instruction that fall in one cache
        // line as one request (global atomics are bound by requests, not lanes).
        const float inv_m = 1.f / scale_m, inv_p = 1.f / scale_p;
        constexpr int NC = DIM + 1;
        // only nodes inside the grid receive tile adds; for a tile reaching past a
        // wall each entry is tested as well, so that an entry outside the grid never
        // addresses memory (the test is uniform, and skipped, for interior tiles)
        bool tile_in = true;
        {
            const int lim[3] = {ph.W, ph.H, ph.D};
#pragma unroll
            for (int a = 0; a < DIM; ++a) tile_in &= o[a] >= 0 && o[a] + T - 1 <= lim[a];
        }
        for (int pe = tid; pe < TN * NC; pe += 256) {
            const int e = pe / NC, c = pe - e * NC;
            int acc, mass;
            if constexpr (PACKED) {
                const unsigned long long sum = s_acc[c >> 1][e];
                acc = (c & 1) ? unpack_hi(sum) : unpack_lo(sum);
                mass = c == 0 ? acc : unpack_lo(s_acc[0][e]);
            } else {
                acc = s_acc32[c][e];
                mass = c == 0 ? acc : s_acc32[0][e];
            }
            // a node whose mass rounded to zero flushes no momentum either: mass and
            // momentum round independently, and a momentum unit over a float-path mass
            // far below the mass unit (a direct scatter's tiny stencil weight) would
            // give the node any velocity
            if (acc == 0 || mass == 0) continue;
            int t[3];
            tile_coords<DIM>(e, t);
            int i = o[0] + t[0], j = o[1] + t[1], k = o[2] + t[2];
            if (!tile_in && !in_grid<DIM>(ph, i, j, k)) continue;
            float* dstp = reinterpret_cast<float*>(&dst.a[node_addr<DIM>(dst, i, j, k)]) + c;
            unsafeAtomicAdd(dstp, (float)acc * (c == 0 ? inv_m : inv_p));
            int rbx = (i >> BlockShape<DIM>::SHIFT) - (bc[0] - 1);
            int rby = (j >> BlockShape<DIM>::SHIFT) - (bc[1] - 1);
            int rbz = DIM == 3 ? (k >> BlockShape<DIM>::SHIFT) - (bc[2] - 1) : 0;
            s_touch[(rbz * 3 + rby) * 3 + rbx] = 1;
        }
        lds_barrier();
        int ftid = tid;                       // recomputed per item, as ltid above
        asm volatile("" : "+v"(ftid));
        if (ftid < TS::NB && s_touch[ftid]) {
            int rbx = ftid % 3, rby = (ftid / 3) % 3, rbz = ftid / 9;
            int bx = bc[0] - 1 + rbx, by = bc[1] - 1 + rby, bz = DIM == 3 ? bc[2] - 1 + rbz : 0;
            int b = DIM == 3 ? (bz * dst.nby + by) * dst.nbx + bx : by * dst.nbx + bx;
            dst.flags[b] = 1u;
        }
        lds_barrier();
    }
}

// Grid of the tiled kernels: a multiple of 8 (one share per XCD), at most 8192 (the
// item count is on the device; workgroups past it exit at once).  On the falling
// cube (~7,100 items) 8192 measured 224 us per launch against 230 us for 4096 (half
// the workgroups taking two items), 246 us for 2048 and 240 us for 768; 16384 the
// same as 8192.
inline int tiled_grid(const TileWork& w)
{
    int g = w.max_items < 8192 ? w.max_items : 8192;
    g = (g + 7) & ~7;
    return g > 0 ? g : 8;
}

// The 3D cubic-stencil, Coulomb-wall, single-scene fused kernel lives in its own
// translation unit (mpm_fused_cubic.hip), built with a higher full-unroll limit:
// there its 64-node stencil loops unroll completely, which measured 726 -> 643 us
// per launch on the falling cube (the same limit slowed the cubic sticky-wall
// kernel, 522 -> 636 us, and left the quadratic kernels unchanged).
hipError_t launch_g2p2g_cubic_coulomb3d(dim3 grid, const DevParticles& p, const DevGrid& src, const DevGrid& dst,
                                        const TileWork& w, const Phys& ph, hipStream_t st);

}  // namespace mpm
