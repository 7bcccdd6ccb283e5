// mpm_super.hip — the 3D fused substep over super-block work items (mpm_options
// fused = 2 or 3): one workgroup takes the particles binned into 2x2x2 grid blocks
// (8^3 cells) and runs G2P (shaders/3d/g2p.glsl:22-105) and the P2G of the updated
// particle (shaders/3d/p2g.glsl:31-97) in ONE pass per particle, through an LDS tile
// of the 8^3 cells plus the stencil reach and a drift margin of MG cells.
//
// Against the per-block kernel k_tiled (mpm_fused_kernel.hpp):
//   * the tile is staged and flushed once per ~8x more particles (a 4^3 block's
//     9^3-node tile carries ~2.6 staged and ~1.2 flushed nodes per particle at 4.4
//     particles per cell; an 8^3 super-block's 11^3 tile 0.6 and 0.6);
//   * no barrier between G2P and the scatter: k_tiled must finish the G2P of all
//     its particles before it knows the item's fixed-point scale, and keeps their
//     scatter inputs in registers across that barrier; here the scale comes from
//     the bound the same particles had one substep earlier (written by this kernel
//     at the end of each item, and after each binning by k_item_bounds from the
//     stored state, which is exactly what the previous substep scattered), plus the
//     momentum gravity adds in one substep.  A particle whose bound exceeds the
//     scale's range this substep scatters with global float atomics instead, so
//     results never depend on how well the bound predicted it.
//
// Fixed-point accumulation as in k_tiled: per component an integer sum in LDS, two
// components per 64-bit add.  With cnt particles in the item a node receives at
// most cnt contributions, each |w * value| * 2^e <= wmax * 2^cap rounded (w <=
// 0.75^3 for the quadratic, (2/3)^3 for the cubic stencil); cap (fixed_cap) is the
// largest for which cnt of them stay below 2^31, so each 32-bit half of a node's sum
// cannot overflow, and every contribution is below 2^22 as the rounding trick of
// fixed_bits needs.  For 512 to 4096 particles cap is 23 to 20, and the rounding of
// a contribution is at most 2^-(cap+1) of the scale's power of two, 2^-cap of the
// item's bound: the per-block kernel's resolution for the same item sizes.
#include "mpm_fused_kernel.hpp"
#include "mpm_super.hpp"

namespace mpm {

namespace {

// |m|, the scatter's inputs m v and A = m C - (V D^-1 dt) tau, and the bound on the
// magnitude of any one of its momentum contributions, |m v_r| + reach dx sum_c |A_rc|
// (the same quantities k_tiled computes for its scale).
template <int ORDER>
__device__ inline float scatter_inputs(const Phys& ph, float pm, float vol, const float* v, const float* C,
                                       const float* tau, float* mvp, float* A)
{
    constexpr float REACH = ORDER == 3 ? 2.f : 1.5f;
    float dinv = ph.Dinv;
    asm volatile("" : "+v"(dinv));           // formed per particle, not hoisted and kept live
    const float vfd = vol * (dinv * ph.dt);
    float bp = 0.f;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        mvp[r] = pm * v[r];
        float rowsum = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float a = pm * C[r * 3 + c] - vfd * tau[r * 3 + c];
            A[r * 3 + c] = a;
            rowsum += fabsf(a);
        }
        bp = fmaxf(bp, fabsf(mvp[r]) + REACH * ph.dx * rowsum);
    }
    return bp;
}

// 2^e with bound * 2^e in [2^(cap-1), 2^cap): the power-of-two rounding leaves the
// item between 1x and 2x of headroom over the bound.  A zero or non-finite bound
// gives 2^126, so that every particle with a nonzero contribution takes the float
// path.
__device__ inline float headroom_scale(float bound, int cap)
{
    if (!(bound > 0.f) || !(bound < 1.0e37f)) return ldexpf(1.f, 126);
    int k;
    frexpf(bound, &k);                       // bound < 2^k
    return ldexpf(1.f, min(max(cap - k, -126), 126));
}

// The largest cap <= 23 for which cnt contributions of at most wmax * 2^cap + 1/2
// each (a stencil weight times a value within 2^cap, rounded) sum below 2^31 (with a
// margin of 2^-10 for the float evaluation); 23 keeps every contribution below 2^22
// as fixed_bits needs.
__device__ inline int fixed_cap(int cnt, float wmax)
{
    int cap = 23;
    while (cap > 0 && (float)cnt * (wmax * ldexpf(1.f, cap) + 0.5f) >= 2145386496.f) --cap;
    return cap;
}

}  // namespace

// Largest |m| and largest momentum bound of each item's particles in the stored
// state (shaders/3d/p2g.glsl:44-59 stress), i.e. what a scatter of that state
// would need, their mean position (the reference of the tile's drift) and the
// z-range of their stencil bases, with a zero tile offset: written after every
// binning.
template <int ORDER>
__global__ void __launch_bounds__(256) k_item_bounds(DevParticles P, SuperWork work, Phys ph)
{
    __shared__ float s_red[7][4];
    const int n = P.n;
    const int nit = *work.n_items;
    for (int it = blockIdx.x; it < nit; it += gridDim.x) {
        const int4 item = work.items[it];
        const int cnt = min(item.z, SB_MAX_PARTICLES);
        int sc[3];
        super_coords(item.x, work.nsx, work.nsy, sc);
        float bm = 0.f, bpm = 0.f, msum[3] = {0.f, 0.f, 0.f};
        int zlo = 1 << 20, zhi = -(1 << 20);
        for (int q = threadIdx.x; q < cnt; q += 256) {
            const int p = item.y + q;
#pragma unroll
            for (int d = 0; d < 3; ++d) msum[d] += P.x[d * n + p] * ph.inv_dx - (float)(sc[d] * SB_CELLS);
            const int bz = Spline<ORDER>::base(P.x[2 * n + p] * ph.inv_dx);
            zlo = min(zlo, bz);
            zhi = max(zhi, bz);
            const float vol = P.vol[p];
            if (vol == 0.f) continue;                                // p2g.glsl:42
            const float pm = P.mass[p], Jp = P.Jp[p];
            float v[3], C[9], F[9], tau[9], mvp[3], A[9];
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] = P.v[d * n + p];
#pragma unroll
            for (int e = 0; e < 9; ++e) { C[e] = P.C[e * n + p]; F[e] = P.F[e * n + p]; }
            // the stress of an elastic state from its left stretch, as the fused
            // kernel takes it, otherwise through the SVD
            float elo2, ehi2, J;
            elastic_bounds(ph.theta0, ph.theta1, &elo2, &ehi2);
            Sym3 vmi, bmv;
            if (elastic_stretch(F, elo2, ehi2, &vmi, &bmv, &J)) {
                const float harden = expf(ph.xi * (1.f - Jp));
                elastic_kirchhoff(bmv, J, 2.f * (ph.mu * harden), ph.lambda * harden, tau);
            } else {
                float U[9], sig[3], V[9];
                svd3(F, U, sig, V);
                kirchhoff<3>(ph, F, U, V, det<3>(F), Jp, tau);
            }
            const float bp = scatter_inputs<ORDER>(ph, pm, vol, v, C, tau, mvp, A);
            if (bp < 1.0e37f && fabsf(pm) < 1.0e37f) {
                bm = fmaxf(bm, fabsf(pm));
                bpm = fmaxf(bpm, bp);
            }
        }
        bm = wave_max(bm);
        bpm = wave_max(bpm);
#pragma unroll
        for (int d = 0; d < 3; ++d) msum[d] = wave_sum(msum[d]);
        const float nzlo = wave_max(-(float)zlo), fzhi = wave_max((float)zhi);
        if ((threadIdx.x & 63) == 0) {
            s_red[0][threadIdx.x >> 6] = bm;
            s_red[1][threadIdx.x >> 6] = bpm;
#pragma unroll
            for (int d = 0; d < 3; ++d) s_red[2 + d][threadIdx.x >> 6] = msum[d];
            s_red[5][threadIdx.x >> 6] = nzlo;
            s_red[6][threadIdx.x >> 6] = fzhi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            work.bound[it] = make_float2(fmaxf(fmaxf(s_red[0][0], s_red[0][1]), fmaxf(s_red[0][2], s_red[0][3])),
                                         fmaxf(fmaxf(s_red[1][0], s_red[1][1]), fmaxf(s_red[1][2], s_red[1][3])));
            const float inv = cnt > 0 ? 1.f / (float)cnt : 0.f;
            float mean[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) mean[d] = ((s_red[2 + d][0] + s_red[2 + d][1]) + (s_red[2 + d][2] + s_red[2 + d][3])) * inv;
            work.ref[it] = make_float4(mean[0], mean[1], mean[2], 0.f);
            const int lo = -(int)fmaxf(fmaxf(s_red[5][0], s_red[5][1]), fmaxf(s_red[5][2], s_red[5][3]));
            const int hi = (int)fmaxf(fmaxf(s_red[6][0], s_red[6][1]), fmaxf(s_red[6][2], s_red[6][3]));
            work.shift[it] = make_int4(0, 0, 0, pack_zrange(lo, hi));
        }
        __syncthreads();
    }
}

template <int ORDER, bool COULOMB, bool G2P, int MG>
__global__ void __launch_bounds__(256, 3) k_super(DevParticles P, DevGrid src, DevGrid dst, SuperWork work,
                                                  Phys ph, bool store_vc)
{
    constexpr int T = SB_CELLS + 3 + 2 * MG;          // tile edge (nodes)
    constexpr int TN = T * T * T;
    static_assert(T <= 96, "walk_axis takes tile bases below 96");
    // grid blocks a tile overlaps per axis: ((o & 3) + T - 1) / 4 + 1 <= 4 for T <= 13
    constexpr int NF = SB_BLOCKS + 2;
    static_assert((3 + T - 1) / 4 + 1 <= NF, "tile overlaps more grid blocks than s_touch holds");
    constexpr int W = Spline<ORDER>::W;
    constexpr int NVR = COULOMB ? 6 : 3;
    // accumulators: per tile node (mass, p_x) and (p_y, p_z) as packed fixed point
    __shared__ unsigned long long s_acc[2][TN];
    // gather velocities, component-major: friction velocity in rows 0..2, collision
    // velocity in rows 3..5 with Coulomb walls
    __shared__ float s_vel[G2P ? NVR : 1][G2P ? TN : 1];
    __shared__ int s_touch[NF * NF * NF];
    __shared__ float s_red[7][4];

    const int n = P.n;
    const int tid = threadIdx.x;
    const int nit = *work.n_items;

    // XCD-aware: each XCD takes runs of consecutive items (xcd_item_slot)
    const int G = gridDim.x;
    const int slot = xcd_item_slot();
    for (int r0 = 0; r0 < nit; r0 += G) {
        const int it = r0 + slot;
        if (it >= nit) continue;
        const int4 item = work.items[it];
        const int first = item.y, cnt = min(item.z, SB_MAX_PARTICLES);
        int sc[3];
        super_coords(item.x, work.nsx, work.nsy, sc);
        // the tile follows the item's particles: its origin moves by the rounded mean
        // drift since binning that the previous substep measured, so that a body
        // moving as a whole stays in its tiles between binnings
        const int4 shift = work.shift[it];
        const int sh[3] = {shift.x, shift.y, shift.z};
        int o[3], tb[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o[a] = sc[a] * SB_CELLS - SB_CELLS - MG + sh[a];
            tb[a] = o[a] >> 2;                        // first grid block the tile overlaps
        }
        // node planes (tile z) in use: the gather reads [g_lo, g_hi], the stencils
        // of the positions the previous substep scattered from; the scatter adds to
        // [r_lo, r_hi], a plane more each way for this substep's motion.  Particles
        // beyond them take the per-node gather and the float scatter.
        const int g_lo = max(zrange_lo(shift.w) - o[2], 0), g_hi = min(zrange_hi(shift.w) - o[2] + W - 1, T - 1);
        const int r_lo = max(zrange_lo(shift.w) - o[2] - 1, 0), r_hi = min(zrange_hi(shift.w) - o[2] + W, T - 1);

        // ---- stage the tile's planes in use (all loads issued before the first
        // LDS store)
        int ltid = tid;
        asm volatile("" : "+v"(ltid));
        const int r_beg = r_lo * (T * T), r_end = (r_hi + 1) * (T * T);
        {
            constexpr int NR = (TN + 255) / 256;
            float4 va[NR], vb[NR];
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                va[k] = make_float4(0.f, 0.f, 0.f, 0.f);
                vb[k] = va[k];
                const int e = r_beg + ltid + 256 * k;
                if (G2P && e < r_end) {
                    const int pz = e / (T * T);
                    const int i = o[0] + e % T, j = o[1] + (e / T) % T, kk = o[2] + pz;
                    if (pz >= g_lo && pz <= g_hi && in_grid<3>(ph, i, j, kk)) {
                        const int addr = node_addr<3>(src, i, j, kk);
                        va[k] = src.a[addr];
                        if constexpr (COULOMB) vb[k] = src.vc[addr];
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const int e = r_beg + ltid + 256 * k;
                if (e >= r_end) continue;
                s_acc[0][e] = 0ull;
                s_acc[1][e] = 0ull;
                if constexpr (G2P) {
                    s_vel[0][e] = va[k].y;
                    s_vel[1][e] = va[k].z;
                    s_vel[2][e] = va[k].w;
                    if constexpr (COULOMB) {
                        s_vel[3 % NVR][e] = vb[k].x;
                        s_vel[4 % NVR][e] = vb[k].y;
                        s_vel[5 % NVR][e] = vb[k].z;
                    }
                }
            }
        }
        if (tid < NF * NF * NF) s_touch[tid] = 0;

        // ---- the item's fixed-point scales, from its particles' bounds one substep
        // earlier plus the momentum gravity adds in a substep (so that a body at
        // rest has a scale for its first substep)
        // largest stencil weight: 0.75^3 (quadratic), (2/3)^3 (cubic), rounded up
        const int cap = fixed_cap(cnt, ORDER == 3 ? 0.2963f : 0.4219f);
        const float lim = ldexpf(1.f, cap);
        const float2 prev = work.bound[it];
        const float scale_m = headroom_scale(prev.x, cap);
        const float scale_p = headroom_scale(prev.y + prev.x * (fabsf(ph.gravity) * ph.dt), cap);
        lds_barrier();

        float bound_m = 0.f, bound_p = 0.f, msum[3] = {0.f, 0.f, 0.f};
        int zmn = 1 << 20, zmx = -(1 << 20);
        const int npass = (cnt + 63) >> 6;
        for (int pass = ((tid >> 6) + it) & 3; pass < npass; pass += 4) {
            const int q = pass * 64 + (tid & 63);
            if (q >= cnt) continue;
            const int p = first + q;
            float x[3], v[3], C[9], tau[9], Jp;
#pragma unroll
            for (int d = 0; d < 3; ++d) x[d] = P.x[d * n + p];
            Jp = P.Jp[p];
            const float vol = P.vol[p];
            const float pm = P.mass[p];

            if constexpr (G2P) {
                // g2p.glsl: gather, B (C = D^-1 B), position, F_E, plasticity
                float xg[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) xg[d] = x[d] * ph.inv_dx;
                Stencil<3, ORDER> st;
                st.init(xg);
                bool tile_ok = true, inside;
                int rb[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    rb[a] = st.base[a] - o[a];
                    tile_ok &= rb[a] >= 0 && rb[a] + W <= T;
                }
                tile_ok &= rb[2] >= g_lo && rb[2] + W - 1 <= g_hi;
                inside = tile_ok;
                {
                    const int lim3[3] = {ph.W, ph.H, ph.D};
#pragma unroll
                    for (int a = 0; a < 3; ++a) inside &= st.base[a] >= 0 && st.base[a] + W - 1 <= lim3[a];
                }
                float nv[3] = {0.f, 0.f, 0.f}, nB[9], nx[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int e = 0; e < 9; ++e) nB[e] = 0.f;
                if (inside) {
                    // the node weight is a product of per-axis weights, so the sums
                    // over the W^3 nodes factor axis by axis (as in k_tiled).  Along z,
                    // sum_s w_s (d_0 - s) v_s = d_0 sum_s w_s v_s - sum_s s w_s v_s: the
                    // second sum (one term fewer) is carried instead of the weighted
                    // distances, and the first is the velocity sum itself
                    float wd[2][W], kw[W];
#pragma unroll
                    for (int a = 0; a < 2; ++a)
#pragma unroll
                        for (int s2 = 0; s2 < W; ++s2) wd[a][s2] = -st.w[a][s2] * (st.d[a][s2] * ph.dx);
#pragma unroll
                    for (int s2 = 0; s2 < W; ++s2) kw[s2] = (float)s2 * st.w[2][s2];
                    float nc[3] = {0.f, 0.f, 0.f};
                    // the stencil's first node in the tile; the others sit at constant
                    // offsets from it (no per-node index arithmetic)
                    const float* sv = &s_vel[0][0] + ((rb[2] * T + rb[1]) * T + rb[0]);
#pragma unroll
                    for (int a0 = 0; a0 < W; ++a0) {
                        float S1[3] = {0.f, 0.f, 0.f}, D1[3] = {0.f, 0.f, 0.f}, D2[3] = {0.f, 0.f, 0.f},
                              C1[3] = {0.f, 0.f, 0.f};
#pragma unroll
                        for (int a1 = 0; a1 < W; ++a1) {
                            float S2[3] = {0.f, 0.f, 0.f}, E2[3] = {0.f, 0.f, 0.f}, C2[3] = {0.f, 0.f, 0.f};
#pragma unroll
                            for (int a2 = 0; a2 < W; ++a2) {
                                const int off = (a2 * T + a1) * T + a0;
                                const float w2 = st.w[2][a2];
#pragma unroll
                                for (int d = 0; d < 3; ++d) {
                                    const float vF = sv[d * TN + off];
                                    S2[d] += w2 * vF;
                                    if (a2 > 0) E2[d] += kw[a2] * vF;
                                    if constexpr (COULOMB) C2[d] += w2 * sv[((3 + d) % NVR) * TN + off];
                                }
                            }
                            const float w1 = st.w[1][a1], wd1 = wd[1][a1];
#pragma unroll
                            for (int d = 0; d < 3; ++d) {
                                S1[d] += w1 * S2[d];
                                D1[d] += wd1 * S2[d];
                                D2[d] += w1 * E2[d];
                                if constexpr (COULOMB) C1[d] += w1 * C2[d];
                            }
                        }
                        const float w0 = st.w[0][a0], wd0 = wd[0][a0];
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            nv[d] += w0 * S1[d];
                            nB[d * 3 + 0] += wd0 * S1[d];
                            nB[d * 3 + 1] += w0 * D1[d];
                            nB[d * 3 + 2] += w0 * D2[d];
                            if constexpr (COULOMB) nc[d] += w0 * C1[d];
                        }
                    }
                    // z: -dx (d_0 sum w v - sum s w v)
#pragma unroll
                    for (int d = 0; d < 3; ++d) nB[d * 3 + 2] = (nB[d * 3 + 2] - st.d[2][0] * nv[d]) * ph.dx;
                    float sw[3], mi[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        sw[a] = 0.f;
                        mi[a] = 0.f;
#pragma unroll
                        for (int s2 = 0; s2 < W; ++s2) {
                            sw[a] += st.w[a][s2];
                            mi[a] += st.w[a][s2] * (float)(st.base[a] + s2);
                        }
                    }
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        float m = mi[d];
#pragma unroll
                        for (int a = 0; a < 3; ++a)
                            if (a != d) m *= sw[a];
                        nx[d] = m * ph.dx + ph.dt * (COULOMB ? nc[d] : nv[d]);
                    }
                } else {
                    // stencil reaching past the grid or the tile: per-node tests,
                    // from the tile where it holds the node, else from HBM
#pragma unroll
                    for (int a0 = 0; a0 < W; ++a0)
#pragma unroll
                        for (int a1 = 0; a1 < W; ++a1)
#pragma unroll
                            for (int a2 = 0; a2 < W; ++a2) {
                                const int idx[3] = {st.base[0] + a0, st.base[1] + a1, st.base[2] + a2};
                                if (!in_grid<3>(ph, idx[0], idx[1], idx[2])) continue;
                                const float w = st.w[0][a0] * st.w[1][a1] * st.w[2][a2];
                                const float dp[3] = {st.d[0][a0] * ph.dx, st.d[1][a1] * ph.dx, st.d[2][a2] * ph.dx};
                                float vF[3], vC[3];
                                if (tile_ok) {
                                    const int e = ((rb[2] + a2) * T + (rb[1] + a1)) * T + (rb[0] + a0);
#pragma unroll
                                    for (int d = 0; d < 3; ++d) {
                                        vF[d] = s_vel[d][e];
                                        vC[d] = COULOMB ? s_vel[(3 + d) % NVR][e] : vF[d];
                                    }
                                } else {
                                    const int addr = node_addr<3>(src, idx[0], idx[1], idx[2]);
                                    const float4 a4 = src.a[addr];
                                    vF[0] = a4.y; vF[1] = a4.z; vF[2] = a4.w;
                                    if constexpr (COULOMB) {
                                        const float4 c4 = src.vc[addr];
                                        vC[0] = c4.x; vC[1] = c4.y; vC[2] = c4.z;
                                    } else {
                                        vC[0] = vF[0]; vC[1] = vF[1]; vC[2] = vF[2];
                                    }
                                }
#pragma unroll
                                for (int r = 0; r < 3; ++r) {
                                    nv[r] += vF[r] * w;
#pragma unroll
                                    for (int c = 0; c < 3; ++c) nB[r * 3 + c] -= vF[r] * (w * dp[c]);
                                    nx[r] += (ph.dt * vC[r] + (float)idx[r] * ph.dx) * w;
                                }
                            }
                }
                float M[9], Fo[9], Fh[9], F[9];
#pragma unroll
                for (int e = 0; e < 9; ++e) {
                    C[e] = ph.Dinv * nB[e];
                    M[e] = ((e % 4) == 0 ? 1.f : 0.f) + ph.dt * C[e];
                    Fo[e] = P.F[e * n + p];
                }
                matmul<3>(M, Fo, Fh);
                // an elastic F^ (every singular value inside the clamp range, det > 0)
                // skips the SVD: F_E' = F^, J_P' = J_P, stress from the left stretch
                float elo2, ehi2, Jh;
                elastic_bounds(ph.theta0, ph.theta1, &elo2, &ehi2);
                Sym3 vmi, bmv;
                if (elastic_stretch(Fh, elo2, ehi2, &vmi, &bmv, &Jh)) {
#pragma unroll
                    for (int e = 0; e < 9; ++e) F[e] = Fh[e];
                    const float harden = expf(ph.xi * (1.f - Jp));
                    elastic_kirchhoff(bmv, Jh, 2.f * (ph.mu * harden), ph.lambda * harden, tau);
                } else {
                    float U[9], sig[3], V[9], scv[3];
                    svd3(Fh, U, sig, V);
                    float ratio = 1.f, J = 1.f;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        float sv = fminf(fmaxf(sig[d], ph.theta0), ph.theta1);
                        sv = fmaxf(sv, 1e-6f);
                        scv[d] = sv;
                        ratio *= sig[d] / sv;
                        J *= sv;
                    }
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            float sacc = (U[r * 3] * scv[0]) * V[c * 3];
#pragma unroll
                            for (int qq = 1; qq < 3; ++qq) sacc += (U[r * 3 + qq] * scv[qq]) * V[c * 3 + qq];
                            F[r * 3 + c] = sacc;
                        }
                    Jp = Jp * ratio;
                    // stress of the new state: U diag(2 mu (s-1) s + la (J-1) J) U^T
                    const float harden = expf(ph.xi * (1.f - Jp));
                    const float tmu = 2.f * (ph.mu * harden);
                    const float sI = (ph.lambda * harden) * (J - 1.f) * J;
                    float tdiag[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) tdiag[d] = tmu * (scv[d] - 1.f) * scv[d] + sI;
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            float sacc = U[r * 3] * tdiag[0] * U[c * 3];
#pragma unroll
                            for (int qq = 1; qq < 3; ++qq) sacc += U[r * 3 + qq] * tdiag[qq] * U[c * 3 + qq];
                            tau[r * 3 + c] = sacc;
                        }
                }
#pragma unroll
                for (int d = 0; d < 3; ++d) { x[d] = nx[d]; v[d] = nv[d]; }
                // v and C only when a reader follows (the next substep's gather
                // recomputes both from the grid)
#pragma unroll
                for (int d = 0; d < 3; ++d) P.x[d * n + p] = x[d];
#pragma unroll
                for (int e = 0; e < 9; ++e) P.F[e * n + p] = F[e];
                P.Jp[p] = Jp;
                if (store_vc) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) P.v[d * n + p] = v[d];
#pragma unroll
                    for (int e = 0; e < 9; ++e) P.C[e * n + p] = C[e];
                }
            } else {
                // stress from the stored state (p2g.glsl:44-59)
                float F[9], U[9], sig[3], V[9];
#pragma unroll
                for (int d = 0; d < 3; ++d) v[d] = P.v[d * n + p];
#pragma unroll
                for (int e = 0; e < 9; ++e) { C[e] = P.C[e * n + p]; F[e] = P.F[e * n + p]; }
                svd3(F, U, sig, V);
                kirchhoff<3>(ph, F, U, V, det<3>(F), Jp, tau);
            }

#pragma unroll
            for (int d = 0; d < 3; ++d) msum[d] += x[d] * ph.inv_dx - (float)(sc[d] * SB_CELLS);
            {
                const int bz = Spline<ORDER>::base(x[2] * ph.inv_dx);
                zmn = min(zmn, bz);
                zmx = max(zmx, bz);
            }
            if (vol == 0.f) continue;                                 // p2g.glsl:42
            // ---- scatter (p2g.glsl): into the tile when the stencil lies in it and
            // the contributions fit the item's scale, else global float atomics
            float mvp[3], A[9];
            const float bp = scatter_inputs<ORDER>(ph, pm, vol, v, C, tau, mvp, A);
            const bool finite = bp < 1.0e37f && fabsf(pm) < 1.0e37f;
            if (finite) {
                bound_m = fmaxf(bound_m, fabsf(pm));
                bound_p = fmaxf(bound_p, bp);
            }
            float xg[3];
            int base[3], rb[3];
            bool tile_ok = finite && bp * scale_p <= lim && fabsf(pm) * scale_m <= lim;
            bool inside = true;
            {
                const int lim3[3] = {ph.W, ph.H, ph.D};
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    xg[d] = x[d] * ph.inv_dx;
                    base[d] = Spline<ORDER>::base(xg[d]);
                    rb[d] = base[d] - o[d];
                    tile_ok &= rb[d] >= 0 && rb[d] + W <= T;
                    inside &= base[d] >= 0 && base[d] + W - 1 <= lim3[d];
                }
                tile_ok &= rb[2] >= r_lo && rb[2] + W - 1 <= r_hi;
            }
            float wa[3][W], xa[3][W];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                float dd[W];
                Spline<ORDER>::weights(xg[d], base[d], wa[d], dd);
#pragma unroll
                for (int s2 = 0; s2 < W; ++s2) xa[d][s2] = -dd[s2] * ph.dx;
            }
            if (tile_ok && inside) {
                // whole stencil in the tile and the grid: scaled inputs, one rounding
                // per component, two packed 64-bit LDS adds per node.  Each lane walks
                // its stencil from a rotated start: at step t it adds to node
                // b + ((t + rho) mod W) per axis, whose residue mod W, b + rho + t, moves
                // in step for every lane; rho is chosen so that b + rho mod W is the
                // lane's own phase (its index's base-W digits), and lanes of different
                // phases then never add to the same word at the same step, whatever
                // their bases (27 phases for the quadratic stencil: lanes 27 apart share
                // one)
                const float pms = pm * scale_m;
                float mvs[3], As[9];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    mvs[r] = mvp[r] * scale_p;
#pragma unroll
                    for (int c = 0; c < 3; ++c) As[r * 3 + c] = A[r * 3 + c] * scale_p;
                }
                float wr[3][W], xr[3][W];
                int er[3][W];
                {
                    const int strd[3] = {1, T, T * T};
                    unsigned lr = tid & 63;
                    asm volatile("" : "+v"(lr));
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const unsigned q = small_div<W>(lr);      // rb[d] >= 0 in the tile
                        walk_axis<W>((int)(lr - q * W), rb[d], strd[d], wa[d], xa[d], wr[d], xr[d], er[d]);
                        lr = q;
                    }
                }
#pragma unroll
                for (int a0 = 0; a0 < W; ++a0)
#pragma unroll
                    for (int a1 = 0; a1 < W; ++a1) {
                        const float w01 = wr[0][a0] * wr[1][a1];
                        float sb[3];
#pragma unroll
                        for (int r = 0; r < 3; ++r)
                            sb[r] = (mvs[r] + As[r * 3] * xr[0][a0]) + As[r * 3 + 1] * xr[1][a1];
#pragma unroll
                        for (int a2 = 0; a2 < W; ++a2) {
                            const float w = w01 * wr[2][a2];
                            int qb[4];
                            qb[0] = fixed_bits(w, pms);
#pragma unroll
                            for (int r = 0; r < 3; ++r) qb[1 + r] = fixed_bits(w, sb[r] + As[r * 3 + 2] * xr[2][a2]);
                            const int e = er[0][a0] + er[1][a1] + er[2][a2];
                            atomicAdd(&s_acc[0][e], pack_fixed(qb[0], qb[1]));
                            atomicAdd(&s_acc[1][e], pack_fixed(qb[2], qb[3]));
                        }
                    }
            } else {
#pragma unroll
                for (int a0 = 0; a0 < W; ++a0)
#pragma unroll
                    for (int a1 = 0; a1 < W; ++a1)
#pragma unroll
                        for (int a2 = 0; a2 < W; ++a2) {
                            const int idx[3] = {base[0] + a0, base[1] + a1, base[2] + a2};
                            if (!in_grid<3>(ph, idx[0], idx[1], idx[2])) continue;
                            const float w = wa[0][a0] * wa[1][a1] * wa[2][a2];
                            const float xi[3] = {xa[0][a0], xa[1][a1], xa[2][a2]};
                            float vals[4];
                            node_contrib<3>(w, pm, mvp, A, xi, vals);
                            if (tile_ok) {
                                const int e = ((rb[2] + a2) * T + (rb[1] + a1)) * T + (rb[0] + a0);
                                const int q0 = fixed_bits(vals[0], scale_m), q1 = fixed_bits(vals[1], scale_p);
                                const int q2 = fixed_bits(vals[2], scale_p), q3 = fixed_bits(vals[3], scale_p);
                                atomicAdd(&s_acc[0][e], pack_fixed(q0, q1));
                                atomicAdd(&s_acc[1][e], pack_fixed(q2, q3));
                            } else {
                                float* dstp = reinterpret_cast<float*>(&dst.a[node_addr<3>(dst, idx[0], idx[1], idx[2])]);
#pragma unroll
                                for (int c = 0; c < 4; ++c) unsafeAtomicAdd(dstp + c, vals[c]);
                                dst.flags[block_of<3>(dst, idx[0], idx[1], idx[2])] = 1u;
                            }
                        }
            }
        }

        // ---- this substep's bounds, for the next substep's scales
        bound_m = wave_max(bound_m);
        bound_p = wave_max(bound_p);
#pragma unroll
        for (int d = 0; d < 3; ++d) msum[d] = wave_sum(msum[d]);
        const float nzmn = wave_max(-(float)zmn), fzmx = wave_max((float)zmx);
        if ((tid & 63) == 0) {
            s_red[0][tid >> 6] = bound_m;
            s_red[1][tid >> 6] = bound_p;
#pragma unroll
            for (int d = 0; d < 3; ++d) s_red[2 + d][tid >> 6] = msum[d];
            s_red[5][tid >> 6] = nzmn;
            s_red[6][tid >> 6] = fzmx;
        }
        lds_barrier();
        // z-range of the stencil bases this substep scattered from (every particle's
        // new position): the next substep's gather planes, and with [r_lo, r_hi] the
        // planes to flush
        const int zlo = -(int)fmaxf(fmaxf(s_red[5][0], s_red[5][1]), fmaxf(s_red[5][2], s_red[5][3]));
        const int zhi = (int)fmaxf(fmaxf(s_red[6][0], s_red[6][1]), fmaxf(s_red[6][2], s_red[6][3]));
        if (tid == 0) {
            work.bound[it] = make_float2(fmaxf(fmaxf(s_red[0][0], s_red[0][1]), fmaxf(s_red[0][2], s_red[0][3])),
                                         fmaxf(fmaxf(s_red[1][0], s_red[1][1]), fmaxf(s_red[1][2], s_red[1][3])));
            // the next substep's tile offset, from the mean of the positions this
            // substep scattered from
            const float4 ref = work.ref[it];
            const float rf[3] = {ref.x, ref.y, ref.z};
            const float inv = 1.f / (float)max(cnt, 1);
            int ns[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
                ns[d] = drift_shift(((s_red[2 + d][0] + s_red[2 + d][1]) + (s_red[2 + d][2] + s_red[2 + d][3])) * inv - rf[d]);
            work.shift[it] = make_int4(ns[0], ns[1], ns[2], pack_zrange(zlo, zhi));
        }

        // ---- flush: fixed point back to FP32, one global atomic per touched node
        // component, lanes over (node, component) pairs with the component fastest
        const float inv_m = 1.f / scale_m, inv_p = 1.f / scale_p;
        bool tile_in = true;
        {
            const int lim3[3] = {ph.W, ph.H, ph.D};
#pragma unroll
            for (int a = 0; a < 3; ++a) tile_in &= o[a] >= 0 && o[a] + T - 1 <= lim3[a];
        }
        // four passes' LDS words read together, so that their latencies overlap
        constexpr int FU = 4;
        const int c = tid & 3;
        const int f_beg = max(zlo - o[2], r_lo) * (T * T * 4), f_end = (min(zhi - o[2] + W - 1, r_hi) + 1) * (T * T * 4);
        for (int pe0 = f_beg + tid; pe0 < f_end; pe0 += 256 * FU) {
            int acc[FU], mass[FU];
#pragma unroll
            for (int u = 0; u < FU; ++u) {
                const int e = min(pe0 + 256 * u, f_end - 1) >> 2;
                const unsigned long long sum = s_acc[c >> 1][e];
                acc[u] = (c & 1) ? unpack_hi(sum) : unpack_lo(sum);
                mass[u] = unpack_lo(s_acc[0][e]);
            }
#pragma unroll
            for (int u = 0; u < FU; ++u) {
                const int pe = pe0 + 256 * u;
                // a node whose mass rounded to zero flushes no momentum either: mass
                // and momentum round independently, and a momentum unit over a
                // float-path mass far below the mass unit would give the node any
                // velocity
                if (pe >= f_end || acc[u] == 0 || mass[u] == 0) continue;
                const int e = pe >> 2;
                const int i = o[0] + e % T, j = o[1] + (e / T) % T, k = o[2] + e / (T * T);
                if (!tile_in && !in_grid<3>(ph, i, j, k)) continue;
                float* dstp = reinterpret_cast<float*>(&dst.a[node_addr<3>(dst, i, j, k)]) + c;
                unsafeAtomicAdd(dstp, (float)acc[u] * (c == 0 ? inv_m : inv_p));
                s_touch[(((k >> 2) - tb[2]) * NF + ((j >> 2) - tb[1])) * NF + ((i >> 2) - tb[0])] = 1;
            }
        }
        lds_barrier();
        int ftid = tid;
        asm volatile("" : "+v"(ftid));
        if (ftid < NF * NF * NF && s_touch[ftid]) {
            const int bx = tb[0] + ftid % NF;
            const int by = tb[1] + (ftid / NF) % NF;
            const int bz = tb[2] + ftid / (NF * NF);
            dst.flags[(bz * dst.nby + by) * dst.nbx + bx] = 1u;
        }
        lds_barrier();
    }
}

// Super-block binning keys: the cell containing the particle, clamped into the
// grid, numbered super-block-major: key = ((sz * nsy + sy) * nsx + sx) * 512 +
// (cell within the super-block, z slowest and x fastest), or with
// cells = false the super-block number alone.
__global__ void __launch_bounds__(256) k_cell_keys_super(const float* x, int n, float inv_dx, int W, int H, int D,
                                                         int nsx, int nsy, bool cells, uint32_t* keys)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int lim[3] = {W, H, D};
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] = min(max(cell_floor(x[d * n + p] * inv_dx), 0), lim[d]);
    keys[p] = super_key(c[0], c[1], c[2], nsx, nsy, cells);
}

// ---------------------------------------------------------------------------
hipError_t launch_cell_keys_super(const float* x, int n, const Phys& ph, int nsx, int nsy, bool cells,
                                  uint32_t* keys, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    k_cell_keys_super<<<(n + 255) / 256, 256, 0, st>>>(x, n, ph.inv_dx, ph.W, ph.H, ph.D, nsx, nsy, cells, keys);
    return hipGetLastError();
}

static int super_grid(const SuperWork& w)
{
    constexpr int R = 8 * XCD_RUN;           // k_super maps items in rounds of R (xcd_item_slot)
    int g = w.max_items < 4096 ? w.max_items : 4096;
    g = (g + R - 1) / R * R;
    return g > 0 ? g : R;
}

hipError_t launch_item_bounds(int order, const DevParticles& p, const SuperWork& w, const Phys& ph,
                              hipStream_t st)
{
    if (p.n == 0) return hipSuccess;
    if (order == 3) k_item_bounds<3><<<super_grid(w), 256, 0, st>>>(p, w, ph);
    else            k_item_bounds<2><<<super_grid(w), 256, 0, st>>>(p, w, ph);
    return hipGetLastError();
}

hipError_t launch_super(const LaunchCfg& c, int margin, bool g2p, bool store_vc, const DevParticles& p,
                        const DevGrid& src, const DevGrid& dst, const SuperWork& w, const Phys& ph, hipStream_t st)
{
    if (p.n == 0) return hipSuccess;
    if (c.dim != 3) return hipErrorInvalidValue;
    const dim3 grid(super_grid(w)), blk(256);
    // the P2G-only pass gathers nothing, so the wall kind does not enter it
    const bool coul = g2p && c.boundary == MPM_BOUNDARY_COULOMB_MARGIN;
#define SCASE(O, CL, G, M) \
    if (c.order == O && coul == CL && g2p == G && margin == M) { k_super<O, CL, G, M><<<grid, blk, 0, st>>>(p, src, dst, w, ph, store_vc); return hipGetLastError(); }
    SCASE(2, false, true, 0) SCASE(2, false, true, 1) SCASE(2, true, true, 0) SCASE(2, true, true, 1)
    SCASE(2, false, false, 0) SCASE(2, false, false, 1)
    SCASE(3, false, true, 0) SCASE(3, false, true, 1) SCASE(3, true, true, 0) SCASE(3, true, true, 1)
    SCASE(3, false, false, 0) SCASE(3, false, false, 1)
#undef SCASE
    return hipErrorInvalidValue;
}

}  // namespace mpm
