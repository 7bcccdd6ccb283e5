// mpm_kernels.hip — the substep stages as separate kernels, one per dispatch of
// the reference's GPUSolver::Step (src/gpu/GPUSolver.cpp:63-80):
//   k_p2g           shaders/{2d,3d}/p2g.glsl          per particle, global float atomics
//   k_grid_update   shaders/{2d,3d}/grid_update.glsl  per node of every touched block
//   k_g2p           shaders/{2d,3d}/g2p.glsl          per particle gather + F_E update
//   k_reset         shaders/{2d,3d}/reset.glsl        clears touched blocks only
// plus k_grid_update_clear, the grid pass of the fused pipeline.
//
// Layout: particles SoA component-major (coalesced 4-byte loads per field), grid
// block-major float4 nodes (mass, momentum|velocity), 64 nodes per block so a
// 64-lane wavefront owns one block in the grid passes.
#include "mpm_kernels.hpp"

namespace mpm {

template <int DIM> struct Walls;
// shaders/2d/grid_update.glsl:22-33 — right, bottom, left, top.
template <> struct Walls<2> {
    static constexpr int NW = 4;
    __device__ static inline void get(const Phys& P, int b, float* pt, float* nrm)
    {
        float w = (float)P.W * P.dx, h = (float)P.H * P.dx, m = P.margin * P.dx;
        const float px[4] = {w - m, m, m, w - m}, py[4] = {m, m, h - m, h - m};
        const float nx[4] = {-1.f, 0.f, 1.f, 0.f}, ny[4] = {0.f, 1.f, 0.f, -1.f};
        pt[0] = px[b]; pt[1] = py[b]; nrm[0] = nx[b]; nrm[1] = ny[b];
    }
};
// shaders/3d/grid_update.glsl:24-39 — +X, -X, +Y, -Y, +Z, -Z.
template <> struct Walls<3> {
    static constexpr int NW = 6;
    __device__ static inline void get(const Phys& P, int b, float* pt, float* nrm)
    {
        float w = (float)P.W * P.dx, h = (float)P.H * P.dx, d = (float)P.D * P.dx;
        float m = P.margin * P.dx;
        int axis = b >> 1;
        bool hi = (b & 1) == 0;
        float ext = axis == 0 ? w : (axis == 1 ? h : d);
#pragma unroll
        for (int a = 0; a < 3; ++a) { pt[a] = 0.f; nrm[a] = 0.f; }
        pt[axis] = hi ? ext - m : m;
        nrm[axis] = hi ? -1.f : 1.f;
    }
};

// shaders/2d/grid_update.glsl:17-80: sequential predictive projection against each
// wall, then Coulomb friction for every wall that was hit.
template <int DIM>
__device__ inline void collide(const Phys& P, const float* loc, const float* vel, float* vcol,
                               float* vfric)
{
    constexpr int NW = Walls<DIM>::NW;
    float hitn[NW][DIM];
    int nh = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) vcol[d] = vel[d];
#pragma unroll
    for (int b = 0; b < NW; ++b) {
        float pt[DIM], nrm[DIM];
        Walls<DIM>::get(P, b, pt, nrm);
        float dist = 0.f, test = 0.f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            dist += nrm[d] * (loc[d] - pt[d]);
            test += nrm[d] * ((loc[d] + P.dt * vcol[d]) - pt[d]);
        }
        float delta = test - fminf(dist, 0.f);
        if (delta < 0.f) {
            float q = delta / P.dt;
#pragma unroll
            for (int d = 0; d < DIM; ++d) vcol[d] -= q * nrm[d];
#pragma unroll
            for (int d = 0; d < DIM; ++d) hitn[nh][d] = nrm[d];
            ++nh;
        }
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) vfric[d] = vcol[d];
    for (int c = 0; c < nh; ++c) {
        float nd = 0.f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) nd += hitn[c][d] * vfric[d];
        float vt[DIM], vn2 = 0.f, dc2 = 0.f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            vt[d] = vcol[d] - hitn[c][d] * nd;
            vn2 += vt[d] * vt[d];
            float df = vcol[d] - vel[d];
            dc2 += df * df;
        }
        float vt_norm = sqrtf(vn2);
        if (vt_norm <= 1e-7f) continue;
        float k = fminf(vt_norm, P.friction * sqrtf(dc2));
#pragma unroll
        for (int d = 0; d < DIM; ++d) vfric[d] -= k * (vt[d] / vt_norm);
    }
}

// Grid update of one node (shaders/2d/grid_update.glsl:82-109): velocity from
// momentum, gravity along y, boundary.  Returns friction velocity in `vf` and
// collision velocity in `vc`; light nodes keep zero velocity (the reference skips
// them, leaving reset's zeros).
// APIC: the CPU path's update v = mv/m + dt (g - f/m) with the explicit grid force
// f (src/cpu/Grid.cpp:165-184) instead of v = mv/m + dt g.
template <int DIM, bool COULOMB, bool APIC = false>
__device__ inline void node_update(const Phys& P, int i, int j, int k, float4 a, float* vf,
                                   float* vc, float4 force = make_float4(0.f, 0.f, 0.f, 0.f))
{
    float m = a.x;
#pragma unroll
    for (int d = 0; d < DIM; ++d) { vf[d] = 0.f; vc[d] = 0.f; }
    if (m < P.mass_eps || !(m > 0.f)) return;
    float vel[DIM];
    vel[0] = a.y / m;
    vel[1] = a.z / m;
    if constexpr (DIM == 3) vel[2] = a.w / m;
    if constexpr (APIC) {
        const float fc[3] = {force.x, force.y, force.z};
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            float acc = (d == 1 ? P.gravity : 0.f) + ((-fc[d]) / m);
            vel[d] += acc * P.dt;
        }
    } else {
        vel[1] += P.gravity * P.dt;
    }
    if constexpr (COULOMB) {
        float loc[DIM];
        loc[0] = (float)i * P.dx;
        loc[1] = (float)j * P.dx;
        if constexpr (DIM == 3) loc[2] = (float)k * P.dx;
        collide<DIM>(P, loc, vel, vc, vf);
    } else {
        int idx[3] = {i, j, k};
        int lim[3] = {P.W, P.H, P.D};
        bool stick = false;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            stick |= ((float)idx[d] < P.margin) || ((float)idx[d] > (float)lim[d] - P.margin);
#pragma unroll
        for (int d = 0; d < DIM; ++d) vf[d] = vc[d] = stick ? 0.f : vel[d];
    }
}

template <int DIM>
__device__ inline void block_node(const DevGrid& g, int blk, int lane, int& i, int& j, int& k)
{
    constexpr int S = BlockShape<DIM>::SHIFT;
    constexpr int M = (1 << S) - 1;
    int bi = blk % g.nbx;
    int rest = blk / g.nbx;
    if constexpr (DIM == 2) {
        i = (bi << S) | (lane & M);
        j = (rest << S) | (lane >> S);
        k = 0;
    } else {
        int bj = rest % g.nby, bk = rest / g.nby;
        i = (bi << S) | (lane & M);
        j = (bj << S) | ((lane >> S) & M);
        k = (bk << S) | (lane >> (2 * S));
    }
}

// ---------------------------------------------------------------------------
// P2G (shaders/{2d,3d}/p2g.glsl): one thread per particle, stress from the SVD of
// F_E, MLS scatter with hardware float atomics.  APIC: the CPU path's rasterisation
// (src/cpu/Grid.cpp:118-163) — APIC momentum without stress, and the explicit force
// grad w * A_p with A_p = V tau (Particles.cpp:21-62) into the force grid.
template <int DIM, int ORDER, bool APIC>
__global__ void __launch_bounds__(256) k_p2g(DevParticles P, DevGrid g, Phys ph)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const int n = P.n;
    float vol = P.vol[p];
    if (!APIC && vol == 0.f) return;                            // p2g.glsl:42
    float pm = P.mass[p];
    float x[DIM], v[DIM], C[DIM * DIM], F[DIM * DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) { x[d] = P.x[d * n + p]; v[d] = P.v[d * n + p]; }
#pragma unroll
    for (int e = 0; e < DIM * DIM; ++e) { C[e] = P.C[e * n + p]; F[e] = P.F[e * n + p]; }
    float Jp = P.Jp[p];

    float U[DIM * DIM], sig[DIM], V[DIM * DIM], tau[DIM * DIM];
    if constexpr (APIC) svd_nonneg<DIM>(F, U, sig, V);
    else svd<DIM>(F, U, sig, V);
    kirchhoff<DIM>(ph, F, U, V, det<DIM>(F), Jp, tau);
    float Vt[DIM * DIM];
#pragma unroll
    for (int e = 0; e < DIM * DIM; ++e) Vt[e] = (APIC && vol == 0.f) ? 0.f : vol * tau[e];

    float xg[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) xg[d] = x[d] * ph.inv_dx;
    Stencil<DIM, ORDER> st;
    st.init(xg);
    constexpr int W = Spline<ORDER>::W;
    const float fdt = ph.Dinv * ph.dt;
#pragma unroll
    for (int a0 = 0; a0 < W; ++a0)
#pragma unroll
        for (int a1 = 0; a1 < W; ++a1)
#pragma unroll
            for (int a2 = 0; a2 < (DIM == 3 ? W : 1); ++a2) {
                int i = st.base[0] + a0, j = st.base[1] + a1;
                int k = DIM == 3 ? st.base[DIM - 1] + a2 : 0;
                if (i < 0 || i > ph.W || j < 0 || j > ph.H) continue;
                if (DIM == 3 && (k < 0 || k > ph.D)) continue;
                float dp[DIM];
                dp[0] = st.d[0][a0] * ph.dx;
                dp[1] = st.d[1][a1] * ph.dx;
                float w = st.w[0][a0] * st.w[1][a1];
                if constexpr (DIM == 3) {
                    dp[DIM - 1] = st.d[DIM - 1][a2] * ph.dx;
                    w = w * st.w[DIM - 1][a2];
                }
                float m = w * pm;
                float mv[DIM];
#pragma unroll
                for (int r = 0; r < DIM; ++r) {
                    float av = 0.f, fv = 0.f;
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        av += C[r * DIM + c] * (-dp[c]);
                        fv += Vt[r * DIM + c] * dp[c];
                    }
                    mv[r] = APIC ? m * (v[r] + av) : m * (v[r] + av) + w * (fv * fdt);
                }
                const int addr = node_addr<DIM>(g, i, j, k);
                float* dst = reinterpret_cast<float*>(&g.a[addr]);
                unsafeAtomicAdd(dst + 0, m);
#pragma unroll
                for (int d = 0; d < DIM; ++d) unsafeAtomicAdd(dst + 1 + d, mv[d]);
                if constexpr (APIC) {
                    // grad w with respect to x_p (Grid.h:77-83), physical units
                    const int aa[3] = {a0, a1, a2};
                    float gw[DIM];
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        float gc = Spline<ORDER>::dN(st.d[c][aa[c]]);
#pragma unroll
                        for (int q = 0; q < DIM; ++q)
                            if (q != c) gc *= st.w[q][aa[q]];
                        gw[c] = gc * ph.inv_dx;
                    }
                    float* fd = reinterpret_cast<float*>(&g.f[addr]);
#pragma unroll
                    for (int r = 0; r < DIM; ++r) {
                        float f = 0.f;
#pragma unroll
                        for (int c = 0; c < DIM; ++c) f += gw[c] * Vt[c * DIM + r];   // glm vec * mat
                        unsafeAtomicAdd(fd + r, f);
                    }
                }
            }
    // mark the (at most 2 per axis) blocks the stencil spans
#pragma unroll
    for (int c0 = 0; c0 < 2; ++c0)
#pragma unroll
        for (int c1 = 0; c1 < 2; ++c1)
#pragma unroll
            for (int c2 = 0; c2 < (DIM == 3 ? 2 : 1); ++c2) {
                int i = min(max(st.base[0] + c0 * (W - 1), 0), ph.W);
                int j = min(max(st.base[1] + c1 * (W - 1), 0), ph.H);
                int k = DIM == 3 ? min(max(st.base[DIM - 1] + c2 * (W - 1), 0), ph.D) : 0;
                g.flags[block_of<DIM>(g, i, j, k)] = 1u;
            }
}

// Grid update (shaders/{2d,3d}/grid_update.glsl) over touched blocks; one wave per block.
template <int DIM, bool COULOMB, bool APIC>
__global__ void __launch_bounds__(64) k_grid_update(DevGrid g, Phys ph)
{
    int blk = blockIdx.x;
    if (g.flags[blk] == 0u) return;
    int lane = threadIdx.x;
    int i, j, k;
    block_node<DIM>(g, blk, lane, i, j, k);
    int addr = (blk << 6) | lane;
    float4 a = g.a[addr];
    float vf[DIM], vc[DIM];
    if constexpr (APIC) node_update<DIM, COULOMB, true>(ph, i, j, k, a, vf, vc, g.f[addr]);
    else node_update<DIM, COULOMB>(ph, i, j, k, a, vf, vc);
    g.a[addr] = make_float4(a.x, vf[0], vf[1], DIM == 3 ? vf[DIM - 1] : 0.f);
    if constexpr (COULOMB) g.vc[addr] = make_float4(vc[0], vc[1], DIM == 3 ? vc[DIM - 1] : 0.f, 0.f);
}

// Reset (shaders/{2d,3d}/reset.glsl) of the touched blocks; clears their flags.
template <int DIM>
__global__ void __launch_bounds__(64) k_reset(DevGrid g)
{
    int blk = blockIdx.x;
    if (g.flags[blk] == 0u) return;
    int addr = (blk << 6) | threadIdx.x;
    g.a[addr] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (g.vc) g.vc[addr] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (g.f) g.f[addr] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (threadIdx.x == 0) g.flags[blk] = 0u;
}

// G2P (shaders/{2d,3d}/g2p.glsl): gather, APIC/MLS C, position, F_E with the
// clamped SVD, plastic determinant.  APIC: the CPU path (Particles.cpp:133-216)
// advances F_E with grad v = sum_i v_col (x) grad w instead of C, decomposed with
// non-negative singular values.
template <int DIM, int ORDER, bool COULOMB, bool APIC>
__global__ void __launch_bounds__(256) k_g2p(DevParticles P, DevGrid g, Phys ph)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const int n = P.n;
    float x[DIM], xg[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) { x[d] = P.x[d * n + p]; xg[d] = x[d] * ph.inv_dx; }
    Stencil<DIM, ORDER> st;
    st.init(xg);
    constexpr int W = Spline<ORDER>::W;
    float nv[DIM], nB[DIM * DIM], nx[DIM], gv[DIM * DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) { nv[d] = 0.f; nx[d] = 0.f; }
#pragma unroll
    for (int e = 0; e < DIM * DIM; ++e) { nB[e] = 0.f; gv[e] = 0.f; }
#pragma unroll
    for (int a0 = 0; a0 < W; ++a0)
#pragma unroll
        for (int a1 = 0; a1 < W; ++a1)
#pragma unroll
            for (int a2 = 0; a2 < (DIM == 3 ? W : 1); ++a2) {
                int i = st.base[0] + a0, j = st.base[1] + a1;
                int k = DIM == 3 ? st.base[DIM - 1] + a2 : 0;
                if (i < 0 || i > ph.W || j < 0 || j > ph.H) continue;
                if (DIM == 3 && (k < 0 || k > ph.D)) continue;
                int idx[3] = {i, j, k};
                const int aa[3] = {a0, a1, a2};
                float dp[DIM];
                dp[0] = st.d[0][a0] * ph.dx;
                dp[1] = st.d[1][a1] * ph.dx;
                float w = st.w[0][a0] * st.w[1][a1];
                if constexpr (DIM == 3) {
                    dp[DIM - 1] = st.d[DIM - 1][a2] * ph.dx;
                    w = w * st.w[DIM - 1][a2];
                }
                int addr = node_addr<DIM>(g, i, j, k);
                float4 a = g.a[addr];
                float vF[3] = {a.y, a.z, a.w};
                float vC[3];
                if constexpr (COULOMB) {
                    float4 c = g.vc[addr];
                    vC[0] = c.x; vC[1] = c.y; vC[2] = c.z;
                } else {
                    vC[0] = vF[0]; vC[1] = vF[1]; vC[2] = vF[2];
                }
#pragma unroll
                for (int r = 0; r < DIM; ++r) {
                    nv[r] += vF[r] * w;
#pragma unroll
                    for (int c = 0; c < DIM; ++c) nB[r * DIM + c] += w * (vF[r] * (-dp[c]));
                    nx[r] += (ph.dt * vC[r] + (float)idx[r] * ph.dx) * w;
                }
                if constexpr (APIC) {
                    float gw[DIM];
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        float gc = Spline<ORDER>::dN(st.d[c][aa[c]]);
#pragma unroll
                        for (int q = 0; q < DIM; ++q)
                            if (q != c) gc *= st.w[q][aa[q]];
                        gw[c] = gc * ph.inv_dx;
                    }
#pragma unroll
                    for (int r = 0; r < DIM; ++r)
#pragma unroll
                        for (int c = 0; c < DIM; ++c) gv[r * DIM + c] += vC[r] * gw[c];   // Particles.cpp:210
                }
            }
    float Cn[DIM * DIM], M[DIM * DIM], Fo[DIM * DIM], Fh[DIM * DIM];
#pragma unroll
    for (int e = 0; e < DIM * DIM; ++e) {
        Cn[e] = ph.Dinv * nB[e];
        M[e] = ((e % (DIM + 1)) == 0 ? 1.f : 0.f) + ph.dt * (APIC ? gv[e] : Cn[e]);
        Fo[e] = P.F[e * n + p];
    }
    matmul<DIM>(M, Fo, Fh);
    float U[DIM * DIM], sig[DIM], V[DIM * DIM], sc[DIM];
    if constexpr (APIC) svd_nonneg<DIM>(Fh, U, sig, V);
    else svd<DIM>(Fh, U, sig, V);
    float ratio = 1.f;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        float s = fminf(fmaxf(sig[d], ph.theta0), ph.theta1);
        s = fmaxf(s, 1e-6f);
        sc[d] = s;
        ratio *= sig[d] / s;
    }
    float Fn[DIM * DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            float s = (U[r * DIM] * sc[0]) * V[c * DIM];
#pragma unroll
            for (int q = 1; q < DIM; ++q) s += (U[r * DIM + q] * sc[q]) * V[c * DIM + q];
            Fn[r * DIM + c] = s;
        }
#pragma unroll
    for (int d = 0; d < DIM; ++d) { P.x[d * n + p] = nx[d]; P.v[d * n + p] = nv[d]; }
#pragma unroll
    for (int e = 0; e < DIM * DIM; ++e) { P.C[e * n + p] = Cn[e]; P.F[e * n + p] = Fn[e]; }
    P.Jp[p] = P.Jp[p] * ratio;
}

// Fused-pipeline grid pass: converts the touched blocks of `cur` to velocities
// and clears the touched blocks of `old` (the grid the last G2P read) so the next
// scatter starts from zero.  One workgroup scans 64 consecutive block flags with one
// coalesced load, compacts the touched ones in LDS, and its four waves then take one
// 64-node block each — the launch is nblocks/64 workgroups, not one per block.
__device__ inline uint64_t lanemask_lt64()
{
    int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <int DIM, bool COULOMB, bool BATCH>
__global__ void __launch_bounds__(256) k_grid_update_clear(DevGrid cur, DevGrid old, Phys ph_one, int nblocks,
                                                          const SceneDesc* scenes, const int* block_scene)
{
    __shared__ int s_list[64];
    __shared__ int s_cnt;
    const int tid = threadIdx.x;
    if (tid < 64) {
        int b = blockIdx.x * 64 + tid;
        int f = 0;
        if (b < nblocks) f = (cur.flags[b] ? 1 : 0) | (old.flags[b] ? 2 : 0);
        uint64_t m = __ballot(f != 0);
        if (f) s_list[__popcll(m & lanemask_lt64())] = (b << 2) | f;
        if (tid == 0) s_cnt = __popcll(m);
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    for (int q = wave; q < s_cnt; q += 4) {
        int e = s_list[q];
        int blk = e >> 2;
        int addr = (blk << 6) | lane;
        if (e & 2) {
            old.a[addr] = make_float4(0.f, 0.f, 0.f, 0.f);
            if constexpr (COULOMB) old.vc[addr] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (lane == 0) old.flags[blk] = 0u;
        }
        if (e & 1) {
            // node coordinates inside the block's own scene
            Phys ph = ph_one;
            DevGrid gs = cur;
            int lb = blk;
            if constexpr (BATCH) {
                const SceneDesc sd = scenes[block_scene[blk]];
                ph = sd.ph;
                gs.nbx = sd.nbx; gs.nby = sd.nby; gs.nbz = sd.nbz;
                lb = blk - sd.block_off;
            }
            int i, j, k;
            block_node<DIM>(gs, lb, lane, i, j, k);
            float4 a = cur.a[addr];
            float vf[DIM], vc[DIM];
            node_update<DIM, COULOMB>(ph, i, j, k, a, vf, vc);
            cur.a[addr] = make_float4(a.x, vf[0], vf[1], DIM == 3 ? vf[DIM - 1] : 0.f);
            if constexpr (COULOMB)
                cur.vc[addr] = make_float4(vc[0], vc[1], DIM == 3 ? vc[DIM - 1] : 0.f, 0.f);
        }
    }
}

template <int DIM>
__global__ void __launch_bounds__(256) k_check_finite(DevParticles P, unsigned long long* bad)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = true;
    if (p < P.n) {
        const int n = P.n;
#pragma unroll
        for (int d = 0; d < DIM; ++d) ok &= isfinite(P.x[d * n + p]) && isfinite(P.v[d * n + p]);
#pragma unroll
        for (int e = 0; e < DIM * DIM; ++e) ok &= isfinite(P.C[e * n + p]) && isfinite(P.F[e * n + p]);
        ok &= isfinite(P.Jp[p]);
    }
    unsigned long long m = __ballot(!ok);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(bad, (unsigned long long)__popcll(m));
}

// ---------------------------------------------------------------------------
// Launchers.
static inline int nblocks_total(const DevGrid& g) { return g.nbx * g.nby * g.nbz; }

hipError_t launch_p2g(const LaunchCfg& c, const DevParticles& p, const DevGrid& g,
                      const Phys& ph, hipStream_t st)
{
    if (p.n == 0) return hipSuccess;
    dim3 grid((p.n + 255) / 256), blk(256);
    const bool apic = ph.scheme == MPM_SCHEME_APIC_FORCE;
#define P2G_CASE(D, O)                                                                   \
    if (c.dim == D && c.order == O) {                                                    \
        if (apic) k_p2g<D, O, true><<<grid, blk, 0, st>>>(p, g, ph);                     \
        else      k_p2g<D, O, false><<<grid, blk, 0, st>>>(p, g, ph);                    \
        return hipGetLastError();                                                        \
    }
    P2G_CASE(2, 3) P2G_CASE(2, 2) P2G_CASE(3, 3) P2G_CASE(3, 2)
#undef P2G_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_grid_update(const LaunchCfg& c, const DevGrid& g, const Phys& ph, hipStream_t st)
{
    dim3 grid(nblocks_total(g)), blk(64);
    bool coul = c.boundary == MPM_BOUNDARY_COULOMB_MARGIN;
    const bool apic = ph.scheme == MPM_SCHEME_APIC_FORCE;
#define GU_CASE(D, CO, AP)                                                               \
    if (c.dim == D && coul == CO && apic == AP) {                                        \
        k_grid_update<D, CO, AP><<<grid, blk, 0, st>>>(g, ph);                           \
        return hipGetLastError();                                                        \
    }
    GU_CASE(2, true, false) GU_CASE(2, false, false) GU_CASE(3, true, false) GU_CASE(3, false, false)
    GU_CASE(2, true, true) GU_CASE(2, false, true) GU_CASE(3, true, true) GU_CASE(3, false, true)
#undef GU_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_g2p(const LaunchCfg& c, const DevParticles& p, const DevGrid& g,
                      const Phys& ph, hipStream_t st)
{
    if (p.n == 0) return hipSuccess;
    dim3 grid((p.n + 255) / 256), blk(256);
    bool coul = c.boundary == MPM_BOUNDARY_COULOMB_MARGIN;
    const bool apic = ph.scheme == MPM_SCHEME_APIC_FORCE;
#define G2P_CASE(D, O)                                                                   \
    if (c.dim == D && c.order == O) {                                                    \
        if (apic) {                                                                      \
            if (coul) k_g2p<D, O, true, true><<<grid, blk, 0, st>>>(p, g, ph);           \
            else      k_g2p<D, O, false, true><<<grid, blk, 0, st>>>(p, g, ph);          \
        } else {                                                                         \
            if (coul) k_g2p<D, O, true, false><<<grid, blk, 0, st>>>(p, g, ph);          \
            else      k_g2p<D, O, false, false><<<grid, blk, 0, st>>>(p, g, ph);         \
        }                                                                                \
        return hipGetLastError();                                                        \
    }
    G2P_CASE(2, 3) G2P_CASE(2, 2) G2P_CASE(3, 3) G2P_CASE(3, 2)
#undef G2P_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_reset(const LaunchCfg& c, const DevGrid& g, int nblocks, hipStream_t st)
{
    if (nblocks <= 0) return hipSuccess;
    dim3 grid(nblocks), blk(64);
    if (c.dim == 2) k_reset<2><<<grid, blk, 0, st>>>(g);
    else            k_reset<3><<<grid, blk, 0, st>>>(g);
    return hipGetLastError();
}

hipError_t launch_grid_update_clear(const LaunchCfg& c, const DevGrid& cur, const DevGrid& old,
                                    const Phys& ph, hipStream_t st)
{
    const int nb = nblocks_total(cur);
    dim3 grid((nb + 63) / 64), blk(256);
    bool coul = c.boundary == MPM_BOUNDARY_COULOMB_MARGIN;
    if (c.dim == 2) {
        if (coul) k_grid_update_clear<2, true, false><<<grid, blk, 0, st>>>(cur, old, ph, nb, nullptr, nullptr);
        else      k_grid_update_clear<2, false, false><<<grid, blk, 0, st>>>(cur, old, ph, nb, nullptr, nullptr);
    } else {
        if (coul) k_grid_update_clear<3, true, false><<<grid, blk, 0, st>>>(cur, old, ph, nb, nullptr, nullptr);
        else      k_grid_update_clear<3, false, false><<<grid, blk, 0, st>>>(cur, old, ph, nb, nullptr, nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_grid_update_clear_batch(const LaunchCfg& c, const DevGrid& cur, const DevGrid& old,
                                          const SceneDesc* scenes, const int* block_scene, int nblocks,
                                          hipStream_t st)
{
    dim3 grid((nblocks + 63) / 64), blk(256);
    bool coul = c.boundary == MPM_BOUNDARY_COULOMB_MARGIN;
    Phys none{};
    if (c.dim == 2) {
        if (coul) k_grid_update_clear<2, true, true><<<grid, blk, 0, st>>>(cur, old, none, nblocks, scenes, block_scene);
        else      k_grid_update_clear<2, false, true><<<grid, blk, 0, st>>>(cur, old, none, nblocks, scenes, block_scene);
    } else {
        if (coul) k_grid_update_clear<3, true, true><<<grid, blk, 0, st>>>(cur, old, none, nblocks, scenes, block_scene);
        else      k_grid_update_clear<3, false, true><<<grid, blk, 0, st>>>(cur, old, none, nblocks, scenes, block_scene);
    }
    return hipGetLastError();
}

hipError_t launch_check_finite(int dim, const DevParticles& p, unsigned long long* bad,
                               hipStream_t st)
{
    if (p.n == 0) return hipSuccess;
    dim3 grid((p.n + 255) / 256), blk(256);
    if (dim == 2) k_check_finite<2><<<grid, blk, 0, st>>>(p, bad);
    else          k_check_finite<3><<<grid, blk, 0, st>>>(p, bad);
    return hipGetLastError();
}

}  // namespace mpm
