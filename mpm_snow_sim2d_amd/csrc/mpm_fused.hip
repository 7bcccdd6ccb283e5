// mpm_fused.hip — the tiled, LDS-staged particle kernels.
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
// block (at most tile_p of them, sorted by cell).  The workgroup stages the block's
// node tile plus a halo (stencil reach + a drift margin of MG cells for particles
// that moved since the last binning) in LDS: velocities for the gather and a zeroed
// accumulator for the scatter.  Particle scatters are LDS float atomics; the tile is
// flushed to HBM with one global atomic per touched node component.  A particle
// whose stencil leaves the tile takes the direct global path, so the result never
// depends on how stale the binning is.
#include "mpm_kernels.hpp"

namespace mpm {

template <int DIM> struct TileShape {
    static constexpr int BS = BlockShape<DIM>::BS;
    static constexpr int MG = DIM == 3 ? 1 : 2;          // drift margin (cells)
    static constexpr int T = BS + 3 + 2 * MG;            // tile edge (nodes)
    static constexpr int TN = DIM == 3 ? T * T * T : T * T;
    static constexpr int NB = DIM == 3 ? 27 : 9;         // grid blocks a tile overlaps
};

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

// Element a (0 <= a < W) of a register array, through selects (no scratch).
template <int W>
__device__ inline float pick(const float (&arr)[W], int a)
{
    float r = arr[0];
#pragma unroll
    for (int s = 1; s < W; ++s) r = (a == s) ? arr[s] : r;
    return r;
}

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

template <int DIM, int ORDER, bool COULOMB, bool G2P>
__global__ void __launch_bounds__(256) k_tiled(DevParticles P, DevGrid src, DevGrid dst, TileWork work,
                                               Phys ph)
{
    using TS = TileShape<DIM>;
    constexpr int BS = TS::BS, MG = TS::MG, T = TS::T, TN = TS::TN;
    constexpr int W = Spline<ORDER>::W;
    constexpr int NV = G2P ? (COULOMB ? 2 * DIM : DIM) : 1;
    constexpr int TV = G2P ? TN : 1;
    __shared__ float s_acc[DIM + 1][TN];
    __shared__ float s_vel[NV][TV];
    __shared__ int s_touch[TS::NB];

    const int n = P.n;
    const int tid = threadIdx.x;
    const int nit = *work.n_items;
    const float fdt = ph.Dinv * ph.dt;

    for (int it = blockIdx.x; it < nit; it += gridDim.x) {
        int4 item = work.items[it];
        int blk = item.x, first = item.y, cnt = item.z;
        int bc[3];
        bc[0] = blk % dst.nbx;
        int rest = blk / dst.nbx;
        if constexpr (DIM == 2) { bc[1] = rest; bc[2] = 0; }
        else { bc[1] = rest % dst.nby; bc[2] = rest / dst.nby; }
        int o[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) o[a] = (a < DIM) ? bc[a] * BS - 1 - MG : 0;

        // ---- stage the tile
        for (int e = tid; e < TN; e += 256) {
#pragma unroll
            for (int c = 0; c <= DIM; ++c) s_acc[c][e] = 0.f;
            if constexpr (G2P) {
                int t[3];
                tile_coords<DIM>(e, t);
                int i = o[0] + t[0], j = o[1] + t[1], k = o[2] + t[2];
                float4 a = make_float4(0.f, 0.f, 0.f, 0.f), c4 = a;
                if (in_grid<DIM>(ph, i, j, k)) {
                    int addr = node_addr<DIM>(src, i, j, k);
                    a = src.a[addr];
                    if constexpr (COULOMB) c4 = src.vc[addr];
                }
                s_vel[0][e] = a.y;
                s_vel[1][e] = a.z;
                if constexpr (DIM == 3) s_vel[2][e] = a.w;
                if constexpr (COULOMB) {
                    s_vel[DIM][e] = c4.x;
                    s_vel[DIM + 1][e] = c4.y;
                    if constexpr (DIM == 3) s_vel[DIM + 2][e] = c4.z;
                }
            }
        }
        if (tid < TS::NB) s_touch[tid] = 0;
        __syncthreads();

        for (int q = tid; q < cnt; q += 256) {
            const int p = first + q;
            float x[DIM], v[DIM], C[DIM * DIM], F[DIM * DIM], Jp;
            float U[DIM * DIM], sc[DIM], tdiag[DIM];
            float J;
#pragma unroll
            for (int d = 0; d < DIM; ++d) x[d] = P.x[d * n + p];
            Jp = P.Jp[p];
            const float vol = P.vol[p];
            const float pm = P.mass[p];
            float tau[DIM * DIM];

            if constexpr (G2P) {
                // ---------------- G2P (g2p.glsl)
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
#pragma unroll
                for (int a0 = 0; a0 < W; ++a0)
#pragma unroll
                    for (int a1 = 0; a1 < W; ++a1)
#pragma unroll
                        for (int a2 = 0; a2 < (DIM == 3 ? W : 1); ++a2) {
                            int idx[3] = {st.base[0] + a0, st.base[1] + a1,
                                          DIM == 3 ? st.base[DIM - 1] + a2 : 0};
                            if (!in_grid<DIM>(ph, idx[0], idx[1], idx[2])) continue;
                            float dp[DIM];
                            dp[0] = st.d[0][a0] * ph.dx;
                            dp[1] = st.d[1][a1] * ph.dx;
                            float w = st.w[0][a0] * st.w[1][a1];
                            if constexpr (DIM == 3) {
                                dp[DIM - 1] = st.d[DIM - 1][a2] * ph.dx;
                                w = w * st.w[DIM - 1][a2];
                            }
                            float vF[3], vC[3];
                            if (tile_ok) {
                                int e = DIM == 3 ? ((rb[2] + a2) * T + (rb[1] + a1)) * T + (rb[0] + a0)
                                                 : (rb[1] + a1) * T + (rb[0] + a0);
#pragma unroll
                                for (int d = 0; d < DIM; ++d) {
                                    vF[d] = s_vel[d][e];
                                    vC[d] = COULOMB ? s_vel[(DIM + d) % NV][e] : vF[d];
                                }
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
                            for (int r = 0; r < DIM; ++r) {
                                nv[r] += vF[r] * w;
#pragma unroll
                                for (int c = 0; c < DIM; ++c) nB[r * DIM + c] += w * (vF[r] * (-dp[c]));
                                nx[r] += (ph.dt * vC[r] + (float)idx[r] * ph.dx) * w;
                            }
                        }
                float M[DIM * DIM], Fo[DIM * DIM], Fh[DIM * DIM];
#pragma unroll
                for (int e = 0; e < DIM * DIM; ++e) {
                    C[e] = ph.Dinv * nB[e];
                    M[e] = ((e % (DIM + 1)) == 0 ? 1.f : 0.f) + ph.dt * C[e];
                    Fo[e] = P.F[e * n + p];
                }
                matmul<DIM>(M, Fo, Fh);
                float sig[DIM], V[DIM * DIM];
                svd<DIM>(Fh, U, sig, V);
                float ratio = 1.f;
                J = 1.f;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    float s = fminf(fmaxf(sig[d], ph.theta0), ph.theta1);
                    s = fmaxf(s, 1e-6f);
                    sc[d] = s;
                    ratio *= sig[d] / s;
                    J *= s;
                }
#pragma unroll
                for (int r = 0; r < DIM; ++r)
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        float s = (U[r * DIM] * sc[0]) * V[c * DIM];
#pragma unroll
                        for (int qq = 1; qq < DIM; ++qq) s += (U[r * DIM + qq] * sc[qq]) * V[c * DIM + qq];
                        F[r * DIM + c] = s;
                    }
                Jp = Jp * ratio;
#pragma unroll
                for (int d = 0; d < DIM; ++d) { x[d] = nx[d]; v[d] = nv[d]; }
#pragma unroll
                for (int d = 0; d < DIM; ++d) { P.x[d * n + p] = x[d]; P.v[d * n + p] = v[d]; }
#pragma unroll
                for (int e = 0; e < DIM * DIM; ++e) { P.C[e * n + p] = C[e]; P.F[e * n + p] = F[e]; }
                P.Jp[p] = Jp;
                // stress of the new state: U diag(2 mu (s-1) s + la (J-1) J) U^T
                float harden = expf(ph.xi * (1.f - Jp));
                float tmu = 2.f * (ph.mu * harden);
                float sI = (ph.lambda * harden) * (J - 1.f) * J;
#pragma unroll
                for (int d = 0; d < DIM; ++d) tdiag[d] = tmu * (sc[d] - 1.f) * sc[d] + sI;
#pragma unroll
                for (int r = 0; r < DIM; ++r)
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        float s = U[r * DIM] * tdiag[0] * U[c * DIM];
#pragma unroll
                        for (int qq = 1; qq < DIM; ++qq) s += U[r * DIM + qq] * tdiag[qq] * U[c * DIM + qq];
                        tau[r * DIM + c] = s;
                    }
            } else {
                // ---------------- stress from the stored state (p2g.glsl:44-59)
#pragma unroll
                for (int d = 0; d < DIM; ++d) v[d] = P.v[d * n + p];
#pragma unroll
                for (int e = 0; e < DIM * DIM; ++e) { C[e] = P.C[e * n + p]; F[e] = P.F[e * n + p]; }
                float sig[DIM], V[DIM * DIM];
                svd<DIM>(F, U, sig, V);
                kirchhoff<DIM>(ph, F, U, V, det<DIM>(F), Jp, tau);
                (void)sc; (void)tdiag; (void)J;
            }

            if (vol == 0.f) continue;                                 // p2g.glsl:42
            // ---------------- P2G (p2g.glsl) of the particle into dst
            // momentum of node i: w (m v + A (x_i - x_p)), A = m C - (V D^-1 dt) tau
            float A[DIM * DIM];
            const float vfd = vol * fdt;
#pragma unroll
            for (int e = 0; e < DIM * DIM; ++e) A[e] = pm * C[e] - vfd * tau[e];
            float mvp[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) mvp[d] = pm * v[d];
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
            if (tile_ok) {
                // LDS scatter in a per-lane rotated node order: lanes that hold
                // particles of the same cell (neighbours after binning) start at
                // different stencil nodes, so one ds_add never sees several lanes on
                // one address.
                constexpr int NS = DIM == 3 ? W * W * W : W * W;
                int r0 = (tid & 63) % NS;
                int a2 = DIM == 3 ? r0 % W : 0;
                int a1 = DIM == 3 ? (r0 / W) % W : r0 % W;
                int a0 = DIM == 3 ? r0 / (W * W) : r0 / W;
#pragma unroll 3
                for (int sn = 0; sn < NS; ++sn) {
                    int idx[3] = {st.base[0] + a0, st.base[1] + a1, DIM == 3 ? st.base[DIM - 1] + a2 : 0};
                    if (in_grid<DIM>(ph, idx[0], idx[1], idx[2])) {
                        float xi[DIM];
                        xi[0] = -pick<W>(st.d[0], a0) * ph.dx;
                        xi[1] = -pick<W>(st.d[1], a1) * ph.dx;
                        float w = pick<W>(st.w[0], a0) * pick<W>(st.w[1], a1);
                        if constexpr (DIM == 3) {
                            xi[DIM - 1] = -pick<W>(st.d[DIM - 1], a2) * ph.dx;
                            w = w * pick<W>(st.w[DIM - 1], a2);
                        }
                        float vals[DIM + 1];
                        node_contrib<DIM>(w, pm, mvp, A, xi, vals);
                        int e = DIM == 3 ? ((rb[2] + a2) * T + (rb[1] + a1)) * T + (rb[0] + a0)
                                         : (rb[1] + a1) * T + (rb[0] + a0);
#pragma unroll
                        for (int c = 0; c <= DIM; ++c) atomicAdd(&s_acc[c][e], vals[c]);
                    }
                    if constexpr (DIM == 3) {
                        if (++a2 == W) { a2 = 0; if (++a1 == W) { a1 = 0; if (++a0 == W) a0 = 0; } }
                    } else {
                        if (++a1 == W) { a1 = 0; if (++a0 == W) a0 = 0; }
                    }
                }
            } else {
                // stencil outside the tile (particle drifted since binning): direct
                // global atomics
#pragma unroll
                for (int a0 = 0; a0 < W; ++a0)
#pragma unroll
                    for (int a1 = 0; a1 < W; ++a1)
#pragma unroll
                        for (int a2 = 0; a2 < (DIM == 3 ? W : 1); ++a2) {
                            int idx[3] = {st.base[0] + a0, st.base[1] + a1, DIM == 3 ? st.base[DIM - 1] + a2 : 0};
                            if (!in_grid<DIM>(ph, idx[0], idx[1], idx[2])) continue;
                            float xi[DIM];
                            xi[0] = -st.d[0][a0] * ph.dx;
                            xi[1] = -st.d[1][a1] * ph.dx;
                            float w = st.w[0][a0] * st.w[1][a1];
                            if constexpr (DIM == 3) {
                                xi[DIM - 1] = -st.d[DIM - 1][a2] * ph.dx;
                                w = w * st.w[DIM - 1][a2];
                            }
                            float vals[DIM + 1];
                            node_contrib<DIM>(w, pm, mvp, A, xi, vals);
                            float* dstp = reinterpret_cast<float*>(&dst.a[node_addr<DIM>(dst, idx[0], idx[1], idx[2])]);
#pragma unroll
                            for (int c = 0; c <= DIM; ++c) unsafeAtomicAdd(dstp + c, vals[c]);
                            dst.flags[block_of<DIM>(dst, idx[0], idx[1], idx[2])] = 1u;
                        }
            }
        }
        __syncthreads();

        // ---- flush the tile
        for (int e = tid; e < TN; e += 256) {
            float vals[DIM + 1];
            bool any = false;
#pragma unroll
            for (int c = 0; c <= DIM; ++c) { vals[c] = s_acc[c][e]; any |= vals[c] != 0.f; }
            if (!any) continue;
            int t[3];
            tile_coords<DIM>(e, t);
            int i = o[0] + t[0], j = o[1] + t[1], k = o[2] + t[2];
            float* dstp = reinterpret_cast<float*>(&dst.a[node_addr<DIM>(dst, i, j, k)]);
#pragma unroll
            for (int c = 0; c <= DIM; ++c) unsafeAtomicAdd(dstp + c, vals[c]);
            int rbx = (i >> BlockShape<DIM>::SHIFT) - (bc[0] - 1);
            int rby = (j >> BlockShape<DIM>::SHIFT) - (bc[1] - 1);
            int rbz = DIM == 3 ? (k >> BlockShape<DIM>::SHIFT) - (bc[2] - 1) : 0;
            s_touch[(rbz * 3 + rby) * 3 + rbx] = 1;
        }
        __syncthreads();
        if (tid < TS::NB && s_touch[tid]) {
            int rbx = tid % 3, rby = (tid / 3) % 3, rbz = tid / 9;
            int bx = bc[0] - 1 + rbx, by = bc[1] - 1 + rby, bz = DIM == 3 ? bc[2] - 1 + rbz : 0;
            int b = DIM == 3 ? (bz * dst.nby + by) * dst.nbx + bx : by * dst.nbx + bx;
            dst.flags[b] = 1u;
        }
        __syncthreads();
    }
}

static int tiled_grid(const TileWork& w)
{
    int g = w.max_items < 4096 ? w.max_items : 4096;
    return g > 0 ? g : 1;
}

hipError_t launch_g2p2g(const LaunchCfg& c, const DevParticles& p, const DevGrid& src, const DevGrid& dst,
                        const TileWork& w, const Phys& ph, hipStream_t st)
{
    if (p.n == 0) return hipSuccess;
    dim3 grid(tiled_grid(w)), blk(256);
    bool coul = c.boundary == MPM_BOUNDARY_COULOMB_MARGIN;
#define TCASE(D, O)                                                                          \
    if (c.dim == D && c.order == O) {                                                        \
        if (coul) k_tiled<D, O, true, true><<<grid, blk, 0, st>>>(p, src, dst, w, ph);       \
        else      k_tiled<D, O, false, true><<<grid, blk, 0, st>>>(p, src, dst, w, ph);      \
        return hipGetLastError();                                                            \
    }
    TCASE(2, 3) TCASE(2, 2) TCASE(3, 3) TCASE(3, 2)
#undef TCASE
    return hipErrorInvalidValue;
}

hipError_t launch_p2g_tiled(const LaunchCfg& c, const DevParticles& p, const DevGrid& dst,
                            const TileWork& w, const Phys& ph, hipStream_t st)
{
    if (p.n == 0) return hipSuccess;
    dim3 grid(tiled_grid(w)), blk(256);
#define TCASE(D, O)                                                                          \
    if (c.dim == D && c.order == O) {                                                        \
        k_tiled<D, O, false, false><<<grid, blk, 0, st>>>(p, dst, dst, w, ph);               \
        return hipGetLastError();                                                            \
    }
    TCASE(2, 3) TCASE(2, 2) TCASE(3, 3) TCASE(3, 2)
#undef TCASE
    return hipErrorInvalidValue;
}

}  // namespace mpm
