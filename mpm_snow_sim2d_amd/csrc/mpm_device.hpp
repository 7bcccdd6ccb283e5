// mpm_device.hpp — device-side math of the MLS-MPM substep for gfx950.
//
// B-spline weights, the reference's SVDs (2x2 analytic, 3x3 cyclic Jacobi),
// Kirchhoff stress with hardening, and the block-major grid addressing used by
// every kernel.  All arithmetic is FP32, matrices are register arrays in row-major
// order, and every loop is fully unrolled so they stay in VGPRs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpm/params.h"

namespace mpm {

// ---------------------------------------------------------------------------
// Parameters as the kernels consume them (derived once on the host).
struct Phys {
    float dx, inv_dx, dt, Dinv;      // Dinv = D^-1: 3/dx^2 cubic, 4/dx^2 quadratic
    float gravity, friction, margin; // margin in cells (BORDER_MARGIN)
    float theta0, theta1;            // 1 - THETA_C, 1 + THETA_S
    float xi, mu, lambda;
    float mass_eps;
    int   W, H, D;                   // cells per axis; nodes 0..W
    int   boundary;
    int   scheme;                    // MPM_SCHEME_MLS or MPM_SCHEME_APIC_FORCE
};

// Grid: block-major storage, 64 nodes per block (4x4x4 in 3D, 8x8 in 2D), one
// float4 per node: (mass, momentum) after P2G, (mass, velocity) after the grid
// update.  vc holds the collision velocity when it differs from the friction
// velocity (Coulomb walls); f accumulates the explicit grid force of the CPU
// path's scheme (NULL otherwise); flags marks the blocks a scatter touched.
struct DevGrid {
    float4*   a;
    float4*   vc;
    float4*   f;
    uint32_t* flags;
    int       nbx, nby, nbz;
};

// One scene of a batch: its parameters and where its grid blocks start in the
// batch-wide block numbering (block-major addresses are offset by block_off*64).
struct SceneDesc {
    Phys ph;
    int  block_off;
    int  nbx, nby, nbz;
    int  pad;
};

__host__ __device__ inline DevGrid scene_grid(const DevGrid& g, const SceneDesc& sd)
{
    DevGrid r;
    const size_t off = (size_t)sd.block_off * 64;
    r.a = g.a + off;
    r.vc = g.vc ? g.vc + off : nullptr;
    r.f = g.f ? g.f + off : nullptr;
    r.flags = g.flags + sd.block_off;
    r.nbx = sd.nbx; r.nby = sd.nby; r.nbz = sd.nbz;
    return r;
}

struct DevParticles {
    float*       x;     // [DIM][n]
    float*       v;     // [DIM][n]
    float*       C;     // [DIM*DIM][n]
    float*       F;     // [DIM*DIM][n]
    float*       Jp;    // [n]
    const float* mass;  // [n]
    const float* vol;   // [n]
    uint32_t*    pid;   // [n] original particle index of each slot
    int          n;
};

template <int DIM> struct BlockShape;
template <> struct BlockShape<2> { static constexpr int BS = 8; static constexpr int SHIFT = 3; };
template <> struct BlockShape<3> { static constexpr int BS = 4; static constexpr int SHIFT = 2; };

template <int DIM>
__host__ __device__ inline int node_addr(const DevGrid& g, int i, int j, int k)
{
    constexpr int S = BlockShape<DIM>::SHIFT;
    constexpr int M = (1 << S) - 1;
    if constexpr (DIM == 2) {
        int blk = (j >> S) * g.nbx + (i >> S);
        return (blk << 6) | ((j & M) << S) | (i & M);
    } else {
        int blk = ((k >> S) * g.nby + (j >> S)) * g.nbx + (i >> S);
        return (blk << 6) | ((k & M) << (2 * S)) | ((j & M) << S) | (i & M);
    }
}

template <int DIM>
__host__ __device__ inline int block_of(const DevGrid& g, int i, int j, int k)
{
    constexpr int S = BlockShape<DIM>::SHIFT;
    if constexpr (DIM == 2) return (j >> S) * g.nbx + (i >> S);
    else return ((k >> S) * g.nby + (j >> S)) * g.nbx + (i >> S);
}

// ---------------------------------------------------------------------------
// Cell index of a position in grid units.  Within +-2^30 it is floor (or int()
// truncation) exactly; larger magnitudes and infinities become +-2^30, far outside
// any grid, so that an index plus a stencil or tile offset never overflows the int
// range (a plain conversion saturates at INT_MAX, and INT_MAX + 1 wraps to a
// negative index that bounds tests can take for an inside one); NaN gives 0, as the
// hardware conversion does.
__host__ __device__ inline int cell_floor(float v)
{
    const float f = floorf(v);
    return f != f ? 0 : (int)fminf(fmaxf(f, -1073741824.0f), 1073741824.0f);
}

__host__ __device__ inline int cell_trunc(float v)
{
    const float f = truncf(v);
    return f != f ? 0 : (int)fminf(fmaxf(f, -1073741824.0f), 1073741824.0f);
}

// B-spline weights along one axis (shaders/2d/common.glsl:43-55 for cubic).
template <int ORDER> struct Spline;

template <> struct Spline<3> {
    static constexpr int W = 4;
    __host__ __device__ static inline int base(float xg) { return cell_trunc(xg) - 1; }   // int() truncation
    __device__ static inline float N(float x)
    {
        float a = fabsf(x);
        if (a < 1.0f) return 0.5f * (a * a * a) - (a * a) + (2.0f / 3.0f);
        if (a < 2.0f) { float t = 2.0f - a; return t * t * t / 6.0f; }
        return 0.0f;
    }
    // N_grad as coded (shaders/2d/common.glsl:57-69, src/cpu/Grid.h:102-116)
    __device__ static inline float dN(float x)
    {
        float a = fabsf(x);
        if (a < 1.0f) return (1.5f * a * x) - (2.0f * x);
        if (a < 2.0f) return (-0.5f * a * x) + (2.0f * x) - ((2.0f * x) / a);
        return 0.0f;
    }
    // The stencil's weights w[s] = N(d[s]) and distances d[s] = xg - (base + s).
    __device__ static inline void weights(float xg, int base, float* w, float* d)
    {
#pragma unroll
        for (int s = 0; s < W; ++s) {
            d[s] = xg - (float)(base + s);
            w[s] = N(d[s]);
        }
    }
};

template <> struct Spline<2> {
    static constexpr int W = 3;
    __host__ __device__ static inline int base(float xg) { return cell_floor(xg - 0.5f); }
    __device__ static inline float N(float x)
    {
        float a = fabsf(x);
        if (a < 0.5f) return 0.75f - a * a;
        if (a < 1.5f) { float t = 1.5f - a; return 0.5f * t * t; }
        return 0.0f;
    }
    __device__ static inline float dN(float x)
    {
        float a = fabsf(x);
        if (a < 0.5f) return -2.0f * x;
        if (a < 1.5f) return (x > 0.0f ? -1.0f : 1.0f) * (1.5f - a);
        return 0.0f;
    }
    // The stencil's weights w[s] = N(d[s]) and distances d[s] = xg - (base + s) for
    // base = base(xg): d[0] lies in [0.5, 1.5), so each node's piece of N is known
    // and no branch is needed (N is C1 where the pieces meet, so a base that the
    // rounding of xg - 0.5 moves by one changes a weight only in its last bits).  A
    // non-finite xg, or one so large that its base was clamped, gets N's zeros
    // beyond the support, as N itself gives.
    __device__ static inline void weights(float xg, int base, float* w, float* d)
    {
        const float d0 = xg - (float)base;
        const float d1 = d0 - 1.0f;
        d[0] = d0;
        d[1] = d1;
        d[2] = d0 - 2.0f;
        const float t0 = 0.5f - d1, t2 = 0.5f + d1;
        const bool ok = fabsf(d1) <= 1.0f;
        w[0] = ok ? 0.5f * t0 * t0 : 0.0f;
        w[1] = ok ? 0.75f - d1 * d1 : 0.0f;
        w[2] = ok ? 0.5f * t2 * t2 : 0.0f;
    }
};

// Per-particle stencil: base node, per-axis weights and grid-unit distances
// dist = xg - node (the reference's pPos - gi).
template <int DIM, int ORDER>
struct Stencil {
    static constexpr int W = Spline<ORDER>::W;
    int   base[DIM];
    float w[DIM][W];
    float d[DIM][W];
    __device__ inline void init(const float* xg)
    {
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            base[a] = Spline<ORDER>::base(xg[a]);
            Spline<ORDER>::weights(xg[a], base[a], w[a], d[a]);
        }
    }
};

// ---------------------------------------------------------------------------
// Small matrix helpers (row-major, fully unrolled).
template <int DIM>
__host__ __device__ inline float det(const float* A)
{
    if constexpr (DIM == 2) return A[0] * A[3] - A[2] * A[1];
    else
        return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
               A[2] * (A[3] * A[7] - A[4] * A[6]);
}

template <int DIM>
__device__ inline void matmul(const float* A, const float* B, float* C)
{
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            float s = A[r * DIM] * B[c];
#pragma unroll
            for (int k = 1; k < DIM; ++k) s += A[r * DIM + k] * B[k * DIM + c];
            C[r * DIM + c] = s;
        }
}

// ---------------------------------------------------------------------------
// SVD, the reference's algorithms and conventions: sigma unsorted, det U > 0 by
// flipping the last column of U and the last sigma.
//
// 2x2: shaders/2d/common.glsl:2-37 (eigenvectors of M^T M from one angle).
__device__ inline void svd2(const float* M, float* U, float* sig, float* V)
{
    float a = M[0] * M[0] + M[2] * M[2];
    float b = M[0] * M[1] + M[2] * M[3];
    float d = M[1] * M[1] + M[3] * M[3];
    float theta = (fabsf(b) < 1e-10f && fabsf(a - d) < 1e-10f) ? 0.0f
                                                               : 0.5f * atan2f(2.0f * b, a - d);
    float st, ct;
    sincosf(theta, &st, &ct);
    V[0] = ct; V[1] = -st; V[2] = st; V[3] = ct;
    // diag(V^T (M^T M) V)
    float t00 = ct * a + st * b, t01 = ct * b + st * d;
    float t10 = -st * a + ct * b, t11 = -st * b + ct * d;
    float d0 = t00 * ct + t01 * st;
    float d1 = -t10 * st + t11 * ct;
    sig[0] = sqrtf(fmaxf(d0, 1e-12f));
    sig[1] = sqrtf(fmaxf(d1, 1e-12f));
    float i0 = 1.0f / sig[0], i1 = 1.0f / sig[1];
    U[0] = (M[0] * ct + M[1] * st) * i0;
    U[1] = (-M[0] * st + M[1] * ct) * i1;
    U[2] = (M[2] * ct + M[3] * st) * i0;
    U[3] = (-M[2] * st + M[3] * ct) * i1;
    if (U[0] * U[3] - U[2] * U[1] < 0.0f) {
        U[1] = -U[1]; U[3] = -U[3]; sig[1] = -sig[1];
    }
}

// 3x3: shaders/3d/common.glsl:8-105 — cyclic Jacobi eigen-decomposition of the
// symmetric S = M^T M (Golub & Van Loan, Matrix Computations, sec. 8.5: symmetric
// Schur pair, six sweeps of (0,1), (0,2), (1,2)), then U = M V diag(1/sigma).
__device__ inline void jacobi_pair(float& spp, float& sqq, float& spq, float& sr_p, float& sr_q,
                                   float* V, int p, int q)
{
    // (c, s): symmetric Schur decomposition of [[spp, spq], [spq, sqq]] with the
    // reference's choice of root: t = tan(theta) = sign(tau) / (|tau| + sqrt(1 + tau^2)),
    // tau = (s_qq - s_pp) / (2 s_pq), written without the division by 2 s_pq as
    // |t| = |2 s_pq| / (|d| + sqrt(d^2 + 4 s_pq^2)), d = s_qq - s_pp (one reciprocal
    // fewer); the rotation zeroes s_pq and moves the diagonal by t s_pq, the form of
    // Golub & Van Loan, Matrix Computations, sec. 8.5.2, equal in exact arithmetic
    // to the reference's c^2 s_pp - 2 s c s_pq + s^2 s_qq.  The reciprocals and roots
    // use the hardware's single-instruction approximations (v_rcp_f32, v_sqrt_f32,
    // v_rsq_f32, about 1 ulp).  A skipped rotation (c = 1, s = 0) leaves every value
    // as it is except s_pq = 0, so it is applied as that alone: a wave whose lanes
    // have all converged passes the pair at once.  Besides the reference's absolute
    // test, a pair is taken as converged once |s_pq| <= 2^-24 sqrt(s_pp s_qq): the
    // rotation would then move the singular values by less than FP32 resolves, and
    // M = U diag(sigma) V^T with U = M V diag(1/sigma) holds for any V, so the
    // reconstructed F_E, the rotation U V^T and the stress change only at round-off
    // level.  (Past the absolute test |2 s_pq| >= 2e-10, so d^2 + 4 s_pq^2 cannot
    // underflow and |d| + sqrt(...) > 0.)
    if (fabsf(spq) < 1e-10f || spq * spq <= 3.5527137e-15f * (spp * sqq)) {
        spq = 0.0f;
        return;
    }
    const float d = sqq - spp, e2 = 2.0f * spq;
    const float r = __builtin_amdgcn_sqrtf(fmaf(d, d, e2 * e2));
    float t = fabsf(e2) * __builtin_amdgcn_rcpf(fabsf(d) + r);
    if (d != 0.0f && ((d < 0.0f) != (e2 < 0.0f))) t = -t;     // tau < 0
    const float c = __builtin_amdgcn_rsqf(fmaf(t, t, 1.0f));
    const float s = t * c;
    spp = spp - t * spq;
    sqq = sqq + t * spq;
    spq = 0.0f;
    // the remaining off-diagonal pair (r,p), (r,q) with r the third index
    float Srp = sr_p, Srq = sr_q;
    sr_p = c * Srp - s * Srq;
    sr_q = s * Srp + c * Srq;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float vp = V[r * 3 + p], vq = V[r * 3 + q];
        V[r * 3 + p] = c * vp - s * vq;
        V[r * 3 + q] = s * vp + c * vq;
    }
}

__device__ inline void svd3(const float* M, float* U, float* sig, float* V)
{
    float s00 = M[0] * M[0] + M[3] * M[3] + M[6] * M[6];
    float s01 = M[0] * M[1] + M[3] * M[4] + M[6] * M[7];
    float s02 = M[0] * M[2] + M[3] * M[5] + M[6] * M[8];
    float s11 = M[1] * M[1] + M[4] * M[4] + M[7] * M[7];
    float s12 = M[1] * M[2] + M[4] * M[5] + M[7] * M[8];
    float s22 = M[2] * M[2] + M[5] * M[5] + M[8] * M[8];
#pragma unroll
    for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0f : 0.0f;
#pragma unroll
    for (int sweep = 0; sweep < 6; ++sweep) {
        jacobi_pair(s00, s11, s01, s02, s12, V, 0, 1);   // third index 2: (2,0), (2,1)
        jacobi_pair(s00, s22, s02, s01, s12, V, 0, 2);   // third index 1: (1,0), (1,2)
        jacobi_pair(s11, s22, s12, s01, s02, V, 1, 2);   // third index 0: (0,1), (0,2)
    }
    sig[0] = __builtin_amdgcn_sqrtf(fmaxf(s00, 1e-12f));
    sig[1] = __builtin_amdgcn_sqrtf(fmaxf(s11, 1e-12f));
    sig[2] = __builtin_amdgcn_sqrtf(fmaxf(s22, 1e-12f));
    float inv[3] = {__builtin_amdgcn_rcpf(sig[0]), __builtin_amdgcn_rcpf(sig[1]), __builtin_amdgcn_rcpf(sig[2])};
    float MV[9];
    matmul<3>(M, V, MV);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) U[r * 3 + c] = MV[r * 3 + c] * inv[c];
    if (det<3>(U) < 0.0f) {
        U[2] = -U[2]; U[5] = -U[5]; U[8] = -U[8];
        sig[2] = -sig[2];
    }
}

template <int DIM>
__device__ inline void svd(const float* M, float* U, float* sig, float* V)
{
    if constexpr (DIM == 2) svd2(M, U, sig, V);
    else svd3(M, U, sig, V);
}

// The CPU path's decomposition (Eigen JacobiSVD, src/cpu/Particles.cpp:31,73): singular
// values non-negative, no sign fix.  The reference SVD with its last-column flip
// undone gives the same products U diag(s) V^T, U V^T.
template <int DIM>
__device__ inline void svd_nonneg(const float* M, float* U, float* sig, float* V)
{
    svd<DIM>(M, U, sig, V);
    if (sig[DIM - 1] < 0.0f) {
        sig[DIM - 1] = -sig[DIM - 1];
#pragma unroll
        for (int r = 0; r < DIM; ++r) U[r * DIM + DIM - 1] = -U[r * DIM + DIM - 1];
    }
}

// ---------------------------------------------------------------------------
// Elastic states without an SVD.  G2P clamps the singular values of
// F^ = (I + dt C) F_E to [theta0, theta1] (g2p.glsl:80-93), and the stress is
// 2 mu (F - R) F^T + lambda (J - 1) J I (p2g.glsl:44-59).  When every singular value
// already lies in the clamp range and det F > 0, the clamp changes nothing:
// F_E' = F^, J_P' = J_P, and with F = V R (V = sqrt(F F^T), the left stretch),
// 2 mu (F - R) F^T = 2 mu (V - I) V, so neither U, Sigma nor V of the SVD is needed.
// The test: every eigenvalue of B = F F^T (the sigma^2) lies in [lo2, hi2], i.e.
// B - lo2 I and hi2 I - B are positive definite (their leading principal minors are
// positive, Sylvester's criterion), and det F > 0 (a reflection has a negative
// sigma, which the clamp moves).  With lo2 >= 0.9 and hi2 <= 1.1, E = B - I has |E| <= 0.1, and
// V - I = sqrt(I + E) - I from the Taylor series of sqrt(1 + x) to x^5 is accurate
// to 21/1024 |E|^6 < 3e-8.  Symmetric 3x3 matrices as (xx, yy, zz, xy, xz, yz).
struct Sym3 {
    float xx, yy, zz, xy, xz, yz;
};

// a b for symmetric a, b that commute (polynomials in one matrix): symmetric too
__host__ __device__ inline Sym3 sym_mul(const Sym3& a, const Sym3& b)
{
    Sym3 c;
    c.xx = a.xx * b.xx + a.xy * b.xy + a.xz * b.xz;
    c.yy = a.xy * b.xy + a.yy * b.yy + a.yz * b.yz;
    c.zz = a.xz * b.xz + a.yz * b.yz + a.zz * b.zz;
    c.xy = a.xx * b.xy + a.xy * b.yy + a.xz * b.yz;
    c.xz = a.xx * b.xz + a.xy * b.yz + a.xz * b.zz;
    c.yz = a.xy * b.xz + a.yy * b.yz + a.yz * b.zz;
    return c;
}

__host__ __device__ inline Sym3 sym_add_diag(Sym3 a, float d)
{
    a.xx += d;
    a.yy += d;
    a.zz += d;
    return a;
}

// s (B - c I) positive definite (s = +1) or s (B - c I) with s = -1, i.e. c I - B:
// the leading principal minors of that symmetric matrix are all positive
__host__ __device__ inline bool positive_definite(const Sym3& B, float c, float s)
{
    const float a11 = s * (B.xx - c), a22 = s * (B.yy - c), a33 = s * (B.zz - c);
    const float a12 = s * B.xy, a13 = s * B.xz, a23 = s * B.yz;
    const float m2 = a11 * a22 - a12 * a12;
    const float m3 = a11 * (a22 * a33 - a23 * a23) - a12 * (a12 * a33 - a23 * a13) + a13 * (a12 * a23 - a22 * a13);
    return a11 > 0.f && m2 > 0.f && m3 > 0.f;
}

// The clamp bounds of the test: [theta0^2, theta1^2] within [0.9, 1.1], pulled in
// by 2^-20 so that the FP32 rounding of B cannot admit a singular value the clamp
// would move.
__host__ __device__ inline void elastic_bounds(float theta0, float theta1, float* lo2, float* hi2)
{
    *lo2 = fmaxf(theta0 * theta0, 0.9f) * (1.f + 9.5367431640625e-7f);
    *hi2 = fminf(theta1 * theta1, 1.1f) * (1.f - 9.5367431640625e-7f);
}

// True when F (row-major 3x3) passes the test; then *vmi = V - I, *bmv = B - V =
// (V - I) V (what the stress needs: V^2 = B) and *J = det F.
__host__ __device__ inline bool elastic_stretch(const float* F, float lo2, float hi2, Sym3* vmi, Sym3* bmv,
                                                float* J)
{
    Sym3 B;
    B.xx = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
    B.yy = F[3] * F[3] + F[4] * F[4] + F[5] * F[5];
    B.zz = F[6] * F[6] + F[7] * F[7] + F[8] * F[8];
    B.xy = F[0] * F[3] + F[1] * F[4] + F[2] * F[5];
    B.xz = F[0] * F[6] + F[1] * F[7] + F[2] * F[8];
    B.yz = F[3] * F[6] + F[4] * F[7] + F[5] * F[8];
    *J = det<3>(F);
    if (!(*J > 0.f) || !positive_definite(B, lo2, 1.f) || !positive_definite(B, hi2, -1.f)) return false;
    const Sym3 E = sym_add_diag(B, -1.f);
    // sqrt(1 + x) - 1 = x (1/2 + x (-1/8 + x (1/16 + x (-5/128 + x 7/256))))
    Sym3 P = E;
    P.xx *= 7.f / 256.f; P.yy *= 7.f / 256.f; P.zz *= 7.f / 256.f;
    P.xy *= 7.f / 256.f; P.xz *= 7.f / 256.f; P.yz *= 7.f / 256.f;
    P = sym_add_diag(P, -5.f / 128.f);
    P = sym_add_diag(sym_mul(E, P), 1.f / 16.f);
    P = sym_add_diag(sym_mul(E, P), -1.f / 8.f);
    P = sym_add_diag(sym_mul(E, P), 0.5f);
    const Sym3 v = sym_mul(E, P);
    *vmi = v;
    bmv->xx = E.xx - v.xx; bmv->yy = E.yy - v.yy; bmv->zz = E.zz - v.zz;
    bmv->xy = E.xy - v.xy; bmv->xz = E.xz - v.xz; bmv->yz = E.yz - v.yz;
    return true;
}

// Kirchhoff stress of an elastic state from its left stretch: 2 mu h (V - I) V +
// lambda h (J - 1) J I with (V - I) V = B - V (bmv from elastic_stretch), h =
// exp(xi (1 - J_P)).
__host__ __device__ inline void elastic_kirchhoff(const Sym3& bmv, float J, float tmu, float la, float* tau)
{
    const float sI = la * (J - 1.f) * J;
    tau[0] = tmu * bmv.xx + sI;
    tau[4] = tmu * bmv.yy + sI;
    tau[8] = tmu * bmv.zz + sI;
    tau[1] = tau[3] = tmu * bmv.xy;
    tau[2] = tau[6] = tmu * bmv.xz;
    tau[5] = tau[7] = tmu * bmv.yz;
}

// Kirchhoff stress of F with polar rotation R = U V^T (shaders/2d/p2g.glsl:44-59):
//   tau = 2 mu (F - R) F^T + lambda (J - 1) J I,  mu, lambda scaled by exp(XI (1 - Jp)).
template <int DIM>
__device__ inline void kirchhoff(const Phys& P, const float* F, const float* U, const float* V,
                                 float J, float Jp, float* tau)
{
    float Rm[DIM * DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            float s = U[r * DIM] * V[c * DIM];
#pragma unroll
            for (int k = 1; k < DIM; ++k) s += U[r * DIM + k] * V[c * DIM + k];
            Rm[r * DIM + c] = s;
        }
    float harden = expf(P.xi * (1.0f - Jp));
    float tmu = 2.0f * (P.mu * harden);
    float la = P.lambda * harden;
    float A[DIM * DIM];
#pragma unroll
    for (int i = 0; i < DIM * DIM; ++i) A[i] = tmu * (F[i] - Rm[i]);
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            float s = A[r * DIM] * F[c * DIM];
#pragma unroll
            for (int k = 1; k < DIM; ++k) s += A[r * DIM + k] * F[c * DIM + k];
            tau[r * DIM + c] = s;
        }
    float sI = la * (J - 1.0f) * J;
#pragma unroll
    for (int i = 0; i < DIM; ++i) tau[i * DIM + i] += sI;
}

}  // namespace mpm
