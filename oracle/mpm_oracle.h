/*
 * oracle/mpm_oracle.h — CPU restatement of the reference's MLS-MPM substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared with;
 * it is never linked into the product library.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py load it.
 *
 * Every function follows the reference's GLSL compute shaders statement by
 * statement (src/gpu/shaders/{2d,3d}/{common,p2g,grid_update,g2p,reset}.glsl),
 * in FP32 (suffix _f32) and in an FP64 twin of the same algorithm (suffix _f64).
 * With p->scheme == MPM_SCHEME_APIC_FORCE the stages restate the CPU path instead
 * (src/cpu/Grid.cpp, src/cpu/Particles.cpp).
 *
 * Layouts (all SoA, component-major, N particles, G grid nodes):
 *   x[d*N+p], v[d*N+p]             position, velocity
 *   B[(r*D+c)*N+p]                 APIC matrix B (row-major), C = D^-1 B
 *   Fe[(r*D+c)*N+p], Fp[...]       elastic / plastic deformation gradients
 *   mass[p], vol[p]
 *   grid node id = k*(W+1)*(H+1) + j*(W+1) + i   (shaders/3d/p2g.glsl:75)
 *   gmass[G], gmom[d*G+id], gvcol[d*G+id], gvfric[d*G+id]
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#include "../include/mpm/params.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Number of grid nodes for the parameter set: (W+1)(H+1)[(D+1)]. */
long orc_num_nodes(const mpm_params* p);

/* B-spline weight N(x) and derivative, x in grid units; order 3 = cubic (as coded in
 * shaders/2d/common.glsl:43-69), order 2 = quadratic. */
float  orc_weight_f32(float x, int order);
double orc_weight_f64(double x, int order);
float  orc_dweight_f32(float x, int order);
double orc_dweight_f64(double x, int order);

/* SVD as coded in the reference: 2x2 analytic (shaders/2d/common.glsl:2-37),
 * 3x3 cyclic Jacobi, 6 sweeps (shaders/3d/common.glsl:8-105).  Row-major M, U, V;
 * M = U diag(sig) V^T, det U >= 0 enforced by flipping the last column and sigma. */
void orc_svd2_f32(const float* M, float* U, float* sig, float* V);
void orc_svd2_f64(const double* M, double* U, double* sig, double* V);
void orc_svd3_f32(const float* M, float* U, float* sig, float* V);
void orc_svd3_f64(const double* M, double* U, double* sig, double* V);

/* Kirchhoff stress tau = 2 mu (F-R) F^T + lambda (J-1) J I with hardening
 * exp(XI (1-Jp)) (shaders/2d/p2g.glsl:44-59).  F row-major DxD. */
void orc_stress_f32(const mpm_params* p, const float* F, float Jp, float* tau);
void orc_stress_f64(const mpm_params* p, const double* F, double Jp, double* tau);

/* Stage functions (dispatch on p->dim and p->scheme).  gforce[d*G+id] is the CPU
 * scheme's grid force (src/cpu/Grid.h:49); unused (may be NULL) for the MLS scheme. */
void orc_p2g_f32(const mpm_params* p, long n, const float* x, const float* v,
                 const float* mass, const float* vol, const float* B,
                 const float* Fe, const float* Fp, float* gmass, float* gmom, float* gforce);
void orc_p2g_f64(const mpm_params* p, long n, const double* x, const double* v,
                 const double* mass, const double* vol, const double* B,
                 const double* Fe, const double* Fp, double* gmass, double* gmom, double* gforce);

void orc_grid_update_f32(const mpm_params* p, const float* gmass, const float* gmom,
                         const float* gforce, float* gvcol, float* gvfric);
void orc_grid_update_f64(const mpm_params* p, const double* gmass, const double* gmom,
                         const double* gforce, double* gvcol, double* gvfric);

void orc_g2p_f32(const mpm_params* p, long n, float* x, float* v, float* B,
                 float* Fe, float* Fp, const float* gvcol, const float* gvfric);
void orc_g2p_f64(const mpm_params* p, long n, double* x, double* v, double* B,
                 double* Fe, double* Fp, const double* gvcol, const double* gvfric);

/* nsteps full substeps: p2g -> grid_update -> g2p -> reset (GPUSolver.cpp:63-80).
 * grid scratch: gmass[G], gmom/gvcol/gvfric[D*G]; zero on entry, zero on exit. */
void orc_run_f32(const mpm_params* p, long n, int nsteps, float* x, float* v,
                 const float* mass, const float* vol, float* B, float* Fe, float* Fp,
                 float* gmass, float* gmom, float* gvcol, float* gvfric);
void orc_run_f64(const mpm_params* p, long n, int nsteps, double* x, double* v,
                 const double* mass, const double* vol, double* B, double* Fe, double* Fp,
                 double* gmass, double* gmom, double* gvcol, double* gvfric);

#ifdef __cplusplus
}
#endif

#endif
