/*
 * oracle/mpm_oracle.c — instantiates the oracle body (mpm_oracle_impl.inc) for
 * float/double x 2D/3D and exports the C entry points of mpm_oracle.h.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference's GLSL substep
 * (src/gpu/shaders/{2d,3d}/{p2g,grid_update,g2p,reset}.glsl and
 * src/gpu/GPUSolver.cpp:63-80) that the HIP path
 * is checked against.  Built with -ffp-contract=off so each GLSL operation rounds
 * once, as written.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "mpm_oracle.h"

long orc_num_nodes(const mpm_params* p)
{
    long n = (long)(p->WIDTH + 1) * (long)(p->HEIGHT + 1);
    if (p->dim == 3) n *= (long)(p->DEPTH + 1);
    return n;
}

/* ---- float ---- */
#define R float
#define FABS fabsf
#define SQRT sqrtf
#define ATAN2 atan2f
#define COS cosf
#define SIN sinf
#define EXP expf
#define FLOOR floorf

#define DIM 2
#define NM(x) x##_f32_2
#include "mpm_oracle_impl.inc"
#undef NM
#undef DIM

#define DIM 3
#define NM(x) x##_f32_3
#include "mpm_oracle_impl.inc"
#undef NM
#undef DIM

#undef R
#undef FABS
#undef SQRT
#undef ATAN2
#undef COS
#undef SIN
#undef EXP
#undef FLOOR

/* ---- double ---- */
#define R double
#define FABS fabs
#define SQRT sqrt
#define ATAN2 atan2
#define COS cos
#define SIN sin
#define EXP exp
#define FLOOR floor

#define DIM 2
#define NM(x) x##_f64_2
#include "mpm_oracle_impl.inc"
#undef NM
#undef DIM

#define DIM 3
#define NM(x) x##_f64_3
#include "mpm_oracle_impl.inc"
#undef NM
#undef DIM

#undef R

/* ---- exported entry points ---- */

float  orc_weight_f32(float x, int order)   { return spline_f32_2(x, order); }
double orc_weight_f64(double x, int order)  { return spline_f64_2(x, order); }
float  orc_dweight_f32(float x, int order)  { return dspline_f32_2(x, order); }
double orc_dweight_f64(double x, int order) { return dspline_f64_2(x, order); }

void orc_svd2_f32(const float* M, float* U, float* s, float* V)    { svd_f32_2(M, U, s, V); }
void orc_svd2_f64(const double* M, double* U, double* s, double* V) { svd_f64_2(M, U, s, V); }
void orc_svd3_f32(const float* M, float* U, float* s, float* V)    { svd_f32_3(M, U, s, V); }
void orc_svd3_f64(const double* M, double* U, double* s, double* V) { svd_f64_3(M, U, s, V); }

void orc_stress_f32(const mpm_params* p, const float* F, float Jp, float* tau)
{
    if (p->dim == 3) stress_f32_3(p, F, Jp, tau); else stress_f32_2(p, F, Jp, tau);
}
void orc_stress_f64(const mpm_params* p, const double* F, double Jp, double* tau)
{
    if (p->dim == 3) stress_f64_3(p, F, Jp, tau); else stress_f64_2(p, F, Jp, tau);
}

#define DEFINE_STAGES(T, S)                                                                  \
void orc_p2g_##S(const mpm_params* p, long n, const T* x, const T* v, const T* mass,         \
                 const T* vol, const T* B, const T* Fe, const T* Fp, T* gmass, T* gmom,      \
                 T* gforce)                                                                  \
{                                                                                            \
    if (p->scheme == MPM_SCHEME_APIC_FORCE) {                                                \
        if (p->dim == 3) p2g_apic_##S##_3(p, n, x, v, mass, vol, B, Fe, Fp, gmass, gmom, gforce); \
        else             p2g_apic_##S##_2(p, n, x, v, mass, vol, B, Fe, Fp, gmass, gmom, gforce); \
        return;                                                                              \
    }                                                                                        \
    if (p->dim == 3) p2g_##S##_3(p, n, x, v, mass, vol, B, Fe, Fp, gmass, gmom);             \
    else             p2g_##S##_2(p, n, x, v, mass, vol, B, Fe, Fp, gmass, gmom);             \
}                                                                                            \
void orc_grid_update_##S(const mpm_params* p, const T* gmass, const T* gmom, const T* gforce,\
                         T* gvcol, T* gvfric)                                                \
{                                                                                            \
    if (p->scheme == MPM_SCHEME_APIC_FORCE) {                                                \
        if (p->dim == 3) grid_update_apic_##S##_3(p, gmass, gmom, gforce, gvcol, gvfric);    \
        else             grid_update_apic_##S##_2(p, gmass, gmom, gforce, gvcol, gvfric);    \
        return;                                                                              \
    }                                                                                        \
    if (p->dim == 3) grid_update_##S##_3(p, gmass, gmom, gvcol, gvfric);                     \
    else             grid_update_##S##_2(p, gmass, gmom, gvcol, gvfric);                     \
}                                                                                            \
void orc_g2p_##S(const mpm_params* p, long n, T* x, T* v, T* B, T* Fe, T* Fp,                \
                 const T* gvcol, const T* gvfric)                                            \
{                                                                                            \
    if (p->scheme == MPM_SCHEME_APIC_FORCE) {                                                \
        if (p->dim == 3) g2p_apic_##S##_3(p, n, x, v, B, Fe, Fp, gvcol, gvfric);             \
        else             g2p_apic_##S##_2(p, n, x, v, B, Fe, Fp, gvcol, gvfric);             \
        return;                                                                              \
    }                                                                                        \
    if (p->dim == 3) g2p_##S##_3(p, n, x, v, B, Fe, Fp, gvcol, gvfric);                      \
    else             g2p_##S##_2(p, n, x, v, B, Fe, Fp, gvcol, gvfric);                      \
}                                                                                            \
void orc_run_##S(const mpm_params* p, long n, int nsteps, T* x, T* v, const T* mass,         \
                 const T* vol, T* B, T* Fe, T* Fp, T* gmass, T* gmom, T* gvcol, T* gvfric)   \
{                                                                                            \
    long G = orc_num_nodes(p);                                                               \
    size_t gb = (size_t)G * sizeof(T), vb = (size_t)p->dim * G * sizeof(T);                  \
    T* gforce = p->scheme == MPM_SCHEME_APIC_FORCE ? (T*)calloc((size_t)p->dim * G, sizeof(T)) : NULL; \
    for (int s = 0; s < nsteps; ++s) {                                                       \
        orc_p2g_##S(p, n, x, v, mass, vol, B, Fe, Fp, gmass, gmom, gforce);                  \
        orc_grid_update_##S(p, gmass, gmom, gforce, gvcol, gvfric);                          \
        orc_g2p_##S(p, n, x, v, B, Fe, Fp, gvcol, gvfric);                                   \
        memset(gmass, 0, gb); memset(gmom, 0, vb);          /* reset.glsl, Grid::Reset */    \
        memset(gvcol, 0, vb); memset(gvfric, 0, vb);                                         \
        if (gforce) memset(gforce, 0, vb);                                                   \
    }                                                                                        \
    free(gforce);                                                                            \
}

DEFINE_STAGES(float, f32)
DEFINE_STAGES(double, f64)
