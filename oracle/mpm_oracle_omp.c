/*
 * oracle/mpm_oracle_omp.c — the FP32 oracle substep on every host core (OpenMP),
 * for the CPU-baseline leg of bench.py only.
 *
 * TEST INFRASTRUCTURE ONLY.  The same body as the serial oracle
 * (mpm_oracle_impl.inc, the restatement of src/gpu/shaders/{2d,3d}/{p2g,grid_update,g2p}.glsl and of
 * src/gpu/GPUSolver.cpp:63-80), with the loops over particles and grid nodes run
 * in parallel and the P2G grid sums as atomic adds.  The reference's CPU path is
 * serial (src/cpu/Grid.cpp:120, src/cpu/Particles.cpp:135,175); this build is the
 * all-cores baseline beside it.  Its grid sums add in thread order, so it is not
 * bit-reproducible and is never used as the checker.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "mpm_oracle.h"

#define ORC_PAR_FOR _Pragma("omp parallel for schedule(static)")
#define ORC_ATOMIC _Pragma("omp atomic")

long orc_num_nodes(const mpm_params* p)
{
    long n = (long)(p->WIDTH + 1) * (long)(p->HEIGHT + 1);
    if (p->dim == 3) n *= (long)(p->DEPTH + 1);
    return n;
}

#define R float
#define FABS fabsf
#define SQRT sqrtf
#define ATAN2 atan2f
#define COS cosf
#define SIN sinf
#define EXP expf
#define FLOOR floorf

#define DIM 2
#define NM(x) x##_omp_2
#include "mpm_oracle_impl.inc"
#undef NM
#undef DIM

#define DIM 3
#define NM(x) x##_omp_3
#include "mpm_oracle_impl.inc"
#undef NM
#undef DIM

/* Threads a parallel region of this library runs on. */
int orc_omp_threads(void) { return omp_get_max_threads(); }

static void zero_par(float* a, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)((n + 65535) / 65536); ++c) {
        size_t lo = (size_t)c * 65536, hi = lo + 65536 < n ? lo + 65536 : n;
        memset(a + lo, 0, (hi - lo) * sizeof(float));
    }
}

/* nsteps substeps (p2g -> grid update -> g2p -> reset) of the MLS scheme, or of the
 * CPU path's APIC scheme with the explicit grid force (src/cpu/Particles.cpp:115-131,
 * 2D), with the layout and grid scratch of orc_run_f32 (the force grid is
 * allocated here). */
void orc_run_omp_f32(const mpm_params* p, long n, int nsteps, float* x, float* v, const float* mass,
                     const float* vol, float* B, float* Fe, float* Fp, float* gmass, float* gmom, float* gvcol,
                     float* gvfric)
{
    const long G = orc_num_nodes(p);
    const size_t vn = (size_t)p->dim * G;
    if (p->scheme == MPM_SCHEME_APIC_FORCE) {
        float* gforce = (float*)calloc(vn, sizeof(float));
        if (!gforce) return;
        for (int s = 0; s < nsteps; ++s) {
            if (p->dim == 3) {
                p2g_apic_omp_3(p, n, x, v, mass, vol, B, Fe, Fp, gmass, gmom, gforce);
                grid_update_apic_omp_3(p, gmass, gmom, gforce, gvcol, gvfric);
                g2p_apic_omp_3(p, n, x, v, B, Fe, Fp, gvcol, gvfric);
            } else {
                p2g_apic_omp_2(p, n, x, v, mass, vol, B, Fe, Fp, gmass, gmom, gforce);
                grid_update_apic_omp_2(p, gmass, gmom, gforce, gvcol, gvfric);
                g2p_apic_omp_2(p, n, x, v, B, Fe, Fp, gvcol, gvfric);
            }
            zero_par(gmass, (size_t)G);
            zero_par(gmom, vn);
            zero_par(gforce, vn);
            zero_par(gvcol, vn);
            zero_par(gvfric, vn);
        }
        free(gforce);
        return;
    }
    for (int s = 0; s < nsteps; ++s) {
        if (p->dim == 3) {
            p2g_omp_3(p, n, x, v, mass, vol, B, Fe, Fp, gmass, gmom);
            grid_update_omp_3(p, gmass, gmom, gvcol, gvfric);
            g2p_omp_3(p, n, x, v, B, Fe, Fp, gvcol, gvfric);
        } else {
            p2g_omp_2(p, n, x, v, mass, vol, B, Fe, Fp, gmass, gmom);
            grid_update_omp_2(p, gmass, gmom, gvcol, gvfric);
            g2p_omp_2(p, n, x, v, B, Fe, Fp, gvcol, gvfric);
        }
        zero_par(gmass, (size_t)G);   /* reset.glsl, Grid::Reset */
        zero_par(gmom, vn);
        zero_par(gvcol, vn);
        zero_par(gvfric, vn);
    }
}
