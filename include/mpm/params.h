/*
 * mpm/params.h — the solver's run-time parameter set.
 *
 * Mirrors the compile-time constants of the reference's src/shared/Parameters.h
 * (same names, same meaning) and extends them with the fields the reference hard
 * codes or fixes by build mode:
 *
 *   WIDTH, HEIGHT, DEPTH       Parameters.h:30-36   grid size in cells (nodes 0..WIDTH)
 *   BORDER_MARGIN              Parameters.h:33      wall inset in cells
 *   GRAVITY, FRICTION          Parameters.h:58-60   gravity along +y, Coulomb mu
 *   DELTA_TIME                 Parameters.h:68      substep length
 *   THETA_C, THETA_S, XI       Parameters.h:74-76   Stomakhin clamp and hardening
 *   NU, E, RHO                 Parameters.h:77-79
 *   MU, LAMBDA                 Parameters.h:82-84   Lame coefficients (derived from E, NU)
 *
 * Extensions (the reference fixes these in code):
 *   dim            2 or 3 (reference: CMake USE_3D, CMakeLists.txt:17-24)
 *   spline_order   3 = cubic B-spline, D^-1 = 3/dx^2 (reference, shaders/2d/common.glsl:43-69,
 *                  shaders/2d/p2g.glsl:82); 2 = quadratic B-spline, D^-1 = 4/dx^2
 *   boundary       MPM_BOUNDARY_COULOMB_MARGIN: the reference's predictive half-space walls
 *                  with Coulomb friction (shaders/2d/grid_update.glsl:17-80);
 *                  MPM_BOUNDARY_STICKY: nodes closer than BORDER_MARGIN cells to the domain
 *                  edge get zero velocity
 *   dx             grid spacing (reference: 1, grid units)
 *   mass_eps       grid nodes lighter than this are skipped by the grid update
 *                  (reference: 1e-6, shaders/2d/grid_update.glsl:90)
 *   scheme         MPM_SCHEME_MLS: the GPU path's MLS-MPM, stress folded into the P2G
 *                  momentum, F_E advanced with C (shaders/2d/p2g.glsl:84-86,
 *                  g2p.glsl:74-75);
 *                  MPM_SCHEME_APIC_FORCE: the CPU path's APIC with an explicit grid
 *                  force f_i = sum_p V_p tau_p grad w (src/cpu/Grid.cpp:148-159,
 *                  175-179) and F_E advanced with grad v = sum_i v_col grad w
 *                  (src/cpu/Particles.cpp:200-214)
 */
#ifndef MPM_PARAMS_H
#define MPM_PARAMS_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MPM_BOUNDARY_COULOMB_MARGIN = 0,
    MPM_BOUNDARY_STICKY = 1
};

enum {
    MPM_SCHEME_MLS = 0,
    MPM_SCHEME_APIC_FORCE = 1
};

typedef struct mpm_params {
    int   dim;
    int   WIDTH, HEIGHT, DEPTH;
    int   spline_order;
    int   boundary;
    float BORDER_MARGIN;
    float dx;
    float GRAVITY;
    float FRICTION;
    float DELTA_TIME;
    float THETA_C, THETA_S, XI;
    float NU, E, RHO;
    float MU, LAMBDA;
    float mass_eps;
    int   scheme;
} mpm_params;

#ifdef __cplusplus
}
#endif

#endif /* MPM_PARAMS_H */
