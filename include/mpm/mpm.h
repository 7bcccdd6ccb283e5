/*
 * mpm/mpm.h — C-ABI of the MI355X MLS-MPM snow solver (libmpm_hip.so).
 *
 * The reference exposes its solver as a compile-time, duck-typed class with a
 * default constructor, void Step() and GetParticlePositions()
 * (src/gpu/GPUSolver.h:10-17, src/cpu/CPUSolver.h:10-27, selected at
 * src/main.cpp:22-29).  This header is the drop-in for that boundary, with the
 * stages of GPUSolver::Step (src/gpu/GPUSolver.cpp:63-80) exported one by one.
 *
 * Conventions
 *   - every entry point returns 0 on success and a negative MPM_E* code on error;
 *     mpm_last_error() returns the message of the last error on the calling thread;
 *     no C++ exception crosses this boundary;
 *   - a solver handle owns all of its device memory and one HIP stream; it is not
 *     thread-safe: callers serialise the calls on one handle;
 *   - host arrays are FP32, structure-of-arrays, component-major:
 *       x[d*n + p]            d < dim
 *       C[(r*dim + c)*n + p]  row-major dim x dim  (C = D^-1 B, the MLS affine matrix)
 *       F[(r*dim + c)*n + p]  elastic deformation gradient F_E
 *       Jp[p]                 det of the plastic deformation gradient F_P
 *     and always in the particle order given to mpm_create, whatever order the
 *     solver keeps internally;
 *   - grid arrays use the reference's node numbering k*(W+1)*(H+1) + j*(W+1) + i
 *     (src/gpu/shaders/3d/p2g.glsl:75), component-major for vectors.
 */
#ifndef MPM_MPM_H
#define MPM_MPM_H

#include <stdint.h>
#include "params.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_OK            0
#define MPM_E_INVALID    -1   /* bad argument or parameter                 */
#define MPM_E_HIP        -2   /* HIP runtime error (allocation, launch)    */
#define MPM_E_STATE      -3   /* call not valid in the solver's state      */
#define MPM_E_NONFINITE  -4   /* NaN/Inf detected by mpm_check_finite      */

typedef struct mpm_solver mpm_solver;

/* Initial particle set (host memory).  NULL optional fields take the reference's
 * initial values (GPUSolver.cpp:187-193): v = 0, C = 0, F = I, Jp = 1.
 * mass/vol: per-particle arrays; the reference computes them per shape as
 * vol = (shape area or volume)/count and mass = vol*RHO/100 (GPUSolver.cpp:117-119,147-148). */
typedef struct mpm_particles {
    int64_t      n;
    const float* x;
    const float* v;
    const float* mass;
    const float* vol;
    const float* C;
    const float* F;
    const float* Jp;
} mpm_particles;

/* Solver options beyond the physics parameters (the reference has none). */
typedef struct mpm_options {
    int sort_every;     /* re-bin and reorder particles every k substeps of mpm_step; 0 = never */
    int fused;          /* 0: the four reference stages; 1: the fused G2P+P2G pipeline over work
                         * items of one grid block; 2 / 3: the fused pipeline over work items of
                         * 2x2x2 grid blocks (one 3D scene only; MPM_E_INVALID otherwise) whose
                         * tile follows the item's mean drift since the last binning and holds
                         * particles up to 0.5 / 1.5 cells away from it (farther ones scatter
                         * with float atomics) */
    int tile_particles; /* particles per work item at most: 1..512 for fused 1 (0 = 512),
                         * 1..4096 for fused 2 and 3 (0 = 4096); longer runs split evenly     */
} mpm_options;

/* Parameters.h defaults (src/shared/Parameters.h:30-84) for dim 2 or 3. */
int mpm_params_reference(int dim, mpm_params* out);
/* Recompute MU and LAMBDA from E and NU as Parameters.h:82-84 does. */
int mpm_params_derive_lame(mpm_params* p);
/* Default options. */
int mpm_options_default(mpm_options* out);

/* GPUSolver::GPUSolver (GPUSolver.cpp:7-51): allocate and upload; opts may be NULL. */
int mpm_create(const mpm_params* params, const mpm_particles* init,
               const mpm_options* opts, mpm_solver** out);
/* GPUSolver::~GPUSolver (GPUSolver.cpp:53-60). */
int mpm_destroy(mpm_solver* s);

/* GPUSolver::Step (GPUSolver.cpp:63-80), n_substeps times, asynchronous on the
 * solver's stream.  MPM_E_STATE on a solver that has taken a learned-surrogate
 * macro step (mpm_surrogate_advance): from then on mpm_step, the stage calls,
 * mpm_sort, mpm_reset_grid and mpm_set_state are refused, and the state stays
 * readable (a solver created from mpm_get_state's output steps again). */
int mpm_step(mpm_solver* s, int n_substeps);

/* Stage-level drop-ins, one per dispatch of GPUSolver::Step:
 *   mpm_p2g          shaders/{2d,3d}/p2g.glsl           (adds into the grid)
 *   mpm_grid_update  shaders/{2d,3d}/grid_update.glsl
 *   mpm_g2p          shaders/{2d,3d}/g2p.glsl
 *   mpm_reset_grid   shaders/{2d,3d}/reset.glsl
 * Sequencing them p2g, grid_update, g2p, reset is one reference substep. */
int mpm_p2g(mpm_solver* s);
int mpm_grid_update(mpm_solver* s);
int mpm_g2p(mpm_solver* s);
int mpm_reset_grid(mpm_solver* s);

/* Particle binning: key = cell id in block-major order (with fused 2 or 3:
 * super-block-major, super-block s of an axis holding cells 8s-7 .. 8s, then the
 * cell within the super-block, z slowest; the periodic re-binning of mpm_step
 * then sorts by the super-block number alone, keeping the previous order inside
 * each super-block), stable LSD radix sort, reorder of every particle array.
 * mpm_get_sort returns the keys of the last binning.  No reference counterpart
 * (the reference scatters in particle order, shaders/2d/p2g.glsl:88-90). */
int mpm_sort(mpm_solver* s);
/* After mpm_sort: the sorted keys and, for each sorted slot, the original
 * particle index (n entries each; either pointer may be NULL). */
int mpm_get_sort(mpm_solver* s, uint32_t* keys, uint32_t* perm);

/* Wait for all work on the solver's stream (the reference's glFinish,
 * GPUSolver.cpp:278). */
int mpm_sync(mpm_solver* s);

/* GPUSolver::GetParticlePositions (GPUSolver.cpp:276-295): synchronises and copies
 * positions, interleaved [p*dim + d] as std::vector<Vec> holds them. */
int mpm_get_positions(mpm_solver* s, float* xyz);

/* Full state readback / upload (SoA layout above; NULL pointers are skipped). */
int mpm_get_state(mpm_solver* s, float* x, float* v, float* C, float* F, float* Jp);
int mpm_set_state(mpm_solver* s, const float* x, const float* v, const float* C,
                  const float* F, const float* Jp);

/* Grid readback in the reference's node numbering: after mpm_p2g mass and
 * momentum; after mpm_grid_update the collision and friction velocities
 * (gVelCol / gVelFric).  Any pointer may be NULL. */
int mpm_get_grid(mpm_solver* s, float* mass, float* mom, float* vcol, float* vfric);

/* CPU-path scheme only: the explicit grid force f_i = sum_p grad w_ip * A_p after
 * mpm_p2g (the force member of src/cpu/Grid.h:49), component-major. */
int mpm_get_grid_force(mpm_solver* s, float* force);

int mpm_num_particles(const mpm_solver* s, int64_t* n);
int mpm_num_nodes(const mpm_params* p, int64_t* nodes);

/* Device-side timing of the last mpm_step call (HIP events around it on the
 * solver's stream), milliseconds; and per-kernel average durations of the kernels
 * mpm_step launched since mpm_reset_timers (kernel ids: MPM_KERNEL_*). */
enum {
    MPM_KERNEL_P2G = 0,
    MPM_KERNEL_GRID_UPDATE = 1,
    MPM_KERNEL_G2P = 2,
    MPM_KERNEL_RESET = 3,
    MPM_KERNEL_G2P2G = 4,
    MPM_KERNEL_SORT = 5,
    MPM_KERNEL_COUNT = 6
};
int mpm_set_timing(mpm_solver* s, int enabled);
/* Timing of the kernels whose bit (1 << MPM_KERNEL_*) is set only (0: off): fewer
 * event records between launches when one kernel is measured.  No reference
 * counterpart (instrumentation of this library's own kernels). */
int mpm_set_timing_mask(mpm_solver* s, uint32_t kernels);
/* Time only every k-th launch of each timed kernel (1, the default: every launch),
 * counting from this call: fewer event records between launches, the same average
 * per launch.  No reference counterpart. */
int mpm_set_timing_every(mpm_solver* s, int every);
int mpm_reset_timers(mpm_solver* s);
int mpm_get_kernel_time(mpm_solver* s, int kernel, double* total_ms, int64_t* launches);

/* Scan the particle state for NaN/Inf (the reference's NaN trap,
 * src/cpu/Particles.cpp:50-61, as a status instead of a spin): returns
 * MPM_E_NONFINITE and the count of bad particles. */
int mpm_check_finite(mpm_solver* s, int64_t* bad_particles);

/* Batch of independent scenes in one solver (no reference counterpart: the
 * reference simulates one scene per process, src/main.cpp:22-29).  Scene k has its
 * own parameters params[k] and particles init[k]; all scenes share dim, spline
 * order, boundary kind and the MLS scheme.  Every scene evolves as it would in a
 * solver of its own (up to the run-to-run rounding order of floating-point grid
 * accumulation, which a single scene has too); one mpm_step advances all of them
 * with the same kernels.
 * Particle arrays of the batch (mpm_get_state, mpm_set_state, mpm_get_positions)
 * hold the scenes one after another, scene k at [first, first+count) as returned by
 * mpm_scene_range.  The stage-level calls and the grid readbacks are single-scene
 * only (MPM_E_STATE on a batch). */
int mpm_batch_create(int nscenes, const mpm_params* params, const mpm_particles* init,
                     const mpm_options* opts, mpm_solver** out);
int mpm_num_scenes(const mpm_solver* s, int* nscenes);
int mpm_scene_range(const mpm_solver* s, int scene, int64_t* first, int64_t* count);

/* Graph replay (no reference counterpart): with enabled != 0 mpm_step records one
 * hipGraph of 2*sort_every fused substeps (binning included) the first time it can
 * and replays it for every whole period of later calls; the remainder runs eagerly.
 * The replay launches the same kernels on the same buffers as eager stepping.
 * Per-kernel timers do not see replayed substeps. */
int mpm_set_graph(mpm_solver* s, int enabled);

/* Autotuner (no reference counterpart): times trial_substeps of the fused pipeline
 * (rounded up to whole binning periods) on fresh solvers built from the same input,
 * over tile_particles in {128, 256, 512} x sort_every in {4, 8, 16, 32, 64} with
 * fused = 1, and for one 3D scene also fused in {2, 3} x tile_particles in
 * {1024, 2048, 4096} x sort_every in {8, 16, 32, 64, 128}, and returns the fastest options and
 * their time per substep.  nscenes = 1 tunes a single scene (params/init point to
 * one element).  The input is not modified. */
int mpm_autotune(int nscenes, const mpm_params* params, const mpm_particles* init, int trial_substeps,
                 int use_graph, mpm_options* best, double* best_ms_per_substep);

/* ---------------------------------------------------------------------------
 * Learned grid surrogate (no reference counterpart).  A 4-level BF16 U-Net
 * (32-64-128-128 channels, 3x3 convolutions, GELU; FP32 accumulation on the matrix
 * cores) over the grid nodes of a 2D quadratic-spline scene of height x width cells
 * predicts the grid velocity (and, optionally, the grid displacement: a second
 * image of the batch, told apart by a constant mode plane) k substeps ahead from
 * node features
 * [mass, velocity, mean J_P, mean log singular value of F_E].  Its training set is
 * captured from solvers into device-resident slots (targets: transfers, optionally
 * least-squares fits, of the particles' velocities and displacements), training
 * runs on the device
 * (forward, backward, Adam), and mpm_surrogate_advance moves a single-scene solver
 * k substeps with the predicted fields or, when the residual test fails, with k
 * plain substeps.
 */
typedef struct mpm_surrogate mpm_surrogate;

enum { MPM_ADVANCE_PLAIN = 0, MPM_ADVANCE_NETWORK = 1, MPM_ADVANCE_BALLISTIC = 2 };

/* height, width: image of grid nodes (multiples of 8; a scene must have WIDTH =
 * width and HEIGHT = height cells); max_batch: slots per training step (one or two
 * network images each; the host-data calls below take up to max_batch images and
 * mpm_surrogate_time up to 2 max_batch); capacity:
 * training-set slots; seed: weight initialisation. */
int mpm_surrogate_create(int height, int width, int max_batch, int capacity, uint64_t seed,
                         mpm_surrogate** out);
int mpm_surrogate_destroy(mpm_surrogate* g);
/* Parameters, FP32: per convolution layer (e0 e1 e2 e3 d2 d1 d0) W[9][cout][cin]
 * (tap = 3*(dy+1) + (dx+1)) then b[cout]; then the head wh[2][32], bh[2]. */
int mpm_surrogate_num_params(const mpm_surrogate* g, int64_t* n);
int mpm_surrogate_get_params(mpm_surrogate* g, float* host);
int mpm_surrogate_set_params(mpm_surrogate* g, const float* host);
/* Velocity scale of the network output (m/s; default 0.1). */
int mpm_surrogate_set_scale(mpm_surrogate* g, float d_ref);
/* Kernel of the forward 3x3 convolutions (no reference counterpart: the surrogate
 * is this library's own): MPM_CONV_AUTO (default) chooses by layer shape,
 * MPM_CONV_STREAM the register-streaming kernel, MPM_CONV_LDS the LDS-tiled one.
 * Results agree up to the order of FP32 accumulation. */
enum { MPM_CONV_AUTO = 0, MPM_CONV_STREAM = 1, MPM_CONV_LDS = 2 };
int mpm_surrogate_set_conv(mpm_surrogate* g, int kernel);
/* Network on host data: input [nb][h][w][32] BF16 bit patterns, wm [nb][h*w]
 * projection weights (summing to one per image), target [nb][h*w][2].  forward
 * returns the projected output [nb][h*w][2]; gradient returns the loss
 * (1/nb) sum_i wm_i |out_i - target_i|^2 and, when grad != NULL, dloss/dparams. */
int mpm_surrogate_forward(mpm_surrogate* g, int nb, const uint16_t* input, const float* wm, float* out);
int mpm_surrogate_gradient(mpm_surrogate* g, int nb, const uint16_t* input, const float* wm,
                           const float* target, float* loss, float* grad);
/* Capture scene `scene` of solver s into training slot `slot`: what = 0 the node
 * features now (and the particle positions), what = 1 the target velocity and
 * displacement now (since the last what = 0 capture), as grid fields at the
 * positions of that capture (refined per mpm_surrogate_set_targets); k_substeps =
 * the horizon. */
int mpm_surrogate_record(mpm_surrogate* g, mpm_solver* s, int scene, int slot, int what, int k_substeps);
/* Raw sample of a slot: host [9][h*w] (mass, vx, vy, J_P, log sigma, target vx, vy,
 * target displacement x, y)
 * and scalars [6] (particle mass, grid mass, wall-band mass, horizon T, gravity,
 * m_ref); either may be NULL.  slot == capacity reads the advance call's sample. */
int mpm_surrogate_get_sample(mpm_surrogate* g, int slot, float* host, float* scalars);
/* One Adam step (lr > 0) on the batch of slots, or the loss only (lr == 0). */
int mpm_surrogate_train(mpm_surrogate* g, const int* slots, int nb, float lr, float* loss);
/* Loss of the zero increment (ballistic grid velocity and displacement) on the
 * same samples. */
int mpm_surrogate_baseline_loss(mpm_surrogate* g, const int* slots, int nb, float* loss);
/* Which predicted field the macro step of mpm_surrogate_advance takes each update
 * from (no reference counterpart): MPM_MACRO_DISPLACEMENT_X moves the particles by
 * the gathered displacement (else x' = x + T (v + v') / 2), MPM_MACRO_DISPLACEMENT_F
 * advances F_E by the displacement gradient (else by T C').  v' and C' always come
 * from the predicted velocity.  With either flag the network runs on two images per
 * sample (velocity and displacement) in training and in the macro step; set it
 * before training.  Default 0: velocity only, one image. */
enum { MPM_MACRO_DISPLACEMENT_X = 1, MPM_MACRO_DISPLACEMENT_F = 2 };
int mpm_surrogate_set_macro(mpm_surrogate* g, int flags);
/* Jacobi steps (0..64) of the mass-weighted least-squares fit applied to the target
 * fields of later what = 1 captures (no reference counterpart); 0 (default) keeps
 * the transfers sum_p w_ip m_p y_p / m_i. */
int mpm_surrogate_set_targets(mpm_surrogate* g, int refine_iters);
/* Advance a single-scene solver by k_substeps: mode MPM_ADVANCE_NETWORK or
 * MPM_ADVANCE_BALLISTIC (zero increment) when the residual (share of the mass in
 * the wall band or outside the image) is <= tol, else k plain substeps;
 * MPM_ADVANCE_PLAIN always steps.  *used: the mode that ran; *residual: the test's
 * value.  Plain substeps are taken only before the solver's first macro step:
 * afterwards MPM_ADVANCE_PLAIN and a failed residual test return MPM_E_STATE (see
 * mpm_step). */
int mpm_surrogate_advance(mpm_surrogate* g, mpm_solver* s, int k_substeps, int mode, float tol, int* used,
                          float* residual);

/* n_macro macro steps of k_substeps each, as n_macro calls of mpm_surrogate_advance
 * would take them (no reference counterpart): until the solver has taken a macro
 * step, each goes through that call (a failed residual test runs plain substeps);
 * the rest replay one recorded hipGraph of the macro step whose residual test runs
 * on the device, so the host does not wait between them.  The first of those whose
 * test fails leaves the particles as they were before it and the call returns
 * MPM_E_STATE.  *n_network: macro steps taken (network or ballistic); *n_plain:
 * plain fallbacks; *max_residual: the largest residual of the macro steps taken.
 * MPM_ADVANCE_PLAIN runs k_substeps * n_macro plain substeps. */
int mpm_surrogate_rollout(mpm_surrogate* g, mpm_solver* s, int k_substeps, int n_macro, int mode, float tol,
                          int* n_network, int* n_plain, float* max_residual);

/* Average device time (ms) of `iters` network passes on nb images of the batch
 * buffers: forward only (train = 0) or forward + loss + backward (train = 1). */
int mpm_surrogate_time(mpm_surrogate* g, int nb, int iters, int train, double* ms);

/* ---------------------------------------------------------------------------
 * State snapshots (the state GPUSolver::ReadbackParticles and GetParticlePositions
 * read back, src/gpu/GPUSolver.cpp:204-225, 276-295), single-scene solvers.  File:
 * "MPMSNAP1", uint32 version 1, uint32 sizeof(mpm_params), mpm_params, int64 n,
 * int64 substeps, then FP32 x[d][n], v[d][n], C[d*d][n], F[d*d][n], Jp[n], mass[n],
 * vol[n] in the caller's particle order (little-endian). */
int mpm_save_snapshot(mpm_solver* s, const char* path);
/* Header only (no device needed); any pointer may be NULL. */
int mpm_snapshot_info(const char* path, mpm_params* params, int64_t* n, int64_t* substeps);
/* A new solver holding the snapshot's parameters and state (opts may be NULL). */
int mpm_load_snapshot(const char* path, const mpm_options* opts, mpm_solver** out);

/* Headless frames (the reference's Visualize + saveFrame, src/main.cpp:56-160,
 * 177-205): white points on black, a grey border (2D) or wireframe box (3D) at
 * BORDER_MARGIN.  2D maps the domain with glOrtho(0, WIDTH, 0, HEIGHT); 3D looks at
 * the grid centre from a camera orbiting at `distance` (domain units: 300 for the
 * reference's 200x100x100 grid units) with a perspective projection.  rgb receives
 * width*height*3 bytes, top row first. */
typedef struct mpm_camera {
    float angle_deg;       /* horizontal orbit angle (main.cpp:51)  */
    float elevation_deg;   /* vertical angle         (main.cpp:52)  */
    float distance;        /* from the grid centre   (main.cpp:53)  */
    float fov_deg;         /* vertical field of view (main.cpp:62)  */
    float z_near, z_far;
} mpm_camera;
int mpm_camera_default(mpm_camera* cam);
int mpm_render_frame(mpm_solver* s, const mpm_camera* cam, int width, int height, int point_size,
                     unsigned char* rgb);
/* 8-bit RGB PNG (stored deflate blocks), top row first. */
int mpm_write_png(const char* path, int width, int height, const unsigned char* rgb);

/* Stable sort of n 32-bit keys on the device (the binning sort on its own, for
 * callers with their own keys).  Host arrays; out_perm[i] = index of the i-th
 * smallest key, ties in input order. key_bits <= 32. */
int mpm_sort_pairs_host(const uint32_t* keys, int64_t n, int key_bits,
                        uint32_t* out_keys, uint32_t* out_perm);

/* No counterpart in the reference: mpm_sort_pairs_host for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors; it calls no hipMalloc, hipFree or stream
 * creation: its device keys, values and radix scratch live in a fixed device array
 * of the library (about 17 MB, held on each GPU it runs on until the process ends)
 * and it runs on the null stream.  n above MPM_SORT_NOALLOC_MAX_KEYS returns
 * MPM_E_INVALID.  Calls from several host threads are serialised; every call
 * overwrites the array. */
#define MPM_SORT_NOALLOC_MAX_KEYS (1 << 20)
int mpm_sort_pairs_host_noalloc(const uint32_t* keys, int64_t n, int key_bits,
                                uint32_t* out_keys, uint32_t* out_perm);

const char* mpm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MPM_MPM_H */
