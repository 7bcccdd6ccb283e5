// mpm_kernels.hpp — host-side launchers of the substep kernels (mpm_kernels.hip,
// mpm_fused.hip, mpm_sort.hip).  Each launcher dispatches on (dim, spline order,
// boundary kind) to a template instantiation and enqueues on the given stream.
#pragma once

#include <hip/hip_runtime.h>
#include "mpm_device.hpp"

namespace mpm {

struct LaunchCfg {
    int dim, order, boundary;
};

// The reference's four dispatches (shaders/{2d,3d}/{p2g,grid_update,g2p,reset}.glsl).
hipError_t launch_p2g(const LaunchCfg& c, const DevParticles& p, const DevGrid& g,
                      const Phys& ph, hipStream_t st);
hipError_t launch_grid_update(const LaunchCfg& c, const DevGrid& g, const Phys& ph,
                              hipStream_t st);
hipError_t launch_g2p(const LaunchCfg& c, const DevParticles& p, const DevGrid& g,
                      const Phys& ph, hipStream_t st);
// nblocks: the blocks to scan — every scene's, one after another, for a batch.
hipError_t launch_reset(const LaunchCfg& c, const DevGrid& g, int nblocks, hipStream_t st);

// Grid update of `cur` fused with the clear of `old` (the grid the previous G2P
// read): both driven by the block flags, no dense sweep.
hipError_t launch_grid_update_clear(const LaunchCfg& c, const DevGrid& cur, const DevGrid& old,
                                    const Phys& ph, hipStream_t st);

// Fused G2P (from grid `src`) + P2G of the updated particles (into grid `dst`),
// tiled through LDS per particle block (mpm_fused.hip).
struct TileWork {
    const int4* items;   // (block id, first particle, count, 0) per work item
    const int*  n_items;
    int         max_items;
    int         store_vc;   // the substep stores v and C (x, F_E, J_P always): the next
                            // fused substep's gather recomputes both, so only a substep
                            // whose state the caller reads needs to
};
// scenes == nullptr: one scene with parameters ph; otherwise a batch whose items
// carry their scene index (item.w) into the scene table.
hipError_t launch_g2p2g(const LaunchCfg& c, const DevParticles& p, const DevGrid& src,
                        const DevGrid& dst, const TileWork& w, const Phys& ph, const SceneDesc* scenes,
                        hipStream_t st);
hipError_t launch_p2g_tiled(const LaunchCfg& c, const DevParticles& p, const DevGrid& dst,
                            const TileWork& w, const Phys& ph, const SceneDesc* scenes, hipStream_t st);
// Batch grid pass: like launch_grid_update_clear over the batch-wide block range,
// each block taking its scene's parameters (block_scene[b] indexes scenes).
hipError_t launch_grid_update_clear_batch(const LaunchCfg& c, const DevGrid& cur, const DevGrid& old,
                                          const SceneDesc* scenes, const int* block_scene, int nblocks,
                                          hipStream_t st);

// NaN/Inf scan of the particle state; count accumulated into *bad.
hipError_t launch_check_finite(int dim, const DevParticles& p, unsigned long long* bad,
                               hipStream_t st);

}  // namespace mpm
