// mpm_super.hpp — super-block work items of the 3D fused substep (mpm_super.hip):
// binning keys, per-item bounds and the launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_kernels.hpp"

namespace mpm {

constexpr int SB_BLOCKS = 2;                 // grid blocks per axis of a super-block
constexpr int SB_CELLS = 4 * SB_BLOCKS;      // cells per axis of a super-block
constexpr int SB_KEY_SHIFT = 9;              // key bits below the super-block id: 8 blocks x 64 cells
constexpr int SB_MAX_PARTICLES = 4096;       // particles per work item at most

struct SuperWork {
    const int4* items;   // (super-block id, first particle, count, 0) per work item
    const int*  n_items;
    int         max_items;
    float2*     bound;   // per item: largest |m| and largest momentum bound of its particles
    float4*     ref;     // per item: its particles' mean position at binning, in cells from
                         // the super-block's first cell
    int4*       shift;   // per item: in x, y, z the tile's offset in cells, the rounded mean
                         // drift of its particles since binning; in w the z-range of their
                         // stencil bases (pack_zrange), from the positions the last substep
                         // scattered from, i.e. the ones the next gather reads at
    int         nsx, nsy;   // super-blocks along x and y
};

// Coordinates (in super-blocks) of super-block id sb.
__host__ __device__ inline void super_coords(int sb, int nsx, int nsy, int* sc)
{
    sc[0] = sb % nsx;
    const int rest = sb / nsx;
    sc[1] = rest % nsy;
    sc[2] = rest / nsy;
}

// Tile offset from a mean drift of d cells: its rounding, 0 where d is not finite or
// beyond 64 cells (the particles are then far from any one tile).
__host__ __device__ inline int drift_shift(float d)
{
    return fabsf(d) < 64.f ? (int)rintf(d) : 0;
}

// Lowest and highest z of an item's stencil bases (grid nodes) in one int, 16 bits
// each, clamped to the 16-bit range.  A tile stages and flushes only the node planes
// that range needs; an empty or stale range costs speed, never results (a particle
// outside the staged planes gathers from HBM and scatters with global atomics).
__host__ __device__ inline int pack_zrange(int lo, int hi)
{
    lo = lo < -32768 ? -32768 : (lo > 32767 ? 32767 : lo);
    hi = hi < -32768 ? -32768 : (hi > 32767 ? 32767 : hi);
    return (int)(((unsigned)hi << 16) | ((unsigned)lo & 0xffffu));
}
__host__ __device__ inline int zrange_lo(int packed) { return (int)(short)(packed & 0xffff); }
__host__ __device__ inline int zrange_hi(int packed) { return packed >> 16; }

// Super-block s along an axis holds cells [8 s - 7, 8 s]: the quadratic stencils of
// its particles span nodes [8 s - 8, 8 s + 2], so its tile starts on a grid-block
// boundary and each tile row covers 3 (not 4) blocks' 64-byte node runs.
constexpr int SB_CELL_SHIFT = SB_CELLS - 1;

// Super-block of cell c, and super-blocks per axis for cells 0..cells.
__host__ __device__ constexpr int super_of_cell(int c) { return (c + SB_CELL_SHIFT) / SB_CELLS; }
__host__ __device__ constexpr int super_count(int cells) { return super_of_cell(cells) + 1; }

// Binning key of cell (cx, cy, cz) (inside the grid): super-block-major, then the
// cell within the super-block in z-plane order (z slowest, x fastest), so that an
// item holding part of a super-block spans few of its cell planes; with cells = false
// the super-block alone (key >> SB_KEY_SHIFT of the full key), so that a stable sort
// keeps the previous order inside each super-block.
__host__ __device__ inline uint32_t super_key(int cx, int cy, int cz, int nsx, int nsy, bool cells = true)
{
    const uint32_t sb = ((uint32_t)super_of_cell(cz) * (uint32_t)nsy + (uint32_t)super_of_cell(cy)) * (uint32_t)nsx +
                        (uint32_t)super_of_cell(cx);
    if (!cells) return sb;
    const int lx = (cx + SB_CELL_SHIFT) & 7, ly = (cy + SB_CELL_SHIFT) & 7, lz = (cz + SB_CELL_SHIFT) & 7;
    return (sb << SB_KEY_SHIFT) | (uint32_t)((lz << 6) | (ly << 3) | lx);
}

hipError_t launch_cell_keys_super(const float* x, int n, const Phys& ph, int nsx, int nsy, bool cells,
                                  uint32_t* keys, hipStream_t st);
// Per item of w: the bound of its particles' stored state (what their scatter needs).
hipError_t launch_item_bounds(int order, const DevParticles& p, const SuperWork& w, const Phys& ph,
                              hipStream_t st);
// g2p: one fused substep (G2P from src, P2G into dst); otherwise the P2G of the
// stored state into dst.  margin: drift margin of the tile in cells (0 or 1).
// store_vc: the substep also stores each particle's v and C (x, F_E and J_P always);
// the next fused substep recomputes v and C by its gather, so only a substep that a
// reader of them follows (binning's item bounds, the caller) needs to.
hipError_t launch_super(const LaunchCfg& c, int margin, bool g2p, bool store_vc, const DevParticles& p,
                        const DevGrid& src, const DevGrid& dst, const SuperWork& w, const Phys& ph,
                        hipStream_t st);

}  // namespace mpm
