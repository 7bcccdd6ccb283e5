// mpm_sort.hpp — launchers of the binning kernels (mpm_sort.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "mpm_device.hpp"

namespace mpm {

// Radix-sort tiling: 256-thread workgroups of 16 rounds (4096 keys per tile), and
// the block of the exclusive scan (1024 entries).
constexpr int RS_THREADS = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER = 4;
constexpr int SCAN_BLOCK = SCAN_THREADS * SCAN_PER;

// uint32 words of scratch radix_sort_pairs needs for n keys: per-tile digit counts,
// the scan's block sums, and slack.
constexpr size_t radix_scratch_words(int n)
{
    const size_t ntiles = ((size_t)n + RS_TILE - 1) / RS_TILE;
    const size_t counts = 256 * ntiles;
    const size_t sums = (counts + SCAN_BLOCK - 1) / SCAN_BLOCK + 1;
    return counts + sums + 64;
}

// Stable LSD radix sort of n (key, val) pairs over the low key_bits bits.  Ping-pongs
// between (keys, vals) and (keys_alt, vals_alt); *result_in_alt says where the
// sorted pairs ended.  init_vals: fill vals with 0..n-1 first.
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                            int n, int key_bits, bool init_vals, uint32_t* scratch, hipStream_t st,
                            bool* result_in_alt);

hipError_t launch_cell_keys(int dim, const float* x, int n, const Phys& ph, const DevGrid& g,
                            uint32_t* keys, hipStream_t st);
hipError_t launch_gather_fields(const float* src, float* dst, const uint32_t* perm, int n, int ncomp,
                                hipStream_t st);
hipError_t launch_gather_u32(const uint32_t* src, uint32_t* dst, const uint32_t* perm, int n,
                             hipStream_t st);
// Reorder a whole particle buffer (2 dim + 2 dim^2 + 3 float rows of n, then the u32
// row of original indices, contiguous from src / dst) by perm in one launch.
hipError_t launch_gather_particles(int dim, const float* src, float* dst, const uint32_t* perm, int n,
                                   hipStream_t st);
// Work items (block id = key >> shift, first particle, count, scene) from sorted
// keys, in key order; tmp_runlen and tmp_counts hold n words each, scan_sums
// n/1024 + 1; block_scene (batch-wide block -> scene) may be NULL for one scene.
hipError_t launch_emit_items(const uint32_t* keys, int n, int tile_p, int shift, int4* items, int* n_items,
                             int max_items, uint32_t* tmp_runlen, uint32_t* tmp_counts, uint32_t* scan_sums,
                             const int* block_scene, hipStream_t st);
hipError_t launch_cell_keys_batch(int dim, const float* x, int n, const SceneDesc* scenes, const int* p_off,
                                  int nscenes, uint32_t* keys, hipStream_t st);

}  // namespace mpm
