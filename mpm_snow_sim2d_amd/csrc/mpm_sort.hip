// mpm_sort.hip — particle binning: cell keys, a stable LSD radix sort of
// (key, slot) pairs, the reorder of every particle field, and the work list of
// (block, particle range) items the tiled kernels consume.
//
// No reference counterpart: the reference scatters in particle order with float
// atomics (src/gpu/shaders/2d/p2g.glsl:88-90).  The sort is integer work and
// deterministic: equal keys keep their input order (checked bit-exact against a
// CPU stable sort).
//
// Radix pass (8-bit digits, tiles of 4096 keys per 256-thread workgroup):
//   k_rs_hist     per-tile digit counts           -> counts[digit * ntiles + tile]
//   k_scan_*      exclusive scan of counts (digit-major, so the result is the
//                 global output offset of each (digit, tile))
//   k_rs_scatter  per tile, 16 rounds of 256 keys in input order; the stable
//                 rank of a key inside a round comes from a wavefront match on its
//                 digit (8 ballots) plus per-wave counts in LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mpm_sort.hpp"

namespace mpm {

static constexpr int RS_WAVES = RS_THREADS / 64;

__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint32_t* keys, int n, int shift,
                                                       uint32_t* counts, int ntiles)
{
    __shared__ uint32_t h[256];
    int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    int base = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        int i = base + r * RS_THREADS + t;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[t * ntiles + blockIdx.x] = h[t];
}

// Exclusive scan, three kernels: per-block scan of 1024 entries, scan of the block
// sums (one workgroup, any count), add-back.

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total)
{
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        uint32_t add = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    total = sh[SCAN_THREADS - 1];
    uint32_t incl = sh[t];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_local(uint32_t* data, int n, uint32_t* sums)
{
    __shared__ uint32_t sh[SCAN_THREADS];
    int t = threadIdx.x;
    int base = blockIdx.x * SCAN_BLOCK + t * SCAN_PER;
    uint32_t v[SCAN_PER], s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q) {
        v[q] = (base + q < n) ? data[base + q] : 0u;
        s += v[q];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan(s, sh, total);
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q) {
        if (base + q < n) data[base + q] = ex;
        ex += v[q];
    }
    if (t == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_sums(uint32_t* sums, int nsums)
{
    __shared__ uint32_t sh[SCAN_THREADS];
    uint32_t carry = 0;
    for (int base = 0; base < nsums; base += SCAN_BLOCK) {
        int t = threadIdx.x;
        int i0 = base + t * SCAN_PER;
        uint32_t v[SCAN_PER], s = 0;
#pragma unroll
        for (int q = 0; q < SCAN_PER; ++q) {
            v[q] = (i0 + q < nsums) ? sums[i0 + q] : 0u;
            s += v[q];
        }
        uint32_t total;
        uint32_t ex = block_exclusive_scan(s, sh, total) + carry;
#pragma unroll
        for (int q = 0; q < SCAN_PER; ++q) {
            if (i0 + q < nsums) sums[i0 + q] = ex;
            ex += v[q];
        }
        carry += total;
    }
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_add(uint32_t* data, int n, const uint32_t* sums)
{
    int base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER;
    uint32_t add = sums[blockIdx.x];
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q)
        if (base + q < n) data[base + q] += add;
}

__device__ inline uint64_t lanemask_lt()
{
    int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint32_t* keys_in, const uint32_t* vals_in,
                                                          uint32_t* keys_out, uint32_t* vals_out, int n,
                                                          int shift, const uint32_t* offsets, int ntiles)
{
    __shared__ uint32_t gbase[256];               // global offset of (digit, this tile)
    __shared__ uint32_t running[256];             // keys of each digit already placed
    __shared__ uint32_t wcount[RS_WAVES][256];    // this round's count per wave and digit
    int t = threadIdx.x, wave = t >> 6;
    gbase[t] = offsets[t * ntiles + blockIdx.x];
    running[t] = 0;
    int base = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) wcount[w][t] = 0;
        __syncthreads();
        int i = base + r * RS_THREADS + t;
        bool valid = i < n;
        uint32_t key = valid ? keys_in[i] : 0u;
        uint32_t val = valid ? vals_in[i] : 0u;
        uint32_t dg = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bits = __ballot((dg >> b) & 1u);
            peers &= ((dg >> b) & 1u) ? bits : ~bits;
        }
        uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
        if (valid && below == 0) wcount[wave][dg] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = gbase[dg] + running[dg] + below;
            for (int w = 0; w < wave; ++w) pos += wcount[w][dg];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t tot = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) tot += wcount[w][t];
        running[t] += tot;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_iota(uint32_t* v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
static hipError_t exclusive_scan(uint32_t* data, int n, uint32_t* sums, hipStream_t st)
{
    int nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb == 0) return hipSuccess;
    k_scan_local<<<nb, SCAN_THREADS, 0, st>>>(data, n, sums);
    k_scan_sums<<<1, SCAN_THREADS, 0, st>>>(sums, nb);
    k_scan_add<<<nb, SCAN_THREADS, 0, st>>>(data, n, sums);
    return hipGetLastError();
}

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                            int n, int key_bits, bool init_vals, uint32_t* scratch, hipStream_t st,
                            bool* result_in_alt)
{
    *result_in_alt = false;
    if (init_vals && n > 0) k_iota<<<(n + 255) / 256, 256, 0, st>>>(vals, n);
    if (n <= 0) return hipGetLastError();
    int ntiles = (n + RS_TILE - 1) / RS_TILE;
    uint32_t* counts = scratch;
    uint32_t* sums = scratch + 256 * (size_t)ntiles;
    uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
    int passes = (key_bits + 7) / 8;
    for (int ps = 0; ps < passes; ++ps) {
        int shift = ps * 8;
        k_rs_hist<<<ntiles, RS_THREADS, 0, st>>>(ki, n, shift, counts, ntiles);
        hipError_t e = exclusive_scan(counts, 256 * ntiles, sums, st);
        if (e != hipSuccess) return e;
        k_rs_scatter<<<ntiles, RS_THREADS, 0, st>>>(ki, vi, ko, vo, n, shift, counts, ntiles);
        uint32_t* tk = ki; ki = ko; ko = tk;
        uint32_t* tv = vi; vi = vo; vo = tv;
        *result_in_alt = !*result_in_alt;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Cell keys: the cell containing the particle, numbered like the block-major grid
// nodes (key = node address of the cell's lower corner), clamped into the grid.
template <int DIM>
__global__ void __launch_bounds__(256) k_cell_keys(const float* x, int n, float inv_dx, int W, int H,
                                                  int D, DevGrid g, uint32_t* keys)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int c[3] = {0, 0, 0};
    int lim[3] = {W, H, D};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        float xg = x[d * n + p] * inv_dx;
        int ci = cell_floor(xg);
        c[d] = min(max(ci, 0), lim[d]);
    }
    keys[p] = (uint32_t)node_addr<DIM>(g, c[0], c[1], c[2]);
}

// Batch: the slot's scene by binary search in the scene particle offsets (binning
// never moves a particle across scenes: keys are scene-major), then the scene-local
// key shifted by the scene's first block.
template <int DIM>
__global__ void __launch_bounds__(256) k_cell_keys_batch(const float* x, int n, const SceneDesc* scenes,
                                                        const int* p_off, int nscenes, uint32_t* keys)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int lo = 0, hi = nscenes - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (p_off[mid] <= p) lo = mid; else hi = mid - 1;
    }
    const SceneDesc sd = scenes[lo];
    DevGrid g;
    g.nbx = sd.nbx; g.nby = sd.nby; g.nbz = sd.nbz;
    int c[3] = {0, 0, 0};
    int lim[3] = {sd.ph.W, sd.ph.H, sd.ph.D};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        float xg = x[d * n + p] * sd.ph.inv_dx;
        int ci = cell_floor(xg);
        c[d] = min(max(ci, 0), lim[d]);
    }
    keys[p] = (uint32_t)(sd.block_off * 64 + node_addr<DIM>(g, c[0], c[1], c[2]));
}

// Gather every particle field into sorted order: dst[s] = src[perm[s]].
__global__ void __launch_bounds__(256) k_gather_fields(const float* __restrict__ src, float* __restrict__ dst,
                                                      const uint32_t* perm, int n, int ncomp)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t q = perm[s];
    for (int c = 0; c < ncomp; ++c) dst[(size_t)c * n + s] = src[(size_t)c * n + q];
}

// Every particle array in one pass: the solver's particle buffer holds `rows` float
// rows of n (x, v, C, F, J_P, mass, volume) followed by the u32 row of original
// indices; dst[r][s] = src[r][perm[s]] for every row (one index load per particle,
// all of its field loads in flight together).
template <int ROWS>
__global__ void __launch_bounds__(256) k_gather_rows(const float* __restrict__ src, float* __restrict__ dst,
                                                    const uint32_t* __restrict__ perm, int n)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const size_t q = perm[s];
    float v[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) v[r] = src[(size_t)r * n + q];
    const uint32_t id = reinterpret_cast<const uint32_t*>(src + (size_t)ROWS * n)[q];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) dst[(size_t)r * n + s] = v[r];
    reinterpret_cast<uint32_t*>(dst + (size_t)ROWS * n)[s] = id;
}

__global__ void __launch_bounds__(256) k_gather_u32(const uint32_t* __restrict__ src,
                                                   uint32_t* __restrict__ dst, const uint32_t* perm, int n)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) dst[s] = src[perm[s]];
}

// Work items: every run of equal block id (key >> shift: 6 for grid blocks, 9 for
// the super-blocks of mpm_super.hip) in the sorted keys, split into the fewest
// chunks of at most tile_p particles, of equal size (within one particle); item.w
// is the block's scene (0 for one scene).  Deterministic: a run start writes its run length and chunk count, an
// exclusive scan of the counts places the chunks, so the items come out in key
// (spatial) order.
__global__ void __launch_bounds__(256) k_item_counts(const uint32_t* keys, int n, int tile_p, int shift,
                                                    uint32_t* runlen, uint32_t* counts)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t b = keys[s] >> shift;
    if (s > 0 && (keys[s - 1] >> shift) == b) { runlen[s] = 0; counts[s] = 0; return; }
    // end of the run: first index with block id > b (binary search, keys sorted)
    uint32_t next = (b + 1u) << shift;
    int lo = s + 1, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (keys[mid] < next) lo = mid + 1; else hi = mid;
    }
    uint32_t len = (uint32_t)(lo - s);
    runlen[s] = len;
    counts[s] = (len + tile_p - 1) / tile_p;
}

__global__ void __launch_bounds__(256) k_item_write(const uint32_t* keys, int n, int tile_p, int shift,
                                                   const uint32_t* runlen, const uint32_t* offsets,
                                                   int4* items, int* n_items, int max_items,
                                                   const int* block_scene)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int len = (int)runlen[s];
    int nch = (len + tile_p - 1) / tile_p;
    int at = (int)offsets[s];
    if (s == n - 1) *n_items = min(at + nch, max_items);
    if (len == 0) return;
    int b = (int)(keys[s] >> shift);
    int scene = block_scene ? block_scene[b] : 0;
    // nch chunks of equal size (within one particle), each at most tile_p: chunk c
    // holds run positions [c len / nch, (c + 1) len / nch)
    for (int c = 0; c < nch && at + c < max_items; ++c) {
        const int lo = (int)((long long)c * len / nch), hi = (int)((long long)(c + 1) * len / nch);
        items[at + c] = make_int4(b, s + lo, hi - lo, scene);
    }
}

__global__ void k_zero_int(int* v) { *v = 0; }

hipError_t launch_cell_keys(int dim, const float* x, int n, const Phys& ph, const DevGrid& g,
                            uint32_t* keys, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    dim3 grid((n + 255) / 256), blk(256);
    if (dim == 2) k_cell_keys<2><<<grid, blk, 0, st>>>(x, n, ph.inv_dx, ph.W, ph.H, ph.D, g, keys);
    else          k_cell_keys<3><<<grid, blk, 0, st>>>(x, n, ph.inv_dx, ph.W, ph.H, ph.D, g, keys);
    return hipGetLastError();
}

hipError_t launch_cell_keys_batch(int dim, const float* x, int n, const SceneDesc* scenes, const int* p_off,
                                  int nscenes, uint32_t* keys, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    dim3 grid((n + 255) / 256), blk(256);
    if (dim == 2) k_cell_keys_batch<2><<<grid, blk, 0, st>>>(x, n, scenes, p_off, nscenes, keys);
    else          k_cell_keys_batch<3><<<grid, blk, 0, st>>>(x, n, scenes, p_off, nscenes, keys);
    return hipGetLastError();
}

hipError_t launch_gather_fields(const float* src, float* dst, const uint32_t* perm, int n, int ncomp,
                                hipStream_t st)
{
    if (n == 0) return hipSuccess;
    k_gather_fields<<<(n + 255) / 256, 256, 0, st>>>(src, dst, perm, n, ncomp);
    return hipGetLastError();
}

hipError_t launch_gather_particles(int dim, const float* src, float* dst, const uint32_t* perm, int n,
                                   hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const dim3 grid((n + 255) / 256), blk(256);
    if (dim == 3) k_gather_rows<27><<<grid, blk, 0, st>>>(src, dst, perm, n);
    else          k_gather_rows<15><<<grid, blk, 0, st>>>(src, dst, perm, n);
    return hipGetLastError();
}

hipError_t launch_gather_u32(const uint32_t* src, uint32_t* dst, const uint32_t* perm, int n,
                             hipStream_t st)
{
    if (n == 0) return hipSuccess;
    k_gather_u32<<<(n + 255) / 256, 256, 0, st>>>(src, dst, perm, n);
    return hipGetLastError();
}

hipError_t launch_emit_items(const uint32_t* keys, int n, int tile_p, int shift, int4* items, int* n_items,
                             int max_items, uint32_t* tmp_runlen, uint32_t* tmp_counts, uint32_t* scan_sums,
                             const int* block_scene, hipStream_t st)
{
    k_zero_int<<<1, 1, 0, st>>>(n_items);
    if (n <= 0) return hipGetLastError();
    dim3 grid((n + 255) / 256), blk(256);
    k_item_counts<<<grid, blk, 0, st>>>(keys, n, tile_p, shift, tmp_runlen, tmp_counts);
    hipError_t e = exclusive_scan(tmp_counts, n, scan_sums, st);
    if (e != hipSuccess) return e;
    k_item_write<<<grid, blk, 0, st>>>(keys, n, tile_p, shift, tmp_runlen, tmp_counts, items, n_items, max_items,
                                      block_scene);
    return hipGetLastError();
}

}  // namespace mpm
