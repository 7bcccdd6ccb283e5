// mpm_solver_impl.hpp — the solver handle behind the C-ABI (include/mpm/mpm.h):
// device particle buffers (two, for reordering), double-buffered block-major grids,
// binning work list, the scene table of a batch, graph state and timers.  Shared by
// the C-ABI (mpm_capi.hip) and the learned surrogate (mpm_surrogate.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/mpm/mpm.h"
#include "mpm_kernels.hpp"

namespace mpm {
namespace detail {

enum GridState { GRID_CLEAN = 0, GRID_READY = 1, GRID_DIRTY = 2 };

// One set of particle arrays on the device (the solver keeps two for reordering).
struct ParticleBuffers {
    void* base = nullptr;
    float *x = nullptr, *v = nullptr, *C = nullptr, *F = nullptr, *Jp = nullptr;
    float *mass = nullptr, *vol = nullptr;
    uint32_t* pid = nullptr;
};

// Records msg as the calling thread's last error (mpm_last_error) and returns code.
int set_error(int code, const std::string& msg);

struct KernelTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0.0;
    int64_t launches = 0;
};

}  // namespace detail
}  // namespace mpm

struct mpm_solver {
    mpm_params params{};
    mpm_options opts{};
    mpm::Phys ph{};
    mpm::LaunchCfg cfg{};
    int dim = 2;
    int n = 0;
    hipStream_t stream = nullptr;

    mpm::detail::ParticleBuffers pb[2];
    int cur_pb = 0;

    // grids: a[2], vc[2], flags[2]
    int nbx = 0, nby = 0, nbz = 0, nblocks = 0;
    float4* ga[2] = {nullptr, nullptr};
    float4* gvc[2] = {nullptr, nullptr};
    float4* gf[2] = {nullptr, nullptr};
    uint32_t* gflags[2] = {nullptr, nullptr};
    int cur_grid = 0;
    int grid_state = mpm::detail::GRID_CLEAN;

    // binning
    uint32_t *keys = nullptr, *vals = nullptr, *keys_alt = nullptr, *vals_alt = nullptr;
    uint32_t* radix_scratch = nullptr;
    int key_bits = 0;
    int4* items = nullptr;
    int* n_items = nullptr;
    int max_items = 0;
    // super-block work items (mpm_options.fused 2 or 3, 3D single scene): keys and
    // items per 2x2x2 grid blocks, and each item's bound for its fixed-point scale
    bool super_items = false;
    int super_margin = 0;
    int nsx = 0, nsy = 0, nsz = 0;
    int key_bits_super = 0;          // bits of a super-block number (periodic re-binning)
    float2* item_bound = nullptr;
    float4* item_ref = nullptr;      // mean particle position per item at binning
    int4* item_shift = nullptr;      // tile offset per item (drift since binning)
    bool items_valid = false;
    bool sorted_keys_in_alt = false;
    int since_sort = 0;
    int64_t substeps_done = 0;   // substeps advanced since creation (snapshots)
    // a learned-surrogate macro step has moved the particles (mpm_surrogate_advance):
    // binning, grid and substep calls are refused from then on (MPM_E_STATE)
    bool after_macro = false;

    unsigned long long* d_bad = nullptr;

    // batch of independent scenes (nscenes > 1): scene table, block -> scene,
    // scene particle offsets; a single scene uses none of them
    int nscenes = 1;
    std::vector<mpm::SceneDesc> scenes_h;
    std::vector<mpm_params> scene_params;
    std::vector<int64_t> p_off;
    mpm::SceneDesc* scenes_d = nullptr;
    int* block_scene_d = nullptr;
    int* p_off_d = nullptr;

    // hipGraph of one binning period pair (2 * sort_every substeps)
    bool use_graph = false;
    hipGraphExec_t graph_exec = nullptr;
    int graph_grid = -1, graph_pb = -1, graph_period = 0;

    bool timing = false;
    uint32_t timing_mask = ~0u;   // kernels timed while timing is on (bit MPM_KERNEL_*)
    int timing_every = 1;           // time every k-th launch of each timed kernel
    int64_t timing_seen[MPM_KERNEL_COUNT] = {};   // launches of each kernel while timed
    mpm::detail::KernelTimer timers[MPM_KERNEL_COUNT];
    std::vector<hipEvent_t> event_pool;

    mpm::DevParticles dev_particles() const
    {
        const mpm::detail::ParticleBuffers& b = pb[cur_pb];
        mpm::DevParticles p;
        p.x = b.x; p.v = b.v; p.C = b.C; p.F = b.F; p.Jp = b.Jp;
        p.mass = b.mass; p.vol = b.vol; p.pid = b.pid; p.n = n;
        return p;
    }
    mpm::DevGrid dev_grid(int which) const
    {
        mpm::DevGrid g;
        g.a = ga[which];
        g.vc = gvc[which];
        g.f = gf[which];
        g.flags = gflags[which];
        g.nbx = nbx; g.nby = nby; g.nbz = nbz;
        return g;
    }
    hipEvent_t take_event()
    {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
};
