// mpm_capi.hip — the C-ABI of libmpm_hip.so (include/mpm/mpm.h).
//
// Host side of the reference's GPUSolver (src/gpu/GPUSolver.{h,cpp}): owns the
// device buffers and one HIP stream, uploads the initial particle set, sequences
// the substep kernels and reads state back.  No exception crosses the ABI; every
// entry point returns an MPM_* status and records a message for mpm_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpm/mpm.h"
#include "mpm_kernels.hpp"
#include "mpm_solver_impl.hpp"
#include "mpm_sort.hpp"
#include "mpm_super.hpp"

using namespace mpm;
using namespace mpm::detail;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(MPM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

}  // namespace

int mpm::detail::set_error(int code, const std::string& msg) { return fail(code, msg); }

namespace {

// Wraps one kernel launch with HIP events when timing is on.
template <typename Fn>
int timed(mpm_solver* s, int kernel, Fn&& fn)
{
    hipEvent_t a = nullptr, b = nullptr;
    const bool on = s->timing && ((s->timing_mask >> kernel) & 1u) && s->timing_seen[kernel]++ % s->timing_every == 0;
    if (on) {
        a = s->take_event();
        b = s->take_event();
        if (a) HIP_TRY(hipEventRecord(a, s->stream));
    }
    hipError_t e = fn();
    if (e != hipSuccess) return fail(MPM_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    if (on && a && b) {
        HIP_TRY(hipEventRecord(b, s->stream));
        s->timers[kernel].pending.emplace_back(a, b);
    }
    return MPM_OK;
}

int collect_timers(mpm_solver* s)
{
    for (int k = 0; k < MPM_KERNEL_COUNT; ++k) {
        KernelTimer& t = s->timers[k];
        for (auto& pr : t.pending) {
            HIP_TRY(hipEventSynchronize(pr.second));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second));
            t.total_ms += ms;
            t.launches += 1;
            s->event_pool.push_back(pr.first);
            s->event_pool.push_back(pr.second);
        }
        t.pending.clear();
    }
    return MPM_OK;
}

Phys make_phys(const mpm_params& p)
{
    Phys ph{};
    ph.dx = p.dx;
    ph.inv_dx = 1.0f / p.dx;
    ph.dt = p.DELTA_TIME;
    ph.Dinv = (p.spline_order == 3 ? 3.0f : 4.0f) * ph.inv_dx * ph.inv_dx;
    ph.gravity = p.GRAVITY;
    ph.friction = p.FRICTION;
    ph.margin = p.BORDER_MARGIN;
    ph.theta0 = 1.0f - p.THETA_C;
    ph.theta1 = 1.0f + p.THETA_S;
    ph.xi = p.XI;
    ph.mu = p.MU;
    ph.lambda = p.LAMBDA;
    ph.mass_eps = p.mass_eps;
    ph.W = p.WIDTH;
    ph.H = p.HEIGHT;
    ph.D = p.dim == 3 ? p.DEPTH : 0;
    ph.boundary = p.boundary;
    ph.scheme = p.scheme;
    return ph;
}

int validate_params(const mpm_params* p)
{
    if (!p) return fail(MPM_E_INVALID, "params is NULL");
    if (p->dim != 2 && p->dim != 3) return fail(MPM_E_INVALID, "dim must be 2 or 3");
    if (p->spline_order != 2 && p->spline_order != 3)
        return fail(MPM_E_INVALID, "spline_order must be 2 (quadratic) or 3 (cubic)");
    if (p->boundary != MPM_BOUNDARY_COULOMB_MARGIN && p->boundary != MPM_BOUNDARY_STICKY)
        return fail(MPM_E_INVALID, "unknown boundary kind");
    if (p->scheme != MPM_SCHEME_MLS && p->scheme != MPM_SCHEME_APIC_FORCE)
        return fail(MPM_E_INVALID, "unknown scheme");
    if (p->WIDTH < 1 || p->HEIGHT < 1 || (p->dim == 3 && p->DEPTH < 1))
        return fail(MPM_E_INVALID, "grid size must be positive");
    if (!(p->dx > 0.f) || !(p->DELTA_TIME > 0.f)) return fail(MPM_E_INVALID, "dx and DELTA_TIME must be > 0");
    int64_t nodes = 0;
    mpm_num_nodes(p, &nodes);
    // block-major node addresses and sort keys are 32-bit
    int bs = p->dim == 3 ? 4 : 8;
    int64_t nb = (int64_t)((p->WIDTH + bs) / bs) * ((p->HEIGHT + bs) / bs) * (p->dim == 3 ? (p->DEPTH + bs) / bs : 1);
    if (nb * 64 >= ((int64_t)1 << 31)) return fail(MPM_E_INVALID, "grid too large for 32-bit node addressing");
    return MPM_OK;
}

int alloc_particles(mpm_solver* s, ParticleBuffers& b)
{
    const int d = s->dim;
    const size_t n = (size_t)std::max(s->n, 1);
    size_t nf = (size_t)(d + d + d * d + d * d + 3) * n;
    size_t bytes = nf * sizeof(float) + n * sizeof(uint32_t) + 256;
    HIP_TRY(hipMalloc(&b.base, bytes));
    float* f = static_cast<float*>(b.base);
    b.x = f; f += d * n;
    b.v = f; f += d * n;
    b.C = f; f += d * d * n;
    b.F = f; f += d * d * n;
    b.Jp = f; f += n;
    b.mass = f; f += n;
    b.vol = f; f += n;
    b.pid = reinterpret_cast<uint32_t*>(f);
    return MPM_OK;
}

int reset_grid(mpm_solver* s, int which)
{
    // every scene's blocks (a batch numbers them one after another)
    return timed(s, MPM_KERNEL_RESET,
                 [&] { return launch_reset(s->cfg, s->dev_grid(which), s->nblocks, s->stream); });
}

int reset_all_grids(mpm_solver* s)
{
    int r = reset_grid(s, 0);
    if (r) return r;
    r = reset_grid(s, 1);
    if (r) return r;
    s->grid_state = GRID_CLEAN;
    return MPM_OK;
}

SuperWork super_work(const mpm_solver* s);

int do_sort(mpm_solver* s)
{
    const int n = s->n;
    // Super-block items: a binning of a state the items no longer describe (the first,
    // or after mpm_set_state or stage calls) sorts by the full cell key; the periodic
    // re-binning by the super-block number alone, a stable sort that keeps the order
    // the last full sort gave inside each super-block (fewer radix passes).
    const bool cells = !s->super_items || !s->items_valid;
    auto body = [&]() -> hipError_t {
        DevParticles p = s->dev_particles();
        hipError_t e = s->scenes_d      ? launch_cell_keys_batch(s->dim, p.x, n, s->scenes_d, s->p_off_d, s->nscenes,
                                                                 s->keys, s->stream)
                       : s->super_items ? launch_cell_keys_super(p.x, n, s->ph, s->nsx, s->nsy, cells, s->keys,
                                                                 s->stream)
                                        : launch_cell_keys(s->dim, p.x, n, s->ph, s->dev_grid(0), s->keys, s->stream);
        if (e != hipSuccess) return e;
        bool in_alt = false;
        e = radix_sort_pairs(s->keys, s->vals, s->keys_alt, s->vals_alt, n, cells ? s->key_bits : s->key_bits_super,
                             true, s->radix_scratch, s->stream, &in_alt);
        if (e != hipSuccess) return e;
        s->sorted_keys_in_alt = in_alt;
        const uint32_t* perm = in_alt ? s->vals_alt : s->vals;
        const uint32_t* skeys = in_alt ? s->keys_alt : s->keys;
        // x, v, C, F, J_P, mass, volume and original index sit one after another
        // from x in each particle buffer (alloc_particles)
        e = launch_gather_particles(s->dim, s->pb[s->cur_pb].x, s->pb[1 - s->cur_pb].x, perm, n, s->stream);
        if (e != hipSuccess) return e;
        s->cur_pb = 1 - s->cur_pb;
        // the pair of key/value buffers the sort did not end in is free now
        uint32_t* tmp_a = in_alt ? s->keys : s->keys_alt;
        uint32_t* tmp_b = in_alt ? s->vals : s->vals_alt;
        e = launch_emit_items(skeys, n, s->opts.tile_particles, s->super_items ? (cells ? SB_KEY_SHIFT : 0) : 6, s->items,
                              s->n_items, s->max_items, tmp_a, tmp_b, s->radix_scratch, s->block_scene_d, s->stream);
        if (e != hipSuccess || !s->super_items) return e;
        // the new items' fixed-point bounds from the reordered state
        return launch_item_bounds(s->cfg.order, s->dev_particles(), super_work(s), s->ph, s->stream);
    };
    int r = timed(s, MPM_KERNEL_SORT, body);
    if (r) return r;
    s->items_valid = true;
    s->since_sort = 0;
    return MPM_OK;
}

SuperWork super_work(const mpm_solver* s)
{
    SuperWork w;
    w.items = s->items;
    w.n_items = s->n_items;
    w.max_items = s->max_items;
    w.bound = s->item_bound;
    w.ref = s->item_ref;
    w.shift = s->item_shift;
    w.nsx = s->nsx;
    w.nsy = s->nsy;
    return w;
}

TileWork tile_work(const mpm_solver* s)
{
    TileWork w;
    w.items = s->items;
    w.n_items = s->n_items;
    w.max_items = s->max_items;
    w.store_vc = 1;
    return w;
}

// One fused substep: binning when due, grid update + clear, fused G2P/P2G.  last:
// the substep ends the caller's step (it stores v and C, which the other substeps
// leave to the next gather; with super-block items also the substep before a
// binning, whose item bounds read them).
int fused_substep(mpm_solver* s, bool last)
{
    int r;
    if (s->opts.sort_every > 0 && s->since_sort >= s->opts.sort_every && (r = do_sort(s))) return r;
    DevGrid cur = s->dev_grid(s->cur_grid), old = s->dev_grid(1 - s->cur_grid);
    r = timed(s, MPM_KERNEL_GRID_UPDATE, [&] {
        return s->scenes_d ? launch_grid_update_clear_batch(s->cfg, cur, old, s->scenes_d, s->block_scene_d,
                                                            s->nblocks, s->stream)
                           : launch_grid_update_clear(s->cfg, cur, old, s->ph, s->stream);
    });
    if (r) return r;
    DevParticles p = s->dev_particles();
    if (s->super_items) {
        const SuperWork w = super_work(s);
        // v and C are read after this substep by the next binning's item bounds
        // (due when the next substep starts) or by the caller
        const bool store_vc = last || (s->opts.sort_every > 0 && s->since_sort + 1 >= s->opts.sort_every);
        r = timed(s, MPM_KERNEL_G2P2G, [&] {
            return launch_super(s->cfg, s->super_margin, true, store_vc, p, cur, old, w, s->ph, s->stream);
        });
    } else {
        TileWork w = tile_work(s);
        w.store_vc = last ? 1 : 0;
        r = timed(s, MPM_KERNEL_G2P2G,
                  [&] { return launch_g2p2g(s->cfg, p, cur, old, w, s->ph, s->scenes_d, s->stream); });
    }
    if (r) return r;
    s->cur_grid = 1 - s->cur_grid;
    s->since_sort += 1;
    return MPM_OK;
}

void drop_graph(mpm_solver* s)
{
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    s->graph_exec = nullptr;
    s->graph_grid = s->graph_pb = -1;
}

// A graph replays `period` = 2 * sort_every substeps recorded from a point where the
// next substep bins; after them the grid parity, the particle-buffer parity and the
// binning counter are back where they started, so the same graph (same pointers)
// applies again.  Timing events are not recorded inside it.
int run_graph_period(mpm_solver* s, int period)
{
    if (s->graph_exec && (s->graph_grid != s->cur_grid || s->graph_pb != s->cur_pb || s->graph_period != period))
        drop_graph(s);
    if (!s->graph_exec) {
        bool timing = s->timing;
        s->timing = false;
        HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        const int grid0 = s->cur_grid, pb0 = s->cur_pb;
        int r = MPM_OK;
        for (int k = 0; k < period && r == MPM_OK; ++k) r = fused_substep(s, k == period - 1);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s->stream, &graph);
        s->timing = timing;
        if (r) { if (graph) (void)hipGraphDestroy(graph); return r; }
        if (e != hipSuccess) return fail(MPM_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&s->graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) { s->graph_exec = nullptr; return fail(MPM_E_HIP, "hipGraphInstantiate failed"); }
        if (s->cur_grid != grid0 || s->cur_pb != pb0) { drop_graph(s); return fail(MPM_E_STATE, "graph period not closed"); }
        s->graph_grid = grid0;
        s->graph_pb = pb0;
        s->graph_period = period;
    }
    HIP_TRY(hipGraphLaunch(s->graph_exec, s->stream));
    return MPM_OK;
}

int step_fused(mpm_solver* s, int nsub)
{
    int r;
    if (!s->items_valid && (r = do_sort(s))) return r;
    if (s->grid_state == GRID_DIRTY && (r = reset_all_grids(s))) return r;
    if (s->grid_state != GRID_READY) {
        DevParticles p = s->dev_particles();
        DevGrid g = s->dev_grid(s->cur_grid);
        if (s->super_items) {
            // bounds of the state as it is now (stage calls may have changed it)
            const SuperWork w = super_work(s);
            r = timed(s, MPM_KERNEL_P2G, [&] {
                hipError_t e = launch_item_bounds(s->cfg.order, p, w, s->ph, s->stream);
                return e != hipSuccess ? e
                                       : launch_super(s->cfg, s->super_margin, false, true, p, g, g, w, s->ph, s->stream);
            });
        } else {
            TileWork w = tile_work(s);
            r = timed(s, MPM_KERNEL_P2G,
                      [&] { return launch_p2g_tiled(s->cfg, p, g, w, s->ph, s->scenes_d, s->stream); });
        }
        if (r) return r;
        s->grid_state = GRID_READY;
    }
    const int period = s->opts.sort_every > 0 ? 2 * s->opts.sort_every : 2;
    int k = 0;
    while (k < nsub) {
        const bool aligned = s->opts.sort_every > 0 ? s->since_sort >= s->opts.sort_every : true;
        if (s->use_graph && aligned && nsub - k >= period) {
            if ((r = run_graph_period(s, period))) return r;
            k += period;
            continue;
        }
        if ((r = fused_substep(s, k == nsub - 1))) return r;
        ++k;
    }
    return MPM_OK;
}

int step_staged(mpm_solver* s, int nsub)
{
    int r;
    if (s->grid_state != GRID_CLEAN && (r = reset_all_grids(s))) return r;
    for (int k = 0; k < nsub; ++k) {
        if (s->opts.sort_every > 0 && s->since_sort >= s->opts.sort_every && (r = do_sort(s))) return r;
        DevParticles p = s->dev_particles();
        DevGrid g = s->dev_grid(s->cur_grid);
        if ((r = timed(s, MPM_KERNEL_P2G, [&] { return launch_p2g(s->cfg, p, g, s->ph, s->stream); }))) return r;
        if ((r = timed(s, MPM_KERNEL_GRID_UPDATE,
                       [&] { return launch_grid_update(s->cfg, g, s->ph, s->stream); }))) return r;
        if ((r = timed(s, MPM_KERNEL_G2P, [&] { return launch_g2p(s->cfg, p, g, s->ph, s->stream); }))) return r;
        if ((r = reset_grid(s, s->cur_grid))) return r;
        s->since_sort += 1;
    }
    s->grid_state = GRID_CLEAN;
    return MPM_OK;
}

// Stage calls see the reference's state between Steps: a zero grid.
int prepare_stage_call(mpm_solver* s)
{
    if (s->grid_state == GRID_READY) return reset_all_grids(s);
    return MPM_OK;
}

int copy_to_host(mpm_solver* s, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0) return MPM_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream));
    return MPM_OK;
}

}  // namespace

extern "C" {

int mpm_params_derive_lame(mpm_params* p)
{
    if (!p) return fail(MPM_E_INVALID, "params is NULL");
    // Parameters.h:82-84, in float as there
    p->MU = p->E / (2.f * (1.f + p->NU));
    p->LAMBDA = (p->E * p->NU) / ((1.f + p->NU) * (1.f - (2.f * p->NU)));
    return MPM_OK;
}

int mpm_params_reference(int dim, mpm_params* out)
{
    if (!out) return fail(MPM_E_INVALID, "out is NULL");
    if (dim != 2 && dim != 3) return fail(MPM_E_INVALID, "dim must be 2 or 3");
    mpm_params p{};
    p.dim = dim;
    p.WIDTH = 200;           // Parameters.h:31-33
    p.HEIGHT = 100;
    p.DEPTH = 100;           // Parameters.h:36
    p.BORDER_MARGIN = 2.f;
    p.spline_order = 3;      // cubic B-spline, shaders/2d/common.glsl:43-69
    p.boundary = MPM_BOUNDARY_COULOMB_MARGIN;
    p.dx = 1.f;              // grid units
    p.GRAVITY = -9.81f;      // Parameters.h:58-60
    p.FRICTION = 0.3f;
    p.DELTA_TIME = 1e-3f;    // Parameters.h:68
    p.THETA_C = 2.0e-2f;     // Parameters.h:74-79
    p.THETA_S = 6.0e-3f;
    p.XI = 10.f;
    p.NU = 2.0e-1f;
    p.E = 1.4e5f;
    p.RHO = 400.f;
    p.mass_eps = 1e-6f;      // shaders/2d/grid_update.glsl:90
    mpm_params_derive_lame(&p);
    *out = p;
    return MPM_OK;
}

int mpm_options_default(mpm_options* out)
{
    if (!out) return fail(MPM_E_INVALID, "out is NULL");
    out->sort_every = 16;
    out->fused = 1;
    out->tile_particles = 512;
    return MPM_OK;
}

int mpm_num_nodes(const mpm_params* p, int64_t* nodes)
{
    if (!p || !nodes) return fail(MPM_E_INVALID, "NULL argument");
    int64_t g = (int64_t)(p->WIDTH + 1) * (p->HEIGHT + 1);
    if (p->dim == 3) g *= (p->DEPTH + 1);
    *nodes = g;
    return MPM_OK;
}

}  // extern "C"

namespace {

int create_impl(int nscenes, const mpm_params* params, const mpm_particles* init, const mpm_options* opts,
                mpm_solver** out)
{
    if (!out) return fail(MPM_E_INVALID, "out is NULL");
    *out = nullptr;
    if (nscenes < 1) return fail(MPM_E_INVALID, "nscenes must be >= 1");
    if (!params || !init) return fail(MPM_E_INVALID, "params/init is NULL");
    int r;
    int64_t ntot = 0;
    for (int sc = 0; sc < nscenes; ++sc) {
        if ((r = validate_params(&params[sc]))) return r;
        const mpm_particles& in = init[sc];
        if (in.n < 0) return fail(MPM_E_INVALID, "particle count out of range");
        if (in.n > 0 && (!in.x || !in.mass || !in.vol)) return fail(MPM_E_INVALID, "x, mass and vol are required");
        ntot += in.n;
        if (sc > 0 && (params[sc].dim != params[0].dim || params[sc].spline_order != params[0].spline_order ||
                       params[sc].boundary != params[0].boundary || params[sc].scheme != params[0].scheme))
            return fail(MPM_E_INVALID, "scenes of a batch must share dim, spline order, boundary and scheme");
    }
    if (nscenes > 1 && params[0].scheme != MPM_SCHEME_MLS)
        return fail(MPM_E_INVALID, "a batch runs the MLS scheme");
    if (ntot > (int64_t)INT32_MAX / 16) return fail(MPM_E_INVALID, "particle count out of range");
    if (opts && (opts->fused < 0 || opts->fused > 3)) return fail(MPM_E_INVALID, "fused must be 0, 1, 2 or 3");
    if (opts && opts->fused >= 2 && (params[0].dim != 3 || nscenes != 1))
        return fail(MPM_E_INVALID, "fused 2 and 3 (super-block work items) need one 3D scene");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1)
        return fail(MPM_E_HIP, "no HIP device available");

    mpm_solver* s = new mpm_solver();
    s->params = params[0];
    if (opts) s->opts = *opts; else mpm_options_default(&s->opts);
    // fused 2 / 3: super-block work items with a drift margin of 0 / 1 cells, for one
    // 3D scene; the per-block tiled kernels hold at most two particles per thread of a
    // 256-thread group, the super-block kernel streams up to SB_MAX_PARTICLES
    s->super_items = s->opts.fused >= 2;
    s->super_margin = s->opts.fused == 3 ? 1 : 0;
    const int tile_max = s->super_items ? SB_MAX_PARTICLES : 512;
    if (s->opts.tile_particles <= 0 || s->opts.tile_particles > tile_max) s->opts.tile_particles = tile_max;
    if (s->opts.sort_every < 0) s->opts.sort_every = 0;
    s->dim = params[0].dim;
    s->n = (int)ntot;
    s->nscenes = nscenes;
    s->ph = make_phys(params[0]);
    s->cfg.dim = params[0].dim;
    s->cfg.order = params[0].spline_order;
    s->cfg.boundary = params[0].boundary;

    auto cleanup_fail = [&](int code) {
        mpm_destroy(s);
        return code;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup_fail(fail(MPM_E_HIP, "hipStreamCreate failed"));

    // grid blocks of every scene, one after another
    const int bs = s->dim == 3 ? 4 : 8;
    std::vector<int> block_scene;
    int64_t nb_total = 0;
    s->p_off.assign(1, 0);
    for (int sc = 0; sc < nscenes; ++sc) {
        const mpm_params& P = params[sc];
        SceneDesc sd{};
        sd.ph = make_phys(P);
        sd.nbx = (P.WIDTH + 1 + bs - 1) / bs;
        sd.nby = (P.HEIGHT + 1 + bs - 1) / bs;
        sd.nbz = s->dim == 3 ? (P.DEPTH + 1 + bs - 1) / bs : 1;
        sd.block_off = (int)nb_total;
        int64_t nb = (int64_t)sd.nbx * sd.nby * sd.nbz;
        nb_total += nb;
        if (nb_total * 64 >= ((int64_t)1 << 31)) return cleanup_fail(fail(MPM_E_INVALID, "grids too large for 32-bit node addressing"));
        if (nscenes > 1) block_scene.insert(block_scene.end(), (size_t)nb, sc);
        s->scenes_h.push_back(sd);
        s->scene_params.push_back(P);
        s->p_off.push_back(s->p_off.back() + init[sc].n);
    }
    s->nbx = s->scenes_h[0].nbx;
    s->nby = s->scenes_h[0].nby;
    s->nbz = s->scenes_h[0].nbz;
    s->nblocks = (int)nb_total;
    if (nscenes > 1) {
        std::vector<int> poff32(s->p_off.begin(), s->p_off.end());
        if (hipMalloc(&s->scenes_d, sizeof(SceneDesc) * nscenes) != hipSuccess ||
            hipMalloc(&s->block_scene_d, sizeof(int) * block_scene.size()) != hipSuccess ||
            hipMalloc(&s->p_off_d, sizeof(int) * poff32.size()) != hipSuccess)
            return cleanup_fail(fail(MPM_E_HIP, "scene table allocation failed"));
        if (hipMemcpyAsync(s->scenes_d, s->scenes_h.data(), sizeof(SceneDesc) * nscenes, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipMemcpyAsync(s->block_scene_d, block_scene.data(), sizeof(int) * block_scene.size(), hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipMemcpyAsync(s->p_off_d, poff32.data(), sizeof(int) * poff32.size(), hipMemcpyHostToDevice, s->stream) != hipSuccess)
            return cleanup_fail(fail(MPM_E_HIP, "scene table upload failed"));
    }

    const size_t gbytes = (size_t)s->nblocks * 64 * sizeof(float4);
    const bool coul = params[0].boundary == MPM_BOUNDARY_COULOMB_MARGIN;
    for (int g = 0; g < 2; ++g) {
        if (hipMalloc(&s->ga[g], gbytes) != hipSuccess) return cleanup_fail(fail(MPM_E_HIP, "grid allocation failed"));
        if (hipMemsetAsync(s->ga[g], 0, gbytes, s->stream) != hipSuccess) return cleanup_fail(fail(MPM_E_HIP, "memset"));
        if (coul) {
            if (hipMalloc(&s->gvc[g], gbytes) != hipSuccess) return cleanup_fail(fail(MPM_E_HIP, "grid allocation failed"));
            if (hipMemsetAsync(s->gvc[g], 0, gbytes, s->stream) != hipSuccess) return cleanup_fail(fail(MPM_E_HIP, "memset"));
        }
        if (params[0].scheme == MPM_SCHEME_APIC_FORCE) {
            if (hipMalloc(&s->gf[g], gbytes) != hipSuccess) return cleanup_fail(fail(MPM_E_HIP, "grid allocation failed"));
            if (hipMemsetAsync(s->gf[g], 0, gbytes, s->stream) != hipSuccess) return cleanup_fail(fail(MPM_E_HIP, "memset"));
        }
        if (hipMalloc(&s->gflags[g], (size_t)s->nblocks * sizeof(uint32_t)) != hipSuccess)
            return cleanup_fail(fail(MPM_E_HIP, "flag allocation failed"));
        if (hipMemsetAsync(s->gflags[g], 0, (size_t)s->nblocks * sizeof(uint32_t), s->stream) != hipSuccess)
            return cleanup_fail(fail(MPM_E_HIP, "memset"));
    }

    for (int b = 0; b < 2; ++b)
        if ((r = alloc_particles(s, s->pb[b]))) return cleanup_fail(r);

    // binning buffers
    const size_t nn = (size_t)std::max(s->n, 1);
    uint64_t maxkey = (uint64_t)s->nblocks * 64;
    int64_t nsuper = 0;
    if (s->super_items) {
        s->nsx = super_count(s->params.WIDTH);
        s->nsy = super_count(s->params.HEIGHT);
        s->nsz = super_count(s->params.DEPTH);
        nsuper = (int64_t)s->nsx * s->nsy * s->nsz;
        maxkey = (uint64_t)nsuper << SB_KEY_SHIFT;
        if (maxkey > ((uint64_t)1 << 32)) return cleanup_fail(fail(MPM_E_INVALID, "grid too large for 32-bit super-block keys"));
        s->key_bits_super = 1;
        while (((uint64_t)1 << s->key_bits_super) < (uint64_t)nsuper) ++s->key_bits_super;
    }
    s->key_bits = 1;
    while (s->key_bits < 32 && ((uint64_t)1 << s->key_bits) < maxkey) ++s->key_bits;
    if (hipMalloc(&s->keys, nn * 4) != hipSuccess || hipMalloc(&s->vals, nn * 4) != hipSuccess ||
        hipMalloc(&s->keys_alt, nn * 4) != hipSuccess || hipMalloc(&s->vals_alt, nn * 4) != hipSuccess ||
        hipMalloc(&s->radix_scratch, radix_scratch_words(s->n) * 4) != hipSuccess)
        return cleanup_fail(fail(MPM_E_HIP, "binning allocation failed"));
    s->max_items = (int)std::min<int64_t>(s->n, s->super_items ? nsuper : s->nblocks) + s->n / s->opts.tile_particles + 1;
    if (s->super_items && (hipMalloc(&s->item_bound, (size_t)s->max_items * sizeof(float2)) != hipSuccess ||
                           hipMemsetAsync(s->item_bound, 0, (size_t)s->max_items * sizeof(float2), s->stream) != hipSuccess ||
                           hipMalloc(&s->item_ref, (size_t)s->max_items * sizeof(float4)) != hipSuccess ||
                           hipMemsetAsync(s->item_ref, 0, (size_t)s->max_items * sizeof(float4), s->stream) != hipSuccess ||
                           hipMalloc(&s->item_shift, (size_t)s->max_items * sizeof(int4)) != hipSuccess ||
                           hipMemsetAsync(s->item_shift, 0, (size_t)s->max_items * sizeof(int4), s->stream) != hipSuccess))
        return cleanup_fail(fail(MPM_E_HIP, "work-list allocation failed"));
    if (hipMalloc(&s->items, (size_t)s->max_items * sizeof(int4)) != hipSuccess ||
        hipMalloc(&s->n_items, sizeof(int)) != hipSuccess || hipMalloc(&s->d_bad, sizeof(unsigned long long)) != hipSuccess)
        return cleanup_fail(fail(MPM_E_HIP, "work-list allocation failed"));
    if (hipMemsetAsync(s->n_items, 0, sizeof(int), s->stream) != hipSuccess)
        return cleanup_fail(fail(MPM_E_HIP, "memset"));

    // initial state (GPUSolver.cpp:100-159 defaults for omitted fields), scenes
    // concatenated in order
    const int d = s->dim, n = s->n;
    if (n > 0) {
        std::vector<float> host((size_t)(d + d + d * d + d * d + 3) * n);
        float* x = host.data();
        float* v = x + (size_t)d * n;
        float* C = v + (size_t)d * n;
        float* F = C + (size_t)d * d * n;
        float* Jp = F + (size_t)d * d * n;
        float* m = Jp + n;
        float* vol = m + n;
        for (int sc = 0; sc < nscenes; ++sc) {
            const mpm_particles& in = init[sc];
            const int64_t ns = in.n, o = s->p_off[sc];
            for (int c = 0; c < d; ++c) {
                std::memcpy(x + (size_t)c * n + o, in.x + (size_t)c * ns, sizeof(float) * ns);
                if (in.v) std::memcpy(v + (size_t)c * n + o, in.v + (size_t)c * ns, sizeof(float) * ns);
            }
            for (int e = 0; e < d * d; ++e) {
                if (in.C) std::memcpy(C + (size_t)e * n + o, in.C + (size_t)e * ns, sizeof(float) * ns);
                if (in.F) std::memcpy(F + (size_t)e * n + o, in.F + (size_t)e * ns, sizeof(float) * ns);
                else std::fill(F + (size_t)e * n + o, F + (size_t)e * n + o + ns, (e % (d + 1)) == 0 ? 1.f : 0.f);
            }
            if (in.Jp) std::memcpy(Jp + o, in.Jp, sizeof(float) * ns);
            else std::fill(Jp + o, Jp + o + ns, 1.f);
            std::memcpy(m + o, in.mass, sizeof(float) * ns);
            std::memcpy(vol + o, in.vol, sizeof(float) * ns);
        }
        std::vector<uint32_t> pid(n);
        for (int i = 0; i < n; ++i) pid[i] = (uint32_t)i;
        ParticleBuffers& b = s->pb[0];
        if (hipMemcpyAsync(b.x, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipMemcpyAsync(b.pid, pid.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s->stream) != hipSuccess)
            return cleanup_fail(fail(MPM_E_HIP, "upload failed"));
        if (hipStreamSynchronize(s->stream) != hipSuccess) return cleanup_fail(fail(MPM_E_HIP, "upload sync failed"));
    } else if (hipStreamSynchronize(s->stream) != hipSuccess) {
        return cleanup_fail(fail(MPM_E_HIP, "sync failed"));
    }
    *out = s;
    return MPM_OK;
}

int reject_batch(mpm_solver* s)
{
    if (s->nscenes > 1) return fail(MPM_E_STATE, "not available on a batch of scenes");
    return MPM_OK;
}

// Binning, grid and substep calls on particles a learned-surrogate macro step has
// moved are not enabled: two runs that took plain substeps after macro steps
// ended in illegal memory accesses whose cause is not established.  The state
// stays readable; a solver created from it (mpm_get_state, mpm_create) steps.
int refuse_after_macro(mpm_solver* s)
{
    if (!s->after_macro) return MPM_OK;
    return fail(MPM_E_STATE,
                "plain substeps, binning and stage calls are not enabled on a solver after a surrogate macro "
                "step; read its state out (mpm_get_state) and create a new solver from it");
}

}  // namespace

extern "C" {

int mpm_create(const mpm_params* params, const mpm_particles* init, const mpm_options* opts,
               mpm_solver** out)
{
    if (!params || !init) {
        if (out) *out = nullptr;
        int r = validate_params(params);
        if (r) return r;
        return fail(MPM_E_INVALID, "init is NULL");
    }
    return create_impl(1, params, init, opts, out);
}

int mpm_batch_create(int nscenes, const mpm_params* params, const mpm_particles* init,
                     const mpm_options* opts, mpm_solver** out)
{
    return create_impl(nscenes, params, init, opts, out);
}

int mpm_num_scenes(const mpm_solver* s, int* nscenes)
{
    if (!s || !nscenes) return fail(MPM_E_INVALID, "NULL argument");
    *nscenes = s->nscenes;
    return MPM_OK;
}

int mpm_scene_range(const mpm_solver* s, int scene, int64_t* first, int64_t* count)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    if (scene < 0 || scene >= s->nscenes) return fail(MPM_E_INVALID, "scene out of range");
    if (first) *first = s->p_off[scene];
    if (count) *count = s->p_off[scene + 1] - s->p_off[scene];
    return MPM_OK;
}

int mpm_set_graph(mpm_solver* s, int enabled)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    s->use_graph = enabled != 0;
    if (!s->use_graph) drop_graph(s);
    return MPM_OK;
}

int mpm_autotune(int nscenes, const mpm_params* params, const mpm_particles* init, int trial_substeps,
                 int use_graph, mpm_options* best, double* best_ms_per_substep)
{
    if (!best) return fail(MPM_E_INVALID, "best is NULL");
    if (trial_substeps < 1) return fail(MPM_E_INVALID, "trial_substeps must be >= 1");
    static const int tiles[] = {128, 256, 512};
    static const int periods[] = {4, 8, 16, 32, 64};
    // candidates (fused, tile_particles, sort_every): the per-block work items, and
    // for one 3D scene the super-block items with either drift margin
    std::vector<mpm_options> cands;
    for (int tp : tiles)
        for (int se : periods) cands.push_back(mpm_options{se, 1, tp});
    if (nscenes == 1 && params && params[0].dim == 3)
        for (int f : {2, 3})
            for (int tp : {SB_MAX_PARTICLES / 4, SB_MAX_PARTICLES / 2, SB_MAX_PARTICLES})
                for (int se : {8, 16, 32, 64, 128}) cands.push_back(mpm_options{se, f, tp});
    double best_ms = -1.0;
    mpm_options chosen{};
    for (const mpm_options& o : cands) {
        {
            mpm_solver* s = nullptr;
            int r = create_impl(nscenes, params, init, &o, &s);
            if (r) return r;
            // whole binning periods, so that every candidate pays its binning share
            const int period = 2 * o.sort_every;
            const int trial = ((trial_substeps + period - 1) / period) * period;
            hipEvent_t a = nullptr, b = nullptr;
            float ms = 0.f;
            if (use_graph) s->use_graph = true;
            r = mpm_step(s, period);  // warm-up: first binning, first P2G, graph capture
            if (!r && (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess))
                r = fail(MPM_E_HIP, "hipEventCreate failed");
            if (!r && hipEventRecord(a, s->stream) != hipSuccess) r = fail(MPM_E_HIP, "hipEventRecord failed");
            if (!r) r = mpm_step(s, trial);
            if (!r && (hipEventRecord(b, s->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                       hipEventElapsedTime(&ms, a, b) != hipSuccess))
                r = fail(MPM_E_HIP, "trial timing failed");
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
            mpm_destroy(s);
            if (r) return r;
            const double per = (double)ms / trial;
            if (best_ms < 0.0 || per < best_ms) {
                best_ms = per;
                chosen = o;
            }
        }
    }
    *best = chosen;
    if (best_ms_per_substep) *best_ms_per_substep = best_ms;
    return MPM_OK;
}

int mpm_destroy(mpm_solver* s)
{
    if (!s) return MPM_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& t : s->timers)
        for (auto& pr : t.pending) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    for (int g = 0; g < 2; ++g) {
        (void)hipFree(s->ga[g]);
        (void)hipFree(s->gvc[g]);
        (void)hipFree(s->gf[g]);
        (void)hipFree(s->gflags[g]);
        (void)hipFree(s->pb[g].base);
    }
    (void)hipFree(s->keys); (void)hipFree(s->vals); (void)hipFree(s->keys_alt); (void)hipFree(s->vals_alt);
    (void)hipFree(s->radix_scratch); (void)hipFree(s->items); (void)hipFree(s->n_items); (void)hipFree(s->d_bad);
    (void)hipFree(s->item_bound);
    (void)hipFree(s->item_ref);
    (void)hipFree(s->item_shift);
    (void)hipFree(s->scenes_d); (void)hipFree(s->block_scene_d); (void)hipFree(s->p_off_d);
    drop_graph(s);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return MPM_OK;
}

int mpm_step(mpm_solver* s, int n_substeps)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    if (n_substeps < 0) return fail(MPM_E_INVALID, "n_substeps < 0");
    if (n_substeps > 0) {
        const int r0 = refuse_after_macro(s);
        if (r0) return r0;
    }
    if (s->n == 0 || n_substeps == 0) return MPM_OK;
    // the fused pipeline implements the MLS scheme; the CPU path's scheme runs the
    // four stages
    const bool fused = s->opts.fused && s->params.scheme == MPM_SCHEME_MLS;
    if (!fused && s->nscenes > 1) return fail(MPM_E_STATE, "a batch of scenes runs the fused pipeline only");
    const int r = fused ? step_fused(s, n_substeps) : step_staged(s, n_substeps);
    if (r == MPM_OK) s->substeps_done += n_substeps;
    return r;
}

int mpm_p2g(mpm_solver* s)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    int r = reject_batch(s);
    if (r) return r;
    if ((r = refuse_after_macro(s))) return r;
    r = prepare_stage_call(s);
    if (r) return r;
    DevParticles p = s->dev_particles();
    DevGrid g = s->dev_grid(s->cur_grid);
    r = timed(s, MPM_KERNEL_P2G, [&] { return launch_p2g(s->cfg, p, g, s->ph, s->stream); });
    s->grid_state = GRID_DIRTY;
    return r;
}

int mpm_grid_update(mpm_solver* s)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    int r = reject_batch(s);
    if (r) return r;
    if ((r = refuse_after_macro(s))) return r;
    r = prepare_stage_call(s);
    if (r) return r;
    DevGrid g = s->dev_grid(s->cur_grid);
    r = timed(s, MPM_KERNEL_GRID_UPDATE, [&] { return launch_grid_update(s->cfg, g, s->ph, s->stream); });
    s->grid_state = GRID_DIRTY;
    return r;
}

int mpm_g2p(mpm_solver* s)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    int r = reject_batch(s);
    if (r) return r;
    if ((r = refuse_after_macro(s))) return r;
    r = prepare_stage_call(s);
    if (r) return r;
    DevParticles p = s->dev_particles();
    DevGrid g = s->dev_grid(s->cur_grid);
    r = timed(s, MPM_KERNEL_G2P, [&] { return launch_g2p(s->cfg, p, g, s->ph, s->stream); });
    s->grid_state = GRID_DIRTY;
    return r;
}

int mpm_reset_grid(mpm_solver* s)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    const int r = refuse_after_macro(s);
    if (r) return r;
    return reset_all_grids(s);
}

int mpm_sort(mpm_solver* s)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    const int r = refuse_after_macro(s);
    if (r) return r;
    return do_sort(s);
}

int mpm_get_sort(mpm_solver* s, uint32_t* keys, uint32_t* perm)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    if (!s->items_valid) return fail(MPM_E_STATE, "no sort has run");
    const uint32_t* sk = s->sorted_keys_in_alt ? s->keys_alt : s->keys;
    int r;
    if (keys && (r = copy_to_host(s, keys, sk, (size_t)s->n * 4))) return r;
    if (perm && (r = copy_to_host(s, perm, s->pb[s->cur_pb].pid, (size_t)s->n * 4))) return r;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MPM_OK;
}

int mpm_sync(mpm_solver* s)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MPM_OK;
}

int mpm_get_state(mpm_solver* s, float* x, float* v, float* C, float* F, float* Jp)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    const int n = s->n, d = s->dim;
    if (n == 0) return MPM_OK;
    const ParticleBuffers& b = s->pb[s->cur_pb];
    std::vector<uint32_t> pid(n);
    int r;
    if ((r = copy_to_host(s, pid.data(), b.pid, (size_t)n * 4))) return r;
    struct { float* out; const float* src; int comps; } f[] = {
        {x, b.x, d}, {v, b.v, d}, {C, b.C, d * d}, {F, b.F, d * d}, {Jp, b.Jp, 1}};
    std::vector<std::vector<float>> tmp(5);
    for (int i = 0; i < 5; ++i) {
        if (!f[i].out) continue;
        tmp[i].resize((size_t)f[i].comps * n);
        if ((r = copy_to_host(s, tmp[i].data(), f[i].src, tmp[i].size() * sizeof(float)))) return r;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int i = 0; i < 5; ++i) {
        if (!f[i].out) continue;
        for (int c = 0; c < f[i].comps; ++c) {
            const float* src = tmp[i].data() + (size_t)c * n;
            float* dst = f[i].out + (size_t)c * n;
            for (int sl = 0; sl < n; ++sl) dst[pid[sl]] = src[sl];
        }
    }
    return MPM_OK;
}

int mpm_set_state(mpm_solver* s, const float* x, const float* v, const float* C, const float* F,
                  const float* Jp)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    int r = refuse_after_macro(s);
    if (r) return r;
    const int n = s->n, d = s->dim;
    if (n == 0) return MPM_OK;
    const ParticleBuffers& b = s->pb[s->cur_pb];
    std::vector<uint32_t> pid(n);
    if ((r = copy_to_host(s, pid.data(), b.pid, (size_t)n * 4))) return r;
    HIP_TRY(hipStreamSynchronize(s->stream));
    struct { const float* in; float* dst; int comps; } f[] = {
        {x, b.x, d}, {v, b.v, d}, {C, b.C, d * d}, {F, b.F, d * d}, {Jp, b.Jp, 1}};
    std::vector<float> tmp;
    for (auto& fi : f) {
        if (!fi.in) continue;
        tmp.resize((size_t)fi.comps * n);
        for (int c = 0; c < fi.comps; ++c)
            for (int sl = 0; sl < n; ++sl) tmp[(size_t)c * n + sl] = fi.in[(size_t)c * n + pid[sl]];
        HIP_TRY(hipMemcpyAsync(fi.dst, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    // the grid no longer matches the particles
    if ((r = reset_all_grids(s))) return r;
    if (x) s->items_valid = false;
    return MPM_OK;
}

int mpm_get_positions(mpm_solver* s, float* xyz)
{
    if (!s || !xyz) return fail(MPM_E_INVALID, "NULL argument");
    const int n = s->n, d = s->dim;
    std::vector<float> x((size_t)d * n);
    int r = mpm_get_state(s, x.data(), nullptr, nullptr, nullptr, nullptr);
    if (r) return r;
    for (int p = 0; p < n; ++p)
        for (int c = 0; c < d; ++c) xyz[(size_t)p * d + c] = x[(size_t)c * n + p];
    return MPM_OK;
}

int mpm_get_grid(mpm_solver* s, float* mass, float* mom, float* vcol, float* vfric)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    if (reject_batch(s)) return MPM_E_STATE;
    const mpm_params& P = s->params;
    const int d = s->dim;
    const size_t nodes_alloc = (size_t)s->nblocks * 64;
    std::vector<float4> a(nodes_alloc), c;
    int r;
    if ((r = copy_to_host(s, a.data(), s->ga[s->cur_grid], nodes_alloc * sizeof(float4)))) return r;
    const bool coul = s->gvc[s->cur_grid] != nullptr;
    if (coul && vcol) {
        c.resize(nodes_alloc);
        if ((r = copy_to_host(s, c.data(), s->gvc[s->cur_grid], nodes_alloc * sizeof(float4)))) return r;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    DevGrid g = s->dev_grid(s->cur_grid);
    const int W1 = P.WIDTH + 1, H1 = P.HEIGHT + 1, D1 = d == 3 ? P.DEPTH + 1 : 1;
    const size_t G = (size_t)W1 * H1 * D1;
    for (int k = 0; k < D1; ++k)
        for (int j = 0; j < H1; ++j)
            for (int i = 0; i < W1; ++i) {
                size_t id = ((size_t)k * H1 + j) * W1 + i;
                int addr = d == 3 ? node_addr<3>(g, i, j, k) : node_addr<2>(g, i, j, 0);
                const float4 q = a[addr];
                const float comp[3] = {q.y, q.z, q.w};
                if (mass) mass[id] = q.x;
                for (int cc = 0; cc < d; ++cc) {
                    if (mom) mom[cc * G + id] = comp[cc];
                    if (vfric) vfric[cc * G + id] = comp[cc];
                    if (vcol) {
                        if (coul) {
                            const float cv[3] = {c[addr].x, c[addr].y, c[addr].z};
                            vcol[cc * G + id] = cv[cc];
                        } else {
                            vcol[cc * G + id] = comp[cc];
                        }
                    }
                }
            }
    return MPM_OK;
}

int mpm_get_grid_force(mpm_solver* s, float* force)
{
    if (!s || !force) return fail(MPM_E_INVALID, "NULL argument");
    if (reject_batch(s)) return MPM_E_STATE;
    if (!s->gf[s->cur_grid]) return fail(MPM_E_STATE, "the solver's scheme has no grid force");
    const mpm_params& P = s->params;
    const int d = s->dim;
    const size_t nodes_alloc = (size_t)s->nblocks * 64;
    std::vector<float4> f(nodes_alloc);
    int r;
    if ((r = copy_to_host(s, f.data(), s->gf[s->cur_grid], nodes_alloc * sizeof(float4)))) return r;
    HIP_TRY(hipStreamSynchronize(s->stream));
    DevGrid g = s->dev_grid(s->cur_grid);
    const int W1 = P.WIDTH + 1, H1 = P.HEIGHT + 1, D1 = d == 3 ? P.DEPTH + 1 : 1;
    const size_t G = (size_t)W1 * H1 * D1;
    for (int k = 0; k < D1; ++k)
        for (int j = 0; j < H1; ++j)
            for (int i = 0; i < W1; ++i) {
                size_t id = ((size_t)k * H1 + j) * W1 + i;
                int addr = d == 3 ? node_addr<3>(g, i, j, k) : node_addr<2>(g, i, j, 0);
                const float comp[3] = {f[addr].x, f[addr].y, f[addr].z};
                for (int cc = 0; cc < d; ++cc) force[cc * G + id] = comp[cc];
            }
    return MPM_OK;
}

int mpm_num_particles(const mpm_solver* s, int64_t* n)
{
    if (!s || !n) return fail(MPM_E_INVALID, "NULL argument");
    *n = s->n;
    return MPM_OK;
}

int mpm_set_timing(mpm_solver* s, int enabled)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    s->timing = enabled != 0;
    s->timing_mask = ~0u;
    return MPM_OK;
}

int mpm_set_timing_mask(mpm_solver* s, uint32_t kernels)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    s->timing = kernels != 0;
    s->timing_mask = kernels;
    return MPM_OK;
}

int mpm_set_timing_every(mpm_solver* s, int every)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    if (every < 1) return fail(MPM_E_INVALID, "every must be >= 1");
    s->timing_every = every;
    for (auto& c : s->timing_seen) c = 0;
    return MPM_OK;
}

int mpm_reset_timers(mpm_solver* s)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    int r = collect_timers(s);
    if (r) return r;
    for (auto& t : s->timers) { t.total_ms = 0.0; t.launches = 0; }
    return MPM_OK;
}

int mpm_get_kernel_time(mpm_solver* s, int kernel, double* total_ms, int64_t* launches)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    if (kernel < 0 || kernel >= MPM_KERNEL_COUNT) return fail(MPM_E_INVALID, "unknown kernel id");
    int r = collect_timers(s);
    if (r) return r;
    if (total_ms) *total_ms = s->timers[kernel].total_ms;
    if (launches) *launches = s->timers[kernel].launches;
    return MPM_OK;
}

int mpm_check_finite(mpm_solver* s, int64_t* bad_particles)
{
    if (!s) return fail(MPM_E_INVALID, "solver is NULL");
    HIP_TRY(hipMemsetAsync(s->d_bad, 0, sizeof(unsigned long long), s->stream));
    HIP_TRY(launch_check_finite(s->dim, s->dev_particles(), s->d_bad, s->stream));
    unsigned long long bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, s->d_bad, sizeof(bad), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (bad_particles) *bad_particles = (int64_t)bad;
    if (bad) return fail(MPM_E_NONFINITE, std::to_string(bad) + " particles hold NaN/Inf");
    return MPM_OK;
}

int mpm_sort_pairs_host(const uint32_t* keys, int64_t n, int key_bits, uint32_t* out_keys,
                        uint32_t* out_perm)
{
    if (n < 0 || n > (int64_t)INT32_MAX / 4) return fail(MPM_E_INVALID, "n out of range");
    if (key_bits < 1 || key_bits > 32) return fail(MPM_E_INVALID, "key_bits must be in [1, 32]");
    if (n > 0 && !keys) return fail(MPM_E_INVALID, "keys is NULL");
    if (n == 0) return MPM_OK;
    if (key_bits < 32)
        for (int64_t i = 0; i < n; ++i)
            if (keys[i] >> key_bits) return fail(MPM_E_INVALID, "key exceeds key_bits");
    const int nn = (int)n;
    uint32_t* buf = nullptr;
    const size_t words = 4 * (size_t)nn + radix_scratch_words(nn);
    HIP_TRY(hipMalloc(&buf, words * 4));
    hipStream_t st = nullptr;
    int rc = MPM_OK;
    do {
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) { rc = fail(MPM_E_HIP, "stream"); break; }
        uint32_t *k = buf, *v = buf + nn, *ka = buf + 2 * (size_t)nn, *va = buf + 3 * (size_t)nn;
        uint32_t* scratch = buf + 4 * (size_t)nn;
        if (hipMemcpyAsync(k, keys, (size_t)nn * 4, hipMemcpyHostToDevice, st) != hipSuccess) { rc = fail(MPM_E_HIP, "upload"); break; }
        bool in_alt = false;
        hipError_t e = radix_sort_pairs(k, v, ka, va, nn, key_bits, true, scratch, st, &in_alt);
        if (e != hipSuccess) { rc = fail(MPM_E_HIP, std::string("radix sort: ") + hipGetErrorString(e)); break; }
        if (out_keys && hipMemcpyAsync(out_keys, in_alt ? ka : k, (size_t)nn * 4, hipMemcpyDeviceToHost, st) != hipSuccess) { rc = fail(MPM_E_HIP, "download"); break; }
        if (out_perm && hipMemcpyAsync(out_perm, in_alt ? va : v, (size_t)nn * 4, hipMemcpyDeviceToHost, st) != hipSuccess) { rc = fail(MPM_E_HIP, "download"); break; }
        if (hipStreamSynchronize(st) != hipSuccess) { rc = fail(MPM_E_HIP, "sync"); break; }
    } while (0);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(buf);
    return rc;
}

const char* mpm_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
