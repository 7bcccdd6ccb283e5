// mpm_fused.hip — launchers of the tiled, LDS-staged particle kernels
// (mpm_fused_kernel.hpp): the fused G2P/P2G substep and the P2G-only first pass.
#include "mpm_fused_kernel.hpp"

namespace mpm {

hipError_t launch_g2p2g(const LaunchCfg& c, const DevParticles& p, const DevGrid& src, const DevGrid& dst,
                        const TileWork& w, const Phys& ph, const SceneDesc* scenes, hipStream_t st)
{
    if (p.n == 0) return hipSuccess;
    dim3 grid(tiled_grid(w)), blk(256);
    bool coul = c.boundary == MPM_BOUNDARY_COULOMB_MARGIN;
    const bool batch = scenes != nullptr;
#define TCASE(D, O)                                                                          \
    if (c.dim == D && c.order == O) {                                                        \
        if (batch) {                                                                         \
            if (coul) k_tiled<D, O, true, true, true><<<grid, blk, 0, st>>>(p, src, dst, w, ph, scenes);   \
            else      k_tiled<D, O, false, true, true><<<grid, blk, 0, st>>>(p, src, dst, w, ph, scenes);  \
        } else {                                                                             \
            if (coul) k_tiled<D, O, true, true, false><<<grid, blk, 0, st>>>(p, src, dst, w, ph, scenes);  \
            else      k_tiled<D, O, false, true, false><<<grid, blk, 0, st>>>(p, src, dst, w, ph, scenes); \
        }                                                                                    \
        return hipGetLastError();                                                            \
    }
    TCASE(2, 3) TCASE(2, 2) TCASE(3, 2)
#undef TCASE
    if (c.dim == 3 && c.order == 3) {
        if (batch) {
            if (coul) k_tiled<3, 3, true, true, true><<<grid, blk, 0, st>>>(p, src, dst, w, ph, scenes);
            else      k_tiled<3, 3, false, true, true><<<grid, blk, 0, st>>>(p, src, dst, w, ph, scenes);
            return hipGetLastError();
        }
        if (coul) return launch_g2p2g_cubic_coulomb3d(grid, p, src, dst, w, ph, st);
        k_tiled<3, 3, false, true, false><<<grid, blk, 0, st>>>(p, src, dst, w, ph, scenes);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

hipError_t launch_p2g_tiled(const LaunchCfg& c, const DevParticles& p, const DevGrid& dst,
                            const TileWork& w, const Phys& ph, const SceneDesc* scenes, hipStream_t st)
{
    if (p.n == 0) return hipSuccess;
    dim3 grid(tiled_grid(w)), blk(256);
    const bool batch = scenes != nullptr;
#define TCASE(D, O)                                                                          \
    if (c.dim == D && c.order == O) {                                                        \
        if (batch) k_tiled<D, O, false, false, true><<<grid, blk, 0, st>>>(p, dst, dst, w, ph, scenes);  \
        else       k_tiled<D, O, false, false, false><<<grid, blk, 0, st>>>(p, dst, dst, w, ph, scenes); \
        return hipGetLastError();                                                            \
    }
    TCASE(2, 3) TCASE(2, 2) TCASE(3, 3) TCASE(3, 2)
#undef TCASE
    return hipErrorInvalidValue;
}

}  // namespace mpm
