// mpm_fused_cubic.hip — the fused G2P/P2G kernel for 3D scenes with the cubic
// stencil and Coulomb walls (the reference's own scheme, Parameters.h and
// shaders/3d/*.glsl), single scene.  Built with a higher full-unroll limit
// (-pragma-unroll-threshold, set in build.py for this file only) so that the
// 64-node stencil loops unroll completely instead of indexing stencil arrays at run
// time (726 -> 643 us per launch on the 2M-particle falling cube).
#include "mpm_fused_kernel.hpp"

namespace mpm {

hipError_t launch_g2p2g_cubic_coulomb3d(dim3 grid, const DevParticles& p, const DevGrid& src, const DevGrid& dst,
                                        const TileWork& w, const Phys& ph, hipStream_t st)
{
    k_tiled<3, 3, true, true, false><<<grid, dim3(256), 0, st>>>(p, src, dst, w, ph, nullptr);
    return hipGetLastError();
}

}  // namespace mpm
