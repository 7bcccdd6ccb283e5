"""Index contracts of the kernels that run on particles a surrogate macro step has
moved (CPU).

Two GPU runs that took plain substeps after macro steps ended in illegal memory
accesses whose cause is not established; the library therefore refuses plain
substeps on such a solver (tests/test_gpu_surrogate.py).  This test pins, on the
host, every index those kernels form from particle positions, for post-macro
states that a bad network can produce: particles pushed past the walls and out of
the image, huge finite values, infinities and NaN.  A host program built from the
same headers as the kernels (csrc/mpm_device.hpp: cell_floor, Spline<2>::base,
node_addr, block_of; csrc/mpm_fused_kernel.hpp: TileShape) follows the index
arithmetic of

  k_cell_keys, k_item_counts / k_item_write   (csrc/mpm_sort.hip)
  k_tiled: tile test, tile entries, direct scatter, flush, block flags
                                              (csrc/mpm_fused_kernel.hpp)
  k_capture box and image writes, k_macro_g2p (csrc/mpm_surrogate.hip)

and checks each index against the allocation it addresses (nblocks * 64 grid
nodes, nblocks flags, the tile's TN entries and 9 touched-block flags, max_items
work items, the h * w image, the capture box).  Every check passing says that no
such state reaches outside a buffer through this arithmetic; it does not name the
cause of the two faults.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))

PROGRAM = r"""
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "mpm_fused_kernel.hpp"

using namespace mpm;
static long violations = 0;
#define CHECK(cond, what) do { if (!(cond)) { if (violations < 20) std::printf("VIOLATION %s\n", what); ++violations; } } while (0)

int main(int argc, char** argv)
{
    // stdin: W H dx tile_p n, then x[0][n], x[1][n]
    int W, H, tile_p, n; float dx;
    if (std::scanf("%d %d %f %d %d", &W, &H, &dx, &tile_p, &n) != 5) return 2;
    std::vector<float> x(2 * (size_t)n);
    for (auto& v : x) if (std::scanf("%f", &v) != 1) return 2;
    const float inv_dx = 1.0f / dx;
    using TS = TileShape<2>;
    DevGrid g{};
    g.nbx = (W + 1 + 7) / 8; g.nby = (H + 1 + 7) / 8; g.nbz = 1;
    const long nblocks = (long)g.nbx * g.nby, nodes = nblocks * 64;
    auto in_grid = [&](int i, int j) { return i >= 0 && i <= W && j >= 0 && j <= H; };

    // k_cell_keys: cell of the position clamped into the grid
    std::vector<uint32_t> keys(n);
    for (int p = 0; p < n; ++p) {
        int c[2];
        const int lim[2] = {W, H};
        for (int d = 0; d < 2; ++d) c[d] = std::min(std::max(cell_floor(x[d * (size_t)n + p] * inv_dx), 0), lim[d]);
        const int a = node_addr<2>(g, c[0], c[1], 0);
        CHECK(a >= 0 && a < nodes, "cell key outside the grid");
        keys[p] = (uint32_t)a;
    }
    // stable sort by key (the radix sort is checked bit-exact against this elsewhere)
    std::vector<int> perm(n);
    for (int p = 0; p < n; ++p) perm[p] = p;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return keys[a] < keys[b]; });
    std::vector<uint32_t> sk(n);
    for (int s = 0; s < n; ++s) sk[s] = keys[perm[s]];
    // k_item_counts / k_item_write
    const long max_items = std::min<long>(n, nblocks) + n / tile_p + 1;
    std::vector<int> it_blk, it_first, it_cnt;
    for (int s = 0; s < n; ++s) {
        const uint32_t b = sk[s] >> 6;
        if (s > 0 && (sk[s - 1] >> 6) == b) continue;
        int e = s + 1;
        while (e < n && (sk[e] >> 6) == b) ++e;
        const int len = e - s;
        const int nch = (len + tile_p - 1) / tile_p;
        for (int c = 0; c < nch; ++c) {
            const int lo = (int)((long long)c * len / nch), hi = (int)((long long)(c + 1) * len / nch);
            CHECK(hi - lo <= tile_p, "work item larger than tile_p");
            it_blk.push_back((int)b); it_first.push_back(s + lo); it_cnt.push_back(hi - lo);
        }
    }
    CHECK((long)it_blk.size() <= max_items, "more work items than max_items");
    // k_tiled over every item
    for (size_t it = 0; it < it_blk.size(); ++it) {
        const int blk = it_blk[it];
        CHECK(blk >= 0 && blk < nblocks, "item block outside the grid");
        CHECK(it_first[it] >= 0 && it_first[it] + it_cnt[it] <= n, "item particles outside [0, n)");
        const int bc[2] = {blk % g.nbx, blk / g.nbx};
        const int o[2] = {bc[0] * TS::BS - 1 - TS::MG, bc[1] * TS::BS - 1 - TS::MG};
        for (int q = 0; q < it_cnt[it]; ++q) {
            const int p = perm[it_first[it] + q];
            float xg[2] = {x[p] * inv_dx, x[(size_t)n + p] * inv_dx};
            int base[2], rb[2];
            bool tile_ok = true;
            for (int d = 0; d < 2; ++d) {
                base[d] = Spline<2>::base(xg[d]);
                rb[d] = base[d] - o[d];
                tile_ok &= rb[d] >= 0 && rb[d] + 3 <= TS::T;
            }
            for (int a0 = 0; a0 < 3; ++a0)
                for (int a1 = 0; a1 < 3; ++a1) {
                    const int i = base[0] + a0, j = base[1] + a1;
                    CHECK(std::abs((long long)i) <= (1ll << 30) + 64 && std::abs((long long)j) <= (1ll << 30) + 64 &&
                          std::abs((long long)rb[0]) < (1ll << 31) - 64 && std::abs((long long)rb[1]) < (1ll << 31) - 64,
                          "stencil or tile index beyond the bounded cell range");
                    if (!in_grid(i, j)) continue;
                    if (tile_ok) {
                        const int e = (rb[1] + a1) * TS::T + (rb[0] + a0);
                        CHECK(e >= 0 && e < TS::TN, "tile entry outside the tile");
                    } else {
                        const int a = node_addr<2>(g, i, j, 0);
                        CHECK(a >= 0 && a < nodes, "direct scatter outside the grid");
                        const int b = block_of<2>(g, i, j, 0);
                        CHECK(b >= 0 && b < nblocks, "direct flag outside the flags");
                    }
                }
        }
        // flush: every tile entry inside the grid, and its touched-block flag
        for (int e = 0; e < TS::TN; ++e) {
            int t[3];
            t[0] = e % TS::T; t[1] = e / TS::T;
            const int i = o[0] + t[0], j = o[1] + t[1];
            if (!in_grid(i, j)) continue;
            const int a = node_addr<2>(g, i, j, 0);
            CHECK(a >= 0 && a < nodes, "flush outside the grid");
            const int rbx = (i >> 3) - (bc[0] - 1), rby = (j >> 3) - (bc[1] - 1);
            CHECK(rbx >= 0 && rbx < 3 && rby >= 0 && rby < 3, "touched-block flag outside its 3x3");
        }
    }
    // k_capture (per workgroup of 256 sorted particles) and k_macro_g2p: image of
    // w = W, h = H nodes, CAP_NODES = 1024
    const int w = W, h = H, CAP = 1024;
    for (int first = 0; first < n; first += 256) {
        const int cnt = std::min(256, n - first);
        int bx0 = INT_MAX, by0 = INT_MAX, bx1 = INT_MIN, by1 = INT_MIN;
        std::vector<int> b0(cnt), b1(cnt);
        for (int k = 0; k < cnt; ++k) {
            const int p = perm[first + k];
            b0[k] = Spline<2>::base(x[p] * inv_dx);
            b1[k] = Spline<2>::base(x[(size_t)n + p] * inv_dx);
            bx0 = std::min(bx0, b0[k]); bx1 = std::max(bx1, b0[k]);
            by0 = std::min(by0, b1[k]); by1 = std::max(by1, b1[k]);
        }
        const long long bwl = (long long)bx1 - bx0 + 3, bhl = (long long)by1 - by0 + 3;
        const bool boxed = bwl <= CAP && bhl <= CAP && bwl * bhl <= CAP;
        for (int k = 0; k < cnt; ++k)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const int i = b0[k] + a, j = b1[k] + b;
                    if (boxed) {
                        const long long qn = (long long)(b1[k] + b - by0) * bwl + (b0[k] + a - bx0);
                        CHECK(qn >= 0 && qn < bwl * bhl, "capture box entry outside the box");
                    }
                    if (i < 0 || i >= w || j < 0 || j >= h) continue;
                    const long long gq = (long long)j * w + i;
                    CHECK(gq >= 0 && gq < (long long)h * w, "image entry outside the image");
                }
        if (boxed)
            for (long long qn = 0; qn < bwl * bhl; ++qn) {
                const long long i = bx0 + qn % bwl, j = by0 + qn / bwl;
                if (i < 0 || i >= w || j < 0 || j >= h) continue;
                CHECK(j * w + i < (long long)h * w, "capture flush outside the image");
            }
    }
    std::printf("checked %d particles, %zu items, violations %ld\n", n, it_blk.size(), violations);
    return violations ? 1 : 0;
}
"""


def _post_macro_states(rng):
    """Positions after macro steps of a bad network: the restated macro step
    (tests/surrogate_floor.py, k_macro_g2p) with large random grid increments, and
    explicit extreme values."""
    import surrogate_floor as SF
    from mpm_snow_sim2d_amd.params import snow_params
    from mpm_snow_sim2d_amd.scenes import colliding_scene
    from oracle import oracle as orc
    P = snow_params(2, 64, 2e-4)
    ps = colliding_scene(11, 3000)
    st = orc.OracleState.from_particles(ps, np.float64)
    img = SF.Image(P, 64, 64)
    out = []
    k = 8
    for scale in (1.0, 30.0, 1e4):
        s = orc.OracleState(st.x.copy(), st.v.copy(), st.B.copy(), st.Fe.copy(), st.Fp.copy(), st.mass, st.vol)
        for _ in range(3):
            sten = img.stencil(s.x)
            vf, _, _ = img.transfer(sten, s.mass, s.v)
            vf = vf + rng.normal(scale=scale, size=vf.shape)
            SF.macro_step(P, img, s, k, vf, None)
        out.append(s.x.astype(np.float32))
    x = st.x.astype(np.float32).copy()
    special = np.array([0.0, -0.0, 1.0, 64 / 64, 63.99 / 64, -1e-7, 1e30, -1e30, np.inf, -np.inf, np.nan,
                        3.0e38, -3.0e38, 2.0 ** 24, -(2.0 ** 24), 1e-45], np.float32)
    for d in range(2):
        xs = x.copy()
        xs[d, :special.size] = special
        xs[1 - d, :special.size] = special[::-1]
        out.append(xs)
    return P, out


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available")
def test_post_macro_indices_stay_in_their_buffers(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    src = tmp_path / "post_macro.hip"
    src.write_text(PROGRAM)
    exe = tmp_path / "post_macro"
    cmd = [hipcc, "-std=c++17", "-O1", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "mpm_snow_sim2d_amd", "csrc"), str(src), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    P, states = _post_macro_states(np.random.default_rng(3))
    moved_out = 0
    for x in states:
        with np.errstate(invalid="ignore"):
            moved_out += int((~((x >= 0) & (x <= 1.0))).any(axis=0).sum())
        for tile_p in (128, 512):
            inp = f"{P.WIDTH} {P.HEIGHT} {float(np.float32(P.dx))!r} {tile_p} {x.shape[1]}\n" + \
                  " ".join(repr(float(v)) if np.isfinite(v) else ("nan" if np.isnan(v) else ("inf" if v > 0 else "-inf"))
                           for v in x.ravel())
            out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, timeout=120)
            assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
            assert "violations 0" in out.stdout
    # the states do reach outside the domain (the case the checks are about)
    assert moved_out > 100
