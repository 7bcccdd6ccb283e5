// mpm_io.hip — state snapshots and headless frames.
//
// Snapshots: a little-endian binary file holding the parameter set, the substep
// count and the particle state in the caller's particle order (the state the
// reference reads back in GPUSolver::ReadbackParticles, src/gpu/GPUSolver.cpp:204-225,
// and GetParticlePositions, :276-295):
//   char[8] "MPMSNAP1" | uint32 version (1) | uint32 sizeof(mpm_params) |
//   mpm_params | int64 n | int64 substeps |
//   float x[d][n], v[d][n], C[d*d][n], F[d*d][n], Jp[n], mass[n], vol[n]
//
// Frames: the reference draws a grey border (or wireframe box) and white points and
// saves the window with stbi_write_png (src/main.cpp:56-160, :177-205).  Here the
// particles are splatted by a HIP kernel into an RGB image with the same camera
// (2D: glOrtho(0, WIDTH, 0, HEIGHT); 3D: 45-degree perspective orbiting the grid
// centre, main.cpp:50-79), the border is drawn on the host, and the PNG writer
// emits stored (uncompressed) deflate blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mpm/mpm.h"
#include "mpm_solver_impl.hpp"

using namespace mpm;

namespace {

int io_fail(int code, const std::string& m) { return mpm::detail::set_error(code, m); }

const char kMagic[8] = {'M', 'P', 'M', 'S', 'N', 'A', 'P', '1'};
constexpr uint32_t kVersion = 1;

struct FileCloser {
    FILE* f;
    ~FileCloser()
    {
        if (f) std::fclose(f);
    }
};

int read_header(FILE* f, mpm_params* p, int64_t* n, int64_t* substeps)
{
    char magic[8];
    uint32_t ver = 0, psz = 0;
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kMagic, 8) != 0)
        return io_fail(MPM_E_INVALID, "not a snapshot file");
    if (std::fread(&ver, 4, 1, f) != 1 || ver != kVersion) return io_fail(MPM_E_INVALID, "unsupported snapshot version");
    if (std::fread(&psz, 4, 1, f) != 1 || psz != sizeof(mpm_params))
        return io_fail(MPM_E_INVALID, "snapshot parameter block has the wrong size");
    if (std::fread(p, sizeof(mpm_params), 1, f) != 1 || std::fread(n, 8, 1, f) != 1 ||
        std::fread(substeps, 8, 1, f) != 1)
        return io_fail(MPM_E_INVALID, "truncated snapshot header");
    if (*n < 0 || (p->dim != 2 && p->dim != 3)) return io_fail(MPM_E_INVALID, "corrupt snapshot header");
    return MPM_OK;
}

// ---------------------------------------------------------------------------
// frames
struct Camera {
    float m[16];   // column-major view-projection (clip = m * (x, y, z, 1))
};

__global__ void __launch_bounds__(256) k_splat(const float* __restrict__ x, int n, int dim, Camera cam, int w,
                                              int h, int psize, unsigned char* __restrict__ rgb)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const float px = x[p], py = x[n + p], pz = dim == 3 ? x[2 * n + p] : 0.f;
    const float* m = cam.m;
    const float cx = m[0] * px + m[4] * py + m[8] * pz + m[12];
    const float cy = m[1] * px + m[5] * py + m[9] * pz + m[13];
    const float cw = m[3] * px + m[7] * py + m[11] * pz + m[15];
    if (!(cw > 0.f)) return;
    // normalised device coordinates -> window pixels (y up, as glViewport)
    const float sx = (cx / cw * 0.5f + 0.5f) * (float)w;
    const float sy = (cy / cw * 0.5f + 0.5f) * (float)h;
    const int ix = cell_floor(sx), iy = cell_floor(sy);
    const int r0 = -(psize - 1) / 2, r1 = psize / 2;
    for (int dy = r0; dy <= r1; ++dy)
        for (int dx = r0; dx <= r1; ++dx) {
            const int X = ix + dx, Y = iy + dy;
            if (X < 0 || X >= w || Y < 0 || Y >= h) continue;
            // top-down rows in the file, as the reference's flip before saving
            unsigned char* o = rgb + ((size_t)(h - 1 - Y) * w + X) * 3;
            o[0] = 255;
            o[1] = 255;
            o[2] = 255;
        }
}

void mat_mul(const float* a, const float* b, float* o)   // column-major 4x4: o = a * b
{
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float s = 0.f;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + r] * b[c * 4 + k];
            o[c * 4 + r] = s;
        }
}

// glm::perspective / glm::lookAt / glOrtho in column-major form
Camera make_camera(const mpm_params& P, const mpm_camera& cam, int w, int h)
{
    Camera out{};
    const float Wd = (float)P.WIDTH * P.dx, Hd = (float)P.HEIGHT * P.dx, Dd = (float)P.DEPTH * P.dx;
    if (P.dim == 2) {
        // glOrtho(0, W, 0, H, -1, 1)
        float* m = out.m;
        std::memset(m, 0, sizeof(out.m));
        m[0] = 2.f / Wd;
        m[5] = 2.f / Hd;
        m[10] = -1.f;
        m[12] = -1.f;
        m[13] = -1.f;
        m[15] = 1.f;
        return out;
    }
    const float fov = cam.fov_deg * 3.14159265358979f / 180.f, aspect = (float)w / (float)h;
    const float zn = cam.z_near, zf = cam.z_far;
    const float f = 1.f / std::tan(fov * 0.5f);
    float proj[16] = {0};
    proj[0] = f / aspect;
    proj[5] = f;
    proj[10] = -(zf + zn) / (zf - zn);
    proj[11] = -1.f;
    proj[14] = -(2.f * zf * zn) / (zf - zn);
    const float ctr[3] = {Wd * 0.5f, Hd * 0.5f, Dd * 0.5f};
    const float rad = cam.angle_deg * 3.14159265358979f / 180.f, el = cam.elevation_deg * 3.14159265358979f / 180.f;
    const float eye[3] = {ctr[0] + cam.distance * std::cos(el) * std::sin(rad), ctr[1] + cam.distance * std::sin(el),
                          ctr[2] + cam.distance * std::cos(el) * std::cos(rad)};
    float fwd[3] = {ctr[0] - eye[0], ctr[1] - eye[1], ctr[2] - eye[2]};
    float ln = std::sqrt(fwd[0] * fwd[0] + fwd[1] * fwd[1] + fwd[2] * fwd[2]);
    for (float& v : fwd) v /= ln;
    const float up[3] = {0.f, 1.f, 0.f};
    float side[3] = {fwd[1] * up[2] - fwd[2] * up[1], fwd[2] * up[0] - fwd[0] * up[2], fwd[0] * up[1] - fwd[1] * up[0]};
    ln = std::sqrt(side[0] * side[0] + side[1] * side[1] + side[2] * side[2]);
    for (float& v : side) v /= ln;
    const float u[3] = {side[1] * fwd[2] - side[2] * fwd[1], side[2] * fwd[0] - side[0] * fwd[2],
                        side[0] * fwd[1] - side[1] * fwd[0]};
    float view[16] = {side[0], u[0], -fwd[0], 0.f, side[1], u[1], -fwd[1], 0.f, side[2], u[2], -fwd[2], 0.f,
                      -(side[0] * eye[0] + side[1] * eye[1] + side[2] * eye[2]),
                      -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]),
                      fwd[0] * eye[0] + fwd[1] * eye[1] + fwd[2] * eye[2], 1.f};
    mat_mul(proj, view, out.m);
    return out;
}

bool project(const Camera& c, const float* p, int w, int h, float* sx, float* sy)
{
    const float* m = c.m;
    const float cx = m[0] * p[0] + m[4] * p[1] + m[8] * p[2] + m[12];
    const float cy = m[1] * p[0] + m[5] * p[1] + m[9] * p[2] + m[13];
    const float cw = m[3] * p[0] + m[7] * p[1] + m[11] * p[2] + m[15];
    if (!(cw > 0.f)) return false;
    *sx = (cx / cw * 0.5f + 0.5f) * (float)w;
    *sy = (cy / cw * 0.5f + 0.5f) * (float)h;
    return true;
}

// grey line of width 2 (glLineWidth(2), glColor3f(0.5, 0.5, 0.5)), DDA in window
// pixels; the reference draws the border before the points, so white pixels stay
void draw_line(unsigned char* rgb, int w, int h, float x0, float y0, float x1, float y1)
{
    const int steps = (int)std::ceil(std::max(std::fabs(x1 - x0), std::fabs(y1 - y0))) + 1;
    for (int i = 0; i <= steps; ++i) {
        const float t = steps ? (float)i / (float)steps : 0.f;
        const int X = (int)std::floor(x0 + t * (x1 - x0)), Y = (int)std::floor(y0 + t * (y1 - y0));
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int xx = X + dx, yy = Y + dy;
                if (xx < 0 || xx >= w || yy < 0 || yy >= h) continue;
                unsigned char* o = rgb + ((size_t)(h - 1 - yy) * w + xx) * 3;
                if (o[0] == 255 && o[1] == 255 && o[2] == 255) continue;
                o[0] = o[1] = o[2] = 128;
            }
    }
}

uint32_t crc32_update(uint32_t crc, const unsigned char* d, size_t n)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ d[i]) & 0xFF] ^ (crc >> 8);
    return crc;
}

void put_be32(std::vector<unsigned char>& v, uint32_t x)
{
    v.push_back((unsigned char)(x >> 24));
    v.push_back((unsigned char)(x >> 16));
    v.push_back((unsigned char)(x >> 8));
    v.push_back((unsigned char)x);
}

bool write_chunk(FILE* f, const char* type, const std::vector<unsigned char>& data)
{
    std::vector<unsigned char> buf;
    put_be32(buf, (uint32_t)data.size());
    buf.insert(buf.end(), type, type + 4);
    buf.insert(buf.end(), data.begin(), data.end());
    uint32_t crc = crc32_update(0xFFFFFFFFu, buf.data() + 4, buf.size() - 4) ^ 0xFFFFFFFFu;
    put_be32(buf, crc);
    return std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
}

}  // namespace

extern "C" {

int mpm_save_snapshot(mpm_solver* s, const char* path)
{
    if (!s || !path) return io_fail(MPM_E_INVALID, "NULL argument");
    if (s->nscenes > 1) return io_fail(MPM_E_STATE, "snapshots hold one scene");
    const int d = s->dim;
    const int64_t n = s->n;
    std::vector<float> buf((size_t)(2 * d + 2 * d * d + 1) * n);
    float* x = buf.data();
    float* v = x + (size_t)d * n;
    float* C = v + (size_t)d * n;
    float* F = C + (size_t)d * d * n;
    float* Jp = F + (size_t)d * d * n;
    int r = mpm_get_state(s, x, v, C, F, Jp);
    if (r) return r;
    // mass and volume are constants: read them in caller order through pid
    std::vector<float> mass(n), vol(n), mtmp(n), vtmp(n);
    std::vector<uint32_t> pid(n);
    const detail::ParticleBuffers& b = s->pb[s->cur_pb];
    if (n > 0) {
        if (hipMemcpy(mtmp.data(), b.mass, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(vtmp.data(), b.vol, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(pid.data(), b.pid, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return io_fail(MPM_E_HIP, "snapshot readback failed");
        for (int64_t i = 0; i < n; ++i) {
            mass[pid[i]] = mtmp[i];
            vol[pid[i]] = vtmp[i];
        }
    }
    FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) return io_fail(MPM_E_INVALID, std::string("cannot open ") + path);
    const uint32_t ver = kVersion, psz = sizeof(mpm_params);
    const int64_t steps = s->substeps_done;
    bool ok = std::fwrite(kMagic, 1, 8, fc.f) == 8 && std::fwrite(&ver, 4, 1, fc.f) == 1 &&
              std::fwrite(&psz, 4, 1, fc.f) == 1 && std::fwrite(&s->params, sizeof(mpm_params), 1, fc.f) == 1 &&
              std::fwrite(&n, 8, 1, fc.f) == 1 && std::fwrite(&steps, 8, 1, fc.f) == 1;
    ok = ok && std::fwrite(buf.data(), 4, buf.size(), fc.f) == buf.size();
    ok = ok && std::fwrite(mass.data(), 4, n, fc.f) == (size_t)n && std::fwrite(vol.data(), 4, n, fc.f) == (size_t)n;
    if (!ok) return io_fail(MPM_E_INVALID, std::string("write failed: ") + path);
    return MPM_OK;
}

int mpm_snapshot_info(const char* path, mpm_params* params, int64_t* n, int64_t* substeps)
{
    if (!path) return io_fail(MPM_E_INVALID, "path is NULL");
    FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) return io_fail(MPM_E_INVALID, std::string("cannot open ") + path);
    mpm_params p{};
    int64_t nn = 0, st = 0;
    int r = read_header(fc.f, &p, &nn, &st);
    if (r) return r;
    if (params) *params = p;
    if (n) *n = nn;
    if (substeps) *substeps = st;
    return MPM_OK;
}

int mpm_load_snapshot(const char* path, const mpm_options* opts, mpm_solver** out)
{
    if (!path || !out) return io_fail(MPM_E_INVALID, "NULL argument");
    *out = nullptr;
    FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) return io_fail(MPM_E_INVALID, std::string("cannot open ") + path);
    mpm_params p{};
    int64_t n = 0, steps = 0;
    int r = read_header(fc.f, &p, &n, &steps);
    if (r) return r;
    const int d = p.dim;
    std::vector<float> buf((size_t)(2 * d + 2 * d * d + 3) * n);
    if (std::fread(buf.data(), 4, buf.size(), fc.f) != buf.size()) return io_fail(MPM_E_INVALID, "truncated snapshot");
    float* x = buf.data();
    float* v = x + (size_t)d * n;
    float* C = v + (size_t)d * n;
    float* F = C + (size_t)d * d * n;
    float* Jp = F + (size_t)d * d * n;
    float* mass = Jp + n;
    float* vol = mass + n;
    mpm_particles init{n, x, v, mass, vol, C, F, Jp};
    r = mpm_create(&p, &init, opts, out);
    if (r) return r;
    (*out)->substeps_done = steps;
    return MPM_OK;
}

int mpm_render_frame(mpm_solver* s, const mpm_camera* cam, int width, int height, int point_size,
                     unsigned char* rgb)
{
    if (!s || !rgb) return io_fail(MPM_E_INVALID, "NULL argument");
    if (s->nscenes > 1) return io_fail(MPM_E_STATE, "frames show one scene");
    if (width < 1 || height < 1 || width > 16384 || height > 16384) return io_fail(MPM_E_INVALID, "image size out of range");
    if (point_size < 1 || point_size > 16) return io_fail(MPM_E_INVALID, "point_size must be in [1, 16]");
    mpm_camera c{};
    if (cam) c = *cam;
    else mpm_camera_default(&c);
    const mpm_params& P = s->params;
    const Camera m = make_camera(P, c, width, height);
    const size_t bytes = (size_t)width * height * 3;
    unsigned char* d_img = nullptr;
    if (hipMalloc(&d_img, bytes) != hipSuccess) return io_fail(MPM_E_HIP, "frame allocation failed");
    int rc = MPM_OK;
    if (hipMemsetAsync(d_img, 0, bytes, s->stream) != hipSuccess) rc = io_fail(MPM_E_HIP, "memset failed");
    if (rc == MPM_OK && s->n > 0) {
        k_splat<<<(s->n + 255) / 256, 256, 0, s->stream>>>(s->pb[s->cur_pb].x, s->n, s->dim, m, width, height,
                                                           point_size, d_img);
        if (hipGetLastError() != hipSuccess) rc = io_fail(MPM_E_HIP, "splat launch failed");
    }
    if (rc == MPM_OK && (hipMemcpyAsync(rgb, d_img, bytes, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
                         hipStreamSynchronize(s->stream) != hipSuccess))
        rc = io_fail(MPM_E_HIP, "frame readback failed");
    (void)hipFree(d_img);
    if (rc) return rc;
    // border (2D) or wireframe box (3D) at BORDER_MARGIN cells, under the points
    const float mg = P.BORDER_MARGIN * P.dx;
    const float X0 = mg, Y0 = mg, Z0 = mg;
    const float X1 = P.WIDTH * P.dx - mg, Y1 = P.HEIGHT * P.dx - mg, Z1 = P.DEPTH * P.dx - mg;
    std::vector<std::pair<std::array<float, 3>, std::array<float, 3>>> edges;
    if (P.dim == 2) {
        const float q[4][3] = {{X0, Y0, 0}, {X1, Y0, 0}, {X1, Y1, 0}, {X0, Y1, 0}};
        for (int i = 0; i < 4; ++i)
            edges.push_back({{q[i][0], q[i][1], 0.f}, {q[(i + 1) % 4][0], q[(i + 1) % 4][1], 0.f}});
    } else {
        const float xs[2] = {X0, X1}, ys[2] = {Y0, Y1}, zs[2] = {Z0, Z1};
        for (int a = 0; a < 2; ++a)
            for (int b2 = 0; b2 < 2; ++b2) {
                edges.push_back({{X0, ys[a], zs[b2]}, {X1, ys[a], zs[b2]}});
                edges.push_back({{xs[a], Y0, zs[b2]}, {xs[a], Y1, zs[b2]}});
                edges.push_back({{xs[a], ys[b2], Z0}, {xs[a], ys[b2], Z1}});
            }
    }
    for (auto& e : edges) {
        float ax, ay, bx, by;
        if (project(m, e.first.data(), width, height, &ax, &ay) && project(m, e.second.data(), width, height, &bx, &by))
            draw_line(rgb, width, height, ax, ay, bx, by);
    }
    return MPM_OK;
}

int mpm_camera_default(mpm_camera* cam)
{
    if (!cam) return io_fail(MPM_E_INVALID, "cam is NULL");
    cam->angle_deg = 0.f;        // main.cpp:51-53
    cam->elevation_deg = 30.f;
    cam->distance = 300.f;
    cam->fov_deg = 45.f;         // main.cpp:62
    cam->z_near = 1.f;
    cam->z_far = 1000.f;
    return MPM_OK;
}

int mpm_write_png(const char* path, int width, int height, const unsigned char* rgb)
{
    if (!path || !rgb) return io_fail(MPM_E_INVALID, "NULL argument");
    if (width < 1 || height < 1) return io_fail(MPM_E_INVALID, "image size out of range");
    FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) return io_fail(MPM_E_INVALID, std::string("cannot open ") + path);
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    bool ok = std::fwrite(sig, 1, 8, fc.f) == 8;
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, (uint32_t)width);
    put_be32(ihdr, (uint32_t)height);
    ihdr.push_back(8);   // bit depth
    ihdr.push_back(2);   // colour type RGB
    ihdr.push_back(0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    ok = ok && write_chunk(fc.f, "IHDR", ihdr);
    // zlib stream of stored deflate blocks over the filtered rows (filter 0)
    const size_t row = (size_t)width * 3 + 1;
    std::vector<unsigned char> raw(row * height);
    for (int y = 0; y < height; ++y) {
        raw[y * row] = 0;
        std::memcpy(&raw[y * row + 1], rgb + (size_t)y * width * 3, (size_t)width * 3);
    }
    std::vector<unsigned char> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((unsigned char)(len & 0xFF));
        z.push_back((unsigned char)(len >> 8));
        z.push_back((unsigned char)(~len & 0xFF));
        z.push_back((unsigned char)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (unsigned char c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    ok = ok && write_chunk(fc.f, "IDAT", z);
    ok = ok && write_chunk(fc.f, "IEND", {});
    if (!ok) return io_fail(MPM_E_INVALID, std::string("write failed: ") + path);
    return MPM_OK;
}

}  // extern "C"
