"""The C++ face (include/mpm/GPUSolver.hpp) compiles and links against
libmpm_hip.so with a plain host compiler, the way the reference's main.cpp would
use it (Solver solver; solver.Step(); solver.GetParticlePositions()).  Build only —
running needs a GPU (the GPU suite drives the same entry points through ctypes)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <vector>
#include "mpm/GPUSolver.hpp"

// src/main.cpp:164-174: substeps until 1/SAVE_FPS of simulated time has passed.
template <class Solver>
void Simulate(Solver& solver, float dt, float save_fps)
{
    float cum = 0.f;
    while (cum < 1.0f / save_fps) { solver.Step(); cum += dt; }
}

int main()
{
    mpm_params p;
    if (mpm_params_reference(2, &p) != MPM_OK) return 1;
    std::vector<float> x = {50.f, 51.f, 45.f, 45.5f}, v = {40.f, 40.f, 0.f, 0.f};
    std::vector<float> m = {1.f, 1.f}, vol = {0.25f, 0.25f};
    try {
        mpm::GPUSolver<2> solver(p, 2, x.data(), v.data(), m.data(), vol.data());
        Simulate(solver, p.DELTA_TIME, 30.f);
        const auto& pos = solver.GetParticlePositions();
        std::printf("%f %f\n", pos[0][0], pos[0][1]);
        mpm::CPUSolver<2> cpu(p, 2, x.data(), v.data(), m.data(), vol.data());
        Simulate(cpu, p.DELTA_TIME, 30.f);
        const auto& q = cpu.GetParticlePositions();
        std::printf("%f %f\n", q[0][0], q[0][1]);
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 2;
    }
    return 0;
}
"""


def _build(tmp_path):
    lib_dir = os.path.join(ROOT, "mpm_snow_sim2d_amd")
    if not os.path.exists(os.path.join(lib_dir, "libmpm_hip.so")):
        pytest.fail("libmpm_hip.so not built")
    src = tmp_path / "main.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "demo"
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L", lib_dir, "-lmpm_hip", f"-Wl,-rpath,{lib_dir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.gpu
@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_cpp_program_runs(tmp_path):
    """The reference-shaped program runs both solver faces: 34 substeps each, the
    snowball particles move right at 40 grid units/s under gravity."""
    exe = _build(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    for line in lines[:2]:
        x, y = map(float, line.split())
        assert 50.0 < x < 52.0 and 40.0 < y < 45.5


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_cpp_header_compiles_and_links(tmp_path):
    lib_dir = os.path.join(ROOT, "mpm_snow_sim2d_amd")
    if not os.path.exists(os.path.join(lib_dir, "libmpm_hip.so")):
        pytest.fail("libmpm_hip.so not built")
    src = tmp_path / "main.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "demo"
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L", lib_dir, "-lmpm_hip", f"-Wl,-rpath,{lib_dir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert exe.exists()
