// mpm/GPUSolver.hpp — header-only C++ face of libmpm_hip.so with the reference's
// solver interface.
//
// The reference selects its solver at compile time (src/main.cpp:22-29) and uses it
// as
//     Solver solver;                                  // src/main.cpp:268
//     solver.Step();                                  // src/main.cpp:169, 34 per frame
//     const auto& pos = solver.GetParticlePositions(); // src/main.cpp:114,146
// (src/gpu/GPUSolver.h:10-17, src/cpu/CPUSolver.h:10-27).  mpm::GPUSolver keeps those
// three members with the same meaning: the constructor uploads the particle set and
// throws std::runtime_error on failure (the reference throws from its constructor,
// GPUSolver.cpp:39-50), Step() enqueues one substep asynchronously, and
// GetParticlePositions() synchronises and returns a solver-owned vector valid until
// the next call (GPUSolver.h:19).  The stage calls and state access of mpm.h are
// exposed too.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "mpm.h"

namespace mpm {

template <int DIM>
struct Vec {
    float v[DIM];
    float& operator[](int i) { return v[i]; }
    const float& operator[](int i) const { return v[i]; }
};

template <int DIM>
class GPUSolver {
public:
    // params.dim must equal DIM.  x, v: component-major [d*n + p]; mass, vol: [n].
    GPUSolver(const mpm_params& params, int64_t n, const float* x, const float* v, const float* mass,
              const float* vol, const mpm_options* opts = nullptr)
    {
        if (params.dim != DIM) throw std::runtime_error("mpm::GPUSolver: params.dim does not match DIM");
        mpm_particles init{};
        init.n = n;
        init.x = x;
        init.v = v;
        init.mass = mass;
        init.vol = vol;
        check(mpm_create(&params, &init, opts, &h_));
        n_ = n;
    }
    ~GPUSolver() { mpm_destroy(h_); }
    GPUSolver(const GPUSolver&) = delete;
    GPUSolver& operator=(const GPUSolver&) = delete;

    // GPUSolver::Step (GPUSolver.cpp:63-80).
    void Step(int n_substeps = 1) { check(mpm_step(h_, n_substeps)); }

    // GPUSolver::GetParticlePositions (GPUSolver.cpp:276-295).
    const std::vector<Vec<DIM>>& GetParticlePositions()
    {
        cached_.resize((size_t)n_);
        if (n_ > 0) check(mpm_get_positions(h_, &cached_[0][0]));
        return cached_;
    }

    void P2G() { check(mpm_p2g(h_)); }
    void GridUpdate() { check(mpm_grid_update(h_)); }
    void G2P() { check(mpm_g2p(h_)); }
    void Reset() { check(mpm_reset_grid(h_)); }
    void Sync() { check(mpm_sync(h_)); }
    mpm_solver* handle() { return h_; }

private:
    static void check(int code)
    {
        if (code != MPM_OK) throw std::runtime_error(std::string("mpm: ") + mpm_last_error());
    }
    mpm_solver* h_ = nullptr;
    int64_t n_ = 0;
    std::vector<Vec<DIM>> cached_;
};

// The CPU path's solver (src/cpu/CPUSolver.h:10-27: APIC transfer with an explicit
// grid force, src/cpu/Grid.cpp and src/cpu/Particles.cpp) on the device: the same
// interface, with the parameter set's scheme switched to MPM_SCHEME_APIC_FORCE.
template <int DIM>
class CPUSolver : public GPUSolver<DIM> {
public:
    CPUSolver(const mpm_params& params, int64_t n, const float* x, const float* v, const float* mass,
              const float* vol, const mpm_options* opts = nullptr)
        : GPUSolver<DIM>(with_cpu_scheme(params), n, x, v, mass, vol, opts)
    {
    }

private:
    static mpm_params with_cpu_scheme(mpm_params p)
    {
        p.scheme = MPM_SCHEME_APIC_FORCE;
        return p;
    }
};

}  // namespace mpm
