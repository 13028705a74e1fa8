#include "Simulation.h"

#include <chrono>
#include <cstdio>

Simulation::Simulation(SubDomain& sd, double tau, double pressure, double del_t)
    : SubDomain_(sd), m_relaxation(tau), m_omega(1.0 / tau), pressure_(pressure), m_dt(del_t) {}

void Simulation::WriteOutput(int iter, const std::string& dir, bool raw) {
    OutputFile file(iter, dir, raw);
    if (!file.WriteVTK_LocalData(SubDomain_) || !file.WriteWallShearStressVTK(SubDomain_, iter))
        error = "cannot write the output of iteration " + std::to_string(iter);
}

bool Simulation::CollectNodeFields() {
    SubDomain& sd = SubDomain_;
    const size_t ns = (size_t)sd.my_Nx_ * sd.my_Ny_ * sd.my_Nz_;
    sd.density.resize(ns);
    sd.velocity.resize(3 * ns);
    sd.wall_shear_stress.resize(ns);
    return lbm_get_macros(sd.dom, sd.density.data(), sd.velocity.data()) == LBM_OK &&
           lbm_get_wss(sd.dom, sd.wall_shear_stress.data()) == LBM_OK;
}

bool Simulation::Advance(long long n) {
    if (n <= 0) return true;
    const auto t0 = std::chrono::steady_clock::now();
    if (lbm_step(SubDomain_.dom, n) != LBM_OK || lbm_sync(SubDomain_.dom) != LBM_OK) {
        error = "lbm_step failed";
        return false;
    }
    seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

bool Simulation::Iterate(long long steps, int output_every, const std::string& dir, bool raw) {
    long long iter = 0;
    while (iter < steps) {
        // the next output iteration n (a multiple of output_every), if any
        long long n = steps;
        if (output_every > 0) n = (iter + output_every - 1) / output_every * output_every;
        if (n >= steps) return Advance(steps - iter);
        // state S_n: the moments iteration n's collision starts from
        if (!Advance(n - iter)) return false;
        if (!CollectNodeFields()) {
            error = "reading the state back failed";
            return false;
        }
        // iteration n itself, then its output (Simulation.cpp:909-922)
        if (!Advance(1)) return false;
        iter = n + 1;
        WriteOutput((int)n, dir, raw);
        if (!error.empty()) return false;
        double mass = 0.0;
        const int rc = lbm_reduce(SubDomain_.dom, &mass, nullptr);
        std::printf("Iteration %lld has been performed successfully (mass %.12g)\n", n, mass);
        std::fflush(stdout);
        if (rc == LBM_ENAN) {
            error = "non-finite state";
            return false;
        }
    }
    return true;
}
