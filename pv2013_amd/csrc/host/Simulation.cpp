#include "Simulation.h"

#include <chrono>
#include <cstdio>

Simulation::Simulation(SubDomain& sd, double tau, double pressure, double del_t)
    : SubDomain_(sd), m_relaxation(tau), m_omega(1.0 / tau), pressure_(pressure), m_dt(del_t) {}

void Simulation::WriteOutput(int iter, const std::string& dir) {
    OutputFile file(iter, dir);
    file.WriteVTK_LocalData(SubDomain_);
    file.WriteWallShearStressVTK(SubDomain_, iter);
}

bool Simulation::Iterate(long long steps, int output_every, const std::string& dir) {
    long long iter = 0;
    while (iter < steps) {
        // run up to and including the next output iteration
        long long next = steps;
        if (output_every > 0) {
            const long long k = (iter + output_every - 1) / output_every * output_every;
            next = std::min(steps, k + 1);
        }
        const auto t0 = std::chrono::steady_clock::now();
        if (lbm_step(SubDomain_.dom, next - iter) != LBM_OK || lbm_sync(SubDomain_.dom) != LBM_OK) {
            error = "lbm_step failed";
            return false;
        }
        seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        iter = next;
        const long long done = iter - 1;  // index of the iteration just completed
        if (output_every > 0 && done % output_every == 0) {
            WriteOutput((int)done, dir);
            double mass = 0.0;
            const int rc = lbm_reduce(SubDomain_.dom, &mass, nullptr);
            std::printf("Iteration %lld has been performed successfully (mass %.12g)\n", done, mass);
            std::fflush(stdout);
            if (rc == LBM_ENAN) {
                error = "non-finite state";
                return false;
            }
        }
    }
    return true;
}
