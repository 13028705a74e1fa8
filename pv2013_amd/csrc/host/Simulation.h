// Simulation (Simulation.h:23-42 interface).  The reference's constructor runs
// the whole time loop (Simulation.cpp:3-13); here Iterate() is explicit and
// each step is one fused kernel launch (lbm_step).
#pragma once

#include "OutputFile.h"
#include "Subdomain.h"

class Simulation {
public:
    Simulation(SubDomain& subdomain, double tau, double pressure, double del_t);
    // Simulation.cpp:903-932: `steps` iterations; after iteration `iter`
    // (0-based) with iter % output_every == 0 the fields and the wall shear
    // stress are written, as the reference writes them after that iteration:
    // Velocity, Density and WSS from the moments iteration `iter` collided
    // (state S_iter), Pressure from the state after it (S_iter+1).  Returns
    // false on an engine or output error.
    bool Iterate(long long steps, int output_every, const std::string& output_dir, bool raw = false);
    void WriteOutput(int iter, const std::string& output_dir, bool raw = false);
    double seconds = 0.0;  // wall time of the stepping (outputs excluded)
    std::string error;

private:
    // SubDomain's node fields from the current state (what the next
    // collision computes before it relaxes)
    bool CollectNodeFields();
    bool Advance(long long nsteps);
    SubDomain& SubDomain_;
    double m_relaxation, m_omega, pressure_, m_dt;
};
