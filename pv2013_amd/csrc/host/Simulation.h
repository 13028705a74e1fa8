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
    // stress are written.  Returns false on an engine error.
    bool Iterate(long long steps, int output_every, const std::string& output_dir);
    void WriteOutput(int iter, const std::string& output_dir);
    double seconds = 0.0;  // wall time of the stepping (outputs excluded)
    std::string error;

private:
    SubDomain& SubDomain_;
    double m_relaxation, m_omega, pressure_, m_dt;
};
