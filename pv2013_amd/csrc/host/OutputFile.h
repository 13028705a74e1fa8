// OutputFile (OutputFile.h:48-53 interface): VTK XML unstructured-grid files
// written without the VTK library (ASCII .vtu, VTK_VERTEX cells).
#pragma once

#include <string>

#include "Subdomain.h"

class OutputFile {
public:
    OutputFile(int outputIdx, std::string dir = ".");
    // Fluid_Variables/<idx>_rank_0.vtu: Velocity, Pressure, VelocityMagnitude,
    // Density at every fluid site (OutputFile.cpp:16-87)
    bool WriteVTK_LocalData(SubDomain& subdomain);
    // Wall_Shear_Stress_VTK/<iter>_rank_0_wall_shear_stress.vtu (OutputFile.cpp:90-137)
    bool WriteWallShearStressVTK(SubDomain& subdomain, int iter);

private:
    int m_outputIdx;
    std::string dir_;
};
