// OutputFile (OutputFile.h:48-53 interface): VTK XML unstructured-grid files
// written without the VTK library (ASCII .vtu, VTK_VERTEX cells).
#pragma once

#include <string>

#include "Subdomain.h"

class OutputFile {
public:
    // raw: data arrays in one appended raw-binary block instead of text
    OutputFile(int outputIdx, std::string dir = ".", bool raw = false);
    // Fluid_Variables/<idx>_rank_0.vtu at every fluid site (OutputFile.cpp:16-87):
    // Velocity, VelocityMagnitude and Density from the node fields the last
    // collision set (SubDomain::density / velocity), Pressure = c_s^2 Density()
    // of the current state (Node.cpp:112-118; 1/3 at inlet nodes, whose
    // Boundary::Density() is 1, Boundary.cpp:19-22)
    bool WriteVTK_LocalData(SubDomain& subdomain);
    // Wall_Shear_Stress_VTK/<iter>_rank_0_wall_shear_stress.vtu
    // (OutputFile.cpp:90-137): SubDomain::wall_shear_stress in wall-list order
    bool WriteWallShearStressVTK(SubDomain& subdomain, int iter);

private:
    int m_outputIdx;
    std::string dir_;
    bool raw_;
};
