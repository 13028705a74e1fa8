// Geometry (ComplexGeometry.h:58 interface): STL -> grid -> voxel flags ->
// boundary classification, on the GPU through lbm_voxelize / lbm_classify.
#pragma once

#include <string>
#include <vector>

#include "Case.h"
#include "GeometryReader.h"
#include "lbm_capi.h"

class Geometry {
public:
    Geometry(const CaseConfig& cfg, lbm_ctx* ctx);
    // ComplexGeometry.cpp:7-31 pipeline with the reference's sizing rules
    void GetBoundingbox();
    void labelPoints();
    void classify();
    // The reference's boundary-point files (ComplexGeometry.cpp:23-31, format of
    // Writer::writePointsToFile, GeometryWriter.cpp:35-68): one legacy VTK file
    // per inlet/outlet plane, <name>boundaryPointsFromGeometry.vtk, and
    // WallboundaryPointsFromGeometry.vtk, points in i-j-k order (i outermost)
    // at the grid coordinates origin + index * spacing.  A point lies in the file
    // of its first matching plane (the node priority of Subdomain.cpp:413-473).
    bool WriteBoundaryPoints(const std::string& dir, std::string& err) const;

    ::Reader reader;
    std::vector<Triangle> triangles;
    lbm_grid_desc grid;
    int GeomSize[3] = {0, 0, 0};
    std::vector<uint8_t> flags;  // 0 outside, 1 fluid, 2 wall, 3 + plane
    double voxelize_ms = 0.0, classify_ms = 0.0;
    const CaseConfig& cfg;
    lbm_ctx* ctx;
    bool ok = true;
    std::string error;
};
