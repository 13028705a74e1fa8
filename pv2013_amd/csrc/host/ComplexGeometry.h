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
