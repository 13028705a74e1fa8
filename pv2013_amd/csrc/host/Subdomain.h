// SubDomain (Subdomain.h:68 interface): the fluid-node lattice of a Geometry,
// held on the GPU (lbm_domain_create).  One GPU holds the whole domain, so the
// reference's cartesian decomposition and halo layer have no counterpart.
#pragma once

#include <vector>

#include "ComplexGeometry.h"

class SubDomain {
public:
    SubDomain(Geometry& geometry, const CaseConfig& cfg, lbm_ctx* ctx);
    ~SubDomain();
    SubDomain(const SubDomain&) = delete;
    SubDomain& operator=(const SubDomain&) = delete;

    Geometry& Geometry_;
    lbm_domain* dom = nullptr;
    int my_Nx_ = 0, my_Ny_ = 0, my_Nz_ = 0;
    int64_t fluid_nodes = 0, boundary_records = 0;
    // i-j-k ordered wall sites (i outermost), the WallBoundaryNode_ order of
    // Subdomain.cpp:464-469
    std::vector<int64_t> wall_sites;
    // The node fields the reference's Collision sets (Simulation.cpp:21-179:
    // Node::density, Node::velocity, the moments of the state the collision
    // starts from) and ComputeWallShearStress (Simulation.cpp:1406-1421) sets
    // from them; dense box arrays, filled by Simulation::CollectNodeFields.
    std::vector<double> density, velocity, wall_shear_stress;
    bool ok = true;
    std::string error;
    // the autotuner's choice when the configuration asked for one (ntried = -1: not tuned)
    lbm_tune_result tuned = {0, -1, 0, -1, 0.0, 0, 0, 0};
};
