// One run described by a configuration file (see LBM.cpp for the keys).
#pragma once

#include <string>
#include <vector>

#include "lbm_capi.h"

struct WaveSpec {
    int slot = 0;
    int kind = LBM_WAVE_NONE;
    double period = 1.0, dt = 1.0;
    std::vector<double> a, phi;
};

struct CaseConfig {
    std::string stl_file;
    double spacing = 0.1;                 // Input.h:9
    bool explicit_box = false;
    double box_origin[3] = {0, 0, 0};
    int box_dims[3] = {0, 0, 0};
    std::vector<std::string> plane_names;  // priority order
    std::vector<lbm_plane_desc> planes;
    std::vector<lbm_bc_desc> bcs;           // bcs[p] drives plane p (flag 3 + p)
    std::vector<WaveSpec> waves;
    int q = 19;
    int storage = LBM_STORE_F64;
    int scheme = LBM_SCHEME_TWO_LATTICE;
    int quirks = 1;
    double tau = 0.66;                      // LBM.cpp:59
    long long steps = 20000;                // Simulation.cpp:908
    int output_every = 100;                 // Simulation.cpp:920
    std::string output_dir = ".";
    // VTU files: ASCII data arrays, or one appended block of raw little-endian
    // binary (the VTK XML "appended, raw" encoding; far smaller and faster for
    // domains of millions of nodes)
    bool output_raw = false;
    // write the boundary-point files of ComplexGeometry.cpp:23-31 into output_dir
    bool boundary_points = false;
    int device = 0;
    // lbm_autotune before building the domain: scheme, node order, kernel
    // layout and block size are timed and the fastest is used.  Every choice
    // gives the same results bit for bit; only the speed changes.
    bool autotune = false;
    std::string autotune_cache;  // tuning cache file ("" = none)
};

// Relative stl_file paths are taken relative to the configuration file's
// directory, so a case directory can be moved as a whole.
bool load_case(const std::string& path, CaseConfig& out, std::string& err);
