#include "OutputFile.h"

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <vector>

OutputFile::OutputFile(int idx, std::string dir) : m_outputIdx(idx), dir_(std::move(dir)) {}

static void write_points_cells(FILE* fp, const std::vector<int64_t>& sites, int nx, int ny) {
    std::fprintf(fp, "      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (int64_t s : sites)
        std::fprintf(fp, "%lld %lld %lld\n", (long long)(s % nx), (long long)((s / nx) % ny),
                     (long long)(s / ((int64_t)nx * ny)));
    std::fprintf(fp, "        </DataArray>\n      </Points>\n      <Cells>\n");
    std::fprintf(fp, "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
    for (size_t i = 0; i < sites.size(); ++i) std::fprintf(fp, "%zu\n", i);
    std::fprintf(fp, "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
    for (size_t i = 0; i < sites.size(); ++i) std::fprintf(fp, "%zu\n", i + 1);
    std::fprintf(fp, "        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
    for (size_t i = 0; i < sites.size(); ++i) std::fprintf(fp, "1\n");
    std::fprintf(fp, "        </DataArray>\n      </Cells>\n");
}

static FILE* open_vtu(const std::string& path, size_t npts) {
    FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return nullptr;
    std::fprintf(fp, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
                     "  <UnstructuredGrid>\n    <Piece NumberOfPoints=\"%zu\" NumberOfCells=\"%zu\">\n", npts, npts);
    return fp;
}

static void close_vtu(FILE* fp) {
    std::fprintf(fp, "    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
    std::fclose(fp);
}

bool OutputFile::WriteVTK_LocalData(SubDomain& sd) {
    const int nx = sd.my_Nx_, ny = sd.my_Ny_, nz = sd.my_Nz_;
    const int64_t ns = (int64_t)nx * ny * nz;
    std::vector<double> rho((size_t)ns), u((size_t)3 * ns);
    if (lbm_get_macros(sd.dom, rho.data(), u.data()) != LBM_OK) return false;
    const std::string folder = dir_ + "/Fluid_Variables";
    mkdir(dir_.c_str(), 0777);
    mkdir(folder.c_str(), 0777);
    // z, y, x loop order as OutputFile.cpp:45-47
    std::vector<int64_t> sites;
    const std::vector<uint8_t>& fl = sd.Geometry_.flags;
    for (int64_t s = 0; s < ns; ++s)
        if (fl[s]) sites.push_back(s);
    FILE* fp = open_vtu(folder + "/" + std::to_string(m_outputIdx) + "_rank_0.vtu", sites.size());
    if (!fp) return false;
    write_points_cells(fp, sites, nx, ny);
    std::fprintf(fp, "      <PointData>\n        <DataArray type=\"Float64\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (int64_t s : sites) std::fprintf(fp, "%.17g %.17g %.17g\n", u[s], u[ns + s], u[2 * ns + s]);
    std::fprintf(fp, "        </DataArray>\n        <DataArray type=\"Float64\" Name=\"Pressure\" format=\"ascii\">\n");
    // Node::Pressure = c_s^2 * Density() (Node.cpp:112-118); the inlet's
    // Boundary::Density() returns 1 (Boundary.cpp:19-22) in the reference
    const std::vector<lbm_bc_desc>& bcs = sd.Geometry_.cfg.bcs;
    for (int64_t s : sites) {
        const int f = fl[s];
        const bool inlet_quirk = sd.Geometry_.cfg.quirks && f >= 3 && bcs[f - 3].kind == LBM_BC_VELOCITY;
        std::fprintf(fp, "%.17g\n", (1.0 / 3.0) * (inlet_quirk ? 1.0 : rho[s]));
    }
    std::fprintf(fp, "        </DataArray>\n        <DataArray type=\"Float64\" Name=\"VelocityMagnitude\" format=\"ascii\">\n");
    for (int64_t s : sites)
        std::fprintf(fp, "%.17g\n", std::sqrt(u[s] * u[s] + u[ns + s] * u[ns + s] + u[2 * ns + s] * u[2 * ns + s]));
    std::fprintf(fp, "        </DataArray>\n        <DataArray type=\"Float64\" Name=\"Density\" format=\"ascii\">\n");
    for (int64_t s : sites) std::fprintf(fp, "%.17g\n", rho[s]);
    std::fprintf(fp, "        </DataArray>\n      </PointData>\n");
    close_vtu(fp);
    return true;
}

bool OutputFile::WriteWallShearStressVTK(SubDomain& sd, int iter) {
    const int nx = sd.my_Nx_, ny = sd.my_Ny_, nz = sd.my_Nz_;
    std::vector<double> wss((size_t)nx * ny * nz);
    if (lbm_get_wss(sd.dom, wss.data()) != LBM_OK) return false;
    const std::string folder = dir_ + "/Wall_Shear_Stress_VTK";
    mkdir(dir_.c_str(), 0777);
    mkdir(folder.c_str(), 0777);
    FILE* fp = open_vtu(folder + "/" + std::to_string(iter) + "_rank_0_wall_shear_stress.vtu", sd.wall_sites.size());
    if (!fp) return false;
    write_points_cells(fp, sd.wall_sites, nx, ny);
    std::fprintf(fp, "      <PointData>\n        <DataArray type=\"Float64\" Name=\"WallShearStress\" format=\"ascii\">\n");
    for (int64_t s : sd.wall_sites) std::fprintf(fp, "%.17g\n", wss[s]);
    std::fprintf(fp, "        </DataArray>\n      </PointData>\n");
    close_vtu(fp);
    return true;
}
