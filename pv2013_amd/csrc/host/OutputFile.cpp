#include "OutputFile.h"

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

OutputFile::OutputFile(int idx, std::string dir, bool raw) : m_outputIdx(idx), dir_(std::move(dir)), raw_(raw) {}

namespace {

// One VTK XML unstructured grid of vertex cells.  Arrays are written inline as
// text, or as one appended block of raw little-endian binary (each array a
// UInt64 byte count followed by its bytes; the DataArray's offset is the
// position of that count after the block's leading '_').
class Vtu {
public:
    explicit Vtu(bool raw) : raw_(raw) {}

    void points(const std::vector<int64_t>& sites, int nx, int ny) {
        std::vector<double> xyz(3 * sites.size());
        for (size_t i = 0; i < sites.size(); ++i) {
            const int64_t s = sites[i];
            xyz[3 * i] = (double)(s % nx);
            xyz[3 * i + 1] = (double)((s / nx) % ny);
            xyz[3 * i + 2] = (double)(s / ((int64_t)nx * ny));
        }
        npts_ = sites.size();
        points_ = add("", "Float64", 3, xyz.data(), xyz.size() * 8);
        std::vector<int64_t> conn(npts_), offs(npts_);
        for (size_t i = 0; i < npts_; ++i) {
            conn[i] = (int64_t)i;
            offs[i] = (int64_t)i + 1;
        }
        std::vector<uint8_t> types(npts_, 1);  // VTK_VERTEX
        conn_ = add("connectivity", "Int64", 1, conn.data(), conn.size() * 8);
        offs_ = add("offsets", "Int64", 1, offs.data(), offs.size() * 8);
        types_ = add("types", "UInt8", 1, types.data(), types.size());
    }
    void point_data(const char* name, int ncomp, const std::vector<double>& v) {
        data_.push_back(add(name, "Float64", ncomp, v.data(), v.size() * 8));
    }
    bool write(const std::string& path) const {
        FILE* fp = std::fopen(path.c_str(), "wb");
        if (!fp) return false;
        std::fprintf(fp, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"%s\" "
                         "byte_order=\"LittleEndian\"%s>\n  <UnstructuredGrid>\n"
                         "    <Piece NumberOfPoints=\"%zu\" NumberOfCells=\"%zu\">\n",
                     raw_ ? "1.0" : "0.1", raw_ ? " header_type=\"UInt64\"" : "", npts_, npts_);
        std::fprintf(fp, "      <Points>\n");
        array(fp, points_);
        std::fprintf(fp, "      </Points>\n      <Cells>\n");
        array(fp, conn_);
        array(fp, offs_);
        array(fp, types_);
        std::fprintf(fp, "      </Cells>\n      <PointData>\n");
        for (size_t a : data_) array(fp, a);
        std::fprintf(fp, "      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n");
        if (raw_) {
            std::fprintf(fp, "  <AppendedData encoding=\"raw\">\n_");
            for (const Arr& a : arrs_) {
                const uint64_t n = a.bytes.size();
                std::fwrite(&n, 8, 1, fp);
                if (n) std::fwrite(a.bytes.data(), 1, n, fp);
            }
            std::fprintf(fp, "\n  </AppendedData>\n");
        }
        std::fprintf(fp, "</VTKFile>\n");
        return std::fclose(fp) == 0;
    }

private:
    struct Arr {
        std::string name, type;
        int ncomp;
        std::vector<char> bytes;
        uint64_t offset;
    };
    size_t add(const char* name, const char* type, int ncomp, const void* p, size_t nbytes) {
        Arr a{name, type, ncomp, std::vector<char>((const char*)p, (const char*)p + nbytes), next_};
        next_ += 8 + nbytes;
        arrs_.push_back(std::move(a));
        return arrs_.size() - 1;
    }
    void array(FILE* fp, size_t i) const {
        const Arr& a = arrs_[i];
        std::fprintf(fp, "        <DataArray type=\"%s\"", a.type.c_str());
        if (!a.name.empty()) std::fprintf(fp, " Name=\"%s\"", a.name.c_str());
        if (a.ncomp > 1) std::fprintf(fp, " NumberOfComponents=\"%d\"", a.ncomp);
        if (raw_) {
            std::fprintf(fp, " format=\"appended\" offset=\"%llu\"/>\n", (unsigned long long)a.offset);
            return;
        }
        std::fprintf(fp, " format=\"ascii\">\n");
        const size_t n = a.bytes.size();
        if (a.type == "Float64") {
            const double* v = (const double*)a.bytes.data();
            const size_t m = n / 8;
            for (size_t k = 0; k < m; k += a.ncomp)
                for (int c = 0; c < a.ncomp; ++c) std::fprintf(fp, c + 1 < a.ncomp ? "%.17g " : "%.17g\n", v[k + c]);
        } else if (a.type == "Int64") {
            const int64_t* v = (const int64_t*)a.bytes.data();
            for (size_t k = 0; k < n / 8; ++k) std::fprintf(fp, "%lld\n", (long long)v[k]);
        } else {
            for (size_t k = 0; k < n; ++k) std::fprintf(fp, "%u\n", (unsigned)(uint8_t)a.bytes[k]);
        }
        std::fprintf(fp, "        </DataArray>\n");
    }
    bool raw_;
    size_t npts_ = 0, points_ = 0, conn_ = 0, offs_ = 0, types_ = 0;
    std::vector<size_t> data_;
    std::vector<Arr> arrs_;
    uint64_t next_ = 0;
};

}  // namespace

bool OutputFile::WriteVTK_LocalData(SubDomain& sd) {
    const int nx = sd.my_Nx_, ny = sd.my_Ny_, nz = sd.my_Nz_;
    const int64_t ns = (int64_t)nx * ny * nz;
    if ((int64_t)sd.density.size() != ns || (int64_t)sd.velocity.size() != 3 * ns) return false;
    // Node::Density() of the current state, for the pressure (Node.cpp:39-47)
    std::vector<double> rho_now((size_t)ns);
    if (lbm_get_macros(sd.dom, rho_now.data(), nullptr) != LBM_OK) return false;
    const std::string folder = dir_ + "/Fluid_Variables";
    mkdir(dir_.c_str(), 0777);
    mkdir(folder.c_str(), 0777);
    // z, y, x loop order as OutputFile.cpp:45-47
    std::vector<int64_t> sites;
    const std::vector<uint8_t>& fl = sd.Geometry_.flags;
    for (int64_t s = 0; s < ns; ++s)
        if (fl[s]) sites.push_back(s);
    const size_t m = sites.size();
    std::vector<double> vel(3 * m), pres(m), vmag(m), dens(m);
    const std::vector<lbm_bc_desc>& bcs = sd.Geometry_.cfg.bcs;
    const std::vector<double>& u = sd.velocity;
    for (size_t i = 0; i < m; ++i) {
        const int64_t s = sites[i];
        vel[3 * i] = u[s];
        vel[3 * i + 1] = u[ns + s];
        vel[3 * i + 2] = u[2 * ns + s];
        vmag[i] = std::sqrt(u[s] * u[s] + u[ns + s] * u[ns + s] + u[2 * ns + s] * u[2 * ns + s]);
        dens[i] = sd.density[s];
        // Node::Pressure = c_s^2 Density() (Node.cpp:112-118); inlet nodes are
        // Boundary objects whose Density() is 1 (Boundary.cpp:19-22)
        const int f = fl[s];
        const bool inlet_quirk = sd.Geometry_.cfg.quirks && f >= 3 && bcs[f - 3].kind == LBM_BC_VELOCITY;
        pres[i] = (1.0 / 3.0) * (inlet_quirk ? 1.0 : rho_now[s]);
    }
    Vtu v(raw_);
    v.points(sites, nx, ny);
    v.point_data("Velocity", 3, vel);
    v.point_data("Pressure", 1, pres);
    v.point_data("VelocityMagnitude", 1, vmag);
    v.point_data("Density", 1, dens);
    return v.write(folder + "/" + std::to_string(m_outputIdx) + "_rank_0.vtu");
}

bool OutputFile::WriteWallShearStressVTK(SubDomain& sd, int iter) {
    const int nx = sd.my_Nx_, ny = sd.my_Ny_, nz = sd.my_Nz_;
    if ((int64_t)sd.wall_shear_stress.size() != (int64_t)nx * ny * nz) return false;
    const std::string folder = dir_ + "/Wall_Shear_Stress_VTK";
    mkdir(dir_.c_str(), 0777);
    mkdir(folder.c_str(), 0777);
    std::vector<double> w(sd.wall_sites.size());
    for (size_t i = 0; i < w.size(); ++i) w[i] = sd.wall_shear_stress[sd.wall_sites[i]];
    Vtu v(raw_);
    v.points(sd.wall_sites, nx, ny);
    v.point_data("WallShearStress", 1, w);
    return v.write(folder + "/" + std::to_string(iter) + "_rank_0_wall_shear_stress.vtu");
}
