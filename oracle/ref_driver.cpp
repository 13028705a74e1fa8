// Test-only driver around the reference's own sources (compiled from
// /root/reference by oracle/build_ref.sh into oracle/_ref/).  It exercises
//   Boundary::ApplyBc (Boundary.cpp:127-148, D3Q19 Zou-He inlet)
//   LatticeModel (LatticeModel.cpp:51-151): weights, directions, opposite
//   directions and c_s^2 of the D3Q19 set the reference compiles
//   Reader::readSTLASCII / readSTLBinary (GeometryReader.cpp:33-176)
// The reference's voxelizer (ComplexGeometry.cpp) is not built: it includes
// mpi_call.h, whose member template at :143 redeclares the class template's
// parameter T, which gcc and clang reject (icc accepted it).
// Usage:
//   ref_driver zouhe <in.bin> <out.bin>
//       reads n then n*19 doubles, applies Boundary::ApplyBc to each node
//   ref_driver lattice <out.bin>
//       writes q, then per direction the weight (double) and cx, cy, cz,
//       opposite (int32), then c_s^2 (double)
//   ref_driver stl ascii|binary <in.stl> <out.bin>
//       writes n, then per facet v1, v2, v3 and the normal (12 doubles)
#include <cstdio>
#include <cstdint>
#include <fstream>
#include <memory>
#include <vector>
#include <string>
#include "LatticeModel.h"
#include "Boundary.h"
#include "GeometryReader.h"

static int zouhe(const char* in, const char* out) {
    std::ifstream i(in, std::ios::binary);
    int64_t n = 0;
    i.read((char*)&n, sizeof n);
    std::vector<double> f((size_t)n * 19);
    i.read((char*)f.data(), f.size() * sizeof(double));
    auto velSet = std::make_shared<LatticeModel>();
    for (int64_t m = 0; m < n; ++m) {
        Boundary b(0, 0, 0, velSet, NodeType::Boundary);
        b.BoundaryType_ = BoundaryType::Inlet;
        b.m_distributions.assign(f.begin() + m * 19, f.begin() + (m + 1) * 19);
        b.ApplyBc();
        std::copy(b.m_distributions.begin(), b.m_distributions.end(), f.begin() + m * 19);
    }
    std::ofstream o(out, std::ios::binary);
    o.write((const char*)f.data(), f.size() * sizeof(double));
    return 0;
}

static int lattice(const char* out) {
    LatticeModel m;
    std::ofstream o(out, std::ios::binary);
    const int32_t q = m.Get_nDirections();
    o.write((const char*)&q, sizeof q);
    for (int d = 0; d < q; ++d) {
        const double w = m.GetWeight(d);
        const std::vector<int> c = m.GetDirection(d);
        const int32_t rec[4] = {c[0], c[1], c[2], m.OppositeDirection(d)};
        o.write((const char*)&w, sizeof w);
        o.write((const char*)rec, sizeof rec);
    }
    const double cs2 = m.Get_c_square(0);
    o.write((const char*)&cs2, sizeof cs2);
    return 0;
}

static int stl(const std::string& kind, const char* in, const char* out) {
    Reader r;
    const std::vector<Triangle> t = kind == "ascii" ? r.readSTLASCII(in) : r.readSTLBinary(in);
    std::ofstream o(out, std::ios::binary);
    const int64_t n = (int64_t)t.size();
    o.write((const char*)&n, sizeof n);
    for (const Triangle& x : t) {
        const double v[12] = {x.x1, x.y1, x.z1, x.x2, x.y2, x.z2, x.x3, x.y3, x.z3, x.nx, x.ny, x.nz};
        o.write((const char*)v, sizeof v);
    }
    return 0;
}

int main(int argc, char** argv) {
    int rc = 2;
    if (argc == 4 && std::string(argv[1]) == "zouhe") rc = zouhe(argv[2], argv[3]);
    else if (argc == 3 && std::string(argv[1]) == "lattice") rc = lattice(argv[2]);
    else if (argc == 5 && std::string(argv[1]) == "stl") rc = stl(argv[2], argv[3], argv[4]);
    else fprintf(stderr, "usage: ref_driver zouhe <in> <out> | lattice <out> | stl ascii|binary <in> <out>\n");
    return rc;
}
