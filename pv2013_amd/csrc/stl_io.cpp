// STL input and the reference's grid sizing (host only).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>

#include "lbm_capi.h"

// Reader::readSTLASCII (GeometryReader.cpp:33-117): line-oriented; a facet is
// kept when "facet normal", "outer loop", three vertex lines, "endloop" and
// "endfacet" all parse.  One facet record is reused, so a short vertex block
// keeps the previous facet's remaining vertices, as in the reference.
extern "C" int64_t lbm_read_stl_ascii(const char* path, double* tri, int64_t cap) {
    if (!path) return LBM_EINVAL;
    std::ifstream in(path);
    if (!in.is_open()) return LBM_EIO;
    std::string line, w1, w2;
    double cur[12] = {0};
    int64_t n = 0;
    auto next = [&](std::istringstream& ss) -> bool {
        if (!std::getline(in, line)) return false;
        ss.clear();
        ss.str(line);
        return true;
    };
    std::istringstream ss;
    while (std::getline(in, line)) {
        ss.clear();
        ss.str(line);
        double nx, ny, nz;
        if (!(ss >> w1 >> w2 >> nx >> ny >> nz) || w1 != "facet" || w2 != "normal") continue;
        cur[9] = nx;
        cur[10] = ny;
        cur[11] = nz;
        if (!next(ss)) break;
        if (!(ss >> w1 >> w2) || w1 != "outer" || w2 != "loop") continue;
        for (int v = 0; v < 3; ++v) {
            if (!next(ss)) break;
            double x, y, z;
            if (!(ss >> w1 >> x >> y >> z) || w1 != "vertex") break;
            cur[3 * v] = x;
            cur[3 * v + 1] = y;
            cur[3 * v + 2] = z;
        }
        if (!next(ss)) break;
        if (!(ss >> w1) || w1 != "endloop") continue;
        if (!next(ss)) break;
        if (!(ss >> w1) || w1 != "endfacet") continue;
        if (tri && n < cap) std::memcpy(tri + 12 * n, cur, sizeof cur);
        ++n;
    }
    return n;
}

// Reader::readSTLBinary (GeometryReader.cpp:119-176): float records, the
// normal renormalised in float.
extern "C" int64_t lbm_read_stl_binary(const char* path, double* tri, int64_t cap) {
    if (!path) return LBM_EINVAL;
    std::ifstream in(path, std::ios::binary);
    if (!in) return LBM_EIO;
    char header[80];
    uint32_t count = 0;
    if (!in.read(header, 80) || !in.read(reinterpret_cast<char*>(&count), 4)) return LBM_EIO;
    for (uint32_t i = 0; i < count; ++i) {
        float v[12];
        uint16_t attr;
        if (!in.read(reinterpret_cast<char*>(v), sizeof v) || !in.read(reinterpret_cast<char*>(&attr), 2))
            return LBM_EIO;
        float nx = v[0], ny = v[1], nz = v[2];
        const float norm = std::sqrt(nx * nx + ny * ny + nz * nz);
        nx /= norm;
        ny /= norm;
        nz /= norm;
        if (tri && (int64_t)i < cap) {
            double* T = tri + 12 * (int64_t)i;
            for (int k = 0; k < 9; ++k) T[k] = v[3 + k];
            T[9] = nx;
            T[10] = ny;
            T[11] = nz;
        }
    }
    return (int64_t)count;
}

// Geometry::GetBoundingbox (ComplexGeometry.cpp:40-73) and the grid of
// generatePointsInBoundingBox (:178-207).
extern "C" int lbm_grid_from_bbox(const double* tri, int64_t ntri, double spacing, lbm_grid_desc* g) {
    if (!g || (ntri > 0 && !tri) || ntri <= 0 || !(spacing > 0)) return LBM_EINVAL;
    double b[6] = {std::numeric_limits<double>::infinity(), -std::numeric_limits<double>::infinity(),
                   std::numeric_limits<double>::infinity(), -std::numeric_limits<double>::infinity(),
                   std::numeric_limits<double>::infinity(), -std::numeric_limits<double>::infinity()};
    for (int64_t t = 0; t < ntri; ++t) {
        const double* T = tri + 12 * t;
        for (int a = 0; a < 3; ++a) {
            b[2 * a] = std::fmin(b[2 * a], std::fmin(T[a], std::fmin(T[3 + a], T[6 + a])));
            b[2 * a + 1] = std::fmax(b[2 * a + 1], std::fmax(T[a], std::fmax(T[3 + a], T[6 + a])));
        }
    }
    std::memset(g, 0, sizeof *g);
    for (int a = 0; a < 3; ++a) {
        g->origin[a] = b[2 * a];
        const double cells = std::ceil((b[2 * a + 1] - b[2 * a]) / spacing);
        if (!(cells < 2.0e9)) return LBM_ERANGE;
        g->dims[a] = (int32_t)cells + 1;
        g->raybox[2 * a] = b[2 * a];
        g->raybox[2 * a + 1] = b[2 * a + 1];
    }
    g->spacing = spacing;
    g->ray_offset = 0.01;
    g->ray_pad = 0.1;
    g->t_eps = 1e-7;
    return LBM_OK;
}
