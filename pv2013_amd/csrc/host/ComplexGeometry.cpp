#include "ComplexGeometry.h"

#include <sys/stat.h>

#include <cstring>
#include <fstream>

Geometry::Geometry(const CaseConfig& c, lbm_ctx* x) : cfg(c), ctx(x) {
    std::memset(&grid, 0, sizeof grid);
    GetBoundingbox();
    if (ok) labelPoints();
    if (ok) classify();
}

void Geometry::GetBoundingbox() {
    if (!cfg.stl_file.empty()) {
        triangles = reader.readSTL(cfg.stl_file);
        if (triangles.empty()) {
            ok = false;
            error = "no facets read from " + cfg.stl_file;
            return;
        }
    }
    if (cfg.explicit_box) {
        // lattice box given in the configuration; the ray spans that box
        for (int a = 0; a < 3; ++a) {
            grid.origin[a] = cfg.box_origin[a];
            grid.dims[a] = cfg.box_dims[a];
            grid.raybox[2 * a] = cfg.box_origin[a];
            grid.raybox[2 * a + 1] = cfg.box_origin[a] + (cfg.box_dims[a] - 1) * cfg.spacing;
        }
        grid.spacing = cfg.spacing;
        grid.ray_offset = 0.01;
        grid.ray_pad = 0.1;
        grid.t_eps = 1e-7;
    } else if (lbm_grid_from_bbox(reinterpret_cast<const double*>(triangles.data()), (int64_t)triangles.size(),
                                  cfg.spacing, &grid) != LBM_OK) {
        ok = false;
        error = "lbm_grid_from_bbox failed";
        return;
    }
    for (int a = 0; a < 3; ++a) GeomSize[a] = grid.dims[a];
}

void Geometry::labelPoints() {
    flags.assign((size_t)GeomSize[0] * GeomSize[1] * GeomSize[2], 0);
    if (lbm_voxelize(ctx, reinterpret_cast<const double*>(triangles.data()), (int64_t)triangles.size(), &grid,
                     flags.data(), &voxelize_ms) != LBM_OK) {
        ok = false;
        error = lbm_last_error(ctx);
    }
}

void Geometry::classify() {
    if (lbm_classify(ctx, &grid, cfg.planes.data(), (int)cfg.planes.size(), flags.data(), &classify_ms) != LBM_OK) {
        ok = false;
        error = lbm_last_error(ctx);
    }
}

bool Geometry::WriteBoundaryPoints(const std::string& dir, std::string& err) const {
    mkdir(dir.c_str(), 0777);
    const int nx = GeomSize[0], ny = GeomSize[1], nz = GeomSize[2];
    const size_t nfiles = cfg.plane_names.size() + 1;  // planes, then the walls
    std::vector<std::vector<int64_t>> sites(nfiles);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t s = i + (int64_t)nx * (j + (int64_t)ny * k);
                const int f = flags[s];
                if (f == 2) sites[nfiles - 1].push_back(s);
                else if (f >= 3 && (size_t)(f - 3) < nfiles - 1) sites[f - 3].push_back(s);
            }
    for (size_t p = 0; p < nfiles; ++p) {
        const std::string name =
            (p + 1 < nfiles ? cfg.plane_names[p] : std::string("Wall")) + "boundaryPointsFromGeometry.vtk";
        std::ofstream out(dir + "/" + name);
        if (!out) {
            err = "cannot write " + dir + "/" + name;
            return false;
        }
        const std::vector<int64_t>& v = sites[p];
        out << "# vtk DataFile Version 2.0\nPoints Inside Complex Geometry\nASCII\nDATASET UNSTRUCTURED_GRID\n";
        out << "POINTS " << v.size() << " double\n";
        for (int64_t s : v) {
            const int64_t idx[3] = {s % nx, (s / nx) % ny, s / ((int64_t)nx * ny)};
            out << grid.origin[0] + idx[0] * grid.spacing << " " << grid.origin[1] + idx[1] * grid.spacing << " "
                << grid.origin[2] + idx[2] * grid.spacing << "\n";
        }
        out << "CELLS " << v.size() << " " << v.size() * 2 << "\n";
        for (size_t i = 0; i < v.size(); ++i) out << "1 " << i << "\n";
        out << "CELL_TYPES " << v.size() << "\n";
        for (size_t i = 0; i < v.size(); ++i) out << "1\n";
        if (!out) {
            err = "write failed: " + dir + "/" + name;
            return false;
        }
    }
    return true;
}
