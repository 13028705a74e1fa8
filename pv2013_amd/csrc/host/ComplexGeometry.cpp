#include "ComplexGeometry.h"

#include <cstring>

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
