// lbm_run: the reference's main() (LBM.cpp:20-84) driven by a configuration
// file instead of compiled-in constants.
//
//   lbm_run case.cfg
//
// Configuration keys ("key value" lines, '#' comments):
//   stl_file <path>            STL mesh (ASCII or binary)
//   spacing <h>                grid spacing in STL units (Input.h:9)
//   box_origin <x> <y> <z>     optional explicit lattice box ...
//   box_dims <nx> <ny> <nz>    ... instead of the STL bounding box
//   plane <name> <x|y|z> <value> [tol <t>] [range <x|y|z> <lo> <hi>]...
//                              inlet/outlet planes in priority order
//                              (ComplexGeometry.cpp:579-633); the value is
//                              compared in long double as the reference does
//   bc <plane> velocity <sigma> <ux> <uy> <uz> [parabolic <c1> <c2> <R>] [waveform <slot>]
//   bc <plane> pressure <sigma> <rho> [typo]
//                              Zou-He condition of a plane; sigma is the sign
//                              of c[axis] of the incoming populations
//   waveform <slot> fourier <period> <a1> <phi1> [<a2> <phi2> ...]
//   waveform <slot> pv2013 <dt>
//   lattice 19|27   storage f64|f32|f16s   scheme two_lattice|aa   quirks 1|0
//   tau <tau>   steps <n>   output_every <n>   output_dir <dir>   device <id>
//   output_format ascii|raw    VTU data arrays as text or appended raw binary
//   boundary_points 1          also write the boundary-point files of
//                              ComplexGeometry.cpp:23-31 (<plane>boundaryPoints
//                              FromGeometry.vtk per plane, WallboundaryPoints
//                              FromGeometry.vtk)
//   autotune 1   autotune_cache <file>
// A relative stl_file is read from the configuration file's directory.
#include <chrono>
#include <cstdio>

#include "Case.h"
#include "ComplexGeometry.h"
#include "Simulation.h"
#include "Subdomain.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <case.cfg>\n", argv[0]);
        return 2;
    }
    const auto start = std::chrono::steady_clock::now();
    CaseConfig cfg;
    std::string err;
    if (!load_case(argv[1], cfg, err)) {
        std::fprintf(stderr, "config error: %s\n", err.c_str());
        return 2;
    }
    lbm_ctx* ctx = nullptr;
    if (lbm_ctx_create(cfg.device, &ctx) != LBM_OK) {
        std::fprintf(stderr, "no GPU device %d\n", cfg.device);
        return 1;
    }
    int rc = 0;
    {
        Geometry geometry(cfg, ctx);
        if (!geometry.ok) {
            std::fprintf(stderr, "geometry: %s\n", geometry.error.c_str());
            lbm_ctx_destroy(ctx);
            return 1;
        }
        std::printf("grid %d x %d x %d, %zu facets, voxelize %.3f ms, classify %.3f ms\n", geometry.GeomSize[0],
                    geometry.GeomSize[1], geometry.GeomSize[2], geometry.triangles.size(), geometry.voxelize_ms,
                    geometry.classify_ms);
        if (cfg.boundary_points) {
            std::string werr;
            if (!geometry.WriteBoundaryPoints(cfg.output_dir, werr)) {
                std::fprintf(stderr, "geometry: %s\n", werr.c_str());
                lbm_ctx_destroy(ctx);
                return 1;
            }
        }
        SubDomain subdomain(geometry, cfg, ctx);
        if (!subdomain.ok) {
            std::fprintf(stderr, "domain: %s\n", subdomain.error.c_str());
            lbm_ctx_destroy(ctx);
            return 1;
        }
        std::printf("fluid nodes %lld, boundary records %lld\n", (long long)subdomain.fluid_nodes,
                    (long long)subdomain.boundary_records);
        if (subdomain.tuned.ntried >= 0)
            std::printf("autotune: scheme %d, fast axis %d, options %d, block %d (%d candidates, %.0f MLUPS)\n",
                        subdomain.tuned.scheme, subdomain.tuned.fast_axis, subdomain.tuned.options,
                        subdomain.tuned.block, subdomain.tuned.ntried, subdomain.tuned.mlups);
        Simulation sim(subdomain, cfg.tau, 1.0, 0.5e-5);
        if (!sim.Iterate(cfg.steps, cfg.output_every, cfg.output_dir, cfg.output_raw)) {
            std::fprintf(stderr, "simulation: %s\n", sim.error.c_str());
            rc = 1;
        } else {
            std::printf("%lld steps in %.3f s: %.1f MLUPS\n", cfg.steps, sim.seconds,
                        sim.seconds > 0 ? subdomain.fluid_nodes * (double)cfg.steps / sim.seconds / 1e6 : 0.0);
        }
    }
    lbm_ctx_destroy(ctx);
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::printf("Elapsed time: %.2f seconds.\n", el);  // LBM.cpp:78-81
    return rc;
}
