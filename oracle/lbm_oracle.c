/*
 * CPU oracle for the PV2013 lattice-Boltzmann hot path (FP64, plain C).
 *
 * TEST INFRASTRUCTURE ONLY: the checker that the HIP engine is compared with,
 * and the timed CPU baseline.  The product never links or calls this file.
 *
 * Built with gcc -O2 -ffp-contract=off (no -march, no -ffast-math) so every
 * expression rounds exactly as written, in the order the reference writes it.
 *
 * What follows which reference code, and what pins it (parity checks that do
 * not just retype the formula):
 *   orc_lattice          LatticeModel.cpp:51-151, 208-334.  Pinned by the closed
 *                        forms sum w = 1, sum w c = 0, sum w c c = I/3, opp(c) = -c
 *                        (tests/test_oracle_lattice.py) and, for D3Q19, by the
 *                        reference's own LatticeModel built from its sources:
 *                        weights bit for bit, directions and opposites exactly
 *                        (tests/test_oracle_bc.py).
 *   orc_read_stl_ascii   GeometryReader.cpp:33-117.  Pinned by the facet count and
 *                        bbox of PV2013.stl, by the PV2013 voxel fixtures, and bit
 *                        for bit by the reference's own Reader built from its
 *                        sources (tests/test_capi_cpu.py).
 *   orc_bbox / grid      ComplexGeometry.cpp:40-73, 178-207.  Pinned by the
 *                        PV2013 fixtures (every fixture point maps to a grid index).
 *   orc_voxelize         ComplexGeometry.cpp:264-357, 431-506.  Pinned by the
 *                        PV2013 fixtures (260,051 boundary points), by mode 0
 *                        (literal brute force) == modes 1/2 on random meshes,
 *                        and by closed shapes (a box and a sphere voxelised
 *                        against their analytic inside test away from the
 *                        surface).  The reference's labelPoints itself does not
 *                        build here (ComplexGeometry.cpp includes mpi_call.h,
 *                        whose template at :143 gcc and clang reject).
 *   orc_classify         ComplexGeometry.cpp:509-638, Subdomain.cpp:375-478.
 *                        Pinned by the PV2013 inlet/outlet/wall fixtures, and,
 *                        through orc_run, by the reference's eight
 *                        boundary-point files byte for byte (SHA-256 digests in
 *                        tests/golden, tests/test_oracle_driver.py).
 *   collision            Simulation.cpp:21-179.  Pinned by: rest state is a fixed
 *                        point, moments of f_eq equal (rho, rho u), BGK relaxes
 *                        the non-equilibrium part by (1 - omega)
 *                        (tests/test_oracle_lbm.py, test_oracle_bc.py), and by a
 *                        textbook BGK evaluation in numpy on every bounced link
 *                        (tests/test_oracle_pins.py).
 *   streaming            Simulation.cpp:184-198, Node.cpp:16-20 (push with the
 *                        f != 0 guard).  Pinned by exact delta shifts with the
 *                        collision switched off, mass conservation in a closed
 *                        box, and the periodic shear wave decaying at the
 *                        lattice viscosity (tests/test_oracle_lbm.py).
 *   Zou-He inlet         Boundary.cpp:127-148 (D3Q19), :85-123 (D3Q27).  Pinned by
 *                        the reference's Boundary::ApplyBc built from its sources
 *                        (oracle/_ref) on random inputs, and by the moment
 *                        identities (the node ends with exactly the imposed u).
 *   Zou-He outlets       Simulation.cpp:952-1348.  Pinned by the moment identities:
 *                        rho = rho_out and zero tangential u after the update, and
 *                        for the x-face typo the closed-form residual
 *                        rho u_z = f16 - f10 that the f10-for-f16 swap leaves.
 *   bounce-back          Simulation.cpp:1353-1370, Node.cpp:121-125.  Pinned by
 *                        one step from a random state: every missing link d of a
 *                        wall node puts the node's own BGK f_new[d] into
 *                        f[opp d]; with the reference's loop from d = 1 (D3Q19)
 *                        f[1] of a node whose +x neighbour is missing keeps its
 *                        pre-step value bit for bit (tests/test_oracle_pins.py);
 *                        and by a resting closed box staying at rest and
 *                        Poiseuille flow (tests/test_oracle_lbm.py).
 *   stale links          Inlet and outlet nodes get no bounce-back
 *                        (WallBoundaryNode_ only): a population whose source is
 *                        missing and that Zou-He does not rewrite keeps its
 *                        pre-step value bit for bit (tests/test_oracle_pins.py).
 *   WSS                  Simulation.cpp:1374-1421.  Pinned by a linear velocity
 *                        field u = u0 + G x at equilibrium: mu |G m| at every
 *                        wall node, m the mean of c_d over the links whose node
 *                        exists, the rest link included, computed from the box
 *                        geometry in numpy (tests/test_oracle_pins.py).
 *   waveform             Simulation.cpp:875-900, pulsatile.h:12-28 (PV2013 mode);
 *                        the Fourier mode has no reference counterpart.  Both
 *                        pinned against long-double evaluation within a rounding
 *                        bound, with the zero gap in [0.81, 0.9) s, the value
 *                        0.1423 at iteration 0 and the period
 *                        (tests/test_oracle_pins.py).
 *   orc_sim_step         Simulation.cpp:903-932: Collision, Streaming, inlet,
 *                        wall, outlets, in that order, over the node lists.  Each
 *                        part is pinned as above and every node belongs to one
 *                        list, so the list order cannot change a value: the three
 *                        boundary passes run separately (orc_sim_step_passes), in
 *                        the reference's order and in the five others, equal the
 *                        per-node loop bit for bit (tests/test_oracle_pins.py); the whole
 *                        is pinned by mass conservation, the resting box, the
 *                        shear wave and config 1's Poiseuille profile with the
 *                        Hagen-Poiseuille pressure gradient (clean bounce-back).
 *                        No independent multi-step run of the reference's own
 *                        Simulation.cpp exists here (it needs the VTK library
 *                        through OutputFile.h), so many steps of the quirk-laden
 *                        step on a real geometry are pinned only through these
 *                        per-step and closed-form checks.
 */
#include "lbm_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ lattice */

static const int32_t C19[19][3] = {
    {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},   {0, 0, -1}, {1, 1, 0},
    {1, -1, 0}, {1, 0, 1},  {1, 0, -1}, {-1, 1, 0}, {-1, -1, 0}, {-1, 0, 1}, {-1, 0, -1},
    {0, 1, 1},  {0, 1, -1}, {0, -1, 1}, {0, -1, -1}, {0, 0, 0}};
static const int32_t OPP19[19] = {1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18};

static const int32_t C27[27][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0}, {1, 0, 1},   {-1, 0, 1}, {1, 0, -1},
    {-1, 0, -1}, {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 1, 1},  {-1, -1, -1},
    {1, 1, -1},  {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1}};
static const int32_t OPP27[27] = {0,  2,  1,  4,  3,  6,  5,  10, 9,  8,  7,  14, 13, 12,
                                  11, 18, 17, 16, 15, 20, 19, 22, 21, 24, 23, 26, 25};

static double weight(int q, int d) {
    if (q == 19) {
        if (d < 6) return 2.0 / 36.0;
        if (d < 18) return 1.0 / 36.0;
        return 12.0 / 36.0;
    }
    if (d == 0) return 8.0 / 27.0;
    if (d < 7) return 2.0 / 27.0;
    if (d < 19) return 1.0 / 54.0;
    return 1.0 / 216.0;
}

static const int32_t (*dirs(int q))[3] { return q == 19 ? C19 : C27; }
static const int32_t* opps(int q) { return q == 19 ? OPP19 : OPP27; }
static int rest_dir(int q) { return q == 19 ? 18 : 0; }

int orc_lattice(int q, int32_t* c, double* w, int32_t* opp) {
    if (q != 19 && q != 27) return -1;
    const int32_t(*cc)[3] = dirs(q);
    for (int d = 0; d < q; ++d) {
        for (int a = 0; a < 3; ++a) c[3 * d + a] = cc[d][a];
        w[d] = weight(q, d);
        opp[d] = opps(q)[d];
    }
    return 0;
}

/* ---------------------------------------------------------------- STL input */

/* GeometryReader.cpp:33-117: line-oriented parse; a facet is kept only when the
 * "facet normal", "outer loop", 3 vertex lines, "endloop" and "endfacet" lines
 * all parse.  A vertex line that fails leaves the remaining vertices of the
 * previous facet in place (the reference reuses one Triangle object). */
static int next_line(FILE* fp, char* buf, size_t n) { return fgets(buf, (int)n, fp) != NULL; }

int64_t orc_read_stl_ascii(const char* path, double* tri, int64_t cap) {
    FILE* fp = fopen(path, "r");
    if (!fp) return -1;
    char line[1024], w1[64], w2[64];
    double cur[12] = {0};
    int64_t n = 0;
    while (next_line(fp, line, sizeof line)) {
        double nx, ny, nz;
        if (sscanf(line, "%63s %63s %lf %lf %lf", w1, w2, &nx, &ny, &nz) != 5) continue;
        if (strcmp(w1, "facet") || strcmp(w2, "normal")) continue;
        cur[9] = nx; cur[10] = ny; cur[11] = nz;
        if (!next_line(fp, line, sizeof line)) break;
        if (sscanf(line, "%63s %63s", w1, w2) != 2 || strcmp(w1, "outer") || strcmp(w2, "loop"))
            continue;
        for (int v = 0; v < 3; ++v) {
            if (!next_line(fp, line, sizeof line)) break;
            double x, y, z;
            if (sscanf(line, "%63s %lf %lf %lf", w1, &x, &y, &z) != 4 || strcmp(w1, "vertex")) break;
            cur[3 * v] = x; cur[3 * v + 1] = y; cur[3 * v + 2] = z;
        }
        if (!next_line(fp, line, sizeof line)) break;
        if (sscanf(line, "%63s", w1) != 1 || strcmp(w1, "endloop")) continue;
        if (!next_line(fp, line, sizeof line)) break;
        if (sscanf(line, "%63s", w1) != 1 || strcmp(w1, "endfacet")) continue;
        if (tri && n < cap) memcpy(tri + 12 * n, cur, sizeof cur);
        ++n;
    }
    fclose(fp);
    return n;
}

/* ComplexGeometry.cpp:40-60 */
void orc_bbox(const double* tri, int64_t ntri, double* b) {
    b[0] = b[2] = b[4] = INFINITY;
    b[1] = b[3] = b[5] = -INFINITY;
    for (int64_t t = 0; t < ntri; ++t) {
        const double* T = tri + 12 * t;
        for (int a = 0; a < 3; ++a) {
            double lo = fmin(T[a], fmin(T[3 + a], T[6 + a]));
            double hi = fmax(T[a], fmax(T[3 + a], T[6 + a]));
            if (lo < b[2 * a]) b[2 * a] = lo;
            if (hi > b[2 * a + 1]) b[2 * a + 1] = hi;
        }
    }
}

/* ComplexGeometry.cpp:68-72, 178-207: GeomSize = ceil(extent/spacing) + 1; the
 * grid origin is the bbox minimum; the ray uses the same bbox. */
void orc_grid_from_bbox(const double* b, double spacing, orc_grid* g) {
    memset(g, 0, sizeof *g);
    for (int a = 0; a < 3; ++a) {
        g->origin[a] = b[2 * a];
        g->dims[a] = (int32_t)ceil((b[2 * a + 1] - b[2 * a]) / spacing) + 1;
    }
    g->spacing = spacing;
    memcpy(g->raybox, b, 6 * sizeof(double));
    g->ray_offset = 0.01;
    g->ray_pad = 0.1;
    g->t_eps = 1e-7;
}

/* ----------------------------------------------------------------- voxelizer */

/* ComplexGeometry.cpp:264-281: inclusive sign test in the XY projection. */
static int inside_xy(double x, double y, const double* T) {
    double x1 = T[0], y1 = T[1], x2 = T[3], y2 = T[4], x3 = T[6], y3 = T[7];
    double d1 = ((x - x2) * (y1 - y2) - (y - y2) * (x1 - x2));
    double d2 = ((x - x3) * (y2 - y3) - (y - y3) * (x2 - x3));
    double d3 = ((x - x1) * (y3 - y1) - (y - y1) * (x3 - x1));
    int neg = (d1 < 0) || (d2 < 0) || (d3 < 0);
    int pos = (d1 > 0) || (d2 > 0) || (d3 > 0);
    return !(neg && pos);
}

typedef struct { double sx, sy, sz, ex, ey, ez; } ray_t;

/* ComplexGeometry.cpp:441-458 (case 0) and :334-357 (adjustLineCoordinates case 0) */
static ray_t make_ray(const orc_grid* g, int i, int j, int k) {
    ray_t r;
    r.sx = g->origin[0] + i * g->spacing;
    r.sy = g->origin[1] + j * g->spacing;
    r.sz = g->origin[2] + k * g->spacing;
    double L = g->raybox[5] - g->raybox[4];
    r.ex = r.sx;
    r.ey = r.sy;
    r.ez = r.sz + L + g->ray_pad;
    if (r.sx == g->raybox[0]) { r.sx += g->ray_offset; r.ex += g->ray_offset; }
    if (r.sx == g->raybox[1]) { r.sx -= g->ray_offset; r.ex -= g->ray_offset; }
    if (r.sy == g->raybox[2]) { r.sy += g->ray_offset; r.ey += g->ray_offset; }
    if (r.sy == g->raybox[3]) { r.sy -= g->ray_offset; r.ey -= g->ray_offset; }
    return r;
}

/* ComplexGeometry.cpp:315-324: t from the plane through v1 with the file normal. */
static int t_window(const ray_t* r, const double* T, double t_eps) {
    double px = r->sx, py = r->sy, pz = r->sz;
    double nx = T[9], ny = T[10], nz = T[11];
    double t = ((T[0] - px) * nx + (T[1] - py) * ny + (T[2] - pz) * nz) /
               ((r->ex - r->sx) * nx + (r->ey - r->sy) * ny + (r->ez - r->sz) * nz);
    return t >= -t_eps && t <= 1 + t_eps;
}

/* ComplexGeometry.cpp:284-330 literally (ix, iy computed from t). */
static int count_literal(const ray_t* r, const double* tri, int64_t ntri, double t_eps) {
    int count = 0;
    for (int64_t n = 0; n < ntri; ++n) {
        const double* T = tri + 12 * n;
        double px = r->sx, py = r->sy, pz = r->sz;
        double nx = T[9], ny = T[10], nz = T[11];
        double t = ((T[0] - px) * nx + (T[1] - py) * ny + (T[2] - pz) * nz) /
                   ((r->ex - r->sx) * nx + (r->ey - r->sy) * ny + (r->ez - r->sz) * nz);
        if (t >= -t_eps && t <= 1 + t_eps) {
            double ix = px + t * (r->ex - r->sx);
            double iy = py + t * (r->ey - r->sy);
            if (inside_xy(ix, iy, T) && t >= -t_eps && t <= 1 + t_eps) count++;
        }
    }
    return count;
}

/* Conservative XY candidate test for mode 2.  A triangle whose projected
 * minimum angle is at least ~1e-6 rad cannot pass the rounded sign test at a
 * point farther than `margin` outside its XY bbox (the rounding error of each
 * d is below 8u|p - v||e| while the separating edge keeps |d| >= |e| m sin(a/2));
 * thinner triangles are tested at every column. */
static int thin_xy(const double* T) {
    double ax = T[3] - T[0], ay = T[4] - T[1], bx = T[6] - T[0], by = T[7] - T[1];
    double cx = T[6] - T[3], cy = T[7] - T[4];
    double area2 = fabs(ax * by - ay * bx);
    double l = fmax(ax * ax + ay * ay, fmax(bx * bx + by * by, cx * cx + cy * cy));
    return !(area2 >= 1e-6 * l) || l == 0.0;
}

int orc_voxelize(const double* tri, int64_t ntri, const orc_grid* g, uint8_t* flags, int mode) {
    const int nx = g->dims[0], ny = g->dims[1], nz = g->dims[2];
    if (nx <= 0 || ny <= 0 || nz <= 0) return -1;
    int64_t* cand = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ntri > 0 ? ntri : 1));
    if (!cand) return -2;
    double ext = 0.0;
    for (int a = 0; a < 3; ++a) {
        double e = fabs(g->raybox[2 * a + 1] - g->raybox[2 * a]) + fabs(g->origin[a]) +
                   g->dims[a] * g->spacing;
        if (e > ext) ext = e;
    }
    const double margin = 1e-7 * (ext > 1.0 ? ext : 1.0);
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            ray_t r0 = make_ray(g, i, j, 0);
            int64_t nc = 0;
            if (mode >= 1) {
                for (int64_t n = 0; n < ntri; ++n) {
                    const double* T = tri + 12 * n;
                    if (mode == 2 && !thin_xy(T)) {
                        double xlo = fmin(T[0], fmin(T[3], T[6])) - margin;
                        double xhi = fmax(T[0], fmax(T[3], T[6])) + margin;
                        double ylo = fmin(T[1], fmin(T[4], T[7])) - margin;
                        double yhi = fmax(T[1], fmax(T[4], T[7])) + margin;
                        if (r0.sx < xlo || r0.sx > xhi || r0.sy < ylo || r0.sy > yhi) continue;
                    }
                    if (inside_xy(r0.sx, r0.sy, T)) cand[nc++] = n;
                }
            }
            for (int k = 0; k < nz; ++k) {
                ray_t r = make_ray(g, i, j, k);
                int count = 0;
                if (mode == 0) {
                    count = count_literal(&r, tri, ntri, g->t_eps);
                } else {
                    for (int64_t c = 0; c < nc; ++c)
                        count += t_window(&r, tri + 12 * cand[c], g->t_eps);
                }
                flags[i + (int64_t)nx * (j + (int64_t)ny * k)] = (uint8_t)(count % 2 == 1);
            }
        }
    }
    free(cand);
    return 0;
}

/* ------------------------------------------------------------ classification */

int orc_split_long_double(const char* text, double* hi, double* lo) {
    char* end = NULL;
    long double v = strtold(text, &end);
    if (end == text) return -1;
    *hi = (double)v;
    *lo = (double)(v - (long double)*hi);
    return 0;
}

static long double plane_value(const orc_plane* p) { return (long double)p->value_hi + (long double)p->value_lo; }

/* ComplexGeometry.cpp:579-638 generalised: |coord - value| <= tol in long
 * double, plus strict open ranges on other coordinates. */
static int plane_match(const orc_plane* p, const double* xyz) {
    long double c = xyz[p->axis];
    long double v = plane_value(p);
    long double tol = p->tol;
    if (!(fabsl(c - v) <= tol)) return 0;
    for (int r = 0; r < p->nranges; ++r) {
        long double x = xyz[p->range_axis[r]];
        if (!((long double)p->range_lo[r] < x && x < (long double)p->range_hi[r])) return 0;
    }
    return 1;
}

/* ComplexGeometry.cpp:509-556 (boundary = fluid with a NotFluid 26-neighbour or
 * on a grid face), :559-575 (wall = boundary minus inlets/outlets) and
 * Subdomain.cpp:413-473 (priority inlet > outlet1..6 > wall). */
int orc_classify(const orc_grid* g, uint8_t* flags, const orc_plane* planes, int nplanes) {
    const int nx = g->dims[0], ny = g->dims[1], nz = g->dims[2];
    const int64_t nxy = (int64_t)nx * ny;
    uint8_t* out = (uint8_t*)malloc((size_t)(nxy * nz));
    if (!out) return -2;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int64_t s = i + nx * (j + (int64_t)ny * k);
                if (flags[s] != 1) { out[s] = 0; continue; }
                int bnd = 0;
                for (int dx = -1; dx <= 1 && !bnd; ++dx)
                    for (int dy = -1; dy <= 1 && !bnd; ++dy)
                        for (int dz = -1; dz <= 1 && !bnd; ++dz) {
                            int a = i + dx, b = j + dy, c = k + dz;
                            if (a < 0 || a >= nx || b < 0 || b >= ny || c < 0 || c >= nz) continue;
                            if (flags[a + nx * (b + (int64_t)ny * c)] == 0) bnd = 1;
                        }
                if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1) bnd = 1;
                if (!bnd) { out[s] = 1; continue; }
                double xyz[3] = {g->origin[0] + i * g->spacing, g->origin[1] + j * g->spacing,
                                 g->origin[2] + k * g->spacing};
                uint8_t f = 2;
                for (int p = 0; p < nplanes; ++p)
                    if (plane_match(&planes[p], xyz)) { f = (uint8_t)(3 + p); break; }
                out[s] = f;
            }
    memcpy(flags, out, (size_t)(nxy * nz));
    free(out);
    return 0;
}

/* ------------------------------------------------------------------ LBM step */

/* Simulation.cpp:32-78 (D3Q19) and :82-152 (D3Q27): rho summed 0..q-1, u from
 * grouped sums (ascending index within the + and - groups) divided by rho. */
static void moments(int q, const double* f, double* rho, double* u) {
    const int32_t(*c)[3] = dirs(q);
    double r = 0.0;
    for (int d = 0; d < q; ++d) r += f[d];
    for (int a = 0; a < 3; ++a) {
        double p = 0.0, m = 0.0;
        int fp = 1, fm = 1;
        for (int d = 0; d < q; ++d) {
            if (c[d][a] == 1) { p = fp ? f[d] : p + f[d]; fp = 0; }
            if (c[d][a] == -1) { m = fm ? f[d] : m + f[d]; fm = 0; }
        }
        u[a] = (p - m) / r;
    }
    *rho = r;
}

/* Simulation.cpp:154-176 */
void orc_collide_node(int q, double omega, const double* f, double* fnew, double* rho, double* u) {
    const int32_t(*c)[3] = dirs(q);
    moments(q, f, rho, u);
    for (int d = 0; d < q; ++d) {
        double du = u[0] * c[d][0] + u[1] * c[d][1] + u[2] * c[d][2];
        double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        double feq = weight(q, d) * *rho * (1.0 + 3.0 * du + 9.0 / 2.0 * du * du - 3.0 / 2.0 * usq);
        fnew[d] = omega * feq + (1.0 - omega) * f[d];
    }
}

/* Sum of f over the ascending index set {d : c[d][a] == va && (t < 0 || c[d][t] == vt)}. */
static double set_sum(int q, const double* f, int a, int va, int t, int vt) {
    const int32_t(*c)[3] = dirs(q);
    double s = 0.0;
    int first = 1;
    for (int d = 0; d < q; ++d) {
        if (c[d][a] != va) continue;
        if (t >= 0 && c[d][t] != vt) continue;
        s = first ? f[d] : s + f[d];
        first = 0;
    }
    return s;
}

/* Zou-He on the face whose unknown populations have c[axis] == sigma.
 *   D3Q19 velocity: Boundary.cpp:131-147 (y face, sigma -1).
 *   D3Q19 pressure: Simulation.cpp:966-1019 (O1 y, -1), :1029-1083 (O2 y, +1),
 *                   :1093-1279 (O3-O5 x, +1, N_xz typo), :1290-1346 (O6 z, +1).
 *   D3Q27 velocity: Boundary.cpp:101-121 (x face, sigma +1).
 * Other faces follow by the lattice symmetry; the D3Q27 pressure face has no
 * reference counterpart and reuses the D3Q19 outlet closure for u_axis. */
void orc_zouhe_node(int q, const orc_bc* bc, const double* uvec, double* f) {
    const int32_t(*c)[3] = dirs(q);
    const int32_t* opp = opps(q);
    const int a = bc->axis, sg = bc->sigma;
    int tang[2], nt = 0;
    for (int t = 0; t < 3; ++t)
        if (t != a) tang[nt++] = t;
    double S0 = set_sum(q, f, a, 0, -1, 0);
    double S1 = set_sum(q, f, a, -sg, -1, 0);
    double rho, u[3];
    if (bc->kind == ORC_BC_VELOCITY) {
        u[0] = uvec[0]; u[1] = uvec[1]; u[2] = uvec[2];
        double un = sg > 0 ? u[a] : -u[a];
        rho = (S0 + 2.0 * S1) / (1 - un);
    } else {
        rho = bc->rho;
        double v = 1 - (1.0 / rho) * (S0 + 2.0 * S1);
        u[0] = u[1] = u[2] = 0.0;
        u[a] = sg > 0 ? v : -v;
    }
    double N[3] = {0, 0, 0};
    for (int k = 0; k < 2; ++k) {
        int t = tang[k];
        if (q == 19) {
            double P = set_sum(q, f, a, 0, t, 1);
            double M = set_sum(q, f, a, 0, t, -1);
            if (bc->typo && bc->kind == ORC_BC_PRESSURE && a == 0 && sg == 1 && t == 2) {
                /* Simulation.cpp:1124-1129: f10 in place of f16 */
                P = f[4] + f[10] + f[14];
            }
            N[t] = 0.5 * (P - M) - (1.0 / 3.0) * rho * u[t];
        } else {
            /* Boundary.cpp:101-105: ((p1+p2+p3) - m1 - m2 - m3) */
            double s = set_sum(q, f, a, 0, t, 1);
            for (int d = 0; d < q; ++d)
                if (c[d][a] == 0 && c[d][t] == -1) s = s - f[d];
            N[t] = (1.0 / 6) * s - (1.0 / 9) * (rho * u[t]);
        }
    }
    double g[27];
    memcpy(g, f, sizeof(double) * q);
    for (int d = 0; d < q; ++d) {
        if (c[d][a] != sg) continue;
        int e = opp[d];
        /* c.u: normal axis first, then tangential axes in ascending order */
        double cu = sg > 0 ? u[a] : -u[a];
        int ntan = 0;
        for (int k = 0; k < 2; ++k) {
            int t = tang[k];
            if (c[d][t] == 1) { cu = cu + u[t]; ntan++; }
            if (c[d][t] == -1) { cu = cu - u[t]; ntan++; }
        }
        double coef;
        if (q == 19) coef = ntan == 0 ? 1.0 / 3.0 : 1.0 / 6.0;
        else coef = ntan == 0 ? 4 / 9.0 : (ntan == 1 ? 1 / 9.0 : 1 / 36.0);
        double v = g[e] + coef * rho * cu;
        for (int k = 0; k < 2; ++k) {
            int t = tang[k];
            if (c[d][t] == 1) v = v - N[t];
            if (c[d][t] == -1) v = v + N[t];
        }
        f[d] = v;
    }
}

#define MAXWAVE 16
#define MAXK 16

typedef struct {
    int kind;
    double period, dt;
    int nk;
    double a[MAXK], phi[MAXK];
} wave_t;

struct orc_sim {
    int q;
    int32_t dims[3];
    int64_t nsites, n;
    int quirks;
    double tau, omega;
    int64_t iter;
    int64_t* site;   /* node -> site */
    int64_t* node;   /* site -> node or -1 */
    int64_t* nb;     /* n*q: node at x + c_d or -1 (Subdomain.cpp:481-505) */
    uint8_t* type;   /* 1 fluid, 2 wall, 3+b boundary condition b */
    double* f;
    double* fnew;
    double* prof;    /* per node profile factor (velocity BC nodes) */
    int nbc;
    orc_bc* bcs;
    wave_t wave[MAXWAVE];
};

int64_t orc_sim_nodes(const orc_sim* s) { return s->n; }
int64_t orc_sim_iter(const orc_sim* s) { return s->iter; }

orc_sim* orc_sim_create(int q, const int32_t dims[3], const uint8_t* flags, double tau,
                        const orc_bc* bcs, int nbc, int quirks) {
    return orc_sim_create_periodic(q, dims, flags, tau, bcs, nbc, quirks, 0);
}

/* Periodic axes have no counterpart in the reference, whose cartesian
 * communicator is non-periodic (Subdomain.cpp:70); a periodic neighbour is the
 * node at the wrapped coordinate. */
orc_sim* orc_sim_create_periodic(int q, const int32_t dims[3], const uint8_t* flags, double tau,
                                 const orc_bc* bcs, int nbc, int quirks, int periodic) {
    if (q != 19 && q != 27) return NULL;
    orc_sim* s = (orc_sim*)calloc(1, sizeof *s);
    s->q = q;
    memcpy(s->dims, dims, sizeof s->dims);
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    s->nsites = (int64_t)nx * ny * nz;
    s->quirks = quirks;
    s->tau = tau;
    s->omega = 1.0 / tau; /* Simulation.cpp:11 */
    s->nbc = nbc;
    s->bcs = (orc_bc*)malloc(sizeof(orc_bc) * (nbc > 0 ? nbc : 1));
    if (nbc) memcpy(s->bcs, bcs, sizeof(orc_bc) * nbc);
    s->node = (int64_t*)malloc(sizeof(int64_t) * s->nsites);
    int64_t n = 0;
    /* Subdomain.cpp:638-651: node list in i-j-k order, i outermost */
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                int64_t st = i + nx * (j + (int64_t)ny * k);
                s->node[st] = flags[st] ? n++ : -1;
            }
    s->n = n;
    s->site = (int64_t*)malloc(sizeof(int64_t) * (n ? n : 1));
    s->type = (uint8_t*)malloc(n ? n : 1);
    s->nb = (int64_t*)malloc(sizeof(int64_t) * (n ? n : 1) * q);
    s->f = (double*)malloc(sizeof(double) * (n ? n : 1) * q);
    s->fnew = (double*)malloc(sizeof(double) * (n ? n : 1) * q);
    s->prof = (double*)malloc(sizeof(double) * (n ? n : 1));
    const int32_t(*c)[3] = dirs(q);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                int64_t st = i + nx * (j + (int64_t)ny * k);
                int64_t m = s->node[st];
                if (m < 0) continue;
                s->site[m] = st;
                s->type[m] = flags[st];
                for (int d = 0; d < q; ++d) {
                    int a = i + c[d][0], b = j + c[d][1], e = k + c[d][2];
                    if (periodic & 1) a = (a + nx) % nx;
                    if (periodic & 2) b = (b + ny) % ny;
                    if (periodic & 4) e = (e + nz) % nz;
                    int64_t nbn = -1;
                    if (a >= 0 && a < nx && b >= 0 && b < ny && e >= 0 && e < nz)
                        nbn = s->node[a + nx * (b + (int64_t)ny * e)];
                    s->nb[m * q + d] = nbn;
                }
                /* Node.cpp:30-37: f = w, f_new = 0 */
                for (int d = 0; d < q; ++d) {
                    s->f[m * q + d] = weight(q, d);
                    s->fnew[m * q + d] = 0.0;
                }
                double pf = 1.0;
                int t = flags[st];
                if (t >= 3 && t - 3 < nbc) {
                    const orc_bc* bc = &s->bcs[t - 3];
                    if (bc->kind == ORC_BC_VELOCITY && bc->profile == ORC_PROFILE_PARABOLIC) {
                        int ta = -1, tb = -1;
                        for (int x = 0; x < 3; ++x)
                            if (x != bc->axis) { if (ta < 0) ta = x; else tb = x; }
                        int ijk[3] = {i, j, k};
                        double d1 = ijk[ta] - bc->center[0];
                        double d2 = ijk[tb] - bc->center[1];
                        double r2 = d1 * d1 + d2 * d2;
                        pf = 1.0 - r2 / (bc->radius * bc->radius);
                        if (pf < 0.0) pf = 0.0;
                    }
                }
                s->prof[m] = pf;
            }
    return s;
}

void orc_sim_destroy(orc_sim* s) {
    if (!s) return;
    free(s->site); free(s->node); free(s->nb); free(s->type);
    free(s->f); free(s->fnew); free(s->prof); free(s->bcs);
    free(s);
}

int orc_sim_set_waveform(orc_sim* s, int slot, int kind, double period, int nk, const double* a,
                         const double* phi, double dt) {
    if (slot < 0 || slot >= MAXWAVE || nk < 0 || nk > MAXK) return -1;
    wave_t* w = &s->wave[slot];
    w->kind = kind;
    w->period = period;
    w->dt = dt;
    w->nk = nk;
    for (int k = 0; k < nk; ++k) { w->a[k] = a[k]; w->phi[k] = phi[k]; }
    return 0;
}

/* pulsatile.h:12-28 */
static double pv_eq1(double x) {
    return (-8.579e+08 * pow(x, 8) + 4.059e+08 * pow(x, 7) - 7.817e+07 * pow(x, 6) +
            7.994e+06 * pow(x, 5) - 4.835e+05 * pow(x, 4) + 1.782e+04 * pow(x, 3) -
            321.6 * pow(x, 2) + 3.215 * x + 0.1423);
}
static double pv_eq2(double x) {
    return (-4.32e+06 * pow(x, 8) + 7.94e+06 * pow(x, 7) - 6.298e+06 * pow(x, 6) +
            2.82e+06 * pow(x, 5) - 7.809e+05 * pow(x, 4) + 1.37e+05 * pow(x, 3) -
            1.484e+04 * pow(x, 2) + 901.9 * x - 22.87);
}
static double pv_eq3(double x) {
    return (-5.123 * pow(x, 4) + 14.15 * pow(x, 3) - 14.21 * pow(x, 2) + 5.865 * x - 0.5696);
}

/* Simulation.cpp:875-900.  The reference's static isFirstCycle flag skips the
 * period reduction on the first call only.  A time loop that calls it once per
 * iteration from iteration 0 (Simulation.cpp:903-932, time starting at 0.0)
 * makes that first call the one at t = 0, so the skip is keyed to iteration 0
 * here and on the device: the value at an iteration does not depend on which
 * evaluation happened first. */
static double pv_velocity(double x, double period, int64_t iter) {
    if (iter != 0) x -= period * floor(x / period);
    double vel = 0.0;
    double nxp = fmod(x, period);
    if (nxp >= 0 && nxp < (0.1 * period)) vel = pv_eq1(nxp);
    else if (nxp >= (0.1 * period) && nxp < (0.3 * period)) vel = pv_eq2(nxp);
    else if (nxp >= (0.3 * period) && nxp < (0.9 * period)) vel = pv_eq3(nxp);
    return vel;
}

static double wave_eval(const orc_sim* s, int slot, int64_t iter) {
    if (slot < 0 || slot >= MAXWAVE) return 1.0;
    const wave_t* w = &s->wave[slot];
    if (w->kind == ORC_WAVE_FOURIER) {
        double acc = 0.0;
        for (int k = 0; k < w->nk; ++k) {
            double arg = 6.283185307179586 * (double)(k + 1) * (double)iter / w->period + w->phi[k];
            acc = acc + w->a[k] * sin(arg);
        }
        return 1.0 + acc;
    }
    if (w->kind == ORC_WAVE_PV2013) return pv_velocity((double)iter * w->dt, w->period, iter);
    return 1.0;
}

double orc_waveform_eval(const orc_sim* s, int slot, int64_t iter) { return wave_eval(s, slot, iter); }

/* One Simulation::Iterate pass (Simulation.cpp:909-918) without output. */
static void step_once(orc_sim* s) {
    const int q = s->q;
    const int32_t* opp = opps(q);
    const int64_t n = s->n;
    double rho, u[3];
    /* Collision, Simulation.cpp:21-179 */
    for (int64_t m = 0; m < n; ++m) orc_collide_node(q, s->omega, s->f + m * q, s->fnew + m * q, &rho, u);
    /* Streaming, Simulation.cpp:184-198: push f_new[d] to the neighbour when it
     * exists and the source population is non-zero at the time of the visit */
    for (int64_t m = 0; m < n; ++m)
        for (int d = 0; d < q; ++d) {
            int64_t t = s->nb[m * q + d];
            if (t >= 0 && s->f[m * q + d] != 0.0) s->f[t * q + d] = s->fnew[m * q + d];
        }
    /* Waveform values for this iteration */
    double wv[MAXWAVE];
    int used[MAXWAVE] = {0};
    for (int b = 0; b < s->nbc; ++b)
        if (s->bcs[b].waveform >= 0 && s->bcs[b].waveform < MAXWAVE) used[s->bcs[b].waveform] = 1;
    for (int w = 0; w < MAXWAVE; ++w) wv[w] = used[w] ? wave_eval(s, w, s->iter) : 1.0;
    /* Bounce-back (Simulation.cpp:1353-1370) and Zou-He (ApplyZouHeBc :941-949,
     * ApplyOutletBc :952-1348).  Each node belongs to one list and every update
     * touches only that node, so the list order of the reference is immaterial. */
    const int rd = rest_dir(q);
    const int d0 = (s->quirks && q == 19) ? 1 : 0; /* D3Q19 loop starts at dir 1 */
    for (int64_t m = 0; m < n; ++m) {
        int t = s->type[m];
        int bb = s->quirks ? (t == 2) : 1;
        if (bb) {
            for (int d = d0; d < q; ++d) {
                if (d == rd) continue;
                if (s->nb[m * q + d] < 0) s->f[m * q + opp[d]] = s->fnew[m * q + d];
            }
        }
        if (t >= 3 && t - 3 < s->nbc) {
            const orc_bc* bc = &s->bcs[t - 3];
            double uv[3] = {0, 0, 0};
            if (bc->kind == ORC_BC_VELOCITY) {
                double w = bc->waveform >= 0 ? wv[bc->waveform] : 1.0;
                double fac = s->prof[m] * w;
                for (int a = 0; a < 3; ++a) uv[a] = bc->u[a] * fac;
            }
            orc_zouhe_node(q, bc, uv, s->f + m * q);
        }
    }
    s->iter++;
}

void orc_sim_step(orc_sim* s, int64_t nsteps) {
    for (int64_t i = 0; i < nsteps; ++i) step_once(s);
}

/* The boundary part of step_once as the reference runs it: three passes over
 * separate node lists, ApplyZouHeBc over the inlet list, ApplyWallBc over the
 * wall list, ApplyOutletBc over the outlet lists (Simulation.cpp:908-918), in
 * the order order[0..2] (0 inlet, 1 wall, 2 outlet; the reference's is 0 1 2).
 * Only for quirks = 1, where the lists are disjoint.  Tests compare every order
 * with step_once's single per-node loop. */
int orc_sim_step_passes(orc_sim* s, int64_t nsteps, const int32_t order[3]) {
    if (!s->quirks) return -1;
    int seen = 0;
    for (int k = 0; k < 3; ++k) {
        if (order[k] < 0 || order[k] > 2) return -1;
        seen |= 1 << order[k];
    }
    if (seen != 7) return -1;
    const int q = s->q;
    const int32_t* opp = opps(q);
    const int64_t n = s->n;
    const int rd = rest_dir(q);
    const int d0 = q == 19 ? 1 : 0;
    for (int64_t i = 0; i < nsteps; ++i) {
        double rho, u[3];
        for (int64_t m = 0; m < n; ++m) orc_collide_node(q, s->omega, s->f + m * q, s->fnew + m * q, &rho, u);
        for (int64_t m = 0; m < n; ++m)
            for (int d = 0; d < q; ++d) {
                int64_t t = s->nb[m * q + d];
                if (t >= 0 && s->f[m * q + d] != 0.0) s->f[t * q + d] = s->fnew[m * q + d];
            }
        double wv[MAXWAVE];
        for (int w = 0; w < MAXWAVE; ++w) wv[w] = 1.0;
        for (int b = 0; b < s->nbc; ++b)
            if (s->bcs[b].waveform >= 0 && s->bcs[b].waveform < MAXWAVE)
                wv[s->bcs[b].waveform] = wave_eval(s, s->bcs[b].waveform, s->iter);
        for (int k = 0; k < 3; ++k) {
            const int pass = order[k];
            for (int64_t m = 0; m < n; ++m) {
                const int t = s->type[m];
                if (pass == 1) {
                    if (t != 2) continue;
                    for (int d = d0; d < q; ++d) {
                        if (d == rd) continue;
                        if (s->nb[m * q + d] < 0) s->f[m * q + opp[d]] = s->fnew[m * q + d];
                    }
                    continue;
                }
                if (t < 3 || t - 3 >= s->nbc) continue;
                const orc_bc* bc = &s->bcs[t - 3];
                if ((pass == 0) != (bc->kind == ORC_BC_VELOCITY)) continue;
                double uv[3] = {0, 0, 0};
                if (bc->kind == ORC_BC_VELOCITY) {
                    double w = bc->waveform >= 0 ? wv[bc->waveform] : 1.0;
                    double fac = s->prof[m] * w;
                    for (int a = 0; a < 3; ++a) uv[a] = bc->u[a] * fac;
                }
                orc_zouhe_node(q, bc, uv, s->f + m * q);
            }
        }
        s->iter++;
    }
    return 0;
}

/* The same step with OpenMP over the nodes of each pass, for the timed CPU
 * baseline.  Collision and the boundary pass touch one node each; each
 * destination slot of the push has exactly one source.  The only read that
 * the sequential pass orders is the f != 0 guard (it may see a value pushed
 * earlier in the same pass): here the guard is taken on the pre-collision f,
 * recorded during collision.  An FP64 population is never exactly 0.0 (f is
 * about w rho), so both read the guard as true and give the same bits; the
 * tests compare the two step functions. */
static void step_once_threads(orc_sim* s, uint8_t* guard) {
    const int q = s->q;
    const int32_t* opp = opps(q);
    const int64_t n = s->n;
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < n; ++m) {
        double rho, u[3];
        for (int d = 0; d < q; ++d) guard[m * q + d] = s->f[m * q + d] != 0.0;
        orc_collide_node(q, s->omega, s->f + m * q, s->fnew + m * q, &rho, u);
    }
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < n; ++m)
        for (int d = 0; d < q; ++d) {
            int64_t t = s->nb[m * q + d];
            if (t >= 0 && guard[m * q + d]) s->f[t * q + d] = s->fnew[m * q + d];
        }
    double wv[MAXWAVE];
    int used[MAXWAVE] = {0};
    for (int b = 0; b < s->nbc; ++b)
        if (s->bcs[b].waveform >= 0 && s->bcs[b].waveform < MAXWAVE) used[s->bcs[b].waveform] = 1;
    for (int w = 0; w < MAXWAVE; ++w) wv[w] = used[w] ? wave_eval(s, w, s->iter) : 1.0;
    const int rd = rest_dir(q);
    const int d0 = (s->quirks && q == 19) ? 1 : 0;
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < n; ++m) {
        int t = s->type[m];
        int bb = s->quirks ? (t == 2) : 1;
        if (bb) {
            for (int d = d0; d < q; ++d) {
                if (d == rd) continue;
                if (s->nb[m * q + d] < 0) s->f[m * q + opp[d]] = s->fnew[m * q + d];
            }
        }
        if (t >= 3 && t - 3 < s->nbc) {
            const orc_bc* bc = &s->bcs[t - 3];
            double uv[3] = {0, 0, 0};
            if (bc->kind == ORC_BC_VELOCITY) {
                double w = bc->waveform >= 0 ? wv[bc->waveform] : 1.0;
                double fac = s->prof[m] * w;
                for (int a = 0; a < 3; ++a) uv[a] = bc->u[a] * fac;
            }
            orc_zouhe_node(q, bc, uv, s->f + m * q);
        }
    }
    s->iter++;
}

int orc_sim_step_threads(orc_sim* s, int64_t nsteps, int nthreads) {
    uint8_t* guard = (uint8_t*)malloc((size_t)(s->n ? s->n : 1) * s->q);
    if (!guard) return -12;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    for (int64_t i = 0; i < nsteps; ++i) step_once_threads(s, guard);
    free(guard);
    return 0;
}

int orc_threads_available(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_sim_get_f_dense(const orc_sim* s, double* f) {
    memset(f, 0, sizeof(double) * s->q * s->nsites);
    for (int64_t m = 0; m < s->n; ++m)
        for (int d = 0; d < s->q; ++d) f[d * s->nsites + s->site[m]] = s->f[m * s->q + d];
}

void orc_sim_set_f_dense(orc_sim* s, const double* f) {
    for (int64_t m = 0; m < s->n; ++m)
        for (int d = 0; d < s->q; ++d) s->f[m * s->q + d] = f[d * s->nsites + s->site[m]];
}

void orc_sim_get_macros_dense(const orc_sim* s, double* rho, double* u) {
    memset(rho, 0, sizeof(double) * s->nsites);
    memset(u, 0, sizeof(double) * 3 * s->nsites);
    for (int64_t m = 0; m < s->n; ++m) {
        double r, uu[3];
        moments(s->q, s->f + m * s->q, &r, uu);
        rho[s->site[m]] = r;
        for (int a = 0; a < 3; ++a) u[a * s->nsites + s->site[m]] = uu[a];
    }
}

/* Simulation.cpp:1374-1421 on the moments of the current state: the mean over
 * present neighbours (self included through the rest link) of u_n - u_x, scaled
 * by mu = (2 tau - 1)/6. */
void orc_sim_get_wss_dense(const orc_sim* s, double* wss) {
    const int q = s->q;
    memset(wss, 0, sizeof(double) * s->nsites);
    double* uu = (double*)malloc(sizeof(double) * 3 * (s->n ? s->n : 1));
    for (int64_t m = 0; m < s->n; ++m) {
        double r;
        moments(q, s->f + m * q, &r, uu + 3 * m);
    }
    double mu = 1.0 / 6.0 * (2.0 * s->tau - 1.0);
    for (int64_t m = 0; m < s->n; ++m) {
        if (s->type[m] != 2) continue;
        double g[3] = {0.0, 0.0, 0.0};
        int cnt = 0;
        for (int d = 0; d < q; ++d) {
            int64_t t = s->nb[m * q + d];
            if (t < 0) continue;
            for (int a = 0; a < 3; ++a) g[a] += (uu[3 * t + a] - uu[3 * m + a]);
            cnt++;
        }
        if (cnt > 0)
            for (int a = 0; a < 3; ++a) g[a] /= cnt;
        double sh[3] = {mu * g[0], mu * g[1], mu * g[2]};
        wss[s->site[m]] = sqrt(sh[0] * sh[0] + sh[1] * sh[1] + sh[2] * sh[2]);
    }
    free(uu);
}

double orc_sim_mass(const orc_sim* s) {
    double m = 0.0;
    for (int64_t i = 0; i < s->n; ++i) {
        double r = 0.0;
        for (int d = 0; d < s->q; ++d) r += s->f[i * s->q + d];
        m += r;
    }
    return m;
}
