/*
 * CPU oracle for the PV2013 lattice-Boltzmann hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C and FP64, what
 * the reference solver computes, so that the HIP engine can be checked against
 * it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the timed CPU baseline; the product
 * (pv2013_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it follows.  Parity pins (what
 * checks each function against something other than itself) are listed in the
 * header comment of lbm_oracle.c.
 */
#ifndef PV2013_LBM_ORACLE_H
#define PV2013_LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- velocity sets (LatticeModel.cpp:51-151 D3Q19, :208-334 D3Q27) ---- */
int orc_lattice(int q, int32_t* c /* q*3 */, double* w /* q */, int32_t* opp /* q */);

/* ---- STL + voxelizer ---- */
/* Triangles are 12 doubles: v1, v2, v3 (xyz each), then the file normal
 * (Struct.h:77-82 field order is x1..z3, nx, ny, nz). */
typedef struct {
    double origin[3];   /* coordinate of grid point (0,0,0) */
    double spacing;
    int32_t dims[3];    /* nx, ny, nz; flags are x-fastest: i + nx*(j + ny*k) */
    int32_t pad_;
    double raybox[6];   /* min_x, max_x, min_y, max_y, min_z, max_z used by the ray */
    double ray_offset;  /* 0.01  (ComplexGeometry.cpp:432) */
    double ray_pad;     /* 0.1   (ComplexGeometry.cpp:458) */
    double t_eps;       /* 1e-7  (ComplexGeometry.cpp:285) */
} orc_grid;

int64_t orc_read_stl_ascii(const char* path, double* tri, int64_t cap);
void orc_bbox(const double* tri, int64_t ntri, double* bbox6);
void orc_grid_from_bbox(const double* bbox6, double spacing, orc_grid* g);
/* mode 0: literal per-voxel brute force; mode 1: per-column exact XY filter
 * over all triangles; mode 2: XY bins with a conservative margin. */
int orc_voxelize(const double* tri, int64_t ntri, const orc_grid* g, uint8_t* flags, int mode);

typedef struct {
    int32_t axis;          /* coordinate compared with the plane value */
    int32_t nranges;
    double value_hi, value_lo; /* long double plane value = hi + lo */
    double tol;
    int32_t range_axis[3];
    int32_t pad_;
    double range_lo[3], range_hi[3];  /* strict: lo < coord < hi */
} orc_plane;

/* Split a decimal literal parsed as long double into hi + lo doubles. */
int orc_split_long_double(const char* text, double* hi, double* lo);

/* flags in: 0 outside / 1 inside.  out: 0 outside, 1 fluid, 2 wall, 3+p plane p */
int orc_classify(const orc_grid* g, uint8_t* flags, const orc_plane* planes, int nplanes);

/* ---- LBM step ---- */
enum { ORC_BC_VELOCITY = 0, ORC_BC_PRESSURE = 1 };
enum { ORC_PROFILE_UNIFORM = 0, ORC_PROFILE_PARABOLIC = 1 };
enum { ORC_WAVE_NONE = 0, ORC_WAVE_FOURIER = 1, ORC_WAVE_PV2013 = 2 };

typedef struct {
    int32_t kind;      /* velocity / pressure Zou-He */
    int32_t axis;      /* face normal axis */
    int32_t sigma;     /* sign of c[axis] of the unknown (incoming) populations */
    int32_t profile;
    double rho;        /* pressure BC density */
    double u[3];       /* velocity BC vector */
    double center[2];  /* parabolic profile centre on the two tangential axes */
    double radius;
    int32_t waveform;  /* -1: none, else waveform slot */
    int32_t typo;      /* reproduce the x-face N_xz index typo */
} orc_bc;

typedef struct orc_sim orc_sim;

orc_sim* orc_sim_create(int q, const int32_t dims[3], const uint8_t* flags, double tau,
                        const orc_bc* bcs, int nbc, int quirks);
/* periodic: bit a makes axis a periodic (neighbours wrap around the box) */
orc_sim* orc_sim_create_periodic(int q, const int32_t dims[3], const uint8_t* flags, double tau,
                                 const orc_bc* bcs, int nbc, int quirks, int periodic);
void orc_sim_destroy(orc_sim* s);
int orc_sim_set_waveform(orc_sim* s, int slot, int kind, double period, int nk,
                         const double* a, const double* phi, double dt);
void orc_sim_step(orc_sim* s, int64_t nsteps);
/* The same step with the boundary lists as three separate passes in a given
 * order (0 inlet, 1 wall, 2 outlet); quirks = 1 only.  0, or -1 on bad input. */
int orc_sim_step_passes(orc_sim* s, int64_t nsteps, const int32_t order[3]);
/* The same steps with OpenMP threads (timed CPU baseline); identical results
 * whenever no population is exactly 0.0.  Returns 0, or -12 without memory. */
int orc_sim_step_threads(orc_sim* s, int64_t nsteps, int nthreads);
int orc_threads_available(void);
int64_t orc_sim_nodes(const orc_sim* s);
int64_t orc_sim_iter(const orc_sim* s);
/* dense arrays over the box, site = i + nx*(j + ny*k); solid sites get 0 */
void orc_sim_get_f_dense(const orc_sim* s, double* f /* q*nsites, [d][site] */);
void orc_sim_set_f_dense(orc_sim* s, const double* f);
void orc_sim_get_macros_dense(const orc_sim* s, double* rho, double* u /* 3*nsites [c][site] */);
void orc_sim_get_wss_dense(const orc_sim* s, double* wss);
double orc_sim_mass(const orc_sim* s);
double orc_waveform_eval(const orc_sim* s, int slot, int64_t iter);

/* single-node kernels exposed for unit checks */
void orc_collide_node(int q, double omega, const double* f, double* fnew, double* rho, double* u);
void orc_zouhe_node(int q, const orc_bc* bc, const double* uvec, double* f);

#ifdef __cplusplus
}
#endif
#endif
