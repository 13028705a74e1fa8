/*
 * lbm_capi.h - C ABI of the MI355X lattice-Boltzmann engine (libpv2013_lbm.so).
 *
 * Plain C types only.  Every call returns 0 on success or a negative errno-style
 * code (LBM_E*); no C++ exception crosses this boundary.  Host buffers passed in
 * are copied; the library owns all device memory and frees it in *_destroy.
 * Calls on one lbm_ctx are stream-ordered and not thread-safe.
 *
 * The reference (Resacd/PV2013) has no FFI layer: its boundary is a handful of
 * C++ classes constructed in main() (LBM.cpp:68-71).  Each entry point below
 * names the reference code it replaces.
 *
 * Layouts shared by all calls:
 *   - box arrays are x-fastest: site = i + nx*(j + ny*k);
 *   - flags: 0 outside, 1 fluid, 2 wall, 3+b boundary condition b (inlet or
 *     outlet, in the order given to lbm_classify / lbm_domain_create);
 *   - triangles: 12 doubles each, v1 v2 v3 (xyz) then the normal as written in
 *     the STL file (Struct.h:77-82);
 *   - per-population arrays are [d][site], d in the reference's index order
 *     (LatticeModel.cpp:51-106 for D3Q19, :208-266 for D3Q27).
 */
#ifndef PV2013_LBM_CAPI_H
#define PV2013_LBM_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_EIO (-5)      /* HIP runtime error, file error */
#define LBM_ENOMEM (-12)  /* device or host allocation failed */
#define LBM_ENODEV (-19)  /* no such device / library built without GPU support */
#define LBM_EINVAL (-22)  /* invalid argument */
#define LBM_ERANGE (-34)  /* size beyond a stated limit */
#define LBM_ENAN (-61)    /* non-finite state found by a health check */

#define LBM_MAX_BCS 253        /* boundary conditions per domain (flag values 3..255) */
#define LBM_MAX_WAVEFORMS 256  /* waveform slots per domain */

typedef struct lbm_ctx lbm_ctx;
typedef struct lbm_domain lbm_domain;

/* ------------------------------------------------------------------ context */

/* Opens `device` and creates the context stream.  Replaces the MPI_Init /
 * cartesian communicator set-up (LBM.cpp:25, Subdomain.cpp:67-230): one GPU
 * holds the whole domain. */
int lbm_ctx_create(int device, lbm_ctx** out);
/* Also destroys every domain still open on the context (their handles become
 * invalid). */
void lbm_ctx_destroy(lbm_ctx* ctx);
/* Text of the last error on this context (never NULL). */
const char* lbm_last_error(const lbm_ctx* ctx);

/* Measured device bandwidth for the roofline next to the published HBM peak:
 * a streaming kernel over a scratch buffer of `bytes` (a multiple of 16),
 * in place (read and write the same buffer) or as a copy between two buffers,
 * repeated `reps` times after one warm-up pass; *gbs = bytes read + written per
 * second, in GB/s, from HIP events.  No reference counterpart. */
int lbm_ctx_bandwidth(lbm_ctx* ctx, int64_t bytes, int in_place, int reps, double* gbs);
/* Practical ceiling of a step's memory traffic: `q` (19 or 27) population
 * arrays of `nodes` elements of `elem_bytes` (2, 4 or 8) bytes each, stored as
 * the step stores them (one array per direction, stride rounded up to 64
 * nodes).  One thread per node reads its q elements, then writes q elements,
 * with no arithmetic beyond an integer add.  `pattern`:
 *   LBM_PROBE_TWO (0)          into a second lattice, aligned (the two-lattice
 *                              pull's stores);
 *   LBM_PROBE_IN_PLACE (1)     in place, into the arrays in reverse order (the
 *                              AA local step);
 *   LBM_PROBE_TWO_SHIFTED (2)  into a second lattice, each read one node along,
 *                              so every 64-lane read run starts off its cache
 *                              line as a neighbour run does (the pull);
 *   LBM_PROBE_IN_PLACE_SHIFTED (3)  in place, reads and writes one node along
 *                              (the AA gather);
 *   LBM_PROBE_AA (4)           the AA step's pair of launches: a shifted pass
 *                              in place walking the lattice downward (the
 *                              gather), then an aligned one walking it upward
 *                              (the local step).
 * Patterns 0-3 walk upward.  Repeated `reps` times (pattern 4: pairs) after one
 * warm-up; *gbs = bytes read + written per second, GB/s, from HIP events.  No
 * reference counterpart. */
#define LBM_PROBE_TWO 0
#define LBM_PROBE_IN_PLACE 1
#define LBM_PROBE_TWO_SHIFTED 2
#define LBM_PROBE_IN_PLACE_SHIFTED 3
#define LBM_PROBE_AA 4
int lbm_ctx_bandwidth_lattice(lbm_ctx* ctx, int q, int elem_bytes, int64_t nodes, int pattern, int reps,
                              double* gbs);
int lbm_version(void);

/* ----------------------------------------------------------------- geometry */

typedef struct {
    double origin[3];  /* coordinate of grid point (0,0,0) */
    double spacing;
    int32_t dims[3];
    int32_t pad_;
    double raybox[6];  /* min_x, max_x, min_y, max_y, min_z, max_z seen by the ray */
    double ray_offset; /* 0.01 */
    double ray_pad;    /* 0.1 */
    double t_eps;      /* 1e-7 */
} lbm_grid_desc;

/* Reader::readSTLASCII (GeometryReader.cpp:33-117).  Returns the facet count;
 * fills at most `cap` facets when `tri` is not NULL. */
int64_t lbm_read_stl_ascii(const char* path, double* tri, int64_t cap);
/* Reader::readSTLBinary (GeometryReader.cpp:119-176), normals renormalised in
 * float as the reference does. */
int64_t lbm_read_stl_binary(const char* path, double* tri, int64_t cap);

/* Geometry::GetBoundingbox + generatePointsInBoundingBox
 * (ComplexGeometry.cpp:40-73, 178-207): origin = bbox min, dims =
 * ceil(extent/spacing)+1, the ray box is the STL bbox. */
int lbm_grid_from_bbox(const double* tri, int64_t ntri, double spacing, lbm_grid_desc* out);

/* Geometry::labelPoints (ComplexGeometry.cpp:431-506): +z ray parity per grid
 * point, bit-exact.  flags_out gets 0 (outside) or 1 (inside).  kernel_ms
 * (may be NULL) receives the device time of the facet-binning and voxelizer
 * kernels (host copies excluded). */
int lbm_voxelize(lbm_ctx* ctx, const double* tri, int64_t ntri, const lbm_grid_desc* grid,
                 uint8_t* flags_out, double* kernel_ms);

/* lbm_voxelize with device buffers, allocating nothing: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture of its own work, tight serving loops, harnesses that forbid
 * hipMalloc).  tri_dev (12 doubles per facet) and flags_dev (one byte per grid
 * point, x fastest) are device pointers on the context's device; the facet
 * bins live in a fixed 64 MiB array per device that the library holds until
 * the process ends.  A case whose bins exceed it (tile arrays of 12 bytes per
 * 8 x 8 column tile, 20 bytes per facet, 4 bytes per facet-tile entry) fails
 * with LBM_ERANGE before any device work.  Calls on one device share that
 * array: calls that overlap on one device race, so serialise them.  Returns
 * when the flags are written (one 32-byte device-to-host read of the bin
 * counts in between).  Same flags as lbm_voxelize, bit for bit. */
int lbm_voxelize_dev(lbm_ctx* ctx, const double* tri_dev, int64_t ntri, const lbm_grid_desc* grid,
                     uint8_t* flags_dev, double* kernel_ms);

typedef struct {
    int32_t axis;      /* coordinate tested against the plane value */
    int32_t nranges;   /* 0..3 extra open-interval conditions */
    double value_hi, value_lo; /* plane value, a long double split as hi + lo */
    double tol;        /* |coord - value| <= tol, evaluated in long double */
    int32_t range_axis[3];
    int32_t pad_;
    double range_lo[3], range_hi[3]; /* lo < coord < hi */
} lbm_plane_desc;

/* Geometry::findBoundaryPoints + getInOutBP + excludeInletOutlet
 * (ComplexGeometry.cpp:509-638) and SubDomain::CreateNodes' priority
 * (Subdomain.cpp:413-473).  flags in: 0/1; out: 0 outside, 1 fluid, 2 wall,
 * 3+p for the first matching plane p. */
int lbm_classify(lbm_ctx* ctx, const lbm_grid_desc* grid, const lbm_plane_desc* planes,
                 int nplanes, uint8_t* flags_inout, double* kernel_ms);

/* ------------------------------------------------------------------ domain */

enum { LBM_STORE_F64 = 0, LBM_STORE_F32 = 1, LBM_STORE_F16S = 2 };
enum { LBM_SCHEME_TWO_LATTICE = 0, LBM_SCHEME_AA = 1 };
enum { LBM_BC_VELOCITY = 0, LBM_BC_PRESSURE = 1 };
enum { LBM_PROFILE_UNIFORM = 0, LBM_PROFILE_PARABOLIC = 1 };
enum { LBM_WAVE_NONE = 0, LBM_WAVE_FOURIER = 1, LBM_WAVE_PV2013 = 2 };

typedef struct {
    int32_t kind;     /* LBM_BC_VELOCITY | LBM_BC_PRESSURE (Zou-He) */
    int32_t axis;     /* face normal axis */
    int32_t sigma;    /* sign of c[axis] of the incoming (unknown) populations */
    int32_t profile;  /* velocity: uniform or parabolic 1 - r^2/R^2 (clamped at 0) */
    double rho;       /* pressure BC density */
    double u[3];      /* velocity BC vector (scaled by profile and waveform) */
    double center[2]; /* profile centre on the two tangential axes (lattice units) */
    double radius;
    int32_t waveform; /* -1 none, else a waveform slot set by lbm_set_waveform */
    int32_t typo;     /* reproduce the x-face N_xz f10/f16 swap (Simulation.cpp:1124) */
} lbm_bc_desc;

typedef struct {
    int32_t dims[3];
    int32_t q;        /* 19 or 27 */
    int32_t storage;  /* LBM_STORE_* */
    int32_t scheme;   /* LBM_SCHEME_* */
    int32_t quirks;   /* 1: the reference's step semantics; 0: clean bounce-back */
    int32_t block;    /* step-kernel threads per block: 64, 128, 256 or 512; 0 = the
                         measured default: 64 for FP16 storage; two-lattice pull 128
                         (FP32) or 256 (FP64); AA 512 while the lattice fits the 256 MiB
                         Infinity Cache, else 256 */
    double tau;
    const uint8_t* flags;    /* host, dims product */
    const lbm_bc_desc* bcs;  /* bcs[b] drives flag 3+b */
    int32_t nbc;
    int32_t fast_axis;       /* node order: fastest-varying axis, -1 = longest runs */
    int32_t options;         /* LBM_OPT_* bits */
} lbm_domain_desc;

/* lbm_domain_desc.options */
/* Step kernels.  Default: one launch per step over row chunks, one wavefront
 * each: up to 64 consecutive node ids of one row whose sources per direction
 * are consecutive ids, read with contiguous loads; the lanes of boundary
 * records (nodes with a missing link or a boundary condition) pick their
 * sources per lane.  The AA local step is one dense launch in node order.
 * LBM_OPT_SPLIT_LAUNCH: records one thread each and row segments, in two
 * launches.  LBM_OPT_GENERIC_KERNEL: one thread per node, sources through the
 * site map, with a scalar fast path for whole interior wavefronts (off with
 * LBM_OPT_NO_ROW_FASTPATH, which implies the generic kernel).  All give
 * identical results; the generic kernel shares no addressing code with the
 * others. */
#define LBM_OPT_NO_ROW_FASTPATH 1  /* generic kernel, every lane through the site map */
#define LBM_OPT_NO_GRAPH 2         /* launch every step eagerly instead of replaying a hipGraph */
/* Periodic axes: the halo of the site map holds the wrapped node ids (a ghost
 * fill done once at build time), so streaming wraps at no cost per step.  The
 * reference's communicator is non-periodic (Subdomain.cpp:70); this replaces
 * its disabled halo exchange (Simulation.cpp:202-872) on one GPU. */
#define LBM_OPT_PERIODIC_X 4
#define LBM_OPT_PERIODIC_Y 8
#define LBM_OPT_PERIODIC_Z 16
#define LBM_OPT_GENERIC_KERNEL 32  /* generic per-node kernel instead of row chunks */
#define LBM_OPT_SPLIT_LAUNCH 64    /* records and segments in two launches per step */
/* FP16 row-chunk kernel: chunks holding boundary records address the lattice
 * through 64-bit pointers instead of one 32-bit buffer descriptor over all Q
 * arrays.  Automatic for FP16 lattices of 4 GiB or more; as an option it lets
 * tests compare the two record paths.  Results are identical. */
#define LBM_OPT_WIDE_RECORDS 128
/* Two-lattice row chunks one per wavefront.  By default the two-lattice pull's
 * chunks (which end at aligned 64-node words) are taken two to a wavefront,
 * each lane addressing through its own chunk, which fills the lanes of words
 * cut by a row end.  AA chunks are always one per wavefront.  Results are
 * identical. */
#define LBM_OPT_ONE_CHUNK_PER_WAVE 256
/* Every step launch walks the lattice upward.  By default launches alternate
 * direction (the AA gather walks downward, and the two-lattice pull when it
 * reads the second lattice), so each launch starts on the cache lines the
 * previous launch touched last.  Results are identical. */
#define LBM_OPT_FORWARD_ONLY 512

/* SubDomain(Geometry&) (Subdomain.cpp:3-33: CreateNodes, Connectivity,
 * Initialize, GetFluidHaloLat) + Simulation's constructor set-up
 * (Simulation.cpp:3-12): builds the packed fluid-node list, site map,
 * boundary records, and initialises f = w (Node.cpp:30-37). */
int lbm_domain_create(lbm_ctx* ctx, const lbm_domain_desc* desc, lbm_domain** out);
void lbm_domain_destroy(lbm_domain* dom);
int lbm_domain_info(const lbm_domain* dom, int64_t* n_nodes, int64_t* n_boundary,
                    int64_t* device_bytes);

typedef struct {
    int64_t nodes, boundary_records, wall_nodes, bc_nodes;
    int64_t waves, regular_waves;  /* generic kernel: 64-node wavefronts; those on the row fast path */
    int64_t segments;              /* row chunks (one-launch layout) or row segments (two-launch layout) */
    int64_t device_bytes;
    int32_t fast_axis;             /* node-order fastest axis */
    int32_t scheme;
    int32_t layout;                /* 0 generic kernel, 1 row chunks (one launch), 2 records + segments (two) */
    int32_t block;                 /* step-kernel threads per block in use */
    int64_t step_waves;            /* wavefronts of one step launch (row chunks: chunk pairs and single chunks) */
} lbm_domain_stats;
int lbm_domain_get_stats(const lbm_domain* dom, lbm_domain_stats* out);

/* Pulsatile inlet waveform for slot (Simulation.cpp:875-900, pulsatile.h).
 * FOURIER: W(t) = 1 + sum_k a_k sin(2 pi k t / period + phi_k), t in steps.
 * PV2013: the reference's piecewise polynomial at t*dt seconds, period 0.9. */
int lbm_set_waveform(lbm_domain* dom, int slot, int kind, double period, int nk,
                     const double* a, const double* phi, double dt);

/* Simulation::Iterate body (Simulation.cpp:908-918) nsteps times: collision,
 * streaming, inlet, wall, outlet.  Enqueued on the context stream. */
int lbm_step(lbm_domain* dom, int64_t nsteps);
/* Same, and returns the device time of the steps (HIP events on the stream).
 * A step graph not yet built is captured before the first event, so the time
 * holds the steps only. */
int lbm_step_timed(lbm_domain* dom, int64_t nsteps, double* ms);
int lbm_sync(lbm_domain* dom);
int64_t lbm_iteration(const lbm_domain* dom);

/* State after the last step, FP64, dense [d][site] (zero outside). */
int lbm_get_f(lbm_domain* dom, double* f);
/* Replace the state (dense [d][site]); the reference's stale populations take
 * these values. */
int lbm_set_f(lbm_domain* dom, const double* f);
/* Moments of the current state (Simulation.cpp:32-78): rho and u [c][site]. */
int lbm_get_macros(lbm_domain* dom, double* rho, double* u);
/* The same restricted to the box [origin, origin + dims) of the domain
 * (arrays sized for that box, x fastest); used to read one case of a batch. */
int lbm_get_macros_box(lbm_domain* dom, const int32_t origin[3], const int32_t dims[3],
                       double* rho, double* u);
int lbm_get_wss_box(lbm_domain* dom, const int32_t origin[3], const int32_t dims[3], double* wss);
/* Wall shear stress of the current state on wall sites
 * (Simulation.cpp:1374-1421), zero elsewhere. */
int lbm_get_wss(lbm_domain* dom, double* wss);
/* FP64 health check: total mass and per-boundary-condition mass flux through
 * its face (flux has nbc entries, may be NULL).  LBM_ENAN on a non-finite state.
 * The sums run in a fixed order (no atomics): the same state gives the same
 * bits on every call and every device. */
int lbm_reduce(lbm_domain* dom, double* mass, double* flux);

/* ---------------------------------------------------------------- autotune */

typedef struct {
    int32_t scheme;     /* best LBM_SCHEME_* */
    int32_t fast_axis;  /* best node-order axis */
    int32_t options;    /* best LBM_OPT_* bits */
    int32_t ntried;     /* candidates timed (0 when read from the cache) */
    double mlups;       /* fluid-node MLUPS of the best candidate */
    uint64_t key;       /* case signature used in the cache */
    int32_t block;      /* best threads per block (lbm_domain_desc.block) */
    int32_t pad_;
} lbm_tune_result;

/* Times each candidate layout of `desc` (streaming scheme two-lattice / AA,
 * node-order fast axis x / y / z, kernel: row chunks in one launch / records
 * and segments in two / generic, at the default block size for the storage and scheme;
 * then the other block sizes of 64 / 128 / 256 / 512 for the winner) for `steps` steps
 * after a warm-up and returns the fastest.  When cache_path is not NULL the
 * result is looked up there first (key: lattice, storage, dims, flag hash) and
 * appended after tuning.  No reference counterpart. */
int lbm_autotune(lbm_ctx* ctx, const lbm_domain_desc* desc, int steps, const char* cache_path,
                 lbm_tune_result* out);

/* ------------------------------------------------------------------- batch */

typedef struct {
    int32_t origin[3];  /* where the case's box starts inside the packed domain */
    int32_t dims[3];
    int32_t bc_base;    /* the case's boundary condition b is b + bc_base of the domain */
    int32_t wave_base;  /* the case's waveform slot s is s + wave_base of the domain */
} lbm_batch_case;

/* Many independent cases of uneven size in one domain: boxes are stacked along
 * `stack_axis` with one solid layer between them, so one launch per step (one
 * graph replay per 64 steps) updates the whole ensemble.  All cases share q,
 * storage, scheme, quirks and tau.  No reference counterpart (the reference
 * runs one geometry per process). */
int lbm_batch_create(lbm_ctx* ctx, const lbm_domain_desc* cases, int ncases, int stack_axis,
                     lbm_domain** out, lbm_batch_case* info);

#ifdef __cplusplus
}
#endif
#endif
