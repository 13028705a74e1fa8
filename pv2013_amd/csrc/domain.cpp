// Host side of the engine: context, domain build (packed node list, padded
// site map, boundary records), time stepping and the read-back entry points
// of include/lbm_capi.h.
//
// Domain build replaces SubDomain::CreateNodes / Connectivity / Initialize /
// GetFluidHaloLat (Subdomain.cpp:375-651):
//   - nodes are the non-zero flags, numbered with the fastest axis along the
//     longest fluid runs (the vessel direction) so a wavefront of 64 nodes is a
//     contiguous run and its pulls from neighbouring rows are coalesced;
//   - a padded int32 site map (one layer of -1 around the box, in the same axis
//     order) turns the reference's neighbour pointers (Node::m_neighbours) into
//     node = map[p(x) - off(c_d)] with no bounds checks;
//   - nodes with a missing link or a boundary condition get a record (found
//     through a bitmap + per-word prefix count): which populations bounce back,
//     the stale values the reference leaves in unlinked populations, and the
//     velocity profile factor.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "lattice.h"
#include "lbm_internal.h"

struct lbm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    std::set<lbm_domain*> domains;  // live domains; destroyed with the context
};

int lbm_ctx_fail(lbm_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}
hipStream_t lbm_ctx_stream(lbm_ctx* ctx) { return ctx->stream; }
int lbm_ctx_device(lbm_ctx* ctx) { return ctx->device; }

#define CHECK(call)                                                                             \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return lbm_ctx_fail(ctx, LBM_EIO, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" int lbm_version(void) { return 1; }

int lbm_vox_scratch_init(int device);  // voxelize.hip

extern "C" int lbm_ctx_create(int device, lbm_ctx** out) {
    if (!out) return LBM_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return LBM_ENODEV;
    if (device < 0 || device >= n) return LBM_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return LBM_ENODEV;
    // the device address of lbm_voxelize_dev's scratch array, with the device current
    if (lbm_vox_scratch_init(device) != LBM_OK) return LBM_EIO;
    lbm_ctx* c = new lbm_ctx;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return LBM_EIO;
    }
    *out = c;
    return LBM_OK;
}

extern "C" void lbm_domain_destroy(lbm_domain* D);

extern "C" void lbm_ctx_destroy(lbm_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    while (!ctx->domains.empty()) lbm_domain_destroy(*ctx->domains.begin());
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char* lbm_last_error(const lbm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int lbm_ctx_bandwidth(lbm_ctx* ctx, int64_t bytes, int in_place, int reps, double* gbs) {
    if (!ctx) return LBM_EINVAL;
    if (bytes <= 0 || (bytes & 15) || reps <= 0 || !gbs)
        return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_ctx_bandwidth: bad argument");
    CHECK(hipSetDevice(ctx->device));
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, (size_t)bytes) != hipSuccess) return lbm_ctx_fail(ctx, LBM_ENOMEM, "probe buffer");
    if (!in_place && hipMalloc((void**)&b, (size_t)bytes) != hipSuccess) {
        (void)hipFree(a);
        return lbm_ctx_fail(ctx, LBM_ENOMEM, "probe buffer");
    }
    const int64_t n = bytes / 16;
    hipError_t e = launch_stream_fill(a, n, ctx->stream);
    if (e == hipSuccess) e = launch_stream_probe(a, b, n, ctx->stream);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_stream_probe(a, b, n, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return lbm_ctx_fail(ctx, LBM_EIO, std::string("lbm_ctx_bandwidth: ") + hipGetErrorString(e));
    *gbs = ms > 0.f ? 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9 : 0.0;
    return LBM_OK;
}

extern "C" int lbm_ctx_bandwidth_lattice(lbm_ctx* ctx, int q, int elem_bytes, int64_t nodes, int pattern, int reps,
                                         double* gbs) {
    if (!ctx) return LBM_EINVAL;
    if ((q != 19 && q != 27) || (elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8) || nodes <= 0 ||
        nodes > ((int64_t)1 << 34) || pattern < LBM_PROBE_TWO || pattern > LBM_PROBE_AA || reps <= 0 ||
        !gbs)
        return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_ctx_bandwidth_lattice: bad argument");
    CHECK(hipSetDevice(ctx->device));
    const bool in_place = pattern == LBM_PROBE_IN_PLACE || pattern >= LBM_PROBE_IN_PLACE_SHIFTED;
    const int shift = pattern >= LBM_PROBE_TWO_SHIFTED ? 1 : 0;  // LBM_PROBE_AA: room for its shifted passes
    const int64_t ns = (nodes + shift + 63) / 64 * 64;
    const size_t bytes = (size_t)q * (size_t)ns * (size_t)elem_bytes;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return lbm_ctx_fail(ctx, LBM_ENOMEM, "probe lattice");
    if (!in_place && hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return lbm_ctx_fail(ctx, LBM_ENOMEM, "probe lattice");
    }
    // one pass of the pattern; LBM_PROBE_AA: the gather's pass (shifted, downward),
    // then the local step's (aligned, upward)
    auto pass = [&]() -> hipError_t {
        if (pattern != LBM_PROBE_AA) return launch_lattice_probe(q, elem_bytes, shift, 0, a, b, nodes, ns, ctx->stream);
        hipError_t r = launch_lattice_probe(q, elem_bytes, 1, 1, a, nullptr, nodes, ns, ctx->stream);
        return r == hipSuccess ? launch_lattice_probe(q, elem_bytes, 0, 0, a, nullptr, nodes, ns, ctx->stream) : r;
    };
    // non-zero contents (the memory system moves zeros faster)
    hipError_t e = launch_stream_fill((float4*)a, (int64_t)(bytes / 16), ctx->stream);
    if (e == hipSuccess && b) e = launch_stream_fill((float4*)b, (int64_t)(bytes / 16), ctx->stream);
    if (e == hipSuccess) e = pass();  // warm-up
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = pass();
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess)
        return lbm_ctx_fail(ctx, LBM_EIO, std::string("lbm_ctx_bandwidth_lattice: ") + hipGetErrorString(e));
    const int passes = pattern == LBM_PROBE_AA ? 2 : 1;
    *gbs = ms > 0.f ? 2.0 * (double)q * (double)elem_bytes * (double)nodes * reps * passes / (ms * 1e-3) / 1e9 : 0.0;
    return LBM_OK;
}

// ------------------------------------------------------------------ domain

struct lbm_domain {
    lbm_ctx* ctx = nullptr;
    int q = 19, storage = 0, scheme = 0, quirks = 1;
    double tau = 1.0;
    int dims[3] = {0, 0, 0};
    int perm[3] = {0, 1, 2};  // perm[0] is the fastest-varying axis of the node order
    int pd[3] = {0, 0, 0};    // padded dims in permuted order
    int64_t nsites = 0, n = 0, nrec = 0;
    int64_t ns = 0;  // population array stride: n rounded up to 64 (arrays 64-node aligned, pairs share words)
    std::vector<int64_t> node_site;  // host: node -> x-fastest site
    std::vector<int32_t> wall_nodes;
    std::vector<int32_t> bc_nodes, bc_face;
    std::vector<FaceTab> faces;
    std::vector<WaveTab> waves;
    std::vector<void*> allocs;
    int64_t device_bytes = 0;
    // device arrays
    int32_t *pnode = nullptr, *map = nullptr, *bpre = nullptr;
    uint64_t* bbits = nullptr;
    uint32_t* wreg = nullptr;  // null when the row fast path is off
    int64_t nreg_waves = 0;
    int layout = LBM_LAYOUT_FUSED;
    int32_t* rsrc = nullptr;   // record/segment layouts: [d][r] record sources
    int32_t* seg = nullptr;    // [s][LBM_SEG_W] row segments
    int64_t nseg = 0;
    int32_t* chunk = nullptr;  // fused layout: row chunks
    int64_t nchunk = 0;
    int32_t* wave = nullptr;   // fused layout: wavefront table (null: one chunk per wavefront)
    int64_t nwave = 0;
    std::vector<int32_t> wused;      // waveform slots referenced by a boundary condition
    int32_t* d_wused = nullptr;
    double* wvals = nullptr;         // [LBM_MAXWAVE][LBM_WCHUNK]
    int64_t wbase = -1;              // first iteration of the eager chunk in wvals (-1: none)
    uint8_t* btype = nullptr;
    uint32_t* bbb = nullptr;
    uint32_t* bmiss = nullptr;
    int32_t* rec_node = nullptr;
    double* bprof = nullptr;
    void* bstale = nullptr;
    FaceTab* d_faces = nullptr;
    WaveTab* d_waves = nullptr;
    int64_t* d_iter = nullptr;
    void* g[2] = {nullptr, nullptr};
    int32_t *d_wall = nullptr, *d_bcnode = nullptr, *d_bcface = nullptr, *d_bcstart = nullptr;
    double *fpre = nullptr, *rho = nullptr, *u = nullptr, *wss = nullptr, *red = nullptr, *red_scratch = nullptr;
    int cur = 0;
    int64_t iter = 0;
    // AA: iteration at which the lattice was last (re)initialised in the
    // gather-ready layout; the step that starts at iteration it is a gather
    // when it - aa_base is even, a local step otherwise
    int64_t aa_base = 0;
    int state = 0;            // 0: f = w (initial), 1: fpre holds f, 2: recompute from g[cur ^ 1]
    int64_t pre_iter = -1;    // iteration for which fpre/rho/u are valid
    int32_t off[27];
    int block = 256;
    int options = 0;
    // hipGraph of GRAPH_STEPS consecutive steps, one per starting lattice
    static constexpr int GRAPH_STEPS = 64;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};

    template <class T>
    hipError_t alloc(T** p, size_t count) {
        void* v = nullptr;
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        hipError_t e = hipMalloc(&v, bytes);
        if (e == hipSuccess) {
            allocs.push_back(v);
            device_bytes += (int64_t)bytes;
        }
        *p = (T*)v;
        return e;
    }
    ~lbm_domain() {
        for (int c = 0; c < 2; ++c) {
            if (exec[c]) (void)hipGraphExecDestroy(exec[c]);
            if (graph[c]) (void)hipGraphDestroy(graph[c]);
        }
        for (void* p : allocs) (void)hipFree(p);
    }
};

static size_t store_bytes(int storage) { return storage == LBM_STORE_F64 ? 8 : (storage == LBM_STORE_F32 ? 4 : 2); }

// Step-kernel threads per block when the description leaves it 0, from
// measurements on the benchmark configurations: FP16 storage 64; the
// two-lattice pull 128 (FP32) or 256 (FP64); AA 512 while its lattice fits the
// 256 MiB Infinity Cache (the 209 MB FP32 benchmark lattice), 256 once it
// streams from HBM (the 2.6 GB FP32 ensemble, the 417 MB FP64 lattice).
static int default_block(int storage, int scheme, uint64_t lattice_bytes) {
    if (storage == LBM_STORE_F16S) return 64;
    if (scheme == LBM_SCHEME_TWO_LATTICE) return storage == LBM_STORE_F32 ? 128 : 256;
    return lattice_bytes <= ((uint64_t)256 << 20) ? 512 : 256;
}

static int choose_fast_axis(const uint8_t* flags, const int* dims) {
    // fewest runs per fluid site = longest runs
    int64_t runs[3] = {0, 0, 0};
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int64_t s = i + (int64_t)nx * (j + (int64_t)ny * k);
                if (!flags[s]) continue;
                if (i == 0 || !flags[s - 1]) runs[0]++;
                if (j == 0 || !flags[s - nx]) runs[1]++;
                if (k == 0 || !flags[s - (int64_t)nx * ny]) runs[2]++;
            }
    int best = 2;
    for (int a = 1; a >= 0; --a)
        if (runs[a] < runs[best]) best = a;
    return best;
}

static FaceTab make_face(const lbm_bc_desc& b) {
    FaceTab F;
    std::memset(&F, 0, sizeof F);
    F.kind = b.kind;
    F.axis = b.axis;
    F.sigma = b.sigma;
    F.waveform = b.waveform;
    F.typo = b.typo;
    F.rho = b.rho;
    F.u[0] = b.u[0];
    F.u[1] = b.u[1];
    F.u[2] = b.u[2];
    return F;
}

// Parabolic factor 1 - r^2/R^2 (clamped at 0), r measured from the profile
// centre on the two tangential axes in lattice units.
static double profile_factor(const lbm_bc_desc& b, const int* ijk) {
    if (b.kind != LBM_BC_VELOCITY || b.profile != LBM_PROFILE_PARABOLIC) return 1.0;
    int ta = -1, tb = -1;
    for (int x = 0; x < 3; ++x)
        if (x != b.axis) {
            if (ta < 0) ta = x;
            else tb = x;
        }
    const double d1 = ijk[ta] - b.center[0];
    const double d2 = ijk[tb] - b.center[1];
    const double r2 = d1 * d1 + d2 * d2;
    double pf = 1.0 - r2 / (b.radius * b.radius);
    if (pf < 0.0) pf = 0.0;
    return pf;
}

static int upload(lbm_domain* D, void* dst, const void* src, size_t bytes) {
    lbm_ctx* ctx = D->ctx;
    if (bytes) CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return LBM_OK;
}

// Stale populations of boundary records, [e][r] in the storage type (FP32 and
// FP16 storage hold f - w, rounded through FP32 as the kernels round).
static int upload_stale(lbm_domain* D, const std::vector<double>& stale_f64) {
    if (D->storage == LBM_STORE_F64) return upload(D, D->bstale, stale_f64.data(), sizeof(double) * stale_f64.size());
    int rc;
    if (D->storage == LBM_STORE_F32) {
        std::vector<float> dev(stale_f64.size());
        for (int e = 0; e < D->q; ++e)
            for (int64_t r = 0; r < D->nrec; ++r)
                dev[e * D->nrec + r] = (float)(stale_f64[e * D->nrec + r] - lat_w(D->q, e));
        rc = upload(D, D->bstale, dev.data(), sizeof(float) * dev.size());
    } else {
        std::vector<_Float16> dev(stale_f64.size());
        for (int e = 0; e < D->q; ++e)
            for (int64_t r = 0; r < D->nrec; ++r)
                dev[e * D->nrec + r] = (_Float16)(float)(stale_f64[e * D->nrec + r] - lat_w(D->q, e));
        rc = upload(D, D->bstale, dev.data(), sizeof(_Float16) * dev.size());
    }
    if (rc) return rc;
    lbm_ctx* ctx = D->ctx;
    CHECK(hipStreamSynchronize(ctx->stream));
    return LBM_OK;
}

static void fill_args(const lbm_domain* D, StepArgs& a) {
    std::memset(&a, 0, sizeof a);
    a.n = D->n;
    a.ns = D->ns;
    a.block = D->block;
    // one 32-bit buffer descriptor spans the lattice only below 4 GiB
    a.wide = (D->options & LBM_OPT_WIDE_RECORDS) ||
             (uint64_t)D->q * (uint64_t)D->ns * store_bytes(D->storage) >= ((uint64_t)1 << 32);
    a.nrec = D->nrec;
    a.q = D->q;
    a.pnode = D->pnode;
    a.map = D->map;
    a.bbits = D->bbits;
    a.wreg = D->wreg;
    a.bpre = D->bpre;
    a.btype = D->btype;
    a.bbb = D->bbb;
    a.bmiss = D->bmiss;
    a.rec_node = D->rec_node;
    a.rsrc = D->rsrc;
    a.seg = D->seg;
    a.nseg = D->nseg;
    a.chunk = D->chunk;
    a.nchunk = D->nchunk;
    a.wave = D->wave;
    a.nwave = D->nwave;
    a.bprof = D->bprof;
    a.bstale = D->bstale;
    a.faces = D->d_faces;
    a.wvals = D->wvals;
    a.omega = 1.0 / D->tau;
    for (int d = 0; d < 27; ++d) a.off[d] = D->off[d];
}

// Streaming scheme of the step `rel` steps after the last (re)initialisation
// of the lattice (AA: gather on even, local on odd).
static int scheme_rel(const lbm_domain* D, int64_t rel) {
    if (D->scheme == LBM_SCHEME_TWO_LATTICE) return 0;
    return (rel & 1) == 0 ? 1 : 2;
}
// Streaming scheme of the step that starts at iteration `it`.
static int scheme_for(const lbm_domain* D, int64_t it) { return scheme_rel(D, it - D->aa_base); }
// AA: the next step is a local step (graphs start with a gather).
static bool aa_odd(const lbm_domain* D) { return D->scheme == LBM_SCHEME_AA && ((D->iter - D->aa_base) & 1); }

// wk: the launch's row of the waveform table (its iteration minus the base of
// the chunk the table holds).
static int launch_step(lbm_domain* D, int mode, int sch, const void* gin, void* gout, int wk) {
    lbm_ctx* ctx = D->ctx;
    StepArgs a;
    fill_args(D, a);
    a.wk = wk;
    a.gin = gin;
    a.gout = gout;
    // Launches alternate direction, so each starts on the lines the previous
    // one touched last: the AA gather (sch 1) walks the lattice downward, and
    // the two-lattice pull walks downward when it reads the second lattice.
    a.rev = !(D->options & LBM_OPT_FORWARD_ONLY) &&
            (D->scheme == LBM_SCHEME_AA ? sch == 1 : gin == D->g[1]);
    // lattices larger than the Infinity Cache stream as one front of 64-wavefront
    // tiles (64-pipe ensemble FP32 1140 -> 1067 us per step), cache-resident
    // ones keep one range per XCD (benchmark case FP32 AA 66.0 against 68.5)
    a.tile = lbm_launch_tile((uint64_t)D->q * (uint64_t)D->ns * store_bytes(D->storage) *
                                 (D->scheme == LBM_SCHEME_TWO_LATTICE ? 2u : 1u),
                             D->block);
    a.fpre = D->fpre;
    a.rho = D->rho;
    a.u = D->u;
    hipError_t e = D->storage == LBM_STORE_F64 ? launch_step_f64(mode, sch, D->layout, a, ctx->stream)
                                               : launch_step_lp(D->storage, mode, sch, D->layout, a, ctx->stream);
    CHECK(e);
    return LBM_OK;
}

static int init_from(lbm_domain* D, const double* fpre_dev) {
    lbm_ctx* ctx = D->ctx;
    InitArgs a;
    std::memset(&a, 0, sizeof a);
    a.n = D->n;
    a.ns = D->ns;
    a.fpre = fpre_dev;
    a.gout = D->g[D->cur];
    a.omega = 1.0 / D->tau;
    const bool swap = D->scheme == LBM_SCHEME_AA;
    hipError_t e = D->storage == LBM_STORE_F64 ? launch_init_f64(D->q, swap, a, ctx->stream)
                                               : launch_init_lp(D->storage, D->q, swap, a, ctx->stream);
    CHECK(e);
    return LBM_OK;
}

static int ensure_pre_buffers(lbm_domain* D) {
    lbm_ctx* ctx = D->ctx;
    if (!D->fpre) CHECK(D->alloc(&D->fpre, (size_t)D->q * D->n));
    if (!D->rho) CHECK(D->alloc(&D->rho, (size_t)D->n));
    if (!D->u) CHECK(D->alloc(&D->u, (size_t)3 * D->n));
    return LBM_OK;
}

extern "C" int lbm_domain_create(lbm_ctx* ctx, const lbm_domain_desc* desc, lbm_domain** out) {
    if (!ctx) return LBM_EINVAL;
    if (!desc || !out || !desc->flags) return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_domain_create: null argument");
    *out = nullptr;
    const lbm_domain_desc& d = *desc;
    if (d.q != 19 && d.q != 27) return lbm_ctx_fail(ctx, LBM_EINVAL, "q must be 19 or 27");
    if (d.storage < 0 || d.storage > 2) return lbm_ctx_fail(ctx, LBM_EINVAL, "bad storage");
    if (d.scheme != LBM_SCHEME_TWO_LATTICE && d.scheme != LBM_SCHEME_AA)
        return lbm_ctx_fail(ctx, LBM_EINVAL, "scheme must be two-lattice or AA");
    if (!(d.tau > 0.5)) return lbm_ctx_fail(ctx, LBM_EINVAL, "tau must exceed 0.5");
    for (int a = 0; a < 3; ++a)
        if (d.dims[a] <= 0) return lbm_ctx_fail(ctx, LBM_EINVAL, "dims must be positive");
    if (d.nbc < 0 || d.nbc > LBM_MAXBC || (d.nbc > 0 && !d.bcs))
        return lbm_ctx_fail(ctx, LBM_EINVAL, "bad boundary-condition list");
    for (int b = 0; b < d.nbc; ++b) {
        const lbm_bc_desc& c = d.bcs[b];
        if ((c.kind != LBM_BC_VELOCITY && c.kind != LBM_BC_PRESSURE) || c.axis < 0 || c.axis > 2 ||
            (c.sigma != 1 && c.sigma != -1) || c.waveform >= LBM_MAXWAVE ||
            (c.kind == LBM_BC_PRESSURE && !(c.rho > 0)) ||
            (c.profile == LBM_PROFILE_PARABOLIC && !(c.radius > 0)))
            return lbm_ctx_fail(ctx, LBM_EINVAL, "bad boundary condition " + std::to_string(b));
    }
    CHECK(hipSetDevice(ctx->device));
    lbm_domain* D = new lbm_domain;
    D->ctx = ctx;
    D->q = d.q;
    D->storage = d.storage;
    D->scheme = d.scheme;
    D->quirks = d.quirks;
    D->tau = d.tau;
    D->block = d.block;  // 0: the default, chosen below once the lattice size is known
    if (D->block != 0 && D->block != 64 && D->block != 128 && D->block != 256 && D->block != 512) {
        delete D;
        return lbm_ctx_fail(ctx, LBM_EINVAL, "block must be 64, 128, 256 or 512 threads");
    }
    D->options = d.options;
    for (int a = 0; a < 3; ++a) D->dims[a] = d.dims[a];
    const int nx = d.dims[0], ny = d.dims[1], nz = d.dims[2];
    D->nsites = (int64_t)nx * ny * nz;
    const uint8_t* flags = d.flags;
    for (int64_t s = 0; s < D->nsites; ++s)
        if (flags[s] >= 3 && flags[s] - 3 >= d.nbc) {
            delete D;
            return lbm_ctx_fail(ctx, LBM_EINVAL, "flag refers to a missing boundary condition");
        }
    const int fast = (d.fast_axis >= 0 && d.fast_axis < 3) ? d.fast_axis : choose_fast_axis(flags, d.dims);
    D->perm[0] = fast;
    for (int a = 0, k = 1; a < 3; ++a)
        if (a != fast) D->perm[k++] = a;
    for (int k = 0; k < 3; ++k) D->pd[k] = d.dims[D->perm[k]] + 2;
    const int64_t mapn = (int64_t)D->pd[0] * D->pd[1] * D->pd[2];
    if (mapn >= (int64_t)INT32_MAX) {
        delete D;
        return lbm_ctx_fail(ctx, LBM_ERANGE, "box too large for the int32 site map");
    }
    const int q = d.q;
    const int rest = lat_rest(q);
    for (int e = 0; e < 27; ++e) D->off[e] = 0;
    for (int e = 0; e < q; ++e) {
        const int c0 = lat_c(q, e, D->perm[0]), c1 = lat_c(q, e, D->perm[1]), c2 = lat_c(q, e, D->perm[2]);
        D->off[e] = c0 + D->pd[0] * (c1 + D->pd[1] * c2);
    }
    // node numbering and padded map
    std::vector<int32_t> map((size_t)mapn, -1);
    std::vector<int32_t> pnode;
    std::vector<int64_t>& site = D->node_site;
    int ext[3] = {d.dims[D->perm[0]], d.dims[D->perm[1]], d.dims[D->perm[2]]};
    int ijk[3];
    for (int s2 = 0; s2 < ext[2]; ++s2)
        for (int s1 = 0; s1 < ext[1]; ++s1)
            for (int s0 = 0; s0 < ext[0]; ++s0) {
                ijk[D->perm[0]] = s0;
                ijk[D->perm[1]] = s1;
                ijk[D->perm[2]] = s2;
                const int64_t st = ijk[0] + (int64_t)nx * (ijk[1] + (int64_t)ny * ijk[2]);
                if (!flags[st]) continue;
                const int64_t p = (s0 + 1) + (int64_t)D->pd[0] * ((s1 + 1) + (int64_t)D->pd[1] * (s2 + 1));
                map[p] = (int32_t)site.size();
                pnode.push_back((int32_t)p);
                site.push_back(st);
            }
    D->n = (int64_t)site.size();
    // periodic axes: fill the map's halo with the node ids of the opposite face
    const int periodic = ((d.options & LBM_OPT_PERIODIC_X) ? 1 : 0) | ((d.options & LBM_OPT_PERIODIC_Y) ? 2 : 0) |
                         ((d.options & LBM_OPT_PERIODIC_Z) ? 4 : 0);
    if (periodic) {
        int pw[3];  // periodic flag per permuted axis
        for (int k = 0; k < 3; ++k) pw[k] = (periodic >> D->perm[k]) & 1;
        for (int s2 = 0; s2 < D->pd[2]; ++s2)
            for (int s1 = 0; s1 < D->pd[1]; ++s1)
                for (int s0 = 0; s0 < D->pd[0]; ++s0) {
                    int sc[3] = {s0, s1, s2};
                    bool halo = false, ok = true;
                    for (int k = 0; k < 3; ++k) {
                        if (sc[k] == 0 || sc[k] == D->pd[k] - 1) {
                            halo = true;
                            if (!pw[k]) ok = false;
                            else sc[k] = sc[k] == 0 ? D->pd[k] - 2 : 1;
                        }
                    }
                    if (!halo || !ok) continue;
                    const int64_t p = s0 + (int64_t)D->pd[0] * (s1 + (int64_t)D->pd[1] * s2);
                    const int64_t w = sc[0] + (int64_t)D->pd[0] * (sc[1] + (int64_t)D->pd[1] * sc[2]);
                    map[p] = map[w];
                }
    }
    // ns (n rounded up to 64) stays below 2^30: FP32/FP16 kernels address a
    // population array with 32-bit byte offsets and size its buffer
    // descriptor in 32 bits (ns * 4 < 2^32), and offset 2^31 lies past every
    // FP16 array.
    if (D->n > ((int64_t)1 << 30) - 64) {
        delete D;
        return lbm_ctx_fail(ctx, LBM_ERANGE, "too many fluid nodes");
    }
    if (D->n == 0) {
        delete D;
        return lbm_ctx_fail(ctx, LBM_EINVAL, "no fluid nodes");
    }
    // boundary records
    const int64_t nwords = (D->n + 63) / 64;
    std::vector<uint64_t> bits((size_t)nwords, 0);
    std::vector<int32_t> pre((size_t)nwords, 0);
    std::vector<uint8_t> btype;
    std::vector<uint32_t> bbb, bmiss;
    std::vector<int32_t> rec_node;
    std::vector<double> bprof;
    std::vector<uint32_t> bstale_mask;
    D->faces.resize(std::max(d.nbc, 1));
    for (int b = 0; b < d.nbc; ++b) D->faces[b] = make_face(d.bcs[b]);
    D->waves.assign(LBM_MAXWAVE, WaveTab());
    for (auto& w : D->waves) std::memset(&w, 0, sizeof w);
    D->layout = (d.options & (LBM_OPT_GENERIC_KERNEL | LBM_OPT_NO_ROW_FASTPATH))
                    ? LBM_LAYOUT_GENERIC
                    : ((d.options & LBM_OPT_SPLIT_LAUNCH) ? LBM_LAYOUT_SPLIT : LBM_LAYOUT_FUSED);
    // missing links of every node
    std::vector<uint32_t> missv((size_t)D->n, 0u);
    for (int64_t m = 0; m < D->n; ++m) {
        const int64_t st = site[m];
        ijk[0] = (int)(st % nx);
        ijk[1] = (int)((st / nx) % ny);
        ijk[2] = (int)(st / ((int64_t)nx * ny));
        uint32_t miss = 0;
        for (int e = 0; e < q; ++e) {
            if (e == rest) continue;
            int x = ijk[0] - lat_c(q, e, 0), y = ijk[1] - lat_c(q, e, 1), z = ijk[2] - lat_c(q, e, 2);
            if (periodic & 1) x = (x + nx) % nx;
            if (periodic & 2) y = (y + ny) % ny;
            if (periodic & 4) z = (z + nz) % nz;
            const bool in = x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
            if (!in || !flags[x + (int64_t)nx * (y + (int64_t)ny * z)]) miss |= 1u << e;
        }
        missv[m] = miss;
    }
    // Nodes with a boundary record: a missing link or a boundary condition.
    std::vector<uint8_t> isrec((size_t)D->n, 0);
    for (int64_t m = 0; m < D->n; ++m) isrec[m] = flags[site[m]] >= 3 || missv[m] != 0;
    for (int64_t m = 0; m < D->n; ++m) {
        const int64_t st = site[m];
        ijk[0] = (int)(st % nx);
        ijk[1] = (int)((st / nx) % ny);
        ijk[2] = (int)(st / ((int64_t)nx * ny));
        const int type = flags[st];
        const uint32_t miss = missv[m];
        if (type == 2) D->wall_nodes.push_back((int32_t)m);
        if (type >= 3) {
            D->bc_nodes.push_back((int32_t)m);
            D->bc_face.push_back(type - 3);
        }
        if (!isrec[m]) continue;
        bits[m >> 6] |= 1ull << (m & 63);
        uint32_t bb;
        if (D->quirks) {
            // Simulation.cpp:1353-1370: only wall nodes bounce back, directions
            // 1..q-1 (for D3Q19 direction 0, the +x link, is never bounced)
            bb = type == 2 ? (miss & ~(q == 19 ? (1u << lat_opp(q, 0)) : 0u)) : 0u;
        } else {
            bb = miss;
        }
        btype.push_back((uint8_t)type);
        bbb.push_back(bb);
        bmiss.push_back(miss);
        rec_node.push_back((int32_t)m);
        bstale_mask.push_back(miss & ~bb);
        bprof.push_back(type >= 3 ? profile_factor(d.bcs[type - 3], ijk) : 1.0);
    }
    D->nrec = (int64_t)btype.size();
    int32_t acc = 0;
    for (int64_t w = 0; w < nwords; ++w) {
        pre[w] = acc;
        acc += __builtin_popcountll(bits[w]);
    }
    // Wavefronts (64 consecutive node ids) made of interior nodes only, all in
    // one row at consecutive positions along the fast axis: then for every
    // direction the 64 sources are fluid, hence one contiguous run of ids.
    const bool fastpath = D->layout == LBM_LAYOUT_GENERIC && !(d.options & LBM_OPT_NO_ROW_FASTPATH);
    std::vector<uint32_t> wreg((size_t)((nwords + 31) / 32), 0u);
    if (fastpath) {
        for (int64_t w = 0; w < nwords; ++w) {
            if ((w + 1) * 64 > D->n || bits[w] != 0) continue;
            const int64_t p0 = pnode[w * 64], p1 = pnode[w * 64 + 63];
            if (p1 - p0 != 63) continue;  // same row, consecutive positions
            const int64_t row0 = p0 / D->pd[0], row1 = p1 / D->pd[0];
            if (row0 != row1) continue;
            wreg[w >> 5] |= 1u << (w & 31);
            D->nreg_waves++;
        }
    }
    // Two-launch layout: record sources and row segments.  A segment
    // grows node by node while the next node is interior, follows the last
    // one along the fast axis, and its source in every direction follows the
    // last one's source; segments also end at multiples of 64 node ids so a
    // wavefront's stores cover whole cache lines.  Every run this accepts is
    // contiguous by construction, and the bounds are checked again below.
    std::vector<int32_t> rsrc, segs, chunks, waves;
    if (D->layout == LBM_LAYOUT_SPLIT) {
        rsrc.assign((size_t)q * std::max<int64_t>(D->nrec, 1), -1);
        for (int64_t r = 0; r < D->nrec; ++r) {
            const int32_t m = rec_node[r];
            for (int e = 0; e < q; ++e) rsrc[e * D->nrec + r] = e == rest ? m : map[pnode[m] - D->off[e]];
        }
        int64_t cur = -1;  // index of the open segment
        for (int64_t m = 0; m < D->n; ++m) {
            if ((bits[m >> 6] >> (m & 63)) & 1ull) {
                cur = -1;
                continue;
            }
            bool extend = cur >= 0 && (m & 63) != 0 && segs[cur * LBM_SEG_W + 1] < 64 && pnode[m] == pnode[m - 1] + 1;
            for (int e = 0; extend && e < q; ++e)
                if (e != rest && map[pnode[m] - D->off[e]] != map[pnode[m - 1] - D->off[e]] + 1) extend = false;
            if (extend) {
                segs[cur * LBM_SEG_W + 1]++;
                continue;
            }
            cur = (int64_t)(segs.size() / LBM_SEG_W);
            segs.resize(segs.size() + LBM_SEG_W, 0);
            int32_t* row = &segs[cur * LBM_SEG_W];
            row[0] = (int32_t)m;
            row[1] = 1;
            for (int e = 0; e < q; ++e) row[2 + e] = e == rest ? (int32_t)m : map[pnode[m] - D->off[e]];
        }
        D->nseg = (int64_t)(segs.size() / LBM_SEG_W);
        for (int64_t sg = 0; sg < D->nseg; ++sg) {
            const int32_t* row = &segs[sg * LBM_SEG_W];
            bool ok = row[0] >= 0 && row[1] >= 1 && row[1] <= 64 && (int64_t)row[0] + row[1] <= D->n;
            for (int e = 0; ok && e < q; ++e) ok = row[2 + e] >= 0 && (int64_t)row[2 + e] + row[1] <= D->n;
            if (!ok) {
                delete D;
                return lbm_ctx_fail(ctx, LBM_EIO, "internal error: row segment out of range");
            }
        }
        if (D->nseg == 0) segs.assign(LBM_SEG_W, 0);
    }
    // Fused layout: row chunks.  A chunk grows node by node, up to 64, while
    // the next node follows the last one along the fast axis and, for every
    // direction, its source either is missing or continues the run of
    // consecutive sources the chunk has so far (a direction whose sources
    // stopped may not start again).  So per direction the lanes with a source
    // form one interval [lo, hi) reading base + lane.  Two-lattice chunks also
    // end at 64-node words, so a full chunk's stores to its own nodes are
    // aligned, and unless LBM_OPT_ONE_CHUNK_PER_WAVE is set the chunks of each
    // word are taken two to a wavefront (the wave table below), so a word cut
    // by a row end still fills its lanes.  AA chunks store to neighbour runs
    // and may span two words (fewer, fuller chunks, one per wavefront).
    const bool word_chunks = D->scheme == LBM_SCHEME_TWO_LATTICE;
    const bool pair_chunks = word_chunks && !(d.options & LBM_OPT_ONE_CHUNK_PER_WAVE);
    if (D->layout == LBM_LAYOUT_FUSED) {
        const int CW = lbm_chunk_w(q);
        std::vector<int32_t> cbase(q, 0), clo(q, -1), chi(q, -1);
        int64_t m = 0;
        while (m < D->n) {
            const int64_t start = m;
            int len = 0;
            for (;;) {
                const int64_t x = start + len;
                if (x >= D->n || len == 64 ||
                    (len > 0 && ((word_chunks && (x & 63) == 0) || pnode[x] != pnode[x - 1] + 1)))
                    break;
                bool ok = true;
                for (int e = 0; ok && len > 0 && e < q; ++e) {
                    if (e == rest) continue;
                    const int32_t sx = map[pnode[x] - D->off[e]];
                    if (sx >= 0 && clo[e] >= 0 && (chi[e] != len || sx != cbase[e] + len)) ok = false;
                }
                if (!ok) break;
                for (int e = 0; e < q; ++e) {
                    if (e == rest) continue;
                    if (len == 0) clo[e] = chi[e] = -1;
                    const int32_t sx = map[pnode[x] - D->off[e]];
                    if (sx < 0) continue;
                    if (clo[e] < 0) {
                        clo[e] = len;
                        chi[e] = len + 1;
                        cbase[e] = sx - len;
                    } else {
                        chi[e] = len + 1;
                    }
                }
                ++len;
            }
            const size_t c0 = chunks.size();
            chunks.resize(c0 + CW, 0);
            int32_t* h = &chunks[c0];
            const uint64_t word = bits[start >> 6];
            const int sh = (int)(start & 63);
            h[0] = (int32_t)start;
            h[1] = len;
            // records before the chunk; records are numbered in node order, so
            // a lane's record index is h[2] + the mask bits below it
            h[2] = pre[start >> 6] + (sh ? __builtin_popcountll(word & ((1ull << sh) - 1ull)) : 0);
            uint64_t rm = word >> sh;  // lane l <-> node start + l, across two words
            if (sh && (start >> 6) + 1 < nwords) rm |= bits[(start >> 6) + 1] << (64 - sh);
            rm &= len == 64 ? ~0ull : ((1ull << len) - 1ull);
            h[4] = (int32_t)(uint32_t)rm;
            h[5] = (int32_t)(uint32_t)(rm >> 32);
            // does any record lane of the chunk keep a stale population?
            for (int l = 0; l < len; ++l)
                if (((rm >> l) & 1ull) && bstale_mask[(size_t)h[2] + __builtin_popcountll(rm & ((1ull << l) - 1ull))])
                    h[3] = 1;
            bool ok = len >= 1 && len <= 64 && (!word_chunks || sh + len <= 64);
            for (int e = 0; ok && e < q; ++e) {
                if (e == rest) continue;
                if (clo[e] >= 0) {
                    h[6 + e] = cbase[e];
                    h[6 + q + e / 2] |= (int32_t)((uint32_t)(clo[e] | (chi[e] << 8)) << (16 * (e & 1)));
                    ok = (int64_t)cbase[e] + clo[e] >= 0 && (int64_t)cbase[e] + chi[e] <= D->n;
                }
                // a lane without a record has every link
                const int lo_e = clo[e] < 0 ? 0 : clo[e], hi_e = clo[e] < 0 ? 0 : chi[e];
                for (int l = 0; ok && l < len; ++l)
                    if (!((rm >> l) & 1ull) && (l < lo_e || l >= hi_e)) ok = false;
            }
            if (!ok) {
                delete D;
                return lbm_ctx_fail(ctx, LBM_EIO, "internal error: bad row chunk");
            }
            m = start + len;
        }
        D->nchunk = (int64_t)(chunks.size() / CW);
        // Wave table: the chunks of one 64-node word, in order, two to a
        // wavefront (a word with k chunks takes ceil(k / 2) wavefronts).  The
        // second chunk of a pair starts where the first ends.
        if (pair_chunks) {
            for (int64_t c = 0; c < D->nchunk;) {
                const int32_t* h = &chunks[c * CW];
                const bool two = c + 1 < D->nchunk && (h[CW] >> 6) == (h[0] >> 6);
                if (c >= LBM_WAVE_PAIR || (two && (h[CW] != h[0] + h[1] || h[1] + h[CW + 1] > 64))) {
                    delete D;
                    return lbm_ctx_fail(ctx, LBM_EIO, "internal error: bad chunk pair");
                }
                waves.push_back((int32_t)c | (two ? LBM_WAVE_PAIR : 0));
                c += two ? 2 : 1;
            }
            D->nwave = (int64_t)waves.size();
        }
    }
    // device allocation
    int rc = LBM_OK;
    auto fail = [&](int code) {
        delete D;
        return code;
    };
#define DCHECK(call)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) {                                                                      \
            lbm_ctx_fail(ctx, e_ == hipErrorOutOfMemory ? LBM_ENOMEM : LBM_EIO,                     \
                         std::string(#call) + ": " + hipGetErrorString(e_));                        \
            return fail(e_ == hipErrorOutOfMemory ? LBM_ENOMEM : LBM_EIO);                          \
        }                                                                                            \
    } while (0)
    const size_t sb = store_bytes(d.storage);
    const size_t stale_sb = store_bytes(d.storage);
    DCHECK(D->alloc(&D->pnode, (size_t)D->n));
    DCHECK(D->alloc(&D->map, (size_t)mapn));
    DCHECK(D->alloc(&D->bbits, (size_t)nwords));
    DCHECK(D->alloc(&D->bpre, (size_t)nwords));
    if (fastpath) DCHECK(D->alloc(&D->wreg, wreg.size()));
    DCHECK(D->alloc(&D->btype, (size_t)D->nrec));
    DCHECK(D->alloc(&D->bbb, (size_t)D->nrec));
    DCHECK(D->alloc(&D->bmiss, (size_t)D->nrec));
    DCHECK(D->alloc(&D->rec_node, (size_t)D->nrec));
    DCHECK(D->alloc(&D->bprof, (size_t)D->nrec));
    if (D->layout == LBM_LAYOUT_SPLIT) {
        DCHECK(D->alloc(&D->rsrc, rsrc.size()));
        DCHECK(D->alloc(&D->seg, segs.size()));
    }
    if (D->layout == LBM_LAYOUT_FUSED) DCHECK(D->alloc(&D->chunk, chunks.size()));
    if (D->nwave > 0) DCHECK(D->alloc(&D->wave, waves.size()));
    uint8_t* stale_raw = nullptr;
    DCHECK(D->alloc(&stale_raw, (size_t)q * std::max<int64_t>(D->nrec, 1) * stale_sb));
    D->bstale = stale_raw;
    DCHECK(D->alloc(&D->d_faces, D->faces.size()));
    DCHECK(D->alloc(&D->d_waves, D->waves.size()));
    DCHECK(D->alloc(&D->d_iter, 1));
    for (int b = 0; b < d.nbc; ++b) {
        const int w = d.bcs[b].waveform;
        if (d.bcs[b].kind == LBM_BC_VELOCITY && w >= 0 &&
            std::find(D->wused.begin(), D->wused.end(), w) == D->wused.end())
            D->wused.push_back(w);
    }
    DCHECK(D->alloc(&D->d_wused, std::max<size_t>(D->wused.size(), 1)));
    DCHECK(D->alloc(&D->wvals, (size_t)LBM_MAXWAVE * LBM_WCHUNK));
    uint8_t *g0 = nullptr, *g1 = nullptr;
    D->ns = (D->n + 63) / 64 * 64;
    if (D->block == 0) D->block = default_block(d.storage, d.scheme, (uint64_t)q * (uint64_t)D->ns * sb);
    DCHECK(D->alloc(&g0, (size_t)q * D->ns * sb));
    if (d.scheme == LBM_SCHEME_TWO_LATTICE) DCHECK(D->alloc(&g1, (size_t)q * D->ns * sb));
    D->g[0] = g0;
    D->g[1] = d.scheme == LBM_SCHEME_TWO_LATTICE ? g1 : g0;  // AA: one lattice, updated in place
    if (d.scheme == LBM_SCHEME_AA) {
        // the last step of every lbm_step call writes the state snapshot here
        DCHECK(D->alloc(&D->fpre, (size_t)q * D->n));
        DCHECK(D->alloc(&D->rho, (size_t)D->n));
        DCHECK(D->alloc(&D->u, (size_t)3 * D->n));
    }
    DCHECK(D->alloc(&D->d_wall, std::max<size_t>(D->wall_nodes.size(), 1)));
    DCHECK(D->alloc(&D->d_bcnode, std::max<size_t>(D->bc_nodes.size(), 1)));
    DCHECK(D->alloc(&D->d_bcface, std::max<size_t>(D->bc_face.size(), 1)));
    DCHECK(D->alloc(&D->red, (size_t)LBM_MAXBC + 2));
    DCHECK(D->alloc(&D->red_scratch, (size_t)lbm_reduce_scratch((int64_t)D->bc_nodes.size())));
    // boundary-condition nodes grouped by condition, node order within one
    std::vector<int32_t> bcstart(D->faces.size() + 1, 0);
    {
        std::vector<size_t> idx(D->bc_nodes.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return D->bc_face[x] < D->bc_face[y]; });
        std::vector<int32_t> nodes(idx.size()), faces(idx.size());
        for (size_t i = 0; i < idx.size(); ++i) {
            nodes[i] = D->bc_nodes[idx[i]];
            faces[i] = D->bc_face[idx[i]];
            bcstart[faces[i] + 1]++;
        }
        for (size_t b = 0; b < D->faces.size(); ++b) bcstart[b + 1] += bcstart[b];
        D->bc_nodes.swap(nodes);
        D->bc_face.swap(faces);
    }
    DCHECK(D->alloc(&D->d_bcstart, bcstart.size()));
    if ((rc = upload(D, D->pnode, pnode.data(), sizeof(int32_t) * pnode.size()))) return fail(rc);
    if ((rc = upload(D, D->map, map.data(), sizeof(int32_t) * map.size()))) return fail(rc);
    if ((rc = upload(D, D->bbits, bits.data(), sizeof(uint64_t) * bits.size()))) return fail(rc);
    if ((rc = upload(D, D->bpre, pre.data(), sizeof(int32_t) * pre.size()))) return fail(rc);
    if (fastpath && (rc = upload(D, D->wreg, wreg.data(), sizeof(uint32_t) * wreg.size()))) return fail(rc);
    if ((rc = upload(D, D->btype, btype.data(), btype.size()))) return fail(rc);
    if ((rc = upload(D, D->bbb, bbb.data(), sizeof(uint32_t) * bbb.size()))) return fail(rc);
    if ((rc = upload(D, D->bmiss, bmiss.data(), sizeof(uint32_t) * bmiss.size()))) return fail(rc);
    if ((rc = upload(D, D->rec_node, rec_node.data(), sizeof(int32_t) * rec_node.size()))) return fail(rc);
    if ((rc = upload(D, D->bprof, bprof.data(), sizeof(double) * bprof.size()))) return fail(rc);
    if (D->layout == LBM_LAYOUT_SPLIT) {
        if ((rc = upload(D, D->rsrc, rsrc.data(), sizeof(int32_t) * rsrc.size()))) return fail(rc);
        if ((rc = upload(D, D->seg, segs.data(), sizeof(int32_t) * segs.size()))) return fail(rc);
    }
    if (D->layout == LBM_LAYOUT_FUSED &&
        (rc = upload(D, D->chunk, chunks.data(), sizeof(int32_t) * chunks.size())))
        return fail(rc);
    if (D->nwave > 0 && (rc = upload(D, D->wave, waves.data(), sizeof(int32_t) * waves.size()))) return fail(rc);
    if ((rc = upload(D, D->d_faces, D->faces.data(), sizeof(FaceTab) * D->faces.size()))) return fail(rc);
    if ((rc = upload(D, D->d_waves, D->waves.data(), sizeof(WaveTab) * D->waves.size()))) return fail(rc);
    if ((rc = upload(D, D->d_wused, D->wused.data(), sizeof(int32_t) * D->wused.size()))) return fail(rc);
    if ((rc = upload(D, D->d_wall, D->wall_nodes.data(), sizeof(int32_t) * D->wall_nodes.size()))) return fail(rc);
    if ((rc = upload(D, D->d_bcnode, D->bc_nodes.data(), sizeof(int32_t) * D->bc_nodes.size()))) return fail(rc);
    if ((rc = upload(D, D->d_bcface, D->bc_face.data(), sizeof(int32_t) * D->bc_face.size()))) return fail(rc);
    if ((rc = upload(D, D->d_bcstart, bcstart.data(), sizeof(int32_t) * bcstart.size()))) return fail(rc);
    const int64_t zero64 = 0;
    if ((rc = upload(D, D->d_iter, &zero64, sizeof zero64))) return fail(rc);
    // Node.cpp:30-37: f = w everywhere, so the stale populations hold w
    std::vector<double> stale((size_t)q * std::max<int64_t>(D->nrec, 1), 0.0);
    for (int e = 0; e < q; ++e)
        for (int64_t r = 0; r < D->nrec; ++r) stale[e * D->nrec + r] = lat_w(q, e);
    if ((rc = upload_stale(D, stale))) return fail(rc);
    D->cur = 0;
    if ((rc = init_from(D, nullptr))) return fail(rc);
    DCHECK(hipStreamSynchronize(ctx->stream));
#undef DCHECK
    D->state = 0;
    ctx->domains.insert(D);
    *out = D;
    return LBM_OK;
}

extern "C" void lbm_domain_destroy(lbm_domain* D) {
    if (!D) return;
    (void)hipSetDevice(D->ctx->device);
    (void)hipStreamSynchronize(D->ctx->stream);
    D->ctx->domains.erase(D);
    delete D;
}

extern "C" int lbm_domain_info(const lbm_domain* D, int64_t* n_nodes, int64_t* n_boundary, int64_t* device_bytes) {
    if (!D) return LBM_EINVAL;
    if (n_nodes) *n_nodes = D->n;
    if (n_boundary) *n_boundary = D->nrec;
    if (device_bytes) *device_bytes = D->device_bytes;
    return LBM_OK;
}

extern "C" int lbm_domain_get_stats(const lbm_domain* D, lbm_domain_stats* o) {
    if (!D || !o) return LBM_EINVAL;
    std::memset(o, 0, sizeof *o);
    o->nodes = D->n;
    o->boundary_records = D->nrec;
    o->wall_nodes = (int64_t)D->wall_nodes.size();
    o->bc_nodes = (int64_t)D->bc_nodes.size();
    o->block = D->block;
    o->waves = (D->n + 63) / 64;
    o->regular_waves = D->nreg_waves;
    o->segments = D->layout == LBM_LAYOUT_FUSED ? D->nchunk : D->nseg;
    o->step_waves = D->layout == LBM_LAYOUT_FUSED ? (D->nwave > 0 ? D->nwave : D->nchunk) : D->nseg;
    o->layout = D->layout;
    o->device_bytes = D->device_bytes;
    o->fast_axis = D->perm[0];
    o->scheme = D->scheme;
    return LBM_OK;
}

extern "C" int lbm_set_waveform(lbm_domain* D, int slot, int kind, double period, int nk, const double* a,
                                const double* phi, double dt) {
    if (!D) return LBM_EINVAL;
    lbm_ctx* ctx = D->ctx;
    if (slot < 0 || slot >= LBM_MAXWAVE || nk < 0 || nk > LBM_MAXK || (nk > 0 && (!a || !phi)) ||
        kind < 0 || kind > 2 || (kind != LBM_WAVE_NONE && !(period > 0)))
        return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_set_waveform: bad argument");
    CHECK(hipSetDevice(ctx->device));
    WaveTab& w = D->waves[slot];
    std::memset(&w, 0, sizeof w);
    w.kind = kind;
    w.period = period;
    w.dt = dt;
    w.nk = kind == LBM_WAVE_FOURIER ? nk : 0;
    for (int k = 0; k < w.nk; ++k) {
        w.a[k] = a[k];
        w.phi[k] = phi[k];
    }
    CHECK(hipMemcpyAsync(&D->d_waves[slot], &w, sizeof w, hipMemcpyHostToDevice, ctx->stream));
    CHECK(hipStreamSynchronize(ctx->stream));
    D->wbase = -1;  // tabulated values are stale
    // the two-lattice read-back recomputes the last step with the waveform
    // table; an AA snapshot was written by that step and stays valid
    if (D->scheme == LBM_SCHEME_TWO_LATTICE) D->pre_iter = -1;
    return LBM_OK;
}

// Captures GRAPH_STEPS steps starting from lattice `c` (two-lattice) or with
// a gather step (AA) into a graph.  Kernel arguments are fixed per node
// (the lattice parity / AA step kind alternates inside the graph); the
// iteration counter lives in device memory, so replays need no updates.
// Nothing captured uses a memset node.
static int build_graph(lbm_domain* D, int c) {
    lbm_ctx* ctx = D->ctx;
    static_assert(lbm_domain::GRAPH_STEPS == LBM_WCHUNK, "one waveform chunk per graph replay");
    CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = LBM_OK;
    // the replay's waveform values, from the device-resident iteration base
    if (launch_wave_chunk(D->d_waves, D->d_wused, (int)D->wused.size(), D->d_iter, 0, 0, D->wvals, ctx->stream) !=
        hipSuccess)
        rc = LBM_EIO;
    for (int i = 0; i < lbm_domain::GRAPH_STEPS && rc == LBM_OK; ++i)
        rc = launch_step(D, LBM_MODE_STEP, scheme_rel(D, i), D->g[(c + i) & 1], D->g[(c + i + 1) & 1], i);
    if (rc == LBM_OK && launch_advance_iter(D->d_iter, lbm_domain::GRAPH_STEPS, ctx->stream) != hipSuccess)
        rc = LBM_EIO;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    CHECK(e);
    D->graph[c] = g;
    CHECK(hipGraphInstantiate(&D->exec[c], g, nullptr, nullptr, 0));
    return LBM_OK;
}

// Row of the waveform table for an eager launch at iteration it, tabulating a
// new chunk from it when the table does not hold it.
static int eager_wave_row(lbm_domain* D, int64_t it, int* wk) {
    lbm_ctx* ctx = D->ctx;
    *wk = 0;
    if (D->wused.empty()) return LBM_OK;
    if (D->wbase < 0 || it < D->wbase || it >= D->wbase + LBM_WCHUNK) {
        CHECK(launch_wave_chunk(D->d_waves, D->d_wused, (int)D->wused.size(), nullptr, it, 1, D->wvals,
                                ctx->stream));
        D->wbase = it;
    }
    *wk = (int)(it - D->wbase);
    return LBM_OK;
}

static int eager_step(lbm_domain* D, int mode) {
    int wk = 0;
    int rc = eager_wave_row(D, D->iter, &wk);
    if (rc) return rc;
    rc = launch_step(D, mode, scheme_for(D, D->iter), D->g[D->cur], D->g[D->cur ^ 1], wk);
    if (rc) return rc;
    if (D->scheme == LBM_SCHEME_TWO_LATTICE) D->cur ^= 1;
    D->iter++;
    return LBM_OK;
}

// Builds the graph that step_n(D, nsteps) would replay, if it is not built
// yet.  Capture and instantiation run nothing on the device.
static int prepare_graph(lbm_domain* D, int64_t nsteps) {
    if (nsteps <= 0 || (D->options & LBM_OPT_NO_GRAPH)) return LBM_OK;
    const bool aa = D->scheme == LBM_SCHEME_AA;
    int64_t body = aa ? nsteps - 1 : nsteps;
    if (aa_odd(D) && body > 0) body--;
    const int c = aa ? 0 : D->cur;
    if (body >= lbm_domain::GRAPH_STEPS && !D->exec[c]) return build_graph(D, c);
    return LBM_OK;
}

static int step_n(lbm_domain* D, int64_t nsteps) {
    lbm_ctx* ctx = D->ctx;
    if (nsteps <= 0) return LBM_OK;
    const bool aa = D->scheme == LBM_SCHEME_AA;
    // AA overwrites its only lattice, so the last step of the call also writes
    // the pre-collision state that lbm_get_* read back.
    int64_t body = aa ? nsteps - 1 : nsteps;
    const int64_t G = lbm_domain::GRAPH_STEPS;
    if (!(D->options & LBM_OPT_NO_GRAPH)) {
        if (aa_odd(D) && body > 0) {  // AA graphs start with a gather step
            int rc = eager_step(D, LBM_MODE_STEP);
            if (rc) return rc;
            body--;
        }
        const int c = aa ? 0 : D->cur;
        if (body >= G) CHECK(launch_set_iter(D->d_iter, D->iter, ctx->stream));  // graph base
        while (body >= G) {
            if (!D->exec[c]) {
                int rc = build_graph(D, c);
                if (rc) return rc;
            }
            CHECK(hipGraphLaunch(D->exec[c], ctx->stream));
            D->wbase = -1;  // the replay re-tabulated wvals for its own chunk
            D->iter += G;  // G is even: the lattice parity and the AA phase are unchanged
            body -= G;
        }
    }
    for (int64_t i = 0; i < body; ++i) {
        int rc = eager_step(D, LBM_MODE_STEP);
        if (rc) return rc;
    }
    if (aa) {
        int rc = eager_step(D, LBM_MODE_SNAP);
        if (rc) return rc;
        D->state = 2;
        D->pre_iter = D->iter;
    } else {
        D->state = 2;
        D->pre_iter = -1;
    }
    return LBM_OK;
}

extern "C" int lbm_step(lbm_domain* D, int64_t nsteps) {
    if (!D) return LBM_EINVAL;
    lbm_ctx* ctx = D->ctx;
    if (nsteps < 0) return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_step: negative step count");
    CHECK(hipSetDevice(ctx->device));
    return step_n(D, nsteps);
}

extern "C" int lbm_step_timed(lbm_domain* D, int64_t nsteps, double* ms) {
    if (!D) return LBM_EINVAL;
    lbm_ctx* ctx = D->ctx;
    if (nsteps < 0) return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_step_timed: negative step count");
    CHECK(hipSetDevice(ctx->device));
    // a graph built on first use is set-up, not a step: build it before the clock starts
    int rc = prepare_graph(D, nsteps);
    if (rc) return rc;
    CHECK(hipEventRecord(ctx->ev0, ctx->stream));
    rc = step_n(D, nsteps);
    if (rc) return rc;
    CHECK(hipEventRecord(ctx->ev1, ctx->stream));
    CHECK(hipEventSynchronize(ctx->ev1));
    float t = 0.f;
    CHECK(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    if (ms) *ms = t;
    return LBM_OK;
}

extern "C" int lbm_sync(lbm_domain* D) {
    if (!D) return LBM_EINVAL;
    lbm_ctx* ctx = D->ctx;
    CHECK(hipSetDevice(ctx->device));
    CHECK(hipStreamSynchronize(ctx->stream));
    return LBM_OK;
}

extern "C" int64_t lbm_iteration(const lbm_domain* D) { return D ? D->iter : -1; }

// f, rho and u of the current state on the device (fpre, rho, u).
__global__ void k_fill_rest(double* fpre, int64_t n, int q) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < q; ++d) fpre[(int64_t)d * n + i] = q == 19 ? Lat<19>::w(d) : Lat<27>::w(d);
}

template <int Q>
__global__ void k_moments_from_f(const double* fpre, int64_t n, double* rho, double* u) {
    using L = Lat<Q>;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double f[Q];
    for (int d = 0; d < Q; ++d) f[d] = fpre[(int64_t)d * n + i];
    double r = f[0];
    for (int d = 1; d < Q; ++d) r = r + f[d];
    for (int a = 0; a < 3; ++a) {
        double p = 0.0, m = 0.0;
        bool fp = true, fm = true;
        for (int d = 0; d < Q; ++d) {
            if (L::c(d, a) == 1) { p = fp ? f[d] : p + f[d]; fp = false; }
            if (L::c(d, a) == -1) { m = fm ? f[d] : m + f[d]; fm = false; }
        }
        u[(int64_t)a * n + i] = (p - m) / r;
    }
    rho[i] = r;
}

// Host copies of device state are ordered on the context stream (a non-blocking
// stream: the legacy NULL stream does not wait for it) and complete on return.
static int copy_d2h(lbm_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return LBM_OK;
    CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    CHECK(hipStreamSynchronize(ctx->stream));
    return LBM_OK;
}

static int copy_h2d(lbm_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return LBM_OK;
    CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    CHECK(hipStreamSynchronize(ctx->stream));
    return LBM_OK;
}

static int ensure_pre(lbm_domain* D) {
    lbm_ctx* ctx = D->ctx;
    if (D->pre_iter == D->iter) {
        // the AA scheme writes the snapshot with its last step: wait for it
        CHECK(hipStreamSynchronize(ctx->stream));
        return LBM_OK;
    }
    int rc = ensure_pre_buffers(D);
    if (rc) return rc;
    const unsigned grid = (unsigned)((D->n + 255) / 256);
    if (D->state == 2) {
        if (D->scheme == LBM_SCHEME_AA) return lbm_ctx_fail(ctx, LBM_EIO, "AA state snapshot missing");
        int wk = 0;
        if ((rc = eager_wave_row(D, D->iter - 1, &wk))) return rc;
        rc = launch_step(D, LBM_MODE_PRE, 0, D->g[D->cur ^ 1], nullptr, wk);
        if (rc) return rc;
    } else {
        if (D->state == 0) hipLaunchKernelGGL(k_fill_rest, dim3(grid), dim3(256), 0, ctx->stream, D->fpre, D->n, D->q);
        if (D->q == 19)
            hipLaunchKernelGGL(k_moments_from_f<19>, dim3(grid), dim3(256), 0, ctx->stream, D->fpre, D->n, D->rho, D->u);
        else
            hipLaunchKernelGGL(k_moments_from_f<27>, dim3(grid), dim3(256), 0, ctx->stream, D->fpre, D->n, D->rho, D->u);
        CHECK(hipGetLastError());
    }
    CHECK(hipStreamSynchronize(ctx->stream));
    D->pre_iter = D->iter;
    return LBM_OK;
}

static int scatter_dense(lbm_domain* D, const double* dev, int ncomp, double* out) {
    lbm_ctx* ctx = D->ctx;
    std::vector<double> h((size_t)ncomp * D->n);
    int rc = copy_d2h(ctx, h.data(), dev, sizeof(double) * h.size());
    if (rc) return rc;
    std::memset(out, 0, sizeof(double) * ncomp * D->nsites);
    for (int c = 0; c < ncomp; ++c)
        for (int64_t m = 0; m < D->n; ++m) out[c * D->nsites + D->node_site[m]] = h[c * D->n + m];
    return LBM_OK;
}

extern "C" int lbm_get_f(lbm_domain* D, double* f) {
    if (!D) return LBM_EINVAL;
    lbm_ctx* ctx = D->ctx;
    if (!f) return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_get_f: null output");
    CHECK(hipSetDevice(ctx->device));
    int rc = ensure_pre(D);
    if (rc) return rc;
    return scatter_dense(D, D->fpre, D->q, f);
}

extern "C" int lbm_get_macros(lbm_domain* D, double* rho, double* u) {
    if (!D) return LBM_EINVAL;
    lbm_ctx* ctx = D->ctx;
    CHECK(hipSetDevice(ctx->device));
    int rc = ensure_pre(D);
    if (rc) return rc;
    if (rho && (rc = scatter_dense(D, D->rho, 1, rho))) return rc;
    if (u && (rc = scatter_dense(D, D->u, 3, u))) return rc;
    return LBM_OK;
}

extern "C" int lbm_set_f(lbm_domain* D, const double* f) {
    if (!D) return LBM_EINVAL;
    lbm_ctx* ctx = D->ctx;
    if (!f) return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_set_f: null input");
    CHECK(hipSetDevice(ctx->device));
    int rc = ensure_pre_buffers(D);
    if (rc) return rc;
    const int q = D->q;
    std::vector<double> h((size_t)q * D->n);
    for (int e = 0; e < q; ++e)
        for (int64_t m = 0; m < D->n; ++m) h[e * D->n + m] = f[e * D->nsites + D->node_site[m]];
    if ((rc = copy_h2d(ctx, D->fpre, h.data(), sizeof(double) * h.size()))) return rc;
    // stale populations take the new values (the reference never rewrites them)
    std::vector<double> stale((size_t)q * std::max<int64_t>(D->nrec, 1), 0.0);
    {
        const int64_t nwords = (D->n + 63) / 64;
        std::vector<uint64_t> bits((size_t)nwords);
        if ((rc = copy_d2h(ctx, bits.data(), D->bbits, sizeof(uint64_t) * nwords))) return rc;
        int64_t r = 0;
        for (int64_t m = 0; m < D->n; ++m)
            if ((bits[m >> 6] >> (m & 63)) & 1ull) {
                for (int e = 0; e < q; ++e) stale[e * D->nrec + r] = h[e * D->n + m];
                ++r;
            }
    }
    if ((rc = upload_stale(D, stale))) return rc;
    if ((rc = init_from(D, D->fpre))) return rc;
    CHECK(hipStreamSynchronize(ctx->stream));
    // init_from wrote the gather-ready layout, whatever the iteration: the AA
    // step sequence restarts with a gather at this iteration
    D->aa_base = D->iter;
    D->state = 1;
    D->pre_iter = -1;
    return LBM_OK;
}

static int compute_wss(lbm_domain* D);

extern "C" int lbm_get_wss(lbm_domain* D, double* wss) {
    if (!D) return LBM_EINVAL;
    lbm_ctx* ctx = D->ctx;
    if (!wss) return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_get_wss: null output");
    CHECK(hipSetDevice(ctx->device));
    int rc = compute_wss(D);
    if (rc) return rc;
    std::vector<double> h(std::max<size_t>(D->wall_nodes.size(), 1));
    if ((rc = copy_d2h(ctx, h.data(), D->wss, sizeof(double) * D->wall_nodes.size()))) return rc;
    std::memset(wss, 0, sizeof(double) * D->nsites);
    for (size_t i = 0; i < D->wall_nodes.size(); ++i) wss[D->node_site[D->wall_nodes[i]]] = h[i];
    return LBM_OK;
}

extern "C" int lbm_reduce(lbm_domain* D, double* mass, double* flux) {
    if (!D) return LBM_EINVAL;
    lbm_ctx* ctx = D->ctx;
    CHECK(hipSetDevice(ctx->device));
    int rc = ensure_pre(D);
    if (rc) return rc;
    CHECK(launch_reduce(D->rho, D->u, D->n, D->d_bcnode, D->d_bcface, D->d_bcstart, (int64_t)D->bc_nodes.size(),
                        D->d_faces, (int)D->faces.size(), D->red_scratch, D->red, ctx->stream));
    std::vector<double> h(D->faces.size() + 2);
    CHECK(hipMemcpyAsync(h.data(), D->red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    CHECK(hipStreamSynchronize(ctx->stream));
    if (mass) *mass = h[0];
    if (flux)
        for (size_t b = 0; b < D->faces.size(); ++b) flux[b] = h[1 + b];
    if (h[D->faces.size() + 1] > 0) return lbm_ctx_fail(ctx, LBM_ENAN, "non-finite density in the state");
    return LBM_OK;
}

// ------------------------------------------------------------ sub-box reads

static int scatter_box(lbm_domain* D, const double* dev, int ncomp, const int32_t* o, const int32_t* bd,
                       double* out, const std::vector<int32_t>* subset) {
    lbm_ctx* ctx = D->ctx;
    const int64_t nb = (int64_t)bd[0] * bd[1] * bd[2];
    std::memset(out, 0, sizeof(double) * ncomp * nb);
    const int nx = D->dims[0], ny = D->dims[1];
    auto inside = [&](int64_t st, int64_t* local) {
        const int x = (int)(st % nx), y = (int)((st / nx) % ny), z = (int)(st / ((int64_t)nx * ny));
        const int lx = x - o[0], ly = y - o[1], lz = z - o[2];
        if (lx < 0 || ly < 0 || lz < 0 || lx >= bd[0] || ly >= bd[1] || lz >= bd[2]) return false;
        *local = lx + (int64_t)bd[0] * (ly + (int64_t)bd[1] * lz);
        return true;
    };
    if (subset) {
        std::vector<double> h(std::max<size_t>(subset->size(), 1));
        int rc = copy_d2h(ctx, h.data(), dev, sizeof(double) * subset->size());
        if (rc) return rc;
        for (size_t i = 0; i < subset->size(); ++i) {
            int64_t l;
            if (inside(D->node_site[(*subset)[i]], &l)) out[l] = h[i];
        }
        return LBM_OK;
    }
    std::vector<double> h((size_t)ncomp * D->n);
    int rc = copy_d2h(ctx, h.data(), dev, sizeof(double) * h.size());
    if (rc) return rc;
    for (int64_t m = 0; m < D->n; ++m) {
        int64_t l;
        if (!inside(D->node_site[m], &l)) continue;
        for (int c = 0; c < ncomp; ++c) out[c * nb + l] = h[c * D->n + m];
    }
    return LBM_OK;
}

static bool box_ok(const lbm_domain* D, const int32_t* o, const int32_t* bd) {
    if (!o || !bd) return false;
    for (int a = 0; a < 3; ++a)
        if (o[a] < 0 || bd[a] <= 0 || (int64_t)o[a] + bd[a] > D->dims[a]) return false;
    return true;
}

extern "C" int lbm_get_macros_box(lbm_domain* D, const int32_t origin[3], const int32_t dims[3], double* rho,
                                  double* u) {
    if (!D) return LBM_EINVAL;
    lbm_ctx* ctx = D->ctx;
    if (!box_ok(D, origin, dims)) return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_get_macros_box: bad box");
    CHECK(hipSetDevice(ctx->device));
    int rc = ensure_pre(D);
    if (rc) return rc;
    if (rho && (rc = scatter_box(D, D->rho, 1, origin, dims, rho, nullptr))) return rc;
    if (u && (rc = scatter_box(D, D->u, 3, origin, dims, u, nullptr))) return rc;
    return LBM_OK;
}

static int compute_wss(lbm_domain* D) {
    lbm_ctx* ctx = D->ctx;
    int rc = ensure_pre(D);
    if (rc) return rc;
    if (!D->wss) CHECK(D->alloc(&D->wss, std::max<size_t>(D->wall_nodes.size(), 1)));
    WssArgs a;
    std::memset(&a, 0, sizeof a);
    a.nwall = (int64_t)D->wall_nodes.size();
    a.n = D->n;
    a.wall = D->d_wall;
    a.pnode = D->pnode;
    a.map = D->map;
    a.u = D->u;
    a.wss = D->wss;
    a.mu = 1.0 / 6.0 * (2.0 * D->tau - 1.0);
    for (int e = 0; e < 27; ++e) a.off[e] = D->off[e];
    CHECK(launch_wss(D->q, a, ctx->stream));
    CHECK(hipStreamSynchronize(ctx->stream));
    return LBM_OK;
}

extern "C" int lbm_get_wss_box(lbm_domain* D, const int32_t origin[3], const int32_t dims[3], double* wss) {
    if (!D) return LBM_EINVAL;
    lbm_ctx* ctx = D->ctx;
    if (!wss || !box_ok(D, origin, dims)) return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_get_wss_box: bad argument");
    CHECK(hipSetDevice(ctx->device));
    int rc = compute_wss(D);
    if (rc) return rc;
    return scatter_box(D, D->wss, 1, origin, dims, wss, &D->wall_nodes);
}

// ------------------------------------------------------------------- batch

extern "C" int lbm_batch_create(lbm_ctx* ctx, const lbm_domain_desc* cases, int ncases, int stack_axis,
                                lbm_domain** out, lbm_batch_case* info) {
    if (!ctx) return LBM_EINVAL;
    if (!cases || ncases <= 0 || !out || stack_axis < 0 || stack_axis > 2)
        return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_batch_create: bad argument");
    const lbm_domain_desc& c0 = cases[0];
    int64_t dims[3] = {0, 0, 0};
    int nbc = 0, nwave = 0;
    std::vector<int> bc_base(ncases), wave_base(ncases);
    std::vector<int64_t> base(ncases);
    // Periodic wrap would happen at the packed box's faces, linking one case's
    // face to another case (along the stack axis) or to solid padding: the
    // cases would no longer be independent.
    const int periodic_bits = LBM_OPT_PERIODIC_X | LBM_OPT_PERIODIC_Y | LBM_OPT_PERIODIC_Z;
    for (int c = 0; c < ncases; ++c) {
        const lbm_domain_desc& d = cases[c];
        if (!d.flags || d.q != c0.q || d.storage != c0.storage || d.scheme != c0.scheme ||
            d.quirks != c0.quirks || d.tau != c0.tau || d.nbc < 0 || (d.nbc > 0 && !d.bcs))
            return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_batch_create: case " + std::to_string(c) +
                                                     " differs in q/storage/scheme/quirks/tau or is empty");
        if (d.options != c0.options || d.fast_axis != c0.fast_axis || d.block != c0.block)
            return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_batch_create: case " + std::to_string(c) +
                                                     " differs in options/fast_axis/block");
        if (d.options & periodic_bits)
            return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_batch_create: periodic axes are not supported in a batch");
        for (int a = 0; a < 3; ++a)
            if (d.dims[a] <= 0) return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_batch_create: bad dims");
        base[c] = dims[stack_axis] + (c > 0 ? 1 : 0);
        for (int a = 0; a < 3; ++a)
            dims[a] = a == stack_axis ? base[c] + d.dims[a] : std::max<int64_t>(dims[a], d.dims[a]);
        bc_base[c] = nbc;
        nbc += d.nbc;
        wave_base[c] = nwave;
        int wmax = -1;
        for (int b = 0; b < d.nbc; ++b) wmax = std::max(wmax, d.bcs[b].waveform);
        nwave += wmax + 1;
    }
    if (nbc > LBM_MAXBC) return lbm_ctx_fail(ctx, LBM_ERANGE, "lbm_batch_create: more than 253 boundary conditions");
    if (nwave > LBM_MAXWAVE) return lbm_ctx_fail(ctx, LBM_ERANGE, "lbm_batch_create: more than 256 waveforms");
    const int64_t nsites = dims[0] * dims[1] * dims[2];
    if (dims[0] > INT32_MAX / 2 || dims[1] > INT32_MAX / 2 || dims[2] > INT32_MAX / 2 || nsites > (int64_t)1 << 34)
        return lbm_ctx_fail(ctx, LBM_ERANGE, "lbm_batch_create: packed box too large");
    std::vector<uint8_t> flags((size_t)nsites, 0);
    std::vector<lbm_bc_desc> bcs((size_t)std::max(nbc, 1));
    for (int c = 0; c < ncases; ++c) {
        const lbm_domain_desc& d = cases[c];
        int64_t o[3] = {0, 0, 0};
        o[stack_axis] = base[c];
        for (int k = 0; k < d.dims[2]; ++k)
            for (int j = 0; j < d.dims[1]; ++j)
                for (int i = 0; i < d.dims[0]; ++i) {
                    const uint8_t f = d.flags[i + (int64_t)d.dims[0] * (j + (int64_t)d.dims[1] * k)];
                    if (f >= 3 && f - 3 >= d.nbc)
                        return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_batch_create: flag refers to a missing BC");
                    const int64_t st = (i + o[0]) + dims[0] * ((j + o[1]) + dims[1] * (k + o[2]));
                    flags[st] = f >= 3 ? (uint8_t)(f + bc_base[c]) : f;
                }
        for (int b = 0; b < d.nbc; ++b) {
            lbm_bc_desc bc = d.bcs[b];
            if (bc.waveform >= 0) bc.waveform += wave_base[c];
            // profile centres are in case coordinates: shift them
            int t = 0;
            for (int a = 0; a < 3; ++a)
                if (a != bc.axis) bc.center[t++] += (double)o[a];
            bcs[bc_base[c] + b] = bc;
        }
        if (info) {
            for (int a = 0; a < 3; ++a) {
                info[c].origin[a] = (int32_t)o[a];
                info[c].dims[a] = d.dims[a];
            }
            info[c].bc_base = bc_base[c];
            info[c].wave_base = wave_base[c];
        }
    }
    lbm_domain_desc all = c0;
    for (int a = 0; a < 3; ++a) all.dims[a] = (int32_t)dims[a];
    all.flags = flags.data();
    all.bcs = bcs.data();
    all.nbc = nbc;
    return lbm_domain_create(ctx, &all, out);
}

// --------------------------------------------------------------- autotuner

static uint64_t case_key(const lbm_domain_desc* d) {
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the case
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    };
    mix(d->dims, sizeof d->dims);
    mix(&d->q, sizeof d->q);
    mix(&d->storage, sizeof d->storage);
    mix(&d->quirks, sizeof d->quirks);
    mix(d->flags, (size_t)d->dims[0] * d->dims[1] * d->dims[2]);
    return h;
}

extern "C" int lbm_autotune(lbm_ctx* ctx, const lbm_domain_desc* desc, int steps, const char* cache_path,
                            lbm_tune_result* out) {
    if (!ctx) return LBM_EINVAL;
    if (!desc || !out || !desc->flags || steps <= 0) return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_autotune: bad argument");
    std::memset(out, 0, sizeof *out);
    const uint64_t key = case_key(desc);
    out->key = key;
    if (cache_path) {
        if (FILE* fp = std::fopen(cache_path, "r")) {
            unsigned long long k;
            int sch, ax, opt, blk;
            double ml;
            while (std::fscanf(fp, "%llx %d %d %d %d %lf", &k, &sch, &ax, &opt, &blk, &ml) == 6) {
                if (k == key) {
                    out->scheme = sch;
                    out->fast_axis = ax;
                    out->options = opt;
                    out->block = blk;
                    out->mlups = ml;
                    std::fclose(fp);
                    return LBM_OK;
                }
            }
            std::fclose(fp);
        }
    }
    double best = -1.0;
    // one candidate: build, warm up for a graph chunk, time `steps` steps
    auto trial = [&](int sch, int ax, int opts, int blk) -> int {
        lbm_domain_desc d = *desc;
        d.scheme = sch;
        d.fast_axis = ax;
        d.options = opts;
        d.block = blk;
        lbm_domain* D = nullptr;
        int rc = lbm_domain_create(ctx, &d, &D);
        if (rc) return rc;
        double ms = 0.0;
        rc = lbm_step(D, 64);
        if (!rc) rc = lbm_step_timed(D, steps, &ms);
        const int64_t n = D->n;
        const int used_block = D->block;  // blk, or the domain's default when blk is 0
        lbm_domain_destroy(D);
        if (rc) return rc;
        const double mlups = ms > 0 ? (double)n * steps / (ms * 1e-3) / 1e6 : 0.0;
        out->ntried++;
        if (mlups > best) {
            best = mlups;
            out->scheme = sch;
            out->fast_axis = ax;
            out->options = opts;
            out->block = used_block;
            out->mlups = mlups;
        }
        return LBM_OK;
    };
    // stage 1: scheme x node order x kernel layout at the default block size
    static const int kern_opt[3] = {0, LBM_OPT_SPLIT_LAUNCH, LBM_OPT_GENERIC_KERNEL};
    const int keep = desc->options & ~(LBM_OPT_NO_ROW_FASTPATH | LBM_OPT_GENERIC_KERNEL | LBM_OPT_SPLIT_LAUNCH);
    for (int sch = 0; sch < 2; ++sch)
        for (int ax = 0; ax < 3; ++ax)
            for (int kern = 0; kern < 3; ++kern)
                if (int rc = trial(sch, ax, keep | kern_opt[kern], 0)) return rc;
    // stage 2: the other block sizes for the winner
    const int bs = out->scheme, ba = out->fast_axis, bo = out->options, bb = out->block;
    for (int blk : {64, 128, 256, 512})
        if (blk != bb)
            if (int rc = trial(bs, ba, bo, blk)) return rc;
    if (cache_path) {
        if (FILE* fp = std::fopen(cache_path, "a")) {
            std::fprintf(fp, "%llx %d %d %d %d %.6f\n", (unsigned long long)key, out->scheme, out->fast_axis,
                         out->options, out->block, out->mlups);
            std::fclose(fp);
        }
    }
    return LBM_OK;
}
