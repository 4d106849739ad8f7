// spbp.hip -- libspbp: Fourier-gridding backprojection (SPIDER `BP 3F` /
// `BP 32F`) for MI355X (gfx950, CDNA4).  C ABI in include/spbp.h.
//
// Pipeline per chunk of B projections (all on one HIP stream):
//   pad_kernel      PADD2 + ASTA_D (bp3f.f:1307-1343, asta.f:61-92): FP64 mean
//                   outside the circle, zero-padded N x N image.
//   hipFFT R2C      batched 2D forward transform (FMRS_2, fmrs_2.f:53); the
//                   SPIDER sign (conjugate, fftw3.f:411-419) and the
//                   checkerboard (bp32f.f:859-864) are folded into the read of
//                   the insertion kernel.
//   insert kernel   ONELINE over every row J of every projection
//                   (oneline.f:38-158): rotate, Friedel flip, nearest voxel,
//                   bounds check, 5x5x5 Kaiser-Bessel scatter with reflected
//                   taps, FP32 atomics into the half-grid X and weights W.
// Finalisation per output volume:
//   normalize       SYMPLANE0 (bp32f.f:1051-1088) evaluated on the fly, halves
//                   summed (ADDADA, bp32f.f:345-346), X (-1)^(i+j+k) / W where
//                   W > 0.1 (NRMW2, bp3f.f:1355-1366), FFTW c2r semantics on
//                   the kx = 0 and kx = N/2 planes, conjugate (fftw3.f:430-437).
//   hipFFT C2R      3D inverse (FMRS_3, fmrs_3.f:54).
//   window          WINDKB2A (bp32f.f:945-1045): crop, x 1/N^3, divide by the
//                   KB window transform, shell mean, sphere mask.
//
// Accumulators stay raw (SYMPLANE0 is never applied in place), so inserts may
// continue after a finalize and partial volumes merge linearly.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "spbp.h"


#include "calib.hpp"
#include "common.hpp"
#include "fft.hpp"
#include "finalize.hpp"
#include "hostmath.hpp"
#include "insert.hpp"
#include "insert_gather.hpp"
#include "project.hpp"
#include "resolve.hpp"
#include "rotate.hpp"

using namespace spbp;

namespace {
// last failure messages: fixed buffers, so that no error path allocates
constexpr int ERR_LEN = 512;
thread_local char g_last_error[ERR_LEN] = {0};
// brick shapes of the LDS-binned strategy (voxels; 3 x 8 B per voxel in LDS)
// {x, y, z, owner-computes}; for owner-computes bricks the LDS region, whose
// core (each sample's owner) is 4 voxels shorter per axis
constexpr int NSHAPES = 8;
constexpr int SHAPES[NSHAPES][4] = {{16, 16, 8, 0},  {16, 8, 8, 0},   {8, 8, 8, 0},    {16, 12, 8, 0},
                                    {16, 16, 10, 0}, {16, 16, 12, 0}, {16, 16, 13, 0}, {16, 16, 13, 1}};
// brick workgroups: 16x16x10, 16x16x12 and 16x16x13 (61.7, 73.8 and 80.2 KB of
// LDS) run two 512-thread groups per CU; the others four (three for 16x16x8)
// 256-thread groups.  The FP64-geometry instantiations need more VGPRs than
// 16 waves per CU allow: those contexts run one 512-thread group per CU, and
// 16x16x10 in place of shared-evaluation 16x16x12 and 16x16x13.
constexpr int NCVALS = 4;
constexpr int CVALS[NCVALS] = {16, 32, 64, 128};   // projections per work item tried by the tuner
}  // namespace

// ------------------------------------------------------------------ context
// RTSF scratch for sub-batches of up to `cap` images of NX^2: the R2C
// spectra and the three derivative spectra per image, the derivative images,
// fp32 copies of BF16 images, and R2C / C2R plans for two batch sizes (the
// full sub-batch and one other).
struct RtsfWork {
    int nx = 0, cap = 0;
    float2* spec = nullptr;   // 4 cap x NX x (NX/2+1)
    float* der = nullptr;     // 3 cap x NX^2
    float* img = nullptr;     // cap x NX^2
    hipfftHandle fwd[2] = {0, 0}, inv[2] = {0, 0};
    int pb[2] = {0, 0};
};

// Re-summation buffers of one finalised volume: the listed voxels, hit and
// term counts and offsets, the terms and their sums, the counters, and a
// pinned word area through which counts come back to the host.
struct Res {
    long long* list = nullptr; size_t list_cap = 0;
    int* cnt = nullptr; size_t cnt_cap = 0;
    long long* off = nullptr; size_t off_cap = 0;
    long long* bsum = nullptr; size_t bsum_cap = 0;
    int2* hits = nullptr; size_t hits_cap = 0;
    int* tcnt = nullptr; size_t tcnt_cap = 0;
    long long* toff = nullptr; size_t toff_cap = 0;
    char* terms = nullptr; size_t terms_cap = 0;
    char* sums = nullptr; size_t sums_cap = 0;
    unsigned long long* ctr = nullptr;   // [listed, changed, inconsistent, -]
    unsigned long long* h = nullptr;     // pinned [4]
};

struct spbp_ctx {
    int nx = 0, N = 0, H = 0, flags = 0, device = 0, batch = 0, nvol = 1;
    size_t nh = 0;
    float2* X[2] = {nullptr, nullptr};
    float* W[2] = {nullptr, nullptr};
    float* Wc[2] = {nullptr, nullptr};   // SPBP_TWOSUM_W: compensation terms of W
    float* tab = nullptr;
    float fltb = 0.f;
    // insertion chunk
    float* padbuf = nullptr;
    float2* spec = nullptr;
    float* vmax = nullptr;    // per chunk image: largest |Re|, |Im| of its spectrum
    Geom* geom_d = nullptr;
    Geom* geom_h = nullptr;   // pinned
    size_t geom_cap = 0;
    hipEvent_t geom_evt = nullptr;
    hipfftHandle fwd = 0;
    int fwd_batch = 0;
    hipfftHandle fwd_tail = 0;
    int tail_batch = 0;
    // projector image synthesis (spbp_gather_images): C2R plans (two batch
    // sizes: a call's full chunks and its tail) and their output
    hipfftHandle c2r = 0, c2r2 = 0;
    int c2r_batch = 0, c2r2_batch = 0;
    float* c2rbuf = nullptr;
    int c2r_cap = 0;
    // own forward transforms (power-of-two N): e^{-2 pi i m / N}, m < N
    float2* tw = nullptr;
    int fft_log2 = 0;   // log2 N, or 0 when hipFFT does the forward transforms
    double* fmean = nullptr;   // per chunk image: PADD2's mean
    // finalisation
    hipfftHandle inv = 0;
    bool final_ok = false;    // every finalisation buffer and the inverse plan exist
    float2* work = nullptr;
    float* real = nullptr;
    float* wkb = nullptr;
    double* psum = nullptr;
    unsigned* pcnt = nullptr;
    float* mean = nullptr;
    int* flag = nullptr;
    // batched finalisation job table (spbp_finalize_batch's first context
    // holds the call's table): device, pinned staging, its upload
    FinalJob* fjobs_d = nullptr;
    FinalJob* fjobs_h = nullptr;
    int fjobs_cap = 0;
    hipEvent_t fjobs_evt = nullptr;
    // page-locked staging of spbp_insert_host's pageable images: two buffers
    // of hstage_cap bytes, filled by host threads while the other uploads
    void* hstage[2] = {nullptr, nullptr};
    size_t hstage_cap = 0;
    // brick strategy
    int strategy = SPBP_STRAT_BRICK;
    int shape = 3;   // index into SHAPES
    int C = 128;
    bool tuned = false;
    bool capturing = false;   // a batch graph is being captured: no host syncs, no timing events
    int graphs = 0;           // live batch graphs that replay work on this context's buffers
    bool filled = false;      // the accumulators hold inserted or imported sums (since the last reset)
    // symmetries (BUILDS): each image is inserted once per matrix, SYMVOL after the window
    int nsym = 1;
    std::vector<float> sm;    // nsym x 9, Fortran (3,3) order
    std::vector<double> smd;  // the same in FP64 (SPBP_GEOM_F64)
    float* sm_d = nullptr;
    float* symtmp = nullptr;  // NX^3
    // RB pre-rotation: rotated chunk and per-image RTSQ parameters
    float* rotbuf = nullptr;
    int rot_cap = 0;
    RtsfParam* fprm_d = nullptr;   // SPBP_FBS: per-image RTSF constants
    RtsfParam* fprm_h = nullptr;   // pinned
    int fprm_cap = 0;
    hipEvent_t fprm_evt = nullptr;
    RtsfWork rtsf;                 // SPBP_FBS: RTSF scratch and transform plans
    RtsqParam* rprm_d = nullptr;
    RtsqParam* rprm_h = nullptr;   // pinned
    int rprm_cap = 0;
    hipEvent_t rprm_evt = nullptr;
    float tune_ms[NSHAPES][NCVALS] = {};   // spbp_tune: ms per projection of each brick plan
    float tune_direct_ms = 0.f;
    float tune_gather_ms = 0.f;
    BrickJob* jobs_d = nullptr;   // brick-kernel job table
    BrickJob* jobs_h = nullptr;   // pinned staging
    int jobs_cap = 0;
    hipEvent_t jobs_evt = nullptr;
    int ncu = 256;
    BrickGrid bg{};
    int* bcnt = nullptr;    // [2 nb] hits per (brick, half)
    int* boff = nullptr;    // [2 nb + 1] list offsets
    int* bioff = nullptr;   // [2 nb + 1] work-item offsets
    int* bcur = nullptr;    // [2 nb] fill cursors
    int2* bscan = nullptr;  // [ceil(2 nb / 1024)] key-scan block totals
    int* bnext = nullptr;   // work-queue head
    float4* bpln = nullptr; // per section: plane normal and half, for the list builders
    size_t bpln_cap = 0;
    int* blist = nullptr;   // projection lists
    int* bikey = nullptr;   // work item -> key
    size_t blist_cap = 0;
    // NRMW2 weight test by the reference's sums: every section inserted since
    // the last reset, in insertion order (valid while the accumulators hold
    // only insertions), its device copy, and the re-summation buffers
    std::vector<Geom> hist;
    bool hist_ok = true;
    Geom* hist_d = nullptr;
    size_t hist_dcap = 0, hist_dn = 0;
    double* tab64 = nullptr;          // FP64 table (FP64 geometry)
    // per voxel, two bits per volume slot (full, half 1, half 2): 0 = GPU sums
    // decide, 1 = zero, 2 = divide
    unsigned char* tmask = nullptr;
    bool tmask_dirty = true;
    Res res[3];                   // re-summation buffers per volume slot
    long long res_stats[3] = {0, 0, 0};
    int stats_slot = -1;          // slot whose counters res[].h receives (-1: res_stats holds all)
    hipEvent_t stats_evt = nullptr;   // their read-back
    // SPBP_TIMING: (start, stop, units) per recorded launch, per kernel kind
    struct Rec { hipEvent_t a, b; long units; };
    std::vector<Rec> recs[7];
    std::vector<hipEvent_t> evpool;
    char err[ERR_LEN] = {0};
};

namespace {

int fail(spbp_ctx* ctx, int code, const char* fmt, ...) {
    char buf[ERR_LEN];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    std::memcpy(g_last_error, buf, sizeof buf);
    if (ctx) std::memcpy(ctx->err, buf, sizeof buf);
    return code;
}

#define HIPCHK(ctx, call)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((ctx), e_ == hipErrorOutOfMemory ? SPBP_ERR_ALLOC : SPBP_ERR_DEVICE,   \
                        "%s: %s", #call, hipGetErrorString(e_));                               \
    } while (0)

#define FFTCHK(ctx, call)                                                                      \
    do {                                                                                       \
        hipfftResult r_ = (call);                                                              \
        if (r_ != HIPFFT_SUCCESS) return fail((ctx), SPBP_ERR_FFT, "%s: hipfft error %d", #call, (int)r_); \
    } while (0)

// Every device allocation of the library goes through dmalloc.  Tests make
// the k-th one from now on fail as if out of memory (spbp_debug_fail_alloc),
// so that the error paths of the lazily allocated buffers can be exercised.
std::atomic<long> g_fail_after{-1};

template <typename T>
hipError_t dmalloc(T** p, size_t bytes) {
    long k = g_fail_after.load();
    while (k >= 0) {
        if (g_fail_after.compare_exchange_weak(k, k - 1)) {
            if (k == 0) { *p = nullptr; return hipErrorOutOfMemory; }
            break;
        }
    }
    return hipMalloc((void**)p, bytes);
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (dev >= 0 && dev != prev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// ---- SPBP_TIMING helpers: events bracket a launch on the stream it runs on
hipEvent_t take_event(spbp_ctx* c) {
    if (!c->evpool.empty()) { hipEvent_t e = c->evpool.back(); c->evpool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
struct Timed {
    spbp_ctx* c;
    int kind;
    long units;
    hipStream_t st;
    hipEvent_t a = nullptr;
    Timed(spbp_ctx* c_, int kind_, long units_, hipStream_t st_) : c(c_), kind(kind_), units(units_), st(st_) {
        if ((c->flags & SPBP_TIMING) && !c->capturing) {
            a = take_event(c);
            if (a) (void)hipEventRecord(a, st);
        }
    }
    ~Timed() {
        if (!a) return;
        hipEvent_t b = take_event(c);
        if (!b) return;
        (void)hipEventRecord(b, st);
        c->recs[kind].push_back({a, b, units});
    }
};

// A captured batch graph (spbp_graph_create) replays launches that read the
// context's chunk buffers, brick lists and forward FFT plans: while one
// exists, nothing may free or replace them.
int refuse_captured(spbp_ctx* c, const char* what) {
    return fail(c, SPBP_ERR_ARG, "%s: a captured batch graph uses this context's buffers "
                "(spbp_graph_destroy it first)", what);
}

// Brick grid of shape `shape` and its per-key arrays.  On failure the
// context holds no brick grid (c->bg.nb = 0) and the next call allocates
// again; the shape and grid are committed only once every array exists.
// Brick grid of shape `shape` over the half grid of edge nx (FP64-geometry
// contexts have no shared-evaluation 16x16x12 / 16x16x13 kernels: 16x16x10).
BrickGrid grid_of_shape(int nx, int flags, int& shape) {
    if ((flags & SPBP_GEOM_F64) && (shape == 5 || shape == 6)) shape = 4;
    const int own = SHAPES[shape][3], halo = own ? 4 : 0;
    const int bx = SHAPES[shape][0] - halo, by = SHAPES[shape][1] - halo, bz = SHAPES[shape][2] - halo;
    BrickGrid g;
    g.bx = bx;
    g.by = by;
    g.bz = bz;
    g.own = own;
    g.nbx = (nx + bx - 1) / bx;
    g.nby = (2 * nx - 1 + by - 1) / by;
    g.nbz = (2 * nx - 1 + bz - 1) / bz;
    g.nb = g.nbx * g.nby * g.nbz;
    return g;
}

int configure_bricks(spbp_ctx* c, int shape) {
    const BrickGrid g = grid_of_shape(c->nx, c->flags, shape);
    if (c->bg.nb > 0 && g.nb == c->bg.nb && shape == c->shape) return SPBP_OK;
    if (c->graphs > 0) return refuse_captured(c, "changing the brick shape");
    c->bg.nb = 0;   // no valid grid until every array below exists
    int* bufs[4] = {c->bcnt, c->boff, c->bioff, c->bcur};
    for (int* q : bufs) if (q) (void)hipFree(q);
    c->bcnt = c->boff = c->bioff = c->bcur = nullptr;
    if (c->bscan) { (void)hipFree(c->bscan); c->bscan = nullptr; }
    if (c->blist) { (void)hipFree(c->blist); c->blist = nullptr; }
    if (c->bikey) { (void)hipFree(c->bikey); c->bikey = nullptr; }
    c->blist_cap = 0;
    const size_t nk = 2 * (size_t)g.nb;
    HIPCHK(c, dmalloc(&c->bcnt, nk * sizeof(int)));
    HIPCHK(c, dmalloc(&c->boff, (nk + 1) * sizeof(int)));
    HIPCHK(c, dmalloc(&c->bioff, (nk + 1) * sizeof(int)));
    HIPCHK(c, dmalloc(&c->bcur, nk * sizeof(int)));
    HIPCHK(c, dmalloc(&c->bscan, ((nk + 1023) / 1024 + 1) * sizeof(int2)));
    c->bg = g;
    c->shape = shape;
    return SPBP_OK;
}

// Projection lists for chunks of up to B projections.  The capacity is
// committed only once both arrays exist.
int ensure_lists(spbp_ctx* c, int B) {
    if (c->bg.nb <= 0) {   // an earlier allocation failed: the last good shape again
        const int rc = configure_bricks(c, c->shape);
        if (rc) return rc;
    }
    const size_t cap = (size_t)B * c->nsym * c->bg.nb;   // B images, nsym sections each
    const size_t ns = (size_t)B * c->nsym;
    if (c->blist && c->bikey && c->blist_cap >= cap && c->bpln && c->bpln_cap >= ns) return SPBP_OK;
    if (c->graphs > 0) return refuse_captured(c, "growing the brick lists");
    c->blist_cap = 0;
    c->bpln_cap = 0;
    if (c->blist) { (void)hipFree(c->blist); c->blist = nullptr; }
    if (c->bikey) { (void)hipFree(c->bikey); c->bikey = nullptr; }
    if (c->bpln) { (void)hipFree(c->bpln); c->bpln = nullptr; }
    HIPCHK(c, dmalloc(&c->blist, cap * sizeof(int)));
    // items <= sum over keys of ceil(cnt / C) <= pairs + keys
    HIPCHK(c, dmalloc(&c->bikey, (cap + 2 * (size_t)c->bg.nb) * sizeof(int)));
    HIPCHK(c, dmalloc(&c->bpln, ns * sizeof(float4)));
    c->blist_cap = cap;
    c->bpln_cap = ns;
    return SPBP_OK;
}

// Chunk buffers (and, for hipFFT transforms, the forward plan) for chunks of
// want_batch projections.  c->fwd_batch is the committed chunk size: zero
// until every buffer of the set exists.
int ensure_chunk(spbp_ctx* c, int want_batch) {
    const bool set = c->spec && c->vmax && (c->fft_log2 ? c->fmean != nullptr : (c->padbuf && c->fwd));
    if (set && c->fwd_batch >= want_batch) return ensure_lists(c, c->fwd_batch);
    if (c->graphs > 0) return refuse_captured(c, "growing the chunk buffers");
    c->fwd_batch = 0;
    if (c->padbuf) { (void)hipFree(c->padbuf); c->padbuf = nullptr; }
    if (c->spec) { (void)hipFree(c->spec); c->spec = nullptr; }
    if (c->fwd) { hipfftDestroy(c->fwd); c->fwd = 0; }
    if (c->fwd_tail) { hipfftDestroy(c->fwd_tail); c->fwd_tail = 0; c->tail_batch = 0; }
    if (c->vmax) { (void)hipFree(c->vmax); c->vmax = nullptr; }
    if (c->fmean) { (void)hipFree(c->fmean); c->fmean = nullptr; }
    const int N = c->N, H = c->H, B = want_batch;
    if (!c->fft_log2) HIPCHK(c, dmalloc(&c->padbuf, (size_t)B * N * N * sizeof(float)));
    HIPCHK(c, dmalloc(&c->spec, (size_t)B * N * H * sizeof(float2)));
    HIPCHK(c, dmalloc(&c->vmax, (size_t)B * sizeof(float)));
    if (c->fft_log2) HIPCHK(c, dmalloc(&c->fmean, (size_t)B * sizeof(double)));
    int rc = ensure_lists(c, B);
    if (rc) return rc;
    if (!c->fft_log2) {   // own transforms need no plan, and the pad buffer is not used
        int n[2] = {N, N};
        int inembed[2] = {N, N};
        int onembed[2] = {N, H};
        hipfftResult r = hipfftPlanMany(&c->fwd, 2, n, inembed, 1, N * N, onembed, 1, N * H, HIPFFT_R2C, B);
        if (r != HIPFFT_SUCCESS) { c->fwd = 0; return fail(c, SPBP_ERR_FFT, "hipfftPlanMany(%d): hipfft error %d", B, (int)r); }
    }
    c->fwd_batch = B;
    return SPBP_OK;
}

hipfftHandle plan_for(spbp_ctx* c, int b, int* rc) {
    *rc = SPBP_OK;
    if (b == c->fwd_batch) return c->fwd;
    if (c->fwd_tail && c->tail_batch == b) return c->fwd_tail;
    if (c->graphs > 0) { *rc = refuse_captured(c, "a new forward FFT plan"); return 0; }
    if (c->fwd_tail) { hipfftDestroy(c->fwd_tail); c->fwd_tail = 0; c->tail_batch = 0; }
    const int N = c->N, H = c->H;
    int n[2] = {N, N};
    int inembed[2] = {N, N};
    int onembed[2] = {N, H};
    hipfftResult r = hipfftPlanMany(&c->fwd_tail, 2, n, inembed, 1, N * N, onembed, 1, N * H, HIPFFT_R2C, b);
    if (r != HIPFFT_SUCCESS) { *rc = fail(c, SPBP_ERR_FFT, "hipfftPlanMany(tail %d): %d", b, (int)r); return 0; }
    c->tail_batch = b;
    return c->fwd_tail;
}

int ensure_geom(spbp_ctx* c, size_t n) {
    if (c->geom_cap >= n) {
        HIPCHK(c, hipEventSynchronize(c->geom_evt));   // previous upload has left the staging buffer
        return SPBP_OK;
    }
    c->geom_cap = 0;   // committed again only once both buffers exist
    if (c->geom_h) { (void)hipEventSynchronize(c->geom_evt); (void)hipHostFree(c->geom_h); c->geom_h = nullptr; }
    if (c->geom_d) { (void)hipFree(c->geom_d); c->geom_d = nullptr; }
    const size_t cap = std::max<size_t>(n, 1024);
    HIPCHK(c, hipHostMalloc((void**)&c->geom_h, cap * sizeof(Geom), hipHostMallocDefault));
    HIPCHK(c, dmalloc(&c->geom_d, cap * sizeof(Geom)));
    c->geom_cap = cap;
    return SPBP_OK;
}

int ensure_final(spbp_ctx* c) {
    if (c->final_ok) return SPBP_OK;
    // a failed earlier attempt may have left part of the set: start over
    void* old[] = {c->work, c->real, c->wkb, c->psum, c->pcnt, c->mean};
    for (void* q : old) if (q) (void)hipFree(q);
    c->work = nullptr; c->real = nullptr; c->wkb = nullptr; c->psum = nullptr; c->pcnt = nullptr; c->mean = nullptr;
    if (c->inv) { hipfftDestroy(c->inv); c->inv = 0; }
    const int N = c->N, nx = c->nx;
    HIPCHK(c, dmalloc(&c->work, c->nh * sizeof(float2)));
    HIPCHK(c, dmalloc(&c->real, (size_t)N * N * N * sizeof(float)));
    std::vector<float> w;
    wkb_table(nx, N, w);
    HIPCHK(c, dmalloc(&c->wkb, w.size() * sizeof(float)));
    HIPCHK(c, hipMemcpy(c->wkb, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    const size_t nb = nblocks((size_t)nx * nx * nx);
    HIPCHK(c, dmalloc(&c->psum, nb * sizeof(double)));
    HIPCHK(c, dmalloc(&c->pcnt, nb * sizeof(unsigned)));
    HIPCHK(c, dmalloc(&c->mean, sizeof(float)));
    if (hipfftPlan3d(&c->inv, N, N, N, HIPFFT_C2R) != HIPFFT_SUCCESS) {
        c->inv = 0;
        return fail(c, SPBP_ERR_FFT, "hipfftPlan3d(%d) failed", N);
    }
    c->final_ok = true;
    return SPBP_OK;
}

// Rotations (and half flags) of nimg projections into c->geom_d, queued on st.
// From dm (nimg x 9 fp32) when given, else from the doc angles (PSI, THETA, PHI).
// With use_sym and symmetries set, image p gets entries p*nsym + k with
// DMS = MATMUL(SM_k, DM) (bp32f.f:866-872), all reading spectrum p.
// Geometry of projection p (its ns symmetry copies) into c->geom_h.  Returns
// 0, or the failure: 1 missing angle, 2 non-finite rotation, 3 half flag not
// 0/1, 4 half 1 without two halves, 5 given rows 1-2 not orthonormal.  Touches only p's entries (the callers
// run it from several threads).
int geom_one(spbp_ctx* c, int p, int ns, const float* dm_in, const double* ang, const int8_t* half, bool exact,
             bool f64) {
    float dm[9];
    double dd[9];
    if (dm_in) {
        std::memcpy(dm, dm_in + 9 * (size_t)p, sizeof dm);
        for (int q = 0; q < 9; ++q) dd[q] = dm[q];
    } else {
        const double* a = ang + 3 * (size_t)p;
        if (!std::isfinite(a[0]) || !std::isfinite(a[1]) || !std::isfinite(a[2])) return 1;
        cang((float)a[2], (float)a[1], (float)a[0], exact, dm);   // CANG(PHI, THETA, PSI)
        if (f64) cang_d(a[2], a[1], a[0], exact, dd);
    }
    for (int q = 0; q < 9; ++q)
        if (!std::isfinite(dm[q])) return 2;
    if (dm_in) {
        // the section's lattice is enumerated along DM's first two rows, which
        // must be orthonormal (a rotation, as CANG builds)
        const double a = (double)dm[0] * dm[0] + (double)dm[1] * dm[1] + (double)dm[2] * dm[2];
        const double b = (double)dm[3] * dm[3] + (double)dm[4] * dm[4] + (double)dm[5] * dm[5];
        const double ab = (double)dm[0] * dm[3] + (double)dm[1] * dm[4] + (double)dm[2] * dm[5];
        if (std::fabs(a - 1.0) > 1e-4 || std::fabs(b - 1.0) > 1e-4 || std::fabs(ab) > 1e-4) return 5;
    }
    int h = 0;
    if (half) {
        if (half[p] != 0 && half[p] != 1) return 3;
        if (half[p] == 1 && c->nvol < 2) return 4;
        h = half[p];
    }
    for (int k = 0; k < ns; ++k) {
        Geom& g = c->geom_h[(size_t)p * ns + k];
        if (ns > 1) sym_matmul(c->sm.data() + 9 * k, dm, g.r);
        else std::memcpy(g.r, dm, sizeof dm);
        g.half = h;
        g.img = p;
        g.pad_ = 0;
        if (f64) {
            double dms[9];
            if (ns > 1) sym_matmul_d(c->smd.data() + 9 * k, dd, dms);
            else std::memcpy(dms, dd, sizeof dms);
            for (int q = 0; q < 6; ++q) g.d[q] = dms[q];
            for (int q = 0; q < 9; ++q) g.r[q] = (float)dms[q];   // for the brick tests
        } else {
            for (int q = 0; q < 6; ++q) g.d[q] = g.r[q];
        }
    }
    return 0;
}

int stage_geom(spbp_ctx* c, int nimg, const float* dm_in, const double* ang, const int8_t* half, hipStream_t st,
               bool use_sym = true) {
    if (!dm_in && !ang) return fail(c, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE (no angles given)");
    const int ns = use_sym ? c->nsym : 1;
    int rc = ensure_geom(c, (size_t)nimg * ns);
    if (rc) return rc;
    const bool exact = (c->flags & SPBP_EXACT_PI) != 0;
    const bool f64 = (c->flags & SPBP_GEOM_F64) != 0;
    // CANG in FP64 costs ~0.1 us per projection: large batches are split over
    // host threads (contiguous ranges; each stops at its first failure, and
    // the first failure in projection order is reported)
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<unsigned>(std::min(hw, 16u), (unsigned)std::max(1, nimg / 2048));
    std::vector<int> bad_p(nt, -1), bad_k(nt, 0);
    auto work = [&](int t) {
        const int p0 = (int)((long long)nimg * t / nt), p1 = (int)((long long)nimg * (t + 1) / nt);
        for (int p = p0; p < p1; ++p) {
            const int k = geom_one(c, p, ns, dm_in, ang, half, exact, f64);
            if (k) { bad_p[t] = p; bad_k[t] = k; return; }
        }
    };
    if (nt > 1) {
        std::vector<std::thread> th;
        th.reserve(nt - 1);
        for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
        work(0);
        for (auto& x : th) x.join();
    } else {
        work(0);
    }
    for (int t = 0; t < nt; ++t) {
        if (bad_p[t] < 0) continue;
        const int p = bad_p[t];
        switch (bad_k[t]) {
            case 1: return fail(c, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE %d", p + 1);
            case 2: return fail(c, SPBP_ERR_ANGLE, "non-finite rotation for image %d", p + 1);
            case 3: return fail(c, SPBP_ERR_ARG, "half flag %d of image %d is not 0/1", half[p], p + 1);
            case 5: return fail(c, SPBP_ERR_ANGLE, "rotation of image %d is not orthonormal", p + 1);
            default: return fail(c, SPBP_ERR_ARG, "half 1 needs SPBP_TWO_HALVES");
        }
    }
    HIPCHK(c, hipMemcpyAsync(c->geom_d, c->geom_h, (size_t)nimg * ns * sizeof(Geom), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipEventRecord(c->geom_evt, st));
    return SPBP_OK;
}

// Insertion targets: half 0 -> (X0, W0, Wc0), half 1 -> (X1, W1, Wc1); W, Wc may be null.
struct Targets {
    float2* X0; float* W0; float2* X1; float* W1; float* Wc0; float* Wc1;
};

// Projections per work item for a launch of nsec sections: the plan's C,
// halved while the launch holds fewer than 8 x C sections (down to 16), so
// the bricks every plane crosses still split into several items.  Measured at
// 128^2 into 256^3 with 256-section launches: 294k proj/s with 16 or 32 per
// item, 212k with 128.
int item_proj(int C, int nsec) {
    int k = C;
    while (k > 16 && 8 * k > nsec) k >>= 1;
    return k;
}

BrickJob brick_job(const spbp_ctx* c, int img0, const Geom* geom_d, int nsec, const Targets& t) {
    BrickJob j;
    j.wscale = WSCALE;
    j.spec = c->spec;
    j.vmax = c->vmax;
    j.img0 = img0;
    j.geom = geom_d;
    j.tab = c->tab;
    j.fltb = c->fltb;
    j.nx = c->nx;
    j.N = c->N;
    j.C = item_proj(c->C, nsec);
    j.bg = c->bg;
    j.cnt = c->bcnt;
    j.off = c->boff;
    j.item_off = c->bioff;
    j.item_key = c->bikey;
    j.list = c->blist;
    j.X0 = t.X0; j.W0 = t.W0; j.X1 = t.X1; j.W1 = t.W1; j.Wc0 = t.Wc0; j.Wc1 = t.Wc1;
    return j;
}

unsigned brick_grid(const spbp_ctx* c) {
    // workgroups resident per CU: the 512-thread shapes (16x16x10, 16x16x12,
    // 16x16x13: 61.7, 73.8 and 80.2 KB of LDS with weights) run two per CU,
    // 16x16x8 (49.2 KB) three, the other 256-thread shapes four (VGPR-bound:
    // 4 waves per SIMD); the FP64-geometry instantiations need ~150-170
    // VGPRs: one 512-thread group or three 256-thread groups per CU
    if (c->shape >= 4) return (unsigned)c->ncu * ((c->flags & SPBP_GEOM_F64) ? 1u : 2u);
    return (unsigned)c->ncu * ((c->shape == 0 || (c->flags & SPBP_GEOM_F64)) ? 3u : 4u);   // 16x12x8: 36.9 KB, 4 fit
}

template <int BX, int BY, int BZ, typename P, int NT = TPB, bool OWN = false>
void launch_shape(bool wgt, unsigned grid, const BrickJob* table_d, int n, int* counter, hipStream_t st) {
    if (wgt) insert_brick_kernel<BX, BY, BZ, P, NT, OWN, true><<<grid, NT, 0, st>>>(table_d, n, counter);
    else insert_brick_kernel<BX, BY, BZ, P, NT, OWN, false><<<grid, NT, 0, st>>>(table_d, n, counter);
}

// One persistent brick-kernel launch over the device job table table_d[0..n),
// all jobs with c's brick shape; `counter` (the queue head) must be zero on st.
// wgt false: no job accumulates weights (W0 and W1 null in every job).
int launch_bricks_dev(spbp_ctx* c, const BrickJob* table_d, int n, int* counter, hipStream_t st, bool wgt = true) {
    const unsigned grid = brick_grid(c);
    const bool f64 = (c->flags & SPBP_GEOM_F64) != 0;
    switch (c->shape * 2 + (f64 ? 1 : 0)) {
        case 0: launch_shape<16, 16, 8, float>(wgt, grid, table_d, n, counter, st); break;
        case 1: launch_shape<16, 16, 8, double>(wgt, grid, table_d, n, counter, st); break;
        case 2: launch_shape<16, 8, 8, float>(wgt, grid, table_d, n, counter, st); break;
        case 3: launch_shape<16, 8, 8, double>(wgt, grid, table_d, n, counter, st); break;
        case 4: launch_shape<8, 8, 8, float>(wgt, grid, table_d, n, counter, st); break;
        case 5: launch_shape<8, 8, 8, double>(wgt, grid, table_d, n, counter, st); break;
        case 6: launch_shape<16, 12, 8, float>(wgt, grid, table_d, n, counter, st); break;
        case 7: launch_shape<16, 12, 8, double>(wgt, grid, table_d, n, counter, st); break;
        case 8: launch_shape<16, 16, 10, float, 2 * TPB>(wgt, grid, table_d, n, counter, st); break;
        case 9: launch_shape<16, 16, 10, double, 2 * TPB>(wgt, grid, table_d, n, counter, st); break;
        case 10: launch_shape<16, 16, 12, float, 2 * TPB>(wgt, grid, table_d, n, counter, st); break;
        case 12: launch_shape<16, 16, 13, float, 2 * TPB>(wgt, grid, table_d, n, counter, st); break;
        case 14: launch_shape<16, 16, 13, float, 2 * TPB, true>(wgt, grid, table_d, n, counter, st); break;
        case 15: launch_shape<16, 16, 13, double, 2 * TPB, true>(wgt, grid, table_d, n, counter, st); break;
        default: return fail(c, SPBP_ERR_ARG, "brick shape %d has no FP64-geometry kernel", c->shape);
    }
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

// Same with a host job table, staged through c's pinned buffer.
int launch_bricks(spbp_ctx* c, const BrickJob* jh, int n, int* counter, hipStream_t st, bool wgt = true) {
    if (n < 1 || n > MAX_JOBS) return fail(c, SPBP_ERR_ARG, "%d jobs in one launch (1..%d)", n, MAX_JOBS);
    if (c->jobs_cap < n) {
        if (c->jobs_h) { (void)hipEventSynchronize(c->jobs_evt); (void)hipHostFree(c->jobs_h); c->jobs_h = nullptr; }
        if (c->jobs_d) { (void)hipFree(c->jobs_d); c->jobs_d = nullptr; }
        HIPCHK(c, hipHostMalloc((void**)&c->jobs_h, MAX_JOBS * sizeof(BrickJob), hipHostMallocDefault));
        HIPCHK(c, dmalloc(&c->jobs_d, MAX_JOBS * sizeof(BrickJob)));
        c->jobs_cap = MAX_JOBS;
    } else {
        HIPCHK(c, hipEventSynchronize(c->jobs_evt));   // the previous upload has left the staging buffer
    }
    std::memcpy(c->jobs_h, jh, n * sizeof(BrickJob));
    HIPCHK(c, hipMemcpyAsync(c->jobs_d, c->jobs_h, n * sizeof(BrickJob), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipEventRecord(c->jobs_evt, st));
    return launch_bricks_dev(c, c->jobs_d, n, counter, st, wgt);
}

// Per-(brick, half) projection lists of b projections (geometry at geom_d);
// also zeroes c's queue head.
int build_lists(spbp_ctx* c, const Geom* geom_d, int b, hipStream_t st) {
    if (c->bg.nb <= 0 || !c->bcnt || !c->blist || !c->bikey || c->blist_cap < (size_t)b * c->bg.nb || !c->bpln ||
        c->bpln_cap < (size_t)b)
        return fail(c, SPBP_ERR_ARG, "brick lists not allocated for %d sections", b);
    const int nkeys = 2 * c->bg.nb;
    const int nsb = ((c->bg.nbx + 1) / 2) * ((c->bg.nby + 1) / 2) * ((c->bg.nbz + 1) / 2);   // superbricks
    const dim3 lgrid(nsb, (b + 1023) / 1024);
    Timed tm(c, SPBP_K_BIN, b, st);
    fill((float*)c->bcnt, nkeys, 0.0f, st);   // zero bits
    plane_pack_kernel<<<nblocks(b), TPB, 0, st>>>(geom_d, b, c->bpln);
    brick_count_kernel<<<lgrid, TPB, 0, st>>>(c->bpln, b, c->bg, c->nx, c->bcnt);
    const int kb = (nkeys + 1023) / 1024;
    key_scan_blocks_kernel<<<kb, 1024, 0, st>>>(c->bcnt, nkeys, item_proj(c->C, b), c->boff, c->bioff, c->bscan);
    key_scan_sums_kernel<<<1, 1024, 0, st>>>(c->bscan, kb, nkeys, c->boff, c->bioff, c->bnext);
    key_scan_add_kernel<<<kb, 1024, 0, st>>>(nkeys, c->bscan, c->boff, c->bioff, c->bcur);
    brick_fill_kernel<<<lgrid, TPB, 0, st>>>(c->bpln, b, c->bg, c->nx, c->boff, c->bcur, c->blist);
    brick_itemkey_kernel<<<nblocks(nkeys), TPB, 0, st>>>(c->bioff, nkeys, c->bikey);
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

// ng sections (geometry at geom_d) of images img0.. whose spectra are in c->spec.
int launch_direct(spbp_ctx* c, int img0, const Geom* geom_d, int ng, const Targets& t, hipStream_t st, long units) {
    Timed tm(c, SPBP_K_INSERT, units, st);
    for (int g0 = 0; g0 < ng; g0 += 65535) {   // grid.y limit
        dim3 grid(nblocks((size_t)c->H * c->N), std::min(65535, ng - g0));
        if (c->flags & SPBP_GEOM_F64)
            insert_direct_kernel<double><<<grid, TPB, 0, st>>>(c->spec, img0, geom_d + g0, c->nx, c->N, c->tab, c->fltb,
                                                               t.X0, t.W0, t.X1, t.W1, t.Wc0, t.Wc1);
        else
            insert_direct_kernel<float><<<grid, TPB, 0, st>>>(c->spec, img0, geom_d + g0, c->nx, c->N, c->tab, c->fltb,
                                                              t.X0, t.W0, t.X1, t.W1, t.Wc0, t.Wc1);
    }
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

// The same with the voxel-driven gather (reference-order sums).
int launch_gather(spbp_ctx* c, int img0, const Geom* geom_d, int ng, const Targets& t, hipStream_t st, long units) {
    Timed tm(c, SPBP_K_INSERT, units, st);
    const dim3 grid((c->H + GBX - 1) / GBX, (c->N + GBY - 1) / GBY, (c->N + GBZ - 1) / GBZ);
    const bool f64 = (c->flags & SPBP_GEOM_F64) != 0, wgt = t.W0 != nullptr, two = t.Wc0 != nullptr;
#define SPBP_GATHER(P, W, T)                                                                                          \
    insert_gather_kernel<P, W, T><<<grid, GTPB, 0, st>>>(c->spec, img0, geom_d, ng, c->nx, c->N, c->tab, c->fltb,   \
                                                       t.X0, t.W0, t.X1, t.W1, t.Wc0, t.Wc1)
    if (f64) {
        if (!wgt) SPBP_GATHER(double, false, false);
        else if (two) SPBP_GATHER(double, true, true);
        else SPBP_GATHER(double, true, false);
    } else {
        if (!wgt) SPBP_GATHER(float, false, false);
        else if (two) SPBP_GATHER(float, true, true);
        else SPBP_GATHER(float, true, false);
    }
#undef SPBP_GATHER
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

// Strategies that run their own kernel per context (not the brick queue).
bool own_kernel(const spbp_ctx* c) { return c->strategy == SPBP_STRAT_DIRECT || c->strategy == SPBP_STRAT_GATHER; }

int launch_own(spbp_ctx* c, int img0, const Geom* geom_d, int ng, const Targets& t, hipStream_t st, long units) {
    return c->strategy == SPBP_STRAT_GATHER ? launch_gather(c, img0, geom_d, ng, t, st, units)
                                            : launch_direct(c, img0, geom_d, ng, t, st, units);
}

// ONELINE over ng sections (geometry at geom_d) of images img0.. whose raw
// R2C spectra are in c->spec, with the context's strategy; `units` = images.
int run_insert(spbp_ctx* c, int img0, const Geom* geom_d, int ng, const Targets& t, hipStream_t st, long units) {
    if (!c->spec || !c->vmax || ng > c->fwd_batch * c->nsym)
        return fail(c, SPBP_ERR_ARG, "chunk buffers not allocated for %d sections", ng);
    if (own_kernel(c)) return launch_own(c, img0, geom_d, ng, t, st, units);
    int rc = build_lists(c, geom_d, ng, st);
    if (rc) return rc;
    const BrickJob job = brick_job(c, img0, geom_d, ng, t);
    Timed tm(c, SPBP_K_INSERT, units, st);
    return launch_bricks(c, &job, 1, c->bnext, st, t.W0 != nullptr || t.W1 != nullptr);
}

// PADD2 + batched forward FFT of b images into c->spec.
// The own two-pass R2C of b zero-padded images minus c->fmean[i] into c->spec
// (power-of-two N); c->vmax receives each spectrum's largest |Re|, |Im| and
// must be zero on entry.
int own_fft(spbp_ctx* c, const void* src, int dtype, int b, hipStream_t st) {
    const size_t lds = fft_lds_bytes(c->N), rlds = fft_lds_bytes(c->N, FFT_ROW_ELEMS);
    const int M = std::min(FFT_ROW_ELEMS / c->N, c->nx / 2);   // >= 1: own transforms need N <= FFT_ROW_ELEMS
    const int cw = FFT_ELEMS / c->N;
    const size_t npix = (size_t)c->nx * c->nx, nspec = (size_t)c->N * c->H;
    const size_t esz = dtype == SPBP_BF16 ? 2 : 4;
    // sub-batches whose row-pass output (~256 MiB, the size of MI355X's
    // last-level cache) is still there when the column pass reads it back.
    // Measured (ms per 20,000 images of 128^2 / 8,000 of 256^2, both passes):
    // 64 MiB 4.75 / 9.00, 128 MiB 4.31 / 8.34, 256 MiB 4.22 / 8.15,
    // 512 MiB 4.04 / 8.64 (fewer launches, but at 256^2 the output no longer fits)
    const int sb = (int)std::max<size_t>(1, ((size_t)256 << 20) / ((size_t)c->nx * c->H * sizeof(float2)));
    for (int s0 = 0; s0 < b; s0 += sb) {
        const int n = std::min(sb, b - s0);
        const void* im = (const char*)src + s0 * npix * esz;
        float2* sp = c->spec + s0 * nspec;
        {
            Timed tm(c, SPBP_K_PAD, n, st);
            const dim3 rgrid((c->nx / 2 + M - 1) / M, n);
            if (dtype == SPBP_BF16)
                row_fft_kernel<uint16_t><<<rgrid, TPB, rlds, st>>>((const uint16_t*)im, c->nx, c->N, c->fft_log2, c->tw,
                                                                  c->fmean + s0, sp);
            else
                row_fft_kernel<float><<<rgrid, TPB, rlds, st>>>((const float*)im, c->nx, c->N, c->fft_log2, c->tw,
                                                               c->fmean + s0, sp);
        }
        {
            Timed tm(c, SPBP_K_FFT2, n, st);
            col_fft_kernel<<<dim3((c->H + cw - 1) / cw, n), FFT_COL_NT, lds, st>>>(c->nx, c->N, c->fft_log2, c->tw, sp,
                                                                           c->vmax + s0);
        }
    }
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

int pad_fft(spbp_ctx* c, const void* src, int dtype, int b, hipStream_t st) {
    if (c->fft_log2) {
        {
            Timed tm(c, SPBP_K_PAD, b, st);   // the means of the whole chunk (one block per image)
            if (dtype == SPBP_BF16)
                pad_mean_kernel<uint16_t><<<b, TPB, 0, st>>>((const uint16_t*)src, c->nx, c->fft_log2 - 1, c->fmean, c->vmax);
            else
                pad_mean_kernel<float><<<b, TPB, 0, st>>>((const float*)src, c->nx, c->fft_log2 - 1, c->fmean, c->vmax);
        }
        return own_fft(c, src, dtype, b, st);
    }
    {
        Timed tm(c, SPBP_K_PAD, b, st);
        if (dtype == SPBP_BF16)
            pad_kernel<uint16_t><<<b, TPB, 0, st>>>((const uint16_t*)src, c->nx, c->N, c->padbuf);
        else
            pad_kernel<float><<<b, TPB, 0, st>>>((const float*)src, c->nx, c->N, c->padbuf);
    }
    HIPCHK(c, hipGetLastError());
    int rc = SPBP_OK;
    hipfftHandle plan = plan_for(c, b, &rc);
    if (rc) return rc;
    FFTCHK(c, hipfftSetStream(plan, st));
    {
        Timed tm(c, SPBP_K_FFT2, b, st);
        FFTCHK(c, hipfftExecR2C(plan, (hipfftReal*)c->padbuf, (hipfftComplex*)c->spec));
    }
    spec_absmax_kernel<<<b, TPB, 0, st>>>(c->spec, c->N * c->H, c->vmax);
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

int check_images(spbp_ctx* c, const void* imgs, int dtype, int nimg, hipStream_t st) {
    if (dtype != SPBP_F32 && dtype != SPBP_BF16) return fail(c, SPBP_ERR_ARG, "unknown image type %d", dtype);
    if (!(c->flags & SPBP_CHECK_FINITE)) return SPBP_OK;
    if (c->capturing) return fail(c, SPBP_ERR_ARG, "SPBP_CHECK_FINITE cannot be used in a captured batch");
    const size_t n = (size_t)nimg * c->nx * c->nx;
    fill((float*)c->flag, 1, 0.0f, st);   // zero bits
    if (dtype == SPBP_BF16)
        finite_kernel<uint16_t><<<nblocks(n), TPB, 0, st>>>((const uint16_t*)imgs, n, c->flag);
    else
        finite_kernel<float><<<nblocks(n), TPB, 0, st>>>((const float*)imgs, n, c->flag);
    int f = 0;
    HIPCHK(c, hipMemcpyAsync(&f, c->flag, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (f) return fail(c, SPBP_ERR_NONFINITE, "NaN or Inf in the images");
    return SPBP_OK;
}

// ------------------------------------------------------------------ RTSF (FBS ON)
void rtsf_release(RtsfWork& w) {
    for (int k = 0; k < 2; ++k) {
        if (w.fwd[k]) hipfftDestroy(w.fwd[k]);
        if (w.inv[k]) hipfftDestroy(w.inv[k]);
        w.fwd[k] = w.inv[k] = 0;
        w.pb[k] = 0;
    }
    void* bufs[] = {w.spec, w.der, w.img};
    for (void* q : bufs) if (q) (void)hipFree(q);
    w.spec = nullptr; w.der = nullptr; w.img = nullptr;
    w.nx = w.cap = 0;
}

// Images per RTSF sub-batch: about 256 MiB of scratch.
int rtsf_cap(int nx, int want) {
    const size_t per = (size_t)nx * (nx / 2 + 1) * 8 * 4 + (size_t)nx * nx * 4 * 4;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)want, ((size_t)256 << 20) / per));
}

// Scratch for sub-batches of `cap` images (committed only once all of it exists).
int rtsf_ensure(spbp_ctx* c, RtsfWork& w, int nx, int cap) {
    if (w.nx == nx && w.cap >= cap && w.spec && w.der && w.img) return SPBP_OK;
    rtsf_release(w);
    const size_t H = nx / 2 + 1;
    float2* spec = nullptr;
    float *der = nullptr, *img = nullptr;
    if (dmalloc(&spec, 4 * (size_t)cap * nx * H * sizeof(float2)) != hipSuccess ||
        dmalloc(&der, 3 * (size_t)cap * nx * nx * sizeof(float)) != hipSuccess ||
        dmalloc(&img, (size_t)cap * nx * nx * sizeof(float)) != hipSuccess) {
        if (spec) (void)hipFree(spec);
        if (der) (void)hipFree(der);
        return fail(c, SPBP_ERR_ALLOC, "RTSF scratch for %d images of %d^2", cap, nx);
    }
    w.spec = spec; w.der = der; w.img = img;
    w.nx = nx;
    w.cap = cap;
    return SPBP_OK;
}

// R2C / C2R plans for a sub-batch of b images: slot 0 holds the full
// sub-batch's, slot 1 any other size.
int rtsf_plans(spbp_ctx* c, RtsfWork& w, int b, hipfftHandle* fwd, hipfftHandle* inv) {
    int k = b == w.pb[0] ? 0 : (b == w.pb[1] ? 1 : -1);
    if (k < 0) {
        k = w.pb[0] == 0 ? 0 : 1;
        if (w.fwd[k]) hipfftDestroy(w.fwd[k]);
        if (w.inv[k]) hipfftDestroy(w.inv[k]);
        w.fwd[k] = w.inv[k] = 0;
        w.pb[k] = 0;
        int n[2] = {w.nx, w.nx};
        if (hipfftPlanMany(&w.fwd[k], 2, n, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_R2C, b) != HIPFFT_SUCCESS) {
            w.fwd[k] = 0;
            return fail(c, SPBP_ERR_FFT, "RTSF forward plan (%d^2 x %d)", w.nx, b);
        }
        if (hipfftPlanMany(&w.inv[k], 2, n, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_C2R, 3 * b) != HIPFFT_SUCCESS) {
            hipfftDestroy(w.fwd[k]);
            w.fwd[k] = w.inv[k] = 0;
            return fail(c, SPBP_ERR_FFT, "RTSF inverse plan (%d^2 x %d)", w.nx, 3 * b);
        }
        w.pb[k] = b;
    }
    *fwd = w.fwd[k];
    *inv = w.inv[k];
    return SPBP_OK;
}

__global__ __launch_bounds__(TPB) void widen_kernel(const uint16_t* __restrict__ in, size_t n, float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = load_img(in, i);
}

// RTSF of nimg NX x NX images (fp32 / BF16) with per-image constants prm_d
// (device) into out (fp32), on st, in sub-batches of w.cap images.
int rtsf_run(spbp_ctx* c, RtsfWork& w, const void* imgs, int dtype, int nimg, int nx, const RtsfParam* prm_d,
             float* out, hipStream_t st) {
    const size_t npix = (size_t)nx * nx, H = nx / 2 + 1, nsp = (size_t)nx * H;
    const size_t esz = dtype == SPBP_BF16 ? 2 : 4;
    for (int s0 = 0; s0 < nimg; s0 += w.cap) {
        const int b = std::min(w.cap, nimg - s0);
        const void* src = (const char*)imgs + (size_t)s0 * npix * esz;
        hipfftHandle fwd, inv;
        int rc = rtsf_plans(c, w, b, &fwd, &inv);
        if (rc) return rc;
        // the transform reads a copy: the caller's images stay untouched whatever the FFT library does with its input
        if (dtype == SPBP_BF16)
            widen_kernel<<<nblocks((size_t)b * npix), TPB, 0, st>>>((const uint16_t*)src, (size_t)b * npix, w.img);
        else
            HIPCHK(c, hipMemcpyAsync(w.img, src, (size_t)b * npix * sizeof(float), hipMemcpyDeviceToDevice, st));
        FFTCHK(c, hipfftSetStream(fwd, st));
        FFTCHK(c, hipfftExecR2C(fwd, (hipfftReal*)w.img, (hipfftComplex*)w.spec));
        float2* dspec = w.spec + (size_t)b * nsp;
        fbs_deriv_kernel<<<nblocks((size_t)b * nsp), TPB, 0, st>>>(w.spec, nx, b, dspec);
        HIPCHK(c, hipGetLastError());
        FFTCHK(c, hipfftSetStream(inv, st));
        FFTCHK(c, hipfftExecC2R(inv, (hipfftComplex*)dspec, (hipfftReal*)w.der));
        const dim3 grid(nblocks(npix), b);
        if (dtype == SPBP_BF16)
            rtsf_kernel<uint16_t><<<grid, TPB, 0, st>>>((const uint16_t*)src, w.der, prm_d + s0, nx, out + (size_t)s0 * npix);
        else
            rtsf_kernel<float><<<grid, TPB, 0, st>>>((const float*)src, w.der, prm_d + s0, nx, out + (size_t)s0 * npix);
        HIPCHK(c, hipGetLastError());
    }
    return SPBP_OK;
}

// RTSF constants of nimg images (checked finite) into h (host).
int rtsf_params(spbp_ctx* c, int nx, int nimg, const float* inplane, RtsfParam* h) {
    for (int p = 0; p < nimg; ++p) {
        const float* q = inplane + 3 * (size_t)p;
        if (!std::isfinite(q[0]) || !std::isfinite(q[1]) || !std::isfinite(q[2]))
            return fail(c, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE %d (in-plane angle or shift)", p + 1);
        float v[12];
        rtsf_param(nx, q[0], q[1], q[2], v);
        h[p] = RtsfParam{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9], v[10], v[11]};
    }
    return SPBP_OK;
}

// SPBP_FBS: the RTSF constants of nimg images into c->fprm_d, queued on st.
int stage_rtsf(spbp_ctx* c, int nimg, const float* inplane, hipStream_t st) {
    if (!c->fprm_evt) HIPCHK(c, hipEventCreateWithFlags(&c->fprm_evt, hipEventDisableTiming));
    if (c->fprm_cap < nimg) {
        c->fprm_cap = 0;   // committed again only once both buffers exist
        if (c->fprm_h) { (void)hipEventSynchronize(c->fprm_evt); (void)hipHostFree(c->fprm_h); c->fprm_h = nullptr; }
        if (c->fprm_d) { (void)hipFree(c->fprm_d); c->fprm_d = nullptr; }
        const int cap = std::max(nimg, 1024);
        HIPCHK(c, hipHostMalloc((void**)&c->fprm_h, cap * sizeof(RtsfParam), hipHostMallocDefault));
        HIPCHK(c, dmalloc(&c->fprm_d, cap * sizeof(RtsfParam)));
        c->fprm_cap = cap;
    } else {
        HIPCHK(c, hipEventSynchronize(c->fprm_evt));
    }
    int rc = rtsf_params(c, c->nx, nimg, inplane, c->fprm_h);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->fprm_d, c->fprm_h, nimg * sizeof(RtsfParam), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipEventRecord(c->fprm_evt, st));
    return SPBP_OK;
}

// RTSQ parameters of nimg images (in-plane angle, x shift, y shift per image)
// into c->rprm_d, queued on st.
int stage_rtsq(spbp_ctx* c, int nimg, const float* inplane, hipStream_t st) {
    // the event exists before any staging buffer does
    if (!c->rprm_evt) HIPCHK(c, hipEventCreateWithFlags(&c->rprm_evt, hipEventDisableTiming));
    if (c->rprm_cap < nimg) {
        c->rprm_cap = 0;   // committed again only once both buffers exist
        if (c->rprm_h) { (void)hipEventSynchronize(c->rprm_evt); (void)hipHostFree(c->rprm_h); c->rprm_h = nullptr; }
        if (c->rprm_d) { (void)hipFree(c->rprm_d); c->rprm_d = nullptr; }
        const int cap = std::max(nimg, 1024);
        HIPCHK(c, hipHostMalloc((void**)&c->rprm_h, cap * sizeof(RtsqParam), hipHostMallocDefault));
        HIPCHK(c, dmalloc(&c->rprm_d, cap * sizeof(RtsqParam)));
        c->rprm_cap = cap;
    } else {
        HIPCHK(c, hipEventSynchronize(c->rprm_evt));
    }
    for (int p = 0; p < nimg; ++p) {
        const float* q = inplane + 3 * (size_t)p;
        if (!std::isfinite(q[0]) || !std::isfinite(q[1]) || !std::isfinite(q[2]))
            return fail(c, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE %d (in-plane angle or shift)", p + 1);
        float v[4];
        rtsq_param(c->nx, q[0], q[1], q[2], v);
        c->rprm_h[p] = RtsqParam{v[0], v[1], v[2], v[3]};
    }
    HIPCHK(c, hipMemcpyAsync(c->rprm_d, c->rprm_h, nimg * sizeof(RtsqParam), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipEventRecord(c->rprm_evt, st));
    return SPBP_OK;
}

// Append the n sections just staged (c->geom_h) to the insertion history.
void record_history(spbp_ctx* c, size_t n) {
    if (c->hist_ok) c->hist.insert(c->hist.end(), c->geom_h, c->geom_h + n);
}

int insert_impl(spbp_ctx* c, const void* imgs, int dtype, int nimg, const float* dm_in,
                const double* ang, const int8_t* half, hipStream_t st, const float* inplane = nullptr) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (nimg < 0) return fail(c, SPBP_ERR_ARG, "negative image count %d", nimg);
    if (nimg == 0) return SPBP_OK;
    if (!imgs) return fail(c, SPBP_ERR_ARG, "NULL image pointer");
    DeviceGuard dg(c->device);
    int rc = check_images(c, imgs, dtype, nimg, st);
    if (rc) return rc;
    if ((rc = stage_geom(c, nimg, dm_in, ang, half, st))) return rc;
    const bool fbs = inplane && (c->flags & SPBP_FBS);
    if (inplane && (rc = fbs ? stage_rtsf(c, nimg, inplane, st) : stage_rtsq(c, nimg, inplane, st))) return rc;
    if ((rc = ensure_chunk(c, std::min(c->batch, nimg)))) return rc;
    if (fbs && (rc = rtsf_ensure(c, c->rtsf, c->nx, rtsf_cap(c->nx, c->fwd_batch)))) return rc;
    const size_t esz = dtype == SPBP_BF16 ? 2 : 4;
    const size_t npix = (size_t)c->nx * c->nx;
    if (inplane && c->rot_cap < c->fwd_batch) {
        c->rot_cap = 0;
        if (c->rotbuf) { (void)hipFree(c->rotbuf); c->rotbuf = nullptr; }
        HIPCHK(c, dmalloc(&c->rotbuf, (size_t)c->fwd_batch * npix * sizeof(float)));
        c->rot_cap = c->fwd_batch;
    }
    const Targets t{c->X[0], c->W[0], c->X[1], c->W[1], c->Wc[0], c->Wc[1]};
    const int ns = c->nsym;
    c->filled = true;
    for (int p0 = 0; p0 < nimg; p0 += c->fwd_batch) {
        const int b = std::min(c->fwd_batch, nimg - p0);
        const char* src = (const char*)imgs + (size_t)p0 * npix * esz;
        if (fbs) {   // RTSF before PADD2 (bp32f.f:811-816)
            if ((rc = rtsf_run(c, c->rtsf, src, dtype, b, c->nx, c->fprm_d + p0, c->rotbuf, st))) return rc;
            if ((rc = pad_fft(c, c->rotbuf, SPBP_F32, b, st))) return rc;
        } else if (inplane) {   // RTSQ before PADD2 (bp32f.f:818-822)
            const dim3 grid(nblocks(npix), b);
            if (dtype == SPBP_BF16)
                rtsq_kernel<uint16_t><<<grid, TPB, 0, st>>>((const uint16_t*)src, c->rprm_d + p0, c->nx, c->rotbuf);
            else
                rtsq_kernel<float><<<grid, TPB, 0, st>>>((const float*)src, c->rprm_d + p0, c->nx, c->rotbuf);
            HIPCHK(c, hipGetLastError());
            if ((rc = pad_fft(c, c->rotbuf, SPBP_F32, b, st))) return rc;
        } else if ((rc = pad_fft(c, src, dtype, b, st))) {
            return rc;
        }
        if ((rc = run_insert(c, p0, c->geom_d + (size_t)p0 * ns, b * ns, t, st, b))) { c->hist_ok = false; return rc; }
    }
    record_history(c, (size_t)nimg * ns);
    return SPBP_OK;
}

// Choose brick shape s for c (and its best C for that shape when tuned).
int apply_shape(spbp_ctx* c, int s) {
    int rc = configure_bricks(c, s);
    if (rc) return rc;
    if (c->tuned) {
        int best = 0;
        for (int k = 1; k < NCVALS; ++k) if (c->tune_ms[s][k] < c->tune_ms[s][best]) best = k;
        c->C = CVALS[best];
    }
    return c->fwd_batch > 0 ? ensure_lists(c, c->fwd_batch) : SPBP_OK;
}

// Time (ms, best of `reps`) one insertion plan over the b projections in
// c->spec into the scratch targets t.
int time_plan(spbp_ctx* c, const Geom* geom_d, int b, const Targets& t, hipStream_t st, int reps, float* ms) {
    const int ng = b * c->nsym;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    const int saved = c->flags;
    c->flags &= ~SPBP_TIMING;
    int rc = run_insert(c, 0, geom_d, ng, t, st, b);   // warm-up
    float best = 3.0e38f;
    for (int r = 0; r < reps && rc == SPBP_OK; ++r) {
        (void)hipEventRecord(e0, st);
        rc = run_insert(c, 0, geom_d, ng, t, st, b);
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) { rc = fail(c, SPBP_ERR_DEVICE, "tuning run failed"); break; }
        float e = 0.f;
        (void)hipEventElapsedTime(&e, e0, e1);
        best = std::min(best, e);
    }
    c->flags = saved;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = best;
    return rc;
}

Acc make_acc(const spbp_ctx* c, int which) {
    Acc a;
    a.X[0] = a.X[1] = nullptr;
    a.W[0] = a.W[1] = nullptr;
    a.Wc[0] = a.Wc[1] = nullptr;
    if (which < 0) {
        a.n = c->nvol;
        for (int h = 0; h < c->nvol; ++h) { a.X[h] = c->X[h]; a.W[h] = c->W[h]; a.Wc[h] = c->Wc[h]; }
    } else {
        a.n = 1;
        a.X[0] = c->X[which];
        a.W[0] = c->W[which];
        a.Wc[0] = c->Wc[which];
    }
    return a;
}

// ------------------------------------------------------------------ NRMW2 weight test
// Half-width of the band around 0.1 whose voxels are re-summed: 2^-11 of 0.1.
// A sequential fp32 sum of n positive terms is within (n - 1) 2^-24 of the
// exact sum, so the decisions are exact for voxels of up to ~4000 weight
// terms, and in practice (errors grow like sqrt(n)) for far more.
constexpr double THRESH_BAND = 0.1 * (1.0 / 2048.0);

template <typename T>
int grow(spbp_ctx* c, T*& p, size_t& cap, size_t n, const char* what) {
    if (n <= cap && p) return SPBP_OK;
    if (p) { (void)hipFree(p); p = nullptr; cap = 0; }
    const size_t m = std::max<size_t>(std::max<size_t>(n, 2 * cap), 1024);
    if (dmalloc((void**)&p, m * sizeof(T)) != hipSuccess) { p = nullptr; return fail(c, SPBP_ERR_ALLOC, "%s: %zu bytes", what, m * sizeof(T)); }
    cap = m;
    return SPBP_OK;
}

// out[0..n] = exclusive scan of n counts (out[n] = total), on st.
template <typename C>
int scan_counts(spbp_ctx* c, Res& r, const C* in, long long n, long long* out, hipStream_t st) {
    const int nb = (int)((n + 1023) / 1024);
    int rc = grow(c, r.bsum, r.bsum_cap, (size_t)nb + 1, "scan");
    if (rc) return rc;
    if (nb > 0) scan_blocks_kernel<C><<<nb, 1024, 0, st>>>(in, n, out, r.bsum);
    scan_sums_kernel<<<1, 1024, 0, st>>>(r.bsum, nb);
    if (nb > 0) scan_add_kernel<<<nb, 1024, 0, st>>>(out, n, r.bsum, nb);
    else HIPCHK(c, hipMemcpyAsync(out, r.bsum, sizeof(long long), hipMemcpyDeviceToDevice, st));
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

// History on the device (only what is new since the last upload).
int upload_history(spbp_ctx* c, hipStream_t st) {
    const size_t nhist = c->hist.size();
    int rc;
    if (c->hist_dn > nhist) c->hist_dn = 0;
    if (nhist > c->hist_dcap) {
        if ((rc = grow(c, c->hist_d, c->hist_dcap, nhist, "insertion history"))) return rc;
        c->hist_dn = 0;
    }
    if (nhist > c->hist_dn)
        HIPCHK(c, hipMemcpyAsync(c->hist_d + c->hist_dn, c->hist.data() + c->hist_dn, (nhist - c->hist_dn) * sizeof(Geom),
                                 hipMemcpyHostToDevice, st));
    c->hist_dn = nhist;
    return SPBP_OK;
}

// One volume on its way through a finalisation: `which` (-1 all halves, else
// one half) of context c, with the re-summation buffers and mask bits of
// volume slot `slot` (0 full, 1 half 1, 2 half 2).  The counts that size each
// next re-summation phase come back through res().h[0].
struct Round {
    spbp_ctx* c = nullptr;
    int which = -1, slot = 0, wave = 0;
    float* out = nullptr;
    Acc a{};
    bool resolve = false;   // the reference's sums decide NRMW2's test near 0.1
    long long nv = 0, nhits = 0, nterms = 0;
    int nchunks = 0, hmask = 0;
    Res& res() const { return c->res[slot]; }
};

Round make_round(spbp_ctx* c, int which, int slot) {
    Round R;
    R.c = c;
    R.which = which;
    R.slot = slot;
    R.a = make_acc(c, which);
    R.hmask = which < 0 ? (1 << c->nvol) - 1 : 1 << which;
    return R;
}

int ensure_res(spbp_ctx* c, Res& r, size_t list_min) {
    if (!r.ctr) HIPCHK(c, dmalloc(&r.ctr, 4 * sizeof(unsigned long long)));
    if (!r.h) {
        HIPCHK(c, hipHostMalloc((void**)&r.h, 4 * sizeof(unsigned long long), hipHostMallocDefault));
        r.h[0] = r.h[1] = r.h[2] = r.h[3] = 0;
    }
    return grow(c, r.list, r.list_cap, list_min, "threshold list");
}

// Blocks of one hit_kernel launch at most (2^30 work-items).
constexpr long long HIT_SLICE = 1ll << 22;

template <int MODE>
int launch_hits(Round& R, long long nblk, hipStream_t st) {
    spbp_ctx* c = R.c;
    Res& r = R.res();
    for (long long b0 = 0; b0 < nblk; b0 += HIT_SLICE) {
        const unsigned nb = (unsigned)std::min(HIT_SLICE, nblk - b0);
        hit_kernel<MODE><<<nb, RES_TPB, 0, st>>>(b0, r.list, 2 * R.nv, R.nchunks, c->hist_d, (int)c->hist.size(), R.hmask,
                                                 c->nx, c->N, r.cnt, r.off, r.hits);
    }
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

template <typename P>
const P* res_table(const spbp_ctx* c) { return sizeof(P) == 4 ? (const P*)c->tab : (const P*)c->tab64; }
template <typename P>
P res_fltb(const spbp_ctx* c) { return sizeof(P) == 4 ? (P)c->fltb : (P)(4999.0 / 3.0); }

// Phase 1: the voxels in the band; their count to res().h[0].
int phase_list(Round& R, hipStream_t st) {
    spbp_ctx* c = R.c;
    Res& r = R.res();
    fill((float*)r.ctr, 8, 0.0f, st);   // zero bits
    thresh_list_kernel<<<nblocks(c->nh), TPB, 0, st>>>(R.a, c->nx, c->N, 1.0 / (c->nx + 1), 1.0 / c->N, THRESH_BAND,
                                                       r.list, (long long)r.list_cap, r.ctr);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(r.h, r.ctr, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    return SPBP_OK;
}

// Phase 2: the (target, section) hits counted per block; their total to res().h[0].
int phase_hit_counts(Round& R, hipStream_t st) {
    spbp_ctx* c = R.c;
    Res& r = R.res();
    R.nchunks = (int)((c->hist.size() + RES_TPB - 1) / RES_TPB);
    const long long nblk = 2 * R.nv * R.nchunks;
    if (nblk > (long long)INT32_MAX) return fail(c, SPBP_ERR_SIZE, "threshold re-summation too large (%lld blocks)", nblk);
    int rc;
    if ((rc = grow(c, r.cnt, r.cnt_cap, (size_t)nblk, "hit counts")) ||
        (rc = grow(c, r.off, r.off_cap, (size_t)nblk + 1, "hit offsets")) || (rc = launch_hits<0>(R, nblk, st)) ||
        (rc = scan_counts(c, r, r.cnt, nblk, r.off, st)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(r.h, r.off + nblk, sizeof(long long), hipMemcpyDeviceToHost, st));
    return SPBP_OK;
}

// Phase 3: the hits written in section order and their weight terms counted;
// the total to res().h[0].
template <typename P>
int phase_term_counts(Round& R, hipStream_t st) {
    spbp_ctx* c = R.c;
    Res& r = R.res();
    if (R.nhits > (long long)INT32_MAX) return fail(c, SPBP_ERR_SIZE, "threshold re-summation too large (%lld hits)", R.nhits);
    const size_t nh1 = (size_t)std::max<long long>(R.nhits, 1);
    int rc;
    if ((rc = grow(c, r.hits, r.hits_cap, nh1, "hits")) || (rc = grow(c, r.tcnt, r.tcnt_cap, nh1, "term counts")) ||
        (rc = grow(c, r.toff, r.toff_cap, (size_t)R.nhits + 1, "term offsets")) ||
        (rc = launch_hits<1>(R, 2 * R.nv * R.nchunks, st)))
        return rc;
    if (R.nhits > 0)
        term_kernel<P, false><<<(unsigned)((R.nhits + RES_TPB - 1) / RES_TPB), RES_TPB, 0, st>>>(
            r.hits, R.nhits, r.list, c->hist_d, res_table<P>(c), res_fltb<P>(c), c->nx, c->N, r.tcnt, nullptr, nullptr);
    if ((rc = scan_counts(c, r, r.tcnt, R.nhits, r.toff, st))) return rc;
    HIPCHK(c, hipMemcpyAsync(r.h, r.toff + R.nhits, sizeof(long long), hipMemcpyDeviceToHost, st));
    return SPBP_OK;
}

// Phase 4: the terms written and summed sequentially per half; the decisions
// into the mask bits of the round's slot, or with wout the combined weights.
template <typename P>
int phase_decide(Round& R, double* wout, hipStream_t st) {
    spbp_ctx* c = R.c;
    Res& r = R.res();
    const long long nitems = 2 * R.nv;
    int rc;
    if ((rc = grow(c, r.terms, r.terms_cap, (size_t)std::max<long long>(R.nterms, 1) * sizeof(P), "weight terms")) ||
        (rc = grow(c, r.sums, r.sums_cap, (size_t)nitems * 2 * sizeof(P), "weight sums")))
        return rc;
    if (R.nhits > 0)
        term_kernel<P, true><<<(unsigned)((R.nhits + RES_TPB - 1) / RES_TPB), RES_TPB, 0, st>>>(
            r.hits, R.nhits, r.list, c->hist_d, res_table<P>(c), res_fltb<P>(c), c->nx, c->N, nullptr, r.toff,
            (P*)r.terms);
    sum_kernel<P><<<(unsigned)((nitems + RES_TPB - 1) / RES_TPB), RES_TPB, 0, st>>>(nitems, R.nchunks, r.off, r.toff,
                                                                                   (const P*)r.terms, (P*)r.sums);
    const unsigned vb = (unsigned)((R.nv + RES_TPB - 1) / RES_TPB);
    if (wout) {
        weights_out_kernel<P><<<vb, RES_TPB, 0, st>>>(r.list, R.nv, R.hmask, c->nx, c->N, (const P*)r.sums, wout);
    } else {
        c->tmask_dirty = true;
        decide_kernel<P><<<vb, RES_TPB, 0, st>>>(r.list, R.nv, R.a, R.hmask, c->nx, c->N, (const P*)r.sums,
                                                 0.1 * (1.0 / 512.0), c->tmask, 2 * R.slot, r.ctr);
    }
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

bool geom64(const spbp_ctx* c) { return (c->flags & SPBP_GEOM_F64) != 0; }

// Phases 2-4 for rounds whose voxel lists are set (nv), with one host
// synchronisation after each of phases 2 and 3 for all of them; wout (one
// round) asks for the combined weights instead of decisions.
int resolve_listed(const std::vector<Round*>& act, double* wout, hipStream_t st) {
    int rc;
    std::vector<Round*> run;
    for (Round* R : act) if (R->nv > 0) run.push_back(R);
    if (run.empty()) return SPBP_OK;
    for (Round* R : run) if ((rc = phase_hit_counts(*R, st))) return rc;
    HIPCHK(run[0]->c, hipStreamSynchronize(st));
    for (Round* R : run) {
        R->nhits = (long long)R->res().h[0];
        rc = geom64(R->c) ? phase_term_counts<double>(*R, st) : phase_term_counts<float>(*R, st);
        if (rc) return rc;
    }
    HIPCHK(run[0]->c, hipStreamSynchronize(st));
    for (Round* R : run) {
        R->nterms = (long long)R->res().h[0];
        rc = geom64(R->c) ? phase_decide<double>(*R, wout, st) : phase_decide<float>(*R, wout, st);
        if (rc) return rc;
    }
    return SPBP_OK;
}

// NRMW2's test decided by the reference's sums for every round with
// R.resolve, phase by phase across all of them: three host synchronisations
// for any number of volumes (four when a band list outgrows its buffer).
int resolve_rounds(std::vector<Round>& rs, hipStream_t st) {
    std::vector<Round*> act;
    for (Round& R : rs) if (R.resolve) act.push_back(&R);
    if (act.empty()) return SPBP_OK;
    int rc;
    for (Round* R : act) if ((rc = phase_list(*R, st))) return rc;
    HIPCHK(act[0]->c, hipStreamSynchronize(st));
    bool again = false;
    for (Round* R : act) {
        Res& r = R->res();
        R->nv = (long long)r.h[0];
        if ((size_t)R->nv > r.list_cap) {   // a list that outgrew its buffer: list again
            if ((rc = grow(R->c, r.list, r.list_cap, (size_t)R->nv, "threshold list")) || (rc = phase_list(*R, st))) return rc;
            again = true;
        }
    }
    if (again) {
        HIPCHK(act[0]->c, hipStreamSynchronize(st));
        for (Round* R : act) {
            R->nv = (long long)R->res().h[0];
            if ((size_t)R->nv > R->res().list_cap)
                return fail(R->c, SPBP_ERR_DEVICE, "threshold list changed between passes (%lld voxels)", R->nv);
        }
    }
    return resolve_listed(act, nullptr, st);
}

// Blocks of one finalisation launch at most (2^31 work-items).
constexpr unsigned FINAL_LAUNCH_BLOCKS = (1u << 31) / TPB;

// A run of jobs one launch of each window-step kernel covers.
struct FinalRun {
    int j0 = 0, n = 0;
    unsigned hb = 0, vb = 0;   // half-grid and volume blocks
};

// Records the job table's last use at scope exit, so the next call reuses the
// table only after every launch that reads it.
struct TableUse {
    spbp_ctx* c;
    hipStream_t st;
    ~TableUse() { if (c->fjobs_evt) (void)hipEventRecord(c->fjobs_evt, st); }
};

// Normalisation, inverse transform and window steps of every round, in waves
// (a context's buffers serve one volume per wave): per wave one normalise
// launch over all its volumes, their inverse transforms, one window, one
// shell-mean and one subtract launch over all of them, then SYMVOL where
// symmetries were inserted.  The job table lives on the first round's context.
int run_waves(std::vector<Round>& rs, hipStream_t st) {
    spbp_ctx* t = rs[0].c;
    int nw = 0;
    for (const Round& R : rs) nw = std::max(nw, R.wave + 1);
    std::vector<FinalJob> fj;
    std::vector<std::vector<FinalRun>> runs(nw);
    std::vector<std::vector<int>> members(nw);
    for (int w = 0; w < nw; ++w) {
        for (int i = 0; i < (int)rs.size(); ++i) {
            const Round& R = rs[i];
            if (R.wave != w) continue;
            const spbp_ctx* c = R.c;
            const size_t nv = (size_t)c->nx * c->nx * c->nx;
            const unsigned hb = (unsigned)((c->nh + (size_t)TPB * NORM_VPT - 1) / ((size_t)TPB * NORM_VPT)), vb = nblocks(nv);
            auto& rw = runs[w];
            if (rw.empty() || (unsigned long long)rw.back().hb + hb > FINAL_LAUNCH_BLOCKS ||
                (unsigned long long)rw.back().vb + vb > FINAL_LAUNCH_BLOCKS) {
                FinalRun fr;
                fr.j0 = (int)fj.size();
                rw.push_back(fr);
            }
            FinalRun& fr = rw.back();
            FinalJob f;
            f.a = R.a;
            f.work = c->work;
            f.real = c->real;
            f.out = R.out;
            f.wkb = c->wkb;
            f.psum = c->psum;
            f.pcnt = c->pcnt;
            f.mean = c->mean;
            f.nx = c->nx;
            f.N = c->N;
            f.pix = 1.0f / (float)((double)c->N * c->N * c->N);   // PIX = 1.0/FLOAT(N^3) (fftw3.f:444)
            f.invH = 1.0 / (c->nx + 1);
            f.invN = 1.0 / c->N;
            f.invL = 1.0 / c->nx;
            f.hb0 = fr.hb;
            f.vb0 = fr.vb;
            f.vnb = vb;
            fr.hb += hb;
            fr.vb += vb;
            ++fr.n;
            fj.push_back(f);
            members[w].push_back(i);
        }
    }
    const int n = (int)fj.size();
    // the event exists before any table does: it marks the table's last use
    if (!t->fjobs_evt) HIPCHK(t, hipEventCreateWithFlags(&t->fjobs_evt, hipEventDisableTiming));
    HIPCHK(t, hipEventSynchronize(t->fjobs_evt));   // the last call's launches have read the table
    if (n > t->fjobs_cap) {
        if (t->fjobs_h) { (void)hipHostFree(t->fjobs_h); t->fjobs_h = nullptr; }
        if (t->fjobs_d) { (void)hipFree(t->fjobs_d); t->fjobs_d = nullptr; }
        t->fjobs_cap = 0;
        const int cap = std::max(n, 16);
        HIPCHK(t, hipHostMalloc((void**)&t->fjobs_h, cap * sizeof(FinalJob), hipHostMallocDefault));
        HIPCHK(t, dmalloc(&t->fjobs_d, cap * sizeof(FinalJob)));
        t->fjobs_cap = cap;
    }
    std::memcpy(t->fjobs_h, fj.data(), n * sizeof(FinalJob));
    TableUse use{t, st};
    HIPCHK(t, hipMemcpyAsync(t->fjobs_d, t->fjobs_h, n * sizeof(FinalJob), hipMemcpyHostToDevice, st));
    for (int w = 0; w < nw; ++w) {
        for (const FinalRun& fr : runs[w]) normalize_kernel<<<fr.hb, TPB, 0, st>>>(t->fjobs_d + fr.j0, fr.n);
        HIPCHK(t, hipGetLastError());
        for (int i : members[w]) {
            spbp_ctx* c = rs[i].c;
            FFTCHK(c, hipfftSetStream(c->inv, st));
            FFTCHK(c, hipfftExecC2R(c->inv, (hipfftComplex*)c->work, (hipfftReal*)c->real));
        }
        for (const FinalRun& fr : runs[w]) {
            window_kernel<<<fr.vb, TPB, 0, st>>>(t->fjobs_d + fr.j0, fr.n);
            shell_mean_kernel<<<fr.n, TPB, 0, st>>>(t->fjobs_d + fr.j0);
            subtract_kernel<<<fr.vb, TPB, 0, st>>>(t->fjobs_d + fr.j0, fr.n);
        }
        HIPCHK(t, hipGetLastError());
        for (int i : members[w]) {   // real-space symmetrisation after WINDKB2A (bp32f.f:278-303)
            spbp_ctx* c = rs[i].c;
            if (c->nsym <= 1) continue;
            const size_t nv = (size_t)c->nx * c->nx * c->nx;
            if (!c->symtmp) HIPCHK(c, dmalloc(&c->symtmp, nv * sizeof(float)));
            HIPCHK(c, hipMemcpyAsync(c->symtmp, rs[i].out, nv * sizeof(float), hipMemcpyDeviceToDevice, st));
            symvol_kernel<<<nblocks(nv), TPB, 0, st>>>(c->symtmp, rs[i].out, c->nx, c->sm_d, c->nsym);
            HIPCHK(c, hipGetLastError());
        }
    }
    return SPBP_OK;
}

// The requested volumes of njobs distinct contexts (validated by the caller).
int finalize_jobs(const spbp_final_job* jobs, int njobs, hipStream_t st) {
    int rc;
    std::vector<Round> rs;
    for (int j = 0; j < njobs; ++j) {
        spbp_ctx* c = jobs[j].ctx;
        if ((rc = ensure_final(c))) return rc;
        // the last call's counters have reached the pinned words they share
        if (c->stats_evt) HIPCHK(c, hipEventSynchronize(c->stats_evt));
        c->res_stats[0] = c->res_stats[1] = c->res_stats[2] = 0;
        c->stats_slot = -1;
        const bool resolve = c->hist_ok && !(c->flags & SPBP_PLAIN_THRESHOLD) && !c->hist.empty() &&
                             c->hist.size() <= (size_t)INT32_MAX;
        float* outs[3] = {jobs[j].out_full, jobs[j].out_h1, jobs[j].out_h2};
        int wave = 0;
        for (int q = 0; q < 3; ++q) {
            if (!outs[q]) continue;
            Round R = make_round(c, q - 1, q);
            R.out = outs[q];
            R.resolve = resolve;
            R.wave = wave++;
            if (resolve && (rc = ensure_res(c, R.res(), 4096))) return rc;
            rs.push_back(R);
        }
        if (wave == 0 || !resolve) continue;
        if ((rc = upload_history(c, st))) return rc;
        if (!c->tmask) {
            HIPCHK(c, dmalloc(&c->tmask, (c->nh + 3) / 4 * 4));
            c->tmask_dirty = true;
        }
        if (c->tmask_dirty) {
            fill((float*)c->tmask, (c->nh + 3) / 4, 0.0f, st);   // zero bits
            HIPCHK(c, hipGetLastError());
            c->tmask_dirty = false;
        }
    }
    if (rs.empty()) return SPBP_OK;
    // every timing context records the batch's spans
    std::vector<std::unique_ptr<Timed>> span;
    for (int j = 0; j < njobs; ++j) span.emplace_back(new Timed(jobs[j].ctx, SPBP_K_RESOLVE, 1, st));
    if ((rc = resolve_rounds(rs, st))) return rc;
    span.clear();
    // the last volume of each context leaves its counters for spbp_threshold_stats
    for (int i = 0; i < (int)rs.size(); ++i) {
        Round& R = rs[i];
        spbp_ctx* c = R.c;
        if (R.resolve && R.nv > 0) {
            R.a.mask = c->tmask;
            R.a.mshift = 2 * R.slot;
        }
        if (i + 1 < (int)rs.size() && rs[i + 1].c == c) continue;
        c->res_stats[0] = R.nv;
        if (R.resolve && R.nv > 0) {
            if (!c->stats_evt) HIPCHK(c, hipEventCreateWithFlags(&c->stats_evt, hipEventDisableTiming));
            HIPCHK(c, hipMemcpyAsync(R.res().h, R.res().ctr, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
            HIPCHK(c, hipEventRecord(c->stats_evt, st));
            c->stats_slot = R.slot;
        }
    }
    for (int j = 0; j < njobs; ++j) span.emplace_back(new Timed(jobs[j].ctx, SPBP_K_FINALIZE, 1, st));
    if ((rc = run_waves(rs, st))) return rc;
    span.clear();
    // clear the decisions for the next finalisation
    bool cleared = true;
    for (const Round& R : rs) {
        if (!(R.resolve && R.nv > 0)) continue;
        mask_clear_kernel<<<(unsigned)((R.nv + RES_TPB - 1) / RES_TPB), RES_TPB, 0, st>>>(R.res().list, R.nv, R.c->tmask);
        cleared = cleared && hipGetLastError() == hipSuccess;
    }
    for (const Round& R : rs)
        if (R.resolve && R.nv > 0) R.c->tmask_dirty = !cleared;
    return SPBP_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* spbp_version(void) { return "spbp 0.1 (gfx950)"; }

const char* spbp_last_error(const spbp_ctx* ctx) {
    return ctx ? ctx->err : g_last_error;
}

size_t spbp_grid_elems(const spbp_ctx* ctx) { return ctx ? ctx->nh : 0; }

int spbp_create(const spbp_opts* o, spbp_ctx** out) {
    if (!o || !out) return fail(nullptr, SPBP_ERR_ARG, "NULL options or output pointer");
    *out = nullptr;
    if (o->nx < 4 || o->nx > 1024) return fail(nullptr, SPBP_ERR_SIZE, "image size %d outside 4..1024", o->nx);
    spbp_ctx* c = new spbp_ctx;
    c->nx = o->nx;
    c->N = 2 * o->nx;
    c->H = o->nx + 1;
    c->flags = o->flags;
    c->nh = (size_t)c->H * c->N * c->N;
    c->nvol = (o->flags & SPBP_TWO_HALVES) ? 2 : 1;
    int dev = o->device;
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) { delete c; return fail(nullptr, SPBP_ERR_DEVICE, "no HIP device"); }
    c->device = dev;
    DeviceGuard dg(dev);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || dev >= ndev) { delete c; return fail(nullptr, SPBP_ERR_DEVICE, "HIP device %d not available", dev); }
    if (o->batch > 0) {
        c->batch = o->batch;
    } else {
        // 8 GiB of chunk buffers (of 288 GB): the spectra, plus the padded
        // images when hipFFT does the transforms.  Larger chunks mean fewer
        // list builds and fewer persistent-kernel tails: 20,000 projections
        // of 128^2 go through one launch (481.8k proj/s for a BP 32F job,
        // 473.5k in three launches of 2 GiB chunks).
        const bool own = (c->N & (c->N - 1)) == 0 && c->N >= 8 && c->N <= std::min(2048, FFT_ROW_ELEMS) &&
                         !(o->flags & SPBP_HIPFFT);
        const size_t per = (own ? 0 : (size_t)c->N * c->N * 4) + (size_t)c->N * c->H * 8;
        c->batch = (int)std::max<size_t>(16, std::min<size_t>(32768, ((size_t)8192 << 20) / per));
    }
    c->strategy = o->strategy == SPBP_STRAT_DIRECT || o->strategy == SPBP_STRAT_GATHER ? o->strategy : SPBP_STRAT_BRICK;
    if (o->strategy < 0 || o->strategy > SPBP_STRAT_GATHER) { delete c; return fail(nullptr, SPBP_ERR_ARG, "unknown strategy %d", o->strategy); }
    if (o->brick_proj > 128) { delete c; return fail(nullptr, SPBP_ERR_ARG, "brick_proj %d above 128", o->brick_proj); }
    c->C = o->brick_proj > 0 ? o->brick_proj : 128;
    {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0) c->ncu = n;
    }
    int rc = SPBP_OK;
    auto bail = [&](int code) { spbp_destroy(c); return code; };
    for (int h = 0; h < c->nvol; ++h) {
        if (dmalloc(&c->X[h], c->nh * sizeof(float2)) != hipSuccess ||
            dmalloc(&c->W[h], c->nh * sizeof(float)) != hipSuccess ||
            ((c->flags & SPBP_TWOSUM_W) && dmalloc(&c->Wc[h], c->nh * sizeof(float)) != hipSuccess))
            return bail(fail(nullptr, SPBP_ERR_ALLOC, "BP 32F; XE, WE, WO, & XO: %zu bytes", c->nh * 12 * (size_t)c->nvol));
    }
    std::vector<float> tab;
    if (c->flags & SPBP_GEOM_F64) {
        kb_table_f64_f32(c->N, tab, c->fltb);
        std::vector<double> t64;
        kb_table_f64(c->N, t64);
        if (dmalloc(&c->tab64, t64.size() * sizeof(double)) != hipSuccess ||
            hipMemcpy(c->tab64, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(nullptr, SPBP_ERR_ALLOC, "FP64 table"));
    } else {
        kb_table(c->N, tab, c->fltb);
    }
    if (dmalloc(&c->tab, tab.size() * sizeof(float)) != hipSuccess ||
        hipMemcpy(c->tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        dmalloc(&c->flag, sizeof(int)) != hipSuccess ||
        hipEventCreateWithFlags(&c->geom_evt, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->jobs_evt, hipEventDisableTiming) != hipSuccess)
        return bail(fail(nullptr, SPBP_ERR_DEVICE, "context set-up failed"));
    if (dmalloc(&c->bnext, sizeof(int)) != hipSuccess) return bail(fail(nullptr, SPBP_ERR_ALLOC, "brick queue"));
    if ((c->N & (c->N - 1)) == 0 && c->N >= 8 && c->N <= std::min(2048, FFT_ROW_ELEMS) && !(o->flags & SPBP_HIPFFT)) {
        int L = 0;
        while ((1 << L) < c->N) ++L;
        std::vector<float2> t(c->N);
        for (int m = 0; m < c->N; ++m) {
            const double a = -2.0 * M_PI * (double)m / (double)c->N;
            t[m] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        if (dmalloc(&c->tw, t.size() * sizeof(float2)) != hipSuccess ||
            hipMemcpy(c->tw, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(nullptr, SPBP_ERR_ALLOC, "twiddle table"));
        c->fft_log2 = L;
    }
    // default plan: owner-computes bricks, 16x16x13 LDS regions around 12x12x9 cores (the largest region
    // that runs two 512-thread groups per CU), 128 projections per item: fastest measured at 128^2 into
    // 256^3 and 256^2 into 512^3; spbp_tune picks per workload
    if ((rc = configure_bricks(c, NSHAPES - 1))) { std::memcpy(g_last_error, c->err, ERR_LEN); return bail(rc); }
    rc = spbp_reset(c, nullptr);
    if (rc) return bail(rc);
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(nullptr, SPBP_ERR_DEVICE, "zeroing accumulators failed"));
    *out = c;
    return SPBP_OK;
}

int spbp_destroy(spbp_ctx* c) {
    if (!c) return SPBP_OK;
    if (c->graphs > 0) return refuse_captured(c, "spbp_destroy");
    DeviceGuard dg(c->device);
    if (c->geom_evt) (void)hipEventSynchronize(c->geom_evt);
    if (c->jobs_evt) (void)hipEventSynchronize(c->jobs_evt);
    for (int h = 0; h < 2; ++h) {
        if (c->X[h]) (void)hipFree(c->X[h]);
        if (c->W[h]) (void)hipFree(c->W[h]);
        if (c->Wc[h]) (void)hipFree(c->Wc[h]);
    }
    void* dev_ptrs[] = {c->tab, c->padbuf, c->spec, c->vmax, c->geom_d, c->work, c->real, c->wkb, c->psum, c->pcnt, c->mean, c->flag,
                        c->bcnt, c->boff, c->bioff, c->bcur, c->bscan, c->bnext, c->blist, c->bikey, c->bpln, c->jobs_d,
                        c->sm_d, c->symtmp, c->rotbuf, c->rprm_d, c->hist_d, c->tab64, c->tmask, c->tw, c->fmean, c->fjobs_d};
    for (void* p : dev_ptrs) if (p) (void)hipFree(p);
    if (c->stats_evt) { (void)hipEventSynchronize(c->stats_evt); (void)hipEventDestroy(c->stats_evt); }
    if (c->fjobs_evt) { (void)hipEventSynchronize(c->fjobs_evt); (void)hipEventDestroy(c->fjobs_evt); }
    if (c->fjobs_h) (void)hipHostFree(c->fjobs_h);
    for (Res& r : c->res) {
        void* rp[] = {r.list, r.cnt, r.off, r.bsum, r.hits, r.tcnt, r.toff, r.terms, r.sums, r.ctr};
        for (void* p : rp) if (p) (void)hipFree(p);
        if (r.h) (void)hipHostFree(r.h);
    }
    if (c->geom_h) (void)hipHostFree(c->geom_h);
    if (c->jobs_h) (void)hipHostFree(c->jobs_h);
    if (c->rprm_evt) (void)hipEventSynchronize(c->rprm_evt);
    if (c->rprm_h) (void)hipHostFree(c->rprm_h);
    if (c->fprm_evt) { (void)hipEventSynchronize(c->fprm_evt); (void)hipEventDestroy(c->fprm_evt); }
    if (c->fprm_h) (void)hipHostFree(c->fprm_h);
    if (c->fprm_d) (void)hipFree(c->fprm_d);
    rtsf_release(c->rtsf);
    if (c->rprm_evt) (void)hipEventDestroy(c->rprm_evt);
    for (void* h : c->hstage)
        if (h) (void)hipHostFree(h);
    if (c->fwd) hipfftDestroy(c->fwd);
    if (c->fwd_tail) hipfftDestroy(c->fwd_tail);
    if (c->inv) hipfftDestroy(c->inv);
    if (c->c2r) hipfftDestroy(c->c2r);
    if (c->c2r2) hipfftDestroy(c->c2r2);
    if (c->c2rbuf) (void)hipFree(c->c2rbuf);
    if (c->geom_evt) (void)hipEventDestroy(c->geom_evt);
    if (c->jobs_evt) (void)hipEventDestroy(c->jobs_evt);
    spbp_reset_stats(c);
    for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
    delete c;
    return SPBP_OK;
}

int spbp_reset(spbp_ctx* c, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    DeviceGuard dg(c->device);
    c->hist.clear();
    c->hist_ok = true;
    c->filled = false;
    hipStream_t st = (hipStream_t)stream;
    for (int h = 0; h < c->nvol; ++h) {
        fill((float*)c->X[h], 2 * c->nh, 0.0f, st);
        fill(c->W[h], c->nh, 0.0f, st);
        if (c->Wc[h]) fill(c->Wc[h], c->nh, 0.0f, st);
    }
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

int spbp_insert(spbp_ctx* c, const void* imgs, int dtype, int nimg, const double* angles_deg,
                const int8_t* half, void* stream) {
    return insert_impl(c, imgs, dtype, nimg, nullptr, angles_deg, half, (hipStream_t)stream);
}

int spbp_insert_dm(spbp_ctx* c, const void* imgs, int dtype, int nimg, const float* dm,
                   const int8_t* half, void* stream) {
    if (!dm) return fail(c, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE (no rotations given)");
    return insert_impl(c, imgs, dtype, nimg, dm, nullptr, half, (hipStream_t)stream);
}

int spbp_gather(spbp_ctx* c, const float* X, int nimg, const double* angles_deg, const float* dm,
                float* out_bi, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (nimg < 0) return fail(c, SPBP_ERR_ARG, "negative projection count %d", nimg);
    if (nimg == 0) return SPBP_OK;
    if (!X || !out_bi) return fail(c, SPBP_ERR_ARG, "NULL grid or output pointer");
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    int rc = stage_geom(c, nimg, dm, angles_deg, nullptr, st, false);
    if (rc) return rc;
    const int N = c->N, H = c->H;
    const unsigned gx = (unsigned)(((H + GT - 1) / GT) * ((N + GT - 1) / GT));
    for (int p0 = 0; p0 < nimg; p0 += 65535) {
        const int b = std::min(65535, nimg - p0);
        Timed tm(c, SPBP_K_GATHER, b, st);
        if (c->flags & SPBP_GEOM_F64)
            gather_kernel<double><<<dim3(gx, b), GT * GT, 0, st>>>((const float2*)X, c->geom_d + p0, c->nx, N, c->tab,
                                                                   c->fltb, (float2*)out_bi + (size_t)p0 * N * H);
        else
            gather_kernel<float><<<dim3(gx, b), GT * GT, 0, st>>>((const float2*)X, c->geom_d + p0, c->nx, N, c->tab,
                                                                  c->fltb, (float2*)out_bi + (size_t)p0 * N * H);
    }
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

int spbp_scatter(spbp_ctx* c, const float* bi, int nimg, const double* angles_deg, const float* dm,
                 float* X, float* W, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (nimg < 0) return fail(c, SPBP_ERR_ARG, "negative projection count %d", nimg);
    if (nimg == 0) return SPBP_OK;
    if (!bi || !X) return fail(c, SPBP_ERR_ARG, "NULL spectrum or grid pointer");
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    int rc = stage_geom(c, nimg, dm, angles_deg, nullptr, st, false);
    if (rc) return rc;
    rc = ensure_chunk(c, std::min(c->batch, nimg));
    if (rc) return rc;
    const int N = c->N, H = c->H;
    const Targets t{(float2*)X, W, nullptr, nullptr, nullptr, nullptr};
    for (int p0 = 0; p0 < nimg; p0 += c->fwd_batch) {
        const int b = std::min(c->fwd_batch, nimg - p0);
        const size_t n = (size_t)b * N * H;
        bi_to_raw_kernel<<<nblocks(n), TPB, 0, st>>>((const float2*)bi + (size_t)p0 * N * H, N, H, n, c->spec);
        spec_absmax_kernel<<<b, TPB, 0, st>>>(c->spec, N * H, c->vmax);
        HIPCHK(c, hipGetLastError());
        if ((rc = run_insert(c, p0, c->geom_d + p0, b, t, st, b))) return rc;
    }
    return SPBP_OK;
}

int spbp_gather_images(spbp_ctx* c, const float* X, int nimg, const double* angles_deg, const float* dm,
                       float* imgs, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (nimg < 0) return fail(c, SPBP_ERR_ARG, "negative projection count %d", nimg);
    if (nimg == 0) return SPBP_OK;
    if (!X || !imgs) return fail(c, SPBP_ERR_ARG, "NULL grid or image pointer");
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    int rc = stage_geom(c, nimg, dm, angles_deg, nullptr, st, false);
    if (rc) return rc;
    if ((rc = ensure_chunk(c, std::min(c->batch, nimg)))) return rc;
    const int N = c->N, H = c->H, B = c->fwd_batch;
    if (c->c2r_cap < B) {
        c->c2r_cap = 0;
        if (c->c2rbuf) { (void)hipFree(c->c2rbuf); c->c2rbuf = nullptr; }
        HIPCHK(c, dmalloc(&c->c2rbuf, (size_t)B * N * N * sizeof(float)));
        c->c2r_cap = B;
    }
    const int ip = (N - c->nx) / 2 + c->nx % 2;
    const unsigned gx = (unsigned)(((H + GT - 1) / GT) * ((N + GT - 1) / GT));
    for (int p0 = 0; p0 < nimg; p0 += B) {
        const int b = std::min(B, nimg - p0);
        // a plan per batch size, two kept: the most recent becomes c2r
        if (c->c2r2 && c->c2r2_batch == b) {
            std::swap(c->c2r, c->c2r2);
            std::swap(c->c2r_batch, c->c2r2_batch);
        } else if (!c->c2r || c->c2r_batch != b) {
            if (c->c2r2) hipfftDestroy(c->c2r2);
            c->c2r2 = c->c2r;
            c->c2r2_batch = c->c2r_batch;
            c->c2r = 0;
            int n2[2] = {N, N}, inembed[2] = {N, H}, onembed[2] = {N, N};
            const hipfftResult r = hipfftPlanMany(&c->c2r, 2, n2, inembed, 1, N * H, onembed, 1, N * N, HIPFFT_C2R, b);
            if (r != HIPFFT_SUCCESS) { c->c2r = 0; return fail(c, SPBP_ERR_FFT, "hipfftPlanMany(C2R %d): %d", b, (int)r); }
            c->c2r_batch = b;
        }
        {
            Timed tm(c, SPBP_K_GATHER, b, st);
            for (int q0 = 0; q0 < b; q0 += 65535) {
                const int bq = std::min(65535, b - q0);
                if (c->flags & SPBP_GEOM_F64)
                    gather_kernel<double><<<dim3(gx, bq), GT * GT, 0, st>>>((const float2*)X, c->geom_d + p0 + q0, c->nx,
                                                                            N, c->tab, c->fltb, c->spec + (size_t)q0 * N * H);
                else
                    gather_kernel<float><<<dim3(gx, bq), GT * GT, 0, st>>>((const float2*)X, c->geom_d + p0 + q0, c->nx,
                                                                           N, c->tab, c->fltb, c->spec + (size_t)q0 * N * H);
            }
        }
        const size_t n = (size_t)b * N * H;
        bi_complete_kernel<<<nblocks(n), TPB, 0, st>>>(c->spec, N, H, c->nx, n);
        HIPCHK(c, hipGetLastError());
        FFTCHK(c, hipfftSetStream(c->c2r, st));
        FFTCHK(c, hipfftExecC2R(c->c2r, (hipfftComplex*)c->spec, (hipfftReal*)c->c2rbuf));
        const size_t m = (size_t)b * c->nx * c->nx;
        crop_images_kernel<<<nblocks(m), TPB, 0, st>>>(c->c2rbuf, N, c->nx, ip, m, imgs + (size_t)p0 * c->nx * c->nx);
        HIPCHK(c, hipGetLastError());
    }
    return SPBP_OK;
}

int spbp_scatter_images(spbp_ctx* c, const float* imgs, int nimg, const double* angles_deg, const float* dm,
                        float* X, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (nimg < 0) return fail(c, SPBP_ERR_ARG, "negative projection count %d", nimg);
    if (nimg == 0) return SPBP_OK;
    if (!imgs || !X) return fail(c, SPBP_ERR_ARG, "NULL image or grid pointer");
    if (!c->fft_log2) return fail(c, SPBP_ERR_SIZE, "the image adjoint needs a power-of-two padded size (N = %d)", c->N);
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    int rc = stage_geom(c, nimg, dm, angles_deg, nullptr, st, false);
    if (rc) return rc;
    if ((rc = ensure_chunk(c, std::min(c->batch, nimg)))) return rc;
    const int N = c->N, H = c->H;
    const size_t npix = (size_t)c->nx * c->nx;
    const Targets t{(float2*)X, nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int p0 = 0; p0 < nimg; p0 += c->fwd_batch) {
        const int b = std::min(c->fwd_batch, nimg - p0);
        const size_t n = (size_t)b * N * H;
        // no PADD2 mean: zero means (all-zero bits are 0.0) and zero maxima
        fill((float*)c->fmean, 2 * (size_t)b, 0.0f, st);
        fill(c->vmax, b, 0.0f, st);
        HIPCHK(c, hipGetLastError());
        if ((rc = own_fft(c, imgs + (size_t)p0 * npix, SPBP_F32, b, st))) return rc;
        adjoint_spectrum_kernel<<<nblocks(n), TPB, 0, st>>>(c->spec, N, H, c->nx, n);
        spec_absmax_kernel<<<b, TPB, 0, st>>>(c->spec, N * H, c->vmax);
        HIPCHK(c, hipGetLastError());
        if ((rc = run_insert(c, p0, c->geom_d + p0, b, t, st, b))) return rc;
    }
    return SPBP_OK;
}

int spbp_tune(spbp_ctx* c, const void* imgs, int dtype, int nimg, const double* angles_deg, const float* dm,
              void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (nimg <= 0 || !imgs) return fail(c, SPBP_ERR_ARG, "tuning needs a sample of images");
    if (dtype != SPBP_F32 && dtype != SPBP_BF16) return fail(c, SPBP_ERR_ARG, "unknown image type %d", dtype);
    if (c->graphs > 0) return refuse_captured(c, "spbp_tune");
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    const int m = std::min(nimg, std::min(c->batch, 1024));
    int rc = stage_geom(c, m, dm, angles_deg, nullptr, st);
    if (rc) return rc;
    if ((rc = ensure_chunk(c, std::min(c->batch, m)))) return rc;
    const int b = std::min(c->fwd_batch, m);
    if ((rc = pad_fft(c, imgs, dtype, b, st))) return rc;
    float2* Xs = nullptr;
    float* Ws = nullptr;
    HIPCHK(c, dmalloc(&Xs, c->nh * sizeof(float2)));
    if (dmalloc(&Ws, c->nh * sizeof(float)) != hipSuccess) { (void)hipFree(Xs); return fail(c, SPBP_ERR_ALLOC, "tuning scratch"); }
    const Targets t{Xs, Ws, nullptr, nullptr, nullptr, nullptr};
    const int s_strat = c->strategy, s_shape = c->shape, s_C = c->C;
    float ms = 0.f;
    c->strategy = SPBP_STRAT_DIRECT;
    rc = time_plan(c, c->geom_d, b, t, st, 1, &ms);
    c->tune_direct_ms = ms / b;
    if (rc == SPBP_OK) {
        c->strategy = SPBP_STRAT_GATHER;
        rc = time_plan(c, c->geom_d, b, t, st, 1, &ms);
        c->tune_gather_ms = ms / b;
    }
    c->strategy = SPBP_STRAT_BRICK;
    for (int sh = 0; sh < NSHAPES && rc == SPBP_OK; ++sh) {
        if ((rc = configure_bricks(c, sh)) || (rc = ensure_lists(c, c->fwd_batch))) break;
        for (int k = 0; k < NCVALS && rc == SPBP_OK; ++k) {
            c->C = CVALS[k];
            rc = time_plan(c, c->geom_d, b, t, st, 2, &ms);
            c->tune_ms[sh][k] = ms / b;
        }
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(Xs);
    (void)hipFree(Ws);
    if (rc) {
        c->strategy = s_strat;
        c->C = s_C;
        (void)configure_bricks(c, s_shape);
        (void)ensure_lists(c, c->fwd_batch);
        return rc;
    }
    int bs = 0, bk = 0;
    for (int sh = 0; sh < NSHAPES; ++sh)
        for (int k = 0; k < NCVALS; ++k)
            if (c->tune_ms[sh][k] < c->tune_ms[bs][bk]) { bs = sh; bk = k; }
    c->tuned = true;
    const float best_brick = c->tune_ms[bs][bk];
    c->strategy = SPBP_STRAT_BRICK;
    if (c->tune_direct_ms < best_brick && c->tune_direct_ms <= c->tune_gather_ms) c->strategy = SPBP_STRAT_DIRECT;
    else if (c->tune_gather_ms < best_brick && c->tune_gather_ms < c->tune_direct_ms) c->strategy = SPBP_STRAT_GATHER;
    c->C = CVALS[bk];
    if ((rc = configure_bricks(c, bs))) return rc;
    return ensure_lists(c, c->fwd_batch);
}

int spbp_get_plan(const spbp_ctx* c, int* strategy, int* shape, int* brick_proj, float* ms_per_proj) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (strategy) *strategy = c->strategy;
    if (shape) *shape = c->shape;
    if (brick_proj) *brick_proj = c->C;
    if (ms_per_proj) {
        // row-major [NSHAPES][NCVALS] brick times, then the direct time; 0 until tuned
        for (int sh = 0; sh < NSHAPES; ++sh)
            for (int k = 0; k < NCVALS; ++k) ms_per_proj[sh * NCVALS + k] = c->tuned ? c->tune_ms[sh][k] : 0.f;
        ms_per_proj[NSHAPES * NCVALS] = c->tuned ? c->tune_direct_ms : 0.f;
        ms_per_proj[NSHAPES * NCVALS + 1] = c->tuned ? c->tune_gather_ms : 0.f;
    }
    return SPBP_OK;
}

int spbp_set_plan(spbp_ctx* c, int strategy, int shape, int brick_proj) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (strategy < SPBP_STRAT_AUTO || strategy > SPBP_STRAT_GATHER) return fail(c, SPBP_ERR_ARG, "unknown strategy %d", strategy);
    if (shape < 0 || shape >= NSHAPES) return fail(c, SPBP_ERR_ARG, "brick shape %d outside 0..%d", shape, NSHAPES - 1);
    if (brick_proj < 0 || brick_proj > 128) return fail(c, SPBP_ERR_ARG, "projections per work item %d outside 0..128", brick_proj);
    DeviceGuard dg(c->device);
    c->strategy = strategy == SPBP_STRAT_DIRECT || strategy == SPBP_STRAT_GATHER ? strategy : SPBP_STRAT_BRICK;
    if (brick_proj > 0) c->C = brick_proj;
    int rc = configure_bricks(c, shape);
    if (rc) return rc;
    return c->fwd_batch > 0 ? ensure_lists(c, c->fwd_batch) : SPBP_OK;
}

}  // extern "C"

namespace {

// Checks a batch and gives every context the shared brick shape: the one
// with the least total tuned time (each job at its own best C), or the first
// context's shape when any is untuned.
int prepare_batch(const spbp_job* jobs, int njobs) {
    if (!jobs || njobs < 1 || njobs > MAX_JOBS) return fail(nullptr, SPBP_ERR_ARG, "batch of %d jobs (1..%d)", njobs, MAX_JOBS);
    for (int j = 0; j < njobs; ++j) {
        spbp_ctx* c = jobs[j].ctx;
        if (!c) return fail(nullptr, SPBP_ERR_ARG, "job %d has no context", j);
        if (c->device != jobs[0].ctx->device) return fail(c, SPBP_ERR_ARG, "job %d is on another device", j);
        for (int k = 0; k < j; ++k)
            if (jobs[k].ctx == c) return fail(c, SPBP_ERR_ARG, "jobs %d and %d share a context", k, j);
        if (jobs[j].nimg < 0) return fail(c, SPBP_ERR_ARG, "job %d: negative image count", j);
        if (jobs[j].nimg > 0 && !jobs[j].imgs) return fail(c, SPBP_ERR_ARG, "job %d: NULL image pointer", j);
        if (jobs[j].dtype != SPBP_F32 && jobs[j].dtype != SPBP_BF16) return fail(c, SPBP_ERR_ARG, "job %d: unknown image type", j);
        if ((c->flags ^ jobs[0].ctx->flags) & SPBP_GEOM_F64)
            return fail(c, SPBP_ERR_ARG, "job %d: SPBP_GEOM_F64 differs from job 0 (one kernel serves the batch)", j);
    }
    bool all_tuned = true;
    for (int j = 0; j < njobs; ++j) all_tuned = all_tuned && jobs[j].ctx->tuned;
    int shape = jobs[0].ctx->shape;
    if (all_tuned) {
        double best = 1e300;
        for (int sh = 0; sh < NSHAPES; ++sh) {
            double tot = 0.0;
            for (int j = 0; j < njobs; ++j) {
                const spbp_ctx* c = jobs[j].ctx;
                float m = c->tune_ms[sh][0];
                for (int k = 1; k < NCVALS; ++k) m = std::min(m, c->tune_ms[sh][k]);
                if (c->strategy == SPBP_STRAT_DIRECT) m = c->tune_direct_ms;
                if (c->strategy == SPBP_STRAT_GATHER) m = c->tune_gather_ms;
                tot += (double)m * jobs[j].nimg;
            }
            if (tot < best) { best = tot; shape = sh; }
        }
    }
    for (int j = 0; j < njobs; ++j) {
        int rc = apply_shape(jobs[j].ctx, shape);
        if (rc) return rc;
    }
    return SPBP_OK;
}

// Host-side job table of round r: the brick jobs of the contexts that still
// have a chunk in that round (direct- and gather-strategy contexts run their own kernels).
// Depends only on device addresses, so it can be built before any work runs.
int round_table(const spbp_job* jobs, int njobs, Geom* const* geom, int r, BrickJob* table, spbp_ctx** lead) {
    int n = 0;
    *lead = nullptr;
    for (int j = 0; j < njobs; ++j) {
        spbp_ctx* c = jobs[j].ctx;
        const int p0 = r * c->fwd_batch;
        if (p0 >= jobs[j].nimg || own_kernel(c)) continue;
        const Targets t{c->X[0], c->W[0], c->X[1], c->W[1], c->Wc[0], c->Wc[1]};
        const int b = std::min(c->fwd_batch, jobs[j].nimg - p0);
        table[n++] = brick_job(c, p0, geom[j] + (size_t)p0 * c->nsym, b * c->nsym, t);
        if (!*lead) *lead = c;   // its list build zeroes its queue head
    }
    return n;
}

// Enqueue the batch on st: per round, every job with projections left pads and
// transforms its next chunk and builds its lists (or runs the direct kernel);
// then one brick-kernel launch serves all jobs of the round.  geom[j] is job
// j's device geometry.  With tables (graph capture), tables[r] must already
// hold round r's job table in device memory: nothing is copied from the host.
int enqueue_batch(const spbp_job* jobs, int njobs, Geom* const* geom, BrickJob* const* tables, hipStream_t st) {
    int rounds = 0;
    for (int j = 0; j < njobs; ++j) {
        spbp_ctx* c = jobs[j].ctx;
        if (jobs[j].nimg > 0) rounds = std::max(rounds, (jobs[j].nimg + c->fwd_batch - 1) / c->fwd_batch);
    }
    for (int r = 0; r < rounds; ++r) {
        long units = 0;
        for (int j = 0; j < njobs; ++j) {
            spbp_ctx* c = jobs[j].ctx;
            const int p0 = r * c->fwd_batch;
            if (p0 >= jobs[j].nimg) continue;
            const int b = std::min(c->fwd_batch, jobs[j].nimg - p0);
            const size_t esz = jobs[j].dtype == SPBP_BF16 ? 2 : 4;
            int rc = pad_fft(c, (const char*)jobs[j].imgs + (size_t)p0 * c->nx * c->nx * esz, jobs[j].dtype, b, st);
            if (rc) return rc;
            const Targets t{c->X[0], c->W[0], c->X[1], c->W[1], c->Wc[0], c->Wc[1]};
            const Geom* gj = geom[j] + (size_t)p0 * c->nsym;
            if (own_kernel(c)) {
                if ((rc = launch_own(c, p0, gj, b * c->nsym, t, st, b))) return rc;
                continue;
            }
            if ((rc = build_lists(c, gj, b * c->nsym, st))) return rc;
            units += b;
        }
        BrickJob table[MAX_JOBS];
        spbp_ctx* lead = nullptr;
        const int n = round_table(jobs, njobs, geom, r, table, &lead);
        if (!n) continue;
        Timed tm(lead, SPBP_K_INSERT, units, st);
        const int rc = tables ? launch_bricks_dev(lead, tables[r], n, lead->bnext, st)
                              : launch_bricks(lead, table, n, lead->bnext, st);
        if (rc) return rc;
    }
    return SPBP_OK;
}

int batch_rounds(const spbp_job* jobs, int njobs) {
    int rounds = 0;
    for (int j = 0; j < njobs; ++j)
        if (jobs[j].nimg > 0) rounds = std::max(rounds, (jobs[j].nimg + jobs[j].ctx->fwd_batch - 1) / jobs[j].ctx->fwd_batch);
    return rounds;
}

int stage_batch(const spbp_job* jobs, int njobs, hipStream_t st) {
    for (int j = 0; j < njobs; ++j) {
        spbp_ctx* c = jobs[j].ctx;
        if (jobs[j].nimg == 0) continue;
        int rc = check_images(c, jobs[j].imgs, jobs[j].dtype, jobs[j].nimg, st);
        if (rc) return rc;
        if ((rc = stage_geom(c, jobs[j].nimg, jobs[j].dm, jobs[j].angles_deg, jobs[j].half, st))) return rc;
        if ((rc = ensure_chunk(c, std::min(c->batch, jobs[j].nimg)))) return rc;
    }
    return SPBP_OK;
}

}  // namespace

// ------------------------------------------------------------------ batch graph
struct spbp_graph {
    int device = 0;
    int reset_first = 0;
    std::vector<spbp_job> jobs;
    std::vector<std::vector<Geom>> hgeom;   // per job, the sections one launch inserts
    std::vector<Geom*> geom;          // per job, device (owned)
    std::vector<BrickJob*> tables;    // per round, device (owned)
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool counted = false;   // the contexts' live-graph counts include this graph
};

extern "C" {

int spbp_insert_batch(const spbp_job* jobs, int njobs, void* stream) {
    int rc = prepare_batch(jobs, njobs);
    if (rc) return rc;
    DeviceGuard dg(jobs[0].ctx->device);
    hipStream_t st = (hipStream_t)stream;
    if ((rc = stage_batch(jobs, njobs, st))) return rc;
    std::vector<Geom*> geom(njobs);
    for (int j = 0; j < njobs; ++j) {
        geom[j] = jobs[j].ctx->geom_d;
        if (jobs[j].nimg > 0) jobs[j].ctx->filled = true;
    }
    rc = enqueue_batch(jobs, njobs, geom.data(), nullptr, st);
    for (int j = 0; j < njobs; ++j) {
        if (rc) jobs[j].ctx->hist_ok = false;
        else record_history(jobs[j].ctx, (size_t)jobs[j].nimg * jobs[j].ctx->nsym);
    }
    return rc;
}

int spbp_graph_destroy(spbp_graph* g) {
    if (!g) return SPBP_OK;
    DeviceGuard dg(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (Geom* p : g->geom) if (p) (void)hipFree(p);
    for (BrickJob* p : g->tables) if (p) (void)hipFree(p);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    if (g->counted)
        for (const spbp_job& j : g->jobs) --j.ctx->graphs;
    delete g;
    return SPBP_OK;
}

int spbp_graph_create(const spbp_job* jobs, int njobs, int reset_first, spbp_graph** out) {
    if (!out) return fail(nullptr, SPBP_ERR_ARG, "NULL output pointer");
    *out = nullptr;
    int rc = prepare_batch(jobs, njobs);
    if (rc) return rc;
    for (int j = 0; j < njobs; ++j)
        if (jobs[j].ctx->flags & SPBP_CHECK_FINITE)
            return fail(jobs[j].ctx, SPBP_ERR_ARG, "SPBP_CHECK_FINITE cannot be used in a captured batch");
    spbp_graph* g = new spbp_graph;
    g->device = jobs[0].ctx->device;
    g->jobs.assign(jobs, jobs + njobs);
    DeviceGuard dg(g->device);
    auto bail = [&](int code) { spbp_graph_destroy(g); return code; };
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(nullptr, SPBP_ERR_DEVICE, "stream creation failed"));
    // everything that allocates or synchronises happens before the capture:
    // geometry into graph-owned buffers, chunk buffers, FFT plans (one dry
    // pass over the chunk sizes), job tables
    if ((rc = stage_batch(jobs, njobs, g->stream))) return bail(rc);
    g->geom.assign(njobs, nullptr);
    g->hgeom.assign(njobs, {});
    g->reset_first = reset_first;
    for (int j = 0; j < njobs; ++j) {
        spbp_ctx* c = jobs[j].ctx;
        if (jobs[j].nimg == 0) continue;
        const size_t ng = (size_t)jobs[j].nimg * c->nsym;
        g->hgeom[j].assign(c->geom_h, c->geom_h + ng);
        if (dmalloc(&g->geom[j], ng * sizeof(Geom)) != hipSuccess) return bail(fail(c, SPBP_ERR_ALLOC, "graph geometry"));
        if (hipMemcpyAsync(g->geom[j], c->geom_d, ng * sizeof(Geom), hipMemcpyDeviceToDevice, g->stream) != hipSuccess)
            return bail(fail(c, SPBP_ERR_DEVICE, "graph geometry copy"));
        // run each FFT plan the batch uses once, outside the capture (lazy
        // set-up inside the FFT library then happens here); the chunk buffers
        // are scratch, so this changes no result
        const int tail = jobs[j].nimg % c->fwd_batch;
        const size_t esz = jobs[j].dtype == SPBP_BF16 ? 2 : 4;
        if ((rc = pad_fft(c, jobs[j].imgs, jobs[j].dtype, std::min(c->fwd_batch, jobs[j].nimg), g->stream))) return bail(rc);
        if (tail && jobs[j].nimg > c->fwd_batch &&
            (rc = pad_fft(c, (const char*)jobs[j].imgs + (size_t)(jobs[j].nimg - tail) * c->nx * c->nx * esz,
                          jobs[j].dtype, tail, g->stream)))
            return bail(rc);
    }
    if (hipStreamSynchronize(g->stream) != hipSuccess) return bail(fail(nullptr, SPBP_ERR_DEVICE, "graph set-up"));
    // every round's job table is uploaded now, synchronously: the captured
    // work only reads these device copies
    const int rounds = batch_rounds(jobs, njobs);
    g->tables.assign(rounds, nullptr);
    for (int r = 0; r < rounds; ++r) {
        BrickJob table[MAX_JOBS];
        spbp_ctx* lead = nullptr;
        const int n = round_table(jobs, njobs, g->geom.data(), r, table, &lead);
        if (!n) continue;
        if (dmalloc(&g->tables[r], MAX_JOBS * sizeof(BrickJob)) != hipSuccess ||
            hipMemcpy(g->tables[r], table, n * sizeof(BrickJob), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(nullptr, SPBP_ERR_ALLOC, "graph job tables"));
    }
    for (int j = 0; j < njobs; ++j) jobs[j].ctx->capturing = true;
    rc = SPBP_OK;
    if (hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        rc = fail(nullptr, SPBP_ERR_DEVICE, "hipStreamBeginCapture failed");
    } else {
        if (reset_first)
            for (int j = 0; j < njobs && rc == SPBP_OK; ++j) {
                spbp_ctx* c = jobs[j].ctx;
                for (int h = 0; h < c->nvol; ++h) {
                    fill((float*)c->X[h], 2 * c->nh, 0.0f, g->stream);
                    fill(c->W[h], c->nh, 0.0f, g->stream);
                    if (c->Wc[h]) fill(c->Wc[h], c->nh, 0.0f, g->stream);
                }
            }
        if (rc == SPBP_OK) rc = enqueue_batch(jobs, njobs, g->geom.data(), g->tables.data(), g->stream);
        const hipError_t e = hipStreamEndCapture(g->stream, &g->graph);
        if (rc == SPBP_OK && e != hipSuccess) rc = fail(nullptr, SPBP_ERR_DEVICE, "stream capture failed: %s", hipGetErrorString(e));
    }
    for (int j = 0; j < njobs; ++j) jobs[j].ctx->capturing = false;
    if (rc) return bail(rc);
    if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        return bail(fail(nullptr, SPBP_ERR_DEVICE, "hipGraphInstantiate failed"));
    // from now on the contexts' buffers, lists and plans must stay in place
    for (int j = 0; j < njobs; ++j) ++jobs[j].ctx->graphs;
    g->counted = true;
    *out = g;
    return SPBP_OK;
}

int spbp_graph_launch(spbp_graph* g, void* stream) {
    if (!g || !g->exec) return fail(nullptr, SPBP_ERR_ARG, "NULL graph");
    DeviceGuard dg(g->device);
    const hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
    for (size_t j = 0; j < g->jobs.size(); ++j) {
        spbp_ctx* c = g->jobs[j].ctx;
        if (e != hipSuccess) { c->hist_ok = false; continue; }
        if (g->reset_first) { c->hist.clear(); c->hist_ok = true; }
        if (g->jobs[j].nimg > 0) c->filled = true;
        if (c->hist_ok) c->hist.insert(c->hist.end(), g->hgeom[j].begin(), g->hgeom[j].end());
    }
    if (e != hipSuccess) return fail(nullptr, SPBP_ERR_DEVICE, "hipGraphLaunch: %s", hipGetErrorString(e));
    return SPBP_OK;
}

int spbp_insert_rb(spbp_ctx* c, const void* imgs, int dtype, int nimg, const double* angles_deg,
                   const float* inplane, const int8_t* half, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (nimg > 0 && !inplane) return fail(c, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE (no in-plane angles / shifts)");
    return insert_impl(c, imgs, dtype, nimg, nullptr, angles_deg, half, (hipStream_t)stream, inplane);
}

int spbp_rtsq(const void* imgs, int dtype, int nimg, int nx, const float* params, float* out, void* stream) {
    if (nimg < 0 || nx < 2) return fail(nullptr, SPBP_ERR_ARG, "bad image count or size");
    if (nimg == 0) return SPBP_OK;
    if (!imgs || !params || !out) return fail(nullptr, SPBP_ERR_ARG, "NULL pointer");
    if (dtype != SPBP_F32 && dtype != SPBP_BF16) return fail(nullptr, SPBP_ERR_ARG, "unknown image type %d", dtype);
    std::vector<RtsqParam> h(nimg);
    for (int p = 0; p < nimg; ++p) {
        const float* q = params + 3 * (size_t)p;
        if (!std::isfinite(q[0]) || !std::isfinite(q[1]) || !std::isfinite(q[2]))
            return fail(nullptr, SPBP_ERR_ANGLE, "non-finite rotation or shift for image %d", p + 1);
        float v[4];
        rtsq_param(nx, q[0], q[1], q[2], v);
        h[p] = RtsqParam{v[0], v[1], v[2], v[3]};
    }
    hipStream_t st = (hipStream_t)stream;
    RtsqParam* d = nullptr;
    HIPCHK(nullptr, dmalloc(&d, nimg * sizeof(RtsqParam)));
    hipError_t e = hipMemcpyAsync(d, h.data(), nimg * sizeof(RtsqParam), hipMemcpyHostToDevice, st);
    const size_t npix = (size_t)nx * nx;
    for (int p0 = 0; p0 < nimg && e == hipSuccess; p0 += 65535) {
        const int b = std::min(65535, nimg - p0);
        const dim3 grid(nblocks(npix), b);
        if (dtype == SPBP_BF16)
            rtsq_kernel<uint16_t><<<grid, TPB, 0, st>>>((const uint16_t*)imgs + p0 * npix, d + p0, nx, out + p0 * npix);
        else
            rtsq_kernel<float><<<grid, TPB, 0, st>>>((const float*)imgs + p0 * npix, d + p0, nx, out + p0 * npix);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(nullptr, SPBP_ERR_DEVICE, "rtsq: %s", hipGetErrorString(e));
    return SPBP_OK;
}

int spbp_rtsq_host(const float* imgs_host, int nimg, int nx, const float* params, float* out_host) {
    if (nimg < 0 || nx < 2) return fail(nullptr, SPBP_ERR_ARG, "bad image count or size");
    if (nimg == 0) return SPBP_OK;
    if (!imgs_host || !params || !out_host) return fail(nullptr, SPBP_ERR_ARG, "NULL pointer");
    const size_t bytes = (size_t)nimg * nx * nx * sizeof(float);
    float *din = nullptr, *dout = nullptr;
    if (dmalloc(&din, bytes) != hipSuccess) return fail(nullptr, SPBP_ERR_ALLOC, "rtsq staging: %zu bytes", bytes);
    if (dmalloc(&dout, bytes) != hipSuccess) {
        (void)hipFree(din);
        return fail(nullptr, SPBP_ERR_ALLOC, "rtsq staging: %zu bytes", bytes);
    }
    int rc = SPBP_OK;
    if (hipMemcpy(din, imgs_host, bytes, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(nullptr, SPBP_ERR_DEVICE, "rtsq: image upload failed");
    if (rc == SPBP_OK) rc = spbp_rtsq(din, SPBP_F32, nimg, nx, params, dout, nullptr);
    if (rc == SPBP_OK && hipMemcpy(out_host, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(nullptr, SPBP_ERR_DEVICE, "rtsq: image download failed");
    (void)hipFree(din);
    (void)hipFree(dout);
    return rc;
}

int spbp_rtsf(const void* imgs, int dtype, int nimg, int nx, const float* params, float* out, void* stream) {
    if (nimg < 0 || nx < 2) return fail(nullptr, SPBP_ERR_ARG, "bad image count or size");
    if (nimg == 0) return SPBP_OK;
    if (!imgs || !params || !out) return fail(nullptr, SPBP_ERR_ARG, "NULL pointer");
    if (dtype != SPBP_F32 && dtype != SPBP_BF16) return fail(nullptr, SPBP_ERR_ARG, "unknown image type %d", dtype);
    std::vector<RtsfParam> h(nimg);
    int rc = rtsf_params(nullptr, nx, nimg, params, h.data());
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    RtsfWork w;
    RtsfParam* d = nullptr;
    if (dmalloc(&d, nimg * sizeof(RtsfParam)) != hipSuccess) return fail(nullptr, SPBP_ERR_ALLOC, "RTSF parameters");
    rc = rtsf_ensure(nullptr, w, nx, rtsf_cap(nx, nimg));
    if (rc == SPBP_OK && hipMemcpyAsync(d, h.data(), nimg * sizeof(RtsfParam), hipMemcpyHostToDevice, st) != hipSuccess)
        rc = fail(nullptr, SPBP_ERR_DEVICE, "RTSF parameter upload failed");
    if (rc == SPBP_OK) rc = rtsf_run(nullptr, w, imgs, dtype, nimg, nx, d, out, st);
    if (rc == SPBP_OK && hipStreamSynchronize(st) != hipSuccess) rc = fail(nullptr, SPBP_ERR_DEVICE, "RTSF failed");
    rtsf_release(w);
    (void)hipFree(d);
    return rc;
}

int spbp_rtsf_host(const float* imgs_host, int nimg, int nx, const float* params, float* out_host) {
    if (nimg < 0 || nx < 2) return fail(nullptr, SPBP_ERR_ARG, "bad image count or size");
    if (nimg == 0) return SPBP_OK;
    if (!imgs_host || !params || !out_host) return fail(nullptr, SPBP_ERR_ARG, "NULL pointer");
    const size_t bytes = (size_t)nimg * nx * nx * sizeof(float);
    float *din = nullptr, *dout = nullptr;
    if (dmalloc(&din, bytes) != hipSuccess) return fail(nullptr, SPBP_ERR_ALLOC, "rtsf staging: %zu bytes", bytes);
    if (dmalloc(&dout, bytes) != hipSuccess) {
        (void)hipFree(din);
        return fail(nullptr, SPBP_ERR_ALLOC, "rtsf staging: %zu bytes", bytes);
    }
    int rc = SPBP_OK;
    if (hipMemcpy(din, imgs_host, bytes, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(nullptr, SPBP_ERR_DEVICE, "rtsf: image upload failed");
    if (rc == SPBP_OK) rc = spbp_rtsf(din, SPBP_F32, nimg, nx, params, dout, nullptr);
    if (rc == SPBP_OK && hipMemcpy(out_host, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(nullptr, SPBP_ERR_DEVICE, "rtsf: image download failed");
    (void)hipFree(din);
    (void)hipFree(dout);
    return rc;
}

}  // extern "C"

namespace {
// Host images staged chunk by chunk through one device buffer, each chunk
// inserted as spbp_insert (or spbp_insert_rb with inplane) would.
// Is p in pageable host memory, unknown to the runtime?  Page-locked host,
// device and managed memory are copied as they are.  (A failed query is
// pageable memory; its error is cleared.)
bool host_pageable(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return a.type == hipMemoryTypeUnregistered;
}

// dst[0 .. n) = src[0 .. n) on up to 8 host threads (contiguous slices).
void parallel_copy(void* dst, const void* src, size_t n) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<size_t>(std::min(hw, 8u), std::max<size_t>(1, n >> 22));   // >= 4 MiB each
    auto work = [&](int t) {
        const size_t a = n * t / nt, b = n * (t + 1) / nt;
        std::memcpy((char*)dst + a, (const char*)src + a, b - a);
    };
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
}

// Host images: uploaded in sub-chunks of about 256 MiB (at most c->batch
// images) into two device buffers by turns, on a stream of their own, so
// that the upload of one sub-chunk overlaps the insertion of the one before
// (the insertion of sub-chunk k waits for its upload; the upload of k + 2
// waits for that insertion to be done with the buffer).  Pageable images
// are first copied by host threads into one of two page-locked staging
// buffers (kept by the context), so that the upload is an asynchronous DMA
// and the copy of sub-chunk k + 1 overlaps the upload and insertion of k;
// when page-locked memory cannot be had, the runtime's pageable copies
// serve instead.  Page-locked, device and managed images are copied as they
// are (the runtime tells the direction).
int insert_host_impl(spbp_ctx* c, const void* imgs_host, int dtype, int nimg, const double* angles_deg,
                     const float* inplane, const int8_t* half, hipStream_t st) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (nimg <= 0) return nimg == 0 ? SPBP_OK : fail(c, SPBP_ERR_ARG, "negative image count %d", nimg);
    if (!imgs_host) return fail(c, SPBP_ERR_ARG, "NULL image pointer");
    DeviceGuard dg(c->device);
    const size_t esz = dtype == SPBP_BF16 ? 2 : 4;
    const size_t per = (size_t)c->nx * c->nx * esz;
    const int sub = (int)std::max<size_t>(1, std::min<size_t>({(size_t)nimg, (size_t)std::max(1, c->batch),
                                                                ((size_t)256 << 20) / per}));
    const int nbuf = nimg > sub ? 2 : 1;
    const bool pageable = host_pageable(imgs_host);
    bool staged = pageable;
    if (staged && c->hstage_cap < per * sub) {   // grow the page-locked staging (kept for later calls)
        for (void*& h : c->hstage)
            if (h) { (void)hipHostFree(h); h = nullptr; }
        c->hstage_cap = 0;
        bool ok = true;
        for (void*& h : c->hstage)
            if (ok && hipHostMalloc(&h, per * sub, hipHostMallocDefault) != hipSuccess) { h = nullptr; ok = false; }
        if (ok) {
            c->hstage_cap = per * sub;
        } else {   // no page-locked memory: the runtime's pageable copies instead
            (void)hipGetLastError();
            for (void*& h : c->hstage)
                if (h) { (void)hipHostFree(h); h = nullptr; }
            staged = false;
        }
    }
    void* dimg[2] = {nullptr, nullptr};
    hipStream_t cs = nullptr;
    hipEvent_t up[2] = {nullptr, nullptr}, used[2] = {nullptr, nullptr};
    int rc = SPBP_OK;
    auto release = [&]() {
        (void)hipStreamSynchronize(st);
        if (cs) { (void)hipStreamSynchronize(cs); (void)hipStreamDestroy(cs); }
        for (int k = 0; k < 2; ++k) {
            if (up[k]) (void)hipEventDestroy(up[k]);
            if (used[k]) (void)hipEventDestroy(used[k]);
            if (dimg[k]) (void)hipFree(dimg[k]);
        }
    };
    for (int k = 0; k < nbuf && rc == SPBP_OK; ++k) {
        if (dmalloc(&dimg[k], per * sub) != hipSuccess) rc = fail(c, SPBP_ERR_ALLOC, "image upload buffer");
        else if (hipEventCreateWithFlags(&up[k], hipEventDisableTiming) != hipSuccess ||
                 hipEventCreateWithFlags(&used[k], hipEventDisableTiming) != hipSuccess)
            rc = fail(c, SPBP_ERR_DEVICE, "upload events");
    }
    if (rc == SPBP_OK && hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
        cs = nullptr;
        rc = fail(c, SPBP_ERR_DEVICE, "upload stream");
    }
    for (int p0 = 0, k = 0; p0 < nimg && rc == SPBP_OK; p0 += sub, ++k) {
        const int b = std::min(sub, nimg - p0), s = k % nbuf;
        const char* src = (const char*)imgs_host + per * p0;
        hipError_t e = hipSuccess;
        if (staged) {
            // the staging buffer is free once its previous upload (sub-chunk k - 2) is done
            if (k >= 2) e = hipEventSynchronize(up[s]);
            if (e == hipSuccess) {
                parallel_copy(c->hstage[s], src, per * b);
                src = (const char*)c->hstage[s];
            }
        }
        if (e == hipSuccess && k >= nbuf) e = hipStreamWaitEvent(cs, used[s], 0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dimg[s], src, per * b, pageable ? hipMemcpyHostToDevice : hipMemcpyDefault, cs);
        if (e == hipSuccess) e = hipEventRecord(up[s], cs);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, up[s], 0);
        if (e != hipSuccess) { rc = fail(c, SPBP_ERR_DEVICE, "image upload: %s", hipGetErrorString(e)); break; }
        rc = insert_impl(c, dimg[s], dtype, b, nullptr, angles_deg ? angles_deg + 3 * (size_t)p0 : nullptr,
                         half ? half + p0 : nullptr, st, inplane ? inplane + 3 * (size_t)p0 : nullptr);
        if (rc == SPBP_OK && hipEventRecord(used[s], st) != hipSuccess) rc = fail(c, SPBP_ERR_DEVICE, "upload events");
    }
    const hipError_t e = hipStreamSynchronize(st);
    if (rc == SPBP_OK && e != hipSuccess) rc = fail(c, SPBP_ERR_DEVICE, "insert: %s", hipGetErrorString(e));
    release();
    return rc;
}
}  // namespace

extern "C" {

int spbp_insert_host(spbp_ctx* c, const void* imgs_host, int dtype, int nimg, const double* angles_deg,
                     const int8_t* half, void* stream) {
    return insert_host_impl(c, imgs_host, dtype, nimg, angles_deg, nullptr, half, (hipStream_t)stream);
}

int spbp_insert_rb_host(spbp_ctx* c, const void* imgs_host, int dtype, int nimg, const double* angles_deg,
                        const float* inplane, const int8_t* half, void* stream) {
    if (c && nimg > 0 && !inplane) return fail(c, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE (no in-plane angles / shifts)");
    return insert_host_impl(c, imgs_host, dtype, nimg, angles_deg, inplane, half, (hipStream_t)stream);
}

int spbp_finalize(spbp_ctx* c, float* out_full, float* out_h1, float* out_h2, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    spbp_final_job job{c, out_full, out_h1, out_h2};
    return spbp_finalize_batch(&job, 1, stream);
}

int spbp_finalize_batch(const spbp_final_job* jobs, int njobs, void* stream) {
    if (!jobs || njobs < 1 || njobs > MAX_JOBS)
        return fail(nullptr, SPBP_ERR_ARG, "batch of %d finalisations (1..%d)", njobs, MAX_JOBS);
    for (int j = 0; j < njobs; ++j) {
        spbp_ctx* c = jobs[j].ctx;
        if (!c) return fail(nullptr, SPBP_ERR_ARG, "finalisation %d has no context", j);
        if (c->device != jobs[0].ctx->device) return fail(c, SPBP_ERR_ARG, "finalisation %d is on another device", j);
        for (int k = 0; k < j; ++k)
            if (jobs[k].ctx == c) return fail(c, SPBP_ERR_ARG, "finalisations %d and %d share a context", k, j);
        if ((jobs[j].out_h1 || jobs[j].out_h2) && c->nvol < 2)
            return fail(c, SPBP_ERR_ARG, "sample volumes need SPBP_TWO_HALVES");
    }
    DeviceGuard dg(jobs[0].ctx->device);
    return finalize_jobs(jobs, njobs, (hipStream_t)stream);
}

int spbp_finalize_host(spbp_ctx* c, float* out_full, float* out_h1, float* out_h2, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    const size_t nv = (size_t)c->nx * c->nx * c->nx;
    float* outs[3] = {out_full, out_h1, out_h2};
    float* d = nullptr;
    HIPCHK(c, dmalloc(&d, nv * sizeof(float)));
    int rc = SPBP_OK;
    for (int q = 0; q < 3 && rc == SPBP_OK; ++q) {
        if (!outs[q]) continue;
        rc = spbp_finalize(c, q == 0 ? d : nullptr, q == 1 ? d : nullptr, q == 2 ? d : nullptr, st);
        if (rc) break;
        hipError_t e = hipMemcpyAsync(outs[q], d, nv * sizeof(float), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fail(c, SPBP_ERR_DEVICE, "volume download: %s", hipGetErrorString(e));
    }
    (void)hipFree(d);
    return rc;
}

int spbp_export_partial(spbp_ctx* c, int half, float* X, float* W, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (!X || !W) return fail(c, SPBP_ERR_ARG, "NULL output pointer");
    if (half < -1 || half >= c->nvol) return fail(c, SPBP_ERR_ARG, "no half %d", half);
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    export_kernel<<<nblocks(c->nh), TPB, 0, st>>>(make_acc(c, half), c->nx, c->N, (float2*)X, W);
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

int spbp_export_raw(spbp_ctx* c, int half, float* X, float* W, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (!X || !W) return fail(c, SPBP_ERR_ARG, "NULL output pointer");
    if (half < 0 || half >= c->nvol) return fail(c, SPBP_ERR_ARG, "no half %d", half);
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(c, hipMemcpyAsync(X, c->X[half], c->nh * sizeof(float2), hipMemcpyDeviceToDevice, st));
    HIPCHK(c, hipMemcpyAsync(W, c->W[half], c->nh * sizeof(float), hipMemcpyDeviceToDevice, st));
    return SPBP_OK;
}

int spbp_set_symmetries(spbp_ctx* c, int nsym, const double* angles_deg) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (nsym < 0 || (nsym > 1 && !angles_deg)) return fail(c, SPBP_ERR_ARG, "bad symmetry arguments");
    if (c->graphs > 0) return refuse_captured(c, "spbp_set_symmetries");
    // DMS insertion and SYMVOL use one set of symmetries per operation
    // (bp32f.f:866-872, 278-303): it is fixed while the accumulators hold sums
    if (c->filled) return fail(c, SPBP_ERR_ARG, "spbp_set_symmetries on accumulators that hold sums (spbp_reset first)");
    DeviceGuard dg(c->device);
    if (nsym <= 1) {   // one operation (the identity) or none: no symmetrisation (bp32f.f:169)
        c->nsym = 1;
        return SPBP_OK;
    }
    std::vector<float> sm(9 * (size_t)nsym);
    std::vector<double> smd(9 * (size_t)nsym);
    for (int k = 0; k < nsym; ++k) {
        const double* a = angles_deg + 3 * (size_t)k;
        if (!std::isfinite(a[0]) || !std::isfinite(a[1]) || !std::isfinite(a[2]))
            return fail(c, SPBP_ERR_ANGLE, "MISSING SYMMETRY %d", k + 1);   // BUILDS (builds.f:47-52)
        cang((float)a[2], (float)a[1], (float)a[0], (c->flags & SPBP_EXACT_PI) != 0, sm.data() + 9 * k);
        cang_d(a[2], a[1], a[0], (c->flags & SPBP_EXACT_PI) != 0, smd.data() + 9 * k);
    }
    HIPCHK(c, hipDeviceSynchronize());   // no queued work may still read the old tables
    float* smd_new = nullptr;
    HIPCHK(c, dmalloc(&smd_new, sm.size() * sizeof(float)));
    if (hipMemcpy(smd_new, sm.data(), sm.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(smd_new);
        return fail(c, SPBP_ERR_DEVICE, "symmetry table upload failed");
    }
    if (c->sm_d) (void)hipFree(c->sm_d);
    c->sm_d = smd_new;
    c->sm = sm;
    c->smd = smd;
    c->nsym = nsym;
    c->tuned = false;
    return c->fwd_batch > 0 ? ensure_lists(c, c->fwd_batch) : SPBP_OK;
}

int spbp_debug_fail_alloc(long k) {
    g_fail_after.store(k < 0 ? -1 : k);
    return SPBP_OK;
}

int spbp_kb_table(int n2, int flags, float* tab, float* fltb) {
    if (n2 < 2 || !tab || !fltb) return fail(nullptr, SPBP_ERR_ARG, "bad table arguments");
    std::vector<float> t;
    float f = 0.f;
    if (flags & SPBP_GEOM_F64) kb_table_f64_f32(n2, t, f);
    else kb_table(n2, t, f);
    std::memcpy(tab, t.data(), t.size() * sizeof(float));
    *fltb = f;
    return SPBP_OK;
}

int spbp_rotations(int n, const double* ang, int flags, float* dm) {
    if (n < 0 || (n > 0 && (!ang || !dm))) return fail(nullptr, SPBP_ERR_ARG, "bad rotation arguments");
    for (int p = 0; p < n; ++p) {
        const double* a = ang + 3 * (size_t)p;
        if (!std::isfinite(a[0]) || !std::isfinite(a[1]) || !std::isfinite(a[2]))
            return fail(nullptr, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE %d", p + 1);
        cang((float)a[2], (float)a[1], (float)a[0], (flags & SPBP_EXACT_PI) != 0, dm + 9 * (size_t)p);
    }
    return SPBP_OK;
}

int spbp_calibrate(int what, size_t bytes, double* rate, void* stream) {
    if (!rate || (what != SPBP_CAL_LDS_U64 && what != SPBP_CAL_GLOBAL_F32))
        return fail(nullptr, SPBP_ERR_ARG, "bad calibration arguments (what %d)", what);
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu <= 0)
        return fail(nullptr, SPBP_ERR_DEVICE, "no device");
    hipStream_t st = (hipStream_t)stream;
    constexpr int LNT = 2 * TPB;                 // the brick kernel's default group size
    const unsigned lblocks = 8u * (unsigned)ncu;  // four rounds of two groups per CU
    const unsigned gblocks = 8u * (unsigned)ncu;
    const size_t gwidth = (size_t)gblocks * TPB;
    const size_t n = bytes / sizeof(float);
    if (what == SPBP_CAL_GLOBAL_F32 && bytes > ((size_t)1 << 34))
        return fail(nullptr, SPBP_ERR_ARG, "calibration buffer of %zu bytes above 16 GiB", bytes);
    if (what == SPBP_CAL_GLOBAL_F32 && n < gwidth)
        return fail(nullptr, SPBP_ERR_ARG, "calibration buffer of %zu bytes below %zu", bytes, gwidth * sizeof(float));
    void* buf = nullptr;
    const size_t nbytes = what == SPBP_CAL_LDS_U64 ? lblocks * sizeof(unsigned long long) : n * sizeof(float);
    if (dmalloc(&buf, nbytes) != hipSuccess) return fail(nullptr, SPBP_ERR_ALLOC, "calibration buffer");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SPBP_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(nullptr, SPBP_ERR_DEVICE, "events");
    float ms = 0.f;
    if (rc == SPBP_OK && what == SPBP_CAL_GLOBAL_F32) {
        fill((float*)buf, n, 0.f, st);
        cal_global_kernel<<<gblocks, TPB, 0, st>>>((float*)buf, n);   // warm-up
    } else if (rc == SPBP_OK) {
        cal_lds_kernel<LNT><<<lblocks, LNT, 0, st>>>((unsigned long long*)buf);
    }
    if (rc == SPBP_OK && hipEventRecord(e0, st) != hipSuccess) rc = fail(nullptr, SPBP_ERR_DEVICE, "event record");
    if (rc == SPBP_OK) {
        if (what == SPBP_CAL_GLOBAL_F32) cal_global_kernel<<<gblocks, TPB, 0, st>>>((float*)buf, n);
        else cal_lds_kernel<LNT><<<lblocks, LNT, 0, st>>>((unsigned long long*)buf);
        if (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess || !(ms > 0.f))
            rc = fail(nullptr, SPBP_ERR_DEVICE, "calibration launch failed");
    }
    if (rc == SPBP_OK) {
        const double ops = what == SPBP_CAL_LDS_U64 ? (double)lblocks * (LNT / 64) * CAL_LDS_ITER
                                                   : (double)gwidth * CAL_GLOBAL_ITER;
        *rate = ops / (ms * 1e-3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    return rc;
}

int spbp_kernel_stats(spbp_ctx* c, int kernel, double* total_ms, long* launches, long* units) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (kernel < 0 || kernel > SPBP_K_RESOLVE) return fail(c, SPBP_ERR_ARG, "no kernel kind %d", kernel);
    DeviceGuard dg(c->device);
    double ms = 0.0;
    long u = 0;
    for (auto& r : c->recs[kernel]) {
        HIPCHK(c, hipEventSynchronize(r.b));
        float e = 0.f;
        HIPCHK(c, hipEventElapsedTime(&e, r.a, r.b));
        ms += e;
        u += r.units;
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = (long)c->recs[kernel].size();
    if (units) *units = u;
    return SPBP_OK;
}

int spbp_reset_stats(spbp_ctx* c) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    DeviceGuard dg(c->device);
    for (auto& v : c->recs) {
        for (auto& r : v) {
            (void)hipEventSynchronize(r.b);
            c->evpool.push_back(r.a);
            c->evpool.push_back(r.b);
        }
        v.clear();
    }
    return SPBP_OK;
}

int spbp_import_partial(spbp_ctx* c, int half, const float* X, const float* W, int accumulate, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (!X || !W) return fail(c, SPBP_ERR_ARG, "NULL input pointer");
    if (half < 0 || half >= c->nvol) return fail(c, SPBP_ERR_ARG, "no half %d", half);
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    if (!accumulate && c->Wc[half]) fill(c->Wc[half], c->nh, 0.0f, st);
    import_kernel<<<nblocks(c->nh), TPB, 0, st>>>((const float2*)X, W, c->nx, c->N, c->X[half], c->W[half], accumulate);
    HIPCHK(c, hipGetLastError());
    // the accumulators now hold sums the insertion history does not describe
    c->hist_ok = false;
    c->filled = true;
    return SPBP_OK;
}

int spbp_reference_weights(spbp_ctx* c, int which, const long long* voxels, int n, double* w_out, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (n < 0 || (n > 0 && (!voxels || !w_out))) return fail(c, SPBP_ERR_ARG, "bad voxel list");
    if (which < -1 || which >= c->nvol) return fail(c, SPBP_ERR_ARG, "no half %d", which);
    if (!c->hist_ok) return fail(c, SPBP_ERR_ARG, "the accumulators hold imported sums: no insertion history");
    for (int i = 0; i < n; ++i)
        if (voxels[i] < 0 || (size_t)voxels[i] >= c->nh) return fail(c, SPBP_ERR_ARG, "voxel %lld outside the half grid", voxels[i]);
    if (n == 0) return SPBP_OK;
    if (c->hist.empty()) {
        for (int i = 0; i < n; ++i) w_out[i] = 0.0;
        return SPBP_OK;
    }
    if (c->hist.size() > (size_t)INT32_MAX) return fail(c, SPBP_ERR_SIZE, "insertion history too long");
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    int rc;
    // slot 0's buffers; its pinned words may still receive a finalisation's counters
    if (c->stats_evt) HIPCHK(c, hipEventSynchronize(c->stats_evt));
    Round R = make_round(c, which, 0);
    R.nv = n;
    if ((rc = upload_history(c, st)) || (rc = ensure_res(c, R.res(), (size_t)n))) return rc;
    HIPCHK(c, hipMemcpyAsync(R.res().list, voxels, (size_t)n * sizeof(long long), hipMemcpyHostToDevice, st));
    double* d = nullptr;
    HIPCHK(c, dmalloc(&d, (size_t)n * sizeof(double)));
    rc = resolve_listed({&R}, d, st);
    if (rc == SPBP_OK && (hipMemcpyAsync(w_out, d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
                          hipStreamSynchronize(st) != hipSuccess))
        rc = fail(c, SPBP_ERR_DEVICE, "reference weights read-back failed");
    (void)hipFree(d);
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------------ allocation-free adjoint scatter
// spbp_scatter_noalloc keeps its scratch -- KB table, geometry, raw spectra,
// their maxima, brick lists, queue head and job table -- in one file-scope
// device array.  Each GPU the library runs on holds its own copy until the
// process ends.  Host code never uses the array's name as an address: the
// first call on a device (that device current) asks hipGetSymbolAddress for
// the copy's device address, kept per device, and every kernel argument,
// fill and copy uses that address.
namespace {
constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;   // the scratch array: 64 MiB per GPU
constexpr int SCRATCH_DEVICES = 64;                  // devices served
constexpr int SCRATCH_CHUNK = 256;                   // projections per internal chunk at most
}  // namespace

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[SCRATCH_DEVICES] = {};

int scratch_address(int dev, unsigned char** out) {
    if (dev < 0 || dev >= SCRATCH_DEVICES) return fail(nullptr, SPBP_ERR_DEVICE, "device %d beyond the %d served", dev, SCRATCH_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(nullptr, SPBP_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return SPBP_OK;
}

// Consecutive 256-byte-aligned pieces of the scratch array.
struct Carve {
    unsigned char* p;
    size_t left;
    template <typename T>
    T* take(size_t n) {
        const size_t b = (n * sizeof(T) + 255) & ~(size_t)255;
        if (b > left) return nullptr;
        T* r = (T*)p;
        p += b;
        left -= b;
        return r;
    }
};
}  // namespace

extern "C" {

int spbp_scatter_noalloc(int nx, int flags, const float* bi, int nimg, const double* angles_deg, const float* dm,
                         float* X, float* W, void* stream) {
    if (nx < 4 || nx > 1024) return fail(nullptr, SPBP_ERR_SIZE, "image size %d outside 4..1024", nx);
    if (flags & ~(SPBP_EXACT_PI | SPBP_GEOM_F64)) return fail(nullptr, SPBP_ERR_ARG, "flags 0x%x: only SPBP_EXACT_PI and SPBP_GEOM_F64 apply", flags);
    if (nimg < 0) return fail(nullptr, SPBP_ERR_ARG, "negative projection count %d", nimg);
    if (nimg == 0) return SPBP_OK;
    if (!bi || !X) return fail(nullptr, SPBP_ERR_ARG, "NULL spectrum or grid pointer");
    if (!dm && !angles_deg) return fail(nullptr, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE (no angles given)");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(nullptr, SPBP_ERR_DEVICE, "no HIP device");
    unsigned char* base = nullptr;
    int rc = scratch_address(dev, &base);
    if (rc) return rc;
    // a context on the stack whose buffers are pieces of the scratch array
    spbp_ctx c;
    c.nx = nx;
    c.N = 2 * nx;
    c.H = nx + 1;
    c.flags = flags;
    c.device = dev;
    c.nh = (size_t)c.H * c.N * c.N;
    c.strategy = SPBP_STRAT_BRICK;
    {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0) c.ncu = n;
    }
    int shape = NSHAPES - 1;
    c.bg = grid_of_shape(nx, flags, shape);
    c.shape = shape;
    const int N = c.N, H = c.H;
    const size_t nk = 2 * (size_t)c.bg.nb;
    Carve cv{base, SCRATCH_BYTES};
    c.tab = cv.take<float>(LTAB + 1);
    c.bcnt = cv.take<int>(nk);
    c.boff = cv.take<int>(nk + 1);
    c.bioff = cv.take<int>(nk + 1);
    c.bcur = cv.take<int>(nk);
    c.bscan = cv.take<int2>((nk + 1023) / 1024 + 1);
    c.bnext = cv.take<int>(1);
    BrickJob* job_d = cv.take<BrickJob>(1);
    // per projection: geometry, raw spectrum, its maximum, list entries and work items
    const size_t per = sizeof(Geom) + sizeof(float4) + (size_t)N * H * sizeof(float2) + sizeof(float) +
                       2 * (size_t)c.bg.nb * sizeof(int);
    const size_t fixed = nk * sizeof(int) + 7 * 256;   // the work-item keys' extra 2 nb, alignment
    size_t B = 0;
    if (c.tab && job_d && cv.left > fixed) B = std::min<size_t>((cv.left - fixed) / per, SCRATCH_CHUNK);
    B = std::min<size_t>(B, (size_t)nimg);
    if (B < 1)
        return fail(nullptr, SPBP_ERR_SIZE, "NX = %d needs more than the %zu MiB scratch of spbp_scatter_noalloc",
                    nx, SCRATCH_BYTES >> 20);
    Geom* geom_d = cv.take<Geom>(B);
    c.spec = cv.take<float2>(B * N * H);
    c.vmax = cv.take<float>(B);
    c.blist = cv.take<int>(B * c.bg.nb);
    c.bikey = cv.take<int>(B * c.bg.nb + nk);
    c.bpln = cv.take<float4>(B);
    if (!geom_d || !c.spec || !c.vmax || !c.blist || !c.bikey || !c.bpln)
        return fail(nullptr, SPBP_ERR_SIZE, "scratch layout for NX = %d does not fit", nx);
    c.blist_cap = B * c.bg.nb;
    c.bpln_cap = B;
    c.fwd_batch = (int)B;
    // host staging on the stack
    float tabh[LTAB + 1];
    Geom gh[SCRATCH_CHUNK];
    c.geom_h = gh;
    const bool f64 = (flags & SPBP_GEOM_F64) != 0, exact = (flags & SPBP_EXACT_PI) != 0;
    kb_table_into(N, f64, tabh, c.fltb);
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(nullptr, hipMemcpyAsync(c.tab, tabh, sizeof tabh, hipMemcpyHostToDevice, st));
    const Targets t{(float2*)X, W, nullptr, nullptr, nullptr, nullptr};
    for (int p0 = 0; p0 < nimg; p0 += (int)B) {
        const int b = (int)std::min<size_t>(B, (size_t)(nimg - p0));
        for (int p = 0; p < b; ++p) {
            const int k = geom_one(&c, p, 1, dm ? dm + 9 * (size_t)p0 : nullptr,
                                   angles_deg ? angles_deg + 3 * (size_t)p0 : nullptr, nullptr, exact, f64);
            if (k == 1) return fail(nullptr, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE %d", p0 + p + 1);
            if (k) return fail(nullptr, SPBP_ERR_ANGLE, "rotation of image %d is not finite and orthonormal", p0 + p + 1);
        }
        HIPCHK(nullptr, hipMemcpyAsync(geom_d, gh, (size_t)b * sizeof(Geom), hipMemcpyHostToDevice, st));
        const size_t n = (size_t)b * N * H;
        bi_to_raw_kernel<<<nblocks(n), TPB, 0, st>>>((const float2*)bi + (size_t)p0 * N * H, N, H, n, c.spec);
        spec_absmax_kernel<<<b, TPB, 0, st>>>(c.spec, N * H, c.vmax);
        HIPCHK(nullptr, hipGetLastError());
        if ((rc = build_lists(&c, geom_d, b, st))) return fail(nullptr, rc, "%s", c.err);
        const BrickJob job = brick_job(&c, 0, geom_d, b, t);
        HIPCHK(nullptr, hipMemcpyAsync(job_d, &job, sizeof job, hipMemcpyHostToDevice, st));
        if ((rc = launch_bricks_dev(&c, job_d, 1, c.bnext, st, W != nullptr))) return fail(nullptr, rc, "%s", c.err);
        // the stack staging (geometry, job table, KB table) is read by the copies above
        HIPCHK(nullptr, hipStreamSynchronize(st));
    }
    return SPBP_OK;
}

int spbp_threshold_stats(const spbp_ctx* c, long long* listed, long long* changed, long long* inconsistent) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    long long v[3] = {c->res_stats[0], c->res_stats[1], c->res_stats[2]};
    if (c->stats_slot >= 0) {   // the counters' read-back, enqueued by the finalisation
        DeviceGuard dg(c->device);
        if (hipEventSynchronize(c->stats_evt) != hipSuccess)
            return fail(nullptr, SPBP_ERR_DEVICE, "threshold counters read-back failed");
        v[1] = (long long)c->res[c->stats_slot].h[1];
        v[2] = (long long)c->res[c->stats_slot].h[2];
    }
    if (listed) *listed = v[0];
    if (changed) *changed = v[1];
    if (inconsistent) *inconsistent = v[2];
    return SPBP_OK;
}

}  // extern "C"
