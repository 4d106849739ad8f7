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

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "spbp.h"


#include "common.hpp"
#include "finalize.hpp"
#include "hostmath.hpp"
#include "insert.hpp"

using namespace spbp;

namespace {
thread_local std::string g_last_error;
// brick of the LDS-binned strategy (voxels); 3 x 8 B per voxel in LDS
constexpr int BRICK_X = 16, BRICK_Y = 16, BRICK_Z = 8;
}  // namespace

// ------------------------------------------------------------------ context
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
    Geom* geom_d = nullptr;
    Geom* geom_h = nullptr;   // pinned
    size_t geom_cap = 0;
    hipEvent_t geom_evt = nullptr;
    hipfftHandle fwd = 0;
    int fwd_batch = 0;
    hipfftHandle fwd_tail = 0;
    int tail_batch = 0;
    // finalisation
    hipfftHandle inv = 0;
    float2* work = nullptr;
    float* real = nullptr;
    float* wkb = nullptr;
    double* psum = nullptr;
    unsigned* pcnt = nullptr;
    float* mean = nullptr;
    int* flag = nullptr;
    // host staging for _host entry points
    void* stage = nullptr;
    size_t stage_cap = 0;
    // brick strategy
    int strategy = SPBP_STRAT_BRICK;
    int C = 64;
    int ncu = 256;
    BrickGrid bg{};
    int* bcnt = nullptr;    // [2 nb] hits per (brick, half)
    int* boff = nullptr;    // [2 nb + 1] list offsets
    int* bioff = nullptr;   // [2 nb + 1] work-item offsets
    int* bcur = nullptr;    // [2 nb] fill cursors
    int* bnext = nullptr;   // work-queue head
    int* blist = nullptr;   // projection lists
    int* bikey = nullptr;   // work item -> key
    size_t blist_cap = 0;
    // SPBP_TIMING: (start, stop, units) per recorded launch, per kernel kind
    struct Rec { hipEvent_t a, b; long units; };
    std::vector<Rec> recs[5];
    std::vector<hipEvent_t> evpool;
    std::string err;
};

namespace {

int fail(spbp_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    if (ctx) ctx->err = buf;
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
        if (c->flags & SPBP_TIMING) {
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

int ensure_chunk(spbp_ctx* c, int want_batch) {
    if (c->padbuf && c->fwd_batch >= want_batch) return SPBP_OK;
    if (c->padbuf) { (void)hipFree(c->padbuf); c->padbuf = nullptr; }
    if (c->spec) { (void)hipFree(c->spec); c->spec = nullptr; }
    if (c->fwd) { hipfftDestroy(c->fwd); c->fwd = 0; }
    if (c->fwd_tail) { hipfftDestroy(c->fwd_tail); c->fwd_tail = 0; c->tail_batch = 0; }
    const int N = c->N, H = c->H, B = want_batch;
    HIPCHK(c, hipMalloc(&c->padbuf, (size_t)B * N * N * sizeof(float)));
    HIPCHK(c, hipMalloc(&c->spec, (size_t)B * N * H * sizeof(float2)));
    if (c->strategy == SPBP_STRAT_BRICK) {
        const size_t cap = (size_t)B * c->bg.nb;
        if (c->blist) { (void)hipFree(c->blist); c->blist = nullptr; }
        if (c->bikey) { (void)hipFree(c->bikey); c->bikey = nullptr; }
        HIPCHK(c, hipMalloc(&c->blist, cap * sizeof(int)));
        // items <= sum over keys of ceil(cnt / C) <= pairs + keys
        HIPCHK(c, hipMalloc(&c->bikey, (cap + 2 * (size_t)c->bg.nb) * sizeof(int)));
        c->blist_cap = cap;
    }
    int n[2] = {N, N};
    int inembed[2] = {N, N};
    int onembed[2] = {N, H};
    FFTCHK(c, hipfftPlanMany(&c->fwd, 2, n, inembed, 1, N * N, onembed, 1, N * H, HIPFFT_R2C, B));
    c->fwd_batch = B;
    return SPBP_OK;
}

hipfftHandle plan_for(spbp_ctx* c, int b, int* rc) {
    *rc = SPBP_OK;
    if (b == c->fwd_batch) return c->fwd;
    if (c->fwd_tail && c->tail_batch == b) return c->fwd_tail;
    if (c->fwd_tail) { hipfftDestroy(c->fwd_tail); c->fwd_tail = 0; }
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
    if (c->geom_h) { (void)hipEventSynchronize(c->geom_evt); (void)hipHostFree(c->geom_h); c->geom_h = nullptr; }
    if (c->geom_d) { (void)hipFree(c->geom_d); c->geom_d = nullptr; }
    const size_t cap = std::max<size_t>(n, 1024);
    HIPCHK(c, hipHostMalloc((void**)&c->geom_h, cap * sizeof(Geom), hipHostMallocDefault));
    HIPCHK(c, hipMalloc(&c->geom_d, cap * sizeof(Geom)));
    c->geom_cap = cap;
    return SPBP_OK;
}

int ensure_final(spbp_ctx* c) {
    if (c->work) return SPBP_OK;
    const int N = c->N, nx = c->nx;
    HIPCHK(c, hipMalloc(&c->work, c->nh * sizeof(float2)));
    HIPCHK(c, hipMalloc(&c->real, (size_t)N * N * N * sizeof(float)));
    std::vector<float> w;
    wkb_table(nx, N, w);
    HIPCHK(c, hipMalloc(&c->wkb, w.size() * sizeof(float)));
    HIPCHK(c, hipMemcpy(c->wkb, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    const size_t nb = nblocks((size_t)nx * nx * nx);
    HIPCHK(c, hipMalloc(&c->psum, nb * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->pcnt, nb * sizeof(unsigned)));
    HIPCHK(c, hipMalloc(&c->mean, sizeof(float)));
    FFTCHK(c, hipfftPlan3d(&c->inv, N, N, N, HIPFFT_C2R));
    return SPBP_OK;
}

int insert_impl(spbp_ctx* c, const void* imgs, int dtype, int nimg, const float* dm_in,
                const double* ang, const int8_t* half, hipStream_t st) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (nimg < 0) return fail(c, SPBP_ERR_ARG, "negative image count %d", nimg);
    if (nimg == 0) return SPBP_OK;
    if (!imgs) return fail(c, SPBP_ERR_ARG, "NULL image pointer");
    if (dtype != SPBP_F32 && dtype != SPBP_BF16) return fail(c, SPBP_ERR_ARG, "unknown image type %d", dtype);
    if (!dm_in && !ang) return fail(c, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE (no angles given)");
    DeviceGuard dg(c->device);
    int rc = ensure_geom(c, (size_t)nimg);
    if (rc) return rc;
    const bool exact = (c->flags & SPBP_EXACT_PI) != 0;
    for (int p = 0; p < nimg; ++p) {
        Geom& g = c->geom_h[p];
        float dm[9];
        if (dm_in) {
            std::memcpy(dm, dm_in + 9 * (size_t)p, sizeof dm);
        } else {
            const double* a = ang + 3 * (size_t)p;
            if (!std::isfinite(a[0]) || !std::isfinite(a[1]) || !std::isfinite(a[2]))
                return fail(c, SPBP_ERR_ANGLE, "MISSING ANGLE FOR IMAGE %d", p + 1);
            cang((float)a[2], (float)a[1], (float)a[0], exact, dm);   // CANG(PHI, THETA, PSI)
        }
        for (int q = 0; q < 9; ++q) {
            if (!std::isfinite(dm[q])) return fail(c, SPBP_ERR_ANGLE, "non-finite rotation for image %d", p + 1);
            g.r[q] = dm[q];
        }
        g.half = 0;
        if (half) {
            if (half[p] != 0 && half[p] != 1) return fail(c, SPBP_ERR_ARG, "half flag %d of image %d is not 0/1", half[p], p + 1);
            if (half[p] == 1 && c->nvol < 2) return fail(c, SPBP_ERR_ARG, "half 1 needs SPBP_TWO_HALVES");
            g.half = half[p];
        }
        g.pad_[0] = g.pad_[1] = 0;
    }
    HIPCHK(c, hipMemcpyAsync(c->geom_d, c->geom_h, (size_t)nimg * sizeof(Geom), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipEventRecord(c->geom_evt, st));
    const int nx = c->nx, N = c->N, H = c->H;
    const size_t esz = dtype == SPBP_BF16 ? 2 : 4;
    if (c->flags & SPBP_CHECK_FINITE) {
        const size_t n = (size_t)nimg * nx * nx;
        fill_kernel<<<1, TPB, 0, st>>>((float*)c->flag, 1, 0.0f);   // zero bits
        if (dtype == SPBP_BF16)
            finite_kernel<uint16_t><<<nblocks(n), TPB, 0, st>>>((const uint16_t*)imgs, n, c->flag);
        else
            finite_kernel<float><<<nblocks(n), TPB, 0, st>>>((const float*)imgs, n, c->flag);
        int f = 0;
        HIPCHK(c, hipMemcpyAsync(&f, c->flag, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if (f) return fail(c, SPBP_ERR_NONFINITE, "NaN or Inf in the images");
    }
    rc = ensure_chunk(c, std::min(c->batch, nimg));
    if (rc) return rc;
    for (int p0 = 0; p0 < nimg; p0 += c->fwd_batch) {
        const int b = std::min(c->fwd_batch, nimg - p0);
        const char* src = (const char*)imgs + (size_t)p0 * nx * nx * esz;
        {
            Timed t(c, SPBP_K_PAD, b, st);
            if (dtype == SPBP_BF16)
                pad_kernel<uint16_t><<<b, TPB, 0, st>>>((const uint16_t*)src, nx, N, c->padbuf);
            else
                pad_kernel<float><<<b, TPB, 0, st>>>((const float*)src, nx, N, c->padbuf);
        }
        HIPCHK(c, hipGetLastError());
        hipfftHandle plan = plan_for(c, b, &rc);
        if (rc) return rc;
        FFTCHK(c, hipfftSetStream(plan, st));
        {
            Timed t(c, SPBP_K_FFT2, b, st);
            FFTCHK(c, hipfftExecR2C(plan, (hipfftReal*)c->padbuf, (hipfftComplex*)c->spec));
        }
        if (c->strategy == SPBP_STRAT_DIRECT) {
            Timed t(c, SPBP_K_INSERT, b, st);
            dim3 grid(nblocks((size_t)H * N), b);
            insert_direct_kernel<<<grid, TPB, 0, st>>>(c->spec, c->geom_d + p0, nx, N, c->tab, c->fltb,
                                                       c->X[0], c->W[0], c->X[1], c->W[1], c->Wc[0], c->Wc[1]);
        } else {
            const int nkeys = 2 * c->bg.nb;
            const dim3 lgrid(c->bg.nb, (b + 1023) / 1024);
            {
                Timed t(c, SPBP_K_BIN, b, st);
                fill_kernel<<<nblocks(nkeys), TPB, 0, st>>>((float*)c->bcnt, nkeys, 0.0f);   // zero bits
                brick_count_kernel<<<lgrid, TPB, 0, st>>>(c->geom_d + p0, b, c->bg, nx, c->bcnt);
                brick_scan_kernel<<<1, 1024, 0, st>>>(c->bcnt, nkeys, c->C, c->boff, c->bioff, c->bcur, c->bnext);
                brick_fill_kernel<<<lgrid, TPB, 0, st>>>(c->geom_d + p0, b, c->bg, nx, c->boff, c->bcur, c->blist);
                brick_itemkey_kernel<<<nblocks(nkeys), TPB, 0, st>>>(c->bioff, nkeys, c->bikey);
            }
            HIPCHK(c, hipGetLastError());
            Timed t(c, SPBP_K_INSERT, b, st);
            insert_brick_kernel<BRICK_X, BRICK_Y, BRICK_Z><<<c->ncu * 3, TPB, 0, st>>>(
                c->spec, c->geom_d + p0, nx, N, c->tab, c->fltb, c->bg, c->C, c->bcnt, c->boff, c->bioff,
                c->bikey, c->blist, c->bnext, c->X[0], c->W[0], c->X[1], c->W[1], c->Wc[0], c->Wc[1]);
        }
        HIPCHK(c, hipGetLastError());
    }
    return SPBP_OK;
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

int finalize_one(spbp_ctx* c, int which, float* out, hipStream_t st) {
    const int nx = c->nx, N = c->N;
    Timed t(c, SPBP_K_FINALIZE, 1, st);
    normalize_kernel<<<nblocks(c->nh), TPB, 0, st>>>(make_acc(c, which), nx, N, c->work);
    HIPCHK(c, hipGetLastError());
    FFTCHK(c, hipfftSetStream(c->inv, st));
    FFTCHK(c, hipfftExecC2R(c->inv, (hipfftComplex*)c->work, (hipfftReal*)c->real));
    const size_t nv = (size_t)nx * nx * nx;
    const unsigned nb = nblocks(nv);
    const float pix = 1.0f / (float)((double)N * N * N);   // PIX = 1.0/FLOAT(N^3) (fftw3.f:444)
    window_kernel<<<nb, TPB, 0, st>>>(c->real, out, c->wkb, nx, N, pix, c->psum, c->pcnt);
    shell_mean_kernel<<<1, TPB, 0, st>>>(c->psum, c->pcnt, (int)nb, c->mean);
    subtract_kernel<<<nb, TPB, 0, st>>>(out, c->mean, nx);
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* spbp_version(void) { return "spbp 0.1 (gfx950)"; }

const char* spbp_last_error(const spbp_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_last_error.c_str();
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
        const size_t per = (size_t)c->N * c->N * 4 + (size_t)c->N * c->H * 8;
        c->batch = (int)std::max<size_t>(16, std::min<size_t>(4096, ((size_t)512 << 20) / per));
    }
    c->strategy = o->strategy == SPBP_STRAT_DIRECT ? SPBP_STRAT_DIRECT : SPBP_STRAT_BRICK;
    if (o->strategy < 0 || o->strategy > SPBP_STRAT_BRICK) { delete c; return fail(nullptr, SPBP_ERR_ARG, "unknown strategy %d", o->strategy); }
    c->C = o->brick_proj > 0 ? o->brick_proj : 64;
    {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0) c->ncu = n;
    }
    c->bg.bx = BRICK_X;
    c->bg.by = BRICK_Y;
    c->bg.bz = BRICK_Z;
    c->bg.nbx = (c->nx + BRICK_X - 1) / BRICK_X;
    c->bg.nby = (2 * c->nx - 1 + BRICK_Y - 1) / BRICK_Y;
    c->bg.nbz = (2 * c->nx - 1 + BRICK_Z - 1) / BRICK_Z;
    c->bg.nb = c->bg.nbx * c->bg.nby * c->bg.nbz;
    int rc = SPBP_OK;
    auto bail = [&](int code) { spbp_destroy(c); return code; };
    for (int h = 0; h < c->nvol; ++h) {
        if (hipMalloc(&c->X[h], c->nh * sizeof(float2)) != hipSuccess ||
            hipMalloc(&c->W[h], c->nh * sizeof(float)) != hipSuccess ||
            ((c->flags & SPBP_TWOSUM_W) && hipMalloc(&c->Wc[h], c->nh * sizeof(float)) != hipSuccess))
            return bail(fail(nullptr, SPBP_ERR_ALLOC, "BP 32F; XE, WE, WO, & XO: %zu bytes", c->nh * 12 * (size_t)c->nvol));
    }
    std::vector<float> tab;
    kb_table(c->N, tab, c->fltb);
    if (hipMalloc(&c->tab, tab.size() * sizeof(float)) != hipSuccess ||
        hipMemcpy(c->tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&c->flag, sizeof(int)) != hipSuccess ||
        hipEventCreateWithFlags(&c->geom_evt, hipEventDisableTiming) != hipSuccess)
        return bail(fail(nullptr, SPBP_ERR_DEVICE, "context set-up failed"));
    {
        const size_t nk = 2 * (size_t)c->bg.nb;
        if (hipMalloc(&c->bcnt, nk * sizeof(int)) != hipSuccess || hipMalloc(&c->boff, (nk + 1) * sizeof(int)) != hipSuccess ||
            hipMalloc(&c->bioff, (nk + 1) * sizeof(int)) != hipSuccess || hipMalloc(&c->bcur, nk * sizeof(int)) != hipSuccess ||
            hipMalloc(&c->bnext, sizeof(int)) != hipSuccess)
            return bail(fail(nullptr, SPBP_ERR_ALLOC, "brick lists"));
    }
    rc = spbp_reset(c, nullptr);
    if (rc) return bail(rc);
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(nullptr, SPBP_ERR_DEVICE, "zeroing accumulators failed"));
    *out = c;
    return SPBP_OK;
}

int spbp_destroy(spbp_ctx* c) {
    if (!c) return SPBP_OK;
    DeviceGuard dg(c->device);
    if (c->geom_evt) (void)hipEventSynchronize(c->geom_evt);
    for (int h = 0; h < 2; ++h) {
        if (c->X[h]) (void)hipFree(c->X[h]);
        if (c->W[h]) (void)hipFree(c->W[h]);
        if (c->Wc[h]) (void)hipFree(c->Wc[h]);
    }
    void* dev_ptrs[] = {c->tab, c->padbuf, c->spec, c->geom_d, c->work, c->real, c->wkb, c->psum, c->pcnt, c->mean, c->flag,
                        c->bcnt, c->boff, c->bioff, c->bcur, c->bnext, c->blist, c->bikey};
    for (void* p : dev_ptrs) if (p) (void)hipFree(p);
    if (c->geom_h) (void)hipHostFree(c->geom_h);
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->fwd) hipfftDestroy(c->fwd);
    if (c->fwd_tail) hipfftDestroy(c->fwd_tail);
    if (c->inv) hipfftDestroy(c->inv);
    if (c->geom_evt) (void)hipEventDestroy(c->geom_evt);
    spbp_reset_stats(c);
    for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
    delete c;
    return SPBP_OK;
}

int spbp_reset(spbp_ctx* c, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    for (int h = 0; h < c->nvol; ++h) {
        fill_kernel<<<nblocks(2 * c->nh), TPB, 0, st>>>((float*)c->X[h], 2 * c->nh, 0.0f);
        fill_kernel<<<nblocks(c->nh), TPB, 0, st>>>(c->W[h], c->nh, 0.0f);
        if (c->Wc[h]) fill_kernel<<<nblocks(c->nh), TPB, 0, st>>>(c->Wc[h], c->nh, 0.0f);
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

int spbp_insert_host(spbp_ctx* c, const void* imgs_host, int dtype, int nimg, const double* angles_deg,
                     const int8_t* half, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (nimg <= 0) return nimg == 0 ? SPBP_OK : fail(c, SPBP_ERR_ARG, "negative image count %d", nimg);
    if (!imgs_host) return fail(c, SPBP_ERR_ARG, "NULL image pointer");
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    const size_t esz = dtype == SPBP_BF16 ? 2 : 4;
    const size_t per = (size_t)c->nx * c->nx * esz;
    const int chunk = std::max(1, std::min(nimg, c->batch));
    void* dimg = nullptr;
    HIPCHK(c, hipMalloc(&dimg, per * chunk));
    int rc = SPBP_OK;
    for (int p0 = 0; p0 < nimg && rc == SPBP_OK; p0 += chunk) {
        const int b = std::min(chunk, nimg - p0);
        hipError_t e = hipMemcpyAsync(dimg, (const char*)imgs_host + per * p0, per * b, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) { rc = fail(c, SPBP_ERR_DEVICE, "image upload: %s", hipGetErrorString(e)); break; }
        rc = insert_impl(c, dimg, dtype, b, nullptr, angles_deg + 3 * (size_t)p0, half ? half + p0 : nullptr, st);
        e = hipStreamSynchronize(st);
        if (rc == SPBP_OK && e != hipSuccess) rc = fail(c, SPBP_ERR_DEVICE, "insert: %s", hipGetErrorString(e));
    }
    (void)hipFree(dimg);
    return rc;
}

int spbp_finalize(spbp_ctx* c, float* out_full, float* out_h1, float* out_h2, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if ((out_h1 || out_h2) && c->nvol < 2) return fail(c, SPBP_ERR_ARG, "sample volumes need SPBP_TWO_HALVES");
    DeviceGuard dg(c->device);
    int rc = ensure_final(c);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (out_full && (rc = finalize_one(c, -1, out_full, st))) return rc;
    if (out_h1 && (rc = finalize_one(c, 0, out_h1, st))) return rc;
    if (out_h2 && (rc = finalize_one(c, 1, out_h2, st))) return rc;
    return SPBP_OK;
}

int spbp_finalize_host(spbp_ctx* c, float* out_full, float* out_h1, float* out_h2, void* stream) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    DeviceGuard dg(c->device);
    hipStream_t st = (hipStream_t)stream;
    const size_t nv = (size_t)c->nx * c->nx * c->nx;
    float* outs[3] = {out_full, out_h1, out_h2};
    float* d = nullptr;
    HIPCHK(c, hipMalloc(&d, nv * sizeof(float)));
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

int spbp_kernel_stats(spbp_ctx* c, int kernel, double* total_ms, long* launches, long* units) {
    if (!c) return fail(nullptr, SPBP_ERR_ARG, "NULL context");
    if (kernel < 0 || kernel > 4) return fail(c, SPBP_ERR_ARG, "no kernel kind %d", kernel);
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
    if (!accumulate && c->Wc[half]) fill_kernel<<<nblocks(c->nh), TPB, 0, st>>>(c->Wc[half], c->nh, 0.0f);
    import_kernel<<<nblocks(c->nh), TPB, 0, st>>>((const float2*)X, W, c->nx, c->N, c->X[half], c->W[half], accumulate);
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

}  // extern "C"
