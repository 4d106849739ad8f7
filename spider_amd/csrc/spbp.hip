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

namespace {

constexpr int LN2 = 2;       // half width of the 5-tap kernel (bp32f.f:193-194)
constexpr int LTAB = 4999;   // KB table length - 1 (oneline.f:44)
constexpr int TPB = 256;

thread_local std::string g_last_error;

struct Geom {        // per projection: first two rows of CANG's matrix + half flag
    float r[6];
    int half;
    int pad_;
};

// ------------------------------------------------------------------ device helpers
__device__ inline float load_img(const float* p, size_t i) { return p[i]; }
__device__ inline float load_img(const uint16_t* p, size_t i) {
    return __uint_as_float(((unsigned)p[i]) << 16);   // bf16 -> fp32 (exact)
}

template <typename T>
__device__ inline T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block reduction of (double, unsigned) pairs; result valid in thread 0.
__device__ inline void block_sum2(double& s, unsigned& c) {
    __shared__ double ss[TPB / 64];
    __shared__ unsigned sc[TPB / 64];
    s = wave_sum(s);
    c = wave_sum(c);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { ss[w] = s; sc[w] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) { s += ss[k]; c += sc[k]; }
    }
}

// ------------------------------------------------------------------ PADD2
template <typename T>
__global__ __launch_bounds__(TPB) void pad_kernel(const T* __restrict__ imgs, int nx, int N,
                                                  float* __restrict__ out) {
    const int b = blockIdx.x;
    const T* img = imgs + (size_t)b * nx * nx;
    const int c = nx / 2;                 // NC - 1 (asta.f:75)
    const float r2 = (float)c * (float)c; // RI = L/2 (bp3f.f:1316)
    double s = 0.0;
    unsigned cnt = 0;
    for (int idx = threadIdx.x; idx < nx * nx; idx += blockDim.x) {
        const int j = idx / nx, i = idx - j * nx;
        const float tj = (float)(j - c), ti = (float)(i - c);
        if (tj * tj + ti * ti > r2) { s += (double)load_img(img, idx); ++cnt; }
    }
    block_sum2(s, cnt);
    __shared__ double qs_sh;
    if (threadIdx.x == 0) qs_sh = s / (double)cnt;
    __syncthreads();
    const double qs = qs_sh;
    const int ip = (N - nx) / 2 + (nx % 2);   // bp3f.f:1336
    float* o = out + (size_t)b * N * N;
    for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
        const int y = idx / N, x = idx - y * N;
        const int jj = y - ip, ii = x - ip;
        float v = 0.0f;
        if ((unsigned)jj < (unsigned)nx && (unsigned)ii < (unsigned)nx)
            v = (float)((double)load_img(img, (size_t)jj * nx + ii) - qs);
        o[idx] = v;
    }
}

// ------------------------------------------------------------------ ONELINE scatter
// One thread per sample point (I, J) of one projection.  Positions, rounding
// and table indices use round-to-nearest fp32 operations in the reference's
// order (no contraction), so the geometry matches ONELINE bit for bit.
__global__ __launch_bounds__(TPB) void insert_direct_kernel(
    const float2* __restrict__ spec, const Geom* __restrict__ geom, int nx, int N,
    const float* __restrict__ tab, float fltb,
    float2* __restrict__ X0, float* __restrict__ W0, float2* __restrict__ X1, float* __restrict__ W1) {
    const int H = nx + 1;
    const int p = blockIdx.y;
    const int lin = blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= H * N) return;
    const int jrow = lin / H;
    const int I = lin - jrow * H;
    const int J = jrow - nx + 1;                                   // J = -NX+1 .. NX
    if (!(I * I + J * J < (N * N) / 4) || (I == 0 && J < 0)) return;   // oneline.f:53-55
    const Geom g = geom[p];
    const float fI = (float)I, fJ = (float)J;
    float x = __fadd_rn(__fmul_rn(fI, g.r[0]), __fmul_rn(fJ, g.r[3]));   // oneline.f:56-58
    float y = __fadd_rn(__fmul_rn(fI, g.r[1]), __fmul_rn(fJ, g.r[4]));
    float z = __fadd_rn(__fmul_rn(fI, g.r[2]), __fmul_rn(fJ, g.r[5]));
    const int jp = J >= 0 ? J : N + J;
    const float2 s = spec[((size_t)p * N + jp) * H + I];
    const float sg = ((I + jp) & 1) ? -1.0f : 1.0f;   // (-1)^(I+J+1), J 1-based
    float vr = s.x * sg;
    float vi = -s.y * sg;                             // SPIDER sign: conjugate
    if (x < 0.0f) { x = -x; y = -y; z = -z; vi = -vi; }   // oneline.f:60-67
    const float fN = (float)N;
    const int ix = (int)__fadd_rn(__fadd_rn(x, 0.5f), fN) - N;   // IFIX(X+0.5+N)-N
    const int iy = (int)__fadd_rn(__fadd_rn(y, 0.5f), fN) - N;
    const int iz = (int)__fadd_rn(__fadd_rn(z, 0.5f), fN) - N;
    if (!(ix <= nx - LN2 - 1 && iy >= -nx + 2 + LN2 && iy <= nx - LN2 - 1 &&
          iz >= -nx + 2 + LN2 && iz <= nx - LN2 - 1)) return;      // oneline.f:73-75
    float tx[5], ty[5], tz[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) {
        tx[l] = tab[(int)roundf(__fmul_rn(fabsf(__fsub_rn(x, (float)(ix + l - LN2))), fltb))];
        ty[l] = tab[(int)roundf(__fmul_rn(fabsf(__fsub_rn(y, (float)(iy + l - LN2))), fltb))];
        tz[l] = tab[(int)roundf(__fmul_rn(fabsf(__fsub_rn(z, (float)(iz + l - LN2))), fltb))];
    }
    float2* __restrict__ X = g.half ? X1 : X0;
    float* __restrict__ W = g.half ? W1 : W0;
    const size_t plane = (size_t)N * H;
    for (int lz = 0; lz < 5; ++lz) {
        if (tz[lz] == 0.0f) continue;
        const int izp = iz + lz - LN2;
        const size_t za = (size_t)(izp >= 0 ? izp : N + izp) * plane;
        const size_t zt = (size_t)(izp <= 0 ? -izp : N - izp) * plane;   // wrap(-izp)
        for (int ly = 0; ly < 5; ++ly) {
            const float tyz = __fmul_rn(ty[ly], tz[lz]);
            if (tyz == 0.0f) continue;
            const int iyp = iy + ly - LN2;
            const size_t ya = (size_t)(iyp >= 0 ? iyp : N + iyp) * H;
            const size_t yt = (size_t)(iyp <= 0 ? -iyp : N - iyp) * H;
#pragma unroll
            for (int lx = 0; lx < 5; ++lx) {
                const float wg = __fmul_rn(tx[lx], tyz);
                if (wg == 0.0f) continue;
                const int px = ix + lx - LN2;
                if (px >= 0) {                                   // direct taps
                    const size_t o = za + ya + px;
                    unsafeAtomicAdd(&X[o].x, __fmul_rn(vr, wg));
                    unsafeAtomicAdd(&X[o].y, __fmul_rn(vi, wg));
                    unsafeAtomicAdd(&W[o], wg);
                } else {                                         // reflected taps, conj
                    const size_t o = zt + yt + (size_t)(-px);
                    unsafeAtomicAdd(&X[o].x, __fmul_rn(vr, wg));
                    unsafeAtomicAdd(&X[o].y, __fmul_rn(-vi, wg));
                    unsafeAtomicAdd(&W[o], wg);
                }
            }
        }
    }
}

// ------------------------------------------------------------------ accumulator views
struct Acc {
    const float2* X[2];
    const float* W[2];
    int n;   // halves summed: 1 or 2
};

// (j, k) of the kx = 0 plane is paired by SYMPLANE0 with (N-j, N-k) unless it
// lies on the Nyquist row/column or is the origin (bp32f.f:1057-1087).
__device__ inline bool sym_pair(int j, int k, int nx) {
    return j != nx && k != nx && !(j == 0 && k == 0);
}

// Value of SYMPLANE0(half) at (i, j, k), halves added after (bp32f.f:925-930, 345-346).
__device__ inline void load_sym(const Acc& a, int nx, int N, int i, int j, int k, float2& x, float& w) {
    const int H = nx + 1;
    const size_t o = ((size_t)k * N + j) * H + i;
    const bool pr = (i == 0) && sym_pair(j, k, nx);
    const size_t om = pr ? ((size_t)((N - k) % N) * N + (N - j) % N) * H : 0;
    x = make_float2(0.f, 0.f);
    w = 0.f;
    for (int h = 0; h < a.n; ++h) {
        float2 xh = a.X[h][o];
        float wh = a.W[h][o];
        if (pr) {
            const float2 xm = a.X[h][om];
            xh.x = xh.x + xm.x;
            xh.y = xh.y + (-xm.y);
            wh = wh + a.W[h][om];
        }
        if (h == 0) { x = xh; w = wh; }
        else { x.x = x.x + xh.x; x.y = x.y + xh.y; w = w + wh; }
    }
}

__device__ inline float2 norm_val(const Acc& a, int nx, int N, int i, int j, int k) {
    float2 x;
    float w;
    load_sym(a, nx, N, i, j, k, x, w);
    if (w > 0.1f) {                                    // NRMW2 (bp3f.f:1359)
        const float s = ((i + j + k) & 1) ? -1.0f : 1.0f;
        return make_float2(__fdiv_rn(x.x * s, w), __fdiv_rn(x.y * s, w));
    }
    return make_float2(0.f, 0.f);
}

__global__ __launch_bounds__(TPB) void normalize_kernel(Acc a, int nx, int N, float2* __restrict__ out) {
    const int H = nx + 1;
    const size_t total = (size_t)H * N * N;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int i = (int)(idx % H);
    const size_t r = idx / H;
    const int j = (int)(r % N), k = (int)(r / N);
    float2 v = norm_val(a, nx, N, i, j, k);
    if (i == 0 || i == nx) {
        // c2r reads the kx = 0 and kx = N/2 planes as their Hermitian part
        const float2 m = norm_val(a, nx, N, i, (N - j) % N, (N - k) % N);
        v = make_float2((v.x + m.x) * 0.5f, (v.y - m.y) * 0.5f);
    }
    out[idx] = make_float2(v.x, -v.y);                 // SPIDER inverse: conjugate first
}

// ------------------------------------------------------------------ WINDKB2A
__global__ __launch_bounds__(TPB) void window_kernel(const float* __restrict__ real, float* __restrict__ out,
                                                     const float* __restrict__ wkb, int nx, int N, float pix,
                                                     double* __restrict__ psum, unsigned* __restrict__ pcnt) {
    const int L = nx;
    const size_t total = (size_t)L * L * L;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    unsigned c = 0;
    if (idx < total) {
        const int i = (int)(idx % L);
        const size_t t = idx / L;
        const int j = (int)(t % L), k = (int)(t / L);
        const int ip = (N - L) / 2 + (L % 2);
        const int cc = L / 2;
        float v = real[((size_t)(ip + k) * N + (ip + j)) * N + (ip + i)] * pix;   // FMRS_3 scale
        const int lr = (i - cc) * (i - cc) + (j - cc) * (j - cc) + (k - cc) * (k - cc);
        if (lr <= cc * cc) {
            v = __fdiv_rn(v, wkb[lr]);
            if (lr >= (cc - 1) * (cc - 1)) { s = v; c = 1; }
        } else {
            v = 0.0f;
        }
        out[idx] = v;
    }
    block_sum2(s, c);
    if (threadIdx.x == 0) { psum[blockIdx.x] = s; pcnt[blockIdx.x] = c; }
}

__global__ __launch_bounds__(TPB) void shell_mean_kernel(const double* __restrict__ psum, const unsigned* __restrict__ pcnt,
                                                         int nb, float* __restrict__ mean) {
    double s = 0.0;
    unsigned c = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) { s += psum[b]; c += pcnt[b]; }
    block_sum2(s, c);
    if (threadIdx.x == 0) mean[0] = (float)(s / (double)(c ? c : 1));
}

__global__ __launch_bounds__(TPB) void subtract_kernel(float* __restrict__ out, const float* __restrict__ mean, int nx) {
    const int L = nx;
    const size_t total = (size_t)L * L * L;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int i = (int)(idx % L);
    const size_t t = idx / L;
    const int j = (int)(t % L), k = (int)(t / L);
    const int cc = L / 2;
    const int lr = (i - cc) * (i - cc) + (j - cc) * (j - cc) + (k - cc) * (k - cc);
    if (lr <= cc * cc) out[idx] = out[idx] - mean[0];
}

// ------------------------------------------------------------------ partial volumes
__global__ __launch_bounds__(TPB) void export_kernel(Acc a, int nx, int N, float2* __restrict__ X, float* __restrict__ W) {
    const int H = nx + 1;
    const size_t total = (size_t)H * N * N;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int i = (int)(idx % H);
    const size_t r = idx / H;
    float2 x;
    float w;
    load_sym(a, nx, N, i, (int)(r % N), (int)(r / N), x, w);
    X[idx] = x;
    W[idx] = w;
}

// Adds (or sets) a post-SYMPLANE0 half-grid into raw accumulators.  Entries of
// SYMPLANE0 pairs are halved (exactly) so that SYMPLANE0 of the result gives
// back the partial: P(a)/2 + conj(P(-a)/2) = P(a) for Hermitian P.
__global__ __launch_bounds__(TPB) void import_kernel(const float2* __restrict__ Xp, const float* __restrict__ Wp, int nx, int N,
                                                     float2* __restrict__ X, float* __restrict__ W, int accumulate) {
    const int H = nx + 1;
    const size_t total = (size_t)H * N * N;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int i = (int)(idx % H);
    const size_t r = idx / H;
    const int j = (int)(r % N), k = (int)(r / N);
    float2 x = Xp[idx];
    float w = Wp[idx];
    if (i == 0 && sym_pair(j, k, nx)) { x.x *= 0.5f; x.y *= 0.5f; w *= 0.5f; }
    if (accumulate) { x.x += X[idx].x; x.y += X[idx].y; w += W[idx]; }
    X[idx] = x;
    W[idx] = w;
}

__global__ __launch_bounds__(TPB) void fill_kernel(float* __restrict__ p, size_t n, float v) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) p[idx] = v;
}

template <typename T>
__global__ __launch_bounds__(TPB) void finite_kernel(const T* __restrict__ p, size_t n, int* __restrict__ flag) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n && !isfinite(load_img(p, idx))) atomicOr(flag, 1);
}

// ------------------------------------------------------------------ host math
// Modified Bessel I1 by its power series (converges for the 0..6.5 used here).
double bessel_i1(double x) {
    double h = 0.5 * x, term = h, sum = h, h2 = h * h;
    for (int k = 1; k < 200; ++k) {
        term *= h2 / (double(k) * double(k + 1));
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    return sum;
}
// I_{3/2}(x) = sqrt(2/(pi x)) (cosh x - sinh x / x), x > 0.
double bessel_i15(double x) { return std::sqrt(2.0 / (M_PI * x)) * (std::cosh(x) - std::sinh(x) / x); }
double bessel_j15(double x) { return std::sqrt(2.0 / (M_PI * x)) * (std::sin(x) / x - std::cos(x)); }

// KB interpolation table TABI(0:4999) (bp32f.f:193-214), evaluated in FP64.
void kb_table(int n2, std::vector<float>& tab, float& fltb_f) {
    const double alpha = 6.5, V = 4.0 / 2.0 / (double)n2, aaaa = 0.9 * V;
    const double b0 = std::sqrt(alpha) * bessel_i1(alpha);
    const double fltb = (double)LTAB / (double)(LN2 + 1);
    tab.assign(LTAB + 1, 0.0f);
    for (int i = 0; i <= LTAB; ++i) {
        const double s = (double)i / fltb / (double)n2;
        if (s <= aaaa) {
            const double q = s / aaaa, x = std::sqrt(1.0 - q * q);
            tab[i] = (float)(std::sqrt(alpha * x) * bessel_i1(alpha * x) / b0);
        }
    }
    fltb_f = (float)LTAB / (float)(LN2 + 1);   // REAL(LTAB)/REAL(LN2+1)
}

// |WKB| of WINDKB2A per integer squared radius LR = 0..(L/2)^2, FP64.
void wkb_table(int nx, int N, std::vector<float>& t) {
    const int cc = nx / 2;
    const double alpha = 6.5, xnu = 1.5, aaaa = 0.9 * (4.0 / 2.0 / (double)N);
    const double wkb0 = std::pow(alpha, xnu) / bessel_i15(alpha);
    const double qrt = (2.0 * M_PI * aaaa) * (2.0 * M_PI * aaaa);
    t.assign((size_t)cc * cc + 1, 1.0f);
    for (int lr = 0; lr <= cc * cc; ++lr) {
        const double sigma = qrt * lr - alpha * alpha;
        double wkb;
        if (std::fabs(sigma) < 1.0e-7) wkb = 1.0;
        else if (sigma > 0.0) { const double art = std::sqrt(sigma); wkb = wkb0 * bessel_j15(art) / std::pow(art, xnu); }
        else { const double art = std::sqrt(-sigma); wkb = wkb0 * bessel_i15(art) / std::pow(art, xnu); }
        t[lr] = (float)std::fabs(wkb);
    }
}

// CANG (cang.f:44-82) in FP64 from REAL doc registers, stored as REAL.  The
// degree conversion is float32(pi)/180 evaluated in double (cang.f:50-53)
// unless exact_pi.
void cang(float phi, float theta, float psi, bool exact_pi, float* dm) {
    const double dgr = exact_pi ? M_PI / 180.0 : (double)(float)3.14159265358979323846 / 180.0;
    const double cphi = std::cos((double)phi * dgr), sphi = std::sin((double)phi * dgr);
    const double cthe = std::cos((double)theta * dgr), sthe = std::sin((double)theta * dgr);
    const double cpsi = std::cos((double)psi * dgr), spsi = std::sin((double)psi * dgr);
    dm[0] = (float)(cphi * cthe * cpsi - sphi * spsi);
    dm[1] = (float)(sphi * cthe * cpsi + cphi * spsi);
    dm[2] = (float)(-sthe * cpsi);
    dm[3] = (float)(-cphi * cthe * spsi - sphi * cpsi);
    dm[4] = (float)(-sphi * cthe * spsi + cphi * cpsi);
    dm[5] = (float)(sthe * spsi);
    dm[6] = (float)(sthe * cphi);
    dm[7] = (float)(sthe * sphi);
    dm[8] = (float)(cthe);
}

inline unsigned nblocks(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }

}  // namespace

// ------------------------------------------------------------------ context
struct spbp_ctx {
    int nx = 0, N = 0, H = 0, flags = 0, device = 0, batch = 0, nvol = 1;
    size_t nh = 0;
    float2* X[2] = {nullptr, nullptr};
    float* W[2] = {nullptr, nullptr};
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
    // SPBP_TIMING: (start, stop, units) per recorded launch, per kernel kind
    struct Rec { hipEvent_t a, b; long units; };
    std::vector<Rec> recs[4];
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
        for (int q = 0; q < 6; ++q) {
            if (!std::isfinite(dm[q])) return fail(c, SPBP_ERR_ANGLE, "non-finite rotation for image %d", p + 1);
            g.r[q] = dm[q];
        }
        g.half = 0;
        if (half) {
            if (half[p] != 0 && half[p] != 1) return fail(c, SPBP_ERR_ARG, "half flag %d of image %d is not 0/1", half[p], p + 1);
            if (half[p] == 1 && c->nvol < 2) return fail(c, SPBP_ERR_ARG, "half 1 needs SPBP_TWO_HALVES");
            g.half = half[p];
        }
        g.pad_ = 0;
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
        {
            Timed t(c, SPBP_K_INSERT, b, st);
            dim3 grid(nblocks((size_t)H * N), b);
            insert_direct_kernel<<<grid, TPB, 0, st>>>(c->spec, c->geom_d + p0, nx, N, c->tab, c->fltb,
                                                       c->X[0], c->W[0], c->X[1], c->W[1]);
        }
        HIPCHK(c, hipGetLastError());
    }
    return SPBP_OK;
}

Acc make_acc(const spbp_ctx* c, int which) {
    Acc a;
    a.X[0] = a.X[1] = nullptr;
    a.W[0] = a.W[1] = nullptr;
    if (which < 0) {
        a.n = c->nvol;
        for (int h = 0; h < c->nvol; ++h) { a.X[h] = c->X[h]; a.W[h] = c->W[h]; }
    } else {
        a.n = 1;
        a.X[0] = c->X[which];
        a.W[0] = c->W[which];
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
    int rc = SPBP_OK;
    auto bail = [&](int code) { spbp_destroy(c); return code; };
    for (int h = 0; h < c->nvol; ++h) {
        if (hipMalloc(&c->X[h], c->nh * sizeof(float2)) != hipSuccess ||
            hipMalloc(&c->W[h], c->nh * sizeof(float)) != hipSuccess)
            return bail(fail(nullptr, SPBP_ERR_ALLOC, "BP 32F; XE, WE, WO, & XO: %zu bytes", c->nh * 12 * (size_t)c->nvol));
    }
    std::vector<float> tab;
    kb_table(c->N, tab, c->fltb);
    if (hipMalloc(&c->tab, tab.size() * sizeof(float)) != hipSuccess ||
        hipMemcpy(c->tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&c->flag, sizeof(int)) != hipSuccess ||
        hipEventCreateWithFlags(&c->geom_evt, hipEventDisableTiming) != hipSuccess)
        return bail(fail(nullptr, SPBP_ERR_DEVICE, "context set-up failed"));
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
    }
    void* dev_ptrs[] = {c->tab, c->padbuf, c->spec, c->geom_d, c->work, c->real, c->wkb, c->psum, c->pcnt, c->mean, c->flag};
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
    if (kernel < 0 || kernel > 3) return fail(c, SPBP_ERR_ARG, "no kernel kind %d", kernel);
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
    import_kernel<<<nblocks(c->nh), TPB, 0, st>>>((const float2*)X, W, c->nx, c->N, c->X[half], c->W[half], accumulate);
    HIPCHK(c, hipGetLastError());
    return SPBP_OK;
}

}  // extern "C"
