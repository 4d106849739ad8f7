// finalize.hpp -- SYMPLANE0 / NRMW2 / WINDKB2A kernels and partial volumes.
#pragma once
#include "common.hpp"

namespace spbp {

// ------------------------------------------------------------------ accumulator views
struct Acc {
    const float2* X[2];
    const float* W[2];
    const float* Wc[2];   // compensation terms of the weights, or null
    int n;                // halves summed: 1 or 2
};

// (j, k) of the kx = 0 plane is paired by SYMPLANE0 with (N-j, N-k) unless it
// lies on the Nyquist row/column or is the origin (bp32f.f:1057-1087).
__device__ inline bool sym_pair(int j, int k, int nx) {
    return j != nx && k != nx && !(j == 0 && k == 0);
}

// Value of SYMPLANE0(half) at (i, j, k), halves added after (bp32f.f:925-930, 345-346).
__device__ inline void load_sym(const Acc& a, int nx, int N, int i, int j, int k, float2& x, float& w, double& wd) {
    const int H = nx + 1;
    const size_t o = ((size_t)k * N + j) * H + i;
    const bool pr = (i == 0) && sym_pair(j, k, nx);
    const size_t om = pr ? ((size_t)((N - k) % N) * N + (N - j) % N) * H : 0;
    x = make_float2(0.f, 0.f);
    w = 0.f;
    wd = 0.0;   // FP64 weight incl. the compensation terms (when kept)
    for (int h = 0; h < a.n; ++h) {
        float2 xh = a.X[h][o];
        float wh = a.W[h][o];
        double wdh = a.Wc[h] ? (double)wh + (double)a.Wc[h][o] : (double)wh;
        if (pr) {
            const float2 xm = a.X[h][om];
            xh.x = xh.x + xm.x;
            xh.y = xh.y + (-xm.y);
            wh = wh + a.W[h][om];
            wdh += a.Wc[h] ? (double)a.W[h][om] + (double)a.Wc[h][om] : (double)a.W[h][om];
        }
        if (h == 0) { x = xh; w = wh; }
        else { x.x = x.x + xh.x; x.y = x.y + xh.y; w = w + wh; }
        wd += wdh;
    }
}

__device__ inline float2 norm_val(const Acc& a, int nx, int N, int i, int j, int k) {
    float2 x;
    float w;
    double wd;
    load_sym(a, nx, N, i, j, k, x, w, wd);
    const float s = ((i + j + k) & 1) ? -1.0f : 1.0f;
    if (a.Wc[0]) {
        // compensated weights: threshold and division in FP64
        if (wd > (double)0.1f)
            return make_float2((float)((double)(x.x * s) / wd), (float)((double)(x.y * s) / wd));
        return make_float2(0.f, 0.f);
    }
    if (w > 0.1f)                                      // NRMW2 (bp3f.f:1359)
        return make_float2(__fdiv_rn(x.x * s, w), __fdiv_rn(x.y * s, w));
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
    double wd;
    load_sym(a, nx, N, i, (int)(r % N), (int)(r / N), x, w, wd);
    X[idx] = x;
    W[idx] = a.Wc[0] ? (float)wd : w;
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

}  // namespace spbp
