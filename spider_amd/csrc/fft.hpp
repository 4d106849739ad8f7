// fft.hpp -- PADD2 + the forward 2D R2C FFT of the padded images
// (fmrs.f / fftw3.f: FFTW's forward sign), for power-of-two N, in two passes
// that skip the padding: three quarters of each padded image are zeros.
//
//   pad_mean_kernel (block per image): the ASTA mean (asta.f, FP64) of the
//                   pixels outside the inscribed circle;
//   row_fft_kernel  (block per image x group of rows): the NX nonzero rows of
//                   the padded image, two real rows per complex FFT of length
//                   N, unpacked to their N/2 + 1 outputs, written to rows
//                   IP .. IP + NX - 1 of the spectrum;
//   col_fft_kernel  (block per image x slab of columns): each column's length-N
//                   transform from those NX rows (the others are zero),
//                   written to all N rows, and the image's largest |Re|, |Im|.
//
// The spectrum layout is the R2C layout of a row-major N x N transform: N
// rows of N/2 + 1 complex values, X[j][k] = sum_y sum_x P[y][x] e^{-2 pi i
// (x k + y j) / N}.  The transforms run in LDS as Stockham radix-4 passes (one
// radix-2 pass first when log2 N is odd), twiddles from an FP64-computed
// table of e^{-2 pi i m / N} copied into LDS.  Global loads are issued in
// groups of FFT_BATCH before their values are used, so that many are in
// flight per thread.
#pragma once
#include "common.hpp"

namespace spbp {

#ifndef SPBP_FFT_ROW_ELEMS
#define SPBP_FFT_ROW_ELEMS 2048
#endif
constexpr int FFT_ELEMS = 4096;             // complex values a columns-pass block transforms at a time
constexpr int FFT_LDS = FFT_ELEMS + 64;     // plus padding of the column layout
constexpr int FFT_BATCH = 8;                // global loads in flight per thread
// The rows pass transforms fewer values per block, so that more blocks share
// a CU and their loads overlap: measured per 20,000 images of 128² (8,000 of
// 256²), rows pass 1.78 ms (2.70) with 2048 values per block against 2.42
// (3.30) with 4096.  The columns pass keeps 4096: its row segments are then
// 16 (8) columns wide, and 2048 made it 2.45 → 3.38 ms (5.57 → 7.91).
constexpr int FFT_ROW_ELEMS = SPBP_FFT_ROW_ELEMS;
// Threads of a columns-pass block: 512 (eight waves over the same LDS slab)
// hide its loads better than 256 (2.38 -> 2.31 ms per 20,000 images of 128^2,
// 5.43 -> 5.14 ms per 8,000 of 256^2).
constexpr int FFT_COL_NT = 512;

// Dynamic LDS of the transform kernels (E values per block): the data, then
// the N twiddles.
inline size_t fft_lds_bytes(int N, int E = FFT_ELEMS) { return (size_t)(E + 64 + N) * sizeof(float2); }
__device__ inline void load_twiddles(float2* stw, const float2* __restrict__ tw, int N) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) stw[i] = tw[i];
}

__device__ inline float2 c_mul(float2 a, float2 b) {
    return make_float2(__fsub_rn(__fmul_rn(a.x, b.x), __fmul_rn(a.y, b.y)),
                       __fadd_rn(__fmul_rn(a.x, b.y), __fmul_rn(a.y, b.x)));
}
__device__ inline float2 c_add(float2 a, float2 b) { return make_float2(__fadd_rn(a.x, b.x), __fadd_rn(a.y, b.y)); }
__device__ inline float2 c_sub(float2 a, float2 b) { return make_float2(__fsub_rn(a.x, b.x), __fsub_rn(a.y, b.y)); }

// Forward DFTs of M sequences of length N = 2^L in LDS (sequence m at
// buf + m S), in place, by a block of NT threads: every pass reads all its
// inputs into registers, synchronises, then writes.  M N <= E.  Ends
// synchronised.  The per-thread loops have fixed trip counts so that their
// arrays stay in registers.
template <int E, int NT = TPB>
__device__ inline void lds_fft(float2* buf, int M, int N, int L, int S, const float2* __restrict__ tw) {
    static_assert(E >= 4 * NT, "a radix-4 pass gives every thread at least one butterfly slot");
    int Ns = 1;
    if (L & 1) {   // radix-2 pass
        constexpr int U = E / 2 / NT;
        const int half = N >> 1, nb = M * half;
        float2 a[U], b[U];
        int dst[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = threadIdx.x + u * NT;
            dst[u] = -1;
            if (t < nb) {
                const int m = t >> (L - 1), j = t & (half - 1);
                const float2* s = buf + m * S;
                const float2 v0 = s[j], v1 = s[j + half];   // Ns = 1: no twiddle
                a[u] = c_add(v0, v1);
                b[u] = c_sub(v0, v1);
                dst[u] = m * S + 2 * j;
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (dst[u] >= 0) { buf[dst[u]] = a[u]; buf[dst[u] + 1] = b[u]; }
        __syncthreads();
        Ns = 2;
    }
    for (; Ns < N; Ns <<= 2) {   // radix-4 passes
        constexpr int U = E / 4 / NT;
        const int q = N >> 2, nb = M * q, step = N / (4 * Ns);
        float2 x0[U], x1[U], x2[U], x3[U];
        int dst[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = threadIdx.x + u * NT;
            dst[u] = -1;
            if (t < nb) {
                const int m = t >> (L - 2), j = t & (q - 1);
                const int k = j & (Ns - 1);
                const float2* s = buf + m * S;
                const float2 v0 = s[j];
                const float2 v1 = c_mul(s[j + q], tw[k * step]);
                const float2 v2 = c_mul(s[j + 2 * q], tw[2 * k * step]);
                const float2 v3 = c_mul(s[j + 3 * q], tw[3 * k * step]);
                const float2 a0 = c_add(v0, v2), a1 = c_sub(v0, v2), a2 = c_add(v1, v3);
                const float2 d = c_sub(v1, v3);
                const float2 a3 = make_float2(d.y, -d.x);   // (v1 - v3) (-i)
                x0[u] = c_add(a0, a2);
                x1[u] = c_add(a1, a3);
                x2[u] = c_sub(a0, a2);
                x3[u] = c_sub(a1, a3);
                dst[u] = m * S + (j - k) * 4 + k;
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (dst[u] >= 0) {
                buf[dst[u]] = x0[u];
                buf[dst[u] + Ns] = x1[u];
                buf[dst[u] + 2 * Ns] = x2[u];
                buf[dst[u] + 3 * Ns] = x3[u];
            }
        __syncthreads();
    }
}

// PADD2's mean: FP64 mean of the pixels outside the inscribed circle
// (asta.f, bp3f.f:1316); block per image.  Also zeroes the image's vmax.
// Each thread reads 16 bytes at a time (4 fp32 or 8 BF16 pixels), FFT_BATCH /
// 2 such loads in flight; an image that does not start 16-byte aligned, or
// whose pixel count is not a multiple of the vector (NX < 4), takes the
// scalar loop.
template <typename T>
__global__ __launch_bounds__(TPB) void pad_mean_kernel(const T* __restrict__ imgs, int nx, int lnx,
                                                       double* __restrict__ mean, float* __restrict__ vmax) {
    // nx = 2^lnx
    constexpr int VW = 16 / sizeof(T);   // pixels per 16-byte load
    const int b = blockIdx.x;
    const T* img = imgs + (size_t)b * nx * nx;
    const int c = nx / 2, n = nx * nx;
    const float r2 = (float)c * (float)c;
    double s = 0.0;
    unsigned cnt = 0;
    auto add = [&](int idx, float v) {
        const int j = idx >> lnx, i = idx & (nx - 1);
        const float tj = (float)(j - c), ti = (float)(i - c);
        if (tj * tj + ti * ti > r2) { s += (double)v; ++cnt; }
    };
    if (n % VW == 0 && ((uintptr_t)img & 15) == 0) {
        constexpr int MB = FFT_BATCH / 2;
        const int nv = n / VW;
        const uint4* img4 = reinterpret_cast<const uint4*>(img);
        for (int i0 = threadIdx.x; i0 < nv; i0 += MB * TPB) {
            uint4 q[MB];
#pragma unroll
            for (int u = 0; u < MB; ++u) {
                const int k = i0 + u * TPB;
                q[u] = k < nv ? img4[k] : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int u = 0; u < MB; ++u) {
                const int k = i0 + u * TPB;
                if (k >= nv) continue;
                const unsigned w[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if constexpr (sizeof(T) == 4) {
                        add(k * VW + e, __uint_as_float(w[e]));
                    } else {   // two BF16 pixels per word, the lower address in the low half
                        add(k * VW + 2 * e, __uint_as_float(w[e] << 16));
                        add(k * VW + 2 * e + 1, __uint_as_float(w[e] & 0xFFFF0000u));
                    }
                }
            }
        }
    } else {
        for (int idx = threadIdx.x; idx < n; idx += TPB) add(idx, load_img(img, idx));
    }
    block_sum2(s, cnt);
    if (threadIdx.x == 0) { mean[b] = s / (double)cnt; vmax[b] = 0.0f; }
}

// Four consecutive pixels (an aligned 16- or 8-byte load) as fp32.
__device__ inline float4 load_img4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ inline float4 load_img4(const uint16_t* p) {
    const uint2 w = *reinterpret_cast<const uint2*>(p);   // pixel 0 in the low half of w.x
    return make_float4(__uint_as_float(w.x << 16), __uint_as_float(w.x & 0xFFFF0000u), __uint_as_float(w.y << 16),
                       __uint_as_float(w.y & 0xFFFF0000u));
}

// Rows pass.  Block per (image, group of FFT_ROW_ELEMS / N row pairs); the
// spectrum's rows IP .. IP + NX - 1 receive the row transforms of the padded
// image (PADD2: image minus the FP64 mean, bp3f.f:1316-1345).
template <typename T>
__global__ __launch_bounds__(TPB) void row_fft_kernel(const T* __restrict__ imgs, int nx, int N, int L,
                                                      const float2* __restrict__ tw, const double* __restrict__ mean,
                                                      float2* __restrict__ spec) {
    extern __shared__ float2 lds[];
    float2* buf = lds;
    float2* stw = lds + FFT_ROW_ELEMS + 64;
    load_twiddles(stw, tw, N);
    const int b = blockIdx.y;
    const T* img = imgs + (size_t)b * nx * nx;
    const double qs = mean[b];
    const int H = nx + 1, ip = (N - nx) / 2 + (nx % 2);
    const int pairs = nx / 2, M = min(FFT_ROW_ELEMS / N, pairs);
    const int p0 = blockIdx.x * M, mc = min(M, pairs - p0);
    const int total = mc * N;
    if ((nx & 3) == 0 && ((uintptr_t)img & (4 * sizeof(T) - 1)) == 0) {
        // four pixels of each of the pair's rows per load (N = 2 NX: the
        // padding is NX columns per sequence, zeroed separately)
        const int lq = L - 3, nq = 1 << lq;   // quads per row: NX / 4
        const int totq = mc * nq;
        for (int i = threadIdx.x; i < mc * nx; i += TPB) {
            const int m = i >> (L - 1), n = i & (nx - 1);
            buf[m * N + (n < ip ? n : n + nx)] = make_float2(0.f, 0.f);
        }
        constexpr int BQ = FFT_BATCH / 2;
        for (int q0 = threadIdx.x; q0 < totq; q0 += BQ * TPB) {
            float4 a[BQ], c[BQ];
#pragma unroll
            for (int u = 0; u < BQ; ++u) {
                const int q = q0 + u * TPB;
                const int m = q >> lq, xq = q & (nq - 1);
                const int r = 2 * (p0 + m);
                if (q < totq) {
                    a[u] = load_img4(img + (size_t)r * nx + 4 * xq);
                    c[u] = load_img4(img + (size_t)(r + 1) * nx + 4 * xq);
                }
            }
#pragma unroll
            for (int u = 0; u < BQ; ++u) {
                const int q = q0 + u * TPB;
                if (q >= totq) continue;
                const int m = q >> lq, xq = q & (nq - 1);
                float2* o = buf + m * N + ip + 4 * xq;
                o[0] = make_float2((float)((double)a[u].x - qs), (float)((double)c[u].x - qs));
                o[1] = make_float2((float)((double)a[u].y - qs), (float)((double)c[u].y - qs));
                o[2] = make_float2((float)((double)a[u].z - qs), (float)((double)c[u].z - qs));
                o[3] = make_float2((float)((double)a[u].w - qs), (float)((double)c[u].w - qs));
            }
        }
    } else {
        for (int i0 = threadIdx.x; i0 < total; i0 += FFT_BATCH * TPB) {
            float2 v[FFT_BATCH];
#pragma unroll
            for (int u = 0; u < FFT_BATCH; ++u) {
                const int idx = i0 + u * TPB;
                const int m = idx >> L, x = (idx & (N - 1)) - ip;
                const int r = 2 * (p0 + m);
                v[u] = make_float2(0.f, 0.f);
                if (idx < total && (unsigned)x < (unsigned)nx) {
                    v[u].x = (float)((double)load_img(img, (size_t)r * nx + x) - qs);
                    v[u].y = (float)((double)load_img(img, (size_t)(r + 1) * nx + x) - qs);
                }
            }
#pragma unroll
            for (int u = 0; u < FFT_BATCH; ++u)
                if (i0 + u * TPB < total) buf[i0 + u * TPB] = v[u];
        }
    }
    __syncthreads();
    lds_fft<FFT_ROW_ELEMS>(buf, mc, N, L, N, stw);
    // Z = FFT(x_r + i x_{r+1}): X_r[k] = (Z[k] + conj Z[N-k]) / 2,
    // X_{r+1}[k] = -i (Z[k] - conj Z[N-k]) / 2
    float2* out = spec + (size_t)b * N * H;
    const int hn = N >> 1;   // k = 0 .. N/2 - 1 by shifts, then the N/2 column of each pair
    for (int idx = threadIdx.x; idx < mc * (hn + 1); idx += blockDim.x) {
        int m, k;
        if (idx < mc * hn) { m = idx >> (L - 1); k = idx & (hn - 1); }
        else { m = idx - mc * hn; k = hn; }
        const float2 z = buf[m * N + k], w = buf[m * N + ((N - k) & (N - 1))];
        const int r = 2 * (p0 + m);
        out[(size_t)(ip + r) * H + k] = make_float2(0.5f * __fadd_rn(z.x, w.x), 0.5f * __fsub_rn(z.y, w.y));
        out[(size_t)(ip + r + 1) * H + k] = make_float2(0.5f * __fadd_rn(z.y, w.y), 0.5f * __fsub_rn(w.x, z.x));
    }
}

// Columns pass.  Block per (image, slab of FFT_ELEMS / N columns): the
// columns' transforms over all N rows, from rows IP .. IP + NX - 1; the
// image's largest |Re|, |Im| into vmax (zeroed by pad_mean_kernel).
__global__ __launch_bounds__(FFT_COL_NT) void col_fft_kernel(int nx, int N, int L, const float2* __restrict__ tw,
                                                      float2* __restrict__ spec, float* __restrict__ vmax) {
    extern __shared__ float2 lds[];
    float2* buf = lds;
    float2* stw = lds + FFT_LDS;
    __shared__ float wmax[FFT_COL_NT / 64];
    load_twiddles(stw, tw, N);
    const int b = blockIdx.y;
    const int H = nx + 1, ip = (N - nx) / 2 + (nx % 2);
    const int CW = FFT_ELEMS / N, c0 = blockIdx.x * CW, cw = min(CW, H - c0);
    const int S = N + 1;   // odd column stride: a row of the slab spreads over the banks
    float2* sp = spec + (size_t)b * N * H + c0;
    // element idx of the slab: row idx >> lc, column idx & (cp - 1), with cp
    // the power of two >= cw (columns >= cw idle)
    int lc = 0;
    while ((1 << lc) < cw) ++lc;
    const int cp = 1 << lc;
    const int total = cp * nx;   // only rows IP .. IP + NX - 1 are nonzero
    for (int i0 = threadIdx.x; i0 < total; i0 += FFT_BATCH * FFT_COL_NT) {
        float2 v[FFT_BATCH];
#pragma unroll
        for (int u = 0; u < FFT_BATCH; ++u) {
            const int idx = i0 + u * FFT_COL_NT;
            const int n = idx >> lc, m = idx & (cp - 1);
            v[u] = (idx < total && m < cw) ? sp[(size_t)(ip + n) * H + m] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < FFT_BATCH; ++u) {
            const int idx = i0 + u * FFT_COL_NT;
            const int n = idx >> lc, m = idx & (cp - 1);
            if (idx < total && m < cw) buf[m * S + ip + n] = v[u];
        }
    }
    for (int idx = threadIdx.x; idx < cp * (N - nx); idx += blockDim.x) {   // the zero rows
        const int n = idx >> lc, m = idx & (cp - 1);
        if (m < cw) buf[m * S + (n < ip ? n : n + nx)] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    lds_fft<FFT_ELEMS, FFT_COL_NT>(buf, cw, N, L, S, stw);
    float mx = 0.0f;
    for (int idx = threadIdx.x; idx < cp * N; idx += blockDim.x) {
        const int n = idx >> lc, m = idx & (cp - 1);
        if (m >= cw) continue;
        const float2 v = buf[m * S + n];
        sp[(size_t)n * H + m] = v;
        mx = fmaxf(mx, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < FFT_COL_NT / 64; ++k) mx = fmaxf(mx, wmax[k]);
        atomicMax((unsigned*)&vmax[b], __float_as_uint(mx));   // non-negative floats order as their bits
    }
}

}  // namespace spbp
