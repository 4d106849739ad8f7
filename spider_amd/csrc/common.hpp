// common.hpp -- shared device definitions of libspbp (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace spbp {

constexpr int LN2 = 2;       // half width of the 5-tap kernel (bp32f.f:193-194)
constexpr int LTAB = 4999;   // KB table length - 1 (oneline.f:44)
constexpr int TPB = 256;


struct Geom {        // per inserted section: CANG's matrix DM(1..9) (times a symmetry), half flag
    float r[9];
    int half;
    int img;         // image (spectrum) index, counted from the call's first image
    int pad_;
    double d[6];     // DM(1..6) in FP64 (used by the FP64-geometry instantiations only)
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


// p[0 .. n) = v, four values per thread (one 16-byte store when p is
// 16-byte aligned).  Launched by fill(): a kernel, not hipMemsetAsync, so that
// it may be captured in a graph.
__global__ __launch_bounds__(TPB) void fill_kernel(float* __restrict__ p, size_t n, float v) {
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i + 4 <= n && ((uintptr_t)p & 15) == 0) {
        *reinterpret_cast<float4*>(p + i) = make_float4(v, v, v, v);
    } else {
        for (size_t k = i; k < n && k < i + 4; ++k) p[k] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(TPB) void finite_kernel(const T* __restrict__ p, size_t n, int* __restrict__ flag) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n && !isfinite(load_img(p, idx))) atomicOr(flag, 1);
}

inline unsigned nblocks(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }
inline void fill(float* p, size_t n, float v, hipStream_t st) {
    if (n) fill_kernel<<<nblocks((n + 3) / 4), TPB, 0, st>>>(p, n, v);
}

// q = n / d and r = n mod d for n < 2^52 and d >= 1, given inv = 1.0 / d
// rounded to double: the product n * inv is within 1 of n / d, so one
// correction step makes the quotient exact (no 64-bit integer division).
__host__ __device__ inline void fast_divmod(unsigned long long n, unsigned d, double inv, unsigned long long& q,
                                            unsigned& r) {
    unsigned long long qq = (unsigned long long)((double)n * inv);
    long long rr = (long long)(n - qq * d);
    if (rr < 0) { --qq; rr += d; }
    else if (rr >= (long long)d) { ++qq; rr -= d; }
    q = qq;
    r = (unsigned)rr;
}

// Half-grid index ((k N + j) H + i, H = NX + 1) -> (i, j, k).
__host__ __device__ inline void half_coords(unsigned long long idx, int H, int N, double invH, double invN, int& i,
                                            int& j, int& k) {
    unsigned long long r, kk;
    unsigned ii, jj;
    fast_divmod(idx, (unsigned)H, invH, r, ii);
    fast_divmod(r, (unsigned)N, invN, kk, jj);
    i = (int)ii;
    j = (int)jj;
    k = (int)kk;
}

// ------------------------------------------------------------------ compensated weights
// W += x with the rounding error of that very addition kept in Wc: the
// returning atomic gives the exact prior value, so the error of
// s = fl(old + x) follows from Knuth's two-sum (needs round-to-nearest,
// no contraction); `extra` is a further error term to carry (e.g. the
// rounding of an FP64 partial sum to fp32).
__device__ inline void twosum_atomic_add(float* __restrict__ W, float* __restrict__ Wc, size_t o, float x, float extra) {
    const float old = atomicAdd(&W[o], x);
    const float s = __fadd_rn(old, x);
    const float bb = __fsub_rn(s, old);
    const float e = __fadd_rn(__fsub_rn(old, __fsub_rn(s, bb)), __fsub_rn(x, bb));
    const float c = __fadd_rn(e, extra);
    if (c != 0.0f) unsafeAtomicAdd(&Wc[o], c);
}

// ------------------------------------------------------------------ ONELINE sample geometry
// Shared by every insertion strategy so that all of them compute the sample
// positions, nearest voxels and table indices with the same round-to-nearest
// fp32 operations as the reference (oneline.f:56-105).

// Position of lattice sample (I, J): I*DM(1:3) + J*DM(4:6).
__device__ inline void sample_pos(const Geom& g, int I, int J, float& x, float& y, float& z) {
    const float fI = (float)I, fJ = (float)J;
    x = __fadd_rn(__fmul_rn(fI, g.r[0]), __fmul_rn(fJ, g.r[3]));
    y = __fadd_rn(__fmul_rn(fI, g.r[1]), __fmul_rn(fJ, g.r[4]));
    z = __fadd_rn(__fmul_rn(fI, g.r[2]), __fmul_rn(fJ, g.r[5]));
}

// IFIX(V + 0.5 + N) - N (oneline.f:69-71)
__device__ inline int nearest_voxel(float v, int N) {
    return (int)__fadd_rn(__fadd_rn(v, 0.5f), (float)N) - N;
}

// oneline.f:73-75
__device__ inline bool in_bounds(int ix, int iy, int iz, int nx) {
    return ix <= nx - LN2 - 1 && iy >= -nx + 2 + LN2 && iy <= nx - LN2 - 1 &&
           iz >= -nx + 2 + LN2 && iz <= nx - LN2 - 1;
}

// TABI(NINT(ABS(D) * FLTB))
__device__ inline float kb_weight(const float* __restrict__ tab, float d, float fltb) {
    return tab[(int)roundf(__fmul_rn(fabsf(d), fltb))];
}

// FP64-geometry variants (SPBP_GEOM_F64): the same steps on FP64 positions
// from the FP64 rotation, FLTB = 4999/3 in FP64.  They depart from the
// reference's fp32 geometry on purpose, to agree with an FP64 computation.
__device__ inline void sample_pos(const Geom& g, int I, int J, double& x, double& y, double& z) {
    const double fI = (double)I, fJ = (double)J;
    x = __dadd_rn(__dmul_rn(fI, g.d[0]), __dmul_rn(fJ, g.d[3]));
    y = __dadd_rn(__dmul_rn(fI, g.d[1]), __dmul_rn(fJ, g.d[4]));
    z = __dadd_rn(__dmul_rn(fI, g.d[2]), __dmul_rn(fJ, g.d[5]));
}
__device__ inline int nearest_voxel(double v, int N) {
    return (int)__dadd_rn(__dadd_rn(v, 0.5), (double)N) - N;
}
__device__ inline float kb_weight(const float* __restrict__ tab, double d, double fltb) {
    return tab[(int)round(__dmul_rn(fabs(d), fltb))];
}
// position minus an integer voxel coordinate, in the position's precision
__device__ inline float dsub(float a, int b) { return __fsub_rn(a, (float)b); }
__device__ inline double dsub(double a, int b) { return __dsub_rn(a, (double)b); }
// FLTB in the position's precision
template <typename P> __device__ inline P fltb_of(float f);
template <> __device__ inline float fltb_of<float>(float f) { return f; }
template <> __device__ inline double fltb_of<double>(float) { return 4999.0 / 3.0; }

// BI(I, JP) of the raw R2C output with the SPIDER sign (conjugate,
// fftw3.f:411-419) and the checkerboard (-1)^(I+J+1), J 1-based (bp32f.f:859-864).
__device__ inline float2 spectrum_value(const float2* __restrict__ spec_p, int I, int J, int N, int H) {
    const int jp = J >= 0 ? J : N + J;
    const float2 s = spec_p[(size_t)jp * H + I];
    const float sg = ((I + jp) & 1) ? -1.0f : 1.0f;
    return make_float2(s.x * sg, -s.y * sg);
}

}  // namespace spbp
