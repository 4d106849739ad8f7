// resolve.hpp -- NRMW2's weight test (W > 0.1, bp3f.f:1359) decided by the
// reference's own sums.
//
// The test is discontinuous: where W lies within rounding distance of 0.1,
// two correct summation orders can land on opposite sides, and the Fourier
// voxel is then divided by W in one and zeroed in the other (after the 3D FFT
// one such voxel moves the whole volume by ~1e-4 relative).  The reference
// sums each voxel's weights sequentially in fp32: projections in input order,
// symmetries in order (bp32f.f:866-898), rows J ascending, I ascending within
// a row, and within a sample its direct taps before its reflected ones
// (oneline.f:56-152).  For every voxel whose GPU weight lies in a narrow band
// around 0.1, these kernels recompute that exact sequence from the context's
// insertion history and record the reference's decision in a byte mask that
// normalize_kernel reads (two bits per volume of the context):
//   1. thresh_list_kernel  -- voxels in the band (two targets each: the voxel
//                             and, on the kx = 0 plane, its SYMPLANE0 partner);
//   2. hit_kernel          -- (target, section) pairs whose plane passes within
//                             reach of the target, kept in section order;
//   3. term_kernel         -- each pair's ONELINE weight terms on the target,
//                             in row / column / tap order (count, then write);
//   4. sum_kernel          -- each target's terms summed sequentially per half;
//   5. decide_kernel       -- SYMPLANE0 pair and halves combined as load_sym
//                             does, compared with 0.1 as NRMW2 does.
// With FP64 geometry the same steps follow the FP64 computation (FP64 table,
// positions and sums) instead.
#pragma once
#include "common.hpp"
#include "finalize.hpp"

namespace spbp {

constexpr int RES_TPB = 256;

// ------------------------------------------------------------------ scans (int64)
// Exclusive scan of n counts into out[0..n] (out[n] = total), three kernels.
__device__ inline long long block_exclusive_scan(long long v, long long* total) {
    __shared__ long long wsum[1024 / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    long long inc = v;
    for (int d = 1; d < 64; d <<= 1) {
        const long long u = __shfl_up(inc, d, 64);
        if (lane >= d) inc += u;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    long long base = 0, all = 0;
    for (int k = 0; k < nw; ++k) {
        if (k < w) base += wsum[k];
        all += wsum[k];
    }
    __syncthreads();
    *total = all;
    return base + inc - v;
}

template <typename C>
__global__ __launch_bounds__(1024) void scan_blocks_kernel(const C* __restrict__ in, long long n,
                                                           long long* __restrict__ out, long long* __restrict__ bsum) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long v = i < n ? (long long)in[i] : 0;
    long long tot;
    const long long ex = block_exclusive_scan(v, &tot);
    if (i < n) out[i] = ex;
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void scan_sums_kernel(long long* __restrict__ bsum, int nb) {
    long long carry = 0;
    for (int b0 = 0; b0 < nb; b0 += blockDim.x) {
        const int b = b0 + threadIdx.x;
        const long long v = b < nb ? bsum[b] : 0;
        long long tot;
        const long long ex = block_exclusive_scan(v, &tot);
        if (b < nb) bsum[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ __launch_bounds__(1024) void scan_add_kernel(long long* __restrict__ out, long long n,
                                                        const long long* __restrict__ bsum, int nb) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += bsum[blockIdx.x];
    if (i == 0) out[n] = bsum[nb];
}

// ------------------------------------------------------------------ targets
// Half-grid index -> signed frequency (kx, ky, kz); ky, kz in (-NX, NX].
__device__ inline void signed_coords(long long o, int nx, int N, int& x, int& y, int& z) {
    const int H = nx + 1;
    x = (int)(o % H);
    const long long r = o / H;
    const int j = (int)(r % N), k = (int)(r / N);
    y = j <= nx ? j : j - N;
    z = k <= nx ? k : k - N;
}

// Target t of item `it`: item 2s is listed voxel s, item 2s + 1 its SYMPLANE0
// partner on the kx = 0 plane (or none).
__device__ inline bool item_target(const long long* __restrict__ list, long long it, int nx, int N, int& x, int& y,
                                   int& z) {
    const long long o = list[it >> 1];
    signed_coords(o, nx, N, x, y, z);
    if ((it & 1) == 0) return true;
    const int H = nx + 1;
    const int j = (int)((o / H) % N), k = (int)(o / ((long long)H * N));
    if (!(x == 0 && sym_pair(j, k, nx))) return false;
    y = -y;
    z = -z;
    return true;
}

// Voxels whose weight (as the NRMW2 test reads it) lies within `band` of 0.1.
__global__ __launch_bounds__(RES_TPB) void thresh_list_kernel(Acc a, int nx, int N, double invH, double invN, double band,
                                                              long long* __restrict__ list, long long cap,
                                                              unsigned long long* __restrict__ count) {
    const int H = nx + 1;
    const size_t total = (size_t)H * N * N;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    int i, j, k;
    half_coords(idx, H, N, invH, invN, i, j, k);
    float2 x;
    float w;
    double wd;
    load_sym(a, nx, N, i, j, k, x, w, wd);
    const double v = a.Wc[0] ? wd : (double)w;
    if (fabs(v - (double)0.1f) <= band) {
        const unsigned long long s = atomicAdd(count, 1ull);
        if ((long long)s < cap) list[s] = (long long)idx;
    }
}

// ------------------------------------------------------------------ hits
// Can section g put weight on target t?  Every tap lies within 2.5 voxels per
// axis of its sample, and samples lie in the section's plane, so the plane
// must pass within 2.5 sqrt(3) of t (plus slack for the fp32 matrix).
__device__ inline bool plane_near(const Geom& g, int x, int y, int z) {
    const double d = (double)g.r[6] * x + (double)g.r[7] * y + (double)g.r[8] * z;
    return fabs(d) <= 4.45;
}

// Block (item, chunk of RES_TPB sections): MODE 0 counts the hits of the
// chunk into cnt[block]; MODE 1 writes them, in section order, at off[block].
// A launch covers blocks b0 .. b0 + gridDim.x - 1.
template <int MODE>
__global__ __launch_bounds__(RES_TPB) void hit_kernel(long long b0, const long long* __restrict__ list, long long nitems,
                                                      int nchunks, const Geom* __restrict__ hist, int nhist, int hmask,
                                                      int nx, int N, int* __restrict__ cnt,
                                                      const long long* __restrict__ off, int2* __restrict__ hits) {
    const long long b = b0 + blockIdx.x;
    const long long it = b / nchunks;
    const int e = (int)(b % nchunks) * RES_TPB + threadIdx.x;
    int x = 0, y = 0, z = 0;
    const bool valid = it < nitems && item_target(list, it, nx, N, x, y, z);
    bool hit = false;
    if (valid && e < nhist) {
        const Geom& g = hist[e];
        hit = ((hmask >> g.half) & 1) && plane_near(g, x, y, z);
    }
    if (MODE == 0) {
        const int c = __syncthreads_count(hit);
        if (threadIdx.x == 0) cnt[b] = c;
        return;
    }
    __shared__ int wcnt[RES_TPB / 64];
    const unsigned long long bal = __ballot(hit);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wcnt[w] = __popcll(bal);
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; ++k) base += wcnt[k];
    if (hit) {
        const int rank = base + __popcll(bal & ((1ull << lane) - 1ull));
        hits[off[b] + rank] = make_int2(e, (int)it);
    }
}

// ------------------------------------------------------------------ terms
__device__ inline float res_weight(const float* __restrict__ tab, float d, float fltb) { return kb_weight(tab, d, fltb); }
__device__ inline double res_weight(const double* __restrict__ tab, double d, double fltb) {
    return tab[(int)round(__dmul_rn(fabs(d), fltb))];
}
__device__ inline float res_mul(float a, float b) { return __fmul_rn(a, b); }
__device__ inline double res_mul(double a, double b) { return __dmul_rn(a, b); }

// WG = TABI(X) * (TABI(Y) * TABI(Z)) for the tap at voxel p of the sample at
// (x, y, z), in the precision P (oneline.f:80-100).
template <typename P>
__device__ inline P tap_weight(const P* __restrict__ tab, P fltb, P x, P y, P z, int px, int py, int pz) {
    const P tz = res_weight(tab, dsub(z, pz), fltb);
    const P ty = res_mul(res_weight(tab, dsub(y, py), fltb), tz);
    return res_mul(res_weight(tab, dsub(x, px), fltb), ty);
}

// ONELINE's weight terms of lattice sample (I, J) on target t, in its order:
// the direct tap first, then the reflected one (oneline.f:56-152).
template <typename P>
__device__ inline int sample_terms(const Geom& g, int I, int J, int nx, int N, const P* __restrict__ tab, P fltb,
                                   int tx, int ty, int tz, P out[2]) {
    if (!((I * I + J * J) < nx * nx) || (I == 0 && J < 0)) return 0;   // oneline.f:53-55
    {
        // cheap test first: a term needs the sample (at q or -q) within 2.5
        // voxels per axis of t, i.e. q within 2.5 of t or of -t
        const float fI = (float)I, fJ = (float)J;
        const float qx = fI * g.r[0] + fJ * g.r[3], qy = fI * g.r[1] + fJ * g.r[4], qz = fI * g.r[2] + fJ * g.r[5];
        const float dm = fmaxf(fabsf(qx - tx), fmaxf(fabsf(qy - ty), fabsf(qz - tz)));
        const float dp = fmaxf(fabsf(qx + tx), fmaxf(fabsf(qy + ty), fabsf(qz + tz)));
        if (fminf(dm, dp) > 2.6f) return 0;
    }
    P x, y, z;
    sample_pos(g, I, J, x, y, z);
    if (x < (P)0) { x = -x; y = -y; z = -z; }
    const int ix = nearest_voxel(x, N), iy = nearest_voxel(y, N), iz = nearest_voxel(z, N);
    if (!in_bounds(ix, iy, iz, nx)) return 0;
    int n = 0;
    // direct taps: x in [max(0, ix - 2), ix + 2]
    if (tx >= ix - LN2 && tx >= 0 && tx <= ix + LN2 && abs(ty - iy) <= LN2 && abs(tz - iz) <= LN2) {
        const P wg = tap_weight(tab, fltb, x, y, z, tx, ty, tz);
        if (wg != (P)0) out[n++] = wg;
    }
    // reflected taps p, x in [ix - 2, -1], written at -p
    if (ix < LN2) {
        const int px = -tx, py = -ty, pz = -tz;
        if (px >= ix - LN2 && px <= -1 && abs(py - iy) <= LN2 && abs(pz - iz) <= LN2) {
            const P wg = tap_weight(tab, fltb, x, y, z, px, py, pz);
            if (wg != (P)0) out[n++] = wg;
        }
    }
    return n;
}

// Lattice samples that can reach t lie within 2.5 sqrt(3) of t or of -t in
// the plane (in-plane coordinates (a, b) and (-a, -b)); visited rows J
// ascending and I ascending within a row, as ONELINE's callers visit them.
template <typename P, typename F>
__device__ inline void for_candidates(const Geom& g, int nx, int tx, int ty, int tz, F&& f) {
    const double a = (double)g.r[0] * tx + (double)g.r[1] * ty + (double)g.r[2] * tz;
    const double b = (double)g.r[3] * tx + (double)g.r[4] * ty + (double)g.r[5] * tz;
    constexpr double R2 = 19.8;   // (2.5 sqrt(3))^2 plus slack
    const double R = sqrt(R2);
    const double ca[2] = {a, -a}, cb[2] = {b, -b};
    int jlo[2], jhi[2];
    for (int w = 0; w < 2; ++w) {
        const bool any = ca[w] + R >= 0.0 && ca[w] - R <= (double)nx;   // I in [0, NX]
        jlo[w] = any ? max(-nx + 1, (int)ceil(cb[w] - R)) : 1;
        jhi[w] = any ? min(nx, (int)floor(cb[w] + R)) : 0;
    }
    // the rows to visit: the union of the two ranges, ascending
    int r0lo = jlo[0], r0hi = jhi[0], r1lo = jlo[1], r1hi = jhi[1];
    if (r1lo <= r1hi && (r0lo > r0hi || r1lo < r0lo)) {
        int t = r0lo; r0lo = r1lo; r1lo = t;
        t = r0hi; r0hi = r1hi; r1hi = t;
    }
    if (r0lo <= r0hi && r1lo <= r1hi && r1lo <= r0hi + 1) { r0hi = max(r0hi, r1hi); r1lo = 1; r1hi = 0; }
    for (int seg = 0; seg < 2; ++seg) {
        const int lo = seg ? r1lo : r0lo, hi = seg ? r1hi : r0hi;
        for (int J = lo; J <= hi; ++J) {
            int ilo[2], ihi[2];
            for (int w = 0; w < 2; ++w) {
                const double dj = (double)J - cb[w];
                const double h2 = R2 - dj * dj;
                if (J < jlo[w] || J > jhi[w] || h2 < 0.0) { ilo[w] = 1; ihi[w] = 0; continue; }
                const double h = sqrt(h2);
                ilo[w] = max(0, (int)ceil(ca[w] - h));
                ihi[w] = min(nx, (int)floor(ca[w] + h));
            }
            if (ilo[1] <= ihi[1] && (ilo[0] > ihi[0] || ilo[1] < ilo[0])) {
                int t = ilo[0]; ilo[0] = ilo[1]; ilo[1] = t;
                t = ihi[0]; ihi[0] = ihi[1]; ihi[1] = t;
            }
            if (ilo[0] <= ihi[0] && ilo[1] <= ihi[1] && ilo[1] <= ihi[0] + 1) { ihi[0] = max(ihi[0], ihi[1]); ilo[1] = 1; ihi[1] = 0; }
            for (int s = 0; s < 2; ++s)
                for (int I = ilo[s]; I <= ihi[s]; ++I) f(I, J);
        }
    }
}

// Thread per hit: WRITE = false counts its terms into tcnt, WRITE = true
// writes them in order at toff, negated for sections of half 1 (weights are
// positive, so the sign carries the half to the summation).
template <typename P, bool WRITE>
__global__ __launch_bounds__(RES_TPB) void term_kernel(const int2* __restrict__ hits, long long nhits,
                                                       const long long* __restrict__ list, const Geom* __restrict__ hist,
                                                       const P* __restrict__ tab, P fltb, int nx, int N,
                                                       int* __restrict__ tcnt, const long long* __restrict__ toff,
                                                       P* __restrict__ terms) {
    const long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nhits) return;
    const int2 hv = hits[h];
    int tx, ty, tz;
    item_target(list, hv.y, nx, N, tx, ty, tz);
    const Geom g = hist[hv.x];
    long long q = WRITE ? toff[h] : 0;
    int n = 0;
    const bool neg = g.half != 0;
    for_candidates<P>(g, nx, tx, ty, tz, [&](int I, int J) {
        P t[2];
        const int k = sample_terms<P>(g, I, J, nx, N, tab, fltb, tx, ty, tz, t);
        for (int u = 0; u < k; ++u) {
            if (WRITE) terms[q++] = neg ? -t[u] : t[u];
            ++n;
        }
    });
    if (!WRITE) tcnt[h] = n;
}

// ------------------------------------------------------------------ sums, decisions
// Thread per item: its terms in section order (contiguous, since hits and
// terms are both laid out item-major), one sequential sum per half.
template <typename P>
__global__ __launch_bounds__(RES_TPB) void sum_kernel(long long nitems, int nchunks, const long long* __restrict__ hoff,
                                                      const long long* __restrict__ toff, const P* __restrict__ terms,
                                                      P* __restrict__ sums) {
    const long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nitems) return;
    P a0 = (P)0, a1 = (P)0;
    long long q = toff[hoff[it * nchunks]];
    const long long q1 = toff[hoff[(it + 1) * nchunks]];
    // the loads of a group are independent; only the additions are serial
    for (; q + 8 <= q1; q += 8) {
        P v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = terms[q + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (v[k] > (P)0) a0 = a0 + v[k];
            else a1 = a1 + (-v[k]);
        }
    }
    for (; q < q1; ++q) {
        const P v = terms[q];
        if (v > (P)0) a0 = a0 + v;
        else a1 = a1 + (-v);
    }
    sums[2 * it] = a0;
    sums[2 * it + 1] = a1;
}

template <typename P>
__device__ inline P combined_weight(const long long* __restrict__ list, long long s, int hmask, int nx, int N,
                                    const P* __restrict__ sums) {
    int x, y, z;
    const bool pr = item_target(list, 2 * s + 1, nx, N, x, y, z);
    P w = (P)0;
    bool first = true;
    for (int h = 0; h < 2; ++h) {
        if (!((hmask >> h) & 1)) continue;
        P wh = sums[2 * (2 * s) + h];
        if (pr) wh = wh + sums[2 * (2 * s + 1) + h];
        w = first ? wh : w + wh;
        first = false;
    }
    return w;
}

// The combined weights themselves (spbp_reference_weights).
template <typename P>
__global__ __launch_bounds__(RES_TPB) void weights_out_kernel(const long long* __restrict__ list, long long nv, int hmask,
                                                              int nx, int N, const P* __restrict__ sums,
                                                              double* __restrict__ out) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nv) out[s] = (double)combined_weight(list, s, hmask, nx, N, sums);
}

template <typename P>
__global__ __launch_bounds__(RES_TPB) void decide_kernel(const long long* __restrict__ list, long long nv, Acc a,
                                                         int hmask, int nx, int N, const P* __restrict__ sums,
                                                         double tol, unsigned char* __restrict__ mask, int shift,
                                                         unsigned long long* __restrict__ ctr) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nv) return;
    const P w = combined_weight(list, s, hmask, nx, N, sums);
    const long long o = list[s];
    const int H = nx + 1;
    float2 xv;
    float wf;
    double wd;
    load_sym(a, nx, N, (int)(o % H), (int)((o / H) % N), (int)(o / ((long long)H * N)), xv, wf, wd);
    const double gv = a.Wc[0] ? wd : (double)wf;
    if (fabs((double)w - gv) > tol) {
        atomicAdd(&ctr[2], 1ull);
        return;
    }
    const bool divide = w > (P)0.1f;
    // listed voxels are distinct, and the volumes sharing the byte decide in
    // launches of their own
    mask[o] = (unsigned char)(mask[o] | ((divide ? 2 : 1) << shift));
    if ((gv > (double)0.1f) != divide) atomicAdd(&ctr[1], 1ull);
}

__global__ __launch_bounds__(RES_TPB) void mask_clear_kernel(const long long* __restrict__ list, long long nv,
                                                             unsigned char* __restrict__ mask) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nv) mask[list[s]] = 0;
}

}  // namespace spbp
