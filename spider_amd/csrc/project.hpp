// project.hpp -- central-slice KB gather: the exact adjoint of the ONELINE
// scatter (oneline.f:38-158), and the BI -> raw spectrum conversion that lets
// BI-layout spectra go through the insertion kernels.
#pragma once
#include "common.hpp"

namespace spbp {

// four floats at an 8-byte aligned address (global loads need only dword
// alignment on gfx950)
typedef float float4a8 __attribute__((ext_vector_type(4), aligned(8)));

// For lattice sample (I, J) of projection p, ONELINE adds w * v to X at each
// direct tap t (t_x >= 0) and w * conj(v) at -t for each reflected tap
// (t_x < 0), where v is BI(I, J), conjugated when the sample is Friedel-
// flipped (x < 0).  Viewing complex numbers as pairs of reals, the transpose
// of that map is
//     out(I, J) = c( sum_direct w X[t] + conj(sum_reflected w X[-t]) ),
// c = conj when flipped, identity otherwise; samples ONELINE skips give 0.
// The taps, the table lookups and the fp32 weights w are computed exactly as
// the insertion kernels compute them, and the sum is accumulated in FP64.
// One thread per output entry; out is nproj x N x H in the BI layout (row
// jp = J mod N, I fastest).  A workgroup covers a GT x GT tile of (I, jp), so
// its samples form a patch of the section and share most of their taps in
// the caches; blockIdx.x = tile (I tiles fastest), blockIdx.y = projection.
constexpr int GT = 16;
// P: sample-position precision, float (the reference's geometry) or double
// (contexts with FP64 geometry), as in the insertion kernels it transposes.
template <typename P>
__global__ __launch_bounds__(GT * GT) void gather_kernel(const float2* __restrict__ X, const Geom* __restrict__ geom,
                                                         int nx, int N, const float* __restrict__ tab, float fltb_f,
                                                         float2* __restrict__ out) {
    const int H = nx + 1;
    const int p = blockIdx.y;
    const int ti = (H + GT - 1) / GT;
    const int tj = blockIdx.x / ti;
    const int I = (blockIdx.x - tj * ti) * GT + (threadIdx.x & (GT - 1));
    const int jp = tj * GT + threadIdx.x / GT;
    if (I >= H || jp >= N) return;
    const int J = jp <= nx ? jp : jp - N;
    float2* o = out + ((size_t)p * N + jp) * H + I;
    if (!(I * I + J * J < (N * N) / 4) || (I == 0 && J < 0)) { *o = make_float2(0.f, 0.f); return; }
    const Geom g = geom[p];
    const P fltb = fltb_of<P>(fltb_f);
    P x, y, z;
    sample_pos(g, I, J, x, y, z);
    const bool flip = x < (P)0;
    if (flip) { x = -x; y = -y; z = -z; }
    const int ix = nearest_voxel(x, N), iy = nearest_voxel(y, N), iz = nearest_voxel(z, N);
    if (!in_bounds(ix, iy, iz, nx)) { *o = make_float2(0.f, 0.f); return; }
    const size_t plane = (size_t)N * H;
    if (ix >= LN2) {
        // no reflected taps (every tap has p_x >= ix - 2 >= 0); the nonzero
        // taps of each axis are a run of at most 4 (the table vanishes beyond
        // 1.8 voxels): from i - 2 when its weight is nonzero, else from i - 1.
        // Same nonzero terms in the same order as the general loop below.
        float tx[4], ty[4], tz[4];
        const float wx0 = kb_weight(tab, dsub(x, ix - 2), fltb);
        const float wy0 = kb_weight(tab, dsub(y, iy - 2), fltb);
        const float wz0 = kb_weight(tab, dsub(z, iz - 2), fltb);
        const int sx = wx0 != 0.0f ? ix - 2 : ix - 1;
        const int sy = wy0 != 0.0f ? iy - 2 : iy - 1;
        const int sz = wz0 != 0.0f ? iz - 2 : iz - 1;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            tx[t] = (t == 0 && sx == ix - 2) ? wx0 : kb_weight(tab, dsub(x, sx + t), fltb);
            ty[t] = (t == 0 && sy == iy - 2) ? wy0 : kb_weight(tab, dsub(y, sy + t), fltb);
            tz[t] = (t == 0 && sz == iz - 2) ? wz0 : kb_weight(tab, dsub(z, sz + t), fltb);
        }
        // each tap row is read whole (columns sx .. sx + 3 <= ix + 2 <= nx - 1,
        // inside the row) as two 16-byte loads; zero-weight taps add exact
        // zeros, so the sums equal those over the nonzero taps alone
        double dre = 0.0, dim = 0.0;
#pragma unroll
        for (int lz = 0; lz < 4; ++lz) {
            const int izp = sz + lz;
            const float2* Xz = X + (size_t)(izp >= 0 ? izp : N + izp) * plane;
#pragma unroll
            for (int ly = 0; ly < 4; ++ly) {
                const float tyz = __fmul_rn(ty[ly], tz[lz]);
                const int iyp = sy + ly;
                const float2* Xy = Xz + (size_t)(iyp >= 0 ? iyp : N + iyp) * H + sx;
                const float4a8 a = *(const float4a8*)Xy, b = *(const float4a8*)(Xy + 2);
                const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
                for (int lx = 0; lx < 4; ++lx) {
                    const double w = (double)__fmul_rn(tx[lx], tyz);
                    dre += w * (double)v[2 * lx];
                    dim += w * (double)v[2 * lx + 1];
                }
            }
        }
        // the general form adds the (empty) reflected sums: + 0 and - 0
        const double re = dre + 0.0, im = dim - 0.0;
        *o = make_float2((float)re, (float)(flip ? -im : im));
        return;
    }
    float tx[5], ty[5], tz[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) {
        tx[l] = kb_weight(tab, dsub(x, ix + l - LN2), fltb);
        ty[l] = kb_weight(tab, dsub(y, iy + l - LN2), fltb);
        tz[l] = kb_weight(tab, dsub(z, iz + l - LN2), fltb);
    }
    double dre = 0.0, dim = 0.0, mre = 0.0, mim = 0.0;
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
                const double w = (double)wg;
                if (px >= 0) {
                    const float2 v = X[za + ya + px];
                    dre += w * (double)v.x;
                    dim += w * (double)v.y;
                } else {
                    const float2 v = X[zt + yt + (size_t)(-px)];
                    mre += w * (double)v.x;
                    mim += w * (double)v.y;
                }
            }
        }
    }
    const double re = dre + mre, im = dim - mim;
    *o = make_float2((float)re, (float)(flip ? -im : im));
}

// The projector's image synthesis from gathered BI spectra, before the C2R
// transform: in place, C(I, jp) = conj(BI(I, jp) (-1)^(I + jp)), and column 0
// completed by Hermitian symmetry, C(0, N - jp) = conj(C(0, jp)) for
// 0 < jp < NX (the rows J < 0 of column 0 are not gathered).  One thread per
// entry; column 0's pairs by the thread of jp.
__global__ __launch_bounds__(TPB) void bi_complete_kernel(float2* __restrict__ spec, int N, int H, int nx, size_t n) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const size_t per = (size_t)N * H;
    const size_t r = idx % per;
    const int jp = (int)(r / H), I = (int)(r - (size_t)jp * H);
    if (I == 0 && jp > nx) return;   // written by the thread of N - jp
    const float sg = ((I + jp) & 1) ? -1.0f : 1.0f;
    const float2 v = spec[idx];
    const float2 c = make_float2(v.x * sg, -(v.y * sg));
    spec[idx] = c;
    if (I == 0 && jp > 0 && jp < nx) spec[(idx - r) + (size_t)(N - jp) * H] = make_float2(c.x, -c.y);
}

// Crop of the C2R outputs (N x N each, unnormalised) to the NX x NX images at
// offset IP, scaled by 1 / N^2 (the inverse transform's normalisation).
__global__ __launch_bounds__(TPB) void crop_images_kernel(const float* __restrict__ full, int N, int nx, int ip,
                                                          size_t n, float* __restrict__ out) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const size_t npix = (size_t)nx * nx;
    const size_t img = idx / npix, q = idx - img * npix;
    const int y = (int)(q / nx), x = (int)(q - (size_t)y * nx);
    const float inv = 1.0f / (float)((double)N * (double)N);
    out[idx] = full[(img * N + (size_t)(ip + y)) * N + (size_t)(ip + x)] * inv;
}

// Spectra for the projector's adjoint applied to real images r (the transpose
// of images = crop(irfft2(BI completed by Hermitian symmetry in column 0))):
// in place on the R2C spectra R of the zero-padded images, the raw values the
// insertion takes are R(I, jp) cx(I) / N^2 with cx = 2 for 0 < I < NX and 1 at
// I = 0 and NX (the columns irfft2 counts twice), and in column 0 the
// transpose of the completion: R(0, jp) + conj(R(0, N - jp)) for 0 < jp < NX,
// 0 for jp > NX.  One thread per entry; column 0's pairs (jp, N - jp) by the
// thread of jp.
__global__ __launch_bounds__(TPB) void adjoint_spectrum_kernel(float2* __restrict__ spec, int N, int H, int nx,
                                                               size_t n) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const size_t per = (size_t)N * H;
    const size_t r = idx % per;
    const int jp = (int)(r / H), I = (int)(r - (size_t)jp * H);
    const float inv = 1.0f / (float)((double)N * (double)N);   // a power of two: exact
    if (I > 0) {
        const float2 v = spec[idx];
        const float sc = I < nx ? 2.0f * inv : inv;
        spec[idx] = make_float2(v.x * sc, v.y * sc);
    } else if (jp == 0 || jp == nx) {
        const float2 v = spec[idx];
        spec[idx] = make_float2(v.x * inv, v.y * inv);
    } else if (jp < nx) {
        float2* m = spec + (idx - r) + (size_t)(N - jp) * H;
        const float2 a = spec[idx], b = *m;
        spec[idx] = make_float2(__fadd_rn(a.x * inv, b.x * inv), __fsub_rn(a.y * inv, b.y * inv));
        *m = make_float2(0.0f, 0.0f);
    }
}

// BI(I, J) -> the raw R2C value s that spectrum_value() maps to it:
// BI = c conj(s), c = (-1)^(I + jp), so s = c conj(BI).
__global__ __launch_bounds__(TPB) void bi_to_raw_kernel(const float2* __restrict__ bi, int N, int H, size_t n,
                                                        float2* __restrict__ raw) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const size_t r = idx % ((size_t)N * H);
    const int jp = (int)(r / H), I = (int)(r - (size_t)jp * H);
    const float sg = ((I + jp) & 1) ? -1.0f : 1.0f;
    const float2 v = bi[idx];
    raw[idx] = make_float2(v.x * sg, -v.y * sg);
}

}  // namespace spbp
