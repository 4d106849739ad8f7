// rotate.hpp -- RB 3F / RB 32F pre-rotation: RTSQ (rtsq.f:56-150) with QUADRI
// interpolation (quadri.f:83-150), and with FBS ON, RTSF (rtsf.f:106-205) with
// FBS2 interpolation (fbs2.f:95-245); REAL arithmetic in the reference's order.
#pragma once
#include "common.hpp"

namespace spbp {

// Per image: the angle's cosine and sine over the scale, and -(centre + shift).
struct RtsqParam {
    float coddscli, siddscli, fixcenmshx, fiycenmshy;
};

// QUADRI of image f (NX x NY, x fastest, 1-based coordinates) at (xx, yy),
// wrapping coordinates outside [1, N] as the reference does.
template <typename T>
__device__ inline float quadri(const T* __restrict__ f, int nx, int ny, float xx, float yy) {
    float x = xx, y = yy;
    if (x < 1.0f) x = __fadd_rn(x, (float)((1 - (int)x / nx) * nx));
    if (x > __fadd_rn((float)nx, 0.5f)) x = __fadd_rn(fmodf(__fsub_rn(x, 1.0f), (float)nx), 1.0f);
    if (y < 1.0f) y = __fadd_rn(y, (float)((1 - (int)y / ny) * ny));
    if (y > __fadd_rn((float)ny, 0.5f)) y = __fadd_rn(fmodf(__fsub_rn(y, 1.0f), (float)ny), 1.0f);
    const int i = (int)x, j = (int)y;
    const float dx0 = __fsub_rn(x, (float)i), dy0 = __fsub_rn(y, (float)j);
    int ip1 = i + 1, im1 = i - 1, jp1 = j + 1, jm1 = j - 1;
    if (ip1 > nx) ip1 -= nx;
    if (im1 < 1) im1 += nx;
    if (jp1 > ny) jp1 -= ny;
    if (jm1 < 1) jm1 += ny;
    auto F = [&](int a, int b) { return load_img(f, (size_t)(b - 1) * nx + (a - 1)); };
    const float f0 = F(i, j);
    const float c1 = __fsub_rn(F(ip1, j), f0);
    const float c2 = __fmul_rn(__fadd_rn(__fsub_rn(c1, f0), F(im1, j)), 0.5f);
    const float c3 = __fsub_rn(F(i, jp1), f0);
    const float c4 = __fmul_rn(__fadd_rn(__fsub_rn(c3, f0), F(i, jm1)), 0.5f);
    const float dxb = __fsub_rn(dx0, 1.0f), dyb = __fsub_rn(dy0, 1.0f);
    const float hxc = signbit(dx0) ? -1.0f : 1.0f;   // INT(SIGN(1.0, DX0)) as REAL
    const float hyc = signbit(dy0) ? -1.0f : 1.0f;
    int ic = (int)__fadd_rn((float)i, hxc), jc = (int)__fadd_rn((float)j, hyc);
    if (ic > nx) ic -= nx;
    else if (ic < 1) ic += nx;
    if (jc > ny) jc -= ny;
    else if (jc < 1) jc += ny;
    float c5 = __fsub_rn(F(ic, jc), f0);
    c5 = __fsub_rn(c5, __fmul_rn(hxc, c1));
    c5 = __fsub_rn(c5, __fmul_rn(__fmul_rn(hxc, __fsub_rn(hxc, 1.0f)), c2));
    c5 = __fsub_rn(c5, __fmul_rn(hyc, c3));
    c5 = __fsub_rn(c5, __fmul_rn(__fmul_rn(hyc, __fsub_rn(hyc, 1.0f)), c4));
    c5 = __fmul_rn(c5, __fmul_rn(hxc, hyc));
    const float tx = __fadd_rn(__fadd_rn(c1, __fmul_rn(dxb, c2)), __fmul_rn(dy0, c5));
    const float ty = __fadd_rn(c3, __fmul_rn(dyb, c4));
    return __fadd_rn(__fadd_rn(f0, __fmul_rn(dx0, tx)), __fmul_rn(dy0, ty));
}

// One thread per output pixel of b images (grid.y = image).
template <typename T>
__global__ __launch_bounds__(TPB) void rtsq_kernel(const T* __restrict__ imgs, const RtsqParam* __restrict__ prm,
                                                   int nx, float* __restrict__ out) {
    const int p = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * nx) return;
    const int iy = idx / nx + 1, ix = idx % nx + 1;
    const RtsqParam q = prm[p];
    const int cen = nx / 2 + 1;
    const float r1 = (float)(-(nx / 2));
    const float r2 = (nx % 2 == 0) ? __fsub_rn((float)(nx / 2), 1.0f) : (float)(nx / 2);
    const float re1 = (nx % 2 == 0) ? (float)(-nx) : (float)(-nx - 1);
    const float re2 = (nx % 2 == 0) ? (float)nx : (float)(nx + 1);
    float yi = __fadd_rn((float)iy, q.fiycenmshy);
    if (yi < r1) yi = fminf(__fadd_rn(yi, re2), r2);
    if (yi > r2) yi = fmaxf(__fadd_rn(yi, re1), r1);
    const float ycod = __fadd_rn(__fmul_rn(yi, q.coddscli), (float)cen);
    const float ysid = __fadd_rn(__fmul_rn(-yi, q.siddscli), (float)cen);
    float xi = __fadd_rn((float)ix, q.fixcenmshx);
    if (xi < r1) xi = fminf(__fadd_rn(xi, re2), r2);
    if (xi > r2) xi = fmaxf(__fadd_rn(xi, re1), r1);
    const float yold = __fadd_rn(__fmul_rn(xi, q.siddscli), ycod);
    const float xold = __fadd_rn(__fmul_rn(xi, q.coddscli), ysid);
    out[(size_t)p * nx * nx + idx] = quadri(imgs + (size_t)p * nx * nx, nx, nx, xold, yold);
}

// ------------------------------------------------------------------ RTSF (FBS ON)
// Per image: cosine and sine of the angle, the shift taken MODULO the size,
// and the output-coordinate offsets of rtsf.f:147-159.
struct RtsfParam {
    float cth, sth, shx, shy, shxpnx, shypny;
    float fx0, fx1, fx2, fy0, fy1, fy2;
};

// The three derivative spectra of FBS2_PREP (fbs2.f:200-235) from the FFTW-sign
// R2C spectrum S of an NX x NX image, as inputs of a C2R: i wx S, i wy S and
// -wx wy S over NX^2 (the reference's conjugations before and after cancel;
// its 1/NX^2 after the inverse is folded in).  wx = I (2 pi / NX) for the
// stored I = 0 .. NX/2, wy = J (2 pi / NX) with J - NX above NX/2, factors in
// REAL.  The I = 0 and I = NX/2 columns are replaced by their Hermitian parts,
// which is how FFTW's C2R reads them.  Thread per (image, J, I); out holds the
// three spectra of each image one after another.
__global__ __launch_bounds__(TPB) void fbs_deriv_kernel(const float2* __restrict__ spec, int nx, int nimg,
                                                        float2* __restrict__ out) {
    const int H = nx / 2 + 1;
    const size_t per = (size_t)nx * H;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= per * nimg) return;
    const int p = (int)(idx / per);
    const int r = (int)(idx - (size_t)p * per);
    const int J = r / H, I = r - J * H;
    const float2* sp = spec + (size_t)p * per;
    const float a4 = __fdiv_rn((float)6.28318530717958647692, (float)nx);
    const float wx = __fmul_rn((float)I, a4);
    const float wy = __fmul_rn((float)(J <= nx / 2 ? J : J - nx), a4);
    const float pix = (float)(1.0 / ((double)nx * nx));
    float2 s = sp[r];
    float2 d[3];
    auto mult = [&](float2 v, int k) {
        if (k == 0) return make_float2(-v.y * wx, v.x * wx);
        if (k == 1) return make_float2(-v.y * wy, v.x * wy);
        const float m = -(wx * wy);
        return make_float2(v.x * m, v.y * m);
    };
    if (I == 0 || 2 * I == nx) {
        // Hermitian part of the column: (G(J) + conj G(-J)) / 2, with the
        // multiplier of -J (wy changes sign, except at J = NX/2)
        const int Jm = (nx - J) % nx;
        const float2 sm = sp[(size_t)Jm * H + I];
        const float wym = __fmul_rn((float)(Jm <= nx / 2 ? Jm : Jm - nx), a4);
        for (int k = 0; k < 3; ++k) {
            const float2 g = mult(s, k);
            float2 gm;
            if (k == 0) gm = make_float2(-sm.y * wx, sm.x * wx);
            else if (k == 1) gm = make_float2(-sm.y * wym, sm.x * wym);
            else { const float m = -(wx * wym); gm = make_float2(sm.x * m, sm.y * m); }
            d[k] = make_float2(0.5f * (g.x + gm.x), 0.5f * (g.y - gm.y));
        }
    } else {
        for (int k = 0; k < 3; ++k) d[k] = mult(s, k);
    }
    float2* o = out + (size_t)p * 3 * per;
    for (int k = 0; k < 3; ++k) o[(size_t)k * per + r] = make_float2(d[k].x * pix, d[k].y * pix);
}

// FBS2 (fbs2.f:95-160) with CHKBOUND: cubic Hermite interpolation along x on
// the rows J and J + 1 around (X, Y), then along y, from the image and its
// derivative images dx, dy, dxy (NX x NX, x fastest).
template <typename T>
__device__ inline float fbs2(const T* __restrict__ f, const float* __restrict__ dx_, const float* __restrict__ dy_,
                             const float* __restrict__ dxy, int nx, float x, float y) {
    const int i = (int)floorf(x), j = (int)floorf(y);
    const float dx = __fsub_rn(x, (float)i), dy = __fsub_rn(y, (float)j);
    const float dxsqr = __fmul_rn(dx, dx), dxcub = __fmul_rn(__fmul_rn(dx, dx), dx);
    auto wrap = [&](int a) { const int r = a % nx; return r < 0 ? r + nx : r; };
    const int i2 = wrap(i - 1), j2 = wrap(j - 1), i3 = wrap(i), j3 = wrap(j);   // 0-based
    auto at = [&](int a, int b) { return (size_t)b * nx + a; };
    auto hermite = [&](float a0, float a1, float f3, float d3) {
        const float a2 = __fsub_rn(__fsub_rn(__fmul_rn(3.0f, __fsub_rn(f3, a0)), __fmul_rn(2.0f, a1)), d3);
        const float a3 = __fadd_rn(__fadd_rn(__fmul_rn(2.0f, __fadd_rn(-f3, a0)), a1), d3);
        return __fadd_rn(__fadd_rn(__fadd_rn(a0, __fmul_rn(a1, dx)), __fmul_rn(a2, dxsqr)), __fmul_rn(a3, dxcub));
    };
    const float adx = hermite(load_img(f, at(i2, j2)), dx_[at(i2, j2)], load_img(f, at(i3, j2)), dx_[at(i3, j2)]);
    const float bdx = hermite(load_img(f, at(i2, j3)), dx_[at(i2, j3)], load_img(f, at(i3, j3)), dx_[at(i3, j3)]);
    const float dadx = hermite(dy_[at(i2, j2)], dxy[at(i2, j2)], dy_[at(i3, j2)], dxy[at(i3, j2)]);
    const float dbdx = hermite(dy_[at(i2, j3)], dxy[at(i2, j3)], dy_[at(i3, j3)], dxy[at(i3, j3)]);
    const float a2 = __fsub_rn(__fsub_rn(__fmul_rn(3.0f, __fsub_rn(bdx, adx)), __fmul_rn(2.0f, dadx)), dbdx);
    const float a3 = __fadd_rn(__fadd_rn(__fmul_rn(2.0f, __fsub_rn(adx, bdx)), dadx), dbdx);
    const float dy2 = __fmul_rn(dy, dy), dy3 = __fmul_rn(__fmul_rn(dy, dy), dy);
    return __fadd_rn(__fadd_rn(__fadd_rn(adx, __fmul_rn(dadx, dy)), __fmul_rn(a2, dy2)), __fmul_rn(a3, dy3));
}

// RTSF (rtsf.f:142-175, scale 1): one thread per output pixel of b images
// (grid.y = image); der holds each image's dF/dx, dF/dy, d2F/dxdy.
template <typename T>
__global__ __launch_bounds__(TPB) void rtsf_kernel(const T* __restrict__ imgs, const float* __restrict__ der,
                                                   const RtsfParam* __restrict__ prm, int nx, float* __restrict__ out) {
    const int p = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * nx) return;
    const int iy = idx / nx + 1, ix = idx % nx + 1;
    const RtsfParam q = prm[p];
    const int c = nx / 2 + 1;
    float fy = __fadd_rn((float)iy, q.fy0);
    if ((float)(iy - 1) < q.shy) fy = __fadd_rn((float)iy, q.fy1);
    if ((float)(iy - 1) >= q.shypny) fy = __fadd_rn((float)iy, q.fy2);
    const float ycod = __fadd_rn(__fmul_rn(q.cth, fy), (float)c);
    const float ysid = __fadd_rn(-__fmul_rn(q.sth, fy), (float)c);
    float fx = __fadd_rn((float)ix, q.fx0);
    if ((float)(ix - 1) < q.shx) fx = __fadd_rn((float)ix, q.fx1);
    if ((float)(ix - 1) >= q.shxpnx) fx = __fadd_rn((float)ix, q.fx2);
    const float xold = __fadd_rn(__fmul_rn(q.cth, fx), ysid);
    const float yold = __fadd_rn(__fmul_rn(q.sth, fx), ycod);
    const size_t npix = (size_t)nx * nx;
    const float* d = der + (size_t)p * 3 * npix;
    out[(size_t)p * npix + idx] = fbs2(imgs + (size_t)p * npix, d, d + npix, d + 2 * npix, nx, xold, yold);
}

}  // namespace spbp
