// rotate.hpp -- RB 3F / RB 32F pre-rotation: RTSQ (rtsq.f:56-150) with QUADRI
// interpolation (quadri.f:83-150), REAL arithmetic in the reference's order.
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

}  // namespace spbp
