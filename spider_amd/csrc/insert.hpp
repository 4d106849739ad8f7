// insert.hpp -- PADD2 and the ONELINE scatter kernels.
#pragma once
#include <type_traits>

#include "common.hpp"

namespace spbp {

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

// ------------------------------------------------------------------ strategy 1: direct scatter
// One thread per sample point (I, J) of one projection, FP32 atomics straight
// into the half grid (the reference's ONELINE loop, oneline.f:53-156).
template <typename P>   // position precision: float (the reference's) or double (SPBP_GEOM_F64)
__global__ __launch_bounds__(TPB) void insert_direct_kernel(
    const float2* __restrict__ spec, int img0, const Geom* __restrict__ geom, int nx, int N,
    const float* __restrict__ tab, float fltb_f,
    float2* __restrict__ X0, float* __restrict__ W0, float2* __restrict__ X1, float* __restrict__ W1,
    float* __restrict__ Wc0, float* __restrict__ Wc1) {
    // W0/W1 may be null (no weights wanted: scatter of BI spectra)
    const int H = nx + 1;
    const int p = blockIdx.y;
    const int lin = blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= H * N) return;
    const int jrow = lin / H;
    const int I = lin - jrow * H;
    const int J = jrow - nx + 1;                                   // J = -NX+1 .. NX
    if (!(I * I + J * J < (N * N) / 4) || (I == 0 && J < 0)) return;   // oneline.f:53-55
    const Geom g = geom[p];
    const P fltb = fltb_of<P>(fltb_f);
    P x, y, z;
    sample_pos(g, I, J, x, y, z);
    float2 v = spectrum_value(spec + (size_t)(g.img - img0) * N * H, I, J, N, H);
    if (x < (P)0) { x = -x; y = -y; z = -z; v.y = -v.y; }          // oneline.f:60-67
    const int ix = nearest_voxel(x, N), iy = nearest_voxel(y, N), iz = nearest_voxel(z, N);
    if (!in_bounds(ix, iy, iz, nx)) return;
    float tx[5], ty[5], tz[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) {
        tx[l] = kb_weight(tab, dsub(x, ix + l - LN2), fltb);
        ty[l] = kb_weight(tab, dsub(y, iy + l - LN2), fltb);
        tz[l] = kb_weight(tab, dsub(z, iz + l - LN2), fltb);
    }
    float2* __restrict__ X = g.half ? X1 : X0;
    float* __restrict__ W = g.half ? W1 : W0;
    float* __restrict__ Wc = g.half ? Wc1 : Wc0;   // null unless compensated weights
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
                    unsafeAtomicAdd(&X[o].x, __fmul_rn(v.x, wg));
                    unsafeAtomicAdd(&X[o].y, __fmul_rn(v.y, wg));
                    if (Wc) twosum_atomic_add(W, Wc, o, wg, 0.0f);
                    else if (W) unsafeAtomicAdd(&W[o], wg);
                } else {                                         // reflected taps, conj (oneline.f:118-152)
                    const size_t o = zt + yt + (size_t)(-px);
                    unsafeAtomicAdd(&X[o].x, __fmul_rn(v.x, wg));
                    unsafeAtomicAdd(&X[o].y, __fmul_rn(-v.y, wg));
                    if (Wc) twosum_atomic_add(W, Wc, o, wg, 0.0f);
                    else if (W) unsafeAtomicAdd(&W[o], wg);
                }
            }
        }
    }
}

// ------------------------------------------------------------------ strategy 2: brick-binned
// The half grid (signed kx in [0, NX), ky, kz in [-NX+1, NX)) is cut into
// bricks of BX x BY x BZ voxels.  A work item is (brick, half, up to C
// projections whose central plane crosses the brick).  Its workgroup
// accumulates, in LDS, every tap of those projections that lands in the
// brick, then adds the brick to the half grid once.  A sample (I, J) with
// effective position E (after the Friedel flip) writes a tap p of its 5x5x5
// window at p when p_x >= 0 and, conjugated, at -p when p_x < 0
// (oneline.f:77-152); its taps with nonzero weight satisfy |E - p| <= 1.8
// per axis.  So the samples feeding brick B are those with E within 1.8 of
// B (all their taps are checked against B) plus, for bricks at kx < 3,
// those with E within 1.8 of -B (mirrored taps only).

constexpr float BRICK_MARGIN = 1.86f;   // > 1.8: support of the KB table in voxels

struct BrickGrid {
    int bx, by, bz;      // brick edge (voxels); the core's with owner-computes bricks
    int nbx, nby, nbz;   // bricks per axis
    int nb;              // nbx * nby * nbz
    int own;             // owner-computes bricks (LDS region = core + 2 per side)
};

__device__ inline void brick_origin(const BrickGrid& bg, int b, int nx, int& x0, int& y0, int& z0) {
    const int i = b % bg.nbx, j = (b / bg.nbx) % bg.nby, k = b / (bg.nbx * bg.nby);
    x0 = i * bg.bx;
    y0 = -nx + 1 + j * bg.by;
    z0 = -nx + 1 + k * bg.bz;
}

// The list builders read, per section, only its plane's normal DM(7:9) and its
// half: packed into 16 bytes (x, y, z, half) so that the (brick, section)
// tests read a sixth of the section records' bytes.
__global__ __launch_bounds__(TPB) void plane_pack_kernel(const Geom* __restrict__ geom, int n, float4* __restrict__ pln) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) pln[p] = make_float4(geom[p].r[6], geom[p].r[7], geom[p].r[8], geom[p].half ? 1.0f : 0.0f);
}

// A section's plane (through the origin, normal n = DM(7:9)) feeds a brick
// when it passes through the brick's expanded box and that box reaches into
// the sample disk of radius NX.  Owner-computes bricks expand their core by
// 0.52 (samples whose nearest voxel lies in the core), and at kx = 0 by 2.52
// (the mirrored samples within 2.5 of -core: a plane through the origin
// crosses -B exactly when it crosses B); shared-evaluation bricks by the
// table's support.
//
// The list builders take bricks in groups of 2 x 2 x 2 ("superbricks"): a
// section is tested against the group's box first, and against its member
// bricks only where that box is hit.  The group's box holds every member's
// expanded box (with the largest member margin, plus 0.05 for rounding), so
// the first test never drops a pair the brick test keeps.  The members'
// boxes, and the disk condition, which does not depend on the section, are
// set up once per block.  (Host-callable as well, so that they can be
// checked on the CPU.)
struct SuperBrick {
    int b[8];       // member bricks, -1 past the grid's edge or wholly outside the sample disk
    float mc[8][3]; // member expanded boxes: centres and half extents
    float mh[8][3];
    float cx, cy, cz, hx, hy, hz;
};

__host__ __device__ __forceinline__ SuperBrick superbrick(const BrickGrid& bg, int s, int nx) {
    const int nsx = (bg.nbx + 1) / 2, nsy = (bg.nby + 1) / 2;
    const int si = s % nsx, sj = (s / nsx) % nsy, sk = s / (nsx * nsy);
    SuperBrick g;
    const float r = (float)nx + 0.02f;
#pragma unroll
    for (int d = 0; d < 8; ++d) {   // unrolled: the arrays stay in registers
        const int i = 2 * si + (d & 1), j = 2 * sj + ((d >> 1) & 1), k = 2 * sk + (d >> 2);
        g.b[d] = (i < bg.nbx && j < bg.nby && k < bg.nbz) ? i + bg.nbx * (j + bg.nby * k) : -1;
        // the member's expanded box and its disk condition
        const int x0 = i * bg.bx, y0 = -nx + 1 + j * bg.by, z0 = -nx + 1 + k * bg.bz;
        const float m = bg.own ? (x0 == 0 ? 2.52f : 0.52f) : BRICK_MARGIN;
        const float hx = 0.5f * (bg.bx - 1) + m, hy = 0.5f * (bg.by - 1) + m, hz = 0.5f * (bg.bz - 1) + m;
        const float cx = x0 + 0.5f * (bg.bx - 1), cy = y0 + 0.5f * (bg.by - 1), cz = z0 + 0.5f * (bg.bz - 1);
        const float qx = fmaxf(fabsf(cx) - hx, 0.f), qy = fmaxf(fabsf(cy) - hy, 0.f), qz = fmaxf(fabsf(cz) - hz, 0.f);
        if (!(qx * qx + qy * qy + qz * qz <= r * r)) g.b[d] = -1;
        g.mc[d][0] = cx;
        g.mc[d][1] = cy;
        g.mc[d][2] = cz;
        g.mh[d][0] = hx;
        g.mh[d][1] = hy;
        g.mh[d][2] = hz;
    }
    const int i1 = min(2 * si + 1, bg.nbx - 1), j1 = min(2 * sj + 1, bg.nby - 1), k1 = min(2 * sk + 1, bg.nbz - 1);
    const float m = (bg.own ? (si == 0 ? 2.52f : 0.52f) : BRICK_MARGIN) + 0.05f;
    const float xlo = 2 * si * bg.bx - m, xhi = i1 * bg.bx + bg.bx - 1 + m;
    const float ylo = -nx + 1 + 2 * sj * bg.by - m, yhi = -nx + 1 + j1 * bg.by + bg.by - 1 + m;
    const float zlo = -nx + 1 + 2 * sk * bg.bz - m, zhi = -nx + 1 + k1 * bg.bz + bg.bz - 1 + m;
    g.cx = 0.5f * (xlo + xhi);
    g.cy = 0.5f * (ylo + yhi);
    g.cz = 0.5f * (zlo + zhi);
    g.hx = 0.5f * (xhi - xlo);
    g.hy = 0.5f * (yhi - ylo);
    g.hz = 0.5f * (zhi - zlo);
    return g;
}

// Member d of superbrick s for a thread-dependent d, recomputed (indexing
// the record's array with it would move the whole record to scratch
// memory); -1 past the grid's edge.
__host__ __device__ __forceinline__ int member(const BrickGrid& bg, int s, int d) {
    const int nsx = (bg.nbx + 1) / 2, nsy = (bg.nby + 1) / 2;
    const int i = 2 * (s % nsx) + (d & 1), j = 2 * ((s / nsx) % nsy) + ((d >> 1) & 1), k = 2 * (s / (nsx * nsy)) + (d >> 2);
    return (i < bg.nbx && j < bg.nby && k < bg.nbz) ? i + bg.nbx * (j + bg.nby * k) : -1;
}

// Does the plane pass through member d's expanded box (its disk condition
// already applied)?
__host__ __device__ __forceinline__ bool member_hit(const SuperBrick& sb, int d, const float4& g) {
    const float dd = fabsf(g.x * sb.mc[d][0] + g.y * sb.mc[d][1] + g.z * sb.mc[d][2]);
    const float reach = sb.mh[d][0] * fabsf(g.x) + sb.mh[d][1] * fabsf(g.y) + sb.mh[d][2] * fabsf(g.z) + 0.02f;
    return dd <= reach;
}

// The group test: both conditions on the group's box.
__host__ __device__ __forceinline__ bool superbrick_hit(const SuperBrick& sb, int nx, const float4& g) {
    const float d = fabsf(g.x * sb.cx + g.y * sb.cy + g.z * sb.cz);
    const float reach = sb.hx * fabsf(g.x) + sb.hy * fabsf(g.y) + sb.hz * fabsf(g.z) + 0.05f;
    if (d > reach) return false;
    const float qx = fmaxf(fabsf(sb.cx) - sb.hx, 0.f), qy = fmaxf(fabsf(sb.cy) - sb.hy, 0.f),
                qz = fmaxf(fabsf(sb.cz) - sb.hz, 0.f);
    const float r = (float)nx + 0.1f;
    return qx * qx + qy * qy + qz * qz <= r * r;
}

// Count (brick, half) hits.  Block: one superbrick x 1024 projections.
__global__ __launch_bounds__(TPB) void brick_count_kernel(const float4* __restrict__ pln, int nproj, BrickGrid bg, int nx,
                                                          int* __restrict__ cnt) {
    __shared__ int sc[TPB / 64][16];
    const SuperBrick sb = superbrick(bg, blockIdx.x, nx);
    const int p0 = blockIdx.y * 1024;
    unsigned c[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) c[k] = 0;
    for (int p = p0 + threadIdx.x; p < min(nproj, p0 + 1024); p += TPB) {
        const float4 g = pln[p];
        if (!superbrick_hit(sb, nx, g)) continue;
        const bool h1 = g.w != 0.0f;
#pragma unroll
        for (int d = 0; d < 8; ++d)
            if (sb.b[d] >= 0 && member_hit(sb, d, g)) {
                if (h1) ++c[8 + d];
                else ++c[d];
            }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const unsigned v = wave_sum(c[k]);
        if (lane == 0) sc[w][k] = (int)v;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        const int d = threadIdx.x & 7, h = threadIdx.x >> 3;
        int t = 0;
        for (int k = 0; k < TPB / 64; ++k) t += sc[k][threadIdx.x];
        const int b = member(bg, blockIdx.x, d);   // a member outside the disk counted nothing
        if (t && b >= 0) atomicAdd(&cnt[b + h * bg.nb], t);
    }
}

// Exclusive scans of the counts and of the work items per key, zeroed
// cursors, queue head zeroed: three kernels (block per 1024 keys; one block
// over the block totals; offsets added).  off[nkeys] and item_off[nkeys] are
// the totals.
__device__ inline int2 block_scan_pair(int2 v, int2* total) {
    __shared__ int2 wsum[1024 / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int2 inc = v;
    for (int d = 1; d < 64; d <<= 1) {
        const int ux = __shfl_up(inc.x, d, 64), uy = __shfl_up(inc.y, d, 64);
        if (lane >= d) { inc.x += ux; inc.y += uy; }
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int2 base = make_int2(0, 0), all = make_int2(0, 0);
    for (int k = 0; k < nw; ++k) {
        if (k < w) { base.x += wsum[k].x; base.y += wsum[k].y; }
        all.x += wsum[k].x;
        all.y += wsum[k].y;
    }
    __syncthreads();
    *total = all;
    return make_int2(base.x + inc.x - v.x, base.y + inc.y - v.y);
}

__global__ __launch_bounds__(1024) void key_scan_blocks_kernel(const int* __restrict__ cnt, int nkeys, int C,
                                                               int* __restrict__ off, int* __restrict__ item_off,
                                                               int2* __restrict__ bsum) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = k < nkeys ? cnt[k] : 0;
    int2 tot;
    const int2 ex = block_scan_pair(make_int2(c, (c + C - 1) / C), &tot);
    if (k < nkeys) { off[k] = ex.x; item_off[k] = ex.y; }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void key_scan_sums_kernel(int2* __restrict__ bsum, int nb, int nkeys,
                                                             int* __restrict__ off, int* __restrict__ item_off,
                                                             int* __restrict__ next_item) {
    int2 carry = make_int2(0, 0);
    for (int b0 = 0; b0 < nb; b0 += blockDim.x) {
        const int b = b0 + threadIdx.x;
        int2 tot;
        const int2 ex = block_scan_pair(b < nb ? bsum[b] : make_int2(0, 0), &tot);
        if (b < nb) bsum[b] = make_int2(carry.x + ex.x, carry.y + ex.y);
        carry.x += tot.x;
        carry.y += tot.y;
    }
    if (threadIdx.x == 0) { off[nkeys] = carry.x; item_off[nkeys] = carry.y; *next_item = 0; }
}

__global__ __launch_bounds__(1024) void key_scan_add_kernel(int nkeys, const int2* __restrict__ bsum,
                                                            int* __restrict__ off, int* __restrict__ item_off,
                                                            int* __restrict__ cursor) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    const int2 b = bsum[blockIdx.x];
    off[k] += b.x;
    item_off[k] += b.y;
    cursor[k] = 0;
}

// Write the projection lists of every (brick, half) key.  Block: one
// superbrick x 1024 projections; each key's hits of a round of TPB
// projections take one cursor reservation per block (wave ballots give the
// ranks).
__global__ __launch_bounds__(TPB) void brick_fill_kernel(const float4* __restrict__ pln, int nproj, BrickGrid bg, int nx,
                                                         const int* __restrict__ off, int* __restrict__ cursor,
                                                         int* __restrict__ list) {
    __shared__ int wc[16][TPB / 64];
    __shared__ int base[16];
    const SuperBrick sb = superbrick(bg, blockIdx.x, nx);
    const int p0 = blockIdx.y * 1024;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int r = p0; r < min(nproj, p0 + 1024); r += TPB) {
        const int p = r + threadIdx.x;
        unsigned hits = 0;   // bit d: member d hit; this section's half is h
        int h = 0;
        if (p < nproj) {
            const float4 g = pln[p];
            h = g.w != 0.0f ? 1 : 0;
            if (superbrick_hit(sb, nx, g)) {
#pragma unroll
                for (int d = 0; d < 8; ++d)
                    if (sb.b[d] >= 0 && member_hit(sb, d, g)) hits |= 1u << d;
            }
        }
        unsigned long long m[8];
        const bool any = __ballot(hits != 0) != 0ull;
        if (any) {
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                const unsigned long long m0 = __ballot((hits >> d & 1u) && h == 0), m1 = __ballot((hits >> d & 1u) && h == 1);
                m[d] = h ? m1 : m0;
                if (lane == 0) { wc[d][w] = __popcll(m0); wc[8 + d][w] = __popcll(m1); }
            }
        } else if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 16; ++k) wc[k][w] = 0;
        }
        __syncthreads();
        if (threadIdx.x < 16) {
            const int d = threadIdx.x & 7, hh = threadIdx.x >> 3;
            int t = 0;
            for (int k = 0; k < TPB / 64; ++k) t += wc[threadIdx.x][k];
            const int b = member(bg, blockIdx.x, d);
            base[threadIdx.x] = (t && b >= 0) ? atomicAdd(&cursor[b + hh * bg.nb], t) : 0;
        }
        __syncthreads();
        if (hits) {
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                if (!(hits >> d & 1u)) continue;
                const int key = sb.b[d] + h * bg.nb;
                int rank = __popcll(m[d] & below);
                for (int k = 0; k < w; ++k) rank += wc[h * 8 + d][k];
                list[off[key] + base[h * 8 + d] + rank] = p;
            }
        }
        __syncthreads();
    }
}

// item -> key map: item_key[i] = key for i in [item_off[k], item_off[k+1]).
__global__ __launch_bounds__(TPB) void brick_itemkey_kernel(const int* __restrict__ item_off, int nkeys,
                                                            int* __restrict__ item_key) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    for (int i = item_off[k]; i < item_off[k + 1]; ++i) item_key[i] = k;
}

// x rotation of the LDS brick rows (A y + B z); 0, 0 selects the padded
// layout.  Measured at 128^2 into 256^3: (5, 3) 12.29 ms per 4000-projection
// launch, (3, 5) 12.34, (7, 11) 12.31, (5, 9) 12.31, padded 12.57.
#ifndef SPBP_LDS_ROT_A
#define SPBP_LDS_ROT_A 5
#endif
#ifndef SPBP_LDS_ROT_B
#define SPBP_LDS_ROT_B 3
#endif

template <int BX, int BY, int BZ, int NT = TPB, bool WGT = true>
struct BrickLds {
    static_assert(BX <= 31 && BY <= 31 && BZ <= 31, "brick edge too long for the 64-lane line table");
    // Layout of the brick's voxels.  Padded: rows of BX + 1, layers of
    // BY (BX + 1) + 1.  Rotated: rows of exactly BX, x rotated by
    // (RA y + RB z) mod BX within its row, so that voxels one step apart in
    // y or z fall in other banks than a step in x.
    static constexpr int RA = SPBP_LDS_ROT_A, RB = SPBP_LDS_ROT_B;
    static constexpr bool ROT = (RA != 0 || RB != 0) && (BX & (BX - 1)) == 0;
    static constexpr int SY = ROT ? BX : BX + 1;
    static constexpr int SZ = ROT ? BY * BX : BY * SY + 1;
    static constexpr int NS = BZ * SZ;
    __device__ static int rowbase(int y, int z) { return z * SZ + y * SY; }
    __device__ static int rot(int y, int z) { return ROT ? RA * y + RB * z : 0; }
    __device__ static int at(int base, int r, int x) { return ROT ? base + ((x + r) & (BX - 1)) : base + x; }
    // the same in bytes (8-byte elements): the tap loop's addresses need no scaling
    __device__ static int rowbase8(int y, int z) { return (z * SZ + y * SY) * 8; }
    __device__ static int rot8(int y, int z) { return ROT ? (RA * y + RB * z) * 8 : 0; }
    // rotated rows start at multiples of BX * 8 bytes, so the row offset is
    // OR-ed onto the rotated x (one v_and_or_b32)
    __device__ static int at8(int base8, int r8, int x8) { return ROT ? base8 | ((x8 + r8) & ((BX - 1) * 8)) : base8 + x8; }
    // brick-local coordinates of the v-th voxel (x fastest) and its element
    __device__ static int coords(int v, int& x, int& y, int& z) {
        x = v % BX;
        y = (v / BX) % BY;
        z = v / (BX * BY);
        return at(rowbase(y, z), rot(y, z), x);
    }
    // gfx950's LDS runs ds_add_u64 at 0.14 cycles per lane, ds_add_f64 at
    // 0.26 and ds_add_f32 at ~3 (measured).  X partial sums are exact 64-bit
    // fixed point: each per-tap term is the exact product of the fp32
    // spectrum value and the fp32 weight (two 24-bit mantissas fit FP64),
    // scaled by a power of two and rounded to an integer unit of 2^-VBITS of
    // the work item's largest spectrum magnitude; the brick sum is exact and
    // only that rounding and the final conversion round.  Weights are exact
    // 64-bit fixed point too, in units of 2^-WBITS: every weight term is a
    // product of three table values, at most 1, rounded to 2^-(WBITS+1)
    // absolute (relative 10^-14 where NRMW2 divides, W > 0.1).  Each LDS
    // word holds the sum of its terms' raw encodings (fixed_term); the
    // flush decodes them (brick_sums).  Without weights (WGT false: the
    // adjoint scatter) the w array is not used.
    unsigned long long re[NS], im[NS], w[WGT ? NS : 1];
    int item;
    float vmax[NT / 64];
};

// LDS layout of the owner-computes regions, in 4 x 4 tiles of (x, y):
// element ((z (BY/4) + y/4) (BX/4) + x/4) 16 + 4 (y mod 4) + x mod 4.
// gfx950's LDS has 32 four-byte banks and serves a 64-bit access for 16
// lanes per cycle, so 16 lanes whose elements differ mod 16 -- here: whose
// (x mod 4, y mod 4) differ -- never conflict (measured for ds_add_u64: 6.2
// CU-cycles per wave instruction, against 12.9 for random elements and 6.35
// for consecutive ones).  own_taps orders each lane's taps so that at every
// tap slot the 16 lanes of a group hold the 16 distinct (x mod 4, y mod 4).
// The element index is the sum of one part per axis (disjoint bit fields of
// the tile index and the in-tile position).  Sums as in BrickLds.
template <int BX, int BY, int BZ, int NT = TPB, bool WGT = true>
struct TileLds {
    static_assert(BX % 4 == 0 && BY % 4 == 0, "the tile layout needs x and y edges divisible by 4");
    static constexpr int NS = BX * BY * BZ;
    static constexpr int TX = BX / 4, TY = BY / 4;
    __device__ static int xpart(int x) { return ((x >> 2) << 4) | (x & 3); }
    __device__ static int ypart(int y) { return (((y >> 2) * TX) << 4) | ((y & 3) << 2); }
    __device__ static int zpart(int z) { return (z * TY * TX) << 4; }
    __device__ static int coords(int v, int& x, int& y, int& z) {
        const int t = v >> 4;
        x = ((t % TX) << 2) | (v & 3);
        y = (((t / TX) % TY) << 2) | ((v >> 2) & 3);
        z = t / (TX * TY);
        return v;
    }
    unsigned long long re[NS], im[NS], w[WGT ? NS : 1];
    int item;
    int pnext[2];   // next section pair of the item's direct and mirror passes
    float vmax[NT / 64];
};

#ifndef SPBP_LANE_SPREAD
#define SPBP_LANE_SPREAD 2
#endif
// power of two <= 64.  Measured at 128^2 into 256^3 (insertion ms per 20k
// projections): 1 (consecutive lanes, consecutive candidates) 55.0, 2 53.2,
// 4 53.6, 8 54.2, 16 54.3; with the current kernel (compaction, tile layout,
// counter-assigned section pairs) 1, 2 and 4 are within 0.3% (22.74-22.78)
constexpr int LANE_SPREAD = SPBP_LANE_SPREAD;
// Fixed-point accumulation in LDS.  A term t (an integer, |t| < 2^51) is
// encoded as the bit pattern of the double 1.5 * 2^52 + t, which is the
// 64-bit integer FIXED_OFF + t: one fused multiply-add forms it from the
// exact product (fixed_term).  A voxel's LDS word accumulates the encodings of
// its n terms: n FIXED_OFF + sum(t) mod 2^64.  FIXED_OFF's low 48 bits are
// zero, so a sum with |sum(t)| < 2^47 is its low 48 bits, sign-extended, and
// the high 16 bits then hold n * 0x4338 mod 2^16 = 8 (n * 2151 mod 2^13),
// which gives n back (n < 2^13; 2151 is odd).  The weights' word, whose terms
// need all 63 bits, is decoded by subtracting n FIXED_OFF.
//
// Bound on n: a work item holds at most 128 sections; one section reaches a
// voxel p through the lattice samples of its plane within 1.8 voxels of p
// (direct taps) or of -p (reflected taps), at most 28 each (a plane section
// of a cube of edge 3.6 has area <= 18.4 and perimeter <= 17.4, and a convex
// set holds at most area + perimeter / 2 + 1 lattice points), so n <= 7168.
// With terms below 2^VBITS, |sum(t)| < 7168 * 2^34 < 2^47.
constexpr int VBITS = 34;                   // |term| < 2^34 units; item sums < 2^47
constexpr int WBITS = 50;                   // weight terms <= 1 < 2^51 / 2^WBITS; item sums < 2^63
constexpr double WSCALE = 1125899906842624.0;            // 2^WBITS
constexpr double WINV = 1.0 / 1125899906842624.0;        // 2^-WBITS
constexpr unsigned long long FIXED_OFF = 0x4338000000000000ull;   // bits of 1.5 * 2^52
constexpr unsigned INV_2151 = 4951;         // 2151 * 4951 = 1 mod 2^13

// Encoding of round(a b) (|a b| < 2^51): the bits of fma(a, b, 1.5 * 2^52).
// The callers' products are exact in FP64 (two fp32 mantissas, one scaled by
// a power of two), so the fused multiply-add rounds once, exactly as the
// rounded product plus the offset would.
__device__ inline unsigned long long fixed_term(double a, double b) {
    return (unsigned long long)__double_as_longlong(__fma_rn(a, b, 6755399441055744.0));
}
// sum(t) of an accumulated word whose |sum(t)| < 2^47
__device__ inline long long fixed_sum48(unsigned long long s) { return (long long)(s << 16) >> 16; }
// n of an accumulated word with n < 2^13 terms, given its sum(t)
__device__ inline unsigned fixed_count(unsigned long long s, long long t) {
    const unsigned k = (unsigned)((s - (unsigned long long)t) >> 48);   // n * 0x4338 mod 2^16
    return ((k >> 3) * INV_2151) & 0x1FFFu;
}
// LDS add at a byte offset into one of the brick arrays
__device__ inline void lds_add(unsigned long long* arr, int ob, unsigned long long v) {
    atomicAdd((unsigned long long*)((char*)arr + ob), v);
}

// Per image of a chunk: the largest |Re| or |Im| of its raw R2C spectrum
// (the magnitudes the insertion's terms are formed from).  Block per image.
__global__ __launch_bounds__(TPB) void spec_absmax_kernel(const float2* __restrict__ spec, int nper,
                                                          float* __restrict__ out) {
    const float2* s = spec + (size_t)blockIdx.x * nper;
    float m = 0.0f;
    for (int i = threadIdx.x; i < nper; i += blockDim.x) {
        const float2 v = s[i];
        m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
    __shared__ float sm[TPB / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = fmaxf(m, sm[k]);
        out[blockIdx.x] = m;
    }
}

// One wave feeds the taps of the effective lattice samples of one projection
// whose E lies in the box [lo, hi] (and, with `exclude`, not in [elo, ehi])
// into the LDS brick.  Lattice lines are rows J (b) of the sample lattice,
// one per lane; candidates along a line are consecutive I (a), so
// neighbouring lanes read neighbouring spectrum entries.
struct Box {
    float lo[3], hi[3];
};

// MIR: the brick starts at kx <= LN2, so taps with p_x < 0 can land in it
// mirrored (at -p_x <= LN2); elsewhere they never do and the tap loop drops
// the per-tap mirror selects.
template <int BX, int BY, int BZ, bool MIR, typename P, int NT, bool WGT>
__device__ __forceinline__ void wave_project(BrickLds<BX, BY, BZ, NT, WGT>& L, const Geom& g, const float2* __restrict__ spec_p,
                                             int nx, int N, int x0, int y0, int z0, const Box box, bool exclude,
                                             const Box ex, const float* __restrict__ tab, P fltb, double vscale,
                                             double wscale) {
    const float* lo = box.lo;
    const float* hi = box.hi;
    const float* elo = ex.lo;
    const float* ehi = ex.hi;
    const int H = nx + 1;
    const int lane = threadIdx.x & 63;
    // b range: extent of r2 . E over the box
    float bmn = 0.f, bmx = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float c = g.r[3 + k];
        bmn += fminf(c * lo[k], c * hi[k]);
        bmx += fmaxf(c * lo[k], c * hi[k]);
    }
    const int b0 = (int)ceilf(bmn - 0.01f), b1 = (int)floorf(bmx + 0.01f);
    const int nl = b1 - b0 + 1;   // <= 64: edge + 2 margins, times sqrt(3), plus 2
    // this lane's line b = b0 + lane: a interval with lo <= a r1 + b r2 <= hi
    int cnt = 0, alo = 0;
    if (lane < nl) {
        const float bb = (float)(b0 + lane);
        float amin = -1e30f, amax = 1e30f;
        bool empty = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float r1 = g.r[k], off = bb * g.r[3 + k];
            if (fabsf(r1) > 1e-6f) {
                const float u = (lo[k] - off) / r1, w = (hi[k] - off) / r1;
                amin = fmaxf(amin, fminf(u, w));
                amax = fminf(amax, fmaxf(u, w));
            } else if (off < lo[k] - 0.01f || off > hi[k] + 0.01f) {
                empty = true;
            }
        }
        if (!empty) {
            const int a0 = (int)ceilf(amin - 0.01f), a1 = (int)floorf(amax + 0.01f);
            if (a1 >= a0) { alo = a0; cnt = a1 - a0 + 1; }
        }
    }
    int pref = cnt;   // inclusive scan over the wave
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(pref, d, 64);
        if (lane >= d) pref += u;
    }
    const int total = __shfl(pref, 63, 64);
    for (int q0 = 0; q0 < total; q0 += 64) {
        // lanes of one 16-lane LDS pass take candidates LANE_SPREAD apart
        // (samples further apart in the brick: fewer bank and address conflicts)
        const int q = q0 + (lane & (64 / LANE_SPREAD - 1)) * LANE_SPREAD + lane / (64 / LANE_SPREAD);
        // line of candidate q: smallest l with pref[l] > q (every lane searches)
        int l = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const int t = __shfl(pref, l + step - 1, 64);
            if (t <= q) l += step;
        }
        const int lpref = __shfl(pref, l, 64), lcnt = __shfl(cnt, l, 64), lalo = __shfl(alo, l, 64);
        if (q >= total) continue;
        const int bb = b0 + l;
        const int aa = lalo + (q - (lpref - lcnt));
        // canonical sample (I, J) of lattice point (aa, bb): the half plane I > 0 or I = 0, J >= 0
        const bool canon = aa > 0 || (aa == 0 && bb >= 0);
        const int I = canon ? aa : -aa, J = canon ? bb : -bb;
        if (I > nx || J < -nx + 1 || J > nx || !(I * I + J * J < nx * nx)) continue;   // oneline.f:53-55
        P x, y, z;
        sample_pos(g, I, J, x, y, z);
        const bool flip = x < (P)0;                                    // oneline.f:60-67
        if (flip == canon) continue;   // (aa, bb) must be the sample's effective lattice point
        if (flip) { x = -x; y = -y; z = -z; }
        if (x < lo[0] || x > hi[0] || y < lo[1] || y > hi[1] || z < lo[2] || z > hi[2]) continue;
        if (exclude && x >= elo[0] && x <= ehi[0] && y >= elo[1] && y <= ehi[1] && z >= elo[2] && z <= ehi[2]) continue;
        const int ix = nearest_voxel(x, N), iy = nearest_voxel(y, N), iz = nearest_voxel(z, N);
        if (!in_bounds(ix, iy, iz, nx)) continue;                      // oneline.f:73-75
        float2 v = spectrum_value(spec_p, I, J, N, H);
        if (flip) v.y = -v.y;
        // The nonzero taps of each axis are a run of at most 4 of the 5
        // reference taps i-2..i+2 (the table vanishes beyond 1.8 voxels, and
        // i-2 and i+2 are 4 apart): start at i-2 when its weight is nonzero.
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
        // the value in fixed-point units (exact: a power-of-two scale)
        const double vxs = __dmul_rn((double)v.x, vscale), vys = __dmul_rn((double)v.y, vscale);
        if constexpr (MIR) {
            // Targets inside the brick, per axis: a tap p with p_x >= 0 is written
            // at p, a tap with p_x < 0 at -p (conjugated).  Offsets are
            // brick-local; out-of-brick positions get a flag instead of a branch.
            int xo[4], xo8[4];
            bool xok[4], xmir[4];
            double vim[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int px = sx + t;
                xmir[t] = px < 0;
                xo[t] = (xmir[t] ? -px : px) - x0;
                xo8[t] = xo[t] * 8;
                xok[t] = (unsigned)xo[t] < (unsigned)BX;
                vim[t] = xmir[t] ? -vys : vys;
            }
#pragma unroll
            for (int lz = 0; lz < 4; ++lz) {
                const int pz = sz + lz;
                const int zd = pz - z0, zm = -pz - z0;
                const bool zdok = (unsigned)zd < (unsigned)BZ, zmok = (unsigned)zm < (unsigned)BZ;
#pragma unroll
                for (int ly = 0; ly < 4; ++ly) {
                    const int py = sy + ly;
                    const int yd = py - y0, ym = -py - y0;
                    const bool dok = zdok && (unsigned)yd < (unsigned)BY;
                    const bool mok = zmok && (unsigned)ym < (unsigned)BY;
                    const int offd = L.rowbase8(yd, zd), offm = L.rowbase8(ym, zm);
                    const int rd = L.rot8(yd, zd), rm = L.rot8(ym, zm);
                    const float tyz = __fmul_rn(ty[ly], tz[lz]);       // TY = TABI(Y) * TZ
#pragma unroll
                    for (int lx = 0; lx < 4; ++lx) {
                        const float wg = __fmul_rn(tx[lx], tyz);       // WG = TABI(X) * TY
                        const bool ok = xok[lx] && (xmir[lx] ? mok : dok) && wg != 0.0f;
                        if (ok) {
                            const int ob = xmir[lx] ? L.at8(offm, rm, xo8[lx]) : L.at8(offd, rd, xo8[lx]);
                            const double wd = (double)wg;
                            lds_add(L.re, ob, fixed_term(vxs, wd));
                            lds_add(L.im, ob, fixed_term(vim[lx], wd));
                            if constexpr (WGT) lds_add(L.w, ob, fixed_term(wd, wscale));
                        }
                    }
                }
            }
        } else {
            int xo8[4];
            bool xok[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int xo = sx + t - x0;
                xo8[t] = xo * 8;
                xok[t] = (unsigned)xo < (unsigned)BX;
            }
#pragma unroll
            for (int lz = 0; lz < 4; ++lz) {
                const int zd = sz + lz - z0;
                const bool zdok = (unsigned)zd < (unsigned)BZ;
#pragma unroll
                for (int ly = 0; ly < 4; ++ly) {
                    const int yd = sy + ly - y0;
                    const bool dok = zdok && (unsigned)yd < (unsigned)BY;
                    const int offd = L.rowbase8(yd, zd), rd = L.rot8(yd, zd);
                    const float tyz = __fmul_rn(ty[ly], tz[lz]);       // TY = TABI(Y) * TZ
#pragma unroll
                    for (int lx = 0; lx < 4; ++lx) {
                        const float wg = __fmul_rn(tx[lx], tyz);       // WG = TABI(X) * TY
                        if (xok[lx] && dok && wg != 0.0f) {
                            const int ob = L.at8(offd, rd, xo8[lx]);
                            const double wd = (double)wg;
                            lds_add(L.re, ob, fixed_term(vxs, wd));
                            lds_add(L.im, ob, fixed_term(vys, wd));
                            if constexpr (WGT) lds_add(L.w, ob, fixed_term(wd, wscale));
                        }
                    }
                }
            }
        }
    }
}

// Lattice lines (rows b of the sample lattice) of projection g crossing the box: [b0, b0 + nl).
__device__ inline void box_lines(const Geom& g, const Box& box, int& b0, int& nl) {
    float bmn = 0.f, bmx = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float c = g.r[3 + k];
        bmn += fminf(c * box.lo[k], c * box.hi[k]);
        bmx += fmaxf(c * box.lo[k], c * box.hi[k]);
    }
    b0 = (int)ceilf(bmn - 0.01f);
    nl = (int)floorf(bmx + 0.01f) - b0 + 1;
}

// Owner-computes bricks.  The LDS region is BX x BY x BZ voxels; its core
// (BX - 4) x (BY - 4) x (BZ - 4) starts 2 voxels in from the region's origin.
// A sample belongs to the brick whose core holds its nearest voxel (ix, iy,
// iz) and is evaluated there only: its direct taps (p_x >= 0, p within 2 of
// the nearest voxel per axis) all land in the region, so no tap is masked by
// the brick's edge.  Reflected taps (p_x < 0, written at -p) go to the brick
// whose core holds -p: bricks at kx = 0 run a second pass over the samples
// whose nearest voxel lies within 2 of -core with ix <= 1.  Every (sample,
// tap) pair is written once; regions overlap, and the flush adds each region
// to the half grid.
//
// MIRPASS false: owned samples, direct taps (XCHK: drop taps with p_x < 0,
// needed only at kx = 0).  MIRPASS true: the reflected taps whose target lies
// in the core.  (x0, y0, z0): the core's origin.

template <int BX, int BY, int BZ, bool MIRPASS, bool XCHK, typename P, int NT, bool WGT>
__device__ __forceinline__ void own_taps(TileLds<BX, BY, BZ, NT, WGT>& L, const Geom& gc,
                                         const float2* __restrict__ spec_p, int I, int J, int nx, int N, int x0,
                                         int y0, int z0, const float* __restrict__ tab, P fltb, double vscale,
                                         double wscale) {
    using Lds = TileLds<BX, BY, BZ, NT, WGT>;
    constexpr int CX = BX - 4, CY = BY - 4, CZ = BZ - 4;
    const int H = nx + 1;
    P x, y, z;
    sample_pos(gc, I, J, x, y, z);
    const bool flip = x < (P)0;                                        // oneline.f:60-67
    if (flip) { x = -x; y = -y; z = -z; }
    const int ix = nearest_voxel(x, N), iy = nearest_voxel(y, N), iz = nearest_voxel(z, N);
    float2 v = spectrum_value(spec_p, I, J, N, H);
    if (flip) v.y = -v.y;
    // The nonzero taps of each axis are a run of at most 4 of the 5
    // reference taps i-2..i+2 (the table vanishes beyond 1.8 voxels, and
    // i-2 and i+2 are 4 apart): start at i-2 when its weight is nonzero.
    const int sx = kb_weight(tab, dsub(x, ix - 2), fltb) != 0.0f ? ix - 2 : ix - 1;
    const int sy = kb_weight(tab, dsub(y, iy - 2), fltb) != 0.0f ? iy - 2 : iy - 1;
    const int sz = kb_weight(tab, dsub(z, iz - 2), fltb) != 0.0f ? iz - 2 : iz - 1;
    float tz[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) tz[t] = kb_weight(tab, dsub(z, sz + t), fltb);
    const double vxs = __dmul_rn((double)v.x, vscale), vys = __dmul_rn((double)v.y, vscale);
    if constexpr (MIRPASS) {
        // target -p, region-local: (-p) - (core origin - 2)
        const double vim = -vys;
        float tx[2], ty[4];
        int mx8[2], my8[4], mz8[4];
        bool mxok[2], myok[4], mzok[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            ty[t] = kb_weight(tab, dsub(y, sy + t), fltb);
            const int tyl = -(sy + t) - y0, tzl = -(sz + t) - z0;
            myok[t] = (unsigned)tyl < (unsigned)CY;
            mzok[t] = (unsigned)tzl < (unsigned)CZ;
            my8[t] = Lds::ypart(tyl + 2) * 8;
            mz8[t] = Lds::zpart(tzl + 2) * 8;
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {   // p_x = sx + t < 0 only for the first two taps
            tx[t] = kb_weight(tab, dsub(x, sx + t), fltb);
            const int px = sx + t;
            mxok[t] = px < 0 && -px - x0 < CX;
            mx8[t] = Lds::xpart(-px - x0 + 2) * 8;
        }
#pragma unroll
        for (int lz = 0; lz < 4; ++lz) {
#pragma unroll
            for (int ly = 0; ly < 4; ++ly) {
                const bool ok = mzok[lz] && myok[ly];
                const int row = my8[ly] + mz8[lz];
                const float tyz = __fmul_rn(ty[ly], tz[lz]);
#pragma unroll
                for (int lx = 0; lx < 2; ++lx) {
                    const float wg = __fmul_rn(tx[lx], tyz);
                    if (ok && mxok[lx] && wg != 0.0f) {
                        const int ob = row + mx8[lx];
                        const double wd = (double)wg;
                        lds_add(L.re, ob, fixed_term(vxs, wd));
                        lds_add(L.im, ob, fixed_term(vim, wd));
                        if constexpr (WGT) lds_add(L.w, ob, fixed_term(wd, wscale));
                    }
                }
            }
        }
    } else {
        // every direct tap is inside the region: p - (core origin - 2) in [0, B).
        // This lane visits its x and y taps in the cyclic order that puts
        // (x mod 4, y mod 4) = ((lane + kx) mod 4, (lane / 4 + ky) mod 4) at
        // tap slot (kx, ky): the 16 lanes of a group then hold 16 distinct
        // banks at every slot.
        const int lane = threadIdx.x & 63;
        const int xl = sx - x0 + 2, yl = sy - y0 + 2, zl = sz - z0 + 2;
        const int rx = (lane - xl) & 3, ry = ((lane >> 2) - yl) & 3;
        float tx[4], ty[4];
        int x8[4], y8[4], z8[4];
        bool xok[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int dx = (k + rx) & 3, dy = (k + ry) & 3;
            tx[k] = kb_weight(tab, dsub(x, sx + dx), fltb);
            ty[k] = kb_weight(tab, dsub(y, sy + dy), fltb);
            x8[k] = Lds::xpart(xl + dx) * 8;
            y8[k] = Lds::ypart(yl + dy) * 8;
            z8[k] = Lds::zpart(zl + k) * 8;
            xok[k] = !XCHK || sx + dx >= 0;
        }
#pragma unroll
        for (int lz = 0; lz < 4; ++lz) {
#pragma unroll
            for (int ly = 0; ly < 4; ++ly) {
                const int row = y8[ly] + z8[lz];
                const float tyz = __fmul_rn(ty[ly], tz[lz]);           // TY = TABI(Y) * TZ
#pragma unroll
                for (int lx = 0; lx < 4; ++lx) {
                    const float wg = __fmul_rn(tx[lx], tyz);           // WG = TABI(X) * TY
                    if (xok[lx] && wg != 0.0f) {
                        const int ob = row + x8[lx];
                        const double wd = (double)wg;
                        lds_add(L.re, ob, fixed_term(vxs, wd));
                        lds_add(L.im, ob, fixed_term(vys, wd));
                        if constexpr (WGT) lds_add(L.w, ob, fixed_term(wd, wscale));
                    }
                }
            }
        }
    }
}

// One insertion job for the brick kernel: a context's chunk of projections
// (spectra, geometry), its table, brick grid and lists, and its half grids.
struct BrickJob {
    double wscale;        // 2^WBITS, the weights' fixed-point scale (a kernel argument: kept in registers)
    const float2* spec;   // spectra of images img0, img0 + 1, ...
    const float* vmax;    // largest |Re|, |Im| of each of those spectra
    int img0;
    const Geom* geom;
    const float* tab;
    float fltb;
    int nx, N, C;
    BrickGrid bg;
    const int* cnt;
    const int* off;
    const int* item_off;
    const int* item_key;
    const int* list;
    float2* X0;
    float* W0;
    float2* X1;
    float* W1;
    float* Wc0;
    float* Wc1;
};
constexpr int MAX_JOBS = 64;

// One pass of a wave over its sections of a work item: the waves take the
// item's sections in pairs (s, s + 1) from a counter in LDS (L.pnext, one per
// pass, zeroed with the item), so that they finish together whatever the
// pairs' work; the sums are exact, so which wave takes which pair does not
// change them.  Two sections, A and B, share each candidate round:
// lanes 0..31 hold A's lattice lines and lanes 32..63 B's (without B, its
// lines are empty).  nlA, nlB <= 32: a rotation's section crosses at most the
// box diagonal + 2 lines, which the kernel's static_assert bounds for its
// region sizes (rotations are checked orthonormal on the host).
//
// Candidates are tested in rounds of 64 (one per lane), and the ones that
// pass are compacted into full batches of 64 before the tap loop runs, across
// rounds and across the wave's sections: every ds_add_u64 wave instruction
// costs the same whatever its active lanes, so the tap loop runs over full
// batches only, plus one partial batch at the end of the pass.  A round's
// accepted candidates are pushed (ds_permute) to the lanes after the pending
// ones; a round that completes a batch leaves its overflow in lanes 0.. as
// the next batch's start.  A candidate travels as its sample (I, J) and its
// section's place s in the item's list (packed in 30 bits); the tap loop
// reloads that section's geometry and recomputes the sample's position.
__device__ __forceinline__ unsigned pack_cand(int I, int J, int s) {
    return (unsigned)I | ((unsigned)(J + 1024) << 11) | ((unsigned)s << 23);
}

template <int BX, int BY, int BZ, bool MIRPASS, bool XCHK, typename P, int NT, bool WGT>
__device__ __forceinline__ void own_pass(TileLds<BX, BY, BZ, NT, WGT>& L, const BrickJob& Jb, const int* __restrict__ lst,
                                         int ns, int x0, int y0, int z0, const Box box,
                                         const float* __restrict__ tab, P fltb, double vscale, double wscale) {
    constexpr int CX = BX - 4, CY = BY - 4, CZ = BZ - 4;
    const int nx = Jb.nx, N = Jb.N, H = nx + 1;
    const float* lo = box.lo;
    const float* hi = box.hi;
    const int lane = threadIdx.x & 63;
    const bool side = lane >= 32;   // this lane's line: section A (0) or B (1), line b0 + (lane & 31)
    const unsigned long long below = (1ull << lane) - 1ull;
    Geom gA, gB;
    int b0A = 0, b0B = 0;
    int pref = 0, cnt = 0, alo = 0, total = 0, q0 = 0;
    int s = -2;                     // current pair
    int sA = -1;                    // the pair whose geometry gA, gB hold (A = sA, B = sA + 1)
    unsigned pend = 0;   // this lane's pending candidate (valid for lane < npend)
    int npend = 0;       // wave-uniform
    for (;;) {
        unsigned cur = 0;
        bool full = false;
        for (;;) {
            if (q0 >= total) {   // next pair of sections: their lattice lines across the box
                int nxt = 0;
                if (lane == 0) nxt = atomicAdd(&L.pnext[MIRPASS ? 1 : 0], 1);
                s = 2 * __shfl(nxt, 0, 64);
                if (s >= ns) break;
                const bool hasB = s + 1 < ns;
                gA = Jb.geom[lst[s]];
                gB = Jb.geom[lst[hasB ? s + 1 : s]];
                sA = s;
                int nlA, nlB;
                box_lines(gA, box, b0A, nlA);
                box_lines(gB, box, b0B, nlB);
                Geom g;
#pragma unroll
                for (int k = 0; k < 6; ++k) g.r[k] = side ? gB.r[k] : gA.r[k];
                const int nl = side ? (hasB ? nlB : 0) : nlA;
                const int b0 = side ? b0B : b0A;
                cnt = 0;
                alo = 0;
                if ((lane & 31) < nl) {
                    const float bb = (float)(b0 + (lane & 31));
                    float amin = -1e30f, amax = 1e30f;
                    bool empty = false;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float r1 = g.r[k], off = bb * g.r[3 + k];
                        if (fabsf(r1) > 1e-6f) {
                            const float u = (lo[k] - off) / r1, w = (hi[k] - off) / r1;
                            amin = fmaxf(amin, fminf(u, w));
                            amax = fminf(amax, fmaxf(u, w));
                        } else if (off < lo[k] - 0.01f || off > hi[k] + 0.01f) {
                            empty = true;
                        }
                    }
                    if (!empty) {
                        const int a0 = (int)ceilf(amin - 0.01f), a1 = (int)floorf(amax + 0.01f);
                        if (a1 >= a0) { alo = a0; cnt = a1 - a0 + 1; }
                    }
                }
                pref = cnt;
                for (int d = 1; d < 64; d <<= 1) {
                    const int u = __shfl_up(pref, d, 64);
                    if (lane >= d) pref += u;
                }
                total = __shfl(pref, 63, 64);
                q0 = 0;
                continue;
            }
            const int q = q0 + (lane & (64 / LANE_SPREAD - 1)) * LANE_SPREAD + lane / (64 / LANE_SPREAD);
            q0 += 64;
            // line of candidate q: smallest l with pref[l] > q (every lane searches)
            int l = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                const int t = __shfl(pref, l + step - 1, 64);
                if (t <= q) l += step;
            }
            const int lpref = __shfl(pref, l, 64), lcnt = __shfl(cnt, l, 64), lalo = __shfl(alo, l, 64);
            bool ok = q < total;
            const bool cside = l >= 32;
            const int bb = (cside ? b0B : b0A) + (l & 31);
            const int aa = lalo + (q - (lpref - lcnt));
            // canonical sample (I, J) of lattice point (aa, bb): I > 0 or I = 0, J >= 0
            const bool canon = aa > 0 || (aa == 0 && bb >= 0);
            const int I = canon ? aa : -aa, J = canon ? bb : -bb;
            ok = ok && !(I > nx || J < -nx + 1 || J > nx || !(I * I + J * J < nx * nx));   // oneline.f:53-55
            if (ok) {
                Geom gc;
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    gc.r[k] = cside ? gB.r[k] : gA.r[k];
                    gc.d[k] = cside ? gB.d[k] : gA.d[k];
                }
                P x, y, z;
                sample_pos(gc, I, J, x, y, z);
                const bool flip = x < (P)0;                            // oneline.f:60-67
                if (flip) { x = -x; y = -y; z = -z; }
                const int ix = nearest_voxel(x, N), iy = nearest_voxel(y, N), iz = nearest_voxel(z, N);
                // (aa, bb) must be the sample's effective lattice point; oneline.f:73-75
                ok = flip != canon && in_bounds(ix, iy, iz, nx);
                if constexpr (MIRPASS) {
                    // reflected taps exist only for ix <= 1; their targets -p lie within 2 of -(ix, iy, iz)
                    ok = ok && !(ix > 1 || -iy + 2 < y0 || -iy - 2 >= y0 + CY || -iz + 2 < z0 || -iz - 2 >= z0 + CZ);
                } else {
                    ok = ok && !(ix < x0 || ix >= x0 + CX || iy < y0 || iy >= y0 + CY || iz < z0 || iz >= z0 + CZ);
                }
            }
            const unsigned long long m = __ballot(ok);
            const int nv = __popcll(m);
            if (nv == 0) continue;
            // a permutation of the lanes: accepted ones first, in lane order
            const int rank = ok ? __popcll(m & below) : nv + __popcll(~m & below);
            const unsigned got = (unsigned)__builtin_amdgcn_ds_permute(((npend + rank) & 63) << 2,
                                                                        (int)pack_cand(I, J, s + (cside ? 1 : 0)));
            if (npend + nv >= 64) {
                cur = lane >= npend ? got : pend;
                pend = got;   // lanes < npend + nv - 64: the overflow
                npend += nv - 64;
                full = true;
                break;
            }
            if (lane >= npend && lane < npend + nv) pend = got;
            npend += nv;
        }
        if (!full) {
            if (npend == 0) break;
            cur = pend;
        }
        if (full || lane < npend) {
            const int I = (int)(cur & 0x7FFu), J = (int)((cur >> 11) & 0xFFFu) - 1024;
            // the section's geometry: the current pair's from registers; a
            // candidate carried over from an earlier pair reloads its own
            const int sc = (int)(cur >> 23);
            const bool inB = sc == sA + 1;
            Geom gc;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                gc.r[k] = inB ? gB.r[k] : gA.r[k];
                gc.d[k] = inB ? gB.d[k] : gA.d[k];
            }
            gc.img = inB ? gB.img : gA.img;
            if (sc != sA && !inB) {
                const Geom& gl = Jb.geom[lst[sc]];
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    gc.r[k] = gl.r[k];
                    gc.d[k] = gl.d[k];
                }
                gc.img = gl.img;
            }
            const float2* spec_p = Jb.spec + (size_t)(gc.img - Jb.img0) * N * H;
            own_taps<BX, BY, BZ, MIRPASS, XCHK, P, NT, WGT>(L, gc, spec_p, I, J, nx, N, x0, y0, z0, tab, fltb, vscale,
                                                            wscale);
        }
        if (!full) break;
    }
}


// Persistent workgroups pull (job, key, projection slice) items from one
// queue spanning every job of the launch (job-major order); the waves of
// a workgroup take the item's projections in turn and share the brick
// through LDS atomics.  The queue head must be zero at launch.
// The fp32-geometry instantiations are held to 128 VGPRs (4 waves per SIMD);
// the FP64-geometry ones need more and run 3.  WGT false: no weights are
// accumulated (every job's W0 and W1 are null: the adjoint scatter).
template <int BX, int BY, int BZ, typename P, int NT = TPB, bool OWN = false, bool WGT = true>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(P) == 4 ? 4 : 3)))
void insert_brick_kernel(const BrickJob* __restrict__ jobs, int njobs, int* __restrict__ next_item) {
    constexpr int NV = BX * BY * BZ;
    // owner-computes regions in the bank-conflict-free tile layout
    using Lds = std::conditional_t<OWN, TileLds<BX, BY, BZ, NT, WGT>, BrickLds<BX, BY, BZ, NT, WGT>>;
    __shared__ Lds L;
    __shared__ int jbase[MAX_JOBS + 1];
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int j = 0; j < njobs; ++j) {
            jbase[j] = acc;
            acc += jobs[j].item_off[2 * jobs[j].bg.nb];
        }
        jbase[njobs] = acc;
    }
    __syncthreads();
    const int total = jbase[njobs];
    const int wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    for (;;) {
        if (threadIdx.x == 0) L.item = atomicAdd(next_item, 1);
        __syncthreads();
        const int gitem = L.item;
        __syncthreads();
        if (gitem >= total) break;
        int j = 0;
        while (gitem >= jbase[j + 1]) ++j;
        const BrickJob& J = jobs[j];
        const int nx = J.nx, N = J.N, H = nx + 1;
        const BrickGrid bg = J.bg;
        const int item = gitem - jbase[j];
        const int key = J.item_key[item];
        const int b = key % bg.nb, half = key / bg.nb;
        const int s0 = (item - J.item_off[key]) * J.C;
        const int ns = min(J.C, J.cnt[key] - s0);
        const int* lst = J.list + J.off[key] + s0;
        int x0, y0, z0;
        brick_origin(bg, b, nx, x0, y0, z0);
        for (int i = threadIdx.x; i < L.NS; i += blockDim.x) {
            L.re[i] = 0ull;
            L.im[i] = 0ull;
            if constexpr (WGT) L.w[i] = 0ull;
        }
        if constexpr (OWN) {
            if (threadIdx.x == 0) { L.pnext[0] = 0; L.pnext[1] = 0; }
        }
        // fixed-point unit of the item: 2^-VBITS of its largest spectrum
        // magnitude (rounded up to a power of two), so every term is below
        // 2^VBITS units and a voxel's sum over the item below 2^47
        float m = 0.0f;
        for (int s = threadIdx.x; s < ns; s += blockDim.x) m = fmaxf(m, J.vmax[J.geom[lst[s]].img - J.img0]);
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
        if ((threadIdx.x & 63) == 0) L.vmax[wave] = m;
        __syncthreads();
        m = L.vmax[0];
        for (int k = 1; k < nwave; ++k) m = fmaxf(m, L.vmax[k]);
        int e = 0;
        if (m > 0.0f && isfinite(m)) (void)frexpf(m, &e);   // m < 2^e
        const double vscale = ldexp(1.0, VBITS - e), vinv = ldexp(1.0, e - VBITS);
        Box bx;
        bx.lo[0] = x0 - BRICK_MARGIN; bx.lo[1] = y0 - BRICK_MARGIN; bx.lo[2] = z0 - BRICK_MARGIN;
        bx.hi[0] = x0 + BX - 1 + BRICK_MARGIN; bx.hi[1] = y0 + BY - 1 + BRICK_MARGIN; bx.hi[2] = z0 + BZ - 1 + BRICK_MARGIN;
        Box mb;
        for (int k = 0; k < 3; ++k) { mb.lo[k] = -bx.hi[k]; mb.hi[k] = -bx.lo[k]; }
        const float* __restrict__ tab = J.tab;
        const double wsc = J.wscale;
        const P fltb = fltb_of<P>(J.fltb);
        if constexpr (OWN) {
            constexpr int CX = BX - 4, CY = BY - 4, CZ = BZ - 4;
            // E whose nearest voxel lies in the core; at kx = 0 also E within
            // 2.5 of -core with ix <= 1 (E_x < 1.5)
            Box oc;
            oc.lo[0] = x0 - 0.51f; oc.lo[1] = y0 - 0.51f; oc.lo[2] = z0 - 0.51f;
            oc.hi[0] = x0 + CX - 0.49f; oc.hi[1] = y0 + CY - 0.49f; oc.hi[2] = z0 + CZ - 0.49f;
            Box om;
            om.lo[0] = -0.01f; om.hi[0] = 1.51f;
            om.lo[1] = -(y0 + CY - 1) - 2.51f; om.hi[1] = -y0 + 2.51f;
            om.lo[2] = -(z0 + CZ - 1) - 2.51f; om.hi[2] = -z0 + 2.51f;
            // each wave takes the item's projections two at a time (s, s + 1).
            // Lines per projection across the core box (edges C + 0.02) or the
            // mirror box (1.52, BY + 0.02, BZ + 0.02): at most the box diagonal
            // + 2 for a rotation (orthonormal rows, checked on the host), <= 32.
            static_assert((BX - 3.98) * (BX - 3.98) + (BY - 3.98) * (BY - 3.98) + (BZ - 3.98) * (BZ - 3.98) <= 900.0 &&
                              1.52 * 1.52 + (BY + 0.02) * (BY + 0.02) + (BZ + 0.02) * (BZ + 0.02) <= 900.0,
                          "owner-computes region too large for 32 lines per projection");
            if (x0 == 0) {
                own_pass<BX, BY, BZ, false, true, P, NT, WGT>(L, J, lst, ns, x0, y0, z0, oc, tab, fltb, vscale, wsc);
                own_pass<BX, BY, BZ, true, false, P, NT, WGT>(L, J, lst, ns, x0, y0, z0, om, tab, fltb, vscale, wsc);
            } else {
                own_pass<BX, BY, BZ, false, false, P, NT, WGT>(L, J, lst, ns, x0, y0, z0, oc, tab, fltb, vscale, wsc);
            }
        }
        if constexpr (!OWN) {
            for (int s = wave; s < ns; s += nwave) {
                const int p = lst[s];
                const Geom g = J.geom[p];
                const float2* spec_p = J.spec + (size_t)(g.img - J.img0) * N * H;
                if (x0 <= LN2) {
                    wave_project<BX, BY, BZ, true, P, NT, WGT>(L, g, spec_p, nx, N, x0, y0, z0, bx, false, bx, tab, fltb, vscale, wsc);
                    wave_project<BX, BY, BZ, true, P, NT, WGT>(L, g, spec_p, nx, N, x0, y0, z0, mb, true, bx, tab, fltb, vscale, wsc);
                } else {
                    wave_project<BX, BY, BZ, false, P, NT, WGT>(L, g, spec_p, nx, N, x0, y0, z0, bx, false, bx, tab, fltb, vscale, wsc);
                }
            }
        }
        __syncthreads();
        // add the brick to the half grid
        float2* __restrict__ X = half ? J.X1 : J.X0;
        [[maybe_unused]] float* __restrict__ W = half ? J.W1 : J.W0;
        [[maybe_unused]] float* __restrict__ Wc = half ? J.Wc1 : J.Wc0;   // null unless compensated weights
        for (int v = threadIdx.x; v < NV; v += blockDim.x) {
            int lx, ly, lz;
            const int i = L.coords(v, lx, ly, lz);
            const unsigned long long sre = L.re[i];
            if (sre == 0ull) continue;   // no tap reached the voxel (n = 0)
            const int rg = OWN ? 2 : 0;   // owner-computes: the region starts 2 voxels before the core
            const int kx = x0 - rg + lx, ky = y0 - rg + ly, kz = z0 - rg + lz;
            // brick overhangs the grid (partial last brick; a region's halo)
            if (kx < 0 || ky < -nx + 1 || kz < -nx + 1 || kx > nx || ky > nx || kz > nx) continue;
            const size_t o = ((size_t)(kz >= 0 ? kz : N + kz) * N + (ky >= 0 ? ky : N + ky)) * H + kx;
            const long long tre = fixed_sum48(sre);
            unsafeAtomicAdd(&X[o].x, (float)__dmul_rn((double)tre, vinv));
            unsafeAtomicAdd(&X[o].y, (float)__dmul_rn((double)fixed_sum48(L.im[i]), vinv));
            if constexpr (WGT) {
                const unsigned long long tw = L.w[i] - (unsigned long long)fixed_count(sre, tre) * FIXED_OFF;
                const double wd = __dmul_rn((double)(long long)tw, WINV);
                const float wi = (float)wd;
                if (Wc) twosum_atomic_add(W, Wc, o, wi, (float)(wd - (double)wi));
                else if (W) unsafeAtomicAdd(&W[o], wi);
            }
        }
        __syncthreads();
    }
}

}  // namespace spbp
