// insert_gather.hpp -- voxel-driven insertion (strategy 3): ONELINE's scatter
// computed as a gather.
//
// Each thread owns one half-grid voxel t and adds, in registers, every term
// the chunk's sections put on it.  Per section (chunk order) it visits the
// lattice samples (I, J) whose 5x5x5 window can cover t or -t, rows J
// ascending and I ascending within a row, the direct tap before the reflected
// one: the order in which the serial reference's loops add to X(t) and W(t)
// (images in input order, symmetries in order, bp32f.f:866-898; rows,
// columns, direct then reflected taps, oneline.f:53-152).  The running sums
// start from the accumulators' values and are written back once per launch,
// so per voxel and half the fp32 sums are the serial reference's bit for bit
// and do not depend on timing: no atomics, no fixed point, no reordering.
//
// The price is the search.  A workgroup is a brick of 8 x 8 x 4 voxels.  It
// lists, in order, the sections whose plane passes within reach of the
// brick (ordered ballot compaction, 256 sections at a time), then every
// thread tests each listed plane against its own voxel and, where it passes
// within reach, walks the disc of candidate samples around t and -t
// (for_candidates, shared with the NRMW2 re-summation).  Most candidates fail
// the cheap window test, so the kernel does several times the arithmetic of
// the scatter plans for the same terms; the tuner times it beside them.
#pragma once
#include "common.hpp"
#include "resolve.hpp"

namespace spbp {

constexpr int GBX = 8, GBY = 8, GBZ = 4;   // voxels per gather brick (signed x, y, z)
constexpr int GTPB = GBX * GBY * GBZ;      // one thread per voxel

// Running sums of one half at one voxel.
struct GatherSum {
    float2 x;
    float w, c;   // weight and, with compensated weights, the sum of its rounding errors
};

template <bool WGT, bool TWOSUM>
__device__ __forceinline__ void gather_add(GatherSum& a, float2 v, float wg, bool conj) {
    a.x.x = __fadd_rn(a.x.x, __fmul_rn(v.x, wg));
    a.x.y = __fadd_rn(a.x.y, __fmul_rn(conj ? -v.y : v.y, wg));
    if (!WGT) return;
    if (TWOSUM) {   // as twosum_atomic_add, on the register copy
        const float old = a.w;
        const float s = __fadd_rn(old, wg);
        const float bb = __fsub_rn(s, old);
        const float e = __fadd_rn(__fsub_rn(old, __fsub_rn(s, bb)), __fsub_rn(wg, bb));
        a.w = s;
        if (e != 0.0f) a.c = __fadd_rn(a.c, e);
    } else {
        a.w = __fadd_rn(a.w, wg);
    }
}

// The terms lattice sample (I, J) of section g puts on voxel t, with the
// arithmetic of insert_direct_kernel: Friedel flip, nearest voxel, bounds,
// then the direct tap at t and the reflected tap at -t (oneline.f:53-152).
template <typename P, bool WGT, bool TWOSUM>
__device__ __forceinline__ void gather_sample(const Geom& g, int I, int J, int nx, int N, int H,
                                              const float2* __restrict__ spec_p, const float* __restrict__ tab, P fltb,
                                              int tx, int ty, int tz, GatherSum& a) {
    if (!((I * I + J * J) < nx * nx) || (I == 0 && J < 0)) return;   // oneline.f:53-55
    {
        // a term needs the sample (at q or -q) within 2.5 voxels per axis of t
        const float fI = (float)I, fJ = (float)J;
        const float qx = fI * g.r[0] + fJ * g.r[3], qy = fI * g.r[1] + fJ * g.r[4], qz = fI * g.r[2] + fJ * g.r[5];
        const float dm = fmaxf(fabsf(qx - tx), fmaxf(fabsf(qy - ty), fabsf(qz - tz)));
        const float dp = fmaxf(fabsf(qx + tx), fmaxf(fabsf(qy + ty), fabsf(qz + tz)));
        if (fminf(dm, dp) > 2.6f) return;
    }
    P x, y, z;
    sample_pos(g, I, J, x, y, z);
    const bool flip = x < (P)0;                                      // oneline.f:60-67
    if (flip) { x = -x; y = -y; z = -z; }
    const int ix = nearest_voxel(x, N), iy = nearest_voxel(y, N), iz = nearest_voxel(z, N);
    if (!in_bounds(ix, iy, iz, nx)) return;
    const bool direct = tx >= ix - LN2 && tx >= 0 && tx <= ix + LN2 && abs(ty - iy) <= LN2 && abs(tz - iz) <= LN2;
    const bool refl = ix < LN2 && -tx >= ix - LN2 && -tx <= -1 && abs(-ty - iy) <= LN2 && abs(-tz - iz) <= LN2;
    if (!direct && !refl) return;
    float2 v = spectrum_value(spec_p, I, J, N, H);
    if (flip) v.y = -v.y;
    if (direct) {
        const float tyz = __fmul_rn(kb_weight(tab, dsub(y, ty), fltb), kb_weight(tab, dsub(z, tz), fltb));
        const float wg = __fmul_rn(kb_weight(tab, dsub(x, tx), fltb), tyz);
        if (wg != 0.0f) gather_add<WGT, TWOSUM>(a, v, wg, false);
    }
    if (refl) {   // tap p = -t, written conjugated at t (oneline.f:118-152)
        const float tyz = __fmul_rn(kb_weight(tab, dsub(y, -ty), fltb), kb_weight(tab, dsub(z, -tz), fltb));
        const float wg = __fmul_rn(kb_weight(tab, dsub(x, -tx), fltb), tyz);
        if (wg != 0.0f) gather_add<WGT, TWOSUM>(a, v, wg, true);
    }
}

// Can section g put weight on some voxel of the brick (centre c, half
// extents h)?  The plane must pass within plane_near's reach of a point of
// the box; the box test is exact for the plane's distance function.
__device__ inline bool gather_brick_near(const Geom& g, double cx, double cy, double cz, double hx, double hy, double hz) {
    const double d = (double)g.r[6] * cx + (double)g.r[7] * cy + (double)g.r[8] * cz;
    const double e = fabs((double)g.r[6]) * hx + fabs((double)g.r[7]) * hy + fabs((double)g.r[8]) * hz;
    return fabs(d) - e <= 4.45;
}

// Grid: (ceil((NX + 1) / 8), ceil(N / 8), ceil(N / 4)) bricks over signed
// kx in [0, NX], ky, kz in [-NX + 1, NX].  ng sections of images img0..,
// their raw R2C spectra at spec.
template <typename P, bool WGT, bool TWOSUM>
__global__ __launch_bounds__(GTPB) void insert_gather_kernel(
    const float2* __restrict__ spec, int img0, const Geom* __restrict__ geom, int ng, int nx, int N,
    const float* __restrict__ tab, float fltb_f, float2* __restrict__ X0, float* __restrict__ W0,
    float2* __restrict__ X1, float* __restrict__ W1, float* __restrict__ Wc0, float* __restrict__ Wc1) {
    __shared__ int lst[GTPB];
    __shared__ int wcnt[GTPB / 64];
    const int H = nx + 1;
    const int x0 = blockIdx.x * GBX, y0 = -nx + 1 + (int)blockIdx.y * GBY, z0 = -nx + 1 + (int)blockIdx.z * GBZ;
    {
        // no sample reaches beyond |t| < NX + 2.5 sqrt(3) (|q| < NX, taps within 2.5 per axis)
        auto lo = [](int a, int b) { return a > 0 ? (double)a : (b < 0 ? (double)-b : 0.0); };
        const double mx = lo(x0, x0 + GBX - 1), my = lo(y0, y0 + GBY - 1), mz = lo(z0, z0 + GBZ - 1);
        if (mx * mx + my * my + mz * mz > (nx + 4.5) * (double)(nx + 4.5)) return;   // whole block
    }
    const int lx = threadIdx.x % GBX, ly = (threadIdx.x / GBX) % GBY, lz = threadIdx.x / (GBX * GBY);
    const int tx = x0 + lx, ty = y0 + ly, tz = z0 + lz;
    const bool active = tx <= nx && ty <= nx && tz <= nx;
    const size_t o = active ? ((size_t)(tz >= 0 ? tz : N + tz) * N + (ty >= 0 ? ty : N + ty)) * H + tx : 0;
    GatherSum a0{}, a1{};
    if (active) {
        a0.x = X0[o];
        a1.x = X1 ? X1[o] : make_float2(0.f, 0.f);
        if (WGT) {
            a0.w = W0[o];
            a1.w = W1 ? W1[o] : 0.f;
            if (TWOSUM) {
                a0.c = Wc0[o];
                a1.c = Wc1 ? Wc1[o] : 0.f;
            }
        }
    }
    bool touched0 = false, touched1 = false;
    const double cx = x0 + 0.5 * (GBX - 1), cy = y0 + 0.5 * (GBY - 1), cz = z0 + 0.5 * (GBZ - 1);
    const P fltb = fltb_of<P>(fltb_f);
    const size_t nspec = (size_t)N * H;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int s0 = 0; s0 < ng; s0 += GTPB) {
        const int s = s0 + threadIdx.x;
        const bool hit = s < ng && gather_brick_near(geom[s], cx, cy, cz, 0.5 * (GBX - 1), 0.5 * (GBY - 1), 0.5 * (GBZ - 1));
        const unsigned long long bal = __ballot(hit);
        if (lane == 0) wcnt[w] = __popcll(bal);
        __syncthreads();
        int base = 0, tot = 0;
        for (int k = 0; k < GTPB / 64; ++k) {
            if (k < w) base += wcnt[k];
            tot += wcnt[k];
        }
        if (hit) lst[base + __popcll(bal & ((1ull << lane) - 1ull))] = s;   // kept in section order
        __syncthreads();
        for (int q = 0; q < tot; ++q) {
            const Geom g = geom[__builtin_amdgcn_readfirstlane(lst[q])];
            if (!active || !plane_near(g, tx, ty, tz)) continue;
            const float2* __restrict__ sp = spec + (size_t)(g.img - img0) * nspec;
            if (g.half) {
                if (!X1) continue;   // one-volume contexts record no second half
                touched1 = true;
                for_candidates<P>(g, nx, tx, ty, tz, [&](int I, int J) {
                    gather_sample<P, WGT, TWOSUM>(g, I, J, nx, N, H, sp, tab, fltb, tx, ty, tz, a1);
                });
            } else {
                touched0 = true;
                for_candidates<P>(g, nx, tx, ty, tz, [&](int I, int J) {
                    gather_sample<P, WGT, TWOSUM>(g, I, J, nx, N, H, sp, tab, fltb, tx, ty, tz, a0);
                });
            }
        }
        __syncthreads();   // lst is rewritten by the next segment
    }
    if (touched0) {
        X0[o] = a0.x;
        if (WGT) {
            W0[o] = a0.w;
            if (TWOSUM) Wc0[o] = a0.c;
        }
    }
    if (touched1) {
        X1[o] = a1.x;
        if (WGT) {
            W1[o] = a1.w;
            if (TWOSUM) Wc1[o] = a1.c;
        }
    }
}

}  // namespace spbp
