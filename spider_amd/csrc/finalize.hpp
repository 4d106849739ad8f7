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
    // NRMW2 decisions fixed by the reference's sums, or null: bits mshift and
    // mshift + 1 of each voxel's byte hold 1 (zero) or 2 (divide), 0 where the
    // GPU sums decide; one byte serves the three volumes of a context
    const unsigned char* mask = nullptr;
    int mshift = 0;
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
    // NRMW2 (bp3f.f:1359): divide where W > 0.1; near 0.1 the decision may
    // come from the reference's own sums
    const unsigned char m = a.mask ? (a.mask[((size_t)k * N + j) * (nx + 1) + i] >> a.mshift) & 3 : 0;
    if (a.Wc[0]) {
        // compensated weights: threshold and division in FP64
        if (m ? m == 2 : wd > (double)0.1f)
            return make_float2((float)((double)(x.x * s) / wd), (float)((double)(x.y * s) / wd));
        return make_float2(0.f, 0.f);
    }
    if (m ? m == 2 : w > 0.1f)
        return make_float2(__fdiv_rn(x.x * s, w), __fdiv_rn(x.y * s, w));
    return make_float2(0.f, 0.f);
}

// ------------------------------------------------------------------ batched finalisation
// One volume of a batched finalisation: its accumulators, transform buffers,
// output and its first block in each of the two grids one launch covers (the
// half-grid grid of the normalisation, the NX^3 grid of the window steps).
// A launch covers a contiguous run of jobs; block offsets are relative to the
// run's first job.
struct FinalJob {
    Acc a;
    float2* work;          // half-complex input of the inverse transform
    const float* real;     // its N^3 output
    float* out;            // NX^3 volume
    const float* wkb;      // WINDKB2A divisor by squared radius
    double* psum;          // per volume block: outer-shell sum and count
    unsigned* pcnt;
    float* mean;
    int nx, N;
    float pix;             // 1 / N^3
    unsigned hb0, vb0;     // first block in the half-grid / volume grid
    unsigned vnb;          // volume blocks
    double invH, invN, invL;   // 1 / (NX + 1), 1 / N, 1 / NX
};

// Job of block b: the last of the run whose first block is <= b.
template <bool VOL>
__device__ inline int job_of_block(const FinalJob* __restrict__ jobs, int nj, unsigned b) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((VOL ? jobs[mid].vb0 : jobs[mid].hb0) <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Half-grid voxels per thread of normalize_kernel: a block's job lookup and
// job record (a serial chain of scalar loads) are paid once per 4 TPB voxels,
// and the voxels' loads are all issued before the first store.
constexpr int NORM_VPT = 4;

__global__ __launch_bounds__(TPB) void normalize_kernel(const FinalJob* __restrict__ jobs, int nj) {
    const FinalJob& f = jobs[job_of_block<false>(jobs, nj, blockIdx.x)];
    const int nx = f.nx, N = f.N, H = nx + 1;
    const size_t total = (size_t)H * N * N;
    const size_t base = (size_t)(blockIdx.x - f.hb0) * (TPB * NORM_VPT) + threadIdx.x;
    float2 v[NORM_VPT];
#pragma unroll
    for (int u = 0; u < NORM_VPT; ++u) {
        const size_t idx = base + (size_t)u * TPB;
        if (idx >= total) continue;
        int i, j, k;
        half_coords(idx, H, N, f.invH, f.invN, i, j, k);
        v[u] = norm_val(f.a, nx, N, i, j, k);
        if (i == 0 || i == nx) {
            // c2r reads the kx = 0 and kx = N/2 planes as their Hermitian part
            const float2 m = norm_val(f.a, nx, N, i, (N - j) % N, (N - k) % N);
            v[u] = make_float2((v[u].x + m.x) * 0.5f, (v[u].y - m.y) * 0.5f);
        }
    }
#pragma unroll
    for (int u = 0; u < NORM_VPT; ++u) {
        const size_t idx = base + (size_t)u * TPB;
        if (idx < total) f.work[idx] = make_float2(v[u].x, -v[u].y);   // SPIDER inverse: conjugate first
    }
}

// ------------------------------------------------------------------ WINDKB2A
__global__ __launch_bounds__(TPB) void window_kernel(const FinalJob* __restrict__ jobs, int nj) {
    const FinalJob& f = jobs[job_of_block<true>(jobs, nj, blockIdx.x)];
    const int L = f.nx, N = f.N;
    const unsigned lb = blockIdx.x - f.vb0;
    const size_t total = (size_t)L * L * L;
    const size_t idx = (size_t)lb * blockDim.x + threadIdx.x;
    double s = 0.0;
    unsigned c = 0;
    if (idx < total) {
        int i, j, k;
        half_coords(idx, L, L, f.invL, f.invL, i, j, k);
        const int ip = (N - L) / 2 + (L % 2);
        const int cc = L / 2;
        float v = f.real[((size_t)(ip + k) * N + (ip + j)) * N + (ip + i)] * f.pix;   // FMRS_3 scale
        const int lr = (i - cc) * (i - cc) + (j - cc) * (j - cc) + (k - cc) * (k - cc);
        if (lr <= cc * cc) {
            v = __fdiv_rn(v, f.wkb[lr]);
            if (lr >= (cc - 1) * (cc - 1)) { s = v; c = 1; }
        } else {
            v = 0.0f;
        }
        f.out[idx] = v;
    }
    block_sum2(s, c);
    if (threadIdx.x == 0) { f.psum[lb] = s; f.pcnt[lb] = c; }
}

// One block per job: the mean over the outer shell.
__global__ __launch_bounds__(TPB) void shell_mean_kernel(const FinalJob* __restrict__ jobs) {
    const FinalJob& f = jobs[blockIdx.x];
    double s = 0.0;
    unsigned c = 0;
    for (unsigned b = threadIdx.x; b < f.vnb; b += blockDim.x) { s += f.psum[b]; c += f.pcnt[b]; }
    block_sum2(s, c);
    if (threadIdx.x == 0) f.mean[0] = (float)(s / (double)(c ? c : 1));
}

__global__ __launch_bounds__(TPB) void subtract_kernel(const FinalJob* __restrict__ jobs, int nj) {
    const FinalJob& f = jobs[job_of_block<true>(jobs, nj, blockIdx.x)];
    const int L = f.nx;
    const size_t total = (size_t)L * L * L;
    const size_t idx = (size_t)(blockIdx.x - f.vb0) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    int i, j, k;
    half_coords(idx, L, L, f.invL, f.invL, i, j, k);
    const int cc = L / 2;
    const int lr = (i - cc) * (i - cc) + (j - cc) * (j - cc) + (k - cc) * (k - cc);
    if (lr <= cc * cc) f.out[idx] = f.out[idx] - f.mean[0];
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

// ------------------------------------------------------------------ SYMVOL
// TRIQUAD (triquad.f): 27-point triquadratic interpolation at offsets
// (R, S, T) from the centre point F[13].  Implicit typing in the reference
// makes RS, ST, RT, RST and the factors 1-R^2, 1-R, 1+R, ... REAL (rounded
// from DOUBLE); a term multiplied by R, S or T itself is evaluated in DOUBLE;
// the sum runs left to right, in REAL until the first DOUBLE term; the result
// is REAL.  Reproduced operation by operation (no contraction).
__device__ inline float triquad(double R, double S, double T, const float* F) {
    const float C2 = 0.5f, C4 = 0.25f, C8 = 0.125f;
    const float RS = (float)(R * S), ST = (float)(S * T), RT = (float)(R * T);
    const float RST = (float)(R * (double)ST);
    const float RSQ = (float)(1.0 - R * R), SSQ = (float)(1.0 - S * S), TSQ = (float)(1.0 - T * T);
    const float RM1 = (float)(1.0 - R), SM1 = (float)(1.0 - S), TM1 = (float)(1.0 - T);
    const float RP1 = (float)(1.0 + R), SP1 = (float)(1.0 + S), TP1 = (float)(1.0 + T);
    float a = (-C8) * RST * RM1 * SM1 * TM1 * F[0];
    a = a + (C4) * ST * RSQ * SM1 * TM1 * F[1];
    a = a + (C8) * RST * RP1 * SM1 * TM1 * F[2];
    a = a + (C4) * RT * RM1 * SSQ * TM1 * F[3];
    double d = (double)a + (double)(-C2) * T * RSQ * SSQ * TM1 * F[4];
    d = d + (double)((-C4) * RT * RP1 * SSQ * TM1 * F[5]);
    d = d + (double)((C8) * RST * RM1 * SP1 * TM1 * F[6]);
    d = d + (double)((-C4) * ST * RSQ * SP1 * TM1 * F[7]);
    d = d + (double)((-C8) * RST * RP1 * SP1 * TM1 * F[8]);
    d = d + (double)((C4) * RS * RM1 * SM1 * TSQ * F[9]);
    d = d + (double)(-C2) * S * RSQ * SM1 * TSQ * F[10];
    d = d + (double)((-C4) * RS * RP1 * SM1 * TSQ * F[11]);
    d = d + (double)(-C2) * R * RM1 * SSQ * TSQ * F[12];
    d = d + (double)(RSQ * SSQ * TSQ * F[13]);
    d = d + (double)(C2) * R * RP1 * SSQ * TSQ * F[14];
    d = d + (double)((-C4) * RS * RM1 * SP1 * TSQ * F[15]);
    d = d + (double)(C2) * S * RSQ * SP1 * TSQ * F[16];
    d = d + (double)((C4) * RS * RP1 * SP1 * TSQ * F[17]);
    d = d + (double)((C8) * RST * RM1 * SM1 * TP1 * F[18]);
    d = d + (double)((-C4) * ST * RSQ * SM1 * TP1 * F[19]);
    d = d + (double)((-C8) * RST * RP1 * SM1 * TP1 * F[20]);
    d = d + (double)((-C4) * RT * RM1 * SSQ * TP1 * F[21]);
    d = d + (double)(C2) * T * RSQ * SSQ * TP1 * F[22];
    d = d + (double)((C4) * RT * RP1 * SSQ * TP1 * F[23]);
    d = d + (double)((-C8) * RST * RM1 * SP1 * TP1 * F[24]);
    d = d + (double)((C4) * ST * RSQ * SP1 * TP1 * F[25]);
    d = d + (double)((C8) * RST * RP1 * SP1 * TP1 * F[26]);
    return (float)d;
}

// SYMVOL (symvol.f:39-219): q2(r) = (1/NSYM) sum_k v_k(r) for r in
// [KLX, KNX]^3, KLX = -NX/2, KNX = NX/2 - 1 (even NX) or NX/2 (odd);
// v_k = q1 sampled at RM_k^T r by TRIQUAD when |r|^2 <= (min|bound| - 1)^2 and
// RM_k is not the identity (diagonal all 1.0), else q1(r).  Summed in the
// reference's order (k = 1..NSYM) in fp32.  One thread per voxel.
__global__ __launch_bounds__(TPB) void symvol_kernel(const float* __restrict__ q1, float* __restrict__ q2, int nx,
                                                     const float* __restrict__ sm, int nsym) {
    const size_t nv = (size_t)nx * nx * nx;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nv) return;
    const int klx = -nx / 2, knx = (nx % 2 == 0) ? nx / 2 - 1 : nx / 2;
    const int mn = min(abs(klx), abs(knx));
    const int iradi = (mn - 1) * (mn - 1);
    const int ix = (int)(idx % nx) + klx, iy = (int)((idx / nx) % nx) + klx, iz = (int)(idx / ((size_t)nx * nx)) + klx;
    const int icrd = ix * ix + iy * iy + iz * iz;
    float acc = 0.0f;
    for (int k = 0; k < nsym; ++k) {
        const float* rm = sm + 9 * k;
        const bool ident = rm[0] == 1.0f && rm[4] == 1.0f && rm[8] == 1.0f;
        float v;
        if (ident || icrd > iradi) {
            v = q1[idx];
        } else {
            double qr[3];
            for (int c = 0; c < 3; ++c) {
                const double qrt = (double)__fadd_rn(__fmul_rn(rm[3 * c + 1], (float)iy), __fmul_rn(rm[3 * c + 2], (float)iz));
                qr[c] = qrt + (double)__fmul_rn(rm[3 * c], (float)ix);
            }
            const int iox = (int)floor(qr[0]), ioy = (int)floor(qr[1]), ioz = (int)floor(qr[2]);
            float F[27];
#pragma unroll
            for (int l = 0; l < 27; ++l) {
                const int x = iox + (l % 3) - 1 - klx, y = ioy + (l / 3) % 3 - 1 - klx, z = ioz + l / 9 - 1 - klx;
                F[l] = q1[((size_t)z * nx + y) * nx + x];
            }
            v = triquad(qr[0] - iox, qr[1] - ioy, qr[2] - ioz, F);
        }
        acc = __fadd_rn(acc, v);
    }
    q2[idx] = acc / (float)nsym;
}

}  // namespace spbp
