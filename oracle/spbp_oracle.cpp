// spbp_oracle.cpp -- CPU oracle for SPIDER's Fourier-gridding backprojection
// (operations `BP 3F` and `BP 32F`).
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
// product path (spider_amd, libspbp) never links, loads or calls it.
//
// Two instantiations of every step:
//   * "f32": a faithful restatement of the reference in single precision, in
//     the reference's loop order and rounding (fp32 positions, NINT table
//     index, fp32 accumulation).  Used to compare with the reference's own
//     routines compiled from source (oracle/_ref) and as the CPU baseline.
//   * "f64": the same discretised algorithm evaluated in double precision
//     (exact pi, FP64 rotation, FP64 KB table, FP64 accumulation and FFTs).
//     This is the "FP64 oracle" that GPU results are measured against.
//
// Citations are to files of the reference tree (spider-em/SPIDER, src/).
// Parity pins (see tests/test_oracle_*.py):
//   cang / kb table / asta mean / ONELINE / BESI1 / RIBSL  -> compared with the
//       reference's own Fortran routines compiled by oracle/Makefile
//   FFTs -> numpy.fft (DFT definition)
//   SYMPLANE0 -> Hermitian invariant + brute-force pairing on tiny grids
//   NRMW2 / WINDKB2A -> closed forms (I_{3/2}) + special cases
//   end to end -> proc/testbp3f.tst known-answer range (12 x 75^2 MO T images)
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int LN2 = 2;        // half width of the 5-tap KB kernel (bp32f.f:193-194)
constexpr int LTAB = 4999;    // KB table length - 1 (bp32f.f:81, oneline.f:44)

// ---------------------------------------------------------------- Bessel
// Modified Bessel I1 by its power series; converges to double precision for
// the arguments used here (0 <= x <= 6.5): the FP64 form of BESI1.
double bessel_i1(double x) {
    double h = 0.5 * x, term = h, sum = h, h2 = h * h;
    for (int k = 1; k < 200; ++k) {
        term *= h2 / (double(k) * double(k + 1));
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    return sum;
}
// I1 in single precision as BESI1 -> CALCI1 computes it for |x| < 15
// (besi1.f:3-230): W. J. Cody's published rational approximation (SPECFUN,
// ACM TOMS Algorithm 715), numerator in x^2 (15 terms) and denominator in
// x^2 - 225 (monic, 5 further terms), both by Horner, value (P / Q) |x|.
// Every step rounds to float as the REAL code does.
float cody_i1_f32(float x) {
    static const float num[15] = {
        -1.9705291802535139930E-19f, -6.5245515583151902910E-16f, -1.1928788903603238754E-12f,
        -1.4831904935994647675E-09f, -1.3466829827635152875E-06f, -9.1746443287817501309E-04f,
        -4.7207090827310162436E-01f, -1.8225946631657315931E+02f, -5.1894091982308017540E+04f,
        -1.0588550724769347106E+07f, -1.4828267606612366099E+09f, -1.3357437682275493024E+11f,
        -6.9876779648010090070E+12f, -1.7732037840791591320E+14f, -1.4577180278143463643E+15f};
    static const float den[5] = {-4.0076864679904189921E+03f, 7.4810580356655069138E+06f,
                                 -8.0059518998619764991E+09f, 4.8544714258273622913E+12f,
                                 -1.3218168307321442305E+15f};
    const float a = fabsf(x);
    float r;
    if (a < 2.98E-8f) {
        r = 0.5f * a;
    } else {
        const float t = a * a;
        float pn = num[0];
        for (float c : std::vector<float>(num + 1, num + 15)) pn = pn * t + c;
        const float u = t - 225.0f;
        float qd = u + den[0];
        for (int k = 1; k < 5; ++k) qd = qd * u + den[k];
        r = (pn / qd) * a;
    }
    return x < 0.0f ? -r : r;
}
// I_{3/2}(x) = sqrt(2/(pi x)) (cosh x - sinh x / x): closed form of the
// half-integer order used by WINDKB2A via RIBSL (besi1.f:793-800).  Small x by
// series to avoid cancellation.
double bessel_i15(double x) {
    if (x < 0.5) {
        // I_v(x) = sum (x/2)^{2k+v} / (k! Gamma(k+v+1)), v=3/2
        double h = 0.5 * x, sum = 0.0, term = std::pow(h, 1.5) / std::tgamma(2.5);
        for (int k = 0; k < 60; ++k) {
            sum += term;
            term *= h * h / (double(k + 1) * (k + 2.5));
        }
        return sum;
    }
    return std::sqrt(2.0 / (M_PI * x)) * (std::cosh(x) - std::sinh(x) / x);
}
// J_{3/2}(x) = sqrt(2/(pi x)) (sin x / x - cos x) (RJBSL, besi1.f:276).  The
// window never takes this branch inside the reconstruction sphere; kept so the
// restatement covers every branch of bp32f.f:984-1000.
double bessel_j15(double x) {
    return std::sqrt(2.0 / (M_PI * x)) * (std::sin(x) / x - std::cos(x));
}

template <typename R> struct Num;
template <> struct Num<float> {
    static long nint(float v) { return lroundf(v); }
    static float fabs_(float v) { return fabsf(v); }
    static float sqrt_(float v) { return sqrtf(v); }
    static float pow_(float a, float b) { return powf(a, b); }
    static float i1(float x) { return cody_i1_f32(x); }
    static float i15(float x) { return (float)bessel_i15((double)x); }
    static float j15(float x) { return (float)bessel_j15((double)x); }
    static float twopi() { return 2.0f * (float)M_PI; }          // REAL parameter (bp32f.f:950-951)
};
template <> struct Num<double> {
    static long nint(double v) { return lround(v); }
    static double fabs_(double v) { return fabs(v); }
    static double sqrt_(double v) { return sqrt(v); }
    static double pow_(double a, double b) { return pow(a, b); }
    static double i1(double x) { return bessel_i1(x); }
    static double i15(double x) { return bessel_i15(x); }
    static double j15(double x) { return bessel_j15(x); }
    static double twopi() { return 2.0 * M_PI; }
};

// ---------------------------------------------------------------- KB table
// bp32f.f:193-214 (== FILLBESSIL, fillbessi0.f:88-117):
//   V = (LN-1)/2/N2, ALPHA = 6.5, AAAA = 0.9 V, B0 = sqrt(ALPHA) I1(ALPHA),
//   FLTB = LTAB/(LN2+1), S = I/FLTB/N2,
//   TABI(I) = S <= AAAA ? sqrt(ALPHA X) I1(ALPHA X) / B0 : 0, X = sqrt(1-(S/AAAA)^2)
template <typename R>
void kb_table(int n2, R* tab, R* fltb_out) {
    typedef Num<R> M;
    const int LN = 5;
    R V = R(LN - 1) / R(2.0) / R(n2);
    R alpha = R(6.5);
    R aaaa = R(0.9) * V;
    R b0 = M::sqrt_(alpha) * M::i1(alpha);
    R fltb = R(LTAB) / R(LN2 + 1);
    for (int i = 0; i <= LTAB; ++i) {
        R s = R(i) / fltb / R(n2);
        if (s <= aaaa) {
            R q = s / aaaa;
            R x = M::sqrt_(R(1.0) - q * q);
            tab[i] = M::sqrt_(alpha * x) * M::i1(alpha * x) / b0;
        } else {
            tab[i] = R(0);
        }
    }
    if (fltb_out) *fltb_out = fltb;
}

// ---------------------------------------------------------------- CANG
// cang.f:44-82.  The degree->radian factor is a DOUBLE PRECISION parameter
// built from a default-real literal, i.e. float32(pi)/180 evaluated in double
// (cang.f:50-53).  cos/sin and products in double, stored as REAL.
void cang_ref(float phi, float theta, float psi, float* dm) {
    const double dgr = (double)(float)3.14159265358979323846 / 180.0;
    double cphi = cos((double)phi * dgr), sphi = sin((double)phi * dgr);
    double cthe = cos((double)theta * dgr), sthe = sin((double)theta * dgr);
    double cpsi = cos((double)psi * dgr), spsi = sin((double)psi * dgr);
    dm[0] = (float)(cphi * cthe * cpsi - sphi * spsi);
    dm[1] = (float)(sphi * cthe * cpsi + cphi * spsi);
    dm[2] = (float)(-sthe * cpsi);
    dm[3] = (float)(-cphi * cthe * spsi - sphi * cpsi);
    dm[4] = (float)(-sphi * cthe * spsi + cphi * cpsi);
    dm[5] = (float)(sthe * spsi);
    dm[6] = (float)(sthe * cphi);
    dm[7] = (float)(sthe * sphi);
    dm[8] = (float)(cthe);
}
// FP64 rotation; `dgr` = degree->radian factor: pi/180 (exact) or the
// reference's float32(pi)/180 (cang.f:52-53).
double deg2rad(int exact_pi) {
    return exact_pi ? M_PI / 180.0 : (double)(float)3.14159265358979323846 / 180.0;
}
void cang_f64(double phi, double theta, double psi, double* dm, double dgr = M_PI / 180.0) {
    double cphi = cos(phi * dgr), sphi = sin(phi * dgr);
    double cthe = cos(theta * dgr), sthe = sin(theta * dgr);
    double cpsi = cos(psi * dgr), spsi = sin(psi * dgr);
    dm[0] = cphi * cthe * cpsi - sphi * spsi;
    dm[1] = sphi * cthe * cpsi + cphi * spsi;
    dm[2] = -sthe * cpsi;
    dm[3] = -cphi * cthe * spsi - sphi * cpsi;
    dm[4] = -sphi * cthe * spsi + cphi * cpsi;
    dm[5] = sthe * spsi;
    dm[6] = sthe * cphi;
    dm[7] = sthe * sphi;
    dm[8] = cthe;
}

// ---------------------------------------------------------------- FFT
// Mixed-radix complex FFT (radix 4, 2, 3, 5, generic prime), double precision.
// Defined by the DFT: y[k] = sum_n x[n] exp(sign * 2 pi i k n / N).
struct FFTPlan {
    int n = 0, sign = -1;
    std::vector<int> radix;
    std::vector<double> tw;   // exp(sign 2 pi i k / n), k = 0..n-1, interleaved
};
FFTPlan make_plan(int n, int sign) {
    FFTPlan P;
    P.n = n; P.sign = sign;
    int m = n;
    while (m % 4 == 0) { P.radix.push_back(4); m /= 4; }
    while (m % 2 == 0) { P.radix.push_back(2); m /= 2; }
    for (int f = 3; m > 1; f += 2) while (m % f == 0) { P.radix.push_back(f); m /= f; }
    P.tw.resize(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        double a = sign * 2.0 * M_PI * (double)k / (double)n;
        P.tw[2 * k] = cos(a);
        P.tw[2 * k + 1] = sin(a);
    }
    return P;
}
void fft_rec(const FFTPlan& P, int level, int n, const double* x, long xs, double* y, long twst) {
    if (n == 1) { y[0] = x[0]; y[1] = x[1]; return; }
    const int p = P.radix[level];
    const int m = n / p;
    for (int r = 0; r < p; ++r) fft_rec(P, level + 1, m, x + 2 * r * xs, xs * p, y + 2 * (long)r * m, twst * p);
    const double* tw = P.tw.data();
    const long N = P.n;
    if (p == 2) {
        for (int k = 0; k < m; ++k) {
            double* a = y + 2 * k; double* b = y + 2 * (k + m);
            long e = (long)k * twst;
            double wr = tw[2 * e], wi = tw[2 * e + 1];
            double br = b[0] * wr - b[1] * wi, bi = b[0] * wi + b[1] * wr;
            double ar = a[0], ai = a[1];
            a[0] = ar + br; a[1] = ai + bi; b[0] = ar - br; b[1] = ai - bi;
        }
        return;
    }
    if (p == 4) {
        const double s = (double)P.sign;  // multiply by (sign * i)
        for (int k = 0; k < m; ++k) {
            double v[4][2];
            for (int r = 0; r < 4; ++r) {
                const double* src = y + 2 * ((long)r * m + k);
                long e = ((long)r * k) * twst % N;
                double wr = tw[2 * e], wi = tw[2 * e + 1];
                v[r][0] = src[0] * wr - src[1] * wi;
                v[r][1] = src[0] * wi + src[1] * wr;
            }
            double t0r = v[0][0] + v[2][0], t0i = v[0][1] + v[2][1];
            double t1r = v[0][0] - v[2][0], t1i = v[0][1] - v[2][1];
            double t2r = v[1][0] + v[3][0], t2i = v[1][1] + v[3][1];
            double t3r = v[1][0] - v[3][0], t3i = v[1][1] - v[3][1];
            // (sign*i) * t3
            double ur = -s * t3i, ui = s * t3r;
            double* o0 = y + 2 * (long)k; double* o1 = y + 2 * ((long)k + m);
            double* o2 = y + 2 * ((long)k + 2 * m); double* o3 = y + 2 * ((long)k + 3 * m);
            o0[0] = t0r + t2r; o0[1] = t0i + t2i;
            o2[0] = t0r - t2r; o2[1] = t0i - t2i;
            o1[0] = t1r + ur;  o1[1] = t1i + ui;
            o3[0] = t1r - ur;  o3[1] = t1i - ui;
        }
        return;
    }
    std::vector<double> t(2 * p), o(2 * p);
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) {
            const double* src = y + 2 * ((long)r * m + k);
            long e = ((long)r * k % n) * twst;
            double wr = tw[2 * e], wi = tw[2 * e + 1];
            t[2 * r] = src[0] * wr - src[1] * wi;
            t[2 * r + 1] = src[0] * wi + src[1] * wr;
        }
        for (int q = 0; q < p; ++q) {
            double sr = 0, si = 0;
            for (int r = 0; r < p; ++r) {
                long e = ((long)r * q % p) * m * twst;
                double wr = tw[2 * e], wi = tw[2 * e + 1];
                sr += t[2 * r] * wr - t[2 * r + 1] * wi;
                si += t[2 * r] * wi + t[2 * r + 1] * wr;
            }
            o[2 * q] = sr; o[2 * q + 1] = si;
        }
        for (int q = 0; q < p; ++q) {
            y[2 * ((long)k + (long)q * m)] = o[2 * q];
            y[2 * ((long)k + (long)q * m) + 1] = o[2 * q + 1];
        }
    }
}
// in-place transform of n complex values with element stride `stride` (complex units)
void fft_strided(const FFTPlan& P, double* data, long stride, double* work_in, double* work_out) {
    const int n = P.n;
    for (int i = 0; i < n; ++i) { work_in[2 * i] = data[2 * i * stride]; work_in[2 * i + 1] = data[2 * i * stride + 1]; }
    fft_rec(P, 0, n, work_in, 1, work_out, 1);
    for (int i = 0; i < n; ++i) { data[2 * i * stride] = work_out[2 * i]; data[2 * i * stride + 1] = work_out[2 * i + 1]; }
}

// 2D real-to-complex, FFTW forward sign (e^{-i}), unnormalised.
// in: N x N real (x fastest); out: (N/2+1) x N complex interleaved (kx fastest).
void fft2_r2c(const double* in, int N, double* out) {
    const int H = N / 2 + 1;
    FFTPlan P = make_plan(N, -1);
    std::vector<double> full(2 * (size_t)N * N);
    #pragma omp parallel
    {
        std::vector<double> a(2 * N), b(2 * N);
        #pragma omp for schedule(static)
        for (int y = 0; y < N; ++y) {
            for (int x = 0; x < N; ++x) { a[2 * x] = in[(size_t)y * N + x]; a[2 * x + 1] = 0.0; }
            fft_rec(P, 0, N, a.data(), 1, b.data(), 1);
            for (int x = 0; x < H; ++x) { full[2 * ((size_t)y * H + x)] = b[2 * x]; full[2 * ((size_t)y * H + x) + 1] = b[2 * x + 1]; }
        }
        #pragma omp for schedule(static)
        for (int x = 0; x < H; ++x) fft_strided(P, full.data() + 2 * x, H, a.data(), b.data());
    }
    std::memcpy(out, full.data(), sizeof(double) * 2 * (size_t)H * N);
}

// 3D complex-to-real, FFTW backward sign (e^{+i}), unnormalised, with FFTW's
// c2r semantics: result = Re sum_k c_kx X(k) e^{+2 pi i k.n/N}, c_0 = c_{N/2} = 1,
// else 2 (equivalent to Hermitian-symmetrising the kx = 0 and kx = N/2 planes).
// in: (N/2+1) x N x N complex (kx fastest); out: N x N x N real (x fastest).
void fft3_c2r(const double* in, int N, double* out) {
    const int H = N / 2 + 1;
    FFTPlan P = make_plan(N, +1);
    std::vector<double> a(in, in + 2 * (size_t)H * N * N);
    #pragma omp parallel
    {
        std::vector<double> w1(2 * N), w2(2 * N);
        // along z: for each (ky, kx)
        #pragma omp for schedule(static)
        for (long yx = 0; yx < (long)N * H; ++yx) fft_strided(P, a.data() + 2 * yx, (long)N * H, w1.data(), w2.data());
        // along y: for each (z, kx)
        #pragma omp for schedule(static)
        for (long zx = 0; zx < (long)N * H; ++zx) {
            long z = zx / H, x = zx % H;
            fft_strided(P, a.data() + 2 * (z * N * H + x), H, w1.data(), w2.data());
        }
        // along x: Hermitian row, real output
        #pragma omp for schedule(static)
        for (long zy = 0; zy < (long)N * N; ++zy) {
            const double* row = a.data() + 2 * zy * H;
            for (int k = 0; k < H; ++k) { w1[2 * k] = row[2 * k]; w1[2 * k + 1] = row[2 * k + 1]; }
            w1[1] = 0.0;
            if (N % 2 == 0) w1[2 * (N / 2) + 1] = 0.0;
            for (int k = 1; k < (N + 1) / 2; ++k) { w1[2 * (N - k)] = row[2 * k]; w1[2 * (N - k) + 1] = -row[2 * k + 1]; }
            fft_rec(P, 0, N, w1.data(), 1, w2.data(), 1);
            for (int x = 0; x < N; ++x) out[zy * N + x] = w2[2 * x];
        }
    }
}

// ---------------------------------------------------------------- PADD2
// bp3f.f:1307-1343 with ASTA_D (asta.f:61-92): subtract the FP64 mean of the
// pixels outside the circle of radius L/2 about (L/2+1, L/2+1) (1-based), then
// place the image at offset IP = (N-L)/2 + mod(L,2) of an N x N zero buffer.
double asta_mean(const float* img, int L) {
    float R = (float)(L / 2);
    float r2 = R * R;
    int nc = L / 2 + 1;
    double aba = 0.0, dklp = 0.0;
    for (int J = 1; J <= L; ++J) {
        float t = (float)(J - nc);
        float xx = t * t;
        for (int I = 1; I <= L; ++I) {
            t = (float)(I - nc);
            if (xx + t * t > r2) { aba += (double)img[(size_t)(J - 1) * L + (I - 1)]; dklp += 1.0; }
        }
    }
    return aba / (double)(float)dklp;
}
template <typename R>
void pad_image(const float* img, int L, int N, R* out) {
    double qs = asta_mean(img, L);
    std::fill(out, out + (size_t)N * N, R(0));
    int ip = (N - L) / 2 + (L % 2);
    for (int j = 0; j < L; ++j)
        for (int i = 0; i < L; ++i)
            out[(size_t)(ip + j) * N + ip + i] = (R)((double)img[(size_t)j * L + i] - qs);
}

// ---------------------------------------------------------------- spectrum
// PADD2 -> FMRS_2 forward (FFTW r2c, then imag = -imag: fftw3.f:409-419) ->
// checkerboard BI(I,J) *= (-1)^(I+J+1), J 1-based (bp32f.f:859-864).
// Output BI(0:NX, 0:N-1) complex, kx fastest.
template <typename R>
void spectrum(const float* img, int nx, R* bi) {
    const int N = 2 * nx, H = nx + 1;
    std::vector<R> pad((size_t)N * N);
    pad_image<R>(img, nx, N, pad.data());
    std::vector<double> pd(pad.begin(), pad.end()), sp(2 * (size_t)H * N);
    fft2_r2c(pd.data(), N, sp.data());
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < H; ++i) {
            size_t o = (size_t)j * H + i;
            double s = ((i + j) & 1) ? -1.0 : 1.0;
            bi[2 * o] = (R)(sp[2 * o] * s);
            bi[2 * o + 1] = (R)(-sp[2 * o + 1] * s);
        }
}

// ---------------------------------------------------------------- ONELINE
// oneline.f:38-158, argument names as in the caller: N = padded size (N2 = 2 NX),
// nx = image size.  X(0:NX,N,N) complex, W(0:NX,N,N) real, BI(0:NX,N) complex,
// dm = first six entries of the CANG matrix.
// A: accumulation type (R for the restatement; double with R = float gives
// the reference's geometry and per-tap terms summed in FP64).
// V: type of the spectrum values (R unless given; V = double with R = float
// forms every term w * v in FP64 from the fp32 weight).
template <typename R, typename A = R, typename V = R>
void oneline(int J, int N, int nx, A* X, A* W, const V* BI, const R* dm, const R* tab, R fltb) {
    typedef Num<R> M;
    const int H = nx + 1;
    const int jp = (J >= 0) ? J : N + J;
    for (int I = 0; I <= nx; ++I) {
        if (!(((I * I + J * J) < (N * N / 4)) && !(I == 0 && J < 0))) continue;
        R x = R(I) * dm[0] + R(J) * dm[3];
        R y = R(I) * dm[1] + R(J) * dm[4];
        R z = R(I) * dm[2] + R(J) * dm[5];
        V vr = BI[2 * ((size_t)jp * H + I)], vi = BI[2 * ((size_t)jp * H + I) + 1];
        if (x < R(0)) { x = -x; y = -y; z = -z; vi = -vi; }
        const int ix = (int)(x + R(0.5) + R(N)) - N;
        const int iy = (int)(y + R(0.5) + R(N)) - N;
        const int iz = (int)(z + R(0.5) + R(N)) - N;
        if (!(ix <= nx - LN2 - 1 && iy >= -nx + 2 + LN2 && iy <= nx - LN2 - 1 &&
              iz >= -nx + 2 + LN2 && iz <= nx - LN2 - 1)) continue;
        if (ix >= 0) {
            const int lb = -std::min(ix, LN2);
            for (int lz = -LN2; lz <= LN2; ++lz) {
                const int izp = iz + lz;
                const int iza = izp >= 0 ? izp : N + izp;
                const R tz = tab[M::nint(M::fabs_(z - R(izp)) * fltb)];
                if (tz == R(0)) continue;
                for (int ly = -LN2; ly <= LN2; ++ly) {
                    const int iyp = iy + ly;
                    const int iya = iyp >= 0 ? iyp : N + iyp;
                    const R ty = tab[M::nint(M::fabs_(y - R(iyp)) * fltb)] * tz;
                    if (ty == R(0)) continue;
                    for (int ixp = lb + ix; ixp <= LN2 + ix; ++ixp) {
                        const R wg = tab[M::nint(M::fabs_(x - R(ixp)) * fltb)] * ty;
                        if (wg != R(0)) {
                            const size_t o = ((size_t)iza * N + iya) * H + ixp;
                            X[2 * o] = X[2 * o] + (A)(vr * wg);
                            X[2 * o + 1] = X[2 * o + 1] + (A)(vi * wg);
                            W[o] = W[o] + (A)wg;
                        }
                    }
                }
            }
        }
        // reflected points (oneline.f:118-152)
        if (ix < LN2) {
            for (int lz = -LN2; lz <= LN2; ++lz) {
                const int izp = iz + lz;
                int izt = -izp;
                if (izp > 0) izt = N + izt;
                const R tz = tab[M::nint(M::fabs_(z - R(izp)) * fltb)];
                if (tz == R(0)) continue;
                for (int ly = -LN2; ly <= LN2; ++ly) {
                    const int iyp = iy + ly;
                    int iyt = -iyp;
                    if (iyp > 0) iyt = iyt + N;
                    const R ty = tab[M::nint(M::fabs_(y - R(iyp)) * fltb)] * tz;
                    if (ty == R(0)) continue;
                    for (int ixp = ix - LN2; ixp <= -1; ++ixp) {
                        const R wg = tab[M::nint(M::fabs_(x - R(ixp)) * fltb)] * ty;
                        if (wg != R(0)) {
                            const size_t o = ((size_t)izt * N + iyt) * H + (-ixp);
                            X[2 * o] = X[2 * o] + (A)(vr * wg);
                            X[2 * o + 1] = X[2 * o + 1] + (A)((-vi) * wg);
                            W[o] = W[o] + (A)wg;
                        }
                    }
                }
            }
        }
    }
}
template <typename R>
void insert_projection(int nx, R* X, R* W, const R* bi, const R* dm, const R* tab, R fltb) {
    const int N = 2 * nx;
    for (int J = -nx + 1; J <= nx; ++J) oneline<R>(J, N, nx, X, W, bi, dm, tab, fltb);
}

// ---------------------------------------------------------------- gather (adjoint of ONELINE)
// Transpose of oneline<float, double, double> (complex numbers as pairs of
// reals): ONELINE adds w v at each direct tap t and w conj(v) at -t for each
// reflected tap, v = BI(I, J) conjugated for a Friedel-flipped sample, so
//   out(I, J) = c( sum_direct w X[t] + conj(sum_reflected w X[-t]) ),
// c = conj for a flipped sample.  Same taps, bounds and fp32 weights as
// oneline.f:53-152 (restated in oneline() above); sums in FP64.
void gather_line(int J, int N, int nx, const double* X, double* out, const float* dm, const float* tab, float fltb) {
    typedef Num<float> M;
    const int H = nx + 1;
    const int jp = (J >= 0) ? J : N + J;
    for (int I = 0; I <= nx; ++I) {
        double* o = out + 2 * ((size_t)jp * H + I);
        o[0] = o[1] = 0.0;
        if (!(((I * I + J * J) < (N * N / 4)) && !(I == 0 && J < 0))) continue;
        float x = float(I) * dm[0] + float(J) * dm[3];
        float y = float(I) * dm[1] + float(J) * dm[4];
        float z = float(I) * dm[2] + float(J) * dm[5];
        const bool flip = x < 0.0f;
        if (flip) { x = -x; y = -y; z = -z; }
        const int ix = (int)(x + 0.5f + float(N)) - N;
        const int iy = (int)(y + 0.5f + float(N)) - N;
        const int iz = (int)(z + 0.5f + float(N)) - N;
        if (!(ix <= nx - LN2 - 1 && iy >= -nx + 2 + LN2 && iy <= nx - LN2 - 1 &&
              iz >= -nx + 2 + LN2 && iz <= nx - LN2 - 1)) continue;
        double dre = 0, dim = 0, mre = 0, mim = 0;
        if (ix >= 0) {
            const int lb = -std::min(ix, LN2);
            for (int lz = -LN2; lz <= LN2; ++lz) {
                const int izp = iz + lz;
                const int iza = izp >= 0 ? izp : N + izp;
                const float tz = tab[M::nint(M::fabs_(z - float(izp)) * fltb)];
                if (tz == 0.0f) continue;
                for (int ly = -LN2; ly <= LN2; ++ly) {
                    const int iyp = iy + ly;
                    const int iya = iyp >= 0 ? iyp : N + iyp;
                    const float ty = tab[M::nint(M::fabs_(y - float(iyp)) * fltb)] * tz;
                    if (ty == 0.0f) continue;
                    for (int ixp = lb + ix; ixp <= LN2 + ix; ++ixp) {
                        const float wg = tab[M::nint(M::fabs_(x - float(ixp)) * fltb)] * ty;
                        if (wg != 0.0f) {
                            const size_t t = ((size_t)iza * N + iya) * H + ixp;
                            dre += (double)wg * X[2 * t];
                            dim += (double)wg * X[2 * t + 1];
                        }
                    }
                }
            }
        }
        if (ix < LN2) {
            for (int lz = -LN2; lz <= LN2; ++lz) {
                const int izp = iz + lz;
                int izt = -izp;
                if (izp > 0) izt = N + izt;
                const float tz = tab[M::nint(M::fabs_(z - float(izp)) * fltb)];
                if (tz == 0.0f) continue;
                for (int ly = -LN2; ly <= LN2; ++ly) {
                    const int iyp = iy + ly;
                    int iyt = -iyp;
                    if (iyp > 0) iyt = iyt + N;
                    const float ty = tab[M::nint(M::fabs_(y - float(iyp)) * fltb)] * tz;
                    if (ty == 0.0f) continue;
                    for (int ixp = ix - LN2; ixp <= -1; ++ixp) {
                        const float wg = tab[M::nint(M::fabs_(x - float(ixp)) * fltb)] * ty;
                        if (wg != 0.0f) {
                            const size_t t = ((size_t)izt * N + iyt) * H + (-ixp);
                            mre += (double)wg * X[2 * t];
                            mim += (double)wg * X[2 * t + 1];
                        }
                    }
                }
            }
        }
        const double im = dim - mim;
        o[0] = dre + mre;
        o[1] = flip ? -im : im;
    }
}

// ---------------------------------------------------------------- SYMPLANE0
// bp32f.f:1051-1088 (argument N2 there is NX).  0-based: mirror of index a is N-a.
template <typename R>
void symplane0(int nx, R* X, R* W) {
    const int N = 2 * nx, H = nx + 1;
    auto at = [&](int iy, int iz) { return ((size_t)iz * N + iy) * H; };
    auto addc = [&](size_t a, size_t b) {   // X(a) += conj(X(b)); W(a) += W(b); X(b) = conj(X(a)); W(b) = W(a)
        X[2 * a] = X[2 * a] + X[2 * b];
        X[2 * a + 1] = X[2 * a + 1] + (-X[2 * b + 1]);
        W[a] = W[a] + W[b];
        X[2 * b] = X[2 * a];
        X[2 * b + 1] = -X[2 * a + 1];
        W[b] = W[a];
    };
    for (int iz = 1; iz <= nx - 1; ++iz)
        for (int iy = 1; iy <= nx - 1; ++iy) {
            addc(at(iy, iz), at(N - iy, N - iz));
            addc(at(N - iy, iz), at(iy, N - iz));
        }
    for (int iy = 1; iy <= nx - 1; ++iy) addc(at(iy, 0), at(N - iy, 0));
    for (int iz = 1; iz <= nx - 1; ++iz) addc(at(0, iz), at(0, N - iz));
}

// ---------------------------------------------------------------- NRMW2 (weighting part)
// bp3f.f:1355-1366: X = W > 0.1 ? X (-1)^(i+j+k) / W : 0 (0-based i,j,k here).
template <typename R>
void normalize(int nx, R* X, const R* W) {
    const int N = 2 * nx, H = nx + 1;
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < H; ++i) {
                size_t o = ((size_t)k * N + j) * H + i;
                if (W[o] > (R)0.1f) {
                    R s = ((i + j + k) & 1) ? R(-1) : R(1);
                    X[2 * o] = (X[2 * o] * s) / W[o];
                    X[2 * o + 1] = (X[2 * o + 1] * s) / W[o];
                } else {
                    X[2 * o] = R(0);
                    X[2 * o + 1] = R(0);
                }
            }
}

// NRMW2 inverse part: FMRS_3 INV=-1 -> conj, c2r, x 1/N^3 (fftw3.f:430-451).
template <typename R>
void inverse3(int nx, const R* X, R* real) {
    const int N = 2 * nx, H = nx + 1;
    const size_t nh = (size_t)H * N * N, nr = (size_t)N * N * N;
    std::vector<double> a(2 * nh), o(nr);
    for (size_t q = 0; q < nh; ++q) { a[2 * q] = (double)X[2 * q]; a[2 * q + 1] = -(double)X[2 * q + 1]; }
    fft3_c2r(a.data(), N, o.data());
    if (sizeof(R) == sizeof(float)) {
        const float pix = 1.0f / (float)((long)N * N * N);
        for (size_t q = 0; q < nr; ++q) real[q] = (R)((float)o[q] * pix);
    } else {
        const double pix = 1.0 / ((double)N * N * N);
        for (size_t q = 0; q < nr; ++q) real[q] = (R)(o[q] * pix);
    }
}

// ---------------------------------------------------------------- WINDKB2A
// bp32f.f:945-1045: crop the central L^3 at IP = (N-L)/2 + mod(L,2); inside the
// sphere LR <= (L/2)^2 divide by the KB window transform
//   WKB = ALPHA^1.5 / I_1.5(ALPHA) * I_1.5(ART) / ART^1.5, ART = sqrt(|QRT LR - ALPHA^2|),
//   QRT = (2 pi AAAA)^2;
// subtract the mean over the shell (L/2-1)^2 <= LR <= (L/2)^2; zero outside.
template <typename R>
void window(int nx, const R* real, R* out) {
    typedef Num<R> M;
    const int L = nx, N = 2 * nx;
    const int ip = (N - L) / 2 + (L % 2);
    for (int k = 0; k < L; ++k)
        for (int j = 0; j < L; ++j)
            for (int i = 0; i < L; ++i)
                out[((size_t)k * L + j) * L + i] = real[((size_t)(ip + k) * N + (ip + j)) * N + (ip + i)];
    const int L2 = (L / 2) * (L / 2);
    const int L2P = (L / 2 - 1) * (L / 2 - 1);
    const int c = L / 2;   // 0-based centre (1-based L/2+1)
    const R alpha = R(6.5);
    const R xnu = R(3) / R(2.0);
    const R V = R(4) / R(2.0) / R(N);
    const R aaaa = R(0.9) * V;
    R ri = M::i15(alpha);
    const R wkb0 = M::pow_(alpha, xnu) / ri;
    const R tq = M::twopi() * aaaa;
    const R qrt = tq * tq;
    R tnr = R(0);
    long m = 0;
    for (int k = 0; k < L; ++k) {
        const int kk = (k - c) * (k - c);
        for (int j = 0; j < L; ++j) {
            const int jj = (j - c) * (j - c) + kk;
            for (int i = 0; i < L; ++i) {
                const int lr = jj + (i - c) * (i - c);
                if (lr > L2) continue;
                const R sigma = qrt * R(lr) - alpha * alpha;
                R wkb;
                if (M::fabs_(sigma) < R(1.0e-7)) {
                    wkb = R(1.0);
                } else if (sigma > R(0)) {
                    const R art = M::sqrt_(sigma);
                    wkb = wkb0 * M::j15(art) / M::pow_(art, xnu);
                } else {
                    const R art = M::sqrt_(M::fabs_(sigma));
                    wkb = wkb0 * M::i15(art) / M::pow_(art, xnu);
                }
                R& v = out[((size_t)k * L + j) * L + i];
                v = v / M::fabs_(wkb);
                if (lr >= L2P && lr <= L2) { tnr = tnr + v; ++m; }
            }
        }
    }
    tnr = tnr / R(m);
    for (int k = 0; k < L; ++k) {
        const int kk = (k - c) * (k - c);
        for (int j = 0; j < L; ++j) {
            const int jj = (j - c) * (j - c) + kk;
            for (int i = 0; i < L; ++i) {
                const int lr = jj + (i - c) * (i - c);
                R& v = out[((size_t)k * L + j) * L + i];
                v = (lr <= L2) ? v - tnr : R(0);
            }
        }
    }
}

template <typename R>
void finalize_volume(int nx, R* X, const R* W, R* out) {
    const int N = 2 * nx;
    std::vector<R> real((size_t)N * N * N);
    normalize<R>(nx, X, W);
    inverse3<R>(nx, X, real.data());
    window<R>(nx, real.data(), out);
}

// ---------------------------------------------------------------- symmetries
// DMS = MATMUL(SM, DM) (bp32f.f:866-872): 3x3 matrices stored as Fortran
// DM(3,3), i.e. element (i, j) at [i + 3 j]; products and sums in R, k = 1..3.
template <typename R>
void sym_matmul(const R* sm, const R* dm, R* out) {
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            R acc = sm[i] * dm[3 * j];
            acc = acc + sm[i + 3] * dm[1 + 3 * j];
            acc = acc + sm[i + 6] * dm[2 + 3 * j];
            out[i + 3 * j] = acc;
        }
}

// TRIQUAD (triquad.f): 27-point triquadratic interpolation at offsets
// (R, S, T) in [0, 1) from the point F(14).  The reference's implicit typing
// makes the products RS, ST, RT, RST and the factors 1-R^2, 1-R, 1+R, ...
// REAL (rounded from DOUBLE), while a term that multiplies by R, S or T
// itself is evaluated in DOUBLE; the sum runs left to right, in REAL until
// the first DOUBLE term, and the result is rounded to REAL.
float triquad_ref(double R, double S, double T, const float* F) {
    const float C2 = 1.0f / 2.0f, C4 = 1.0f / 4.0f, C8 = 1.0f / 8.0f;
    const float RS = (float)(R * S), ST = (float)(S * T), RT = (float)(R * T);
    const float RST = (float)(R * (double)ST);
    const float RSQ = (float)(1.0 - R * R), SSQ = (float)(1.0 - S * S), TSQ = (float)(1.0 - T * T);
    const float RM1 = (float)(1.0 - R), SM1 = (float)(1.0 - S), TM1 = (float)(1.0 - T);
    const float RP1 = (float)(1.0 + R), SP1 = (float)(1.0 + S), TP1 = (float)(1.0 + T);
    // terms 1-4 are REAL; term 5 is the first DOUBLE one
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

// SYMVOL (symvol.f:39-219, the non-SP_MP branch): Q2 = (1/NSYM) sum over
// symmetry matrices RM of Q1 sampled at RM^T r by TRIQUAD, for voxels
// r = (IX, IY, IZ) in [KLX, KNX]^3 (KLX = -NX/2, KNX = NX/2 - 1 for even NX,
// NX/2 for odd) with |r|^2 <= IRADI = (min |bound| - 1)^2; other voxels and
// identity matrices (RM(1,1) = RM(2,2) = RM(3,3) = 1) take Q1(r).  Volumes
// x fastest; sm: nsym x 9 Fortran-ordered matrices.
void symvol_ref(int nx, const float* q1, float* q2, const float* sm, int nsym) {
    const int klx = -nx / 2, knx = (nx % 2 == 0) ? nx / 2 - 1 : nx / 2;
    const int mn = std::min(std::abs(klx), std::abs(knx));
    const int iradi = (mn - 1) * (mn - 1);
    auto at = [&](int ix, int iy, int iz) { return ((size_t)(iz - klx) * nx + (iy - klx)) * nx + (ix - klx); };
    const size_t nv = (size_t)nx * nx * nx;
    std::fill(q2, q2 + nv, 0.0f);
    for (int s = 0; s < nsym; ++s) {
        const float* rm = sm + 9 * s;
        const bool ident = rm[0] == 1.0f && rm[4] == 1.0f && rm[8] == 1.0f;
        for (int iz = klx; iz <= knx; ++iz)
            for (int iy = klx; iy <= knx; ++iy) {
                double qrt[3];
                for (int k = 0; k < 3; ++k) qrt[k] = (double)(rm[3 * k + 1] * (float)iy + rm[3 * k + 2] * (float)iz);
                for (int ix = klx; ix <= knx; ++ix) {
                    float v;
                    const int icrd = ix * ix + iy * iy + iz * iz;
                    if (ident || icrd > iradi) {
                        v = q1[at(ix, iy, iz)];
                    } else {
                        double qr[3];
                        for (int k = 0; k < 3; ++k) qr[k] = qrt[k] + (double)(rm[3 * k] * (float)ix);
                        const int iox = (int)std::floor(qr[0]), ioy = (int)std::floor(qr[1]), ioz = (int)std::floor(qr[2]);
                        float F[27];
                        for (int l = 0; l < 27; ++l)
                            F[l] = q1[at(iox + (l % 3) - 1, ioy + (l / 3) % 3 - 1, ioz + l / 9 - 1)];
                        v = triquad_ref(qr[0] - iox, qr[1] - ioy, qr[2] - ioz, F);
                    }
                    q2[at(ix, iy, iz)] = q2[at(ix, iy, iz)] + v;
                }
            }
    }
    const float fn = (float)nsym;
    for (size_t q = 0; q < nv; ++q) q2[q] = q2[q] / fn;
}

// ---------------------------------------------------------------- RB 3F pre-rotation
// QUADRI (quadri.f:83-150): quadratic interpolation of FDATA(NX, NY) (x fastest,
// 1-based coordinates) at (XX, YY) with wrap-around, in REAL arithmetic and
// the reference's order of operations.
float quadri_ref(float xx, float yy, int nx, int ny, const float* f) {
    float x = xx, y = yy;
    if (x < 1.0f) x = x + (float)((1 - (int)x / nx) * nx);
    if (x > (float)nx + 0.5f) x = std::fmod(x - 1.0f, (float)nx) + 1.0f;
    if (y < 1.0f) y = y + (float)((1 - (int)y / ny) * ny);
    if (y > (float)ny + 0.5f) y = std::fmod(y - 1.0f, (float)ny) + 1.0f;
    const int i = (int)x, j = (int)y;
    const float dx0 = x - (float)i, dy0 = y - (float)j;
    int ip1 = i + 1, im1 = i - 1, jp1 = j + 1, jm1 = j - 1;
    if (ip1 > nx) ip1 -= nx;
    if (im1 < 1) im1 += nx;
    if (jp1 > ny) jp1 -= ny;
    if (jm1 < 1) jm1 += ny;
    auto F = [&](int a, int b) { return f[(size_t)(b - 1) * nx + (a - 1)]; };
    const float f0 = F(i, j);
    const float c1 = F(ip1, j) - f0;
    const float c2 = (c1 - f0 + F(im1, j)) * 0.5f;
    const float c3 = F(i, jp1) - f0;
    const float c4 = (c3 - f0 + F(i, jm1)) * 0.5f;
    const float dxb = dx0 - 1.0f, dyb = dy0 - 1.0f;
    const float hxc = std::signbit(dx0) ? -1.0f : 1.0f;   // INT(SIGN(1.0, DX0)) stored as REAL
    const float hyc = std::signbit(dy0) ? -1.0f : 1.0f;
    int ic = (int)((float)i + hxc), jc = (int)((float)j + hyc);
    if (ic > nx) ic -= nx;
    else if (ic < 1) ic += nx;
    if (jc > ny) jc -= ny;
    else if (jc < 1) jc += ny;
    const float c5 = ((((((F(ic, jc) - f0) - hxc * c1) - (hxc * (hxc - 1.0f)) * c2) - hyc * c3) -
                       (hyc * (hyc - 1.0f)) * c4) * (hxc * hyc));
    return (f0 + dx0 * ((c1 + dxb * c2) + dy0 * c5)) + dy0 * (c3 + dyb * c4);
}

// RTSQ (rtsq.f:56-150) with NXP = NX, NYP = NY: rotate by THETA degrees, scale
// by SCLI, shift by (SHXI, SHYI), sampling with QUADRI; REAL arithmetic.  The
// angle's cosine and sine are single-precision library values.
void rtsq_ref(const float* img, float* out, int nx, int ny, float theta, float scli, float shxi, float shyi) {
    const float dgr_to_rad = (float)3.14159265358979323846 / 180.0f;
    const float shx = std::fmod(shxi, (float)nx), shy = std::fmod(shyi, (float)ny);
    const int iycen = ny / 2 + 1, ixcen = nx / 2 + 1;
    const float rx1 = (float)(-(nx / 2)), ry1 = (float)(-(ny / 2));
    float rx2 = (float)(nx / 2), ry2 = (float)(ny / 2), rxe1, rxe2, rye1, rye2;
    if (nx % 2 == 0) { rx2 = rx2 - 1.0f; rxe1 = (float)(-nx); rxe2 = (float)nx; }
    else { rxe1 = (float)(-nx - 1); rxe2 = (float)(nx + 1); }
    if (ny % 2 == 0) { ry2 = ry2 - 1.0f; rye1 = (float)(-ny); rye2 = (float)ny; }
    else { rye1 = (float)(-ny - 1); rye2 = (float)(ny + 1); }
    const float ang = theta * dgr_to_rad;
    const float cod = std::cos(ang), sid = std::sin(ang);
    const float coddscli = cod / scli, siddscli = sid / scli;
    const float fixcenmshx = (float)(-ixcen) - shx, fiycenmshy = (float)(-iycen) - shy;
    for (int iy = 1; iy <= ny; ++iy) {
        float yi = (float)iy + fiycenmshy;
        if (yi < ry1) yi = std::min(yi + rye2, ry2);
        if (yi > ry2) yi = std::max(yi + rye1, ry1);
        const float ycod = yi * coddscli + (float)iycen;
        const float ysid = (-yi) * siddscli + (float)ixcen;
        for (int ix = 1; ix <= nx; ++ix) {
            float xi = (float)ix + fixcenmshx;
            if (xi < rx1) xi = std::min(xi + rxe2, rx2);
            if (xi > rx2) xi = std::max(xi + rxe1, rx1);
            const float yold = xi * siddscli + ycod;
            const float xold = xi * coddscli + ysid;
            out[(size_t)(iy - 1) * nx + (ix - 1)] = quadri_ref(xold, yold, nx, ny, img);
        }
    }
}

// ---------------------------------------------------------------- RB with FBS ON: RTSF
// FBS2 (fbs2.f:95-160): Fourier-based bicubic interpolation of FDATA(NXP, NY)
// at (X, Y) (1-based) from the image and its derivatives X1 = dF/dx,
// Y1 = dF/dy, XY2 = d2F/dxdy (arrays (NXLD, NY), x fastest); cubic Hermite
// along x on the two rows around Y, then along y.  CHKBOUND wraps the cell's
// corners with MODULO.  REAL arithmetic in the reference's order.
int imodulo(int a, int p) { const int r = a % p; return r < 0 ? r + p : r; }
float fbs2_ref(float x, float y, int nxld, int nx, int ny, const float* fd, int nxp, const float* x1,
               const float* y1, const float* xy2, bool chk) {
    const int i = (int)std::floor(x), j = (int)std::floor(y);
    const float dx = x - (float)i, dy = y - (float)j;
    const float dxsqr = dx * dx, dxcub = dx * dx * dx;
    int i2 = i, j2 = j, i3 = i + 1, j3 = j + 1;
    if (chk) {
        i2 = imodulo(i - 1, nx) + 1;
        j2 = imodulo(j - 1, ny) + 1;
        i3 = imodulo(i, nx) + 1;
        j3 = imodulo(j, ny) + 1;
    }
    auto F = [&](int a, int b) { return fd[(size_t)(b - 1) * nxp + (a - 1)]; };
    auto D = [&](const float* d, int a, int b) { return d[(size_t)(b - 1) * nxld + (a - 1)]; };
    // A0 + A1 DX + A2 DX^2 + A3 DX^3 with A2, A3 from the values and slopes at both ends
    auto hermite = [&](float a0, float a1, float f3, float d3) {
        const float a2 = (3.0f * (f3 - a0) - 2.0f * a1) - d3;
        const float a3 = (2.0f * ((-f3) + a0) + a1) + d3;
        return ((a0 + a1 * dx) + a2 * dxsqr) + a3 * dxcub;
    };
    const float adx = hermite(F(i2, j2), D(x1, i2, j2), F(i3, j2), D(x1, i3, j2));
    const float bdx = hermite(F(i2, j3), D(x1, i2, j3), F(i3, j3), D(x1, i3, j3));
    const float dadx = hermite(D(y1, i2, j2), D(xy2, i2, j2), D(y1, i3, j2), D(xy2, i3, j2));
    const float dbdx = hermite(D(y1, i2, j3), D(xy2, i2, j3), D(y1, i3, j3), D(xy2, i3, j3));
    const float a2 = (3.0f * (bdx - adx) - 2.0f * dadx) - dbdx;
    const float a3 = (2.0f * (adx - bdx) + dadx) + dbdx;
    return ((adx + dadx * dy) + a2 * (dy * dy)) + a3 * (dy * dy * dy);
}

// FBS2_PREP (fbs2.f:180-245): the derivatives as inverse transforms of the
// spectrum times i kx, i ky and (i kx)(i ky), kx = 2 pi I / NX for the stored
// half spectrum I = 0 .. NX/2 and ky = 2 pi J / NY (J <= NY/2, else J - NY),
// the factors in REAL as the reference forms them.  The inverse follows
// FFTW's complex-to-real reading of the half spectrum: the kx = 0 and
// kx = NX/2 columns count by their Hermitian parts.  Transforms in FP64,
// results rounded to fp32.  Arrays NX x NY, x fastest.
void fbs2_prep_ref(const float* img, int nx, int ny, float* x1, float* y1, float* xy2) {
    const int h = nx / 2 + 1;
    const FFTPlan px = make_plan(nx, -1), py = make_plan(ny, -1);
    const FFTPlan qx = make_plan(nx, 1), qy = make_plan(ny, 1);
    std::vector<double> a(2 * (size_t)nx * ny), wi(2 * (size_t)std::max(nx, ny)), wo(2 * (size_t)std::max(nx, ny));
    for (size_t q = 0; q < (size_t)nx * ny; ++q) { a[2 * q] = img[q]; a[2 * q + 1] = 0.0; }
    for (int y = 0; y < ny; ++y) fft_strided(px, a.data() + 2 * (size_t)y * nx, 1, wi.data(), wo.data());
    for (int x = 0; x < nx; ++x) fft_strided(py, a.data() + 2 * x, nx, wi.data(), wo.data());
    const float pi2 = (float)6.28318530717958647692;
    const float ax = pi2 / (float)nx, ay = pi2 / (float)ny;
    auto wxf = [&](int k) { return (double)((float)k * ax); };
    auto wyf = [&](int k) { return (double)((float)(k <= ny / 2 ? k : k - ny) * ay); };
    float* outs[3] = {x1, y1, xy2};
    for (int d = 0; d < 3; ++d) {
        // half-spectrum multipliers: i wx, i wy, -wx wy (FFTW sign: the
        // reference's conjugations cancel on the way back)
        std::vector<double> g(2 * (size_t)h * ny);
        for (int ky = 0; ky < ny; ++ky)
            for (int kx = 0; kx < h; ++kx) {
                const double sr = a[2 * ((size_t)ky * nx + kx)], si = a[2 * ((size_t)ky * nx + kx) + 1];
                double mr = 0.0, mi = 0.0;
                if (d == 0) mi = wxf(kx);
                else if (d == 1) mi = wyf(ky);
                else mr = -wxf(kx) * wyf(ky);
                g[2 * ((size_t)ky * h + kx)] = sr * mr - si * mi;
                g[2 * ((size_t)ky * h + kx) + 1] = sr * mi + si * mr;
            }
        for (int kx = 0; kx < h; ++kx) {
            if (!(kx == 0 || 2 * kx == nx)) continue;
            std::vector<double> col(2 * (size_t)ny);
            for (int ky = 0; ky < ny; ++ky) {
                const int km = (ny - ky) % ny;
                col[2 * ky] = 0.5 * (g[2 * ((size_t)ky * h + kx)] + g[2 * ((size_t)km * h + kx)]);
                col[2 * ky + 1] = 0.5 * (g[2 * ((size_t)ky * h + kx) + 1] - g[2 * ((size_t)km * h + kx) + 1]);
            }
            for (int ky = 0; ky < ny; ++ky) {
                g[2 * ((size_t)ky * h + kx)] = col[2 * ky];
                g[2 * ((size_t)ky * h + kx) + 1] = col[2 * ky + 1];
            }
        }
        // full Hermitian spectrum, then the inverse 2D DFT (real part), / (NX NY)
        std::vector<double> f(2 * (size_t)nx * ny);
        for (int ky = 0; ky < ny; ++ky)
            for (int kx = 0; kx < nx; ++kx) {
                double re, im;
                if (kx < h) {
                    re = g[2 * ((size_t)ky * h + kx)];
                    im = g[2 * ((size_t)ky * h + kx) + 1];
                } else {
                    const int km = (ny - ky) % ny;
                    re = g[2 * ((size_t)km * h + (nx - kx))];
                    im = -g[2 * ((size_t)km * h + (nx - kx)) + 1];
                }
                f[2 * ((size_t)ky * nx + kx)] = re;
                f[2 * ((size_t)ky * nx + kx) + 1] = im;
            }
        for (int y = 0; y < ny; ++y) fft_strided(qx, f.data() + 2 * (size_t)y * nx, 1, wi.data(), wo.data());
        for (int x = 0; x < nx; ++x) fft_strided(qy, f.data() + 2 * x, nx, wi.data(), wo.data());
        const double pix = 1.0 / ((double)nx * ny);
        for (size_t q = 0; q < (size_t)nx * ny; ++q) outs[d][q] = (float)(f[2 * q] * pix);
    }
}

// MODULO(A, P) for REAL A, P > 0: A - FLOOR(A / P) P, from the exact remainder.
float rmodulo(float a, float p) {
    float r = std::fmod(a, p);
    if (r != 0.0f && (r < 0.0f) != (p < 0.0f)) r = r + p;
    return r;
}

// RTSF (rtsf.f:106-205, scale 1): rotate by THETA degrees and shift by
// (SHXI, SHYI) with FBS2 interpolation; the shift is taken MODULO the image
// size and applied by wrapping the output coordinates.  REAL arithmetic; the
// angle's cosine and sine are single-precision library values.
void rtsf_ref(const float* img, float* out, int nx, int ny, float theta, float shxi, float shyi) {
    std::vector<float> x1((size_t)nx * ny), y1((size_t)nx * ny), xy2((size_t)nx * ny);
    fbs2_prep_ref(img, nx, ny, x1.data(), y1.data(), xy2.data());
    const float pi = (float)3.14159265358979323846;
    const float th = theta * pi / 180.0f;
    const float cth = std::cos(th), sth = std::sin(th);
    const int cx = nx / 2 + 1, cy = ny / 2 + 1;
    const float shx = rmodulo(shxi, (float)nx), shy = rmodulo(shyi, (float)ny);
    const float fy0 = (-shy) - (float)cy, fy1 = ((-shy) + (float)ny) - (float)cy, fy2 = ((-shy) - (float)ny) - (float)cy;
    const float fx0 = (-shx) - (float)cx, fx1 = ((-shx) + (float)nx) - (float)cx, fx2 = ((-shx) - (float)nx) - (float)cx;
    const float shypny = shy + (float)ny, shxpnx = shx + (float)nx;
    for (int iy = 1; iy <= ny; ++iy) {
        float fy = (float)iy + fy0;
        if ((float)(iy - 1) < shy) fy = (float)iy + fy1;
        if ((float)(iy - 1) >= shypny) fy = (float)iy + fy2;
        const float ycod = cth * fy + (float)cy;
        const float ysid = -(sth * fy) + (float)cx;
        for (int ix = 1; ix <= nx; ++ix) {
            float fx = (float)ix + fx0;
            if ((float)(ix - 1) < shx) fx = (float)ix + fx1;
            if ((float)(ix - 1) >= shxpnx) fx = (float)ix + fx2;
            const float xold = cth * fx + ysid;
            const float yold = sth * fx + ycod;
            out[(size_t)(iy - 1) * nx + (ix - 1)] =
                fbs2_ref(xold, yold, nx, nx, ny, img, nx, x1.data(), y1.data(), xy2.data(), true);
        }
    }
}

template <typename R> struct Rot;
template <> struct Rot<float> {
    static void make(const double* ang, float* dm, int) { cang_ref((float)ang[2], (float)ang[1], (float)ang[0], dm); }
};
template <> struct Rot<double> {
    static void make(const double* ang, double* dm, int exact_pi) {
        cang_f64(ang[2], ang[1], ang[0], dm, deg2rad(exact_pi));
    }
};

// Whole operation (BP 3F when half == NULL, BP 32F otherwise).
// angles: nimg x 3 doc-file registers (PSI, THETA, PHI) in degrees; the
// rotation is CANG(PHI, THETA, PSI) (bp3f.f:1290-1291).
// exact_pi (f64 only): 0 -> the reference's degree conversion (default), 1 -> exact pi.
// sym: nsym x 3 doc registers of the symmetry operations (BUILDS,
// builds.f:35-63: CANG(PHI, THETA, PSI) of each row); used when nsym > 1:
// each projection is inserted once per DMS = SM DM (bp32f.f:866-872) and every
// output volume is passed through SYMVOL after WINDKB2A (bp32f.f:278-303).
template <typename R>
int bp_run(int nx, int nimg, const float* imgs, const double* angles, const int8_t* half,
           R* out_full, R* out_h1, R* out_h2, R* xw_out, int exact_pi, int nsym = 1, const double* sym = nullptr) {
    if (nx < 2) return 101;
    const int N = 2 * nx, H = nx + 1;
    const size_t nh = (size_t)H * N * N;
    const int nvol = half ? 2 : 1;
    std::vector<R> X(2 * nh * nvol, R(0)), W(nh * nvol, R(0));
    std::vector<R> tab(LTAB + 1);
    R fltb;
    kb_table<R>(N, tab.data(), &fltb);
    std::vector<R> bi(2 * (size_t)H * N);
    R dm[9];
    if (nsym < 1 || (nsym > 1 && !sym)) return 102;
    std::vector<R> sm(9 * (size_t)nsym);
    for (int k = 0; k < nsym && nsym > 1; ++k) Rot<R>::make(sym + 3 * k, sm.data() + 9 * k, exact_pi);
    for (int p = 0; p < nimg; ++p) {
        Rot<R>::make(angles + 3 * p, dm, exact_pi);
        spectrum<R>(imgs + (size_t)p * nx * nx, nx, bi.data());
        const int v = half ? (half[p] ? 1 : 0) : 0;
        for (int k = 0; k < nsym; ++k) {
            R dms[9];
            if (nsym > 1) sym_matmul<R>(sm.data() + 9 * k, dm, dms);
            else std::copy(dm, dm + 9, dms);
            insert_projection<R>(nx, X.data() + 2 * nh * v, W.data() + nh * v, bi.data(), dms, tab.data(), fltb);
        }
    }
    for (int v = 0; v < nvol; ++v) symplane0<R>(nx, X.data() + 2 * nh * v, W.data() + nh * v);
    if (nvol == 2) {
        // full = NRMW2(XE + XO, WE + WO) (bp32f.f:345-351); halves from XE, XO
        std::vector<R> Xs(2 * nh), Ws(nh);
        for (size_t q = 0; q < 2 * nh; ++q) Xs[q] = X[q] + X[2 * nh + q];
        for (size_t q = 0; q < nh; ++q) Ws[q] = W[q] + W[nh + q];
        if (xw_out) { std::copy(Xs.begin(), Xs.end(), xw_out); std::copy(Ws.begin(), Ws.end(), xw_out + 2 * nh); }
        if (out_full) finalize_volume<R>(nx, Xs.data(), Ws.data(), out_full);
        if (out_h1) finalize_volume<R>(nx, X.data(), W.data(), out_h1);
        if (out_h2) finalize_volume<R>(nx, X.data() + 2 * nh, W.data() + nh, out_h2);
    } else {
        if (xw_out) { std::copy(X.begin(), X.end(), xw_out); std::copy(W.begin(), W.end(), xw_out + 2 * nh); }
        if (out_full) finalize_volume<R>(nx, X.data(), W.data(), out_full);
    }
    if (nsym > 1) {
        // SYMVOL in single precision, as the reference's REAL volumes
        const size_t nv = (size_t)nx * nx * nx;
        std::vector<float> q1(nv), q2(nv), smf(sm.begin(), sm.end());
        for (R* out : {out_full, out_h1, out_h2}) {
            if (!out) continue;
            for (size_t q = 0; q < nv; ++q) q1[q] = (float)out[q];
            symvol_ref(nx, q1.data(), q2.data(), smf.data(), nsym);
            for (size_t q = 0; q < nv; ++q) out[q] = (R)q2[q];
        }
    }
    return 0;
}

}  // namespace

extern "C" {

double or_bessel_i1(double x) { return bessel_i1(x); }
double or_bessel_i15(double x) { return bessel_i15(x); }
double or_bessel_j15(double x) { return bessel_j15(x); }

void or_cang_ref(float phi, float theta, float psi, float* dm) { cang_ref(phi, theta, psi, dm); }
void or_cang_f64(double phi, double theta, double psi, double* dm, int exact_pi) {
    cang_f64(phi, theta, psi, dm, deg2rad(exact_pi));
}

void or_kb_table_f32(int n2, float* tab, float* fltb) { kb_table<float>(n2, tab, fltb); }
void or_kb_table_f64(int n2, double* tab, double* fltb) { kb_table<double>(n2, tab, fltb); }

double or_asta_mean(const float* img, int L) { return asta_mean(img, L); }
void or_pad_f32(const float* img, int L, int N, float* out) { pad_image<float>(img, L, N, out); }
void or_pad_f64(const float* img, int L, int N, double* out) { pad_image<double>(img, L, N, out); }

void or_fft1(double* data, int n, int sign) {
    FFTPlan P = make_plan(n, sign);
    std::vector<double> a(2 * n), b(2 * n);
    fft_strided(P, data, 1, a.data(), b.data());
}
void or_fft2_r2c(const double* in, int N, double* out) { fft2_r2c(in, N, out); }
void or_fft3_c2r(const double* in, int N, double* out) { fft3_c2r(in, N, out); }

void or_spectrum_f32(const float* img, int nx, float* bi) { spectrum<float>(img, nx, bi); }
void or_spectrum_f64(const float* img, int nx, double* bi) { spectrum<double>(img, nx, bi); }

void or_oneline_f32(int J, int nx, float* X, float* W, const float* bi, const float* dm, const float* tab, float fltb) {
    oneline<float>(J, 2 * nx, nx, X, W, bi, dm, tab, fltb);
}
void or_insert_f32(int nx, float* X, float* W, const float* bi, const float* dm, const float* tab, float fltb) {
    insert_projection<float>(nx, X, W, bi, dm, tab, fltb);
}
void or_insert_f64(int nx, double* X, double* W, const double* bi, const double* dm, const double* tab, double fltb) {
    insert_projection<double>(nx, X, W, bi, dm, tab, fltb);
}
// reference geometry and fp32 per-tap terms, FP64 accumulation
void or_insert_f32_acc64(int nx, double* X, double* W, const float* bi, const float* dm, const float* tab, float fltb) {
    const int N = 2 * nx;
    for (int J = -nx + 1; J <= nx; ++J) oneline<float, double>(J, N, nx, X, W, bi, dm, tab, fltb);
}
// Gather of nproj central slices from the FP64 half grid X (2 (NX+1) N N
// doubles) into out (nproj x N x (NX+1) complex, BI layout); fp32 table tab
// (5000 entries, step fltb) and rotations dm (nproj x 9), FP64 sums.
void or_gather(int nx, const double* X, int nproj, const float* dm, const float* tab, float fltb, double* out) {
    const int N = 2 * nx, H = nx + 1;
    for (int p = 0; p < nproj; ++p)
        for (int J = -nx + 1; J <= nx; ++J)
            gather_line(J, N, nx, X, out + 2 * (size_t)p * N * H, dm + 9 * (size_t)p, tab, fltb);
}
// ONELINE scatter of FP64 BI spectra with an fp32 table and rotations, FP64
// terms and sums: the transpose of or_gather with the same table.
void or_scatter(int nx, const double* bi, int nproj, const float* dm, const float* tab, float fltb, double* X, double* W) {
    const int N = 2 * nx, H = nx + 1;
    for (int p = 0; p < nproj; ++p)
        for (int J = -nx + 1; J <= nx; ++J)
            oneline<float, double, double>(J, N, nx, X, W, bi + 2 * (size_t)p * N * H, dm + 9 * (size_t)p, tab, fltb);
}
void or_symplane0_f32(int nx, float* X, float* W) { symplane0<float>(nx, X, W); }
void or_symplane0_f64(int nx, double* X, double* W) { symplane0<double>(nx, X, W); }
void or_normalize_f32(int nx, float* X, const float* W) { normalize<float>(nx, X, W); }
void or_normalize_f64(int nx, double* X, const double* W) { normalize<double>(nx, X, W); }
void or_inverse3_f32(int nx, const float* X, float* real) { inverse3<float>(nx, X, real); }
void or_inverse3_f64(int nx, const double* X, double* real) { inverse3<double>(nx, X, real); }
void or_window_f32(int nx, const float* real, float* out) { window<float>(nx, real, out); }
void or_window_f64(int nx, const double* real, double* out) { window<double>(nx, real, out); }
void or_finalize_f32(int nx, float* X, const float* W, float* out) { finalize_volume<float>(nx, X, W, out); }
void or_finalize_f64(int nx, double* X, const double* W, double* out) { finalize_volume<double>(nx, X, W, out); }

int or_bp_f32(int nx, int nimg, const float* imgs, const double* angles, const int8_t* half,
              float* out_full, float* out_h1, float* out_h2, float* xw_out) {
    return bp_run<float>(nx, nimg, imgs, angles, half, out_full, out_h1, out_h2, xw_out, 0);
}
int or_bp_f64(int nx, int nimg, const float* imgs, const double* angles, const int8_t* half,
              double* out_full, double* out_h1, double* out_h2, double* xw_out, int exact_pi) {
    return bp_run<double>(nx, nimg, imgs, angles, half, out_full, out_h1, out_h2, xw_out, exact_pi);
}
int or_bp_sym_f32(int nx, int nimg, const float* imgs, const double* angles, const int8_t* half,
                  float* out_full, float* out_h1, float* out_h2, float* xw_out, int nsym, const double* sym) {
    return bp_run<float>(nx, nimg, imgs, angles, half, out_full, out_h1, out_h2, xw_out, 0, nsym, sym);
}
void or_sym_matmul_f32(const float* sm, const float* dm, float* out) { sym_matmul<float>(sm, dm, out); }
float or_triquad(double r, double s, double t, const float* f) { return triquad_ref(r, s, t, f); }
void or_symvol(int nx, const float* q1, float* q2, const float* sm, int nsym) { symvol_ref(nx, q1, q2, sm, nsym); }
float or_quadri(float x, float y, int nx, int ny, const float* f) { return quadri_ref(x, y, nx, ny, f); }
float or_fbs2(float x, float y, int nxld, int nx, int ny, const float* fd, int nxp, const float* x1, const float* y1,
              const float* xy2, int chk) {
    return fbs2_ref(x, y, nxld, nx, ny, fd, nxp, x1, y1, xy2, chk != 0);
}
void or_fbs2_prep(const float* img, int nx, int ny, float* x1, float* y1, float* xy2) {
    fbs2_prep_ref(img, nx, ny, x1, y1, xy2);
}
void or_rtsf(const float* img, float* out, int nx, int ny, float theta, float shx, float shy) {
    rtsf_ref(img, out, nx, ny, theta, shx, shy);
}
void or_rtsq(const float* img, float* out, int nx, int ny, float theta, float scli, float shx, float shy) {
    rtsq_ref(img, out, nx, ny, theta, scli, shx, shy);
}

// Number of distinct half-grid voxels that receive a nonzero KB weight from one
// projection (Friedel-reflected targets counted once): the unit U(NX) of the
// algorithmic byte count.  Uses the f64 sample geometry of ONELINE.
long or_count_distinct(int nx, const double* dm) {
    const int N = 2 * nx, H = nx + 1;
    const size_t nh = (size_t)H * N * N;
    std::vector<double> X(2 * nh, 0.0), W(nh, 0.0), bi(2 * (size_t)H * N, 1.0);
    std::vector<double> tab(LTAB + 1);
    double fltb;
    kb_table<double>(N, tab.data(), &fltb);
    insert_projection<double>(nx, X.data(), W.data(), bi.data(), dm, tab.data(), fltb);
    long c = 0;
    for (size_t q = 0; q < nh; ++q) c += (W[q] != 0.0);
    return c;
}

// Number of (voxel, weight) updates one projection performs (direct + reflected taps).
long or_count_updates(int nx, const double* dm) {
    const int N = 2 * nx, H = nx + 1;
    const size_t nh = (size_t)H * N * N;
    std::vector<double> X(2 * nh, 0.0), W(nh, 0.0), bi(2 * (size_t)H * N, 0.0);
    std::vector<double> tab(LTAB + 1);
    double fltb;
    kb_table<double>(N, tab.data(), &fltb);
    // count via unit weights: a table of ones where the KB table is nonzero
    for (auto& t : tab) t = (t != 0.0) ? 1.0 : 0.0;
    insert_projection<double>(nx, X.data(), W.data(), bi.data(), dm, tab.data(), fltb);
    double s = 0.0;
    for (size_t q = 0; q < nh; ++q) s += W[q];
    return (long)llround(s);
}

// CPU baseline: the reference's production scheme (SP_MP build): per image,
// spectrum, then six phases of an OpenMP loop over rows J = -NX+JT, ..., step 6
// (bp32f.f:873-888, bp3f.f:468-474), FP32, unsynchronised like the reference.
// dms: nimg x 9 rotation matrices (fp32).  Returns 0.
int or_insert_omp_f32(int nx, int nimg, const float* imgs, const float* dms, float* X, float* W, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const int N = 2 * nx, H = nx + 1;
    std::vector<float> tab(LTAB + 1);
    float fltb;
    kb_table<float>(N, tab.data(), &fltb);
    std::vector<float> bi(2 * (size_t)H * N);
    for (int p = 0; p < nimg; ++p) {
        spectrum<float>(imgs + (size_t)p * nx * nx, nx, bi.data());
        const float* dm = dms + 9 * (size_t)p;
        for (int jt = 1; jt <= 6; ++jt) {
            #pragma omp parallel for schedule(static)
            for (int J = -nx + jt; J <= nx; J += 6)
                oneline<float>(J, N, nx, X, W, bi.data(), dm, tab.data(), fltb);
        }
    }
    return 0;
}

int or_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
