// hostmath.hpp -- FP64 host-side tables: KB interpolant, window, rotations.
#pragma once
#include <cmath>
#include <vector>

namespace spbp {

// ------------------------------------------------------------------ host math
// Modified Bessel I1 by its power series (converges for the 0..6.5 used here).
double bessel_i1(double x) {
    double h = 0.5 * x, term = h, sum = h, h2 = h * h;
    for (int k = 1; k < 200; ++k) {
        term *= h2 / (double(k) * double(k + 1));
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    return sum;
}
// I_{3/2}(x) = sqrt(2/(pi x)) (cosh x - sinh x / x), x > 0.
double bessel_i15(double x) { return std::sqrt(2.0 / (M_PI * x)) * (std::cosh(x) - std::sinh(x) / x); }
double bessel_j15(double x) { return std::sqrt(2.0 / (M_PI * x)) * (std::sin(x) / x - std::cos(x)); }

// KB interpolation table TABI(0:4999) (bp32f.f:193-214), evaluated in FP64
// (used with FP64 geometry, SPBP_GEOM_F64).
void kb_table_f64(int n2, std::vector<double>& tab) {
    const double alpha = 6.5, V = 4.0 / 2.0 / (double)n2, aaaa = 0.9 * V;
    const double b0 = std::sqrt(alpha) * bessel_i1(alpha);
    const double fltb = (double)LTAB / (double)(LN2 + 1);
    tab.assign(LTAB + 1, 0.0);
    for (int i = 0; i <= LTAB; ++i) {
        const double s = (double)i / fltb / (double)n2;
        if (s <= aaaa) {
            const double q = s / aaaa, x = std::sqrt(1.0 - q * q);
            tab[i] = std::sqrt(alpha * x) * bessel_i1(alpha * x) / b0;
        }
    }
}

// I1(x), 0 <= x < 15, in single precision as the reference evaluates it:
// W. J. Cody's rational approximation (SPECFUN routine CALCI1, ACM TOMS 19
// (1993) 22-32, Algorithm 715), reached through BESI1 (besi1.f:3-230).  The
// published coefficients are given as decimal REAL constants; the operation
// order is Cody's: Horner in x^2 for the numerator, in x^2 - 225 for the
// denominator, then (P / Q) x.
float bessel_i1_real(float x) {
    static const float P[15] = {-1.9705291802535139930E-19f, -6.5245515583151902910E-16f, -1.1928788903603238754E-12f,
                                -1.4831904935994647675E-09f, -1.3466829827635152875E-06f, -9.1746443287817501309E-04f,
                                -4.7207090827310162436E-01f, -1.8225946631657315931E+02f, -5.1894091982308017540E+04f,
                                -1.0588550724769347106E+07f, -1.4828267606612366099E+09f, -1.3357437682275493024E+11f,
                                -6.9876779648010090070E+12f, -1.7732037840791591320E+14f, -1.4577180278143463643E+15f};
    static const float Q[5] = {-4.0076864679904189921E+03f, 7.4810580356655069138E+06f, -8.0059518998619764991E+09f,
                               4.8544714258273622913E+12f, -1.3218168307321442305E+15f};
    const float ax = std::fabs(x);
    if (ax < 2.98E-8f) return 0.5f * x;
    float xx = ax * ax;
    float sp = P[0];
    for (int j = 1; j < 15; ++j) sp = sp * xx + P[j];
    xx = xx - 225.0f;
    float sq = xx + Q[0];
    for (int j = 1; j < 5; ++j) sq = sq * xx + Q[j];
    const float r = (sp / sq) * ax;
    return x < 0.0f ? -r : r;
}

// TABI(0:4999) and FLTB in REAL arithmetic, as bp32f.f:193-214 computes them
// (the default table: every per-tap weight is then the reference's).
void kb_table(int n2, std::vector<float>& tab, float& fltb_f) {
    const float alpha = 6.5f, V = (float)(5 - 1) / 2.0f / (float)n2, aaaa = 0.9f * V;
    const float b0 = std::sqrt(alpha) * bessel_i1_real(alpha);
    const float fltb = (float)LTAB / (float)(LN2 + 1);   // REAL(LTAB)/REAL(LN2+1)
    tab.assign(LTAB + 1, 0.0f);
    for (int i = 0; i <= LTAB; ++i) {
        const float s = (float)i / fltb / (float)n2;
        if (s <= aaaa) {
            const float q = s / aaaa;
            const float x = std::sqrt(1.0f - q * q);
            tab[i] = std::sqrt(alpha * x) * bessel_i1_real(alpha * x) / b0;
        }
    }
    fltb_f = fltb;
}

// The FP64 table rounded to fp32 (with FP64 geometry).
void kb_table_f64_f32(int n2, std::vector<float>& tab, float& fltb_f) {
    std::vector<double> t;
    kb_table_f64(n2, t);
    tab.assign(t.begin(), t.end());
    fltb_f = (float)LTAB / (float)(LN2 + 1);
}

// The same tables into a caller's buffer of LTAB + 1 floats (no allocation):
// the REAL evaluation, or with f64 the FP64 one rounded to fp32.
void kb_table_into(int n2, bool f64, float* tab, float& fltb_f) {
    if (!f64) {
        const float alpha = 6.5f, V = (float)(5 - 1) / 2.0f / (float)n2, aaaa = 0.9f * V;
        const float b0 = std::sqrt(alpha) * bessel_i1_real(alpha);
        const float fltb = (float)LTAB / (float)(LN2 + 1);
        for (int i = 0; i <= LTAB; ++i) {
            const float s = (float)i / fltb / (float)n2;
            float t = 0.0f;
            if (s <= aaaa) {
                const float q = s / aaaa;
                const float x = std::sqrt(1.0f - q * q);
                t = std::sqrt(alpha * x) * bessel_i1_real(alpha * x) / b0;
            }
            tab[i] = t;
        }
        fltb_f = fltb;
        return;
    }
    const double alpha = 6.5, V = 4.0 / 2.0 / (double)n2, aaaa = 0.9 * V;
    const double b0 = std::sqrt(alpha) * bessel_i1(alpha);
    const double fltb = (double)LTAB / (double)(LN2 + 1);
    for (int i = 0; i <= LTAB; ++i) {
        const double s = (double)i / fltb / (double)n2;
        double t = 0.0;
        if (s <= aaaa) {
            const double q = s / aaaa, x = std::sqrt(1.0 - q * q);
            t = std::sqrt(alpha * x) * bessel_i1(alpha * x) / b0;
        }
        tab[i] = (float)t;
    }
    fltb_f = (float)LTAB / (float)(LN2 + 1);
}

// |WKB| of WINDKB2A per integer squared radius LR = 0..(L/2)^2, FP64.
void wkb_table(int nx, int N, std::vector<float>& t) {
    const int cc = nx / 2;
    const double alpha = 6.5, xnu = 1.5, aaaa = 0.9 * (4.0 / 2.0 / (double)N);
    const double wkb0 = std::pow(alpha, xnu) / bessel_i15(alpha);
    const double qrt = (2.0 * M_PI * aaaa) * (2.0 * M_PI * aaaa);
    t.assign((size_t)cc * cc + 1, 1.0f);
    for (int lr = 0; lr <= cc * cc; ++lr) {
        const double sigma = qrt * lr - alpha * alpha;
        double wkb;
        if (std::fabs(sigma) < 1.0e-7) wkb = 1.0;
        else if (sigma > 0.0) { const double art = std::sqrt(sigma); wkb = wkb0 * bessel_j15(art) / std::pow(art, xnu); }
        else { const double art = std::sqrt(-sigma); wkb = wkb0 * bessel_i15(art) / std::pow(art, xnu); }
        t[lr] = (float)std::fabs(wkb);
    }
}

// CANG (cang.f:44-82) in FP64 from REAL doc registers, stored as REAL.  The
// degree conversion is float32(pi)/180 evaluated in double (cang.f:50-53)
// unless exact_pi.
void cang(float phi, float theta, float psi, bool exact_pi, float* dm) {
    const double dgr = exact_pi ? M_PI / 180.0 : (double)(float)3.14159265358979323846 / 180.0;
    const double cphi = std::cos((double)phi * dgr), sphi = std::sin((double)phi * dgr);
    const double cthe = std::cos((double)theta * dgr), sthe = std::sin((double)theta * dgr);
    const double cpsi = std::cos((double)psi * dgr), spsi = std::sin((double)psi * dgr);
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


// RTSQ set-up (rtsq.f:93-126) for a square NX image, scale 1: cos/sin of the
// REAL angle in single precision, shifts reduced by AMOD, -(centre + shift).
// Outputs coddscli, siddscli, fixcenmshx, fiycenmshy.
void rtsq_param(int nx, float theta, float shxi, float shyi, float* out4) {
    const float dgr_to_rad = (float)3.14159265358979323846 / 180.0f;
    const float shx = std::fmod(shxi, (float)nx), shy = std::fmod(shyi, (float)nx);
    const int cen = nx / 2 + 1;
    const float ang = theta * dgr_to_rad;
    const float cod = std::cos(ang), sid = std::sin(ang);
    out4[0] = cod / 1.0f;
    out4[1] = sid / 1.0f;
    out4[2] = (float)(-cen) - shx;
    out4[3] = (float)(-cen) - shy;
}

// RTSF's per-image constants (rtsf.f:128-159, scale 1) in REAL: the angle in
// radians as THETAT * PI / 180, its single-precision cosine and sine, the
// shifts MODULO the size and the output-coordinate offsets.
inline float real_modulo(float a, float p) {
    float r = std::fmod(a, p);
    if (r != 0.0f && (r < 0.0f) != (p < 0.0f)) r = r + p;
    return r;
}
void rtsf_param(int nx, float theta, float shxi, float shyi, float* out12) {
    const float pi = (float)3.14159265358979323846;
    const float th = theta * pi / 180.0f;
    const float n = (float)nx, c = (float)(nx / 2 + 1);
    const float shx = real_modulo(shxi, n), shy = real_modulo(shyi, n);
    out12[0] = std::cos(th);
    out12[1] = std::sin(th);
    out12[2] = shx;
    out12[3] = shy;
    out12[4] = shx + n;
    out12[5] = shy + n;
    out12[6] = (-shx) - c;
    out12[7] = ((-shx) + n) - c;
    out12[8] = ((-shx) - n) - c;
    out12[9] = (-shy) - c;
    out12[10] = ((-shy) + n) - c;
    out12[11] = ((-shy) - n) - c;
}

// CANG in FP64 throughout, from FP64 angles (SPBP_GEOM_F64: no fp32 rounding
// of the angles or of the matrix).
void cang_d(double phi, double theta, double psi, bool exact_pi, double* dm) {
    const double dgr = exact_pi ? M_PI / 180.0 : (double)(float)3.14159265358979323846 / 180.0;
    const double cphi = std::cos(phi * dgr), sphi = std::sin(phi * dgr);
    const double cthe = std::cos(theta * dgr), sthe = std::sin(theta * dgr);
    const double cpsi = std::cos(psi * dgr), spsi = std::sin(psi * dgr);
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

void sym_matmul_d(const double* sm, const double* dm, double* out) {
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            double acc = sm[i] * dm[3 * j];
            acc = acc + sm[i + 3] * dm[1 + 3 * j];
            acc = acc + sm[i + 6] * dm[2 + 3 * j];
            out[i + 3 * j] = acc;
        }
}

// DMS = MATMUL(SM, DM) (bp32f.f:866-872) for 3x3 matrices in Fortran (3,3)
// storage (element (i, j) at [i + 3 j]); fp32 products and sums, k = 1..3.
void sym_matmul(const float* sm, const float* dm, float* out) {
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            float acc = sm[i] * dm[3 * j];
            acc = acc + sm[i + 3] * dm[1 + 3 * j];
            acc = acc + sm[i + 6] * dm[2 + 3 * j];
            out[i + 3 * j] = acc;
        }
}

}  // namespace spbp
