// Device-side special functions and Green's-function kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include "bem_bessel_coeffs.h"

namespace bemuse {
namespace dev {

constexpr double kPI = 3.14159265358979323846;
constexpr double kFourPIinv = 0.25 / 3.14159265358979323846;
constexpr double kEUL = 0.5772156649;   // truncated Euler constant of the reference (BEMUse_Math_Types.h:99)

// 1/x: the hardware reciprocal refined by two Newton steps (error within
// about 1 ulp; the arguments here are finite, positive and far from the
// exponent limits). Replaces IEEE divisions, whose scaling and fix-up steps
// cost several times more in the Havelock evaluation's inner loop.
__host__ __device__ inline double frcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}

// sqrt(a) for finite a >= 0 from the hardware reciprocal square root and two
// Newton-Raphson corrections of the root and its half-reciprocal (Goldschmidt
// form). The library sqrt adds an exponent scaling for denormal and huge
// arguments; the arguments here (squared distances scaled by K) are neither.
__host__ __device__ inline double fsqrt(double a) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rsq(a);
    double y = a * r, h = 0.5 * r;
    const double e = fma(-h, y, 0.5);
    y = fma(y, e, y);
    h = fma(h, e, h);
    const double d = fma(-y, y, a);
    y = fma(d, h, y);
    return a == 0.0 ? 0.0 : y;
#else
    return sqrt(a);
#endif
}

// a b + c for a coefficient c known at compile time, as one v_fma_f64 with c
// in an SGPR pair. Left to itself the compiler selects v_fmac_f64, whose
// addend is also its destination, and copies every 64-bit coefficient into
// that VGPR pair first (two v_mov_b32 per term): in the Horner and series
// loops of the Green function those copies were a fifth of all VALU
// instructions. The operation is the same fused multiply-add, so results
// are unchanged.
__host__ __device__ inline double fma_k(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
#else
    return fma(a, b, c);
#endif
}

// Coefficients of the Struve approximations (Kernels.cpp:35-116): for x <= 3
// polynomials in (x/3)^2 (p), above rational functions of (3/x)^2 with
// numerator a and denominator 1 + b. H0: p0..p5, a0..a3, b1..b3; H1: p1..p6,
// a0..a3, b1..b3.
struct StruveCoeffs {
    double h0p[6], h0a[4], h0b[3];
    double h1p[6], h1a[4], h1b[3];
};
constexpr StruveCoeffs kStruve = {
    {1.909859164, -1.909855001, 0.687514637, -0.126164557, 0.013828813, -0.000876918},
    {0.99999906, 4.77228920, 3.85542044, 0.32303607},
    {4.88331068, 4.28957333, 0.52120508},
    {1.909859286, -1.145914713, 0.294656958, -0.042070508, 0.003785727, -0.000207183},
    {1.00000004, 3.92205313, 2.64893033, 0.27450895},
    {3.81095112, 2.26216956, 0.10885141},
};

// Struve H0 and H1 together: below 3 the polynomials; above, the Bessel
// Y0 / Y1 plus the rational functions, whose two denominators share one
// reciprocal.
__host__ __device__ inline void struve_h01(double x, double y0, double y1, double& h0, double& h1) {
    constexpr const double* P = kStruve.h0p;
    constexpr const double* Q = kStruve.h1p;
    if (x <= 3.0) {
        double x3 = x * (1.0 / 3.0), yy = x3 * x3;
        double p = P[5], q = Q[5];
#pragma unroll
        for (int k = 4; k >= 0; k--) {
            p = fma_k(p, yy, P[k]);
            q = fma_k(q, yy, Q[k]);
        }
        h0 = p * x3;
        h1 = q * yy;
        return;
    }
    constexpr const double* a = kStruve.h0a;
    constexpr const double* b = kStruve.h0b;
    constexpr const double* c = kStruve.h1a;
    constexpr const double* d = kStruve.h1b;
    double u = 3.0 * frcp(x), yy = u * u;
    double c1a = 2.0 * fma_k(fma_k(fma_k(a[3], yy, a[2]), yy, a[1]), yy, a[0]);
    double c2a = kPI * x * (1.0 + fma_k(fma_k(b[2], yy, b[1]), yy, b[0]) * yy);
    double c1b = 2.0 * fma_k(fma_k(fma_k(c[3], yy, c[2]), yy, c[1]), yy, c[0]);
    double c2b = kPI * (1.0 + fma_k(fma_k(d[2], yy, d[1]), yy, d[0]) * yy);
    double rab = frcp(c2a * c2b);
    h0 = c1a * c2b * rab + y0;
    h1 = c1b * c2a * rab + y1;
}

// Bessel J0(x), J1(x) for x >= 0 and, for x > 3, Y0(x), Y1(x) (0 below):
// what HavelockGF needs (Kernels.cpp:31-34, 284-295, 327-338; Y only above
// 3, :56-74). x <= 3: power series. x > 3: Hankel form with P0, Q0, P1, Q1
// as their degree-23 Chebyshev interpolants in t = 6/x - 1, rewritten in
// powers of t and evaluated by Horner's rule (coefficients, method and the
// conditioning of the power form in gen_bessel.py) and one shared sin/cos of
// x (three-piece pi/2 reduction,
// Taylor series on |r| <= pi/4). Beyond x = 1.6e6 the reduction is no longer
// exact and the device library functions are used.
__host__ __device__ inline void bessel01_series(double x, double& J0, double& J1) {
    const double t = -0.25 * x * x;
    double a = kBesJ0S[15], b = kBesJ1S[15];
#pragma unroll
    for (int k = 14; k >= 0; k--) {
        a = fma_k(a, t, kBesJ0S[k]);
        b = fma_k(b, t, kBesJ1S[k]);
    }
    J0 = a;
    J1 = 0.5 * x * b;
}

__host__ __device__ inline void bessel01(double x, double& J0, double& J1, double& Y0, double& Y1) {
    if (x <= 3.0) {
        bessel01_series(x, J0, J1);
        Y0 = 0.0;
        Y1 = 0.0;
        return;
    }
    if (x > 1.6e6) {
        J0 = j0(x); J1 = j1(x); Y0 = y0(x); Y1 = y1(x);
        return;
    }
    const double u = 3.0 * frcp(x), t = 2.0 * u - 1.0;
    // Horner for the four series at once
    double P0 = kBesP0[23], Q0 = kBesQ0[23], P1 = kBesP1[23], Q1 = kBesQ1[23];
#pragma unroll
    for (int k = 22; k >= 0; k--) {
        P0 = fma_k(P0, t, kBesP0[k]);
        Q0 = fma_k(Q0, t, kBesQ0[k]);
        P1 = fma_k(P1, t, kBesP1[k]);
        Q1 = fma_k(Q1, t, kBesQ1[k]);
    }
    // x = n pi/2 + r, |r| <= pi/4 (n pi/2 pieces exact for n < 2^21)
    const double n = rint(x * kTwoOverPi);
    double r = fma(-n, kPio2A, x);
    r = fma(-n, kPio2B, r);
    r = fma(-n, kPio2C, r);
    const double rr = r * r;
    double sp = kSinT[8], cp = kCosT[9];
#pragma unroll
    for (int k = 7; k >= 0; k--) sp = fma_k(sp, rr, kSinT[k]);
#pragma unroll
    for (int k = 8; k >= 0; k--) cp = fma_k(cp, rr, kCosT[k]);
    const double sr = r * sp, cr = cp;
    double sx, cx;
    switch (((long long)n) & 3) {
        case 0: sx = sr; cx = cr; break;
        case 1: sx = cr; cx = -sr; break;
        case 2: sx = -sr; cx = -cr; break;
        default: sx = -cr; cx = sr; break;
    }
    // chi0 = x - pi/4: cos chi0 = (cos x + sin x)/sqrt 2, sin chi0 = (sin x - cos x)/sqrt 2
    const double cc = (cx + sx) * kSqrtHalf, sc = (sx - cx) * kSqrtHalf;
    const double A = sqrt(kTwoOver3Pi * u);       // sqrt(2 / (pi x))
    J0 = A * (P0 * cc - Q0 * sc);
    Y0 = A * (P0 * sc + Q0 * cc);
    J1 = A * (P1 * sc + Q1 * cc);                 // chi1 = chi0 - pi/2
    Y1 = A * (Q1 * sc - P1 * cc);
}

// FP32 versions of the special functions above, for the wave term of far
// pairs (a measured option; the exact path stays FP64). Same methods, each
// series cut where its terms fall below FP32 resolution: the Struve power
// series at 6 terms, the Chebyshev interpolants at degree 10 (in powers of t
// like the FP64 ones), sin/cos of the reduced
// argument at degree 9/10; the reduction x = n pi/2 + r stays FP64 (an FP32
// argument of K R ~ 100 would already carry a phase error of 1e-5).
// bessel01_f32 covers x > 3 only (below, the caller takes the FP64 series).
__device__ inline void bessel01_f32(double xd, float& J0, float& J1, float& Y0, float& Y1) {
    const float x = (float)xd;
    const float u = 3.0f / x, t = 2.0f * u - 1.0f;
    float P0 = (float)kBesP0F[10], Q0 = (float)kBesQ0F[10], P1 = (float)kBesP1F[10], Q1 = (float)kBesQ1F[10];
#pragma unroll
    for (int k = 9; k >= 0; k--) {
        P0 = fmaf(P0, t, (float)kBesP0F[k]);
        Q0 = fmaf(Q0, t, (float)kBesQ0F[k]);
        P1 = fmaf(P1, t, (float)kBesP1F[k]);
        Q1 = fmaf(Q1, t, (float)kBesQ1F[k]);
    }
    const double n = rint(xd * kTwoOverPi);
    double rd = fma(-n, kPio2A, xd);
    rd = fma(-n, kPio2B, rd);
    const float r = (float)rd, rr = r * r;
    float sp = (float)kSinT[4], cp = (float)kCosT[5];
#pragma unroll
    for (int k = 3; k >= 0; k--) sp = fmaf(sp, rr, (float)kSinT[k]);
#pragma unroll
    for (int k = 4; k >= 0; k--) cp = fmaf(cp, rr, (float)kCosT[k]);
    const float sr = r * sp, cr = cp;
    float sx, cx;
    switch (((long long)n) & 3) {
        case 0: sx = sr; cx = cr; break;
        case 1: sx = cr; cx = -sr; break;
        case 2: sx = -sr; cx = -cr; break;
        default: sx = -cr; cx = sr; break;
    }
    const float cc = (cx + sx) * (float)kSqrtHalf, sc = (sx - cx) * (float)kSqrtHalf;
    const float A = sqrtf((float)kTwoOver3Pi * u);
    J0 = A * (P0 * cc - Q0 * sc);
    Y0 = A * (P0 * sc + Q0 * cc);
    J1 = A * (P1 * sc + Q1 * cc);
    Y1 = A * (Q1 * sc - P1 * cc);
}

__device__ inline void struve_h01_f32(double xd, float y0, float y1, float& h0, float& h1) {
    constexpr const double* P = kStruve.h0p;
    constexpr const double* Q = kStruve.h1p;
    const float x = (float)xd;
    if (xd <= 3.0) {
        float x3 = x * (1.0f / 3.0f), yy = x3 * x3;
        float p = (float)P[5], q = (float)Q[5];
#pragma unroll
        for (int k = 4; k >= 0; k--) {
            p = fmaf(p, yy, (float)P[k]);
            q = fmaf(q, yy, (float)Q[k]);
        }
        h0 = p * x3;
        h1 = q * yy;
        return;
    }
    constexpr const double* a = kStruve.h0a;
    constexpr const double* b = kStruve.h0b;
    constexpr const double* c = kStruve.h1a;
    constexpr const double* d = kStruve.h1b;
    const float u = 3.0f / x, yy = u * u;
    const float c1a = 2.0f * fmaf(fmaf(fmaf((float)a[3], yy, (float)a[2]), yy, (float)a[1]), yy, (float)a[0]);
    const float c2a = (float)kPI * x * (1.0f + fmaf(fmaf((float)b[2], yy, (float)b[1]), yy, (float)b[0]) * yy);
    const float c1b = 2.0f * fmaf(fmaf(fmaf((float)c[3], yy, (float)c[2]), yy, (float)c[1]), yy, (float)c[0]);
    const float c2b = (float)kPI * (1.0f + fmaf(fmaf((float)d[2], yy, (float)d[1]), yy, (float)d[0]) * yy);
    const float rab = 1.0f / (c2a * c2b);
    h0 = c1a * c2b * rab + y0;
    h1 = c1b * c2a * rab + y1;
}

// Natural logarithm of a finite normal x > 0: x = m 2^e with m in
// [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with s = (m - 1) / (m + 1),
// |s| <= 0.1716, summed as 2 s (1 + s^2 / 3 + s^4 / 5 + ...) to s^22 (the next
// term is below 1e-18 of the sum), and e ln 2 with ln 2 split into a 32-bit
// head and its remainder. The library log carries extra-precision steps that
// cost about twice the instructions.
__host__ __device__ inline double flog(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    constexpr double kLn2Hi = 0.693147180369123816490, kLn2Lo = 1.90821492927058770002e-10;
    int e = __builtin_amdgcn_frexp_exp(x);
    double m = __builtin_amdgcn_frexp_mant(x);       // [0.5, 1)
    if (m < 0.70710678118654752440) { m = 2.0 * m; e -= 1; }
    const double s = (m - 1.0) * frcp(m + 1.0), z = s * s;
    double p = 2.0 / 23.0;
#pragma unroll
    for (int k = 10; k >= 0; k--) p = fma_k(p, z, 2.0 / (2 * k + 1));
    const double ed = (double)e;
    return fma(ed, kLn2Hi, fma(s, p, ed * kLn2Lo));
#else
    return log(x);
#endif
}

// Polynomials A..D (local flow) and dA..dC (its horizontal derivative) of the
// Wu/Zhang/Zhu/Li/Wan/Noblesse global approximation (Eur. J. Mech. B 65 (2017)
// eq. 14; Liang/Wu/Noblesse, Appl. Ocean Res. 74 (2018) eq. 16), coefficients
// as tabulated at Kernels.cpp:120-256, evaluated in Horner form.
__host__ __device__ inline double horner(const double* c, int n, double t) {
    double s = c[n - 1];
    for (int i = n - 2; i >= 0; i--) s = fma_k(s, t, c[i]);
    return s;
}

struct GFCoeffs {
    double A[10], B[10], C[8], D[10], dA[10], dB[10], dC[6];
};

constexpr GFCoeffs kGF = {
    {1.21, -13.328, 215.896, -1763.96, 8418.94, -24314.21, 42002.57, -41592.9, 21859.0, -4838.6},
    {0.938, 5.737, -67.92, 796.534, -4780.77, 17137.74, -36618.81, 44894.06, -29030.24, 7671.22},
    {1.268, -9.747, 209.653, -1397.89, 5155.67, -9844.35, 9136.4, -3272.62},
    {0.632, -40.97, 667.16, -6072.07, 31127.39, -96293.05, 181856.75, -205690.43, 128170.2, -33744.6},
    {2.948, -24.53, 249.69, -754.85, -1187.71, 16370.75, -48811.41, 68220.87, -46688.0, 12622.25},
    {1.11, 2.894, -76.765, 1565.35, -11336.19, 44270.15, -97014.11, 118879.26, -76209.82, 19923.28},
    {14.19, -148.24, 847.8, -2318.58, 3168.35, -1590.27}};

// Non-dimensional free-surface Green function (HavelockGF, Kernels.cpp:342-412):
// in (X, Y, V) = K (x_R - y_q); out GF = L0 + W and GFh = Ls + Wh (complex).
// DGFX = GFh X/h, DGFY = GFh Y/h (0 when h <= 1e-6), DGFZ = GF.
struct GFVal { double gr, gi, hr, hi, h; };

// Quotients are products with frcp reciprocals (1/d shared by alpha and beta,
// rho = d/(1+d) and 1 - rho from one reciprocal).
// S32: exp, log, Bessel, Struve in FP32 (the far-pair option); the
// polynomials, reciprocals and the combination stay FP64 either way (the
// polynomials' large alternating coefficients lose 1e-4 .. 1e-3 of their
// value in FP32). Where h <= 3 the imaginary part's factors e^V, J0 and J1
// stay FP64 too (power series, no Y needed there): at low frequency every
// pair has small h, and the damping of a body small against the wavelength
// is what is left of e^V J0(h) ~ 1 + V + (V^2/2 - h^2/4) after its leading
// terms cancel over the hull, so FP32 rounding of those factors (6e-8 of 1)
// would swamp it.
template <bool S32>
__host__ __device__ inline GFVal havelock_gf_t(double X, double Y, double vv) {
    double hh = fsqrt(X * X + Y * Y);
    double dd = fsqrt(hh * hh + vv * vv);
    double rdd = frcp(dd);
    double al = -vv * rdd;
    double be = hh * rdd;
    double omr = frcp(1.0 + dd);
    double rho = dd * omr;
    double rw = rho * omr * omr * omr;
    double dd3 = dd * dd * dd;
    double ev, lg, emd;
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (S32) {
        ev = hh <= 3.0 ? exp(vv) : (double)expf((float)vv);
        lg = (double)logf((float)(0.5 * (dd - vv)));
        emd = (double)expf((float)-dd);
    } else
#endif
    {
        ev = exp(vv);
        lg = flog(0.5 * (dd - vv));
        emd = exp(-dd);
    }
    double inv1 = frcp(1.0 + dd3);
    // local flow L0 (GF_Func_L0, :258-282)
    double PP = (lg + kEUL - 2.0 * dd * dd) * ev + dd * dd - vv;
    double A = horner(kGF.A, 10, rho), B = horner(kGF.B, 10, rho);
    double C = horner(kGF.C, 8, rho), D = horner(kGF.D, 10, rho);
    double RR = (1.0 - be) * A - be * B - al * C * frcp(1.0 + 6.0 * al * rho * omr) + be * (1.0 - be) * D;
    double L0 = 2.0 * PP * inv1 + 2.0 * rw * RR;
    // derivative local flow Ls (GF_Func_Ls, :297-325)
    double PS = (be + hh) * frcp(dd - vv) - 2.0 * be + 2.0 * dd * ev - hh;
    double QS = emd * (1.0 - be) * (1.0 + dd * inv1);
    double dA = horner(kGF.dA, 10, rho), dB = horner(kGF.dB, 10, rho), dC = horner(kGF.dC, 6, rho);
    double R2 = be * dA - (1.0 - al) * dB + be * (1.0 - be) * rho * (1.0 - 2.0 * rho) * dC;
    double Ls = 2.0 * PS * inv1 - 4.0 * QS + 2.0 * rw * R2;
    // wave parts W, Wh (GF_Func_W / GF_Func_Wh, :284-295, :327-338)
    double j0v, j1v, y0v, y1v, h0v, h1v;
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (S32) {
        float y0f = 0.0f, y1f = 0.0f, h0f, h1f;
        if (hh <= 3.0) {
            bessel01_series(hh, j0v, j1v);
        } else {
            float j0f, j1f;
            bessel01_f32(hh, j0f, j1f, y0f, y1f);
            j0v = j0f; j1v = j1f;
        }
        struve_h01_f32(hh, y0f, y1f, h0f, h1f);
        h0v = h0f; h1v = h1f;
    } else
#endif
    {
        bessel01(hh, j0v, j1v, y0v, y1v);
        struve_h01(hh, y0v, y1v, h0v, h1v);
    }
    double E = 2.0 * kPI * ev;
    GFVal o;
    o.gr = L0 + E * h0v;
    o.gi = -E * j0v;
    o.hr = Ls + E * (2.0 / kPI - h1v);
    o.hi = E * j1v;
    o.h = hh;
    return o;
}
__host__ __device__ inline GFVal havelock_gf(double X, double Y, double vv) { return havelock_gf_t<false>(X, Y, vv); }

// Wave S/D influence of one mirrored panel at receiver (xr, yr, zr): sum over
// its quadrature nodes of w G and w (grad_R G . n_q) (Panel.cpp:601-636,776-794).
// q points to the panel's 32-double quadrature record.
template <bool S32 = false>
__host__ __device__ inline void wave_panel(const double* q, double xr, double yr, double zr, double K,
                                  double& sr, double& si, double& dr, double& di) {
    sr = si = dr = di = 0.0;
    if (K == 0.0) return;
    int nq = (int)q[28];
    double s1 = K * kFourPIinv, s2 = K * K * kFourPIinv;
    for (int k = 0; k < nq; k++) {
        const double* p = q + 7 * k;
        double X = (xr - p[0]) * K, Y = (yr - p[1]) * K, V = (zr - p[2]) * K;
        GFVal g = havelock_gf_t<S32>(X, Y, V);
        double w = p[6];
        sr += w * (g.gr * s1);
        si += w * (g.gi * s1);
        double cx = 0.0, cy = 0.0;
        if (g.h > 1.0e-6) { double rh = frcp(g.h); cx = X * rh; cy = Y * rh; }
        double tr = (g.hr * cx) * p[3] + (g.hr * cy) * p[4] + g.gr * p[5];
        double ti = (g.hi * cx) * p[3] + (g.hi * cy) * p[4] + g.gi * p[5];
        dr += w * (tr * s2);
        di += w * (ti * s2);
    }
}

// Analytic constant-strength source/dipole influence of a flat polygon
// (FlatTriPanel/FlatQuadPanel::Inf_SourceKernel, Panel.cpp:17-156) at point P;
// rec is the panel's 32-double record (centroid, frame, local vertices,
// slopes, edge lengths).
__host__ __device__ inline void rankine_panel(const double* rec, double px, double py, double pz, double& S, double& Dv) {
    double dx = px - rec[0], dy = py - rec[1], dz = pz - rec[2];
    double x = rec[3] * dx + rec[4] * dy + rec[5] * dz;
    double y = rec[6] * dx + rec[7] * dy + rec[8] * dz;
    double z = rec[9] * dx + rec[10] * dy + rec[11] * dz;
    int n = (int)rec[28];
    double Z2 = z * z;
    double DX[4], DY[4], e[4], h[4], r[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        DX[k] = x - rec[12 + k];
        DY[k] = y - rec[16 + k];
        e[k] = DX[k] * DX[k] + Z2;
        h[k] = DX[k] * DY[k];
        r[k] = sqrt(DX[k] * DX[k] + DY[k] * DY[k] + Z2);
    }
    double phi = 0.0, bl = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (k < n) {
            int k1 = (k + 1 == n) ? 0 : k + 1;
            double m = rec[20 + k], d = rec[24 + k];
            phi += atan((m * e[k] - h[k]) / (z * r[k])) - atan((m * e[k1] - h[k1]) / (z * r[k1]));
            double ddx = rec[12 + k1] - rec[12 + k], ddy = rec[16 + k1] - rec[16 + k];
            double B = (DX[k] * ddy - DY[k] * ddx) / d;
            double L = log((r[k] + r[k1] + d) / (r[k] + r[k1] - d));
            bl += B * L;
        }
    }
    Dv = phi * kFourPIinv;
    S = (z * phi - bl) * kFourPIinv;
}

}  // namespace dev
}  // namespace bemuse
