// FP64 CPU oracle of the BEMUse per-frequency hydrodynamic hot path.
//
// TEST INFRASTRUCTURE ONLY. Results are compared against it: it may be linked or
// called only by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
// bench.py. The product (libbemuse_amd.so) never links, loads or calls it.
//
// It restates, in plain C++ (g++, OpenMP, libstdc++ TR1 Bessel functions as
// the reference uses), the arithmetic of the reference path:
//   geometry / quadrature   src/Geometry/Geo_Elements.cpp:45-214, CoordSys.h:62-123
//   mirroring               src/Boundary/Boundary_Base.cpp:51-85
//   BC points               src/Solver/Solver_Base.cpp:71-83 (BCF = 1e-5, Solver_Base.h:55)
//   Rankine S/D integrals   src/Geometry/Panel.cpp:17-156
//   Rankine quadrature twin src/Geometry/Panel.cpp:156-170,582-593 + Kernels.cpp:11-27
//   Havelock GF             src/Kernels.cpp:31-412
//   wave kernel + quad      src/Geometry/Panel.cpp:601-636,676-694,776-794
//   assembly                src/Solver/Hydrodynamic_Solver.cpp:83-115,177-203
//   solve + reductions      src/Solver/Hydrodynamic_Solver.cpp:117-175,400-477,544-596
//   dense LU                Eigen::PartialPivLU (Solver_Base.h:85): restated as a
//                           blocked right-looking LU with partial (row) pivoting on
//                           max |a_ik| (first maximum wins).
//
// Deliberate restatement choices (rounding-level only):
//   * Gauss-Legendre order 2 nodes/weights are the closed form +-1/sqrt(3), w=1
//     (the reference computes them by Golub-Welsch, Quadrule.cpp:14).
//   * Eigen Quaternion::FromTwoVectors for exactly antiparallel vectors takes its
//     rotation axis from a JacobiSVD null vector; here that axis is +x.
//   * polynomial/operation order follows the reference text; no FMA contraction.
//
// Parity status of each function is listed in the header of tests/test_oracle.py.

#include <cmath>
#include <tr1/cmath>
#include <complex>
#include <vector>
#include <cstring>
#include <cstdint>
#include <algorithm>
#include <omp.h>

typedef std::complex<double> cd;

namespace {

const double EUL = 0.5772156649;          // BEMUse_Math_Types.h:99 (truncated literal)
const double PI = M_PI;
const double TwoPIinv = 0.5 / M_PI;
const double FourPIinv = 0.25 / M_PI;
const double Gravity = 9.81;              // BEMUse_Math_Types.h:115
const double BCF = 1.0e-5;                // Solver_Base.h:55

struct V3 { double x, y, z; };
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 scl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(V3 a) { double z = dot(a, a); if (z > 0) { double n = std::sqrt(z); return {a.x / n, a.y / n, a.z / n}; } return a; }

struct Quat { double w; V3 v; };

// Eigen Quaternion::FromTwoVectors(UnitZ, b) (used at Geo_Elements.cpp:66,111,144,205).
Quat from_two_vectors_z(V3 b) {
    V3 v0 = {0.0, 0.0, 1.0};
    V3 v1 = normalized(b);
    double c = dot(v1, v0);
    if (c < -1.0 + 1e-12) {
        c = std::max(c, -1.0);
        V3 axis = {1.0, 0.0, 0.0};
        double w2 = (1.0 + c) * 0.5;
        return {std::sqrt(w2), scl(axis, std::sqrt(1.0 - w2))};
    }
    V3 axis = cross(v0, v1);
    double s = std::sqrt((1.0 + c) * 2.0);
    double invs = 1.0 / s;
    return {s * 0.5, scl(axis, invs)};
}

// Eigen Quaternion * Vector3 (_transformVector): v + w*uv + vec x uv, uv = 2 vec x v.
V3 qrot(const Quat& q, V3 v) {
    V3 uv = cross(q.v, v);
    uv = add(uv, uv);
    return add(add(v, scl(uv, q.w)), cross(q.v, uv));
}

struct Panel {
    int nv;            // 3 or 4 vertices
    V3 P[4];           // global vertex positions
    V3 C;              // centroid
    double area;
    V3 ex, ey, ez;     // rows of OrientMat_Global (CoordSys.h:87-95)
    int nq;
    V3 qp[4], qn[4];   // quadrature node positions / unit normals (Z_Axis_Global)
    double qw[4];      // weight = Jacobian * rule weight
    double lx[4], ly[4]; // vertex coordinates in the panel frame
};

// Tri_Element::Set_Centroid / Set_Quad_Nodes, Geo_Elements.cpp:45-119.
void build_tri(Panel& p) {
    V3 P0 = p.P[0], P1 = p.P[1], P2 = p.P[2];
    p.C = scl(add(add(P0, P1), P2), 1.0 / 3.0);
    V3 X1 = sub(P0, scl(add(P1, P2), 0.5));
    V3 X2 = sub(P2, P1);
    V3 X3 = cross(X1, X2);
    p.area = 0.5 * norm(X3);
    X3 = normalized(X3);
    Quat q = from_two_vectors_z(X3);
    p.ex = qrot(q, {1, 0, 0}); p.ey = qrot(q, {0, 1, 0}); p.ez = qrot(q, {0, 0, 1});
    // Dunavant rule 2 (Quadrule.cpp:26177-26200): truncated 15-digit literals.
    const double a = 0.666666666666667, b = 0.166666666666667, w = 0.333333333333333;
    const double bary[3] = {a, b, b};
    p.nq = 3;
    for (int k = 0; k < 3; k++) {
        double xi = bary[k], eta = bary[(k + 1) % 3];
        double W = w * 0.5;
        V3 P = {0, 0, 0};
        P = add(P, scl(P0, 1. - xi - eta));
        P = add(P, scl(P1, eta));
        P = add(P, scl(P2, xi));
        V3 Rxi = add(scl(P0, -1), scl(P2, 1));
        V3 Reta = add(scl(P0, -1), scl(P1, 1));
        V3 Rn = cross(Rxi, Reta);
        double Jac = norm(Rn);
        Rn = normalized(Rn);
        Quat qq = from_two_vectors_z(Rn);
        p.qp[k] = P;
        p.qn[k] = qrot(qq, {0, 0, 1});
        p.qw[k] = Jac * W;
    }
}

// Quad_Element::Set_Centroid / Set_Quad_Nodes, Geo_Elements.cpp:123-214.
void build_quad(Panel& p) {
    V3 P0 = p.P[0], P1 = p.P[1], P2 = p.P[2], P3 = p.P[3];
    p.C = scl(add(add(add(P0, P1), P2), P3), 0.25);
    V3 X1 = sub(scl(add(P3, P2), 0.5), scl(add(P0, P1), 0.5));
    V3 X2 = sub(scl(add(P2, P1), 0.5), scl(add(P3, P0), 0.5));
    V3 X3 = normalized(cross(X1, X2));
    V3 C1 = sub(P2, P0), C2 = sub(P3, P1);
    p.area = 0.5 * norm(cross(C1, C2));
    Quat q = from_two_vectors_z(X3);
    p.ex = qrot(q, {1, 0, 0}); p.ey = qrot(q, {0, 1, 0}); p.ez = qrot(q, {0, 0, 1});
    const double g = 1.0 / std::sqrt(3.0);
    const double PX[2] = {-g, g}, WX[2] = {1.0, 1.0};
    p.nq = 4;
    int k = 0;
    for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) {
        double xi = PX[i], eta = PX[j];
        V3 P = {0, 0, 0};
        P = add(P, scl(P0, 0.25 * (1 - xi) * (1 - eta)));
        P = add(P, scl(P1, 0.25 * (1 - xi) * (1 + eta)));
        P = add(P, scl(P2, 0.25 * (1 + xi) * (1 + eta)));
        P = add(P, scl(P3, 0.25 * (1 + xi) * (1 - eta)));
        V3 Rxi = {0, 0, 0};
        Rxi = add(Rxi, scl(P0, -0.25 * (1 - eta)));
        Rxi = add(Rxi, scl(P1, -0.25 * (1 + eta)));
        Rxi = add(Rxi, scl(P2, 0.25 * (1 + eta)));
        Rxi = add(Rxi, scl(P3, 0.25 * (1 - eta)));
        V3 Reta = {0, 0, 0};
        Reta = add(Reta, scl(P0, -0.25 * (1 - xi)));
        Reta = add(Reta, scl(P1, 0.25 * (1 - xi)));
        Reta = add(Reta, scl(P2, 0.25 * (1 + xi)));
        Reta = add(Reta, scl(P3, -0.25 * (1 + xi)));
        V3 Rn = cross(Rxi, Reta);
        double Jac = norm(Rn);
        Rn = normalized(Rn);
        Quat qq = from_two_vectors_z(Rn);
        p.qp[k] = P;
        p.qn[k] = qrot(qq, {0, 0, 1});
        p.qw[k] = Jac * WX[i] * WX[j];
        k++;
    }
}

void finish_panel(Panel& p) {
    for (int k = 0; k < p.nv; k++) {
        V3 d = sub(p.P[k], p.C);
        p.lx[k] = dot(p.ex, d);
        p.ly[k] = dot(p.ey, d);
    }
}

// FlatTriPanel / FlatQuadPanel::Inf_SourceKernel, Panel.cpp:17-156 (edge loop form).
void rankine_sd(const Panel& p, V3 P, double& S, double& D) {
    V3 d0 = sub(P, p.C);
    double x = dot(p.ex, d0), y = dot(p.ey, d0), z = dot(p.ez, d0);
    int n = p.nv;
    double m[4], e[4], h[4], r[4], DX[4], DY[4];
    for (int k = 0; k < n; k++) {
        int k1 = (k + 1) % n;
        m[k] = (p.ly[k1] - p.ly[k]) / (p.lx[k1] - p.lx[k]);
        DX[k] = x - p.lx[k];
        DY[k] = y - p.ly[k];
    }
    double Z2 = z * z;
    for (int k = 0; k < n; k++) {
        e[k] = DX[k] * DX[k] + Z2;
        h[k] = DX[k] * DY[k];
        r[k] = std::sqrt(DX[k] * DX[k] + DY[k] * DY[k] + Z2);
    }
    double PhiDip = 0;
    for (int k = 0; k < n; k++) {
        int k1 = (k + 1) % n;
        PhiDip += std::atan((m[k] * e[k] - h[k]) / (z * r[k])) - std::atan((m[k] * e[k1] - h[k1]) / (z * r[k1]));
    }
    double BL = 0;
    for (int k = 0; k < n; k++) {
        int k1 = (k + 1) % n;
        double ddx = p.lx[k1] - p.lx[k], ddy = p.ly[k1] - p.ly[k];
        double dk = std::sqrt(ddx * ddx + ddy * ddy);
        double Bk = (DX[k] * ddy - DY[k] * ddx) / dk;
        double Lk = std::log((r[k] + r[k1] + dk) / (r[k] + r[k1] - dk));
        BL += Bk * Lk;
    }
    D = PhiDip * FourPIinv;
    S = (z * PhiDip - BL) * FourPIinv;
}

// FlatSource*Panel::Inf_SingleDoubleLayerQuad with Source_Kernel / Dipole_Kernel
// (Panel.cpp:156-170,582-593; Kernels.cpp:11-27).
void rankine_sd_quad(const Panel& p, V3 P, double& S, double& D) {
    double So = 0, Do = 0;
    for (int q = 0; q < p.nq; q++) {
        V3 r = sub(P, p.qp[q]);
        double RN = norm(r);
        double sk = (RN != 0) ? -FourPIinv / RN : 0.0;
        double dk = (RN != 0) ? dot(p.qn[q], r) / RN / RN / RN * FourPIinv : 0.0;
        So += p.qw[q] * sk;
        Do += p.qw[q] * dk;
    }
    S = So; D = Do;
}

// ---- Kernels.cpp:31-116 ---------------------------------------------------
inline double BesselJ0(double x) { return std::tr1::cyl_bessel_j(0.0, x); }
inline double BesselJ1(double x) { return std::tr1::cyl_bessel_j(1.0, x); }
inline double BesselY0(double x) { return std::tr1::cyl_neumann(0.0, x); }
inline double BesselY1(double x) { return std::tr1::cyl_neumann(1.0, x); }

// Struve approximation literals (Kernels.cpp:44-49,56-63 for H0; :84-89,96-103
// for H1), in the reference's order: P0..P5, a0..a3, b1..b3 of StruveH0, then
// P1..P6, a0..a3, b1..b3 of StruveH1.
const double kStruveTab[26] = {
    +1.909859164, -1.909855001, +0.687514637, -0.126164557, +0.013828813, -0.000876918,
    0.99999906, 4.77228920, 3.85542044, 0.32303607,
    4.88331068, 4.28957333, 0.52120508,
    +1.909859286, -1.145914713, +0.294656958, -0.042070508, +0.003785727, -0.000207183,
    1.00000004, 3.92205313, 2.64893033, 0.27450895,
    3.81095112, 2.26216956, 0.10885141};

double StruveH0(double xx) {
    const double* P = kStruveTab;
    const double *a = kStruveTab + 6, *b = kStruveTab + 10;
    double SH0;
    if (xx <= 3.0) {
        double yy = (xx / 3.0) * (xx / 3.0);
        SH0 = P[0] + (P[1] + (P[2] + (P[3] + (P[4] + P[5] * yy) * yy) * yy) * yy) * yy;
        SH0 *= xx / 3.0;
    } else {
        double yy = (3.0 / xx) * (3.0 / xx);
        double c1 = 2.0 * (a[0] + (a[1] + (a[2] + a[3] * yy) * yy) * yy);
        double c2 = PI * xx * (1.0 + (b[0] + (b[1] + b[2] * yy) * yy) * yy);
        SH0 = c1 / c2 + BesselY0(xx);
    }
    return SH0;
}

double StruveH1(double xx) {
    const double* P = kStruveTab + 13;
    const double *a = kStruveTab + 19, *b = kStruveTab + 23;
    double SH1;
    if (xx <= 3.0) {
        double yy = (xx / 3.0) * (xx / 3.0);
        SH1 = (P[0] + (P[1] + (P[2] + (P[3] + (P[4] + P[5] * yy) * yy) * yy) * yy) * yy) * yy;
    } else {
        double yy = (3.0 / xx) * (3.0 / xx);
        double c1 = 2.0 * (a[0] + (a[1] + (a[2] + a[3] * yy) * yy) * yy);
        double c2 = PI * (1.0 + (b[0] + (b[1] + b[2] * yy) * yy) * yy);
        SH1 = c1 / c2 + BesselY1(xx);
    }
    return SH1;
}

// Polynomials of the global approximation (Wu et al. 2017 eq. 14; Liang et al.
// 2018 eq. 16), evaluated by running powers as at Kernels.cpp:120-256.
double run_poly(const double* c, int n, double tt) {
    double S = c[0], t = tt;
    for (int i = 1; i < n; i++) { S += c[i] * t; t *= tt; }
    return S;
}
const double cA[10] = {1.21, -13.328, 215.896, -1763.96, 8418.94, -24314.21, 42002.57, -41592.9, 21859.0, -4838.6};
const double cB[10] = {0.938, 5.737, -67.92, 796.534, -4780.77, 17137.74, -36618.81, 44894.06, -29030.24, 7671.22};
const double cC[8] = {1.268, -9.747, 209.653, -1397.89, 5155.67, -9844.35, 9136.4, -3272.62};
const double cD[10] = {0.632, -40.97, 667.16, -6072.07, 31127.39, -96293.05, 181856.75, -205690.43, 128170.2, -33744.6};
const double cdA[10] = {2.948, -24.53, 249.69, -754.85, -1187.71, 16370.75, -48811.41, 68220.87, -46688.0, 12622.25};
const double cdB[10] = {1.11, 2.894, -76.765, 1565.35, -11336.19, 44270.15, -97014.11, 118879.26, -76209.82, 19923.28};
const double cdC[6] = {14.19, -148.24, 847.8, -2318.58, 3168.35, -1590.27};

// HavelockGF, Kernels.cpp:342-412 (with GF_Func_L0/W/Ls/Wh :258-338).
void havelock(double dx, double dy, double vv, cd& GF, cd& DGFX, cd& DGFY, cd& DGFZ) {
    double hh = std::sqrt(dx * dx + dy * dy);
    double dd = std::sqrt(hh * hh + vv * vv);
    double al = -vv / dd;
    double be = hh / dd;
    double rho = dd / (1.0 + dd);
    double dd3 = std::pow(dd, 3);
    double omr3 = std::pow(1.0 - rho, 3);
    double ev = std::exp(vv);
    // L0
    double PP = std::log(0.5 * (dd - vv)) + EUL - 2.0 * dd * dd;
    PP *= ev;
    PP += dd * dd - vv;
    double A = run_poly(cA, 10, rho), B = run_poly(cB, 10, rho), C = run_poly(cC, 8, rho), D = run_poly(cD, 10, rho);
    double RR = (1.0 - be) * A;
    RR -= be * B;
    RR -= al * C / (1.0 + 6.0 * al * rho * (1.0 - rho));
    RR += be * (1.0 - be) * D;
    double Lp = rho * omr3 * RR;
    double L0 = 2.0 * PP / (1.0 + dd3) + 2.0 * Lp;
    // W
    double H0 = StruveH0(hh), J0 = BesselJ0(hh);
    double E = 2.0 * PI * ev;
    cd W(E * H0, -E * J0);
    GF = cd(L0, 0.0) + W;
    // Ls
    double PS = (be + hh) / (dd - vv);
    PS -= 2.0 * be;
    PS += 2.0 * dd * ev;
    PS -= hh;
    double QS = std::exp(-dd) * (1.0 - be);
    QS *= 1.0 + dd / (1.0 + dd3);
    double dA = run_poly(cdA, 10, rho), dB = run_poly(cdB, 10, rho), dC = run_poly(cdC, 6, rho);
    double R2 = be * dA;
    R2 -= (1.0 - al) * dB;
    R2 += be * (1.0 - be) * rho * (1.0 - 2.0 * rho) * dC;
    double Lsp = rho * omr3 * R2;
    double Ls = 2.0 * PS / (1.0 + dd3) - 4.0 * QS + 2.0 * Lsp;
    double H1 = StruveH1(hh), J1 = BesselJ1(hh);
    cd Wh(E * (2.0 / PI - H1), E * J1);
    cd GFh = cd(Ls, 0.0) + Wh;
    if (hh > 1.0e-6) {
        DGFX = GFh * dx / hh;
        DGFY = GFh * dy / hh;
    } else {
        DGFX = cd(0, 0);
        DGFY = cd(0, 0);
    }
    DGFZ = GF;
}

// FreeSurface_Wave_Kernel, Panel.cpp:601-636.
inline void fs_wave_kernel(V3 Pg, V3 Ps, double K, cd& G, cd& dX, cd& dY, cd& dZ) {
    if (K == 0.0) { G = dX = dY = dZ = cd(0, 0); return; }
    V3 R = sub(Pg, Ps);
    havelock(R.x * K, R.y * K, R.z * K, G, dX, dY, dZ);
    G *= K * FourPIinv;
    dX *= K * K * FourPIinv;
    dY *= K * K * FourPIinv;
    dZ *= K * K * FourPIinv;
}

// FlatWave*Panel::CInf_SingleDoubleLayerQuad, Panel.cpp:676-694,776-794.
inline void wave_sd(const Panel& p, V3 P, double K, cd& S, cd& D) {
    cd OS(0, 0), OD(0, 0);
    for (int q = 0; q < p.nq; q++) {
        cd G, dX, dY, dZ;
        fs_wave_kernel(P, p.qp[q], K, G, dX, dY, dZ);
        OS += p.qw[q] * G;
        OD += p.qw[q] * (dX * p.qn[q].x + dY * p.qn[q].y + dZ * p.qn[q].z);
    }
    S = OS; D = OD;
}

struct Body {
    int N;
    int NPA = 0;                    // body panels; N - NPA lid panels (IFR) follow them
    std::vector<Panel> body, refl;
    std::vector<V3> bc;             // BC positions C + BCF*n
    std::vector<double> Ssrc, Srefl, Dsrc, Drefl;   // N x N col-major (R,S)
    std::vector<double> reflnorm;   // 2 n_z(S)
    bool have_rankine = false;
};

// ---- dense complex LU with partial pivoting (Eigen::PartialPivLU semantics) --
void lu_factor(int n, cd* A, int lda, int* piv) {
    const int nb = 64;
    for (int k0 = 0; k0 < n; k0 += nb) {
        int kb = std::min(nb, n - k0);
        // unblocked panel on columns k0..k0+kb, full row swaps
        for (int k = k0; k < k0 + kb; k++) {
            int p = k; double best = std::abs(A[k + (size_t)k * lda]);
            for (int i = k + 1; i < n; i++) {
                double a = std::abs(A[i + (size_t)k * lda]);
                if (a > best) { best = a; p = i; }
            }
            piv[k] = p;
            if (p != k) for (int j = 0; j < n; j++) std::swap(A[k + (size_t)j * lda], A[p + (size_t)j * lda]);
            cd akk = A[k + (size_t)k * lda];
            if (akk != cd(0, 0)) {
                cd inv = 1.0 / akk;
                for (int i = k + 1; i < n; i++) A[i + (size_t)k * lda] *= inv;
            }
            // rank-1 update inside the panel only
            for (int j = k + 1; j < k0 + kb; j++) {
                cd u = A[k + (size_t)j * lda];
                double ur = u.real(), ui = u.imag();
                double* col = reinterpret_cast<double*>(A + (size_t)j * lda);
                const double* l = reinterpret_cast<const double*>(A + (size_t)k * lda);
                for (int i = k + 1; i < n; i++) {
                    double lr = l[2 * i], li = l[2 * i + 1];
                    col[2 * i] -= lr * ur - li * ui;
                    col[2 * i + 1] -= lr * ui + li * ur;
                }
            }
        }
        int j0 = k0 + kb;
        if (j0 >= n) break;
        // U12 = L11^-1 A12 and trailing update A22 -= L21 U12, parallel over columns
        #pragma omp parallel for schedule(dynamic, 4)
        for (int j = j0; j < n; j++) {
            double* col = reinterpret_cast<double*>(A + (size_t)j * lda);
            for (int k = k0; k < k0 + kb; k++) {
                double ur = col[2 * k], ui = col[2 * k + 1];
                const double* l = reinterpret_cast<const double*>(A + (size_t)k * lda);
                for (int i = k + 1; i < n; i++) {
                    double lr = l[2 * i], li = l[2 * i + 1];
                    col[2 * i] -= lr * ur - li * ui;
                    col[2 * i + 1] -= lr * ui + li * ur;
                }
            }
        }
    }
}

void lu_solve(int n, const cd* LU, int lda, const int* piv, cd* B, int ldb, int nrhs) {
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < nrhs; c++) {
        cd* b = B + (size_t)c * ldb;
        for (int k = 0; k < n; k++) if (piv[k] != k) std::swap(b[k], b[piv[k]]);
        for (int k = 0; k < n; k++) {            // L y = b (unit lower)
            cd bk = b[k];
            const cd* l = LU + (size_t)k * lda;
            for (int i = k + 1; i < n; i++) b[i] -= l[i] * bk;
        }
        for (int k = n - 1; k >= 0; k--) {       // U x = y
            const cd* u = LU + (size_t)k * lda;
            b[k] /= u[k];
            cd bk = b[k];
            for (int i = 0; i < k; i++) b[i] -= u[i] * bk;
        }
    }
}

}  // namespace

extern "C" {

// ---- scalar kernels -------------------------------------------------------
void orc_bessel(int n, const double* x, double* j0, double* j1, double* y0, double* y1) {
    for (int i = 0; i < n; i++) {
        j0[i] = BesselJ0(x[i]); j1[i] = BesselJ1(x[i]);
        y0[i] = x[i] > 0 ? BesselY0(x[i]) : NAN; y1[i] = x[i] > 0 ? BesselY1(x[i]) : NAN;
    }
}

void orc_struve(int n, const double* x, double* h0, double* h1) {
    for (int i = 0; i < n; i++) { h0[i] = StruveH0(x[i]); h1[i] = StruveH1(x[i]); }
}

// out: 8 doubles per point: GF re/im, DGFX re/im, DGFY re/im, DGFZ re/im.
void orc_havelock(int n, const double* X, const double* Y, const double* V, double* out) {
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        cd G, dx, dy, dz;
        havelock(X[i], Y[i], V[i], G, dx, dy, dz);
        double* o = out + 8 * (size_t)i;
        o[0] = G.real(); o[1] = G.imag(); o[2] = dx.real(); o[3] = dx.imag();
        o[4] = dy.real(); o[5] = dy.imag(); o[6] = dz.real(); o[7] = dz.imag();
    }
}

void orc_struve_coeffs(double* out /*26*/) { std::memcpy(out, kStruveTab, sizeof(kStruveTab)); }

void orc_poly_coeffs(double* out /*65: 64 coefficients, then EUL*/) {
    int k = 0;
    for (double c : cA) out[k++] = c;
    for (double c : cB) out[k++] = c;
    for (double c : cC) out[k++] = c;
    for (double c : cD) out[k++] = c;
    for (double c : cdA) out[k++] = c;
    for (double c : cdB) out[k++] = c;
    for (double c : cdC) out[k++] = c;
    out[k++] = EUL;
}

// ---- body -----------------------------------------------------------------
// nodes: nn x 3; elems: ne x 4 (elems[4e+3] < 0 for a triangle).
void* orc_body_create(int nn, const double* nodes, int ne, const int* elems) {
    Body* b = new Body;
    b->N = ne;
    b->NPA = ne;
    b->body.resize(ne); b->refl.resize(ne); b->bc.resize(ne);
    for (int e = 0; e < ne; e++) {
        const int* el = elems + 4 * e;
        int nv = el[3] < 0 ? 3 : 4;
        Panel& p = b->body[e];
        Panel& r = b->refl[e];
        p.nv = r.nv = nv;
        for (int k = 0; k < nv; k++) {
            const double* X = nodes + 3 * el[k];
            p.P[k] = {X[0], X[1], X[2]};
            // Generate_Symmetry_Elements: z -> -z and node order reversed.
            const double* Xr = nodes + 3 * el[nv - 1 - k];
            r.P[k] = {Xr[0], Xr[1], -Xr[2]};
        }
        if (nv == 3) { build_tri(p); build_tri(r); } else { build_quad(p); build_quad(r); }
        finish_panel(p); finish_panel(r);
        b->bc[e] = add(p.C, scl(p.ez, BCF));   // Solver_Base.cpp:71-83
    }
    return b;
}

void orc_body_destroy(void* h) { delete static_cast<Body*>(h); }
// The last N - npa elements are the auxiliary lid panels of irregular-frequency
// removal (Create_Panels, Hydrodynamic_Solver.cpp:45-55): they join the
// system; post-processing uses the npa body panels (:145-165, 450-458).
void orc_body_set_npa(void* h, int npa) {
    Body* b = static_cast<Body*>(h);
    if (npa > 0 && npa <= b->N) b->NPA = npa;
}
int orc_body_n(void* h) { return static_cast<Body*>(h)->N; }

// centroid(3) normal(3) area bc(3) for body panels; qnodes of the MIRRORED panels.
void orc_body_geometry(void* h, double* cen, double* nrm, double* area, double* bc,
                       int* nq, double* qpos, double* qnrm, double* qw,
                       double* frame /*N x 9 rows ex,ey,ez*/) {
    Body* b = static_cast<Body*>(h);
    for (int e = 0; e < b->N; e++) {
        const Panel& p = b->body[e];
        const Panel& r = b->refl[e];
        cen[3 * e] = p.C.x; cen[3 * e + 1] = p.C.y; cen[3 * e + 2] = p.C.z;
        nrm[3 * e] = p.ez.x; nrm[3 * e + 1] = p.ez.y; nrm[3 * e + 2] = p.ez.z;
        area[e] = p.area;
        bc[3 * e] = b->bc[e].x; bc[3 * e + 1] = b->bc[e].y; bc[3 * e + 2] = b->bc[e].z;
        nq[e] = r.nq;
        for (int q = 0; q < 4; q++) {
            bool ok = q < r.nq;
            qpos[12 * e + 3 * q] = ok ? r.qp[q].x : 0; qpos[12 * e + 3 * q + 1] = ok ? r.qp[q].y : 0; qpos[12 * e + 3 * q + 2] = ok ? r.qp[q].z : 0;
            qnrm[12 * e + 3 * q] = ok ? r.qn[q].x : 0; qnrm[12 * e + 3 * q + 1] = ok ? r.qn[q].y : 0; qnrm[12 * e + 3 * q + 2] = ok ? r.qn[q].z : 0;
            qw[4 * e + q] = ok ? r.qw[q] : 0;
        }
        if (frame) {
            double* f = frame + 9 * e;
            f[0] = p.ex.x; f[1] = p.ex.y; f[2] = p.ex.z; f[3] = p.ey.x; f[4] = p.ey.y; f[5] = p.ey.z;
            f[6] = p.ez.x; f[7] = p.ez.y; f[8] = p.ez.z;
        }
    }
}

// Body-panel quadrature nodes (not mirrored), for invariant tests.
void orc_body_qnodes(void* h, int* nq, double* qpos, double* qnrm, double* qw) {
    Body* b = static_cast<Body*>(h);
    for (int e = 0; e < b->N; e++) {
        const Panel& p = b->body[e];
        nq[e] = p.nq;
        for (int q = 0; q < 4; q++) {
            bool ok = q < p.nq;
            for (int c = 0; c < 3; c++) {
                qpos[12 * e + 3 * q + c] = ok ? (&p.qp[q].x)[c] : 0;
                qnrm[12 * e + 3 * q + c] = ok ? (&p.qn[q].x)[c] : 0;
            }
            qw[4 * e + q] = ok ? p.qw[q] : 0;
        }
    }
}

// Analytic (mode 0) or quadrature (mode 1) Rankine S/D of panel `s`
// (mirrored panel if refl) at arbitrary points.
void orc_rankine_points(void* h, int s, int refl, int mode, int np, const double* pts, double* S, double* D) {
    Body* b = static_cast<Body*>(h);
    const Panel& p = refl ? b->refl[s] : b->body[s];
    for (int i = 0; i < np; i++) {
        V3 P = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        if (mode == 0) rankine_sd(p, P, S[i], D[i]); else rankine_sd_quad(p, P, S[i], D[i]);
    }
}

// Wave S/D of mirrored panel s at arbitrary points (CInf_SingleDoubleLayerQuad).
void orc_wave_points(void* h, int s, double K, int np, const double* pts, double* S /*2np*/, double* D /*2np*/) {
    Body* b = static_cast<Body*>(h);
    for (int i = 0; i < np; i++) {
        cd ws, wd;
        wave_sd(b->refl[s], {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, K, ws, wd);
        S[2 * i] = ws.real(); S[2 * i + 1] = ws.imag(); D[2 * i] = wd.real(); D[2 * i + 1] = wd.imag();
    }
}

// Prepare_Linear_System, Hydrodynamic_Solver.cpp:83-115. Outputs (optional)
// SSrc=-s, SSrcRefl=-sr, DSrc=d, DSrcRefl=dr as real N x N column-major.
void orc_rankine(void* h, double* Ssrc, double* Srefl, double* Dsrc, double* Drefl, double* reflnorm) {
    Body* b = static_cast<Body*>(h);
    int N = b->N;
    if (!b->have_rankine) {
        size_t NN = (size_t)N * N;
        b->Ssrc.resize(NN); b->Srefl.resize(NN); b->Dsrc.resize(NN); b->Drefl.resize(NN);
        b->reflnorm.resize(N);
        #pragma omp parallel for schedule(dynamic, 8)
        for (int S = 0; S < N; S++) {
            b->reflnorm[S] = b->body[S].ez.z * 2.0;
            for (int R = 0; R < N; R++) {
                double s, d, sr, dr;
                rankine_sd(b->body[S], b->bc[R], s, d);
                rankine_sd(b->refl[S], b->bc[R], sr, dr);
                size_t idx = (size_t)S * N + R;
                b->Ssrc[idx] = -s; b->Srefl[idx] = -sr; b->Dsrc[idx] = d; b->Drefl[idx] = dr;
            }
        }
        b->have_rankine = true;
    }
    size_t NN = (size_t)N * N;
    if (Ssrc) std::memcpy(Ssrc, b->Ssrc.data(), NN * 8);
    if (Srefl) std::memcpy(Srefl, b->Srefl.data(), NN * 8);
    if (Dsrc) std::memcpy(Dsrc, b->Dsrc.data(), NN * 8);
    if (Drefl) std::memcpy(Drefl, b->Drefl.data(), NN * 8);
    if (reflnorm) std::memcpy(reflnorm, b->reflnorm.data(), N * 8);
}

// Prepare_Linear_System_Wave_Terms, Hydrodynamic_Solver.cpp:177-203.
// SWave, DWave: complex N x N column-major, interleaved (re, im).
void orc_wave(void* h, double K, double* SWave, double* DWave) {
    Body* b = static_cast<Body*>(h);
    orc_rankine(h, nullptr, nullptr, nullptr, nullptr, nullptr);
    int N = b->N;
    #pragma omp parallel for schedule(dynamic, 4)
    for (int S = 0; S < N; S++) {
        for (int R = 0; R < N; R++) {
            cd ws, wd;
            wave_sd(b->refl[S], b->bc[R], K, ws, wd);
            size_t idx = (size_t)S * N + R;
            cd sw = -ws;
            cd dw = wd + cd(b->Srefl[idx], 0.0) * cd(b->reflnorm[S], 0.0) * K;
            SWave[2 * idx] = sw.real(); SWave[2 * idx + 1] = sw.imag();
            DWave[2 * idx] = dw.real(); DWave[2 * idx + 1] = dw.imag();
        }
    }
}

// Dense LU + solve on a caller matrix (for LAPACK cross-checks).
// A: n x n complex interleaved column-major, overwritten by LU; piv: n.
void orc_lu_factor(int n, double* A, int* piv) { lu_factor(n, reinterpret_cast<cd*>(A), n, piv); }
void orc_lu_solve(int n, const double* LU, const int* piv, int nrhs, double* B) {
    lu_solve(n, reinterpret_cast<const cd*>(LU), n, piv, reinterpret_cast<cd*>(B), n, nrhs);
}

// Hydrodynamic_Radiation_Solver::Solve (Hydrodynamic_Solver.cpp:400-477) for
// one frequency. omega is the solver's internal Omega (= Frequency/2pi,
// :405), kappa its wavenumber (:409). compose: 0 = wave (:427-428),
// 1 = Omega==0 (:417-418), 2 = Omega==FreqInf (:422-423). rad_only: return
// after A/B (:469). Outputs: A36, B36 (6x6 row-major), F, FK, SC (6 x nb
// complex, row-major, interleaved), optional PhiJ (N x 6) / PhiS (N x nb),
// column-major interleaved. timings (optional, 3 doubles): assemble, factor, rest [s].
// Area factor of the diffraction right-hand side (Hydrodynamic_Solver.cpp:445:
// -DPhi_I_DN o PanAreaLong; the radiation right-hand side at :437 has none).
// 1 (default) restates the reference; 0 solves the physical scattering problem
// D Phi_S = S (-dPhi_I/dn), which the tests use to pin the diffraction solve
// by the Haskind relation (direct pressure integration of Phi_I + Phi_S equals
// the excitation from Phi_J).
static int g_diff_area = 1;
void orc_set_diffraction_area(int on) { g_diff_area = on ? 1 : 0; }

int orc_solve(void* h, double omega, double kappa, int compose, int rad_only,
              int nb, const double* betas, double* A36, double* B36,
              double* F, double* FK, double* SC, double* PhiJ, double* PhiS, double* timings) {
    Body* b = static_cast<Body*>(h);
    int N = b->N;
    double t0 = omp_get_wtime();
    orc_rankine(h, nullptr, nullptr, nullptr, nullptr, nullptr);
    size_t NN = (size_t)N * N;
    std::vector<cd> SMat(NN), DMat(NN);
    bool wave = (compose == 0) && (omega != 0.0);
    #pragma omp parallel for schedule(dynamic, 4)
    for (int S = 0; S < N; S++) {
        for (int R = 0; R < N; R++) {
            size_t idx = (size_t)S * N + R;
            if (compose == 1 || (compose == 0 && omega == 0.0)) {
                SMat[idx] = b->Ssrc[idx] + b->Srefl[idx];
                DMat[idx] = b->Dsrc[idx] + b->Drefl[idx];
            } else if (compose == 2) {
                SMat[idx] = b->Ssrc[idx] - b->Srefl[idx];
                DMat[idx] = b->Dsrc[idx] - b->Drefl[idx];
            } else {
                cd ws, wd;
                wave_sd(b->refl[S], b->bc[R], kappa, ws, wd);
                cd sw = -ws;
                cd dw = wd + cd(b->Srefl[idx], 0.0) * cd(b->reflnorm[S], 0.0) * kappa;
                SMat[idx] = cd(b->Ssrc[idx], 0.0) + cd(b->Srefl[idx], 0.0) + sw;
                DMat[idx] = cd(b->Dsrc[idx], 0.0) + cd(b->Drefl[idx], 0.0) + dw;
            }
        }
    }
    (void)wave;
    double t1 = omp_get_wtime();
    std::vector<int> piv(N);
    lu_factor(N, DMat.data(), N, piv.data());
    double t2 = omp_get_wtime();
    // Set_RHS_Mat (:117-138), Prepare_PostProcessing_Mats (:140-175),
    // Set_Incident_Potential_Mats (:575-596), all at centroids.
    int nr = 6 + nb;
    std::vector<cd> Xin((size_t)N * nr), RHS((size_t)N * nr);
    std::vector<double> M((size_t)6 * N);
    std::vector<cd> PhiI((size_t)N * nb), DPhiI((size_t)N * nb);
    const cd Im(0.0, 1.0);
    for (int i = 0; i < N; i++) {
        V3 P = b->body[i].C, n = b->body[i].ez;
        V3 pcn = cross(P, n);
        double nk[6] = {n.x, n.y, n.z, pcn.x, pcn.y, pcn.z};
        double dS = b->body[i].area;
        for (int j = 0; j < 6; j++) { Xin[(size_t)j * N + i] = nk[j]; M[(size_t)j * N + i] = nk[j] * dS; }
        for (int bb = 0; bb < nb; bb++) {
            double CB = std::cos(betas[bb]), SB = std::sin(betas[bb]);
            cd phi = Im * Gravity / omega * std::exp(Im * kappa * (P.x * CB + P.y * SB));
            double Z = std::exp(kappa * P.z);
            cd pi = Z * phi;
            cd dpx = Im * kappa * CB * pi, dpy = Im * kappa * SB * pi, dpz = kappa * pi;
            PhiI[(size_t)bb * N + i] = pi;
            cd dn = n.x * dpx + n.y * dpy + n.z * dpz;
            DPhiI[(size_t)bb * N + i] = dn;
            Xin[(size_t)(6 + bb) * N + i] = g_diff_area ? -(dn * dS) : -dn;   // -DPhi_I_DN o PanAreaLong (:445)
        }
    }
    // RHS = SMat * Xin (:437, :445)
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < nr; c++) {
        cd* r = RHS.data() + (size_t)c * N;
        const cd* x = Xin.data() + (size_t)c * N;
        for (int S = 0; S < N; S++) {
            cd xs = x[S];
            if (xs == cd(0, 0)) continue;
            const cd* col = SMat.data() + (size_t)S * N;
            for (int R = 0; R < N; R++) r[R] += col[R] * xs;
        }
    }
    lu_solve(N, DMat.data(), N, piv.data(), RHS.data(), N, nr);
    const cd* PJ = RHS.data();
    const cd* PS = RHS.data() + (size_t)6 * N;
    // A = Re(M PhiJ), B = Im(M PhiJ) (:463-465), over the body panels
    const int NPA = b->NPA;
    for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) {
        cd s(0, 0);
        for (int k = 0; k < NPA; k++) s += M[(size_t)i * N + k] * PJ[(size_t)j * N + k];
        A36[6 * i + j] = s.real(); B36[6 * i + j] = s.imag();
    }
    if (!rad_only) {
        for (int i = 0; i < 6; i++) for (int bb = 0; bb < nb; bb++) {
            cd f1(0, 0), f2(0, 0);
            for (int k = 0; k < NPA; k++) {
                f1 += M[(size_t)i * N + k] * PhiI[(size_t)bb * N + k];
                f2 += PJ[(size_t)i * N + k] * (DPhiI[(size_t)bb * N + k] * b->body[k].area);
            }
            cd fk = -Im * omega * f1, sc = Im * omega * f2;
            cd f = std::conj(fk + sc);
            if (omega == 0.0) { f = 0; fk = 0; }
            size_t o = 2 * ((size_t)i * nb + bb);
            F[o] = f.real(); F[o + 1] = f.imag();
            if (FK) { FK[o] = fk.real(); FK[o + 1] = fk.imag(); }
            if (SC) { SC[o] = sc.real(); SC[o + 1] = sc.imag(); }
        }
    }
    // potentials cut to the body panels (:450-458), NPA x 6 and NPA x nb
    for (int c = 0; c < 6 && PhiJ; c++) std::memcpy(PhiJ + 2 * (size_t)c * NPA, PJ + (size_t)c * N, (size_t)NPA * 16);
    for (int c = 0; c < nb && PhiS; c++) std::memcpy(PhiS + 2 * (size_t)c * NPA, PS + (size_t)c * N, (size_t)NPA * 16);
    double t3 = omp_get_wtime();
    if (timings) { timings[0] = t1 - t0; timings[1] = t2 - t1; timings[2] = t3 - t2; }
    return 0;
}

int orc_num_threads() { return omp_get_max_threads(); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

}  // extern "C"
