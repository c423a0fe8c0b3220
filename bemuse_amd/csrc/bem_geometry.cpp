// Host precompute of the per-panel records the assembly kernels read.
//
// Each record holds what the reference re-derives for every (receiver,
// source) pair (Panel.cpp:93-100): the panel frame from
// Quaternion::FromTwoVectors(UnitZ, n) (Geo_Elements.cpp:66,144), the vertex
// coordinates in that frame, the edge slopes m_k = dy/dx (may be +-inf, IEEE
// semantics are relied on, Panel.cpp:102-105) and the edge lengths, plus the
// quadrature nodes of the mirrored panel (Geo_Elements.cpp:69-119,147-214).
#include "bem_internal.h"

#include <cmath>

namespace bemuse {
namespace {

struct V3 { double x, y, z; };
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 crs(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dt(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 unit(V3 a) {
    double s = dt(a, a);
    if (s > 0) { double n = std::sqrt(s); return {a.x / n, a.y / n, a.z / n}; }
    return a;
}

// Rotation taking +z onto unit vector b, applied to the three unit axes:
// quaternion (w, v) = (s/2, (z x b)/s), s = sqrt(2(1 + b_z)); for b = -z the
// rotation is a half turn about +x. Rotating u: u + w t + v x t, t = 2 v x u.
struct Frame { V3 ex, ey, ez; };
Frame frame_from_normal(V3 n) {
    V3 b = unit(n);
    double c = b.z;
    double w; V3 v;
    if (c < -1.0 + 1e-12) {
        c = c < -1.0 ? -1.0 : c;
        double w2 = (1.0 + c) * 0.5;
        w = std::sqrt(w2);
        v = {std::sqrt(1.0 - w2), 0.0, 0.0};
    } else {
        V3 axis = crs({0.0, 0.0, 1.0}, b);
        double s = std::sqrt((1.0 + c) * 2.0);
        v = mul(axis, 1.0 / s);
        w = s * 0.5;
    }
    auto rot = [&](V3 u) {
        V3 t = crs(v, u);
        t = add(t, t);
        return add(add(u, mul(t, w)), crs(v, t));
    };
    return {rot({1, 0, 0}), rot({0, 1, 0}), rot({0, 0, 1})};
}

struct Pan {
    int nv;
    V3 P[4];
    V3 C;
    double area;
    Frame f;
    int nq;
    V3 qp[4], qn[4];
    double qw[4];
};

void panel_geometry(Pan& p) {
    if (p.nv == 3) {
        V3 P0 = p.P[0], P1 = p.P[1], P2 = p.P[2];
        p.C = mul(add(add(P0, P1), P2), 1.0 / 3.0);
        V3 X3 = crs(sub(P0, mul(add(P1, P2), 0.5)), sub(P2, P1));
        p.area = 0.5 * std::sqrt(dt(X3, X3));
        p.f = frame_from_normal(unit(X3));
        // Dunavant degree-2 rule as tabulated (Quadrule.cpp:26177-26200):
        // barycentric (2/3, 1/6, 1/6) and its rotations, weight 1/3 (times the
        // reference triangle area 1/2), 15 significant digits.
        const double a = 0.666666666666667, b = 0.166666666666667, w = 0.333333333333333;
        const double bc[3] = {a, b, b};
        V3 Rn = crs(sub(P2, P0), sub(P1, P0));
        double jac = std::sqrt(dt(Rn, Rn));
        V3 qn = frame_from_normal(unit(Rn)).ez;
        p.nq = 3;
        for (int k = 0; k < 3; k++) {
            double xi = bc[k], eta = bc[(k + 1) % 3];
            V3 P = add(add(mul(P0, 1. - xi - eta), mul(P1, eta)), mul(P2, xi));
            p.qp[k] = P; p.qn[k] = qn; p.qw[k] = jac * (w * 0.5);
        }
    } else {
        V3 P0 = p.P[0], P1 = p.P[1], P2 = p.P[2], P3 = p.P[3];
        p.C = mul(add(add(add(P0, P1), P2), P3), 0.25);
        V3 X1 = sub(mul(add(P3, P2), 0.5), mul(add(P0, P1), 0.5));
        V3 X2 = sub(mul(add(P2, P1), 0.5), mul(add(P3, P0), 0.5));
        V3 d = crs(sub(P2, P0), sub(P3, P1));
        p.area = 0.5 * std::sqrt(dt(d, d));
        p.f = frame_from_normal(unit(crs(X1, X2)));
        // 2x2 Gauss-Legendre on the bilinear map (N0..N3 as Geo_Elements.cpp:179-182).
        const double g = 1.0 / std::sqrt(3.0);
        const double gx[2] = {-g, g};
        p.nq = 4;
        int k = 0;
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
                double xi = gx[i], eta = gx[j];
                V3 P = {0, 0, 0};
                P = add(P, mul(P0, 0.25 * (1 - xi) * (1 - eta)));
                P = add(P, mul(P1, 0.25 * (1 - xi) * (1 + eta)));
                P = add(P, mul(P2, 0.25 * (1 + xi) * (1 + eta)));
                P = add(P, mul(P3, 0.25 * (1 + xi) * (1 - eta)));
                V3 Rx = add(add(mul(P0, -0.25 * (1 - eta)), mul(P1, -0.25 * (1 + eta))),
                            add(mul(P2, 0.25 * (1 + eta)), mul(P3, 0.25 * (1 - eta))));
                V3 Re = add(add(mul(P0, -0.25 * (1 - xi)), mul(P1, 0.25 * (1 - xi))),
                            add(mul(P2, 0.25 * (1 + xi)), mul(P3, -0.25 * (1 + xi))));
                V3 Rn = crs(Rx, Re);
                double jac = std::sqrt(dt(Rn, Rn));
                p.qp[k] = P;
                p.qn[k] = frame_from_normal(unit(Rn)).ez;
                p.qw[k] = jac;      // rule weights are 1 x 1
                k++;
            }
    }
}

void write_panel(const Pan& p, double* r) {
    for (int i = 0; i < PF_STRIDE; i++) r[i] = 0.0;
    r[PF_CX] = p.C.x; r[PF_CY] = p.C.y; r[PF_CZ] = p.C.z;
    r[PF_EX] = p.f.ex.x; r[PF_EX + 1] = p.f.ex.y; r[PF_EX + 2] = p.f.ex.z;
    r[PF_EY] = p.f.ey.x; r[PF_EY + 1] = p.f.ey.y; r[PF_EY + 2] = p.f.ey.z;
    r[PF_EZ] = p.f.ez.x; r[PF_EZ + 1] = p.f.ez.y; r[PF_EZ + 2] = p.f.ez.z;
    double lx[4], ly[4];
    for (int k = 0; k < p.nv; k++) {
        V3 d = sub(p.P[k], p.C);
        lx[k] = dt(p.f.ex, d);
        ly[k] = dt(p.f.ey, d);
    }
    for (int k = 0; k < 4; k++) {
        int kk = k < p.nv ? k : p.nv - 1;
        r[PF_LX + k] = lx[kk];
        r[PF_LY + k] = ly[kk];
    }
    for (int k = 0; k < p.nv; k++) {
        int k1 = (k + 1) % p.nv;
        double ddx = lx[k1] - lx[k], ddy = ly[k1] - ly[k];
        r[PF_M + k] = ddy / ddx;
        r[PF_D + k] = std::sqrt(ddx * ddx + ddy * ddy);
    }
    r[PF_NV] = p.nv;
    r[PF_AREA] = p.area;
}

}  // namespace

bool build_geometry(const Mesh& m, BodyGeometry& g, std::string& err) {
    int N = m.n_elems();
    int nn = m.n_nodes();
    g.N = N;
    g.body.assign((size_t)N * PF_STRIDE, 0.0);
    g.refl.assign((size_t)N * PF_STRIDE, 0.0);
    g.quad.assign((size_t)N * QF_STRIDE, 0.0);
    g.recv.assign((size_t)N * 4, 0.0);
    g.gen.assign((size_t)N * 8, 0.0);
    for (int e = 0; e < N; e++) {
        const int* el = &m.elems[4 * e];
        int nv = el[3] < 0 ? 3 : 4;
        Pan p, r;
        p.nv = r.nv = nv;
        for (int k = 0; k < nv; k++) {
            if (el[k] < 0 || el[k] >= nn || el[nv - 1 - k] < 0 || el[nv - 1 - k] >= nn) {
                err = "element " + std::to_string(e) + " references a node out of range";
                return false;
            }
            const double* X = &m.nodes[3 * el[k]];
            p.P[k] = {X[0], X[1], X[2]};
            // mirror image about z = 0 with reversed node order (Boundary_Base.cpp:51-85)
            const double* Xr = &m.nodes[3 * el[nv - 1 - k]];
            r.P[k] = {Xr[0], Xr[1], -Xr[2]};
        }
        panel_geometry(p);
        panel_geometry(r);
        if (!(p.area > 0.0)) {
            err = "element " + std::to_string(e) + " has zero area";
            return false;
        }
        write_panel(p, &g.body[(size_t)e * PF_STRIDE]);
        write_panel(r, &g.refl[(size_t)e * PF_STRIDE]);
        double* q = &g.quad[(size_t)e * QF_STRIDE];
        for (int k = 0; k < r.nq; k++) {
            double* o = q + QF_POS + 7 * k;
            o[0] = r.qp[k].x; o[1] = r.qp[k].y; o[2] = r.qp[k].z;
            o[3] = r.qn[k].x; o[4] = r.qn[k].y; o[5] = r.qn[k].z;
            o[6] = r.qw[k];
        }
        q[QF_NQ] = r.nq;
        q[QF_REFLNORM] = p.f.ez.z * 2.0;          // ReflNormMat (Hydrodynamic_Solver.cpp:96-97)
        V3 bc = add(p.C, mul(p.f.ez, 1.0e-5));    // BCF = 1e-5 (Solver_Base.h:55)
        g.recv[4 * e] = bc.x; g.recv[4 * e + 1] = bc.y; g.recv[4 * e + 2] = bc.z;
        V3 pcn = crs(p.C, p.f.ez);
        double* gn = &g.gen[8 * (size_t)e];
        gn[0] = p.f.ez.x; gn[1] = p.f.ez.y; gn[2] = p.f.ez.z;
        gn[3] = pcn.x; gn[4] = pcn.y; gn[5] = pcn.z;
        gn[6] = p.area;
    }
    return true;
}

}  // namespace bemuse
