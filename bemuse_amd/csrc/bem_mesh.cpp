// Parametric panel meshes for the benchmark bodies.
//
// Node/element conventions follow the reference generators so that the same
// mesh can be fed to both the FP64 CPU oracle and the GPU path:
//   * volume of revolution from a (r, z) profile whose first point lies on the
//     axis: one axis node, NA nodes per further profile point, a triangle fan
//     around the axis node and quads between rings
//     (src/Boundary/Volume_of_Revolution.cpp:12-37,92-109,151-160);
//   * Half_Cylinder profile with cosine spacing by default
//     (Volume_of_Revolution.cpp:167-199, Boundary_Base.h:38, Boundary_Base.h:86);
//   * panel normals point into the body (node order makes X1 x X2 of
//     Geo_Elements.cpp:135-137 inward), so that the signed volume of
//     Boundary_Base.cpp:177-198 is positive;
//   * boxes duplicate nodes along face edges as the reference Barge does
//     (src/Boundary/Barge.cpp:34-98).
#include "bem_internal.h"

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <istream>
#include <string>
#include <vector>

namespace bemuse {

static inline double pi_cos_fac(int i, int itot) { return 0.5 * (1.0 - std::cos(i * M_PI / (itot - 1))); }

int Mesh::add_node(double x, double y, double z) {
    nodes.push_back(x); nodes.push_back(y); nodes.push_back(z);
    return (int)(nodes.size() / 3) - 1;
}
void Mesh::add_quad(int a, int b, int c, int d) { elems.push_back(a); elems.push_back(b); elems.push_back(c); elems.push_back(d); }
void Mesh::add_tri(int a, int b, int c) { elems.push_back(a); elems.push_back(b); elems.push_back(c); elems.push_back(-1); }
void Mesh::append(const Mesh& o, double dx, double dy, double dz) {
    int base = n_nodes();
    for (int i = 0; i < o.n_nodes(); i++) add_node(o.nodes[3 * i] + dx, o.nodes[3 * i + 1] + dy, o.nodes[3 * i + 2] + dz);
    for (int e = 0; e < o.n_elems(); e++) {
        const int* el = &o.elems[4 * e];
        elems.push_back(el[0] + base); elems.push_back(el[1] + base); elems.push_back(el[2] + base);
        elems.push_back(el[3] < 0 ? -1 : el[3] + base);
    }
}

// Volume of revolution about the z axis from profile points (r_i, z_i);
// profile[0] must lie on the axis (r = 0).
Mesh mesh_revolution(const std::vector<double>& r, const std::vector<double>& z, int NA) {
    Mesh m;
    int NZ = (int)r.size() - 1;
    for (int i = 0; i <= NZ; i++) {
        int NAR = (i == 0) ? 1 : NA;
        for (int j = 0; j < NAR; j++) {
            double u = j * (2.0 * M_PI) / NAR;
            m.add_node(r[i] * std::cos(u), r[i] * std::sin(u), z[i]);
        }
    }
    auto id = [NA](int A, int Z) -> int {
        if (Z == 0) return 0;
        if (A >= NA) return 1 + (Z - 1) * NA + (A - NA);
        if (A < 0) return 1 + (Z - 1) * NA + (A + NA);
        return 1 + (Z - 1) * NA + A;
    };
    for (int i = 0; i < NA; i++) m.add_tri(id(0, 0), id(i + 1, 1), id(i, 1));
    for (int zz = 1; zz < NZ; zz++)
        for (int i = 0; i < NA; i++) m.add_quad(id(i, zz), id(i + 1, zz), id(i + 1, zz + 1), id(i, zz + 1));
    return m;
}

// Interior free-surface lid of a volume of revolution
// (Volume_of_Revolution::Generate_Aux_Nodes / Generate_Aux_Elements,
// Volume_of_Revolution.cpp:43-65, 111-130): rings at r = R i / NRFS (or the
// cosine spacing) in z = 0, a triangle fan round the centre and quads
// outside it, in the reference's node order.
Mesh mesh_lid_disc(double R, int NRFS, int NA, bool cosine) { return mesh_lid_ellipse(R, R, NRFS, NA, cosine); }

// The elliptic lid of a semi-ellipsoid (Semi_Ellipsoid::Generate_Aux_Nodes /
// _Elements, Ellipsoid.cpp:326-362): the disc's pattern, scaled by (a, b).
Mesh mesh_lid_ellipse(double a, double b, int NRFS, int NA, bool cosine) {
    Mesh m;
    for (int i = 0; i <= NRFS; i++) {
        double f = cosine ? pi_cos_fac(i, NRFS + 1) : 1.0 * i / NRFS;
        int NAR = (i == 0) ? 1 : NA;
        for (int j = 0; j < NAR; j++) {
            double u = j * (2.0 * M_PI) / NAR;
            m.add_node(a * f * std::cos(u), b * f * std::sin(u), 0.0);
        }
    }
    auto id = [NA](int A, int Z) -> int {
        if (Z == 0) return 0;
        if (A >= NA) return 1 + (Z - 1) * NA + (A - NA);
        if (A < 0) return 1 + (Z - 1) * NA + (A + NA);
        return 1 + (Z - 1) * NA + A;
    };
    for (int i = 0; i < NA; i++) m.add_tri(id(0, 0), id(i, 1), id(i + 1, 1));
    for (int z = 1; z < NRFS; z++)
        for (int i = 0; i < NA; i++) m.add_quad(id(i, z + 1), id(i + 1, z + 1), id(i + 1, z), id(i, z));
    return m;
}

// Interior free-surface lid of a barge (Barge::Generate_Aux_Nodes /
// Generate_Aux_Elements, Barge.cpp:103-123, 193-204), an NX x NY grid of
// quads in z = 0. The reference computes y from the x index (y = -W/2 +
// W i / NY, :115), which makes every lid quad degenerate; the grid here uses
// the y index.
Mesh mesh_lid_rect(double L, double W, int NX, int NY, bool cosine) {
    Mesh m;
    for (int i = 0; i <= NX; i++)
        for (int j = 0; j <= NY; j++) {
            double x = cosine ? -0.5 * L + L * pi_cos_fac(i, NX + 1) : -0.5 * L + L * i / NX;
            double y = cosine ? -0.5 * W + W * pi_cos_fac(j, NY + 1) : -0.5 * W + W * j / NY;
            m.add_node(x, y, 0.0);
        }
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
            m.add_quad(i * (NY + 1) + j, (i + 1) * (NY + 1) + j, (i + 1) * (NY + 1) + j + 1, i * (NY + 1) + j + 1);
    return m;
}

// Half_Cylinder::Generate_Nodes profile (Volume_of_Revolution.cpp:185-199).
Mesh mesh_half_cylinder(double R, double draft, int NR, int NV, int NA, bool cosine) {
    double Z = -draft;
    std::vector<double> r, z;
    if (cosine) {
        for (int i = 0; i < NR; i++) { r.push_back(R * pi_cos_fac(i, NR + 1)); z.push_back(Z); }
        for (int i = 0; i < NV + 1; i++) { r.push_back(R); z.push_back(Z * (1.0 - pi_cos_fac(i, NV + 1))); }
    } else {
        for (int i = 0; i < NR; i++) { r.push_back(R * i / NR); z.push_back(Z); }
        for (int i = 0; i < NV + 1; i++) { r.push_back(R); z.push_back(Z * (1.0 - 1.0 * i / NV)); }
    }
    return mesh_revolution(r, z, NA);
}

// Tapered_SparBuoy::Generate_Nodes profile (Volume_of_Revolution.cpp:232-251):
// bottom disc, lower column (H1), taper (H2), upper column to the waterline.
// With cosine spacing the disc has NR - 1 rings, as there.
Mesh mesh_tapered_spar(double RB, double RT, double draft, double H1, double H2, int NR, int NV1, int NV2, int NV3,
                       int NA, bool cosine) {
    double Z = -draft, H3 = Z + H1 + H2;
    std::vector<double> r, z;
    auto pt = [&](double rr, double zz) { r.push_back(rr); z.push_back(zz); };
    if (cosine) {
        for (int i = 0; i < NR - 1; i++) pt(RB * pi_cos_fac(i, NR), Z);
        for (int i = 0; i < NV1; i++) pt(RB, Z + H1 * pi_cos_fac(i, NV1 + 1));
        for (int i = 0; i < NV2; i++) { double f = pi_cos_fac(i, NV2 + 1); pt(RB - (RB - RT) * f, Z + H1 + H2 * f); }
        for (int i = 0; i < NV3 + 1; i++) pt(RT, H3 * (1.0 - pi_cos_fac(i, NV3 + 1)));
    } else {
        for (int i = 0; i < NR; i++) pt(RB * i / NR, Z);
        for (int i = 0; i < NV1; i++) pt(RB, Z + H1 * i / NV1);
        for (int i = 0; i < NV2; i++) pt(RB - (RB - RT) * 1.0 * i / NV2, Z + H1 + H2 * 1.0 * i / NV2);
        for (int i = 0; i < NV3 + 1; i++) pt(RT, H3 * (1.0 - 1.0 * i / NV3));
    }
    return mesh_revolution(r, z, NA);
}

// Spar_Leg::Generate_Nodes profile (Volume_of_Revolution.cpp:278-297): base
// disc (N1), heave-plate rim of height H1 (N2), plate top from RB in to RT
// (N3), column to the waterline (N4).
Mesh mesh_spar_leg(double RB, double RT, double draft, double H1, int N1, int N2, int N3, int N4, int NA, bool cosine) {
    double Z = -draft, H2 = Z + H1;
    std::vector<double> r, z;
    auto pt = [&](double rr, double zz) { r.push_back(rr); z.push_back(zz); };
    if (cosine) {
        for (int i = 0; i < N1; i++) pt(RB * pi_cos_fac(i, N1 + 1), Z);
        for (int i = 0; i < N2; i++) pt(RB, Z + H1 * pi_cos_fac(i, N2 + 1));
        for (int i = 0; i < N3; i++) pt(RB - (RB - RT) * pi_cos_fac(i, N3 + 1), Z + H1);
        for (int i = 0; i < N4 + 1; i++) pt(RT, H2 * (1.0 - pi_cos_fac(i, N4 + 1)));
    } else {
        for (int i = 0; i < N1; i++) pt(RB * i / N1, Z);
        for (int i = 0; i < N2; i++) pt(RB, Z + H1 * i / N2);
        for (int i = 0; i < N3; i++) pt(RB - (RB - RT) * 1.0 * i / N3, Z + H1);
        for (int i = 0; i < N4 + 1; i++) pt(RT, H2 * (1.0 - 1.0 * i / N4));
    }
    return mesh_revolution(r, z, NA);
}

// Barge::Generate_Nodes / Generate_Elements (Barge.cpp:33-191): an open-top
// box, x in [-L/2, L/2], y in [-W/2, W/2], z in [-D, 0], each face with its
// own nodes (x faces, then y faces, then the bottom); the two x faces' panels
// alternate, then the two y faces', then the bottom's; cosine or uniform
// spacing in every direction.
Mesh mesh_barge(double L, double W, double D, int NX, int NY, int NZ, bool cosine) {
    Mesh m;
    auto sx = [&](int i) { return cosine ? -0.5 * L + L * pi_cos_fac(i, NX + 1) : -0.5 * L + L * i / NX; };
    auto sy = [&](int j) { return cosine ? -0.5 * W + W * pi_cos_fac(j, NY + 1) : -0.5 * W + W * j / NY; };
    auto sz = [&](int k) { return cosine ? -D + D * pi_cos_fac(k, NZ + 1) : -D + D * k / NZ; };
    const int nxf = (NY + 1) * (NZ + 1), nyf = (NX + 1) * (NZ + 1);
    const int X1 = 0, X2 = nxf, Y1 = 2 * nxf, Y2 = 2 * nxf + nyf, ZF = 2 * nxf + 2 * nyf;
    for (int side = 0; side < 2; side++)
        for (int j = 0; j <= NY; j++)
            for (int k = 0; k <= NZ; k++) m.add_node(side ? 0.5 * L : -0.5 * L, sy(j), sz(k));
    for (int side = 0; side < 2; side++)
        for (int i = 0; i <= NX; i++)
            for (int k = 0; k <= NZ; k++) m.add_node(sx(i), side ? 0.5 * W : -0.5 * W, sz(k));
    for (int i = 0; i <= NX; i++)
        for (int j = 0; j <= NY; j++) m.add_node(sx(i), sy(j), -D);
    auto xf = [NZ](int j, int k) { return j * (NZ + 1) + k; };
    auto zf = [NY](int i, int j) { return i * (NY + 1) + j; };
    for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
            m.add_quad(X1 + xf(j, k), X1 + xf(j, k + 1), X1 + xf(j + 1, k + 1), X1 + xf(j + 1, k));
            m.add_quad(X2 + xf(j, k), X2 + xf(j + 1, k), X2 + xf(j + 1, k + 1), X2 + xf(j, k + 1));
        }
    for (int i = 0; i < NX; i++)
        for (int k = 0; k < NZ; k++) {
            m.add_quad(Y1 + xf(i, k), Y1 + xf(i + 1, k), Y1 + xf(i + 1, k + 1), Y1 + xf(i, k + 1));
            m.add_quad(Y2 + xf(i, k), Y2 + xf(i, k + 1), Y2 + xf(i + 1, k + 1), Y2 + xf(i + 1, k));
        }
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++) m.add_quad(ZF + zf(i, j), ZF + zf(i, j + 1), ZF + zf(i + 1, j + 1), ZF + zf(i + 1, j));
    return m;
}

// Ellipsoid rings: ring 0 and (closed ellipsoid) ring NZ are single nodes on
// the axis, the others NA nodes; with triangular panels the even rings turn by
// half an azimuthal step. Node index of (azimuth A, ring Z) as
// Ellipsoid::Node_ID / Semi_Ellipsoid::Node_ID (Ellipsoid.cpp:150-161, 444-454).
static int ring_id(int A, int Z, int NA, int NZ, bool closed_top) {
    if (closed_top && Z == NZ) return 1 + NA * (NZ - 1);
    if (Z == 0) return 0;
    if (A >= NA) return 1 + (Z - 1) * NA + (A - NA);
    if (A < 0) return 1 + (Z - 1) * NA + (A + NA);
    return 1 + (Z - 1) * NA + A;
}

// Triangles between rings z and z + 1 (two per azimuthal step), the pattern of
// Ellipsoid.cpp:114-137 / 270-293.
static void ring_triangles(Mesh& m, int z, int NA, int NZ, bool closed_top) {
    auto id = [&](int A, int Z) { return ring_id(A, Z, NA, NZ, closed_top); };
    for (int i = 0; i < NA; i++) {
        if (z % 2 == 0) {
            m.add_tri(id(i, z), id(i + 1, z), id(i + 1, z + 1));
            m.add_tri(id(i + 1, z), id(i + 2, z + 1), id(i + 1, z + 1));
        } else {
            m.add_tri(id(i, z), id(i + 1, z), id(i, z + 1));
            m.add_tri(id(i + 1, z), id(i + 1, z + 1), id(i, z + 1));
        }
    }
}

// Ellipsoid::Generate_Nodes / Generate_Elements (Ellipsoid.cpp:31-148): a
// closed ellipsoid of semi-axes a, b, c centred at (0, 0, zc), bottom pole
// first; zenith angle v = pi (1 - PiCosFac) (cosine) or pi (1 - i/NZ).
Mesh mesh_ellipsoid(double a, double b, double c, double zc, int NZ, int NA, bool cosine, bool tri) {
    Mesh m;
    for (int i = 0; i <= NZ; i++) {
        double v = cosine ? M_PI * (1.0 - pi_cos_fac(i, NZ + 1)) : M_PI - M_PI * i / NZ;
        double sv = std::sin(v), cv = std::cos(v);
        if (i == 0) { sv = 0.0; cv = -1.0; }
        else if (i == NZ) { sv = 0.0; cv = 1.0; }
        int NAR = (i == 0 || i == NZ) ? 1 : NA;
        for (int j = 0; j < NAR; j++) {
            double u = j * (2.0 * M_PI) / NAR;
            if (tri && i % 2 == 0) u += 0.5 * (2.0 * M_PI) / NAR;
            m.add_node(a * std::cos(u) * sv, b * std::sin(u) * sv, c * cv + zc);
        }
    }
    auto id = [&](int A, int Z) { return ring_id(A, Z, NA, NZ, true); };
    for (int i = 0; i < NA; i++) m.add_tri(id(0, 0), id(i + 1, 1), id(i, 1));
    for (int z = 1; z < NZ - 1; z++) {
        if (tri) ring_triangles(m, z, NA, NZ, true);
        else
            for (int i = 0; i < NA; i++) m.add_quad(id(i, z), id(i + 1, z), id(i + 1, z + 1), id(i, z + 1));
    }
    for (int i = 0; i < NA; i++) m.add_tri(id(i + 1, NZ - 1), id(i, NZ), id(i, NZ - 1));
    return m;
}

// Semi_Ellipsoid::Generate_Nodes / Generate_Elements (Ellipsoid.cpp:205-310):
// the lower half (z <= 0) of the ellipsoid, open at the waterline;
// v = pi (1 - PiCosFac / 2) (cosine) or pi (1 - i / (2 NZ)).
Mesh mesh_semi_ellipsoid(double a, double b, double c, int NZ, int NA, bool cosine, bool tri) {
    Mesh m;
    for (int i = 0; i <= NZ; i++) {
        double v = cosine ? M_PI * (1.0 - 0.5 * pi_cos_fac(i, NZ + 1)) : M_PI - 0.5 * M_PI * i / NZ;
        double sv = std::sin(v), cv = std::cos(v);
        if (i == 0) { sv = 0.0; cv = -1.0; }
        else if (i == NZ) { sv = 1.0; cv = 0.0; }
        int NAR = (i == 0) ? 1 : NA;
        for (int j = 0; j < NAR; j++) {
            double u = j * (2.0 * M_PI) / NAR;
            if (tri && i % 2 == 0) u += 0.5 * (2.0 * M_PI) / NAR;
            m.add_node(a * std::cos(u) * sv, b * std::sin(u) * sv, c * cv);
        }
    }
    auto id = [&](int A, int Z) { return ring_id(A, Z, NA, NZ, false); };
    for (int i = 0; i < NA; i++) m.add_tri(id(0, 0), id(i + 1, 1), id(i, 1));
    for (int z = 1; z < NZ; z++) {
        if (tri) ring_triangles(m, z, NA, NZ, false);
        else
            for (int i = 0; i < NA; i++) m.add_quad(id(i, z), id(i + 1, z), id(i + 1, z + 1), id(i, z + 1));
    }
    return m;
}

// Wigley_Hull::Hull_Node (Ship_Hulls.cpp:130-142): x = L xn - L/2,
// y = 2 B xn (1 - xn) (1 - zn^2), z = T zn for xn in [0, 1], zn in [-1, 0].
static void wigley_point(double L, double B, double T, double xn, double zn, double p[3]) {
    p[0] = L * xn - 0.5 * L;
    p[1] = 2.0 * B * xn * (1.0 - xn) * (1.0 - zn * zn);
    p[2] = T * zn;
}

// Wigley_Hull::Generate_Nodes / Generate_*_Elements (Ship_Hulls.cpp:12-90,
// 144-200): port-side nodes, then starboard (y mirrored), then, with
// triangular panels, one centre node per cell on each side; port panels
// first, then starboard. Each cell is a quad, or four triangles round its
// centre node. outward reverses every panel (Panels_Outward, Geo_Elements.h:128-159).
Mesh mesh_wigley(double L, double B, double T, int NX, int NZ, bool cosine, bool tri, bool outward) {
    Mesh m;
    double p[3];
    const int side = (NX + 1) * (NZ + 1);
    for (int s = 0; s < 2; s++)
        for (int i = 0; i <= NX; i++) {
            double xf = cosine ? pi_cos_fac(i, NX + 1) : 1.0 * i / NX;
            for (int k = 0; k <= NZ; k++) {
                double zf = cosine ? pi_cos_fac(k, NZ + 1) - 1.0 : -1.0 + 1.0 * k / NZ;
                wigley_point(L, B, T, xf, zf, p);
                m.add_node(p[0], s ? -p[1] : p[1], p[2]);
            }
        }
    if (tri) {
        auto shift = [](int i, int itot) { return 0.5 * (1.0 - std::cos((i + 0.5) * M_PI / (itot - 1))); };
        for (int s = 0; s < 2; s++)
            for (int i = 0; i < NX; i++) {
                double xf = cosine ? shift(i, NX + 1) : (i + 0.5) / NX;
                for (int k = 0; k < NZ; k++) {
                    double zf = cosine ? shift(k, NZ + 1) - 1.0 : -1.0 + (k + 0.5) / NZ;
                    wigley_point(L, B, T, xf, zf, p);
                    m.add_node(p[0], s ? -p[1] : p[1], p[2]);
                }
            }
    }
    auto Lf = [NZ](int i, int j) { return i * (NZ + 1) + j; };
    auto Rf = [NZ, side](int i, int j) { return side + i * (NZ + 1) + j; };
    const int cL = 2 * side, cR = 2 * side + NX * NZ;
    auto quad = [&m, outward](int a, int b, int c, int d) {
        if (outward) m.add_quad(d, c, b, a); else m.add_quad(a, b, c, d);
    };
    auto trig = [&m, outward](int a, int b, int c) {
        if (outward) m.add_tri(c, b, a); else m.add_tri(a, b, c);
    };
    for (int s = 0; s < 2; s++)
        for (int i = 0; i < NX; i++)
            for (int j = 0; j < NZ; j++) {
                int n1, n2, n3, n4, c;
                if (s == 0) { n1 = Lf(i, j); n2 = Lf(i, j + 1); n3 = Lf(i + 1, j + 1); n4 = Lf(i + 1, j); c = cL + i * NZ + j; }
                else { n1 = Rf(i + 1, j); n2 = Rf(i + 1, j + 1); n3 = Rf(i, j + 1); n4 = Rf(i, j); c = cR + i * NZ + j; }
                if (tri) { trig(n1, n2, c); trig(n2, n3, c); trig(n3, n4, c); trig(n4, n1, c); }
                else quad(n1, n2, n3, n4);
            }
    return m;
}

// Rectangular box centred on the z axis, x in [-L/2, L/2], y in [-W/2, W/2],
// z in [z0, z1]. open_top leaves out the z = z1 face (a surface-piercing barge);
// a closed box is a fully submerged pontoon. Spacing is uniform.
Mesh mesh_box(double L, double W, double z0, double z1, int NX, int NY, int NZ, bool open_top) {
    Mesh m;
    // Face with origin o, edge vectors a (n_a panels) and b (n_b panels); the
    // quad order (0,0),(1,0),(1,1),(0,1) gives normal -(a x b), so a x b must be
    // the outward normal for the panel normal to point into the body.
    auto face = [&m](const double o[3], const double a[3], const double b[3], int na, int nb) {
        int base = m.n_nodes();
        for (int i = 0; i <= na; i++)
            for (int j = 0; j <= nb; j++) {
                double s = (double)i / na, t = (double)j / nb;
                m.add_node(o[0] + s * a[0] + t * b[0], o[1] + s * a[1] + t * b[1], o[2] + s * a[2] + t * b[2]);
            }
        auto id = [base, nb](int i, int j) { return base + i * (nb + 1) + j; };
        for (int i = 0; i < na; i++)
            for (int j = 0; j < nb; j++) m.add_quad(id(i, j), id(i + 1, j), id(i + 1, j + 1), id(i, j + 1));
    };
    double xl = -0.5 * L, xh = 0.5 * L, yl = -0.5 * W, yh = 0.5 * W, H = z1 - z0;
    {   // x = xh, outward +x = (+y) x (+z)
        double o[3] = {xh, yl, z0}, a[3] = {0, W, 0}, b[3] = {0, 0, H}; face(o, a, b, NY, NZ);
    }
    {   // x = xl, outward -x = (+z) x (+y)
        double o[3] = {xl, yl, z0}, a[3] = {0, 0, H}, b[3] = {0, W, 0}; face(o, a, b, NZ, NY);
    }
    {   // y = yh, outward +y = (+z) x (+x)
        double o[3] = {xl, yh, z0}, a[3] = {0, 0, H}, b[3] = {L, 0, 0}; face(o, a, b, NZ, NX);
    }
    {   // y = yl, outward -y = (+x) x (+z)
        double o[3] = {xl, yl, z0}, a[3] = {L, 0, 0}, b[3] = {0, 0, H}; face(o, a, b, NX, NZ);
    }
    {   // z = z0 (bottom), outward -z = (+y) x (+x)
        double o[3] = {xl, yl, z0}, a[3] = {0, W, 0}, b[3] = {L, 0, 0}; face(o, a, b, NY, NX);
    }
    if (!open_top) {   // z = z1 (top), outward +z = (+x) x (+y)
        double o[3] = {xl, yl, z1}, a[3] = {L, 0, 0}, b[3] = {0, W, 0}; face(o, a, b, NX, NY);
    }
    return m;
}

static Mesh rotate_z(const Mesh& src, double ang) {
    Mesh m = src;
    double c = std::cos(ang), s = std::sin(ang);
    for (int i = 0; i < m.n_nodes(); i++) {
        double x = src.nodes[3 * i], y = src.nodes[3 * i + 1];
        m.nodes[3 * i] = c * x - s * y;
        m.nodes[3 * i + 1] = s * x + c * y;
    }
    return m;
}

// Three-column semi-submersible: columns of radius rc and draft T at the
// vertices of an equilateral triangle of side `spacing` (centroid at the
// origin), joined pairwise by closed rectangular pontoons of width pw and
// height ph whose bottom is at -T. Pontoon ends stop `gap` short of the column
// wall so no panel lies inside another part.
Mesh mesh_semisub3(double rc, double T, double spacing, double pw, double ph, double gap,
                   int NA, int NV, int NR, int NX, int NY, int NZ) {
    Mesh out;
    Mesh col = mesh_half_cylinder(rc, T, NR, NV, NA, true);
    double Rv = spacing / std::sqrt(3.0);
    double cx[3], cy[3];
    for (int k = 0; k < 3; k++) {
        double a = M_PI / 2 + k * 2.0 * M_PI / 3.0;
        cx[k] = Rv * std::cos(a); cy[k] = Rv * std::sin(a);
        out.append(col, cx[k], cy[k], 0.0);
    }
    double Lp = spacing - 2.0 * (rc + gap);
    Mesh pont = mesh_box(Lp, pw, -T, -T + ph, NX, NY, NZ, false);
    for (int k = 0; k < 3; k++) {
        int k2 = (k + 1) % 3;
        double mx = 0.5 * (cx[k] + cx[k2]), my = 0.5 * (cy[k] + cy[k2]);
        double ang = std::atan2(cy[k2] - cy[k], cx[k2] - cx[k]);
        out.append(rotate_z(pont, ang), mx, my, 0.0);
    }
    return out;
}

// Truncated-cone spar with a heave plate: a single closed volume of
// revolution. Profile: axis point at the plate bottom, plate bottom disc,
// plate rim, plate top annulus, then the cone from radius r_bot up to r_top
// at the waterline.
Mesh mesh_cone_spar(double r_top, double r_bot, double draft, double r_plate, double t_plate,
                    int NA, int n_disc, int n_rim, int n_annulus, int n_cone) {
    std::vector<double> r, z;
    double zb = -draft - t_plate;
    for (int i = 0; i < n_disc; i++) { r.push_back(r_plate * pi_cos_fac(i, n_disc + 1)); z.push_back(zb); }
    for (int i = 0; i < n_rim; i++) { r.push_back(r_plate); z.push_back(zb + t_plate * i / n_rim); }
    for (int i = 0; i < n_annulus; i++) {
        double f = pi_cos_fac(i, n_annulus + 1);
        r.push_back(r_plate + (r_bot - r_plate) * f); z.push_back(-draft);
    }
    for (int i = 0; i <= n_cone; i++) {
        double f = pi_cos_fac(i, n_cone + 1);
        r.push_back(r_bot + (r_top - r_bot) * f); z.push_back(-draft * (1.0 - f));
    }
    return mesh_revolution(r, z, NA);
}

// Deterministic uniform in [0,1) from a 64-bit Mersenne twister draw.
static inline double u01(std::mt19937_64& g) { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }

// Body `index` of the seeded batch: type Bernoulli(1/2) cylinder/barge, plan
// dimensions U[5,40] m, draft U[5,25] m, target panel count U[500,4000]; the
// discretisation is chosen to be as close as possible to the target.
Mesh mesh_random_body(uint64_t seed, int index, double* params /*6: type, a, b, draft, target, n*/) {
    std::mt19937_64 g(seed);
    for (int i = 0; i < index; i++) for (int k = 0; k < 5; k++) g();
    double t = u01(g), a = 5 + 35 * u01(g), b = 5 + 35 * u01(g), T = 5 + 20 * u01(g);
    int target = 500 + (int)(3500 * u01(g));
    Mesh m;
    if (t < 0.5) {
        // cylinder radius a/2 (diameter in [5,40]); panels = NA * (NV + NR)
        double R = 0.5 * a;
        double best = 1e30; int bNA = 8, bNV = 4, bNR = 2;
        for (int NA = 8; NA <= 256; NA += 4) {
            double hs = 2 * M_PI * R / NA;       // azimuthal panel size
            int NV = std::max(2, (int)std::lround(T / hs));
            int NR = std::max(2, (int)std::lround(R / hs));
            double err = std::fabs((double)NA * (NV + NR) - target);
            if (err < best) { best = err; bNA = NA; bNV = NV; bNR = NR; }
        }
        m = mesh_half_cylinder(R, T, bNR, bNV, bNA, true);
    } else {
        // barge L=a, W=b, draft T, open top; panels = 2NY NZ + 2NX NZ + NX NY
        double best = 1e30; int bx = 4, by = 4, bz = 2;
        for (int k = 2; k <= 200; k++) {
            double hs = std::max(a, b) / k;
            int NX = std::max(2, (int)std::lround(a / hs)), NY = std::max(2, (int)std::lround(b / hs));
            int NZ = std::max(2, (int)std::lround(T / hs));
            double n = 2.0 * NY * NZ + 2.0 * NX * NZ + (double)NX * NY;
            double err = std::fabs(n - target);
            if (err < best) { best = err; bx = NX; by = NY; bz = NZ; }
        }
        m = mesh_box(a, b, -T, 0.0, bx, by, bz, true);
    }
    if (params) {
        params[0] = t < 0.5 ? 0 : 1; params[1] = a; params[2] = b; params[3] = T;
        params[4] = target; params[5] = m.n_elems();
    }
    return m;
}

}  // namespace bemuse

// ---------------------------------------------------------------------------
// Mesh files (src/Boundary/GDF_Geo.cpp, MAR_Geo.cpp, PNL_Geo.cpp, STL_Geo.cpp).
// Lines are split into whitespace-separated fields (the reference splits on
// single blanks; tabs and carriage returns are also separators here).
// ---------------------------------------------------------------------------
namespace bemuse {
namespace {

std::vector<std::string> fields(const std::string& line) {
    std::vector<std::string> t;
    std::string cur;
    for (char c : line) {
        if (c == ' ' || c == '\t' || c == '\r' || c == '\n') {
            if (!cur.empty()) { t.push_back(cur); cur.clear(); }
        } else {
            cur.push_back(c);
        }
    }
    if (!cur.empty()) t.push_back(cur);
    return t;
}

bool to_double(const std::string& s, double& v) {
    char* end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end != s.c_str();
}
bool to_int(const std::string& s, int& v) {
    char* end = nullptr;
    long x = std::strtol(s.c_str(), &end, 10);
    v = (int)x;
    return end != s.c_str();
}

}  // namespace

// GDF (WAMIT): every line of exactly three numbers is a vertex, four vertices
// per panel, a fourth vertex equal to the third makes a triangle; nodes are not
// shared between panels (GDF_Geo.cpp:12-103).
bool read_gdf(std::istream& in, Mesh& m, std::string& err) {
    std::string line;
    double p[4][3];
    int count = 0;
    while (std::getline(in, line)) {
        auto f = fields(line);
        if (f.size() != 3) continue;
        double v[3];
        if (!to_double(f[0], v[0]) || !to_double(f[1], v[1]) || !to_double(f[2], v[2])) continue;
        for (int c = 0; c < 3; c++) p[count][c] = v[c];
        if (++count == 4) {
            int id[4];
            for (int k = 0; k < 3; k++) id[k] = m.add_node(p[k][0], p[k][1], p[k][2]);
            bool tri = p[3][0] == p[2][0] && p[3][1] == p[2][1] && p[3][2] == p[2][2];
            if (tri) {
                m.add_tri(id[0], id[1], id[2]);
            } else {
                id[3] = m.add_node(p[3][0], p[3][1], p[3][2]);
                m.add_quad(id[0], id[1], id[2], id[3]);
            }
            count = 0;
        }
    }
    if (m.n_elems() == 0) { err = "no panels found"; return false; }
    return true;
}

// MAR (NEMOH): two-field lines (symmetry) are skipped; four-field lines are
// nodes "id x y z" until a line with id 0, then connectivity "n1 n2 n3 n4"
// (1-based, n3 == n4 for a triangle); the symmetry flag is not applied, as
// in the reference (MAR_Geo.cpp:12-97).
bool read_mar(std::istream& in, Mesh& m, std::string& err) {
    std::string line;
    bool conn = false;
    std::vector<std::array<int, 4>> el;
    while (std::getline(in, line)) {
        auto f = fields(line);
        if (f.size() != 4) continue;
        int id;
        if (!to_int(f[0], id)) continue;
        if (id == 0 && !conn) { conn = true; continue; }
        if (conn) {
            int n[4];
            for (int k = 0; k < 4; k++)
                if (!to_int(f[k], n[k])) { err = "bad connectivity line: " + line; return false; }
            if (n[0] == 0) continue;   // closing "0 0 0 0"
            el.push_back({n[0] - 1, n[1] - 1, n[2] - 1, f[2] == f[3] ? -1 : n[3] - 1});
        } else {
            double v[3];
            for (int k = 0; k < 3; k++)
                if (!to_double(f[k + 1], v[k])) { err = "bad node line: " + line; return false; }
            m.add_node(v[0], v[1], v[2]);
        }
    }
    for (auto& e : el) {
        for (int k = 0; k < 4; k++)
            if ((k < 3 || e[3] >= 0) && (e[k] < 0 || e[k] >= m.n_nodes())) { err = "node index out of range"; return false; }
        if (e[3] < 0) m.add_tri(e[0], e[1], e[2]); else m.add_quad(e[0], e[1], e[2], e[3]);
    }
    if (m.n_elems() == 0) { err = "no panels found"; return false; }
    return true;
}

// PNL (HAMS): sections marked by "Number of Panels", "Start/End Definition of
// Node Coordinates" and "... Node Relations"; node lines "id x y z", panel
// lines "id nv n1 n2 n3 [n4]" (1-based); symmetry flags are read but not
// applied, as in the reference (PNL_Geo.cpp:16-131).
bool read_pnl(std::istream& in, Mesh& m, std::string& err) {
    std::string line;
    int pos = 0;
    std::vector<std::array<int, 4>> el;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        if (line.find("Number of Panels") != std::string::npos) { pos = 1; continue; }
        if (line.find("Start Definition of Node Coordinates") != std::string::npos) { pos = 2; continue; }
        if (line.find("End Definition of Node Coordinates") != std::string::npos) { pos = 3; continue; }
        if (line.find("Start Definition of Node Relations") != std::string::npos) { pos = 4; continue; }
        if (line.find("End Definition of Node Relations") != std::string::npos) { pos = 5; continue; }
        auto f = fields(line);
        if (pos == 2 && f.size() >= 4) {
            double v[3];
            for (int k = 0; k < 3; k++)
                if (!to_double(f[k + 1], v[k])) { err = "bad node line: " + line; return false; }
            m.add_node(v[0], v[1], v[2]);
        } else if (pos == 4 && f.size() >= 5) {
            int nv, n[4] = {0, 0, 0, 0};
            if (!to_int(f[1], nv) || (nv != 3 && nv != 4) || (int)f.size() < 2 + nv) { err = "bad panel line: " + line; return false; }
            for (int k = 0; k < nv; k++)
                if (!to_int(f[2 + k], n[k])) { err = "bad panel line: " + line; return false; }
            el.push_back({n[0] - 1, n[1] - 1, n[2] - 1, nv == 4 ? n[3] - 1 : -1});
        }
    }
    for (auto& e : el) {
        for (int k = 0; k < 4; k++)
            if ((k < 3 || e[3] >= 0) && (e[k] < 0 || e[k] >= m.n_nodes())) { err = "node index out of range"; return false; }
        if (e[3] < 0) m.add_tri(e[0], e[1], e[2]); else m.add_quad(e[0], e[1], e[2], e[3]);
    }
    if (m.n_elems() == 0) { err = "no panels found"; return false; }
    return true;
}

// ASCII STL: three "vertex x y z" lines per facet. As in the reference the
// coordinates are scaled by 100 and the vertex order is reversed
// (STL_Geo.cpp:75, 88-91); nodes are not shared.
bool read_stl(std::istream& in, Mesh& m, std::string& err) {
    std::string line;
    std::vector<std::array<double, 3>> v;
    while (std::getline(in, line)) {
        auto f = fields(line);
        if (f.empty()) continue;
        if (f[0] == "facet") v.clear();
        if (f[0] == "vertex" && f.size() >= 4) {
            std::array<double, 3> p;
            for (int k = 0; k < 3; k++)
                if (!to_double(f[k + 1], p[k])) { err = "bad vertex line: " + line; return false; }
            v.push_back(p);
            if (v.size() == 3) {
                int id[3];
                for (int k = 0; k < 3; k++) id[k] = m.add_node(100.0 * v[k][0], 100.0 * v[k][1], 100.0 * v[k][2]);
                m.add_tri(id[2], id[1], id[0]);
                v.clear();
            }
        }
    }
    if (m.n_elems() == 0) { err = "no facets found"; return false; }
    return true;
}

}  // namespace bemuse
