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
Mesh mesh_lid_disc(double R, int NRFS, int NA, bool cosine) {
    Mesh m;
    for (int i = 0; i <= NRFS; i++) {
        double r = cosine ? R * pi_cos_fac(i, NRFS + 1) : R * i / NRFS;
        int NAR = (i == 0) ? 1 : NA;
        for (int j = 0; j < NAR; j++) {
            double u = j * (2.0 * M_PI) / NAR;
            m.add_node(r * std::cos(u), r * std::sin(u), 0.0);
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
