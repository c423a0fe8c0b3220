// Internal host-side declarations shared by the library's translation units.
#pragma once

#include <cstdint>
#include <random>
#include <string>
#include <istream>
#include <string>
#include <vector>

namespace bemuse {

// Surface mesh: nodes (x,y,z) and elements of 3 or 4 node indices
// (elems[4e+3] < 0 marks a triangle).
struct Mesh {
    std::vector<double> nodes;
    std::vector<int> elems;
    int n_nodes() const { return (int)(nodes.size() / 3); }
    int n_elems() const { return (int)(elems.size() / 4); }
    int add_node(double x, double y, double z);
    void add_quad(int a, int b, int c, int d);
    void add_tri(int a, int b, int c);
    void append(const Mesh& o, double dx, double dy, double dz);
};

// Mesh files (GDF, MAR, PNL, ASCII STL), as the reference's readers parse them.
bool read_gdf(std::istream& in, Mesh& m, std::string& err);
bool read_mar(std::istream& in, Mesh& m, std::string& err);
bool read_pnl(std::istream& in, Mesh& m, std::string& err);
bool read_stl(std::istream& in, Mesh& m, std::string& err);

Mesh mesh_revolution(const std::vector<double>& r, const std::vector<double>& z, int NA);
Mesh mesh_half_cylinder(double R, double draft, int NR, int NV, int NA, bool cosine);
Mesh mesh_lid_disc(double R, int NRFS, int NA, bool cosine);
Mesh mesh_lid_rect(double L, double W, int NX, int NY, bool cosine);
Mesh mesh_lid_ellipse(double a, double b, int NRFS, int NA, bool cosine);
Mesh mesh_tapered_spar(double RB, double RT, double draft, double H1, double H2, int NR, int NV1, int NV2, int NV3,
                       int NA, bool cosine);
Mesh mesh_spar_leg(double RB, double RT, double draft, double H1, int N1, int N2, int N3, int N4, int NA, bool cosine);
Mesh mesh_barge(double L, double W, double D, int NX, int NY, int NZ, bool cosine);
Mesh mesh_ellipsoid(double a, double b, double c, double zc, int NZ, int NA, bool cosine, bool tri);
Mesh mesh_semi_ellipsoid(double a, double b, double c, int NZ, int NA, bool cosine, bool tri);
Mesh mesh_wigley(double L, double B, double T, int NX, int NZ, bool cosine, bool tri, bool outward);
Mesh mesh_box(double L, double W, double z0, double z1, int NX, int NY, int NZ, bool open_top);
Mesh mesh_semisub3(double rc, double T, double spacing, double pw, double ph, double gap,
                   int NA, int NV, int NR, int NX, int NY, int NZ);
Mesh mesh_cone_spar(double r_top, double r_bot, double draft, double r_plate, double t_plate,
                    int NA, int n_disc, int n_rim, int n_annulus, int n_cone);
Mesh mesh_random_body(uint64_t seed, int index, double* params);

// Per-panel record consumed by the assembly kernels (32 doubles, one 256-byte
// line per panel). Offsets are fixed; the kernels index the array directly.
enum PanelField {
    PF_CX = 0, PF_CY, PF_CZ,            // centroid
    PF_EX = 3,                          // frame row x (3)
    PF_EY = 6,                          // frame row y (3)
    PF_EZ = 9,                          // frame row z = unit normal (3)
    PF_LX = 12,                         // vertex local x (4)
    PF_LY = 16,                         // vertex local y (4)
    PF_M = 20,                          // edge slope dy/dx (4)
    PF_D = 24,                          // edge length (4)
    PF_NV = 28,                         // number of vertices (3 or 4), stored as double
    PF_AREA = 29,
    PF_STRIDE = 32
};

// Quadrature record of a mirrored panel (32 doubles): 4 nodes x (pos 3,
// normal 3, weight 1) = 28, then node count and 2 n_z of the body panel.
enum QuadField {
    QF_POS = 0,       // node q at QF_POS + 7 q + {0,1,2}
    QF_NRM = 3,       // node q at QF_POS + 7 q + {3,4,5}
    QF_W = 6,         // node q at QF_POS + 7 q + 6
    QF_NQ = 28,
    QF_REFLNORM = 29,
    QF_STRIDE = 32
};

struct BodyGeometry {
    int N = 0;
    std::vector<double> body;      // N x PF_STRIDE, body panels (Rankine source)
    std::vector<double> refl;      // N x PF_STRIDE, mirrored panels (Rankine image)
    std::vector<double> quad;      // N x QF_STRIDE, mirrored panel quadrature nodes
    std::vector<double> recv;      // N x 4, BC points C + 1e-5 n (x,y,z,0)
    std::vector<double> gen;       // N x 8: n(3), (x x n)(3), area, 0 at centroids
};

// Builds the per-panel records of a mesh (Geo_Elements.cpp:45-214,
// Boundary_Base.cpp:51-85, Solver_Base.cpp:71-83). Returns false on a
// malformed mesh (index out of range, degenerate element).
bool build_geometry(const Mesh& m, BodyGeometry& g, std::string& err);

}  // namespace bemuse
