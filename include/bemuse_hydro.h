/*
 * bemuse_hydro.h: C++ drop-in for the BEMUse hydrodynamic radiation solver
 * (src/Solver/Hydrodynamic_Solver.h:135-160, src/Solver/Solver_Base.h:119-151)
 * running its per-frequency path on an MI355X through the C-ABI of
 * bemuse_amd.h. Header-only: any C++11 compiler, link -lbemuse_amd.
 *
 * What stays as in the reference:
 *   - Parameter (name + int/Real/bool, Boundary_Base.h:47-55) and the solver
 *     keys "IFR", "WaveAngle" (radians, one per heading), "NKochin",
 *     "Density", "Gravity" (Hydrodynamic_Solver.cpp:12-27);
 *   - the call sequence of BEMUse_Console (BEMUser_Console.cpp:227-260):
 *     new solver, Set_Parameters, Set_OutputFilePath, Setup(Boundary*), then
 *     per frequency Set_Real(F) + Solve(), then delete;
 *   - Set_Real takes the reference's Frequency and Solve uses
 *     Omega = Frequency / 2 pi and Kappa = Omega^2 / 9.81 (:405-409), the
 *     Omega == 0 composition (:415-419) and the zeroed F and FK at Omega == 0
 *     (:481-484);
 *   - Setup ends with the solve at Frequency = FreqInf = -1 that stops after
 *     A and B (:295-297, :469), an ordinary wave solve at Omega = -1/2pi;
 *   - RadSolArray / DiffSolArray receive Phi_J (N x 6) and Phi_S (N x NBeta)
 *     per Solve (:512-513), when Keep_Potentials is set (they are large);
 *   - no exceptions: failures are printed to std::cout and leave NaN results.
 *   - output files (Hydrodynamic_Solver_IO.cpp:13-645): Setup writes the
 *     .bem header, .hst and .mmx and clears the per-frequency files; every
 *     Solve appends its block to .bem and its rows to .1, .2, .2fk, .2sc, in
 *     the reference's layout and number format (BEMUse_IO.h:58-73). Boundary
 *     volume, area and centre of buoyancy follow Boundary_Base.cpp:125-228,
 *     the mass matrix Hydrodynamic_Solver.cpp:300-337.
 *   - mesh files through GDF_Geometry, MAR_Geometry, PNL_Geometry and
 *     STL_Geometry (Read_Input_File + Setup), parsed as the reference does;
 * What is added: the device and arithmetic mode (constructor), accessors for
 * A, B, F, FK, SC (private members in the reference, read from its output
 * files), Solve_Frequencies for a pipelined sweep over many frequencies, and
 * Write_Output (default true) to switch the files off.
 *   - Kochin functions (Hydrodynamic_Solver.cpp:489-494, 598-642) at NKochin
 *     angles after every wave Solve, appended to _KochinRad.bem and
 *     _KochinDiff.bem (Hydrodynamic_Solver_IO.cpp:589-621); the exponential E
 *     is the Newman factor exp(K z + i K (x cos + y sin)) (the reference
 *     builds it from an uninitialised value, :626-627); NKochin = 0 skips them;
 *   - exterior free-surface elevation (:499-505) at the boundary's
 *     FreeSurface_Nodes (filled by the caller; the reference takes them from
 *     its numerical-wave-tank boundaries), kept in FS_Rad_SolArray /
 *     FS_Scat_SolArray (:514-515).
 *   - irregular-frequency removal ("IFR"): the boundary's interior lid
 *     (Half_Cylinder and Triple_Column_SemiSub with
 *     NPanels_FreeSurface_Int_Radial > 0, Barge always) joins the system as
 *     in Create_Panels (:45-55); Barge's lid y-index fix is noted at
 *     bem_mesh_lid_rect. Half_Cylinder also makes the exterior free-surface
 *     rings of FreeSurface_Radius / NPanels_FreeSurface_Radial.
 * Not provided: RAOs and the .vtp wave-height export. The file
 * .4 (RAOs, which in the reference use an uninitialised damping matrix) is
 * created empty. The waterplane moments of the hydrostatic stiffness are
 * integrated exactly over the waterline polygon (the reference sums them over
 * its lid panels, which are not generated here).
 */
#ifndef BEMUSE_HYDRO_H
#define BEMUSE_HYDRO_H

#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "bemuse_amd.h"

namespace BEMUse {

typedef double Real;
typedef std::complex<double> CReal;

// Dense column-major matrices (the reference uses Eigen's; only storage and
// element access are needed here).
template <typename T>
class DenseMatrix {
public:
    DenseMatrix() {}
    DenseMatrix(int r, int c, T v = T()) : r_(r), c_(c), d_((size_t)r * c, v) {}
    static DenseMatrix Zero(int r, int c) { return DenseMatrix(r, c, T()); }
    T& operator()(int i, int j) { return d_[(size_t)j * r_ + i]; }
    const T& operator()(int i, int j) const { return d_[(size_t)j * r_ + i]; }
    int rows() const { return r_; }
    int cols() const { return c_; }
    T* data() { return d_.data(); }
    const T* data() const { return d_.data(); }
private:
    int r_ = 0, c_ = 0;
    std::vector<T> d_;
};
typedef DenseMatrix<Real> Matrix;
typedef DenseMatrix<CReal> CMatrix;

struct Vector3 {
    Real x = 0, y = 0, z = 0;
    Vector3() {}
    Vector3(Real a, Real b, Real c) : x(a), y(b), z(c) {}
};

static const Real TwoPI = 2.0 * 3.14159265358979323846;
static const Real TwoPIinv = 1.0 / (2.0 * 3.14159265358979323846);
static const Real R2D = 180.0 / 3.14159265358979323846;
static const Real Gravity = 9.81;      // BEMUse_Math_Types.h:115 (K and Phi_I use this constant)
static const Real Rho_wat = 1025.0;    // BEMUse_Math_Types.h:113
static const Real P_atm = 101325.0;    // BEMUse_Math_Types.h:112

// Fixed-width number and text fields of the output files: default stream
// formatting (6 significant digits), a leading blank for non-negative
// numbers, blank-padded to the width.
inline std::string Field(Real a, int width = 15) {
    std::ostringstream os;
    os << a;
    std::string t = os.str();
    if (a >= 0) t.insert(0, " ");
    if ((int)t.size() < width) t.append(width - t.size(), ' ');
    return t;
}
inline std::string Field(const std::string& str, int width = 15) {
    std::string t = str;
    if ((int)t.size() < width) t.append(width - t.size(), ' ');
    return t;
}
inline int Make_Directory(const std::string& path) { return mkdir(path.c_str(), 0733); }

// Parameter: a named int, Real or bool (Boundary_Base.h:47-55).
struct Parameter {
    Parameter(const std::string& name, int val) : name_(name), i_(val) {}
    Parameter(const std::string& name, Real val) : name_(name), r_(val) {}
    Parameter(const std::string& name, bool val) : name_(name), b_(val) {}
    bool myNameis(const std::string& a) const { return a == name_; }
    template <typename T> T Get_Param() const;
private:
    std::string name_;
    int i_ = 0;
    Real r_ = 0;
    bool b_ = false;
};
template <> inline int Parameter::Get_Param<int>() const { return i_; }
template <> inline Real Parameter::Get_Param<Real>() const { return r_; }
template <> inline bool Parameter::Get_Param<bool>() const { return b_; }

// ---------------------------------------------------------------------------
// Boundaries: a panel mesh. Set_Parameters then Setup() generates it, as in
// the reference (Boundary_Base.h:100-131).
// ---------------------------------------------------------------------------
class Boundary {
public:
    virtual ~Boundary() {
        if (mesh_) bem_mesh_free(mesh_);
        if (lid_) bem_mesh_free(lid_);
    }
    virtual void Set_Parameters(std::vector<Parameter>& Params) { (void)Params; }
    virtual void Setup() {}
    int N_Nodes() const { int n = 0, e = 0; if (mesh_) bem_mesh_size(mesh_, &n, &e); return n; }
    int N_Elements() const { int n = 0, e = 0; if (mesh_) bem_mesh_size(mesh_, &n, &e); return e; }
    // nodes (n x 3) and elements (e x 4, -1 in the 4th slot for a triangle)
    void Get_Mesh(std::vector<Real>& nodes, std::vector<int>& elems) const {
        nodes.assign(3 * (size_t)N_Nodes(), 0.0);
        elems.assign(4 * (size_t)N_Elements(), 0);
        if (mesh_) bem_mesh_export(mesh_, nodes.data(), elems.data());
    }
    const bem_mesh* Get_Device_Mesh() const { return mesh_; }
    // auxiliary (interior free-surface lid) panels for IFR, if the boundary has them
    const bem_mesh* Get_Device_Lid() const { return lid_; }
    // Array forms of the reference's element accessors (Boundary_Base.h:173-179),
    // same layout as Get_Mesh: the lid (Get_Aux_Nodes/Elements) and the images
    // about z = 0 that Generate_Symmetry_Nodes/Elements and their Aux variants
    // build (Boundary_Base.cpp:51-112: z negated, each element's node order
    // reversed so that its normal flips). The solver forms these images itself.
    void Get_Aux_Mesh(std::vector<Real>& nodes, std::vector<int>& elems) const {
        int n = 0, e = 0;
        if (lid_) bem_mesh_size(lid_, &n, &e);
        nodes.assign(3 * (size_t)n, 0.0);
        elems.assign(4 * (size_t)e, 0);
        if (lid_) bem_mesh_export(lid_, nodes.data(), elems.data());
    }
    void Get_Symm_Mesh(std::vector<Real>& nodes, std::vector<int>& elems) const {
        Get_Mesh(nodes, elems);
        Mirror_Z(nodes, elems);
    }
    void Get_Aux_Symm_Mesh(std::vector<Real>& nodes, std::vector<int>& elems) const {
        Get_Aux_Mesh(nodes, elems);
        Mirror_Z(nodes, elems);
    }

    //--- geometric and kinematic properties (Boundary_Base.h:132-156)
    Real Volume = 0, Surface_Area = 0, Max_Dim = 0, Mass = 0;
    Matrix Mass_Mat = Matrix::Zero(6, 6);
    Matrix Stiffness_Mat = Matrix::Zero(6, 6);
    Vector3 Centre_Gravity, Centre_Buoyancy;
    Matrix Radii_Gyration = Matrix::Zero(3, 3);
    std::string FileInfo;
    // waterplane area and moments int x, y, x^2, x y, y^2 dA (exact over the waterline polygon)
    Real WP_Area = 0, WP_Mx = 0, WP_My = 0, WP_Mxx = 0, WP_Mxy = 0, WP_Myy = 0;
    int N_Aux_Elements() const { return N_Aux_Panels(); }
    // exterior free-surface points for the wave elevation (Get_FreeSurface_Nodes)
    std::vector<Vector3> FreeSurface_Nodes;
    void Get_FreeSurface_Nodes(std::vector<Vector3>& N) const { N = FreeSurface_Nodes; }
    int N_Aux_Panels() const { int n = 0, e = 0; if (lid_) bem_mesh_size(lid_, &n, &e); return e; }

protected:
    static void Mirror_Z(std::vector<Real>& nodes, std::vector<int>& elems) {
        for (size_t i = 2; i < nodes.size(); i += 3) nodes[i] = -nodes[i];
        for (size_t e = 0; e + 3 < elems.size(); e += 4) {
            if (elems[e + 3] < 0) std::swap(elems[e], elems[e + 2]);               // triangle: 2 1 0
            else { std::swap(elems[e], elems[e + 3]); std::swap(elems[e + 1], elems[e + 2]); }   // quad: 3 2 1 0
        }
    }
    // Boundary::Set_Kin_Params (Boundary_Base.h:163): bodies with known mass
    // properties set them here
    virtual void Set_Kin_Params() {}
    // the end of Boundary::Setup (Boundary_Base.cpp:35-46)
    void Finish() {
        Set_Kin_Params();
        Report();
    }
    bool Cosine = true;   // Cosine_Disc (Boundary_Base.h:86)
    void Adopt_Lid(int rc, bem_mesh* m) {
        if (lid_) { bem_mesh_free(lid_); lid_ = nullptr; }
        if (rc != BEM_OK) { std::cout << "BEMUse: lid generation failed: " << bem_last_error() << std::endl; return; }
        lid_ = m;
    }
    bem_mesh* lid_ = nullptr;
    void Adopt(int rc, bem_mesh* m) {
        if (mesh_) { bem_mesh_free(mesh_); mesh_ = nullptr; }
        if (rc != BEM_OK) { std::cout << "BEMUse: mesh generation failed: " << bem_last_error() << std::endl; return; }
        mesh_ = m;
        Calculate_Properties();
    }
    // Calculate_GridParams, _Volume, _SurfaceArea, _COB (Boundary_Base.cpp:125-228):
    // sums over the panels with centroid, normal (into the body) and area.
    void Calculate_Properties() {
        std::vector<Real> nodes;
        std::vector<int> elems;
        Get_Mesh(nodes, elems);
        int nn = (int)nodes.size() / 3, ne = (int)elems.size() / 4;
        Real mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
        for (int i = 0; i < nn; i++)
            for (int c = 0; c < 3; c++) { mn[c] = std::min(mn[c], nodes[3 * i + c]); mx[c] = std::max(mx[c], nodes[3 * i + c]); }
        Max_Dim = 0;
        for (int c = 0; c < 3; c++) if (nn && mx[c] - mn[c] > Max_Dim) Max_Dim = mx[c] - mn[c];
        std::vector<double> body(32 * (size_t)ne);
        if (bem_geometry(mesh_, body.data(), nullptr, nullptr, nullptr, nullptr) != BEM_OK) return;
        Real vx = 0, vy = 0, vz = 0, cb[3] = {0, 0, 0};
        Surface_Area = 0;
        for (int e = 0; e < ne; e++) {
            const double* P = &body[32 * (size_t)e];
            Real cx = P[0], cy = P[1], cz = P[2], nx = P[9], ny = P[10], nz = P[11], A = P[29];
            vx += -cx * nx * A; vy += -cy * ny * A; vz += -cz * nz * A;
            cb[0] -= nx * cx * cx * A; cb[1] -= ny * cy * cy * A; cb[2] -= nz * cz * cz * A;
            Surface_Area += A;
        }
        Volume = (vx + vy + vz) / 3.0;
        Real* C[3] = {&Centre_Buoyancy.x, &Centre_Buoyancy.y, &Centre_Buoyancy.z};
        for (int c = 0; c < 3; c++) {
            *C[c] = cb[c] * 0.5 / Volume;
            if (std::fabs(*C[c]) < Max_Dim * 1e-6) *C[c] = 0;
        }
        // waterplane: the panel edges lying in z = 0, in panel node order
        Real a = 0, m1x = 0, m1y = 0, mxx = 0, mxy = 0, myy = 0, tol = 1e-9 * std::max(Max_Dim, (Real)1);
        for (int e = 0; e < ne; e++) {
            int nv = elems[4 * e + 3] < 0 ? 3 : 4;
            for (int k = 0; k < nv; k++) {
                const Real* p = &nodes[3 * (size_t)elems[4 * e + k]];
                const Real* q = &nodes[3 * (size_t)elems[4 * e + (k + 1) % nv]];
                if (std::fabs(p[2]) > tol || std::fabs(q[2]) > tol) continue;
                Real x1 = p[0], y1 = p[1], x2 = q[0], y2 = q[1], cr = x1 * y2 - x2 * y1;
                a += cr / 2;
                m1x += cr * (x1 + x2) / 6;
                m1y += cr * (y1 + y2) / 6;
                mxx += cr * (x1 * x1 + x1 * x2 + x2 * x2) / 12;
                myy += cr * (y1 * y1 + y1 * y2 + y2 * y2) / 12;
                mxy += cr * (x1 * y2 + 2 * x1 * y1 + 2 * x2 * y2 + x2 * y1) / 24;
            }
        }
        Real sg = a < 0 ? -1 : 1;
        WP_Area = sg * a; WP_Mx = sg * m1x; WP_My = sg * m1y; WP_Mxx = sg * mxx; WP_Mxy = sg * mxy; WP_Myy = sg * myy;
    }
    // the summary line of Boundary_Base::Setup, printed once the lid and the
    // exterior free-surface points are known
    void Report() const {
        std::cout << "Geometry generated. " << N_Elements() << " elements, " << N_Aux_Panels()
                  << " auxiliary (free surface) elements. " << FreeSurface_Nodes.size()
                  << " external (free surface) elements. \n";
        std::cout << "Surface area = " << Surface_Area << " m^2. Volume = " << Volume << " m^3" << std::endl;
    }
    bem_mesh* mesh_ = nullptr;
};

// A mesh given as arrays (e.g. read from a GDF/PNL/STL file by the caller).
class Mesh_Boundary : public Boundary {
public:
    Mesh_Boundary(const std::vector<Real>& nodes, const std::vector<int>& elems) {
        bem_mesh* m = nullptr;
        int rc = bem_mesh_from_arrays((int)(nodes.size() / 3), nodes.data(), (int)(elems.size() / 4), elems.data(), &m);
        Adopt(rc, m);   // rc first: the call must complete before m is read
    }
};

// Imported meshes (GDF_Geo.h, MAR_Geo.h, PNL_Geo.h, STL_Geo.h): Read_Input_File
// then Setup, as in the reference; the file is parsed as the reference's
// readers parse it (see bem_mesh_read).
class File_Geometry : public Boundary {
public:
    explicit File_Geometry(int format) : format_(format) {}
    virtual bool Read_Input_File(std::string& FilePath) {
        bem_mesh* m = nullptr;
        int rc = bem_mesh_read(FilePath.c_str(), format_, &m);
        if (rc != BEM_OK) {
            std::cout << FilePath << " . " << bem_last_error() << std::endl;
            return false;
        }
        if (pending_) bem_mesh_free(pending_);
        pending_ = m;
        FileInfo = "Geometry imported from file: " + FilePath;
        return true;
    }
    void Setup() override {
        if (!pending_) { std::cout << "BEMUse: no geometry file read" << std::endl; return; }
        bem_mesh* m = pending_;
        pending_ = nullptr;
        Adopt(BEM_OK, m);
        Finish();
    }
    ~File_Geometry() override { if (pending_) bem_mesh_free(pending_); }
private:
    int format_;
    bem_mesh* pending_ = nullptr;
};
class GDF_Geometry : public File_Geometry { public: GDF_Geometry() : File_Geometry(BEM_MESH_GDF) {} };
class MAR_Geometry : public File_Geometry { public: MAR_Geometry() : File_Geometry(BEM_MESH_MAR) {} };
class PNL_Geometry : public File_Geometry { public: PNL_Geometry() : File_Geometry(BEM_MESH_PNL) {} };
class STL_Geometry : public File_Geometry { public: STL_Geometry() : File_Geometry(BEM_MESH_STL) {} };

// Bodies of revolution (Volume_of_Revolution.cpp): a profile (r, z) swept
// round the z axis, the interior lid of NPanels_FreeSurface_Int_Radial rings
// (:43-65, 111-130) and exterior free-surface rings r = 1.001 R + RFS R i / NRES
// (:69-89; with NRES = 0 the reference divides by zero there, so no rings are
// made then). Subclasses give the hull.
class Volume_of_Revolution : public Boundary {
public:
    void Setup() override {
        bem_mesh *hull = nullptr, *lid = nullptr;
        std::vector<Vector3> fs;
        int rc = Generate_Pieces(&hull, &lid, fs);
        Adopt(rc, hull);
        if (lid) Adopt_Lid(BEM_OK, lid);
        FreeSurface_Nodes = fs;
        Finish();
    }
    // Hull, lid (null without NRFS) and exterior points in the body's own
    // coordinates; composite bodies place them.
    int Generate_Pieces(bem_mesh** hull, bem_mesh** lid, std::vector<Vector3>& fs) {
        int rc = Generate_Hull(hull);
        if (rc != BEM_OK) return rc;
        if (NRFS > 0) {
            rc = bem_mesh_lid_disc(R, NRFS, NA, Cosine ? 1 : 0, lid);
            if (rc != BEM_OK) std::cout << "BEMUse: lid generation failed: " << bem_last_error() << std::endl;
        }
        fs.clear();
        for (int i = 0; NRES > 0 && i <= NRES; i++) {
            Real r = R * 1.001 + RFS * R * i / NRES;
            for (int j = 0; j < NA; j++) {
                Real u = j * TwoPI / NA;
                fs.push_back(Vector3(r * std::cos(u), r * std::sin(u), 0.0));
            }
        }
        return BEM_OK;
    }
protected:
    virtual int Generate_Hull(bem_mesh** m) = 0;
    // keys every body of revolution reads
    void Common_Parameter(const Parameter& P) {
        if (P.myNameis("FreeSurface_Radius")) RFS = P.Get_Param<Real>();
        if (P.myNameis("NPanels_Azimuthal")) NA = P.Get_Param<int>();
        if (P.myNameis("NPanels_FreeSurface_Radial")) NRES = P.Get_Param<int>();
        if (P.myNameis("NPanels_FreeSurface_Int_Radial")) NRFS = P.Get_Param<int>();
        if (P.myNameis("Cosine_Disc")) Cosine = P.Get_Param<bool>();
    }
    Real R = 10, RFS = 0;   // waterline (lid) radius; exterior ring extent in multiples of R
    int NA = 32, NRES = 0, NRFS = 0;
};

// Half_Cylinder (Volume_of_Revolution.cpp:167-199): open-top vertical cylinder
// with a bottom disc (triangle fan at the centre).
class Half_Cylinder : public Volume_of_Revolution {
public:
    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params) {
            if (P.myNameis("Radius")) R = P.Get_Param<Real>();
            if (P.myNameis("Draft")) D = P.Get_Param<Real>();
            if (P.myNameis("NPanels_Radial")) NR = P.Get_Param<int>();
            if (P.myNameis("NPanels_Vertical")) NV = P.Get_Param<int>();
            Common_Parameter(P);
        }
    }
protected:
    int Generate_Hull(bem_mesh** m) override { return bem_mesh_half_cylinder(R, D, NR, NV, NA, Cosine ? 1 : 0, m); }
    Real D = 20;
    int NR = 8, NV = 24;
};

// Tapered_SparBuoy (Volume_of_Revolution.cpp:203-251): lower column of radius
// RB up to Height_Taper_Begin above the keel, a taper over Height_Taper_End to
// RT, the upper column to the waterline; the lid radius is RT.
class Tapered_SparBuoy : public Volume_of_Revolution {
public:
    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params) {
            if (P.myNameis("Lower_Radius")) RB = P.Get_Param<Real>();
            if (P.myNameis("Upper_Radius")) RT = P.Get_Param<Real>();
            if (P.myNameis("Draft")) D = P.Get_Param<Real>();
            if (P.myNameis("Height_Taper_Begin")) H1 = P.Get_Param<Real>();
            if (P.myNameis("Height_Taper_End")) H2 = P.Get_Param<Real>();
            if (P.myNameis("NPanels_Radial")) NR = P.Get_Param<int>();
            if (P.myNameis("NPanels_Vertical_Section1")) NV1 = P.Get_Param<int>();
            if (P.myNameis("NPanels_Vertical_Section2")) NV2 = P.Get_Param<int>();
            if (P.myNameis("NPanels_Vertical_Section3")) NV3 = P.Get_Param<int>();
            Common_Parameter(P);
        }
        R = RT;
    }
protected:
    int Generate_Hull(bem_mesh** m) override {
        return bem_mesh_tapered_spar(RB, RT, D, H1, H2, NR, NV1, NV2, NV3, NA, Cosine ? 1 : 0, m);
    }
    Real RB = 4.7, RT = 3.25, D = 120, H1 = 108, H2 = 8;
    int NR = 8, NV1 = 24, NV2 = 4, NV3 = 8;
};

// OC3_SparBuoy (FOWT_Platforms.h:33-119): the OC3 Hywind spar, a tapered spar
// of fixed dimensions times Scaling_Factor, with its mass and stiffness.
class OC3_SparBuoy : public Tapered_SparBuoy {
public:
    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params)
            if (P.myNameis("Scaling_Factor")) ScaleFac = P.Get_Param<Real>();
        Tapered_SparBuoy::Set_Parameters(Params);
        Set_Dimensions();
    }
    void Set_Dimensions(std::vector<Real>& Dm) { ScaleFac = Dm[0]; Set_Dimensions(); }
    void Set_External_Dimensions(std::vector<Real>& Dm) { RFS = Dm[0]; R = RT; }
protected:
    void Set_Dimensions() {
        RB = 4.7 * ScaleFac; RT = 3.25 * ScaleFac; D = 120 * ScaleFac; H1 = 108 * ScaleFac; H2 = 8 * ScaleFac;
        R = RT;
    }
    void Set_Kin_Params() override {
        Real SF3 = ScaleFac * ScaleFac * ScaleFac, SF5 = SF3 * ScaleFac * ScaleFac;
        Centre_Gravity = Vector3(0, 0, -89.9155 * ScaleFac);
        Mass = 7.46633e+06 * SF3;
        Mass_Mat = Matrix::Zero(6, 6);
        Stiffness_Mat = Matrix::Zero(6, 6);
        for (int i = 0; i < 3; i++) Mass_Mat(i, i) = 7.46633e+06 * SF3;
        Mass_Mat(3, 3) = 4.22923e+09 * SF5; Mass_Mat(4, 4) = 4.22923e+09 * SF5; Mass_Mat(5, 5) = 1.226e+10 * SF5;
        Stiffness_Mat(2, 2) = 3.32941e+05 * SF3;
        Stiffness_Mat(3, 3) = -4.99918e+09 * SF5; Stiffness_Mat(4, 4) = -4.99918e+09 * SF5;
    }
    Real ScaleFac = 1.0;
};

// Softwind_Upscaled (FOWT_Platforms.h:121-153): a tapered spar whose
// dimensions come from Set_Dimensions (the discretisation from the
// parameters). The reference leaves the lid radius at its earlier value
// there; here it follows the new upper radius.
class Softwind_Upscaled : public Tapered_SparBuoy {
public:
    void Set_Dimensions(std::vector<Real>& Dm) {
        (void)Dm;
        RB = 9.0; RT = 5.6; D = 91.4; H1 = 78; H2 = 8; RFS = RT; R = RT;
    }
protected:
    void Set_Kin_Params() override {
        Mass = 1.94E+07;
        Centre_Gravity = Vector3(0, 0, -71.56);
    }
};

// Spar_Leg (Volume_of_Revolution.cpp:253-297): a column of radius RT on a
// heave plate of radius RB and height Heave_Plate_Height.
class Spar_Leg : public Volume_of_Revolution {
public:
    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params) {
            if (P.myNameis("Leg_Lower_Radius")) RB = P.Get_Param<Real>();
            if (P.myNameis("Leg_Upper_Radius")) RT = P.Get_Param<Real>();
            if (P.myNameis("Draft")) D = P.Get_Param<Real>();
            if (P.myNameis("Heave_Plate_Height")) H1 = P.Get_Param<Real>();
            if (P.myNameis("NPanels_Radial_Base")) N1 = P.Get_Param<int>();
            if (P.myNameis("NPanels_Vertical_HeavePlate")) N2 = P.Get_Param<int>();
            if (P.myNameis("NPanels_Horizontal_HeavePlate")) N3 = P.Get_Param<int>();
            if (P.myNameis("NPanels_Vertical")) N4 = P.Get_Param<int>();
            Common_Parameter(P);
        }
        R = RT;
    }
protected:
    int Generate_Hull(bem_mesh** m) override {
        return bem_mesh_spar_leg(RB, RT, D, H1, N1, N2, N3, N4, NA, Cosine ? 1 : 0, m);
    }
    Real RB = 12, RT = 6, D = 20, H1 = 6;
    int N1 = 8, N2 = 4, N3 = 4, N4 = 16;
};

// Triple_Spar (Triple_Spar.cpp:10-215): three identical legs (half
// cylinders) at radius Spar_Radius, azimuths -90, 30 and 150 degrees
// measured from +y, shifted by Spar_Radius along x with Shift_Over_Leg; an
// optional central column comes first. Hull, lid and exterior points are the
// parts' in that order. Every part receives the same parameters.
class Triple_Spar : public Boundary {
public:
    Triple_Spar() : legs_{new Half_Cylinder, new Half_Cylinder, new Half_Cylinder} {}
    ~Triple_Spar() override {
        delete central_;
        for (auto* l : legs_) delete l;
    }
    Triple_Spar(const Triple_Spar&) = delete;
    Triple_Spar& operator=(const Triple_Spar&) = delete;
    void Set_Parameters(std::vector<Parameter>& Params) override {
        if (central_) central_->Set_Parameters(Params);
        for (auto* l : legs_) l->Set_Parameters(Params);
        for (const Parameter& P : Params) {
            if (P.myNameis("Spar_Radius")) R_Leg = P.Get_Param<Real>();
            if (P.myNameis("Shift_Over_Leg")) ShiftRearwards = P.Get_Param<bool>();
        }
    }
    void Setup() override {
        const Real PI = 3.14159265358979323846;
        Real th[3] = {-0.5 * PI, -0.5 * PI + 2.0 * PI / 3, -0.5 * PI + 4.0 * PI / 3};
        Vector3 C[4];   // central (origin), legs 1..3
        for (int k = 0; k < 3; k++) {
            C[k + 1] = Vector3(R_Leg * std::sin(th[k]), R_Leg * std::cos(th[k]), 0.0);
            if (ShiftRearwards) C[k + 1].x -= R_Leg * std::sin(th[0]);
        }
        bem_mesh *hull = nullptr, *lid = nullptr;
        int rc = bem_mesh_from_arrays(0, nullptr, 0, nullptr, &hull);
        if (rc == BEM_OK) rc = bem_mesh_from_arrays(0, nullptr, 0, nullptr, &lid);
        std::vector<Vector3> fs;
        bool any_lid = false;
        Volume_of_Revolution* parts[4] = {central_, legs_[0], legs_[1], legs_[2]};
        for (int k = 0; k < 4 && rc == BEM_OK; k++) {
            if (!parts[k]) continue;
            bem_mesh *h = nullptr, *l = nullptr;
            std::vector<Vector3> f;
            rc = parts[k]->Generate_Pieces(&h, &l, f);
            if (rc == BEM_OK) rc = bem_mesh_append(hull, h, C[k].x, C[k].y, C[k].z);
            if (rc == BEM_OK && l) { rc = bem_mesh_append(lid, l, C[k].x, C[k].y, C[k].z); any_lid = true; }
            for (const Vector3& v : f) fs.push_back(Vector3(v.x + C[k].x, v.y + C[k].y, v.z + C[k].z));
            if (h) bem_mesh_free(h);
            if (l) bem_mesh_free(l);
        }
        if (rc != BEM_OK) {
            if (hull) bem_mesh_free(hull);
            if (lid) bem_mesh_free(lid);
            Adopt(rc, nullptr);
            return;
        }
        Adopt(BEM_OK, hull);
        if (any_lid) Adopt_Lid(BEM_OK, lid);
        else bem_mesh_free(lid);
        FreeSurface_Nodes = fs;
        Finish();
    }
protected:
    Volume_of_Revolution* central_ = nullptr;
    Volume_of_Revolution* legs_[3];
    Real R_Leg = 30;
    bool ShiftRearwards = false;
};

// SemiSub_OC4 (Triple_Spar.cpp:159-242): the OC4 DeepCwind semi-submersible,
// a central half cylinder (radius 3.25 m, draft 20 m) and three spar legs
// (radius 6 m on heave plates of 12 m, 6 m high, draft 20 m) at 28.87 m, with
// its mass and stiffness. The reference reads Scaling_Factor from its
// parameter list before appending the new parameters, so the scale is always
// 1 there; it is 1 here too.
class SemiSub_OC4 : public Triple_Spar {
public:
    SemiSub_OC4() {
        central_ = new Half_Cylinder;
        for (auto*& l : legs_) { delete l; l = new Spar_Leg; }
    }
    void Set_Parameters(std::vector<Parameter>& Params) override {
        std::vector<Parameter> cp = Params;
        cp.push_back(Parameter("Radius", ScaleFac * Real(3.25)));
        cp.push_back(Parameter("Draft", ScaleFac * Real(20.0)));
        central_->Set_Parameters(cp);
        std::vector<Parameter> lp = Params;
        lp.push_back(Parameter("Leg_Lower_Radius", ScaleFac * Real(12.0)));
        lp.push_back(Parameter("Leg_Upper_Radius", ScaleFac * Real(6.0)));
        lp.push_back(Parameter("Draft", ScaleFac * Real(20.0)));
        lp.push_back(Parameter("Heave_Plate_Height", ScaleFac * Real(6.0)));
        for (auto* l : legs_) l->Set_Parameters(lp);
        R_Leg = 28.87 * ScaleFac;
    }
protected:
    void Set_Kin_Params() override {
        Real SF3 = ScaleFac * ScaleFac * ScaleFac, SF5 = SF3 * ScaleFac * ScaleFac;
        Centre_Gravity = Vector3(0, 0, -13.46 * ScaleFac);
        Mass = 1.34730e+07 * SF3;
        Mass_Mat = Matrix::Zero(6, 6);
        Stiffness_Mat = Matrix::Zero(6, 6);
        for (int i = 0; i < 3; i++) Mass_Mat(i, i) = Mass;
        Mass_Mat(3, 3) = 6.82700e+09 * SF5; Mass_Mat(4, 4) = 6.82700e+09 * SF5; Mass_Mat(5, 5) = 1.22600e+10 * SF5;
        Stiffness_Mat(2, 2) = 3.836e+06 * SF3;
        Stiffness_Mat(3, 3) = -3.776e+08 * SF5; Stiffness_Mat(4, 4) = -3.776e+08 * SF5;
    }
    Real ScaleFac = 1.0;
};

// Barge (Barge.cpp:12-204): open-top box of Tank_Length x Tank_Width x
// Tank_Depth, cosine spacing by default (Cosine_Disc, Boundary_Base.h:86); its
// lid always exists; exterior points: the (NXFS+1) x (NYFS+1) grid over
// FreeSurface_Length x FreeSurface_Width, dilated by 1.001 (points inside the
// waterplane included, as there; none without the two panel counts).
class Barge : public Boundary {
public:
    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params) {
            if (P.myNameis("Tank_Length")) L = P.Get_Param<Real>();
            if (P.myNameis("Tank_Width")) W = P.Get_Param<Real>();
            if (P.myNameis("Tank_Depth")) D = P.Get_Param<Real>();
            if (P.myNameis("FreeSurface_Length")) LFS = P.Get_Param<Real>();
            if (P.myNameis("FreeSurface_Width")) WFS = P.Get_Param<Real>();
            if (P.myNameis("NPanels_Tank_Length")) NX = P.Get_Param<int>();
            if (P.myNameis("NPanels_Tank_Width")) NY = P.Get_Param<int>();
            if (P.myNameis("NPanels_Tank_Depth")) NZ = P.Get_Param<int>();
            if (P.myNameis("NPanels_FreeSurface_Length")) NXFS = P.Get_Param<int>();
            if (P.myNameis("NPanels_FreeSurface_Width")) NYFS = P.Get_Param<int>();
            if (P.myNameis("Cosine_Disc")) Cosine = P.Get_Param<bool>();
        }
    }
    void Setup() override {
        bem_mesh* m = nullptr;
        int rc = bem_mesh_barge(L, W, D, NX, NY, NZ, Cosine ? 1 : 0, &m);
        Adopt(rc, m);
        // interior lid (Barge.cpp:103-123, 193-204; y from its own index)
        bem_mesh* l = nullptr;
        int rl = bem_mesh_lid_rect(L, W, NX, NY, Cosine ? 1 : 0, &l);
        Adopt_Lid(rl, l);
        FreeSurface_Nodes.clear();
        for (int i = 0; NXFS > 0 && NYFS > 0 && i <= NXFS; i++)
            for (int j = 0; j <= NYFS; j++)
                FreeSurface_Nodes.push_back(Vector3(1.001 * (-0.5 * LFS + LFS * i / NXFS),
                                                    1.001 * (-0.5 * WFS + WFS * j / NYFS), 0.0));
        Finish();
    }
private:
    Real L = 20, W = 10, D = 5, LFS = 0, WFS = 0;
    int NX = 20, NY = 10, NZ = 5, NXFS = 0, NYFS = 0;
};

// Ellipsoid (Ellipsoid.cpp:12-161): a submerged ellipsoid of semi-axes a, b, c
// centred Ellipsoid_Depth below the free surface; no lid, no exterior points.
class Ellipsoid : public Boundary {
public:
    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params) {
            if (P.myNameis("Semiaxis_a")) a = P.Get_Param<Real>();
            if (P.myNameis("Semiaxis_b")) b = P.Get_Param<Real>();
            if (P.myNameis("Semiaxis_c")) c = P.Get_Param<Real>();
            if (P.myNameis("Ellipsoid_Depth")) Depth = P.Get_Param<Real>();
            if (P.myNameis("NPanels_Axial")) NZ = P.Get_Param<int>();
            if (P.myNameis("NPanels_Azimuthal")) NA = P.Get_Param<int>();
            if (P.myNameis("Cosine_Disc")) Cosine = P.Get_Param<bool>();
            if (P.myNameis("Triangular_Panels")) TriPanels = P.Get_Param<bool>();
        }
    }
    void Setup() override {
        bem_mesh* m = nullptr;
        int rc = bem_mesh_ellipsoid(a, b, c, Depth, NZ, NA, Cosine ? 1 : 0, TriPanels ? 1 : 0, &m);
        Adopt(rc, m);
        Finish();
    }
private:
    Real a = 1, b = 1, c = 1, Depth = 2;
    int NZ = 16, NA = 32;
    bool TriPanels = false;
};

// Semi_Ellipsoid (Ellipsoid.cpp:167-461): the lower half of an ellipsoid
// piercing the free surface (a hemisphere for a = b = c); elliptic lid of
// NPanels_FreeSurface_Int_Radial rings (none with triangular panels, as
// there); exterior rings at f = 1.001 .. RFS times the waterline ellipse
// (RFS <= 1 becomes 2, as there; no rings without NPanels_FreeSurface_Radial).
class Semi_Ellipsoid : public Boundary {
public:
    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params) {
            if (P.myNameis("Semiaxis_a")) a = P.Get_Param<Real>();
            if (P.myNameis("Semiaxis_b")) b = P.Get_Param<Real>();
            if (P.myNameis("Semiaxis_c")) c = P.Get_Param<Real>();
            if (P.myNameis("FreeSurface_Radius")) RFS = P.Get_Param<Real>();
            if (P.myNameis("NPanels_Axial")) NZ = P.Get_Param<int>();
            if (P.myNameis("NPanels_Azimuthal")) NA = P.Get_Param<int>();
            if (P.myNameis("NPanels_FreeSurface_Radial")) NRES = P.Get_Param<int>();
            if (P.myNameis("NPanels_FreeSurface_Int_Radial")) NRFS = P.Get_Param<int>();
            if (P.myNameis("Cosine_Disc")) Cosine = P.Get_Param<bool>();
            if (P.myNameis("Triangular_Panels")) TriPanels = P.Get_Param<bool>();
        }
        if (RFS <= 1.0) RFS = 2.;
    }
    void Setup() override {
        bem_mesh* m = nullptr;
        int rc = bem_mesh_semi_ellipsoid(a, b, c, NZ, NA, Cosine ? 1 : 0, TriPanels ? 1 : 0, &m);
        Adopt(rc, m);
        if (!TriPanels && NRFS > 0) {
            bem_mesh* l = nullptr;
            int rl = bem_mesh_lid_ellipse(a, b, NRFS, NA, Cosine ? 1 : 0, &l);
            Adopt_Lid(rl, l);
        }
        FreeSurface_Nodes.clear();
        const Real FShift = TriPanels ? 0.0 : 0.001;
        for (int i = 0; NRES > 0 && i <= NRES; i++) {
            Real f = Cosine ? 1.0 + FShift + (RFS - 1.0) * 0.5 * (1.0 - std::cos(i * 3.14159265358979323846 / NRES))
                            : 1.0 + FShift + (RFS - 1.0) * i * 1.0 / NRES;
            for (int j = 0; j < NA; j++) {
                Real u = j * TwoPI / NA;
                if (TriPanels && i % 2 == 0) u += 0.5 * TwoPI / NA;
                FreeSurface_Nodes.push_back(Vector3(a * f * std::cos(u), b * f * std::sin(u), 0.0));
            }
        }
        Finish();
    }
private:
    Real a = 1, b = 1, c = 1, RFS = 2;
    int NZ = 16, NA = 32, NRES = 0, NRFS = 0;
    bool TriPanels = false;
};

// Wigley_Hull (Ship_Hulls.cpp:95-360): y = 2 B x'(1 - x')(1 - z'^2) for
// x' = (x + L/2)/L in [0, 1], z' = z/T in [-1, 0]; exterior points on the port
// side over three zones (aft, beside, fore of the hull; :150-330). The
// reference's lid covers only the port half of the waterplane and collapses
// its bow and stern columns to single points (zero-length panel edges); it is
// not built here, so "IFR" solves a Wigley hull without a lid.
class Wigley_Hull : public Boundary {
public:
    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params) {
            if (P.myNameis("Length")) L = P.Get_Param<Real>();
            if (P.myNameis("Beam")) B = P.Get_Param<Real>();
            if (P.myNameis("Draft")) T = P.Get_Param<Real>();
            if (P.myNameis("NPanels_Length")) NX = P.Get_Param<int>();
            if (P.myNameis("NPanels_Depth")) NZ = P.Get_Param<int>();
            if (P.myNameis("NPanels_FS_Length_Bow")) NXS = P.Get_Param<int>();
            if (P.myNameis("NPanels_FS_Length_Stern")) NXK = P.Get_Param<int>();
            if (P.myNameis("NPanels_FS_Width")) NYFS = P.Get_Param<int>();
            if (P.myNameis("Cosine_Disc")) Cosine = P.Get_Param<bool>();
            if (P.myNameis("Triangular_Panels")) TriPanels = P.Get_Param<bool>();
            if (P.myNameis("Panels_Outward")) Outward = P.Get_Param<bool>();
            if (P.myNameis("BowGridFactor")) XFSMax = P.Get_Param<Real>() * L;
            if (P.myNameis("SternGridFactor")) XFSMin = -P.Get_Param<Real>() * L;
            if (P.myNameis("PortGridFactor")) YFS = P.Get_Param<Real>() * L;
        }
        if (NXS == 0) NXS = NX;
        if (NXK == 0) NXK = 2 * NX;
        if (NYFS == 0) NYFS = NZ;
        if (XFSMax == 0) XFSMax = L;
        if (XFSMin == 0) XFSMin = -2.0 * L;
        if (YFS == 0) YFS = L;
    }
    void Setup() override {
        bem_mesh* m = nullptr;
        int rc = bem_mesh_wigley(L, B, T, NX, NZ, Cosine ? 1 : 0, TriPanels ? 1 : 0, Outward ? 1 : 0, &m);
        Adopt(rc, m);
        Generate_FreeSurface_Nodes();
        Finish();
    }
private:
    static Real pcf(int i, int n) { return 0.5 * (1.0 - std::cos(i * 3.14159265358979323846 / (n - 1))); }
    static Real hpcf(int i, int n) { return -std::cos(0.5 * 3.14159265358979323846 + 0.5 * i * 3.14159265358979323846 / (n - 1)); }
    static Real hpcf2(int i, int n) { return 1.0 - std::cos(0.5 * i * 3.14159265358979323846 / (n - 1)); }
    Vector3 FS_Node(Real xn, Real yn, int zone) const {
        if (zone == 0) return Vector3(XFSMin + (-0.5 * L - XFSMin) * xn, YFS * yn, 0);
        if (zone == 1) {
            Real ymin = 2.0 * B * xn * (1.0 - xn);
            return Vector3(-0.5 * L + L * xn, ymin + (YFS - ymin) * yn, 0);
        }
        return Vector3(0.5 * L + (XFSMax - 0.5 * L) * xn, YFS * yn, 0);
    }
    void Generate_FreeSurface_Nodes() {
        FreeSurface_Nodes.clear();
        std::vector<Real> X[3], Y[3];
        for (int i = 0; i < NXK; i++) X[0].push_back(Cosine ? hpcf(i, NXK + 1) : 1.0 * i / NXK);
        for (int i = 0; i < NX + 1; i++) X[1].push_back(Cosine ? pcf(i, NX + 1) : 1.0 * i / NX);
        for (int i = 1; i < NXS + 1; i++) X[2].push_back(Cosine ? hpcf2(i, NXS + 1) : 1.0 * i / NXS);
        for (int j = 0; j < NYFS + 1; j++) {
            Real fy = Cosine ? hpcf2(j, NYFS + 1) : 1.0 * j / NYFS;
            Y[0].push_back(fy); Y[1].push_back(j == 0 ? 0.001 : fy); Y[2].push_back(fy);
        }
        for (int z = 0; z < 3; z++)
            for (Real x : X[z])
                for (Real y : Y[z]) FreeSurface_Nodes.push_back(FS_Node(x, y, z));
        if (!TriPanels) return;
        // cell centres of the same zones (Ship_Hulls.cpp:307-355)
        const Real PI = 3.14159265358979323846;
        for (int z = 0; z < 3; z++) { X[z].clear(); Y[z].clear(); }
        for (int i = 0; i < NXK; i++) X[0].push_back(Cosine ? -std::cos(0.5 * PI + 0.5 * (i + 0.5) * PI / NXK) : (i + 0.5) / NXK);
        for (int i = 0; i < NX; i++) X[1].push_back(Cosine ? 0.5 * (1.0 - std::cos((i + 0.5) * PI / NX)) : (i + 0.5) / NX);
        for (int i = 0; i < NXS; i++) X[2].push_back(Cosine ? 1.0 - std::cos(0.5 * (i + 0.5) * PI / NXS) : (i + 0.5) / NXS);
        for (int j = 0; j < NYFS; j++) {
            Real fy = Cosine ? 1.0 - std::cos(0.5 * (j + 0.5) * PI / NYFS) : (j + 0.5) / NYFS;
            for (int z = 0; z < 3; z++) Y[z].push_back(fy);
        }
        for (int z = 0; z < 3; z++)
            for (Real x : X[z])
                for (Real y : Y[z]) FreeSurface_Nodes.push_back(FS_Node(x, y, z));
    }
    Real L = 10, B = 1, T = 1, XFSMax = 0, XFSMin = 0, YFS = 0;
    int NX = 20, NZ = 8, NXK = 0, NXS = 0, NYFS = 0;
    bool TriPanels = false, Outward = false;
};

// Three columns joined by closed pontoons (the 6k-panel benchmark body).
class Triple_Column_SemiSub : public Boundary {
public:
    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params) {
            if (P.myNameis("Column_Radius")) Rc = P.Get_Param<Real>();
            if (P.myNameis("Draft")) D = P.Get_Param<Real>();
            if (P.myNameis("Column_Spacing")) S = P.Get_Param<Real>();
            if (P.myNameis("Pontoon_Width")) PW = P.Get_Param<Real>();
            if (P.myNameis("Pontoon_Height")) PH = P.Get_Param<Real>();
            if (P.myNameis("NPanels_FreeSurface_Int_Radial")) NRFS = P.Get_Param<int>();
        }
    }
    void Setup() override {
        bem_mesh* m = nullptr;
        int rc = bem_mesh_semisub3(Rc, D, S, PW, PH, 0.5, 44, 24, 8, 31, 5, 4, &m);
        Adopt(rc, m);
        if (NRFS > 0) {   // one disc lid per column, at the columns' centres
            bem_mesh *disc = nullptr, *all = nullptr;
            int r1 = bem_mesh_lid_disc(Rc, NRFS, 44, 1, &disc);
            if (r1 == BEM_OK) r1 = bem_mesh_from_arrays(0, nullptr, 0, nullptr, &all);
            const Real Rv = S / std::sqrt(3.0);
            for (int k = 0; r1 == BEM_OK && k < 3; k++) {
                Real a = 3.14159265358979323846 / 2 + k * 2.0 * 3.14159265358979323846 / 3.0;
                r1 = bem_mesh_append(all, disc, Rv * std::cos(a), Rv * std::sin(a), 0.0);
            }
            if (disc) bem_mesh_free(disc);
            if (r1 != BEM_OK && all) { bem_mesh_free(all); all = nullptr; }
            Adopt_Lid(r1, all);
        }
        Finish();
    }
private:
    Real Rc = 6, D = 20, S = 50, PW = 6, PH = 4;
    int NRFS = 0;
};

// ---------------------------------------------------------------------------
// Solvers
// ---------------------------------------------------------------------------
class Solver {
public:
    virtual ~Solver() {}
    virtual void Set_Parameters(std::vector<Parameter>& Params) { (void)Params; }
    virtual void Set_Real(Real D) { (void)D; }
    virtual void Set_OutputFilePath(std::string& D) { OutputPath = OutputDirectory + "/" + D; }
    virtual void Setup(Boundary* B) { (void)B; }
    virtual void Solve() {}
    virtual void Post_Processing(Boundary* B) { (void)B; }
    virtual void Generate_Output_File(Boundary* B) { (void)B; }
    virtual void Update_Output_File() {}
    // Solver_Base.h:145; the hydrodynamic solver returns its VisMat, which the
    // reference never assigns (Hydrodynamic_Solver.cpp:434 is commented out),
    // so the result is an empty matrix in both
    virtual CMatrix Get_VisMatrix() { return VisMat; }
    std::vector<CMatrix> RadSolArray;     // radiation potential per Solve (N x 6)
    std::vector<CMatrix> DiffSolArray;    // scattering potential per Solve (N x NBeta)
    std::vector<CMatrix> FS_Rad_SolArray;
    std::vector<CMatrix> FS_Scat_SolArray;
protected:
    std::string OutputDirectory = "Output";
    std::string OutputPath = "Output/BEMUse_Outputs";
    CMatrix VisMat;
};

class Hydrodynamic_Radiation_Solver : public Solver {
public:
    // device: HIP device index; mode: BEM_MODE_FP64 (default, the reference's
    // DoublePrec arithmetic) or BEM_MODE_MIXED (FP32 LU + FP64 refinement).
    explicit Hydrodynamic_Radiation_Solver(int device = 0, int mode = BEM_MODE_FP64) : device_(device), mode_(mode) {}
    ~Hydrodynamic_Radiation_Solver() override {
        if (body_) bem_body_destroy(body_);
        if (ctx_) bem_destroy(ctx_);
    }
    Hydrodynamic_Radiation_Solver(const Hydrodynamic_Radiation_Solver&) = delete;
    Hydrodynamic_Radiation_Solver& operator=(const Hydrodynamic_Radiation_Solver&) = delete;

    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params) {
            if (P.myNameis("IFR")) IFR = P.Get_Param<bool>();
            if (P.myNameis("WaveAngle")) BetaArray.push_back(P.Get_Param<Real>());
            if (P.myNameis("NKochin")) NKoch = P.Get_Param<int>();
            if (P.myNameis("Density")) Rho = P.Get_Param<Real>();
            if (P.myNameis("Gravity")) Grav = P.Get_Param<Real>();
        }
        NBeta = (int)BetaArray.size();
    }
    void Set_Real(Real D) override { Frequency = D; }

    void Setup(Boundary* B) override {
        if (!B || !B->Get_Device_Mesh()) { std::cout << "BEMUse: Setup without a generated boundary" << std::endl; return; }
        if (!ctx_ && !check(bem_create(device_, &ctx_))) return;
        if (body_) { bem_body_destroy(body_); body_ = nullptr; }
        // Create_Panels (Hydrodynamic_Solver.cpp:30-81): with IFR the lid panels join the system
        const bem_mesh* lid = IFR ? B->Get_Device_Lid() : nullptr;
        if (IFR && !lid) std::cout << "BEMUse: IFR requested but the boundary has no lid panels; solving without" << std::endl;
        if (!check(bem_body_create_lid(ctx_, B->Get_Device_Mesh(), lid, &body_))) return;
        NPTOT = bem_body_n(body_);
        B->Get_FreeSurface_Nodes(Wave_Nodes);
        if (!check(bem_body_assemble_rankine(ctx_, body_))) return;
        A_ij = Matrix::Zero(6, 6);
        B_ij = Matrix::Zero(6, 6);
        F_i = CMatrix::Zero(6, NBeta);
        FK_i = CMatrix::Zero(6, NBeta);
        SC_i = CMatrix::Zero(6, NBeta);
        Calculate_Mass_Matrix(B);
        Calculate_Hydrostatic_Stiffness_Matrix(B);
        if (Write_Output) Generate_Output_File(B);
        // the "infinite frequency" radiation solve (Hydrodynamic_Solver.cpp:295-297)
        Frequency = FreqInf;
        Solve();
        if (Write_Output) Update_Output_File();
    }

    void Solve() override {
        Omega = Frequency * TwoPIinv;
        if (Omega == 0 && IFR) Omega = 1.0e-3;
        Period = 1.0 / Frequency;
        Kappa = Omega * Omega / Gravity;
        ExtRadMat = CMatrix();   // this frequency's exterior elevation, if any
        if (!body_) { std::cout << "BEMUse: Solve before Setup" << std::endl; set_nan(); return; }
        int compose = (Omega == 0.0) ? BEM_COMPOSE_ZERO : (Omega == FreqInf ? BEM_COMPOSE_INF : BEM_COMPOSE_WAVE);
        int rad_only = (Frequency == FreqInf) ? 1 : 0;
        std::vector<double> A(36), Bm(36), F(12 * (size_t)NBeta), FK(12 * (size_t)NBeta), SC(12 * (size_t)NBeta);
        std::vector<double> phiJ, phiS;
        const bool post = !rad_only && Omega != 0.0 && (NKoch > 0 || !Wave_Nodes.empty());
        const bool want_phi = Keep_Potentials || post;
        if (want_phi) { phiJ.resize(2 * (size_t)NPTOT * 6); phiS.resize(2 * (size_t)NPTOT * NBeta); }
        if (!check(bem_solve_frequency(ctx_, body_, mode_, Omega, Kappa, compose, rad_only, NBeta,
                                       NBeta ? BetaArray.data() : nullptr, A.data(), Bm.data(),
                                       NBeta ? F.data() : nullptr, NBeta ? FK.data() : nullptr,
                                       NBeta ? SC.data() : nullptr, want_phi ? phiJ.data() : nullptr,
                                       want_phi && NBeta ? phiS.data() : nullptr, &Info))) {
            set_nan();
            return;
        }
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) { A_ij(i, j) = A[6 * i + j]; B_ij(i, j) = Bm[6 * i + j]; }
        if (Keep_Potentials) {
            RadSolArray.push_back(unpack(phiJ, NPTOT, 6));
            DiffSolArray.push_back(unpack(phiS, NPTOT, NBeta));
        }
        if (rad_only) return;
        for (int i = 0; i < 6; i++)
            for (int b = 0; b < NBeta; b++) {
                size_t k = 2 * ((size_t)i * NBeta + b);
                F_i(i, b) = CReal(F[k], F[k + 1]);
                FK_i(i, b) = CReal(FK[k], FK[k + 1]);
                SC_i(i, b) = CReal(SC[k], SC[k + 1]);
            }
        KochinRad = CMatrix::Zero(std::max(NKoch, 0), 6);
        KochinDiff = CMatrix::Zero(std::max(NKoch, 0), NBeta);
        if (post && NKoch > 0) {   // Hydrodynamic_Solver.cpp:492-494
            std::vector<double> kr(2 * (size_t)NKoch * 6), kd(2 * (size_t)NKoch * NBeta);
            if (check(bem_kochin(ctx_, body_, Omega, Kappa, NKoch, NBeta, NBeta ? BetaArray.data() : nullptr,
                                 phiJ.data(), NBeta ? phiS.data() : nullptr, kr.data(), NBeta ? kd.data() : nullptr))) {
                KochinRad = unpack_rows(kr, NKoch, 6);
                KochinDiff = unpack_rows(kd, NKoch, NBeta);
            }
        }
        if (post && !Wave_Nodes.empty()) {   // :499-505
            int nw = (int)Wave_Nodes.size();
            std::vector<double> pts(3 * (size_t)nw), er(2 * (size_t)nw * 6), ed(2 * (size_t)nw * NBeta);
            for (int i = 0; i < nw; i++) { pts[3 * i] = Wave_Nodes[i].x; pts[3 * i + 1] = Wave_Nodes[i].y; pts[3 * i + 2] = Wave_Nodes[i].z; }
            if (check(bem_fs_elevation(ctx_, body_, Omega, Kappa, nw, pts.data(), NBeta,
                                       NBeta ? BetaArray.data() : nullptr, phiJ.data(),
                                       NBeta ? phiS.data() : nullptr, er.data(), NBeta ? ed.data() : nullptr))) {
                ExtRadMat = unpack_rows(er, nw, 6);
                FS_Rad_SolArray.push_back(ExtRadMat);
                FS_Scat_SolArray.push_back(unpack_rows(ed, nw, NBeta));
            }
        }
        if (Write_Output) Update_Output_File();   // Hydrodynamic_Solver.cpp:508
    }

    //--- output files (Hydrodynamic_Solver_IO.cpp:13-31)
    void Generate_Output_File(Boundary* B) override {
        if (!B) return;
        Make_Directory(OutputDirectory);
        Make_Directory(OutputPath);
        Generate_Output_File_BEMUse(B);
        Generate_Output_File_WAMIT(B);
    }
    void Update_Output_File() override {
        Update_Output_File_BEMUse();
        Update_Output_File_WAMIT();
    }

    // Many frequencies at once (the console's loop over Set_Real + Solve),
    // pipelined on the device; results per frequency in the same units.
    bool Solve_Frequencies(const std::vector<Real>& Frequencies, std::vector<Matrix>& A, std::vector<Matrix>& Bd,
                           std::vector<CMatrix>& F, int lanes = 4) {
        if (!body_) { std::cout << "BEMUse: Solve before Setup" << std::endl; return false; }
        int nf = (int)Frequencies.size();
        std::vector<double> om(nf), a(36 * (size_t)nf), bd(36 * (size_t)nf), f(12 * (size_t)NBeta * nf);
        for (int k = 0; k < nf; k++) om[k] = Frequencies[k] * TwoPIinv;
        if (!check(bem_solve_sweep(ctx_, body_, mode_, nf, om.data(), nullptr, BEM_COMPOSE_WAVE, NBeta,
                                   NBeta ? BetaArray.data() : nullptr, lanes, a.data(), bd.data(),
                                   NBeta ? f.data() : nullptr, nullptr, nullptr, nullptr)))
            return false;
        A.assign(nf, Matrix::Zero(6, 6)); Bd.assign(nf, Matrix::Zero(6, 6)); F.assign(nf, CMatrix::Zero(6, NBeta));
        for (int k = 0; k < nf; k++)
            for (int i = 0; i < 6; i++) {
                for (int j = 0; j < 6; j++) { A[k](i, j) = a[36 * k + 6 * i + j]; Bd[k](i, j) = bd[36 * k + 6 * i + j]; }
                for (int b = 0; b < NBeta; b++) {
                    size_t o = 12 * (size_t)NBeta * k + 2 * ((size_t)i * NBeta + b);
                    F[k](i, b) = CReal(f[o], f[o + 1]);
                }
            }
        return true;
    }

    // Results of the last Solve (reference scaling: no rho or omega factors).
    const Matrix& Added_Mass() const { return A_ij; }
    const Matrix& Damping() const { return B_ij; }
    const CMatrix& Excitation_Force() const { return F_i; }
    const CMatrix& Froude_Krylov_Force() const { return FK_i; }
    const CMatrix& Scattering_Force() const { return SC_i; }
    const bem_solve_info& Solve_Info() const { return Info; }
    const CMatrix& Kochin_Radiation() const { return KochinRad; }    // NKochin x 6
    const CMatrix& Kochin_Diffraction() const { return KochinDiff; } // NKochin x NBeta
    int N_Panels() const { return NPTOT; }

    bool Keep_Potentials = false;
    bool Write_Output = true;
    const Matrix& Mass_Matrix() const { return M_ij; }
    const Matrix& Hydrostatic_Stiffness() const { return C_ij; }

private:
    // Hydrodynamic_Solver.cpp:300-337
    void Calculate_Mass_Matrix(Boundary* B) {
        bool zero = true;
        for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) zero = zero && B->Mass_Mat(i, j) == 0;
        if (!zero) { M_ij = B->Mass_Mat; return; }
        if (B->Mass == 0) B->Mass = B->Volume * Rho_wat;
        M_ij = Matrix::Zero(6, 6);
        Real m = B->Mass, xg = B->Centre_Gravity.x, yg = B->Centre_Gravity.y, zg = B->Centre_Gravity.z;
        M_ij(0, 0) = m; M_ij(1, 1) = m; M_ij(2, 2) = m;
        M_ij(0, 4) = m * zg; M_ij(1, 3) = -m * zg; M_ij(0, 5) = -m * yg;
        M_ij(2, 3) = m * yg; M_ij(1, 5) = m * xg; M_ij(2, 4) = -m * xg;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) M_ij(3 + i, j) = -M_ij(i, 3 + j);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) M_ij(3 + i, 3 + j) = m * B->Radii_Gyration(i, j) * std::fabs(B->Radii_Gyration(i, j));
    }
    // Hydrodynamic_Solver.cpp:339-398 (WAMIT manual formulas), including the
    // reference's final symmetry assignments C(2,3)=C(3,2), C(2,4)=C(4,2),
    // C(3,4)=C(4,3); waterplane moments from the boundary.
    void Calculate_Hydrostatic_Stiffness_Matrix(Boundary* B) {
        bool zero = true;
        for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) zero = zero && B->Stiffness_Mat(i, j) == 0;
        if (!zero) { C_ij = B->Stiffness_Mat; return; }
        C_ij = Matrix::Zero(6, 6);
        Real g = Grav, rho = Rho_wat, M = B->Mass, V = B->Volume;
        Real xb = B->Centre_Buoyancy.x, yb = B->Centre_Buoyancy.y, zb = B->Centre_Buoyancy.z;
        Real xg = B->Centre_Gravity.x, yg = B->Centre_Gravity.y, zg = B->Centre_Gravity.z;
        C_ij(2, 2) = Rho * g * B->WP_Area;
        C_ij(2, 3) = Rho * g * B->WP_My;
        C_ij(2, 4) = -Rho * g * B->WP_Mx;
        C_ij(3, 3) = Rho * g * B->WP_Myy + rho * g * V * zb - M * g * zg;
        C_ij(3, 4) = -Rho * g * B->WP_Mxy;
        C_ij(3, 5) = -Rho * g * V * xb + M * g * xg;
        C_ij(4, 4) = Rho * g * B->WP_Mxx + rho * g * V * zb - M * g * zg;
        C_ij(4, 5) = -Rho * g * V * yb + M * g * yg;
        C_ij(2, 3) = C_ij(3, 2);
        C_ij(2, 4) = C_ij(4, 2);
        C_ij(3, 4) = C_ij(4, 3);
    }

    std::string beta_row(const char* label) const {
        std::string t = Field(std::string(label), 30);
        for (Real b : BetaArray) t += Field(b * R2D, 30);
        return t;
    }

    // Hydrodynamic_Solver_IO.cpp:107-225
    void Generate_Output_File_BEMUse(Boundary* B) {
        std::ofstream f(OutputPath + ".bem", std::ofstream::out | std::ofstream::trunc);
        if (!f.is_open()) return;
        f << "---------------------------------------------------------------\n";
        f << "------------------------ BEMUse Output ------------------------\n";
        f << "---------------------------------------------------------------\n";
        f << "\n";
        f << "------------ Environmental Parameters ------------\n";
        f << "Gravity:                 " << Gravity << " m s^{-2}.\n";
        f << "Density of Water:        " << Rho_wat << " kg m^{-3}.\n";
        f << "Atmospheric Pressure:    " << P_atm << " Pa.\n";
        f << "Water Depth:             Infinite.\n";
        f << "\n";
        f << "------------ Wave Parameters ------------\n";
        std::string wa = "Incoming wave angles: (deg)        ";
        for (Real b : BetaArray) wa += Field(b * R2D);
        f << wa << std::endl;
        f << "\n";
        f << "------------ Grid Summary ------------\n";
        f << B->FileInfo << std::endl;
        f << "Input Mesh has:          " << B->N_Nodes() << " nodes.\n";
        f << "Input Mesh has:          " << B->N_Elements() << " panels.\n";
        f << "Input Mesh has:          " << B->N_Aux_Elements() << " auxiliary panels.\n";
        f << "\n";
        f << "------------ Geometry Summary ------------\n";
        f << "Nb: This is calculated by BEMUse with the approximation of flat surface panels." << std::endl;
        f << "With increasing mesh refinement the values are expected to approach those of the continuous geometry.\n";
        f << "Volume:                  " << B->Volume << " m^3.\n";
        f << "Surface area:            " << B->Surface_Area << " m^2.\n";
        f << "Mass:                    " << B->Mass << " kg. (Assuming not pre-specified),\n";
        f << "Centre of Mass:          [" << B->Centre_Gravity.x << ", " << B->Centre_Gravity.y << ", "
          << B->Centre_Gravity.z << "] m.\n";
        f << "Centre of Buoyancy:      [" << B->Centre_Buoyancy.x << ", " << B->Centre_Buoyancy.y << ", "
          << B->Centre_Buoyancy.z << "] m.\n";
        f << "\n";
        f << "------------ Inertial Summary ------------\n";
        f << "\n";
        f << "Mass Matrix: \n";
        for (int i = 0; i < 6; i++) {
            f << "|";
            for (int j = 0; j < 6; j++) f << Field(M_ij(i, j));
            f << " |\n";
        }
        f << "\n";
        f << "Hydrostatic stiffness Matrix: \n";
        for (int i = 0; i < 6; i++) {
            f << "|";
            for (int j = 0; j < 6; j++) f << Field(C_ij(i, j));
            f << " |\n";
        }
        f << "\n";
        f << "------------ Output of Hydrodynamic Analysis: ------------\n";
        f << "\n";
        f << "\n";
    }

    // Hydrodynamic_Solver_IO.cpp:227-306
    void Update_Output_File_BEMUse() {
        std::ofstream f(OutputPath + ".bem", std::ios_base::app);
        if (!f.is_open()) return;
        f << "----------\n";
        f << "Frequency:       " << Omega << " s^{-1}." << std::endl;
        f << "Period:          " << Period << " s." << std::endl;
        f << "Wavenumber:      " << Kappa << " m^{-1}." << std::endl;
        f << "\n";
        const char* cols[] = {" A_x", " B_x", " A_y", " B_y", " A_z", " B_z", " A_pitch", " B_pitch", " A_roll",
                              " B_roll", " A_yaw", " B_yaw"};
        f << Field(std::string("Motion"));
        for (const char* c : cols) f << Field(std::string(c));
        f << std::endl;
        const char* rows[] = {"Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"};
        for (int i = 0; i < 6; i++) {
            f << Field(std::string(rows[i]));
            for (int j = 0; j < 6; j++) f << Field(A_ij(i, j)) << Field(B_ij(i, j));
            f << std::endl;
        }
        f << "\n";
        f << "\n";
        // force tables: the reference labels indices 3 and 4 "Pitch" and "Roll" here
        const char* frows[] = {"Surge: ", "Sway: ", "Heave: ", "Pitch: ", "Roll: ", "Yaw: "};
        auto table = [&](const CMatrix& X, const char* mag, const char* arg) {
            std::string hd = Field(std::string("Motion: "), 30);
            for (int b = 0; b < NBeta; b++) hd += Field(std::string(mag), 15) + Field(std::string(arg), 15);
            f << beta_row("Wave Heading: (deg) ") << std::endl;
            f << hd << std::endl;
            for (int i = 0; i < 6; i++) {
                std::string r = Field(std::string(frows[i]), 30);
                for (int b = 0; b < NBeta; b++) r += Field(std::abs(X(i, b)) / Gravity, 15) + Field(std::arg(X(i, b)), 15);
                f << r << std::endl;
            }
            f << "\n";
            f << "\n";
        };
        f << "Froude-Krylov Forces:\n";
        table(FK_i, " Mag(Fh)", " Arg(Fh)");
        f << "Excitation Forces:\n";
        table(F_i, " Mag(Xh)", " Arg(Xh)");
        f << "----------";
        f << "\n";
        f << "\n";
    }

    // Hydrodynamic_Solver_IO.cpp:430-500 (files cleared; .hst, .mmx written)
    void Generate_Output_File_WAMIT(Boundary* B) {
        const char* ext[] = {".1", ".2", ".2fk", ".2sc", ".4", ".6", ".hst", ".mmx", "_KochinRad.bem", "_KochinDiff.bem"};
        for (const char* e : ext) std::ofstream(OutputPath + e, std::ofstream::out | std::ofstream::trunc).close();
        {
            std::ofstream f(OutputPath + ".hst");
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++)
                    f << "     " << i + 1 << "     " << j + 1 << "   " << Field(C_ij(i, j) / Rho / Gravity) << std::endl;
        }
        {
            std::ofstream f(OutputPath + ".mmx");
            f << std::endl;
            f << "Gravity:     " << Grav << "     Length scale:        1.00000" << std::endl;
            f << std::endl;
            f << "Volume     " << B->Volume << std::endl;
            f << "Center of Buoyancy (Xb,Yb,Zb):   " << Field(B->Centre_Buoyancy.x) << Field(B->Centre_Buoyancy.y)
              << Field(B->Centre_Buoyancy.z) << std::endl;
            f << "Center of Gravity  (Xg,Yg,Zg):   " << Field(B->Centre_Gravity.x) << Field(B->Centre_Gravity.y)
              << Field(B->Centre_Gravity.z) << std::endl;
            f << std::endl;
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++)
                    f << "     " << i + 1 << "     " << j + 1 << "   " << Field(M_ij(i, j) / Rho) << std::endl;
        }
        // exterior wave-elevation points (:486-497): index, x, y. As there, the
        // list of files cleared above has ".6", not the ".6r" that
        // Update_Output_File appends to.
        if (!Wave_Nodes.empty()) {
            std::ofstream f(OutputPath + "_WaveElPos.bem");
            for (size_t i = 0; i < Wave_Nodes.size(); i++)
                f << i << "     " << Field(Wave_Nodes[i].x) << Field(Wave_Nodes[i].y) << std::endl;
        }
    }

    // Hydrodynamic_Solver_IO.cpp:502-581 (radiation and Haskind force files)
    void Update_Output_File_WAMIT() {
        {
            std::ofstream f(OutputPath + ".1", std::ios_base::app);
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) {
                    if (Omega == 0.0)
                        f << Field(0.0) << i + 1 << "     " << j + 1 << "     " << Field(A_ij(i, j)) << std::endl;
                    else if (Frequency == FreqInf)
                        f << Field(-1.0) << i + 1 << "     " << j + 1 << "     " << Field(A_ij(i, j)) << std::endl;
                    else
                        f << Field(Frequency) << i + 1 << "     " << j + 1 << "     " << Field(A_ij(i, j))
                          << Field(B_ij(i, j)) << std::endl;
                }
        }
        auto forces = [&](const char* ext, const CMatrix& X, Real sg) {
            std::ofstream f(OutputPath + ext, std::ios_base::app);
            for (int j = 0; j < NBeta; j++)
                for (int i = 0; i < 6; i++) {
                    if (Omega == 0.0 || Frequency == FreqInf) continue;
                    f << Field(Frequency) << Field(BetaArray[j] * R2D) << i + 1 << "  " << Field(std::abs(X(i, j)) / Gravity)
                      << Field(sg * std::arg(X(i, j))) << Field(std::real(X(i, j)) / Gravity)
                      << Field(sg * std::imag(X(i, j)) / Gravity) << std::endl;
                }
        };
        forces(".2", F_i, 1.0);
        forces(".2fk", FK_i, 1.0);
        forces(".2sc", SC_i, -1.0);   // the reference writes -arg and -Im for the scattering force
        // Kochin functions (Hydrodynamic_Solver_IO.cpp:589-621)
        const bool kochin = Omega != 0.0 && Frequency != FreqInf && KochinRad.rows() == NKoch;
        {
            std::ofstream f(OutputPath + "_KochinRad.bem", std::ios_base::app);
            for (int i = 0; kochin && i < NKoch; i++) {
                Real th = (0.5 + i) * TwoPI / NKoch;
                for (int d = 0; d < 6; d++)
                    f << Field(Frequency) << d + 1 << "  " << Field(th * R2D) << Field(std::abs(KochinRad(i, d)))
                      << Field(std::arg(KochinRad(i, d))) << Field(std::real(KochinRad(i, d)))
                      << Field(std::imag(KochinRad(i, d))) << std::endl;
            }
        }
        {
            std::ofstream f(OutputPath + "_KochinDiff.bem", std::ios_base::app);
            for (int d = 0; kochin && d < NBeta; d++)
                for (int i = 0; i < NKoch; i++) {
                    Real th = (0.5 + i) * TwoPI / NKoch;
                    f << Field(Frequency) << Field(BetaArray[d] * R2D) << "  " << Field(th * R2D)
                      << Field(std::abs(KochinDiff(i, d))) << Field(std::arg(KochinDiff(i, d)))
                      << Field(std::real(KochinDiff(i, d))) << Field(std::imag(KochinDiff(i, d))) << std::endl;
                }
        }
        // radiated wave elevation at the exterior points (:624-643): frequency,
        // point index, |eta_j| and arg eta_j for the six modes
        if (!Wave_Nodes.empty() && Omega != 0.0 && Frequency != FreqInf &&
            ExtRadMat.rows() == (int)Wave_Nodes.size()) {
            std::ofstream f(OutputPath + ".6r", std::ios_base::app);
            for (int i = 0; i < ExtRadMat.rows(); i++) {
                f << Field(Frequency) << i << "     ";
                for (int d = 0; d < 6; d++) f << Field(std::abs(ExtRadMat(i, d))) << Field(std::arg(ExtRadMat(i, d)));
                f << std::endl;
            }
        }
    }

    bool check(int rc) {
        if (rc == BEM_OK) return true;
        std::cout << "BEMUse: " << bem_last_error() << std::endl;
        return false;
    }
    void set_nan() {
        const Real n = std::numeric_limits<Real>::quiet_NaN();
        A_ij = Matrix(6, 6, n); B_ij = Matrix(6, 6, n);
        F_i = CMatrix(6, NBeta, CReal(n, n)); FK_i = F_i; SC_i = F_i;
    }
    // r x c complex, row-major interleaved
    static CMatrix unpack_rows(const std::vector<double>& v, int r, int c) {
        CMatrix M = CMatrix::Zero(r, c);
        for (int i = 0; i < r; i++)
            for (int j = 0; j < c; j++) M(i, j) = CReal(v[2 * ((size_t)i * c + j)], v[2 * ((size_t)i * c + j) + 1]);
        return M;
    }
    static CMatrix unpack(const std::vector<double>& v, int r, int c) {
        CMatrix M(r, c);
        for (int j = 0; j < c; j++)
            for (int i = 0; i < r; i++) M(i, j) = CReal(v[2 * ((size_t)j * r + i)], v[2 * ((size_t)j * r + i) + 1]);
        return M;
    }

    int device_ = 0, mode_ = BEM_MODE_FP64;
    bem_ctx* ctx_ = nullptr;
    bem_body* body_ = nullptr;
    int NPTOT = 0;
    Real Kappa = 0, Frequency = 0, Period = 0, Omega = 0;
    bool IFR = false;
    Real Rho = Rho_wat, Grav = Gravity;
    int NBeta = 0, NKoch = 36;
    std::vector<Real> BetaArray;
    Real FreqInf = -1.0;
    Matrix A_ij = Matrix::Zero(6, 6), B_ij = Matrix::Zero(6, 6);
    Matrix M_ij = Matrix::Zero(6, 6), C_ij = Matrix::Zero(6, 6);
    CMatrix F_i, FK_i, SC_i;
    CMatrix KochinRad, KochinDiff;
    std::vector<Vector3> Wave_Nodes;
    CMatrix ExtRadMat;                  // radiated elevation at Wave_Nodes of the last Solve (npts x 6)
    bem_solve_info Info = bem_solve_info();
};

}  // namespace BEMUse

#endif  // BEMUSE_HYDRO_H
