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
        Report();
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

// Half_Cylinder (Volume_of_Revolution.cpp:167-199): open-top vertical cylinder
// with a bottom disc (triangle fan at the centre).
class Half_Cylinder : public Boundary {
public:
    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params) {
            if (P.myNameis("Radius")) R = P.Get_Param<Real>();
            if (P.myNameis("Draft")) D = P.Get_Param<Real>();
            if (P.myNameis("NPanels_Radial")) NR = P.Get_Param<int>();
            if (P.myNameis("NPanels_Vertical")) NV = P.Get_Param<int>();
            if (P.myNameis("NPanels_Azimuthal")) NA = P.Get_Param<int>();
            if (P.myNameis("Cosine_Disc")) Cosine = P.Get_Param<bool>();
            if (P.myNameis("FreeSurface_Radius")) RFS = P.Get_Param<Real>();
            if (P.myNameis("NPanels_FreeSurface_Radial")) NRES = P.Get_Param<int>();
            if (P.myNameis("NPanels_FreeSurface_Int_Radial")) NRFS = P.Get_Param<int>();
        }
    }
    void Setup() override {
        bem_mesh* m = nullptr;
        int rc = bem_mesh_half_cylinder(R, D, NR, NV, NA, Cosine ? 1 : 0, &m);
        Adopt(rc, m);
        // interior lid (Volume_of_Revolution.cpp:43-65, 111-130)
        if (NRFS > 0) {
            bem_mesh* l = nullptr;
            int rl = bem_mesh_lid_disc(R, NRFS, NA, Cosine ? 1 : 0, &l);
            Adopt_Lid(rl, l);
        }
        // exterior free-surface rings r = 1.001 R + RFS R i / NRES
        // (Volume_of_Revolution.cpp:69-89; with NRES = 0 the reference divides
        // by zero there, so no rings are made then)
        FreeSurface_Nodes.clear();
        for (int i = 0; NRES > 0 && i <= NRES; i++) {
            Real r = R * 1.001 + RFS * R * i / NRES;
            for (int j = 0; j < NA; j++) {
                Real u = j * 2.0 * 3.14159265358979323846 / NA;
                FreeSurface_Nodes.push_back(Vector3(r * std::cos(u), r * std::sin(u), 0.0));
            }
        }
        Report();
    }
private:
    Real R = 10, D = 20, RFS = 0;
    int NR = 8, NV = 24, NA = 32, NRES = 0, NRFS = 0;
    bool Cosine = true;
};

// Barge (Barge.cpp:12-191): rectangular box piercing the free surface.
class Barge : public Boundary {
public:
    void Set_Parameters(std::vector<Parameter>& Params) override {
        for (const Parameter& P : Params) {
            if (P.myNameis("Tank_Length")) L = P.Get_Param<Real>();
            if (P.myNameis("Tank_Width")) W = P.Get_Param<Real>();
            if (P.myNameis("Tank_Depth")) D = P.Get_Param<Real>();
            if (P.myNameis("NPanels_Tank_Length")) NX = P.Get_Param<int>();
            if (P.myNameis("NPanels_Tank_Width")) NY = P.Get_Param<int>();
            if (P.myNameis("NPanels_Tank_Depth")) NZ = P.Get_Param<int>();
        }
    }
    void Setup() override {
        bem_mesh* m = nullptr;
        int rc = bem_mesh_box(L, W, -D, 0.0, NX, NY, NZ, 1, &m);
        Adopt(rc, m);
        // interior lid (Barge.cpp:103-123, 193-204; y from its own index)
        bem_mesh* l = nullptr;
        int rl = bem_mesh_lid_rect(L, W, NX, NY, 0, &l);
        Adopt_Lid(rl, l);
        Report();
    }
private:
    Real L = 20, W = 10, D = 5;
    int NX = 20, NY = 10, NZ = 5;
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
        Report();
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
    std::vector<CMatrix> RadSolArray;     // radiation potential per Solve (N x 6)
    std::vector<CMatrix> DiffSolArray;    // scattering potential per Solve (N x NBeta)
    std::vector<CMatrix> FS_Rad_SolArray;
    std::vector<CMatrix> FS_Scat_SolArray;
protected:
    std::string OutputDirectory = "Output";
    std::string OutputPath = "Output/BEMUse_Outputs";
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
                FS_Rad_SolArray.push_back(unpack_rows(er, nw, 6));
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
    bem_solve_info Info = bem_solve_info();
};

}  // namespace BEMUse

#endif  // BEMUSE_HYDRO_H
