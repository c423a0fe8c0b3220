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
 * What is added: the device and arithmetic mode (constructor), accessors for
 * A, B, F, FK, SC (private members in the reference, read from its output
 * files), and Solve_Frequencies for a pipelined sweep over many frequencies.
 * Not provided: IFR lid panels, Kochin functions, exterior free-surface
 * elevation, RAOs and the output files (reported, not computed).
 */
#ifndef BEMUSE_HYDRO_H
#define BEMUSE_HYDRO_H

#include <array>
#include <cmath>
#include <complex>
#include <iostream>
#include <limits>
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

static const Real TwoPIinv = 1.0 / (2.0 * 3.14159265358979323846);
static const Real Gravity = 9.81;      // BEMUse_Math_Types.h:115 (K and Phi_I use this constant)
static const Real Rho_wat = 1025.0;    // BEMUse_Math_Types.h:113

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
    virtual ~Boundary() { if (mesh_) bem_mesh_free(mesh_); }
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
protected:
    void Adopt(int rc, bem_mesh* m) {
        if (mesh_) { bem_mesh_free(mesh_); mesh_ = nullptr; }
        if (rc != BEM_OK) { std::cout << "BEMUse: mesh generation failed: " << bem_last_error() << std::endl; return; }
        mesh_ = m;
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
        }
    }
    void Setup() override {
        bem_mesh* m = nullptr;
        int rc = bem_mesh_half_cylinder(R, D, NR, NV, NA, Cosine ? 1 : 0, &m);
        Adopt(rc, m);
    }
private:
    Real R = 10, D = 20;
    int NR = 8, NV = 24, NA = 32;
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
        }
    }
    void Setup() override {
        bem_mesh* m = nullptr;
        int rc = bem_mesh_semisub3(Rc, D, S, PW, PH, 0.5, 44, 24, 8, 31, 5, 4, &m);
        Adopt(rc, m);
    }
private:
    Real Rc = 6, D = 20, S = 50, PW = 6, PH = 4;
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
        if (IFR) std::cout << "BEMUse: irregular-frequency removal (IFR) is not provided; solving without lid panels" << std::endl;
    }
    void Set_Real(Real D) override { Frequency = D; }

    void Setup(Boundary* B) override {
        if (!B || !B->Get_Device_Mesh()) { std::cout << "BEMUse: Setup without a generated boundary" << std::endl; return; }
        if (!ctx_ && !check(bem_create(device_, &ctx_))) return;
        if (body_) { bem_body_destroy(body_); body_ = nullptr; }
        if (!check(bem_body_create(ctx_, B->Get_Device_Mesh(), &body_))) return;
        NPTOT = bem_body_n(body_);
        if (!check(bem_body_assemble_rankine(ctx_, body_))) return;
        A_ij = Matrix::Zero(6, 6);
        B_ij = Matrix::Zero(6, 6);
        F_i = CMatrix::Zero(6, NBeta);
        FK_i = CMatrix::Zero(6, NBeta);
        SC_i = CMatrix::Zero(6, NBeta);
        // the "infinite frequency" radiation solve (Hydrodynamic_Solver.cpp:295-297)
        Frequency = FreqInf;
        Solve();
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
        if (Keep_Potentials) { phiJ.resize(2 * (size_t)NPTOT * 6); phiS.resize(2 * (size_t)NPTOT * NBeta); }
        if (!check(bem_solve_frequency(ctx_, body_, mode_, Omega, Kappa, compose, rad_only, NBeta,
                                       NBeta ? BetaArray.data() : nullptr, A.data(), Bm.data(),
                                       NBeta ? F.data() : nullptr, NBeta ? FK.data() : nullptr,
                                       NBeta ? SC.data() : nullptr, Keep_Potentials ? phiJ.data() : nullptr,
                                       Keep_Potentials && NBeta ? phiS.data() : nullptr, &Info))) {
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
    int N_Panels() const { return NPTOT; }

    bool Keep_Potentials = false;

private:
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
    CMatrix F_i, FK_i, SC_i;
    bem_solve_info Info = bem_solve_info();
};

}  // namespace BEMUse

#endif  // BEMUSE_HYDRO_H
