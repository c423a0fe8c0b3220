/*
 * bemuse_amd: C-ABI of the MI355X (gfx950) implementation of the BEMUse
 * per-frequency hydrodynamic hot path (Rankine + Havelock influence assembly,
 * dense complex LU, radiation/diffraction solves, added mass / damping /
 * excitation reductions).
 *
 * Conventions
 *   - Every entry point returns 0 (BEM_OK) on success or a BEM_ERR_* code;
 *     bem_last_error() then returns a message (thread-local).
 *   - Buffers are caller-owned host memory. Matrices are column-major.
 *     Complex values are interleaved (re, im) doubles, layout-compatible with
 *     std::complex<double>.
 *   - The reference exposes this path only through C++ virtuals
 *     (src/Solver/Solver_Base.h:119-151, src/Solver/Hydrodynamic_Solver.h:135-160);
 *     the C++ drop-in class in include/bemuse_hydro.h sits on top of these
 *     entry points. Each entry point names the reference seam it replaces.
 *   - One context per host thread. Calls on one context are serialised on the
 *     context's HIP stream.
 */
#ifndef BEMUSE_AMD_H
#define BEMUSE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BEM_OK = 0,
    BEM_ERR_ARG = 1,      /* invalid argument */
    BEM_ERR_HIP = 2,      /* HIP runtime error (no device, launch failure, ...) */
    BEM_ERR_NOMEM = 3,    /* device allocation failed */
    BEM_ERR_NUMERIC = 4,  /* zero or non-finite pivot, or a non-finite result: the outputs are
                             still written (NaN where the system is singular), and a sweep or
                             batch still solves its other frequencies */
    BEM_ERR_STATE = 5     /* call out of order (e.g. solve before Rankine assembly) */
};

/* Arithmetic mode of bem_solve_frequency. */
enum {
    BEM_MODE_FP64 = 0,            /* FP64 assembly, FP64-MFMA LU, FP64 solves */
    BEM_MODE_MIXED = 1,           /* FP64 assembly, FP32-MFMA LU, FP64 iterative refinement */
    BEM_MODE_MIXED_SURROGATE = 2  /* as MIXED, wave term from the trained surrogate inside its domain */
};

/* Influence-matrix composition case (Hydrodynamic_Solver.cpp:415-429). */
enum {
    BEM_COMPOSE_WAVE = 0,  /* S = Ssrc + Srefl + Swave, D = Dsrc + Drefl + Dwave */
    BEM_COMPOSE_ZERO = 1,  /* Omega == 0: S = Ssrc + Srefl, D = Dsrc + Drefl      */
    BEM_COMPOSE_INF = 2    /* Omega == FreqInf: S = Ssrc - Srefl, D = Dsrc - Drefl */
};

typedef struct bem_mesh bem_mesh;
typedef struct bem_ctx bem_ctx;
typedef struct bem_body bem_body;

/* Diagnostics of one frequency solve. */
typedef struct bem_solve_info {
    int refine_iters;        /* iterative-refinement steps taken (MIXED modes) */
    int fell_back_fp64;      /* 1 if refinement stagnated and the FP64 LU was used */
    int pivot_fallback;      /* panels whose speculative (diagonal) pivots were rejected */
    double residual;         /* final max-norm relative residual |b - A x| / (|A| |x|) */
    double ms_assemble;      /* device time of wave assembly + compose [ms] */
    double ms_factor;        /* device time of the LU [ms] */
    double ms_solve;         /* device time of RHS, solves, refinement, reductions [ms] */
    int64_t surrogate_points; /* wave-kernel evaluations served by the surrogate */
    int64_t exact_points;     /* wave-kernel evaluations served by the exact kernel */
    double surrogate_err;     /* largest relative deviation of the surrogate from the exact
                                 kernel over the pairs judged during assembly (one in 61) */
    int lanes;                /* solve lanes the call ran its frequencies on */
} bem_solve_info;

const char* bem_last_error(void);

/* ---- meshes (host only; no GPU needed) -------------------------------------
 * Node/element arrays as consumed by Boundary::Get_Elements
 * (src/Boundary/Boundary_Base.h:168-172): nodes n x 3, elements e x 4 with
 * elems[4e+3] = -1 for a triangle. */
int bem_mesh_from_arrays(int n_nodes, const double* nodes, int n_elems, const int* elems, bem_mesh** out);
/* (n_nodes = n_elems = 0 makes an empty mesh to build on with bem_mesh_append.) */
/* Half_Cylinder (src/Boundary/Volume_of_Revolution.cpp:167-199). */
int bem_mesh_half_cylinder(double radius, double draft, int n_radial, int n_vertical,
                           int n_azimuthal, int cosine, bem_mesh** out);
/* Interior free-surface lids for irregular-frequency removal (the auxiliary
 * panels of Boundary::Generate_Aux_Nodes/_Elements): a disc of radius R in
 * z = 0 (Volume_of_Revolution.cpp:43-65, 111-130; n_radial rings, a triangle
 * fan at the centre) and an L x W rectangle (Barge.cpp:103-123, 193-204; the
 * reference computes the lid's y from the x index, :115, which makes its
 * quads degenerate; here y follows its own index). */
int bem_mesh_lid_disc(double radius, int n_radial, int n_azimuthal, int cosine, bem_mesh** out);
int bem_mesh_lid_rect(double L, double W, int nx, int ny, int cosine, bem_mesh** out);
/* The elliptic lid of a semi-ellipsoid (Ellipsoid.cpp:326-362); a = b = R is
 * bem_mesh_lid_disc. */
int bem_mesh_lid_ellipse(double a, double b, int n_radial, int n_azimuthal, int cosine, bem_mesh** out);
/* Reference parametric bodies, node for node (same node and panel order):
 *  - a volume of revolution from a profile (r_i, z_i), r_0 = 0: one axis node,
 *    n_azimuthal nodes per further point, a triangle fan round the axis node,
 *    quads between rings (Volume_of_Revolution.cpp:12-37, 92-109);
 *  - Tapered_SparBuoy (Volume_of_Revolution.cpp:203-251; keys Lower_Radius,
 *    Upper_Radius, Draft, Height_Taper_Begin, Height_Taper_End, NPanels_Radial,
 *    NPanels_Vertical_Section1..3, NPanels_Azimuthal, Cosine_Disc);
 *  - Spar_Leg (Volume_of_Revolution.cpp:253-297; Leg_Lower_Radius,
 *    Leg_Upper_Radius, Draft, Heave_Plate_Height, NPanels_Radial_Base,
 *    NPanels_Vertical_HeavePlate, NPanels_Horizontal_HeavePlate,
 *    NPanels_Vertical, NPanels_Azimuthal, Cosine_Disc);
 *  - Barge (Barge.cpp:12-191; cosine spacing is the reference default);
 *  - Ellipsoid, centre at z = -depth, and Semi_Ellipsoid (Ellipsoid.cpp:12-310;
 *    triangles = Triangular_Panels);
 *  - Wigley_Hull (Ship_Hulls.cpp:12-200; outward = Panels_Outward). */
int bem_mesh_revolution(int n, const double* r, const double* z, int n_azimuthal, bem_mesh** out);
int bem_mesh_tapered_spar(double r_lower, double r_upper, double draft, double h_taper_begin, double h_taper,
                          int n_radial, int n_vert1, int n_vert2, int n_vert3, int n_azimuthal, int cosine,
                          bem_mesh** out);
int bem_mesh_spar_leg(double r_lower, double r_upper, double draft, double h_plate, int n_base, int n_plate_vert,
                      int n_plate_horiz, int n_vert, int n_azimuthal, int cosine, bem_mesh** out);
int bem_mesh_barge(double L, double W, double D, int nx, int ny, int nz, int cosine, bem_mesh** out);
int bem_mesh_ellipsoid(double a, double b, double c, double depth, int n_axial, int n_azimuthal, int cosine,
                       int triangles, bem_mesh** out);
int bem_mesh_semi_ellipsoid(double a, double b, double c, int n_axial, int n_azimuthal, int cosine, int triangles,
                            bem_mesh** out);
int bem_mesh_wigley(double L, double B, double T, int nx, int nz, int cosine, int triangles, int outward,
                    bem_mesh** out);
/* Appends src, translated by (dx, dy, dz), to dst (e.g. one lid per column). */
int bem_mesh_append(bem_mesh* dst, const bem_mesh* src, double dx, double dy, double dz);
/* Rectangular box x in [-L/2,L/2], y in [-W/2,W/2], z in [z0,z1]
 * (cf. src/Boundary/Barge.cpp:34-191); open_top = 1 for a surface-piercing barge. */
int bem_mesh_box(double L, double W, double z0, double z1, int nx, int ny, int nz, int open_top, bem_mesh** out);
/* Three columns + three closed pontoons (the 6k-panel semi-submersible workload). */
int bem_mesh_semisub3(double col_radius, double draft, double spacing, double pontoon_width,
                      double pontoon_height, double gap, int n_azimuthal, int n_vertical,
                      int n_radial, int nx, int ny, int nz, bem_mesh** out);
/* Truncated-cone spar with heave plate (single volume of revolution). */
int bem_mesh_cone_spar(double r_top, double r_bot, double draft, double r_plate, double t_plate,
                       int n_azimuthal, int n_disc, int n_rim, int n_annulus, int n_cone, bem_mesh** out);
/* Body `index` of the seeded uneven batch; params (6, optional): type (0 cyl,
 * 1 barge), dim a, dim b, draft, target panels, actual panels. */
int bem_mesh_random_body(uint64_t seed, int index, double* params, bem_mesh** out);
/* Mesh files as the reference reads them: GDF (GDF_Geo.cpp:12-103), MAR
 * (MAR_Geo.cpp:12-97), PNL (PNL_Geo.cpp:16-131), ASCII STL (STL_Geo.cpp:12-96,
 * coordinates x100 and reversed vertex order, as there). format: BEM_MESH_*;
 * AUTO picks by extension. */
enum { BEM_MESH_AUTO = 0, BEM_MESH_GDF = 1, BEM_MESH_MAR = 2, BEM_MESH_PNL = 3, BEM_MESH_STL = 4 };
int bem_mesh_read(const char* path, int format, bem_mesh** out);
int bem_mesh_size(const bem_mesh* m, int* n_nodes, int* n_elems);
int bem_mesh_export(const bem_mesh* m, double* nodes, int* elems);
void bem_mesh_free(bem_mesh* m);

/* Per-panel records (host only), for inspection and tests: body panels and
 * mirrored panels (N x 32 doubles each, see bem_internal.h PanelField),
 * mirrored quadrature records (N x 32), BC points (N x 4), generalised
 * normals (N x 8: n, x x n, area, 0). Any pointer may be NULL.
 * Replaces Create_Panels / Specify_BC_Positions / Set_RHS_Mat precompute
 * (Hydrodynamic_Solver.cpp:30-81,117-175; Solver_Base.cpp:71-83). */
int bem_geometry(const bem_mesh* m, double* body, double* refl, double* quad, double* recv, double* gen);

/* ---- context and bodies ----------------------------------------------------- */
int bem_create(int device, bem_ctx** out);
void bem_destroy(bem_ctx* ctx);

/* Frequency graphs (off by default): from the second consecutive frequency
 * of one configuration (body, mode, headings, compose case) on a solve lane,
 * the lane's whole per-frequency sequence (assembly, factorisation, solves,
 * refinement, reductions, result copies) is captured once as a HIP graph and
 * replayed for the following frequencies. Results are identical either way.
 * They pay off when hardware queues are scarce (GPU_MAX_HW_QUEUES = 4:
 * 38 -> 44 frequency-solves/s at 6k panels, 8 lanes) and lose when the lanes
 * have their own queues (32 queues: 60 eager vs 36 graphed), which is why
 * they are off by default. Inside a graph the stages are not timed
 * separately: bem_solve_info then reports the whole frequency in ms_solve
 * (ms_assemble = ms_factor = 0). No counterpart in the reference (its
 * solves are host loops). */
int bem_set_graphs(bem_ctx* ctx, int enable);

/* Iterative refinement of the mixed modes: at most max_steps (0..4, default
 * 4) correction steps per frequency. A frequency whose refinement has not
 * reached the residual target by then (or stagnates earlier) is re-solved
 * with the FP64 LU and reports fell_back_fp64 = 1; max_steps = 0 therefore
 * sends every mixed frequency to the FP64 path (which is how the tests reach
 * that path). No counterpart in the reference (it has no refinement). */
int bem_set_refine_limit(bem_ctx* ctx, int max_steps);

/* Correction steps of the mixed modes' refinement that a solve lane enqueues
 * with its frequency (0..4, default 2): every enqueued step that the device
 * skips once the refinement has converged still costs a launch on the lane's
 * chain, so the lane enqueues these and, when its results come back with the
 * refinement neither converged nor stopped, enqueues the rest (up to the
 * limit of bem_set_refine_limit). Results are the same for any value. No
 * counterpart in the reference. */
int bem_set_refine_eager(bem_ctx* ctx, int steps);

/* Arithmetic of the trailing updates of the FP32 factorisation in the mixed
 * modes (the FP64 refinement that follows makes both give the same results
 * to the refinement tolerance):
 *   BEM_LU_SPLIT_F16 (default): fp16 MFMA on operands split into two scaled
 *     fp16 terms each, three products per real product, FP32 accumulation
 *     (products good to about 2^-22, about 4x the FP32 rounding); the
 *     updates are then bound by their FP32 C traffic;
 *   BEM_LU_FP32: FP32 MFMA (16x16x4);
 *   BEM_LU_SPLIT_F16_TILE: the arithmetic of BEM_LU_SPLIT_F16 with the split
 *     done in every tile of every update instead of once per panel group
 *     (results equal to BEM_LU_SPLIT_F16 bit for bit; slower).
 * No counterpart in the reference (its LU is Eigen's). */
enum { BEM_LU_FP32 = 0, BEM_LU_SPLIT_F16 = 1, BEM_LU_SPLIT_F16_TILE = 2 };
int bem_set_lu_arith(bem_ctx* ctx, int arith);

/* Tile of the BEM_LU_SPLIT_F16 trailing updates (the products on operands
 * split once per panel group): -1 (default) takes the autotuner's choice per
 * update shape; 0, 1, 2 force 64x64, 128x64, 64x128 tiles of 32x32 wave
 * blocks with operand loads one 16-k chunk ahead, 5, 6, 7 the same tiles with
 * loads two chunks ahead; 3, 4 force 128x128 tiles of 64x64 wave blocks with
 * one or two operand chunks in LDS; 8 the 128x128 tile whose operands go
 * from memory straight to LDS (double-buffered). Every tile adds the same
 * products in the same order, so the choice never changes results. No
 * counterpart in the reference. */
int bem_set_lu_tile(bem_ctx* ctx, int tile);

/* How the solve lanes (bem_solve_frequency, bem_solve_sweep,
 * bem_solve_batch) factor: enable = 1 (default) speculatively, every panel
 * with its diagonal as pivot and no pivot-check launch between the panel
 * kernels, a frequency whose factorisation met a multiplier above the pivot
 * threshold (bem_set_pivot_threshold) being run again as with 0; enable = 0:
 * the pivot check after every panel, with the partial-pivoting
 * refactorisation of a rejected one in place. Both give the same results bit
 * for bit wherever no panel is rejected. No counterpart in the reference. */
int bem_set_lu_speculative(bem_ctx* ctx, int enable);

/* Multiplier bound of the factorisations' pivot check, tau in [0, 1]: a panel
 * with a multiplier |l| > tau (or a zero or non-finite pivot) is factored
 * again with partial pivoting (the reference's PartialPivLU, Solver_Base.h:85,
 * pivots by largest modulus). 1 (default): exactly the panels where partial
 * pivoting would choose another pivot than the diagonal; below 1 more panels
 * take the partial-pivoting path (0: every panel). Applies to every solve
 * path (lanes, bem_dense_solve, the no-allocation solve). */
int bem_set_pivot_threshold(bem_ctx* ctx, double tau);

/* Measurement option, off by default: the wave term of receiver / source
 * pairs whose first mirrored quadrature node lies farther than near_radius
 * (metres) from the receiver is evaluated with FP32 exp, log, Bessel, Struve
 * and sin/cos (the Green function's polynomials and combination stay FP64),
 * in the FP64 and mixed modes (the surrogate mode's exact pairs and the
 * field-point evaluation keep FP64). near_radius < 0 turns it off (the exact FP64 wave term
 * everywhere, the default); 0 sends every pair to it. Results then differ
 * from the exact ones by the FP32 error of those functions (about 1e-7
 * relative per term). No counterpart in the reference (its wave term is
 * FP64, or FP32 throughout in its single-precision build). */
int bem_set_wave_far32(bem_ctx* ctx, double near_radius);

/* Work-groups per CU of the wave assembly (the exact FP64 path): n > 0 runs
 * n x CUs work-groups that take the 64 x 16 tiles in turn, 0 one work-group
 * per tile (every CU filled), -1 (the default) n = 2 below padded order 6144
 * and 0 from there: the lanes of small bodies spend most of their time in
 * chains of short factorisation launches, which an assembly filling every
 * CU delays (the 64-body batch runs 2-3% faster capped; the 6018-panel sweep
 * 1% faster uncapped). Results do not depend on it. No counterpart in the
 * reference (a scheduling option). */
int bem_set_wave_wg_per_cu(bem_ctx* ctx, int n);

/* Measured machine peaks of the context's device, for reporting beside the
 * published ones: out[0] FP32 MFMA TFLOP/s, out[1] FP64 MFMA TFLOP/s (dense
 * 16x16x4 chains on registers), out[2] streaming-copy GB/s (1 GiB read +
 * 1 GiB written), out[3] fp16 MFMA TFLOP/s (32x32x16 chains). Allocates 2 GiB
 * for the copy while it runs. No counterpart in the reference (a measurement
 * aid). */
int bem_measure_peaks(bem_ctx* ctx, double* out);

/* Uploads the geometry of a mesh (Hydrodynamic_Radiation_Solver::Setup,
 * Hydrodynamic_Solver.cpp:269-279, geometry part). */
int bem_body_create(bem_ctx* ctx, const bem_mesh* m, bem_body** out);
/* As bem_body_create with irregular-frequency removal (IFR, Create_Panels
 * Hydrodynamic_Solver.cpp:45-55): the lid panels (z = 0) join the system
 * after the body panels, so assembly, factorisation and solves run over all
 * of them, while potentials, A, B, F and the post-processing use the body
 * panels only (:145-165, 450-458). lid may be NULL. */
int bem_body_create_lid(bem_ctx* ctx, const bem_mesh* m, const bem_mesh* lid, bem_body** out);
void bem_body_destroy(bem_body* b);
/* Body panels (the rows of the returned potentials). */
int bem_body_n(const bem_body* b);
/* Order of the linear system: body plus lid panels (the size of the matrices
 * bem_get_rankine and bem_get_system return). */
int bem_body_n_system(const bem_body* b);

/* Rankine source/dipole influence of the body and its image
 * (Prepare_Linear_System, Hydrodynamic_Solver.cpp:83-115), once per body. */
int bem_body_assemble_rankine(bem_ctx* ctx, bem_body* b);

/* One frequency (Hydrodynamic_Radiation_Solver::Solve,
 * Hydrodynamic_Solver.cpp:400-477): wave assembly, compose, LU, radiation and
 * diffraction solves, A/B/F reductions.
 *   omega   the solver's internal Omega (= Frequency / 2 pi, :405)
 *   kappa   wavenumber (Omega^2 / 9.81 in infinite depth, :409)
 *   compose BEM_COMPOSE_*; rad_only = 1 stops after A/B (:469)
 *   betas   nb wave headings [rad]
 * Outputs: A36, B36 row-major 6x6; F, FK, SC row-major 6 x nb complex (may be
 * NULL); phiJ N x 6 and phiS N x nb complex column-major (may be NULL).
 * BEM_ERR_NUMERIC when the system is numerically singular or a result is not
 * finite (A, B, F are written; the potentials are not).
 * BEM_ERR_ARG for a non-finite omega or kappa, or kappa < 0 (omega itself may be
 * negative: the reference's set-up solve runs at Frequency = -1); the same
 * check applies to every frequency of the sweep, batch and post-processing
 * entry points. */
int bem_solve_frequency(bem_ctx* ctx, bem_body* b, int mode, double omega, double kappa,
                        int compose, int rad_only, int nb, const double* betas,
                        double* A36, double* B36, double* F, double* FK, double* SC,
                        double* phiJ, double* phiS, bem_solve_info* info);

/* Several frequencies of one body (the loop of BEMUse_Console::Execute,
 * src/BEMUser_Console.cpp:252-257, each iteration Solve()), pipelined over
 * `lanes` (1..16) concurrent solve lanes, each with its own HIP streams and
 * workspaces, so that one frequency's dense MFMA work overlaps another's
 * latency-bound panel factorisations and triangular solves. lanes = 0 lets
 * the library choose: at most 10, spread so that every round of frequencies
 * fills the lanes evenly, within half the free device memory (infos[f].lanes
 * reports the count). kappas may be
 * NULL (Omega^2 / 9.81). Outputs per frequency f: A36 + 36 f, B36 + 36 f,
 * F/FK/SC + 12 nb f (row-major 6 x nb complex), infos + f (may be NULL).
 * Results are identical to nf calls of bem_solve_frequency. */
int bem_solve_sweep(bem_ctx* ctx, bem_body* b, int mode, int nf, const double* omegas, const double* kappas,
                    int compose, int nb, const double* betas, int lanes, double* A36, double* B36, double* F,
                    double* FK, double* SC, bem_solve_info* infos);

/* Many bodies of uneven size, each with its own frequency list (the reference
 * runs one Solver per body, BEMUser_Console.cpp:227-260): nfreq[i] frequencies
 * of body i, concatenated in omegas; results in the same order (A36 + 36 k,
 * B36 + 36 k, F + 12 nb k). All tasks share a pool of `lanes` solve lanes
 * (0: chosen as for bem_solve_sweep over all tasks)
 * owned by the context, sized for the largest body; larger bodies are started
 * first, and each task goes to the first lane that is free. Same headings for
 * all bodies; compose WAVE. Results equal those of per-body sweeps. */
int bem_solve_batch(bem_ctx* ctx, int nbody, bem_body* const* bodies, const int* nfreq, const double* omegas,
                    int mode, int nb, const double* betas, int lanes, double* A36, double* B36, double* F,
                    bem_solve_info* infos);

/* ---- kernel-level access (parity checks of individual seams) ---------------- */
/* Rankine matrices as assembled on the device: SSrc = -s, SSrcRefl = -sr,
 * DSrc = d, DSrcRefl = dr (N x N real column-major; NULL skips). */
int bem_get_rankine(bem_ctx* ctx, bem_body* b, double* Ssrc, double* Srefl, double* Dsrc, double* Drefl);
/* Composed SMat and DMat for one frequency (complex N x N column-major),
 * as produced by the device assembly in the given mode. */
int bem_get_system(bem_ctx* ctx, bem_body* b, int mode, double kappa, int compose, double* SMat, double* DMat);
/* Device evaluation of HavelockGF (Kernels.cpp:342-412) at n points;
 * out: 8 doubles per point (GF, DGFX, DGFY, DGFZ as re, im). precision: 64
 * (the exact kernel), or 32 for the function the far-pair option of
 * bem_set_wave_far32 evaluates (FP32 special functions). */
int bem_eval_havelock(bem_ctx* ctx, int precision, int n, const double* X, const double* Y,
                      const double* V, double* out);
/* The constant tables the device HavelockGF is compiled with, for checking
 * against the reference literals without a GPU: the coefficients of
 * GF_FuncA..D and GF_dFuncA..C (Kernels.cpp:120-256) in that order (64
 * values), then EUL (BEMUse_Math_Types.h:99). n must be >= 65. */
int bem_gf_coefficients(double* out, int n);
/* The Struve approximation literals the device HavelockGF is compiled with
 * (StruveH0 / StruveH1, Kernels.cpp:35-116), for checking against the
 * reference text without a GPU: P0..P5, a0..a3, b1..b3 of StruveH0, then
 * P1..P6, a0..a3, b1..b3 of StruveH1 (26 values). n must be >= 26. */
int bem_struve_coefficients(double* out, int n);
/* The Bessel evaluator of the device HavelockGF (the reference calls
 * std::tr1::cyl_bessel_j / cyl_neumann, Kernels.cpp:31-34), run on the host
 * for checking without a GPU: out[4 i .. 4 i + 3] = J0, J1, Y0, Y1 at x[i] >= 0,
 * with Y0 = Y1 = 0 for x <= 3 (HavelockGF uses them only above 3). */
int bem_bessel01(int n, const double* x, double* out);
/* Device LU + solve of a caller matrix (complex n x n, column-major):
 * precision 64 (FP64 LU) or 32 (FP32-MFMA LU + FP64 refinement, FP64 LU if
 * the refinement stagnates); X receives the solution of A X = B for nrhs
 * columns. BEM_ERR_ARG for non-finite entries of A or B; BEM_ERR_NUMERIC (X
 * written) for a numerically singular A or a non-finite solution. */
int bem_dense_solve(bem_ctx* ctx, int precision, int n, const double* A, int nrhs, const double* B,
                    double* X, bem_solve_info* info);

/* As bem_dense_solve, allocating nothing on the device. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every device buffer it needs (the padded FP64 matrix and its FP32 copy, the
 * LU workspace, right-hand sides and solution, refinement and
 * triangular-solve state, staging of A, B and X) lives in one fixed device
 * array of BEM_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from
 * the first bem_create on that GPU until the process ends; its streams and
 * events are made by bem_create. Limits: n <= BEM_NOALLOC_MAX_N (BEM_ERR_ARG
 * beyond), nrhs <= 16. Results, errors and info are those of bem_dense_solve.
 * Every call on one GPU shares that GPU's array (it keeps the last call's
 * contents; each call rewrites what it reads), so calls that overlap on one
 * GPU, through the same or different contexts, race: serialise them. */
#define BEM_NOALLOC_MAX_N 512
#define BEM_NOALLOC_SCRATCH_BYTES (16u << 20)
int bem_dense_solve_noalloc(bem_ctx* ctx, int precision, int n, const double* A, int nrhs, const double* B,
                            double* X, bem_solve_info* info);

/* ---- post-processing of a solved frequency -------------------------------------
 * Both take the potentials bem_solve_frequency returns (phiJ N x 6, phiS
 * N x nb, complex column-major) at the same omega, kappa and headings. */

/* Kochin functions (Set_Kochin_Mats, Hydrodynamic_Solver.cpp:598-642, used at
 * :492-494) at theta_i = (i + 1/2) 2 pi / nkoch:
 *   H(theta) = sum_P A_P (E dphi/dn - (n . grad E) phi),
 *   E = exp(K z + i K (x cos theta + y sin theta)) at the centroids,
 * with (dphi/dn, phi) = (N_k, Phi_J) for radiation and (dPhi_I/dn, Phi_S) for
 * diffraction, as the reference combines them. The reference builds E from
 * an uninitialised value (:626-627); E here is the intended Newman (1967)
 * factor. Outputs KochRad nkoch x 6 and KochDiff nkoch x nb, complex
 * row-major (either may be NULL). */
int bem_kochin(bem_ctx* ctx, bem_body* b, double omega, double kappa, int nkoch, int nb, const double* betas,
               const double* phiJ, const double* phiS, double* KochRad, double* KochDiff);

/* Free-surface elevation of the radiated and diffracted waves at npts
 * exterior points pts (npts x 3) (Prepare_FS_Linear_System :205-237,
 * Prepare_FS_Linear_System_Wave_Terms :239-266, ExtRadMat / ExtDiffMat
 * :499-505): Rankine, image and wave influences of every panel at the points,
 *   ExtRad  = -i Omega / g (WS N_k - WD Phi_J),
 *   ExtDiff = -i Omega / g (WS dPhi_I/dn + WD Phi_S).
 * Outputs npts x 6 and npts x nb complex row-major (either may be NULL).
 * The reference takes the points from Boundary::Get_FreeSurface_Nodes (the
 * numerical wave tank boundaries); here the caller passes them. */
int bem_fs_elevation(bem_ctx* ctx, bem_body* b, double omega, double kappa, int npts, const double* pts, int nb,
                     const double* betas, const double* phiJ, const double* phiS, double* ExtRad, double* ExtDiff);

/* ---- wave-term surrogate ------------------------------------------------------
 * No counterpart in the reference: an MLP trained on the device from the exact
 * kernel stands in for HavelockGF (Kernels.cpp:342-412) inside
 * Prepare_Linear_System_Wave_Terms (Hydrodynamic_Solver.cpp:177-203) when a
 * solve runs in BEM_MODE_MIXED_SURROGATE. The exact kernel remains the
 * fallback outside the trained domain (K R in [0, 200], K (z + zeta) in
 * [-50, 0]) and the judge of the surrogate's error. Network, features and
 * targets: bemuse_amd/csrc/bem_surrogate.h. */
typedef struct bem_surrogate_config {
    uint64_t seed;       /* sampling and initialisation seed */
    int64_t n_samples;   /* training samples drawn from the exact kernel on the device */
    int batch;           /* samples per Adam step */
    int steps;           /* Adam steps (a new affine shuffle of the pool every pass) */
    double lr;           /* peak learning rate; cosine decay to lr_final */
    double lr_final;
    int64_t n_valid;     /* held-out samples (indices after the training pool) judged exactly */
    int sampling;        /* 0: uniform in (K R, K (z + zeta)); 1: uniform in the log-scaled
                            features log1p(K R), log1p(-K (z + zeta)) (default) */
    int steps_per_launch; /* 0 (default): up to 500 Adam steps per dispatch, one cooperative
                             launch with a grid barrier between gradient and update, when the
                             device holds all of the batch's work-groups at once (else as 1);
                             1: a gradient and an Adam dispatch per step; n > 1: n steps per
                             cooperative launch. The parameters are the same bit for bit. */
} bem_surrogate_config;

typedef struct bem_surrogate_report {
    double loss_initial;    /* mean squared error of the normalised targets, first 10 steps */
    double loss_final;      /* same, last 100 steps */
    double valid_rms[2];    /* validation RMS error per output channel (normalised) */
    double valid_max[2];    /* validation max error per output channel (normalised) */
    double rel_rms;         /* validation relative error of (GF, GFh): RMS over samples */
    double rel_p99;         /* 99th percentile */
    double rel_max;         /* maximum */
    double ms_sample;       /* device time: drawing the pools from the exact kernel */
    double ms_train;        /* device time: all Adam steps */
    double ms_valid;        /* device time: validation forward pass */
} bem_surrogate_report;

/* Defaults: seed 20261203, 20,000,000 samples, batch 8192, 24,000 steps,
 * lr 2e-3 -> 2e-5, 1,000,000 validation samples, log-scaled sampling,
 * steps_per_launch 0. */
int bem_surrogate_default_config(bem_surrogate_config* cfg);
/* Number of network parameters (floats). */
int bem_surrogate_param_count(void);
/* Samples, trains and validates a surrogate on the context's device and
 * keeps it on the context (replacing any earlier one). rep may be NULL. */
int bem_surrogate_train(bem_ctx* ctx, const bem_surrogate_config* cfg, bem_surrogate_report* rep);
/* Copies the context's surrogate parameters out / loads parameters into the
 * context (n = bem_surrogate_param_count()). */
int bem_surrogate_get(bem_ctx* ctx, float* params, int n);
int bem_surrogate_set(bem_ctx* ctx, const float* params, int n);
/* Kernel-level access for checks. Samples `first` .. first+n-1 of the
 * sampler with this seed and sampling: features X (n x 4) and normalised
 * targets T (n x 2). */
int bem_surrogate_sample(bem_ctx* ctx, uint64_t seed, int sampling, int64_t first, int n, float* X, float* T);
/* Network outputs Y (n x 2) for features X (n x 4); params NULL uses the
 * context's. */
int bem_surrogate_forward(bem_ctx* ctx, const float* params, int n, const float* X, float* Y);
/* Loss (sum of squared errors / (2 n)) and its gradient with respect to
 * params (host, kSurParams floats) over the batch (X, T), by the training
 * kernels. */
int bem_surrogate_grad(bem_ctx* ctx, const float* params, int n, const float* X, const float* T, float* grad,
                       double* loss);
/* The context's surrogate at n points (h, v): out n x 4 doubles
 * (Re GF, Im GF, Re GFh, Im GFh); used[i] = 1 where the network served the
 * point, 0 where the exact kernel did (outside the domain). used may be NULL. */
int bem_surrogate_eval(bem_ctx* ctx, int n, const double* h, const double* v, double* out, int* used);

/* Instrumented factorisation for measurement: assembles the system at kappa
 * (compose WAVE) and factors it in the mode's precision (and the context's
 * LU arithmetic) with HIP events around every trailing-update GEMM launch.
 * out (9 doubles): total LU ms, summed GEMM ms (event pair per launch),
 * useful GEMM flops (unpadded order N), GEMM launches, mean GEMM ms per
 * launch (event pairs), mean useful flops per launch, useful GEMM bytes (C
 * read and written, A and B read once, unpadded), mean useful bytes per
 * launch, mean GEMM ms per launch when the same launches are issued again
 * back to back inside one event pair (kernel time without the per-launch
 * event packets; this replay overwrites the factors, which the call
 * discards). */
int bem_profile_lu(bem_ctx* ctx, bem_body* b, int mode, double kappa, double* out);

#ifdef __cplusplus
}
#endif

#endif /* BEMUSE_AMD_H */
