// Launch wrappers of the device kernels (bem_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

namespace bemuse {

constexpr int kPad = 128;       // Np = roundup(N, kPad)
constexpr int kLuNb = 64;       // LU panel width
constexpr int kMaxRhs = 16;     // columns of the right-hand-side block (6 + nb <= 16)
constexpr int kMaxPanels = 1024; // largest padded order / kLuNb (Np <= 65536)

void launch_rankine(hipStream_t s, int N, int Np, const double* body, const double* refl, const double* recv,
                    double* Ssrc, double* Srefl, double* Dsrc, double* Drefl);
// freq: device pointer to (omega, kappa) of the frequency. wg_per_cu > 0
// (default path, far_r2 < 0) launches wg_per_cu x CUs work-groups that take
// the tiles in turn (about that many resident per CU); 0: one work-group per
// tile.
void launch_wave_compose(hipStream_t s, int N, int Np, const double* freq, int compose, const double* quad,
                         const double* recv, const double* Ssrc, const double* Srefl, const double* Dsrc,
                         const double* Drefl, double* Sr, double* Si, double* Dr, double* Di, float* D32r,
                         float* D32i, double far_r2 = -1.0, int wg_per_cu = 0);
void launch_havelock_eval(hipStream_t s, int n, const double* X, const double* Y, const double* V, double* out,
                          bool far32 = false);
// FP32 and FP64 MFMA TFLOP/s and streaming-copy GB/s measured on the current device; 0 on success.
int measure_peaks(hipStream_t s, double* out);
void launch_incident(hipStream_t s, int N, int Np, int nb, const double* body, const double* gen, const double* betas,
                     const double* freq, double* Xr, double* Xi, int ldx, double* PhiIr, double* PhiIi,
                     double* DAr, double* DAi);
int matvec_kslabs(int Np);
// Y = M X (mode 0) or Y = B - M X (mode 1).
template <typename TM>
void launch_matvec(hipStream_t s, int Np, int nr, const TM* Mr, const TM* Mi, const double* Xr, const double* Xi,
                   double* Pr, double* Pi, const double* Br, const double* Bi, int mode, double* Yr, double* Yi,
                   const int* skip = nullptr);
// The partial products of launch_matvec only (and the row sums of |re| +
// |im| per slab into Pabs, nks x Np, when given); the reduction is left to the
// caller.
template <typename TM>
void launch_matvec_partials(hipStream_t s, int Np, int nr, const TM* Mr, const TM* Mi, const double* Xr,
                            const double* Xi, double* Pr, double* Pi, double* Pabs = nullptr,
                            const int* skip = nullptr);
// Optional timing of the trailing-update GEMM launches of one factorisation:
// events bracket every GEMM launch; useful_flops counts 8 M N K over the
// unpadded part (order N) of each launch.
struct LuProfile {
    static constexpr int kMax = 1024;
    hipEvent_t ev[2 * kMax] = {};
    int n = 0;
    int N = 0;
    double useful_flops = 0.0;
    double useful_bytes = 0.0;   // C read and written, A and B read once, over the unpadded part
    // every timed launch, as issued (operand pointers, offsets, tile choice),
    // for replay_lu_gemms
    struct Call {
        int dev, Np, M, N, K, ra, ca, rb, cb, rc, cc, critical, cfg, split;   // split 2: operands split beforehand
        void *Ar, *Ai;
        const float* umax;
        const _Float16 *AS, *BS;
        int nb0, ldn;
    };
    Call call[kMax] = {};
};

// Launches every GEMM the profile recorded again, back to back on stream s
// with one pair of HIP events around the run, and returns the mean ms per
// launch: the kernels' own durations without the event packets between them.
// The launches overwrite the factors (a measurement, not a solve).
template <typename T>
double replay_lu_gemms(const LuProfile& prof, hipStream_t s);

// Work buffers of one factorisation in precision T: saved diagonal block
// (NB x NB), saved A21 (Np x NB, ld Np) and the per-panel inverses of the
// diagonal-block triangles (Np/NB blocks of NB x NB, column-major).
template <typename T>
struct LuWork {
    T *sd_r = nullptr, *sd_i = nullptr, *s21_r = nullptr, *s21_i = nullptr;
    T *Linv_r = nullptr, *Linv_i = nullptr, *Uinv_r = nullptr, *Uinv_i = nullptr;
    T *PL_r = nullptr, *PL_i = nullptr, *PU_r = nullptr, *PU_i = nullptr;   // block-solve coefficients
    float* umax = nullptr;   // per panel: max |Re|, |Im| of its U12 rows (FP32 split-arithmetic updates)
    // FP32 split arithmetic: a panel group's update operands split into fp16
    // once (4 parts x Np x 4 NB halves each); null: split in every tile
    _Float16 *opA = nullptr, *opB = nullptr;
};

// Arithmetic of the FP32 factorisation's trailing updates. The bits above
// kLuArithMask carry a forced tile configuration of the updates on pre-split
// operands: (cfg + 1) << 4, 0 = autotuned.
enum { kLuArithF32 = 0, kLuArithSplitF16 = 1, kLuArithSplitF16Tile = 2 };
constexpr int kLuArithMask = 15, kLuTileShift = 4, kLuTileMask = 255, kLuTileCfgs = 9;
// With this bit in arith the factorisation runs without its pivoting
// fallback: no per-panel pivot-check launch and no interchanges (the panel
// flags still record multipliers with |l| > 1). The caller then re-runs a
// flagged system with the fallback; an unflagged one has the LU of the full
// path bit for bit.
constexpr int kLuSpeculative = 1 << 12;
// Split-arithmetic factorisations of padded order >= kSplitGroup8Np defer
// their trailing updates over groups of eight panels (K = 512: half the
// passes over C, longer thin updates inside the panel kernels); the split
// operand buffers hold kSplitGroupMax panels. Measured: at 20736 panels the
// mixed sweep gains 8% (7.41 against 6.88 solves/s), at 6144 it loses 3%.
constexpr int kSplitGroup8Np = 12288, kSplitGroupMax = 8;

// s2: panel (lookahead) stream; ev: 2 events per panel (Np / NB panels).
// arith (FP32 only): kLuArithF32 (FP32 MFMA) or kLuArithSplitF16 (fp16 MFMA on
// split operands, w.umax required; operands split once per panel group when
// w.opA, w.opB are there) or kLuArithSplitF16Tile (the same, split in every
// tile); a forced tile configuration in the bits above kLuArithMask;
// kLuSpeculative (both precisions).
// latency != 0: the factorisation has the
// device to itself (one frequency in flight), so the panel kernels take the
// variant with the shorter critical path and more registers; the results are
// the same bit for bit.
// pivot_tau (0..1): a panel with a multiplier |l| > pivot_tau takes the
// partial-pivoting fallback (1: exactly where partial pivoting would pick
// another pivot than the diagonal).
template <typename T>
void launch_lu(hipStream_t s, hipStream_t s2, hipEvent_t* ev, int Np, T* Ar, T* Ai, const LuWork<T>& w, int* flag,
               int* ipiv, LuProfile* prof = nullptr, int arith = kLuArithF32, int latency = 1,
               double pivot_tau = 1.0);
// Times the trailing-update tile shapes on the GEMM shapes of a factorisation
// of padded order Np (once per device, precision and Np) using the Np x Np
// scratch matrix Ar/Ai; returns the number of shape classes tuned.
// umax (FP32: Np / 64 floats, overwritten): the panel maxima the split-
// arithmetic tiles read; with it the split tiles are tuned too.
template <typename T>
int tune_lu_gemms(hipStream_t s, int Np, T* Ar, T* Ai, float* umax = nullptr, const _Float16* AS = nullptr,
                  const _Float16* BS = nullptr);
// Tile chosen for an M x N x K update on the current device: 1000 * rows + cols.
template <typename T>
int gemm_tile_choice(int M, int N, int K);
// flag_any: the factorisation's any-interchange flag (the row interchanges
// are applied to X when it is set); nullptr for a speculative factorisation,
// which has none (no launch for them).
// X64r/X64i (FP32 solves, optional): the solution also into FP64 arrays as
// the upper-triangular solve writes it, stored (x64_add = 0) or added to
// them (x64_add = 1).
template <typename T, typename TX>
void launch_lu_solve(hipStream_t s, int Np, int nr, const T* Ar, const T* Ai, const LuWork<T>& w, const int* ipiv,
                     const int* flag_any, TX* Xr, TX* Xi, unsigned long long* counter, unsigned long long* pub,
                     const int* skip = nullptr, double* X64r = nullptr, double* X64i = nullptr, int x64_add = 0);
// Words of the triangular-solve publication buffer (enough for FP64).
inline size_t lu_solve_pub_words(int Np) { return (size_t)(Np / kLuNb) * kMaxRhs * kLuNb * 2 * 2; }
struct RefineCtrl;
// reset (optional): a refinement state to clear on the way.
void launch_to_f32(hipStream_t s, size_t n, const double* a, const double* b, float* fa, float* fb,
                   const int* skip = nullptr, RefineCtrl* reset = nullptr);
// Interleaved complex n_rows x ncol (ld n_rows) -> planar Np x ncolp (ld Np),
// zero beyond the source except a unit diagonal when ident = 1.
void launch_pack_planar(hipStream_t s, int n_rows, int ncol, int Np, int ncolp, const double* src, double* re,
                        double* im, int ident);
// Planar (ld Np) -> interleaved complex n_rows x ncol (ld n_rows).
void launch_unpack_planar(hipStream_t s, int n_rows, int ncol, int Np, const double* re, const double* im,
                          double* dst);
// Triangular-solve state as a new lane starts it: counter at epoch 1,
// nwords publication words at epoch 0.
void launch_solve_state_init(hipStream_t s, unsigned long long* counter, unsigned long long* pub, size_t nwords);

// Device-resident refinement state (one per solve lane); `done` is the first
// member so kernels can take &ctrl->done as their skip flag.
struct RefineCtrl {
    int done, iters, converged, pad;
    double rel, prev;
    // posted by the work-groups of k_refine_reduce, cleared by its last one:
    // per column the maxima |r|, |x|, |Im r|, |Im x| and the matrix norm, as
    // the bit patterns of non-negative doubles; the finished work-groups
    unsigned long long mx[4][kMaxRhs];
    unsigned long long an;
    unsigned count, pad2;
};
// One refinement residual and convergence check after
// launch_matvec_partials(M = A, X): R = B - A X, R32 = fp32(R), and the
// decision in ctrl (done, converged, iters, rel) from the max-norm relative
// residual against tol; with Pabs (the partials' row sums) anorm[0] =
// |A|_inf first, else anorm[0] is read.
void launch_refine_reduce(hipStream_t s, int N, int Np, int nr, const double* Pr, const double* Pi, const double* Br,
                          const double* Bi, double* Rr, double* Ri, float* R32r, float* R32i, const double* Xr,
                          const double* Xi, const double* Pabs, double* anorm, RefineCtrl* ctrl, double tol,
                          int max_iter);
void launch_copy2(hipStream_t s, size_t n, const double* a, const double* b, double* ca, double* cb);
void launch_colmax(hipStream_t s, int N, int Np, int nr, const double* rr, const double* ri, const double* xr,
                   const double* xi, double* out);
// Kochin functions (nkoch x (6 + nb) complex) and exterior-point potentials
// as free-surface elevation (npts x (6 + nb) complex) from the potentials
// phr/phi (N x (6 + nb), planar, ld N); see bem_postproc.hip.
void launch_kochin(hipStream_t s, int N, int nkoch, int nb, const double* body, const double* gen,
                   const double* betas, double omega, double kappa, const double* phr, const double* phi,
                   double* out);
void launch_fs_field(hipStream_t s, int N, int npts, const double* pts, int nb, const double* betas, double omega,
                     double kappa, const double* body, const double* refl, const double* quad, const double* gen,
                     const double* phr, const double* phi, double* out);
void launch_reduce_coeffs(hipStream_t s, int N, int Np, int nb, const double* gen, const double* Xr, const double* Xi,
                          const double* PhiIr, const double* PhiIi, const double* DAr, const double* DAi,
                          double* out);

}  // namespace bemuse
