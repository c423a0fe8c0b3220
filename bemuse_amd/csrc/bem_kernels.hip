// HIP kernels of the per-frequency path for gfx950 (MI355X).
//
// Matrices live in HBM as planar complex (separate re / im arrays),
// column-major, padded to Np = roundup(N, 128) rows and columns. The padding
// block of the system matrix is the identity and of every other operand zero,
// so no GEMM / TRSM kernel needs edge handling and the padded solve equals the
// unpadded one.
#include "bem_kernels.h"
#include "bem_device.hpp"

#include <array>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include <hip/hip_runtime.h>

#include <algorithm>

namespace bemuse {

using namespace dev;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Assembly
// ---------------------------------------------------------------------------

// Tile: 64 receivers (one per lane) x 16 source panels (4 per wave).
constexpr int AS_TR = 64, AS_TS = 16;

// Rankine S/D of every body panel S and its mirror image at every BC point R
// (Prepare_Linear_System, Hydrodynamic_Solver.cpp:83-115). Outputs, column
// major (R, S): Ssrc = -s, Srefl = -sr, Dsrc = d, Drefl = dr; zero in padding.
__global__ __launch_bounds__(256) void k_rankine(int N, int Np, const double* __restrict__ body,
                                                 const double* __restrict__ refl, const double* __restrict__ recv,
                                                 double* Ssrc, double* Srefl, double* Dsrc, double* Drefl) {
    __shared__ double sb[AS_TS][32], sr[AS_TS][32];
    int r0 = blockIdx.x * AS_TR, s0 = blockIdx.y * AS_TS;
    for (int i = threadIdx.x; i < AS_TS * 32; i += 256) {
        int s = s0 + i / 32;
        sb[i / 32][i % 32] = s < N ? body[(size_t)s * 32 + i % 32] : 0.0;
        sr[i / 32][i % 32] = s < N ? refl[(size_t)s * 32 + i % 32] : 0.0;
    }
    __syncthreads();
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int R = r0 + lane;
    double px = 0, py = 0, pz = 0;
    if (R < N) { px = recv[4 * (size_t)R]; py = recv[4 * (size_t)R + 1]; pz = recv[4 * (size_t)R + 2]; }
    for (int c = 0; c < AS_TS / 4; c++) {
        int sl = wave * (AS_TS / 4) + c;
        int S = s0 + sl;
        size_t idx = (size_t)S * Np + R;
        double s = 0, d = 0, srr = 0, drr = 0;
        if (R < N && S < N) {
            rankine_panel(sb[sl], px, py, pz, s, d);
            rankine_panel(sr[sl], px, py, pz, srr, drr);
        }
        Ssrc[idx] = -s; Srefl[idx] = -srr; Dsrc[idx] = d; Drefl[idx] = drr;
    }
}

// Wave-term assembly fused with composition (Prepare_Linear_System_Wave_Terms,
// Hydrodynamic_Solver.cpp:177-203, and compose :414-429). For compose WAVE:
//   S = Ssrc + Srefl - ws
//   D = Dsrc + Drefl + wd + Srefl * 2 n_z(S) * K
// ZERO: S = Ssrc + Srefl, D = Dsrc + Drefl; INF: S = Ssrc - Srefl, D = Dsrc - Drefl.
// Writes planar FP64 S and D, and optionally an FP32 copy of D. Padding: S = 0,
// D = identity.
// freq: device (omega, kappa) of the frequency (read on the device, so a
// captured graph replays with the frequency written before each launch).
// FAR32: the instantiation with the FP32 far-pair branch (far_r2 >= 0). The
// default one leaves it out: the branch's inlined FP32 functions raise the
// kernel from 120 to 173 VGPRs, which halves the waves per SIMD (4 to 2).
// One tile: 64 receivers from r0 x 16 sources from s0.
template <bool FAR32>
__device__ __forceinline__ void wave_compose_tile(int N, int Np, double K, int compose, int r0, int s0,
                                                  double (*sq)[32], const double* __restrict__ quad,
                                                  const double* __restrict__ recv, const double* __restrict__ Ssrc,
                                                  const double* __restrict__ Srefl, const double* __restrict__ Dsrc,
                                                  const double* __restrict__ Drefl, double* Sr, double* Si, double* Dr,
                                                  double* Di, float* D32r, float* D32i, double far_r2) {
    for (int i = threadIdx.x; i < AS_TS * 32; i += 256) {
        int s = s0 + i / 32;
        sq[i / 32][i % 32] = s < N ? quad[(size_t)s * 32 + i % 32] : 0.0;
    }
    __syncthreads();
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int R = r0 + lane;
    double px = 0, py = 0, pz = 0;
    if (R < N) { px = recv[4 * (size_t)R]; py = recv[4 * (size_t)R + 1]; pz = recv[4 * (size_t)R + 2]; }
    for (int c = 0; c < AS_TS / 4; c++) {
        int sl = wave * (AS_TS / 4) + c;
        int S = s0 + sl;
        size_t idx = (size_t)S * Np + R;
        double sre = 0, sim = 0, dre = 0, dim = 0;
        if (R < N && S < N) {
            double ss = Ssrc[idx], sf = Srefl[idx], ds = Dsrc[idx], df = Drefl[idx];
            if (compose == 0) {
                double wsr, wsi, wdr, wdi;
                // far_r2 >= 0: pairs whose first mirrored node lies farther than
                // sqrt(far_r2) from the receiver take the FP32 special functions
                const double ex = px - sq[sl][0], ey = py - sq[sl][1], ez = pz - sq[sl][2];
                if (FAR32 && ex * ex + ey * ey + ez * ez > far_r2)
                    wave_panel<true>(sq[sl], px, py, pz, K, wsr, wsi, wdr, wdi);
                else
                    wave_panel<false>(sq[sl], px, py, pz, K, wsr, wsi, wdr, wdi);
                sre = (ss + sf) - wsr;
                sim = -wsi;
                dre = (ds + df) + (wdr + sf * sq[sl][29] * K);
                dim = wdi;
            } else if (compose == 1) {
                sre = ss + sf; dre = ds + df;
            } else {
                sre = ss - sf; dre = ds - df;
            }
        } else if (R == S) {
            dre = 1.0;
        }
        Sr[idx] = sre; Si[idx] = sim; Dr[idx] = dre; Di[idx] = dim;
        if (D32r) { D32r[idx] = (float)dre; D32i[idx] = (float)dim; }
    }
}

// One work-group per tile (grid Np/64 x Np/16).
template <bool FAR32>
__global__ __launch_bounds__(256) void k_wave_compose(int N, int Np, const double* __restrict__ freq, int compose,
                                                      const double* __restrict__ quad, const double* __restrict__ recv,
                                                      const double* __restrict__ Ssrc, const double* __restrict__ Srefl,
                                                      const double* __restrict__ Dsrc, const double* __restrict__ Drefl,
                                                      double* Sr, double* Si, double* Dr, double* Di,
                                                      float* D32r, float* D32i, double far_r2) {
    __shared__ double sq[AS_TS][32];
    wave_compose_tile<FAR32>(N, Np, freq[1], compose, blockIdx.x * AS_TR, blockIdx.y * AS_TS, sq, quad, recv, Ssrc,
                             Srefl, Dsrc, Drefl, Sr, Si, Dr, Di, D32r, D32i, far_r2);
}

// The same over a 1-D grid smaller than the tile count: work-group g takes
// tiles g, g + grid, ... (tile t at receiver block t % (Np/64)), which bounds
// the work-groups resident at a time. A separate kernel because the loop
// raises the register count (120 to 130 VGPRs: 3 waves per SIMD instead of 4).
__global__ __launch_bounds__(256) void k_wave_compose_strided(
    int N, int Np, const double* __restrict__ freq, int compose, const double* __restrict__ quad,
    const double* __restrict__ recv, const double* __restrict__ Ssrc, const double* __restrict__ Srefl,
    const double* __restrict__ Dsrc, const double* __restrict__ Drefl, double* Sr, double* Si, double* Dr, double* Di,
    float* D32r, float* D32i) {
    __shared__ double sq[AS_TS][32];
    const double K = freq[1];
    const int tiles_r = Np / AS_TR, ntiles = tiles_r * (Np / AS_TS);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        wave_compose_tile<false>(N, Np, K, compose, (tile % tiles_r) * AS_TR, (tile / tiles_r) * AS_TS, sq, quad,
                                 recv, Ssrc, Srefl, Dsrc, Drefl, Sr, Si, Dr, Di, D32r, D32i, -1.0);
        __syncthreads();    // the tile's reads of sq are done before the next tile's loads
    }
}

template <bool S32>
__global__ void k_havelock_eval(int n, const double* X, const double* Y, const double* V, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    GFVal g = havelock_gf_t<S32>(X[i], Y[i], V[i]);
    double cx = 0.0, cy = 0.0;
    if (g.h > 1.0e-6) { double rh = frcp(g.h); cx = X[i] * rh; cy = Y[i] * rh; }
    double* o = out + 8 * (size_t)i;
    o[0] = g.gr; o[1] = g.gi;
    o[2] = g.hr * cx; o[3] = g.hi * cx;
    o[4] = g.hr * cy; o[5] = g.hi * cy;
    o[6] = g.gr; o[7] = g.gi;
}

// Incident potential at centroids (Set_Incident_Potential_Mats /
// Calc_Phi_Incident, Hydrodynamic_Solver.cpp:544-596) and the right-hand-side
// block Xin = [N_k (6 columns, :117-138) | -dPhi_I/dn * area (nb columns, :445)].
// PhiI, DPhiIA (= dPhi_I/dn * area) are N x nb complex planar with ld Np.
__global__ void k_incident(int N, int Np, int nb, const double* __restrict__ body, const double* __restrict__ gen,
                           const double* __restrict__ betas, const double* __restrict__ freq,
                           double* Xr, double* Xi, int ldx, double* PhiIr, double* PhiIi, double* DAr, double* DAi) {
    const double omega = freq[0], kappa = freq[1];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    const double G = 9.81;
    if (i >= N) {
        for (int j = 0; j < ldx; j++) { Xr[(size_t)j * Np + i] = 0; Xi[(size_t)j * Np + i] = 0; }
        for (int b = 0; b < nb; b++) {
            PhiIr[(size_t)b * Np + i] = 0; PhiIi[(size_t)b * Np + i] = 0;
            DAr[(size_t)b * Np + i] = 0; DAi[(size_t)b * Np + i] = 0;
        }
        return;
    }
    const double* p = body + (size_t)i * 32;
    double x = p[0], y = p[1], z = p[2];
    const double* g = gen + (size_t)i * 8;
    double nx = g[0], ny = g[1], nz = g[2], area = g[6];
    for (int j = 0; j < 6; j++) { Xr[(size_t)j * Np + i] = g[j]; Xi[(size_t)j * Np + i] = 0.0; }
    for (int b = 0; b < nb; b++) {
        double CB = cos(betas[b]), SB = sin(betas[b]);
        // Phi = i g/Omega exp(i K (x cos + y sin)) exp(K z)
        double ph = kappa * (x * CB + y * SB);
        double c0 = G / omega;
        double er = cos(ph), ei = sin(ph);
        double Z = exp(kappa * z);
        double pr = -c0 * ei * Z, pi = c0 * er * Z;          // (i c0)(er + i ei) Z
        // dPhi/dn = (n_x i K cos + n_y i K sin + n_z K) Phi
        double ar = nz * kappa, ai = (nx * CB + ny * SB) * kappa;
        double dr = ar * pr - ai * pi, di = ar * pi + ai * pr;
        PhiIr[(size_t)b * Np + i] = pr; PhiIi[(size_t)b * Np + i] = pi;
        DAr[(size_t)b * Np + i] = dr * area; DAi[(size_t)b * Np + i] = di * area;
        Xr[(size_t)(6 + b) * Np + i] = -(dr * area); Xi[(size_t)(6 + b) * Np + i] = -(di * area);
    }
    for (int j = 6 + nb; j < ldx; j++) { Xr[(size_t)j * Np + i] = 0; Xi[(size_t)j * Np + i] = 0; }
}

// ---------------------------------------------------------------------------
// Tall-skinny products Y = M X for an Np x Np complex matrix M and nr <= 16
// columns X, split over K slabs (partials reduced in a fixed order by
// k_matvec_reduce); the kernel, k_matvec_mfma, is defined after the MFMA
// helpers below.
// ---------------------------------------------------------------------------
constexpr int MV_ROWS = 64, MV_MAXR = 16, MV_KC = 64;

// Y = sum_ks P (mode 0) or Y = B - sum_ks P (mode 1).
__global__ void k_matvec_reduce(int Np, int nr, int nks, const double* __restrict__ Pr, const double* __restrict__ Pi,
                                const double* Br, const double* Bi, int mode, double* Yr, double* Yi, const int* skip) {
    if (skip && *skip) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nr * Np) return;
    int j = t / Np, r = t % Np;
    double sr = 0, si = 0;
    for (int ks = 0; ks < nks; ks++) {
        size_t o = ((size_t)ks * MV_MAXR + j) * Np + r;
        sr += Pr[o]; si += Pi[o];
    }
    size_t y = (size_t)j * Np + r;
    if (mode == 1) { sr = Br[y] - sr; si = Bi[y] - si; }
    Yr[y] = sr; Yi[y] = si;
}

// ---------------------------------------------------------------------------
// Blocked LU, right-looking, panel width LU_NB, planar complex.
//
// Pivoting: every panel is first factored with the diagonal as pivot. Partial
// pivoting (max |a_ik|, first maximum) chooses exactly these pivots iff every
// multiplier satisfies |l_ik| <= 1, so the kernels check that bound; a panel
// that violates it is restored from its saved copy and refactored with row
// interchanges (k_lu_panel_gepp), which also applies the interchanges to the
// other columns. Results therefore follow partial pivoting.
//
// Per panel p the factorisation also keeps the inverses of the unit-lower and
// upper triangles of its diagonal block (Linv_p, Uinv_p, NB x NB column-major
// planar): L21 = A21 Uinv_p and U12 = Linv_p A12 are then small MFMA products,
// and the triangular solves multiply by them instead of substituting.
// ---------------------------------------------------------------------------
constexpr int LU_NB = 64;
constexpr int NB2 = LU_NB * LU_NB;
// Panels per deferred trailing update (K = group NB). Four halves the passes
// over C but moves twice the flops into the thin updates of the group's own
// panels; measured on MI355X, that pays in FP64 and for large orders, while
// FP32 at N ~ 6000 is faster with pairs (74.9 vs 71.4 frequency-solves/s).
template <typename T>
inline int lu_group(int Np, bool split = false) {
    if (split && Np >= kSplitGroup8Np) return 8;
    return (sizeof(T) == 8 || split || Np >= 12288) ? 4 : 2;
}

template <typename T>
__device__ inline T cabs2(T r, T i) { return r * r + i * i; }

// Broadcast lane `l` (uniform) of v to the wave.
__device__ inline float bcast(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ inline double bcast(double v, int l) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Inverses of the unit-lower and upper triangles of a factored NB x NB block
// held in LDS as lu[col][row] (rdr/rdi: LDS scratch of NB reciprocal pivots).
// Run by waves 0..3 of the work-group; wave w computes columns 16w .. 16w+15
// one at a time (lane = row) and writes them column-major.
template <typename T>
__device__ void diag_inverses(const T (*lur)[LU_NB + 1], const T (*lui)[LU_NB + 1], T* rdr, T* rdi, T* Li_r,
                              T* Li_i, T* Ui_r, T* Ui_i) {
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x < LU_NB) {
        T pr = lur[threadIdx.x][threadIdx.x], pi = lui[threadIdx.x][threadIdx.x];
        T den = cabs2(pr, pi);
        rdr[threadIdx.x] = pr / den; rdi[threadIdx.x] = -pi / den;
    }
    __syncthreads();
    if (wave >= 4) return;
    for (int cl = 0; cl < 16; cl++) {
        int c = 16 * wave + cl;
        // column c of L^-1: forward substitution on e_c (unit diagonal)
        T xr = (lane == c) ? (T)1 : (T)0, xi = 0;
        for (int k = c; k < LU_NB - 1; k++) {
            T br = bcast(xr, k), bi = bcast(xi, k);
            if (lane > k) {
                T lr = lur[k][lane], li = lui[k][lane];
                xr -= lr * br - li * bi; xi -= lr * bi + li * br;
            }
        }
        Li_r[c * LU_NB + lane] = xr; Li_i[c * LU_NB + lane] = xi;
        // column c of U^-1: back substitution on e_c
        xr = (lane == c) ? (T)1 : (T)0; xi = 0;
        for (int k = c; k >= 0; k--) {
            T b0r = bcast(xr, k), b0i = bcast(xi, k);
            T rr = rdr[k], ri = rdi[k];
            T br = b0r * rr - b0i * ri, bi = b0r * ri + b0i * rr;
            if (lane == k) { xr = br; xi = bi; }
            if (lane < k) {
                T ur = lur[k][lane], ui = lui[k][lane];
                xr -= ur * br - ui * bi; xi -= ur * bi + ui * br;
            }
        }
        Ui_r[c * LU_NB + lane] = xr; Ui_i[c * LU_NB + lane] = xi;
    }
}

// ---------------------------------------------------------------------------
// MFMA 16x16x4 (lane l = i + 16 g): A supplies A[i][k(g)], B supplies
// B[k(g)][j]. Accumulator register r of lane l holds D[row][j] with
// row = 4g + r for the f32 instruction and row = g + 4r for the f64
// instruction on gfx950 (measured). Within a chunk of 16 k, lane group g
// covers k = 4g + t at step t, so each lane reads 4 consecutive k of a B
// column as one vector.
// ---------------------------------------------------------------------------
template <typename T> struct MfmaT;
template <> struct MfmaT<double> {
    typedef d4 acc;
    static __device__ inline acc mma(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ inline int row(int g, int r) { return g + 4 * r; }
    static constexpr int WM = 2, WN = 2;   // wave tile 32 x 32
};
template <> struct MfmaT<float> {
    typedef f4 acc;
    static __device__ inline acc mma(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ inline int row(int g, int r) { return 4 * g + r; }
    static constexpr int WM = 4, WN = 4;   // wave tile 64 x 64
};

// Complex 16x16 block product step over 16 k: acc += A(16 x 16k) B(16k x 16)
// with A element (i, k) at a[k * lda + i] and B element (k, j) at b[j * ldb + k].
template <typename T>
__device__ inline void cmma16(const T* ar_, const T* ai_, size_t lda, const T* br_, const T* bi_, size_t ldb,
                              typename MfmaT<T>::acc& cr, typename MfmaT<T>::acc& ci, int i, int g) {
    typedef MfmaT<T> MT;
    T br[4], bi[4];
#pragma unroll
    for (int t = 0; t < 4; t++) { br[t] = br_[(size_t)i * ldb + 4 * g + t]; bi[t] = bi_[(size_t)i * ldb + 4 * g + t]; }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        T a_r = ar_[(size_t)(4 * g + t) * lda + i], a_i = ai_[(size_t)(4 * g + t) * lda + i];
        cr = MT::mma(a_r, br[t], cr);
        cr = MT::mma(-a_i, bi[t], cr);
        ci = MT::mma(a_r, bi[t], ci);
        ci = MT::mma(a_i, br[t], ci);
    }
}

// The same product over 64 k (four steps of 16): all operand loads of the
// four steps are issued before the first MFMA, so the chain waits for one
// memory round trip instead of four (the panel kernels are latency-bound).
// B's four k of a lane (4 g .. 4 g + 3 of a step) are one vector load:
// callers keep b and ldb multiples of 4 elements.
template <typename T>
struct Ops64 {
    typedef T v4 __attribute__((ext_vector_type(4)));
    T a_r[4][4], a_i[4][4];
    v4 b_r[4], b_i[4];
};

template <typename T>
__device__ inline void load64(Ops64<T>& o, const T* ar_, const T* ai_, size_t lda, const T* br_, const T* bi_,
                              size_t ldb, int i, int g) {
    typedef typename Ops64<T>::v4 v4;
#pragma unroll
    for (int st = 0; st < 4; st++) {
        const size_t ob = (size_t)i * ldb + 16 * st + 4 * g;
        o.b_r[st] = *reinterpret_cast<const v4*>(br_ + ob);
        o.b_i[st] = *reinterpret_cast<const v4*>(bi_ + ob);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int k = 16 * st + 4 * g + t;
            o.a_r[st][t] = ar_[(size_t)k * lda + i]; o.a_i[st][t] = ai_[(size_t)k * lda + i];
        }
    }
}

template <typename T>
__device__ inline void mma64(const Ops64<T>& o, typename MfmaT<T>::acc& cr, typename MfmaT<T>::acc& ci) {
    typedef MfmaT<T> MT;
#pragma unroll
    for (int st = 0; st < 4; st++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            cr = MT::mma(o.a_r[st][t], o.b_r[st][t], cr);
            cr = MT::mma(-o.a_i[st][t], o.b_i[st][t], cr);
            ci = MT::mma(o.a_r[st][t], o.b_i[st][t], ci);
            ci = MT::mma(o.a_i[st][t], o.b_r[st][t], ci);
        }
}

template <typename T>
__device__ inline void cmma64(const T* ar_, const T* ai_, size_t lda, const T* br_, const T* bi_, size_t ldb,
                              typename MfmaT<T>::acc& cr, typename MfmaT<T>::acc& ci, int i, int g) {
    Ops64<T> o;
    load64<T>(o, ar_, ai_, lda, br_, bi_, ldb, i, g);
    mma64<T>(o, cr, ci);
}

// acc += A B over K (a multiple of 64) in chunks of 64; A advances by 64
// columns (64 lda), B by 64 rows. PIPE (FP32 only): the next chunk's loads
// are in flight while the current one feeds the MFMAs, two register stages
// (FP64 would need twice the registers per stage). The sums are the same
// either way, in the same order; PIPE shortens a lone factorisation's
// critical path, but its 256 VGPRs per wave crowd the other lanes' work
// when several frequencies share the device.
template <typename T, bool PIPE>
__device__ inline void cmma_k(const T* ar_, const T* ai_, size_t lda, const T* br_, const T* bi_, size_t ldb, int K,
                              typename MfmaT<T>::acc& cr, typename MfmaT<T>::acc& ci, int i, int g) {
    const int nch = K / LU_NB;
    if constexpr (PIPE && sizeof(T) == 4) {
        Ops64<T> x, y;
        if (nch > 0) load64<T>(x, ar_, ai_, lda, br_, bi_, ldb, i, g);
        for (int c = 0; c < nch; c += 2) {
            if (c + 1 < nch)
                load64<T>(y, ar_ + (size_t)(c + 1) * LU_NB * lda, ai_ + (size_t)(c + 1) * LU_NB * lda, lda,
                          br_ + (c + 1) * LU_NB, bi_ + (c + 1) * LU_NB, ldb, i, g);
            mma64<T>(x, cr, ci);
            if (c + 1 < nch) {
                if (c + 2 < nch)
                    load64<T>(x, ar_ + (size_t)(c + 2) * LU_NB * lda, ai_ + (size_t)(c + 2) * LU_NB * lda, lda,
                              br_ + (c + 2) * LU_NB, bi_ + (c + 2) * LU_NB, ldb, i, g);
                mma64<T>(y, cr, ci);
            }
        }
    } else {
        for (int c = 0; c < nch; c++)
            cmma64<T>(ar_ + (size_t)c * LU_NB * lda, ai_ + (size_t)c * LU_NB * lda, lda, br_ + c * LU_NB,
                      bi_ + c * LU_NB, ldb, cr, ci, i, g);
    }
}

// ---------------------------------------------------------------------------
// Factor the NB x NB diagonal block at (k0, k0) without interchanges and form
// the inverses of its unit-lower and upper triangles, blocked by DB = 16 so
// that most of the work is 16 x 16 MFMA block products and the work-group
// meets at 18 barriers (the unblocked form needs one per column, 128 in all).
// The block lives in LDS, [col][row]. For sub-block column kb = 0..3:
//   wave 0 factors the 16 x 16 diagonal block in registers (lane = row, the
//   pivot row broadcast with readlane); then waves 0 and 1 form Li_kk and
//   Ui_kk side by side from the factor in LDS (lane = column of the inverse);
//   the waves share L_ik = A_ik Ui_kk (i > kb), U_kj = Li_kk A_kj (j > kb),
//   then the trailing blocks A_ij -= L_ik U_kj.
// Then wave j forms block column j of both inverses, keeping its blocks in
// MFMA accumulators: X_ij = -Li_ii sum_{k=j}^{i-1} L_ik X_kj (L^-1) and
// Y_ij = -Ui_ii sum_{k=i+1}^{j} U_ik Y_kj (U^-1). Saves the block (save_r
// non-null), writes L\U, L^-1 and U^-1 (NB x NB column-major), and sets
// flag[panel] bit 1 if a multiplier exceeds the bound in modulus (tau2 its
// square; 1|2 on a zero pivot).
// ---------------------------------------------------------------------------
constexpr int DB = 16;
constexpr int DLD = LU_NB + 4;       // LDS column stride of the block

template <typename T>
struct DiagLds {
    T ar[LU_NB][DLD], ai[LU_NB][DLD];               // the block, [col][row]
    T lr[4][DB][DB + 1], li[4][DB][DB + 1];         // Li_kk, [col][row]
    T ur[4][DB][DB + 1], ui[4][DB][DB + 1];         // Ui_kk
};

// acc (cr, ci) += A B for 16 x 16 blocks, A (i, k) at a[k lda + i], B (k, j) at
// b[j ldb + k]; lane l = i + 16 g; the result (row MT::row(g, r), column i) is
// in register r.
template <typename T>
__device__ inline void blk_mma(const T* ar_, const T* ai_, int lda, const T* br_, const T* bi_, int ldb,
                               typename MfmaT<T>::acc& cr, typename MfmaT<T>::acc& ci, int i, int g) {
    typedef MfmaT<T> MT;
    T br[4], bi[4], a_r[4], a_i[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        br[t] = br_[i * ldb + 4 * g + t]; bi[t] = bi_[i * ldb + 4 * g + t];
        a_r[t] = ar_[(4 * g + t) * lda + i]; a_i[t] = ai_[(4 * g + t) * lda + i];
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        cr = MT::mma(a_r[t], br[t], cr);
        ci = MT::mma(a_r[t], bi[t], ci);
        cr = MT::mma(-a_i[t], bi[t], cr);
        ci = MT::mma(a_i[t], br[t], ci);
    }
}

// acc (cr, ci) += A B with B a 16 x 16 block held in accumulator layout (lane
// (i, g), register t: B(MT::row(g, t), i)). Step t pairs lane group g with
// k = MT::row(g, t), so B is read straight from the registers.
template <typename T>
__device__ inline void blk_mma_rb(const T* ar_, const T* ai_, int lda, const typename MfmaT<T>::acc& Br,
                                  const typename MfmaT<T>::acc& Bi, typename MfmaT<T>::acc& cr,
                                  typename MfmaT<T>::acc& ci, int i, int g) {
    typedef MfmaT<T> MT;
    T a_r[4], a_i[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        int k = MT::row(g, t);
        a_r[t] = ar_[k * lda + i]; a_i[t] = ai_[k * lda + i];
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        cr = MT::mma(a_r[t], Br[t], cr);
        ci = MT::mma(a_r[t], Bi[t], ci);
        cr = MT::mma(-a_i[t], Bi[t], cr);
        ci = MT::mma(a_i[t], Br[t], ci);
    }
}

// 1 / x for a pivot's squared modulus: the hardware reciprocal refined by
// Newton steps (one in FP32, two in FP64), about 1 ulp, in place of the IEEE
// division sequence on each pivot step's critical path. x = 0 gives inf and
// x = NaN gives NaN, as the division would (the pivot checks flag both).
__device__ inline float rcp_pivot(float x) {
    float r = __builtin_amdgcn_rcpf(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
}
__device__ inline double rcp_pivot(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// Wave 0: the 16 x 16 diagonal block at (o, o) factored in registers (lane
// rho & 15 holds row rho, the pivot row broadcast with readlane), written
// back to LDS.
// |pivot|^2 usable: positive and finite (zero, NaN and inf are not).
template <typename T>
__device__ inline bool pivot_ok(T den) { return den > (T)0 && den < (T)INFINITY; }

template <typename T>
__device__ inline void factor16(DiagLds<T>& S, int o, int lane, int& mybad, T (&ar)[DB], T (&ai)[DB], T tau2) {
    const int rho = lane & 15;
    const bool writer = lane < DB;
#pragma unroll
    for (int c = 0; c < DB; c++) { ar[c] = S.ar[o + c][o + rho]; ai[c] = S.ai[o + c][o + rho]; }
#pragma unroll
    for (int j = 0; j < DB; j++) {
        T pr = bcast(ar[j], j), pi = bcast(ai[j], j);
        T den = cabs2(pr, pi);
        // a zero or non-finite pivot: the pivoting fallback decides (it flags
        // the panel numerically singular if no usable pivot exists)
        if (!pivot_ok(den)) mybad |= 1;
        T inv = rcp_pivot(den);
        T qr = pr * inv, qi = -pi * inv;
        const bool below = rho > j;
        T mr = ar[j] * qr - ai[j] * qi, mi = ar[j] * qi + ai[j] * qr;
        if (below && !(cabs2(mr, mi) <= tau2)) mybad |= 1;   // also NaN
        mr = below ? mr : (T)0;
        mi = below ? mi : (T)0;
        ar[j] = below ? mr : ar[j];
        ai[j] = below ? mi : ai[j];
#pragma unroll
        for (int c = j + 1; c < DB; c++) {
            T ur = bcast(ar[c], j), ui = bcast(ai[c], j);
            ar[c] -= mr * ur - mi * ui;
            ai[c] -= mr * ui + mi * ur;
        }
    }
    if (writer) {
#pragma unroll
        for (int c = 0; c < DB; c++) { S.ar[o + c][o + rho] = ar[c]; S.ai[o + c][o + rho] = ai[c]; }
    }
}

// The unit-lower inverse X = L^-1 of the factored 16 x 16 block, row by row
// in the registers factor16 left (lane rho: row rho of the factor, L(rho, k)
// for k < rho): for k = 0 .. 14 row k of X is final, is broadcast (readlane),
// and every row below takes X(r, c) -= L(r, k) X(k, c), c <= k. Writes
// Li_kk to S.lr[kb] ([col][row]).
template <typename T>
__device__ inline void linv16_rows(DiagLds<T>& S, int kb, int lane, const T (&ar)[DB], const T (&ai)[DB]) {
    const int rho = lane & 15;
    T lr[DB], li[DB], xr[DB], xi[DB];
#pragma unroll
    for (int k = 0; k < DB; k++) {
        lr[k] = rho > k ? ar[k] : (T)0;
        li[k] = rho > k ? ai[k] : (T)0;
        xr[k] = k == rho ? (T)1 : (T)0;
        xi[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < DB - 1; k++) {
#pragma unroll
        for (int c = 0; c <= k; c++) {
            T br = bcast(xr[c], k), bi = bcast(xi[c], k);
            xr[c] -= lr[k] * br - li[k] * bi;
            xi[c] -= lr[k] * bi + li[k] * br;
        }
    }
    if (lane < DB) {
#pragma unroll
        for (int c = 0; c < DB; c++) { S.lr[kb][c][rho] = xr[c]; S.li[kb][c][rho] = xi[c]; }
    }
}

// The upper inverse Y = U^-1 of the factored 16 x 16 block, row by row (lane
// rho: row rho, read from LDS): with D = diag U and N = D^-1 U - I, U^-1 =
// (I + N)^-1 D^-1. Lane rho scales its row by 1 / U(rho, rho); then for
// k = 15 .. 1 row k of W = (I + N)^-1 is final, is broadcast, and every row
// above takes W(r, c) -= N(r, k) W(k, c), c >= k; last, column c of W is
// scaled by 1 / U(c, c). Writes Ui_kk to S.ur[kb] ([col][row]).
template <typename T>
__device__ inline void uinv16_rows(DiagLds<T>& S, int o, int kb, int lane) {
    const int rho = lane & 15;
    T ur[DB], ui[DB];
#pragma unroll
    for (int c = 0; c < DB; c++) { ur[c] = S.ar[o + c][o + rho]; ui[c] = S.ai[o + c][o + rho]; }
    T dr = 0, di = 0;
#pragma unroll
    for (int c = 0; c < DB; c++) {
        dr = c == rho ? ur[c] : dr;
        di = c == rho ? ui[c] : di;
    }
    const T den = rcp_pivot(cabs2(dr, di));
    const T qr = dr * den, qi = -di * den;           // 1 / U(rho, rho)
    T nr_[DB], ni_[DB], wr[DB], wi[DB];
#pragma unroll
    for (int k = 0; k < DB; k++) {
        T a = ur[k] * qr - ui[k] * qi, b = ur[k] * qi + ui[k] * qr;
        nr_[k] = k > rho ? a : (T)0;
        ni_[k] = k > rho ? b : (T)0;
        wr[k] = k == rho ? (T)1 : (T)0;
        wi[k] = 0;
    }
#pragma unroll
    for (int k = DB - 1; k >= 1; k--) {
#pragma unroll
        for (int c = k; c < DB; c++) {
            T br = bcast(wr[c], k), bi = bcast(wi[c], k);
            wr[c] -= nr_[k] * br - ni_[k] * bi;
            wi[c] -= nr_[k] * bi + ni_[k] * br;
        }
    }
#pragma unroll
    for (int c = 0; c < DB; c++) {
        T sr = bcast(qr, c), si = bcast(qi, c);      // 1 / U(c, c)
        T a = wr[c] * sr - wi[c] * si, b = wr[c] * si + wi[c] * sr;
        wr[c] = a;
        wi[c] = b;
    }
    if (lane < DB) {
#pragma unroll
        for (int c = 0; c < DB; c++) { S.ur[kb][c][rho] = wr[c]; S.ui[kb][c][rho] = wi[c]; }
    }
}

// Column (lane & 15) of the unit-lower inverse of the factored diagonal
// block by forward substitution, L(r, k) read from LDS as broadcasts (the
// factor is final here, so register broadcasts would all be hoisted and
// spill).
template <typename T>
__device__ inline void linv16(DiagLds<T>& S, int o, int kb, int lane) {
    const int col = lane & 15;
    T xr[DB], xi[DB];
#pragma unroll
    for (int r = 0; r < DB; r++) { xr[r] = (r == col) ? (T)1 : (T)0; xi[r] = 0; }
#pragma unroll
    for (int k = 0; k < DB - 1; k++) {
#pragma unroll
        for (int r = k + 1; r < DB; r++) {
            T lr = S.ar[o + k][o + r], li = S.ai[o + k][o + r];
            xr[r] -= lr * xr[k] - li * xi[k];
            xi[r] -= lr * xi[k] + li * xr[k];
        }
    }
    if (lane < DB) {
#pragma unroll
        for (int r = 0; r < DB; r++) { S.lr[kb][col][r] = xr[r]; S.li[kb][col][r] = xi[r]; }
    }
}

// Column (lane & 15) of the upper inverse by back substitution, U(r, k) read
// from LDS as broadcasts; runs on a second wave beside linv16.
template <typename T>
__device__ inline void uinv16(DiagLds<T>& S, int o, int kb, int lane) {
    const int col = lane & 15;
    T xr[DB], xi[DB];
#pragma unroll
    for (int r = 0; r < DB; r++) { xr[r] = (r == col) ? (T)1 : (T)0; xi[r] = 0; }
#pragma unroll
    for (int k = DB - 1; k >= 0; k--) {
        T dr = S.ar[o + k][o + k], di = S.ai[o + k][o + k];
        T inv = (T)1 / cabs2(dr, di);
        T qr = dr * inv, qi = -di * inv;
        T yr = xr[k] * qr - xi[k] * qi, yi = xr[k] * qi + xi[k] * qr;
        xr[k] = yr; xi[k] = yi;
#pragma unroll
        for (int r = 0; r < k; r++) {
            T u_r = S.ar[o + k][o + r], u_i = S.ai[o + k][o + r];
            xr[r] -= u_r * yr - u_i * yi;
            xi[r] -= u_r * yi + u_i * yr;
        }
    }
    if (lane < DB) {
#pragma unroll
        for (int r = 0; r < DB; r++) { S.ur[kb][col][r] = xr[r]; S.ui[kb][col][r] = xi[r]; }
    }
}

template <typename T>
__device__ inline void load_blk(const T (*br)[DB + 1], const T (*bi)[DB + 1], typename MfmaT<T>::acc& r_,
                                typename MfmaT<T>::acc& i_, int i, int g) {
#pragma unroll
    for (int r = 0; r < 4; r++) { r_[r] = br[i][MfmaT<T>::row(g, r)]; i_[r] = bi[i][MfmaT<T>::row(g, r)]; }
}

template <typename T>
__device__ inline void store_blk(T* pr, T* pi, const typename MfmaT<T>::acc& r_, const typename MfmaT<T>::acc& i_,
                                 int i, int g) {
#pragma unroll
    for (int r = 0; r < 4; r++) { pr[i * LU_NB + MfmaT<T>::row(g, r)] = r_[r]; pi[i * LU_NB + MfmaT<T>::row(g, r)] = i_[r]; }
}

// Wave J: block column J of L^-1 (blocks J..3) and of U^-1 (blocks 0..J).
template <typename T, int J>
__device__ inline void inv_column(DiagLds<T>& S, T* Li_r, T* Li_i, T* Ui_r, T* Ui_i, int i, int g) {
    typedef typename MfmaT<T>::acc acc_t;
    const acc_t z = acc_t{0, 0, 0, 0};
    acc_t Xr[4], Xi[4];
    load_blk<T>(S.lr[J], S.li[J], Xr[J], Xi[J], i, g);
#pragma unroll
    for (int ib = J + 1; ib < 4; ib++) {
        acc_t tr = z, ti = z;
#pragma unroll
        for (int k = J; k < ib; k++) blk_mma_rb<T>(&S.ar[DB * k][DB * ib], &S.ai[DB * k][DB * ib], DLD, Xr[k], Xi[k], tr, ti, i, g);
        acc_t xr = z, xi = z;
        blk_mma_rb<T>(&S.lr[ib][0][0], &S.li[ib][0][0], DB + 1, tr, ti, xr, xi, i, g);
        Xr[ib] = -xr; Xi[ib] = -xi;
    }
#pragma unroll
    for (int ib = J; ib < 4; ib++)
        store_blk<T>(Li_r + DB * J * LU_NB + DB * ib, Li_i + DB * J * LU_NB + DB * ib, Xr[ib], Xi[ib], i, g);
    acc_t Yr[4], Yi[4];
    load_blk<T>(S.ur[J], S.ui[J], Yr[J], Yi[J], i, g);
#pragma unroll
    for (int ib = J - 1; ib >= 0; ib--) {
        acc_t tr = z, ti = z;
#pragma unroll
        for (int k = ib + 1; k <= J; k++) blk_mma_rb<T>(&S.ar[DB * k][DB * ib], &S.ai[DB * k][DB * ib], DLD, Yr[k], Yi[k], tr, ti, i, g);
        acc_t yr = z, yi = z;
        blk_mma_rb<T>(&S.ur[ib][0][0], &S.ui[ib][0][0], DB + 1, tr, ti, yr, yi, i, g);
        Yr[ib] = -yr; Yi[ib] = -yi;
    }
#pragma unroll
    for (int ib = 0; ib <= J; ib++)
        store_blk<T>(Ui_r + DB * J * LU_NB + DB * ib, Ui_i + DB * J * LU_NB + DB * ib, Yr[ib], Yi[ib], i, g);
}

template <typename T>
__global__ __launch_bounds__(256) void k_lu_diag(int Np, int k0, T* Ar, T* Ai, T* save_r, T* save_i, T* Li_r,
                                                 T* Li_i, T* Ui_r, T* Ui_i, int* flag, T tau2) {
    // on the factorisation's critical path: issue ahead of other lanes' GEMM waves on the CU
    __builtin_amdgcn_s_setprio(3);
    typedef MfmaT<T> MT;
    typedef typename MT::acc acc_t;
    __shared__ DiagLds<T> S;
    __shared__ int bad;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, g = lane >> 4;       // MFMA lane coordinates
    if (tid == 0) bad = 0;
    {
        // all loads first: the save stores below may alias them as far as
        // the compiler knows, so interleaving would make every load wait for
        // the previous store (one memory round trip per element pair)
        constexpr int PER = LU_NB * LU_NB / 256;
        T vr[PER], vi[PER];
#pragma unroll
        for (int n = 0; n < PER; n++) {
            int t = tid + 256 * n, c = t / LU_NB, r = t % LU_NB;
            size_t gi = (size_t)(k0 + c) * Np + k0 + r;
            vr[n] = Ar[gi]; vi[n] = Ai[gi];
        }
#pragma unroll
        for (int n = 0; n < PER; n++) {
            int t = tid + 256 * n, c = t / LU_NB, r = t % LU_NB;
            S.ar[c][r] = vr[n]; S.ai[c][r] = vi[n];
            if (save_r) { save_r[t] = vr[n]; save_i[t] = vi[n]; }
        }
    }
    __syncthreads();
    int mybad = 0;
    for (int kb = 0; kb < 4; kb++) {
        const int o = DB * kb;
        T far_[DB], fai_[DB];                        // wave 0: its rows of the factored block
        if (wave == 0) factor16<T>(S, o, lane, mybad, far_, fai_, tau2);
        __syncthreads();
        // the two triangle inverses side by side, each in one wave's registers
        if (wave == 0) linv16_rows<T>(S, kb, lane, far_, fai_);
        else if (wave == 1) uinv16_rows<T>(S, o, kb, lane);
        __syncthreads();
        // L_ik = A_ik Ui_kk (i > kb), U_kj = Li_kk A_kj (j > kb), one block per wave in turn
        const int nrest = 3 - kb;
        for (int job = wave; job < 2 * nrest; job += 4) {
            acc_t cr = acc_t{0, 0, 0, 0}, ci = acc_t{0, 0, 0, 0};
            int b = kb + 1 + (job % nrest);
            if (job < nrest) {
                blk_mma<T>(&S.ar[o][DB * b], &S.ai[o][DB * b], DLD, &S.ur[kb][0][0], &S.ui[kb][0][0], DB + 1, cr, ci,
                           i, g);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    int row = MT::row(g, r);
                    if (!(cabs2(cr[r], ci[r]) <= tau2)) mybad |= 1;
                    S.ar[o + i][DB * b + row] = cr[r]; S.ai[o + i][DB * b + row] = ci[r];
                }
            } else {
                blk_mma<T>(&S.lr[kb][0][0], &S.li[kb][0][0], DB + 1, &S.ar[DB * b][o], &S.ai[DB * b][o], DLD, cr, ci,
                           i, g);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    int row = MT::row(g, r);
                    S.ar[DB * b + i][o + row] = cr[r]; S.ai[DB * b + i][o + row] = ci[r];
                }
            }
        }
        __syncthreads();
        // trailing blocks A_ij -= L_ik U_kj
        for (int job = wave; job < nrest * nrest; job += 4) {
            int bi_ = kb + 1 + job / nrest, bj = kb + 1 + job % nrest;
            acc_t cr = acc_t{0, 0, 0, 0}, ci = acc_t{0, 0, 0, 0};
            blk_mma<T>(&S.ar[o][DB * bi_], &S.ai[o][DB * bi_], DLD, &S.ar[DB * bj][o], &S.ai[DB * bj][o], DLD, cr, ci,
                       i, g);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int row = MT::row(g, r);
                S.ar[DB * bj + i][DB * bi_ + row] -= cr[r];
                S.ai[DB * bj + i][DB * bi_ + row] -= ci[r];
            }
        }
        __syncthreads();
    }
    if (mybad) atomicOr(&bad, mybad);
    // L\U out, and the zero triangles of the inverses (strictly upper blocks
    // of L^-1, strictly lower of U^-1); the other blocks come from inv_column
    for (int t = tid; t < LU_NB * LU_NB; t += 256) {
        int c = t / LU_NB, r = t % LU_NB;
        size_t gi = (size_t)(k0 + c) * Np + k0 + r;
        Ar[gi] = S.ar[c][r]; Ai[gi] = S.ai[c][r];
        int bc = c / DB, br = r / DB;
        if (br < bc) { Li_r[t] = 0; Li_i[t] = 0; }
        if (br > bc) { Ui_r[t] = 0; Ui_i[t] = 0; }
    }
    switch (wave) {
        case 0: inv_column<T, 0>(S, Li_r, Li_i, Ui_r, Ui_i, i, g); break;
        case 1: inv_column<T, 1>(S, Li_r, Li_i, Ui_r, Ui_i, i, g); break;
        case 2: inv_column<T, 2>(S, Li_r, Li_i, Ui_r, Ui_i, i, g); break;
        default: inv_column<T, 3>(S, Li_r, Li_i, Ui_r, Ui_i, i, g); break;
    }
    __syncthreads();
    if (tid == 0 && bad) atomicOr(flag + k0 / LU_NB, bad);
}

// L21 = A21 Uinv_p in place for rows k0+NB .. Np-1. A work-group owns 16
// rows and wave b the 16 columns 16 b .. 16 b + 15 of them (a quarter of the
// MFMA chain each, four times the work-groups of one wave per 16 rows); all
// waves read all NB columns of the rows, so they meet before overwriting.
// Saves A21 (rows relative to k0+NB, ld Np) for the pivoting fallback
// (save_r non-null) and flags multipliers with |l| above the bound (tau2 its
// square).
// kp < k0: the panel is not the first of its group (whose first column is
// kp), and its columns still lack the group's deferred update. The
// work-group applies it to its own rows first,
//   A(rows, k0..k0+NB) -= L(rows, kp..k0) U(kp..k0, k0..k0+NB),
// keeps the updated rows in LDS, and saves and factors those.
template <typename T, bool PIPE>
__global__ __launch_bounds__(256) void k_lu_l21(int Np, int k0, int kp, T* Ar, T* Ai, const T* Ui_r, const T* Ui_i,
                                                T* save_r, T* save_i, int* flag, T tau2) {
    // on the factorisation's critical path: issue ahead of other lanes' GEMM waves on the CU
    __builtin_amdgcn_s_setprio(3);
    typedef MfmaT<T> MT;
    typedef typename MT::acc acc_t;
    __shared__ T wr[LU_NB][16 + 1], wi[LU_NB][16 + 1];   // updated rows, [col][row]
    int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int rowl = blockIdx.x * 16;                          // relative to k0 + NB
    int row0 = k0 + LU_NB + rowl;
    int i = lane & 15, g = lane >> 4, b = wave;
    acc_t cr = acc_t{0, 0, 0, 0}, ci = acc_t{0, 0, 0, 0};
    if (kp < k0) {
        acc_t ur = acc_t{0, 0, 0, 0}, ui = acc_t{0, 0, 0, 0};
        cmma_k<T, PIPE>(Ar + (size_t)kp * Np + row0, Ai + (size_t)kp * Np + row0, Np, Ar + (size_t)(k0 + 16 * b) * Np + kp,
                  Ai + (size_t)(k0 + 16 * b) * Np + kp, Np, k0 - kp, ur, ui, i, g);
        const int col = 16 * b + i;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int row = MT::row(g, r);
            size_t o = (size_t)(k0 + col) * Np + row0 + row;
            T vr = Ar[o] - ur[r], vi = Ai[o] - ui[r];
            wr[col][row] = vr; wi[col][row] = vi;
            if (save_r) {
                save_r[(size_t)col * Np + rowl + row] = vr;
                save_i[(size_t)col * Np + rowl + row] = vi;
            }
        }
        __syncthreads();
        cmma64<T>(&wr[0][0], &wi[0][0], 17, Ui_r + (16 * b) * LU_NB, Ui_i + (16 * b) * LU_NB, LU_NB, cr, ci, i, g);
    } else {
        {
            constexpr int PER = 16 * LU_NB / 256;        // loads first (see k_lu_diag)
            T vr[PER], vi[PER];
#pragma unroll
            for (int n = 0; n < PER; n++) {
                int t = tid + 256 * n, r = t % 16, c = t / 16;
                size_t o = (size_t)(k0 + c) * Np + row0 + r;
                vr[n] = Ar[o]; vi[n] = Ai[o];
            }
            if (save_r) {
#pragma unroll
                for (int n = 0; n < PER; n++) {
                    int t = tid + 256 * n, r = t % 16, c = t / 16;
                    save_r[(size_t)c * Np + rowl + r] = vr[n];
                    save_i[(size_t)c * Np + rowl + r] = vi[n];
                }
            }
        }
        cmma64<T>(Ar + (size_t)k0 * Np + row0, Ai + (size_t)k0 * Np + row0, Np, Ui_r + (16 * b) * LU_NB,
                  Ui_i + (16 * b) * LU_NB, LU_NB, cr, ci, i, g);
        __syncthreads();                                 // every wave has read the rows
    }
    int mybad = 0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        size_t o = (size_t)(k0 + 16 * b + i) * Np + row0 + MT::row(g, r);
        T lr = cr[r], li = ci[r];
        Ar[o] = lr; Ai[o] = li;
        if (!(cabs2(lr, li) <= tau2)) mybad = 1;
    }
    if (mybad) atomicOr(flag + k0 / LU_NB, 1);
}

// U12 = Linv_p A12 in place for columns k0+NB .. Np-1. A work-group owns 16
// columns and wave a rows 16 a .. 16 a + 15 of them; all waves read all NB
// rows of the columns, so they meet before overwriting.
// kp < k0: the panel is not the first of its group (first column kp); the
// work-group first applies the group's deferred update to its columns of the
// panel's rows, A(k0..k0+NB, cols) -= L(k0..k0+NB, kp..k0) U(kp..k0, cols),
// in LDS.
// next != 0: the next panel belongs to the same group. The four work-groups
// whose columns are that panel's then apply the group's update, this panel's
// included, to its diagonal block,
//   A(k0+NB..k0+2NB, cols) -= L(k0+NB..k0+2NB, kp..k0+NB) U(kp..k0+NB, cols),
// with this panel's U12 rows from LDS, so that the diagonal-block kernel of
// the next panel can start at once.
template <typename T, bool PIPE>
__global__ __launch_bounds__(256) void k_lu_u12(int Np, int k0, int kp, int next, T* Ar, T* Ai, const T* Li_r,
                                                const T* Li_i, float* umax) {
    // on the factorisation's critical path: issue ahead of other lanes' GEMM waves on the CU
    __builtin_amdgcn_s_setprio(3);
    typedef MfmaT<T> MT;
    typedef typename MT::acc acc_t;
    constexpr int SLD = LU_NB + 4;
    __shared__ __align__(32) T sr[16][SLD], si[16][SLD]; // the columns' NB rows, [col][row] (vector reads)
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int col0 = k0 + LU_NB + blockIdx.x * 16;
    int i = lane & 15, g = lane >> 4, a = wave;
    acc_t cr = acc_t{0, 0, 0, 0}, ci = acc_t{0, 0, 0, 0};
    if (kp < k0) {
        acc_t ur = acc_t{0, 0, 0, 0}, ui = acc_t{0, 0, 0, 0};
        cmma_k<T, PIPE>(Ar + (size_t)kp * Np + k0 + 16 * a, Ai + (size_t)kp * Np + k0 + 16 * a, Np, Ar + (size_t)col0 * Np + kp,
                  Ai + (size_t)col0 * Np + kp, Np, k0 - kp, ur, ui, i, g);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int row = 16 * a + MT::row(g, r);
            size_t o = (size_t)(col0 + i) * Np + k0 + row;
            sr[i][row] = Ar[o] - ur[r]; si[i][row] = Ai[o] - ui[r];
        }
        __syncthreads();
        cmma64<T>(Li_r + 16 * a, Li_i + 16 * a, LU_NB, &sr[0][0], &si[0][0], SLD, cr, ci, i, g);
    } else {
        cmma64<T>(Li_r + 16 * a, Li_i + 16 * a, LU_NB, Ar + (size_t)col0 * Np + k0, Ai + (size_t)col0 * Np + k0, Np,
                  cr, ci, i, g);
    }
    __syncthreads();                                     // every wave has read the columns
    float m = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        size_t o = (size_t)(col0 + i) * Np + k0 + 16 * a + MT::row(g, r);
        Ar[o] = cr[r]; Ai[o] = ci[r];
        m = fmaxf(m, fmaxf(fabsf((float)cr[r]), fabsf((float)ci[r])));
    }
    if (next && blockIdx.x < LU_NB / 16) {
        // the next panel's diagonal block, these 16 of its columns
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int row = 16 * a + MT::row(g, r);
            sr[i][row] = cr[r]; si[i][row] = ci[r];
        }
        __syncthreads();
        const int rn = k0 + LU_NB + 16 * a;
        acc_t dr = acc_t{0, 0, 0, 0}, di = acc_t{0, 0, 0, 0};
        cmma_k<T, PIPE>(Ar + (size_t)kp * Np + rn, Ai + (size_t)kp * Np + rn, Np, Ar + (size_t)col0 * Np + kp,
                  Ai + (size_t)col0 * Np + kp, Np, k0 - kp, dr, di, i, g);
        cmma64<T>(Ar + (size_t)k0 * Np + rn, Ai + (size_t)k0 * Np + rn, Np, &sr[0][0], &si[0][0], SLD, dr, di, i, g);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            size_t o = (size_t)(col0 + i) * Np + rn + MT::row(g, r);
            Ar[o] -= dr[r]; Ai[o] -= di[r];
        }
    }
    if (umax) {   // the panel's max |Re|, |Im| of U12 (non-negative floats order as their bits)
        // one atomic per work-group: every wave's atomic on the one address
        // serialised and doubled the kernel's duration
        __shared__ float wmax[4];
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        if (lane == 0) wmax[wave] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
            atomicMax(reinterpret_cast<unsigned*>(umax + k0 / LU_NB), __float_as_uint(m));
        }
    }
}

// ---------------------------------------------------------------------------
// Wide variants of the L21 / U12 panel kernels for factorisations that share
// the device (several frequencies in flight). The narrow kernels give every
// 16 rows (columns) a work-group of their own, which reloads the operands all
// of them share (Uinv or Linv, and the group's deferred-update block): up to
// 160 KB of loads for a 16 x 64 result, and a work-group on every CU for the
// kernel's duration. Here a work-group owns 64 rows (columns), and each
// fragment of a shared operand that a wave loads feeds four 16 x 16 MFMA
// blocks: a quarter of the work-groups and of the shared-operand loads. Each
// block takes the products of the narrow kernels in the same order, so the
// results are equal bit for bit.
// ---------------------------------------------------------------------------
constexpr int WIDE = 4;   // 16-row (column) blocks per work-group
// Padded orders from which the shared-device factorisation takes the wide
// variants. Measured (mixed sweeps): 20736 panels 7.41 -> 7.92 solves/s, the
// 6018-panel semi-sub unchanged (103.5 / 103.2), the 64-body batch of orders
// 500-4200 3.5% slower.
constexpr int kWidePanelNp = 8192;

// acc[q] += A_q B_q over K (a multiple of 64) for WIDE block pairs, where
// either A is shared and B_q is column block q of B (ROWS = false: B_q at
// b + 16 q ldb) or B is shared and A_q is row block q of A (ROWS = true: A_q
// at a + 16 q); element layouts as in cmma16. Per 64-k chunk and block, the
// MFMA sequence of mma64.
template <typename T, bool ROWS>
__device__ inline void cmma_k_wide(const T* ar_, const T* ai_, size_t lda, const T* br_, const T* bi_, size_t ldb,
                                   int K, typename MfmaT<T>::acc (&cr)[WIDE], typename MfmaT<T>::acc (&ci)[WIDE],
                                   int i, int g) {
    typedef MfmaT<T> MT;
    typedef typename Ops64<T>::v4 v4;
    constexpr int NA = ROWS ? WIDE : 1, NB_ = ROWS ? 1 : WIDE;
    for (int c = 0; c < K / LU_NB; c++) {
        T a_r[NA][4][4], a_i[NA][4][4];
        v4 b_r[NB_][4], b_i[NB_][4];
#pragma unroll
        for (int q = 0; q < NB_; q++)
#pragma unroll
            for (int st = 0; st < 4; st++) {
                const size_t ob = (size_t)(16 * q + i) * ldb + (size_t)c * LU_NB + 16 * st + 4 * g;
                b_r[q][st] = *reinterpret_cast<const v4*>(br_ + ob);
                b_i[q][st] = *reinterpret_cast<const v4*>(bi_ + ob);
            }
#pragma unroll
        for (int q = 0; q < NA; q++)
#pragma unroll
            for (int st = 0; st < 4; st++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const size_t k = (size_t)c * LU_NB + 16 * st + 4 * g + t;
                    a_r[q][st][t] = ar_[k * lda + 16 * q + i];
                    a_i[q][st][t] = ai_[k * lda + 16 * q + i];
                }
#pragma unroll
        for (int q = 0; q < WIDE; q++) {
            const int qa = ROWS ? q : 0, qb = ROWS ? 0 : q;
#pragma unroll
            for (int st = 0; st < 4; st++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    cr[q] = MT::mma(a_r[qa][st][t], b_r[qb][st][t], cr[q]);
                    cr[q] = MT::mma(-a_i[qa][st][t], b_i[qb][st][t], cr[q]);
                    ci[q] = MT::mma(a_r[qa][st][t], b_i[qb][st][t], ci[q]);
                    ci[q] = MT::mma(a_i[qa][st][t], b_r[qb][st][t], ci[q]);
                }
        }
    }
}

// k_lu_l21 for 64 rows per work-group: wave b owns columns 16 b .. 16 b + 15
// of all four 16-row blocks.
template <typename T>
__global__ __launch_bounds__(256) void k_lu_l21_wide(int Np, int k0, int kp, T* Ar, T* Ai, const T* Ui_r,
                                                     const T* Ui_i, T* save_r, T* save_i, int* flag, T tau2) {
    __builtin_amdgcn_s_setprio(3);
    typedef MfmaT<T> MT;
    typedef typename MT::acc acc_t;
    constexpr int RW = 16 * WIDE, WLD = RW + 1;
    __shared__ T wr[LU_NB][WLD], wi[LU_NB][WLD];         // updated rows, [col][row]
    const int tid = threadIdx.x, lane = tid & 63, b = tid >> 6;
    const int rowl = blockIdx.x * RW;                    // relative to k0 + NB
    const int row0 = k0 + LU_NB + rowl;
    const int i = lane & 15, g = lane >> 4;
    acc_t cr[WIDE], ci[WIDE];
#pragma unroll
    for (int q = 0; q < WIDE; q++) { cr[q] = acc_t{0, 0, 0, 0}; ci[q] = acc_t{0, 0, 0, 0}; }
    if (kp < k0) {
        acc_t ur[WIDE], ui[WIDE];
#pragma unroll
        for (int q = 0; q < WIDE; q++) { ur[q] = acc_t{0, 0, 0, 0}; ui[q] = acc_t{0, 0, 0, 0}; }
        cmma_k_wide<T, true>(Ar + (size_t)kp * Np + row0, Ai + (size_t)kp * Np + row0, Np,
                             Ar + (size_t)(k0 + 16 * b) * Np + kp, Ai + (size_t)(k0 + 16 * b) * Np + kp, Np, k0 - kp,
                             ur, ui, i, g);
        const int col = 16 * b + i;
#pragma unroll
        for (int q = 0; q < WIDE; q++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * q + MT::row(g, r);
                const size_t o = (size_t)(k0 + col) * Np + row0 + row;
                const T vr = Ar[o] - ur[q][r], vi = Ai[o] - ui[q][r];
                wr[col][row] = vr; wi[col][row] = vi;
                if (save_r) {
                    save_r[(size_t)col * Np + rowl + row] = vr;
                    save_i[(size_t)col * Np + rowl + row] = vi;
                }
            }
        __syncthreads();
        cmma_k_wide<T, true>(&wr[0][0], &wi[0][0], WLD, Ui_r + (16 * b) * LU_NB, Ui_i + (16 * b) * LU_NB, LU_NB,
                             LU_NB, cr, ci, i, g);
    } else {
        {
            constexpr int PER = RW * LU_NB / 256;        // loads first (see k_lu_diag)
            T vr[PER], vi[PER];
#pragma unroll
            for (int n = 0; n < PER; n++) {
                const int t = tid + 256 * n, r = t % RW, c = t / RW;
                const size_t o = (size_t)(k0 + c) * Np + row0 + r;
                vr[n] = Ar[o]; vi[n] = Ai[o];
            }
            if (save_r) {
#pragma unroll
                for (int n = 0; n < PER; n++) {
                    const int t = tid + 256 * n, r = t % RW, c = t / RW;
                    save_r[(size_t)c * Np + rowl + r] = vr[n];
                    save_i[(size_t)c * Np + rowl + r] = vi[n];
                }
            }
        }
        cmma_k_wide<T, true>(Ar + (size_t)k0 * Np + row0, Ai + (size_t)k0 * Np + row0, Np, Ui_r + (16 * b) * LU_NB,
                             Ui_i + (16 * b) * LU_NB, LU_NB, LU_NB, cr, ci, i, g);
        __syncthreads();                                 // every wave has read the rows
    }
    int mybad = 0;
#pragma unroll
    for (int q = 0; q < WIDE; q++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const size_t o = (size_t)(k0 + 16 * b + i) * Np + row0 + 16 * q + MT::row(g, r);
            const T lr = cr[q][r], li = ci[q][r];
            Ar[o] = lr; Ai[o] = li;
            if (!(cabs2(lr, li) <= tau2)) mybad = 1;
        }
    if (mybad) atomicOr(flag + k0 / LU_NB, 1);
}

// k_lu_u12 for 64 columns per work-group: wave a owns rows 16 a .. 16 a + 15
// of all four 16-column blocks; with `next`, work-group 0 (the next panel's
// columns) applies the group's update to that panel's diagonal block.
template <typename T>
__global__ __launch_bounds__(256) void k_lu_u12_wide(int Np, int k0, int kp, int next, T* Ar, T* Ai, const T* Li_r,
                                                     const T* Li_i, float* umax) {
    __builtin_amdgcn_s_setprio(3);
    typedef MfmaT<T> MT;
    typedef typename MT::acc acc_t;
    constexpr int CW = 16 * WIDE, SLD = LU_NB + 4;
    __shared__ __align__(32) T sr[CW][SLD], si[CW][SLD]; // the columns' NB rows, [col][row]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col0 = k0 + LU_NB + blockIdx.x * CW;
    const int i = lane & 15, g = lane >> 4, a = wave;
    acc_t cr[WIDE], ci[WIDE];
#pragma unroll
    for (int q = 0; q < WIDE; q++) { cr[q] = acc_t{0, 0, 0, 0}; ci[q] = acc_t{0, 0, 0, 0}; }
    if (kp < k0) {
        acc_t ur[WIDE], ui[WIDE];
#pragma unroll
        for (int q = 0; q < WIDE; q++) { ur[q] = acc_t{0, 0, 0, 0}; ui[q] = acc_t{0, 0, 0, 0}; }
        cmma_k_wide<T, false>(Ar + (size_t)kp * Np + k0 + 16 * a, Ai + (size_t)kp * Np + k0 + 16 * a, Np,
                              Ar + (size_t)col0 * Np + kp, Ai + (size_t)col0 * Np + kp, Np, k0 - kp, ur, ui, i, g);
#pragma unroll
        for (int q = 0; q < WIDE; q++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * a + MT::row(g, r);
                const size_t o = (size_t)(col0 + 16 * q + i) * Np + k0 + row;
                sr[16 * q + i][row] = Ar[o] - ur[q][r];
                si[16 * q + i][row] = Ai[o] - ui[q][r];
            }
        __syncthreads();
        cmma_k_wide<T, false>(Li_r + 16 * a, Li_i + 16 * a, LU_NB, &sr[0][0], &si[0][0], SLD, LU_NB, cr, ci, i, g);
    } else {
        cmma_k_wide<T, false>(Li_r + 16 * a, Li_i + 16 * a, LU_NB, Ar + (size_t)col0 * Np + k0,
                              Ai + (size_t)col0 * Np + k0, Np, LU_NB, cr, ci, i, g);
    }
    __syncthreads();                                     // every wave has read the columns
    float m = 0.0f;
#pragma unroll
    for (int q = 0; q < WIDE; q++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const size_t o = (size_t)(col0 + 16 * q + i) * Np + k0 + 16 * a + MT::row(g, r);
            Ar[o] = cr[q][r]; Ai[o] = ci[q][r];
            m = fmaxf(m, fmaxf(fabsf((float)cr[q][r]), fabsf((float)ci[q][r])));
        }
    if (next && blockIdx.x == 0) {
        // the next panel's diagonal block: these 64 columns
#pragma unroll
        for (int q = 0; q < WIDE; q++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * a + MT::row(g, r);
                sr[16 * q + i][row] = cr[q][r]; si[16 * q + i][row] = ci[q][r];
            }
        __syncthreads();
        const int rn = k0 + LU_NB + 16 * a;
        acc_t dr[WIDE], di[WIDE];
#pragma unroll
        for (int q = 0; q < WIDE; q++) { dr[q] = acc_t{0, 0, 0, 0}; di[q] = acc_t{0, 0, 0, 0}; }
        cmma_k_wide<T, false>(Ar + (size_t)kp * Np + rn, Ai + (size_t)kp * Np + rn, Np, Ar + (size_t)col0 * Np + kp,
                              Ai + (size_t)col0 * Np + kp, Np, k0 - kp, dr, di, i, g);
        cmma_k_wide<T, false>(Ar + (size_t)k0 * Np + rn, Ai + (size_t)k0 * Np + rn, Np, &sr[0][0], &si[0][0], SLD,
                              LU_NB, dr, di, i, g);
#pragma unroll
        for (int q = 0; q < WIDE; q++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const size_t o = (size_t)(col0 + 16 * q + i) * Np + rn + MT::row(g, r);
                Ar[o] -= dr[q][r]; Ai[o] -= di[q][r];
            }
    }
    if (umax) {   // one atomic per work-group, as in k_lu_u12
        __shared__ float wmax[4];
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        if (lane == 0) wmax[wave] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
            atomicMax(reinterpret_cast<unsigned*>(umax + k0 / LU_NB), __float_as_uint(m));
        }
    }
}

// Fallback for a rejected panel: restore it and factor it with partial
// pivoting in one work-group, streaming through global memory; then rebuild
// the triangle inverses of its diagonal block and apply the interchanges to
// all other columns. ipiv receives absolute row indices. Runs only if
// flag[panel] is set (rare), so one launch per panel serves both steps.
template <typename T>
__global__ __launch_bounds__(1024) void k_lu_panel_gepp(int Np, int k0, T* Ar, T* Ai, const T* sd_r, const T* sd_i,
                                                       const T* s21_r, const T* s21_i, T* Li_r, T* Li_i, T* Ui_r,
                                                       T* Ui_i, int* flag, int* ipiv) {
    // on the factorisation's critical path: issue ahead of other lanes' GEMM waves on the CU
    __builtin_amdgcn_s_setprio(3);
    int f = flag[k0 / LU_NB];
    if (!(f & 1)) return;
    int tid = threadIdx.x;
    for (int t = tid; t < NB2; t += 1024) {
        int c = t / LU_NB, r = t % LU_NB;
        size_t g = (size_t)(k0 + c) * Np + k0 + r;
        Ar[g] = sd_r[(size_t)c * LU_NB + r]; Ai[g] = sd_i[(size_t)c * LU_NB + r];
    }
    int M2 = Np - k0 - LU_NB;
    for (int c = 0; c < LU_NB; c++)
        for (int r = tid; r < M2; r += 1024) {
            size_t g = (size_t)(k0 + c) * Np + k0 + LU_NB + r;
            Ar[g] = s21_r[(size_t)c * Np + r]; Ai[g] = s21_i[(size_t)c * Np + r];
        }
    __syncthreads();
    __shared__ T bv[1024];
    __shared__ int bi[1024];
    __shared__ int piv_s;
    __shared__ T prow_r[LU_NB], prow_i[LU_NB];
    __shared__ T lur[LU_NB][LU_NB + 1], lui[LU_NB][LU_NB + 1];
    __shared__ T rdr[LU_NB], rdi[LU_NB];
    for (int j = 0; j < LU_NB; j++) {
        int col = k0 + j;
        T best = (T)-1; int brow = Np;
        for (int r = k0 + j + tid; r < Np; r += 1024) {
            size_t g = (size_t)col * Np + r;
            T v = sqrt(cabs2(Ar[g], Ai[g]));
            if (v > best) { best = v; brow = r; }
        }
        bv[tid] = best; bi[tid] = brow;
        __syncthreads();
        for (int s = 512; s > 0; s >>= 1) {
            if (tid < s) {
                T o = bv[tid + s]; int oi = bi[tid + s];
                if (o > bv[tid] || (o == bv[tid] && oi < bi[tid])) { bv[tid] = o; bi[tid] = oi; }
            }
            __syncthreads();
        }
        // no finite candidate (every |a| NaN): keep the diagonal, flagged below
        if (tid == 0) { int p0 = bi[0] < Np ? bi[0] : col; piv_s = p0; ipiv[col] = p0; }
        __syncthreads();
        int p = piv_s;
        if (p != col && tid < LU_NB) {
            size_t a = (size_t)(k0 + tid) * Np;
            T tr = Ar[a + col], ti = Ai[a + col];
            Ar[a + col] = Ar[a + p]; Ai[a + col] = Ai[a + p];
            Ar[a + p] = tr; Ai[a + p] = ti;
        }
        __syncthreads();
        if (tid < LU_NB) { size_t a = (size_t)(k0 + tid) * Np + col; prow_r[tid] = Ar[a]; prow_i[tid] = Ai[a]; }
        __syncthreads();
        T pr = prow_r[j], pi = prow_i[j];
        T den = cabs2(pr, pi);
        if (!pivot_ok(den)) { if (tid == 0) flag[k0 / LU_NB] |= 2; continue; }   // numerically singular
        for (int r = col + 1 + tid; r < Np; r += 1024) {
            size_t g = (size_t)col * Np + r;
            T ar = Ar[g], ai = Ai[g];
            T mr = (ar * pr + ai * pi) / den, mi = (ai * pr - ar * pi) / den;
            Ar[g] = mr; Ai[g] = mi;
            for (int c = j + 1; c < LU_NB; c++) {
                size_t h = (size_t)(k0 + c) * Np + r;
                T u_r = prow_r[c], u_i = prow_i[c];
                Ar[h] -= mr * u_r - mi * u_i;
                Ai[h] -= mr * u_i + mi * u_r;
            }
        }
        __syncthreads();
    }
    for (int t = tid; t < NB2; t += 1024) {
        int c = t / LU_NB, r = t % LU_NB;
        size_t g = (size_t)(k0 + c) * Np + k0 + r;
        lur[c][r] = Ar[g]; lui[c][r] = Ai[g];
    }
    __syncthreads();
    diag_inverses<T>(lur, lui, rdr, rdi, Li_r, Li_i, Ui_r, Ui_i);
    // the panel's interchanges on every column outside it, in pivot order
    // (one column per thread; ipiv was written before the barriers above)
    for (int c = tid; c < Np; c += 1024) {
        if (c >= k0 && c < k0 + LU_NB) continue;
        size_t a = (size_t)c * Np;
        for (int j = k0; j < k0 + LU_NB; j++) {
            int p = ipiv[j];
            if (p != j) {
                T tr = Ar[a + j], ti = Ai[a + j];
                Ar[a + j] = Ar[a + p]; Ai[a + j] = Ai[a + p];
                Ar[a + p] = tr; Ai[a + p] = ti;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Complex GEMM on MFMA: C -= A B (LU trailing update) for planar complex
// operands in one column-major array with leading dimension ld.
// A: M x K at (ra, ca), B: K x N at (rb, cb), C: M x N at (rc, cc); M, N
// multiples of 64, K multiple of 16.
//
// Chunks of 16 k of the A and B tiles are staged in LDS (double buffered:
// the global loads of chunk c+1 are in flight while chunk c feeds the MFMAs)
// with 16-byte coalesced loads; both are stored k-contiguous so every MFMA
// operand quad (4 consecutive k) is one LDS read. 32 KB (FP32) / 32 KB (FP64)
// of LDS per stage lets two work-groups share a CU.
//
// FP32 (the factorisation that FP64 refinement corrects) takes three real
// products per k instead of four (Gauss; Higham, "Stability of a method for
// multiplying complex matrices with three real matrix multiplications",
// SIAM J. Matrix Anal. Appl. 13, 1992): with T1 = Ar Br, T2 = Ai Bi,
// T3 = (Ar + Ai)(Br + Bi),  Re(A B) = T1 - T2,  Im(A B) = T3 - T1 - T2.
// It is normwise as accurate as the four-product form but not componentwise:
// the error of Im(A B) scales with |Ar| |Br|, and at low frequency the
// system's imaginary part (the wave term, O(K)) is orders of magnitude
// below its real part, which is where the damping comes from. FP64, the
// unrefined factorisation, therefore keeps the four-product form.
// A is staged negated and the accumulators start from zero; C is read into
// registers during the first chunk and added at the end:
//   three products: P = -T1,  Q = T2,  R = -T3, and
//                   Cr - Re(A B) = Cr + (Q + P),  Ci - Im(A B) = Ci + ((R - P) + Q);
//   four products:  Q = -Re(A B),  R = -Im(A B),  C - A B = (Cr + Q, Ci + R).
// ---------------------------------------------------------------------------
constexpr int GK = 16;           // k per LDS chunk
constexpr int GKP = GK + 4;      // padded row length in LDS (bank spread)

template <typename T> struct Vec4;
template <> struct Vec4<float> { typedef float4 type; };
template <> struct Vec4<double> { typedef double4 type; };

// WM x WN MFMA tiles of 16 x 16 per wave; the four waves form a 2 x 2 grid, so
// the work-group tile is (32 WM) x (32 WN). Every element accumulates its k in
// the same order for any (WM, WN), so all tile shapes give identical results.
template <typename T, int WM, int WN>
__global__ __launch_bounds__(256, 2) void k_cgemm_sub(int M, int N, int K, T* __restrict__ Pr, T* __restrict__ Pi,
                                                      int ld, int ra, int ca, int rb, int cb, int rc, int cc,
                                                      int critical) {
    // the factorisation's thin updates (panel columns and rows) gate the
    // next panel: their waves issue ahead of the trailing updates on a CU
    if (critical) __builtin_amdgcn_s_setprio(3);
    typedef MfmaT<T> MT;
    typedef typename MT::acc acc_t;
    typedef typename Vec4<T>::type v4;
    constexpr int TM = 32 * WM, TN = 32 * WN;      // work-group tile
    static_assert(TM >= 64 && TN >= 64, "a chunk must give every thread at least one 4-element load");
    constexpr int NLA = TM * GK / 4 / 256;          // 4-element A loads per thread per chunk (re or im)
    constexpr int NLB = TN * GK / 4 / 256;
    // A is kept k-major as in global memory (one vector store per load,
    // conflict-free); rows padded by 4 so the four lane groups of an MFMA
    // operand read hit different banks. B is kept column-major.
    constexpr int TMP = TM + 4;
    __shared__ T As[2][2][GK][TMP];                 // [buf][re/im][k][row]
    __shared__ T Bs[2][2][TN][GKP];                 // [buf][re/im][col][k]
    int tiles_m = (M + TM - 1) / TM, tiles_n = (N + TN - 1) / TN;
    int ntiles = tiles_m * tiles_n;
    // XCD-aware order: work-groups are dealt round-robin to the 8 XCDs, so
    // each XCD gets a contiguous run of a grouped tile order: blocks of
    // GROUP_M tile rows, walked column by column, so that an XCD's resident
    // tiles share GROUP_M A slices and each B column slice in its L2 (in
    // row-major order every tile row re-read the whole B panel).
    constexpr int GROUP_M = 8;
    int bid = blockIdx.x;
    int per = (ntiles + 7) / 8;
    int tile = (bid % 8) * per + bid / 8;
    if (tile >= ntiles) return;
    int group = tile / (GROUP_M * tiles_n), first_m = group * GROUP_M;
    int gsize = min(GROUP_M, tiles_m - first_m), in_group = tile - group * GROUP_M * tiles_n;
    int tm = first_m + in_group % gsize, tn = in_group / gsize;
    int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int wm = wave >> 1, wn = wave & 1;
    int rowt = tm * TM, colt = tn * TN;
    int mrows = min(TM, M - rowt), ncols = min(TN, N - colt);
    const size_t ldl = (size_t)ld;

    v4 ga[2][NLA], gb[2][NLB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int l = 0; l < NLA; l++) {
            int idx = tid + 256 * l;
            int kk = idx / (TM / 4), rq = idx % (TM / 4);
            int r = min(4 * rq, mrows - 4);
            size_t o = (size_t)(ca + k0 + kk) * ldl + ra + rowt + r;
            ga[0][l] = *reinterpret_cast<const v4*>(Pr + o);
            ga[1][l] = *reinterpret_cast<const v4*>(Pi + o);
        }
#pragma unroll
        for (int l = 0; l < NLB; l++) {
            int idx = tid + 256 * l;
            int c = idx / (GK / 4), kq = idx % (GK / 4);
            int cl = min(c, ncols - 1);
            size_t o = (size_t)(cb + colt + cl) * ldl + rb + k0 + 4 * kq;
            gb[0][l] = *reinterpret_cast<const v4*>(Pr + o);
            gb[1][l] = *reinterpret_cast<const v4*>(Pi + o);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int l = 0; l < NLA; l++) {
            int idx = tid + 256 * l;
            int kk = idx / (TM / 4), rq = idx % (TM / 4);
#pragma unroll
            for (int p = 0; p < 2; p++) {
                v4 v = ga[p][l];
                v.x = -v.x; v.y = -v.y; v.z = -v.z; v.w = -v.w;
                *reinterpret_cast<v4*>(&As[buf][p][kk][4 * rq]) = v;
            }
        }
#pragma unroll
        for (int l = 0; l < NLB; l++) {
            int idx = tid + 256 * l;
            int c = idx / (GK / 4), kq = idx % (GK / 4);
#pragma unroll
            for (int p = 0; p < 2; p++) {
                Bs[buf][p][c][4 * kq + 0] = gb[p][l].x;
                Bs[buf][p][c][4 * kq + 1] = gb[p][l].y;
                Bs[buf][p][c][4 * kq + 2] = gb[p][l].z;
                Bs[buf][p][c][4 * kq + 3] = gb[p][l].w;
            }
        }
    };

    int i = lane & 15, g = lane >> 4;
    int row0 = wm * 16 * WM, col0 = wn * 16 * WN;           // within the tile
    bool active = (row0 < mrows) && (col0 < ncols);
    constexpr bool kGauss = sizeof(T) == 4;
    acc_t p1[WM][WN], q2[WM][WN], r3[WM][WN];
    // C is added after the last product, so its loads (issued behind those
    // of the second chunk) overlap the first chunks instead of holding up
    // the first MFMA; the work-groups of one dispatch run in step, and C
    // read up front made every CU wait on the same burst.
    acc_t c_re[WM][WN], c_im[WM][WN];
    int nchunk = K / GK;
    gload(0);
#pragma unroll
    for (int a = 0; a < WM; a++)
#pragma unroll
        for (int b = 0; b < WN; b++) {
            p1[a][b] = acc_t{0, 0, 0, 0};
            q2[a][b] = acc_t{0, 0, 0, 0};
            r3[a][b] = acc_t{0, 0, 0, 0};
        }
    sstore(0);
    __syncthreads();
    for (int c = 0; c < nchunk; c++) {
        int buf = c & 1;
        if (c + 1 < nchunk) gload((c + 1) * GK);
        if (c == 0 && active) {
#pragma unroll
            for (int a = 0; a < WM; a++)
#pragma unroll
                for (int b = 0; b < WN; b++) {
                    size_t off = (size_t)(cc + colt + col0 + 16 * b + i) * ldl + rc + rowt + row0 + 16 * a;
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        c_re[a][b][r] = Pr[off + MT::row(g, r)];
                        c_im[a][b][r] = Pi[off + MT::row(g, r)];
                    }
                }
        }
        if (active) {
            T ar[WM][4], ai[WM][4], br[WN][4], bi[WN][4];
#pragma unroll
            for (int a = 0; a < WM; a++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    ar[a][t] = As[buf][0][4 * g + t][row0 + 16 * a + i];
                    ai[a][t] = As[buf][1][4 * g + t][row0 + 16 * a + i];
                }
#pragma unroll
            for (int b = 0; b < WN; b++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    br[b][t] = Bs[buf][0][col0 + 16 * b + i][4 * g + t];
                    bi[b][t] = Bs[buf][1][col0 + 16 * b + i][4 * g + t];
                }
            // three (four) passes over the blocks per k step, so consecutive
            // MFMAs write different accumulators (each element accumulates
            // its products in the same order for every tile shape)
#pragma unroll
            for (int t = 0; t < 4 && !kGauss; t++) {
#pragma unroll
                for (int a = 0; a < WM; a++)
#pragma unroll
                    for (int b = 0; b < WN; b++) q2[a][b] = MT::mma(ar[a][t], br[b][t], q2[a][b]);
#pragma unroll
                for (int a = 0; a < WM; a++)
#pragma unroll
                    for (int b = 0; b < WN; b++) r3[a][b] = MT::mma(ar[a][t], bi[b][t], r3[a][b]);
#pragma unroll
                for (int a = 0; a < WM; a++)
#pragma unroll
                    for (int b = 0; b < WN; b++) q2[a][b] = MT::mma(-ai[a][t], bi[b][t], q2[a][b]);
#pragma unroll
                for (int a = 0; a < WM; a++)
#pragma unroll
                    for (int b = 0; b < WN; b++) r3[a][b] = MT::mma(ai[a][t], br[b][t], r3[a][b]);
            }
#pragma unroll
            for (int t = 0; t < 4 && kGauss; t++) {
                T sa[WM], sb[WN];
#pragma unroll
                for (int a = 0; a < WM; a++) sa[a] = ar[a][t] + ai[a][t];   // -(Ar + Ai)
#pragma unroll
                for (int b = 0; b < WN; b++) sb[b] = br[b][t] + bi[b][t];
#pragma unroll
                for (int a = 0; a < WM; a++)
#pragma unroll
                    for (int b = 0; b < WN; b++) p1[a][b] = MT::mma(ar[a][t], br[b][t], p1[a][b]);
#pragma unroll
                for (int a = 0; a < WM; a++)
#pragma unroll
                    for (int b = 0; b < WN; b++) q2[a][b] = MT::mma(-ai[a][t], bi[b][t], q2[a][b]);
#pragma unroll
                for (int a = 0; a < WM; a++)
#pragma unroll
                    for (int b = 0; b < WN; b++) r3[a][b] = MT::mma(sa[a], sb[b], r3[a][b]);
            }
        }
        if (c + 1 < nchunk) {
            sstore(buf ^ 1);
            __syncthreads();
        }
    }
    if (!active) return;
#pragma unroll
    for (int a = 0; a < WM; a++)
#pragma unroll
        for (int b = 0; b < WN; b++) {
            size_t off = (size_t)(cc + colt + col0 + 16 * b + i) * ldl + rc + rowt + row0 + 16 * a;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int rr = MT::row(g, r);
                if (kGauss) {
                    Pr[off + rr] = c_re[a][b][r] + (q2[a][b][r] + p1[a][b][r]);
                    Pi[off + rr] = c_im[a][b][r] + ((r3[a][b][r] - p1[a][b][r]) + q2[a][b][r]);
                } else {
                    Pr[off + rr] = c_re[a][b][r] + q2[a][b][r];
                    Pi[off + rr] = c_im[a][b][r] + r3[a][b][r];
                }
            }
        }
}

// ---------------------------------------------------------------------------
// The same update for FP32 operands on fp16 MFMA (v_mfma_f32_32x32x16_f16,
// FP32 accumulation): every FP32 operand x is scaled by a power of two and
// split into two fp16 terms, x s = h + l with h = fp16(x s) and
// l = fp16(x s - h), and each real product takes three MFMAs,
//   a b s_a s_b ~ h_a h_b + h_a l_b + l_a h_b,
// dropping l_a l_b. The pair carries 22 significant bits, so a product is
// good to about 2^-22 of |a| |b| (FP32 FMA: 2^-24); the accumulation is FP32
// as in the FP32 kernel. This is the split-operand emulation of single
// precision on half-precision matrix units (Markidis et al., IPDPSW 2018;
// Ootomo and Yokota, IJHPCA 2022), here with both scales chosen so that no
// term overflows fp16 and small entries stay well above its subnormal range
// relative to the largest: A holds LU multipliers (|l| <= 1 under partial
// pivoting), scaled by 2^12; B holds rows of U, scaled by 2^(10 - e) with
// 2^e > max |Re|, |Im| over the panels the launch reads (recorded by the U12
// kernel in umax). The refinement against the FP64 matrix that follows the
// factorisation absorbs the 2^-22 products; a non-finite term (growth beyond
// the margin) makes the refinement fail and the frequency fall back to FP64.
//
// Four real products per complex product (Re = Ar Br - Ai Bi, Im = Ar Bi +
// Ai Br), twelve MFMAs per 32 x 32 block and k chunk; the kernel is bound by
// its C traffic, not by these. A is staged negated, so C + (-A) B; the
// result is C + acc / (s_a s_b), an exact power-of-two scaling.
//
// Tiles (64 WM) x (64 WN), four waves in a 2 x 2 grid, each with WM x WN
// blocks of 32 x 32. The chunk of 16 k is staged in LDS as the four fp16
// parts (Re h, Re l, Im h, Im l) of A [row][k] and of B [col][k], so an MFMA
// operand (8 consecutive k of one row or column) is one 16-byte LDS read.
// Every element accumulates its products in the same order for any (WM, WN).
// ---------------------------------------------------------------------------
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef _Float16 h4_t __attribute__((ext_vector_type(4)));
typedef float f16x_t __attribute__((ext_vector_type(16)));
constexpr int HK = 16;          // k per LDS chunk: one MFMA depth
constexpr int HKP = HK + 8;     // padded row length (halves) in LDS
constexpr int kSplitExpA = 12;  // A scale 2^12
constexpr int kSplitExpB = 10;  // max |B| s_b < 2^10

__device__ inline h8_t neg8(h8_t v) { return -v; }


__device__ inline f16x_t mfma32(h8_t a, h8_t b, f16x_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

template <int WM, int WN>
__global__ __launch_bounds__(256, 2) void k_cgemm_sub_h2(int M, int N, int K, float* __restrict__ Pr,
                                                         float* __restrict__ Pi, int ld, int ra, int ca, int rb,
                                                         int cb, int rc, int cc, int critical,
                                                         const float* __restrict__ umax) {
    if (critical) __builtin_amdgcn_s_setprio(3);
    constexpr int TM = 64 * WM, TN = 64 * WN;
    constexpr int NQA = TM * HK / 4 / 256;        // quads of 4 k per thread per chunk (A)
    constexpr int NQB = TN * HK / 4 / 256;
    __shared__ __align__(16) _Float16 As[2][4][TM][HKP];   // [buf][Re h, Re l, Im h, Im l][row][k]
    __shared__ __align__(16) _Float16 Bs[2][4][TN][HKP];   // [buf][part][col][k]
    int tiles_m = (M + TM - 1) / TM, tiles_n = (N + TN - 1) / TN;
    int ntiles = tiles_m * tiles_n;
    constexpr int GROUP_M = 8;                     // XCD-aware grouped order, as in k_cgemm_sub
    int bid = blockIdx.x;
    int per = (ntiles + 7) / 8;
    int tile = (bid % 8) * per + bid / 8;
    if (tile >= ntiles) return;
    int group = tile / (GROUP_M * tiles_n), first_m = group * GROUP_M;
    int gsize = min(GROUP_M, tiles_m - first_m), in_group = tile - group * GROUP_M * tiles_n;
    int tm = first_m + in_group % gsize, tn = in_group / gsize;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int rowt = tm * TM, colt = tn * TN;
    const int mrows = min(TM, M - rowt), ncols = min(TN, N - colt);
    const size_t ldl = (size_t)ld;
    // operand scales
    float um = 0.0f;
    for (int q = 0; q < K / LU_NB; q++) um = fmaxf(um, umax[rb / LU_NB + q]);
    int eb = 0;
    if (um > 0.0f && um < INFINITY) (void)frexpf(um, &eb);   // um < 2^eb
    const float sa = -(float)(1 << kSplitExpA);              // negated A
    const float sb = ldexpf(1.0f, kSplitExpB - eb);
    const float inv = ldexpf(1.0f, eb - kSplitExpB - kSplitExpA);

    float ga[2][NQA][4];
    float4 gb[2][NQB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int l = 0; l < NQA; l++) {
            int q = tid + 256 * l, row = min(q % TM, mrows - 1), kq = q / TM;
            size_t o = (size_t)(ca + k0 + 4 * kq) * ldl + ra + rowt + row;
#pragma unroll
            for (int t = 0; t < 4; t++) { ga[0][l][t] = Pr[o + t * ldl]; ga[1][l][t] = Pi[o + t * ldl]; }
        }
#pragma unroll
        for (int l = 0; l < NQB; l++) {
            int q = tid + 256 * l, col = min(q / 4, ncols - 1), kq = q % 4;
            size_t o = (size_t)(cb + colt + col) * ldl + rb + k0 + 4 * kq;
            gb[0][l] = *reinterpret_cast<const float4*>(Pr + o);
            gb[1][l] = *reinterpret_cast<const float4*>(Pi + o);
        }
    };
    auto split4 = [](const float* x, float s, h4_t& hi, h4_t& lo) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
            float v = x[t] * s;
            _Float16 h = (_Float16)v;
            hi[t] = h;
            lo[t] = (_Float16)(v - (float)h);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int l = 0; l < NQA; l++) {
            int q = tid + 256 * l, row = q % TM, kq = q / TM;
#pragma unroll
            for (int p = 0; p < 2; p++) {
                h4_t hi, lo;
                split4(ga[p][l], sa, hi, lo);
                *reinterpret_cast<h4_t*>(&As[buf][2 * p][row][4 * kq]) = hi;
                *reinterpret_cast<h4_t*>(&As[buf][2 * p + 1][row][4 * kq]) = lo;
            }
        }
#pragma unroll
        for (int l = 0; l < NQB; l++) {
            int q = tid + 256 * l, col = q / 4, kq = q % 4;
#pragma unroll
            for (int p = 0; p < 2; p++) {
                float x[4] = {gb[p][l].x, gb[p][l].y, gb[p][l].z, gb[p][l].w};
                h4_t hi, lo;
                split4(x, sb, hi, lo);
                *reinterpret_cast<h4_t*>(&Bs[buf][2 * p][col][4 * kq]) = hi;
                *reinterpret_cast<h4_t*>(&Bs[buf][2 * p + 1][col][4 * kq]) = lo;
            }
        }
    };

    const int row0 = wm * 32 * WM, col0 = wn * 32 * WN;     // the wave's blocks within the tile
    const bool active = row0 < mrows && col0 < ncols;
    const int li = lane & 31, lk = 8 * (lane >> 5);          // operand row / column and k offset
    f16x_t accR[WM][WN], accI[WM][WN];
    float4 cR[WM][WN][4], cI[WM][WN][4];
#pragma unroll
    for (int a = 0; a < WM; a++)
#pragma unroll
        for (int b = 0; b < WN; b++) {
            accR[a][b] = f16x_t{};
            accI[a][b] = f16x_t{};
        }
    const int nchunk = K / HK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int c = 0; c < nchunk; c++) {
        const int buf = c & 1;
        if (c + 1 < nchunk) gload((c + 1) * HK);
        if (c == 0 && active) {   // C, added at the end: its loads overlap the first chunk
#pragma unroll
            for (int a = 0; a < WM; a++)
#pragma unroll
                for (int b = 0; b < WN; b++) {
                    size_t off = (size_t)(cc + colt + col0 + 32 * b + li) * ldl + rc + rowt + row0 + 32 * a +
                                 4 * (lane >> 5);
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        cR[a][b][q] = *reinterpret_cast<const float4*>(Pr + off + 8 * q);
                        cI[a][b][q] = *reinterpret_cast<const float4*>(Pi + off + 8 * q);
                    }
                }
        }
        if (active) {
            h8_t arh[WM], arl[WM], aih[WM], ail[WM], nih[WM], nil[WM];
            h8_t brh[WN], brl[WN], bih[WN], bil[WN];
#pragma unroll
            for (int a = 0; a < WM; a++) {
                int r = row0 + 32 * a + li;
                arh[a] = *reinterpret_cast<const h8_t*>(&As[buf][0][r][lk]);
                arl[a] = *reinterpret_cast<const h8_t*>(&As[buf][1][r][lk]);
                aih[a] = *reinterpret_cast<const h8_t*>(&As[buf][2][r][lk]);
                ail[a] = *reinterpret_cast<const h8_t*>(&As[buf][3][r][lk]);
                nih[a] = neg8(aih[a]);
                nil[a] = neg8(ail[a]);
            }
#pragma unroll
            for (int b = 0; b < WN; b++) {
                int cl = col0 + 32 * b + li;
                brh[b] = *reinterpret_cast<const h8_t*>(&Bs[buf][0][cl][lk]);
                brl[b] = *reinterpret_cast<const h8_t*>(&Bs[buf][1][cl][lk]);
                bih[b] = *reinterpret_cast<const h8_t*>(&Bs[buf][2][cl][lk]);
                bil[b] = *reinterpret_cast<const h8_t*>(&Bs[buf][3][cl][lk]);
            }
            // Re += Ar Br - Ai Bi,  Im += Ar Bi + Ai Br  (three split products each),
            // alternating the two accumulators
#pragma unroll
            for (int a = 0; a < WM; a++)
#pragma unroll
                for (int b = 0; b < WN; b++) {
                    f16x_t R = accR[a][b], I = accI[a][b];
                    R = mfma32(arh[a], brh[b], R);
                    I = mfma32(arh[a], bih[b], I);
                    R = mfma32(arh[a], brl[b], R);
                    I = mfma32(arh[a], bil[b], I);
                    R = mfma32(arl[a], brh[b], R);
                    I = mfma32(arl[a], bih[b], I);
                    R = mfma32(nih[a], bih[b], R);
                    I = mfma32(aih[a], brh[b], I);
                    R = mfma32(nih[a], bil[b], R);
                    I = mfma32(aih[a], brl[b], I);
                    R = mfma32(nil[a], bih[b], R);
                    I = mfma32(ail[a], brh[b], I);
                    accR[a][b] = R;
                    accI[a][b] = I;
                }
        }
        if (c + 1 < nchunk) {
            sstore(buf ^ 1);
            __syncthreads();
        }
    }
    if (!active) return;
    // accumulator element e of a 32 x 32 block: row 8 (e / 4) + 4 (lane / 32) + e % 4, column lane % 32
#pragma unroll
    for (int a = 0; a < WM; a++)
#pragma unroll
        for (int b = 0; b < WN; b++) {
            size_t off = (size_t)(cc + colt + col0 + 32 * b + li) * ldl + rc + rowt + row0 + 32 * a + 4 * (lane >> 5);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                float4 r = cR[a][b][q], m = cI[a][b][q];
                r.x = fmaf(accR[a][b][4 * q + 0], inv, r.x);
                r.y = fmaf(accR[a][b][4 * q + 1], inv, r.y);
                r.z = fmaf(accR[a][b][4 * q + 2], inv, r.z);
                r.w = fmaf(accR[a][b][4 * q + 3], inv, r.w);
                m.x = fmaf(accI[a][b][4 * q + 0], inv, m.x);
                m.y = fmaf(accI[a][b][4 * q + 1], inv, m.y);
                m.z = fmaf(accI[a][b][4 * q + 2], inv, m.z);
                m.w = fmaf(accI[a][b][4 * q + 3], inv, m.w);
                *reinterpret_cast<float4*>(Pr + off + 8 * q) = r;
                *reinterpret_cast<float4*>(Pi + off + 8 * q) = m;
            }
        }
}

// ---------------------------------------------------------------------------
// Operands of a panel group's K = g NB updates in fp16 split form, made once
// per group instead of in every tile of every update that reads them (the
// split's VALU work had made the update kernel VALU-bound). For the group's
// rows kg.. of its L21 columns (A) and columns kg.. of its U12 rows (B):
//   AS[p][c][q][m][i] = part p of -2^12 A(ra + m, ca + 16 c + 8 q + i),   m < M
//   BS[p][c][q][n][i] = part p of 2^(10 - e) A(rb + 16 c + 8 q + i, cb + n), n < N
// with parts Re h, Re l, Im h, Im l (h = fp16(x s), l = fp16(x s - h)) and 2^e
// above the largest |Re|, |Im| of the K / NB panels' U12 rows, the scales of
// k_cgemm_sub_h2. Chunk-major (16 k per chunk), then the chunk's two 8-k
// halves q: an MFMA operand (8 consecutive k of one row or column) is one
// 16-byte unit, and the units of consecutive rows of one (part, chunk, half)
// are contiguous, so a wave's 64 16-byte loads read 1 KB in one piece. The
// staged values equal those k_cgemm_sub_h2 forms, so the update kernels on
// them compute the same sums.
// Grid: cdiv(M, 64) K/64 tiles of A, then cdiv(N, 64) K/64 tiles of B.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_split_ops(int M, int N, int K, const float* __restrict__ Pr,
                                                   const float* __restrict__ Pi, int ld, int ra, int ca, int rb,
                                                   int cb, const float* __restrict__ umax, _Float16* __restrict__ AS,
                                                   _Float16* __restrict__ BS) {
    __shared__ float tr[LU_NB][LU_NB + 1], ti[LU_NB][LU_NB + 1];   // A tile, [k][m]
    const int tid = threadIdx.x;
    const int kt = K / LU_NB, ta = (M + LU_NB - 1) / LU_NB * kt;
    float um = 0.0f;
    for (int q = 0; q < kt; q++) um = fmaxf(um, umax[rb / LU_NB + q]);
    int eb = 0;
    if (um > 0.0f && um < INFINITY) (void)frexpf(um, &eb);
    auto split16 = [](const float* x, float s, h8_t* hi, h8_t* lo) {
#pragma unroll
        for (int t = 0; t < 16; t++) {
            float v = x[t] * s;
            _Float16 h = (_Float16)v;
            hi[t / 8][t % 8] = h;
            lo[t / 8][t % 8] = (_Float16)(v - (float)h);
        }
    };
    const int kq = tid / 64, j = tid % 64;   // output: chunk kq of the tile, row (column) j
    if ((int)blockIdx.x < ta) {
        const int m0 = (blockIdx.x / kt) * LU_NB, k0 = (blockIdx.x % kt) * LU_NB;
        for (int e = tid; e < LU_NB * LU_NB; e += 256) {
            int m = e % LU_NB, k = e / LU_NB;
            int mm = min(m0 + m, M - 1);
            size_t o = (size_t)(ca + k0 + k) * ld + ra + mm;
            tr[k][m] = Pr[o]; ti[k][m] = Pi[o];
        }
        __syncthreads();
        if (m0 + j >= M) return;
        const float sa = -(float)(1 << kSplitExpA);
        float xr[16], xi[16];
        for (int t = 0; t < 16; t++) { xr[t] = tr[16 * kq + t][j]; xi[t] = ti[16 * kq + t][j]; }
        h8_t h[2], l[2];
        // unit (chunk, half 0, row) and the half-1 unit M rows later
        const size_t pa = (size_t)M * K, o = ((size_t)(k0 / HK + kq) * 2 * M + m0 + j) * 8, o1 = o + (size_t)M * 8;
        split16(xr, sa, h, l);
        *reinterpret_cast<h8_t*>(AS + o) = h[0]; *reinterpret_cast<h8_t*>(AS + o1) = h[1];
        *reinterpret_cast<h8_t*>(AS + pa + o) = l[0]; *reinterpret_cast<h8_t*>(AS + pa + o1) = l[1];
        split16(xi, sa, h, l);
        *reinterpret_cast<h8_t*>(AS + 2 * pa + o) = h[0]; *reinterpret_cast<h8_t*>(AS + 2 * pa + o1) = h[1];
        *reinterpret_cast<h8_t*>(AS + 3 * pa + o) = l[0]; *reinterpret_cast<h8_t*>(AS + 3 * pa + o1) = l[1];
    } else {
        const int b = blockIdx.x - ta;
        const int n0 = (b / kt) * LU_NB, k0 = (b % kt) * LU_NB;
        if (n0 + j >= N) return;
        const float sb = ldexpf(1.0f, kSplitExpB - eb);
        const size_t src = (size_t)(cb + n0 + j) * ld + rb + k0 + 16 * kq;
        float xr[16], xi[16];
        for (int t = 0; t < 16; t += 4) {
            float4 vr = *reinterpret_cast<const float4*>(Pr + src + t);
            float4 vi = *reinterpret_cast<const float4*>(Pi + src + t);
            xr[t] = vr.x; xr[t + 1] = vr.y; xr[t + 2] = vr.z; xr[t + 3] = vr.w;
            xi[t] = vi.x; xi[t + 1] = vi.y; xi[t + 2] = vi.z; xi[t + 3] = vi.w;
        }
        h8_t h[2], l[2];
        const size_t pb = (size_t)N * K, o = ((size_t)(k0 / HK + kq) * 2 * N + n0 + j) * 8, o1 = o + (size_t)N * 8;
        split16(xr, sb, h, l);
        *reinterpret_cast<h8_t*>(BS + o) = h[0]; *reinterpret_cast<h8_t*>(BS + o1) = h[1];
        *reinterpret_cast<h8_t*>(BS + pb + o) = l[0]; *reinterpret_cast<h8_t*>(BS + pb + o1) = l[1];
        split16(xi, sb, h, l);
        *reinterpret_cast<h8_t*>(BS + 2 * pb + o) = h[0]; *reinterpret_cast<h8_t*>(BS + 2 * pb + o1) = h[1];
        *reinterpret_cast<h8_t*>(BS + 3 * pb + o) = l[0]; *reinterpret_cast<h8_t*>(BS + 3 * pb + o1) = l[1];
    }
}

// The update of k_cgemm_sub_h2 on operands split beforehand by k_split_ops:
// the chunk's halves go to LDS with 16-byte loads and no conversion. A rows
// are AS rows 0 .. M-1 (C rows rc ..); B columns are BS columns nb0 .. nb0+N-1
// of ldn (C columns cc ..); the MFMA sequence and the epilogue are those of
// k_cgemm_sub_h2, so the results are the same bit for bit.
template <int WM, int WN, int PD>
__global__ __launch_bounds__(256, 2) void k_cgemm_sub_hs(int M, int N, int K, float* __restrict__ Pr,
                                                         float* __restrict__ Pi, int ld, int rb, int rc, int cc,
                                                         int critical, const float* __restrict__ umax,
                                                         const _Float16* __restrict__ AS,
                                                         const _Float16* __restrict__ BS, int nb0, int ldn) {
    if (critical) __builtin_amdgcn_s_setprio(3);
    constexpr int TM = 64 * WM, TN = 64 * WN;
    __shared__ __align__(16) _Float16 As[2][4][TM][HKP];   // [buf][Re h, Re l, Im h, Im l][row][k]
    __shared__ __align__(16) _Float16 Bs[2][4][TN][HKP];   // [buf][part][col][k]
    int tiles_m = (M + TM - 1) / TM, tiles_n = (N + TN - 1) / TN;
    int ntiles = tiles_m * tiles_n;
    constexpr int GROUP_M = 8;                     // XCD-aware grouped order, as in k_cgemm_sub
    int bid = blockIdx.x;
    int per = (ntiles + 7) / 8;
    int tile = (bid % 8) * per + bid / 8;
    if (tile >= ntiles) return;
    int group = tile / (GROUP_M * tiles_n), first_m = group * GROUP_M;
    int gsize = min(GROUP_M, tiles_m - first_m), in_group = tile - group * GROUP_M * tiles_n;
    int tm = first_m + in_group % gsize, tn = in_group / gsize;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int rowt = tm * TM, colt = tn * TN;
    const int mrows = min(TM, M - rowt), ncols = min(TN, N - colt);
    const size_t ldl = (size_t)ld;
    // operand scales
    float um = 0.0f;
    for (int q = 0; q < K / LU_NB; q++) um = fmaxf(um, umax[rb / LU_NB + q]);
    int eb = 0;
    if (um > 0.0f && um < INFINITY) (void)frexpf(um, &eb);   // um < 2^eb
    const float inv = ldexpf(1.0f, eb - kSplitExpB - kSplitExpA);

    // staged halves: (part, row or column, k half of the chunk), 16 bytes each
    constexpr int NA16 = 4 * TM * 2 / 256, NB16 = 4 * TN * 2 / 256;
    // PD register sets: chunk c waits in set c % PD from its load, issued PD
    // chunks ahead, until its store to LDS at the end of chunk c - 1
    h8_t ha[PD][NA16], hb[PD][NB16];
    const size_t pa = (size_t)M * K, pb = (size_t)ldn * K;
    // load q: row (column) q % TM of k half (q / TM) % 2 of part q / (2 TM):
    // consecutive threads read consecutive 16-byte units of the split layout
    auto gload = [&](int k0, int set) {
#pragma unroll
        for (int l = 0; l < NA16; l++) {
            int q = tid + 256 * l, row = q % TM, hf = (q / TM) & 1, part = q / (2 * TM);
            size_t o = part * pa + ((size_t)(k0 / HK) * 2 * M + (size_t)hf * M + rowt + min(row, mrows - 1)) * 8;
            ha[set][l] = *reinterpret_cast<const h8_t*>(AS + o);
        }
#pragma unroll
        for (int l = 0; l < NB16; l++) {
            int q = tid + 256 * l, col = q % TN, hf = (q / TN) & 1, part = q / (2 * TN);
            size_t o = part * pb + ((size_t)(k0 / HK) * 2 * ldn + (size_t)hf * ldn + nb0 + colt + min(col, ncols - 1)) * 8;
            hb[set][l] = *reinterpret_cast<const h8_t*>(BS + o);
        }
    };
    auto sstore = [&](int buf, int set) {
#pragma unroll
        for (int l = 0; l < NA16; l++) {
            int q = tid + 256 * l, row = q % TM, hf = (q / TM) & 1, part = q / (2 * TM);
            *reinterpret_cast<h8_t*>(&As[buf][part][row][8 * hf]) = ha[set][l];
        }
#pragma unroll
        for (int l = 0; l < NB16; l++) {
            int q = tid + 256 * l, col = q % TN, hf = (q / TN) & 1, part = q / (2 * TN);
            *reinterpret_cast<h8_t*>(&Bs[buf][part][col][8 * hf]) = hb[set][l];
        }
    };

    const int row0 = wm * 32 * WM, col0 = wn * 32 * WN;     // the wave's blocks within the tile
    const bool active = row0 < mrows && col0 < ncols;
    const int li = lane & 31, lk = 8 * (lane >> 5);          // operand row / column and k offset
    f16x_t accR[WM][WN], accI[WM][WN];
    float4 cR[WM][WN][4], cI[WM][WN][4];
#pragma unroll
    for (int a = 0; a < WM; a++)
#pragma unroll
        for (int b = 0; b < WN; b++) {
            accR[a][b] = f16x_t{};
            accI[a][b] = f16x_t{};
        }
    const int nchunk = K / HK;
#pragma unroll
    for (int u = 0; u < PD; u++)
        if (u < nchunk) gload(u * HK, u);
    sstore(0, 0);
    __syncthreads();
    for (int c0 = 0; c0 < nchunk; c0 += PD)
#pragma unroll
    for (int u = 0; u < PD; u++) {
        const int c = c0 + u;   // its halves came from set u
        if (c >= nchunk) break;
        const int buf = c & 1;
        if (c + PD < nchunk) gload((c + PD) * HK, u);
        if (c == 0 && active) {   // C, added at the end: its loads overlap the first chunk
#pragma unroll
            for (int a = 0; a < WM; a++)
#pragma unroll
                for (int b = 0; b < WN; b++) {
                    size_t off = (size_t)(cc + colt + col0 + 32 * b + li) * ldl + rc + rowt + row0 + 32 * a +
                                 4 * (lane >> 5);
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        cR[a][b][q] = *reinterpret_cast<const float4*>(Pr + off + 8 * q);
                        cI[a][b][q] = *reinterpret_cast<const float4*>(Pi + off + 8 * q);
                    }
                }
        }
        if (active) {
            h8_t arh[WM], arl[WM], aih[WM], ail[WM], nih[WM], nil[WM];
            h8_t brh[WN], brl[WN], bih[WN], bil[WN];
#pragma unroll
            for (int a = 0; a < WM; a++) {
                int r = row0 + 32 * a + li;
                arh[a] = *reinterpret_cast<const h8_t*>(&As[buf][0][r][lk]);
                arl[a] = *reinterpret_cast<const h8_t*>(&As[buf][1][r][lk]);
                aih[a] = *reinterpret_cast<const h8_t*>(&As[buf][2][r][lk]);
                ail[a] = *reinterpret_cast<const h8_t*>(&As[buf][3][r][lk]);
                nih[a] = neg8(aih[a]);
                nil[a] = neg8(ail[a]);
            }
#pragma unroll
            for (int b = 0; b < WN; b++) {
                int cl = col0 + 32 * b + li;
                brh[b] = *reinterpret_cast<const h8_t*>(&Bs[buf][0][cl][lk]);
                brl[b] = *reinterpret_cast<const h8_t*>(&Bs[buf][1][cl][lk]);
                bih[b] = *reinterpret_cast<const h8_t*>(&Bs[buf][2][cl][lk]);
                bil[b] = *reinterpret_cast<const h8_t*>(&Bs[buf][3][cl][lk]);
            }
            // Re += Ar Br - Ai Bi,  Im += Ar Bi + Ai Br  (three split products each),
            // alternating the two accumulators
#pragma unroll
            for (int a = 0; a < WM; a++)
#pragma unroll
                for (int b = 0; b < WN; b++) {
                    f16x_t R = accR[a][b], I = accI[a][b];
                    R = mfma32(arh[a], brh[b], R);
                    I = mfma32(arh[a], bih[b], I);
                    R = mfma32(arh[a], brl[b], R);
                    I = mfma32(arh[a], bil[b], I);
                    R = mfma32(arl[a], brh[b], R);
                    I = mfma32(arl[a], bih[b], I);
                    R = mfma32(nih[a], bih[b], R);
                    I = mfma32(aih[a], brh[b], I);
                    R = mfma32(nih[a], bil[b], R);
                    I = mfma32(aih[a], brl[b], I);
                    R = mfma32(nil[a], bih[b], R);
                    I = mfma32(ail[a], brh[b], I);
                    accR[a][b] = R;
                    accI[a][b] = I;
                }
        }
        if (c + 1 < nchunk) {
            sstore(buf ^ 1, (u + 1) % PD);
            __syncthreads();
        }
    }
    if (!active) return;
    // accumulator element e of a 32 x 32 block: row 8 (e / 4) + 4 (lane / 32) + e % 4, column lane % 32
#pragma unroll
    for (int a = 0; a < WM; a++)
#pragma unroll
        for (int b = 0; b < WN; b++) {
            size_t off = (size_t)(cc + colt + col0 + 32 * b + li) * ldl + rc + rowt + row0 + 32 * a + 4 * (lane >> 5);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                float4 r = cR[a][b][q], m = cI[a][b][q];
                r.x = fmaf(accR[a][b][4 * q + 0], inv, r.x);
                r.y = fmaf(accR[a][b][4 * q + 1], inv, r.y);
                r.z = fmaf(accR[a][b][4 * q + 2], inv, r.z);
                r.w = fmaf(accR[a][b][4 * q + 3], inv, r.w);
                m.x = fmaf(accI[a][b][4 * q + 0], inv, m.x);
                m.y = fmaf(accI[a][b][4 * q + 1], inv, m.y);
                m.z = fmaf(accI[a][b][4 * q + 2], inv, m.z);
                m.w = fmaf(accI[a][b][4 * q + 3], inv, m.w);
                *reinterpret_cast<float4*>(Pr + off + 8 * q) = r;
                *reinterpret_cast<float4*>(Pi + off + 8 * q) = m;
            }
        }
}

// Wide tiles for k_cgemm_sub_hs's updates: 128 x 128 per work-group, four
// waves in a 2 x 2 grid with 2 x 2 blocks of 32 x 32 each (64 x 64 per wave).
// Every 16-byte operand read from LDS then feeds two blocks' MFMAs, half the
// LDS reads per MFMA of the 32 x 64 wave tile, and the work-group reads 32 KB
// of operands per 192 MFMAs instead of 24 KB per 96.
// - C is read after the products: its registers do not fit beside the four
//   blocks' accumulators. The other work-group on the CU runs meanwhile.
// - DB = false: the LDS holds one chunk (49 KB, two work-groups per CU); the
//   next chunk's halves wait in registers while the current one is
//   multiplied, with a barrier before and after the store. DB = true: two
//   chunks in LDS (98 KB, one work-group per CU), one barrier per chunk.
// - The products of each accumulator are those of k_cgemm_sub_hs in the same
//   order, and the epilogue is the same: equal results bit for bit. The four
//   accumulators of a 32-row block pair take their products interleaved, so
//   no MFMA waits on the one before it.
// M and N are multiples of 64, so a wave's 64 x 64 block lies wholly inside
// the update or wholly outside (the host checks).
template <bool DB>
__global__ __launch_bounds__(256, DB ? 1 : 2) void k_cgemm_sub_hw(int M, int N, int K, float* __restrict__ Pr,
                                                                  float* __restrict__ Pi, int ld, int rb, int rc,
                                                                  int cc, int critical,
                                                                  const float* __restrict__ umax,
                                                                  const _Float16* __restrict__ AS,
                                                                  const _Float16* __restrict__ BS, int nb0, int ldn) {
    if (critical) __builtin_amdgcn_s_setprio(3);
    constexpr int TM = 128, TN = 128, NBUF = DB ? 2 : 1;
    __shared__ __align__(16) _Float16 As[NBUF][4][TM][HKP];   // [buf][Re h, Re l, Im h, Im l][row][k]
    __shared__ __align__(16) _Float16 Bs[NBUF][4][TN][HKP];   // [buf][part][col][k]
    const int tiles_m = (M + TM - 1) / TM, tiles_n = (N + TN - 1) / TN, ntiles = tiles_m * tiles_n;
    constexpr int GROUP_M = 8;                     // XCD-aware grouped order, as in k_cgemm_sub
    const int bid = blockIdx.x, per = (ntiles + 7) / 8;
    const int tile = (bid % 8) * per + bid / 8;
    if (tile >= ntiles) return;
    const int group = tile / (GROUP_M * tiles_n), first_m = group * GROUP_M;
    const int gsize = min(GROUP_M, tiles_m - first_m), in_group = tile - group * GROUP_M * tiles_n;
    const int tm = first_m + in_group % gsize, tn = in_group / gsize;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rowt = tm * TM, colt = tn * TN;
    const int mrows = min(TM, M - rowt), ncols = min(TN, N - colt);
    float um = 0.0f;
    for (int q = 0; q < K / LU_NB; q++) um = fmaxf(um, umax[rb / LU_NB + q]);
    int eb = 0;
    if (um > 0.0f && um < INFINITY) (void)frexpf(um, &eb);   // um < 2^eb
    const float inv = ldexpf(1.0f, eb - kSplitExpB - kSplitExpA);

    // staging: load l is part l of row (column) tid % 128, k half tid / 128 of the chunk
    const int srow = tid & 127, shf = tid >> 7;
    const size_t pa = (size_t)M * K, pb = (size_t)ldn * K;
    const size_t arow = (size_t)(rowt + min(srow, mrows - 1)), bcol = (size_t)(nb0 + colt + min(srow, ncols - 1));
    h8_t ha[4], hb[4];
    auto gload = [&](int c) {
#pragma unroll
        for (int l = 0; l < 4; l++) {
            ha[l] = *reinterpret_cast<const h8_t*>(AS + l * pa + ((size_t)c * 2 * M + (size_t)shf * M + arow) * 8);
            hb[l] = *reinterpret_cast<const h8_t*>(BS + l * pb + ((size_t)c * 2 * ldn + (size_t)shf * ldn + bcol) * 8);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int l = 0; l < 4; l++) {
            *reinterpret_cast<h8_t*>(&As[buf][l][srow][8 * shf]) = ha[l];
            *reinterpret_cast<h8_t*>(&Bs[buf][l][srow][8 * shf]) = hb[l];
        }
    };

    const int row0 = (wave >> 1) * 64, col0 = (wave & 1) * 64;
    const bool active = row0 < mrows && col0 < ncols;
    const int li = lane & 31, lk = 8 * (lane >> 5);
    f16x_t accR[2][2], accI[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            accR[a][b] = f16x_t{};
            accI[a][b] = f16x_t{};
        }
    const int nchunk = K / HK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int c = 0; c < nchunk; c++) {
        const int buf = DB ? (c & 1) : 0;
        if (c + 1 < nchunk) gload(c + 1);
        if (active) {
            h8_t bo[2][4];   // [b][Re h, Re l, Im h, Im l]
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int p = 0; p < 4; p++)
                    bo[b][p] = *reinterpret_cast<const h8_t*>(&Bs[buf][p][col0 + 32 * b + li][lk]);
#pragma unroll
            for (int a = 0; a < 2; a++) {
                h8_t ao[4];
#pragma unroll
                for (int p = 0; p < 4; p++)
                    ao[p] = *reinterpret_cast<const h8_t*>(&As[buf][p][row0 + 32 * a + li][lk]);
                const h8_t nih = neg8(ao[2]), nil = neg8(ao[3]);
                // Re: arh brh, arh brl, arl brh, -aih bih, -aih bil, -ail bih
                // Im: arh bih, arh bil, arl bih,  aih brh,  aih brl,  ail brh
                const h8_t ra[6] = {ao[0], ao[0], ao[1], nih, nih, nil};
                const h8_t ia[6] = {ao[0], ao[0], ao[1], ao[2], ao[2], ao[3]};
                constexpr int rbp[6] = {0, 1, 0, 2, 3, 2}, ibp[6] = {2, 3, 2, 0, 1, 0};
#pragma unroll
                for (int s = 0; s < 6; s++)
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        accR[a][b] = mfma32(ra[s], bo[b][rbp[s]], accR[a][b]);
                        accI[a][b] = mfma32(ia[s], bo[b][ibp[s]], accI[a][b]);
                    }
            }
        }
        if (c + 1 < nchunk) {
            if constexpr (DB) {
                sstore(buf ^ 1);
                __syncthreads();
            } else {
                __syncthreads();
                sstore(0);
                __syncthreads();
            }
        }
    }
    if (!active) return;
    // accumulator element e of a 32 x 32 block: row 8 (e / 4) + 4 (lane / 32) + e % 4, column lane % 32
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const size_t off =
                (size_t)(cc + colt + col0 + 32 * b + li) * ld + rc + rowt + row0 + 32 * a + 4 * (lane >> 5);
            float4 cr[4], ci[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                cr[q] = *reinterpret_cast<const float4*>(Pr + off + 8 * q);
                ci[q] = *reinterpret_cast<const float4*>(Pi + off + 8 * q);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                float4 r = cr[q], m = ci[q];
                r.x = fmaf(accR[a][b][4 * q + 0], inv, r.x);
                r.y = fmaf(accR[a][b][4 * q + 1], inv, r.y);
                r.z = fmaf(accR[a][b][4 * q + 2], inv, r.z);
                r.w = fmaf(accR[a][b][4 * q + 3], inv, r.w);
                m.x = fmaf(accI[a][b][4 * q + 0], inv, m.x);
                m.y = fmaf(accI[a][b][4 * q + 1], inv, m.y);
                m.z = fmaf(accI[a][b][4 * q + 2], inv, m.z);
                m.w = fmaf(accI[a][b][4 * q + 3], inv, m.w);
                *reinterpret_cast<float4*>(Pr + off + 8 * q) = r;
                *reinterpret_cast<float4*>(Pi + off + 8 * q) = m;
            }
        }
}

// 128 x 128 tiles of the pre-split updates with the operands copied to LDS by
// the memory path itself: global_load_lds_dwordx4 writes each lane's 16 bytes
// at the wave's LDS base + 16 lane, so the copy takes no registers and no LDS
// store instructions.
// - LDS per buffer: A and B each as [part][k half][128 rows][8 halves]
//   (16 KB); two buffers (64 KB), two work-groups per CU.
// - A copy instruction moves one (part, k half, 64-row block): 1 KB in one
//   piece in the split layout and in LDS. Each wave issues four for A and four
//   for B per chunk.
// - Chunk c + 1 is copied while chunk c is multiplied; each wave waits for its
//   own copies (vmcnt) before the one barrier per chunk, which makes them
//   visible and frees the buffer that the copies after it overwrite.
// - MFMA operand reads of 32 consecutive rows are 32 consecutive 16-byte
//   units: no bank conflicts.
// Waves, products and epilogue as k_cgemm_sub_hw: equal results bit for bit.
// M and N multiples of 64 (the host checks).
typedef _Float16 OpBuf[2][4][2][128][8];   // [A, B][part][k half][row][8 halves]

__device__ inline void lds_copy16(const _Float16* src, _Float16* dst) {
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}

__global__ __launch_bounds__(256, 2) void k_cgemm_sub_hd(int M, int N, int K, float* __restrict__ Pr,
                                                         float* __restrict__ Pi, int ld, int rb, int rc, int cc,
                                                         int critical, const float* __restrict__ umax,
                                                         const _Float16* __restrict__ AS,
                                                         const _Float16* __restrict__ BS, int nb0, int ldn) {
    if (critical) __builtin_amdgcn_s_setprio(3);
    constexpr int TM = 128, TN = 128;
    __shared__ __align__(1024) OpBuf L0;
    __shared__ __align__(1024) OpBuf L1;
    const int tiles_m = (M + TM - 1) / TM, tiles_n = (N + TN - 1) / TN, ntiles = tiles_m * tiles_n;
    constexpr int GROUP_M = 8;                     // XCD-aware grouped order, as in k_cgemm_sub
    const int bid = blockIdx.x, per = (ntiles + 7) / 8;
    const int tile = (bid % 8) * per + bid / 8;
    if (tile >= ntiles) return;
    const int group = tile / (GROUP_M * tiles_n), first_m = group * GROUP_M;
    const int gsize = min(GROUP_M, tiles_m - first_m), in_group = tile - group * GROUP_M * tiles_n;
    const int tm = first_m + in_group % gsize, tn = in_group / gsize;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rowt = tm * TM, colt = tn * TN;
    const int mrows = min(TM, M - rowt), ncols = min(TN, N - colt);
    float um = 0.0f;
    for (int q = 0; q < K / LU_NB; q++) um = fmaxf(um, umax[rb / LU_NB + q]);
    int eb = 0;
    if (um > 0.0f && um < INFINITY) (void)frexpf(um, &eb);   // um < 2^eb
    const float inv = ldexpf(1.0f, eb - kSplitExpB - kSplitExpA);

    // the wave's copies i = 0..3: j = 4 wave + i is (part j / 4, k half j / 2 % 2,
    // rows 64 (j % 2) ..); lane l takes row 64 (j % 2) + l (clamped into the update)
    const size_t pa = (size_t)M * K, pb = (size_t)ldn * K;
    const size_t astep = (size_t)16 * M, bstep = (size_t)16 * ldn;   // halves per chunk in one part
    size_t asrc[4], bsrc[4];
    int lofs[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int j = 4 * wave + i, p = j >> 2, h = (j >> 1) & 1, r = 64 * (j & 1) + lane;
        asrc[i] = p * pa + ((size_t)h * M + rowt + min(r, mrows - 1)) * 8;
        bsrc[i] = p * pb + ((size_t)h * ldn + nb0 + colt + min(r, ncols - 1)) * 8;
        lofs[i] = ((p * 2 + h) * TM + 64 * (j & 1)) * 8;   // halves; the lane's unit is added by the copy
    }
    auto copy = [&](OpBuf& buf, int c) {
        _Float16* a = &buf[0][0][0][0][0];
        _Float16* b = &buf[1][0][0][0][0];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            lds_copy16(AS + asrc[i] + c * astep, a + lofs[i]);
            lds_copy16(BS + bsrc[i] + c * bstep, b + lofs[i]);
        }
    };
    auto copies_done = [] { __builtin_amdgcn_s_waitcnt(0x0F70); };   // vmcnt(0)

    const int row0 = (wave >> 1) * 64, col0 = (wave & 1) * 64;
    const bool active = row0 < mrows && col0 < ncols;
    const int li = lane & 31, lh = lane >> 5;
    f16x_t accR[2][2], accI[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            accR[a][b] = f16x_t{};
            accI[a][b] = f16x_t{};
        }
    auto multiply = [&](const OpBuf& buf) {
        if (!active) return;
        h8_t bo[2][4];   // [b][Re h, Re l, Im h, Im l]
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int p = 0; p < 4; p++)
                bo[b][p] = *reinterpret_cast<const h8_t*>(&buf[1][p][lh][col0 + 32 * b + li][0]);
#pragma unroll
        for (int a = 0; a < 2; a++) {
            h8_t ao[4];
#pragma unroll
            for (int p = 0; p < 4; p++) ao[p] = *reinterpret_cast<const h8_t*>(&buf[0][p][lh][row0 + 32 * a + li][0]);
            const h8_t nih = neg8(ao[2]), nil = neg8(ao[3]);
            const h8_t ra[6] = {ao[0], ao[0], ao[1], nih, nih, nil};
            const h8_t ia[6] = {ao[0], ao[0], ao[1], ao[2], ao[2], ao[3]};
            constexpr int rbp[6] = {0, 1, 0, 2, 3, 2}, ibp[6] = {2, 3, 2, 0, 1, 0};
#pragma unroll
            for (int s = 0; s < 6; s++)
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    accR[a][b] = mfma32(ra[s], bo[b][rbp[s]], accR[a][b]);
                    accI[a][b] = mfma32(ia[s], bo[b][ibp[s]], accI[a][b]);
                }
        }
    };
    const int nchunk = K / HK;
    copy(L0, 0);
    copies_done();
    __syncthreads();
    for (int c = 0; c < nchunk; c += 2) {
        if (c + 1 < nchunk) copy(L1, c + 1);
        multiply(L0);
        copies_done();
        __syncthreads();
        if (c + 1 >= nchunk) break;
        if (c + 2 < nchunk) copy(L0, c + 2);
        multiply(L1);
        copies_done();
        __syncthreads();
    }
    if (!active) return;
    // accumulator element e of a 32 x 32 block: row 8 (e / 4) + 4 (lane / 32) + e % 4, column lane % 32
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const size_t off =
                (size_t)(cc + colt + col0 + 32 * b + li) * ld + rc + rowt + row0 + 32 * a + 4 * (lane >> 5);
            float4 cr[4], ci[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                cr[q] = *reinterpret_cast<const float4*>(Pr + off + 8 * q);
                ci[q] = *reinterpret_cast<const float4*>(Pi + off + 8 * q);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                float4 r = cr[q], m = ci[q];
                r.x = fmaf(accR[a][b][4 * q + 0], inv, r.x);
                r.y = fmaf(accR[a][b][4 * q + 1], inv, r.y);
                r.z = fmaf(accR[a][b][4 * q + 2], inv, r.z);
                r.w = fmaf(accR[a][b][4 * q + 3], inv, r.w);
                m.x = fmaf(accI[a][b][4 * q + 0], inv, m.x);
                m.y = fmaf(accI[a][b][4 * q + 1], inv, m.y);
                m.z = fmaf(accI[a][b][4 * q + 2], inv, m.z);
                m.w = fmaf(accI[a][b][4 * q + 3], inv, m.w);
                *reinterpret_cast<float4*>(Pr + off + 8 * q) = r;
                *reinterpret_cast<float4*>(Pi + off + 8 * q) = m;
            }
        }
}

// ---------------------------------------------------------------------------
// Triangular solves with NB-row blocks, one persistent launch per solve.
//
// Block form of the lower solve (the upper one mirrors it with c > b):
//   X_b = Inv_b (B_b - sum_{c<b} M_bc X_c) = Z_b - P_b X_{b-1},
//   Z_b = Inv_b (B_b - sum_{c<b-1} M_bc X_c),   P_b = Inv_b M_{b,b-1},
// with Inv_b the stored inverse of the diagonal triangle. P_b is formed once
// per factorisation (k_trsm_pcoef), and a work-group computes Z_b while the
// blocks before it are still being solved, so once X_{b-1} is out a single
// 64 x 64 by 64 x nr product stands between it and X_b.
//
// Publication: X_b goes to X (for the kernels after the solve) and to `pub`,
// where every 32 bits of a value travel in one 64-bit word tagged with the
// solve's epoch (value bits | epoch << 32), stored with agent-scope relaxed
// atomics. Readers poll exactly the words they need with agent-scope relaxed
// loads until all carry the current epoch, so no flag, release fence or L2
// write-back is involved, and the buffer never needs clearing (each launch
// takes a new epoch). Work-groups take blocks in dependency order from a
// ticket counter, so a block only waits for blocks whose work-groups are
// already running. Ticket and epoch share one 64-bit counter word
// (epoch << 32 | ticket): every work-group of a launch reads the same epoch
// with its ticket, and the work-group taking the last ticket starts the next
// epoch at ticket 0. The epoch therefore lives on the device, and a captured
// graph replays the solve unchanged. The word starts at epoch 1 (the
// publication buffer starts zeroed, epoch 0).
// ---------------------------------------------------------------------------
template <typename TX> struct PubWords;
template <> struct PubWords<float> { static constexpr int W = 1; };
template <> struct PubWords<double> { static constexpr int W = 2; };

__device__ inline void pub_store(unsigned long long* p, float v, unsigned epoch) {
    unsigned long long w = ((unsigned long long)epoch << 32) | (unsigned long long)__float_as_uint(v);
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void pub_store(unsigned long long* p, double v, unsigned epoch) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v), e = (unsigned long long)epoch << 32;
    __hip_atomic_store(p, e | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, e | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// One published value; false while any of its words is from an older epoch.
__device__ inline bool pub_load(unsigned long long* p, unsigned epoch, float& v) {
    unsigned long long w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v = __uint_as_float((unsigned)w);
    return (unsigned)(w >> 32) == epoch;
}
__device__ inline bool pub_load(unsigned long long* p, unsigned epoch, double& v) {
    unsigned long long h = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long l = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v = __longlong_as_double((long long)((h << 32) | (l & 0xffffffffull)));
    return (unsigned)(h >> 32) == epoch && (unsigned)(l >> 32) == epoch;
}
// Spins (one lane) until the word at p carries the epoch.
__device__ inline void pub_wait(unsigned long long* p, unsigned epoch) {
    while ((unsigned)(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) != epoch)
        __builtin_amdgcn_s_sleep(2);
}
// word offset of element (row r, rhs j, part p: 0 re / 1 im) of block c
template <typename TX>
__device__ inline size_t pub_off(int c, int j, int r, int p) {
    return ((((size_t)c * MV_MAXR + j) * LU_NB + r) * 2 + p) * PubWords<TX>::W;
}

// P coefficients of the block solves: PL_b = Linv_b L_{b,b-1} (b >= 1) and
// PU_b = Uinv_b U_{b,b+1} (b <= nblk-2), 64 x 64 column-major per block,
// from the factor in place (L strictly below, U on and above the diagonal).
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_pcoef(int Np, const T* __restrict__ Ar, const T* __restrict__ Ai,
                                                    const T* Li_r, const T* Li_i, const T* Ui_r, const T* Ui_i,
                                                    T* PL_r, T* PL_i, T* PU_r, T* PU_i) {
    typedef MfmaT<T> MT;
    typedef typename MT::acc acc_t;
    int b = blockIdx.x, nblk = Np / LU_NB;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, i = lane & 15, g = lane >> 4;
    for (int which = 0; which < 2; which++) {
        int c = which == 0 ? b - 1 : b + 1;
        if (c < 0 || c >= nblk) continue;
        const T* Ir = (which == 0 ? Li_r : Ui_r) + (size_t)b * NB2;
        const T* Ii = (which == 0 ? Li_i : Ui_i) + (size_t)b * NB2;
        T* Or = (which == 0 ? PL_r : PU_r) + (size_t)b * NB2;
        T* Oi = (which == 0 ? PL_i : PU_i) + (size_t)b * NB2;
        for (int tj = 0; tj < 4; tj++) {
            acc_t cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
            for (int kc = 0; kc < 4; kc++) {
                size_t ao = (size_t)(kc * 16) * LU_NB + wave * 16;
                size_t bo = (size_t)(c * LU_NB + tj * 16) * Np + b * LU_NB + kc * 16;
                cmma16<T>(Ir + ao, Ii + ao, LU_NB, Ar + bo, Ai + bo, (size_t)Np, cr, ci, i, g);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                size_t o = (size_t)(tj * 16 + i) * LU_NB + wave * 16 + MT::row(g, r);
                Or[o] = cr[r];
                Oi[o] = ci[r];
            }
        }
    }
}

// lower = 1: L (unit diagonal) forward with Inv = Linv, P = PL; lower = 0: U
// backward with Inv = Uinv, P = PU. X is Np x nr planar (ld Np); the factor
// and X have the same type T.
//
// All products run on MFMA with wave w owning rows 16w..16w+15 of the block
// and lane column i = right-hand side: the accumulator starts at B_b, every
// dependency c other than `prev` subtracts M_bc X_c, then Z = Inv_b y and
// X_b = Z - P_b X_prev. A block of X is staged once per work-group into LDS
// (one batched poll: all loads are issued before any epoch is checked) and
// each lane then reads its 16 operands in one burst.
template <typename T>
__device__ inline void cmac(typename MfmaT<T>::acc& cr, typename MfmaT<T>::acc& ci, T ar, T ai, T br, T bi) {
    typedef MfmaT<T> MT;
    cr = MT::mma(ar, br, cr);
    cr = MT::mma(-ai, bi, cr);
    ci = MT::mma(ar, bi, ci);
    ci = MT::mma(ai, br, ci);
}

// Stages block c of the published X (rows 0..63, right-hand sides < nr) into
// sr/si [rhs][row]; returns once every value carries the epoch.
template <typename TX>
__device__ inline void stage_block(unsigned long long* pub, unsigned epoch, int c, int nr, int tid, int lane,
                                   TX (*sr)[LU_NB + 1], TX (*si)[LU_NB + 1]) {
    constexpr int W = PubWords<TX>::W;
    for (;;) {
        TX v[8];
        int ok = 1;
#pragma unroll
        for (int q = 0; q < 8; q++) {          // value t: rhs t / 128, row (t % 128) / 2, part t % 2
            int t = tid + 256 * q;
            v[q] = 0;
            if (t < 128 * nr) {
                size_t o = ((size_t)(c * MV_MAXR + (t >> 7)) * LU_NB) * 2 * W + (size_t)(t & 127) * W;
                ok &= pub_load(pub + o, epoch, v[q]) ? 1 : 0;
            }
        }
        if (__all(ok)) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                int t = tid + 256 * q;
                if (t < 128 * nr) {
                    if (t & 1) si[t >> 7][(t & 127) >> 1] = v[q];
                    else sr[t >> 7][(t & 127) >> 1] = v[q];
                }
            }
            return;
        }
        // not out yet: the batch is read again. Only the work-group that waits
        // for the block finishing its own (the chain) polls a block that is not
        // out; the others find theirs published when they get to them.
        __builtin_amdgcn_s_sleep(1);
    }
}

template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_trsm_persist(int Np, int nr, int lower, const T* __restrict__ Ar,
                                                      const T* __restrict__ Ai, const T* __restrict__ Inv_r,
                                                      const T* __restrict__ Inv_i, const T* __restrict__ P_r,
                                                      const T* __restrict__ P_i, TX* Xr, TX* Xi,
                                                      unsigned long long* pub, unsigned long long* counter,
                                                      const int* skip, double* X64r, double* X64i, int x64_add) {
    static_assert(sizeof(T) == sizeof(TX), "factor and solution share one precision");
    typedef MfmaT<T> MT;
    typedef typename MT::acc acc_t;
    if (skip && *skip) return;   // uniform: no work-group takes a ticket
    __shared__ TX sr[MV_MAXR][LU_NB + 1], si[MV_MAXR][LU_NB + 1];   // a staged block of X / y  [rhs][row]
    __shared__ unsigned ticket, epoch_s;
    int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int i = lane & 15, g = lane >> 4;
    int nblk = Np / LU_NB;
    if (tid == 0) {
        const unsigned long long w = atomicAdd(counter, 1ULL);
        ticket = (unsigned)(w & 0xffffffffULL);
        epoch_s = (unsigned)(w >> 32);
        if (ticket == (unsigned)nblk - 1) atomicExch(counter, (unsigned long long)(epoch_s + 1u) << 32);
    }
    __syncthreads();
    const unsigned epoch = epoch_s;
    int b = lower ? (int)ticket : nblk - 1 - (int)ticket;
    int r0 = b * LU_NB, w0 = r0 + wave * 16;
    int prev = lower ? b - 1 : b + 1;               // the block whose X finishes X_b
    bool has_prev = prev >= 0 && prev < nblk;
    // accumulator = B_b (rows of this wave's tile, column i)
    acc_t cr, ci;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        size_t x = (size_t)i * Np + w0 + MT::row(g, r);
        cr[r] = i < nr ? Xr[x] : (TX)0;
        ci[r] = i < nr ? Xi[x] : (TX)0;
    }
    int ndep = (lower ? b : nblk - 1 - b) - 1;      // all dependencies but `prev`
    // Dependency d: -M_bc (A(i, k = 4s + g)) and the published words of X_c.
    // Both are loaded one dependency ahead, while the current one feeds the
    // MFMAs, so each step of this loop waits for no memory round trip of its
    // own; words that were not out yet at the early load are polled again.
    struct Dep {
        T ma_r[16], ma_i[16];
        TX xv[8];
        int ok;
    };
    auto load_dep = [&](int d, Dep& q) {
        const int c = lower ? d : nblk - 1 - d;
#pragma unroll
        for (int s = 0; s < 16; s++) {
            size_t o = (size_t)(c * LU_NB + 4 * s + g) * Np + w0 + i;
            q.ma_r[s] = -Ar[o]; q.ma_i[s] = -Ai[o];
        }
        constexpr int W = PubWords<TX>::W;
        q.ok = 1;
#pragma unroll
        for (int u = 0; u < 8; u++) {               // value t: rhs t / 128, row (t % 128) / 2, part t % 2
            int t = tid + 256 * u;
            q.xv[u] = 0;
            if (t < 128 * nr) {
                size_t o = ((size_t)(c * MV_MAXR + (t >> 7)) * LU_NB) * 2 * W + (size_t)(t & 127) * W;
                q.ok &= pub_load(pub + o, epoch, q.xv[u]) ? 1 : 0;
            }
        }
    };
    auto use_dep = [&](int d, const Dep& q) {
        const int c = lower ? d : nblk - 1 - d;
        __syncthreads();                            // the previous block's operands are read
        if (__all(q.ok)) {                          // this wave's words were all out at the early load
#pragma unroll
            for (int u = 0; u < 8; u++) {
                int t = tid + 256 * u;
                if (t < 128 * nr) {
                    if (t & 1) si[t >> 7][(t & 127) >> 1] = q.xv[u];
                    else sr[t >> 7][(t & 127) >> 1] = q.xv[u];
                }
            }
        } else {
            stage_block<TX>(pub, epoch, c, nr, tid, lane, sr, si);
        }
        __syncthreads();
        T xb_r[16], xb_i[16];
#pragma unroll
        for (int s = 0; s < 16; s++) {
            xb_r[s] = i < nr ? sr[i][4 * s + g] : (TX)0;
            xb_i[s] = i < nr ? si[i][4 * s + g] : (TX)0;
        }
#pragma unroll
        for (int s = 0; s < 16; s++) cmac<T>(cr, ci, q.ma_r[s], q.ma_i[s], xb_r[s], xb_i[s]);
    };
    {
        Dep q0, q1;
        if (ndep > 0) load_dep(0, q0);
        for (int d = 0; d < ndep; d += 2) {
            if (d + 1 < ndep) load_dep(d + 1, q1);
            use_dep(d, q0);
            if (d + 1 < ndep) {
                if (d + 2 < ndep) load_dep(d + 2, q0);
                use_dep(d + 1, q1);
            }
        }
    }
    // y = B_b - sum into LDS as the B operand of Inv_b
    T va_r[16], va_i[16];
#pragma unroll
    for (int s = 0; s < 16; s++) {
        size_t o = (size_t)b * NB2 + (size_t)(4 * s + g) * LU_NB + wave * 16 + i;
        va_r[s] = Inv_r[o]; va_i[s] = Inv_i[o];
    }
    __syncthreads();
    if (i < nr) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            sr[i][wave * 16 + MT::row(g, r)] = cr[r];
            si[i][wave * 16 + MT::row(g, r)] = ci[r];
        }
    }
    __syncthreads();
    T xb_r[16], xb_i[16];
#pragma unroll
    for (int s = 0; s < 16; s++) {
        xb_r[s] = i < nr ? sr[i][4 * s + g] : (TX)0;
        xb_i[s] = i < nr ? si[i][4 * s + g] : (TX)0;
    }
    acc_t zr = {0, 0, 0, 0}, zi = {0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 16; s++) cmac<T>(zr, zi, va_r[s], va_i[s], xb_r[s], xb_i[s]);
    if (has_prev) {
#pragma unroll
        for (int s = 0; s < 16; s++) {
            size_t o = (size_t)b * NB2 + (size_t)(4 * s + g) * LU_NB + wave * 16 + i;
            va_r[s] = -P_r[o]; va_i[s] = -P_i[o];
        }
        __syncthreads();                           // y is read
        stage_block<TX>(pub, epoch, prev, nr, tid, lane, sr, si);
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 16; s++) {
            xb_r[s] = i < nr ? sr[i][4 * s + g] : (TX)0;
            xb_i[s] = i < nr ? si[i][4 * s + g] : (TX)0;
        }
#pragma unroll
        for (int s = 0; s < 16; s++) cmac<T>(zr, zi, va_r[s], va_i[s], xb_r[s], xb_i[s]);
    }
    if (i < nr) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int row = wave * 16 + MT::row(g, r);
            pub_store(pub + pub_off<TX>(b, i, row, 0), zr[r], epoch);
            pub_store(pub + pub_off<TX>(b, i, row, 1), zi[r], epoch);
            const size_t o = (size_t)i * Np + r0 + row;
            const TX xr = zr[r], xi = zi[r];
            Xr[o] = xr;
            Xi[o] = xi;
            if (X64r) {   // the FP64 solution: stored, or the correction added to it
                if (x64_add) { X64r[o] += (double)xr; X64i[o] += (double)xi; }
                else { X64r[o] = (double)xr; X64i[o] = (double)xi; }
            }
        }
    }
}

// Row permutation of the right-hand sides for the interchanges recorded in
// ipiv (applied sequentially, one thread per column); only if any panel fell
// back to pivoting.
template <typename TX>
__global__ void k_apply_ipiv(int Np, int nr, const int* ipiv, const int* any_flag, TX* Xr, TX* Xi, const int* skip) {
    if (skip && *skip) return;
    if (!any_flag[0]) return;
    int j = threadIdx.x;
    if (j >= nr) return;
    size_t a = (size_t)j * Np;
    for (int k = 0; k < Np; k++) {
        int p = ipiv[k];
        if (p != k) {
            TX tr = Xr[a + k], ti = Xi[a + k];
            Xr[a + k] = Xr[a + p]; Xi[a + k] = Xi[a + p];
            Xr[a + p] = tr; Xi[a + p] = ti;
        }
    }
}

// OR of all panel flags into flag[npanel] (any interchange / zero pivot).
__global__ void k_flags_any(int npanel, int* flag) {
    __shared__ int s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    int v = 0;
    for (int i = threadIdx.x; i < npanel; i += blockDim.x) v |= flag[i];
    if (v) atomicOr(&s, v);
    __syncthreads();
    if (threadIdx.x == 0) flag[npanel] = s;
}

__global__ void k_init_lu(int Np, int npanel, int* flag, int* ipiv, float* umax) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Np) ipiv[i] = i;
    if (i <= npanel) flag[i] = 0;
    if (umax && i < npanel) umax[i] = 0.0f;
}

// ---------------------------------------------------------------------------
// Refinement helpers
// ---------------------------------------------------------------------------
__device__ inline void refine_clear(RefineCtrl* ctrl) {
    ctrl->done = 0; ctrl->iters = 0; ctrl->converged = 0;
    ctrl->rel = 1e300; ctrl->prev = 1e300;
    for (int q = 0; q < 4; q++)
        for (int j = 0; j < kMaxRhs; j++) ctrl->mx[q][j] = 0ULL;
    ctrl->an = 0ULL;
    ctrl->count = 0u;
}
__global__ void k_to_f32(size_t n, const double* a, const double* b, float* fa, float* fb, const int* skip,
                         RefineCtrl* reset) {
    if (skip && *skip) return;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (reset && i == 0) refine_clear(reset);
    if (i < n) { fa[i] = (float)a[i]; fb[i] = (float)b[i]; }
}
// Interleaved complex n_rows x ncol (column-major, ld n_rows) to planar
// Np x ncolp (ld Np): rows and columns beyond the source are zero, except the
// diagonal when ident = 1 (the identity padding of a system matrix).
__global__ void k_pack_planar(int n_rows, int ncol, int Np, int ncolp, const double* __restrict__ src,
                              double* __restrict__ re, double* __restrict__ im, int ident) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)Np * ncolp) return;
    int i = (int)(t % Np), j = (int)(t / Np);
    double vr = 0.0, vi = 0.0;
    if (i < n_rows && j < ncol) {
        size_t o = 2 * ((size_t)j * n_rows + i);
        vr = src[o]; vi = src[o + 1];
    } else if (ident && i == j) {
        vr = 1.0;
    }
    re[t] = vr; im[t] = vi;
}

// Planar Np-row block (ld Np) to interleaved n_rows x ncol (ld n_rows).
__global__ void k_unpack_planar(int n_rows, int ncol, int Np, const double* __restrict__ re,
                                const double* __restrict__ im, double* __restrict__ dst) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n_rows * ncol) return;
    int i = (int)(t % n_rows), j = (int)(t / n_rows);
    size_t o = (size_t)j * Np + i;
    dst[2 * t] = re[o]; dst[2 * t + 1] = im[o];
}

// Fresh triangular-solve state: ticket counter at epoch 1, publication words
// at epoch 0 (nwords of them).
__global__ void k_solve_state_init(unsigned long long* counter, unsigned long long* pub, size_t nwords) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nwords) pub[t] = 0ull;
    if (t == 0) *counter = 1ull << 32;
}

__global__ void k_copy2(size_t n, const double* a, const double* b, double* ca, double* cb) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { ca[i] = a[i]; cb[i] = b[i]; }
}

// Max-norm of complex vectors r (residual) and x per column: out[2j] = max|r|,
// out[2j+1] = max|x| over rows < N. One work-group per column.
__global__ void k_colmax(int N, int Np, const double* rr, const double* ri, const double* xr, const double* xi, double* out) {
    int j = blockIdx.x;
    __shared__ double s1[256], s2[256];
    double a = 0, b = 0;
    for (int r = threadIdx.x; r < N; r += 256) {
        size_t o = (size_t)j * Np + r;
        a = fmax(a, sqrt(rr[o] * rr[o] + ri[o] * ri[o]));
        b = fmax(b, sqrt(xr[o] * xr[o] + xi[o] * xi[o]));
    }
    s1[threadIdx.x] = a; s2[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) { s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + s]); s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + s]); }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[2 * j] = s1[0]; out[2 * j + 1] = s2[0]; }
}

// Device-side control of iterative refinement (no host round trip inside a
// frequency solve), after a residual R = B - A X, per right-hand side j over
// rows < N:
//   rel_j = max( max|r_j| / (|A| max|x_j|),
//                max|Im r_j| / (|A| max(max|Im x_j|, 0.01 max|x_j|)) ).
// The second term asks the imaginary part for the same relative accuracy: at
// low frequency it is orders of magnitude below the real part (the wave term
// is O(K)), the damping comes from it, and a residual small against |x| alone
// can leave it inaccurate. The 0.01 floor keeps its roundoff floor (about
// 1e2 u |A| |x|) under the tolerance. Converged when max_j rel_j <= tol;
// stops (not converged) on stagnation (rel > 0.5 prev) or after max_iter
// corrections; any NaN in r or x makes rel NaN, read as not converged.

// max that keeps NaN (fmax drops it): a NaN residual or solution must read as
// not converged, never as a zero residual
__device__ inline double nmax(double a, double b) { return (b > a || b != b) ? b : a; }

// rel = max(rel, rel_j) for the column maxima a = max|r|, b = max|x|,
// c = max|Im r|, d = max|Im x|, and |A| = an (any NaN: 1e300).
__device__ inline double refine_column(double rel, double a, double b, double c, double d, double an) {
    double den = an * b, deni = an * fmax(d, 0.01 * b);
    double v = den > 0 ? a / den : (a > 0 ? 1e300 : 0.0);
    double vi = deni > 0 ? c / deni : (c > 0 ? 1e300 : 0.0);
    const bool finite = a == a && b == b && c == c && d == d && an == an;
    return finite ? nmax(rel, nmax(v, vi)) : 1e300;
}
__device__ inline void refine_decide(RefineCtrl* ctrl, double rel, double tol, int max_iter) {
    if (!(rel == rel)) rel = 1e300;
    ctrl->prev = ctrl->rel;
    ctrl->rel = rel;
    if (rel <= tol) { ctrl->done = 1; ctrl->converged = 1; }
    else if (ctrl->iters > 0 && rel > 0.5 * ctrl->prev) ctrl->done = 1;
    else if (ctrl->iters >= max_iter) ctrl->done = 1;
    else ctrl->iters += 1;
}

// The refinement's residual and convergence check in one launch, after the
// partial products of k_matvec_mfma (slab sums, as for k_matvec_reduce):
//   r = b - sum_ks P into R, and its FP32 copy into R32 (the right-hand side
//   of the next correction solve);
//   the column maxima of the criterion above: each wave (64 rows of one
//   column: Np is a multiple of 128) reduces them and posts them to the
//   refinement state with atomicMax on their bit patterns (non-negative
//   doubles order as their bits, and a NaN's bits lie above those of +inf,
//   so a NaN survives as with nmax);
//   with Pabs (the first iteration) also |A|_inf: every column-0 thread sums
//   its row over the slabs (four running sums, so that several loads are in
//   flight), and a NaN row sum is dropped;
//   the last work-group to finish (a counter taken after a device-scope
//   fence) takes the decision and clears the slots for the next iteration.
__global__ __launch_bounds__(256) void k_refine_reduce(int N, int Np, int nr, int nks, const double* __restrict__ Pr,
                                                       const double* __restrict__ Pi, const double* Br,
                                                       const double* Bi, double* Rr, double* Ri, float* R32r,
                                                       float* R32i, const double* Xr, const double* Xi,
                                                       const double* Pabs, double* anorm, RefineCtrl* ctrl,
                                                       double tol, int max_iter) {
    if (ctrl->done) return;   // uniform over the grid
    __shared__ int last;
    const int tid = threadIdx.x, lane = tid & 63;
    const size_t total = (size_t)nr * Np;
    const size_t t = (size_t)blockIdx.x * 256 + tid;
    double mr = 0.0, mx = 0.0, mri = 0.0, mxi = 0.0, an = 0.0;
    if (t < total) {
        const int j = (int)(t / Np), r = (int)(t % Np);
        double sr = 0.0, si = 0.0;
        for (int ks = 0; ks < nks; ks++) {
            size_t o = ((size_t)ks * MV_MAXR + j) * Np + r;
            sr += Pr[o]; si += Pi[o];
        }
        const size_t y = (size_t)j * Np + r;
        sr = Br[y] - sr; si = Bi[y] - si;
        Rr[y] = sr; Ri[y] = si;
        R32r[y] = (float)sr; R32i[y] = (float)si;
        if (r < N) {
            const double xr = Xr[y], xi = Xi[y];
            mr = sqrt(sr * sr + si * si);
            mx = sqrt(xr * xr + xi * xi);
            mri = fabs(si);
            mxi = fabs(xi);
        }
        if (Pabs && j == 0) {
            double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
            int ks = 0;
            for (; ks + 4 <= nks; ks += 4) {
                v0 += Pabs[(size_t)ks * Np + r];
                v1 += Pabs[(size_t)(ks + 1) * Np + r];
                v2 += Pabs[(size_t)(ks + 2) * Np + r];
                v3 += Pabs[(size_t)(ks + 3) * Np + r];
            }
            for (; ks < nks; ks++) v0 += Pabs[(size_t)ks * Np + r];
            an = fmax(0.0, (v0 + v1) + (v2 + v3));
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        mr = nmax(mr, __shfl_xor(mr, off));
        mx = nmax(mx, __shfl_xor(mx, off));
        mri = nmax(mri, __shfl_xor(mri, off));
        mxi = nmax(mxi, __shfl_xor(mxi, off));
        an = fmax(an, __shfl_xor(an, off));
    }
    const size_t tw = t - lane;                    // the wave's first element
    if (lane == 0 && tw < total) {
        const int jw = (int)(tw / Np);
        atomicMax(&ctrl->mx[0][jw], (unsigned long long)__double_as_longlong(mr));
        atomicMax(&ctrl->mx[1][jw], (unsigned long long)__double_as_longlong(mx));
        atomicMax(&ctrl->mx[2][jw], (unsigned long long)__double_as_longlong(mri));
        atomicMax(&ctrl->mx[3][jw], (unsigned long long)__double_as_longlong(mxi));
        if (Pabs) atomicMax(&ctrl->an, (unsigned long long)__double_as_longlong(an));
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) last = atomicAdd(&ctrl->count, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (tid == 0) {
        auto rd = [](unsigned long long* p) {
            return __longlong_as_double((long long)__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        };
        double a_n;
        if (Pabs) { a_n = rd(&ctrl->an); anorm[0] = a_n; }
        else a_n = anorm[0];
        double rel = 0.0;
        for (int j = 0; j < nr; j++)
            rel = refine_column(rel, rd(&ctrl->mx[0][j]), rd(&ctrl->mx[1][j]), rd(&ctrl->mx[2][j]), rd(&ctrl->mx[3][j]),
                                a_n);
        refine_decide(ctrl, rel, tol, max_iter);
        for (int q = 0; q < 4; q++)
            for (int j = 0; j < kMaxRhs; j++) ctrl->mx[q][j] = 0ULL;
        ctrl->an = 0ULL;
        ctrl->count = 0u;
    }
}

// ---------------------------------------------------------------------------
// Coefficient reductions (Hydrodynamic_Solver.cpp:463-477): one work-group per
// dot product over panels, fixed-order tree (deterministic).
//   q < 36:            (M Phi_J)_{i,j}, i = q/6, j = q%6; M_i = gen_i * area
//   36 <= q < 36+6nb:  F1_{i,b} = sum M_i Phi_I(:,b)
//   36+6nb <= q:       F2_{i,b} = sum Phi_J(:,i) dPhi_I/dn(:,b) area
// out: complex (2 doubles) per q.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_reduce_coeffs(int N, int Np, int nb, const double* __restrict__ gen,
                                                       const double* Xr, const double* Xi,
                                                       const double* PhiIr, const double* PhiIi,
                                                       const double* DAr, const double* DAi, double* out) {
    int q = blockIdx.x;
    double sr = 0, si = 0;
    for (int k = threadIdx.x; k < N; k += 256) {
        double ar, ai, br, bi;
        if (q < 36) {
            int i = q / 6, j = q % 6;
            ar = gen[(size_t)k * 8 + i] * gen[(size_t)k * 8 + 6]; ai = 0;
            br = Xr[(size_t)j * Np + k]; bi = Xi[(size_t)j * Np + k];
        } else if (q < 36 + 6 * nb) {
            int t = q - 36, i = t / nb, b = t % nb;
            ar = gen[(size_t)k * 8 + i] * gen[(size_t)k * 8 + 6]; ai = 0;
            br = PhiIr[(size_t)b * Np + k]; bi = PhiIi[(size_t)b * Np + k];
        } else {
            int t = q - 36 - 6 * nb, i = t / nb, b = t % nb;
            ar = Xr[(size_t)i * Np + k]; ai = Xi[(size_t)i * Np + k];
            br = DAr[(size_t)b * Np + k]; bi = DAi[(size_t)b * Np + k];
        }
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
    }
    __shared__ double rr[256], ri[256];
    rr[threadIdx.x] = sr; ri[threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) { rr[threadIdx.x] += rr[threadIdx.x + s]; ri[threadIdx.x] += ri[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[2 * q] = rr[0]; out[2 * q + 1] = ri[0]; }
}

// ===========================================================================
// Launchers
// ===========================================================================
static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

void launch_rankine(hipStream_t s, int N, int Np, const double* body, const double* refl, const double* recv,
                    double* Ssrc, double* Srefl, double* Dsrc, double* Drefl) {
    dim3 grid(Np / AS_TR, Np / AS_TS);
    hipLaunchKernelGGL(k_rankine, grid, dim3(256), 0, s, N, Np, body, refl, recv, Ssrc, Srefl, Dsrc, Drefl);
}

void launch_wave_compose(hipStream_t s, int N, int Np, const double* freq, int compose, const double* quad,
                         const double* recv, const double* Ssrc, const double* Srefl, const double* Dsrc,
                         const double* Drefl, double* Sr, double* Si, double* Dr, double* Di, float* D32r,
                         float* D32i, double far_r2, int wg_per_cu) {
    const int ntiles = (Np / AS_TR) * (Np / AS_TS);
    int dev = 0, ncu = 0;
    if (far_r2 < 0.0 && wg_per_cu > 0 && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0 &&
        wg_per_cu * ncu < ntiles) {
        hipLaunchKernelGGL(k_wave_compose_strided, dim3(wg_per_cu * ncu), dim3(256), 0, s, N, Np, freq, compose, quad,
                           recv, Ssrc, Srefl, Dsrc, Drefl, Sr, Si, Dr, Di, D32r, D32i);
        return;
    }
    dim3 grid(Np / AS_TR, Np / AS_TS);
    auto k = far_r2 >= 0.0 ? k_wave_compose<true> : k_wave_compose<false>;
    hipLaunchKernelGGL(k, grid, dim3(256), 0, s, N, Np, freq, compose, quad, recv, Ssrc, Srefl, Dsrc, Drefl, Sr, Si,
                       Dr, Di, D32r, D32i, far_r2);
}

void launch_havelock_eval(hipStream_t s, int n, const double* X, const double* Y, const double* V, double* out,
                          bool far32) {
    auto k = far32 ? k_havelock_eval<true> : k_havelock_eval<false>;
    hipLaunchKernelGGL(k, dim3(cdiv(n, 256)), dim3(256), 0, s, n, X, Y, V, out);
}

void launch_incident(hipStream_t s, int N, int Np, int nb, const double* body, const double* gen, const double* betas,
                     const double* freq, double* Xr, double* Xi, int ldx, double* PhiIr, double* PhiIi,
                     double* DAr, double* DAi) {
    hipLaunchKernelGGL(k_incident, dim3(cdiv(Np, 256)), dim3(256), 0, s, N, Np, nb, body, gen, betas, freq,
                       Xr, Xi, ldx, PhiIr, PhiIi, DAr, DAi);
}

// Y_partial(ks) = M(:, slab ks) X(slab ks, :) on FP64 MFMA 16x16x4: wave w of
// a work-group owns rows 16w..16w+15 of the group's 64, all 16 right-hand
// sides in the accumulators (lane column = right-hand side, columns >= nr are
// zero). Chunks of 64 k of X are staged in LDS; each lane loads its 16 M
// operands of a chunk before the MFMAs, so a wave has 16 KB of M in flight.
// ROWABS: also the row sums of |re| + |im| of M (for |A|_inf), per slab.
template <typename TM, bool ROWABS>
__global__ __launch_bounds__(256) void k_matvec_mfma(int Np, int nr, int kslab, const TM* __restrict__ Mr,
                                                     const TM* __restrict__ Mi, const double* __restrict__ Xr,
                                                     const double* __restrict__ Xi, double* Pr, double* Pi,
                                                     double* Pabs, const int* skip) {
    typedef MfmaT<double> MT;
    typedef MT::acc acc_t;
    if (skip && *skip) return;
    __shared__ double xs[MV_KC][MV_MAXR][2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, g = lane >> 4;
    const int row = blockIdx.x * MV_ROWS + wave * 16 + i, ks = blockIdx.y;
    const int kb = ks * kslab, ke = min(Np, kb + kslab);
    acc_t cr = acc_t{0, 0, 0, 0}, ci = acc_t{0, 0, 0, 0};
    double sabs = 0.0;
    for (int kc = kb; kc < ke; kc += MV_KC) {
        double mr[16], mi[16];
#pragma unroll
        for (int t = 0; t < 16; t++) {
            size_t o = (size_t)(kc + 4 * t + g) * Np + row;
            mr[t] = (double)Mr[o]; mi[t] = (double)Mi[o];
        }
        __syncthreads();                 // the previous chunk's operands are read
        for (int t = tid; t < MV_KC * MV_MAXR; t += 256) {
            int k = t % MV_KC, j = t / MV_KC;
            bool ok = j < nr;
            xs[k][j][0] = ok ? Xr[(size_t)j * Np + kc + k] : 0.0;
            xs[k][j][1] = ok ? Xi[(size_t)j * Np + kc + k] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 16; t++) {
            double xr = xs[4 * t + g][i][0], xi = xs[4 * t + g][i][1];
            cr = MT::mma(mr[t], xr, cr);
            cr = MT::mma(-mi[t], xi, cr);
            ci = MT::mma(mr[t], xi, ci);
            ci = MT::mma(mi[t], xr, ci);
            if (ROWABS) sabs += fabs(mr[t]) + fabs(mi[t]);
        }
    }
    if (ROWABS) {
        sabs += __shfl_xor(sabs, 16);
        sabs += __shfl_xor(sabs, 32);
        if (g == 0) Pabs[(size_t)ks * Np + row] = sabs;
    }
    if (i < nr) {
        const int rb = blockIdx.x * MV_ROWS + wave * 16;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            size_t o = ((size_t)ks * MV_MAXR + i) * Np + rb + MT::row(g, r);
            Pr[o] = cr[r]; Pi[o] = ci[r];
        }
    }
}

// ---------------------------------------------------------------------------
// Machine-peak microbenchmarks (reported beside the published peaks): MFMA
// 16x16x4 chains on registers (8 independent accumulators per wave, operands
// from the thread index so nothing folds), and a streaming copy.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_peak_mfma(int iters, T* sink) {
    typedef MfmaT<T> MT;
    typedef typename MT::acc acc_t;
    T a = (T)(threadIdx.x & 7) * (T)0.125, b = (T)(threadIdx.x >> 3) * (T)0.0625;
    acc_t c[8];
#pragma unroll
    for (int q = 0; q < 8; q++) c[q] = acc_t{(T)q, 0, 0, 0};
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int q = 0; q < 8; q++) c[q] = MT::mma(a, b, c[q]);
    }
    T v = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) v += c[q][0] + c[q][1] + c[q][2] + c[q][3];
    if (v == (T)-1.2345) sink[blockIdx.x] = v;    // never true; keeps the chains live
}

// fp16 MFMA 32x32x16 chains (the split-operand updates' instruction): eight
// independent accumulators per wave.
__global__ __launch_bounds__(256) void k_peak_mfma_f16(int iters, float* sink) {
    h8_t a, b;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        a[q] = (_Float16)((threadIdx.x + q) & 7) * (_Float16)0.125f;
        b[q] = (_Float16)((threadIdx.x >> 3) & 7) * (_Float16)0.0625f;
    }
    f16x_t c[8];
#pragma unroll
    for (int q = 0; q < 8; q++) c[q] = f16x_t{(float)q};
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int q = 0; q < 8; q++) c[q] = mfma32(a, b, c[q]);
    }
    float v = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) v += c[q][0] + c[q][15];
    if (v == -1.2345f) sink[blockIdx.x] = v;    // never true; keeps the chains live
}

__global__ __launch_bounds__(256) void k_peak_copy(size_t n4, const float4* __restrict__ src, float4* dst) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n4; t += stride) dst[t] = src[t];
}

int measure_peaks(hipStream_t s, double* out) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 1;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return 1; }
    double* sink = nullptr;
    float4 *a = nullptr, *b = nullptr;
    const size_t bytes = (size_t)1 << 30;
    int rc = 0;
    if (hipMalloc(&sink, 4096 * sizeof(double)) != hipSuccess || hipMalloc(&a, bytes) != hipSuccess ||
        hipMalloc(&b, bytes) != hipSuccess) {
        rc = 1;
    } else {
        const int wgs = ncu * 8, iters = 8192;
        auto time_launch = [&](auto launch) {
            launch();   // warm-up
            (void)hipEventRecord(e0, s);
            launch();
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            return (double)ms * 1e-3;
        };
        const double mfma_flops = (double)wgs * 4 * iters * 8 * (16.0 * 16 * 4 * 2);
        double t32 = time_launch([&] {
            hipLaunchKernelGGL(k_peak_mfma<float>, dim3(wgs), dim3(256), 0, s, iters, (float*)sink);
        });
        double t64 = time_launch([&] {
            hipLaunchKernelGGL(k_peak_mfma<double>, dim3(wgs), dim3(256), 0, s, iters, sink);
        });
        (void)hipMemsetAsync(a, 0, bytes, s);
        const size_t n4 = bytes / sizeof(float4);
        double tc = time_launch([&] {
            hipLaunchKernelGGL(k_peak_copy, dim3(ncu * 16), dim3(256), 0, s, n4, (const float4*)a, b);
        });
        const int iters16 = 2048;
        double t16 = time_launch([&] {
            hipLaunchKernelGGL(k_peak_mfma_f16, dim3(wgs), dim3(256), 0, s, iters16, (float*)sink);
        });
        out[0] = mfma_flops / t32 / 1e12;
        out[1] = mfma_flops / t64 / 1e12;
        out[2] = 2.0 * (double)bytes / tc / 1e9;
        out[3] = (double)wgs * 4 * iters16 * 8 * (32.0 * 32 * 16 * 2) / t16 / 1e12;
        if (hipStreamSynchronize(s) != hipSuccess) rc = 1;
    }
    if (sink) (void)hipFree(sink);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int matvec_kslabs(int Np) { return Np >= 4096 ? 32 : (Np >= 1024 ? 16 : 4); }

// K slabs of the tall-skinny products (nks) and their width (kslab).
int matvec_slabs(int Np, int& kslab) {
    int nks = matvec_kslabs(Np);
    kslab = (Np / nks + 63) / 64 * 64;
    return (Np + kslab - 1) / kslab;
}

template <typename TM>
void launch_matvec_partials(hipStream_t s, int Np, int nr, const TM* Mr, const TM* Mi, const double* Xr,
                            const double* Xi, double* Pr, double* Pi, double* Pabs, const int* skip) {
    int kslab = 0;
    const int nks = matvec_slabs(Np, kslab);
    dim3 grid(Np / MV_ROWS, nks);
    if (Pabs)
        hipLaunchKernelGGL((k_matvec_mfma<TM, true>), grid, dim3(256), 0, s, Np, nr, kslab, Mr, Mi, Xr, Xi, Pr, Pi,
                           Pabs, skip);
    else
        hipLaunchKernelGGL((k_matvec_mfma<TM, false>), grid, dim3(256), 0, s, Np, nr, kslab, Mr, Mi, Xr, Xi, Pr,
                           Pi, Pabs, skip);
}
template void launch_matvec_partials<double>(hipStream_t, int, int, const double*, const double*, const double*,
                                             const double*, double*, double*, double*, const int*);

template <typename TM>
void launch_matvec(hipStream_t s, int Np, int nr, const TM* Mr, const TM* Mi, const double* Xr, const double* Xi,
                   double* Pr, double* Pi, const double* Br, const double* Bi, int mode, double* Yr, double* Yi,
                   const int* skip) {
    int kslab = 0;
    const int nks = matvec_slabs(Np, kslab);
    launch_matvec_partials<TM>(s, Np, nr, Mr, Mi, Xr, Xi, Pr, Pi, nullptr, skip);
    hipLaunchKernelGGL(k_matvec_reduce, dim3(cdiv((size_t)nr * Np, 256)), dim3(256), 0, s, Np, nr, nks, Pr, Pi, Br, Bi,
                       mode, Yr, Yi, skip);
}
template void launch_matvec<double>(hipStream_t, int, int, const double*, const double*, const double*, const double*,
                                    double*, double*, const double*, const double*, int, double*, double*, const int*);

// ---------------------------------------------------------------------------
// GEMM tile autotuner. The tile shapes of k_cgemm_sub are timed on the device
// per (M class, N class, K) with M, N classes floor(log2(x / 64)), for the
// shapes a factorisation of a given padded order launches, and the fastest is
// kept per device and precision. Untuned shapes use a size rule.
// ---------------------------------------------------------------------------
namespace {
struct TileCfg { int wm, wn; };
// 128 x 128 tiles would need 3 x 16 accumulator blocks per wave, more
// registers than two waves per SIMD allow, so the shapes stop at 128 x 64.
constexpr TileCfg kCfg[] = {{2, 2}, {4, 2}, {2, 4}};
template <typename T> constexpr int n_cfg() { return 3; }
template <typename T> constexpr TileCfg cfg_of(int i) { return kCfg[i]; }

constexpr int kMaxDev = 16, kClasses = 11;
struct TuneTable {
    signed char pick[2][kClasses][kClasses][3];   // [prec][M class][N class][K class: 64, 128, 192-256]
    signed char split[kClasses][kClasses][3];     // split-arithmetic FP32 updates, split in the tiles
    signed char pre[kClasses][kClasses][3];       // split-arithmetic FP32 updates on pre-split operands
    bool tuned_np[4][kMaxPanels / 2 + 1];          // padded orders already tuned (Np / 128): FP32, FP64, split, pre
    TuneTable() {
        std::memset(pick, -1, sizeof(pick));
        std::memset(split, -1, sizeof(split));
        std::memset(pre, -1, sizeof(pre));
        std::memset(tuned_np, 0, sizeof(tuned_np));
    }
};
TuneTable g_tune[kMaxDev];
std::mutex g_tune_mu;

inline int k_class(int K) { return K <= 64 ? 0 : (K <= 128 ? 1 : 2); }

inline int size_class(int x) {
    int c = 0;
    while ((64 << (c + 1)) <= x && c < kClasses - 1) c++;
    return c;
}

template <typename T, int WM, int WN>
void launch_tile(hipStream_t s, int M, int N, int K, T* Ar, T* Ai, int ld, int ra, int ca, int rb, int cb, int rc,
                 int cc, int critical = 0) {
    constexpr int TM = 32 * WM, TN = 32 * WN;
    int tiles = (int)(cdiv(M, TM) * cdiv(N, TN));
    int per = (tiles + 7) / 8;
    hipLaunchKernelGGL((k_cgemm_sub<T, WM, WN>), dim3(per * 8), dim3(256), 0, s, M, N, K, Ar, Ai, ld, ra, ca, rb, cb,
                       rc, cc, critical);
}

template <typename T>
void launch_cfg(int cfg, hipStream_t s, int M, int N, int K, T* Ar, T* Ai, int ld, int ra, int ca, int rb, int cb,
                int rc, int cc, int critical = 0) {
    switch (cfg) {
    case 1: launch_tile<T, 4, 2>(s, M, N, K, Ar, Ai, ld, ra, ca, rb, cb, rc, cc, critical); return;
    case 2: launch_tile<T, 2, 4>(s, M, N, K, Ar, Ai, ld, ra, ca, rb, cb, rc, cc, critical); return;
    default: launch_tile<T, 2, 2>(s, M, N, K, Ar, Ai, ld, ra, ca, rb, cb, rc, cc, critical); return;
    }
}

// Split-arithmetic tiles: 64 x 64, 128 x 64, 64 x 128 (WM, WN blocks of 64).
template <int WM, int WN>
void launch_split_tile(hipStream_t s, int M, int N, int K, float* Ar, float* Ai, int ld, int ra, int ca, int rb,
                       int cb, int rc, int cc, int critical, const float* umax) {
    constexpr int TM = 64 * WM, TN = 64 * WN;
    int tiles = (int)(cdiv(M, TM) * cdiv(N, TN));
    int per = (tiles + 7) / 8;
    hipLaunchKernelGGL((k_cgemm_sub_h2<WM, WN>), dim3(per * 8), dim3(256), 0, s, M, N, K, Ar, Ai, ld, ra, ca, rb, cb,
                       rc, cc, critical, umax);
}
constexpr int kSplitCfgs = 3;
void launch_split(int cfg, hipStream_t s, int M, int N, int K, float* Ar, float* Ai, int ld, int ra, int ca, int rb,
                  int cb, int rc, int cc, int critical, const float* umax) {
    switch (cfg) {
    case 1: launch_split_tile<2, 1>(s, M, N, K, Ar, Ai, ld, ra, ca, rb, cb, rc, cc, critical, umax); return;
    case 2: launch_split_tile<1, 2>(s, M, N, K, Ar, Ai, ld, ra, ca, rb, cb, rc, cc, critical, umax); return;
    default: launch_split_tile<1, 1>(s, M, N, K, Ar, Ai, ld, ra, ca, rb, cb, rc, cc, critical, umax); return;
    }
}
int default_split(int M, int N) { return cdiv(M, 128) * cdiv(N, 64) >= 512 ? 1 : 0; }

// The same updates on operands split beforehand (k_split_ops): the tiles of
// launch_split with loads one chunk ahead (configurations 0-2) or two (5-7),
// and the 128 x 128 tiles of k_cgemm_sub_hw with one LDS chunk (3) or two (4)
// and of k_cgemm_sub_hd (8), for M and N multiples of 64 (else configuration 0).
template <int WM, int WN, int PD>
void launch_split_pre_tile(hipStream_t s, int M, int N, int K, float* Ar, float* Ai, int ld, int rb, int rc, int cc,
                           int critical, const float* umax, const _Float16* AS, const _Float16* BS, int nb0,
                           int ldn) {
    constexpr int TM = 64 * WM, TN = 64 * WN;
    int tiles = (int)(cdiv(M, TM) * cdiv(N, TN));
    int per = (tiles + 7) / 8;
    hipLaunchKernelGGL((k_cgemm_sub_hs<WM, WN, PD>), dim3(per * 8), dim3(256), 0, s, M, N, K, Ar, Ai, ld, rb, rc, cc,
                       critical, umax, AS, BS, nb0, ldn);
}
template <bool DB>
void launch_split_wide(hipStream_t s, int M, int N, int K, float* Ar, float* Ai, int ld, int rb, int rc, int cc,
                       int critical, const float* umax, const _Float16* AS, const _Float16* BS, int nb0, int ldn) {
    int tiles = (int)(cdiv(M, 128) * cdiv(N, 128));
    int per = (tiles + 7) / 8;
    hipLaunchKernelGGL(k_cgemm_sub_hw<DB>, dim3(per * 8), dim3(256), 0, s, M, N, K, Ar, Ai, ld, rb, rc, cc, critical,
                       umax, AS, BS, nb0, ldn);
}
void launch_split_direct(hipStream_t s, int M, int N, int K, float* Ar, float* Ai, int ld, int rb, int rc, int cc,
                         int critical, const float* umax, const _Float16* AS, const _Float16* BS, int nb0, int ldn) {
    int tiles = (int)(cdiv(M, 128) * cdiv(N, 128));
    int per = (tiles + 7) / 8;
    hipLaunchKernelGGL(k_cgemm_sub_hd, dim3(per * 8), dim3(256), 0, s, M, N, K, Ar, Ai, ld, rb, rc, cc, critical,
                       umax, AS, BS, nb0, ldn);
}
constexpr int kSplitPreCfgs = 9;
void launch_split_pre(int cfg, hipStream_t s, int M, int N, int K, float* Ar, float* Ai, int ld, int rb, int rc,
                      int cc, int critical, const float* umax, const _Float16* AS, const _Float16* BS, int nb0,
                      int ldn) {
    if (cfg >= 3 && (M % 64 != 0 || N % 64 != 0)) cfg = 0;
    switch (cfg) {
    case 3: launch_split_wide<false>(s, M, N, K, Ar, Ai, ld, rb, rc, cc, critical, umax, AS, BS, nb0, ldn); return;
    case 4: launch_split_wide<true>(s, M, N, K, Ar, Ai, ld, rb, rc, cc, critical, umax, AS, BS, nb0, ldn); return;
    case 5: launch_split_pre_tile<1, 1, 2>(s, M, N, K, Ar, Ai, ld, rb, rc, cc, critical, umax, AS, BS, nb0, ldn); return;
    case 6: launch_split_pre_tile<2, 1, 2>(s, M, N, K, Ar, Ai, ld, rb, rc, cc, critical, umax, AS, BS, nb0, ldn); return;
    case 7: launch_split_pre_tile<1, 2, 2>(s, M, N, K, Ar, Ai, ld, rb, rc, cc, critical, umax, AS, BS, nb0, ldn); return;
    case 8: launch_split_direct(s, M, N, K, Ar, Ai, ld, rb, rc, cc, critical, umax, AS, BS, nb0, ldn); return;
    case 1: launch_split_pre_tile<2, 1, 1>(s, M, N, K, Ar, Ai, ld, rb, rc, cc, critical, umax, AS, BS, nb0, ldn); return;
    case 2: launch_split_pre_tile<1, 2, 1>(s, M, N, K, Ar, Ai, ld, rb, rc, cc, critical, umax, AS, BS, nb0, ldn); return;
    default: launch_split_pre_tile<1, 1, 1>(s, M, N, K, Ar, Ai, ld, rb, rc, cc, critical, umax, AS, BS, nb0, ldn); return;
    }
}

// Size rule for untuned shapes: the larger FP32 tile where it still gives 2
// work-groups per CU, else 64 x 64.
template <typename T>
int default_cfg(int M, int N) {
    if constexpr (sizeof(T) == 4) {
        if (cdiv(M, 128) * cdiv(N, 64) >= 512) return 1;
    }
    (void)M; (void)N;
    return 0;
}

template <typename T>
int pick_cfg(int dev, int M, int N, int K) {
    int v = (dev >= 0 && dev < kMaxDev) ? g_tune[dev].pick[sizeof(T) == 8][size_class(M)][size_class(N)][k_class(K)] : -1;
    return v >= 0 ? v : default_cfg<T>(M, N);
}

int pick_split(int dev, int M, int N, int K) {
    int v = (dev >= 0 && dev < kMaxDev) ? g_tune[dev].split[size_class(M)][size_class(N)][k_class(K)] : -1;
    return v >= 0 ? v : default_split(M, N);
}

int pick_split_pre(int dev, int M, int N, int K) {
    int v = (dev >= 0 && dev < kMaxDev) ? g_tune[dev].pre[size_class(M)][size_class(N)][k_class(K)] : -1;
    return v >= 0 ? v : default_split(M, N);
}

// BEM_TUNE_LOG=1 in the environment: the tuner prints each shape's timings
// to stderr.
bool tune_log() {
    static const bool on = [] {
        const char* e = getenv("BEM_TUNE_LOG");
        return e && *e && *e != '0';
    }();
    return on;
}

// The (M, N, K) of every update launch of launch_lu at padded order Np, in
// launch order (the same loop structure as launch_lu).
std::vector<std::array<int, 3>> lu_gemm_shapes(int Np, int group) {
    std::vector<std::array<int, 3>> v;
    int npanel = Np / LU_NB;
    for (int p = 0; p < npanel; p += group) {
        int g = std::min(group, npanel - p), k0 = p * LU_NB;
        for (int j = 1; j < g; j++) {
            int kj = k0 + j * LU_NB;
            v.push_back({Np - kj, LU_NB, j * LU_NB});
            if (Np - kj - LU_NB > 0) v.push_back({LU_NB, Np - kj - LU_NB, j * LU_NB});
        }
        int kg = k0 + g * LU_NB, restg = Np - kg;
        if (restg <= 0) break;
        v.push_back({restg, LU_NB, g * LU_NB});
        if (restg > LU_NB) v.push_back({restg, restg - LU_NB, g * LU_NB});
    }
    return v;
}
}  // namespace

namespace {
// The split-arithmetic tiles timed like the FP32 ones (tune_lu_gemms), with
// the panel maxima cleared (scale 1: the timing does not depend on it).
int tune_split(hipStream_t s, int Np, float* Ar, float* Ai, float* umax) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return 0;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (g_tune[dev].tuned_np[2][Np / 128]) return 0;
    }
    if (hipMemsetAsync(umax, 0, sizeof(float) * (Np / LU_NB), s) != hipSuccess) return 0;
    std::map<std::array<int, 3>, std::array<int, 3>> reps;
    for (auto& sh : lu_gemm_shapes(Np, lu_group<float>(Np, true))) {
        std::array<int, 3> key = {size_class(sh[0]), size_class(sh[1]), k_class(sh[2])};
        auto it = reps.find(key);
        if (it == reps.end() || (double)sh[0] * sh[1] > (double)it->second[0] * it->second[1]) reps[key] = sh;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 0;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return 0; }
    std::vector<std::pair<std::array<int, 3>, int>> picks;
    for (auto& kv : reps) {
        int M = std::min(kv.second[0], Np - 128), N = std::min(kv.second[1], Np - 128), K = kv.second[2];
        if (M < 64 || N < 64) continue;
        float best = 1e30f;
        int bi = default_split(M, N);
        for (int c = 0; c < kSplitCfgs; c++) {
            launch_split(c, s, M, N, K, Ar, Ai, Np, 128, 0, 0, 128, 128, 128, 0, umax);   // warm-up
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < 3; r++) launch_split(c, s, M, N, K, Ar, Ai, Np, 128, 0, 0, 128, 128, 128, 0, umax);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) { best = ms; bi = c; }
        }
        picks.push_back({kv.first, bi});
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& pk : picks) g_tune[dev].split[pk.first[0]][pk.first[1]][pk.first[2]] = (signed char)pk.second;
    g_tune[dev].tuned_np[2][Np / 128] = true;
    return (int)picks.size();
}
}  // namespace

namespace {
// The pre-split update kernels (launch_split_pre) timed on the trailing-update
// shapes of a split-arithmetic factorisation of order Np, with operands from
// the lane's split buffers AS, BS (their contents do not change the timing).
int tune_split_pre(hipStream_t s, int Np, float* Ar, float* Ai, float* umax, const _Float16* AS,
                   const _Float16* BS) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return 0;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (g_tune[dev].tuned_np[3][Np / 128]) return 0;
    }
    if (hipMemsetAsync(umax, 0, sizeof(float) * (Np / LU_NB), s) != hipSuccess) return 0;
    const int group = lu_group<float>(Np, true);
    std::map<std::array<int, 3>, std::array<int, 3>> reps;
    for (int p = 0; p < Np / LU_NB; p += group) {   // the trailing updates of launch_lu
        int kg = (p + group) * LU_NB, restg = Np - kg;
        if (restg <= LU_NB) break;
        std::array<int, 3> sh = {restg, restg - LU_NB, group * LU_NB};
        std::array<int, 3> key = {size_class(sh[0]), size_class(sh[1]), k_class(sh[2])};
        auto it = reps.find(key);
        if (it == reps.end() || (double)sh[0] * sh[1] > (double)it->second[0] * it->second[1]) reps[key] = sh;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 0;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return 0; }
    std::vector<std::pair<std::array<int, 3>, int>> picks;
    for (auto& kv : reps) {
        int M = std::min(kv.second[0], Np - 128), N = std::min(kv.second[1], Np - 128), K = kv.second[2];
        if (M < 64 || N < 64) continue;
        float best = 1e30f;
        int bi = default_split(M, N);
        for (int c = 0; c < kSplitPreCfgs; c++) {
            launch_split_pre(c, s, M, N, K, Ar, Ai, Np, 0, 128, 128, 0, umax, AS, BS, 0, N);   // warm-up
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < 3; r++) launch_split_pre(c, s, M, N, K, Ar, Ai, Np, 0, 128, 128, 0, umax, AS, BS, 0, N);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (tune_log())
                fprintf(stderr, "[bem tune] pre-split M=%d N=%d K=%d cfg %d: %.1f us\n", M, N, K, c, ms * 1e3f / 3);
            if (ms < best) { best = ms; bi = c; }
        }
        picks.push_back({kv.first, bi});
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& pk : picks) g_tune[dev].pre[pk.first[0]][pk.first[1]][pk.first[2]] = (signed char)pk.second;
    g_tune[dev].tuned_np[3][Np / 128] = true;
    return (int)picks.size();
}
}  // namespace

// Times every tile shape on the shapes a factorisation of order Np launches
// (one representative per class, the largest of the class) and records the
// fastest. Ar, Ai: an Np x Np scratch matrix (its contents are overwritten).
template <typename T>
int tune_lu_gemms(hipStream_t s, int Np, T* Ar, T* Ai, float* umax, const _Float16* AS, const _Float16* BS) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return 0;
    int prec = sizeof(T) == 8;
    int tuned = 0;
    if constexpr (sizeof(T) == 4) {
        if (umax) tuned += tune_split(s, Np, Ar, Ai, umax);
        if (umax && AS && BS) tuned += tune_split_pre(s, Np, Ar, Ai, umax, AS, BS);
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (g_tune[dev].tuned_np[prec][Np / 128]) return tuned;
    }
    std::map<std::array<int, 3>, std::array<int, 3>> reps;   // (mc, nc, kb) -> largest (M, N, K)
    for (auto& sh : lu_gemm_shapes(Np, lu_group<T>(Np))) {
        std::array<int, 3> key = {size_class(sh[0]), size_class(sh[1]), k_class(sh[2])};
        auto it = reps.find(key);
        if (it == reps.end() || (double)sh[0] * sh[1] > (double)it->second[0] * it->second[1]) reps[key] = sh;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 0;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return 0; }
    std::vector<std::pair<std::array<int, 3>, int>> picks;
    for (auto& kv : reps) {
        const std::array<int, 3>& sh = kv.second;
        int M = sh[0], N = sh[1], K = sh[2];
        // A at (128, 0), B at (0, 128), C at (128, 128): inside the Np x Np scratch
        if (M > Np - 128 || N > Np - 128) { M = std::min(M, Np - 128); N = std::min(N, Np - 128); }
        if (M < 64 || N < 64) continue;
        float best = 1e30f;
        int bi = default_cfg<T>(M, N);
        for (int c = 0; c < n_cfg<T>(); c++) {
            launch_cfg<T>(c, s, M, N, K, Ar, Ai, Np, 128, 0, 0, 128, 128, 128);   // warm-up
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < 3; r++) launch_cfg<T>(c, s, M, N, K, Ar, Ai, Np, 128, 0, 0, 128, 128, 128);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) { best = ms; bi = c; }
        }
        picks.push_back({kv.first, bi});
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& pk : picks) g_tune[dev].pick[prec][pk.first[0]][pk.first[1]][pk.first[2]] = (signed char)pk.second;
    g_tune[dev].tuned_np[prec][Np / 128] = true;
    return (int)picks.size();
}
template int tune_lu_gemms<float>(hipStream_t, int, float*, float*, float*, const _Float16*, const _Float16*);
template int tune_lu_gemms<double>(hipStream_t, int, double*, double*, float*, const _Float16*, const _Float16*);

template <typename T>
int gemm_tile_choice(int M, int N, int K) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    int c = pick_cfg<T>(dev, M, N, K);
    TileCfg t = cfg_of<T>(c);
    return 1000 * (32 * t.wm) + 32 * t.wn;
}
template int gemm_tile_choice<float>(int, int, int);
template int gemm_tile_choice<double>(int, int, int);

// C(rc.., cc..) -= A(ra.., ca..) B(rb.., cb..), M x N x K (profile: HIP events
// around the launch and the flops of the unpadded part).
template <typename T>
static void gemm_sub(hipStream_t s, int dev, int Np, int M, int N, int K, T* Ar, T* Ai, int ra, int ca, int rb,
                     int cb, int rc, int cc, LuProfile* prof, int critical, const float* umax,
                     const _Float16* AS = nullptr, const _Float16* BS = nullptr, int nb0 = 0, int ldn = 0,
                     int force_cfg = -1) {
    bool timed = prof && prof->n < LuProfile::kMax;
    if (timed) {
        if (!prof->ev[2 * prof->n]) {
            (void)hipEventCreate(&prof->ev[2 * prof->n]);
            (void)hipEventCreate(&prof->ev[2 * prof->n + 1]);
        }
        (void)hipEventRecord(prof->ev[2 * prof->n], s);
    }
    const bool split = sizeof(T) == 4 && umax;
    const int mode = split ? (AS ? 2 : 1) : 0;
    int cfg = split ? (mode == 2 ? (force_cfg >= 0 ? force_cfg : pick_split_pre(dev, M, N, K))
                                 : pick_split(dev, M, N, K))
                    : pick_cfg<T>(dev, M, N, K);
    if constexpr (sizeof(T) == 4) {
        if (mode == 2) launch_split_pre(cfg, s, M, N, K, Ar, Ai, Np, rb, rc, cc, critical, umax, AS, BS, nb0, ldn);
        else if (mode == 1) launch_split(cfg, s, M, N, K, Ar, Ai, Np, ra, ca, rb, cb, rc, cc, critical, umax);
        else launch_cfg<T>(cfg, s, M, N, K, Ar, Ai, Np, ra, ca, rb, cb, rc, cc, critical);
    } else {
        launch_cfg<T>(cfg, s, M, N, K, Ar, Ai, Np, ra, ca, rb, cb, rc, cc, critical);
    }
    if (timed) {
        (void)hipEventRecord(prof->ev[2 * prof->n + 1], s);
        prof->call[prof->n] = {dev, Np, M, N, K, ra, ca, rb, cb, rc, cc, critical, cfg, mode, Ar, Ai, umax, AS, BS,
                               nb0, ldn};
        double mu = (double)std::max(0, std::min(M, prof->N - rc));
        double nu = (double)std::max(0, std::min(N, prof->N - cc));
        double ku = (double)std::max(0, std::min(K, prof->N - ca));
        prof->useful_flops += 8.0 * mu * nu * ku;
        prof->useful_bytes += 2.0 * sizeof(T) * (2.0 * mu * nu + mu * ku + ku * nu);
        prof->n++;
    }
}

template <typename T>
double replay_lu_gemms(const LuProfile& prof, hipStream_t s) {
    if (prof.n == 0) return 0.0;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < prof.n; i++) {
        const LuProfile::Call& c = prof.call[i];
        T* Ar = static_cast<T*>(c.Ar);
        T* Ai = static_cast<T*>(c.Ai);
        if constexpr (sizeof(T) == 4) {
            if (c.split == 2)
                launch_split_pre(c.cfg, s, c.M, c.N, c.K, Ar, Ai, c.Np, c.rb, c.rc, c.cc, c.critical, c.umax, c.AS,
                                 c.BS, c.nb0, c.ldn);
            else if (c.split)
                launch_split(c.cfg, s, c.M, c.N, c.K, Ar, Ai, c.Np, c.ra, c.ca, c.rb, c.cb, c.rc, c.cc, c.critical, c.umax);
            else
                launch_cfg<T>(c.cfg, s, c.M, c.N, c.K, Ar, Ai, c.Np, c.ra, c.ca, c.rb, c.cb, c.rc, c.cc, c.critical);
        } else {
            launch_cfg<T>(c.cfg, s, c.M, c.N, c.K, Ar, Ai, c.Np, c.ra, c.ca, c.rb, c.cb, c.rc, c.cc, c.critical);
        }
    }
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / prof.n;
}
template double replay_lu_gemms<float>(const LuProfile&, hipStream_t);
template double replay_lu_gemms<double>(const LuProfile&, hipStream_t);

// Right-looking blocked LU over groups of lu_group<T>(Np) panels, with lookahead.
// For the group p .. p+g-1 at k0 = p NB (panel p already factored on the
// panel stream s2, every column right of it up to date with the groups
// before):
//   for j = 0 .. g-1, panel q = p + j at kj = k0 + j NB, on the main stream:
//     (j > 0) factor panel q (diagonal block; L21 after the K = j NB update
//             of its rows by the group's panels before it, [L21_p ..
//             L21_{q-1}] [U12 ...]): nothing else could run beside them, and a
//             hand-off between streams costs about 10 us each way;
//     fix q (pivot check / fallback + row swaps elsewhere);
//     U12_q = Linv_q A12 after the K = j NB update of panel q's rows right of
//     it, and (j + 1 < g) the K = (j+1) NB update of panel q+1's diagonal
//     block. These thin updates ride in the panel kernels that consume them
//     instead of taking GEMM launches of their own on the critical path;
//   main stream: the K = g NB update of the next group's first panel, which
//   the panel stream then factors while the main stream applies the same
//   K = g NB update to every column right of it.
// Deferring the group's trailing update to one K = g NB product divides the
// passes over C by g. Row interchanges of panel q (fallback
// pivoting) hit the trailing rows not yet updated by the group's earlier
// panels and those panels' L21 alike, which commutes with the update. Events
// ev[2 q], ev[2 q + 1] (q the first panel of a group) order the two streams.
template <typename T>
void launch_lu(hipStream_t s, hipStream_t s2, hipEvent_t* ev, int Np, T* Ar, T* Ai, const LuWork<T>& w, int* flag,
               int* ipiv, LuProfile* prof, int arith, int latency, double pivot_tau) {
    int npanel = Np / LU_NB;    // even: Np is a multiple of 128
    int dev = 0;
    (void)hipGetDevice(&dev);
    const int force_cfg = ((arith >> kLuTileShift) & kLuTileMask) - 1;   // forced tile of the pre-split updates, -1: tuned
    const bool spec = (arith & kLuSpeculative) != 0;    // no pivoting fallback (the caller re-runs flagged systems)
    const T tau2 = (T)(pivot_tau * pivot_tau);           // multipliers with |l|^2 above it flag their panel
    arith &= kLuArithMask;
    const bool split = sizeof(T) == 4 && (arith == kLuArithSplitF16 || arith == kLuArithSplitF16Tile) && w.umax;
    const bool presplit = split && arith == kLuArithSplitF16 && w.opA && w.opB;
    float* umax = split ? w.umax : nullptr;
    hipLaunchKernelGGL(k_init_lu, dim3(cdiv(Np + 1, 256)), dim3(256), 0, s, Np, npanel, flag, ipiv, umax);
    // kp: first column of the panel's group (kp < k0: the group's deferred
    // update of the panel's rows below the diagonal block is applied by L21)
    // the copies that the pivoting fallback restores a panel from (none in a
    // speculative factorisation, which has no fallback)
    T* sd_r = spec ? nullptr : w.sd_r;
    T* sd_i = spec ? nullptr : w.sd_i;
    T* s21_r = spec ? nullptr : w.s21_r;
    T* s21_i = spec ? nullptr : w.s21_i;
    auto panel = [&](hipStream_t st, int p, int kp) {
        int k0 = p * LU_NB, rest = Np - k0 - LU_NB;
        T *Lir = w.Linv_r + (size_t)p * NB2, *Lii = w.Linv_i + (size_t)p * NB2;
        T *Uir = w.Uinv_r + (size_t)p * NB2, *Uii = w.Uinv_i + (size_t)p * NB2;
        hipLaunchKernelGGL(k_lu_diag<T>, dim3(1), dim3(256), 0, st, Np, k0, Ar, Ai, sd_r, sd_i, Lir, Lii, Uir, Uii,
                           flag, tau2);
        auto l21k = latency ? k_lu_l21<T, true> : k_lu_l21<T, false>;
        if constexpr (sizeof(T) == 4) {
            if (rest > 0 && !latency && Np >= kWidePanelNp) {   // shared device: 64 rows per work-group
                hipLaunchKernelGGL(k_lu_l21_wide<T>, dim3(rest / (16 * WIDE)), dim3(256), 0, st, Np, k0, kp, Ar, Ai,
                                   Uir, Uii, s21_r, s21_i, flag, tau2);
                return;
            }
        }
        if (rest > 0)
            hipLaunchKernelGGL(l21k, dim3(rest / 16), dim3(256), 0, st, Np, k0, kp, Ar, Ai, Uir, Uii, s21_r, s21_i,
                               flag, tau2);
    };
    auto fix = [&](int p) {      // pivot check and fallback of panel p, row swaps elsewhere
        int k0 = p * LU_NB;
        T *Lir = w.Linv_r + (size_t)p * NB2, *Lii = w.Linv_i + (size_t)p * NB2;
        T *Uir = w.Uinv_r + (size_t)p * NB2, *Uii = w.Uinv_i + (size_t)p * NB2;
        hipLaunchKernelGGL(k_lu_panel_gepp<T>, dim3(1), dim3(1024), 0, s, Np, k0, Ar, Ai, w.sd_r, w.sd_i, w.s21_r,
                           w.s21_i, Lir, Lii, Uir, Uii, flag, ipiv);
    };
    // kp as for panel; next: the next panel is in the same group (its
    // diagonal block takes the group's update here)
    auto u12 = [&](int p, int kp, int next) {
        int k0 = p * LU_NB, rest = Np - k0 - LU_NB;
        auto u12k = latency ? k_lu_u12<T, true> : k_lu_u12<T, false>;
        if constexpr (sizeof(T) == 4) {
            if (rest > 0 && !latency && Np >= kWidePanelNp) {   // shared device: 64 columns per work-group
                hipLaunchKernelGGL(k_lu_u12_wide<T>, dim3(rest / (16 * WIDE)), dim3(256), 0, s, Np, k0, kp, next, Ar,
                                   Ai, w.Linv_r + (size_t)p * NB2, w.Linv_i + (size_t)p * NB2, umax);
                return;
            }
        }
        if (rest > 0)
            hipLaunchKernelGGL(u12k, dim3(rest / 16), dim3(256), 0, s, Np, k0, kp, next, Ar, Ai,
                               w.Linv_r + (size_t)p * NB2, w.Linv_i + (size_t)p * NB2, umax);
    };
    const int NB = LU_NB, group = lu_group<T>(Np, split);
    // A factorisation that shares the device with others (several lanes)
    // below kWidePanelNp factors every panel on the main stream: the
    // lookahead's overlap is what the other lanes' work gives anyway, and
    // each active stream is one more hardware queue for the command
    // processor to serve (64 seeded bodies x 16 frequencies: 512 -> 531
    // mixed solves/s; the semi-sub sweep unchanged).
    const bool one_stream = !latency && Np < kWidePanelNp;
    if (one_stream) s2 = s;
    panel(s, 0, 0);
    for (int p = 0; p < npanel; p += group) {
        const int g = std::min(group, npanel - p), k0 = p * NB;
        if (p > 0 && !one_stream) (void)hipStreamWaitEvent(s, ev[2 * p + 1], 0);   // panel p factored on s2
        for (int j = 0; j < g; j++) {
            const int q = p + j;
            // j > 0: the group's K = j NB update of panel q's columns was applied
            // to its diagonal block by U12 of panel q-1 and is applied to the
            // rows below by L21 of panel q; the main stream factors it (nothing
            // to overlap: no cross-stream hand-off)
            if (j > 0) panel(s, q, k0);
            if (!spec) fix(q);
            // U12 of panel q after the K = j NB update of its rows, and the
            // K = (j+1) NB update of panel q+1's diagonal block
            u12(q, k0, j + 1 < g);
        }
        const int kg = k0 + g * NB, restg = Np - kg;
        if (restg <= 0) break;
        // the next group's first panel (thin: its operands split in the tiles),
        // then hand it to the panel stream
        gemm_sub<T>(s, dev, Np, restg, NB, g * NB, Ar, Ai, kg, k0, k0, kg, kg, kg, prof, 1, umax);
        if (!one_stream) {
            (void)hipEventRecord(ev[2 * (p + g)], s);
            (void)hipStreamWaitEvent(s2, ev[2 * (p + g)], 0);
        }
        panel(s2, p + g, kg);
        if (!one_stream) (void)hipEventRecord(ev[2 * (p + g) + 1], s2);
        if (restg > NB) {   // trailing update of everything right of it
            // split arithmetic: its operands (rows kg.. of the group's L21,
            // columns kg+NB.. of its U12) in fp16 split form once, beside the
            // panel stream's work
            const _Float16 *AS = nullptr, *BS = nullptr;
            if constexpr (sizeof(T) == 4) {
                if (presplit) {
                    const int grid = (cdiv(restg, NB) + cdiv(restg - NB, NB)) * g;
                    hipLaunchKernelGGL(k_split_ops, dim3(grid), dim3(256), 0, s, restg, restg - NB, g * NB, Ar, Ai, Np,
                                       kg, k0, k0, kg + NB, umax, w.opA, w.opB);
                    AS = w.opA;
                    BS = w.opB;
                }
            }
            gemm_sub<T>(s, dev, Np, restg, restg - NB, g * NB, Ar, Ai, kg, k0, k0, kg + NB, kg, kg + NB, prof, 0, umax,
                        AS, BS, 0, restg - NB, force_cfg);
        }
    }
    if (!spec) hipLaunchKernelGGL(k_flags_any, dim3(1), dim3(256), 0, s, npanel, flag);
    hipLaunchKernelGGL(k_trsm_pcoef<T>, dim3(npanel), dim3(256), 0, s, Np, Ar, Ai, w.Linv_r, w.Linv_i, w.Uinv_r,
                       w.Uinv_i, w.PL_r, w.PL_i, w.PU_r, w.PU_i);
}
template void launch_lu<double>(hipStream_t, hipStream_t, hipEvent_t*, int, double*, double*, const LuWork<double>&,
                                 int*, int*, LuProfile*, int, int, double);
template void launch_lu<float>(hipStream_t, hipStream_t, hipEvent_t*, int, float*, float*, const LuWork<float>&, int*,
                                int*, LuProfile*, int, int, double);

template <typename T, typename TX>
void launch_lu_solve(hipStream_t s, int Np, int nr, const T* Ar, const T* Ai, const LuWork<T>& w, const int* ipiv,
                     const int* flag_any, TX* Xr, TX* Xi, unsigned long long* counter, unsigned long long* pub,
                     const int* skip, double* X64r, double* X64i, int x64_add) {
    if (flag_any) hipLaunchKernelGGL(k_apply_ipiv<TX>, dim3(1), dim3(64), 0, s, Np, nr, ipiv, flag_any, Xr, Xi, skip);
    int nblk = Np / LU_NB;
    hipLaunchKernelGGL((k_trsm_persist<T, TX>), dim3(nblk), dim3(256), 0, s, Np, nr, 1, Ar, Ai, w.Linv_r, w.Linv_i,
                       w.PL_r, w.PL_i, Xr, Xi, pub, counter, skip, nullptr, nullptr, 0);
    hipLaunchKernelGGL((k_trsm_persist<T, TX>), dim3(nblk), dim3(256), 0, s, Np, nr, 0, Ar, Ai, w.Uinv_r, w.Uinv_i,
                       w.PU_r, w.PU_i, Xr, Xi, pub, counter, skip, sizeof(TX) == 4 ? X64r : nullptr,
                       sizeof(TX) == 4 ? X64i : nullptr, x64_add);
}
template void launch_lu_solve<double, double>(hipStream_t, int, int, const double*, const double*,
                                              const LuWork<double>&, const int*, const int*, double*, double*,
                                              unsigned long long*, unsigned long long*, const int*, double*, double*,
                                              int);
template void launch_lu_solve<float, float>(hipStream_t, int, int, const float*, const float*, const LuWork<float>&,
                                            const int*, const int*, float*, float*, unsigned long long*,
                                            unsigned long long*, const int*, double*, double*, int);

void launch_to_f32(hipStream_t s, size_t n, const double* a, const double* b, float* fa, float* fb, const int* skip,
                   RefineCtrl* reset) {
    hipLaunchKernelGGL(k_to_f32, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, b, fa, fb, skip, reset);
}
void launch_refine_reduce(hipStream_t s, int N, int Np, int nr, const double* Pr, const double* Pi, const double* Br,
                          const double* Bi, double* Rr, double* Ri, float* R32r, float* R32i, const double* Xr,
                          const double* Xi, const double* Pabs, double* anorm, RefineCtrl* ctrl, double tol,
                          int max_iter) {
    int kslab = 0;
    const int nks = matvec_slabs(Np, kslab);
    hipLaunchKernelGGL(k_refine_reduce, dim3(cdiv((size_t)nr * Np, 256)), dim3(256), 0, s, N, Np, nr, nks, Pr, Pi, Br,
                       Bi, Rr, Ri, R32r, R32i, Xr, Xi, Pabs, anorm, ctrl, tol, max_iter);
}
void launch_pack_planar(hipStream_t s, int n_rows, int ncol, int Np, int ncolp, const double* src, double* re,
                        double* im, int ident) {
    size_t n = (size_t)Np * ncolp;
    hipLaunchKernelGGL(k_pack_planar, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n_rows, ncol, Np, ncolp, src,
                       re, im, ident);
}
void launch_unpack_planar(hipStream_t s, int n_rows, int ncol, int Np, const double* re, const double* im,
                          double* dst) {
    size_t n = (size_t)n_rows * ncol;
    hipLaunchKernelGGL(k_unpack_planar, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n_rows, ncol, Np, re, im,
                       dst);
}
void launch_solve_state_init(hipStream_t s, unsigned long long* counter, unsigned long long* pub, size_t nwords) {
    size_t n = nwords > 0 ? nwords : 1;
    hipLaunchKernelGGL(k_solve_state_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, counter, pub, nwords);
}
void launch_copy2(hipStream_t s, size_t n, const double* a, const double* b, double* ca, double* cb) {
    hipLaunchKernelGGL(k_copy2, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, b, ca, cb);
}
void launch_colmax(hipStream_t s, int N, int Np, int nr, const double* rr, const double* ri, const double* xr,
                   const double* xi, double* out) {
    hipLaunchKernelGGL(k_colmax, dim3(nr), dim3(256), 0, s, N, Np, rr, ri, xr, xi, out);
}
void launch_reduce_coeffs(hipStream_t s, int N, int Np, int nb, const double* gen, const double* Xr, const double* Xi,
                          const double* PhiIr, const double* PhiIi, const double* DAr, const double* DAi,
                          double* out) {
    hipLaunchKernelGGL(k_reduce_coeffs, dim3(36 + 12 * nb), dim3(256), 0, s, N, Np, nb, gen, Xr, Xi, PhiIr, PhiIi,
                       DAr, DAi, out);
}

}  // namespace bemuse
