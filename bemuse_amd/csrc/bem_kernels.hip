// HIP kernels of the per-frequency path for gfx950 (MI355X).
//
// Matrices live in HBM as planar complex (separate re / im arrays),
// column-major, padded to Np = roundup(N, 128) rows and columns. The padding
// block of the system matrix is the identity and of every other operand zero,
// so no GEMM / TRSM kernel needs edge handling and the padded solve equals the
// unpadded one.
#include "bem_kernels.h"
#include "bem_device.hpp"

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
__global__ __launch_bounds__(256) void k_wave_compose(int N, int Np, double K, int compose,
                                                      const double* __restrict__ quad, const double* __restrict__ recv,
                                                      const double* __restrict__ Ssrc, const double* __restrict__ Srefl,
                                                      const double* __restrict__ Dsrc, const double* __restrict__ Drefl,
                                                      double* Sr, double* Si, double* Dr, double* Di,
                                                      float* D32r, float* D32i) {
    __shared__ double sq[AS_TS][32];
    int r0 = blockIdx.x * AS_TR, s0 = blockIdx.y * AS_TS;
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
                wave_panel(sq[sl], px, py, pz, K, wsr, wsi, wdr, wdi);
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

__global__ void k_havelock_eval(int n, const double* X, const double* Y, const double* V, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    GFVal g = havelock_gf(X[i], Y[i], V[i]);
    double cx = 0.0, cy = 0.0;
    if (g.h > 1.0e-6) { cx = X[i] / g.h; cy = Y[i] / g.h; }
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
                           const double* __restrict__ betas, double omega, double kappa,
                           double* Xr, double* Xi, int ldx, double* PhiIr, double* PhiIi, double* DAr, double* DAi) {
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
// Tall-skinny products: Y = op(M X) for an Np x Np complex matrix and nr <= 16
// columns, split over K slabs; partials are reduced in a fixed order.
// ---------------------------------------------------------------------------
constexpr int MV_ROWS = 64, MV_MAXR = 16;

template <typename TM>
__global__ __launch_bounds__(256) void k_matvec_partial(int Np, int nr, int kslab, const TM* __restrict__ Mr,
                                                        const TM* __restrict__ Mi, const double* __restrict__ Xr,
                                                        const double* __restrict__ Xi, double* Pr, double* Pi) {
    int r0 = blockIdx.x * MV_ROWS, ks = blockIdx.y;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int row = r0 + lane;
    double ar[MV_MAXR], ai[MV_MAXR];
#pragma unroll
    for (int j = 0; j < MV_MAXR; j++) { ar[j] = 0; ai[j] = 0; }
    int kb = ks * kslab, ke = min(Np, kb + kslab);
    for (int k = kb + wave; k < ke; k += 4) {
        double mr = (double)Mr[(size_t)k * Np + row], mi = (double)Mi[(size_t)k * Np + row];
#pragma unroll
        for (int j = 0; j < MV_MAXR; j++) {
            if (j < nr) {
                double xr = Xr[(size_t)j * Np + k], xi = Xi[(size_t)j * Np + k];
                ar[j] += mr * xr - mi * xi;
                ai[j] += mr * xi + mi * xr;
            }
        }
    }
    __shared__ double red[4][MV_MAXR][2][64];
#pragma unroll
    for (int j = 0; j < MV_MAXR; j++) { red[wave][j][0][lane] = ar[j]; red[wave][j][1][lane] = ai[j]; }
    __syncthreads();
    for (int t = threadIdx.x; t < nr * 64; t += 256) {
        int j = t / 64, l = t % 64;
        double sr = red[0][j][0][l] + red[1][j][0][l] + red[2][j][0][l] + red[3][j][0][l];
        double si = red[0][j][1][l] + red[1][j][1][l] + red[2][j][1][l] + red[3][j][1][l];
        size_t o = ((size_t)ks * MV_MAXR + j) * Np + r0 + l;
        Pr[o] = sr; Pi[o] = si;
    }
}

// Y = sum_ks P (mode 0) or Y = B - sum_ks P (mode 1).
__global__ void k_matvec_reduce(int Np, int nr, int nks, const double* __restrict__ Pr, const double* __restrict__ Pi,
                                const double* Br, const double* Bi, int mode, double* Yr, double* Yi) {
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
// interchanges (k_lu_panel_gepp), and the interchanges are applied to the
// other columns (k_laswp). Results therefore follow partial pivoting.
// ---------------------------------------------------------------------------
constexpr int LU_NB = 64;

template <typename T>
__device__ inline T cabs2(T r, T i) { return r * r + i * i; }

// Factor the NB x NB diagonal block at (k0, k0) without interchanges, in LDS.
// Sets flag[panel] bit 1 if a multiplier exceeds 1 in modulus, bit 2 on a zero
// pivot. Saves the block to `save` (NB x NB planar, ld NB).
template <typename T>
__global__ __launch_bounds__(256) void k_lu_diag(int Np, int k0, T* Ar, T* Ai, T* save_r, T* save_i, int* flag) {
    __shared__ T lr[LU_NB][LU_NB + 1], li[LU_NB][LU_NB + 1];   // [col][row]
    __shared__ int bad;
    int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    for (int t = tid; t < LU_NB * LU_NB; t += 256) {
        int c = t / LU_NB, r = t % LU_NB;
        size_t g = (size_t)(k0 + c) * Np + k0 + r;
        T vr = Ar[g], vi = Ai[g];
        lr[c][r] = vr; li[c][r] = vi;
        save_r[(size_t)c * LU_NB + r] = vr; save_i[(size_t)c * LU_NB + r] = vi;
    }
    __syncthreads();
    for (int j = 0; j < LU_NB; j++) {
        T pr = lr[j][j], pi = li[j][j];
        T den = cabs2(pr, pi);
        // multipliers l_ij = a_ij / a_jj
        if (tid > j && tid < LU_NB) {
            if (den == (T)0) {
                atomicOr(&bad, 3);
            } else {
                T ar = lr[j][tid], ai = li[j][tid];
                T mr = (ar * pr + ai * pi) / den, mi = (ai * pr - ar * pi) / den;
                lr[j][tid] = mr; li[j][tid] = mi;
                if (cabs2(mr, mi) > (T)1) atomicOr(&bad, 1);
            }
        }
        __syncthreads();
        // rank-1 update of the trailing block
        int m = LU_NB - j - 1;
        for (int t = tid; t < m * m; t += 256) {
            int c = j + 1 + t / m, r = j + 1 + t % m;
            T ur = lr[c][j], ui = li[c][j];
            T mr = lr[j][r], mi = li[j][r];
            lr[c][r] -= mr * ur - mi * ui;
            li[c][r] -= mr * ui + mi * ur;
        }
        __syncthreads();
    }
    for (int t = tid; t < LU_NB * LU_NB; t += 256) {
        int c = t / LU_NB, r = t % LU_NB;
        size_t g = (size_t)(k0 + c) * Np + k0 + r;
        Ar[g] = lr[c][r]; Ai[g] = li[c][r];
    }
    if (tid == 0 && bad) atomicOr(flag + k0 / LU_NB, bad);
}

// L21 = A21 U11^-1 for rows k0+NB .. Np-1 (row-parallel: one wave per row
// group of 64, lanes over columns). Saves A21 to `save` before overwriting
// (rows relative to k0+NB, ld = Np). Flags multipliers with |l| > 1.
template <typename T>
__global__ __launch_bounds__(256) void k_lu_l21(int Np, int k0, T* Ar, T* Ai, T* save_r, T* save_i, int* flag) {
    __shared__ T ur[LU_NB][LU_NB + 1], ui[LU_NB][LU_NB + 1];   // U11 [row][col]
    __shared__ T xr[LU_NB][LU_NB + 1], xi[LU_NB][LU_NB + 1];   // tile rows [row][col]
    __shared__ int bad;
    int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    int row0 = k0 + LU_NB + blockIdx.x * LU_NB;
    for (int t = tid; t < LU_NB * LU_NB; t += 256) {
        int c = t / LU_NB, r = t % LU_NB;
        size_t g = (size_t)(k0 + c) * Np + k0 + r;
        ur[r][c] = (r <= c) ? Ar[g] : (T)0;
        ui[r][c] = (r <= c) ? Ai[g] : (T)0;
        size_t a = (size_t)(k0 + c) * Np + row0 + r;
        T vr = Ar[a], vi = Ai[a];
        xr[r][c] = vr; xi[r][c] = vi;
        save_r[(size_t)c * Np + row0 - k0 - LU_NB + r] = vr;
        save_i[(size_t)c * Np + row0 - k0 - LU_NB + r] = vi;
    }
    __syncthreads();
    // each wave solves 16 rows: x U = a, lanes over columns
    int lane = tid & 63, wave = tid >> 6;
    for (int rr = 0; rr < LU_NB / 4; rr++) {
        int r = wave * (LU_NB / 4) + rr;
        T ar = xr[r][lane], ai = xi[r][lane];
        for (int j = 0; j < LU_NB; j++) {
            // lane j finalises x_j = a_j / u_jj, then broadcast
            T pr = ur[j][j], pi = ui[j][j];
            T den = cabs2(pr, pi);
            T vr = (ar * pr + ai * pi) / den, vi = (ai * pr - ar * pi) / den;
            T br = __shfl(vr, j), bi = __shfl(vi, j);
            if (lane == j) { ar = br; ai = bi; if (cabs2(br, bi) > (T)1) bad = 1; }
            if (lane > j) {
                T u_r = ur[j][lane], u_i = ui[j][lane];
                ar -= br * u_r - bi * u_i;
                ai -= br * u_i + bi * u_r;
            }
        }
        xr[r][lane] = ar; xi[r][lane] = ai;
    }
    __syncthreads();
    for (int t = tid; t < LU_NB * LU_NB; t += 256) {
        int c = t / LU_NB, r = t % LU_NB;
        size_t a = (size_t)(k0 + c) * Np + row0 + r;
        Ar[a] = xr[r][c]; Ai[a] = xi[r][c];
    }
    if (tid == 0 && bad) atomicOr(flag + k0 / LU_NB, 1);
}

// Fallback for a rejected panel: restore it and factor it with partial
// pivoting in one work-group, streaming through global memory. ipiv receives
// absolute row indices. Runs only if flag[panel] is set.
template <typename T>
__global__ __launch_bounds__(1024) void k_lu_panel_gepp(int Np, int k0, T* Ar, T* Ai, const T* sd_r, const T* sd_i,
                                                       const T* s21_r, const T* s21_i, int* flag, int* ipiv) {
    int f = flag[k0 / LU_NB];
    if (!(f & 1)) return;
    int tid = threadIdx.x;
    // restore
    for (int t = tid; t < LU_NB * LU_NB; t += 1024) {
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
        if (tid == 0) { piv_s = bi[0]; ipiv[col] = bi[0]; }
        __syncthreads();
        int p = piv_s;
        // swap rows col <-> p within the panel columns
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
        if (den == (T)0) { if (tid == 0) flag[k0 / LU_NB] |= 2; continue; }
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
}

// Apply the panel's interchanges to all columns outside the panel
// (sequentially in pivot order, parallel over columns). Only if flagged.
template <typename T>
__global__ void k_laswp(int Np, int k0, T* Ar, T* Ai, const int* flag, const int* ipiv) {
    if (!(flag[k0 / LU_NB] & 1)) return;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= Np || (c >= k0 && c < k0 + LU_NB)) return;
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

// U12 = L11^-1 A12 for columns k0+NB .. Np-1: one wave per column, lanes over
// the NB rows, forward substitution with unit-lower L11 held in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_lu_u12(int Np, int k0, T* Ar, T* Ai) {
    __shared__ T lr[LU_NB][LU_NB + 1], li[LU_NB][LU_NB + 1];   // L11 [col][row]
    int tid = threadIdx.x;
    for (int t = tid; t < LU_NB * LU_NB; t += 256) {
        int c = t / LU_NB, r = t % LU_NB;
        size_t g = (size_t)(k0 + c) * Np + k0 + r;
        lr[c][r] = Ar[g]; li[c][r] = Ai[g];
    }
    __syncthreads();
    int lane = tid & 63, wave = tid >> 6;
    int col = k0 + LU_NB + blockIdx.x * 16 + wave * 4;
    for (int cc = 0; cc < 4; cc++) {
        int c = col + cc;
        if (c >= Np) break;
        size_t g = (size_t)c * Np + k0 + lane;
        T ar = Ar[g], ai = Ai[g];
        for (int j = 0; j < LU_NB - 1; j++) {
            T br = __shfl(ar, j), bi = __shfl(ai, j);
            if (lane > j) {
                T mr = lr[j][lane], mi = li[j][lane];
                ar -= mr * br - mi * bi;
                ai -= mr * bi + mi * br;
            }
        }
        Ar[g] = ar; Ai[g] = ai;
    }
}

// ---------------------------------------------------------------------------
// Complex GEMM on MFMA: C -= A B (LU trailing update) for planar complex
// operands in one column-major array with leading dimension ld.
// A: M x K at (ra, ca), B: K x N at (rb, cb), C: M x N at (rc, cc); M, N
// multiples of 64, K multiple of 16.
//
// MFMA 16x16x4 operand layout (lane l = i + 16 g): A supplies A[i][k(g)],
// B supplies B[k(g)][j]. The accumulator register r of lane l holds D[row][j]
// with row = 4g + r for the f32 instruction and row = g + 4r for the f64
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

template <typename T>
__global__ __launch_bounds__(256) void k_cgemm_sub(int M, int N, int K, T* __restrict__ Pr, T* __restrict__ Pi, int ld,
                                                   int ra, int ca, int rb, int cb, int rc, int cc) {
    typedef MfmaT<T> MT;
    typedef typename MT::acc acc_t;
    constexpr int WM = MT::WM, WN = MT::WN;
    constexpr int TM = 32 * WM, TN = 32 * WN;      // work-group tile
    int tiles_m = M / TM + (M % TM ? 1 : 0);
    int tiles_n = N / TN + (N % TN ? 1 : 0);
    int ntiles = tiles_m * tiles_n;
    // XCD-aware order: work-groups are dealt round-robin to the 8 XCDs, so
    // give each XCD a contiguous run of tiles (row-major), sharing A rows in L2.
    int bid = blockIdx.x;
    int per = (ntiles + 7) / 8;
    int xcd = bid % 8, loc = bid / 8;
    int tile = xcd * per + loc;
    if (tile >= ntiles) return;
    int tm = tile / tiles_n, tn = tile % tiles_n;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int wm = wave >> 1, wn = wave & 1;
    int row0 = tm * TM + wm * 16 * WM, col0 = tn * TN + wn * 16 * WN;
    if (row0 >= M || col0 >= N) return;
    int i = lane & 15, g = lane >> 4;
    acc_t cr[WM][WN], ci[WM][WN];
#pragma unroll
    for (int a = 0; a < WM; a++)
#pragma unroll
        for (int b = 0; b < WN; b++) { cr[a][b] = acc_t{0, 0, 0, 0}; ci[a][b] = acc_t{0, 0, 0, 0}; }
    const size_t ldl = (size_t)ld;
    for (int k0 = 0; k0 < K; k0 += 16) {
        T br[WN][4], bi[WN][4];
#pragma unroll
        for (int b = 0; b < WN; b++) {
            size_t off = (size_t)(cb + col0 + 16 * b + i) * ldl + rb + k0 + 4 * g;
#pragma unroll
            for (int t = 0; t < 4; t++) { br[b][t] = Pr[off + t]; bi[b][t] = Pi[off + t]; }
        }
#pragma unroll
        for (int t = 0; t < 4; t++) {
            T ar[WM], ai[WM];
            size_t offa = (size_t)(ca + k0 + 4 * g + t) * ldl + ra + row0 + i;
#pragma unroll
            for (int a = 0; a < WM; a++) { ar[a] = Pr[offa + 16 * a]; ai[a] = Pi[offa + 16 * a]; }
#pragma unroll
            for (int a = 0; a < WM; a++)
#pragma unroll
                for (int b = 0; b < WN; b++) {
                    cr[a][b] = MT::mma(ar[a], br[b][t], cr[a][b]);
                    cr[a][b] = MT::mma(-ai[a], bi[b][t], cr[a][b]);
                    ci[a][b] = MT::mma(ar[a], bi[b][t], ci[a][b]);
                    ci[a][b] = MT::mma(ai[a], br[b][t], ci[a][b]);
                }
        }
    }
#pragma unroll
    for (int a = 0; a < WM; a++)
#pragma unroll
        for (int b = 0; b < WN; b++) {
            size_t off = (size_t)(cc + col0 + 16 * b + i) * ldl + rc + row0 + 16 * a;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int rr = MT::row(g, r);
                Pr[off + rr] -= cr[a][b][r];
                Pi[off + rr] -= ci[a][b][r];
            }
        }
}

// ---------------------------------------------------------------------------
// Triangular solves with NB-row blocks, one persistent launch per solve.
// Work-groups take blocks in dependency order from a ticket counter; block b
// waits for the published results of the blocks it depends on (epoch flags),
// accumulates their contributions, solves its diagonal block and publishes.
// lower = 1: L (unit diagonal) forward; lower = 0: U backward.
// X is Np x nr planar (ld Np), type TX; the factor has type T.
// ---------------------------------------------------------------------------
template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_trsm_persist(int Np, int nr, int lower, const T* __restrict__ Ar,
                                                      const T* __restrict__ Ai, TX* Xr, TX* Xi, unsigned* counter,
                                                      unsigned* flags, unsigned epoch) {
    __shared__ TX accr[MV_MAXR][LU_NB], acci[MV_MAXR][LU_NB];
    __shared__ T mr[LU_NB][LU_NB + 1], mi[LU_NB][LU_NB + 1];   // [col][row]
    __shared__ unsigned ticket;
    int tid = threadIdx.x;
    int nblk = Np / LU_NB;
    if (tid == 0) {
        ticket = atomicAdd(counter, 1u);
        if (ticket == (unsigned)nblk - 1) atomicExch(counter, 0u);
    }
    __syncthreads();
    int b = lower ? (int)ticket : nblk - 1 - (int)ticket;
    int r0 = b * LU_NB;
    for (int t = tid; t < nr * LU_NB; t += 256) {
        int j = t / LU_NB, r = t % LU_NB;
        accr[j][r] = Xr[(size_t)j * Np + r0 + r];
        acci[j][r] = Xi[(size_t)j * Np + r0 + r];
    }
    int ndep = lower ? b : nblk - 1 - b;
    for (int d = 0; d < ndep; d++) {
        int c = lower ? d : nblk - 1 - d;
        int c0 = c * LU_NB;
        // off-diagonal block M(r0.., c0..)
        for (int t = tid; t < LU_NB * LU_NB; t += 256) {
            int cc = t / LU_NB, r = t % LU_NB;
            size_t g = (size_t)(c0 + cc) * Np + r0 + r;
            mr[cc][r] = Ar[g]; mi[cc][r] = Ai[g];
        }
        if (tid == 0) {
            while (__hip_atomic_load(flags + c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch)
                __builtin_amdgcn_s_sleep(1);
        }
        __syncthreads();
        // acc -= M X_c (X_c read from global after the acquire)
        for (int t = tid; t < nr * LU_NB; t += 256) {
            int j = t / LU_NB, r = t % LU_NB;
            TX sr = 0, si = 0;
            const TX* xr = Xr + (size_t)j * Np + c0;
            const TX* xi = Xi + (size_t)j * Np + c0;
            for (int k = 0; k < LU_NB; k++) {
                TX a_r = (TX)mr[k][r], a_i = (TX)mi[k][r];
                TX vr = __builtin_nontemporal_load(xr + k), vi = __builtin_nontemporal_load(xi + k);
                sr += a_r * vr - a_i * vi;
                si += a_r * vi + a_i * vr;
            }
            accr[j][r] -= sr; acci[j][r] -= si;
        }
        __syncthreads();
    }
    // diagonal block
    for (int t = tid; t < LU_NB * LU_NB; t += 256) {
        int cc = t / LU_NB, r = t % LU_NB;
        size_t g = (size_t)(r0 + cc) * Np + r0 + r;
        mr[cc][r] = Ar[g]; mi[cc][r] = Ai[g];
    }
    __syncthreads();
    // substitution: one wave per right-hand side column, lanes over rows
    int lane = tid & 63, wave = tid >> 6;
    for (int j = wave; j < nr; j += 4) {
        TX vr = accr[j][lane], vi = acci[j][lane];
        if (lower) {
            for (int k = 0; k < LU_NB - 1; k++) {
                TX br = __shfl(vr, k), bi = __shfl(vi, k);
                if (lane > k) {
                    TX a_r = (TX)mr[k][lane], a_i = (TX)mi[k][lane];
                    vr -= a_r * br - a_i * bi;
                    vi -= a_r * bi + a_i * br;
                }
            }
        } else {
            for (int k = LU_NB - 1; k >= 0; k--) {
                TX pr = (TX)mr[k][k], pi = (TX)mi[k][k];
                TX den = pr * pr + pi * pi;
                TX qr = (vr * pr + vi * pi) / den, qi = (vi * pr - vr * pi) / den;
                TX br = __shfl(qr, k), bi = __shfl(qi, k);
                if (lane == k) { vr = br; vi = bi; }
                if (lane < k) {
                    TX a_r = (TX)mr[k][lane], a_i = (TX)mi[k][lane];
                    vr -= a_r * br - a_i * bi;
                    vi -= a_r * bi + a_i * br;
                }
            }
        }
        Xr[(size_t)j * Np + r0 + lane] = vr;
        Xi[(size_t)j * Np + r0 + lane] = vi;
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        __hip_atomic_store(flags + b, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Row permutation of the right-hand sides: X(perm order) for the
// interchanges recorded in ipiv (applied sequentially, one thread per column).
template <typename TX>
__global__ void k_apply_ipiv(int Np, int nr, const int* ipiv, const int* any_flag, TX* Xr, TX* Xi) {
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

__global__ void k_init_lu(int Np, int npanel, int* flag, int* ipiv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Np) ipiv[i] = i;
    if (i <= npanel) flag[i] = 0;
}

// ---------------------------------------------------------------------------
// Refinement helpers
// ---------------------------------------------------------------------------
__global__ void k_to_f32(size_t n, const double* a, const double* b, float* fa, float* fb) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { fa[i] = (float)a[i]; fb[i] = (float)b[i]; }
}
__global__ void k_add_f32(size_t n, double* a, double* b, const float* fa, const float* fb) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { a[i] += (double)fa[i]; b[i] += (double)fb[i]; }
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

// Max row sum of |A| (infinity norm), rows < N: per-row sums then a max.
__global__ void k_rowabs(int N, int Np, const double* Ar, const double* Ai, double* rowsum) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    double s = 0;
    for (int c = 0; c < N; c++) { size_t o = (size_t)c * Np + r; s += sqrt(Ar[o] * Ar[o] + Ai[o] * Ai[o]); }
    rowsum[r] = s;
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

void launch_wave_compose(hipStream_t s, int N, int Np, double K, int compose, const double* quad, const double* recv,
                         const double* Ssrc, const double* Srefl, const double* Dsrc, const double* Drefl,
                         double* Sr, double* Si, double* Dr, double* Di, float* D32r, float* D32i) {
    dim3 grid(Np / AS_TR, Np / AS_TS);
    hipLaunchKernelGGL(k_wave_compose, grid, dim3(256), 0, s, N, Np, K, compose, quad, recv, Ssrc, Srefl, Dsrc, Drefl,
                       Sr, Si, Dr, Di, D32r, D32i);
}

void launch_havelock_eval(hipStream_t s, int n, const double* X, const double* Y, const double* V, double* out) {
    hipLaunchKernelGGL(k_havelock_eval, dim3(cdiv(n, 256)), dim3(256), 0, s, n, X, Y, V, out);
}

void launch_incident(hipStream_t s, int N, int Np, int nb, const double* body, const double* gen, const double* betas,
                     double omega, double kappa, double* Xr, double* Xi, int ldx, double* PhiIr, double* PhiIi,
                     double* DAr, double* DAi) {
    hipLaunchKernelGGL(k_incident, dim3(cdiv(Np, 256)), dim3(256), 0, s, N, Np, nb, body, gen, betas, omega, kappa,
                       Xr, Xi, ldx, PhiIr, PhiIi, DAr, DAi);
}

int matvec_kslabs(int Np) { return Np >= 4096 ? 16 : (Np >= 1024 ? 8 : 4); }

template <typename TM>
void launch_matvec(hipStream_t s, int Np, int nr, const TM* Mr, const TM* Mi, const double* Xr, const double* Xi,
                   double* Pr, double* Pi, const double* Br, const double* Bi, int mode, double* Yr, double* Yi) {
    int nks = matvec_kslabs(Np);
    int kslab = (Np / nks + 63) / 64 * 64;
    nks = (Np + kslab - 1) / kslab;
    dim3 grid(Np / MV_ROWS, nks);
    hipLaunchKernelGGL(k_matvec_partial<TM>, grid, dim3(256), 0, s, Np, nr, kslab, Mr, Mi, Xr, Xi, Pr, Pi);
    hipLaunchKernelGGL(k_matvec_reduce, dim3(cdiv((size_t)nr * Np, 256)), dim3(256), 0, s, Np, nr, nks, Pr, Pi, Br, Bi,
                       mode, Yr, Yi);
}
template void launch_matvec<double>(hipStream_t, int, int, const double*, const double*, const double*, const double*,
                                    double*, double*, const double*, const double*, int, double*, double*);

template <typename T>
void launch_lu(hipStream_t s, int Np, T* Ar, T* Ai, T* sd_r, T* sd_i, T* s21_r, T* s21_i, int* flag, int* ipiv,
               LuProfile* prof) {
    int npanel = Np / LU_NB;
    hipLaunchKernelGGL(k_init_lu, dim3(cdiv(Np + 1, 256)), dim3(256), 0, s, Np, npanel, flag, ipiv);
    for (int p = 0; p < npanel; p++) {
        int k0 = p * LU_NB;
        int rest = Np - k0 - LU_NB;
        hipLaunchKernelGGL(k_lu_diag<T>, dim3(1), dim3(256), 0, s, Np, k0, Ar, Ai, sd_r, sd_i, flag);
        if (rest > 0)
            hipLaunchKernelGGL(k_lu_l21<T>, dim3(rest / LU_NB), dim3(256), 0, s, Np, k0, Ar, Ai, s21_r, s21_i, flag);
        hipLaunchKernelGGL(k_lu_panel_gepp<T>, dim3(1), dim3(1024), 0, s, Np, k0, Ar, Ai, sd_r, sd_i, s21_r, s21_i,
                           flag, ipiv);
        hipLaunchKernelGGL(k_laswp<T>, dim3(cdiv(Np, 256)), dim3(256), 0, s, Np, k0, Ar, Ai, flag, ipiv);
        if (rest > 0) {
            hipLaunchKernelGGL(k_lu_u12<T>, dim3(cdiv(rest, 16)), dim3(256), 0, s, Np, k0, Ar, Ai);
            constexpr int TM = 32 * MfmaT<T>::WM, TN = 32 * MfmaT<T>::WN;
            int tiles = (int)(cdiv(rest, TM) * cdiv(rest, TN));
            int per = (tiles + 7) / 8;
            bool timed = prof && prof->n < LuProfile::kMax;
            if (timed) {
                if (!prof->ev[2 * prof->n]) {
                    (void)hipEventCreate(&prof->ev[2 * prof->n]);
                    (void)hipEventCreate(&prof->ev[2 * prof->n + 1]);
                }
                (void)hipEventRecord(prof->ev[2 * prof->n], s);
            }
            hipLaunchKernelGGL(k_cgemm_sub<T>, dim3(per * 8), dim3(256), 0, s, rest, rest, LU_NB, Ar, Ai, Np,
                               k0 + LU_NB, k0, k0, k0 + LU_NB, k0 + LU_NB, k0 + LU_NB);
            if (timed) {
                (void)hipEventRecord(prof->ev[2 * prof->n + 1], s);
                double u = (double)std::max(0, prof->N - (k0 + LU_NB));
                prof->useful_flops += 8.0 * u * u * LU_NB;
                prof->n++;
            }
        }
    }
    hipLaunchKernelGGL(k_flags_any, dim3(1), dim3(256), 0, s, npanel, flag);
}
template void launch_lu<double>(hipStream_t, int, double*, double*, double*, double*, double*, double*, int*, int*,
                                 LuProfile*);
template void launch_lu<float>(hipStream_t, int, float*, float*, float*, float*, float*, float*, int*, int*,
                                LuProfile*);

template <typename T, typename TX>
void launch_lu_solve(hipStream_t s, int Np, int nr, const T* Ar, const T* Ai, const int* ipiv, const int* flag_any,
                     TX* Xr, TX* Xi, unsigned* counter, unsigned* flags, unsigned& epoch) {
    hipLaunchKernelGGL(k_apply_ipiv<TX>, dim3(1), dim3(64), 0, s, Np, nr, ipiv, flag_any, Xr, Xi);
    int nblk = Np / LU_NB;
    ++epoch;
    hipLaunchKernelGGL((k_trsm_persist<T, TX>), dim3(nblk), dim3(256), 0, s, Np, nr, 1, Ar, Ai, Xr, Xi, counter, flags,
                       epoch);
    ++epoch;
    hipLaunchKernelGGL((k_trsm_persist<T, TX>), dim3(nblk), dim3(256), 0, s, Np, nr, 0, Ar, Ai, Xr, Xi, counter, flags,
                       epoch);
}
template void launch_lu_solve<double, double>(hipStream_t, int, int, const double*, const double*, const int*,
                                              const int*, double*, double*, unsigned*, unsigned*, unsigned&);
template void launch_lu_solve<float, float>(hipStream_t, int, int, const float*, const float*, const int*, const int*,
                                            float*, float*, unsigned*, unsigned*, unsigned&);

void launch_to_f32(hipStream_t s, size_t n, const double* a, const double* b, float* fa, float* fb) {
    hipLaunchKernelGGL(k_to_f32, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, b, fa, fb);
}
void launch_add_f32(hipStream_t s, size_t n, double* a, double* b, const float* fa, const float* fb) {
    hipLaunchKernelGGL(k_add_f32, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, b, fa, fb);
}
void launch_copy2(hipStream_t s, size_t n, const double* a, const double* b, double* ca, double* cb) {
    hipLaunchKernelGGL(k_copy2, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, b, ca, cb);
}
void launch_colmax(hipStream_t s, int N, int Np, int nr, const double* rr, const double* ri, const double* xr,
                   const double* xi, double* out) {
    hipLaunchKernelGGL(k_colmax, dim3(nr), dim3(256), 0, s, N, Np, rr, ri, xr, xi, out);
}
void launch_rowabs(hipStream_t s, int N, int Np, const double* Ar, const double* Ai, double* rowsum) {
    hipLaunchKernelGGL(k_rowabs, dim3(cdiv(N, 256)), dim3(256), 0, s, N, Np, Ar, Ai, rowsum);
}
void launch_reduce_coeffs(hipStream_t s, int N, int Np, int nb, const double* gen, const double* Xr, const double* Xi,
                          const double* PhiIr, const double* PhiIi, const double* DAr, const double* DAi,
                          double* out) {
    hipLaunchKernelGGL(k_reduce_coeffs, dim3(36 + 12 * nb), dim3(256), 0, s, N, Np, nb, gen, Xr, Xi, PhiIr, PhiIi,
                       DAr, DAi, out);
}

}  // namespace bemuse
