// Post-processing of a solved frequency: Kochin functions of the radiated and
// diffracted waves, and the radiated/diffracted potentials (as free-surface
// elevation) at exterior field points. Both act on the potentials a solve
// returned (N x (6 + nb) complex, planar, leading dimension N).
#include "bem_device.hpp"
#include "bem_kernels.h"

namespace bemuse {
using namespace dev;

namespace {

constexpr double kGravity = 9.81;   // the constant Gravity of the reference (BEMUse_Math_Types.h:115)

// dPhi_I/dn at a centroid without the area factor (Calc_Phi_Incident,
// Hydrodynamic_Solver.cpp:544-573; Set_Incident_Potential_Mats :575-596):
// Phi_I = i (g / Omega) e^{K z} e^{i K (x cos b + y sin b)},
// dPhi_I/dn = n . (i K cos b, i K sin b, K) Phi_I.
__device__ inline void dphi_incident_dn(double x, double y, double z, double nx, double ny, double nz, double beta,
                                        double omega, double kappa, double& dr, double& di) {
    const double CB = cos(beta), SB = sin(beta);
    const double ph = kappa * (x * CB + y * SB);
    const double c0 = kGravity / omega, Z = exp(kappa * z);
    const double pr = -c0 * sin(ph) * Z, pi = c0 * cos(ph) * Z;
    const double ar = nz * kappa, ai = (nx * CB + ny * SB) * kappa;
    dr = ar * pr - ai * pi;
    di = ar * pi + ai * pr;
}

// Kochin functions (Set_Kochin_Mats, Hydrodynamic_Solver.cpp:598-642, and
// :492-494), one work-group per angle theta_i = (i + 1/2) 2 pi / nkoch:
//   H_c(theta) = sum_P A_P (E_P q_{P,c} - (n_P . grad E_P) phi_{P,c}),
//   E = exp(K z + i K (x cos theta + y sin theta)) at the centroid,
// with (q, phi) = (N_k, Phi_J) for the 6 radiation columns and
// (dPhi_I/dn, Phi_S) for the nb diffraction columns. out: nkoch x (6 + nb)
// complex interleaved.
__global__ __launch_bounds__(256) void k_kochin(int N, int nkoch, int nb, const double* __restrict__ body,
                                                const double* __restrict__ gen, const double* __restrict__ betas,
                                                double omega, double kappa, const double* __restrict__ phr,
                                                const double* __restrict__ phi, double* __restrict__ out) {
    constexpr int MC = 16;
    __shared__ double red[2 * MC][256];
    const int it = blockIdx.x, tid = threadIdx.x, nc = 6 + nb;
    const double th = (0.5 + it) * 2.0 * kPI / nkoch;
    const double CT = cos(th), ST = sin(th);
    double ar[MC], ai[MC];
#pragma unroll
    for (int c = 0; c < MC; c++) { ar[c] = 0.0; ai[c] = 0.0; }
    for (int P = tid; P < N; P += 256) {
        const double x = body[(size_t)P * 32], y = body[(size_t)P * 32 + 1], z = body[(size_t)P * 32 + 2];
        const double* g = gen + (size_t)P * 8;
        const double nx = g[0], ny = g[1], nz = g[2], A = g[6];
        const double mag = exp(kappa * z), ph = kappa * (x * CT + y * ST);
        const double er = mag * cos(ph), ei = mag * sin(ph);
        // n . grad E = (n_z K + i K (n_x cos + n_y sin)) E
        const double gr = nz * kappa, gi = (nx * CT + ny * ST) * kappa;
        const double dr = gr * er - gi * ei, di = gr * ei + gi * er;
#pragma unroll
        for (int c = 0; c < MC; c++) {
            if (c < nc) {
                double qr, qi;
                if (c < 6) { qr = g[c]; qi = 0.0; }
                else dphi_incident_dn(x, y, z, nx, ny, nz, betas[c - 6], omega, kappa, qr, qi);
                const double fr = phr[(size_t)c * N + P], fi = phi[(size_t)c * N + P];
                ar[c] += A * ((er * qr - ei * qi) - (dr * fr - di * fi));
                ai[c] += A * ((er * qi + ei * qr) - (dr * fi + di * fr));
            }
        }
    }
#pragma unroll
    for (int c = 0; c < MC; c++) { red[2 * c][tid] = ar[c]; red[2 * c + 1][tid] = ai[c]; }
    __syncthreads();
    if (tid < 2 * nc) {
        double s = 0.0;
        for (int k = 0; k < 256; k++) s += red[tid][k];
        out[2 * ((size_t)it * nc + tid / 2) + (tid & 1)] = s;
    }
}

constexpr int FS_TS = 16;   // panels staged per step of the field kernel

// Potential at exterior field points (Prepare_FS_Linear_System :205-237,
// Prepare_FS_Linear_System_Wave_Terms :239-266, and :499-505), one field
// point per thread:
//   WS = -s - s_r - ws, WD = d + d_r + wd - s_r (2 n_z) K   (panel S, point R)
//   Rad_j  = -i Omega / g sum_S (WS N_k(S, j) - WD Phi_J(S, j))
//   Diff_b = -i Omega / g sum_S (WS dPhi_I/dn(S, b) + WD Phi_S(S, b))
// out: npts x (6 + nb) complex interleaved.
__global__ __launch_bounds__(256) void k_fs_field(int N, int npts, const double* __restrict__ pts, int nb,
                                                  const double* __restrict__ betas, double omega, double kappa,
                                                  const double* __restrict__ body, const double* __restrict__ refl,
                                                  const double* __restrict__ quad, const double* __restrict__ gen,
                                                  const double* __restrict__ phr, const double* __restrict__ phi,
                                                  double* __restrict__ out) {
    constexpr int MC = 16;
    __shared__ double sb[FS_TS][32], sr[FS_TS][32], sq[FS_TS][32];
    __shared__ double sqr[FS_TS][MC], sqi[FS_TS][MC], sfr[FS_TS][MC], sfi[FS_TS][MC];
    const int tid = threadIdx.x, nc = 6 + nb;
    const int R = blockIdx.x * 256 + tid;
    const bool live = R < npts;
    const double px = live ? pts[3 * (size_t)R] : 0.0, py = live ? pts[3 * (size_t)R + 1] : 0.0,
                 pz = live ? pts[3 * (size_t)R + 2] : -1.0;
    double ar[MC], ai[MC];
#pragma unroll
    for (int c = 0; c < MC; c++) { ar[c] = 0.0; ai[c] = 0.0; }
    for (int s0 = 0; s0 < N; s0 += FS_TS) {
        __syncthreads();
        for (int k = tid; k < FS_TS * 32; k += 256) {
            const int s = s0 + k / 32;
            const bool ok = s < N;
            sb[k / 32][k % 32] = ok ? body[(size_t)s * 32 + k % 32] : 0.0;
            sr[k / 32][k % 32] = ok ? refl[(size_t)s * 32 + k % 32] : 0.0;
            sq[k / 32][k % 32] = ok ? quad[(size_t)s * 32 + k % 32] : 0.0;
        }
        for (int k = tid; k < FS_TS * MC; k += 256) {
            const int sl = k / MC, c = k % MC, s = s0 + sl;
            double qr = 0.0, qi = 0.0, fr = 0.0, fi = 0.0;
            if (s < N && c < nc) {
                const double* g = gen + (size_t)s * 8;
                if (c < 6) qr = g[c];
                else dphi_incident_dn(body[(size_t)s * 32], body[(size_t)s * 32 + 1], body[(size_t)s * 32 + 2],
                                      g[0], g[1], g[2], betas[c - 6], omega, kappa, qr, qi);
                fr = phr[(size_t)c * N + s];
                fi = phi[(size_t)c * N + s];
            }
            sqr[sl][c] = qr; sqi[sl][c] = qi; sfr[sl][c] = fr; sfi[sl][c] = fi;
        }
        __syncthreads();
        if (!live) continue;
        const int ns = min(FS_TS, N - s0);
        for (int sl = 0; sl < ns; sl++) {
            double s, d, srf, drf, wsr, wsi, wdr, wdi;
            rankine_panel(sb[sl], px, py, pz, s, d);
            rankine_panel(sr[sl], px, py, pz, srf, drf);
            wave_panel(sq[sl], px, py, pz, kappa, wsr, wsi, wdr, wdi);
            const double WSr = -s - srf - wsr, WSi = -wsi;
            const double WDr = d + drf + wdr - srf * sq[sl][29] * kappa, WDi = wdi;
#pragma unroll
            for (int c = 0; c < MC; c++) {
                if (c < nc) {
                    const double qr = sqr[sl][c], qi = sqi[sl][c], fr = sfr[sl][c], fi = sfi[sl][c];
                    const double tr = WSr * qr - WSi * qi, ti = WSr * qi + WSi * qr;
                    const double ur = WDr * fr - WDi * fi, ui = WDr * fi + WDi * fr;
                    if (c < 6) { ar[c] += tr - ur; ai[c] += ti - ui; }
                    else { ar[c] += tr + ur; ai[c] += ti + ui; }
                }
            }
        }
    }
    if (!live) return;
    const double f = omega / kGravity;   // -i (Omega/g) (a + i b) = (Omega/g) (b - i a)
    for (int c = 0; c < nc; c++) {
        out[2 * ((size_t)R * nc + c)] = f * ai[c];
        out[2 * ((size_t)R * nc + c) + 1] = -f * ar[c];
    }
}

inline int cdivp(int a, int b) { return (a + b - 1) / b; }

}  // namespace

void launch_kochin(hipStream_t s, int N, int nkoch, int nb, const double* body, const double* gen,
                   const double* betas, double omega, double kappa, const double* phr, const double* phi,
                   double* out) {
    hipLaunchKernelGGL(k_kochin, dim3(nkoch), dim3(256), 0, s, N, nkoch, nb, body, gen, betas, omega, kappa, phr, phi,
                       out);
}

void launch_fs_field(hipStream_t s, int N, int npts, const double* pts, int nb, const double* betas, double omega,
                     double kappa, const double* body, const double* refl, const double* quad, const double* gen,
                     const double* phr, const double* phi, double* out) {
    hipLaunchKernelGGL(k_fs_field, dim3(cdivp(npts, 256)), dim3(256), 0, s, N, npts, pts, nb, betas, omega, kappa,
                       body, refl, quad, gen, phr, phi, out);
}

}  // namespace bemuse
