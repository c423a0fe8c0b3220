// Learned surrogate of the regular part of the free-surface (Havelock) Green
// function, and the kernels that sample, train, evaluate and use it.
//
// Network: 4 features -> 128 -> 128 -> 128 -> 128 -> 2, tanh on the hidden
// layers, FP32 weights. For h = K R (horizontal) and v = K (z + zeta) <= 0,
// with dd = sqrt(h^2 + v^2):
//   features  f0 = 2 log1p(h) / log1p(200) - 1, f1 = 2 log1p(-v) / log1p(50) - 1,
//             f2 = sin h, f3 = cos h
//   scale     s  = 1 / (1 + dd) + e^v / sqrt(1 + h)
//   singular  SG  = 2 e^v ln((dd - v) / 2) / (1 + dd^3)        (log term of L0)
//             SGh = 2 (h / dd + h) / ((dd - v) (1 + dd^3))      (pole term of Ls)
//   outputs   y = (Re GF - SG, Re GFh - SGh) / s
// where GF = L0 + W and GFh = Ls + Wh are HavelockGF's outputs
// (Kernels.cpp:342-412). DGFZ = GF there (:411), so (GF, GFh) is the whole
// kernel: the network's outputs are the regular parts of G and dG/dR (with
// dG/dz = G). The imaginary parts are closed forms, Im GF = -2 pi e^v J0(h)
// and Im GFh = 2 pi e^v J1(h) (:284-295, :327-338), and are evaluated
// exactly: damping at low frequency comes from their near cancellation,
// which a fitted approximation does not preserve. The trained domain is
// h in [0, 200], v in [-50, 0]; outside it the exact kernel is used.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace bemuse {

constexpr int kSurIn = 4, kSurHid = 128, kSurOut = 2;
constexpr double kSurHmax = 200.0, kSurVmin = -50.0;
// Flat parameter layout (floats). W1 is 128 x 4, W2..W4 are 128 x 128, W5 is
// 4 x 128, all row-major (output, input); each followed by its bias.
constexpr int kSurW1 = 0, kSurB1 = kSurW1 + kSurHid * kSurIn;
constexpr int kSurW2 = kSurB1 + kSurHid, kSurB2 = kSurW2 + kSurHid * kSurHid;
constexpr int kSurW3 = kSurB2 + kSurHid, kSurB3 = kSurW3 + kSurHid * kSurHid;
constexpr int kSurW4 = kSurB3 + kSurHid, kSurB4 = kSurW4 + kSurHid * kSurHid;
constexpr int kSurW5 = kSurB4 + kSurHid, kSurB5 = kSurW5 + kSurOut * kSurHid;
constexpr int kSurParams = kSurB5 + kSurOut;
constexpr int kSurTile = 32;        // samples per work-group in training
constexpr int kSurStepsPerLaunch = 500;   // Adam steps per cooperative training launch (default)
constexpr int kSurJudgeStride = 61; // one pair in ~61 is also evaluated exactly during assembly

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Uniform in (0, 1] from the top 53 bits of splitmix64(key).
__host__ __device__ inline double unit_open0(uint64_t key) {
    return (double)((splitmix64(key) >> 11) + 1) * (1.0 / 9007199254740992.0);
}

// Input features, scale and closed-form singular terms at (h, v).
struct SurPoint {
    double s, sg, sgh;
    float f[4];
};

__host__ __device__ inline SurPoint sur_point(double h, double v) {
    SurPoint p;
    double dd = sqrt(h * h + v * v);
    double ev = exp(v);
    double inv1 = 1.0 / (1.0 + dd * dd * dd);
    p.s = 1.0 / (1.0 + dd) + ev / sqrt(1.0 + h);
    p.sg = 2.0 * ev * log(0.5 * (dd - v)) * inv1;
    p.sgh = 2.0 * (h / dd + h) / (dd - v) * inv1;
    p.f[0] = (float)(2.0 * log1p(h) / log1p(kSurHmax) - 1.0);
    p.f[1] = (float)(2.0 * log1p(-v) / log1p(-kSurVmin) - 1.0);
    p.f[2] = (float)sin(h);
    p.f[3] = (float)cos(h);
    return p;
}

// Maps two uniforms in (0, 1] to a sample point. sampling 0: uniform in
// the domain, h = 200 u0, v = -50 u1. sampling 1: uniform in the log-scaled
// features, h = 201^u0 - 1, v = 1 - 51^u1 (the same density the network's
// inputs f0, f1 see; near-field pairs at low frequency lie close to the
// origin, which uniform sampling leaves almost empty).
__host__ __device__ inline void sur_sample_point(int sampling, double u0, double u1, double& h, double& v) {
    if (sampling == 1) {
        h = expm1(u0 * log1p(kSurHmax));
        v = -expm1(u1 * log1p(-kSurVmin));
    } else {
        h = kSurHmax * u0;
        v = kSurVmin * u1;
    }
}

// Device counters of one surrogate assembly: surrogate pairs, exact pairs,
// max judged relative error (float bits).
struct SurCounters { unsigned long long n_sur, n_exact; unsigned err_bits, pad; };

// Glorot-uniform initial weights, zero biases, from a splitmix64 stream.
void surrogate_init_params(uint64_t seed, std::vector<float>& p);

// Training pool: n samples (X features n x 4, T normalised targets n x 2
// floats) drawn on the device from the exact kernel, index offset `first`.
void launch_sur_sample(hipStream_t s, uint64_t seed, int sampling, long long first, long long n, float* X, float* T,
                       double* hv);
// One forward + backward over a batch: sample j of the batch is pool row
// map(p0 + j) (affine shuffle per epoch, table `perm` of 16 (a, c) pairs; perm
// NULL means row p0 + j). Writes per-work-group gradient partials
// (nwg x kSurParams) and squared-error sums (nwg).
void launch_sur_grad(hipStream_t s, const float* params, const float* X, const float* T, long long pool,
                     long long p0, int batch, const unsigned long long* perm, float* gpart, float* lpart);
int sur_grad_groups(int batch);
// Sums the partials into the gradient (grad may be NULL) and, if m is not
// NULL, takes one Adam step (lr_t includes the bias correction). The summed
// squared error is stored in loss[0].
void launch_sur_adam(hipStream_t s, int nwg, const float* gpart, const float* lpart, float* params, float* m,
                     float* v, float* grad, double* loss, float lr_t, float beta1, float beta2, float eps);
// Training steps t0 .. t0 + nsteps - 1 in one cooperative launch of nwg
// work-groups (gradient, grid barrier, Adam step, grid barrier per step;
// lr_tab[t]: step t's learning rate with the bias correction). bar: two
// counters at bar[0] and bar[32], zero before the first launch; *err is set
// if a barrier wait times out. Returns 0, or -1 if the launch was refused.
int launch_sur_train(hipStream_t s, int nwg, int t0, int nsteps, const float* lr_tab, float beta1, float beta2,
                     float eps, const float* X, const float* T, long long pool, int batch,
                     const unsigned long long* perm, float* params, float* m, float* v, float* gpart, float* lpart,
                     double* loss, unsigned* bar, int* err);
// Work-groups of k_sur_train that the device holds at once (0: no cooperative launch).
int sur_train_coop_capacity(int device);
// Network outputs y (n x 2) for features X (n x 4).
void launch_sur_forward(hipStream_t s, const float* params, long long n, const float* X, float* Y);
// Reconstructed GF, GFh (n x 4 doubles: Re GF, Im GF, Re GFh, Im GFh) at (h, v);
// exact kernel outside the domain (used[i] = 0 there, 1 inside).
void launch_sur_eval(hipStream_t s, const float* params, int n, const double* h, const double* v, double* out,
                     int* used);
// Wave assembly fused with compose (as launch_wave_compose with compose =
// wave), the regular Green function from the network inside its domain.
void launch_wave_compose_surr(hipStream_t s, const float* params, int N, int Np, const double* freq, const double* quad,
                              const double* recv, const double* Ssrc, const double* Srefl, const double* Dsrc,
                              const double* Drefl, double* Sr, double* Si, double* Dr, double* Di, float* D32r,
                              float* D32i, SurCounters* cnt);
void launch_sur_counters_clear(hipStream_t s, SurCounters* cnt);

}  // namespace bemuse
