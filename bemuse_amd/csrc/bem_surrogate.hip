// Wave-term surrogate: sampling from the exact kernel, MLP training
// (forward, backward, Adam) and inference fused into the wave assembly.
// The network and its input/output maps are defined in bem_surrogate.h.
//
// Matrix products run on MFMA v_mfma_f32_16x16x4f32. Lane l = i + 16 g
// supplies A[i][k] and B[k][i] for its k; accumulator register r holds
// D[4 g + r][i]. Within a chunk of 16 k, lane group g covers k = 16 t + 4 g + j
// at step j, so a lane reads four consecutive k as one float4.
#include "bem_surrogate.h"
#include "bem_device.hpp"

#include <cmath>

namespace bemuse {
using namespace dev;

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int LDA = kSurHid + 4;   // LDS row stride of activations (floats): conflict-free float4 row reads
constexpr int SA_TS = 16;          // source panels per work-group of the fused assembly

__device__ inline bool sur_in_domain(double h, double v) {
    return h <= kSurHmax && v >= kSurVmin && v <= 0.0 && (h > 0.0 || v < 0.0);
}

// tanh(x) = 1 - 2 / (e^{2x} + 1); exact limits +-1 for large |x|.
__device__ inline float sur_tanh(float x) { return 1.0f - 2.0f / (__expf(2.0f * x) + 1.0f); }

// J0(x), J1(x) out of line: inlined into the fused assembly they would raise
// its register count above what two work-groups per CU allow.
struct J01 { double j0, j1; };
__device__ __attribute__((noinline)) J01 bessel_j01(double x) {
    J01 o;
    double y0v, y1v;
    bessel01(x, o.j0, o.j1, y0v, y1v);
    return o;
}

__device__ inline f4 mfma4(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// Aout[m][n] = tanh(sum_k Ain[m][k] W[n][k] + bias[n]), m < M, n < 128.
// Wave w computes columns [32 w, 32 w + 32) for all M rows.
template <int M>
__device__ inline void hidden_fwd(const float* __restrict__ W, const float* __restrict__ bias, const float* Ain,
                                  float* Aout, int lane, int wave) {
    constexpr int MB = M / 16;
    const int i = lane & 15, g = lane >> 4, n0 = 32 * wave;
    f4 acc[MB][2];
#pragma unroll
    for (int a = 0; a < MB; a++) { acc[a][0] = f4{0, 0, 0, 0}; acc[a][1] = f4{0, 0, 0, 0}; }
#pragma unroll 1
    for (int t = 0; t < kSurHid / 16; t++) {
        const int k = 16 * t + 4 * g;
        const f4 w0 = *reinterpret_cast<const f4*>(W + (n0 + i) * kSurHid + k);
        const f4 w1 = *reinterpret_cast<const f4*>(W + (n0 + 16 + i) * kSurHid + k);
#pragma unroll
        for (int a = 0; a < MB; a++) {
            const f4 x = *reinterpret_cast<const f4*>(Ain + (16 * a + i) * LDA + k);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc[a][0] = mfma4(x[j], w0[j], acc[a][0]);
                acc[a][1] = mfma4(x[j], w1[j], acc[a][1]);
            }
        }
    }
#pragma unroll
    for (int a = 0; a < MB; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int n = n0 + 16 * c + i;
            const float bn = bias[n];
#pragma unroll
            for (int r = 0; r < 4; r++) Aout[(16 * a + 4 * g + r) * LDA + n] = sur_tanh(acc[a][c][r] + bn);
        }
}

// gW[o][k] = sum_m D[m][o] Ain[m][k] over the M rows of the tile (written to
// global memory). Wave w covers o in [32 w, 32 w + 32), all k.
template <int M>
__device__ inline void hidden_dw(const float* D, const float* Ain, float* __restrict__ gW, int lane, int wave) {
    const int i = lane & 15, g = lane >> 4, o0 = 32 * wave;
    f4 acc[2][8];
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 8; c++) acc[b][c] = f4{0, 0, 0, 0};
#pragma unroll 2
    for (int m0 = 0; m0 < M; m0 += 4) {
        const int m = m0 + g;
        const float d0 = D[m * LDA + o0 + i], d1 = D[m * LDA + o0 + 16 + i];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const float x = Ain[m * LDA + 16 * c + i];
            acc[0][c] = mfma4(d0, x, acc[0][c]);
            acc[1][c] = mfma4(d1, x, acc[1][c]);
        }
    }
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 8; c++)
#pragma unroll
            for (int r = 0; r < 4; r++) gW[(o0 + 16 * b + 4 * g + r) * kSurHid + 16 * c + i] = acc[b][c][r];
}

// acc = D W restricted to columns [32 w, 32 w + 32): acc[a][c][r] is
// sum_o D[m][o] W[o][k] at m = 16 a + 4 g + r, k = 32 w + 16 c + i.
template <int M>
__device__ inline void hidden_bwd(const float* __restrict__ W, const float* D, f4 (&acc)[M / 16][2], int lane,
                                  int wave) {
    constexpr int MB = M / 16;
    const int i = lane & 15, g = lane >> 4, k0 = 32 * wave;
#pragma unroll
    for (int a = 0; a < MB; a++) { acc[a][0] = f4{0, 0, 0, 0}; acc[a][1] = f4{0, 0, 0, 0}; }
#pragma unroll 2
    for (int t = 0; t < kSurHid / 16; t++) {
        const int o = 16 * t + 4 * g;
        float w0[4], w1[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            w0[j] = W[(o + j) * kSurHid + k0 + i];
            w1[j] = W[(o + j) * kSurHid + k0 + 16 + i];
        }
#pragma unroll
        for (int a = 0; a < MB; a++) {
            const f4 d = *reinterpret_cast<const f4*>(D + (16 * a + i) * LDA + o);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc[a][0] = mfma4(d[j], w0[j], acc[a][0]);
                acc[a][1] = mfma4(d[j], w1[j], acc[a][1]);
            }
        }
    }
}

// Backward through one 128 x 128 layer: D (this layer's pre-activation
// gradient, LDS) and Ain (its input activations, LDS, overwritten by the
// input layer's pre-activation gradient D W * (1 - Ain^2)).
template <int M>
__device__ inline void hidden_backward(const float* __restrict__ W, float* __restrict__ gW, float* __restrict__ gb,
                                       const float* D, float* Ain, int tid, int lane, int wave) {
    hidden_dw<M>(D, Ain, gW, lane, wave);
    if (tid < kSurHid) {
        float s = 0.0f;
        for (int m = 0; m < M; m++) s += D[m * LDA + tid];
        gb[tid] = s;
    }
    f4 acc[M / 16][2];
    hidden_bwd<M>(W, D, acc, lane, wave);
    __syncthreads();   // every wave has finished reading Ain
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int a = 0; a < M / 16; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int o = (16 * a + 4 * g + r) * LDA + 32 * wave + 16 * c + i;
                const float av = Ain[o];
                Ain[o] = acc[a][c][r] * (1.0f - av * av);
            }
    __syncthreads();
}

// Small layers staged in LDS by every kernel that runs the network.
struct SmallLayers {
    float w1[kSurHid * kSurIn], b1[kSurHid], w5[kSurOut * kSurHid], b5[kSurOut];
};

__device__ inline void stage_small(const float* __restrict__ P, SmallLayers& L, int tid) {
    for (int k = tid; k < kSurHid * kSurIn; k += 256) L.w1[k] = P[kSurW1 + k];
    for (int k = tid; k < kSurHid; k += 256) L.b1[k] = P[kSurB1 + k];
    for (int k = tid; k < kSurOut * kSurHid; k += 256) L.w5[k] = P[kSurW5 + k];
    if (tid < kSurOut) L.b5[tid] = P[kSurB5 + tid];
}

// Input layer: A[m][n] = tanh(W1[n] . x[m] + b1[n]) for the M rows.
template <int M>
__device__ inline void input_fwd(const SmallLayers& L, const float (*x)[4], float* A, int tid) {
#pragma unroll 2
    for (int idx = tid; idx < M * kSurHid; idx += 256) {
        const int m = idx >> 7, n = idx & 127;
        const float* w = L.w1 + n * kSurIn;
        float z = L.b1[n];
        z += x[m][0] * w[0];
        z += x[m][1] * w[1];
        z += x[m][2] * w[2];
        z += x[m][3] * w[3];
        A[m * LDA + n] = sur_tanh(z);
    }
}

// Output layer for row m, channel c.
__device__ inline float output_unit(const SmallLayers& L, const float* A, int m, int c) {
    const float* a = A + m * LDA;
    const float* w = L.w5 + c * kSurHid;
    float y = L.b5[c];
#pragma unroll 8
    for (int o = 0; o < kSurHid; o++) y += a[o] * w[o];
    return y;
}

// Forward pass of 64 rows: features x -> outputs y (both LDS). Uses act0
// and act1 as ping-pong buffers. Ends synchronised.
__device__ __attribute__((always_inline)) inline void forward64(const float* __restrict__ P, const SmallLayers& L, const float (*x)[4], float* act0,
                                 float* act1, float (*y)[kSurOut], int tid, int lane, int wave) {
    input_fwd<64>(L, x, act0, tid);
    __syncthreads();
    hidden_fwd<64>(P + kSurW2, P + kSurB2, act0, act1, lane, wave);
    __syncthreads();
    hidden_fwd<64>(P + kSurW3, P + kSurB3, act1, act0, lane, wave);
    __syncthreads();
    hidden_fwd<64>(P + kSurW4, P + kSurB4, act0, act1, lane, wave);
    __syncthreads();
    if (tid < 64 * kSurOut) y[tid / kSurOut][tid % kSurOut] = output_unit(L, act1, tid / kSurOut, tid % kSurOut);
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Sampling: sample i draws u0 = u(2 i), u1 = u(2 i + 1) with
// u(k) = unit_open0(seed + k) and maps them by sur_sample_point; targets
// from the exact kernel.
// ---------------------------------------------------------------------------
__global__ void k_sur_sample(uint64_t seed, int sampling, long long first, long long n, float* __restrict__ X,
                             float* __restrict__ T, double* __restrict__ hv) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint64_t idx = (uint64_t)(first + j);
    double h, v;
    sur_sample_point(sampling, unit_open0(seed + 2 * idx), unit_open0(seed + 2 * idx + 1), h, v);
    const SurPoint p = sur_point(h, v);
    const GFVal g = havelock_gf(h, 0.0, v);
    const double is = 1.0 / p.s;
    float* x = X + 4 * j;
    float* t = T + kSurOut * j;
    x[0] = p.f[0]; x[1] = p.f[1]; x[2] = p.f[2]; x[3] = p.f[3];
    t[0] = (float)((g.gr - p.sg) * is);
    t[1] = (float)((g.hr - p.sgh) * is);
    if (hv) { hv[2 * j] = h; hv[2 * j + 1] = v; }
}

// ---------------------------------------------------------------------------
// Training: forward and backward of one 32-sample tile per work-group. The
// loss is sum over samples and channels of (y - t)^2 / (kSurOut batch).
// ---------------------------------------------------------------------------
struct GradLds {
    SmallLayers L;
    float sx[kSurTile][4], st[kSurTile][kSurOut], sdy[kSurTile][kSurOut];
    __align__(16) float act[4][kSurTile * LDA];
    float sred[256];
};

// Tile `tile` of the batch starting at pool position p0: writes its gradient
// partial (kSurParams floats at gpart + tile kSurParams) and its loss
// partial lpart[tile].
__device__ void sur_grad_tile(GradLds& S, int tile, const float* __restrict__ P, const float* __restrict__ X,
                              const float* __restrict__ T, long long pool, long long p0, int batch,
                              const unsigned long long* __restrict__ perm, float dscale, float* __restrict__ gpart,
                              float* __restrict__ lpart) {
    constexpr int M = kSurTile;
    SmallLayers& L = S.L;
    auto& sx = S.sx;
    auto& st = S.st;
    auto& sdy = S.sdy;
    auto& act = S.act;
    auto& sred = S.sred;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j0 = tile * M;
    float* gp = gpart + (size_t)tile * kSurParams;
    stage_small(P, L, tid);
    if (tid < M) {
        const int j = j0 + tid;
        float x[4] = {0.0f, 0.0f, 0.0f, 0.0f}, t[kSurOut] = {};
        if (j < batch) {
            const long long p = p0 + j;
            long long row = p;
            if (perm) {
                const long long e = p / pool, q = p - e * pool;
                const unsigned long long a = perm[2 * (e & 15)], b = perm[2 * (e & 15) + 1];
                row = (long long)((a * (unsigned long long)q + b) % (unsigned long long)pool);
            }
            for (int c = 0; c < 4; c++) x[c] = X[4 * row + c];
            for (int c = 0; c < kSurOut; c++) t[c] = T[kSurOut * row + c];
        }
        for (int c = 0; c < 4; c++) sx[tid][c] = x[c];
        for (int c = 0; c < kSurOut; c++) st[tid][c] = t[c];
    }
    __syncthreads();
    input_fwd<M>(L, sx, act[0], tid);
    __syncthreads();
    hidden_fwd<M>(P + kSurW2, P + kSurB2, act[0], act[1], lane, wave);
    __syncthreads();
    hidden_fwd<M>(P + kSurW3, P + kSurB3, act[1], act[2], lane, wave);
    __syncthreads();
    hidden_fwd<M>(P + kSurW4, P + kSurB4, act[2], act[3], lane, wave);
    __syncthreads();
    float sq = 0.0f;
    if (tid < M * kSurOut) {
        const int m = tid / kSurOut, c = tid % kSurOut;
        const float y = output_unit(L, act[3], m, c);
        float d = 0.0f;
        if (j0 + m < batch) {
            const float e = y - st[m][c];
            sq = e * e;
            d = e * dscale;
        }
        sdy[m][c] = d;
    }
    sred[tid] = sq;
    __syncthreads();
    for (int k = tid; k < kSurOut * kSurHid; k += 256) {
        const int c = k >> 7, o = k & 127;
        float s = 0.0f;
        for (int m = 0; m < M; m++) s += sdy[m][c] * act[3][m * LDA + o];
        gp[kSurW5 + k] = s;
    }
    if (tid < kSurOut) {
        float s = 0.0f;
        for (int m = 0; m < M; m++) s += sdy[m][tid];
        gp[kSurB5 + tid] = s;
    }
    if (tid == 0) {
        float s = 0.0f;
        for (int k = 0; k < 256; k++) s += sred[k];
        lpart[tile] = s;
    }
    __syncthreads();
    for (int k = tid; k < M * kSurHid; k += 256) {
        const int m = k >> 7, o = k & 127;
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < kSurOut; c++) s += sdy[m][c] * L.w5[c * kSurHid + o];
        const float a = act[3][m * LDA + o];
        act[3][m * LDA + o] = s * (1.0f - a * a);
    }
    __syncthreads();
    hidden_backward<M>(P + kSurW4, gp + kSurW4, gp + kSurB4, act[3], act[2], tid, lane, wave);
    hidden_backward<M>(P + kSurW3, gp + kSurW3, gp + kSurB3, act[2], act[1], tid, lane, wave);
    hidden_backward<M>(P + kSurW2, gp + kSurW2, gp + kSurB2, act[1], act[0], tid, lane, wave);
    for (int k = tid; k < kSurHid * kSurIn; k += 256) {
        const int n = k >> 2, c = k & 3;
        float s = 0.0f;
        for (int m = 0; m < M; m++) s += act[0][m * LDA + n] * sx[m][c];
        gp[kSurW1 + k] = s;
    }
    if (tid < kSurHid) {
        float s = 0.0f;
        for (int m = 0; m < M; m++) s += act[0][m * LDA + tid];
        gp[kSurB1 + tid] = s;
    }
}

__global__ __launch_bounds__(256) void k_sur_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                  const float* __restrict__ T, long long pool, long long p0,
                                                  int batch, const unsigned long long* __restrict__ perm,
                                                  float dscale, float* __restrict__ gpart,
                                                  float* __restrict__ lpart) {
    __shared__ GradLds S;
    sur_grad_tile(S, blockIdx.x, P, X, T, pool, p0, batch, perm, dscale, gpart, lpart);
}

// Parameter p: the sum of the nwg work-group partials (FP64, a fixed order:
// eight interleaved running sums over w = j mod 8, then combined pairwise, so
// that eight loads are in flight instead of one chain) and one Adam step.
__device__ inline void sur_adam_param(int p, int nwg, const float* __restrict__ gpart, float* __restrict__ P,
                                      float* __restrict__ m, float* __restrict__ v, float* __restrict__ grad,
                                      float lr_t, float beta1, float beta2, float eps) {
    double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int w = 0;
    for (; w + 8 <= nwg; w += 8) {
        float x[8];
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = gpart[(size_t)(w + j) * kSurParams + p];
#pragma unroll
        for (int j = 0; j < 8; j++) acc[j] += x[j];
    }
    for (int j = 0; w < nwg; w++, j++) acc[j] += gpart[(size_t)w * kSurParams + p];
    const double gs = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    const float g = (float)gs;
    if (grad) grad[p] = g;
    if (m) {
        const float mm = beta1 * m[p] + (1.0f - beta1) * g;
        const float vv = beta2 * v[p] + (1.0f - beta2) * g * g;
        m[p] = mm;
        v[p] = vv;
        P[p] -= lr_t * mm / (sqrtf(vv) + eps);
    }
}

__device__ inline double sur_loss_sum(int nwg, const float* __restrict__ lpart) {
    double s = 0.0;
    for (int w = 0; w < nwg; w++) s += lpart[w];
    return s;
}

__global__ void k_sur_adam(int nwg, const float* __restrict__ gpart, const float* __restrict__ lpart,
                           float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                           float* __restrict__ grad, double* __restrict__ loss, float lr_t, float beta1,
                           float beta2, float eps) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && loss) loss[0] = sur_loss_sum(nwg, lpart);
    if (p >= kSurParams) return;
    sur_adam_param(p, nwg, gpart, P, m, v, grad, lr_t, beta1, beta2, eps);
}

// Barrier of all work-groups of a cooperative launch (every work-group is
// resident, which the cooperative launch guarantees or refuses). bar[0]
// counts arrivals, bar[kBarGen] is the generation the last arrival advances.
// The wait is bounded: after kBarSpins polls the work-group sets *err and
// gives up, so a broken barrier ends the kernel instead of hanging it.
constexpr int kBarGen = 32;                 // the generation word on its own 128-byte line
constexpr long long kBarSpins = 1LL << 21;  // about a second of polling; a step takes ~0.1 ms

__device__ bool grid_barrier(unsigned* bar, unsigned nblocks, int* err) {
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        ok = 1;
        unsigned* count = bar;
        unsigned* gen = bar + kBarGen;
        const unsigned g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // release: this work-group's writes (and the read of g) before its arrival
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned arrived = atomicAdd(count, 1u);
        if (arrived == nblocks - 1) {
            __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // acquire the arrivals' writes, release them (and the reset) with the new generation
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            atomicAdd(gen, 1u);
        } else {
            long long spins = 0;
            while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > kBarSpins) {
                    atomicOr(err, 1);
                    ok = 0;
                    break;
                }
            }
        }
        // acquire: the other work-groups' writes after the wait
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    return ok != 0;
}

// Steps t0 .. t0 + nsteps - 1 of the training in one cooperative launch of
// one work-group per 32-sample tile: each step is the gradient of every tile,
// a barrier, the Adam step of all parameters spread over the work-groups
// (the same fixed-order sums as k_sur_adam, so the parameters equal those of
// one gradient and one Adam dispatch per step, bit for bit), and a barrier.
__global__ __launch_bounds__(256) void k_sur_train(int t0, int nsteps, const float* __restrict__ lr_tab,
                                                   float beta1, float beta2, float eps, const float* __restrict__ X,
                                                   const float* __restrict__ T, long long pool, int batch,
                                                   const unsigned long long* __restrict__ perm, float dscale,
                                                   float* P, float* m, float* v, float* gpart, float* lpart,
                                                   double* loss, unsigned* bar, int* err) {
    __shared__ GradLds S;
    const int nwg = gridDim.x;
    for (int s = 0; s < nsteps; s++) {
        const int t = t0 + s;
        sur_grad_tile(S, blockIdx.x, P, X, T, pool, (long long)t * batch, batch, perm, dscale, gpart, lpart);
        if (!grid_barrier(bar, nwg, err)) return;
        if (blockIdx.x == 0 && threadIdx.x == 0) loss[t] = sur_loss_sum(nwg, lpart);
        const float lr_t = lr_tab[t];
        for (int p = blockIdx.x * 256 + threadIdx.x; p < kSurParams; p += nwg * 256)
            sur_adam_param(p, nwg, gpart, P, m, v, nullptr, lr_t, beta1, beta2, eps);
        if (!grid_barrier(bar, nwg, err)) return;
    }
}

// Network outputs for n feature rows, 64 rows per work-group.
__global__ __launch_bounds__(256) void k_sur_forward(const float* __restrict__ P, long long n,
                                                     const float* __restrict__ X, float* __restrict__ Y) {
    __shared__ SmallLayers L;
    __shared__ float sx[64][4], sy[64][kSurOut];
    __shared__ __align__(16) float act[2][64 * LDA];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long j0 = (long long)blockIdx.x * 64;
    stage_small(P, L, tid);
    {
        const int m = tid >> 2, c = tid & 3;
        sx[m][c] = (j0 + m < n) ? X[4 * (j0 + m) + c] : 0.0f;
    }
    __syncthreads();
    forward64(P, L, sx, act[0], act[1], sy, tid, lane, wave);
    if (tid < 64 * kSurOut) {
        const int m = tid / kSurOut, c = tid % kSurOut;
        if (j0 + m < n) Y[kSurOut * (j0 + m) + c] = sy[m][c];
    }
}

// Reconstructed (Re GF, Im GF, Re GFh, Im GFh) at n points (h, v).
__global__ __launch_bounds__(256) void k_sur_eval(const float* __restrict__ P, int n, const double* __restrict__ H,
                                                  const double* __restrict__ V, double* __restrict__ out,
                                                  int* __restrict__ used) {
    __shared__ SmallLayers L;
    __shared__ float sx[64][4], sy[64][kSurOut];
    __shared__ __align__(16) float act[2][64 * LDA];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = blockIdx.x * 64 + lane;
    stage_small(P, L, tid);
    double h = 0.0, v = -1.0;
    bool in = false;
    SurPoint sp{};
    if (wave == 0) {
        if (j < n) { h = H[j]; v = V[j]; }
        in = j < n && sur_in_domain(h, v);
        sp = sur_point(in ? h : 1.0, in ? v : -1.0);
        for (int c = 0; c < 4; c++) sx[lane][c] = sp.f[c];
    }
    __syncthreads();
    forward64(P, L, sx, act[0], act[1], sy, tid, lane, wave);
    if (wave == 0 && j < n) {
        double* o = out + 4 * (size_t)j;
        if (in) {
            const double E = 2.0 * kPI * exp(v);
            const J01 jj = bessel_j01(h);
            o[0] = sp.sg + sp.s * (double)sy[lane][0];
            o[1] = -E * jj.j0;
            o[2] = sp.sgh + sp.s * (double)sy[lane][1];
            o[3] = E * jj.j1;
        } else {
            const GFVal g = havelock_gf(h, 0.0, v);
            o[0] = g.gr; o[1] = g.gi; o[2] = g.hr; o[3] = g.hi;
        }
        if (used) used[j] = in ? 1 : 0;
    }
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long x) {
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}
__device__ inline float wave_max_f(float x) {
    for (int d = 32; d >= 1; d >>= 1) x = fmaxf(x, __shfl_xor(x, d, 64));
    return x;
}

// ---------------------------------------------------------------------------
// Wave assembly fused with compose (k_wave_compose with compose = wave),
// HavelockGF replaced inside the trained domain by the network. Work-group:
// 64 receivers (one per lane of wave 0) x SA_TS source panels; each
// (panel, quadrature node) is one 64-row network pass shared by the four
// waves. One pair in kSurJudgeStride is also evaluated by the exact kernel,
// which records the largest relative deviation.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256, 2) void k_wave_compose_surr(
    const float* __restrict__ P, int N, int Np, const double* __restrict__ freq, const double* __restrict__ quad,
    const double* __restrict__ recv, const double* __restrict__ Ssrc, const double* __restrict__ Srefl,
    const double* __restrict__ Dsrc, const double* __restrict__ Drefl, double* Sr, double* Si, double* Dr,
    double* Di, float* D32r, float* D32i, SurCounters* cnt) {
    __shared__ SmallLayers L;
    __shared__ double sq[SA_TS][32];
    __shared__ float sx[64][4], sy[64][kSurOut];
    __shared__ __align__(16) float act[2][64 * LDA];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = blockIdx.x * 64, s0 = blockIdx.y * SA_TS;
    const double K = freq[1];
    stage_small(P, L, tid);
    for (int k = tid; k < SA_TS * 32; k += 256) {
        const int s = s0 + k / 32;
        sq[k / 32][k % 32] = s < N ? quad[(size_t)s * 32 + k % 32] : 0.0;
    }
    const int R = r0 + lane;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (wave == 0 && R < N) { px = recv[4 * (size_t)R]; py = recv[4 * (size_t)R + 1]; pz = recv[4 * (size_t)R + 2]; }
    const double s1 = K * kFourPIinv, s2 = K * K * kFourPIinv;
    unsigned long long nsur = 0, nexact = 0;
    float errmax = 0.0f;
    __syncthreads();
    for (int sl = 0; sl < SA_TS; sl++) {
        const int S = s0 + sl;
        const size_t idx = (size_t)S * Np + R;
        if (S >= N) {   // padding column: identity in D, zero elsewhere
            if (wave == 0) {
                const double dre = (R == S) ? 1.0 : 0.0;
                Sr[idx] = 0.0; Si[idx] = 0.0; Dr[idx] = dre; Di[idx] = 0.0;
                if (D32r) { D32r[idx] = (float)dre; D32i[idx] = 0.0f; }
            }
            continue;
        }
        const double* q = sq[sl];
        const int nq = (int)q[28];
        double wsr = 0.0, wsi = 0.0, wdr = 0.0, wdi = 0.0;
        for (int k = 0; k < nq; k++) {
            const double* p = q + 7 * k;
            double X = 0.0, Y = 0.0, V = 0.0, h = 0.0;
            bool in = false;
            SurPoint sp{};
            if (wave == 0) {
                X = (px - p[0]) * K; Y = (py - p[1]) * K; V = (pz - p[2]) * K;
                h = sqrt(X * X + Y * Y);
                in = R < N && sur_in_domain(h, V);
                sp = sur_point(in ? h : 1.0, in ? V : -1.0);
                for (int c = 0; c < 4; c++) sx[lane][c] = sp.f[c];
            }
            __syncthreads();
            forward64(P, L, sx, act[0], act[1], sy, tid, lane, wave);
            if (wave == 0 && R < N) {
                double gr = 0.0, gi = 0.0, hr = 0.0, hi = 0.0;
                if (in) {
                    const double E = 2.0 * kPI * exp(V);
                    const J01 jj = bessel_j01(h);
                    gr = sp.sg + sp.s * (double)sy[lane][0];
                    gi = -E * jj.j0;
                    hr = sp.sgh + sp.s * (double)sy[lane][1];
                    hi = E * jj.j1;
                    nsur++;
                } else {
                    nexact++;
                }
                const bool judge = in && (unsigned)(R * 7 + S * 13 + k) % kSurJudgeStride == 0;
                if (!in || judge) {   // one call site of the exact kernel
                    const GFVal g = havelock_gf(X, Y, V);
                    if (judge) {
                        const double den = fmax(fmax(fabs(g.gr), fabs(g.gi)), fmax(fabs(g.hr), fabs(g.hi)));
                        const double num = fmax(fmax(fabs(gr - g.gr), fabs(gi - g.gi)), fmax(fabs(hr - g.hr), fabs(hi - g.hi)));
                        if (den > 0.0) errmax = fmaxf(errmax, (float)(num / den));
                    } else {
                        gr = g.gr; gi = g.gi; hr = g.hr; hi = g.hi;
                    }
                }
                // one quadrature node of wave_panel (Panel.cpp:601-636,776-794)
                const double w = p[6];
                wsr += w * (gr * s1);
                wsi += w * (gi * s1);
                double cx = 0.0, cy = 0.0;
                if (h > 1.0e-6) { cx = X / h; cy = Y / h; }
                const double tr = (hr * cx) * p[3] + (hr * cy) * p[4] + gr * p[5];
                const double ti = (hi * cx) * p[3] + (hi * cy) * p[4] + gi * p[5];
                wdr += w * (tr * s2);
                wdi += w * (ti * s2);
            }
        }
        if (wave == 0) {
            double sre = 0.0, sim = 0.0, dre = 0.0, dim = 0.0;
            if (R < N) {
                const double ss = Ssrc[idx], sf = Srefl[idx], ds = Dsrc[idx], df = Drefl[idx];
                sre = (ss + sf) - wsr;
                sim = -wsi;
                dre = (ds + df) + (wdr + sf * q[29] * K);
                dim = wdi;
            }
            Sr[idx] = sre; Si[idx] = sim; Dr[idx] = dre; Di[idx] = dim;
            if (D32r) { D32r[idx] = (float)dre; D32i[idx] = (float)dim; }
        }
    }
    if (wave == 0) {
        nsur = wave_sum_u64(nsur);
        nexact = wave_sum_u64(nexact);
        errmax = wave_max_f(errmax);
        if (lane == 0) {
            if (nsur) atomicAdd(&cnt->n_sur, nsur);
            if (nexact) atomicAdd(&cnt->n_exact, nexact);
            atomicMax(&cnt->err_bits, __float_as_uint(errmax));
        }
    }
}

__global__ void k_sur_counters_clear(SurCounters* c) {
    c->n_sur = 0;
    c->n_exact = 0;
    c->err_bits = 0;
    c->pad = 0;
}

inline int cdivll(long long a, int b) { return (int)((a + b - 1) / b); }

}  // namespace

void surrogate_init_params(uint64_t seed, std::vector<float>& p) {
    p.assign(kSurParams, 0.0f);
    uint64_t k = 0;
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double a = sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < fan_out * fan_in; i++) p[off + i] = (float)(a * (2.0 * unit_open0(seed + k++) - 1.0));
    };
    fill(kSurW1, kSurHid, kSurIn);
    fill(kSurW2, kSurHid, kSurHid);
    fill(kSurW3, kSurHid, kSurHid);
    fill(kSurW4, kSurHid, kSurHid);
    fill(kSurW5, kSurOut, kSurHid);
}

void launch_sur_sample(hipStream_t s, uint64_t seed, int sampling, long long first, long long n, float* X, float* T,
                       double* hv) {
    hipLaunchKernelGGL(k_sur_sample, dim3(cdivll(n, 256)), dim3(256), 0, s, seed, sampling, first, n, X, T, hv);
}

int sur_grad_groups(int batch) { return (batch + kSurTile - 1) / kSurTile; }

void launch_sur_grad(hipStream_t s, const float* params, const float* X, const float* T, long long pool,
                     long long p0, int batch, const unsigned long long* perm, float* gpart, float* lpart) {
    const float dscale = 2.0f / ((float)kSurOut * (float)batch);
    hipLaunchKernelGGL(k_sur_grad, dim3(sur_grad_groups(batch)), dim3(256), 0, s, params, X, T, pool, p0, batch,
                       perm, dscale, gpart, lpart);
}

int launch_sur_train(hipStream_t s, int nwg, int t0, int nsteps, const float* lr_tab, float beta1, float beta2,
                     float eps, const float* X, const float* T, long long pool, int batch,
                     const unsigned long long* perm, float* params, float* m, float* v, float* gpart, float* lpart,
                     double* loss, unsigned* bar, int* err) {
    float dscale = 2.0f / ((float)kSurOut * (float)batch);
    void* args[] = {&t0, &nsteps, &lr_tab, &beta1, &beta2, &eps, &X, &T, &pool, &batch, &perm, &dscale,
                    &params, &m, &v, &gpart, &lpart, &loss, &bar, &err};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_sur_train), dim3(nwg), dim3(256), args, 0, s) ==
                   hipSuccess
               ? 0
               : -1;
}

int sur_train_coop_capacity(int device) {
    int coop = 0, cus = 0, per = 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess || !coop) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(&k_sur_train), 256, 0) !=
        hipSuccess)
        return 0;
    return per * cus;
}

void launch_sur_adam(hipStream_t s, int nwg, const float* gpart, const float* lpart, float* params, float* m,
                     float* v, float* grad, double* loss, float lr_t, float beta1, float beta2, float eps) {
    hipLaunchKernelGGL(k_sur_adam, dim3(cdivll(kSurParams, 256)), dim3(256), 0, s, nwg, gpart, lpart, params, m, v,
                       grad, loss, lr_t, beta1, beta2, eps);
}

void launch_sur_forward(hipStream_t s, const float* params, long long n, const float* X, float* Y) {
    hipLaunchKernelGGL(k_sur_forward, dim3(cdivll(n, 64)), dim3(256), 0, s, params, n, X, Y);
}

void launch_sur_eval(hipStream_t s, const float* params, int n, const double* h, const double* v, double* out,
                     int* used) {
    hipLaunchKernelGGL(k_sur_eval, dim3(cdivll(n, 64)), dim3(256), 0, s, params, n, h, v, out, used);
}

void launch_wave_compose_surr(hipStream_t s, const float* params, int N, int Np, const double* freq, const double* quad,
                              const double* recv, const double* Ssrc, const double* Srefl, const double* Dsrc,
                              const double* Drefl, double* Sr, double* Si, double* Dr, double* Di, float* D32r,
                              float* D32i, SurCounters* cnt) {
    dim3 grid(Np / 64, Np / SA_TS);
    hipLaunchKernelGGL(k_wave_compose_surr, grid, dim3(256), 0, s, params, N, Np, freq, quad, recv, Ssrc, Srefl, Dsrc,
                       Drefl, Sr, Si, Dr, Di, D32r, D32i, cnt);
}

void launch_sur_counters_clear(hipStream_t s, SurCounters* cnt) {
    hipLaunchKernelGGL(k_sur_counters_clear, dim3(1), dim3(1), 0, s, cnt);
}

}  // namespace bemuse
