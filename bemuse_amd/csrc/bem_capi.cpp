// C-ABI of the library (include/bemuse_amd.h): meshes, contexts, bodies, the
// per-frequency solve and the pipelined frequency sweep. Host orchestration
// only; arithmetic is on the GPU.
#include "../../include/bemuse_amd.h"
#include "bem_device.hpp"
#include "bem_internal.h"
#include "bem_kernels.h"
#include "bem_surrogate.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cctype>
#include <cstring>
#include <fstream>
#include <mutex>
#include <string>
#include <thread>
#include <chrono>
#include <vector>

using namespace bemuse;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(BEM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

int dalloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(BEM_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e));
    }
    return BEM_OK;
}

template <typename T>
int dalloc_t(T** p, size_t n) { return dalloc(reinterpret_cast<void**>(p), n * sizeof(T)); }

inline int round_up(int n, int m) { return (n + m - 1) / m * m; }

// Temporary device buffers of one call, freed on every exit path.
struct SurBuffers {
    std::vector<void*> ptrs;
    ~SurBuffers() {
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
    template <typename T>
    int get(T** p, size_t n) {
        int r = dalloc_t(p, n);
        if (!r) ptrs.push_back(*p);
        return r;
    }
};

}  // namespace

struct bem_mesh {
    Mesh m;
};

// Wave-term surrogate held by a context: its network parameters on the
// device and the report of the training that produced them.
struct SurState {
    float* params = nullptr;
    bem_surrogate_report rep{};
    ~SurState() {
        if (params) (void)hipFree(params);
    }
};

// Iterative refinement of the mixed modes: at most kRefineMaxIter correction
// steps (bem_set_refine_limit lowers it), stopping at a max-norm relative
// residual of kRefineTol: with the system's condition numbers this keeps A, B,
// F orders of magnitude inside the 1e-6 bar while saving correction sweeps.
constexpr int kRefineMaxIter = 4;
constexpr double kRefineTol = 1e-13;
// Correction steps a solve lane enqueues with the frequency; the rest, when
// its refinement needs them, follow once the lane's results are in (every
// enqueued step that the device skips still costs a launch on the lane's
// chain). The benchmark systems converge in two.
constexpr int kRefineEager = 2;
// Below this padded order the wave assembly runs about two work-groups per CU
// (wg_per_cu of launch_wave_compose) instead of one per tile: the small
// bodies' lanes spend most of their time in chains of short LU launches, and
// an assembly that fills every CU delays those chains. Measured on the 64-body
// batch (orders 640-4096): 620 frequency-solves/s uncapped, 634-638 capped;
// on the 6018-panel sweep, uncapped is 1% faster.
constexpr int kWaveCapNp = 6144;
constexpr int kWaveCapWgPerCu = 2;

// Largest order of bem_dense_solve_noalloc and its padded order.
constexpr int kNaMaxN = BEM_NOALLOC_MAX_N;
constexpr int kNaNp = (kNaMaxN + 127) / 128 * 128;
constexpr int kNaEvents = 2 * (kNaNp / 64) + 2;    // LU lookahead events at that order

struct bem_ctx {
    int device = 0;
    int refine_max = kRefineMaxIter;   // correction steps allowed (bem_set_refine_limit)
    int lu_arith = kLuArithSplitF16;   // arithmetic of the FP32 factorisation's updates (bem_set_lu_arith)
    double wave_far_r2 = -1.0;         // bem_set_wave_far32: squared radius beyond which the wave term runs FP32 special functions
    int wave_wg_per_cu = -1;           // bem_set_wave_wg_per_cu: -1 by order (kWaveCapNp), 0 one work-group per tile
    // streams and events of bem_dense_solve_noalloc, made here so that the
    // entry point itself creates nothing
    hipStream_t na_s2 = nullptr;
    hipEvent_t na_ev[kNaEvents] = {};
    hipEvent_t na_t[3] = {};
    SurState* surrogate = nullptr;     // trained wave-term surrogate, if any
    bool graphs = false;               // replay captured frequency graphs on the lanes (bem_set_graphs)
    int lu_spec = 1;                   // solve lanes factor speculatively (bem_set_lu_speculative)
    double pivot_tau = 1.0;            // multiplier bound of the pivot check (bem_set_pivot_threshold)
    int refine_eager = kRefineEager;   // correction steps a lane enqueues with its frequency (bem_set_refine_eager)
    hipStream_t stream = nullptr;      // setup work (uploads, Rankine assembly)
    hipEvent_t ev[2] = {};
    std::vector<struct Lane*> pool;    // lanes shared by the bodies of bem_solve_batch
};

// Dense system workspace. Buffers are allocated for a capacity capNp; N and
// Np are those of the body currently bound (Np <= capNp), and every kernel
// uses Np as the leading dimension, so one workspace serves any smaller body.
struct LinSys {
    int N = 0, Np = 0, capNp = 0;
    double *Dr = nullptr, *Di = nullptr;          // FP64 system matrix (LU in place in FP64 mode)
    float *D32r = nullptr, *D32i = nullptr;       // FP32 copy factored in mixed mode
    int *ipiv = nullptr, *flag = nullptr;
    int *ipiv32 = nullptr, *flag32 = nullptr;
    LuWork<double> w64;
    LuWork<float> w32;
    double *Br = nullptr, *Bi = nullptr, *Xr = nullptr, *Xi = nullptr, *Rr = nullptr, *Ri = nullptr;
    float *X32r = nullptr, *X32i = nullptr;
    double *Pr = nullptr, *Pi = nullptr, *Pabs = nullptr;   // matvec partials
    double* anorm = nullptr;
    bool have32 = false;
    std::vector<void*> owned;

    LinSys() = default;
    LinSys(const LinSys&) = delete;
    LinSys& operator=(const LinSys&) = delete;
    ~LinSys() {
        for (void* p : owned)
            if (p) (void)hipFree(p);
    }
    template <typename T>
    int get(T** p, size_t n) {
        int r = dalloc_t(p, n);
        if (!r) owned.push_back(*p);
        return r;
    }
    template <typename T>
    int work(LuWork<T>& w) {
        size_t pan = (size_t)capNp * kLuNb;
        int r = BEM_OK;
        if ((r = get(&w.sd_r, kLuNb * kLuNb)) || (r = get(&w.sd_i, kLuNb * kLuNb)) || (r = get(&w.s21_r, pan)) ||
            (r = get(&w.s21_i, pan)) || (r = get(&w.Linv_r, pan)) || (r = get(&w.Linv_i, pan)) ||
            (r = get(&w.Uinv_r, pan)) || (r = get(&w.Uinv_i, pan)) || (r = get(&w.PL_r, pan)) ||
            (r = get(&w.PL_i, pan)) || (r = get(&w.PU_r, pan)) || (r = get(&w.PU_i, pan)))
            return r;
        if (sizeof(T) == 4 && (r = get(&w.umax, (size_t)capNp / kLuNb))) return r;
        if (sizeof(T) == 4 && ((r = get(&w.opA, (size_t)4 * capNp * kSplitGroupMax * kLuNb)) ||
                               (r = get(&w.opB, (size_t)4 * capNp * kSplitGroupMax * kLuNb))))
            return r;
        return BEM_OK;
    }
    int init(int n) {
        N = n; Np = round_up(n, kPad); capNp = Np;
        size_t NN = (size_t)Np * Np, NR = (size_t)Np * kMaxRhs;
        int r = BEM_OK;
        if ((r = get(&Dr, NN)) || (r = get(&Di, NN))) return r;
        if ((r = get(&ipiv, Np)) || (r = get(&flag, Np / kLuNb + 1))) return r;
        if ((r = work(w64))) return r;
        if ((r = get(&Br, NR)) || (r = get(&Bi, NR)) || (r = get(&Xr, NR)) || (r = get(&Xi, NR)) ||
            (r = get(&Rr, NR)) || (r = get(&Ri, NR)))
            return r;
        size_t nks = (size_t)matvec_kslabs(Np);
        if ((r = get(&Pr, nks * NR)) || (r = get(&Pi, nks * NR)) || (r = get(&Pabs, nks * Np))) return r;
        if ((r = get(&anorm, 1))) return r;
        return BEM_OK;
    }
    int init32() {
        if (have32) return BEM_OK;
        size_t NN = (size_t)capNp * capNp, NR = (size_t)capNp * kMaxRhs;
        int r = BEM_OK;
        if ((r = get(&D32r, NN)) || (r = get(&D32i, NN))) return r;
        if ((r = get(&ipiv32, capNp)) || (r = get(&flag32, capNp / kLuNb + 1))) return r;
        if ((r = work(w32))) return r;
        if ((r = get(&X32r, NR)) || (r = get(&X32i, NR))) return r;
        have32 = true;
        return BEM_OK;
    }
};

// Configuration a captured frequency graph is valid for: the body (by id,
// never reused), its padded order, mode, headings, rad_only, compose case and
// surrogate use.
struct GraphKey {
    unsigned long long body = 0;
    int Np = 0, mode = -1, nb = -1, rad_only = -1, compose = -1, sur = -1, refine = -1, arith = -1, wave_cap = -2;
    double far_r2 = -2.0;
    bool operator==(const GraphKey& o) const {
        return body == o.body && Np == o.Np && mode == o.mode && nb == o.nb && rad_only == o.rad_only &&
               compose == o.compose && sur == o.sur && refine == o.refine && arith == o.arith &&
               wave_cap == o.wave_cap && far_r2 == o.far_r2;
    }
};

// Results of one frequency as copied back to the host (pinned).
struct LaneOut {
    double coeffs[2 * (36 + 12 * kMaxRhs)];
    RefineCtrl ctrl;
    SurCounters sur;
    double freq[2];          // (omega, kappa) of the frequency, host -> device
    int flags[kMaxPanels + 1];
};

// One solve lane: the streams, events, solve-chain state and per-frequency
// buffers needed to run a frequency independently of the other lanes.
struct Lane {
    hipStream_t s1 = nullptr, s2 = nullptr;     // main + LU lookahead stream
    std::vector<hipEvent_t> lu_events;
    hipEvent_t t[4] = {};                      // timing
    hipEvent_t done = nullptr;
    unsigned long long* counter = nullptr;     // triangular-solve ticket | epoch << 32
    unsigned long long* pub = nullptr;         // triangular-solve publication words
    RefineCtrl* ctrl = nullptr;
    LaneOut* hout = nullptr;                   // pinned
    LaneOut* dout = nullptr;
    LinSys sys;
    double *Sr = nullptr, *Si = nullptr;            // composed SMat
    double *PhiIr = nullptr, *PhiIi = nullptr, *DAr = nullptr, *DAi = nullptr;
    double *Xinr = nullptr, *Xini = nullptr;        // RHS input block
    double* d_betas = nullptr;
    std::vector<void*> owned;
    // bookkeeping of the frequency in flight
    int pending = -1;
    int mode = 0, prec = 64, nb = 0, rad_only = 0, compose = 0;
    int refine_max = kRefineMaxIter;
    int lu_arith = kLuArithSplitF16;
    int lu_latency = 1;    // set per run: 1 when this lane is the only one in flight
    // set per frequency by issue_frequency (other users of a lane keep the
    // full path): the factorisation without its pivoting fallback
    // (speculative), and the correction steps enqueued with the solve
    int lu_spec = 0;
    int refine_eager = kRefineMaxIter;
    double pivot_tau = 1.0;   // the pivot check's multiplier bound
    bool sur_used = false;
    double omega = 0, kappa = 0;
    // captured frequency graph and the configuration it was captured for
    hipGraphExec_t gexec = nullptr;
    GraphKey gkey{}, last_key{};
    bool graphed = false;

    ~Lane() {
        if (s1) (void)hipStreamSynchronize(s1);
        if (s2) (void)hipStreamSynchronize(s2);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (void* p : owned)
            if (p) (void)hipFree(p);
        if (hout) (void)hipHostFree(hout);
        for (auto& e : lu_events) (void)hipEventDestroy(e);
        for (auto& e : t)
            if (e) (void)hipEventDestroy(e);
        if (done) (void)hipEventDestroy(done);
        if (s2) (void)hipStreamDestroy(s2);
        if (s1) (void)hipStreamDestroy(s1);
    }
    template <typename T>
    int get(T** p, size_t n) {
        int r = dalloc_t(p, n);
        if (!r) owned.push_back(*p);
        return r;
    }
    int init(int N) {
        int r = BEM_OK;
        HIP_TRY(hipStreamCreateWithFlags(&s1, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        for (auto& e : t) HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        if ((r = sys.init(N))) return r;
        int Np = sys.Np;
        int nblk = Np / kLuNb;
        size_t npub = lu_solve_pub_words(Np);
        if ((r = get(&counter, 1)) || (r = get(&pub, npub)) || (r = get(&ctrl, 1)) || (r = get(&dout, 1))) return r;
        const unsigned long long first_epoch = 1ULL << 32;
        HIP_TRY(hipMemcpy(counter, &first_epoch, sizeof(first_epoch), hipMemcpyHostToDevice));
        HIP_TRY(hipMemset(pub, 0, npub * sizeof(unsigned long long)));
        HIP_TRY(hipHostMalloc((void**)&hout, sizeof(LaneOut), hipHostMallocDefault));
        size_t NN = (size_t)Np * Np, NR = (size_t)Np * kMaxRhs;
        if ((r = get(&Sr, NN)) || (r = get(&Si, NN))) return r;
        if ((r = get(&PhiIr, NR)) || (r = get(&PhiIi, NR)) || (r = get(&DAr, NR)) || (r = get(&DAi, NR)) ||
            (r = get(&Xinr, NR)) || (r = get(&Xini, NR)) || (r = get(&d_betas, kMaxRhs)))
            return r;
        for (size_t i = 0; i < 2 * (size_t)nblk + 2; i++) {
            hipEvent_t e;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            lu_events.push_back(e);
        }
        HIP_TRY(hipDeviceSynchronize());
        return BEM_OK;
    }
};

struct bem_body {
    bem_ctx* ctx = nullptr;
    int NPA = 0;                        // body panels; N - NPA lid panels follow them
    unsigned long long id = 0;          // unique per body (graph keys)
    int N = 0, Np = 0;
    BodyGeometry geo;
    double *d_body = nullptr, *d_refl = nullptr, *d_quad = nullptr, *d_recv = nullptr, *d_gen = nullptr;
    double *Ssrc = nullptr, *Srefl = nullptr, *Dsrc = nullptr, *Drefl = nullptr;
    bool have_rankine = false;
    long long nq_total = 0;             // quadrature nodes over all panels
    std::vector<Lane*> lanes;

    ~bem_body() {
        for (Lane* l : lanes) delete l;
        void* ps[] = {d_body, d_refl, d_quad, d_recv, d_gen, Ssrc, Srefl, Dsrc, Drefl};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
};

namespace {

// What a factorisation and solve run on: the streams (main, LU lookahead),
// the lookahead events (2 per panel + 2), the system workspace, the
// triangular-solve ticket counter and publication words, the refinement
// state and its step limit. A solve lane provides one (Lane::view); the
// no-allocation dense solve builds one over its fixed device array.
struct SolveView {
    hipStream_t s1 = nullptr, s2 = nullptr;
    hipEvent_t* ev = nullptr;
    LinSys* sys = nullptr;
    unsigned long long* counter = nullptr;
    unsigned long long* pub = nullptr;
    RefineCtrl* ctrl = nullptr;
    int refine_max = kRefineMaxIter;
    int lu_arith = kLuArithSplitF16;
    int lu_latency = 1;    // one frequency in flight (launch_lu's latency)
    int lu_spec = 0;       // speculative factorisation (kLuSpeculative): the caller re-runs flagged systems
    int refine_eager = kRefineMaxIter;   // correction steps enqueued with the solve (the rest: refine_tail)
    double pivot_tau = 1.0;              // the pivot check's multiplier bound
};

SolveView lane_view(Lane& L) {
    SolveView v;
    v.s1 = L.s1; v.s2 = L.s2; v.ev = L.lu_events.data(); v.sys = &L.sys;
    v.counter = L.counter; v.pub = L.pub; v.ctrl = L.ctrl; v.refine_max = L.refine_max; v.lu_arith = L.lu_arith;
    v.lu_latency = L.lu_latency;
    v.lu_spec = L.lu_spec;
    v.refine_eager = L.refine_eager;
    v.pivot_tau = L.pivot_tau;
    return v;
}

// Device LU in the given precision (FP64 in place, or FP32 on the FP32 copy).
void factor_system(const SolveView& V, int precision, LuProfile* prof = nullptr) {
    LinSys& s = *V.sys;
    const int spec = V.lu_spec ? kLuSpeculative : 0;
    if (precision == 64)
        launch_lu<double>(V.s1, V.s2, V.ev, s.Np, s.Dr, s.Di, s.w64, s.flag, s.ipiv, prof, kLuArithF32 | spec,
                          V.lu_latency, V.pivot_tau);
    else
        launch_lu<float>(V.s1, V.s2, V.ev, s.Np, s.D32r, s.D32i, s.w32, s.flag32, s.ipiv32, prof, V.lu_arith | spec,
                         V.lu_latency, V.pivot_tau);
}
void factor_system(Lane& L, int precision, LuProfile* prof = nullptr) { factor_system(lane_view(L), precision, prof); }

// Refinement iterations it0 .. it1 of an FP32 solve: each is a correction
// (it > 0: d = LU32^-1 r, x += d) followed by the residual r = b - A x in
// FP64 (the first also yields |A|_inf) and the convergence check; all of
// them are no-ops once the device marked the solve done.
// The residual launch also leaves fp32(r) in X32 (the correction solve's
// right-hand side), and the correction's upper-triangular solve adds d to x
// as it writes d.
void refine_steps(const SolveView& V, int nr, int it0, int it1) {
    LinSys& s = *V.sys;
    hipStream_t st = V.s1;
    const int* fany = V.lu_spec ? nullptr : s.flag32 + s.Np / kLuNb;
    const int* skip = &V.ctrl->done;
    const int max_iter = std::max(0, std::min(V.refine_max, kRefineMaxIter));
    for (int it = it0; it <= it1; it++) {
        if (it > 0)
            launch_lu_solve<float, float>(st, s.Np, nr, s.D32r, s.D32i, s.w32, s.ipiv32, fany, s.X32r, s.X32i,
                                          V.counter, V.pub, skip, s.Xr, s.Xi, 1);
        double* pabs = it == 0 ? s.Pabs : nullptr;
        launch_matvec_partials<double>(st, s.Np, nr, s.Dr, s.Di, s.Xr, s.Xi, s.Pr, s.Pi, pabs, skip);
        launch_refine_reduce(st, s.N, s.Np, nr, s.Pr, s.Pi, s.Br, s.Bi, s.Rr, s.Ri, s.X32r, s.X32i, s.Xr, s.Xi, pabs,
                             s.anorm, V.ctrl, kRefineTol, max_iter);
    }
}

// Iterations that V.refine_eager left for later (refine_tail): from
// min(refine_eager, max) + 1 to max.
int refine_last(const SolveView& V) { return std::max(0, std::min(V.refine_max, kRefineMaxIter)); }
int refine_first_tail(const SolveView& V) { return std::min(V.refine_eager, refine_last(V)) + 1; }

// Solve D X = B for nr columns (B in the workspace, X out). FP32 mode:
// FP32 LU solve followed by FP64 iterative refinement against the FP64
// matrix s.Dr/s.Di, controlled on the device (no host round trip), with at
// most V.refine_max correction steps, of which the first V.refine_eager are
// enqueued here (refine_tail enqueues the others).
void solve_system(const SolveView& V, int precision, int nr) {
    LinSys& s = *V.sys;
    hipStream_t st = V.s1;
    size_t n = (size_t)s.Np * nr;
    if (precision == 64) {
        launch_copy2(st, n, s.Br, s.Bi, s.Xr, s.Xi);
        launch_lu_solve<double, double>(st, s.Np, nr, s.Dr, s.Di, s.w64, s.ipiv,
                                        V.lu_spec ? nullptr : s.flag + s.Np / kLuNb, s.Xr, s.Xi, V.counter, V.pub);
        return;
    }
    // x = LU32^-1 b (the refinement state cleared on the way), into X32 and X
    launch_to_f32(st, n, s.Br, s.Bi, s.X32r, s.X32i, nullptr, V.ctrl);
    launch_lu_solve<float, float>(st, s.Np, nr, s.D32r, s.D32i, s.w32, s.ipiv32,
                                  V.lu_spec ? nullptr : s.flag32 + s.Np / kLuNb, s.X32r, s.X32i, V.counter, V.pub,
                                  nullptr, s.Xr, s.Xi, 0);
    refine_steps(V, nr, 0, refine_first_tail(V) - 1);
}
// The refinement iterations solve_system left out (refine_eager below the limit).
void refine_tail(const SolveView& V, int nr) { refine_steps(V, nr, refine_first_tail(V), refine_last(V)); }
void solve_system(Lane& L, int precision, int nr) { solve_system(lane_view(L), precision, nr); }

void planar_to_interleaved(const std::vector<double>& re, const std::vector<double>& im, int N, int Np, int ncol,
                           double* out) {
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < N; i++) {
            out[2 * ((size_t)j * N + i)] = re[(size_t)j * Np + i];
            out[2 * ((size_t)j * N + i) + 1] = im[(size_t)j * Np + i];
        }
}

}  // namespace

extern "C" {

const char* bem_last_error(void) { return g_err.c_str(); }

// ---- meshes -----------------------------------------------------------------
int bem_mesh_from_arrays(int n_nodes, const double* nodes, int n_elems, const int* elems, bem_mesh** out) {
    // an empty mesh (0 nodes, 0 elements) is allowed as a start for bem_mesh_append
    if (!out || n_nodes < 0 || n_elems < 0 || (n_nodes > 0 && !nodes) || (n_elems > 0 && !elems) ||
        (n_elems > 0 && n_nodes == 0))
        return fail(BEM_ERR_ARG, "bem_mesh_from_arrays: bad arguments");
    bem_mesh* m = new bem_mesh;
    m->m.nodes.assign(nodes, nodes + 3 * (size_t)n_nodes);
    m->m.elems.assign(elems, elems + 4 * (size_t)n_elems);
    for (int e = 0; e < n_elems; e++)
        for (int k = 0; k < 4; k++) {
            int v = elems[4 * e + k];
            if ((k < 3 && (v < 0 || v >= n_nodes)) || (k == 3 && v >= n_nodes)) {
                delete m;
                return fail(BEM_ERR_ARG, "bem_mesh_from_arrays: element " + std::to_string(e) + " node index out of range");
            }
        }
    *out = m;
    return BEM_OK;
}

static int wrap_mesh(Mesh&& mm, bem_mesh** out);

int bem_mesh_read(const char* path, int format, bem_mesh** out) {
    if (!path || !out) return fail(BEM_ERR_ARG, "bem_mesh_read: bad arguments");
    std::string p(path);
    if (format == BEM_MESH_AUTO) {
        std::string ext = p.size() >= 4 ? p.substr(p.size() - 4) : "";
        for (auto& ch : ext) ch = (char)std::tolower((unsigned char)ch);
        format = ext == ".gdf" ? BEM_MESH_GDF : ext == ".mar" ? BEM_MESH_MAR : ext == ".pnl" ? BEM_MESH_PNL
               : ext == ".stl" ? BEM_MESH_STL : -1;
        if (format < 0) return fail(BEM_ERR_ARG, "bem_mesh_read: unknown extension of " + p);
    }
    std::ifstream in(p);
    if (!in.is_open()) return fail(BEM_ERR_ARG, "bem_mesh_read: cannot open " + p);
    Mesh m;
    std::string err;
    bool ok = format == BEM_MESH_GDF ? read_gdf(in, m, err) : format == BEM_MESH_MAR ? read_mar(in, m, err)
            : format == BEM_MESH_PNL ? read_pnl(in, m, err) : format == BEM_MESH_STL ? read_stl(in, m, err)
            : (err = "unknown format", false);
    if (!ok) return fail(BEM_ERR_ARG, "bem_mesh_read: " + p + ": " + err);
    return wrap_mesh(std::move(m), out);
}

static int wrap_mesh(Mesh&& mm, bem_mesh** out) {
    if (!out) return fail(BEM_ERR_ARG, "null output");
    if (mm.n_elems() <= 0) return fail(BEM_ERR_ARG, "mesh generator produced no panels");
    bem_mesh* m = new bem_mesh;
    m->m = std::move(mm);
    *out = m;
    return BEM_OK;
}

int bem_mesh_half_cylinder(double radius, double draft, int n_radial, int n_vertical, int n_azimuthal, int cosine,
                           bem_mesh** out) {
    if (!(radius > 0) || !(draft > 0) || n_radial < 1 || n_vertical < 1 || n_azimuthal < 3)
        return fail(BEM_ERR_ARG, "bem_mesh_half_cylinder: bad dimensions");
    return wrap_mesh(mesh_half_cylinder(radius, draft, n_radial, n_vertical, n_azimuthal, cosine != 0), out);
}

int bem_mesh_lid_disc(double radius, int n_radial, int n_azimuthal, int cosine, bem_mesh** out) {
    if (!out || !(radius > 0) || n_radial < 1 || n_azimuthal < 3) return fail(BEM_ERR_ARG, "bem_mesh_lid_disc: bad arguments");
    return wrap_mesh(mesh_lid_disc(radius, n_radial, n_azimuthal, cosine != 0), out);
}

int bem_mesh_lid_rect(double L, double W, int nx, int ny, int cosine, bem_mesh** out) {
    if (!out || !(L > 0) || !(W > 0) || nx < 1 || ny < 1) return fail(BEM_ERR_ARG, "bem_mesh_lid_rect: bad arguments");
    return wrap_mesh(mesh_lid_rect(L, W, nx, ny, cosine != 0), out);
}

int bem_mesh_lid_ellipse(double a, double b, int n_radial, int n_azimuthal, int cosine, bem_mesh** out) {
    if (!out || !(a > 0) || !(b > 0) || n_radial < 1 || n_azimuthal < 3)
        return fail(BEM_ERR_ARG, "bem_mesh_lid_ellipse: bad arguments");
    return wrap_mesh(mesh_lid_ellipse(a, b, n_radial, n_azimuthal, cosine != 0), out);
}

int bem_mesh_revolution(int n, const double* r, const double* z, int n_azimuthal, bem_mesh** out) {
    if (!out || !r || !z || n < 2 || n_azimuthal < 3 || r[0] != 0.0)
        return fail(BEM_ERR_ARG, "bem_mesh_revolution: need >= 2 profile points starting on the axis (r = 0)");
    return wrap_mesh(mesh_revolution(std::vector<double>(r, r + n), std::vector<double>(z, z + n), n_azimuthal), out);
}

int bem_mesh_tapered_spar(double r_lower, double r_upper, double draft, double h_taper_begin, double h_taper,
                          int n_radial, int n_vert1, int n_vert2, int n_vert3, int n_azimuthal, int cosine,
                          bem_mesh** out) {
    if (!out || !(r_lower > 0) || !(r_upper > 0) || !(draft > 0) || h_taper_begin < 0 || h_taper < 0 ||
        h_taper_begin + h_taper > draft || n_radial < (cosine ? 2 : 1) || n_vert1 < 0 || n_vert2 < 0 || n_vert3 < 1 ||
        n_azimuthal < 3)
        return fail(BEM_ERR_ARG, "bem_mesh_tapered_spar: bad dimensions");
    return wrap_mesh(mesh_tapered_spar(r_lower, r_upper, draft, h_taper_begin, h_taper, n_radial, n_vert1, n_vert2,
                                       n_vert3, n_azimuthal, cosine != 0), out);
}

int bem_mesh_spar_leg(double r_lower, double r_upper, double draft, double h_plate, int n_base, int n_plate_vert,
                      int n_plate_horiz, int n_vert, int n_azimuthal, int cosine, bem_mesh** out) {
    if (!out || !(r_lower > 0) || !(r_upper > 0) || !(draft > 0) || h_plate < 0 || h_plate >= draft || n_base < 1 ||
        n_plate_vert < 0 || n_plate_horiz < 0 || n_vert < 1 || n_azimuthal < 3)
        return fail(BEM_ERR_ARG, "bem_mesh_spar_leg: bad dimensions");
    return wrap_mesh(mesh_spar_leg(r_lower, r_upper, draft, h_plate, n_base, n_plate_vert, n_plate_horiz, n_vert,
                                   n_azimuthal, cosine != 0), out);
}

int bem_mesh_barge(double L, double W, double D, int nx, int ny, int nz, int cosine, bem_mesh** out) {
    if (!out || !(L > 0) || !(W > 0) || !(D > 0) || nx < 1 || ny < 1 || nz < 1)
        return fail(BEM_ERR_ARG, "bem_mesh_barge: bad dimensions");
    return wrap_mesh(mesh_barge(L, W, D, nx, ny, nz, cosine != 0), out);
}

int bem_mesh_ellipsoid(double a, double b, double c, double depth, int n_axial, int n_azimuthal, int cosine,
                       int triangles, bem_mesh** out) {
    if (!out || !(a > 0) || !(b > 0) || !(c > 0) || !(depth > c) || n_axial < 2 || n_azimuthal < 3)
        return fail(BEM_ERR_ARG, "bem_mesh_ellipsoid: bad dimensions (the ellipsoid must lie below z = 0)");
    return wrap_mesh(mesh_ellipsoid(a, b, c, -depth, n_axial, n_azimuthal, cosine != 0, triangles != 0), out);
}

int bem_mesh_semi_ellipsoid(double a, double b, double c, int n_axial, int n_azimuthal, int cosine, int triangles,
                            bem_mesh** out) {
    if (!out || !(a > 0) || !(b > 0) || !(c > 0) || n_axial < 1 || n_azimuthal < 3)
        return fail(BEM_ERR_ARG, "bem_mesh_semi_ellipsoid: bad dimensions");
    return wrap_mesh(mesh_semi_ellipsoid(a, b, c, n_axial, n_azimuthal, cosine != 0, triangles != 0), out);
}

int bem_mesh_wigley(double L, double B, double T, int nx, int nz, int cosine, int triangles, int outward,
                    bem_mesh** out) {
    if (!out || !(L > 0) || !(B > 0) || !(T > 0) || nx < 1 || nz < 1)
        return fail(BEM_ERR_ARG, "bem_mesh_wigley: bad dimensions");
    return wrap_mesh(mesh_wigley(L, B, T, nx, nz, cosine != 0, triangles != 0, outward != 0), out);
}

int bem_mesh_append(bem_mesh* dst, const bem_mesh* src, double dx, double dy, double dz) {
    if (!dst || !src || dst == src) return fail(BEM_ERR_ARG, "bem_mesh_append: bad arguments");
    dst->m.append(src->m, dx, dy, dz);
    return BEM_OK;
}

int bem_mesh_box(double L, double W, double z0, double z1, int nx, int ny, int nz, int open_top, bem_mesh** out) {
    if (!(L > 0) || !(W > 0) || !(z1 > z0) || nx < 1 || ny < 1 || nz < 1)
        return fail(BEM_ERR_ARG, "bem_mesh_box: bad dimensions");
    return wrap_mesh(mesh_box(L, W, z0, z1, nx, ny, nz, open_top != 0), out);
}

int bem_mesh_semisub3(double col_radius, double draft, double spacing, double pontoon_width, double pontoon_height,
                      double gap, int n_azimuthal, int n_vertical, int n_radial, int nx, int ny, int nz,
                      bem_mesh** out) {
    if (!(col_radius > 0) || !(draft > pontoon_height) || !(spacing > 2 * (col_radius + gap)) || n_azimuthal < 3)
        return fail(BEM_ERR_ARG, "bem_mesh_semisub3: bad dimensions");
    return wrap_mesh(mesh_semisub3(col_radius, draft, spacing, pontoon_width, pontoon_height, gap, n_azimuthal,
                                   n_vertical, n_radial, nx, ny, nz),
                     out);
}

int bem_mesh_cone_spar(double r_top, double r_bot, double draft, double r_plate, double t_plate, int n_azimuthal,
                       int n_disc, int n_rim, int n_annulus, int n_cone, bem_mesh** out) {
    if (!(r_top > 0) || !(r_bot > 0) || !(draft > 0) || !(r_plate > r_bot) || !(t_plate > 0) || n_azimuthal < 3 ||
        n_disc < 1 || n_rim < 1 || n_annulus < 1 || n_cone < 1)
        return fail(BEM_ERR_ARG, "bem_mesh_cone_spar: bad dimensions");
    return wrap_mesh(mesh_cone_spar(r_top, r_bot, draft, r_plate, t_plate, n_azimuthal, n_disc, n_rim, n_annulus, n_cone),
                     out);
}

int bem_mesh_random_body(uint64_t seed, int index, double* params, bem_mesh** out) {
    if (index < 0) return fail(BEM_ERR_ARG, "bem_mesh_random_body: negative index");
    return wrap_mesh(mesh_random_body(seed, index, params), out);
}

int bem_mesh_size(const bem_mesh* m, int* n_nodes, int* n_elems) {
    if (!m) return fail(BEM_ERR_ARG, "null mesh");
    if (n_nodes) *n_nodes = m->m.n_nodes();
    if (n_elems) *n_elems = m->m.n_elems();
    return BEM_OK;
}

int bem_mesh_export(const bem_mesh* m, double* nodes, int* elems) {
    if (!m) return fail(BEM_ERR_ARG, "null mesh");
    if (nodes) std::memcpy(nodes, m->m.nodes.data(), m->m.nodes.size() * sizeof(double));
    if (elems) std::memcpy(elems, m->m.elems.data(), m->m.elems.size() * sizeof(int));
    return BEM_OK;
}

void bem_mesh_free(bem_mesh* m) { delete m; }

int bem_geometry(const bem_mesh* m, double* body, double* refl, double* quad, double* recv, double* gen) {
    if (!m) return fail(BEM_ERR_ARG, "null mesh");
    BodyGeometry g;
    std::string err;
    if (!build_geometry(m->m, g, err)) return fail(BEM_ERR_ARG, err);
    if (body) std::memcpy(body, g.body.data(), g.body.size() * sizeof(double));
    if (refl) std::memcpy(refl, g.refl.data(), g.refl.size() * sizeof(double));
    if (quad) std::memcpy(quad, g.quad.data(), g.quad.size() * sizeof(double));
    if (recv) std::memcpy(recv, g.recv.data(), g.recv.size() * sizeof(double));
    if (gen) std::memcpy(gen, g.gen.data(), g.gen.size() * sizeof(double));
    return BEM_OK;
}

// ---- context -----------------------------------------------------------------
// ---------------------------------------------------------------------------
// No-allocation dense solve: every device buffer of bem_dense_solve_noalloc
// lives in one fixed array per GPU, `scratch`. Host code never uses the
// array's name as an address: at set-up (bem_create, once per device, with
// that device current) hipGetSymbolAddress yields the device's own address of
// it, kept per device in g_na_ptr, and every kernel argument and copy goes
// through that pointer.
// ---------------------------------------------------------------------------
namespace {

constexpr int kNaRhs = kMaxRhs;
constexpr int kNaKslabs = 4;                       // matvec_kslabs(Np) for Np < 1024
constexpr size_t kNaAlign = 256;

// Byte offsets of the buffers in the array, laid out for the largest order
// (smaller systems use the same places with their own leading dimension).
struct NaLayout {
    size_t Dr, Di, D32r, D32i, ipiv, flag, ipiv32, flag32;
    size_t w64[12], w32[12], umax;
    size_t Br, Bi, Xr, Xi, Rr, Ri, X32r, X32i, Pr, Pi, Pabs, anorm;
    size_t counter, pub, ctrl, stageA, stageB;
    size_t total;
};

constexpr NaLayout na_layout() {
    NaLayout L{};
    size_t o = 0;
    auto take = [&o](size_t bytes) {
        size_t at = o;
        o = (o + bytes + kNaAlign - 1) / kNaAlign * kNaAlign;
        return at;
    };
    const size_t NN = (size_t)kNaNp * kNaNp, NR = (size_t)kNaNp * kNaRhs, pan = (size_t)kNaNp * kLuNb;
    const size_t nflag = kNaNp / kLuNb + 1;
    L.Dr = take(NN * 8); L.Di = take(NN * 8);
    L.D32r = take(NN * 4); L.D32i = take(NN * 4);
    L.ipiv = take(kNaNp * 4); L.flag = take(nflag * 4);
    L.ipiv32 = take(kNaNp * 4); L.flag32 = take(nflag * 4);
    // LuWork: sd_r, sd_i (NB x NB), then s21, Linv, Uinv, PL, PU (Np x NB) re and im
    L.w64[0] = take(kLuNb * kLuNb * 8); L.w64[1] = take(kLuNb * kLuNb * 8);
    for (int k = 2; k < 12; k++) L.w64[k] = take(pan * 8);
    L.w32[0] = take(kLuNb * kLuNb * 4); L.w32[1] = take(kLuNb * kLuNb * 4);
    for (int k = 2; k < 12; k++) L.w32[k] = take(pan * 4);
    L.umax = take((size_t)(kNaNp / kLuNb) * 4);
    L.Br = take(NR * 8); L.Bi = take(NR * 8); L.Xr = take(NR * 8); L.Xi = take(NR * 8);
    L.Rr = take(NR * 8); L.Ri = take(NR * 8);
    L.X32r = take(NR * 4); L.X32i = take(NR * 4);
    L.Pr = take(kNaKslabs * NR * 8); L.Pi = take(kNaKslabs * NR * 8); L.Pabs = take((size_t)kNaKslabs * kNaNp * 8);
    L.anorm = take(8);
    L.counter = take(8);
    L.pub = take((size_t)(kNaNp / kLuNb) * kNaRhs * kLuNb * 2 * 2 * 8);   // lu_solve_pub_words(kNaNp) words
    L.ctrl = take(sizeof(RefineCtrl));
    L.stageA = take((size_t)2 * kNaMaxN * kNaMaxN * 8);                    // interleaved A as the caller passes it
    L.stageB = take((size_t)2 * kNaMaxN * kNaRhs * 8);                     // interleaved B in, X out
    L.total = o;
    return L;
}

constexpr size_t kNaBytes = BEM_NOALLOC_SCRATCH_BYTES;
static_assert(na_layout().total <= kNaBytes, "the no-allocation layout must fit its array");

constexpr int kNaMaxDev = 64;
std::mutex g_na_mu;
void* g_na_ptr[kNaMaxDev] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {

// Once per device, with the device current: the array's address on this
// device, and the triangular-solve state a new solve lane starts from.
int noalloc_setup(bem_ctx* c) {
    if (c->device < 0 || c->device >= kNaMaxDev) return fail(BEM_ERR_ARG, "bem_create: device index too large");
    HIP_TRY(hipStreamCreateWithFlags(&c->na_s2, hipStreamNonBlocking));
    for (auto& e : c->na_ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : c->na_t) HIP_TRY(hipEventCreate(&e));
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (g_na_ptr[c->device]) return BEM_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(BEM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    constexpr NaLayout L = na_layout();
    char* base = static_cast<char*>(p);
    launch_solve_state_init(c->stream, reinterpret_cast<unsigned long long*>(base + L.counter),
                            reinterpret_cast<unsigned long long*>(base + L.pub), lu_solve_pub_words(kNaNp));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    g_na_ptr[c->device] = p;
    return BEM_OK;
}

}  // namespace

int bem_create(int device, bem_ctx** out) {
    if (!out) return fail(BEM_ERR_ARG, "null output");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(BEM_ERR_ARG, "bem_create: device " + std::to_string(device) + " not present");
    HIP_TRY(hipSetDevice(device));
    bem_ctx* c = new bem_ctx;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(BEM_ERR_HIP, "hipStreamCreate failed");
    }
    for (auto& e : c->ev) (void)hipEventCreate(&e);
    int r = noalloc_setup(c);
    if (r) { bem_destroy(c); return r; }
    *out = c;
    return BEM_OK;
}

void bem_destroy(bem_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    for (Lane* l : c->pool) delete l;
    c->pool.clear();
    delete c->surrogate;
    c->surrogate = nullptr;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->na_s2) (void)hipStreamSynchronize(c->na_s2);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->na_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->na_t)
        if (e) (void)hipEventDestroy(e);
    if (c->na_s2) (void)hipStreamDestroy(c->na_s2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int bem_body_create(bem_ctx* c, const bem_mesh* m, bem_body** out) { return bem_body_create_lid(c, m, nullptr, out); }

int bem_body_create_lid(bem_ctx* c, const bem_mesh* m, const bem_mesh* lid, bem_body** out) {
    if (!c || !m || !out) return fail(BEM_ERR_ARG, "bem_body_create: null argument");
    if (m->m.n_elems() < 1) return fail(BEM_ERR_ARG, "bem_body_create: empty mesh");
    HIP_TRY(hipSetDevice(c->device));
    bem_body* b = new bem_body;
    static std::atomic<unsigned long long> next_body_id{1};
    b->id = next_body_id++;
    b->ctx = c;
    std::string err;
    // Create_Panels (Hydrodynamic_Solver.cpp:30-81): body panels first, then
    // the auxiliary (lid) panels when irregular-frequency removal is on
    Mesh all = m->m;
    if (lid) {
        for (int k = 0; k < lid->m.n_nodes(); k++)
            if (std::fabs(lid->m.nodes[3 * (size_t)k + 2]) > 1e-9) {
                delete b;
                return fail(BEM_ERR_ARG, "bem_body_create_lid: lid nodes must lie in z = 0");
            }
        all.append(lid->m, 0.0, 0.0, 0.0);
    }
    if (!build_geometry(all, b->geo, err)) { delete b; return fail(BEM_ERR_ARG, err); }
    b->N = b->geo.N;
    b->NPA = m->m.n_elems();
    b->Np = round_up(b->N, kPad);
    for (int i = 0; i < b->N; i++) b->nq_total += (long long)b->geo.quad[(size_t)i * QF_STRIDE + QF_NQ];
    if (b->Np / kLuNb > kMaxPanels) { delete b; return fail(BEM_ERR_ARG, "body too large"); }
    int N = b->N;
    int r = BEM_OK;
    if ((r = dalloc_t(&b->d_body, (size_t)N * 32)) || (r = dalloc_t(&b->d_refl, (size_t)N * 32)) ||
        (r = dalloc_t(&b->d_quad, (size_t)N * 32)) || (r = dalloc_t(&b->d_recv, (size_t)N * 4)) ||
        (r = dalloc_t(&b->d_gen, (size_t)N * 8))) {
        delete b;
        return r;
    }
    hipStream_t s = c->stream;
    if (hipMemcpyAsync(b->d_body, b->geo.body.data(), b->geo.body.size() * 8, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(b->d_refl, b->geo.refl.data(), b->geo.refl.size() * 8, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(b->d_quad, b->geo.quad.data(), b->geo.quad.size() * 8, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(b->d_recv, b->geo.recv.data(), b->geo.recv.size() * 8, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(b->d_gen, b->geo.gen.data(), b->geo.gen.size() * 8, hipMemcpyHostToDevice, s) ||
        hipStreamSynchronize(s)) {
        delete b;
        return fail(BEM_ERR_HIP, "geometry upload failed");
    }
    *out = b;
    return BEM_OK;
}

void bem_body_destroy(bem_body* b) {
    if (!b) return;
    if (b->ctx) { (void)hipSetDevice(b->ctx->device); (void)hipStreamSynchronize(b->ctx->stream); }
    delete b;
}

int bem_body_n(const bem_body* b) { return b ? b->NPA : -1; }
int bem_body_n_system(const bem_body* b) { return b ? b->N : -1; }

int bem_body_assemble_rankine(bem_ctx* c, bem_body* b) {
    if (!c || !b || b->ctx != c) return fail(BEM_ERR_ARG, "bem_body_assemble_rankine: bad arguments");
    HIP_TRY(hipSetDevice(c->device));
    size_t NN = (size_t)b->Np * b->Np;
    int r = BEM_OK;
    if (!b->Ssrc) {
        if ((r = dalloc_t(&b->Ssrc, NN)) || (r = dalloc_t(&b->Srefl, NN)) || (r = dalloc_t(&b->Dsrc, NN)) ||
            (r = dalloc_t(&b->Drefl, NN)))
            return r;
    }
    launch_rankine(c->stream, b->N, b->Np, b->d_body, b->d_refl, b->d_recv, b->Ssrc, b->Srefl, b->Dsrc, b->Drefl);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    b->have_rankine = true;
    return BEM_OK;
}

namespace {

int ensure_lanes(bem_body* b, int n, bool need32) {
    while ((int)b->lanes.size() < n) {
        Lane* L = new Lane;
        int r = L->init(b->N);
        if (r) { delete L; return r; }
        b->lanes.push_back(L);
    }
    if (need32)
        for (int i = 0; i < n; i++) {
            int r = b->lanes[i]->sys.init32();
            if (r) return r;
        }
    // GEMM tile autotuning for this padded order (once per device, precision
    // and order), on lane 0's matrix buffers before any frequency uses them
    Lane& L0 = *b->lanes[0];
    if (L0.pending < 0) {
        tune_lu_gemms<double>(L0.s1, b->Np, L0.sys.Dr, L0.sys.Di);
        if (need32)
            tune_lu_gemms<float>(L0.s1, b->Np, L0.sys.D32r, L0.sys.D32i, L0.sys.w32.umax, L0.sys.w32.opA,
                                 L0.sys.w32.opB);
        HIP_TRY(hipStreamSynchronize(L0.s1));
    }
    return BEM_OK;
}

// Copies the frequency (omega, kappa) to the lane's device block; kernels
// read it there. The host value must stay until the copy has run (the lane's
// next frequency is set only after its previous one completed).
void set_frequency(Lane& L, double omega, double kappa) {
    L.hout->freq[0] = omega;
    L.hout->freq[1] = kappa;
    (void)hipMemcpyAsync(L.dout->freq, L.hout->freq, sizeof(L.hout->freq), hipMemcpyHostToDevice, L.s1);
}

// Work-groups per CU of the body's wave assembly (0: one per tile).
int wave_cap(const bem_body* b) {
    const int c = b->ctx->wave_wg_per_cu;
    return c >= 0 ? c : (b->Np < kWaveCapNp ? kWaveCapWgPerCu : 0);
}

// Wave assembly + compose at the lane's device frequency; kappa (the same
// value) selects the surrogate path on the host.
void assemble(bem_body* b, Lane& L, int mode, double kappa, int compose) {
    bool f32 = (mode != BEM_MODE_FP64);
    L.sur_used = mode == BEM_MODE_MIXED_SURROGATE && compose == BEM_COMPOSE_WAVE && kappa != 0.0;
    if (L.sur_used) {
        launch_sur_counters_clear(L.s1, &L.dout->sur);
        launch_wave_compose_surr(L.s1, b->ctx->surrogate->params, b->N, b->Np, L.dout->freq, b->d_quad, b->d_recv, b->Ssrc,
                                 b->Srefl, b->Dsrc, b->Drefl, L.Sr, L.Si, L.sys.Dr, L.sys.Di, L.sys.D32r, L.sys.D32i,
                                 &L.dout->sur);
        return;
    }
    launch_wave_compose(L.s1, b->N, b->Np, L.dout->freq, compose, b->d_quad, b->d_recv, b->Ssrc, b->Srefl, b->Dsrc, b->Drefl,
                        L.Sr, L.Si, L.sys.Dr, L.sys.Di, f32 ? L.sys.D32r : nullptr, f32 ? L.sys.D32i : nullptr,
                        b->ctx->wave_far_r2, wave_cap(b));
}

int enqueue_results(bem_body* b, Lane& L, bool timing);

// Everything one frequency puts on the lane's streams, from the frequency
// copy to the result copies. timing: per-stage events (not inside a graph).
int enqueue_frequency(bem_body* b, Lane& L, bool timing) {
    LinSys& s = L.sys;
    hipStream_t st = L.s1;
    const int nb = L.nb, nr = 6 + nb;
    (void)hipMemcpyAsync(L.dout->freq, L.hout->freq, sizeof(L.hout->freq), hipMemcpyHostToDevice, st);
    if (timing) HIP_TRY(hipEventRecord(L.t[0], st));
    assemble(b, L, L.mode, L.kappa, L.compose);
    if (timing) HIP_TRY(hipEventRecord(L.t[1], st));
    factor_system(L, L.prec);
    if (timing) HIP_TRY(hipEventRecord(L.t[2], st));
    launch_incident(st, b->N, b->Np, nb, b->d_body, b->d_gen, L.d_betas, L.dout->freq, L.Xinr, L.Xini, kMaxRhs,
                    L.PhiIr, L.PhiIi, L.DAr, L.DAi);
    launch_matvec<double>(st, b->Np, nr, L.Sr, L.Si, L.Xinr, L.Xini, s.Pr, s.Pi, nullptr, nullptr, 0, s.Br, s.Bi);
    solve_system(L, L.prec, nr);
    return enqueue_results(b, L, timing);
}

// The reductions and the result copies of the lane's frequency.
int enqueue_results(bem_body* b, Lane& L, bool timing) {
    LinSys& s = L.sys;
    hipStream_t st = L.s1;
    const int nb = L.nb;
    // reductions over the body panels only (Prepare_PostProcessing_Mats :145-165, potentials cut to NPA :450-458)
    launch_reduce_coeffs(st, b->NPA, b->Np, nb, b->d_gen, s.Xr, s.Xi, L.PhiIr, L.PhiIi, L.DAr, L.DAi, L.dout->coeffs);
    if (timing) HIP_TRY(hipEventRecord(L.t[3], st));
    int np = b->Np / kLuNb;
    HIP_TRY(hipMemcpyAsync(L.hout->coeffs, L.dout->coeffs, sizeof(double) * 2 * (36 + 12 * nb), hipMemcpyDeviceToHost, st));
    if (L.prec == 32)
        HIP_TRY(hipMemcpyAsync(&L.hout->ctrl, L.ctrl, sizeof(RefineCtrl), hipMemcpyDeviceToHost, st));
    if (L.sur_used)
        HIP_TRY(hipMemcpyAsync(&L.hout->sur, &L.dout->sur, sizeof(SurCounters), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(L.hout->flags, L.prec == 64 ? s.flag : s.flag32, sizeof(int) * (np + 1),
                           hipMemcpyDeviceToHost, st));
    return BEM_OK;
}

// Issues one frequency on a lane (all asynchronous on the lane's streams).
// The betas must already be on the device (L.d_betas). With graphs enabled
// on the context, the second consecutive frequency of the same configuration
// (body, mode, headings, compose case, surrogate use) on a lane captures the
// enqueue sequence into a graph, and later ones replay it: the frequency
// reaches the kernels through the captured host-to-device copy, and every
// other input is at a fixed device address.
// pivot: factor with the pivoting fallback inside the LU (a re-run of a
// frequency whose speculative factorisation flagged a panel); otherwise the
// LU is speculative and harvest re-runs flagged frequencies this way.
int issue_frequency(bem_body* b, Lane& L, int index, int mode, double omega, double kappa, int compose, int rad_only,
                    int nb, bool pivot = false) {
    if (b->Np > L.sys.capNp) return fail(BEM_ERR_STATE, "lane workspace smaller than the body");
    L.lu_spec = (pivot || !b->ctx->lu_spec) ? 0 : 1;
    L.refine_eager = b->ctx->refine_eager;
    L.pivot_tau = b->ctx->pivot_tau;
    LinSys& s = L.sys;
    s.N = b->N;
    s.Np = b->Np;
    hipStream_t st = L.s1;
    L.pending = index;
    L.mode = mode; L.prec = (mode == BEM_MODE_FP64) ? 64 : 32;
    L.omega = omega; L.kappa = kappa; L.compose = compose; L.rad_only = rad_only; L.nb = nb;
    L.refine_max = b->ctx->refine_max;
    L.lu_arith = b->ctx->lu_arith;
    L.hout->freq[0] = omega;
    L.hout->freq[1] = kappa;
    const bool sur = mode == BEM_MODE_MIXED_SURROGATE && compose == BEM_COMPOSE_WAVE && kappa != 0.0;
    const GraphKey key{b->id, b->Np, mode, nb, rad_only, compose, sur ? 1 : 0, L.refine_max, L.lu_arith,
                       wave_cap(b), b->ctx->wave_far_r2};
    const bool use_graph = !pivot && b->ctx->graphs && (L.gexec && L.gkey == key ? true : L.last_key == key);
    L.last_key = key;
    L.graphed = use_graph;
    if (!use_graph) {
        int r = enqueue_frequency(b, L, true);
        if (r) return r;
    } else {
        if (!(L.gexec && L.gkey == key)) {
            if (L.gexec) { (void)hipGraphExecDestroy(L.gexec); L.gexec = nullptr; }
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            int r = enqueue_frequency(b, L, false);
            hipError_t e = hipStreamEndCapture(st, &g);
            if (r) { if (g) (void)hipGraphDestroy(g); return r; }
            if (e != hipSuccess) return fail(BEM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&L.gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { L.gexec = nullptr; return fail(BEM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e)); }
            L.gkey = key;
        }
        L.sur_used = sur;
        HIP_TRY(hipEventRecord(L.t[0], st));
        HIP_TRY(hipGraphLaunch(L.gexec, st));
        HIP_TRY(hipEventRecord(L.t[3], st));
    }
    HIP_TRY(hipEventRecord(L.done, st));
    HIP_TRY(hipGetLastError());
    return BEM_OK;
}

// Waits for the lane's frequency and writes its results (FP64 re-solve if
// the refinement did not converge). BEM_ERR_NUMERIC, with the results still
// written, when the final factorisation met a zero or non-finite pivot or a
// result is not finite (the reference's Eigen LU would return the same NaN
// silently).
int harvest(bem_body* b, Lane& L, double* A36, double* B36, double* F, double* FK, double* SC, bem_solve_info* info) {
    HIP_TRY(hipEventSynchronize(L.done));
    int np = b->Np / kLuNb;
    if (L.lu_spec) {
        // a panel of the speculative factorisation wanted interchanges: the
        // frequency again, with the pivoting fallback
        bool flagged = false;
        for (int i = 0; i < np; i++) flagged = flagged || (L.hout->flags[i] & 1);
        if (flagged) {
            int r = issue_frequency(b, L, L.pending, L.mode, L.omega, L.kappa, L.compose, L.rad_only, L.nb, true);
            if (r) return r;
            HIP_TRY(hipEventSynchronize(L.done));
        }
    }
    if (L.prec == 32 && !L.hout->ctrl.done) {
        // the refinement needs the correction steps not enqueued with it
        refine_tail(lane_view(L), 6 + L.nb);
        int r = enqueue_results(b, L, !L.graphed);
        if (r) return r;
        HIP_TRY(hipEventRecord(L.done, L.s1));
        HIP_TRY(hipEventSynchronize(L.done));
    }
    bem_solve_info loc{};
    int fell_back = 0;
    if (L.prec == 32) {
        loc.refine_iters = L.hout->ctrl.iters;
        loc.residual = L.hout->ctrl.rel;
        if (!L.hout->ctrl.converged) {
            // the FP64 re-solve takes the checked factorisation (its flags
            // are not looked at again)
            int r = issue_frequency(b, L, L.pending, BEM_MODE_FP64, L.omega, L.kappa, L.compose, L.rad_only, L.nb,
                                    true);
            if (r) return r;
            HIP_TRY(hipEventSynchronize(L.done));
            fell_back = 1;
        }
    }
    loc.fell_back_fp64 = fell_back;
    if (L.sur_used) {
        loc.surrogate_points = (int64_t)L.hout->sur.n_sur;
        loc.exact_points = (int64_t)L.hout->sur.n_exact;
        float e = 0.0f;
        std::memcpy(&e, &L.hout->sur.err_bits, sizeof(e));
        loc.surrogate_err = e;
    } else if (L.compose == BEM_COMPOSE_WAVE && L.kappa != 0.0) {
        loc.exact_points = (int64_t)b->N * b->nq_total;
    }
    for (int i = 0; i < np; i++) loc.pivot_fallback += (L.hout->flags[i] & 1) ? 1 : 0;
    float t01 = 0, t12 = 0, t23 = 0;
    if (L.graphed) {   // stages are not timed inside a graph: the whole frequency is in ms_solve
        (void)hipEventElapsedTime(&t23, L.t[0], L.t[3]);
    } else {
        (void)hipEventElapsedTime(&t01, L.t[0], L.t[1]);
        (void)hipEventElapsedTime(&t12, L.t[1], L.t[2]);
        (void)hipEventElapsedTime(&t23, L.t[2], L.t[3]);
    }
    loc.ms_assemble = t01; loc.ms_factor = t12; loc.ms_solve = t23;
    const double* out = L.hout->coeffs;
    int nb = L.nb;
    double omega = L.omega;
    // A = Re(M Phi_J), B = Im(M Phi_J) (Hydrodynamic_Solver.cpp:463-465)
    for (int q = 0; q < 36; q++) { A36[q] = out[2 * q]; B36[q] = out[2 * q + 1]; }
    if (!L.rad_only && nb > 0) {
        typedef std::complex<double> cd;
        const cd Im(0.0, 1.0);
        for (int i = 0; i < 6; i++)
            for (int bb = 0; bb < nb; bb++) {
                int q1 = 36 + i * nb + bb, q2 = 36 + 6 * nb + i * nb + bb;
                cd F1(out[2 * q1], out[2 * q1 + 1]), F2(out[2 * q2], out[2 * q2 + 1]);
                cd fk = -Im * omega * F1, sc = Im * omega * F2;   // :474-475
                cd f = std::conj(fk + sc);                          // :476-477
                if (omega == 0.0) { f = 0; fk = 0; }                // :480-483
                size_t o = 2 * ((size_t)i * nb + bb);
                if (F) { F[o] = f.real(); F[o + 1] = f.imag(); }
                if (FK) { FK[o] = fk.real(); FK[o + 1] = fk.imag(); }
                if (SC) { SC[o] = sc.real(); SC[o + 1] = sc.imag(); }
            }
    }
    if (info) *info = loc;
    L.pending = -1;
    const bool singular = (L.hout->flags[np] & 2) != 0;
    bool finite = true;
    const int ncoef = 2 * (36 + ((L.rad_only || nb == 0) ? 0 : 12 * nb));
    for (int q = 0; q < ncoef; q++) finite = finite && std::isfinite(out[q]);
    if (singular || !finite) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "%s at omega = %.17g", singular ?
                      "numerically singular system (zero or non-finite pivot)" : "non-finite A, B or F", omega);
        return fail(BEM_ERR_NUMERIC, msg);
    }
    return BEM_OK;
}

// Frequencies: omega may be negative (the reference's set-up solve runs at
// Frequency = -1), but must be finite; kappa = omega^2 / g must be finite
// and >= 0. Anything else would run the whole path on NaN.
bool freq_ok(double omega, double kappa) { return std::isfinite(omega) && std::isfinite(kappa) && kappa >= 0.0; }

int check_solve_args(bem_ctx* c, bem_body* b, int mode, int compose, int nb, const double* betas) {
    if (!c || !b || b->ctx != c) return fail(BEM_ERR_ARG, "bad context/body");
    if (nb < 0 || nb > kMaxRhs - 6 || (nb > 0 && !betas)) return fail(BEM_ERR_ARG, "nb must be in [0, 10]");
    if (mode < BEM_MODE_FP64 || mode > BEM_MODE_MIXED_SURROGATE) return fail(BEM_ERR_ARG, "bad mode");
    if (mode == BEM_MODE_MIXED_SURROGATE && (!b->ctx->surrogate || !b->ctx->surrogate->params))
        return fail(BEM_ERR_STATE, "BEM_MODE_MIXED_SURROGATE needs a trained surrogate on the context");
    if (compose < 0 || compose > 2) return fail(BEM_ERR_ARG, "bad compose case");
    if (!b->have_rankine) return fail(BEM_ERR_STATE, "Rankine matrices not assembled");
    return BEM_OK;
}

// One (body, frequency) of a sweep or batch; `out` is its result slot.
struct Task {
    bem_body* b;
    int out;
    double omega, kappa;
};

// Runs the tasks in order on the lanes: each goes to a free lane, else to the
// first lane whose previous task has completed (polled while all are busy),
// after that lane's results are collected.
int run_tasks(const std::vector<Lane*>& lanes, const std::vector<Task>& tasks, int mode, int compose, int nb,
              double* A36, double* B36, double* F, double* FK, double* SC, bem_solve_info* infos) {
    size_t fs = (size_t)6 * nb * 2;
    int nl = (int)lanes.size();
    // one lane: its factorisations have the device to themselves
    for (Lane* L : lanes) L->lu_latency = nl == 1;
    std::vector<int> busy(nl, -1);
    size_t next = 0;
    int numeric = BEM_OK;
    std::string numeric_msg;
    // on any other error no lane may be left with a frequency in flight
    auto drain = [&](int r) {
        std::string msg = g_err;
        for (int l = 0; l < nl; l++)
            if (busy[l] >= 0) { (void)hipEventSynchronize(lanes[l]->done); lanes[l]->pending = -1; }
        return fail(r, msg);
    };
    for (;;) {
        int pick = -1;
        if (next < tasks.size())
            for (int l = 0; l < nl && pick < 0; l++)
                if (busy[l] < 0) pick = l;
        // every lane busy: the first to finish (polled, so that no finished
        // lane waits while the host blocks on another)
        bool any_busy = false;
        for (int l = 0; l < nl; l++) any_busy = any_busy || busy[l] >= 0;
        while (pick < 0 && any_busy) {
            for (int l = 0; l < nl && pick < 0; l++)
                if (busy[l] >= 0 && hipEventQuery(lanes[l]->done) == hipSuccess) pick = l;
            if (pick < 0) std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        if (pick < 0) break;   // nothing in flight and nothing left
        int r;
        if (busy[pick] >= 0) {
            const Task& t = tasks[busy[pick]];
            size_t f = (size_t)t.out;
            r = harvest(t.b, *lanes[pick], A36 + 36 * f, B36 + 36 * f, F ? F + fs * f : nullptr,
                        FK ? FK + fs * f : nullptr, SC ? SC + fs * f : nullptr, infos ? infos + f : nullptr);
            if (infos) infos[f].lanes = nl;
            busy[pick] = -1;
            if (r == BEM_ERR_NUMERIC) {   // that frequency's results are NaN; the others go on
                if (numeric == BEM_OK) { numeric = r; numeric_msg = g_err; }
            } else if (r) {
                return drain(r);
            }
        }
        if (next < tasks.size()) {
            const Task& t = tasks[next];
            if ((r = issue_frequency(t.b, *lanes[pick], (int)next, mode, t.omega, t.kappa, compose, 0, nb))) {
                (void)hipEventSynchronize(lanes[pick]->done);
                lanes[pick]->pending = -1;
                return drain(r);
            }
            busy[pick] = (int)next;
            next++;
        }
    }
    return numeric ? fail(numeric, numeric_msg) : BEM_OK;
}

int upload_betas(Lane& L, int nb, const double* betas) {
    if (nb > 0) {
        HIP_TRY(hipMemcpyAsync(L.d_betas, betas, nb * sizeof(double), hipMemcpyHostToDevice, L.s1));
        HIP_TRY(hipStreamSynchronize(L.s1));
    }
    return BEM_OK;
}

}  // namespace

int bem_set_refine_limit(bem_ctx* c, int max_steps) {
    if (!c || max_steps < 0 || max_steps > kRefineMaxIter)
        return fail(BEM_ERR_ARG, "bem_set_refine_limit: max_steps must be in [0, 4]");
    c->refine_max = max_steps;
    return BEM_OK;
}

int bem_set_wave_far32(bem_ctx* c, double near_radius) {
    if (!c || std::isnan(near_radius)) return fail(BEM_ERR_ARG, "bem_set_wave_far32: null context or NaN radius");
    c->wave_far_r2 = near_radius < 0.0 ? -1.0 : near_radius * near_radius;
    return BEM_OK;
}

int bem_set_wave_wg_per_cu(bem_ctx* c, int n) {
    if (!c || n < -1 || n > 64) return fail(BEM_ERR_ARG, "bem_set_wave_wg_per_cu: null context or n not in [-1, 64]");
    c->wave_wg_per_cu = n;
    return BEM_OK;
}

int bem_set_lu_arith(bem_ctx* c, int arith) {
    if (!c || (arith != BEM_LU_FP32 && arith != BEM_LU_SPLIT_F16 && arith != BEM_LU_SPLIT_F16_TILE))
        return fail(BEM_ERR_ARG, "bem_set_lu_arith: BEM_LU_FP32, BEM_LU_SPLIT_F16 or BEM_LU_SPLIT_F16_TILE");
    const int base = arith == BEM_LU_FP32 ? kLuArithF32 : arith == BEM_LU_SPLIT_F16 ? kLuArithSplitF16 : kLuArithSplitF16Tile;
    c->lu_arith = (c->lu_arith & ~kLuArithMask) | base;
    return BEM_OK;
}

int bem_set_refine_eager(bem_ctx* c, int steps) {
    if (!c || steps < 0 || steps > kRefineMaxIter)
        return fail(BEM_ERR_ARG, "bem_set_refine_eager: steps must be in [0, 4]");
    c->refine_eager = steps;
    return BEM_OK;
}

int bem_set_pivot_threshold(bem_ctx* c, double tau) {
    if (!c || !(tau >= 0.0 && tau <= 1.0)) return fail(BEM_ERR_ARG, "bem_set_pivot_threshold: tau in [0, 1]");
    c->pivot_tau = tau;
    return BEM_OK;
}

int bem_set_lu_speculative(bem_ctx* c, int enable) {
    if (!c) return fail(BEM_ERR_ARG, "bem_set_lu_speculative: null context");
    c->lu_spec = enable ? 1 : 0;
    return BEM_OK;
}

int bem_set_lu_tile(bem_ctx* c, int tile) {
    if (!c || tile < -1 || tile >= kLuTileCfgs)
        return fail(BEM_ERR_ARG, "bem_set_lu_tile: tile in [-1, 8]");
    c->lu_arith = (c->lu_arith & kLuArithMask) | ((tile + 1) << kLuTileShift);
    return BEM_OK;
}

int bem_solve_frequency(bem_ctx* c, bem_body* b, int mode, double omega, double kappa, int compose, int rad_only,
                        int nb, const double* betas, double* A36, double* B36, double* F, double* FK, double* SC,
                        double* phiJ, double* phiS, bem_solve_info* info) {
    int r = check_solve_args(c, b, mode, compose, nb, betas);
    if (r) return r;
    if (!A36 || !B36) return fail(BEM_ERR_ARG, "bem_solve_frequency: A36/B36 required");
    if (!freq_ok(omega, kappa)) return fail(BEM_ERR_ARG, "bem_solve_frequency: omega and kappa must be finite, kappa >= 0");
    HIP_TRY(hipSetDevice(c->device));
    if ((r = ensure_lanes(b, 1, mode != BEM_MODE_FP64))) return r;
    Lane& L = *b->lanes[0];
    L.lu_latency = 1;   // the one frequency in flight
    if ((r = upload_betas(L, nb, betas))) return r;
    if ((r = issue_frequency(b, L, 0, mode, omega, kappa, compose, rad_only, nb))) return r;
    if ((r = harvest(b, L, A36, B36, F, FK, SC, info))) return r;
    if (phiJ || phiS) {
        int Np = b->Np, nr = 6 + nb;
        std::vector<double> xr((size_t)Np * nr), xi((size_t)Np * nr);
        HIP_TRY(hipMemcpy(xr.data(), L.sys.Xr, xr.size() * 8, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(xi.data(), L.sys.Xi, xi.size() * 8, hipMemcpyDeviceToHost));
        if (phiJ) planar_to_interleaved(xr, xi, b->NPA, Np, 6, phiJ);
        if (phiS && nb > 0) {
            std::vector<double> sr(xr.begin() + (size_t)6 * Np, xr.end()), si(xi.begin() + (size_t)6 * Np, xi.end());
            planar_to_interleaved(sr, si, b->NPA, Np, nb, phiS);
        }
    }
    return BEM_OK;
}

// Lanes for ntask tasks of padded order up to Np when the caller leaves the
// choice to the library: at most `pool`, spread so that every round of tasks
// fills the lanes evenly (20 tasks: 10 lanes twice, not 8 + 8 + 4; the last,
// partly filled round runs with the device half idle).
constexpr int kAutoLanes = 10;
int auto_lanes(int ntask, int pool) {
    int lanes = std::max(1, std::min(pool, ntask));
    const int rounds = (ntask + lanes - 1) / lanes;
    return std::max(1, (ntask + rounds - 1) / rounds);
}
// Lanes the library keeps for such calls: kAutoLanes, or fewer where half the
// free device memory would not hold them (a lane keeps about 48 Np^2 bytes:
// FP64 S and D, the FP32 copy of D, workspaces), counting lanes already
// there. They are made on the first call that leaves the choice to the
// library, whatever its task count, so that a short first sweep (a warm-up)
// does not leave the allocation of the rest to the next one.
int auto_pool(int Np, int have) {
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) return std::max(1, std::min(kAutoLanes, have));
    const size_t per = (size_t)48 * Np * Np;
    return std::max(1, std::min(kAutoLanes, have + (int)(freeb / 2 / per)));
}

int bem_solve_sweep(bem_ctx* c, bem_body* b, int mode, int nf, const double* omegas, const double* kappas,
                    int compose, int nb, const double* betas, int lanes, double* A36, double* B36, double* F,
                    double* FK, double* SC, bem_solve_info* infos) {
    int r = check_solve_args(c, b, mode, compose, nb, betas);
    if (r) return r;
    if (nf < 0 || !omegas || !A36 || !B36 || lanes < 0 || lanes > 16)
        return fail(BEM_ERR_ARG, "bem_solve_sweep: bad arguments (lanes in [0, 16])");
    HIP_TRY(hipSetDevice(c->device));
    if (lanes == 0) {
        const int pool = auto_pool(b->Np, (int)b->lanes.size());
        if ((r = ensure_lanes(b, pool, mode != BEM_MODE_FP64))) return r;
        lanes = auto_lanes(nf, pool);
    }
    if ((r = ensure_lanes(b, lanes, mode != BEM_MODE_FP64))) return r;
    for (int l = 0; l < lanes; l++)
        if ((r = upload_betas(*b->lanes[l], nb, betas))) return r;
    std::vector<Task> tasks(nf);
    for (int f = 0; f < nf; f++) {
        tasks[f] = {b, f, omegas[f], kappas ? kappas[f] : omegas[f] * omegas[f] / 9.81};
        if (!freq_ok(tasks[f].omega, tasks[f].kappa))
            return fail(BEM_ERR_ARG, "bem_solve_sweep: omega and kappa must be finite, kappa >= 0");
    }
    std::vector<Lane*> ls(b->lanes.begin(), b->lanes.begin() + lanes);
    return run_tasks(ls, tasks, mode, compose, nb, A36, B36, F, FK, SC, infos);
}

int bem_solve_batch(bem_ctx* c, int nbody, bem_body* const* bodies, const int* nfreq, const double* omegas, int mode,
                    int nb, const double* betas, int lanes, double* A36, double* B36, double* F,
                    bem_solve_info* infos) {
    if (!c || nbody < 1 || !bodies || !nfreq || !omegas || !A36 || !B36 || lanes < 0 || lanes > 16)
        return fail(BEM_ERR_ARG, "bem_solve_batch: bad arguments (lanes in [0, 16])");
    int capNp = 0, total = 0, r = BEM_OK;
    for (int i = 0; i < nbody; i++) {
        if ((r = check_solve_args(c, bodies[i], mode, BEM_COMPOSE_WAVE, nb, betas))) return r;
        if (nfreq[i] < 0) return fail(BEM_ERR_ARG, "bem_solve_batch: negative frequency count");
        capNp = std::max(capNp, bodies[i]->Np);
        total += nfreq[i];
    }
    for (int t = 0; t < total; t++)
        if (!freq_ok(omegas[t], omegas[t] * omegas[t] / 9.81))
            return fail(BEM_ERR_ARG, "bem_solve_batch: omegas must be finite");
    HIP_TRY(hipSetDevice(c->device));
    bool need32 = mode != BEM_MODE_FP64;
    int make = lanes;
    if (lanes == 0) {
        const bool fits = !c->pool.empty() && c->pool[0]->sys.capNp >= capNp;
        if (!fits) {   // the pool is remade for this batch: its memory is free then
            for (Lane* l : c->pool) delete l;
            c->pool.clear();
        }
        make = auto_pool(capNp, (int)c->pool.size());
        lanes = auto_lanes(total, make);
    }
    // lane pool of the context, sized for the largest body of the batch
    if ((int)c->pool.size() < make || (!c->pool.empty() && c->pool[0]->sys.capNp < capNp)) {
        for (Lane* l : c->pool) delete l;
        c->pool.clear();
        while ((int)c->pool.size() < make) {
            Lane* L = new Lane;
            if ((r = L->init(capNp))) { delete L; return r; }
            c->pool.push_back(L);
        }
    }
    std::vector<Lane*> ls(c->pool.begin(), c->pool.begin() + lanes);
    for (Lane* L : ls) {
        if (need32 && (r = L->sys.init32())) return r;
        if ((r = upload_betas(*L, nb, betas))) return r;
    }
    // tile tuning for every padded order of the batch (cached per device)
    std::vector<int> nps;
    for (int i = 0; i < nbody; i++) nps.push_back(bodies[i]->Np);
    std::sort(nps.begin(), nps.end());
    nps.erase(std::unique(nps.begin(), nps.end()), nps.end());
    for (int np : nps) {
        tune_lu_gemms<double>(ls[0]->s1, np, ls[0]->sys.Dr, ls[0]->sys.Di);
        if (need32)
            tune_lu_gemms<float>(ls[0]->s1, np, ls[0]->sys.D32r, ls[0]->sys.D32i, ls[0]->sys.w32.umax,
                                 ls[0]->sys.w32.opA, ls[0]->sys.w32.opB);
    }
    HIP_TRY(hipStreamSynchronize(ls[0]->s1));
    // tasks, largest body first (the long factorisations start early)
    std::vector<Task> tasks;
    tasks.reserve(total);
    for (int i = 0, o = 0; i < nbody; i++)
        for (int k = 0; k < nfreq[i]; k++, o++) tasks.push_back({bodies[i], o, omegas[o], omegas[o] * omegas[o] / 9.81});
    std::stable_sort(tasks.begin(), tasks.end(), [](const Task& x, const Task& y) { return x.b->Np > y.b->Np; });
    return run_tasks(ls, tasks, mode, BEM_COMPOSE_WAVE, nb, A36, B36, F, nullptr, nullptr, infos);
    return BEM_OK;
}

// ---- kernel-level access -------------------------------------------------------
static int copy_matrix_out(const double* d, int N, int Np, double* h) {
    std::vector<double> tmp((size_t)Np * N);
    HIP_TRY(hipMemcpy(tmp.data(), d, tmp.size() * 8, hipMemcpyDeviceToHost));
    for (int j = 0; j < N; j++) std::memcpy(h + (size_t)j * N, tmp.data() + (size_t)j * Np, N * sizeof(double));
    return BEM_OK;
}

int bem_get_rankine(bem_ctx* c, bem_body* b, double* Ssrc, double* Srefl, double* Dsrc, double* Drefl) {
    if (!c || !b || !b->have_rankine) return fail(BEM_ERR_STATE, "bem_get_rankine: Rankine matrices not assembled");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    int r = BEM_OK;
    if (Ssrc && (r = copy_matrix_out(b->Ssrc, b->N, b->Np, Ssrc))) return r;
    if (Srefl && (r = copy_matrix_out(b->Srefl, b->N, b->Np, Srefl))) return r;
    if (Dsrc && (r = copy_matrix_out(b->Dsrc, b->N, b->Np, Dsrc))) return r;
    if (Drefl && (r = copy_matrix_out(b->Drefl, b->N, b->Np, Drefl))) return r;
    return BEM_OK;
}

int bem_get_system(bem_ctx* c, bem_body* b, int mode, double kappa, int compose, double* SMat, double* DMat) {
    if (!c || !b || !b->have_rankine) return fail(BEM_ERR_STATE, "bem_get_system: Rankine matrices not assembled");
    if (compose < 0 || compose > 2) return fail(BEM_ERR_ARG, "bem_get_system: bad compose case");
    if (!freq_ok(0.0, kappa)) return fail(BEM_ERR_ARG, "bem_get_system: kappa must be finite and >= 0");
    if (b->ctx != c || mode < BEM_MODE_FP64 || mode > BEM_MODE_MIXED_SURROGATE) return fail(BEM_ERR_ARG, "bem_get_system: bad context or mode");
    if (mode == BEM_MODE_MIXED_SURROGATE && (!c->surrogate || !c->surrogate->params))
        return fail(BEM_ERR_STATE, "bem_get_system: BEM_MODE_MIXED_SURROGATE needs a surrogate on the context");
    HIP_TRY(hipSetDevice(c->device));
    int r = ensure_lanes(b, 1, mode != BEM_MODE_FP64);
    if (r) return r;
    Lane& L = *b->lanes[0];
    set_frequency(L, 0.0, kappa);
    assemble(b, L, mode, kappa, compose);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(L.s1));
    int N = b->N, Np = b->Np;
    std::vector<double> re((size_t)Np * N), im((size_t)Np * N);
    auto out = [&](const double* dr, const double* di, double* h) -> int {
        HIP_TRY(hipMemcpy(re.data(), dr, re.size() * 8, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(im.data(), di, im.size() * 8, hipMemcpyDeviceToHost));
        planar_to_interleaved(re, im, N, Np, N, h);
        return BEM_OK;
    };
    if (SMat && (r = out(L.Sr, L.Si, SMat))) return r;
    if (DMat && (r = out(L.sys.Dr, L.sys.Di, DMat))) return r;
    return BEM_OK;
}

int bem_profile_lu(bem_ctx* c, bem_body* b, int mode, double kappa, double* out) {
    if (!c || !b || !b->have_rankine || !out) return fail(BEM_ERR_STATE, "bem_profile_lu: Rankine matrices not assembled");
    if (b->ctx != c || mode < BEM_MODE_FP64 || mode > BEM_MODE_MIXED_SURROGATE) return fail(BEM_ERR_ARG, "bem_profile_lu: bad context or mode");
    if (mode == BEM_MODE_MIXED_SURROGATE && (!c->surrogate || !c->surrogate->params))
        return fail(BEM_ERR_STATE, "bem_profile_lu: BEM_MODE_MIXED_SURROGATE needs a surrogate on the context");
    if (!freq_ok(0.0, kappa)) return fail(BEM_ERR_ARG, "bem_profile_lu: kappa must be finite and >= 0");
    HIP_TRY(hipSetDevice(c->device));
    int r = ensure_lanes(b, 1, mode != BEM_MODE_FP64);
    if (r) return r;
    Lane& L = *b->lanes[0];
    L.lu_arith = c->lu_arith;
    L.lu_latency = 1;   // one factorisation alone on the device
    L.lu_spec = 0;      // the full path, pivoting fallback included
    set_frequency(L, 0.0, kappa);
    assemble(b, L, mode, kappa, BEM_COMPOSE_WAVE);
    LuProfile* prof = new LuProfile;
    prof->N = b->N;
    HIP_TRY(hipEventRecord(L.t[0], L.s1));
    factor_system(L, mode == BEM_MODE_FP64 ? 64 : 32, prof);
    HIP_TRY(hipEventRecord(L.t[1], L.s1));
    hipError_t e = hipStreamSynchronize(L.s1);
    double sum = 0.0;
    for (int i = 0; i < prof->n; i++) {
        float t = 0;
        (void)hipEventElapsedTime(&t, prof->ev[2 * i], prof->ev[2 * i + 1]);
        sum += t;
    }
    float tot = 0;
    (void)hipEventElapsedTime(&tot, L.t[0], L.t[1]);
    out[0] = tot;
    out[1] = sum;
    out[2] = prof->useful_flops;
    out[3] = prof->n;
    out[4] = prof->n ? sum / prof->n : 0.0;
    out[5] = prof->n ? prof->useful_flops / prof->n : 0.0;
    out[6] = prof->useful_bytes;
    out[7] = prof->n ? prof->useful_bytes / prof->n : 0.0;
    out[8] = 0.0;
    if (e == hipSuccess && prof->n) {
        out[8] = mode == BEM_MODE_FP64 ? replay_lu_gemms<double>(*prof, L.s1) : replay_lu_gemms<float>(*prof, L.s1);
        e = hipStreamSynchronize(L.s1);
    }
    for (auto& ev : prof->ev)
        if (ev) (void)hipEventDestroy(ev);
    delete prof;
    if (e != hipSuccess) return fail(BEM_ERR_HIP, std::string("bem_profile_lu: ") + hipGetErrorString(e));
    return BEM_OK;
}

}  // extern "C"

namespace {

// Uploads the potentials (interleaved complex, N x 6 and N x nb) as one planar
// N x (6 + nb) block, and the headings.
int upload_potentials(bem_body* b, int nb, const double* betas, const double* phiJ, const double* phiS,
                      SurBuffers& B, double** phr, double** phi, double** dbetas) {
    const int N = b->NPA, nc = 6 + nb;
    std::vector<double> hr((size_t)N * nc), hi((size_t)N * nc);
    for (int c = 0; c < nc; c++) {
        const double* src = c < 6 ? phiJ + 2 * (size_t)c * N : phiS + 2 * (size_t)(c - 6) * N;
        for (int i = 0; i < N; i++) { hr[(size_t)c * N + i] = src[2 * i]; hi[(size_t)c * N + i] = src[2 * i + 1]; }
    }
    int r = BEM_OK;
    if ((r = B.get(phr, hr.size())) || (r = B.get(phi, hi.size())) || (r = B.get(dbetas, std::max(nb, 1)))) return r;
    HIP_TRY(hipMemcpy(*phr, hr.data(), hr.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(*phi, hi.data(), hi.size() * 8, hipMemcpyHostToDevice));
    if (nb > 0) HIP_TRY(hipMemcpy(*dbetas, betas, nb * 8, hipMemcpyHostToDevice));
    return BEM_OK;
}

// Splits an n x (6 + nb) complex block into the 6-column and nb-column outputs.
void split_columns(const std::vector<double>& h, int n, int nb, double* A, double* B) {
    const int nc = 6 + nb;
    for (int i = 0; i < n; i++)
        for (int c = 0; c < nc; c++) {
            double* dst = c < 6 ? (A ? A + 2 * ((size_t)i * 6 + c) : nullptr)
                                : (B ? B + 2 * ((size_t)i * nb + (c - 6)) : nullptr);
            if (dst) { dst[0] = h[2 * ((size_t)i * nc + c)]; dst[1] = h[2 * ((size_t)i * nc + c) + 1]; }
        }
}

}  // namespace

extern "C" {

int bem_kochin(bem_ctx* c, bem_body* b, double omega, double kappa, int nkoch, int nb, const double* betas,
               const double* phiJ, const double* phiS, double* KochRad, double* KochDiff) {
    if (!c || !b || b->ctx != c || nkoch < 1 || nkoch > 65536 || nb < 0 || nb > kMaxRhs - 6 || !phiJ ||
        (nb > 0 && (!betas || !phiS)))
        return fail(BEM_ERR_ARG, "bem_kochin: bad arguments (nkoch in [1, 65536], nb in [0, 10])");
    if (nb > 0 && omega == 0.0) return fail(BEM_ERR_ARG, "bem_kochin: diffraction needs omega != 0");
    if (!freq_ok(omega, kappa)) return fail(BEM_ERR_ARG, "bem_kochin: omega and kappa must be finite, kappa >= 0");
    HIP_TRY(hipSetDevice(c->device));
    SurBuffers B;
    double *phr = nullptr, *phi = nullptr, *dbetas = nullptr, *dout = nullptr;
    int r = upload_potentials(b, nb, betas, phiJ, phiS, B, &phr, &phi, &dbetas);
    if (r) return r;
    const int nc = 6 + nb;
    if ((r = B.get(&dout, 2 * (size_t)nkoch * nc))) return r;
    launch_kochin(c->stream, b->NPA, nkoch, nb, b->d_body, b->d_gen, dbetas, omega, kappa, phr, phi, dout);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<double> h(2 * (size_t)nkoch * nc);
    HIP_TRY(hipMemcpy(h.data(), dout, h.size() * 8, hipMemcpyDeviceToHost));
    split_columns(h, nkoch, nb, KochRad, KochDiff);
    return BEM_OK;
}

int bem_fs_elevation(bem_ctx* c, bem_body* b, double omega, double kappa, int npts, const double* pts, int nb,
                     const double* betas, const double* phiJ, const double* phiS, double* ExtRad, double* ExtDiff) {
    if (!c || !b || b->ctx != c || npts < 1 || !pts || nb < 0 || nb > kMaxRhs - 6 || !phiJ ||
        (nb > 0 && (!betas || !phiS)))
        return fail(BEM_ERR_ARG, "bem_fs_elevation: bad arguments (npts >= 1, nb in [0, 10])");
    if (nb > 0 && omega == 0.0) return fail(BEM_ERR_ARG, "bem_fs_elevation: diffraction needs omega != 0");
    if (!freq_ok(omega, kappa)) return fail(BEM_ERR_ARG, "bem_fs_elevation: omega and kappa must be finite, kappa >= 0");
    HIP_TRY(hipSetDevice(c->device));
    SurBuffers B;
    double *phr = nullptr, *phi = nullptr, *dbetas = nullptr, *dpts = nullptr, *dout = nullptr;
    int r = upload_potentials(b, nb, betas, phiJ, phiS, B, &phr, &phi, &dbetas);
    if (r) return r;
    const int nc = 6 + nb;
    if ((r = B.get(&dpts, 3 * (size_t)npts)) || (r = B.get(&dout, 2 * (size_t)npts * nc))) return r;
    HIP_TRY(hipMemcpy(dpts, pts, 24 * (size_t)npts, hipMemcpyHostToDevice));
    launch_fs_field(c->stream, b->NPA, npts, dpts, nb, dbetas, omega, kappa, b->d_body, b->d_refl, b->d_quad, b->d_gen,
                    phr, phi, dout);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<double> h(2 * (size_t)npts * nc);
    HIP_TRY(hipMemcpy(h.data(), dout, h.size() * 8, hipMemcpyDeviceToHost));
    split_columns(h, npts, nb, ExtRad, ExtDiff);
    return BEM_OK;
}

int bem_bessel01(int n, const double* x, double* out) {
    if (n < 0 || (n > 0 && (!x || !out))) return fail(BEM_ERR_ARG, "bem_bessel01: bad arguments");
    for (int i = 0; i < n; i++) {
        if (!(x[i] >= 0.0)) return fail(BEM_ERR_ARG, "bem_bessel01: x must be >= 0");
        dev::bessel01(x[i], out[4 * i], out[4 * i + 1], out[4 * i + 2], out[4 * i + 3]);
    }
    return BEM_OK;
}

int bem_gf_coefficients(double* out, int n) {
    if (!out || n < 65) return fail(BEM_ERR_ARG, "bem_gf_coefficients: need room for 65 values");
    int k = 0;
    for (double v : dev::kGF.A) out[k++] = v;
    for (double v : dev::kGF.B) out[k++] = v;
    for (double v : dev::kGF.C) out[k++] = v;
    for (double v : dev::kGF.D) out[k++] = v;
    for (double v : dev::kGF.dA) out[k++] = v;
    for (double v : dev::kGF.dB) out[k++] = v;
    for (double v : dev::kGF.dC) out[k++] = v;
    out[k++] = dev::kEUL;
    return BEM_OK;
}

int bem_struve_coefficients(double* out, int n) {
    if (!out || n < 26) return fail(BEM_ERR_ARG, "bem_struve_coefficients: need room for 26 values");
    int k = 0;
    for (double v : dev::kStruve.h0p) out[k++] = v;
    for (double v : dev::kStruve.h0a) out[k++] = v;
    for (double v : dev::kStruve.h0b) out[k++] = v;
    for (double v : dev::kStruve.h1p) out[k++] = v;
    for (double v : dev::kStruve.h1a) out[k++] = v;
    for (double v : dev::kStruve.h1b) out[k++] = v;
    return BEM_OK;
}

int bem_eval_havelock(bem_ctx* c, int precision, int n, const double* X, const double* Y, const double* V,
                      double* out) {
    if (!c || n <= 0 || !X || !Y || !V || !out) return fail(BEM_ERR_ARG, "bem_eval_havelock: bad arguments");
    if (precision != 64 && precision != 32) return fail(BEM_ERR_ARG, "bem_eval_havelock: precision must be 64 or 32");
    HIP_TRY(hipSetDevice(c->device));
    double *dx = nullptr, *dy = nullptr, *dv = nullptr, *dout = nullptr;
    int r = BEM_OK;
    if ((r = dalloc_t(&dx, n)) || (r = dalloc_t(&dy, n)) || (r = dalloc_t(&dv, n)) || (r = dalloc_t(&dout, 8 * (size_t)n))) {
        (void)hipFree(dx); (void)hipFree(dy); (void)hipFree(dv); (void)hipFree(dout);
        return r;
    }
    (void)hipMemcpy(dx, X, n * 8, hipMemcpyHostToDevice);
    (void)hipMemcpy(dy, Y, n * 8, hipMemcpyHostToDevice);
    (void)hipMemcpy(dv, V, n * 8, hipMemcpyHostToDevice);
    launch_havelock_eval(c->stream, n, dx, dy, dv, dout, precision == 32);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(out, dout, 8 * (size_t)n * 8, hipMemcpyDeviceToHost);
    (void)hipFree(dx); (void)hipFree(dy); (void)hipFree(dv); (void)hipFree(dout);
    if (e != hipSuccess) return fail(BEM_ERR_HIP, std::string("bem_eval_havelock: ") + hipGetErrorString(e));
    return BEM_OK;
}

}  // extern "C"

// ---- wave-term surrogate -----------------------------------------------------
namespace {

// 16 (a, c) pairs of the per-pass affine shuffle row = (a q + c) mod pool:
// primes a above the pool (so coprime with it) and seeded offsets c.
std::vector<unsigned long long> shuffle_table(uint64_t seed, unsigned long long pool) {
    std::vector<unsigned long long> t;
    unsigned long long a = std::max<unsigned long long>(pool, 1ULL << 31) + 1;
    while (t.size() < 32) {
        bool prime = a > 1;
        for (unsigned long long d = 2; d * d <= a && prime; d++) prime = (a % d) != 0;
        if (prime) {
            t.push_back(a);
            t.push_back(splitmix64(seed ^ (0xA5A5A5A5ULL + t.size())) % pool);
        }
        a += 2;
    }
    return t;
}

int ensure_surrogate(bem_ctx* c) {
    if (!c->surrogate) c->surrogate = new SurState;
    if (!c->surrogate->params) return dalloc_t(&c->surrogate->params, kSurParams);
    return BEM_OK;
}

}  // namespace

extern "C" {

int bem_surrogate_default_config(bem_surrogate_config* cfg) {
    if (!cfg) return fail(BEM_ERR_ARG, "bem_surrogate_default_config: null config");
    cfg->seed = 20261203ULL;
    cfg->n_samples = 20000000;
    cfg->batch = 8192;
    cfg->steps = 24000;
    cfg->lr = 2e-3;
    cfg->lr_final = 2e-5;
    cfg->n_valid = 1000000;
    cfg->sampling = 1;
    cfg->steps_per_launch = 0;
    return BEM_OK;
}

int bem_surrogate_param_count(void) { return kSurParams; }

int bem_measure_peaks(bem_ctx* c, double* out) {
    if (!c || !out) return fail(BEM_ERR_ARG, "bem_measure_peaks: null argument");
    HIP_TRY(hipSetDevice(c->device));
    if (measure_peaks(c->stream, out) != 0) return fail(BEM_ERR_HIP, "bem_measure_peaks: a HIP call failed");
    return BEM_OK;
}

int bem_set_graphs(bem_ctx* c, int enable) {
    if (!c) return fail(BEM_ERR_ARG, "bem_set_graphs: null context");
    c->graphs = enable != 0;
    return BEM_OK;
}

int bem_surrogate_train(bem_ctx* c, const bem_surrogate_config* cfg, bem_surrogate_report* rep) {
    if (!c || !cfg) return fail(BEM_ERR_ARG, "bem_surrogate_train: null argument");
    const long long n = cfg->n_samples, nv = cfg->n_valid;
    if (n < 1 || n > (1LL << 30) || cfg->batch < 1 || cfg->batch > (1 << 20) || cfg->steps < 0 || nv < 0 ||
        nv > (1LL << 26) || !(cfg->lr >= 0.0) || !(cfg->lr_final >= 0.0) || cfg->sampling < 0 || cfg->sampling > 1 ||
        cfg->steps_per_launch < 0)
        return fail(BEM_ERR_ARG, "bem_surrogate_train: n_samples in [1, 2^30], batch in [1, 2^20], steps >= 0, "
                                 "n_valid in [0, 2^26], learning rates >= 0, sampling 0 or 1, steps_per_launch >= 0");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int nwg = sur_grad_groups(cfg->batch);
    SurBuffers B;
    float *X = nullptr, *T = nullptr, *P = nullptr, *m = nullptr, *v = nullptr, *gpart = nullptr, *lpart = nullptr;
    float *Xv = nullptr, *Tv = nullptr, *Yv = nullptr;
    double *loss = nullptr, *hv = nullptr;
    unsigned long long* perm = nullptr;
    int r = BEM_OK;
    if ((r = B.get(&X, 4 * (size_t)n)) || (r = B.get(&T, kSurOut * (size_t)n)) || (r = B.get(&P, kSurParams)) ||
        (r = B.get(&m, kSurParams)) || (r = B.get(&v, kSurParams)) || (r = B.get(&gpart, (size_t)nwg * kSurParams)) ||
        (r = B.get(&lpart, nwg)) || (r = B.get(&loss, std::max(cfg->steps, 1))) || (r = B.get(&perm, 32)))
        return r;
    if (nv > 0 && ((r = B.get(&Xv, 4 * (size_t)nv)) || (r = B.get(&Tv, kSurOut * (size_t)nv)) ||
                   (r = B.get(&Yv, kSurOut * (size_t)nv)) || (r = B.get(&hv, 2 * (size_t)nv))))
        return r;
    std::vector<float> p0;
    surrogate_init_params(cfg->seed ^ 0x5EED5EEDULL, p0);
    std::vector<unsigned long long> tab = shuffle_table(cfg->seed, (unsigned long long)n);
    HIP_TRY(hipMemcpyAsync(P, p0.data(), sizeof(float) * kSurParams, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(perm, tab.data(), sizeof(unsigned long long) * 32, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(m, 0, sizeof(float) * kSurParams, st));
    HIP_TRY(hipMemsetAsync(v, 0, sizeof(float) * kSurParams, st));
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
    bem_surrogate_report out{};
    float ms = 0.0f;
    HIP_TRY(hipEventRecord(e0, st));
    launch_sur_sample(st, cfg->seed, cfg->sampling, 0, n, X, T, nullptr);
    if (nv > 0) launch_sur_sample(st, cfg->seed, cfg->sampling, n, nv, Xv, Tv, hv);
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    out.ms_sample = ms;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    // step t's learning rate (cosine decay) with Adam's bias correction
    std::vector<float> lrt(std::max(cfg->steps, 1));
    for (int t = 0; t < cfg->steps; t++) {
        const double lr = cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(M_PI * t / cfg->steps));
        lrt[t] = (float)(lr * std::sqrt(1.0 - std::pow((double)b2, t + 1)) / (1.0 - std::pow((double)b1, t + 1)));
    }
    // several steps per dispatch: one cooperative launch holds every
    // work-group of the batch and meets at a grid barrier twice per step
    const bool coop = cfg->steps_per_launch != 1 && cfg->steps > 0 && sur_train_coop_capacity(c->device) >= nwg;
    const int chunk = cfg->steps_per_launch > 1 ? cfg->steps_per_launch : kSurStepsPerLaunch;
    float* lr_tab = nullptr;
    unsigned* bar = nullptr;
    int* berr = nullptr;
    if (coop) {
        if ((r = B.get(&lr_tab, (size_t)cfg->steps)) || (r = B.get(&bar, 64)) || (r = B.get(&berr, 1))) return r;
        HIP_TRY(hipMemcpyAsync(lr_tab, lrt.data(), sizeof(float) * cfg->steps, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemsetAsync(bar, 0, sizeof(unsigned) * 64, st));
        HIP_TRY(hipMemsetAsync(berr, 0, sizeof(int), st));
    }
    HIP_TRY(hipEventRecord(e0, st));
    if (coop) {
        for (int t0 = 0; t0 < cfg->steps; t0 += chunk)
            if (launch_sur_train(st, nwg, t0, std::min(chunk, cfg->steps - t0), lr_tab, b1, b2, eps, X, T, n,
                                 cfg->batch, perm, P, m, v, gpart, lpart, loss, bar, berr))
                return fail(BEM_ERR_HIP, "bem_surrogate_train: cooperative launch refused");
    } else {
        for (int t = 0; t < cfg->steps; t++) {
            launch_sur_grad(st, P, X, T, n, (long long)t * cfg->batch, cfg->batch, perm, gpart, lpart);
            launch_sur_adam(st, nwg, gpart, lpart, P, m, v, nullptr, loss + t, lrt[t], b1, b2, eps);
        }
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    out.ms_train = ms;
    if (coop) {
        int berr_h = 0;
        HIP_TRY(hipMemcpy(&berr_h, berr, sizeof(int), hipMemcpyDeviceToHost));
        if (berr_h) return fail(BEM_ERR_HIP, "bem_surrogate_train: a grid barrier of the training kernel timed out");
    }
    if (cfg->steps > 0) {
        std::vector<double> lh(cfg->steps);
        HIP_TRY(hipMemcpy(lh.data(), loss, sizeof(double) * cfg->steps, hipMemcpyDeviceToHost));
        const double norm = 1.0 / ((double)kSurOut * cfg->batch);
        int k0 = std::min(cfg->steps, 10), k1 = std::min(cfg->steps, 100);
        for (int k = 0; k < k0; k++) out.loss_initial += lh[k] * norm / k0;
        for (int k = cfg->steps - k1; k < cfg->steps; k++) out.loss_final += lh[k] * norm / k1;
    }
    if (nv > 0) {
        HIP_TRY(hipEventRecord(e0, st));
        launch_sur_forward(st, P, nv, Xv, Yv);
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        out.ms_valid = ms;
        std::vector<float> y(kSurOut * (size_t)nv), tg(kSurOut * (size_t)nv);
        std::vector<double> pts(2 * (size_t)nv);
        HIP_TRY(hipMemcpy(y.data(), Yv, y.size() * sizeof(float), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(tg.data(), Tv, tg.size() * sizeof(float), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(pts.data(), hv, pts.size() * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<double> rel((size_t)nv);
        double ss[kSurOut] = {}, rs = 0.0;
        for (long long j = 0; j < nv; j++) {
            const double h = pts[2 * j], vv = pts[2 * j + 1];
            const SurPoint sp = sur_point(h, vv);
            double num = 0.0, g[kSurOut];
            for (int ch = 0; ch < kSurOut; ch++) {
                const double e = std::fabs((double)y[kSurOut * j + ch] - (double)tg[kSurOut * j + ch]);
                ss[ch] += e * e;
                out.valid_max[ch] = std::max(out.valid_max[ch], e);
                num = std::max(num, e * sp.s);
                g[ch] = sp.s * tg[kSurOut * j + ch];
            }
            g[0] += sp.sg;
            g[1] += sp.sgh;
            // relative to the largest component of (GF, GFh), imaginary parts exact
            const double E = 2.0 * M_PI * std::exp(vv);
            double J0, J1, Y0, Y1;
            dev::bessel01(h, J0, J1, Y0, Y1);
            const double den = std::max(std::max(std::fabs(g[0]), std::fabs(E * J0)),
                                        std::max(std::fabs(g[1]), std::fabs(E * J1)));
            rel[j] = den > 0.0 ? num / den : 0.0;
            rs += rel[j] * rel[j];
            out.rel_max = std::max(out.rel_max, rel[j]);
        }
        for (int ch = 0; ch < kSurOut; ch++) out.valid_rms[ch] = std::sqrt(ss[ch] / nv);
        out.rel_rms = std::sqrt(rs / nv);
        size_t k99 = (size_t)(0.99 * (nv - 1));
        std::nth_element(rel.begin(), rel.begin() + k99, rel.end());
        out.rel_p99 = rel[k99];
    }
    if ((r = ensure_surrogate(c))) return r;
    HIP_TRY(hipMemcpyAsync(c->surrogate->params, P, sizeof(float) * kSurParams, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    c->surrogate->rep = out;
    if (rep) *rep = out;
    return BEM_OK;
}

int bem_surrogate_get(bem_ctx* c, float* params, int n) {
    if (!c || !params || n != kSurParams) return fail(BEM_ERR_ARG, "bem_surrogate_get: need bem_surrogate_param_count() floats");
    if (!c->surrogate || !c->surrogate->params) return fail(BEM_ERR_STATE, "bem_surrogate_get: no surrogate on the context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(params, c->surrogate->params, sizeof(float) * kSurParams, hipMemcpyDeviceToHost));
    return BEM_OK;
}

int bem_surrogate_set(bem_ctx* c, const float* params, int n) {
    if (!c || !params || n != kSurParams) return fail(BEM_ERR_ARG, "bem_surrogate_set: need bem_surrogate_param_count() floats");
    for (int i = 0; i < n; i++)
        if (!std::isfinite(params[i])) return fail(BEM_ERR_ARG, "bem_surrogate_set: non-finite parameter");
    HIP_TRY(hipSetDevice(c->device));
    int r = ensure_surrogate(c);
    if (r) return r;
    HIP_TRY(hipMemcpy(c->surrogate->params, params, sizeof(float) * kSurParams, hipMemcpyHostToDevice));
    c->surrogate->rep = bem_surrogate_report{};
    return BEM_OK;
}

int bem_surrogate_sample(bem_ctx* c, uint64_t seed, int sampling, int64_t first, int n, float* X, float* T) {
    if (!c || n < 1 || first < 0 || !X || !T || sampling < 0 || sampling > 1)
        return fail(BEM_ERR_ARG, "bem_surrogate_sample: bad arguments");
    HIP_TRY(hipSetDevice(c->device));
    SurBuffers B;
    float *dx = nullptr, *dt = nullptr;
    int r = BEM_OK;
    if ((r = B.get(&dx, 4 * (size_t)n)) || (r = B.get(&dt, kSurOut * (size_t)n))) return r;
    launch_sur_sample(c->stream, seed, sampling, first, n, dx, dt, nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(X, dx, 16 * (size_t)n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(T, dt, sizeof(float) * kSurOut * (size_t)n, hipMemcpyDeviceToHost));
    return BEM_OK;
}

int bem_surrogate_forward(bem_ctx* c, const float* params, int n, const float* X, float* Y) {
    if (!c || n < 1 || !X || !Y) return fail(BEM_ERR_ARG, "bem_surrogate_forward: bad arguments");
    if (!params && (!c->surrogate || !c->surrogate->params))
        return fail(BEM_ERR_STATE, "bem_surrogate_forward: no parameters given and none on the context");
    HIP_TRY(hipSetDevice(c->device));
    SurBuffers B;
    float *dp = nullptr, *dx = nullptr, *dy = nullptr;
    int r = BEM_OK;
    if ((r = B.get(&dx, 4 * (size_t)n)) || (r = B.get(&dy, kSurOut * (size_t)n))) return r;
    if (params) {
        if ((r = B.get(&dp, kSurParams))) return r;
        HIP_TRY(hipMemcpy(dp, params, sizeof(float) * kSurParams, hipMemcpyHostToDevice));
    } else {
        dp = c->surrogate->params;
    }
    HIP_TRY(hipMemcpy(dx, X, 16 * (size_t)n, hipMemcpyHostToDevice));
    launch_sur_forward(c->stream, dp, n, dx, dy);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(Y, dy, sizeof(float) * kSurOut * (size_t)n, hipMemcpyDeviceToHost));
    return BEM_OK;
}

int bem_surrogate_grad(bem_ctx* c, const float* params, int n, const float* X, const float* T, float* grad,
                       double* loss) {
    if (!c || !params || n < 1 || n > (1 << 20) || !X || !T || !grad)
        return fail(BEM_ERR_ARG, "bem_surrogate_grad: bad arguments (n in [1, 2^20])");
    HIP_TRY(hipSetDevice(c->device));
    const int nwg = sur_grad_groups(n);
    SurBuffers B;
    float *dp = nullptr, *dx = nullptr, *dt = nullptr, *gpart = nullptr, *lpart = nullptr, *dg = nullptr;
    double* dl = nullptr;
    int r = BEM_OK;
    if ((r = B.get(&dp, kSurParams)) || (r = B.get(&dx, 4 * (size_t)n)) || (r = B.get(&dt, kSurOut * (size_t)n)) ||
        (r = B.get(&gpart, (size_t)nwg * kSurParams)) || (r = B.get(&lpart, nwg)) || (r = B.get(&dg, kSurParams)) ||
        (r = B.get(&dl, 1)))
        return r;
    HIP_TRY(hipMemcpy(dp, params, sizeof(float) * kSurParams, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dx, X, 16 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dt, T, sizeof(float) * kSurOut * (size_t)n, hipMemcpyHostToDevice));
    launch_sur_grad(c->stream, dp, dx, dt, n, 0, n, nullptr, gpart, lpart);
    launch_sur_adam(c->stream, nwg, gpart, lpart, dp, nullptr, nullptr, dg, dl, 0.0f, 0.0f, 0.0f, 0.0f);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(grad, dg, sizeof(float) * kSurParams, hipMemcpyDeviceToHost));
    double l = 0.0;
    HIP_TRY(hipMemcpy(&l, dl, sizeof(double), hipMemcpyDeviceToHost));
    if (loss) *loss = l / ((double)kSurOut * n);
    return BEM_OK;
}

int bem_surrogate_eval(bem_ctx* c, int n, const double* h, const double* v, double* out, int* used) {
    if (!c || n < 1 || !h || !v || !out) return fail(BEM_ERR_ARG, "bem_surrogate_eval: bad arguments");
    if (!c->surrogate || !c->surrogate->params) return fail(BEM_ERR_STATE, "bem_surrogate_eval: no surrogate on the context");
    HIP_TRY(hipSetDevice(c->device));
    SurBuffers B;
    double *dh = nullptr, *dv = nullptr, *dout = nullptr;
    int* du = nullptr;
    int r = BEM_OK;
    if ((r = B.get(&dh, n)) || (r = B.get(&dv, n)) || (r = B.get(&dout, 4 * (size_t)n)) || (r = B.get(&du, n))) return r;
    HIP_TRY(hipMemcpy(dh, h, 8 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dv, v, 8 * (size_t)n, hipMemcpyHostToDevice));
    launch_sur_eval(c->stream, c->surrogate->params, n, dh, dv, dout, du);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(out, dout, 32 * (size_t)n, hipMemcpyDeviceToHost));
    if (used) HIP_TRY(hipMemcpy(used, du, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return BEM_OK;
}

int bem_dense_solve(bem_ctx* c, int precision, int n, const double* A, int nrhs, const double* B, double* X,
                    bem_solve_info* info) {
    if (!c || n <= 0 || !A || !B || !X || nrhs < 1 || nrhs > kMaxRhs || (precision != 64 && precision != 32))
        return fail(BEM_ERR_ARG, "bem_dense_solve: bad arguments (nrhs in [1,16], precision 64 or 32)");
    if (round_up(n, kPad) / kLuNb > kMaxPanels) return fail(BEM_ERR_ARG, "bem_dense_solve: matrix too large");
    for (size_t q = 0; q < 2 * (size_t)n * n; q++)
        if (!std::isfinite(A[q])) return fail(BEM_ERR_ARG, "bem_dense_solve: A has a non-finite entry");
    for (size_t q = 0; q < 2 * (size_t)n * nrhs; q++)
        if (!std::isfinite(B[q])) return fail(BEM_ERR_ARG, "bem_dense_solve: B has a non-finite entry");
    HIP_TRY(hipSetDevice(c->device));
    Lane L;
    int r = L.init(n);
    if (r) return r;
    L.refine_max = c->refine_max;
    L.lu_arith = c->lu_arith;
    L.pivot_tau = c->pivot_tau;
    LinSys& s = L.sys;
    int Np = s.Np;
    {
        std::vector<double> re((size_t)Np * Np, 0.0), im((size_t)Np * Np, 0.0);
        for (int j = 0; j < Np; j++)
            for (int i = 0; i < Np; i++) {
                size_t o = (size_t)j * Np + i;
                if (i < n && j < n) { re[o] = A[2 * ((size_t)j * n + i)]; im[o] = A[2 * ((size_t)j * n + i) + 1]; }
                else if (i == j) re[o] = 1.0;
            }
        HIP_TRY(hipMemcpy(s.Dr, re.data(), re.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s.Di, im.data(), im.size() * 8, hipMemcpyHostToDevice));
        std::vector<double> br((size_t)Np * kMaxRhs, 0.0), bi((size_t)Np * kMaxRhs, 0.0);
        for (int j = 0; j < nrhs; j++)
            for (int i = 0; i < n; i++) { br[(size_t)j * Np + i] = B[2 * ((size_t)j * n + i)]; bi[(size_t)j * Np + i] = B[2 * ((size_t)j * n + i) + 1]; }
        HIP_TRY(hipMemcpy(s.Br, br.data(), br.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s.Bi, bi.data(), bi.size() * 8, hipMemcpyHostToDevice));
    }
    bem_solve_info loc{};
    if (precision == 32) {
        if ((r = s.init32())) return r;
        launch_to_f32(L.s1, (size_t)Np * Np, s.Dr, s.Di, s.D32r, s.D32i);
    }
    HIP_TRY(hipEventRecord(L.t[0], L.s1));
    factor_system(L, precision);
    HIP_TRY(hipEventRecord(L.t[1], L.s1));
    solve_system(L, precision, nrhs);
    HIP_TRY(hipEventRecord(L.t[2], L.s1));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(L.s1));
    RefineCtrl ctrl{};
    if (precision == 32) HIP_TRY(hipMemcpy(&ctrl, L.ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
    int used = precision;
    if (precision == 32 && !ctrl.converged) {
        // refinement stagnated: factor and solve the FP64 matrix (still intact:
        // mixed mode factors the FP32 copy)
        factor_system(L, 64);
        solve_system(L, 64, nrhs);
        HIP_TRY(hipEventRecord(L.t[2], L.s1));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(L.s1));
        used = 64;
    }
    std::vector<int> flags(Np / kLuNb + 1);
    HIP_TRY(hipMemcpy(flags.data(), used == 64 ? s.flag : s.flag32, flags.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < Np / kLuNb; i++) loc.pivot_fallback += (flags[i] & 1) ? 1 : 0;
    loc.refine_iters = ctrl.iters;
    loc.residual = precision == 32 ? ctrl.rel : 0.0;
    loc.fell_back_fp64 = (precision == 32 && used == 64) ? 1 : 0;
    float t1 = 0, t2 = 0;
    (void)hipEventElapsedTime(&t1, L.t[0], L.t[1]);
    (void)hipEventElapsedTime(&t2, L.t[1], L.t[2]);
    loc.ms_factor = t1; loc.ms_solve = t2;
    std::vector<double> xr((size_t)Np * kMaxRhs), xi((size_t)Np * kMaxRhs);
    HIP_TRY(hipMemcpy(xr.data(), s.Xr, xr.size() * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(xi.data(), s.Xi, xi.size() * 8, hipMemcpyDeviceToHost));
    planar_to_interleaved(xr, xi, n, Np, nrhs, X);
    if (info) *info = loc;
    if (flags[Np / kLuNb] & 2) return fail(BEM_ERR_NUMERIC, "bem_dense_solve: numerically singular matrix (zero or non-finite pivot)");
    for (size_t q = 0; q < 2 * (size_t)n * nrhs; q++)
        if (!std::isfinite(X[q])) return fail(BEM_ERR_NUMERIC, "bem_dense_solve: non-finite solution");
    return BEM_OK;
}

int bem_dense_solve_noalloc(bem_ctx* c, int precision, int n, const double* A, int nrhs, const double* B,
                            double* X, bem_solve_info* info) {
    if (!c || n <= 0 || !A || !B || !X || nrhs < 1 || nrhs > kMaxRhs || (precision != 64 && precision != 32))
        return fail(BEM_ERR_ARG, "bem_dense_solve_noalloc: bad arguments (nrhs in [1,16], precision 64 or 32)");
    if (n > kNaMaxN)
        return fail(BEM_ERR_ARG, "bem_dense_solve_noalloc: n = " + std::to_string(n) + " exceeds the maximum " +
                                     std::to_string(kNaMaxN) + " of its fixed device array (bem_dense_solve has no limit)");
    for (size_t q = 0; q < 2 * (size_t)n * n; q++)
        if (!std::isfinite(A[q])) return fail(BEM_ERR_ARG, "bem_dense_solve_noalloc: A has a non-finite entry");
    for (size_t q = 0; q < 2 * (size_t)n * nrhs; q++)
        if (!std::isfinite(B[q])) return fail(BEM_ERR_ARG, "bem_dense_solve_noalloc: B has a non-finite entry");
    HIP_TRY(hipSetDevice(c->device));
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_na_mu);
        base = static_cast<char*>(g_na_ptr[c->device]);
    }
    if (!base) return fail(BEM_ERR_STATE, "bem_dense_solve_noalloc: device array not set up");
    const int Np = round_up(n, kPad);
    if (matvec_kslabs(Np) > kNaKslabs) return fail(BEM_ERR_ARG, "bem_dense_solve_noalloc: order too large");
    constexpr NaLayout L = na_layout();
    auto at = [base](size_t off) { return static_cast<void*>(base + off); };
    LinSys s;    // a view: owns nothing, frees nothing
    s.N = n; s.Np = Np; s.capNp = kNaNp; s.have32 = true;
    s.Dr = (double*)at(L.Dr); s.Di = (double*)at(L.Di);
    s.D32r = (float*)at(L.D32r); s.D32i = (float*)at(L.D32i);
    s.ipiv = (int*)at(L.ipiv); s.flag = (int*)at(L.flag); s.ipiv32 = (int*)at(L.ipiv32); s.flag32 = (int*)at(L.flag32);
    double** w64[12] = {&s.w64.sd_r, &s.w64.sd_i, &s.w64.s21_r, &s.w64.s21_i, &s.w64.Linv_r, &s.w64.Linv_i,
                        &s.w64.Uinv_r, &s.w64.Uinv_i, &s.w64.PL_r, &s.w64.PL_i, &s.w64.PU_r, &s.w64.PU_i};
    float** w32[12] = {&s.w32.sd_r, &s.w32.sd_i, &s.w32.s21_r, &s.w32.s21_i, &s.w32.Linv_r, &s.w32.Linv_i,
                       &s.w32.Uinv_r, &s.w32.Uinv_i, &s.w32.PL_r, &s.w32.PL_i, &s.w32.PU_r, &s.w32.PU_i};
    for (int k = 0; k < 12; k++) { *w64[k] = (double*)at(L.w64[k]); *w32[k] = (float*)at(L.w32[k]); }
    s.w32.umax = (float*)at(L.umax);
    s.Br = (double*)at(L.Br); s.Bi = (double*)at(L.Bi); s.Xr = (double*)at(L.Xr); s.Xi = (double*)at(L.Xi);
    s.Rr = (double*)at(L.Rr); s.Ri = (double*)at(L.Ri);
    s.X32r = (float*)at(L.X32r); s.X32i = (float*)at(L.X32i);
    s.Pr = (double*)at(L.Pr); s.Pi = (double*)at(L.Pi); s.Pabs = (double*)at(L.Pabs); s.anorm = (double*)at(L.anorm);
    double* stA = (double*)at(L.stageA);
    double* stB = (double*)at(L.stageB);
    SolveView V;
    V.s1 = c->stream; V.s2 = c->na_s2; V.ev = c->na_ev; V.sys = &s;
    V.counter = (unsigned long long*)at(L.counter); V.pub = (unsigned long long*)at(L.pub);
    V.ctrl = (RefineCtrl*)at(L.ctrl); V.refine_max = c->refine_max; V.lu_arith = c->lu_arith;
    V.pivot_tau = c->pivot_tau;
    hipStream_t st = c->stream;
    HIP_TRY(hipMemcpyAsync(stA, A, 16 * (size_t)n * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(stB, B, 16 * (size_t)n * nrhs, hipMemcpyHostToDevice, st));
    // every element the solve reads is rewritten here: the padded matrix
    // (identity on the padding diagonal) and the padded right-hand sides
    launch_pack_planar(st, n, n, Np, Np, stA, s.Dr, s.Di, 1);
    launch_pack_planar(st, n, nrhs, Np, nrhs, stB, s.Br, s.Bi, 0);
    if (precision == 32) launch_to_f32(st, (size_t)Np * Np, s.Dr, s.Di, s.D32r, s.D32i);
    HIP_TRY(hipEventRecord(c->na_t[0], st));
    factor_system(V, precision);
    HIP_TRY(hipEventRecord(c->na_t[1], st));
    solve_system(V, precision, nrhs);
    HIP_TRY(hipEventRecord(c->na_t[2], st));
    HIP_TRY(hipGetLastError());
    RefineCtrl ctrl{};
    if (precision == 32) HIP_TRY(hipMemcpyAsync(&ctrl, V.ctrl, sizeof(ctrl), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    int used = precision;
    if (precision == 32 && !ctrl.converged) {   // the FP64 matrix is intact: mixed mode factors the copy
        factor_system(V, 64);
        solve_system(V, 64, nrhs);
        HIP_TRY(hipEventRecord(c->na_t[2], st));
        HIP_TRY(hipGetLastError());
        used = 64;
    }
    int flags[kNaNp / kLuNb + 1] = {};
    HIP_TRY(hipMemcpyAsync(flags, used == 64 ? s.flag : s.flag32, (Np / kLuNb + 1) * sizeof(int),
                           hipMemcpyDeviceToHost, st));
    launch_unpack_planar(st, n, nrhs, Np, s.Xr, s.Xi, stB);
    HIP_TRY(hipMemcpyAsync(X, stB, 16 * (size_t)n * nrhs, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    bem_solve_info loc{};
    for (int i = 0; i < Np / kLuNb; i++) loc.pivot_fallback += (flags[i] & 1) ? 1 : 0;
    loc.refine_iters = ctrl.iters;
    loc.residual = precision == 32 ? ctrl.rel : 0.0;
    loc.fell_back_fp64 = (precision == 32 && used == 64) ? 1 : 0;
    float t1 = 0, t2 = 0;
    (void)hipEventElapsedTime(&t1, c->na_t[0], c->na_t[1]);
    (void)hipEventElapsedTime(&t2, c->na_t[1], c->na_t[2]);
    loc.ms_factor = t1; loc.ms_solve = t2;
    if (info) *info = loc;
    if (flags[Np / kLuNb] & 2)
        return fail(BEM_ERR_NUMERIC, "bem_dense_solve_noalloc: numerically singular matrix (zero or non-finite pivot)");
    for (size_t q = 0; q < 2 * (size_t)n * nrhs; q++)
        if (!std::isfinite(X[q])) return fail(BEM_ERR_NUMERIC, "bem_dense_solve_noalloc: non-finite solution");
    return BEM_OK;
}

}  // extern "C"
