// C-ABI of the library (include/bemuse_amd.h): meshes, contexts, bodies and
// the per-frequency solve. Host orchestration only; arithmetic is on the GPU.
#include "../../include/bemuse_amd.h"
#include "bem_internal.h"
#include "bem_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
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

}  // namespace

struct bem_mesh {
    Mesh m;
};

struct bem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t panel_stream = nullptr;  // LU lookahead stream (same priority)
    std::vector<hipEvent_t> lu_events;   // 2 per LU panel, grown on demand
    hipEvent_t ev[6] = {};
    unsigned* counter = nullptr;     // ticket counter of the persistent solves
    unsigned* flags = nullptr;       // per-block epoch flags
    unsigned epoch = 0;
    int max_blocks = 0;
};

// Dense system workspace for one padded size Np.
struct LinSys {
    int N = 0, Np = 0;
    double *Dr = nullptr, *Di = nullptr;          // FP64 system matrix (LU in place in FP64 mode)
    float *D32r = nullptr, *D32i = nullptr;       // FP32 copy factored in mixed mode
    int *ipiv = nullptr, *flag = nullptr;
    int *ipiv32 = nullptr, *flag32 = nullptr;
    LuWork<double> w64;
    LuWork<float> w32;
    double *Br = nullptr, *Bi = nullptr, *Xr = nullptr, *Xi = nullptr, *Rr = nullptr, *Ri = nullptr;
    float *X32r = nullptr, *X32i = nullptr;
    double *Pr = nullptr, *Pi = nullptr, *Pabs = nullptr;   // matvec partials
    double *colmax = nullptr, *anorm = nullptr;
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
        size_t pan = (size_t)Np * kLuNb;
        int r = BEM_OK;
        if ((r = get(&w.sd_r, kLuNb * kLuNb)) || (r = get(&w.sd_i, kLuNb * kLuNb)) || (r = get(&w.s21_r, pan)) ||
            (r = get(&w.s21_i, pan)) || (r = get(&w.Linv_r, pan)) || (r = get(&w.Linv_i, pan)) ||
            (r = get(&w.Uinv_r, pan)) || (r = get(&w.Uinv_i, pan)))
            return r;
        return BEM_OK;
    }
    int init(int n) {
        N = n; Np = round_up(n, kPad);
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
        if ((r = get(&colmax, 2 * kMaxRhs)) || (r = get(&anorm, 1))) return r;
        return BEM_OK;
    }
    int init32() {
        if (have32) return BEM_OK;
        size_t NN = (size_t)Np * Np, NR = (size_t)Np * kMaxRhs;
        int r = BEM_OK;
        if ((r = get(&D32r, NN)) || (r = get(&D32i, NN))) return r;
        if ((r = get(&ipiv32, Np)) || (r = get(&flag32, Np / kLuNb + 1))) return r;
        if ((r = work(w32))) return r;
        if ((r = get(&X32r, NR)) || (r = get(&X32i, NR))) return r;
        have32 = true;
        return BEM_OK;
    }
};

struct bem_body {
    bem_ctx* ctx = nullptr;
    int N = 0, Np = 0;
    BodyGeometry geo;
    double *d_body = nullptr, *d_refl = nullptr, *d_quad = nullptr, *d_recv = nullptr, *d_gen = nullptr;
    double *Ssrc = nullptr, *Srefl = nullptr, *Dsrc = nullptr, *Drefl = nullptr;
    bool have_rankine = false;
    double *Sr = nullptr, *Si = nullptr;            // composed SMat (per frequency)
    double *PhiIr = nullptr, *PhiIi = nullptr, *DAr = nullptr, *DAi = nullptr;
    double *Xinr = nullptr, *Xini = nullptr;        // RHS input block
    double *d_betas = nullptr, *d_out = nullptr;
    LinSys sys;

    ~bem_body() {
        void* ps[] = {d_body, d_refl, d_quad, d_recv, d_gen, Ssrc, Srefl, Dsrc, Drefl, Sr, Si,
                      PhiIr, PhiIi, DAr, DAi, Xinr, Xini, d_betas, d_out};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
};

namespace {

int lu_events(bem_ctx* c, int Np) {
    size_t need = 2 * (size_t)(Np / kLuNb) + 2;
    while (c->lu_events.size() < need) {
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->lu_events.push_back(e);
    }
    return BEM_OK;
}

// Device LU in the given precision (FP64, or FP32 on the FP32 copy).
int factor_system(bem_ctx* c, LinSys& s, int precision, LuProfile* prof = nullptr) {
    int r = lu_events(c, s.Np);
    if (r) return r;
    if (precision == 64) {
        launch_lu<double>(c->stream, c->panel_stream, c->lu_events.data(), s.Np, s.Dr, s.Di, s.w64, s.flag, s.ipiv,
                          prof);
    } else {
        if ((r = s.init32())) return r;
        launch_lu<float>(c->stream, c->panel_stream, c->lu_events.data(), s.Np, s.D32r, s.D32i, s.w32, s.flag32,
                         s.ipiv32, prof);
    }
    HIP_TRY(hipGetLastError());
    return BEM_OK;
}

int count_flags(bem_ctx* c, LinSys& s, int precision, int* fallback) {
    int np = s.Np / kLuNb;
    std::vector<int> f(np + 1);
    HIP_TRY(hipMemcpyAsync(f.data(), precision == 64 ? s.flag : s.flag32, (np + 1) * sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    int n = 0;
    for (int i = 0; i < np; i++) n += (f[i] & 1) ? 1 : 0;
    *fallback = n;
    return (f[np] & 2) ? 1 : 0;
}

// Solve D X = B for nr columns (B, X in the workspace). In FP32 mode the
// FP64 matrix s.Dr/s.Di must hold the unfactored system.
int solve_system(bem_ctx* c, LinSys& s, int precision, int nr, bem_solve_info* info) {
    size_t n = (size_t)s.Np * nr;
    if (precision == 64) {
        launch_copy2(c->stream, n, s.Br, s.Bi, s.Xr, s.Xi);
        launch_lu_solve<double, double>(c->stream, s.Np, nr, s.Dr, s.Di, s.w64, s.ipiv, s.flag + s.Np / kLuNb, s.Xr,
                                        s.Xi, c->counter, c->flags, c->epoch);
        HIP_TRY(hipGetLastError());
        if (info) { info->refine_iters = 0; info->residual = 0.0; }
        return BEM_OK;
    }
    // FP32 LU solve, then refine: r = b - A x (FP64), d = LU32^-1 r, x += d.
    const int* fany = s.flag32 + s.Np / kLuNb;
    launch_to_f32(c->stream, n, s.Br, s.Bi, s.X32r, s.X32i);
    launch_lu_solve<float, float>(c->stream, s.Np, nr, s.D32r, s.D32i, s.w32, s.ipiv32, fany, s.X32r, s.X32i,
                                  c->counter, c->flags, c->epoch);
    HIP_TRY(hipMemsetAsync(s.Xr, 0, n * sizeof(double), c->stream));
    HIP_TRY(hipMemsetAsync(s.Xi, 0, n * sizeof(double), c->stream));
    launch_add_f32(c->stream, n, s.Xr, s.Xi, s.X32r, s.X32i);
    // Stop at a max-norm relative residual of 1e-13: with the system's
    // condition numbers this keeps A, B, F orders of magnitude inside the
    // 1e-6 bar while saving a refinement sweep.
    const int max_iter = 12;
    const double tol = 1e-13;
    double prev = 1e300, rel = 1e300, anorm = 0.0;
    int it = 0;
    std::vector<double> cm(2 * kMaxRhs + 1);
    for (; it <= max_iter; it++) {
        // the first residual also yields |A|_inf (row sums of |re| + |im|)
        launch_matvec<double>(c->stream, s.Np, nr, s.Dr, s.Di, s.Xr, s.Xi, s.Pr, s.Pi, s.Br, s.Bi, 1, s.Rr, s.Ri,
                              it == 0 ? s.Pabs : nullptr, it == 0 ? s.anorm : nullptr);
        launch_colmax(c->stream, s.N, s.Np, nr, s.Rr, s.Ri, s.Xr, s.Xi, s.colmax);
        HIP_TRY(hipMemcpyAsync(cm.data(), s.colmax, 2 * nr * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if (it == 0) HIP_TRY(hipMemcpyAsync(&anorm, s.anorm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        rel = 0;
        for (int j = 0; j < nr; j++) {
            double den = anorm * cm[2 * j + 1];
            double v = den > 0 ? cm[2 * j] / den : (cm[2 * j] > 0 ? 1e300 : 0.0);
            rel = std::max(rel, v);
        }
        if (!(rel == rel)) break;                 // NaN
        if (rel <= tol) break;
        if (it > 0 && rel > 0.5 * prev) break;    // stagnation
        if (it == max_iter) break;
        prev = rel;
        launch_to_f32(c->stream, n, s.Rr, s.Ri, s.X32r, s.X32i);
        launch_lu_solve<float, float>(c->stream, s.Np, nr, s.D32r, s.D32i, s.w32, s.ipiv32, fany, s.X32r, s.X32i,
                                      c->counter, c->flags, c->epoch);
        launch_add_f32(c->stream, n, s.Xr, s.Xi, s.X32r, s.X32i);
    }
    HIP_TRY(hipGetLastError());
    if (info) { info->refine_iters = it; info->residual = rel; }
    return BEM_OK;
}

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
    if (!out || n_nodes <= 0 || n_elems <= 0 || !nodes || !elems) return fail(BEM_ERR_ARG, "bem_mesh_from_arrays: bad arguments");
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
    if (hipStreamCreateWithFlags(&c->panel_stream, hipStreamNonBlocking) != hipSuccess) {
        bem_destroy(c);
        return fail(BEM_ERR_HIP, "hipStreamCreate failed");
    }
    c->max_blocks = 4096;
    int r = dalloc_t(&c->counter, 1);
    if (!r) r = dalloc_t(&c->flags, c->max_blocks);
    if (r) { bem_destroy(c); return r; }
    (void)hipMemset(c->counter, 0, sizeof(unsigned));
    (void)hipMemset(c->flags, 0, c->max_blocks * sizeof(unsigned));
    (void)hipDeviceSynchronize();
    *out = c;
    return BEM_OK;
}

void bem_destroy(bem_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->counter) (void)hipFree(c->counter);
    if (c->flags) (void)hipFree(c->flags);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->lu_events)
        if (e) (void)hipEventDestroy(e);
    if (c->panel_stream) { (void)hipStreamSynchronize(c->panel_stream); (void)hipStreamDestroy(c->panel_stream); }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int bem_body_create(bem_ctx* c, const bem_mesh* m, bem_body** out) {
    if (!c || !m || !out) return fail(BEM_ERR_ARG, "bem_body_create: null argument");
    HIP_TRY(hipSetDevice(c->device));
    bem_body* b = new bem_body;
    b->ctx = c;
    std::string err;
    if (!build_geometry(m->m, b->geo, err)) { delete b; return fail(BEM_ERR_ARG, err); }
    b->N = b->geo.N;
    b->Np = round_up(b->N, kPad);
    if (b->Np / kLuNb > c->max_blocks) { delete b; return fail(BEM_ERR_ARG, "body too large"); }
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

int bem_body_n(const bem_body* b) { return b ? b->N : -1; }

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

static int ensure_freq_buffers(bem_body* b) {
    if (b->Sr) return BEM_OK;
    size_t NN = (size_t)b->Np * b->Np, NR = (size_t)b->Np * kMaxRhs;
    int r = BEM_OK;
    if ((r = dalloc_t(&b->Sr, NN)) || (r = dalloc_t(&b->Si, NN))) return r;
    if ((r = dalloc_t(&b->PhiIr, NR)) || (r = dalloc_t(&b->PhiIi, NR)) || (r = dalloc_t(&b->DAr, NR)) ||
        (r = dalloc_t(&b->DAi, NR)) || (r = dalloc_t(&b->Xinr, NR)) || (r = dalloc_t(&b->Xini, NR)))
        return r;
    if ((r = dalloc_t(&b->d_betas, kMaxRhs)) || (r = dalloc_t(&b->d_out, 2 * (36 + 12 * kMaxRhs)))) return r;
    return b->sys.init(b->N);
}

static void assemble(bem_ctx* c, bem_body* b, int mode, double kappa, int compose) {
    bool f32 = (mode != BEM_MODE_FP64);
    launch_wave_compose(c->stream, b->N, b->Np, kappa, compose, b->d_quad, b->d_recv, b->Ssrc, b->Srefl, b->Dsrc,
                        b->Drefl, b->Sr, b->Si, b->sys.Dr, b->sys.Di, f32 ? b->sys.D32r : nullptr,
                        f32 ? b->sys.D32i : nullptr);
}

int bem_solve_frequency(bem_ctx* c, bem_body* b, int mode, double omega, double kappa, int compose, int rad_only,
                        int nb, const double* betas, double* A36, double* B36, double* F, double* FK, double* SC,
                        double* phiJ, double* phiS, bem_solve_info* info) {
    if (!c || !b || b->ctx != c) return fail(BEM_ERR_ARG, "bem_solve_frequency: bad context/body");
    if (nb < 0 || nb > kMaxRhs - 6 || (nb > 0 && !betas)) return fail(BEM_ERR_ARG, "bem_solve_frequency: nb must be in [0, 10]");
    if (mode < BEM_MODE_FP64 || mode > BEM_MODE_MIXED_SURROGATE) return fail(BEM_ERR_ARG, "bem_solve_frequency: bad mode");
    if (compose < 0 || compose > 2) return fail(BEM_ERR_ARG, "bem_solve_frequency: bad compose case");
    if (!A36 || !B36) return fail(BEM_ERR_ARG, "bem_solve_frequency: A36/B36 required");
    if (!b->have_rankine) return fail(BEM_ERR_STATE, "bem_solve_frequency: Rankine matrices not assembled");
    HIP_TRY(hipSetDevice(c->device));
    int r = ensure_freq_buffers(b);
    if (r) return r;
    if (mode != BEM_MODE_FP64 && (r = b->sys.init32())) return r;
    hipStream_t s = c->stream;
    LinSys& sys = b->sys;
    int nr = 6 + nb;
    bem_solve_info loc{};
    if (nb > 0) HIP_TRY(hipMemcpyAsync(b->d_betas, betas, nb * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(c->ev[0], s));
    assemble(c, b, mode, kappa, compose);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev[1], s));
    int prec = (mode == BEM_MODE_FP64) ? 64 : 32;
    if ((r = factor_system(c, sys, prec))) return r;
    HIP_TRY(hipEventRecord(c->ev[2], s));
    launch_incident(s, b->N, b->Np, nb, b->d_body, b->d_gen, b->d_betas, omega, kappa, b->Xinr, b->Xini, kMaxRhs,
                    b->PhiIr, b->PhiIi, b->DAr, b->DAi);
    launch_matvec<double>(s, b->Np, nr, b->Sr, b->Si, b->Xinr, b->Xini, sys.Pr, sys.Pi, nullptr, nullptr, 0, sys.Br,
                          sys.Bi);
    HIP_TRY(hipGetLastError());
    if ((r = solve_system(c, sys, prec, nr, &loc))) return r;
    if (prec == 32 && !(loc.residual <= 1e-11)) {
        // refinement did not reach FP64 accuracy: refactor in FP64
        assemble(c, b, BEM_MODE_FP64, kappa, compose);
        if ((r = factor_system(c, sys, 64)) || (r = solve_system(c, sys, 64, nr, nullptr))) return r;
        loc.fell_back_fp64 = 1;
        prec = 64;
    }
    launch_reduce_coeffs(s, b->N, b->Np, nb, b->d_gen, sys.Xr, sys.Xi, b->PhiIr, b->PhiIi, b->DAr, b->DAi, b->d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev[3], s));
    std::vector<double> out(2 * (36 + 12 * nb));
    HIP_TRY(hipMemcpyAsync(out.data(), b->d_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    int zero_piv = count_flags(c, sys, prec, &loc.pivot_fallback);
    // A = Re(M Phi_J), B = Im(M Phi_J) (Hydrodynamic_Solver.cpp:463-465)
    for (int q = 0; q < 36; q++) { A36[q] = out[2 * q]; B36[q] = out[2 * q + 1]; }
    if (!rad_only && nb > 0) {
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
    if (phiJ || phiS) {
        std::vector<double> xr((size_t)b->Np * nr), xi((size_t)b->Np * nr);
        HIP_TRY(hipMemcpy(xr.data(), sys.Xr, xr.size() * 8, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(xi.data(), sys.Xi, xi.size() * 8, hipMemcpyDeviceToHost));
        if (phiJ) planar_to_interleaved(xr, xi, b->N, b->Np, 6, phiJ);
        if (phiS && nb > 0) {
            std::vector<double> sr(xr.begin() + (size_t)6 * b->Np, xr.end()), si(xi.begin() + (size_t)6 * b->Np, xi.end());
            planar_to_interleaved(sr, si, b->N, b->Np, nb, phiS);
        }
    }
    float t01 = 0, t12 = 0, t23 = 0;
    (void)hipEventElapsedTime(&t01, c->ev[0], c->ev[1]);
    (void)hipEventElapsedTime(&t12, c->ev[1], c->ev[2]);
    (void)hipEventElapsedTime(&t23, c->ev[2], c->ev[3]);
    loc.ms_assemble = t01; loc.ms_factor = t12; loc.ms_solve = t23;
    if (info) *info = loc;
    (void)zero_piv;
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
    HIP_TRY(hipSetDevice(c->device));
    int r = ensure_freq_buffers(b);
    if (r) return r;
    if (mode != BEM_MODE_FP64 && (r = b->sys.init32())) return r;
    assemble(c, b, mode, kappa, compose);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    int N = b->N, Np = b->Np;
    std::vector<double> re((size_t)Np * N), im((size_t)Np * N);
    auto out = [&](const double* dr, const double* di, double* h) -> int {
        HIP_TRY(hipMemcpy(re.data(), dr, re.size() * 8, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(im.data(), di, im.size() * 8, hipMemcpyDeviceToHost));
        planar_to_interleaved(re, im, N, Np, N, h);
        return BEM_OK;
    };
    if (SMat && (r = out(b->Sr, b->Si, SMat))) return r;
    if (DMat && (r = out(b->sys.Dr, b->sys.Di, DMat))) return r;
    return BEM_OK;
}

int bem_profile_lu(bem_ctx* c, bem_body* b, int mode, double kappa, double* out) {
    if (!c || !b || !b->have_rankine || !out) return fail(BEM_ERR_STATE, "bem_profile_lu: Rankine matrices not assembled");
    HIP_TRY(hipSetDevice(c->device));
    int r = ensure_freq_buffers(b);
    if (r) return r;
    if (mode != BEM_MODE_FP64 && (r = b->sys.init32())) return r;
    assemble(c, b, mode, kappa, BEM_COMPOSE_WAVE);
    LinSys& s = b->sys;
    LuProfile* prof = new LuProfile;
    prof->N = b->N;
    HIP_TRY(hipEventRecord(c->ev[0], c->stream));
    if ((r = factor_system(c, s, mode == BEM_MODE_FP64 ? 64 : 32, prof))) { delete prof; return r; }
    HIP_TRY(hipEventRecord(c->ev[1], c->stream));
    hipError_t e = hipStreamSynchronize(c->stream);
    double sum = 0.0;
    for (int i = 0; i < prof->n; i++) {
        float t = 0;
        (void)hipEventElapsedTime(&t, prof->ev[2 * i], prof->ev[2 * i + 1]);
        sum += t;
    }
    float tot = 0;
    (void)hipEventElapsedTime(&tot, c->ev[0], c->ev[1]);
    out[0] = tot;
    out[1] = sum;
    out[2] = prof->useful_flops;
    out[3] = prof->n;
    out[4] = prof->n ? sum / prof->n : 0.0;
    out[5] = prof->n ? prof->useful_flops / prof->n : 0.0;
    for (auto& ev : prof->ev)
        if (ev) (void)hipEventDestroy(ev);
    delete prof;
    if (e != hipSuccess) return fail(BEM_ERR_HIP, std::string("bem_profile_lu: ") + hipGetErrorString(e));
    return BEM_OK;
}

int bem_eval_havelock(bem_ctx* c, int precision, int n, const double* X, const double* Y, const double* V,
                      double* out) {
    if (!c || n <= 0 || !X || !Y || !V || !out) return fail(BEM_ERR_ARG, "bem_eval_havelock: bad arguments");
    if (precision != 64) return fail(BEM_ERR_ARG, "bem_eval_havelock: precision must be 64");
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
    launch_havelock_eval(c->stream, n, dx, dy, dv, dout);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(out, dout, 8 * (size_t)n * 8, hipMemcpyDeviceToHost);
    (void)hipFree(dx); (void)hipFree(dy); (void)hipFree(dv); (void)hipFree(dout);
    if (e != hipSuccess) return fail(BEM_ERR_HIP, std::string("bem_eval_havelock: ") + hipGetErrorString(e));
    return BEM_OK;
}

int bem_dense_solve(bem_ctx* c, int precision, int n, const double* A, int nrhs, const double* B, double* X,
                    bem_solve_info* info) {
    if (!c || n <= 0 || !A || !B || !X || nrhs < 1 || nrhs > kMaxRhs || (precision != 64 && precision != 32))
        return fail(BEM_ERR_ARG, "bem_dense_solve: bad arguments (nrhs in [1,16], precision 64 or 32)");
    HIP_TRY(hipSetDevice(c->device));
    LinSys s;
    int r = s.init(n);
    if (r) return r;
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
        launch_to_f32(c->stream, (size_t)Np * Np, s.Dr, s.Di, s.D32r, s.D32i);
    }
    HIP_TRY(hipEventRecord(c->ev[0], c->stream));
    if ((r = factor_system(c, s, precision))) return r;
    HIP_TRY(hipEventRecord(c->ev[1], c->stream));
    if ((r = solve_system(c, s, precision, nrhs, &loc))) return r;
    HIP_TRY(hipEventRecord(c->ev[2], c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    count_flags(c, s, precision, &loc.pivot_fallback);
    float t1 = 0, t2 = 0;
    (void)hipEventElapsedTime(&t1, c->ev[0], c->ev[1]);
    (void)hipEventElapsedTime(&t2, c->ev[1], c->ev[2]);
    loc.ms_factor = t1; loc.ms_solve = t2;
    std::vector<double> xr((size_t)Np * kMaxRhs), xi((size_t)Np * kMaxRhs);
    HIP_TRY(hipMemcpy(xr.data(), s.Xr, xr.size() * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(xi.data(), s.Xi, xi.size() * 8, hipMemcpyDeviceToHost));
    planar_to_interleaved(xr, xi, n, Np, nrhs, X);
    if (info) *info = loc;
    return BEM_OK;
}

}  // extern "C"
