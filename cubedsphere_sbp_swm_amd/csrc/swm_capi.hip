// C ABI of the device path: context set-up (geometry, neighbour tables, work buffers),
// kernel launches, HIP-graph captured RK4 stepping and FP64 diagnostics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/swm_gpu.h"
#include "geometry.h"
#include "swm_aux.hip"
#include "swm_launch.h"

namespace {

thread_local std::string g_err;

constexpr int RK4_GRAPH_STEPS = 50;  // steps per replay of the multi-step graph

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return fail(SWM_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int NDEPTH = 6;  // neighbour-table depths kept on the host (>= widest closure row)

struct AdjWork;  // adjoint work space (swm_adjoint.hip), created on first use
void free_adj_work(AdjWork *w);
struct HdWork;   // hyperdiffusion work space (swm_hordiff.hip), created on first use
void free_hd_work(HdWork *w);

}  // namespace

struct swm_ctx {
    int N = 0, n = 0, pitch = 0;
    long long fs = 0;
    int grad_order = 42, div_kind = 42, quad_order = 42, prec = 64;
    double scale = 1, omega = 1, grav = 1, hx = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // device geometry in the context precision
    void *J = nullptr, *Qi11 = nullptr, *Qi12 = nullptr, *Qi22 = nullptr, *f = nullptr;
    void *q[4] = {nullptr, nullptr, nullptr, nullptr};
    int4 *nbr = nullptr;
    int *einfo = nullptr;
    double *tanA = nullptr;
    bool otf = false;   // FP64: metric terms evaluated on the fly from tan(alpha)
    int tile_tx = 64, tile_ty = 8;  // tile shape of the stage kernel (swm::tile_shape_ok)
    void *hs = nullptr;
    unsigned hs_gen = 0;  // bumped whenever hs is replaced (ensemble tables hold the pointer)
    // FP64 copies used by the diagnostics
    double *J64 = nullptr, *Qi64[3] = {nullptr, nullptr, nullptr}, *f64 = nullptr, *A64 = nullptr;
    double *hs64 = nullptr;
    // RK4 work buffers (3 fields each) and diagnostics scratch
    void *yA = nullptr, *yB = nullptr, *acc = nullptr, *curl = nullptr;
    double *red = nullptr;
    int red_blocks = 0;
    // captured RK4 steps: one step, and RK4_GRAPH_STEPS steps in a row (fewer graph launches)
    hipGraphExec_t graph = nullptr, graph_multi = nullptr;
    void *graph_state = nullptr;
    double graph_dt = 0.0;
    std::vector<int> host_nbr;
    AdjWork *adj = nullptr;
    HdWork *hd = nullptr;
    // ECS cubic halo: weights / base indices of the last halo width used, neighbour rows
    int ecs_hw = 0;
    void *ecs_w = nullptr, *ecs_ind = nullptr, *ecs_src = nullptr;
    // ECS cubic halo on O points: weight rows set up so far, weights, base indices, edge map;
    // transform matrices of the vector halo and their rows
    int ecs_o_rows = 0, ecs_tm_rows = 0;
    void *ecs_o_w = nullptr, *ecs_o_ind = nullptr, *ecs_o_edge = nullptr, *ecs_tm = nullptr;
    // every panel holds bitwise the same J, Qi11, Qi12, Qi22 (checked at creation)
    bool metric_shared = false;

    size_t esz() const { return prec == 64 ? 8 : 4; }
    size_t field_bytes() const { return (size_t)fs * esz(); }
};

namespace {

template <typename T>
swm::DevGeo<T> make_geo(const swm_ctx *c) {
    swm::DevGeo<T> g;
    g.J = (const T *)c->J;
    g.Qi11 = (const T *)c->Qi11;
    g.Qi12 = (const T *)c->Qi12;
    g.Qi22 = (const T *)c->Qi22;
    g.f = (const T *)c->f;
    g.qb = (const T *)c->q[0];
    g.qt = (const T *)c->q[1];
    g.ql = (const T *)c->q[2];
    g.qr = (const T *)c->q[3];
    g.einfo = c->einfo;
    g.tanA = c->tanA;
    g.n = c->n;
    g.pitch = c->pitch;
    g.fs = c->fs;
    g.mps = c->metric_shared ? 0 : (long long)c->n * c->pitch;
    g.hxs = (T)(c->hx * c->scale);
    g.hx = (T)c->hx;
    g.scale = (T)c->scale;
    g.grav = (T)c->grav;
    return g;
}

int run_stage(swm_ctx *c, int mode, const void *y, void *out, const void *v0, void *acc, double cy, double cacc,
              void *zeta = nullptr) {
    swm::StageLaunch r{};
    r.prec = c->prec;
    r.mode = mode;
    r.tx = c->tile_tx;
    r.ty = c->tile_ty;
    r.otf = c->otf && c->prec == 64;
    r.stream = c->stream;
    if (c->prec == 64) {
        r.gd = make_geo<double>(c);
        r.ad = {(const double *)y, (double *)out, (const double *)v0, (double *)acc, (const double *)c->hs, cy, cacc,
                (double *)zeta};
    } else {
        r.gf = make_geo<float>(c);
        r.af = {(const float *)y, (float *)out, (const float *)v0, (float *)acc, (const float *)c->hs, (float)cy,
                (float)cacc, (float *)zeta};
    }
    hipError_t e;
    switch (c->grad_order) {
    case 21: e = swm::launch_stage_o21(r); break;
    case 42: e = swm::launch_stage_o42(r); break;
    case 43: e = swm::launch_stage_o43(r); break;
    case 63: e = swm::launch_stage_o63(r); break;
    default: return fail(SWM_ENAME, "operator set not available on the device path");
    }
    if (e != hipSuccess) return fail(SWM_EHIP, std::string("stage launch: ") + hipGetErrorString(e));
    return SWM_OK;
}

// one RK4 step = 4 fused stage launches (rk4_mod.f90:28-40); `out` = state, or another buffer
// for a timing run that leaves the state unchanged
int enqueue_rk4_step(swm_ctx *c, void *state, double dt, void *out = nullptr) {
    const double c1 = 0.5 * dt, c2 = dt / 6.0, c3 = dt / 3.0;
    int rc;
    if ((rc = run_stage(c, swm::MODE_STAGE1, state, c->yA, nullptr, c->acc, c1, c2))) return rc;
    if ((rc = run_stage(c, swm::MODE_STAGE23, c->yA, c->yB, state, c->acc, c1, c3))) return rc;
    if ((rc = run_stage(c, swm::MODE_STAGE23, c->yB, c->yA, state, c->acc, dt, c3))) return rc;
    if ((rc = run_stage(c, swm::MODE_STAGE4, c->yA, out ? out : state, nullptr, c->acc, c2, 0.0))) return rc;
    return SWM_OK;
}

template <typename T>
int upload_plane(swm_ctx *c, void *dst, const double *src, int hw) {
    const int n = c->n, ld = n + 2 * hw;
    std::vector<T> buf((size_t)c->fs, T(0));
    for (int p = 0; p < 6; ++p)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                buf[((size_t)p * n + j) * c->pitch + i] = (T)src[((size_t)p * ld + (j + hw)) * ld + (i + hw)];
    HIP_TRY(hipMemcpyAsync(dst, buf.data(), buf.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SWM_OK;
}

template <typename T>
int download_plane(swm_ctx *c, const void *src, double *dst, int hw) {
    const int n = c->n, ld = n + 2 * hw;
    std::vector<T> buf((size_t)c->fs);
    HIP_TRY(hipMemcpyAsync(buf.data(), src, buf.size() * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int p = 0; p < 6; ++p)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                dst[((size_t)p * ld + (j + hw)) * ld + (i + hw)] = (double)buf[((size_t)p * n + j) * c->pitch + i];
    return SWM_OK;
}

template <typename T>
int upload_device_plane(swm_ctx *c, void **dst, const std::vector<double> &src, size_t count) {
    std::vector<T> buf(count);
    for (size_t k = 0; k < count; ++k) buf[k] = (T)src[k];
    HIP_TRY(hipMalloc(dst, count * sizeof(T)));
    HIP_TRY(hipMemcpy(*dst, buf.data(), count * sizeof(T), hipMemcpyHostToDevice));
    return SWM_OK;
}

int upload_f64(double **dst, const std::vector<double> &src) {
    HIP_TRY(hipMalloc((void **)dst, src.size() * sizeof(double)));
    HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice));
    return SWM_OK;
}

int parse_names(swm_ctx *c, const char *div, const char *grad, const char *curl, const char *quad) {
    std::string d = div ? div : "", gr = grad ? grad : "", cu = curl ? curl : "", q = quad ? quad : "";
    int dk = 0, go = 0, qo = 0;
    if (d == "divergence_ah42_sbp") dk = 42;
    else if (d == "divergence_ah43_sbp") dk = 43;
    else if (d == "divergence_ah63_sbp") dk = 63;
    else if (d == "divergence_ah2") dk = 2;
    else return fail(SWM_ENAME, "unknown divergence operator: " + d);
    if (gr == "gradient_ah42_sbp_ecs") go = 42;
    else if (gr == "gradient_ah43_sbp_ecs") go = 43;
    else if (gr == "gradient_ah63_sbp_ecs") go = 63;
    else if (gr == "gradient_ah21_sbp_ecs") go = 21;
    else return fail(SWM_ENAME, "unknown gradient operator: " + gr);
    // curl_factory_mod.f90:17-18: "curl_div..." names wrap the named divergence
    if (cu.rfind("curl_", 0) != 0 || cu.substr(5) != d)
        return fail(SWM_ENAME, "curl operator must be curl_ + divergence name: " + cu);
    if (q == "SBP_Ah21_quadrature") qo = 21;
    else if (q == "SBP_Ah42_quadrature") qo = 42;
    else if (q == "SBP_Ah63_quadrature") qo = 63;
    else return fail(SWM_ENAME, "unknown quadrature: " + q);
    // the fused stage kernel pairs gradient and divergence of the same SBP family
    const bool ok = (dk == go) || (dk == 2 && go == 21);
    if (!ok) return fail(SWM_ENAME, "unsupported gradient/divergence pair: " + gr + " + " + d);
    c->div_kind = dk;
    c->grad_order = go;
    c->quad_order = qo;
    return SWM_OK;
}

void drop_graphs(swm_ctx *c) {
    if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    if (c->graph_multi) { (void)hipGraphExecDestroy(c->graph_multi); c->graph_multi = nullptr; }
    c->graph_state = nullptr;
}

void free_ctx(swm_ctx *c) {
    if (!c) return;
    drop_graphs(c);
    void *ptrs[] = {c->J, c->Qi11, c->Qi12, c->Qi22, c->f, c->q[0], c->q[1], c->q[2], c->q[3], c->nbr, c->einfo, c->tanA, c->hs,
                    c->J64, c->Qi64[0], c->Qi64[1], c->Qi64[2], c->f64, c->A64, c->hs64, c->yA, c->yB, c->acc,
                    c->curl, c->red, c->ecs_w, c->ecs_ind, c->ecs_src, c->ecs_o_w, c->ecs_o_ind, c->ecs_o_edge,
                    c->ecs_tm};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    free_adj_work(c->adj);
    free_hd_work(c->hd);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
extern "C" const char *swm_last_error(void) { return g_err.c_str(); }

extern "C" int swm_ctx_create(swm_ctx **out, int N, const char *div, const char *grad, const char *curl,
                              const char *quad, double scale, double omega, double grav, int precision_bits) {
    if (!out) return fail(SWM_EARG, "ctx pointer is NULL");
    *out = nullptr;
    if (N < 8 || N > 8192) return fail(SWM_EARG, "N must be in [8, 8192]");  // n = N+1 > tile height
    if (precision_bits != 64 && precision_bits != 32) return fail(SWM_EARG, "precision_bits must be 64 or 32");
    if (!(scale > 0) || !(grav > 0)) return fail(SWM_EARG, "scale and grav must be positive");
    swm_ctx *c = new swm_ctx();
    int rc = parse_names(c, div, grad, curl, quad);
    if (rc) { delete c; return rc; }
    c->N = N;
    c->n = N + 1;
    c->pitch = (c->n + 7) / 8 * 8;
    c->fs = 6LL * c->n * c->pitch;
    c->prec = precision_bits;
    c->scale = scale;
    c->omega = omega;
    c->grav = grav;
    c->hx = (0.25 * std::acos(-1.0) - (-0.25 * std::acos(-1.0))) / (double)N;
    auto bail = [&](int code) { free_ctx(c); return code; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SWM_EHIP, "hipStreamCreate failed"));
    c->own_stream = true;

    const size_t F = (size_t)c->fs;
    std::vector<double> J(F, 0.0), Qi11(F, 0.0), Qi12(F, 0.0), Qi22(F, 0.0), fc(F, 0.0);
    const double axis[3] = {0.0, 0.0, 1.0};
    swmh_geometry(N, c->pitch, omega, axis, J.data(), Qi11.data(), Qi12.data(), Qi22.data(), fc.data(), nullptr,
                  nullptr, nullptr);
    {
        // the ECS metric depends on (alpha, beta) only: when all six panels agree bit for bit
        // the stage kernels read panel 0's copy for every panel, a sixth of the metric traffic
        const size_t slab = (size_t)c->n * c->pitch;
        bool same = true;
        for (const auto *m : {&J, &Qi11, &Qi12, &Qi22})
            for (int p = 1; p < 6 && same; ++p)
                same = std::memcmp(m->data(), m->data() + p * slab, slab * sizeof(double)) == 0;
        c->metric_shared = same;
    }
    std::vector<double> qw[4];
    for (auto &v : qw) v.assign(c->n, 0.0);
    swmh_sync_weights(N, qw[0].data(), qw[1].data(), qw[2].data(), qw[3].data());
    std::vector<double> A(c->n);
    swmh_quad_weights(N, c->quad_order, A.data());
    c->host_nbr.assign((size_t)6 * 4 * c->n * NDEPTH * 4, 0);
    swmh_halo_table(N, NDEPTH, c->host_nbr.data());

    if (c->prec == 64) {
        if ((rc = upload_device_plane<double>(c, &c->J, J, F)) || (rc = upload_device_plane<double>(c, &c->Qi11, Qi11, F)) ||
            (rc = upload_device_plane<double>(c, &c->Qi12, Qi12, F)) || (rc = upload_device_plane<double>(c, &c->Qi22, Qi22, F)) ||
            (rc = upload_device_plane<double>(c, &c->f, fc, F)))
            return bail(rc);
        for (int k = 0; k < 4; ++k)
            if ((rc = upload_device_plane<double>(c, &c->q[k], qw[k], c->n))) return bail(rc);
    } else {
        if ((rc = upload_device_plane<float>(c, &c->J, J, F)) || (rc = upload_device_plane<float>(c, &c->Qi11, Qi11, F)) ||
            (rc = upload_device_plane<float>(c, &c->Qi12, Qi12, F)) || (rc = upload_device_plane<float>(c, &c->Qi22, Qi22, F)) ||
            (rc = upload_device_plane<float>(c, &c->f, fc, F)))
            return bail(rc);
        for (int k = 0; k < 4; ++k)
            if ((rc = upload_device_plane<float>(c, &c->q[k], qw[k], c->n))) return bail(rc);
    }
    if ((rc = upload_f64(&c->J64, J)) || (rc = upload_f64(&c->Qi64[0], Qi11)) || (rc = upload_f64(&c->Qi64[1], Qi12)) ||
        (rc = upload_f64(&c->Qi64[2], Qi22)) || (rc = upload_f64(&c->f64, fc)) || (rc = upload_f64(&c->A64, A)))
        return bail(rc);
    // per (panel, edge) orientation of the neighbour panel: q, its normal axis (x?), its index
    // of the shared edge along that axis, and tangential index = a_base + s*k
    std::vector<int> einfo(6 * 4 * 8, 0);
    for (int p = 0; p < 6; ++p)
        for (int e = 0; e < 4; ++e) {
            auto at = [&](int k, int d) { return &c->host_nbr[((((size_t)p * 4 + e) * c->n + k) * NDEPTH + d) * 4]; };
            const int *t0 = at(0, 0), *t1 = at(0, 1), *tk = at(1, 0);
            const int nu_x = t0[1] != t1[1] ? 1 : 0;
            const int tnu = nu_x ? t0[1] : t0[2];
            const int abase = nu_x ? t0[2] : t0[1];
            const int sgn = (nu_x ? tk[2] : tk[1]) - abase;
            for (int k = 0; k < c->n; ++k)
                for (int d = 0; d < NDEPTH; ++d) {
                    const int *t = at(k, d);
                    const int a_ = nu_x ? t[2] : t[1], nu_ = nu_x ? t[1] : t[2];
                    const int nu_expect = tnu == 0 ? d : c->n - 1 - d;
                    if (t[0] != t0[0] || a_ != abase + sgn * k || nu_ != nu_expect || (sgn != 1 && sgn != -1))
                        return bail(fail(SWM_EARG, "neighbour table is not an edge strip"));
                }
            int *ei = &einfo[(p * 4 + e) * 8];
            for (int k = 0; k < c->n; ++k)
                for (int d = 0; d < NDEPTH; ++d)
                    if (at(k, d)[3] != t0[3]) return bail(fail(SWM_EARG, "vector code varies along an edge"));
            ei[0] = t0[0]; ei[1] = nu_x; ei[2] = tnu; ei[3] = abase; ei[4] = sgn; ei[5] = t0[3];
        }
    if (hipMalloc((void **)&c->einfo, einfo.size() * sizeof(int)) != hipSuccess ||
        hipMemcpy(c->einfo, einfo.data(), einfo.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(SWM_EHIP, "edge info upload failed"));
    {
        // tan(alpha_i) of the grid coordinate alpha_i = -pi/4 + (i-1)*hx (mesh_mod.f90:83-92)
        std::vector<double> ta(c->n);
        for (int i = 0; i < c->n; ++i) ta[i] = std::tan(-0.25 * std::acos(-1.0) + ((double)i + 0.0) * c->hx);
        if ((rc = upload_f64(&c->tanA, ta))) return bail(rc);
    }
    if (hipMalloc((void **)&c->nbr, c->host_nbr.size() * sizeof(int)) != hipSuccess ||
        hipMemcpy(c->nbr, c->host_nbr.data(), c->host_nbr.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(SWM_EHIP, "neighbour table upload failed"));
    const size_t sb = 3 * c->field_bytes();
    if (hipMalloc(&c->yA, sb) != hipSuccess || hipMalloc(&c->yB, sb) != hipSuccess || hipMalloc(&c->acc, sb) != hipSuccess ||
        hipMalloc(&c->curl, c->field_bytes()) != hipSuccess)
        return bail(fail(SWM_EHIP, "work buffer allocation failed"));
    // zero the padding columns once so that every buffer holds defined values
    if (hipMemset(c->yA, 0, sb) != hipSuccess || hipMemset(c->yB, 0, sb) != hipSuccess ||
        hipMemset(c->acc, 0, sb) != hipSuccess || hipMemset(c->curl, 0, c->field_bytes()) != hipSuccess)
        return bail(fail(SWM_EHIP, "work buffer clear failed"));
    c->red_blocks = 1024;
    if (hipMalloc((void **)&c->red, sizeof(double) * 16 * (c->red_blocks + 1)) != hipSuccess)
        return bail(fail(SWM_EHIP, "reduction scratch allocation failed"));
    // the copies and clears above run on the null stream, which the context's non-blocking
    // stream does not wait for: finish them before any kernel of the context can start
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(SWM_EHIP, "set-up synchronisation failed"));
    *out = c;
    return SWM_OK;
}

extern "C" int swm_ctx_destroy(swm_ctx *c) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_ctx(c);
    return SWM_OK;
}

extern "C" int swm_ctx_info(const swm_ctx *c, int *n, int *pitch, long long *fe, int *eb) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    if (n) *n = c->n;
    if (pitch) *pitch = c->pitch;
    if (fe) *fe = c->fs;
    if (eb) *eb = (int)c->esz();
    return SWM_OK;
}

extern "C" int swm_ctx_set_stream(swm_ctx *c, void *s) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize(c->stream));
    drop_graphs(c);
    if (s) {
        if (c->own_stream) { (void)hipStreamDestroy(c->stream); c->own_stream = false; }
        c->stream = (hipStream_t)s;
    } else if (!c->own_stream) {
        HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    }
    return SWM_OK;
}

extern "C" int swm_ctx_synchronize(swm_ctx *c) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SWM_OK;
}

extern "C" int swm_state_alloc(swm_ctx *c, void **state) {
    if (!c || !state) return fail(SWM_EARG, "NULL argument");
    const size_t sb = 3 * c->field_bytes();
    HIP_TRY(hipMalloc(state, sb));
    // cleared on the context stream: a null-stream clear could land after a later upload
    HIP_TRY(hipMemsetAsync(*state, 0, sb, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SWM_OK;
}

extern "C" int swm_state_free(swm_ctx *c, void *state) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    if (!state) return SWM_OK;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->graph_state == state) drop_graphs(c);
    HIP_TRY(hipFree(state));
    return SWM_OK;
}

extern "C" int swm_state_upload(swm_ctx *c, void *state, const double *h, const double *u, const double *v, int hw) {
    if (!c || !state || !h || !u || !v) return fail(SWM_EARG, "NULL argument");
    if (hw < 0 || hw > 64) return fail(SWM_EARG, "halo_width out of range");
    const double *src[3] = {h, u, v};
    for (int k = 0; k < 3; ++k) {
        char *dst = (char *)state + k * c->field_bytes();
        int rc = c->prec == 64 ? upload_plane<double>(c, dst, src[k], hw) : upload_plane<float>(c, dst, src[k], hw);
        if (rc) return rc;
    }
    return SWM_OK;
}

extern "C" int swm_state_download(swm_ctx *c, const void *state, double *h, double *u, double *v, int hw) {
    if (!c || !state || !h || !u || !v) return fail(SWM_EARG, "NULL argument");
    if (hw < 0 || hw > 64) return fail(SWM_EARG, "halo_width out of range");
    double *dst[3] = {h, u, v};
    for (int k = 0; k < 3; ++k) {
        const char *src = (const char *)state + k * c->field_bytes();
        int rc = c->prec == 64 ? download_plane<double>(c, src, dst[k], hw) : download_plane<float>(c, src, dst[k], hw);
        if (rc) return rc;
    }
    return SWM_OK;
}

extern "C" int swm_rhs(swm_ctx *c, const void *vin, void *vout) {
    if (!c || !vin || !vout) return fail(SWM_EARG, "NULL argument");
    if (vin == vout) return fail(SWM_EARG, "vin and vout must differ");
    return run_stage(c, swm::MODE_RHS, vin, vout, nullptr, nullptr, 0.0, 0.0);
}

template <typename T>
static int axpby(swm_ctx *c, bool update, void *out, double s1, const void *v1, double s2, const void *v2) {
    const int threads = 256, blocks = 2048;
    if (update)
        hipLaunchKernelGGL((swm::axpby_kernel<T, true>), dim3(blocks), dim3(threads), 0, c->stream, (T *)out, (T)s1,
                           (const T *)v1, (T)s2, (const T *)v2, c->n, c->pitch, c->fs);
    else
        hipLaunchKernelGGL((swm::axpby_kernel<T, false>), dim3(blocks), dim3(threads), 0, c->stream, (T *)out, (T)s1,
                           (const T *)v1, (T)s2, (const T *)v2, c->n, c->pitch, c->fs);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

extern "C" int swm_assign(swm_ctx *c, void *out, double s1, const void *v1, double s2, const void *v2) {
    if (!c || !out || !v1 || !v2) return fail(SWM_EARG, "NULL argument");
    return c->prec == 64 ? axpby<double>(c, false, out, s1, v1, s2, v2) : axpby<float>(c, false, out, s1, v1, s2, v2);
}

extern "C" int swm_update(swm_ctx *c, void *out, double s1, const void *v1, double s2, const void *v2) {
    if (!c || !out || !v1 || !v2) return fail(SWM_EARG, "NULL argument");
    return c->prec == 64 ? axpby<double>(c, true, out, s1, v1, s2, v2) : axpby<float>(c, true, out, s1, v1, s2, v2);
}

extern "C" int swm_rk4_steps(swm_ctx *c, void *state, double dt, int nsteps) {
    if (!c || !state) return fail(SWM_EARG, "NULL argument");
    if (nsteps < 0) return fail(SWM_EARG, "nsteps must be >= 0");
    if (nsteps == 0) return SWM_OK;
    if (c->graph_state != state || c->graph_dt != dt) drop_graphs(c);
    // whole runs of RK4_GRAPH_STEPS steps replay the multi-step graph (4 * RK4_GRAPH_STEPS
    // kernel nodes, one graph launch), the rest the one-step graph
    const int nmulti = nsteps / RK4_GRAPH_STEPS, nsingle = nsteps - nmulti * RK4_GRAPH_STEPS;
    for (int which = 0; which < 2; ++which) {
        hipGraphExec_t &ge = which == 0 ? c->graph : c->graph_multi;
        const int per = which == 0 ? 1 : RK4_GRAPH_STEPS;
        if (ge || (which == 0 ? nsingle : nmulti) == 0) continue;
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = SWM_OK;
        for (int k = 0; k < per && !rc; ++k) rc = enqueue_rk4_step(c, state, dt);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return fail(SWM_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) { ge = nullptr; return fail(SWM_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e)); }
        c->graph_state = state;
        c->graph_dt = dt;
    }
    for (int s = 0; s < nmulti; ++s) HIP_TRY(hipGraphLaunch(c->graph_multi, c->stream));
    for (int s = 0; s < nsingle; ++s) HIP_TRY(hipGraphLaunch(c->graph, c->stream));
    return SWM_OK;
}

extern "C" int swm_ctx_set_tile_shape(swm_ctx *c, int tile_cols, int tile_rows, int metric_on_the_fly) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    if (!swm::tile_shape_ok(tile_cols, tile_rows))
        return fail(SWM_EARG, "tile shape must be 64 x 8, 32 x 16, 16 x 32 or 32 x 8 points");
    if (metric_on_the_fly && c->prec != 64) return fail(SWM_EARG, "metric on the fly is FP64 only");
    // a tile must not touch both the j=1 and the j=n edge of a panel
    if (c->n <= tile_rows) return fail(SWM_EARG, "tile_rows must be smaller than N+1");
    HIP_TRY(hipStreamSynchronize(c->stream));
    drop_graphs(c);
    c->tile_tx = tile_cols;
    c->tile_ty = tile_rows;
    c->otf = metric_on_the_fly != 0;
    return SWM_OK;
}

extern "C" int swm_ctx_get_tile_shape(const swm_ctx *c, int *tile_cols, int *tile_rows, int *metric_on_the_fly) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    if (tile_cols) *tile_cols = c->tile_tx;
    if (tile_rows) *tile_rows = c->tile_ty;
    if (metric_on_the_fly) *metric_on_the_fly = c->otf ? 1 : 0;
    return SWM_OK;
}

extern "C" int swm_ctx_set_tiling(swm_ctx *c, int tile_rows, int metric_on_the_fly) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    if (tile_rows != 8 && tile_rows != 16 && tile_rows != 32)
        return fail(SWM_EARG, "tile_rows must be 8, 16 or 32 (tiles of 64 x 8, 32 x 16, 16 x 32 points)");
    return swm_ctx_set_tile_shape(c, swm::tile_width(tile_rows), tile_rows, metric_on_the_fly);
}

extern "C" int swm_ctx_get_tiling(const swm_ctx *c, int *tile_rows, int *metric_on_the_fly) {
    return swm_ctx_get_tile_shape(c, nullptr, tile_rows, metric_on_the_fly);
}

// Autotuner: times `reps` RK4 steps on `state` (the four stage kernels with the buffers of the
// step, the last writing yB so that the state is unchanged) for every tile shape and, for FP64,
// metric read or evaluated on the fly, with HIP events on the context stream; keeps the fastest
// and returns its time per step.
extern "C" int swm_ctx_autotune(swm_ctx *c, void *state, int reps, double *best_us) {
    if (!c || !state || reps < 1) return fail(SWM_EARG, "bad argument");
    struct Cand { int tx, ty; bool otf; };
    std::vector<Cand> cands;
    for (int otf = 0; otf < (c->prec == 64 ? 2 : 1); ++otf)
        for (auto sh : {std::make_pair(64, 8), std::make_pair(32, 16), std::make_pair(16, 32), std::make_pair(32, 8)})
            cands.push_back({sh.first, sh.second, otf != 0});
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // two interleaved rounds over the candidates, the minimum of each: configurations within a
    // few per cent of each other are not decided by one noisy measurement
    std::vector<double> t(cands.size(), 1e300);
    for (int round = 0; round < 2; ++round)
        for (size_t q = 0; q < cands.size(); ++q) {
            const Cand &cd = cands[q];
            if (c->n <= cd.ty) continue;
            c->tile_tx = cd.tx;
            c->tile_ty = cd.ty;
            c->otf = cd.otf;
            int rc = 0;
            for (int r = 0; r < 2 && !rc; ++r) rc = enqueue_rk4_step(c, state, 1.0, c->yB);
            if (rc) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
            HIP_TRY(hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps && !rc; ++r) rc = enqueue_rk4_step(c, state, 1.0, c->yB);
            HIP_TRY(hipEventRecord(e1, c->stream));
            HIP_TRY(hipEventSynchronize(e1));
            if (rc) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            t[q] = std::min(t[q], 1e3 * ms / reps);
        }
    double best = 1e300;
    Cand pick = cands[0];
    for (size_t q = 0; q < cands.size(); ++q)
        if (t[q] < best) { best = t[q]; pick = cands[q]; }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->tile_tx = pick.tx;
    c->tile_ty = pick.ty;
    c->otf = pick.otf;
    drop_graphs(c);
    if (best_us) *best_us = best;
    return SWM_OK;
}

// Measurement helper: `reps` back-to-back launches of one fused stage kernel on the context
// stream (stage 1, 2 (= 2 and 3) or 4) with the buffers the RK4 step gives it: stage 1 reads
// `state` and writes yA, acc; stage 2 reads yA, v0 = state, acc and writes yB, acc; stage 4
// reads yA, acc and writes yB (in place of the state).  yA starts as a copy of `state`.
extern "C" int swm_bench_stage(swm_ctx *c, void *state, int stage_id, int reps) {
    if (!c || !state || reps < 0) return fail(SWM_EARG, "bad argument");
    if (stage_id != 1 && stage_id != 2 && stage_id != 4) return fail(SWM_EARG, "stage_id must be 1, 2 or 4");
    if (stage_id != 1) HIP_TRY(hipMemcpyAsync(c->yA, state, 3 * c->field_bytes(), hipMemcpyDeviceToDevice, c->stream));
    for (int r = 0; r < reps; ++r) {
        int rc;
        if (stage_id == 1) rc = run_stage(c, swm::MODE_STAGE1, state, c->yA, nullptr, c->acc, 1.0, 1.0);
        else if (stage_id == 2) rc = run_stage(c, swm::MODE_STAGE23, c->yA, c->yB, state, c->acc, 1.0, 0.0);
        else rc = run_stage(c, swm::MODE_STAGE4, c->yA, c->yB, nullptr, c->acc, 1.0, 0.0);
        if (rc) return rc;
    }
    return SWM_OK;
}

extern "C" int swm_set_orography(swm_ctx *c, const double *hs, int hw) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize(c->stream));
    drop_graphs(c);
    if (c->hs) { HIP_TRY(hipFree(c->hs)); c->hs = nullptr; }
    if (c->hs64) { HIP_TRY(hipFree(c->hs64)); c->hs64 = nullptr; }
    ++c->hs_gen;
    if (!hs) return SWM_OK;
    HIP_TRY(hipMalloc(&c->hs, c->field_bytes()));
    HIP_TRY(hipMalloc((void **)&c->hs64, (size_t)c->fs * sizeof(double)));
    int rc = c->prec == 64 ? upload_plane<double>(c, c->hs, hs, hw) : upload_plane<float>(c, c->hs, hs, hw);
    if (rc) return rc;
    return upload_plane<double>(c, c->hs64, hs, hw);
}

extern "C" int swm_ic_spherical_harmonics(swm_ctx *c, unsigned long long seed, int lmax, double h0, double h_amp,
                                          double v_max, double *h, double *u, double *v, int hw) {
    if (!c || !h || !u || !v) return fail(SWM_EARG, "NULL argument");
    if (lmax < 1 || lmax > 512) return fail(SWM_EARG, "lmax out of range");
    const int n = c->n, ld = n + 2 * hw;
    std::vector<double> H((size_t)c->fs), U((size_t)c->fs), V((size_t)c->fs);
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (swmh_ic_sh(c->N, c->pitch, seed, lmax, h0, h_amp, v_max, H.data(), U.data(), V.data(), (int)nt))
        return fail(SWM_EARG, "initial-state generator failed");
    for (int p = 0; p < 6; ++p)
        for (int j = 0; j < ld; ++j)
            for (int i = 0; i < ld; ++i) {
                const size_t o = ((size_t)p * ld + j) * ld + i;
                const int jj = j - hw, ii = i - hw;
                const bool in = ii >= 0 && ii < n && jj >= 0 && jj < n;
                const size_t s = in ? ((size_t)p * n + jj) * c->pitch + ii : 0;
                h[o] = in ? H[s] : 0.0;
                u[o] = in ? U[s] : 0.0;
                v[o] = in ? V[s] : 0.0;
            }
    return SWM_OK;
}

extern "C" int swm_debug_geometry(swm_ctx *c, int which, double *out, int hw) {
    if (!c || !out) return fail(SWM_EARG, "NULL argument");
    void *src = which == 0 ? c->J : which == 1 ? c->Qi11 : which == 2 ? c->Qi12 : which == 3 ? c->Qi22 : which == 4 ? c->f : nullptr;
    if (!src) return fail(SWM_EARG, "which out of range");
    return c->prec == 64 ? download_plane<double>(c, src, out, hw) : download_plane<float>(c, src, out, hw);
}

extern "C" int swm_debug_halo_table(swm_ctx *c, int *table) {
    if (!c || !table) return fail(SWM_EARG, "NULL argument");
    std::memcpy(table, c->host_nbr.data(), c->host_nbr.size() * sizeof(int));
    return SWM_OK;
}

#include "swm_diag.hip"
#include "swm_adjoint.hip"
#include "swm_ensemble.hip"
#include "swm_ecs.hip"
#include "swm_hordiff.hip"
#include "swm_output.hip"
#include "swm_cgrid.hip"
#include "swm_cgrid_ecs.hip"
