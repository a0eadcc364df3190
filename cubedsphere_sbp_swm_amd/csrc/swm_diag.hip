// Global diagnostics of operator_swm_t%get_diagnostics (operator_swm_mod.f90:115-260):
// hmin, hmax, mass, TE, KE, PE, enstrophy.  Integrands follow the reference expressions in
// FP64 (FP32 states are widened first); the quadrature sums are compensated (Neumaier) in
// two levels: per block, then one block over the partials.
#include <hip/hip_runtime.h>

namespace swm {

constexpr int NQ = 5;  // mass, KE_u, KE_v, PE, enstrophy

__device__ __forceinline__ void neu_add(double &s, double &c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

template <typename T>
__global__ void __launch_bounds__(256) diag_partial(const T *y, const T *curl, const double *J, const double *Qi11,
                                                    const double *Qi12, const double *Qi22, const double *f,
                                                    const double *A, const double *hs, int n, int pitch,
                                                    long long fs, double hx, double grav, double *part) {
    double s[NQ], c[NQ];
    for (int q = 0; q < NQ; ++q) s[q] = c[q] = 0.0;
    double hmin = 1e300, hmax = -1e300;
    const long long total = 6LL * n * n;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < total; k += (long long)gridDim.x * blockDim.x) {
        const long long row = k / n;
        const int i = (int)(k - row * n);
        const int j = (int)(row % n);
        const long long id = row * pitch + i;
        const double h = (double)y[id], u = (double)y[id + fs], v = (double)y[id + 2 * fs];
        const double w1 = J[id], ai = A[i], aj = A[j];
        auto term = [&](double val) { return val * w1 * hx * hx * ai * aj; };
        hmin = fmin(hmin, h);
        hmax = fmax(hmax, h);
        // quadrature_h%mass(h) (sbp_quadrature_mod.f90:36)
        neu_add(s[0], c[0], term(h));
        // calc_energy (operator_swm_mod.f90:211-236)
        const double ut = Qi11[id] * u + Qi12[id] * v;
        const double vt = Qi12[id] * u + Qi22[id] * v;
        const double hu = h * ut, hv = h * vt;
        neu_add(s[1], c[1], term(0.5 * hu * u));
        neu_add(s[2], c[2], term(0.5 * hv * v));
        const double pe = h + (hs ? hs[id] : 0.0);
        neu_add(s[3], c[3], term(0.5 * grav * pe * pe));
        // calc_enstrophy (operator_swm_mod.f90:238-260)
        double z = (double)curl[id] + f[id];
        z = 0.5 * z * z;
        neu_add(s[4], c[4], term(z / h));
    }
    __shared__ double sh[256][2 * NQ + 2];
    const int t = threadIdx.x;
    for (int q = 0; q < NQ; ++q) {
        sh[t][2 * q] = s[q];
        sh[t][2 * q + 1] = c[q];
    }
    sh[t][2 * NQ] = hmin;
    sh[t][2 * NQ + 1] = hmax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            for (int q = 0; q < NQ; ++q) {
                double a = sh[t][2 * q], ac = sh[t][2 * q + 1];
                neu_add(a, ac, sh[t + w][2 * q]);
                ac += sh[t + w][2 * q + 1];
                sh[t][2 * q] = a;
                sh[t][2 * q + 1] = ac;
            }
            sh[t][2 * NQ] = fmin(sh[t][2 * NQ], sh[t + w][2 * NQ]);
            sh[t][2 * NQ + 1] = fmax(sh[t][2 * NQ + 1], sh[t + w][2 * NQ + 1]);
        }
        __syncthreads();
    }
    if (t < 2 * NQ + 2) part[blockIdx.x * 16 + t] = sh[0][t];
}

__global__ void __launch_bounds__(64) diag_final(const double *part, int nblocks, double *out) {
    if (threadIdx.x != 0) return;
    double s[NQ], c[NQ];
    for (int q = 0; q < NQ; ++q) s[q] = c[q] = 0.0;
    double hmin = 1e300, hmax = -1e300;
    for (int b = 0; b < nblocks; ++b) {
        for (int q = 0; q < NQ; ++q) {
            neu_add(s[q], c[q], part[b * 16 + 2 * q]);
            c[q] += part[b * 16 + 2 * q + 1];
        }
        hmin = fmin(hmin, part[b * 16 + 2 * NQ]);
        hmax = fmax(hmax, part[b * 16 + 2 * NQ + 1]);
    }
    for (int q = 0; q < NQ; ++q) out[q] = s[q] + c[q];
    out[NQ] = hmin;
    out[NQ + 1] = hmax;
}

// the seven diagnostics in the order of swm_diagnostics from the sums of diag_final (the same
// FP64 additions swm_diagnostics makes on the host)
__global__ void __launch_bounds__(64) diag_order(const double *r, double *out) {
    if (threadIdx.x != 0) return;
    const double ke = r[1] + r[2];
    out[0] = r[5];
    out[1] = r[6];
    out[2] = r[0];
    out[3] = ke + r[3];
    out[4] = ke;
    out[5] = r[3];
    out[6] = r[4];
}

// Scratch of swm_diagnostics_noalloc: the per-block partial sums of diag_partial (16 doubles per
// block, at most NOALLOC_BLOCKS blocks) and the 8 sums of diag_final, 128.1 KiB per device.  Every
// call on a device shares it.
constexpr int NOALLOC_BLOCKS = 1024;
constexpr int NOALLOC_DOUBLES = 16 * NOALLOC_BLOCKS + 16;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

}  // namespace swm

namespace {

// device address of swm::scratch on each device, taken once per device with that device current
constexpr int NOALLOC_MAX_DEVICES = 64;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

int noalloc_scratch(double **ptr) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(SWM_EARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(swm::scratch)));
        if (!p) return fail(SWM_EHIP, "hipGetSymbolAddress returned NULL for the diagnostics scratch");
        g_scratch_ptr[dev] = (double *)p;
    }
    *ptr = g_scratch_ptr[dev];
    return SWM_OK;
}

}  // namespace

// The diagnostics with the partial sums in the device scratch above: the same kernels as
// swm_diagnostics, no allocation and no synchronisation; the seven values are written to the
// device array `out` in stream order on the context stream.
extern "C" int swm_diagnostics_noalloc(swm_ctx *c, const void *state, double *out) {
    if (!c || !state || !out) return fail(SWM_EARG, "NULL argument");
    if (c->red_blocks > swm::NOALLOC_BLOCKS)
        return fail(SWM_EARG, "diagnostics use more partial sums than the no-allocation scratch holds (1024 blocks)");
    double *scr = nullptr;
    int rc = noalloc_scratch(&scr);
    if (rc) return rc;
    rc = run_stage(c, swm::MODE_CURL, state, c->curl, nullptr, nullptr, 0.0, 0.0);
    if (rc) return rc;
    double *part = scr, *res = scr + 16 * swm::NOALLOC_BLOCKS;
    if (c->prec == 64)
        hipLaunchKernelGGL(swm::diag_partial<double>, dim3(c->red_blocks), dim3(256), 0, c->stream, (const double *)state,
                           (const double *)c->curl, c->J64, c->Qi64[0], c->Qi64[1], c->Qi64[2], c->f64, c->A64, c->hs64,
                           c->n, c->pitch, c->fs, c->hx, c->grav, part);
    else
        hipLaunchKernelGGL(swm::diag_partial<float>, dim3(c->red_blocks), dim3(256), 0, c->stream, (const float *)state,
                           (const float *)c->curl, c->J64, c->Qi64[0], c->Qi64[1], c->Qi64[2], c->f64, c->A64, c->hs64,
                           c->n, c->pitch, c->fs, c->hx, c->grav, part);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(swm::diag_final, dim3(1), dim3(64), 0, c->stream, part, c->red_blocks, res);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(swm::diag_order, dim3(1), dim3(64), 0, c->stream, res, out);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

extern "C" int swm_diagnostics(swm_ctx *c, const void *state, double out[7]) {
    if (!c || !state || !out) return fail(SWM_EARG, "NULL argument");
    int rc = run_stage(c, swm::MODE_CURL, state, c->curl, nullptr, nullptr, 0.0, 0.0);
    if (rc) return rc;
    double *part = c->red, *res = c->red + 16 * c->red_blocks;
    if (c->prec == 64)
        hipLaunchKernelGGL(swm::diag_partial<double>, dim3(c->red_blocks), dim3(256), 0, c->stream, (const double *)state,
                           (const double *)c->curl, c->J64, c->Qi64[0], c->Qi64[1], c->Qi64[2], c->f64, c->A64, c->hs64,
                           c->n, c->pitch, c->fs, c->hx, c->grav, part);
    else
        hipLaunchKernelGGL(swm::diag_partial<float>, dim3(c->red_blocks), dim3(256), 0, c->stream, (const float *)state,
                           (const float *)c->curl, c->J64, c->Qi64[0], c->Qi64[1], c->Qi64[2], c->f64, c->A64, c->hs64,
                           c->n, c->pitch, c->fs, c->hx, c->grav, part);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(swm::diag_final, dim3(1), dim3(64), 0, c->stream, part, c->red_blocks, res);
    HIP_TRY(hipGetLastError());
    double r[8];
    HIP_TRY(hipMemcpyAsync(r, res, sizeof(double) * 7, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double ke = r[1] + r[2];
    out[0] = r[5];
    out[1] = r[6];
    out[2] = r[0];
    out[3] = ke + r[3];
    out[4] = ke;
    out[5] = r[3];
    out[6] = r[4];
    return SWM_OK;
}
