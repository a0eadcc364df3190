// ECS halos on the C grid and the operators built on them (included into swm_capi.hip after
// swm_cgrid.hip and swm_ecs.hip).
//
// Vector halo: get_halo_vector of ecs_halo_vec_c_t ("ecs_C_vec", contravariant components) and
// ecs_halo_vec_c_cov_t ("ecs_C_vec_covariant"), equiang_cs/ecs_halo_vec_c_mod.f90:76-595 with the
// tables of ecs_halo_vec_c_factory_mod.f90:18-449.  The reference exchanges 7 rows beyond every
// edge (create_symmetric_halo_vec_exchange_C), averages the normal component's shared faces
// (contravariant only), shifts its rows by one towards the panel and then, edge by edge,
// interpolates each halo row along the edge and mixes in the other component interpolated to the
// same points (4-point rows across the edge).  Every value it writes is a function of the panel's
// own edge lines and the exchanged rows, so here one gather collects the rows of every edge
// (normal component at face depths 0..7, tangential at centre depths 0..7) and one launch
// computes each output point on its own.  Output: u (6, N+2hw, N+1+2hw) and v (6, N+1+2hw,
// N+2hw) for i, j from 1-hw, NaN in the corner squares, which the procedure leaves undefined.
//
// ECS_O halo on the C grid's o points (for gradient_c2_ecs, grad_c2_ecs_mod.f90:33-51, and
// massflux_c_up4, massflux_Cgrid_mod.f90:87-104): the same procedure as swm_ecs_halo_o
// (swm::ecs_o_fill), reading the exchanged cells from the gathered strips.

namespace swmc {

constexpr int DE = 8;   // rows gathered beyond each edge

__global__ void k_gather(long long n, const long long *__restrict__ src, const signed char *__restrict__ sign,
                         const double *__restrict__ base, double *__restrict__ out) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double v = base[src[t]];
    out[t] = (sign && sign[t] < 0) ? -v : v;
}

// ECS_O procedure on o points stored in the C layout; the exchanged cells come from the strips
// S[p][e][d][t-1] (e: 0 i < 1, 1 i > N, 2 j < 1, 3 j > N; d = 0 the neighbour's first cell row)
struct EcsOC {
    using value_type = double;
    const double *f, *S, *w;
    const int *ind;
    int N, hw, rows, pitch;

    __device__ double at(int p, int i, int j) const { return f[((long long)p * rows + j - 1 + H) * pitch + i - 1 + H]; }
    __device__ double raw(int p, int e, int d, int t) const { return S[(((long long)p * 4 + e) * DE + d) * N + t - 1]; }
    __device__ double pf(int p, int I, int J) const {
        if (I >= 1 && I <= N && J >= 1 && J <= N) return at(p, I, J);
        if (J < 1) return raw(p, 2, -J, I);
        if (J > N) return raw(p, 3, J - N - 1, I);
        if (I < 1) return raw(p, 0, -I, J);
        return raw(p, 1, I - N - 1, J);
    }
    __device__ double W(int s, int t, int r) const { return w[((r - 1) * (N + 4) + t + 1) * 4 + s + 1]; }
};

struct CVecArgs {
    const double *u, *v;        // the state's u and v fields (stored C layout)
    const double *En, *Et;      // strips: normal [p][e][D][N], tangential [p][e][D][N+1]
    const double *w0, *w1, *w2, *q1, *q2, *ct;
    const int *i0, *i1, *i2;
    double *out_u, *out_v;
    int N, hw, rows, pitch;
    long long nu, nv;
};

__device__ __forceinline__ double c_dot4(const double *w, const double r0, const double r1, const double r2,
                                         const double r3) {
    double s = w[0] * r0;
    s = s + w[1] * r1;
    s = s + w[2] * r2;
    s = s + w[3] * r3;
    return s;
}

// the panel's own normal component on edge e's line at position k
__device__ __forceinline__ double c_own_normal(const CVecArgs &g, int p, int e, int k) {
    const int N = g.N;
    const long long pb = (long long)p * g.rows;
    if (e == 0) return g.u[(pb + k - 1 + H) * g.pitch + H];
    if (e == 1) return g.u[(pb + k - 1 + H) * g.pitch + N + H];
    if (e == 2) return g.v[(pb + H) * g.pitch + k - 1 + H];
    return g.v[(pb + N + H) * g.pitch + k - 1 + H];
}

// halo row jr (1..hw) of edge e at position k: the normal component (interp_edge, then for the
// covariant kind normal_comp_transform) or the tangential one (interp_edge, then
// tangent_comp_transform or the covariant stretch)
template <bool COV>
__device__ double c_vec_strip(const CVecArgs &g, int p, int e, int jr, int k, bool normal) {
    const int N = g.N, hw = g.hw, o = (k - 1) * hw + (jr - 1);
    const bool low = e == 0 || e == 2;
    const double *En = g.En + ((long long)p * 4 + e) * DE * N;
    const double *Et = g.Et + ((long long)p * 4 + e) * DE * (N + 1);
    if (normal) {
        const double *row = En + (long long)jr * N - 1;            // shifted row jr = depth jr
        const int b = g.i1[o];
        const double x = c_dot4(g.w1 + 4 * o, row[b - 1], row[b], row[b + 1], row[b + 2]);
        if (!COV) return x;
        // the tangential component at the point: rows at centre depths d0..d0+3
        const int d0 = jr == 1 ? 0 : jr - 2, bb = g.i2[o];
        double s[4];
        for (int m = 0; m < 4; ++m) {
            const double *r = Et + (long long)(d0 + m) * (N + 1) - 1;
            s[m] = c_dot4(g.w2 + 4 * o, r[bb - 1], r[bb], r[bb + 1], r[bb + 2]);
        }
        const double s0 = jr == 1 ? (s[3] - 5.0 * s[2] + 15.0 * s[1] + 5.0 * s[0]) / 16.0
                                  : (-s[0] + 9.0 * s[1] + 9.0 * s[2] - s[3]) / 16.0;
        return low ? g.q1[o] * x - g.q2[o] * s0 : g.q1[o] * x + g.q2[o] * s0;
    }
    const double *row = Et + (long long)(jr - 1) * (N + 1) - 1;  // centre depth jr - 1
    const int b = g.i0[o];
    const double x = c_dot4(g.w0 + 4 * o, row[b - 1], row[b], row[b + 1], row[b + 2]);
    if (COV) return x * g.ct[o];
    // the normal component at the point: its rows at distances m0..m0+3 from the edge line, after
    // the average of the shared faces (m = 0) and the shift (depth m up to hw + 1, then m - 1)
    const int m0 = jr == 1 ? 0 : jr - 2, bb = g.i2[o];
    double s[4];
    for (int m = 0; m < 4; ++m) {
        const int mm = m0 + m;
        double r4[4];
        for (int q = 0; q < 4; ++q) {
            const int kk = bb - 1 + q;
            if (mm == 0) r4[q] = 0.5 * (c_own_normal(g, p, e, kk) + En[kk - 1]);
            else r4[q] = En[(long long)(mm <= hw + 1 ? mm : mm - 1) * N + kk - 1];
        }
        s[m] = c_dot4(g.w2 + 4 * o, r4[0], r4[1], r4[2], r4[3]);
    }
    const double s0 = jr == 1 ? (s[3] - 5.0 * s[2] + 15.0 * s[1] + 5.0 * s[0]) / 16.0
                              : (-s[0] + 9.0 * s[1] + 9.0 * s[2] - s[3]) / 16.0;
    return low ? g.q1[o] * x - g.q2[o] * s0 : g.q1[o] * x + g.q2[o] * s0;
}

// one thread per output point: u (6, N+2hw, N+1+2hw) first, then v (6, N+1+2hw, N+2hw)
template <bool COV>
__global__ void k_cvec_fill(CVecArgs g) {
    // 32-bit index arithmetic (nu + nv < 2^32 for every panel size the library accepts)
    unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)(g.nu + g.nv)) return;
    const int N = g.N, hw = g.hw;
    const bool isu = t < (unsigned)g.nu;
    if (!isu) t -= (unsigned)g.nu;
    const int nx = isu ? N + 1 : N, ny = isu ? N : N + 1;
    const int Lx = nx + 2 * hw, Ly = ny + 2 * hw;
    const unsigned per = (unsigned)Lx * (unsigned)Ly;
    const int p = (int)(t / per), r = (int)(t - (unsigned)p * per);
    const int jj = r / Lx, ii = r - jj * Lx;
    const int i = ii + 1 - hw, j = jj + 1 - hw;
    const bool in_i = i >= 1 && i <= nx, in_j = j >= 1 && j <= ny;
    double val;
    if (in_i && in_j) {
        const double *f = isu ? g.u : g.v;
        val = f[((long long)p * g.rows + j - 1 + H) * g.pitch + i - 1 + H];
        if (!COV) {
            // the normal component's faces on the panel edges: mean with the neighbour's copy
            const int e = isu ? (i == 1 ? 0 : i == nx ? 1 : -1) : (j == 1 ? 2 : j == ny ? 3 : -1);
            if (e >= 0) val = 0.5 * (val + g.En[((long long)p * 4 + e) * DE * N + (isu ? j : i) - 1]);
        }
    } else if (in_i || in_j) {
        int e, jr, k;
        if (in_j) { e = i < 1 ? 0 : 1; jr = i < 1 ? 1 - i : i - nx; k = j; }
        else { e = j < 1 ? 2 : 3; jr = j < 1 ? 1 - j : j - ny; k = i; }
        const bool normal = isu ? e < 2 : e >= 2;
        val = c_vec_strip<COV>(g, p, e, jr, k, normal);
    } else {
        val = __builtin_nan("");
    }
    (isu ? g.out_u : g.out_v)[((long long)p * Ly + jj) * Lx + ii] = val;
}

// gradient_c2_ecs from the extended o field F (6, N+2, N+2): x points i = 1..N+1, j = 0..N+1,
// then y points i = 0..N+1, j = 1..N+1
__global__ void k_grad_c2(const double *__restrict__ F, int N, int rows, int pitch, double d, double *__restrict__ gx,
                          double *__restrict__ gy) {
    const unsigned per = (unsigned)(N + 1) * (unsigned)(N + 2), t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 >= 12u * per) return;
    const bool isx = t0 < 6u * per;
    const unsigned t = isx ? t0 : t0 - 6u * per;
    const int p = (int)(t / per), r = (int)(t - (unsigned)p * per);
    const int L = N + 2;
    int i, j;
    if (isx) { i = r % (N + 1) + 1; j = r / (N + 1); }
    else { i = r % (N + 2); j = r / (N + 2) + 1; }
    const double *Fp = F + (long long)p * L * L;
    const double a = Fp[(long long)j * L + i];      // F(i, j) at [j-1+1][i-1+1]
    const double b = isx ? Fp[(long long)j * L + i - 1] : Fp[(long long)(j - 1) * L + i];
    (isx ? gx : gy)[((long long)p * rows + j - 1 + H) * pitch + i - 1 + H] = (a - b) / d;
}

// massflux_c_up4 from the extended o field F (6, N+6, N+6): x points, then y points
__global__ void k_up4(const double *__restrict__ F, const double *__restrict__ u, const double *__restrict__ v, int N,
                      int rows, int pitch, double *__restrict__ fx, double *__restrict__ fy) {
    const unsigned per = (unsigned)(N + 1) * (unsigned)N, t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 >= 12u * per) return;
    const bool isx = t0 < 6u * per;
    const unsigned t = isx ? t0 : t0 - 6u * per;
    const int p = (int)(t / per), r = (int)(t - (unsigned)p * per);
    const int hw = 3, L = N + 2 * hw;
    int i, j;
    if (isx) { i = r % (N + 1) + 1; j = r / (N + 1) + 1; }
    else { i = r % N + 1; j = r / N + 1; }
    const double *Fp = F + (long long)p * L * L;
    auto FE = [&](int a, int b) { return Fp[(long long)(b - 1 + hw) * L + a - 1 + hw]; };
    const long long off = ((long long)p * rows + j - 1 + H) * pitch + i - 1 + H;
    const double w = isx ? u[off] : v[off];
    const double zl = 0.5 + copysign(0.5, w), zr = 1.0 - zl;
    double A, B;
    if (isx) {
        A = 5.0 * FE(i, j) + 15.0 * FE(i - 1, j) - 5.0 * FE(i - 2, j) + FE(i - 3, j);
        B = 5.0 * FE(i - 1, j) + 15.0 * FE(i, j) - 5.0 * FE(i + 1, j) + FE(i + 2, j);
    } else {
        A = 5.0 * FE(i, j) + 15.0 * FE(i, j - 1) - 5.0 * FE(i, j - 2) + FE(i, j - 3);
        B = 5.0 * FE(i, j - 1) + 15.0 * FE(i, j) - 5.0 * FE(i, j + 1) + FE(i, j + 2);
    }
    (isx ? fx : fy)[off] = w * (zl * A + zr * B) / 16.0;
}

}  // namespace swmc

namespace {

// strip sources and buffers (once per context)
int cecs_setup_strips(swm_cctx *c) {
    if (c->ecs_src) return SWM_OK;
    const int N = c->N;
    const size_t nn = (size_t)24 * swmc::DE * N, nt = (size_t)24 * swmc::DE * (N + 1);
    std::vector<long long> src(nn + nt + nn);
    std::vector<signed char> sg(nn + nt);
    if (swmh_c_ecs_strip_sources(N, swmc::H, c->rows, c->pitch, c->fs, swmc::DE, src.data(), sg.data(),
                                 src.data() + nn, sg.data() + nn, src.data() + nn + nt))
        return fail(SWM_EARG, "C-grid ECS strips do not match the cube topology");
    for (size_t k = 0; k < nn + nt; ++k)
        if (src[k] < 0 || src[k] >= 2 * c->fs) return fail(SWM_EARG, "C-grid ECS strip offset out of range");
    for (size_t k = nn + nt; k < src.size(); ++k)
        if (src[k] < 0 || src[k] >= c->fs) return fail(SWM_EARG, "C-grid ECS strip offset out of range");
    HIP_TRY(hipMalloc((void **)&c->ecs_src, src.size() * sizeof(long long)));
    HIP_TRY(hipMalloc((void **)&c->ecs_sign, sg.size()));
    HIP_TRY(hipMalloc((void **)&c->ecs_strips, src.size() * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(c->ecs_src, src.data(), src.size() * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->ecs_sign, sg.data(), sg.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SWM_OK;
}

// gather the normal and tangential strips of the state's (u, v) (which = 0) or the cell strips
// of an o field (which = 1)
int cecs_gather(swm_cctx *c, int which, const double *base) {
    const int N = c->N;
    const long long nn = 24LL * swmc::DE * N, nt = 24LL * swmc::DE * (N + 1);
    const long long n = which == 0 ? nn + nt : nn, first = which == 0 ? 0 : nn + nt;
    hipLaunchKernelGGL(swmc::k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                       c->ecs_src + first, which == 0 ? c->ecs_sign : nullptr, base, c->ecs_strips + first);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

int cecs_setup_o(swm_cctx *c, int rows) {
    const int N = c->N, TT = N + 4;
    int rc = cecs_setup_strips(c);
    if (rc) return rc;
    if (!c->ecs_ext) HIP_TRY(hipMalloc((void **)&c->ecs_ext, (size_t)6 * (N + 6) * (N + 6) * sizeof(double)));
    if (c->ecs_orows >= rows) return SWM_OK;
    std::vector<double> w((size_t)4 * rows * TT);
    std::vector<int> ind((size_t)rows * TT);
    swmh_ecs_o_weights(N, rows, w.data(), ind.data());
    for (int r = 1; r <= rows; ++r)
        for (int t = r == 1 ? 2 : 1; t <= (r == 1 ? N - 1 : N); ++t) {
            const int b = ind[(size_t)(r - 1) * TT + t + 1];
            if (b - 1 < 1 || b + 2 > N) return fail(SWM_EARG, "ECS O-point halo stencil leaves the panel");
        }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->ecs_ow) { HIP_TRY(hipFree(c->ecs_ow)); c->ecs_ow = nullptr; }
    if (c->ecs_oind) { HIP_TRY(hipFree(c->ecs_oind)); c->ecs_oind = nullptr; }
    HIP_TRY(hipMalloc((void **)&c->ecs_ow, w.size() * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->ecs_oind, ind.size() * sizeof(int)));
    HIP_TRY(hipMemcpyAsync(c->ecs_ow, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->ecs_oind, ind.data(), ind.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->ecs_orows = rows;
    return SWM_OK;
}

// the ECS_O halo of width hw (1..3) of the o field f into c->ecs_ext
int cecs_o_halo(swm_cctx *c, const double *f, int hw) {
    int rc = cecs_setup_o(c, hw > 2 ? hw : 2);
    if (rc || (rc = cecs_gather(c, 1, f))) return rc;
    const int N = c->N;
    const long long nn = 24LL * swmc::DE * N, nt = 24LL * swmc::DE * (N + 1);
    const swmc::EcsOC o{f, c->ecs_strips + nn + nt, c->ecs_ow, c->ecs_oind, N, hw, c->rows, c->pitch};
    const long long total = 6LL * (N + 2 * hw) * (N + 2 * hw);
    hipLaunchKernelGGL(swm::ecs_o_fill<swmc::EcsOC>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream, o,
                       c->ecs_ext);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

int cecs_setup_tables(swm_cctx *c, int hw, int cov) {
    if (c->ecs_vt && c->ecs_vt_hw == hw && c->ecs_vt_cov == cov) return SWM_OK;
    const int N = c->N;
    const size_t np = (size_t)(N + 1) * hw;
    std::vector<double> d(15 * np, 0.0);   // w0, w1, w2 (4 np each), q1, q2, ct
    std::vector<int> k(3 * np, 2);         // i0, i1, i2
    if (swmh_c_ecs_vec_tables(N, hw, cov, d.data(), k.data(), d.data() + 4 * np, k.data() + np, d.data() + 8 * np,
                              k.data() + 2 * np, d.data() + 12 * np, d.data() + 13 * np, d.data() + 14 * np))
        return fail(SWM_EARG, "C-grid ECS vector halo tables failed");
    // every stencil stays on the rows the strips hold: positions 1..N (normal, and the other
    // component's rows), 1..N+1 (tangential)
    for (size_t q = 0; q < np; ++q) {
        if (k[q] < 2 || k[q] + 2 > N + 1) return fail(SWM_EARG, "C-grid ECS stencil out of range");
        if (k[2 * np + q] < 2 || k[2 * np + q] + 2 > N) return fail(SWM_EARG, "C-grid ECS stencil out of range");
        if (q < (size_t)N * hw && (k[np + q] < 2 || k[np + q] + 2 > N)) return fail(SWM_EARG, "C-grid ECS stencil out of range");
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!c->ecs_vt) {
        const size_t cap = (size_t)(N + 1) * 5;   // the widest tables (hw = 5)
        HIP_TRY(hipMalloc((void **)&c->ecs_vt, 15 * cap * sizeof(double)));
        HIP_TRY(hipMalloc((void **)&c->ecs_vti, 3 * cap * sizeof(int)));
    }
    HIP_TRY(hipMemcpyAsync(c->ecs_vt, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->ecs_vti, k.data(), k.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->ecs_vt_hw = hw;
    c->ecs_vt_cov = cov;
    return SWM_OK;
}

}  // namespace

extern "C" int swm_c_ecs_halo_vec(swm_cctx *c, const char *halo_procedure_name, const void *state, int halo_width,
                                  void *out_u, void *out_v) {
    if (!c || !halo_procedure_name || !state || !out_u || !out_v) return fail(SWM_EARG, "NULL argument");
    const std::string name(halo_procedure_name);
    int cov;
    if (name == "ecs_C_vec") cov = 0;
    else if (name == "ecs_C_vec_covariant") cov = 1;
    else return fail(SWM_ENAME, "unsupported C-grid vector halo procedure: " + name);
    if (halo_width < 1 || halo_width > 5)
        return fail(SWM_EARG, "halo_width must be in [1, 5] (the procedure reads halo_width + 2 of the 7 exchanged rows)");
    int rc = cecs_setup_strips(c);
    if (rc || (rc = cecs_setup_tables(c, halo_width, cov))) return rc;
    const double *u = (const double *)state + c->fs, *v = (const double *)state + 2 * c->fs;
    if ((rc = cecs_gather(c, 0, u))) return rc;
    const int N = c->N, hw = halo_width;
    const size_t np = (size_t)(N + 1) * hw;
    const long long nn = 24LL * swmc::DE * N;
    swmc::CVecArgs g;
    g.u = u;
    g.v = v;
    g.En = c->ecs_strips;
    g.Et = c->ecs_strips + nn;
    g.w0 = c->ecs_vt;
    g.w1 = c->ecs_vt + 4 * np;
    g.w2 = c->ecs_vt + 8 * np;
    g.q1 = c->ecs_vt + 12 * np;
    g.q2 = c->ecs_vt + 13 * np;
    g.ct = c->ecs_vt + 14 * np;
    g.i0 = c->ecs_vti;
    g.i1 = c->ecs_vti + np;
    g.i2 = c->ecs_vti + 2 * np;
    g.out_u = (double *)out_u;
    g.out_v = (double *)out_v;
    g.N = N;
    g.hw = hw;
    g.rows = c->rows;
    g.pitch = c->pitch;
    g.nu = 6LL * (N + 2 * hw) * (N + 1 + 2 * hw);
    g.nv = g.nu;
    const dim3 grid((unsigned)((g.nu + g.nv + 255) / 256)), block(256);
    if (cov) hipLaunchKernelGGL(swmc::k_cvec_fill<true>, grid, block, 0, c->stream, g);
    else hipLaunchKernelGGL(swmc::k_cvec_fill<false>, grid, block, 0, c->stream, g);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

extern "C" int swm_c_grad(swm_cctx *c, const char *grad_op_name, const void *f, void *gx, void *gy) {
    if (!c || !grad_op_name || !f || !gx || !gy) return fail(SWM_EARG, "NULL argument");
    if (std::string(grad_op_name) != "gradient_c2_ecs")
        return fail(SWM_ENAME, std::string("unsupported C-grid gradient for this entry point: ") + grad_op_name);
    int rc = cecs_o_halo(c, (const double *)f, 1);
    if (rc) return rc;
    const int N = c->N;
    const long long n = 12LL * (N + 1) * (N + 2);
    hipLaunchKernelGGL(swmc::k_grad_c2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->ecs_ext, N,
                       c->rows, c->pitch, c->hx * c->scale, (double *)gx, (double *)gy);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

extern "C" int swm_c_massflux(swm_cctx *c, const char *massflux_op_name, const void *f, const void *u, const void *v,
                              void *fx, void *fy) {
    if (!c || !massflux_op_name || !f || !u || !v || !fx || !fy) return fail(SWM_EARG, "NULL argument");
    if (std::string(massflux_op_name) != "massflux_c_up4")
        return fail(SWM_ENAME, std::string("unsupported C-grid massflux for this entry point: ") + massflux_op_name);
    int rc = cecs_o_halo(c, (const double *)f, 3);
    if (rc) return rc;
    const int N = c->N;
    const long long n = 12LL * (N + 1) * N;
    hipLaunchKernelGGL(swmc::k_up4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->ecs_ext,
                       (const double *)u, (const double *)v, N, c->rows, c->pitch, (double *)fx, (double *)fy);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}
