// ECS cubic halo on xy (Ah) points (included into swm_capi.hip).
//
// get_ecs_halo / ext_ecs_tile_halo (equiang_cs/ecs_halo_xy_mod.f90:68-219) for a field of the
// context: the panel is extended by `hw` rows of virtual points beyond each edge.  The
// reference exchanges the neighbours' rows at depth 0..hw into a halo, averages the cube
// corners (1/3) and the duplicated edge lines (1/2), sets the halo diagonal at each cube corner
// to the mean of the two neighbours' values there (and substitutes it into both source rows),
// and fills halo row j by 1-D cubic Lagrange interpolation along the neighbour's row at depth
// j (ecs_ext_halo_1e, :221-235) with the weights of init_halo_xy_interp.  Here the same values
// come from two gathers: ecs_sources collects the neighbours' rows (depth 0..hw) of every edge
// through the edge descriptors, and ecs_fill computes every point of the extended panel from
// the field and those rows, with the corner substitutions evaluated where they are read.
//
// Output layout: [panel][jj][ii], jj, ii = 0 .. n+2hw-1 for 1-based j, i = jj+1-hw, ii+1-hw.

namespace swm {

template <typename T>
__global__ void ecs_sources(DevGeo<T> g, const T *__restrict__ f, int hw, T *__restrict__ S) {
    // S[((p*4 + e)*(hw+1) + d)*n + k]: field of edge e's neighbour at depth d, position k
    const int n = g.n;
    const int total = 6 * 4 * (hw + 1) * n;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int k = t % n;
    const int d = (t / n) % (hw + 1);
    const int pe = t / (n * (hw + 1));
    const int *ei = g.einfo + pe * 8;
    const int nidx = ei[2] == 0 ? d : n - 1 - d, tidx = ei[3] + ei[4] * k;
    const int qi = ei[1] ? nidx : tidx, qj = ei[1] ? tidx : nidx;
    S[t] = f[((long long)ei[0] * n + qj) * g.pitch + qi];
}

template <typename T>
__global__ void ecs_fill(DevGeo<T> g, const T *__restrict__ f, int hw, const T *__restrict__ S,
                         const T *__restrict__ W, const int *__restrict__ IND, T *__restrict__ out) {
    const int n = g.n, L = n + 2 * hw;
    const long long total = 6LL * L * L;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int ii = (int)(t % L), jj = (int)((t / L) % L), p = (int)(t / ((long long)L * L));
    const int i = ii + 1 - hw, j = jj + 1 - hw;  // 1-based panel indices
    // neighbour row of edge e (0 west, 1 east, 2 south, 3 north) at depth d, 1-based position k
    auto src = [&](int e, int d, int k) { return S[(((long long)p * 4 + e) * (hw + 1) + d) * n + (k - 1)]; };
    // cube-corner diagonal values at depth d (corners 1..4 = (1,1), (n,1), (n,n), (1,n))
    auto diag = [&](int c, int d) {
        if (c == 1) return T(0.5) * (src(0, d, 1) + src(2, d, 1));
        if (c == 2) return T(0.5) * (src(1, d, 1) + src(2, d, n));
        if (c == 3) return T(0.5) * (src(1, d, n) + src(3, d, n));
        return T(0.5) * (src(3, d, 1) + src(0, d, n));
    };
    T v;
    const bool in_i = i >= 1 && i <= n, in_j = j >= 1 && j <= n;
    if (in_i && in_j) {
        const T c = f[((long long)p * n + (j - 1)) * g.pitch + (i - 1)];
        const bool ei = i == 1 || i == n, ej = j == 1 || j == n;
        if (ei && ej) {
            // cube corners (ecs_halo_xy_mod.f90:114,125,136,147)
            if (i == 1 && j == 1) v = (c + src(2, 0, 1) + src(0, 0, 1)) / T(3);
            else if (i == n && j == 1) v = (c + src(2, 0, n) + src(1, 0, 1)) / T(3);
            else if (i == n && j == n) v = (c + src(3, 0, n) + src(1, 0, n)) / T(3);
            else v = (c + src(0, 0, n) + src(3, 0, 1)) / T(3);
        } else if (j == 1) v = T(0.5) * (c + src(2, 0, i));
        else if (j == n) v = T(0.5) * (c + src(3, 0, i));
        else if (i == 1) v = T(0.5) * (c + src(0, 0, j));
        else if (i == n) v = T(0.5) * (c + src(1, 0, j));
        else v = c;
    } else {
        // halo point: row (depth d) of the edge whose interpolation writes it, position along
        // that edge; row d of an edge covers positions 2-d .. n+d-1, the rest are diagonals
        int e = -1, d = 0, pos = 0;
        if (j <= 0 && i >= 1 + j && i <= n - j) { e = 2; d = 1 - j; pos = i; }
        else if (j > n && i >= 2 - (j - n) && i <= j - 1) { e = 3; d = j - n; pos = i; }
        else if (i <= 0 && j >= 1 + i && j <= n - i) { e = 0; d = 1 - i; pos = j; }
        else if (i > n && j >= 2 - (i - n) && j <= i - 1) { e = 1; d = i - n; pos = j; }
        if (e < 0) {
            // halo diagonals of the cube corners (ecs_halo_xy_mod.f90:116,127,138,149)
            if (i < 1 && j < 1) v = diag(1, 1 - i);
            else if (i > n && j < 1) v = diag(2, i - n);
            else if (i > n && j > n) v = diag(3, i - n);
            else v = diag(4, 1 - i);
        } else {
            // corner-adjacent source positions carry the diagonal value (:117-118,128-129,...)
            const int c_lo = e == 0 ? 1 : e == 1 ? 2 : e == 2 ? 1 : 4;  // corner at position 1
            const int c_hi = e == 0 ? 4 : e == 1 ? 3 : e == 2 ? 2 : 3;  // corner at position n
            auto zb = [&](int k) { return k == 1 ? diag(c_lo, d) : k == n ? diag(c_hi, d) : src(e, d, k); };
            const int o = (d - 1) * L + (pos - 1 + hw);
            const int b = IND[o];
            const T *w = W + 4 * o;
            T s = w[0] * zb(b - 1);
            s = s + w[1] * zb(b);
            s = s + w[2] * zb(b + 1);
            s = s + w[3] * zb(b + 2);
            v = s;
        }
    }
    out[t] = v;
}

}  // namespace swm

namespace {

template <typename T>
int ecs_halo_run(swm_ctx *c, const void *field, int hw, void *out) {
    const int n = c->n, L = n + 2 * hw;
    if (c->ecs_hw != hw) {
        std::vector<double> w((size_t)4 * hw * L);
        std::vector<int> ind((size_t)hw * L);
        if (swmh_ecs_xy_weights(c->N, hw, w.data(), ind.data())) return fail(SWM_EARG, "ECS halo weights failed");
        std::vector<T> wt(w.begin(), w.end());
        if (c->ecs_w) { HIP_TRY(hipFree(c->ecs_w)); c->ecs_w = nullptr; }
        if (c->ecs_ind) { HIP_TRY(hipFree(c->ecs_ind)); c->ecs_ind = nullptr; }
        if (c->ecs_src) { HIP_TRY(hipFree(c->ecs_src)); c->ecs_src = nullptr; }
        HIP_TRY(hipMalloc(&c->ecs_w, wt.size() * sizeof(T)));
        HIP_TRY(hipMalloc((void **)&c->ecs_ind, ind.size() * sizeof(int)));
        HIP_TRY(hipMalloc(&c->ecs_src, (size_t)6 * 4 * (hw + 1) * n * sizeof(T)));
        HIP_TRY(hipMemcpy(c->ecs_w, wt.data(), wt.size() * sizeof(T), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(c->ecs_ind, ind.data(), ind.size() * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipDeviceSynchronize());  // null-stream copies before the context stream reads them
        c->ecs_hw = hw;
    }
    const swm::DevGeo<T> g = make_geo<T>(c);
    const int ns = 6 * 4 * (hw + 1) * n;
    hipLaunchKernelGGL(swm::ecs_sources<T>, dim3((ns + 255) / 256), dim3(256), 0, c->stream, g, (const T *)field, hw,
                       (T *)c->ecs_src);
    const long long total = 6LL * L * L;
    hipLaunchKernelGGL(swm::ecs_fill<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream, g,
                       (const T *)field, hw, (const T *)c->ecs_src, (const T *)c->ecs_w, (const int *)c->ecs_ind,
                       (T *)out);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

}  // namespace

extern "C" int swm_ecs_halo_xy(swm_ctx *c, const void *field, int halo_width, void *out) {
    if (!c || !field || !out) return fail(SWM_EARG, "NULL argument");
    if (halo_width < 1 || halo_width > 8) return fail(SWM_EARG, "halo_width must be in [1, 8]");
    if (c->n < halo_width + 4) return fail(SWM_EARG, "the ECS halo needs N >= halo_width + 3");
    return c->prec == 64 ? ecs_halo_run<double>(c, field, halo_width, out) : ecs_halo_run<float>(c, field, halo_width, out);
}

extern "C" int swm_ecs_halo_xy_host(swm_ctx *c, const void *field, int halo_width, double *out) {
    if (!c || !field || !out) return fail(SWM_EARG, "NULL argument");
    if (halo_width < 1 || halo_width > 8) return fail(SWM_EARG, "halo_width must be in [1, 8]");
    const size_t L = (size_t)c->n + 2 * halo_width, cnt = 6 * L * L;
    void *dev = nullptr;
    HIP_TRY(hipMalloc(&dev, cnt * c->esz()));
    int rc = swm_ecs_halo_xy(c, field, halo_width, dev);
    if (rc) { (void)hipFree(dev); return rc; }
    if (c->prec == 64) {
        HIP_TRY(hipMemcpyAsync(out, dev, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> buf(cnt);
        HIP_TRY(hipMemcpyAsync(buf.data(), dev, cnt * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < cnt; ++k) out[k] = buf[k];
    }
    HIP_TRY(hipFree(dev));
    return SWM_OK;
}
