// ECS cubic halo on xy (Ah) points and on O points (included into swm_capi.hip).
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

// ---- ECS cubic halo on O points (cell centres of the A grid) ----
// get_ecs_halo / ext_ecs_tile_halo (equiang_cs/ecs_halo_mod.f90:71-166) after the width-8 'full'
// O-point exchange (parallel/exchange_factory_mod.f90:13-28,45-155), which fills the strip beyond
// each panel edge with the edge neighbour's cells and leaves the halo corner squares alone.  Every
// value the procedure writes is a function of the panel and those strips only:
//  - halo row r of an edge at position t along it: 4-point Lagrange interpolation along the
//    neighbour's row at depth r-1 (ecs_ext_halo_1e, :451-475), positions 2..N-1 in row 1 and
//    1..N beyond;
//  - the special points at each cube corner (ecs_halo_1corner, :215-285, placed by
//    ecs_halo_corners_put, :287-334): row 1 at positions 1 and N, the corner-square diagonal and,
//    for hw >= 2, the rest of the 2 x 2 block next to the corner.
// One thread computes one output point from f directly.  Points of the corner squares the
// procedure does not write (only for hw >= 3) are set to NaN.
template <typename T>
struct EcsO {
    using value_type = T;
    const T *f;        // [6][N][N], i fastest
    const T *w;        // [rows][N+4][4]: positions -1..N+2, stencil offsets -1..2
    const int *ind;    // [rows][N+4]: 1-based stencil base
    const int *edge;   // [6][4][8]: neighbour panel, normal axis is x, far side, orientation, vector code
    int N, hw;

    __device__ T at(int p, int i, int j) const { return f[((long long)p * N + j - 1) * N + i - 1]; }
    // exchanged value beyond edge e (0 W, 1 E, 2 S, 3 N) at depth d (0 = first cell row of the
    // neighbour), position t = 1..N along the edge
    __device__ T raw(int p, int e, int d, int t) const {
        const int *ed = edge + (p * 4 + e) * 8;
        const int nidx = ed[2] == 0 ? d : N - 1 - d, tidx = ed[3] > 0 ? t - 1 : N - t;
        const int qi = ed[1] ? nidx : tidx, qj = ed[1] ? tidx : nidx;
        return f[((long long)ed[0] * N + qj) * N + qi];
    }
    // the exchanged array pf(I, J) where one of I, J lies in 1..N
    __device__ T pf(int p, int I, int J) const {
        if (I >= 1 && I <= N && J >= 1 && J <= N) return at(p, I, J);
        if (J < 1) return raw(p, 2, -J, I);
        if (J > N) return raw(p, 3, J - N - 1, I);
        if (I < 1) return raw(p, 0, -I, J);
        return raw(p, 1, I - N - 1, J);
    }
    __device__ T W(int s, int t, int r) const { return w[((r - 1) * (N + 4) + t + 1) * 4 + s + 1]; }
};

// Vector halo on O points (A grid, contravariant components; create_ecs_A_vec_halo_procedure,
// ecs_halo_vec_a_mod.f90:66-192): the exchanged neighbour components (u, v), swapped and negated
// by the panel pair's code (pack_to_buf_vec, parallel/buffer_mod.f90:171-277), times the 2 x 2
// matrix TM of the edge at rows 1..hw (positions ish..ieh) and at rows hw+1..5 near the corners
// (positions ish..2, N-1..ieh); then the scalar procedure on component `comp` (0 u, 1 v).
template <typename T>
struct EcsOVec {
    using value_type = T;
    const T *fu, *fv, *w;
    const int *ind, *edge;
    const T *tm;       // [edge 1..4][tmrows][N][4]
    int N, hw, tmrows, ish, ieh, comp;

    __device__ T at(int p, int i, int j) const {
        const long long o = ((long long)p * N + j - 1) * N + i - 1;
        return comp ? fv[o] : fu[o];
    }
    __device__ T raw(int p, int e, int d, int t) const {
        const int *ed = edge + (p * 4 + e) * 8;
        const int nidx = ed[2] == 0 ? d : N - 1 - d, tidx = ed[3] > 0 ? t - 1 : N - t;
        const int qi = ed[1] ? nidx : tidx, qj = ed[1] ? tidx : nidx;
        const long long o = ((long long)ed[0] * N + qj) * N + qi;
        T zu, zv;
        map_vec(ed[4], fu[o], fv[o], zu, zv);
        const int r = d + 1;
        const bool in = t >= ish && t <= ieh;
        if ((r <= hw && in) || (r > hw && r <= 5 && in && (t <= 2 || t >= N - 1))) {
            const int en = e == 2 ? 1 : e == 3 ? 2 : e == 0 ? 3 : 4;   // the reference's edge number
            const T *m = tm + (((long long)(en - 1) * tmrows + r - 1) * N + t - 1) * 4;
            return comp ? m[2] * zu + m[3] * zv : m[0] * zu + m[1] * zv;
        }
        return comp ? zv : zu;
    }
    __device__ T pf(int p, int I, int J) const {
        if (I >= 1 && I <= N && J >= 1 && J <= N) return at(p, I, J);
        if (J < 1) return raw(p, 2, -J, I);
        if (J > N) return raw(p, 3, J - N - 1, I);
        if (I < 1) return raw(p, 0, -I, J);
        return raw(p, 1, I - N - 1, J);
    }
    __device__ T W(int s, int t, int r) const { return w[((r - 1) * (N + 4) + t + 1) * 4 + s + 1]; }
};

// special value k (0..5) of corner c (0..3: (1,1), (N,1), (N,N), (1,N) sides), ecs_halo_1corner.
// Inlined: a call would make every thread of ecs_o_fill copy its argument struct to scratch
// memory to pass it by reference.
template <typename S>
__device__ __forceinline__ typename S::value_type ecs_o_special(const S &o, int p, int c, int k) {
    using T = typename S::value_type;
    const int N = o.N;
    const int ist = (c == 1 || c == 2) ? N + 1 : 0, idir = (c == 1 || c == 2) ? -1 : 1;
    const int jst = c >= 2 ? N + 1 : 0, jdir = c >= 2 ? -1 : 1;
    auto FC = [&](int a, int b) { return o.pf(p, ist + idir * a, jst + jdir * b); };
    // weights of corner 1; mirrored, w(2:-1:-1, N+2:N-1:-1, :), where the corner is on the right
    auto PW = [&](int s, int t, int r) { return idir == 1 ? o.W(s, t, r) : o.W(1 - s, N + 1 - t, r); };
    const T z0 = PW(0, 1, 1), z1 = PW(1, 1, 1), z2 = PW(2, 1, 1);
    const T zw = PW(-1, 1, 1), zw2 = zw * zw, zw3 = zw2 * zw;
    if (k == 0 || k == 2) {
        const T zx = z0 * FC(0, 1) + z1 * FC(0, 2) + z2 * FC(0, 3);
        const T zy = z0 * FC(1, 0) + z1 * FC(1, -1) + z2 * FC(1, -2);
        const T zz = z0 * FC(1, 1) + z1 * FC(2, 1) + z2 * FC(3, 1);
        if (k == 0) return (zx + zw * zy + zw2 * zz) / (T(1) - zw3);
        const T zfy1 = (zy + zw * zz + zw2 * zx) / (T(1) - zw3);
        const T xx = z0 * FC(1, 0) + z1 * FC(2, 0) + z2 * FC(3, 0);
        const T yy = z0 * FC(0, 1) + z1 * FC(-1, 1) + z2 * FC(-2, 1);
        const T zz2 = z0 * FC(1, 1) + z1 * FC(1, 2) + z2 * FC(1, 3);
        const T zfy2 = (yy + zw * zz2 + zw2 * xx) / (T(1) - zw3);
        const T zff1 = PW(-1, 1, 2) * FC(2, 0) + PW(0, 1, 2) * FC(2, -1) + PW(1, 1, 2) * FC(2, -2) + PW(2, 1, 2) * FC(2, -3);
        const T zff2 = PW(-1, 1, 2) * FC(0, 2) + PW(0, 1, 2) * FC(-1, 2) + PW(1, 1, 2) * FC(-2, 2) + PW(2, 1, 2) * FC(-3, 2);
        return T(0.5) * (PW(-1, 0, 1) * zff1 + PW(0, 0, 1) * zfy1 + PW(1, 0, 1) * FC(0, 1) + PW(2, 0, 1) * FC(0, 2) +
                         PW(-1, 0, 1) * zff2 + PW(0, 0, 1) * zfy2 + PW(1, 0, 1) * FC(1, 0) + PW(2, 0, 1) * FC(2, 0));
    }
    if (k == 1) {
        const T zx = z0 * FC(1, 0) + z1 * FC(2, 0) + z2 * FC(3, 0);
        const T zy = z0 * FC(0, 1) + z1 * FC(-1, 1) + z2 * FC(-2, 1);
        const T zz = z0 * FC(1, 1) + z1 * FC(1, 2) + z2 * FC(1, 3);
        return (zx + zw * zy + zw2 * zz) / (T(1) - zw3);
    }
    const T zff1 = PW(-1, 2, 1) * FC(1, 0) + PW(0, 2, 1) * FC(1, -1) + PW(1, 2, 1) * FC(1, -2) + PW(2, 2, 1) * FC(1, -3);
    const T zff2 = PW(-1, 2, 1) * FC(0, 1) + PW(0, 2, 1) * FC(-1, 1) + PW(1, 2, 1) * FC(-2, 1) + PW(2, 2, 1) * FC(-3, 1);
    if (k == 3) return PW(-1, 0, 2) * zff1 + PW(0, 0, 2) * FC(-1, 1) + PW(1, 0, 2) * FC(-1, 2) + PW(2, 0, 2) * FC(-1, 3);
    if (k == 4) return PW(-1, 0, 2) * zff2 + PW(0, 0, 2) * FC(1, -1) + PW(1, 0, 2) * FC(2, -1) + PW(2, 0, 2) * FC(3, -1);
    const T zff11 = PW(-1, 2, 2) * FC(2, -1) + PW(0, 2, 2) * FC(2, -2) + PW(1, 2, 2) * FC(2, -3) + PW(2, 2, 2) * FC(2, -4);
    const T zff22 = PW(-1, 2, 2) * FC(-1, 2) + PW(0, 2, 2) * FC(-2, 2) + PW(1, 2, 2) * FC(-3, 2) + PW(2, 2, 2) * FC(-4, 2);
    return T(0.5) * (PW(-1, -1, 2) * zff11 + PW(0, -1, 2) * zff1 + PW(1, -1, 2) * FC(-1, 1) + PW(2, -1, 2) * FC(-1, 2) +
                     PW(-1, -1, 2) * zff22 + PW(0, -1, 2) * zff2 + PW(1, -1, 2) * FC(1, -1) + PW(2, -1, 2) * FC(2, -1));
}

// out[(p*L + jj)*L + ii], L = N+2hw, for i, j = ii+1-hw, jj+1-hw
template <typename S>
__global__ void ecs_o_fill(S o, typename S::value_type *__restrict__ out) {
    using T = typename S::value_type;
    const int N = o.N, hw = o.hw, L = N + 2 * hw;
    // 32-bit index arithmetic: 6 L^2 < 2^32 for every panel size the library accepts
    const unsigned LL = (unsigned)L * (unsigned)L;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6u * LL) return;
    const int p = (int)(t / LL), r = (int)(t - (unsigned)p * LL);
    const int jj = r / L, ii = r - jj * L;
    const int i = ii + 1 - hw, j = jj + 1 - hw;
    const bool in_i = i >= 1 && i <= N, in_j = j >= 1 && j <= N;
    // corner and special index of a point next to a cube corner (-1 if not special)
    auto special = [&](int &c, int &k) {
        const bool left = i <= 1, bottom = j <= 1;
        c = bottom ? (left ? 0 : 1) : (left ? 3 : 2);
        const int ist = left ? 0 : N + 1, idir = left ? 1 : -1, jst = bottom ? 0 : N + 1, jdir = bottom ? 1 : -1;
        const int ci = idir * (i - ist), cj = jdir * (j - jst);
        k = -1;
        if (ci == 0 && cj == 1) k = 0;
        else if (ci == 1 && cj == 0) k = 1;
        else if (ci == 0 && cj == 0) k = 2;
        else if (hw > 1 && ci == -1 && cj == 0) k = 3;
        else if (hw > 1 && ci == 0 && cj == -1) k = 4;
        else if (hw > 1 && ci == -1 && cj == -1) k = 5;
    };
    T v;
    if (in_i && in_j) {
        v = o.at(p, i, j);
    } else if (in_i || in_j) {
        int e, r, pos;
        if (j < 1) { e = 2; r = 1 - j; pos = i; }
        else if (j > N) { e = 3; r = j - N; pos = i; }
        else if (i < 1) { e = 0; r = 1 - i; pos = j; }
        else { e = 1; r = i - N; pos = j; }
        if (r == 1 && (pos == 1 || pos == N)) {
            int c, k;
            special(c, k);
            v = ecs_o_special(o, p, c, k);
        } else {
            const int b = o.ind[(r - 1) * (N + 4) + pos + 1];
            T s = o.W(-1, pos, r) * o.raw(p, e, r - 1, b - 1);
            s = s + o.W(0, pos, r) * o.raw(p, e, r - 1, b);
            s = s + o.W(1, pos, r) * o.raw(p, e, r - 1, b + 1);
            s = s + o.W(2, pos, r) * o.raw(p, e, r - 1, b + 2);
            v = s;
        }
    } else {
        int c, k;
        special(c, k);
        v = k >= 0 ? ecs_o_special(o, p, c, k) : T(__builtin_nan(""));
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

namespace {

// weights and base indices for halo rows 1..rows, edge map; cached on the context
template <typename T>
int ecs_o_setup(swm_ctx *c, int rows) {
    const int N = c->N, TT = N + 4;
    if (c->ecs_o_rows >= rows) return SWM_OK;
    std::vector<double> w((size_t)4 * rows * TT);
    std::vector<int> ind((size_t)rows * TT);
    swmh_ecs_o_weights(N, rows, w.data(), ind.data());
    // the positions the procedure interpolates take their stencils inside the panel
    for (int r = 1; r <= rows; ++r)
        for (int t = r == 1 ? 2 : 1; t <= (r == 1 ? N - 1 : N); ++t) {
            const int b = ind[(size_t)(r - 1) * TT + t + 1];
            if (b - 1 < 1 || b + 2 > N) return fail(SWM_EARG, "ECS O-point halo stencil leaves the panel");
        }
    std::vector<T> wt(w.begin(), w.end());
    if (c->ecs_o_w) { HIP_TRY(hipFree(c->ecs_o_w)); c->ecs_o_w = nullptr; }
    if (c->ecs_o_ind) { HIP_TRY(hipFree(c->ecs_o_ind)); c->ecs_o_ind = nullptr; }
    HIP_TRY(hipMalloc(&c->ecs_o_w, wt.size() * sizeof(T)));
    HIP_TRY(hipMalloc(&c->ecs_o_ind, ind.size() * sizeof(int)));
    HIP_TRY(hipMemcpy(c->ecs_o_w, wt.data(), wt.size() * sizeof(T), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->ecs_o_ind, ind.data(), ind.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!c->ecs_o_edge) {
        // O-point edge map from the Ah edge descriptors: cell k between Ah points k and k+1
        // maps to cell k (orientation +1) or N-1-k (-1); depth d to d or N-1-d; the vector code
        // of the panel pair is the same as for the Ah points
        std::vector<int> ei((size_t)6 * 4 * 8), eo((size_t)6 * 4 * 8, 0);
        HIP_TRY(hipMemcpy(ei.data(), c->einfo, ei.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int k = 0; k < 24; ++k) {
            eo[8 * k + 0] = ei[8 * k + 0];
            eo[8 * k + 1] = ei[8 * k + 1];
            eo[8 * k + 2] = ei[8 * k + 2] == 0 ? 0 : 1;
            eo[8 * k + 3] = ei[8 * k + 4];
            eo[8 * k + 4] = ei[8 * k + 5];
        }
        HIP_TRY(hipMalloc(&c->ecs_o_edge, eo.size() * sizeof(int)));
        HIP_TRY(hipMemcpy(c->ecs_o_edge, eo.data(), eo.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipDeviceSynchronize());  // null-stream copies before the context stream reads them
    c->ecs_o_rows = rows;
    return SWM_OK;
}

template <typename T>
int ecs_o_run(swm_ctx *c, const void *field, int hw, void *out) {
    const int N = c->N;
    int rc = ecs_o_setup<T>(c, hw > 2 ? hw : 2);
    if (rc) return rc;
    const swm::EcsO<T> o{(const T *)field, (const T *)c->ecs_o_w, (const int *)c->ecs_o_ind,
                         (const int *)c->ecs_o_edge, N, hw};
    const long long total = 6LL * (N + 2 * hw) * (N + 2 * hw);
    hipLaunchKernelGGL(swm::ecs_o_fill<swm::EcsO<T>>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream,
                       o, (T *)out);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

template <typename T>
int ecs_vec_a_run(swm_ctx *c, const void *u, const void *v, int hw, void *out_u, void *out_v) {
    const int N = c->N, TT = N + 4, rows = hw > 2 ? hw : 2, tmrows = hw > 5 ? hw : 5;
    int rc = ecs_o_setup<T>(c, rows);
    if (rc) return rc;
    if (c->ecs_tm_rows < tmrows) {
        std::vector<double> tm((size_t)4 * tmrows * N * 4);
        swmh_ecs_vec_a_tm(N, tmrows, tm.data());
        std::vector<T> tt(tm.begin(), tm.end());
        if (c->ecs_tm) { HIP_TRY(hipFree(c->ecs_tm)); c->ecs_tm = nullptr; }
        HIP_TRY(hipMalloc(&c->ecs_tm, tt.size() * sizeof(T)));
        HIP_TRY(hipMemcpy(c->ecs_tm, tt.data(), tt.size() * sizeof(T), hipMemcpyHostToDevice));
        HIP_TRY(hipDeviceSynchronize());
        c->ecs_tm_rows = tmrows;
    }
    // ish..ieh of init_ecs_tile_halo_vect (one tile per panel), from the stencil bases
    std::vector<double> w((size_t)4 * rows * TT);
    std::vector<int> ind((size_t)rows * TT);
    swmh_ecs_o_weights(N, rows, w.data(), ind.data());
    const int s0 = std::max(1, 1 - hw), e0 = std::min(N, N + hw);
    int mn = 1 << 30, mx = -(1 << 30);
    for (int r = 1; r <= hw; ++r) {
        mn = std::min(mn, ind[(size_t)(r - 1) * TT + s0 + 1]);
        mx = std::max(mx, ind[(size_t)(r - 1) * TT + e0 + 1]);
    }
    const int ish = std::max(1, mn - 1), ieh = std::min(N, mx + 2);
    swm::EcsOVec<T> o{(const T *)u, (const T *)v, (const T *)c->ecs_o_w, (const int *)c->ecs_o_ind,
                      (const int *)c->ecs_o_edge, (const T *)c->ecs_tm, N, hw, c->ecs_tm_rows, ish, ieh, 0};
    const long long total = 6LL * (N + 2 * hw) * (N + 2 * hw);
    const dim3 grd((unsigned)((total + 255) / 256)), blk(256);
    hipLaunchKernelGGL(swm::ecs_o_fill<swm::EcsOVec<T>>, grd, blk, 0, c->stream, o, (T *)out_u);
    o.comp = 1;
    hipLaunchKernelGGL(swm::ecs_o_fill<swm::EcsOVec<T>>, grd, blk, 0, c->stream, o, (T *)out_v);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

int ecs_o_check(swm_ctx *c, int hw) {
    if (hw < 1 || hw > 8) return fail(SWM_EARG, "halo_width must be in [1, 8]");
    if (c->N < 8) return fail(SWM_EARG, "the ECS O-point halo exchanges 8 cell rows: N >= 8 required");
    return SWM_OK;
}

}  // namespace

extern "C" int swm_ecs_halo_o(swm_ctx *c, const void *field_o, int halo_width, void *out) {
    if (!c || !field_o || !out) return fail(SWM_EARG, "NULL argument");
    int rc = ecs_o_check(c, halo_width);
    if (rc) return rc;
    return c->prec == 64 ? ecs_o_run<double>(c, field_o, halo_width, out) : ecs_o_run<float>(c, field_o, halo_width, out);
}

extern "C" int swm_ecs_halo_o_host(swm_ctx *c, const double *field_o, int halo_width, double *out) {
    if (!c || !field_o || !out) return fail(SWM_EARG, "NULL argument");
    int rc = ecs_o_check(c, halo_width);
    if (rc) return rc;
    const size_t N = (size_t)c->N, L = N + 2 * halo_width, nin = 6 * N * N, nout = 6 * L * L, es = c->esz();
    void *din = nullptr, *dout = nullptr;
    HIP_TRY(hipMalloc(&din, nin * es));
    if (hipMalloc(&dout, nout * es) != hipSuccess) {
        (void)hipFree(din);
        return fail(SWM_EHIP, "hipMalloc failed");
    }
    auto done = [&](int r) {
        (void)hipFree(din);
        (void)hipFree(dout);
        return r;
    };
    if (c->prec == 64) {
        if (hipMemcpyAsync(din, field_o, nin * es, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return done(fail(SWM_EHIP, "upload failed"));
        if ((rc = swm_ecs_halo_o(c, din, halo_width, dout))) return done(rc);
        if (hipMemcpyAsync(out, dout, nout * es, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return done(fail(SWM_EHIP, "download failed"));
    } else {
        std::vector<float> bin(field_o, field_o + nin), bout(nout);
        if (hipMemcpyAsync(din, bin.data(), nin * es, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return done(fail(SWM_EHIP, "upload failed"));
        if ((rc = swm_ecs_halo_o(c, din, halo_width, dout))) return done(rc);
        if (hipMemcpyAsync(bout.data(), dout, nout * es, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return done(fail(SWM_EHIP, "download failed"));
        for (size_t k = 0; k < nout; ++k) out[k] = bout[k];
    }
    return done(SWM_OK);
}

// The device path of this entry point is switched off: its first GPU run (N = 8, halo width 2)
// ended in a memory aperture violation inside ecs_o_fill<EcsOVec<double>> whose cause has not been
// found (a CPU build of the same device code with AddressSanitizer over the same tables runs
// clean).  Until a cause is shown the entry point returns SWM_ESTATE without launching anything.
constexpr bool kEcsVecADeviceEnabled = false;

extern "C" int swm_ecs_halo_vec_a(swm_ctx *c, const void *u_o, const void *v_o, int halo_width, void *out_u,
                                  void *out_v) {
    if (!c || !u_o || !v_o || !out_u || !out_v) return fail(SWM_EARG, "NULL argument");
    int rc = ecs_o_check(c, halo_width);
    if (rc) return rc;
    if (!kEcsVecADeviceEnabled)
        return fail(SWM_ESTATE, "the ECS vector halo on O points is switched off on the device (unresolved GPU fault)");
    return c->prec == 64 ? ecs_vec_a_run<double>(c, u_o, v_o, halo_width, out_u, out_v)
                         : ecs_vec_a_run<float>(c, u_o, v_o, halo_width, out_u, out_v);
}

extern "C" int swm_ecs_halo_vec_a_host(swm_ctx *c, const double *u_o, const double *v_o, int halo_width,
                                       double *out_u, double *out_v) {
    if (!c || !u_o || !v_o || !out_u || !out_v) return fail(SWM_EARG, "NULL argument");
    int rc = ecs_o_check(c, halo_width);
    if (rc) return rc;
    const size_t N = (size_t)c->N, L = N + 2 * halo_width, nin = 6 * N * N, nout = 6 * L * L, es = c->esz();
    void *d = nullptr;
    HIP_TRY(hipMalloc(&d, (2 * nin + 2 * nout) * es));
    char *base = (char *)d;
    void *du = base, *dv = base + nin * es, *dou = base + 2 * nin * es, *dov = base + (2 * nin + nout) * es;
    auto done = [&](int r) {
        (void)hipFree(d);
        return r;
    };
    std::vector<float> bu, bv, bou, bov;
    const void *su = u_o, *sv = v_o;
    if (c->prec != 64) {
        bu.assign(u_o, u_o + nin);
        bv.assign(v_o, v_o + nin);
        su = bu.data();
        sv = bv.data();
    }
    if (hipMemcpyAsync(du, su, nin * es, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(dv, sv, nin * es, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return done(fail(SWM_EHIP, "upload failed"));
    if ((rc = swm_ecs_halo_vec_a(c, du, dv, halo_width, dou, dov))) return done(rc);
    if (c->prec == 64) {
        if (hipMemcpyAsync(out_u, dou, nout * es, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(out_v, dov, nout * es, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return done(fail(SWM_EHIP, "download failed"));
    } else {
        bou.resize(nout);
        bov.resize(nout);
        if (hipMemcpyAsync(bou.data(), dou, nout * es, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(bov.data(), dov, nout * es, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return done(fail(SWM_EHIP, "download failed"));
        for (size_t k = 0; k < nout; ++k) {
            out_u[k] = bou[k];
            out_v[k] = bov[k];
        }
    }
    return done(SWM_OK);
}
