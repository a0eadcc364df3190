// C grid (staggering 'C') vector-invariant shallow-water right-hand side and RK4, SBP42 family:
// divergence_c_sbp42, gradient_c_sbp42, curl_c_sbp42, coriolis_Cgrid_sbp42, KE_Cgrid_sbp42,
// co2contra_c_sbp42, massflux_c_sbp42 (the &shallow_water_model_c set, namelist_swm:80-102),
// composed as operator_swm_t%apply (operator_swm_mod.f90:61-113), FP64.
//
// Points per panel: h at cell centres o (N x N), u at x faces (N+1 x N), v at y faces
// (N x N+1), the curl at corners q (N+1 x N+1).  Every field is stored as 6 panels of
// rows x pitch doubles with a halo of H = 4 on each side; point (i, j) of any type (1-based)
// sits at ((p*rows + j-1+H)*pitch + i-1+H).
//
// The staggered operators are one-dimensional SBP stencils applied along x or y inside a
// panel (sbp_operator_mod.f90:152-228); the panels couple only through
//   - the tangential halo of u and v (curl penalties, curl_c_sbp_mod.f90:73-118),
//   - the face halo of the mass fluxes J*hu, J*hv (divergence penalties,
//     sbp_operator_mod.f90:344-413),
//   - the O-point halo of the kinetic-energy potential (gradient, grad_c_sbp42_mod.f90:53-139),
//   - the corner sync of the curl (halo_Ah_scalar_sync_mod.f90:52-93),
// the first three read on the fly through precomputed neighbour lists by the threads on the panel
// edges, the sync one small launch over the shared points.  One RHS is eight launches of memory-bound stencil sweeps with one thread per output point (16 x 16
// thread blocks, coalesced along i); the RK4 stage updates are applied by the kernels that produce
// the tendencies.  Intermediate values at another point type (the J Qi12-weighted components at o
// points, the curl at x and y points) are stored once and read by the next kernel: recomputing
// them inside the consumer (4-6 nested stencil rows per output) measured slower than the extra
// field traffic.  The RK4 step is captured into a HIP graph.
//
// Every expression keeps the reference's operation order with no fused multiply-adds
// (-ffp-contract=off), so results match an FP64 CPU evaluation of the same algorithm bit for bit.

namespace swmc {

constexpr int H = 4;
constexpr int BX = 16, BY = 16;

// ---- staggered SBP42 operators, sbp_operators_collection_mod.f90:110-191 ----
// W42 interpolation interfaces -> centres, optimised weights (W42_staggered_i2c_opt1)
struct W42i2c {
    static constexpr int NW = 6, NE = 4, INSHIFT = -1, DNX = -1;
    static constexpr double SIGN = 1.0;
    static constexpr int last[4] = {4, 4, 5, 6};
    static constexpr double Ql[4][6] = {
        {0.4892964885611128, 0.48123107353632727, 0.0696483872440123, -0.040175949341450676, 0.0, 0.0},
        {-0.07913099957454726, 0.6127279550862234, 0.5119370885511979, -0.045534044062872786, 0.0, 0.0},
        {-0.0869982983587545, 0.1691760860218048, 0.3626427230326682, 0.6151794893042863, -0.06, 0.0},
        {0.018680545032460714, -0.10869571330417979, 0.09884979151099299, 0.4911653767607321, 0.5625, -0.0625}};
    static constexpr double Qin[4] = {-1.0 / 16.0, 9.0 / 16.0, 9.0 / 16.0, -1.0 / 16.0};
};
// W42 interpolation centres -> interfaces (W42_staggered_c2i_opt1)
struct W42c2i {
    static constexpr int NW = 5, NE = 4, INSHIFT = -2, DNX = 1;
    static constexpr double SIGN = 1.0;
    static constexpr int last[4] = {4, 4, 4, 5};
    static constexpr double Ql[4][5] = {
        {1.3630402181345285, -0.17804474904273135, -0.23303115631809243, 0.048035687226327554, 0.0},
        {0.46340770044238916, 0.4765661872892848, 0.15664452409426372, -0.09661841182593758, 0.0},
        {0.07545241951434666, 0.44794495248229815, 0.37775283649236274, 0.09884979151099299, 0.0},
        {-0.04413695843145285, -0.04040344754874627, 0.649837488701711, 0.4980831989686297, -0.06338028169014084}};
    static constexpr double Qin[4] = {-1.0 / 16.0, 9.0 / 16.0, 9.0 / 16.0, -1.0 / 16.0};
};
// staggered first differences, centres -> interfaces (D42_staggered_c2i)
struct D42c2i {
    static constexpr int NW = 5, NE = 4, INSHIFT = -2, DNX = 1;
    static constexpr double SIGN = -1.0;
    static constexpr int last[4] = {3, 2, 4, 5};
    static constexpr double Ql[4][5] = {{-2.0, 3.0, -1.0, 0.0, 0.0},
                                        {-1.0, 1.0, 0.0, 0.0, 0.0},
                                        {1.0 / 24.0, -9.0 / 8.0, 9.0 / 8.0, -1.0 / 24.0, 0.0},
                                        {-1.0 / 71.0, 6.0 / 71.0, -83.0 / 71.0, 81.0 / 71.0, -3.0 / 71.0}};
    static constexpr double Qin[4] = {1.0 / 24.0, -9.0 / 8.0, 9.0 / 8.0, -1.0 / 24.0};
};
// staggered first differences, interfaces -> centres (D42_staggered_i2c)
struct D42i2c {
    static constexpr int NW = 5, NE = 3, INSHIFT = -1, DNX = -1;
    static constexpr double SIGN = -1.0;
    static constexpr int last[3] = {4, 4, 5};
    static constexpr double Ql[3][5] = {{-79.0 / 78.0, 27.0 / 26.0, -1.0 / 26.0, 1.0 / 78.0, 0.0},
                                        {2.0 / 21.0, -9.0 / 7.0, 9.0 / 7.0, -2.0 / 21.0, 0.0},
                                        {1.0 / 75.0, 0.0, -27.0 / 25.0, 83.0 / 75.0, -1.0 / 25.0}};
    static constexpr double Qin[4] = {1.0 / 24.0, -9.0 / 8.0, 9.0 / 8.0, -1.0 / 24.0};
};

// D42 boundary projection and norms (D42_boundary_proj, D42_A_interfaces, D42_A_centers)
constexpr double PROJ[3] = {15.0 / 8.0, -5.0 / 4.0, 3.0 / 8.0};
constexpr double A_IF1 = 7.0 / 18.0;
// add_penalty_center weights: proj_l(k)/Al_out(k) on the left, proj_r(k)/Ar_out(1+k) on the right
// (Ar_out = A_centers reversed = [1, 25/24, 7/8, 13/12], proj_r = proj_l reversed)
constexpr double PEN_L[3] = {(15.0 / 8.0) / (13.0 / 12.0), (-5.0 / 4.0) / (7.0 / 8.0), (3.0 / 8.0) / (25.0 / 24.0)};
constexpr double PEN_R[3] = {(3.0 / 8.0) / (25.0 / 24.0), (-5.0 / 4.0) / (7.0 / 8.0), (15.0 / 8.0) / (13.0 / 12.0)};

// one output row i (1-based) of sbp_apply along a line with values f(k), k = 1..nsrc: left rows
// sum(Ql(1:last,i) f(1:last)), inner rows sum(Qin f(i+inshift:...)), right rows the left block
// reflected (sign SIGN), each accumulated left to right from the first product
template <class O, class F>
__device__ __forceinline__ double sopf(int i, int nsrc, F &&f) {
    const int ntarg = nsrc + O::DNX;
    double r;
    if (i <= O::NE) {
        const int L = O::last[i - 1];
        r = O::Ql[i - 1][0] * f(1);
#pragma unroll
        for (int m = 2; m <= O::NW; ++m)
            if (m <= L) r = r + O::Ql[i - 1][m - 1] * f(m);
    } else if (i <= ntarg - O::NE) {
        const int b = i + O::INSHIFT;
        r = O::Qin[0] * f(b);
        r = r + O::Qin[1] * f(b + 1);
        r = r + O::Qin[2] * f(b + 2);
        r = r + O::Qin[3] * f(b + 3);
    } else {
        const int l = ntarg - i + 1;  // mirrored left row
        const int L = O::last[l - 1];
        const int b = nsrc + 1 - L;
        r = (O::SIGN * O::Ql[l - 1][L - 1]) * f(b);
#pragma unroll
        for (int t = 1; t < O::NW; ++t)
            if (t < L) r = r + (O::SIGN * O::Ql[l - 1][L - 1 - t]) * f(b + t);
    }
    return r;
}

// the same on a strided array line, f(k) = f0[k*s]
template <class O>
__device__ __forceinline__ double sop(int i, int nsrc, const double *__restrict__ f0, long long s) {
    return sopf<O>(i, nsrc, [&](int k) { return f0[(long long)k * s]; });
}

// metric terms J, Qi at the four point types depend on (alpha, beta) only: one panel plane each,
// read for every panel at the point's offset within its panel (mid); f differs per panel.
// Cube-edge neighbour lists (offsets in the stored layout, signs +-1), indexed by panel p, edge
// e (0: i = 1 side, 1: i = nx side, 2: j = 1 side, 3: j = ny side), depth d and position k:
//   vsrc/vsg  the tangential vector component (v beyond x edges, u beyond y edges), offsets from
//             the u field (+fs for v), [((p*4+e)*3 + d-1)*(N+1) + k-1]
//   fsrc/fsg  the neighbour's copy of a shared face (normal component), from the first field of
//             the pair, [(p*4+e)*N + k-1]
//   osrc      O points, [((p*4+e)*3 + d-1)*N + k-1]
//   qnb       q points on an edge: the neighbour's copy, [(p*4+e)*(N+1) + k-1]
struct CGeo {
    const double *Jo, *Jx, *Jy, *Jq, *Qi12o, *Qi11x, *Qi22y, *fo;
    const double *hs;  // orography at o points, or NULL
    const long long *vsrc, *fsrc, *osrc, *qnb;
    const signed char *vsg, *fsg;
    int N, pitch;
    int nbx, nby, ntiles, per;  // tiles per row, row strips per panel, tiles, tiles per XCD
    long long plane, fs;
    double hx, scale, grav, dx1;
};

__device__ __forceinline__ long long cid(const CGeo &g, int p, int i, int j) {
    return (long long)p * g.plane + (long long)(j - 1 + H) * g.pitch + (i - 1 + H);
}

// tangential vector component beyond edge e at depth d, position k (pair = u field)
__device__ __forceinline__ double vhalo(const CGeo &g, const double *pair, int p, int e, int d, int k) {
    const long long x = (((long long)p * 4 + e) * 3 + (d - 1)) * (g.N + 1) + (k - 1);
    const double val = pair[g.vsrc[x]];
    return g.vsg[x] < 0 ? -val : val;
}
// the neighbour's copy of the shared face beyond edge e at position k (pair = first field)
__device__ __forceinline__ double fhalo(const CGeo &g, const double *pair, int p, int e, int k) {
    const long long x = ((long long)p * 4 + e) * g.N + (k - 1);
    const double val = pair[g.fsrc[x]];
    return g.fsg[x] < 0 ? -val : val;
}
__device__ __forceinline__ double ohalo(const CGeo &g, const double *f, int p, int e, int d, int k) {
    return f[g.osrc[(((long long)p * 4 + e) * 3 + (d - 1)) * g.N + (k - 1)]];
}

// RK4 stage handling of the kernels that produce the tendencies (rk4_mod.f90:20-42):
// RK_RHS writes the tendency k; RK_STAGE writes k and the next stage input 1.0*v0 + b*k
// (stvec_swm_t%assign); RK_FINAL writes v0 = (v0 + c2 k0 + c3 k1) + c3 k2 + c2 k (k = the last
// stage's tendency, not stored)
enum { RK_RHS = 0, RK_STAGE = 1, RK_FINAL = 2 };
struct RkArgs {
    double *k;          // tendency out (RK_RHS, RK_STAGE)
    const double *v0;   // stage-1 state (RK_STAGE, RK_FINAL)
    double *yn;         // next stage input (RK_STAGE)
    double *v0w;        // state written (RK_FINAL; the same buffer as v0)
    const double *k0, *k1, *k2;
    double b, c2, c3;
};

template <int RKM>
__device__ __forceinline__ void rk_out(const RkArgs &r, long long off, double k) {
    if (RKM == RK_RHS) {
        r.k[off] = k;
    } else if (RKM == RK_STAGE) {
        r.k[off] = k;
        r.yn[off] = 1.0 * r.v0[off] + r.b * k;
    } else {
        double v = r.v0[off] + r.c2 * r.k0[off] + r.c3 * r.k1[off];
        v = v + r.c3 * r.k2[off] + r.c2 * k;
        r.v0w[off] = v;
    }
}

// Block -> tile mapping: tiles of BX x BY points in the order (column strip fastest, row strip,
// panel); the hardware hands consecutive block indices to the 8 XCDs in turn, so block b takes
// tile (b % 8) * per + b / 8 and every XCD walks a contiguous band of row strips, keeping the
// rows its stencils share in its own L2
#define C_POINT                                                        \
    const int tile_ = (int)(blockIdx.x & 7) * g.per + (int)(blockIdx.x >> 3); \
    if (tile_ >= g.ntiles) return;                                     \
    const int tx_ = tile_ % g.nbx, ty_ = (tile_ / g.nbx) % g.nby;       \
    const int i = tx_ * BX + threadIdx.x + 1;                          \
    const int j = ty_ * BY + threadIdx.y + 1;                          \
    const int p = tile_ / (g.nbx * g.nby);

// co2contra_c_sbp42 first half (co2contra_Cgrid_mod.f90:361-378): at o points
// a = J Qi12 * (v interpolated along y), b = J Qi12 * (u interpolated along x)
__global__ void __launch_bounds__(256) k_contra_o(CGeo g, const double *__restrict__ u, const double *__restrict__ v,
                                                  double *__restrict__ A, double *__restrict__ B) {
    C_POINT
    if (i > g.N || j > g.N) return;
    const long long id = cid(g, p, i, j), mid = id - (long long)p * g.plane;
    const double dy = sop<W42i2c>(j, g.N + 1, v + cid(g, p, i, 0), g.pitch);
    const double dx = sop<W42i2c>(i, g.N + 1, u + cid(g, p, 0, j), 1);
    A[id] = g.Jo[mid] * g.Qi12o[mid] * dy;
    B[id] = g.Jo[mid] * g.Qi12o[mid] * dx;
}

// at x and y points: second half of co2contra (:370-398), massflux_c_sbp42
// (massflux_Cgrid_mod.f90:330-355), the J-weighted fluxes of the divergence and Coriolis terms
// (div_c_sbp42_mod.f90:101-130, coriolis_Cgrid_mod.f90:69-84) and the halves of KE_Cgrid
// (KE_Cgrid_mod.f90:41-57): G = (J hu, J hv), KEu = J u ut / 2, KEv = J v vt / 2
__global__ void __launch_bounds__(256) k_flux(CGeo g, const double *__restrict__ h, const double *__restrict__ u,
                                              const double *__restrict__ v, const double *__restrict__ A,
                                              const double *__restrict__ B, double *__restrict__ G,
                                              double *__restrict__ KEu, double *__restrict__ KEv) {
    C_POINT
    const int N = g.N;
    if (i <= N + 1 && j <= N) {
        const long long id = cid(g, p, i, j), mid = id - (long long)p * g.plane;
        const double t = sop<W42c2i>(i, N, A + cid(g, p, 0, j), 1);
        const double ut = g.Qi11x[mid] * u[id] + t / g.Jx[mid];
        const double hh = sop<W42c2i>(i, N, h + cid(g, p, 0, j), 1);
        const double hu = ut * hh;
        G[id] = g.Jx[mid] * hu;
        KEu[id] = g.Jx[mid] * (u[id] * ut) / 2.0;
    }
    if (i <= N && j <= N + 1) {
        const long long id = cid(g, p, i, j), mid = id - (long long)p * g.plane;
        const double t = sop<W42c2i>(j, N, B + cid(g, p, i, 0), g.pitch);
        const double vt = g.Qi22y[mid] * v[id] + t / g.Jy[mid];
        const double hh = sop<W42c2i>(j, N, h + cid(g, p, i, 0), g.pitch);
        const double hv = vt * hh;
        G[g.fs + id] = g.Jy[mid] * hv;
        KEv[id] = g.Jy[mid] * (v[id] * vt) / 2.0;
    }
}

// at o points: KE = (I_x KEu + I_y KEv)/J + g h + g hs (KE_Cgrid_mod.f90:59-75 and the update of
// operator_swm_mod.f90:85), and dh = -div(J hu, J hv) with the centre penalties
// (div_c_sbp42_mod.f90:70-99, sbp_operator_mod.f90:344-413), the neighbour's copies of the shared
// faces read through fsrc
template <int RKM>
__global__ void __launch_bounds__(256) k_ke_div(CGeo g, const double *__restrict__ h, const double *__restrict__ KEu,
                                                const double *__restrict__ KEv, const double *__restrict__ G,
                                                double *__restrict__ KE, RkArgs r) {
    C_POINT
    const int N = g.N;
    if (i > N || j > N) return;
    const long long id = cid(g, p, i, j), mid = id - (long long)p * g.plane;
    const double kuh = sop<W42i2c>(i, N + 1, KEu + cid(g, p, 0, j), 1);
    const double kvh = sop<W42i2c>(j, N + 1, KEv + cid(g, p, i, 0), g.pitch);
    double ke = (kuh + kvh) / g.Jo[mid];
    ke = ke + g.grav * h[id] + g.grav * (g.hs ? g.hs[id] : 0.0);
    KE[id] = ke;
    const double *Gu = G, *Gv = G + g.fs;
    double dx = sop<D42i2c>(i, N + 1, Gu + cid(g, p, 0, j), 1);
    if (i <= 3) {
        const double df = Gu[cid(g, p, 1, j)] - fhalo(g, G, p, 0, j);
        dx = dx + 0.5 * PEN_L[i - 1] * df;
    }
    if (i >= N - 2) {
        const double df = fhalo(g, G, p, 1, j) - Gu[cid(g, p, N + 1, j)];
        dx = dx + 0.5 * PEN_R[i - N + 2] * df;
    }
    double dy = sop<D42i2c>(j, N + 1, Gv + cid(g, p, i, 0), g.pitch);
    if (j <= 3) {
        const double df = Gv[cid(g, p, i, 1)] - fhalo(g, G, p, 2, i);
        dy = dy + 0.5 * PEN_L[j - 1] * df;
    }
    if (j >= N - 2) {
        const double df = fhalo(g, G, p, 3, i) - Gv[cid(g, p, i, N + 1)];
        dy = dy + 0.5 * PEN_R[j - N + 2] * df;
    }
    const double div = (dx + dy) / (g.Jo[mid] * g.hx * g.scale);
    rk_out<RKM>(r, id, -1.0 * div);
}

// curl_c_sbp42 at q points before the corner sync (curl_c_sbp_mod.f90:73-118): staggered
// differences of v along x and u along y with the interface penalties, the neighbour's
// tangential components read through vsrc
__global__ void __launch_bounds__(256) k_curl(CGeo g, const double *__restrict__ u, double *__restrict__ W) {
    C_POINT
    const int N = g.N, n = N + 1;
    if (i > n || j > n) return;
    const double *v = u + g.fs;
    const long long id = cid(g, p, i, j), mid = id - (long long)p * g.plane;
    double dvx = sop<D42c2i>(i, N, v + cid(g, p, 0, j), 1);
    if (i == 1) {
        double dv = 0.0;
        for (int k = 1; k <= 3; ++k) dv = dv + PROJ[k - 1] * (v[cid(g, p, k, j)] - vhalo(g, u, p, 0, k, j));
        dvx = dvx + 0.5 * dv / A_IF1;
    }
    if (i == n) {
        double dv = 0.0;
        for (int t = 0; t <= 2; ++t) dv = dv + PROJ[t] * (vhalo(g, u, p, 1, t + 1, j) - v[cid(g, p, n - t - 1, j)]);
        dvx = dvx + 0.5 * dv / A_IF1;
    }
    double duy = sop<D42c2i>(j, N, u + cid(g, p, i, 0), g.pitch);
    if (j == 1) {
        double du = 0.0;
        for (int k = 1; k <= 3; ++k) du = du + PROJ[k - 1] * (u[cid(g, p, i, k)] - vhalo(g, u, p, 2, k, i));
        duy = duy + 0.5 * du / A_IF1;
    }
    if (j == n) {
        double du = 0.0;
        for (int t = 0; t <= 2; ++t) du = du + PROJ[t] * (vhalo(g, u, p, 3, t + 1, i) - u[cid(g, p, i, n - t - 1)]);
        duy = duy + 0.5 * du / A_IF1;
    }
    W[id] = (dvx - duy) / (g.Jq[mid] * g.hx * g.scale);
}

// halo_Ah_scalar_sync of the curl at q points (halo_Ah_scalar_sync_mod.f90:52-93), in place on
// the panel-edge points only.  Every shared point is one group handled by one thread: a point on
// a cube edge has two copies whose mean 0.5 (a + b) is the same value on both panels; a cube
// corner has three copies, copy c = (own + first neighbour + second neighbour) / 3 in the
// order of its own panel's formula.  pairs[2k], pairs[2k+1]: offsets of the two copies;
// corners[6g..6g+2]: offsets of the three copies, corners[6g+3+c]: the neighbour order of copy
// c packed as 3 * first + second.
__global__ void __launch_bounds__(256) k_qsync_edges(int npairs, int ncorners, const long long *__restrict__ pairs,
                                                     const long long *__restrict__ corners, double *W) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < npairs) {
        const long long a = pairs[2 * k], b = pairs[2 * k + 1];
        const double v = 0.5 * (W[a] + W[b]);
        W[a] = v;
        W[b] = v;
    } else if (k < npairs + ncorners) {
        const long long *cg = corners + 6 * (k - npairs);
        const double r[3] = {W[cg[0]], W[cg[1]], W[cg[2]]};
        double v[3];
        for (int c = 0; c < 3; ++c) {
            const int ord = (int)cg[3 + c];
            v[c] = (r[c] + r[ord / 3] + r[ord % 3]) / 3.0;
        }
        for (int c = 0; c < 3; ++c) W[cg[c]] = v[c];
    }
}

// interpolator_w2h first half (interpolator_w2h_mod.f90:50-68): the synced curl at x points
// (interpolated along y) and at y points (along x)
__global__ void __launch_bounds__(256) k_w2uv(CGeo g, const double *__restrict__ W, double *__restrict__ fu,
                                              double *__restrict__ fv) {
    C_POINT
    const int N = g.N;
    if (i <= N + 1 && j <= N) fu[cid(g, p, i, j)] = sop<W42i2c>(j, N + 1, W + cid(g, p, i, 0), g.pitch);
    if (i <= N && j <= N + 1) fv[cid(g, p, i, j)] = sop<W42i2c>(i, N + 1, W + cid(g, p, 0, j), 1);
}

// at o points: curl_p = (I_x fu + I_y fv)/2 (interpolator_w2h_mod.f90:40-46), the fluxes J hu, J hv
// interpolated to o points, and the vector-invariant Coriolis terms (coriolis_Cgrid_mod.f90:86-103)
__global__ void __launch_bounds__(256) k_cor_o(CGeo g, const double *__restrict__ h, const double *__restrict__ fu,
                                               const double *__restrict__ fv, const double *__restrict__ G,
                                               double *__restrict__ cup, double *__restrict__ cvp) {
    C_POINT
    const int N = g.N;
    if (i > N || j > N) return;
    const long long id = cid(g, p, i, j);
    const double fuh = sop<W42i2c>(i, N + 1, fu + cid(g, p, 0, j), 1);
    const double fvh = sop<W42i2c>(j, N + 1, fv + cid(g, p, i, 0), g.pitch);
    const double ghup = sop<W42i2c>(i, N + 1, G + cid(g, p, 0, j), 1);
    const double ghvp = sop<W42i2c>(j, N + 1, G + g.fs + cid(g, p, i, 0), g.pitch);
    const double cp = 0.5 * fuh + 0.5 * fvh;
    cup[id] = (g.fo[id] + cp) * ghvp / h[id];
    cvp[id] = -(g.fo[id] + cp) * ghup / h[id];
}

// gradient_c_sbp42 along one line (grad_c_sbp42_mod.f90:61-139), F(k) for k = -2..N+3
template <class F>
__device__ __forceinline__ double grad_line(int i, int N, F &&Fv, double dx1) {
    constexpr double dedge[4][5] = {{-2.0, 3.0, -1.0, 0.0, 0.0},
                                    {-1.0, 1.0, 0.0, 0.0, 0.0},
                                    {1.0 / 24.0, -9.0 / 8.0, 9.0 / 8.0, -1.0 / 24.0, 0.0},
                                    {-1.0 / 71.0, 6.0 / 71.0, -83.0 / 71.0, 81.0 / 71.0, -3.0 / 71.0}};
    constexpr int dlast[4] = {3, 2, 4, 5};
    constexpr double din[4] = {1.0 / 24.0, -9.0 / 8.0, 9.0 / 8.0, -1.0 / 24.0};
    const int n = N + 1;
    double r;
    if (i <= 4) {
        r = dedge[i - 1][0] * Fv(1);
#pragma unroll
        for (int m = 2; m <= 5; ++m)
            if (m <= dlast[i - 1]) r = r + dedge[i - 1][m - 1] * Fv(m);
        r = r * dx1;
        if (i == 1) {
            const double dh = (PROJ[0] * (Fv(1) - Fv(0)) + PROJ[1] * (Fv(2) - Fv(-1)) + PROJ[2] * (Fv(3) - Fv(-2))) * dx1;
            r = r + 0.5 * dh / A_IF1;
        }
    } else if (i <= N + 1 - 4) {
        r = Fv(i - 2) * din[0];
        r = r + Fv(i - 1) * din[1];
        r = r + Fv(i) * din[2];
        r = r + Fv(i + 1) * din[3];
        r = r * dx1;
    } else {
        const int i2 = n - i + 1, i1 = dlast[i2 - 1];
        double t = dedge[i2 - 1][0] * Fv(n - 1);
#pragma unroll
        for (int m = 2; m <= 5; ++m)
            if (m <= i1) t = t + dedge[i2 - 1][m - 1] * Fv(n - m);
        r = -t * dx1;
        if (i == n) {
            const double dh =
                (PROJ[0] * (Fv(n) - Fv(n - 1)) + PROJ[1] * (Fv(n + 1) - Fv(n - 2)) + PROJ[2] * (Fv(n + 2) - Fv(n - 3))) * dx1;
            r = r + 0.5 * dh / A_IF1;
        }
    }
    return r;
}

// momentum tendencies at x and y points: du = -grad_x(KE) + cor_u, dv = -grad_y(KE) + cor_v with
// cor interpolated from o points (interp2d_pvec2uv_C_sbp42, interpolator_p2uv_mod.f90:90-113;
// operator_swm_mod.f90:94-95); KE beyond the panel edges read through osrc
template <int RKM>
__global__ void __launch_bounds__(256) k_mom(CGeo g, const double *__restrict__ KE, const double *__restrict__ cup,
                                             const double *__restrict__ cvp, RkArgs r) {
    C_POINT
    const int N = g.N;
    if (i <= N + 1 && j <= N) {
        const long long id = cid(g, p, i, j);
        const double *row = KE + cid(g, p, 0, j);
        const double gx = grad_line(i, N, [&](int k) {
            return k < 1 ? ohalo(g, KE, p, 0, 1 - k, j) : k > N ? ohalo(g, KE, p, 1, k - N, j) : row[k];
        }, g.dx1);
        const double cu = sop<W42c2i>(i, N, cup + cid(g, p, 0, j), 1);
        rk_out<RKM>(r, g.fs + id, -1.0 * gx + 1.0 * cu);
    }
    if (i <= N && j <= N + 1) {
        const long long id = cid(g, p, i, j);
        const double *col = KE + cid(g, p, i, 0);
        const double gy = grad_line(j, N, [&](int k) {
            return k < 1 ? ohalo(g, KE, p, 2, 1 - k, i) : k > N ? ohalo(g, KE, p, 3, k - N, i)
                                                               : col[(long long)k * g.pitch];
        }, g.dx1);
        const double cv = sop<W42c2i>(j, N, cvp + cid(g, p, i, 0), g.pitch);
        rk_out<RKM>(r, 2 * g.fs + id, -1.0 * gy + 1.0 * cv);
    }
}

// ---- get_diagnostics on the C grid (operator_swm_mod.f90:115-155, calc_energy :211-236) ----
// Per-point terms f*J*hx*hy*Ax(i)*Ay(j) of SBP_C42_quadrature (sbp_quadrature_mod.f90:21-40;
// Ax, Ay = A_centers along centre directions, A_interfaces along face directions) summed with
// compensated (Neumaier) additions per block: mass, KE_u, KE_v, PE, then hmin, hmax.
constexpr int CD_NQ = 4;

__device__ __forceinline__ void cd_add(double &s, double &c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

__global__ void __launch_bounds__(256) k_cdiag_partial(CGeo g, const double *__restrict__ y,
                                                       const double *__restrict__ Acen,
                                                       const double *__restrict__ Aint, double *__restrict__ part) {
    C_POINT
    const int N = g.N;
    const double *h = y, *u = y + g.fs, *v = y + 2 * g.fs;
    double s[CD_NQ] = {0.0, 0.0, 0.0, 0.0}, cc[CD_NQ] = {0.0, 0.0, 0.0, 0.0};
    double hmin = 1e300, hmax = -1e300;
    const double hx = g.hx;
    if (i <= N && j <= N) {
        const long long id = cid(g, p, i, j), mid = id - (long long)p * g.plane;
        const double hh = h[id];
        hmin = hh;
        hmax = hh;
        cd_add(s[0], cc[0], hh * g.Jo[mid] * hx * hx * Acen[i] * Acen[j]);
        const double pe = 1.0 * hh + 1.0 * (g.hs ? g.hs[id] : 0.0);
        const double PE = 0.5 * g.grav * pe * pe;
        cd_add(s[3], cc[3], PE * g.Jo[mid] * hx * hx * Acen[i] * Acen[j]);
    }
    if (i <= N + 1 && j <= N) {  // KE_diag_u = 0.5 * hu * u at x points, hu = massflux(h, ut)
        const long long id = cid(g, p, i, j), mid = id - (long long)p * g.plane;
        const double t = sopf<W42c2i>(i, N, [&](int ip) {
            const long long m = cid(g, 0, ip, j);
            const double dy = sop<W42i2c>(j, N + 1, v + cid(g, p, ip, 0), g.pitch);
            return g.Jo[m] * g.Qi12o[m] * dy;
        });
        const double ut = g.Qi11x[mid] * u[id] + t / g.Jx[mid];
        const double hu = ut * sop<W42c2i>(i, N, h + cid(g, p, 0, j), 1);
        const double ke = 0.5 * hu * u[id];
        cd_add(s[1], cc[1], ke * g.Jx[mid] * hx * hx * Aint[i] * Acen[j]);
    }
    if (i <= N && j <= N + 1) {
        const long long id = cid(g, p, i, j), mid = id - (long long)p * g.plane;
        const double t = sopf<W42c2i>(j, N, [&](int jp) {
            const long long m = cid(g, 0, i, jp);
            const double dx = sop<W42i2c>(i, N + 1, u + cid(g, p, 0, jp), 1);
            return g.Jo[m] * g.Qi12o[m] * dx;
        });
        const double vt = g.Qi22y[mid] * v[id] + t / g.Jy[mid];
        const double hv = vt * sop<W42c2i>(j, N, h + cid(g, p, i, 0), g.pitch);
        const double ke = 0.5 * hv * v[id];
        cd_add(s[2], cc[2], ke * g.Jy[mid] * hx * hx * Acen[i] * Aint[j]);
    }
    __shared__ double sh[256][2 * CD_NQ + 2];
    const int t = threadIdx.y * BX + threadIdx.x;
    for (int q = 0; q < CD_NQ; ++q) {
        sh[t][2 * q] = s[q];
        sh[t][2 * q + 1] = cc[q];
    }
    sh[t][2 * CD_NQ] = hmin;
    sh[t][2 * CD_NQ + 1] = hmax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            for (int q = 0; q < CD_NQ; ++q) {
                double a = sh[t][2 * q], ac = sh[t][2 * q + 1];
                cd_add(a, ac, sh[t + w][2 * q]);
                ac += sh[t + w][2 * q + 1];
                sh[t][2 * q] = a;
                sh[t][2 * q + 1] = ac;
            }
            sh[t][2 * CD_NQ] = fmin(sh[t][2 * CD_NQ], sh[t + w][2 * CD_NQ]);
            sh[t][2 * CD_NQ + 1] = fmax(sh[t][2 * CD_NQ + 1], sh[t + w][2 * CD_NQ + 1]);
        }
        __syncthreads();
    }
    if (t < 2 * CD_NQ + 2) part[(long long)tile_ * 10 + t] = sh[0][t];
}

// one thread over the block partials: {mass, KE_u, KE_v, PE, hmin, hmax}
__global__ void __launch_bounds__(64) k_cdiag_final(const double *__restrict__ part, long long nblocks,
                                                    double *__restrict__ out) {
    if (threadIdx.x != 0) return;
    double s[CD_NQ] = {0.0, 0.0, 0.0, 0.0}, cc[CD_NQ] = {0.0, 0.0, 0.0, 0.0};
    double hmin = 1e300, hmax = -1e300;
    for (long long b = 0; b < nblocks; ++b) {
        for (int q = 0; q < CD_NQ; ++q) {
            cd_add(s[q], cc[q], part[b * 10 + 2 * q]);
            cc[q] += part[b * 10 + 2 * q + 1];
        }
        hmin = fmin(hmin, part[b * 10 + 2 * CD_NQ]);
        hmax = fmax(hmax, part[b * 10 + 2 * CD_NQ + 1]);
    }
    for (int q = 0; q < CD_NQ; ++q) out[q] = s[q] + cc[q];
    out[CD_NQ] = hmin;
    out[CD_NQ + 1] = hmax;
}

#undef C_POINT

}  // namespace swmc

// ------------------------------------------------------------------------------------------------
struct swm_cctx {
    int N = 0, rows = 0, pitch = 0;
    long long plane = 0, fs = 0;
    double scale = 1, omega = 1, grav = 1, hx = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double *geo = nullptr;  // one panel plane each of Jo, Jx, Jy, Jq, Qi12o, Qi11x, Qi22y, then fo (6 panels)
    double *hs = nullptr;
    // cube-edge neighbour lists (swmc::CGeo): vsrc, fsrc, osrc, qnb; signs vsg, fsg
    long long *lists = nullptr;
    signed char *signs = nullptr;
    long long *vsrc = nullptr, *fsrc = nullptr, *osrc = nullptr, *qnb = nullptr;
    signed char *vsg = nullptr, *fsg = nullptr;
    // groups of the q-point edge sync (k_qsync_edges)
    long long *qpairs = nullptr, *qcorners = nullptr;
    int nqpairs = 0, nqcorners = 0;
    // work fields: G (pair), KEu, KEv, KE, W, cup, cvp, RK tendencies k0..k2, stage inputs yA, yB,
    // A, B (J Qi12-weighted components at o points), fu, fv (the synced curl at x and y points)
    double *work = nullptr;
    // diagnostics: quadrature weights A_centers (1..N) and A_interfaces (1..N+1) per index, block
    // partial sums (10 per block) and the 6 results
    double *qw = nullptr, *dpart = nullptr;
    long long dblocks = 0;
    hipGraphExec_t graph = nullptr, graph_multi = nullptr;
    void *graph_state = nullptr;
    double graph_dt = 0.0;
    // ECS halos of the C grid (the vector halo and the ECS_O halo of gradient_c2_ecs and
    // massflux_c_up4), set up on first use: strip sources (normal, tangential, cells; element
    // offsets) and signs, the gathered strips, ECS_O weights, the vector tables of one (width,
    // kind) and the extended o-point array
    long long *ecs_src = nullptr;
    signed char *ecs_sign = nullptr;
    double *ecs_strips = nullptr;
    double *ecs_ow = nullptr;
    int *ecs_oind = nullptr;
    int ecs_orows = 0;
    double *ecs_vt = nullptr;
    int *ecs_vti = nullptr;
    int ecs_vt_hw = 0, ecs_vt_cov = -1;
    double *ecs_ext = nullptr;

    size_t field_bytes() const { return (size_t)fs * sizeof(double); }
    double *wf(int k) const { return work + (size_t)k * fs; }
};

namespace {

constexpr int CW_G = 0 /* and 1 */, CW_KEU = 2, CW_KEV = 3, CW_KE = 4, CW_W = 5, CW_CUP = 6, CW_CVP = 7,
              CW_K = 8 /* 3 stages x 3 fields */, CW_YA = 17, CW_YB = 20, CW_A = 23, CW_B = 24, CW_FU = 25,
              CW_FV = 26, CW_COUNT = 27;

swmc::CGeo make_cgeo(const swm_cctx *c) {
    swmc::CGeo g;
    const double *G = c->geo;
    const size_t P = (size_t)c->plane;
    g.Jo = G;
    g.Jx = G + P;
    g.Jy = G + 2 * P;
    g.Jq = G + 3 * P;
    g.Qi12o = G + 4 * P;
    g.Qi11x = G + 5 * P;
    g.Qi22y = G + 6 * P;
    g.fo = G + 7 * P;
    g.hs = c->hs;
    g.vsrc = c->vsrc;
    g.fsrc = c->fsrc;
    g.osrc = c->osrc;
    g.qnb = c->qnb;
    g.vsg = c->vsg;
    g.fsg = c->fsg;
    g.N = c->N;
    g.pitch = c->pitch;
    g.nbx = (c->N + 1 + swmc::BX - 1) / swmc::BX;
    g.nby = (c->N + 1 + swmc::BY - 1) / swmc::BY;
    g.ntiles = g.nbx * g.nby * 6;
    g.per = (g.ntiles + 7) / 8;
    g.plane = c->plane;
    g.fs = c->fs;
    g.hx = c->hx;
    g.scale = c->scale;
    g.grav = c->grav;
    g.dx1 = 1.0 / (c->scale * c->hx);
    return g;
}

// operator_swm_t%apply on the C grid for the input y (h, u, v), eight launches; the tendencies go
// where RKM says (RK_RHS: r.k; RK_STAGE: r.k and the next stage input; RK_FINAL: the state)
template <int RKM>
int c_enqueue_rhs(swm_cctx *c, const double *y, const swmc::RkArgs &r) {
    const swmc::CGeo g = make_cgeo(c);
    const dim3 blk(swmc::BX, swmc::BY), grd(8 * g.per);
    const size_t F = (size_t)c->fs;
    const double *h = y, *u = y + F, *v = y + 2 * F;
    hipLaunchKernelGGL(swmc::k_contra_o, grd, blk, 0, c->stream, g, u, v, c->wf(CW_A), c->wf(CW_B));
    hipLaunchKernelGGL(swmc::k_flux, grd, blk, 0, c->stream, g, h, u, v, c->wf(CW_A), c->wf(CW_B), c->wf(CW_G),
                       c->wf(CW_KEU), c->wf(CW_KEV));
    hipLaunchKernelGGL(swmc::k_ke_div<RKM>, grd, blk, 0, c->stream, g, h, c->wf(CW_KEU), c->wf(CW_KEV), c->wf(CW_G),
                       c->wf(CW_KE), r);
    hipLaunchKernelGGL(swmc::k_curl, grd, blk, 0, c->stream, g, u, c->wf(CW_W));
    const int nsync = c->nqpairs + c->nqcorners;
    hipLaunchKernelGGL(swmc::k_qsync_edges, dim3((nsync + 255) / 256), dim3(256), 0, c->stream, c->nqpairs,
                       c->nqcorners, c->qpairs, c->qcorners, c->wf(CW_W));
    hipLaunchKernelGGL(swmc::k_w2uv, grd, blk, 0, c->stream, g, c->wf(CW_W), c->wf(CW_FU), c->wf(CW_FV));
    hipLaunchKernelGGL(swmc::k_cor_o, grd, blk, 0, c->stream, g, h, c->wf(CW_FU), c->wf(CW_FV), c->wf(CW_G),
                       c->wf(CW_CUP), c->wf(CW_CVP));
    hipLaunchKernelGGL(swmc::k_mom<RKM>, grd, blk, 0, c->stream, g, c->wf(CW_KE), c->wf(CW_CUP), c->wf(CW_CVP), r);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

// rk4_t%step (rk4_mod.f90:20-42): stage inputs state -> yA -> yB -> yA, the stage updates and the
// final combination inside the tendency kernels
int c_enqueue_rk4_step(swm_cctx *c, double *state, double dt) {
    double *k[3] = {c->wf(CW_K), c->wf(CW_K + 3), c->wf(CW_K + 6)};
    const double *in[4] = {state, c->wf(CW_YA), c->wf(CW_YB), c->wf(CW_YA)};
    double *next[3] = {c->wf(CW_YA), c->wf(CW_YB), c->wf(CW_YA)};
    const double b[3] = {0.5 * dt, 0.5 * dt, dt};
    int rc;
    for (int s = 0; s < 3; ++s) {
        swmc::RkArgs r{};
        r.k = k[s];
        r.v0 = state;
        r.yn = next[s];
        r.b = b[s];
        if ((rc = c_enqueue_rhs<swmc::RK_STAGE>(c, in[s], r))) return rc;
    }
    swmc::RkArgs r{};
    r.v0 = state;
    r.v0w = state;
    r.k0 = k[0];
    r.k1 = k[1];
    r.k2 = k[2];
    r.c2 = dt / 6.0;
    r.c3 = dt / 3.0;
    return c_enqueue_rhs<swmc::RK_FINAL>(c, in[3], r);
}

void c_drop_graphs(swm_cctx *c) {
    if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    if (c->graph_multi) { (void)hipGraphExecDestroy(c->graph_multi); c->graph_multi = nullptr; }
    c->graph_state = nullptr;
}

void c_free(swm_cctx *c) {
    if (!c) return;
    c_drop_graphs(c);
    for (void *p : {(void *)c->geo, (void *)c->hs, (void *)c->lists, (void *)c->signs, (void *)c->work, (void *)c->qw,
                    (void *)c->dpart, (void *)c->qpairs, (void *)c->ecs_src, (void *)c->ecs_sign,
                    (void *)c->ecs_strips, (void *)c->ecs_ow, (void *)c->ecs_oind, (void *)c->ecs_vt,
                    (void *)c->ecs_vti, (void *)c->ecs_ext})
        if (p) (void)hipFree(p);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// host <-> device copy of one field of point type t (nx x ny points) between the host tile layout
// (6, N+1+2hw, N+1+2hw) and the stored layout
int c_copy_field(swm_cctx *c, int t, double *dev, double *host, int hw, bool upload) {
    const int N = c->N, nx = N + (t == 1 || t == 3), ny = N + (t == 2 || t == 3), ld = N + 1 + 2 * hw;
    std::vector<double> buf((size_t)c->fs, 0.0);
    if (upload) {
        for (int p = 0; p < 6; ++p)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i)
                    buf[((size_t)p * c->rows + (j - 1 + swmc::H)) * c->pitch + (i - 1 + swmc::H)] =
                        host[((size_t)p * ld + (j - 1 + hw)) * ld + (i - 1 + hw)];
        HIP_TRY(hipMemcpyAsync(dev, buf.data(), c->field_bytes(), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    } else {
        HIP_TRY(hipMemcpyAsync(buf.data(), dev, c->field_bytes(), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        std::fill(host, host + (size_t)6 * ld * ld, 0.0);
        for (int p = 0; p < 6; ++p)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i)
                    host[((size_t)p * ld + (j - 1 + hw)) * ld + (i - 1 + hw)] =
                        buf[((size_t)p * c->rows + (j - 1 + swmc::H)) * c->pitch + (i - 1 + swmc::H)];
    }
    return SWM_OK;
}

}  // namespace

extern "C" int swm_c_create(swm_cctx **out, int N, const char *div, const char *grad, const char *curl,
                            const char *coriolis, const char *ke, const char *co2contra, const char *massflux,
                            const char *quad, double scale, double omega, double grav) {
    if (!out) return fail(SWM_EARG, "ctx pointer is NULL");
    *out = nullptr;
    if (N < 8 || N > 8192) return fail(SWM_EARG, "N must be in [8, 8192]");
    if (!(scale > 0) || !(grav > 0)) return fail(SWM_EARG, "scale and grav must be positive");
    const std::pair<const char *, const char *> names[] = {
        {div, "divergence_c_sbp42"}, {grad, "gradient_c_sbp42"}, {curl, "curl_c_sbp42"},
        {coriolis, "coriolis_Cgrid_sbp42"}, {ke, "KE_Cgrid_sbp42"}, {co2contra, "co2contra_c_sbp42"},
        {massflux, "massflux_c_sbp42"}, {quad, "SBP_C42_quadrature"}};
    for (const auto &nm : names)
        if (!nm.first || std::string(nm.first) != nm.second)
            return fail(SWM_ENAME, std::string("unsupported C-grid operator name: ") + (nm.first ? nm.first : "(null)") +
                                       " (expected " + nm.second + ")");
    swm_cctx *c = new swm_cctx();
    c->N = N;
    c->rows = N + 1 + 2 * swmc::H;
    c->pitch = (c->rows + 7) / 8 * 8;
    c->plane = (long long)c->rows * c->pitch;
    c->fs = 6LL * c->plane;
    c->scale = scale;
    c->omega = omega;
    c->grav = grav;
    c->hx = (0.25 * std::acos(-1.0) - (-0.25 * std::acos(-1.0))) / (double)N;
    auto bail = [&](int code) { c_free(c); return code; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SWM_EHIP, "hipStreamCreate failed"));
    c->own_stream = true;

    const size_t F = (size_t)c->fs;
    std::vector<double> geo(8 * F, 0.0);
    const double axis[3] = {0.0, 0.0, 1.0};
    double *G = geo.data();
    swmh_c_geometry(N, 0, swmc::H, c->rows, c->pitch, omega, axis, G, nullptr, G + 4 * F, nullptr, G + 7 * F);
    swmh_c_geometry(N, 1, swmc::H, c->rows, c->pitch, omega, axis, G + F, G + 5 * F, nullptr, nullptr, nullptr);
    swmh_c_geometry(N, 2, swmc::H, c->rows, c->pitch, omega, axis, G + 2 * F, nullptr, nullptr, G + 6 * F, nullptr);
    swmh_c_geometry(N, 3, swmc::H, c->rows, c->pitch, omega, axis, G + 3 * F, nullptr, nullptr, nullptr, nullptr);
    // device copy: panel 0 of the seven metric planes (every panel holds the same values, checked
    // here bit for bit), then f on all six panels
    const size_t P = (size_t)c->plane;
    std::vector<double> dgeo(7 * P + F);
    for (int w = 0; w < 7; ++w) {
        const double *m = G + (size_t)w * F;
        for (int p = 1; p < 6; ++p)
            if (std::memcmp(m, m + p * P, P * sizeof(double)) != 0)
                return bail(fail(SWM_EARG, "C-grid metric differs between panels"));
        std::copy(m, m + P, dgeo.begin() + (size_t)w * P);
    }
    std::copy(G + 7 * F, G + 8 * F, dgeo.begin() + 7 * P);
    if (hipMalloc((void **)&c->geo, dgeo.size() * sizeof(double)) != hipSuccess ||
        hipMemcpy(c->geo, dgeo.data(), dgeo.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(SWM_EHIP, "C-grid geometry upload failed"));

    const int dvec = 3, dO = 3;
    const size_t mv = (size_t)6 * 4 * (N + 1) * dvec, mf = (size_t)6 * 4 * N, mo = (size_t)6 * 4 * N * dO,
                 mq = (size_t)6 * 4 * (N + 1);
    std::vector<long long> vd(mv), vs(mv), fd(mf), fsrc(mf), od(mo), os(mo), qo(mq), qn(mq);
    std::vector<signed char> vsg(mv), fsg(mf);
    int counts[4] = {0, 0, 0, 0};
    const int hr = swmh_c_halo_lists(N, swmc::H, c->rows, c->pitch, c->fs, dvec, dO, vd.data(), vs.data(), vsg.data(),
                                     fd.data(), fsrc.data(), fsg.data(), od.data(), os.data(), qo.data(), qn.data(), counts);
    if (hr != 0 || (size_t)counts[0] != mv || (size_t)counts[1] != mf || (size_t)counts[2] != mo || (size_t)counts[3] != mq)
        return bail(fail(SWM_EARG, "C-grid halo lists do not match the cube topology"));
    // every offset stays inside one field pair (the pair for vectors, one field otherwise)
    for (size_t k = 0; k < mv; ++k)
        if (vd[k] < 0 || vd[k] >= 2 * c->fs || vs[k] < 0 || vs[k] >= 2 * c->fs) return bail(fail(SWM_EARG, "vector halo offset out of range"));
    for (size_t k = 0; k < mf; ++k)
        if (fd[k] < 0 || fd[k] >= 2 * c->fs || fsrc[k] < 0 || fsrc[k] >= 2 * c->fs) return bail(fail(SWM_EARG, "face halo offset out of range"));
    for (size_t k = 0; k < mo; ++k)
        if (od[k] < 0 || od[k] >= c->fs || os[k] < 0 || os[k] >= c->fs) return bail(fail(SWM_EARG, "O halo offset out of range"));
    for (size_t k = 0; k < mq; ++k)
        if (qn[k] < 0 || qn[k] >= c->fs) return bail(fail(SWM_EARG, "q sync offset out of range"));
    // device lists: sources in the order the kernels index them (swmc::CGeo)
    std::vector<long long> L;
    L.reserve(mv + mf + mo + mq);
    for (auto *v : {&vs, &fsrc, &os, &qn}) L.insert(L.end(), v->begin(), v->end());
    std::vector<signed char> S;
    S.insert(S.end(), vsg.begin(), vsg.end());
    S.insert(S.end(), fsg.begin(), fsg.end());
    if (hipMalloc((void **)&c->lists, L.size() * sizeof(long long)) != hipSuccess ||
        hipMemcpy(c->lists, L.data(), L.size() * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc((void **)&c->signs, S.size()) != hipSuccess ||
        hipMemcpy(c->signs, S.data(), S.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(SWM_EHIP, "C-grid halo list upload failed"));
    c->vsrc = c->lists;
    c->fsrc = c->vsrc + mv;
    c->osrc = c->fsrc + mf;
    c->qnb = c->osrc + mo;
    c->vsg = c->signs;
    c->fsg = c->signs + mv;
    {
        // q-point sync groups from the neighbour copies qn[(p*4+e)*(N+1) + k-1] of edge point k of
        // edge e (0: i = 1, 1: i = N+1, 2: j = 1, 3: j = N+1) of panel p
        const int n = N + 1;
        auto own = [&](int p, int i, int j) { return ((long long)p * c->rows + (j - 1 + swmc::H)) * c->pitch + (i - 1 + swmc::H); };
        auto nb = [&](int p, int e, int k) { return qn[((size_t)p * 4 + e) * n + (k - 1)]; };
        std::vector<long long> pairs, corners;
        for (int p = 0; p < 6; ++p)
            for (int e = 0; e < 4; ++e)
                for (int k = 2; k <= n - 1; ++k) {
                    const int i = e == 0 ? 1 : e == 1 ? n : k, j = e == 2 ? 1 : e == 3 ? n : k;
                    const long long a = own(p, i, j), b = nb(p, e, k);
                    if (a < b) { pairs.push_back(a); pairs.push_back(b); }
                }
        // corner copies: (own, first, second) in the order of halo_Ah_scalar_sync_mod.f90:56-72
        struct Copy { long long o, a, b; };
        std::vector<Copy> cps;
        for (int p = 0; p < 6; ++p) {
            cps.push_back({own(p, 1, 1), nb(p, 0, 1), nb(p, 2, 1)});
            cps.push_back({own(p, 1, n), nb(p, 0, n), nb(p, 3, 1)});
            cps.push_back({own(p, n, 1), nb(p, 1, 1), nb(p, 2, n)});
            cps.push_back({own(p, n, n), nb(p, 1, n), nb(p, 3, n)});
        }
        std::vector<bool> used(cps.size(), false);
        for (size_t x = 0; x < cps.size(); ++x) {
            if (used[x]) continue;
            long long g3[3] = {cps[x].o, cps[x].a, cps[x].b};
            size_t idx[3] = {x, 0, 0};
            for (int m = 1; m < 3; ++m) {
                idx[m] = cps.size();
                for (size_t y = 0; y < cps.size(); ++y)
                    if (cps[y].o == g3[m]) idx[m] = y;
                if (idx[m] == cps.size()) return bail(fail(SWM_EARG, "cube corner copies do not form a group"));
            }
            long long ords[3];
            for (int m = 0; m < 3; ++m) {
                const Copy &cp = cps[idx[m]];
                int fa = -1, fb = -1;
                for (int t = 0; t < 3; ++t) {
                    if (g3[t] == cp.a) fa = t;
                    if (g3[t] == cp.b) fb = t;
                }
                if (fa < 0 || fb < 0 || fa == m || fb == m || fa == fb)
                    return bail(fail(SWM_EARG, "cube corner copies do not form a group"));
                used[idx[m]] = true;
                ords[m] = 3 * fa + fb;
            }
            for (int m = 0; m < 3; ++m) corners.push_back(g3[m]);
            for (int m = 0; m < 3; ++m) corners.push_back(ords[m]);
        }
        if (pairs.size() != (size_t)2 * 12 * (n - 2) || corners.size() != (size_t)6 * 8)
            return bail(fail(SWM_EARG, "q-point sync groups do not match the cube topology"));
        c->nqpairs = (int)(pairs.size() / 2);
        c->nqcorners = (int)(corners.size() / 6);
        std::vector<long long> Q(pairs);
        Q.insert(Q.end(), corners.begin(), corners.end());
        if (hipMalloc((void **)&c->qpairs, Q.size() * sizeof(long long)) != hipSuccess ||
            hipMemcpy(c->qpairs, Q.data(), Q.size() * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(SWM_EHIP, "q-point sync group upload failed"));
        c->qcorners = c->qpairs + pairs.size();
    }

    {
        // create_mass_coefficients (quadrature_factory_mod.f90:156-175): A_centers (13/12, 7/8,
        // 25/24, 1, ...) and A_interfaces (7/18, 9/8, 1, 71/72, 1, ...) from both ends
        const double Ac[4] = {13.0 / 12.0, 7.0 / 8.0, 25.0 / 24.0, 1.0};
        const double Ai[5] = {7.0 / 18.0, 9.0 / 8.0, 1.0, 71.0 / 72.0, 1.0};
        std::vector<double> q((size_t)2 * (N + 2), 1.0);
        double *acen = q.data(), *aint = q.data() + (N + 2);
        for (int i = 1; i <= std::min(4, N); ++i) acen[i] = Ac[i - 1];
        for (int i = std::max(N - 3, 1); i <= N; ++i) acen[i] = Ac[N - i];
        for (int i = 1; i <= std::min(5, N + 1); ++i) aint[i] = Ai[i - 1];
        for (int i = std::max(N + 1 - 4, 1); i <= N + 1; ++i) aint[i] = Ai[N + 1 - i];
        c->dblocks = (long long)((N + 1 + swmc::BX - 1) / swmc::BX) * ((N + 1 + swmc::BY - 1) / swmc::BY) * 6;
        if (hipMalloc((void **)&c->qw, q.size() * sizeof(double)) != hipSuccess ||
            hipMemcpy(c->qw, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMalloc((void **)&c->dpart, (size_t)(c->dblocks * 10 + 8) * sizeof(double)) != hipSuccess)
            return bail(fail(SWM_EHIP, "C-grid diagnostics set-up failed"));
    }
    const size_t wb = (size_t)CW_COUNT * c->field_bytes();
    if (hipMalloc((void **)&c->work, wb) != hipSuccess || hipMemset(c->work, 0, wb) != hipSuccess)
        return bail(fail(SWM_EHIP, "C-grid work buffer allocation failed"));
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(SWM_EHIP, "set-up synchronisation failed"));
    *out = c;
    return SWM_OK;
}

extern "C" int swm_c_destroy(swm_cctx *c) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c_free(c);
    return SWM_OK;
}

extern "C" int swm_c_info(const swm_cctx *c, int *N, int *halo, int *rows, int *pitch, long long *field_elems) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    if (N) *N = c->N;
    if (halo) *halo = swmc::H;
    if (rows) *rows = c->rows;
    if (pitch) *pitch = c->pitch;
    if (field_elems) *field_elems = c->fs;
    return SWM_OK;
}

extern "C" int swm_c_synchronize(swm_cctx *c) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SWM_OK;
}

extern "C" int swm_c_state_alloc(swm_cctx *c, void **state) {
    if (!c || !state) return fail(SWM_EARG, "NULL argument");
    const size_t sb = 3 * c->field_bytes();
    HIP_TRY(hipMalloc(state, sb));
    HIP_TRY(hipMemsetAsync(*state, 0, sb, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SWM_OK;
}

extern "C" int swm_c_state_free(swm_cctx *c, void *state) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    if (!state) return SWM_OK;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->graph_state == state) c_drop_graphs(c);
    HIP_TRY(hipFree(state));
    return SWM_OK;
}

extern "C" int swm_c_state_upload(swm_cctx *c, void *state, const double *h, const double *u, const double *v, int hw) {
    if (!c || !state || !h || !u || !v) return fail(SWM_EARG, "NULL argument");
    if (hw < 0 || hw > 64) return fail(SWM_EARG, "halo_width out of range");
    const double *src[3] = {h, u, v};
    for (int k = 0; k < 3; ++k) {
        int rc = c_copy_field(c, k, (double *)state + (size_t)k * c->fs, const_cast<double *>(src[k]), hw, true);
        if (rc) return rc;
    }
    return SWM_OK;
}

extern "C" int swm_c_state_download(swm_cctx *c, const void *state, double *h, double *u, double *v, int hw) {
    if (!c || !state || !h || !u || !v) return fail(SWM_EARG, "NULL argument");
    if (hw < 0 || hw > 64) return fail(SWM_EARG, "halo_width out of range");
    double *dst[3] = {h, u, v};
    for (int k = 0; k < 3; ++k) {
        int rc = c_copy_field(c, k, (double *)state + (size_t)k * c->fs, dst[k], hw, false);
        if (rc) return rc;
    }
    return SWM_OK;
}

extern "C" int swm_c_set_orography(swm_cctx *c, const double *hs, int hw) {
    if (!c) return fail(SWM_EARG, "ctx is NULL");
    if (hw < 0 || hw > 64) return fail(SWM_EARG, "halo_width out of range");
    HIP_TRY(hipStreamSynchronize(c->stream));
    c_drop_graphs(c);
    if (!hs) {
        if (c->hs) HIP_TRY(hipFree(c->hs));
        c->hs = nullptr;
        return SWM_OK;
    }
    if (!c->hs) {
        HIP_TRY(hipMalloc((void **)&c->hs, c->field_bytes()));
        HIP_TRY(hipMemsetAsync(c->hs, 0, c->field_bytes(), c->stream));
    }
    return c_copy_field(c, 0, c->hs, const_cast<double *>(hs), hw, true);
}

extern "C" int swm_c_rhs(swm_cctx *c, const void *vin, void *vout) {
    if (!c || !vin || !vout) return fail(SWM_EARG, "NULL argument");
    if (vin == vout) return fail(SWM_EARG, "vin and vout must differ");
    swmc::RkArgs r{};
    r.k = (double *)vout;
    return c_enqueue_rhs<swmc::RK_RHS>(c, (const double *)vin, r);
}

extern "C" int swm_c_rk4_steps(swm_cctx *c, void *state, double dt, int nsteps) {
    if (!c || !state) return fail(SWM_EARG, "NULL argument");
    if (nsteps < 0) return fail(SWM_EARG, "nsteps must be >= 0");
    if (nsteps == 0) return SWM_OK;
    if (c->graph_state != state || c->graph_dt != dt) c_drop_graphs(c);
    const int nmulti = nsteps / RK4_GRAPH_STEPS, nsingle = nsteps - nmulti * RK4_GRAPH_STEPS;
    for (int which = 0; which < 2; ++which) {
        hipGraphExec_t &ge = which == 0 ? c->graph : c->graph_multi;
        const int per = which == 0 ? 1 : RK4_GRAPH_STEPS;
        if (ge || (which == 0 ? nsingle : nmulti) == 0) continue;
        hipGraph_t gr = nullptr;
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = SWM_OK;
        for (int k = 0; k < per && !rc; ++k) rc = c_enqueue_rk4_step(c, (double *)state, dt);
        hipError_t e = hipStreamEndCapture(c->stream, &gr);
        if (rc) { if (gr) (void)hipGraphDestroy(gr); return rc; }
        if (e != hipSuccess) return fail(SWM_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) { ge = nullptr; return fail(SWM_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e)); }
        c->graph_state = state;
        c->graph_dt = dt;
    }
    for (int s = 0; s < nmulti; ++s) HIP_TRY(hipGraphLaunch(c->graph_multi, c->stream));
    for (int s = 0; s < nsingle; ++s) HIP_TRY(hipGraphLaunch(c->graph, c->stream));
    return SWM_OK;
}

extern "C" int swm_c_diagnostics(swm_cctx *c, const void *state, double out[7]) {
    if (!c || !state || !out) return fail(SWM_EARG, "NULL argument");
    const swmc::CGeo g = make_cgeo(c);
    const dim3 blk(swmc::BX, swmc::BY), grd(8 * g.per);
    double *res = c->dpart + c->dblocks * 10;
    hipLaunchKernelGGL(swmc::k_cdiag_partial, grd, blk, 0, c->stream, g, (const double *)state, c->qw,
                       c->qw + (c->N + 2), c->dpart);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(swmc::k_cdiag_final, dim3(1), dim3(64), 0, c->stream, c->dpart, c->dblocks, res);
    HIP_TRY(hipGetLastError());
    double r[6];
    HIP_TRY(hipMemcpyAsync(r, res, sizeof(r), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double ke = r[1] + r[2];
    out[0] = r[4];
    out[1] = r[5];
    out[2] = r[0];
    out[3] = ke + r[3];
    out[4] = ke;
    out[5] = r[3];
    out[6] = std::nan("");
    return SWM_OK;
}
