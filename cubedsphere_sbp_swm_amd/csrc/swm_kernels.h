// Device kernels of the Ah-grid vector-invariant shallow-water step (gfx950).
//
// One launch per RK stage evaluates the whole right-hand side of
// operator_swm_t%apply (models/shallow_water/operator/operator_swm_mod.f90:61-113) for a
// tile of points and applies the RK4 stage update (time_schemes/rk4_mod.f90:20-42):
//   * the five stencil inputs K = (u.ut+v.vt)/2 + g h + g hs, J*hu, J*hv, J*(v/J), J*(-u/J)
//     are formed once per point in LDS from the state and the metric;
//   * SBP derivative rows (interior stencil or boundary closure) are applied along x and y;
//     the tiles compute every point off the panel edges;
//   * blocks of their own in the same launch compute the panel-edge points: they load both
//     panels' points along a segment of a shared edge (depth 0 .. closure width) into LDS,
//     evaluate both one-sided derivative rows at the shared points, and combine them with
//     the reference's scalar and covariant-vector edge averaging
//     (halo_Ah_scalar_sync_mod.f90:52-93, ecs_Ah_vec_sync_mod.f90:77-124) -- this replaces
//     the MPI halo exchange;
//   * the Coriolis term, tendencies and the RK stage combination finish in registers.
// FP64 builds use -ffp-contract=off and the reference's operation order, so results agree
// bit for bit with an FP64 CPU evaluation of the same algorithm.
#pragma once

#include <hip/hip_runtime.h>

#include "sbp_tables.h"

namespace swm {

// ----------------------------------------------------------------------------------------
// Geometry and per-launch arguments
template <typename T>
struct DevGeo {
    const T *J, *Qi11, *Qi12, *Qi22, *f;   // [panel][j][i], row pitch `pitch`
    // panel stride of J, Qi11, Qi12, Qi22 in elements: n*pitch, or 0 where every panel holds the
    // same values (the ECS metric depends on alpha, beta only) and panel 0's copy serves all six
    long long mps;
    const T *qb, *qt, *ql, *qr;            // covariant sync weights, n values each
    const int *einfo;                      // per (panel, edge): q, nu_is_x, tnu, a_base, s, code
    const double *tanA;                    // tan(alpha_i), i = 0..n-1 (metric evaluated on the fly)
    int n, pitch;
    long long fs;                          // elements per field (6*n*pitch)
    T hxs;                                 // hx*scale
    T hx, scale, grav;
};

enum StageMode : int {
    MODE_RHS = 0,     // out = RHS(y)
    MODE_STAGE1 = 1,  // out = y + cy*k ; acc = y + cacc*k          (y is the step's v0)
    MODE_STAGE23 = 2, // out = v0 + cy*k ; acc = acc + cacc*k
    MODE_STAGE4 = 3,  // out = acc + cy*k                           (out is v0)
    MODE_CURL = 4     // out(h plane) = synced relative vorticity of y
};

template <typename T>
struct StageArgs {
    const T *y;    // input state: h at y, u at y+fs, v at y+2fs
    T *out;
    const T *v0;
    T *acc;
    const T *hs;   // orography, may be null
    T cy, cacc;
    T *zeta;       // optional: the synced relative vorticity of y is also written here
};

// Tile shapes (columns x rows of points, one thread per point) of the stage kernels: 64 x 8,
// 32 x 16, 16 x 32 (512 threads) and 32 x 8 (256 threads).  tile_width(ty) is the 512-thread
// shape of `ty` rows.
SWM_HD constexpr int tile_width(int ty) { return 512 / ty; }
SWM_HD constexpr bool tile_shape_ok(int tx, int ty) {
    return (tx == 64 && ty == 8) || (tx == 32 && ty == 16) || (tx == 16 && ty == 32) || (tx == 32 && ty == 8);
}
// The tiles cover the points 1 .. n-2 of a panel line (the edge segments own 0 and n-1):
// tiles_along(n, t) tiles of t points, the b-th starting at point 1 + b*t.
SWM_HD constexpr int tiles_along(int n, int t) { return (n - 2 + t - 1) / t; }

// The panel-edge points are computed by edge segments of edge_seg_len(n) consecutive points:
// per panel the x-edges (i = 0 and i = n-1, positions j = 0..n-1, cube corners included) and
// the y-edges (j = 0 and j = n-1, positions i = 1..n-2).  A single cube with small panels
// takes shorter segments, i.e. more blocks for the same edge work; the persistent ensemble
// loop, where every task costs a fixed overhead, keeps the longest.
constexpr int EDGE_SEG = 64;  // longest segment (LDS sizing)
SWM_HD constexpr int edge_seg_len(int n, bool persist = false) { return (persist || n > 256) ? EDGE_SEG : 48; }
SWM_HD constexpr int edge_segs_x(int n, bool persist = false) {
    return (n + edge_seg_len(n, persist) - 1) / edge_seg_len(n, persist);
}
SWM_HD constexpr int edge_segs_y(int n, bool persist = false) {
    return (n - 2 + edge_seg_len(n, persist) - 1) / edge_seg_len(n, persist);
}
SWM_HD constexpr int edge_segs(int n, bool persist = false) {
    return 6 * 2 * (edge_segs_x(n, persist) + edge_segs_y(n, persist));
}

template <typename T>
struct EnsMember {
    DevGeo<T> g;
    StageArgs<T> a;
};

// ----------------------------------------------------------------------------------------
// Pointwise stencil inputs of one point (co2contra_colocated_mod.f90:55-56,
// massflux_colocated_mod.f90:26-27, KE_colocated_mod.f90:44, operator_swm_mod.f90:85,
// div_ah_sbp_mod.f90:80-91, curl_div_based_mod.f90:58-59)
template <typename T>
struct Derived {
    T K, Gxd, Gyd, Gxc, Gyc;
    T hu, hv;  // h*ut, h*vt (kept in registers for the point's own output)
};

// WHICH: 0 all five inputs, 2 only those of the y-derivative (K, J*hv, J*(-u/J))
template <typename T, int WHICH = 0>
__device__ __forceinline__ Derived<T> make_derived(T h, T u, T v, T J, T q11, T q12, T q22, T hsv, T grav) {
    Derived<T> d;
    const T ut = q11 * u + q12 * v;
    const T vt = q12 * u + q22 * v;
    const T hu = h * ut;
    const T hv = h * vt;
    T K = (u * ut + v * vt) / T(2);
    K = K + grav * h + grav * hsv;
    d.K = K;
    d.hu = hu;
    d.hv = hv;
    if constexpr (WHICH != 2) {
        d.Gxd = J * hu;
        d.Gxc = J * (v / J);
    }
    d.Gyd = J * hv;
    d.Gyc = J * (-u / J);
    return d;
}

template <typename T>
__device__ __forceinline__ long long pidx(const DevGeo<T> &g, int p, int i, int j) {
    return ((long long)p * g.n + j) * g.pitch + i;
}
// index of the metric terms J, Qi of point (i, j) of panel p
template <typename T>
__device__ __forceinline__ long long midx(const DevGeo<T> &g, int p, int i, int j) {
    return (long long)p * g.mps + (long long)j * g.pitch + i;
}

// Metric terms of Ah point (i,j) from tan(alpha_i), tan(beta_j), with the operation order of
// the host set-up (ecs_metric_mod.f90:413-463); FP64 only.
struct Metric {
    double J, Qi11, Qi12, Qi22;
};
__device__ __forceinline__ Metric metric_otf(const double *tanA, int i, int j) {
    const double ta = tanA[i], tb = tanA[j];
    const double sg = sqrt(1.0 + ta * ta + tb * tb);
    Metric m;
    m.J = (1.0 + ta * ta) * (1.0 + tb * tb) / (sg * sg * sg);
    m.Qi11 = sg * sg / (1.0 + ta * ta);
    m.Qi12 = sg * sg * ta * tb / ((1 + ta * ta) * (1 + tb * tb));
    m.Qi22 = sg * sg / (1.0 + tb * tb);
    return m;
}

template <bool OTF, typename T>
__device__ __forceinline__ T metric_J(const DevGeo<T> &g, int p, int i, int j) {
    if constexpr (OTF) return metric_otf(g.tanA, i, j).J;
    else return g.J[midx(g, p, i, j)];
}

template <bool OTF, typename T, int WHICH = 0>
__device__ __forceinline__ Derived<T> derived_global(const DevGeo<T> &g, const T *y, const T *hs, int p, int i, int j) {
    const long long id = pidx(g, p, i, j);
    if constexpr (OTF) {
        const Metric m = metric_otf(g.tanA, i, j);
        return make_derived<T, WHICH>(y[id], y[id + g.fs], y[id + 2 * g.fs], (T)m.J, (T)m.Qi11, (T)m.Qi12,
                                      (T)m.Qi22, hs ? hs[id] : T(0), g.grav);
    } else {
        const long long mid = midx(g, p, i, j);
        return make_derived<T, WHICH>(y[id], y[id + g.fs], y[id + 2 * g.fs], g.J[mid], g.Qi11[mid], g.Qi12[mid],
                                      g.Qi22[mid], hs ? hs[id] : T(0), g.grav);
    }
}

// Raw inputs of one point (state, metric, orography) and the derived stencil inputs formed
// from them; split so that the loads of several points can be issued before any is used.
template <typename T>
struct RawPoint {
    T h, u, v, J, q11, q12, q22, hs;
};

template <bool OTF, typename T>
__device__ __forceinline__ RawPoint<T> load_point(const DevGeo<T> &g, const T *y, const T *hs, int p, int i, int j) {
    const long long id = pidx(g, p, i, j);
    RawPoint<T> r;
    r.h = y[id];
    r.u = y[id + g.fs];
    r.v = y[id + 2 * g.fs];
    if constexpr (OTF) {
        const Metric m = metric_otf(g.tanA, i, j);
        r.J = (T)m.J; r.q11 = (T)m.Qi11; r.q12 = (T)m.Qi12; r.q22 = (T)m.Qi22;
    } else {
        const long long mid = midx(g, p, i, j);
        r.J = g.J[mid]; r.q11 = g.Qi11[mid]; r.q12 = g.Qi12[mid]; r.q22 = g.Qi22[mid];
    }
    r.hs = hs ? hs[id] : T(0);
    return r;
}

template <typename T, int WHICH = 0>
__device__ __forceinline__ Derived<T> derived_of(const DevGeo<T> &g, const RawPoint<T> &r) {
    return make_derived<T, WHICH>(r.h, r.u, r.v, r.J, r.q11, r.q12, r.q22, r.hs, g.grav);
}

// ----------------------------------------------------------------------------------------
// SBP rows (sbp_operator_mod.f90:187-223).  `get(a)` returns the input at 0-based index a
// along the line; t is the 0-based output index, n the line length.
template <int ORD, int R, typename T, typename G>
__device__ __forceinline__ T sbp_left_row(const G &get) {
    using S = Sbp<ORD>;
    T s = T(S::Ql[R][0]) * get(0);
#pragma unroll
    for (int m = 1; m < S::last[R]; ++m) s = s + T(S::Ql[R][m]) * get(m);
    return s;
}
// right-boundary row ir (0-based): the left row RR = NE-1-ir reflected, sign flipped
template <int ORD, int IR, typename T, typename G>
__device__ __forceinline__ T sbp_right_row(int n, const G &get) {
    using S = Sbp<ORD>;
    constexpr int RR = S::NE - 1 - IR;
    constexpr int LAST = S::last[RR];
    const int base = n - LAST;
    T s = T(-1.0 * S::Ql[RR][LAST - 1]) * get(base);
#pragma unroll
    for (int c = 1; c < LAST; ++c) s = s + T(-1.0 * S::Ql[RR][LAST - 1 - c]) * get(base + c);
    return s;
}

template <int ORD, typename T, typename G, int... Is>
__device__ __forceinline__ T sbp_left_dispatch(int t, const G &get, std::integer_sequence<int, Is...>) {
    T r = T(0);
    ((t == Is ? (void)(r = sbp_left_row<ORD, Is, T>(get)) : (void)0), ...);
    return r;
}
template <int ORD, typename T, typename G, int... Is>
__device__ __forceinline__ T sbp_right_dispatch(int ir, int n, const G &get, std::integer_sequence<int, Is...>) {
    T r = T(0);
    ((ir == Is ? (void)(r = sbp_right_row<ORD, Is, T>(n, get)) : (void)0), ...);
    return r;
}

// interior row of the SBP first derivative at t (NE <= t < n - NE)
template <int ORD, typename T, typename G>
__device__ __forceinline__ T sbp_row_in(int t, const G &get) {
    using S = Sbp<ORD>;
    T s = T(S::Qin[0]) * get(t + S::INSHIFT);
#pragma unroll
    for (int k = 1; k < S::WIN; ++k) s = s + T(S::Qin[k]) * get(t + S::INSHIFT + k);
    return s;
}

template <int ORD, typename T, typename G>
__device__ __forceinline__ T sbp_row(int t, int n, const G &get) {
    using S = Sbp<ORD>;
    if (t >= S::NE && t < n - S::NE) return sbp_row_in<ORD, T>(t, get);
    if (t < S::NE) return sbp_left_dispatch<ORD, T>(t, get, std::make_integer_sequence<int, S::NE>{});
    return sbp_right_dispatch<ORD, T>(t - (n - S::NE), n, get, std::make_integer_sequence<int, S::NE>{});
}

// Vector halo convention: neighbour components (a,b) -> local (u,v)
template <typename T>
__device__ __forceinline__ void map_vec(int code, T a, T b, T &u, T &v) {
    const T x = (code & 1) ? b : a;
    const T z = (code & 1) ? a : b;
    u = (code & 2) ? -x : x;
    v = (code & 4) ? -z : z;
}

// XCD-aware mapping: consecutive block ids go round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so that tiles sharing halo rows share an L2.
__device__ __forceinline__ int xcd_tile(int bid, int nblocks) {
    const int q = nblocks / 8, r = nblocks % 8;
    const int xcd = bid % 8, local = bid / 8;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

}  // namespace swm
