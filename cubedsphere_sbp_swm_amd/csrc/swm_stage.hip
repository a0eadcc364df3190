// Fused RK-stage kernel of the Ah-grid vector-invariant shallow-water model (see
// swm_kernels.h for the structure) and the small kernels around it.
#include <hip/hip_runtime.h>

#include "swm_kernels.h"

namespace swm {


// Tile of TX x TY output points per workgroup of TX*TY threads (one point per thread).  Every
// tiling has 512 threads: 64 x 8, 32 x 16 or 16 x 32 points.
template <int ORD, typename T, int TX, int TY>
struct Tile {
    static constexpr int R = Sbp<ORD>::REACH;    // halo of the stencil inputs
    static constexpr int D = Sbp<ORD>::last[0];  // points of the boundary row 0 (normal depth)
    static constexpr int LX = TX + 2 * R;
    static constexpr int LY = TY + 2 * R;
    static constexpr int NT = TX * TY;
};

// LDS of one tile.  K lives on the cross-shaped region; the x-derivative inputs only on the
// tile rows, the y-derivative inputs only on the tile columns.
template <int ORD, typename T, int TX, int TY>
struct TileSmem {
    using C = Tile<ORD, T, TX, TY>;
    static constexpr int R = C::R, LX = C::LX, LY = C::LY;
    T K[LY][LX];
    T Gxd[TY][LX], Gxc[TY][LX];
    T Gyd[LY][TX], Gyc[LY][TX];
};

// LDS of one edge segment: the stencil inputs along the edge of the own panel (side 0) and of
// the neighbour across the edge (side 1), and at the cube corners the neighbour across the
// other edge (slot 0: position 0, slot 1: position n-1).
template <int ORD, typename T>
struct EdgeSmem {
    static constexpr int R = Sbp<ORD>::REACH, D = Sbp<ORD>::last[0];
    T e0[2][EDGE_SEG + 2 * R][6];   // depth 0: K, Gxd, Gyd, Gxc, Gyc, J
    T en[2][D - 1][EDGE_SEG][3];    // depth >= 1: K and the two inputs along the normal axis
    T c0[2][2 * R + 1][6];
    T cn[2][D - 1][3];
};

template <int ORD, typename T, int TX, int TY>
union Smem {
    TileSmem<ORD, T, TX, TY> t;
    EdgeSmem<ORD, T> e;
};

// Raw values (gx, gy, div, curl before the edge syncs) of a panel at the point k of one of
// its edges, from strips of stencil inputs along that edge.  ei describes the edge in the
// neighbour-table form (panel, normal axis is x, normal index, tangential base and step); for
// a neighbour panel it maps the local position k to the neighbour's tangential index.
// d0(kk, f) returns field f (0 K, 1 Gxd, 2 Gyd, 3 Gxc, 4 Gyc, 5 J) at tangential index kk,
// depth 0; dn(d, f3) the normal-axis inputs (0 K, 1 Gnd, 2 Gnc) at depth d >= 1.
// One panel edge in neighbour-table form: panel, normal axis is x, normal index of the edge
// line, tangential index = a_base + s * (local position), component code of the vector halo.
struct EdgeDesc {
    int q, nu_x, tnu, a_base, s, code;
};
__device__ __forceinline__ EdgeDesc edge_desc(const int *ei) { return {ei[0], ei[1], ei[2], ei[3], ei[4], ei[5]}; }

template <int ORD, typename T, typename D0, typename DN>
__device__ __forceinline__ void neighbour_raw(const DevGeo<T> &g, const EdgeDesc &ed, int k, const D0 &d0,
                                              const DN &dn, T out[4]) {
    const int n = g.n;
    const int nu_x = ed.nu_x, tnu = ed.tnu, a0 = ed.a_base + ed.s * k, s = ed.s;
    auto nu_get = [&](int fld3, int a) {  // along the normal axis of the neighbour
        const int d = tnu == 0 ? a : n - 1 - a;
        if (d == 0) return d0(k, fld3 == 0 ? 0 : (nu_x ? (fld3 == 1 ? 1 : 3) : (fld3 == 1 ? 2 : 4)));
        return dn(d, fld3);
    };
    auto tau_get = [&](int fld, int a) { return d0(k + s * (a - a0), fld); };
    const T nK = sbp_row<ORD, T>(tnu, n, [&](int a) { return nu_get(0, a); });
    const T nD = sbp_row<ORD, T>(tnu, n, [&](int a) { return nu_get(1, a); });
    const T nC = sbp_row<ORD, T>(tnu, n, [&](int a) { return nu_get(2, a); });
    const int fD = nu_x ? 2 : 1, fC = nu_x ? 4 : 3;  // inputs along the tangential axis
    const T tK = sbp_row<ORD, T>(a0, n, [&](int a) { return tau_get(0, a); });
    const T tD = sbp_row<ORD, T>(a0, n, [&](int a) { return tau_get(fD, a); });
    const T tC = sbp_row<ORD, T>(a0, n, [&](int a) { return tau_get(fC, a); });
    const T dxK = nu_x ? nK : tK, dyK = nu_x ? tK : nK;
    const T dxD = nu_x ? nD : tD, dyD = nu_x ? tD : nD;
    const T dxC = nu_x ? nC : tC, dyC = nu_x ? tC : nC;
    const T Jd = d0(k, 5) * g.hx * g.scale;
    out[0] = dxK / g.hxs;
    out[1] = dyK / g.hxs;
    out[2] = (dxD + dyD) / Jd;
    out[3] = (dxC + dyC) / Jd;
}

// Values of the output point that only the final combination reads (Coriolis parameter, RK
// stage vectors); loaded at the start of the tile so that their latency overlaps the stencil
// phases.
template <typename T>
struct OutPre {
    T f, v0[3], acc[3];
};

template <typename T, int MODE>
__device__ __forceinline__ OutPre<T> prefetch_point(const DevGeo<T> &g, const StageArgs<T> &a, long long id) {
    OutPre<T> o{};
    if constexpr (MODE == MODE_CURL) return o;
    o.f = g.f[id];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if constexpr (MODE == MODE_STAGE23) o.v0[k] = a.v0[id + k * g.fs];
        if constexpr (MODE == MODE_STAGE23 || MODE == MODE_STAGE4) o.acc[k] = a.acc[id + k * g.fs];
    }
    return o;
}

// Coriolis term, tendencies and RK stage combination of one point
// (coriolis_colocated_mod.f90:97-98, operator_swm_mod.f90:95-101, rk4_mod.f90:28-40)
// hu = h*ut, hv = h*vt: the mass fluxes of the point (u, v are read by stage 1 only)
template <typename T, int MODE>
__device__ __forceinline__ void finish_point_hv(const DevGeo<T> &g, const StageArgs<T> &a, long long id, T h, T u, T v,
                                                T hu, T hv, T Jv, T gx, T gy, T dv, T cu, const OutPre<T> &pre) {
    if constexpr (MODE == MODE_CURL) {
        a.out[id] = cu;
        return;
    }
    if (a.zeta) a.zeta[id] = cu;
    const T fz = pre.f + cu;
    const T cor_u = fz * hv / h * Jv;
    const T cor_v = -(fz * hu / h * Jv);
    const T kh = -dv;
    const T ku = -gx + cor_u;
    const T kv = -gy + cor_v;

    if constexpr (MODE == MODE_RHS) {
        a.out[id] = kh;
        a.out[id + g.fs] = ku;
        a.out[id + 2 * g.fs] = kv;
    } else if constexpr (MODE == MODE_STAGE1) {
        a.out[id] = h + a.cy * kh;
        a.out[id + g.fs] = u + a.cy * ku;
        a.out[id + 2 * g.fs] = v + a.cy * kv;
        a.acc[id] = h + a.cacc * kh;
        a.acc[id + g.fs] = u + a.cacc * ku;
        a.acc[id + 2 * g.fs] = v + a.cacc * kv;
    } else if constexpr (MODE == MODE_STAGE23) {
        a.out[id] = pre.v0[0] + a.cy * kh;
        a.out[id + g.fs] = pre.v0[1] + a.cy * ku;
        a.out[id + 2 * g.fs] = pre.v0[2] + a.cy * kv;
        a.acc[id] = pre.acc[0] + a.cacc * kh;
        a.acc[id + g.fs] = pre.acc[1] + a.cacc * ku;
        a.acc[id + 2 * g.fs] = pre.acc[2] + a.cacc * kv;
    } else if constexpr (MODE == MODE_STAGE4) {
        a.out[id] = pre.acc[0] + a.cy * kh;
        a.out[id + g.fs] = pre.acc[1] + a.cy * ku;
        a.out[id + 2 * g.fs] = pre.acc[2] + a.cy * kv;
    }
}

template <typename T, int MODE>
__device__ __forceinline__ void finish_point(const DevGeo<T> &g, const StageArgs<T> &a, long long id, T h, T u, T v,
                                             T Jv, T q11, T q12, T q22, T gx, T gy, T dv, T cu, const OutPre<T> &pre) {
    const T ut = q11 * u + q12 * v;
    const T vt = q12 * u + q22 * v;
    finish_point_hv<T, MODE>(g, a, id, h, u, v, h * ut, h * vt, Jv, gx, gy, dv, cu, pre);
}

// One TX x TY tile of panel p starting at (tx0, ty0) >= (1, 1): the whole RHS + RK stage update
// of its points off the panel edges (the panel-edge points belong to the edge segments).
// INNER: the tile lies inside NE <= i, j < n - NE (inner_tile), so every point of it and of its
// halo exists, none is on a panel edge and every derivative row is an interior one; its guards
// and closure branches are then known at compile time.
template <int ORD, typename T, int MODE, int TX, int TY, bool OTF, bool PERSIST = false, bool INNER = false>
__device__ __forceinline__ void stage_tile(const DevGeo<T> &g, const StageArgs<T> &a, int p, int tx0, int ty0,
                                           TileSmem<ORD, T, TX, TY> &sm) {
    using C = Tile<ORD, T, TX, TY>;
    constexpr int R = C::R, NT = C::NT;
    const int n = g.n;
    // in the persistent ensemble loop the thread index is made opaque per tile, so that the
    // thread-index arithmetic is recomputed per tile instead of being hoisted out of the loop
    // and held in registers
    int tid = threadIdx.x;
    if constexpr (PERSIST) asm volatile("" : "+v"(tid));
    const T *y = a.y;
    OutPre<T> pre{};

    // ---- load phase: every thread loads its own tile point (kept in registers for the output
    // phase), plus at most NXH points of the x-halo columns and NYH points of the y-halo rows;
    // all global loads are issued before any derived value is formed, so that they are in flight
    // together. ----
    constexpr int NXH = (2 * R * TY + NT - 1) / NT, NYH = (2 * R * TX + NT - 1) / NT;
    const int tx = tid % TX, ty = tid / TX;
    const int i = tx0 + tx, j = ty0 + ty;
    const bool own_ok = INNER || (i < n && j < n);
    RawPoint<T> own{};
    if (own_ok) own = load_point<OTF>(g, y, a.hs, p, i, j);
    RawPoint<T> rx[NXH], ry[NYH];
    bool okx[NXH], oky[NYH];
#pragma unroll
    for (int k = 0; k < NXH; ++k) {
        const int c = tid + k * NT;
        const int r = c / (2 * R), q = c - r * (2 * R);
        const int lx = q < R ? q : q + TX;
        const int xi = tx0 - R + lx, xj = ty0 + r;
        okx[k] = c < 2 * R * TY && (INNER || (xi >= 0 && xi < n && xj < n));
        if (okx[k]) rx[k] = load_point<OTF>(g, y, a.hs, p, xi, xj);
    }
#pragma unroll
    for (int k = 0; k < NYH; ++k) {
        const int c = tid + k * NT;
        const int r = c / TX, cx = c - r * TX;
        const int ly = r < R ? r : r + TY;
        const int yi = tx0 + cx, yj = ty0 - R + ly;
        oky[k] = c < 2 * R * TX && (INNER || (yi < n && yj >= 0 && yj < n));
        if (oky[k]) ry[k] = load_point<OTF>(g, y, a.hs, p, yi, yj);
    }
    // of the own point only h, u, v, J and the mass fluxes h*ut, h*vt stay live for the output
    // phase (the same products the Coriolis term forms)
    T own_hu = T(0), own_hv = T(0);
    if (own_ok) {
        const Derived<T> d = derived_of(g, own);
        own_hu = d.hu;
        own_hv = d.hv;
        sm.K[ty + R][tx + R] = d.K;
        sm.Gxd[ty][tx + R] = d.Gxd;
        sm.Gxc[ty][tx + R] = d.Gxc;
        sm.Gyd[ty + R][tx] = d.Gyd;
        sm.Gyc[ty + R][tx] = d.Gyc;
    }
#pragma unroll
    for (int k = 0; k < NXH; ++k) {
        if (!okx[k]) continue;
        const int c = tid + k * NT;
        const int r = c / (2 * R), q = c - r * (2 * R);
        const int lx = q < R ? q : q + TX;
        const Derived<T> d = derived_of(g, rx[k]);
        sm.K[r + R][lx] = d.K;
        sm.Gxd[r][lx] = d.Gxd;
        sm.Gxc[r][lx] = d.Gxc;
    }
#pragma unroll
    for (int k = 0; k < NYH; ++k) {
        if (!oky[k]) continue;
        const int c = tid + k * NT;
        const int r = c / TX, cx = c - r * TX;
        const int ly = r < R ? r : r + TY;
        const Derived<T> d = derived_of<T, 2>(g, ry[k]);
        sm.K[ly][cx + R] = d.K;
        sm.Gyd[ly][cx] = d.Gyd;
        sm.Gyc[ly][cx] = d.Gyc;
    }
    // the output point's Coriolis parameter and RK vectors: issued after the load phase (issued
    // with it, they raised its register peak from 74 to 88 VGPRs at FP64, one wave per SIMD
    // less), in flight across the barrier and the stencil phase
    if (INNER || (i > 0 && j > 0 && i < n - 1 && j < n - 1)) pre = prefetch_point<T, MODE>(g, a, pidx(g, p, i, j));
    __syncthreads();

    // ---- output phase: the thread's own point ----
    if (!INNER && (i <= 0 || j <= 0 || i >= n - 1 || j >= n - 1)) return;
    const int lx = tx + R, ly = ty + R;
    auto d1 = [&](int t, const auto &get) {
        if constexpr (INNER) return sbp_row_in<ORD, T>(t, get);
        else return sbp_row<ORD, T>(t, n, get);
    };
    const T dxK = d1(i, [&](int ai) { return sm.K[ly][ai - tx0 + R]; });
    const T dyK = d1(j, [&](int aj) { return sm.K[aj - ty0 + R][lx]; });
    const T dxD = d1(i, [&](int ai) { return sm.Gxd[ty][ai - tx0 + R]; });
    const T dyD = d1(j, [&](int aj) { return sm.Gyd[aj - ty0 + R][tx]; });
    const T dxC = d1(i, [&](int ai) { return sm.Gxc[ty][ai - tx0 + R]; });
    const T dyC = d1(j, [&](int aj) { return sm.Gyc[aj - ty0 + R][tx]; });
    const long long id = pidx(g, p, i, j);
    const T Jd = own.J * g.hx * g.scale;
    const T gx = dxK / g.hxs;
    const T gy = dyK / g.hxs;
    const T dv = (dxD + dyD) / Jd;
    const T cu = (dxC + dyC) / Jd;

    finish_point_hv<T, MODE>(g, a, id, own.h, own.u, own.v, own_hu, own_hv, own.J, gx, gy, dv, cu, pre);
}

// Edge segment `seg` of the cube (edge_segs(n, PERSIST) per cube): positions k0..k1-1 of one
// panel edge, at most edge_seg_len(n, PERSIST) points.  The own panel's and the neighbour's stencil inputs along the edge go to LDS; each
// point then forms both panels' raw values, applies the scalar and covariant vector edge syncs
// (ecs_Ah_vec_sync_mod.f90:77-124 with (U, V) = (gy, gx), halo_Ah_scalar_sync_mod.f90:52-93,
// in the reference's operand order) and the RK stage update.
template <int ORD, typename T, int MODE, bool OTF, bool PERSIST = false>
__device__ __forceinline__ void stage_edge(const DevGeo<T> &g, const StageArgs<T> &a, int seg, EdgeSmem<ORD, T> &sm,
                                           int nthreads) {
    constexpr int R = Sbp<ORD>::REACH, D = Sbp<ORD>::last[0];
    const int n = g.n, S = edge_seg_len(n, PERSIST);
    int tid = threadIdx.x;
    if constexpr (PERSIST) asm volatile("" : "+v"(tid));
    const int nsx = edge_segs_x(n, PERSIST), nsy = edge_segs_y(n, PERSIST), per = 2 * (nsx + nsy);
    const int p = seg / per, r = seg - p * per;
    int e, si;
    if (r < 2 * nsx) { e = r / nsx; si = r - e * nsx; }
    else { e = 2 + (r - 2 * nsx) / nsy; si = r - 2 * nsx - (e - 2) * nsy; }
    const int k0 = e < 2 ? si * S : 1 + si * S;
    const int k1 = e < 2 ? min(n, k0 + S) : min(n - 1, k0 + S);
    // side 0: the own panel in neighbour-table form (identity orientation); side 1: across e
    const EdgeDesc own{p, e < 2 ? 1 : 0, (e == 0 || e == 2) ? 0 : n - 1, 0, 1, 0};
    const EdgeDesc nbr = edge_desc(g.einfo + (p * 4 + e) * 8);
    // cube corners of the x-edges: the neighbour across the y-edge at i = 0 / n-1
    const bool c0 = e < 2 && k0 == 0, c1 = e < 2 && k1 == n;
    const int ci = e == 0 ? 0 : n - 1;
    const T *y = a.y;

    const int k = k0 + tid;
    const bool act = tid < k1 - k0;

    // ---- strips: per side S + 2R positions at depth 0, S positions at depths 1..D-1; at a
    // corner 2R+1 positions at depth 0 and the corner position at depths 1..D-1 ----
    const int N0 = S + 2 * R, NS = N0 + (D - 1) * S;
    constexpr int NC = 2 * R + 1 + (D - 1);
    const int total = 2 * NS + 2 * NC;
    for (int c = tid; c < total; c += nthreads) {
        EdgeDesc ei;
        int d, kk, slot, side, idx;
        if (c < 2 * NS) {
            side = c / NS;
            const int q = c - side * NS;
            ei = side == 0 ? own : nbr;
            slot = -1;
            if (q < N0) { d = 0; idx = q; kk = k0 - R + q; }
            else { d = 1 + (q - N0) / S; idx = (q - N0) % S; kk = k0 + idx; if (kk >= k1) continue; }
        } else {
            const int q = c - 2 * NS;
            slot = q / NC;
            if (!(slot == 0 ? c0 : c1)) continue;
            const int w = q - slot * NC;
            ei = edge_desc(g.einfo + (p * 4 + (slot == 0 ? 2 : 3)) * 8);
            side = 0;
            if (w < 2 * R + 1) { d = 0; idx = w; kk = ci - R + w; }
            else { d = 1 + (w - (2 * R + 1)); idx = 0; kk = ci; }
        }
        if (kk < 0 || kk >= n) continue;
        const int nidx = ei.tnu == 0 ? d : n - 1 - d, tidx = ei.a_base + ei.s * kk;
        const int qi = ei.nu_x ? nidx : tidx, qj = ei.nu_x ? tidx : nidx;
        const Derived<T> dv = derived_global<OTF>(g, y, a.hs, ei.q, qi, qj);
        if (d == 0) {
            T *dst = slot < 0 ? sm.e0[side][idx] : sm.c0[slot][idx];
            dst[0] = dv.K; dst[1] = dv.Gxd; dst[2] = dv.Gyd; dst[3] = dv.Gxc; dst[4] = dv.Gyc;
            dst[5] = metric_J<OTF>(g, ei.q, qi, qj);
        } else {
            const bool nu_x = ei.nu_x != 0;
            T *dst = slot < 0 ? sm.en[side][d - 1][idx] : sm.cn[slot][d - 1];
            dst[0] = dv.K;
            dst[1] = nu_x ? dv.Gxd : dv.Gyd;
            dst[2] = nu_x ? dv.Gxc : dv.Gyc;
        }
    }
    __syncthreads();
    if (!act) return;
    int i, j;
    if (e < 2) { i = e == 0 ? 0 : n - 1; j = k; }
    else { i = k; j = e == 2 ? 0 : n - 1; }
    const long long id = pidx(g, p, i, j);
    const OutPre<T> pre = prefetch_point<T, MODE>(g, a, id);

    T ro[4], rn[4];
    neighbour_raw<ORD, T>(g, own, k, [&](int kk, int f) { return sm.e0[0][kk - (k0 - R)][f]; },
                          [&](int d, int f3) { return sm.en[0][d - 1][k - k0][f3]; }, ro);
    neighbour_raw<ORD, T>(g, nbr, k, [&](int kk, int f) { return sm.e0[1][kk - (k0 - R)][f]; },
                          [&](int d, int f3) { return sm.en[1][d - 1][k - k0][f3]; }, rn);
    T gx = ro[0], gy = ro[1], dv = ro[2], cu = ro[3];
    const T U = gy, V = gx;
    // the neighbour's raw values mapped to this panel's components (vector halo convention)
    T Vn, Un;
    map_vec(nbr.code, rn[0], rn[1], Vn, Un);
    const T Dn = rn[2], Cn = rn[3];
    const bool corner = e < 2 && (k == 0 || k == n - 1);
    if (corner) {
        const int slot = k == 0 ? 0 : 1;
        const EdgeDesc ec = edge_desc(g.einfo + (p * 4 + (slot == 0 ? 2 : 3)) * 8);
        T rc[4];
        neighbour_raw<ORD, T>(g, ec, ci, [&](int kk, int f) { return sm.c0[slot][kk - (ci - R)][f]; },
                              [&](int d, int f3) { return sm.cn[slot][d - 1][f3]; }, rc);
        T Vy, Uy;
        map_vec(ec.code, rc[0], rc[1], Vy, Uy);
        const T qy = slot == 0 ? g.qb[i] : g.qt[i];
        const T qx = e == 0 ? g.ql[j] : g.qr[j];
        dv = (dv + Dn + rc[2]) / T(3);
        cu = (cu + Cn + rc[3]) / T(3);
        gy = (U + Un + Uy + qy * Vy) / T(3);
        gx = (V + Vn + Vy + qx * Un) / T(3);
    } else if (e == 0) {
        dv = T(0.5) * (Dn + dv);
        cu = T(0.5) * (Cn + cu);
        gy = T(0.5) * (Un + U);
        gx = T(0.5) * (V + Vn + g.ql[j] * Un);
    } else if (e == 1) {
        dv = T(0.5) * (dv + Dn);
        cu = T(0.5) * (cu + Cn);
        gy = T(0.5) * (U + Un);
        gx = T(0.5) * (V + Vn + g.qr[j] * Un);
    } else if (e == 2) {
        dv = T(0.5) * (dv + Dn);
        cu = T(0.5) * (cu + Cn);
        gx = T(0.5) * (Vn + V);
        gy = T(0.5) * (U + Un + g.qb[i] * Vn);
    } else {
        dv = T(0.5) * (dv + Dn);
        cu = T(0.5) * (cu + Cn);
        gx = T(0.5) * (V + Vn);
        gy = T(0.5) * (U + Un + g.qt[i] * Vn);
    }
    T Jv, q11, q12, q22;
    if constexpr (OTF) {
        const Metric m = metric_otf(g.tanA, i, j);
        Jv = (T)m.J; q11 = (T)m.Qi11; q12 = (T)m.Qi12; q22 = (T)m.Qi22;
    } else {
        const long long mid = midx(g, p, i, j);
        Jv = g.J[mid]; q11 = g.Qi11[mid]; q12 = g.Qi12[mid]; q22 = g.Qi22[mid];
    }
    finish_point<T, MODE>(g, a, id, y[id], y[id + g.fs], y[id + 2 * g.fs], Jv, q11, q12, q22, gx, gy, dv, cu, pre);
}

// a tile starting at (tx0, ty0) whose derivative rows are all interior ones (NE <= i, j < n - NE)
// and whose halo of REACH points lies inside the panel: M <= i, j < n - M with M the larger
template <int ORD, int TX, int TY>
__device__ __forceinline__ bool inner_tile(int tx0, int ty0, int n) {
    constexpr int NE = Sbp<ORD>::NE, RE = Sbp<ORD>::REACH, M = NE > RE ? NE : RE;
    return tx0 >= M && tx0 + TX <= n - M && ty0 >= M && ty0 + TY <= n - M;
}

// edge blocks first (a multiple of 8, so that the tiles keep their XCD-contiguous mapping)
SWM_HD constexpr int edge_blocks(int n) { return (edge_segs(n) + 7) / 8 * 8; }

template <int ORD, typename T, int MODE, int TX, int TY, bool OTF>
__global__ void __launch_bounds__(TX * TY) stage_kernel(DevGeo<T> g, StageArgs<T> a) {
    __shared__ Smem<ORD, T, TX, TY> sm;
    const int ne = edge_blocks(g.n);
    if ((int)blockIdx.x < ne) {
        if ((int)blockIdx.x < edge_segs(g.n)) stage_edge<ORD, T, MODE, OTF>(g, a, blockIdx.x, sm.e, TX * TY);
        return;
    }
    const int ntx = tiles_along(g.n, TX), nty = tiles_along(g.n, TY);
    const int tile = xcd_tile(blockIdx.x - ne, 6 * ntx * nty);
    const int p = tile / (ntx * nty);
    const int rem = tile - p * ntx * nty;
    const int tx0 = 1 + (rem % ntx) * TX, ty0 = 1 + (rem / ntx) * TY;
    if (inner_tile<ORD, TX, TY>(tx0, ty0, g.n))
        stage_tile<ORD, T, MODE, TX, TY, OTF, false, true>(g, a, p, tx0, ty0, sm.t);
    else
        stage_tile<ORD, T, MODE, TX, TY, OTF>(g, a, p, tx0, ty0, sm.t);
}

// Copy of a block-uniform struct into scalar registers (word by word through readfirstlane):
// the stores of the stage body keep the compiler from proving the table unchanged, which it
// would need to use scalar loads by itself.
template <typename S>
__device__ __forceinline__ S uniform_copy(const S *src) {
    static_assert(sizeof(S) % 4 == 0, "word-sized struct");
    S out;
    const unsigned *w = reinterpret_cast<const unsigned *>(src);
    unsigned *o = reinterpret_cast<unsigned *>(&out);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(S) / 4); ++k) o[k] = __builtin_amdgcn_readfirstlane(w[k]);
    return out;
}

// Ensemble of cubes of different sizes in one launch: the blocks walk the concatenated task list
// of all members (per member its 6*ntx*nty tiles (tiles_along), then its edge_segs(n, true) edge
// segments), XCD-contiguous; a persistent grid loops over it, a grid of one block per task runs
// each block once.  tasks[t] = (member, task index within the
// member): one load gives a block its member, whose geometry and stage arguments it then takes
// from the member table.
template <int ORD, typename T, int MODE, int TX, int TY, bool OTF>
__global__ void __launch_bounds__(TX * TY) ens_stage_kernel(const EnsMember<T> *__restrict__ members,
                                                            const int2 *__restrict__ tasks, int ntiles) {
    __shared__ Smem<ORD, T, TX, TY> sm;
    for (int it = blockIdx.x; it < ntiles; it += gridDim.x) {
        const int tile = xcd_tile(it, ntiles);
        const int2 tk = tasks[tile];
        const int m = __builtin_amdgcn_readfirstlane(tk.x);   // block-uniform member index
        const int lt = __builtin_amdgcn_readfirstlane(tk.y);
        const DevGeo<T> g = uniform_copy(&members[m].g);
        const StageArgs<T> a = uniform_copy(&members[m].a);
        const int ntx = tiles_along(g.n, TX), nty = tiles_along(g.n, TY);
        __syncthreads();  // the previous task's LDS reads are done
        if (lt >= 6 * ntx * nty) {
            stage_edge<ORD, T, MODE, OTF, true>(g, a, lt - 6 * ntx * nty, sm.e, TX * TY);
            continue;
        }
        const int p = lt / (ntx * nty);
        const int rem = lt - p * ntx * nty;
        const int tx0 = 1 + (rem % ntx) * TX, ty0 = 1 + (rem / ntx) * TY;
        if (inner_tile<ORD, TX, TY>(tx0, ty0, g.n))
            stage_tile<ORD, T, MODE, TX, TY, OTF, true, true>(g, a, p, tx0, ty0, sm.t);
        else
            stage_tile<ORD, T, MODE, TX, TY, OTF, true>(g, a, p, tx0, ty0, sm.t);
    }
}

}  // namespace swm

namespace swm {

template <int ORD, typename T, int TX, int TY, bool OTF>
hipError_t launch_one(int mode, const DevGeo<T> &g, const StageArgs<T> &a, hipStream_t s) {
    const int ntx = tiles_along(g.n, TX), nty = tiles_along(g.n, TY);
    const dim3 grid(edge_blocks(g.n) + 6 * ntx * nty), block(TX * TY);
    switch (mode) {
    case MODE_RHS: hipLaunchKernelGGL((stage_kernel<ORD, T, MODE_RHS, TX, TY, OTF>), grid, block, 0, s, g, a); break;
    case MODE_STAGE1: hipLaunchKernelGGL((stage_kernel<ORD, T, MODE_STAGE1, TX, TY, OTF>), grid, block, 0, s, g, a); break;
    case MODE_STAGE23: hipLaunchKernelGGL((stage_kernel<ORD, T, MODE_STAGE23, TX, TY, OTF>), grid, block, 0, s, g, a); break;
    case MODE_STAGE4: hipLaunchKernelGGL((stage_kernel<ORD, T, MODE_STAGE4, TX, TY, OTF>), grid, block, 0, s, g, a); break;
    case MODE_CURL: hipLaunchKernelGGL((stage_kernel<ORD, T, MODE_CURL, TX, TY, OTF>), grid, block, 0, s, g, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int ORD, typename T, int TX, int TY, bool OTF, int MODE>
hipError_t launch_ens_mode(const EnsLaunch &r) {
    auto kern = ens_stage_kernel<ORD, T, MODE, TX, TY, OTF>;
    static int per_cu = 0;  // resident blocks per CU of this instantiation
    if (per_cu == 0) {
        int nb = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, TX * TY, 0);
        if (e != hipSuccess) return e;
        per_cu = nb > 0 ? nb : 1;
    }
    // resident blocks x the requested waves of blocks (0: one block per task); a multiple of 8
    // keeps a block's tiles on its XCD
    long long grid = r.blocks_per_slot > 0 ? (long long)r.num_cu * per_cu * r.blocks_per_slot : (long long)r.ntiles;
    grid = (grid + 7) / 8 * 8;
    if (grid > r.ntiles) grid = (r.ntiles + 7) / 8 * 8;
    if (grid < 8) grid = 8;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(TX * TY), 0, r.stream, (const EnsMember<T> *)r.members,
                       r.tasks, r.ntiles);
    return hipGetLastError();
}

template <int ORD, typename T, int TX, int TY, bool OTF>
hipError_t launch_ens_one(const EnsLaunch &r) {
    switch (r.mode) {
    case MODE_STAGE1: return launch_ens_mode<ORD, T, TX, TY, OTF, MODE_STAGE1>(r);
    case MODE_STAGE23: return launch_ens_mode<ORD, T, TX, TY, OTF, MODE_STAGE23>(r);
    case MODE_STAGE4: return launch_ens_mode<ORD, T, TX, TY, OTF, MODE_STAGE4>(r);
    default: return hipErrorInvalidValue;
    }
}

// the tile shapes of the stage kernels (tile_shape_ok): 64 x 8, 32 x 16, 16 x 32 (512 threads)
// and 32 x 8 (256 threads)
template <int ORD, typename T, bool OTF>
hipError_t launch_ens_shape(const EnsLaunch &r) {
    if (r.tx == 64 && r.ty == 8) return launch_ens_one<ORD, T, 64, 8, OTF>(r);
    if (r.tx == 32 && r.ty == 16) return launch_ens_one<ORD, T, 32, 16, OTF>(r);
    if (r.tx == 16 && r.ty == 32) return launch_ens_one<ORD, T, 16, 32, OTF>(r);
    if (r.tx == 32 && r.ty == 8) return launch_ens_one<ORD, T, 32, 8, OTF>(r);
    return hipErrorInvalidValue;
}

template <int ORD>
hipError_t launch_ens_f64(const EnsLaunch &r) {
    return r.otf ? launch_ens_shape<ORD, double, true>(r) : launch_ens_shape<ORD, double, false>(r);
}

template <int ORD>
hipError_t launch_ens_f32(const EnsLaunch &r) {
    return launch_ens_shape<ORD, float, false>(r);
}

template <int ORD, typename T, bool OTF>
hipError_t launch_stage_shape(const StageLaunch &r, const DevGeo<T> &g, const StageArgs<T> &a) {
    if (r.tx == 64 && r.ty == 8) return launch_one<ORD, T, 64, 8, OTF>(r.mode, g, a, r.stream);
    if (r.tx == 32 && r.ty == 16) return launch_one<ORD, T, 32, 16, OTF>(r.mode, g, a, r.stream);
    if (r.tx == 16 && r.ty == 32) return launch_one<ORD, T, 16, 32, OTF>(r.mode, g, a, r.stream);
    if (r.tx == 32 && r.ty == 8) return launch_one<ORD, T, 32, 8, OTF>(r.mode, g, a, r.stream);
    return hipErrorInvalidValue;
}

template <int ORD>
hipError_t launch_stage_f64(const StageLaunch &r) {
    return r.otf ? launch_stage_shape<ORD, double, true>(r, r.gd, r.ad) : launch_stage_shape<ORD, double, false>(r, r.gd, r.ad);
}

template <int ORD>
hipError_t launch_stage_f32(const StageLaunch &r) {
    return launch_stage_shape<ORD, float, false>(r, r.gf, r.af);
}

}  // namespace swm
