// Hyperdiffusion of the Ah state with the narrow-stencil SBP Laplacians sbp21, sbp42 and sbp63,
// FP64 (included into swm_capi.hip).  sbp21 (three-point rows, laplace_Ah_sbp21_narrow_mod.f90)
// runs as one point per thread (hd_lap21); the table-driven sbp42 / sbp63 on LDS tiles.
//
// Laplacian, laplace_Ah_sbp42_narrow_t%calc_laplace (differential_operators/horizontal/
// laplace/laplace_Ah_sbp42_narrow_mod.f90:23-763) and laplace_Ah_sbp63_narrow_t%calc_laplace
// (laplace_Ah_sbp63_narrow_mod.f90:494-1717), the same steps with their own tables:
//   d1x = Dx(f) J Qi12,  d1y = Dy(f) J Qi12                      (first-derivative SBP d42 / d63)
//   d2x = D2x(Qi11 J) f, d2y = D2y(Qi22 J) f                     (narrow variable-coefficient D2)
//   d2x -= 0.5 * pc * (-+2 d1y) on the x edges, then (-+2 d1x) on the y edges
//        (pc = 48/17 for sbp42, 43200/13649 for sbp63)
//   L f = ((d2x + d2y) + Dx(d1y) + Dy(d1x)) / (hx a)^2 / J, then the scalar edge sync.
// Hyperdiffusion of one field, hordiff_scalar_t%calc_diff (hordiff/hordiff_scalar_mod.f90:
// 34-59): tend = -(c a hx)^4 L(L f).  Velocity, hordiff_colocated_xyz_t%calc_diff_vec
// (hordiff/hordiff_colocated_mod.f90:39-58): Cartesian components V = b1 u + b2 v, the scalar
// hyperdiffusion of each, back as (V.b1, V.b2), covariant vector edge sync.
// Step, operator_swm_diff_t%apply + explicit_Eul1 (operator_swm_diff_mod.f90:25-46,
// time_schemes/explicit_Eul1_mod.f90:20-30): y += dt * tend.
//
// One Laplacian is two launches, hd_lap_tile (one flat grid over tiles x fields) and hd_sync
// (fields on blockIdx.y), over several fields:
// h and the three Cartesian velocity components share every launch.  Panel-edge results are
// written to a compact edge buffer first and synchronised from there, so that no kernel reads
// a value another block of the same launch overwrites.

#include "laplace_sbp42_narrow.h"
#include "laplace_sbp63_narrow.h"

namespace swm {

// The generated coefficient tables of one narrow Laplacian family, bound by name.
template <int ORD> struct Ln;
template <> struct Ln<42> {
    static constexpr int NB = LN42_NB, NCOL = LN42_NCOL, R = LN42_R;
    static constexpr bool NEG_INV_H = LN42_ROW_NEG_INV_H != 0;
    static constexpr auto &H = LN42_H;
    static constexpr auto &NCOLS = LN42_NCOLS;
    static constexpr auto &MSTART = LN42_MSTART;
    static constexpr auto &TCOEF = LN42_TCOEF;
    static constexpr auto &TSIGN = LN42_TSIGN;
    static constexpr auto &TB = LN42_TB;
    static constexpr auto &ISTART = LN42_ISTART;
    static constexpr auto &ICOEF = LN42_ICOEF;
    static constexpr auto &ISIGN = LN42_ISIGN;
    static constexpr auto &IB = LN42_IB;
    static constexpr double PC = LN42_PC;
};
template <> struct Ln<63> {
    static constexpr int NB = LN63_NB, NCOL = LN63_NCOL, R = LN63_R;
    static constexpr bool NEG_INV_H = LN63_ROW_NEG_INV_H != 0;
    static constexpr auto &H = LN63_H;
    static constexpr auto &NCOLS = LN63_NCOLS;
    static constexpr auto &MSTART = LN63_MSTART;
    static constexpr auto &TCOEF = LN63_TCOEF;
    static constexpr auto &TSIGN = LN63_TSIGN;
    static constexpr auto &TB = LN63_TB;
    static constexpr auto &ISTART = LN63_ISTART;
    static constexpr auto &ICOEF = LN63_ICOEF;
    static constexpr auto &ISIGN = LN63_ISIGN;
    static constexpr auto &IB = LN63_IB;
    static constexpr double PC = LN63_PC;
};

struct HdGeo {
    const double *J, *Qi11, *Qi12, *Qi22;
    const double *qb, *qt, *ql, *qr;
    const int *einfo;
    const double *b1, *b2;   // 3 planes each
    int n, pitch, es;
    long long fs;
    double hs;               // hx * scale
};

__device__ __forceinline__ void slot_point(int n, int sl, int &i, int &j) {
    if (sl < n) { i = sl; j = 0; }
    else if (sl < 2 * n) { i = sl - n; j = n - 1; }
    else if (sl < 3 * n - 2) { i = 0; j = sl - 2 * n + 1; }
    else { i = n - 1; j = sl - (3 * n - 2) + 1; }
}

// slot (global over panels) of the neighbour's duplicate of point k of edge e of panel p
__device__ __forceinline__ int nbr_slot(const HdGeo &g, int p, int e, int k) {
    const int *ei = g.einfo + (p * 4 + e) * 8;
    const int n = g.n;
    const int nidx = ei[2] == 0 ? 0 : n - 1, tidx = ei[3] + ei[4] * k;
    const int qi = ei[1] ? nidx : tidx, qj = ei[1] ? tidx : nidx;
    return ei[0] * g.es + edge_slot(n, qi, qj);
}

// Closure row RW of the narrow D2 with b[k], fv[c] taken k, c points from the panel edge, and
// the interior row, expanded at compile time (every index and weight a constant): each weight
// is the table's constant expression folded to a double, products and sums in table order.
template <int ORD, int T>
__device__ __forceinline__ double ln_term(const double *b) {
    constexpr double w = Ln<ORD>::TCOEF[T];
    return w * b[Ln<ORD>::TB[T]];
}
template <int ORD, int T0, int... Is>
__device__ __forceinline__ double ln_sum(const double *b, std::integer_sequence<int, Is...>) {
    using L = Ln<ORD>;
    double m = 0.0;
    ((m = Is == 0 ? (L::TSIGN[T0 + Is] < 0 ? -ln_term<ORD, T0 + Is>(b) : ln_term<ORD, T0 + Is>(b))
                  : (L::TSIGN[T0 + Is] < 0 ? m - ln_term<ORD, T0 + Is>(b) : m + ln_term<ORD, T0 + Is>(b))),
     ...);
    return m;
}
template <int ORD, int RW, int C>
__device__ __forceinline__ double ln_m(const double *b) {
    using L = Ln<ORD>;
    constexpr int t0 = L::MSTART[RW][C], cnt = L::MSTART[RW][C + 1] - L::MSTART[RW][C];
    return ln_sum<ORD, t0>(b, std::make_integer_sequence<int, cnt>{});
}
template <int ORD, int RW, int... Cs>
__device__ __forceinline__ double ln_row_cols(const double *b, const double *fv, std::integer_sequence<int, Cs...>) {
    double s = 0.0;
    ((s = Cs == 0 ? ln_m<ORD, RW, Cs>(b) * fv[Cs] : s + ln_m<ORD, RW, Cs>(b) * fv[Cs]), ...);
    if constexpr (Ln<ORD>::NEG_INV_H) {
        constexpr double h = Ln<ORD>::H[RW];
        return -((1.0 / h) * s);   // -1/h_rr * (...)
    } else {
        return s;                  // 1.0 * (...), exact
    }
}
template <int ORD, int RW>
__device__ __forceinline__ double ln_row(const double *b, const double *fv) {
    return ln_row_cols<ORD, RW>(b, fv, std::make_integer_sequence<int, Ln<ORD>::NCOLS[RW]>{});
}
template <int ORD, int... RWs>
__device__ __forceinline__ double ln_row_any(int r, const double *b, const double *fv,
                                             std::integer_sequence<int, RWs...>) {
    double out = 0.0;
    ((r == RWs ? (void)(out = ln_row<ORD, RWs>(b, fv)) : (void)0), ...);
    return out;
}

template <int ORD, int T>
__device__ __forceinline__ double ln_iterm(const double *b) {
    constexpr double w = Ln<ORD>::ICOEF[T];
    return w * b[Ln<ORD>::IB[T] + Ln<ORD>::R];
}
template <int ORD, int T0, int... Is>
__device__ __forceinline__ double ln_isum(const double *b, std::integer_sequence<int, Is...>) {
    using L = Ln<ORD>;
    double m = 0.0;
    ((m = Is == 0 ? (L::ISIGN[T0 + Is] < 0 ? -ln_iterm<ORD, T0 + Is>(b) : ln_iterm<ORD, T0 + Is>(b))
                  : (L::ISIGN[T0 + Is] < 0 ? m - ln_iterm<ORD, T0 + Is>(b) : m + ln_iterm<ORD, T0 + Is>(b))),
     ...);
    return m;
}
template <int ORD, int O>
__device__ __forceinline__ double ln_icoef(const double *b) {
    using L = Ln<ORD>;
    return ln_isum<ORD, L::ISTART[O]>(b, std::make_integer_sequence<int, L::ISTART[O + 1] - L::ISTART[O]>{});
}
template <int ORD, int... Os>
__device__ __forceinline__ double ln_interior_seq(const double *b, const double *fv, std::integer_sequence<int, Os...>) {
    double s = 0.0;
    ((s = Os == 0 ? ln_icoef<ORD, Os>(b) * fv[Os] : s + ln_icoef<ORD, Os>(b) * fv[Os]), ...);
    return s;
}
template <int ORD>
__device__ __forceinline__ double ln_interior(const double *b, const double *fv) {
    return ln_interior_seq<ORD>(b, fv, std::make_integer_sequence<int, 2 * Ln<ORD>::R + 1>{});
}

// V = b1 u + b2 v (vec_math_mod.f90:107-126); W[0] = h, W[1..3] = Vx, Vy, Vz
__global__ void hd_to_cart(HdGeo g, const double *__restrict__ y, double *__restrict__ W) {
    const int n = g.n;
    const long long total = 6LL * n * n;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int i = (int)(t % n);
    const long long pj = t / n;
    const long long o = pj * g.pitch + i;
    const double u = y[g.fs + o], v = y[2 * g.fs + o];
    W[o] = 1.0 * y[o];
#pragma unroll
    for (int c = 0; c < 3; ++c) W[(1 + c) * g.fs + o] = g.b1[c * g.fs + o] * u + g.b2[c * g.fs + o] * v;
}

// The Laplacian of field k = blockIdx.y on LDS tiles of TX x TY points, before the edge sync:
// off-edge points to Wout, panel-edge points to E (ns slots per field).  A tile loads f on the
// box of lines its rows read (+-R points in the interior, points 0..NR-1 / n-NR..n-1 where it
// holds closure rows: the narrow D2 closure reads NR = 8 (sbp42) / 12 (sbp63) points, the
// first-derivative closure 6 / 9), and b = Qi11 J on the tile rows, b = Qi22 J on the tile
// columns.  It then forms the first derivatives d1y = Dy(f) J Qi12 on the tile rows and
// d1x = Dx(f) J Qi12 on the tile columns (each at the positions whose J and Qi12 the same
// thread loaded), and finally the D2 rows, the cross-term penalties and the cross derivatives at
// its points.
template <int ORD, int TX, int TY>
struct HdTile {
    static constexpr int R = Ln<ORD>::R, B = Ln<ORD>::NB, NR = Ln<ORD>::NCOL;
    // the first-derivative operator of the family reads within the same box
    static_assert(-Sbp<ORD>::INSHIFT <= R && Sbp<ORD>::WIN + Sbp<ORD>::INSHIFT - 1 <= R, "d1 reach");
    static_assert(Sbp<ORD>::NE <= B && Sbp<ORD>::NW <= NR, "d1 closure");
    // box extent along a tile side of t points: closure rows at one edge give at most
    // t + R + B - 1 lines (or NR); a tile holding closure rows of both edges (a0 < B, a0 a
    // multiple of t, and a0 + t > n - B) covers the whole line, n <= a0 + t + B - 1
    static constexpr int both(int t) { return ((B - 1) / t) * t + t + B - 1; }
    static constexpr int one(int t) { return t + R + B - 1 > NR ? t + R + B - 1 : NR; }
    static constexpr int span(int t) { return one(t) > both(t) ? one(t) : both(t); }
    static constexpr int LW = span(TX), LH = span(TY), NT = TX * TY;
};

template <int ORD>
__device__ __forceinline__ void hd_rows(int a0, int ae, int n, int &lo, int &hi) {
    constexpr int R = Ln<ORD>::R, B = Ln<ORD>::NB, NR = Ln<ORD>::NCOL;
    lo = a0 - R > 0 ? a0 - R : 0;
    hi = ae + R < n ? ae + R : n;
    if (a0 < B) {
        lo = 0;
        const int m = NR < n ? NR : n;
        hi = hi > m ? hi : m;
    }
    if (ae > n - B) {
        hi = n;
        const int m = n - NR > 0 ? n - NR : 0;
        lo = lo < m ? lo : m;
    }
}

// One tile of hd_lap_tile.  INNER: the tile holds no closure row and no panel-edge point in
// either direction (B <= x0, x0 + TX <= n - B, the same in y), so its box is the tile +-R,
// every row of it is an interior row and its guards and branches are known at compile time.
template <int ORD, int TX, int TY, bool INNER>
__device__ __forceinline__ void hd_lap_run(const HdGeo &g, const double *__restrict__ W, double *__restrict__ Wout,
                                           double *__restrict__ E, int k, int p, int x0, int y0,
                                           double (&sf)[HdTile<ORD, TX, TY>::LH][HdTile<ORD, TX, TY>::LW],
                                           double (&sbx)[TY][HdTile<ORD, TX, TY>::LW],
                                           double (&sdy)[TY][HdTile<ORD, TX, TY>::LW],
                                           double (&sby)[HdTile<ORD, TX, TY>::LH][TX],
                                           double (&sdx)[HdTile<ORD, TX, TY>::LH][TX]) {
    using C = HdTile<ORD, TX, TY>;
    using L = Ln<ORD>;
    const int n = g.n;
    const int xe = x0 + TX < n ? x0 + TX : n, ye = y0 + TY < n ? y0 + TY : n;
    int lx, hx, ly, hy;
    if constexpr (INNER) {
        lx = x0 - C::R; hx = x0 + TX + C::R;
        ly = y0 - C::R; hy = y0 + TY + C::R;
    } else {
        hd_rows<ORD>(x0, xe, n, lx, hx);
        hd_rows<ORD>(y0, ye, n, ly, hy);
    }
    const long long pb = (long long)p * n;
    const int pitch = g.pitch;
    // panel bases; offsets within a panel are 32-bit (the 24-bit multiply: rows, pitch < 2^24)
    const double *fb = W + k * g.fs + pb * pitch;
    const double *Jb = g.J + pb * pitch, *Q11b = g.Qi11 + pb * pitch, *Q12b = g.Qi12 + pb * pitch,
                 *Q22b = g.Qi22 + pb * pitch;
    // box extents (compile-time in an inner tile)
    constexpr int WAI = TX + 2 * C::R, HBI = TY + 2 * C::R;
    const int wa = INNER ? WAI : hx - lx, hb = INNER ? HBI : hy - ly;
    const int nc = INNER ? TX : xe - x0, nr = INNER ? TY : ye - y0;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    constexpr int RB = (C::LH + TY - 1) / TY, CB = (C::LW + TX - 1) / TX;
    // row ty + a*TY / column tx + b*TX inside the box: always so in an inner tile when the whole
    // pass fits (after unrolling these are constants)
    auto row_in = [&](int a) { return (INNER && (a + 1) * TY <= HBI) || ty + a * TY < hb; };
    auto col_in = [&](int b) { return (INNER && (b + 1) * TX <= WAI) || tx + b * TX < wa; };
    const bool own_r = INNER || ty < nr, own_c = INNER || tx < nc;

    // ---- loads, each thread on rows ty + a*TY and columns tx + b*TX of its ranges: f on the box;
    // Qi, J, Qi12 at the cross positions (tile rows x line range: kind 1, line range x tile
    // columns: kind 2) ----
    double fv[RB][CB];
#pragma unroll
    for (int a = 0; a < RB; ++a)
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int r = ty + a * TY, c = tx + b * TX;
            if (row_in(a) && col_in(b)) fv[a][b] = fb[__umul24(ly + r, pitch) + lx + c];
        }
    double q1v[CB], j1v[CB], q121[CB];   // kind 1: row y0 + ty, column lx + tx + b*TX
#pragma unroll
    for (int b = 0; b < CB; ++b) {
        const int c = tx + b * TX;
        if (own_r && col_in(b)) {
            const int o = __umul24(y0 + ty, pitch) + lx + c;
            q1v[b] = Q11b[o];
            j1v[b] = Jb[o];
            q121[b] = Q12b[o];
        }
    }
    double q2v[RB], j2v[RB], q122[RB];   // kind 2: row ly + ty + a*TY, column x0 + tx
#pragma unroll
    for (int a = 0; a < RB; ++a) {
        const int r = ty + a * TY;
        if (row_in(a) && own_c) {
            const int o = __umul24(ly + r, pitch) + x0 + tx;
            q2v[a] = Q22b[o];
            j2v[a] = Jb[o];
            q122[a] = Q12b[o];
        }
    }
#pragma unroll
    for (int a = 0; a < RB; ++a)
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int r = ty + a * TY, c = tx + b * TX;
            if (row_in(a) && col_in(b)) sf[r][c] = fv[a][b];
        }
#pragma unroll
    for (int b = 0; b < CB; ++b)
        if (own_r && col_in(b)) sbx[ty][tx + b * TX] = q1v[b] * j1v[b];
#pragma unroll
    for (int a = 0; a < RB; ++a)
        if (row_in(a) && own_c) sby[ty + a * TY][tx] = q2v[a] * j2v[a];
    __syncthreads();

    // ---- first derivatives at the cross positions (calc_tile_d1x_d1y: sbp42 :60-93,
    // sbp63 :531-564) ----
    auto d1 = [&](int t, const auto &get) {
        if constexpr (INNER) return sbp_row_in<ORD, double>(t, get);
        else return sbp_row<ORD, double>(t, n, get);
    };
#pragma unroll
    for (int b = 0; b < CB; ++b) {
        const int c = tx + b * TX;
        if (own_r && col_in(b)) {
            const double d = d1(y0 + ty, [&](int a) { return sf[a - ly][c]; });
            sdy[ty][c] = d * j1v[b] * q121[b];
        }
    }
#pragma unroll
    for (int a = 0; a < RB; ++a) {
        const int r = ty + a * TY;
        if (row_in(a) && own_c) {
            const double d = d1(x0 + tx, [&](int b) { return sf[r][b - lx]; });
            sdx[r][tx] = d * j2v[a] * q122[a];
        }
    }
    __syncthreads();

    const int i = x0 + tx, j = y0 + ty;
    if (!INNER && (i >= n || j >= n)) return;
    const int o = __umul24(j, pitch) + i;
    auto d2 = [&](int t, const auto &bget, const auto &fget) {
        double bb[L::NCOL], ff[L::NCOL];
        if (!INNER && (t < L::NB || t >= n - L::NB)) {
            const bool left = t < L::NB;
#pragma unroll
            for (int q = 0; q < L::NCOL; ++q) {
                const int a = left ? q : n - 1 - q;
                bb[q] = bget(a);
                ff[q] = fget(a);
            }
            return ln_row_any<ORD>(left ? t : n - 1 - t, bb, ff, std::make_integer_sequence<int, L::NB>{});
        }
#pragma unroll
        for (int q = 0; q < 2 * L::R + 1; ++q) {
            bb[q] = bget(t - L::R + q);
            ff[q] = fget(t - L::R + q);
        }
        return ln_interior<ORD>(bb, ff);
    };
    double d2x = d2(i, [&](int a) { return sbx[ty][a - lx]; }, [&](int a) { return sf[j - ly][a - lx]; });
    const double d2y = d2(j, [&](int a) { return sby[a - ly][tx]; }, [&](int a) { return sf[a - ly][i - lx]; });
    if constexpr (!INNER) {
        // penalties of the cross terms (sbp42 :716-742, sbp63 :1671-1697), x edges first
        constexpr double pc = L::PC;
        if (i == 0) d2x = d2x - 0.5 * (pc * (-2.0 * sdy[ty][i - lx]));
        if (i == n - 1) d2x = d2x - 0.5 * (pc * (2.0 * sdy[ty][i - lx]));
        if (j == 0) d2x = d2x - 0.5 * (pc * (-2.0 * sdx[j - ly][tx]));
        if (j == n - 1) d2x = d2x - 0.5 * (pc * (2.0 * sdx[j - ly][tx]));
    }
    const double d2yx = d1(j, [&](int a) { return sdx[a - ly][tx]; });
    const double d2xy = d1(i, [&](int a) { return sdy[ty][a - lx]; });
    double v = 1.0 * d2x + 1.0 * d2y;
    v = v + 1.0 * d2xy + 1.0 * d2yx;
    v = v / (g.hs * g.hs);
    v = v / Jb[o];
    if (!INNER && (i == 0 || j == 0 || i == n - 1 || j == n - 1))
        E[(long long)k * 6 * g.es + p * g.es + edge_slot(n, i, j)] = v;
    else
        (Wout + k * g.fs + pb * pitch)[o] = v;
}

template <int ORD, int TX, int TY>
__global__ void __launch_bounds__(TX * TY) hd_lap_tile(HdGeo g, const double *__restrict__ W,
                                                       double *__restrict__ Wout, double *__restrict__ E, int nf) {
    using C = HdTile<ORD, TX, TY>;
    __shared__ double sf[C::LH][C::LW];                  // f on the box
    __shared__ double sbx[TY][C::LW], sdy[TY][C::LW];    // tile rows: Qi11 J, d1y
    __shared__ double sby[C::LH][TX], sdx[C::LH][TX];    // tile columns: Qi22 J, d1x
    const int n = g.n;
    const int ntx = (n + TX - 1) / TX, nty = (n + TY - 1) / TY;
    // one flat block index over (tile, field), field fastest: the nf blocks of a tile sit in the
    // same XCD's contiguous range, so the metric they all read is fetched into one L2
    const int u = xcd_tile(blockIdx.x, 6 * ntx * nty * nf);
    const int tile = u / nf, k = u - tile * nf;
    const int p = tile / (ntx * nty), rem = tile - p * ntx * nty;
    const int x0 = (rem % ntx) * TX, y0 = (rem / ntx) * TY;
    const bool inner = x0 >= C::B && x0 + TX <= n - C::B && y0 >= C::B && y0 + TY <= n - C::B;
    if (inner)
        hd_lap_run<ORD, TX, TY, true>(g, W, Wout, E, k, p, x0, y0, sf, sbx, sdy, sby, sdx);
    else
        hd_lap_run<ORD, TX, TY, false>(g, W, Wout, E, k, p, x0, y0, sf, sbx, sdy, sby, sdx);
}

// The sbp21 narrow Laplacian (laplace_Ah_sbp21_narrow_mod.f90:22-245) of field k = blockIdx.y,
// one thread per point, before the edge sync (the output convention of hd_lap_tile).  Three-point
// rows: D2(b) with b = Qi11 J along x, Qi22 J along y, edge rows -(b0+b1) f0 + (b0+b1) f1,
// interior (bp+bc)/2 f+ - (bm+2bc+bp)/2 f + (bm+bc)/2 f-; the cross-term penalties with
// coefficient 2, all into d2x; first derivatives d21 times J Qi12, formed where they are read.
__global__ void __launch_bounds__(256) hd_lap21(HdGeo g, const double *__restrict__ W, double *__restrict__ Wout,
                                                double *__restrict__ E) {
    const int n = g.n;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= 6LL * n * n) return;
    const int k = blockIdx.y;
    const int i = (int)(t % n);
    const long long pj = t / n;
    const int j = (int)(pj % n), p = (int)(pj / n);
    const long long pb = (long long)p * n * g.pitch;
    const double *f = W + k * g.fs + pb;
    auto at = [&](int a, int b) { return (long long)b * g.pitch + a; };
    auto d1x = [&](int a, int b) {
        const double d = sbp_row<21, double>(a, n, [&](int m) { return f[at(m, b)]; });
        return d * g.J[pb + at(a, b)] * g.Qi12[pb + at(a, b)];
    };
    auto d1y = [&](int a, int b) {
        const double d = sbp_row<21, double>(b, n, [&](int m) { return f[at(a, m)]; });
        return d * g.J[pb + at(a, b)] * g.Qi12[pb + at(a, b)];
    };
    auto d2 = [&](int t0, auto bget, auto fget) {
        if (t0 == 0) {
            const double bc = bget(0), bp = bget(1);
            const double cc = -(bc + bp), cp = +(bc + bp);
            return cc * fget(0) + cp * fget(1);
        }
        if (t0 == n - 1) {
            const double bc = bget(n - 1), bm = bget(n - 2);
            const double cc = -(bc + bm), cm = +(bc + bm);
            return cm * fget(n - 2) + cc * fget(n - 1);
        }
        const double bm = bget(t0 - 1), bc = bget(t0), bp = bget(t0 + 1);
        const double cm = (bm + bc) / 2, cc = -(bm + 2 * bc + bp) / 2, cp = (bp + bc) / 2;
        return cp * fget(t0 + 1) + cc * fget(t0) + cm * fget(t0 - 1);
    };
    double d2x = d2(i, [&](int a) { return g.Qi11[pb + at(a, j)] * g.J[pb + at(a, j)]; },
                    [&](int a) { return f[at(a, j)]; });
    const double d2y = d2(j, [&](int b) { return g.Qi22[pb + at(i, b)] * g.J[pb + at(i, b)]; },
                          [&](int b) { return f[at(i, b)]; });
    // laplace_Ah_sbp21_narrow_mod.f90:189-215, x edges first
    constexpr double pc = 2.0;
    if (i == 0) d2x = d2x - 0.5 * (pc * (-2.0 * d1y(i, j)));
    if (i == n - 1) d2x = d2x - 0.5 * (pc * (2.0 * d1y(i, j)));
    if (j == 0) d2x = d2x - 0.5 * (pc * (-2.0 * d1x(i, j)));
    if (j == n - 1) d2x = d2x - 0.5 * (pc * (2.0 * d1x(i, j)));
    const double d2yx = sbp_row<21, double>(j, n, [&](int b) { return d1x(i, b); });
    const double d2xy = sbp_row<21, double>(i, n, [&](int a) { return d1y(a, j); });
    double v = 1.0 * d2x + 1.0 * d2y;
    v = v + 1.0 * d2xy + 1.0 * d2yx;
    v = v / (g.hs * g.hs);
    const long long o = pb + at(i, j);
    v = v / g.J[o];
    if (i == 0 || j == 0 || i == n - 1 || j == n - 1)
        E[(long long)k * 6 * g.es + p * g.es + edge_slot(n, i, j)] = v;
    else
        Wout[k * g.fs + o] = v;
}

// scalar edge sync (halo_Ah_scalar_sync_mod.f90:52-93) of field blockIdx.y from E into Wout
__global__ void hd_sync(HdGeo g, const double *__restrict__ E, double *__restrict__ Wout) {
    const int n = g.n, ns = 6 * g.es;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns) return;
    const int k = blockIdx.y;
    const double *Ek = E + (long long)k * ns;
    const int p = t / g.es;
    int i, j;
    slot_point(n, t - p * g.es, i, j);
    const bool L = i == 0, Rt = i == n - 1, B = j == 0, Tp = j == n - 1;
    const double own = Ek[t];
    double v;
    if ((L || Rt) && (B || Tp)) {
        const double Dx = Ek[nbr_slot(g, p, L ? 0 : 1, j)];
        const double Dy = Ek[nbr_slot(g, p, B ? 2 : 3, i)];
        v = (own + Dx + Dy) / 3.0;
    } else if (L) {
        v = 0.5 * (Ek[nbr_slot(g, p, 0, j)] + own);
    } else if (Rt) {
        v = 0.5 * (own + Ek[nbr_slot(g, p, 1, j)]);
    } else {
        v = 0.5 * (own + Ek[nbr_slot(g, p, B ? 2 : 3, i)]);
    }
    Wout[k * g.fs + ((long long)p * n + j) * g.pitch + i] = v;
}

// tendencies ht = ch W0 and (ut, vt) = (V.b1, V.b2) of V = cuv W1..3; yo = y0 + dt * tend
// (explicit_Eul1) or, with y0 = NULL, yo = tend (operator apply).  Velocity tendencies of
// panel-edge points go to E2 (x slots, then y slots) for the vector sync.
__global__ void hd_final(HdGeo g, const double *y0, double *yo, const double *__restrict__ W, double ch, double cuv,
                         double dt, double *__restrict__ E2) {
    const int n = g.n;
    const long long total = 6LL * n * n;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int i = (int)(t % n);
    const long long pj = t / n;
    const int j = (int)(pj % n);
    const int p = (int)(pj / n);
    const long long o = pj * g.pitch + i;
    const double ht = ch * W[o];
    yo[o] = y0 ? y0[o] + dt * ht : ht;
    const double tx = cuv * W[g.fs + o], ty = cuv * W[2 * g.fs + o], tz = cuv * W[3 * g.fs + o];
    const double ut = tx * g.b1[o] + ty * g.b1[g.fs + o] + tz * g.b1[2 * g.fs + o];
    const double vt = tx * g.b2[o] + ty * g.b2[g.fs + o] + tz * g.b2[2 * g.fs + o];
    if (i == 0 || j == 0 || i == n - 1 || j == n - 1) {
        const int sl = p * g.es + edge_slot(n, i, j);
        E2[sl] = ut;
        E2[6 * g.es + sl] = vt;
    } else {
        yo[g.fs + o] = y0 ? y0[g.fs + o] + dt * ut : ut;
        yo[2 * g.fs + o] = y0 ? y0[2 * g.fs + o] + dt * vt : vt;
    }
}

// covariant vector edge sync (ecs_Ah_vec_sync_mod.f90:51-124, components entering as
// (U, V) = (vt, ut)) of the edge velocity tendencies, then yo as in hd_final
__global__ void hd_vsync(HdGeo g, const double *__restrict__ E2, double dt, const double *y0, double *yo) {
    const int n = g.n, ns = 6 * g.es;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns) return;
    const int p = t / g.es;
    int i, j;
    slot_point(n, t - p * g.es, i, j);
    const bool L = i == 0, Rt = i == n - 1, B = j == 0, Tp = j == n - 1;
    auto nbr = [&](int e, int k, double &V_, double &U_) {
        const int s = nbr_slot(g, p, e, k);
        map_vec(g.einfo[(p * 4 + e) * 8 + 5], E2[s], E2[ns + s], V_, U_);
    };
    const double V = E2[t], U = E2[ns + t];
    double gx, gy, Vx, Ux, Vy, Uy;
    if ((L || Rt) && (B || Tp)) {
        nbr(L ? 0 : 1, j, Vx, Ux);
        nbr(B ? 2 : 3, i, Vy, Uy);
        const double qy = B ? g.qb[i] : g.qt[i];
        const double qx = L ? g.ql[j] : g.qr[j];
        gy = (U + Ux + Uy + qy * Vy) / 3.0;
        gx = (V + Vx + Vy + qx * Ux) / 3.0;
    } else if (L) {
        nbr(0, j, Vx, Ux);
        gy = 0.5 * (Ux + U);
        gx = 0.5 * (V + Vx + g.ql[j] * Ux);
    } else if (Rt) {
        nbr(1, j, Vx, Ux);
        gy = 0.5 * (U + Ux);
        gx = 0.5 * (V + Vx + g.qr[j] * Ux);
    } else if (B) {
        nbr(2, i, Vy, Uy);
        gx = 0.5 * (Vy + V);
        gy = 0.5 * (U + Uy + g.qb[i] * Vy);
    } else {
        nbr(3, i, Vy, Uy);
        gx = 0.5 * (V + Vy);
        gy = 0.5 * (U + Uy + g.qt[i] * Vy);
    }
    const long long o = ((long long)p * n + j) * g.pitch + i;
    yo[g.fs + o] = y0 ? y0[g.fs + o] + dt * gx : gx;
    yo[2 * g.fs + o] = y0 ? y0[2 * g.fs + o] + dt * gy : gy;
}

}  // namespace swm

// ------------------------------------------------------------------------------------------------
// Host side
namespace {

struct HdWork {
    double *b1 = nullptr, *b2 = nullptr;          // 3 planes each
    double *WA = nullptr, *WB = nullptr;          // 4 fields each (h, Vx, Vy, Vz)
    double *E = nullptr, *E2 = nullptr;           // 4 x ns, 2 x ns edge values
};

// closure rows of both panel edges disjoint: N+1 >= 12 (sbp42), 18 (sbp63); sbp21 has one-point
// closures (any context size)
int hd_min_n(int order) { return order == 21 ? 3 : 2 * (order == 63 ? swm::Ln<63>::NB : swm::Ln<42>::NB); }

int ensure_hd(swm_ctx *c, int order) {
    if (c->prec != 64) return fail(SWM_EARG, "the hyperdiffusion runs in FP64 contexts only");
    if (c->n < hd_min_n(order))
        return fail(SWM_EARG, "the sbp" + std::to_string(order) + " narrow Laplacian needs N+1 >= " +
                                  std::to_string(hd_min_n(order)));
    if (c->hd) return SWM_OK;
    HdWork *w = new HdWork();
    c->hd = w;
    const size_t fb = c->field_bytes(), ns = (size_t)24 * c->N;
    std::vector<double> b1((size_t)3 * c->fs, 0.0), b2((size_t)3 * c->fs, 0.0);
    swmh_basis_b(c->N, c->pitch, b1.data(), b2.data());
    HIP_TRY(hipMalloc((void **)&w->b1, 3 * fb));
    HIP_TRY(hipMalloc((void **)&w->b2, 3 * fb));
    HIP_TRY(hipMalloc((void **)&w->WA, 4 * fb));
    HIP_TRY(hipMalloc((void **)&w->WB, 4 * fb));
    HIP_TRY(hipMalloc((void **)&w->E, 4 * ns * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&w->E2, 2 * ns * sizeof(double)));
    HIP_TRY(hipMemcpy(w->b1, b1.data(), 3 * fb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(w->b2, b2.data(), 3 * fb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(w->WA, 0, 4 * fb));
    HIP_TRY(hipMemset(w->WB, 0, 4 * fb));
    HIP_TRY(hipDeviceSynchronize());  // null-stream copies and clears before the context stream uses them
    return SWM_OK;
}

swm::HdGeo hd_geo(const swm_ctx *c) {
    swm::HdGeo g;
    g.J = c->J64; g.Qi11 = c->Qi64[0]; g.Qi12 = c->Qi64[1]; g.Qi22 = c->Qi64[2];
    g.qb = (const double *)c->q[0]; g.qt = (const double *)c->q[1];
    g.ql = (const double *)c->q[2]; g.qr = (const double *)c->q[3];
    g.einfo = c->einfo;
    g.b1 = c->hd->b1; g.b2 = c->hd->b2;
    g.n = c->n; g.pitch = c->pitch; g.es = 4 * c->N; g.fs = c->fs;
    g.hs = c->hx * c->scale;
    return g;
}

// fields k0 .. k0+nf-1 of `in` (4-field work layout) -> L of each in the same fields of `out`,
// synced; order 21, 42 or 63 picks the narrow Laplacian
int enqueue_laplace(swm_ctx *c, int order, int k0, int nf, const double *in, double *out) {
    HdWork *w = c->hd;
    const swm::HdGeo g = hd_geo(c);
    const dim3 blk(256);
    const int ns = 24 * c->N;
    constexpr int LTX = 32, LTY = 8;
    const int ltiles = 6 * ((c->n + LTX - 1) / LTX) * ((c->n + LTY - 1) / LTY);
    const double *ik = in + (long long)k0 * c->fs;
    double *ok = out + (long long)k0 * c->fs, *ek = w->E + (long long)k0 * ns;
    if (order == 21)
        hipLaunchKernelGGL(swm::hd_lap21, dim3((unsigned)((6LL * c->n * c->n + 255) / 256), nf), blk, 0, c->stream, g,
                           ik, ok, ek);
    else if (order == 63)
        hipLaunchKernelGGL((swm::hd_lap_tile<63, LTX, LTY>), dim3(ltiles * nf), dim3(LTX * LTY), 0, c->stream, g, ik,
                           ok, ek, nf);
    else
        hipLaunchKernelGGL((swm::hd_lap_tile<42, LTX, LTY>), dim3(ltiles * nf), dim3(LTX * LTY), 0, c->stream, g, ik,
                           ok, ek, nf);
    hipLaunchKernelGGL(swm::hd_sync, dim3((ns + 255) / 256, nf), blk, 0, c->stream, g, (const double *)ek, ok);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

// hordiff_factory_mod.f90:25-50: the narrow-Laplacian hyperdiffusion names and their families
int hordiff_order(const char *name, const char *prefix, int &order) {
    const std::string s(name);
    for (int o : {21, 42, 63})
        if (s == prefix + std::to_string(o) + "_narrow") {
            order = o;
            return SWM_OK;
        }
    return fail(SWM_ENAME, std::string("unsupported hyperdiffusion: ") + name + " (supported: " + prefix +
                               "21_narrow, " + prefix + "42_narrow, " + prefix + "63_narrow)");
}

int parse_hordiff(const char *uv_name, const char *h_name, int &uv_order, int &h_order) {
    if (!uv_name || !h_name) return fail(SWM_EARG, "NULL hyperdiffusion name");
    int rc = hordiff_order(uv_name, "hordiff_vec_xyz_Ah_sbp_", uv_order);
    if (rc) return rc;
    return hordiff_order(h_name, "hordiff_scalar_Ah_sbp_", h_order);
}

// laplace_factory_mod.f90:20-30: the narrow-Laplacian names
int parse_laplace(const char *name, int &order) {
    if (!name) return fail(SWM_EARG, "NULL Laplacian name");
    const std::string s(name);
    if (s == "laplace_ah_sbp21_narrow") order = 21;
    else if (s == "laplace_ah_sbp42_narrow") order = 42;
    else if (s == "laplace_ah_sbp63_narrow") order = 63;
    else return fail(SWM_ENAME, "unsupported Laplacian: " + s + " (supported: laplace_ah_sbp21_narrow, "
                                    "laplace_ah_sbp42_narrow, laplace_ah_sbp63_narrow)");
    return SWM_OK;
}

// yo = y0 + dt * diff(yin) (y0 = yin = yo for the Euler step), or yo = diff(yin) with y0 = NULL
int enqueue_diff(swm_ctx *c, int uv_order, int h_order, const double *yin, const double *y0, double *yo,
                 double dt, double uv_coeff, double h_coeff) {
    HdWork *w = c->hd;
    const swm::HdGeo g = hd_geo(c);
    const long long total = 6LL * c->n * c->n;
    const dim3 blk(256), grd((unsigned)((total + 255) / 256));
    const int ns = 24 * c->N;
    int rc;
    hipLaunchKernelGGL(swm::hd_to_cart, grd, blk, 0, c->stream, g, yin, w->WA);
    // h and the three Cartesian velocity components share the launches of one family
    for (int pass = 0; pass < 2; ++pass) {
        const double *src = pass ? w->WB : w->WA;
        double *dst = pass ? w->WA : w->WB;
        if (uv_order == h_order) {
            if ((rc = enqueue_laplace(c, h_order, 0, 4, src, dst))) return rc;
        } else if ((rc = enqueue_laplace(c, h_order, 0, 1, src, dst)) ||
                   (rc = enqueue_laplace(c, uv_order, 1, 3, src, dst))) {
            return rc;
        }
    }
    // hordiff_factory_mod.f90:158-188: diff_coeff = hordiff_coeff*scale*hx, tend coefficient
    // (-1)**3 * diff_coeff**4 (the power as (c*c)*(c*c))
    const double du = uv_coeff * c->scale * c->hx, dh = h_coeff * c->scale * c->hx;
    const double cuv = -1.0 * ((du * du) * (du * du)), ch = -1.0 * ((dh * dh) * (dh * dh));
    hipLaunchKernelGGL(swm::hd_final, grd, blk, 0, c->stream, g, y0, yo, (const double *)w->WA, ch, cuv, dt, w->E2);
    hipLaunchKernelGGL(swm::hd_vsync, dim3((ns + 255) / 256), blk, 0, c->stream, g, (const double *)w->E2, dt, y0, yo);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

void free_hd_work(HdWork *w) {
    if (!w) return;
    void *ptrs[] = {w->b1, w->b2, w->WA, w->WB, w->E, w->E2};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete w;
}

}  // namespace

extern "C" int swm_laplace(swm_ctx *c, const char *laplace_name, const void *f, void *out) {
    if (!c || !f || !out) return fail(SWM_EARG, "NULL argument");
    if (f == out) return fail(SWM_EARG, "out must differ from f");
    int order = 0;
    int rc = parse_laplace(laplace_name, order);
    if (rc || (rc = ensure_hd(c, order))) return rc;
    HIP_TRY(hipMemcpyAsync(c->hd->WA, f, c->field_bytes(), hipMemcpyDeviceToDevice, c->stream));
    if ((rc = enqueue_laplace(c, order, 0, 1, c->hd->WA, c->hd->WB))) return rc;
    HIP_TRY(hipMemcpyAsync(out, c->hd->WB, c->field_bytes(), hipMemcpyDeviceToDevice, c->stream));
    return SWM_OK;
}

extern "C" int swm_diff_step(swm_ctx *c, void *state, double dt, const char *hordiff_uv_name, double uv_diff_coeff,
                             const char *hordiff_h_name, double h_diff_coeff) {
    if (!c || !state) return fail(SWM_EARG, "NULL argument");
    int uv_order = 0, h_order = 0;
    int rc = parse_hordiff(hordiff_uv_name, hordiff_h_name, uv_order, h_order);
    if (rc || (rc = ensure_hd(c, uv_order)) || (rc = ensure_hd(c, h_order))) return rc;
    double *y = (double *)state;
    return enqueue_diff(c, uv_order, h_order, y, y, y, dt, uv_diff_coeff, h_diff_coeff);
}

extern "C" int swm_diff_apply(swm_ctx *c, const void *vin, void *vout, const char *hordiff_uv_name,
                              double uv_diff_coeff, const char *hordiff_h_name, double h_diff_coeff) {
    if (!c || !vin || !vout) return fail(SWM_EARG, "NULL argument");
    if (vin == vout) return fail(SWM_EARG, "vout must differ from vin");
    int uv_order = 0, h_order = 0;
    int rc = parse_hordiff(hordiff_uv_name, hordiff_h_name, uv_order, h_order);
    if (rc || (rc = ensure_hd(c, uv_order)) || (rc = ensure_hd(c, h_order))) return rc;
    return enqueue_diff(c, uv_order, h_order, (const double *)vin, nullptr, (double *)vout, 0.0, uv_diff_coeff,
                        h_diff_coeff);
}
