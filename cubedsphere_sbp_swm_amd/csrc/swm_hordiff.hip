// Hyperdiffusion of the Ah state with the narrow-stencil SBP Laplacian (sbp42), FP64
// (included into swm_capi.hip).
//
// Laplacian, laplace_Ah_sbp42_narrow_t%calc_laplace (differential_operators/horizontal/
// laplace/laplace_Ah_sbp42_narrow_mod.f90:23-763):
//   d1x = Dx(f) J Qi12,  d1y = Dy(f) J Qi12                      (first-derivative SBP d42)
//   d2x = D2x(Qi11 J) f, d2y = D2y(Qi22 J) f                     (narrow variable-coefficient D2)
//   d2x -= 0.5 * (48/17) * (-+2 d1y) on the x edges, then (-+2 d1x) on the y edges
//   L f = ((d2x + d2y) + Dx(d1y) + Dy(d1x)) / (hx a)^2 / J, then the scalar edge sync.
// Hyperdiffusion of one field, hordiff_scalar_t%calc_diff (hordiff/hordiff_scalar_mod.f90:
// 34-59): tend = -(c a hx)^4 L(L f).  Velocity, hordiff_colocated_xyz_t%calc_diff_vec
// (hordiff/hordiff_colocated_mod.f90:39-58): Cartesian components V = b1 u + b2 v, the scalar
// hyperdiffusion of each, back as (V.b1, V.b2), covariant vector edge sync.
// Step, operator_swm_diff_t%apply + explicit_Eul1 (operator_swm_diff_mod.f90:25-46,
// time_schemes/explicit_Eul1_mod.f90:20-30): y += dt * tend.
//
// One Laplacian is two launches, hd_lap_tile and hd_sync, over several fields (blockIdx.y):
// h and the three Cartesian velocity components share every launch.  Panel-edge results are
// written to a compact edge buffer first and synchronised from there, so that no kernel reads
// a value another block of the same launch overwrites.

#include "laplace_sbp42_narrow.h"

namespace swm {

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

// Closure row R of the narrow D2 with b[k], fv[c] taken k, c points from the panel edge, and
// the interior row, expanded at compile time (every index and weight a constant); each weight
// p/q is the correctly rounded quotient, as the run-time division of the oracle.
template <int T>
__device__ __forceinline__ double ln42_term(const double *b) {
    constexpr double w = LN42_TNUM[T] / LN42_TDEN[T];
    return w * b[LN42_TB[T]];
}
template <int T0, int... Is>
__device__ __forceinline__ double ln42_sum(const double *b, std::integer_sequence<int, Is...>) {
    double m = 0.0;
    ((m = Is == 0 ? (LN42_TSIGN[T0 + Is] < 0 ? -ln42_term<T0 + Is>(b) : ln42_term<T0 + Is>(b))
                  : (LN42_TSIGN[T0 + Is] < 0 ? m - ln42_term<T0 + Is>(b) : m + ln42_term<T0 + Is>(b))),
     ...);
    return m;
}
template <int R, int C>
__device__ __forceinline__ double ln42_m(const double *b) {
    constexpr int t0 = LN42_MSTART[R][C], cnt = LN42_MSTART[R][C + 1] - LN42_MSTART[R][C];
    return ln42_sum<t0>(b, std::make_integer_sequence<int, cnt>{});
}
template <int R, int... Cs>
__device__ __forceinline__ double ln42_row_cols(const double *b, const double *fv, std::integer_sequence<int, Cs...>) {
    double s = 0.0;
    ((s = Cs == 0 ? ln42_m<R, Cs>(b) * fv[Cs] : s + ln42_m<R, Cs>(b) * fv[Cs]), ...);
    constexpr double h = LN42_HNUM[R] / LN42_HDEN[R];
    return -((1.0 / h) * s);
}
template <int R>
__device__ __forceinline__ double ln42_row(const double *b, const double *fv) {
    return ln42_row_cols<R>(b, fv, std::make_integer_sequence<int, LN42_NCOLS[R]>{});
}

template <int T>
__device__ __forceinline__ double ln42_iterm(const double *b) {
    constexpr double w = LN42_INUM[T] / LN42_IDEN[T];
    return w * b[LN42_IB[T] + 2];
}
template <int T0, int... Is>
__device__ __forceinline__ double ln42_isum(const double *b, std::integer_sequence<int, Is...>) {
    double m = 0.0;
    ((m = Is == 0 ? (LN42_ISIGN[T0 + Is] < 0 ? -ln42_iterm<T0 + Is>(b) : ln42_iterm<T0 + Is>(b))
                  : (LN42_ISIGN[T0 + Is] < 0 ? m - ln42_iterm<T0 + Is>(b) : m + ln42_iterm<T0 + Is>(b))),
     ...);
    return m;
}
template <int O>
__device__ __forceinline__ double ln42_icoef(const double *b) {
    return ln42_isum<LN42_ISTART[O]>(b, std::make_integer_sequence<int, LN42_ISTART[O + 1] - LN42_ISTART[O]>{});
}
template <int... Os>
__device__ __forceinline__ double ln42_interior_seq(const double *b, const double *fv, std::integer_sequence<int, Os...>) {
    double s = 0.0;
    ((s = Os == 0 ? ln42_icoef<Os>(b) * fv[Os] : s + ln42_icoef<Os>(b) * fv[Os]), ...);
    return s;
}
__device__ __forceinline__ double ln42_interior(const double *b, const double *fv) {
    return ln42_interior_seq(b, fv, std::make_integer_sequence<int, 5>{});
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
// box of lines its rows read (+-2 points in the interior, points 0..7 / n-8..n-1 where it holds
// closure rows: the narrow D2 closure reads 8 points, the first-derivative closure 6), and
// b = Qi11 J on the tile rows, b = Qi22 J on the tile columns.  It then forms the first
// derivatives d1y = Dy(f) J Qi12 on the tile rows and d1x = Dx(f) J Qi12 on the tile columns
// (each at the positions whose J and Qi12 the same thread loaded), and finally the D2 rows,
// the cross-term penalties and the cross derivatives at its points.
template <int TX, int TY>
struct HdTile {
    static constexpr int R = 2, B = 6, NR = 8;
    static constexpr int span(int t) { return t + R + B - 1 > NR ? t + R + B - 1 : NR; }
    static constexpr int LW = span(TX), LH = span(TY), NT = TX * TY;
    static constexpr int KF = (LH * LW + NT - 1) / NT;              // box positions per thread
    static constexpr int KX = (TY * LW + LH * TX + NT - 1) / NT;    // cross positions per thread
};

__device__ __forceinline__ void hd_rows(int a0, int ae, int n, int &lo, int &hi) {
    constexpr int R = 2, B = 6, NR = 8;
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

template <int TX, int TY>
__global__ void __launch_bounds__(TX * TY) hd_lap_tile(HdGeo g, const double *__restrict__ W,
                                                       double *__restrict__ Wout, double *__restrict__ E) {
    using C = HdTile<TX, TY>;
    __shared__ double sf[C::LH][C::LW];                  // f on the box
    __shared__ double sbx[TY][C::LW], sdy[TY][C::LW];    // tile rows: Qi11 J, d1y
    __shared__ double sby[C::LH][TX], sdx[C::LH][TX];    // tile columns: Qi22 J, d1x
    const int n = g.n, k = blockIdx.y;
    const int ntx = (n + TX - 1) / TX, nty = (n + TY - 1) / TY;
    const int tile = xcd_tile(blockIdx.x, 6 * ntx * nty);
    const int p = tile / (ntx * nty), rem = tile - p * ntx * nty;
    const int x0 = (rem % ntx) * TX, y0 = (rem / ntx) * TY;
    const int xe = x0 + TX < n ? x0 + TX : n, ye = y0 + TY < n ? y0 + TY : n;
    int lx, hx, ly, hy;
    hd_rows(x0, xe, n, lx, hx);
    hd_rows(y0, ye, n, ly, hy);
    const double *f = W + k * g.fs;
    const long long pb = (long long)p * n;
    const int wa = hx - lx, nbox = (hy - ly) * wa;
    const int nA = (ye - y0) * wa, nc = xe - x0, nB = (hy - ly) * nc;
    const int tid = threadIdx.x;

    // ---- loads: f on the box; Qi, J, Qi12 at the cross positions (tile rows x line range:
    // kind 1; line range x tile columns: kind 2) ----
    double fv[C::KF];
#pragma unroll
    for (int u = 0; u < C::KF; ++u) {
        const int q = tid + u * C::NT;
        if (q < nbox) {
            const int r = q / wa;
            fv[u] = f[(pb + ly + r) * g.pitch + lx + (q - r * wa)];
        }
    }
    double qv[C::KX], jv[C::KX], q12[C::KX];
    int kind[C::KX], px[C::KX], py[C::KX];
#pragma unroll
    for (int u = 0; u < C::KX; ++u) {
        const int q = tid + u * C::NT;
        kind[u] = 0;
        px[u] = 0;
        py[u] = 0;
        if (q < nA) {
            const int r = q / wa;
            py[u] = y0 + r;
            px[u] = lx + (q - r * wa);
            kind[u] = 1;
        } else if (q < nA + nB) {
            const int q2 = q - nA, r = q2 / nc;
            py[u] = ly + r;
            px[u] = x0 + (q2 - r * nc);
            kind[u] = 2;
        }
        if (kind[u]) {
            const long long o = (pb + py[u]) * g.pitch + px[u];
            qv[u] = kind[u] == 1 ? g.Qi11[o] : g.Qi22[o];
            jv[u] = g.J[o];
            q12[u] = g.Qi12[o];
        }
    }
#pragma unroll
    for (int u = 0; u < C::KF; ++u) {
        const int q = tid + u * C::NT;
        if (q < nbox) {
            const int r = q / wa;
            sf[r][q - r * wa] = fv[u];
        }
    }
#pragma unroll
    for (int u = 0; u < C::KX; ++u) {
        if (kind[u] == 1) sbx[py[u] - y0][px[u] - lx] = qv[u] * jv[u];
        else if (kind[u] == 2) sby[py[u] - ly][px[u] - x0] = qv[u] * jv[u];
    }
    __syncthreads();

    // ---- first derivatives at the cross positions (calc_tile_d1x_d1y, :60-93) ----
#pragma unroll
    for (int u = 0; u < C::KX; ++u) {
        if (kind[u] == 1) {
            const int c = px[u] - lx;
            const double d = sbp_row<42, double>(py[u], n, [&](int a) { return sf[a - ly][c]; });
            sdy[py[u] - y0][c] = d * jv[u] * q12[u];
        } else if (kind[u] == 2) {
            const int r = py[u] - ly;
            const double d = sbp_row<42, double>(px[u], n, [&](int a) { return sf[r][a - lx]; });
            sdx[r][px[u] - x0] = d * jv[u] * q12[u];
        }
    }
    __syncthreads();

    const int tx = tid % TX, ty = tid / TX, i = x0 + tx, j = y0 + ty;
    if (i >= n || j >= n) return;
    const long long o = (pb + j) * g.pitch + i;
    auto d2 = [&](int t, const auto &bget, const auto &fget) {
        double bb[LN42_NCOL], ff[LN42_NCOL];
        if (t < LN42_NB || t >= n - LN42_NB) {
            const bool left = t < LN42_NB;
#pragma unroll
            for (int q = 0; q < LN42_NCOL; ++q) {
                const int a = left ? q : n - 1 - q;
                bb[q] = bget(a);
                ff[q] = fget(a);
            }
            switch (left ? t : n - 1 - t) {
            case 0: return ln42_row<0>(bb, ff);
            case 1: return ln42_row<1>(bb, ff);
            case 2: return ln42_row<2>(bb, ff);
            case 3: return ln42_row<3>(bb, ff);
            case 4: return ln42_row<4>(bb, ff);
            default: return ln42_row<5>(bb, ff);
            }
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            bb[q] = bget(t - 2 + q);
            ff[q] = fget(t - 2 + q);
        }
        return ln42_interior(bb, ff);
    };
    double d2x = d2(i, [&](int a) { return sbx[ty][a - lx]; }, [&](int a) { return sf[j - ly][a - lx]; });
    const double d2y = d2(j, [&](int a) { return sby[a - ly][tx]; }, [&](int a) { return sf[a - ly][i - lx]; });
    // penalties of the cross terms (laplace_Ah_sbp42_narrow_mod.f90:716-742), x edges first
    const double pc = LN42_PNUM / LN42_PDEN;
    if (i == 0) d2x = d2x - 0.5 * (pc * (-2.0 * sdy[ty][i - lx]));
    if (i == n - 1) d2x = d2x - 0.5 * (pc * (2.0 * sdy[ty][i - lx]));
    if (j == 0) d2x = d2x - 0.5 * (pc * (-2.0 * sdx[j - ly][tx]));
    if (j == n - 1) d2x = d2x - 0.5 * (pc * (2.0 * sdx[j - ly][tx]));
    const double d2yx = sbp_row<42, double>(j, n, [&](int a) { return sdx[a - ly][tx]; });
    const double d2xy = sbp_row<42, double>(i, n, [&](int a) { return sdy[ty][a - lx]; });
    double v = 1.0 * d2x + 1.0 * d2y;
    v = v + 1.0 * d2xy + 1.0 * d2yx;
    v = v / (g.hs * g.hs);
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

constexpr int HD_MIN_N = 2 * LN42_NB;  // closure rows of both edges disjoint

int ensure_hd(swm_ctx *c) {
    if (c->hd) return SWM_OK;
    if (c->prec != 64) return fail(SWM_EARG, "the hyperdiffusion runs in FP64 contexts only");
    if (c->n < HD_MIN_N) return fail(SWM_EARG, "the sbp42 narrow Laplacian needs N+1 >= " + std::to_string(HD_MIN_N));
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

// nf fields of `in` (4-field work layout) -> L of each in `out`, synced
int enqueue_laplace(swm_ctx *c, int nf, const double *in, double *out) {
    HdWork *w = c->hd;
    const swm::HdGeo g = hd_geo(c);
    const dim3 blk(256);
    const int ns = 24 * c->N;
    constexpr int LTX = 32, LTY = 8;
    const int ltiles = 6 * ((c->n + LTX - 1) / LTX) * ((c->n + LTY - 1) / LTY);
    hipLaunchKernelGGL((swm::hd_lap_tile<LTX, LTY>), dim3(ltiles, nf), dim3(LTX * LTY), 0, c->stream, g, in, out,
                       w->E);
    hipLaunchKernelGGL(swm::hd_sync, dim3((ns + 255) / 256, nf), blk, 0, c->stream, g, (const double *)w->E, out);
    HIP_TRY(hipGetLastError());
    return SWM_OK;
}

int parse_hordiff(const char *uv_name, const char *h_name) {
    if (!uv_name || !h_name) return fail(SWM_EARG, "NULL hyperdiffusion name");
    if (std::string(uv_name) != "hordiff_vec_xyz_Ah_sbp_42_narrow")
        return fail(SWM_ENAME, std::string("unsupported velocity hyperdiffusion: ") + uv_name);
    if (std::string(h_name) != "hordiff_scalar_Ah_sbp_42_narrow")
        return fail(SWM_ENAME, std::string("unsupported height hyperdiffusion: ") + h_name);
    return SWM_OK;
}

// yo = y0 + dt * diff(yin) (y0 = yin = yo for the Euler step), or yo = diff(yin) with y0 = NULL
int enqueue_diff(swm_ctx *c, const double *yin, const double *y0, double *yo, double dt, double uv_coeff,
                 double h_coeff) {
    HdWork *w = c->hd;
    const swm::HdGeo g = hd_geo(c);
    const long long total = 6LL * c->n * c->n;
    const dim3 blk(256), grd((unsigned)((total + 255) / 256));
    const int ns = 24 * c->N;
    int rc;
    hipLaunchKernelGGL(swm::hd_to_cart, grd, blk, 0, c->stream, g, yin, w->WA);
    if ((rc = enqueue_laplace(c, 4, w->WA, w->WB)) || (rc = enqueue_laplace(c, 4, w->WB, w->WA))) return rc;
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

extern "C" int swm_laplace(swm_ctx *c, const void *f, void *out) {
    if (!c || !f || !out) return fail(SWM_EARG, "NULL argument");
    if (f == out) return fail(SWM_EARG, "out must differ from f");
    int rc = ensure_hd(c);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->hd->WA, f, c->field_bytes(), hipMemcpyDeviceToDevice, c->stream));
    if ((rc = enqueue_laplace(c, 1, c->hd->WA, c->hd->WB))) return rc;
    HIP_TRY(hipMemcpyAsync(out, c->hd->WB, c->field_bytes(), hipMemcpyDeviceToDevice, c->stream));
    return SWM_OK;
}

extern "C" int swm_diff_step(swm_ctx *c, void *state, double dt, const char *hordiff_uv_name, double uv_diff_coeff,
                             const char *hordiff_h_name, double h_diff_coeff) {
    if (!c || !state) return fail(SWM_EARG, "NULL argument");
    int rc = parse_hordiff(hordiff_uv_name, hordiff_h_name);
    if (rc || (rc = ensure_hd(c))) return rc;
    double *y = (double *)state;
    return enqueue_diff(c, y, y, y, dt, uv_diff_coeff, h_diff_coeff);
}

extern "C" int swm_diff_apply(swm_ctx *c, const void *vin, void *vout, const char *hordiff_uv_name,
                              double uv_diff_coeff, const char *hordiff_h_name, double h_diff_coeff) {
    if (!c || !vin || !vout) return fail(SWM_EARG, "NULL argument");
    if (vin == vout) return fail(SWM_EARG, "vout must differ from vin");
    int rc = parse_hordiff(hordiff_uv_name, hordiff_h_name);
    if (rc || (rc = ensure_hd(c))) return rc;
    return enqueue_diff(c, (const double *)vin, nullptr, (double *)vout, 0.0, uv_diff_coeff, h_diff_coeff);
}
