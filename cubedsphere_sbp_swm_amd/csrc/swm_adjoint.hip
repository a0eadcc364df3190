// Hand-written adjoint of the Ah-grid SWM right-hand side and of the RK4 step, FP64, SBP
// families 42 / 43 / 63 (included into swm_capi.hip).
//
// Forward map at a state y = (h, u, v) (operator_swm_mod.f90:61-113):
//   ut, vt = Qi (u, v);  K = (u ut + v vt)/2 + g h (+ g hs)
//   G  = Vsync( Dx K, Dy K ) / (hx a)                 gradient + covariant vector sync
//   z  = Ssync( Dx(J v/J) + Dy(J (-u/J)) ) / (J hx a) curl through the divergence
//   dh = -Ssync( Dx(J h ut) + Dy(J h vt) ) / (J hx a)
//   du = -Gx + (f + z) vt J,   dv = -Gy - (f + z) ut J
// The vector-Jacobian product g = (dF/dy)^T mu is evaluated by two launches:
//   1. adj_pre_edge: pointwise seeds of the three linear operators
//                 Ldiv = -mu_h, Lz = J (vt mu_u - ut mu_v), LG = -(mu_u, mu_v)/(hx a),
//                 divided by J hx a, at the off-panel-edge points, and, in blocks of their
//                 own, the transposed edge syncs of the seeds at the panel-edge points into a
//                 compact buffer (the scalar sync is an averaging projector, hence symmetric;
//                 the transposed covariant vector sync is a sparse matrix built on the host),
//                 then the division;
//   2. adj_main:  transposed SBP derivatives D^T along x and y, reading panel-edge points from
//                 the compact buffer (boundary columns from the
//                 closure blocks, interior columns -D), the pointwise chain rule back to
//                 (h, u, v), and the RK4 reverse-stage combination.
// RK4 reverse step for y1 = y0 + dt/2 k1, y2 = y0 + dt/2 k2, y3 = y0 + dt k3,
// y+ = y0 + dt/6 k1 + dt/3 k2 + dt/3 k3 + dt/6 k4 (rk4_mod.f90:28-40), given L = dJ/dy+:
//   g3 = J3^T (dt/6 L);  MU = dt/3 L + dt g3;    ACC = L + g3
//   g2 = J2^T MU;        MU = dt/3 L + dt/2 g2;  ACC += g2
//   g1 = J1^T MU;        MU = dt/6 L + dt/2 g1;  ACC += g1
//   g0 = J0^T MU;        L  = ACC + g0
// Step states come from checkpoints every C steps: the reverse sweep recomputes one segment
// of C steps from its checkpoint with the forward stage kernels, which also store each
// step's stage inputs y1..y3 and the synced vorticity of y0..y3 for the adjoint kernels.

namespace swm {

enum AdjMode : int { ADJ_VJP = 0, ADJ_S4 = 1, ADJ_S32 = 2, ADJ_S1 = 3 };

struct V3 {
    double a, b, c;
};
__device__ __forceinline__ V3 operator+(const V3 &x, const V3 &y) { return {x.a + y.a, x.b + y.b, x.c + y.c}; }
__device__ __forceinline__ V3 operator*(double w, const V3 &x) { return {w * x.a, w * x.b, w * x.c}; }

template <int ORD>
struct SbpT {
    using S = Sbp<ORD>;
    static constexpr int R = -S::INSHIFT;
    static constexpr int maxlast() {
        int m = 0;
        for (int k = 0; k < S::NE; ++k) m = S::last[k] > m ? S::last[k] : m;
        return m;
    }
    static constexpr int W = (maxlast() - 1 > R) ? maxlast() - 1 : R;        // column bandwidth of D
    static constexpr int B = (S::NE + R > maxlast()) ? S::NE + R : maxlast(); // boundary columns
    static constexpr int MIN_N = B + W + S::NE;                               // left/right blocks disjoint
    // D[r][c] of the left block and the interior (valid for r < n - NE)
    static constexpr double left(int r, int c) {
        if (r < S::NE) return (c >= 0 && c < S::last[r]) ? S::Ql[r][c] : 0.0;
        const int k = c - r - S::INSHIFT;
        return (k >= 0 && k < S::WIN) ? S::Qin[k] : 0.0;
    }
};

template <int ORD, int C, typename G, int... Rs>
__device__ __forceinline__ V3 dT_col(const G &get, std::integer_sequence<int, Rs...>) {
    V3 s{0.0, 0.0, 0.0};
    auto term = [&](auto rI) {
        constexpr int r = decltype(rI)::value;
        constexpr double w = SbpT<ORD>::left(r, C);
        if constexpr (w != 0.0) s = s + w * get(r);
    };
    (term(std::integral_constant<int, Rs>{}), ...);
    return s;
}

template <int ORD, typename G, int... Cs>
__device__ __forceinline__ V3 dT_left(int c, const G &get, std::integer_sequence<int, Cs...>) {
    V3 out{0.0, 0.0, 0.0};
    constexpr int NR = SbpT<ORD>::B + SbpT<ORD>::W;
    ((c == Cs ? (out = dT_col<ORD, Cs>(get, std::make_integer_sequence<int, NR>{}), 0) : 0), ...);
    return out;
}

// (D^T y)_c for a line of n points; get(r) returns the three transported fields at row r
template <int ORD, typename G>
__device__ __forceinline__ V3 sbp_col_T(int c, int n, const G &get) {
    using S = Sbp<ORD>;
    using T = SbpT<ORD>;
    if (c >= T::B && c < n - T::B) {
        // interior columns: sum_k Qin[k] y[c + R - k]
        V3 s{0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < S::WIN; ++k)
            if (S::Qin[k] != 0.0) s = s + S::Qin[k] * get(c + T::R - k);
        return s;
    }
    if (c < T::B) return dT_left<ORD>(c, get, std::make_integer_sequence<int, T::B>{});
    // right block: D[n-1-r][n-1-c] = -D[r][c]
    const V3 m = dT_left<ORD>(n - 1 - c, [&](int r) { return get(n - 1 - r); }, std::make_integer_sequence<int, T::B>{});
    return {-m.a, -m.b, -m.c};
}

struct AdjGeo {
    const double *J, *Qi11, *Qi12, *Qi22, *f;
    int n, pitch, es;
    long long fs;
    double hxs, grav;
};

// One launch for the pointwise seeds and the transposed edge syncs.  Blocks below nb_edge
// evaluate the transposed syncs into E = [S Ldiv | S Lz | (V^T LG)_x | (V^T LG)_y] (ns slots
// each): task t < ns is scalar row t for Ldiv and Lz, divided by J hx a at the slot's point;
// task t >= ns is row t - ns of V^T (x rows, then y rows).  They compute the seeds they gather
// from (y, M) at the panel-edge points themselves, so they do not wait for W.  The remaining
// blocks write W = [Ldiv | Lz | LGx | LGy] for the off-edge points read by adj_main.
// s_off: field-relative point offsets; v_off: point offset + comp * fs (comp 0 = x, 1 = y).
__global__ void adj_pre_edge(AdjGeo g, const double *__restrict__ y, const double *__restrict__ M, double alpha,
                             double *__restrict__ W, int nb_edge, const int *__restrict__ s_ptr,
                             const long long *__restrict__ s_off, const double *__restrict__ s_w,
                             const int *__restrict__ v_ptr, const long long *__restrict__ v_off,
                             const double *__restrict__ v_w, const long long *__restrict__ r_off,
                             double *__restrict__ E) {
    const int n = g.n;
    if ((int)blockIdx.x < nb_edge) {
        const int ns = 6 * g.es;
        const int t = blockIdx.x * blockDim.x + threadIdx.x;
        if (t >= 3 * ns) return;
        if (t < ns) {
            double a = 0.0, b = 0.0;
            for (int k = s_ptr[t]; k < s_ptr[t + 1]; ++k) {
                const long long o = s_off[k];
                const double u = y[g.fs + o], v = y[2 * g.fs + o];
                const double ut = g.Qi11[o] * u + g.Qi12[o] * v;
                const double vt = g.Qi12[o] * u + g.Qi22[o] * v;
                const double mh = alpha * M[o], mu = alpha * M[g.fs + o], mv = alpha * M[2 * g.fs + o];
                const double z = g.J[o] * (vt * mu - ut * mv);
                a += s_w[k] * -mh;
                b += s_w[k] * z;
            }
            const double dnm = g.J[r_off[t]] * g.hxs;
            E[t] = a / dnm;
            E[ns + t] = b / dnm;
            return;
        }
        const int row = t - ns;
        double s = 0.0;
        for (int k = v_ptr[row]; k < v_ptr[row + 1]; ++k) s += v_w[k] * (-(alpha * M[g.fs + v_off[k]]) / g.hxs);
        E[2 * ns + row] = s;
        return;
    }
    const long long total = 6LL * n * n;
    const long long stride = (long long)(gridDim.x - nb_edge) * blockDim.x;
    for (long long t = (blockIdx.x - nb_edge) * (long long)blockDim.x + threadIdx.x; t < total; t += stride) {
        const int i = (int)(t % n);
        const long long pj = t / n;
        const int j = (int)(pj % n);
        const int p = (int)(pj / n);
        const long long o = ((long long)p * n + j) * g.pitch + i;
        const double u = y[g.fs + o], v = y[2 * g.fs + o];
        const double Jv = g.J[o];
        const double ut = g.Qi11[o] * u + g.Qi12[o] * v;
        const double vt = g.Qi12[o] * u + g.Qi22[o] * v;
        const double mh = alpha * M[o], mu = alpha * M[g.fs + o], mv = alpha * M[2 * g.fs + o];
        const double z = Jv * (vt * mu - ut * mv);
        const double dnm = Jv * g.hxs;
        W[o] = -mh / dnm;
        W[g.fs + o] = z / dnm;
        W[2 * g.fs + o] = -mu / g.hxs;
        W[3 * g.fs + o] = -mv / g.hxs;
    }
}

__device__ __forceinline__ int edge_slot(int n, int i, int j) {
    if (j == 0) return i;
    if (j == n - 1) return n + i;
    if (i == 0) return 2 * n + j - 1;
    return 3 * n + j - 3;
}

struct AdjArgs {
    const double *y, *M, *zeta, *W, *E;
    double alpha;
    double *out;         // ADJ_VJP
    double *L, *MU, *ACC;
    double a, b;         // MU = a L + b g
};

template <int ORD, int MODE>
__global__ void __launch_bounds__(256) adj_main(AdjGeo g, AdjArgs A) {
    const int n = g.n;
    const long long total = 6LL * n * n;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int i = (int)(t % n);
    const long long pj = t / n;
    const int j = (int)(pj % n);
    const int p = (int)(pj / n);
    const long long row = ((long long)p * n + j) * g.pitch;
    const long long o = row + i;
    // off-edge points read W, panel-edge points the synced values in E
    const double *W0 = A.W, *W1 = A.W + g.fs, *W2 = A.W + 2 * g.fs, *W3 = A.W + 3 * g.fs;
    const int ns = 6 * g.es;
    const double *E0 = A.E + (long long)p * g.es, *E1 = E0 + ns, *E2 = E0 + 2 * ns, *E3 = E0 + 3 * ns;
    const bool jedge = j == 0 || j == n - 1, iedge = i == 0 || i == n - 1;
    const V3 dx = sbp_col_T<ORD>(i, n, [&](int r) {
        if (jedge || r == 0 || r == n - 1) {
            const int sl = edge_slot(n, r, j);
            return V3{E0[sl], E1[sl], E2[sl]};
        }
        return V3{W0[row + r], W1[row + r], W2[row + r]};
    });
    const long long col = (long long)p * n * g.pitch + i;
    const V3 dy = sbp_col_T<ORD>(j, n, [&](int r) {
        if (iedge || r == 0 || r == n - 1) {
            const int sl = edge_slot(n, i, r);
            return V3{E0[sl], E1[sl], E3[sl]};
        }
        const long long q = col + (long long)r * g.pitch;
        return V3{W0[q], W1[q], W3[q]};
    });
    const double Jv = g.J[o];
    const double lFx = Jv * dx.a, lFy = Jv * dy.a, lAx = Jv * dx.b, lAy = Jv * dy.b, lK = dx.c + dy.c;
    const double h = A.y[o], u = A.y[g.fs + o], v = A.y[2 * g.fs + o];
    const double q11 = g.Qi11[o], q12 = g.Qi12[o], q22 = g.Qi22[o];
    const double ut = q11 * u + q12 * v, vt = q12 * u + q22 * v;
    const double qz = g.f[o] + A.zeta[o];
    const double mu = A.alpha * A.M[g.fs + o], mv = A.alpha * A.M[2 * g.fs + o];
    const double lut = h * lFx - qz * Jv * mv;
    const double lvt = h * lFy + qz * Jv * mu;
    const double gh = ut * lFx + vt * lFy + g.grav * lK;
    const double gu = ut * lK - lAy / Jv + q11 * lut + q12 * lvt;
    const double gv = vt * lK + lAx / Jv + q12 * lut + q22 * lvt;
    const double gg[3] = {gh, gu, gv};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const long long ok = k * g.fs + o;
        if constexpr (MODE == ADJ_VJP) {
            A.out[ok] = gg[k];
        } else if constexpr (MODE == ADJ_S4) {
            const double l = A.L[ok];
            A.MU[ok] = A.a * l + A.b * gg[k];
            A.ACC[ok] = l + gg[k];
        } else if constexpr (MODE == ADJ_S32) {
            A.MU[ok] = A.a * A.L[ok] + A.b * gg[k];
            A.ACC[ok] = A.ACC[ok] + gg[k];
        } else {
            A.L[ok] = A.ACC[ok] + gg[k];
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Fused transposed RHS: one launch per evaluation.  A tile of TX x TY points of one panel
// evaluates the seeds W = [Ldiv | Lz | LGx | LGy] of its points and of the points its D^T
// columns read into LDS, then applies D^T along x (fields 0, 1, 2) and y (fields 0, 1, 3), the
// chain rule and the RK reverse-stage combination at its own points.  Panel-edge entries hold
// the transposed edge syncs of the seeds instead (as the compact buffer E of the two-launch
// path), gathered through fixed-width tables so that all loads of a point issue together.
// The rows of one D^T line held in LDS: interior columns read +-R rows; the closure columns
// c < B read rows 0 .. NR-1 and the mirrored ones rows n-NR .. n-1, so a tile near a panel
// boundary extends its range to those rows.

constexpr int ELL_S = 3, ELL_V = 4;  // widest row of the scalar sync / of the transposed vector sync

// k-major: entry k of row r at [k * rows + r]; padding entries repeat the row's first offset
// with weight 0.  s_off: field-relative point offset; v_off: point offset + comp * fs.
struct AdjEll {
    const int *s_off;
    const double *s_w;   // ELL_S x ns
    const int *v_off;
    const double *v_w;   // ELL_V x 2ns (x rows, then y rows)
};

template <int ORD, int TX, int TY>
struct AdjTile {
    static constexpr int R = SbpT<ORD>::R, B = SbpT<ORD>::B, NR = SbpT<ORD>::B + SbpT<ORD>::W;
    static constexpr int span(int t) { return (t + R + B - 1 > NR) ? t + R + B - 1 : NR; }
    static constexpr int LW = span(TX), LH = span(TY), NT = TX * TY;
    static_assert(TX >= B && TY >= B, "a tile holds the closure columns of at most one boundary");
};

// range [lo, hi) of rows that the D^T columns a0 .. ae-1 of a line of n points read
template <int ORD, int TX, int TY>
__device__ __forceinline__ void adj_rows(int a0, int ae, int n, int &lo, int &hi) {
    using Tl = AdjTile<ORD, TX, TY>;
    lo = a0 - Tl::R > 0 ? a0 - Tl::R : 0;
    hi = ae + Tl::R < n ? ae + Tl::R : n;
    if (a0 < Tl::B) {
        lo = 0;
        const int m = Tl::NR < n ? Tl::NR : n;
        hi = hi > m ? hi : m;
    }
    if (ae > n - Tl::B) {
        hi = n;
        const int m = n - Tl::NR > 0 ? n - Tl::NR : 0;
        lo = lo < m ? lo : m;
    }
}

template <int ORD, int MODE, int TX, int TY>
__global__ void __launch_bounds__(TX * TY) adj_tile(AdjGeo g, AdjArgs A, AdjEll ell) {
    using Tl = AdjTile<ORD, TX, TY>;
    __shared__ double box[4][Tl::LH][Tl::LW];
    const int n = g.n, ns = 6 * g.es;
    const int ntx = (n + TX - 1) / TX, nty = (n + TY - 1) / TY;
    const int tile = xcd_tile(blockIdx.x, 6 * ntx * nty);
    const int p = tile / (ntx * nty), rem = tile - p * (ntx * nty);
    const int x0 = (rem % ntx) * TX, y0 = (rem / ntx) * TY;
    const int xe = x0 + TX < n ? x0 + TX : n, ye = y0 + TY < n ? y0 + TY : n;
    int lx, hx, ly, hy;
    adj_rows<ORD, TX, TY>(x0, xe, n, lx, hx);
    adj_rows<ORD, TX, TY>(y0, ye, n, ly, hy);
    const double alpha = A.alpha, hxs = g.hxs;
    const double *__restrict__ y = A.y;
    const double *__restrict__ M = A.M;

    // seeds at the point (px, py) of panel p into LDS
    auto seed = [&](int px, int py) {
        const long long o = ((long long)p * n + py) * g.pitch + px;
        double w0, w1, w2, w3;
        if (px == 0 || py == 0 || px == n - 1 || py == n - 1) {
            const int sl = p * g.es + edge_slot(n, px, py);
            int so[ELL_S], vx[ELL_V], vy[ELL_V];
            double sw[ELL_S], wx[ELL_V], wy[ELL_V];
#pragma unroll
            for (int k = 0; k < ELL_S; ++k) {
                so[k] = ell.s_off[k * ns + sl];
                sw[k] = ell.s_w[k * ns + sl];
            }
#pragma unroll
            for (int k = 0; k < ELL_V; ++k) {
                vx[k] = ell.v_off[k * 2 * ns + sl];
                wx[k] = ell.v_w[k * 2 * ns + sl];
                vy[k] = ell.v_off[k * 2 * ns + ns + sl];
                wy[k] = ell.v_w[k * 2 * ns + ns + sl];
            }
            double a = 0.0, b = 0.0;
#pragma unroll
            for (int k = 0; k < ELL_S; ++k) {
                const long long q = so[k];
                const double u = y[g.fs + q], v = y[2 * g.fs + q];
                const double ut = g.Qi11[q] * u + g.Qi12[q] * v;
                const double vt = g.Qi12[q] * u + g.Qi22[q] * v;
                const double mh = alpha * M[q], mu = alpha * M[g.fs + q], mv = alpha * M[2 * g.fs + q];
                const double z = g.J[q] * (vt * mu - ut * mv);
                a += sw[k] * -mh;
                b += sw[k] * z;
            }
            const double dnm = g.J[o] * hxs;
            w0 = a / dnm;
            w1 = b / dnm;
            w2 = 0.0;
            w3 = 0.0;
#pragma unroll
            for (int k = 0; k < ELL_V; ++k) {
                w2 += wx[k] * (-(alpha * M[g.fs + vx[k]]) / hxs);
                w3 += wy[k] * (-(alpha * M[g.fs + vy[k]]) / hxs);
            }
        } else {
            const double u = y[g.fs + o], v = y[2 * g.fs + o];
            const double Jv = g.J[o];
            const double ut = g.Qi11[o] * u + g.Qi12[o] * v;
            const double vt = g.Qi12[o] * u + g.Qi22[o] * v;
            const double mh = alpha * M[o], mu = alpha * M[g.fs + o], mv = alpha * M[2 * g.fs + o];
            const double z = Jv * (vt * mu - ut * mv);
            const double dnm = Jv * hxs;
            w0 = -mh / dnm;
            w1 = z / dnm;
            w2 = -mu / hxs;
            w3 = -mv / hxs;
        }
        box[0][py - ly][px - lx] = w0;
        box[1][py - ly][px - lx] = w1;
        box[2][py - ly][px - lx] = w2;
        box[3][py - ly][px - lx] = w3;
    };

    const int tid = threadIdx.x;
    const int gx = x0 + tid % TX, gy = y0 + tid / TX;
    const bool own = gx < n && gy < n;
    if (own) seed(gx, gy);
    // halo: tile rows outside [x0, xe), then tile columns outside [y0, ye)
    const int xl = x0 - lx, xw = xl + (hx - xe), yl = y0 - ly, ncol = xe - x0;
    const int nxh = (ye - y0) * xw, nyh = (y0 - ly + hy - ye) * ncol;
    for (int q = tid; q < nxh + nyh; q += Tl::NT) {
        if (q < nxh) {
            const int r = q / xw, k = q - r * xw;
            seed(k < xl ? lx + k : xe + (k - xl), y0 + r);
        } else {
            const int q2 = q - nxh, r = q2 / ncol, k = q2 - r * ncol;
            seed(x0 + k, r < yl ? ly + r : ye + (r - yl));
        }
    }
    __syncthreads();
    if (!own) return;

    const int cx = gx - lx, cy = gy - ly;
    const V3 dx = sbp_col_T<ORD>(gx, n, [&](int r) { return V3{box[0][cy][r - lx], box[1][cy][r - lx], box[2][cy][r - lx]}; });
    const V3 dy = sbp_col_T<ORD>(gy, n, [&](int r) { return V3{box[0][r - ly][cx], box[1][r - ly][cx], box[3][r - ly][cx]}; });
    const long long o = ((long long)p * n + gy) * g.pitch + gx;
    const double Jv = g.J[o];
    const double lFx = Jv * dx.a, lFy = Jv * dy.a, lAx = Jv * dx.b, lAy = Jv * dy.b, lK = dx.c + dy.c;
    const double h = y[o], u = y[g.fs + o], v = y[2 * g.fs + o];
    const double q11 = g.Qi11[o], q12 = g.Qi12[o], q22 = g.Qi22[o];
    const double ut = q11 * u + q12 * v, vt = q12 * u + q22 * v;
    const double qz = g.f[o] + A.zeta[o];
    const double mu = alpha * M[g.fs + o], mv = alpha * M[2 * g.fs + o];
    const double lut = h * lFx - qz * Jv * mv;
    const double lvt = h * lFy + qz * Jv * mu;
    const double gh = ut * lFx + vt * lFy + g.grav * lK;
    const double gu = ut * lK - lAy / Jv + q11 * lut + q12 * lvt;
    const double gv = vt * lK + lAx / Jv + q12 * lut + q22 * lvt;
    const double gg[3] = {gh, gu, gv};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const long long ok = k * g.fs + o;
        if constexpr (MODE == ADJ_VJP) {
            A.out[ok] = gg[k];
        } else if constexpr (MODE == ADJ_S4) {
            const double l = A.L[ok];
            A.MU[ok] = A.a * l + A.b * gg[k];
            A.ACC[ok] = l + gg[k];
        } else if constexpr (MODE == ADJ_S32) {
            A.MU[ok] = A.a * A.L[ok] + A.b * gg[k];
            A.ACC[ok] = A.ACC[ok] + gg[k];
        } else {
            A.L[ok] = A.ACC[ok] + gg[k];
        }
    }
}

}  // namespace swm

// ------------------------------------------------------------------------------------------------
// Host side
namespace {

struct AdjWork {
    int es = 0, ns = 0;
    int *s_ptr = nullptr, *v_ptr = nullptr;
    long long *s_off = nullptr, *v_off = nullptr, *r_off = nullptr;
    double *s_w = nullptr, *v_w = nullptr;
    double *W = nullptr, *E = nullptr, *Z = nullptr;  // 4 fields, 4*ns, 4 fields
    int *e_soff = nullptr, *e_voff = nullptr;          // fixed-width gather tables of the fused kernel
    double *e_sw = nullptr, *e_vw = nullptr;
    int tile_x = 16;                                   // fused kernel tile width (16 rows); 0 = two launches
};

void free_adj_work(AdjWork *w) {
    if (!w) return;
    void *ptrs[] = {w->s_ptr, w->v_ptr, w->s_off, w->v_off, w->r_off, w->s_w, w->v_w,
                    w->W,     w->E,     w->Z,     w->e_soff, w->e_voff, w->e_sw, w->e_vw};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete w;
}

// CSR of the scalar sync and of the TRANSPOSED covariant vector sync
int build_csr(int N, int vector, bool transpose, int nrows, std::vector<int> &ptr, std::vector<int> &col,
              std::vector<double> &w) {
    int nnz = swmh_edge_sync_matrix(N, vector, nullptr, nullptr, nullptr, 0);
    std::vector<int> r(nnz), c(nnz);
    std::vector<double> v(nnz);
    if (swmh_edge_sync_matrix(N, vector, r.data(), c.data(), v.data(), nnz) != nnz) return -1;
    if (transpose) std::swap(r, c);
    ptr.assign(nrows + 1, 0);
    for (int k = 0; k < nnz; ++k) {
        if (r[k] < 0 || r[k] >= nrows || c[k] < 0 || c[k] >= nrows) return -1;
        ptr[r[k] + 1]++;
    }
    for (int k = 0; k < nrows; ++k) ptr[k + 1] += ptr[k];
    col.assign(nnz, 0);
    w.assign(nnz, 0.0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int k = 0; k < nnz; ++k) {
        const int at = fill[r[k]]++;
        col[at] = c[k];
        w[at] = v[k];
    }
    return 0;
}

template <typename V>
int upload_vec(V **dst, const std::vector<V> &src) {
    HIP_TRY(hipMalloc((void **)dst, src.size() * sizeof(V)));
    HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice));
    return SWM_OK;
}

int adj_min_n(int order) {
    switch (order) {
    case 42: return swm::SbpT<42>::MIN_N;
    case 43: return swm::SbpT<43>::MIN_N;
    case 63: return swm::SbpT<63>::MIN_N;
    default: return 1 << 30;
    }
}

int ensure_adj(swm_ctx *c) {
    if (c->adj) return SWM_OK;
    if (c->prec != 64) return fail(SWM_EARG, "the adjoint runs in FP64 contexts only");
    if (c->grad_order != 42 && c->grad_order != 43 && c->grad_order != 63)
        return fail(SWM_EARG, "the adjoint covers the SBP families 42, 43 and 63");
    if (c->n < adj_min_n(c->grad_order))
        return fail(SWM_EARG, "the adjoint needs N+1 >= " + std::to_string(adj_min_n(c->grad_order)) + " for this operator");
    AdjWork *w = new AdjWork();
    w->es = 4 * c->N;
    w->ns = 6 * w->es;
    std::vector<int> sp, sc, vp, vc;
    std::vector<double> sw, vw;
    int rc = SWM_OK;
    if (build_csr(c->N, 0, false, w->ns, sp, sc, sw) || build_csr(c->N, 1, true, 2 * w->ns, vp, vc, vw)) {
        free_adj_work(w);
        return fail(SWM_EARG, "edge sync matrix construction failed");
    }
    // memory offsets of the slot points (field-relative) and of the vector entries (into W)
    auto point_off = [&](int slot) {
        const int n = c->n, p = slot / w->es, sl = slot % w->es;
        int i, j;
        if (sl < n) { i = sl; j = 0; }
        else if (sl < 2 * n) { i = sl - n; j = n - 1; }
        else if (sl < 3 * n - 2) { i = 0; j = sl - 2 * n + 1; }
        else { i = n - 1; j = sl - (3 * n - 2) + 1; }
        return ((long long)p * n + j) * c->pitch + i;
    };
    std::vector<long long> so(sc.size()), vo(vc.size()), ro(w->ns);
    for (size_t k = 0; k < sc.size(); ++k) so[k] = point_off(sc[k]);
    for (size_t k = 0; k < vc.size(); ++k) {
        const int comp = vc[k] >= w->ns ? 1 : 0;
        vo[k] = comp * c->fs + point_off(vc[k] - comp * w->ns);
    }
    for (int t = 0; t < w->ns; ++t) ro[t] = point_off(t);
    // fixed-width (k-major) copies of both tables for the fused kernel
    auto ell = [](const std::vector<int> &ptr, const std::vector<long long> &off, const std::vector<double> &wt,
                  int width, std::vector<int> &eo, std::vector<double> &ew) {
        const int rows = (int)ptr.size() - 1;
        eo.assign((size_t)width * rows, 0);
        ew.assign((size_t)width * rows, 0.0);
        for (int r = 0; r < rows; ++r) {
            const int cnt = ptr[r + 1] - ptr[r];
            if (cnt < 1 || cnt > width) return -1;
            for (int k = 0; k < width; ++k) {
                const int at = ptr[r] + (k < cnt ? k : 0);
                eo[(size_t)k * rows + r] = (int)off[at];
                ew[(size_t)k * rows + r] = k < cnt ? wt[at] : 0.0;
            }
        }
        return 0;
    };
    std::vector<int> eso, evo;
    std::vector<double> esw, evw;
    if (2 * (long long)c->fs > (1LL << 31) - 1 || ell(sp, so, sw, swm::ELL_S, eso, esw) ||
        ell(vp, vo, vw, swm::ELL_V, evo, evw)) {
        free_adj_work(w);
        return fail(SWM_EARG, "edge sync gather tables do not fit the fused adjoint kernel");
    }
    if ((rc = upload_vec(&w->s_ptr, sp)) || (rc = upload_vec(&w->s_off, so)) || (rc = upload_vec(&w->s_w, sw)) ||
        (rc = upload_vec(&w->v_ptr, vp)) || (rc = upload_vec(&w->v_off, vo)) || (rc = upload_vec(&w->v_w, vw)) ||
        (rc = upload_vec(&w->r_off, ro)) || (rc = upload_vec(&w->e_soff, eso)) || (rc = upload_vec(&w->e_sw, esw)) ||
        (rc = upload_vec(&w->e_voff, evo)) || (rc = upload_vec(&w->e_vw, evw))) {
        free_adj_work(w);
        return rc;
    }
    const size_t fb = c->field_bytes();
    if (hipMalloc((void **)&w->W, 4 * fb) != hipSuccess || hipMalloc((void **)&w->E, 4 * sizeof(double) * w->ns) != hipSuccess ||
        hipMalloc((void **)&w->Z, 4 * fb) != hipSuccess || hipMemset(w->W, 0, 4 * fb) != hipSuccess ||
        hipMemset(w->Z, 0, 4 * fb) != hipSuccess) {
        free_adj_work(w);
        return fail(SWM_EHIP, "adjoint work allocation failed");
    }
    // null-stream copies and clears complete before the context stream uses the buffers
    if (hipDeviceSynchronize() != hipSuccess) {
        free_adj_work(w);
        return fail(SWM_EHIP, "adjoint set-up synchronisation failed");
    }
    c->adj = w;
    return SWM_OK;
}

swm::AdjGeo adj_geo(const swm_ctx *c) {
    swm::AdjGeo g;
    g.J = c->J64; g.Qi11 = c->Qi64[0]; g.Qi12 = c->Qi64[1]; g.Qi22 = c->Qi64[2]; g.f = c->f64;
    g.n = c->n; g.pitch = c->pitch; g.es = 4 * c->N; g.fs = c->fs;
    g.hxs = c->hx * c->scale; g.grav = c->grav;
    return g;
}

template <int ORD>
hipError_t launch_adj_main(int mode, const swm::AdjGeo &g, const swm::AdjArgs &a, hipStream_t s) {
    const long long total = 6LL * g.n * g.n;
    const dim3 blk(256), grd((unsigned)((total + 255) / 256));
    switch (mode) {
    case swm::ADJ_VJP: hipLaunchKernelGGL((swm::adj_main<ORD, swm::ADJ_VJP>), grd, blk, 0, s, g, a); break;
    case swm::ADJ_S4: hipLaunchKernelGGL((swm::adj_main<ORD, swm::ADJ_S4>), grd, blk, 0, s, g, a); break;
    case swm::ADJ_S32: hipLaunchKernelGGL((swm::adj_main<ORD, swm::ADJ_S32>), grd, blk, 0, s, g, a); break;
    default: hipLaunchKernelGGL((swm::adj_main<ORD, swm::ADJ_S1>), grd, blk, 0, s, g, a); break;
    }
    return hipGetLastError();
}

template <int ORD, int TX>
hipError_t launch_adj_tile(int mode, const swm::AdjGeo &g, const swm::AdjArgs &a, const swm::AdjEll &e, hipStream_t s) {
    constexpr int TY = 16;
    const int nb = 6 * ((g.n + TX - 1) / TX) * ((g.n + TY - 1) / TY);
    const dim3 blk(TX * TY), grd(nb);
    switch (mode) {
    case swm::ADJ_VJP: hipLaunchKernelGGL((swm::adj_tile<ORD, swm::ADJ_VJP, TX, TY>), grd, blk, 0, s, g, a, e); break;
    case swm::ADJ_S4: hipLaunchKernelGGL((swm::adj_tile<ORD, swm::ADJ_S4, TX, TY>), grd, blk, 0, s, g, a, e); break;
    case swm::ADJ_S32: hipLaunchKernelGGL((swm::adj_tile<ORD, swm::ADJ_S32, TX, TY>), grd, blk, 0, s, g, a, e); break;
    default: hipLaunchKernelGGL((swm::adj_tile<ORD, swm::ADJ_S1, TX, TY>), grd, blk, 0, s, g, a, e); break;
    }
    return hipGetLastError();
}

template <int ORD>
hipError_t launch_adj_tile_x(int tx, int mode, const swm::AdjGeo &g, const swm::AdjArgs &a, const swm::AdjEll &e,
                             hipStream_t s) {
    return tx == 32 ? launch_adj_tile<ORD, 32>(mode, g, a, e, s) : launch_adj_tile<ORD, 16>(mode, g, a, e, s);
}

// one transposed RHS evaluation at state y with seed alpha*M; zeta = synced vorticity of y.
// The fused kernel reads M at its neighbours' points: M must not be one of its outputs.
int enqueue_vjp(swm_ctx *c, int mode, const void *y, const double *zeta, const double *M, double alpha,
                swm::AdjArgs a) {
    AdjWork *w = c->adj;
    const swm::AdjGeo g = adj_geo(c);
    hipError_t e;
    if (w->tile_x) {
        a.y = (const double *)y;
        a.M = M;
        a.alpha = alpha;
        a.zeta = zeta;
        const swm::AdjEll ell{w->e_soff, w->e_sw, w->e_voff, w->e_vw};
        switch (c->grad_order) {
        case 42: e = launch_adj_tile_x<42>(w->tile_x, mode, g, a, ell, c->stream); break;
        case 43: e = launch_adj_tile_x<43>(w->tile_x, mode, g, a, ell, c->stream); break;
        default: e = launch_adj_tile_x<63>(w->tile_x, mode, g, a, ell, c->stream); break;
        }
        if (e != hipSuccess) return fail(SWM_EHIP, std::string("adjoint launch: ") + hipGetErrorString(e));
        return SWM_OK;
    }
    const int nb_edge = (3 * w->ns + 255) / 256;
    const int nb_pre = (int)std::min<long long>((6LL * g.n * g.n + 255) / 256, 2048);
    hipLaunchKernelGGL(swm::adj_pre_edge, dim3(nb_edge + nb_pre), dim3(256), 0, c->stream, g, (const double *)y, M,
                       alpha, w->W, nb_edge, w->s_ptr, w->s_off, w->s_w, w->v_ptr, w->v_off, w->v_w, w->r_off, w->E);
    a.y = (const double *)y;
    a.E = w->E;
    a.M = M;
    a.alpha = alpha;
    a.zeta = zeta;
    a.W = w->W;
    switch (c->grad_order) {
    case 42: e = launch_adj_main<42>(mode, g, a, c->stream); break;
    case 43: e = launch_adj_main<43>(mode, g, a, c->stream); break;
    default: e = launch_adj_main<63>(mode, g, a, c->stream); break;
    }
    if (e != hipSuccess) return fail(SWM_EHIP, std::string("adjoint launch: ") + hipGetErrorString(e));
    return SWM_OK;
}

}  // namespace

struct swm_adj {
    swm_ctx *c = nullptr;
    int nsteps = 0, every = 0, nck = 0;
    double dt = 0.0;
    bool have_forward = false;
    void *work = nullptr;           // forward state
    std::vector<void *> ck;         // checkpoints: state at steps 0, every, 2*every, ...
    // one recomputed segment: per step s the start state, the stage inputs y1, y2, y3 and the
    // synced vorticity of y0..y3 (4 fields), all written by the forward stage kernels
    std::vector<void *> seg, ys, zs;
    int held = -1;                  // index of the segment whose data seg / ys / zs hold
    void *sacc = nullptr;           // accumulator of the recomputed stages
    void *tail = nullptr;           // end state of a segment's last step (not kept)
    void *L = nullptr, *MU = nullptr, *MU2 = nullptr, *ACC = nullptr;  // MU / MU2 alternate between stages
};

namespace {

void free_tape(swm_adj *t) {
    if (!t) return;
    std::vector<void *> all = {t->work, t->sacc, t->tail, t->L, t->MU, t->MU2, t->ACC};
    for (auto *v : {&t->ck, &t->seg, &t->ys, &t->zs}) all.insert(all.end(), v->begin(), v->end());
    for (void *p : all)
        if (p) (void)hipFree(p);
    delete t;
}

int copy_state(swm_ctx *c, void *dst, const void *src) {
    HIP_TRY(hipMemcpyAsync(dst, src, 3 * c->field_bytes(), hipMemcpyDeviceToDevice, c->stream));
    return SWM_OK;
}

// recompute `len` steps from seg[0], keeping every step's stage inputs and vorticities
int recompute_segment(swm_adj *t, int len) {
    swm_ctx *c = t->c;
    const double dt = t->dt, c1 = 0.5 * dt, c2 = dt / 6.0, c3 = dt / 3.0;
    const size_t fe = (size_t)c->fs;
    int rc;
    for (int s = 0; s < len; ++s) {
        void *y0 = t->seg[s], **y = &t->ys[3 * s];
        double *Z = (double *)t->zs[s];
        void *next = s + 1 < len ? t->seg[s + 1] : t->tail;
        if ((rc = run_stage(c, swm::MODE_STAGE1, y0, y[0], nullptr, t->sacc, c1, c2, Z))) return rc;
        if ((rc = run_stage(c, swm::MODE_STAGE23, y[0], y[1], y0, t->sacc, c1, c3, Z + fe))) return rc;
        if ((rc = run_stage(c, swm::MODE_STAGE23, y[1], y[2], y0, t->sacc, dt, c3, Z + 2 * fe))) return rc;
        if ((rc = run_stage(c, swm::MODE_STAGE4, y[2], next, nullptr, t->sacc, c2, 0.0, Z + 3 * fe))) return rc;
    }
    return SWM_OK;
}

// reverse of step s of the recomputed segment; L holds dJ/dy(step end) on entry and
// dJ/dy(step start) on exit
int reverse_step(swm_adj *t, int s) {
    swm_ctx *c = t->c;
    const double dt = t->dt, c1 = 0.5 * dt, c2 = dt / 6.0, c3 = dt / 3.0;
    const size_t fe = (size_t)c->fs;
    const double *Z = (const double *)t->zs[s];
    const void *yy[4] = {t->seg[s], t->ys[3 * s], t->ys[3 * s + 1], t->ys[3 * s + 2]};
    int rc;
    swm::AdjArgs a{};
    double *mu0 = (double *)t->MU, *mu1 = (double *)t->MU2;
    a.L = (double *)t->L;
    a.ACC = (double *)t->ACC;
    a.MU = mu0;
    a.a = c3; a.b = dt;
    if ((rc = enqueue_vjp(c, swm::ADJ_S4, yy[3], Z + 3 * fe, (const double *)t->L, c2, a))) return rc;
    a.MU = mu1;
    a.a = c3; a.b = c1;
    if ((rc = enqueue_vjp(c, swm::ADJ_S32, yy[2], Z + 2 * fe, mu0, 1.0, a))) return rc;
    a.MU = mu0;
    a.a = c2; a.b = c1;
    if ((rc = enqueue_vjp(c, swm::ADJ_S32, yy[1], Z + 1 * fe, mu1, 1.0, a))) return rc;
    return enqueue_vjp(c, swm::ADJ_S1, yy[0], Z, mu0, 1.0, a);
}

}  // namespace

extern "C" int swm_rhs_vjp(swm_ctx *c, const void *y, const void *mu, void *out) {
    if (!c || !y || !mu || !out) return fail(SWM_EARG, "NULL argument");
    if (out == y || out == mu) return fail(SWM_EARG, "out must differ from y and mu");
    int rc = ensure_adj(c);
    if (rc) return rc;
    double *Z = c->adj->Z;
    if ((rc = run_stage(c, swm::MODE_CURL, y, Z, nullptr, nullptr, 0.0, 0.0))) return rc;
    swm::AdjArgs a{};
    a.out = (double *)out;
    return enqueue_vjp(c, swm::ADJ_VJP, y, Z, (const double *)mu, 1.0, a);
}

extern "C" int swm_adj_set_tiling(swm_ctx *c, int tile_x) {
    if (!c) return fail(SWM_EARG, "NULL argument");
    if (tile_x != 0 && tile_x != 16 && tile_x != 32) return fail(SWM_EARG, "adjoint tile width must be 0, 16 or 32");
    int rc = ensure_adj(c);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->adj->tile_x = tile_x;
    return SWM_OK;
}

extern "C" int swm_adj_create(swm_ctx *c, int nsteps, double dt, int checkpoint_every, swm_adj **out) {
    if (!c || !out) return fail(SWM_EARG, "NULL argument");
    *out = nullptr;
    if (nsteps < 1 || checkpoint_every < 1) return fail(SWM_EARG, "nsteps and checkpoint_every must be >= 1");
    int rc = ensure_adj(c);
    if (rc) return rc;
    swm_adj *t = new swm_adj();
    t->c = c;
    t->nsteps = nsteps;
    t->dt = dt;
    t->every = checkpoint_every < nsteps ? checkpoint_every : nsteps;
    t->nck = (nsteps + t->every - 1) / t->every;
    const size_t sb = 3 * c->field_bytes(), zb = 4 * c->field_bytes();
    auto alloc = [&](void **p, size_t b) { return hipMalloc(p, b) == hipSuccess && hipMemset(*p, 0, b) == hipSuccess; };
    bool ok = alloc(&t->work, sb) && alloc(&t->sacc, sb) && alloc(&t->tail, sb) && alloc(&t->L, sb) &&
              alloc(&t->MU, sb) && alloc(&t->MU2, sb) && alloc(&t->ACC, sb);
    t->ck.assign(t->nck, nullptr);
    for (int k = 0; k < t->nck && ok; ++k) ok = alloc(&t->ck[k], sb);
    t->seg.assign(t->every, nullptr);
    t->ys.assign(3 * t->every, nullptr);
    t->zs.assign(t->every, nullptr);
    for (int k = 0; k < t->every && ok; ++k) ok = alloc(&t->seg[k], sb) && alloc(&t->zs[k], zb);
    for (int k = 0; k < 3 * t->every && ok; ++k) ok = alloc(&t->ys[k], sb);
    if (!ok || hipDeviceSynchronize() != hipSuccess) {  // null-stream clears complete first
        free_tape(t);
        return fail(SWM_EHIP, "adjoint tape allocation failed");
    }
    *out = t;
    return SWM_OK;
}

extern "C" int swm_adj_destroy(swm_adj *t) {
    if (!t) return fail(SWM_EARG, "tape is NULL");
    (void)hipStreamSynchronize(t->c->stream);
    free_tape(t);
    return SWM_OK;
}

extern "C" int swm_adj_forward(swm_adj *t, const void *y0, void *yT) {
    if (!t || !y0 || !yT) return fail(SWM_EARG, "NULL argument");
    swm_ctx *c = t->c;
    int rc;
    t->held = -1;
    if ((rc = copy_state(c, t->work, y0))) return rc;
    for (int k = 0; k < t->nck; ++k) {
        if ((rc = copy_state(c, t->ck[k], t->work))) return rc;
        const int len = std::min(t->every, t->nsteps - k * t->every);
        if (k + 1 < t->nck) {
            if ((rc = swm_rk4_steps(c, t->work, t->dt, len))) return rc;
            continue;
        }
        // the last segment is the first one the reverse sweep needs: record its stage
        // inputs and vorticities now instead of recomputing them there
        if ((rc = copy_state(c, t->seg[0], t->work))) return rc;
        if ((rc = recompute_segment(t, len))) return rc;
        if ((rc = copy_state(c, t->work, t->tail))) return rc;
        t->held = k;
    }
    if ((rc = copy_state(c, yT, t->work))) return rc;
    t->have_forward = true;
    return SWM_OK;
}

extern "C" int swm_adj_backward(swm_adj *t, const void *lamT, void *lam0) {
    if (!t || !lamT || !lam0) return fail(SWM_EARG, "NULL argument");
    if (!t->have_forward) return fail(SWM_ESTATE, "swm_adj_backward needs swm_adj_forward first");
    swm_ctx *c = t->c;
    int rc;
    if ((rc = copy_state(c, t->L, lamT))) return rc;
    for (int k = t->nck - 1; k >= 0; --k) {
        const int len = std::min(t->every, t->nsteps - k * t->every);
        if (t->held != k) {
            t->held = -1;
            if ((rc = copy_state(c, t->seg[0], t->ck[k]))) return rc;
            if ((rc = recompute_segment(t, len))) return rc;
            t->held = k;
        }
        for (int s = len - 1; s >= 0; --s)
            if ((rc = reverse_step(t, s))) return rc;
    }
    return copy_state(c, lam0, t->L);
}

// dense (6, n, n) device arrays of the context's element type <-> pitched state
extern "C" int swm_state_copy_from_device(swm_ctx *c, void *state, const void *h, const void *u, const void *v) {
    if (!c || !state || !h || !u || !v) return fail(SWM_EARG, "NULL argument");
    const void *src[3] = {h, u, v};
    const size_t e = c->esz();
    for (int k = 0; k < 3; ++k)
        HIP_TRY(hipMemcpy2DAsync((char *)state + k * c->field_bytes(), c->pitch * e, src[k], c->n * e, c->n * e,
                                 (size_t)6 * c->n, hipMemcpyDeviceToDevice, c->stream));
    return SWM_OK;
}

extern "C" int swm_state_copy_to_device(swm_ctx *c, const void *state, void *h, void *u, void *v) {
    if (!c || !state || !h || !u || !v) return fail(SWM_EARG, "NULL argument");
    void *dst[3] = {h, u, v};
    const size_t e = c->esz();
    for (int k = 0; k < 3; ++k)
        HIP_TRY(hipMemcpy2DAsync(dst[k], c->n * e, (const char *)state + k * c->field_bytes(), c->pitch * e, c->n * e,
                                 (size_t)6 * c->n, hipMemcpyDeviceToDevice, c->stream));
    return SWM_OK;
}
