// Stage kernel of the sbp21 operator family: gradient_ah21_sbp_ecs (d21 rows + covariant
// edge sync) with divergence_ah2 / curl_divergence_ah2, the hand-written second-order Ah
// divergence whose edge and corner formulas read the neighbour panel's vector components at
// halo depths 0 and 1 (differential_operators/horizontal/divergence/div_ah2_mod.f90:46-139).
// Halo values follow the vector exchange convention (parallel/buffer_mod.f90:171-277).
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {

template <typename T, int TX, int TY>
struct Smem21 {
    // in-panel cells of the tile and a 1-point halo: K, J*?, U_d = h*ut, V_d = h*vt,
    // U_c = v/J, V_c = -u/J, J
    T K[TY + 2][TX + 2], Ud[TY + 2][TX + 2], Vd[TY + 2][TX + 2], Uc[TY + 2][TX + 2], Vc[TY + 2][TX + 2],
        J[TY + 2][TX + 2];
    // ghost strips (depth 0/1): K and the vector pairs mapped to this panel's components
    T gx[2][2][TY + 2][5];  // [side L/R][depth][row - ty0 + 1][K, Ud, Vd, Uc, Vc]
    T gy[2][TX + 2][5];     // [depth][col - tx0 + 1][...]
    T nrx[2][TY][2], nry[TX][2];  // neighbour raw gx, gy (gradient sync)
};

template <typename T>
struct Pt21 {
    T K, Ud, Vd, Uc, Vc, J;
};

template <bool OTF, typename T>
__device__ __forceinline__ Pt21<T> point21(const DevGeo<T> &g, const T *y, const T *hs, int p, int i, int j) {
    const long long id = pidx(g, p, i, j);
    T J, q11, q12, q22;
    if constexpr (OTF) {
        const Metric m = metric_otf(g.tanA, i, j);
        J = (T)m.J; q11 = (T)m.Qi11; q12 = (T)m.Qi12; q22 = (T)m.Qi22;
    } else {
        J = g.J[id]; q11 = g.Qi11[id]; q12 = g.Qi12[id]; q22 = g.Qi22[id];
    }
    const T h = y[id], u = y[id + g.fs], v = y[id + 2 * g.fs];
    const T ut = q11 * u + q12 * v;
    const T vt = q12 * u + q22 * v;
    Pt21<T> r;
    T K = (u * ut + v * vt) / T(2);
    r.K = K + g.grav * h + g.grav * (hs ? hs[id] : T(0));
    r.Ud = h * ut;
    r.Vd = h * vt;
    r.Uc = v / J;
    r.Vc = -u / J;
    r.J = J;
    return r;
}

template <typename T, int MODE, int TX, int TY, bool OTF>
__global__ void __launch_bounds__(TX * TY) stage21_kernel(DevGeo<T> g, StageArgs<T> a) {
    constexpr int NT = TX * TY;
    __shared__ Smem21<T, TX, TY> sm;
    const int n = g.n, last = n - 1;
    const int ntx = (n + TX - 1) / TX, nty = (n + TY - 1) / TY;
    const int tile = xcd_tile(blockIdx.x, 6 * ntx * nty);
    const int p = tile / (ntx * nty);
    const int rem = tile - p * ntx * nty;
    const int ty0 = (rem / ntx) * TY, tx0 = (rem % ntx) * TX;
    const int tid = threadIdx.x;
    const T *y = a.y;

    // in-panel cross region (1-point halo)
    for (int c = tid; c < (TY + 2) * (TX + 2); c += NT) {
        const int ly = c / (TX + 2), lx = c - ly * (TX + 2);
        const bool row = ly >= 1 && ly <= TY, col = lx >= 1 && lx <= TX;
        const int i = tx0 - 1 + lx, j = ty0 - 1 + ly;
        if (!(row || col) || i < 0 || i >= n || j < 0 || j >= n) continue;
        const Pt21<T> q = point21<OTF>(g, y, a.hs, p, i, j);
        sm.K[ly][lx] = q.K; sm.Ud[ly][lx] = q.Ud; sm.Vd[ly][lx] = q.Vd;
        sm.Uc[ly][lx] = q.Uc; sm.Vc[ly][lx] = q.Vc; sm.J[ly][lx] = q.J;
    }
    const bool eL = tx0 == 0, eR = tx0 + TX >= n, eB = ty0 == 0, eT = ty0 + TY >= n;
    const bool edge_tile = eL || eR || eB || eT;
    const int ey = eB ? 2 : 3;
    if (edge_tile) {
        constexpr int NXS = 2 * (TY + 2), NYS = 2 * (TX + 2);
        for (int c = tid; c < 2 * NXS + NYS; c += NT) {
            int slot, e, d, idx, kk;
            if (c < 2 * NXS) {
                slot = c / NXS;
                const int r = c - slot * NXS;
                if (!(slot == 0 ? eL : eR)) continue;
                e = slot; d = r / (TY + 2); idx = r % (TY + 2); kk = ty0 - 1 + idx;
            } else {
                const int r = c - 2 * NXS;
                if (!(eB || eT)) continue;
                slot = 2; e = ey; d = r / (TX + 2); idx = r % (TX + 2); kk = tx0 - 1 + idx;
            }
            if (kk < 0 || kk >= n) continue;
            const int *ei = g.einfo + (p * 4 + e) * 8;
            const int nidx = ei[2] == 0 ? d : n - 1 - d, tidx = ei[3] + ei[4] * kk;
            const int qi = ei[1] ? nidx : tidx, qj = ei[1] ? tidx : nidx;
            const Pt21<T> q = point21<OTF>(g, y, a.hs, ei[0], qi, qj);
            T *dst = slot < 2 ? sm.gx[slot][d][idx] : sm.gy[d][idx];
            dst[0] = q.K;
            map_vec(ei[5], q.Ud, q.Vd, dst[1], dst[2]);
            map_vec(ei[5], q.Uc, q.Vc, dst[3], dst[4]);
        }
    }
    __syncthreads();

    // neighbour gradient rows at the shared points (d21), from the ghost K
    if (edge_tile) {
        const int task = tid;
        if (task < 2 * TY + TX) {
            int slot, k, kloc;
            if (task < 2 * TY) { slot = task / TY; kloc = task % TY; k = ty0 + kloc; }
            else { slot = 2; kloc = task - 2 * TY; k = tx0 + kloc; }
            const bool active = slot == 0 ? eL : slot == 1 ? eR : (eB || eT);
            if (active && k < n) {
                const int e = slot < 2 ? slot : ey;
                const int *ei = g.einfo + (p * 4 + e) * 8;
                const int k0 = (slot < 2 ? ty0 : tx0) - 1;
                auto d0 = [&](int kk, int f) -> T {
                    if (f == 5) return T(1);
                    if (f != 0) return T(0);
                    return slot < 2 ? sm.gx[slot][0][kk - k0][0] : sm.gy[0][kk - k0][0];
                };
                auto dn = [&](int d, int f3) -> T {
                    if (f3 != 0) return T(0);
                    return slot < 2 ? sm.gx[slot][d][k - k0][0] : sm.gy[d][k - k0][0];
                };
                T r[4];
                neighbour_raw<21, T>(g, edge_desc(ei), k, d0, dn, r);
                if (slot < 2) { sm.nrx[slot][kloc][0] = r[0]; sm.nrx[slot][kloc][1] = r[1]; }
                else { sm.nry[kloc][0] = r[0]; sm.nry[kloc][1] = r[1]; }
            }
        }
        __syncthreads();
    }

    const int tx = tid % TX, ty = tid / TX;
    const int i = tx0 + tx, j = ty0 + ty;
    if (i >= n || j >= n) return;
    const int lx = tx + 1, ly = ty + 1;
    // gradient_ah21_sbp_ecs raw values
    const T dxK = sbp_row<21, T>(i, n, [&](int ai) { return sm.K[ly][ai - tx0 + 1]; });
    const T dyK = sbp_row<21, T>(j, n, [&](int aj) { return sm.K[aj - ty0 + 1][lx]; });
    T gx = dxK / g.hxs;
    T gy = dyK / g.hxs;
    const bool L = i == 0, Rt = i == last, B = j == 0, Tp = j == last;
    if (L || Rt || B || Tp) {
        auto hx_ = [&](int slot, T &V_, T &U_) {
            const T *r = sm.nrx[slot][ty];
            map_vec(g.einfo[(p * 4 + slot) * 8 + 5], r[0], r[1], V_, U_);
        };
        auto hy_ = [&](T &V_, T &U_) {
            const T *r = sm.nry[tx];
            map_vec(g.einfo[(p * 4 + ey) * 8 + 5], r[0], r[1], V_, U_);
        };
        const T U = gy, V = gx;  // (ecs_Ah_vec_sync_mod.f90:51-53)
        T Ux, Vx, Uy, Vy;
        if ((L || Rt) && (B || Tp)) {
            hx_(L ? 0 : 1, Vx, Ux);
            hy_(Vy, Uy);
            const T qy = B ? g.qb[i] : g.qt[i];
            const T qx = L ? g.ql[j] : g.qr[j];
            gy = (U + Ux + Uy + qy * Vy) / T(3);
            gx = (V + Vx + Vy + qx * Ux) / T(3);
        } else if (L) {
            hx_(0, Vx, Ux);
            gy = T(0.5) * (Ux + U);
            gx = T(0.5) * (V + Vx + g.ql[j] * Ux);
        } else if (Rt) {
            hx_(1, Vx, Ux);
            gy = T(0.5) * (U + Ux);
            gx = T(0.5) * (V + Vx + g.qr[j] * Ux);
        } else if (B) {
            hy_(Vy, Uy);
            gx = T(0.5) * (Vy + V);
            gy = T(0.5) * (U + Uy + g.qb[i] * Vy);
        } else {
            hy_(Vy, Uy);
            gx = T(0.5) * (V + Vy);
            gy = T(0.5) * (U + Uy + g.qt[i] * Vy);
        }
    }

    // divergence_ah2 of a vector pair (fU, fV select the field), div_ah2_mod.f90:59-133
    const T hxs2 = g.hx, sc = g.scale;
    auto div_ah2 = [&](int fU, int fV) -> T {
        auto P = [&](int f, int ai, int aj) -> T {
            const int yy = aj - ty0 + 1, xx = ai - tx0 + 1;
            return f == 1 ? sm.Ud[yy][xx] : f == 2 ? sm.Vd[yy][xx] : f == 3 ? sm.Uc[yy][xx] : sm.Vc[yy][xx];
        };
        auto Jp = [&](int ai, int aj) -> T { return sm.J[aj - ty0 + 1][ai - tx0 + 1]; };
        auto HX = [&](int side, int d, int aj, int f) -> T { return sm.gx[side][d][aj - ty0 + 1][f]; };
        auto HY = [&](int d, int ai, int f) -> T { return sm.gy[d][ai - tx0 + 1][f]; };
        const int U = fU, V = fV;
        if (L && B) {
            return Jp(1, 0) * (P(U, 1, 0) - HX(0, 1, 0, U) + P(V, 0, 1) - HY(1, 0, V) + HX(0, 0, 1, V) + HY(0, 1, U)) /
                   (T(3) * Jp(0, 0) * hxs2 * sc);
        }
        if (Rt && B) {
            return Jp(last - 1, 0) *
                   (HX(1, 1, 0, U) - P(U, last - 1, 0) + P(V, last, 1) - HY(1, last, V) + HX(1, 0, 1, V) -
                    HY(0, last - 1, U)) /
                   (T(3) * Jp(last, 0) * hxs2 * sc);
        }
        if (Rt && Tp) {
            return Jp(last - 1, last) *
                   (-HY(0, last - 1, U) + HY(1, last, V) + -P(U, last - 1, last) - P(V, last, last - 1) +
                    HX(1, 1, last, U) - HX(1, 0, last - 1, V)) /
                   (T(3) * Jp(last, last) * hxs2 * sc);
        }
        if (L && Tp) {
            return Jp(1, last) *
                   (HY(0, 1, U) + HY(1, 0, V) + P(U, 1, last) - P(V, 0, last - 1) + -HX(0, 1, last, U) -
                    HX(0, 0, last - 1, V)) /
                   (T(3) * Jp(0, last) * hxs2 * sc);
        }
        if (B) {
            return (T(0.5) * Jp(i + 1, 0) * (P(U, i + 1, 0) + HY(0, i + 1, U)) -
                    T(0.5) * Jp(i - 1, 0) * (P(U, i - 1, 0) + HY(0, i - 1, U)) +
                    Jp(i, 1) * (P(V, i, 1) - HY(1, i, V))) /
                   (T(2) * Jp(i, 0) * hxs2 * sc);
        }
        if (Tp) {
            return (T(0.5) * Jp(i + 1, last) * (P(U, i + 1, last) + HY(0, i + 1, U)) -
                    T(0.5) * Jp(i - 1, last) * (P(U, i - 1, last) + HY(0, i - 1, U)) +
                    Jp(i, last - 1) * (HY(1, i, V) - P(V, i, last - 1))) /
                   (T(2) * Jp(i, last) * hxs2 * sc);
        }
        if (L) {
            return (T(0.5) * Jp(0, j + 1) * (P(V, 0, j + 1) + HX(0, 0, j + 1, V)) -
                    T(0.5) * Jp(0, j - 1) * (P(V, 0, j - 1) + HX(0, 0, j - 1, V)) +
                    Jp(1, j) * (P(U, 1, j) - HX(0, 1, j, U))) /
                   (T(2) * Jp(0, j) * hxs2 * sc);
        }
        if (Rt) {
            return (T(0.5) * Jp(last, j + 1) * (P(V, last, j + 1) + HX(1, 0, j + 1, V)) -
                    T(0.5) * Jp(last, j - 1) * (P(V, last, j - 1) + HX(1, 0, j - 1, V)) +
                    Jp(last - 1, j) * (HX(1, 1, j, U) - P(U, last - 1, j))) /
                   (T(2) * Jp(last, j) * hxs2 * sc);
        }
        return (Jp(i + 1, j) * P(U, i + 1, j) - Jp(i - 1, j) * P(U, i - 1, j) + Jp(i, j + 1) * P(V, i, j + 1) -
                Jp(i, j - 1) * P(V, i, j - 1)) /
               (T(2) * Jp(i, j) * hxs2 * sc);
    };
    const T dv = div_ah2(1, 2);
    const T cu = div_ah2(3, 4);
    const long long id = pidx(g, p, i, j);
    T q11, q12, q22;
    if constexpr (OTF) {
        const Metric m = metric_otf(g.tanA, i, j);
        q11 = (T)m.Qi11; q12 = (T)m.Qi12; q22 = (T)m.Qi22;
    } else {
        q11 = g.Qi11[id]; q12 = g.Qi12[id]; q22 = g.Qi22[id];
    }
    finish_point<T, MODE>(g, a, id, a.y[id], a.y[id + g.fs], a.y[id + 2 * g.fs], sm.J[ly][lx], q11, q12, q22, gx, gy, dv,
                          cu, prefetch_point<T, MODE>(g, a, id));
}

template <typename T, int TY, bool OTF>
hipError_t launch21(int mode, const DevGeo<T> &g, const StageArgs<T> &a, hipStream_t s) {
    constexpr int TX = 64;
    const int ntx = (g.n + TX - 1) / TX, nty = (g.n + TY - 1) / TY;
    const dim3 grid(6 * ntx * nty), block(TX * TY);
    switch (mode) {
    case MODE_RHS: hipLaunchKernelGGL((stage21_kernel<T, MODE_RHS, TX, TY, OTF>), grid, block, 0, s, g, a); break;
    case MODE_STAGE1: hipLaunchKernelGGL((stage21_kernel<T, MODE_STAGE1, TX, TY, OTF>), grid, block, 0, s, g, a); break;
    case MODE_STAGE23: hipLaunchKernelGGL((stage21_kernel<T, MODE_STAGE23, TX, TY, OTF>), grid, block, 0, s, g, a); break;
    case MODE_STAGE4: hipLaunchKernelGGL((stage21_kernel<T, MODE_STAGE4, TX, TY, OTF>), grid, block, 0, s, g, a); break;
    case MODE_CURL: hipLaunchKernelGGL((stage21_kernel<T, MODE_CURL, TX, TY, OTF>), grid, block, 0, s, g, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_stage_o21(const StageLaunch &r) {
    if (r.prec == 64) {
        if (r.ty == 16)
            return r.otf ? launch21<double, 16, true>(r.mode, r.gd, r.ad, r.stream)
                         : launch21<double, 16, false>(r.mode, r.gd, r.ad, r.stream);
        return r.otf ? launch21<double, 8, true>(r.mode, r.gd, r.ad, r.stream)
                     : launch21<double, 8, false>(r.mode, r.gd, r.ad, r.stream);
    }
    if (r.ty == 16) return launch21<float, 16, false>(r.mode, r.gf, r.af, r.stream);
    return launch21<float, 8, false>(r.mode, r.gf, r.af, r.stream);
}

}  // namespace swm
