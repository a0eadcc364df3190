// Fused RK-stage kernel of the Ah-grid vector-invariant shallow-water model (see
// swm_kernels.h for the structure) and the small kernels around it.
#include <hip/hip_runtime.h>

#include "swm_kernels.h"

namespace swm {

// Tile of TX x TY output points per workgroup of TX*(TY/RPT) threads; LDS holds the five
// stencil inputs on the tile plus a halo of REACH points in x and y (cross-shaped region).
template <int ORD, typename T, int TX, int TY, int RPT>
struct TileCfg {
    static constexpr int R = Sbp<ORD>::REACH;
    static constexpr int LX = TX + 2 * R;
    static constexpr int LY = TY + 2 * R;
    static constexpr int THREADS = TX * (TY / RPT);
};

// XCD-aware mapping: consecutive block ids go round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so that tiles sharing halo rows share an L2.
__device__ __forceinline__ int xcd_tile(int bid, int nblocks) {
    const int q = nblocks / 8, r = nblocks % 8;
    const int xcd = bid % 8, local = bid / 8;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

template <int ORD, typename T>
__device__ __forceinline__ void edge_sync(const DevGeo<T> &g, const T *y, const T *hs, int p, int i, int j,
                                          T &gx, T &gy, T &dv, T &cu) {
    const int n = g.n;
    // neighbour entry (depth 0) of edge e at position k
    auto nb = [&](int e, int k) { return g.nbr[(((long long)p * 4 + e) * n + k) * g.ndepth]; };
    auto halo = [&](int e, int k, T &hx_, T &hy_, T &hd, T &hc) {
        const int4 t = nb(e, k);
        const Raw4<ORD, T> r = raw_global<ORD, T>(g, y, hs, t.x, t.y, t.z);
        map_vec(t.w, r.gx, r.gy, hx_, hy_);
        hd = r.dv;
        hc = r.cu;
    };
    // covariant components enter the sync routine as (U,V) = (gy,gx)
    // (ecs_Ah_vec_sync_mod.f90:51-53)
    const T U = gy, V = gx;
    const bool L = i == 0, Rt = i == n - 1, B = j == 0, Tp = j == n - 1;
    T Ux, Vx, Dx, Cx, Uy, Vy, Dy, Cy;  // halo values across the x-edge and the y-edge
    if (L && B) {
        halo(0, 0, Vx, Ux, Dx, Cx);
        halo(2, 0, Vy, Uy, Dy, Cy);
        dv = (dv + Dx + Dy) / T(3);
        cu = (cu + Cx + Cy) / T(3);
        gy = (U + Ux + Uy + g.qb[0] * Vy) / T(3);
        gx = (V + Vx + Vy + g.ql[0] * Ux) / T(3);
    } else if (Rt && B) {
        halo(1, 0, Vx, Ux, Dx, Cx);
        halo(2, n - 1, Vy, Uy, Dy, Cy);
        dv = (dv + Dx + Dy) / T(3);
        cu = (cu + Cx + Cy) / T(3);
        gy = (U + Ux + Uy + g.qb[n - 1] * Vy) / T(3);
        gx = (V + Vx + Vy + g.qr[0] * Ux) / T(3);
    } else if (L && Tp) {
        halo(0, n - 1, Vx, Ux, Dx, Cx);
        halo(3, 0, Vy, Uy, Dy, Cy);
        dv = (dv + Dx + Dy) / T(3);
        cu = (cu + Cx + Cy) / T(3);
        gy = (U + Ux + Uy + g.qt[0] * Vy) / T(3);
        gx = (V + Vx + Vy + g.ql[n - 1] * Ux) / T(3);
    } else if (Rt && Tp) {
        halo(1, n - 1, Vx, Ux, Dx, Cx);
        halo(3, n - 1, Vy, Uy, Dy, Cy);
        dv = (dv + Dx + Dy) / T(3);
        cu = (cu + Cx + Cy) / T(3);
        gy = (U + Ux + Uy + g.qt[n - 1] * Vy) / T(3);
        gx = (V + Vx + Vy + g.qr[n - 1] * Ux) / T(3);
    } else if (L) {
        halo(0, j, Vx, Ux, Dx, Cx);
        dv = T(0.5) * (Dx + dv);
        cu = T(0.5) * (Cx + cu);
        gy = T(0.5) * (Ux + U);
        gx = T(0.5) * (V + Vx + g.ql[j] * Ux);
    } else if (Rt) {
        halo(1, j, Vx, Ux, Dx, Cx);
        dv = T(0.5) * (dv + Dx);
        cu = T(0.5) * (cu + Cx);
        gy = T(0.5) * (U + Ux);
        gx = T(0.5) * (V + Vx + g.qr[j] * Ux);
    } else if (B) {
        halo(2, i, Vy, Uy, Dy, Cy);
        dv = T(0.5) * (dv + Dy);
        cu = T(0.5) * (cu + Cy);
        gx = T(0.5) * (Vy + V);
        gy = T(0.5) * (U + Uy + g.qb[i] * Vy);
    } else if (Tp) {
        halo(3, i, Vy, Uy, Dy, Cy);
        dv = T(0.5) * (dv + Dy);
        cu = T(0.5) * (cu + Cy);
        gx = T(0.5) * (V + Vy);
        gy = T(0.5) * (U + Uy + g.qt[i] * Vy);
    }
}

template <int ORD, typename T, int MODE, int TX, int TY, int RPT>
__global__ void __launch_bounds__(TX * (TY / RPT))
stage_kernel(DevGeo<T> g, StageArgs<T> a) {
    using C = TileCfg<ORD, T, TX, TY, RPT>;
    constexpr int R = C::R, LX = C::LX, LY = C::LY, NT = C::THREADS;
    __shared__ T sK[LY][LX], sGxd[LY][LX], sGyd[LY][LX], sGxc[LY][LX], sGyc[LY][LX];

    const int n = g.n;
    const int ntx = (n + TX - 1) / TX, nty = (n + TY - 1) / TY;
    const int nblocks = 6 * ntx * nty;
    const int tile = xcd_tile(blockIdx.x, nblocks);
    const int p = tile / (ntx * nty);
    const int rem = tile - p * ntx * nty;
    const int ty0 = (rem / ntx) * TY, tx0 = (rem % ntx) * TX;
    const int tid = threadIdx.x;
    const T *y = a.y;

    // ---- load phase: stencil inputs on the cross-shaped region of the tile ----
    for (int c = tid; c < LX * LY; c += NT) {
        const int ly = c / LX, lx = c - ly * LX;
        const bool cross = (ly >= R && ly < R + TY) || (lx >= R && lx < R + TX);
        const int i = tx0 - R + lx, j = ty0 - R + ly;
        if (!cross || i < 0 || i >= n || j < 0 || j >= n) continue;
        const Derived<T> d = derived_global(g, y, a.hs, p, i, j);
        sK[ly][lx] = d.K;
        sGxd[ly][lx] = d.Gxd;
        sGyd[ly][lx] = d.Gyd;
        sGxc[ly][lx] = d.Gxc;
        sGyc[ly][lx] = d.Gyc;
    }
    __syncthreads();

    // ---- output phase ----
    const int tx = tid % TX, tyb = tid / TX;
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
        const int ty = tyb + rr * (TY / RPT);
        const int i = tx0 + tx, j = ty0 + ty;
        if (i >= n || j >= n) continue;
        const int lx = tx + R, ly = ty + R;
        auto xK = [&](int ai) { return sK[ly][ai - tx0 + R]; };
        auto yK = [&](int aj) { return sK[aj - ty0 + R][lx]; };
        auto xD = [&](int ai) { return sGxd[ly][ai - tx0 + R]; };
        auto yD = [&](int aj) { return sGyd[aj - ty0 + R][lx]; };
        auto xC = [&](int ai) { return sGxc[ly][ai - tx0 + R]; };
        auto yC = [&](int aj) { return sGyc[aj - ty0 + R][lx]; };
        const T dxK = sbp_row<ORD, T>(i, n, xK);
        const T dyK = sbp_row<ORD, T>(j, n, yK);
        const T dxD = sbp_row<ORD, T>(i, n, xD);
        const T dyD = sbp_row<ORD, T>(j, n, yD);
        const T dxC = sbp_row<ORD, T>(i, n, xC);
        const T dyC = sbp_row<ORD, T>(j, n, yC);
        const long long id = pidx(g, p, i, j);
        const T Jv = g.J[id];
        const T Jd = Jv * g.hx * g.scale;
        T gx = dxK / g.hxs;
        T gy = dyK / g.hxs;
        T dv = (dxD + dyD) / Jd;
        T cu = (dxC + dyC) / Jd;
        if (i == 0 || j == 0 || i == n - 1 || j == n - 1) edge_sync<ORD, T>(g, y, a.hs, p, i, j, gx, gy, dv, cu);

        if constexpr (MODE == MODE_CURL) {
            a.out[id] = cu;
            continue;
        }
        // Coriolis and tendencies (coriolis_colocated_mod.f90:97-98, operator_swm_mod.f90:95-101)
        const T h = y[id], u = y[id + g.fs], v = y[id + 2 * g.fs];
        const T q11 = g.Qi11[id], q12 = g.Qi12[id], q22 = g.Qi22[id];
        const T ut = q11 * u + q12 * v;
        const T vt = q12 * u + q22 * v;
        const T hu = h * ut, hv = h * vt;
        const T fz = g.f[id] + cu;
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
            a.out[id] = a.v0[id] + a.cy * kh;
            a.out[id + g.fs] = a.v0[id + g.fs] + a.cy * ku;
            a.out[id + 2 * g.fs] = a.v0[id + 2 * g.fs] + a.cy * kv;
            a.acc[id] = a.acc[id] + a.cacc * kh;
            a.acc[id + g.fs] = a.acc[id + g.fs] + a.cacc * ku;
            a.acc[id + 2 * g.fs] = a.acc[id + 2 * g.fs] + a.cacc * kv;
        } else if constexpr (MODE == MODE_STAGE4) {
            a.out[id] = a.acc[id] + a.cy * kh;
            a.out[id + g.fs] = a.acc[id + g.fs] + a.cy * ku;
            a.out[id + 2 * g.fs] = a.acc[id + 2 * g.fs] + a.cy * kv;
        }
    }
}

// out = s1*v1 + s2*v2 (assign) or out = out + s1*v1 + s2*v2 (update) over the mesh points of
// all three fields (grid_field_mod.f90:430-485 via stvec_swm_mod.f90:38-106)
template <typename T, bool UPDATE>
__global__ void axpby_kernel(T *out, T s1, const T *v1, T s2, const T *v2, int n, int pitch, long long fs) {
    const long long total = 3LL * 6 * n * (long long)n;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < total; k += (long long)gridDim.x * blockDim.x) {
        const long long row = k / n;
        const int i = (int)(k - row * n);
        const long long id = row * pitch + i;
        if (UPDATE)
            out[id] = out[id] + s1 * v1[id] + s2 * v2[id];
        else
            out[id] = s1 * v1[id] + s2 * v2[id];
    }
}

}  // namespace swm
