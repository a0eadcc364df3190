"""ECS cubic halo on O points (cell centres, the reference's ECS_O halo procedure) on the CPU: the
oracle's restatement pinned by what the mathematics fixes.  The halo value at a virtual point is
an approximation of the field at that point of the panel's extended gnomonic coordinates, so for
a smooth field the error shrinks by more than 10x per doubling of N (4th order: about 14-16x
measured) on the edge strips (cubic interpolation along the neighbour's rows) and at the corner
special points (a 3 x 3 linear solve and averages of such interpolations).  The weights are 4-point Lagrange weights (sum 1, exact
on cubics).  Each panel's frame is recovered from the Ah geometry (r = M_p r_proto), so no
topology table is restated here.
"""
import numpy as np
import pytest


def _frames(G):
    """3 x 3 map per panel from prototype-face coordinates to the sphere, fitted on Ah points."""
    N, n = G.N, G.n
    a = -0.25 * np.pi + np.arange(n) * (0.5 * np.pi / N)
    A, B = np.meshgrid(a, a)          # [j, i]
    rp = np.stack([np.tan(A), np.tan(B), np.ones_like(A)], -1)
    rp /= np.linalg.norm(rp, axis=-1, keepdims=True)
    r = np.stack([G.interior(c) for c in G.geometry("r")], -1)   # (6, n, n, 3)
    Ms = []
    for p in range(6):
        M, *_ = np.linalg.lstsq(rp.reshape(-1, 3), r[p].reshape(-1, 3), rcond=None)
        assert np.abs(rp.reshape(-1, 3) @ M - r[p].reshape(-1, 3)).max() < 1e-13
        Ms.append(M)
    return Ms


def _o_points(G, hw):
    """sphere positions (6, L, L, 3) of the O points and virtual points of the extended panels"""
    N = G.N
    idx = np.arange(1 - hw, N + hw + 1)
    a = -0.25 * np.pi + (idx - 0.5) * (0.5 * np.pi / N)
    A, B = np.meshgrid(a, a)
    rp = np.stack([np.tan(A), np.tan(B), np.ones_like(A)], -1)
    rp /= np.linalg.norm(rp, axis=-1, keepdims=True)
    return np.stack([rp @ M for M in _frames(G)])


def _field(r):
    x, y, z = r[..., 0], r[..., 1], r[..., 2]
    return np.sin(2.0 * x) * y + z * z * x + np.cos(y + z)


def _corner_mask(N, hw):
    L = N + 2 * hw
    k = np.arange(L)
    outside = (k < hw) | (k >= N + hw)
    return np.outer(outside, outside)


@pytest.mark.parametrize("hw", [1, 2, 3])
def test_halo_converges_to_the_field(oracle_lib, hw):
    errs = []
    for N in (16, 32, 64):
        G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
        r = _o_points(G, hw)
        exact = _field(r)
        f = exact[:, hw:hw + N, hw:hw + N].copy()
        out = G.ecs_halo_o(f, hw)
        assert np.array_equal(out[:, hw:hw + N, hw:hw + N], f)
        cm = _corner_mask(N, hw)
        defined = ~np.isnan(out)
        # everything but the corner squares is written; in the corner squares the special
        # points: (0,0) for hw = 1, the 2 x 2 block next to the corner for hw >= 2
        assert defined[:, ~cm].all()
        assert defined[:, cm].sum() == 6 * 4 * (1 if hw == 1 else 4)
        e = np.abs(out - exact)
        strip = e[:, ~cm & ~np.pad(np.ones((N, N), bool), hw)]
        corner = e[:, cm & defined[0]]
        errs.append((strip.max(), corner.max()))
    print(f"hw={hw}: strips / corner points max errors", errs)
    (s0, c0), (s1, c1), (s2, c2) = errs
    assert s0 / s1 > 10.0 and s1 / s2 > 12.0
    assert c0 / c1 > 10.0 and c1 / c2 > 10.0
    assert s2 < 1e-6


def test_weights_are_cubic_lagrange(oracle_lib):
    G = oracle_lib.Grid(24, 42, 42, 1.0, 0.0, 1.0)
    w, ind = G.ecs_o_weights(3)
    assert np.abs(w.sum(-1) - 1.0).max() < 1e-13
    k = np.arange(-1.0, 3.0)
    z = (w * k).sum(-1)
    for m in (1, 2, 3):
        assert np.abs((w * (k - z[..., None]) ** m).sum(-1)).max() < 1e-12
    # targets lie inside the middle interval of their stencil, sources inside the panel for the
    # positions the procedure interpolates
    assert z.min() >= -1e-12 and z.max() <= 1.0 + 1e-12
    N = 24
    for r in (1, 2, 3):
        t = np.arange(2, N) if r == 1 else np.arange(1, N + 1)
        base = ind[r - 1, t + 1]
        assert base.min() - 1 >= 1 and base.max() + 2 <= N


def test_rejects_bad_widths(oracle_lib):
    G = oracle_lib.Grid(16, 42, 42, 1.0, 0.0, 1.0)
    with pytest.raises(ValueError):
        G.ecs_halo_o(np.zeros((6, 16, 16)), 0)
    with pytest.raises(ValueError):
        G.ecs_halo_o(np.zeros((6, 16, 16)), 9)


@pytest.mark.parametrize("N,rows", [(8, 2), (33, 4)])
def test_host_weights_equal_oracle(oracle_lib, N, rows):
    import ctypes
    from cubedsphere_sbp_swm_amd import _capi
    G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
    w_ref, ind_ref = G.ecs_o_weights(rows)
    w = np.zeros_like(w_ref)
    ind = np.zeros_like(ind_ref)
    _capi.lib().swmh_ecs_o_weights(N, rows, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   ind.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert np.array_equal(w, w_ref) and np.array_equal(ind, ind_ref)


# ---- vector halo on O points (A grid, contravariant components) ----

def _contravariant(G, hw, V):
    """contravariant components (u, v) of the tangent field V(r) at the O points and virtual
    points of the extended panels: V = u a1 + v a2 with a1 = dr/dalpha, a2 = dr/dbeta of the
    panel's gnomonic map (derivatives taken analytically, components from the 2 x 2 Gram system)"""
    N = G.N
    idx = np.arange(1 - hw, N + hw + 1)
    a = -0.25 * np.pi + (idx - 0.5) * (0.5 * np.pi / N)
    A, B = np.meshgrid(a, a)
    X, Y = np.tan(A), np.tan(B)
    d = np.sqrt(1 + X * X + Y * Y)
    P = np.stack([X, Y, np.ones_like(X)], -1) / d[..., None]
    e1, e2 = np.zeros(P.shape), np.zeros(P.shape)
    e1[..., 0], e2[..., 1] = 1.0, 1.0
    a1p = (1 + X * X)[..., None] * (e1 / d[..., None] - X[..., None] * P / (d * d)[..., None])
    a2p = (1 + Y * Y)[..., None] * (e2 / d[..., None] - Y[..., None] * P / (d * d)[..., None])
    us, vs = [], []
    for M in _frames(G):
        r, a1, a2 = P @ M, a1p @ M, a2p @ M
        Vr = V(r)
        g11, g12, g22 = (a1 * a1).sum(-1), (a1 * a2).sum(-1), (a2 * a2).sum(-1)
        r1, r2 = (Vr * a1).sum(-1), (Vr * a2).sum(-1)
        det = g11 * g22 - g12 * g12
        us.append((g22 * r1 - g12 * r2) / det)
        vs.append((g11 * r2 - g12 * r1) / det)
    return np.stack(us), np.stack(vs)


def _flow(r):
    x, y, z = r[..., 0], r[..., 1], r[..., 2]
    rot = np.stack([-y, x, np.zeros_like(x)], -1) + 0.5 * np.stack([np.zeros_like(x), -z, y], -1)
    # plus the tangential part of a gradient field: grad(x y + z^2) projected on the sphere
    gr = np.stack([y, x, 2 * z], -1)
    gr = gr - (gr * r).sum(-1)[..., None] * r
    return rot + 0.3 * gr


@pytest.mark.parametrize("hw", [1, 2, 3])
def test_vector_halo_converges_to_the_field(oracle_lib, hw):
    errs = []
    for N in (16, 32, 64):
        G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
        eu, ev = _contravariant(G, hw, _flow)
        u, v = eu[:, hw:hw + N, hw:hw + N].copy(), ev[:, hw:hw + N, hw:hw + N].copy()
        ou, ov = G.ecs_halo_vec_a(u, v, hw)
        assert np.array_equal(ou[:, hw:hw + N, hw:hw + N], u) and np.array_equal(ov[:, hw:hw + N, hw:hw + N], v)
        assert np.array_equal(np.isnan(ou), np.isnan(ov))
        cm = _corner_mask(N, hw)
        defined = ~np.isnan(ou)
        assert defined[:, ~cm].all()
        e = np.maximum(np.abs(ou - eu), np.abs(ov - ev))
        strip = e[:, ~cm & ~np.pad(np.ones((N, N), bool), hw)]
        corner = e[:, cm & defined[0]]
        errs.append((strip.max(), corner.max()))
    print(f"vector hw={hw}: strips / corner points max errors", errs)
    (s0, c0), (s1, c1), (s2, c2) = errs
    assert s0 / s1 > 10.0 and s1 / s2 > 10.0
    assert c0 / c1 > 8.0 and c1 / c2 > 8.0
    assert s2 < 1e-6


@pytest.mark.parametrize("N,rows", [(8, 5), (33, 6)])
def test_host_transform_matrices_equal_oracle(oracle_lib, N, rows):
    import ctypes
    from cubedsphere_sbp_swm_amd import _capi
    G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
    ref = G.ecs_vec_a_tm(rows)
    tm = np.zeros_like(ref)
    _capi.lib().swmh_ecs_vec_a_tm(N, rows, tm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert np.array_equal(tm, ref)
