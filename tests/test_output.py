"""Lat-lon regrid of Ah fields (the reference's latlon outputers, cubic interpolation) on the CPU:
the product's host target table against the oracle's restatement, and the oracle pinned by what
the mathematics fixes: weights of a 4-point Lagrange stencil (sum 1, exact on cubics in the
panel's native coordinate), 4th-order convergence on a smooth field, and the rotation of
covariant components to lat-lon components on two solid-body rotations, whose zonal and
meridional components are known in closed form (about the z axis: (cos lat, 0); about the x
axis: (-sin lat cos lon, sin lon)).
"""
import ctypes

import numpy as np
import pytest


def _table_host(N, Nlat, Nlon):
    from cubedsphere_sbp_swm_amd import _capi
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    panel, x0, y0 = (np.zeros((Nlat, Nlon), np.int32) for _ in range(3))
    wx, wy, rot = np.zeros((Nlat, Nlon, 4)), np.zeros((Nlat, Nlon, 4)), np.zeros((Nlat, Nlon, 4))
    _capi.lib().swmh_latlon_table(N, Nlat, Nlon, ip(panel), ip(x0), ip(y0), dp(wx), dp(wy), dp(rot))
    return panel, x0, y0, wx, wy, rot


@pytest.mark.parametrize("N", [8, 24])
def test_host_table_equals_oracle(oracle_lib, N):
    Nlat, Nlon = 2 * N + 1, 4 * N
    G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
    _, _, panel, x0, y0, wx, wy = G.latlon_regrid(Nlat, Nlon, G.zeros(), G.zeros(), table=True)
    hp, hx0, hy0, hwx, hwy, _ = _table_host(N, Nlat, Nlon)
    for a, b in ((hp, panel), (hx0, x0), (hy0, y0), (hwx, wx), (hwy, wy)):
        assert np.array_equal(a, b)
    n = N + 1
    assert x0.min() >= 0 and x0.max() <= n - 4 and y0.min() >= 0 and y0.max() <= n - 4
    assert set(np.unique(panel)) == set(range(6))


def test_weights_are_cubic_lagrange(oracle_lib):
    N = 16
    G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
    _, _, _, x0, y0, wx, wy = G.latlon_regrid(33, 64, G.zeros(), G.zeros(), table=True)
    assert np.abs(wx.sum(-1) - 1.0).max() < 1e-13 and np.abs(wy.sum(-1) - 1.0).max() < 1e-13
    # exact on cubics of the stencil offset: sum_k w_k (k - z)^m = 0 for m = 1..3, z = sum w_k k
    k = np.arange(4.0)
    for w in (wx, wy):
        z = (w * k).sum(-1)
        for m in (1, 2, 3):
            assert np.abs((w * (k - z[..., None]) ** m).sum(-1)).max() < 1e-12
        # the target lies inside the stencil, shifted one-sided only at panel edges
        assert z.min() >= -1e-12 and z.max() <= 3.0 + 1e-12


def test_scalar_regrid_converges(oracle_lib):
    errs = []
    for N in (16, 32, 64):
        G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
        x, y, z = G.geometry("r")
        f = np.sin(2.0 * x) * y + z * z * x + np.cos(y + z)
        Nlat, Nlon = 2 * N + 1, 4 * N
        (out,) = G.latlon_regrid(Nlat, Nlon, f)
        lat = -0.5 * np.pi + np.pi / (Nlat - 1) * np.arange(Nlat)
        lon = 2.0 * np.pi / Nlon * np.arange(Nlon)
        X = np.cos(lat)[:, None] * np.cos(lon)[None]
        Y = np.cos(lat)[:, None] * np.sin(lon)[None]
        Z = np.sin(lat)[:, None] * np.ones(Nlon)[None]
        errs.append(np.abs(out - (np.sin(2.0 * X) * Y + Z * Z * X + np.cos(Y + Z))).max())
        c = G.latlon_regrid(Nlat, Nlon, np.full_like(x, 3.5))[0]
        assert np.abs(c - 3.5).max() < 1e-13
    print("scalar regrid max errors", errs)
    assert errs[0] / errs[1] > 10.0 and errs[1] / errs[2] > 12.0


@pytest.mark.parametrize("axis", ["z", "x"])
def test_vector_regrid_of_solid_rotation(oracle_lib, axis):
    errs = []
    for N in (16, 32, 64):
        G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
        x, y, z = G.geometry("r")
        a1, a2 = G.geometry("a1"), G.geometry("a2")
        V = (-y, x, np.zeros_like(x)) if axis == "z" else (np.zeros_like(x), -z, y)
        u = V[0] * a1[0] + V[1] * a1[1] + V[2] * a1[2]
        v = V[0] * a2[0] + V[1] * a2[1] + V[2] * a2[2]
        Nlat, Nlon = 2 * N + 1, 4 * N
        zon, mer = G.latlon_regrid(Nlat, Nlon, u, v)
        lat = (-0.5 * np.pi + np.pi / (Nlat - 1) * np.arange(Nlat))[:, None]
        lon = (2.0 * np.pi / Nlon * np.arange(Nlon))[None]
        if axis == "z":
            ez, em = np.cos(lat) * np.ones_like(lon), np.zeros((Nlat, Nlon))
        else:
            ez, em = -np.sin(lat) * np.cos(lon), np.sin(lon) * np.ones_like(lat)
        # the poles have no defined zonal direction: compare away from them
        inner = slice(1, Nlat - 1)
        errs.append(max(np.abs(zon - ez)[inner].max(), np.abs(mer - em)[inner].max()))
    print(f"rotation about {axis}: max errors", errs)
    assert errs[2] < 1e-5
    assert errs[0] / errs[1] > 8.0 and errs[1] / errs[2] > 9.0
