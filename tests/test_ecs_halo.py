"""ECS cubic halo on xy (Ah) points, oracle side (CPU): known answers and the weights.

The reference's test (src/test/test_halo/test_generic_halo_mod.f90, called with nh=32,
halo_width=3, "ECS_xy") extends the analytic fields x, y, z (test_fields_mod.f90:463-480)
into the halo and prints the errors without a pass/fail bound for this procedure; the
checks here are what the mathematics fixes:
  * cubic Lagrange weights reproduce 1, s, s^2, s^3 (moments, 1e-13);
  * 4th-order convergence of the cross-edge error (N=32 -> 64: ratio > 12), max error
    < 3e-7 at N=32 for the unit-sphere fields;
  * the cube-corner halo diagonals and the averaged edge lines are exact for continuous
    fields (< 1e-14).
"""
import ctypes

import numpy as np
import pytest


def _band_err(G, f, true, hw):
    n, H = G.n, G.hw
    d = np.abs(f - true)
    band = max(max(d[p, H - hw:H, H:H + n].max(), d[p, H + n:H + n + hw, H:H + n].max(),
                   d[p, H:H + n, H - hw:H].max(), d[p, H:H + n, H + n:H + n + hw].max()) for p in range(6))
    diag = 0.0
    for p in range(6):
        for j in range(1, hw + 1):
            for ii, jj in ((1 - j, 1 - j), (n + j, 1 - j), (n + j, n + j), (1 - j, n + j)):
                diag = max(diag, d[p, jj - 1 + H, ii - 1 + H])
    inner = d[:, H:H + n, H:H + n].max()
    return band, diag, inner


def _run(oracle_lib, N, hw, level):
    G = oracle_lib.Grid(N)
    true = G.geometry("r")[level].copy()
    f = np.full_like(true, -9999999.0)
    G.interior(f)[...] = G.interior(true)
    G.ecs_halo_xy(f, hw)
    return G, f, true


@pytest.mark.parametrize("level", [0, 1, 2])
def test_known_answer_and_convergence(oracle_lib, level):
    G, f, t = _run(oracle_lib, 32, 3, level)
    b32, diag, inner = _band_err(G, f, t, 3)
    G2, f2, t2 = _run(oracle_lib, 64, 3, level)
    b64, _, _ = _band_err(G2, f2, t2, 3)
    print(f"  level {level}: band max N=32 {b32:.3e}, N=64 {b64:.3e}, ratio {b32 / b64:.1f}; diag {diag:.1e}")
    assert b32 < 3e-7
    assert b32 / b64 > 12.0
    assert diag < 1e-14 and inner < 1e-14


def test_weights_host_equals_oracle_and_reproduce_cubics(oracle_lib):
    from cubedsphere_sbp_swm_amd import _capi
    for N, hw in ((32, 3), (17, 2), (48, 8)):
        G = oracle_lib.Grid(N)
        w_ref, ind_ref = G.ecs_xy_weights(hw)
        L = G.n + 2 * hw
        w = np.zeros((hw, L, 4))
        ind = np.zeros((hw, L), dtype=np.int32)
        assert _capi.lib().swmh_ecs_xy_weights(N, hw, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                               ind.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == 0
        assert np.array_equal(w, w_ref) and np.array_equal(ind, ind_ref)
        # moments: sum_s w_s s^m = t^m for the nodes s = -1, 0, 1, 2 and t = position of the
        # projected point relative to node 0 (recomputed here from the geometry)
        hx = 0.5 * np.pi / N
        s = np.array([-1.0, 0.0, 1.0, 2.0])
        for j in range(1, hw + 1):
            zb = 0.25 * np.pi + j * hx
            i = np.arange(1 - hw, G.n + hw + 1)
            za = np.arctan(np.tan(-0.25 * np.pi + (i - 1.0) * hx) / np.tan(zb))
            t = (za + 0.25 * np.pi) / hx - (ind[j - 1] - 1.0)
            for m in range(4):
                assert np.max(np.abs(w[j - 1] @ s ** m - t ** m)) < 1e-12
        assert ind.min() >= 2 and ind.max() <= G.n - 2
