"""Output writers on the device (swm_latlon_write, swm_latlon_vec_write, swm_paneled_write):
the files hold REAL*4 records equal, value for value, to the oracle's FP64 cubic lat-lon regrid
(scalar; covariant vector rotated to zonal / meridional) rounded to REAL*4, and to the REAL*4
rounding of the mesh points for the paneled writer; record r starts at byte (r-1)*4*W, other
records of an existing file are left as they were.  Lat-lon grid (2N+1) x 4N as the reference's
drivers use it (ts2_mod.f90:108-112)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EARTH_A, EARTH_G = 6.371229e6, 9.80616
EARTH_OM = 2.0 * np.pi / (23 * 3600 + 56 * 60 + 4)


def _setup(oracle_lib, N, order=42, seed=3, precision=64):
    import cubedsphere_sbp_swm_amd as swm
    op = swm.create_swm_operator(N, **swm.operator_names(order), scale=EARTH_A, omega=EARTH_OM, grav=EARTH_G,
                                 precision=precision)
    G = oracle_lib.Grid(N, grad_order=42, div_kind=42, scale=EARTH_A, omega=EARTH_OM, grav=EARTH_G)
    h, u, v = op.initial_state(seed=seed, lmax=8, halo_width=G.hw)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v, G.hw)
    if precision == 32:   # the oracle sees the values the FP32 state holds
        h, u, v = (a.astype(np.float32).astype(np.float64) for a in (h, u, v))
    return swm, op, G, x, h, u, v


def _records(path, W):
    raw = np.fromfile(path, dtype="<f4")
    assert raw.size % W == 0
    return raw.reshape(-1, W)


@pytest.mark.parametrize("N,precision", [(16, 64), (48, 64), (24, 32)])
def test_latlon_writers_match_oracle(oracle_lib, tmp_path, N, precision):
    swm, op, G, x, h, u, v = _setup(oracle_lib, N, precision=precision)
    Nlat, Nlon = 2 * N + 1, 4 * N
    W = Nlat * Nlon
    print(f"[latlon] N={N} precision={precision}", flush=True)
    out = swm.create_latlon_outputer(op, Nlat, Nlon, "Ah")
    outv = swm.create_latlon_vec_outputer(op, Nlat, Nlon, "Ah", "covariant")
    fh, fu, fv = (str(tmp_path / f"{c}.dat") for c in "huv")
    out.write(x.field("h"), fh, 1)
    outv.write_vec(x, fu, fv, 1)
    # record 3 after a gap, from a changed state
    ts = swm.create_timescheme(op, "rk4")
    ts.steps(x, 300.0, 2)
    out.write(x.field("h"), fh, 3)
    outv.write_vec(x, fu, fv, 3)
    h2, u2, v2 = x.download(G.hw)
    (rh,) = G.latlon_regrid(Nlat, Nlon, h)
    ru, rv = G.latlon_regrid(Nlat, Nlon, u, v)
    (rh2,) = G.latlon_regrid(Nlat, Nlon, h2)
    ru2, rv2 = G.latlon_regrid(Nlat, Nlon, u2, v2)
    for path, r1, r3 in ((fh, rh, rh2), (fu, ru, ru2), (fv, rv, rv2)):
        recs = _records(path, W)
        assert recs.shape[0] == 3
        assert np.array_equal(recs[0], r1.astype(np.float32).ravel())
        assert np.array_equal(recs[2], r3.astype(np.float32).ravel())
        assert not recs[1].any()
    # rewriting record 1 keeps record 3
    out.write(x.field("h"), fh, 1)
    recs = _records(fh, W)
    assert np.array_equal(recs[0], rh2.astype(np.float32).ravel())
    assert np.array_equal(recs[2], rh2.astype(np.float32).ravel())


def test_paneled_writer(oracle_lib, tmp_path):
    N = 20
    swm, op, G, x, h, u, v = _setup(oracle_lib, N)
    n = N + 1
    print(f"[paneled] N={N}", flush=True)
    out = swm.create_master_paneled_outputer(op, "xy")
    path = str(tmp_path / "u_paneled.dat")
    out.write(x.field("u"), path, 2)
    recs = _records(path, 6 * n * n)
    assert recs.shape[0] == 2 and not recs[0].any()
    assert np.array_equal(recs[1].reshape(6, n, n), G.interior(u).astype(np.float32))


def test_writers_reject_unsupported(tmp_path):
    import cubedsphere_sbp_swm_amd as swm
    op = swm.create_swm_operator(16, **swm.operator_names(42))
    x = swm.create_stvec_swm(op)
    with pytest.raises(swm.SwmError):
        swm.create_latlon_outputer(op, 33, 64, "A")
    with pytest.raises(swm.SwmError):
        swm.create_latlon_vec_outputer(op, 33, 64, "Ah", "contravariant")
    with pytest.raises(swm.SwmError):
        swm.create_master_paneled_outputer(op, "o")
    out = swm.create_latlon_outputer(op, 33, 64, "Ah")
    with pytest.raises(swm.SwmError):
        out.write(x.field("h"), str(tmp_path / "h.dat"), 0)
    with pytest.raises(swm.SwmError):
        out.write(x.field("h"), str(tmp_path / "missing_dir" / "h.dat"), 1)
    assert not os.path.exists(tmp_path / "h.dat")
