"""Hyperdiffusion on the device (swm_laplace, swm_diff_apply, swm_diff_step) against the oracle's
restatement of laplace_Ah_sbp{21,42,63}_narrow_t, hordiff_scalar_t /
hordiff_colocated_xyz_t and explicit_Eul1: FP64, the same coefficient tables and operation
order, so every point is expected bit for bit; the tests allow 1e-13 relative (max abs
difference / max abs value).  Panel sizes from the smallest the closures allow (N+1 = 12 for
sbp42, 18 for sbp63; 23-25 put one tile over the closures of both edges for sbp63; sbp21 from
the smallest context, N = 8) to 49;
seeded spherical-harmonic states.  The diffusion step is checked inside an RK4 loop as the
reference drivers call it (ts2_mod.f90:145-146), also with different families for the velocity
and the height."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EARTH_A, EARTH_G = 6.371229e6, 9.80616
EARTH_OM = 2.0 * np.pi / (23 * 3600 + 56 * 60 + 4)
TOL = 1e-13


def _setup(oracle_lib, N, seed=1, lmax=8):
    import cubedsphere_sbp_swm_amd as swm
    op = swm.create_swm_operator(N, **swm.operator_names(42), scale=EARTH_A, omega=EARTH_OM, grav=EARTH_G)
    G = oracle_lib.Grid(N, grad_order=42, div_kind=42, scale=EARTH_A, omega=EARTH_OM, grav=EARTH_G)
    h, u, v = op.initial_state(seed=seed, lmax=lmax, halo_width=G.hw)
    return swm, op, G, h, u, v


def _rel(G, a, b):
    ai, bi = G.interior(a), G.interior(b)
    return float(np.max(np.abs(ai - bi)) / max(np.max(np.abs(bi)), 1e-300))


LAPLACE_CASES = [(21, 8), (21, 16), (21, 33), (42, 11), (42, 16), (42, 33), (42, 48), (63, 17), (63, 22),
                 (63, 23), (63, 24), (63, 33), (63, 48)]


@pytest.mark.parametrize("order,N", LAPLACE_CASES)
def test_laplace_matches_oracle(oracle_lib, order, N):
    swm, op, G, h, u, v = _setup(oracle_lib, N, seed=N)
    print(f"[laplace] order={order} N={N}", flush=True)
    x, out = swm.create_stvec_swm(op), swm.create_stvec_swm(op)
    x.upload(h, u, v, G.hw)
    op.laplace(out, x, f"laplace_ah_sbp{order}_narrow")
    got = out.download(G.hw)[0]
    ref = G.laplace(h, order)
    err = _rel(G, got, ref)
    print(f"  rel err {err:.2e}, bitwise {np.array_equal(G.interior(got), G.interior(ref))}")
    assert err < TOL


def _names(uv_order, h_order):
    return f"hordiff_vec_xyz_Ah_sbp_{uv_order}_narrow", f"hordiff_scalar_Ah_sbp_{h_order}_narrow"


@pytest.mark.parametrize("uv_order,h_order,N", [(21, 21, 16), (21, 63, 24), (42, 42, 16), (42, 42, 33),
                                                (63, 63, 17), (63, 63, 33), (63, 42, 24)])
def test_diff_apply_matches_oracle(oracle_lib, uv_order, h_order, N):
    swm, op, G, h, u, v = _setup(oracle_lib, N, seed=2 * N)
    print(f"[diff_apply] uv={uv_order} h={h_order} N={N}", flush=True)
    x, t = swm.create_stvec_swm(op), swm.create_stvec_swm(op)
    x.upload(h, u, v, G.hw)
    op.diff_apply(t, x, 0.03, 0.01, *_names(uv_order, h_order))
    gh, gu, gv = t.download(G.hw)
    rh = G.hordiff_scalar(0.01, h, h_order)
    ru, rv = G.hordiff_vec(0.03, u, v, uv_order)
    errs = [_rel(G, a, b) for a, b in ((gh, rh), (gu, ru), (gv, rv))]
    print(f"  rel err h,u,v = {errs}")
    assert max(errs) < TOL


@pytest.mark.parametrize("uv_order,h_order", [(21, 21), (42, 42), (63, 63), (42, 63)])
def test_rk4_with_diff_steps_matches_oracle(oracle_lib, uv_order, h_order):
    """Three steps of RK4 (sbp42 operators) followed by the explicit hyperdiffusion step, C24."""
    N, dt, steps = 24, 600.0, 3
    swm, op, G, h, u, v = _setup(oracle_lib, N, seed=5)
    print(f"[rk4+diff] uv={uv_order} h={h_order} N={N} steps={steps}", flush=True)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v, G.hw)
    ts = swm.create_timescheme(op, "rk4")
    for _ in range(steps):
        ts.steps(x, dt, 1)
        op.diff_step(x, dt, 0.03, 0.01, *_names(uv_order, h_order))
        G.rk4(h, u, v, dt, 1)
        G.diff_step(h, u, v, dt, 0.03, 0.01, uv_order, h_order)
    errs = [_rel(G, a, b) for a, b in zip(x.download(G.hw), (h, u, v))]
    print(f"  rel err h,u,v = {errs}")
    assert max(errs) < 1e-12


def test_diff_rejects_unsupported():
    import cubedsphere_sbp_swm_amd as swm
    op = swm.create_swm_operator(16, **swm.operator_names(42))
    x, t = swm.create_stvec_swm(op), swm.create_stvec_swm(op)
    with pytest.raises(swm.SwmError):
        op.diff_apply(t, x, 0.03, 0.01, hordiff_uv_name="hordiff_vec_xyz_Ah_sbp_43_narrow")
    with pytest.raises(swm.SwmError):
        op.laplace(t, x, "laplace_ah_sbp43_narrow")
    with pytest.raises(swm.SwmError):   # N+1 = 17 < 18: too small for the sbp63 closures
        op.diff_step(x, 1.0, 0.03, 0.01, hordiff_h_name="hordiff_scalar_Ah_sbp_63_narrow")
    with pytest.raises(swm.SwmError):
        op.diff_step(x, 1.0, 0.03, 0.01, hordiff_h_name="hordiff_scalar_C")
    with pytest.raises(swm.SwmError):
        op.diff_apply(x, x, 0.03, 0.01)
    small = swm.create_swm_operator(10, **swm.operator_names(42))
    y = swm.create_stvec_swm(small)
    with pytest.raises(swm.SwmError):
        small.diff_step(y, 1.0, 0.03, 0.01)
    f32 = swm.create_swm_operator(16, **swm.operator_names(42), precision=32)
    z = swm.create_stvec_swm(f32)
    with pytest.raises(swm.SwmError):
        f32.diff_step(z, 1.0, 0.03, 0.01)
