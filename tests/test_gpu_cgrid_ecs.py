"""C-grid ECS halos on the device against the oracle, FP64, bit for bit.

* swm_c_ecs_halo_vec ("ecs_C_vec" and "ecs_C_vec_covariant") at N = 8 (the smallest panel the
  ECS halos allow), 16, 33 and 384, halo widths 1..5, random components: every written point
  equal to the oracle's restatement of get_halo_vector, NaN in the same corner squares.
* swm_c_grad with "gradient_c2_ecs" (also on the halo rows it computes) and swm_c_massflux with
  "massflux_c_up4" (random velocities with exact zeros of both signs, so both upwind branches
  and the sign of zero are exercised) at N = 8, 32, 65.
* One context serving several widths and both kinds in turn (the cached tables are replaced).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _op(N):
    import cubedsphere_sbp_swm_amd as swm
    return swm.create_swm_operator(N, **swm.operator_names("c42"))


def _fields(C, seed):
    N = C.N
    rng = np.random.default_rng(seed)
    u = C.embed("x", rng.standard_normal((6, N, N + 1)))
    v = C.embed("y", rng.standard_normal((6, N + 1, N)))
    f = C.embed("o", rng.standard_normal((6, N, N)))
    return f, u, v


@pytest.mark.parametrize("N,hw", [(8, 1), (8, 2), (16, 3), (16, 5), (33, 2), (33, 4)])
@pytest.mark.parametrize("procedure", ["ecs_C_vec", "ecs_C_vec_covariant"])
def test_ecs_halo_vec_matches_oracle(oracle_lib, N, hw, procedure):
    print(f"[c_ecs_halo_vec] N={N} hw={hw} {procedure}", flush=True)
    op = _op(N)
    C = oracle_lib.CGrid(N)
    _, u, v = _fields(C, 100 + N + hw)
    gu, gv = op.ecs_halo_vector(u, v, hw, procedure, halo_width_in=C.hw)
    ru, rv = C.ecs_halo_vec(u, v, hw, procedure.endswith("covariant"))
    for got, ref in ((gu, ru), (gv, rv)):
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        assert np.array_equal(got, ref, equal_nan=True)


def test_ecs_halo_vec_c384(oracle_lib):
    print("[c_ecs_halo_vec] N=384 hw=3", flush=True)
    N = 384
    op = _op(N)
    C = oracle_lib.CGrid(N)
    _, u, v = _fields(C, 7)
    for procedure in ("ecs_C_vec", "ecs_C_vec_covariant"):
        gu, gv = op.ecs_halo_vector(u, v, 3, procedure, halo_width_in=C.hw)
        ru, rv = C.ecs_halo_vec(u, v, 3, procedure.endswith("covariant"))
        assert np.array_equal(gu, ru, equal_nan=True) and np.array_equal(gv, rv, equal_nan=True)


def test_ecs_halo_vec_one_context_many_widths(oracle_lib):
    N = 16
    op = _op(N)
    C = oracle_lib.CGrid(N)
    _, u, v = _fields(C, 3)
    for hw, procedure in ((3, "ecs_C_vec"), (1, "ecs_C_vec_covariant"), (5, "ecs_C_vec"), (3, "ecs_C_vec"),
                          (2, "ecs_C_vec_covariant")):
        print(f"[c_ecs_halo_vec] same context hw={hw} {procedure}", flush=True)
        gu, gv = op.ecs_halo_vector(u, v, hw, procedure, halo_width_in=C.hw)
        ru, rv = C.ecs_halo_vec(u, v, hw, procedure.endswith("covariant"))
        assert np.array_equal(gu, ru, equal_nan=True) and np.array_equal(gv, rv, equal_nan=True)


@pytest.mark.parametrize("N", [8, 32, 65])
def test_grad_c2_ecs_matches_oracle(oracle_lib, N):
    print(f"[c_grad_c2_ecs] N={N}", flush=True)
    op = _op(N)
    C = oracle_lib.CGrid(N)
    f, _, _ = _fields(C, 11 + N)
    gx, gy = op.calc_grad(f, "gradient_c2_ecs", halo_width=C.hw)
    rx, ry = C.grad_c2_ecs(f)
    hw = C.hw
    # x points i = 1..N+1, j = 0..N+1; y points i = 0..N+1, j = 1..N+1
    sx = (slice(None), slice(hw - 1, hw + N + 1), slice(hw, hw + N + 1))
    sy = (slice(None), slice(hw, hw + N + 1), slice(hw - 1, hw + N + 1))
    assert np.array_equal(gx[sx], rx[sx])
    assert np.array_equal(gy[sy], ry[sy])


@pytest.mark.parametrize("N", [8, 32, 65])
def test_massflux_up4_matches_oracle(oracle_lib, N):
    print(f"[c_massflux_up4] N={N}", flush=True)
    op = _op(N)
    C = oracle_lib.CGrid(N)
    f, u, v = _fields(C, 23 + N)
    # exact zeros of both signs in the velocities
    hw = C.hw
    u[:, hw:hw + N, hw + 2] = 0.0
    u[:, hw:hw + N, hw + 3] = -0.0
    v[:, hw + 1, hw:hw + N] = -0.0
    fx, fy = op.calc_massflux(f, u, v, "massflux_c_up4", halo_width=hw)
    rx, ry = C.massflux_up4(f, u, v)
    ix, iy = C.interior("x", fx), C.interior("y", fy)
    assert np.array_equal(ix, C.interior("x", rx)) and np.array_equal(iy, C.interior("y", ry))
    assert np.array_equal(np.signbit(ix), np.signbit(C.interior("x", rx)))
    assert np.array_equal(np.signbit(iy), np.signbit(C.interior("y", ry)))


def test_c_ecs_rejects():
    import cubedsphere_sbp_swm_amd as swm
    N = 8
    op = _op(N)
    z = np.zeros(op.host_shape(0))
    with pytest.raises(swm.SwmError):
        op.ecs_halo_vector(z, z, 6)
    with pytest.raises(swm.SwmError):
        op.ecs_halo_vector(z, z, 2, "ecs_A_vec")
    with pytest.raises(swm.SwmError):
        op.calc_grad(z, "gradient_c_sbp42", halo_width=0)
    with pytest.raises(swm.SwmError):
        op.calc_massflux(z, z, z, "massflux_c_sbp42", halo_width=0)
