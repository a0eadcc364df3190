"""ECS cubic halo on the device (swm_ecs_halo_xy) against the oracle's restatement of
ecs_halo_xy_t%get_halo_scalar: every point of the extended panels bit for bit in FP64 (same
weights, same operation order), random and analytic fields, halo widths 1..3; FP32 contexts
to 1e-6 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ext_from_oracle(G, f, hw):
    H, n = G.hw, G.n
    return f[:, H - hw:H + n + hw, H - hw:H + n + hw]


@pytest.mark.parametrize("N,hw", [(16, 1), (16, 2), (16, 3), (32, 3), (21, 2)])
@pytest.mark.parametrize("kind", ["random", "xyz"])
def test_ecs_halo_matches_oracle(oracle_lib, N, hw, kind):
    import cubedsphere_sbp_swm_amd as swm
    op = swm.create_swm_operator(N, **swm.operator_names(42))
    G = oracle_lib.Grid(N)
    n = G.n
    if kind == "random":
        vals = np.random.default_rng(N * 10 + hw).standard_normal((6, n, n))
    else:
        vals = G.interior(G.geometry("r")[0]).copy()
    print(f"[ecs] N={N} hw={hw} {kind}", flush=True)
    x = swm.create_stvec_swm(op)
    x.upload(vals, vals, vals)
    got = op.ecs_halo_xy(x, 1, hw)
    f = G.embed(vals)
    G.ecs_halo_xy(f, hw)
    ref = _ext_from_oracle(G, f, hw)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


def test_ecs_halo_fp32(oracle_lib):
    import cubedsphere_sbp_swm_amd as swm
    N, hw = 32, 3
    op = swm.create_swm_operator(N, **swm.operator_names(42), precision=32)
    G = oracle_lib.Grid(N)
    vals = G.interior(G.geometry("r")[2]).copy()
    x = swm.create_stvec_swm(op)
    x.upload(vals, vals, vals)
    got = op.ecs_halo_xy(x, 0, hw)
    f = G.embed(vals)
    G.ecs_halo_xy(f, hw)
    ref = _ext_from_oracle(G, f, hw)
    assert np.max(np.abs(got - ref)) < 1e-6


def test_ecs_halo_rejects_bad_width():
    import cubedsphere_sbp_swm_amd as swm
    op = swm.create_swm_operator(8, **swm.operator_names(42))
    x = swm.create_stvec_swm(op)
    with pytest.raises(swm.SwmError):
        op.ecs_halo_xy(x, 0, 9)
    with pytest.raises(swm.SwmError):
        op.ecs_halo_xy(x, 0, 7)  # N must be >= halo_width + 3
