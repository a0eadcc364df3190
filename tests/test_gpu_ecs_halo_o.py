"""ECS cubic halo on O points on the device (swm_ecs_halo_o) against the oracle's restatement of
the reference's ECS_O procedure with its width-8 O-point exchange: FP64, the same weights and
operation order, so every written point is expected bit for bit and the unwritten corner-square
points (NaN, hw >= 3) in the same places.  Panel sizes N = 8 (the smallest the width-8 exchange
allows), 16 and 33; halo widths 1, 2 (the width the reference's lat-lon regrid uses), 3 and 8;
random fields and a smooth field.  FP32 contexts: 1e-5 relative to the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(G, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((6, G.N, G.N))


@pytest.mark.parametrize("N,hw", [(8, 1), (8, 2), (16, 2), (16, 3), (33, 2), (33, 8)])
def test_ecs_halo_o_matches_oracle(oracle_lib, N, hw):
    import cubedsphere_sbp_swm_amd as swm
    op = swm.create_swm_operator(N, **swm.operator_names(42))
    G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
    print(f"[ecs_halo_o] N={N} hw={hw}", flush=True)
    for seed in (1, 2):
        f = _field(G, seed + N)
        got = op.ecs_halo_o(f, hw)
        ref = G.ecs_halo_o(f, hw)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        assert np.array_equal(got, ref, equal_nan=True)
    assert (np.isnan(ref).sum() == 0) == (hw <= 2)


def test_ecs_halo_o_fp32(oracle_lib):
    import cubedsphere_sbp_swm_amd as swm
    N, hw = 16, 2
    op = swm.create_swm_operator(N, **swm.operator_names(42), precision=32)
    G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
    f = _field(G, 7).astype(np.float32).astype(np.float64)
    got, ref = op.ecs_halo_o(f, hw), G.ecs_halo_o(f, hw)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def test_ecs_halo_o_rejects():
    import cubedsphere_sbp_swm_amd as swm
    op = swm.create_swm_operator(16, **swm.operator_names(42))
    with pytest.raises(swm.SwmError):
        op.ecs_halo_o(np.zeros((6, 16, 16)), 0)
    with pytest.raises(swm.SwmError):
        op.ecs_halo_o(np.zeros((6, 16, 16)), 9)
    with pytest.raises(ValueError):
        op.ecs_halo_o(np.zeros((6, 15, 15)), 2)
