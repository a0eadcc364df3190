"""FP32 form: stencils, edge syncs and RK stages in FP32 with the metric precomputed in FP64
(rounded once to FP32) and FP64 compensated integrals.  Errors are reported per field
against the FP64 oracle; tolerances are stated per test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A, G0, OM = 0.6371229e7, 9.80616, 2.0 * np.pi / (23 * 60 * 60 + 56 * 60 + 4)


def _pair(oracle_lib, N, order, seed, lmax):
    import cubedsphere_sbp_swm_amd as swm
    op = swm.create_swm_operator(N, **swm.operator_names(order), scale=A, omega=OM, grav=G0, precision=32)
    G = oracle_lib.Grid(N, order, order if order != 21 else 2, A, OM, G0)
    h, u, v = op.initial_state(seed=seed, lmax=lmax, halo_width=G.hw)
    return swm, op, G, h, u, v


def _errs(G, a, b, sub=0.0):
    ai, bi = G.interior(a) - sub, G.interior(b) - sub
    l2 = np.sqrt(np.sum((ai - bi) ** 2) / np.sum(bi ** 2))
    linf = np.max(np.abs(ai - bi)) / np.max(np.abs(bi))
    return float(l2), float(linf)


@pytest.mark.parametrize("order", [42, 21, 63])
def test_fp32_rhs_against_fp64_oracle(oracle_lib, order):
    swm, op, G, h, u, v = _pair(oracle_lib, 32, order, seed=4, lmax=8)
    x = swm.create_stvec_swm(op)
    k = x.create_similar()
    x.upload(h, u, v, G.hw)
    op.apply(k, x)
    ours = k.download(G.hw)
    ref = G.rhs(h.copy(), u.copy(), v.copy())
    errs = [_errs(G, a, b) for a, b in zip(ours, ref)]
    print(f"  FP32 RHS order {order}: (L2, Linf) rel per field h,u,v = {errs}")
    # The initial velocity is non-divergent (V = r x grad psi), so the h tendency -div(h V) is
    # ~1e-3 of its terms h0*div(V) and V.grad(h'): FP32 round-off in h*u (eps*|h u|*C/dx) is
    # amplified by that cancellation to ~1e-4 of the tendency; the u, v tendencies stay ~1e-5.
    for l2, linf in errs:
        assert l2 < 5e-4 and linf < 2e-3


def test_fp32_rk4_trajectory(oracle_lib):
    swm, op, G, h, u, v = _pair(oracle_lib, 48, 42, seed=6, lmax=12)
    dt, nsteps = 600.0, 40
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v, G.hw)
    d0 = op.get_diagnostics(x)
    swm.create_timescheme(op, "rk4").steps(x, dt, nsteps)
    gh, gu, gv = x.download(G.hw)
    d1 = op.get_diagnostics(x)
    G.rk4(h, u, v, dt, nsteps)
    eh = _errs(G, gh, h, sub=1.0e4)
    eu, ev = _errs(G, gu, u), _errs(G, gv, v)
    drift = (d1["mass"] - d0["mass"]) / d0["mass"]
    print(f"  FP32 40 steps: h' {eh}, u {eu}, v {ev}, mass drift {drift:.3e}")
    assert eh[0] < 1e-3 and eu[0] < 1e-3 and ev[0] < 1e-3
    assert abs(drift) < 1e-6
