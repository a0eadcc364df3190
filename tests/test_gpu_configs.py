"""GPU parity at the sizes of the configured workloads, against the oracle.

* C384 sbp42 FP64 (the throughput workload): 3 RK4 steps at dt = 75 s from the degree-32
  seeded spherical-harmonic state, for the autotuned tiling and four fixed ones.  Panels of
  more than 256 points take the 64-point edge segments, which the smaller parity cases do not
  reach.  Bound: bit for bit (asserted as < 1e-13 relative per field, exact equality
  reported).
* C768 sbp42 FP32 (the precision form): 3 RK4 steps at dt = 37.5 s (gravity-wave Courant
  number 0.9) from the degree-64 state, against the FP64 oracle.  Bounds per field (relative
  L2 / max norm; h measured as h - 1e4 against the oracle's h - 1e4): 1e-4 / 1e-3 for h',
  u and v, and a mass drift against the oracle below 1e-7.
* The 64-member ensemble (16 members each at C48, C96, C144, C192; sbp42 FP64; member seeds
  0..63, degree 16; per-member dt at Courant 0.5 of gravity wave plus 40 m/s): 2 ensemble
  steps, every member against the oracle, bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A, G0, OM = 0.6371229e7, 9.80616, 2.0 * np.pi / (23 * 60 * 60 + 56 * 60 + 4)


def _op(N, precision=64):
    import cubedsphere_sbp_swm_amd as swm
    return swm, swm.create_swm_operator(N, **swm.operator_names(42), scale=A, omega=OM, grav=G0,
                                        precision=precision)


def _relerr(G, a, b):
    ai, bi = G.interior(a), G.interior(b)
    return float(np.max(np.abs(ai - bi)) / np.max(np.abs(bi)))


@pytest.fixture(scope="module")
def c384(oracle_lib):
    swm, op = _op(384)
    G = oracle_lib.Grid(384, 42, 42, A, OM, G0)
    ic = op.initial_state(seed=0, lmax=32, halo_width=G.hw)
    ref = [f.copy() for f in ic]
    G.rk4(*ref, 75.0, 3)
    return swm, op, G, ic, ref


@pytest.mark.parametrize("tiling", ["auto", (16, False), (8, True), (32, False), (8, False, 32)])
def test_c384_sbp42_fp64_three_steps_bitwise(c384, tiling):
    swm, op, G, ic, ref = c384
    x = swm.create_stvec_swm(op)
    x.upload(*ic, G.hw)
    if tiling == "auto":
        op.autotune(x, reps=5)
        x.upload(*ic, G.hw)
    else:
        op.set_tiling(*tiling)
    print(f"[c384] tiling {op.get_tiling()}", flush=True)
    swm.create_timescheme(op, "rk4").steps(x, 75.0, 3)
    got = x.download(G.hw)
    errs = [_relerr(G, a, b) for a, b in zip(got, ref)]
    same = [bool(np.array_equal(G.interior(a), G.interior(b))) for a, b in zip(got, ref)]
    print(f"  rel err h,u,v = {errs}, bitwise = {same}")
    assert max(errs) < 1e-13


def test_c768_sbp42_fp32_against_fp64_oracle(oracle_lib):
    swm, op = _op(768, precision=32)
    G = oracle_lib.Grid(768, 42, 42, A, OM, G0)
    ic = op.initial_state(seed=0, lmax=64, halo_width=G.hw)
    dt, nsteps = 37.5, 3
    print("[c768 fp32] oracle", flush=True)
    ref = [f.copy() for f in ic]
    G.rk4(*ref, dt, nsteps)
    x = swm.create_stvec_swm(op)
    x.upload(*ic, G.hw)
    swm.create_timescheme(op, "rk4").steps(x, dt, nsteps)
    got = x.download(G.hw)
    errs = {}
    for name, a, b, sub in (("h", got[0], ref[0], 1.0e4), ("u", got[1], ref[1], 0.0), ("v", got[2], ref[2], 0.0)):
        ai, bi = G.interior(a) - sub, G.interior(b) - sub
        errs[name] = (float(np.sqrt(np.sum((ai - bi) ** 2) / np.sum(bi ** 2))),
                      float(np.max(np.abs(ai - bi)) / np.max(np.abs(bi))))
    drift = (G.quad_mass(42, got[0]) - G.quad_mass(42, ref[0])) / G.quad_mass(42, ref[0])
    print(f"  FP32 C768 {nsteps} steps (L2, Linf) rel: {errs}, mass vs oracle {drift:.3e}")
    for l2, linf in errs.values():
        assert l2 < 1e-4 and linf < 1e-3
    assert abs(drift) < 1e-7


def test_ensemble_64_members_against_oracle(oracle_lib):
    import cubedsphere_sbp_swm_amd as swm
    sizes = [48] * 16 + [96] * 16 + [144] * 16 + [192] * 16
    grids = {N: oracle_lib.Grid(N, 42, 42, A, OM, G0) for N in set(sizes)}
    hw = grids[48].hw
    ops, xs, ics = [], [], []
    for k, N in enumerate(sizes):
        op = swm.create_swm_operator(N, **swm.operator_names(42), scale=A, omega=OM, grav=G0)
        ic = op.initial_state(seed=k, lmax=16, halo_width=hw)
        x = swm.create_stvec_swm(op)
        x.upload(*ic, hw)
        ops.append(op), xs.append(x), ics.append(ic)
    ens = swm.Ensemble(ops, xs)
    us = ens.autotune(reps=3)
    for x, ic in zip(xs, ics):
        x.upload(*ic, hw)
    print(f"[ensemble 64] tiling {ens.get_tiling()} ({us:.1f} us per stage-2 launch)", flush=True)
    dts = [0.5 * A * (np.pi / 2) / N / (np.sqrt(G0 * 1.0e4) + 40.0) for N in sizes]
    ens.steps(dts, 2)
    worst = 0.0
    for k, (N, x, ic, dt) in enumerate(zip(sizes, xs, ics, dts)):
        G = grids[N]
        ref = [f.copy() for f in ic]
        G.rk4(*ref, dt, 2)
        got = x.download(hw)
        for a, b in zip(got, ref):
            worst = max(worst, _relerr(G, a, b))
            assert np.array_equal(G.interior(a), G.interior(b)), (k, N)
    print(f"  64 members bit for bit (max rel err {worst:.1e})")
