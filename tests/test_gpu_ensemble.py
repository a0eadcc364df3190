"""Batched ensemble form: cubes of different resolutions and time steps advanced by one
launch per RK stage (a persistent grid, or one block per task; hipGraph per step) give, member by member, the same states as
the single-cube path (which equals the oracle bit for bit): bitwise equality asserted."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A, G0, OM = 0.6371229e7, 9.80616, 2.0 * np.pi / (23 * 60 * 60 + 56 * 60 + 4)


def _members(Ns, order, seed0=0):
    import cubedsphere_sbp_swm_amd as swm
    ops, xs, ics = [], [], []
    for k, N in enumerate(Ns):
        op = swm.create_swm_operator(N, **swm.operator_names(order), scale=A, omega=OM, grav=G0)
        ic = op.initial_state(seed=seed0 + k, lmax=8)
        x = swm.create_stvec_swm(op)
        x.upload(*ic)
        ops.append(op)
        xs.append(x)
        ics.append(ic)
    return swm, ops, xs, ics


def _dt(N):
    # gravity-wave + advective Courant number 0.5 (the ensemble configuration's rule)
    return 0.5 * A * (np.pi / 2) / N / (np.sqrt(G0 * 1.0e4) + 40.0)


@pytest.mark.parametrize("tiling", [(8, False, 1), (16, True, 2), (16, False, 4), (8, False, 2, 32), (8, True, 1, 32),
                                    (8, True, 0, 32), (16, False, 0), "auto"])
def test_ensemble_equals_single_cube_path(tiling):
    Ns = [16, 24, 33, 48]
    swm, ops, xs, ics = _members(Ns, 42)
    ens = swm.Ensemble(ops, xs)
    if tiling == "auto":
        us = ens.autotune(reps=3)
        for x, ic in zip(xs, ics):
            x.upload(*ic)
        print("  autotuned", ens.get_tiling(), us)
    else:
        ens.set_tiling(*tiling)
    dts = [_dt(N) for N in Ns]
    print(f"[ensemble] tiling={tiling}", flush=True)
    ens.steps(dts, 6)
    got = [x.download() for x in xs]
    for op, ic, dt, g in zip(ops, ics, dts, got):
        y = swm.create_stvec_swm(op)
        y.upload(*ic)
        swm.create_timescheme(op, "rk4").steps(y, dt, 6)
        for a, b in zip(g, y.download()):
            assert np.array_equal(a, b)


def test_ensemble_sbp63_and_repeated_calls():
    Ns = [24, 40, 17]
    swm, ops, xs, ics = _members(Ns, 63, seed0=10)
    ens = swm.Ensemble(ops, xs)
    dts = [_dt(N) for N in Ns]
    ens.steps(dts, 2)
    ens.steps(dts, 3)
    for op, ic, dt, x in zip(ops, ics, dts, xs):
        y = swm.create_stvec_swm(op)
        y.upload(*ic)
        swm.create_timescheme(op, "rk4").steps(y, dt, 5)
        for a, b in zip(x.download(), y.download()):
            assert np.array_equal(a, b)


def test_ensemble_rejects_mixed_families_and_shared_members():
    import cubedsphere_sbp_swm_amd as swm
    a = swm.create_swm_operator(20, **swm.operator_names(42))
    b = swm.create_swm_operator(20, **swm.operator_names(63))
    xa, xb = swm.create_stvec_swm(a), swm.create_stvec_swm(b)
    with pytest.raises(swm.SwmError):
        swm.Ensemble([a, b], [xa, xb])
    with pytest.raises(swm.SwmError):
        swm.Ensemble([a, a], [xa, swm.create_stvec_swm(a)])


def test_ensemble_follows_orography_changes():
    """swm_set_orography replaces a member's orography buffer: ensemble steps before and after
    the change use the member's current orography (the single-cube path with the same
    orography sequence gives the same states)."""
    Ns = [20, 28]
    swm, ops, xs, ics = _members(Ns, 42, seed0=30)
    ens = swm.Ensemble(ops, xs)
    dts = [_dt(N) for N in Ns]

    def mountain(op, amp):
        f = np.zeros(op.host_shape())
        f[2, op.n // 3: 2 * op.n // 3, op.n // 4: op.n // 2] = amp
        return f

    ens.steps(dts, 2)
    ops[1].set_orography(mountain(ops[1], 1500.0))
    ens.steps(dts, 2)
    ops[1].set_orography(mountain(ops[1], 700.0))
    ops[0].set_orography(mountain(ops[0], 300.0))
    ens.steps(dts, 2)
    refs = []
    for k, (op, ic, dt) in enumerate(zip(ops, ics, dts)):
        op.set_orography(None)
        y = swm.create_stvec_swm(op)
        y.upload(*ic)
        ts = swm.create_timescheme(op, "rk4")
        ts.steps(y, dt, 2)
        if k == 1:
            op.set_orography(mountain(op, 1500.0))
        ts.steps(y, dt, 2)
        op.set_orography(mountain(op, 700.0 if k == 1 else 300.0))
        ts.steps(y, dt, 2)
        refs.append(y.download())
    for x, ref in zip(xs, refs):
        for a, b in zip(x.download(), ref):
            assert np.array_equal(a, b)
