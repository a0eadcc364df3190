"""GPU parity of the C-grid SBP42 path (swm_c_*) against the oracle, bit for bit.

States: test case 2 (solid-body rotation in geostrophic balance, Earth parameters) plus a
Gaussian height bump and a random velocity perturbation on every face, so that the SAT
penalties (which vanish for smooth fields) and every cube edge and corner carry non-trivial
values.  Bound: exact equality of every mesh point of h, u, v (asserted as max |diff| == 0,
the relative error printed for reference).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A, G0 = 0.6371229e7, 9.80616
OM = 2.0 * np.pi / (23 * 60 * 60 + 56 * 60 + 4)
HM = 29400 / G0
U0 = np.pi * A / 6.0 / (23 * 60 * 60 + 56 * 60 + 4)


def _state(o, C, seed=0, noise=2.0):
    def unit(t):
        x, y, z = C.geometry(t, "r")
        r = np.sqrt(x * x + y * y + z * z)
        return x / r, y / r, z / r
    xo, yo, zo = unit("o")
    c2 = (A * OM * U0 + 0.5 * U0 ** 2) / G0
    cc = (np.cos(0.3) * np.cos(0.7), np.cos(0.3) * np.sin(0.7), np.sin(0.3))
    d = np.arccos(np.clip(xo * cc[0] + yo * cc[1] + zo * cc[2], -1, 1))
    h = C.embed("o", C.interior("o", HM - c2 * zo ** 2 + 120.0 * np.exp(-(d / 0.35) ** 2)))
    rng = np.random.default_rng(seed)
    out = [h]
    for t, comp in (("x", 0), ("y", 1)):
        x, y, z = unit(t)
        f = C.project(t, "covariant", -U0 * y, U0 * x, 0 * z, comp)
        f = C.interior(t, f) + noise * rng.standard_normal(C.interior(t, f).shape)
        out.append(C.embed(t, f))
    return out


def _op(N):
    import cubedsphere_sbp_swm_amd as swm
    return swm, swm.create_swm_operator(N, **swm.operator_names("c42"), scale=A, omega=OM, grav=G0)


def _compare(C, got, ref, label):
    worst = 0.0
    for t, a, b in zip("oxy", got, ref):
        ai, bi = C.interior(t, a), C.interior(t, b)
        diff = float(np.max(np.abs(ai - bi)))
        worst = max(worst, diff / float(np.max(np.abs(bi))))
        assert diff == 0.0, (label, t, diff)
    print(f"  {label}: bitwise (max rel {worst:.1e})", flush=True)


@pytest.mark.parametrize("N", [8, 32, 65])
def test_c_rhs_bitwise(oracle_lib, N):
    print(f"[c rhs N={N}]", flush=True)
    swm, op = _op(N)
    C = oracle_lib.CGrid(N, A, OM, G0)
    h, u, v = _state(oracle_lib, C, seed=N)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v, C.hw)
    k = x.create_similar()
    op.apply(k, x)
    got = k.download(C.hw)
    ref = C.rhs(h.copy(), u.copy(), v.copy())
    _compare(C, got, ref, f"RHS N={N}")


def test_c_rhs_with_orography_bitwise(oracle_lib):
    N = 24
    print(f"[c rhs orography N={N}]", flush=True)
    swm, op = _op(N)
    C = oracle_lib.CGrid(N, A, OM, G0)
    h, u, v = _state(oracle_lib, C, seed=5)
    xo, yo, zo = C.geometry("o", "r")
    hs = C.embed("o", C.interior("o", 500.0 * np.cos(3 * np.arctan2(yo, xo)) * (1 - zo ** 2)))
    op.set_orography(hs, C.hw)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v, C.hw)
    k = x.create_similar()
    op.apply(k, x)
    ref = C.rhs(h.copy(), u.copy(), v.copy(), hs)
    _compare(C, k.download(C.hw), ref, "RHS with orography")
    op.set_orography(None)
    op.apply(k, x)
    _compare(C, k.download(C.hw), C.rhs(h.copy(), u.copy(), v.copy()), "RHS orography removed")


@pytest.mark.parametrize("N,nsteps", [(32, 3), (96, 2)])
def test_c_rk4_bitwise(oracle_lib, N, nsteps):
    print(f"[c rk4 N={N} steps={nsteps}]", flush=True)
    swm, op = _op(N)
    C = oracle_lib.CGrid(N, A, OM, G0)
    h, u, v = _state(oracle_lib, C, seed=1)
    dt = 0.4 * (A * np.pi / 2 / N) / np.sqrt(G0 * HM)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v, C.hw)
    swm.create_timescheme(op, "rk4").steps(x, dt, nsteps)
    ref = [a.copy() for a in (h, u, v)]
    C.rk4(*ref, dt, nsteps)
    _compare(C, x.download(C.hw), ref, f"RK4 N={N} x{nsteps}")


def test_c_rk4_multi_step_graph_bitwise(oracle_lib):
    """53 steps: one replay of the 50-step graph and three of the one-step graph, then 2 more
    steps on the cached graphs."""
    N = 12
    print(f"[c rk4 graphs N={N}]", flush=True)
    swm, op = _op(N)
    C = oracle_lib.CGrid(N, A, OM, G0)
    h, u, v = _state(oracle_lib, C, seed=2, noise=0.5)
    dt = 0.1 * (A * np.pi / 2 / N) / np.sqrt(G0 * HM)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v, C.hw)
    ts = swm.create_timescheme(op, "rk4")
    ts.steps(x, dt, 53)
    ts.steps(x, dt, 2)
    ref = [a.copy() for a in (h, u, v)]
    C.rk4(*ref, dt, 55)
    _compare(C, x.download(C.hw), ref, "RK4 55 steps")


@pytest.mark.parametrize("N,with_hs", [(24, False), (64, True)])
def test_c_diagnostics_match_oracle(oracle_lib, N, with_hs):
    """hmin, hmax exactly; mass, TE, KE, PE within 1e-12 relative (compensated device sums against
    the oracle's serial sums of the same per-point terms); enstrophy NaN on both sides."""
    print(f"[c diagnostics N={N} hs={with_hs}]", flush=True)
    swm, op = _op(N)
    C = oracle_lib.CGrid(N, A, OM, G0)
    h, u, v = _state(oracle_lib, C, seed=7)
    hs = None
    if with_hs:
        xo, yo, zo = C.geometry("o", "r")
        hs = C.embed("o", C.interior("o", 300.0 * np.sin(2 * np.arctan2(yo, xo)) * (1 - zo ** 2)))
        op.set_orography(hs, C.hw)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v, C.hw)
    got = op.get_diagnostics(x)
    ref = C.diagnostics(h, u, v, hs)
    assert got["hmin"] == ref["hmin"] and got["hmax"] == ref["hmax"]
    for k in ("mass", "TE", "KE", "PE"):
        assert abs(got[k] - ref[k]) <= 1e-12 * abs(ref[k]), (k, got[k], ref[k])
    assert np.isnan(got["Enstrophy"]) and np.isnan(ref["Enstrophy"])
    print(f"  diagnostics agree: {got}")


def test_c384_cgrid_two_steps_bitwise(oracle_lib):
    """The C-grid measurement's size: C384, 2 RK4 steps at dt = 37.5 s (gravity-wave Courant number
    0.45; the C-grid scheme's RK4 limit lies near 0.5), every point against the oracle."""
    N = 384
    print(f"[c384 c-grid]", flush=True)
    swm, op = _op(N)
    C = oracle_lib.CGrid(N, A, OM, G0)
    h, u, v = _state(oracle_lib, C, seed=11, noise=1.0)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v, C.hw)
    swm.create_timescheme(op, "rk4").steps(x, 37.5, 2)
    ref = [a.copy() for a in (h, u, v)]
    C.rk4(*ref, 37.5, 2)
    _compare(C, x.download(C.hw), ref, "C384 RK4 x2")
