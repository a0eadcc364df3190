"""Pins of the oracle's C-grid SBP42 family (staggering 'C': h at cell centres, u and v at the
cell faces, curl at the cell corners).

* Each operator reproduces the reference's printed error norms (test_diffops_all, N=32, three
  levels) to all 7 digits: divergence_c_sbp42, gradient_c_sbp42, curl_c_sbp42,
  coriolis_Cgrid_sbp42 (vector-invariant form, test_diffops_mod.f90:723-789), KE_Cgrid_sbp42
  and co2contra_c_sbp42.  These fix the staggered SBP weights, the C vector halo (shifted faces,
  component signs, corner rows), the O halo, the q-point sync and the metric factors.
* What smooth fields leave open -- the interface penalties of the divergence and the curl,
  which vanish where the neighbour's copy of a shared face equals the panel's own -- is fixed by
  the SBP conservation identities on random fields: sum A_i A_j J div = 0 with the centre
  norm A = [13/12, 7/8, 25/24, 1, ...], and sum A_i A_j J curl = 0 with the interface norm
  [7/18, 9/8, 1, 71/72, 1, ...] over every copy of the corner points.
* The composition of operator_swm_t%apply with these operators is fixed by the test case 2
  steady state (the RHS vanishes as the grid is refined), the RK4 combination by its temporal
  order, and the sign of the continuity equation by stability of the gravity-wave system.
"""
import os

import numpy as np
import pytest

from oracle import test_fields as tf

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "diffops_reference_results_c.txt")

A_EARTH = 0.6371229e7                      # const_mod.f90:7
G_EARTH = 9.80616                          # const_mod.f90:6
T_SID = 23 * 60 * 60 + 56 * 60 + 4         # const_mod.f90:8
OMEGA = 2.0 * np.pi / T_SID                # const_mod.f90:9
H_MEAN = 29400 / G_EARTH                   # config_ts2_mod.f90:14
U0 = np.pi * A_EARTH / 6.0 / T_SID         # config_ts2_mod.f90:16
C_TS2 = (A_EARTH * OMEGA * U0 + 0.5 * U0 ** 2) / G_EARTH


def _golden():
    out = {}
    for line in open(GOLDEN):
        if line.startswith("#") or "|" not in line:
            continue
        k, v = line.split("|")
        out[k.strip()] = [float(x) for x in v.split()]
    return out


def _same_print(ours, ref):
    """True when `ours` prints as `ref` in Fortran E15.7 (7 significant digits)."""
    for a, b in zip(ours, ref):
        e = np.floor(np.log10(abs(b))) + 1
        if abs(a - b) > 0.5 * 10.0 ** (e - 7) * 1.0001:
            return False
    return True


def _levels(C, t, fields):
    return max(C.maxabs(t, f) for f in fields), np.sqrt(sum(C.l2norm(t, f) ** 2 for f in fields))


def _vec(C, t, gen, kind, level):
    x, y, z = C.geometry(t, "r")
    vx, vy, vz = gen(x, y, z, level)
    return C.embed(t, C.interior(t, C.project(t, kind, vx, vy, vz, 0 if t == "x" else 1)))


def _uv(C, gen, kind, level):
    return _vec(C, "x", gen, kind, level), _vec(C, "y", gen, kind, level)


@pytest.fixture(scope="module")
def c32(oracle_lib):
    return oracle_lib.CGrid(32)


def test_golden_divergence_c_sbp42(c32):
    # test_div, test_diffops_mod.f90:71-125
    C = c32
    xo, yo, zo = C.geometry("o", "r")
    ds = [C.div(*_uv(C, tf.solid_rotation, "contravariant", k)) for k in (1, 2, 3)]
    res = list(_levels(C, "o", ds))
    ds = [C.embed("o", C.interior("o", 1.0 * C.div(*_uv(C, tf.cross_polar, "contravariant", k))
                                   + (-1.0) * tf.cross_polar_div(xo, yo, zo))) for k in (1, 2, 3)]
    res += _levels(C, "o", ds)
    assert _same_print(res, _golden()["divergence_c_sbp42"]), res


def test_golden_gradient_c_sbp42(c32):
    # test_grad, test_diffops_mod.f90:127-178
    C = c32
    xo, yo, zo = C.geometry("o", "r")
    ex, ey = [], []
    for k in (1, 2, 3):
        gx, gy = C.grad(C.embed("o", C.interior("o", tf.xyz_scalar(xo, yo, zo, k))))
        tx, ty = _uv(C, tf.xyz_grad, "covariant", k)
        ex.append(C.embed("x", C.interior("x", 1.0 * gx + (-1.0) * tx)))
        ey.append(C.embed("y", C.interior("y", 1.0 * gy + (-1.0) * ty)))
    (m1, l1), (m2, l2) = _levels(C, "x", ex), _levels(C, "y", ey)
    assert _same_print([m1 + m2, l1 + l2], _golden()["gradient_c_sbp42"]), (m1 + m2, l1 + l2)


def test_golden_curl_c_sbp42(c32):
    # test_curl, test_diffops_mod.f90:466-522
    C = c32
    xq, yq, zq = C.geometry("q", "r")
    cs = [C.embed("q", C.interior("q", C.curl(*_uv(C, tf.vsh_curl_free_10, "covariant", k))))
          for k in (1, 2, 3)]
    res = list(_levels(C, "q", cs))
    cs = [C.embed("q", C.interior("q", 1.0 * C.curl(*_uv(C, tf.cross_polar, "covariant", k))
                                   + (-1.0) * tf.cross_polar_zeta(xq, yq, zq))) for k in (1, 2, 3)]
    res += _levels(C, "q", cs)
    assert _same_print(res, _golden()["curl_c_sbp42"]), res


def test_golden_coriolis_cgrid_sbp42(c32):
    # test_coriolis_vec_inv, test_diffops_mod.f90:723-789: h = 1, curl = 0, hu, hv = the
    # contravariant VSH_curl_free_10 field; exact answer f x V as covariant components
    C = c32
    h = C.embed("o", np.ones((6, 32, 32)))
    curl = C.zeros()
    eu, ev = [], []
    for k in (1, 2, 3):
        cu, cv = C.coriolis(*_uv(C, tf.vsh_curl_free_10, "contravariant", k), h, curl)
        tu, tv = _uv(C, tf.coriolis_force(tf.vsh_curl_free_10), "covariant", k)
        eu.append(C.embed("x", C.interior("x", cu + (-1.0) * tu)))
        ev.append(C.embed("y", C.interior("y", cv + (-1.0) * tv)))
    (m1, l1), (m2, l2) = _levels(C, "x", eu), _levels(C, "y", ev)
    assert _same_print([m1 + m2, l1 + l2], _golden()["coriolis_Cgrid_sbp42"]), (m1 + m2, l1 + l2)


def test_golden_ke_cgrid_sbp42(c32):
    # test_KE, test_diffops_mod.f90:523-593
    C = c32
    xo, yo, zo = C.geometry("o", "r")
    res = []
    for gen in (tf.vsh_curl_free_10, tf.cross_polar):
        es = []
        for k in (1, 2, 3):
            KE = C.KE(*_uv(C, gen, "covariant", k), *_uv(C, gen, "contravariant", k))
            es.append(C.embed("o", C.interior("o", 1.0 * KE + (-1.0) * tf.kinetic_energy(gen)(xo, yo, zo, k))))
        res += _levels(C, "o", es)
    assert _same_print(res, _golden()["KE_Cgrid_sbp42"]), res


def test_golden_co2contra_c_sbp42(c32):
    # test_co2contra, test_diffops_mod.f90:231-278
    C = c32
    eu, ev = [], []
    for k in (1, 2, 3):
        ut, vt = C.co2contra(*_uv(C, tf.cross_polar, "covariant", k))
        tu, tv = _uv(C, tf.cross_polar, "contravariant", k)
        eu.append(C.embed("x", C.interior("x", ut + (-1.0) * tu)))
        ev.append(C.embed("y", C.interior("y", vt + (-1.0) * tv)))
    (m1, l1), (m2, l2) = _levels(C, "x", eu), _levels(C, "y", ev)
    assert _same_print([m1 + m2, l1 + l2], _golden()["co2contra_c_sbp42"]), (m1 + m2, l1 + l2)


def _norm_weights(n, edge):
    w = np.ones(n)
    w[:len(edge)] = edge
    w[-len(edge):] = edge[::-1]
    return w


def test_sbp_conservation_identities_on_random_fields(oracle_lib):
    """The interface penalties vanish for smooth fields; on random face values they carry the
    jump between a panel's face value and its neighbour's copy.  Conservation of the SBP
    divergence and curl holds exactly only with the reference's penalty weights."""
    N = 20
    C = oracle_lib.CGrid(N)
    rng = np.random.default_rng(1)
    u = C.embed("x", rng.standard_normal((6, N, N + 1)))
    v = C.embed("y", rng.standard_normal((6, N + 1, N)))
    Ac = _norm_weights(N, [13 / 12, 7 / 8, 25 / 24])
    W = Ac[None, :, None] * Ac[None, None, :] * C.interior("o", C.geometry("o", "J"))
    d = C.interior("o", C.div(u, v))
    assert abs(np.sum(W * d)) / np.sum(W * np.abs(d)) < 1e-14
    Ai = _norm_weights(N + 1, [7 / 18, 9 / 8, 1.0, 71 / 72])
    Wq = Ai[None, :, None] * Ai[None, None, :] * C.interior("q", C.geometry("q", "J"))
    c = C.interior("q", C.curl(u.copy(), v.copy()))
    assert abs(np.sum(Wq * c)) / np.sum(Wq * np.abs(c)) < 1e-14
    # the curl is single-valued at shared corners and edges after the q-point sync
    assert np.all(np.isfinite(c))


def _unit(C, t):
    x, y, z = C.geometry(t, "r")
    r = np.sqrt(x * x + y * y + z * z)
    return x / r, y / r, z / r


def ts2_state_c(C):
    """TS2 height (o points) and covariant solid-rotation velocity (u at x, v at y points)."""
    xo, yo, zo = _unit(C, "o")
    h = C.embed("o", C.interior("o", H_MEAN - C_TS2 * zo ** 2))
    xx, yx, zx = _unit(C, "x")
    xy, yy, zy = _unit(C, "y")
    u = C.embed("x", C.interior("x", C.project("x", "covariant", -U0 * yx, U0 * xx, 0 * zx, 0)))
    v = C.embed("y", C.interior("y", C.project("y", "covariant", -U0 * yy, U0 * xy, 0 * zy, 1)))
    return h, u, v


def bump_c(C, amp=120.0, width=0.35):
    xo, yo, zo = _unit(C, "o")
    c = (np.cos(0.3) * np.cos(0.7), np.cos(0.3) * np.sin(0.7), np.sin(0.3))
    d = np.arccos(np.clip(xo * c[0] + yo * c[1] + zo * c[2], -1.0, 1.0))
    return C.embed("o", C.interior("o", amp * np.exp(-(d / width) ** 2)))


def _ts2_residuals(o, N):
    C = o.CGrid(N, A_EARTH, OMEGA, G_EARTH)
    h, u, v = ts2_state_c(C)
    dh, du, dv = C.rhs(h.copy(), u.copy(), v.copy())
    gx, gy = C.grad(C.embed("o", C.interior("o", G_EARTH * h)))
    s2 = np.sqrt(C.l2norm("x", gx) ** 2 + C.l2norm("y", gy) ** 2)
    si = max(C.maxabs("x", gx), C.maxabs("y", gy))
    sh = C.l2norm("o", h) * U0 / A_EARTH
    return (np.sqrt(C.l2norm("x", du) ** 2 + C.l2norm("y", dv) ** 2) / s2,
            max(C.maxabs("x", du), C.maxabs("y", dv)) / si, C.l2norm("o", dh) / sh)


def test_ts2_steady_state_c_rhs_converges(oracle_lib):
    """RHS(TS2) -> 0 with N (measured momentum L2 rate 2.35-2.41, max norm 1.96-1.98; the
    continuity residual, the divergence of the interpolated mass flux, L2 rate 1.34-1.42).
    A wrong Coriolis sign, a dropped 1/2 of the kinetic energy or a missing g*h leaves an O(1)
    momentum residual."""
    r = [_ts2_residuals(oracle_lib, N) for N in (16, 32, 64)]
    for k in range(2):
        assert np.log2(r[k][0] / r[k + 1][0]) > 2.1, r
        assert np.log2(r[k][1] / r[k + 1][1]) > 1.7, r
        assert np.log2(r[k][2] / r[k + 1][2]) > 1.15, r
    assert r[2][0] < 1e-3 and r[2][2] < 1e-2, r[2]


def _state_with_bump(o, N):
    C = o.CGrid(N, A_EARTH, OMEGA, G_EARTH)
    h, u, v = ts2_state_c(C)
    return C, h + bump_c(C), u, v


def test_c_rk4_stable_and_mass_conserving(oracle_lib):
    """40 RK4 steps at gravity-wave Courant number 0.5: the bump disperses without growth and
    the centre-norm mass is kept to round-off.  With dh = +div (a flipped continuity sign) the
    gravity-wave system is elliptic and the same run overflows."""
    N = 16
    C, h, u, v = _state_with_bump(oracle_lib, N)
    Ac = _norm_weights(N, [13 / 12, 7 / 8, 25 / 24])
    W = Ac[None, :, None] * Ac[None, None, :] * C.interior("o", C.geometry("o", "J"))
    m0 = np.sum(W * C.interior("o", h))
    h0 = h.copy()
    dt = 0.5 * (A_EARTH * np.pi / 2 / N) / np.sqrt(G_EARTH * H_MEAN)
    C.rk4(h, u, v, dt, 40)
    dev = C.interior("o", h - h0)
    assert np.all(np.isfinite(dev)) and np.max(np.abs(dev)) < 200.0
    assert abs(np.sum(W * C.interior("o", h)) - m0) / m0 < 1e-14


@pytest.mark.parametrize("courant", [0.125, 0.0625])
def test_c_rk4_temporal_order_is_four(oracle_lib, courant):
    """Fixed grid (C12), fixed interval of 8 steps at dt0: error against dt0/16 falls 16-fold
    per halving of dt (measured 16.2-16.5 at gravity-wave Courant number 0.125; the C grid's
    fastest discrete modes leave the asymptotic range above about 0.2)."""
    N = 12
    dt0 = courant * (A_EARTH * np.pi / 2 / N) / np.sqrt(G_EARTH * H_MEAN)

    def run(dt, n):
        C, h, u, v = _state_with_bump(oracle_lib, N)
        C.rk4(h, u, v, dt, n)
        return C, h, u, v

    C, *ref = run(dt0 / 16, 128)
    errs = []
    for m in (1, 2, 4):
        _, *s = run(dt0 / m, 8 * m)
        errs.append(max(np.max(np.abs(C.interior(t, a - b))) for t, a, b in zip("oxy", s, ref)))
    for k in range(2):
        assert 12.0 < errs[k] / errs[k + 1] < 20.0, errs


def test_c_quadrature_converges_to_sphere_area(oracle_lib):
    """SBP_C42_quadrature of 1 on o, x and y points -> 4 pi at fourth order (measured ratios
    15.1-16.3 per doubling); a wrong edge weight leaves an O(h) error."""
    errs = []
    for N in (16, 32, 64):
        C = oracle_lib.CGrid(N)
        errs.append([abs(C.quad_mass(t, C.embed(t, np.ones(C.interior(t, C.zeros()).shape))) - 4 * np.pi)
                     for t in "oxy"])
    for k in range(2):
        for t in range(3):
            assert errs[k][t] / errs[k + 1][t] > 13.0, errs


def test_c_diagnostics_conserve_mass_and_energy(oracle_lib):
    """The C-grid SBP scheme conserves mass exactly and energy up to the RK4 time error:
    100 steps at gravity-wave Courant number 0.2 move TE by ~1e-9 relative (measured 3.6e-10 -
    1.6e-9); a kinetic energy with a wrong factor or the wrong quadrature is not conserved."""
    N = 24
    C, h, u, v = _state_with_bump(oracle_lib, N)
    d0 = C.diagnostics(h, u, v)
    assert np.isnan(d0["Enstrophy"])
    assert d0["hmin"] < d0["hmax"] and d0["KE"] > 0 and d0["PE"] > 0
    assert d0["TE"] == d0["KE"] + d0["PE"]
    dt = 0.2 * (A_EARTH * np.pi / 2 / N) / np.sqrt(G_EARTH * H_MEAN)
    C.rk4(h, u, v, dt, 100)
    d1 = C.diagnostics(h, u, v)
    assert abs(d1["mass"] - d0["mass"]) / d0["mass"] < 1e-14
    assert abs(d1["TE"] - d0["TE"]) / d0["TE"] < 1e-8
