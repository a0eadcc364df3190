"""Pins of the oracle's composite shallow-water RHS, its RK4 trajectory and its diagnostics.

The single operators of the oracle are pinned by the reference's golden error norms
(tests/test_oracle.py).  What those leave open -- the composition of operator_swm_t%apply
(operator_swm_mod.f90:61-113), the RK4 combination of rk4_t%step (rk4_mod.f90:20-42) and
get_diagnostics (operator_swm_mod.f90:115-260) -- is pinned here by what the mathematics
fixes, independently of the oracle's own formulas:

* The reference's test case 2 (ts2_mod.f90:36-205): solid-body rotation about the Earth's
  axis (test_fields_mod.f90:553-580) with the height of geostrophic balance
  h = h_mean - (a*Omega*u0 + u0^2/2)/g * sin^2(phi) (test_fields_mod.f90:826-859) is an exact
  steady state, so the discrete RHS must vanish at the order of the SBP operators as the grid
  is refined.  A wrong Coriolis sign, a dropped 1/2 of the kinetic energy, a missing g*h term
  or a wrong metric factor leaves an O(1) residual.
* Classical RK4 has temporal order 4: on a fixed grid, halving dt cuts the error of a fixed
  time interval 16-fold.  Swapping the weights dt/6 and dt/3 leaves a second-order scheme
  (error ratio 4).
* Conservation: the SBP divergence with the reference's edge averaging conserves mass to
  round-off (sum of A_i A_j J dh = 0), and the discrete circulation over the closed sphere
  vanishes (sum of A_i A_j J zeta = 0).
* The diagnostics of the TS2 state have closed forms (integrals of polynomials in sin(phi)
  over the unit sphere, the enstrophy by Gauss-Legendre quadrature) that the discrete
  quadrature must approach with N.
"""
import numpy as np
import pytest

A_EARTH = 0.6371229e7                      # const_mod.f90:7
G_EARTH = 9.80616                          # const_mod.f90:6
T_SID = 23 * 60 * 60 + 56 * 60 + 4         # const_mod.f90:8
OMEGA = 2.0 * np.pi / T_SID                # const_mod.f90:9
H_MEAN = 29400 / G_EARTH                   # config_ts2_mod.f90:14
U0 = np.pi * A_EARTH / 6.0 / T_SID         # config_ts2_mod.f90:16
C_TS2 = (A_EARTH * OMEGA * U0 + 0.5 * U0 ** 2) / G_EARTH

# (gradient order, divergence kind) of the Ah operator families, with the order of the
# boundary closure of the SBP first derivative (1 for d21, 2 for d42, 3 for d43 / d63)
FAMILIES = {21: (21, 2, 1), 42: (42, 42, 2), 43: (43, 43, 3), 63: (63, 63, 3)}


def _unit_r(G):
    x, y, z = G.geometry("r")
    r = np.sqrt(x * x + y * y + z * z)
    return x / r, y / r, z / r


def ts2_state(G):
    """TS2 height and covariant solid-rotation velocity (axis z) on the mesh points."""
    xn, yn, zn = _unit_r(G)
    h = H_MEAN - C_TS2 * zn ** 2
    u, v = G.project("covariant", -U0 * yn, U0 * xn, 0.0 * zn)
    return tuple(G.embed(G.interior(f)) for f in (h, u, v))


def _bump(G, amp=120.0, width=0.35):
    """Gaussian hill in h centred off the cube edges and the poles."""
    xn, yn, zn = _unit_r(G)
    c = (np.cos(0.3) * np.cos(0.7), np.cos(0.3) * np.sin(0.7), np.sin(0.3))
    d = np.arccos(np.clip(xn * c[0] + yn * c[1] + zn * c[2], -1.0, 1.0))
    return G.embed(G.interior(amp * np.exp(-(d / width) ** 2)))


def _grid(o, N, fam):
    go, dk, _ = FAMILIES[fam]
    return o.Grid(N, go, dk, A_EARTH, OMEGA, G_EARTH)


def _ts2_residuals(o, N, fam):
    G = _grid(o, N, fam)
    h, u, v = ts2_state(G)
    dh, du, dv = G.rhs(h.copy(), u.copy(), v.copy())
    # scales of the balanced terms: the gradient of g*h, and h*|V|/a for the mass flux
    gx, gy = G.grad(G.embed(G.interior(G_EARTH * h)))
    s_uv2 = np.sqrt(G.l2norm(gx) ** 2 + G.l2norm(gy) ** 2)
    s_uvi = max(G.maxabs(gx), G.maxabs(gy))
    s_h = G.l2norm(h) * U0 / A_EARTH
    l2_uv = np.sqrt(G.l2norm(du) ** 2 + G.l2norm(dv) ** 2) / s_uv2
    li_uv = max(G.maxabs(du), G.maxabs(dv)) / s_uvi
    l2_h = G.l2norm(dh) / s_h
    return l2_uv, li_uv, l2_h


@pytest.mark.parametrize("fam", [21, 42, 43, 63])
def test_ts2_steady_state_rhs_converges(oracle_lib, fam):
    """RHS(TS2) -> 0 at the SBP order: L2 at (closure order + 1/2), max norm at the closure
    order (measured: sbp21 1.5 / 1.0, sbp42 2.4 / 2.0, sbp43 and sbp63 3.4 / 2.7-2.8)."""
    q = FAMILIES[fam][2]
    r = [_ts2_residuals(oracle_lib, N, fam) for N in (16, 32, 64)]
    for k in range(2):
        rate_l2_uv = np.log2(r[k][0] / r[k + 1][0])
        rate_li_uv = np.log2(r[k][1] / r[k + 1][1])
        rate_l2_h = np.log2(r[k][2] / r[k + 1][2])
        assert rate_l2_uv > q + 0.5 - 0.35, (fam, k, rate_l2_uv)
        assert rate_l2_h > q + 0.5 - 0.35, (fam, k, rate_l2_h)
        assert rate_li_uv > q - 0.4, (fam, k, rate_li_uv)
    # and the residual is small in absolute terms at N = 64
    assert r[2][0] < 2e-3 and r[2][2] < 2e-3, r[2]


@pytest.mark.parametrize("fam", [21, 42, 63])
def test_rhs_conserves_mass_and_circulation(oracle_lib, fam):
    """sum A_i A_j J dh = 0 and sum A_i A_j J zeta = 0 to round-off, on an unbalanced state,
    with the diagonal SBP norm A of the family as quadrature (d43's closure rows have no
    diagonal norm of their own, so it has no such identity)."""
    G = _grid(oracle_lib, 24, fam)
    h, u, v = ts2_state(G)
    h = h + _bump(G)
    xn, yn, zn = _unit_r(G)
    # an extra velocity that is neither zonal nor divergence-free
    du_, dv_ = G.project("covariant", 5.0 * zn * xn, 3.0 * yn, -2.0 * xn * yn)
    u = u + G.embed(G.interior(du_))
    v = v + G.embed(G.interior(dv_))
    qo = fam
    dh, _, _ = G.rhs(h.copy(), u.copy(), v.copy())
    tot = G.quad_mass(qo, np.abs(dh))
    assert tot > 0
    assert abs(G.quad_mass(qo, dh)) / tot < 1e-13
    zeta = G.curl(u.copy(), v.copy())
    tot = G.quad_mass(qo, np.abs(zeta))
    assert abs(G.quad_mass(qo, zeta)) / tot < 1e-13


def test_rk4_conserves_mass(oracle_lib):
    G = _grid(oracle_lib, 16, 42)
    h, u, v = ts2_state(G)
    h = h + _bump(G)
    m0 = G.quad_mass(42, h)
    dt = 0.5 * (A_EARTH * np.pi / 2 / 16) / np.sqrt(G_EARTH * H_MEAN)
    G.rk4(h, u, v, dt, 10)
    assert abs(G.quad_mass(42, h) - m0) / m0 < 1e-14


@pytest.mark.parametrize("courant", [0.5, 0.25])
def test_rk4_temporal_order_is_four(oracle_lib, courant):
    """Fixed grid (C16 sbp42), fixed interval of 8 steps at dt0: errors against dt0/16 fall
    16-fold per halving of dt (measured 15.8-17.2)."""
    N = 16
    G = _grid(oracle_lib, N, 42)
    h, u, v = ts2_state(G)
    h = h + _bump(G)
    dt0 = courant * (A_EARTH * np.pi / 2 / N) / np.sqrt(G_EARTH * H_MEAN)
    res = {}
    for k in (1, 2, 4, 16):
        s = [h.copy(), u.copy(), v.copy()]
        G.rk4(*s, dt0 / k, 8 * k)
        res[k] = [G.interior(f) for f in s]
    err = {k: max(np.max(np.abs(res[k][c] - res[16][c])) / np.max(np.abs(res[16][c])) for c in range(3))
           for k in (1, 2, 4)}
    assert 12.0 < err[1] / err[2] < 20.0, err
    assert 12.0 < err[2] / err[4] < 20.0, err


def _ts2_closed_forms():
    """hmin, hmax, mass, TE, KE, PE, enstrophy of the TS2 state on the unit sphere (the
    reference's quadrature carries no a^2), as integrals over s = sin(phi) in [-1, 1]."""
    s, w = np.polynomial.legendre.leggauss(200)
    hh = H_MEAN - C_TS2 * s ** 2
    mass = 2 * np.pi * np.sum(w * hh)
    ke = 2 * np.pi * np.sum(w * 0.5 * hh * U0 ** 2 * (1 - s ** 2))
    pe = 2 * np.pi * np.sum(w * 0.5 * G_EARTH * hh ** 2)
    # absolute vorticity of solid rotation: zeta + f = 2 (u0/a + Omega) sin(phi)
    ens = 2 * np.pi * np.sum(w * 0.5 * (2 * (U0 / A_EARTH + OMEGA) * s) ** 2 / hh)
    assert abs(mass - (4 * np.pi * H_MEAN - C_TS2 * 4 * np.pi / 3)) < 1e-9 * mass
    return np.array([H_MEAN - C_TS2, H_MEAN, mass, ke + pe, ke, pe, ens])


def test_ts2_diagnostics_match_closed_forms(oracle_lib):
    ex = _ts2_closed_forms()
    errs = []
    for N in (16, 32, 64):
        G = _grid(oracle_lib, N, 42)
        h, u, v = ts2_state(G)
        d = G.diagnostics(42, h, u, v)
        # the poles and the equator are mesh points for even N
        assert abs(d[0] - ex[0]) < 1e-12 * ex[0] and abs(d[1] - ex[1]) < 1e-12 * ex[1]
        errs.append(np.abs(d[2:] - ex[2:]) / np.abs(ex[2:]))
    assert np.all(errs[2] < 1e-7), errs[2]
    assert np.all(errs[0] / errs[1] > 8.0) and np.all(errs[1] / errs[2] > 8.0), errs
    # energy is the sum of its parts
    d = G.diagnostics(42, h, u, v)
    assert d[3] == d[4] + d[5]
