"""Narrow-stencil SBP Laplacian (sbp42) and the hyperdiffusion built on it: the generated
coefficient table, the oracle restatement, and the host basis vectors, on the CPU.

The table (cubedsphere_sbp_swm_amd/csrc/laplace_sbp42_narrow.h) is generated from the
reference's laplace_Ah_sbp42_narrow_mod.f90; it is pinned here by what the mathematics fixes
for a narrow-stencil SBP second derivative D2(b) = H^-1 (-M + boundary term): H D2 symmetric
(every m_rc, parsed separately, equals m_cr), negative semi-definite, zero row sums, exactness
of (b f')' on low-degree polynomials, and an edge-row defect on linear f of exactly the omitted
boundary term -b0 f'(0) / h_11 (the cube assembly averages the two panels' edge rows instead).
The oracle's use of the table is pinned by the spherical-harmonic eigenfunctions of the
Laplace-Beltrami operator (x y z: -12, x^2 - y^2: -6, z: -2 on the unit sphere): error ratios
per doubling of N above 6 away from panel edges (4th-order interior stencil, polluted by the
edges), above 2.8 on the closure rows (2nd order), above 1.8 on the shared edge points (the
edge-averaged closures, 1st order), and L(1) = 0 to rounding.
"""
import ctypes
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "..", "cubedsphere_sbp_swm_amd", "csrc", "laplace_sbp42_narrow.h")


def _table():
    src = open(TABLE).read()

    def arr(name, conv=float):
        m = re.search(name + r"(?:\[[^\]]*\])+\s*=\s*\{(.*?)\};", src, re.S)
        body = m.group(1).replace("{", "").replace("}", "")
        return [conv(v) for v in body.replace("\n", " ").split(",") if v.strip()]

    return dict(HN=arr("LN42_HNUM"), HD=arr("LN42_HDEN"), NC=arr("LN42_NCOLS", int),
                MS=np.array(arr("LN42_MSTART", int)).reshape(6, 9), TN=arr("LN42_TNUM"), TD=arr("LN42_TDEN"),
                TS=arr("LN42_TSIGN", int), TB=arr("LN42_TB", int), IS=arr("LN42_ISTART", int),
                IN=arr("LN42_INUM"), ID=arr("LN42_IDEN"), ISG=arr("LN42_ISIGN", int), IB=arr("LN42_IB", int))


def _d2_matrix(T, b):
    """Matrix of the 1-D operator the table defines on n = len(b) points, and the norm H."""
    n = len(b)
    A, H = np.zeros((n, n)), np.ones(n)
    for i in range(n):
        if i < 6 or i >= n - 6:
            left = i < 6
            r = i if left else n - 1 - i
            bb = [b[k] if left else b[n - 1 - k] for k in range(8)]
            H[i] = T["HN"][r] / T["HD"][r]
            for c in range(T["NC"][r]):
                m = sum(T["TS"][t] * T["TN"][t] / T["TD"][t] * bb[T["TB"][t]]
                        for t in range(T["MS"][r][c], T["MS"][r][c + 1]))
                A[i, c if left else n - 1 - c] = -m / H[i]
        else:
            for o in range(5):
                cf = sum(T["ISG"][t] * T["IN"][t] / T["ID"][t] * b[i + T["IB"][t]]
                         for t in range(T["IS"][o], T["IS"][o + 1]))
                A[i, i - 2 + o] = cf
    return A, H


def test_table_is_a_dissipative_sbp_operator():
    T = _table()
    rng = np.random.default_rng(0)
    for n in (12, 13, 30):
        b = 1.0 + rng.random(n)
        A, H = _d2_matrix(T, b)
        HA = H[:, None] * A
        assert np.abs(HA - HA.T).max() < 1e-14
        assert np.linalg.eigvalsh(0.5 * (HA + HA.T)).max() < 1e-13
        assert np.abs(A.sum(axis=1)).max() < 1e-13


def test_table_accuracy_on_polynomials():
    T = _table()
    n = 30
    x = np.arange(n, dtype=float)
    for bdeg in (0, 1):
        b = 1.0 + 0.01 * x ** bdeg if bdeg else np.ones(n)
        bp = 0.01 * bdeg * np.ones(n) if bdeg else np.zeros(n)
        A, H = _d2_matrix(T, b)
        for deg in range(0, 4 if bdeg == 0 else 3):
            f = x ** deg
            fp = deg * x ** (deg - 1) if deg else np.zeros(n)
            fpp = deg * (deg - 1) * x ** (deg - 2) if deg > 1 else np.zeros(n)
            err = A @ f - (bp * fp + b * fpp)
            # the edge rows omit the boundary term (-b0 f'(0) / h_11 at the left edge, mirrored)
            err[0] -= b[0] * fp[0] / H[0]
            err[-1] += b[-1] * fp[-1] / H[-1]
            assert np.abs(err).max() < 1e-10 * max(1.0, np.abs(f).max()), (bdeg, deg)
    # interior rows: 4th order, exact for quintic f when b is constant
    A, _ = _d2_matrix(T, np.ones(n))
    for deg in (4, 5):
        f = (x / n) ** deg
        fpp = deg * (deg - 1) * (x / n) ** (deg - 2) / n ** 2
        assert np.abs((A @ f - fpp)[6:-6]).max() < 1e-12


def _edge_distance(n):
    a = np.minimum(np.arange(n), n - 1 - np.arange(n))
    return np.minimum.outer(a, a)


def test_oracle_laplace_eigenfunctions(oracle_lib):
    errs = {}
    for N in (24, 48, 96):
        G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
        x, y, z = G.geometry("r")
        L1 = G.interior(G.laplace42(np.ones_like(x)))
        assert np.abs(L1).max() < 1e-14 * N ** 2 * 100
        d = _edge_distance(N + 1)
        for f, lam in ((x * y * z, -12.0), (x * x - y * y, -6.0), (z, -2.0)):
            e = np.abs(G.interior(G.laplace42(f) - lam * f))
            errs.setdefault(lam, []).append((e[:, d == 0].max(), e[:, (d >= 1) & (d <= 3)].max(), e[:, d >= 8].max()))
    for lam, rows in errs.items():
        (e0a, e1a, e8a), (e0b, e1b, e8b), (e0c, e1c, e8c) = rows
        print(f"lambda {lam}: edge {e0a:.2e} {e0b:.2e} {e0c:.2e}; rows 1-3 {e1a:.2e} {e1b:.2e} {e1c:.2e}; "
              f"interior {e8a:.2e} {e8b:.2e} {e8c:.2e}")
        assert e0b / e0c > 1.8 and e1b / e1c > 2.8 and e8b / e8c > 6.0
        assert e8c < 1e-6


def test_oracle_hyperdiffusion_of_eigenfunctions(oracle_lib):
    """tend = -(c a hx)^4 L(L f): for f = x y z, -(c hx)^4 * 144 f on the unit sphere (interior
    points); the velocity of a solid-body rotation V = e_z x r (Cartesian components -y, x, 0,
    all l = 1) gives (V.b1, V.b2) * -(c hx)^4 * 4."""
    N, c = 96, 0.03
    G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
    x, y, z = G.geometry("r")
    b1, b2 = G.geometry("b1"), G.geometry("b2")
    a1, a2 = G.geometry("a1"), G.geometry("a2")
    d = _edge_distance(N + 1)
    k4 = (c * (0.5 * np.pi / N)) ** 4
    f = x * y * z
    t = G.interior(G.hordiff_scalar42(c, f))
    ref = G.interior(-k4 * 144.0 * f)
    assert np.abs((t - ref)[:, d >= 8]).max() < 1e-4 * np.abs(ref).max()
    V = (-y, x, np.zeros_like(x))
    u = V[0] * a1[0] + V[1] * a1[1] + V[2] * a1[2]
    v = V[0] * a2[0] + V[1] * a2[1] + V[2] * a2[2]
    ut, vt = G.hordiff_vec42(c, u, v)
    ru = -k4 * 4.0 * (V[0] * b1[0] + V[1] * b1[1] + V[2] * b1[2])
    rv = -k4 * 4.0 * (V[0] * b2[0] + V[1] * b2[1] + V[2] * b2[2])
    for got, want in ((ut, ru), (vt, rv)):
        g_, w_ = G.interior(got), G.interior(want)
        assert np.abs((g_ - w_)[:, d >= 8]).max() < 1e-4 * np.abs(w_).max()


def test_oracle_diff_step_is_dissipative(oracle_lib):
    """One explicit hyperdiffusion step lowers the quadrature-weighted norm of a mean-free
    height field and of the velocity, and leaves a constant height unchanged."""
    N = 32
    G = oracle_lib.Grid(N, 42, 42, 6.371229e6, 0.0, 9.80616)
    x, y, z = G.geometry("r")
    a1, a2 = G.geometry("a1"), G.geometry("a2")
    A = G.quad_weights(42)
    w = G.interior(G.geometry("J")) * np.outer(A, A)[None]
    hp = 100.0 * (x * y * z + x * x - y * y + np.sin(3.0 * x) * z)
    V = (-y * z, x * z + np.cos(2.0 * y), np.zeros_like(x))
    u = 20.0 * (V[0] * a1[0] + V[1] * a1[1] + V[2] * a1[2])
    v = 20.0 * (V[0] * a2[0] + V[1] * a2[1] + V[2] * a2[2])
    h1, u1, v1 = hp.copy(), u.copy(), v.copy()
    G.diff_step42(h1, u1, v1, 600.0, 0.03, 0.03)
    norm = lambda *fs: float(sum(np.sum(w * G.interior(f) ** 2) for f in fs))
    assert norm(h1) < norm(hp)
    assert norm(u1, v1) < norm(u, v)
    c = np.full_like(x, 1.0e4)
    zu, zv = np.zeros_like(u), np.zeros_like(v)
    G.diff_step42(c, zu, zv, 600.0, 0.03, 0.03)
    assert np.abs(G.interior(c) - 1.0e4).max() < 1e-6


def test_basis_b_host_equals_oracle(oracle_lib):
    from cubedsphere_sbp_swm_amd import _capi
    N = 20
    n = N + 1
    G = oracle_lib.Grid(N, 42, 42, 1.0, 0.0, 1.0)
    pitch = n
    b1 = np.zeros(3 * 6 * n * pitch)
    b2 = np.zeros_like(b1)
    _capi.lib().swmh_basis_b(N, pitch, b1.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             b2.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    for host, name in ((b1, "b1"), (b2, "b2")):
        ref = G.geometry(name)
        for c in range(3):
            got = host[c * 6 * n * pitch:(c + 1) * 6 * n * pitch].reshape(6, n, pitch)
            assert np.array_equal(got, G.interior(ref[c]))
