"""Pin the oracle (CPU FP64 restatement) before trusting it as the GPU checker.

Every check here is independent of the oracle's own formulas: the reference's
printed error norms (tests/golden), closed-form known answers of the
reference's own tests (quadrature exactness, RK4 amplification polynomial), and
mathematical invariants (SBP summation-by-parts identity, polynomial
exactness, cube-edge connectivity, sync agreement of duplicated points).
"""
import os

import numpy as np
import pytest

from oracle import test_fields as tf

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "diffops_reference_results_ah.txt")


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
        e = np.floor(np.log10(abs(b))) + 1          # b = 0.ddddddd * 10^e
        ulp = 10.0 ** (e - 7)
        if abs(a - b) > 0.5 * ulp * 1.0001:
            return False
    return True


def _levels(G, fields):
    return max(G.maxabs(f) for f in fields), np.sqrt(sum(G.l2norm(f) ** 2 for f in fields))


def _div_errs(o, kind):
    # test_div, test_diff_ops/test_diffops_mod.f90:71-125 (N=32, nz=3 levels)
    G = o.Grid(32, grad_order=42, div_kind=kind)
    x, y, z = G.geometry("r")
    res, ds = [], []
    for k in (1, 2, 3):
        u, v = G.project("contravariant", *tf.solid_rotation(x, y, z, k))
        ds.append(G.div(G.embed(G.interior(u)), G.embed(G.interior(v))))
    res += _levels(G, ds)
    ds = []
    for k in (1, 2, 3):
        u, v = G.project("contravariant", *tf.cross_polar(x, y, z, k))
        d = G.div(G.embed(G.interior(u)), G.embed(G.interior(v)))
        ds.append(G.embed(G.interior(1.0 * d + (-1.0) * tf.cross_polar_div(x, y, z))))
    res += _levels(G, ds)
    return res


def _grad_errs(o, order):
    # test_grad, test_diffops_mod.f90:127-178
    G = o.Grid(32, grad_order=order, div_kind=42)
    x, y, z = G.geometry("r")
    ex, ey = [], []
    for k in (1, 2, 3):
        gx, gy = G.grad(G.embed(G.interior(tf.xyz_scalar(x, y, z, k))))
        tx, ty = G.project("covariant", *tf.xyz_grad(x, y, z, k))
        ex.append(G.embed(G.interior(1.0 * gx + (-1.0) * tx)))
        ey.append(G.embed(G.interior(1.0 * gy + (-1.0) * ty)))
    m1, l1 = _levels(G, ex)
    m2, l2 = _levels(G, ey)
    return [m1 + m2, l1 + l2]


def _curl_errs(o, kind):
    # test_curl, test_diffops_mod.f90:466-522
    G = o.Grid(32, grad_order=42, div_kind=kind)
    x, y, z = G.geometry("r")
    res, cs = [], []
    for k in (1, 2, 3):
        u, v = G.project("covariant", *tf.vsh_curl_free_10(x, y, z, k))
        cs.append(G.curl(G.embed(G.interior(u)), G.embed(G.interior(v))))
    res += _levels(G, cs)
    cs = []
    for k in (1, 2, 3):
        u, v = G.project("covariant", *tf.cross_polar(x, y, z, k))
        c = G.curl(G.embed(G.interior(u)), G.embed(G.interior(v)))
        cs.append(G.embed(G.interior(1.0 * c + (-1.0) * tf.cross_polar_zeta(x, y, z))))
    res += _levels(G, cs)
    return res


@pytest.mark.parametrize("name,kind", [("divergence_ah2", 2), ("divergence_ah42_sbp", 42),
                                       ("divergence_ah43_sbp", 43)])
def test_golden_divergence(oracle_lib, name, kind):
    assert _same_print(_div_errs(oracle_lib, kind), _golden()[name])


@pytest.mark.parametrize("name,order", [("gradient_ah21_sbp_ecs", 21), ("gradient_ah42_sbp_ecs", 42),
                                        ("gradient_ah43_sbp_ecs", 43)])
def test_golden_gradient(oracle_lib, name, order):
    assert _same_print(_grad_errs(oracle_lib, order), _golden()[name])


@pytest.mark.parametrize("name,kind", [("curl_divergence_ah42_sbp", 42), ("curl_divergence_ah43_sbp", 43)])
def test_golden_curl(oracle_lib, name, kind):
    assert _same_print(_curl_errs(oracle_lib, kind), _golden()[name])


def test_golden_co2contra_roundoff(oracle_lib):
    """co2contra_colocated on the Ah grid (test_diffops_mod.f90:231-278): a round-off level norm,
    matched to all 7 printed digits -> the metric Qi and the a/b basis vectors are bit-faithful."""
    G = oracle_lib.Grid(32)
    x, y, z = G.geometry("r")
    uc, vc = G.project("covariant", *tf.cross_polar(x, y, z))
    ut, vt = G.project("contravariant", *tf.cross_polar(x, y, z))
    Qi1, Qi2, Qi3 = G.geometry("Qi11"), G.geometry("Qi12"), G.geometry("Qi22")
    eu = G.embed(G.interior(Qi1 * uc + Qi2 * vc + (-1.0) * ut))
    ev = G.embed(G.interior(Qi2 * uc + Qi3 * vc + (-1.0) * vt))
    ours = [G.maxabs(eu) + G.maxabs(ev), np.sqrt(3.0) * (G.l2norm(eu) + G.l2norm(ev))]
    assert _same_print(ours, _golden()["co2contra_colocated Ah grid"])


def test_contra2co_roundoff_bound(oracle_lib):
    """contra2co (Q tensor, off the hot path): the reference prints 1.7e-15/1.8e-15; the
    restatement's integer-power rounding differs in the last bit, so only the round-off
    level is asserted."""
    G = oracle_lib.Grid(32)
    x, y, z = G.geometry("r")
    uc, vc = G.project("covariant", *tf.cross_polar(x, y, z))
    ut, vt = G.project("contravariant", *tf.cross_polar(x, y, z))
    Q1, Q2, Q3 = G.geometry("Q11"), G.geometry("Q12"), G.geometry("Q22")
    eu = G.embed(G.interior(Q1 * ut + Q2 * vt + (-1.0) * uc))
    ev = G.embed(G.interior(Q2 * ut + Q3 * vt + (-1.0) * vc))
    assert G.maxabs(eu) + G.maxabs(ev) < 3e-15


@pytest.mark.parametrize("order,pmax", [(21, 1), (42, 3), (63, 5)])
def test_quadrature_exactness(oracle_lib, order, pmax):
    """test_quadrature/test_quadrature_mod.f90:18-81: integral of a^p b^p / J over the cube
    (= 6((pi/4)^(p+1) - (-pi/4)^(p+1))^2/(p+1)^2) is exact to 1e-12 for p <= pmax."""
    G = oracle_lib.Grid(20)
    J = G.geometry("J")
    n, hw = G.n, G.hw
    pi = np.arccos(-1.0)
    al = -0.25 * pi + np.arange(-hw, n + hw) * ((0.25 * pi + 0.25 * pi) / 20)
    A, B = np.meshgrid(al, al)
    for p in range(pmax + 1):
        f = 1.0 / J * A[None] ** p * B[None] ** p
        exact = 6.0 * ((0.25 * pi) ** (p + 1) - (-0.25 * pi) ** (p + 1)) ** 2 / (p + 1) ** 2
        assert abs(G.quad_mass(order, f) - exact) < 1e-12, p


def test_rk4_amplification_iomega(oracle_lib):
    """test_time_steping/test_ts_mod.f90:7-73: one RK4 step on df/dt = w f equals
    1 + z + z^2/2 + z^3/6 + z^4/24 (z = w dt).  The reference asserts 1e-12 absolute; here
    |f| reaches 2.4e3, so the bound is taken relative to |f| (a wrong RK weight gives O(1))."""
    day = 24.0 * 3600.0
    dt = 6.0 * 3600.0
    i = np.arange(1, 11)
    w = 1.0 / (day * i) + 1j * 2.0 * np.pi * i / day
    z = dt * w
    f = oracle_lib.rk4_iomega(w, np.ones(10), dt)
    true = 1 + z + 0.5 * z ** 2 + z ** 3 / 6 + z ** 4 / 24
    assert np.max(np.abs(f - true) / np.abs(true)) < 1e-14


def _dmatrix(o, order, N=24):
    G = o.Grid(N, grad_order=order if order != 2 else 21)
    n = G.n
    f = G.zeros()
    I = np.eye(n)
    G.interior(f)[0] = I  # row j holds e_j along i, so output row j is column j of D
    D = G.interior(G.sbp(order, "x", f))[0].T.copy()
    return D, G


@pytest.mark.parametrize("order,qorder", [(21, 21), (42, 42), (63, 63)])
def test_sbp_summation_by_parts(oracle_lib, order, qorder):
    """H D + (H D)^T = diag(-1, 0, ..., 0, 1) with H = diag of the SBP quadrature weights."""
    D, G = _dmatrix(oracle_lib, order)
    H = np.diag(G.quad_weights(qorder))
    B = np.zeros_like(D)
    B[0, 0], B[-1, -1] = -1.0, 1.0
    assert np.max(np.abs(H @ D + (H @ D).T - B)) < 2e-14


@pytest.mark.parametrize("order,p_in,p_bd", [(21, 2, 1), (42, 4, 2), (43, 4, 3), (63, 6, 3)])
def test_sbp_polynomial_exactness(oracle_lib, order, p_in, p_bd):
    D, G = _dmatrix(oracle_lib, order)
    n = D.shape[0]
    x = np.arange(n, dtype=float) / n
    ne = {21: 1, 42: 4, 43: 4, 63: 6}[order]
    for k in range(p_in + 1):
        d = D @ (x ** k) * n
        ex = k * x ** (k - 1) if k > 0 else 0 * x
        err = np.abs(d - ex)
        assert err[ne:n - ne].max() < 1e-9, (order, k)
        if k <= p_bd:
            assert err.max() < 1e-9, (order, k)


def test_halo_map_connectivity(oracle_lib):
    """Depth-0 halo points hold the neighbour's copy of the shared edge point (same position);
    depth-d halo points hold the neighbour's d-th interior line (d grid steps away)."""
    G = oracle_lib.Grid(16)
    src, vec = G.halo_map()
    r = np.stack(G.geometry("r"))  # (3, 6, ld, ld)
    rf = r.reshape(3, -1)
    n, hw = G.n, G.hw
    hx = 0.5 * np.pi / 16
    count = 0
    for p in range(6):
        for jj in range(G.ld):
            for ii in range(G.ld):
                s = src[p, jj, ii]
                if s < 0:
                    continue
                count += 1
                i, j = ii - hw + 1, jj - hw + 1
                # nearest panel-edge point of this halo point
                ie, je = min(max(i, 1), n), min(max(j, 1), n)
                depth = max(1 - i, i - n, 1 - j, j - n, 0) - 1
                d = np.linalg.norm(rf[:, s] - r[:, p, je - 1 + hw, ie - 1 + hw])
                if depth == 0:
                    assert d < 1e-15
                else:
                    assert 0.3 * depth * hx < d < 1.2 * depth * hx
    assert count == 6 * 4 * n * hw


def test_scalar_sync_agreement(oracle_lib):
    """After the Ah scalar sync duplicated edge points agree bitwise (corners: operand order
    differs per panel, so to round-off)."""
    G = oracle_lib.Grid(12)
    rng = np.random.default_rng(0)
    f = G.embed(rng.standard_normal((6, G.n, G.n)))
    div_like = f.copy()
    oracle_lib.lib().ora_scalar_sync(G.h, oracle_lib._p(div_like))
    src, _ = G.halo_map()
    flat = div_like.reshape(-1)
    n, hw = G.n, G.hw
    for p in range(6):
        for j in range(1, n + 1):
            for i in (0, n + 1):
                s = src[p, j - 1 + hw, i - 1 + hw]
                own = div_like[p, j - 1 + hw, (1 if i == 0 else n) - 1 + hw]
                if j in (1, n):
                    assert abs(flat[s] - own) < 1e-15 * 4
                else:
                    assert flat[s] == own
