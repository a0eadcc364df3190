"""The oracle's C-grid ECS vector halo ("ecs_C_vec", "ecs_C_vec_covariant") and its consumers
gradient_c2_ecs and massflux_c_up4, pinned by the reference's own checks and by what the
mathematics fixes.

* The reference's halo test (test_ecs_cvec_halo, src/test/test_halo/test_ecs_halo_c_mod.f90:13-95):
  N = 64, halo width 3, solid rotation about each axis, max error on the strips beyond the edges
  below 3e-7 for both component kinds.  The true values are the components at the virtual points
  of each panel's extended gnomonic coordinates, computed here from first principles (analytic
  tangent vectors of the prototype face, the panel's frame fitted on the corner points, the dual
  basis from the Gram system); no table of the procedure is restated.
* 4th-order convergence of every strip (>10x per doubling) on a second flow (cross-polar), so
  a wrong transform sign, extrapolation stencil or stencil offset fails.
* gradient_c2_ecs: the reference's golden error line at N = 32 (src/test/test_diff_ops/
  reference_results.txt:37-38), and 2nd-order convergence of the halo rows it also computes.
* massflux_c_up4: exact on cubics in the interior (fixes the 5, 15, -5, 1 / 16 stencil), the
  upwind stencil support (faces i0..i0+3 for u > 0, i0-2..i0+1 for u < 0 see cell i0), a constant
  field times u, and convergence near the panel edges where the ECS_O halo of width 3 enters.
"""
import numpy as np
import pytest

from oracle import test_fields as tf

HW_MAX = 5


def _golden():
    import os
    out = {}
    for line in open(os.path.join(os.path.dirname(__file__), "golden", "diffops_reference_results_c.txt")):
        if line.startswith("#") or "|" not in line:
            continue
        k, v = line.split("|")
        out[k.strip()] = [float(x) for x in v.split()]
    return out


def _same_print(ours, ref):
    for a, b in zip(ours, ref):
        e = np.floor(np.log10(abs(b))) + 1
        if abs(a - b) > 0.5 * 10.0 ** (e - 7) * 1.0001:
            return False
    return True


def _frames(C):
    """r = r_proto @ M per panel, fitted on the corner (q) points of the C grid."""
    N = C.N
    r = np.stack([C.interior("q", a) for a in C.geometry("q", "r")], -1)
    a = -0.25 * np.pi + np.arange(N + 1) * (0.5 * np.pi / N)
    A, B = np.meshgrid(a, a)
    rp = np.stack([np.tan(A), np.tan(B), np.ones_like(A)], -1)
    rp /= np.linalg.norm(rp, axis=-1, keepdims=True)
    Ms = []
    for p in range(6):
        M, *_ = np.linalg.lstsq(rp.reshape(-1, 3), r[p].reshape(-1, 3), rcond=None)
        assert np.abs(rp.reshape(-1, 3) @ M - r[p].reshape(-1, 3)).max() < 1e-13
        Ms.append(M)
    return Ms


def _basis(A, B, M):
    """position, covariant (a1, a2) and contravariant (b1, b2) basis at gnomonic (A, B)"""
    X, Y = np.tan(A), np.tan(B)
    rho = np.sqrt(1 + X * X + Y * Y)
    P = np.stack([X, Y, np.ones_like(X)], -1)
    ex, ey = np.zeros_like(P), np.zeros_like(P)
    ex[..., 0], ey[..., 1] = 1, 1
    a1 = (ex / rho[..., None] - P * (X / rho ** 3)[..., None]) / np.cos(A)[..., None] ** 2
    a2 = (ey / rho[..., None] - P * (Y / rho ** 3)[..., None]) / np.cos(B)[..., None] ** 2
    r, a1, a2 = (P / rho[..., None]) @ M, a1 @ M, a2 @ M
    g11, g12, g22 = (a1 * a1).sum(-1), (a1 * a2).sum(-1), (a2 * a2).sum(-1)
    det = g11 * g22 - g12 * g12
    b1 = (g22[..., None] * a1 - g12[..., None] * a2) / det[..., None]
    b2 = (-g12[..., None] * a1 + g11[..., None] * a2) / det[..., None]
    return r, a1, a2, b1, b2


def _coords(N, t, hw):
    hx = 0.5 * np.pi / N
    nx, ny = (N + 1, N) if t == "x" else (N, N + 1)
    si, sj = (0.0, 0.5) if t == "x" else (0.5, 0.0)
    ii, jj = np.arange(1 - hw, nx + hw + 1), np.arange(1 - hw, ny + hw + 1)
    return np.meshgrid(-0.25 * np.pi + (ii - 1 + si) * hx, -0.25 * np.pi + (jj - 1 + sj) * hx)


def _truth(C, Ms, gen, level, hw, covariant):
    out = []
    for t in ("x", "y"):
        A, B = _coords(C.N, t, hw)
        res = []
        for p in range(6):
            r, a1, a2, b1, b2 = _basis(A, B, Ms[p])
            V = np.stack(gen(r[..., 0], r[..., 1], r[..., 2], level), -1)
            e = (a1 if t == "x" else a2) if covariant else (b1 if t == "x" else b2)
            res.append((V * e).sum(-1))
        out.append(np.stack(res))
    return out


def _inputs(C, gen, level, covariant):
    kind = "covariant" if covariant else "contravariant"

    def one(t, comp):
        x, y, z = C.geometry(t, "r")
        return C.embed(t, C.interior(t, C.project(t, kind, *gen(x, y, z, level), comp)))
    return one("x", 0), one("y", 1)


def _interior_masks(N, hw):
    mu = np.zeros((N + 2 * hw, N + 1 + 2 * hw), bool)
    mu[hw:hw + N, hw:hw + N + 1] = True
    mv = np.zeros((N + 1 + 2 * hw, N + 2 * hw), bool)
    mv[hw:hw + N + 1, hw:hw + N] = True
    return mu, mv


def _strip_errors(C, gen, level, hw, covariant, Ms=None):
    Ms = Ms or _frames(C)
    u, v = _inputs(C, gen, level, covariant)
    ou, ov = C.ecs_halo_vec(u, v, hw, covariant)
    tu, tv = _truth(C, Ms, gen, level, hw, covariant)
    mu, mv = _interior_masks(C.N, hw)
    su, sv = ~mu & ~np.isnan(ou[0]), ~mv & ~np.isnan(ov[0])
    return (np.abs(ou - tu)[:, su].max(), np.abs(ov - tv)[:, sv].max(),
            max(np.abs(ou - tu)[:, mu].max(), np.abs(ov - tv)[:, mv].max()))


@pytest.mark.parametrize("covariant", [False, True])
def test_reference_cvec_halo_tolerance(oracle_lib, covariant):
    # test_ecs_cvec_halo: nh = 64, halo_width = 3, solid rotation on 3 levels, tolerance 3e-7
    C = oracle_lib.CGrid(64)
    Ms = _frames(C)
    worst = max(max(_strip_errors(C, tf.solid_rotation, lev, 3, covariant, Ms)[:2]) for lev in (1, 2, 3))
    print("max strip error", worst)
    assert worst < 3.0e-7


@pytest.mark.parametrize("covariant", [False, True])
@pytest.mark.parametrize("hw", [2, 3, HW_MAX])
def test_cvec_halo_converges_4th_order(oracle_lib, covariant, hw):
    errs = []
    for N in (16, 32, 64):
        C = oracle_lib.CGrid(N)
        eu, ev, ei = _strip_errors(C, tf.cross_polar, 1, hw, covariant)
        errs.append(max(eu, ev))
        assert ei < 1e-13            # interior (and the averaged shared faces) untouched up to rounding
    print(hw, covariant, errs)
    assert errs[0] / errs[1] > 10.0 and errs[1] / errs[2] > 12.0


def test_cvec_halo_width_1(oracle_lib):
    """Halo width 1: the covariant procedure converges at 4th order; the contravariant one shifts
    hw + 1 = 2 rows of the normal component but its tangential transform reads 3 rows beyond the
    edge (ecs_halo_vec_c_mod.f90:144-147,411-418), the last one unshifted, so the tangential strip
    converges at first order while the normal strip keeps 4th order."""
    res = {}
    for cov in (False, True):
        errs = []
        for N in (16, 32, 64):
            C = oracle_lib.CGrid(N)
            Ms = _frames(C)
            u, v = _inputs(C, tf.cross_polar, 1, cov)
            ou, ov = C.ecs_halo_vec(u, v, 1, cov)
            tu, tv = _truth(C, Ms, tf.cross_polar, 1, 1, cov)
            eu, ev = np.abs(ou - tu), np.abs(ov - tv)
            normal = max(eu[:, 1:-1, [0, -1]].max(), ev[:, [0, -1], 1:-1].max())
            tangential = max(eu[:, [0, -1], 1:-1].max(), ev[:, 1:-1, [0, -1]].max())
            errs.append((normal, tangential))
        res[cov] = errs
    print(res)
    for cov in (False, True):
        (n0, t0), (n1, t1), (n2, t2) = res[cov]
        assert n0 / n1 > 10.0 and n1 / n2 > 12.0
        if cov:
            assert t0 / t1 > 10.0 and t1 / t2 > 12.0
        else:
            assert 1.6 < t0 / t1 < 2.5 and 1.6 < t1 / t2 < 2.5


def test_cvec_halo_defines_the_strips_only(oracle_lib):
    N, hw = 16, 2
    C = oracle_lib.CGrid(N)
    rng = np.random.default_rng(5)
    u = C.embed("x", rng.standard_normal((6, N, N + 1)))
    v = C.embed("y", rng.standard_normal((6, N + 1, N)))
    for cov in (False, True):
        ou, ov = C.ecs_halo_vec(u, v, hw, cov)
        mu, mv = _interior_masks(N, hw)
        for o, m in ((ou, mu), (ov, mv)):
            rows, cols = np.where(np.isnan(o[0]))
            # NaN exactly in the four hw x hw corner squares
            assert np.isnan(o).sum() == 6 * 4 * hw * hw
            H, W = m.shape
            assert all((r < hw or r >= H - hw) and (c < hw or c >= W - hw) for r, c in zip(rows, cols))
        # the covariant procedure leaves the interior alone; the contravariant one averages the
        # normal component's shared faces with the neighbour's copies
        if cov:
            assert np.array_equal(ou[:, hw:hw + N, hw:hw + N + 1], C.interior("x", u))
            assert np.array_equal(ov[:, hw:hw + N + 1, hw:hw + N], C.interior("y", v))
        else:
            inner_u = ou[:, hw:hw + N, hw + 1:hw + N]
            assert np.array_equal(inner_u, C.interior("x", u)[:, :, 1:N])
    with pytest.raises(ValueError):
        C.ecs_halo_vec(u, v, HW_MAX + 1)


def test_cvec_tables_are_cubic_lagrange(oracle_lib):
    N, hw = 24, 4
    C = oracle_lib.CGrid(N)
    for cov in (False, True):
        t = C.ecs_vec_tables(hw, cov)
        # positions 1..N+1 (tangential, cross-component), 1..N (normal)
        for w, i, nx, n in ((t["w0"], t["i0"], N + 1, (N + 1) * hw), (t["w1"], t["i1"], N, N * hw),
                            (t["w2"], t["i2"], N, (N + 1) * hw)):
            w = w.reshape(-1, 4)
            assert np.allclose(w[:n].sum(1), 1.0, atol=1e-14)
            assert i[:n].min() >= 2 and i[:n].max() <= nx - 2
        # the tables are mirror-symmetric along the edge (the prototype face is)
        ct = t["ct"].reshape(N + 1, hw) if cov else None
        if cov:
            assert np.allclose(ct, ct[::-1], rtol=1e-13, atol=0)


def test_golden_gradient_c2_ecs(oracle_lib):
    # test_grad (test_diffops_mod.f90:127-178) with gradient_c2_ecs on the C grid, N = 32
    C = oracle_lib.CGrid(32)
    xo, yo, zo = C.geometry("o", "r")
    ex, ey = [], []
    for k in (1, 2, 3):
        gx, gy = C.grad_c2_ecs(C.embed("o", C.interior("o", tf.xyz_scalar(xo, yo, zo, k))))
        tx, ty = [C.embed(t, C.interior(t, C.project(t, "covariant", *tf.xyz_grad(*C.geometry(t, "r"), k), c)))
                  for t, c in (("x", 0), ("y", 1))]
        ex.append(C.embed("x", C.interior("x", 1.0 * gx + (-1.0) * tx)))
        ey.append(C.embed("y", C.interior("y", 1.0 * gy + (-1.0) * ty)))
    m = max(C.maxabs("x", f) for f in ex) + max(C.maxabs("y", f) for f in ey)
    l2 = np.sqrt(sum(C.l2norm("x", f) ** 2 for f in ex)) + np.sqrt(sum(C.l2norm("y", f) ** 2 for f in ey))
    assert _same_print([m, l2], _golden()["gradient_c2_ecs"]), (m, l2)


def test_grad_c2_ecs_halo_rows_converge(oracle_lib):
    """gx on the rows j = 0, N+1 and gy on the columns i = 0, N+1 approximate the covariant
    gradient at those virtual points of the extended panel (2nd order)."""
    errs = []
    for N in (16, 32, 64):
        C = oracle_lib.CGrid(N)
        Ms = _frames(C)
        xo, yo, zo = C.geometry("o", "r")
        f = C.embed("o", C.interior("o", np.sin(2 * xo) * yo + zo * zo))
        gx, gy = C.grad_c2_ecs(f)
        e = 0.0
        for t, g in (("x", gx), ("y", gy)):
            A, B = _coords(N, t, 1)
            for p in range(6):
                r, a1, a2, _, _ = _basis(A, B, Ms[p])
                x, y, z = r[..., 0], r[..., 1], r[..., 2]
                grad = np.stack([2 * np.cos(2 * x) * y, np.sin(2 * x), 2 * z], -1)
                exact = (grad * (a1 if t == "x" else a2)).sum(-1)
                G = g[p, C.hw - 1:C.hw + exact.shape[0] - 1, C.hw - 1:C.hw + exact.shape[1] - 1]
                if t == "x":   # rows j = 0 and N+1, i = 1..N+1
                    e = max(e, np.abs(G[[0, -1], 1:-1] - exact[[0, -1], 1:-1]).max())
                else:          # columns i = 0 and N+1, j = 1..N+1
                    e = max(e, np.abs(G[1:-1, [0, -1]] - exact[1:-1, [0, -1]]).max())
        errs.append(e)
    print(errs)
    assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5


def test_massflux_up4_constant_and_cubic(oracle_lib):
    N = 24
    C = oracle_lib.CGrid(N)
    rng = np.random.default_rng(3)
    u = C.embed("x", rng.standard_normal((6, N, N + 1)))
    v = C.embed("y", rng.standard_normal((6, N + 1, N)))
    fx, fy = C.massflux_up4(C.embed("o", np.full((6, N, N), 2.5)), u, v)
    assert np.abs(C.interior("x", fx) - 2.5 * C.interior("x", u)).max() < 1e-14 * 2.5 * np.abs(u).max() * 8
    assert np.abs(C.interior("y", fy) - 2.5 * C.interior("y", v)).max() < 1e-14 * 2.5 * np.abs(v).max() * 8
    # a cubic in the cell index along x (and, separately, along y): every interior face whose
    # stencils stay inside the panel gets u * P(face) exactly, for either flow direction
    P = lambda s: 0.3 - 1.1 * s + 0.07 * s ** 2 + 0.002 * s ** 3
    i = np.arange(1, N + 1)
    fxc = np.broadcast_to(P(i.astype(float))[None, None, :], (6, N, N)).copy()
    fx, _ = C.massflux_up4(C.embed("o", fxc), u, v)
    faces = np.arange(4, N - 1)                      # i-3 >= 1 and i+2 <= N
    want = C.interior("x", u)[:, :, faces - 1] * P(faces - 0.5)
    assert np.abs(C.interior("x", fx)[:, :, faces - 1] - want).max() < 1e-12
    fyc = np.broadcast_to(P(i.astype(float))[None, :, None], (6, N, N)).copy()
    _, fy = C.massflux_up4(C.embed("o", fyc), u, v)
    want = C.interior("y", v)[:, faces - 1, :] * P(faces - 0.5)[None, :, None]
    assert np.abs(C.interior("y", fy)[:, faces - 1, :] - want).max() < 1e-12


def test_massflux_up4_upwind_support(oracle_lib):
    N, i0, j0 = 20, 9, 11
    C = oracle_lib.CGrid(N)
    rng = np.random.default_rng(8)
    f = C.embed("o", rng.standard_normal((6, N, N)))
    g = f.copy()
    g[2, C.hw + j0 - 1, C.hw + i0 - 1] += 1.0
    for sgn, want in ((1.0, {i0, i0 + 1, i0 + 2, i0 + 3}), (-1.0, {i0 - 2, i0 - 1, i0, i0 + 1})):
        u = C.embed("x", np.full((6, N, N + 1), sgn))
        v = C.embed("y", np.full((6, N + 1, N), sgn))
        a, _ = C.massflux_up4(f, u, v)
        b, _ = C.massflux_up4(g, u, v)
        d = C.interior("x", b - a)
        rows, cols = np.nonzero(d[2])
        assert set(rows) == {j0 - 1} and set(cols + 1) == want
        assert np.count_nonzero(d[[0, 1, 3, 4, 5]]) == 0
        _, a = C.massflux_up4(f, u, v)
        _, b = C.massflux_up4(g, u, v)
        d = C.interior("y", b - a)
        rows, cols = np.nonzero(d[2])
        assert set(cols) == {i0 - 1} and set(rows + 1) == {j0 - i0 + k for k in want}


def test_massflux_up4_converges_at_the_edges(oracle_lib):
    errs = []
    for N in (16, 32, 64):
        C = oracle_lib.CGrid(N)
        xo, yo, zo = C.geometry("o", "r")
        fun = lambda x, y, z: 1.0 + 0.3 * np.sin(2 * x) * y + 0.2 * z * z
        f = C.embed("o", C.interior("o", fun(xo, yo, zo)))
        xs, ys, zs = C.geometry("x", "r")
        xt, yt, zt = C.geometry("y", "r")
        u = C.embed("x", C.interior("x", np.cos(ys + zs) - 0.2))
        v = C.embed("y", C.interior("y", np.sin(xt - zt) + 0.1))
        fx, fy = C.massflux_up4(f, u, v)
        ex = np.abs(C.interior("x", fx) - C.interior("x", u * fun(xs, ys, zs)))
        ey = np.abs(C.interior("y", fy) - C.interior("y", v * fun(xt, yt, zt)))
        # the faces next to the panel edges read the ECS_O halo
        errs.append(max(ex[:, :, :3].max(), ex[:, :, -3:].max(), ey[:, :3, :].max(), ey[:, -3:, :].max()))
    print(errs)
    assert errs[0] / errs[1] > 7.0 and errs[1] / errs[2] > 7.0
