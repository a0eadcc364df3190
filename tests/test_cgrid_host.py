"""CPU checks of the C-grid set-up of the device library against the oracle (no device calls).

* The metric terms and the Coriolis parameter at o, x, y and q points, computed by the host
  code that fills the device arrays, are bitwise those of the oracle.
* The cube-edge gather lists (tangential vector halo to depth 3, face halo of the normal
  component, O-point halo to depth 3) reproduce the oracle's C-grid exchanges point for point,
  signs included, on fields of distinct random values.
* The neighbour copies listed for the q-point sync are the points the Ah exchange of the same
  N puts into halo index 0 / n+1 (the neighbour's copy of a shared edge point).
"""
import ctypes

import numpy as np
import pytest

N = 12
H = 4


def _lib():
    from cubedsphere_sbp_swm_amd import _capi
    return _capi.lib()


def _layout():
    rows = N + 1 + 2 * H
    pitch = (rows + 7) // 8 * 8
    return rows, pitch, 6 * rows * pitch


def _to_oracle(a, rows, pitch):
    return np.ascontiguousarray(a.reshape(6, rows, pitch)[:, :, :rows])


def _from_oracle(f, rows, pitch):
    out = np.zeros((6, rows, pitch))
    out[:, :, :rows] = f
    return out.reshape(-1)


@pytest.fixture(scope="module")
def cgrid(oracle_lib):
    return oracle_lib.CGrid(N, 6.4e6, 7.3e-5, 9.8)


@pytest.mark.parametrize("t,name", [("o", "J"), ("o", "Qi12"), ("o", "f"), ("x", "J"), ("x", "Qi11"),
                                    ("y", "J"), ("y", "Qi22"), ("q", "J")])
def test_c_geometry_bitwise(cgrid, t, name):
    rows, pitch, F = _layout()
    outs = {k: np.zeros(F) for k in ("J", "Qi11", "Qi12", "Qi22", "f")}
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    axis = (ctypes.c_double * 3)(0.0, 0.0, 1.0)
    tt = "oxyq".index(t)
    _lib().swmh_c_geometry(N, tt, H, rows, pitch, 7.3e-5, axis, dp(outs["J"]), dp(outs["Qi11"]),
                           dp(outs["Qi12"]), dp(outs["Qi22"]), dp(outs["f"]))
    ours = cgrid.interior(t, _to_oracle(outs[name], rows, pitch))
    ref = cgrid.interior(t, cgrid.geometry(t, name))
    assert np.array_equal(ours, ref)


def _lists():
    rows, pitch, F = _layout()
    mv, mf, mo, mq = 24 * (N + 1) * 3, 24 * N, 24 * N * 3, 24 * (N + 1)
    ll = lambda m: np.zeros(m, dtype=np.int64)
    vd, vs, fd, fsrc, od, os_, qo, qn = ll(mv), ll(mv), ll(mf), ll(mf), ll(mo), ll(mo), ll(mq), ll(mq)
    vsg, fsg = np.zeros(mv, dtype=np.int8), np.zeros(mf, dtype=np.int8)
    counts = np.zeros(4, dtype=np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
    rc = _lib().swmh_c_halo_lists(N, H, rows, pitch, F, 3, 3, p(vd), p(vs), vsg.ctypes.data, p(fd), p(fsrc),
                                  fsg.ctypes.data, p(od), p(os_), p(qo), p(qn),
                                  counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert rc == 0 and list(counts) == [mv, mf, mo, mq]
    return dict(vd=vd, vs=vs, vsg=vsg, fd=fd, fs=fsrc, fsg=fsg, od=od, os=os_, qo=qo, qn=qn, F=F)


def test_c_vector_halo_lists_match_oracle_exchange(cgrid):
    rows, pitch, F = _layout()
    L = _lists()
    rng = np.random.default_rng(3)
    u = cgrid.embed("x", rng.standard_normal((6, N, N + 1)))
    v = cgrid.embed("y", rng.standard_normal((6, N + 1, N)))
    pair = np.concatenate([_from_oracle(u, rows, pitch), _from_oracle(v, rows, pitch)])
    for d, s, sg in ((L["vd"], L["vs"], L["vsg"]), (L["fd"], L["fs"], L["fsg"])):
        pair[d] = np.where(sg < 0, -pair[s], pair[s])
    uo, vo = u.copy(), v.copy()
    cgrid.exchange_vec(uo, vo, 3)
    ref = np.concatenate([_from_oracle(uo, rows, pitch), _from_oracle(vo, rows, pitch)])
    for d in (L["vd"], L["fd"]):
        assert np.array_equal(pair[d], ref[d])
    # every listed target is a halo point and every source a mesh point of its field
    assert len(np.unique(np.concatenate([L["vd"], L["fd"]]))) == len(L["vd"]) + len(L["fd"])


def test_c_o_halo_lists_match_oracle_exchange(cgrid):
    rows, pitch, F = _layout()
    L = _lists()
    f = cgrid.embed("o", np.random.default_rng(4).standard_normal((6, N, N)))
    ours = _from_oracle(f, rows, pitch)
    ours[L["od"]] = ours[L["os"]]
    fo = f.copy()
    cgrid.exchange_o(fo, 3)
    assert np.array_equal(ours[L["od"]], _from_oracle(fo, rows, pitch)[L["od"]])


def test_c_q_sync_neighbours_are_the_ah_exchange_copies(oracle_lib):
    rows, pitch, F = _layout()
    L = _lists()
    G = oracle_lib.Grid(N)
    src, _ = G.halo_map()
    n = N + 1
    # map oracle flat indices (6, ld, ld), ld = rows, into the stored layout
    conv = lambda k: (k // (rows * rows)) * rows * pitch + ((k // rows) % rows) * pitch + k % rows
    k = 0
    for p in range(6):
        for e in range(4):
            for t in range(1, n + 1):
                i, j = {0: (0, t), 1: (n + 1, t), 2: (t, 0), 3: (t, n + 1)}[e]
                s = src[p, j - 1 + H, i - 1 + H]
                assert s >= 0 and conv(int(s)) == L["qn"][k], (p, e, t)
                k += 1


@pytest.mark.parametrize("hw", [1, 2, 3, 5])
@pytest.mark.parametrize("cov", [0, 1])
def test_c_ecs_vec_tables_bitwise(cgrid, hw, cov):
    """Tables of the ECS vector halo on the C grid (interpolation weights and bases, basis
    transforms, covariant stretch) computed by the host code = the oracle's, bit for bit."""
    ref = cgrid.ecs_vec_tables(hw, cov)
    n = (N + 1) * hw
    d = {k: np.zeros(4 * n) for k in ("w0", "w1", "w2")}
    d.update({k: np.zeros(n) for k in ("q1", "q2", "ct")})
    k = {x: np.zeros(n, dtype=np.int32) for x in ("i0", "i1", "i2")}
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    assert _lib().swmh_c_ecs_vec_tables(N, hw, cov, dp(d["w0"]), ip(k["i0"]), dp(d["w1"]), ip(k["i1"]), dp(d["w2"]),
                                        ip(k["i2"]), dp(d["q1"]), dp(d["q2"]), dp(d["ct"])) == 0
    nn = N * hw   # the normal component's tables cover positions 1..N
    for name, m in (("w0", 4 * n), ("w1", 4 * nn), ("w2", 4 * n), ("q1", n), ("q2", n), ("i0", n), ("i1", nn),
                    ("i2", n)) + ((("ct", n),) if cov else ()):
        ours = d[name] if name in d else k[name]
        assert np.array_equal(ours[:m], ref[name][:m]), name
    assert _lib().swmh_c_ecs_vec_tables(N, 6, cov, dp(d["w0"]), ip(k["i0"]), dp(d["w1"]), ip(k["i1"]), dp(d["w2"]),
                                        ip(k["i2"]), dp(d["q1"]), dp(d["q2"]), dp(d["ct"])) != 0


def test_c_ecs_strip_sources_match_oracle_exchange(cgrid):
    """The rows gathered beyond each edge for the ECS halos (normal component at face depth D,
    D = 0 the neighbour's copy of the shared face; tangential at centre depth D; cells at depth D)
    are the values the oracle's C-grid exchanges put there, signs included."""
    rows, pitch, F = _layout()
    DE = 8
    nn, nt = 24 * DE * N, 24 * DE * (N + 1)
    src = np.zeros(nn + nt + nn, dtype=np.int64)
    sg = np.zeros(nn + nt, dtype=np.int8)
    lp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
    s_n, s_t, s_o = src[:nn], src[nn:nn + nt], src[nn + nt:]
    assert _lib().swmh_c_ecs_strip_sources(N, H, rows, pitch, F, DE, lp(s_n), sg[:nn].ctypes.data_as(ctypes.c_void_p),
                                           lp(s_t), sg[nn:].ctypes.data_as(ctypes.c_void_p), lp(s_o)) == 0
    rng = np.random.default_rng(17)
    u = cgrid.embed("x", rng.standard_normal((6, N, N + 1)))
    v = cgrid.embed("y", rng.standard_normal((6, N + 1, N)))
    f = cgrid.embed("o", rng.standard_normal((6, N, N)))
    uv = np.concatenate([_from_oracle(u, rows, pitch), _from_oracle(v, rows, pitch)])
    sgn = np.where(sg < 0, -1.0, 1.0)
    En = (uv[s_n] * sgn[:nn]).reshape(6, 4, DE, N)
    Et = (uv[s_t] * sgn[nn:]).reshape(6, 4, DE, N + 1)
    Eo = _from_oracle(f, rows, pitch)[s_o].reshape(6, 4, DE, N)
    ue, ve, fe = u.copy(), v.copy(), f.copy()
    cgrid.exchange_vec(ue, ve, 3)
    cgrid.exchange_o(fe, 3)
    hw = cgrid.hw
    at = lambda a, p, i, j: a[p, j - 1 + hw, i - 1 + hw]
    for p in range(6):
        for D in range(3):
            for kk in range(1, N + 1):
                assert En[p, 0, D, kk - 1] == at(ue, p, -D, kk)
                assert En[p, 1, D, kk - 1] == at(ue, p, N + 2 + D, kk)
                assert En[p, 2, D, kk - 1] == at(ve, p, kk, -D)
                assert En[p, 3, D, kk - 1] == at(ve, p, kk, N + 2 + D)
                assert Eo[p, 0, D, kk - 1] == at(fe, p, -D, kk)
                assert Eo[p, 1, D, kk - 1] == at(fe, p, N + 1 + D, kk)
                assert Eo[p, 2, D, kk - 1] == at(fe, p, kk, -D)
                assert Eo[p, 3, D, kk - 1] == at(fe, p, kk, N + 1 + D)
            for kk in range(1, N + 2):
                assert Et[p, 0, D, kk - 1] == at(ve, p, -D, kk)
                assert Et[p, 1, D, kk - 1] == at(ve, p, N + 1 + D, kk)
                assert Et[p, 2, D, kk - 1] == at(ue, p, kk, -D)
                assert Et[p, 3, D, kk - 1] == at(ue, p, kk, N + 1 + D)
    # every value gathered is a distinct interior value of the neighbour (no halo, no repeats
    # along a row)
    assert np.all(np.abs(En) > 0) and np.all(np.abs(Et) > 0) and np.all(np.abs(Eo) > 0)
