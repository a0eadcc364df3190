"""ctypes wrapper of the oracle library (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Fields are
numpy float64 arrays of shape (6, ld, ld) = (panel, j, i) with halo width hw,
i.e. reference indices i,j in [1-hw, n+hw] map to array index [.., j-1+hw, i-1+hw].
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            from oracle import build as _b  # noqa: F401  (local import: build on demand)
            _b.build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ora_create.restype = ctypes.c_void_p
        L.ora_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.ora_destroy.argtypes = [ctypes.c_void_p]
        for f in ("ora_n", "ora_hw", "ora_ld"):
            getattr(L, f).restype = ctypes.c_int
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.ora_num_threads.restype = ctypes.c_int
        L.ora_num_threads.argtypes = []
        L.ora_field_len.restype = ctypes.c_long
        L.ora_field_len.argtypes = [ctypes.c_void_p]
        L.ora_geometry.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.ora_sync_weights.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.ora_quad_weights.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.ora_halo_map.argtypes = [ctypes.c_void_p, _ip, _ip]
        L.ora_exchange.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        L.ora_exchange_vec.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.ora_sbp_apply.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char, _dp, _dp]
        L.ora_div.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ora_grad.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ora_curl.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ora_scalar_sync.argtypes = [ctypes.c_void_p, _dp]
        L.ora_vec_sync_covariant.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.ora_ecs_xy_weights.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _ip]
        L.ora_ecs_halo_xy.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        L.ora_ecs_halo_xy.restype = ctypes.c_int
        L.ora_rhs.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.ora_rk4.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int]
        L.ora_diagnostics.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.ora_quad_mass.restype = ctypes.c_double
        L.ora_quad_mass.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.ora_l2norm.restype = ctypes.c_double
        L.ora_l2norm.argtypes = [ctypes.c_void_p, _dp]
        L.ora_maxabs.restype = ctypes.c_double
        L.ora_maxabs.argtypes = [ctypes.c_void_p, _dp]
        L.ora_project_vector.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.ora_narrow_min_n.restype = ctypes.c_int
        L.ora_narrow_min_n.argtypes = [ctypes.c_int]
        L.ora_laplace_narrow.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.ora_hordiff_scalar_narrow.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.ora_hordiff_vec_narrow.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp,
                                             _dp]
        L.ora_diff_step_narrow.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.ora_latlon_regrid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _ip,
                                         _ip, _dp]
        L.ora_ecs_o_weights.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _ip]
        L.ora_ecs_halo_o.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
        L.ora_ecs_halo_o.restype = ctypes.c_int
        L.ora_ecs_vec_a_tm.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.ora_ecs_halo_vec_a.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.ora_ecs_halo_vec_a.restype = ctypes.c_int
        L.ora_rk4_iomega.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double]
        L.ora_cgrid_create.restype = ctypes.c_void_p
        L.ora_cgrid_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.ora_cgrid_destroy.argtypes = [ctypes.c_void_p]
        L.ora_cgrid_grid.restype = ctypes.c_void_p
        L.ora_cgrid_grid.argtypes = [ctypes.c_void_p]
        L.ora_cgrid_geometry.restype = ctypes.c_int
        L.ora_cgrid_geometry.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.ora_c_co2contra.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.ora_c_massflux.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.ora_c_KE.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.ora_c_grad.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ora_c_curl.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ora_c_coriolis.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.ora_c_div.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ora_c_exchange_o.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        L.ora_c_exchange_vec.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.ora_c_rhs.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.ora_c_rk4.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int]
        L.ora_c_l2norm.restype = ctypes.c_double
        L.ora_c_l2norm.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.ora_c_maxabs.restype = ctypes.c_double
        L.ora_c_maxabs.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.ora_c_quad_mass.restype = ctypes.c_double
        L.ora_c_quad_mass.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.ora_c_diagnostics.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.ora_c_ecs_vec_tables.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _ip, _dp, _ip, _dp,
                                           _ip, _dp, _dp, _dp]
        L.ora_c_ecs_halo_vec.restype = ctypes.c_int
        L.ora_c_ecs_halo_vec.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.ora_c_grad_c2_ecs.restype = ctypes.c_int
        L.ora_c_grad_c2_ecs.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ora_c_massflux_up4.restype = ctypes.c_int
        L.ora_c_massflux_up4.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _pn(a):
    return None if a is None else _p(a)


def _pi(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


GEO = {"J": 0, "Qi11": 1, "Qi12": 2, "Qi22": 3, "f": 4, "r": 5, "a1": 8, "a2": 11,
       "b1": 14, "b2": 17, "Q11": 20, "Q12": 21, "Q22": 22}


class Grid:
    """One equiangular cubed sphere with Ah (corner-point) storage."""

    def __init__(self, N, grad_order=42, div_kind=42, scale=1.0, omega=1.0, grav=1.0):
        self.N = N
        self.scale, self.omega, self.grav = scale, omega, grav
        self.grad_order = grad_order
        self.div_kind = div_kind
        self.h = lib().ora_create(N, grad_order, div_kind, scale, omega, grav)
        if not self.h:
            raise RuntimeError("ora_create failed")
        self.n = lib().ora_n(self.h)
        self.hw = lib().ora_hw(self.h)
        self.ld = lib().ora_ld(self.h)
        self.shape = (6, self.ld, self.ld)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_destroy(self.h)
            self.h = None

    def zeros(self):
        return np.zeros(self.shape)

    def geometry(self, name):
        w = GEO[name]
        if name in ("r", "a1", "a2", "b1", "b2"):
            out = [self.zeros() for _ in range(3)]
            for c in range(3):
                lib().ora_geometry(self.h, w + c, _p(out[c]))
            return out
        out = self.zeros()
        lib().ora_geometry(self.h, w, _p(out))
        return out

    def sync_weights(self):
        out = {}
        for k, name in enumerate(("qb", "qt", "ql", "qr")):
            a = np.zeros(self.n)
            lib().ora_sync_weights(self.h, k, _p(a))
            out[name] = a
        return out

    def quad_weights(self, order):
        a = np.zeros(self.n)
        lib().ora_quad_weights(self.h, order, _p(a))
        return a

    def halo_map(self):
        src = np.zeros(self.shape, dtype=np.int32)
        vec = np.zeros(self.shape, dtype=np.int32)
        lib().ora_halo_map(self.h, src.ctypes.data_as(_ip), vec.ctypes.data_as(_ip))
        return src, vec

    def interior(self, f):
        """(6, n, n) view of the mesh points."""
        hw, n = self.hw, self.n
        return f[:, hw:hw + n, hw:hw + n]

    def embed(self, a):
        """(6, n, n) mesh values -> stored field with zero halo."""
        f = self.zeros()
        self.interior(f)[...] = a
        return f

    # ---- operators ----
    def sbp(self, order, direction, f):
        out = self.zeros()
        lib().ora_sbp_apply(self.h, order, direction.encode(), _p(f), _p(out))
        return out

    def scalar_sync(self, f):
        """In place: halo_Ah_scalar_sync (after a width-1 exchange)."""
        lib().ora_scalar_sync(self.h, _p(f))

    def vec_sync_covariant(self, gx, gy):
        """In place: covariant Ah vector sync of (gx, gy) (after a width-1 vector exchange)."""
        lib().ora_vec_sync_covariant(self.h, _p(gx), _p(gy))

    def ecs_halo_xy(self, f, hw):
        """In place: ECS cubic halo of width hw on the stored field f."""
        if lib().ora_ecs_halo_xy(self.h, _p(f), int(hw)) != 0:
            raise ValueError("halo width out of range")

    def ecs_xy_weights(self, hw):
        L = self.n + 2 * hw
        w = np.zeros((hw, L, 4))
        ind = np.zeros((hw, L), dtype=np.int32)
        lib().ora_ecs_xy_weights(self.h, int(hw), _p(w), ind.ctypes.data_as(_ip))
        return w, ind

    def div(self, u, v):
        out = self.zeros()
        lib().ora_div(self.h, _p(u), _p(v), _p(out))
        return out

    def grad(self, f):
        gx, gy = self.zeros(), self.zeros()
        lib().ora_grad(self.h, _p(f), _p(gx), _p(gy))
        return gx, gy

    def curl(self, u, v):
        out = self.zeros()
        lib().ora_curl(self.h, _p(u), _p(v), _p(out))
        return out

    def rhs(self, h, u, v, hs=None):
        dh, du, dv = self.zeros(), self.zeros(), self.zeros()
        lib().ora_rhs(self.h, _p(h), _p(u), _p(v), _pn(hs), _p(dh), _p(du), _p(dv))
        return dh, du, dv

    def rk4(self, h, u, v, dt, nsteps, hs=None):
        lib().ora_rk4(self.h, _p(h), _p(u), _p(v), _pn(hs), dt, nsteps)

    # narrow-stencil SBP Laplacians (order 21, 42 or 63) and hyperdiffusion; N+1 >= 2 (21), 12 (42),
    # 18 (63)
    def _check_narrow(self, order):
        if order not in (21, 42, 63) or self.n < lib().ora_narrow_min_n(order):
            raise ValueError(f"narrow Laplacian of order {order} needs order 21/42/63 and N+1 large enough")

    def laplace(self, f, order=42):
        self._check_narrow(order)
        out = self.zeros()
        lib().ora_laplace_narrow(self.h, order, _p(f), _p(out))
        return out

    def hordiff_scalar(self, coeff, f, order=42):
        self._check_narrow(order)
        out = self.zeros()
        lib().ora_hordiff_scalar_narrow(self.h, order, float(coeff), _p(f), _p(out))
        return out

    def hordiff_vec(self, coeff, u, v, order=42):
        self._check_narrow(order)
        ut, vt = self.zeros(), self.zeros()
        lib().ora_hordiff_vec_narrow(self.h, order, float(coeff), _p(u), _p(v), _p(ut), _p(vt))
        return ut, vt

    def diff_step(self, h, u, v, dt, uv_coeff, h_coeff, uv_order=42, h_order=42):
        self._check_narrow(uv_order)
        self._check_narrow(h_order)
        lib().ora_diff_step_narrow(self.h, uv_order, h_order, _p(h), _p(u), _p(v), float(dt),
                                   float(uv_coeff), float(h_coeff))

    def ecs_o_weights(self, rows):
        """init_halo_interp weights (rows, N+4, 4) for positions -1..N+2 and stencil base indices."""
        T = self.N + 4
        w = np.zeros((rows, T, 4))
        ind = np.zeros((rows, T), dtype=np.int32)
        lib().ora_ecs_o_weights(self.h, int(rows), _p(w), ind.ctypes.data_as(_ip))
        return w, ind

    def ecs_halo_o(self, f, hw):
        """ECS_O halo of a cell-centre field f (6, N, N) -> (6, N+2hw, N+2hw); NaN where the
        procedure writes nothing."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == (6, self.N, self.N)
        L = self.N + 2 * hw
        out = np.zeros((6, L, L))
        if lib().ora_ecs_halo_o(self.h, _p(f), int(hw), _p(out)) != 0:
            raise ValueError("halo width out of range (1..8, N >= 8)")
        return out

    def ecs_vec_a_tm(self, rows):
        """init_transform_matrix (4 edges, rows, N positions, 4 entries)."""
        tm = np.zeros((4, rows, self.N, 4))
        lib().ora_ecs_vec_a_tm(self.h, int(rows), _p(tm))
        return tm

    def ecs_halo_vec_a(self, u, v, hw):
        """ECS vector halo (A grid, contravariant) of cell-centre (u, v) (6, N, N) each."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        L = self.N + 2 * hw
        ou, ov = np.zeros((6, L, L)), np.zeros((6, L, L))
        if lib().ora_ecs_halo_vec_a(self.h, _p(u), _p(v), int(hw), _p(ou), _p(ov)) != 0:
            raise ValueError("halo width out of range (1..8, N >= 8)")
        return ou, ov

    def latlon_regrid(self, Nlat, Nlon, f, fv=None, table=False):
        """Cubic lat-lon regrid (Nlat, Nlon) of an Ah scalar f, or of covariant (f, fv) to (zonal,
        meridional) components; with table=True also (panel, x0, y0, wx, wy) per target point."""
        a = np.zeros((Nlat, Nlon))
        b = np.zeros((Nlat, Nlon)) if fv is not None else None
        ip = lambda x: x.ctypes.data_as(_ip)
        tab = None
        if table:
            tab = (np.zeros((Nlat, Nlon), np.int32), np.zeros((Nlat, Nlon), np.int32),
                   np.zeros((Nlat, Nlon), np.int32), np.zeros((Nlat, Nlon, 8)))
        lib().ora_latlon_regrid(self.h, int(Nlat), int(Nlon), _p(f), _pn(fv), _p(a), _pn(b),
                                ip(tab[0]) if tab else None, ip(tab[1]) if tab else None,
                                ip(tab[2]) if tab else None, _p(tab[3]) if tab else None)
        out = (a,) if fv is None else (a, b)
        if table:
            out = out + (tab[0], tab[1], tab[2], tab[3][..., :4], tab[3][..., 4:])
        return out

    def diagnostics(self, quad_order, h, u, v, hs=None):
        out = np.zeros(7)
        lib().ora_diagnostics(self.h, quad_order, _p(h), _p(u), _p(v), _pn(hs), _p(out))
        return out

    def quad_mass(self, order, f):
        return lib().ora_quad_mass(self.h, order, _p(f))

    def l2norm(self, f):
        return lib().ora_l2norm(self.h, _p(f))

    def maxabs(self, f):
        return lib().ora_maxabs(self.h, _p(f))

    def project(self, kind, vx, vy, vz):
        """kind 'covariant' or 'contravariant'; returns (u, v) over the stored extent."""
        u, v = self.zeros(), self.zeros()
        lib().ora_project_vector(self.h, 0 if kind == "covariant" else 1,
                                 _p(np.ascontiguousarray(vx)), _p(np.ascontiguousarray(vy)),
                                 _p(np.ascontiguousarray(vz)), _p(u), _p(v))
        return u, v


def rk4_iomega(w, f, dt):
    w = np.asarray(w, dtype=np.complex128)
    f = np.asarray(f, dtype=np.complex128)
    wr, wi = np.ascontiguousarray(w.real), np.ascontiguousarray(w.imag)
    fr, fi = np.ascontiguousarray(f.real), np.ascontiguousarray(f.imag)
    lib().ora_rk4_iomega(len(w), _p(wr), _p(wi), _p(fr), _p(fi), dt)
    return fr + 1j * fi


C_TYPES = {"o": 0, "x": 1, "y": 2, "q": 3}
C_GEO = {"J": 0, "Qi11": 1, "Qi12": 2, "Qi22": 3, "f": 4, "r": 5, "a1": 6, "a2": 7, "b1": 8, "b2": 9}


class CGrid:
    """C-grid (staggering 'C') storage on one cubed sphere of N cells per panel side: h at o
    points (N x N), u at x points (N+1 x N), v at y points (N x N+1), curl at q points
    (N+1 x N+1).  Arrays have the (6, ld, ld) layout of the Ah grid of the same N; index (i, j)
    of every point type sits at [.., j-1+hw, i-1+hw]."""

    def __init__(self, N, scale=1.0, omega=1.0, grav=1.0):
        self.N = N
        self.scale, self.omega, self.grav = scale, omega, grav
        self.h = lib().ora_cgrid_create(N, scale, omega, grav)
        if not self.h:
            raise RuntimeError("ora_cgrid_create failed")
        g = lib().ora_cgrid_grid(self.h)
        self.hw = lib().ora_hw(g)
        self.ld = lib().ora_ld(g)
        self.shape = (6, self.ld, self.ld)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_cgrid_destroy(self.h)
            self.h = None

    def zeros(self):
        return np.zeros(self.shape)

    def dims(self, t):
        t = C_TYPES[t]
        return self.N + (t in (1, 3)), self.N + (t in (2, 3))   # (nx, ny)

    def interior(self, t, f):
        nx, ny = self.dims(t)
        return f[:, self.hw:self.hw + ny, self.hw:self.hw + nx]

    def embed(self, t, a):
        f = self.zeros()
        self.interior(t, f)[...] = a
        return f

    def geometry(self, t, name):
        tt, w = C_TYPES[t], C_GEO[name]
        if name in ("r", "a1", "a2", "b1", "b2"):
            out = [self.zeros() for _ in range(3)]
            for c in range(3):
                lib().ora_cgrid_geometry(self.h, tt, w, c, _p(out[c]))
            return out
        out = self.zeros()
        lib().ora_cgrid_geometry(self.h, tt, w, 0, _p(out))
        return out

    def project(self, t, kind, vx, vy, vz, comp):
        """Component `comp` (0 along x, 1 along y) of the cartesian vectors at points of type t:
        sum([vx,vy,vz]*a1) for covariant, *b1 for contravariant (test_fields_mod.f90:397-433)."""
        e = self.geometry(t, ("a1" if comp == 0 else "a2") if kind == "covariant" else
                          ("b1" if comp == 0 else "b2"))
        return vx * e[0] + vy * e[1] + vz * e[2]

    def co2contra(self, u, v):
        ut, vt = self.zeros(), self.zeros()
        lib().ora_c_co2contra(self.h, _p(u), _p(v), _p(ut), _p(vt))
        return ut, vt

    def massflux(self, h, ut, vt):
        hu, hv = self.zeros(), self.zeros()
        lib().ora_c_massflux(self.h, _p(h), _p(ut), _p(vt), _p(hu), _p(hv))
        return hu, hv

    def KE(self, u, v, ut, vt):
        out = self.zeros()
        lib().ora_c_KE(self.h, _p(u), _p(v), _p(ut), _p(vt), _p(out))
        return out

    def grad(self, f):
        gx, gy = self.zeros(), self.zeros()
        lib().ora_c_grad(self.h, _p(f), _p(gx), _p(gy))
        return gx, gy

    def curl(self, u, v):
        out = self.zeros()
        lib().ora_c_curl(self.h, _p(u), _p(v), _p(out))
        return out

    def coriolis(self, hu, hv, h, curl):
        cu, cv = self.zeros(), self.zeros()
        lib().ora_c_coriolis(self.h, _p(hu), _p(hv), _p(h), _p(curl), _p(cu), _p(cv))
        return cu, cv

    def div(self, u, v):
        out = self.zeros()
        lib().ora_c_div(self.h, _p(u), _p(v), _p(out))
        return out

    def exchange_o(self, f, depth):
        lib().ora_c_exchange_o(self.h, _p(f), int(depth))

    def exchange_vec(self, u, v, depth):
        lib().ora_c_exchange_vec(self.h, _p(u), _p(v), int(depth))

    def rhs(self, h, u, v, hs=None):
        dh, du, dv = self.zeros(), self.zeros(), self.zeros()
        lib().ora_c_rhs(self.h, _p(h), _p(u), _p(v), _pn(hs), _p(dh), _p(du), _p(dv))
        return dh, du, dv

    def rk4(self, h, u, v, dt, nsteps, hs=None):
        lib().ora_c_rk4(self.h, _p(h), _p(u), _p(v), _pn(hs), dt, nsteps)

    def l2norm(self, t, f):
        return lib().ora_c_l2norm(self.h, C_TYPES[t], _p(f))

    def maxabs(self, t, f):
        return lib().ora_c_maxabs(self.h, C_TYPES[t], _p(f))

    def quad_mass(self, t, f):
        return lib().ora_c_quad_mass(self.h, C_TYPES[t], _p(f))

    def ecs_vec_tables(self, hw, covariant):
        """Interpolation tables of the C-grid ECS vector halo: dict of w0, i0 (tangential), w1, i1
        (normal), w2, i2 (cross-component), q1, q2 (basis transform), ct (covariant stretch)."""
        N = self.N
        np_ = (N + 1) * hw
        t = {k: np.zeros(4 * np_) for k in ("w0", "w1", "w2")}
        t.update({k: np.zeros(np_, dtype=np.int32) for k in ("i0", "i1", "i2")})
        t.update({k: np.zeros(np_) for k in ("q1", "q2", "ct")})
        lib().ora_c_ecs_vec_tables(self.h, int(hw), int(bool(covariant)), _p(t["w0"]), _pi(t["i0"]), _p(t["w1"]),
                                   _pi(t["i1"]), _p(t["w2"]), _pi(t["i2"]), _p(t["q1"]), _p(t["q2"]), _p(t["ct"]))
        return t

    def ecs_halo_vec(self, u, v, hw, covariant=False):
        """ECS vector halo of (u, v) (the reference's ecs_C_vec or ecs_C_vec_covariant procedure):
        arrays (6, N+2hw, N+1+2hw) for u and (6, N+1+2hw, N+2hw) for v, i and j from 1-hw."""
        N = self.N
        ou = np.zeros((6, N + 2 * hw, N + 1 + 2 * hw))
        ov = np.zeros((6, N + 1 + 2 * hw, N + 2 * hw))
        rc = lib().ora_c_ecs_halo_vec(self.h, _p(u), _p(v), int(hw), int(bool(covariant)), _p(ou), _p(ov))
        if rc:
            raise ValueError("ora_c_ecs_halo_vec: 1 <= hw <= 5 and N >= 8 required")
        return ou, ov

    def grad_c2_ecs(self, f):
        gx, gy = self.zeros(), self.zeros()
        if lib().ora_c_grad_c2_ecs(self.h, _p(f), _p(gx), _p(gy)):
            raise ValueError("ora_c_grad_c2_ecs: N >= 8 required")
        return gx, gy

    def massflux_up4(self, f, u, v):
        fx, fy = self.zeros(), self.zeros()
        if lib().ora_c_massflux_up4(self.h, _p(f), _p(u), _p(v), _p(fx), _p(fy)):
            raise ValueError("ora_c_massflux_up4: N >= 8 required")
        return fx, fy

    def diagnostics(self, h, u, v, hs=None):
        out = np.zeros(7)
        lib().ora_c_diagnostics(self.h, _p(h), _p(u), _p(v), _pn(hs), _p(out))
        return dict(zip(("hmin", "hmax", "mass", "TE", "KE", "PE", "Enstrophy"), out.tolist()))
