"""MI355X (gfx950) shallow-water dynamical core on the equiangular cubed sphere.

Host-side mirror of the reference's operator / stvec / timescheme interface for the
Ah-grid vector-invariant shallow-water model, backed by the HIP library libswm_gpu.so
(C ABI in include/swm_gpu.h):

    op = create_swm_operator(N, **operator_names(42), scale=a, omega=om, grav=g)
    v = create_stvec_swm(op); v.upload(h, u, v_)          # stvec_swm_t
    k = v.create_similar(); op.apply(k, v)                # operator_t%apply
    ts = create_timescheme(op, "rk4"); ts.step(v, dt)     # timescheme_t%step
    op.get_diagnostics(v)                                 # operator_t%get_diagnostics

Host arrays use the reference tile layout: shape (6, n+2*hw, n+2*hw), index [panel, j, i]
with reference indices i,j = 1..n stored at [.., j-1+hw, i-1+hw], n = N+1.
"""
import ctypes
import os

import numpy as np

from . import _capi
from ._capi import SwmError, check

__all__ = ["SwmError", "operator_names", "create_latlon_outputer", "create_latlon_vec_outputer",
           "create_master_paneled_outputer", "create_swm_operator", "create_stvec_swm", "create_timescheme",
           "OperatorSWM", "StvecSWM", "RK4", "Ensemble", "EARTH", "OperatorSWMC", "StvecSWMC", "RK4C",
           "C_GRID_NAMES"]

# const_mod.f90:6-9
EARTH = {"radius": 0.6371229e7, "grav": 9.80616, "omega": 2.0 * np.pi / (23 * 60 * 60 + 56 * 60 + 4)}


# the &shallow_water_model_c namelist group (namelist_swm:80-102)
C_GRID_NAMES = {"div_op_name": "divergence_c_sbp42", "grad_op_name": "gradient_c_sbp42",
                "curl_op_name": "curl_c_sbp42", "quadrature_name": "SBP_C42_quadrature",
                "coriolis_op_name": "coriolis_Cgrid_sbp42", "KE_op_name": "KE_Cgrid_sbp42",
                "co2contra_op_name": "co2contra_c_sbp42", "massflux_op_name": "massflux_c_sbp42"}


def operator_names(order):
    """Operator names of the reference's sbp{21,42,43,63} run configurations
    (models/shallow_water/test/scripts/ts2_ah/gen_namelist.sh:1-32); "c42" gives the C-grid
    SBP42 set of namelist_swm:80-102."""
    if order == "c42":
        return dict(C_GRID_NAMES)
    table = {
        21: ("divergence_ah2", "gradient_ah21_sbp_ecs", "SBP_Ah21_quadrature"),
        42: ("divergence_ah42_sbp", "gradient_ah42_sbp_ecs", "SBP_Ah42_quadrature"),
        43: ("divergence_ah43_sbp", "gradient_ah43_sbp_ecs", "SBP_Ah63_quadrature"),
        63: ("divergence_ah63_sbp", "gradient_ah63_sbp_ecs", "SBP_Ah63_quadrature"),
    }
    d, g, q = table[order]
    return {"div_op_name": d, "grad_op_name": g, "curl_op_name": "curl_" + d, "quadrature_name": q}


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OperatorSWM:
    """Domain + vector-invariant SWM operator (create_domain + create_swm_operator)."""

    def __init__(self, N, div_op_name, grad_op_name, curl_op_name, quadrature_name,
                 scale=1.0, omega=1.0, grav=1.0, precision=64):
        L = _capi.lib()
        h = ctypes.c_void_p()
        check(L.swm_ctx_create(ctypes.byref(h), int(N), div_op_name.encode(), grad_op_name.encode(),
                               curl_op_name.encode(), quadrature_name.encode(), float(scale), float(omega),
                               float(grav), int(precision)))
        self.ctx = h
        self.N = N
        self.precision = precision
        self.quadrature_name = quadrature_name
        n, pitch, fe, eb = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int()
        check(L.swm_ctx_info(self.ctx, ctypes.byref(n), ctypes.byref(pitch), ctypes.byref(fe), ctypes.byref(eb)))
        self.n, self.pitch, self.field_elems, self.elem_bytes = n.value, pitch.value, fe.value, eb.value

    def close(self):
        if getattr(self, "ctx", None):
            _capi.lib().swm_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(_capi.lib().swm_ctx_synchronize(self.ctx))

    def set_stream(self, hip_stream_handle):
        check(_capi.lib().swm_ctx_set_stream(self.ctx, ctypes.c_void_p(hip_stream_handle)))

    # operator_t%apply (operator_swm_mod.f90:61-113)
    def apply(self, vout, vin):
        check(_capi.lib().swm_rhs(self.ctx, ctypes.c_void_p(vin.ptr), ctypes.c_void_p(vout.ptr)))

    # operator_t%get_diagnostics (operator_swm_mod.f90:115-155)
    def get_diagnostics(self, v):
        out = np.zeros(7)
        check(_capi.lib().swm_diagnostics(self.ctx, ctypes.c_void_p(v.ptr), _dp(out)))
        return dict(zip(("hmin", "hmax", "mass", "TE", "KE", "PE", "Enstrophy"), out.tolist()))

    def set_orography(self, hs, halo_width=0):
        if hs is None:
            check(_capi.lib().swm_set_orography(self.ctx, None, 0))
            return
        hs = np.ascontiguousarray(hs, dtype=np.float64)
        self._check_host(hs, halo_width)
        check(_capi.lib().swm_set_orography(self.ctx, _dp(hs), int(halo_width)))

    def host_shape(self, halo_width=0):
        ld = self.n + 2 * halo_width
        return (6, ld, ld)

    def _check_host(self, a, halo_width):
        if a.shape != self.host_shape(halo_width):
            raise ValueError(f"host array shape {a.shape} != {self.host_shape(halo_width)}")

    def initial_state(self, seed, lmax, h0=1.0e4, h_amp=100.0, v_max=40.0, halo_width=0):
        """Seeded spherical-harmonic state (h, u, v) as host arrays (swm_ic_spherical_harmonics)."""
        shp = self.host_shape(halo_width)
        h, u, v = np.zeros(shp), np.zeros(shp), np.zeros(shp)
        check(_capi.lib().swm_ic_spherical_harmonics(self.ctx, int(seed), int(lmax), float(h0), float(h_amp),
                                                     float(v_max), _dp(h), _dp(u), _dp(v), int(halo_width)))
        return h, u, v

    def set_tiling(self, tile_rows=8, metric_on_the_fly=False, tile_cols=None):
        """Tile shape of the fused stage kernel: tile_cols x tile_rows = 64 x 8, 32 x 16, 16 x 32
        (512 threads) or 32 x 8 (256 threads) points; tile_cols=None takes the 512-thread shape of
        tile_rows.  FP64 may evaluate the metric terms on the fly instead of reading them
        (identical values)."""
        cols = 512 // int(tile_rows) if tile_cols is None else int(tile_cols)
        check(_capi.lib().swm_ctx_set_tile_shape(self.ctx, cols, int(tile_rows), int(bool(metric_on_the_fly))))

    def get_tiling(self):
        c, r, m = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(_capi.lib().swm_ctx_get_tile_shape(self.ctx, ctypes.byref(c), ctypes.byref(r), ctypes.byref(m)))
        return {"tile_cols": c.value, "tile_rows": r.value, "metric_on_the_fly": bool(m.value)}

    def autotune(self, v, reps=50):
        """Time every tile configuration with RK4 steps on state `v` (left unchanged) and keep the
        fastest (returns us per step)."""
        best = ctypes.c_double()
        check(_capi.lib().swm_ctx_autotune(self.ctx, ctypes.c_void_p(v.ptr), int(reps), ctypes.byref(best)))
        return best.value

    def quadrature_weights(self):
        """(6, n, n) weights w = J hx^2 A_i A_j of the operator's SBP quadrature on the unit sphere
        (sbp_quadrature_mod.f90:21-61): sum(w * f) is the quadrature of f."""
        L = _capi.lib()
        n, N = self.n, self.N
        pitch = self.pitch
        J = np.zeros(6 * n * pitch)
        L.swmh_geometry(N, pitch, 1.0, (ctypes.c_double * 3)(0.0, 0.0, 1.0), _dp(J), None, None, None, None,
                        None, None, None)
        A = np.zeros(n)
        order = {"SBP_Ah21_quadrature": 21, "SBP_Ah42_quadrature": 42, "SBP_Ah63_quadrature": 63}[self.quadrature_name]
        check(0 if L.swmh_quad_weights(N, order, _dp(A)) == 0 else -1)
        hx = (np.pi / 2) / N
        J = J.reshape(6, n, pitch)[:, :, :n]
        return J * hx * hx * A[None, :, None] * A[None, None, :]

    def ecs_halo_xy(self, x, component, halo_width):
        """ECS cubic halo (ecs_halo_xy_t%get_halo_scalar) of field `component` (0 h, 1 u, 2 v)
        of state `x`: host array (6, n+2hw, n+2hw), [panel, j, i] with i, j from 1-hw to n+hw."""
        L = self.n + 2 * halo_width
        out = np.zeros((6, L, L))
        field = ctypes.c_void_p(x.ptr + int(component) * self.field_elems * self.elem_bytes)
        check(_capi.lib().swm_ecs_halo_xy_host(self.ctx, field, int(halo_width), _dp(out)))
        return out

    def ecs_halo_o(self, f_o, halo_width):
        """ECS cubic halo on O points (the reference's "ECS_O" halo procedure) of a cell-centre
        field f_o (6, N, N) [panel, j, i]: host array (6, N+2hw, N+2hw) with i, j from 1-hw to
        N+hw; NaN at the corner-square points the procedure does not write (hw >= 3)."""
        f_o = np.ascontiguousarray(f_o, dtype=np.float64)
        if f_o.shape != (6, self.N, self.N):
            raise ValueError(f"expected a (6, {self.N}, {self.N}) cell-centre field")
        L = self.N + 2 * halo_width
        out = np.zeros((6, L, L))
        check(_capi.lib().swm_ecs_halo_o_host(self.ctx, _dp(f_o), int(halo_width), _dp(out)))
        return out

    def ecs_halo_vec_a(self, u_o, v_o, halo_width):
        """ECS vector halo on O points (the reference's A-grid vector halo procedure) of cell-centre
        contravariant components u_o, v_o (6, N, N): host arrays (6, N+2hw, N+2hw) each."""
        u_o = np.ascontiguousarray(u_o, dtype=np.float64)
        v_o = np.ascontiguousarray(v_o, dtype=np.float64)
        for a in (u_o, v_o):
            if a.shape != (6, self.N, self.N):
                raise ValueError(f"expected (6, {self.N}, {self.N}) cell-centre fields")
        L = self.N + 2 * halo_width
        ou, ov = np.zeros((6, L, L)), np.zeros((6, L, L))
        check(_capi.lib().swm_ecs_halo_vec_a_host(self.ctx, _dp(u_o), _dp(v_o), int(halo_width), _dp(ou), _dp(ov)))
        return ou, ov

    # transposed linearised right-hand side (adjoint form)
    def rhs_vjp(self, out, y, mu):
        """out = (dF/dy)^T mu at state y (FP64, SBP families 42/43/63)."""
        check(_capi.lib().swm_rhs_vjp(self.ctx, ctypes.c_void_p(y.ptr), ctypes.c_void_p(mu.ptr),
                                      ctypes.c_void_p(out.ptr)))

    def laplace(self, out, x, laplace_name="laplace_ah_sbp42_narrow"):
        """h plane of `out` = narrow-stencil SBP Laplacian of the h plane of `x`
        (laplace_Ah_sbp{21,42,63}_narrow_t %calc_laplace, picked by the reference's names
        laplace_ah_sbp21_narrow / laplace_ah_sbp42_narrow / laplace_ah_sbp63_narrow); FP64,
        N+1 >= 12 (sbp42) or 18 (sbp63)."""
        check(_capi.lib().swm_laplace(self.ctx, laplace_name.encode(), ctypes.c_void_p(x.ptr),
                                      ctypes.c_void_p(out.ptr)))

    def diff_apply(self, vout, vin, uv_diff_coeff, h_diff_coeff,
                   hordiff_uv_name="hordiff_vec_xyz_Ah_sbp_42_narrow",
                   hordiff_h_name="hordiff_scalar_Ah_sbp_42_narrow"):
        """vout = hyperdiffusion tendency of vin (operator_swm_diff_t%apply)."""
        check(_capi.lib().swm_diff_apply(self.ctx, ctypes.c_void_p(vin.ptr), ctypes.c_void_p(vout.ptr),
                                         hordiff_uv_name.encode(), float(uv_diff_coeff),
                                         hordiff_h_name.encode(), float(h_diff_coeff)))

    def diff_step(self, x, dt, uv_diff_coeff, h_diff_coeff,
                  hordiff_uv_name="hordiff_vec_xyz_Ah_sbp_42_narrow",
                  hordiff_h_name="hordiff_scalar_Ah_sbp_42_narrow"):
        """x += dt * hyperdiffusion tendency of x (explicit_Eul1 on operator_swm_diff_t)."""
        check(_capi.lib().swm_diff_step(self.ctx, ctypes.c_void_p(x.ptr), float(dt), hordiff_uv_name.encode(),
                                        float(uv_diff_coeff), hordiff_h_name.encode(), float(h_diff_coeff)))

    def set_adjoint_tiling(self, tile_x):
        """Transposed RHS as one fused launch on tile_x x 16 tiles (16 or 32), or as two launches (0)."""
        check(_capi.lib().swm_adj_set_tiling(self.ctx, int(tile_x)))

    def debug_geometry(self, which, halo_width=0):
        out = np.zeros(self.host_shape(halo_width))
        check(_capi.lib().swm_debug_geometry(self.ctx, int(which), _dp(out), int(halo_width)))
        return out


class Outputer:
    """Writers of the reference's on-disk formats (outputer_factory_mod.f90): REAL*4 direct-access
    records, record r at byte (r-1) * 4 * (values per record).

    create_latlon_outputer(op, Nlat, Nlon, "Ah")                 -> .write(x.field("h"), file, rec)
    create_latlon_vec_outputer(op, Nlat, Nlon, "Ah", "covariant") -> .write_vec(x, file_u, file_v, rec)
    create_master_paneled_outputer(op, "xy")                     -> .write(x.field("h"), file, rec)
    """

    def __init__(self, op, handle=None, paneled=False):
        self.op, self.handle, self.paneled = op, handle, paneled

    def write(self, field, file_name, rec_num=1):
        if self.paneled:
            check(_capi.lib().swm_paneled_write(self.op.ctx, field, os.fsencode(file_name), int(rec_num)))
        else:
            check(_capi.lib().swm_latlon_write(self.handle, field, os.fsencode(file_name), int(rec_num)))

    def write_vec(self, x, file_name_u, file_name_v, rec_num=1):
        check(_capi.lib().swm_latlon_vec_write(self.handle, x.field("u"), x.field("v"), os.fsencode(file_name_u),
                                               os.fsencode(file_name_v), int(rec_num)))

    def __del__(self):
        try:
            if self.handle:
                _capi.lib().swm_outputer_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def create_latlon_outputer(op, Nlat, Nlon, scalar_grid_type="Ah"):
    h = ctypes.c_void_p()
    check(_capi.lib().swm_latlon_outputer_create(op.ctx, int(Nlat), int(Nlon), scalar_grid_type.encode(),
                                                 ctypes.byref(h)))
    return Outputer(op, h)


def create_latlon_vec_outputer(op, Nlat, Nlon, staggering="Ah", components_type="covariant"):
    h = ctypes.c_void_p()
    check(_capi.lib().swm_latlon_vec_outputer_create(op.ctx, int(Nlat), int(Nlon), staggering.encode(),
                                                     components_type.encode(), ctypes.byref(h)))
    return Outputer(op, h)


def create_master_paneled_outputer(op, points_type="xy"):
    if points_type != "xy":
        raise SwmError(f"the device path writes Ah (xy) points only, not {points_type}")
    return Outputer(op, paneled=True)


class StvecSWM:
    """Device-resident state vector (h, u, v) of one operator's domain (stvec_swm_t)."""

    def __init__(self, op):
        self.op = op
        p = ctypes.c_void_p()
        check(_capi.lib().swm_state_alloc(op.ctx, ctypes.byref(p)))
        self.ptr = p.value

    def free(self):
        if getattr(self, "ptr", None) and getattr(self.op, "ctx", None):
            _capi.lib().swm_state_free(self.op.ctx, ctypes.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def create_similar(self):
        return StvecSWM(self.op)

    def field(self, name):
        """Device pointer of the h, u or v field (as the reference's state%h, state%u, state%v)."""
        k = {"h": 0, "u": 1, "v": 2}[name]
        return ctypes.c_void_p(self.ptr + k * self.op.field_elems * self.op.elem_bytes)

    def upload(self, h, u, v, halo_width=0):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (h, u, v)]
        for a in arrs:
            self.op._check_host(a, halo_width)
        check(_capi.lib().swm_state_upload(self.op.ctx, ctypes.c_void_p(self.ptr), *(_dp(a) for a in arrs),
                                           int(halo_width)))

    def download(self, halo_width=0):
        shp = self.op.host_shape(halo_width)
        h, u, v = np.zeros(shp), np.zeros(shp), np.zeros(shp)
        check(_capi.lib().swm_state_download(self.op.ctx, ctypes.c_void_p(self.ptr), _dp(h), _dp(u), _dp(v),
                                             int(halo_width)))
        return h, u, v

    def from_tensors(self, h, u, v):
        """Copy three contiguous device tensors of shape (6, n, n) into this state (enqueued on
        the operator's stream after the tensors' producers have finished)."""
        import torch
        ts = [h, u, v]
        for t in ts:
            if t.shape != (6, self.op.n, self.op.n) or not t.is_contiguous() or not t.is_cuda:
                raise ValueError("expected contiguous device tensors of shape (6, n, n)")
        torch.cuda.current_stream().synchronize()
        check(_capi.lib().swm_state_copy_from_device(self.op.ctx, ctypes.c_void_p(self.ptr),
                                                     *(ctypes.c_void_p(t.data_ptr()) for t in ts)))

    def to_tensors(self, h, u, v):
        """Copy this state into three contiguous device tensors of shape (6, n, n); returns after
        the copy has completed."""
        ts = [h, u, v]
        for t in ts:
            if t.shape != (6, self.op.n, self.op.n) or not t.is_contiguous() or not t.is_cuda:
                raise ValueError("expected contiguous device tensors of shape (6, n, n)")
        check(_capi.lib().swm_state_copy_to_device(self.op.ctx, ctypes.c_void_p(self.ptr),
                                                   *(ctypes.c_void_p(t.data_ptr()) for t in ts)))
        self.op.synchronize()

    # stvec_t%assign / update (stvec_swm_mod.f90:38-80)
    def assign(self, s1, v1, s2, v2):
        check(_capi.lib().swm_assign(self.op.ctx, ctypes.c_void_p(self.ptr), float(s1), ctypes.c_void_p(v1.ptr),
                                     float(s2), ctypes.c_void_p(v2.ptr)))

    def update(self, s1, v1, s2, v2):
        check(_capi.lib().swm_update(self.op.ctx, ctypes.c_void_p(self.ptr), float(s1), ctypes.c_void_p(v1.ptr),
                                     float(s2), ctypes.c_void_p(v2.ptr)))


class RK4:
    """Classical RK4 (time_schemes/rk4_mod.f90) with fused stage kernels."""

    def __init__(self, op):
        self.op = op

    def step(self, v0, dt):
        self.steps(v0, dt, 1)

    def steps(self, v0, dt, nsteps):
        check(_capi.lib().swm_rk4_steps(self.op.ctx, ctypes.c_void_p(v0.ptr), float(dt), int(nsteps)))


class Ensemble:
    """Cubes of different resolutions (one operator and one state each) advanced together:
    one launch per RK stage over all members (a persistent grid or one block per task, as the
    autotuner picks), one captured hipGraph per step."""

    def __init__(self, ops, states):
        if len(ops) != len(states) or not ops:
            raise ValueError("one state per operator")
        self.ops, self.states = list(ops), list(states)
        m = len(ops)
        ctxs = (ctypes.c_void_p * m)(*[o.ctx.value for o in ops])
        sts = (ctypes.c_void_p * m)(*[s.ptr for s in states])
        h = ctypes.c_void_p()
        check(_capi.lib().swm_ens_create(ctxs, sts, m, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _capi.lib().swm_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def steps(self, dts, nsteps):
        d = np.ascontiguousarray(dts, dtype=np.float64)
        if d.shape != (len(self.ops),):
            raise ValueError("one time step per member")
        check(_capi.lib().swm_ens_rk4_steps(self.h, _dp(d), int(nsteps)))

    def set_tiling(self, tile_rows=8, metric_on_the_fly=False, blocks_per_slot=1, tile_cols=None):
        cols = 512 // int(tile_rows) if tile_cols is None else int(tile_cols)
        check(_capi.lib().swm_ens_set_tile_shape(self.h, cols, int(tile_rows), int(bool(metric_on_the_fly)),
                                                 int(blocks_per_slot)))

    def get_tiling(self):
        c, r, m, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(_capi.lib().swm_ens_get_tile_shape(self.h, ctypes.byref(c), ctypes.byref(r), ctypes.byref(m),
                                                 ctypes.byref(b)))
        return {"tile_cols": c.value, "tile_rows": r.value, "metric_on_the_fly": bool(m.value),
                "blocks_per_slot": b.value}

    def autotune(self, reps=20):
        best = ctypes.c_double()
        check(_capi.lib().swm_ens_autotune(self.h, int(reps), ctypes.byref(best)))
        return best.value

    def bench_stage(self, reps):
        check(_capi.lib().swm_ens_bench_stage(self.h, int(reps)))

    def synchronize(self):
        self.ops[0].synchronize()


class OperatorSWMC:
    """C-grid (staggering 'C') domain + vector-invariant SWM operator with the SBP42 C-grid
    operators (create_domain + create_swm_operator with the &shallow_water_model_c names), FP64.

    Host arrays have the shape of the Ah grid of the same N, (6, N+1+2hw, N+1+2hw): h uses the
    points i, j = 1..N, u (x faces) i = 1..N+1, j = 1..N, v (y faces) i = 1..N, j = 1..N+1."""

    precision = 64

    def __init__(self, N, div_op_name, grad_op_name, curl_op_name, quadrature_name, coriolis_op_name,
                 KE_op_name, co2contra_op_name, massflux_op_name, scale=1.0, omega=1.0, grav=1.0):
        L = _capi.lib()
        h = ctypes.c_void_p()
        names = (div_op_name, grad_op_name, curl_op_name, coriolis_op_name, KE_op_name, co2contra_op_name,
                 massflux_op_name, quadrature_name)
        check(L.swm_c_create(ctypes.byref(h), int(N), *(n.encode() for n in names), float(scale), float(omega),
                             float(grav)))
        self.ctx = h
        self.N = N
        self.n = N + 1
        Nv, hw, rows, pitch, fe = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        check(L.swm_c_info(self.ctx, ctypes.byref(Nv), ctypes.byref(hw), ctypes.byref(rows), ctypes.byref(pitch),
                           ctypes.byref(fe)))
        self.halo, self.rows, self.pitch, self.field_elems = hw.value, rows.value, pitch.value, fe.value
        self.elem_bytes = 8

    def close(self):
        if getattr(self, "ctx", None):
            _capi.lib().swm_c_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(_capi.lib().swm_c_synchronize(self.ctx))

    def host_shape(self, halo_width=0):
        ld = self.n + 2 * halo_width
        return (6, ld, ld)

    def _check_host(self, a, halo_width):
        if a.shape != self.host_shape(halo_width):
            raise ValueError(f"host array shape {a.shape} != {self.host_shape(halo_width)}")

    # operator_t%apply (operator_swm_mod.f90:61-113)
    def apply(self, vout, vin):
        check(_capi.lib().swm_c_rhs(self.ctx, ctypes.c_void_p(vin.ptr), ctypes.c_void_p(vout.ptr)))

    # operator_t%get_diagnostics (operator_swm_mod.f90:115-236); enstrophy is NaN on the C grid
    def get_diagnostics(self, v):
        out = np.zeros(7)
        check(_capi.lib().swm_c_diagnostics(self.ctx, ctypes.c_void_p(v.ptr), _dp(out)))
        return dict(zip(("hmin", "hmax", "mass", "TE", "KE", "PE", "Enstrophy"), out.tolist()))

    def set_orography(self, hs, halo_width=0):
        if hs is None:
            check(_capi.lib().swm_c_set_orography(self.ctx, None, 0))
            return
        hs = np.ascontiguousarray(hs, dtype=np.float64)
        self._check_host(hs, halo_width)
        check(_capi.lib().swm_c_set_orography(self.ctx, _dp(hs), int(halo_width)))

    # ---- ECS halos of the C grid and the operators on them (host arrays in, host arrays out) ----
    def _device_buffer(self, n):
        import torch
        return torch.zeros(int(n), dtype=torch.float64, device="cuda")

    def _from_stored(self, buf, halo_width):
        """A stored field (field_elems doubles) as a host array (6, N+1+2hw, N+1+2hw)."""
        H, n = self.halo, self.n
        st = buf.reshape(6, self.rows, self.pitch)
        ld = n + 2 * halo_width
        out = np.zeros((6, ld, ld))
        m = min(halo_width, H)
        out[:, halo_width - m:halo_width + n + m, halo_width - m:halo_width + n + m] = st[:, H - m:H + n + m,
                                                                                        H - m:H + n + m]
        return out

    def _state_of(self, h, u, v, halo_width):
        x = StvecSWMC(self)
        z = np.zeros(self.host_shape(halo_width))
        x.upload(z if h is None else h, z if u is None else u, z if v is None else v, halo_width)
        return x

    def ecs_halo_vector(self, u, v, halo_width, procedure="ecs_C_vec", halo_width_in=0):
        """get_halo_vector of the "ecs_C_vec" (contravariant) or "ecs_C_vec_covariant" halo
        procedure for host u (x points), v (y points): arrays (6, N+2hw, N+1+2hw) and
        (6, N+1+2hw, N+2hw) for i, j from 1-hw, NaN in the corner squares."""
        import torch
        N, hw = self.N, int(halo_width)
        x = self._state_of(None, u, v, halo_width_in)
        n = 6 * (N + 2 * hw) * (N + 1 + 2 * hw)
        ou, ov = self._device_buffer(n), self._device_buffer(n)
        torch.cuda.current_stream().synchronize()
        check(_capi.lib().swm_c_ecs_halo_vec(self.ctx, procedure.encode(), ctypes.c_void_p(x.ptr), hw,
                                             ctypes.c_void_p(ou.data_ptr()), ctypes.c_void_p(ov.data_ptr())))
        self.synchronize()
        x.free()
        return (ou.cpu().numpy().reshape(6, N + 2 * hw, N + 1 + 2 * hw),
                ov.cpu().numpy().reshape(6, N + 1 + 2 * hw, N + 2 * hw))

    def calc_grad(self, f, grad_op_name="gradient_c2_ecs", halo_width=1):
        """calc_grad of "gradient_c2_ecs" for a host o-point field f: (gx, gy) host arrays with
        halo halo_width (gx also on the rows j = 0, N+1 and gy on the columns i = 0, N+1)."""
        import torch
        x = self._state_of(f, None, None, halo_width)
        gx, gy = self._device_buffer(self.field_elems), self._device_buffer(self.field_elems)
        torch.cuda.current_stream().synchronize()
        check(_capi.lib().swm_c_grad(self.ctx, grad_op_name.encode(), ctypes.c_void_p(x.ptr),
                                     ctypes.c_void_p(gx.data_ptr()), ctypes.c_void_p(gy.data_ptr())))
        self.synchronize()
        x.free()
        return self._from_stored(gx.cpu().numpy(), halo_width), self._from_stored(gy.cpu().numpy(), halo_width)

    def calc_massflux(self, f, u, v, massflux_op_name="massflux_c_up4", halo_width=0):
        """calc_massflux of "massflux_c_up4": (fx, fy) host arrays for host f (o points), u, v."""
        import torch
        x = self._state_of(f, u, v, halo_width)
        fs = self.field_elems
        fx, fy = self._device_buffer(fs), self._device_buffer(fs)
        torch.cuda.current_stream().synchronize()
        base = x.ptr
        check(_capi.lib().swm_c_massflux(self.ctx, massflux_op_name.encode(), ctypes.c_void_p(base),
                                         ctypes.c_void_p(base + 8 * fs), ctypes.c_void_p(base + 16 * fs),
                                         ctypes.c_void_p(fx.data_ptr()), ctypes.c_void_p(fy.data_ptr())))
        self.synchronize()
        x.free()
        return self._from_stored(fx.cpu().numpy(), halo_width), self._from_stored(fy.cpu().numpy(), halo_width)


class StvecSWMC:
    """Device-resident C-grid state vector (h at o points, u at x faces, v at y faces)."""

    def __init__(self, op):
        self.op = op
        p = ctypes.c_void_p()
        check(_capi.lib().swm_c_state_alloc(op.ctx, ctypes.byref(p)))
        self.ptr = p.value

    def free(self):
        if getattr(self, "ptr", None) and getattr(self.op, "ctx", None):
            _capi.lib().swm_c_state_free(self.op.ctx, ctypes.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def create_similar(self):
        return StvecSWMC(self.op)

    def upload(self, h, u, v, halo_width=0):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (h, u, v)]
        for a in arrs:
            self.op._check_host(a, halo_width)
        check(_capi.lib().swm_c_state_upload(self.op.ctx, ctypes.c_void_p(self.ptr), *(_dp(a) for a in arrs),
                                             int(halo_width)))

    def download(self, halo_width=0):
        shp = self.op.host_shape(halo_width)
        h, u, v = np.zeros(shp), np.zeros(shp), np.zeros(shp)
        check(_capi.lib().swm_c_state_download(self.op.ctx, ctypes.c_void_p(self.ptr), _dp(h), _dp(u), _dp(v),
                                               int(halo_width)))
        return h, u, v


class RK4C:
    """Classical RK4 (time_schemes/rk4_mod.f90) on the C-grid operator, HIP-graph captured."""

    def __init__(self, op):
        self.op = op

    def step(self, v0, dt):
        self.steps(v0, dt, 1)

    def steps(self, v0, dt, nsteps):
        check(_capi.lib().swm_c_rk4_steps(self.op.ctx, ctypes.c_void_p(v0.ptr), float(dt), int(nsteps)))


def create_swm_operator(N, div_op_name, grad_op_name, curl_op_name, quadrature_name, scale=1.0, omega=1.0,
                        grav=1.0, precision=64, coriolis_op_name=None, KE_op_name=None, co2contra_op_name=None,
                        massflux_op_name=None):
    """create_swm_operator (operator_swm_factory_mod.f90:14-41): the Ah-grid operator for the
    &shallow_water_model names, the C-grid operator for the &shallow_water_model_c names."""
    if div_op_name.startswith("divergence_c_"):
        if precision != 64:
            raise SwmError("the C-grid operators run in FP64 only")
        return OperatorSWMC(N, div_op_name, grad_op_name, curl_op_name, quadrature_name,
                            coriolis_op_name or "", KE_op_name or "", co2contra_op_name or "",
                            massflux_op_name or "", scale, omega, grav)
    return OperatorSWM(N, div_op_name, grad_op_name, curl_op_name, quadrature_name, scale, omega, grav, precision)


def create_stvec_swm(op):
    return StvecSWMC(op) if isinstance(op, OperatorSWMC) else StvecSWM(op)


def create_timescheme(op, name="rk4"):
    if name != "rk4":
        raise SwmError(f"unknown or unsupported time scheme on the device path: {name}")
    return RK4C(op) if isinstance(op, OperatorSWMC) else RK4(op)
