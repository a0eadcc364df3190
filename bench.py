#!/usr/bin/env python3
"""Benchmark: RK4 steps of the Ah-grid vector-invariant shallow-water model on one MI355X.

Workload (default): C384 (6 x 384 x 384 cells, 6 x 385^2 Ah points), sbp42 operators
(divergence_ah42_sbp, gradient_ah42_sbp_ecs, curl_divergence_ah42_sbp), FP64, RK4 with
dt = 75 s (gravity-wave Courant number 0.9; 150 s is beyond RK4's stability limit), Earth radius / gravity / rotation, seeded spherical-harmonic initial state
(h = 1e4 m + perturbation of degree <= 32 with max 100 m, max |V| = 40 m/s).

A "step" is one RK4 time step (4 fused stage kernels) over the whole sphere; the metric
is cell-steps/s = 6 N^2 * steps / time, inputs resident in HBM before the timed region.
The dominant kernel (RK stages 2/3) is timed separately with HIP events on its stream
and priced against the HBM roofline with its algorithmic bytes per launch.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s at C384 sbp42 vs HBM roofline; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
# Algorithmic traffic per Ah point of one RK stage-2/3 launch: read y(3) + metric J, Qi11,
# Qi12, Qi22, f (5) + v0(3) + acc(3), write y(3) + acc(3) = 20 values.
STAGE23_VALUES = 20
STEP_VALUES = 71  # per point per RK4 step (14 + 20 + 20 + 17)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=384)
    ap.add_argument("--order", type=int, default=42)
    # 150 s (gravity-wave CFL 1.8) is beyond the stability limit of RK4 with these SBP operators
    # (the FP64 state overflows within ~10 steps, on the CPU oracle as on the GPU); 75 s
    # (CFL 0.9) keeps the same grid, operators and work per step
    ap.add_argument("--dt", type=float, default=75.0)
    ap.add_argument("--lmax", type=int, default=32)
    ap.add_argument("--precision", type=int, default=64, choices=(32, 64))
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--autotune-reps", type=int, default=30, help="RK4 steps per tiling candidate in the autotuner")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--c-steps", type=int, default=200, help="C-grid RK4 steps timed")
    ap.add_argument("--c-dt", type=float, default=37.5,
                    help="C-grid time step (the C-grid scheme's RK4 limit is near gravity-wave Courant number 0.5)")
    ap.add_argument("--forms", default="fp32,adjoint,batched,hyperdiffusion,cgrid",
                    help="comma-separated extra measurements beside the FP64 headline: fp32, adjoint, batched, "
                         "hyperdiffusion, cgrid")
    ap.add_argument("--diff-steps", type=int, default=100, help="hyperdiffusion steps timed")
    # precision form: C768 sbp42, RK4 at gravity-wave Courant 0.9 (37.5 s), degree <= 64
    ap.add_argument("--fp32-N", type=int, default=768)
    ap.add_argument("--fp32-steps", type=int, default=1000)
    ap.add_argument("--fp32-dt", type=float, default=37.5)
    ap.add_argument("--fp32-lmax", type=int, default=64)
    # adjoint form: C192 sbp42 FP64, 200 RK4 steps at Courant 0.9 (150 s), checkpoints every 20
    ap.add_argument("--adj-N", type=int, default=192)
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--adj-dt", type=float, default=150.0)
    ap.add_argument("--adj-every", type=int, default=20)
    ap.add_argument("--adj-fd", type=int, default=4, help="finite-difference directions checked")
    ap.add_argument("--adj-full", type=int, default=1,
                    help="also time the gradient with one segment (every stage input stored, no recomputation)")
    ap.add_argument("--adj-tiling", type=int, default=16, choices=(0, 16, 32),
                    help="transposed RHS: fused launch on tiles of this width x 16, or two launches (0)")
    # batched form: 16 members each at C48, C96, C144, C192, sbp42 FP64, 200 RK4 steps at
    # Courant 0.5 per member, degree <= 16, seed = member id
    ap.add_argument("--ens-sizes", default="48,96,144,192")
    ap.add_argument("--ens-per-size", type=int, default=16)
    ap.add_argument("--ens-steps", type=int, default=200)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--tiling", default="auto",
                    help="stage-kernel tiling: auto (autotuner) or COLSxROWS[,otf] with COLSxROWS one of 64x8, "
                         "32x16, 16x32, 32x8")
    return ap.parse_args()


def profiled_traffic(N, order, prec, tiling):
    """HBM bytes per stage-2/3 launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (copy-kernel
    calibrated) for this workload and tiling, from the committed profile summary; None if absent."""
    import re
    path = os.path.join(ROOT, "profiles", f"stage23_c{N}_sbp{order}_f{prec}.md")
    if not os.path.exists(path):
        return None
    key = (f"C{N} sbp{order} f{prec} tile={tiling['tile_cols']}x{tiling['tile_rows']} "
           f"otf={int(tiling['metric_on_the_fly'])}")
    for line in open(path):
        m = re.match(r"traffic_bytes_per_launch (.*): (\d+)\s*$", line)
        if m and m.group(1) == key:
            return int(m.group(2))
    return None


def _timed_steps(ts, x, dt, steps, stream):
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    ts.steps(x, dt, steps)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3


def fp32_form(args, stream):
    """FP32 stencils / halo / RK stages (metric precomputed in FP64, rounded once; FP64
    compensated integrals) against the FP64 device path on the same initial state."""
    import numpy as np
    import cubedsphere_sbp_swm_amd as swm
    N, dt, steps = args.fp32_N, args.fp32_dt, args.fp32_steps
    ops, res = {}, {}  # ops keeps both contexts alive until the comparison is done
    h0 = u0 = v0 = None
    for prec in (64, 32):
        op = swm.create_swm_operator(N, **swm.operator_names(42), scale=swm.EARTH["radius"],
                                     omega=swm.EARTH["omega"], grav=swm.EARTH["grav"], precision=prec)
        op.set_stream(stream.cuda_stream)
        if h0 is None:
            h0, u0, v0 = op.initial_state(seed=0, lmax=args.fp32_lmax)
        x = swm.create_stvec_swm(op)
        x.upload(h0, u0, v0)
        op.autotune(x, reps=20)
        x.upload(h0, u0, v0)
        d0 = op.get_diagnostics(x)
        ts = swm.create_timescheme(op, "rk4")
        ts.steps(x, dt, 5)  # warm-up / graph capture
        x.upload(h0, u0, v0)
        secs = _timed_steps(ts, x, dt, steps, stream)
        d1 = op.get_diagnostics(x)
        res[prec] = {"secs": secs, "fields": x.download(), "d0": d0, "d1": d1, "tiling": op.get_tiling(),
                     "eb": op.elem_bytes}
        ops[prec] = (op, x)
    cells, points = 6 * N * N, 6 * (N + 1) ** 2
    errs = {}
    for name, a, b in zip(("h", "u", "v"), res[32]["fields"], res[64]["fields"]):
        d = a - b
        errs[name] = {"rel_l2": float(np.sqrt(np.sum(d * d) / np.sum(b * b))),
                      "rel_linf": float(np.max(np.abs(d)) / np.max(np.abs(b)))}
    hp = res[64]["fields"][0] - 1.0e4
    dh = res[32]["fields"][0] - res[64]["fields"][0]
    errs["h_minus_1e4"] = {"rel_l2": float(np.sqrt(np.sum(dh * dh) / np.sum(hp * hp))),
                           "rel_linf": float(np.max(np.abs(dh)) / np.max(np.abs(hp)))}
    drift = {p: (res[p]["d1"]["mass"] - res[p]["d0"]["mass"]) / res[p]["d0"]["mass"] for p in (32, 64)}
    rate = {p: cells * steps / res[p]["secs"] for p in (32, 64)}
    step_gbs32 = STEP_VALUES * 4 * points / (res[32]["secs"] / steps) / 1e9
    return {"workload": f"C{N} sbp42 RK4 dt={dt:g}s x {steps} steps, degree <= {args.fp32_lmax}",
            "fp32_cell_steps_per_s": rate[32], "fp64_cell_steps_per_s": rate[64],
            "speedup_vs_fp64": rate[32] / rate[64],
            "fp32_ms_per_step": res[32]["secs"] * 1e3 / steps, "fp64_ms_per_step": res[64]["secs"] * 1e3 / steps,
            "fp32_step_effective_GBs": step_gbs32, "fp32_step_frac": step_gbs32 / HBM_PEAK_GBS,
            "error_vs_fp64_gpu": errs, "mass_drift": {"fp32": drift[32], "fp64": drift[64]},
            "energy_drift": {f"fp{p}": (res[p]["d1"]["TE"] - res[p]["d0"]["TE"]) / res[p]["d0"]["TE"]
                             for p in (32, 64)},
            "tiling": {"fp32": res[32]["tiling"], "fp64": res[64]["tiling"]}}


def adjoint_form(args, stream):
    """Gradient of 0.5*sum(w (h_T - h_target)^2) w.r.t. the initial state by the hand-written
    adjoint (checkpointed forward + reverse sweep) vs the plain FP64 forward of the same steps."""
    import ctypes
    import numpy as np
    import torch
    import cubedsphere_sbp_swm_amd as swm
    from cubedsphere_sbp_swm_amd import _capi
    from cubedsphere_sbp_swm_amd.autograd import AdjointTape
    N, steps, dt, every = args.adj_N, args.adj_steps, args.adj_dt, args.adj_every
    op = swm.create_swm_operator(N, **swm.operator_names(42), scale=swm.EARTH["radius"],
                                 omega=swm.EARTH["omega"], grav=swm.EARTH["grav"])
    op.set_stream(stream.cuda_stream)
    y0 = op.initial_state(seed=0, lmax=16)
    x = swm.create_stvec_swm(op)
    x.upload(*op.initial_state(seed=1, lmax=16))
    op.autotune(x, reps=20)
    op.set_adjoint_tiling(args.adj_tiling)
    ts = swm.create_timescheme(op, "rk4")
    ts.steps(x, dt, steps)
    tgt = x.download()[0]
    w = op.quadrature_weights()
    tape = AdjointTape(op, steps, dt, every)
    L = _capi.lib()
    cells = 6 * N * N

    def ev():
        return torch.cuda.Event(enable_timing=True)

    def timed(fn, reps=3):
        fn()  # warm-up (graph capture, lazy work buffers)
        best = 1e30
        for _ in range(reps):
            e0, e1 = ev(), ev()
            torch.cuda.synchronize()
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) / 1e3)
        return best

    x.upload(*y0)
    xT, lam, g0 = (swm.create_stvec_swm(op) for _ in range(3))
    t_fwd = timed(lambda: ts.steps(x, dt, steps))
    tape.x0.upload(*y0)
    fwd = lambda: _capi.check(L.swm_adj_forward(tape.tape, ctypes.c_void_p(tape.x0.ptr), ctypes.c_void_p(xT.ptr)))
    fwd()
    hT = xT.download()[0]
    loss = 0.5 * float(np.sum(w * (hT - tgt) ** 2))
    z = np.zeros_like(hT)
    lam.upload(w * (hT - tgt), z, z)
    bwd = lambda: _capi.check(L.swm_adj_backward(tape.tape, ctypes.c_void_p(lam.ptr), ctypes.c_void_p(g0.ptr)))
    bwd()
    # one gradient = checkpointed forward then reverse sweep, timed as a pair (the forward
    # records the last segment, which the reverse sweep then does not recompute)
    def timed_pair(fwd, bwd):
        t_f, t_b = 1e30, 1e30
        for _ in range(3):
            e0, e1, e2 = ev(), ev(), ev()
            torch.cuda.synchronize()
            e0.record(stream)
            fwd()
            e1.record(stream)
            bwd()
            e2.record(stream)
            torch.cuda.synchronize()
            if e0.elapsed_time(e2) < (t_f + t_b) * 1e3:
                t_f, t_b = e0.elapsed_time(e1) / 1e3, e1.elapsed_time(e2) / 1e3
        return t_f, t_b

    t_tape_fwd, t_bwd = timed_pair(fwd, bwd)
    grads = g0.download()

    def loss_at(y):
        x.upload(*y)
        ts.steps(x, dt, steps)
        return 0.5 * float(np.sum(w * (x.download()[0] - tgt) ** 2))

    rng = np.random.default_rng(42)
    eps, ads, fds = 1e-4, [], []
    for _ in range(args.adj_fd):
        d = [rng.standard_normal(y0[0].shape) * sc for sc in (1.0, 0.1, 0.1)]
        ads.append(sum(float(np.sum(g * dd)) for g, dd in zip(grads, d)))
        fds.append((loss_at([a + eps * dd for a, dd in zip(y0, d)]) -
                    loss_at([a - eps * dd for a, dd in zip(y0, d)])) / (2 * eps))
    ads, fds = np.array(ads), np.array(fds)
    t_grad = t_tape_fwd + t_bwd
    full = _adjoint_single_segment(args, op, tape, y0, lam, grads, timed_pair, stream) if args.adj_full else None
    if full:
        full["adjoint_forward_time_ratio"] = full["gradient_ms"] / (t_fwd * 1e3)
    return {"workload": f"C{N} sbp42 FP64 RK4 dt={dt:g}s x {steps} steps, checkpoints every {every}, "
                        "loss 0.5*sum(w (h_T - h_target)^2)",
            "transposed_rhs": f"fused, {args.adj_tiling} x 16 tiles" if args.adj_tiling else "two launches",
            "adjoint_forward_time_ratio": t_grad / t_fwd,
            "forward_ms": t_fwd * 1e3, "gradient_ms": t_grad * 1e3,
            "checkpointed_forward_ms": t_tape_fwd * 1e3, "reverse_sweep_ms": t_bwd * 1e3,
            "forward_cell_steps_per_s": cells * steps / t_fwd, "loss": loss,
            "fd_check": {"directions": int(args.adj_fd), "step": eps,
                         "max_rel_err": float(np.max(np.abs(ads - fds) / np.abs(fds))) if len(fds) else None},
            "single_segment": full}


def _adjoint_single_segment(args, op, tape, y0, lam, grads, timed_pair, stream):
    """The same gradient with one checkpoint segment: the forward records every RK4 stage input
    of the whole trajectory, so the reverse sweep recomputes nothing. Reports its device memory
    (free memory before and after the tape is made), its time ratio and whether its gradient is
    bit for bit the checkpointed one."""
    import ctypes
    import numpy as np
    import torch
    import cubedsphere_sbp_swm_amd as swm
    from cubedsphere_sbp_swm_amd import _capi
    from cubedsphere_sbp_swm_amd.autograd import AdjointTape
    L = _capi.lib()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    full = AdjointTape(op, tape.nsteps, tape.dt, tape.nsteps)
    torch.cuda.synchronize()
    tape_bytes = free0 - torch.cuda.mem_get_info()[0]
    xT, g0 = swm.create_stvec_swm(op), swm.create_stvec_swm(op)
    full.x0.upload(*y0)
    fwd = lambda: _capi.check(L.swm_adj_forward(full.tape, ctypes.c_void_p(full.x0.ptr), ctypes.c_void_p(xT.ptr)))
    bwd = lambda: _capi.check(L.swm_adj_backward(full.tape, ctypes.c_void_p(lam.ptr), ctypes.c_void_p(g0.ptr)))
    fwd()
    bwd()
    t_f, t_b = timed_pair(fwd, bwd)
    same = all(np.array_equal(a, b) for a, b in zip(g0.download(), grads))
    full.close()
    return {"tape_bytes": int(tape_bytes), "forward_with_tape_ms": t_f * 1e3, "reverse_sweep_ms": t_b * 1e3,
            "gradient_ms": (t_f + t_b) * 1e3, "gradient_bitwise_equal_to_checkpointed": bool(same)}


# algorithmic FP64 values per point of one explicit hyperdiffusion step (4 fields: h and the
# Cartesian velocity): to Cartesian 9 read (h, u, v, b1, b2) / 4 written; per Laplacian
# application the 4 fields and J, Qi11, Qi12, Qi22 read once / 4 written; tendencies and Euler
# step 13 read / 3 written (edge syncs O(N))
DIFF_VALUES = (9 + 4) + 2 * ((4 + 4) + 4) + (13 + 3)


def hyperdiffusion_form(args, stream, op, x):
    """One explicit_Eul1 hyperdiffusion step (sbp42 narrow Laplacian, coefficients 0.03 / 0.01)
    on the headline state, timed over --diff-steps steps, against the RK4 step; the same for the
    sbp63 narrow pair (the one the reference's Ah namelist uses)."""
    import torch
    ev = lambda: torch.cuda.Event(enable_timing=True)

    def timed(order):
        names = (f"hordiff_vec_xyz_Ah_sbp_{order}_narrow", f"hordiff_scalar_Ah_sbp_{order}_narrow")
        op.diff_step(x, args.dt, 0.03, 0.01, *names)
        torch.cuda.synchronize()
        e0, e1 = ev(), ev()
        e0.record(stream)
        for _ in range(args.diff_steps):
            op.diff_step(x, args.dt, 0.03, 0.01, *names)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e3 / args.diff_steps

    us = timed(42)
    us63 = timed(63)
    points = 6 * (args.N + 1) ** 2
    gbs = DIFF_VALUES * 8 * points / (us * 1e-6) / 1e9
    return {"workload": f"C{args.N} sbp42 FP64 explicit_Eul1 hyperdiffusion step, hordiff_vec_xyz_Ah_sbp_42_narrow "
                        "0.03 + hordiff_scalar_Ah_sbp_42_narrow 0.01 (7 launches)",
            "us_per_step": us, "algorithmic_bytes_per_step": DIFF_VALUES * 8 * points,
            "effective_GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
            "sbp63_us_per_step": us63}


# Field passes of one C-grid RK4 step as implemented (fields read or written once per kernel,
# shared metric planes not counted): RHS = 4 (contra_o: u, v; A, B) + 9 (flux: A, B, h, u, v;
# G pair, KEu, KEv) + 6 (ke_div: KEu, KEv, h, G pair; KE) + 3 (curl: u, v; W) + 3 (w2uv: W; fu,
# fv) + 7 (cor_o: fu, fv, G pair, h, f; cup, cvp) + 3 (mom: KE, cup, cvp) = 35, plus the
# tendency outputs: stages 1-3 write k (3), read v0 (3) and write the next input (3), stage 4
# reads v0, k0, k1, k2 (12) and writes v0 (3): step = 4 x 35 + 3 x 9 + 15.
C_STEP_FIELD_PASSES = 182


def cgrid_ecs_timings(args, op, xs, reps=200):
    """The C-grid ECS halo operators on the device state (h, u, v) at the form's N: the vector halo
    of width 3 (both kinds), gradient_c2_ecs and massflux_c_up4; time per call over `reps` calls
    enqueued back to back (launch included), and the operator's minimum bytes (its inputs read
    once, its outputs written once) per call time."""
    import ctypes
    import torch
    from cubedsphere_sbp_swm_amd import _capi
    L, N, fs, hw = _capi.lib(), op.N, op.field_elems, 3
    nvec = 6 * (N + 2 * hw) * (N + 1 + 2 * hw)
    ou, ov = (torch.empty(nvec, dtype=torch.float64, device="cuda") for _ in range(2))
    gx, gy = (torch.empty(fs, dtype=torch.float64, device="cuda") for _ in range(2))
    torch.cuda.synchronize()
    vp = ctypes.c_void_p
    h_, u_, v_ = vp(xs.ptr), vp(xs.ptr + 8 * fs), vp(xs.ptr + 16 * fs)
    calls = {
        "ecs_C_vec_hw3": (lambda: L.swm_c_ecs_halo_vec(op.ctx, b"ecs_C_vec", vp(xs.ptr), hw, vp(ou.data_ptr()),
                                                       vp(ov.data_ptr())),
                          2 * 6 * N * (N + 1) + 2 * nvec),
        "ecs_C_vec_covariant_hw3": (lambda: L.swm_c_ecs_halo_vec(op.ctx, b"ecs_C_vec_covariant", vp(xs.ptr), hw,
                                                                 vp(ou.data_ptr()), vp(ov.data_ptr())),
                                    2 * 6 * N * (N + 1) + 2 * nvec),
        "gradient_c2_ecs": (lambda: L.swm_c_grad(op.ctx, b"gradient_c2_ecs", h_, vp(gx.data_ptr()), vp(gy.data_ptr())),
                            6 * N * N + 2 * 6 * (N + 1) * (N + 2)),
        "massflux_c_up4": (lambda: L.swm_c_massflux(op.ctx, b"massflux_c_up4", h_, u_, v_, vp(gx.data_ptr()),
                                                    vp(gy.data_ptr())),
                           6 * N * N + 4 * 6 * N * (N + 1)),
    }
    res = {}
    for name, (fn, values) in calls.items():
        for _ in range(5):
            _capi.check(fn())
        op.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            _capi.check(fn())
        op.synchronize()
        us = (time.perf_counter() - t0) * 1e6 / reps
        res[name] = {"us_per_call": us, "min_bytes": 8 * values, "min_bytes_GBs": 8 * values / (us * 1e-6) / 1e9}
    return res


def cgrid_form(args, stream):
    """C-grid SBP42 family (divergence_c_sbp42, gradient_c_sbp42, curl_c_sbp42,
    coriolis_Cgrid_sbp42, KE_Cgrid_sbp42, co2contra_c_sbp42, massflux_c_sbp42), FP64 RK4 on the
    same sphere and dt as the headline; the oracle's C-grid RK4 on the host as its CPU baseline."""
    import numpy as np
    import torch
    import cubedsphere_sbp_swm_amd as swm
    from oracle.oracle import CGrid, lib as oracle_lib
    N, dt, steps = args.N, args.c_dt, args.c_steps
    op = swm.create_swm_operator(N, **swm.operator_names("c42"), scale=swm.EARTH["radius"],
                                 omega=swm.EARTH["omega"], grav=swm.EARTH["grav"])
    C = CGrid(N, swm.EARTH["radius"], swm.EARTH["omega"], swm.EARTH["grav"])
    # solid-body rotation in geostrophic balance plus a 100 m Gaussian bump (test case 2 + bump)
    om, a, g = swm.EARTH["omega"], swm.EARTH["radius"], swm.EARTH["grav"]
    u0 = 40.0
    fields = []
    for t in "oxy":
        x, y, z = C.geometry(t, "r")
        r = np.sqrt(x * x + y * y + z * z)
        fields.append((x / r, y / r, z / r))
    xo, yo, zo = fields[0]
    h = 1.0e4 - (a * om * u0 + 0.5 * u0 ** 2) / g * zo ** 2 + 100.0 * np.exp(-((xo - 0.5) ** 2 + yo ** 2) / 0.05)
    h = C.embed("o", C.interior("o", h))
    u = C.embed("x", C.interior("x", C.project("x", "covariant", -u0 * fields[1][1], u0 * fields[1][0],
                                               0 * fields[1][2], 0)))
    v = C.embed("y", C.interior("y", C.project("y", "covariant", -u0 * fields[2][1], u0 * fields[2][0],
                                               0 * fields[2][2], 1)))
    xs = swm.create_stvec_swm(op)
    xs.upload(h, u, v, C.hw)
    ts = swm.create_timescheme(op, "rk4")
    ts.steps(xs, dt, 5)
    op.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ts.steps(xs, dt, steps)
    op.synchronize()
    secs = time.perf_counter() - t0
    cells, points = 6 * N * N, 6 * (N + 1) ** 2
    out = {"workload": f"C{N} C-grid SBP42 FP64 RK4 dt={dt:g}s x {steps} steps (test case 2 + 100 m bump)",
           "cell_steps_per_s": cells * steps / secs, "ms_per_step": secs * 1e3 / steps,
           "field_passes_per_step": C_STEP_FIELD_PASSES,
           "implementation_GBs": C_STEP_FIELD_PASSES * 8 * points / (secs / steps) / 1e9,
           "implementation_frac": C_STEP_FIELD_PASSES * 8 * points / (secs / steps) / 1e9 / HBM_PEAK_GBS,
           "fused_minimum_GBs_equiv": STEP_VALUES * 8 * points / (secs / steps) / 1e9}
    hd, _, _ = xs.download(C.hw)
    out["h_range_after"] = [float(C.interior("o", hd).min()), float(C.interior("o", hd).max())]
    out["ecs"] = cgrid_ecs_timings(args, op, xs)
    if not args.no_cpu:
        hc, uc, vc = h.copy(), u.copy(), v.copy()
        t0 = time.perf_counter()
        C.rk4(hc, uc, vc, dt, 1)
        one = time.perf_counter() - t0
        extra = max(0, int(args.cpu_seconds / 2 / max(one, 1e-6)) - 1)
        t1 = time.perf_counter()
        if extra:
            C.rk4(hc, uc, vc, dt, extra)
        total = one + time.perf_counter() - t1
        threads = int(oracle_lib().ora_num_threads())
        out["cpu_baseline"] = {"value": cells * (1 + extra) / total, "unit": "cell-steps/s", "cores": threads,
                               "kind": "port", "sample": f"C{N} C-grid RK4, {1 + extra} steps ({total:.1f} s) "
                                                         f"of the oracle (C restatement, OpenMP over {threads} threads)"}
    op.close()
    return out


def batched_form(args, stream):
    """One launch per RK stage over all members (hipGraph per step) vs the plain
    single-cube path run member after member; same states bit for bit."""
    import numpy as np
    import torch
    import cubedsphere_sbp_swm_amd as swm
    a, g = swm.EARTH["radius"], swm.EARTH["grav"]
    sizes = [int(v) for v in args.ens_sizes.split(",")]
    Ns = [N for N in sizes for _ in range(args.ens_per_size)]
    steps = args.ens_steps
    dts = [0.5 * a * (np.pi / 2) / N / (np.sqrt(g * 1.0e4) + 40.0) for N in Ns]
    ops, xs, ics = [], [], []
    for k, N in enumerate(Ns):
        op = swm.create_swm_operator(N, **swm.operator_names(42), scale=a, omega=swm.EARTH["omega"], grav=g)
        ic = op.initial_state(seed=k, lmax=16)
        x = swm.create_stvec_swm(op)
        x.upload(*ic)
        ops.append(op)
        xs.append(x)
        ics.append(ic)
    ops[0].set_stream(stream.cuda_stream)
    ens = swm.Ensemble(ops, xs)
    tuned_us = ens.autotune(reps=10)
    for x, ic in zip(xs, ics):
        x.upload(*ic)

    def ev():
        return torch.cuda.Event(enable_timing=True)

    ens.steps(dts, 2)  # tables, graph capture
    for x, ic in zip(xs, ics):
        x.upload(*ic)
    torch.cuda.synchronize()
    e0, e1 = ev(), ev()
    e0.record(stream)
    ens.steps(dts, steps)
    e1.record(stream)
    torch.cuda.synchronize()
    t_ens = e0.elapsed_time(e1) / 1e3
    finals = [x.download() for x in xs]
    # dominant kernel: the ensemble stage-2/3 launch, after 40 untimed launches (about 10 ms),
    # which bring the clocks back up after the host-side downloads left the GPU idle
    ens.bench_stage(40)
    torch.cuda.synchronize()
    k0, k1 = ev(), ev()
    k0.record(stream)
    reps = 40
    ens.bench_stage(reps)
    k1.record(stream)
    torch.cuda.synchronize()
    k_us = k0.elapsed_time(k1) * 1e3 / reps
    points = sum(6 * (N + 1) ** 2 for N in Ns)
    cells = sum(6 * N * N for N in Ns)
    stage_bytes = STAGE23_VALUES * 8 * points
    # plain path: the same members one after another (each its own autotuned tiling, graph)
    for op, x, ic in zip(ops, xs, ics):
        op.set_stream(stream.cuda_stream)
        x.upload(*ic)
        op.autotune(x, reps=5)
        x.upload(*ic)
    for op, x, ic, dt in zip(ops, xs, ics, dts):
        swm.create_timescheme(op, "rk4").steps(x, dt, 1)  # graph capture for this state and dt
        x.upload(*ic)
    torch.cuda.synchronize()
    p0, p1 = ev(), ev()
    p0.record(stream)
    for op, x, dt in zip(ops, xs, dts):
        swm.create_timescheme(op, "rk4").steps(x, dt, steps)
    p1.record(stream)
    torch.cuda.synchronize()
    t_plain = p0.elapsed_time(p1) / 1e3
    same = all(np.array_equal(f, q) for fin, x in zip(finals, xs) for f, q in zip(fin, x.download()))
    tiling = ens.get_tiling()
    ens.close()
    return {"workload": f"{len(Ns)} members ({args.ens_per_size} each at C{', C'.join(map(str, sizes))}), sbp42 FP64, "
                        f"RK4 x {steps} steps, dt per member at Courant 0.5, degree <= 16, seed = member id",
            "cell_steps_per_s": cells * steps / t_ens, "ms_per_step": t_ens * 1e3 / steps,
            "plain_cell_steps_per_s": cells * steps / t_plain, "speedup_vs_plain_fp64": t_plain / t_ens,
            "same_states_as_plain_path": bool(same), "tiling": tiling,
            "autotune_stage_us": tuned_us,
            "roofline": {"kernel": "ens_stage_kernel<42, double, MODE_STAGE23> (all members)",
                         "kernel_avg_us": k_us, "algorithmic_bytes_per_launch": stage_bytes,
                         "achieved_GBs": stage_bytes / (k_us * 1e-6) / 1e9,
                         "frac": stage_bytes / (k_us * 1e-6) / 1e9 / HBM_PEAK_GBS}}


def copy_bandwidth(stream, mib):
    """Measured device-to-device copy rate (read + write bytes / time) of torch's copy kernel."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.empty(n, dtype=torch.float64, device="cuda:0").fill_(1.0)
    dst = torch.empty_like(src)
    with torch.cuda.stream(stream):
        for _ in range(3):
            dst.copy_(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        reps = 10
        for _ in range(reps):
            dst.copy_(src)
        e1.record(stream)
    torch.cuda.synchronize()
    secs = e0.elapsed_time(e1) / 1e3 / reps
    return {"bytes_per_copy": 2 * n * 8, "us_per_copy": secs * 1e6, "GBs": 2 * n * 8 / secs / 1e9}


def cpu_baseline(N, order, dt, lmax, budget_s, op):
    """The oracle (FP64 C restatement, point loops under OpenMP) on the host, bounded sample."""
    from oracle.oracle import Grid, lib as oracle_lib
    import cubedsphere_sbp_swm_amd as swm
    G = Grid(N, order, order, swm.EARTH["radius"], swm.EARTH["omega"], swm.EARTH["grav"])
    h, u, v = op.initial_state(seed=0, lmax=lmax, halo_width=G.hw)
    t0 = time.perf_counter()
    G.rk4(h, u, v, dt, 1)
    one = time.perf_counter() - t0
    steps = 1
    extra = max(0, int(budget_s / max(one, 1e-6)) - 1)
    if extra:
        t1 = time.perf_counter()
        G.rk4(h, u, v, dt, extra)
        steps += extra
        total = one + (time.perf_counter() - t1)
    else:
        total = one
    cells = 6 * N * N
    threads = int(oracle_lib().ora_num_threads())
    return {"value": cells * steps / total, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"C{N} sbp{order} FP64 RK4 dt={dt:g}s, {steps} steps ({total:.1f} s) of the oracle "
                      f"(C restatement, gcc -O2, OpenMP over {threads} threads) on the GPU box host"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the panel-sharded multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    import cubedsphere_sbp_swm_amd as swm
    from cubedsphere_sbp_swm_amd import _capi

    torch.cuda.init()
    dev = torch.device("cuda:0")
    N, order, dt = args.N, args.order, args.dt
    op = swm.create_swm_operator(N, **swm.operator_names(order), scale=swm.EARTH["radius"],
                                 omega=swm.EARTH["omega"], grav=swm.EARTH["grav"], precision=args.precision)
    stream = torch.cuda.Stream(dev)  # default priority; the library's kernels and the events share it
    op.set_stream(stream.cuda_stream)
    h, u, v = op.initial_state(seed=0, lmax=args.lmax)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v)
    if args.tiling == "auto":
        tuned_us = op.autotune(x, reps=args.autotune_reps)
    else:
        shape, _, otf = args.tiling.partition(",")
        cols, _, rows = shape.partition("x")
        op.set_tiling(int(rows), otf == "otf", tile_cols=int(cols))
        tuned_us = None
    tiling = op.get_tiling()
    ts = swm.create_timescheme(op, "rk4")
    ts.steps(x, dt, args.warmup)
    torch.cuda.synchronize()

    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    ts.steps(x, dt, args.steps)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end)
    cells = 6 * N * N
    points = 6 * (N + 1) ** 2
    value = cells * args.steps / (ms / 1e3)
    diag = op.get_diagnostics(x)

    # dominant kernel: RK stage 2/3, back-to-back launches timed with events on the same stream
    y = swm.create_stvec_swm(op)
    y.upload(h, u, v)
    L = _capi.lib()
    import ctypes
    _capi.check(L.swm_bench_stage(op.ctx, ctypes.c_void_p(y.ptr), 2, 5))
    torch.cuda.synchronize()
    k0, k1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    k0.record(stream)
    _capi.check(L.swm_bench_stage(op.ctx, ctypes.c_void_p(y.ptr), 2, args.kernel_reps))
    k1.record(stream)
    torch.cuda.synchronize()
    k_us = k0.elapsed_time(k1) * 1e3 / args.kernel_reps
    eb = op.elem_bytes
    stage_bytes = STAGE23_VALUES * eb * points
    achieved = stage_bytes / (k_us * 1e-6) / 1e9
    step_bytes = STEP_VALUES * eb * points
    step_gbs = step_bytes / ((ms / 1e3) / args.steps) / 1e9

    forms = {}
    wanted = [f for f in args.forms.split(",") if f]
    if "fp32" in wanted:
        forms["fp32"] = fp32_form(args, stream)
    if "adjoint" in wanted:
        forms["adjoint"] = adjoint_form(args, stream)
    if "batched" in wanted:
        forms["batched"] = batched_form(args, stream)
    if "cgrid" in wanted and args.precision == 64:
        forms["cgrid"] = cgrid_form(args, stream)
    if "hyperdiffusion" in wanted and args.precision == 64:
        forms["hyperdiffusion"] = hyperdiffusion_form(args, stream, op, y)
        forms["hyperdiffusion"]["rk4_step_us"] = ms * 1e3 / args.steps
    copy = copy_bandwidth(stream, args.copy_mib) if args.copy_mib > 0 else None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(N, order, dt, args.lmax, args.cpu_seconds, op)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 64 else "f32",
        "data": "synthetic (seeded spherical-harmonic initial state)",
        "config": {"workload": f"C{N} sbp{order} RK4 dt={dt:g}s, {'FP64' if args.precision == 64 else 'FP32'}, "
                               f"one RK4 step per step", "N": N, "points": points, "cells": cells,
                   "parallelism": "single GPU, all six panels resident", "dt": dt,
                   "tiling": tiling, "autotune_stage_us": tuned_us},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(N, order, args.precision, tiling),
                     "kernel": "stage_kernel<%d, %s, MODE_STAGE23, %d, %d, %s> (RK stages 2 and 3)" % (
                         order, "double" if args.precision == 64 else "float", tiling["tile_cols"],
                         tiling["tile_rows"], str(tiling["metric_on_the_fly"]).lower()),
                     "kernel_avg_us": k_us, "algorithmic_bytes_per_launch": stage_bytes,
                     "step_effective_GBs": step_gbs, "step_frac": step_gbs / HBM_PEAK_GBS,
                     "measured_copy": copy},
        "cpu_baseline": cpu,
        "adjoint_forward_time_ratio": forms.get("adjoint", {}).get("adjoint_forward_time_ratio"),
        "forms": forms,
        "diagnostics_after": {k: diag[k] for k in ("mass", "TE", "hmin", "hmax")},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
