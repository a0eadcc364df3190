"""Fortran host layer: the namelist driver (fortran/swm_gpu_driver.f90) over the ISO_C_BINDING
module reads the reference's &domain / &metric / &shallow_water_model groups, runs rk4 on the
device and writes the final state; the result must equal the Python host path bit for bit."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cubedsphere_sbp_swm_amd", "_build", "swm_gpu_driver")

# the reference's ts2_ah namelist layout (models/shallow_water/test/scripts/ts2_ah/ts2_run.sh)
NAMELIST = """
&domain
    N = {N},
    Nz = 1,
    topology_type    = 'cube',
    staggering_type  = '{stag}',
    metric_type      = 'ecs'
/
&metric
    scale = 0.6371229e7,
    omega = {omega!r}
/
&shallow_water_model
    swm_op_type       = 'vector_invariant'
    v_components_type = 'covariant'
    div_op_name       = '{div}',
    grad_op_name      = '{grad}',
    coriolis_op_name  = 'coriolis_colocated',
    curl_op_name      = '{curl}',
    KE_op_name        = 'KE_colocated',
    co2contra_op_name = 'co2contra_colocated',
    massflux_op_name  = 'massflux_colocated',
    quadrature_name   = '{quad}',
    diff_time_scheme  = 'explicit_Eul1'
    uv_diff_coeff     =  {uvc},
    hordiff_uv_name   = '{uvn}',
    h_diff_coeff      =  {hc},
    hordiff_h_name    = '{hn}',
    dt={dt},
    tau_write = {twrite},
    tau_diagnostics = {tau}
    simulation_time_days  = 0.0,
    simulation_time_hours = 0.0,
    simulation_time_min   = 0.0,
    simulation_time_sec   = {tsec},
/
&swm_gpu_run
    seed = {seed}, lmax = 8, grav = 9.80616, output_file = '{out}', latlon_output = {llo}
/
"""


def _need_driver():
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (run __graft_entry__.build())")


def _write_nml(tmp_path, N, order, dt, nsteps, seed, stag="Ah", uvc=0.0, hc=0.0,
               uvn="hordiff_empty", hn="hordiff_empty", twrite=86400.0, llo=".false."):
    import cubedsphere_sbp_swm_amd as swm
    names = swm.operator_names(order)
    out = tmp_path / "state.bin"
    text = NAMELIST.format(N=N, stag=stag, omega=swm.EARTH["omega"], div=names["div_op_name"],
                           grad=names["grad_op_name"], curl=names["curl_op_name"], quad=names["quadrature_name"],
                           dt=dt, tau=dt * 3, tsec=dt * nsteps, seed=seed, out=out, uvc=uvc, hc=hc, uvn=uvn,
                           hn=hn, twrite=twrite, llo=llo)
    p = tmp_path / "namelist_swm"
    p.write_text(text)
    return p, out


def test_driver_rejects_unsupported_staggering(tmp_path):
    """Settings outside the device path stop the driver before any device call."""
    _need_driver()
    nml, _ = _write_nml(tmp_path, 16, 42, 900.0, 2, 1, stag="C")
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "staggering" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("N,order,nsteps", [(24, 42, 7), (20, 63, 4), (16, 21, 5)])
def test_driver_equals_python_path(tmp_path, N, order, nsteps):
    _need_driver()
    import cubedsphere_sbp_swm_amd as swm
    dt, seed = 600.0, 5
    nml, out = _write_nml(tmp_path, N, order, dt, nsteps, seed)
    print(f"[fortran driver] N={N} order={order} steps={nsteps}", flush=True)
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    n = N + 1
    raw = np.fromfile(out, dtype=np.float64)
    assert raw.size == 3 * 6 * n * n
    fh, fu, fv = (a.reshape(6, n, n) for a in np.split(raw, 3))

    op = swm.create_swm_operator(N, **swm.operator_names(order), scale=swm.EARTH["radius"],
                                 omega=swm.EARTH["omega"], grav=swm.EARTH["grav"])
    h, u, v = op.initial_state(seed=seed, lmax=8)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v)
    swm.create_timescheme(op, "rk4").steps(x, dt, nsteps)
    ph, pu, pv = x.download()
    for a, b in ((fh, ph), (fu, pu), (fv, pv)):
        assert np.array_equal(a, b)

    # last diagnostics line printed by the driver = Python diagnostics of the same state
    d = op.get_diagnostics(x)
    last = [ln for ln in r.stdout.splitlines() if "mass =" in ln][-1]
    vals = dict((k, float(val)) for k, val in re.findall(r"(\w+) = +([-+0-9.Ee]+)", last))
    for k in ("hmin", "hmax", "mass", "TE", "KE", "PE", "Enstrophy"):
        assert vals[k] == pytest.approx(d[k], rel=1e-15, abs=0.0), k


@pytest.mark.gpu
@pytest.mark.parametrize("order", [21, 42, 63])
def test_driver_with_hyperdiffusion_equals_python_path(tmp_path, order):
    """&shallow_water_model with the sbp21 / sbp42 / sbp63 narrow-stencil hyperdiffusion (the operator
    family of the same order): the driver's RK4 + explicit_Eul1 loop equals the Python path step
    for step, bit for bit."""
    _need_driver()
    import cubedsphere_sbp_swm_amd as swm
    N, nsteps, dt, seed = 24, 5, 600.0, 3
    uvn, hn = f"hordiff_vec_xyz_Ah_sbp_{order}_narrow", f"hordiff_scalar_Ah_sbp_{order}_narrow"
    nml, out = _write_nml(tmp_path, N, order, dt, nsteps, seed, uvc=0.03, hc=0.01, uvn=uvn, hn=hn)
    print(f"[fortran driver + diffusion] N={N} order={order} steps={nsteps}", flush=True)
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    n = N + 1
    raw = np.fromfile(out, dtype=np.float64)
    fh, fu, fv = (a.reshape(6, n, n) for a in np.split(raw, 3))
    op = swm.create_swm_operator(N, **swm.operator_names(order), scale=swm.EARTH["radius"],
                                 omega=swm.EARTH["omega"], grav=swm.EARTH["grav"])
    h, u, v = op.initial_state(seed=seed, lmax=8)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v)
    ts = swm.create_timescheme(op, "rk4")
    for _ in range(nsteps):
        ts.steps(x, dt, 1)
        op.diff_step(x, dt, 0.03, 0.01, uvn, hn)
    for a, b in zip((fh, fu, fv), x.download()):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_driver_latlon_output_equals_python_outputers(tmp_path):
    """latlon_output = .true.: the driver writes h.dat, u.dat, v.dat as run_ts2 does (record 1 at
    the start, then one every tau_write), equal byte for byte to the Python outputers on the same
    states."""
    _need_driver()
    import cubedsphere_sbp_swm_amd as swm
    N, order, nsteps, dt, seed = 16, 42, 4, 600.0, 2
    nml, _ = _write_nml(tmp_path, N, order, dt, nsteps, seed, twrite=2 * dt, llo=".true.")
    print(f"[fortran driver latlon output] N={N} steps={nsteps}", flush=True)
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    op = swm.create_swm_operator(N, **swm.operator_names(order), scale=swm.EARTH["radius"],
                                 omega=swm.EARTH["omega"], grav=swm.EARTH["grav"])
    h, u, v = op.initial_state(seed=seed, lmax=8)
    x = swm.create_stvec_swm(op)
    x.upload(h, u, v)
    ref = tmp_path / "ref"
    ref.mkdir()
    out = swm.create_latlon_outputer(op, 2 * N + 1, 4 * N, "Ah")
    outv = swm.create_latlon_vec_outputer(op, 2 * N + 1, 4 * N, "Ah", "covariant")
    ts = swm.create_timescheme(op, "rk4")
    for rec in (1, 2, 3):
        if rec > 1:
            ts.steps(x, dt, 2)
        out.write(x.field("h"), str(ref / "h.dat"), rec)
        outv.write_vec(x, str(ref / "u.dat"), str(ref / "v.dat"), rec)
    for name in ("h.dat", "u.dat", "v.dat"):
        got, want = (tmp_path / name).read_bytes(), (ref / name).read_bytes()
        assert len(got) == 3 * 4 * (2 * N + 1) * 4 * N
        assert got == want, name


def test_driver_rejects_other_hyperdiffusion(tmp_path):
    _need_driver()
    nml, _ = _write_nml(tmp_path, 16, 42, 600.0, 2, 1, uvc=0.03, hc=0.01,
                        uvn="hordiff_c_biharm_curl_div", hn="hordiff_scalar_Ah_sbp_42_narrow")
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode != 0
    assert "hyperdiffusion" in r.stdout
