! Shallow-water run driver on the device path, reading the reference's namelist groups.
!
! Usage: swm_gpu_driver [namelist_file]   (default "namelist_swm", as the reference's mains read it,
! models/shallow_water/test/ts2/ts2_mod.f90:69).
!
! Groups read: &domain and &metric (domain/config_domain_mod.f90:41, metric/config_metric_mod.f90:51),
! &shallow_water_model (models/shallow_water/config/config_swm_mod.f90:70), all with the reference's
! variable names, plus &swm_gpu_run for what the reference takes from its test-case modules: the
! seeded spherical-harmonic initial state, gravity, arithmetic precision and the output file.
! Settings the device path does not implement (another staggering, topology, metric, rotated
! frame, operator family, or diffusion other than the sbp21 / sbp42 / sbp63 narrow-stencil hyperdiffusion
! with explicit_Eul1) stop the run with a message.
!
! Loop as in run_ts2 (ts2_mod.f90:141-156): an rk4 step, then, when a diffusion coefficient is
! nonzero, the explicit hyperdiffusion step (:145-146); key/value diagnostics every tau_diagnostics.
! The final state is written to output_file as raw FP64 h, u, v, each real(8) f(n, n, 6).  With
! latlon_output = .true. the run also writes h.dat, u.dat, v.dat as run_ts2 does (ts2_mod.f90:
! 108-112,134-138,174-179): REAL*4 records of the cubic (2N+1) x 4N lat-lon regrid, record 1 at
! the start and record it/nstep_write + 1 every nstep_write = nint(tau_write/dt) steps.
program swm_gpu_driver

use swm_gpu_mod
implicit none

! &domain
integer(kind=4)    :: N = 32, Nz = 1
character(len=255) :: staggering_type = "Ah", topology_type = "cube", metric_type = "ecs"
character(len=255) :: vertical_staggering = "None"
real(kind=8)       :: h_top = 1.0_8
! &metric
real(kind=8)       :: scale = 0.6371229e7_8, omega = 7.292115855e-5_8
real(kind=8)       :: rotation_matrix(3,3) = reshape([1,0,0, 0,1,0, 0,0,1], [3,3])
real(kind=8)       :: rotation_axis(3) = [0, 0, 1]
character(len=255) :: metric_2d_type = "ecs", vertical_transform_name = "vertical_transform_default"
! &shallow_water_model
character(len=255) :: swm_op_type = "vector_invariant", v_components_type = "covariant"
character(len=255) :: div_op_name = "divergence_ah42_sbp", grad_op_name = "gradient_ah42_sbp_ecs"
character(len=255) :: curl_op_name = "curl_divergence_ah42_sbp", coriolis_op_name = "coriolis_colocated"
character(len=255) :: KE_op_name = "KE_colocated", massflux_op_name = "massflux_colocated"
character(len=255) :: co2contra_op_name = "co2contra_colocated", vector_adv_op_name = "vector_advection_Ah"
character(len=255) :: quadrature_name = "SBP_Ah42_quadrature", diff_time_scheme = "explicit_Eul1"
character(len=255) :: hordiff_uv_name = "hordiff_empty", hordiff_h_name = "hordiff_empty"
real(kind=8)       :: uv_diff_coeff = 0.0_8, h_diff_coeff = 0.0_8
real(kind=8)       :: dt = 180.0_8, tau_write = 180.0_8, tau_diagnostics = 180.0_8
real(kind=8)       :: simulation_time_days = 0.0_8, simulation_time_hours = 0.0_8
real(kind=8)       :: simulation_time_min = 0.0_8, simulation_time_sec = 3600.0_8
! &swm_gpu_run
integer(kind=8)    :: seed = 0
integer(kind=4)    :: lmax = 8, precision = 64, autotune_reps = 0
real(kind=8)       :: h0 = 1.0e4_8, h_amp = 100.0_8, v_max = 40.0_8, grav = 9.80616_8
character(len=255) :: output_file = ""
logical            :: latlon_output = .false.

namelist /domain/ N, Nz, staggering_type, topology_type, metric_type, vertical_staggering, h_top
namelist /metric/ scale, omega, rotation_matrix, rotation_axis, metric_2d_type, vertical_transform_name
namelist /shallow_water_model/ swm_op_type, v_components_type, div_op_name, grad_op_name, curl_op_name, &
                               coriolis_op_name, KE_op_name, massflux_op_name, co2contra_op_name, &
                               vector_adv_op_name, quadrature_name, hordiff_uv_name, diff_time_scheme, &
                               uv_diff_coeff, hordiff_h_name, h_diff_coeff, dt, tau_write, tau_diagnostics, &
                               simulation_time_sec, simulation_time_min, simulation_time_hours, &
                               simulation_time_days
namelist /swm_gpu_run/ seed, lmax, h0, h_amp, v_max, grav, precision, autotune_reps, output_file, &
                       latlon_output

type(operator_swm_gpu_t) :: operator
type(stvec_swm_gpu_t)    :: state
type(rk4_gpu_t)          :: timescheme
type(operator_swm_diff_gpu_t) :: operator_diff
type(explicit_eul1_gpu_t)     :: timescheme_diff
logical                       :: use_diff
type(outputer_gpu_t)            :: outputer
type(latlon_vec_outputer_gpu_t) :: outputer_vec
type(key_value_r8_gpu_t) :: diagnostics
character(len=1024) :: fname
real(kind=8), allocatable :: h(:,:,:), u(:,:,:), v(:,:,:)
real(kind=8) :: simulation_time, best_us
integer(kind=4) :: nsteps, nstep_diagnostics, nstep_write, it, chunk, unit_nml, ios, n_pts, k

fname = "namelist_swm"
if (command_argument_count() >= 1) call get_command_argument(1, fname)
open(newunit=unit_nml, file=trim(fname), status="old", action="read", iostat=ios)
if (ios /= 0) call fail("cannot open namelist file "//trim(fname))
call read_group(1)
call read_group(2)
call read_group(3)
call read_group(4)
close(unit_nml)

if (trim(staggering_type) /= "Ah" .or. trim(topology_type) /= "cube" .or. trim(metric_type) /= "ecs" &
    .or. trim(metric_2d_type) /= "ecs") &
    call fail("device path implements staggering 'Ah', topology 'cube', metric 'ecs' only")
if (any(rotation_axis /= [0.0_8, 0.0_8, 1.0_8]) .or. &
    any(rotation_matrix /= reshape([1.0_8,0.0_8,0.0_8, 0.0_8,1.0_8,0.0_8, 0.0_8,0.0_8,1.0_8], [3,3]))) &
    call fail("device path implements the unrotated frame only")
if (trim(swm_op_type) /= "vector_invariant" .or. trim(v_components_type) /= "covariant" .or. &
    trim(coriolis_op_name) /= "coriolis_colocated" .or. trim(KE_op_name) /= "KE_colocated" .or. &
    trim(massflux_op_name) /= "massflux_colocated" .or. trim(co2contra_op_name) /= "co2contra_colocated") &
    call fail("device path implements the covariant vector-invariant colocated operator only")
use_diff = uv_diff_coeff /= 0.0_8 .or. h_diff_coeff /= 0.0_8
if (use_diff .and. trim(diff_time_scheme) /= "explicit_Eul1") &
    call fail("device path implements diff_time_scheme = explicit_Eul1 only")
if (use_diff .and. ((trim(hordiff_uv_name) /= "hordiff_vec_xyz_Ah_sbp_21_narrow" .and. &
                     trim(hordiff_uv_name) /= "hordiff_vec_xyz_Ah_sbp_42_narrow" .and. &
                     trim(hordiff_uv_name) /= "hordiff_vec_xyz_Ah_sbp_63_narrow") .or. &
                    (trim(hordiff_h_name) /= "hordiff_scalar_Ah_sbp_21_narrow" .and. &
                     trim(hordiff_h_name) /= "hordiff_scalar_Ah_sbp_42_narrow" .and. &
                     trim(hordiff_h_name) /= "hordiff_scalar_Ah_sbp_63_narrow"))) &
    call fail("device path implements the hyperdiffusion names hordiff_vec_xyz_Ah_sbp_{21,42,63}_narrow / " // &
              "hordiff_scalar_Ah_sbp_{21,42,63}_narrow only")

simulation_time = simulation_time_days*86400.0_8 + simulation_time_hours*3600.0_8 + &
                  simulation_time_min*60.0_8 + simulation_time_sec
nsteps = int(simulation_time/dt)
nstep_diagnostics = max(1, nint(tau_diagnostics/dt))
nstep_write = max(1, nint(tau_write/dt))

print '(A,I6,A,A,A,A,A,A,A,I3)', "N = ", N, "  div = ", trim(div_op_name), "  grad = ", trim(grad_op_name), &
      "  quadrature = ", trim(quadrature_name), "  precision = ", precision
print '(A,ES12.4,A,I10)', "dt = ", dt, "  steps = ", nsteps
print '(A,F8.4)', "Gravity Wave CFL = ", real(dt*sqrt(h0*grav)/(2.0_8*acos(-1.0_8)/4/N)/scale, 4)

call create_swm_gpu_operator(operator, N, div_op_name, grad_op_name, curl_op_name, quadrature_name, &
                             scale, omega, grav, precision)
n_pts = operator%n_pts
allocate(h(n_pts, n_pts, 6), u(n_pts, n_pts, 6), v(n_pts, n_pts, 6))
call operator%initial_state(seed, lmax, h0, h_amp, v_max, h, u, v, 0)
call create_stvec_swm_gpu(state, operator)
call state%upload(h, u, v, 0)
call create_rk4_gpu(timescheme)
if (use_diff) then
    call create_swm_diff_gpu_operator(operator_diff, operator, hordiff_uv_name, uv_diff_coeff, &
                                      hordiff_h_name, h_diff_coeff)
    call create_explicit_eul1_gpu(timescheme_diff)
end if
if (autotune_reps > 0) then
    best_us = operator%autotune(state, autotune_reps)
    print '(A,F10.2,A)', "autotuned stage kernel: ", best_us, " us"
    call state%upload(h, u, v, 0)
end if

diagnostics = operator%get_diagnostics(state)
call diagnostics%print()
if (latlon_output) then
    ! lat-lon output as run_ts2 writes it (ts2_mod.f90:108-112,134-138,174-179)
    call create_latlon_outputer_gpu(outputer, operator, 2*N+1, 4*N, "Ah")
    call create_latlon_vec_outputer_gpu(outputer_vec, operator, 2*N+1, 4*N, "Ah", v_components_type)
    call outputer%write(state, "h", "h.dat", 1)
    call outputer_vec%write(state, "u.dat", "v.dat", 1)
end if
it = 0
do while (it < nsteps)
    chunk = min(nstep_diagnostics - mod(it, nstep_diagnostics), nsteps - it)
    if (latlon_output) chunk = min(chunk, nstep_write - mod(it, nstep_write))
    if (use_diff) then
        do k = 1, chunk
            call timescheme%step(state, operator, dt)
            call timescheme_diff%step(state, operator_diff, dt)
        end do
    else
        call timescheme%steps(state, operator, dt, chunk)
    end if
    it = it + chunk
    if (mod(it, nstep_diagnostics) == 0 .or. it == nsteps) then
        diagnostics = operator%get_diagnostics(state)
        print '(A,F12.4)', "Hours = ", real(it*dt/3600.0_8, 4)
        call diagnostics%print()
    end if
    if (latlon_output .and. mod(it, nstep_write) == 0) then
        call outputer%write(state, "h", "h.dat", it/nstep_write + 1)
        call outputer_vec%write(state, "u.dat", "v.dat", it/nstep_write + 1)
    end if
end do
if (latlon_output) then
    call outputer%destroy()
    call outputer_vec%destroy()
end if

if (len_trim(output_file) > 0) then
    call state%download(h, u, v, 0)
    open(newunit=unit_nml, file=trim(output_file), access="stream", form="unformatted", &
         status="replace", action="write", iostat=ios)
    if (ios /= 0) call fail("cannot open output file "//trim(output_file))
    write(unit_nml) h, u, v
    close(unit_nml)
    print '(A,A)', "wrote ", trim(output_file)
end if

call state%free()
call operator%destroy()

contains

subroutine read_group(which)
    integer, intent(in) :: which
    rewind(unit_nml)
    select case (which)
    case (1)
        read(unit_nml, nml=domain, iostat=ios)
    case (2)
        read(unit_nml, nml=metric, iostat=ios)
    case (3)
        read(unit_nml, nml=shallow_water_model, iostat=ios)
    case (4)
        read(unit_nml, nml=swm_gpu_run, iostat=ios)
    end select
    ! a missing group keeps the defaults; a malformed one is an error
    if (ios > 0) call fail("malformed namelist group in "//trim(fname))
end subroutine read_group

subroutine fail(msg)
    character(len=*), intent(in) :: msg
    print '(A,A)', "swm_gpu_driver: ", msg
    error stop 1
end subroutine fail

end program swm_gpu_driver
