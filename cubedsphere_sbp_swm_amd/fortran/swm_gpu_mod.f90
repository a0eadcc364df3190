! Fortran host binding of the MI355X shallow-water path (C ABI: include/swm_gpu.h).
!
! The types below mirror the reference's abstract interfaces for the Ah-grid SWM hot path
! (operator_mod.f90:11-25 operator_t%apply / get_diagnostics, stvec_mod.f90:10-28
! stvec_t%assign_s1v1s2v2 / update_s1v1s2v2 / create_similar, timescheme_mod.f90:9-24
! timescheme_t%step) with the device context in place of domain_t.  A negative status from
! the library stops the program with the library's message, as parcomm_global%abort does
! in the reference (parallel/parcomm_mod.f90:73-83).
!
! Host arrays: real(8) :: f(1-hw:n+hw, 1-hw:n+hw, 6), n = N+1 (i fastest, panels last).
module swm_gpu_mod

use iso_c_binding
implicit none

private
public :: operator_swm_gpu_t, stvec_swm_gpu_t, rk4_gpu_t, key_value_r8_gpu_t
public :: create_swm_gpu_operator, create_stvec_swm_gpu, create_rk4_gpu, swm_gpu_check
public :: operator_swm_diff_gpu_t, explicit_eul1_gpu_t, create_swm_diff_gpu_operator, create_explicit_eul1_gpu
public :: outputer_gpu_t, latlon_vec_outputer_gpu_t, create_latlon_outputer_gpu, create_latlon_vec_outputer_gpu
public :: create_master_paneled_outputer_gpu

interface
    function c_swm_last_error() bind(C, name="swm_last_error") result(p)
        import :: c_ptr
        type(c_ptr) :: p
    end function
    integer(c_int) function c_swm_ctx_create(ctx, N, div, grad, curl, quad, scale, omega, grav, prec) &
                                             bind(C, name="swm_ctx_create")
        import :: c_ptr, c_int, c_char, c_double
        type(c_ptr), intent(out) :: ctx
        integer(c_int), value :: N, prec
        character(kind=c_char), intent(in) :: div(*), grad(*), curl(*), quad(*)
        real(c_double), value :: scale, omega, grav
    end function
    integer(c_int) function c_swm_ctx_destroy(ctx) bind(C, name="swm_ctx_destroy")
        import :: c_ptr, c_int
        type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_swm_ctx_info(ctx, n, pitch, fe, eb) bind(C, name="swm_ctx_info")
        import :: c_ptr, c_int, c_long_long
        type(c_ptr), value :: ctx
        integer(c_int), intent(out) :: n, pitch, eb
        integer(c_long_long), intent(out) :: fe
    end function
    integer(c_int) function c_swm_ctx_synchronize(ctx) bind(C, name="swm_ctx_synchronize")
        import :: c_ptr, c_int
        type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_swm_ctx_autotune(ctx, state, reps, best_us) bind(C, name="swm_ctx_autotune")
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx, state
        integer(c_int), value :: reps
        real(c_double), intent(out) :: best_us
    end function
    integer(c_int) function c_swm_state_alloc(ctx, state) bind(C, name="swm_state_alloc")
        import :: c_ptr, c_int
        type(c_ptr), value :: ctx
        type(c_ptr), intent(out) :: state
    end function
    integer(c_int) function c_swm_state_free(ctx, state) bind(C, name="swm_state_free")
        import :: c_ptr, c_int
        type(c_ptr), value :: ctx, state
    end function
    integer(c_int) function c_swm_state_upload(ctx, state, h, u, v, hw) bind(C, name="swm_state_upload")
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx, state
        real(c_double), intent(in) :: h(*), u(*), v(*)
        integer(c_int), value :: hw
    end function
    integer(c_int) function c_swm_state_download(ctx, state, h, u, v, hw) bind(C, name="swm_state_download")
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx, state
        real(c_double), intent(out) :: h(*), u(*), v(*)
        integer(c_int), value :: hw
    end function
    integer(c_int) function c_swm_rhs(ctx, vin, vout) bind(C, name="swm_rhs")
        import :: c_ptr, c_int
        type(c_ptr), value :: ctx, vin, vout
    end function
    integer(c_int) function c_swm_assign(ctx, out, s1, v1, s2, v2) bind(C, name="swm_assign")
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx, out, v1, v2
        real(c_double), value :: s1, s2
    end function
    integer(c_int) function c_swm_update(ctx, out, s1, v1, s2, v2) bind(C, name="swm_update")
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx, out, v1, v2
        real(c_double), value :: s1, s2
    end function
    integer(c_int) function c_swm_rk4_steps(ctx, state, dt, nsteps) bind(C, name="swm_rk4_steps")
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx, state
        real(c_double), value :: dt
        integer(c_int), value :: nsteps
    end function
    integer(c_int) function c_swm_diagnostics(ctx, state, out) bind(C, name="swm_diagnostics")
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx, state
        real(c_double), intent(out) :: out(7)
    end function
    integer(c_int) function c_swm_ic_sh(ctx, seed, lmax, h0, hamp, vmax, h, u, v, hw) &
                                        bind(C, name="swm_ic_spherical_harmonics")
        import :: c_ptr, c_int, c_double, c_long_long
        type(c_ptr), value :: ctx
        integer(c_long_long), value :: seed
        integer(c_int), value :: lmax, hw
        real(c_double), value :: h0, hamp, vmax
        real(c_double), intent(out) :: h(*), u(*), v(*)
    end function
    integer(c_int) function c_swm_diff_apply(ctx, vin, vout, uv_name, uv_coeff, h_name, h_coeff) &
                                              bind(C, name="swm_diff_apply")
        import :: c_ptr, c_int, c_char, c_double
        type(c_ptr), value :: ctx, vin, vout
        character(kind=c_char), intent(in) :: uv_name(*), h_name(*)
        real(c_double), value :: uv_coeff, h_coeff
    end function
    integer(c_int) function c_swm_diff_step(ctx, state, dt, uv_name, uv_coeff, h_name, h_coeff) &
                                             bind(C, name="swm_diff_step")
        import :: c_ptr, c_int, c_char, c_double
        type(c_ptr), value :: ctx, state
        real(c_double), value :: dt
        character(kind=c_char), intent(in) :: uv_name(*), h_name(*)
        real(c_double), value :: uv_coeff, h_coeff
    end function
    integer(c_int) function c_swm_latlon_outputer_create(ctx, nlat, nlon, grid, out) &
                                                          bind(C, name="swm_latlon_outputer_create")
        import :: c_ptr, c_int, c_char
        type(c_ptr), value :: ctx
        integer(c_int), value :: nlat, nlon
        character(kind=c_char), intent(in) :: grid(*)
        type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function c_swm_latlon_vec_outputer_create(ctx, nlat, nlon, stag, comp, out) &
                                                              bind(C, name="swm_latlon_vec_outputer_create")
        import :: c_ptr, c_int, c_char
        type(c_ptr), value :: ctx
        integer(c_int), value :: nlat, nlon
        character(kind=c_char), intent(in) :: stag(*), comp(*)
        type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function c_swm_outputer_destroy(o) bind(C, name="swm_outputer_destroy")
        import :: c_ptr, c_int
        type(c_ptr), value :: o
    end function
    integer(c_int) function c_swm_latlon_write(o, f, file_name, rec) bind(C, name="swm_latlon_write")
        import :: c_ptr, c_int, c_char
        type(c_ptr), value :: o, f
        character(kind=c_char), intent(in) :: file_name(*)
        integer(c_int), value :: rec
    end function
    integer(c_int) function c_swm_latlon_vec_write(o, u, v, file_u, file_v, rec) bind(C, name="swm_latlon_vec_write")
        import :: c_ptr, c_int, c_char
        type(c_ptr), value :: o, u, v
        character(kind=c_char), intent(in) :: file_u(*), file_v(*)
        integer(c_int), value :: rec
    end function
    integer(c_int) function c_swm_paneled_write(ctx, f, file_name, rec) bind(C, name="swm_paneled_write")
        import :: c_ptr, c_int, c_char
        type(c_ptr), value :: ctx, f
        character(kind=c_char), intent(in) :: file_name(*)
        integer(c_int), value :: rec
    end function
    function c_strlen(s) bind(C, name="strlen") result(n)
        import :: c_ptr, c_size_t
        type(c_ptr), value :: s
        integer(c_size_t) :: n
    end function
end interface

! key/value diagnostics (as stuff/key_value_mod.f90:9-43)
type key_value_r8_gpu_t
    character(len=16) :: keys(7)
    real(kind=8)      :: values(7)
contains
    procedure, public :: print => print_key_value
end type key_value_r8_gpu_t

type operator_swm_gpu_t
    type(c_ptr)     :: ctx = c_null_ptr
    integer(kind=4) :: N = 0, n_pts = 0, precision = 64
contains
    procedure, public :: apply => apply_swm_gpu
    procedure, public :: get_diagnostics => get_diagnostics_swm_gpu
    procedure, public :: autotune => autotune_swm_gpu
    procedure, public :: initial_state => initial_state_swm_gpu
    procedure, public :: destroy => destroy_swm_gpu
end type operator_swm_gpu_t

type stvec_swm_gpu_t
    type(c_ptr) :: ctx = c_null_ptr
    type(c_ptr) :: state = c_null_ptr
contains
    procedure, public :: assign_s1v1s2v2 => assign_s1v1s2v2_gpu
    procedure, public :: update_s1v1s2v2 => update_s1v1s2v2_gpu
    procedure, public :: create_similar => create_similar_gpu
    procedure, public :: upload => upload_gpu
    procedure, public :: download => download_gpu
    procedure, public :: free => free_gpu
end type stvec_swm_gpu_t

! operator_swm_diff_t (operator_swm_diff_mod.f90:14-21): hyperdiffusion tendency operator
type operator_swm_diff_gpu_t
    type(c_ptr)         :: ctx = c_null_ptr
    character(len=255)  :: hordiff_uv_name = "", hordiff_h_name = ""
    real(kind=8)        :: uv_diff_coeff = 0.0_8, h_diff_coeff = 0.0_8
contains
    procedure, public :: apply => apply_swm_diff_gpu
end type operator_swm_diff_gpu_t

! explicit_Eul1_t (time_schemes/explicit_Eul1_mod.f90:12-15)
type explicit_eul1_gpu_t
contains
    procedure, public :: step => step_explicit_eul1_gpu
end type explicit_eul1_gpu_t

! latlon_outputer_t (outputer/latlon_outputer_mod.f90:16-24) and master_paneled_outputer_t
! (master_paneled_outputer_mod.f90:11-21): write(state, field, file_name, rec_num) with field
! 'h', 'u' or 'v' of the state in place of a grid_field_t
type outputer_gpu_t
    type(c_ptr) :: ctx = c_null_ptr, handle = c_null_ptr
    logical     :: paneled = .false.
contains
    procedure, public :: write => write_outputer_gpu
    procedure, public :: destroy => destroy_outputer_gpu
end type outputer_gpu_t

! latlon_vec_outputer_t (latlon_outputer_mod.f90:26-38): write(state, file_name_u, file_name_v, rec_num)
type latlon_vec_outputer_gpu_t
    type(c_ptr) :: ctx = c_null_ptr, handle = c_null_ptr
contains
    procedure, public :: write => write_latlon_vec_gpu
    procedure, public :: destroy => destroy_latlon_vec_gpu
end type latlon_vec_outputer_gpu_t

type rk4_gpu_t
contains
    procedure, public :: step => step_rk4_gpu
    procedure, public :: steps => steps_rk4_gpu
end type rk4_gpu_t

contains

subroutine swm_gpu_check(ierr, where)
    integer(c_int),   intent(in) :: ierr
    character(len=*), intent(in) :: where
    type(c_ptr) :: p
    character(kind=c_char), pointer :: msg(:)
    integer :: k, l
    character(len=512) :: text
    if (ierr == 0) return
    p = c_swm_last_error()
    text = ""
    if (c_associated(p)) then
        l = int(c_strlen(p))
        call c_f_pointer(p, msg, [l])
        do k = 1, min(l, 512)
            text(k:k) = msg(k)
        end do
    end if
    print *, "swm_gpu error in ", where, " (status ", ierr, "): ", trim(text)
    error stop 1
end subroutine swm_gpu_check

! create_domain + create_swm_operator (operator_swm_factory_mod.f90:43-131)
subroutine create_swm_gpu_operator(op, N, div_op_name, grad_op_name, curl_op_name, quadrature_name, &
                                   scale, omega, grav, precision)
    type(operator_swm_gpu_t), intent(out) :: op
    integer(kind=4),  intent(in) :: N, precision
    character(len=*), intent(in) :: div_op_name, grad_op_name, curl_op_name, quadrature_name
    real(kind=8),     intent(in) :: scale, omega, grav
    integer(c_int) :: n_, pitch, eb
    integer(c_long_long) :: fe
    call swm_gpu_check(c_swm_ctx_create(op%ctx, int(N, c_int), trim(div_op_name)//c_null_char, &
                                        trim(grad_op_name)//c_null_char, trim(curl_op_name)//c_null_char, &
                                        trim(quadrature_name)//c_null_char, scale, omega, grav, &
                                        int(precision, c_int)), "create_swm_gpu_operator")
    call swm_gpu_check(c_swm_ctx_info(op%ctx, n_, pitch, fe, eb), "create_swm_gpu_operator")
    op%N = N
    op%n_pts = n_
    op%precision = precision
end subroutine create_swm_gpu_operator

subroutine destroy_swm_gpu(this)
    class(operator_swm_gpu_t), intent(inout) :: this
    if (c_associated(this%ctx)) call swm_gpu_check(c_swm_ctx_destroy(this%ctx), "destroy")
    this%ctx = c_null_ptr
end subroutine destroy_swm_gpu

! create_stvec_swm (stvec_swm_factory_mod.f90:7-28)
subroutine create_stvec_swm_gpu(v, op)
    type(stvec_swm_gpu_t),    intent(out) :: v
    type(operator_swm_gpu_t), intent(in)  :: op
    v%ctx = op%ctx
    call swm_gpu_check(c_swm_state_alloc(op%ctx, v%state), "create_stvec_swm_gpu")
end subroutine create_stvec_swm_gpu

subroutine create_rk4_gpu(ts)
    type(rk4_gpu_t), intent(out) :: ts
end subroutine create_rk4_gpu

! operator_t%apply: vout = RHS(vin)
subroutine apply_swm_gpu(this, vout, vin)
    class(operator_swm_gpu_t), intent(inout) :: this
    class(stvec_swm_gpu_t),    intent(inout) :: vout, vin
    call swm_gpu_check(c_swm_rhs(this%ctx, vin%state, vout%state), "operator%apply")
end subroutine apply_swm_gpu

! operator_t%get_diagnostics (operator_swm_mod.f90:115-155)
function get_diagnostics_swm_gpu(this, v) result(kv)
    class(operator_swm_gpu_t), intent(inout) :: this
    class(stvec_swm_gpu_t),    intent(inout) :: v
    type(key_value_r8_gpu_t) :: kv
    real(c_double) :: out(7)
    call swm_gpu_check(c_swm_diagnostics(this%ctx, v%state, out), "operator%get_diagnostics")
    kv%keys = [character(len=16) :: "hmin", "hmax", "mass", "TE", "KE", "PE", "Enstrophy"]
    kv%values = out
end function get_diagnostics_swm_gpu

function autotune_swm_gpu(this, v, reps) result(best_us)
    class(operator_swm_gpu_t), intent(inout) :: this
    class(stvec_swm_gpu_t),    intent(inout) :: v
    integer(kind=4),           intent(in)    :: reps
    real(kind=8) :: best_us
    call swm_gpu_check(c_swm_ctx_autotune(this%ctx, v%state, int(reps, c_int), best_us), "autotune")
end function autotune_swm_gpu

subroutine initial_state_swm_gpu(this, seed, lmax, h0, h_amp, v_max, h, u, v, halo_width)
    class(operator_swm_gpu_t), intent(inout) :: this
    integer(kind=8), intent(in) :: seed
    integer(kind=4), intent(in) :: lmax, halo_width
    real(kind=8),    intent(in) :: h0, h_amp, v_max
    real(kind=8),    intent(out) :: h(:,:,:), u(:,:,:), v(:,:,:)
    call swm_gpu_check(c_swm_ic_sh(this%ctx, int(seed, c_long_long), int(lmax, c_int), h0, h_amp, v_max, &
                                   h, u, v, int(halo_width, c_int)), "initial_state")
end subroutine initial_state_swm_gpu

! stvec_t%assign_s1v1s2v2: this = s1*v1 + s2*v2
subroutine assign_s1v1s2v2_gpu(this, scalar1, v1, scalar2, v2)
    class(stvec_swm_gpu_t), intent(inout) :: this
    real(kind=8),           intent(in)    :: scalar1, scalar2
    class(stvec_swm_gpu_t), intent(in)    :: v1, v2
    call swm_gpu_check(c_swm_assign(this%ctx, this%state, scalar1, v1%state, scalar2, v2%state), "assign")
end subroutine assign_s1v1s2v2_gpu

! stvec_t%update_s1v1s2v2: this = this + s1*v1 + s2*v2
subroutine update_s1v1s2v2_gpu(this, scalar1, v1, scalar2, v2)
    class(stvec_swm_gpu_t), intent(inout) :: this
    real(kind=8),           intent(in)    :: scalar1, scalar2
    class(stvec_swm_gpu_t), intent(in)    :: v1, v2
    call swm_gpu_check(c_swm_update(this%ctx, this%state, scalar1, v1%state, scalar2, v2%state), "update")
end subroutine update_s1v1s2v2_gpu

subroutine create_similar_gpu(this, destination)
    class(stvec_swm_gpu_t), intent(in)  :: this
    type(stvec_swm_gpu_t),  intent(out) :: destination
    destination%ctx = this%ctx
    call swm_gpu_check(c_swm_state_alloc(this%ctx, destination%state), "create_similar")
end subroutine create_similar_gpu

subroutine upload_gpu(this, h, u, v, halo_width)
    class(stvec_swm_gpu_t), intent(inout) :: this
    real(kind=8),           intent(in)    :: h(:,:,:), u(:,:,:), v(:,:,:)
    integer(kind=4),        intent(in)    :: halo_width
    call swm_gpu_check(c_swm_state_upload(this%ctx, this%state, h, u, v, int(halo_width, c_int)), "upload")
end subroutine upload_gpu

subroutine download_gpu(this, h, u, v, halo_width)
    class(stvec_swm_gpu_t), intent(inout) :: this
    real(kind=8),           intent(out)   :: h(:,:,:), u(:,:,:), v(:,:,:)
    integer(kind=4),        intent(in)    :: halo_width
    call swm_gpu_check(c_swm_state_download(this%ctx, this%state, h, u, v, int(halo_width, c_int)), "download")
end subroutine download_gpu

subroutine free_gpu(this)
    class(stvec_swm_gpu_t), intent(inout) :: this
    if (c_associated(this%state)) call swm_gpu_check(c_swm_state_free(this%ctx, this%state), "free")
    this%state = c_null_ptr
end subroutine free_gpu

! timescheme_t%step for rk4 (time_schemes/rk4_mod.f90:20-42)
subroutine step_rk4_gpu(this, v0, operator, dt)
    class(rk4_gpu_t),          intent(inout) :: this
    class(stvec_swm_gpu_t),    intent(inout) :: v0
    class(operator_swm_gpu_t), intent(inout) :: operator
    real(kind=8),              intent(in)    :: dt
    call swm_gpu_check(c_swm_rk4_steps(operator%ctx, v0%state, dt, 1_c_int), "rk4%step")
end subroutine step_rk4_gpu

subroutine steps_rk4_gpu(this, v0, operator, dt, nsteps)
    class(rk4_gpu_t),          intent(inout) :: this
    class(stvec_swm_gpu_t),    intent(inout) :: v0
    class(operator_swm_gpu_t), intent(inout) :: operator
    real(kind=8),              intent(in)    :: dt
    integer(kind=4),           intent(in)    :: nsteps
    call swm_gpu_check(c_swm_rk4_steps(operator%ctx, v0%state, dt, int(nsteps, c_int)), "rk4%steps")
end subroutine steps_rk4_gpu

! create_swm_diff_operator (operator_swm_diff_factory_mod.f90:11-34) for the names of the
! &shallow_water_model namelist; the device implements the narrow-stencil names
! "hordiff_vec_xyz_Ah_sbp_{21,42,63}_narrow" and "hordiff_scalar_Ah_sbp_{21,42,63}_narrow"
subroutine create_swm_diff_gpu_operator(op_diff, op, hordiff_uv_name, uv_diff_coeff, hordiff_h_name, h_diff_coeff)
    type(operator_swm_diff_gpu_t), intent(out) :: op_diff
    type(operator_swm_gpu_t),      intent(in)  :: op
    character(len=*),              intent(in)  :: hordiff_uv_name, hordiff_h_name
    real(kind=8),                  intent(in)  :: uv_diff_coeff, h_diff_coeff
    op_diff%ctx = op%ctx
    op_diff%hordiff_uv_name = hordiff_uv_name
    op_diff%hordiff_h_name = hordiff_h_name
    op_diff%uv_diff_coeff = uv_diff_coeff
    op_diff%h_diff_coeff = h_diff_coeff
end subroutine create_swm_diff_gpu_operator

! operator_swm_diff_t%apply (operator_swm_diff_mod.f90:25-46): vout = diffusion tendency of vin
subroutine apply_swm_diff_gpu(this, vout, vin)
    class(operator_swm_diff_gpu_t), intent(inout) :: this
    class(stvec_swm_gpu_t),         intent(inout) :: vout, vin
    call swm_gpu_check(c_swm_diff_apply(this%ctx, vin%state, vout%state, trim(this%hordiff_uv_name)//c_null_char, &
                                        this%uv_diff_coeff, trim(this%hordiff_h_name)//c_null_char, &
                                        this%h_diff_coeff), "operator_diff%apply")
end subroutine apply_swm_diff_gpu

subroutine create_explicit_eul1_gpu(ts)
    type(explicit_eul1_gpu_t), intent(out) :: ts
end subroutine create_explicit_eul1_gpu

! explicit_Eul1_t%step (explicit_Eul1_mod.f90:20-30): v0 = v0 + dt * operator(v0)
subroutine step_explicit_eul1_gpu(this, v0, operator, dt)
    class(explicit_eul1_gpu_t),     intent(inout) :: this
    class(stvec_swm_gpu_t),         intent(inout) :: v0
    class(operator_swm_diff_gpu_t), intent(inout) :: operator
    real(kind=8),                   intent(in)    :: dt
    call swm_gpu_check(c_swm_diff_step(operator%ctx, v0%state, dt, trim(operator%hordiff_uv_name)//c_null_char, &
                                       operator%uv_diff_coeff, trim(operator%hordiff_h_name)//c_null_char, &
                                       operator%h_diff_coeff), "timescheme_diff%step")
end subroutine step_explicit_eul1_gpu

! device pointer of field 'h', 'u' or 'v' of a state vector
function field_ptr(ctx, state, field) result(p)
    type(c_ptr),      intent(in) :: ctx, state
    character(len=*), intent(in) :: field
    type(c_ptr) :: p
    integer(c_int) :: n_, pitch, eb, k
    integer(c_long_long) :: fe
    call swm_gpu_check(c_swm_ctx_info(ctx, n_, pitch, fe, eb), "outputer%write")
    select case (field)
    case ("h"); k = 0
    case ("u"); k = 1
    case ("v"); k = 2
    case default
        k = -1
        call swm_gpu_check(-1_c_int, "outputer%write: field must be 'h', 'u' or 'v'")
    end select
    p = transfer(transfer(state, 0_c_intptr_t) + int(k, c_intptr_t) * int(fe, c_intptr_t) * int(eb, c_intptr_t), &
                 c_null_ptr)
end function field_ptr

! create_latlon_outputer(outputer, Nlat, Nlon, scalar_grid_type, domain)
! (outputer/outputer_factory_mod.f90:45-136); scalar_grid_type "Ah"
subroutine create_latlon_outputer_gpu(outputer, op, Nlat, Nlon, scalar_grid_type)
    type(outputer_gpu_t),     intent(out) :: outputer
    type(operator_swm_gpu_t), intent(in)  :: op
    integer(kind=4),          intent(in)  :: Nlat, Nlon
    character(len=*),         intent(in)  :: scalar_grid_type
    outputer%ctx = op%ctx
    call swm_gpu_check(c_swm_latlon_outputer_create(op%ctx, int(Nlat, c_int), int(Nlon, c_int), &
                                                    trim(scalar_grid_type)//c_null_char, outputer%handle), &
                       "create_latlon_outputer")
end subroutine create_latlon_outputer_gpu

! create_master_paneled_outputer(outputer, points_type, domain) (outputer_factory_mod.f90:11-43);
! points_type "xy"
subroutine create_master_paneled_outputer_gpu(outputer, op, points_type)
    type(outputer_gpu_t),     intent(out) :: outputer
    type(operator_swm_gpu_t), intent(in)  :: op
    character(len=*),         intent(in)  :: points_type
    if (trim(points_type) /= "xy") &
        call swm_gpu_check(-2_c_int, "create_master_paneled_outputer: the device path writes xy points only")
    outputer%ctx = op%ctx
    outputer%paneled = .true.
end subroutine create_master_paneled_outputer_gpu

! create_latlon_vec_outputer(outputer, Nlat, Nlon, staggering, components_type, domain)
! (outputer_factory_mod.f90:138-190); staggering "Ah", components_type "covariant"
subroutine create_latlon_vec_outputer_gpu(outputer, op, Nlat, Nlon, staggering, components_type)
    type(latlon_vec_outputer_gpu_t), intent(out) :: outputer
    type(operator_swm_gpu_t),        intent(in)  :: op
    integer(kind=4),                 intent(in)  :: Nlat, Nlon
    character(len=*),                intent(in)  :: staggering, components_type
    outputer%ctx = op%ctx
    call swm_gpu_check(c_swm_latlon_vec_outputer_create(op%ctx, int(Nlat, c_int), int(Nlon, c_int), &
                                                        trim(staggering)//c_null_char, &
                                                        trim(components_type)//c_null_char, outputer%handle), &
                       "create_latlon_vec_outputer")
end subroutine create_latlon_vec_outputer_gpu

! latlon_outputer_t%write / master_paneled_outputer_t%write (f, domain, file_name, rec_num)
subroutine write_outputer_gpu(this, state, field, file_name, rec_num)
    class(outputer_gpu_t),  intent(inout) :: this
    class(stvec_swm_gpu_t), intent(in)    :: state
    character(len=*),       intent(in)    :: field, file_name
    integer(kind=4),        intent(in)    :: rec_num
    if (this%paneled) then
        call swm_gpu_check(c_swm_paneled_write(this%ctx, field_ptr(this%ctx, state%state, field), &
                                               trim(file_name)//c_null_char, int(rec_num, c_int)), "outputer%write")
    else
        call swm_gpu_check(c_swm_latlon_write(this%handle, field_ptr(this%ctx, state%state, field), &
                                              trim(file_name)//c_null_char, int(rec_num, c_int)), "outputer%write")
    end if
end subroutine write_outputer_gpu

subroutine destroy_outputer_gpu(this)
    class(outputer_gpu_t), intent(inout) :: this
    integer(c_int) :: ierr
    if (c_associated(this%handle)) ierr = c_swm_outputer_destroy(this%handle)
    this%handle = c_null_ptr
end subroutine destroy_outputer_gpu

! latlon_vec_outputer_t%write(u, v, domain, file_name_u, file_name_v, rec_num)
subroutine write_latlon_vec_gpu(this, state, file_name_u, file_name_v, rec_num)
    class(latlon_vec_outputer_gpu_t), intent(inout) :: this
    class(stvec_swm_gpu_t),           intent(in)    :: state
    character(len=*),                 intent(in)    :: file_name_u, file_name_v
    integer(kind=4),                  intent(in)    :: rec_num
    call swm_gpu_check(c_swm_latlon_vec_write(this%handle, field_ptr(this%ctx, state%state, "u"), &
                                              field_ptr(this%ctx, state%state, "v"), &
                                              trim(file_name_u)//c_null_char, trim(file_name_v)//c_null_char, &
                                              int(rec_num, c_int)), "outputer_vec%write")
end subroutine write_latlon_vec_gpu

subroutine destroy_latlon_vec_gpu(this)
    class(latlon_vec_outputer_gpu_t), intent(inout) :: this
    integer(c_int) :: ierr
    if (c_associated(this%handle)) ierr = c_swm_outputer_destroy(this%handle)
    this%handle = c_null_ptr
end subroutine destroy_latlon_vec_gpu

subroutine print_key_value(this)
    class(key_value_r8_gpu_t), intent(in) :: this
    integer :: i
    character(len=1024) :: line
    character(len=64) :: item
    line = ""
    do i = 1, size(this%keys)
        write(item, "(A,' = ',ES24.16E3)") trim(this%keys(i)), this%values(i)
        line = trim(line) // " " // trim(item)
    end do
    print "(A)", trim(line)
end subroutine print_key_value

end module swm_gpu_mod
