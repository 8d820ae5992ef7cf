! lbmib_mod.f90 -- iso_c_binding interface of liblbmib.so (include/lbmib.h)
! for Fortran callers.  A caller keeps the reference's main loop, input files
! and output formats and replaces the body of `call evolution(mesh,n,nforce,h)`
! (main.f90:359) with lbmib_step; see lbmib_driver.f90 for a complete program.
!
! Array conventions are the reference's: lattice fields are rho(nx,ny),
! populations f(nx,ny,9); the motion table is passed as motion(ncols,nrows)
! (one table row per column) and the force history comes back as
! rows(4+12*n_wings, n_rows), one force.dat row per column.
module lbmib
    use iso_c_binding
    implicit none
    private

    integer(c_int32_t), parameter, public :: LBMIB_OK = 0
    integer(c_int32_t), parameter, public :: LBMIB_E_ARG = -1
    integer(c_int32_t), parameter, public :: LBMIB_E_HIP = -2
    integer(c_int32_t), parameter, public :: LBMIB_E_GEOMETRY = -3
    integer(c_int32_t), parameter, public :: LBMIB_E_CAPACITY = -4
    integer(c_int32_t), parameter, public :: LBMIB_E_SOLVE = -5
    integer(c_int32_t), parameter, public :: LBMIB_E_STATE = -6
    integer(c_int32_t), parameter, public :: LBMIB_E_NAN = -7
    integer(c_int32_t), parameter, public :: LBMIB_PREC_FP64 = 0
    integer(c_int32_t), parameter, public :: LBMIB_PREC_FP32DEV = 1
    integer(c_int32_t), parameter, public :: LBMIB_NTIMINGS = 16
    integer(c_int32_t), parameter, public :: LBMIB_GEOMETRY_BATCH = 16
    integer(c_int32_t), parameter, public :: LBMIB_SOLVER_AUTO = 0
    integer(c_int32_t), parameter, public :: LBMIB_SOLVER_CHAIN = 1
    integer(c_int32_t), parameter, public :: LBMIB_SOLVER_SEPARATOR = 2

    ! struct lbmib_params (include/lbmib.h)
    type, bind(C), public :: lbmib_params
        integer(c_int32_t) :: nx = 0, ny = 0
        real(c_double) :: ox = 0.0d0, oy = 0.0d0, scale = 0.0d0, U = 0.0d0
        real(c_double) :: xc = 0.0d0, yc = 0.0d0, xp = 0.0d0, yp = 0.0d0
        real(c_double) :: s(9) = 0.0d0
        integer(c_int32_t) :: smooth = 1, precision = 0, n_wings = 1, pts_per_wing = 0
        integer(c_int32_t) :: device = 0, nan_guard = 0, ib_solver = 0
        integer(c_int32_t) :: reserved(5) = 0
    end type

    ! struct lbmib_dynamics (include/lbmib.h): rigid-body dynamics of wing 1,
    ! defaults as the reference's parameters (main.f90:76, 101-129)
    type, bind(C) :: lbmib_dynamics
        integer(c_int32_t) :: enabled = 0, rk4 = 1, closed_boundary = 1, reserved = 0
        real(c_double) :: solid_den = 10.0d0, area_airfoil = 6.807946476488794d-2
        real(c_double) :: par1 = 0.0d0, par2 = 0.0d0
        real(c_double) :: alphad = 10.0d0 * 3.141592653589793d0 / 180.0d0
    end type

    ! struct lbmib_launch_shape (include/lbmib.h): how steps are launched
    type, bind(C), public :: lbmib_launch_shape
        integer(c_int32_t) :: nodes_per_thread = 1, tile_order = 0, persistent = 0
        integer(c_int32_t) :: ib_cus = 16, graph = 0, ib_collide_shared = 0, ib_rhs_reserved = 1
        integer(c_int32_t) :: solve_unit = 0
        integer(c_int32_t) :: fuse_steps = 1
    end type

    public :: lbmib_create, lbmib_set_motion_table, lbmib_step, lbmib_steps_done, lbmib_max_steps
    public :: lbmib_get_forces, lbmib_get_history, lbmib_download, lbmib_upload_state
    public :: lbmib_residual, lbmib_ib_nodes, lbmib_set_profiling, lbmib_get_timings, lbmib_set_step
    public :: lbmib_markers, lbmib_vorticity, lbmib_surface_pressure, lbmib_surface_pressure_noalloc
    public :: lbmib_lattice_window
    ! rigid-body dynamics
    public :: lbmib_dynamics, lbmib_set_dynamics, lbmib_get_body_state, lbmib_set_body_state
    public :: lbmib_get_motion_rows, lbmib_get_edge_f0, lbmib_set_edge_f0
    public :: lbmib_destroy, lbmib_error_message, lbmib_version_string, lbmib_check
    ! batched form (many cases advanced together) and footprints
    public :: lbmib_set_shape, lbmib_get_shape, lbmib_device_bytes, lbmib_mem_info
    public :: lbmib_batch_create, lbmib_batch_add, lbmib_batch_remove, lbmib_batch_size, lbmib_batch_step
    public :: lbmib_batch_set_shape, lbmib_batch_get_shape, lbmib_batch_autotune, lbmib_batch_destroy

    interface
        integer(c_int) function lbmib_create(p, xb0, yb0, h) bind(C, name="lbmib_create")
            import :: c_int, c_double, c_ptr, lbmib_params
            type(lbmib_params), intent(in) :: p
            real(c_double), intent(in) :: xb0(*), yb0(*)
            type(c_ptr), intent(out) :: h
        end function

        integer(c_int) function lbmib_set_motion_table(h, rows, nrows, ncols) bind(C, name="lbmib_set_motion_table")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(in) :: rows(*)
            integer(c_int32_t), value :: nrows, ncols
        end function

        integer(c_int) function lbmib_step(h, n_steps) bind(C, name="lbmib_step")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: n_steps
        end function

        integer(c_int) function lbmib_steps_done(h, n) bind(C, name="lbmib_steps_done")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: n
        end function

        integer(c_int) function lbmib_max_steps(h, n) bind(C, name="lbmib_max_steps")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: n
        end function

        integer(c_int) function lbmib_get_forces(h, wing_force, wing_moment) bind(C, name="lbmib_get_forces")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: wing_force(*), wing_moment(*)
        end function

        integer(c_int) function lbmib_get_history(h, rows, max_rows, n_rows) bind(C, name="lbmib_get_history")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: rows(*)
            integer(c_int32_t), value :: max_rows
            integer(c_int32_t), intent(out) :: n_rows
        end function

        integer(c_int) function lbmib_download(h, rho, u, v, f) bind(C, name="lbmib_download")
            import :: c_int, c_ptr
            type(c_ptr), value :: h, rho, u, v, f
        end function

        integer(c_int) function lbmib_upload_state(h, f) bind(C, name="lbmib_upload_state")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(in) :: f(*)
        end function

        integer(c_int) function lbmib_markers(h, xb1, yb1) bind(C, name="lbmib_markers")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: xb1(*), yb1(*)
        end function

        integer(c_int) function lbmib_set_dynamics(h, d) bind(C, name="lbmib_set_dynamics")
            import :: c_int, c_ptr, lbmib_dynamics
            type(c_ptr), value :: h
            type(lbmib_dynamics), intent(in) :: d
        end function

        integer(c_int) function lbmib_get_body_state(h, state, n) bind(C, name="lbmib_get_body_state")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: state(*)
            integer(c_int32_t), value :: n
        end function

        integer(c_int) function lbmib_set_body_state(h, state, n) bind(C, name="lbmib_set_body_state")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(in) :: state(*)
            integer(c_int32_t), value :: n
        end function

        ! first, last: f0(1,:,:) and f0(Nx,:,:) as (Ny, 9) arrays, c_loc or c_null_ptr
        integer(c_int) function lbmib_get_edge_f0(h, first, last) bind(C, name="lbmib_get_edge_f0")
            import :: c_int, c_ptr
            type(c_ptr), value :: h, first, last
        end function

        integer(c_int) function lbmib_set_edge_f0(h, first, last) bind(C, name="lbmib_set_edge_f0")
            import :: c_int, c_ptr
            type(c_ptr), value :: h, first, last
        end function

        ! rows(1 + 3*wing_count, max_rows)
        integer(c_int) function lbmib_get_motion_rows(h, rows, max_rows, n_rows) bind(C, name="lbmib_get_motion_rows")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: rows(*)
            integer(c_int32_t), value :: max_rows
            integer(c_int32_t), intent(out) :: n_rows
        end function

        ! vor, dvordt: c_loc of nx*ny arrays or c_null_ptr
        integer(c_int) function lbmib_vorticity(h, rate, vor, dvordt) bind(C, name="lbmib_vorticity")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h, vor, dvordt
            integer(c_int32_t), value :: rate
        end function

        integer(c_int) function lbmib_surface_pressure(h, p, n) bind(C, name="lbmib_surface_pressure")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: p(*)
            integer(c_int32_t), value :: n
        end function

        ! no counterpart in the reference: the same, staged through the
        ! library's fixed per-device buffer (for callers that cannot allocate)
        integer(c_int) function lbmib_lattice_window(h, i0, j0, w, hgt) bind(C, name="lbmib_lattice_window")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: i0, j0, w, hgt
        end function

        integer(c_int) function lbmib_surface_pressure_noalloc(h, p, n) bind(C, name="lbmib_surface_pressure_noalloc")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: p(*)
            integer(c_int32_t), value :: n
        end function

        integer(c_int) function lbmib_set_step(h, n) bind(C, name="lbmib_set_step")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: n
        end function

        integer(c_int) function lbmib_residual(h, err) bind(C, name="lbmib_residual")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: err
        end function

        integer(c_int) function lbmib_ib_nodes(h, n, pi, pj, wing, glx, gly) bind(C, name="lbmib_ib_nodes")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), intent(inout) :: n
            type(c_ptr), value :: pi, pj, wing, glx, gly
        end function

        integer(c_int) function lbmib_set_profiling(h, every) bind(C, name="lbmib_set_profiling")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: every
        end function

        integer(c_int) function lbmib_get_timings(h, t, n) bind(C, name="lbmib_get_timings")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: t(*)
            integer(c_int32_t), value :: n
        end function

        integer(c_int) function lbmib_destroy(h) bind(C, name="lbmib_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function lbmib_set_shape(h, s) bind(C, name="lbmib_set_shape")
            import :: c_int, c_ptr, lbmib_launch_shape
            type(c_ptr), value :: h
            type(lbmib_launch_shape), intent(in) :: s
        end function

        integer(c_int) function lbmib_get_shape(h, s) bind(C, name="lbmib_get_shape")
            import :: c_int, c_ptr, lbmib_launch_shape
            type(c_ptr), value :: h
            type(lbmib_launch_shape), intent(out) :: s
        end function

        integer(c_int) function lbmib_device_bytes(h, nbytes) bind(C, name="lbmib_device_bytes")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int64_t), intent(out) :: nbytes
        end function

        integer(c_int) function lbmib_mem_info(device, free_bytes, total_bytes) bind(C, name="lbmib_mem_info")
            import :: c_int, c_int32_t, c_int64_t
            integer(c_int32_t), value :: device
            integer(c_int64_t), intent(out) :: free_bytes, total_bytes
        end function

        integer(c_int) function lbmib_batch_create(device, b) bind(C, name="lbmib_batch_create")
            import :: c_int, c_int32_t, c_ptr
            integer(c_int32_t), value :: device
            type(c_ptr), intent(out) :: b
        end function

        integer(c_int) function lbmib_batch_add(b, h) bind(C, name="lbmib_batch_add")
            import :: c_int, c_ptr
            type(c_ptr), value :: b, h
        end function

        integer(c_int) function lbmib_batch_remove(b, h) bind(C, name="lbmib_batch_remove")
            import :: c_int, c_ptr
            type(c_ptr), value :: b, h
        end function

        integer(c_int) function lbmib_batch_size(b, n) bind(C, name="lbmib_batch_size")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: b
            integer(c_int32_t), intent(out) :: n
        end function

        integer(c_int) function lbmib_batch_step(b, n_steps) bind(C, name="lbmib_batch_step")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: b
            integer(c_int32_t), value :: n_steps
        end function

        integer(c_int) function lbmib_batch_set_shape(b, s) bind(C, name="lbmib_batch_set_shape")
            import :: c_int, c_ptr, lbmib_launch_shape
            type(c_ptr), value :: b
            type(lbmib_launch_shape), intent(in) :: s
        end function

        integer(c_int) function lbmib_batch_get_shape(b, s) bind(C, name="lbmib_batch_get_shape")
            import :: c_int, c_ptr, lbmib_launch_shape
            type(c_ptr), value :: b
            type(lbmib_launch_shape), intent(out) :: s
        end function

        integer(c_int) function lbmib_batch_autotune(b, trial_steps, best, report, report_len) &
                bind(C, name="lbmib_batch_autotune")
            import :: c_int, c_int32_t, c_ptr, lbmib_launch_shape
            type(c_ptr), value :: b
            integer(c_int32_t), value :: trial_steps
            type(lbmib_launch_shape), intent(out) :: best
            type(c_ptr), value :: report
            integer(c_int32_t), value :: report_len
        end function

        integer(c_int) function lbmib_batch_destroy(b) bind(C, name="lbmib_batch_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: b
        end function

        type(c_ptr) function c_lbmib_last_error() bind(C, name="lbmib_last_error")
            import :: c_ptr
        end function

        type(c_ptr) function c_lbmib_version() bind(C, name="lbmib_version")
            import :: c_ptr
        end function

        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

contains

    function c_string(p) result(s)
        type(c_ptr), intent(in) :: p
        character(len=:), allocatable :: s
        character(kind=c_char), pointer :: chars(:)
        integer :: i, n
        if (.not. c_associated(p)) then
            s = ''
            return
        endif
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: s)
        do i = 1, n
            s(i:i) = chars(i)
        enddo
    end function

    ! message of the calling thread's last failure
    function lbmib_error_message() result(s)
        character(len=:), allocatable :: s
        s = c_string(c_lbmib_last_error())
    end function

    function lbmib_version_string() result(s)
        character(len=:), allocatable :: s
        s = c_string(c_lbmib_version())
    end function

    ! The reference stops the program on errors (e.g. mode.f90:276-277); the
    ! library returns codes, and this helper restores that behaviour for
    ! callers that want it.
    subroutine lbmib_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc .ne. LBMIB_OK) then
            write(*, '(a, a, i0, a, a)') trim(what), ' failed (', rc, '): ', lbmib_error_message()
            stop 1
        endif
    end subroutine

end module lbmib
