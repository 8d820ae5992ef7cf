! lbmib_driver.f90 -- the reference's main loop (main.f90:351-418) over the
! GPU hot path.  It keeps the reference's inputs and outputs:
!   boundary.dat, boundary<k>.dat   marker polygons, '(2f14.8)' (airfoil.f90:207-242)
!   motion_params.dat               list-directed rows, 1+9*wing_count columns (mode.f90:68-142)
!   force.dat                       one '(1000f15.8)' row per step (solution.f90:349-376)
!   airfoil_motion.dat              one '(1000f15.8)' row per step (solution.f90:378-390)
!   F<n>.dat                        Tecplot flow field and airfoil zones (output, solution.f90:129-238)
!   saved_flow_feild.dat            binary restart state and its copy
!   saved_flow_feild_copy.dat       (saveflowfield / readflowfield, solution.f90:1604-1839;
!                                   lbmib_restart.f90)
! and takes the parameters that the reference sets in ini_mesh / main from a
! namelist file (first argument, default lbmib.nml).
program lbmib_driver
    use iso_c_binding
    use lbmib
    use lbmib_restart
    implicit none

    integer :: nx = 0, ny = 0, smooth = 1, wing_count = 1, points_per_wing = 0
    integer :: chunk = 100, device = 0, report_every = 100
    real(c_double) :: ox = 0.0d0, oy = 0.0d0, scale = 0.0d0, UU = 0.05d0
    real(c_double) :: xc = 0.0d0, yc = 0.0d0, xp = 0.0d0, yp = 0.0d0, s(9) = 0.0d0
    character(len=256) :: boundary_dir = '.', motion_file = 'motion_params.dat'
    character(len=256) :: force_file = 'force.dat', motion_out_file = 'airfoil_motion.dat'
    character(len=256) :: fields_file = ''
    ! flow-field output every output_every steps (0: none) at every
    ! output_stride-th node (output_stride_x/y, main.f90:33-34); binary
    ! restart state every restart_every steps (0: none); restart_in resumes
    ! from such a file
    integer :: output_every = 0, output_stride = 2, restart_every = 0
    ! vorticity_rate = 1 calls vor_mom_rate after every step (the call
    ! main.f90:391 leaves commented out) and fills the dvordt column
    integer :: vorticity_rate = 0
    ! rigid-body dynamics of wing 1 (main.f90:76, 101-129): off by default,
    ! as the reference's parameter solve_dynamic_equations
    integer :: solve_dynamic_equations = 0, if_rk4 = 1, closed_boundary = 1
    real(c_double) :: solid_den = 10.0d0, area_airfoil = 6.807946476488794d-2
    real(c_double) :: par1 = 0.0d0, par2 = 0.0d0, alphad = 10.0d0 * 3.141592653589793d0 / 180.0d0
    character(len=256) :: output_dir = '.', restart_out = 'saved_flow_feild.dat', restart_in = ''
    namelist /lbmib_case/ nx, ny, ox, oy, scale, UU, xc, yc, xp, yp, s, smooth, wing_count, &
        points_per_wing, boundary_dir, motion_file, force_file, motion_out_file, fields_file, &
        chunk, device, report_every, output_every, output_stride, restart_every, output_dir, &
        restart_out, restart_in, vorticity_rate, solve_dynamic_equations, if_rk4, closed_boundary, &
        solid_den, area_airfoil, par1, par2, alphad

    type(lbmib_params) :: p
    type(lbmib_dynamics) :: dyn
    real(c_double), allocatable :: body_state(:), mrows(:)
    real(c_double), allocatable, target :: edge_first(:, :), edge_last(:, :)
    integer :: movenum = 0
    type(c_ptr) :: h
    character(len=256) :: nml_file, fname
    character(len=32) :: idx_text
    real(c_double), allocatable :: xb0(:), yb0(:), motion(:, :), hist(:, :), row(:)
    real(c_double), allocatable, target :: rho(:, :), u(:, :), v(:, :), f(:, :, :)
    real(c_double), allocatable, target :: vor(:, :), dvr(:, :)
    real(c_double) :: err, feq0(9)
    integer :: unit_id, ios, wing, i0, ps, ncols, nrows, nmax, n, k, cols, col, n0
    integer(c_int32_t) :: got
    character(len=1024) :: line

    nml_file = 'lbmib.nml'
    if (command_argument_count() .ge. 1) call get_command_argument(1, nml_file)
    open(newunit=unit_id, file=trim(nml_file), status='old', action='read', iostat=ios)
    if (ios .ne. 0) then
        write(*, *) 'cannot open parameter file: ', trim(nml_file)
        stop 1
    endif
    read(unit_id, nml=lbmib_case, iostat=ios)
    close(unit_id)
    if (ios .ne. 0) then
        write(*, *) 'bad namelist lbmib_case in ', trim(nml_file)
        stop 1
    endif
    if (wing_count .lt. 1 .or. wing_count .gt. 8 .or. points_per_wing .lt. 4) then
        write(*, *) 'wing_count must be 1..8 and points_per_wing >= 4'
        stop 1
    endif

    ! read_boundary: boundary.dat, boundary1.dat, ... with '(2f14.8)'
    allocate(xb0(wing_count * points_per_wing), yb0(wing_count * points_per_wing))
    do wing = 1, wing_count
        if (wing .eq. 1) then
            fname = 'boundary.dat'
        else
            write(idx_text, '(i0)') wing - 1
            fname = 'boundary' // trim(idx_text) // '.dat'
        endif
        fname = trim(boundary_dir) // '/' // trim(fname)
        open(newunit=unit_id, file=trim(fname), status='old', action='read', iostat=ios)
        if (ios .ne. 0) then
            write(*, *) 'cannot open boundary file: ', trim(fname)
            stop 1
        endif
        ps = (wing - 1) * points_per_wing
        read(unit_id, '(2f14.8)', iostat=ios) (xb0(i0), yb0(i0), i0 = ps + 1, ps + points_per_wing)
        close(unit_id)
        if (ios .ne. 0) then
            write(*, *) 'bad boundary file: ', trim(fname)
            stop 1
        endif
    enddo

    ! load_motion_table: list-directed rows, blank lines skipped
    ncols = 1 + 9 * wing_count
    open(newunit=unit_id, file=trim(motion_file), status='old', action='read', iostat=ios)
    if (ios .ne. 0) then
        write(*, *) 'cannot open motion file: ', trim(motion_file)
        stop 1
    endif
    nrows = 0
    do
        read(unit_id, '(a)', iostat=ios) line
        if (ios .ne. 0) exit
        if (len_trim(line) .gt. 0) nrows = nrows + 1
    enddo
    rewind(unit_id)
    allocate(motion(ncols, nrows))
    k = 0
    do
        read(unit_id, '(a)', iostat=ios) line
        if (ios .ne. 0) exit
        if (len_trim(line) .eq. 0) cycle
        k = k + 1
        read(line, *, iostat=ios) motion(:, k)
        if (ios .ne. 0) then
            write(*, *) 'motion file rows need ', ncols, ' columns: ', trim(motion_file)
            stop 1
        endif
    enddo
    close(unit_id)

    ! ini_mesh + init + read_boundary's post-processing on the device
    p%nx = nx
    p%ny = ny
    p%ox = ox
    p%oy = oy
    p%scale = scale
    p%U = UU
    p%xc = xc
    p%yc = yc
    p%xp = xp
    p%yp = yp
    p%s = s
    p%smooth = smooth
    p%precision = LBMIB_PREC_FP64
    p%n_wings = wing_count
    p%pts_per_wing = points_per_wing
    p%device = device
    call lbmib_check(lbmib_create(p, xb0, yb0, h), 'lbmib_create')
    if (solve_dynamic_equations .ne. 0) then
        dyn%enabled = 1
        dyn%rk4 = if_rk4
        dyn%closed_boundary = closed_boundary
        dyn%solid_den = solid_den
        dyn%area_airfoil = area_airfoil
        dyn%par1 = par1
        dyn%par2 = par2
        dyn%alphad = alphad
        call lbmib_check(lbmib_set_dynamics(h, dyn), 'lbmib_set_dynamics')
    endif
    allocate(body_state(11 * wing_count + 1), edge_first(ny, 9), edge_last(ny, 9))
    call lbmib_check(lbmib_set_motion_table(h, motion, int(nrows, c_int32_t), int(ncols, c_int32_t)), &
                     'lbmib_set_motion_table')
    call lbmib_check(lbmib_max_steps(h, got), 'lbmib_max_steps')
    nmax = got
    write(*, '(a, a, i0, a)') trim(lbmib_version_string()), ': ', nmax, ' steps'
    allocate(rho(nx, ny), u(nx, ny), v(nx, ny), vor(nx, ny), dvr(nx, ny))
    dvr = 0.0d0
    if (restart_every .gt. 0 .or. len_trim(restart_in) .gt. 0) allocate(f(nx, ny, 9))
    ! the stale ring populations f0 and the coarse-to-fine copies the
    ! restart file carries stay at the initial equilibrium: f of any node now
    if (restart_every .gt. 0) then
        call lbmib_check(lbmib_download(h, c_loc(rho), c_loc(u), c_loc(v), c_loc(f)), 'lbmib_download')
        feq0 = f(1, 1, :)
    endif

    ! readflowfield (main.f90:320-324): resume after the saved step
    n = 0
    if (len_trim(restart_in) .gt. 0) then
        call read_restart(trim(restart_in), n)
        call lbmib_check(lbmib_set_step(h, int(n, c_int32_t)), 'lbmib_set_step')
        call lbmib_check(lbmib_upload_state(h, f), 'lbmib_upload_state')
        call lbmib_check(lbmib_set_body_state(h, body_state, int(size(body_state), c_int32_t)), &
                         'lbmib_set_body_state')
        call lbmib_check(lbmib_set_edge_f0(h, c_loc(edge_first), c_loc(edge_last)), 'lbmib_set_edge_f0')
        write(*, '(a, i0)') 'resumed after step ', n
    endif
    n0 = n

    ! the main loop: evolution(mesh, n, nforce, h) for n = n0+1 .. nmax, the
    ! residual reported every report_every steps, the flow field and the
    ! restart state written on their schedules
    do while (n .lt. nmax)
        k = min(chunk, nmax - n)
        if (report_every .gt. 0) k = min(k, report_every - mod(n, report_every))
        if (output_every .gt. 0) k = min(k, output_every - mod(n, output_every))
        if (restart_every .gt. 0) k = min(k, restart_every - mod(n, restart_every))
        if (vorticity_rate .ne. 0) k = 1
        call lbmib_check(lbmib_step(h, int(k, c_int32_t)), 'lbmib_step')
        n = n + k
        if (vorticity_rate .ne. 0) then
            if (output_every .gt. 0) then
                if (mod(n, output_every) .eq. 0) then
                    call lbmib_check(lbmib_vorticity(h, 1_c_int32_t, c_loc(vor), c_loc(dvr)), 'vor_mom_rate')
                else
                    call lbmib_check(lbmib_vorticity(h, 1_c_int32_t, c_null_ptr, c_null_ptr), 'vor_mom_rate')
                endif
            else
                call lbmib_check(lbmib_vorticity(h, 1_c_int32_t, c_null_ptr, c_null_ptr), 'vor_mom_rate')
            endif
        endif
        if (report_every .gt. 0) then
            if (mod(n, report_every) .eq. 0 .and. n .ge. n0 + 2) then
                call lbmib_check(lbmib_residual(h, err), 'lbmib_residual')
                write(*, '(a, i0, a, es14.6)') 'step ', n, ' err ', err
            endif
        endif
        if (output_every .gt. 0) then
            if (mod(n, output_every) .eq. 0) call write_tecplot(n)
        endif
        if (restart_every .gt. 0) then
            if (mod(n, restart_every) .eq. 0) call write_restart(n)
        endif
    enddo

    ! force.dat and airfoil_motion.dat rows (appended after a restart, as
    ! the reference continues its force file)
    cols = 4 + 12 * wing_count
    allocate(hist(cols, max(nmax, 1)), row(cols))
    call lbmib_check(lbmib_get_history(h, hist, int(nmax, c_int32_t), got), 'lbmib_get_history')
    if (n0 .gt. 0) then
        open(newunit=unit_id, file=trim(force_file), status='unknown', position='append', action='write')
    else
        open(newunit=unit_id, file=trim(force_file), status='replace', action='write')
    endif
    do k = n0 + 1, nmax
        write(unit_id, '(1000f15.8)') hist(:, k)
    enddo
    close(unit_id)
    if (n0 .gt. 0) then
        open(newunit=unit_id, file=trim(motion_out_file), status='unknown', position='append', action='write')
    else
        open(newunit=unit_id, file=trim(motion_out_file), status='replace', action='write')
    endif
    ! the body's extra velocity and shift per wing (zero with the dynamics off)
    col = 1 + 3 * wing_count
    allocate(mrows(col * max(nmax, 1)))
    call lbmib_check(lbmib_get_motion_rows(h, mrows, int(nmax, c_int32_t), got), 'lbmib_get_motion_rows')
    do k = n0 + 1, nmax
        write(unit_id, '(1000f15.8)') mrows((k - 1) * col + 1:k * col)
    enddo
    close(unit_id)

    if (len_trim(fields_file) .gt. 0) then
        call lbmib_check(lbmib_download(h, c_loc(rho), c_loc(u), c_loc(v), c_null_ptr), 'lbmib_download')
        open(newunit=unit_id, file=trim(fields_file), status='replace', access='stream', form='unformatted')
        write(unit_id) rho, u, v
        close(unit_id)
    endif
    call lbmib_check(lbmib_destroy(h), 'lbmib_destroy')

contains

    ! output (solution.f90:129-238) for one mesh without domain shifting:
    ! a point zone of X, Y, U/UU, V/UU, the pressure rho/3, the vorticity
    ! (central differences, one-sided on the edges, computed on the device),
    ! the IB-node mask and the vorticity rate (zero unless vorticity_rate is
    ! set: the reference leaves its routine uncalled, main.f90:391) at every
    ! output_stride-th node; a triangle zone of the marker polygons with one
    ! centre point per wing.
    subroutine write_tecplot(cycle)
        integer, intent(in) :: cycle
        integer :: i, j, a, o, out_nx, out_ny, npts, kk, jp
        integer(c_int32_t) :: na
        integer, allocatable :: body(:, :)
        integer(c_int32_t), allocatable, target :: pi(:), pj(:)
        real(c_double), allocatable :: xb1(:), yb1(:)
        character(len=32) :: nm
        call lbmib_check(lbmib_download(h, c_loc(rho), c_loc(u), c_loc(v), c_null_ptr), 'lbmib_download')
        ! with vorticity_rate the step's vor_mom_rate call has filled vor and dvr
        if (vorticity_rate .eq. 0) &
            call lbmib_check(lbmib_vorticity(h, 0_c_int32_t, c_loc(vor), c_null_ptr), 'lbmib_vorticity')
        allocate(body(nx, ny))
        body = 0
        na = 0
        call lbmib_check(lbmib_ib_nodes(h, na, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), &
                         'lbmib_ib_nodes')
        if (na .gt. 0) then
            allocate(pi(na), pj(na))
            call lbmib_check(lbmib_ib_nodes(h, na, c_loc(pi), c_loc(pj), c_null_ptr, c_null_ptr, c_null_ptr), &
                             'lbmib_ib_nodes')
            do a = 1, na
                body(pi(a) + 1, pj(a) + 1) = 1
            enddo
        endif
        npts = wing_count * points_per_wing
        allocate(xb1(npts), yb1(npts))
        call lbmib_check(lbmib_markers(h, xb1, yb1), 'lbmib_markers')
        ! the domain has followed the body movenum nodes (solution.f90:210)
        call lbmib_check(lbmib_get_body_state(h, body_state, int(size(body_state), c_int32_t)), &
                         'lbmib_get_body_state')
        movenum = nint(body_state(size(body_state)))
        write(nm, '(i0)') cycle
        open(newunit=o, file=trim(output_dir) // '/F' // trim(nm) // '.dat', status='replace', action='write')
        write(o, '(a)') 'VARIABLES = "X" "Y" "U" "V" "Presure" "Vorticity" "body" "dvordt" '
        out_nx = (nx - 1) / output_stride + 1
        out_ny = (ny - 1) / output_stride + 1
        ! the reference writes the zone headers list-directed; an explicit
        ! format keeps each on one line whatever the compiler's record length
        write(o, '(a, i0, a, i0, a, es24.16)') ' ZONE T="Flow feild", I=', out_nx, ',J=', out_ny, ',SOLUTIONTIME=', &
            cycle * UU * scale
        do j = 1, ny, output_stride
            do i = 1, nx, output_stride
                write(o, '(25e16.8)') ox + dble(i - 1) * scale + dble(movenum) * scale, oy + dble(j - 1) * scale, &
                    u(i, j) / UU, &
                    v(i, j) / UU, rho(i, j) / 3.0d0, vor(i, j), dble(body(i, j)), dvr(i, j)
            enddo
        enddo
        write(o, '(a, i0, a, i0, a, a, es24.16)') ' ZONE T="airfoil", NODES=', npts + wing_count, ' ELEMENTS,=', &
            wing_count * (points_per_wing - 1), ', DATAPACKING=POINT, ZONETYPE=FETRIANGLE', ',SOLUTIONTIME=', &
            cycle * UU * scale
        do i = 1, npts
            write(o, '(8f14.8)') xb1(i), yb1(i), 255d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0
        enddo
        do i = 1, wing_count
            j = (i - 1) * points_per_wing + 1
            kk = j + (points_per_wing - 1) / 2
            write(o, '(8f14.8)') 0.5d0 * (xb1(j) + xb1(kk)), 0.5d0 * (yb1(j) + yb1(kk)), 255d0, 0.0d0, 0.0d0, &
                0.0d0, 0.0d0, 0.0d0
        enddo
        do i = 1, wing_count
            j = (i - 1) * points_per_wing + 1
            do jp = j, j + points_per_wing - 2
                write(o, '(3I5.3)') jp, jp + 1, npts + i
            enddo
        enddo
        close(o)
    end subroutine

    ! saveflowfield (solution.f90:1604-1707): the step, the wing count, the
    ! rigid-body state and the domain shift count, then per node u, v and
    ! f(1:9), then the edge records: the coarse-to-fine copies (the initial
    ! equilibrium for one mesh) and the stale f0 of the edge columns (from
    ! the device) and rows (the initial equilibrium: no step reads them, so
    ! the device does not keep them).  The primary file, then its copy
    ! (saved_flow_feild_copy.dat beside it, solution.f90:1677-1706).
    subroutine write_restart(cycle)
        integer, intent(in) :: cycle
        integer :: nb
        character(len=512) :: path
        call lbmib_check(lbmib_download(h, c_loc(rho), c_loc(u), c_loc(v), c_loc(f)), 'lbmib_download')
        nb = 11 * wing_count
        call lbmib_check(lbmib_get_body_state(h, body_state, int(nb + 1, c_int32_t)), 'lbmib_get_body_state')
        call lbmib_check(lbmib_get_edge_f0(h, c_loc(edge_first), c_loc(edge_last)), 'lbmib_get_edge_f0')
        path = trim(output_dir) // '/' // trim(restart_out)
        call write_flow_field(trim(path), cycle, wing_count, body_state(1:nb), nint(body_state(nb + 1)), nx, ny, &
                              u, v, f, feq0, edge_first, edge_last)
        call write_flow_field(restart_copy_name(trim(path)), cycle, wing_count, body_state(1:nb), &
                              nint(body_state(nb + 1)), nx, ny, u, v, f, feq0, edge_first, edge_last)
    end subroutine

    ! readflowfield (solution.f90:1709-1839): the step, the rigid-body state,
    ! the populations and the stale f0 of the edge columns; the velocity
    ! fields are recomputed by the next step.  A primary file that cannot be
    ! opened or read to its end gives way to its copy (solution.f90:1737-1802).
    subroutine read_restart(fname, cycle)
        character(len=*), intent(in) :: fname
        integer, intent(out) :: cycle
        integer :: nb, st, nwf, mv
        nb = 11 * wing_count
        call read_flow_field(fname, cycle, wing_count, body_state(1:nb), mv, nx, ny, f, edge_first, edge_last, &
                             st, nwf)
        if (st .eq. RESTART_UNREADABLE) then
            write(*, *) 'restart file unreadable, reading its copy: ', restart_copy_name(fname)
            call read_flow_field(restart_copy_name(fname), cycle, wing_count, body_state(1:nb), mv, nx, ny, f, &
                                 edge_first, edge_last, st, nwf)
            if (st .eq. RESTART_UNREADABLE) then
                write(*, *) 'cannot read the restart file or its copy: ', fname
                stop 1
            endif
        endif
        if (st .eq. RESTART_WING_COUNT) then
            write(*, *) 'restart file wing count ', nwf, ' differs from wing_count ', wing_count
            stop 1
        endif
        body_state(nb + 1) = dble(mv)
    end subroutine
end program lbmib_driver
