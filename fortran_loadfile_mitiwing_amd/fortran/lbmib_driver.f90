! lbmib_driver.f90 -- the reference's main loop (main.f90:351-418) over the
! GPU hot path.  It keeps the reference's inputs and outputs:
!   boundary.dat, boundary<k>.dat   marker polygons, '(2f14.8)' (airfoil.f90:207-242)
!   motion_params.dat               list-directed rows, 1+9*wing_count columns (mode.f90:68-142)
!   force.dat                       one '(1000f15.8)' row per step (solution.f90:349-376)
!   airfoil_motion.dat              one '(1000f15.8)' row per step (solution.f90:378-390)
! and takes the parameters that the reference sets in ini_mesh / main from a
! namelist file (first argument, default lbmib.nml).
program lbmib_driver
    use iso_c_binding
    use lbmib
    implicit none

    integer :: nx = 0, ny = 0, smooth = 1, wing_count = 1, points_per_wing = 0
    integer :: chunk = 100, device = 0, report_every = 100
    real(c_double) :: ox = 0.0d0, oy = 0.0d0, scale = 0.0d0, UU = 0.05d0
    real(c_double) :: xc = 0.0d0, yc = 0.0d0, xp = 0.0d0, yp = 0.0d0, s(9) = 0.0d0
    character(len=256) :: boundary_dir = '.', motion_file = 'motion_params.dat'
    character(len=256) :: force_file = 'force.dat', motion_out_file = 'airfoil_motion.dat'
    character(len=256) :: fields_file = ''
    namelist /lbmib_case/ nx, ny, ox, oy, scale, UU, xc, yc, xp, yp, s, smooth, wing_count, &
        points_per_wing, boundary_dir, motion_file, force_file, motion_out_file, fields_file, &
        chunk, device, report_every

    type(lbmib_params) :: p
    type(c_ptr) :: h
    character(len=256) :: nml_file, fname
    character(len=32) :: idx_text
    real(c_double), allocatable :: xb0(:), yb0(:), motion(:, :), hist(:, :), row(:)
    real(c_double), allocatable, target :: rho(:, :), u(:, :), v(:, :)
    real(c_double) :: err
    integer :: unit_id, ios, wing, i0, ps, ncols, nrows, nmax, n, k, cols, col
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
    call lbmib_check(lbmib_set_motion_table(h, motion, int(nrows, c_int32_t), int(ncols, c_int32_t)), &
                     'lbmib_set_motion_table')
    call lbmib_check(lbmib_max_steps(h, got), 'lbmib_max_steps')
    nmax = got
    write(*, '(a, a, i0, a)') trim(lbmib_version_string()), ': ', nmax, ' steps'

    ! the main loop: evolution(mesh, n, nforce, h) for n = 1 .. nmax, the
    ! residual reported every report_every steps
    n = 0
    do while (n .lt. nmax)
        k = min(chunk, nmax - n)
        if (report_every .gt. 0) k = min(k, report_every - mod(n, report_every))
        call lbmib_check(lbmib_step(h, int(k, c_int32_t)), 'lbmib_step')
        n = n + k
        if (report_every .gt. 0) then
            if (mod(n, report_every) .eq. 0 .and. n .ge. 2) then
                call lbmib_check(lbmib_residual(h, err), 'lbmib_residual')
                write(*, '(a, i0, a, es14.6)') 'step ', n, ' err ', err
            endif
        endif
    enddo

    ! force.dat and airfoil_motion.dat rows
    cols = 4 + 12 * wing_count
    allocate(hist(cols, max(nmax, 1)), row(cols))
    call lbmib_check(lbmib_get_history(h, hist, int(nmax, c_int32_t), got), 'lbmib_get_history')
    open(newunit=unit_id, file=trim(force_file), status='replace', action='write')
    do k = 1, nmax
        write(unit_id, '(1000f15.8)') hist(:, k)
    enddo
    close(unit_id)
    open(newunit=unit_id, file=trim(motion_out_file), status='replace', action='write')
    do k = 1, nmax
        ! with the dynamics off the body's extra velocity and shift are zero
        row(1) = hist(1, k)
        col = 2
        do wing = 1, wing_count
            row(col:col + 2) = 0.0d0
            col = col + 3
        enddo
        write(unit_id, '(1000f15.8)') row(1:col - 1)
    enddo
    close(unit_id)

    if (len_trim(fields_file) .gt. 0) then
        allocate(rho(nx, ny), u(nx, ny), v(nx, ny))
        call lbmib_check(lbmib_download(h, c_loc(rho), c_loc(u), c_loc(v), c_null_ptr), 'lbmib_download')
        open(newunit=unit_id, file=trim(fields_file), status='replace', access='stream', form='unformatted')
        write(unit_id) rho, u, v
        close(unit_id)
    endif
    call lbmib_check(lbmib_destroy(h), 'lbmib_destroy')
end program lbmib_driver
