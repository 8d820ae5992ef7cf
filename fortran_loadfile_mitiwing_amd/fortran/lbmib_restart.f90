! lbmib_restart.f90 -- the reference's binary restart state, saved_flow_feild.dat
! (saveflowfield / readflowfield, solution.f90:1604-1839), as plain Fortran I/O
! with no library calls, so that the file logic is testable without a device.
!
! Layout (raw stream, as form='binary' writes it): the step (int32), the wing
! count (int32), the 11 rigid-body arrays of wing_count values each, the
! domain-shift count (int32), then per node (x fastest) u, v and f(1:9), then
! the edge records: for rows 2..Ny-1 the coarse-to-fine copy and the stale f0
! of column 1, then the same of column Nx; for columns 1..Nx the copy and the
! stale f0 of row Ny, then the same of row 1 (solution.f90:1651-1675).
!
! Resilience as in the reference: every save writes the primary file and then
! a copy beside it (saved_flow_feild_copy.dat, solution.f90:1677-1706); a
! resume reads the primary and, when it cannot be opened or any record of it
! cannot be read (a save cut short), reads the copy instead (readflowfield's
! fallback, solution.f90:1737-1802).  A wing count that differs from the
! case's stops the run in either file, as it does there.
module lbmib_restart
    use iso_c_binding, only: c_double, c_int32_t
    implicit none
    private
    public :: restart_copy_name, write_flow_field, read_flow_field
    public :: RESTART_OK, RESTART_UNREADABLE, RESTART_WING_COUNT

    integer, parameter :: RESTART_OK = 0, RESTART_UNREADABLE = 1, RESTART_WING_COUNT = 2

contains

    ! The copy's name: 'X.dat' -> 'X_copy.dat', anything else gets '_copy'.
    function restart_copy_name(path) result(copy)
        character(len=*), intent(in) :: path
        character(len=len_trim(path) + 5) :: copy
        integer :: n
        n = len_trim(path)
        if (n .ge. 4) then
            if (path(n - 3:n) .eq. '.dat') then
                copy = path(1:n - 4) // '_copy.dat'
                return
            endif
        endif
        copy = path(1:n) // '_copy'
    end function

    subroutine write_flow_field(path, cycle, wing_count, body, movenum, nx, ny, u, v, f, feq0, &
                                edge_first, edge_last)
        character(len=*), intent(in) :: path
        integer, intent(in) :: cycle, wing_count, movenum, nx, ny
        real(c_double), intent(in) :: body(11 * wing_count), u(nx, ny), v(nx, ny), f(nx, ny, 9), feq0(9)
        real(c_double), intent(in) :: edge_first(ny, 9), edge_last(ny, 9)
        integer :: i, j, o
        integer(c_int32_t) :: i4
        open(newunit=o, file=path, status='replace', access='stream', form='unformatted', action='write')
        i4 = cycle
        write(o) i4
        i4 = wing_count
        write(o) i4
        write(o) body
        i4 = movenum
        write(o) i4
        do j = 1, ny
            do i = 1, nx
                write(o) u(i, j), v(i, j), f(i, j, :)
            enddo
        enddo
        do j = 2, ny - 1
            write(o) feq0, edge_first(j, :), feq0, edge_last(j, :)
        enddo
        do j = 1, nx
            if (j .eq. 1) then
                write(o) feq0, edge_first(ny, :), feq0, edge_first(1, :)
            else if (j .eq. nx) then
                write(o) feq0, edge_last(ny, :), feq0, edge_last(1, :)
            else
                write(o) feq0, feq0, feq0, feq0
            endif
        enddo
        close(o)
    end subroutine

    ! Reads one file; stat is RESTART_OK, RESTART_UNREADABLE (cannot open, or
    ! a record is missing or unreadable) or RESTART_WING_COUNT (file_wings
    ! then holds the file's count).  Nothing is taken from a file that fails.
    subroutine read_flow_field(path, cycle, wing_count, body, movenum, nx, ny, f, edge_first, edge_last, &
                               stat, file_wings)
        character(len=*), intent(in) :: path
        integer, intent(in) :: wing_count, nx, ny
        integer, intent(out) :: cycle, movenum, stat, file_wings
        real(c_double), intent(inout) :: body(11 * wing_count), f(nx, ny, 9)
        real(c_double), intent(inout) :: edge_first(ny, 9), edge_last(ny, 9)
        real(c_double), allocatable :: ft(:, :, :), ef(:, :), el(:, :)
        real(c_double) :: bt(11 * wing_count), uu1, vv1, skip(9)
        integer :: i, j, o, ios
        integer(c_int32_t) :: i4, nwf, mv
        stat = RESTART_UNREADABLE
        file_wings = -1
        cycle = 0
        movenum = 0
        open(newunit=o, file=path, status='old', access='stream', form='unformatted', action='read', iostat=ios)
        if (ios .ne. 0) return
        allocate(ft(nx, ny, 9), ef(ny, 9), el(ny, 9))
        ef = edge_first
        el = edge_last
        read(o, iostat=ios) i4
        if (ios .ne. 0) goto 90
        read(o, iostat=ios) nwf
        if (ios .ne. 0) goto 90
        file_wings = nwf
        if (nwf .ne. wing_count) then
            stat = RESTART_WING_COUNT
            goto 90
        endif
        read(o, iostat=ios) bt
        if (ios .ne. 0) goto 90
        read(o, iostat=ios) mv
        if (ios .ne. 0) goto 90
        do j = 1, ny
            do i = 1, nx
                read(o, iostat=ios) uu1, vv1, ft(i, j, :)
                if (ios .ne. 0) goto 90
            enddo
        enddo
        do j = 2, ny - 1
            read(o, iostat=ios) skip, ef(j, :), skip, el(j, :)
            if (ios .ne. 0) goto 90
        enddo
        do j = 1, nx
            if (j .eq. 1) then
                read(o, iostat=ios) skip, ef(ny, :), skip, ef(1, :)
            else if (j .eq. nx) then
                read(o, iostat=ios) skip, el(ny, :), skip, el(1, :)
            else
                read(o, iostat=ios) skip, skip, skip, skip
            endif
            if (ios .ne. 0) goto 90
        enddo
        cycle = i4
        movenum = mv
        body = bt
        f = ft
        edge_first = ef
        edge_last = el
        stat = RESTART_OK
90      close(o)
    end subroutine
end module lbmib_restart
