! filemode_check.f90 -- the file mode of a FAIR tools main program (main.f90:222-275), written
! against the drop-in modules as that program would be: read_data, the angle cull (-c skip), the
! grid_size rescale, fanlineartomo scaled by grid_size/N, bt = -log(T), then 20 SART iterations
! with a lower bound 0.  The right-hand side is also formed on the device (fair_extensions'
! absorbance) for comparison.  Also reads a row order and a reference image when named.
!
! Arguments: data file, skip, output file, [ordering file, reference-image file, image size].
! Output (stream): nth, m, nnz (default integers), grid_size, r, dw, sd, theta(nth), bt(m),
! bt from the device(m), IA(0:m), JA(nnz), vA(nnz), x(n*n) [, l, ordering(l), xref].
program filemode_check
    use system_generation
    use sparse_matrices
    use AIRMs
    use fair_extensions
    implicit none
    type(data) :: dat
    type(csr_matrix) :: A
    type(opts) :: options
    character(len=512) :: fin, fout, arg, ford, fref
    integer :: skip, cull, u, nimg
    real(kind=8), allocatable :: data_temp(:, :), theta_temp(:), bt(:), bd(:), xout(:), xref(:)
    integer, allocatable :: ordering(:)
    real(kind=8) :: grid_size

    call get_command_argument(1, fin)
    call get_command_argument(2, arg)
    read (arg, *) skip
    call get_command_argument(3, fout)

    call dat%read(trim(fin))
    if (skip > 0) then
        allocate (data_temp(dat%p, dat%nth))
        data_temp = reshape(dat%T, [dat%p, dat%nth])
        cull = size(dat%theta(1:dat%nth:skip))
        theta_temp = dat%theta(1:dat%nth:skip)
        call move_alloc(theta_temp, dat%theta)
        deallocate (dat%T)
        allocate (dat%T(cull*dat%p))
        dat%T = reshape(data_temp(:, 1:dat%nth:skip), [cull*dat%p])
        dat%nth = cull
    end if

    grid_size = dat%dw*dat%r/dat%sd
    dat%r = dat%r/grid_size
    dat%dw = dat%dw/grid_size
    dat%sd = dat%sd/grid_size

    A = fanlineartomo(dat)
    call A%multiply_by_scalar(grid_size/dat%n)
    allocate (bt(A%m))
    bt = -log(dat%T)
    bd = absorbance(dat%T, dat%p, dat%nth, 0)

    open (newunit=u, file=trim(fout), access='stream', form='unformatted', action='write', status='replace')
    write (u) dat%nth, A%m, A%nnz
    write (u) grid_size, dat%r, dat%dw, dat%sd
    write (u) dat%theta, bt, bd
    write (u) A%IA, A%JA, A%vA

    call options%initialise()
    call options%set_lbound(0d0)
    allocate (xout(A%n))
    call sirt(xout, 'sart', A, bt, [20], options=options)
    write (u) xout

    if (command_argument_count() >= 6) then
        call get_command_argument(4, ford)
        call get_command_argument(5, fref)
        call get_command_argument(6, arg)
        read (arg, *) nimg
        call read_ordering(ordering, trim(ford))
        allocate (xref(nimg))
        call read_ref(xref, trim(fref), nimg)
        write (u) size(ordering), ordering, xref
    end if
    close (u)
end program filemode_check
