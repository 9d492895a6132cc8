! dropin_check.f90 -- a reference-style Fortran caller of the drop-in modules.
!
! Reads a problem from the stream file named by argument 1 (n, p, nth, r, dw, sd, theta(nth),
! m, b(m), x probe(n*n)), then uses the reference's module API exactly as a FAIR tools program
! would: fanlineartomo, matvecmul, transvecmul, sirt('sart'), sirt('cimmino') and
! art('symkaczmarz'), all with box [0,1].  Writes the results as a stream file named by
! argument 2.
program dropin_check
    use system_generation
    use sparse_matrices
    use AIRMs
    implicit none
    type(data) :: dat
    type(csr_matrix) :: A
    type(opts) :: options
    character(len=512) :: fin, fout
    integer :: u, m, nn
    real(kind=8), allocatable :: b(:), probe(:), y(:), w(:), xs(:), xc(:), X(:, :)

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', action='read')
    read (u) dat%n, dat%p, dat%nth, dat%r, dat%dw, dat%sd
    allocate (dat%theta(dat%nth))
    read (u) dat%theta
    read (u) m
    allocate (b(m), probe(dat%n**2))
    read (u) b, probe
    close (u)

    A = fanlineartomo(dat)
    nn = A%nnz
    y = A%matvecmul(probe)
    w = A%transvecmul(b)

    call options%initialise()
    call options%set_lbound(0d0)
    call options%set_ubound(1d0)
    allocate (xs(A%n), xc(A%n), X(A%n, 1))
    call art(X, 'symkaczmarz', A, b, [4], options=options)
    call sirt(xc, 'cimmino', A, b, [100], options=options)

    open (newunit=u, file=trim(fout), access='stream', form='unformatted', action='write', status='replace')
    write (u) A%m, A%n, A%nnz
    write (u) A%IA, A%JA, A%vA
    write (u) y, w, X(:, 1), xc
    call sirt(xs, 'sart', A, b, [100], options=options)
    write (u) xs
    write (u) A%JA
    close (u)
end program dropin_check
