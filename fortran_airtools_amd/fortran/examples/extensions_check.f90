! extensions_check.f90 -- a Fortran caller of fair_extensions: FP32 sirt, batched sirt and the
! matrix-free fan-beam projector, next to the drop-in modules.
!
! Reads (stream file, argument 1): n, p, nth, r, dw, sd, theta(nth), m, k, B(m, k) as real(4),
! x probe(n*n) as real(4).  Writes (argument 2): sirt_f32('sart', B(:,1), 20 iterations),
! sirt_batch('sart', B, 20 iterations) Xout(n, k), the projector's A*probe and A^T*B(:,1), its
! segment count, and 10 gradient-descent iterations on B(:,1); all with box [0,1].
program extensions_check
    use system_generation
    use sparse_matrices
    use AIRMs
    use fair_extensions
    implicit none
    type(data) :: dat
    type(csr_matrix) :: A
    type(opts) :: options
    type(fan_projector) :: P
    character(len=512) :: fin, fout
    integer :: u, m, k
    integer(kind=8) :: nseg
    real(kind=4), allocatable :: B(:, :), probe(:), xs(:), X(:, :), y(:), w(:), xg(:)

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', action='read')
    read (u) dat%n, dat%p, dat%nth, dat%r, dat%dw, dat%sd
    allocate (dat%theta(dat%nth))
    read (u) dat%theta
    read (u) m, k
    allocate (B(m, k), probe(dat%n**2))
    read (u) B, probe
    close (u)

    call options%initialise()
    call options%set_lbound(0d0)
    call options%set_ubound(1d0)

    A = fanlineartomo(dat)
    allocate (xs(A%n), X(A%n, k), xg(A%n))
    call sirt_f32(xs, 'sart', A, B(:, 1), [20], options=options)
    call sirt_batch(X, 'sart', A, B, [20], options=options)

    call P%create(dat)
    y = P%matvecmul(probe)
    w = P%transvecmul(B(:, 1))
    nseg = P%segments()
    call P%gradient_descent(xg, B(:, 1), max_iter=10, tolerance=0d0, options=options)
    call P%free()

    open (newunit=u, file=trim(fout), access='stream', form='unformatted', action='write', status='replace')
    write (u) xs, X, y, w, nseg, xg
    close (u)
end program extensions_check
