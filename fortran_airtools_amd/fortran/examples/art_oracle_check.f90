! art_oracle_check.f90 -- a reference-style Fortran caller of art with xref (AIRMs), the
! runtime's random_number stream seeded as the golden-vector script seeds the reference.
!
! Reads (stream file, argument 1): n, p, nth, r, dw, sd, theta(nth), xref(n*n), x0(n*n), m, and
! three right-hand sides b(m).  Runs symkaczmarz (Kin 4,60, start x0, box [0,1]), randkaczmarz
! (Kin 2,30, lower bound 0, seed 304) and randkaczmarz with row probabilities (Kin 30, upper
! bound 1, relaxation 0.8, seed 305).  Writes (argument 2) per case: Xout, and the position in a
! fresh stream of the next draw after art (draws consumed + 1).
program art_oracle_check
    use system_generation
    use sparse_matrices
    use AIRMs
    implicit none
    type(data) :: dat
    type(csr_matrix) :: A
    type(opts) :: options
    character(len=512) :: fin, fout
    integer :: u, m, nn
    real(kind=8), allocatable :: xref(:), x0(:), b(:, :), X2(:, :), X1(:, :)

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', action='read')
    read (u) dat%n, dat%p, dat%nth, dat%r, dat%dw, dat%sd
    allocate (dat%theta(dat%nth))
    read (u) dat%theta
    nn = dat%n**2
    allocate (xref(nn), x0(nn))
    read (u) xref, x0
    read (u) m
    allocate (b(m, 3))
    read (u) b
    close (u)
    A = fanlineartomo(dat)
    allocate (X2(A%n, 2), X1(A%n, 1))
    open (newunit=u, file=trim(fout), access='stream', form='unformatted', action='write', status='replace')

    call options%initialise()
    call options%set_lbound(0d0)
    call options%set_ubound(1d0)
    call seed(303)
    call art(X2, 'symkaczmarz', A, b(:, 1), [4, 60], x0=x0, xref=xref, options=options)
    write (u) X2, next_draw(303)

    call options%initialise()
    call options%set_lbound(0d0)
    call seed(304)
    call art(X2, 'randkaczmarz', A, b(:, 2), [2, 30], xref=xref, options=options)
    write (u) X2, next_draw(304)

    call options%initialise()
    call options%set_ubound(1d0)
    options%relaxpar = 0.8d0
    options%rowprobability = .true.
    call seed(305)
    call art(X1, 'randkaczmarz', A, b(:, 3), [30], xref=xref, options=options)
    write (u) X1, next_draw(305)
    close (u)
contains
    subroutine seed(s0)
        integer, intent(in) :: s0
        integer :: k, i
        integer, allocatable :: s(:)
        call random_seed(size=k)
        allocate (s(k))
        do i = 1, k
            s(i) = s0 + 7919*i
        end do
        call random_seed(put=s)
    end subroutine seed

    ! Position of the next draw in a fresh stream of seed s0 (the draws are distinct in practice).
    integer function next_draw(s0)
        integer, intent(in) :: s0
        real(kind=8) :: after, v
        integer :: q
        call random_number(after)
        call seed(s0)
        next_draw = 0
        do q = 1, 2000000
            call random_number(v)
            if (v == after) then
                next_draw = q
                return
            end if
        end do
    end function next_draw
end program art_oracle_check
