! twin_check.f90 -- a reference-style Fortran caller of twin_art (AIRMs), with the runtime's
! random_number stream seeded as the golden-vector script seeds the reference.
!
! Reads (stream file, argument 1): n, p, nth, r, dw, sd, theta(nth), m, b(m).  Runs three cases:
! kaczmarz/mutualstep (Kin 10,20,40), randkaczmarz/errorgauge (Kin 10,30) and randkaczmarz with
! row probabilities/mutualstep (Kin 10,30), box [0,1], seeds 201, 204, 205.  Writes (argument 2)
! per case: final_it, xout, and 1 if the generator continues exactly where the draws twin_art
! consumed end (else 0).
program twin_check
    use system_generation
    use sparse_matrices
    use AIRMs
    implicit none
    type(data) :: dat
    type(csr_matrix) :: A
    type(opts) :: options
    character(len=512) :: fin, fout
    integer :: u, m, c, fi, ok
    real(kind=8), allocatable :: b(:), xout(:), Xs(:, :), Ys(:, :), EG(:)
    character(len=12) :: meth(3), rule(3)
    integer :: seeds(3)
    logical :: rp(3)
    meth = [character(len=12) :: 'kaczmarz', 'randkaczmarz', 'randkaczmarz']
    rule = [character(len=12) :: 'mutualstep', 'errorgauge', 'mutualstep']
    seeds = [201, 204, 205]
    rp = [.false., .false., .true.]

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', action='read')
    read (u) dat%n, dat%p, dat%nth, dat%r, dat%dw, dat%sd
    allocate (dat%theta(dat%nth))
    read (u) dat%theta
    read (u) m
    allocate (b(m))
    read (u) b
    close (u)
    A = fanlineartomo(dat)
    allocate (xout(A%n))
    open (newunit=u, file=trim(fout), access='stream', form='unformatted', action='write', status='replace')
    do c = 1, 3
        call options%initialise()
        call options%set_lbound(0d0)
        call options%set_ubound(1d0)
        options%rowprobability = rp(c)
        call seed(seeds(c))
        if (c == 1) then
            allocate (Xs(A%n, 3), Ys(A%n, 3), EG(40))
            call twin_art(xout, fi, Xs, Ys, EG, trim(meth(c)), trim(rule(c)), A, b, [10, 20, 40], options=options)
        else
            allocate (Xs(A%n, 2), Ys(A%n, 2), EG(30))
            call twin_art(xout, fi, Xs, Ys, EG, trim(meth(c)), trim(rule(c)), A, b, [10, 30], options=options)
        end if
        ok = next_draw_matches(seeds(c))
        write (u) fi, xout, ok
        deallocate (Xs, Ys, EG)
    end do
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

    ! The next draw after twin_art equals draw number q+1 of a fresh stream, where q is the
    ! smallest count whose draw q+1 matches (the draws are distinct in practice).
    integer function next_draw_matches(s0)
        integer, intent(in) :: s0
        real(kind=8) :: after, v
        integer :: q
        call random_number(after)
        call seed(s0)
        next_draw_matches = 0
        do q = 1, 2000000
            call random_number(v)
            if (v == after) then
                next_draw_matches = q
                return
            end if
        end do
    end function next_draw_matches
end program twin_check
