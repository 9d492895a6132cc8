! phantom_check.f90 -- a reference-style Fortran caller of the phantom gallery and of
! add_white_noise, with the runtime generator seeded as put(seed + 7919*i), i = 1..size.
!
!   phantom_check draws   <seed> <count> <out>     out: count random_number draws (stream)
!   phantom_check phantom <name> <N> <pin> <steps> <v> <gauss> <seed> <out>
!                 out: vectorise(<name>(N, ...)); pin / steps / v <= 0 leave that optional
!                 argument absent, gauss = 1 selects randomtype 'gaussian'
!   phantom_check noise   <in> <eta> <seed> <out>  in: m, b(m) (stream); out: bt, e
program phantom_check
    use system_generation
    use phantomgallery
    implicit none
    character(len=512) :: mode, a2, a3, a4, a5, a6, a7, a8, a9
    integer :: n, pin, steps, gauss, seed, cnt, u, m, k
    real(kind=8) :: v, eta
    real(kind=8), allocatable :: x(:, :), w(:), b(:), bt(:), e(:)
    character(8) :: rt
    call get_command_argument(1, mode)
    call get_command_argument(2, a2)
    call get_command_argument(3, a3)
    call get_command_argument(4, a4)
    call get_command_argument(5, a5)
    call get_command_argument(6, a6)
    call get_command_argument(7, a7)
    call get_command_argument(8, a8)
    call get_command_argument(9, a9)
    select case (trim(mode))
    case ('draws')
        read (a2, *) seed
        read (a3, *) cnt
        call seed_runtime(seed)
        allocate (w(cnt))
        do k = 1, cnt
            call random_number(w(k))
        end do
        call write_out(a4, w)
    case ('phantom')
        read (a3, *) n
        read (a4, *) pin
        read (a5, *) steps
        read (a6, *) v
        read (a7, *) gauss
        read (a8, *) seed
        rt = 'uniform'
        if (gauss /= 0) rt = 'gaussian'
        call seed_runtime(seed)
        allocate (x(n, n))
        select case (trim(a2))
        case ('shepplogan')
            x = shepplogan(n)
        case ('smooth')
            if (pin > 0) then
                x = smooth(n, pin)
            else
                x = smooth(n)
            end if
        case ('threephases')
            if (pin > 0) then
                x = threephases(n, pin)
            else
                x = threephases(n)
            end if
        case ('threephases_smooth')
            if (pin > 0 .and. v > 0) then
                x = threephases_smooth(n, pin, v)
            else if (pin > 0) then
                x = threephases_smooth(n, pin)
            else
                x = threephases_smooth(n)
            end if
        case ('binary')
            if (pin > 0 .and. steps > 0) then
                x = binary(n, pin, steps)
            else if (pin > 0) then
                x = binary(n, pin)
            else
                x = binary(n)
            end if
        case ('fourphases')
            x = fourphases(n)
        case ('mst')
            if (pin > 0) then
                x = mst(n, pin)
            else
                x = mst(n)
            end if
        case ('grains')
            if (pin > 0) then
                x = grains(n, rt, pin)
            else
                x = grains(n, rt)
            end if
        case ('bubbles')
            if (pin > 0) then
                x = bubbles(n, rt, pin)
            else
                x = bubbles(n, rt)
            end if
        case default
            x = choose_phantom(trim(a2), n)
        end select
        call write_out(a9, vectorise(x))
    case ('noise')
        read (a3, *) eta
        read (a4, *) seed
        open (newunit=u, file=trim(a2), access='stream', form='unformatted', action='read')
        read (u) m
        allocate (b(m), bt(m), e(m))
        read (u) b
        close (u)
        call seed_runtime(seed)
        call add_white_noise(bt, e, b, eta)
        call write_out(a5, [bt, e])
    case default
        stop 'usage: phantom_check draws|phantom|noise ...'
    end select
contains
    subroutine seed_runtime(s0)
        integer, intent(in) :: s0
        integer :: k2, ns
        integer, allocatable :: s(:)
        call random_seed(size=ns)
        allocate (s(ns))
        do k2 = 1, ns
            s(k2) = s0 + 7919*k2
        end do
        call random_seed(put=s)
    end subroutine seed_runtime

    subroutine write_out(path, vals)
        character(*), intent(in) :: path
        real(kind=8), intent(in) :: vals(:)
        integer :: uu
        open (newunit=uu, file=trim(path), access='stream', form='unformatted', action='write', status='replace')
        write (uu) vals
        close (uu)
    end subroutine write_out
end program phantom_check
