! airms.f90 -- drop-in for the reference module AIRMs (AIRMs.f90 of FAIR tools): type(opts),
! art, sirt, gradient_descent, normest and scramble, with the iterations on the gfx950 device.
!
! Random draws stay on the host and follow the reference's order, so with the same compiler
! runtime the row sequences of randomized Kaczmarz and normest's start vector are those of the
! reference: art draws 2*nord uniforms for its set-up scramble, then per iteration 2*nord for
! scramble and, with rowprobability, one per row step (AIRMs.f90:297, 320, 830).
!
! Differences: verbose mode prints no per-iteration log; sirt without options uses relax 1
! (the reference leaves it undefined, AIRMs.f90:1050-1057); gradient_descent's sigma comes
! from the device power iteration, equal to the reference's to normest's tolerance (1e-5).
module AIRMs
    use iso_c_binding
    use fair_capi
    use sparse_matrices
    implicit none
    private :: initialise_options, set_upperbound, set_lowerbound
    public :: to_c

    ! type(opts), AIRMs.f90:6-13
    type opts
        real(kind=8) :: relaxpar, ubound, lbound, damp
        logical :: verbose, upb, lwrb, rowprobability
    contains
        procedure :: initialise => initialise_options
        procedure :: set_ubound => set_upperbound
        procedure :: set_lbound => set_lowerbound
    end type opts

contains

    ! AIRMs.f90:17-37 (placeholders ubound = 0, lbound = 1 with both bounds off)
    subroutine initialise_options(options)
        class(opts) :: options
        options%upb = .false.
        options%lwrb = .false.
        options%ubound = 0d0
        options%lbound = 1d0
        options%damp = 0d0
        options%relaxpar = 1d0
        options%verbose = .false.
        options%rowprobability = .false.
    end subroutine initialise_options

    subroutine set_upperbound(options, ubound)
        class(opts), intent(inout) :: options
        real(kind=8) :: ubound
        options%upb = .true.
        options%ubound = ubound
    end subroutine set_upperbound

    subroutine set_lowerbound(options, lbound)
        class(opts), intent(inout) :: options
        real(kind=8) :: lbound
        options%lwrb = .true.
        options%lbound = lbound
    end subroutine set_lowerbound

    function to_c(options) result(o)
        class(opts), intent(in), optional :: options
        type(fair_opts_c) :: o
        o%relaxpar = 1d0
        o%ubound = 0d0
        o%lbound = 1d0
        o%damp = 0d0
        o%verbose = 0
        o%upb = 0
        o%lwrb = 0
        o%rowprobability = 0
        if (present(options)) then
            o%relaxpar = options%relaxpar
            o%ubound = options%ubound
            o%lbound = options%lbound
            o%damp = options%damp
            if (options%verbose) o%verbose = 1
            if (options%upb) o%upb = 1
            if (options%lwrb) o%lwrb = 1
            if (options%rowprobability) o%rowprobability = 1
        end if
    end function to_c

    integer function art_code(art_method)
        character(*), intent(in) :: art_method
        select case (trim(art_method))
        case ('kaczmarz')
            art_code = 0
        case ('symkaczmarz', 'sym')
            art_code = 1
        case ('randkaczmarz', 'rand')
            art_code = 2
        case default
            stop "Unknown ART method. Use kaczmarz, symkaczmarz or randkaczmarz"
        end select
    end function art_code

    ! art, AIRMs.f90:59-369.  With xref the device stops once the distance to xref grows
    ! (AIRMs.f90:322, 330-333); the random_number draws of randomized Kaczmarz are then taken as
    ! twin_art takes them: an upper bound drawn from a saved generator state, the state restored
    ! and advanced by exactly the number consumed.
    subroutine art(Xout, art_method, A, b, Kin, x0, xref, order, options)
        character(*) :: art_method
        class(csr_matrix), intent(in) :: A
        real(kind=8), intent(in) :: b(A%m)
        integer, intent(in) :: Kin(:)
        real(kind=8), intent(in), optional :: x0(A%n), xref(A%n)
        integer, optional :: order(:)
        class(opts), intent(in), optional :: options
        real(kind=8), intent(out) :: Xout(A%n, size(Kin))
        type(fair_opts_c) :: o
        type(c_ptr) :: h, px0, pord, pu
        logical :: temp
        integer(c_int32_t) :: meth, nord, ex
        integer(c_int32_t), allocatable, target :: ordc(:)
        integer(c_int32_t), allocatable :: kc(:)
        integer(c_int64_t) :: cnt, used
        real(c_double), allocatable, target :: u(:), xs(:)
        integer, allocatable :: state(:)
        integer :: ns
        meth = art_code(art_method)
        o = to_c(options)
        call borrow_dev(A, h, temp)
        pord = c_null_ptr
        nord = 0
        if (present(order)) then
            ordc = order
            nord = size(order)
            pord = c_loc(ordc)
        end if
        kc = Kin
        call fair_check(c_fair_art_uniform_count(h, meth, int(size(Kin), c_int32_t), kc, pord, nord, o, cnt))
        allocate (u(max(cnt, 1_c_int64_t)))
        if (present(xref)) then
            call random_seed(size=ns)
            allocate (state(ns))
            call random_seed(get=state)
        end if
        if (cnt > 0) call random_number(u(1:cnt))
        pu = c_null_ptr
        if (cnt > 0) pu = c_loc(u)
        px0 = c_null_ptr
        if (present(x0)) then
            xs = x0
            px0 = c_loc(xs)
        end if
        if (.not. present(xref)) then
            call fair_check(c_fair_art(h, meth, b, int(size(Kin), c_int32_t), kc, px0, pord, nord, o, pu, cnt, Xout))
        else
            call fair_check(c_fair_art_oracle(h, meth, b, int(size(Kin), c_int32_t), kc, px0, xref, pord, nord, o, &
                                              pu, cnt, Xout, ex, used))
            call random_seed(put=state)
            if (used > 0) call random_number(u(1:used))
        end if
        call release_dev(h, temp)
    end subroutine art

    ! twin_art, AIRMs.f90:371-811 ('errorgauge' | 'mutualstep').  Randomized Kaczmarz draws
    ! random_number values in the reference's order; the stopping rule decides how many are used,
    ! so the generator state is saved, an upper bound of draws is taken, and afterwards the state
    ! is restored and advanced by exactly the number consumed.
    subroutine twin_art(xout, final_it, Xstore, Ystore, EG, art_method, stoprule, A, b, Kin, x0, y0, &
                        order, slack, options)
        character(*) :: art_method, stoprule
        class(csr_matrix), intent(in) :: A
        real(kind=8), intent(in) :: b(A%m)
        integer, intent(in) :: Kin(:)
        real(kind=8), intent(in), optional, target :: x0(A%n), y0(A%n)
        integer, optional :: order(:), slack
        class(opts), intent(in), optional :: options
        real(kind=8), intent(out) :: xout(A%n)
        integer :: final_it
        real(kind=8), dimension(A%n, size(Kin)), intent(out) :: Xstore, Ystore
        real(kind=8), intent(out) :: EG(maxval(Kin))
        type(fair_opts_c) :: o
        type(c_ptr) :: h, px0, py0, pord, pu
        logical :: temp
        integer(c_int32_t) :: meth, rule, nord, sl, fi
        integer(c_int32_t), allocatable, target :: ordc(:)
        integer(c_int32_t), allocatable :: kc(:)
        integer(c_int64_t) :: cnt, used
        real(c_double), allocatable, target :: u(:), xs(:), ys(:)
        integer, allocatable :: state(:)
        integer :: ns
        meth = art_code(art_method)
        select case (stoprule)
        case ('errorgauge')
            rule = 0
        case ('mutualstep')
            rule = 1
        case default
            stop "Unknown stopping rule. Use errorgauge or mutualstep."
        end select
        sl = 7
        if (present(slack)) sl = slack
        o = to_c(options)
        call borrow_dev(A, h, temp)
        pord = c_null_ptr
        nord = 0
        if (present(order)) then
            ordc = order
            nord = size(order)
            pord = c_loc(ordc)
        end if
        kc = Kin
        call fair_check(c_fair_twin_art_uniform_bound(h, meth, int(size(Kin), c_int32_t), kc, pord, nord, o, cnt))
        call random_seed(size=ns)
        allocate (state(ns), u(max(cnt, 1_c_int64_t)))
        call random_seed(get=state)
        if (cnt > 0) call random_number(u(1:cnt))
        pu = c_null_ptr
        if (cnt > 0) pu = c_loc(u)
        px0 = c_null_ptr
        py0 = c_null_ptr
        if (present(x0)) then
            xs = x0
            px0 = c_loc(xs)
        end if
        if (present(y0)) then
            ys = y0
            py0 = c_loc(ys)
        end if
        call fair_check(c_fair_twin_art(h, meth, rule, b, int(size(Kin), c_int32_t), kc, px0, py0, pord, nord, sl, &
                                        o, pu, cnt, xout, fi, Xstore, Ystore, EG, used))
        final_it = fi
        call random_seed(put=state)
        if (used > 0) call random_number(u(1:used))
        call release_dev(h, temp)
    end subroutine twin_art

    ! basic_sweep, AIRMs.f90:813-858: the row steps over ord(1:nord), or with isrand and rowprob
    ! one random_number draw per step and the first row whose cumul reaches it (a binary search
    ! over the non-decreasing table finds the row the reference's linear search finds).  The rows
    ! are chosen here, in the reference's draw order; the steps run on the device as one launch.
    subroutine basic_sweep(x, A, b, ord, nord, isrand, rowprob, omg, nrA, cumul, upb, lwrb, ub, lb)
        class(csr_matrix), intent(in) :: A
        real(kind=8), intent(inout) :: x(A%n)
        real(kind=8), intent(in) :: b(A%m), cumul(A%m), nrA(A%m), omg, ub, lb
        integer, intent(in) :: ord(2*A%m), nord
        logical, intent(in) :: isrand, rowprob, upb, lwrb
        type(fair_opts_c) :: o
        type(c_ptr) :: h
        logical :: temp
        integer(c_int32_t), allocatable :: rows(:)
        real(c_double), allocatable :: xs(:)
        real(kind=8) :: u
        integer :: ii, lo, hi, mid
        allocate (rows(max(nord, 1)))
        do ii = 1, nord
            if (isrand .and. rowprob) then
                call random_number(u)
                lo = 1
                hi = A%m
                do while (lo < hi)
                    mid = (lo + hi)/2
                    if (cumul(mid) < u) then
                        lo = mid + 1
                    else
                        hi = mid
                    end if
                end do
                rows(ii) = lo
            else
                rows(ii) = ord(ii)
            end if
        end do
        o%relaxpar = omg
        o%ubound = ub
        o%lbound = lb
        o%damp = 0d0
        o%verbose = 0
        o%upb = merge(1, 0, upb)
        o%lwrb = merge(1, 0, lwrb)
        o%rowprobability = 0
        xs = x
        call borrow_dev(A, h, temp)
        call fair_check(c_fair_art_steps(h, rows, int(max(nord, 0), c_int64_t), b, nrA, o, xs))
        call release_dev(h, temp)
        x = xs
    end subroutine basic_sweep

    ! sirt, AIRMs.f90:997-1127 ('cimmino' | 'sart'); SART leaves A's rows column-sorted.
    subroutine sirt(xout, sirt_method, A, b, Kin, x0, options)
        character(*) :: sirt_method
        class(sparse_matrix) :: A
        real(kind=8), intent(in) :: b(A%m)
        integer, intent(in) :: Kin(:)
        real(kind=8), intent(in), optional :: x0(A%n)
        class(opts), intent(in), optional :: options
        real(kind=8) :: xout(A%n)
        type(fair_opts_c) :: o
        integer(c_int32_t) :: meth
        real(c_double), allocatable, target :: xs(:)
        type(c_ptr) :: px0
        select case (sirt_method)
        case ('cimmino')
            meth = 0
        case ('sart', 'sirt')
            meth = 1
        case default
            stop "Not a recognised SIRT method."
        end select
        o = to_c(options)
        px0 = c_null_ptr
        if (present(x0)) then
            if (size(x0) /= A%n) stop "Size of x0 should be equal to n (#columns)."
            xs = x0
            px0 = c_loc(xs)
        end if
        call fair_check(c_fair_sirt(A%device_handle(), meth, b, int(maxval(Kin), c_int32_t), px0, o, xout))
        if (meth == 1) call A%sync_host()
    end subroutine sirt

    ! normest, AIRMs.f90:1129-1173: random_number start vector, power iteration on the device
    function normest(A, tolerance, max_its) result(sigma)
        class(sparse_matrix) :: A
        real(kind=8), optional :: tolerance
        integer, optional :: max_its
        real(kind=8) :: sigma
        real(kind=8) :: v(A%n), tol
        integer :: maxit
        real(c_double) :: s
        tol = 1d-5
        if (present(tolerance)) tol = tolerance
        maxit = 100
        if (present(max_its)) maxit = max_its
        call random_number(v)
        call fair_check(c_fair_normest(A%device_handle(), v, real(tol, c_double), int(maxit, c_int32_t), s))
        sigma = s
    end function normest

    ! gradient_descent, AIRMs.f90:860-995
    subroutine gradient_descent(xout, A, b, max_iter, x0, rho_in, tolerance, options)
        class(sparse_matrix), intent(in) :: A
        real(kind=8), intent(in) :: b(A%m)
        real(kind=8), optional :: x0(A%n), rho_in, tolerance
        integer, optional :: max_iter
        class(opts), optional :: options
        real(kind=8), intent(out) :: xout(A%n)
        type(fair_opts_c) :: o
        type(c_ptr) :: h, px0
        logical :: temp
        integer :: maxiter
        real(kind=8) :: rho, tol, sigma, v(A%n)
        real(c_double) :: s
        real(c_double), allocatable, target :: xs(:)
        integer(c_int32_t) :: its
        maxiter = 100
        if (present(max_iter)) maxiter = max_iter
        rho = 1.5d0
        if (present(rho_in)) rho = rho_in
        tol = 1d-6
        if (present(tolerance)) then
            tol = tolerance
            if (tol < 0d0) stop "Must use positive tolerance."
        end if
        o = to_c(options)
        call borrow_dev(A, h, temp)
        call random_number(v)
        call fair_check(c_fair_normest(h, v, 1d-5, 100_c_int32_t, s))
        sigma = s
        px0 = c_null_ptr
        if (present(x0)) then
            xs = x0
            px0 = c_loc(xs)
        end if
        call fair_check(c_fair_gradient_descent(h, b, int(maxiter, c_int32_t), px0, real(rho, c_double), &
            real(tol, c_double), real(sigma, c_double), o, xout, its))
        call release_dev(h, temp)
    end subroutine gradient_descent

    ! scramble, AIRMs.f90:1175-1212: two passes of "swap i with a uniform j", values 1..N
    function scramble(number_of_values) result(array)
        integer, intent(in) :: number_of_values
        integer, allocatable :: array(:)
        integer :: i, j, k, tmp
        real(kind=8) :: u
        array = [(i, i=1, number_of_values)]
        do k = 1, 2
            do i = 1, number_of_values
                call random_number(u)
                j = 1 + floor(number_of_values*u)
                tmp = array(j)
                array(j) = array(i)
                array(i) = tmp
            end do
        end do
    end function scramble
end module AIRMs
