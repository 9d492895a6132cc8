! fair_extensions.f90 -- Fortran access to the additions that stand beside the drop-in path:
! the FP32 precision variant of sirt, sirt over many right-hand sides at once (batched, with
! an autotuner), the matrix-free fan-beam projector, and bt = -log(T) of measured data on the
! device.  None has a counterpart in the reference's modules; the first three answer its to-dos
! "blocking support" and "make matrix-free implementations" (README.md:6-7), and all keep the
! reference's argument conventions.
module fair_extensions
    use iso_c_binding
    use fair_capi
    use sparse_matrices
    use system_generation
    use AIRMs
    implicit none
    private
    public :: sirt_f32, sirt_batch, autotune_batch, fan_projector, absorbance

    ! A of fanlineartomo(dat), applied without storing it (FP32 values, FP64 ray geometry)
    type fan_projector
        type(c_ptr) :: h = c_null_ptr
        integer :: m = 0, n = 0
    contains
        procedure :: create => mf_create
        procedure :: matvecmul => mf_matvecmul
        procedure :: transvecmul => mf_transvecmul
        procedure :: segments => mf_segments
        procedure :: normest => mf_normest
        procedure :: gradient_descent => mf_gradient_descent
        procedure :: free => mf_free
    end type fan_projector

    interface
        function c_fair_absorbance(T, p, nth, skip, b) bind(C, name='fair_absorbance') result(s)
            import :: c_int, c_int64_t, c_double
            real(c_double), intent(in) :: T(*)
            integer(c_int64_t), value :: p, nth, skip
            real(c_double) :: b(*)
            integer(c_int) :: s
        end function
        function c_fair_sirt_f32(A, method, b, maxits, x0, o, xout) bind(C, name='fair_sirt_f32') result(s)
            import :: c_ptr, c_int, c_int32_t, c_float, fair_opts_c
            type(c_ptr), value :: A, x0
            integer(c_int32_t), value :: method, maxits
            real(c_float), intent(in) :: b(*)
            type(fair_opts_c), intent(in) :: o
            real(c_float) :: xout(*)
            integer(c_int) :: s
        end function
        function c_fair_sirt_batch(A, method, B, k, maxits, X0, o, use_graph, Xout) &
            bind(C, name='fair_sirt_batch') result(s)
            import :: c_ptr, c_int, c_int32_t, c_float, fair_opts_c
            type(c_ptr), value :: A, X0
            integer(c_int32_t), value :: method, k, maxits, use_graph
            real(c_float), intent(in) :: B(*)
            type(fair_opts_c), intent(in) :: o
            real(c_float) :: Xout(*)
            integer(c_int) :: s
        end function
        function c_fair_autotune_batch(A, k, kc) bind(C, name='fair_autotune_batch') result(s)
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: A
            integer(c_int32_t), value :: k
            integer(c_int32_t) :: kc(2)
            integer(c_int) :: s
        end function
        function c_fair_mf_create(g, theta, h) bind(C, name='fair_mf_create') result(s)
            import :: c_ptr, c_int, c_double, fair_geom_c
            type(fair_geom_c), intent(in) :: g
            real(c_double), intent(in) :: theta(*)
            type(c_ptr), intent(out) :: h
            integer(c_int) :: s
        end function
        function c_fair_mf_free(h) bind(C, name='fair_mf_free') result(s)
            import :: c_ptr, c_int
            type(c_ptr), value :: h
            integer(c_int) :: s
        end function
        function c_fair_mf_forward(h, x, y) bind(C, name='fair_mf_forward') result(s)
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: h
            real(c_float), intent(in) :: x(*)
            real(c_float) :: y(*)
            integer(c_int) :: s
        end function
        function c_fair_mf_adjoint(h, y, x) bind(C, name='fair_mf_adjoint') result(s)
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: h
            real(c_float), intent(in) :: y(*)
            real(c_float) :: x(*)
            integer(c_int) :: s
        end function
        function c_fair_mf_segments(h, c) bind(C, name='fair_mf_segments') result(s)
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: h
            integer(c_int64_t) :: c
            integer(c_int) :: s
        end function
        function c_fair_mf_normest(h, v0, tol, maxit, sigma) bind(C, name='fair_mf_normest') result(s)
            import :: c_ptr, c_int, c_int32_t, c_float, c_double
            type(c_ptr), value :: h
            real(c_float), intent(in) :: v0(*)
            real(c_double), value :: tol
            integer(c_int32_t), value :: maxit
            real(c_double) :: sigma
            integer(c_int) :: s
        end function
        function c_fair_mf_gd(h, b, maxiter, x0, rho, tol, sigma, o, xout, iters) &
            bind(C, name='fair_mf_gd') result(s)
            import :: c_ptr, c_int, c_int32_t, c_float, c_double, fair_opts_c
            type(c_ptr), value :: h, x0
            real(c_float), intent(in) :: b(*)
            integer(c_int32_t), value :: maxiter
            real(c_double), value :: rho, tol, sigma
            type(fair_opts_c), intent(in) :: o
            real(c_float) :: xout(*)
            integer(c_int32_t) :: iters
            integer(c_int) :: s
        end function
    end interface

contains

    integer(c_int32_t) function sirt_code(sirt_method)
        character(*), intent(in) :: sirt_method
        select case (sirt_method)
        case ('cimmino')
            sirt_code = 0
        case ('sart', 'sirt')
            sirt_code = 1
        case default
            stop "Not a recognised SIRT method."
        end select
    end function sirt_code

    ! sirt (AIRMs.f90:997-1127) with FP32 matrix values, iterate and accumulation; the weights
    ! M, D stay FP64.  Arguments as sirt's, vectors real(kind=4).
    subroutine sirt_f32(xout, sirt_method, A, b, Kin, x0, options)
        character(*) :: sirt_method
        class(sparse_matrix) :: A
        real(kind=4), intent(in) :: b(A%m)
        integer, intent(in) :: Kin(:)
        real(kind=4), intent(in), optional, target :: x0(A%n)
        class(opts), intent(in), optional :: options
        real(kind=4) :: xout(A%n)
        type(c_ptr) :: px0
        integer(c_int32_t) :: meth
        meth = sirt_code(sirt_method)
        px0 = c_null_ptr
        if (present(x0)) px0 = c_loc(x0)
        call fair_check(c_fair_sirt_f32(A%device_handle(), meth, b, int(maxval(Kin), c_int32_t), px0, &
            to_c(options), xout))
        if (meth == 1) call A%sync_host()
    end subroutine sirt_f32

    ! sirt over the k columns of B(m, k) at once (k slices sharing A), FP32 as sirt_f32;
    ! Xout(n, k).  use_graph replays the whole run from a captured HIP graph.
    subroutine sirt_batch(Xout, sirt_method, A, B, Kin, X0, options, use_graph)
        character(*) :: sirt_method
        class(sparse_matrix) :: A
        real(kind=4), intent(in) :: B(:, :)
        integer, intent(in) :: Kin(:)
        real(kind=4), intent(in), optional :: X0(:, :)
        class(opts), intent(in), optional :: options
        logical, intent(in), optional :: use_graph
        real(kind=4) :: Xout(:, :)
        real(c_float), allocatable, target :: Bt(:, :), Xt(:, :), X0t(:, :)
        type(c_ptr) :: px0
        integer(c_int32_t) :: meth, k, g
        meth = sirt_code(sirt_method)
        k = size(B, 2)
        if (size(B, 1) /= A%m) stop "Vector input size incorrect."
        if (size(Xout, 1) /= A%n .or. size(Xout, 2) /= k) stop "Vector input size incorrect."
        ! the library's panels are row-major (entry (i, j) at i*k + j): transposed Fortran arrays
        Bt = transpose(B)
        allocate (Xt(k, A%n))
        px0 = c_null_ptr
        if (present(X0)) then
            X0t = transpose(X0)
            px0 = c_loc(X0t)
        end if
        g = 0
        if (present(use_graph)) then
            if (use_graph) g = 1
        end if
        call fair_check(c_fair_sirt_batch(A%device_handle(), meth, Bt, k, int(maxval(Kin), c_int32_t), px0, &
            to_c(options), g, Xt))
        Xout = transpose(Xt)
        if (meth == 1) call A%sync_host()
    end subroutine sirt_batch

    ! Times the batched products for k right-hand sides and keeps the fastest shape for A;
    ! kc receives the chosen right-hand sides per wavefront for A*X and A^T*R.
    subroutine autotune_batch(A, k, kc)
        class(sparse_matrix) :: A
        integer, intent(in) :: k
        integer, intent(out) :: kc(2)
        integer(c_int32_t) :: c(2)
        call fair_check(c_fair_autotune_batch(A%device_handle(), int(k, c_int32_t), c))
        kc = c
    end subroutine autotune_batch

    subroutine mf_create(P, dat)
        class(fan_projector), intent(inout) :: P
        class(data), intent(in) :: dat
        real(c_double), allocatable :: th(:)
        type(fair_geom_c) :: g
        call fair_ensure_init()
        if (c_associated(P%h)) call P%free()
        g = geom_of(dat)
        th = dat%theta
        call fair_check(c_fair_mf_create(g, th, P%h))
        P%m = dat%nth * dat%p
        P%n = dat%n**2
    end subroutine mf_create

    function mf_matvecmul(P, x) result(y)
        class(fan_projector), intent(in) :: P
        real(kind=4), intent(in) :: x(:)
        real(kind=4) :: y(P%m)
        if (size(x) /= P%n) stop "Vector input size incorrect."
        call fair_check(c_fair_mf_forward(P%h, x, y))
    end function mf_matvecmul

    function mf_transvecmul(P, y) result(x)
        class(fan_projector), intent(in) :: P
        real(kind=4), intent(in) :: y(:)
        real(kind=4) :: x(P%n)
        if (size(y) /= P%m) stop "Vector input size incorrect."
        call fair_check(c_fair_mf_adjoint(P%h, y, x))
    end function mf_transvecmul

    function mf_segments(P) result(c)
        class(fan_projector), intent(in) :: P
        integer(kind=8) :: c
        integer(c_int64_t) :: cc
        call fair_check(c_fair_mf_segments(P%h, cc))
        c = cc
    end function mf_segments

    ! normest (AIRMs.f90:1129-1173) on the operator, random_number start vector
    function mf_normest(P, tolerance, max_its) result(sigma)
        class(fan_projector), intent(in) :: P
        real(kind=8), optional :: tolerance
        integer, optional :: max_its
        real(kind=8) :: sigma
        real(kind=4) :: v(P%n)
        real(c_double) :: s, tol
        integer(c_int32_t) :: maxit
        tol = 1d-5
        if (present(tolerance)) tol = tolerance
        maxit = 100
        if (present(max_its)) maxit = max_its
        call random_number(v)
        call fair_check(c_fair_mf_normest(P%h, v, tol, maxit, s))
        sigma = s
    end function mf_normest

    ! gradient_descent (AIRMs.f90:860-995) on the operator: same arguments and control flow
    ! (r = -b first, exit before accepting a step shorter than tolerance), FP32 vectors.
    subroutine mf_gradient_descent(P, xout, b, max_iter, x0, rho_in, tolerance, options)
        class(fan_projector), intent(in) :: P
        real(kind=4), intent(in) :: b(:)
        real(kind=4), intent(out) :: xout(:)
        integer, optional :: max_iter
        real(kind=4), optional, target :: x0(:)
        real(kind=8), optional :: rho_in, tolerance
        class(opts), optional :: options
        integer(c_int32_t) :: maxiter, its
        real(c_double) :: rho, tol
        type(c_ptr) :: px0
        if (size(b) /= P%m .or. size(xout) /= P%n) stop "Vector input size incorrect."
        maxiter = 100
        if (present(max_iter)) maxiter = max_iter
        rho = 1.5d0
        if (present(rho_in)) rho = rho_in
        tol = 1d-6
        if (present(tolerance)) then
            tol = tolerance
            if (tol < 0d0) stop "Must use positive tolerance."
        end if
        px0 = c_null_ptr
        if (present(x0)) px0 = c_loc(x0)
        call fair_check(c_fair_mf_gd(P%h, b, maxiter, px0, rho, tol, real(P%normest(), c_double), &
            to_c(options), xout, its))
    end subroutine mf_gradient_descent

    subroutine mf_free(P)
        class(fan_projector), intent(inout) :: P
        if (c_associated(P%h)) call fair_check(c_fair_mf_free(P%h))
        P%h = c_null_ptr
    end subroutine mf_free


    ! bt = -log(T) of main.f90:270 on the device, after the angle cull of main.f90:229-245 when
    ! skip > 0 (T holds p values per angle for nth angles).
    function absorbance(T, p, nth, skip) result(b)
        real(kind=8), intent(in) :: T(:)
        integer, intent(in) :: p, nth, skip
        real(kind=8), allocatable :: b(:)
        integer :: keep
        if (size(T) /= p*nth) stop "Data has the wrong size, or p and nth are incorrect."
        keep = nth
        if (skip > 0 .and. nth > 0) keep = (nth - 1)/skip + 1
        allocate (b(keep*p))
        call fair_ensure_init()
        call fair_check(c_fair_absorbance(T, int(p, c_int64_t), int(nth, c_int64_t), int(skip, c_int64_t), b))
    end function absorbance
end module fair_extensions
