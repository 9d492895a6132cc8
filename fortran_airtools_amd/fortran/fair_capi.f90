! fair_capi.f90 -- ISO_C_BINDING interface to libfair_hip.so (include/fair_hip.h).
!
! The drop-in modules sparse_matrices, system_generation and airms call the gfx950 library
! only through this interface block.  Status codes: 0 = ok; fair_stop turns any other status
! into the reference's `stop "<message>"` behaviour with the library's message text.
module fair_capi
    use iso_c_binding
    implicit none
    private
    public :: fair_geom_c, fair_opts_c, fair_check, fair_ensure_init
    public :: c_fair_init, c_fair_fanlineartomo, c_fair_paralleltomo, c_fair_csr_create, &
              c_fair_csr_info, c_fair_csr_download, c_fair_csr_free, c_fair_matvecmul, &
              c_fair_transvecmul, c_fair_rowvecmult, c_fair_row_norms, c_fair_column_norms, &
              c_fair_transpose, c_fair_multiply_by_scalar, c_fair_add_scalar_mult_row, &
              c_fair_sirt, c_fair_normest, c_fair_gradient_descent, c_fair_art, &
              c_fair_art_uniform_count, c_fair_linspace, c_fair_twin_art, c_fair_twin_art_uniform_bound, &
              c_fair_art_oracle, c_fair_art_steps, c_fair_phantom_ellipses, c_fair_phantom_bumps, &
              c_fair_phantom_nearest, c_fair_phantom_last_disk, c_fair_phantom_stamp_disks, &
              c_fair_phantom_last_band, c_fair_normal_newton

    ! bind(C) mirror of type(data) without its allocatables (system_generation.f90:6-12)
    type, bind(C) :: fair_geom_c
        integer(c_int32_t) :: n, p, nth
        real(c_double) :: r, dw, sd
    end type fair_geom_c

    ! bind(C) mirror of type(opts) (AIRMs.f90:6-13)
    type, bind(C) :: fair_opts_c
        real(c_double) :: relaxpar, ubound, lbound, damp
        integer(c_int32_t) :: verbose, upb, lwrb, rowprobability
    end type fair_opts_c

    logical, save :: initialised = .false.

    interface
        function c_fair_last_error() bind(C, name='fair_last_error') result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function
        function c_fair_init(dev) bind(C, name='fair_init') result(s)
            import :: c_int
            integer(c_int), value :: dev
            integer(c_int) :: s
        end function
        ! phantom images (column-major n x n), include/fair_hip.h
        function c_fair_phantom_ellipses(n, X, Y, ne, ell, img) bind(C, name='fair_phantom_ellipses') result(s)
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: n, ne
            real(c_double), intent(in) :: X(*), Y(*), ell(*)
            real(c_double) :: img(*)
            integer(c_int) :: s
        end function
        function c_fair_phantom_bumps(n, X, Y, nb, bumps, power, img) bind(C, name='fair_phantom_bumps') result(s)
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: n, nb, power
            real(c_double), intent(in) :: X(*), Y(*), bumps(*)
            real(c_double) :: img(*)
            integer(c_int) :: s
        end function
        function c_fair_phantom_nearest(n, X, Y, np, pts, img) bind(C, name='fair_phantom_nearest') result(s)
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: n, np
            real(c_double), intent(in) :: X(*), Y(*), pts(*)
            real(c_double) :: img(*)
            integer(c_int) :: s
        end function
        function c_fair_phantom_last_disk(n, X, Y, nd, disks, img) bind(C, name='fair_phantom_last_disk') result(s)
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: n, nd
            real(c_double), intent(in) :: X(*), Y(*), disks(*)
            real(c_double) :: img(*)
            integer(c_int) :: s
        end function
        function c_fair_phantom_stamp_disks(n, X, Y, nd, disks, value, img) &
            bind(C, name='fair_phantom_stamp_disks') result(s)
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: n, nd
            real(c_double), intent(in) :: X(*), Y(*), disks(*)
            real(c_double), value :: value
            real(c_double) :: img(*)
            integer(c_int) :: s
        end function
        function c_fair_phantom_last_band(n, X, Y, ne, edges, img) bind(C, name='fair_phantom_last_band') result(s)
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: n, ne
            real(c_double), intent(in) :: X(*), Y(*), edges(*)
            real(c_double) :: img(*)
            integer(c_int) :: s
        end function
        function c_fair_normal_newton(m, u, e) bind(C, name='fair_normal_newton') result(s)
            import :: c_int, c_int64_t, c_double
            integer(c_int64_t), value :: m
            real(c_double), intent(in) :: u(*)
            real(c_double) :: e(*)
            integer(c_int) :: s
        end function
        function c_fair_linspace(a, b, n, x) bind(C, name='fair_linspace') result(s)
            import :: c_int, c_int32_t, c_double
            real(c_double), value :: a, b
            integer(c_int32_t), value :: n
            real(c_double) :: x(*)
            integer(c_int) :: s
        end function
        function c_fair_fanlineartomo(g, theta, h) bind(C, name='fair_fanlineartomo') result(s)
            import :: c_int, c_double, c_ptr, fair_geom_c
            type(fair_geom_c), intent(in) :: g
            real(c_double), intent(in) :: theta(*)
            type(c_ptr), intent(out) :: h
            integer(c_int) :: s
        end function
        function c_fair_paralleltomo(g, theta, h) bind(C, name='fair_paralleltomo') result(s)
            import :: c_int, c_double, c_ptr, fair_geom_c
            type(fair_geom_c), intent(in) :: g
            real(c_double), intent(in) :: theta(*)
            type(c_ptr), intent(out) :: h
            integer(c_int) :: s
        end function
        function c_fair_csr_create(m, n, nnz, IA, JA, vA, h) bind(C, name='fair_csr_create') result(s)
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            integer(c_int32_t), value :: m, n
            integer(c_int64_t), value :: nnz
            integer(c_int32_t), intent(in) :: IA(*), JA(*)
            real(c_double), intent(in) :: vA(*)
            type(c_ptr), intent(out) :: h
            integer(c_int) :: s
        end function
        function c_fair_csr_info(h, m, n, nnz) bind(C, name='fair_csr_info') result(s)
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: m, n
            integer(c_int64_t), intent(out) :: nnz
            integer(c_int) :: s
        end function
        function c_fair_csr_download(h, IA, JA, vA) bind(C, name='fair_csr_download') result(s)
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t) :: IA(*), JA(*)
            real(c_double) :: vA(*)
            integer(c_int) :: s
        end function
        function c_fair_csr_free(h) bind(C, name='fair_csr_free') result(s)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int) :: s
        end function
        function c_fair_matvecmul(h, v, nv, w, nw) bind(C, name='fair_matvecmul') result(s)
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(in) :: v(*)
            integer(c_int32_t), value :: nv, nw
            real(c_double) :: w(*)
            integer(c_int) :: s
        end function
        function c_fair_transvecmul(h, v, nv, w, nw) bind(C, name='fair_transvecmul') result(s)
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(in) :: v(*)
            integer(c_int32_t), value :: nv, nw
            real(c_double) :: w(*)
            integer(c_int) :: s
        end function
        function c_fair_rowvecmult(h, k, v, nv, f) bind(C, name='fair_rowvecmult') result(s)
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: k, nv
            real(c_double), intent(in) :: v(*)
            real(c_double), intent(out) :: f
            integer(c_int) :: s
        end function
        function c_fair_row_norms(h, p, nrA, nm) bind(C, name='fair_row_norms') result(s)
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: p, nm
            real(c_double) :: nrA(*)
            integer(c_int) :: s
        end function
        function c_fair_column_norms(h, p, ncA, nn) bind(C, name='fair_column_norms') result(s)
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: p, nn
            real(c_double) :: ncA(*)
            integer(c_int) :: s
        end function
        function c_fair_transpose(h) bind(C, name='fair_transpose') result(s)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int) :: s
        end function
        function c_fair_multiply_by_scalar(h, c) bind(C, name='fair_multiply_by_scalar') result(s)
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), value :: c
            integer(c_int) :: s
        end function
        function c_fair_add_scalar_mult_row(h, k, c, w, nw, lwrb, upb) &
            bind(C, name='fair_add_scalar_mult_row') result(s)
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: k, nw
            real(c_double), value :: c
            real(c_double) :: w(*)
            type(c_ptr), value :: lwrb, upb
            integer(c_int) :: s
        end function
        function c_fair_sirt(h, method, b, maxits, x0, opts, xout) bind(C, name='fair_sirt') result(s)
            import :: c_int, c_int32_t, c_double, c_ptr, fair_opts_c
            type(c_ptr), value :: h
            integer(c_int32_t), value :: method, maxits
            real(c_double), intent(in) :: b(*)
            type(c_ptr), value :: x0
            type(fair_opts_c), intent(in) :: opts
            real(c_double) :: xout(*)
            integer(c_int) :: s
        end function
        function c_fair_normest(h, v0, tol, maxit, sigma) bind(C, name='fair_normest') result(s)
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(in) :: v0(*)
            real(c_double), value :: tol
            integer(c_int32_t), value :: maxit
            real(c_double), intent(out) :: sigma
            integer(c_int) :: s
        end function
        function c_fair_gradient_descent(h, b, maxiter, x0, rho, tol, sigma, opts, xout, iters) &
            bind(C, name='fair_gradient_descent') result(s)
            import :: c_int, c_int32_t, c_double, c_ptr, fair_opts_c
            type(c_ptr), value :: h
            real(c_double), intent(in) :: b(*)
            integer(c_int32_t), value :: maxiter
            type(c_ptr), value :: x0
            real(c_double), value :: rho, tol, sigma
            type(fair_opts_c), intent(in) :: opts
            real(c_double) :: xout(*)
            integer(c_int32_t), intent(out) :: iters
            integer(c_int) :: s
        end function
        function c_fair_art(h, method, b, nk, kin, x0, order, nord, opts, u, nu, Xout) &
            bind(C, name='fair_art') result(s)
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, fair_opts_c
            type(c_ptr), value :: h
            integer(c_int32_t), value :: method, nk, nord
            real(c_double), intent(in) :: b(*)
            integer(c_int32_t), intent(in) :: kin(*)
            type(c_ptr), value :: x0, order, u
            type(fair_opts_c), intent(in) :: opts
            integer(c_int64_t), value :: nu
            real(c_double) :: Xout(*)
            integer(c_int) :: s
        end function
        function c_fair_art_steps(h, rows, nsteps, b, nrA, opts, x) bind(C, name='fair_art_steps') result(s)
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, fair_opts_c
            type(c_ptr), value :: h
            integer(c_int32_t), intent(in) :: rows(*)
            integer(c_int64_t), value :: nsteps
            real(c_double), intent(in) :: b(*), nrA(*)
            type(fair_opts_c), intent(in) :: opts
            real(c_double) :: x(*)
            integer(c_int) :: s
        end function
        function c_fair_art_oracle(h, method, b, nk, kin, x0, xref, order, nord, opts, u, nu, Xout, exit_iter, &
                                   consumed) bind(C, name='fair_art_oracle') result(s)
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, fair_opts_c
            type(c_ptr), value :: h
            integer(c_int32_t), value :: method, nk, nord
            real(c_double), intent(in) :: b(*), xref(*)
            integer(c_int32_t), intent(in) :: kin(*)
            type(c_ptr), value :: x0, order, u
            type(fair_opts_c), intent(in) :: opts
            integer(c_int64_t), value :: nu
            real(c_double) :: Xout(*)
            integer(c_int32_t) :: exit_iter
            integer(c_int64_t) :: consumed
            integer(c_int) :: s
        end function
        function c_fair_twin_art(h, method, stoprule, b, nk, kin, x0, y0, order, nord, slack, opts, u, nu, &
                                 xout, final_it, xstore, ystore, eg, consumed) bind(C, name='fair_twin_art') result(s)
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, fair_opts_c
            type(c_ptr), value :: h, x0, y0, order, u
            integer(c_int32_t), value :: method, stoprule, nk, nord, slack
            real(c_double), intent(in) :: b(*)
            integer(c_int32_t), intent(in) :: kin(*)
            type(fair_opts_c), intent(in) :: opts
            integer(c_int64_t), value :: nu
            real(c_double) :: xout(*), xstore(*), ystore(*), eg(*)
            integer(c_int32_t) :: final_it
            integer(c_int64_t) :: consumed
            integer(c_int) :: s
        end function
        function c_fair_twin_art_uniform_bound(h, method, nk, kin, order, nord, opts, cnt) &
            bind(C, name='fair_twin_art_uniform_bound') result(s)
            import :: c_int, c_int32_t, c_int64_t, c_ptr, fair_opts_c
            type(c_ptr), value :: h
            integer(c_int32_t), value :: method, nk, nord
            integer(c_int32_t), intent(in) :: kin(*)
            type(c_ptr), value :: order
            type(fair_opts_c), intent(in) :: opts
            integer(c_int64_t), intent(out) :: cnt
            integer(c_int) :: s
        end function
        function c_fair_art_uniform_count(h, method, nk, kin, order, nord, opts, cnt) &
            bind(C, name='fair_art_uniform_count') result(s)
            import :: c_int, c_int32_t, c_int64_t, c_ptr, fair_opts_c
            type(c_ptr), value :: h
            integer(c_int32_t), value :: method, nk, nord
            integer(c_int32_t), intent(in) :: kin(*)
            type(c_ptr), value :: order
            type(fair_opts_c), intent(in) :: opts
            integer(c_int64_t), intent(out) :: cnt
            integer(c_int) :: s
        end function
    end interface

contains

    ! First call: select device 0 (the environment variable FAIR_DEVICE may name another).
    subroutine fair_ensure_init()
        character(len=16) :: buf
        integer :: dev, ln, st
        if (initialised) return
        dev = 0
        call get_environment_variable('FAIR_DEVICE', buf, ln, st)
        if (st == 0 .and. ln > 0) read (buf, *) dev
        call fair_check(c_fair_init(int(dev, c_int)))
        initialised = .true.
    end subroutine fair_ensure_init

    ! A non-zero status stops the program with the library's message, as the reference's
    ! `stop "<message>"` does.
    subroutine fair_check(status)
        integer(c_int), intent(in) :: status
        character(kind=c_char), pointer :: msg(:)
        character(len=:), allocatable :: text
        integer :: i, ln
        type(c_ptr) :: p
        if (status == 0) return
        p = c_fair_last_error()
        call c_f_pointer(p, msg, [4096])
        ln = 0
        do i = 1, 4096
            if (msg(i) == c_null_char) exit
            ln = i
        end do
        allocate (character(len=ln) :: text)
        do i = 1, ln
            text(i:i) = msg(i)
        end do
        stop text
    end subroutine fair_check
end module fair_capi
