! oracle/ref_harness.f90 -- C-callable driver around the reference modules.
!
! TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the reference's own
! sparse_matrices / system_generation / AIRMs / phantomgallery sources (compiled where they
! lie, never copied) into oracle/_ref/libfair_ref.so.  Golden-vector scripts under
! tests/golden/ and bench.py's CPU-baseline leg load it; the product never does.
!
! Matrices live in module-level slots so that Python can drive the reference API
! (fanlineartomo, paralleltomo, matvecmul, transvecmul, row_norms, transpose, sirt,
! gradient_descent, normest, art) exactly as a Fortran caller would.
module ref_harness
    use iso_c_binding
    use sparse_matrices
    use system_generation
    use airms
    use phantomgallery
    implicit none
    integer, parameter :: NSLOT = 8
    type(csr_matrix), save :: mats(0:NSLOT-1)   ! slots 0 .. NSLOT-1, as the Python side numbers them
    ! data read by ref_read_data / ref_file_mode, and the file mode's right-hand side
    type(data), save :: fm_dat
    real(c_double), allocatable, save :: fm_b(:)
contains

    subroutine make_geom(dat, n, p, nth, r, dw, sd, theta)
        type(data), intent(out) :: dat
        integer(c_int), value :: n, p, nth
        real(c_double), value :: r, dw, sd
        real(c_double), intent(in) :: theta(nth)
        dat%n = n
        dat%p = p
        dat%nth = nth
        dat%r = r
        dat%dw = dw
        dat%sd = sd
        allocate (dat%theta(nth))
        dat%theta = theta
    end subroutine make_geom

    ! kind 0 = fanlineartomo, 1 = paralleltomo.  Returns nnz.
    function ref_build(slot, kind, n, p, nth, r, dw, sd, theta) bind(C, name='ref_build') result(nnz)
        integer(c_int), value :: slot, kind, n, p, nth
        real(c_double), value :: r, dw, sd
        real(c_double), intent(in) :: theta(nth)
        integer(c_int) :: nnz
        type(data) :: dat
        type(csr_matrix) :: tmp
        call make_geom(dat, n, p, nth, r, dw, sd, theta)
        if (kind == 0) then
            tmp = fanlineartomo(dat)
        else
            tmp = paralleltomo(dat)
        end if
        call take(slot, tmp)
        nnz = mats(slot)%nnz
    end function ref_build

    subroutine take(slot, tmp)
        integer(c_int), value :: slot
        type(csr_matrix), intent(inout) :: tmp
        if (allocated(mats(slot)%IA)) deallocate (mats(slot)%IA)
        if (allocated(mats(slot)%JA)) deallocate (mats(slot)%JA)
        if (allocated(mats(slot)%vA)) deallocate (mats(slot)%vA)
        mats(slot)%m = tmp%m
        mats(slot)%n = tmp%n
        mats(slot)%nnz = tmp%nnz
        call move_alloc(tmp%IA, mats(slot)%IA)
        call move_alloc(tmp%JA, mats(slot)%JA)
        call move_alloc(tmp%vA, mats(slot)%vA)
    end subroutine take

    subroutine ref_info(slot, m, n, nnz) bind(C, name='ref_info')
        integer(c_int), value :: slot
        integer(c_int), intent(out) :: m, n, nnz
        m = mats(slot)%m
        n = mats(slot)%n
        nnz = mats(slot)%nnz
    end subroutine ref_info

    subroutine ref_copy_out(slot, nia, IA, JA, vA) bind(C, name='ref_copy_out')
        integer(c_int), value :: slot, nia
        integer(c_int), intent(out) :: IA(nia), JA(*)
        real(c_double), intent(out) :: vA(*)
        integer :: nnz
        nnz = mats(slot)%nnz
        IA = mats(slot)%IA
        JA(1:nnz) = mats(slot)%JA
        vA(1:nnz) = mats(slot)%vA
    end subroutine ref_copy_out

    subroutine ref_set(slot, m, n, nnz, IA, JA, vA) bind(C, name='ref_set')
        integer(c_int), value :: slot, m, n, nnz
        integer(c_int), intent(in) :: IA(m + 1), JA(nnz)
        real(c_double), intent(in) :: vA(nnz)
        if (allocated(mats(slot)%IA)) deallocate (mats(slot)%IA)
        if (allocated(mats(slot)%JA)) deallocate (mats(slot)%JA)
        if (allocated(mats(slot)%vA)) deallocate (mats(slot)%vA)
        mats(slot)%m = m
        mats(slot)%n = n
        mats(slot)%nnz = nnz
        allocate (mats(slot)%IA(0:m), mats(slot)%JA(nnz), mats(slot)%vA(nnz))
        mats(slot)%IA(0:m) = IA
        mats(slot)%JA = JA
        mats(slot)%vA = vA
    end subroutine ref_set

    subroutine ref_matvec(slot, n, x, m, y) bind(C, name='ref_matvec')
        integer(c_int), value :: slot, n, m
        real(c_double), intent(in) :: x(n)
        real(c_double), intent(out) :: y(m)
        y = mats(slot)%matvecmul(x)
    end subroutine ref_matvec

    subroutine ref_transvec(slot, m, y, n, x) bind(C, name='ref_transvec')
        integer(c_int), value :: slot, m, n
        real(c_double), intent(in) :: y(m)
        real(c_double), intent(out) :: x(n)
        x = mats(slot)%transvecmul(y)
    end subroutine ref_transvec

    subroutine ref_rowvecmult(slot, k, n, v, f) bind(C, name='ref_rowvecmult')
        integer(c_int), value :: slot, k, n
        real(c_double), intent(in) :: v(n)
        real(c_double), intent(out) :: f
        f = mats(slot)%rowvecmult(k, v)
    end subroutine ref_rowvecmult

    subroutine ref_row_norms(slot, p, m, out) bind(C, name='ref_row_norms')
        integer(c_int), value :: slot, p, m
        real(c_double), intent(out) :: out(m)
        out = mats(slot)%row_norms(p)
    end subroutine ref_row_norms

    subroutine ref_transpose(slot) bind(C, name='ref_transpose')
        integer(c_int), value :: slot
        call mats(slot)%transpose()
    end subroutine ref_transpose

    subroutine ref_scale(slot, c) bind(C, name='ref_scale')
        integer(c_int), value :: slot
        real(c_double), value :: c
        real(kind=8) :: cc
        cc = c
        call mats(slot)%multiply_by_scalar(cc)
    end subroutine ref_scale

    subroutine ref_add_mult_row(slot, k, c, n, x, lwrb, lb, upb, ub) bind(C, name='ref_add_mult_row')
        integer(c_int), value :: slot, k, n, lwrb, upb
        real(c_double), value :: c, lb, ub
        real(c_double), intent(inout) :: x(n)
        real(kind=8) :: cc
        cc = c
        if (lwrb /= 0 .and. upb /= 0) then
            call mats(slot)%add_scalar_mult_row(k, cc, x, lwrb=lb, upb=ub)
        else if (lwrb /= 0) then
            call mats(slot)%add_scalar_mult_row(k, cc, x, lwrb=lb)
        else if (upb /= 0) then
            call mats(slot)%add_scalar_mult_row(k, cc, x, upb=ub)
        else
            call mats(slot)%add_scalar_mult_row(k, cc, x)
        end if
    end subroutine ref_add_mult_row

    subroutine fill_opts(options, upb, lwrb, ub, lb, relax, damp, rowprob)
        type(opts), intent(out) :: options
        integer(c_int), value :: upb, lwrb, rowprob
        real(c_double), value :: ub, lb, relax, damp
        call options%initialise()
        if (upb /= 0) call options%set_ubound(ub)
        if (lwrb /= 0) call options%set_lbound(lb)
        options%relaxpar = relax
        options%damp = damp
        options%rowprobability = (rowprob /= 0)
    end subroutine fill_opts

    ! method 0 = cimmino, 1 = sart.  Note: SART re-sorts the slot's matrix in place, as the
    ! reference does.
    subroutine ref_sirt(slot, method, m, b, maxits, hasx0, n, x0, upb, lwrb, ub, lb, relax, xout) &
        bind(C, name='ref_sirt')
        integer(c_int), value :: slot, method, m, maxits, hasx0, n, upb, lwrb
        real(c_double), value :: ub, lb, relax
        real(c_double), intent(in) :: b(m), x0(n)
        real(c_double), intent(out) :: xout(n)
        type(opts) :: options
        real(kind=8) :: xo(n)
        call fill_opts(options, upb, lwrb, ub, lb, relax, 0d0, 0)
        if (method == 0) then
            if (hasx0 /= 0) then
                call sirt(xo, 'cimmino', mats(slot), b, (/maxits/), x0=x0, options=options)
            else
                call sirt(xo, 'cimmino', mats(slot), b, (/maxits/), options=options)
            end if
        else
            if (hasx0 /= 0) then
                call sirt(xo, 'sart', mats(slot), b, (/maxits/), x0=x0, options=options)
            else
                call sirt(xo, 'sart', mats(slot), b, (/maxits/), options=options)
            end if
        end if
        xout = xo
    end subroutine ref_sirt

    subroutine ref_gd(slot, m, b, maxiter, hasx0, n, x0, rho, tol, upb, lwrb, ub, lb, xout) &
        bind(C, name='ref_gd')
        integer(c_int), value :: slot, m, maxiter, hasx0, n, upb, lwrb
        real(c_double), value :: rho, tol, ub, lb
        real(c_double), intent(in) :: b(m), x0(n)
        real(c_double), intent(out) :: xout(n)
        type(opts) :: options
        real(kind=8) :: xo(n), rr, tt
        integer :: mi
        call fill_opts(options, upb, lwrb, ub, lb, 1d0, 0d0, 0)
        rr = rho
        tt = tol
        mi = maxiter
        if (hasx0 /= 0) then
            call gradient_descent(xo, mats(slot), b, max_iter=mi, x0=x0, rho_in=rr, tolerance=tt, options=options)
        else
            call gradient_descent(xo, mats(slot), b, max_iter=mi, rho_in=rr, tolerance=tt, options=options)
        end if
        xout = xo
    end subroutine ref_gd

    function ref_normest(slot, tol, maxit) bind(C, name='ref_normest') result(sigma)
        integer(c_int), value :: slot, maxit
        real(c_double), value :: tol
        real(c_double) :: sigma
        real(kind=8) :: tt
        integer :: mi
        tt = tol
        mi = maxit
        sigma = normest(mats(slot), tt, mi)
    end function ref_normest

    ! method 0 = kaczmarz, 1 = symkaczmarz, 2 = randkaczmarz.
    subroutine ref_art(slot, method, m, b, nk, kin, hasx0, n, x0, hasorder, nord, order, &
                       upb, lwrb, ub, lb, relax, damp, rowprob, xout) bind(C, name='ref_art')
        integer(c_int), value :: slot, method, m, nk, hasx0, n, hasorder, nord, upb, lwrb, rowprob
        real(c_double), value :: ub, lb, relax, damp
        real(c_double), intent(in) :: b(m), x0(n)
        integer(c_int), intent(in) :: kin(nk), order(*)
        real(c_double), intent(out) :: xout(n, nk)
        type(opts) :: options
        character(len=12) :: meth
        real(kind=8) :: xo(n, nk)
        integer, allocatable :: ordv(:)
        call fill_opts(options, upb, lwrb, ub, lb, relax, damp, rowprob)
        select case (method)
        case (0)
            meth = 'kaczmarz'
        case (1)
            meth = 'symkaczmarz'
        case default
            meth = 'randkaczmarz'
        end select
        xo = 0d0
        if (hasorder /= 0) then
            allocate (ordv(nord))
            ordv = order(1:nord)
            if (hasx0 /= 0) then
                call art(xo, trim(meth), mats(slot), b, kin, x0=x0, order=ordv, options=options)
            else
                call art(xo, trim(meth), mats(slot), b, kin, order=ordv, options=options)
            end if
        else
            if (hasx0 /= 0) then
                call art(xo, trim(meth), mats(slot), b, kin, x0=x0, options=options)
            else
                call art(xo, trim(meth), mats(slot), b, kin, options=options)
            end if
        end if
        xout = xo
    end subroutine ref_art

    ! basic_sweep (AIRMs.f90:813-858) on a slot's matrix; x updated in place.
    subroutine ref_basic_sweep(slot, m, n, x, b, nord, ord, isrand, rowprob, omg, nrA, cumul, upb, lwrb, ub, lb) &
        bind(C, name='ref_basic_sweep')
        integer(c_int), value :: slot, m, n, nord, isrand, rowprob, upb, lwrb
        real(c_double), value :: omg, ub, lb
        real(c_double), intent(inout) :: x(n)
        real(c_double), intent(in) :: b(m), nrA(m), cumul(m)
        integer(c_int), intent(in) :: ord(2*m)
        integer :: ordv(2*m)
        real(kind=8) :: xv(n)
        ordv = ord
        xv = x
        call basic_sweep(xv, mats(slot), b, ordv, nord, isrand /= 0, rowprob /= 0, omg, nrA, cumul, &
                         upb /= 0, lwrb /= 0, ub, lb)
        x = xv
    end subroutine ref_basic_sweep

    ! quicksort (system_generation.f90:37-77) of a(1:nn) with IND alongside.
    subroutine ref_quicksort(nn, a, ind) bind(C, name='ref_quicksort')
        integer(c_int), value :: nn
        real(c_double), intent(inout) :: a(nn)
        integer(c_int), intent(inout) :: ind(nn)
        real(kind=8) :: av(nn)
        integer :: iv(nn)
        av = a
        iv = ind
        call quicksort(av, iv)
        a = av
        ind = iv
    end subroutine ref_quicksort

    ! art with xref (AIRMs.f90:322, 330-333), natural row order.
    subroutine ref_art_xref(slot, method, m, b, nk, kin, hasx0, n, x0, xref, upb, lwrb, ub, lb, relax, damp, &
                            rowprob, xout) bind(C, name='ref_art_xref')
        integer(c_int), value :: slot, method, m, nk, hasx0, n, upb, lwrb, rowprob
        real(c_double), value :: ub, lb, relax, damp
        real(c_double), intent(in) :: b(m), x0(n), xref(n)
        integer(c_int), intent(in) :: kin(nk)
        real(c_double), intent(out) :: xout(n, nk)
        type(opts) :: options
        character(len=12) :: meth
        real(kind=8) :: xo(n, nk)
        call fill_opts(options, upb, lwrb, ub, lb, relax, damp, rowprob)
        select case (method)
        case (0)
            meth = 'kaczmarz'
        case (1)
            meth = 'symkaczmarz'
        case default
            meth = 'randkaczmarz'
        end select
        xo = 0d0
        if (hasx0 /= 0) then
            call art(xo, trim(meth), mats(slot), b, kin, x0=x0, xref=xref, options=options)
        else
            call art(xo, trim(meth), mats(slot), b, kin, xref=xref, options=options)
        end if
        xout = xo
    end subroutine ref_art_xref

    ! twin_art (AIRMs.f90:371-811); stoprule 0 = 'errorgauge', 1 = 'mutualstep'.
    subroutine ref_twin_art(slot, method, stoprule, m, b, nk, kin, hasx0, hasy0, n, x0, y0, slack, &
                            upb, lwrb, ub, lb, relax, damp, rowprob, xout, final_it, xs, ys, eg, neg) &
        bind(C, name='ref_twin_art')
        integer(c_int), value :: slot, method, stoprule, m, nk, hasx0, hasy0, n, slack, upb, lwrb, rowprob, neg
        real(c_double), value :: ub, lb, relax, damp
        real(c_double), intent(in) :: b(m), x0(n), y0(n)
        integer(c_int), intent(in) :: kin(nk)
        real(c_double), intent(out) :: xout(n), xs(n, nk), ys(n, nk), eg(neg)
        integer(c_int), intent(out) :: final_it
        type(opts) :: options
        character(len=12) :: meth, rule
        real(kind=8) :: xo(n), xst(n, nk), yst(n, nk), egv(maxval(kin))
        integer :: fi, sl
        call fill_opts(options, upb, lwrb, ub, lb, relax, damp, rowprob)
        select case (method)
        case (0)
            meth = 'kaczmarz'
        case (1)
            meth = 'symkaczmarz'
        case default
            meth = 'randkaczmarz'
        end select
        rule = 'errorgauge'
        if (stoprule == 1) rule = 'mutualstep'
        xst = 0d0
        yst = 0d0
        egv = 0d0
        fi = 0
        sl = slack
        if (hasx0 /= 0 .and. hasy0 /= 0) then
            call twin_art(xo, fi, xst, yst, egv, trim(meth), trim(rule), mats(slot), b, kin, x0=x0, y0=y0, &
                          slack=sl, options=options)
        else if (hasx0 /= 0) then
            call twin_art(xo, fi, xst, yst, egv, trim(meth), trim(rule), mats(slot), b, kin, x0=x0, &
                          slack=sl, options=options)
        else if (hasy0 /= 0) then
            call twin_art(xo, fi, xst, yst, egv, trim(meth), trim(rule), mats(slot), b, kin, y0=y0, &
                          slack=sl, options=options)
        else
            call twin_art(xo, fi, xst, yst, egv, trim(meth), trim(rule), mats(slot), b, kin, slack=sl, &
                          options=options)
        end if
        xout = xo
        final_it = fi
        xs = xst
        ys = yst
        eg = egv(1:neg)
    end subroutine ref_twin_art

    ! store_matrix (system_generation.f90:595-625) of a slot's matrix.
    subroutine ref_store(slot, path, plen) bind(C, name='ref_store')
        integer(c_int), value :: slot, plen
        character(kind=c_char), intent(in) :: path(plen)
        character(len=plen) :: fname
        integer :: i
        do i = 1, plen
            fname(i:i) = path(i)
        end do
        call store_matrix(mats(slot), fname)
    end subroutine ref_store

    ! Seed the runtime generator deterministically from one integer.
    ! OpenMP threads of the reference's parallel regions (MMRS, MMCS, compute_row_norms) from now on.
    subroutine ref_set_threads(nt) bind(C, name='ref_set_threads')
        use omp_lib, only: omp_set_num_threads
        integer(c_int), value :: nt
        call omp_set_num_threads(nt)
    end subroutine ref_set_threads

    subroutine ref_rng_seed(seed) bind(C, name='ref_rng_seed')
        integer(c_int), value :: seed
        integer :: k, i
        integer, allocatable :: s(:)
        call random_seed(size=k)
        allocate (s(k))
        do i = 1, k
            s(i) = seed + 7919*i
        end do
        call random_seed(put=s)
    end subroutine ref_rng_seed

    subroutine ref_rng_draw(cnt, u) bind(C, name='ref_rng_draw')
        integer(c_int), value :: cnt
        real(c_double), intent(out) :: u(cnt)
        integer :: i
        real(kind=8) :: v
        do i = 1, cnt
            call random_number(v)
            u(i) = v
        end do
    end subroutine ref_rng_draw

    subroutine ref_shepplogan(nn, x) bind(C, name='ref_shepplogan')
        integer(c_int), value :: nn
        real(c_double), intent(out) :: x(nn*nn)
        x = vectorise(shepplogan(nn))
    end subroutine ref_shepplogan

    ! A phantom of the reference's gallery (phantomgallery.f90) after ref_rng_seed(seed), as the
    ! vectorised image.  kind: 1 smooth, 2 threephases, 3 threephases_smooth, 4 binary,
    ! 5 fourphases, 6 mst, 7 grains, 8 bubbles.  pin / steps / v <= 0 leave that optional argument
    ! absent; gauss selects randomtype 'gaussian' (grains, bubbles).
    subroutine ref_phantom(kind, nn, pin, steps, v, gauss, seed, x) bind(C, name='ref_phantom')
        integer(c_int), value :: kind, nn, pin, steps, gauss, seed
        real(c_double), value :: v
        real(c_double), intent(out) :: x(nn*nn)
        character(8) :: rt
        real(kind=8) :: vv
        rt = 'uniform'
        if (gauss /= 0) rt = 'gaussian'
        vv = v
        call ref_rng_seed(seed)
        select case (kind)
        case (1)
            if (pin > 0) then
                x = vectorise(smooth(nn, pin))
            else
                x = vectorise(smooth(nn))
            end if
        case (2)
            if (pin > 0) then
                x = vectorise(threephases(nn, pin))
            else
                x = vectorise(threephases(nn))
            end if
        case (3)
            if (pin > 0 .and. v > 0) then
                x = vectorise(threephases_smooth(nn, pin, vv))
            else if (pin > 0) then
                x = vectorise(threephases_smooth(nn, pin))
            else
                x = vectorise(threephases_smooth(nn))
            end if
        case (4)
            if (pin > 0 .and. steps > 0) then
                x = vectorise(binary(nn, pin, steps))
            else if (pin > 0) then
                x = vectorise(binary(nn, pin))
            else
                x = vectorise(binary(nn))
            end if
        case (5)
            x = vectorise(fourphases(nn))
        case (6)
            if (pin > 0) then
                x = vectorise(mst(nn, pin))
            else
                x = vectorise(mst(nn))
            end if
        case (7)
            if (pin > 0) then
                x = vectorise(grains(nn, rt, pin))
            else
                x = vectorise(grains(nn, rt))
            end if
        case (8)
            if (pin > 0) then
                x = vectorise(bubbles(nn, rt, pin))
            else
                x = vectorise(bubbles(nn, rt))
            end if
        case default
            x = 0d0
        end select
    end subroutine ref_phantom

    ! add_white_noise (system_generation.f90:709-745) after ref_rng_seed(seed).
    subroutine ref_white_noise(m, b, eta, seed, bt, e) bind(C, name='ref_white_noise')
        integer(c_int), value :: m, seed
        real(c_double), intent(in) :: b(m)
        real(c_double), value :: eta
        real(c_double), intent(out) :: bt(m), e(m)
        real(kind=8) :: w(m), f(m), ee
        ee = eta
        call ref_rng_seed(seed)
        call add_white_noise(w, f, b, ee)
        bt = w
        e = f
    end subroutine ref_white_noise

    subroutine ref_linspace(a, b, nn, x) bind(C, name='ref_linspace')
        real(c_double), value :: a, b
        integer(c_int), value :: nn
        real(c_double), intent(out) :: x(nn)
        real(kind=8) :: aa, bb
        aa = a
        bb = b
        x = linspace(aa, bb, nn)
    end subroutine ref_linspace

    ! Fortran SUM / NORM2 / MIN / MAX probes used to pin the oracle's reductions.
    function ref_sum(nn, v) bind(C, name='ref_sum') result(s)
        integer(c_int), value :: nn
        real(c_double), intent(in) :: v(nn)
        real(c_double) :: s
        s = sum(v)
    end function ref_sum
    ! NORM2 of a 2-vector difference and integer powers of a real, in the forms the phantom gallery
    ! uses them (phantomgallery.f90:555, 584, 612; :230, 283), for pinning the device kernels.
    subroutine ref_norm2_diff(cnt, a, b, d) bind(C, name='ref_norm2_diff')
        integer(c_int), value :: cnt
        real(c_double), intent(in) :: a(2, cnt), b(2, cnt)
        real(c_double), intent(out) :: d(cnt)
        real(kind=8) :: xloc(2)
        integer :: k
        do k = 1, cnt
            xloc = a(:, k)
            d(k) = norm2(xloc - b(:, k))
        end do
    end subroutine ref_norm2_diff

    subroutine ref_powers(cnt, v, p2, p3) bind(C, name='ref_powers')
        integer(c_int), value :: cnt
        real(c_double), intent(in) :: v(cnt)
        real(c_double), intent(out) :: p2(cnt), p3(cnt)
        real(kind=8) :: t
        integer :: k
        do k = 1, cnt
            t = v(k)
            p2(k) = t**2
            p3(k) = abs(t)**3
        end do
    end subroutine ref_powers

    ! A path as the reference's character(30) file-name arguments take it (blank padded).
    function name30(path, plen) result(fname)
        integer(c_int), intent(in) :: plen
        character(kind=c_char), intent(in) :: path(plen)
        character(30) :: fname
        integer :: i
        fname = ' '
        do i = 1, min(plen, 30)
            fname(i:i) = path(i)
        end do
    end function name30

    ! read_data (system_generation.f90:494-539) into fm_dat; ig = n, p, nth; rg = r, dw, sd.
    subroutine ref_read_data(path, plen, ig, rg) bind(C, name='ref_read_data')
        integer(c_int), value :: plen
        character(kind=c_char), intent(in) :: path(plen)
        integer(c_int), intent(out) :: ig(3)
        real(c_double), intent(out) :: rg(3)
        call read_data(fm_dat, name30(path, plen))
        ig = [fm_dat%n, fm_dat%p, fm_dat%nth]
        rg = [fm_dat%r, fm_dat%dw, fm_dat%sd]
    end subroutine ref_read_data

    ! theta (nth) and T (nt) of fm_dat.
    subroutine ref_data_arrays(nth, theta, nt, T) bind(C, name='ref_data_arrays')
        integer(c_int), value :: nth, nt
        real(c_double), intent(out) :: theta(nth), T(nt)
        theta = fm_dat%theta
        T = fm_dat%T
    end subroutine ref_data_arrays

    ! The file mode of main.f90:222-270 driven through the reference's own procedures: read_data,
    ! the angle cull for skip > 0, the grid_size rescale, A = fanlineartomo(dat) in the slot,
    ! scaled by multiply_by_scalar(grid_size/N), and bt = -log(T).  ig = n, p, nth and
    ! rg = r, dw, sd, grid_size after these steps; ref_data_arrays and ref_file_mode_b read the
    ! culled data and bt.
    subroutine ref_file_mode(slot, path, plen, skip, ig, rg) bind(C, name='ref_file_mode')
        integer(c_int), value :: slot, plen, skip
        character(kind=c_char), intent(in) :: path(plen)
        integer(c_int), intent(out) :: ig(3)
        real(c_double), intent(out) :: rg(4)
        real(kind=8), allocatable :: grid(:, :), keep(:)
        real(kind=8) :: grid_size
        integer :: cnt
        call read_data(fm_dat, name30(path, plen))
        if (skip > 0) then
            grid = reshape(fm_dat%T, [fm_dat%p, fm_dat%nth])
            cnt = size(fm_dat%theta(1:fm_dat%nth:skip))
            keep = fm_dat%theta(1:fm_dat%nth:skip)
            call move_alloc(keep, fm_dat%theta)
            fm_dat%T = reshape(grid(:, 1:fm_dat%nth:skip), [cnt*fm_dat%p])
            fm_dat%nth = cnt
        end if
        grid_size = fm_dat%dw*fm_dat%r/fm_dat%sd
        fm_dat%r = fm_dat%r/grid_size
        fm_dat%dw = fm_dat%dw/grid_size
        fm_dat%sd = fm_dat%sd/grid_size
        mats(slot) = fanlineartomo(fm_dat)
        call mats(slot)%multiply_by_scalar(grid_size/fm_dat%n)
        fm_b = -log(fm_dat%T)
        ig = [fm_dat%n, fm_dat%p, fm_dat%nth]
        rg = [fm_dat%r, fm_dat%dw, fm_dat%sd, grid_size]
    end subroutine ref_file_mode

    subroutine ref_file_mode_b(m, b) bind(C, name='ref_file_mode_b')
        integer(c_int), value :: m
        real(c_double), intent(out) :: b(m)
        b = fm_b
    end subroutine ref_file_mode_b

    ! read_ref (system_generation.f90:541-565).
    subroutine ref_read_ref(path, plen, n, x) bind(C, name='ref_read_ref')
        integer(c_int), value :: plen, n
        character(kind=c_char), intent(in) :: path(plen)
        real(c_double), intent(out) :: x(n)
        real(kind=8), allocatable :: xr(:)
        allocate (xr(n))
        call read_ref(xr, name30(path, plen), n)
        x = xr
    end subroutine ref_read_ref

    ! read_ordering (system_generation.f90:567-593): l entries, at most cap copied to order.
    subroutine ref_read_ordering(path, plen, cap, l, order) bind(C, name='ref_read_ordering')
        integer(c_int), value :: plen, cap
        character(kind=c_char), intent(in) :: path(plen)
        integer(c_int), intent(out) :: l
        integer(c_int), intent(out) :: order(cap)
        integer, allocatable :: ordering(:)
        call read_ordering(ordering, name30(path, plen))
        l = size(ordering)
        order(1:min(l, cap)) = ordering(1:min(l, cap))
    end subroutine ref_read_ordering
end module ref_harness
