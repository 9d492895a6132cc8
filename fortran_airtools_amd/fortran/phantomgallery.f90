! phantomgallery.f90 -- drop-in for the reference's module phantomgallery: the same functions,
! arguments, optional arguments and results.  The random parts (bump centres, random walks, the
! spanning tree, seed points) draw from random_number in the reference's order and run here; the
! per-pixel evaluation, O(N^2) per shape, runs on the device through fair_phantom_* (one thread
! per pixel, the reference's expression for every shape in the reference's order).
! draw_line and swapcoord (phantomgallery.f90:164-249), which nothing in the reference calls,
! are not provided.
module phantomgallery
    use iso_c_binding
    use fair_capi
    use system_generation, only: pi, linspace
    implicit none
    private :: axes, int_axes, run_bumps, walk
contains

    ! Column and row coordinates of meshgrid over [x0,x1] x [y0,y1]: column j at cx(j), row i at
    ! ry(i) (rows run from y1 down to y0), phantomgallery.f90:17-56.
    subroutine axes(N, x0, x1, y0, y1, cx, ry)
        integer, intent(in) :: N
        real(kind=8), intent(in) :: x0, x1, y0, y1
        real(kind=8), intent(out) :: cx(N), ry(N)
        cx = linspace(x0, x1, N)
        ry = linspace(y1, y0, N)
    end subroutine axes

    subroutine int_axes(N, cx, ry)
        integer, intent(in) :: N
        real(kind=8), intent(out) :: cx(N), ry(N)
        integer :: k
        cx = [(dble(k), k=1, N)]
        ry = cx
    end subroutine int_axes

    subroutine meshgrid(x, y, N, corners)
        integer, intent(in) :: N
        real(kind=8), dimension(0:N-1, 0:N-1), intent(out) :: x, y
        real(kind=8), optional :: corners(4)
        real(kind=8) :: cx(N), ry(N)
        integer :: k
        if (present(corners)) then
            call axes(N, corners(1), corners(2), corners(3), corners(4), cx, ry)
        else
            call axes(N, -1d0, 1d0, -1d0, 1d0, cx, ry)
        end if
        do k = 0, N - 1
            x(k, :) = cx
            y(:, k) = ry
        end do
    end subroutine meshgrid

    subroutine intmeshgrid(x, y, N)
        integer, intent(in) :: N
        integer, dimension(0:N-1, 0:N-1), intent(out) :: x, y
        integer :: i, j
        do j = 0, N - 1
            do i = 0, N - 1
                x(i, j) = j + 1
                y(i, j) = i + 1
            end do
        end do
    end subroutine intmeshgrid

    function choose_phantom(phantom, N) result(x)
        character(*), intent(in) :: phantom
        integer, intent(in) :: N
        real(kind=8) :: x(N, N)
        select case (phantom)
        case ('shepplogan')
            x = shepplogan(N)
        case ('smooth')
            x = smooth(N)
        case ('3phases')
            x = threephases(N)
        case ('3phasesmooth')
            x = threephases_smooth(N)
        case ('binary')
            x = binary(N)
        case ('4phases')
            x = fourphases(N)
        case ('mst')
            x = mst(N)
        case ('grains')
            x = grains(N)
        case ('bubbles')
            x = bubbles(N)
        case default
            stop 'Not a recognised phantom.'
        end select
    end function choose_phantom

    ! shepplogan, phantomgallery.f90:116-164 (Toft's high-contrast head): ellipses
    ! (A, a, b, x0, y0, phi in degrees).
    function shepplogan(N) result(x)
        integer, intent(in) :: N
        real(kind=8) :: x(N, N)
        real(kind=8) :: t(6, 10), ell(7, 10), cx(N), ry(N), phi
        integer :: k
        t = reshape([1d0, .6900d0, .9200d0, 0d0, 0d0, 0d0, &
                     -.8d0, .6624d0, .8740d0, 0d0, -.0184d0, 0d0, &
                     -.2d0, .1100d0, .3100d0, .22d0, 0d0, -18d0, &
                     -.2d0, .1600d0, .4100d0, -.22d0, 0d0, 18d0, &
                     .1d0, .2100d0, .2500d0, 0d0, .35d0, 0d0, &
                     .1d0, .0460d0, .0460d0, 0d0, .1d0, 0d0, &
                     .1d0, .0460d0, .0460d0, 0d0, -.1d0, 0d0, &
                     .1d0, .0460d0, .0230d0, -.08d0, -.605d0, 0d0, &
                     .1d0, .0230d0, .0230d0, 0d0, -.606d0, 0d0, &
                     .1d0, .0230d0, .0460d0, .06d0, -.605d0, 0d0], [6, 10])
        do k = 1, 10
            phi = t(6, k)*pi/180d0
            ell(:, k) = [t(1, k), t(2, k)**2, t(3, k)**2, t(4, k), t(5, k), cos(phi), sin(phi)]
        end do
        call axes(N, -1d0, 1d0, -1d0, 1d0, cx, ry)
        call fair_ensure_init()
        call fair_check(c_fair_phantom_ellipses(int(N, c_int32_t), cx, ry, 10_c_int32_t, ell, x))
    end function shepplogan

    ! x = sum over bumps of w*exp(-(c*P(X-cx))/dx - (c*P(Y-cy))/dy); bumps(:, k) = (w, cx, cy, dx, dy, c)
    subroutine run_bumps(N, cx, ry, bumps, power, x)
        integer, intent(in) :: N, power
        real(kind=8), intent(in) :: cx(N), ry(N), bumps(:, :)
        real(kind=8), intent(out) :: x(N, N)
        call fair_ensure_init()
        call fair_check(c_fair_phantom_bumps(int(N, c_int32_t), cx, ry, int(size(bumps, 2), c_int32_t), bumps, &
                                             int(power, c_int32_t), x))
    end subroutine run_bumps

    function smooth(N, pin) result(x)
        integer, intent(in) :: N
        real(kind=8) :: x(N, N)
        integer, optional :: pin
        integer :: i, p
        real(kind=8) :: sigma, c(4, 2), a(4), cx(N), ry(N)
        real(kind=8), allocatable :: bumps(:, :)
        p = 4
        if (present(pin)) p = pin
        sigma = 0.25d0*N
        c(1, :) = (/0.6d0*N, 0.6d0*N/)
        c(2, :) = (/0.5*N, 0.3*N/)        ! default-real products, as in the reference
        c(3, :) = (/0.2*N, 0.7*N/)
        c(4, :) = (/0.8*N, 0.2*N/)
        a = (/1d0, 0.5d0, 0.7d0, 0.9d0/)
        allocate (bumps(6, p))
        do i = 1, p
            bumps(:, i) = [a(i), c(i, 1), c(i, 2), (1.2d0*sigma)**2, sigma**2, 1d0]
        end do
        call int_axes(N, cx, ry)
        call run_bumps(N, cx, ry, bumps, 2, x)
        x = x/maxval(x)
    end function smooth

    ! p cubic bumps of width den at centres drawn as random_number(c(p,2))*N
    subroutine cubic_image(N, p, cx, ry, den, img)
        integer, intent(in) :: N, p
        real(kind=8), intent(in) :: cx(N), ry(N), den
        real(kind=8), intent(out) :: img(N, N)
        real(kind=8) :: c(p, 2), bumps(6, p)
        integer :: i
        call random_number(c)
        c = c*N
        do i = 1, p
            bumps(:, i) = [1d0, c(i, 1), c(i, 2), den, den, 1d0]
        end do
        call run_bumps(N, cx, ry, bumps, 3, img)
    end subroutine cubic_image

    function threephases(N, pin, seed) result(x)
        integer, intent(in) :: N
        real(kind=8) :: x(N, N)
        integer, optional :: pin, seed(:)
        integer :: p
        real(kind=8) :: cx(N), ry(N), im1(N, N), im2(N, N)
        p = 100
        if (present(pin)) p = pin
        if (present(seed)) call random_seed(put=seed)
        call axes(N, 1d0, N*1d0, N*1d0, 1d0, cx, ry)
        call cubic_image(N, p, cx, ry, (2.5d0*(0.025d0*N))**3, im1)
        where (im1 < .35d0) im1 = 0
        where (im1 >= .35d0) im1 = 2d0
        call cubic_image(N, p, cx, ry, (2.5d0*(0.025d0*N))**2, im2)
        where (im2 < 0.55d0) im2 = 0d0
        where (im2 >= 0.55d0) im2 = 1d0
        x = im1 + im2
        where (x == 3d0) x = 1d0
        x = x/maxval(x)
    end function threephases

    function threephases_smooth(N, pin, vin, seed) result(x)
        integer, intent(in) :: N
        real(kind=8) :: x(N, N)
        integer, optional :: pin, seed(:)
        real(kind=8), optional :: vin
        integer :: i, p
        real(kind=8) :: cx(N), ry(N), im1(N, N), im2(N, N), bg(N, N), v, sb
        real(kind=8), allocatable :: c(:, :), bumps(:, :)
        p = 100
        if (present(pin)) p = pin
        if (present(seed)) call random_seed(put=seed)
        v = 1.8d0
        if (present(vin)) v = vin
        call axes(N, 1d0, N*1d0, N*1d0, 1d0, cx, ry)
        call cubic_image(N, p, cx, ry, (2.5d0*(0.025d0*N))**3, im1)
        where (im1 < .35d0) im1 = 0
        where (im1 >= .35d0) im1 = (im1 - minval(im1))/maxval(im1)*v + .8d0
        call cubic_image(N, p, cx, ry, (2.5d0*(0.025d0*N))**2, im2)
        where (im2 < .55d0) im2 = 0d0
        where (im2 >= .55d0) im2 = (im2 - minval(im2))/maxval(im2)*v + .3d0
        allocate (c(p, 2), bumps(6, p))
        call random_number(c)
        c = c*N
        do i = 1, p
            call random_number(sb)
            sb = N*(5d0*sb + .5d0)
            bumps(:, i) = [1d-1/p, c(i, 1), c(i, 2), sb**2, sb**2, .5d0]
        end do
        call run_bumps(N, cx, ry, bumps, 2, bg)
        x = bg
        where (im1 > 0) x = im1
        where (im2 > 0) x = im2
        x = x/maxval(x)
    end function threephases_smooth

    ! The walks of binary (phantomgallery.f90:358-409): disks(:, k) = (centre, stepsize^2) of every
    ! step taken, nd of them.
    subroutine walk(p, steps, psize, disks, nd)
        integer, intent(in) :: p, steps
        real(kind=8), intent(in) :: psize
        real(kind=8), intent(inout) :: disks(:, :)
        integer, intent(inout) :: nd
        integer :: i, j
        real(kind=8) :: step, vert(3), horz(3), prob, pos(2)
        do j = 1, p
            call random_number(step)
            call random_number(vert)
            vert(2) = (10d0 + 10d0/step)*vert(2)
            vert = (/vert(1), vert(1) + vert(2), sum(vert)/)
            vert = vert/maxval(vert)
            call random_number(horz)
            horz(2) = 0d0
            horz = (/horz(1), horz(1) + horz(2), sum(horz)/)
            horz = horz/maxval(horz)
            step = .75d0*psize*(step + 1)
            call random_number(pos)
            pos = 2d0*pos - 1d0
            do i = 1, steps
                call random_number(prob)
                if (prob < vert(1)) then
                    pos(2) = pos(2) - step
                elseif (prob > vert(2)) then
                    pos(2) = pos(2) + step
                end if
                call random_number(prob)
                if (prob < horz(1)) then
                    pos(1) = pos(1) - step
                elseif (prob > horz(2)) then
                    pos(1) = pos(1) + step
                end if
                if (abs(pos(1)) > 1d0 .or. abs(pos(2)) > 1d0) exit
                nd = nd + 1
                disks(:, nd) = [pos(1), pos(2), step**2]
            end do
        end do
    end subroutine walk

    function binary(N, pin, stepsin, seed) result(x)
        integer, intent(in) :: N
        real(kind=8) :: x(N, N)
        integer, optional :: pin, stepsin, seed(:)
        integer :: p, steps, nd
        real(kind=8) :: cx(N), ry(N)
        real(kind=8), allocatable :: disks(:, :)
        p = 100
        if (present(pin)) p = pin
        steps = N
        if (present(stepsin)) steps = stepsin
        if (present(seed)) call random_seed(put=seed)
        x = 0d0
        call axes(N, -1d0, 1d0, -1d0, 1d0, cx, ry)
        allocate (disks(3, max(1, p*steps)))
        nd = 0
        call walk(p, steps, cx(3) - cx(2), disks, nd)
        call fair_ensure_init()
        call fair_check(c_fair_phantom_stamp_disks(int(N, c_int32_t), cx, ry, int(nd, c_int32_t), disks, 1d0, x))
    end function binary

    function fourphases(N, seed) result(x)
        integer, intent(in) :: N
        real(kind=8) :: x(N, N)
        integer, optional :: seed(:)
        integer :: i
        if (present(seed)) call random_seed(put=seed)
        x = 0d0
        do i = 1, 8
            x = x + (-1d0)**i*binary(N, 80)
        end do
        where (x < 0d0) x = 0d0
        x = x/maxval(x)
    end function fourphases

    ! mst, phantomgallery.f90:439-527: Prim's tree over p normal points.  Each loose point keeps
    ! its smallest distance to the tree and the first tree point giving it; the next edge is the
    ! smallest distance, then the earliest tree point, then the earliest loose point -- the pair
    ! the reference's scan (tree outer, loose inner, strict <) finds.
    function mst(N, pin, seed) result(img)
        integer, intent(in) :: N
        real(kind=8) :: img(N, N)
        integer, optional :: seed(:), pin
        integer :: i, j, k, p, pow, ii, jj, t
        integer, allocatable :: from(:), at(:)
        logical, allocatable :: done(:)
        real(kind=8) :: xm(2), r, th, u, m, cx(N), ry(N)
        real(kind=8), allocatable :: dist(:, :), x0(:, :), best(:), edges(:, :)
        if (present(pin)) then
            p = pin
        else
            p = 30*sqrt(N*1d0)
        end if
        if (present(seed)) call random_seed(put=seed)
        pow = 1
        allocate (x0(2, p))
        do i = 1, 2
            do j = 1, p
                x0(i, j) = .2d0*N*randn() + .5d0*N
            end do
        end do
        allocate (dist(p, p))
        do i = 1, p
            do j = 1, p
                dist(i, j) = norm2(x0(:, i) - x0(:, j))
            end do
        end do
        allocate (best(p), from(p), at(p), done(p), edges(6, max(1, p - 1)))
        done = .false.
        done(1) = .true.
        at = 0
        at(1) = 1
        best = dist(1, :)
        from = 1
        do t = 2, p
            ii = 0
            m = huge(m)
            do k = 1, p
                if (done(k)) cycle
                if (best(k) < m) then
                    m = best(k)
                    ii = k
                elseif (best(k) == m .and. ii > 0) then
                    if (at(from(k)) < at(from(ii))) ii = k
                end if
            end do
            jj = from(ii)
            done(ii) = .true.
            at(ii) = t
            do k = 1, p
                if (dist(ii, k) < best(k)) then
                    best(k) = dist(ii, k)
                    from(k) = ii
                end if
            end do
            xm = .5d0*(x0(:, ii) + x0(:, jj))
            th = atan2(x0(2, ii) - x0(2, jj), x0(1, ii) - x0(1, jj))
            r = sum(abs(xm - x0(:, jj))**pow)
            call random_number(u)
            edges(:, t - 1) = [xm(1), xm(2), cos(th), sin(th), r, 1d0 + u]
        end do
        img = 0d0
        call int_axes(N, cx, ry)
        call fair_ensure_init()
        call fair_check(c_fair_phantom_last_band(int(N, c_int32_t), cx, ry, int(max(p - 1, 0), c_int32_t), edges, img))
        img = img/maxval(img)
    end function mst

    ! Seed points of grains / bubbles: 'gaussian' (scale*randn, redrawn until inside (-1, 1),
    ! x coordinates first) or 'uniform' (random_number over x0(2, p), then 1.5(2 x0 - 1)).
    subroutine seed_points(rt, scale, x0)
        character(8), intent(in) :: rt
        real(kind=8), intent(in) :: scale
        real(kind=8), intent(out) :: x0(:, :)
        integer :: i, j
        if (rt == 'gaussian') then
            do i = 1, 2
                do j = 1, size(x0, 2)
                    x0(i, j) = -2d0
                    do while (.not. (-1d0 < x0(i, j) .and. x0(i, j) < 1d0))
                        x0(i, j) = scale*randn()
                    end do
                end do
            end do
        elseif (rt == 'uniform') then
            call random_number(x0)
            x0 = 1.5d0*(2d0*x0 - 1d0)
        else
            stop "Not a recognised input for randomtype. Use 'gaussian' or 'uniform'."
        end if
    end subroutine seed_points

    function grains(N, randomtype, pin, seed) result(img)
        integer, intent(in) :: N
        real(kind=8) :: img(N, N)
        character(8), optional :: randomtype
        integer, optional :: seed(:), pin
        integer :: p
        real(kind=8) :: cx(N), ry(N)
        real(kind=8), allocatable :: x0(:, :), vals(:)
        character(8) :: rt
        if (present(pin)) then
            p = pin
        else
            p = 3*sqrt(N*1d0)
        end if
        rt = 'uniform'
        if (present(randomtype)) rt = randomtype
        if (present(seed)) call random_seed(put=seed)
        allocate (x0(2, p), vals(p))
        call seed_points(rt, 1d0, x0)
        call random_number(vals)
        call axes(N, -1d0, 1d0, -1d0, 1d0, cx, ry)
        call fair_ensure_init()
        call fair_check(c_fair_phantom_nearest(int(N, c_int32_t), cx, ry, int(p, c_int32_t), x0, img))
        img = img/maxval(img)
    end function grains

    function bubbles(N, randomtype, pin, seed) result(img)
        integer, intent(in) :: N
        real(kind=8) :: img(N, N)
        character(8), optional :: randomtype
        integer, optional :: pin, seed(:)
        integer :: i, j, kmin, p
        real(kind=8) :: dmin, d, u, cx(N), ry(N)
        real(kind=8), allocatable :: x0(:, :), disks(:, :)
        character(8) :: rt
        if (present(pin)) then
            p = pin
        else
            p = 15*sqrt(N*1d0)
        end if
        rt = 'uniform'
        if (present(randomtype)) rt = randomtype
        if (present(seed)) call random_seed(put=seed)
        call axes(N, -1d0, 1d0, -1d0, 1d0, cx, ry)
        allocate (x0(2, p), disks(4, max(1, p)))
        call seed_points(rt, .5d0, x0)
        kmin = 0
        do i = 1, p
            dmin = huge(dmin)
            do j = 1, p
                if (i == j) cycle
                d = norm2(x0(:, i) - x0(:, j))
                if (d < dmin) then
                    dmin = d
                    kmin = j
                end if
            end do
            call random_number(u)
            disks(:, i) = [x0(1, i), x0(2, i), (1d0*dmin)**2, dble(kmin)]
        end do
        img = 0d0
        call fair_ensure_init()
        call fair_check(c_fair_phantom_last_disk(int(N, c_int32_t), cx, ry, int(p, c_int32_t), disks, img))
        img = img/maxval(img)
    end function bubbles

    ! vectorise, phantomgallery.f90:144-162: the columns one after another.
    function vectorise(x) result(v)
        real(kind=8), intent(in) :: x(:, :)
        real(kind=8) :: v(size(x))
        v = reshape(x, [size(x)])
    end function vectorise

    ! randn, phantomgallery.f90:251-283: one random_number draw, Newton from 0 on Phi(x) = u.
    function randn(maxits, tolerance) result(x)
        integer, optional :: maxits
        real(kind=8), optional :: tolerance
        real(kind=8) :: x
        integer :: i, imax
        real(kind=8) :: f, df, tol, u
        imax = 100
        if (present(maxits)) imax = maxits
        tol = 1d-14
        if (present(tolerance)) tol = tolerance
        x = 0d0
        call random_number(u)
        do i = 1, imax
            f = .5d0*(1 + erf(x/sqrt(2d0))) - u
            df = sqrt(2d0/pi)*exp(-.5d0*x**2)
            if (abs(f/df) < tol) exit
            if (i == imax .and. abs(f/df) > tol) stop "Newton method did not converge."
            x = x - f/df
        end do
    end function randn
end module phantomgallery
