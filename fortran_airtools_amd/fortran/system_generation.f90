! system_generation.f90 -- drop-in for the reference module of the same name
! (system_generation.f90 of FAIR tools): the geometry record, linspace, the fan / parallel
! beam system-matrix builders (run on the gfx950 device, bit-identical to the reference's
! serial build), the text data reader and the white-noise generator.
module system_generation
    use iso_c_binding
    use fair_capi
    use sparse_matrices
    implicit none
    real(kind=8), parameter :: pi = 4d0*atan(1d0)

    ! type(data), system_generation.f90:6-12
    type data
        integer :: n, p, nth
        real(kind=8) :: r, dw, sd
        real(kind=8), allocatable :: T(:), theta(:)
    contains
        procedure :: read => read_data
    end type data

contains

    ! system_generation.f90:17-35 (x(i) = c*i + d), evaluated by the library
    function linspace(a, b, n) result(x)
        real(kind=8), intent(in) :: a, b
        integer, intent(in) :: n
        real(kind=8) :: x(n)
        call fair_check(c_fair_linspace(real(a, c_double), real(b, c_double), int(n, c_int32_t), x))
    end function linspace

    function geom_of(dat) result(g)
        class(data), intent(in) :: dat
        type(fair_geom_c) :: g
        if (.not. allocated(dat%theta)) stop "Angle vector is not allocated."
        if (size(dat%theta) /= dat%nth) stop "Angle vector has the wrong length, nth is incorrect."
        g%n = dat%n
        g%p = dat%p
        g%nth = dat%nth
        g%r = dat%r
        g%dw = dat%dw
        g%sd = dat%sd
    end function geom_of

    ! quicksort, system_generation.f90:37-77: Hoare partition around the middle element of
    ! a(first:last), IND permuted alongside; the same comparisons and swaps, so ties end in the
    ! same order (the sort is not stable).
    recursive subroutine quicksort(a, IND, first, last)
        real(kind=8) :: a(:)
        integer :: IND(:)
        integer, optional :: first, last
        real(kind=8) :: pivot, t
        integer :: i, j, lo, hi, k
        lo = 1
        hi = size(a)
        if (present(first)) lo = first
        if (present(last)) hi = last
        pivot = a((lo + hi)/2)
        i = lo
        j = hi
        do
            do while (a(i) < pivot)
                i = i + 1
            end do
            do while (pivot < a(j))
                j = j - 1
            end do
            if (i >= j) exit
            t = a(i)
            a(i) = a(j)
            a(j) = t
            k = IND(i)
            IND(i) = IND(j)
            IND(j) = k
            i = i + 1
            j = j - 1
        end do
        if (lo < i - 1) call quicksort(a, IND, lo, i - 1)
        if (j + 1 < hi) call quicksort(a, IND, j + 1, hi)
    end subroutine quicksort

    ! fanlineartomo, system_generation.f90:79-307
    function fanlineartomo(dat) result(Amat)
        class(data), intent(in) :: dat
        type(csr_matrix) :: Amat
        type(fair_geom_c) :: g
        real(c_double), allocatable :: th(:)
        call fair_ensure_init()
        g = geom_of(dat)
        th = dat%theta
        call fair_check(c_fair_fanlineartomo(g, th, Amat%dev))
        call Amat%sync_host()
    end function fanlineartomo

    ! paralleltomo, system_generation.f90:309-492
    function paralleltomo(dat) result(Amat)
        class(data), intent(in) :: dat
        type(csr_matrix) :: Amat
        type(fair_geom_c) :: g
        real(c_double), allocatable :: th(:)
        call fair_ensure_init()
        g = geom_of(dat)
        th = dat%theta
        call fair_check(c_fair_paralleltomo(g, th, Amat%dev))
        call Amat%sync_host()
    end function paralleltomo

    ! read_data, system_generation.f90:494-539: "label value" lines for n, p, nth, r, dw, sd,
    ! then "label count" + angles in degrees, then "label count" + transmission data.
    subroutine read_data(dat, inputfile, verbose)
        class(data), intent(out) :: dat
        character(*), intent(in) :: inputfile
        logical, optional :: verbose
        character(30) :: label
        integer :: cnt, i, u
        logical :: verb
        verb = .false.
        if (present(verbose)) verb = verbose
        if (verb) write (*, *) "Opening file: ", trim(inputfile)
        open (newunit=u, file=trim(inputfile), action='read')
        read (u, *) label, dat%n
        read (u, *) label, dat%p
        read (u, *) label, dat%nth
        read (u, *) label, dat%r
        read (u, *) label, dat%dw
        read (u, *) label, dat%sd
        read (u, *) label, cnt
        if (cnt /= dat%nth) stop "Angle vector has the wrong length, nth is incorrect."
        allocate (dat%theta(cnt))
        do i = 1, cnt
            read (u, *) dat%theta(i)
        end do
        dat%theta = dat%theta/180d0*pi
        read (u, *) label, cnt
        if (cnt /= dat%nth*dat%p) stop "Data has the wrong size, or p and nth are incorrect."
        allocate (dat%T(cnt))
        do i = 1, cnt
            read (u, *) dat%T(i)
        end do
        close (u)
        if (verb) write (*, *) "Closing file: ", trim(inputfile)
    end subroutine read_data

    ! read_ref, system_generation.f90:541-565: n records of one real each (a reference image).
    subroutine read_ref(xref, inputfile, n, verbose)
        integer :: n
        real(kind=8) :: xref(n)
        character(*), intent(in) :: inputfile
        logical, optional :: verbose
        integer :: i, u
        logical :: verb
        verb = .false.
        if (present(verbose)) verb = verbose
        if (verb) write (*, *) "Opening file: ", trim(inputfile)
        open (newunit=u, file=trim(inputfile), action='read')
        do i = 1, n
            read (u, *) xref(i)
        end do
        close (u)
        if (verb) write (*, *) "Closing file: ", trim(inputfile)
    end subroutine read_ref

    ! read_ordering, system_generation.f90:567-593: a "label l" record, then l records of one
    ! integer each (a row order for art).
    subroutine read_ordering(ordering, inputfile, verbose)
        character(*), intent(in) :: inputfile
        integer, allocatable :: ordering(:)
        logical, optional :: verbose
        character(30) :: label
        integer :: cnt, i, u
        logical :: verb
        verb = .false.
        if (present(verbose)) verb = verbose
        if (verb) write (*, *) "Opening file: ", trim(inputfile)
        open (newunit=u, file=trim(inputfile), action='read')
        read (u, *) label, cnt
        if (allocated(ordering)) deallocate (ordering)
        allocate (ordering(cnt))
        do i = 1, cnt
            read (u, *) ordering(i)
        end do
        close (u)
    end subroutine read_ordering

    ! store_matrix, system_generation.f90:595-625: the records m, n, nnz, vA, JA, IA (Fortran
    ! unformatted sequential), from the host components (device copies synchronised first).
    subroutine store_matrix(A, inputfile, verbose)
        character(*), intent(in) :: inputfile
        class(sparse_matrix) :: A
        logical, optional :: verbose
        integer :: k
        logical :: verb
        verb = .false.
        if (present(verbose)) verb = verbose
        if (c_associated(A%dev)) call A%sync_host()
        open (newunit=k, file=inputfile, form='unformatted')
        if (verb) write (*, *) "Opening file: ", inputfile
        write (k) A%m
        write (k) A%n
        write (k) A%nnz
        write (k) A%vA
        write (k) A%JA
        write (k) A%IA
        close (k)
        if (verb) write (*, *) "done! Closing file."
    end subroutine store_matrix

    ! read_csr_matrix, system_generation.f90:627-665 (with a unit of its own: the reference opens
    ! an unset unit number, :637).  Any device copy is dropped; the next product uploads.
    subroutine read_csr_matrix(A, inputfile, verbose)
        character(*), intent(in) :: inputfile
        class(csr_matrix) :: A
        logical, optional :: verbose
        call read_dump(A, inputfile, .false., verbose)
    end subroutine read_csr_matrix

    ! read_csc_matrix, system_generation.f90:667-707
    subroutine read_csc_matrix(A, inputfile, verbose)
        character(*), intent(in) :: inputfile
        class(csc_matrix) :: A
        logical, optional :: verbose
        call read_dump(A, inputfile, .true., verbose)
    end subroutine read_csc_matrix

    subroutine read_dump(A, inputfile, csc, verbose)
        class(sparse_matrix) :: A
        character(*), intent(in) :: inputfile
        logical, intent(in) :: csc
        logical, optional :: verbose
        integer :: k
        logical :: verb
        verb = .false.
        if (present(verbose)) verb = verbose
        if (verb) write (*, *) "Opening file: ", inputfile
        call A%free_device()
        open (newunit=k, form="unformatted", file=inputfile, action="read")
        read (k) A%m
        read (k) A%n
        read (k) A%nnz
        if (allocated(A%vA)) deallocate (A%vA)
        if (allocated(A%JA)) deallocate (A%JA)
        if (allocated(A%IA)) deallocate (A%IA)
        allocate (A%vA(A%nnz), A%JA(A%nnz))
        if (csc) then
            allocate (A%IA(0:A%n))
        else
            allocate (A%IA(0:A%m))
        end if
        read (k) A%vA
        read (k) A%JA
        read (k) A%IA
        close (k)
        if (verb) write (*, *) "Data collected. Closing file."
    end subroutine read_dump

    ! add_white_noise, system_generation.f90:709-745: one uniform draw per sample, in order; the
    ! Newton solves of Phi(x) = u (from x = 0, step f/df, df = sqrt(2/pi)exp(-x^2/2)) run on the
    ! device, one thread per sample; then e = eta*||b||*e/||e|| and bt = b + e.
    subroutine add_white_noise(bt, e, b, eta)
        real(kind=8), intent(in) :: b(:), eta
        real(kind=8), dimension(size(b)) :: bt, e
        real(kind=8), allocatable :: u(:)
        integer :: k
        allocate (u(size(b)))
        do k = 1, size(b)
            call random_number(u(k))
        end do
        call fair_ensure_init()
        call fair_check(c_fair_normal_newton(int(size(b), c_int64_t), u, e))
        e = eta*norm2(b)*e/norm2(e)
        bt = b + e
    end subroutine add_white_noise
end module system_generation
