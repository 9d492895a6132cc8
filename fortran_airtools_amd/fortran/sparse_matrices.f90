! sparse_matrices.f90 -- drop-in for the reference module of the same name
! (sparse_matrices.f90 of FAIR tools), backed by the gfx950 library.
!
! The public components m, n, nnz, vA, IA(0:m), JA keep the reference's meaning and stay valid
! on the host (they are refreshed from the device after every operation that changes the
! matrix).  `dev` holds the device copy.  Matrices from fanlineartomo / paralleltomo have one;
! for a matrix assembled on the host, call A%to_device() once -- procedures whose matrix
! argument is intent(in) in the reference (matvecmul, rowvecmult, add_scalar_mult_row)
! otherwise upload a temporary copy per call.  After editing vA / IA / JA on the host, call
! A%to_device() again.  Device copies are released by A%free_device() (there is no
! finaliser: assignment copies share one device copy).
!
! CSC_matrix: the stored arrays are the CSR of A^T, so A*v is the device's transposed product
! and vice versa (sparse_matrices.f90:279-296, 310-345).
module sparse_matrices
    use iso_c_binding
    use fair_capi
    implicit none
    private
    public :: sparse_matrix, CSR_matrix, CSC_matrix, borrow_dev, release_dev

    real(kind=8), parameter, public :: TOL = 1d-15

    type sparse_matrix
        integer :: m = 0, n = 0, nnz = 0
        real(kind=8), allocatable :: vA(:)
        integer, allocatable :: IA(:), JA(:)
        type(c_ptr) :: dev = c_null_ptr
    contains
        procedure :: matvecmul => sparsematmul
        procedure :: transvecmul => sparsetransmul
        procedure :: rowvecmult => csr_matrowvec
        procedure :: transpose => transposition
        procedure :: row_norms => compute_row_norms
        procedure :: column_norms => compute_column_norms
        procedure :: multiply_by_scalar => scalar_mult
        procedure :: to_device
        procedure :: sync_host
        procedure :: free_device
        procedure :: device_handle
    end type sparse_matrix

    type, extends(sparse_matrix) :: CSR_matrix
    contains
        procedure :: add_scalar_mult_row => add_mult_row
    end type CSR_matrix

    type, extends(sparse_matrix) :: CSC_matrix
    end type CSC_matrix

contains

    logical function is_csc(A)
        class(sparse_matrix), intent(in) :: A
        select type (A)
        type is (sparse_matrix)
            stop "Matrix type must be either CSR or CSC."
        class is (CSC_matrix)
            is_csc = .true.
        class default
            is_csc = .false.
        end select
    end function is_csc

    ! Upload the host components (replacing any device copy).
    subroutine to_device(A)
        class(sparse_matrix), intent(inout) :: A
        type(sparse_matrix) :: B
        call fair_ensure_init()
        call A%free_device()
        if (.not. allocated(A%IA)) stop "Matrix components are not allocated."
        B%m = A%m
        B%n = A%n
        B%nnz = A%nnz
        call move_alloc(A%IA, B%IA)
        call move_alloc(A%JA, B%JA)
        call move_alloc(A%vA, B%vA)
        call upload_as(B, is_csc(A))
        call move_alloc(B%IA, A%IA)
        call move_alloc(B%JA, A%JA)
        call move_alloc(B%vA, A%vA)
        A%dev = B%dev
    end subroutine to_device

    subroutine ensure_dev(A)
        class(sparse_matrix), intent(inout) :: A
        call fair_ensure_init()
        if (.not. c_associated(A%dev)) call A%to_device()
    end subroutine ensure_dev

    ! Device handle for an intent(in) matrix: its own copy, or a temporary upload that the
    ! caller releases with release_dev.
    subroutine borrow_dev(A, h, temp)
        class(sparse_matrix), intent(in) :: A
        type(c_ptr), intent(out) :: h
        logical, intent(out) :: temp
        type(sparse_matrix) :: B
        call fair_ensure_init()
        temp = .not. c_associated(A%dev)
        if (temp) then
            B%m = A%m
            B%n = A%n
            B%nnz = A%nnz
            B%IA = A%IA
            B%JA = A%JA
            B%vA = A%vA
            call upload_as(B, is_csc(A))
            h = B%dev
        else
            h = A%dev
        end if
    end subroutine borrow_dev

    subroutine release_dev(h, temp)
        type(c_ptr), intent(inout) :: h
        logical, intent(in) :: temp
        if (temp) call fair_check(c_fair_csr_free(h))
    end subroutine release_dev

    subroutine upload_as(B, csc)
        type(sparse_matrix), intent(inout) :: B
        logical, intent(in) :: csc
        integer :: rows, cols
        integer(c_int32_t), target :: dummyi(1)
        real(c_double), target :: dummyd(1)
        if (csc) then
            rows = B%n
            cols = B%m
        else
            rows = B%m
            cols = B%n
        end if
        if (B%nnz > 0) then
            call fair_check(c_fair_csr_create(int(rows, c_int32_t), int(cols, c_int32_t), &
                int(B%nnz, c_int64_t), B%IA, B%JA, B%vA, B%dev))
        else
            call fair_check(c_fair_csr_create(int(rows, c_int32_t), int(cols, c_int32_t), &
                0_c_int64_t, B%IA, dummyi, dummyd, B%dev))
        end if
    end subroutine upload_as

    ! Refresh the host components from the device copy.
    subroutine sync_host(A)
        class(sparse_matrix), intent(inout) :: A
        integer(c_int32_t) :: rows, cols
        integer(c_int64_t) :: nnz
        call fair_check(c_fair_csr_info(A%dev, rows, cols, nnz))
        if (nnz > huge(0)) stop "nnz exceeds the int32 range of the reference's components."
        if (is_csc(A)) then
            A%n = rows
            A%m = cols
        else
            A%m = rows
            A%n = cols
        end if
        A%nnz = int(nnz)
        if (allocated(A%IA)) deallocate (A%IA)
        if (allocated(A%JA)) deallocate (A%JA)
        if (allocated(A%vA)) deallocate (A%vA)
        allocate (A%IA(0:rows), A%JA(max(A%nnz, 1)), A%vA(max(A%nnz, 1)))
        call fair_check(c_fair_csr_download(A%dev, A%IA, A%JA, A%vA))
        if (A%nnz == 0) then
            deallocate (A%JA, A%vA)
            allocate (A%JA(0), A%vA(0))
        end if
    end subroutine sync_host

    subroutine free_device(A)
        class(sparse_matrix), intent(inout) :: A
        if (c_associated(A%dev)) call fair_check(c_fair_csr_free(A%dev))
        A%dev = c_null_ptr
    end subroutine free_device

    ! Device handle (for callers that drive the C ABI directly).
    function device_handle(A) result(h)
        class(sparse_matrix), intent(inout) :: A
        type(c_ptr) :: h
        call ensure_dev(A)
        h = A%dev
    end function device_handle

    ! csr_matrowvec, sparse_matrices.f90:252-270
    function csr_matrowvec(A, k, v) result(f)
        class(sparse_matrix), intent(in) :: A
        real(kind=8), intent(in) :: v(:)
        integer, intent(in) :: k
        real(kind=8) :: f
        real(c_double) :: ff
        real(c_double), allocatable :: vc(:)
        type(c_ptr) :: h
        logical :: temp
        call borrow_dev(A, h, temp)
        vc = v
        call fair_check(c_fair_rowvecmult(h, int(k, c_int32_t), vc, int(size(v), c_int32_t), ff))
        call release_dev(h, temp)
        f = ff
    end function csr_matrowvec

    ! sparsematmul / MMRS, sparse_matrices.f90:272-298, 349-368
    function sparsematmul(A, v) result(w)
        class(sparse_matrix), intent(in) :: A
        real(kind=8), intent(in) :: v(:)
        real(kind=8) :: w(A%m)
        real(c_double), allocatable :: vc(:)
        type(c_ptr) :: h
        logical :: temp
        if (size(v) /= A%n) stop "Vector input size incorrect."
        call borrow_dev(A, h, temp)
        vc = v
        if (is_csc(A)) then
            call fair_check(c_fair_transvecmul(h, vc, int(A%n, c_int32_t), w, int(A%m, c_int32_t)))
        else
            call fair_check(c_fair_matvecmul(h, vc, int(A%n, c_int32_t), w, int(A%m, c_int32_t)))
        end if
        call release_dev(h, temp)
    end function sparsematmul

    ! sparsetransmul / MMCS, sparse_matrices.f90:300-347, 370-393
    function sparsetransmul(A, v) result(w)
        class(sparse_matrix) :: A
        real(kind=8), intent(in) :: v(:)
        real(kind=8) :: w(A%n)
        real(c_double), allocatable :: vc(:)
        if (size(v) /= A%m) stop "Vector input size incorrect."
        call ensure_dev(A)
        vc = v
        if (is_csc(A)) then
            call fair_check(c_fair_matvecmul(A%dev, vc, int(A%m, c_int32_t), w, int(A%n, c_int32_t)))
        else
            call fair_check(c_fair_transvecmul(A%dev, vc, int(A%m, c_int32_t), w, int(A%n, c_int32_t)))
        end if
    end function sparsetransmul

    ! transposition / reorder, sparse_matrices.f90:395-411, 492-570
    subroutine transposition(A)
        class(sparse_matrix), intent(inout) :: A
        call ensure_dev(A)
        call fair_check(c_fair_transpose(A%dev))
        call A%sync_host()
    end subroutine transposition

    ! compute_row_norms, sparse_matrices.f90:740-780
    function compute_row_norms(A, p) result(nrA)
        class(sparse_matrix) :: A
        integer, intent(in) :: p
        real(kind=8) :: nrA(A%m)
        call ensure_dev(A)
        if (is_csc(A)) then
            call fair_check(c_fair_column_norms(A%dev, int(p, c_int32_t), nrA, int(A%m, c_int32_t)))
        else
            call fair_check(c_fair_row_norms(A%dev, int(p, c_int32_t), nrA, int(A%m, c_int32_t)))
        end if
    end function compute_row_norms

    ! compute_column_norms, sparse_matrices.f90:782-793
    function compute_column_norms(A, p) result(ncA)
        class(sparse_matrix) :: A
        integer, intent(in) :: p
        real(kind=8) :: ncA(A%n)
        call ensure_dev(A)
        if (is_csc(A)) then
            call fair_check(c_fair_row_norms(A%dev, int(p, c_int32_t), ncA, int(A%n, c_int32_t)))
        else
            call fair_check(c_fair_column_norms(A%dev, int(p, c_int32_t), ncA, int(A%n, c_int32_t)))
        end if
    end function compute_column_norms

    ! add_mult_row, sparse_matrices.f90:795-822
    subroutine add_mult_row(A, k, c, w, lwrb, upb)
        class(CSR_matrix), intent(in) :: A
        integer, intent(in) :: k
        real(kind=8), intent(inout) :: w(A%n), c
        real(kind=8), intent(in), optional :: lwrb, upb
        real(c_double), target :: lb, ub
        type(c_ptr) :: plb, pub, h
        logical :: temp
        call borrow_dev(A, h, temp)
        plb = c_null_ptr
        pub = c_null_ptr
        if (present(lwrb)) then
            lb = lwrb
            plb = c_loc(lb)
        end if
        if (present(upb)) then
            ub = upb
            pub = c_loc(ub)
        end if
        call fair_check(c_fair_add_scalar_mult_row(h, int(k, c_int32_t), real(c, c_double), w, &
                                                   int(A%n, c_int32_t), plb, pub))
        call release_dev(h, temp)
    end subroutine add_mult_row

    ! scalar_mult, sparse_matrices.f90:824-831
    subroutine scalar_mult(A, c)
        class(sparse_matrix), intent(inout) :: A
        real(kind=8) :: c
        call ensure_dev(A)
        call fair_check(c_fair_multiply_by_scalar(A%dev, real(c, c_double)))
        if (allocated(A%vA)) A%vA = c*A%vA
    end subroutine scalar_mult
end module sparse_matrices
