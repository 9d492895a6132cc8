! sweep_check.f90 -- a reference-style Fortran caller of quicksort (system_generation) and
! basic_sweep (AIRMs).
!
!   sweep_check sort  <in> <out>   in: nn, arr(nn) (stream); out: sorted arr, IND (host only)
!   sweep_check sweep <in> <out>   in: m, n, nnz, IA(0:m), JA, vA, b(m), nrA(m), cumul(m), x0(n),
!                                  then per case: nord, ord(nord), rowprob, omg, haslb, lb, hasub, ub
!                                  (3 cases; runtime generator seeded with 77 before each);
!                                  out: x after basic_sweep per case
program sweep_check
    use system_generation
    use sparse_matrices
    use AIRMs
    implicit none
    character(len=512) :: mode, fin, fout
    integer :: u, v, nn, c, nord, rp, hl, hu, k, ns
    integer, allocatable :: ind(:), ordv(:), seed(:)
    real(kind=8), allocatable :: arr(:), b(:), nrA(:), cumul(:), x0(:), x(:)
    real(kind=8) :: omg, lb, ub
    type(csr_matrix) :: A
    call get_command_argument(1, mode)
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', action='read')
    open (newunit=v, file=trim(fout), access='stream', form='unformatted', action='write', status='replace')
    if (trim(mode) == 'sort') then
        read (u) nn
        allocate (arr(nn), ind(nn))
        read (u) arr
        ind = [(k, k=1, nn)]
        call quicksort(arr, ind)
        write (v) arr, ind
    else
        read (u) A%m, A%n, A%nnz
        allocate (A%IA(0:A%m), A%JA(A%nnz), A%vA(A%nnz))
        read (u) A%IA, A%JA, A%vA
        allocate (b(A%m), nrA(A%m), cumul(A%m), x0(A%n), x(A%n), ordv(2*A%m))
        read (u) b, nrA, cumul, x0
        call random_seed(size=ns)
        allocate (seed(ns))
        do c = 1, 3
            read (u) nord
            ordv = 0
            read (u) ordv(1:nord)
            read (u) rp, omg, hl, lb, hu, ub
            do k = 1, ns
                seed(k) = 77 + 7919*k
            end do
            call random_seed(put=seed)
            x = x0
            call basic_sweep(x, A, b, ordv, nord, rp /= 0, rp /= 0, omg, nrA, cumul, hu /= 0, hl /= 0, ub, lb)
            write (v) x
        end do
    end if
    close (u)
    close (v)
end program sweep_check
