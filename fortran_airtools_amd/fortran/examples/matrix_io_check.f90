! matrix_io_check.f90 -- a Fortran caller of the drop-in's read_csr_matrix and store_matrix
! (system_generation): reads the dump named by argument 1 into a csr_matrix and stores it again
! under the name in argument 2.  No device work.
program matrix_io_check
    use system_generation
    use sparse_matrices
    implicit none
    type(csr_matrix) :: A
    character(len=512) :: fin, fout
    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    call read_csr_matrix(A, trim(fin))
    call store_matrix(A, trim(fout))
end program matrix_io_check
