! Driver-side demonstration of the Fortran binding: LS defaults, the LHSA
! pattern of a two-tet mesh and, when a GPU is present, an FSILS_SOLVE of a
! block-diagonal system with a Dirichlet face.
program fsils_gpu_demo
  use iso_c_binding
  use fsils_gpu
  implicit none
  type(fsils_ls_c) :: ls
  integer(c_int) :: rc, nnz, rowPtr(6), i, stat, nd
  integer(c_int), target :: colPtr(32)
  integer(c_int) :: IEN(4,2), gNodes(5), dnode(1)
  real(c_double), target :: relTol, absTol
  integer(c_int), target :: maxItr
  type(c_ptr) :: lhs
  real(c_double) :: Ri(2,5), Val(4,25)

  rc = fsils_gpu_ls_create(ls, LS_TYPE_GMRES, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
       c_null_ptr, c_null_ptr, c_null_ptr)
  print '(a,i0,1x,f4.2,1x,i0,1x,i0)', 'LS ', rc, ls%RI%relTol, ls%RI%mItr, ls%RI%sD

  IEN(:,1) = [1, 2, 3, 4]
  IEN(:,2) = [2, 3, 4, 5]
  rc = svfsi_lhsa(5, 2, 4, IEN, nnz, rowPtr, c_null_ptr)
  rc = svfsi_lhsa(5, 2, 4, IEN, nnz, rowPtr, c_loc(colPtr))
  print '(a,i0,1x,i0,a,6(1x,i0))', 'LHSA ', rc, nnz, ' rowPtr', rowPtr

  nd = svfsi_device_count()
  print '(a,i0)', 'DEVICES ', nd
  if (nd == 0) stop

  ! block-diagonal 2x2 system diag(i, 2i) per node, node 1 clamped
  do i = 1, 5
     gNodes(i) = i
  end do
  rc = fsils_gpu_lhs_create(lhs, 5, 5, nnz, gNodes, rowPtr, colPtr, 1)
  dnode(1) = 1
  rc = fsils_gpu_bc_create(lhs, 1, 1, 2, BC_TYPE_Dir, dnode, c_null_ptr)
  Val = 0.0d0
  do i = 1, 5
     Val(1, diag(i)) = dble(i)
     Val(4, diag(i)) = 2.0d0 * dble(i)
  end do
  Ri = 1.0d0
  relTol = 1.0d-10
  absTol = 1.0d-30
  maxItr = 10
  ls%foC = 0
  rc = fsils_gpu_ls_create(ls, LS_TYPE_GMRES, c_loc(relTol), c_loc(absTol), c_loc(maxItr), c_null_ptr, &
       c_null_ptr, c_null_ptr, c_null_ptr)
  call FSILS_GPU_SOLVE_F(lhs, ls, 2, Ri, Val, PRECOND_FSILS, stat)
  if (stat /= 0) print *, trim(last_error())
  print '(a,i0,1x,i0,10(1x,f8.5))', 'SOLVE ', stat, ls%RI%suc, Ri
  rc = fsils_gpu_lhs_free(lhs)

contains

  integer function diag(r)
    integer(c_int), intent(in) :: r
    integer :: k
    diag = -1
    do k = rowPtr(r), rowPtr(r + 1) - 1
       if (colPtr(k) == r) diag = k
    end do
  end function diag

end program fsils_gpu_demo
