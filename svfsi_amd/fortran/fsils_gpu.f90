! fsils_gpu.f90 -- Fortran binding of libsvfsi_amd (include/svfsi_amd.h).
!
! The svFSI driver keeps its FSILS_* call sequence (svFSILS/FSILS_API.h:51-135)
! and reaches the GPU through these iso_c_binding interfaces.  The bind(C)
! types have the memory layout of svFSILS/FSILS_struct_c.h:115-140, i.e. of
! the reference's SEQUENCE types with default (4-byte) LOGICAL, so a driver
! may pass its own eq%FSILS through C_LOC as well.
module fsils_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: LS_TYPE_CG = 798, LS_TYPE_GMRES = 797
  integer(c_int), parameter :: PRECOND_FSILS = 701
  integer(c_int), parameter :: BC_TYPE_Dir = 0, BC_TYPE_Neu = 1

  type, bind(C) :: fsils_sub_ls_c
     integer(c_int) :: suc, mItr, sD, itr, cM, cN, cD, reserve
     real(c_double) :: absTol, relTol, iNorm, fNorm, dB, callD
  end type fsils_sub_ls_c

  type, bind(C) :: fsils_ls_c
     integer(c_int) :: foC = 0, LS_type = 0, Resm = 0, Resc = 0
     type(fsils_sub_ls_c) :: GM, CG, RI
  end type fsils_ls_c

  interface
     integer(c_int) function svfsi_device_count() bind(C, name="svfsi_device_count")
       import :: c_int
     end function svfsi_device_count

     type(c_ptr) function svfsi_last_error() bind(C, name="svfsi_last_error")
       import :: c_ptr
     end function svfsi_last_error

     integer(c_int) function fsils_gpu_ls_create(ls, LS_type, relTol, absTol, maxItr, dimKry, &
          relTolIn, absTolIn, maxItrIn) bind(C, name="fsils_gpu_ls_create")
       import :: c_int, c_ptr, fsils_ls_c
       type(fsils_ls_c), intent(inout) :: ls
       integer(c_int), value :: LS_type
       type(c_ptr), value :: relTol, absTol, maxItr, dimKry, relTolIn, absTolIn, maxItrIn
     end function fsils_gpu_ls_create

     integer(c_int) function svfsi_lhsa(nNo, nEl, eNoN, IEN, nnz, rowPtr, colPtr) bind(C, name="svfsi_lhsa")
       import :: c_int, c_ptr
       integer(c_int), value :: nNo, nEl, eNoN
       integer(c_int), intent(in) :: IEN(*)
       integer(c_int), intent(out) :: nnz
       integer(c_int), intent(out) :: rowPtr(*)
       type(c_ptr), value :: colPtr
     end function svfsi_lhsa

     integer(c_int) function fsils_gpu_lhs_create(lhs, gnNo, nNo, nnz, gNodes, rowPtr, colPtr, nFaces) &
          bind(C, name="fsils_gpu_lhs_create")
       import :: c_int, c_ptr
       type(c_ptr), intent(out) :: lhs
       integer(c_int), value :: gnNo, nNo, nnz, nFaces
       integer(c_int), intent(in) :: gNodes(*), rowPtr(*), colPtr(*)
     end function fsils_gpu_lhs_create

     integer(c_int) function fsils_gpu_lhs_free(lhs) bind(C, name="fsils_gpu_lhs_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: lhs
     end function fsils_gpu_lhs_free

     integer(c_int) function fsils_gpu_bc_create(lhs, faIn, nNo, dof, BC_type, gNodes, Val) &
          bind(C, name="fsils_gpu_bc_create")
       import :: c_int, c_ptr
       type(c_ptr), value :: lhs
       integer(c_int), value :: faIn, nNo, dof, BC_type
       integer(c_int), intent(in) :: gNodes(*)
       type(c_ptr), value :: Val
     end function fsils_gpu_bc_create

     integer(c_int) function fsils_gpu_solve(lhs, ls, dof, Ri, Val, prec, incL, res) bind(C, name="fsils_gpu_solve")
       import :: c_int, c_ptr, c_double, fsils_ls_c
       type(c_ptr), value :: lhs
       type(fsils_ls_c), intent(inout) :: ls
       integer(c_int), value :: dof, prec
       real(c_double), intent(inout) :: Ri(*), Val(*)
       type(c_ptr), value :: incL, res
     end function fsils_gpu_solve

     ! CHECKIEN (svFSI/READMSH.f:1010-1185) for tet4 / tet10: ien(eNoN, nEl)
     ! 1-based, reordered in place; x(3, nNo)
     integer(c_int) function svfsi_checkien(eNoN, nEl, nNo, ien, x, nReordered) bind(C, name="svfsi_checkien")
       import :: c_int, c_double
       integer(c_int), value :: eNoN, nEl, nNo
       integer(c_int), intent(inout) :: ien(*)
       real(c_double), intent(in) :: x(*)
       integer(c_int), intent(out) :: nReordered
     end function svfsi_checkien

     ! .vtu / .vtp meshes and faces (svFSI/VTKXML.f): sizes, then the data
     integer(c_int) function svfsi_read_vtk_sizes(path, nPoints, nCells, nodesPerCell, hasNodeIds, hasElementIds) &
          bind(C, name="svfsi_read_vtk_sizes")
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), intent(out) :: nPoints, nCells, nodesPerCell, hasNodeIds, hasElementIds
     end function svfsi_read_vtk_sizes

     integer(c_int) function svfsi_read_vtk(path, nPoints, nCells, nodesPerCell, x, conn, nodeIds, elementIds) &
          bind(C, name="svfsi_read_vtk")
       import :: c_int, c_char, c_double, c_ptr
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), value :: nPoints, nCells, nodesPerCell
       real(c_double), intent(out) :: x(*)
       integer(c_int), intent(out) :: conn(*)
       type(c_ptr), value :: nodeIds, elementIds
     end function svfsi_read_vtk
  end interface

contains

  ! FSILS_SOLVE argument list (FSILS_API.h:125-134) on the GPU: Ri(dof,nNo)
  ! in = RHS, out = solution; Val(dof*dof,nnz) is left scaled in place.
  subroutine FSILS_GPU_SOLVE_F(lhs, ls, dof, Ri, Val, prec, stat)
    type(c_ptr), intent(in) :: lhs
    type(fsils_ls_c), intent(inout) :: ls
    integer(c_int), intent(in) :: dof, prec
    real(c_double), intent(inout) :: Ri(:,:), Val(:,:)
    integer(c_int), intent(out) :: stat
    stat = fsils_gpu_solve(lhs, ls, dof, Ri, Val, prec, c_null_ptr, c_null_ptr)
  end subroutine FSILS_GPU_SOLVE_F

  ! the same with the optional incL(nFaces) and res(nFaces) of FSILS_SOLVE
  ! (coupled Neumann faces, SOLVE.f:66-86)
  subroutine FSILS_GPU_SOLVE_RES_F(lhs, ls, dof, Ri, Val, prec, incL, res, stat)
    type(c_ptr), intent(in) :: lhs
    type(fsils_ls_c), intent(inout) :: ls
    integer(c_int), intent(in) :: dof, prec
    real(c_double), intent(inout) :: Ri(:,:), Val(:,:)
    integer(c_int), intent(in), target :: incL(:)
    real(c_double), intent(in), target :: res(:)
    integer(c_int), intent(out) :: stat
    stat = fsils_gpu_solve(lhs, ls, dof, Ri, Val, prec, c_loc(incL), c_loc(res))
  end subroutine FSILS_GPU_SOLVE_RES_F

  function last_error() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ' '
    call c_f_pointer(svfsi_last_error(), p, [256])
    do i = 1, 256
       if (p(i) == c_null_char) exit
       msg(i:i) = p(i)
    end do
  end function last_error

end module fsils_gpu
