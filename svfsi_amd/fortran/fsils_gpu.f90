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

  ! codes of svFSILS/FSILS_STRUCT.h:53-62 and of include/svfsi_amd.h
  integer(c_int), parameter :: LS_TYPE_CG = 798, LS_TYPE_GMRES = 797, LS_TYPE_NS = 796, LS_TYPE_BICGS = 795
  integer(c_int), parameter :: PRECOND_FSILS = 701, PRECOND_RCS = 709
  integer(c_int), parameter :: BC_TYPE_Dir = 0, BC_TYPE_Neu = 1
  integer(c_int), parameter :: SVFSI_PHYS_FLUID = 1, SVFSI_PHYS_LELAS = 2, SVFSI_PHYS_STRUCT = 3
  integer(c_int), parameter :: SVFSI_VISC_CONST = 698, SVFSI_VISC_CY = 697, SVFSI_VISC_CASS = 696
  integer(c_int), parameter :: SVFSI_STRUCT_AUTO = 0, SVFSI_STRUCT_VALU = 256, SVFSI_STRUCT_MFMA64 = 512, &
       SVFSI_STRUCT_MFMA32 = 768

  ! FSILS_commuType of svFSILS/FSILS_struct_c.h:60-68 (comm is an MPI_Aint)
  type, bind(C) :: fsils_commu_c
     integer(c_int) :: foC = 0, masF = 0, master = 0, task = 0, tF = 0, nTasks = 0
     integer(c_intptr_t) :: comm = 0
  end type fsils_commu_c

  ! svfsi_eq_params of include/svfsi_amd.h: the eq(cEq) / dmn props of
  ! COMMOD that the element routines read (generalized-alpha constants of
  ! INITIALIZE.f:76-139, material and viscosity-model parameters)
  type, bind(C) :: svfsi_eq_params
     real(c_double) :: am = 0, af = 0, beta = 0, gam = 0, dt = 0
     real(c_double) :: rho = 0
     real(c_double) :: E = 0, nu = 0
     real(c_double) :: fb(3) = 0
     real(c_double) :: mu = 0
     real(c_double) :: bstab = 0
     real(c_double) :: C10 = 0, Kpen = 0
     integer(c_int) :: volType = 2, pad = 0
     integer(c_int) :: viscType = 0, pad2 = 0
     real(c_double) :: mu_o = 0, lam = 0, visc_a = 0, visc_n = 0
  end type svfsi_eq_params

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

     integer(c_int) function fsils_gpu_bc_free(lhs, faIn) bind(C, name="fsils_gpu_bc_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: lhs
       integer(c_int), value :: faIn
     end function fsils_gpu_bc_free

     ! FSILS_SOLVE on DEVICE arrays dRi(dof,nNo), dVal(dof*dof,nnz)
     integer(c_int) function fsils_gpu_solve_dev(lhs, ls, dof, dRi, dVal, prec, incL, res, stream) &
          bind(C, name="fsils_gpu_solve_dev")
       import :: c_int, c_ptr, fsils_ls_c
       type(c_ptr), value :: lhs
       type(fsils_ls_c), intent(inout) :: ls
       integer(c_int), value :: dof, prec
       type(c_ptr), value :: dRi, dVal, incL, res, stream
     end function fsils_gpu_solve_dev

     ! mixed-precision form (FP32 inner, FP64 outer flexible GMRES)
     integer(c_int) function fsils_gpu_solve_mixed_dev(lhs, ls, dof, dRi, dVal, prec, incL, res, stream) &
          bind(C, name="fsils_gpu_solve_mixed_dev")
       import :: c_int, c_ptr, fsils_ls_c
       type(c_ptr), value :: lhs
       type(fsils_ls_c), intent(inout) :: ls
       integer(c_int), value :: dof, prec
       type(c_ptr), value :: dRi, dVal, incL, res, stream
     end function fsils_gpu_solve_mixed_dev

     ! ---- the reference's C API (svFSILS/FSILS_api_c.h:52-64): every
     ! argument by reference; misuse stops the program as the reference does
     subroutine fsils_c_commu_create(commu, commi) bind(C, name="fsils_commu_create_")
       import :: c_intptr_t, fsils_commu_c
       type(fsils_commu_c), intent(inout) :: commu
       integer(c_intptr_t), intent(in) :: commi
     end subroutine fsils_c_commu_create

     subroutine fsils_c_commu_free(commu) bind(C, name="fsils_commu_free_")
       import :: fsils_commu_c
       type(fsils_commu_c), intent(inout) :: commu
     end subroutine fsils_c_commu_free

     subroutine fsils_c_lhs_create(lhs, commu, gnNo, nNo, nnz, gNodes, rowPtr, colPtr, nFaces) &
          bind(C, name="fsils_lhs_create_c_")
       import :: c_int, c_ptr, fsils_commu_c
       type(c_ptr), intent(out) :: lhs
       type(fsils_commu_c), intent(in) :: commu
       integer(c_int), intent(in) :: gnNo, nNo, nnz, nFaces
       integer(c_int), intent(in) :: gNodes(*), rowPtr(*), colPtr(*)
     end subroutine fsils_c_lhs_create

     subroutine fsils_c_lhs_free(lhs) bind(C, name="fsils_lhs_free_")
       import :: c_ptr
       type(c_ptr), value :: lhs
     end subroutine fsils_c_lhs_free

     subroutine fsils_c_ls_create(ls, LS_type, relTol, absTol, maxItr, dimKry, relTolIn, absTolIn, maxItrIn) &
          bind(C, name="fsils_ls_create_")
       import :: c_int, c_double, fsils_ls_c
       type(fsils_ls_c), intent(inout) :: ls
       integer(c_int), intent(in) :: LS_type
       real(c_double), intent(in), optional :: relTol, absTol, relTolIn(2), absTolIn(2)
       integer(c_int), intent(in), optional :: maxItr, dimKry, maxItrIn(2)
     end subroutine fsils_c_ls_create

     subroutine fsils_c_ls_free(ls) bind(C, name="fsils_ls_free_")
       import :: fsils_ls_c
       type(fsils_ls_c), intent(inout) :: ls
     end subroutine fsils_c_ls_free

     subroutine fsils_c_bc_create(lhs, faIn, nNo, dof, BC_type, gNodes, Val) bind(C, name="fsils_bc_create_")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: lhs
       integer(c_int), intent(in) :: faIn, nNo, dof, BC_type
       integer(c_int), intent(in) :: gNodes(*)
       real(c_double), intent(in), optional :: Val(*)
     end subroutine fsils_c_bc_create

     subroutine fsils_c_bc_free(lhs, faIn) bind(C, name="fsils_bc_free_")
       import :: c_int, c_ptr
       type(c_ptr), value :: lhs
       integer(c_int), intent(in) :: faIn
     end subroutine fsils_c_bc_free

     subroutine fsils_c_solve(lhs, ls, dof, Ri, Val, prec, incL, res) bind(C, name="fsils_solve_")
       import :: c_int, c_ptr, c_double, fsils_ls_c
       type(c_ptr), value :: lhs
       type(fsils_ls_c), intent(inout) :: ls
       integer(c_int), intent(in) :: dof, prec
       real(c_double), intent(inout) :: Ri(*), Val(*)
       integer(c_int), intent(in), optional :: incL(*)
       real(c_double), intent(in), optional :: res(*)
     end subroutine fsils_c_solve

     ! ---- assembly: the element loop of GLOBALEQASSEM (EQASSEM.f:45-84)
     integer(c_int) function svfsi_mesh_create(mesh, lhs, nNo, nEl, eNoN, IEN, x) bind(C, name="svfsi_mesh_create")
       import :: c_int, c_ptr, c_double
       type(c_ptr), intent(out) :: mesh
       type(c_ptr), value :: lhs
       integer(c_int), value :: nNo, nEl, eNoN
       integer(c_int), intent(in) :: IEN(*)
       real(c_double), intent(in) :: x(*)
     end function svfsi_mesh_create

     integer(c_int) function svfsi_mesh_free(mesh) bind(C, name="svfsi_mesh_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: mesh
     end function svfsi_mesh_free

     ! tri3 boundary face: fIEN(3,nFe) 1-based nodes, fEl(nFe) 1-based element
     integer(c_int) function svfsi_mesh_add_face(mesh, nFe, fIEN, fEl, faceId) bind(C, name="svfsi_mesh_add_face")
       import :: c_int, c_ptr
       type(c_ptr), value :: mesh
       integer(c_int), value :: nFe
       integer(c_int), intent(in) :: fIEN(*), fEl(*)
       integer(c_int), intent(out) :: faceId
     end function svfsi_mesh_add_face

     ! LSALLOC (svFSI/LS.f:338-366): zero dVal and dR
     integer(c_int) function svfsi_zero_dev(mesh, dof, dVal, dR, stream) bind(C, name="svfsi_zero_dev")
       import :: c_int, c_ptr
       type(c_ptr), value :: mesh
       integer(c_int), value :: dof
       type(c_ptr), value :: dVal, dR, stream
     end function svfsi_zero_dev

     integer(c_int) function svfsi_assemble_dev(mesh, phys, p, tDof, dAg, dYg, dDg, dprop_el, quad, dVal, dR, &
          stream) bind(C, name="svfsi_assemble_dev")
       import :: c_int, c_ptr, svfsi_eq_params
       type(c_ptr), value :: mesh
       integer(c_int), value :: phys, tDof, quad
       type(svfsi_eq_params), intent(in) :: p
       type(c_ptr), value :: dAg, dYg, dDg, dprop_el, dVal, dR, stream
     end function svfsi_assemble_dev

     ! SETBCTRACL (svFSI/SETBC.f:319-416), uniform traction h(3)
     integer(c_int) function svfsi_assemble_traction_dev(mesh, faceId, dof, h, dR, stream) &
          bind(C, name="svfsi_assemble_traction_dev")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: mesh
       integer(c_int), value :: faceId, dof
       real(c_double), intent(in) :: h(3)
       type(c_ptr), value :: dR, stream
     end function svfsi_assemble_traction_dev

     ! BASSEMNEUBC + BFLUID with h = 0 (EQASSEM.f:90-192, FLUID.f:1280-1336)
     integer(c_int) function svfsi_assemble_backflow_dev(mesh, faceId, p, tDof, dYg, dVal, dR, stream) &
          bind(C, name="svfsi_assemble_backflow_dev")
       import :: c_int, c_ptr, svfsi_eq_params
       type(c_ptr), value :: mesh
       integer(c_int), value :: faceId, tDof
       type(svfsi_eq_params), intent(in) :: p
       type(c_ptr), value :: dYg, dVal, dR, stream
     end function svfsi_assemble_backflow_dev

     ! ---- device memory for a driver without a HIP binding
     integer(c_int) function svfsi_dev_malloc(dptr, bytes) bind(C, name="svfsi_dev_malloc")
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), intent(out) :: dptr
       integer(c_size_t), value :: bytes
     end function svfsi_dev_malloc

     integer(c_int) function svfsi_dev_free(dptr) bind(C, name="svfsi_dev_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: dptr
     end function svfsi_dev_free

     integer(c_int) function svfsi_memcpy_h2d(dst, src, bytes) bind(C, name="svfsi_memcpy_h2d")
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), value :: dst, src
       integer(c_size_t), value :: bytes
     end function svfsi_memcpy_h2d

     integer(c_int) function svfsi_memcpy_d2h(dst, src, bytes) bind(C, name="svfsi_memcpy_d2h")
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), value :: dst, src
       integer(c_size_t), value :: bytes
     end function svfsi_memcpy_d2h

     ! ---- mesh files (READCCNE, svFSI/LOADMSH.f:125-239): sizes, then data
     integer(c_int) function svfsi_read_ccne_sizes(connPath, coordPath, nEl, eNoN, nNo, nsd) &
          bind(C, name="svfsi_read_ccne_sizes")
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: connPath(*), coordPath(*)
       integer(c_int), intent(out) :: nEl, eNoN, nNo, nsd
     end function svfsi_read_ccne_sizes

     integer(c_int) function svfsi_read_ccne(connPath, coordPath, nEl, eNoN, nNo, nsd, ien, x) &
          bind(C, name="svfsi_read_ccne")
       import :: c_int, c_char, c_double
       character(kind=c_char), intent(in) :: connPath(*), coordPath(*)
       integer(c_int), value :: nEl, eNoN, nNo, nsd
       integer(c_int), intent(out) :: ien(*)
       real(c_double), intent(out) :: x(*)
     end function svfsi_read_ccne

     integer(c_int) function svfsi_read_ebc_sizes(path, nFe, feNoN) bind(C, name="svfsi_read_ebc_sizes")
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), intent(out) :: nFe, feNoN
     end function svfsi_read_ebc_sizes

     integer(c_int) function svfsi_read_ebc(path, nFe, feNoN, gE, fien) bind(C, name="svfsi_read_ebc")
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), value :: nFe, feNoN
       integer(c_int), intent(out) :: gE(*), fien(*)
     end function svfsi_read_ebc

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
