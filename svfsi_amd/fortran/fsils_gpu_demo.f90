! Driver-side demonstration of the Fortran binding.
!
! Without arguments: the LS defaults, the LHSA pattern of a two-tet mesh and,
! when a GPU is present, an FSILS_SOLVE of a block-diagonal system with a
! Dirichlet face (host arrays).
!
! With a directory argument: the hooks a maintainer adds to svFSI, in the
! order the driver reaches them, on a linear-elasticity mesh read from that
! directory (coordinate/connectivity files and two face files, the READCCNE
! formats): mesh files and CHECKIEN (READMSH), LHSA, FSILS_COMMU_CREATE and
! FSILS_LHS_CREATE (INITIALIZE.f:191-208), FSILS_BC_CREATE (BAFINI.f:510),
! the element loop of GLOBALEQASSEM with the traction face (EQASSEM.f:45-84,
! SETBC.f:319-416) and FSILS_SOLVE (LS.f:397-398), the solve on device arrays.
! x = 0 is clamped (x0.ebc), a traction of 1e3 along x loads x = 1 (x1.ebc).
program fsils_gpu_demo
  use iso_c_binding
  use fsils_gpu
  implicit none
  character(len=512) :: dir

  if (command_argument_count() == 0) then
     call small_checks()
  else
     call get_command_argument(1, dir)
     call lelas_hooks(trim(dir))
  end if

contains

  subroutine small_checks()
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
    if (nd == 0) return

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

  end subroutine small_checks

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
       print '(a,a,a,i0,a,a)', 'ERROR ', what, ' rc=', rc, ' ', trim(last_error())
       stop 1
    end if
  end subroutine check

  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len(s) + 1) :: c
    c = s // c_null_char
  end function cstr

  subroutine lelas_hooks(d)
    character(len=*), intent(in) :: d
    integer(c_int), parameter :: dof = 3, tDof = 3
    integer(c_int) :: nEl, eNoN, nNo, nsd, nnz, nRe, nF0, nF1, fN, faceId, a, k
    integer(c_int), allocatable, target :: IEN(:,:), rowPtr(:), colPtr(:), gNodes(:), gE0(:), f0(:,:), gE1(:), &
         f1(:,:), dirNodes(:)
    real(c_double), allocatable, target :: x(:,:), zero(:,:), R(:,:)
    logical, allocatable :: onFace(:)
    type(c_ptr) :: lhs, msh, dAg, dYg, dDg, dVal, dR
    type(fsils_commu_c) :: commu
    type(fsils_ls_c) :: ls
    type(svfsi_eq_params) :: prm
    integer(c_intptr_t) :: commi
    integer(c_size_t) :: nbR, nbV
    real(c_double) :: h(3)

    ! READMSH: coordinate/connectivity files, then CHECKIEN
    call check(svfsi_read_ccne_sizes(cstr(d // '/conn.txt'), cstr(d // '/coord.txt'), nEl, eNoN, nNo, nsd), 'sizes')
    allocate(IEN(eNoN, nEl), x(nsd, nNo))
    call check(svfsi_read_ccne(cstr(d // '/conn.txt'), cstr(d // '/coord.txt'), nEl, eNoN, nNo, nsd, IEN, x), 'ccne')
    call check(svfsi_checkien(eNoN, nEl, nNo, IEN, x, nRe), 'checkien')
    call check(svfsi_read_ebc_sizes(cstr(d // '/x0.ebc'), nF0, fN), 'ebc sizes')
    allocate(gE0(nF0), f0(fN, nF0))
    call check(svfsi_read_ebc(cstr(d // '/x0.ebc'), nF0, fN, gE0, f0), 'ebc x0')
    call check(svfsi_read_ebc_sizes(cstr(d // '/x1.ebc'), nF1, fN), 'ebc sizes')
    allocate(gE1(nF1), f1(fN, nF1))
    call check(svfsi_read_ebc(cstr(d // '/x1.ebc'), nF1, fN, gE1, f1), 'ebc x1')
    print '(a,4(1x,i0))', 'MESH', nNo, nEl, eNoN, nRe

    ! INITIALIZE.f:191-208: LHSA, FSILS_COMMU_CREATE, FSILS_LHS_CREATE
    allocate(rowPtr(nNo + 1), gNodes(nNo))
    call check(svfsi_lhsa(nNo, nEl, eNoN, IEN, nnz, rowPtr, c_null_ptr), 'lhsa')
    allocate(colPtr(nnz))
    call check(svfsi_lhsa(nNo, nEl, eNoN, IEN, nnz, rowPtr, c_loc(colPtr)), 'lhsa')
    gNodes = [(a, a = 1, nNo)]
    commi = 0
    call fsils_c_commu_create(commu, commi)
    call fsils_c_lhs_create(lhs, commu, nNo, nNo, nnz, gNodes, rowPtr, colPtr, 1)
    call check(svfsi_mesh_create(msh, lhs, nNo, nEl, eNoN, IEN, x), 'mesh')

    ! BAFINI.f:510: the clamped face's nodes as a Dirichlet BC face
    allocate(onFace(nNo))
    onFace = .false.
    do k = 1, nF0
       onFace(f0(:, k)) = .true.
    end do
    dirNodes = pack([(a, a = 1, nNo)], onFace)
    call fsils_c_bc_create(lhs, 1, size(dirNodes), dof, BC_TYPE_Dir, dirNodes)
    call check(svfsi_mesh_add_face(msh, nF1, f1, gE1, faceId), 'face')

    ! the state and the system stay in device memory across Newton iterations
    nbR = int(8, c_size_t) * dof * nNo
    nbV = int(8, c_size_t) * dof * dof * nnz
    allocate(zero(tDof, nNo), R(dof, nNo))
    zero = 0.0d0
    call check(svfsi_dev_malloc(dAg, nbR), 'malloc')
    call check(svfsi_dev_malloc(dYg, nbR), 'malloc')
    call check(svfsi_dev_malloc(dDg, nbR), 'malloc')
    call check(svfsi_dev_malloc(dVal, nbV), 'malloc')
    call check(svfsi_dev_malloc(dR, nbR), 'malloc')
    call check(svfsi_memcpy_h2d(dAg, c_loc(zero), nbR), 'h2d')
    call check(svfsi_memcpy_h2d(dYg, c_loc(zero), nbR), 'h2d')
    call check(svfsi_memcpy_h2d(dDg, c_loc(zero), nbR), 'h2d')

    ! EQASSEM.f:45-84 and SETBC.f:319-416: one call each per Newton iteration
    prm%rho = 1.0d0
    prm%E = 1.0d6
    prm%nu = 0.3d0
    prm%dt = 1.0d0
    prm%am = (2.0d0 - 0.2d0) / 1.2d0
    prm%af = 1.0d0 / 1.2d0
    prm%beta = 0.25d0 * (1.0d0 + prm%am - prm%af)**2
    prm%gam = 0.5d0 + prm%am - prm%af
    h = [1.0d3, 0.0d0, 0.0d0]
    call check(svfsi_zero_dev(msh, dof, dVal, dR, c_null_ptr), 'zero')
    call check(svfsi_assemble_dev(msh, SVFSI_PHYS_LELAS, prm, tDof, dAg, dYg, dDg, c_null_ptr, 0, dVal, dR, &
         c_null_ptr), 'assemble')
    call check(svfsi_assemble_traction_dev(msh, faceId, dof, h, dR, c_null_ptr), 'traction')

    ! READFILES.f:1210 and LS.f:397-398: FSILS_LS_CREATE, FSILS_SOLVE
    call fsils_c_ls_create(ls, LS_TYPE_CG, relTol=1.0d-10, absTol=1.0d-30, maxItr=2000)
    call check(fsils_gpu_solve_dev(lhs, ls, dof, dR, dVal, PRECOND_FSILS, c_null_ptr, c_null_ptr, c_null_ptr), &
         'solve')
    call check(svfsi_memcpy_d2h(c_loc(R), dR, nbR), 'd2h')
    print '(a,2(1x,i0),2(1x,es24.16))', 'STAT', ls%RI%suc, ls%RI%itr, ls%RI%iNorm, ls%RI%fNorm
    do a = 1, nNo
       print '(a,1x,i0,3(1x,es24.16))', 'SOL', a, R(:, a)
    end do

    call fsils_c_ls_free(ls)
    call fsils_c_bc_free(lhs, 1)
    call check(svfsi_mesh_free(msh), 'mesh free')
    call fsils_c_lhs_free(lhs)
    call fsils_c_commu_free(commu)
    call check(svfsi_dev_free(dAg), 'free')
    call check(svfsi_dev_free(dYg), 'free')
    call check(svfsi_dev_free(dDg), 'free')
    call check(svfsi_dev_free(dVal), 'free')
    call check(svfsi_dev_free(dR), 'free')
  end subroutine lelas_hooks

end program fsils_gpu_demo
