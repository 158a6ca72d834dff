! TEST INFRASTRUCTURE ONLY.  Harness around the reference's own Fortran
! routines, compiled from the reference sources by oracle/ref.mk into
! oracle/_ref/.  It writes the fixtures that pin the C restatement in
! oracle/svfsi_oracle.c; nothing in the product path runs it.
!
! Modes (argument 1), with an input and an output file (stream, native
! little-endian):
!   asm    element loop + DOASSEM of one equation on one mesh:
!          svFSI/LHSA.f:38-262 (pattern), svFSI/NN.f:41-174 (SELECTELE),
!          svFSI/FS.f:39-97 (INITFSMSH), then CONSTRUCT_FLUID
!          (svFSI/FLUID.f:40-190), CONSTRUCT_LELAS (svFSI/LELAS.f:38-128)
!          or CONSTRUCT_dSOLID (svFSI/STRUCT.f:39-162).
!   solve  svFSILS on one rank: FSILS_COMMU_CREATE, FSILS_LHS_CREATE,
!          FSILS_BC_CREATE (one Dirichlet face), FSILS_LS_CREATE and
!          FSILS_SOLVE (svFSILS/SOLVE.f:51-143).
!
! asm input:  int32 hdr(16), real64 par(24), x(3,nNo), int32 IEN(eNoN,nEl)
!             (1-based), Ag, Yg, Dg (tDof,nNo), then pS0(6,nNo) if hdr(9)
!             and Bf(3,nNo) if hdr(10).
!   hdr: phys, nNo, nEl, eNoN, nFs, viscType, volType, isoType, hasPS0,
!        hasBf, nrep
!   par: dt, am, af, beta, gam, rho, E, nu, mu_i, mu_o, lam, visc_a,
!        visc_n, C10, Kpen, f_x, f_y, f_z, solid_viscosity, damping
! asm output: int32 nnz, rowPtr(nNo+1), colPtr(nnz) (1-based);
!             real64 Val(dof*dof,nnz), R(dof,nNo), seconds per assembly.
!
! solve input: int32 hdr(16), real64 par(16), rowPtr(nNo+1), colPtr(nnz)
!              (1-based), real64 Val(dof*dof,nnz), R(dof,nNo), int32
!              dirNodes(nDir), real64 dirVal(dof,nDir).
!   hdr: dof, nNo, nnz, LS_type, prec, nDir, nrep, RI%mItr, RI%sD,
!        GM%mItr, GM%sD, CG%mItr
!   par: RI%relTol, RI%absTol, GM%relTol, GM%absTol, CG%relTol, CG%absTol
!        (a field left at 0 keeps FSILS_LS_CREATE's default)
! solve output: real64 x(dof,nNo); int32 suc, itr, cM, cN, cD and the
!               same for GM and CG; real64 RI iNorm, fNorm, dB, callD;
!               seconds per solve.
program svfsi_ref_harness
   use commod
   use allfun
   implicit none
   character(len=1024) :: mode, fin, fout
   integer :: ierr

   call get_command_argument(1, mode)
   call get_command_argument(2, fin)
   call get_command_argument(3, fout)
   call MPI_INIT(ierr)
   select case (trim(mode))
   case ("asm")
      call run_asm(trim(fin), trim(fout))
   case ("solve")
      call run_solve(trim(fin), trim(fout))
   case default
      print *, "usage: svfsi_ref_harness asm|solve <in> <out>"
      stop 2
   end select
   call MPI_FINALIZE(ierr)

contains

   real(8) function wall()
      integer(8) :: c, r
      call system_clock(c, r)
      wall = real(c, 8)/real(r, 8)
   end function wall

   subroutine run_asm(fin, fout)
      character(len=*), intent(in) :: fin, fout
      integer(4) :: hdr(16), nnzl, i, rep, nrep
      real(8) :: par(24), t0, secs
      real(8), allocatable :: Ag(:,:), Yg(:,:), Dg(:,:)
      integer :: u

      open(newunit=u, file=fin, access="stream", form="unformatted", status="old")
      read(u) hdr, par
      nsd = 3
      nsymd = 6
      nMsh = 1
      nEq = 1
      cEq = 1
      call cm%new(MPI_COMM_WORLD)
      shlEq = .false.
      pstEq = .false.
      ecCpld = .false.
      tnNo = hdr(2)
      gtnNo = tnNo
      select case (hdr(1))
      case (phys_fluid)
         dof = 4
      case default
         dof = 3
      end select
      tDof = dof
      dt = par(1)
      mvMsh = .false.

      allocate(msh(1))
      msh(1)%eNoN = hdr(4)
      msh(1)%nEl = hdr(3)
      msh(1)%gnEl = hdr(3)
      msh(1)%nFs = max(hdr(5), 1)
      msh(1)%nNo = tnNo
      msh(1)%gnNo = tnNo
      allocate(x(3, tnNo), msh(1)%IEN(msh(1)%eNoN, msh(1)%nEl))
      read(u) x
      read(u) msh(1)%IEN
      allocate(Ag(tDof, tnNo), Yg(tDof, tnNo), Dg(tDof, tnNo))
      read(u) Ag, Yg, Dg
      allocate(Bf(3, tnNo))
      Bf = 0d0
      if (hdr(9) /= 0) then
         allocate(pS0(6, tnNo))
         read(u) pS0
      end if
      if (hdr(10) /= 0) read(u) Bf
      close(u)

      call SELECTELE(msh(1))
      call INITFSMSH(msh(1))

      allocate(eq(1))
      eq(1)%nDmn = 1
      eq(1)%nBc = 0
      eq(1)%s = 1
      eq(1)%dof = dof
      eq(1)%am = par(2)
      eq(1)%af = par(3)
      eq(1)%beta = par(4)
      eq(1)%gam = par(5)
      allocate(eq(1)%dmn(1))
      eq(1)%dmn(1)%Id = -1
      eq(1)%dmn(1)%phys = hdr(1)
      eq(1)%phys = hdr(1)
      eq(1)%dmn(1)%prop = 0d0
      eq(1)%dmn(1)%prop(fluid_density) = par(6)
      eq(1)%dmn(1)%prop(solid_density) = par(6)
      eq(1)%dmn(1)%prop(elasticity_modulus) = par(7)
      eq(1)%dmn(1)%prop(poisson_ratio) = par(8)
      eq(1)%dmn(1)%prop(f_x) = par(16)
      eq(1)%dmn(1)%prop(f_y) = par(17)
      eq(1)%dmn(1)%prop(f_z) = par(18)
      eq(1)%dmn(1)%prop(solid_viscosity) = par(19)
      eq(1)%dmn(1)%prop(damping) = par(20)
      eq(1)%dmn(1)%visc%viscType = hdr(6)
      eq(1)%dmn(1)%visc%mu_i = par(9)
      eq(1)%dmn(1)%visc%mu_o = par(10)
      eq(1)%dmn(1)%visc%lam = par(11)
      eq(1)%dmn(1)%visc%a = par(12)
      eq(1)%dmn(1)%visc%n = par(13)
      eq(1)%dmn(1)%stM%isoType = hdr(8)
      eq(1)%dmn(1)%stM%volType = hdr(7)
      eq(1)%dmn(1)%stM%C10 = par(14)
      eq(1)%dmn(1)%stM%Kpen = par(15)

      call LHSA(nnzl)
      allocate(R(dof, tnNo), Val(dof*dof, nnzl))

      nrep = max(hdr(11), 1)
      t0 = wall()
      do rep = 1, nrep
         R = 0d0
         Val = 0d0
         select case (hdr(1))
         case (phys_fluid)
            call CONSTRUCT_FLUID(msh(1), Ag, Yg)
            ! P2P1 (nFs = 2): the pressure rows of the mid-edge nodes,
            ! as svFSI's MAIN does after the assembly (MAIN.f:167-173)
            if (msh(1)%nFs == 2) call THOOD_ValRC()
         case (phys_lElas)
            call CONSTRUCT_LELAS(msh(1), Ag, Dg)
         case (phys_struct)
            call CONSTRUCT_dSOLID(msh(1), Ag, Yg, Dg)
         case default
            print *, "unknown phys", hdr(1)
            stop 3
         end select
      end do
      secs = (wall() - t0)/nrep

      open(newunit=u, file=fout, access="stream", form="unformatted", status="replace")
      write(u) nnzl
      write(u) (rowPtr(i), i = 1, tnNo + 1)
      write(u) (colPtr(i), i = 1, nnzl)
      write(u) Val, R, secs
      close(u)
   end subroutine run_asm

   subroutine run_solve(fin, fout)
      character(len=*), intent(in) :: fin, fout
      integer(4) :: hdr(16), d, nNo, nnzl, nDir, rep, nrep, i, stats(15)
      real(8) :: par(16), t0, secs, res(1)
      integer(4), allocatable :: rp(:), cp(:), gN(:), dirN(:)
      integer(4) :: incL(1)
      real(8), allocatable :: V0(:,:), R0(:,:), V(:,:), Ri(:,:), dirV(:,:)
      type(FSILS_commuType) :: commu
      type(FSILS_lhsType) :: lhs
      type(FSILS_lsType) :: ls
      integer :: u

      open(newunit=u, file=fin, access="stream", form="unformatted", status="old")
      read(u) hdr, par
      d = hdr(1)
      nNo = hdr(2)
      nnzl = hdr(3)
      nDir = hdr(6)
      allocate(rp(nNo + 1), cp(nnzl), V0(d*d, nnzl), R0(d, nNo), dirN(max(nDir, 1)), dirV(d, max(nDir, 1)))
      read(u) rp, cp, V0, R0
      if (nDir > 0) read(u) dirN(1:nDir), dirV(:, 1:nDir)
      close(u)

      allocate(gN(nNo))
      do i = 1, nNo
         gN(i) = i
      end do
      call FSILS_COMMU_CREATE(commu, MPI_COMM_WORLD)
      call FSILS_LHS_CREATE(lhs, commu, nNo, nNo, nnzl, gN, rp, cp, 1)
      if (nDir > 0) then
         call FSILS_BC_CREATE(lhs, 1, nDir, d, BC_TYPE_Dir, dirN(1:nDir), dirV(:, 1:nDir))
      else
         call FSILS_BC_CREATE(lhs, 1, 0, d, BC_TYPE_Dir, dirN(1:0))
      end if
      incL = 1
      res = 0d0

      allocate(V(d*d, nnzl), Ri(d, nNo))
      nrep = max(hdr(7), 1)
      secs = 0d0
      do rep = 1, nrep
         call FSILS_LS_CREATE(ls, hdr(4))
         if (hdr(8) > 0) ls%RI%mItr = hdr(8)
         if (hdr(9) > 0) ls%RI%sD = hdr(9)
         if (hdr(10) > 0) ls%GM%mItr = hdr(10)
         if (hdr(11) > 0) ls%GM%sD = hdr(11)
         if (hdr(12) > 0) ls%CG%mItr = hdr(12)
         if (par(1) > 0d0) ls%RI%relTol = par(1)
         if (par(2) > 0d0) ls%RI%absTol = par(2)
         if (par(3) > 0d0) ls%GM%relTol = par(3)
         if (par(4) > 0d0) ls%GM%absTol = par(4)
         if (par(5) > 0d0) ls%CG%relTol = par(5)
         if (par(6) > 0d0) ls%CG%absTol = par(6)
         V = V0
         Ri = R0
         t0 = wall()
         call FSILS_SOLVE(lhs, ls, d, Ri, V, hdr(5), incL, res)
         secs = secs + (wall() - t0)
         if (rep < nrep) call FSILS_LS_FREE(ls)
      end do
      secs = secs/nrep

      stats = 0
      stats(1) = merge(1, 0, ls%RI%suc)
      stats(2) = ls%RI%itr
      stats(3) = ls%RI%cM
      stats(4) = ls%RI%cN
      stats(5) = ls%RI%cD
      stats(6) = merge(1, 0, ls%GM%suc)
      stats(7) = ls%GM%itr
      stats(8) = ls%GM%cM
      stats(9) = ls%GM%cN
      stats(10) = ls%GM%cD
      stats(11) = merge(1, 0, ls%CG%suc)
      stats(12) = ls%CG%itr
      stats(13) = ls%CG%cM
      stats(14) = ls%CG%cN
      stats(15) = ls%CG%cD
      open(newunit=u, file=fout, access="stream", form="unformatted", status="replace")
      write(u) Ri
      write(u) stats
      write(u) ls%RI%iNorm, ls%RI%fNorm, ls%RI%dB, ls%RI%callD, secs
      close(u)
   end subroutine run_solve

end program svfsi_ref_harness
