! mforces_gpu.f90 -- a replacement for the reference's module MForces
! (src/MForces.f90) whose ComputeAccelerations runs on the GPU through the
! C-ABI of include/dem3d.h (dem_compute_accelerations_aos).  It keeps the
! public names, types and signatures of src/MForces.f90:11-48 and of its
! procedures (src/MForces.f90:54-150), so the reference's own MDEM.f90 and
! DEM.f90 compile and run against it unchanged.
!
! Division of the work:
!  * the driver integrates on the host as before (MoveParticles,
!    MoveParticlesEuler, ApplyPeriodicBoundaryConditions, MoveWalls);
!  * every ComputeAccelerations call hands Particle(1:NP) -- 128-byte
!    TParticle records, the layout of dem_particle_t -- and the walls as they
!    stand to the library, which keeps the contact history and the
!    reference's cell lists resident between calls and returns Acceleration
!    and Alpha; Ncol follows the library's count;
!  * the host's cell lists (MCells) and contact lists (MContacts) go to the
!    library at the first call after InitializeCellDistribution or
!    ReadCellDistribution (a new run or a resumed state file), and come back
!    from it, with CellIndex, when the driver saves a state
!    (SaveCellDistribution, which SaveState calls before SaveContactList).
!
! Limits: pouring and particle removal (PourParticle, UpdateParticleList)
! edit the host's lists between calls, which the library does not see, so
! a change of NP after the first call stops the run with a message (runs
! that pour or lose particles use the resident-state entry points dem_step,
! dem_pour and dem_remove_escaped); fluid flow stops the run as well.  A
! resumed state passes the stored pairs of particles in the same or
! neighbouring cells (the pairs a cell loop can visit); stale pairs of cells
! further apart are not passed.
module MForces

use MTypes
use MDimension
use MCells
use MContacts
use MConfig
use MFluid
use, intrinsic :: iso_c_binding
use dem3d_gpu

implicit none

real(rb) :: Gravity(dim)

real(rb), allocatable, target :: En(:,:), Enw(:,:), &
                                 Et(:,:), Etw(:,:), &
                                 Kn(:,:), Knw(:,:), &
                                 Kt(:,:), Ktw(:,:), &
                                 Yn(:,:), Ynw(:,:), &
                                 Yt(:,:), Ytw(:,:), &
                                 Mu(:,:), Muw(:,:)

real(rb), allocatable, target :: MinDij(:,:), MinDijSq(:,:),  &
                                 InvMass(:), InvInertia(:),   &
                                 EffectiveMass(:),            &
                                 MuSq(:,:), MuwSq(:,:)

integer, parameter, private :: MaxNeighbours = 50
integer(ib), parameter, private :: MaxNCell = 30000_ib

logical :: PBC(dim)
integer(ib), allocatable :: CellIndex(:)
integer(ib) :: MCellx, MCelly, Mcellz, NCell, &
               Neighbour(MaxNeighbours)

integer(ib) :: Ncol = 0_ib

! The library's system, its particle count (-1: the host's lists are to be
! sent at the next call) and the part of Ncol counted before the upload.
type(c_ptr), private :: Gpu = c_null_ptr
integer(ib), private :: GpuNP = -1_ib
integer(ib), private :: NcolBase = 0_ib
real(rb), allocatable, target, private :: GRadius(:), GMass(:), GInertia(:)
type(dem_wall_t), allocatable, private :: GWall(:)

private :: CreateSystem, PushWalls, SendHostLists, FetchHostLists, Check

contains

!===============================================================================

  ! The tables of src/MForces.f90:54-87, in the same arithmetic.
  subroutine InitializeForce

    integer(ib) :: i, j

    allocate( MinDij(NC,NC), MinDijSq(NC,NC) )
    do j = 1_ib, NC
      do i = 1_ib, NC
        MinDij(i,j) = Component(i)%Radius + Component(j)%Radius
        MinDijSq(i,j) = MinDij(i,j)**2
      end do
    end do

    allocate( InvMass(NC), InvInertia(NC), EffectiveMass(NC) )
    InvMass = 1.0_rb/Component%Mass
    InvInertia = 1.0_rb/Component%Inertia
    EffectiveMass = Component%Mass*(Component%Density - FluidDensity)/Component%Density

    allocate( MuSq(NC,NC), MuwSq(NC,NW) )
    MuSq  = Mu**2
    MuwSq = Muw**2

    if (FluidFlow) then
      write(*,'("MForces (GPU): fluid flow is not supported by this module")')
      stop
    end if

  end subroutine InitializeForce

!===============================================================================

  subroutine InitializeCellList

    call InitCellStructure3D( MCellx, MCelly, Mcellz, NCell, Lbox(1), Lbox(2), Lbox(3), &
                              2.0_rb*maxval(Component%Radius), MaxNCell )
    call FullNeighbourCells3D( MolecularDynamics, PBC(1), PBC(2), PBC(3) )

  end subroutine InitializeCellList

!===============================================================================

  subroutine InitializeCellDistribution

    allocate( CellIndex(NP) )
    call DistributeParticles3D( CellIndex, Particle(1:NP)%Position(1), &
                                           Particle(1:NP)%Position(2), &
                                           Particle(1:NP)%Position(3)  )
    GpuNP = -1_ib

  end subroutine InitializeCellDistribution

!===============================================================================

  subroutine ReadCellDistribution( Unit )

    integer, intent(in) :: Unit

    allocate( CellIndex(NP) )
    CellIndex(1:NP) = CurrentCell3D( Particle(1:NP)%Position(1), &
                                     Particle(1:NP)%Position(2), &
                                     Particle(1:NP)%Position(3)  )
    call ReadCellStructure( Unit )
    GpuNP = -1_ib

  end subroutine ReadCellDistribution

!===============================================================================

  ! SaveState (src/DEM.f90:947-1014) writes the cell lists here and the
  ! contact lists right after: both are brought back from the library first.
  subroutine SaveCellDistribution( Unit )

    integer, intent(in) :: Unit

    if (GpuNP == NP .and. NP > 0_ib) call FetchHostLists
    call SaveCellStructure( Unit )

  end subroutine SaveCellDistribution

!===============================================================================

  subroutine ComputeAccelerations( Particle, NP, TimeStep, Update )

    integer(ib), intent(in) :: NP
    type(TParticle), intent(inout), target :: Particle(NP)
    real(rb), intent(in) :: TimeStep
    logical, intent(in) :: Update

    type(dem_stats_t) :: St

    if (.not. c_associated(Gpu)) call CreateSystem( TimeStep )
    call PushWalls
    if (GpuNP /= NP) then
      if (GpuNP >= 0_ib) then
        write(*,'("MForces (GPU): the number of particles changed from ",I0," to ",I0, &
                  &"; pouring and removal need the resident-state entry points")') GpuNP, NP
        stop
      end if
      call SendHostLists( Particle, NP )
    end if
    if (NP == 0_ib) return
    call Check( dem_compute_accelerations_aos(Gpu, int(NP, c_int32_t), c_loc(Particle), &
                                              TimeStep, merge(1_c_int32_t, 0_c_int32_t, Update)), &
                'ComputeAccelerations' )
    call Check( dem_stats(Gpu, St), 'dem_stats' )
    Ncol = NcolBase + int(St%ncol, ib)

  end subroutine ComputeAccelerations

!===============================================================================

  subroutine Check( Rc, What )

    integer(c_int), intent(in) :: Rc
    character(*), intent(in) :: What

    if (Rc /= DEM_OK) then
      write(*,'("MForces (GPU): ",A,": ",A)') What, dem_error_message(Gpu)
      stop
    end if

  end subroutine Check

!===============================================================================

  ! The library's system from the tables of InitializeForce and the driver's
  ! scaled variables (reduced units).
  subroutine CreateSystem( TimeStep )

    real(rb), intent(in) :: TimeStep

    type(dem_params_t) :: P
    integer :: k

    allocate( GRadius(NC), GMass(NC), GInertia(NC), GWall(max(NW,1_ib)) )
    GRadius = Component%Radius
    GMass = Component%Mass
    GInertia = Component%Inertia
    P%nc = int(NC, c_int32_t)
    P%nw = int(NW, c_int32_t)
    do k = 1, dim
      P%pbc(k) = merge(1_c_int32_t, 0_c_int32_t, PBC(k))
    end do
    P%max_ncell = int(MaxNCell, c_int32_t)
    P%lbox = Lbox
    P%gravity = Gravity
    P%time_step = TimeStep
    P%radius = c_loc(GRadius)
    P%mass = c_loc(GMass)
    P%inertia = c_loc(GInertia)
    P%inv_mass = c_loc(InvMass)
    P%inv_inertia = c_loc(InvInertia)
    P%eff_mass = c_loc(EffectiveMass)
    P%min_dij = c_loc(MinDij)
    P%min_dij_sq = c_loc(MinDijSq)
    P%kn = c_loc(Kn)
    P%yn = c_loc(Yn)
    P%kt = c_loc(Kt)
    P%yt = c_loc(Yt)
    P%musq = c_loc(MuSq)
    if (NW > 0_ib) then
      P%knw = c_loc(Knw)
      P%ynw = c_loc(Ynw)
      P%ktw = c_loc(Ktw)
      P%ytw = c_loc(Ytw)
      P%muwsq = c_loc(MuwSq)
    end if
    call Check( dem_create(P, Gpu), 'dem_create' )

  end subroutine CreateSystem

!===============================================================================

  ! Wall(1:NW) as they stand (MoveWalls moves them between calls).
  subroutine PushWalls

    integer(ib) :: j

    if (NW == 0_ib) return
    do j = 1_ib, NW
      GWall(j)%shape = int(Wall(j)%Shape, c_int32_t)
      GWall(j)%order = int(Wall(j)%Order, c_int32_t)
      GWall(j)%l1 = Wall(j)%L1
      GWall(j)%l2 = Wall(j)%L2
      GWall(j)%l3 = Wall(j)%L3
      GWall(j)%stiffness = Wall(j)%Stiffness
      GWall(j)%shear_stiffness = Wall(j)%ShearStiffness
      GWall(j)%u = Wall(j)%U
      GWall(j)%w = Wall(j)%W
      GWall(j)%n = Wall(j)%N
      GWall(j)%position = Wall(j)%Position
      GWall(j)%velocity = Wall(j)%Velocity
      GWall(j)%omega = Wall(j)%Omega
    end do
    call Check( dem_set_walls(Gpu, int(NW, c_int32_t), GWall), 'dem_set_walls' )

  end subroutine PushWalls

!===============================================================================

  ! The particles, the order of the host's cell lists (head first) and its
  ! stored pairs and wall contacts, to the library.
  subroutine SendHostLists( Particle, NP )

    integer(ib), intent(in) :: NP
    type(TParticle), intent(in) :: Particle(NP)

    integer(c_int32_t), allocatable, target :: Comp(:), Rank(:)
    real(c_double), allocatable, target :: X(:,:), V(:,:), W(:,:)
    integer(c_int64_t), allocatable, target :: Keys(:), WKeys(:)
    real(c_double), allocatable, target :: FT(:,:), VT(:,:), WFT(:,:), WVT(:,:)
    integer(ib), allocatable :: Buf(:)
    integer(ib) :: iCell, ii, jj, i, j, NPCC, NPSC, Pass, NPair, NWall
    logical :: Contacting
    real(rb) :: V1(dim), V2(dim)
    type(dem_stats_t) :: St

    allocate( Comp(max(NP,1_ib)), Rank(max(NP,1_ib)), X(max(NP,1_ib),dim), &
              V(max(NP,1_ib),dim), W(max(NP,1_ib),dim), Buf(max(NP,1_ib)) )
    Rank = 0
    do i = 1_ib, NP
      Comp(i) = int(Particle(i)%Index, c_int32_t)
      X(i,:) = Particle(i)%Position
      V(i,:) = Particle(i)%Velocity
      W(i,:) = Particle(i)%Omega
    end do
    ! two passes over the host's lists: count, then fill
    NPair = 0_ib
    NWall = 0_ib
    do Pass = 1, 2
      if (Pass == 2) then
        allocate( Keys(max(NPair,1_ib)), FT(dim,max(NPair,1_ib)), VT(dim,max(NPair,1_ib)), &
                  WKeys(max(NWall,1_ib)), WFT(dim,max(NWall,1_ib)), WVT(dim,max(NWall,1_ib)) )
        NPair = 0_ib
        NWall = 0_ib
      end if
      do iCell = 1_ib, NCell
        call NeighbourParticles( iCell, NPCC, NPSC, Buf )
        if (NPCC == 0_ib) cycle
        do ii = 1_ib, NPCC
          Rank(Buf(ii)) = int(NPCC - ii, c_int32_t)
          do jj = ii + 1_ib, NPSC
            call VerifyPairContact( Buf(ii), Buf(jj), Contacting, V1, V2 )
            if (.not. Contacting) cycle
            NPair = NPair + 1_ib
            if (Pass == 2) then
              i = min(Buf(ii), Buf(jj))
              j = max(Buf(ii), Buf(jj))
              Keys(NPair) = ior(ishft(int(i, c_int64_t), 32), int(j, c_int64_t))
              FT(:,NPair) = V1
              VT(:,NPair) = V2
            end if
          end do
        end do
      end do
      do i = 1_ib, NP
        do j = 1_ib, NW
          call VerifyWallContact( i, j, Contacting, V1, V2 )
          if (.not. Contacting) cycle
          NWall = NWall + 1_ib
          if (Pass == 2) then
            WKeys(NWall) = ior(ishft(int(i, c_int64_t), 32), int(j, c_int64_t))
            WFT(:,NWall) = V1
            WVT(:,NWall) = V2
          end if
        end do
      end do
    end do
    call Check( dem_upload_state(Gpu, int(NP, c_int32_t), c_loc(Comp), c_loc(X), c_loc(V), &
                                 c_loc(W), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                                 c_loc(Rank)), 'dem_upload_state' )
    if (NP > 0_ib) then
      call Check( dem_upload_contacts(Gpu, int(NPair, c_int64_t), c_loc(Keys), c_loc(FT), &
                                      c_loc(VT), int(NWall, c_int64_t), c_loc(WKeys), &
                                      c_loc(WFT), c_loc(WVT)), 'dem_upload_contacts' )
    end if
    call Check( dem_stats(Gpu, St), 'dem_stats' )
    NcolBase = Ncol - int(St%ncol, ib)
    GpuNP = NP

  end subroutine SendHostLists

!===============================================================================

  ! The library's cell lists (CellIndex and the order within each cell) and
  ! its stored pairs and wall contacts, into the host's MCells and MContacts
  ! lists.
  subroutine FetchHostLists

    integer(c_int32_t), allocatable :: Cells(:), Rank(:)
    integer(c_int64_t), allocatable, target :: Keys(:), WKeys(:)
    real(c_double), allocatable, target :: FT(:,:), VT(:,:), WFT(:,:), WVT(:,:)
    integer(c_int64_t) :: NPair, NWall, k
    integer(ib) :: i, j, r
    logical :: Contacting
    real(rb) :: V1(dim), V2(dim)

    allocate( Cells(NP), Rank(NP) )
    call Check( dem_download_cells(Gpu, int(NP, c_int32_t), Cells), 'dem_download_cells' )
    call Check( dem_download_cell_rank(Gpu, int(NP, c_int32_t), Rank), 'dem_download_cell_rank' )
    do i = 1_ib, NP
      call RemoveParticleFromCell( i, CellIndex(i) )
    end do
    CellIndex(1:NP) = int(Cells, ib)
    ! ascending rank: each particle is pushed to the head of its cell's list,
    ! so the largest rank ends at the head
    do r = 0_ib, int(maxval(Rank), ib)
      do i = 1_ib, NP
        if (Rank(i) == r) call AddParticle( i, CellIndex(i) )
      end do
    end do

    call Check( dem_download_contacts(Gpu, 0_c_int64_t, NPair, c_null_ptr, c_null_ptr, &
                                      c_null_ptr, 0_c_int64_t, NWall, c_null_ptr, c_null_ptr, &
                                      c_null_ptr), 'dem_download_contacts' )
    allocate( Keys(max(NPair,1_c_int64_t)), FT(dim,max(NPair,1_c_int64_t)), &
              VT(dim,max(NPair,1_c_int64_t)), WKeys(max(NWall,1_c_int64_t)), &
              WFT(dim,max(NWall,1_c_int64_t)), WVT(dim,max(NWall,1_c_int64_t)) )
    call Check( dem_download_contacts(Gpu, NPair, NPair, c_loc(Keys), c_loc(FT), c_loc(VT), &
                                      NWall, NWall, c_loc(WKeys), c_loc(WFT), c_loc(WVT)), &
                'dem_download_contacts' )
    call DestroyContactList
    call InitializeContacts( NP )
    do k = 1_c_int64_t, NPair
      i = int(ishft(Keys(k), -32), ib)
      j = int(iand(Keys(k), int(z'FFFFFFFF', c_int64_t)), ib)
      call VerifyPairContact( i, j, Contacting, V1, V2 )
      call UpdatePairContact( FT(:,k), VT(:,k), Contacting )
    end do
    do k = 1_c_int64_t, NWall
      i = int(ishft(WKeys(k), -32), ib)
      j = int(iand(WKeys(k), int(z'FFFFFFFF', c_int64_t)), ib)
      call VerifyWallContact( i, j, Contacting, V1, V2 )
      call UpdateWallContact( WFT(:,k), WVT(:,k), Contacting )
    end do

  end subroutine FetchHostLists

!===============================================================================

end module MForces
