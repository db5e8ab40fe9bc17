!
! ref_harness.f90 -- C-callable harness around the reference DEM3D modules.
!
! TEST INFRASTRUCTURE ONLY.  oracle/build_ref.sh compiles this file together
! with the reference's own modules (taken unmodified from src/ of the reference
! checkout, except for the two capacity constants of the "big" variant) into
! oracle/_ref/libdem3d_ref*.so.  Tests and the golden-fixture script load it to
! pin the C restatement (dem3d_oracle.c) to the reference, and bench.py times
! it as the CPU baseline.  Nothing in the product links it.
!
! The harness sets the reference's module variables directly (the state that
! src/DEM.f90:67-83 would build from a specification file), then calls the
! reference routines:
!   InitializeAuxiliaries          src/MDEM.f90:110-162
!   ConfigureWalls                 src/MConfig.f90:636-688
!   InitializeCellDistribution     src/MForces.f90:110-117
!   InitializeIntegrationVariables src/MDEM.f90:93-106
!   InitializeContacts             src/MContacts.f90:43-49
!   ComputeAccelerations           src/MForces.f90:145-335
!   MoveParticles / MoveParticlesEuler src/MDEM.f90:258-369
!   KineticEnergy                  src/MDEM.f90:199-213
! and moves private state (cell lists, ABM ring, contact lists) through the
! reference's own binary save/read routines (src/MForces.f90:121-141,
! src/MDEM.f90:52-89, src/MContacts.f90:75-169).
!
module ref_harness

use iso_c_binding
use MTypes
use MDimension
use MConfig
use MCells
use MContacts
use MFluid
use MForces
use MDEM
use MUtil
use BinFiles

implicit none

! set by ref_prepare_fluid: ref_init keeps FluidDensity and FluidFlow
logical, private :: keep_fluid = .false.

contains

  subroutine c2f_string( cstr, n, fstr )
    character(kind=c_char), intent(in) :: cstr(*)
    integer(c_int32_t), intent(in) :: n
    character(len=*), intent(out) :: fstr
    integer :: k
    fstr = ' '
    do k = 1, n
      fstr(k:k) = cstr(k)
    end do
  end subroutine c2f_string

  ! Constants of the system in reduced units (what ReadSpecifications,
  ! CalculateProperties and ScaleVariables leave in the modules).
  subroutine ref_init( nc_, nw_, pbc_, lbox_, gravity_, dt_, order_, neuler_, &
                       radius, mass, inertia, density,                      &
                       kn_, yn_, kt_, yt_, mu_,                              &
                       knw_, ynw_, ktw_, ytw_, muw_ ) bind(C, name="ref_init")
    integer(c_int32_t), value :: nc_, nw_, order_, neuler_
    integer(c_int32_t), intent(in) :: pbc_(3)
    real(c_double), intent(in) :: lbox_(3), gravity_(3)
    real(c_double), value :: dt_
    real(c_double), intent(in) :: radius(nc_), mass(nc_), inertia(nc_), density(nc_)
    real(c_double), intent(in) :: kn_(nc_,nc_), yn_(nc_,nc_), kt_(nc_,nc_), &
                                  yt_(nc_,nc_), mu_(nc_,nc_)
    real(c_double), intent(in) :: knw_(nc_,max(nw_,1)), ynw_(nc_,max(nw_,1)), &
                                  ktw_(nc_,max(nw_,1)), ytw_(nc_,max(nw_,1)), &
                                  muw_(nc_,max(nw_,1))
    integer :: i

    NC = nc_
    NW = nw_
    allocate( Component(NC), Wall(NW) )
    do i = 1, NC
      Component(i)%Number = 0
      Component(i)%Radius = radius(i)
      Component(i)%Density = density(i)
      Component(i)%Mass = mass(i)
      Component(i)%Inertia = inertia(i)
      Component(i)%Stiffness = 0.0_rb
      Component(i)%ShearStiffness = 0.0_rb
    end do
    Lbox = lbox_
    InvLbox = 1.0_rb/Lbox
    PBC = pbc_ /= 0
    Gravity = gravity_
    if (.not.keep_fluid) then
      FluidDensity = 0.0_rb
      FluidFlow = .false.
    end if
    allocate( Kn(NC,NC), Yn(NC,NC), Kt(NC,NC), Yt(NC,NC), Mu(NC,NC) )
    Kn = kn_ ; Yn = yn_ ; Kt = kt_ ; Yt = yt_ ; Mu = mu_
    allocate( Knw(NC,NW), Ynw(NC,NW), Ktw(NC,NW), Ytw(NC,NW), Muw(NC,NW) )
    if (NW > 0) then
      Knw = knw_(:,1:NW) ; Ynw = ynw_(:,1:NW) ; Ktw = ktw_(:,1:NW)
      Ytw = ytw_(:,1:NW) ; Muw = muw_(:,1:NW)
    end if
    TimeStep = dt_
    MethodOrder = order_
    NumberOfEulerIntervals = neuler_
    WallsInMovement = .false.
    call InitializeAuxiliaries
  end subroutine ref_init

  ! Fluid flow before ref_init, in the order of the driver: FluidDensity
  ! (reduced), ReadFluidSpecifications on the text of the specification's
  ! fluid block (SI), ScaleFluidProperties with the given scales
  ! (src/DEM.f90:596,737,858-864).  ref_init then runs InitializeForce,
  ! which calls InitializeFluid.
  subroutine ref_prepare_fluid( rho, path, n, L, T, M ) bind(C, name="ref_prepare_fluid")
    real(c_double), value :: rho, L, T, M
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int32_t), value :: n
    character(len=1024) :: f
    integer, parameter :: u = 74
    call c2f_string( path, n, f )
    FluidDensity = rho
    LengthScale = L ; TimeScale = T ; MassScale = M
    open( unit=u, file=trim(f), status='old' )
    call ReadFluidSpecifications( u, 'Error reading the fluid block:' )
    close( u )
    if (FluidFlow) call ScaleFluidProperties
    keep_fluid = .true.
  end subroutine ref_prepare_fluid

  ! One wall (TWall, lengths already halved); ConfigureWalls after the last.
  subroutine ref_set_wall( j, shape, l1, l2, l3, u, w, n, pos, vel, omg ) &
      bind(C, name="ref_set_wall")
    integer(c_int32_t), value :: j, shape
    real(c_double), value :: l1, l2, l3
    real(c_double), intent(in) :: u(3), w(3), n(3), pos(3), vel(3), omg(3)
    Wall(j)%Shape = shape
    Wall(j)%L1 = l1 ; Wall(j)%L2 = l2 ; Wall(j)%L3 = l3
    Wall(j)%Stiffness = 1.0_rb ; Wall(j)%ShearStiffness = 1.0_rb
    Wall(j)%U = u ; Wall(j)%W = w ; Wall(j)%N = n
    Wall(j)%Position = pos ; Wall(j)%Velocity = vel ; Wall(j)%Omega = omg
  end subroutine ref_set_wall

  subroutine ref_configure_walls() bind(C, name="ref_configure_walls")
    call ConfigureWalls
  end subroutine ref_configure_walls

  ! WallsInMovement = .true. and WM(1:NW) in reduced units (the values
  ! ReadSpecifications and ScaleVariables leave, src/DEM.f90:719-732,841-846).
  subroutine ref_set_wall_motion( nw_, cf, amp, vamp, tpf, phase ) &
      bind(C, name="ref_set_wall_motion")
    integer(c_int32_t), value :: nw_
    real(c_double), intent(in) :: cf(3,nw_), amp(3,nw_), vamp(3,nw_), tpf(3,nw_), phase(3,nw_)
    integer :: j
    if (allocated(WM)) deallocate(WM)
    allocate( WM(nw_) )
    do j = 1, nw_
      WM(j)%ConstantFactor = cf(:,j)
      WM(j)%Amplitude = amp(:,j)
      WM(j)%VelocityAmplitude = vamp(:,j)
      WM(j)%TwoPiFrequency = tpf(:,j)
      WM(j)%PhaseAngle = phase(:,j)
    end do
    WallsInMovement = .true.
  end subroutine ref_set_wall_motion

  subroutine ref_get_walls( nw_, pos, vel ) bind(C, name="ref_get_walls")
    integer(c_int32_t), value :: nw_
    real(c_double), intent(out) :: pos(3,nw_), vel(3,nw_)
    integer :: j
    do j = 1, nw_
      pos(:,j) = Wall(j)%Position
      vel(:,j) = Wall(j)%Velocity
    end do
  end subroutine ref_get_walls

  function ref_get_time() result( t ) bind(C, name="ref_get_time")
    real(c_double) :: t
    t = Time
  end function ref_get_time

  ! What ReadSpecifications and ScaleVariables leave in MConfig for
  ! SaveConfiguration to write: the scales, Time, the component and wall
  ! stiffnesses, and Component%Number counted from the particles.
  subroutine ref_set_meta( L, T, M, time_, stiff, shear, wstiff, wshear ) &
      bind(C, name="ref_set_meta")
    real(c_double), value :: L, T, M, time_
    real(c_double), intent(in) :: stiff(NC), shear(NC), wstiff(max(NW,1)), wshear(max(NW,1))
    integer :: i
    LengthScale = L ; TimeScale = T ; MassScale = M ; Time = time_
    do i = 1, NC
      Component(i)%Stiffness = stiff(i)
      Component(i)%ShearStiffness = shear(i)
      Component(i)%Number = count( Particle(1:NP)%Index == i )
    end do
    do i = 1, NW
      Wall(i)%Stiffness = wstiff(i)
      Wall(i)%ShearStiffness = wshear(i)
    end do
  end subroutine ref_set_meta

  ! SaveConfiguration( File, Form ) (src/MConfig.f90:87-105).
  subroutine ref_save_configuration( path, n, form ) bind(C, name="ref_save_configuration")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int32_t), value :: n, form
    character(len=1024) :: f
    call c2f_string( path, n, f )
    call SaveConfiguration( trim(f), int(form,1) )
  end subroutine ref_save_configuration

  ! ReadConfiguration( File ) (src/MConfig.f90:250-262), then the module
  ! variables it set through the getters below.
  subroutine ref_read_configuration( path, n ) bind(C, name="ref_read_configuration")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int32_t), value :: n
    character(len=1024) :: f
    call c2f_string( path, n, f )
    call ReadConfiguration( trim(f) )
  end subroutine ref_read_configuration

  subroutine ref_get_config_scalars( s, counts ) bind(C, name="ref_get_config_scalars")
    real(c_double), intent(out) :: s(7)
    integer(c_int32_t), intent(out) :: counts(3)
    s(1) = LengthScale ; s(2) = TimeScale ; s(3) = MassScale ; s(4) = Time
    s(5:7) = Lbox
    counts = [NC, NW, NP]
  end subroutine ref_get_config_scalars

  subroutine ref_get_components( nc_, number, props ) bind(C, name="ref_get_components")
    integer(c_int32_t), value :: nc_
    integer(c_int32_t), intent(out) :: number(nc_)
    real(c_double), intent(out) :: props(6,nc_)
    integer :: i
    do i = 1, nc_
      number(i) = Component(i)%Number
      props(:,i) = [Component(i)%Radius, Component(i)%Density, Component(i)%Stiffness, &
                    Component(i)%ShearStiffness, Component(i)%Mass, Component(i)%Inertia]
    end do
  end subroutine ref_get_components

  ! Wall(j): shape, then L1 L2 L3 Stiffness ShearStiffness U W N Position
  ! Velocity Omega.
  subroutine ref_get_wall( j, shape, v ) bind(C, name="ref_get_wall")
    integer(c_int32_t), value :: j
    integer(c_int32_t), intent(out) :: shape
    real(c_double), intent(out) :: v(23)
    shape = Wall(j)%Shape
    v(1:5) = [Wall(j)%L1, Wall(j)%L2, Wall(j)%L3, Wall(j)%Stiffness, Wall(j)%ShearStiffness]
    v(6:8) = Wall(j)%U ; v(9:11) = Wall(j)%W ; v(12:14) = Wall(j)%N
    v(15:17) = Wall(j)%Position ; v(18:20) = Wall(j)%Velocity ; v(21:23) = Wall(j)%Omega
  end subroutine ref_get_wall

  ! ranf (src/MUtil.f90:16-44) n times from the generator state (aam, Dummy,
  ! iix, iiy, iik), which is returned advanced.
  subroutine ref_ranf_seq( n, aam_, st, out ) bind(C, name="ref_ranf_seq")
    integer(c_int32_t), value :: n
    real(c_double), intent(inout) :: aam_
    integer(c_int32_t), intent(inout) :: st(4)
    real(c_double), intent(out) :: out(max(n,1))
    integer :: k
    aam = aam_ ; Dummy = st(1) ; iix = st(2) ; iiy = st(3) ; iik = st(4)
    do k = 1, n
      out(k) = ranf( Dummy )
    end do
    aam_ = aam ; st = [Dummy, iix, iiy, iik]
  end subroutine ref_ranf_seq

  ! PourParticle (src/MDEM.f90:412-536) with the generator state given and
  ! returned, as PouringAttempt calls it (src/DEM.f90:115-123).
  subroutine ref_pour( target, shape, psize, maxatt, aam_, st ) bind(C, name="ref_pour")
    integer(c_int32_t), intent(in) :: target(NC)
    integer(c_int32_t), value :: shape, maxatt
    real(c_double), intent(in) :: psize(2)
    real(c_double), intent(inout) :: aam_
    integer(c_int32_t), intent(inout) :: st(4)
    integer(ib) :: tn(NC)
    tn = target
    aam = aam_ ; Dummy = st(1) ; iix = st(2) ; iiy = st(3) ; iik = st(4)
    if (shape == 1) then
      call PourParticle( tn, int(shape,1), psize(1:2), int(maxatt,ib) )
    else
      call PourParticle( tn, int(shape,1), psize(1:1), int(maxatt,ib) )
    end if
    aam_ = aam ; st = [Dummy, iix, iiy, iik]
  end subroutine ref_pour

  subroutine ref_update_particle_list() bind(C, name="ref_update_particle_list")
    call UpdateParticleList
  end subroutine ref_update_particle_list

  function ref_get_np() result( n ) bind(C, name="ref_get_np")
    integer(c_int32_t) :: n
    n = NP
  end function ref_get_np

  ! SaveState (src/DEM.f90:947-1014), whose run records are variables of the
  ! main program: here they come in as arguments; the configuration, cell
  ! lists, ring and contact lists are the reference's own routines.
  subroutine ref_save_state_file( path, n, fin, start_time, nint, inter, ret, finish_time, &
                                  hours, minutes, seconds, milisecs, aam_, dummy_, iix_, &
                                  iiy_, iik_, conf, next_step ) &
      bind(C, name="ref_save_state_file")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int32_t), value :: n, fin, nint, hours, minutes, seconds, milisecs
    integer(c_int32_t), value :: dummy_, iix_, iiy_, iik_, conf, next_step
    real(c_double), value :: aam_
    character(kind=c_char), intent(in) :: start_time(28), finish_time(28)
    character(kind=c_char), intent(in) :: inter(28,max(nint,1)), ret(28,max(nint,1))
    character(len=1024) :: f
    character(28) :: s
    logical(lb) :: finished_
    integer :: u, i
    call c2f_string( path, n, f )
    u = BinOpen( trim(f), 'write' )
    finished_ = fin /= 0
    call BinWrite( u, finished_ )
    s = transfer( start_time, s ) ; call BinWrite( u, s )
    call BinWrite( u, int(nint,ib) )
    do i = 1, nint
      s = transfer( inter(:,i), s ) ; call BinWrite( u, s )
      s = transfer( ret(:,i), s ) ; call BinWrite( u, s )
    end do
    s = transfer( finish_time, s ) ; call BinWrite( u, s )
    call BinWrite( u, int(hours,4) )
    call BinWrite( u, int(minutes,1) )
    call BinWrite( u, int(seconds,1) )
    call BinWrite( u, int(milisecs,2) )
    call SaveConfiguration( u, Complete )
    call SaveCellDistribution( u )
    call SaveIntegrationVariables( u )
    call SaveContactList( u )
    call BinWrite( u, real(aam_,8) )
    call BinWrite( u, int(dummy_,4) )
    call BinWrite( u, int(iix_,4) )
    call BinWrite( u, int(iiy_,4) )
    call BinWrite( u, int(iik_,4) )
    call BinWrite( u, int(conf,ib) )
    call BinWrite( u, int(next_step,ib) )
    call BinWrite( u, Ncol )
    call BinClose( u )
  end subroutine ref_save_state_file

  ! ReadState (src/DEM.f90:1016-1110) after ref_init: the run records are
  ! returned (ints: finished, n_interruptions, hours, minutes, seconds,
  ! milisecs, dummy, iix, iiy, iik, conf, next_step; aam), the rest goes into
  ! the modules through the reference's own routines.
  subroutine ref_read_state_file( path, n, ints, aam_, start_time, finish_time ) &
      bind(C, name="ref_read_state_file")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int32_t), value :: n
    integer(c_int32_t), intent(out) :: ints(12)
    real(c_double), intent(out) :: aam_
    character(kind=c_char), intent(out) :: start_time(28), finish_time(28)
    character(len=1024) :: f
    character(28) :: s
    logical(lb) :: finished_
    integer(ib) :: nint_, i4
    integer(1) :: i1a, i1b
    integer(2) :: i2
    integer(4) :: d4(4)
    real(8) :: a8
    integer :: u, i
    call c2f_string( path, n, f )
    u = BinOpen( trim(f), 'read' )
    call BinRead( u, finished_ )
    call BinRead( u, s ) ; start_time = transfer( s, start_time )
    call BinRead( u, nint_ )
    do i = 1, nint_
      call BinRead( u, s )
      call BinRead( u, s )
    end do
    call BinRead( u, s ) ; finish_time = transfer( s, finish_time )
    call BinRead( u, i4 ) ; call BinRead( u, i1a ) ; call BinRead( u, i1b ) ; call BinRead( u, i2 )
    call ReadConfiguration( u )
    MaxNP = NP
    call ConfigureWalls
    call ReadCellDistribution( u )
    call ReadIntegrationVariables( u )
    call ReadContactList( u )
    call BinRead( u, a8 )
    do i = 1, 4
      call BinRead( u, d4(i) )
    end do
    call BinRead( u, ints(11) )
    call BinRead( u, ints(12) )
    call BinRead( u, Ncol )
    call BinClose( u )
    ints(1) = merge(1, 0, finished_)
    ints(2) = nint_
    ints(3) = i4 ; ints(4) = i1a ; ints(5) = i1b ; ints(6) = i2
    ints(7:10) = d4
    aam_ = a8
  end subroutine ref_read_state_file

  ! Tables derived by InitializeForce (src/MForces.f90:54-87).
  subroutine ref_get_tables( mindij_, mindijsq_, invmass_, invinertia_, effmass_, &
                             musq_, muwsq_ ) bind(C, name="ref_get_tables")
    real(c_double), intent(out) :: mindij_(NC,NC), mindijsq_(NC,NC), invmass_(NC), &
                                   invinertia_(NC), effmass_(NC), musq_(NC,NC),    &
                                   muwsq_(NC,max(NW,1))
    mindij_ = MinDij ; mindijsq_ = MinDijSq
    invmass_ = InvMass ; invinertia_ = InvInertia ; effmass_ = EffectiveMass
    musq_ = MuSq
    if (NW > 0) muwsq_(:,1:NW) = MuwSq
  end subroutine ref_get_tables

  subroutine ref_get_grid( ncell ) bind(C, name="ref_get_grid")
    integer(c_int32_t), intent(out) :: ncell(3)
    ncell = (/ MCellx, MCelly, Mcellz /)
  end subroutine ref_get_grid

  ! Particle(1:np) from SoA arrays X(np,3).
  subroutine ref_set_particles( np_, comp, x, v, w, a, alpha ) &
      bind(C, name="ref_set_particles")
    integer(c_int32_t), value :: np_
    integer(c_int32_t), intent(in) :: comp(np_)
    real(c_double), intent(in) :: x(np_,3), v(np_,3), w(np_,3), a(np_,3), alpha(np_,3)
    integer :: i
    if (.not.allocated(Particle)) then
      NP = np_
      allocate( Particle(NP) )
    end if
    do i = 1, NP
      Particle(i)%Index = comp(i)
      Particle(i)%Position = x(i,:)
      Particle(i)%Velocity = v(i,:)
      Particle(i)%Omega = w(i,:)
      Particle(i)%Acceleration = a(i,:)
      Particle(i)%Alpha = alpha(i,:)
    end do
  end subroutine ref_set_particles

  subroutine ref_get_particles( np_, comp, x, v, w, a, alpha ) &
      bind(C, name="ref_get_particles")
    integer(c_int32_t), value :: np_
    integer(c_int32_t), intent(out) :: comp(np_)
    real(c_double), intent(out) :: x(np_,3), v(np_,3), w(np_,3), a(np_,3), alpha(np_,3)
    integer :: i
    do i = 1, np_
      comp(i) = Particle(i)%Index
      x(i,:) = Particle(i)%Position
      v(i,:) = Particle(i)%Velocity
      w(i,:) = Particle(i)%Omega
      a(i,:) = Particle(i)%Acceleration
      alpha(i,:) = Particle(i)%Alpha
    end do
  end subroutine ref_get_particles

  ! New simulation (InitializeState, src/DEM.f90:910-922).
  subroutine ref_fresh_state() bind(C, name="ref_fresh_state")
    MaxNP = NP
    call InitializeCellDistribution
    call InitializeIntegrationVariables
    call InitializeContacts( NP )
  end subroutine ref_fresh_state

  ! Resume (ReadState, src/DEM.f90:1087-1093): cell lists, ABM ring and
  ! contact lists from one file in the reference's binary layout.
  subroutine ref_load_state( path, n ) bind(C, name="ref_load_state")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int32_t), value :: n
    character(len=1024) :: f
    integer :: u
    call c2f_string( path, n, f )
    MaxNP = NP
    u = BinOpen( trim(f), 'read' )
    call ReadCellDistribution( u )
    call ReadIntegrationVariables( u )
    call ReadContactList( u )
    call BinClose( u )
  end subroutine ref_load_state

  ! SaveState's private parts (src/DEM.f90:987-993), same order as load.
  subroutine ref_save_state( path, n ) bind(C, name="ref_save_state")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int32_t), value :: n
    character(len=1024) :: f
    integer :: u
    call c2f_string( path, n, f )
    u = BinOpen( trim(f), 'write' )
    call SaveCellDistribution( u )
    call SaveIntegrationVariables( u )
    call SaveContactList( u )
    call BinClose( u )
  end subroutine ref_save_state

  subroutine ref_get_cell_index( np_, ci ) bind(C, name="ref_get_cell_index")
    integer(c_int32_t), value :: np_
    integer(c_int32_t), intent(out) :: ci(np_)
    ci = CellIndex(1:np_)
  end subroutine ref_get_cell_index

  ! ComputeAccelerations on a caller-provided particle set (the drop-in
  ! signature of src/MForces.f90:145); accelerations are returned in a/alpha.
  subroutine ref_compute_accelerations_soa( np_, comp, x, v, w, a, alpha, dt_, upd ) &
      bind(C, name="ref_compute_accelerations_soa")
    integer(c_int32_t), value :: np_, upd
    integer(c_int32_t), intent(in) :: comp(np_)
    real(c_double), intent(in) :: x(np_,3), v(np_,3), w(np_,3)
    real(c_double), intent(out) :: a(np_,3), alpha(np_,3)
    real(c_double), value :: dt_
    type(TParticle), allocatable :: P(:)
    integer :: i
    allocate( P(np_) )
    do i = 1, np_
      P(i)%Index = comp(i)
      P(i)%Position = x(i,:)
      P(i)%Velocity = v(i,:)
      P(i)%Omega = w(i,:)
      P(i)%Acceleration = 0.0_rb
      P(i)%Alpha = 0.0_rb
    end do
    call ComputeAccelerations( P, np_, dt_, upd /= 0 )
    do i = 1, np_
      a(i,:) = P(i)%Acceleration
      alpha(i,:) = P(i)%Alpha
    end do
    deallocate( P )
  end subroutine ref_compute_accelerations_soa

  subroutine ref_compute_accelerations( upd ) bind(C, name="ref_compute_accelerations")
    integer(c_int32_t), value :: upd
    call ComputeAccelerations( Particle, NP, TimeStep, upd /= 0 )
  end subroutine ref_compute_accelerations

  subroutine ref_move_particles_euler() bind(C, name="ref_move_particles_euler")
    call MoveParticlesEuler
  end subroutine ref_move_particles_euler

  ! Steps first..first+n-1 of the time loop (src/DEM.f90:94-99) without I/O
  ! and without UpdateParticleList.
  subroutine ref_move_particles( first, n ) bind(C, name="ref_move_particles")
    integer(c_int32_t), value :: first, n
    integer(ib) :: s
    do s = first, first + n - 1
      call MoveParticles( s )
    end do
  end subroutine ref_move_particles

  function ref_ncol() result( n ) bind(C, name="ref_ncol")
    integer(c_int32_t) :: n
    n = Ncol
  end function ref_ncol

  function ref_kinetic_energy() result( e ) bind(C, name="ref_kinetic_energy")
    real(c_double) :: e
    e = KineticEnergy()
  end function ref_kinetic_energy

  function ref_potential_energy() result( e ) bind(C, name="ref_potential_energy")
    real(c_double) :: e
    e = PotentialEnergy()
  end function ref_potential_energy

end module ref_harness
