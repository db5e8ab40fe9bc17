! dem3d_gpu.f90 -- Fortran 2008 bindings (iso_c_binding) of include/dem3d.h.
!
! The reference driver (src/DEM.f90, src/MDEM.f90) uses this module to call
! the GPU library in place of its own MForces/MCells/MContacts procedures:
!
!   level 1  ComputeAccelerations(Particle, NP, TimeStep, Update)
!            (src/MForces.f90:145-150)  ->  dem_compute_accelerations_aos
!   level 2  MoveParticlesEuler / MoveParticles (src/MDEM.f90:227-369) on a
!            state kept on the GPU  ->  dem_euler_startup / dem_step
!
! Types mirror the C structs field by field; bind(c) gives them the C layout
! (the test tests/test_abi.py compares every size with the C compiler's).
module dem3d_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: DEM_OK = 0, DEM_ERR_ARG = -1, DEM_ERR_DEVICE = -2, &
       DEM_ERR_CAPACITY = -3, DEM_ERR_STATE = -4, DEM_ERR_UNSUPPORTED = -5, DEM_ERR_ESCAPE = -6, &
       DEM_ERR_IO = -7
  integer(c_int32_t), parameter, public :: DEM_CFG_COMPLETE = 72, DEM_CFG_COMPACT = 28
  integer(c_int32_t), parameter, public :: DEM_FP64 = 0, DEM_MIXED = 1
  integer(c_int32_t), parameter, public :: DEM_SHAPE_RECTANGLE = 1, DEM_SHAPE_CONE = 2
  integer(c_int32_t), parameter, public :: DEM_HIST_REFERENCE = 0, DEM_HIST_CANONICAL = 1

  ! dem_params_t: the module variables of MForces/MConfig/MDEM after
  ! InitializeForce and ScaleVariables (src/MForces.f90:11-48).
  type, bind(c), public :: dem_params_t
    integer(c_int32_t) :: nc = 0, nw = 0
    integer(c_int32_t) :: pbc(3) = 0
    integer(c_int32_t) :: method_order = 4, n_euler = 10, max_ncell = 30000
    integer(c_int32_t) :: ref_grid(3) = 0
    integer(c_int32_t) :: precision = DEM_FP64, history_mode = DEM_HIST_REFERENCE
    integer(c_int32_t) :: max_neighbours = 0, max_contacts = 0
    real(c_double) :: lbox(3) = 0, gravity(3) = 0
    real(c_double) :: time_step = 0, skin = 0
    type(c_ptr) :: radius = c_null_ptr, mass = c_null_ptr, inertia = c_null_ptr
    type(c_ptr) :: inv_mass = c_null_ptr, inv_inertia = c_null_ptr, eff_mass = c_null_ptr
    type(c_ptr) :: min_dij = c_null_ptr, min_dij_sq = c_null_ptr
    type(c_ptr) :: kn = c_null_ptr, yn = c_null_ptr, kt = c_null_ptr, yt = c_null_ptr
    type(c_ptr) :: musq = c_null_ptr
    type(c_ptr) :: knw = c_null_ptr, ynw = c_null_ptr, ktw = c_null_ptr, ytw = c_null_ptr
    type(c_ptr) :: muwsq = c_null_ptr
  end type dem_params_t

  ! TWall (src/MDimension3D.f90:45-59), half lengths.
  type, bind(c), public :: dem_wall_t
    integer(c_int32_t) :: shape = DEM_SHAPE_RECTANGLE, order = 0
    real(c_double) :: l1 = 0, l2 = 0, l3 = 0
    real(c_double) :: stiffness = 0, shear_stiffness = 0
    real(c_double) :: u(3) = 0, w(3) = 0, n(3) = 0
    real(c_double) :: position(3) = 0, velocity(3) = 0, omega(3) = 0
  end type dem_wall_t

  ! TWallMovement (src/MDEM.f90:39-45), reduced units.
  type, bind(c), public :: dem_wall_motion_t
    real(c_double) :: constant_factor(3) = 0, amplitude(3) = 0, velocity_amplitude(3) = 0
    real(c_double) :: two_pi_frequency(3) = 0, phase_angle(3) = 0
  end type dem_wall_motion_t

  ! Byte image of TParticle (src/MDimension3D.f90:28-35): 128 bytes.
  type, bind(c), public :: dem_particle_t
    integer(c_int32_t) :: index = 0, pad_ = 0
    real(c_double) :: position(3) = 0, velocity(3) = 0, acceleration(3) = 0
    real(c_double) :: omega(3) = 0, alpha(3) = 0
  end type dem_particle_t

  type, bind(c), public :: dem_stats_t
    integer(c_int64_t) :: np, pair_contacts, wall_contacts, stored_pairs, ncol
    integer(c_int64_t) :: evaluations, steps, rebuilds, neighbour_pairs, stale_dropped
    integer(c_int64_t) :: ref_ncell(3), nb_ncell(3)
    real(c_double) :: kinetic_energy, max_overlap
    real(c_double) :: ms_pass1, ms_pass2, ms_rebuild, ms_step
    real(c_double) :: mean_pair_contacts, time
    integer(c_int64_t) :: next_step
    integer(c_int64_t) :: max_halo, halo_overflow, device_bytes
  end type dem_stats_t

  ! dem_tune_t: the autotuner's report (launch shapes of the force passes).
  integer(c_int), parameter, public :: DEM_NSHAPES = 8
  type, bind(c), public :: dem_tune_t
    integer(c_int32_t) :: nshapes = 0, chosen = 0
    integer(c_int64_t) :: steps = 0
    integer(c_int32_t) :: block(DEM_NSHAPES) = 0, waves(DEM_NSHAPES) = 0
    integer(c_int32_t) :: front(DEM_NSHAPES) = 0
    real(c_double) :: ms(DEM_NSHAPES) = 0
    integer(c_int32_t) :: block_map = 0, pad_ = 0
    real(c_double) :: ms_map(2) = 0
  end type dem_tune_t

  ! dem_observables_t: energies (KineticEnergy, PotentialEnergy,
  ! src/MDEM.f90:183-213), contact counts, coordination number and bulk
  ! packing fraction of the resident state.
  integer(c_int), parameter, public :: DEM_OBS_MAX_LAYERS = 256
  type, bind(c), public :: dem_observables_t
    real(c_double) :: kinetic_energy = 0, potential_energy = 0
    integer(c_int64_t) :: pair_contacts = 0, wall_contacts = 0
    real(c_double) :: coordination = 0, packing_fraction = 0
    real(c_double) :: region(2) = 0
    real(c_double) :: solid_volume = 0, top = 0
  end type dem_observables_t

  ! dem_config_t: a configuration file (SaveConfiguration, src/MConfig.f90:109-236).
  type, bind(c), public :: dem_config_t
    integer(c_int32_t) :: form = DEM_CFG_COMPLETE, version = 2
    real(c_double) :: length_scale = 0, time_scale = 0, mass_scale = 0, time = 0
    real(c_double) :: lbox(3) = 0
    integer(c_int32_t) :: nc = 0, nw = 0, np = 0
    type(c_ptr) :: number = c_null_ptr, radius = c_null_ptr, density = c_null_ptr
    type(c_ptr) :: stiffness = c_null_ptr, shear_stiffness = c_null_ptr
    type(c_ptr) :: mass = c_null_ptr, inertia = c_null_ptr, walls = c_null_ptr
    type(c_ptr) :: comp = c_null_ptr, x = c_null_ptr, v = c_null_ptr, w = c_null_ptr
  end type dem_config_t

  ! dem_run_t: the driver's records of a state file (SaveState, src/DEM.f90:947-1014).
  type, bind(c), public :: dem_run_t
    integer(c_int32_t) :: finished = 0
    character(kind=c_char) :: start_time(28) = ' '
    integer(c_int32_t) :: n_interruptions = 0
    type(c_ptr) :: interruption = c_null_ptr, retaking = c_null_ptr
    character(kind=c_char) :: finish_time(28) = ' '
    integer(c_int32_t) :: hours = 0, minutes = 0, seconds = 0, milisecs = 0
    real(c_double) :: aam = 0
    integer(c_int32_t) :: dummy = 0, iix = 0, iiy = 0, iik = 0, conf = 0, next_step = 0, ncol = 0
  end type dem_run_t

  ! dem_state_sections_t: cell lists, ABM ring and contact lists of a state file.
  type, bind(c), public :: dem_state_sections_t
    integer(c_int32_t) :: nprev = 0, ipc(3) = 0
    type(c_ptr) :: ring = c_null_ptr
    integer(c_int32_t) :: ncell = 0
    type(c_ptr) :: cell_len = c_null_ptr, cell_members = c_null_ptr
    integer(c_int64_t) :: npairs = 0
    type(c_ptr) :: keys = c_null_ptr, ft = c_null_ptr, vt = c_null_ptr
    integer(c_int64_t) :: nwall = 0
    type(c_ptr) :: wkeys = c_null_ptr, wft = c_null_ptr, wvt = c_null_ptr
  end type dem_state_sections_t

  ! dem_fluid_t: FluidFlow of mFluid, reduced.
  type, bind(c), public :: dem_fluid_t
    real(c_double) :: fluid_density = 0, viscosity = 0, flow_rate = 0
    integer(c_int32_t) :: nl = 0, pad_ = 0
    type(c_ptr) :: density = c_null_ptr
  end type dem_fluid_t

  ! dem_rng_t: the generator state of MUtil (aam, Dummy, iix, iiy, iik).
  type, bind(c), public :: dem_rng_t
    real(c_double) :: aam = 0
    integer(c_int32_t) :: dummy = 0, iix = -1, iiy = -1, iik = 0
  end type dem_rng_t

  integer(c_int32_t), parameter, public :: DEM_POUR_RECTANGLE = 1, DEM_POUR_CIRCLE = 2
  ! dem_pour_t: the pouring specification (.prg), reduced.
  type, bind(c), public :: dem_pour_t
    type(c_ptr) :: target_number = c_null_ptr
    integer(c_int32_t) :: shape = DEM_POUR_CIRCLE, max_attempts = 10
    real(c_double) :: size(2) = 0, length_scale = 0, time_scale = 0
  end type dem_pour_t

  public :: dem_create, dem_destroy, dem_last_error, dem_set_walls
  public :: dem_ranf, dem_pour, dem_remove_escaped, dem_set_fluid, dem_download_fluid_field
  public :: dem_config_write, dem_config_read, dem_state_file_write, dem_state_file_read
  public :: dem_save_state_file, dem_load_state_file, dem_download_cell_rank
  public :: dem_set_wall_motion, dem_download_walls
  public :: dem_upload_state, dem_download_state, dem_download_predicted
  public :: dem_upload_contacts, dem_download_contacts, dem_download_cells
  public :: dem_compute_accelerations_aos, dem_compute_accelerations
  public :: dem_euler_startup, dem_step, dem_synchronize, dem_stats, dem_set_option
  public :: dem_autotune, dem_observables, dem_observables_noalloc
  public :: dem_error_message

  interface
    integer(c_int) function dem_create(params, out) bind(c, name="dem_create")
      import :: c_int, c_ptr, dem_params_t
      type(dem_params_t), intent(in) :: params
      type(c_ptr), intent(out) :: out
    end function
    subroutine dem_destroy(h) bind(c, name="dem_destroy")
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine
    type(c_ptr) function dem_last_error(h) bind(c, name="dem_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function dem_set_walls(h, nw, walls) bind(c, name="dem_set_walls")
      import :: c_int, c_int32_t, c_ptr, dem_wall_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nw
      type(dem_wall_t), intent(in) :: walls(*)
    end function
    ! WM(1:nw) of MDEM; nw = 0 stops the walls.
    integer(c_int) function dem_set_wall_motion(h, nw, motion) &
        bind(c, name="dem_set_wall_motion")
      import :: c_int, c_int32_t, c_ptr, dem_wall_motion_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nw
      type(dem_wall_motion_t), intent(in) :: motion(*)
    end function
    integer(c_int) function dem_download_walls(h, nw, walls) bind(c, name="dem_download_walls")
      import :: c_int, c_int32_t, c_ptr, dem_wall_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nw
      type(dem_wall_t), intent(inout) :: walls(*)
    end function
    ! SoA arrays are Fortran X(np,3); optional arrays are passed as c_null_ptr.
    integer(c_int) function dem_upload_state(h, np, comp, x, v, w, a, alpha, ring, ipc, &
                                             cell_rank) bind(c, name="dem_upload_state")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: np
      type(c_ptr), value :: comp, x, v, w, a, alpha, ring, ipc, cell_rank
    end function
    integer(c_int) function dem_download_state(h, np, comp, x, v, w, a, alpha, ring, ipc) &
        bind(c, name="dem_download_state")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: np
      type(c_ptr), value :: comp, x, v, w, a, alpha, ring, ipc
    end function
    integer(c_int) function dem_download_predicted(h, np, x, v, w, a, alpha) &
        bind(c, name="dem_download_predicted")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: np
      type(c_ptr), value :: x, v, w, a, alpha
    end function
    integer(c_int) function dem_upload_contacts(h, npairs, keys, ft, vt, nwall, wkeys, wft, &
                                                wvt) bind(c, name="dem_upload_contacts")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int64_t), value :: npairs, nwall
      type(c_ptr), value :: keys, ft, vt, wkeys, wft, wvt
    end function
    integer(c_int) function dem_download_contacts(h, cap, npairs, keys, ft, vt, wcap, nwall, &
                                                  wkeys, wft, wvt) &
        bind(c, name="dem_download_contacts")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int64_t), value :: cap, wcap
      integer(c_int64_t), intent(out) :: npairs, nwall
      type(c_ptr), value :: keys, ft, vt, wkeys, wft, wvt
    end function
    integer(c_int) function dem_download_cells(h, np, cell_index) &
        bind(c, name="dem_download_cells")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: np
      integer(c_int32_t), intent(out) :: cell_index(*)
    end function
    ! Particle(1:NP) of the driver, passed as c_loc(Particle(1)).
    integer(c_int) function dem_compute_accelerations_aos(h, np, p, time_step, update) &
        bind(c, name="dem_compute_accelerations_aos")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: np
      type(c_ptr), value :: p
      real(c_double), value :: time_step
      integer(c_int32_t), value :: update
    end function
    integer(c_int) function dem_compute_accelerations(h, which, update) &
        bind(c, name="dem_compute_accelerations")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: which, update
    end function
    integer(c_int) function dem_euler_startup(h) bind(c, name="dem_euler_startup")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function dem_step(h, nsteps, first_step, steps_done, escaped) &
        bind(c, name="dem_step")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int64_t), value :: nsteps, first_step
      integer(c_int64_t), intent(out) :: steps_done
      integer(c_int32_t), intent(out) :: escaped
    end function
    integer(c_int) function dem_synchronize(h) bind(c, name="dem_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function dem_stats(h, out) bind(c, name="dem_stats")
      import :: c_int, c_ptr, dem_stats_t
      type(c_ptr), value :: h
      type(dem_stats_t), intent(out) :: out
    end function
    integer(c_int) function dem_set_option(h, name, value) bind(c, name="dem_set_option")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), value :: value
    end function
    ! Times every launch shape of the force passes while stepping
    ! (steps_per_shape + 1 MoveParticles steps per shape, numbered from
    ! first_step) and keeps the fastest; report%steps = steps advanced.
    integer(c_int) function dem_autotune(h, steps_per_shape, first_step, report) &
        bind(c, name="dem_autotune")
      import :: c_int, c_int64_t, c_ptr, dem_tune_t
      type(c_ptr), value :: h
      integer(c_int64_t), value :: steps_per_shape, first_step
      type(dem_tune_t), intent(inout) :: report
    end function
    ! Observables of the resident state; profile(4, nlayers) when nlayers > 0
    ! (pass a one-element array when nlayers = 0).
    integer(c_int) function dem_observables(h, nlayers, out, profile) &
        bind(c, name="dem_observables")
      import :: c_int, c_int32_t, c_double, c_ptr, dem_observables_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nlayers
      type(dem_observables_t), intent(out) :: out
      real(c_double), intent(out) :: profile(*)
    end function
    ! The same with no device allocation (nlayers <= DEM_OBS_MAX_LAYERS).
    integer(c_int) function dem_observables_noalloc(h, nlayers, out, profile) &
        bind(c, name="dem_observables_noalloc")
      import :: c_int, c_int32_t, c_double, c_ptr, dem_observables_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nlayers
      type(dem_observables_t), intent(out) :: out
      real(c_double), intent(out) :: profile(*)
    end function
    ! FluidFlow with the specification's fluid (the C entry also takes a
    ! null pointer to turn it off).
    integer(c_int) function dem_set_fluid(h, f) bind(c, name="dem_set_fluid")
      import :: c_int, c_ptr, dem_fluid_t
      type(c_ptr), value :: h
      type(dem_fluid_t), intent(in) :: f
    end function
    integer(c_int) function dem_download_fluid_field(h, nl, vf, ef) &
        bind(c, name="dem_download_fluid_field")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nl
      real(c_double), intent(out) :: vf(0:*), ef(0:*)
    end function
    real(c_double) function dem_ranf(rng) bind(c, name="dem_ranf")
      import :: c_double, dem_rng_t
      type(dem_rng_t), intent(inout) :: rng
    end function
    integer(c_int) function dem_pour(h, pour, rng, inserted) bind(c, name="dem_pour")
      import :: c_int, c_int32_t, c_ptr, dem_pour_t, dem_rng_t
      type(c_ptr), value :: h
      type(dem_pour_t), intent(in) :: pour
      type(dem_rng_t), intent(inout) :: rng
      integer(c_int32_t), intent(out) :: inserted
    end function
    integer(c_int) function dem_remove_escaped(h, cap, removed, nremoved) &
        bind(c, name="dem_remove_escaped")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: cap
      integer(c_int32_t), intent(out) :: removed(*), nremoved
    end function
    ! File names are C strings: pass trim(name)//c_null_char.
    integer(c_int) function dem_config_write(path, cfg) bind(c, name="dem_config_write")
      import :: c_int, c_char, dem_config_t
      character(kind=c_char), intent(in) :: path(*)
      type(dem_config_t), intent(in) :: cfg
    end function
    integer(c_int) function dem_config_read(path, cfg) bind(c, name="dem_config_read")
      import :: c_int, c_char, dem_config_t
      character(kind=c_char), intent(in) :: path(*)
      type(dem_config_t), intent(inout) :: cfg
    end function
    integer(c_int) function dem_state_file_write(path, run, cfg, sec) &
        bind(c, name="dem_state_file_write")
      import :: c_int, c_char, dem_run_t, dem_config_t, dem_state_sections_t
      character(kind=c_char), intent(in) :: path(*)
      type(dem_run_t), intent(in) :: run
      type(dem_config_t), intent(in) :: cfg
      type(dem_state_sections_t), intent(in) :: sec
    end function
    integer(c_int) function dem_state_file_read(path, run, cfg, sec) &
        bind(c, name="dem_state_file_read")
      import :: c_int, c_char, dem_run_t, dem_config_t, dem_state_sections_t
      character(kind=c_char), intent(in) :: path(*)
      type(dem_run_t), intent(inout) :: run
      type(dem_config_t), intent(inout) :: cfg
      type(dem_state_sections_t), intent(inout) :: sec
    end function
    integer(c_int) function dem_save_state_file(h, path, run, meta) &
        bind(c, name="dem_save_state_file")
      import :: c_int, c_char, c_ptr, dem_run_t, dem_config_t
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
      type(dem_run_t), intent(in) :: run
      type(dem_config_t), intent(in) :: meta
    end function
    integer(c_int) function dem_load_state_file(h, path, run, meta) &
        bind(c, name="dem_load_state_file")
      import :: c_int, c_char, c_ptr, dem_run_t, dem_config_t
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
      type(dem_run_t), intent(inout) :: run
      type(dem_config_t), intent(inout) :: meta
    end function
    integer(c_int) function dem_download_cell_rank(h, np, rank) &
        bind(c, name="dem_download_cell_rank")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: np
      integer(c_int32_t), intent(out) :: rank(*)
    end function
    integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! dem_last_error as a Fortran string (for the driver's write + stop).
  function dem_error_message(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, k
    p = dem_last_error(h)
    if (.not. c_associated(p)) then
      msg = ""
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do k = 1, n
      msg(k:k) = chars(k)
    end do
  end function dem_error_message

end module dem3d_gpu
