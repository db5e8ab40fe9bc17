! drive_steps.f90 -- the time loop of the reference driver (DEM.f90: the
! MoveParticlesEuler start-up, then MoveParticles steps, src/DEM.f90:87-100)
! run on the GPU through the iso_c_binding module dem3d_gpu, as the Fortran
! driver would call it.
!
! Usage: drive_steps <input> <output>
!
! <input> (unformatted stream, native byte order) holds what the driver has
! in memory after ScaleVariables and InitializeForce:
!   int32  nc, nw, np, method_order, n_euler, max_ncell, pbc(3), ref_grid(3),
!          precision, history_mode, nsteps, pad
!   real64 lbox(3), gravity(3), time_step, skin
!   real64 radius, mass, inertia, inv_mass, inv_inertia, eff_mass     (nc each)
!   real64 min_dij, min_dij_sq, kn, yn, kt, yt, musq                  (nc*nc each)
!   real64 knw, ynw, ktw, ytw, muwsq                                  (nc*max(nw,1) each)
!   dem_wall_t walls(max(nw,1))
!   int32  comp(np); real64 x(np,3), v(np,3), w(np,3)
! <output> (unformatted stream):
!   int64  rc, steps_done, ncol, pair_contacts, wall_contacts
!   real64 x(np,3), v(np,3), w(np,3), a(np,3), alpha(np,3)
program drive_steps
  use, intrinsic :: iso_c_binding
  use dem3d_gpu
  implicit none
  integer(c_int32_t) :: hdr(16)
  integer :: nc, nw, np, nwt, u, k
  type(dem_params_t) :: p
  type(dem_wall_t), allocatable, target :: walls(:)
  real(c_double), allocatable, target :: radius(:), mass(:), inertia(:), inv_mass(:), &
      inv_inertia(:), eff_mass(:), min_dij(:), min_dij_sq(:), kn(:), yn(:), kt(:), yt(:), &
      musq(:), knw(:), ynw(:), ktw(:), ytw(:), muwsq(:)
  integer(c_int32_t), allocatable, target :: comp(:)
  real(c_double), allocatable, target :: x(:, :), v(:, :), w(:, :), a(:, :), alpha(:, :)
  real(c_double) :: r8(8)
  type(c_ptr) :: h
  integer(c_int) :: rc
  integer(c_int64_t) :: done, nsteps
  integer(c_int32_t) :: escaped
  type(dem_stats_t) :: st
  character(len=4096) :: inp, outp

  h = c_null_ptr
  call get_command_argument(1, inp)
  call get_command_argument(2, outp)
  open(newunit=u, file=trim(inp), access="stream", form="unformatted", status="old")
  read(u) hdr
  nc = hdr(1); nw = hdr(2); np = hdr(3)
  nwt = max(nw, 1)
  p%nc = nc
  p%nw = nw
  p%method_order = hdr(4)
  p%n_euler = hdr(5)
  p%max_ncell = hdr(6)
  p%pbc = hdr(7:9)
  p%ref_grid = hdr(10:12)
  p%precision = hdr(13)
  p%history_mode = hdr(14)
  nsteps = int(hdr(15), c_int64_t)
  read(u) r8
  p%lbox = r8(1:3)
  p%gravity = r8(4:6)
  p%time_step = r8(7)
  p%skin = r8(8)
  allocate(radius(nc), mass(nc), inertia(nc), inv_mass(nc), inv_inertia(nc), eff_mass(nc))
  read(u) radius, mass, inertia, inv_mass, inv_inertia, eff_mass
  allocate(min_dij(nc * nc), min_dij_sq(nc * nc), kn(nc * nc), yn(nc * nc), kt(nc * nc), &
           yt(nc * nc), musq(nc * nc))
  read(u) min_dij, min_dij_sq, kn, yn, kt, yt, musq
  allocate(knw(nc * nwt), ynw(nc * nwt), ktw(nc * nwt), ytw(nc * nwt), muwsq(nc * nwt))
  read(u) knw, ynw, ktw, ytw, muwsq
  allocate(walls(nwt))
  read(u) walls
  allocate(comp(np), x(np, 3), v(np, 3), w(np, 3), a(np, 3), alpha(np, 3))
  read(u) comp, x, v, w
  close(u)

  p%radius = c_loc(radius); p%mass = c_loc(mass); p%inertia = c_loc(inertia)
  p%inv_mass = c_loc(inv_mass); p%inv_inertia = c_loc(inv_inertia); p%eff_mass = c_loc(eff_mass)
  p%min_dij = c_loc(min_dij); p%min_dij_sq = c_loc(min_dij_sq)
  p%kn = c_loc(kn); p%yn = c_loc(yn); p%kt = c_loc(kt); p%yt = c_loc(yt); p%musq = c_loc(musq)
  p%knw = c_loc(knw); p%ynw = c_loc(ynw); p%ktw = c_loc(ktw); p%ytw = c_loc(ytw)
  p%muwsq = c_loc(muwsq)

  done = 0
  rc = dem_create(p, h)
  if (rc == DEM_OK .and. nw > 0) rc = dem_set_walls(h, int(nw, c_int32_t), walls)
  if (rc == DEM_OK) rc = dem_upload_state(h, int(np, c_int32_t), c_loc(comp), c_loc(x), &
                                          c_loc(v), c_loc(w), c_null_ptr, c_null_ptr, &
                                          c_null_ptr, c_null_ptr, c_null_ptr)
  ! MoveParticlesEuler, MethodOrder-1 times (src/DEM.f90:87-92)
  do k = 1, int(p%method_order) - 1
    if (rc == DEM_OK) rc = dem_euler_startup(h)
  end do
  ! MoveParticles(Step) for Step = MethodOrder, MethodOrder+nsteps-1
  if (rc == DEM_OK) rc = dem_step(h, nsteps, int(p%method_order, c_int64_t), done, escaped)
  if (rc == DEM_OK) rc = dem_download_state(h, int(np, c_int32_t), c_loc(comp), c_loc(x), &
                                            c_loc(v), c_loc(w), c_loc(a), c_loc(alpha), &
                                            c_null_ptr, c_null_ptr)
  if (rc == DEM_OK) rc = dem_stats(h, st)
  if (rc /= DEM_OK) write(*, '(a, i0, a, a)') 'drive_steps: rc ', rc, ': ', dem_error_message(h)

  open(newunit=u, file=trim(outp), access="stream", form="unformatted", status="replace")
  write(u) int(rc, c_int64_t), done, st%ncol, st%pair_contacts, st%wall_contacts
  write(u) x, v, w, a, alpha
  close(u)
  if (c_associated(h)) call dem_destroy(h)
  if (rc /= DEM_OK) stop 1
end program drive_steps
