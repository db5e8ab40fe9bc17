! check_bindings.f90 -- prints the sizes of the bind(c) types of dem3d_gpu
! and calls dem_create once through the bindings.  Built and run by
! tests/test_abi.py (no GPU needed: without a device dem_create returns
! DEM_ERR_DEVICE and a message, which exercises the error path of the ABI).
program check_bindings
  use, intrinsic :: iso_c_binding
  use dem3d_gpu
  implicit none
  type(dem_params_t) :: p
  type(dem_wall_t) :: w
  type(dem_particle_t) :: q
  type(dem_stats_t) :: s
  type(dem_wall_motion_t) :: m
  type(dem_config_t) :: cf
  type(dem_run_t) :: rn
  type(dem_state_sections_t) :: sc
  type(dem_rng_t) :: rg
  type(dem_pour_t) :: po
  type(dem_fluid_t) :: fl
  type(dem_tune_t) :: tu
  type(dem_observables_t) :: ob
  type(c_ptr) :: h
  real(c_double), target :: radius(1), tab(1)
  integer(c_int) :: rc

  write(*, '(a, i0)') 'sizeof dem_params_t ', c_sizeof(p)
  write(*, '(a, i0)') 'sizeof dem_wall_t ', c_sizeof(w)
  write(*, '(a, i0)') 'sizeof dem_particle_t ', c_sizeof(q)
  write(*, '(a, i0)') 'sizeof dem_stats_t ', c_sizeof(s)
  write(*, '(a, i0)') 'sizeof dem_wall_motion_t ', c_sizeof(m)
  write(*, '(a, i0)') 'sizeof dem_config_t ', c_sizeof(cf)
  write(*, '(a, i0)') 'sizeof dem_run_t ', c_sizeof(rn)
  write(*, '(a, i0)') 'sizeof dem_state_sections_t ', c_sizeof(sc)
  write(*, '(a, i0)') 'sizeof dem_rng_t ', c_sizeof(rg)
  write(*, '(a, i0)') 'sizeof dem_pour_t ', c_sizeof(po)
  write(*, '(a, i0)') 'sizeof dem_fluid_t ', c_sizeof(fl)
  write(*, '(a, i0)') 'sizeof dem_tune_t ', c_sizeof(tu)
  write(*, '(a, i0)') 'sizeof dem_observables_t ', c_sizeof(ob)

  radius = 0.5d0
  tab = 1.0d0
  p%nc = 1
  p%nw = 0
  p%pbc = [1, 1, 1]
  p%lbox = [10.0d0, 10.0d0, 10.0d0]
  p%time_step = 1.0d-3
  p%radius = c_loc(radius)
  p%mass = c_loc(tab)
  p%inertia = c_loc(tab)
  p%inv_mass = c_loc(tab)
  p%inv_inertia = c_loc(tab)
  p%eff_mass = c_loc(tab)
  p%min_dij = c_loc(tab)
  p%min_dij_sq = c_loc(tab)
  p%kn = c_loc(tab)
  p%yn = c_loc(tab)
  p%kt = c_loc(tab)
  p%yt = c_loc(tab)
  p%musq = c_loc(tab)
  rc = dem_create(p, h)
  write(*, '(a, i0)') 'dem_create ', rc
  if (rc /= DEM_OK) then
    write(*, '(a, a)') 'message ', dem_error_message(c_null_ptr)
  else
    call dem_destroy(h)
  end if
end program check_bindings
