! romspath_hip.f90 -- ISO_C_BINDING interface to librp.so (include/rp_abi.h).
!
! The reference driver (ROMSPath.f90) calls update_particles() once per
! internal step (ROMSPath.f90:529).  With this module it instead hands its
! state to the GPU once and calls rp_step per internal step; see the
! rp_driver program next to this file for the call sequence.
module romspath_hip
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: RP_MAXGRID = 5
  integer(c_int32_t), parameter, public :: RP_PREC_FP32_FIELDS = 0, RP_PREC_FP64 = 1

  ! rp_config: namelist scalars of PARAM_MOD on the hot path plus build knobs
  type, bind(C), public :: rp_config
    integer(c_int32_t) :: idt, dt
    real(c_double) :: deltat, serr, smth, sub, aksback
    integer(c_int32_t) :: vturb_on, hturb_on
    real(c_double) :: kh(RP_MAXGRID)
    integer(c_int32_t) :: scheme, nsb
    real(c_double) :: vertdist
    integer(c_int32_t) :: open_boundary, ngrid
    integer(c_int64_t) :: seed
    integer(c_int32_t) :: precision, block, sort_interval
    integer(c_int32_t) :: reserved(5)
  end type rp_config

  ! rp_nc_meta (include/rp_ncout.h): global attributes of the trajectory file
  integer(c_int32_t), parameter, public :: RP_NC_WATERDEPTH = 1, RP_NC_ZETA = 2, RP_NC_BATH = 4, &
                                           RP_NC_SALTTEMP = 8, RP_NC_BOTTOM = 16, &
                                           RP_NC_COLLISIONS = 32
  type, bind(C), public :: rp_nc_meta
    type(c_ptr) :: time_units, run_name
    integer(c_int32_t) :: idt
    real(c_float) :: days
    integer(c_int32_t) :: behavior, open_boundary, vturb_on, hturb_on, seed
    real(c_double) :: deltat, serr, smth, sub, aksback
    real(c_double) :: swimfast, temp_offset
    integer(c_int32_t) :: salt_temp_mean
    type(c_ptr) :: executable, output_loc, administrator, institution
  end type rp_nc_meta

  ! rp_grid: GRIDDATA after InitGrid.  C's scl[5][2] is Fortran scl(2,5):
  ! scl(:,i) is the reference's GRIDS(ng)%scl(i,:).
  type, bind(C), public :: rp_grid
    integer(c_int32_t) :: L, M, N, vtransform
    real(c_double) :: hc
    type(c_ptr) :: s_rho, cs_r, s_w, cs_w
    type(c_ptr) :: h, pm, pn
    type(c_ptr) :: mask_rho, mask_u, mask_v
    real(c_double) :: scl(2, RP_MAXGRID), off(2, RP_MAXGRID)
  end type rp_grid

  public :: rp_create, rp_set_grid, rp_set_hydro_slot, rp_set_particles, rp_step, &
            rp_get_particles, rp_set_deviates, rp_sample_deviates, rp_active_steps, &
            rp_secant_capped, rp_synchronize, rp_kernel_timing, rp_phase_timing, rp_modal_basis, &
            rp_last_error, &
            rp_destroy, rp_set_tuning, rp_set_scenarios, rp_step_scenario, rp_step_batch, &
            rp_set_perturbation, rp_adjoint_record, rp_adjoint, &
            rp_set_lonlat, rp_sample_lonlat, rp_nc_create, rp_nc_write, rp_nc_close, &
            rp_set_grid_angle, rp_locate_particles, rp_set_hydro_tracers, rp_sample_tracers, &
            rp_set_hydro_bstress, rp_set_hydro_wetdry, rp_bottom_begin, rp_bottom_update, rp_bottom_get, &
            rp_tracers_begin, rp_tracers_update, rp_tracers_get, &
            rp_ncin_open, rp_ncin_dim, rp_ncin_var, rp_ncin_get, rp_ncin_close, &
            rp_ncin_att_text, rp_ncin_error, rp_nc_records, rp_nc_error, &
            rp_error_text

  interface
    integer(c_int) function rp_create(ctx, cfg) bind(C, name="rp_create")
      import :: c_int, c_ptr, rp_config
      type(c_ptr), intent(out) :: ctx
      type(rp_config), intent(in) :: cfg
    end function

    integer(c_int) function rp_set_grid(ctx, ng, g) bind(C, name="rp_set_grid")
      import :: c_int, c_int32_t, c_ptr, rp_grid
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ng
      type(rp_grid), intent(in) :: g
    end function

    integer(c_int) function rp_set_hydro_slot(ctx, ng, slot, zeta, u, v, w, aks, su, sv) &
        bind(C, name="rp_set_hydro_slot")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ng, slot
      real(c_double), intent(in) :: zeta(*), u(*), v(*), w(*), aks(*)
      type(c_ptr), value :: su, sv
    end function

    integer(c_int) function rp_set_particles(ctx, n, x, y, z, age, dob, status, gid, isin) &
        bind(C, name="rp_set_particles")
      import :: c_int, c_int32_t, c_int64_t, c_int8_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: x(*), y(*), z(*), age(*), dob(*), status(*)
      integer(c_int32_t), intent(in) :: gid(*)
      integer(c_int8_t), intent(in) :: isin(*)
    end function

    integer(c_int) function rp_step(ctx, ex, ix, ets, t_b, t_c, t_f, step_index, want_diag) &
        bind(C, name="rp_step")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ex(3), ix(3)
      integer(c_int32_t), value :: ets, t_b, t_c, t_f
      integer(c_int64_t), value :: step_index
      integer(c_int32_t), value :: want_diag
    end function

    ! rp_step without any device allocation (no reference counterpart): for
    ! hosts that cannot allocate, e.g. inside a captured graph
    integer(c_int) function rp_step_noalloc(ctx, ex, ix, ets, t_b, t_c, t_f, step_index, &
        want_diag) bind(C, name="rp_step_noalloc")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ex(3), ix(3)
      integer(c_int32_t), value :: ets, t_b, t_c, t_f
      integer(c_int64_t), value :: step_index
      integer(c_int32_t), value :: want_diag
    end function

    integer(c_int) function rp_get_particles(ctx, n, x, y, z, age, status, gid, isin, wd, zeta, &
        bath) bind(C, name="rp_get_particles")
      import :: c_int, c_int32_t, c_int64_t, c_int8_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: x(*), y(*), z(*), age(*), status(*), wd(*), zeta(*), bath(*)
      integer(c_int32_t), intent(out) :: gid(*)
      integer(c_int8_t), intent(out) :: isin(*)
    end function

    integer(c_int) function rp_set_deviates(ctx, normals, n, ndev) bind(C, name="rp_set_deviates")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: ctx, normals
      integer(c_int64_t), value :: n
      integer(c_int32_t), value :: ndev
    end function

    integer(c_int) function rp_sample_deviates(ctx, step_index, n, ndev, out) &
        bind(C, name="rp_sample_deviates")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: step_index, n
      integer(c_int32_t), value :: ndev
      real(c_double), intent(out) :: out(*)
    end function

    integer(c_int) function rp_active_steps(ctx, count) bind(C, name="rp_active_steps")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: count
    end function

    integer(c_int) function rp_secant_capped(ctx, count) bind(C, name="rp_secant_capped")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: count
    end function

    integer(c_int) function rp_synchronize(ctx) bind(C, name="rp_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function rp_kernel_timing(ctx, avg_ms, launches) bind(C, name="rp_kernel_timing")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: avg_ms
      integer(c_int64_t), intent(out) :: launches
    end function

    integer(c_int) function rp_modal_basis(n, basis, lam) bind(C, name="rp_modal_basis")
      import :: c_int, c_int32_t, c_float, c_double
      integer(c_int32_t), value :: n
      real(c_float), intent(out) :: basis(64, 64, 2)
      real(c_double), intent(out) :: lam(64, 2)
    end function

    integer(c_int) function rp_phase_timing(ctx, advect_ms, walk_ms) bind(C, name="rp_phase_timing")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: advect_ms, walk_ms
    end function

    integer(c_int) function rp_set_tuning(ctx, block, sort_interval, tile_x, tile_y) &
        bind(C, name="rp_set_tuning")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: block, sort_interval, tile_x, tile_y
    end function

    integer(c_int) function rp_set_scenarios(ctx, nscen, counts) bind(C, name="rp_set_scenarios")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nscen
      integer(c_int64_t), intent(in) :: counts(*)
    end function

    integer(c_int) function rp_step_scenario(ctx, scen, ex, ix, ets, t_b, t_c, t_f, step_index, &
        want_diag) bind(C, name="rp_step_scenario")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: scen, ets, t_b, t_c, t_f, want_diag
      real(c_double), intent(in) :: ex(3), ix(3)
      integer(c_int64_t), value :: step_index
    end function

    integer(c_int) function rp_set_perturbation(ctx, nx, ny, coef) bind(C, name="rp_set_perturbation")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nx, ny
      real(c_double), intent(in) :: coef(*)
    end function

    integer(c_int) function rp_adjoint_record(ctx, ckpt_every) bind(C, name="rp_adjoint_record")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ckpt_every
    end function

    integer(c_int) function rp_adjoint(ctx, n, lam_x, lam_y, lam_z, g_x, g_y, g_z, g_pert) &
        bind(C, name="rp_adjoint")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: lam_x(*), lam_y(*), lam_z(*)
      real(c_double), intent(out) :: g_x(*), g_y(*), g_z(*)
      type(c_ptr), value :: g_pert
    end function

    integer(c_int) function rp_step_batch(ctx, nsteps, ex, ix, ets, slots, step_index, want_diag, &
        use_graph) bind(C, name="rp_step_batch")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps, use_graph
      real(c_double), intent(in) :: ex(3, *), ix(3, *)
      integer(c_int32_t), intent(in) :: ets(*), slots(3, *), want_diag(*)
      integer(c_int64_t), intent(in) :: step_index(*)
    end function

    ! GRIDS(ng)%lon_rho / lat_rho (x_rho / y_rho on Cartesian grids)
    integer(c_int) function rp_set_lonlat(ctx, ng, lon, lat) bind(C, name="rp_set_lonlat")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ng
      real(c_double), intent(in) :: lon(*), lat(*)
    end function

    ! GRIDS(ng)%angle and GRIDS(1)%spherical, for LL2ij
    integer(c_int) function rp_set_grid_angle(ctx, ng, spherical, angle) bind(C, name="rp_set_grid_angle")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ng, spherical
      type(c_ptr), value :: angle
    end function

    ! particle placement of ini_ROMSPath (ROMSPath.f90:322-403): LL2ij, bounds, zbounds
    integer(c_int) function rp_locate_particles(ctx, n, lon, lat, z, t_b, x, y, status, gid, isin) &
        bind(C, name="rp_locate_particles")
      import :: c_int, c_int32_t, c_int64_t, c_int8_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: lon(*), lat(*)
      real(c_double), intent(inout) :: z(*)
      integer(c_int32_t), value :: t_b
      real(c_double), intent(out) :: x(*), y(*), status(*)
      integer(c_int32_t), intent(out) :: gid(*)
      integer(c_int8_t), intent(out) :: isin(*)
    end function

    ! getInterp2D("lon"/"lat") at every particle (dataOutput, ROMSPath.f90:1012-1025)
    integer(c_int) function rp_sample_lonlat(ctx, n, lon, lat) bind(C, name="rp_sample_lonlat")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: lon(*), lat(*)
    end function

    ! createNetCDF(par(:,pDOB)) (ROMSPath.f90:449, 1140-1868); path NUL-terminated
    integer(c_int) function rp_nc_create(nc, path, numpar, ngrid, flags, meta, dob) &
        bind(C, name="rp_nc_create")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_char, c_double, rp_nc_meta
      type(c_ptr), intent(out) :: nc
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), value :: numpar
      integer(c_int32_t), value :: ngrid, flags
      type(rp_nc_meta), intent(in) :: meta
      real(c_double), intent(in) :: dob(*)
    end function

    ! writeNetCDF (ROMSPath.f90:1870-2157): one record
    integer(c_int) function rp_nc_write(nc, model_time, age, lon, lat, zp, status, gid, wd, zeta, &
        bath, salt, temp, hob, bustr, bvstr) bind(C, name="rp_nc_write")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: nc
      real(c_double), value :: model_time
      real(c_double), intent(in) :: age(*), lon(*), lat(*), zp(*), status(*)
      integer(c_int32_t), intent(in) :: gid(*)
      type(c_ptr), value :: wd, zeta, bath, salt, temp, hob, bustr, bvstr
    end function

    ! SaltTempOn: salt/temp of slot tind (updateHydro) and P_salt/P_temp at the particles
    integer(c_int) function rp_set_hydro_tracers(ctx, ng, slot, salt, temp) &
        bind(C, name="rp_set_hydro_tracers")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ng, slot
      real(c_double), intent(in) :: salt(*), temp(*)
    end function

    integer(c_int) function rp_sample_tracers(ctx, n, ex, ix, t_b, t_c, t_f, salt, temp) &
        bind(C, name="rp_sample_tracers")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: ex(3), ix(3)
      integer(c_int32_t), value :: t_b, t_c, t_f
      real(c_double), intent(out) :: salt(*), temp(*)
    end function

    ! netCDF classic input (include/rp_ncin.h): NF90_OPEN / NF90_INQUIRE_DIMENSION /
    ! NF90_INQUIRE_VARIABLE / NF90_GET_VAR of one record / NF90_CLOSE without netCDF-Fortran
    integer(c_int) function rp_ncin_open(f, path) bind(C, name="rp_ncin_open")
      import :: c_int, c_ptr, c_char
      type(c_ptr), intent(out) :: f
      character(kind=c_char), intent(in) :: path(*)
    end function

    integer(c_int) function rp_ncin_dim(f, name, len) bind(C, name="rp_ncin_dim")
      import :: c_int, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: f
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int64_t), intent(out) :: len
    end function

    integer(c_int) function rp_ncin_var(f, name, ndims, shape, is_record) bind(C, name="rp_ncin_var")
      import :: c_int, c_ptr, c_char, c_int32_t, c_int64_t
      type(c_ptr), value :: f
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int32_t), intent(out) :: ndims, is_record
      integer(c_int64_t), intent(out) :: shape(8)
    end function

    integer(c_int) function rp_ncin_get(f, name, rec, out, count) bind(C, name="rp_ncin_get")
      import :: c_int, c_ptr, c_char, c_int64_t, c_double
      type(c_ptr), value :: f
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int64_t), value :: rec, count
      real(c_double), intent(out) :: out(*)
    end function

    integer(c_int) function rp_ncin_close(f) bind(C, name="rp_ncin_close")
      import :: c_int, c_ptr
      type(c_ptr), value :: f
    end function

    ! NF90_GET_ATT of a text attribute (var: c_null_ptr for a global one)
    integer(c_int) function rp_ncin_att_text(f, var, att, buf, size, len) bind(C, name="rp_ncin_att_text")
      import :: c_int, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: f, var
      character(kind=c_char), intent(in) :: att(*)
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int64_t), value :: size
      integer(c_int64_t), intent(out) :: len
    end function

    ! message of the reader's last failure (f: c_null_ptr for rp_ncin_open), a C string
    type(c_ptr) function rp_ncin_error(f) bind(C, name="rp_ncin_error")
      import :: c_ptr
      type(c_ptr), value :: f
    end function

    ! records written so far, and the writer's last failure (a C string)
    integer(c_int64_t) function rp_nc_records(nc) bind(C, name="rp_nc_records")
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: nc
    end function

    type(c_ptr) function rp_nc_error(nc) bind(C, name="rp_nc_error")
      import :: c_ptr
      type(c_ptr), value :: nc
    end function

    ! SaltTempOn / SaltTempMean kept per step: P_Salt, P_Temp (or their means)
    integer(c_int) function rp_tracers_begin(ctx, mean, t_c) bind(C, name="rp_tracers_begin")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mean, t_c
    end function

    integer(c_int) function rp_tracers_update(ctx, ex, ix, t_b, t_c, t_f) &
        bind(C, name="rp_tracers_update")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ex(3), ix(3)
      integer(c_int32_t), value :: t_b, t_c, t_f
    end function

    integer(c_int) function rp_tracers_get(ctx, n, salt, temp, reset_mean) &
        bind(C, name="rp_tracers_get")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: salt(*), temp(*)
      integer(c_int32_t), value :: reset_mean
    end function

    ! WriteBottom: bustr/bvstr of slot tind, P_HOB/P_bustr/P_bvstr at the particles
    integer(c_int) function rp_set_hydro_bstress(ctx, ng, slot, bustr, bvstr) &
        bind(C, name="rp_set_hydro_bstress")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ng, slot
      real(c_double), intent(in) :: bustr(*), bvstr(*)
    end function

    ! -DWETDRY: wetdry_mask_rho (xi_rho, eta_rho), wetdry_mask_u (xi_u, eta_u),
    ! wetdry_mask_v (xi_v, eta_v) of time slot 1..3 (hydrodynamic_module.f90:510-570)
    integer(c_int) function rp_set_hydro_wetdry(ctx, ng, slot, mask_rho, mask_u, mask_v) &
        bind(C, name="rp_set_hydro_wetdry")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ng, slot
      real(c_double), intent(in) :: mask_rho(*), mask_u(*), mask_v(*)
    end function

    integer(c_int) function rp_bottom_begin(ctx) bind(C, name="rp_bottom_begin")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function rp_bottom_update(ctx, ex, ix, t_b, t_c, t_f) &
        bind(C, name="rp_bottom_update")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ex(3), ix(3)
      integer(c_int32_t), value :: t_b, t_c, t_f
    end function

    integer(c_int) function rp_bottom_get(ctx, n, hob, bustr, bvstr, reset_hob) &
        bind(C, name="rp_bottom_get")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: hob(*), bustr(*), bvstr(*)
      integer(c_int32_t), value :: reset_hob
    end function

    integer(c_int) function rp_nc_close(nc) bind(C, name="rp_nc_close")
      import :: c_int, c_ptr
      type(c_ptr), value :: nc
    end function

    type(c_ptr) function rp_last_error(ctx) bind(C, name="rp_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function rp_destroy(ctx) bind(C, name="rp_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! rp_last_error as a Fortran string, for errorHandler-style reporting
  function rp_error_text(ctx) result(msg)
    type(c_ptr), value :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n, i
    p = rp_last_error(ctx)
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = s(i)
    end do
  end function rp_error_text

end module romspath_hip
