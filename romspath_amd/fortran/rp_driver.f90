! rp_driver.f90 -- Fortran host driving librp.so through romspath_hip.
!
! Same loop structure as the reference main program: run_ROMSPath ->
! run_External_Timestep -> run_Internal_Timestep (ROMSPath.f90:151-546),
! with `call update_particles()` replaced by rp_step.  Namelist groups and
! names are the reference's (ROMSPath.h:69-258).  Grid and hydro records are
! the analytic flat-bottom double gyre of the CPU-runnable case (no netCDF
! library exists in this environment); particles come from the reference's
! release file, "lon lat Z DOB" per line read list-directed
! (ROMSPath.f90:319-332), placed as ini_ROMSPath places them (LL2ij, bounds,
! zbounds: rp_locate_particles); on this Cartesian grid lon/lat are x_rho/y_rho
! in metres, as InitGrid reads them.
!
! usage: rp_driver <namelist file> <particle file> <output file> [<netCDF file>]
! With a fourth argument the run also writes the reference's trajectory file,
! one record whenever printdt reaches iprint (ROMSPath.f90:532-543).
program rp_driver
  use iso_c_binding
  use romspath_hip
  implicit none

  ! --- namelist groups (names as in ROMSPath.h) ---
  integer :: numpar
  real :: days
  integer :: iprint, dt, idt
  logical :: VTurbOn, HTurbOn
  double precision :: serr, smth, sub, deltat, AKSback
  double precision :: ConstantHTurb(5)
  integer :: scheme, nsb
  double precision :: vertdist
  integer :: Ngrid
  double precision :: refine(5)
  namelist /numparticles/ numpar
  namelist /timeparam/ days, iprint, dt, idt
  namelist /turbparam/ VTurbOn, serr, smth, sub, deltat, AKSback, HTurbOn, ConstantHTurb
  namelist /advectparam/ scheme, nsb, vertdist
  namelist /romsgrid/ Ngrid, refine

  integer, parameter :: L = 100, M = 80, N = 20
  double precision, parameter :: dx = 1000.d0, h0 = 50.d0, U0 = 0.3d0, pi = 3.14159265358979323846d0
  double precision, target :: s_rho(N), cs_r(N), s_w(N + 1), cs_w(N + 1)
  double precision, target :: h(L, M), pm(L, M), pn(L, M)
  integer(c_int32_t), target :: mrho(L, M), mu(L - 1, M), mv(L, M - 1)
  double precision, allocatable :: zeta(:, :), u(:, :, :), v(:, :, :), w(:, :, :), aks(:, :, :)
  double precision, allocatable :: px(:), py(:), pz(:), page(:), pdob(:), pst(:), wd(:), pzeta(:), bath(:)
  integer(c_int32_t), allocatable :: pgid(:)
  integer(c_int8_t), allocatable :: pin(:)

  type(rp_config) :: cfg
  type(rp_grid) :: g
  type(c_ptr) :: ctx
  character(len=256) :: nmlfile, parfile, outfile, ncfile
  type(c_ptr) :: nc
  type(rp_nc_meta), target :: meta
  character(kind=c_char, len=32), target :: tunits
  double precision, allocatable :: plon(:), plat(:), xr(:, :), yr(:, :)
  integer :: printdt
  logical :: write_nc
  integer :: rc, ets, its, stepT, stepIT, t_b, t_c, t_f, tstep, ip, i, j, k, iu
  double precision :: ex(3), ix(3), x, y, Lx, Ly, th
  integer(c_int64_t) :: step_index, nact
  double precision :: ms
  integer(c_int64_t) :: nl

  call get_command_argument(1, nmlfile)
  call get_command_argument(2, parfile)
  call get_command_argument(3, outfile)
  call get_command_argument(4, ncfile)
  write_nc = len_trim(ncfile) > 0
  ConstantHTurb = 0.d0
  refine = 1.d0
  open (newunit=iu, file=trim(nmlfile), status='old', action='read')
  read (iu, nml=numparticles)
  read (iu, nml=timeparam)
  read (iu, nml=turbparam)
  read (iu, nml=advectparam)
  read (iu, nml=romsgrid)
  close (iu)

  ! --- grid: Vtransform 2 / Vstretching 4 (theta_s 5, theta_b 0.4, hc 10) ---
  do k = 1, N
    s_rho(k) = (dble(k) - N - 0.5d0) / N
    cs_r(k) = stretch(s_rho(k))
  end do
  do k = 1, N + 1
    s_w(k) = (dble(k - 1) - N) / N
    cs_w(k) = stretch(s_w(k))
  end do
  cs_w(1) = -1.d0
  cs_w(N + 1) = 0.d0
  h = h0
  pm = 1.d0 / dx
  pn = 1.d0 / dx
  mrho = 1
  mu = 1
  mv = 1

  cfg%idt = idt
  cfg%dt = dt
  cfg%deltat = deltat
  cfg%serr = serr
  cfg%smth = smth
  cfg%sub = sub
  cfg%aksback = AKSback
  cfg%vturb_on = merge(1, 0, VTurbOn)
  cfg%hturb_on = merge(1, 0, HTurbOn)
  cfg%kh = ConstantHTurb
  cfg%scheme = scheme
  cfg%nsb = nsb
  cfg%vertdist = vertdist
  cfg%open_boundary = 1
  cfg%ngrid = Ngrid
  cfg%seed = 1_c_int64_t
  cfg%precision = RP_PREC_FP64
  cfg%block = 0
  cfg%sort_interval = 0
  cfg%reserved = 0
  rc = rp_create(ctx, cfg)
  if (rc /= 0) call die('rp_create', rc)

  g%L = L; g%M = M; g%N = N; g%vtransform = 2; g%hc = 10.d0
  g%s_rho = c_loc(s_rho); g%cs_r = c_loc(cs_r); g%s_w = c_loc(s_w); g%cs_w = c_loc(cs_w)
  g%h = c_loc(h); g%pm = c_loc(pm); g%pn = c_loc(pn)
  g%mask_rho = c_loc(mrho); g%mask_u = c_loc(mu); g%mask_v = c_loc(mv)
  g%scl = 1.d0
  g%off = 0.d0
  rc = rp_set_grid(ctx, 1_c_int32_t, g)
  if (rc /= 0) call die('rp_set_grid', rc)

  ! --- steady double gyre, depth-uniform (u at u points, v at v points) ---
  allocate (zeta(L, M), u(L - 1, M, N), v(L, M - 1, N), w(L, M, N + 1), aks(L, M, N + 1))
  Lx = (L - 1) * dx
  Ly = (M - 1) * dx
  zeta = 0.d0
  w = 0.d0
  aks = 0.d0
  do j = 1, M
    do i = 1, L - 1
      x = (i - 0.5d0) * dx
      y = (j - 1) * dx
      u(i, j, :) = -U0 * sin(pi * 2.d0 * x / Lx) * cos(pi * y / Ly)
    end do
  end do
  do j = 1, M - 1
    do i = 1, L
      x = (i - 1) * dx
      y = (j - 0.5d0) * dx
      v(i, j, :) = U0 * (2.d0 * Ly / Lx) * cos(pi * 2.d0 * x / Lx) * sin(pi * y / Ly)
    end do
  end do
  ! records 1, 2, 3 into slots 1, 2, 3 (ROMSPath.f90:307-309); the field is steady
  do k = 1, 3
    rc = rp_set_hydro_slot(ctx, 1_c_int32_t, int(k, c_int32_t), zeta, u, v, w, aks, c_null_ptr, c_null_ptr)
    if (rc /= 0) call die('rp_set_hydro_slot', rc)
  end do

  ! --- particles (par(:,pX..pDOB) of ini_ROMSPath) ---
  allocate (px(numpar), py(numpar), pz(numpar), page(numpar), pdob(numpar), pst(numpar), pgid(numpar), &
            pin(numpar), wd(numpar), pzeta(numpar), bath(numpar))
  allocate (plon(numpar), plat(numpar), xr(L, M), yr(L, M))
  open (newunit=iu, file=trim(parfile), status='old', action='read')
  do ip = 1, numpar
    pdob(ip) = -9.d0
    read (iu, *) plon(ip), plat(ip), pz(ip), pdob(ip)
  end do
  close (iu)
  ! lon_rho/lat_rho of the grid (x_rho/y_rho on a Cartesian grid, InitGrid)
  do j = 1, M
    do i = 1, L
      xr(i, j) = (i - 1) * dx
      yr(i, j) = (j - 1) * dx
    end do
  end do
  rc = rp_set_lonlat(ctx, 1_c_int32_t, xr, yr)
  if (rc /= 0) call die('rp_set_lonlat', rc)
  rc = rp_set_grid_angle(ctx, 1_c_int32_t, 0_c_int32_t, c_null_ptr)
  if (rc /= 0) call die('rp_set_grid_angle', rc)
  nl = numpar
  ! ini_ROMSPath's placement with the hydro of slot t_b = 1 (ROMSPath.f90:366-403)
  rc = rp_locate_particles(ctx, nl, plon, plat, pz, 1_c_int32_t, px, py, pst, pgid, pin)
  if (rc /= 0) call die('rp_locate_particles', rc)
  page = 0.d0
  rc = rp_set_particles(ctx, nl, px, py, pz, page, pdob, pst, pgid, pin)
  if (rc /= 0) call die('rp_set_particles', rc)

  ! --- trajectory file (initNetCDF / createNetCDF, ROMSPath.f90:446-449) ---
  if (write_nc) then
    tunits = 'seconds since model start' // c_null_char
    meta%time_units = c_loc(tunits)
    meta%run_name = c_null_ptr
    meta%idt = idt
    meta%days = days
    meta%behavior = 0
    meta%open_boundary = 1
    meta%vturb_on = merge(1, 0, VTurbOn)
    meta%hturb_on = merge(1, 0, HTurbOn)
    meta%seed = 1
    meta%deltat = deltat; meta%serr = serr; meta%smth = smth; meta%sub = sub; meta%aksback = AKSback
    rc = rp_nc_create(nc, trim(ncfile) // c_null_char, nl, 1_c_int32_t, 0_c_int32_t, meta, pdob)
    if (rc /= 0) call die('rp_nc_create', rc)
  end if
  printdt = 0

  ! --- time loop (ROMSPath.f90:168-174, 465-546) ---
  t_b = 1; t_c = 2; t_f = 3; tstep = 1
  stepT = int(days * 86400.0) / dt
  stepIT = int(dt / idt)
  step_index = 0
  do ets = 1, stepT
    if (ets > 2) then
      t_b = mod(t_b, 3) + 1
      t_c = mod(t_c, 3) + 1
      t_f = mod(t_f, 3) + 1
      rc = rp_set_hydro_slot(ctx, 1_c_int32_t, int(t_f, c_int32_t), zeta, u, v, w, aks, c_null_ptr, c_null_ptr)
      if (rc /= 0) call die('rp_set_hydro_slot', rc)
    end if
    ex(1) = (ets - 2) * dt
    ex(2) = (ets - 1) * dt
    ex(3) = ets * dt
    do its = 1, stepIT
      ix(1) = ex(2) + dble((its - 2) * idt)
      ix(2) = ex(2) + dble((its - 1) * idt)
      ix(3) = ex(2) + dble(its * idt)
      rc = rp_step(ctx, ex, ix, int(ets, c_int32_t), int(t_b, c_int32_t), int(t_c, c_int32_t), &
                   int(t_f, c_int32_t), step_index, 0_c_int32_t)
      if (rc /= 0) call die('rp_step', rc)
      step_index = step_index + 1
      ! PRINT OUTPUT TO FILE (ROMSPath.f90:532-543)
      printdt = printdt + idt
      if (printdt >= iprint) then
        if (write_nc) call data_output(int(ix(3)))
        printdt = 0
      end if
    end do
    tstep = tstep + 1
  end do

  rc = rp_get_particles(ctx, nl, px, py, pz, page, pst, pgid, pin, wd, pzeta, bath)
  if (rc /= 0) call die('rp_get_particles', rc)
  rc = rp_active_steps(ctx, nact)
  rc = rp_kernel_timing(ctx, ms, nl)
  open (newunit=iu, file=trim(outfile), status='replace', action='write')
  do ip = 1, numpar
    write (iu, '(3ES25.17,F6.2)') px(ip), py(ip), pz(ip), pst(ip)
  end do
  close (iu)
  write (*, '(A,I0,A,I0,A,F9.4,A)') 'rp_driver: ', step_index, ' steps, ', nact, &
       ' particle-steps, ', ms, ' ms per step kernel'
  if (write_nc) then
    rc = rp_nc_close(nc)
    if (rc /= 0) call die('rp_nc_close', rc)
  end if
  rc = rp_destroy(ctx)

contains

  ! dataOutput (ROMSPath.f90:1012-1025): lon/lat at the particles, one record
  subroutine data_output(time)
    integer, intent(in) :: time
    rc = rp_get_particles(ctx, nl, px, py, pz, page, pst, pgid, pin, wd, pzeta, bath)
    if (rc /= 0) call die('rp_get_particles', rc)
    rc = rp_sample_lonlat(ctx, nl, plon, plat)
    if (rc /= 0) call die('rp_sample_lonlat', rc)
    rc = rp_nc_write(nc, dble(time), page, plon, plat, pz, pst, pgid, c_null_ptr, c_null_ptr, c_null_ptr, &
                     c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr)
    if (rc /= 0) call die('rp_nc_write', rc)
  end subroutine data_output

  double precision function stretch(s)
    double precision, intent(in) :: s
    double precision :: csur
    csur = (1.d0 - cosh(5.d0 * s)) / (cosh(5.d0) - 1.d0)
    stretch = (exp(0.4d0 * csur) - 1.d0) / (1.d0 - exp(-0.4d0))
  end function stretch

  subroutine die(what, code)
    character(len=*), intent(in) :: what
    integer, intent(in) :: code
    write (*, '(A,A,I0,A,A)') what, ' failed: ', code, ' ', rp_error_text(ctx)
    stop 1
  end subroutine die

end program rp_driver
