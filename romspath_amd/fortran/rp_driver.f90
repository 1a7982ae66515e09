! rp_driver.f90 -- Fortran host driving librp.so through romspath_hip.
!
! Same loop structure as the reference main program: run_ROMSPath ->
! run_External_Timestep -> run_Internal_Timestep (ROMSPath.f90:151-546),
! with `call update_particles()` replaced by rp_step.  Namelist groups and
! names are the reference's (ROMSPath.h:69-258).  Grid and hydro records are
! the analytic flat-bottom double gyre of the CPU-runnable case (no netCDF
! library exists in this environment), or, when the namelist has a
! $romsoutput group with a prefix, from ROMS history files in the netCDF
! classic format read through rp_ncin (InitGrid and updateHydro, one nest,
! getFileNames and the multifile roll-over); particles come from the reference's
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
  ! history files (ROMSPath.h: prefix, suffix, filenum, numdigits, multifile, time names)
  character(len=200) :: prefix(5), suffix, time_vname, time_dname
  integer :: filenum, numdigits
  logical :: multifile
  namelist /romsoutput/ prefix, suffix, time_vname, time_dname, filenum, numdigits, multifile

  integer :: L, M, N, vtr, ios, tdim, open_num
  logical :: from_files
  double precision :: hcv, sph(1), tmp1(1)
  double precision, parameter :: dx = 1000.d0, h0 = 50.d0, U0 = 0.3d0, pi = 3.14159265358979323846d0
  double precision, allocatable, target :: s_rho(:), cs_r(:), s_w(:), cs_w(:)
  double precision, allocatable, target :: h(:, :), pm(:, :), pn(:, :), tmp2(:, :)
  integer(c_int32_t), allocatable, target :: mrho(:, :), mu(:, :), mv(:, :)
  type(c_ptr) :: fh
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
  prefix = ' '
  suffix = '.nc'
  time_vname = 'ocean_time'
  time_dname = 'ocean_time'
  filenum = 1
  numdigits = 4
  multifile = .false.
  rewind (iu)
  read (iu, nml=romsoutput, iostat=ios)
  close (iu)
  from_files = ios == 0 .and. len_trim(prefix(1)) > 0
  open_num = -1
  fh = c_null_ptr

  if (from_files) then
    ! --- InitGrid (grid_module.f90:41-422) from the first history file ---
    call open_hist(filenum)
    L = int(dimlen('xi_rho')); M = int(dimlen('eta_rho')); N = int(dimlen('s_rho'))
    allocate (s_rho(N), cs_r(N), s_w(N + 1), cs_w(N + 1), h(L, M), pm(L, M), pn(L, M))
    allocate (mrho(L, M), mu(L - 1, M), mv(L, M - 1), xr(L, M), yr(L, M), tmp2(L, M))
    call getv('s_rho', s_rho, N); call getv('Cs_r', cs_r, N)
    call getv('s_w', s_w, N + 1); call getv('Cs_w', cs_w, N + 1)
    call getv('h', h, L * M); call getv('pm', pm, L * M); call getv('pn', pn, L * M)
    call getv('mask_rho', tmp2, L * M)
    mrho = nint(tmp2)
    deallocate (tmp2)
    allocate (tmp2(L - 1, M))
    call getv('mask_u', tmp2, (L - 1) * M)
    mu = nint(tmp2)
    deallocate (tmp2)
    allocate (tmp2(L, M - 1))
    call getv('mask_v', tmp2, L * (M - 1))
    mv = nint(tmp2)
    deallocate (tmp2)
    call getv('Vtransform', tmp1, 1)
    vtr = nint(tmp1(1))
    call getv('hc', tmp1, 1)
    hcv = tmp1(1)
    call getv('spherical', sph, 1)
    if (nint(sph(1)) == 1) then
      call getv('lon_rho', xr, L * M); call getv('lat_rho', yr, L * M)
    else
      call getv('x_rho', xr, L * M); call getv('y_rho', yr, L * M)
    end if
  else
    ! --- grid: Vtransform 2 / Vstretching 4 (theta_s 5, theta_b 0.4, hc 10) ---
    L = 100; M = 80; N = 20; vtr = 2; hcv = 10.d0
    allocate (s_rho(N), cs_r(N), s_w(N + 1), cs_w(N + 1), h(L, M), pm(L, M), pn(L, M))
    allocate (mrho(L, M), mu(L - 1, M), mv(L, M - 1), xr(L, M), yr(L, M))
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
    ! lon_rho/lat_rho of the grid (x_rho/y_rho on a Cartesian grid, InitGrid)
    do j = 1, M
      do i = 1, L
        xr(i, j) = (i - 1) * dx
        yr(i, j) = (j - 1) * dx
      end do
    end do
  end if

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

  g%L = L; g%M = M; g%N = N; g%vtransform = vtr; g%hc = hcv
  g%s_rho = c_loc(s_rho); g%cs_r = c_loc(cs_r); g%s_w = c_loc(s_w); g%cs_w = c_loc(cs_w)
  g%h = c_loc(h); g%pm = c_loc(pm); g%pn = c_loc(pn)
  g%mask_rho = c_loc(mrho); g%mask_u = c_loc(mu); g%mask_v = c_loc(mv)
  g%scl = 1.d0
  g%off = 0.d0
  rc = rp_set_grid(ctx, 1_c_int32_t, g)
  if (rc /= 0) call die('rp_set_grid', rc)

  allocate (zeta(L, M), u(L - 1, M, N), v(L, M - 1, N), w(L, M, N + 1), aks(L, M, N + 1))
  if (from_files) then
    ! records 1, 2, 3 of the first file into slots 1, 2, 3 (ROMSPath.f90:307-309)
    do k = 1, 3
      call load_slot(k, k, filenum)
    end do
  else
  ! --- steady double gyre, depth-uniform (u at u points, v at v points) ---
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
  end if

  ! --- particles (par(:,pX..pDOB) of ini_ROMSPath) ---
  allocate (px(numpar), py(numpar), pz(numpar), page(numpar), pdob(numpar), pst(numpar), pgid(numpar), &
            pin(numpar), wd(numpar), pzeta(numpar), bath(numpar))
  allocate (plon(numpar), plat(numpar))
  open (newunit=iu, file=trim(parfile), status='old', action='read')
  do ip = 1, numpar
    pdob(ip) = -9.d0
    read (iu, *) plon(ip), plat(ip), pz(ip), pdob(ip)
  end do
  close (iu)
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
    ! file roll-over (ROMSPath.f90:473-480)
    if (from_files) then
      if (tstep > tdim) then
        tstep = 1
        filenum = filenum + 1
        if (.not. multifile) then
          write (*, '(A)') 'NOT ENOUGH TIMES STEP TO CONTINUE:'
          stop 1
        end if
      end if
    end if
    if (ets > 2) then
      t_b = mod(t_b, 3) + 1
      t_c = mod(t_c, 3) + 1
      t_f = mod(t_f, 3) + 1
      if (from_files) then
        call load_slot(t_f, tstep, filenum)
      else
        rc = rp_set_hydro_slot(ctx, 1_c_int32_t, int(t_f, c_int32_t), zeta, u, v, w, aks, c_null_ptr, c_null_ptr)
        if (rc /= 0) call die('rp_set_hydro_slot', rc)
      end if
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
  if (open_num >= 0) rc = rp_ncin_close(fh)
  rc = rp_destroy(ctx)

contains

  ! getFileNames (hydrodynamic_module.f90:774-808) and NF90_OPEN of file fnum;
  ! tdim is the record count of the file last opened, as updateHydro sets it
  subroutine open_hist(fnum)
    integer, intent(in) :: fnum
    character(len=400) :: fname
    character(len=16) :: fmt
    if (fnum == open_num) return
    if (open_num >= 0) rc = rp_ncin_close(fh)
    if (multifile) then
      write (fmt, '(A,I1,A,I1,A)') '(A,I', numdigits, '.', numdigits, ',A)'
      write (fname, fmt) trim(prefix(1)), fnum, trim(suffix)
    else
      fname = prefix(1)
    end if
    rc = rp_ncin_open(fh, trim(fname) // c_null_char)
    if (rc /= 0) then
      write (*, '(A,A)') 'cannot read ', trim(fname)
      stop 1
    end if
    open_num = fnum
    tdim = int(dimlen(trim(time_dname)))
  end subroutine open_hist

  integer(c_int64_t) function dimlen(name)
    character(len=*), intent(in) :: name
    rc = rp_ncin_dim(fh, trim(name) // c_null_char, dimlen)
    if (rc /= 0) call die_nc(name)
  end function dimlen

  ! NF90_GET_VAR of a whole fixed variable
  subroutine getv(name, a, cnt)
    character(len=*), intent(in) :: name
    double precision, intent(out) :: a(*)
    integer, intent(in) :: cnt
    rc = rp_ncin_get(fh, trim(name) // c_null_char, -1_c_int64_t, a, int(cnt, c_int64_t))
    if (rc /= 0) call die_nc(name)
  end subroutine getv

  ! NF90_GET_VAR of record rec (1-based) of a record variable
  subroutine getr(name, rec, a, cnt)
    character(len=*), intent(in) :: name
    integer, intent(in) :: rec, cnt
    double precision, intent(out) :: a(*)
    rc = rp_ncin_get(fh, trim(name) // c_null_char, int(rec - 1, c_int64_t), a, int(cnt, c_int64_t))
    if (rc /= 0) call die_nc(name)
  end subroutine getr

  ! updateHydro(.., rec, slot) for the one nest: zeta, u, v, w, AKs of record
  ! rec of file fnum (hydrodynamic_module.f90:232-512)
  subroutine load_slot(slot, rec, fnum)
    integer, intent(in) :: slot, rec, fnum
    call open_hist(fnum)
    call getr('zeta', rec, zeta, L * M)
    call getr('u', rec, u, (L - 1) * M * N)
    call getr('v', rec, v, L * (M - 1) * N)
    call getr('w', rec, w, L * M * (N + 1))
    call getr('AKs', rec, aks, L * M * (N + 1))
    rc = rp_set_hydro_slot(ctx, 1_c_int32_t, int(slot, c_int32_t), zeta, u, v, w, aks, c_null_ptr, c_null_ptr)
    if (rc /= 0) call die('rp_set_hydro_slot', rc)
  end subroutine load_slot

  subroutine die_nc(what)
    character(len=*), intent(in) :: what
    write (*, '(A,A)') 'netCDF read failed: ', what
    stop 1
  end subroutine die_nc

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
