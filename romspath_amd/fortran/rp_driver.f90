! rp_driver.f90 -- the reference's main program as a Fortran host of librp.so
! through the romspath_hip module (ISO_C_BINDING).
!
! getParams (parameter_module.f90:22-150) reads the namelist file named on the
! command line group by group, in the reference's order and with its error
! codes and messages.  Every group and every variable of ROMSPath.h:69-345 is
! declared here, so the reference's own ROMSPath.data reads unchanged.  Then
! ini_ROMSPath (ROMSPath.f90:190-455): InitGrid of every nest from its first
! history file, prefix(ng) (grid_module.f90:41-422), the nest maps scl/off
! (grid_module.f90:437-481), records 1, 2, 3 of every nest into slots 1, 2, 3
! (updateHydro, hydrodynamic_module.f90:112-760, with the read*/const* pairs,
! TempOffset, salt/temp under SaltTempOn and bustr/bvstr under WriteBottom),
! the release file parfile placed by rp_locate_particles, the SaltTempOn and
! WriteBottom samplers, createNetCDF with the file name of $output and the
! release record writeNetCDF(0, pLon, pLat).  Then run_ROMSPath ->
! run_External_Timestep -> run_Internal_Timestep (ROMSPath.f90:151-546) with
! update_particles replaced by rp_step and the sampler updates, the multifile
! roll-over, and a record whenever printdt reaches iprint; model_time is
! time + reftime of the last nest (ROMSPath.f90:1931).
!
! usage: rp_driver [-n] [--stokes] [--wetdry] [--fp32] <namelist file> [<final-state file>]
!   -n         read the namelist, print what was read, stop (touches no GPU)
!   --stokes   add the Stokes drift of stokesprefix(ng) to u, v: the reference
!              built with -DSTOKES (hydrodynamic_module.f90:573-660)
!   --wetdry   read each record's wetdry_mask_rho/_u/_v into its slot: the
!              reference built with -DWETDRY (hydrodynamic_module.f90:510-570)
!   --fp32     the FP32-field form of the step (default: FP64, the reference's)
! The final-state file, when named, gets "X Y Z status GID age" per particle.
!
! With prefix(1) blank the grid and the records are the analytic flat-bottom
! double gyre of the CPU-runnable case (one nest; salt, temp = constSalt,
! constTemp; bustr = bvstr = 0), since the reference itself needs history
! files.  Larval behaviour (Behavior /= 0, Process_VA, Process_WA) and
! settlement are modules outside the particle step: the driver stops with a
! message when they are switched on.  A failed library call stops with status 1.
program rp_driver
  use iso_c_binding
  use romspath_hip
  implicit none

  integer, parameter :: MAXNgrid = 5

  ! ---- namelist variables, groups and types as in ROMSPath.h:69-345 ----
  integer :: numpar
  real :: days
  integer :: iprint, dt, idt
  logical :: readZeta, readSalt, readTemp, readU, readV, readW, readAks, readDens
  double precision :: constZeta, constSalt, constTemp, constU, constV, constW, constAks, constDens
  character(len=200) :: stokesprefix(MAXNgrid), turbstd_v_a_prefix(MAXNgrid), wavestd_prefix(MAXNgrid)
  logical :: Process_VA, Process_WA
  logical :: VTurbOn, HTurbOn
  double precision :: serr, smth, sub, deltat, AKSback, ConstantHTurb(MAXNgrid)
  integer :: scheme, nsb
  double precision :: vertdist
  integer :: Behavior
  logical :: OpenOceanBoundary
  double precision :: pediage, swimstart, swimslow, swimfast, Sgradient, sink, Hswimspeed, Swimdepth
  double precision :: vort_cr, vort_sat, b0pv, b1pv, b0wv, b1w, acc_cr, acc_sat, b0pa, b1pa, b0wa
  integer :: va_flag
  integer :: Growth
  logical :: mortality
  double precision :: deadage, initsize, maxsize, tempcut, a0, a1, a2, a3, a4, a5, a6, a7, a8
  double precision :: twistart, twiend, daylength, Em, Kd, thresh
  logical :: settlementon, holesExist
  integer :: minpolyid, maxpolyid, minholeid, maxholeid, pedges, hedges
  integer :: Ngrid
  double precision :: refine(MAXNgrid)
  character(len=200) :: prefix(MAXNgrid), suffix
  character(len=100) :: time_vname, time_dname
  integer :: filenum, numdigits
  logical :: multifile
  character(len=200) :: parfile
  character(len=200) :: habitatfile, holefile
  character(len=200) :: outpath
  character(len=100) :: NCOutFile
  logical :: outpathGiven
  integer :: NCtime
  character(len=100) :: RunName
  character(len=200) :: ExeDir, OutDir
  character(len=100) :: RunBy, Institution
  character(len=200) :: StartedOn
  integer :: seed, ErrorFlag
  logical :: SaltTempOn, SaltTempMean, WriteBottom, WriteWaterDepth, WriteZeta, WriteBath
  double precision :: TempOffset
  logical :: TrackCollisions, WriteHeaders, WriteModelTiming, WriteProblemFile, FreeSlip
  integer :: ijbuff

  namelist /numparticles/ numpar
  namelist /timeparam/ days, iprint, dt, idt
  namelist /hydroparam/ readZeta, constZeta, readSalt, constSalt, readTemp, constTemp, readU, &
    constU, readV, constV, readW, constW, readAks, constAks, readDens, constDens, stokesprefix, &
    turbstd_v_a_prefix, wavestd_prefix, Process_VA, Process_WA
  namelist /turbparam/ VTurbOn, serr, smth, sub, deltat, AKSback, HTurbOn, ConstantHTurb
  namelist /advectparam/ scheme, nsb, vertdist
  namelist /behavparam/ Behavior, OpenOceanBoundary, pediage, swimstart, swimslow, swimfast, &
    Sgradient, sink, Hswimspeed, Swimdepth
  namelist /fuchsparam/ vort_cr, vort_sat, b0pv, b1pv, b0wv, b1w, acc_cr, acc_sat, b0pa, b1pa, &
    b0wa, va_flag
  namelist /growparam/ Growth, mortality, deadage, initsize, maxsize, tempcut, a0, a1, a2, a3, &
    a4, a5, a6, a7, a8
  namelist /dvmparam/ twistart, twiend, daylength, Em, Kd, thresh
  namelist /settleparam/ settlementon, holesExist, minpolyid, maxpolyid, minholeid, maxholeid, &
    pedges, hedges
  namelist /romsgrid/ Ngrid, refine
  namelist /romsoutput/ prefix, suffix, filenum, numdigits, multifile, time_vname, time_dname
  namelist /parloc/ parfile
  namelist /HabPolyLoc/ habitatfile, holefile
  namelist /output/ outpath, NCOutFile, outpathGiven, NCtime, RunName, ExeDir, OutDir, RunBy, &
    Institution, StartedOn
  namelist /other/ seed, SaltTempOn, TrackCollisions, WriteHeaders, SaltTempMean, WriteBottom, &
    WriteModelTiming, WriteProblemFile, ijbuff, ErrorFlag, FreeSlip, TempOffset, &
    WriteWaterDepth, WriteZeta, WriteBath

  ! ---- one nest: GRIDS(ng) of InitGrid and its open history files ----
  type nest_t
    integer :: L = 0, M = 0, N = 0, vtr = 0, sph = 0
    double precision :: hc = 0.d0
    double precision, allocatable :: s_rho(:), cs_r(:), s_w(:), cs_w(:)
    double precision, allocatable :: h(:, :), pm(:, :), pn(:, :), lon(:, :), lat(:, :), angle(:, :)
    integer(c_int32_t), allocatable :: mrho(:, :), mu(:, :), mv(:, :)
    type(c_ptr) :: fh = c_null_ptr, sfh = c_null_ptr
    integer :: fnum = -1, sfnum = -1
  end type nest_t
  type(nest_t), allocatable, target :: nests(:)

  double precision, parameter :: dx = 1000.d0, h0 = 50.d0, U0 = 0.3d0, pi = 3.14159265358979323846d0
  double precision :: sclv(2, MAXNgrid, MAXNgrid), offv(2, MAXNgrid, MAXNgrid)
  double precision :: reftime
  character(len=100) :: time_units
  logical :: from_files, use_stokes, use_wetdry, check_only, fp32
  ! analytic records (prefix(1) blank)
  double precision, allocatable :: gz(:, :), gu(:, :, :), gv(:, :, :), gw(:, :, :), gaks(:, :, :)

  double precision, allocatable, target :: px(:), py(:), pz(:), page(:), pdob(:), pst(:), &
    wd(:), pzeta(:), bath(:), psalt(:), ptemp(:), phob(:), pbus(:), pbvs(:), plon(:), plat(:), &
    rlon(:), rlat(:)
  integer(c_int32_t), allocatable :: pgid(:)
  integer(c_int8_t), allocatable :: pin(:)

  type(rp_config) :: cfg
  type(rp_grid) :: g
  type(c_ptr) :: ctx, nc
  type(rp_nc_meta), target :: meta
  character(kind=c_char, len=101), target :: c_units, c_run, c_exe, c_out, c_by, c_inst
  character(len=256) :: nmlfile, statefile, arg
  character(len=420) :: ncfile
  integer :: printdt, tstep, t_b, t_c, t_f, ets, its, stepT, stepIT, ng, ip, iu, rc, nflags, k, nargs
  double precision :: ex(3), ix(3), ms
  integer(c_int64_t) :: step_index, nact, nl, nlaunch
  integer(c_int32_t) :: want

  ! ---- command line ----
  check_only = .false.
  use_stokes = .false.
  use_wetdry = .false.
  fp32 = .false.
  nmlfile = ' '
  statefile = ' '
  nargs = 0
  do k = 1, command_argument_count()
    call get_command_argument(k, arg)
    select case (trim(arg))
    case ('-n')
      check_only = .true.
    case ('--stokes')
      use_stokes = .true.
    case ('--wetdry')
      use_wetdry = .true.
    case ('--fp32')
      fp32 = .true.
    case default
      nargs = nargs + 1
      if (nargs == 1) nmlfile = arg
      if (nargs == 2) statefile = arg
    end select
  end do
  if (len_trim(nmlfile) == 0) then
    write (*, '(A)') 'usage: rp_driver [-n] [--stokes] [--wetdry] [--fp32] <namelist file> [<final-state file>]'
    stop 1
  end if

  call get_params(nmlfile)
  if (check_only) then
    call print_params()
    stop
  end if
  if (Ngrid < 1 .or. Ngrid > MAXNgrid) call fatal('Ngrid must be 1 .. 5')
  if (Behavior /= 0 .or. Process_VA .or. Process_WA) &
    call fatal('larval behaviour (Behavior /= 0, Process_VA, Process_WA) is outside the particle step')
  if (settlementon) call fatal('the settlement module is outside the particle step')
  from_files = len_trim(prefix(1)) > 0
  if (.not. from_files .and. Ngrid /= 1) call fatal('the analytic double gyre is one nest: give prefix(ng)')
  if (use_stokes .and. .not. from_files) call fatal('--stokes needs history files')
  if (use_wetdry .and. .not. from_files) call fatal('--wetdry needs history files')

  ! ---- InitGrid for every nest, then the nest maps ----
  allocate (nests(Ngrid))
  reftime = 0.d0
  time_units = 'seconds since model start'
  do ng = 1, Ngrid
    if (from_files) then
      call init_grid_file(nests(ng), ng)
    else
      call init_grid_gyre(nests(ng))
    end if
  end do
  call nest_maps()

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
  cfg%open_boundary = merge(1, 0, OpenOceanBoundary)
  cfg%ngrid = Ngrid
  if (seed == 0) then
    ! init_random_seed: a seed from the clock when the namelist gives 0
    call system_clock(count=k)
    seed = max(1, abs(k))
  end if
  cfg%seed = int(seed, c_int64_t)
  cfg%precision = merge(RP_PREC_FP32_FIELDS, RP_PREC_FP64, fp32)
  cfg%block = 0
  cfg%sort_interval = 0
  cfg%reserved = 0
  rc = rp_create(ctx, cfg)
  if (rc /= 0) call die('rp_create', rc)

  do ng = 1, Ngrid
    g%L = nests(ng)%L; g%M = nests(ng)%M; g%N = nests(ng)%N
    g%vtransform = nests(ng)%vtr; g%hc = nests(ng)%hc
    g%s_rho = c_loc(nests(ng)%s_rho); g%cs_r = c_loc(nests(ng)%cs_r)
    g%s_w = c_loc(nests(ng)%s_w); g%cs_w = c_loc(nests(ng)%cs_w)
    g%h = c_loc(nests(ng)%h); g%pm = c_loc(nests(ng)%pm); g%pn = c_loc(nests(ng)%pn)
    g%mask_rho = c_loc(nests(ng)%mrho); g%mask_u = c_loc(nests(ng)%mu); g%mask_v = c_loc(nests(ng)%mv)
    g%scl = sclv(:, :, ng)
    g%off = offv(:, :, ng)
    rc = rp_set_grid(ctx, int(ng, c_int32_t), g)
    if (rc /= 0) call die('rp_set_grid', rc)
    rc = rp_set_lonlat(ctx, int(ng, c_int32_t), nests(ng)%lon, nests(ng)%lat)
    if (rc /= 0) call die('rp_set_lonlat', rc)
    ! LL2ij takes spherical from GRIDS(1) for every nest
    rc = rp_set_grid_angle(ctx, int(ng, c_int32_t), int(nests(1)%sph, c_int32_t), c_loc(nests(ng)%angle))
    if (rc /= 0) call die('rp_set_grid_angle', rc)
  end do

  ! ---- updateHydro(.TRUE.,1,t_b), (2,t_c), (3,t_f) (ROMSPath.f90:307-309) ----
  t_b = 1; t_c = 2; t_f = 3
  tstep = 1
  do k = 1, 3
    call load_slot(k, k, filenum)
  end do

  ! ---- the release file and the placement (ROMSPath.f90:316-403) ----
  nl = numpar
  allocate (px(numpar), py(numpar), pz(numpar), page(numpar), pdob(numpar), pst(numpar), &
            pgid(numpar), pin(numpar), wd(numpar), pzeta(numpar), bath(numpar), psalt(numpar), &
            ptemp(numpar), phob(numpar), pbus(numpar), pbvs(numpar), plon(numpar), plat(numpar), &
            rlon(numpar), rlat(numpar))
  open (newunit=iu, file=trim(parfile), status='old', action='read', iostat=rc)
  if (rc /= 0) call fatal('cannot open the particle locations file ' // trim(parfile))
  do ip = 1, numpar
    pdob(ip) = -9.d0
    read (iu, *) rlon(ip), rlat(ip), pz(ip), pdob(ip)
  end do
  close (iu)
  rc = rp_locate_particles(ctx, nl, rlon, rlat, pz, int(t_b, c_int32_t), px, py, pst, pgid, pin)
  if (rc /= 0) call die('rp_locate_particles', rc)
  page = 0.d0
  rc = rp_set_particles(ctx, nl, px, py, pz, page, pdob, pst, pgid, pin)
  if (rc /= 0) call die('rp_set_particles', rc)

  ! ---- the samplers' initial state (ROMSPath.f90:247-266, 408-424) ----
  if (WriteBottom) then
    rc = rp_bottom_begin(ctx)
    if (rc /= 0) call die('rp_bottom_begin', rc)
  end if
  if (SaltTempOn) then
    rc = rp_tracers_begin(ctx, merge(1_c_int32_t, 0_c_int32_t, SaltTempMean), int(t_c, c_int32_t))
    if (rc /= 0) call die('rp_tracers_begin', rc)
  end if

  ! ---- initNetCDF / createNetCDF / writeNetCDF(0, pLon, pLat) (ROMSPath.f90:446-453) ----
  call create_netcdf()
  call data_output(0, .true.)
  printdt = 0

  ! ---- time loop (ROMSPath.f90:168-174, 465-546) ----
  stepT = int(days * 86400.0) / dt
  stepIT = int(dt / idt)
  step_index = 0
  do ets = 1, stepT
    ! file roll-over on nest 1's record count (ROMSPath.f90:473-480)
    if (from_files) then
      if (tstep > tdim_of(filenum)) then
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
      call load_slot(t_f, tstep, filenum)
    end if
    ex(1) = (ets - 2) * dt
    ex(2) = (ets - 1) * dt
    ex(3) = ets * dt
    do its = 1, stepIT
      ix(1) = ex(2) + dble((its - 2) * idt)
      ix(2) = ex(2) + dble((its - 1) * idt)
      ix(3) = ex(2) + dble(its * idt)
      ! the step computes pWD/pZeta/pBath for the record that follows it
      want = merge(1_c_int32_t, 0_c_int32_t, printdt + idt >= iprint)
      rc = rp_step(ctx, ex, ix, int(ets, c_int32_t), int(t_b, c_int32_t), int(t_c, c_int32_t), &
                   int(t_f, c_int32_t), step_index, want)
      if (rc /= 0) call die('rp_step', rc)
      if (WriteBottom) then
        rc = rp_bottom_update(ctx, ex, ix, int(t_b, c_int32_t), int(t_c, c_int32_t), int(t_f, c_int32_t))
        if (rc /= 0) call die('rp_bottom_update', rc)
      end if
      if (SaltTempOn) then
        rc = rp_tracers_update(ctx, ex, ix, int(t_b, c_int32_t), int(t_c, c_int32_t), int(t_f, c_int32_t))
        if (rc /= 0) call die('rp_tracers_update', rc)
      end if
      step_index = step_index + 1
      ! PRINT OUTPUT TO FILE (ROMSPath.f90:532-543)
      printdt = printdt + idt
      if (printdt >= iprint) then
        call data_output(int(ix(3)), .false.)
        printdt = 0
      end if
    end do
    tstep = tstep + 1
  end do

  rc = rp_get_particles(ctx, nl, px, py, pz, page, pst, pgid, pin, wd, pzeta, bath)
  if (rc /= 0) call die('rp_get_particles', rc)
  rc = rp_active_steps(ctx, nact)
  rc = rp_kernel_timing(ctx, ms, nlaunch)
  if (len_trim(statefile) > 0) then
    open (newunit=iu, file=trim(statefile), status='replace', action='write')
    do ip = 1, numpar
      write (iu, '(4ES25.17,I3,ES25.17)') px(ip), py(ip), pz(ip), pst(ip), pgid(ip), page(ip)
    end do
    close (iu)
  end if
  write (*, '(A,I0,A,I0,A,F9.4,A,I0,A)') 'rp_driver: ', step_index, ' steps, ', nact, &
       ' particle-steps, ', ms, ' ms per step kernel, ', rp_nc_records(nc), ' records'
  rc = rp_nc_close(nc)
  if (rc /= 0) call die('rp_nc_close', rc)
  do ng = 1, Ngrid
    if (c_associated(nests(ng)%fh)) rc = rp_ncin_close(nests(ng)%fh)
    if (c_associated(nests(ng)%sfh)) rc = rp_ncin_close(nests(ng)%sfh)
  end do
  rc = rp_destroy(ctx)

contains

  ! ---- getParams (parameter_module.f90:22-150) ----
  subroutine get_params(fname)
    character(len=*), intent(in) :: fname
    integer :: istat, err, u
    character(len=120) :: header
    call set_defaults()
    err = 0
    open (newunit=u, file=trim(fname), status='old', action='read', iostat=istat)
    if (istat /= 0) then
      write (*, '(A,A)') 'cannot open ', trim(fname)
      stop 1
    end if
    if (err == 0) then
      read (u, nml=numparticles, iostat=istat); if (istat /= 0) err = 10
    end if
    if (err == 0) then
      read (u, nml=timeparam, iostat=istat); if (istat /= 0) err = 20
    end if
    if (err == 0) then
      read (u, nml=hydroparam, iostat=istat); if (istat /= 0) err = 30
    end if
    if (err == 0) then
      read (u, nml=turbparam, iostat=istat); if (istat /= 0) err = 40
    end if
    if (err == 0) then
      read (u, nml=advectparam, iostat=istat); if (istat /= 0) err = 45
    end if
    if (err == 0) then
      read (u, nml=behavparam, iostat=istat); if (istat /= 0) err = 50
    end if
    if (err == 0) then
      read (u, nml=fuchsparam, iostat=istat); if (istat /= 0) err = 55
    end if
    if (err == 0) then
      read (u, nml=growparam, iostat=istat); if (istat /= 0) err = 60
    end if
    if (err == 0) then
      read (u, nml=dvmparam, iostat=istat); if (istat /= 0) err = 60
    end if
    if (err == 0) then
      read (u, nml=settleparam, iostat=istat); if (istat /= 0) err = 70
    end if
    if (err == 0) then
      read (u, nml=romsgrid, iostat=istat); if (istat /= 0) err = 90
    end if
    if (err == 0) then
      read (u, nml=romsoutput, iostat=istat); if (istat /= 0) err = 100
    end if
    if (err == 0) then
      read (u, nml=parloc, iostat=istat); if (istat /= 0) err = 110
    end if
    if (err == 0) then
      read (u, nml=HabPolyLoc, iostat=istat); if (istat /= 0) err = 120
    end if
    if (err == 0) then
      read (u, nml=output, iostat=istat); if (istat /= 0) err = 130
    end if
    if (err == 0) then
      read (u, nml=other, iostat=istat); if (istat /= 0) err = 140
    end if
    close (u)
    select case (err)
    case (0)
      header = 'No Errors'
    case (10)
      header = 'Error when reading numparticles, pls check ROMSPath.data'
    case (20)
      header = 'Error when reading timeparam, pls check ROMSPath.data'
    case (30)
      header = 'Error when reading hydroparam, pls check ROMSPath.data'
    case (40)
      header = 'Error when reading turbparam, pls check ROMSPath.data'
    case (45)
      header = 'Error when reading advectparam, pls check ROMSPath.data'
    case (50)
      header = 'Error when reading behavparam, pls check ROMSPath.data'
    case (55)
      header = 'Error when reading fuchsparam, pls check ROMSPath.data'
    case (60)
      header = 'Error when reading growthparam, pls check ROMSPath.data'
    case (70)
      header = 'Error when reading settleparam, pls check ROMSPath.data'
    case (90)
      header = 'Error when reading romsgrid, pls check ROMSPath.data'
    case (100)
      header = 'Error when reading romsoutput, pls check ROMSPath.data'
    case (110)
      header = 'Error when reading parloc, pls check ROMSPath.data'
    case (120)
      header = 'Error when reading HabPolLoc, pls check ROMSPath.data'
    case (130)
      header = 'Error when reading output, pls check ROMSPath.data'
    case (140)
      header = 'Error when reading other, pls check ROMSPath.data'
    case default
      header = 'Error: unexpected err number'
    end select
    if (err /= 0) then
      write (*, '(A)') trim(header)
      stop 1
    end if
  end subroutine get_params

  ! values for the variables a namelist file leaves out (the reference leaves
  ! them undefined); the arrays' unset entries matter for refine and prefix
  subroutine set_defaults()
    numpar = 0; days = 0.0; iprint = 3600; dt = 3600; idt = 60
    readZeta = .true.; readSalt = .true.; readTemp = .true.; readU = .true.; readV = .true.
    readW = .true.; readAks = .true.; readDens = .false.
    constZeta = 0.d0; constSalt = 0.d0; constTemp = 0.d0; constU = 0.d0; constV = 0.d0
    constW = 0.d0; constAks = 0.d0; constDens = 0.d0
    stokesprefix = ' '; turbstd_v_a_prefix = ' '; wavestd_prefix = ' '
    Process_VA = .false.; Process_WA = .false.
    VTurbOn = .false.; HTurbOn = .false.
    serr = 1.d-4; smth = 0.6d0; sub = 4.d0; deltat = 1.d0; AKSback = 1.d-8; ConstantHTurb = 0.d0
    scheme = 1; nsb = 0; vertdist = 0.d0
    Behavior = 0; OpenOceanBoundary = .true.
    pediage = 0.d0; swimstart = 0.d0; swimslow = 0.d0; swimfast = 0.d0; Sgradient = 0.d0
    sink = 0.d0; Hswimspeed = 0.d0; Swimdepth = 0.d0
    vort_cr = 0.d0; vort_sat = 0.d0; b0pv = 0.d0; b1pv = 0.d0; b0wv = 0.d0; b1w = 0.d0
    acc_cr = 0.d0; acc_sat = 0.d0; b0pa = 0.d0; b1pa = 0.d0; b0wa = 0.d0; va_flag = 0
    Growth = 0; mortality = .false.; deadage = 0.d0; initsize = 0.d0; maxsize = 0.d0
    tempcut = 0.d0; a0 = 0.d0; a1 = 0.d0; a2 = 0.d0; a3 = 0.d0; a4 = 0.d0; a5 = 0.d0
    a6 = 0.d0; a7 = 0.d0; a8 = 0.d0
    twistart = 0.d0; twiend = 0.d0; daylength = 0.d0; Em = 0.d0; Kd = 0.d0; thresh = 0.d0
    settlementon = .false.; holesExist = .false.
    minpolyid = 0; maxpolyid = 0; minholeid = 0; maxholeid = 0; pedges = 0; hedges = 0
    Ngrid = 1; refine = 1.d0
    prefix = ' '; suffix = '.nc'; filenum = 1; numdigits = 4; multifile = .false.
    time_vname = 'ocean_time'; time_dname = 'ocean_time'
    parfile = ' '; habitatfile = ' '; holefile = ' '
    outpath = ' '; NCOutFile = 'ROMSPath'; outpathGiven = .false.; NCtime = 0
    RunName = ' '; ExeDir = ' '; OutDir = ' '; RunBy = ' '; Institution = ' '; StartedOn = ' '
    seed = 1; ErrorFlag = 0
    SaltTempOn = .false.; SaltTempMean = .false.; WriteBottom = .false.
    WriteWaterDepth = .false.; WriteZeta = .false.; WriteBath = .false.; TempOffset = 0.d0
    TrackCollisions = .false.; WriteHeaders = .false.; WriteModelTiming = .false.
    WriteProblemFile = .false.; ijbuff = 4; FreeSlip = .false.
  end subroutine set_defaults

  ! what getParams read, one "name = value" per line, for checking a file
  subroutine print_params()
    integer :: i
    write (*, '(A,I0)') 'numpar = ', numpar
    write (*, '(A,ES24.16)') 'days = ', dble(days)
    write (*, '(A,I0,A,I0,A,I0)') 'iprint = ', iprint, ' dt = ', dt, ' idt = ', idt
    write (*, '(A,7L2)') 'read zeta salt temp u v w aks = ', readZeta, readSalt, readTemp, readU, &
      readV, readW, readAks
    write (*, '(A,L2,L2)') 'Process_VA Process_WA = ', Process_VA, Process_WA
    write (*, '(A,L2,L2)') 'VTurbOn HTurbOn = ', VTurbOn, HTurbOn
    write (*, '(A,5ES24.16)') 'serr smth sub deltat AKSback = ', serr, smth, sub, deltat, AKSback
    write (*, '(A,5ES24.16)') 'ConstantHTurb = ', ConstantHTurb
    write (*, '(A,I0,A,I0,A,ES24.16)') 'scheme = ', scheme, ' nsb = ', nsb, ' vertdist = ', vertdist
    write (*, '(A,I0,A,L2,A,ES24.16)') 'Behavior = ', Behavior, ' OpenOceanBoundary = ', &
      OpenOceanBoundary, ' swimfast = ', swimfast
    write (*, '(A,I0,A,L2)') 'va_flag = ', va_flag, ' mortality = ', mortality
    write (*, '(A,ES24.16)') 'thresh = ', thresh
    write (*, '(A,L2,A,I0)') 'settlementon = ', settlementon, ' hedges = ', hedges
    write (*, '(A,I0)') 'Ngrid = ', Ngrid
    write (*, '(A,5ES24.16)') 'refine = ', refine
    do i = 1, MAXNgrid
      write (*, '(A,I0,A,A)') 'prefix(', i, ') = ', trim(prefix(i))
    end do
    do i = 1, MAXNgrid
      write (*, '(A,I0,A,A)') 'stokesprefix(', i, ') = ', trim(stokesprefix(i))
    end do
    write (*, '(A,A,A,I0,A,I0,A,L2)') 'suffix = ', trim(suffix), ' filenum = ', filenum, &
      ' numdigits = ', numdigits, ' multifile = ', multifile
    write (*, '(A,A,A,A)') 'time_vname = ', trim(time_vname), ' time_dname = ', trim(time_dname)
    write (*, '(A,A)') 'parfile = ', trim(parfile)
    write (*, '(A,A)') 'habitatfile = ', trim(habitatfile)
    write (*, '(A,A,A,A,A,L2,A,I0)') 'outpath = ', trim(outpath), ' NCOutFile = ', trim(NCOutFile), &
      ' outpathGiven = ', outpathGiven, ' NCtime = ', NCtime
    write (*, '(A,A)') 'RunName = ', trim(RunName)
    write (*, '(A,A)') 'RunBy = ', trim(RunBy)
    write (*, '(A,A)') 'Institution = ', trim(Institution)
    write (*, '(A,I0,A,I0)') 'seed = ', seed, ' ErrorFlag = ', ErrorFlag
    write (*, '(A,6L2)') 'SaltTempOn SaltTempMean WriteBottom WriteWaterDepth WriteZeta WriteBath = ', &
      SaltTempOn, SaltTempMean, WriteBottom, WriteWaterDepth, WriteZeta, WriteBath
    write (*, '(A,ES24.16)') 'TempOffset = ', TempOffset
    write (*, '(A,L2,A,I0,A,L2)') 'TrackCollisions = ', TrackCollisions, ' ijbuff = ', ijbuff, &
      ' FreeSlip = ', FreeSlip
  end subroutine print_params

  ! ---- getFileNames (hydrodynamic_module.f90:774-808) ----
  function file_name(pfx, fnum) result(fname)
    character(len=*), intent(in) :: pfx
    integer, intent(in) :: fnum
    character(len=420) :: fname
    character(len=20) :: fmt
    if (multifile) then
      if (numdigits < 1 .or. numdigits > 8) then
        write (*, *) 'Model presently does not support numdigits of ', numdigits
        write (*, *) 'Please use numdigit value from 1 to 8'
        stop 1
      end if
      write (fmt, '(A,I1,A,I1,A)') '(A,I', numdigits, '.', numdigits, ',A)'
      write (fname, fmt) trim(pfx), fnum, trim(suffix)
    else
      fname = trim(pfx)
    end if
  end function file_name

  function open_nc(fname) result(f)
    character(len=*), intent(in) :: fname
    type(c_ptr) :: f
    rc = rp_ncin_open(f, trim(fname) // c_null_char)
    if (rc /= 0) call fatal('Problem with NF90_OPEN: ' // trim(fname))
  end function open_nc

  ! the history file fnum of nest ng (hydro) or of its Stokes series
  function nest_file(ng, fnum, stokes) result(f)
    integer, intent(in) :: ng, fnum
    logical, intent(in) :: stokes
    type(c_ptr) :: f
    if (stokes) then
      if (nests(ng)%sfnum /= fnum) then
        if (c_associated(nests(ng)%sfh)) rc = rp_ncin_close(nests(ng)%sfh)
        nests(ng)%sfh = open_nc(file_name(stokesprefix(ng), fnum))
        nests(ng)%sfnum = fnum
      end if
      f = nests(ng)%sfh
    else
      if (nests(ng)%fnum /= fnum) then
        if (c_associated(nests(ng)%fh)) rc = rp_ncin_close(nests(ng)%fh)
        nests(ng)%fh = open_nc(file_name(prefix(ng), fnum))
        nests(ng)%fnum = fnum
      end if
      f = nests(ng)%fh
    end if
  end function nest_file

  ! tdim(1): the record count of nest 1's file fnum
  integer function tdim_of(fnum)
    integer, intent(in) :: fnum
    tdim_of = int(dimlen(nest_file(1, fnum, .false.), time_dname))
  end function tdim_of

  integer(c_int64_t) function dimlen(f, name)
    type(c_ptr), intent(in) :: f
    character(len=*), intent(in) :: name
    rc = rp_ncin_dim(f, trim(name) // c_null_char, dimlen)
    if (rc /= 0) call fatal('netCDF dimension missing: ' // trim(name))
  end function dimlen

  ! NF90_GET_VAR of a whole fixed variable
  subroutine getv(f, name, a, cnt)
    type(c_ptr), intent(in) :: f
    character(len=*), intent(in) :: name
    double precision, intent(out) :: a(*)
    integer, intent(in) :: cnt
    rc = rp_ncin_get(f, trim(name) // c_null_char, -1_c_int64_t, a, int(cnt, c_int64_t))
    if (rc /= 0) call fatal('netCDF read failed: ' // trim(name))
  end subroutine getv

  ! NF90_GET_VAR of record rec (1-based) of a record variable
  subroutine getr(f, name, rec, a, cnt)
    type(c_ptr), intent(in) :: f
    character(len=*), intent(in) :: name
    integer, intent(in) :: rec, cnt
    double precision, intent(out) :: a(*)
    rc = rp_ncin_get(f, trim(name) // c_null_char, int(rec - 1, c_int64_t), a, int(cnt, c_int64_t))
    if (rc /= 0) call fatal('netCDF read failed: ' // trim(name))
  end subroutine getr

  ! the first value of a variable, rank 0, fixed or record
  double precision function first_value(f, name)
    type(c_ptr), intent(in) :: f
    character(len=*), intent(in) :: name
    integer(c_int32_t) :: nd, isrec
    integer(c_int64_t) :: shp(8), cnt
    double precision, allocatable :: a(:)
    integer :: i
    rc = rp_ncin_var(f, trim(name) // c_null_char, nd, shp, isrec)
    if (rc /= 0) call fatal('netCDF variable missing: ' // trim(name))
    cnt = 1
    do i = merge(2, 1, isrec /= 0), nd
      cnt = cnt * shp(i)
    end do
    allocate (a(max(cnt, 1_c_int64_t)))
    rc = rp_ncin_get(f, trim(name) // c_null_char, merge(0_c_int64_t, -1_c_int64_t, isrec /= 0), a, cnt)
    if (rc /= 0) call fatal('netCDF read failed: ' // trim(name))
    first_value = a(1)
  end function first_value

  ! ---- InitGrid (grid_module.f90:41-422) for nest ng from its first file ----
  subroutine init_grid_file(nt, inest)
    type(nest_t), intent(inout) :: nt
    integer, intent(in) :: inest
    type(c_ptr) :: f
    integer :: L, M, N
    double precision, allocatable :: tmp(:, :)
    double precision :: sv
    character(kind=c_char) :: buf(256)
    character(kind=c_char, len=101), target :: cvar
    character(len=256) :: units
    integer(c_int64_t) :: alen
    integer :: i
    f = open_nc(file_name(prefix(inest), filenum))
    L = int(dimlen(f, 'xi_rho')); M = int(dimlen(f, 'eta_rho')); N = int(dimlen(f, 's_rho'))
    if (dimlen(f, 'xi_u') /= L - 1 .or. dimlen(f, 'eta_v') /= M - 1 .or. dimlen(f, 's_w') /= N + 1) &
      call fatal('staggered dimensions do not match xi_rho, eta_rho, s_rho')
      nt%L = L; nt%M = M; nt%N = N
      allocate (nt%s_rho(N), nt%cs_r(N), nt%s_w(N + 1), nt%cs_w(N + 1), nt%h(L, M), nt%pm(L, M), &
                nt%pn(L, M), nt%lon(L, M), nt%lat(L, M), nt%angle(L, M), nt%mrho(L, M), &
                nt%mu(L - 1, M), nt%mv(L, M - 1))
      call getv(f, 's_rho', nt%s_rho, N); call getv(f, 'Cs_r', nt%cs_r, N)
      call getv(f, 's_w', nt%s_w, N + 1); call getv(f, 'Cs_w', nt%cs_w, N + 1)
      call getv(f, 'h', nt%h, L * M); call getv(f, 'pm', nt%pm, L * M); call getv(f, 'pn', nt%pn, L * M)
      call getv(f, 'angle', nt%angle, L * M)
      allocate (tmp(L, M))
      call getv(f, 'mask_rho', tmp, L * M)
      nt%mrho = nint(tmp)
      deallocate (tmp)
      allocate (tmp(L - 1, M))
      call getv(f, 'mask_u', tmp, (L - 1) * M)
      nt%mu = nint(tmp)
      deallocate (tmp)
      allocate (tmp(L, M - 1))
      call getv(f, 'mask_v', tmp, L * (M - 1))
      nt%mv = nint(tmp)
      deallocate (tmp)
      nt%vtr = nint(first_value(f, 'Vtransform'))
      nt%hc = first_value(f, 'hc')
      ! spherical: an int flag, or the character T/F in older files
      sv = first_value(f, 'spherical')
      nt%sph = merge(1, 0, nint(sv) == 1 .or. nint(sv) == ichar('T') .or. nint(sv) == ichar('t'))
      if (nt%sph == 1) then
        call getv(f, 'lon_rho', nt%lon, L * M); call getv(f, 'lat_rho', nt%lat, L * M)
      else
        call getv(f, 'x_rho', nt%lon, L * M); call getv(f, 'y_rho', nt%lat, L * M)
      end if
    ! reftime and time_units (grid_module.f90:374-415); the last nest's stay
    reftime = first_value(f, time_vname)
    cvar = trim(time_vname) // c_null_char
    rc = rp_ncin_att_text(f, c_loc(cvar), 'units' // c_null_char, buf, &
                          int(size(buf), c_int64_t), alen)
    if (rc /= 0) call fatal('NO time units')
    units = ' '
    do i = 1, int(min(alen, int(len(units), c_int64_t)))
      units(i:i) = buf(i)
    end do
    if (index(units, 'days') > 0) then
      reftime = reftime * dble(86400.0)
      time_units = 'seconds' // units(5:)
    else if (index(units, 'hours') > 0) then
      reftime = reftime * dble(3600.0)
      time_units = 'seconds' // units(6:)
    else if (index(units, 'minutes') > 0) then
      reftime = reftime * dble(60.0)
      time_units = 'seconds' // units(8:)
    else if (index(units, 'seconds') > 0) then
      time_units = units
    else
      call fatal('Not a recognized time unit')
    end if
    rc = rp_ncin_close(f)
  end subroutine init_grid_file

  ! the analytic grid: Vtransform 2 / Vstretching 4 (theta_s 5, theta_b 0.4, hc 10)
  subroutine init_grid_gyre(nt)
    type(nest_t), intent(inout) :: nt
    integer :: i, j, kk, L, M, N
    double precision :: x, y, Lx, Ly
    L = 100; M = 80; N = 20
    nt%L = L; nt%M = M; nt%N = N; nt%vtr = 2; nt%hc = 10.d0; nt%sph = 0
    allocate (nt%s_rho(N), nt%cs_r(N), nt%s_w(N + 1), nt%cs_w(N + 1), nt%h(L, M), nt%pm(L, M), &
              nt%pn(L, M), nt%lon(L, M), nt%lat(L, M), nt%angle(L, M), nt%mrho(L, M), &
              nt%mu(L - 1, M), nt%mv(L, M - 1))
    do kk = 1, N
      nt%s_rho(kk) = (dble(kk) - N - 0.5d0) / N
      nt%cs_r(kk) = stretch(nt%s_rho(kk))
    end do
    do kk = 1, N + 1
      nt%s_w(kk) = (dble(kk - 1) - N) / N
      nt%cs_w(kk) = stretch(nt%s_w(kk))
    end do
    nt%cs_w(1) = -1.d0
    nt%cs_w(N + 1) = 0.d0
    nt%h = h0
    nt%pm = 1.d0 / dx
    nt%pn = 1.d0 / dx
    nt%mrho = 1
    nt%mu = 1
    nt%mv = 1
    nt%angle = 0.d0
    ! x_rho / y_rho, which InitGrid keeps in lon_rho / lat_rho on a Cartesian grid
    do j = 1, M
      do i = 1, L
        nt%lon(i, j) = (i - 1) * dx
        nt%lat(i, j) = (j - 1) * dx
      end do
    end do
    ! steady, depth-uniform double gyre (u at u points, v at v points)
    allocate (gz(L, M), gu(L - 1, M, N), gv(L, M - 1, N), gw(L, M, N + 1), gaks(L, M, N + 1))
    Lx = (L - 1) * dx
    Ly = (M - 1) * dx
    gz = 0.d0
    gw = 0.d0
    gaks = 0.d0
    do j = 1, M
      do i = 1, L - 1
        x = (i - 0.5d0) * dx
        y = (j - 1) * dx
        gu(i, j, :) = -U0 * sin(pi * 2.d0 * x / Lx) * cos(pi * y / Ly)
      end do
    end do
    do j = 1, M - 1
      do i = 1, L
        x = (i - 1) * dx
        y = (j - 0.5d0) * dx
        gv(i, j, :) = U0 * (2.d0 * Ly / Lx) * cos(pi * 2.d0 * x / Lx) * sin(pi * y / Ly)
      end do
    end do
  end subroutine init_grid_gyre

  double precision function stretch(s)
    double precision, intent(in) :: s
    double precision :: csur
    csur = (1.d0 - cosh(5.d0 * s)) / (cosh(5.d0) - 1.d0)
    stretch = (exp(0.4d0 * csur) - 1.d0) / (1.d0 - exp(-0.4d0))
  end function stretch

  ! ---- nest maps X_other = scl*X + off (grid_module.f90:437-481) ----
  ! For every ordered pair (ng2 outer, ng inner) the first rho point of nest
  ! ng (i outer, j inner) whose coordinates equal a rho point of nest ng2
  ! (the first such, ii outer, jj inner) fixes the offsets; later pairs
  ! overwrite what earlier pairs wrote.  sclv(:, ng2, ng) is GRIDS(ng)%scl(ng2,:).
  subroutine nest_maps()
    double precision :: gref(MAXNgrid, MAXNgrid)
    integer :: a, b, d, kk, na, nb, i, j, ii, jj
    do a = 1, Ngrid
      do b = 1, Ngrid
        gref(a, b) = 1.d0
        d = abs(a - b)
        do kk = 1, d
          gref(a, b) = refine(d) * gref(a, b)
        end do
      end do
    end do
    sclv = 1.d0
    offv = 0.d0
    if (Ngrid < 2) return
    do nb = 1, Ngrid
      do na = 1, Ngrid
        call first_match(nests(na), nests(nb), i, j, ii, jj)
        sclv(:, nb, na) = 1.d0 / gref(na, nb)
        sclv(:, na, nb) = gref(nb, na)
        offv(1, nb, na) = dble(ii) - sclv(1, nb, na) * dble(i)
        offv(2, nb, na) = dble(jj) - sclv(2, nb, na) * dble(j)
        offv(1, na, nb) = dble(i) - sclv(1, na, nb) * dble(ii)
        offv(2, na, nb) = dble(j) - sclv(2, na, nb) * dble(jj)
      end do
    end do
  end subroutine nest_maps

  ! the first point (i, j) of nest a that coincides with a point (ii, jj) of
  ! nest b; one past the last indices of each when none does (the DO
  ! variables' values after loops that complete).  b's points are sorted by
  ! (x, y, position in ii-outer order) so each search is a bisection.
  subroutine first_match(na_, nb_, i, j, ii, jj)
    type(nest_t), intent(in) :: na_, nb_
    integer, intent(out) :: i, j, ii, jj
    integer, allocatable :: idx(:), tmpi(:)
    double precision, allocatable :: bx(:), by(:)
    integer :: nbp, p, lo, hi, mid, q, w, s, e1, e2, l1, l2, o, q1, q2
    double precision :: x, y
    logical :: take2
    nbp = nb_%L * nb_%M
    allocate (idx(nbp), tmpi(nbp), bx(nbp), by(nbp))
    ! position q = (ii-1)*M + jj, ii outer as the reference's search runs
    do ii = 1, nb_%L
      do jj = 1, nb_%M
        q = (ii - 1) * nb_%M + jj
        bx(q) = nb_%lon(ii, jj)
        by(q) = nb_%lat(ii, jj)
      end do
    end do
    do p = 1, nbp
      idx(p) = p
    end do
    ! bottom-up merge sort of idx by (x, y, q)
    w = 1
    do while (w < nbp)
      s = 1
      do while (s <= nbp)
        e1 = min(s + w - 1, nbp)
        e2 = min(s + 2 * w - 1, nbp)
        l1 = s; l2 = e1 + 1; o = s
        do while (l1 <= e1 .or. l2 <= e2)
          if (l2 > e2) then
            take2 = .false.
          else if (l1 > e1) then
            take2 = .true.
          else
            q1 = idx(l1); q2 = idx(l2)
            if (bx(q2) /= bx(q1)) then
              take2 = bx(q2) < bx(q1)
            else if (by(q2) /= by(q1)) then
              take2 = by(q2) < by(q1)
            else
              take2 = q2 < q1
            end if
          end if
          if (take2) then
            tmpi(o) = idx(l2); l2 = l2 + 1
          else
            tmpi(o) = idx(l1); l1 = l1 + 1
          end if
          o = o + 1
        end do
        s = s + 2 * w
      end do
      idx = tmpi
      w = 2 * w
    end do
    do i = 1, na_%L
      do j = 1, na_%M
        x = na_%lon(i, j)
        y = na_%lat(i, j)
        ! lower bound of (x, y) among b's sorted points: the smallest q
        lo = 1; hi = nbp + 1
        do while (lo < hi)
          mid = (lo + hi) / 2
          q = idx(mid)
          if (bx(q) < x .or. (bx(q) == x .and. by(q) < y)) then
            lo = mid + 1
          else
            hi = mid
          end if
        end do
        if (lo <= nbp) then
          q = idx(lo)
          if (bx(q) == x .and. by(q) == y) then
            ii = (q - 1) / nb_%M + 1
            jj = q - (ii - 1) * nb_%M
            return
          end if
        end if
      end do
    end do
    i = na_%L + 1; j = na_%M + 1; ii = nb_%L + 1; jj = nb_%M + 1
  end subroutine first_match

  ! ---- updateHydro(.., rec, slot) of every nest from file fnum ----
  subroutine load_slot(slot, rec, fnum)
    integer, intent(in) :: slot, rec, fnum
    integer :: ng_, L, M, N
    double precision, allocatable, target :: zeta(:), u(:), v(:), w(:), aks(:), su(:), sv(:), &
      salt(:), temp(:), bu(:), bv(:), wdr(:), wdu(:), wdv(:)
    type(c_ptr) :: f, sf
    do ng_ = 1, Ngrid
      L = nests(ng_)%L; M = nests(ng_)%M; N = nests(ng_)%N
      allocate (zeta(L * M), u((L - 1) * M * N), v(L * (M - 1) * N), w(L * M * (N + 1)), &
                aks(L * M * (N + 1)), salt(L * M * N), temp(L * M * N), bu((L - 1) * M), bv(L * (M - 1)))
      if (from_files) then
        f = nest_file(ng_, fnum, .false.)
        if (rec < 1 .or. rec > dimlen(f, time_dname)) call fatal('record past the end of a history file')
        call field(f, 'zeta', readZeta, constZeta, rec, zeta)
        call field(f, 'u', readU, constU, rec, u)
        call field(f, 'v', readV, constV, rec, v)
        call field(f, 'w', readW, constW, rec, w)
        call field(f, 'AKs', readAks, constAks, rec, aks)
        if (SaltTempOn) then
          call field(f, 'salt', readSalt, constSalt, rec, salt)
          if (readTemp) then
            call getr(f, 'temp', rec, temp, size(temp))
            temp = temp + TempOffset
          else
            temp = constTemp
          end if
        end if
        if (WriteBottom) then
          call getr(f, 'bustr', rec, bu, size(bu))
          call getr(f, 'bvstr', rec, bv, size(bv))
        end if
      else
        zeta = reshape(gz, [size(zeta)])
        u = reshape(gu, [size(u)])
        v = reshape(gv, [size(v)])
        w = reshape(gw, [size(w)])
        aks = reshape(gaks, [size(aks)])
        salt = constSalt
        temp = constTemp
        bu = 0.d0
        bv = 0.d0
      end if
      if (use_stokes) then
        allocate (su(size(u)), sv(size(v)))
        sf = nest_file(ng_, fnum, .true.)
        call getr(sf, 'ustokes', rec, su, size(su))
        call getr(sf, 'vstokes', rec, sv, size(sv))
        rc = rp_set_hydro_slot(ctx, int(ng_, c_int32_t), int(slot, c_int32_t), zeta, u, v, w, aks, &
                               c_loc(su), c_loc(sv))
        deallocate (su, sv)
      else
        rc = rp_set_hydro_slot(ctx, int(ng_, c_int32_t), int(slot, c_int32_t), zeta, u, v, w, aks, &
                               c_null_ptr, c_null_ptr)
      end if
      if (rc /= 0) call die('rp_set_hydro_slot', rc)
      if (SaltTempOn) then
        rc = rp_set_hydro_tracers(ctx, int(ng_, c_int32_t), int(slot, c_int32_t), salt, temp)
        if (rc /= 0) call die('rp_set_hydro_tracers', rc)
      end if
      if (WriteBottom) then
        rc = rp_set_hydro_bstress(ctx, int(ng_, c_int32_t), int(slot, c_int32_t), bu, bv)
        if (rc /= 0) call die('rp_set_hydro_bstress', rc)
      end if
      if (use_wetdry) then
        allocate (wdr(L * M), wdu((L - 1) * M), wdv(L * (M - 1)))
        call getr(f, 'wetdry_mask_rho', rec, wdr, size(wdr))
        call getr(f, 'wetdry_mask_u', rec, wdu, size(wdu))
        call getr(f, 'wetdry_mask_v', rec, wdv, size(wdv))
        rc = rp_set_hydro_wetdry(ctx, int(ng_, c_int32_t), int(slot, c_int32_t), wdr, wdu, wdv)
        if (rc /= 0) call die('rp_set_hydro_wetdry', rc)
        deallocate (wdr, wdu, wdv)
      end if
      deallocate (zeta, u, v, w, aks, salt, temp, bu, bv)
    end do
  end subroutine load_slot

  ! one field of a record, or the namelist's constant when read<name> is off
  subroutine field(f, name, doread, const, rec, a)
    type(c_ptr), intent(in) :: f
    character(len=*), intent(in) :: name
    logical, intent(in) :: doread
    double precision, intent(in) :: const
    integer, intent(in) :: rec
    double precision, intent(out) :: a(:)
    if (doread) then
      call getr(f, name, rec, a, size(a))
    else
      a = const
    end if
  end subroutine field

  ! ---- createNetCDF (ROMSPath.f90:1140-1868) with initNetCDF's counters ----
  subroutine create_netcdf()
    if (outpathGiven) then
      if (NCtime == 0) then
        ncfile = trim(outpath) // trim(NCOutFile) // '.nc'
      else
        write (ncfile, '(A,A,A,I3.3,A)') trim(outpath), trim(NCOutFile), '_', 1, '.nc'
      end if
    else
      if (NCtime == 0) then
        ncfile = trim(NCOutFile) // '.nc'
      else
        write (ncfile, '(A,A,I3.3,A)') trim(NCOutFile), '_', 1, '.nc'
      end if
    end if
    write (*, *) 'Creating NetCDF Output File: ', trim(ncfile)
    c_units = trim(time_units) // c_null_char
    c_run = trim(RunName) // c_null_char
    c_exe = trim(ExeDir(1:100)) // c_null_char
    c_out = trim(OutDir(1:100)) // c_null_char
    c_by = trim(RunBy) // c_null_char
    c_inst = trim(Institution) // c_null_char
    meta%time_units = c_loc(c_units)
    meta%run_name = c_loc(c_run)
    meta%idt = idt
    meta%days = days
    meta%behavior = Behavior
    meta%open_boundary = merge(1, 0, OpenOceanBoundary)
    meta%vturb_on = merge(1, 0, VTurbOn)
    meta%hturb_on = merge(1, 0, HTurbOn)
    meta%seed = seed
    meta%deltat = deltat; meta%serr = serr; meta%smth = smth; meta%sub = sub; meta%aksback = AKSback
    meta%swimfast = swimfast
    meta%temp_offset = TempOffset
    meta%salt_temp_mean = merge(1, 0, SaltTempMean)
    meta%executable = c_loc(c_exe)
    meta%output_loc = c_loc(c_out)
    meta%administrator = c_loc(c_by)
    meta%institution = c_loc(c_inst)
    nflags = 0
    if (WriteWaterDepth) nflags = ior(nflags, RP_NC_WATERDEPTH)
    if (WriteZeta) nflags = ior(nflags, RP_NC_ZETA)
    if (WriteBath) nflags = ior(nflags, RP_NC_BATH)
    if (SaltTempOn) nflags = ior(nflags, RP_NC_SALTTEMP)
    if (WriteBottom) nflags = ior(nflags, RP_NC_BOTTOM)
    if (TrackCollisions) nflags = ior(nflags, RP_NC_COLLISIONS)
    rc = rp_nc_create(nc, trim(ncfile) // c_null_char, nl, int(Ngrid, c_int32_t), &
                      int(nflags, c_int32_t), meta, pdob)
    if (rc /= 0) call die('rp_nc_create', rc)
  end subroutine create_netcdf

  ! ---- dataOutput / writeNetCDF (ROMSPath.f90:1012-1025, 1870-2157) ----
  ! the release record (release) writes the release file's lon/lat
  subroutine data_output(time, release)
    integer, intent(in) :: time
    logical, intent(in) :: release
    type(c_ptr) :: cwd, czeta, cbath, csalt, ctemp, chob, cbus, cbvs
    rc = rp_get_particles(ctx, nl, px, py, pz, page, pst, pgid, pin, wd, pzeta, bath)
    if (rc /= 0) call die('rp_get_particles', rc)
    if (release) then
      plon = rlon
      plat = rlat
    else
      rc = rp_sample_lonlat(ctx, nl, plon, plat)
      if (rc /= 0) call die('rp_sample_lonlat', rc)
    end if
    cwd = c_null_ptr; czeta = c_null_ptr; cbath = c_null_ptr
    csalt = c_null_ptr; ctemp = c_null_ptr; chob = c_null_ptr; cbus = c_null_ptr; cbvs = c_null_ptr
    if (WriteWaterDepth) cwd = c_loc(wd)
    if (WriteZeta) czeta = c_loc(pzeta)
    if (WriteBath) cbath = c_loc(bath)
    if (SaltTempOn) then
      ! the means are written and set back to 0 (ROMSPath.f90:2009-2035)
      rc = rp_tracers_get(ctx, nl, psalt, ptemp, 1_c_int32_t)
      if (rc /= 0) call die('rp_tracers_get', rc)
      csalt = c_loc(psalt); ctemp = c_loc(ptemp)
    end if
    if (WriteBottom) then
      ! P_HOB goes back to 9999 once written (ROMSPath.f90:2063-2065)
      rc = rp_bottom_get(ctx, nl, phob, pbus, pbvs, 1_c_int32_t)
      if (rc /= 0) call die('rp_bottom_get', rc)
      chob = c_loc(phob); cbus = c_loc(pbus); cbvs = c_loc(pbvs)
    end if
    rc = rp_nc_write(nc, dble(time) + reftime, page, plon, plat, pz, pst, pgid, cwd, czeta, cbath, &
                     csalt, ctemp, chob, cbus, cbvs)
    if (rc /= 0) call die('rp_nc_write', rc)
  end subroutine data_output

  subroutine fatal(msg)
    character(len=*), intent(in) :: msg
    write (*, '(A)') trim(msg)
    stop 1
  end subroutine fatal

  subroutine die(what, code)
    character(len=*), intent(in) :: what
    integer, intent(in) :: code
    write (*, '(A,A,I0,A,A)') what, ' failed: ', code, ' ', rp_error_text(ctx)
    stop 1
  end subroutine die

end program rp_driver
