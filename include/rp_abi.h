/*
 * rp_abi.h -- C-ABI of librp.so, the MI355X implementation of ROMSPath's
 * per-particle time step.
 *
 * The reference has no plugin or FFI layer: the replaced region is the body
 * of update_particles (ROMSPath.f90:608-1008), called once per internal step
 * from run_Internal_Timestep (ROMSPath.f90:529), which communicates through
 * module globals (par, isIn, GRIDS, HYDRODATA, PARAM_MOD, t_b/t_c/t_f,
 * ex/ix/ets).  Each entry point below replaces one of those hand-over points;
 * its comment cites the reference code it stands in for.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative RP_E* code on
 *     failure; rp_last_error() gives the message.  Nothing aborts the host.
 *   - Arrays are passed exactly as the reference holds them: FP64, Fortran
 *     column-major (xi fastest), positions as fractional 1-based rho indices.
 *   - The caller owns all host arrays (copied in and out); the library owns
 *     device memory.  One context = one HIP stream = one host thread.
 *   - Grid numbers ng and hydro slots are 1-based, as in the reference.
 */
#ifndef RP_ABI_H
#define RP_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RP_MAXGRID 5

enum {
    RP_OK = 0,
    RP_E_ARG = -1,        /* invalid argument */
    RP_E_STATE = -2,      /* call out of order (grid/slot/particles missing) */
    RP_E_HIP = -3,        /* HIP runtime error */
    RP_E_ALLOC = -4,      /* device allocation failed */
    RP_E_RANGE = -5,      /* size beyond a limit */
    RP_E_NODEV = -6,      /* no usable GPU */
    RP_E_IO = -7          /* file could not be written */
};

enum { RP_PREC_FP32_FIELDS = 0, RP_PREC_FP64 = 1 };

/* Namelist scalars on the hot path (ROMSPath.h:77-258) plus build knobs.
 * Stand-in for PARAM_MOD after getParams (parameter_module.f90:22-152). */
typedef struct {
    int32_t idt;                 /* timeparam: internal step (s) */
    int32_t dt;                  /* timeparam: external step (s) */
    double deltat;               /* turbparam: RDM sub-step (s) */
    double serr, smth, sub;      /* turbparam: spline error, smoothing, resolution */
    double aksback;              /* turbparam: AKs background */
    int32_t vturb_on, hturb_on;  /* turbparam */
    double kh[RP_MAXGRID];       /* turbparam: ConstantHTurb(ng) */
    int32_t scheme, nsb;         /* advectparam */
    double vertdist;             /* advectparam */
    int32_t open_boundary;       /* behavparam: OpenOceanBoundary */
    int32_t ngrid;               /* romsgrid: Ngrid */
    uint64_t seed;               /* seed of the counter-based RNG */
    /* build knobs (no namelist counterpart) */
    int32_t precision;           /* RP_PREC_FP32_FIELDS or RP_PREC_FP64 */
    int32_t block;               /* threads per block, 0 = default */
    int32_t sort_interval;       /* re-sort particles by cell every k steps, 0 = never */
    int32_t reserved[5];
} rp_config;

/* One grid, as InitGrid leaves GRIDDATA (grid_module.f90:14-36, 41-481). */
typedef struct {
    int32_t L, M, N;             /* xi_rho, eta_rho, s_rho */
    int32_t vtransform;
    double hc;
    const double *s_rho, *cs_r;  /* N   */
    const double *s_w, *cs_w;    /* N+1 */
    const double *h, *pm, *pn;   /* (xi_rho, eta_rho) */
    const int32_t *mask_rho;     /* (xi_rho, eta_rho) */
    const int32_t *mask_u;       /* (xi_u, eta_u) */
    const int32_t *mask_v;       /* (xi_v, eta_v) */
    double scl[RP_MAXGRID][2];   /* scl(i,:), off(i,:): affine map to grid i */
    double off[RP_MAXGRID][2];
} rp_grid;

/* after getParams: replaces PARAM_MOD's hot-path globals */
int rp_create(void **ctx, const rp_config *cfg);

/* after InitGrid (grid_module.f90:41-499): one call per grid ng=1..Ngrid */
int rp_set_grid(void *ctx, int32_t ng, const rp_grid *g);

/* end of updateHydro (hydrodynamic_module.f90:112-760), which fills time
 * slot tind=1..3 of HYDRODATA(ng).  zeta (xi_rho,eta_rho); u (xi_u,eta_u,s_rho);
 * v (xi_v,eta_v,s_rho); w and aks (xi_rho,eta_rho,s_w).  su/sv (Stokes drift,
 * may be NULL) are added to u/v in FP64 before storage, as under -DSTOKES
 * (hydrodynamic_module.f90:608-657).  Asynchronous on the context stream; the
 * host arrays may be reused when the call returns.  The arrays are copied to
 * pinned host memory before the call returns; the host-to-device copies run on
 * a second stream, overlapping the steps already queued, and the FP64 -> field
 * precision conversion runs on the context stream after those steps. */
int rp_set_hydro_slot(void *ctx, int32_t ng, int32_t slot, const double *zeta, const double *u,
                      const double *v, const double *w, const double *aks, const double *su,
                      const double *sv);

/* end of ini_ROMSPath (ROMSPath.f90:319-433): the par(:,k) columns for
 * X, Y, Z, Age, DOB, Status, GID and isIn of n particles. */
int rp_set_particles(void *ctx, int64_t n, const double *x, const double *y, const double *z,
                     const double *age, const double *dob, const double *status,
                     const int32_t *gid, const uint8_t *isin);

/* replaces call update_particles() (ROMSPath.f90:529): enqueues one internal
 * step for all particles.  ex/ix and the slot numbers are the host schedule's
 * values (ROMSPath.f90:488-524); step_index counts internal steps from 0 and
 * keys the random deviates.  want_diag != 0 also stores the water depth,
 * zeta and bathymetry columns (pWD, pZeta, pBath; ROMSPath.f90:845-847),
 * which only the output reads.  While the WriteBottom or SaltTempOn sampler
 * is active (rp_bottom_begin, rp_tracers_begin) pBath, the step's own
 * -tdepth, is stored on every step, since those samplers read it.
 * Asynchronous. */
int rp_step(void *ctx, const double ex[3], const double ix[3], int32_t ets, int32_t t_b,
            int32_t t_c, int32_t t_f, int64_t step_index, int32_t want_diag);

/* No counterpart in the reference.  rp_step for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments, same results, and no
 * hipMalloc, hipFree or allocating library call.  Its scratch (the hand-over
 * rows between the step's two kernels, the vertical walk's work-queue
 * counter and per-lane work rows) lives in one 64 MiB device array that
 * every GPU the library runs on holds until the process ends, instead of
 * the context's own buffers.  Particles are stepped in launch pairs of at
 * most 131,072; the array's 60 MiB row region bounds the vertical walk's
 * lanes (fewer resident lanes than rp_step uses), and a spline whose work
 * rows for one block of lanes exceed it returns RP_E_ARG.  No re-sort
 * happens (results never depend on the order), the step is not timed, and a
 * recorded run (rp_adjoint_record) returns RP_E_STATE.  All calls on one
 * GPU share the array: calls that overlap on one GPU (other contexts,
 * other streams) race.  Asynchronous. */
int rp_step_noalloc(void *ctx, const double ex[3], const double ix[3], int32_t ets, int32_t t_b,
                    int32_t t_c, int32_t t_f, int64_t step_index, int32_t want_diag);

/* dataOutput (ROMSPath.f90:1012-1072): copies the particle columns back in
 * release order.  Any pointer may be NULL.  Synchronises the stream. */
int rp_get_particles(void *ctx, int64_t n, double *x, double *y, double *z, double *age,
                     double *status, int32_t *gid, uint8_t *isin, double *wd, double *zeta,
                     double *bath);

/* Trajectory output support (dataOutput, ROMSPath.f90:1012-1025).
 * rp_set_lonlat: GRIDS(ng)%lon_rho / lat_rho, (xi_rho, eta_rho) FP64, which
 * InitGrid reads from lon_rho/lat_rho, or x_rho/y_rho on Cartesian grids
 * (grid_module.f90:266-300).  rp_sample_lonlat: getInterp2D("lon"/"lat")
 * (interpolation_module.f90:421-507) at every particle's X, Y on its grid,
 * in release order; synchronises the stream.  The corners are clamped into
 * the grid as in the step (the reference would index outside its arrays). */
int rp_set_lonlat(void *ctx, int32_t ng, const double *lon, const double *lat);
int rp_sample_lonlat(void *ctx, int64_t n, double *lon, double *lat);

/* The reference built with -DWETDRY (makefile:12-13; off by default):
 * updateHydro reads wetdry_mask_rho (xi_rho, eta_rho), wetdry_mask_u
 * (xi_u, eta_u) and wetdry_mask_v (xi_v, eta_v) of each record into its slot
 * (hydrodynamic_module.f90:510-570), and getInterp3d's u, v, w, salt, temp
 * and getInterpStr's bustr, bvstr use the slot's masks in place of the grid's
 * static masks (interpolation_module.f90:628-923, 1022-1140); getInterpAKs
 * keeps the static mask.  rp_set_hydro_wetdry stores the three masks (FP64,
 * each value 0 or 1, else RP_E_ARG) of time slot 1..3 of grid ng, after
 * waiting for the steps already queued; from the first call on, the grid's
 * gathers use the masks, and slots not given masks are dry (the reference
 * allocates them as 0).  In the FP32-field form the current u and v fields
 * take each slot's value only where that slot's mask is wet.  w uses the
 * current field where the step's slots agree on a cell's four corners, and
 * weighs the FP32 slot fields slot by slot where they differ. */
int rp_set_hydro_wetdry(void *ctx, int32_t ng, int32_t slot, const double *mask_rho,
                        const double *mask_u, const double *mask_v);

/* SaltTempOn (ROMSPath.f90:958-981): rp_set_hydro_tracers stores salt and
 * temp (xi_rho, eta_rho, s_rho) of time slot 1..3, as updateHydro reads them
 * (hydrodynamic_module.f90:663-745); asynchronous like rp_set_hydro_slot.
 * rp_sample_tracers gives salt and temp at every particle's current
 * position (release order): zeta of slots b, c, f, getInterp3d("salt"/"temp")
 * on rho levels with each slot's zeta (interpolation_module.f90:509-589,
 * 769-830), polintd in time at ix(2).  Synchronises the stream.  The
 * reference's P_salt keeps a particle's value from the last step it took;
 * rp_tracers_* below keep that state. */
int rp_set_hydro_tracers(void *ctx, int32_t ng, int32_t slot, const double *salt,
                         const double *temp);
int rp_sample_tracers(void *ctx, int64_t n, const double ex[3], const double ix[3], int32_t t_b,
                      int32_t t_c, int32_t t_f, double *salt, double *temp);

/* P_Salt / P_Temp kept as the reference keeps them, per step.
 * rp_tracers_begin (after rp_set_particles and rp_locate_particles' isIn,
 * with slot t_c's salt/temp and hydro loaded; ROMSPath.f90:247-258,
 * 408-424): P = 0, then for the particles in a grid P = getInterp3d at t_c;
 * with mean set (SaltTempMean) also mean = P and mI = 1 when any particle
 * is in a grid.  rp_tracers_update (after each rp_step, with that step's
 * ex, ix and slots; queued on the stream; ROMSPath.f90:958-981, 533): for
 * each particle that took the step (its age changed) P = the sample
 * rp_sample_tracers would give, mean += P; then mI += 1.
 * rp_tracers_get: P (or, with mean, mean / mI as writeNetCDF writes it,
 * ROMSPath.f90:2010-2014) in release order; reset_mean puts the means and mI
 * back to 0 (ROMSPath.f90:2030-2035).  Synchronises the stream. */
int rp_tracers_begin(void *ctx, int32_t mean, int32_t t_c);
int rp_tracers_update(void *ctx, const double ex[3], const double ix[3], int32_t t_b, int32_t t_c,
                      int32_t t_f);
int rp_tracers_get(void *ctx, int64_t n, double *salt, double *temp, int32_t reset_mean);

/* WriteBottom (ROMSPath.f90:260-266, 353-355, 935-951, 2056-2078).
 * rp_set_hydro_bstress stores bustr (xi_u, eta_u) and bvstr (xi_v, eta_v)
 * of time slot 1..3; asynchronous like rp_set_hydro_slot.
 * rp_bottom_begin (after rp_set_particles, before the first step): P_HOB =
 * 9999, P_bustr = P_bvstr = 0 for every particle.
 * rp_bottom_update (after each rp_step, with that step's ex, ix and slots;
 * queued on the stream): for each particle that took the step (its age
 * changed), btemp = -tdepth + Z, P_HOB = min(btemp, P_HOB), and where btemp
 * equals the new P_HOB, P_bustr / P_bvstr = polintd at ix(2) of getInterpStr
 * (interpolation_module.f90:964-1174) over slots t_b, t_c, t_f.
 * rp_bottom_get copies P_HOB, P_bustr, P_bvstr out in release order and,
 * when reset_hob is set, puts P_HOB back to 9999 as writeNetCDF does after
 * writing it; synchronises the stream. */
int rp_set_hydro_bstress(void *ctx, int32_t ng, int32_t slot, const double *bustr,
                         const double *bvstr);
int rp_bottom_begin(void *ctx);
int rp_bottom_update(void *ctx, const double ex[3], const double ix[3], int32_t t_b, int32_t t_c,
                     int32_t t_f);
int rp_bottom_get(void *ctx, int64_t n, double *hob, double *bustr, double *bvstr,
                  int32_t reset_hob);

/* ini_ROMSPath's placement of the released particles (ROMSPath.f90:322-403):
 * for grid ng = 1..Ngrid in order, LL2ij (interpolation_module.f90:110-225:
 * bisection of the index box with inbox/inside, then the Cartesian linear
 * map or the spherical local-triangle map) and bounds; the first grid that
 * holds a particle takes it (X, Y, GID = ng, status 0, isIn), and zbounds
 * (boundary_module.f90:145-179, zeta of slot t_b) may lift z to zeta - 0.01
 * or set status 9 again.  Others keep X = Y = 1, GID = Ngrid, status 9,
 * isIn 0.  lon, lat, z are the release file's columns; z is updated in
 * place.  Needs rp_set_lonlat for every grid and slot t_b loaded.  The
 * results go to rp_set_particles; synchronises the stream.
 * rp_set_grid_angle: GRIDS(ng)%angle (may be NULL: 0) and spherical
 * (GRIDS(1)%spherical decides for all grids, as in LL2ij). */
int rp_set_grid_angle(void *ctx, int32_t ng, int32_t spherical, const double *angle);
int rp_locate_particles(void *ctx, int64_t n, const double *lon, const double *lat, double *z,
                        int32_t t_b, double *x, double *y, double *status, int32_t *gid,
                        uint8_t *isin);

/* Test hook: use these normal deviates instead of the RNG for the following
 * steps.  normals is [n][ndev] in release order, ndev = 2 + idt/int(deltat):
 * columns 0,1 feed HTurb (x, y), the rest VTurb's sub-steps in order.
 * normals == NULL returns to the RNG. */
int rp_set_deviates(void *ctx, const double *normals, int64_t n, int32_t ndev);

/* Test hook: the RNG's normal deviates for particles [0,n) at step
 * step_index, in the layout of rp_set_deviates. */
int rp_sample_deviates(void *ctx, int64_t step_index, int64_t n, int32_t ndev, double *out);

/* Tuning of the step, no reference counterpart: threads per block (0 =
 * 256, else a multiple of 64 up to 256), re-sort of the stored particles by
 * cell every sort_interval steps (0 = never) and the sort tile in cells
 * (tile_x x tile_y; 1 x 1 sorts by cell).  Results do not depend on any of
 * them.  The next step re-sorts. */
int rp_set_tuning(void *ctx, int32_t block, int32_t sort_interval, int32_t tile_x,
                  int32_t tile_y);

/* Release ensembles, no reference counterpart (the reference runs one
 * release per job): the particles of rp_set_particles, in their order, are
 * nscen consecutive scenarios of counts[s] particles (sum = particle count).
 * Every scenario is still stepped by rp_step / rp_step_batch together; this
 * only enables rp_step_scenario. */
int rp_set_scenarios(void *ctx, int32_t nscen, const int64_t *counts);

/* One internal step of scenario scen (0-based) only: the per-ensemble launch
 * form of rp_step, with the same arguments.  Each scenario keeps its own
 * re-sort count and is sorted among its own particles. */
int rp_step_scenario(void *ctx, int32_t scen, const double ex[3], const double ix[3], int32_t ets,
                     int32_t t_b, int32_t t_c, int32_t t_f, int64_t step_index, int32_t want_diag);

/* nsteps consecutive internal steps of all particles (the rp_step calls of
 * one external step, ROMSPath.f90:515-546, with no hydro load between them):
 * ex, ix [nsteps][3], ets [nsteps], slots [nsteps][3] = (t_b, t_c, t_f),
 * step_index [nsteps], want_diag [nsteps] or NULL.  use_graph = 1 captures
 * the launches of all steps (re-sorts included) into one HIP graph and
 * launches it once; 0 launches them as rp_step does.  No reference
 * counterpart. */
int rp_step_batch(void *ctx, int32_t nsteps, const double *ex, const double *ix,
                  const int32_t *ets, const int32_t *slots, const int64_t *step_index,
                  const int32_t *want_diag, int32_t use_graph);

/* Velocity perturbation, no reference counterpart: du, dv (m/s) added to the
 * interpolated u, v of every RK stage, from uniform cubic B-splines on an
 * nx x ny control net (4..64 each way) spread over each grid's rho-index
 * domain [1, L] x [1, M].  coef [2][ny][nx] (du then dv, x fastest); NULL
 * removes it.  Zero coefficients leave every result unchanged. */
int rp_set_perturbation(void *ctx, int32_t nx, int32_t ny, const double *coef);

/* Adjoint recording, no reference counterpart.  ckpt_every > 0 starts a
 * recorded run at the current particle state: every following rp_step runs
 * the whole step in one kernel with the vertical walk's spline on its
 * reference-order FP64 path (for both field precisions), stores the state
 * every ckpt_every steps, a copy of the hydro slots whenever they changed,
 * and the injected deviates of each step if any; no re-sort happens while
 * recording.  0 ends recording and frees what it kept. */
int rp_adjoint_record(void *ctx, int32_t ckpt_every);

/* Reverse sweep over the recorded steps, no reference counterpart.  Given
 * lam = dL/d(x, y, z) of the particles' current state (release order; x, y
 * in grid-index units of each particle's grid, z in metres), returns
 * g = dL/d(x, y, z) of the state when recording started and, if g_pert is
 * not NULL, dL/dcoef of the perturbation ([2][ny][nx]).  Segments between
 * checkpoints are recomputed forward, then each step is differentiated
 * exactly (forward-mode dual numbers seeded at the step's start) and
 * lambda <- J^T lambda.  The particle state is left as it was. */
int rp_adjoint(void *ctx, int64_t n, const double *lam_x, const double *lam_y,
               const double *lam_z, double *g_x, double *g_y, double *g_z, double *g_pert);

/* Sum over all steps so far of the particles that took the step (those that
 * pass the skip tests of ROMSPath.f90:672-690).  Synchronises the stream. */
int rp_active_steps(void *ctx, int64_t *count);

/* Number of smoothing-spline secant searches stopped at the iteration cap
 * (the reference's search, pppack.f90:5270-5292, has none; a GPU lane must
 * end).  Synchronises the stream. */
int rp_secant_capped(void *ctx, int64_t *count);

/* Wait for all work queued on the context stream. */
int rp_synchronize(void *ctx);

/* Average device time per rp_step (ms) of its kernels since the last call,
 * measured with HIP events on the context stream.  No reference counterpart
 * (the reference has no timers on this path). */
int rp_kernel_timing(void *ctx, double *avg_ms, int64_t *launches);

/* Split of the averages of the last rp_kernel_timing call: the advection
 * kernel (RK stages, horizontal walk; the whole step when vertical
 * turbulence is off) and the vertical-walk kernel (spline fit, sub-steps,
 * grid hand-off).  No reference counterpart. */
int rp_phase_timing(void *ctx, double *advect_ms, double *walk_ms);

/* The eigenbasis tables of the FP32-field vertical walk for n interior
 * unknowns of the resampled AKs column (n = int((N+1)*sub) - 2, even, 64..128):
 * the generalized eigenvectors of chol1d's matrix pencil at unit spacing
 * (pentadiag(1,-4,6,-4,1), tridiag(1,4,1); pppack.f90:1509-1512), split into
 * n/2 symmetric (par 0) and n/2 antisymmetric (par 1) ones.
 * basis[par][k][j] (2 x 64 x 64 floats): component j of eigenvector k over
 * sqrt(2) (zero past n/2); lam[par][k] (2 x 64 doubles): its eigenvalue,
 * ascending (1 past n/2).  Host only, no context.  Returns RP_E_ARG for n
 * outside the range.  No reference counterpart: it replaces the pentadiagonal
 * eliminations of the reference's secant search (pppack.f90:5236-5298) with
 * a change of basis computed once. */
int rp_modal_basis(int32_t n, float *basis, double *lam);

const char *rp_last_error(void *ctx);
int rp_destroy(void *ctx);

#ifdef __cplusplus
}
#endif
#endif
