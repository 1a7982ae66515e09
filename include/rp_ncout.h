/*
 * rp_ncout.h -- trajectory output of librp.so: the reference's netCDF file
 * (createNetCDF / writeNetCDF, ROMSPath.f90:1140-2157), written by the
 * library itself because no netCDF library is present.
 *
 * The file is netCDF classic with 64-bit offsets (CDF-2), readable by every
 * netCDF-3/-4 reader.  The reference creates NF90_NETCDF4 (HDF5) files with
 * deflate; the schema is the same: dimensions numpar and time (unlimited),
 * and in the reference's order the variables
 *   model_time(time)            double  "Model time", units = time_units
 *   dob(numpar)                 float   date of birth (s), written at create
 *   age(time, numpar)           double  par(:,pAge)
 *   lon, lat(time, numpar)      float   getInterp2D("lon"/"lat") at the particle
 *   zp(time, numpar)            float   par(:,pZ)
 *   status(time, numpar)        float   par(:,pStatus)
 *   GID(time, numpar)           float   par(:,pGID), only when Ngrid > 1
 *   hitBottom, hitLand          float   0, when asked for (TrackCollisions)
 *   salinity, temperature       float   P_salt, P_temp, when asked for (SaltTempOn)
 *   HOB, bustr, bvstr           float   P_HOB, P_bustr, P_bvstr, when asked for (WriteBottom)
 *   WaterDepth, zeta, h         float   par(:,pWD/pZeta/pBath), when asked for
 * with the reference's long_name / units / field attributes and the global
 * attributes of createNetCDF that apply to the hot path.  Values are rounded
 * to float exactly as nf90_put_var converts double arrays into NF_FLOAT
 * variables (round to nearest).
 *
 * Host code only: nothing here touches the GPU.  Arrays are in release
 * order (as rp_get_particles returns them).  Every function returns 0 or a
 * negative RP_E* code (rp_abi.h); rp_nc_error() gives the message.
 */
#ifndef RP_NCOUT_H
#define RP_NCOUT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Optional variables: WriteWaterDepth, WriteZeta, WriteBath, SaltTempOn,
 * WriteBottom, TrackCollisions (hitBottom / hitLand, written as zeros: the
 * reference never increments them, advection_module.f90:173). */
enum { RP_NC_WATERDEPTH = 1, RP_NC_ZETA = 2, RP_NC_BATH = 4, RP_NC_SALTTEMP = 8, RP_NC_BOTTOM = 16,
       RP_NC_COLLISIONS = 32 };

/* Global attributes (createNetCDF, ROMSPath.f90:1676-1836), written in the
 * reference's order: title, author, adv_time_step, days_run, Behavior,
 * OpenOceanBoundary (only when set), swimfast, Vertical_Turb (+ seed,
 * turb_time_step, spline_*, subset_scale, AKs_bakcground when on),
 * Horiz_Turb (+ seed), run_name, executable, output_loc, administrator,
 * institution, Start_datetime (local wall clock at creation, the
 * reference's "YYYY/MM/DD hh:mm:ss"), SaltTempMean, TempOffset.  String
 * fields may be NULL: run_name is then left out, the others are written
 * empty.  The behaviour-10 / Process_VA and GROWTH attributes belong to
 * modules outside the particle step and are not written. */
typedef struct {
    const char *time_units;      /* model_time units (GRID_MOD time_units) */
    const char *run_name;        /* RunName, may be NULL */
    int32_t idt;                 /* adv_time_step */
    float days;                  /* days_run */
    int32_t behavior;            /* Behavior */
    int32_t open_boundary;       /* OpenOceanBoundary -> "TRUE" when set */
    int32_t vturb_on, hturb_on;  /* Vertical_Turb / Horiz_Turb */
    int32_t seed;
    double deltat, serr, smth, sub, aksback;
    double swimfast;             /* $behavparam swimfast */
    double temp_offset;          /* $other TempOffset */
    int32_t salt_temp_mean;      /* $other SaltTempMean -> "TRUE" / "FALSE" */
    const char *executable;      /* $output ExeDir */
    const char *output_loc;      /* $output OutDir */
    const char *administrator;   /* $output RunBy */
    const char *institution;     /* $output Institution */
} rp_nc_meta;

/* createNetCDF(par(:,pDOB)) (ROMSPath.f90:449, 1140-1868): creates the file
 * for numpar particles and writes dob.  flags: RP_NC_* optional variables;
 * GID is defined when ngrid > 1. */
int rp_nc_create(void **nc, const char *path, int64_t numpar, int32_t ngrid, int32_t flags,
                 const rp_nc_meta *meta, const double *dob);

/* writeNetCDF (ROMSPath.f90:1870-2157): appends one time record.
 * model_time = DBLE(time) + reftime.  gid, wd, zeta, bath, salt, temp, hob,
 * bustr and bvstr may be NULL when the file has no such variable.  The
 * reference resets P_HOB to 9999 after writing it (ROMSPath.f90:2063-2065);
 * here that is rp_bottom_get's reset flag (rp_abi.h). */
int rp_nc_write(void *nc, double model_time, const double *age, const double *lon,
                const double *lat, const double *zp, const double *status, const int32_t *gid,
                const double *wd, const double *zeta, const double *bath, const double *salt,
                const double *temp, const double *hob, const double *bustr, const double *bvstr);

/* Records written so far. */
int64_t rp_nc_records(void *nc);

/* Message of the last failure of this writer. */
const char *rp_nc_error(void *nc);

/* Closes the file and frees the writer. */
int rp_nc_close(void *nc);

#ifdef __cplusplus
}
#endif

#endif /* RP_NCOUT_H */
