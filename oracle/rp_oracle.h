/*
 * rp_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64, single-threaded restatement of the reference ROMSPath
 * per-particle time step (ROMSPath.f90:608-1008 and everything it calls).
 * It exists to check the HIP implementation and to serve as the timed CPU
 * baseline.  Nothing in the product library links, loads or calls it.
 *
 * Conventions mirror the reference Fortran:
 *   - positions X, Y are fractional 1-based rho indices (rho i sits at X=i);
 *   - 2-D arrays are Fortran column-major (xi fastest), 3-D arrays are
 *     (xi, eta, s) column-major, one pointer per hydro slot 1..3;
 *   - slots t_b, t_c, t_f are 1-based like the reference's.
 */
#ifndef RP_ORACLE_H
#define RP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OG_MAXGRID 5

typedef struct {
    int L, M, N;                 /* xi_rho, eta_rho, s_rho (s_w = N+1) */
    int vtransform;
    double hc;
    const double *s_rho, *cs_r;  /* N   */
    const double *s_w, *cs_w;    /* N+1 */
    const double *h, *pm, *pn;   /* L*M */
    const int32_t *mask_rho;     /* L*M */
    const int32_t *mask_u;       /* (L-1)*M */
    const int32_t *mask_v;       /* L*(M-1) */
    double scl[OG_MAXGRID][2];   /* scl(i,:) toward grid i (grid_module.f90:437-481) */
    double off[OG_MAXGRID][2];
    const double *zeta[3];       /* L*M per slot */
    const double *u[3];          /* (L-1)*M*N */
    const double *v[3];          /* L*(M-1)*N */
    const double *w[3];          /* L*M*(N+1) */
    const double *aks[3];        /* L*M*(N+1) */
    const double *salt[3];       /* L*M*N, may be NULL (SaltTempOn off) */
    const double *temp[3];       /* L*M*N */
    const double *bustr[3];      /* (L-1)*M, may be NULL (WriteBottom off) */
    const double *bvstr[3];      /* L*(M-1) */
    /* The reference built with -DWETDRY: wet/dry masks of each slot
     * (hydrodynamic_module.f90:510-570) in place of the static masks in
     * getInterp3d's u, v, w, salt, temp and getInterpStr's bustr, bvstr.
     * NULL: the static masks (the default build). */
    const int32_t *wd_rho[3];    /* L*M */
    const int32_t *wd_u[3];      /* (L-1)*M */
    const int32_t *wd_v[3];      /* L*(M-1) */
} og_grid;

typedef struct {
    int idt;
    double deltat, serr, smth, sub, aksback;
    int vturb_on, hturb_on;
    double kh[OG_MAXGRID];       /* ConstantHTurb(ng) */
    int scheme, nsb;
    double vertdist;
    int open_boundary;
    int ngrid;
} og_params;

typedef struct {
    double ex[3], ix[3];
    int ets;
    int tb, tc, tf;              /* 1-based slot indices */
} og_time;

typedef struct {
    double *x, *y, *z, *age, *dob, *status, *wd, *zeta, *bath;
    int32_t *gid;                /* 1-based */
    uint8_t *isin;
} og_particles;

/* ---- unit-level restatements (each cites its reference routine) ---- */
double og_polintd(const double *xa, const double *ya, double x);
void   og_linint(const double *xa, const double *ya, int n, double x, double *y, double *m);
double og_getslevel(const og_grid *g, double zeta, double depth, int k);
double og_getwlevel(const og_grid *g, double zeta, double depth, int k);
void   og_getcoeff(double X, double Y, const double m[4], double cff[4]);
double og_interp2d(const og_grid *g, int var, int slot, double X, double Y);
double og_interp3d(const og_grid *g, int var, double X, double Y, double Z, int slot,
                   int ztype, double zeta, double depth);
double og_interp_aks(const og_grid *g, double X, double Y, int slot, int k);
int    og_inside(int nb, const double *xb, const double *yb, double xo, double yo);
void   og_bounds(const og_grid *g, double X, double Y, int *nmask, int *ingrid, int *obound);
void   og_find_currents(const og_grid *g, const og_time *T, double X, double Y, double *Z,
                        double *U, double *V, double *W, double *tdepth, double *zeta);
void   og_rkadvect(const og_grid *g, const og_params *P, const og_time *T, double X, double Y,
                   double *Z, double pm, double pn, double *ax, double *ay, double *az);
double og_smooth(const double *x, const double *y, const double *dy, int np, double s,
                 double *v, double *a);
double og_ppvalu(const double *brk, const double *coef, int l, int k, double x, int jderiv);
int    og_interv(const double *xt, int lxt, double x, int *mflag);
double og_vturb(const og_grid *g, const og_params *P, const og_time *T, double X, double Y,
                double Zpar, const double *dev);

/* ---- full step (ROMSPath.f90:662-1003) ----
 * dev: injected normal deviates, [n][ndev] with ndev = 2 + idt/int(deltat);
 * columns 0,1 feed HTurb (x then y), columns 2.. feed VTurb's sub-steps.
 * Returns the number of particles that took the step (the active count). */
void og_set_threads(int n);
double og_interp_str(const og_grid *g, int var, int slot, double Ipar, double Jpar);
void og_bottom_update(const og_grid *grids, const og_time *T, int64_t n, const double *x,
                      const double *y, const double *z, const int32_t *gid, const double *age,
                      const double *bath, double *last_age, double *hob, double *bustr,
                      double *bvstr);
void og_sample_tracers(const og_grid *grids, const og_time *T, int64_t n, const double *x,
                       const double *y, const double *z, const int32_t *gid, const double *bath,
                       double *salt, double *temp);
void og_ll2ij(const double *lon, const double *lat, const double *angle, int L, int M,
              int spherical, double inlon, double inlat, double *Ipar, double *Jpar);
void og_zbounds(const og_grid *g, double Ipar, double Jpar, double *Zpar, int *inz, int t);
void og_locate(const og_grid *grids, int ngrid, const double *const *lon, const double *const *lat,
               const double *const *angle, int spherical, int tb, int64_t n, const double *inlon,
               const double *inlat, double *z, double *x, double *y, double *status, int32_t *gid,
               uint8_t *isin);
int64_t og_update_particles(const og_grid *grids, const og_params *P, const og_time *T,
                            og_particles *p, int64_t n, const double *dev, int ndev);

#ifdef __cplusplus
}
#endif
#endif
