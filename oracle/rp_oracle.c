/*
 * rp_oracle.c -- TEST INFRASTRUCTURE ONLY (see rp_oracle.h).
 *
 * CPU FP64 restatement of the reference per-particle step.  Every routine
 * follows the reference routine it names, statement for statement, including
 * operation order (compiled with -ffp-contract=off so no FMA is formed) and
 * the precision of Fortran literals: a default-REAL literal such as 0.001 or
 * 0.1 is single precision, so DBLE(0.001) is the float 0.001 widened.
 *
 * Deliberate differences from the reference, each forced by undefined
 * behaviour there:
 *   - dy(1) and dy(np) of the smoothing spline are set to serr
 *     (ver_turb_module.f90:68,131-136 leaves them uninitialised);
 *   - normal deviates are injected by the caller instead of drawn from the
 *     global MT19937 stream (pdf_module.f90:26-40);
 *   - grid reads at the post-step position clamp indices into the grid (the
 *     reference reads out of bounds after a failed nested hand-off,
 *     ROMSPath.f90:837-843);
 *   - the w-level scan guards its lower index (interpolation_module.f90:572
 *     would read s_w(0) if z sat exactly on the bed);
 *   - getInterpStr clamps its node reads into the bustr/bvstr arrays
 *     (interpolation_module.f90:1012-1174 reads outside them only for a
 *     particle on the grid's last row or column).
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include "rp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define F_0P001 ((double)0.001f)   /* DBLE(0.001), advection_module.f90:172,183 */
#define F_0P1   ((double)0.1f)     /* ecut=0.1,   boundary_module.f90:58 */

/* Fortran column-major accessors, 1-based indices */
#define R2(a, i, j)      ((a)[((i) - 1) + (int64_t)g->L * ((j) - 1)])
#define U3(a, i, j, k)   ((a)[((i) - 1) + (int64_t)(g->L - 1) * (((j) - 1) + (int64_t)g->M * ((k) - 1))])
#define V3(a, i, j, k)   ((a)[((i) - 1) + (int64_t)g->L * (((j) - 1) + (int64_t)(g->M - 1) * ((k) - 1))])
#define W3(a, i, j, k)   ((a)[((i) - 1) + (int64_t)g->L * (((j) - 1) + (int64_t)g->M * ((k) - 1))])
#define MU(i, j)         (g->mask_u[((i) - 1) + (int64_t)(g->L - 1) * ((j) - 1)])
#define MV(i, j)         (g->mask_v[((i) - 1) + (int64_t)g->L * ((j) - 1)])

/* interpolation_module.f90:72-109 */
double og_polintd(const double *xa, const double *ya, double x)
{
    int ns = 1;
    double dif = fabs(x - xa[0]);
    for (int i = 1; i <= 3; i++) {
        double dift = fabs(x - xa[i - 1]);
        if (dift < dif) {
            ns = i;
            dif = dift;
        }
    }
    double c = (xa[1] - x) * ((ya[2] - ya[1]) / (xa[1] - xa[2]));
    c = c - (xa[1] - x) * ((ya[1] - ya[0]) / (xa[0] - xa[1]));
    c = c / (xa[0] - xa[2]);
    double a, b;
    if (ns == 3) {
        a = (ya[2] - ya[1]) / (xa[1] - xa[2]);
        b = xa[2] - x;
    } else {
        a = (ya[1] - ya[0]) / (xa[0] - xa[1]);
        b = xa[0] - x;
    }
    return ya[ns - 1] + (xa[ns - 1] - x) * a + b * c;
}

/* interpolation_module.f90:27-61 */
void og_linint(const double *xa, const double *ya, int n, double x, double *y, double *m)
{
    int jlo = 1, jhi = n;
    for (;;) {
        int k = (jhi + jlo) / 2;
        if (xa[k - 1] > x)
            jhi = k;
        else
            jlo = k;
        if (jhi - jlo == 1)
            break;
    }
    *m = (ya[jlo - 1] - ya[jhi - 1]) / (xa[jlo - 1] - xa[jhi - 1]);
    double b = ya[jlo - 1] - *m * xa[jlo - 1];
    *y = *m * x + b;
}

/* grid_module.f90:513-554 */
double og_getslevel(const og_grid *g, double zeta, double depth, int k)
{
    double h = -1.0 * depth, S;
    switch (g->vtransform) {
    case 1:
        S = g->hc * g->s_rho[k - 1] + (h - g->hc) * g->cs_r[k - 1];
        return S + zeta * (1.0 + S / h);
    case 2:
        S = (g->hc * g->s_rho[k - 1] + h * g->cs_r[k - 1]) / (g->hc + h);
        return zeta + (zeta + h) * S;
    default:
        return zeta * (1.0 + g->s_rho[k - 1]) + g->hc * g->s_rho[k - 1] + (h - g->hc) * g->cs_r[k - 1];
    }
}

/* grid_module.f90:557-598 */
double og_getwlevel(const og_grid *g, double zeta, double depth, int k)
{
    double h = -1.0 * depth, S;
    switch (g->vtransform) {
    case 1:
        S = g->hc * g->s_w[k - 1] + (h - g->hc) * g->cs_w[k - 1];
        return S + zeta * (1.0 + S / h);
    case 2:
        S = (g->hc * g->s_w[k - 1] + h * g->cs_w[k - 1]) / (g->hc + h);
        return zeta + (zeta + h) * S;
    default:
        return zeta * (1.0 + g->s_w[k - 1]) + g->hc * g->s_w[k - 1] + (h - g->hc) * g->cs_w[k - 1];
    }
}

/* interpolation_module.f90:1230-1273 */
void og_getcoeff(double X, double Y, const double m[4], double cff[4])
{
    double dist[4], W[4], WT;
    dist[0] = sqrt(X * X + Y * Y);
    W[0] = m[0] / dist[0];
    dist[1] = sqrt((1 - X) * (1 - X) + Y * Y);
    W[1] = m[1] / dist[1];
    dist[2] = sqrt((1 - X) * (1 - X) + (1 - Y) * (1 - Y));
    W[2] = m[2] / dist[2];
    dist[3] = sqrt(X * X + (1 - Y) * (1 - Y));
    W[3] = m[3] / dist[3];
    WT = W[0] + W[1] + W[2] + W[3];
    for (int i = 0; i < 4; i++)
        cff[i] = (WT == 0.0) ? 0.0 : W[i] / WT;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* interpolation_module.f90:421-507.  var: 0 depth, 1 pm, 2 pn, 3 zeta */
double og_interp2d(const og_grid *g, int var, int slot, double X, double Y)
{
    int I = (int)floor(X), J = (int)floor(Y);
    double x = X - I, y = Y - J;
    /* In-range positions are untouched by the clamp; it only keeps the
     * post-step reads of a failed nested hand-off inside the arrays. */
    int I0 = clampi(I, 1, g->L), I1 = clampi(I + 1, 1, g->L);
    int J0 = clampi(J, 1, g->M), J1 = clampi(J + 1, 1, g->M);
    double m[4], cff[4], v[4];
    m[0] = R2(g->mask_rho, I0, J0);
    m[1] = R2(g->mask_rho, I1, J0);
    m[2] = R2(g->mask_rho, I1, J1);
    m[3] = R2(g->mask_rho, I0, J1);
    cff[0] = (1 - x) * (1 - y);
    cff[1] = x * (1 - y);
    cff[2] = x * y;
    cff[3] = (1 - x) * y;
    const double *f = var == 0 ? g->h : var == 1 ? g->pm : var == 2 ? g->pn : g->zeta[slot - 1];
    v[0] = R2(f, I0, J0);
    v[1] = R2(f, I1, J0);
    v[2] = R2(f, I1, J1);
    v[3] = R2(f, I0, J1);
    if (var == 3)
        og_getcoeff(x, y, m, cff);
    return cff[0] * v[0] + cff[1] * v[1] + cff[2] * v[2] + cff[3] * v[3];
}

/* interpolation_module.f90:509-962 for var u (0), v (1), w (2) */
double og_interp3d(const og_grid *g, int var, double X, double Y, double Z, int slot,
                   int ztype, double zeta, double depth)
{
    int nz, itop, ibot, i;
    double zlev[2] = {0.0, 0.0}, zcff[2], Zr;
    const int s = slot - 1;
    /* -DWETDRY: this slot's wet/dry masks replace the static ones
     * (interpolation_module.f90:628-665, 710-747, 774-923) */
    const int32_t *mk_u = g->wd_u[s] ? g->wd_u[s] : g->mask_u;
    const int32_t *mk_v = g->wd_v[s] ? g->wd_v[s] : g->mask_v;
    const int32_t *mk_r = g->wd_rho[s] ? g->wd_rho[s] : g->mask_rho;
#define MKU(i, j) (mk_u[((i) - 1) + (int64_t)(g->L - 1) * ((j) - 1)])
#define MKV(i, j) (mk_v[((i) - 1) + (int64_t)g->L * ((j) - 1)])
    if (ztype == 1) {
        nz = g->N;
        itop = nz;
        ibot = nz;
        for (i = 1; i <= nz; i++) {
            zlev[1] = og_getslevel(g, zeta, depth, i);
            if (zlev[1] > Z) {
                itop = i;
                ibot = i - 1;
                zlev[0] = (i == 1) ? depth : og_getslevel(g, zeta, depth, i - 1);
                break;
            }
        }
        if (i > nz) {
            itop = nz;
            ibot = nz;
            zcff[0] = 0.5;
            zcff[1] = 0.5;
        } else if (itop == 1) {
            itop = 1;
            ibot = 1;
            zcff[0] = 0.5;
            zcff[1] = 0.5;
        } else {
            Zr = (Z - zlev[0]) / (zlev[1] - zlev[0]);
            zcff[0] = 1.0 - Zr;
            zcff[1] = Zr;
        }
    } else {
        nz = g->N + 1;
        itop = nz;
        ibot = nz;
        int found = 0;
        for (i = 1; i <= nz; i++) {
            zlev[1] = og_getwlevel(g, zeta, depth, i);
            if (zlev[1] > Z) {
                zlev[0] = og_getwlevel(g, zeta, depth, i > 1 ? i - 1 : 1);
                itop = i;
                ibot = i > 1 ? i - 1 : 1;
                found = 1;
                break;
            }
        }
        if (found) {
            Zr = (Z - zlev[0]) / (zlev[1] - zlev[0]);
            zcff[0] = 1.0 - Zr;
            zcff[1] = Zr;
        } else {
            /* z at or above the top w level: the reference blends two
             * copies of level nz with an undefined weight; both give the
             * same value, taken here with weights 1/2, 1/2. */
            zcff[0] = 0.5;
            zcff[1] = 0.5;
        }
    }

    int Inode = (int)floor(X), Jnode = (int)floor(Y);
    double x = X - Inode, y = Y - Jnode;
    double cff[4], m[4];
    double vb[4], vt[4];   /* corner values at levels ibot, itop */
    cff[0] = (1 - x) * (1 - y);
    cff[1] = x * (1 - y);
    cff[2] = x * y;
    cff[3] = (1 - x) * y;

    if (var == 0) {
        double uvI = X - 0.5;
        Inode = (int)floor(uvI);
        x = uvI - Inode;
        cff[0] = (1 - x) * (1 - y);
        cff[1] = x * (1 - y);
        cff[2] = x * y;
        cff[3] = (1 - x) * y;
        const int xiu = g->L - 1;
        int ia, ib;
        if (Inode >= xiu) {
            ia = ib = xiu;
            m[0] = MKU(xiu, Jnode); m[1] = 0; m[2] = 0; m[3] = MKU(xiu, Jnode + 1);
        } else if (Inode < 1) {
            ia = ib = 1;
            m[0] = 0; m[1] = MKU(1, Jnode); m[2] = MKU(1, Jnode + 1); m[3] = 0;
        } else {
            ia = Inode;
            ib = Inode + 1;
            m[0] = MKU(Inode, Jnode); m[1] = MKU(Inode + 1, Jnode);
            m[2] = MKU(Inode + 1, Jnode + 1); m[3] = MKU(Inode, Jnode + 1);
        }
        const double *f = g->u[s];
        vb[0] = U3(f, ia, Jnode, ibot);     vt[0] = U3(f, ia, Jnode, itop);
        vb[1] = U3(f, ib, Jnode, ibot);     vt[1] = U3(f, ib, Jnode, itop);
        vb[2] = U3(f, ib, Jnode + 1, ibot); vt[2] = U3(f, ib, Jnode + 1, itop);
        vb[3] = U3(f, ia, Jnode + 1, ibot); vt[3] = U3(f, ia, Jnode + 1, itop);
        double nwater = m[0] + m[1] + m[2] + m[3];
        if (nwater < 4.0 && nwater > 0.0) {
            for (int c = 0; c < 4; c++)
                if (m[c] == 0.0) { vb[c] = 0.0; vt[c] = 0.0; }
        } else if (nwater == 0.0) {
            for (int c = 0; c < 4; c++) { cff[c] = 0.0; vb[c] = 0.0; vt[c] = 0.0; }
        }
    } else if (var == 1) {
        double uvJ = Y - 0.5;
        Jnode = (int)floor(uvJ);
        if (Jnode < 1) Jnode = 1;
        const int etav = g->M - 1;
        if (Jnode > etav) Inode = etav;     /* interpolation_module.f90:696 assigns Inode */
        y = uvJ - Jnode;
        cff[0] = (1 - x) * (1 - y);
        cff[1] = x * (1 - y);
        cff[2] = x * y;
        cff[3] = (1 - x) * y;
        int ja, jb;
        if (Jnode >= etav) {
            ja = jb = etav;
            m[0] = MKV(Inode, etav); m[1] = MKV(Inode + 1, etav); m[2] = 0; m[3] = 0;
        } else if (Jnode < 1) {
            ja = jb = 1;
            m[0] = 0; m[1] = 0; m[2] = MKV(Inode + 1, 1); m[3] = MKV(Inode, 1);
        } else {
            ja = Jnode;
            jb = Jnode + 1;
            m[0] = MKV(Inode, Jnode); m[1] = MKV(Inode + 1, Jnode);
            m[2] = MKV(Inode + 1, Jnode + 1); m[3] = MKV(Inode, Jnode + 1);
        }
        const double *f = g->v[s];
        vb[0] = V3(f, Inode, ja, ibot);     vt[0] = V3(f, Inode, ja, itop);
        vb[1] = V3(f, Inode + 1, ja, ibot); vt[1] = V3(f, Inode + 1, ja, itop);
        vb[2] = V3(f, Inode + 1, jb, ibot); vt[2] = V3(f, Inode + 1, jb, itop);
        vb[3] = V3(f, Inode, jb, ibot);     vt[3] = V3(f, Inode, jb, itop);
        double nwater = m[0] + m[1] + m[2] + m[3];
        if (nwater < 4.0 && nwater > 0.0) {
            for (int c = 0; c < 4; c++)
                if (m[c] == 0.0) { vb[c] = 0.0; vt[c] = 0.0; }
        } else if (nwater == 0.0) {
            for (int c = 0; c < 4; c++) { cff[c] = 0.0; vb[c] = 0.0; vt[c] = 0.0; }
        }
    } else if (var == 3 || var == 4) {
        /* salt, temp: rho points and rho levels, the w/AKs mask rule
         * (interpolation_module.f90:769-830) */
        const double *f = var == 3 ? g->salt[s] : g->temp[s];
        vb[0] = W3(f, Inode, Jnode, ibot);         vt[0] = W3(f, Inode, Jnode, itop);
        vb[1] = W3(f, Inode + 1, Jnode, ibot);     vt[1] = W3(f, Inode + 1, Jnode, itop);
        vb[2] = W3(f, Inode + 1, Jnode + 1, ibot); vt[2] = W3(f, Inode + 1, Jnode + 1, itop);
        vb[3] = W3(f, Inode, Jnode + 1, ibot);     vt[3] = W3(f, Inode, Jnode + 1, itop);
        m[0] = R2(mk_r, Inode, Jnode);
        m[1] = R2(mk_r, Inode + 1, Jnode);
        m[2] = R2(mk_r, Inode + 1, Jnode + 1);
        m[3] = R2(mk_r, Inode, Jnode + 1);
        double nwater = m[0] + m[1] + m[2] + m[3];
        if (nwater < 4.0)
            og_getcoeff(x, y, m, cff);
    } else {
        const double *f = g->w[s];
        vb[0] = W3(f, Inode, Jnode, ibot);         vt[0] = W3(f, Inode, Jnode, itop);
        vb[1] = W3(f, Inode + 1, Jnode, ibot);     vt[1] = W3(f, Inode + 1, Jnode, itop);
        vb[2] = W3(f, Inode + 1, Jnode + 1, ibot); vt[2] = W3(f, Inode + 1, Jnode + 1, itop);
        vb[3] = W3(f, Inode, Jnode + 1, ibot);     vt[3] = W3(f, Inode, Jnode + 1, itop);
        m[0] = R2(mk_r, Inode, Jnode);
        m[1] = R2(mk_r, Inode + 1, Jnode);
        m[2] = R2(mk_r, Inode + 1, Jnode + 1);
        m[3] = R2(mk_r, Inode, Jnode + 1);
        double nwater = m[0] + m[1] + m[2] + m[3];
        if (nwater < 4.0)
            og_getcoeff(x, y, m, cff);
    }
#undef MKU
#undef MKV
    double zb = cff[0] * vb[0] + cff[1] * vb[1] + cff[2] * vb[2] + cff[3] * vb[3];
    double zt = cff[0] * vt[0] + cff[1] * vt[1] + cff[2] * vt[2] + cff[3] * vt[3];
    return zcff[0] * zb + zcff[1] * zt;
}

/* interpolation_module.f90:1175-1229 */
double og_interp_aks(const og_grid *g, double X, double Y, int slot, int k)
{
    int Inode = (int)floor(X), Jnode = (int)floor(Y);
    double x = X - Inode, y = Y - Jnode;
    double cff[4], m[4], v[4];
    cff[0] = (1 - x) * (1 - y);
    cff[1] = x * (1 - y);
    cff[2] = x * y;
    cff[3] = (1 - x) * y;
    const double *f = g->aks[slot - 1];
    v[0] = W3(f, Inode, Jnode, k);
    v[1] = W3(f, Inode + 1, Jnode, k);
    v[2] = W3(f, Inode + 1, Jnode + 1, k);
    v[3] = W3(f, Inode, Jnode + 1, k);
    m[0] = R2(g->mask_rho, Inode, Jnode);
    m[1] = R2(g->mask_rho, Inode + 1, Jnode);
    m[2] = R2(g->mask_rho, Inode + 1, Jnode + 1);
    m[3] = R2(g->mask_rho, Inode, Jnode + 1);
    double nwater = m[0] + m[1] + m[2] + m[3];
    if (nwater < 4.0)
        og_getcoeff(x, y, m, cff);
    return cff[0] * v[0] + cff[1] * v[1] + cff[2] * v[2] + cff[3] * v[3];
}

/* interpolation_module.f90:269-420 (Reid 1969 crossing count) */
int og_inside(int nb, const double *xb_in, const double *yb_in, double xo, double yo)
{
    double xb[16], yb[16];
    int crossings = 0;
    for (int k = 0; k < nb; k++) { xb[k] = xb_in[k]; yb[k] = yb_in[k]; }
    xb[nb] = xb[0];
    yb[nb] = yb[0];
    int sidx[16], nc = 0;
    for (int k = 0; k < nb; k++)
        if (((xb[k + 1] - xo) * (xo - xb[k]) >= 0.0) && (xb[k] != xb[k + 1]))
            sidx[nc++] = k;
    for (int q = 0; q < nc; q++) {
        int k = sidx[q];
        if (xb[k] != xb[k + 1]) {
            double dx1 = xo - xb[k];
            double dx2 = xb[k + 1] - xo;
            double dxy = dx2 * (yo - yb[k]) - dx1 * (yb[k + 1] - yo);
            int inc = 0;
            if ((xb[k] == xo) && (yb[k] == yo)) {
                crossings = 1;
                break;
            } else if (((dx1 == 0.0) && (yo >= yb[k])) || ((dx2 == 0.0) && (yo >= yb[k + 1]))) {
                inc = 1;
            } else if ((dx1 * dx2 > 0.0) && ((xb[k + 1] - xb[k]) * dxy >= 0.0)) {
                inc = 2;
            }
            if (xb[k + 1] > xb[k])
                crossings += inc;
            else
                crossings -= inc;
        }
    }
    return crossings != 0;
}

/* inside() (interpolation_module.f90:261-420) for a boundary of any length:
 * xb, yb hold nb vertices and one more slot for the closing vertex. */
static int og_inside_n(int nb, double *xb, double *yb, double xo, double yo)
{
    int crossings = 0;
    xb[nb] = xb[0];
    yb[nb] = yb[0];
    for (int k = 0; k < nb; k++) {
        if (!(((xb[k + 1] - xo) * (xo - xb[k]) >= 0.0) && (xb[k] != xb[k + 1]))) continue;
        double dx1 = xo - xb[k];
        double dx2 = xb[k + 1] - xo;
        double dxy = dx2 * (yo - yb[k]) - dx1 * (yb[k + 1] - yo);
        int inc = 0;
        if ((xb[k] == xo) && (yb[k] == yo)) {
            crossings = 1;
            break;
        } else if (((dx1 == 0.0) && (yo >= yb[k])) || ((dx2 == 0.0) && (yo >= yb[k + 1]))) {
            inc = 1;
        } else if ((dx1 * dx2 > 0.0) && ((xb[k + 1] - xb[k]) * dxy >= 0.0)) {
            inc = 2;
        }
        if (xb[k + 1] > xb[k]) crossings += inc;
        else crossings -= inc;
    }
    return crossings != 0;
}

/* inbox (interpolation_module.f90:227-259): the boundary of the index box
 * [Li,Ui] x [Lj,Uj] of the (L, M) column-major lon/lat arrays, in the
 * reference's vertex order, then inside(). */
static int og_inbox(int Li, int Ui, int Lj, int Uj, const double *lon, const double *lat, int L,
                    double xo, double yo)
{
    const int nb = 2 * (Uj - Lj + Ui - Li);
    double *xb = (double *)malloc(sizeof(double) * (size_t)(nb + 1));
    double *yb = (double *)malloc(sizeof(double) * (size_t)(nb + 1));
#define LL(a, i, j) (a)[(size_t)((j) - 1) * L + ((i) - 1)]
    int shft = 1 - Li;
    for (int i = Li; i <= Ui - 1; i++) { xb[i + shft - 1] = LL(lon, i, Lj); yb[i + shft - 1] = LL(lat, i, Lj); }
    shft = 1 - Lj + Ui - Li;
    for (int j = Lj; j <= Uj - 1; j++) { xb[j + shft - 1] = LL(lon, Ui, j); yb[j + shft - 1] = LL(lat, Ui, j); }
    shft = 1 + Uj - Lj + 2 * Ui - Li;
    for (int i = Ui; i >= Li + 1; i--) { xb[shft - i - 1] = LL(lon, i, Uj); yb[shft - i - 1] = LL(lat, i, Uj); }
    shft = 1 + 2 * Uj - Lj + 2 * (Ui - Li);
    for (int j = Uj; j >= Lj + 1; j--) { xb[shft - j - 1] = LL(lon, Li, j); yb[shft - j - 1] = LL(lat, Li, j); }
#undef LL
    const int in = og_inside_n(nb, xb, yb, xo, yo);
    free(xb);
    free(yb);
    return in;
}

/* LL2ij (interpolation_module.f90:110-225) for one point: bisection of the
 * index box with inbox, then the Cartesian linear map or the spherical
 * local-triangle map (spherical = GRIDS(1)%spherical). */
void og_ll2ij(const double *lon, const double *lat, const double *angle, int L, int M,
              int spherical, double inlon, double inlat, double *Ipar, double *Jpar)
{
    const double pi = 3.14159265358979323846, deg2rad = pi / 180.0, Eradius = 6371315.0;
#define LL(a, i, j) (a)[(size_t)((j) - 1) * L + ((i) - 1)]
    int Imin = 1, Imax = L, Jmin = 1, Jmax = M;
    int found = og_inbox(Imin, Imax, Jmin, Jmax, lon, lat, L, inlon, inlat);
    if (found) {
        while (((Imax - Imin) > 1) || ((Jmax - Jmin) > 1)) {
            if ((Imax - Imin) > 1) {
                int i0 = (Imin + Imax) / 2;
                if (og_inbox(Imin, i0, Jmin, Jmax, lon, lat, L, inlon, inlat)) Imax = i0;
                else Imin = i0;
            }
            if ((Jmax - Jmin) > 1) {
                int j0 = (Jmin + Jmax) / 2;
                if (og_inbox(Imin, Imax, Jmin, j0, lon, lat, L, inlon, inlat)) Jmax = j0;
                else Jmin = j0;
            }
        }
    }
    found = og_inbox(Imin, Imax, Jmin, Jmax, lon, lat, L, inlon, inlat);
    if (spherical == 0) {
        double xmin = LL(lon, Imin, Jmin), xmax = LL(lon, Imax, Jmax);
        double ymin = LL(lat, Imin, Jmin), ymax = LL(lat, Imax, Jmax);
        double xfac = ((double)Imax - (double)Imin) / (xmax - xmin);
        double yfac = ((double)Jmax - (double)Jmin) / (ymax - ymin);
        double dx = (inlon - xmin) * xfac;
        double dy = (inlat - ymin) * yfac;
        *Ipar = (double)Imin + dx;
        *Jpar = (double)Jmin + dy;
    } else if (found) {
        double yfac = Eradius * deg2rad;
        double xfac = yfac * cos(inlat * deg2rad);
        double xpp = (inlon - LL(lon, Imin, Jmin)) * xfac;
        double ypp = (inlat - LL(lat, Imin, Jmin)) * yfac;
        double t1 = (LL(lon, Imin + 1, Jmin) - LL(lon, Imin, Jmin + 1)) * xfac;
        double t2 = (LL(lat, Imin + 1, Jmin) - LL(lat, Imin, Jmin + 1)) * yfac;
        double diag2 = t1 * t1 + t2 * t2;
        t1 = (LL(lon, Imin, Jmin) - LL(lon, Imin + 1, Jmin)) * xfac;
        t2 = (LL(lat, Imin, Jmin) - LL(lat, Imin + 1, Jmin)) * yfac;
        double aa2 = t1 * t1 + t2 * t2;
        t1 = (LL(lon, Imin, Jmin) - LL(lon, Imin, Jmin + 1)) * xfac;
        t2 = (LL(lat, Imin, Jmin) - LL(lat, Imin, Jmin + 1)) * yfac;
        double bb2 = t1 * t1 + t2 * t2;
        double phi = asin((diag2 - aa2 - bb2) / (2.0 * sqrt(aa2 * bb2)));
        double ang = angle ? LL(angle, Imin, Jmin) : 0.0;
        double dx = xpp * cos(ang) + ypp * sin(ang);
        double dy = ypp * cos(ang) - xpp * sin(ang);
        dx = dx + dy * tan(phi);
        dy = dy / cos(phi);
        dx = fmin(fmax(0.0, dx / sqrt(aa2)), 1.0);
        dy = fmin(fmax(0.0, dy / sqrt(bb2)), 1.0);
        *Ipar = (double)Imin + dx;
        *Jpar = (double)Jmin + dy;
    } else {
        *Ipar = -1.0;
        *Jpar = -1.0;
    }
#undef LL
}

/* zbounds (boundary_module.f90:145-179): depth bilinear, zeta (IDW) at slot t */
void og_zbounds(const og_grid *g, double Ipar, double Jpar, double *Zpar, int *inz, int t)
{
    *inz = 0;
    int I = (int)floor(Ipar);
    if (I > 0) {
        double tdepth = -1.0 * og_interp2d(g, 0, 1, Ipar, Jpar);
        double tzeta = og_interp2d(g, 3, t, Ipar, Jpar);
        if ((*Zpar < tzeta) && (*Zpar > tdepth)) *inz = 1;
        if (*Zpar >= tzeta) {
            *Zpar = tzeta - 0.01;
            *inz = 1;
        }
    }
}

/* ini_ROMSPath's placement of the released particles (ROMSPath.f90:322-403):
 * status 9, GID Ngrid, X = Y = 1 and isIn false to start; then for each grid
 * in order, LL2ij and bounds, and the first grid that holds a particle takes
 * it (status 0, zbounds at slot t_b, status 9 again if z is outside). */
void og_locate(const og_grid *grids, int ngrid, const double *const *lon, const double *const *lat,
               const double *const *angle, int spherical, int tb, int64_t n, const double *inlon,
               const double *inlat, double *z, double *x, double *y, double *status, int32_t *gid,
               uint8_t *isin)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t q = 0; q < n; q++) {
        x[q] = 1.0;
        y[q] = 1.0;
        status[q] = 9.0;
        gid[q] = ngrid;
        isin[q] = 0;
        for (int ng = 1; ng <= ngrid; ng++) {
            const og_grid *g = &grids[ng - 1];
            double I, J;
            og_ll2ij(lon[ng - 1], lat[ng - 1], angle ? angle[ng - 1] : NULL, g->L, g->M, spherical,
                     inlon[q], inlat[q], &I, &J);
            int nmask, ingrid, obound;
            og_bounds(g, I, J, &nmask, &ingrid, &obound);
            if (!isin[q] && ingrid) {
                x[q] = I;
                y[q] = J;
                gid[q] = ng;
                status[q] = 0.0;
                isin[q] = 1;
                int inz;
                og_zbounds(g, I, J, &z[q], &inz, tb);
                if (!inz) status[q] = 9.0;
            }
        }
    }
}

/* The SaltTempOn block of update_particles (ROMSPath.f90:958-981) at the
 * particles' positions: zeta of slots b, c, f (IDW), getInterp3d("salt"/
 * "temp", ztype 1) with each slot's own zeta, then polintd in time at ix(2).
 * The depth passed to getInterp3d is the step's local tdepth
 * (ROMSPath.f90:849, 964-971): bath (= -tdepth, the pBath column the step
 * wrote) when given.  For a rejected move that depth belongs to the
 * attempted position (tempX, tempY) on the old grid while X, Y stay.
 * bath == NULL: the depth at (x, y) at slot t_c, as the release-time
 * sample computes it (ROMSPath.f90:411-414). */
void og_sample_tracers(const og_grid *grids, const og_time *T, int64_t n, const double *x,
                       const double *y, const double *z, const int32_t *gid, const double *bath,
                       double *salt, double *temp)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t q = 0; q < n; q++) {
        const og_grid *g = &grids[gid[q] - 1];
        double tdepth = bath ? -1.0 * bath[q] : -1.0 * og_interp2d(g, 0, T->tc, x[q], y[q]);
        double zb = og_interp2d(g, 3, T->tb, x[q], y[q]);
        double zc = og_interp2d(g, 3, T->tc, x[q], y[q]);
        double zf = og_interp2d(g, 3, T->tf, x[q], y[q]);
        double ey[3];
        ey[0] = og_interp3d(g, 3, x[q], y[q], z[q], T->tb, 1, zb, tdepth);
        ey[1] = og_interp3d(g, 3, x[q], y[q], z[q], T->tc, 1, zc, tdepth);
        ey[2] = og_interp3d(g, 3, x[q], y[q], z[q], T->tf, 1, zf, tdepth);
        salt[q] = og_polintd(T->ex, ey, T->ix[1]);
        ey[0] = og_interp3d(g, 4, x[q], y[q], z[q], T->tb, 1, zb, tdepth);
        ey[1] = og_interp3d(g, 4, x[q], y[q], z[q], T->tc, 1, zc, tdepth);
        ey[2] = og_interp3d(g, 4, x[q], y[q], z[q], T->tf, 1, zf, tdepth);
        temp[q] = og_polintd(T->ex, ey, T->ix[1]);
    }
}

/* getInterpStr (interpolation_module.f90:964-1174) without WETDRY: var 0
 * bustr at u points, var 1 bvstr at v points, slot 1..3.  The Jnode < 1
 * branch of bvstr is unreachable after the reference's clamp and is left
 * out.  Node indices are clamped into the arrays for the reads only (no
 * effect for a particle inside the grid; the reference would read outside). */
static int og_clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

double og_interp_str(const og_grid *g, int var, int slot, double Ipar, double Jpar)
{
    const int s = slot - 1;
    /* -DWETDRY: this slot's wet/dry masks (interpolation_module.f90:1022-1060, 1103-1140) */
    const int32_t *mk_u = g->wd_u[s] ? g->wd_u[s] : g->mask_u;
    const int32_t *mk_v = g->wd_v[s] ? g->wd_v[s] : g->mask_v;
    int Inode = (int)floor(Ipar), Jnode = (int)floor(Jpar);
    double X = Ipar - Inode, Y = Jpar - Jnode;
    double cff[4], v[4], m[4];
    if (var == 0) {
        const int xiu = g->L - 1;
        const double *f = g->bustr[s];
#define BU(a, i, j) ((double)(a)[((i) - 1) + (int64_t)xiu * ((j) - 1)])
        double uvI = Ipar - 0.5;
        Inode = (int)floor(uvI);
        X = uvI - Inode;
        const int j0 = og_clampi(Jnode, 1, g->M), j1 = og_clampi(Jnode + 1, 1, g->M);
        if (Inode >= xiu) {
            v[0] = BU(f, xiu, j0); v[1] = BU(f, xiu, j0);
            v[2] = BU(f, xiu, j1); v[3] = BU(f, xiu, j1);
            m[0] = BU(mk_u, xiu, j0); m[1] = 0; m[2] = 0; m[3] = BU(mk_u, xiu, j1);
        } else if (Inode < 1) {
            v[0] = BU(f, 1, j0); v[1] = BU(f, 1, j0);
            v[2] = BU(f, 1, j1); v[3] = BU(f, 1, j1);
            m[0] = 0; m[1] = BU(mk_u, 1, j0); m[2] = BU(mk_u, 1, j1); m[3] = 0;
        } else {
            v[0] = BU(f, Inode, j0); v[1] = BU(f, Inode + 1, j0);
            v[2] = BU(f, Inode + 1, j1); v[3] = BU(f, Inode, j1);
            m[0] = BU(mk_u, Inode, j0); m[1] = BU(mk_u, Inode + 1, j0);
            m[2] = BU(mk_u, Inode + 1, j1); m[3] = BU(mk_u, Inode, j1);
        }
#undef BU
    } else {
        const int etav = g->M - 1;
        const double *f = g->bvstr[s];
#define BV(a, i, j) ((double)(a)[((i) - 1) + (int64_t)g->L * ((j) - 1)])
        double uvJ = Jpar - 0.5;
        Jnode = (int)floor(uvJ);
        if (Jnode < 1) Jnode = 1;
        if (Jnode > etav) Inode = etav;      /* interpolation_module.f90:1089 assigns Inode */
        Y = uvJ - Jnode;
        const int i0 = og_clampi(Inode, 1, g->L), i1 = og_clampi(Inode + 1, 1, g->L);
        if (Jnode >= etav) {
            v[0] = BV(f, i0, etav); v[1] = BV(f, i1, etav);
            v[2] = BV(f, i1, etav); v[3] = BV(f, i0, etav);
            m[0] = BV(mk_v, i0, etav); m[1] = BV(mk_v, i1, etav); m[2] = 0; m[3] = 0;
        } else {
            v[0] = BV(f, i0, Jnode); v[1] = BV(f, i1, Jnode);
            v[2] = BV(f, i1, Jnode + 1); v[3] = BV(f, i0, Jnode + 1);
            m[0] = BV(mk_v, i0, Jnode); m[1] = BV(mk_v, i1, Jnode);
            m[2] = BV(mk_v, i1, Jnode + 1); m[3] = BV(mk_v, i0, Jnode + 1);
        }
#undef BV
    }
    cff[0] = (1 - X) * (1 - Y); cff[1] = X * (1 - Y); cff[2] = X * Y; cff[3] = (1 - X) * Y;
    double nwater = m[0] + m[1] + m[2] + m[3];
    if (nwater < 4.0 && nwater > 0.0) {
        for (int c = 0; c < 4; c++)
            if (m[c] == 0.0) v[c] = 0.0;
    } else if (nwater == 0.0) {
        for (int c = 0; c < 4; c++) { cff[c] = 0.0; v[c] = 0.0; }
    }
    return cff[0] * v[0] + cff[1] * v[1] + cff[2] * v[2] + cff[3] * v[3];
}

/* The WriteBottom block of update_particles (ROMSPath.f90:935-951) for the
 * particles that took the last step (their age changed): height above the
 * bed from the post-step depth (bath = -tdepth), its running minimum, and
 * bustr/bvstr (polintd at ix(2)) where the minimum was just reached. */
void og_bottom_update(const og_grid *grids, const og_time *T, int64_t n, const double *x,
                      const double *y, const double *z, const int32_t *gid, const double *age,
                      const double *bath, double *last_age, double *hob, double *bustr,
                      double *bvstr)
{
    for (int64_t q = 0; q < n; q++) {
        if (age[q] == last_age[q]) continue;
        last_age[q] = age[q];
        const og_grid *g = &grids[gid[q] - 1];
        const double btemp = bath[q] + z[q];
        hob[q] = fmin(btemp, hob[q]);
        if (btemp == hob[q]) {
            double ey[3];
            ey[0] = og_interp_str(g, 0, T->tb, x[q], y[q]);
            ey[1] = og_interp_str(g, 0, T->tc, x[q], y[q]);
            ey[2] = og_interp_str(g, 0, T->tf, x[q], y[q]);
            bustr[q] = og_polintd(T->ex, ey, T->ix[1]);
            ey[0] = og_interp_str(g, 1, T->tb, x[q], y[q]);
            ey[1] = og_interp_str(g, 1, T->tc, x[q], y[q]);
            ey[2] = og_interp_str(g, 1, T->tf, x[q], y[q]);
            bvstr[q] = og_polintd(T->ex, ey, T->ix[1]);
        }
    }
}

/* boundary_module.f90:43-143 */
void og_bounds(const og_grid *g, double X, double Y, int *nmask, int *ingrid, int *obound)
{
    const double ecut = F_0P1, fcut = 1.0 - ecut;
    *ingrid = 0;
    *obound = 0;
    *nmask = 0;
    int I = (int)floor(X), J = (int)floor(Y);
    double x = X - (double)I, y = Y - (double)J;
    if ((I > 1) && (J > 1) && (I < g->L - 1) && (J < g->M - 1)) {
        double m11 = R2(g->mask_rho, I, J), m21 = R2(g->mask_rho, I + 1, J);
        double m22 = R2(g->mask_rho, I + 1, J + 1), m12 = R2(g->mask_rho, I, J + 1);
        int nm = (int)(m11 + m21 + m22 + m12);
        *nmask = nm;
        if (nm == 4)
            *ingrid = 1;
        if (nm == 3) {
            double tX[4] = {0, 0, 0, 0}, tY[4] = {0, 0, 0, 0};
            double mm[2][2] = {{m11, m12}, {m21, m22}};   /* mm[it-1][jt-1] = m(it,jt) */
            int n = 0;
            for (int it = 1; it <= 2; it++)
                for (int jt = 1; jt <= 2; jt++)
                    if (mm[it - 1][jt - 1] == 1) {
                        tX[n] = (double)it - 1.0;
                        tY[n] = (double)jt - 1.0;
                        n++;
                    }
            if (og_inside(3, tX, tY, x, y))
                *ingrid = 1;
        }
        if (nm == 2) {
            if (m11 == 1 && (x < ecut) && (y < ecut)) *ingrid = 1;
            if (m21 == 1 && (x > fcut) && (y < ecut)) *ingrid = 1;
            if (m22 == 1 && (x > fcut) && (y > fcut)) *ingrid = 1;
            if (m12 == 1 && (x < ecut) && (y > fcut)) *ingrid = 1;
        }
        if ((I <= 1) || (J <= 1) || (I >= g->L - 1) || (J >= g->M - 1))
            *obound = 1;
    } else {
        *obound = 1;
    }
}

/* advection_module.f90:129-350 (version=2: time frozen at ix(2)) */
void og_find_currents(const og_grid *g, const og_time *T, double X, double Y, double *Z,
                      double *U, double *V, double *W, double *tdepth, double *zeta)
{
    double ey[3];
    *tdepth = -1.0 * og_interp2d(g, 0, T->tc, X, Y);
    double zetab = og_interp2d(g, 3, T->tb, X, Y);
    double zetac = og_interp2d(g, 3, T->tc, X, Y);
    double zetaf = og_interp2d(g, 3, T->tf, X, Y);
    if (*Z < *tdepth)
        *Z = *tdepth + F_0P001;
    ey[0] = zetab; ey[1] = zetac; ey[2] = zetaf;
    *zeta = og_polintd(T->ex, ey, T->ix[1]);
    if (*Z > *zeta)
        *Z = *zeta - F_0P001;

    double Uc = og_interp3d(g, 0, X, Y, *Z, T->tc, 1, *zeta, *tdepth);
    double Uf = og_interp3d(g, 0, X, Y, *Z, T->tf, 1, *zeta, *tdepth);
    double Ub = og_interp3d(g, 0, X, Y, *Z, T->tb, 1, *zeta, *tdepth);
    double Vb = og_interp3d(g, 1, X, Y, *Z, T->tb, 1, *zeta, *tdepth);
    double Vc = og_interp3d(g, 1, X, Y, *Z, T->tc, 1, *zeta, *tdepth);
    double Vf = og_interp3d(g, 1, X, Y, *Z, T->tf, 1, *zeta, *tdepth);
    double Wb = og_interp3d(g, 2, X, Y, *Z, T->tb, 2, *zeta, *tdepth);
    double Wc = og_interp3d(g, 2, X, Y, *Z, T->tc, 2, *zeta, *tdepth);
    double Wf = og_interp3d(g, 2, X, Y, *Z, T->tf, 2, *zeta, *tdepth);

    if (T->ets == 1) { ey[0] = Ub; ey[1] = Ub; ey[2] = Uc; }
    else             { ey[0] = Ub; ey[1] = Uc; ey[2] = Uf; }
    *U = og_polintd(T->ex, ey, T->ix[1]);
    if (T->ets == 1) { ey[0] = Vb; ey[1] = Vb; ey[2] = Vc; }
    else             { ey[0] = Vb; ey[1] = Vc; ey[2] = Vf; }
    *V = og_polintd(T->ex, ey, T->ix[1]);
    if (T->ets == 1) { ey[0] = Wb; ey[1] = Wb; ey[2] = Wc; }
    else             { ey[0] = Wb; ey[1] = Wc; ey[2] = Wf; }
    *W = og_polintd(T->ex, ey, T->ix[1]);
}

/* advection_module.f90:20-125; Z is clamped in place by the first stage */
void og_rkadvect(const og_grid *g, const og_params *P, const og_time *T, double X, double Y,
                 double *Z, double pm, double pn, double *ax, double *ay, double *az)
{
    double Uad, Vad, Wad, dep, zet;
    const double idt = (double)P->idt;
    og_find_currents(g, T, X, Y, Z, &Uad, &Vad, &Wad, &dep, &zet);
    double k1u = Uad * pm, k1v = Vad * pn, k1w = Wad;
    double x1 = X + Uad * pm * idt / 2.0;
    double y1 = Y + Vad * pn * idt / 2.0;
    double z1 = *Z + Wad * idt / 2.0;
    og_find_currents(g, T, x1, y1, &z1, &Uad, &Vad, &Wad, &dep, &zet);
    double k2u = Uad * pm, k2v = Vad * pn, k2w = Wad;
    double x2 = X + Uad * pm * idt / 2.0;
    double y2 = Y + Vad * pn * idt / 2.0;
    double z2 = *Z + Wad * idt / 2.0;
    og_find_currents(g, T, x2, y2, &z2, &Uad, &Vad, &Wad, &dep, &zet);
    double k3u = Uad * pm, k3v = Vad * pn, k3w = Wad;
    double x3 = X + Uad * pm * idt / 2.0;
    double y3 = Y + Vad * pn * idt / 2.0;
    double z3 = *Z + Wad * idt / 2.0;
    og_find_currents(g, T, x3, y3, &z3, &Uad, &Vad, &Wad, &dep, &zet);
    double k4u = Uad * pm, k4v = Vad * pn, k4w = Wad;
    double PU = (k1u + 2.0 * k2u + 2.0 * k3u + k4u) / 6.0;
    double PV = (k1v + 2.0 * k2v + 2.0 * k3v + k4v) / 6.0;
    double PW = (k1w + 2.0 * k2w + 2.0 * k3w + k4w) / 6.0;
    *ax = idt * PU;
    *ay = idt * PV;
    *az = idt * PW;
}

/* ---------------- PPPACK (pppack.f90) ---------------- */
#define V7(i, c)  v[((i) - 1) + (int64_t)np * ((c) - 1)]
#define A4(i, c)  a[((i) - 1) + (int64_t)np * ((c) - 1)]

/* pppack.f90:4726-4825 */
static void og_setupq(const double *x, const double *dx, const double *y, int np, double *v,
                      double *qty)
{
    for (int i = 1; i <= np - 1; i++)
        V7(i, 4) = x[i] - x[i - 1];
    for (int i = 2; i <= np - 1; i++)
        V7(i, 1) = dx[i - 2] / V7(i - 1, 4);
    V7(np, 1) = 0.0;
    for (int i = 2; i <= np - 1; i++)
        V7(i, 2) = -dx[i - 1] / V7(i, 4) - dx[i - 1] / V7(i - 1, 4);
    for (int i = 2; i <= np - 1; i++)
        V7(i, 3) = dx[i] / V7(i, 4);
    for (int i = 2; i <= np - 1; i++)
        V7(i, 5) = V7(i, 1) * V7(i, 1) + V7(i, 2) * V7(i, 2) + V7(i, 3) * V7(i, 3);
    for (int i = 2; i <= np - 2; i++)
        V7(i, 6) = V7(i, 2) * V7(i + 1, 1) + V7(i, 3) * V7(i + 1, 2);
    V7(np - 1, 6) = 0.0;
    for (int i = 2; i <= np - 3; i++)
        V7(i, 7) = V7(i, 3) * V7(i + 2, 1);
    V7(np - 2, 7) = 0.0;
    V7(np - 1, 7) = 0.0;
    double prev = (y[1] - y[0]) / V7(1, 4);
    for (int i = 2; i <= np - 1; i++) {
        double diff = (y[i] - y[i - 1]) / V7(i, 4);
        qty[i - 1] = diff - prev;
        prev = diff;
    }
}

/* pppack.f90:1428-1566 (ncol=1) */
static void og_chol1d(double p, double *v, const double *qty, int np, double *u, double *qu)
{
    double six1mp = 6.0 * (1.0 - p);
    double twop = 2.0 * p;
    for (int i = 2; i <= np - 1; i++)
        V7(i, 1) = six1mp * V7(i, 5) + twop * (V7(i - 1, 4) + V7(i, 4));
    for (int i = 2; i <= np - 1; i++)
        V7(i, 2) = six1mp * V7(i, 6) + p * V7(i, 4);
    for (int i = 2; i <= np - 1; i++)
        V7(i, 3) = six1mp * V7(i, 7);
    if (np < 4) {
        u[0] = 0.0;
        u[1] = qty[1] / V7(2, 1);
        u[2] = 0.0;
    } else {
        for (int i = 2; i <= np - 2; i++) {
            double ratio = V7(i, 2) / V7(i, 1);
            V7(i + 1, 1) = V7(i + 1, 1) - ratio * V7(i, 2);
            V7(i + 1, 2) = V7(i + 1, 2) - ratio * V7(i, 3);
            V7(i, 2) = ratio;
            ratio = V7(i, 3) / V7(i, 1);
            V7(i + 2, 1) = V7(i + 2, 1) - ratio * V7(i, 3);
            V7(i, 3) = ratio;
        }
        u[0] = 0.0;
        V7(1, 3) = 0.0;
        u[1] = qty[1];
        for (int i = 2; i <= np - 2; i++)
            u[i] = qty[i] - V7(i, 2) * u[i - 1] - V7(i - 1, 3) * u[i - 2];
        u[np - 1] = 0.0;
        u[np - 2] = u[np - 2] / V7(np - 1, 1);
        for (int i = np - 2; i >= 2; i--)
            u[i - 1] = u[i - 1] / V7(i, 1) - u[i] * V7(i, 2) - u[i + 1] * V7(i, 3);
    }
    double prev = 0.0;
    for (int i = 2; i <= np; i++) {
        qu[i - 1] = (u[i - 1] - u[i - 2]) / V7(i - 1, 4);
        qu[i - 2] = qu[i - 1] - prev;
        prev = qu[i - 1];
    }
    qu[np - 1] = -qu[np - 1];
}

static double og_sfq(const double *a, const double *dy, int np)
{
    double s = 0.0;
    for (int i = 1; i <= np; i++)
        s += (A4(i, 1) * A4(i, 1)) * (dy[i - 1] * dy[i - 1]);
    return 36.0 * s;
}

/* pppack.f90:5108-5317; v is np*7 and a is np*4, both column-major */
double og_smooth(const double *x, const double *y, const double *dy, int np, double s,
                 double *v, double *a)
{
    double p, q, sfq, utru, ooss, oosf, prevq, prevsf, change;
    og_setupq(x, dy, y, np, v, &A4(1, 4));
    if (s <= 0.0) {
        p = 1.0;
        og_chol1d(p, v, &A4(1, 4), np, &A4(1, 3), &A4(1, 1));
        sfq = 0.0;
    } else {
        p = 0.0;
        og_chol1d(p, v, &A4(1, 4), np, &A4(1, 3), &A4(1, 1));
        sfq = og_sfq(a, dy, np);
        if (s < sfq) {
            utru = 0.0;
            for (int i = 2; i <= np; i++)
                utru = utru + V7(i - 1, 4) * (A4(i - 1, 3) * (A4(i - 1, 3) + A4(i, 3)) + A4(i, 3) * A4(i, 3));
            ooss = 1.0 / sqrt(s);
            oosf = 1.0 / sqrt(sfq);
            q = -(oosf - ooss) * sfq / (6.0 * utru * oosf);
            prevq = 0.0;
            prevsf = oosf;
            for (;;) {
                og_chol1d(q / (1.0 + q), v, &A4(1, 4), np, &A4(1, 3), &A4(1, 1));
                sfq = og_sfq(a, dy, np) / ((1.0 + q) * (1.0 + q));
                if (fabs(sfq - s) <= 0.01 * s)
                    break;
                oosf = 1.0 / sqrt(sfq);
                change = (q - prevq) / (oosf - prevsf) * (oosf - ooss);
                prevq = q;
                q = q - change;
                prevsf = oosf;
            }
            p = q / (1.0 + q);
        }
    }
    double smooth = sfq;
    for (int i = 1; i <= np; i++)
        A4(i, 1) = y[i - 1] - 6.0 * (1.0 - p) * (dy[i - 1] * dy[i - 1]) * A4(i, 1);
    for (int i = 1; i <= np; i++)
        A4(i, 3) = A4(i, 3) * 6.0 * p;
    for (int i = 1; i <= np - 1; i++) {
        A4(i, 4) = (A4(i + 1, 3) - A4(i, 3)) / V7(i, 4);
        A4(i, 2) = (A4(i + 1, 1) - A4(i, 1)) / V7(i, 4)
                   - (A4(i, 3) + A4(i, 4) / 3.0 * V7(i, 4)) / 2.0 * V7(i, 4);
    }
    return smooth;
}

/* pppack.f90:3363-3576.  For strictly increasing xt the result does not
 * depend on the saved search hint, so the hint is not kept. */
int og_interv(const double *xt, int lxt, double x, int *mflag)
{
    if (x < xt[0]) {
        *mflag = -1;
        return 1;
    }
    if (xt[lxt - 1] <= x) {
        *mflag = (x == xt[lxt - 1]) ? 0 : 1;
        int left;
        for (left = lxt - 1; left >= 1; left--)
            if (xt[left - 1] < xt[lxt - 1])
                return left;
        return left;
    }
    int ilo = 1, ihi = lxt;
    for (;;) {
        int middle = (ilo + ihi) / 2;
        if (middle == ilo) {
            *mflag = 0;
            return ilo;
        }
        if (xt[middle - 1] <= x)
            ilo = middle;
        else
            ihi = middle;
    }
}

/* pppack.f90:4246-4362; coef is (k, l) column-major */
double og_ppvalu(const double *brk, const double *coef, int l, int k, double x, int jderiv)
{
    double value = 0.0;
    double fmmjdr = k - jderiv;
    if (k <= jderiv)
        return value;
    int mflag;
    int i = og_interv(brk, l + 1, x, &mflag);
    double h = x - brk[i - 1];
    int m = k;
    for (;;) {
        value = (value / fmmjdr) * h + coef[(m - 1) + (int64_t)k * (i - 1)];
        m = m - 1;
        fmmjdr = fmmjdr - 1.0;
        if (fmmjdr <= 0.0)
            break;
    }
    return value;
}

/* ver_turb_module.f90:31-329 */
double og_vturb(const og_grid *g, const og_params *P, const og_time *T, double X, double Y,
                double Zpar, const double *dev)
{
    const int nw = g->N + 1;
    const int np = (int)((double)nw * P->sub);
    double *z = (double *)malloc(sizeof(double) * nw);
    double *pkh = (double *)malloc(sizeof(double) * nw);
    double *fitz = (double *)calloc(np, sizeof(double));
    double *fitkh = (double *)calloc(np, sizeof(double));
    double *dy = (double *)malloc(sizeof(double) * np);
    double *v = (double *)calloc((size_t)np * 7, sizeof(double));
    double *a = (double *)calloc((size_t)np * 4, sizeof(double));
    double *coefsm = (double *)calloc((size_t)np * 4, sizeof(double));
    double ey[3];

    double tdepth = -1.0 * og_interp2d(g, 0, T->tc, X, Y);
    double zetab = og_interp2d(g, 3, T->tb, X, Y);
    double zetac = og_interp2d(g, 3, T->tc, X, Y);
    double zetaf = og_interp2d(g, 3, T->tf, X, Y);
    ey[0] = zetab; ey[1] = zetac; ey[2] = zetaf;
    double zeta = og_polintd(T->ex, ey, T->ix[1]);
    for (int i = 1; i <= nw; i++) {
        ey[0] = og_getwlevel(g, zetab, tdepth, i);
        ey[1] = og_getwlevel(g, zetac, tdepth, i);
        ey[2] = og_getwlevel(g, zetaf, tdepth, i);
        z[i - 1] = og_polintd(T->ex, ey, T->ix[1]);
        ey[0] = og_interp_aks(g, X, Y, T->tb, i);
        ey[1] = og_interp_aks(g, X, Y, T->tc, i);
        ey[2] = og_interp_aks(g, X, Y, T->tf, i);
        pkh[i - 1] = og_polintd(T->ex, ey, T->ix[1]);
    }
    double dz = (z[nw - 1] - z[0]) / (double)(np - 1);
    fitz[0] = z[0];
    fitz[np - 1] = z[nw - 1];
    fitkh[0] = pkh[0];
    fitkh[np - 1] = pkh[nw - 1];
    double m;
    for (int i = 2; i <= np - 1; i++) {
        fitz[i - 1] = fitz[i - 2] + dz;
        og_linint(z, pkh, nw, fitz[i - 1], &fitkh[i - 1], &m);
        dy[i - 1] = P->serr;
    }
    dy[0] = P->serr;
    dy[np - 1] = P->serr;
    og_smooth(fitz, fitkh, dy, np, P->smth, v, a);
    for (int i = 1; i <= np; i++)
        for (int c = 1; c <= 4; c++)
            coefsm[(c - 1) + 4 * (int64_t)(i - 1)] = A4(i, c);

    const int loop = P->idt / (int)P->deltat;
    const double deltat = P->deltat;
    double ParZc = Zpar;
    for (int i = 1; i <= loop; i++) {
        double Kprimec = 0.0, KprimeZc, Z3rdc, KH3rdc;
        if (ParZc <= tdepth || ParZc >= zeta)
            Kprimec = 0.0;
        else
            Kprimec = og_ppvalu(fitz, coefsm, np - 1, 4, ParZc, 1);
        KprimeZc = Kprimec * deltat;
        Z3rdc = ParZc + 0.5 * KprimeZc;
        if (Z3rdc < tdepth || Z3rdc > zeta) {
            KH3rdc = P->aksback;
        } else {
            KH3rdc = og_ppvalu(fitz, coefsm, np - 1, 4, Z3rdc, 0);
            if (KH3rdc < P->aksback)
                KH3rdc = P->aksback;
        }
        double DEV = dev[i - 1];
        double r = 1.0;
        double dTz = KprimeZc + DEV * sqrt(2.0 / r * KH3rdc * deltat);
        ParZc = ParZc + dTz;
        if (ParZc <= tdepth)
            ParZc = tdepth + fabs(ParZc - tdepth);
        if (ParZc >= zeta)
            ParZc = zeta - fabs(ParZc - zeta);
    }
    free(z); free(pkh); free(fitz); free(fitkh); free(dy); free(v); free(a); free(coefsm);
    return ParZc - Zpar;
}

/* Threads of the particle loop (OpenMP builds only; the reference is
 * serial, and each particle's step reads and writes only its own columns, so
 * the result does not depend on the count). */
void og_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

/* ROMSPath.f90:662-1003, behaviour off (Behavior=0: behave() returns zeros) */
int64_t og_update_particles(const og_grid *grids, const og_params *P, const og_time *T,
                            og_particles *p, int64_t n, const double *dev, int ndev)
{
    int64_t active = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : active) schedule(dynamic, 64)
#endif
    for (int64_t q = 0; q < n; q++) {
        if (T->ix[2] <= p->dob[q]) continue;
        if (p->status[q] == 9.0) continue;
        if (P->open_boundary && !p->isin[q]) continue;
        active++;
        p->age[q] = p->age[q] + (double)(float)P->idt;
        double Xpar = p->x[q], Ypar = p->y[q], Zpar = p->z[q];
        int gid = p->gid[q];
        const og_grid *g = &grids[gid - 1];
        double pm = og_interp2d(g, 1, 1, Xpar, Ypar);
        double pn = og_interp2d(g, 2, 1, Xpar, Ypar);
        double AdvectX = 0.0, AdvectY = 0.0, AdvectZ = 0.0, TurbHx = 0.0, TurbHy = 0.0, TurbV = 0.0;
        if (P->scheme == 1)
            og_rkadvect(g, P, T, Xpar, Ypar, &Zpar, pm, pn, &AdvectX, &AdvectY, &AdvectZ);
        if (P->hturb_on) {
            double KM = P->kh[gid - 1], r = 1.0;
            TurbHx = dev[q * ndev + 0] * sqrt(2.0 / r * KM * (double)P->idt);
            TurbHy = dev[q * ndev + 1] * sqrt(2.0 / r * KM * (double)P->idt);
        }
        TurbHx = TurbHx * pm;
        TurbHy = TurbHy * pn;
        if (P->vturb_on)
            TurbV = og_vturb(g, P, T, Xpar, Ypar, Zpar, dev + q * ndev + 2);
        const double Zbehav = 0.0;
        double newX = p->x[q] + AdvectX + TurbHx;
        double newY = p->y[q] + AdvectY + TurbHy;
        double tempX = newX, tempY = newY;
        int nmask, ingrid = 0, obound = 0;
        for (int i = 1; i <= P->ngrid; i++) {
            const og_grid *gc = &grids[gid - 1];
            tempX = gc->scl[i - 1][0] * newX + gc->off[i - 1][0];
            tempY = gc->scl[i - 1][1] * newY + gc->off[i - 1][1];
            og_bounds(&grids[i - 1], tempX, tempY, &nmask, &ingrid, &obound);
            if (ingrid) {
                gid = i;
                break;
            }
        }
        const og_grid *gn = &grids[gid - 1];
        double ey[3];
        double tdepth = -1.0 * og_interp2d(gn, 0, T->tc, tempX, tempY);
        ey[0] = 1.0 * og_interp2d(gn, 3, T->tb, tempX, tempY);
        ey[1] = 1.0 * og_interp2d(gn, 3, T->tc, tempX, tempY);
        ey[2] = 1.0 * og_interp2d(gn, 3, T->tf, tempX, tempY);
        double Pzeta = og_polintd(T->ex, ey, T->ix[1]);
        p->wd[q] = (-1.0 * tdepth) + Pzeta;
        p->zeta[q] = Pzeta;
        p->bath[q] = -1.0 * tdepth;
        int hbot = 0, htop = 0;
        double newZ;
        if (P->nsb == 1) {
            newZ = Pzeta - P->vertdist;
        } else if (P->nsb == 2) {
            newZ = tdepth + P->vertdist;
        } else {
            newZ = p->z[q] + AdvectZ + TurbV + Zbehav;
            if (newZ < tdepth) {
                newZ = tdepth + fabs(newZ - tdepth);
                hbot = 1;
            }
            if (newZ > Pzeta) {
                newZ = Pzeta - fabs(newZ - Pzeta);
                htop = 1;
            }
        }
        if (ingrid) {
            p->isin[q] = 1;
            p->x[q] = tempX;
            p->y[q] = tempY;
            p->z[q] = newZ;
            p->status[q] = 1.0;
            /* default-REAL literals 1.1 and 1.2, widened (ROMSPath.f90:880-881) */
            if (htop) p->status[q] = (double)1.1f;
            if (hbot) p->status[q] = (double)1.2f;
        } else if (obound) {
            p->isin[q] = 0;
            gid = P->ngrid;
            p->status[q] = -8.0;
        } else {
            p->isin[q] = 1;
            p->status[q] = -9.0;
        }
        p->gid[q] = gid;
    }
    return active;
}
