// rp_locate.hip -- placement of released particles on the grids, and the
// optional salt/temp sampler at the particles.
//
// Placement of released particles on the grids
// (ini_ROMSPath, ROMSPath.f90:322-403): LL2ij's bisection of each grid's
// index box with inbox/inside (interpolation_module.f90:110-420), bounds
// (boundary_module.f90:43-143) and zbounds (boundary_module.f90:145-179),
// the first grid that holds a particle taking it.  One lane per particle;
// the boundary polygons of inbox are walked vertex by vertex from the lon/lat
// arrays instead of being copied out.
#include "rp_step_dev.h"

namespace rp {

namespace {

// vertex t (0-based, t == nb closes the polygon) of inbox's boundary of the
// index box [Li,Ui] x [Lj,Uj], in the reference's order: the lower side
// eastward, the right side northward, the upper side westward, the left side
// southward (interpolation_module.f90:238-257)
struct BoxWalk {
    const double *lon, *lat;
    int L, Li, Ui, Lj, Uj, a, b, nb;
    __device__ __forceinline__ BoxWalk(const double* lo, const double* la, int L_, int li, int ui,
                                       int lj, int uj)
        : lon(lo), lat(la), L(L_), Li(li), Ui(ui), Lj(lj), Uj(uj)
    {
        a = Ui - Li;
        b = Uj - Lj;
        nb = 2 * (a + b);
    }
    __device__ __forceinline__ void vertex(int t, double& x, double& y) const
    {
        if (t >= nb) t = 0;
        int i, j;
        if (t < a) { i = Li + t; j = Lj; }
        else if (t < a + b) { i = Ui; j = Lj + (t - a); }
        else if (t < 2 * a + b) { i = Ui - (t - a - b); j = Uj; }
        else { i = Li; j = Uj - (t - 2 * a - b); }
        const int64_t o = (int64_t)(j - 1) * L + (i - 1);
        x = lon[o];
        y = lat[o];
    }
};

// inside() (interpolation_module.f90:261-420) over the box boundary: the
// edges are taken in order, so the crossing count and the early exit on a
// vertex hit match the reference's chunked candidate list
__device__ bool inbox(const BoxWalk& w, double xo, double yo)
{
    double xk, yk;
    w.vertex(0, xk, yk);
    int crossings = 0;
    for (int k = 0; k < w.nb; k++) {
        double x1, y1;
        w.vertex(k + 1, x1, y1);
        if (((x1 - xo) * (xo - xk) >= 0.0) && (xk != x1)) {
            const double dx1 = xo - xk;
            const double dx2 = x1 - xo;
            const double dxy = dx2 * (yo - yk) - dx1 * (y1 - yo);
            int inc = 0;
            if ((xk == xo) && (yk == yo)) return true;
            if (((dx1 == 0.0) && (yo >= yk)) || ((dx2 == 0.0) && (yo >= y1))) inc = 1;
            else if ((dx1 * dx2 > 0.0) && ((x1 - xk) * dxy >= 0.0)) inc = 2;
            if (x1 > xk) crossings += inc;
            else crossings -= inc;
        }
        xk = x1;
        yk = y1;
    }
    return crossings != 0;
}

__device__ __forceinline__ bool inbox(const double* lon, const double* lat, int L, int Li, int Ui,
                                      int Lj, int Uj, double xo, double yo)
{
    return inbox(BoxWalk(lon, lat, L, Li, Ui, Lj, Uj), xo, yo);
}

// LL2ij (interpolation_module.f90:110-225) for one point
__device__ void ll2ij(const double* lon, const double* lat, const double* angle, int L, int M,
                      int spherical, double inlon, double inlat, double& Ipar, double& Jpar)
{
    const double pi = 3.14159265358979323846, deg2rad = pi / 180.0, Eradius = 6371315.0;
    int Imin = 1, Imax = L, Jmin = 1, Jmax = M;
    bool found = inbox(lon, lat, L, Imin, Imax, Jmin, Jmax, inlon, inlat);
    if (found) {
        while (((Imax - Imin) > 1) || ((Jmax - Jmin) > 1)) {
            if ((Imax - Imin) > 1) {
                const int i0 = (Imin + Imax) / 2;
                if (inbox(lon, lat, L, Imin, i0, Jmin, Jmax, inlon, inlat)) Imax = i0;
                else Imin = i0;
            }
            if ((Jmax - Jmin) > 1) {
                const int j0 = (Jmin + Jmax) / 2;
                if (inbox(lon, lat, L, Imin, Imax, Jmin, j0, inlon, inlat)) Jmax = j0;
                else Jmin = j0;
            }
        }
    }
    found = inbox(lon, lat, L, Imin, Imax, Jmin, Jmax, inlon, inlat);
    auto at = [&](const double* f, int i, int j) { return f[(int64_t)(j - 1) * L + (i - 1)]; };
    if (spherical == 0) {
        const double xmin = at(lon, Imin, Jmin), xmax = at(lon, Imax, Jmax);
        const double ymin = at(lat, Imin, Jmin), ymax = at(lat, Imax, Jmax);
        const double xfac = ((double)Imax - (double)Imin) / (xmax - xmin);
        const double yfac = ((double)Jmax - (double)Jmin) / (ymax - ymin);
        const double dx = (inlon - xmin) * xfac;
        const double dy = (inlat - ymin) * yfac;
        Ipar = (double)Imin + dx;
        Jpar = (double)Jmin + dy;
    } else if (found) {
        const double yfac = Eradius * deg2rad;
        const double xfac = yfac * cos(inlat * deg2rad);
        const double xpp = (inlon - at(lon, Imin, Jmin)) * xfac;
        const double ypp = (inlat - at(lat, Imin, Jmin)) * yfac;
        double t1 = (at(lon, Imin + 1, Jmin) - at(lon, Imin, Jmin + 1)) * xfac;
        double t2 = (at(lat, Imin + 1, Jmin) - at(lat, Imin, Jmin + 1)) * yfac;
        const double diag2 = t1 * t1 + t2 * t2;
        t1 = (at(lon, Imin, Jmin) - at(lon, Imin + 1, Jmin)) * xfac;
        t2 = (at(lat, Imin, Jmin) - at(lat, Imin + 1, Jmin)) * yfac;
        const double aa2 = t1 * t1 + t2 * t2;
        t1 = (at(lon, Imin, Jmin) - at(lon, Imin, Jmin + 1)) * xfac;
        t2 = (at(lat, Imin, Jmin) - at(lat, Imin, Jmin + 1)) * yfac;
        const double bb2 = t1 * t1 + t2 * t2;
        const double phi = asin((diag2 - aa2 - bb2) / (2.0 * sqrt(aa2 * bb2)));
        const double ang = angle ? at(angle, Imin, Jmin) : 0.0;
        double dx = xpp * cos(ang) + ypp * sin(ang);
        double dy = ypp * cos(ang) - xpp * sin(ang);
        dx = dx + dy * tan(phi);
        dy = dy / cos(phi);
        dx = fmin(fmax(0.0, dx / sqrt(aa2)), 1.0);
        dy = fmin(fmax(0.0, dy / sqrt(bb2)), 1.0);
        Ipar = (double)Imin + dx;
        Jpar = (double)Jmin + dy;
    } else {
        Ipar = -1.0;
        Jpar = -1.0;
    }
}

}  // namespace

template <typename F>
__global__ void __launch_bounds__(256) locate_kernel(const DGrid* __restrict__ grids, LocateArgs A)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < A.n;
         q += (int64_t)gridDim.x * blockDim.x) {
        double x = 1.0, y = 1.0, z = A.z[q], st = 9.0;
        int gid = A.ngrid;
        bool isin = false;
        for (int ng = 1; ng <= A.ngrid; ng++) {
            const DGrid& g = grids[ng - 1];
            double I, J;
            ll2ij(A.lon[ng - 1], A.lat[ng - 1], A.angle[ng - 1], g.L, g.M, A.spherical, A.inlon[q],
                  A.inlat[q], I, J);
            bool ingrid, obound;
            bounds(g, I, J, ingrid, obound);
            if (!isin && ingrid) {
                x = I;
                y = J;
                gid = ng;
                st = 0.0;
                isin = true;
                // zbounds: depth bilinear, zeta (IDW) at slot t_b
                bool inz = false;
                if ((int)floor(I) > 0) {
                    const RhoCell c = rho_cell(g, I, J);
                    const double tdepth = -1.0 * bilinear2d(g.h, c);
                    SlotSel S;
                    S.s[0] = S.s[1] = S.s[2] = A.tb_lane;
                    double z3[3];
                    zeta3<F>(g, c, S, z3);
                    const double tzeta = z3[0];
                    if ((z < tzeta) && (z > tdepth)) inz = true;
                    if (z >= tzeta) {
                        z = tzeta - 0.01;
                        inz = true;
                    }
                }
                if (!inz) st = 9.0;
            }
        }
        A.x[q] = x;
        A.y[q] = y;
        A.z[q] = z;
        A.status[q] = st;
        A.gid[q] = gid;
        A.isin[q] = isin ? 1 : 0;
    }
}

// The SaltTempOn block of update_particles (ROMSPath.f90:958-981): zeta of
// slots b, c, f, getInterp3d("salt"/"temp", ztype 1) with each slot's own
// zeta and vertical bracket (interpolation_module.f90:509-589, 769-830),
// polintd at ix(2).
template <typename F>
__global__ void __launch_bounds__(256) tracer_kernel(const DGrid* __restrict__ grids, TracerArgs A)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < A.n;
         q += (int64_t)gridDim.x * blockDim.x) {
        uint32_t r = 0;
        if (A.mode == 1) {
            r = A.id[q];
            A.last_age[r] = A.age[q];
            if (!A.isin[q]) continue;
        } else if (A.mode == 2) {
            r = A.id[q];
            const double age = A.age[q];
            if (age == A.last_age[r]) continue;
            A.last_age[r] = age;
        }
        const int gi = (int)A.gid[q] - 1;
        const DGrid& g = grids[gi];
        const double X = A.x[q], Y = A.y[q], Z = A.z[q];
        const RhoCell c = rho_cell(g, X, Y);
        const double tdepth = A.mode == 2 ? -1.0 * A.bath[q] : -1.0 * bilinear2d(g.h, c);
        double zt[3];
        zeta3<F>(g, c, A.slots, zt);
        double cff[4];
        rho_weights(g, c, cff);
        double es[3], et[3];
        for (int t = 0; t < 3; t++) {
            const VBracket vb = bracket_rho<false>(g, Z, zt[t], tdepth);
            SlotSel S;
            S.s[0] = S.s[1] = S.s[2] = A.slots.s[t];
            // -DWETDRY: the slot's wet/dry mask in place of the static one
            // (interpolation_module.f90:774-830)
            double cw[4] = {cff[0], cff[1], cff[2], cff[3]};
            if (g.wdr) rho_weights_wet(g, c, A.slots.s[t], cw);
            double o[3];
            blend3<F, 1>(A.salt[gi], nullptr, c.o, g.N, cw, 0u, vb, S, o);
            es[t] = o[0];
            blend3<F, 1>(A.temp[gi], nullptr, c.o, g.N, cw, 0u, vb, S, o);
            et[t] = o[0];
        }
        const double ps = polintd(A.ex, es[0], es[1], es[2], A.ix2);
        const double pt = polintd(A.ex, et[0], et[1], et[2], A.ix2);
        if (A.mode == 1) {
            // the release sample also sets pWD, pZeta, pBath at t_c
            A.wd[q] = (-1.0 * tdepth) + zt[1];
            A.pzeta[q] = zt[1];
            A.bath[q] = -1.0 * tdepth;
        }
        if (A.mode == 0) {
            A.out_salt[q] = ps;
            A.out_temp[q] = pt;
            continue;
        }
        A.p_salt[r] = ps;
        A.p_temp[r] = pt;
        if (A.mean_salt) {
            A.mean_salt[r] = A.mode == 1 ? ps : A.mean_salt[r] + ps;
            A.mean_temp[r] = A.mode == 1 ? pt : A.mean_temp[r] + pt;
        }
    }
}

template __global__ void tracer_kernel<float>(const DGrid*, TracerArgs);
template __global__ void tracer_kernel<double>(const DGrid*, TracerArgs);

// getInterpStr (interpolation_module.f90:964-1174) at slot s, with the
// slot's wet/dry masks in place of the static ones when the grid has them
// (-DWETDRY, :1022-1060, 1103-1140):
// var 0 bustr on u points (I shifted by 1/2), var 1 bvstr on v points (J
// shifted, Jnode clamped to >= 1, and the reference's Inode = eta_v when
// Jnode > eta_v kept).  Masked corners give 0, an all-land cell 0.  Node
// indices are clamped into the arrays for the reads only, which changes
// nothing for a particle inside the grid.
template <typename F>
__device__ double interp_str(const DGrid& g, const F* __restrict__ f, int var, int s, double Ipar,
                             double Jpar)
{
    int Inode = (int)floor(Ipar), Jnode = (int)floor(Jpar);
    double X = Ipar - Inode, Y = Jpar - Jnode;
    double v[4], m[4];
    if (var == 0) {
        const int xiu = g.L - 1;
        const double uvI = Ipar - 0.5;
        Inode = (int)floor(uvI);
        X = uvI - Inode;
        const int j0 = min(max(Jnode, 1), g.M), j1 = min(max(Jnode + 1, 1), g.M);
        auto F_ = [&](int i, int j) { return (double)f[4 * ((int64_t)(i - 1) + (int64_t)xiu * (j - 1)) + s]; };
        auto M_ = [&](int i, int j) {
            const int64_t p = (int64_t)(i - 1) + (int64_t)xiu * (j - 1);
            return (double)(g.wdu ? g.wdu[(int64_t)s * g.wdp + p] : g.mu[p]);
        };
        if (Inode >= xiu) {
            v[0] = v[1] = F_(xiu, j0);
            v[2] = v[3] = F_(xiu, j1);
            m[0] = M_(xiu, j0); m[1] = 0.0; m[2] = 0.0; m[3] = M_(xiu, j1);
        } else if (Inode < 1) {
            v[0] = v[1] = F_(1, j0);
            v[2] = v[3] = F_(1, j1);
            m[0] = 0.0; m[1] = M_(1, j0); m[2] = M_(1, j1); m[3] = 0.0;
        } else {
            v[0] = F_(Inode, j0); v[1] = F_(Inode + 1, j0);
            v[2] = F_(Inode + 1, j1); v[3] = F_(Inode, j1);
            m[0] = M_(Inode, j0); m[1] = M_(Inode + 1, j0);
            m[2] = M_(Inode + 1, j1); m[3] = M_(Inode, j1);
        }
    } else {
        const int etav = g.M - 1;
        const double uvJ = Jpar - 0.5;
        Jnode = (int)floor(uvJ);
        if (Jnode < 1) Jnode = 1;
        if (Jnode > etav) Inode = etav;
        Y = uvJ - Jnode;
        const int i0 = min(max(Inode, 1), g.L), i1 = min(max(Inode + 1, 1), g.L);
        auto F_ = [&](int i, int j) { return (double)f[4 * ((int64_t)(i - 1) + (int64_t)g.L * (j - 1)) + s]; };
        auto M_ = [&](int i, int j) {
            const int64_t p = (int64_t)(i - 1) + (int64_t)g.L * (j - 1);
            return (double)(g.wdv ? g.wdv[(int64_t)s * g.wdp + p] : g.mv[p]);
        };
        if (Jnode >= etav) {
            v[0] = v[3] = F_(i0, etav);
            v[1] = v[2] = F_(i1, etav);
            m[0] = M_(i0, etav); m[1] = M_(i1, etav); m[2] = 0.0; m[3] = 0.0;
        } else {
            v[0] = F_(i0, Jnode); v[1] = F_(i1, Jnode);
            v[2] = F_(i1, Jnode + 1); v[3] = F_(i0, Jnode + 1);
            m[0] = M_(i0, Jnode); m[1] = M_(i1, Jnode);
            m[2] = M_(i1, Jnode + 1); m[3] = M_(i0, Jnode + 1);
        }
    }
    double cff[4] = {(1 - X) * (1 - Y), X * (1 - Y), X * Y, (1 - X) * Y};
    const double nwater = m[0] + m[1] + m[2] + m[3];
    if (nwater < 4.0 && nwater > 0.0) {
        for (int k = 0; k < 4; k++)
            if (m[k] == 0.0) v[k] = 0.0;
    } else if (nwater == 0.0) {
        for (int k = 0; k < 4; k++) { cff[k] = 0.0; v[k] = 0.0; }
    }
    return cff[0] * v[0] + cff[1] * v[1] + cff[2] * v[2] + cff[3] * v[3];
}

// The WriteBottom block of update_particles (ROMSPath.f90:935-951) for the
// particles that took the last step, i.e. whose age changed since the last
// call (the loop adds idt to the age of every particle it does not skip and
// skips none after that): btemp = -tdepth + Z from the step's depth and
// position, P_HOB = min(btemp, P_HOB), and where btemp is the new minimum
// P_bustr / P_bvstr = polintd over slots b, c, f of getInterpStr at ix(2).
template <typename F>
__global__ void __launch_bounds__(256) bottom_kernel(const DGrid* __restrict__ grids, BottomArgs A)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < A.n;
         q += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t r = A.id[q];
        const double age = A.age[q];
        if (age == A.last_age[r]) continue;
        A.last_age[r] = age;
        const double btemp = A.bath[q] + A.z[q];
        const double hob = fmin(btemp, A.hob[r]);
        A.hob[r] = hob;
        if (btemp != hob) continue;
        const int gi = (int)A.gid[q] - 1;
        const DGrid& g = grids[gi];
        const double X = A.x[q], Y = A.y[q];
        const F* bu = (const F*)A.bustr[gi];
        const F* bv = (const F*)A.bvstr[gi];
        const double u0 = interp_str(g, bu, 0, A.slots.s[0], X, Y);
        const double u1 = interp_str(g, bu, 0, A.slots.s[1], X, Y);
        const double u2 = interp_str(g, bu, 0, A.slots.s[2], X, Y);
        A.bustr_out[r] = polintd(A.ex, u0, u1, u2, A.ix2);
        const double v0 = interp_str(g, bv, 1, A.slots.s[0], X, Y);
        const double v1 = interp_str(g, bv, 1, A.slots.s[1], X, Y);
        const double v2 = interp_str(g, bv, 1, A.slots.s[2], X, Y);
        A.bvstr_out[r] = polintd(A.ex, v0, v1, v2, A.ix2);
    }
}

template __global__ void bottom_kernel<float>(const DGrid*, BottomArgs);
template __global__ void bottom_kernel<double>(const DGrid*, BottomArgs);

template __global__ void locate_kernel<float>(const DGrid*, LocateArgs);
template __global__ void locate_kernel<double>(const DGrid*, LocateArgs);

}  // namespace rp
