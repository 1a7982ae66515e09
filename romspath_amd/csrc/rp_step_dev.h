// rp_step_dev.h -- device code of the particle step, shared by the step
// kernels (rp_step.hip) and the adjoint kernels (rp_adjoint.hip).  Replaces
// update_particles (ROMSPath.f90:662-1003) and everything it calls for
// Behavior=0.
#pragma once

#include "rp_dual.h"
#include "rp_kernels.h"


namespace rp {

// polintd(ex, y, ix(2)) (interpolation_module.f90:72-109).  The FP64 form
// evaluates the reference's Newton formula; the FP32-field form uses the
// Lagrange weights of the same quadratic, computed once per launch.
template <typename F, class R>
__device__ __forceinline__ R tinterp(const StepArgs& A, R y0, R y1, R y2)
{
    if constexpr (is_cur<F>()) return y0;   // current fields: already at ix(2)
    if (sizeof(F) == 8) return polintd(A.ex, y0, y1, y2, A.ix[1]);
    return A.pw[0] * y0 + A.pw[1] * y1 + A.pw[2] * y2;
}

// ---------------------------------------------------------------------------
// horizontal stencil of a rho-point field: floor cell, bilinear weights, the
// four corner offsets (I,J) (I+1,J) (I+1,J+1) (I,J+1) in an (L, M) array.
// ---------------------------------------------------------------------------
template <class R = double>
struct RhoCellT {
    int I, J;
    R x, y;
    int64_t o[4];
};
using RhoCell = RhoCellT<double>;

// Corner indices are clamped into the grid.  For positions inside the grid
// this changes nothing; it keeps every read in bounds where the reference
// indexes outside its arrays (after a failed nested hand-off,
// ROMSPath.f90:837-843, or for a non-finite position).
template <class R>
__device__ __forceinline__ RhoCellT<R> rho_cell(const DGrid& g, R X, R Y)
{
    RhoCellT<R> c;
    c.I = (int)floor(val(X));
    c.J = (int)floor(val(Y));
    c.x = X - (double)c.I;
    c.y = Y - (double)c.J;
    const int i0 = min(max(c.I, 1), g.L), i1 = min(max(c.I + 1, 1), g.L);
    const int j0 = min(max(c.J, 1), g.M), j1 = min(max(c.J + 1, 1), g.M);
    c.o[0] = (int64_t)(j0 - 1) * g.L + (i0 - 1);
    c.o[1] = (int64_t)(j0 - 1) * g.L + (i1 - 1);
    c.o[2] = (int64_t)(j1 - 1) * g.L + (i1 - 1);
    c.o[3] = (int64_t)(j1 - 1) * g.L + (i0 - 1);
    return c;
}

// getInterp2D("depth"|"pm"|"pn") -- plain bilinear (interpolation_module.f90:421-507)
template <class R>
__device__ __forceinline__ R bilinear2d(const double* f, const RhoCellT<R>& c)
{
    R w[4];
    bilin(c.x, c.y, w);
    return w[0] * f[c.o[0]] + w[1] * f[c.o[1]] + w[2] * f[c.o[2]] + w[3] * f[c.o[3]];
}

// getInterp2D("zeta") at slots b, c, f: always inverse-distance weighted
// (getcoeff is called unconditionally, interpolation_module.f90:486).
template <typename F, class R>
__device__ __forceinline__ void zeta3(const DGrid& g, const RhoCellT<R>& c, const SlotSel& S,
                                      R out[3])
{
    double m[4];
    R w[4];
    for (int k = 0; k < 4; k++) m[k] = (double)g.mrho[c.o[k]];
    getcoeff(c.x, c.y, m, w);
    if constexpr (is_cur<F>()) {
        const float* z = g.czeta;
        out[0] = w[0] * (double)z[c.o[0]] + w[1] * (double)z[c.o[1]] + w[2] * (double)z[c.o[2]] +
                 w[3] * (double)z[c.o[3]];
        out[1] = out[2] = out[0];
        return;
    } else {
    const F* z = (const F*)g.zeta;
    Slots4<F> v[4];
    for (int k = 0; k < 4; k++) ld4(z + c.o[k] * 4, v[k]);
    for (int t = 0; t < 3; t++) {
        const int s = S.s[t];
        out[t] = w[0] * pick(v[0], s) + w[1] * pick(v[1], s) + w[2] * pick(v[2], s) +
                 w[3] * pick(v[3], s);
    }
    }
}

// ---------------------------------------------------------------------------
// getInterp3d for u, v, w at the slots b, c, f (interpolation_module.f90:509-962)
// The vertical bracket is shared by every slot: the reference recomputes it
// per call from the same (z, zeta, depth), so the values are identical.
// ---------------------------------------------------------------------------
// zero: bit 4t+k set where corner k of slot t (S.s[t]) reads as 0 (a dry
// or masked corner); cff: the corner weights, the same for every slot
// (CS = 0) or slot t's at cff + 4t (CS = 4, per-slot wet/dry masks)
template <typename F, int NSLOT, int CS = 0, class R>
__device__ __forceinline__ void blend3(const void* fv, const float* fc, const int64_t col[4], int nk,
                                       const R* cff, uint32_t zero, const VBracketT<R>& vb,
                                       const SlotSel& S, R out[3])
{
    if constexpr (is_cur<F>()) {
        // one value per point: the current field fc
        R zb = 0.0, zt = 0.0;
        for (int k = 0; k < 4; k++) {
            const bool z = (zero >> k) & 1u;
            const double b = z ? 0.0 : (double)fc[col[k] * nk + (vb.ibot - 1)];
            const double a = z ? 0.0 : (double)fc[col[k] * nk + (vb.itop - 1)];
            zb = zb + cff[k] * b;
            zt = zt + cff[k] * a;
        }
        out[0] = out[1] = out[2] = vb.c0 * zb + vb.c1 * zt;
        return;
    } else {
    const F* f = (const F*)fv;
    R zb[3] = {0.0, 0.0, 0.0}, zt[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 4; k++) {
        Slots4<F> lb, lt;
        ld4(f + (col[k] * nk + (vb.ibot - 1)) * 4, lb);
        ld4(f + (col[k] * nk + (vb.itop - 1)) * 4, lt);
#pragma unroll
        for (int t = 0; t < NSLOT; t++) {
            const int s = S.s[t];
            const R* c = cff + CS * t;
            const bool z = (zero >> (4 * t + k)) & 1u;
            const double b = z ? 0.0 : pick(lb, s);
            const double a = z ? 0.0 : pick(lt, s);
            if (k == 0) {
                zb[t] = c[0] * b;
                zt[t] = c[0] * a;
            } else {
                zb[t] = zb[t] + c[k] * b;
                zt[t] = zt[t] + c[k] * a;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NSLOT; t++) out[t] = vb.c0 * zb[t] + vb.c1 * zt[t];
    }
}

// the same corner flags for every slot
__device__ __forceinline__ uint32_t zero_all_slots(const bool zero[4])
{
    const uint32_t z = (uint32_t)zero[0] | (uint32_t)zero[1] << 1 | (uint32_t)zero[2] << 2 | (uint32_t)zero[3] << 3;
    return z | z << 4 | z << 8;
}

// The reference built with -DWETDRY (interpolation_module.f90:628-665,
// 710-747, 1022-1140): each slot's wet/dry mask in place of the static one.
// Corner k of slot t reads as 0 where the slot's mask at point pt[k] is 0,
// or where pt[k] < 0 (the corners the edge rows zero whatever the masks).
// Zeroing the dry corners with the weights kept is the reference's rule for
// 0 < nwater < 4, and gives 0 for nwater = 0 as its cff = 0 does.
template <int NSLOT>
__device__ __forceinline__ uint32_t zero_wet(const uint8_t* wd, int64_t wdp, const SlotSel& S,
                                             const int64_t pt[4])
{
    uint32_t z = 0;
#pragma unroll
    for (int t = 0; t < NSLOT; t++)
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (pt[k] < 0 || wd[(int64_t)S.s[t] * wdp + pt[k]] == 0) z |= 1u << (4 * t + k);
    return z;
}

template <typename F, int NSLOT, class R>
__device__ __forceinline__ void interp_u(const DGrid& g, R X, R Y, const VBracketT<R>& vb,
                                        const SlotSel& S, R out[3])
{
    const int Jf = (int)floor(val(Y));
    const R y = Y - (double)Jf;
    const int Jnode = min(max(Jf, 1), g.M - 1);   // in-range positions: unchanged
    const R uvI = X - 0.5;
    const int Inode = (int)floor(val(uvI));
    const R x = uvI - (double)Inode;
    R cff[4];
    bilin(x, y, cff);
    const int xiu = g.L - 1;
    int ia, ib;
    double m[4];
    auto MU = [&](int i, int j) { return (double)g.mu[(int64_t)(j - 1) * xiu + (i - 1)]; };
    if (Inode >= xiu) {
        ia = ib = xiu;
        m[0] = MU(xiu, Jnode); m[1] = 0; m[2] = 0; m[3] = MU(xiu, Jnode + 1);
    } else if (Inode < 1) {
        ia = ib = 1;
        m[0] = 0; m[1] = MU(1, Jnode); m[2] = MU(1, Jnode + 1); m[3] = 0;
    } else {
        ia = Inode; ib = Inode + 1;
        m[0] = MU(ia, Jnode); m[1] = MU(ib, Jnode); m[2] = MU(ib, Jnode + 1); m[3] = MU(ia, Jnode + 1);
    }
    const double nwater = m[0] + m[1] + m[2] + m[3];
    bool zero[4];
    if (nwater < 4.0 && nwater > 0.0) {
        for (int k = 0; k < 4; k++) zero[k] = (m[k] == 0.0);
    } else if (nwater == 0.0) {
        for (int k = 0; k < 4; k++) { zero[k] = true; cff[k] = 0.0; }
    } else {
        for (int k = 0; k < 4; k++) zero[k] = false;
    }
    int64_t col[4];
    col[0] = (int64_t)(Jnode - 1) * xiu + (ia - 1);
    col[1] = (int64_t)(Jnode - 1) * xiu + (ib - 1);
    col[2] = (int64_t)(Jnode) * xiu + (ib - 1);
    col[3] = (int64_t)(Jnode) * xiu + (ia - 1);
    if (wet_code<F>() && g.wdu) {
        // the slots' masks alone (the static ones take no part)
        const bool east = Inode >= xiu, west = Inode < 1;
        R cw[4];
        bilin(x, y, cw);
        if constexpr (is_cur<F>()) {
            // the current field holds each slot's value where its mask is
            // wet and 0 where dry (zeroing a slot's dry corners is linear in
            // its values); the edge corners are zeroed here
            const bool edge[4] = {west, east, east, west};
            blend3<F, NSLOT>(g.u, g.cu, col, g.N, cw, zero_all_slots(edge), vb, S, out);
        } else {
            const int64_t pt[4] = {west ? -1 : col[0], east ? -1 : col[1], east ? -1 : col[2],
                                   west ? -1 : col[3]};
            blend3<F, NSLOT>(g.u, g.cu, col, g.N, cw, zero_wet<NSLOT>(g.wdu, g.wdp, S, pt), vb, S, out);
        }
        return;
    }
    blend3<F, NSLOT>(g.u, g.cu, col, g.N, cff, zero_all_slots(zero), vb, S, out);
}

template <typename F, int NSLOT, class R>
__device__ __forceinline__ void interp_v(const DGrid& g, R X, R Y, const VBracketT<R>& vb,
                                        const SlotSel& S, R out[3])
{
    int Inode = (int)floor(val(X));
    const R x = X - (double)Inode;
    const R uvJ = Y - 0.5;
    int Jnode = (int)floor(val(uvJ));
    if (Jnode < 1) Jnode = 1;
    const int etav = g.M - 1;
    if (Jnode > etav) Inode = etav;   // interpolation_module.f90:696 assigns Inode
    Inode = min(max(Inode, 1), g.L - 1);   // in-range positions: unchanged
    const R y = uvJ - (double)Jnode;
    R cff[4];
    bilin(x, y, cff);
    int ja, jb;
    double m[4];
    auto MV = [&](int i, int j) { return (double)g.mv[(int64_t)(j - 1) * g.L + (i - 1)]; };
    if (Jnode >= etav) {
        ja = jb = etav;
        m[0] = MV(Inode, etav); m[1] = MV(Inode + 1, etav); m[2] = 0; m[3] = 0;
    } else {
        ja = Jnode; jb = Jnode + 1;
        m[0] = MV(Inode, ja); m[1] = MV(Inode + 1, ja); m[2] = MV(Inode + 1, jb); m[3] = MV(Inode, jb);
    }
    const double nwater = m[0] + m[1] + m[2] + m[3];
    bool zero[4];
    if (nwater < 4.0 && nwater > 0.0) {
        for (int k = 0; k < 4; k++) zero[k] = (m[k] == 0.0);
    } else if (nwater == 0.0) {
        for (int k = 0; k < 4; k++) { zero[k] = true; cff[k] = 0.0; }
    } else {
        for (int k = 0; k < 4; k++) zero[k] = false;
    }
    int64_t col[4];
    col[0] = (int64_t)(ja - 1) * g.L + (Inode - 1);
    col[1] = (int64_t)(ja - 1) * g.L + (Inode);
    col[2] = (int64_t)(jb - 1) * g.L + (Inode);
    col[3] = (int64_t)(jb - 1) * g.L + (Inode - 1);
    if (wet_code<F>() && g.wdv) {
        const bool north = Jnode >= etav;
        R cw[4];
        bilin(x, y, cw);
        if constexpr (is_cur<F>()) {
            // as in interp_u: dry corners are 0 in the current field
            const bool edge[4] = {false, false, north, north};
            blend3<F, NSLOT>(g.v, g.cv, col, g.N, cw, zero_all_slots(edge), vb, S, out);
        } else {
            const int64_t pt[4] = {col[0], col[1], north ? -1 : col[2], north ? -1 : col[3]};
            blend3<F, NSLOT>(g.v, g.cv, col, g.N, cw, zero_wet<NSLOT>(g.wdv, g.wdp, S, pt), vb, S, out);
        }
        return;
    }
    blend3<F, NSLOT>(g.v, g.cv, col, g.N, cff, zero_all_slots(zero), vb, S, out);
}

// masked rho-point weights for w and AKs: bilinear, or IDW when a corner is dry
template <class R>
__device__ __forceinline__ void rho_weights(const DGrid& g, const RhoCellT<R>& c, R cff[4])
{
    double m[4];
    for (int k = 0; k < 4; k++) m[k] = (double)g.mrho[c.o[k]];
    bilin(c.x, c.y, cff);
    if (m[0] + m[1] + m[2] + m[3] < 4.0) getcoeff(c.x, c.y, m, cff);
}

// the same weights from slot s's wet/dry mask of the rho points (-DWETDRY,
// interpolation_module.f90:837-923)
template <class R>
__device__ __forceinline__ void rho_weights_wet(const DGrid& g, const RhoCellT<R>& c, int s, R cff[4])
{
    double m[4];
    for (int k = 0; k < 4; k++) m[k] = (double)g.wdr[(int64_t)s * g.wdp + c.o[k]];
    bilin(c.x, c.y, cff);
    if (m[0] + m[1] + m[2] + m[3] < 4.0) getcoeff(c.x, c.y, m, cff);
}

// w in the FP32 current-field form where the slots' wet/dry masks differ at
// the cell's corners: each slot weighed with its own mask from the FP32 slot
// fields, the slots combined with the current fields' time weights pw
// (slots (b, b, c) when NSLOT is 2).  One slot at a time, in a loop kept
// rolled: such cells are few, and the plain path's registers stay its own.
template <int NSLOT>
__device__ __forceinline__ double w_wet_slots(const DGrid& g, const RhoCell& c, const VBracketT<double>& vb,
                                              const SlotSel& S, const double* pw)
{
    double v = 0.0;
#pragma unroll 1
    for (int t = 0; t < NSLOT; t++) {
        const int s = t == 0 ? S.s[0] : (t == 1 ? S.s[1] : S.s[2]);
        const double wt = NSLOT == 2 ? (t == 0 ? pw[0] + pw[1] : pw[2]) : (t == 0 ? pw[0] : (t == 1 ? pw[1] : pw[2]));
        double cw[4], o[3];
        rho_weights_wet(g, c, s, cw);
        const SlotSel one{{s, s, s}};
        blend3<float, 1>(g.w, g.cw, c.o, g.N + 1, cw, 0u, vb, one, o);
        v += wt * o[0];
    }
    return v;
}

// pw: the time weights of the current fields (w_wet_slots)
template <typename F, int NSLOT, class R>
__device__ __forceinline__ void interp_w(const DGrid& g, const RhoCellT<R>& c,
                                        const VBracketT<R>& vb, const SlotSel& S, const double pw[3],
                                        R out[3])
{
    if (wet_code<F>() && g.wdr) {
        if constexpr (is_cur<F>()) {
            // Where the slots' masks agree at the four corners every slot
            // has the same weights, and the current field (the slots
            // combined) gives the weighted sum of the slots; elsewhere
            // w_wet_slots weighs each slot with its own mask.
            bool same = true;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint8_t m0 = g.wdr[(int64_t)S.s[0] * g.wdp + c.o[k]];
#pragma unroll
                for (int t = 1; t < NSLOT; t++) same = same && g.wdr[(int64_t)S.s[t] * g.wdp + c.o[k]] == m0;
            }
            if (same) {
                R cw[4];
                rho_weights_wet(g, c, S.s[0], cw);
                blend3<F, NSLOT>(g.w, g.cw, c.o, g.N + 1, cw, 0u, vb, S, out);
            } else {
                out[0] = out[1] = out[2] = w_wet_slots<NSLOT>(g, c, vb, S, pw);
            }
        } else {
            R cw[4 * NSLOT];
#pragma unroll
            for (int t = 0; t < NSLOT; t++) rho_weights_wet(g, c, S.s[t], cw + 4 * t);
            blend3<F, NSLOT, 4>(g.w, g.cw, c.o, g.N + 1, cw, 0u, vb, S, out);
        }
        return;
    }
    R cff[4];
    rho_weights(g, c, cff);
    blend3<F, NSLOT>(g.w, g.cw, c.o, g.N + 1, cff, 0u, vb, S, out);
}

// ---------------------------------------------------------------------------
// Velocity perturbation (no reference counterpart; zero unless a caller sets
// it): du, dv (m/s) from uniform cubic B-splines on an nx x ny control net
// spread over the grid's rho-index domain [1, L] x [1, M], added to the
// interpolated u, v of every RK stage.  Coefficients [2][ny][nx] (u then v).
// ---------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ void bspline4(R t, R b[4])
{
    const R t2 = t * t, t3 = t2 * t, omt = 1.0 - t;
    b[0] = omt * omt * omt / 6.0;
    b[1] = (3.0 * t3 - 6.0 * t2 + 4.0) / 6.0;
    b[2] = (-3.0 * t3 + 3.0 * t2 + 3.0 * t + 1.0) / 6.0;
    b[3] = t3 / 6.0;
}

// knot span (first of the four coefficients) and local coordinate of X
template <class R>
__device__ __forceinline__ int bspline_span(R X, int L, int nc, R& t)
{
    const R xi = (X - 1.0) * ((double)(nc - 3) / (double)(L - 1));
    int sp = (int)floor(val(xi));
    sp = min(max(sp, 0), nc - 4);
    t = xi - (double)sp;
    return sp;
}

template <class R>
__device__ __forceinline__ void add_perturbation(const DGrid& g, const StepArgs& A, R X, R Y, R& U,
                                                 R& V)
{
    R tx, ty, bx[4], by[4];
    const int sx = bspline_span(X, g.L, A.pert_nx, tx);
    const int sy = bspline_span(Y, g.M, A.pert_ny, ty);
    bspline4(tx, bx);
    bspline4(ty, by);
    const double* cu = A.pert;
    const double* cv = A.pert + (int64_t)A.pert_nx * A.pert_ny;
    R du = 0.0, dv = 0.0;
    for (int b = 0; b < 4; b++)
        for (int a = 0; a < 4; a++) {
            const R w = bx[a] * by[b];
            const int64_t k = (int64_t)(sy + b) * A.pert_nx + (sx + a);
            du = du + w * cu[k];
            dv = dv + w * cv[k];
        }
    U = U + du;
    V = V + dv;
}

// ---------------------------------------------------------------------------
// find_currents (advection_module.f90:129-350), version=2: all stages use
// time ix(2).  Z is clamped in place, as the reference's INTENT(INOUT) Zpar.
// ---------------------------------------------------------------------------
template <class R = double>
struct SurfaceT {
    R tdepth, zb, zc, zf, zeta;
};
using Surface = SurfaceT<double>;

template <typename F, class R>
__device__ __forceinline__ SurfaceT<R> surface_at(const DGrid& g, const StepArgs& A,
                                                  const RhoCellT<R>& c)
{
    SurfaceT<R> s;
    s.tdepth = -1.0 * bilinear2d(g.h, c);
    R z3[3];
    zeta3<F>(g, c, A.slots, z3);
    s.zb = z3[0]; s.zc = z3[1]; s.zf = z3[2];
    s.zeta = tinterp<F>(A, s.zb, s.zc, s.zf);
    return s;
}

// NSLOT=2 on the first external step, whose u, v, w use slots (b,b,c)
// (advection_module.f90:277-334); 3 otherwise.
// hint (may be null): the rho bracket's level index of the previous stage,
// in and out; 0 = none.  The w bracket starts from this stage's rho index.
template <typename F, int NSLOT, class R>
__device__ __forceinline__ void find_currents(const DGrid& g, const StepArgs& A, R X, R Y, R& Z,
                                              R& U, R& V, R& W, SurfaceT<R>* surf_out,
                                              int* hint = nullptr)
{
    const RhoCellT<R> c = rho_cell(g, X, Y);
    const SurfaceT<R> s = surface_at<F>(g, A, c);
    if (surf_out) *surf_out = s;
    if (Z < s.tdepth) Z = s.tdepth + (double)0.001f;
    if (Z > s.zeta) Z = s.zeta - (double)0.001f;
    const VBracketT<R> br = bracket_rho<is_fast<F>()>(g, Z, s.zeta, s.tdepth, hint ? *hint : 0);
    const VBracketT<R> bw = bracket_w<is_fast<F>()>(g, Z, s.zeta, s.tdepth, hint ? br.k : 0);
    if (hint) *hint = br.k;
    R q[3];
    interp_u<F, NSLOT>(g, X, Y, br, A.slots, q);
    U = NSLOT == 2 ? tinterp<F>(A, q[0], q[0], q[1]) : tinterp<F>(A, q[0], q[1], q[2]);
    interp_v<F, NSLOT>(g, X, Y, br, A.slots, q);
    V = NSLOT == 2 ? tinterp<F>(A, q[0], q[0], q[1]) : tinterp<F>(A, q[0], q[1], q[2]);
    interp_w<F, NSLOT>(g, c, bw, A.slots, A.pw, q);
    W = NSLOT == 2 ? tinterp<F>(A, q[0], q[0], q[1]) : tinterp<F>(A, q[0], q[1], q[2]);
    if (A.pert) add_perturbation(g, A, X, Y, U, V);
}

// ---------------------------------------------------------------------------
// deviates: injected (test hook) or Philox, in the fixed per-particle layout
// [HTurb x, HTurb y, VTurb sub-step 1..loop]
// ---------------------------------------------------------------------------
struct DevSrc {
    const double* inj;
    uint64_t seed, id, step;
    int blk;
    double buf[4];
    __device__ __forceinline__ double get(int j)
    {
        if (inj) return inj[j];
        const int b = j >> 2;
        if (b != blk) {
            normals4(seed, id, step, (uint32_t)b, buf);
            blk = b;
        }
        const int r = j & 3;
        return r == 0 ? buf[0] : (r == 1 ? buf[1] : (r == 2 ? buf[2] : buf[3]));
    }
};

// ---------------------------------------------------------------------------
// ver_turb_module.f90:31-329 with pppack.f90 smooth/setupq/chol1d/ppvalu.
// Per-lane work arrays live in a global scratch area, one column per lane
// (row r of lane t at r*T + t), so every access of a wavefront is one
// contiguous 512-byte segment.
// ---------------------------------------------------------------------------
template <class R>
struct ScrT {
    R* p;
    int64_t T;
    __device__ __forceinline__ R& operator()(int row) const { return p[(int64_t)row * T]; }
};
using Scr = ScrT<double>;

// the same lane column viewed as FP32 rows (the FP32-field form's work arrays)
// The same scratch viewed as FP32 rows of T lanes, read and written by
// buffer instructions over one descriptor: the lane's column is its vector
// offset, and for a row index that is the same across the wave, S(row), the
// row's byte offset is a scalar offset, so a row access needs no address
// arithmetic per lane.  S.at(row) takes a per-lane row.  Accesses beyond
// the descriptor's extent read 0 and are not written.
// The first kLdsQ rows of the FP32 work rows (the qty rows, which every
// secant sweep reads) are also kept in LDS, kLdsQ consecutive words per
// lane of the block (dynamic LDS of blockDim x kLdsQ x 4 bytes; the odd
// stride puts the 64 lanes of a row in distinct banks).
extern __shared__ float rp_qlds[];

struct Scr32 {
    __amdgpu_buffer_rsrc_t r;
    uint32_t voff;   // lane * 4
    uint32_t rowb;   // T * 4
    float* lq;       // this lane's LDS rows: qty row j < kLdsQ at lq[j]
    __device__ __forceinline__ Scr32(float* base, uint32_t bytes, uint32_t lane, uint32_t T)
        : r(__builtin_amdgcn_make_buffer_rsrc(base, 0, (int)bytes, 0x00020000)), voff(lane * 4u),
          rowb(T * 4u), lq(rp_qlds + threadIdx.x * kLdsQ)
    {
    }
    __device__ __forceinline__ float lq_ld(int j) const { return lq[j]; }
    struct Ref {
        const Scr32& s;
        uint32_t vo, so;
        __device__ __forceinline__ operator float() const
        {
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(s.r, vo, so, 0));
        }
        __device__ __forceinline__ const Ref& operator=(float v) const
        {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), s.r, vo, so, 0);
            return *this;
        }
    };
    __device__ __forceinline__ Ref operator()(int row) const { return Ref{*this, voff, (uint32_t)row * rowb}; }
    __device__ __forceinline__ Ref at(int row) const { return Ref{*this, voff + (uint32_t)row * rowb, 0u}; }
    // rows written once and read once later: non-temporal (cache policy NT),
    // so they do not displace the qty row the secant sweeps read eight times
    __device__ __forceinline__ void st_nt(int row, float v) const
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, (uint32_t)row * rowb, 2);
    }
    __device__ __forceinline__ void st_nt_at(int row, float v) const
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff + (uint32_t)row * rowb, 0u, 2);
    }
    __device__ __forceinline__ float ld_nt(int row) const
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, (uint32_t)row * rowb, 2));
    }
};

struct VtLayout {
    int np, nw;
    int fitz, y, h, r, qty, d, l1, l2, u, z, pkh;
    __device__ __forceinline__ VtLayout(int np_, int nw_) : np(np_), nw(nw_)
    {
        fitz = 0; y = np; h = 2 * np; r = 3 * np; qty = 4 * np; d = 5 * np; l1 = 6 * np;
        l2 = 7 * np; u = 8 * np; z = 9 * np; pkh = 9 * np + nw;
    }
};

// chol1d (pppack.f90:1428-1566) for dy(:)=serr: the setupq products are
// rebuilt row by row from r(i)=serr/h(i) (bit-identical to setupq's
// v(:,1..3,5..7)), elimination and forward substitution share one sweep.
// Returns sum_i qu(i)^2 dy(i)^2 (the dot_product of smooth) and, when asked,
// utru of the p=0 solve.
template <class R>
__device__ __forceinline__ R chol1d_dev(R p, const VtLayout& Lo, const ScrT<R>& S, double dy2,
                                        bool want_utru, R* utru)
{
    const int np = Lo.np;
    const R six1mp = 6.0 * (1.0 - p);
    const R twop = 2.0 * p;
    auto Rv = [&](int i) { return S(Lo.r + i - 1); };
    auto H = [&](int i) { return S(Lo.h + i - 1); };
    auto row = [&](int i, R& d0, R& e0, R& f0) {
        // setupq (pppack.f90:4793-4813) then chol1d's matrix (pppack.f90:1509-1512)
        const R ri = Rv(i), rim1 = Rv(i - 1);
        const R v1 = rim1, v3 = ri;
        const R v2 = (-ri) - rim1;
        const R v5 = v1 * v1 + v2 * v2 + v3 * v3;
        R v6 = 0.0, v7 = 0.0;
        if (i <= np - 2) {
            const R rip1 = Rv(i + 1);
            const R v1n = ri, v2n = (-rip1) - ri;
            v6 = v2 * v1n + v3 * v2n;
            if (i <= np - 3) v7 = v3 * rip1;
        }
        d0 = six1mp * v5 + twop * (H(i - 1) + H(i));
        e0 = six1mp * v6 + p * H(i);
        f0 = six1mp * v7;
    };
    if (np < 4) {
        R d0, e0, f0;
        row(2, d0, e0, f0);
        S(Lo.u + 0) = 0.0;
        S(Lo.u + 1) = S(Lo.qty + 1) / d0;
        S(Lo.u + 2) = 0.0;
    } else {
        R di, ei, fi, dn, en, fn, dnn, enn, fnn;
        row(2, di, ei, fi);
        row(3, dn, en, fn);
        R u_im1 = 0.0, u_i = S(Lo.qty + 1), l2_im1 = 0.0;
        S(Lo.u + 0) = 0.0;
        S(Lo.u + 1) = u_i;
        for (int i = 2; i <= np - 2; i++) {
            R rat = ei / di;
            dn = dn - rat * ei;
            en = en - rat * fi;
            const R l1 = rat;
            rat = fi / di;
            if (i + 2 <= np - 1) {
                row(i + 2, dnn, enn, fnn);
            } else {
                dnn = 0.0; enn = 0.0; fnn = 0.0;
            }
            dnn = dnn - rat * fi;
            const R l2 = rat;
            S(Lo.d + i - 1) = di;
            S(Lo.l1 + i - 1) = l1;
            S(Lo.l2 + i - 1) = l2;
            const R u_ip1 = S(Lo.qty + i) - l1 * u_i - l2_im1 * u_im1;
            S(Lo.u + i) = u_ip1;
            di = dn; ei = en; fi = fn;
            dn = dnn; en = enn; fn = fnn;
            u_im1 = u_i; u_i = u_ip1; l2_im1 = l2;
        }
        // back substitution (pppack.f90:1545-1550); di is v(np-1,1)
        S(Lo.u + np - 1) = 0.0;
        R u_p2 = 0.0;
        R u_p1 = u_i / di;
        S(Lo.u + np - 2) = u_p1;
        for (int i = np - 2; i >= 2; i--) {
            const R ui = S(Lo.u + i - 1) / S(Lo.d + i - 1) - u_p1 * S(Lo.l1 + i - 1) -
                              u_p2 * S(Lo.l2 + i - 1);
            S(Lo.u + i - 1) = ui;
            u_p2 = u_p1;
            u_p1 = ui;
        }
    }
    // qu = Q*u (pppack.f90:1556-1563) fused with smooth's dot_product and utru
    R sum = 0.0, ut = 0.0, prev = 0.0;
    R um1 = S(Lo.u + 0);
    for (int i = 2; i <= np; i++) {
        const R ui = S(Lo.u + i - 1);
        const R hi = H(i - 1);
        const R raw = (ui - um1) / hi;
        const R qu = raw - prev;
        sum = sum + (qu * qu) * dy2;
        if (want_utru) ut = ut + hi * (um1 * (um1 + ui) + ui * ui);
        prev = raw;
        um1 = ui;
    }
    const R qn = -prev;
    sum = sum + (qn * qn) * dy2;
    if (want_utru) *utru = ut;
    return sum;
}

// interv (pppack.f90:3363-3576) on the nearly uniform breakpoints fitz: a
// guess from the mean spacing, corrected against the stored breakpoints, so
// the result is the reference's: the largest i<=np-1 with fitz(i)<=x, else 1.
template <class R>
__device__ __forceinline__ int interv_dev(const VtLayout& Lo, const ScrT<R>& S, double x, double f1,
                                          double dz)
{
    const int np = Lo.np;
    const double gss = (x - f1) / dz;
    int i;
    if (!(gss >= 0.0)) i = 1;
    else if (gss >= (double)(np - 1)) i = np - 1;
    else i = (int)gss + 1;
    while (i > 1 && val(S(Lo.fitz + i - 1)) > x) i--;
    while (i < np - 1 && val(S(Lo.fitz + i)) <= x) i++;
    return i;
}

template <class R = double>
struct PieceCacheT {
    int i;
    R c[4];
    R brk;
};
using PieceCache = PieceCacheT<double>;

// Coefficients of piece i of the smoothing spline (smooth's final loop,
// pppack.f90:5305-5314), formed on demand from u, h and y.
template <class R>
__device__ __forceinline__ void spline_piece(const VtLayout& Lo, const ScrT<R>& S, int i, R p, R k2,
                                             PieceCacheT<R>& pc)
{
    const int np = Lo.np;
    auto U = [&](int j) { return S(Lo.u + j - 1); };
    auto H = [&](int j) { return S(Lo.h + j - 1); };
    auto raw = [&](int j) { return (U(j) - U(j - 1)) / H(j - 1); };   // j >= 2
    const R r_i = (i >= 2) ? raw(i) : 0.0;
    const R r_i1 = raw(i + 1);
    const R qu_i = r_i1 - r_i;
    R qu_i1;
    if (i + 1 == np) qu_i1 = -r_i1;
    else qu_i1 = raw(i + 2) - r_i1;
    const R a1 = S(Lo.y + i - 1) - k2 * qu_i;
    const R a1n = S(Lo.y + i) - k2 * qu_i1;
    const R a3 = U(i) * 6.0 * p;
    const R a3n = U(i + 1) * 6.0 * p;
    const R hi = H(i);
    const R a4 = (a3n - a3) / hi;
    const R a2 = (a1n - a1) / hi - (a3 + a4 / 3.0 * hi) / 2.0 * hi;
    pc.i = i;
    pc.c[0] = a1; pc.c[1] = a2; pc.c[2] = a3; pc.c[3] = a4;
    pc.brk = S(Lo.fitz + i - 1);
}

// The Horner loop of ppvalu (pppack.f90:4349-4357) for k=4, written out:
// value=(value/fmmjdr)*h+coef(m,i) for fmmjdr=4-jderiv..1; its first step
// yields coef(4,i) exactly and the last divides by 1.
template <class R>
__device__ __forceinline__ R horner(const PieceCacheT<R>& pc, R h, int jderiv)
{
    if (jderiv == 0) return ((pc.c[3] / 3.0 * h + pc.c[2]) / 2.0 * h + pc.c[1]) * h + pc.c[0];
    return (pc.c[3] / 2.0 * h + pc.c[2]) * h + pc.c[1];
}

// ppvalu (pppack.f90:4246-4362), k=4.  The pieces are formed in R (the
// spline's type) and evaluated in RZ (the argument's): in the adjoint the
// spline depends on the step's start X, Y only and the argument on Z too,
// so the fit runs in two-direction duals and only the sub-steps in three.
template <class R, class RZ>
__device__ __forceinline__ RZ ppvalu_dev(const VtLayout& Lo, const ScrT<R>& S, PieceCacheT<RZ>& pc,
                                         RZ x, int jderiv, R p, R k2, R f1, R dz)
{
    const int i = interv_dev(Lo, S, val(x), val(f1), val(dz));
    if (i != pc.i) {
        PieceCacheT<R> pr;
        spline_piece(Lo, S, i, p, k2, pr);
        pc.i = i;
#pragma unroll
        for (int k = 0; k < 4; k++) pc.c[k] = Conv<RZ>::from(pr.c[k]);
        pc.brk = Conv<RZ>::from(pr.brk);
    }
    const RZ h = x - pc.brk;
    return horner(pc, h, jderiv);
}

// R: the type of the column and the spline fit; RZ: of the sub-steps
template <typename F, class R, class RZ = R>
__device__ __forceinline__ RZ vturb(const DGrid& g, const StepArgs& A, int np, const RhoCellT<R>& c,
                                    const SurfaceT<R>& sf, RZ Zpar, DevSrc& dv, const ScrT<R>& S,
                                    unsigned* capped)
{
    const int nw = g.N + 1;
    const VtLayout Lo(np, nw);
    const R tdepth = sf.tdepth, zeta = sf.zeta;
    const R h = -1.0 * tdepth;
    // AKs column at the particle (getInterpAKs, interpolation_module.f90:1175-1229)
    R cff[4];
    rho_weights(g, c, cff);
    const F* aks = (const F*)g.aks;
    for (int k = 1; k <= nw; k++) {
        const R Sk = level_S(g, g.s_w, g.cs_w, k, h);
        const R zb = level_z(g, g.s_w, g.cs_w, k, sf.zb, h, Sk);
        const R zc = level_z(g, g.s_w, g.cs_w, k, sf.zc, h, Sk);
        const R zf = level_z(g, g.s_w, g.cs_w, k, sf.zf, h, Sk);
        S(Lo.z + k - 1) = polintd(A.ex, zb, zc, zf, A.ix[1]);
        Slots4<F> v[4];
        for (int q = 0; q < 4; q++) ld4(aks + (c.o[q] * nw + (k - 1)) * 4, v[q]);
        R kh[3];
        for (int t = 0; t < 3; t++) {
            const int s = A.slots.s[t];
            kh[t] = cff[0] * pick(v[0], s) + cff[1] * pick(v[1], s) + cff[2] * pick(v[2], s) +
                    cff[3] * pick(v[3], s);
        }
        S(Lo.pkh + k - 1) = polintd(A.ex, kh[0], kh[1], kh[2], A.ix[1]);
    }
    // resample to np uniform points with linint (ver_turb_module.f90:126-136;
    // interpolation_module.f90:27-61).  fitz increases, so the bracket is
    // found by a forward merge; it equals linint's bisection result.
    const R z1 = S(Lo.z + 0), zn = S(Lo.z + nw - 1);
    const R dz = (zn - z1) / (double)(np - 1);
    S(Lo.fitz + 0) = z1;
    S(Lo.fitz + np - 1) = zn;
    S(Lo.y + 0) = S(Lo.pkh + 0);
    S(Lo.y + np - 1) = S(Lo.pkh + nw - 1);
    {
        R fz = z1;
        int jlo = 1;
        R zlo = z1, zhi = S(Lo.z + 1);
        for (int i = 2; i <= np - 1; i++) {
            fz = fz + dz;
            S(Lo.fitz + i - 1) = fz;
            while (jlo < nw - 1 && zhi <= fz) {
                jlo++;
                zlo = zhi;
                zhi = S(Lo.z + jlo);
            }
            const R ylo = S(Lo.pkh + jlo - 1), yhi = S(Lo.pkh + jlo);
            const R m = (ylo - yhi) / (zlo - zhi);
            const R b = ylo - m * zlo;
            S(Lo.y + i - 1) = m * fz + b;
        }
    }
    // setupq (pppack.f90:4726-4825) with dy(:)=serr
    const double serr = A.serr;
    {
        R prev = 0.0;
        for (int i = 1; i <= np - 1; i++) {
            const R hi = S(Lo.fitz + i) - S(Lo.fitz + i - 1);
            S(Lo.h + i - 1) = hi;
            S(Lo.r + i - 1) = serr / hi;
            const R diff = (S(Lo.y + i) - S(Lo.y + i - 1)) / hi;
            if (i >= 2) S(Lo.qty + i - 1) = diff - prev;
            prev = diff;
        }
    }
    // smooth (pppack.f90:5236-5298)
    const double dy2 = serr * serr;
    const double s = A.smth;
    R p, q, sfq;
    if (s <= 0.0) {
        p = 1.0;
        chol1d_dev<R>(p, Lo, S, dy2, false, nullptr);
        sfq = 0.0;
    } else {
        p = 0.0;
        R utru = 0.0;
        sfq = 36.0 * chol1d_dev(p, Lo, S, dy2, true, &utru);
        if (s < sfq) {
            const double ooss = 1.0 / sqrt(s);
            R oosf = 1.0 / sqrt(sfq);
            q = -(oosf - ooss) * sfq / (6.0 * utru * oosf);
            R prevq = 0.0, prevsf = oosf;
            int it = 0;
            for (;;) {
                sfq = 36.0 * chol1d_dev<R>(q / (1.0 + q), Lo, S, dy2, false, nullptr) /
                      ((1.0 + q) * (1.0 + q));
                if (fabs(sfq - s) <= 0.01 * s) break;
                if (++it >= kSecantCap) {   // the reference has no cap
                    atomicAdd(capped, 1u);
                    break;
                }
                oosf = 1.0 / sqrt(sfq);
                const R change = (q - prevq) / (oosf - prevsf) * (oosf - ooss);
                prevq = q;
                q = q - change;
                prevsf = oosf;
            }
            p = q / (1.0 + q);
        }
    }
    const R k2 = 6.0 * (1.0 - p) * dy2;
    // RDM sub-steps (ver_turb_module.f90:231-293)
    PieceCacheT<RZ> pc;
    pc.i = -1;
    const double deltat = A.deltat;
    const RZ td = Conv<RZ>::from(tdepth), ze = Conv<RZ>::from(zeta);
    RZ ParZc = Zpar;
    for (int i = 1; i <= A.loop; i++) {
        RZ Kprimec;
        if (ParZc <= td || ParZc >= ze) Kprimec = 0.0;
        else Kprimec = ppvalu_dev(Lo, S, pc, ParZc, 1, p, k2, z1, dz);
        const RZ KprimeZc = Kprimec * deltat;
        const RZ Z3rdc = ParZc + 0.5 * KprimeZc;
        RZ KH3rdc;
        if (Z3rdc < td || Z3rdc > ze) {
            KH3rdc = A.aksback;
        } else {
            KH3rdc = ppvalu_dev(Lo, S, pc, Z3rdc, 0, p, k2, z1, dz);
            if (KH3rdc < A.aksback) KH3rdc = A.aksback;
        }
        const double DEV = dv.get(1 + i);
        const double r = 1.0;
        const RZ dTz = KprimeZc + DEV * sqrt(2.0 / r * KH3rdc * deltat);
        ParZc = ParZc + dTz;
        if (ParZc <= td) ParZc = td + fabs(ParZc - td);
        if (ParZc >= ze) ParZc = ze - fabs(ParZc - ze);
    }
    return ParZc - Zpar;
}

// ---------------------------------------------------------------------------
// VTurb, FP32-field form.  Same model and the same secant search for the
// smoothing parameter as above, evaluated without per-iteration work arrays:
//
//   * the resampled points are taken as exactly uniform (spacing dz; the
//     reference accumulates fitz(i)=fitz(i-1)+dz, which differs in the last
//     bits), so chol1d's matrix T(p) = 6(1-p) V + p W is pentadiagonal
//     Toeplitz, V = r^2 (Q'Q), W = h tridiag(1,4,1), r = serr/dz;
//   * smooth's sum_i (a(i,1) dy(i))^2 = u'V u for the solution of T u = qty.
//     With G(p) = qty' T(p)^-1 qty (= sum w_i^2/d_i of an LDL' sweep) and
//     u'T u = G, dG/dp = -u'(dT/dp)u = 6 u'Vu - u'Wu, it follows that
//     u'Vu = (G + p G')/6 and, at p=0, utru = u'Wu/2 = (G - G')/2.  One
//     forward LDL' sweep carrying derivatives in p therefore gives every
//     quantity the secant search needs, with no back substitution;
//   * only the accepted p gets a full solve, whose u builds the spline
//     pieces the random-displacement sub-steps visit.
// Work arrays are FP32 rows (the arithmetic stays FP64): qty, formed in FP64
// from the FP64 resampled values and rounded once, so each sweep sees qty to
// a relative 6e-8 per entry; y for the pieces; the factor rows and u.  They
// are read in chunks of kCh rows so their loads overlap.
// ---------------------------------------------------------------------------
constexpr int kCh = 8;

// Secant sweeps (0-based) from which the factor rows are kept.  With the
// spline parameters of ROMSPath.data (serr 1e-4, smth 0.6, sub 4) the search
// on resampled AKs profiles took exactly 7 sweeps in every case sampled, so
// storing from the 7th makes the final solve a back substitution.  A search
// that ends earlier falls back to the full solve and one that ends later
// stores again; the result is the same either way.
constexpr int kStoreFrom = 6;

struct FastLayout {
    int np;
    int q, y, rd, l1, w, u;
    __device__ __forceinline__ FastLayout(int np_) : np(np_)
    {
        q = 0; y = np; rd = 2 * np; l1 = 3 * np; w = 4 * np; u = 5 * np;
    }
};

// 1/d to 2.2e-15 relative: the hardware estimate (4.6e-8) and one Newton
// step (both bounds measured on gfx950 over 4M random doubles)
__device__ __forceinline__ double rcp64(double d)
{
    const double r = __builtin_amdgcn_rcp(d);
    const double e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// Toeplitz entries of chol1d's matrix (pppack.f90:1509-1512) for uniform h
struct Band {
    double a, b, c, da, db, dc;
};

__device__ __forceinline__ Band band_of(double p, double h, double serr)
{
    const double r = serr / h;
    const double v5 = 6.0 * r * r, v6 = -4.0 * r * r, v7 = r * r;
    const double six1mp = 6.0 * (1.0 - p), twop = 2.0 * p;
    Band B;
    B.a = six1mp * v5 + twop * (h + h);
    B.b = six1mp * v6 + p * h;
    B.c = six1mp * v7;
    B.da = -6.0 * v5 + 4.0 * h;
    B.db = -6.0 * v6 + h;
    B.dc = -6.0 * v7;
    return B;
}

// The stored 1/d and l1 rows in FP32.  The LDL' pivots of the Toeplitz
// band converge along the rows, and their FP32 roundings stop changing
// after a few to a few dozen rows (earlier the larger p is).  A storing
// sweep notes the first row m from which every stored pair equals the one
// before it; rows from m on then all hold the last row's pair, and the back
// substitution takes it from registers instead of reading those rows.
struct FactorTail {
    int m;
    float rd, l1;
};
struct TailTrack {
    int m;
    float rd, l1;
    __device__ __forceinline__ void init()
    {
        m = 0; rd = 0.0f; l1 = 0.0f;
    }
    __device__ __forceinline__ void put(const Scr32& S, const FastLayout& Lo, int r, double rdv,
                                        double l1v)
    {
        const float frd = (float)rdv, fl1 = (float)l1v;
        S.st_nt(Lo.rd + r, frd);
        S.st_nt(Lo.l1 + r, fl1);
        m = (frd == rd && fl1 == l1) ? m : r + 1;
        rd = frd;
        l1 = fl1;
    }
    __device__ __forceinline__ FactorTail tail() const { return FactorTail{m, rd, l1}; }
};

// One row of the sweep for G(p) and dG/dp (qn: qty of the next row)
struct GSweep {
    double d, dd, e, de, dn, ddn, en, den;
    double w, dw, wp, dwp, l2p, dl2p, g, gp;
    TailTrack tt;
    // STORE: also keep 1/d, w and l1 of the row, exactly as the forward half
    // of the full solve at this p would (same operations, same order)
    template <bool STORE = false>
    __device__ __forceinline__ void row(const Band& B, double qn, const Scr32* S = nullptr,
                                        const FastLayout* Lo = nullptr, int m = 0)
    {
#pragma clang fp contract(fast)
        const double rd = rcp64(d);
        if constexpr (STORE) {
            S->st_nt(Lo->w + m, (float)w);
            tt.put(*S, *Lo, m, rd, e * rd);
        }
        const double wrd = w * rd;
        g += w * wrd;
        // d(w^2/d)/dp = (2 w dw - w^2 dd / d) / d = wrd (2 dw - wrd dd)
        gp += wrd * ((dw + dw) - wrd * dd);
        const double rat = e * rd, drat = (de - rat * dd) * rd;
        const double rat2 = B.c * rd, drat2 = (B.dc - rat2 * dd) * rd;
        const double d1 = dn - rat * e, dd1 = ddn - drat * e - rat * de;
        const double e1 = en - rat * B.c, de1 = den - drat * B.c - rat * B.dc;
        const double d2 = B.a - rat2 * B.c, dd2 = B.da - drat2 * B.c - rat2 * B.dc;
        const double wn = qn - rat * w - l2p * wp;
        const double dwn = -drat * w - rat * dw - dl2p * wp - l2p * dwp;
        wp = w; dwp = dw; w = wn; dw = dwn;
        l2p = rat2; dl2p = drat2;
        d = d1; dd = dd1; e = e1; de = de1;
        dn = d2; ddn = dd2; en = B.b; den = B.db;
    }
};

// qty row j (wave-uniform j) from LDS or global memory
__device__ __forceinline__ float qty_at(const Scr32& S, int oq, int j)
{
    return j < kLdsQ ? S.lq_ld(j) : (float)S(oq + j);
}

// G(p) and dG/dp by one LDL' sweep over the n interior unknowns.  qty row n
// holds 0, so every chunk of rows reads the next row's qty without a guard.
// STORE: keep the factor rows for back_u (tail: where they stop changing).
template <bool STORE = false>
__device__ __forceinline__ void stream_G(const Band& B, int n, const Scr32& S, const FastLayout& Lo,
                                         double& G, double& Gp, FactorTail* tail = nullptr)
{
    const int oq = Lo.q;
    GSweep sw;
    sw.d = B.a; sw.dd = B.da; sw.e = B.b; sw.de = B.db;
    sw.dn = B.a; sw.ddn = B.da; sw.en = B.b; sw.den = B.db;
    sw.w = (double)S.lq_ld(0); sw.dw = 0.0; sw.wp = 0.0; sw.dwp = 0.0; sw.l2p = 0.0; sw.dl2p = 0.0;
    sw.g = 0.0; sw.gp = 0.0;
    if constexpr (STORE) sw.tt.init();
    int m0 = 0;
    for (; m0 + kCh < kLdsQ && m0 + kCh <= n; m0 += kCh) {
        float qn[kCh];
#pragma unroll
        for (int k = 0; k < kCh; k++) qn[k] = S.lq_ld(m0 + k + 1);
#pragma unroll
        for (int k = 0; k < kCh; k++) sw.template row<STORE>(B, (double)qn[k], &S, &Lo, m0 + k);
    }
    for (; m0 + kCh <= n; m0 += kCh) {
        float qn[kCh];
#pragma unroll
        for (int k = 0; k < kCh; k++) qn[k] = S(oq + m0 + k + 1);
#pragma unroll
        for (int k = 0; k < kCh; k++) sw.template row<STORE>(B, (double)qn[k], &S, &Lo, m0 + k);
    }
    for (; m0 < n; m0++) sw.template row<STORE>(B, (double)qty_at(S, oq, m0 + 1), &S, &Lo, m0);
    G = sw.g;
    Gp = sw.gp;
    if constexpr (STORE) *tail = sw.tt.tail();
}

// The p = 0 sweep from universal tables.  At p = 0 the band is
// s (6, -4, 1) with s = 6 r^2, and its p-derivative is -s (6, -4, 1) +
// h (4, 1, 0).  The LDL' pivots of a banded Toeplitz matrix are homogeneous
// (degree 1 for d, 0 for l1 and l2) and row m of the forward elimination
// depends only on the leading m x m block, so one table for (6, -4, 1) and
// its derivative along (4, 1, 0) serves every s, h and n:
//   rd = T1 / s, l1 = rat, l2 = rat2, d(l1)/dp = k drat, d(l2)/dp = k drat2,
//   dd/dp = -s d^ + h dd^, with k = h / s.
// With omega = (dw/dp) / k the sweep accumulates A = sum w^2 T1,
// B = sum w omega T1, C = sum w^2 T2 (T2 = rd^2 dd^), and
//   G = A / s,  dG/dp = (A + k (2 B - C)) / s.
constexpr int kP0Rows = 1024;
static __constant__ double kP0Tab[kP0Rows * 6];   // T1, T2, rat, rat2, drat, drat2

__device__ __forceinline__ void stream_G_p0(double h, double serr, int n, const Scr32& S, int oq,
                                            double& G, double& Gp)
{
#pragma clang fp contract(fast)
    const double r = serr / h;
    const double s = 6.0 * r * r;
    const double kap = h / s;
    double w = (double)S.lq_ld(0), om = 0.0, wp = 0.0, omp = 0.0;
    double A = 0.0, B = 0.0, C = 0.0, r2p = 0.0, dr2p = 0.0;
    auto row = [&](int m, double qn) {
        const double* T = kP0Tab + 6 * m;
        const double ww = w * w;
        A += ww * T[0];
        B += (w * om) * T[0];
        C += ww * T[1];
        const double wn = qn - T[2] * w - r2p * wp;
        const double omn = -(T[4] * w + dr2p * wp) - T[2] * om - r2p * omp;
        wp = w; w = wn; omp = om; om = omn;
        r2p = T[3]; dr2p = T[5];
    };
    int m0 = 0;
    for (; m0 + kCh < kLdsQ && m0 + kCh <= n; m0 += kCh) {
        float qn[kCh];
#pragma unroll
        for (int k = 0; k < kCh; k++) qn[k] = S.lq_ld(m0 + k + 1);
#pragma unroll
        for (int k = 0; k < kCh; k++) row(m0 + k, (double)qn[k]);
    }
    for (; m0 + kCh <= n; m0 += kCh) {
        float qn[kCh];
#pragma unroll
        for (int k = 0; k < kCh; k++) qn[k] = S(oq + m0 + k + 1);
#pragma unroll
        for (int k = 0; k < kCh; k++) row(m0 + k, (double)qn[k]);
    }
    for (; m0 < n; m0++) row(m0, (double)qty_at(S, oq, m0 + 1));
    G = A / s;
    Gp = (A + kap * (2.0 * B - C)) / s;
}

// full solve at the accepted p: u(2..np-1) into Lo.u (u(1)=u(np)=0).  The
// forward sweep keeps 1/d, l1 and w per row; l2 = c/d is formed again on
// the way back.
struct FSweep {
    double d, e, dn, en, w, wp, l2p;
    TailTrack tt;
    __device__ __forceinline__ void row(const Band& B, const FastLayout& Lo, const Scr32& S, int m,
                                        double qn)
    {
#pragma clang fp contract(fast)
        const double rd = rcp64(d);
        const double rat = e * rd, rat2 = B.c * rd;
        S.st_nt(Lo.w + m, (float)w);
        tt.put(S, Lo, m, rd, rat);
        const double d1 = dn - rat * e;
        const double e1 = en - rat * B.c;
        const double d2 = B.a - rat2 * B.c;
        const double wn = qn - rat * w - l2p * wp;
        wp = w; w = wn; l2p = rat2;
        d = d1; e = e1; dn = d2; en = B.b;
    }
};

struct BSweep {
    double up1, up2, c;
    int n;
    __device__ __forceinline__ void row(const FastLayout& Lo, const Scr32& S, int m, float wv,
                                        float rv, float a1)
    {
#pragma clang fp contract(fast)
        const double r = (double)rv;
        double um = (double)wv * r;
        if (m < n - 1) um = um - up1 * (double)a1 - up2 * (c * r);
        S.st_nt(Lo.u + m + 1, (float)um);
        up2 = up1;
        up1 = um;
    }
};

__device__ __forceinline__ void back_u(const Band& B, const FastLayout& Lo, const Scr32& S,
                                       const FactorTail& tail);

__device__ __forceinline__ void solve_u(const Band& B, const FastLayout& Lo, const Scr32& S)
{
    const int n = Lo.np - 2;
    FSweep fw;
    fw.tt.init();
    fw.d = B.a; fw.e = B.b; fw.dn = B.a; fw.en = B.b;
    fw.w = (double)S.lq_ld(0); fw.wp = 0.0; fw.l2p = 0.0;
    int m0 = 0;
    for (; m0 + kCh < kLdsQ && m0 + kCh <= n; m0 += kCh) {
        float qn[kCh];
#pragma unroll
        for (int k = 0; k < kCh; k++) qn[k] = S.lq_ld(m0 + k + 1);
#pragma unroll
        for (int k = 0; k < kCh; k++) fw.row(B, Lo, S, m0 + k, (double)qn[k]);
    }
    for (; m0 + kCh <= n; m0 += kCh) {
        float qn[kCh];
#pragma unroll
        for (int k = 0; k < kCh; k++) qn[k] = S(Lo.q + m0 + k + 1);
#pragma unroll
        for (int k = 0; k < kCh; k++) fw.row(B, Lo, S, m0 + k, (double)qn[k]);
    }
    for (; m0 < n; m0++) fw.row(B, Lo, S, m0, (double)qty_at(S, Lo.q, m0 + 1));
    back_u(B, Lo, S, fw.tt.tail());
}

// smallest x in [0, hi] with pred(x) true in every active lane (pred
// monotone in x, true at hi), by bisection on wave votes: wave-uniform
template <class P>
__device__ __forceinline__ int wave_first(int hi, P pred)
{
    int lo = 0;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (__all(pred(mid))) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// back substitution from the stored 1/d, w, l1 rows: u(2..np-1) into Lo.u.
// Rows from tail.m on hold the tail's pair, taken from registers.  The
// lanes of a wave have different m: rows from the largest m on take the
// tail in every lane, rows below the smallest m are read in every lane, and
// rows in between are read and then replaced by the tail (the same value)
// where the lane's tail starts at or before them.
__device__ __forceinline__ void back_u(const Band& B, const FastLayout& Lo, const Scr32& S,
                                       const FactorTail& tail)
{
    const int n = Lo.np - 2;
    BSweep bw;
    bw.up1 = 0.0; bw.up2 = 0.0; bw.c = B.c; bw.n = n;
    const int tm = tail.m;
    const int mhi = wave_first(n, [&](int x) { return tm <= x; });
    const int mlo = n - wave_first(n, [&](int x) { return tm >= n - x; });
    int m1 = n - 1;
    for (; m1 - kCh + 1 >= mhi; m1 -= kCh) {
        float wv[kCh];
#pragma unroll
        for (int k = 0; k < kCh; k++) wv[k] = S.ld_nt(Lo.w + m1 - k);
#pragma unroll
        for (int k = 0; k < kCh; k++) bw.row(Lo, S, m1 - k, wv[k], tail.rd, tail.l1);
    }
    for (; m1 >= mhi; m1--) bw.row(Lo, S, m1, S.ld_nt(Lo.w + m1), tail.rd, tail.l1);
    auto mixed = [&](int m, float r, float l) {
        const bool st = m < tm;
        bw.row(Lo, S, m, S.ld_nt(Lo.w + m), st ? r : tail.rd, st ? l : tail.l1);
    };
    for (; m1 >= mlo; m1--) mixed(m1, S.ld_nt(Lo.rd + m1), S.ld_nt(Lo.l1 + m1));
    for (; m1 - kCh + 1 >= 0; m1 -= kCh) {
        float wv[kCh], rv[kCh], a1[kCh];
#pragma unroll
        for (int k = 0; k < kCh; k++) {
            wv[k] = S.ld_nt(Lo.w + m1 - k);
            rv[k] = S.ld_nt(Lo.rd + m1 - k);
            a1[k] = S.ld_nt(Lo.l1 + m1 - k);
        }
#pragma unroll
        for (int k = 0; k < kCh; k++) bw.row(Lo, S, m1 - k, wv[k], rv[k], a1[k]);
    }
    for (; m1 >= 0; m1--) bw.row(Lo, S, m1, S.ld_nt(Lo.w + m1), S.ld_nt(Lo.rd + m1), S.ld_nt(Lo.l1 + m1));
    S.st_nt(Lo.u + 0, 0.0f);
    S.st_nt(Lo.u + Lo.np - 1, 0.0f);
}

// The sub-steps' spline pieces from a window of the u and y rows: four
// intervals w..w+3 (u(w-1)..u(w+5), y(w)..y(w+4)) arrive in one round of
// loads, and a piece inside the window is formed from registers.  The
// coefficients of smooth's final loop (pppack.f90:5305-5314), formed from
// u, h and y.
struct PieceWindow {
    int w;           // first interval of the window (0: none loaded)
    float u[7], y[5];
};

__device__ __forceinline__ void window_load(const FastLayout& Lo, const Scr32& S, int i, PieceWindow& W)
{
    const int w = i <= 1 ? 1 : (i - 1 > Lo.np - 4 ? Lo.np - 4 : i - 1);
    W.w = w;
#pragma unroll
    for (int k = 0; k < 7; k++) W.u[k] = S.at(Lo.u + w - 2 + k);   // u(w-1+k); u(0), u(np+1) unused
#pragma unroll
    for (int k = 0; k < 5; k++) W.y[k] = S.at(Lo.y + w - 1 + k);   // y(w+k)
}

__device__ __forceinline__ void piece_window(const FastLayout& Lo, const PieceWindow& W, int i, double p,
                                             double k2, double h, double rh, double z1, PieceCache& pc)
{
#pragma clang fp contract(fast)
    const int np = Lo.np;
    const int k = i - W.w;   // 0..3
    auto sel = [&](const float* a, int o) {   // a[k + o] for k in 0..3
        return k == 0 ? a[o] : (k == 1 ? a[o + 1] : (k == 2 ? a[o + 2] : a[o + 3]));
    };
    const double um1 = (i >= 2) ? (double)sel(W.u, 0) : 0.0;
    const double u0 = (float)sel(W.u, 1), u1 = (float)sel(W.u, 2);
    const double u2 = (i + 2 <= np) ? (double)sel(W.u, 3) : 0.0;
    const double y0 = (float)sel(W.y, 0), y1 = (float)sel(W.y, 1);
    const double r_i = (i >= 2) ? (u0 - um1) * rh : 0.0;
    const double r_i1 = (u1 - u0) * rh;
    const double qu_i = r_i1 - r_i;
    const double qu_i1 = (i + 1 == np) ? -r_i1 : (u2 - u1) * rh - r_i1;
    const double a1 = y0 - k2 * qu_i;
    const double a1n = y1 - k2 * qu_i1;
    const double a3 = u0 * 6.0 * p;
    const double a3n = u1 * 6.0 * p;
    const double a4 = (a3n - a3) * rh;
    const double a2 = (a1n - a1) * rh - (a3 + a4 / 3.0 * h) / 2.0 * h;
    pc.i = i;
    pc.c[0] = a1; pc.c[1] = a2; pc.c[2] = a3; pc.c[3] = a4;
    pc.brk = z1 + (double)(i - 1) * h;
}

__device__ __forceinline__ double ppvalu_window(const FastLayout& Lo, const Scr32& S, PieceWindow& W,
                                               PieceCache& pc, double x, int jderiv, double p,
                                               double k2, double h, double rh, double z1)
{
    const int np = Lo.np;
    const double gss = (x - z1) * rh;
    int i;
    if (!(gss >= 0.0)) i = 1;
    else if (gss >= (double)(np - 1)) i = np - 1;
    else i = (int)gss + 1;
    if (i != pc.i) {
        if (W.w == 0 || i < W.w || i > W.w + 3) window_load(Lo, S, i, W);
        piece_window(Lo, W, i, p, k2, h, rh, z1, pc);
    }
    return horner(pc, x - pc.brk, jderiv);
}


// AKs at w level k of the particle's column, time-interpolated, and the
// level's depth (getInterpAKs + getWlevel + polintd, ver_turb_module.f90:100-124)
template <typename F>
__device__ __forceinline__ void column_level(const DGrid& g, const StepArgs& A, const RhoCell& c,
                                             const double cff[4], const Surface& sf, double hh,
                                             int k, double& z, double& kh)
{
#pragma clang fp contract(fast)
    const int nw = g.N + 1;
    if constexpr (is_cur<F>()) {
        // zeta and AKs already at ix(2): one level depth, one value per corner
        const float* ak = g.caks;
        const int nwc = g.nwc;
        const double a0 = ak[c.o[0] * nwc + (k - 1)], a1 = ak[c.o[1] * nwc + (k - 1)];
        const double a2 = ak[c.o[2] * nwc + (k - 1)], a3 = ak[c.o[3] * nwc + (k - 1)];
        const double Sk = level_S(g, g.s_w, g.cs_w, k, hh);
        z = level_z(g, g.s_w, g.cs_w, k, sf.zeta, hh, Sk);
        kh = cff[0] * a0 + cff[1] * a1 + cff[2] * a2 + cff[3] * a3;
        return;
    } else {
    const F* aks = (const F*)g.aks;
    Slots4<F> v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) ld4(aks + (c.o[q] * nw + (k - 1)) * 4, v[q]);
    const double Sk = level_S(g, g.s_w, g.cs_w, k, hh);
    const double zb = level_z(g, g.s_w, g.cs_w, k, sf.zb, hh, Sk);
    const double zc = level_z(g, g.s_w, g.cs_w, k, sf.zc, hh, Sk);
    const double zf = level_z(g, g.s_w, g.cs_w, k, sf.zf, hh, Sk);
    z = tinterp<F>(A, zb, zc, zf);
    double kk[3];
#pragma unroll
    for (int t = 0; t < 3; t++) {
        const int s = A.slots.s[t];
        kk[t] = cff[0] * pick(v[0], s) + cff[1] * pick(v[1], s) + cff[2] * pick(v[2], s) +
                cff[3] * pick(v[3], s);
    }
    kh = tinterp<F>(A, kk[0], kk[1], kk[2]);
    }
}

// The AKs column at the particle resampled on np uniform points: y rows and
// qty rows (qty row n = np-2 holds 0) into the work rows.  NT_Y: the y rows
// are stored non-temporal (the band sweeps re-read qty eight times and keep
// it cached; the sine-basis kernel reads every row once, soon).
struct VtCol {
    double z1, h, rh;
};

// KLV: levels per batch of loads (8 where the kernel has registers to spare
// in this phase, 4 under the band form's tighter budget)
template <typename F, bool NT_Y = true, int KLV = 4>
__device__ __forceinline__ VtCol vt_column_fast(const DGrid& g, const StepArgs& A, int np,
                                                const RhoCell& c, const Surface& sf, const Scr32& S)
{
#pragma clang fp contract(fast)
    const int nw = g.N + 1;
    const FastLayout Lo(np);
    const double tdepth = sf.tdepth, zeta = sf.zeta;
    const double hh = -1.0 * tdepth;
    double cff[4];
    rho_weights(g, c, cff);
    // the bottom and top levels fix the resampling grid; the column is then
    // walked once, kLv levels per batch of loads, emitting y (linint: the
    // point fz uses the segment of the largest level j<=nw-1 with z(j)<=fz)
    double z1, k1, zn, kn;
    constexpr int kLv = KLV;
    constexpr int kQ = kLv / 4;
    // the four corners' AKs of levels kb..kb+kLv-1 (kb-1 is a multiple of
    // four: kQ aligned 16-byte loads per corner; past the padded column, zeros)
    struct Quad {
        float4 a[4][kQ];
    };
    auto load_quad = [&](int kb) {
        Quad Q;
        if constexpr (is_cur<F>()) {
            const int nwc = g.nwc;
#pragma unroll
            for (int j = 0; j < kQ; j++)
#pragma unroll
                for (int q = 0; q < 4; q++)
                    Q.a[q][j] = (kb - 1 + 4 * j < nwc)
                                    ? *reinterpret_cast<const float4*>(g.caks + c.o[q] * nwc + (kb - 1 + 4 * j))
                                    : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        return Q;
    };
    // levels kb..kb+kLv-1 of the column (those in 2..nw are used); Q: their
    // AKs corners, loaded one batch ahead
    auto batch = [&](int kb, const Quad& Q, double zk[kLv], double kk[kLv]) {
        if constexpr (is_cur<F>()) {
            float v[4][kLv];
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int j = 0; j < kQ; j++) {
                    v[q][4 * j] = Q.a[q][j].x; v[q][4 * j + 1] = Q.a[q][j].y;
                    v[q][4 * j + 2] = Q.a[q][j].z; v[q][4 * j + 3] = Q.a[q][j].w;
                }
#pragma unroll
            for (int t = 0; t < kLv; t++) {
                const int k = kb + t;
                if (k >= 2 && k <= nw) {
                    const double Sk = level_S(g, g.s_w, g.cs_w, k, hh);
                    zk[t] = level_z(g, g.s_w, g.cs_w, k, zeta, hh, Sk);
                    kk[t] = cff[0] * (double)v[0][t] + cff[1] * (double)v[1][t] + cff[2] * (double)v[2][t] +
                            cff[3] * (double)v[3][t];
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < kLv; t++) {
                const int k = kb + t;
                if (k >= 2 && k <= nw) column_level<F>(g, A, c, cff, sf, hh, k, zk[t], kk[t]);
            }
        }
    };
    column_level<F>(g, A, c, cff, sf, hh, 1, z1, k1);
    column_level<F>(g, A, c, cff, sf, hh, nw, zn, kn);
    const double h = (zn - z1) / (double)(np - 1);
    const double rh = 1.0 / h;
    {
        double zlo = z1, klo = k1;
        double ym1 = 0.0, y0 = k1;
        if constexpr (NT_Y) S.st_nt(Lo.y + 0, (float)k1);
        else S(Lo.y + 0) = (float)k1;
        int i = 2;                                   // next point to emit
        Quad Qn = load_quad(1);
        for (int kb = 1; kb <= nw; kb += kLv) {
            const Quad Q = Qn;
            if (kb + kLv <= nw) Qn = load_quad(kb + kLv);
            double zk[kLv], kk[kLv];
            batch(kb, Q, zk, kk);
#pragma unroll
            for (int t = 0; t < kLv; t++) {
                const int k = kb + t;
                if (k >= 2 && k <= nw) {
                    const double zhi = zk[t], khi = kk[t];
                    const double m = (klo - khi) / (zlo - zhi);
                    const double b = klo - m * zlo;
                    for (; i <= np; i++) {
                        double yi;
                        if (i == np) {
                            yi = kn;
                        } else {
                            const double fz = z1 + (double)(i - 1) * h;
                            if (k < nw && !(fz < zhi)) break;
                            yi = m * fz + b;
                        }
                        if constexpr (NT_Y) S.st_nt_at(Lo.y + i - 1, (float)yi);
                        else S.at(Lo.y + i - 1) = (float)yi;
                        // qty row i-3 needs y(i-2): defined for i >= 3 (rows 0..np-3)
                        if (i >= 3) S.at(Lo.q + i - 3) = (float)(((yi - y0) - (y0 - ym1)) * rh);
                        ym1 = y0;
                        y0 = yi;
                    }
                    zlo = zhi;
                    klo = khi;
                }
            }
        }
    }
    S(Lo.q + (np - 2)) = 0.0f;   // read as the next row's qty by the last row of a sweep
    return VtCol{z1, h, rh};
}

// Below this t = p h^3 / (6 (1-p) serr^2) the knot values are rebuilt from
// their second differences (knots_from_curvature).  Smoothing weights of
// order 1 give t of 1e-2 and above on metre-spaced columns, where the direct
// form is the more accurate; t falls to 1e-4 and below as the weight grows.
#ifndef RP_KNOT_T0
#define RP_KNOT_T0 1.0e-3
#endif

// The spline's knot values f = y - 6(1-p) serr^2 Q u (smooth's final loop,
// pppack.f90:5305-5314) tend to the least-squares line of y as p -> 0, and
// the difference of the two nearly equal vectors y and 6(1-p) serr^2 Q u is
// lost in the FP32 rows of u.  From the system smooth solves,
// (6(1-p) Q'D^2 Q + p h tridiag(1,4,1)) u = Q'y, the second differences of f
// are Q'f = p h tridiag(1,4,1) u, small when p is, and f's least-squares line
// is y's (the columns of Q are orthogonal to linear functions).  So f is
// built in FP64 from its second differences, g(j+1) = 2 g(j) - g(j-1) +
// p h^2 T(j) with T(j) = u(j-1) + 4 u(j) + u(j+1) from g(1) = g(2) = 0, that
// is g(i) = p h^2 sum_j (i - j) T(j) over j < i, and given y's line.  The
// first pass sums y's line and, in closed form, g's; the second runs the
// recurrence and writes f over the y rows.  The pieces are then formed with
// k2 = 0 (a1 = the row).  u(1) = u(np) = 0.
__device__ __forceinline__ void knots_from_curvature(const FastLayout& Lo, const Scr32& S, double h, double p)
{
    constexpr int kB = 8;                            // rows loaded per batch
    const int np = Lo.np, last = np - 1;
    const double c = 0.5 * (double)(np + 1);
    // y's line about the centre knot, and with m = np - j the sums
    // sum_i g(i) = w sum_j T(j) m(m+1)/2 and
    // sum_i (i - c) g(i) = w sum_j T(j) [(j - c) m(m+1)/2 + m(m+1)(2m+1)/6]
    const double y1 = S.at(Lo.y + 0), yn = S.at(Lo.y + np - 1);
    double sy = y1 + yn, sjy = (1.0 - c) * (y1 - yn);
    double a0 = 0.0, a1 = 0.0;
    double um = 0.0, u0 = S.at(Lo.u + 1);            // u(1) = 0, u(2)
    for (int j0 = 2; j0 <= last; j0 += kB) {
        float yb[kB], ub[kB];
#pragma unroll
        for (int t = 0; t < kB; t++) {               // rows past the last are clamped, not used
            const int j = min(j0 + t, last);
            yb[t] = S.at(Lo.y + j - 1);
            ub[t] = S.at(Lo.u + j);                  // u(j+1)
        }
#pragma unroll
        for (int t = 0; t < kB; t++) {
            const int j = j0 + t;
            if (j <= last) {
                const double y = yb[t], up = ub[t];
                const double dj = (double)j - c;
                sy += y;
                sjy += dj * y;
                const double T = um + 4.0 * u0 + up;
                const double m = (double)(np - j);
                const double m2 = 0.5 * m * (m + 1.0);
                a0 += T * m2;
                a1 += T * (dj * m2 + m2 * (2.0 * m + 1.0) * (1.0 / 3.0));
                um = u0;
                u0 = up;
            }
        }
    }
    const double w = p * h * h;
    const double sjj = (double)np * ((double)np * (double)np - 1.0) / 12.0;
    const double da = (sy - w * a0) / (double)np, db = (sjy - w * a1) / sjj;
    S.at(Lo.y + 0) = (float)(da + db * (1.0 - c));
    S.at(Lo.y + 1) = (float)(da + db * (2.0 - c));
    double gm = 0.0, g0 = 0.0;
    um = 0.0;
    u0 = S.at(Lo.u + 1);
    for (int j0 = 2; j0 <= last; j0 += kB) {
        float ub[kB];
#pragma unroll
        for (int t = 0; t < kB; t++) ub[t] = S.at(Lo.u + min(j0 + t, last));
#pragma unroll
        for (int t = 0; t < kB; t++) {
            const int j = j0 + t;
            if (j <= last) {
                const double up = ub[t];
                const double g1 = 2.0 * g0 - gm + w * (um + 4.0 * u0 + up);
                S.at(Lo.y + j) = (float)(g1 + da + db * ((double)j + 1.0 - c));
                gm = g0;
                g0 = g1;
                um = u0;
                u0 = up;
            }
        }
    }
}

// the random-displacement sub-steps (ver_turb_module.f90:240-293) on the
// spline of p whose u and y rows are in the work rows
__device__ __forceinline__ double vt_substeps_fast(const StepArgs& A, const FastLayout& Lo,
                                                   const Scr32& S, const VtCol& col,
                                                   const Surface& sf, double Zpar, double p,
                                                   DevSrc& dv)
{
#pragma clang fp contract(fast)
    const double tdepth = sf.tdepth, zeta = sf.zeta;
    const double h = col.h, rh = col.rh, z1 = col.z1;
    double k2 = 6.0 * (1.0 - p) * (A.serr * A.serr);
    if (p < 1.0 && p * h * h * h < RP_KNOT_T0 * k2) {
        knots_from_curvature(Lo, S, h, p);
        k2 = 0.0;
    }
    PieceCache pc;
    pc.i = -1;
    PieceWindow W;
    W.w = 0;
    const double deltat = A.deltat;
    double ParZc = Zpar;
    for (int i = 1; i <= A.loop; i++) {
        double Kprimec;
        if (ParZc <= tdepth || ParZc >= zeta) Kprimec = 0.0;
        else Kprimec = ppvalu_window(Lo, S, W, pc, ParZc, 1, p, k2, h, rh, z1);
        const double KprimeZc = Kprimec * deltat;
        const double Z3rdc = ParZc + 0.5 * KprimeZc;
        double KH3rdc;
        if (Z3rdc < tdepth || Z3rdc > zeta) {
            KH3rdc = A.aksback;
        } else {
            KH3rdc = ppvalu_window(Lo, S, W, pc, Z3rdc, 0, p, k2, h, rh, z1);
            if (KH3rdc < A.aksback) KH3rdc = A.aksback;
        }
        const double DEV = dv.get(1 + i);
        const double dTz = KprimeZc + DEV * sqrt(2.0 * KH3rdc * deltat);
        ParZc = ParZc + dTz;
        if (ParZc <= tdepth) ParZc = tdepth + fabs(ParZc - tdepth);
        if (ParZc >= zeta) ParZc = zeta - fabs(ParZc - zeta);
    }
    return ParZc - Zpar;
}

template <typename F>
__device__ __forceinline__ double vturb_fast(const DGrid& g, const StepArgs& A, int np,
                                             const RhoCell& c, const Surface& sf, double Zpar,
                                             DevSrc& dv, const Scr32& S, unsigned* capped)
{
#pragma clang fp contract(fast)
    const FastLayout Lo(np);
    const VtCol col = vt_column_fast<F>(g, A, np, c, sf, S);
    const double h = col.h;
    // secant search for p (pppack.f90:5236-5298), sfq from G and G'
    const double serr = A.serr, s = A.smth;
    const int n = np - 2;
    // the column walk wrote every qty row to global memory; the first kLdsQ
    // go to LDS too, where the sweeps read them
    for (int j0 = 0; j0 < kLdsQ; j0 += 5) {
        float t[5];
#pragma unroll
        for (int j = 0; j < 5; j++) t[j] = S(Lo.q + j0 + j);
#pragma unroll
        for (int j = 0; j < 5; j++) S.lq[j0 + j] = t[j];
    }
    double p = 0.0, q = 0.0, sfq;
    bool stored = false;
    FactorTail tail{0, 0.0f, 0.0f};
    if (s <= 0.0) {
        p = 1.0;
    } else {
        double G, Gp;
        if (n <= kP0Rows) stream_G_p0(h, serr, n, S, Lo.q, G, Gp);
        else stream_G(band_of(0.0, h, serr), n, S, Lo, G, Gp);
        sfq = 6.0 * G;
        if (s < sfq) {
            const double utru = 0.5 * (G - Gp);
            const double ooss = 1.0 / sqrt(s);
            double oosf = 1.0 / sqrt(sfq);
            q = -(oosf - ooss) * sfq / (6.0 * utru * oosf);
            double prevq = 0.0, prevsf = oosf;
            int it = 0;
            for (;;) {
                const double pq = q / (1.0 + q);
                // from the sweep the search usually ends on, keep the sweep's
                // factor rows, so that the accepted p needs only the back
                // substitution
                stored = it >= kStoreFrom;
                if (stored) stream_G<true>(band_of(pq, h, serr), n, S, Lo, G, Gp, &tail);
                else stream_G(band_of(pq, h, serr), n, S, Lo, G, Gp);
                sfq = 6.0 * (G + pq * Gp) / ((1.0 + q) * (1.0 + q));
                if (fabs(sfq - s) <= 0.01 * s) break;
                if (++it >= kSecantCap) {
                    atomicAdd(capped, 1u);
                    break;
                }
                oosf = 1.0 / sqrt(sfq);
                const double change = (q - prevq) / (oosf - prevsf) * (oosf - ooss);
                prevq = q;
                q = q - change;
                prevsf = oosf;
            }
            p = q / (1.0 + q);
        }
    }
    if (stored) back_u(band_of(p, h, serr), Lo, S, tail);
    else solve_u(band_of(p, h, serr), Lo, S);
    return vt_substeps_fast(A, Lo, S, col, sf, Zpar, p, dv);
}

// boundary_module.f90:43-143 with the 3-point inside() of
// interpolation_module.f90:269-420 specialised to the unit cell.
__device__ __forceinline__ bool inside3(const double xb[3], const double yb[3], double xo, double yo)
{
    double X[4] = {xb[0], xb[1], xb[2], xb[0]};
    double Y[4] = {yb[0], yb[1], yb[2], yb[0]};
    int crossings = 0;
    bool sel[3];
    for (int k = 0; k < 3; k++)
        sel[k] = ((X[k + 1] - xo) * (xo - X[k]) >= 0.0) && (X[k] != X[k + 1]);
    for (int k = 0; k < 3; k++) {
        if (!sel[k]) continue;
        const double dx1 = xo - X[k];
        const double dx2 = X[k + 1] - xo;
        const double dxy = dx2 * (yo - Y[k]) - dx1 * (Y[k + 1] - yo);
        int inc = 0;
        if ((X[k] == xo) && (Y[k] == yo)) {
            return true;
        } else if (((dx1 == 0.0) && (yo >= Y[k])) || ((dx2 == 0.0) && (yo >= Y[k + 1]))) {
            inc = 1;
        } else if ((dx1 * dx2 > 0.0) && ((X[k + 1] - X[k]) * dxy >= 0.0)) {
            inc = 2;
        }
        if (X[k + 1] > X[k]) crossings += inc;
        else crossings -= inc;
    }
    return crossings != 0;
}

__device__ __forceinline__ void bounds(const DGrid& g, double Xp, double Yp, bool& ingrid, bool& obound)
{
    const double ecut = (double)0.1f, fcut = 1.0 - ecut;
    ingrid = false;
    obound = false;
    const int I = (int)floor(Xp), J = (int)floor(Yp);
    const double X = Xp - (double)I, Y = Yp - (double)J;
    if ((I > 1) && (J > 1) && (I < g.L - 1) && (J < g.M - 1)) {
        const int64_t b = (int64_t)(J - 1) * g.L + (I - 1);
        const double m11 = g.mrho[b], m21 = g.mrho[b + 1];
        const double m22 = g.mrho[b + g.L + 1], m12 = g.mrho[b + g.L];
        const int nm = (int)(m11 + m21 + m22 + m12);
        if (nm == 4) ingrid = true;
        if (nm == 3) {
            double tX[3] = {0, 0, 0}, tY[3] = {0, 0, 0};
            const double mm[4] = {m11, m12, m21, m22};   // (it,jt) = (1,1) (1,2) (2,1) (2,2)
            int n = 0;
            for (int q = 0; q < 4; q++) {
                if (mm[q] == 1.0 && n < 3) {
                    tX[n] = (double)(q >> 1);
                    tY[n] = (double)(q & 1);
                    n++;
                }
            }
            if (inside3(tX, tY, X, Y)) ingrid = true;
        }
        if (nm == 2) {
            if (m11 == 1.0 && (X < ecut) && (Y < ecut)) ingrid = true;
            if (m21 == 1.0 && (X > fcut) && (Y < ecut)) ingrid = true;
            if (m22 == 1.0 && (X > fcut) && (Y > fcut)) ingrid = true;
            if (m12 == 1.0 && (X < ecut) && (Y > fcut)) ingrid = true;
        }
        if ((I <= 1) || (J <= 1) || (I >= g.L - 1) || (J >= g.M - 1)) obound = true;
    } else {
        obound = true;
    }
}

// ---------------------------------------------------------------------------
// the step: update_particles (ROMSPath.f90:662-1003), one lane per particle,
// grid-stride over the particles.  PHASE 0 does everything in one pass (used
// without vertical turbulence); with it, PHASE 1 (advection + horizontal walk)
// hands newX, newY, AdvectZ and the clamped z to PHASE 2 (vertical walk,
// grid hand-off, reflection, status), so each kernel keeps a smaller
// register footprint.  Lane t of PHASE 2 uses scratch column t.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool skips(const StepArgs& A, int64_t q)
{
    if (A.ix[2] <= A.dob[q]) return true;                    // ROMSPath.f90:672
    if (A.status[q] == ST_UNSET) return true;                // ROMSPath.f90:676
    if (A.open_boundary && !A.isin[q]) return true;          // ROMSPath.f90:688-690
    return false;
}

__device__ __forceinline__ DevSrc dev_of(const StepArgs& A, int64_t q)
{
    DevSrc dv;
    dv.inj = A.dev ? A.dev + A.id[q] * (int64_t)A.ndev : nullptr;
    dv.seed = A.seed;
    dv.id = A.id[q];
    dv.step = A.step_index;
    dv.blk = -1;
    return dv;
}

// RKAdvect (advection_module.f90:20-125) + HTurb (hor_turb_module.f90:30-51,
// ROMSPath.f90:766-768).  Returns newX/newY (ROMSPath.f90:821-822).
// find_currents inputs of the four RK stages (start point and the three
// half-step points) and the start-point metrics, for the adjoint
template <class R>
struct StageRec {
    R x[4], y[4], z[4];
    R pm, pn;
};

template <typename F, int NSLOT, class R>
__device__ __forceinline__ void advect(const DGrid& g, const StepArgs& A, DevSrc& dv, R X0, R Y0,
                                       R& Zpar, R& newX, R& newY, R& AdvectZ,
                                       StageRec<R>* rec = nullptr)
{
    const RhoCellT<R> c0 = rho_cell(g, X0, Y0);
    const R pm = bilinear2d(g.pm, c0);
    const R pn = bilinear2d(g.pn, c0);
    R AdvectX = 0.0, AdvectY = 0.0, TurbHx = 0.0, TurbHy = 0.0;
    AdvectZ = 0.0;
    if (rec) {
        rec->pm = pm;
        rec->pn = pn;
    }
    if (A.scheme == 1) {
        // every stage starts from the step's start point with half a step (idt/2)
        const double idt = (double)A.idt;
        R Uad, Vad, Wad;
        int hint = 0;     // rho level index of the previous stage (the stages are close)
        if (rec) { rec->x[0] = X0; rec->y[0] = Y0; rec->z[0] = Zpar; }
        find_currents<F, NSLOT>(g, A, X0, Y0, Zpar, Uad, Vad, Wad, (SurfaceT<R>*)nullptr, &hint);
        const R k1u = Uad * pm, k1v = Vad * pn, k1w = Wad;
        R x1 = X0 + Uad * pm * idt / 2.0;
        R y1 = Y0 + Vad * pn * idt / 2.0;
        R z1 = Zpar + Wad * idt / 2.0;
        if (rec) { rec->x[1] = x1; rec->y[1] = y1; rec->z[1] = z1; }
        find_currents<F, NSLOT>(g, A, x1, y1, z1, Uad, Vad, Wad, (SurfaceT<R>*)nullptr, &hint);
        const R k2u = Uad * pm, k2v = Vad * pn, k2w = Wad;
        R x2 = X0 + Uad * pm * idt / 2.0;
        R y2 = Y0 + Vad * pn * idt / 2.0;
        R z2 = Zpar + Wad * idt / 2.0;
        if (rec) { rec->x[2] = x2; rec->y[2] = y2; rec->z[2] = z2; }
        find_currents<F, NSLOT>(g, A, x2, y2, z2, Uad, Vad, Wad, (SurfaceT<R>*)nullptr, &hint);
        const R k3u = Uad * pm, k3v = Vad * pn, k3w = Wad;
        R x3 = X0 + Uad * pm * idt / 2.0;
        R y3 = Y0 + Vad * pn * idt / 2.0;
        R z3 = Zpar + Wad * idt / 2.0;
        if (rec) { rec->x[3] = x3; rec->y[3] = y3; rec->z[3] = z3; }
        find_currents<F, NSLOT>(g, A, x3, y3, z3, Uad, Vad, Wad, (SurfaceT<R>*)nullptr, &hint);
        const R k4u = Uad * pm, k4v = Vad * pn, k4w = Wad;
        const R PU = (k1u + 2.0 * k2u + 2.0 * k3u + k4u) / 6.0;
        const R PV = (k1v + 2.0 * k2v + 2.0 * k3v + k4v) / 6.0;
        const R PW = (k1w + 2.0 * k2w + 2.0 * k3w + k4w) / 6.0;
        AdvectX = idt * PU;
        AdvectY = idt * PV;
        AdvectZ = idt * PW;
    }
    if (A.hturb_on) {
        const double KM = g.kh, r = 1.0;
        TurbHx = dv.get(0) * sqrt(2.0 / r * KM * (double)A.idt);
        TurbHy = dv.get(1) * sqrt(2.0 / r * KM * (double)A.idt);
    }
    TurbHx = TurbHx * pm;
    TurbHy = TurbHy * pn;
    newX = X0 + AdvectX + TurbHx;
    newY = Y0 + AdvectY + TurbHy;
}

// grid hand-off, new-position depth and zeta, vertical reflection and status
// (ROMSPath.f90:828-892).  A rejected or exiting move keeps (X0, Y0, Z0).
template <class R>
struct StepOut {
    R X, Y, Z, tdepth, zeta;
    int gid;
    int8_t st;
    bool ingrid, obound;
};

template <typename F, bool NESTED, class R>
__device__ __forceinline__ StepOut<R> finish_math(const DGrid* __restrict__ grids,
                                                  const StepArgs& A, int gid, R X0, R Y0, R Z0,
                                                  R newX, R newY, R AdvectZ, R TurbV)
{
    const double Zbehav = 0.0;
    R tempX = newX, tempY = newY;
    bool ingrid = false, obound = false;
    const int ngrid = NESTED ? A.ngrid : 1;
    for (int i = 1; i <= ngrid; i++) {
        const DGrid& gc = grids[NESTED ? gid - 1 : 0];
        tempX = gc.scl[i - 1][0] * newX + gc.off[i - 1][0];
        tempY = gc.scl[i - 1][1] * newY + gc.off[i - 1][1];
        bounds(grids[NESTED ? i - 1 : 0], val(tempX), val(tempY), ingrid, obound);
        if (ingrid) {
            gid = i;
            break;
        }
    }
    const DGrid& gn = grids[NESTED ? gid - 1 : 0];
    const RhoCellT<R> cn = rho_cell(gn, tempX, tempY);
    const SurfaceT<R> sn = surface_at<F>(gn, A, cn);
    const R tdepth = sn.tdepth, Pzeta = sn.zeta;
    bool hbot = false, htop = false;
    R newZ;
    if (A.nsb == 1) {
        newZ = Pzeta - A.vertdist;
    } else if (A.nsb == 2) {
        newZ = tdepth + A.vertdist;
    } else {
        newZ = Z0 + AdvectZ + TurbV + Zbehav;
        if (newZ < tdepth) {
            newZ = tdepth + fabs(newZ - tdepth);
            hbot = true;
        }
        if (newZ > Pzeta) {
            newZ = Pzeta - fabs(newZ - Pzeta);
            htop = true;
        }
    }
    StepOut<R> o;
    o.tdepth = tdepth;
    o.zeta = Pzeta;
    o.ingrid = ingrid;
    o.obound = obound;
    if (ingrid) {
        o.X = tempX;
        o.Y = tempY;
        o.Z = newZ;
        o.st = hbot ? ST_BOT : (htop ? ST_TOP : ST_MOVED);
    } else {
        o.X = X0;
        o.Y = Y0;
        o.Z = Z0;
        if (obound) {
            gid = A.ngrid;
            o.st = ST_EXIT;
        } else {
            o.st = ST_LAND;
        }
    }
    o.gid = gid;
    return o;
}

template <typename F, bool NESTED>
__device__ __forceinline__ void finish(const DGrid* __restrict__ grids, const StepArgs& A,
                                       int64_t q, int gid, double Z0, double newX, double newY,
                                       double AdvectZ, double TurbV)
{
    const StepOut<double> o = finish_math<F, NESTED>(grids, A, gid, A.x[q], A.y[q], Z0, newX, newY,
                                                     AdvectZ, TurbV);
    // want_diag 1: pWD, pZeta and pBath (dataOutput reads them); 2: pBath
    // only, which the per-step samplers read (the step's own tdepth)
    if (A.want_diag) {
        if (A.want_diag == 1) {
            A.wd[q] = (-1.0 * o.tdepth) + o.zeta;
            A.zeta[q] = o.zeta;
        }
        A.bath[q] = -1.0 * o.tdepth;
    }
    A.isin[q] = o.obound && !o.ingrid ? 0 : 1;
    if (o.ingrid) {
        A.x[q] = o.X;
        A.y[q] = o.Y;
        A.z[q] = o.Z;
    }
    A.status[q] = o.st;
    A.gid[q] = (int8_t)o.gid;
}

// One particle of the step.  Returns whether it took the step.  mrow: the
// particle's row of the hand-over rows (its position in the launch's list);
// tid, T: its vertical-walk scratch column and the row stride.
template <typename F, bool NESTED, int PHASE, bool FAITHFUL = false>
__device__ __forceinline__ bool step_one(const DGrid* __restrict__ grids, const StepArgs& A,
                                         int64_t q, int64_t mrow, int64_t tid, int64_t T)
{
    {
        if (skips(A, q)) return false;
        const double X0 = A.x[q], Y0 = A.y[q], Z0 = A.z[q];
        const int gid = NESTED ? (int)A.gid[q] : 1;
        const DGrid& g = grids[NESTED ? gid - 1 : 0];
        DevSrc dv = dev_of(A, q);
        double Zpar = Z0, newX, newY, AdvectZ, TurbV = 0.0;
        if (PHASE != 2) {
            A.age[q] = A.age[q] + (double)(float)A.idt;
            if (A.ets == 1) advect<F, 2>(g, A, dv, X0, Y0, Zpar, newX, newY, AdvectZ);
            else advect<F, 3>(g, A, dv, X0, Y0, Zpar, newX, newY, AdvectZ);
            if (PHASE == 1) {
                const int64_t r = mrow;
                A.mid[r] = newX;
                A.mid[A.nmid + r] = newY;
                A.mid[2 * A.nmid + r] = AdvectZ;
                A.mid[3 * A.nmid + r] = Zpar;
                return true;
            }
        } else {
            const int64_t r = mrow;
            newX = A.mid[r];
            newY = A.mid[A.nmid + r];
            AdvectZ = A.mid[2 * A.nmid + r];
            Zpar = A.mid[3 * A.nmid + r];
        }
        if (A.vturb_on) {
            const RhoCell c0 = rho_cell(g, X0, Y0);
            const Surface s0 = surface_at<F>(g, A, c0);
            // np: int(s_w(ng) * sub) of the particle's grid.  It is the
            // launch's (uniform) value unless the grids' N differ, in which
            // case a second copy of the walk takes it per lane.
            if constexpr (is_fast<F>() && !FAITHFUL) {
                const Scr32 S((float*)A.scratch, A.scratch_bytes, (uint32_t)tid, (uint32_t)T);
                if (NESTED && A.np_mixed) TurbV = vturb_fast<F>(g, A, g.np, c0, s0, Zpar, dv, S, A.capped);
                else TurbV = vturb_fast<F>(g, A, A.np, c0, s0, Zpar, dv, S, A.capped);
            } else {
                const Scr S{A.scratch + tid, T};
                if (NESTED && A.np_mixed) TurbV = vturb<F>(g, A, g.np, c0, s0, Zpar, dv, S, A.capped);
                else TurbV = vturb<F>(g, A, A.np, c0, s0, Zpar, dv, S, A.capped);
            }
        }
        finish<F, NESTED>(grids, A, q, gid, Z0, newX, newY, AdvectZ, TurbV);
    }
    return true;
}

}  // namespace rp
