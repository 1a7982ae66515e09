// rp_device.h -- device-side building blocks of the particle step (gfx950).
//
// Each function restates one reference routine; the citation is on the
// function.  FP64 arithmetic follows the reference's operation order and the
// translation unit is built with -ffp-contract=off, so with FP64 fields the
// results equal the reference's to the last bit wherever the reference's
// arithmetic is IEEE (no FMA, correctly rounded +-*/ and sqrt).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rp_dual.h"

namespace rp {

constexpr int kMaxGrid = 5;

// Status codes held on the device (int8); the host maps them to the
// reference's FP64 par(:,pStatus) values {9, 0, 1, 1.1, 1.2, -8, -9}; 1.1
// and 1.2 are single-precision literals there, so they read back as
// (double)1.1f and (double)1.2f.
enum : int8_t {
    ST_NEW = 0,       // 0.0  released, not yet moved
    ST_UNSET = 9,     // 9.0  below the bed at release: never moves
    ST_MOVED = 10,    // 1.0
    ST_TOP = 11,      // 1.1  reflected at the surface
    ST_BOT = 12,      // 1.2  reflected at the bed
    ST_EXIT = -8,     // -8.0 left through an open boundary
    ST_LAND = -9      // -9.0 move rejected: landed on a dry cell
};

// Device view of one grid.  Hydro fields hold four physical time slots
// interleaved innermost: element (i, j, k) slot s of a field with nx points
// along xi and nk levels sits at (((j-1)*nx + (i-1))*nk + (k-1))*4 + s, so the
// three slots the step reads (back, centre, forward) arrive in one 16-byte
// (FP32) or 32-byte (FP64) load, and levels k and k+1 are adjacent.
struct DGrid {
    int L, M, N, vtransform;
    double hc;
    const double *s_rho, *cs_r, *s_w, *cs_w;
    const double *h, *pm, *pn;           // (xi_rho, eta_rho) column-major
    const int8_t *mrho, *mu, *mv;        // masks
    const void *zeta, *u, *v, *w, *aks;  // FP32 or FP64, slot-interleaved
    // FP32 "current" fields of the step: the three slots already combined at
    // the step's time ix(2), one value per point (same point order, no slot);
    // an AKs column holds nwc >= N+1 values (padded to a multiple of four so
    // that four levels of a corner arrive in one aligned 16-byte load)
    const float *czeta, *cu, *cv, *cw, *caks;
    int nwc;
    int np;                              // VTurb resampling points, int(s_w(ng) * sub)
    double kh;                           // ConstantHTurb(ng)
    double scl[kMaxGrid][2], off[kMaxGrid][2];
    // the reference built with -DWETDRY: wet/dry masks (0/1 bytes) of each
    // slot s at wdr/wdu/wdv[s * wdp + point] for rho, u and v points, in
    // place of mrho/mu/mv in the u, v, w, salt, temp and bottom-stress
    // gathers; null: the static masks
    const uint8_t *wdr, *wdu, *wdv;
    int64_t wdp;
};

// Field kinds of the step code: double and float read the slot-interleaved
// fields (FP64 parity form, and the FP32 form's reference-order spline path
// used by the adjoint); Cur reads the FP32 current fields.  Every stage and
// the spline column interpolate in time at the same ix(2), and the temporal
// interpolation is linear in the field values, so combining the slots once
// per point and step is the same interpolation in another order.
struct Cur {};
// the same current fields under the -DWETDRY build's wet/dry masks: u and v
// carry each slot's mask (current_fields), w weighs the corners by them
struct CurWet {};
template <typename F>
constexpr bool is_cur()
{
    return __is_same(F, Cur) || __is_same(F, CurWet);
}
// the gathers take the per-slot wet/dry masks where the grid has them (the
// slot fields' kinds, and CurWet; plain Cur kernels carry no mask code)
template <typename F>
constexpr bool wet_code()
{
    return !__is_same(F, Cur);
}
// the FP32-field form's arithmetic economies (Lagrange time weights, one
// reciprocal per column, the streaming spline) apply to float and Cur
template <typename F>
constexpr bool is_fast()
{
    return is_cur<F>() || sizeof(F) == 4;
}

// ---------------------------------------------------------------------------
// loads
// ---------------------------------------------------------------------------
// The four slots of one point arrive in one load and stay in the storage
// type until one of them is picked (widened to FP64 only then).
template <typename F> struct Slots4;
template <> struct Slots4<float> {
    float v[4];
};
template <> struct Slots4<double> {
    double v[4];
};
__device__ __forceinline__ void ld4(const float* p, Slots4<float>& o)
{
    const float4 v = *reinterpret_cast<const float4*>(p);
    o.v[0] = v.x; o.v[1] = v.y; o.v[2] = v.z; o.v[3] = v.w;
}
__device__ __forceinline__ void ld4(const double* p, Slots4<double>& o)
{
    const double2 a = reinterpret_cast<const double2*>(p)[0];
    const double2 b = reinterpret_cast<const double2*>(p)[1];
    o.v[0] = a.x; o.v[1] = a.y; o.v[2] = b.x; o.v[3] = b.y;
}
template <typename F>
__device__ __forceinline__ double pick(const Slots4<F>& o, int s)
{
    return (double)(s == 0 ? o.v[0] : (s == 1 ? o.v[1] : (s == 2 ? o.v[2] : o.v[3])));
}

// ---------------------------------------------------------------------------
// interpolation_module.f90:72-109  (exact 3-point quadratic, Newton form
// based at the node nearest x)
// ---------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ R polintd(const double xa[3], R y0, R y1, R y2, double x)
{
    int ns = 1;
    double dif = fabs(x - xa[0]);
    double dift = fabs(x - xa[1]);
    if (dift < dif) { ns = 2; dif = dift; }
    dift = fabs(x - xa[2]);
    if (dift < dif) { ns = 3; dif = dift; }
    R c = (xa[1] - x) * ((y2 - y1) / (xa[1] - xa[2]));
    c = c - (xa[1] - x) * ((y1 - y0) / (xa[0] - xa[1]));
    c = c / (xa[0] - xa[2]);
    R a;
    double b;
    if (ns == 3) {
        a = (y2 - y1) / (xa[1] - xa[2]);
        b = xa[2] - x;
    } else {
        a = (y1 - y0) / (xa[0] - xa[1]);
        b = xa[0] - x;
    }
    const R yn = ns == 1 ? y0 : (ns == 2 ? y1 : y2);
    const double xn = ns == 1 ? xa[0] : (ns == 2 ? xa[1] : xa[2]);
    return yn + (xn - x) * a + b * c;
}

// interpolation_module.f90:1230-1273 (masked inverse-distance weights)
template <class R>
__device__ __forceinline__ void getcoeff(R X, R Y, const double m[4], R cff[4])
{
    const R d0 = sqrt(X * X + Y * Y);
    const R d1 = sqrt((1.0 - X) * (1.0 - X) + Y * Y);
    const R d2 = sqrt((1.0 - X) * (1.0 - X) + (1.0 - Y) * (1.0 - Y));
    const R d3 = sqrt(X * X + (1.0 - Y) * (1.0 - Y));
    const R w0 = m[0] / d0, w1 = m[1] / d1, w2 = m[2] / d2, w3 = m[3] / d3;
    const R wt = w0 + w1 + w2 + w3;
    if (wt == 0.0) {
        cff[0] = cff[1] = cff[2] = cff[3] = 0.0;
    } else {
        cff[0] = w0 / wt; cff[1] = w1 / wt; cff[2] = w2 / wt; cff[3] = w3 / wt;
    }
}

template <class R>
__device__ __forceinline__ void bilin(R x, R y, R c[4])
{
    c[0] = (1.0 - x) * (1.0 - y);
    c[1] = x * (1.0 - y);
    c[2] = x * y;
    c[3] = (1.0 - x) * y;
}

// grid_module.f90:513-598.  S(k) does not depend on zeta, so callers that
// need several slots at one level compute it once (same value bit for bit).
template <class R>
__device__ __forceinline__ R level_S(const DGrid& g, const double* s, const double* cs, int k, R h)
{
    if (g.vtransform == 2) return (g.hc * s[k - 1] + h * cs[k - 1]) / (g.hc + h);
    if (g.vtransform == 1) return g.hc * s[k - 1] + (h - g.hc) * cs[k - 1];
    return R(0.0);
}
template <class R>
__device__ __forceinline__ R level_z(const DGrid& g, const double* s, const double* cs, int k,
                                     R zeta, R h, R S)
{
    if (g.vtransform == 2) return zeta + (zeta + h) * S;
    if (g.vtransform == 1) return S + zeta * (1.0 + S / h);
    return zeta * (1.0 + s[k - 1]) + g.hc * s[k - 1] + (h - g.hc) * cs[k - 1];
}
template <class R>
__device__ __forceinline__ R level(const DGrid& g, const double* s, const double* cs, int k, R zeta,
                                   R h)
{
    return level_z(g, s, cs, k, zeta, h, level_S(g, s, cs, k, h));
}

// Level z(k) for one column: the FP32-field form multiplies by 1/(hc+h)
// computed once instead of dividing at every level.
template <bool FAST, class R = double>
struct Levels {
    const DGrid& g;
    const double *s, *cs;
    R zeta, h, inv;
    __device__ __forceinline__ Levels(const DGrid& g_, const double* s_, const double* cs_, R zeta_,
                                      R h_)
        : g(g_), s(s_), cs(cs_), zeta(zeta_), h(h_)
    {
        inv = (FAST && g.vtransform == 2) ? R(1.0 / (g.hc + h)) : R(0.0);
    }
    __device__ __forceinline__ R operator()(int k) const
    {
        if (FAST && g.vtransform == 2) return zeta + (zeta + h) * ((g.hc * s[k - 1] + h * cs[k - 1]) * inv);
        return level(g, s, cs, k, zeta, h);
    }
};

// First k in [1, nz] with level(k) > Z (nz+1 if none).  The reference scans
// linearly (interpolation_module.f90:533-577); the levels are non-decreasing
// in k, so the predicate is monotone and bisection returns the same index.
template <class Lev, class R>
__device__ __forceinline__ int first_above(const Lev& lev, int nz, R Z)
{
    int lo = 1, hi = nz + 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (lev(mid) > Z) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// The same index found by walking from a guess (the index of the previous RK
// stage, or of the rho bracket for the w bracket): down while the level below
// is above Z, then up while level k is not.  With a guess off by one or two
// this takes two or three level evaluations instead of the bisection's five.
template <class Lev, class R>
__device__ __forceinline__ int first_above_near(const Lev& lev, int nz, R Z, int guess)
{
    int k = min(max(guess, 1), nz + 1);
    while (k > 1 && lev(k - 1) > Z) k--;
    while (k <= nz && !(lev(k) > Z)) k++;
    return k;
}

// Vertical bracket of getInterp3d (interpolation_module.f90:528-589).
// k: the first level above Z (the scan's result), kept as the next guess.
template <class R = double>
struct VBracketT {
    int ibot, itop;
    R c0, c1;
    int k;
};
using VBracket = VBracketT<double>;

template <bool FAST, class R>
__device__ __forceinline__ VBracketT<R> bracket_rho(const DGrid& g, R Z, R zeta, R depth,
                                                    int guess = 0)
{
    const R h = -1.0 * depth;
    const int nz = g.N;
    const Levels<FAST, R> lev(g, g.s_rho, g.cs_r, zeta, h);
    const int i = guess > 0 ? first_above_near(lev, nz, Z, guess) : first_above(lev, nz, Z);
    VBracketT<R> b;
    b.k = i;
    if (i > nz) {
        b.ibot = b.itop = nz; b.c0 = 0.5; b.c1 = 0.5;
    } else if (i == 1) {
        b.ibot = b.itop = 1; b.c0 = 0.5; b.c1 = 0.5;
    } else {
        const R z1 = lev(i - 1);
        const R z2 = lev(i);
        const R Zr = (Z - z1) / (z2 - z1);
        b.ibot = i - 1; b.itop = i; b.c0 = 1.0 - Zr; b.c1 = Zr;
    }
    return b;
}

template <bool FAST, class R>
__device__ __forceinline__ VBracketT<R> bracket_w(const DGrid& g, R Z, R zeta, R depth,
                                                  int guess = 0)
{
    const R h = -1.0 * depth;
    const int nz = g.N + 1;
    const Levels<FAST, R> lev(g, g.s_w, g.cs_w, zeta, h);
    const int i = guess > 0 ? first_above_near(lev, nz, Z, guess) : first_above(lev, nz, Z);
    VBracketT<R> b;
    b.k = i;
    if (i > nz) {
        // z at or above the top w level: two copies of level nz (the
        // reference's weight is undefined there; the value is not).
        b.ibot = b.itop = nz; b.c0 = 0.5; b.c1 = 0.5;
    } else {
        const int ib = i > 1 ? i - 1 : 1;
        const R z1 = lev(ib);
        const R z2 = lev(i);
        const R Zr = (Z - z1) / (z2 - z1);
        b.ibot = ib; b.itop = i; b.c0 = 1.0 - Zr; b.c1 = Zr;
    }
    return b;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11), the counter-based generator that replaces the global
// serial MT19937 stream of random_module.f90 / pdf_module.f90:26-40.
// ---------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

__host__ __device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)M0 * c.x;
        const uint64_t p1 = (uint64_t)M1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        U4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// sin and cos of 2 pi (x + 1/2) 2^-32 for a 32-bit integer x.  The angle is
// reduced exactly on the integer: the top two bits give the quadrant, and
// the rest is folded to an angle t in (0, pi/4] (its complement where it is
// past pi/4), where the Taylor series of sin and cos through t^17 and t^18
// leave a truncation error below 1e-19.
__host__ __device__ __forceinline__ void sincos_2pi_u32(uint32_t x, double& s, double& c)
{
    const uint32_t q = x >> 30, m = x & 0x3FFFFFFFu;
    const bool fold = m >= 0x20000000u;
    const double f = fold ? (double)(0x40000000u - m) - 0.5 : (double)m + 0.5;
    const double t = f * (6.283185307179586 * 2.3283064365386963e-10);   // 2 pi 2^-32
    const double t2 = t * t;
    double ps = 1.0 / 355687428096000.0;                 // 1/17!
    ps = ps * t2 - 1.0 / 1307674368000.0;                // 1/15!
    ps = ps * t2 + 1.0 / 6227020800.0;                   // 1/13!
    ps = ps * t2 - 1.0 / 39916800.0;                     // 1/11!
    ps = ps * t2 + 1.0 / 362880.0;                       // 1/9!
    ps = ps * t2 - 1.0 / 5040.0;
    ps = ps * t2 + 1.0 / 120.0;
    ps = ps * t2 - 1.0 / 6.0;
    const double st = t + t * (t2 * ps);
    double pc = 1.0 / 6402373705728000.0;                // 1/18!
    pc = pc * t2 - 1.0 / 20922789888000.0;               // 1/16!
    pc = pc * t2 + 1.0 / 87178291200.0;                  // 1/14!
    pc = pc * t2 - 1.0 / 479001600.0;                    // 1/12!
    pc = pc * t2 + 1.0 / 3628800.0;                      // 1/10!
    pc = pc * t2 - 1.0 / 40320.0;
    pc = pc * t2 + 1.0 / 720.0;
    pc = pc * t2 - 1.0 / 24.0;
    pc = pc * t2 + 0.5;
    const double ct = 1.0 - t2 * pc;
    const double s1 = fold ? ct : st, c1 = fold ? st : ct;
    switch (q) {
    case 0: s = s1; c = c1; break;
    case 1: s = c1; c = -s1; break;
    case 2: s = -s1; c = -c1; break;
    default: s = -c1; c = s1; break;
    }
}

// Four standard normals from one Philox block: two Box-Muller pairs on
// uniforms (x + 1/2) 2^-32 in (0,1), the resolution of genrand_real3
// (random_module.f90:240-247).  The reference keeps only the cosine branch
// (pdf_module.f90:38); both branches are independent normals, so both are used.
__device__ __forceinline__ void normals4(uint64_t seed, uint64_t id, uint64_t step, uint32_t blk,
                                         double out[4])
{
    U4 c;
    c.x = blk;
    c.y = (uint32_t)id;
    c.z = (uint32_t)(id >> 32);
    c.w = (uint32_t)step;
    const U4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32) ^ (uint32_t)(step >> 32));
    const double s = 2.3283064365386963e-10;  // 2^-32
    const double u0 = ((double)r.x + 0.5) * s, u2 = ((double)r.z + 0.5) * s;
    const double ra = sqrt(-2.0 * log(u0)), rb = sqrt(-2.0 * log(u2));
    double sa, ca, sb, cb;
    sincos_2pi_u32(r.y, sa, ca);     // angle 2 pi u1, u1 = (r.y + 1/2) 2^-32
    sincos_2pi_u32(r.w, sb, cb);
    out[0] = ra * ca; out[1] = ra * sa; out[2] = rb * cb; out[3] = rb * sb;
}

}  // namespace rp
