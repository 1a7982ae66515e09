// rp_adjoint.hip -- adjoint of the particle step (no reference counterpart).
//
// The forward steps of a recorded run are the FP64-arithmetic step (the
// smoothing spline of the vertical walk on its reference-order path for
// either field precision), so the trajectory map is piecewise smooth.  The
// backward pass walks the recorded steps in reverse.  For each particle and
// step it evaluates the step once more in Dual<3> numbers seeded at the
// step's start (X, Y, Z): the derivative parts are the exact 3x3 Jacobian J
// of the step as implemented (RK stages with the in-place z clamp,
// interpolation, the vertical walk with its spline fit and secant search,
// hand-off and reflection), and lambda <- J^T lambda.  For the velocity
// perturbation the RK combination is reversed stage by stage:
//   A = idt (k1 + 2 k2 + 2 k3 + k4) / 6,  x_s = base + idt/2 k_{s-1},
//   nu4 = idt/6 eta, nu3 = 2 idt/6 eta + idt/2 J4^T nu4, ...,
//   dL/dc += sum_s nu_s . dk_s/dc(x_s),
// with eta = (d out / d(newX, newY, newZ))^T lambda from finish_math in
// Dual<3> and the stage Jacobians J_s = dk_s/dx_s from find_currents in
// Dual<3> at the recorded stage inputs.
#include "rp_step_dev.h"

namespace rp {

using D3 = Dual<3>;

// forward step of a recorded run: the whole step in one kernel, spline on
// the reference-order path; persistent grid with the vertical walk's queue.
// Held to 2 waves per SIMD (the compiler's choice, 1, waits on memory more
// than the spill costs: C5 recorded forward 1.57 s against 2.28 s).
template <typename F, bool NESTED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    adj_forward_kernel(const DGrid* __restrict__ grids, StepArgs A)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    unsigned long long nact = 0;
    for (;;) {
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(A.work, 64ull);
        b = __shfl(b, 0, 64);
        if (b >= (unsigned long long)A.n) break;
        const int64_t r = (int64_t)b + lane;
        if (r < A.n && step_one<F, NESTED, 0, true>(grids, A, A.q0 + r, r, tid, T)) nact++;
    }
    for (int off = 32; off > 0; off >>= 1) nact += __shfl_down(nact, off, 64);
    if (lane == 0 && nact) atomicAdd(A.active, nact);
}

// state of every particle at the start of a step, for the backward sweep
__global__ void __launch_bounds__(256) adj_save_kernel(AdjState dst, const double* x,
                                                       const double* y, const double* z,
                                                       const int8_t* status, const int8_t* gid,
                                                       const uint8_t* isin, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        dst.x[i] = x[i];
        dst.y[i] = y[i];
        dst.z[i] = z[i];
        dst.status[i] = status[i];
        dst.gid[i] = gid[i];
        dst.isin[i] = isin[i];
    }
}

__global__ void __launch_bounds__(256) adj_fill(double* a, int64_t n, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// k_s = (U pm, V pn, W) at stage input x (z clamped inside find_currents)
template <typename F, int NSLOT>
__device__ __forceinline__ void stage_jacobian(const DGrid& g, const StepArgs& A, double x, double y,
                                               double z, double pm, double pn, double J[3][3])
{
    D3 X = D3::seed(x, 0), Y = D3::seed(y, 1), Z = D3::seed(z, 2);
    D3 U, V, W;
    find_currents<F, NSLOT>(g, A, X, Y, Z, U, V, W, (SurfaceT<D3>*)nullptr);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        J[0][j] = U.d[j] * pm;
        J[1][j] = V.d[j] * pn;
        J[2][j] = W.d[j];
    }
}

__device__ __forceinline__ void mtv(const double J[3][3], const double v[3], double out[3])
{
#pragma unroll
    for (int j = 0; j < 3; j++) out[j] = J[0][j] * v[0] + J[1][j] * v[1] + J[2][j] * v[2];
}

template <typename F, bool NESTED, int NSLOT>
__device__ __forceinline__ void adj_particle(const DGrid* __restrict__ grids, const StepArgs& A,
                                             const AdjArgs& J, int64_t q, int64_t tid, int64_t T,
                                             double* gp_lds)
{
    const double X0 = A.x[q], Y0 = A.y[q], Z0 = A.z[q];
    const int gid = NESTED ? (int)A.gid[q] : 1;
    const DGrid& g = grids[NESTED ? gid - 1 : 0];
    DevSrc dv = dev_of(A, q);
    const D3 X = D3::seed(X0, 0), Y = D3::seed(Y0, 1), Z = D3::seed(Z0, 2);
    D3 Zpar = Z, newX, newY, AdvectZ, TurbV = 0.0;
    StageRec<D3> rec;
    advect<F, NSLOT>(g, A, dv, X, Y, Zpar, newX, newY, AdvectZ, &rec);
    if (A.vturb_on) {
        // the column and the spline fit depend on the start X, Y only: they
        // run in two-direction duals, the sub-steps in three
        const Dual<2> X2 = Dual<2>::seed(X0, 0), Y2 = Dual<2>::seed(Y0, 1);
        const RhoCellT<Dual<2>> c0 = rho_cell(g, X2, Y2);
        const SurfaceT<Dual<2>> s0 = surface_at<F>(g, A, c0);
        const ScrT<Dual<2>> S{(Dual<2>*)A.scratch + tid, T};
        TurbV = vturb<F, Dual<2>, D3>(g, A, (NESTED && A.np_mixed) ? g.np : A.np, c0, s0, Zpar, dv, S,
                                      A.capped);
    }
    const StepOut<D3> o = finish_math<F, NESTED>(grids, A, gid, X, Y, Z, newX, newY, AdvectZ, TurbV);
    const double lam[3] = {J.lx[q], J.ly[q], J.lz[q]};
    double Jx[3][3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        Jx[0][j] = o.X.d[j];
        Jx[1][j] = o.Y.d[j];
        Jx[2][j] = o.Z.d[j];
    }
    double ln[3];
    mtv(Jx, lam, ln);
    J.lx[q] = ln[0];
    J.ly[q] = ln[1];
    J.lz[q] = ln[2];
    if (!gp_lds || !o.ingrid || A.scheme != 1) return;
    // eta: sensitivity of the outputs to (newX, newY, z before reflection)
    const double zpre = (Z0 + AdvectZ.v) + TurbV.v;
    const StepOut<D3> of = finish_math<F, NESTED>(grids, A, gid, D3(X0), D3(Y0), D3::seed(zpre, 2),
                                                  D3::seed(newX.v, 0), D3::seed(newY.v, 1), D3(0.0),
                                                  D3(0.0));
    double Jf[3][3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        Jf[0][j] = of.X.d[j];
        Jf[1][j] = of.Y.d[j];
        Jf[2][j] = of.Z.d[j];
    }
    double eta[3];
    mtv(Jf, lam, eta);
    // reverse through the four stages
    const double idt = (double)A.idt, h = idt / 2.0;
    const double pm = rec.pm.v, pn = rec.pn.v;
    const double w[4] = {idt / 6.0, 2.0 * idt / 6.0, 2.0 * idt / 6.0, idt / 6.0};
    double nu[4][3];
#pragma unroll
    for (int i = 0; i < 3; i++) nu[3][i] = w[3] * eta[i];
    for (int s = 3; s >= 1; s--) {
        double Js[3][3], t[3];
        stage_jacobian<F, NSLOT>(g, A, rec.x[s].v, rec.y[s].v, rec.z[s].v, pm, pn, Js);
        mtv(Js, nu[s], t);
#pragma unroll
        for (int i = 0; i < 3; i++) nu[s - 1][i] = w[s - 1] * eta[i] + h * t[i];
    }
    // dk_s/dc at the stage inputs: pm B for u, pn B for v
    const int nc = A.pert_nx * A.pert_ny;
    for (int s = 0; s < 4; s++) {
        double tx, ty, bx[4], by[4];
        const int sx = bspline_span(rec.x[s].v, g.L, A.pert_nx, tx);
        const int sy = bspline_span(rec.y[s].v, g.M, A.pert_ny, ty);
        bspline4(tx, bx);
        bspline4(ty, by);
        const double au = nu[s][0] * pm, av = nu[s][1] * pn;
        for (int b = 0; b < 4; b++)
            for (int a = 0; a < 4; a++) {
                const double bw = bx[a] * by[b];
                const int k = (sy + b) * A.pert_nx + (sx + a);
                atomicAdd(&gp_lds[k], au * bw);
                atomicAdd(&gp_lds[nc + k], av * bw);
            }
    }
}

// One recorded step backwards: lambda <- J^T lambda for every particle that
// took the step (the others keep lambda), and the perturbation gradient.
// Held to 2 waves per SIMD like the forward (C5 adjoint 5.04 s against 6.23
// at the compiler's 1; 3 waves spill too much).
template <typename F, bool NESTED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    adj_step_kernel(const DGrid* __restrict__ grids, StepArgs A, AdjArgs J)
{
    extern __shared__ double gp_lds[];
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    const int nc2 = J.gpert ? 2 * A.pert_nx * A.pert_ny : 0;
    for (int k = threadIdx.x; k < nc2; k += blockDim.x) gp_lds[k] = 0.0;
    __syncthreads();
    for (;;) {
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(A.work, 64ull);
        b = __shfl(b, 0, 64);
        if (b >= (unsigned long long)A.n) break;
        const int64_t q = A.q0 + (int64_t)b + lane;
        if ((int64_t)b + lane < A.n && !skips(A, q)) {
            if (A.ets == 1) adj_particle<F, NESTED, 2>(grids, A, J, q, tid, T, nc2 ? gp_lds : nullptr);
            else adj_particle<F, NESTED, 3>(grids, A, J, q, tid, T, nc2 ? gp_lds : nullptr);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nc2; k += blockDim.x)
        if (gp_lds[k] != 0.0) atomicAdd(&J.gpert[k], gp_lds[k]);
}

#define RP_ADJ_INST(F, N)                                                                 \
    template __global__ void adj_forward_kernel<F, N>(const DGrid*, StepArgs);            \
    template __global__ void adj_step_kernel<F, N>(const DGrid*, StepArgs, AdjArgs);
RP_ADJ_INST(float, false) RP_ADJ_INST(float, true) RP_ADJ_INST(double, false) RP_ADJ_INST(double, true)
#undef RP_ADJ_INST

}  // namespace rp
