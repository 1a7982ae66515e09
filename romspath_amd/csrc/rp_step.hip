// rp_step.hip -- the particle-step kernels and the field-upload kernels.
#include <vector>

#include "rp_step_dev.h"

namespace rp {

__device__ __forceinline__ void count_active(const StepArgs& A, unsigned long long nact)
{
    // wave reduction, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) nact += __shfl_down(nact, off, 64);
    if ((threadIdx.x & 63) == 0 && nact) atomicAdd(A.active, nact);
}

// particles in the launch: the compacted active set, or all n.  When every
// particle is active the list is the identity and is not read.
__device__ __forceinline__ int64_t launch_count(const StepArgs& A)
{
    return A.act ? (int64_t)*A.act_n : A.n;
}
__device__ __forceinline__ int64_t launch_particle(const StepArgs& A, int64_t ne, int64_t r)
{
    return A.q0 + ((A.act && ne != A.n) ? (int64_t)A.act[r] : r);
}

// one lane per particle (the grid covers n, the upper bound of the count)
template <typename F, bool NESTED, int PHASE>
__device__ __forceinline__ void step_body(const DGrid* __restrict__ grids, const StepArgs& A)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ne = launch_count(A);
    unsigned long long nact = 0;
    if (r < ne && step_one<F, NESTED, PHASE>(grids, A, launch_particle(A, ne, r), r, r, 0)) nact = 1;
    if (PHASE != 1) count_active(A, nact);
}

// The launch's active set in particle order, in three launches:
// compact_count counts each block's active particles, sort_scan turns the
// counts into block offsets, compact_scatter writes the list and its length.
__device__ __forceinline__ bool active_in_launch(const StepArgs& A, int64_t r)
{
    return r < A.n && !skips(A, A.q0 + r);
}

__global__ void __launch_bounds__(256) compact_count(StepArgs A, uint32_t* bcnt)
{
    __shared__ uint32_t wsum[4];
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const unsigned long long m = __ballot(active_in_launch(A, r));
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) bcnt[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ void __launch_bounds__(256) compact_scatter(StepArgs A, const uint32_t* boff, uint32_t* act,
                                                       unsigned long long* act_n)
{
    __shared__ uint32_t wsum[4];
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool a = active_in_launch(A, r);
    const unsigned long long m = __ballot(a);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wsum[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = boff[blockIdx.x];
    for (int k = 0; k < w; k++) off += wsum[k];
    if (a) act[off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)r;
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        *act_n = (unsigned long long)boff[blockIdx.x] + wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Persistent form for the vertical walk: a grid sized to the resident
// waves; each wave claims the next 64 particles from a counter until none
// are left (every wave reaches the exit once the counter passes n).  Lane t
// keeps scratch column t.  Dynamic claiming evens out the per-particle cost,
// which varies with the number of secant iterations.
template <typename F, bool NESTED>
__device__ __forceinline__ void vwalk_body(const DGrid* __restrict__ grids, const StepArgs& A)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    unsigned long long nact = 0;
    const int64_t ne = launch_count(A);
    for (;;) {
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(A.work, 64ull);
        b = __shfl(b, 0, 64);
        if (b >= (unsigned long long)ne) break;
        const int64_t r = (int64_t)b + lane;
        if (r < ne && step_one<F, NESTED, 2>(grids, A, launch_particle(A, ne, r), r, tid, T)) nact++;
    }
    count_active(A, nact);
}

// ---------------------------------------------------------------------------
// wave priority (0-3) of the eigenbasis walk's memory-bound phases
#ifndef RP_PRIO_MEM
#define RP_PRIO_MEM 1
#endif

// The FP32-field vertical walk with smooth's secant search in the eigenbasis
// of chol1d's matrix pencil (rp_modal.h).  Same persistent wave-granular
// claiming as vwalk_body, but each 64-particle claim runs in phases so that
// the wave is converged where the matrix instructions run:
//   1. per lane: the column walk (y and qty work rows);
//   2. wave: d = Phi' qty (lanes without a column contribute zeros);
//   3. per lane: the secant search on the modes in registers;
//   4. wave: u = Phi vh into the u work rows, vh formed from d on the way;
//   5. per lane: the random-displacement sub-steps and the rest of the step.
// ---------------------------------------------------------------------------

// a double in two FP32 work rows (its bits, unchanged)
__device__ __forceinline__ void st_dbl(const Scr32& S, int row, double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    S(row) = __builtin_bit_cast(float, (uint32_t)u);
    S(row + 1) = __builtin_bit_cast(float, (uint32_t)(u >> 32));
}
__device__ __forceinline__ double ld_dbl(const Scr32& S, int row)
{
    const uint32_t lo = __builtin_bit_cast(uint32_t, (float)S(row));
    const uint32_t hi = __builtin_bit_cast(uint32_t, (float)S(row + 1));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// The accepted solution in the eigenbasis: mode k is d_k R scale with
// R = 1/(lambda_k + t), or, at p = 1, d_k scale.
struct ModalSol {
    double t, scale;
    bool p1;
};

// smooth's search for p (pppack.f90:5236-5298) on the modal sums: sfq and
// utru as in rp_modal.h.  Returns p and the accepted solution's form.
__device__ __forceinline__ double modal_search(const ModalTab* __restrict__ MT, float (&co)[kMh],
                                               float (&ce)[kMh], double h, const StepArgs& A,
                                               ModalSol& sol)
{
#pragma clang fp contract(fast)
    const double serr = A.serr, s = A.smth;
    const double r2 = (serr / h) * (serr / h);
    double p = 0.0, t = 0.0;
    ModalSums Sm;
    if (s <= 0.0) {
        p = 1.0;
    } else {
        // one loop body for every evaluation (the first at p = 0), so that
        // the mode loop is compiled once and never with t known
        double alpha = 6.0 * r2, q = 0.0, prevq = 0.0, prevsf = 0.0, oosf = 0.0, ooss = 0.0;
        int it = -1;
        for (;;) {
            modal_sums(MT, co, ce, t, Sm);
            const double g = Sm.g, gp = -Sm.gp2;
            if (it < 0) {
                const double sfq = 36.0 * r2 * g / (alpha * alpha);
                if (!(s < sfq)) break;                      // p = 0
                const double utru = -h * gp / (2.0 * alpha * alpha);
                ooss = 1.0 / sqrt(s);
                oosf = 1.0 / sqrt(sfq);
                q = -(oosf - ooss) * sfq / (6.0 * utru * oosf);
                prevq = 0.0;
                prevsf = oosf;
                it = 0;
            } else {
                const double sfq = 36.0 * r2 * (g + t * gp) / (alpha * alpha * ((1.0 + q) * (1.0 + q)));
                if (fabs(sfq - s) <= 0.01 * s) {
                    p = q / (1.0 + q);
                    break;
                }
                if (++it >= kSecantCap) {
                    atomicAdd(A.capped, 1u);
                    p = q / (1.0 + q);
                    break;
                }
                oosf = 1.0 / sqrt(sfq);
                const double change = (q - prevq) / (oosf - prevsf) * (oosf - ooss);
                prevq = q;
                q = q - change;
                prevsf = oosf;
            }
            const double pq = q / (1.0 + q);
            alpha = 6.0 * (1.0 - pq) * r2;
            t = pq * h / alpha;
        }
    }
    // the last evaluation was at the accepted p
    sol.t = t;
    sol.p1 = p >= 1.0;
    sol.scale = sol.p1 ? 1.0 / h : 1.0 / (6.0 * (1.0 - p) * r2);
    return p;
}

template <bool NESTED>
__device__ __forceinline__ void vwalk_modal_body(const DGrid* __restrict__ grids, const StepArgs& A)
{
    __shared__ float tab[2 * kMh * kMStride];
    const ModalTab* __restrict__ MT = A.modal;
    for (int i = threadIdx.x; i < 2 * kMh * kMh; i += blockDim.x) {
        const int pm = i / kMh, j = i % kMh;      // pm = par * kMh + m
        tab[pm * kMStride + j] = (&MT->basis[0][0][0])[i];
    }
    __syncthreads();
    const float* tab_e = tab + kMh * kMStride;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    const int n0 = __builtin_amdgcn_readfirstlane(MT->n);
    unsigned long long nact = 0;
    const int64_t ne = launch_count(A);
    for (;;) {
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(A.work, 64ull);
        // lane 0's claim, as a wave-uniform (scalar) value
        b = ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32)) << 32) |
            (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
        if (b >= (unsigned long long)ne) break;
        // The hundreds of work-row offsets below are invariant across the
        // claims; hoisted out of this loop they overflow the scalar
        // registers.  Opaque copies of n and the row stride keep them
        // computed where they are used.
        int n = n0;
        uint32_t Tu = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)T);
        asm volatile("" : "+s"(n), "+s"(Tu));
        const int half = n / 2;
        const FastLayout Lo(n + 2);
        const Scr32 S((float*)A.scratch, A.scratch_bytes, (uint32_t)tid, Tu);
        const int64_t r = (int64_t)b + lane;
        // 1. the column.  Only what the next phases use stays in registers
        // (the 128 mode values need the room): h; z1, 1/h and the surface
        // go to spare work rows, and the particle's state is read again for
        // the sub-steps.
        // Wave priority: the stretches that wait on memory (the column walk,
        // each parity's qty row reads and the sub-steps) issue ahead of the
        // other wave's transforms and search, whose arithmetic fills the
        // gaps (C3 vertical walk 5.29 -> 5.14 ms for the column walk and the
        // sub-steps, 4.73 ms with the row reads too).
        __builtin_amdgcn_s_setprio(RP_PRIO_MEM);
        bool go = false;
        double h = 0.0;
        if (r < ne) {
            const int64_t q = launch_particle(A, ne, r);
            go = !skips(A, q);
            if (go) {
                const double X0 = A.x[q], Y0 = A.y[q];
                const int gid = NESTED ? (int)A.gid[q] : 1;
                const DGrid& g = grids[NESTED ? gid - 1 : 0];
                const RhoCell c0 = rho_cell(g, X0, Y0);
                const Surface s0 = surface_at<Cur>(g, A, c0);
                const VtCol col = vt_column_fast<Cur, false, 8>(g, A, n + 2, c0, s0, S);
                st_dbl(S, Lo.w + 0, col.z1);
                st_dbl(S, Lo.w + 2, col.rh);
                st_dbl(S, Lo.w + 4, s0.tdepth);
                st_dbl(S, Lo.w + 6, s0.zeta);
                h = col.h;
            }
        }
        __builtin_amdgcn_s_setprio(0);
        // 2. modes of qty: symmetric ones from q+ = q_j + q_{n-1-j},
        // antisymmetric ones from q- (the rows are read once per parity,
        // which keeps one folded set live)
        // Every lane reads its own column (n >= kMh: rows j and n-1-j lie in
        // 0..n); rows past n/2 read qty row n, which holds 0, and lanes
        // without a column are zeroed after the load, so the reads need no
        // branch and go out together.
        float co[kMh], ce[kMh];
#pragma unroll
        for (int par = 0; par < 2; par++) {
            // a fence: the second parity's reads would otherwise be hoisted
            // into the first parity's transform, with 128 more live values
            asm volatile("" ::: "memory");
            float qp[kMh], qm[kMh];
            __builtin_amdgcn_s_setprio(RP_PRIO_MEM);   // the row reads, as in the column walk
#pragma unroll
            for (int j = 0; j < kMh; j++) {
                qp[j] = S(Lo.q + (j < half ? j : n));
                qm[j] = S(Lo.q + (j < half ? n - 1 - j : n));
            }
            // the folded rows: q+ for the odd modes, q- for the even ones
            // (the rows are read once per parity, so one folded set is live)
#pragma unroll
            for (int j = 0; j < kMh; j++) qp[j] = go ? (par == 0 ? qp[j] + qm[j] : qp[j] - qm[j]) : 0.0f;
            __builtin_amdgcn_s_setprio(0);
            if (par == 0) modal_gemm<false>(tab, [&](int j) { return qp[j]; }, co);
            else modal_gemm<false>(tab_e, [&](int j) { return qp[j]; }, ce);
        }
        asm volatile("" ::: "memory");
        // 3. the search and the accepted solution's modes
        double p = 0.0;
        ModalSol sol{0.0, 0.0, true};
        if (go) p = modal_search(MT, co, ce, h, A, sol);
        // 4. u = Phi vh: rows j and n-1-j from the symmetric part +/- the
        // antisymmetric part
        // (plain stores: the sub-steps read them soon).  Every lane writes
        // its own column; the pairs past n/2 go to qty rows, which are no
        // longer read.
        {
            // scalar reads of the constants (as in modal_sums)
            const __attribute__((address_space(4))) ModalTab* T =
                (const __attribute__((address_space(4))) ModalTab*)MT;
            asm volatile("" : "+s"(T));
            // mode m of the solution, formed where the transform takes it;
            // lanes without a column give zeros
            auto vh = [&](int par, float c, int m) -> float {
                const double R = sol.p1 ? 1.0 : modal_rcp(T->lam[par][m] + sol.t);
                const double v = (double)c * R * sol.scale;
                return go ? (float)v : 0.0f;
            };
            float uo[kMh], ue[kMh];
            modal_gemm<true>(tab, [&](int m) { return vh(0, co[m], m); }, uo);
            modal_gemm<true>(tab_e, [&](int m) { return vh(1, ce[m], m); }, ue);
#pragma unroll
            for (int j = 0; j < kMh; j++) {
                const bool in = j < half;
                S(in ? Lo.u + 1 + j : Lo.q + (j - half)) = uo[j] + ue[j];
                S(in ? Lo.u + (n - j) : Lo.q + (j - half)) = uo[j] - ue[j];
            }
            S(Lo.u + 0) = 0.0f;
            S(Lo.u + n + 1) = 0.0f;
        }
        __builtin_amdgcn_s_setprio(RP_PRIO_MEM);
        // 5. sub-steps and the rest of the step
        if (go) {
            const int64_t q = launch_particle(A, ne, r);
            VtCol col;
            col.z1 = ld_dbl(S, Lo.w + 0);
            col.rh = ld_dbl(S, Lo.w + 2);
            col.h = h;
            Surface s0;
            s0.tdepth = ld_dbl(S, Lo.w + 4);
            s0.zeta = ld_dbl(S, Lo.w + 6);
            DevSrc dv = dev_of(A, q);
            const double newX = A.mid[r], newY = A.mid[A.nmid + r];
            const double AdvectZ = A.mid[2 * A.nmid + r], Zpar = A.mid[3 * A.nmid + r];
            const double TurbV = vt_substeps_fast(A, Lo, S, col, s0, Zpar, p, dv);
            finish<Cur, NESTED>(grids, A, q, NESTED ? (int)A.gid[q] : 1, A.z[q], newX, newY, AdvectZ,
                                TurbV);
            nact++;
        }
    }
    count_active(A, nact);
}

// Occupancy targets (waves per SIMD) of the three step kernels; 0 leaves the
// choice to the compiler.
#ifndef RP_WAVES_ALL
#define RP_WAVES_ALL 0
#endif
#ifndef RP_WAVES_ADVECT
#define RP_WAVES_ADVECT 0
#endif
#ifndef RP_WAVES_VWALK
#define RP_WAVES_VWALK 0
#endif
#define RP_WAVES(n) __attribute__((amdgpu_waves_per_eu((n) > 0 ? (n) : 1)))

template <typename F, bool NESTED>
__global__ void __launch_bounds__(256) RP_WAVES(RP_WAVES_ALL)
    step_kernel(const DGrid* __restrict__ grids, StepArgs A)
{
    step_body<F, NESTED, 0>(grids, A);
}

// Waves per SIMD of the advection kernels: their RK stages wait on field
// gathers, which more waves hide, as long as the stages' state fits their
// registers.  The FP32-field form fits 3 waves (168 VGPRs) with a spill of
// 20 B a lane on one grid, 140 B with nesting.  Measured: FP64 form on C3
// 2.47 ms at 2 waves against 3.63 at 1 and 2.66 at 3; FP32 form on C3
// (nested) 1.41 ms advection phase at 2 against 1.51 at 3 and 1.76 at 4; on
// C4 (one grid) 17.55 ms per batched step at 3 against 18.10 at 2.
template <typename F, bool NESTED>
constexpr int advect_waves()
{
    return RP_WAVES_ADVECT > 0 ? RP_WAVES_ADVECT : (is_fast<F>() && !NESTED ? 3 : 2);
}

template <typename F, bool NESTED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(advect_waves<F, NESTED>())))
    advect_kernel(const DGrid* __restrict__ grids, StepArgs A)
{
    step_body<F, NESTED, 1>(grids, A);
}

// The FP32-field vertical walk is held to 5 waves per SIMD (96 VGPRs, a
// spill of about 250 B), the FP64 one to 6: their sweeps and gathers wait on
// memory, and more resident waves hide more of that than the spill costs
// (measured on C3: FP64 form 98.8 ms at 6 waves against 143.7 at the
// compiler's 2, 109.4 at 4 and 100.6 at 8).
template <typename F>
constexpr int vwalk_waves()
{
    return RP_WAVES_VWALK > 0 ? RP_WAVES_VWALK : (is_fast<F>() ? 5 : 6);
}

template <typename F, bool NESTED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(vwalk_waves<F>())))
    vwalk_kernel(const DGrid* __restrict__ grids, StepArgs A)
{
    vwalk_body<F, NESTED>(grids, A);
}

// Eigenbasis form: the mode registers (128 FP32 per lane) and the
// transforms' tiles need 2 waves per SIMD.
template <bool NESTED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    vwalk_modal_kernel(const DGrid* __restrict__ grids, StepArgs A)
{
    vwalk_modal_body<NESTED>(grids, A);
}
template __global__ void vwalk_modal_kernel<false>(const DGrid*, StepArgs);
template __global__ void vwalk_modal_kernel<true>(const DGrid*, StepArgs);

#define RP_INST(F, N)                                                              \
    template __global__ void step_kernel<F, N>(const DGrid*, StepArgs);            \
    template __global__ void advect_kernel<F, N>(const DGrid*, StepArgs);          \
    template __global__ void vwalk_kernel<F, N>(const DGrid*, StepArgs);
RP_INST(Cur, false) RP_INST(Cur, true) RP_INST(CurWet, false) RP_INST(CurWet, true)
RP_INST(double, false) RP_INST(double, true)
#undef RP_INST

// ---------------------------------------------------------------------------
// upload: Fortran (nx, ny, nk) FP64 column-major -> slot-interleaved layout
// dst[(((j*nx + i)*nk + k)*4 + slot], optional FP64 addend (Stokes drift)
// ---------------------------------------------------------------------------
template <typename F>
__global__ void __launch_bounds__(256) pack_field(const double* __restrict__ src,
                                                  const double* __restrict__ add, F* dst,
                                                  int nx, int ny, int nk, int slot)
{
    const int64_t total = (int64_t)nx * ny * nk;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        // e enumerates the destination order (k fastest) so stores coalesce
        const int k = (int)(e % nk);
        const int64_t col = e / nk;            // j*nx + i
        const int64_t s = col + (int64_t)k * nx * ny;
        double v = src[s];
        if (add) v = v + add[s];
        dst[e * 4 + slot] = (F)v;
    }
}

template __global__ void pack_field<float>(const double*, const double*, float*, int, int, int, int);
template __global__ void pack_field<double>(const double*, const double*, double*, int, int, int, int);

__global__ void sample_deviates(uint64_t seed, int64_t step, int64_t n, int ndev, double* out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    DevSrc dv;
    dv.inj = nullptr;
    dv.seed = seed;
    dv.id = (uint64_t)q;
    dv.step = (uint64_t)step;
    dv.blk = -1;
    for (int j = 0; j < ndev; j++) out[q * ndev + j] = dv.get(j);
}

// Tables of the p = 0 spline sweep (stream_G_p0): the forward LDL' sweep of
// the Toeplitz band (6, -4, 1) with derivatives along (4, 1, 0), per row
// 1/d, (1/d)^2 dd, l1, l2, dl1, dl2.  Uploaded once per device.
int upload_p0_tables()
{
    static bool done[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
    if (done[dev]) return 0;
    std::vector<double> T(6 * (size_t)kP0Rows);
    const double a = 6.0, b = -4.0, c = 1.0, da = 4.0, db = 1.0, dc = 0.0;
    double d = a, dd = da, e = b, de = db, dn = a, ddn = da, en = b, den = db;
    for (int m = 0; m < kP0Rows; m++) {
        const double rd = 1.0 / d;
        const double rat = e * rd, drat = (de - rat * dd) * rd;
        const double rat2 = c * rd, drat2 = (dc - rat2 * dd) * rd;
        double* t = &T[6 * (size_t)m];
        t[0] = rd; t[1] = rd * rd * dd; t[2] = rat; t[3] = rat2; t[4] = drat; t[5] = drat2;
        const double d1 = dn - rat * e, dd1 = ddn - drat * e - rat * de;
        const double e1 = en - rat * c, de1 = den - drat * c - rat * dc;
        const double d2 = a - rat2 * c, dd2 = da - drat2 * c - rat2 * dc;
        d = d1; dd = dd1; e = e1; de = de1;
        dn = d2; ddn = dd2; en = b; den = db;
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(kP0Tab), T.data(), T.size() * sizeof(double)) != hipSuccess)
        return -1;
    done[dev] = true;
    return 0;
}

__global__ void __launch_bounds__(256) sample_lonlat(LonLatArgs A)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < A.n;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int gi = (int)A.gid[q] - 1;
        DGrid g;
        g.L = A.L[gi];
        g.M = A.M[gi];
        const RhoCell c = rho_cell(g, A.x[q], A.y[q]);
        A.out_lon[q] = bilinear2d(A.lon[gi], c);
        A.out_lat[q] = bilinear2d(A.lat[gi], c);
    }
}

__global__ void __launch_bounds__(256) current_fields(CurJobs J)
{
    const int j = blockIdx.y;
    if (j >= J.njobs) return;
    const float* src = J.src[j];
    float* dst = J.dst[j];
    const int s0 = J.slot[j][0], s1 = J.slot[j][1], s2 = J.slot[j][2];
    const double p0 = J.pw[0], p1 = J.pw[1], p2 = J.pw[2];
    const int kin = J.kin[j], kout = J.kout[j];
    const uint8_t* wd = J.wd[j];
    const int64_t wdp = J.wdp[j];
    const int nk = J.nk[j];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < J.n[j];
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t si = e;
        if (kin != kout) {
            const int64_t col = e / kout;
            const int k = (int)(e - col * kout);
            if (k >= kin) {
                dst[e] = 0.0f;
                continue;
            }
            si = col * kin + k;
        }
        const float4 v = reinterpret_cast<const float4*>(src)[si];
        auto pk = [&](int t) { return (double)(t == 0 ? v.x : (t == 1 ? v.y : (t == 2 ? v.z : v.w))); };
        if (wd) {
            // each slot's value where that slot's mask is wet, 0 where dry
            const int64_t pt = si / nk;
            const double a0 = wd[s0 * wdp + pt] ? pk(s0) : 0.0;
            const double a1 = wd[s1 * wdp + pt] ? pk(s1) : 0.0;
            const double a2 = wd[s2 * wdp + pt] ? pk(s2) : 0.0;
            dst[e] = (float)(p0 * a0 + p1 * a1 + p2 * a2);
        } else {
            dst[e] = (float)(p0 * pk(s0) + p1 * pk(s1) + p2 * pk(s2));
        }
    }
}

}  // namespace rp
