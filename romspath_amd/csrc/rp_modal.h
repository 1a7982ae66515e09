// rp_modal.h -- the smoothing-spline secant search of the FP32-field form in
// the eigenbasis of chol1d's matrix pencil, with its two basis changes on
// the matrix cores.
//
// chol1d's matrix for n = np-2 interior unknowns on the uniform resampling
// grid (spacing h, r = serr/h; pppack.f90:1509-1512 with setupq's
// v(:,1..7), pppack.f90:4793-4813) is
//     T(p) = alpha (P + t Tri),   alpha = 6 (1-p) r^2,   t = p h / alpha,
// with P = pentadiag(1,-4,6,-4,1) and Tri = tridiag(1,4,1), neither
// depending on the particle.  The pencil (P, Tri) has a Tri-orthonormal
// eigenbasis Phi (Phi' Tri Phi = I, Phi' P Phi = diag(lambda)), so with
// d = Phi' qty
//     g(t) = qty'(P + t Tri)^-1 qty = sum_k d_k^2 / (lambda_k + t),
//     g'(t) = -sum_k d_k^2 / (lambda_k + t)^2,
// and for v = (P + t Tri)^-1 qty, v'P v = g + t g'.  smooth's secant search
// (pppack.f90:5236-5298) needs only
//     sfq = 36 u'V u / (1+q)^2 = 36 r^2 (g + t g') / (alpha^2 (1+q)^2),
//     utru (p = 0) = u'W u / 2 = -h g'(0) / (2 alpha^2),
// so each of its evaluations is one pass over the n modes held in registers
// (no band elimination, no work rows), and the accepted p's solution is
//     u = Phi vh,  vh_k = d_k / ((lambda_k + t) alpha)
// (at p = 1, T = h Tri and vh_k = d_k / h).  Both matrices commute with the
// reversal of the rows, so for even n the basis splits into n/2 symmetric and
// n/2 antisymmetric vectors: the first see only q+ (q_j + q_{n-1-j}), the
// second only q- (q_j - q_{n-1-j}).  The host computes each half's pencil in
// long double (Cholesky of the Tri block, cyclic Jacobi on L^-1 P L^-T).
//
// The transforms d = Phi' qty and u = Phi vh act on the 64 particles of a
// wave at once: per parity a (64 x 64) x (64 x 64) product in FP32 on
// v_mfma_f32_32x32x2f32 (2 x 2 output tiles of 32 x 32, 32 steps of K = 2),
// with the basis in LDS.  A lane holds its particle's rows in registers;
// v_permlane32_swap turns rows 2s and 2s+1 of every lane into the B operand
// of both column tiles, and turns the accumulators back into one particle
// per lane.  The FP32 rounding of d (about 1e-7 relative, the same as the
// FP32 rounding of qty that the form stores) moves the accepted p by about
// 5e-8 relative; the sums themselves run in FP64.
#pragma once

#include <cstdint>

namespace rp {

constexpr int kMh = 64;        // modes (and folded rows) per parity, padded
constexpr int kMStride = 65;   // LDS row stride of a basis table (odd: no bank conflicts)
constexpr int kModalLds = 2 * kMh * kMStride * 4;   // bytes of LDS for both tables

// Tables of one resampling size n (even, kMh <= n <= 2 kMh), built on the
// host (rp_modal_basis).
struct ModalTab {
    // basis[par][k][j]: component j (j < n/2) of eigenvector k of parity par
    // (0 symmetric, 1 antisymmetric), divided by sqrt(2), so that
    // d_k = sum_j basis[par][k][j] q+-_j; zero for k or j >= n/2
    float basis[2][kMh][kMh];
    // lambda_k, ascending; padded modes hold 1 (their d is 0)
    double lam[2][kMh];
    int n, half;
};

typedef float rp_f32x16 __attribute__((ext_vector_type(16)));

// (lo, hi) halves of a wave: returns {[x_lo, y_lo], [x_hi, y_hi]}
__device__ __forceinline__ void swap32(float x, float y, float& xo, float& yo)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(int, x),
                                                    __builtin_bit_cast(int, y), false, false);
    // the elements go through plain ints: a bit cast of a vector element
    // lvalue reads element 0 whatever the index
    const int r0 = r[0], r1 = r[1];
    xo = __builtin_bit_cast(float, r0);
    yo = __builtin_bit_cast(float, r1);
}

// out[j] (lane = particle) = sum_m M[j][m] in(m), for the 64 particles of the
// wave, with M read from the LDS table tab ([64][kMStride]):
//   TRANS = false: M[j][m] = tab[j][m]  (forward: modes from folded rows)
//   TRANS = true:  M[j][m] = tab[m][j]  (inverse: folded rows from modes)
// Every lane of the wave must execute this (the matrix instructions and the
// lane swaps use all 64 lanes).
// IN(j) gives row j of the lane's input (called once per j, in order, just
// before its matrix instruction).
template <bool TRANS, class IN>
__device__ __forceinline__ void modal_gemm(const float* __restrict__ tab, IN in, float (&out)[kMh])
{
    const int l = (int)(threadIdx.x & 63);
    const int lm = l & 31, lk = l >> 5;
    rp_f32x16 acc[2][2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[mt][nt][i] = 0.0f;
#pragma unroll
    for (int s = 0; s < kMh / 2; s++) {
        // B operand (K = 2 rows x 32 particles) of column tiles 0 and 1
        float b0, b1;
        swap32(in(2 * s), in(2 * s + 1), b0, b1);
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            // A operand: row 32 mt + lm of M, column 2 s + lk
            const int row = 32 * mt + lm, col = 2 * s + lk;
            const float a = TRANS ? tab[col * kMStride + row] : tab[row * kMStride + col];
            acc[mt][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b0, acc[mt][0], 0, 0, 0);
            acc[mt][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b1, acc[mt][1], 0, 0, 0);
        }
    }
    // accumulator register i of lane l holds row 8(i/4) + 4(l/32) + i%4 of
    // its tile, particle 32 nt + l%32; the swap gives each lane its own
    // particle: the first value rows 8(i/4) + i%4, the second 8(i/4) + 4 + i%4
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int r0 = 32 * mt + 8 * (i / 4) + (i % 4);
            swap32(acc[mt][0][i], acc[mt][1][i], out[r0], out[r0 + 4]);
        }
}

// the sums of one evaluation at t: sum d^2/(lambda + t) and
// sum d^2/(lambda + t)^2
struct ModalSums {
    double g, gp2;
};

// 1/d to 2.2e-15 relative (hardware estimate and one Newton step)
__device__ __forceinline__ double modal_rcp(double d)
{
    const double r = __builtin_amdgcn_rcp(d);
    const double e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ void modal_sums(const ModalTab* __restrict__ T0, float (&co)[kMh],
                                           float (&ce)[kMh], double t, ModalSums& S)
{
#pragma clang fp contract(fast)
    // The table reads and the widening of the modes to FP64 are the same on
    // every evaluation; left visible, the compiler hoists them out of the
    // secant loop and keeps the constants and 128 doubles live (and spills
    // them).  An opaque copy of the table pointer (in the constant address
    // space, so the reads are scalar) and an empty asm on each mode make
    // every evaluation read the constants through the scalar cache and widen
    // the modes again (the values do not change).
    const __attribute__((address_space(4))) ModalTab* T =
        (const __attribute__((address_space(4))) ModalTab*)T0;
    asm volatile("" : "+s"(T));
    double g = 0.0, gp2 = 0.0;
#pragma unroll
    for (int par = 0; par < 2; par++) {
#pragma unroll
        for (int m = 0; m < kMh; m++) {
            // the hardware reciprocal alone (4.6e-8 relative): the sums then
            // carry about the FP32 rounding of the modes themselves
            const double R = __builtin_amdgcn_rcp(T->lam[par][m] + t);
            float& cm = par == 0 ? co[m] : ce[m];
            asm("" : "+v"(cm));
            const double d = (double)cm;
            const double dR = d * R;
            g = fma(d, dR, g);
            gp2 = fma(dR, dR, gp2);
        }
    }
    S.g = g;
    S.gp2 = gp2;
}

}  // namespace rp
