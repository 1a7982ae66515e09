// rp_modal_host.cpp -- host construction of the eigenbasis tables of the
// FP32-field vertical walk (rp_modal.h): for n interior unknowns of the
// resampled column, the generalized eigenvectors of the pencil
// (P, Tri), P = pentadiag(1,-4,6,-4,1), Tri = tridiag(1,4,1) (chol1d's
// matrix at unit spacing, pppack.f90:1509-1512), split into the n/2
// symmetric and n/2 antisymmetric ones.  Computed in long double: Cholesky
// Tri_h = L L', cyclic Jacobi on C = L^-1 P_h L^-T = W diag(lambda) W',
// Phi = L^-T W (Phi' Tri_h Phi = I).  The smallest eigenvalue is about
// 3.5e-7 for n = 122, so the absolute accuracy of the Jacobi sweeps in long
// double (about 1e-18) leaves it to about 1e-11 relative.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "rp_abi.h"

namespace {

using LD = long double;

// eigenvalues w and orthonormal eigenvectors (columns of V) of the m x m
// symmetric A (row-major, overwritten), by cyclic Jacobi rotations
void jacobi(int m, std::vector<LD>& A, std::vector<LD>& V, std::vector<LD>& w)
{
    V.assign((size_t)m * m, 0.0L);
    for (int i = 0; i < m; i++) V[(size_t)i * m + i] = 1.0L;
    for (int sweep = 0; sweep < 100; sweep++) {
        LD off = 0.0L, diag = 0.0L;
        for (int i = 0; i < m; i++) {
            diag += A[(size_t)i * m + i] * A[(size_t)i * m + i];
            for (int j = i + 1; j < m; j++) off += A[(size_t)i * m + j] * A[(size_t)i * m + j];
        }
        if (off <= diag * 1e-40L) break;
        for (int p = 0; p < m - 1; p++)
            for (int q = p + 1; q < m; q++) {
                const LD apq = A[(size_t)p * m + q];
                if (apq == 0.0L) continue;
                const LD theta = (A[(size_t)q * m + q] - A[(size_t)p * m + p]) / (2.0L * apq);
                const LD t = (theta >= 0.0L ? 1.0L : -1.0L) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0L));
                const LD c = 1.0L / std::sqrt(t * t + 1.0L), s = t * c;
                for (int k = 0; k < m; k++) {
                    const LD akp = A[(size_t)k * m + p], akq = A[(size_t)k * m + q];
                    A[(size_t)k * m + p] = c * akp - s * akq;
                    A[(size_t)k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; k++) {
                    const LD apk = A[(size_t)p * m + k], aqk = A[(size_t)q * m + k];
                    A[(size_t)p * m + k] = c * apk - s * aqk;
                    A[(size_t)q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; k++) {
                    const LD vkp = V[(size_t)k * m + p], vkq = V[(size_t)k * m + q];
                    V[(size_t)k * m + p] = c * vkp - s * vkq;
                    V[(size_t)k * m + q] = s * vkp + c * vkq;
                }
            }
    }
    w.resize(m);
    for (int i = 0; i < m; i++) w[i] = A[(size_t)i * m + i];
}

// entry (i, j) of the full n x n P (which = 0) or Tri (which = 1)
LD band(int which, int i, int j)
{
    const int d = std::abs(i - j);
    if (which == 0) return d == 0 ? 6.0L : (d == 1 ? -4.0L : (d == 2 ? 1.0L : 0.0L));
    return d == 0 ? 4.0L : (d == 1 ? 1.0L : 0.0L);
}

}  // namespace

extern "C" int rp_modal_basis(int32_t n, float* basis, double* lam)
{
    constexpr int kMh = 64;
    if (!basis || !lam || n < kMh || n % 2 != 0 || n / 2 > kMh) return RP_E_ARG;
    const int m = n / 2;
    std::memset(basis, 0, sizeof(float) * 2 * kMh * kMh);
    for (int par = 0; par < 2; par++) {
        // the pencil restricted to the (anti)symmetric vectors
        // (e_j +- e_{n-1-j}) / sqrt(2), j < m
        const LD sg = par == 0 ? 1.0L : -1.0L;
        std::vector<LD> Ph((size_t)m * m), Th((size_t)m * m);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                Ph[(size_t)i * m + j] = band(0, i, j) + sg * band(0, i, n - 1 - j);
                Th[(size_t)i * m + j] = band(1, i, j) + sg * band(1, i, n - 1 - j);
            }
        // Th = L L'
        std::vector<LD> L((size_t)m * m, 0.0L);
        for (int j = 0; j < m; j++) {
            LD d = Th[(size_t)j * m + j];
            for (int k = 0; k < j; k++) d -= L[(size_t)j * m + k] * L[(size_t)j * m + k];
            if (!(d > 0.0L)) return RP_E_ARG;
            L[(size_t)j * m + j] = std::sqrt(d);
            for (int i = j + 1; i < m; i++) {
                LD v = Th[(size_t)i * m + j];
                for (int k = 0; k < j; k++) v -= L[(size_t)i * m + k] * L[(size_t)j * m + k];
                L[(size_t)i * m + j] = v / L[(size_t)j * m + j];
            }
        }
        // X = L^-1 Ph (columns by forward substitution), C = L^-1 X'
        auto fwd = [&](std::vector<LD>& B) {   // B <- L^-1 B, B m x m row-major
            for (int c = 0; c < m; c++)
                for (int i = 0; i < m; i++) {
                    LD v = B[(size_t)i * m + c];
                    for (int k = 0; k < i; k++) v -= L[(size_t)i * m + k] * B[(size_t)k * m + c];
                    B[(size_t)i * m + c] = v / L[(size_t)i * m + i];
                }
        };
        std::vector<LD> X = Ph;
        fwd(X);
        std::vector<LD> C((size_t)m * m);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) C[(size_t)i * m + j] = X[(size_t)j * m + i];
        fwd(C);
        for (int i = 0; i < m; i++)
            for (int j = i + 1; j < m; j++) {
                const LD a = 0.5L * (C[(size_t)i * m + j] + C[(size_t)j * m + i]);
                C[(size_t)i * m + j] = C[(size_t)j * m + i] = a;
            }
        std::vector<LD> W, w;
        jacobi(m, C, W, w);
        // Phi = L^-T W (back substitution), columns in ascending lambda
        std::vector<LD> Phi = W;
        for (int c = 0; c < m; c++)
            for (int i = m - 1; i >= 0; i--) {
                LD v = Phi[(size_t)i * m + c];
                for (int k = i + 1; k < m; k++) v -= L[(size_t)k * m + i] * Phi[(size_t)k * m + c];
                Phi[(size_t)i * m + c] = v / L[(size_t)i * m + i];
            }
        std::vector<int> ord(m);
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return w[a] < w[b]; });
        const LD r2 = 1.0L / std::sqrt(2.0L);
        for (int k = 0; k < kMh; k++) {
            lam[par * kMh + k] = k < m ? (double)w[ord[k]] : 1.0;
            if (k < m)
                for (int j = 0; j < m; j++)
                    basis[((size_t)par * kMh + k) * kMh + j] = (float)(Phi[(size_t)j * m + ord[k]] * r2);
        }
    }
    return RP_OK;
}
