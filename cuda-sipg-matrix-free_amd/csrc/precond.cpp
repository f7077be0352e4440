// precond.cpp -- host tables of the block-Jacobi preconditioner.
//
// The diagonal block of A for an interior element (both neighbours in each
// direction present) is the Kronecker sum  U -> K U W + W U K, where W =
// diag(w) and K = K1 + F is the 1-D stiffness plus the element's own face
// terms of flux_term6a/6b (sipg_sem_2d_gpu_kernels.hpp:212-247, 279-312)
// with both faces interior (the 1/2 averages and the 2 pen weight):
//
//   F[i][k] = 1/2 D_i(0) [k = 0] - 1/2 D_i(N) [k = N]            (symmetry)
//           + [i = 0] (1/2 D_k(0) + pen [k = 0])                (consistency +
//           + [i = N] (-1/2 D_k(N) + pen [k = N])                 penalty)
//
// with D_i(k) = phi'_i(z_k).  The generalized eigenproblem K S = W S Lambda
// (S^T W S = I) diagonalises the block: its inverse is
// U = S [ (S^T R S) ./ (lambda_a + lambda_c) ] S^T.  It is solved through the
// symmetric matrix W^-1/2 K W^-1/2 with cyclic Jacobi rotations (n <= 17).
#include "sipg_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace sipg_internal {

namespace {

// Eigen-decomposition of the symmetric n x n matrix A (row-major, destroyed):
// A = V diag(lam) V^T, V orthogonal (columns are eigenvectors).
void jacobi_eigen(int n, std::vector<double> &A, std::vector<double> &V, std::vector<double> &lam) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    double scale = 0.0;
    for (double v : A) scale += v * v;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
        if (off <= 1e-34 * scale) break;
        for (int p = 0; p < n; ++p) {
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                // rotation angle that zeroes A[p][q]: tan(2 phi) = 2 apq / (aqq - app)
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {           // columns p, q of A
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {           // rows p, q of A
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {           // columns p, q of V
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
        }
    }
    lam.resize(n);
    for (int i = 0; i < n; ++i) lam[i] = A[(size_t)i * n + i];
}

}  // namespace

int build_block_jacobi(int p, double pen, std::vector<double> &tab) {
    if (p < 1 || !(pen > 0.0)) return SIPG_ERR_INVALID;
    if (p > kMaxDegree) return SIPG_ERR_UNSUPPORTED;     // no compiled preconditioner kernel
    const int n = p + 1, N = p;
    std::vector<double> z(n), w(n), D((size_t)n * n), K1((size_t)n * n);
    gll(p, z.data(), w.data());
    dphi(p, z.data(), D.data());
    stiffness_1d(p, w.data(), D.data(), K1.data());
    auto d = [&](int mode, int node) { return D[(size_t)mode * n + node]; };
    std::vector<double> K(K1);
    for (int i = 0; i < n; ++i) {
        K[(size_t)i * n + 0] += 0.5 * d(i, 0);
        K[(size_t)i * n + N] -= 0.5 * d(i, N);
    }
    for (int k = 0; k < n; ++k) {
        K[(size_t)0 * n + k] += 0.5 * d(k, 0);
        K[(size_t)N * n + k] -= 0.5 * d(k, N);
    }
    K[0] += pen;
    K[(size_t)N * n + N] += pen;
    // W^-1/2 K W^-1/2, symmetrised against rounding
    std::vector<double> A((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k)
            A[(size_t)i * n + k] = 0.5 * (K[(size_t)i * n + k] + K[(size_t)k * n + i]) / std::sqrt(w[i] * w[k]);
    std::vector<double> V, lam;
    jacobi_eigen(n, A, V, lam);
    for (double l : lam)
        if (!(l > 0.0) || !std::isfinite(l)) return SIPG_ERR_INVALID;
    tab.assign((size_t)2 * n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < n; ++a) tab[(size_t)i * n + a] = V[(size_t)i * n + a] / std::sqrt(w[i]);
    for (int a = 0; a < n; ++a)
        for (int c = 0; c < n; ++c) tab[(size_t)n * n + (size_t)a * n + c] = 1.0 / (lam[a] + lam[c]);
    return SIPG_OK;
}

// Coarse level (kernels_coarse.hip): the sine transform of size n = dim - 1,
// Q[j][k] = sqrt(2/dim) sin((j+1)(k+1) pi / dim) (symmetric, orthogonal),
// and the eigenvalues of the 1-D Q1 stiffness tridiag(-1, 2, -1) and mass
// tridiag(m1, 2 m0, m1) it diagonalises, the mass by the GLL quadrature of
// degree p on [0, 1]: m0 = sum w/2 (1-xi)^2, m1 = sum w/2 xi (1-xi).
void coarse_dst(int dim, std::vector<double> &Q, std::vector<double> &lamK, std::vector<double> &lamM, int p) {
    const int n = dim - 1;
    Q.assign((size_t)std::max(n, 0) * std::max(n, 0), 0.0);
    lamK.assign(std::max(n, 0), 0.0);
    lamM.assign(std::max(n, 0), 0.0);
    if (n < 1) return;
    std::vector<double> z(p + 1), w(p + 1);
    gll(p, z.data(), w.data());
    double m0 = 0.0, m1 = 0.0;
    for (int q = 0; q <= p; ++q) {
        const double xi = 0.5 * (z[q] + 1.0);
        m0 += 0.5 * w[q] * (1.0 - xi) * (1.0 - xi);
        m1 += 0.5 * w[q] * xi * (1.0 - xi);
    }
    const double pi = 3.14159265358979323846, sc = std::sqrt(2.0 / dim);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) Q[(size_t)j * n + k] = sc * std::sin((double)(j + 1) * (k + 1) * pi / dim);
    for (int k = 0; k < n; ++k) {
        const double c = std::cos((k + 1) * pi / dim);
        lamK[k] = 2.0 - 2.0 * c;
        lamM[k] = 2.0 * m0 + 2.0 * m1 * c;
    }
}

}  // namespace sipg_internal
