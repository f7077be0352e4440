// basis.cpp -- host-side spectral-element numerics that feed the device
// tables: Legendre-Gauss-Lobatto nodes and weights, the Lagrange derivative
// table and the 1-D stiffness matrix.
//
// Reference counterparts: LGL_quadrature_table.hpp:84-146 (nodes, weights),
// sem_function.hpp:86-110 (derivative table), device_const.hpp:115-127 (its
// layout), laplacian_operator.hpp:31-43 (1-D stiffness).  The node solver
// here is Newton on q = P_{N+1} - P_{N-1} (whose interior roots are the LGL
// points) from the asymptotic guesses of Kopriva, "Implementing Spectral
// Methods for PDEs" (2009), Alg. 25; the reference uses Newton on
// (1-x^2)P_N' from Lether's guesses.  Both converge to the same roots.
#include "sipg_internal.hpp"

#include <cmath>
#include <limits>

namespace sipg_internal {

namespace {

// q = P_{N+1} - P_{N-1} and q' at x, from the recurrences
// (k+1) P_{k+1} = (2k+1) x P_k - k P_{k-1},  P'_{k+1} = P'_{k-1} + (2k+1) P_k.
void legendre_q(int N, double x, double &q, double &dq) {
    std::vector<double> L(N + 2), dL(N + 2);
    L[0] = 1.0; dL[0] = 0.0;
    L[1] = x;   dL[1] = 1.0;
    for (int k = 1; k <= N; ++k) {
        L[k + 1] = ((2.0 * k + 1.0) * x * L[k] - k * L[k - 1]) / (k + 1.0);
        dL[k + 1] = dL[k - 1] + (2.0 * k + 1.0) * L[k];
    }
    q = L[N + 1] - L[N - 1];
    dq = dL[N + 1] - dL[N - 1];
}

double legendre_value(int N, double x) {
    if (N == 0) return 1.0;
    double L0 = 1.0, L1 = x;
    for (int k = 2; k <= N; ++k) {
        const double L2 = ((2.0 * k - 1.0) * x * L1 - (k - 1.0) * L0) / k;
        L0 = L1;
        L1 = L2;
    }
    return L1;
}

}  // namespace

void gll(int p, double *z, double *w) {
    const int N = p;
    const double eps = std::numeric_limits<double>::epsilon();
    z[0] = -1.0;
    z[N] = 1.0;
    for (int j = 1; 2 * j < N; ++j) {
        const double t = (j + 0.25) * M_PI / N - 3.0 / (8.0 * N * M_PI * (j + 0.25));
        double x = -std::cos(t);
        for (int it = 0; it < 100; ++it) {
            double q, dq;
            legendre_q(N, x, q, dq);
            const double d = -q / dq;
            x += d;
            if (std::fabs(d) <= 4.0 * eps * std::fabs(x)) break;
        }
        z[j] = x;
        z[N - j] = -x;
    }
    if (N % 2 == 0) z[N / 2] = 0.0;
    for (int j = 0; j <= N; ++j) {
        const double L = legendre_value(N, z[j]);
        w[j] = 2.0 / (N * (N + 1.0) * L * L);
    }
}

void dphi(int p, const double *z, double *D) {
    const int n1 = p + 1;
    std::vector<double> L(n1);
    for (int i = 0; i < n1; ++i) L[i] = legendre_value(p, z[i]);
    for (int mode = 0; mode < n1; ++mode)
        for (int node = 0; node < n1; ++node) {
            double v;
            if (mode != node)
                v = L[node] / (L[mode] * (z[node] - z[mode]));
            else if (node == 0)
                v = -0.25 * p * (p + 1.0);
            else if (node == p)
                v = 0.25 * p * (p + 1.0);
            else
                v = 0.0;
            D[mode * n1 + node] = v;
        }
}

void stiffness_1d(int p, const double *w, const double *D, double *K1) {
    const int n1 = p + 1;
    for (int a = 0; a < n1; ++a)
        for (int b = 0; b < n1; ++b) {
            double r = 0.0;
            for (int k = 0; k < n1; ++k) r += w[k] * D[a * n1 + k] * D[b * n1 + k];
            K1[a * n1 + b] = r;
        }
}

void build_tables(int p, std::vector<double> &tab) {
    const int n1 = p + 1;
    std::vector<double> z(n1), w(n1), D(n1 * n1), K(n1 * n1);
    gll(p, z.data(), w.data());
    dphi(p, z.data(), D.data());
    stiffness_1d(p, w.data(), D.data(), K.data());
    tab.assign(table_len(p), 0.0);
    double *t = tab.data();
    for (int i = 0; i < n1; ++i) t[i] = w[i];
    for (int i = 0; i < n1 * n1; ++i) t[n1 + i] = K[i];
    for (int i = 0; i < n1; ++i) {
        t[n1 + n1 * n1 + i] = D[i * n1 + 0];
        t[2 * n1 + n1 * n1 + i] = D[i * n1 + p];
        t[3 * n1 + n1 * n1 + i] = z[i];
    }
}

}  // namespace sipg_internal
