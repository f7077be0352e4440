// basis_constexpr.hpp -- the degree-p spectral-element tables evaluated at
// compile time, so the fused kernels (templated on p) see their
// coefficients as constants: the compiler rematerialises them instead of
// keeping ~(p+1)^2 table values live in registers.
//
// Same mathematics as basis.cpp (reference: LGL_quadrature_table.hpp:84-146,
// sem_function.hpp:86-110, laplacian_operator.hpp:31-43): LGL nodes are the
// endpoints plus the roots of P_p', found by Newton from Chebyshev-Gauss-
// Lobatto guesses; weights 2/(p(p+1) P_p(z)^2); D(mode,node) the Lagrange
// derivative; K1 = D W D^T.  The tests compare the tables the kernels use
// with the host ones and with the oracle.
#pragma once

namespace sipg_internal {
namespace cx {

constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double cabs(double x) { return x < 0 ? -x : x; }

// cos by argument reduction to [-pi, pi] and a Taylor series (enough terms
// for full double accuracy); only used for Newton starting points.
constexpr double ccos(double x) {
    while (x > kPi) x -= 2 * kPi;
    while (x < -kPi) x += 2 * kPi;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 30; ++k) {
        term *= -x * x / ((2 * k - 1) * (2 * k));
        sum += term;
    }
    return sum;
}

// P_n(x) and P_n'(x)
constexpr void legendre(int n, double x, double &P, double &dP) {
    double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
    if (n == 0) { P = 1.0; dP = 0.0; return; }
    for (int k = 1; k < n; ++k) {
        const double p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
        const double d2 = d0 + (2.0 * k + 1.0) * p1;
        p0 = p1; p1 = p2;
        d0 = d1; d1 = d2;
    }
    P = p1;
    dP = d1;
}

template <int P>
struct Tables {
    static constexpr int N1 = P + 1;
    double z[N1] = {};
    double w[N1] = {};
    double D[N1][N1] = {};    // D[mode][node] = phi'_mode(z_node)
    double K1[N1][N1] = {};   // sum_k w_k D[a][k] D[b][k]
};

template <int P>
constexpr Tables<P> make_tables() {
    Tables<P> t{};
    constexpr int N = P;
    t.z[0] = -1.0;
    t.z[N] = 1.0;
    for (int j = 1; 2 * j < N; ++j) {
        double x = -ccos(kPi * j / N);
        for (int it = 0; it < 100; ++it) {
            double Pv = 0, dP = 0;
            legendre(N, x, Pv, dP);
            const double d2P = (2.0 * x * dP - N * (N + 1.0) * Pv) / (1.0 - x * x);
            const double dx = dP / d2P;
            x -= dx;
            if (cabs(dx) <= 1e-17) break;
        }
        t.z[j] = x;
        t.z[N - j] = -x;
    }
    if (N % 2 == 0) t.z[N / 2] = 0.0;
    double Lz[P + 1] = {};
    for (int i = 0; i <= N; ++i) {
        double Pv = 0, dP = 0;
        legendre(N, t.z[i], Pv, dP);
        Lz[i] = Pv;
        t.w[i] = 2.0 / (N * (N + 1.0) * Pv * Pv);
    }
    for (int m = 0; m <= N; ++m)
        for (int q = 0; q <= N; ++q) {
            double v = 0.0;
            if (m != q) v = Lz[q] / (Lz[m] * (t.z[q] - t.z[m]));
            else if (q == 0) v = -0.25 * N * (N + 1.0);
            else if (q == N) v = 0.25 * N * (N + 1.0);
            t.D[m][q] = v;
        }
    for (int a = 0; a <= N; ++a)
        for (int b = 0; b <= N; ++b) {
            double r = 0.0;
            for (int k = 0; k <= N; ++k) r += t.w[k] * t.D[a][k] * t.D[b][k];
            t.K1[a][b] = r;
        }
    return t;
}

// Even-odd split of the centrosymmetric K1 (K1[a][b] = K1[N-a][N-b] on the
// symmetric LGL nodes): with ue[i] = u[i] + u[N-i], uo[i] = u[i] - u[N-i]
// (i < H = (p+1)/2, and ue[H] = u[H] for odd p+1),
//   (K1 u)[a]   = sum_i E[a][i] ue[i] + sum_i O[a][i] uo[i]
//   (K1 u)[N-a] = sum_i E[a][i] ue[i] - sum_i O[a][i] uo[i]
// with E[a][i] = (K1[a][i] + K1[a][N-i]) / 2, O[a][i] = (K1[a][i] - K1[a][N-i]) / 2
// and E[a][H] = K1[a][H].  (p+1)^2 multiply-adds become about half.
template <int P>
struct EvenOdd {
    static constexpr int N1 = P + 1;
    static constexpr int H = N1 / 2;            // pairs
    static constexpr int ME = (N1 + 1) / 2;     // even size (H, plus the middle node)
    double E[ME][ME] = {};
    double O[H > 0 ? H : 1][H > 0 ? H : 1] = {};
    // face derivative at node 0, D0[i] = D[i][0], split the same way:
    // DA[i] = (D0[i] + D0[p-i]) / 2, DB[i] = (D0[i] - D0[p-i]) / 2, and the
    // middle node's D0[p/2] when p is even.  With DP[i] = -D0[p-i] these
    // give sum D0 u = S + B and sum DP u = B - S (S over pair sums, B over
    // pair differences).
    double DA[H > 0 ? H : 1] = {};
    double DB[H > 0 ? H : 1] = {};
    double DM = 0.0;
};

template <int P>
constexpr EvenOdd<P> make_even_odd() {
    const Tables<P> t = make_tables<P>();
    EvenOdd<P> eo{};
    constexpr int N = P, H = EvenOdd<P>::H, ME = EvenOdd<P>::ME;
    for (int a = 0; a < ME; ++a) {
        for (int i = 0; i < H; ++i) {
            eo.E[a][i] = 0.5 * (t.K1[a][i] + t.K1[a][N - i]);
            if (a < H) eo.O[a][i] = 0.5 * (t.K1[a][i] - t.K1[a][N - i]);
        }
        if (ME > H) eo.E[a][H] = t.K1[a][H];
    }
    for (int i = 0; i < H; ++i) {
        eo.DA[i] = 0.5 * (t.D[i][0] + t.D[N - i][0]);
        eo.DB[i] = 0.5 * (t.D[i][0] - t.D[N - i][0]);
    }
    if (ME > H) eo.DM = t.D[H][0];
    return eo;
}

constexpr double csqrt(double x) {
    if (!(x > 0.0)) return 0.0;
    double g = x > 1.0 ? x : 1.0;
    for (int i = 0; i < 200; ++i) {
        const double ng = 0.5 * (g + x / g);
        if (ng == g) break;
        g = ng;
    }
    return g;
}

// Block-Jacobi tables for the standard penalty pen = 5 (p+1)^2, the same
// construction as the host's build_block_jacobi (precond.cpp): S with
// K S = W S diag(lambda), S^T W S = I, for K = K1 + the interior element's
// own face terms, and Linv[a][c] = 1 / (lambda_a + lambda_c).  Constants let
// the preconditioner kernel take its coefficients as immediates.
template <int P>
struct PrecTables {
    double S[P + 1][P + 1] = {};
    double Linv[P + 1][P + 1] = {};
};

template <int P>
constexpr PrecTables<P> make_prec() {
    const Tables<P> t = make_tables<P>();
    constexpr int n = P + 1, N = P;
    const double pen = 5.0 * n * n;
    double K[n][n] = {};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) K[i][k] = t.K1[i][k];
    for (int i = 0; i < n; ++i) {
        K[i][0] += 0.5 * t.D[i][0];
        K[i][N] -= 0.5 * t.D[i][N];
    }
    for (int k = 0; k < n; ++k) {
        K[0][k] += 0.5 * t.D[k][0];
        K[N][k] -= 0.5 * t.D[k][N];
    }
    K[0][0] += pen;
    K[N][N] += pen;
    double A[n][n] = {}, V[n][n] = {};
    double scale = 0.0;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            A[i][k] = 0.5 * (K[i][k] + K[k][i]) / csqrt(t.w[i] * t.w[k]);
            scale += A[i][k] * A[i][k];
        }
    for (int i = 0; i < n; ++i) V[i][i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {          // cyclic Jacobi rotations
        double off = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) off += A[i][j] * A[i][j];
        if (off <= 1e-34 * scale) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p][q];
                if (cabs(apq) < 1e-300) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                const double tt = (theta >= 0 ? 1.0 : -1.0) / (cabs(theta) + csqrt(theta * theta + 1.0));
                const double c = 1.0 / csqrt(tt * tt + 1.0), s = tt * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    PrecTables<P> r{};
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < n; ++a) r.S[i][a] = V[i][a] / csqrt(t.w[i]);
    for (int a = 0; a < n; ++a)
        for (int c = 0; c < n; ++c) r.Linv[a][c] = 1.0 / (A[a][a] + A[c][c]);
    return r;
}

// The interior block commutes with the reflection i -> p - i (centro-
// symmetric K, symmetric nodes and weights), so each eigenvector is even or
// odd: S[p - i][a] = sig[a] S[i][a].  The Jacobi sweeps keep that to
// rounding; this table has it exactly (each mirror pair replaced by its
// signed mean), so a contraction with S can work on the pair sums and
// differences of its operand, half the multiplies.  asym: the largest
// departure of the Jacobi table S0 from the symmetry, smax its largest entry.
// A degenerate eigenvalue lets the sweeps return a basis without parities
// (p = 1: K is a multiple of the identity and S0 stays the identity); ok()
// is false there and S0 is the table to use.
template <int P>
struct PrecEvenOdd {
    double S[P + 1][P + 1] = {};
    double S0[P + 1][P + 1] = {};
    double Linv[P + 1][P + 1] = {};
    int sig[P + 1] = {};
    double asym = 0.0, smax = 0.0;
    constexpr bool ok() const { return asym <= 1e-10 * smax; }
};

template <int P>
constexpr PrecEvenOdd<P> make_prec_even_odd() {
    const PrecTables<P> t = make_prec<P>();
    constexpr int n = P + 1, N = P;
    PrecEvenOdd<P> r{};
    for (int a = 0; a < n; ++a) {
        double d = 0.0;
        for (int i = 0; i < n; ++i) d += t.S[i][a] * t.S[N - i][a];
        const int sg = d >= 0.0 ? 1 : -1;
        r.sig[a] = sg;
        for (int i = 0; 2 * i <= N; ++i) {
            const double m = 0.5 * (t.S[i][a] + sg * t.S[N - i][a]);
            r.S[i][a] = m;
            r.S[N - i][a] = sg * m;             // the middle row of an odd vector: 0
            const double dev = cabs(t.S[N - i][a] - sg * t.S[i][a]);
            if (dev > r.asym) r.asym = dev;
        }
        for (int i = 0; i < n; ++i) {
            r.S0[i][a] = t.S[i][a];
            if (cabs(t.S[i][a]) > r.smax) r.smax = cabs(t.S[i][a]);
        }
        for (int c = 0; c < n; ++c) r.Linv[a][c] = t.Linv[a][c];
    }
    return r;
}

// One instance per degree, usable inside device code as constants.
template <int P>
struct Basis {
    static constexpr Tables<P> T = make_tables<P>();
};

}  // namespace cx
}  // namespace sipg_internal
