/*
 * sipg_oracle.c -- CPU restatement of the reference's SIPG/SEM Poisson path.
 * TEST INFRASTRUCTURE ONLY (checker + CPU baseline); see sipg_oracle.h for
 * the pinning status of every function.  Citations are reference file:line.
 */
#include "sipg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* The reference's analytical solutions use a truncated pi
 * (validation_tests/analytical_solutions.hpp:11,34,56,78). */
#define REF_PI 3.14159265359

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* legendre_poly.hpp:26-50 (value) and the derivative recurrence
 * P'_{k+1} = P'_{k-1} + (2k+1) P_k, valid on the closed interval. */
void orc_legendre(int n, double z, double *P, double *dP)
{
    double p0 = 1.0, p1 = z, d0 = 0.0, d1 = 1.0;
    if (n == 0) { *P = 1.0; *dP = 0.0; return; }
    for (int k = 1; k < n; ++k) {
        double p2 = ((2.0 * k + 1.0) * z * p1 - k * p0) / (k + 1.0);
        double d2 = d0 + (2.0 * k + 1.0) * p1;
        p0 = p1; p1 = p2;
        d0 = d1; d1 = d2;
    }
    *P = p1;
    *dP = d1;
}

/* LGL nodes: the endpoints plus the roots of P_p' (LGL_quadrature_table.hpp:
 * 84-146 finds them by Newton from Lether's guesses).  Here: Newton on P_p'
 * with P_p'' from the Legendre equation, started from the Chebyshev-Gauss-
 * Lobatto points -cos(pi j/p), which interlace the LGL points.  Only the
 * negative half is iterated; the positive half is its mirror image.
 * Weights 2/(p(p+1) P_p(z)^2) (LGL_quadrature_table.hpp:144). */
int orc_gll(int p, double *nodes, double *weights)
{
    if (p < 1) return -1;
    const int N = p;
    nodes[0] = -1.0;
    nodes[N] = 1.0;
    for (int j = 1; 2 * j < N; ++j) {
        double x = -cos(M_PI * j / N);
        for (int it = 0; it < 100; ++it) {
            double P, dP;
            orc_legendre(N, x, &P, &dP);
            double d2P = (2.0 * x * dP - N * (N + 1.0) * P) / (1.0 - x * x);
            double dx = dP / d2P;
            x -= dx;
            if (fabs(dx) <= 1e-17) break;
        }
        nodes[j] = x;
        nodes[N - j] = -x;
    }
    if (N % 2 == 0) nodes[N / 2] = 0.0;
    for (int i = 0; i <= N; ++i) {
        double P, dP;
        orc_legendre(N, nodes[i], &P, &dP);
        weights[i] = 2.0 / (N * (N + 1.0) * P * P);
    }
    return 0;
}

/* sem_function.hpp:86-110: phi'_mode(z_node) =
 *   -p(p+1)/4 at mode = node = 0, +p(p+1)/4 at mode = node = p, 0 on the
 *   rest of the diagonal, P_p(z_node) / (P_p(z_mode)(z_node - z_mode)) off it.
 * Stored D[mode*(p+1)+node] (device_const.hpp:123-125). */
void orc_dphi(int p, const double *z, double *D)
{
    const int n1 = p + 1;
    double *Pz = (double *)malloc(sizeof(double) * n1);
    for (int i = 0; i < n1; ++i) {
        double dP;
        orc_legendre(p, z[i], &Pz[i], &dP);
    }
    for (int mode = 0; mode < n1; ++mode)
        for (int node = 0; node < n1; ++node) {
            double v;
            if (mode == node) {
                if (node == 0) v = -0.25 * (p + 1) * p;
                else if (node == p) v = 0.25 * (p + 1) * p;
                else v = 0.0;
            } else {
                v = Pz[node] / (Pz[mode] * (z[node] - z[mode]));
            }
            D[mode * n1 + node] = v;
        }
    free(Pz);
}

/* laplacian_operator.hpp:31-43 (laplace_1d): K1(a,b) = sum_k w_k D_a(k) D_b(k). */
void orc_k1(int p, const double *w, const double *D, double *K1)
{
    const int n1 = p + 1;
    for (int a = 0; a < n1; ++a)
        for (int b = 0; b < n1; ++b) {
            double r = 0.0;
            for (int k = 0; k < n1; ++k) r += w[k] * D[a * n1 + k] * D[b * n1 + k];
            K1[a * n1 + b] = r;
        }
}

/* ------------------------------------------------------------------------ */
/* Operator apply, generic over the float type through a macro body so that
 * the double and float instantiations are the same text (the reference's
 * template<typename FLOAT_TYPE>).                                          */

#define DEFINE_APPLY(NAME, T)                                                          \
void NAME(int dimx, int dimy, int p, T pen, const T *x, T *y)                          \
{                                                                                      \
    const int N = p, n1 = p + 1;                                                       \
    const long noe = (long)dimx * dimy;                                                \
    double zd[64], wd[64];                                                             \
    double *Dd = (double *)malloc(sizeof(double) * n1 * n1);                           \
    double *Kd = (double *)malloc(sizeof(double) * n1 * n1);                           \
    orc_gll(p, zd, wd);                                                                \
    orc_dphi(p, zd, Dd);                                                               \
    orc_k1(p, wd, Dd, Kd);                                                             \
    T w[64];                                                                           \
    T *D = (T *)malloc(sizeof(T) * n1 * n1);                                           \
    T *K1 = (T *)malloc(sizeof(T) * n1 * n1);                                          \
    for (int i = 0; i < n1; ++i) w[i] = (T)wd[i];                                      \
    for (int i = 0; i < n1 * n1; ++i) { D[i] = (T)Dd[i]; K1[i] = (T)Kd[i]; }           \
    /* in(i1,i2,e) = x[(i1*n1+i2)*noe+e]  (mode_vector.hpp:119) */                     \
    _Pragma("omp parallel for schedule(static)")                                       \
    for (long e = 0; e < noe; ++e) {                                                   \
        /* volume_gpu_kernels.hpp:46-83: Y(i1,i2) = sum_i w_i2 K1(i,i1) U(i,i2)         \
         *                                          + w_i1 K1(i,i2) U(i1,i)   */        \
        for (int i1 = 0; i1 < n1; ++i1)                                                \
            for (int i2 = 0; i2 < n1; ++i2) {                                          \
                T result = 0;                                                          \
                for (int i = 0; i < n1; ++i)                                           \
                    result += w[i2] * K1[i * n1 + i1] * x[(i * n1 + i2) * noe + e]     \
                            + w[i1] * K1[i * n1 + i2] * x[(i1 * n1 + i) * noe + e];    \
                y[(i1 * n1 + i2) * noe + e] = result;                                  \
            }                                                                          \
    }                                                                                  \
    /* flux_term6a: x-faces (sipg_sem_2d_gpu_kernels.hpp:189-253) */                   \
    _Pragma("omp parallel for schedule(static)")                                       \
    for (long e = 0; e < noe; ++e) {                                                   \
        const int xx = (int)(e % dimx), yy = (int)(e / dimx);                          \
        const long L = (xx == 0) ? -4 : e - 1;          /* mesh_info.hpp:305-309 */    \
        const long R = (xx == dimx - 1) ? -2 : e + 1;   /* mesh_info.hpp:287-291 */    \
        (void)yy;                                                                      \
        for (int j = 0; j < n1; ++j) {                                                 \
            T a = 0, b = 0;                                                            \
            const T ev = w[j] * (L < 0 ? x[(0 * n1 + j) * noe + e]                     \
                                       : (T)0.5 * (x[(0 * n1 + j) * noe + e]           \
                                                 - x[(N * n1 + j) * noe + L]));        \
            const T fv = w[j] * (R < 0 ? -x[(N * n1 + j) * noe + e]                    \
                                       : (T)0.5 * (x[(0 * n1 + j) * noe + R]           \
                                                 - x[(N * n1 + j) * noe + e]));        \
            for (int i = 0; i < n1; ++i) {                                             \
                const T d0 = D[i * n1 + 0], dN = D[i * n1 + N];                        \
                y[(i * n1 + j) * noe + e] += d0 * ev + dN * fv;                        \
                a += L < 0 ? d0 * x[(i * n1 + j) * noe + e]                            \
                           : (T)0.5 * (d0 * x[(i * n1 + j) * noe + e]                  \
                                     + dN * x[(i * n1 + j) * noe + L]);                \
                b += R < 0 ? -dN * x[(i * n1 + j) * noe + e]                           \
                           : (T)-0.5 * (dN * x[(i * n1 + j) * noe + e]                 \
                                      + d0 * x[(i * n1 + j) * noe + R]);               \
            }                                                                          \
            y[(0 * n1 + j) * noe + e] += w[j] * a + (L < 0 ? pen : 2 * pen) * ev;      \
            y[(N * n1 + j) * noe + e] += w[j] * b - (R < 0 ? pen : 2 * pen) * fv;      \
        }                                                                              \
    }                                                                                  \
    /* flux_term6b: y-faces (sipg_sem_2d_gpu_kernels.hpp:256-316) */                   \
    _Pragma("omp parallel for schedule(static)")                                       \
    for (long e = 0; e < noe; ++e) {                                                   \
        const int yy = (int)(e / dimx);                                                \
        const long Dn = (yy == dimy - 1) ? -1 : e + dimx; /* mesh_info.hpp:278-282 */  \
        const long Up = (yy == 0) ? -3 : e - dimx;        /* mesh_info.hpp:296-300 */  \
        for (int j = 0; j < n1; ++j) {                                                 \
            T c = 0, d = 0;                                                            \
            const T g = w[j] * (Dn < 0 ? x[(j * n1 + 0) * noe + e]                     \
                                       : (T)0.5 * (x[(j * n1 + 0) * noe + e]           \
                                                 - x[(j * n1 + N) * noe + Dn]));       \
            const T h = w[j] * (Up < 0 ? -x[(j * n1 + N) * noe + e]                    \
                                       : (T)0.5 * (x[(j * n1 + 0) * noe + Up]          \
                                                 - x[(j * n1 + N) * noe + e]));        \
            for (int i = 0; i < n1; ++i) {                                             \
                const T d0 = D[i * n1 + 0], dN = D[i * n1 + N];                        \
                y[(j * n1 + i) * noe + e] += d0 * g + dN * h;                          \
                c += Dn < 0 ? d0 * x[(j * n1 + i) * noe + e]                           \
                            : (T)0.5 * (d0 * x[(j * n1 + i) * noe + e]                 \
                                      + dN * x[(j * n1 + i) * noe + Dn]);              \
                d += Up < 0 ? -dN * x[(j * n1 + i) * noe + e]                          \
                            : (T)-0.5 * (dN * x[(j * n1 + i) * noe + e]                \
                                       + d0 * x[(j * n1 + i) * noe + Up]);             \
            }                                                                          \
            y[(j * n1 + 0) * noe + e] += w[j] * c + (Dn < 0 ? pen : 2 * pen) * g;      \
            y[(j * n1 + N) * noe + e] += w[j] * d - (Up < 0 ? pen : 2 * pen) * h;      \
        }                                                                              \
    }                                                                                  \
    free(Dd); free(Kd); free(D); free(K1);                                             \
}

DEFINE_APPLY(orc_apply, double)
DEFINE_APPLY(orc_apply_f32, float)

/* ------------------------------------------------------------------------ */
/* Problems (validation_tests/analytical_solutions.hpp:7-96) and the seeded
 * Fourier forcing (u = sum a_k sin(m pi x) sin(n pi y), -lap u = f).        */

double orc_u_ex(const orc_problem *pb, double x, double y)
{
    const double PI = REF_PI;
    switch (pb->kind) {
    case 1: return x * y * (y - 1) * (x - 1) * exp(x - y);
    case 2: return x * (1 - x) * exp(3 * x) * sin(2 * PI * y);
    case 3: return y * (1 - y) * x * x * x;
    case 4: return x * x * y * y;
    case 5: return cos(8 * PI * x) + cos(8 * PI * y);
    default: {
        double s = 0;
        for (int k = 0; k < pb->nmodes; ++k)
            s += pb->a[k] * sin(pb->m[k] * M_PI * x) * sin(pb->n[k] * M_PI * y);
        return s;
    }
    }
}

double orc_f(const orc_problem *pb, double x, double y)
{
    const double PI = REF_PI;
    switch (pb->kind) {
    case 1: return -2 * x * (y - 1) * (y - 2 * x + x * y + 2) * exp(x - y);
    case 2: return -sin(2 * PI * y) * exp(3 * x) * ((4 - 3 * x - 9 * x * x) - 4 * PI * PI * (x - x * x));
    case 3: return 2 * x * x * x - 6 * x * y * (1 - y);
    case 4: return -2 * y * y - 2 * x * x;
    case 5: return 64 * PI * PI * cos(8 * PI * x) + 64 * PI * PI * cos(8 * PI * y);
    default: {
        double s = 0;
        for (int k = 0; k < pb->nmodes; ++k) {
            const double mm = pb->m[k], nn = pb->n[k];
            s += pb->a[k] * M_PI * M_PI * (mm * mm + nn * nn)
                 * sin(mm * M_PI * x) * sin(nn * M_PI * y);
        }
        return s;
    }
    }
}

double orc_dx_u_ex(const orc_problem *pb, double x, double y)
{
    const double PI = REF_PI;
    switch (pb->kind) {
    case 1: return (x * x + x - 1) * (y - 1) * y * exp(x - y);
    case 2: return (3 * (1 - x) * x - x + (1 - x)) * exp(3 * x) * sin(2 * PI * y);
    case 3: return y * (1 - y) * x * x * 3;
    case 4: return 2 * x * y * y;
    case 5: return -8 * PI * sin(8 * PI * x);
    default: {
        double s = 0;
        for (int k = 0; k < pb->nmodes; ++k)
            s += pb->a[k] * pb->m[k] * M_PI * cos(pb->m[k] * M_PI * x) * sin(pb->n[k] * M_PI * y);
        return s;
    }
    }
}

double orc_dy_u_ex(const orc_problem *pb, double x, double y)
{
    const double PI = REF_PI;
    switch (pb->kind) {
    case 1: return -1 * (y * y - 3 * y + 1) * (x - 1) * x * exp(x - y);
    case 2: return 2 * PI * (1 - x) * x * exp(3 * x) * cos(2 * PI * y);
    case 3: return (1 - 2 * y) * x * x * x;
    case 4: return 2 * y * x * x;
    case 5: return -8 * PI * sin(8 * PI * y);
    default: {
        double s = 0;
        for (int k = 0; k < pb->nmodes; ++k)
            s += pb->a[k] * pb->n[k] * M_PI * sin(pb->m[k] * M_PI * x) * cos(pb->n[k] * M_PI * y);
        return s;
    }
    }
}

/* build_square_mesh.hpp:250-288: corner 0 of element i*dim+j, accumulated. */
void orc_mesh_coords(int dim, double *xe, double *ye)
{
    const double delta = 1. / (double)dim;
    double xx = 0, yy = 1;
    for (int i = 0; i < dim; ++i) {
        xx = 0;
        yy -= delta;
        for (int j = 0; j < dim; ++j) {
            xe[i * dim + j] = xx;
            ye[i * dim + j] = yy;
            xx += delta;
        }
    }
}

/* sipg_sem_2d.hpp:232-305 */
void orc_rhs(int dim, int p, double pen, const orc_problem *pb, double *b)
{
    const int N = p, n1 = p + 1;
    const long noe = (long)dim * dim;
    double z[64], wq[64], nodes[64], weights[64], w[64];
    double *D = (double *)malloc(sizeof(double) * n1 * n1);
    double *xe = (double *)malloc(sizeof(double) * noe);
    double *ye = (double *)malloc(sizeof(double) * noe);
    orc_gll(p, z, wq);
    orc_dphi(p, z, D);
    orc_mesh_coords(dim, xe, ye);
    for (int i = 0; i <= N; ++i) {
        nodes[i] = (z[i] * .5 + .5) / dim;
        weights[i] = wq[i];
        w[i] = wq[i] * .5 / dim;
    }
#define RHS(i1, i2, e) b[((long)(i1) * n1 + (i2)) * noe + (e)]
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < noe; ++e)
        for (int i1 = 0; i1 <= N; ++i1)
            for (int i2 = 0; i2 <= N; ++i2)
                RHS(i1, i2, e) = w[i1] * w[i2] * orc_f(pb, xe[e] + nodes[i1], ye[e] + nodes[i2]);
    /* d_phi(j, node) = D[j*n1+node] */
    for (int xx = 0; xx < dim; ++xx) {
        const long eD = (long)(dim - 1) * dim + xx, eU = xx, eL = (long)xx * dim, eR = (long)xx * dim + dim - 1;
        for (int i = 0; i <= N; ++i) {
            RHS(i, 0, eD) += pen * weights[i] * orc_u_ex(pb, xe[eD] + nodes[i], ye[eD] + nodes[0]);
            RHS(i, N, eU) += pen * weights[i] * orc_u_ex(pb, xe[eU] + nodes[i], ye[eU] + nodes[N]);
            RHS(0, i, eL) += pen * weights[i] * orc_u_ex(pb, xe[eL] + nodes[0], ye[eL] + nodes[i]);
            RHS(N, i, eR) += pen * weights[i] * orc_u_ex(pb, xe[eR] + nodes[N], ye[eR] + nodes[i]);
            for (int j = 0; j <= N; ++j) {
                RHS(i, j, eD) -= -1 * weights[i] * D[j * n1 + 0] * orc_u_ex(pb, xe[eD] + nodes[i], ye[eD] + nodes[0]);
                RHS(i, j, eU) -= weights[i] * D[j * n1 + N] * orc_u_ex(pb, xe[eU] + nodes[i], ye[eU] + nodes[N]);
                RHS(j, i, eL) -= -1 * weights[i] * D[j * n1 + 0] * orc_u_ex(pb, xe[eL] + nodes[0], ye[eL] + nodes[i]);
                RHS(j, i, eR) -= weights[i] * D[j * n1 + N] * orc_u_ex(pb, xe[eR] + nodes[N], ye[eR] + nodes[i]);
            }
        }
    }
#undef RHS
    free(D); free(xe); free(ye);
}

static double dot(long n, const double *a, const double *b)
{
    double s = 0;
    #pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* conjugate_gradient_gpu.hpp:18-109 with conjugate_gradient_kernels.hpp:5-38 */
int orc_cg(int dimx, int dimy, int p, double pen, const double *b, double *x,
           double tol, int tol_mode, int maxit, double *rr_out)
{
    const long n = (long)dimx * dimy * (p + 1) * (p + 1);
    double *r = (double *)malloc(sizeof(double) * n);
    double *pv = (double *)malloc(sizeof(double) * n);
    double *Ap = (double *)malloc(sizeof(double) * n);
    const double norm_b = sqrt(dot(n, b, b));
    orc_apply(dimx, dimy, p, pen, x, r);
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) { r[i] = b[i] + (-1.0) * r[i]; pv[i] = r[i]; }
    double rr_old = dot(n, r, r);
    const double stop = tol_mode == 1 ? tol * norm_b : tol;
    const long max_it = maxit > 0 ? maxit : n;
    long it = 0;
    while (sqrt(rr_old) > stop && it < max_it) {
        orc_apply(dimx, dimy, p, pen, pv, Ap);
        const double pAp = dot(n, pv, Ap);
        const double alpha = rr_old / pAp;
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            x[i] = x[i] + alpha * pv[i];
            r[i] = r[i] - alpha * Ap[i];
        }
        const double rr_new = dot(n, r, r);
        const double beta = rr_new / rr_old;
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) pv[i] = r[i] + beta * pv[i];
        rr_old = rr_new;
        ++it;
    }
    if (rr_out) *rr_out = rr_old;
    free(r); free(pv); free(Ap);
    return (int)it;
}

/* sipg_sem_2d.hpp:311-370.  phi_j(z_q) = delta_jq, so the reference's
 * double sums over (i,j) reduce to single sums without changing a bit. */
void orc_err_norms(int dim, int p, const double *u, const orc_problem *pb, double *out)
{
    const int n1 = p + 1;
    const long noe = (long)dim * dim;
    double z[64], wq[64], nodes[64], weights[64];
    double *D = (double *)malloc(sizeof(double) * n1 * n1);
    double *xe = (double *)malloc(sizeof(double) * noe);
    double *ye = (double *)malloc(sizeof(double) * noe);
    orc_gll(p, z, wq);
    orc_dphi(p, z, D);
    orc_mesh_coords(dim, xe, ye);
    for (int i = 0; i <= p; ++i) {
        nodes[i] = (z[i] * .5 + .5) / dim;
        weights[i] = wq[i] * .5 / dim;
    }
    double max_err = -1, L2 = 0, dxs = 0, dys = 0;
    for (long e = 0; e < noe; ++e)
        for (int pp = 0; pp <= p; ++pp)
            for (int q = 0; q <= p; ++q) {
                double dx_u = 0, dy_u = 0;
                const double w2 = weights[pp] * weights[q];
                for (int i = 0; i <= p; ++i) {
                    dx_u += u[((long)i * n1 + q) * noe + e] * 2 * dim * D[i * n1 + pp];
                    dy_u += u[((long)pp * n1 + i) * noe + e] * 2 * dim * D[i * n1 + q];
                }
                const double X = xe[e] + nodes[pp], Y = ye[e] + nodes[q];
                const double err = fabs(u[((long)pp * n1 + q) * noe + e] - orc_u_ex(pb, X, Y));
                if (err > max_err) max_err = err;
                L2 += w2 * err * err;
                const double ex = dx_u - orc_dx_u_ex(pb, X, Y), ey = dy_u - orc_dy_u_ex(pb, X, Y);
                dxs += w2 * ex * ex;
                dys += w2 * ey * ey;
            }
    out[0] = max_err;
    out[1] = sqrt(L2);
    out[2] = sqrt(L2 + dxs + dys);
    free(D); free(xe); free(ye);
}

/* sipg_sem_2d.hpp:375-403 */
int orc_print_result(int dim, int p, const double *u, const char *path)
{
    const int n1 = p + 1;
    const long noe = (long)dim * dim;
    double z[64], wq[64], nodes[64];
    double *xe = (double *)malloc(sizeof(double) * noe);
    double *ye = (double *)malloc(sizeof(double) * noe);
    FILE *f = fopen(path, "w");
    if (!f) { free(xe); free(ye); return -1; }
    orc_gll(p, z, wq);
    orc_mesh_coords(dim, xe, ye);
    for (int i = 0; i <= p; ++i) nodes[i] = (z[i] * .5 + .5) / dim;
    for (long e = 0; e < noe; ++e) {
        for (int i = 0; i <= p; ++i)
            for (int j = 0; j <= p; ++j)
                fprintf(f, "%10g\t%10g\t%10g\n", xe[e] + nodes[i], ye[e] + nodes[j], u[((long)i * n1 + j) * noe + e]);
        fprintf(f, "\n");
    }
    free(xe); free(ye);
    return fclose(f) == 0 ? 0 : -1;
}

int orc_solve(int dim, int p, double pen, const orc_problem *pb, double tol,
              int tol_mode, double *u, double *norms_out)
{
    const long n = (long)dim * dim * (p + 1) * (p + 1);
    double *b = (double *)malloc(sizeof(double) * n);
    orc_rhs(dim, p, pen, pb, b);
    memset(u, 0, sizeof(double) * n);
    double rr;
    int it = orc_cg(dim, dim, p, pen, b, u, tol, tol_mode, 0, &rr);
    if (norms_out) orc_err_norms(dim, p, u, pb, norms_out);
    free(b);
    return it;
}
