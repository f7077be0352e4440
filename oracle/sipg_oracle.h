/*
 * sipg_oracle.h -- CPU restatement of the reference's matrix-free SIPG
 * spectral-element Poisson path.  TEST INFRASTRUCTURE ONLY: this library is
 * the checker for the HIP product path and the CPU baseline timed beside it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  It is never part of the shipped solve path.
 *
 * Every function names the reference file:line it restates.  Index
 * conventions follow the reference exactly:
 *   - element e = dimx*y + x, y = 0 is the TOP row    (mesh_info.hpp:231-235)
 *   - dof (i1,i2,e) lives at ((i1*(p+1)+i2)*noe + e)   (mode_vector.hpp:119)
 *     i1 = x-node, i2 = y-node, i2 = 0 is the element's bottom edge
 *   - D[mode*(p+1)+node] = phi'_mode(z_node)           (device_const.hpp:115-127)
 *
 * Pinning status of each function (what checks it, see tests/):
 *   orc_gll            closed forms p=2..6, numpy Legendre roots, quadrature
 *                      exactness, and the reference's own host code (oracle/_ref)
 *   orc_dphi           exact differentiation of monomials, row sums, reference host code
 *   orc_k1             closed forms p=2,3 (reference test), D^T W D identity
 *   orc_apply          independent dense assembly of the SIPG bilinear form
 *                      (oracle/dense.py), symmetry, constants in kernel
 *   orc_rhs/orc_solve/orc_err_norms
 *                      thesis Tables 4.6/4.8/4.9/4.10 error norms, exact
 *                      reproduction of polynomial solutions, dense LAPACK solve
 *   orc_cg             dense LAPACK solve on small meshes
 */
#ifndef SIPG_ORACLE_H
#define SIPG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Legendre polynomial P_n(z) and P_n'(z) by the three-term recurrence
 * (legendre_poly.hpp:26-83). */
void orc_legendre(int n, double z, double *P, double *dP);

/* Legendre-Gauss-Lobatto nodes (ascending, -1..1) and weights on [-1,1]
 * for degree p, i.e. p+1 points (LGL_quadrature_table.hpp:84-146).
 * Returns 0 on success. */
int orc_gll(int p, double *nodes, double *weights);

/* Lagrange-basis derivative table D[mode*(p+1)+node] = phi'_mode(z_node)
 * (sem_function.hpp:86-110, laid out as device_const.hpp:115-127). */
void orc_dphi(int p, const double *nodes, double *D);

/* 1-D stiffness K1[a*(p+1)+b] = sum_k w_k D_a(k) D_b(k)
 * (laplacian_operator.hpp:31-43). */
void orc_k1(int p, const double *w, const double *D, double *K1);

/* y = A x, the reference's three-kernel operator apply in its order:
 * volume (writes, volume_gpu_kernels.hpp:46-83), then flux_term6a (+=,
 * sipg_sem_2d_gpu_kernels.hpp:189-253), then flux_term6b (+=, 256-316).
 * x and y must not alias.  OpenMP-parallel over elements. */
void orc_apply(int dimx, int dimy, int p, double pen, const double *x, double *y);

/* Same formulas evaluated in single precision (the reference's template
 * instantiated with float). */
void orc_apply_f32(int dimx, int dimy, int p, float pen, const float *x, float *y);

/* Problem definitions for the right-hand side and error norms.
 *   kind 1..5: the reference's analytical solutions A..E
 *              (validation_tests/analytical_solutions.hpp:7-96, with its
 *              truncated pi = 3.14159265359)
 *   kind 0   : seeded Fourier forcing u = sum_k a_k sin(m_k pi x) sin(n_k pi y)
 *              with M_PI, nmodes entries in a/m/n. */
typedef struct {
    int kind;
    int nmodes;
    const double *a;
    const int *m;
    const int *n;
} orc_problem;

double orc_u_ex(const orc_problem *pb, double x, double y);
double orc_f(const orc_problem *pb, double x, double y);
double orc_dx_u_ex(const orc_problem *pb, double x, double y);
double orc_dy_u_ex(const orc_problem *pb, double x, double y);

/* Element lower-left corners, accumulated like build_square_mesh.hpp:250-288
 * (xe[e], ye[e] for e = dim*row + col). */
void orc_mesh_coords(int dim, double *xe, double *ye);

/* Right-hand side incl. weak Dirichlet data g = u_ex on the boundary
 * (sipg_sem_2d.hpp:232-305).  b has noe*(p+1)^2 entries. */
void orc_rhs(int dim, int p, double pen, const orc_problem *pb, double *b);

/* Unpreconditioned CG (conjugate_gradient_gpu.hpp:18-109).  x is the
 * initial guess on entry and the solution on exit.  tol_mode 0 stops on
 * sqrt(r.r) <= tol (the reference's absolute test), 1 on
 * sqrt(r.r) <= tol*||b||.  maxit <= 0 means n (the reference's cap).
 * Returns the iteration count; *rr_out receives the final r.r. */
int orc_cg(int dimx, int dimy, int p, double pen, const double *b, double *x,
           double tol, int tol_mode, int maxit, double *rr_out);

/* Nodal GLL error norms against the exact solution (sipg_sem_2d.hpp:311-370).
 * out[0] = max error, out[1] = L2, out[2] = H1. */
void orc_err_norms(int dim, int p, const double *u, const orc_problem *pb, double *out);

/* print_result (sipg_sem_2d.hpp:375-403) into the file `path`: per element
 * and node "x\ty\tu" with setw(10) at the stream's default precision (the
 * printf %10g), an empty line after each element.  u in the reference
 * order.  Returns 0, or -1 if the file cannot be written. */
int orc_print_result(int dim, int p, const double *u, const char *path);

/* Convenience: rhs + cg from zero + norms.  Returns iterations. */
int orc_solve(int dim, int p, double pen, const orc_problem *pb, double tol,
              int tol_mode, double *u, double *norms_out);

/* Number of OpenMP threads the oracle uses. */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
