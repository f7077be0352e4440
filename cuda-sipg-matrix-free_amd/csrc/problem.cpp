// problem.cpp -- the general problem set-up on the op's square mesh of the
// unit square: the physical coordinates of every GLL node, the load vector
// of -lap u = f with Dirichlet data g (sipg_sem_2d::compute_rhs,
// sipg_sem_2d.hpp:232-305) and the error norms of a solution against any
// exact solution and its gradient (sipg_sem_2d::err_norms,
// sipg_sem_2d.hpp:311-370).
//
// The functions are given either as nodal samples (values at the points
// sipg_quad_points returns, element-major) or as callbacks evaluated here
// on the host at those points, as the reference evaluates its function
// pointers.  The load vector is assembled on the host in double in the
// reference's order of terms (so it matches the reference's host loop term
// for term) and copied to the device in the op's dtype; the norms reduce on
// the device.
#include "sipg_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace sipg_internal;

void sipg_internal::mesh_corners(int dim, std::vector<double> &xe, std::vector<double> &ye) {
    const int64_t noe = (int64_t)dim * dim;
    xe.resize(noe);
    ye.resize(noe);
    const double delta = 1. / (double)dim;
    double yy = 1;
    for (int row = 0; row < dim; ++row) {
        double xx = 0;
        yy -= delta;
        for (int col = 0; col < dim; ++col) {
            xe[(int64_t)row * dim + col] = xx;
            ye[(int64_t)row * dim + col] = yy;
            xx += delta;
        }
    }
}

namespace {

// GLL nodes mapped to one element's side, (z + 1) / (2 dim)
// (sipg_sem_2d.hpp:250-255), and the reference quadrature weights.
struct ElementNodes {
    std::vector<double> z, w, nodes;
    ElementNodes(int p, int dim) : z(p + 1), w(p + 1), nodes(p + 1) {
        gll(p, z.data(), w.data());
        for (int i = 0; i <= p; ++i) nodes[i] = (z[i] * .5 + .5) / dim;
    }
};

bool square_op(const sipg_op *op) { return op && op->dimx == op->dimy && op->dimx > 0; }

// One host array of the op's length in double to a device vector of the
// op's dtype (same element-major order).
int to_device(sipg_op *op, const std::vector<double> &h, void *dev) {
    if (op->dtype == SIPG_F64) {
        SIPG_HIP_TRY(hipMemcpyAsync(dev, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, op->stream));
        SIPG_HIP_TRY(hipStreamSynchronize(op->stream));
    } else {
        std::vector<float> f(h.size());
        for (size_t k = 0; k < h.size(); ++k) f[k] = (float)h[k];
        SIPG_HIP_TRY(hipMemcpyAsync(dev, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, op->stream));
        SIPG_HIP_TRY(hipStreamSynchronize(op->stream));
    }
    return SIPG_OK;
}

// Device index of node (i1, i2) of element e.
inline int64_t at(int64_t e, int i1, int i2, int n1) { return e * n1 * n1 + (int64_t)i1 * n1 + i2; }

}  // namespace

extern "C" {

int sipg_quad_points(sipg_op *op, double *x, double *y) {
    if (!square_op(op) || !x || !y) return SIPG_ERR_INVALID;
    const int dim = op->dimx, p = op->p, n1 = p + 1;
    const ElementNodes N(p, dim);
    std::vector<double> xe, ye;
    mesh_corners(dim, xe, ye);
    const int64_t noe = (int64_t)dim * dim;
    for (int64_t e = 0; e < noe; ++e)
        for (int i1 = 0; i1 < n1; ++i1)
            for (int i2 = 0; i2 < n1; ++i2) {
                x[at(e, i1, i2, n1)] = xe[e] + N.nodes[i1];
                y[at(e, i1, i2, n1)] = ye[e] + N.nodes[i2];
            }
    return SIPG_OK;
}

int sipg_rhs_nodal(sipg_op *op, const double *f, const double *g, void *b) {
    if (!square_op(op) || !f || !b) return SIPG_ERR_INVALID;
    const int dim = op->dimx, N = op->p, n1 = N + 1;
    const int64_t noe = (int64_t)dim * dim;
    const ElementNodes Q(N, dim);
    std::vector<double> w(n1), D((size_t)n1 * n1);
    for (int i = 0; i < n1; ++i) w[i] = Q.w[i] * .5 / dim;
    dphi(N, Q.z.data(), D.data());                 // D[j * n1 + node] = phi_j'(z_node)
    std::vector<double> r(op->n);
    // volume: (w_i1 h/2)(w_i2 h/2) f at each node
    for (int64_t e = 0; e < noe; ++e)
        for (int i1 = 0; i1 < n1; ++i1)
            for (int i2 = 0; i2 < n1; ++i2) r[at(e, i1, i2, n1)] = w[i1] * w[i2] * f[at(e, i1, i2, n1)];
    if (g) {
        // weak Dirichlet data on the four sides of the square, side by side
        // in the reference's order: the penalty term on the side's nodes,
        // then the consistency term -/+ w_i phi_j'(side) g on every line
        // through them (eD bottom row, y = 0; eU top row, y = 1; eL left
        // column; eR right column)
        const double pen = op->pen;
        const double *wq = Q.w.data();
        for (int xx = 0; xx < dim; ++xx) {
            const int64_t eD = (int64_t)(dim - 1) * dim + xx, eU = xx, eL = (int64_t)xx * dim,
                          eR = (int64_t)xx * dim + dim - 1;
            for (int i = 0; i <= N; ++i) {
                const double gD = g[at(eD, i, 0, n1)], gU = g[at(eU, i, N, n1)];
                const double gL = g[at(eL, 0, i, n1)], gR = g[at(eR, N, i, n1)];
                r[at(eD, i, 0, n1)] += pen * wq[i] * gD;
                r[at(eU, i, N, n1)] += pen * wq[i] * gU;
                r[at(eL, 0, i, n1)] += pen * wq[i] * gL;
                r[at(eR, N, i, n1)] += pen * wq[i] * gR;
                for (int j = 0; j <= N; ++j) {
                    r[at(eD, i, j, n1)] -= -1 * wq[i] * D[(size_t)j * n1 + 0] * gD;
                    r[at(eU, i, j, n1)] -= wq[i] * D[(size_t)j * n1 + N] * gU;
                    r[at(eL, j, i, n1)] -= -1 * wq[i] * D[(size_t)j * n1 + 0] * gL;
                    r[at(eR, j, i, n1)] -= wq[i] * D[(size_t)j * n1 + N] * gR;
                }
            }
        }
    }
    return to_device(op, r, b);
}

int sipg_rhs(sipg_op *op, sipg_fn2 f, sipg_fn2 g, void *ctx, void *b) {
    if (!square_op(op) || !f || !b) return SIPG_ERR_INVALID;
    std::vector<double> x(op->n), y(op->n), fv(op->n), gv;
    int rc = sipg_quad_points(op, x.data(), y.data());
    if (rc) return rc;
    for (int64_t k = 0; k < op->n; ++k) fv[k] = f(x[k], y[k], ctx);
    if (g) {
        // g is needed on the boundary nodes of the boundary elements only
        const int dim = op->dimx, N = op->p, n1 = N + 1;
        gv.assign(op->n, 0.0);
        auto put = [&](int64_t e, int i1, int i2) {
            const int64_t k = at(e, i1, i2, n1);
            gv[k] = g(x[k], y[k], ctx);
        };
        for (int xx = 0; xx < dim; ++xx)
            for (int i = 0; i <= N; ++i) {
                put((int64_t)(dim - 1) * dim + xx, i, 0);
                put(xx, i, N);
                put((int64_t)xx * dim, 0, i);
                put((int64_t)xx * dim + dim - 1, N, i);
            }
    }
    return sipg_rhs_nodal(op, fv.data(), g ? gv.data() : nullptr, b);
}

int sipg_err_norms_nodal(sipg_op *op, const void *x, const double *u, const double *ux, const double *uy,
                         double *out) {
    if (!square_op(op) || !x || !u || !ux || !uy || !out) return SIPG_ERR_INVALID;
    const int p = op->p, n1 = p + 1;
    std::vector<double> z(n1), w(n1), D((size_t)n1 * n1);
    gll(p, z.data(), w.data());
    dphi(p, z.data(), D.data());
    const size_t bytes = (size_t)op->n * sizeof(double);
    double *d = nullptr;
    SIPG_HIP_TRY(hipMalloc((void **)&d, 3 * bytes));
    // D in the head of the p.Ap partials (as the Fourier norms); the block
    // partials in the four regions after it
    const parts P = part_layout(op);
    hipError_t e = hipMemcpyAsync(d, u, bytes, hipMemcpyHostToDevice, op->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d + op->n, ux, bytes, hipMemcpyHostToDevice, op->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d + 2 * op->n, uy, bytes, hipMemcpyHostToDevice, op->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(P.pAp, D.data(), D.size() * sizeof(double), hipMemcpyHostToDevice, op->stream);
    int rc = SIPG_OK;
    if (e != hipSuccess) {
        set_last_hip_error(e);
        rc = SIPG_ERR_HIP;
    }
    const int nb = norms_num_blocks(op);
    if (rc == SIPG_OK)
        rc = launch_err_norms_nodal(op->dtype, p, op->dimx, (const double *)op->d_tab64, P.pAp, d, d + op->n,
                                    d + 2 * op->n, x, P.rr, nb, op->stream);
    if (rc == SIPG_OK) rc = norms_finish(op, nb, out);     // synchronises the stream
    else (void)hipStreamSynchronize(op->stream);
    (void)hipFree(d);
    return rc;
}

int sipg_err_norms(sipg_op *op, const void *x, sipg_fn2 u, sipg_fn2 ux, sipg_fn2 uy, void *ctx, double *out) {
    if (!square_op(op) || !x || !u || !ux || !uy || !out) return SIPG_ERR_INVALID;
    std::vector<double> px(op->n), py(op->n), a(op->n), ax(op->n), ay(op->n);
    int rc = sipg_quad_points(op, px.data(), py.data());
    if (rc) return rc;
    for (int64_t k = 0; k < op->n; ++k) {
        a[k] = u(px[k], py[k], ctx);
        ax[k] = ux(px[k], py[k], ctx);
        ay[k] = uy(px[k], py[k], ctx);
    }
    return sipg_err_norms_nodal(op, x, a.data(), ax.data(), ay.data(), out);
}

}  // extern "C"
