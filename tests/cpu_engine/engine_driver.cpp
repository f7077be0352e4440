// CPU test driver for the solver engine (host orchestration compiled
// against the mock runtime).  Runs the C-ABI solvers on small problems and
// prints one JSON line per case for tests/test_cpu_engine.py.
#include "../../include/sipg_hip.h"

#include <cmath>
#include <cstdio>
#include <vector>

extern "C" {
#include "../../oracle/sipg_oracle.h"
}

namespace sipg_internal {
int &mock_peeks();       // mock_kernels.cpp: solves finished by a peek
int &mock_p_updates();   // mock_kernels.cpp: iterations enqueued
}
using sipg_internal::mock_p_updates;
using sipg_internal::mock_peeks;

namespace sipg_internal {
void coarse_dst(int dim, std::vector<double> &Q, std::vector<double> &lamK, std::vector<double> &lamM, int p);
}

// The coarse operator from first principles against the DST tables:
// A0 = P0 A P0^T with A the oracle's operator and P0^T the bilinear
// interpolation of interior vertex values at the GLL nodes (reference order,
// element e = row * dimx + col, row 0 at the top, i1 the x node, i2 the y
// node), compared with Qy diag(lamMy lamKx + lamKy lamMx) Qx applied to unit
// vectors.  Prints the largest relative difference.
static void coarse_check(int dx, int dy, int p) {
    const int n1 = p + 1, nvx = dx - 1, nvy = dy - 1, nv = nvx * nvy;
    const int64_t noe = (int64_t)dx * dy, n = noe * n1 * n1;
    std::vector<double> z(n1), w(n1);
    orc_gll(p, z.data(), w.data());
    std::vector<double> xi(n1);
    for (int i = 0; i < n1; ++i) xi[i] = 0.5 * (z[i] + 1.0);
    auto interp = [&](const std::vector<double> &V, std::vector<double> &u) {   // V on (dy+1) x (dx+1)
        u.assign(n, 0.0);
        for (int row = 0; row < dy; ++row)
            for (int col = 0; col < dx; ++col)
                for (int i1 = 0; i1 < n1; ++i1)
                    for (int i2 = 0; i2 < n1; ++i2) {
                        const double X = xi[i1], Y = xi[i2];   // from the left, from the bottom
                        const double tl = V[(size_t)row * (dx + 1) + col], tr = V[(size_t)row * (dx + 1) + col + 1];
                        const double bl = V[(size_t)(row + 1) * (dx + 1) + col],
                                     br = V[(size_t)(row + 1) * (dx + 1) + col + 1];
                        u[((int64_t)i1 * n1 + i2) * noe + (int64_t)row * dx + col] =
                            (1 - X) * (1 - Y) * bl + X * (1 - Y) * br + (1 - X) * Y * tl + X * Y * tr;
                    }
    };
    std::vector<double> A0((size_t)nv * nv), V, u, Au(n), u2;
    for (int j = 0; j < nv; ++j) {
        V.assign((size_t)(dx + 1) * (dy + 1), 0.0);
        V[(size_t)(j / nvx + 1) * (dx + 1) + j % nvx + 1] = 1.0;
        interp(V, u);
        orc_apply(dx, dy, p, 1e3, u.data(), Au.data());    // the penalty drops out on V0
        for (int i = 0; i < nv; ++i) {
            V.assign((size_t)(dx + 1) * (dy + 1), 0.0);
            V[(size_t)(i / nvx + 1) * (dx + 1) + i % nvx + 1] = 1.0;
            interp(V, u2);
            double s = 0.0;
            for (int64_t k = 0; k < n; ++k) s += u2[k] * Au[k];
            A0[(size_t)i * nv + j] = s;
        }
    }
    std::vector<double> Qx, Qy, kx, mx, ky, my;
    sipg_internal::coarse_dst(dx, Qx, kx, mx, p);
    sipg_internal::coarse_dst(dy, Qy, ky, my, p);
    double worst = 0.0, scale = 0.0;
    for (double v : A0) scale = std::fmax(scale, std::fabs(v));
    for (int i = 0; i < nv; ++i)
        for (int j = 0; j < nv; ++j) {
            const int ai = i / nvx, bi = i % nvx, aj = j / nvx, bj = j % nvx;
            double s = 0.0;
            for (int a = 0; a < nvy; ++a)
                for (int b = 0; b < nvx; ++b)
                    s += Qy[(size_t)ai * nvy + a] * Qx[(size_t)bi * nvx + b] * (my[a] * kx[b] + ky[a] * mx[b]) *
                         Qy[(size_t)a * nvy + aj] * Qx[(size_t)b * nvx + bj];
            worst = std::fmax(worst, std::fabs(s - A0[(size_t)i * nv + j]));
        }
    printf("{\"case\":\"coarse_tables_%dx%d_p%d\",\"rel\":%.3e,\"center\":%.6f}\n", dx, dy, p, worst / scale,
           A0[(size_t)(nv / 2) * nv + nv / 2]);
}

static std::vector<double> rhs(int64_t n, unsigned seed) {
    std::vector<double> b(n);
    unsigned long long s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    for (auto &v : b) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        v = ((s >> 11) * 0x1.0p-53) * 2.0 - 1.0;
    }
    return b;
}

static double relres(int d, int p, double pen, const std::vector<double> &x, const std::vector<double> &b) {
    std::vector<double> Ax(b.size());
    orc_apply(d, d, p, pen, x.data(), Ax.data());
    double rr = 0, bb = 0;
    for (size_t i = 0; i < b.size(); ++i) { rr += (b[i] - Ax[i]) * (b[i] - Ax[i]); bb += b[i] * b[i]; }
    return std::sqrt(rr / bb);
}

static void run(const char *name, int d, int p, double pen, int mode, double tol, int maxit, double rf,
                int restart = 1, int true_stop = 1) {
    sipg_op *op = nullptr;
    int rc = sipg_op_create(&op, d, d, p, pen, SIPG_F64, nullptr);
    if (rc) { printf("{\"case\":\"%s\",\"create_rc\":%d}\n", name, rc); return; }
    int64_t n;
    sipg_op_size(op, &n);
    std::vector<double> bh = rhs(n, 7), xh(n, 0.0);
    void *x, *b;
    sipg_malloc(&x, n * 8);
    sipg_malloc(&b, n * 8);
    sipg_upload(op, bh.data(), b);  // reference order -> element-major device order
    sipg_upload(op, xh.data(), x);
    sipg_cg_opts o{tol, SIPG_TOL_REL, maxit, rf, true_stop, restart};
    sipg_cg_info info{};
    mock_peeks() = 0;
    mock_p_updates() = 0;
    if (mode == 0) {
        int it = 0;
        double res = 0;
        rc = sipg_cg(op, x, b, tol, SIPG_TOL_REL, maxit, &it, &res);
        info.iters = it;
        info.res_norm = res;
    } else if (mode == 1) {
        rc = sipg_cg_ex(op, x, b, &o, &info);
    } else if (mode == 2) {
        rc = sipg_cg_mixed(op, x, b, &o, &info);
    } else if (mode == 3) {
        rc = sipg_pcg(op, x, b, &o, &info);
    } else if (mode == 5) {
        rc = sipg_pcg2(op, x, b, &o, &info);
    } else {
        rc = sipg_pcg_mixed(op, x, b, &o, &info);
    }
    const int launched = mock_p_updates();
    sipg_download(op, x, xh.data());
    std::vector<double> xo(n, 0.0);
    double rro = 0;
    const int ito = orc_cg(d, d, p, pen, bh.data(), xo.data(), tol, 1, maxit, &rro);
    double dx = 0, nx = 0;
    for (int64_t i = 0; i < n; ++i) { dx += (xh[i] - xo[i]) * (xh[i] - xo[i]); nx += xo[i] * xo[i]; }
    printf("{\"case\":\"%s\",\"rc\":%d,\"iters\":%d,\"reps\":%d,\"oracle_iters\":%d,"
           "\"relres\":%.6e,\"diff\":%.6e,\"peeks\":%d,\"true_res\":%.6e,\"b_norm\":%.6e,\"launched\":%d}\n",
           name, rc, info.iters, info.replacements, ito, relres(d, p, pen, xh, bh), std::sqrt(dx / nx), mock_peeks(),
           info.true_res_norm, info.b_norm, launched);
    sipg_free(x);
    sipg_free(b);
    sipg_op_destroy(op);
}

// The general set-up's host path against the oracle: sipg_rhs_nodal fed
// with orc_f / orc_u_ex at the library's node points vs orc_rhs, and
// sipg_err_norms_nodal vs orc_err_norms on a perturbed interpolant.
static void problem_check(int kind, int dim, int p) {
    const double pen = 100.0 * p * p;
    sipg_op *op = nullptr;
    if (sipg_op_create(&op, dim, dim, p, pen, SIPG_F64, nullptr)) return;
    int64_t n;
    sipg_op_size(op, &n);
    orc_problem pb{kind, 0, nullptr, nullptr, nullptr};
    std::vector<double> X(n), Y(n), f(n), g(n), u(n), ux(n), uy(n);
    int rc = sipg_quad_points(op, X.data(), Y.data());
    for (int64_t k = 0; k < n; ++k) {
        f[k] = orc_f(&pb, X[k], Y[k]);
        g[k] = u[k] = orc_u_ex(&pb, X[k], Y[k]);
        ux[k] = orc_dx_u_ex(&pb, X[k], Y[k]);
        uy[k] = orc_dy_u_ex(&pb, X[k], Y[k]);
    }
    void *b, *x;
    sipg_malloc(&b, n * 8);
    sipg_malloc(&x, n * 8);
    if (!rc) rc = sipg_rhs_nodal(op, f.data(), g.data(), b);
    std::vector<double> bh(n), bo(n);
    sipg_download(op, b, bh.data());
    orc_rhs(dim, p, pen, &pb, bo.data());
    double db = 0, sb = 0;
    for (int64_t k = 0; k < n; ++k) { db = std::fmax(db, std::fabs(bh[k] - bo[k])); sb = std::fmax(sb, std::fabs(bo[k])); }
    // x = the interpolant of u_ex plus a smooth perturbation (device order)
    std::vector<double> xd(n), xr(n);
    for (int64_t k = 0; k < n; ++k) xd[k] = u[k] + 1e-3 * std::sin(3.0 * X[k] + 2.0 * Y[k]);
    sipg_memcpy_h2d(x, xd.data(), n * 8);
    double ours[3] = {0, 0, 0}, want[3];
    if (!rc) rc = sipg_err_norms_nodal(op, x, u.data(), ux.data(), uy.data(), ours);
    sipg_download(op, x, xr.data());
    orc_err_norms(dim, p, xr.data(), &pb, want);
    double dn = 0;
    for (int k = 0; k < 3; ++k) dn = std::fmax(dn, std::fabs(ours[k] - want[k]) / want[k]);
    printf("{\"case\":\"problem_%d_%d_p%d\",\"rc\":%d,\"rhs_rel\":%.3e,\"norms_rel\":%.3e}\n", kind, dim, p,
           rc, db / sb, dn);
    sipg_free(b);
    sipg_free(x);
    sipg_op_destroy(op);
}

int main() {
    run("plain", 6, 3, 80.0, 0, 1e-10, 0, 0.0);
    run("plain_cap", 6, 3, 80.0, 0, 1e-12, 5, 0.0);
    run("ex_replace", 8, 4, 125.0, 1, 1e-10, 0, 0.1, 0);
    run("ex_noreplace", 8, 4, 125.0, 1, 1e-10, 0, 0.0, 0);
    run("mixed", 8, 4, 125.0, 2, 1e-10, 0, 0.1);
    run("mixed_cap", 8, 4, 125.0, 2, 1e-10, 30, 0.1);
    run("mixed_keep_p", 8, 4, 125.0, 2, 1e-10, 0, 0.1, 0);
    run("pcg_replace", 8, 4, 125.0, 3, 1e-10, 0, 0.1, 0);
    run("pcg_true_stop", 8, 4, 125.0, 3, 1e-10, 0, 0.0, 0);
    run("pcg_plain", 8, 4, 125.0, 3, 1e-10, 0, 0.0, 0, 0);
    run("pcg_other_pen", 8, 4, 100.0, 3, 1e-10, 0, 0.1, 0);
    run("pcg_cap", 8, 4, 125.0, 3, 1e-12, 5, 0.0, 0, 0);
    run("pcg_mixed", 8, 4, 125.0, 4, 1e-10, 0, 0.1);
    run("pcg2_replace", 8, 4, 125.0, 5, 1e-10, 0, 0.1, 0);
    run("pcg2_true_stop", 12, 3, 80.0, 5, 1e-10, 0, 0.0, 0);
    run("pcg2_plain", 12, 3, 80.0, 5, 1e-10, 0, 0.0, 0, 0);
    run("pcg2_vs_pcg", 12, 3, 80.0, 3, 1e-10, 0, 0.0, 0, 0);
    run("pcg2_one_element", 1, 3, 80.0, 5, 1e-10, 0, 0.0, 0, 0);
    // a tolerance below the FP64 floor: the solvers end at the floor
    run("ex_floor", 8, 4, 125.0, 1, 1e-17, 0, 0.1, 0);
    run("pcg_floor", 8, 4, 125.0, 3, 1e-17, 0, 0.1, 0);
    run("pcg2_floor", 12, 3, 80.0, 5, 1e-17, 0, 0.1, 0);
    // tolerances a little above the floor (about 1.26e-14): between the
    // finishing checks, peeks at the true residual end the solve
    run("ex_near_floor", 8, 4, 125.0, 1, 2e-14, 0, 0.1, 0);
    run("ex_nearer_floor", 8, 4, 125.0, 1, 1.3e-14, 0, 0.1, 0);
    run("pcg_near_floor", 8, 4, 125.0, 3, 2e-14, 0, 0.1, 0);
    run("pcg2_near_floor", 12, 3, 80.0, 5, 2e-14, 0, 0.1, 0);
    coarse_check(5, 5, 3);
    coarse_check(6, 4, 2);
    coarse_check(4, 5, 1);
    problem_check(1, 4, 2);
    problem_check(2, 5, 3);
    problem_check(3, 3, 4);
    problem_check(4, 6, 2);
    problem_check(5, 8, 3);
    problem_check(5, 1, 5);
    return 0;
}
