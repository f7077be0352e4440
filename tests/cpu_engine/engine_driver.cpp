// CPU test driver for the solver engine (host orchestration compiled
// against the mock runtime).  Runs the C-ABI solvers on small problems and
// prints one JSON line per case for tests/test_cpu_engine.py.
#include "../../include/sipg_hip.h"

#include <cmath>
#include <cstdio>
#include <vector>

extern "C" int orc_cg(int, int, int, double, const double *, double *, double, int, int, double *);
extern "C" void orc_apply(int, int, int, double, const double *, double *);

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
    } else {
        rc = sipg_pcg_mixed(op, x, b, &o, &info);
    }
    sipg_download(op, x, xh.data());
    std::vector<double> xo(n, 0.0);
    double rro = 0;
    const int ito = orc_cg(d, d, p, pen, bh.data(), xo.data(), tol, 1, maxit, &rro);
    double dx = 0, nx = 0;
    for (int64_t i = 0; i < n; ++i) { dx += (xh[i] - xo[i]) * (xh[i] - xo[i]); nx += xo[i] * xo[i]; }
    printf("{\"case\":\"%s\",\"rc\":%d,\"iters\":%d,\"reps\":%d,\"oracle_iters\":%d,"
           "\"relres\":%.6e,\"diff\":%.6e}\n",
           name, rc, info.iters, info.replacements, ito, relres(d, p, pen, xh, bh), std::sqrt(dx / nx));
    sipg_free(x);
    sipg_free(b);
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
    return 0;
}
