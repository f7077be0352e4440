// CPU test driver for the batched solver (batch.cpp compiled against the
// mock runtime and the CPU stand-ins of its launchers).  Prints one JSON
// line per instance and case for tests/test_cpu_engine.py.
#include "../../include/sipg_hip.h"

#include <hip/hip_runtime.h>   // the mock runtime (its execution model)

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

static void run(const char *name, const std::vector<int> &dim, const std::vector<int> &p, std::vector<double> pen,
                int maxit, bool pcg = false) {
    const int ni = (int)dim.size();
    if (pen.empty())
        for (int k : p) pen.push_back(5.0 * (k + 1) * (k + 1));
    sipg_batch *B = nullptr;
    int rc = sipg_batch_create(&B, ni, dim.data(), dim.data(), p.data(), pen.data(), nullptr);
    if (rc) { printf("{\"case\":\"%s\",\"create_rc\":%d}\n", name, rc); return; }
    int64_t n = 0;
    sipg_batch_size(B, &n);
    void *x, *b, *y;
    sipg_malloc(&x, n * 8);
    sipg_malloc(&b, n * 8);
    sipg_malloc(&y, n * 8);
    std::vector<std::vector<double>> bh(ni);
    for (int i = 0; i < ni; ++i) {
        bh[i] = rhs((int64_t)dim[i] * dim[i] * (p[i] + 1) * (p[i] + 1), 11 + i);
        sipg_batch_upload(B, i, bh[i].data(), b);
        std::vector<double> z(bh[i].size(), 0.0);
        sipg_batch_upload(B, i, z.data(), x);
    }
    // operator: batched apply of b against the oracle's
    double apply_err = 0.0;
    sipg_batch_apply(B, b, y);
    for (int i = 0; i < ni; ++i) {
        std::vector<double> got(bh[i].size()), ref(bh[i].size());
        sipg_batch_download(B, i, y, got.data());
        orc_apply(dim[i], dim[i], p[i], pen[i], bh[i].data(), ref.data());
        for (size_t k = 0; k < ref.size(); ++k) apply_err = std::fmax(apply_err, std::fabs(got[k] - ref[k]));
    }
    sipg_cg_opts o{1e-10, SIPG_TOL_REL, maxit, 0.0, 1, 0};
    std::vector<sipg_cg_info> info(ni);
    rc = pcg ? sipg_batch_pcg(B, x, b, &o, info.data()) : sipg_batch_cg(B, x, b, &o, info.data());
    for (int i = 0; i < ni; ++i) {
        std::vector<double> xh(bh[i].size()), xo(bh[i].size(), 0.0);
        sipg_batch_download(B, i, x, xh.data());
        double rro = 0;
        const int ito = pen[i] > 0 ? orc_cg(dim[i], dim[i], p[i], pen[i], bh[i].data(), xo.data(), 1e-10, 1, maxit, &rro) : -1;
        double dx = 0, nx = 0;
        for (size_t k = 0; k < xh.size(); ++k) { dx += (xh[k] - xo[k]) * (xh[k] - xo[k]); nx += xo[k] * xo[k]; }
        printf("{\"case\":\"%s\",\"inst\":%d,\"rc\":%d,\"status\":%d,\"iters\":%d,\"reps\":%d,\"oracle_iters\":%d,"
               "\"relres\":%.6e,\"diff\":%.6e,\"apply_err\":%.3e}\n",
               name, i, rc, info[i].status, info[i].iters, info[i].replacements, ito,
               pen[i] > 0 ? relres(dim[i], p[i], pen[i], xh, bh[i]) : -1.0, nx > 0 ? std::sqrt(dx / nx) : -1.0,
               apply_err);
    }
    sipg_free(x);
    sipg_free(b);
    sipg_free(y);
    sipg_batch_destroy(B);
}

int main() {
    // asynchronous work runs only when the engine synchronises (the
    // runtime's ordering, with host buffers and graphs read late)
    mock_deferred() = true;
    run("batch", {6, 4, 5, 3, 7, 4, 2}, {3, 2, 4, 5, 2, 13, 16}, {}, 0);
    run("batch_cap", {6, 4, 5}, {3, 2, 4}, {}, 5);
    // maxit = INT_MAX as "no cap": the replay bound must not overflow
    run("batch_maxit_intmax", {6, 4, 5}, {3, 2, 4}, {}, 2147483647);
    run("batch_breakdown", {6, 6, 5}, {3, 3, 4}, {80.0, -500.0, 125.0}, 0);
    run("batch_pcg", {6, 4, 5, 3, 7, 4, 2}, {3, 2, 4, 5, 2, 13, 16}, {}, 0, true);
    run("batch_pcg_other_pen", {6, 4}, {3, 2}, {81.0, 45.0}, 0, true);
    // nothing may be left queued once every call has returned
    printf("{\"case\":\"deferred_queue\",\"pending\":%zu}\n", mock_pending());
    return 0;
}
