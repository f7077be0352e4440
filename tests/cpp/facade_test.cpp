// Test program for the C++ facade (include/sipg_hip.hpp), written as code
// against the reference's API would be: sipg_sem_2d with function pointers,
// conjugate_gradient / preconditioned_conjugate_gradient over abs_mvm,
// mode_vector / host_mode_vector.  Links libsipg_hip.so only (the C-ABI).
// Built twice, with and without USE_PRECONDITIONER, as the reference's
// drivers choose the solver.
//
//   facade_test OUTDIR a1 m1 n1 [a2 m2 n2 ...]
//
// runs the 8x8, p = 3, pen 80 problem with the Fourier solution
// u = sum_k a_k sin(m_k pi x) sin(n_k pi y), writes solutions and the load
// vector (reference order, raw doubles) under OUTDIR, and prints one JSON
// line per case; tests/test_gpu_facade.py checks them against the dense
// LAPACK solve and the thesis tables.
#include "sipg_hip.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

std::vector<double> A;
std::vector<int> M, N;

template <typename T> T u_four(T x, T y) {
    double s = 0;
    for (size_t k = 0; k < A.size(); ++k) s += A[k] * std::sin(M[k] * M_PI * (double)x) * std::sin(N[k] * M_PI * (double)y);
    return (T)s;
}
template <typename T> T f_four(T x, T y) {
    double s = 0;
    for (size_t k = 0; k < A.size(); ++k) {
        const double m = M[k], n = N[k];
        s += A[k] * M_PI * M_PI * (m * m + n * n) * std::sin(m * M_PI * (double)x) * std::sin(n * M_PI * (double)y);
    }
    return (T)s;
}
template <typename T> T dx_four(T x, T y) {
    double s = 0;
    for (size_t k = 0; k < A.size(); ++k)
        s += A[k] * M[k] * M_PI * std::cos(M[k] * M_PI * (double)x) * std::sin(N[k] * M_PI * (double)y);
    return (T)s;
}
template <typename T> T dy_four(T x, T y) {
    double s = 0;
    for (size_t k = 0; k < A.size(); ++k)
        s += A[k] * N[k] * M_PI * std::sin(M[k] * M_PI * (double)x) * std::cos(N[k] * M_PI * (double)y);
    return (T)s;
}

// solution E of the reference's validation tests, with its PI
const double PI_E = 3.14159265359;
double u_E(double x, double y) { return std::cos(8 * PI_E * x) + std::cos(8 * PI_E * y); }
double f_E(double x, double y) { return 64 * PI_E * PI_E * std::cos(8 * PI_E * x) + 64 * PI_E * PI_E * std::cos(8 * PI_E * y); }
double dx_E(double x, double) { return -8 * PI_E * std::sin(8 * PI_E * x); }
double dy_E(double, double y) { return -8 * PI_E * std::sin(8 * PI_E * y); }

std::string outdir;

template <typename T>
void write(const std::string &name, const host_mode_vector<T, int> &h) {
    const std::string path = outdir + "/" + name;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    std::vector<double> d(h.data(), h.data() + h.size());
    std::fwrite(d.data(), sizeof(double), d.size(), f);
    std::fclose(f);
}

// A user operator written against abs_mvm: the SIPG operator through the
// C-ABI, with the block-Jacobi preconditioner, so the solvers run their
// generic loop (no native_op).
class user_operator : public abs_mvm<double> {
public:
    user_operator(int dim, int p, double pen) { sipg::check(sipg_op_create(&op, dim, dim, p, pen, SIPG_F64, nullptr), "op"); }
    ~user_operator() override { sipg_op_destroy(op); }
    int _mvm(mode_vector<double, int> in, mode_vector<double, int> out) const override {
        return sipg_apply(op, in.data(), out.data());
    }
    int _prec_mvm(mode_vector<double, int> in, mode_vector<double, int> out) const override {
        return sipg_prec_apply(op, in.data(), out.data());
    }

private:
    sipg_op *op = nullptr;
};

#ifdef USE_PRECONDITIONER
const char *kTag = "pcg";
#else
const char *kTag = "cg";
#endif

}  // namespace

int main(int argc, char **argv) {
    if (argc < 5 || (argc - 2) % 3) {
        std::fprintf(stderr, "usage: %s OUTDIR a1 m1 n1 [a2 m2 n2 ...]\n", argv[0]);
        return 2;
    }
    outdir = argv[1];
    for (int k = 2; k < argc; k += 3) {
        A.push_back(std::atof(argv[k]));
        M.push_back(std::atoi(argv[k + 1]));
        N.push_back(std::atoi(argv[k + 2]));
    }
    try {
        const int dim = 8, p = 3;
        const double pen = 80.0, tol = 1e-13;
        square_mesh<double> mesh(dim);
        sipg_sem_2d<double> sem(p, mesh, f_four<double>, u_four<double>, dx_four<double>, dy_four<double>, pen, tol);
        write(std::string("c1_") + kTag + "_u.bin", sem.solution);
        write(std::string("c1_") + kTag + "_rhs.bin", host_mode_vector<double, int>(sem.d_rhs));
        std::printf("{\"case\":\"c1_sem_%s\",\"iterations\":%d,\"max_err\":%.17g,\"L2_err\":%.17g,\"H1_err\":%.17g,"
                    "\"ms\":%.3f}\n",
                    kTag, sem.iterations, sem.max_err, sem.L2_err, sem.H1_err, sem.solution_time());

        // the same system through the free functions: the library operator
        // (native solvers) and a user operator (the generic loops)
        const int noe = dim * dim;
        user_operator user(dim, p, pen);
        struct { const char *name; const abs_mvm<double> *op; bool prec; } runs[] = {
            {"native_cg", &sem, false}, {"native_pcg", &sem, true},
            {"generic_cg", &user, false}, {"generic_pcg", &user, true}};
        for (const auto &r : runs) {
            host_mode_vector<double, int> zero(noe, p + 1);
            mode_vector<double, int> x(zero);
            const int it = r.prec ? preconditioned_conjugate_gradient(*r.op, x, sem.d_rhs, tol)
                                  : conjugate_gradient(*r.op, x, sem.d_rhs, tol);
            host_mode_vector<double, int> h(x);
            write(std::string("c1_") + kTag + "_" + r.name + ".bin", h);
            std::printf("{\"case\":\"c1_%s_%s\",\"iterations\":%d}\n", kTag, r.name, it);
            x.free();
        }

        // host <-> device round trip of the reference order
        host_mode_vector<double, int> h0(5, 4);
        for (int e = 0; e < 5; ++e)
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) h0(i, j, e) = 100 * e + 10 * i + j;
        mode_vector<double, int> d0(h0);
        host_mode_vector<double, int> h1(d0);
        bool same = h1.get_noe() == 5 && h1.get_nompe() == 4 && h1.size() == h0.size();
        for (int k = 0; same && k < h0.size(); ++k) same = h0.data()[k] == h1.data()[k];
        // the device array is element-major: element 1's first dof at 16
        double probe = 0;
        sipg::check(sipg_memcpy_d2h(&probe, d0.data() + 16, sizeof(double)), "probe");
        std::printf("{\"case\":\"round_trip_%s\",\"same\":%s,\"element_major\":%s}\n", kTag, same ? "true" : "false",
                    probe == 100.0 ? "true" : "false");
        d0.free();

        // FP32 instance of the facade on the same problem
        square_mesh<float> mesh32(dim);
        sipg_sem_2d<float> sem32(p, mesh32, f_four<float>, u_four<float>, dx_four<float>, dy_four<float>, 80.f, 1e-5f);
        write(std::string("c1_") + kTag + "_u_f32.bin", sem32.solution);
        std::printf("{\"case\":\"c1_sem_f32_%s\",\"iterations\":%d,\"L2_err\":%.9g}\n", kTag, sem32.iterations,
                    (double)sem32.L2_err);

        // h-convergence on solution E (the reference's sipg_2d_h_adaptivity
        // driver loop, thesis penalty 100 p^2, tolerance 1e-15)
        for (int deg = 2; deg <= 3; ++deg)
            for (int d = 4; d <= 32; d *= 2) {
                square_mesh<double> sq(d);
                sipg_sem_2d<double> s(deg, sq, f_E, u_E, dx_E, dy_E, 100.0 * deg * deg, 1e-15);
                std::printf("{\"case\":\"E_%s_%d_%d\",\"dim\":%d,\"p\":%d,\"iterations\":%d,\"max_err\":%.9g,"
                            "\"L2_err\":%.9g,\"H1_err\":%.9g}\n",
                            kTag, d, deg, d, deg, s.iterations, s.max_err, s.L2_err, s.H1_err);
                sq.device_info.free();
            }
    } catch (const sipg::error &e) {
        std::fprintf(stderr, "sipg error: %s\n", e.what());
        return 1;
    }
    return 0;
}
