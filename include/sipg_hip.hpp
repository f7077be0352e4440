/*
 * sipg_hip.hpp -- C++ facade of libsipg_hip with the reference's class and
 * function names, so code written against the reference's sipg_sem_2d /
 * abs_mvm / conjugate_gradient / mode_vector API builds against this
 * library unchanged.  Header only, over the C-ABI of sipg_hip.h: it needs
 * no HIP headers and links against libsipg_hip.so alone.
 *
 *   sipg_sem_2d<T>        sipg_sem_2d.hpp:31-230 (solve in the constructor;
 *                         max_err, L2_err, H1_err, iterations, solution,
 *                         d_u, d_rhs, solution_time(), print_result(),
 *                         _mvm, _prec_mvm)
 *   abs_mvm<T>            abs_mvm.hpp:10-19
 *   conjugate_gradient    conjugate_gradient_gpu.hpp:18-109
 *   preconditioned_conjugate_gradient
 *                         conjugate_gradient_gpu.hpp:116-213
 *   mode_vector<T,I>      mode_vector.hpp:37-198 (device vector handle)
 *   host_mode_vector<T,I> mode_vector.hpp:207-304
 *   copy(...)             mode_vector.hpp:313-359
 *   square_mesh<T>        build_square_mesh.hpp (dim, noe, element corners)
 *
 * Differences a caller can see:
 *  - mode_vector's device array is element-major (this library's layout:
 *    dof (i1,i2,e) at e*nompe^2 + i1*nompe + i2); host_mode_vector keeps the
 *    reference order ((i1*nompe+i2)*noe + e) and copy() converts.  No
 *    __device__ operator() is offered on mode_vector.
 *  - errors (a failed allocation, an invalid argument, a HIP error, CG
 *    breakdown) throw sipg::error instead of printing; reaching the
 *    iteration cap is not an error (the iteration count is returned, as in
 *    the reference).
 *  - abs_mvm gains one virtual with a default, native_op(): an operator of
 *    this library returns its handle there and the two CG functions then
 *    run the library's device-resident solvers (sipg_cg, sipg_pcg) with the
 *    reference's stopping rule; any other abs_mvm runs the reference's CG
 *    loop over its _mvm / _prec_mvm with this library's vector kernels.
 *  - as in the reference, defining USE_PRECONDITIONER before including
 *    this header makes sipg_sem_2d solve with preconditioned CG.
 *
 * All work is on the default stream; mode_vector is a shallow handle (copies
 * share the device array; free() releases it), as the reference's.
 */
#ifndef SIPG_HIP_HPP
#define SIPG_HIP_HPP

#include "sipg_hip.h"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace sipg {

class error : public std::runtime_error {
public:
    error(int status, const std::string &what)
        : std::runtime_error(what + ": " + sipg_strerror(status)), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int status, const char *what) {
    if (status != SIPG_OK) throw error(status, what);
}

template <typename T> struct dtype_of;
template <> struct dtype_of<double> { static constexpr int value = SIPG_F64; };
template <> struct dtype_of<float> { static constexpr int value = SIPG_F32; };

template <typename FLOAT_TYPE, typename INT_TYPE = int> class host_mode_vector;

// Device vector handle (mode_vector.hpp:37-198).
template <typename FLOAT_TYPE, typename INT_TYPE = int>
class mode_vector {
public:
    mode_vector() : noe(0), nompe(0), device_array(nullptr) {}
    // allocates noe * nompe^2 values on the device (contents undefined)
    mode_vector(INT_TYPE _noe, INT_TYPE _nompe) : noe(_noe), nompe(_nompe), device_array(nullptr) {
        void *p = nullptr;
        check(sipg_malloc(&p, bytes()), "mode_vector");
        device_array = static_cast<FLOAT_TYPE *>(p);
    }
    // maps device memory already present (element-major)
    mode_vector(INT_TYPE _noe, INT_TYPE _nompe, FLOAT_TYPE *_d_a) : noe(_noe), nompe(_nompe), device_array(_d_a) {}
    // a new device copy of a host vector
    mode_vector(const host_mode_vector<FLOAT_TYPE, INT_TYPE> &hmv) : noe(0), nompe(0), device_array(nullptr) {
        copy(hmv, *this);
    }

    const FLOAT_TYPE *data() const { return device_array; }
    FLOAT_TYPE *data() { return device_array; }
    INT_TYPE size() const { return noe * nompe * nompe; }
    INT_TYPE get_noe() const { return noe; }
    INT_TYPE get_nompe() const { return nompe; }
    void free() {
        if (device_array) sipg_free(device_array);
        device_array = nullptr;
    }

private:
    size_t bytes() const { return (size_t)noe * nompe * nompe * sizeof(FLOAT_TYPE); }
    INT_TYPE noe, nompe;
    FLOAT_TYPE *device_array;

    template <typename F, typename I>
    friend void copy(const host_mode_vector<F, I> &from, mode_vector<F, I> &to);
};

// Host vector in the reference order (mode_vector.hpp:207-304).
template <typename FLOAT_TYPE, typename INT_TYPE>
class host_mode_vector {
public:
    host_mode_vector() : noe(0), nompe(0) {}
    host_mode_vector(INT_TYPE _noe, INT_TYPE _nompe)
        : noe(_noe), nompe(_nompe), host_vector((size_t)_noe * _nompe * _nompe, 0) {}
    host_mode_vector(const mode_vector<FLOAT_TYPE, INT_TYPE> &mv) : noe(0), nompe(0) { copy(mv, *this); }

    INT_TYPE get_noe() const { return noe; }
    INT_TYPE get_nompe() const { return nompe; }
    INT_TYPE size() const { return (INT_TYPE)host_vector.size(); }
    const FLOAT_TYPE *data() const { return host_vector.data(); }
    // dof of mode i1 (x), mode i2 (y) on element e_idx
    FLOAT_TYPE &operator()(INT_TYPE i1, INT_TYPE i2, INT_TYPE e_idx) {
        return host_vector[((size_t)i1 * nompe + i2) * noe + e_idx];
    }
    const FLOAT_TYPE &operator()(INT_TYPE i1, INT_TYPE i2, INT_TYPE e_idx) const {
        return host_vector[((size_t)i1 * nompe + i2) * noe + e_idx];
    }

private:
    INT_TYPE noe, nompe;
    std::vector<FLOAT_TYPE> host_vector;

    template <typename F, typename I>
    friend void copy(const mode_vector<F, I> &from, host_mode_vector<F, I> &to);
    template <typename F, typename I>
    friend void copy(const host_mode_vector<F, I> &from, mode_vector<F, I> &to);
};

// device -> host, element-major -> reference order (mode_vector.hpp:313-333)
template <typename F, typename I>
void copy(const mode_vector<F, I> &from, host_mode_vector<F, I> &to) {
    const int64_t noe = from.get_noe(), nq = (int64_t)from.get_nompe() * from.get_nompe();
    std::vector<F> tmp((size_t)(noe * nq));
    check(sipg_device_synchronize(), "copy");
    if (!tmp.empty()) check(sipg_memcpy_d2h(tmp.data(), from.data(), tmp.size() * sizeof(F)), "copy");
    to.noe = from.get_noe();
    to.nompe = from.get_nompe();
    to.host_vector.resize(tmp.size());
    for (int64_t e = 0; e < noe; ++e)
        for (int64_t q = 0; q < nq; ++q) to.host_vector[(size_t)(q * noe + e)] = tmp[(size_t)(e * nq + q)];
}

// host -> device, reference order -> element-major (mode_vector.hpp:
// 336-359); allocates the device array unless `to` already holds one of
// the same shape
template <typename F, typename I>
void copy(const host_mode_vector<F, I> &from, mode_vector<F, I> &to) {
    const int64_t noe = from.get_noe(), nq = (int64_t)from.get_nompe() * from.get_nompe();
    if (!to.device_array || to.noe != from.get_noe() || to.nompe != from.get_nompe()) {
        to.noe = from.get_noe();
        to.nompe = from.get_nompe();
        void *p = nullptr;
        check(sipg_malloc(&p, to.bytes()), "copy");
        to.device_array = static_cast<F *>(p);
    }
    std::vector<F> tmp((size_t)(noe * nq));
    for (int64_t e = 0; e < noe; ++e)
        for (int64_t q = 0; q < nq; ++q) tmp[(size_t)(e * nq + q)] = from.host_vector[(size_t)(q * noe + e)];
    if (!tmp.empty()) check(sipg_memcpy_h2d(to.device_array, tmp.data(), tmp.size() * sizeof(F)), "copy");
}

// Square mesh of the unit square, dim x dim elements (build_square_mesh.hpp):
// element e = row * dim + col, row 0 at the top; x_coord()/y_coord() are
// the corners the reference's builder places (x summed along each row from
// 0, y stepped down from 1 once per row).
template <typename FLOAT_TYPE>
class square_mesh {
public:
    // the reference keeps a device copy of the neighbour codes here; this
    // library computes them inline, so there is nothing to free
    struct device_mesh_info {
        int dimx = 0, dimy = 0;
        int get_dimx() const { return dimx; }
        int get_dimy() const { return dimy; }
        void free() {}
    };

    square_mesh() : dim_(0) {}
    explicit square_mesh(int _dim) : dim_(_dim) {
        device_info.dimx = device_info.dimy = _dim;
        const FLOAT_TYPE delta = FLOAT_TYPE(1) / FLOAT_TYPE(_dim);
        FLOAT_TYPE yy = 1;
        for (int row = 0; row < _dim; ++row) {
            FLOAT_TYPE xx = 0;
            yy -= delta;
            for (int col = 0; col < _dim; ++col) {
                x_.push_back(xx);
                y_.push_back(yy);
                xx += delta;
            }
        }
    }
    int dim() const { return dim_; }
    int noe() const { return dim_ * dim_; }
    const std::vector<FLOAT_TYPE> &x_coord() const { return x_; }
    const std::vector<FLOAT_TYPE> &y_coord() const { return y_; }

    device_mesh_info device_info;

private:
    int dim_;
    std::vector<FLOAT_TYPE> x_, y_;
};

// Operator interface of the iterative solvers (abs_mvm.hpp:10-19).
template <typename FLOAT_TYPE>
class abs_mvm {
public:
    virtual ~abs_mvm() {}
    virtual int _mvm(mode_vector<FLOAT_TYPE, int>, mode_vector<FLOAT_TYPE, int>) const = 0;
    virtual int _prec_mvm(mode_vector<FLOAT_TYPE, int>, mode_vector<FLOAT_TYPE, int>) const = 0;
    // a library operator behind this problem, if any (see the file comment)
    virtual sipg_op *native_op() const { return nullptr; }
};

namespace detail {

template <typename T>
double dot(const mode_vector<T, int> &a, const mode_vector<T, int> &b) {
    double r = 0;
    check(sipg_vec_dot(dtype_of<T>::value, a.size(), a.data(), b.data(), &r, nullptr), "dot");
    return r;
}

// The reference's CG loop (conjugate_gradient_gpu.hpp:18-109) over any
// operator: stop on sqrt(r.r) <= tol, at most b.size() iterations.
template <typename T>
int cg_loop(const abs_mvm<T> &problem, mode_vector<T, int> x, mode_vector<T, int> b, T tol, bool prec) {
    const int dt = dtype_of<T>::value;
    const int noe = x.get_noe(), nompe = x.get_nompe();
    const int64_t n = b.size();
    mode_vector<T, int> r(noe, nompe), p(noe, nompe), Ap(noe, nompe), z;
    if (prec) z = mode_vector<T, int>(noe, nompe);
    int it = 0;
    try {
        problem._mvm(x, r);
        check(sipg_vec_xpby(dt, n, b.data(), T(-1), r.data(), nullptr), "cg");        // r = b - A x
        T rr_old, rr = 1, pAp;
        if (prec) {
            problem._prec_mvm(r, z);
            check(sipg_vec_copy(dt, n, z.data(), p.data(), nullptr), "cg");           // p = z
            rr_old = (T)dot(r, z);
        } else {
            check(sipg_vec_copy(dt, n, r.data(), p.data(), nullptr), "cg");           // p = r
            rr_old = (T)dot(r, r);
            rr = rr_old;
        }
        const T stop = tol;
        const int64_t max_it = n;
        while (std::sqrt(rr) > stop && it < max_it) {
            problem._mvm(p, Ap);
            pAp = (T)dot(p, Ap);
            const T alpha = rr_old / pAp;
            check(sipg_vec_alpha_update(dt, n, alpha, x.data(), p.data(), r.data(), Ap.data(), nullptr), "cg");
            rr = (T)dot(r, r);
            T rr_new = rr;
            if (prec) {
                problem._prec_mvm(r, z);
                rr_new = (T)dot(z, r);
            }
            const T beta = rr_new / rr_old;
            check(sipg_vec_xpby(dt, n, prec ? z.data() : r.data(), beta, p.data(), nullptr), "cg");
            rr_old = rr_new;
            ++it;
        }
        check(sipg_device_synchronize(), "cg");
    } catch (...) {
        r.free(); p.free(); Ap.free(); z.free();
        throw;
    }
    r.free(); p.free(); Ap.free(); z.free();
    return it;
}

// The library's device-resident solver with the reference's stopping rule
// (absolute sqrt(r.r) <= tol on the recursive residual, cap n).
template <typename T>
int native_solve(sipg_op *op, mode_vector<T, int> x, mode_vector<T, int> b, T tol, bool prec) {
    sipg_cg_opts o{(double)tol, SIPG_TOL_ABS, 0, 0.0, 0, 0};
    sipg_cg_info info{};
    const int st = prec ? sipg_pcg(op, x.data(), b.data(), &o, &info) : sipg_cg_ex(op, x.data(), b.data(), &o, &info);
    if (st != SIPG_OK && st != SIPG_NOT_CONVERGED) throw error(st, prec ? "preconditioned_conjugate_gradient"
                                                                       : "conjugate_gradient");
    return info.iters;
}

// Up to three functions of (x, y) in the reference's form behind one
// sipg_fn2 context; call<T, K> evaluates the K-th.
template <typename T>
struct fns {
    T (*fn[3])(T, T);
};
template <typename T, int K>
double call(double x, double y, void *ctx) {
    return (double)static_cast<const fns<T> *>(ctx)->fn[K]((T)x, (T)y);
}

}  // namespace detail

// CG on problem x = b, x in/out (conjugate_gradient_gpu.hpp:18-109):
// stops on sqrt(r.r) <= tol, at most b.size() iterations; returns the
// iteration count.
template <typename FLOAT_TYPE>
int conjugate_gradient(const abs_mvm<FLOAT_TYPE> &problem, mode_vector<FLOAT_TYPE, int> x,
                       mode_vector<FLOAT_TYPE, int> b, FLOAT_TYPE tol) {
    if (sipg_op *op = problem.native_op()) return detail::native_solve(op, x, b, tol, false);
    return detail::cg_loop(problem, x, b, tol, false);
}

// Preconditioned CG (conjugate_gradient_gpu.hpp:116-213): M^-1 = _prec_mvm,
// stopping test on sqrt(r.r) <= tol, at least one iteration.
template <typename FLOAT_TYPE>
int preconditioned_conjugate_gradient(const abs_mvm<FLOAT_TYPE> &problem, mode_vector<FLOAT_TYPE, int> x,
                                      mode_vector<FLOAT_TYPE, int> b, FLOAT_TYPE tol) {
    if (sipg_op *op = problem.native_op()) return detail::native_solve(op, x, b, tol, true);
    return detail::cg_loop(problem, x, b, tol, true);
}

// The SIPG Poisson problem on the unit square with Dirichlet data u_ex,
// solved in the constructor (sipg_sem_2d.hpp:31-230).
template <typename FLOAT_TYPE>
class sipg_sem_2d : public abs_mvm<FLOAT_TYPE> {
public:
    FLOAT_TYPE max_err = 0;
    FLOAT_TYPE L2_err = 0;
    FLOAT_TYPE H1_err = 0;
    int iterations = 0;

    mode_vector<FLOAT_TYPE, int> d_u;
    mode_vector<FLOAT_TYPE, int> d_rhs;

    host_mode_vector<FLOAT_TYPE, int> solution;

    sipg_sem_2d(int _order, const square_mesh<FLOAT_TYPE> &_mesh, FLOAT_TYPE (*f)(FLOAT_TYPE, FLOAT_TYPE),
                FLOAT_TYPE (*u_ex)(FLOAT_TYPE, FLOAT_TYPE), FLOAT_TYPE (*dx_u_ex)(FLOAT_TYPE, FLOAT_TYPE),
                FLOAT_TYPE (*dy_u_ex)(FLOAT_TYPE, FLOAT_TYPE), FLOAT_TYPE _pen, FLOAT_TYPE _tol)
        : order(_order), dim(_mesh.dim()) {
        check(sipg_op_create(&op, dim, dim, order, (double)_pen, dtype_of<FLOAT_TYPE>::value, nullptr),
              "sipg_sem_2d");
        try {
            const int noe = dim * dim;
            d_u = mode_vector<FLOAT_TYPE, int>(noe, order + 1);
            d_rhs = mode_vector<FLOAT_TYPE, int>(noe, order + 1);
            check(sipg_memset_zero(d_u.data(), (size_t)d_u.size() * sizeof(FLOAT_TYPE)), "sipg_sem_2d");
            // compute_rhs(f, u_ex): Dirichlet data g = u_ex
            detail::fns<FLOAT_TYPE> fg{{f, u_ex, nullptr}};
            check(sipg_rhs(op, detail::call<FLOAT_TYPE, 0>, detail::call<FLOAT_TYPE, 1>, &fg, d_rhs.data()),
                  "compute_rhs");
            const auto t0 = std::chrono::steady_clock::now();
#ifdef USE_PRECONDITIONER
            iterations = preconditioned_conjugate_gradient(*this, d_u, d_rhs, _tol);
#else
            iterations = conjugate_gradient(*this, d_u, d_rhs, _tol);
#endif
            check(sipg_device_synchronize(), "sipg_sem_2d");
            solve_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
            copy(d_u, solution);
            // err_norms(solution, f, u_ex, dx_u_ex, dy_u_ex)
            detail::fns<FLOAT_TYPE> ex{{u_ex, dx_u_ex, dy_u_ex}};
            double out[3];
            check(sipg_err_norms(op, d_u.data(), detail::call<FLOAT_TYPE, 0>, detail::call<FLOAT_TYPE, 1>,
                                 detail::call<FLOAT_TYPE, 2>, &ex, out),
                  "err_norms");
            max_err = (FLOAT_TYPE)out[0];
            L2_err = (FLOAT_TYPE)out[1];
            H1_err = (FLOAT_TYPE)out[2];
        } catch (...) {
            release();
            throw;
        }
    }

    ~sipg_sem_2d() override { release(); }
    sipg_sem_2d(const sipg_sem_2d &) = delete;
    sipg_sem_2d &operator=(const sipg_sem_2d &) = delete;

    // wall time of the CG solve, ms
    float solution_time() { return solve_ms; }

    // x y u per node as text (sipg_sem_2d.hpp:375-403), to standard error
    void print_result() { check(sipg_write_result(op, d_u.data(), nullptr), "print_result"); }

    int _mvm(mode_vector<FLOAT_TYPE, int> input, mode_vector<FLOAT_TYPE, int> output) const override {
        check(sipg_apply(op, input.data(), output.data()), "_mvm");
        return 0;
    }
    int _prec_mvm(mode_vector<FLOAT_TYPE, int> input, mode_vector<FLOAT_TYPE, int> output) const override {
        check(sipg_prec_apply(op, input.data(), output.data()), "_prec_mvm");
        return 0;
    }
    sipg_op *native_op() const override { return op; }

private:
    void release() {
        d_u.free();
        d_rhs.free();
        if (op) sipg_op_destroy(op);
        op = nullptr;
    }

    int order, dim;
    sipg_op *op = nullptr;
    float solve_ms = 0.f;
};

}  // namespace sipg

#ifndef SIPG_HIP_NO_GLOBAL_NAMES
// the reference's names at global scope, for code written against it
using sipg::abs_mvm;
using sipg::conjugate_gradient;
using sipg::copy;
using sipg::host_mode_vector;
using sipg::mode_vector;
using sipg::preconditioned_conjugate_gradient;
using sipg::sipg_sem_2d;
using sipg::square_mesh;
#endif

#endif /* SIPG_HIP_HPP */
