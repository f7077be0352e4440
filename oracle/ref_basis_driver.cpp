// Exposes the reference's own host basis numerics (compiled from the
// reference sources where they lie, never copied) through a C symbol, so
// tests can check the oracle's GLL tables against them.  TEST
// INFRASTRUCTURE ONLY; built into oracle/_ref/ by oracle/Makefile.
#include <LGL_quadrature_table.hpp>
#include <sem_function.hpp>
#include <laplacian_operator.hpp>

extern "C" int ref_basis(int p, double *nodes, double *weights, double *D, double *K1)
{
    if (p < 1) return -1;
    LGL_quadrature_table<double> qt(p + 1);      // LGL_quadrature_table.hpp:84
    sem_function<double> sf(p);                  // sem_function.hpp:57
    laplace_1d<double> k1(p);                    // laplacian_operator.hpp:17
    for (int i = 0; i <= p; ++i) {
        nodes[i] = qt.node(i);
        weights[i] = qt.weight(i);
    }
    for (int mode = 0; mode <= p; ++mode)
        for (int node = 0; node <= p; ++node)
            D[mode * (p + 1) + node] = sf.d_phi(mode, node);   // device_const.hpp:123-125
    for (int a = 0; a <= p; ++a)
        for (int b = 0; b <= p; ++b)
            K1[a * (p + 1) + b] = k1(a, b);
    return 0;
}
