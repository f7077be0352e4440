"""Regenerates the committed golden fixtures under tests/golden/.

    python tests/golden/make_golden.py

* basis_ref.json   -- LGL nodes/weights, derivative table D and 1-D stiffness
                      K1 for p = 1..16, 20, 24, 32, 40, 48 and 63 (the
                      reference's largest table), produced by the REFERENCE's own host
                      code (LGL_quadrature_table.hpp, sem_function.hpp,
                      laplacian_operator.hpp) compiled from its sources into
                      oracle/_ref by `make -C oracle ref`; stored as hex floats.
* apply_vectors.npz -- y = A x for seeded x ~ U(-1,1) at (dim=4, p=3, pen=80),
                      (dim=3, p=8, pen=405), (dim=2, p=15, pen=1280) and a
                      ragged 5x3 mesh at p=2; y from the oracle's restatement,
                      which the tests cross-check against the independent dense
                      assembly (oracle/dense.py).
* dense_d2p2.npy   -- the dense SIPG matrix for dim=2, p=2, pen=45 (36x36),
                      from the independent assembly.
* c1_solution.npz  -- configuration C1 (8x8, p=3, pen=5(p+1)^2=80, Fourier
                      forcing drawn from seed 0): rhs and the dense LAPACK
                      (Cholesky) solution.
thesis_tables.json is transcribed data (see its "source" field), not generated.
"""
import json
import os
import sys

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle as O   # noqa: E402
import dense         # noqa: E402
from mt19937_64 import forcing_draw   # noqa: E402

BASIS_DEGREES = list(range(1, 17)) + [20, 24, 32, 40, 48, 63]


def main():
    # basis from the reference's own host code
    if O.ref_basis_available():
        out = {"source": "reference host code via oracle/_ref/libref_basis.so", "p": {}}
        for p in BASIS_DEGREES:
            z, w, D, K = O.ref_basis(p)
            out["p"][str(p)] = {k: [float(v).hex() for v in arr.ravel()]
                                for k, arr in (("nodes", z), ("weights", w), ("D", D), ("K1", K))}
        with open(os.path.join(HERE, "basis_ref.json"), "w") as f:
            json.dump(out, f, indent=0)
    else:
        print("oracle/_ref not built: basis_ref.json left as is")

    rng = np.random.default_rng(1)
    vecs = {}
    for name, (dx, dy, p, pen) in {"d4p3": (4, 4, 3, 80.0), "d3p8": (3, 3, 8, 405.0),
                                   "d2p15": (2, 2, 15, 1280.0), "r5x3p2": (5, 3, 2, 45.0)}.items():
        n = dx * dy * (p + 1) ** 2
        x = rng.uniform(-1, 1, n)
        vecs[name + "_x"] = x
        vecs[name + "_y"] = O.apply(dx, dy, p, pen, x)
        vecs[name + "_shape"] = np.array([dx, dy, p, pen])
    np.savez(os.path.join(HERE, "apply_vectors.npz"), **vecs)

    np.save(os.path.join(HERE, "dense_d2p2.npy"), dense.assemble(2, 2, 2, 45.0))

    a, m, nn = forcing_draw(0, 4, 4, 4)
    prob = O.problem(0, (a, m, nn))
    b = O.rhs(8, 3, 80.0, prob)
    A = dense.assemble(8, 8, 3, 80.0)
    c, low = scipy.linalg.cho_factor(A)
    u = scipy.linalg.cho_solve((c, low), b)
    np.savez(os.path.join(HERE, "c1_solution.npz"), rhs=b, u=u, a=a, m=m, n=nn)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
