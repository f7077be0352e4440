"""Pins the oracle's print_result text (sipg_sem_2d.hpp:375-403).

The reference streams each value with std::setw(10) at the stream's default
precision; the oracle and the library print with "%10g".  The two are the
same conversion: a small C++ program compiled here prints a spread of values
both ways.  The file's content is then checked from first principles: the
node coordinates of the mesh builder's convention, the values of u, and the
layout (one line per node, x index outer, an empty line per element)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

_PROBE = r"""
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <sstream>
int main() {
    const double v[] = {0.0, 1.0, -1.0, 0.5, 0.123456789, 1234567.0, 1e-5, -2.77556e-16, 3.0e100,
                        0.999999949, 6.103515625e-05, 100000.0, 999999.5, 1.0 / 3.0, -0.0};
    int bad = 0;
    for (double x : v) {
        std::ostringstream s;
        s << std::setw(10) << x;
        char b[64];
        std::snprintf(b, sizeof b, "%10g", x);
        if (s.str() != b) { std::printf("%s|%s\n", s.str().c_str(), b); ++bad; }
    }
    float f = 0.1f;
    std::ostringstream s;
    s << std::setw(10) << f;
    char b[64];
    std::snprintf(b, sizeof b, "%10g", (double)f);
    if (s.str() != b) ++bad;
    return bad;
}
"""


def test_setw10_is_printf_10g(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    src = tmp_path / "probe.cpp"
    src.write_text(_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-O0", "-o", str(exe), str(src)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout


@pytest.mark.parametrize("dim,p", [(3, 2), (4, 5)])
def test_print_result_content(oracle, tmp_path, dim, p):
    n1, noe = p + 1, dim * dim
    u = np.random.default_rng(dim + p).uniform(-1, 1, noe * n1 * n1)
    path = tmp_path / "u.txt"
    oracle.print_result(dim, p, u, str(path))
    text = path.read_text()
    blocks = text.split("\n\n")
    assert blocks[-1] == "" and len(blocks) == noe + 1            # an empty line after each element
    z, _ = oracle.gll(p)
    for e, blk in enumerate(blocks[:-1]):
        lines = blk.split("\n")
        assert len(lines) == n1 * n1
        row, col = divmod(e, dim)                                # row 0 at the top
        for k, line in enumerate(lines):
            f = line.split("\t")
            assert len(f) == 3 and all(len(x) >= 10 for x in f)
            i, j = divmod(k, n1)                                 # x index outer
            x, y, v = (float(t) for t in f)
            assert abs(x - (col + (z[i] + 1) / 2) / dim) <= 5e-6 * max(1.0, abs(x))
            assert abs(y - (1 - (row + 1 - (z[j] + 1) / 2) / dim)) <= 5e-6 * max(1.0, abs(y)) + 1e-15
            assert abs(v - u[(i * n1 + j) * noe + e]) <= 5e-6 * abs(v) + 1e-300
