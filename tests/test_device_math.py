"""Host checks of the device math helpers in csrc/bd_dev.h.

The helpers marked __host__ __device__ are compiled for the CPU with hipcc's
host pass and measured against 80-bit long double references:
- exp_neg_t (the table form of exp(x), x <= 0, used by the core-core kernel in
  place of the reference's exp of func.cpp:278-371), and its 2^(j/32) table
  against the correctly rounded values from exact decimal arithmetic;
- sincos_small (the sin/cos of the linker mechanics and the rotation,
  func.cpp:951-1289 and 1607-1670).
The squared forms of the reference's distance cut-offs are checked against
numpy's correctly rounded square root on the doubles around each threshold.
"""
import os
import shutil
import subprocess
from decimal import Decimal, getcontext

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "brownian_dynamics_cuda_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

PROG = r"""
#include "bd_dev.h"
#include <cmath>
#include <cstdio>
#include <random>
int main() {
    for (int j = 0; j < 32; j++) printf("tab %d %a\n", j, bd::kExp2Tab.t[j]);
    std::mt19937_64 gen(12345);
    std::uniform_real_distribution<double> ue(-60.0, 0.0), us(-7.0, 7.0);
    long double emax = 0;
    double xe = 0;
    for (int i = 0; i < 2000000; i++) {
        // random points, and points at the edges of the reduction intervals
        double x = (i % 4 == 3) ? -(i / 4) * (0.69314718055994530942 / 64.0) : ue(gen);
        if (x < -60.0) x = ue(gen);
        const long double ref = expl((long double)x);
        const long double e = fabsl(((long double)bd::exp_neg_t(x, bd::kExp2Tab.t, bd::kExpCoef6) - ref) / ref);
        if (e > emax) { emax = e; xe = x; }
    }
    printf("exp %.6Le %a\n", emax, xe);
    long double smax = 0, cmax = 0;
    for (int i = 0; i < 2000000; i++) {
        const double x = us(gen);
        double s, c;
        bd::sincos_small(x, s, c);
        smax = fmaxl(smax, fabsl((long double)s - sinl((long double)x)));
        cmax = fmaxl(cmax, fabsl((long double)c - cosl((long double)x)));
    }
    printf("sincos %.6Le %.6Le\n", smax, cmax);
    return 0;
}
"""


@pytest.fixture(scope="module")
def measured(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("devmath")
    src, exe = d / "devmath.cpp", d / "devmath"
    src.write_text(PROG)
    cmd = [HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O2", "-I" + CSRC, str(src), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    res = {"tab": {}}
    for line in out.splitlines():
        f = line.split()
        if f[0] == "tab":
            res["tab"][int(f[1])] = float.fromhex(f[2])
        elif f[0] == "exp":
            res["exp"] = float(f[1])
        elif f[0] == "sincos":
            res["sin"], res["cos"] = float(f[1]), float(f[2])
    return res


def test_exp_table_correctly_rounded(measured):
    getcontext().prec = 60
    for j in range(32):
        exact = Decimal(2) ** (Decimal(j) / Decimal(32))
        assert measured["tab"][j] == float(exact), j


def test_exp_neg_t_accuracy(measured):
    # table entry (half an ulp), series and product roundings: a few ulps at most
    assert measured["exp"] <= 4e-16


def test_sincos_small_accuracy(measured):
    assert measured["sin"] <= 3e-16
    assert measured["cos"] <= 3e-16


def _doubles_around(t, n=200000):
    import numpy as np
    base = np.array([t], dtype=np.float64).view(np.int64)[0]
    return (base + np.arange(-n, n + 1, dtype=np.int64)).view(np.float64)


def test_squared_cutoffs_exact():
    """The kernels test the reference's cut-offs on d = sqrt(d2) as tests on d2
    (func.cpp:278-371 for 7 and 4 nm; 1291-1605 for 8 and 25 nm): each squared
    threshold must give the same answer as the correctly rounded root for every
    double near it, and bd_dev.h must hold exactly these thresholds."""
    import re
    import numpy as np
    src = open(os.path.join(CSRC, "bd_dev.h")).read()
    d8 = float.fromhex(re.search(r"kD8Sq = (0x[0-9a-fA-Fp.+]+);", src).group(1))
    d25 = float(re.search(r"kD25Sq = ([0-9.]+);", src).group(1))
    assert d8 == np.nextafter(64.0, np.inf) and d25 == 625.0
    assert "if (!(d2 < 49.0)) return false;" in src and "if (d2 < 16.0) {" in src
    for r, t, strict in ((7.0, 49.0, True), (4.0, 16.0, True), (8.0, d8, False), (25.0, d25, False)):
        x = _doubles_around(t)
        root = np.sqrt(x)
        if strict:
            assert np.array_equal(root < r, x < t), r
        else:
            assert np.array_equal(root <= r, x <= t), r
