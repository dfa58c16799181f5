"""The Fortran element-solver layer (cel_v3_amd/fortran/cel_gpu_mod.f90)
driving the library: row-wise assembly, MatrixCompleted (diagonal row
scaling + Jacobi vector), DoIt with the reference's x0 = 1 and absolute
eps = 1e-8, compared with the oracle on the same scaled system."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cel_v3_amd",
                   "fortran", "_build", "elm_poisson")


def test_fortran_elm_driver(gpu):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (no amdflang)")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    first = out.stdout.split("\n")[0].split()
    status, iters, norm, maxerr = int(first[1]), int(first[3]), float(first[5]), float(first[7])
    n = 12
    A = O.gen_poisson7(n, n, n)
    xt = (np.arange(1, A.n + 1) % 7) / 7.0
    b = O.spmv_csr(A, xt)
    d = O.csr_diagonal(A)
    S = O.scale_csr(A, d)
    ref = O.pcg_jacobi(S, O.set_jacobi(d, scaled=True), O.scale_vec(b, d), x0_mode=O.X0_ONE,
                       eps=1e-8, iter_max=5000, workers=3)
    assert status == 0
    assert abs(iters - ref["iters"]) <= 1
    assert abs(norm - ref["norm"]) <= 1e-12 * ref["hist"][0]
    assert maxerr < 1e-6
