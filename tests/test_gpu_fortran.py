"""The Fortran bindings (cel_v3_amd/fortran/cel_gpu_mod.f90) driving the
library: the element-solver layer (row-wise assembly, MatrixCompleted with
diagonal row scaling + Jacobi vector, DoIt with the reference's x0 = 1 and
absolute eps = 1e-8, compared with the oracle on the same scaled system),
and a batched solve through the C-pointer-array interface."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

BUILD = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cel_v3_amd",
                     "fortran", "_build")
EXE = os.path.join(BUILD, "elm_poisson")
BATCH_EXE = os.path.join(BUILD, "batch_poisson")


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


def test_fortran_batch_driver(gpu):
    """A 10^3 system and a 6x7x8 system with 3 right-hand sides solved as one
    batch from Fortran; every lane converges to 1e-10 (true residual from
    the library's FP64 SpMV) in the iteration count of the single-system
    solver (+-1)."""
    if not os.path.exists(BATCH_EXE):
        pytest.skip("Fortran driver not built (no amdflang)")
    out = subprocess.run([BATCH_EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    lanes = [ln.split() for ln in out.stdout.splitlines() if ln.startswith("lane")]
    assert len(lanes) == 4
    n = 1000
    b1 = 1.0 + (np.arange(1, n + 1) % 7) * 0.25
    M1 = gpu.Matrix.poisson7(10, 10, 10)
    M1.set_jacobi()
    M2 = gpu.Matrix.poisson7(6, 7, 8)
    M2.set_jacobi()
    i = np.arange(1, 337)
    rhs = [(M1, b1)] + [(M2, (i * (j + 3) % 11) - 5.0) for j in range(3)]
    for ln, (M, b) in zip(lanes, rhs):
        status, iters, relres = int(ln[3]), int(ln[5]), float(ln[7])
        single = M.pcg(b, tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, alpha_stop=0)
        assert status == 0 and relres <= 1e-10 * 1.01
        assert abs(iters - single["iters"]) <= 1
