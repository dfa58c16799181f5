"""The Fortran bindings (cel_v3_amd/fortran/cel_gpu_mod.f90) driving the
library: the element-solver layer (row-wise assembly, MatrixCompleted with
the reference's sort, diagonal row scaling + Jacobi vector, DoIt/ReDoIt with
the reference's Jacobi-GMRES(3), x0 = 1 and absolute eps = 1e-8, compared
with the oracle on the same scaled system), and a batched solve through the
C-pointer-array interface."""
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
    """tests/fortran/elm_poisson.f90 through the reference's API names:
    a nonsymmetric 12^3 operator solved with the default Jacobi-GMRES(3)
    (DoIt, then ReDoIt with the same b, which the reference does not scale),
    then the Poisson operator with Jacobi-PCG; each against the oracle
    pipeline of MatrixCompleted + DoIt/ReDoIt."""
    from test_gpu_gmres import oracle_elm
    from test_oracle import convdiff_csr

    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (no amdflang)")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    lines = [ln.split() for ln in out.stdout.splitlines() if ln.startswith("status")]
    assert len(lines) == 3
    n = 12
    cases = [(convdiff_csr(n, n, n, 0.5), "gmres", True), (convdiff_csr(n, n, n, 0.5), "gmres", False),
             (O.gen_poisson7(n, n, n), "pcg", True)]
    for ln, (A, solver, scale_b) in zip(lines, cases):
        status, iters, norm, maxerr = int(ln[1]), int(ln[3]), float(ln[5]), float(ln[7])
        xt = (np.arange(1, A.n + 1) % 7) / 7.0
        b = O.spmv_csr(A, xt)
        rows = np.repeat(np.arange(1, A.n + 1), np.diff(A.xadj))
        _, ref = oracle_elm(rows, A.cols, A.vals, A.n, b, solver=solver, scale_b=scale_b)
        assert status == 0
        if solver == "gmres":
            assert iters == ref["iters"]
            assert abs(norm - ref["norm"]) <= 1e-10 * np.linalg.norm(b)
            want = np.abs(ref["x"] - xt).max()
            assert abs(maxerr - want) <= 1e-4 * want + 1e-14  # printed with 5 digits
        else:
            assert abs(iters - ref["iters"]) <= 1
            assert abs(norm - ref["norm"]) <= 1e-12 * ref["hist"][0]
            assert maxerr < 1e-6


def test_fortran_batch_driver(gpu):
    """A 10^3 system and a 6x7x8 system with 3 right-hand sides solved as one
    batch from Fortran, on a configuration fixed through the Fortran binding
    (16 blocks for the SpMV, 8 for the vector kernels, 4 iterations per
    replay) and read back; every lane converges to 1e-10 (true residual from
    the library's FP64 SpMV) in the iteration count of the single-system
    solver (+-1)."""
    if not os.path.exists(BATCH_EXE):
        pytest.skip("Fortran driver not built (no amdflang)")
    out = subprocess.run([BATCH_EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    lanes = [ln.split() for ln in out.stdout.splitlines() if ln.startswith("lane")]
    assert len(lanes) == 4
    config = [ln.split() for ln in out.stdout.splitlines() if ln.startswith("config")]
    assert config == [["config", "16", "8", "4", "1"]]
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
