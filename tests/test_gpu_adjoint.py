"""Adjoint of the PCG solve (cel_gpu_cg_adjoint) and the gradient kernels.

L(kappa, b) = 0.5 ||x - x_obs||^2 with A(kappa) x = b, A the variable-
coefficient 7-point diffusion matrix (harmonic-mean faces).  Checks:
  * lambda (= dL/db) against the FP64 oracle's own PCG solve,
  * dL/da_ij = -lambda_i x_j and dL/dkappa bit-exact with the oracle's
    restatement of the same formulas on the same (x, lambda),
  * directional derivatives against central finite differences (step 1e-6)
    along random directions, to 1e-6 relative.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

# REL_B, x0 = 0, norm test only (the reference's absolute |alpha| test would stop
# these tightly converged solves early)
TIGHT = dict(tol=1e-14, tol_mode=1, x0_mode=1, iter_max=20000, alpha_stop=0)


def problem(nx, ny, nz, seed, gsd):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    kappa = np.exp(rng.normal(0.0, gsd, n))
    b = rng.uniform(0.0, 1.0, n)
    x_obs = rng.normal(size=n)
    return kappa, b, x_obs


def loss_and_grad(gpu, shape, kappa, b, x_obs, precision=0):
    M = gpu.Matrix.diffusion7(*shape, kappa)
    M.set_jacobi()
    fw = M.pcg(b, precision=precision, **TIGHT)
    x = fw["x"]
    lam, dLda, st = M.adjoint(x, x - x_obs, tol=1e-14)
    g = gpu.kappa_grad(*shape, kappa, x, lam)
    return 0.5 * np.sum((x - x_obs) ** 2), x, lam, dLda, g, M


def test_adjoint_matches_oracle_small(gpu):
    shape = (10, 9, 8)
    kappa, b, x_obs = problem(*shape, seed=4, gsd=1.0)
    L, x, lam, dLda, g, M = loss_and_grad(gpu, shape, kappa, b, x_obs)
    A = O.gen_diffusion7(*shape, kappa)
    jac = O.set_jacobi(O.csr_diagonal(A))
    # the oracle keeps the reference's |alpha| stop: solve unit-norm systems
    nb, rhs = np.linalg.norm(b), x - x_obs
    xo = nb * O.pcg_jacobi(A, jac, b / nb, x0_mode=O.X0_ZERO, eps=1e-14, iter_max=20000,
                           workers=4)["x"]
    nr = np.linalg.norm(rhs)
    lo = nr * O.pcg_jacobi(A, jac, rhs / nr, x0_mode=O.X0_ZERO, eps=1e-14, iter_max=20000,
                           workers=4)["x"]
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-11 * np.abs(xo).max())
    np.testing.assert_allclose(lam, lo, rtol=0, atol=1e-10 * np.abs(lo).max())
    # same (x, lambda) -> same formulas -> same bits
    np.testing.assert_array_equal(dLda, O.sddmm_csr(A, lam, x))
    np.testing.assert_array_equal(g, O.diffusion7_kappa_grad(*shape, kappa, x, lam))


def perturbed_delta(gpu, shape, kappa_p, b, x):
    """x(kappa_p) - x by solving A(kappa_p) d = b - A(kappa_p) x to 1e-13
    relative to its own right-hand side (the difference is O(h), so it is
    solved for directly instead of subtracting two full solutions)."""
    M = gpu.Matrix.diffusion7(*shape, kappa_p)
    M.set_jacobi()
    r = b - M.spmv(x)
    return M.pcg(r, tol=1e-13, tol_mode=2, x0_mode=1, iter_max=20000, alpha_stop=0)["x"]


def fd_along(gpu, shape, kappa, b, x, x_obs, d, h):
    """Central difference of L along kappa * exp(t d) at t = 0, from the two
    perturbed solutions: L+ - L- = sum (x+ - x-)((x+ + x-)/2 - x_obs)."""
    dp = perturbed_delta(gpu, shape, kappa * np.exp(h * d), b, x)
    dm = perturbed_delta(gpu, shape, kappa * np.exp(-h * d), b, x)
    return float(np.dot(dp - dm, x + 0.5 * (dp + dm) - x_obs)) / (2 * h)


@pytest.mark.parametrize("shape,ndir", [((12, 10, 9), 8), ((128, 128, 128), 32)])
def test_gradient_central_differences(gpu, shape, ndir):
    """The differentiation workload: kappa = exp(g), g ~ N(0,1) seed 4,
    b ~ U[0,1], x_obs ~ N(0,1); forward solve in FP32 + FP64 residual
    replacement, adjoint, then 32 random directions (step 1e-6) checked by
    central differences built from tightly converged FP64 solves."""
    kappa, b, x_obs = problem(*shape, seed=4, gsd=1.0)
    M = gpu.Matrix.diffusion7(*shape, kappa)
    M.set_jacobi()
    fw = M.pcg(b, tol=1e-13, tol_mode=1, x0_mode=1, iter_max=20000, precision=1, alpha_stop=0)
    print(f"forward: {fw['iters']} inner iterations, {fw['outer_iters']} replacements", flush=True)
    x = fw["x"]
    # polish the mixed-precision solution with one FP64 correction solve
    x = x + M.pcg(b - M.spmv(x), tol=1e-12, tol_mode=2, x0_mode=1, iter_max=20000,
                  alpha_stop=0)["x"]
    lam, dLda, _ = M.adjoint(x, x - x_obs, tol=1e-14)
    g_kappa = gpu.kappa_grad(*shape, kappa, x, lam)
    rng = np.random.default_rng(99)
    h = 1e-6
    worst = 0.0
    for k in range(ndir):
        d = rng.normal(size=kappa.size)
        d /= np.linalg.norm(d)
        fd = fd_along(gpu, shape, kappa, b, x, x_obs, d, h)
        an = float(np.dot(g_kappa * kappa, d))  # chain rule through kappa * exp(t d)
        rel = abs(an - fd) / max(abs(fd), 1e-300)
        worst = max(worst, rel)
        print(f"dir {k}: analytic {an:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
    print(f"worst relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-6, worst
    # dL/db = lambda (L is quadratic in x, x linear in b: exact up to rounding)
    d = rng.normal(size=kappa.size)
    d /= np.linalg.norm(d)
    dx = M.pcg(d, tol=1e-13, tol_mode=1, x0_mode=1, iter_max=20000, alpha_stop=0)["x"]
    assert abs(np.dot(lam, d) - np.dot(dx, x - x_obs)) <= 1e-8 * abs(np.dot(dx, x - x_obs))
