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
from helpers import delta_apply, face_coeffs

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
    # the oracle keeps the reference's absolute |alpha| stop: solve systems
    # scaled to norm 1e8 so that test cannot fire before 1e-14
    S = 1e8
    nb, rhs = np.linalg.norm(b), x - x_obs
    xo = (nb / S) * O.pcg_jacobi(A, jac, b * (S / nb), x0_mode=O.X0_ZERO, eps=1e-14 * S,
                                 iter_max=20000, workers=4)["x"]
    nr = np.linalg.norm(rhs)
    lo = (nr / S) * O.pcg_jacobi(A, jac, rhs * (S / nr), x0_mode=O.X0_ZERO, eps=1e-14 * S,
                                 iter_max=20000, workers=4)["x"]
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-11 * np.abs(xo).max())
    np.testing.assert_allclose(lam, lo, rtol=0, atol=1e-10 * np.abs(lo).max())
    # same (x, lambda) -> same formulas -> same bits
    np.testing.assert_array_equal(dLda, O.sddmm_csr(A, lam, x))
    np.testing.assert_array_equal(g, O.diffusion7_kappa_grad(*shape, kappa, x, lam))


def perturbed_delta(M, shape, kappa, kappa_p, f0, r_base, x):
    """x(kappa_p) - x for (A + dA) x(kappa_p) = b, by the fixed point
    d <- A^-1 (r_base - dA (x + d)) with r_base = b - A x (contraction
    ~ |dA| / |A| ~ 1e-5: three sweeps reach rounding level).

    Why not solve with a matrix generated from kappa_p: a stored matrix
    rounds every diagonal entry (a sum of six face coefficients) to its own
    ulp, independently for kappa + h d and kappa - h d.  The signal of a
    face enters L only through (lambda_i - lambda_j)(x_i - x_j), which on a
    smooth 128^3 field is ~1e-4 of lambda_i x_i, so that rounding alone is a
    relative error of ~1e-16 / (h 1e-4) ~ 1e-6 in the central difference
    at h = 1e-6.  Applying dA from the coefficient changes keeps every
    rounding relative to the change itself."""
    d = np.zeros_like(x)
    for _ in range(4):
        rhs = r_base - delta_apply(shape, kappa, kappa_p, f0, x + d)
        new = M.pcg(rhs, tol=1e-14, tol_mode=1, x0_mode=1, iter_max=20000, alpha_stop=0)["x"]
        change = np.linalg.norm(new - d) / max(np.linalg.norm(new), 1e-300)
        d = new
        if change < 1e-15:
            break
    return d


def fd_along(M, shape, kappa, f0, r_base, x, x_obs, d, h):
    """Central difference of L along kappa * exp(t d) at t = 0, from the two
    perturbed solutions: L+ - L- = sum (x+ - x-)((x+ + x-)/2 - x_obs)."""
    dp = perturbed_delta(M, shape, kappa, kappa * np.exp(h * d), f0, r_base, x)
    dm = perturbed_delta(M, shape, kappa, kappa * np.exp(-h * d), f0, r_base, x)
    return float(np.dot(dp - dm, x + 0.5 * (dp + dm) - x_obs)) / (2 * h)


@pytest.mark.parametrize("shape,ndir", [((12, 10, 9), 8), ((128, 128, 128), 32)])
def test_gradient_central_differences(gpu, shape, ndir):
    """The differentiation workload: kappa = exp(g), g ~ N(0,1) seed 4,
    b ~ U[0,1], x_obs ~ N(0,1); forward solve in FP32 + FP64 residual
    replacement, adjoint, then 32 random directions checked by central
    differences (step 1e-6 in every log(kappa)) built from tightly
    converged FP64 solves of the perturbation itself (perturbed_delta)."""
    kappa, b, x_obs = problem(*shape, seed=4, gsd=1.0)
    M = gpu.Matrix.diffusion7(*shape, kappa)
    M.set_jacobi()
    fw = M.pcg(b, tol=1e-11, tol_mode=1, x0_mode=1, iter_max=20000, precision=1, alpha_stop=0,
               allow_not_converged=False)
    print(f"forward: {fw['iters']} inner iterations, {fw['outer_iters']} replacements", flush=True)
    x = fw["x"]
    # polish the mixed-precision solution with one FP64 correction solve
    x = x + M.pcg(b - M.spmv(x), tol=1e-12, tol_mode=2, x0_mode=1, iter_max=20000,
                  alpha_stop=0)["x"]
    lam, dLda, ast = M.adjoint(x, x - x_obs, tol=1e-14, iter_max=20000)
    print(f"adjoint: {ast['iters']} iterations, status {ast['status']}", flush=True)
    assert ast["status"] == 0
    g_kappa = gpu.kappa_grad(*shape, kappa, x, lam)
    f0 = face_coeffs(shape, kappa)
    r_base = b - M.spmv(x)
    rng = np.random.default_rng(99)
    h = 1e-6
    worst = 0.0
    for k in range(ndir):
        # every cell's log(kappa) moves by h * N(0,1): a unit-norm direction
        # over 2M cells would move each by ~h/1400, where the rounding of the
        # recomputed face coefficients (1 ulp) is a 1e-7 relative noise
        d = rng.normal(size=kappa.size)
        fd = fd_along(M, shape, kappa, f0, r_base, x, x_obs, d, h)
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
