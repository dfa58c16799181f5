"""GPU Jacobi-PCG against the reference-order CPU restatement.

Bar (north star): same iteration count within +-1 and residual histories
within 1e-12 relative in FP64; the only difference between the two is the
summation order of the dot products (deterministic tree on the GPU, worker
chunks on the CPU).  Mixed precision is checked against the FP64 answer.
"""
import numpy as np
import pytest

import oracle as O
from helpers import stencil27

pytestmark = pytest.mark.gpu

HIST_RTOL = 1e-12


def problem(N, seed, dist="uniform", shape=None):
    shape = shape or (N, N, N)
    A = O.gen_poisson7(*shape)
    rng = np.random.default_rng(seed)
    xt = rng.random(A.n) if dist == "uniform" else rng.normal(size=A.n)
    b = O.spmv_csr(A, xt)
    jac = O.set_jacobi(O.csr_diagonal(A))
    return A, xt, b, jac


def compare(ref, got, n_hist=None):
    assert abs(got["iters"] - ref["iters"]) <= 1, (got["iters"], ref["iters"])
    m = min(got["iters"], ref["iters"]) + 1
    h_ref, h_got = ref["hist"][:m], got["hist"][:m]
    rel = np.abs(h_got - h_ref) / h_ref
    assert rel.max() <= HIST_RTOL, (rel.max(), int(np.argmax(rel)))
    return rel.max()


@pytest.mark.parametrize("fmt", ["csr", "jad", "coo"])
def test_config1_poisson32_parity(gpu, fmt):
    """32^3, x_true ~ U[0,1) seed 0, x0 = 0, relative tolerance 1e-8,
    iter_max 1000, reference with 4 threads (3 workers).  COO applies the
    block entry by entry (v(row) += a p from v = 0), which gives the CSR
    solve's bits."""
    A, xt, b, jac = problem(32, 0)
    eps = 1e-8 * np.linalg.norm(b)
    ref = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=eps, iter_max=1000, workers=3)
    M = gpu.Matrix.from_csr(A.xadj, A.cols, A.vals)
    M.set_jacobi()
    f = gpu.FMT_CSR
    if fmt == "jad":
        M.to_jad()
        f = gpu.FMT_JAD
    elif fmt == "coo":
        f = gpu.FMT_COO
    kw = dict(tol=1e-8, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=1000, history=True)
    got = M.pcg(b, fmt=f, **kw)
    print(f"iters gpu {got['iters']} ref {ref['iters']}", flush=True)
    assert got["status"] == 0
    rel = compare(ref, got)
    print(f"max history rel diff {rel:.3e}")
    np.testing.assert_allclose(got["x"], ref["x"], rtol=0, atol=1e-9)
    assert np.abs(got["x"] - xt).max() < 1e-6
    assert got["final_norm"] <= got["eps_abs"]
    if fmt == "coo":
        csr = M.pcg(b, fmt=gpu.FMT_CSR, **kw)
        np.testing.assert_array_equal(got["hist"], csr["hist"])
        np.testing.assert_array_equal(got["x"], csr["x"])
        mc = M.pcg(b, fmt=f, precision=gpu.PREC_MIXED, **kw)
        mr = M.pcg(b, fmt=gpu.FMT_CSR, precision=gpu.PREC_MIXED, **kw)
        np.testing.assert_array_equal(mc["x"], mr["x"])


def test_reference_defaults_x0_one_absolute_eps(gpu):
    """The reference's own settings: x0 = 1, absolute eps = 1e-8."""
    A, xt, b, jac = problem(20, 7, shape=(20, 18, 16))
    ref = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ONE, eps=1e-8, iter_max=1000, workers=3)
    M = gpu.Matrix.from_csr(A.xadj, A.cols, A.vals)
    M.set_jacobi()
    got = M.pcg(b, history=True)  # defaults: x0 = 1, TOL_ABS, 1e-8
    compare(ref, got)
    np.testing.assert_allclose(got["x"], ref["x"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("graph", [0, 1])
def test_poisson64_tol_1e10_graph_and_stream(gpu, graph):
    A, xt, b, jac = problem(64, 1)
    eps = 1e-10 * np.linalg.norm(b)
    ref = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=eps, iter_max=2000, workers=8)
    M = gpu.Matrix.poisson7(64, 64, 64)
    M.set_jacobi()
    got = M.pcg(b, tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=2000,
                history=True, use_graph=graph)
    compare(ref, got)
    assert np.abs(got["x"] - xt).max() < 1e-7


def test_variable_coefficient_parity(gpu):
    nx, ny, nz = 24, 20, 18
    k = np.exp(np.random.default_rng(3).normal(0, 0.5, nx * ny * nz))
    A = O.gen_diffusion7(nx, ny, nz, k)
    b = 1e3 * np.random.default_rng(4).uniform(-1, 1, A.n)
    jac = O.set_jacobi(O.csr_diagonal(A))
    eps = 1e-9 * np.linalg.norm(b)
    ref = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=eps, iter_max=3000, workers=4)
    M = gpu.Matrix.diffusion7(nx, ny, nz, k)
    M.set_jacobi()
    got = M.pcg(b, tol=1e-9, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=3000, history=True)
    compare(ref, got)


def test_row_scaled_solve_matches_oracle(gpu):
    """MATRIX_SCALER_DIAGONAL path: A and b divided row-wise by sqrt|d|,
    jacobi = sqrt|d|/d (cel_sp_mat_distr_format_module.f90:91-203)."""
    nx, ny, nz = 12, 10, 8
    k = np.exp(np.random.default_rng(8).normal(0, 1, nx * ny * nz))
    A = O.gen_diffusion7(nx, ny, nz, k)
    d = O.csr_diagonal(A)
    S = O.scale_csr(A, d)
    b = 1e2 * np.random.default_rng(9).uniform(0, 1, A.n)
    bs = O.scale_vec(b, d)
    jac = O.set_jacobi(d, scaled=True)
    ref = O.pcg_jacobi(S, jac, bs, x0_mode=O.X0_ONE, eps=1e-8, iter_max=2000, workers=3)
    M = gpu.Matrix.diffusion7(nx, ny, nz, k)
    M.set_jacobi(row_scaled=True)
    M.scale_diagonal()
    got = M.pcg(M.scale_vec(b), tol=1e-8, history=True, iter_max=2000)
    compare(ref, got)


def test_iteration_cap_reports_not_converged(gpu):
    A, xt, b, jac = problem(16, 2)
    ref = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=0.0, iter_max=7, workers=1)
    M = gpu.Matrix.poisson7(16, 16, 16)
    M.set_jacobi()
    got = M.pcg(b, tol=0.0, x0_mode=gpu.X0_ZERO, iter_max=7, history=True, check_every=3)
    assert got["status"] == 1 and got["iters"] == 7 == ref["iters"]
    compare(ref, got)


def test_nonfinite_rhs_is_reported(gpu):
    M = gpu.Matrix.poisson7(8, 8, 8)
    M.set_jacobi()
    b = np.ones(M.n)
    b[17] = np.nan
    with pytest.raises(gpu.CelError) as e:
        M.pcg(b)
    assert e.value.code == -7


def test_indefinite_matrix_breakdown(gpu):
    A = O.gen_poisson7(6, 6, 6)
    # half of the diagonal made strongly negative: an indefinite operator
    vals2 = A.vals.copy()
    diag_pos = [j for i in range(A.n) for j in range(A.xadj[i] - 1, A.xadj[i + 1] - 1)
                if A.cols[j] == i + 1 and i % 2 == 0]
    vals2[diag_pos] = -30.0
    M2 = gpu.Matrix.from_csr(A.xadj, A.cols, vals2)
    M2.set_jacobi()
    b = np.random.default_rng(0).normal(size=A.n)
    with pytest.raises(gpu.CelError) as e:
        M2.pcg(b, tol=1e-12, x0_mode=gpu.X0_ZERO, iter_max=500)
    assert e.value.code == -6


@pytest.mark.parametrize("fmt", ["csr", "jad"])
def test_mixed_precision_reaches_fp64_accuracy(gpu, fmt):
    """FP32 inner iteration with FP64 residual replacement reaches 1e-10
    relative residual (checked by the FP64 oracle SpMV) and the FP64
    solution to 1e-8."""
    A, xt, b, jac = problem(48, 2, dist="normal")
    M = gpu.Matrix.poisson7(48, 48, 48)
    M.set_jacobi()
    f = gpu.FMT_CSR
    if fmt == "jad":
        M.to_jad()
        f = gpu.FMT_JAD
    got = M.pcg(b, tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=5000,
                precision=gpu.PREC_MIXED, fmt=f)
    assert got["status"] == 0
    x = got["x"]
    true_rel = np.linalg.norm(b - O.spmv_csr(A, x)) / np.linalg.norm(b)
    ref = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=1e-10 * np.linalg.norm(b), iter_max=5000,
                       workers=8)
    print(f"mixed iters {got['iters']} outer {got['outer_iters']} fp64 iters {ref['iters']} "
          f"true rel {true_rel:.3e}", flush=True)
    assert true_rel <= 1e-10
    assert np.abs(x - ref["x"]).max() <= 1e-8 * max(1.0, np.abs(ref["x"]).max())
    assert np.abs(x - xt).max() <= 1e-7


def test_mixed_restart_variant_converges(gpu):
    A, xt, b, jac = problem(32, 5, dist="normal")
    M = gpu.Matrix.poisson7(32, 32, 32)
    M.set_jacobi()
    got = M.pcg(b, tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=5000,
                precision=gpu.PREC_MIXED, ru_restart=1, ru_delta=1e-4)
    true_rel = np.linalg.norm(b - O.spmv_csr(A, got["x"])) / np.linalg.norm(b)
    assert got["status"] == 0 and true_rel <= 1e-10


def test_mixed_stops_at_attainable_accuracy(gpu):
    """A tolerance below what FP64 residuals can show for this
    variable-coefficient system ends as NOT_CONVERGED with a finite,
    accurate x (two replacement cycles that fail to halve the FP64
    residual), not as an FP32 overflow."""
    n = 64
    rng = np.random.default_rng(4)
    k = np.exp(rng.normal(0, 1, n ** 3))
    b = rng.uniform(0, 1, n ** 3)
    M = gpu.Matrix.diffusion7(n, n, n, k)
    M.set_jacobi()
    r = M.pcg(b, tol=1e-15, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=20000,
              precision=gpu.PREC_MIXED, alpha_stop=0)
    assert r["status"] == 1 and r["iters"] < 20000
    assert np.all(np.isfinite(r["x"]))
    A = O.gen_diffusion7(n, n, n, k)
    assert np.linalg.norm(b - O.spmv_csr(A, r["x"])) <= 1e-12 * np.linalg.norm(b)


def test_device_resident_solve_and_repeat(gpu):
    """Inputs resident in HBM; the same handle solves twice (ReDoIt)."""
    A, xt, b, jac = problem(40, 6)
    M = gpu.Matrix.poisson7(40, 40, 40)
    M.set_jacobi()
    db = gpu.DeviceArray.from_host(b)
    r1 = M.pcg(db, tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=3000)
    r2 = M.pcg(db, tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=3000)
    x1, x2 = r1["x"].to_host(), r2["x"].to_host()
    np.testing.assert_array_equal(x1, x2)  # deterministic reductions
    assert r1["iters"] == r2["iters"]
    assert np.abs(x1 - xt).max() < 1e-7


def test_profile_counts_spmv_launches(gpu):
    M = gpu.Matrix.poisson7(24, 24, 24)
    M.set_jacobi()
    b = np.random.default_rng(0).random(M.n)
    r = M.pcg(b, tol=1e-8, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, profile=1)
    assert r["spmv_launches"] >= r["iters"] and r["spmv_ms"] > 0


def test_config3_mixed_384(gpu):
    """The precision workload: 7-point Poisson on 384^3 (n = 56,623,104,
    nnz = 395,476,992), x_true ~ N(0,1) (numpy default_rng(2)), b = A x_true
    (FP64 device SpMV), x0 = 0, FP32 inner iteration with FP64 residual
    replacement to 1e-10 relative residual.  Checked at full size by what
    does not need a CPU solve: the FP64 true residual, the error against the
    plain FP64 GPU solve of the same system (the FP64 path is pinned against
    the oracle at the smaller sizes), and the error against x_true."""
    n = 384
    M = gpu.Matrix.poisson7(n, n, n)
    M.set_jacobi()
    assert M.nnz == 395476992
    xt = np.random.default_rng(2).normal(size=M.n)
    d_xt = gpu.DeviceArray.from_host(xt)
    d_b = gpu.DeviceArray(M.n)
    M.spmv(d_xt, d_b)
    b = d_b.to_host()
    kw = dict(tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=20000)
    mixed = M.pcg(d_b, precision=gpu.PREC_MIXED, allow_not_converged=False, **kw)
    fp64 = M.pcg(d_b, precision=gpu.PREC_FP64, allow_not_converged=False, **kw)
    xm, xf = mixed["x"].to_host(), fp64["x"].to_host()
    res = gpu.DeviceArray(M.n)
    M.spmv(mixed["x"], res)
    true_rel = np.linalg.norm(b - res.to_host()) / np.linalg.norm(b)
    d_fp64 = np.abs(xm - xf).max() / np.abs(xf).max()
    err = np.abs(xm - xt).max()
    print(f"384^3 mixed {mixed['solve_ms']:.1f} ms ({mixed['iters']} inner, {mixed['outer_iters']} "
          f"replacements), fp64 {fp64['solve_ms']:.1f} ms ({fp64['iters']} iterations); true rel "
          f"{true_rel:.3e}, max|x_mixed - x_fp64|/max|x| {d_fp64:.3e}, max err vs x_true {err:.3e}",
          flush=True)
    assert true_rel <= 1e-10
    assert d_fp64 <= 1e-8  # measured 3.4e-11 (seed 2)
    assert abs(mixed["iters"] - fp64["iters"]) <= 0.05 * fp64["iters"]


def test_zero_rhs_returns_at_once(gpu):
    """b = 0, x0 = 0: the initial residual is 0, so the loop does not run
    (the reference's norm2 > eps test with eps = tol * ||b|| = 0)."""
    M = gpu.Matrix.poisson7(9, 8, 7)
    M.set_jacobi()
    out = M.pcg(np.zeros(M.n), tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO)
    assert out["status"] == 0 and out["iters"] == 0
    np.testing.assert_array_equal(out["x"], np.zeros(M.n))


def test_single_row_and_empty_systems(gpu):
    """1 x 1: one iteration gives x = b / a exactly (Jacobi preconditioning
    makes p = r / a, sigma = 1).  0 x 0: nothing to do, status 0."""
    M = gpu.Matrix.from_csr([1, 2], [1], [4.0])
    M.set_jacobi()
    out = M.pcg(np.array([2.0]), tol=1e-12, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO)
    assert out["status"] == 0 and out["iters"] == 1
    assert out["x"][0] == 0.5
    E = gpu.Matrix.from_csr([1], [], [])
    assert E.n == 0
    E.set_jacobi()
    out = E.pcg(np.zeros(0), tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO)
    assert out["status"] == 0 and out["iters"] == 0


def test_spmv_block_without_columns(gpu):
    """A block with rows but no columns (a rank whose rows couple only to
    neighbours): y = 0, or y unchanged when accumulating."""
    M = gpu.Matrix.from_csr([1, 1, 1], [], [], ncols=0)
    np.testing.assert_array_equal(M.spmv(np.zeros(0)), np.zeros(2))
    y0 = np.array([3.0, -1.0])
    np.testing.assert_array_equal(M.spmv(np.zeros(0), y0, accumulate=True), y0)


@pytest.mark.parametrize("fmt", ["csr", "jad", "dia"])
def test_rows_beyond_the_lds_stage_parity(gpu, fmt):
    """27-point rows (up to 27 entries): every 256-row tile holds more than
    the 2048 entries the K1 stages in LDS, so the CSR K1 reads them from HBM;
    JAD has 27 diagonals and the DIA form 27 diagonals with 32-bit row masks.
    FP64 against the oracle's history; DIA bitwise the CSR solve, FP64 and
    mixed."""
    A = stencil27(14, 13, 12)
    per_row = np.diff(A.indptr)
    assert min(per_row[i:i + 256].sum() for i in range(0, A.shape[0] - 255, 256)) > 2048
    H = O.CSR(A.indptr + 1, A.indices + 1, A.data)
    b = 1e3 * np.random.default_rng(31).uniform(-1.0, 1.0, H.n)
    jac = O.set_jacobi(O.csr_diagonal(H))
    ref = O.pcg_jacobi(H, jac, b, x0_mode=O.X0_ZERO, eps=1e-10 * np.linalg.norm(b), iter_max=2000,
                       workers=3)
    M = gpu.Matrix.from_csr(H.xadj, H.cols, H.vals)
    M.set_jacobi()
    kw = dict(tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=2000)
    f = gpu.FMT_CSR
    if fmt == "jad":
        M.to_jad()
        f = gpu.FMT_JAD
    elif fmt == "dia":
        assert len(M.to_dia()) == 27
        f = gpu.FMT_DIA
    got = M.pcg(b, fmt=f, history=True, **kw)
    print(f"{fmt}: iters gpu {got['iters']} ref {ref['iters']}", flush=True)
    assert got["status"] == 0
    compare(ref, got)
    np.testing.assert_allclose(got["x"], ref["x"], rtol=0, atol=1e-10 * np.abs(ref["x"]).max())
    if fmt == "dia":
        for prec in (gpu.PREC_FP64, gpu.PREC_MIXED):
            a = M.pcg(b, precision=prec, **kw)
            d = M.pcg(b, fmt=gpu.FMT_DIA, precision=prec, **kw)
            assert a["status"] == d["status"] == 0
            assert (a["iters"], a["outer_iters"]) == (d["iters"], d["outer_iters"])
            np.testing.assert_array_equal(a["x"], d["x"])
