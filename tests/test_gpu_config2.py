"""The config-2 workload at its full size against the oracle.

7-point Poisson on 256^3 (n = 16,777,216, nnz = 117,047,296) as one
subdomain, x_true ~ U[0,1) (numpy default_rng(1)), b = A x_true, x0 = 0,
stop at ||r|| <= 1e-10 ||b||.  The oracle's FP64 solve of the same system
(reference op order, OpenMP workers on the host cores) runs once for the
module.

Bars:
  * FP64, CSR and JAD: iteration count within +-1 of the oracle's, residual
    history within 1e-12 relative, solution within 1e-12 max|x| of the
    oracle's (the two differ only in the summation order of the dot
    products);
  * the DIA form: bitwise the CSR solve (FP64 and mixed);
  * the one-subdomain halo model (six halo faces of 65,536 entries, the six
    send faces packed every SpMV, received values zero): bitwise the local
    solve of the same handle, CSR and DIA, which matches the oracle;
  * mixed precision (FP32 inner iteration, FP64 residual replacement), CSR
    and JAD: FP64 true residual <= 1e-10 ||b||, and the solution no further
    from x_true, nor from the oracle's FP64 solution, than twice the
    oracle's own error (the two solves stop at the same residual along
    different rounding trajectories: at 1e-10 both are about 2.5e-8 from
    x_true, and their distance from each other is of that size, 3e-10 to
    3.2e-8 over seeds 1-4 and replacement settings 0.03-0.5 at 128^3 and
    256^3).
"""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

N = 256
TOL = 1e-10
HIST_RTOL = 1e-12


def _workers():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 2
    return max(1, min(cores, 16) - 1)


@pytest.fixture(scope="module")
def c2(gpu):
    A = O.gen_poisson7(N, N, N)
    assert A.n == N ** 3 and A.nnz == 117047296
    xt = np.random.default_rng(1).random(A.n)
    w = _workers()
    b = O.spmv_csr(A, xt, workers=w)
    jac = O.set_jacobi(O.csr_diagonal(A))
    ref = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=TOL * float(np.linalg.norm(b)),
                       iter_max=20000, workers=w)
    print(f"oracle: {ref['iters']} iterations ({w} workers)", flush=True)
    del A, jac
    M = gpu.Matrix.poisson7(N, N, N)
    M.set_jacobi()
    d_b = gpu.DeviceArray.from_host(b)
    # the device generator's matrix and product are the oracle's
    y = gpu.DeviceArray(M.n)
    M.spmv(gpu.DeviceArray.from_host(xt), y)
    np.testing.assert_array_equal(y.to_host(), b)
    return dict(M=M, b=b, d_b=d_b, xt=xt, ref=ref)


def _solve(gpu, c, fmt, precision=None, **kw):
    precision = gpu.PREC_FP64 if precision is None else precision
    return c["M"].pcg(c["d_b"], tol=TOL, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO,
                      iter_max=20000, precision=precision, fmt=fmt, allow_not_converged=False,
                      **kw)


def _check_fp64(c, got, name):
    ref = c["ref"]
    assert abs(got["iters"] - ref["iters"]) <= 1, (got["iters"], ref["iters"])
    m = min(got["iters"], ref["iters"]) + 1
    rel = np.abs(got["hist"][:m] - ref["hist"][:m]) / ref["hist"][:m]
    x = got["x"].to_host()
    dx = np.abs(x - ref["x"]).max() / np.abs(ref["x"]).max()
    print(f"{name}: {got['iters']} iterations (oracle {ref['iters']}), {got['solve_ms']:.1f} ms, "
          f"history {rel.max():.2e}, max|x - x_oracle|/max|x| {dx:.2e}", flush=True)
    assert rel.max() <= HIST_RTOL, (rel.max(), int(np.argmax(rel)))
    assert dx <= 1e-12
    return x


def test_fp64_csr_matches_oracle(gpu, c2):
    got = _solve(gpu, c2, gpu.FMT_CSR, history=True)
    c2["x_csr"] = _check_fp64(c2, got, "fp64 csr")


def test_fp64_jad_matches_oracle(gpu, c2):
    M = c2["M"]
    M.to_jad()
    got = _solve(gpu, c2, gpu.FMT_JAD, history=True)
    _check_fp64(c2, got, "fp64 jad")


def _face_blocks(M):
    """Halo blocks of the centre rank of a 3x3x3 decomposition (neighbour
    ranks 4, 10, 12, 14, 16, 22: z-, y-, x-, x+, y+, z+), each row of a face
    coupled to the facing cell of the neighbour, and the six send lists."""
    idx = np.arange(N * N, dtype=np.int64)
    a, bb = idx % N, idx // N

    def cell(x, y, z):
        return x + N * (y + N * z) + 1

    faces = [(4, cell(a, bb, 0), cell(a, bb, N - 1)), (10, cell(a, 0, bb), cell(a, N - 1, bb)),
             (12, cell(0, a, bb), cell(N - 1, a, bb)), (14, cell(N - 1, a, bb), cell(0, a, bb)),
             (16, cell(a, N - 1, bb), cell(a, 0, bb)), (22, cell(a, bb, N - 1), cell(a, bb, 0))]
    for proc, rows, cols in faces:
        order = np.argsort(rows, kind="stable")
        rows, cols = rows[order].astype(np.int32), cols[order].astype(np.int32)
        M.add_halo_block(proc, proc * N ** 3 + 1, N ** 3, rows, cols, -np.ones(rows.size))
        M.set_send_list(proc, rows)


def test_dia_and_halo_forms_are_bitwise_the_csr_solve(gpu, c2):
    M = c2["M"]
    M.to_dia()
    csr = _solve(gpu, c2, gpu.FMT_CSR)
    dia = _solve(gpu, c2, gpu.FMT_DIA)
    x = csr["x"].to_host()
    assert csr["iters"] == dia["iters"] == c2["ref"]["iters"]
    np.testing.assert_array_equal(dia["x"].to_host(), x)
    # one subdomain with its six halo faces; received values are zero, so
    # the halo solve is the local solve of the same handle (same launch
    # shapes) bit for bit, and that one matches the oracle
    H = gpu.Matrix.poisson7(N, N, N)
    H.set_jacobi()
    _face_blocks(H)
    assert H.ghost_len() == (6 * N * N, 6 * N * N)
    H.to_dia()
    kw = dict(tol=TOL, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=20000,
              allow_not_converged=False)
    loc = H.pcg(c2["d_b"], fmt=gpu.FMT_CSR, history=True, **kw)
    xl = _check_fp64(c2, loc, "local solve of the handle with halo blocks")
    for f, name in ((gpu.FMT_CSR, "csr"), (gpu.FMT_DIA, "dia")):
        h = H.pcg(c2["d_b"], fmt=f, use_halo=1, **kw)
        print(f"halo {name}: {h['iters']} iterations, {h['solve_ms']:.1f} ms", flush=True)
        assert h["iters"] == loc["iters"]
        np.testing.assert_array_equal(h["x"].to_host(), xl)
    H.destroy()
    # mixed: the DIA form is bitwise the CSR form
    mc = _solve(gpu, c2, gpu.FMT_CSR, gpu.PREC_MIXED)
    md = _solve(gpu, c2, gpu.FMT_DIA, gpu.PREC_MIXED)
    assert (mc["iters"], mc["outer_iters"]) == (md["iters"], md["outer_iters"])
    np.testing.assert_array_equal(mc["x"].to_host(), md["x"].to_host())


@pytest.mark.parametrize("fmt", ["csr", "jad"])
def test_mixed_reaches_tolerance_near_the_oracle(gpu, c2, fmt):
    M = c2["M"]
    f = gpu.FMT_CSR
    if fmt == "jad":
        M.to_jad()
        f = gpu.FMT_JAD
    got = _solve(gpu, c2, f, gpu.PREC_MIXED)
    res = gpu.DeviceArray(M.n)
    M.spmv(got["x"], res)
    b = c2["b"]
    true_rel = np.linalg.norm(b - res.to_host()) / np.linalg.norm(b)
    x = got["x"].to_host()
    xo = c2["ref"]["x"]
    dx = np.abs(x - xo).max() / np.abs(xo).max()
    print(f"mixed {fmt}: {got['iters']} inner iterations, {got['outer_iters']} replacements, "
          f"{got['solve_ms']:.1f} ms, true rel residual {true_rel:.3e}, "
          f"max|x - x_oracle|/max|x| {dx:.3e}, max|x - x_true| {np.abs(x - c2['xt']).max():.3e}",
          flush=True)
    assert true_rel <= TOL
    eo = np.abs(xo - c2["xt"]).max()
    assert np.abs(x - c2["xt"]).max() <= 2.0 * eo
    assert np.abs(x - xo).max() <= 2.0 * eo


def test_fused_product_of_the_six_face_block_with_ghost_values(gpu):
    """The 256^3 block with its six halo faces (face, edge and corner rows:
    1, 2 and 3 halo entries) and random received values: the solve's fused
    product (cel_gpu_spmv_dot, CSR and DIA) is bitwise the distributed
    product (cel_gpu_spmv_distr, pinned to the oracle in test_gpu_distr),
    and its x.y agrees with numpy's to 1e-13 of sum |x_i y_i|."""
    H = gpu.Matrix.poisson7(N, N, N)
    _face_blocks(H)
    H.to_dia()
    rng = np.random.default_rng(5)
    x = rng.normal(size=H.n)
    ghost = rng.normal(size=6 * N * N)
    y = H.spmv_distr(x, ghost)
    tol = 1e-13 * float(np.abs(x * y).sum())
    for f in (gpu.FMT_CSR, gpu.FMT_DIA):
        yf, dot = H.spmv_dot(x, ghost, fmt=f)
        np.testing.assert_array_equal(yf, y)
        assert abs(dot - float(np.dot(x, y))) <= tol
    H.destroy()
