"""GPU Jacobi-GMRES (cel_gpu_gmres_solve) against the oracle's restatement
of cel_gmresalg_2jacobi (src/fcgalg/cel_gmresalg_module.f90:570-851), and the
element-solver C ABI (cel_gpu_elm_*) against the oracle pipeline of
cel_elmSolverMatrixCompleted/DoIt/ReDoIt
(src/felmsolver/cel_elmsolver_module.f90:193-788).

The SpMV and every element-wise step match the oracle bit for bit; the dot
products group their sums differently (fixed workgroup chunks vs worker
chunks), so iteration and cycle counts must agree and the solutions agree
to rounding."""
import numpy as np
import pytest

import oracle as O
from helpers import stencil27
from test_oracle import convdiff_csr

pytestmark = pytest.mark.gpu


def matrices():
    rng = np.random.default_rng(51)
    kappa = np.exp(rng.normal(0.0, 0.5, 9 * 9 * 9))
    return [("poisson10", O.gen_poisson7(10, 10, 10)),
            ("convdiff12x10x8", convdiff_csr(12, 10, 8)),
            ("diffusion9", O.gen_diffusion7(9, 9, 9, kappa))]


def dev(gpu, A):
    M = gpu.Matrix.from_csr(A.xadj, A.cols, A.vals)
    M.set_jacobi()
    return M


@pytest.mark.parametrize("m", [1, 3, 5])
@pytest.mark.parametrize("name,A", matrices(), ids=[c[0] for c in matrices()])
def test_gmres_matches_oracle(gpu, name, A, m):
    print(f"case {name} m={m}", flush=True)
    b = np.random.default_rng(52).random(A.n)
    jac = O.set_jacobi(O.csr_diagonal(A))
    ref = O.gmres_jacobi(A, jac, b, eps=1e-8, m=m, iter_max=5000)
    out = dev(gpu, A).gmres(b, eps=1e-8, restart=m, iter_max=5000)
    assert out["status"] == 0 and ref["reason"] == 2
    # a stop test met within rounding of eps may fall one cycle either way
    assert abs(out["iters"] - ref["iters"]) <= 1 and abs(out["cycles"] - ref["cycles"]) <= m + 1, \
        (out["iters"], out["cycles"], ref["iters"], ref["cycles"])
    if out["cycles"] == ref["cycles"]:
        scale = np.abs(ref["x"]).max()
        np.testing.assert_allclose(out["x"], ref["x"], rtol=0, atol=1e-10 * scale)
        assert abs(out["final_norm"] - ref["norm"]) <= 1e-10 * np.linalg.norm(b)
    else:
        assert out["final_norm"] <= 2.0 * ref["norm"] and ref["norm"] <= 2.0 * out["final_norm"]


def test_gmres_device_arrays_jad_and_repeat_are_bitwise(gpu):
    A = convdiff_csr(11, 9, 7)
    b = np.random.default_rng(53).random(A.n)
    M = dev(gpu, A)
    a = M.gmres(b)
    c = M.gmres(b)
    np.testing.assert_array_equal(a["x"], c["x"])
    d_b = gpu.DeviceArray.from_host(b)
    d = M.gmres(d_b)
    np.testing.assert_array_equal(d["x"].to_host(), a["x"])
    M.to_jad()
    j = M.gmres(b, fmt=gpu.FMT_JAD)
    np.testing.assert_array_equal(j["x"], a["x"])
    assert j["iters"] == a["iters"] and j["cycles"] == a["cycles"]


@pytest.mark.parametrize("kind,shape", [("convdiff", (10, 10, 10)), ("convdiff", (24, 22, 20)),
                                        ("stencil27", (14, 13, 12))])
def test_gmres_csr_jad_dia_are_bitwise_and_match_oracle(gpu, kind, shape):
    """The CSR and DIA products form the partials of v_1 . w and r . r
    themselves, over the GMRES row chunks; JAD runs its product and a
    separate dot kernel over the same chunks: the three forms give the same
    bits.  24x22x20: 10560 rows in 6 chunks of 1792 (the last one partial),
    so a fused DIA step (2048 rows) covers rows of the next chunk that it
    must leave out.  The 27-point matrix (2184 rows, 2 chunks) has tiles of
    more than 2048 entries (the fused CSR product reads them from HBM) and
    27 diagonals (32-bit DIA row masks)."""
    if kind == "stencil27":
        S27 = stencil27(*shape)
        A = O.CSR(S27.indptr + 1, S27.indices + 1, S27.data)
    else:
        A = convdiff_csr(*shape)
    b = np.random.default_rng(57).random(A.n)
    M = dev(gpu, A)
    kw = dict(eps=1e-8, restart=3, iter_max=5000)
    c = M.gmres(b, **kw)
    M.to_jad()
    j = M.gmres(b, fmt=gpu.FMT_JAD, **kw)
    M.to_dia()
    d = M.gmres(b, fmt=gpu.FMT_DIA, **kw)
    for o in (j, d):
        assert (o["status"], o["iters"], o["cycles"]) == (c["status"], c["iters"], c["cycles"])
        np.testing.assert_array_equal(o["x"], c["x"])
        assert o["final_norm"] == c["final_norm"]
    jac = O.set_jacobi(O.csr_diagonal(A))
    ref = O.gmres_jacobi(A, jac, b, eps=1e-8, m=3, iter_max=5000)
    print(f"{shape}: gpu {c['iters']}/{c['cycles']} oracle {ref['iters']}/{ref['cycles']}", flush=True)
    assert c["status"] == 0 and ref["reason"] == 2
    assert abs(c["iters"] - ref["iters"]) <= 1 and abs(c["cycles"] - ref["cycles"]) <= 4
    if c["cycles"] == ref["cycles"]:
        np.testing.assert_allclose(c["x"], ref["x"], rtol=0, atol=1e-10 * np.abs(ref["x"]).max())


def test_gmres_iteration_cap_and_errors(gpu):
    A = O.gen_poisson7(12, 12, 12)
    b = np.random.default_rng(54).random(A.n)
    M = dev(gpu, A)
    out = M.gmres(b, eps=1e-14, iter_max=0, restart=3)
    assert out["status"] == 1 and out["iters"] == 1 and out["cycles"] == 4
    jac = O.set_jacobi(O.csr_diagonal(A))
    ref = O.gmres_jacobi(A, jac, b, eps=1e-14, m=3, iter_max=0)
    assert ref["reason"] != 2 and ref["cycles"] == 4
    np.testing.assert_allclose(out["x"], ref["x"], rtol=0, atol=1e-11 * np.abs(ref["x"]).max())
    with pytest.raises(gpu.CelError):
        M.gmres(b, restart=0)
    with pytest.raises(gpu.CelError):
        M.gmres(b, restart=33)
    N = gpu.Matrix.from_csr(A.xadj, A.cols, A.vals)  # no Jacobi vector
    with pytest.raises(gpu.CelError) as e:
        N.gmres(b)
    assert e.value.code == -10


def test_gmres_128_true_residual(gpu):
    """The element solver's settings on a 128^3 Poisson block: the reported
    norm is the true residual of the returned x (checked with the
    library's FP64 SpMV)."""
    import time

    M = gpu.Matrix.poisson7(128, 128, 128)
    M.set_jacobi()
    b = np.random.default_rng(55).random(M.n)
    t0 = time.perf_counter()
    out = M.gmres(b)
    wall = time.perf_counter() - t0
    r = b - M.spmv(out["x"])
    print(f"128^3 GMRES(3): {out['iters']} iterations, {out['cycles']} cycles, "
          f"{out['solve_ms']:.1f} ms device, {wall:.2f} s wall, |r| {out['final_norm']:.3e}", flush=True)
    assert out["status"] == 0
    assert abs(np.linalg.norm(r) - out["final_norm"]) <= 1e-12 * np.linalg.norm(b)


# ---------------------------------------------------------------- element solver API
def oracle_elm(rows, cols, vals, n, b, solver="gmres", scale_b=True):
    """MatrixCompleted + DoIt (or ReDoIt with scale_b False) on the oracle."""
    st, r, c, v = O.coo_sort_cel(n, rows, cols, vals)
    assert st == 0
    A = O.CSR(O.coo_to_csr(n, r), c, v)
    d = O.csr_diagonal(A)
    S = O.scale_csr(A, d)
    jac = O.set_jacobi(d, scaled=True)
    bb = O.scale_vec(b, d) if scale_b else np.asarray(b, np.float64)
    if solver == "gmres":
        out = O.gmres_jacobi(S, jac, bb, x0_mode=O.X0_ONE, eps=1e-8, iter_max=5000, m=3)
    else:
        out = O.pcg_jacobi(S, jac, bb, x0_mode=O.X0_ONE, eps=1e-8, iter_max=5000, workers=3)
    return S, out


def shuffled_rows(A, rng):
    """Entries row by row, each row's entries in random order."""
    rows, cols, vals = [], [], []
    for i in range(A.n):
        lo, hi = A.xadj[i] - 1, A.xadj[i + 1] - 1
        order = rng.permutation(hi - lo)
        rows += [i + 1] * (hi - lo)
        cols += list(A.cols[lo:hi][order])
        vals += list(A.vals[lo:hi][order])
    return np.array(rows, np.int32), np.array(cols, np.int32), np.array(vals)


def test_elm_gmres_matches_oracle_pipeline(gpu):
    A = convdiff_csr(8, 7, 6)
    rng = np.random.default_rng(56)
    rows, cols, vals = shuffled_rows(A, rng)
    b = rng.random(A.n)
    E = gpu.Elm(A.n)
    E.insert_rows(rows, cols, vals)
    E.matrix_completed()
    S, ref = oracle_elm(rows, cols, vals, A.n, b)
    xs, cs, vs = E.matrix_csr()
    np.testing.assert_array_equal(xs, S.xadj)
    np.testing.assert_array_equal(cs, S.cols)
    np.testing.assert_array_equal(vs, S.vals)  # scaled values, bit for bit
    out = E.solve(b)
    assert out["iierr"] == 0 and out["rc"] == 0
    assert out["niter"] == ref["iters"]
    assert abs(out["ftol"] - ref["norm"]) <= 1e-10 * np.linalg.norm(b)
    np.testing.assert_allclose(E.solution(), ref["x"], rtol=0, atol=1e-10 * np.abs(ref["x"]).max())
    # ReDoIt: the new right-hand side is not scaled (as in the reference)
    b2 = rng.random(A.n)
    _, ref2 = oracle_elm(rows, cols, vals, A.n, b2, scale_b=False)
    out2 = E.resolve(b2)
    assert out2["niter"] == ref2["iters"]
    np.testing.assert_allclose(E.solution(), ref2["x"], rtol=0, atol=1e-10 * np.abs(ref2["x"]).max())
    E.destroy()


def test_elm_pcg_and_the_reference_sort_order(gpu):
    """Jacobi-PCG through the element API on a Poisson block whose rows
    arrive shuffled: the last row keeps its insertion order (the
    reference's sort), the solve matches the oracle's PCG."""
    A = O.gen_poisson7(9, 8, 7)
    rng = np.random.default_rng(57)
    rows, cols, vals = shuffled_rows(A, rng)
    b = rng.random(A.n)
    E = gpu.Elm(A.n)
    E.set_solver(gpu.ELM_PCG)
    E.insert_rows(rows, cols, vals)
    E.matrix_completed()
    S, ref = oracle_elm(rows, cols, vals, A.n, b, solver="pcg")
    xs, cs, vs = E.matrix_csr()
    np.testing.assert_array_equal(cs, S.cols)
    last = slice(S.xadj[-2] - 1, S.xadj[-1] - 1)
    np.testing.assert_array_equal(cs[last], cols[rows == A.n])  # insertion order kept
    out = E.solve(b)
    assert abs(out["niter"] - ref["iters"]) <= 1
    np.testing.assert_allclose(E.solution(), ref["x"], rtol=0, atol=1e-8 * np.abs(ref["x"]).max())


def test_elm_reinsertion_and_state_errors(gpu):
    n = 6
    E = gpu.Elm(n)
    with pytest.raises(gpu.CelError):
        E.insert(1, 1.0)                       # no row started
    for i in range(1, n + 1):
        E.row_start(i)
        if i > 1:
            E.insert(i - 1, -1.0)
        if i != 3:
            E.insert(i, 4.0)
        if i < n:
            E.insert(i + 1, -1.0)
        if i == 2:
            E.reinsert(2, 9.0)                 # replaces the diagonal value
            with pytest.raises(gpu.CelError) as e:
                E.reinsert(1, 2.0)
            assert e.value.code == -1          # off-diagonal: the reference stops
        if i == 3:
            with pytest.raises(gpu.CelError) as e:
                E.reinsert(3, 1.0)             # no diagonal in this row
            assert e.value.code == -9
            E.insert(3, 5.0)
        E.row_completed()
    E.matrix_completed()
    xs, cs, vs = E.matrix_csr()
    d = np.array([4.0, 9.0, 5.0, 4.0, 4.0, 4.0])
    diag = vs[[np.flatnonzero(cs[xs[i] - 1:xs[i + 1] - 1] == i + 1)[0] + xs[i] - 1 for i in range(n)]]
    np.testing.assert_array_equal(diag, d / np.sqrt(d))
    with pytest.raises(gpu.CelError) as e:
        E.insert(1, 1.0)
    assert e.value.code == -10
    import ctypes as C
    h = C.c_void_p()
    assert gpu.lib().cel_gpu_elm_create(C.byref(h), 10, 5, 1, None) == -1   # not the whole system
    assert gpu.lib().cel_gpu_elm_create(C.byref(h), 5, 5, 2, None) == -1    # ibl must be 1
