"""Pins the CPU oracle (oracle/) before the GPU results are checked against it.

Each oracle function is checked against something that does not restate its
own formula: the reference's own generator output (tests/golden, from
src/base/cel_discretisation_poisson_3d.c), the worked examples of the
reference converter's behaviour, closed forms, brute force on tiny inputs,
and independent library routines (numpy dense algebra, scipy's CG).
"""
import os

from fractions import Fraction

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_generator.npz")


def dense(A: O.CSR, ncols=None):
    n = A.n
    M = np.zeros((n, ncols or n))
    for i in range(n):
        for j in range(A.xadj[i] - 1, A.xadj[i + 1] - 1):
            M[i, A.cols[j] - 1 - A.col_offset] += A.vals[j]
    return M


def grid_neighbours(nx, ny, nz):
    """Independent brute-force 7-point matrix (dense) for tiny grids."""
    n = nx * ny * nz
    M = np.zeros((n, n))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = x + nx * (y + ny * z)
                M[i, i] = 6.0
                for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    X, Y, Z = x + dx, y + dy, z + dz
                    if 0 <= X < nx and 0 <= Y < ny and 0 <= Z < nz:
                        M[i, X + nx * (Y + ny * Z)] = -1.0
    return M


# ------------------------------------------------------------------ work partition
def test_work_partition_closed_form():
    # 32768 rows, 4 threads = 3 workers -> 10920/10920/10928 (thread 0 idle)
    assert [O.work_distr(32768, 3, w) for w in (1, 2, 3)] == [(1, 10920), (10921, 21840), (21841, 32768)]
    # fewer than 4*W elements: everything goes to the last worker
    assert [O.work_distr(7, 3, w) for w in (1, 2, 3)] == [(1, 0), (1, 0), (1, 7)]
    for L in (0, 1, 5, 99, 1000, 4099):
        for W in (1, 2, 3, 7):
            covered = []
            for w in range(1, W + 1):
                s, e = O.work_distr(L, W, w)
                covered.extend(range(s, e + 1))
                if w < W:
                    assert (e - s + 1) % 4 == 0
            assert covered == list(range(1, L + 1))


# ------------------------------------------------------------------ generators
@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 1, 1), (3, 2, 1), (4, 3, 2), (5, 5, 5), (1, 4, 3)])
def test_poisson7_matches_brute_force(shape):
    A = O.gen_poisson7(*shape)
    nx, ny, nz = shape
    # every boundary face of the box removes one entry from 7 per cell
    assert A.nnz == O.poisson7_nnz(*shape) == 7 * nx * ny * nz - 2 * (nx * ny + ny * nz + nx * nz)
    np.testing.assert_array_equal(dense(A), grid_neighbours(*shape))
    # columns ascending in every row
    for i in range(A.n):
        seg = A.cols[A.xadj[i] - 1:A.xadj[i + 1] - 1]
        assert np.all(np.diff(seg) > 0)


def test_poisson7_nnz_at_32_256_384_cubed():
    # 7 N^3 - 6 N^2 entries
    assert O.poisson7_nnz(32, 32, 32) == 223232
    assert O.poisson7_nnz(256, 256, 256) == 117047296
    assert O.poisson7_nnz(384, 384, 384) == 395476992


def test_diffusion7_unit_kappa_is_poisson7():
    A = O.gen_poisson7(4, 3, 5)
    B = O.gen_diffusion7(4, 3, 5, np.ones(60))
    np.testing.assert_array_equal(A.xadj, B.xadj)
    np.testing.assert_array_equal(A.cols, B.cols)
    np.testing.assert_array_equal(A.vals, B.vals)


def test_diffusion7_harmonic_mean_and_symmetry():
    rng = np.random.default_rng(5)
    k = np.exp(rng.normal(0, 1, 3 * 4 * 2))
    M = dense(O.gen_diffusion7(3, 4, 2, k))
    np.testing.assert_allclose(M, M.T, rtol=0, atol=0)
    # off-diagonal = -2 ki kj / (ki + kj); diagonal = sum of faces incl. boundary faces = ki
    i, j = 0, 1
    assert M[i, j] == -(2 * k[i] * k[j]) / (k[i] + k[j])
    rowsum = M.sum(axis=1)
    # A*1 = sum of boundary-face coefficients (each = kappa_i) of the cell
    cells = [(x, y, z) for z in range(2) for y in range(4) for x in range(3)]
    for idx, (x, y, z) in enumerate(cells):
        b = (x == 0) + (x == 2) + (y == 0) + (y == 3) + (z == 0) + (z == 1)
        assert np.isclose(rowsum[idx], b * k[idx], rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------------ reference generator KAT
def _gold():
    if not os.path.exists(GOLD):
        pytest.skip("golden fixtures missing")
    return np.load(GOLD)


def _coo_csr(d, key):
    rows, cols, vals = d[key + "_rows"], d[key + "_cols"], d[key + "_vals"]
    meta = d[key + "_meta"]
    nrows = int(meta[7])
    row_off = int(rows.min())
    xadj = O.coo_to_csr(nrows, rows + 1, row_off)
    return O.CSR(xadj, cols + 1, vals, col_offset=0, row_offset=row_off), row_off


@pytest.mark.parametrize("name", ["s7_5x4x3", "s7_6x6x6", "s27_5x5x5"])
def test_spmv_known_answer_reference_generator(name):
    """cel_cg.f90:808-837: A x_theory = b, |r_i - b_i| < 1e-7, on the
    reference generator's own matrix (x_theory = x^2+y^2+z^2).  The 7- and
    27-point stencils are the ones cel_sp_mat_generate selects
    (cel_sp_mat_generate_module.f90:73-77); the reference's 19-point stencil
    does not satisfy the identity on non-cubic grids, so it only serves as
    an SpMV input elsewhere."""
    d = _gold()
    A, _ = _coo_csr(d, name + "_r0")
    y = O.spmv_csr(A, d[name + "_r0_xt"])
    assert np.max(np.abs(y - d[name + "_r0_b"])) < 1e-7
    # the same product by dense algebra
    np.testing.assert_allclose(dense(A) @ d[name + "_r0_xt"], y, rtol=1e-13, atol=1e-9)


def test_coo_to_csr_counts_rows():
    rows = np.array([1, 1, 3, 3, 3, 4], np.int32)
    np.testing.assert_array_equal(O.coo_to_csr(4, rows, 0), [1, 3, 3, 6, 7])
    np.testing.assert_array_equal(O.coo_to_csr(2, rows[2:] - 2, 0), [1, 4, 5])


# ------------------------------------------------------------------ SpMV vs brute force
def random_csr(n, m, density, rng, empty_rows=True, long_row=0):
    M = (rng.random((n, m)) < density) * rng.normal(size=(n, m))
    if empty_rows and n > 2:
        M[1, :] = 0
    if long_row:
        M[0, : min(long_row, m)] = rng.normal(size=min(long_row, m))
    xadj = [1]
    cols, vals = [], []
    for i in range(n):
        nzr = np.nonzero(M[i])[0]
        cols.extend(nzr + 1)
        vals.extend(M[i, nzr])
        xadj.append(xadj[-1] + len(nzr))
    return O.CSR(np.array(xadj), np.array(cols), np.array(vals)), M


@pytest.mark.parametrize("workers", [1, 3])
def test_spmv_csr_brute_force(workers):
    rng = np.random.default_rng(1)
    A, M = random_csr(70, 50, 0.1, rng)
    x = rng.normal(size=50)
    np.testing.assert_allclose(O.spmv_csr(A, x, workers=workers), M @ x, rtol=1e-13, atol=1e-13)
    y0 = rng.normal(size=70)
    np.testing.assert_allclose(O.spmv_csr(A, x, y0), y0 + M @ x, rtol=1e-13, atol=1e-13)


def test_spmv_csr_row_order_is_sequential():
    # a row whose sum depends on the order: ((1e16 + 1) - 1e16) = 0 in FP64
    A = O.CSR([1, 4], [1, 2, 3], [1e16, 1.0, -1e16])
    assert O.spmv_csr(A, np.ones(3))[0] == 0.0
    A2 = O.CSR([1, 4], [1, 3, 2], [1e16, -1e16, 1.0])
    assert O.spmv_csr(A2, np.ones(3))[0] == 1.0


def test_spmv_f32_brute_force():
    rng = np.random.default_rng(2)
    A, M = random_csr(40, 40, 0.15, rng)
    x = rng.normal(size=40).astype(np.float32)
    y = O.spmv_csr_f32(A, x)
    np.testing.assert_allclose(y, (M.astype(np.float32) @ x), rtol=2e-5, atol=2e-5)


def test_spmv_coo_brute_force():
    rng = np.random.default_rng(3)
    A, M = random_csr(30, 30, 0.2, rng)
    rows = np.repeat(np.arange(1, 31), np.diff(A.xadj))
    x = rng.normal(size=30)
    y = O.spmv_coo(rows, A.cols, A.vals, 0, 0, x, np.zeros(30))
    np.testing.assert_allclose(y, M @ x, rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------------ JAD conversion
def jad_from_lengths(lens):
    xadj = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    nnz = int(sum(lens))
    cols = np.concatenate([np.arange(1, L + 1) for L in lens]) if nnz else np.zeros(0)
    vals = np.arange(1, nnz + 1, dtype=np.float64)
    return O.CSR(xadj, cols, vals, col_offset=0)


def test_dcsort_is_stable_descending():
    rng = np.random.default_rng(4)
    for _ in range(20):
        lens = rng.integers(2, 9, size=rng.integers(1, 40))
        icnt, perm = O.jad_dcsort(lens, int(lens.min()), int(lens.max()))
        expect = np.argsort(-lens, kind="stable") + 1
        np.testing.assert_array_equal(perm, expect)
        for v in range(lens.min(), lens.max() + 1):
            assert icnt[v - lens.min()] == np.sum(lens > v)


def test_jad_worked_examples():
    # ilo = 0: reference counts equal SPARSKIT's: lengths [0,2,1] -> [2,1]
    for mode in (O.JAD_SPARSKIT, O.JAD_CEL_COMPAT):
        j = O.jad_convert(jad_from_lengths([0, 2, 1]), mode)
        np.testing.assert_array_equal(np.diff(j["begin"]), [2, 1])
        np.testing.assert_array_equal(j["perm"], [2, 3, 1])
    # ilo >= 1: lengths [2,3,1,3] -> reference [3,2,0] vs correct [4,3,2]
    s = O.jad_convert(jad_from_lengths([2, 3, 1, 3]), O.JAD_SPARSKIT)
    c = O.jad_convert(jad_from_lengths([2, 3, 1, 3]), O.JAD_CEL_COMPAT)
    np.testing.assert_array_equal(np.diff(s["begin"]), [4, 3, 2])
    np.testing.assert_array_equal(s["perm"], [2, 4, 1, 3])
    np.testing.assert_array_equal(c["perm"], [2, 4, 1, 3])
    assert c["ndefined"] == 3
    np.testing.assert_array_equal(np.diff(c["begin"][:4]), [3, 2, 0])


def test_jad_4cube_reference_counts():
    A = O.gen_poisson7(4, 4, 4)
    c = O.jad_convert(A, O.JAD_CEL_COMPAT)
    s = O.jad_convert(A, O.JAD_SPARSKIT)
    assert c["ilo"] == 4 and c["mmax"] == 7 and c["ndefined"] == 4
    np.testing.assert_array_equal(np.diff(c["begin"][:5]), [56, 32, 8, 0])
    assert list(c["begin"][5:]) == [-1, -1, -1]
    np.testing.assert_array_equal(np.diff(s["begin"]), [64, 64, 64, 64, 56, 32, 8])
    np.testing.assert_array_equal(s["perm"], c["perm"])


@pytest.mark.parametrize("shape", [(4, 4, 4), (5, 3, 2), (7, 1, 1)])
def test_jad_spmv_equals_csr_bitwise(shape):
    A = O.gen_poisson7(*shape)
    j = O.jad_convert(A)
    x = np.random.default_rng(6).random(A.n)
    np.testing.assert_array_equal(O.spmv_jad(j, x), O.spmv_csr(A, x))
    # perm = stable descending sort of row lengths
    np.testing.assert_array_equal(j["perm"], np.argsort(-np.diff(A.xadj), kind="stable") + 1)


def test_jad_random_rows_with_empty_rows():
    rng = np.random.default_rng(7)
    A, M = random_csr(50, 50, 0.1, rng)
    for mode in (O.JAD_SPARSKIT, O.JAD_CEL_COMPAT):  # ilo == 0 here: modes agree
        j = O.jad_convert(A, mode)
        assert j["ilo"] == 0 and j["ndefined"] == j["mmax"]
        np.testing.assert_allclose(O.spmv_jad(j, np.ones(50)), M.sum(axis=1), rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------------ Jacobi / scaling
def test_jacobi_and_scaling_closed_forms():
    d = np.array([4.0, -9.0, 0.25])
    np.testing.assert_array_equal(O.set_jacobi(d), 1.0 / d)
    np.testing.assert_array_equal(O.set_jacobi(d, scaled=True), np.sqrt(np.abs(d)) / d)
    A = O.CSR([1, 3, 4, 5], [1, 2, 2, 3], [4.0, 2.0, -9.0, 0.25])
    S = O.scale_csr(A, d)
    np.testing.assert_array_equal(S.vals, [2.0, 1.0, -3.0, 0.5])
    # scaled diagonal = sign(d) sqrt|d| and jacobi(scaled) = 1 / scaled diagonal
    np.testing.assert_allclose(1.0 / O.set_jacobi(d, scaled=True), np.sign(d) * np.sqrt(np.abs(d)))
    np.testing.assert_array_equal(O.scale_vec([4.0, 9.0, 1.0], d), [2.0, 3.0, 2.0])


# ------------------------------------------------------------------ PCG
def textbook_pcg(M, jac, b, x0, eps, iter_max):
    """Dense numpy Jacobi-PCG in the reference's op order, written from the
    textbook recurrences (independent of the C restatement)."""
    x = x0.copy()
    r = b - M @ x
    z = jac * r
    p = z.copy()
    a = r @ z
    nrm = [np.sqrt(r @ r)]
    k = 0
    while k < iter_max and abs(a) > 1.1 * np.finfo(float).eps and nrm[-1] > eps:
        v = M @ p
        k += 1
        s = a / (v @ p)
        x = x + s * p
        r = r - s * v
        z = jac * r
        a_new = r @ z
        p = (a_new / a) * p + z
        a = a_new
        nrm.append(np.sqrt(r @ r))
    return x, k, np.array(nrm)


def test_pcg_matches_textbook_dense():
    A = O.gen_poisson7(6, 5, 4)
    rng = np.random.default_rng(0)
    xt = rng.random(A.n)
    b = O.spmv_csr(A, xt)
    jac = O.set_jacobi(O.csr_diagonal(A))
    eps = 1e-10 * np.linalg.norm(b)
    r = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=eps, iter_max=500, workers=1)
    xd, kd, hd = textbook_pcg(dense(A), jac, b, np.zeros(A.n), eps, 500)
    assert abs(r["iters"] - kd) <= 1
    m = min(len(hd), len(r["hist"]))
    np.testing.assert_allclose(r["hist"][:m], hd[:m], rtol=1e-9)
    np.testing.assert_allclose(r["x"], xt, atol=1e-8)


def test_pcg_variable_coefficient_vs_scipy_cg():
    scipy_sparse = pytest.importorskip("scipy.sparse")
    spla = pytest.importorskip("scipy.sparse.linalg")
    rng = np.random.default_rng(3)
    nx, ny, nz = 9, 8, 7
    k = np.exp(rng.normal(0, 0.5, nx * ny * nz))
    A = O.gen_diffusion7(nx, ny, nz, k)
    S = scipy_sparse.csr_matrix((A.vals, A.cols - 1, A.xadj - 1), shape=(A.n, A.n))
    # scaled so the reference's absolute |alpha| > 1.1 eps stop does not fire
    # before the norm criterion (see test_pcg_alpha_stop_is_absolute)
    b = 1e3 * rng.uniform(-1, 1, A.n)
    d = O.csr_diagonal(A)
    jac = O.set_jacobi(d)
    tol = 1e-9
    r = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=tol * np.linalg.norm(b), iter_max=2000, workers=2)
    count = [0]
    xs, info = spla.cg(S, b, rtol=tol, atol=0.0, maxiter=2000,
                       M=scipy_sparse.diags(1.0 / d), callback=lambda _: count.__setitem__(0, count[0] + 1))
    assert info == 0
    assert abs(count[0] - r["iters"]) <= 1
    np.testing.assert_allclose(r["x"], xs, rtol=1e-6, atol=1e-8)
    # true residual meets the tolerance
    assert np.linalg.norm(b - S @ r["x"]) <= 1.05 * tol * np.linalg.norm(b)


def test_pcg_alpha_stop_is_absolute():
    """cel_cgalg_module.f90:247-249 also stops when |alpha| = |r.z| <=
    1.1 eps_mach, an absolute test: a small right-hand side stops early."""
    A = O.gen_poisson7(6, 6, 6)
    b = O.spmv_csr(A, np.random.default_rng(0).random(A.n))
    jac = O.set_jacobi(O.csr_diagonal(A))
    big = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=0.0, iter_max=400)
    small = O.pcg_jacobi(A, jac, 1e-6 * b, x0_mode=O.X0_ZERO, eps=0.0, iter_max=400)
    assert small["iters"] < big["iters"]
    assert abs(small["alpha"][-1]) <= 1.1 * np.finfo(float).eps


def test_pcg_x0_one_equals_shifted_zero_start():
    """The reference hard-codes x0 = 1: solving A x = b + A*1 from x0 = 1
    follows the same residual sequence as A x = b from x0 = 0."""
    A = O.gen_poisson7(8, 8, 8)
    rng = np.random.default_rng(0)
    b = O.spmv_csr(A, rng.random(A.n))
    jac = O.set_jacobi(O.csr_diagonal(A))
    eps = 1e-8 * np.linalg.norm(b)
    r0 = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=eps, iter_max=500)
    r1 = O.pcg_jacobi(A, jac, b + O.spmv_csr(A, np.ones(A.n)), x0_mode=O.X0_ONE, eps=eps, iter_max=500)
    assert r0["iters"] == r1["iters"]
    np.testing.assert_allclose(r1["hist"], r0["hist"], rtol=1e-9)
    np.testing.assert_allclose(r1["x"] - 1.0, r0["x"], atol=1e-9)


def test_pcg_worker_grouping_changes_only_rounding():
    A = O.gen_poisson7(16, 16, 16)
    b = O.spmv_csr(A, np.random.default_rng(0).random(A.n))
    jac = O.set_jacobi(O.csr_diagonal(A))
    eps = 1e-10 * np.linalg.norm(b)
    a = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=eps, iter_max=1000, workers=1)
    c = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=eps, iter_max=1000, workers=3)
    assert abs(a["iters"] - c["iters"]) <= 1
    m = min(a["iters"], c["iters"]) + 1
    rel = np.abs(a["hist"][:m] - c["hist"][:m]) / a["hist"][:m]
    assert rel.max() < 1e-9


def test_pcg_recursive_residual_tracks_true_residual():
    A = O.gen_poisson7(10, 9, 8)
    b = O.spmv_csr(A, np.random.default_rng(1).random(A.n))
    jac = O.set_jacobi(O.csr_diagonal(A))
    r = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=1e-10 * np.linalg.norm(b), iter_max=1000)
    true = np.linalg.norm(b - O.spmv_csr(A, r["x"]))
    assert abs(true - r["norm"]) <= 1e-3 * r["norm"] + 1e-12 * np.linalg.norm(b)


def test_pcg_iteration_cap_and_alpha_stop():
    A = O.gen_poisson7(6, 6, 6)
    b = O.spmv_csr(A, np.ones(A.n))
    jac = O.set_jacobi(O.csr_diagonal(A))
    r = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=0.0, iter_max=5)
    assert r["iters"] == 5
    # exact start: r0 = 0 -> alpha = 0 -> no iteration
    r = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ONE, eps=1e-8, iter_max=5)
    assert r["iters"] == 0 and r["norm"] == 0.0


# ------------------------------------------------------------------ distributed split
def _decomp_case(d, name, nblocks):
    """Global row starts and every rank's COO (1-based, global ids)."""
    parts = []
    for r in range(nblocks):
        key = f"{name}_r{r}"
        parts.append((d[key + "_rows"] + 1, d[key + "_cols"] + 1, d[key + "_vals"], d[key + "_b"]))
    sizes = [p[3].size for p in parts]
    vtx = np.concatenate([[1], 1 + np.cumsum(sizes)]).astype(np.int32)
    return vtx, parts


@pytest.mark.parametrize("name,nblocks", [("s7_8x8x4_d2x2x1", 4), ("s27_6x6x6_d2x1x3", 6)])
def test_distr_split_reassembles_global_spmv(name, nblocks):
    d = _gold()
    vtx, parts = _decomp_case(d, name, nblocks)
    N = int(vtx[-1] - 1)
    x = np.random.default_rng(2).random(N)
    # global product from all ranks' rows (dense brute force)
    G = np.zeros((N, N))
    for rows, cols, vals, _ in parts:
        for r_, c_, v_ in zip(rows, cols, vals):
            G[r_ - 1, c_ - 1] += v_
    yg = G @ x
    for rank, (rows, cols, vals, _) in enumerate(parts):
        s = O.distr_split(vtx, rank, rows, cols, vals)
        blocks = s["blocks"]
        assert blocks[0]["proc"] == rank
        assert [b["proc"] for b in blocks[1:]] == sorted(b["proc"] for b in blocks[1:])
        assert sum(b["rows"].size for b in blocks) == rows.size
        off = s["row_offset"]
        n = s["n"]
        # local block, then halo blocks through full-length neighbour copies
        y = np.zeros(n)
        b0 = blocks[0]
        y = O.spmv_csr(O.CSR(b0["xadj"], b0["cols"], b0["vals"], b0["col_offset"]),
                       x[vtx[rank] - 1:vtx[rank + 1] - 1], y)
        for b in blocks[1:]:
            p = b["proc"]
            xn = np.zeros(vtx[p + 1] - vtx[p])
            perm = b["comm_perm"]
            assert np.all(np.diff(perm) > 0)
            # receive buffer = neighbour's packed values, unpacked at perm
            buf = O.face_pack(perm, x[vtx[p] - 1:vtx[p + 1] - 1])
            xn[perm - 1] = buf
            np.testing.assert_array_equal(np.unique(b["cols"] - b["col_offset"]), perm)
            y = O.spmv_coo(b["rows"], b["cols"], b["vals"], off, b["col_offset"], xn, y)
        np.testing.assert_allclose(y, yg[vtx[rank] - 1:vtx[rank + 1] - 1], rtol=1e-12, atol=1e-12)
        # captured diagonal is the block's diagonal
        np.testing.assert_array_equal(s["diagonal"], np.diag(G)[vtx[rank] - 1:vtx[rank + 1] - 1])


# ------------------------------------------------------------------ gradients
def test_kappa_gradient_central_differences():
    """Pins the oracle's chain rule: dL/dkappa by central differences of
    L = 0.5 ||x(kappa) - x_obs||^2 with dense solves on a tiny grid."""
    rng = np.random.default_rng(4)
    nx, ny, nz = 3, 3, 2
    n = nx * ny * nz
    k = np.exp(rng.normal(0, 1, n))
    b = rng.random(n)
    xo = rng.normal(size=n)

    def loss(kk):
        M = dense(O.gen_diffusion7(nx, ny, nz, kk))
        x = np.linalg.solve(M, b)
        return 0.5 * np.sum((x - xo) ** 2), x, M

    L0, x, M = loss(k)
    lam = np.linalg.solve(M.T, x - xo)
    g = O.diffusion7_kappa_grad(nx, ny, nz, k, x, lam)
    h = 1e-6
    fd = np.array([(loss(k + h * e)[0] - loss(k - h * e)[0]) / (2 * h) for e in np.eye(n)])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)


def test_sddmm_brute_force():
    rng = np.random.default_rng(5)
    A, M = random_csr(20, 20, 0.3, rng)
    lam, x = rng.normal(size=20), rng.normal(size=20)
    g = O.sddmm_csr(A, lam, x)
    full = -np.outer(lam, x)
    rows = np.repeat(np.arange(20), np.diff(A.xadj))
    np.testing.assert_array_equal(g, full[rows, A.cols - 1])


def test_face_difference_operator_matches_generated_matrices():
    """delta_apply (the finite-difference perturbation of the adjoint test)
    equals (A(kappa') - A(kappa)) v built from two generated matrices, to the
    rounding of the change."""
    import scipy.sparse as sp

    from helpers import delta_apply, face_coeffs

    for shape in ((5, 4, 3), (6, 6, 6)):
        n = int(np.prod(shape))
        rng = np.random.default_rng(7)
        kappa = np.exp(rng.normal(0.0, 1.0, n))
        kp = kappa * np.exp(1e-3 * rng.normal(size=n))
        A0 = O.gen_diffusion7(*shape, kappa)
        A1 = O.gen_diffusion7(*shape, kp)
        dA = sp.csr_matrix((A1.vals - A0.vals, A0.cols - 1, A0.xadj - 1), shape=(n, n))
        v = rng.normal(size=n)
        ref = dA @ v
        got = delta_apply(shape, kappa, kp, face_coeffs(shape, kappa), v)
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


# ---------------------------------------------------------------- COO sort
# Hand traces of cel_sp_mat_format_sparse_31_sort
# (src/fspmt/cel_sp_mat_module.f90:518-585).  Values carry the input
# position so every move is visible.  Each expected order was derived by
# stepping the loop (cur_index, last_row, row_counter, start/end_index).
SORT_TRACES = {
    # all rows present, rows(1) == 1: rows 1..n-1 column-sorted, the last
    # row keeps its file order (the loop stops at its first entry)
    "last_row_unsorted": (3, [1, 1, 2, 2, 3, 3], [2, 1, 3, 1, 3, 1],
                          0, [1, 1, 2, 2, 3, 3], [1, 2, 1, 3, 3, 1], [1, 0, 3, 2, 4, 5]),
    # row 1 empty, so rows(1) = 2 != 1: a transition at index 1 closes
    # [1,1]; the first entry is left out of its row's column sort
    "first_entry_excluded": (3, [2, 2, 2, 3, 3], [3, 2, 1, 2, 1],
                             0, [2, 2, 2, 3, 3], [3, 1, 2, 2, 1], [0, 2, 1, 3, 4]),
    # last row of one entry: the transition at index nnz closes
    # [start, nnz], sorting it by column together with the row before it
    "single_last_row_merged": (3, [1, 1, 2, 2, 2, 3], [2, 1, 3, 1, 2, 2],
                               0, [1, 1, 2, 2, 3, 2], [1, 2, 1, 2, 2, 3], [1, 0, 3, 4, 5, 2]),
    # nrows == 1: the loop body never runs; rows sorted (stably) only
    "one_row": (1, [1, 1, 1], [3, 1, 2], 0, [1, 1, 1], [3, 1, 2], [0, 1, 2]),
    # empty middle row: the transition at index nnz (rows unchanged) closes
    # the last row, which is then sorted too
    "empty_middle_row": (3, [3, 1, 3, 1], [2, 2, 1, 1],
                         0, [1, 1, 3, 3], [1, 2, 1, 2], [3, 1, 2, 0]),
    # too few transitions: the loop would read rows(nnz+1)
    "reads_past_end": (3, [1, 3, 1], [2, 1, 1], -1, None, None, None),
    "no_entries": (2, [], [], -1, None, None, None),
}


@pytest.mark.parametrize("name", sorted(SORT_TRACES))
def test_coo_sort_hand_traces(name):
    nrows, rows, cols, st_want, r_want, c_want, p_want = SORT_TRACES[name]
    vals = np.arange(len(rows), dtype=np.float64)
    st, r, c, v = O.coo_sort_cel(nrows, rows, cols, vals)
    assert st == st_want
    if st == 0:
        np.testing.assert_array_equal(r, r_want)
        np.testing.assert_array_equal(c, c_want)
        np.testing.assert_array_equal(v, p_want)
    else:  # left sorted by row (stable)
        order = np.argsort(np.asarray(rows), kind="stable")
        np.testing.assert_array_equal(v, np.arange(len(rows))[order])
    if O.ref_mm_available() and st == 0:
        err, r2, c2, v2 = O.ref_coo_sort(nrows, rows, cols, vals)
        assert err == 0
        np.testing.assert_array_equal(r2, r)
        np.testing.assert_array_equal(c2, c)
        np.testing.assert_array_equal(v2, v)


def random_coo_case(rng, all_rows=False):
    n = int(rng.integers(1, 12))
    nnz = int(rng.integers(0, 60))
    rows = rng.integers(1, n + 1, nnz)
    if all_rows:
        rows = np.concatenate([np.arange(1, n + 1), rows])
        rng.shuffle(rows)
    cols = rng.integers(1, n + 4, rows.size)
    return n, rows.astype(np.int32), cols.astype(np.int32), rng.random(rows.size)


@pytest.mark.skipif(not O.ref_mm_available(), reason="reference sort not built (oracle/_ref)")
def test_coo_sort_matches_compiled_reference():
    """The restatement against the reference's own routine, compiled from
    its sources, on every defined random case (duplicates included; runs of
    32+ entries take the merge path, shorter ones the bubble path)."""
    rng = np.random.default_rng(21)
    checked = undefined = 0
    for t in range(1500):
        n, rows, cols, vals = random_coo_case(rng, all_rows=t % 3 == 0)
        st, r, c, v = O.coo_sort_cel(n, rows, cols, vals)
        if st != 0:
            undefined += 1
            continue
        err, r2, c2, v2 = O.ref_coo_sort(n, rows, cols, vals)
        assert err == 0
        np.testing.assert_array_equal(r, r2)
        np.testing.assert_array_equal(c, c2)
        np.testing.assert_array_equal(v, v2)
        checked += 1
    assert checked > 1000 and undefined > 100


def test_coo_sort_rowcol_is_lexicographic_and_stable():
    rng = np.random.default_rng(22)
    for _ in range(200):
        n, rows, cols, vals = random_coo_case(rng)
        r, c, v = O.coo_sort_rowcol(rows, cols, vals)
        order = np.lexsort((cols, rows))  # stable: ties keep input order
        np.testing.assert_array_equal(r, rows[order])
        np.testing.assert_array_equal(c, cols[order])
        np.testing.assert_array_equal(v, vals[order])


def test_coo_sort_cel_differs_from_rowcol_only_where_the_quirks_say():
    """With every row present and rows(1) == 1, the reference's order equals
    the full (row, col) sort except inside the last row, which keeps the
    row-sorted (file) order; a last row of one entry changes nothing when
    its column is the largest."""
    rng = np.random.default_rng(23)
    for _ in range(200):
        n, rows, cols, vals = random_coo_case(rng, all_rows=True)
        if n < 2:
            continue
        st, r, c, v = O.coo_sort_cel(n, rows, cols, vals)
        assert st == 0
        r0, c0, v0 = O.coo_sort_rowcol(rows, cols, vals)
        last = int((rows == n).sum())
        head = rows.size - last
        if last > 1:
            np.testing.assert_array_equal(r[:head], r0[:head])
            np.testing.assert_array_equal(c[:head], c0[:head])
            file_order = np.flatnonzero(rows == n)
            np.testing.assert_array_equal(v[head:], vals[file_order])
        elif cols[rows == n][0] >= cols[rows == n - 1].max():
            np.testing.assert_array_equal(r, r0)
            np.testing.assert_array_equal(c, c0)
            np.testing.assert_array_equal(v, v0)


# ---------------------------------------------------------------- Jacobi-GMRES
def convdiff_csr(nx, ny, nz, c=0.5):
    """7-point operator with a first-order upwind term in x: diag 6 + c,
    west neighbour -1 - c, the others -1 (nonsymmetric, diagonally
    dominant).  1-based CSR, ascending columns."""
    A = O.gen_poisson7(nx, ny, nz)
    vals = A.vals.copy()
    rows = np.repeat(np.arange(1, A.n + 1), np.diff(A.xadj))
    vals[A.cols == rows] += c
    vals[A.cols == rows - 1] -= c  # west neighbour (same y, z line: x-1)
    # x-1 of the first cell of a line is not a neighbour: the generator holds
    # no entry there, so nothing else changes
    return O.CSR(A.xadj, A.cols, vals)


def dense_of(A):
    D = np.zeros((A.n, A.n))
    rows = np.repeat(np.arange(A.n), np.diff(A.xadj))
    D[rows, A.cols - 1] = A.vals
    return D


def test_gmres_restart1_is_minimal_residual():
    """With one Arnoldi step per cycle no earlier rotation exists, and each
    cycle is the minimal-residual step along z = jac * r: x += a z with
    a = (r . Az) / (Az . Az), stopping once the residual estimate
    |r| sin(angle(r, Az)) <= eps.  The independent numpy loop below counts
    the same steps and lands on the same x."""
    A = convdiff_csr(6, 5, 4)
    D = dense_of(A)
    jac = O.set_jacobi(O.csr_diagonal(A))
    b = np.random.default_rng(41).random(A.n)
    eps = 1e-9
    ref = O.gmres_jacobi(A, jac, b, eps=eps, m=1, iter_max=5000)
    x = np.ones(A.n)
    steps = 0
    while True:
        r = b - D @ x
        beta = np.linalg.norm(r)
        z = jac * r / beta
        Az = D @ z
        v = r / beta
        a = beta * (v @ Az) / (Az @ Az)
        x = x + a * z
        steps += 1
        est = beta * np.sqrt(max(0.0, 1.0 - (v @ Az) ** 2 / (Az @ Az)))
        if est <= eps or steps > 10000:
            break
    assert ref["reason"] == 2
    assert ref["cycles"] == steps
    np.testing.assert_allclose(ref["x"], x, rtol=0, atol=1e-12 * np.abs(x).max())


def test_gmres_diagonal_matrix_is_solved_in_one_step():
    """A diagonal with jac = 1/d: w = A z1 = v1, h11 = 1, h21 = 0, so the
    first estimate is 0 and x = 1 + (b - d)/d = b/d."""
    n = 50
    d = np.random.default_rng(42).uniform(0.5, 3.0, n) * np.where(np.arange(n) % 3 == 0, -1, 1)
    A = O.CSR(np.arange(1, n + 2, dtype=np.int32), np.arange(1, n + 1, dtype=np.int32), d)
    b = np.random.default_rng(43).normal(size=n)
    ref = O.gmres_jacobi(A, 1.0 / d, b, eps=1e-8, m=3)
    assert (ref["iters"], ref["cycles"], ref["reason"]) == (1, 1, 2)
    np.testing.assert_allclose(ref["x"], b / d, rtol=0, atol=1e-14 * np.abs(b / d).max())


def reference_gmres_dense(D, jac, b, eps, iter_max, m):
    """cel_gmresalg_2jacobi written again from the reference text
    (src/fcgalg/cel_gmresalg_module.f90:637-840) over dense numpy arrays,
    1-based dictionaries for the small arrays: a second transcription that
    catches index and loop mistakes in the C restatement (not a misreading
    of the reference, which both would share)."""
    n = b.size
    x = np.ones(n)
    hh, cc, ss, gam, alp = {}, {}, {}, {}, {}
    conv, it, cycles, res = 0, 0, 0, 0.0
    while it <= iter_max and conv == 0:
        it += 1
        r = b - D @ x
        res = np.sqrt(r @ r)
        gam[1] = res
        loc = res
        rc = 0
        while rc <= m and conv == 0:
            v = {1: r * (1.0 / loc)}
            z = {}
            gam[1] = res
            for jj in range(1, m + 1):
                z[jj] = v[jj] * jac
                w = D @ z[jj]
                for ii in range(1, jj + 1):
                    hh[ii, jj] = v[ii] @ w
                    w = w - hh[ii, jj] * v[ii]
                hh[jj + 1, jj] = np.sqrt(w @ w)
                for ii in range(1, jj):
                    hh[ii, jj] = hh[ii, jj] * cc[ii] + hh[ii + 1, jj] * ss[ii]
                    hh[ii, jj] = -hh[ii, jj] * ss[ii] + hh[ii + 1, jj] * cc[ii]
                be = np.sqrt(hh[jj, jj] ** 2 + hh[jj + 1, jj] ** 2)
                ss[jj] = hh[jj + 1, jj] / be
                cc[jj] = hh[jj, jj] / be
                hh[jj, jj] = be
                gam[jj + 1] = -ss[jj] * gam[jj]
                gam[jj] = cc[jj] * gam[jj]
                if abs(gam[jj + 1]) <= eps:
                    for ii in range(jj, 0, -1):
                        alp[ii] = sum((hh[ii, kk] * alp[kk] for kk in range(ii + 1, jj + 1)), 0.0)
                        alp[ii] = (1.0 / hh[ii, ii]) * (gam[ii] - alp[ii])
                    for ii in range(1, jj + 1):
                        x = x + alp[ii] * z[ii]
                    conv = 2
                    break
                v[jj + 1] = w / hh[jj + 1, jj]
            if conv == 0:
                for ii in range(m, 0, -1):
                    alp[ii] = sum((hh[ii, kk] * alp[kk] for kk in range(ii + 1, m + 1)), 0.0)
                    alp[ii] = (1.0 / hh[ii, ii]) * (gam[ii] - alp[ii])
                for ii in range(1, m + 1):
                    x = x + alp[ii] * z[ii]
            r = b - D @ x
            res = np.sqrt(r @ r)
            loc = res
            cycles += 1
            rc += 1
    return dict(iters=it, cycles=cycles, norm=res, x=x, reason=conv)


@pytest.mark.parametrize("m", [2, 3, 5])
def test_gmres_matches_second_transcription(m):
    A = convdiff_csr(5, 4, 4)
    D = dense_of(A)
    jac = O.set_jacobi(O.csr_diagonal(A))
    b = np.random.default_rng(44).random(A.n)
    ref = O.gmres_jacobi(A, jac, b, eps=1e-8, m=m, iter_max=200)
    alt = reference_gmres_dense(D, jac, b, 1e-8, 200, m)
    assert (ref["iters"], ref["cycles"], ref["reason"]) == (alt["iters"], alt["cycles"], alt["reason"])
    np.testing.assert_allclose(ref["x"], alt["x"], rtol=0, atol=1e-9 * np.abs(alt["x"]).max())
    assert abs(ref["norm"] - alt["norm"]) <= 1e-6 * alt["norm"] + 1e-14


def test_gmres_final_norm_is_the_true_residual_and_shows_the_rotation_defect():
    """ftol is ||b - A x|| of the returned x.  With restart >= 2 the
    estimate of the reference's unrotated Hessenberg column meets eps while
    the true residual stays far above it; a textbook rotation (h(i+1,j)
    rotated too) would leave the true residual near eps, so this fails if
    the restatement 'fixes' the reference."""
    A = O.gen_poisson7(8, 8, 8)
    D = dense_of(A)
    jac = O.set_jacobi(O.csr_diagonal(A))
    b = np.random.default_rng(0).random(A.n)
    for m, workers in ((1, 1), (3, 1), (3, 4)):
        ref = O.gmres_jacobi(A, jac, b, eps=1e-8, m=m, workers=workers)
        true = np.linalg.norm(b - D @ ref["x"])
        assert ref["reason"] == 2
        assert abs(ref["norm"] - true) <= 1e-13 * np.linalg.norm(b)
        if m == 1:
            assert ref["norm"] <= 1.01e-8
        else:
            assert ref["norm"] > 100 * 1e-8


# The hand-trace system: A = 2 V H V^T with V = [[1,2,2],[2,1,-2],[2,-2,1]]/3
# (orthonormal columns v1, v2, v3) and H = [[3,1,0],[4,5,-13],[0,12,-5]]
# (upper Hessenberg), so diag(A) = 2, jac = 1/diag(A) = 1/2 and A jac = V H V^T:
# the Arnoldi vectors are v1, v2, v3 and the Arnoldi columns are those of H.
# x0 = 1 and b = A 1 + (1, 2, 2), so r0 = 3 v1.
GM3_A = [[2.0, 86 / 3, -56 / 3], [-20 / 3, 2.0, 34 / 3], [32 / 3, -28 / 3, 2.0]]
GM3_B = [13.0, 26 / 3, 16 / 3]


def gm3_system():
    A = O.CSR(np.array([1, 4, 7, 10], np.int32), np.tile(np.arange(1, 4, dtype=np.int32), 3),
              np.array(GM3_A).ravel())
    return A, np.full(3, 0.5), np.array(GM3_B)


def gm3_x(alphas):
    """x0 + sum_i alpha_i z_i with z_i = jac v_i = v_i / 2, in exact arithmetic."""
    v = [(1, 2, 2), (2, 1, -2), (2, -2, 1)]
    return [1 + sum(a * Fraction(vi[r], 6) for a, vi in zip(alphas, v)) for r in range(3)]


def test_gmres_hand_trace_restart3_converges_with_the_reference_rotation():
    """cel_gmresalg_2jacobi (src/fcgalg/cel_gmresalg_module.f90:637-851) by
    hand on the 3x3 system above, m = 3, in exact arithmetic.
      r0 = (1,2,2): |r0| = 3, gamma = (3).
      jj=1: w = A z1 = 3 v1 + 4 v2; h11 = 3, h21 = 4; beta = 5, c1 = 3/5,
            s1 = 4/5, h11 <- 5; gamma = (9/5, -12/5).
      jj=2: w = A z2 = v1 + 5 v2 + 12 v3; h12 = 1, h22 = 5, h32 = 12.
            Rotation 1 as the reference writes it (:722-727; both statements
            assign h(1,2), h(2,2) is never rotated): t = 1*3/5 + 5*4/5 =
            23/5, h12 <- -t*4/5 + 5*3/5 = -17/25.  beta = sqrt(5^2 + 12^2)
            = 13, c2 = 5/13, s2 = 12/13, h22 <- 13; gamma = (9/5, -12/13,
            144/65).
      jj=3: w = A z3 = -13 v2 - 5 v3; h13 = 0, h23 = -13, h33 = -5, h43 = 0.
            Rotation 1: t = 0*3/5 - 13*4/5 = -52/5, h13 <- 52/5*4/5 - 13*3/5
            = 13/25.  Rotation 2 (reads the unrotated h23 = -13): t = -13*5/13
            - 5*12/13 = -125/13, h23 <- 125/13*12/13 - 5*5/13 = 1175/169.
            beta = |h33| = 5, c3 = -1, s3 = 0, h33 <- 5; gamma4 = 0 <= eps,
            gamma3 <- -144/65.
      back substitution: alpha3 = (-144/65)/5 = -144/325, alpha2 = (-12/13 -
            1175/169*alpha3)/13 = 4740/28561, alpha1 = (9/5 + 17/25*alpha2 -
            13/25*alpha3)/5 = 38258409/89253125.
    One cycle, reason XN.  A textbook rotation would solve this 3x3 system
    exactly in 3 steps; the reference's lands at a residual of 9.437."""
    A, jac, b = gm3_system()
    x = gm3_x([Fraction(38258409, 89253125), Fraction(4740, 28561), Fraction(-144, 325)])
    assert x == [Fraction(174770053, 178506250), Fraction(117656678, 89253125),
                 Fraction(90477428, 89253125)]
    ref = O.gmres_jacobi(A, jac, b, eps=1e-10, m=3, iter_max=50)
    assert (ref["iters"], ref["cycles"], ref["reason"]) == (1, 1, 2)
    np.testing.assert_allclose(ref["x"], [float(v) for v in x], rtol=0, atol=1e-14)
    r = [Fraction(GM3_B[i]) - sum(Fraction(GM3_A[i][j]) * x[j] for j in range(3)) for i in range(3)]
    true = float(sum(v * v for v in r)) ** 0.5
    assert abs(ref["norm"] - true) <= 1e-13 * true and 9.43 < true < 9.44
    assert np.abs(ref["x"] - np.linalg.solve(np.array(GM3_A), b)).max() > 0.1


def test_gmres_hand_trace_restart2_ends_its_cycle_unconverged():
    """The same system with m = 2: jj=1 and jj=2 as in the m = 3 trace;
    |gamma3| = 144/65 > eps, so v3 = w / 12 and the cycle ends with the back
    substitution over restart_max columns (:783-806): alpha2 = (-12/13)/13 =
    -12/169, alpha1 = (9/5 - (-17/25)(-12/169))/5 = 7401/21125.  The
    residual after the cycle (the first history entry) is |b - A x1| =
    sqrt(105339456/17850625).  (The later cycles of the reference's rotation
    diverge on this system; only the first is traced.)"""
    A, jac, b = gm3_system()
    x = gm3_x([Fraction(7401, 21125), Fraction(-12, 169)])
    assert x == [Fraction(43717, 42250), Fraction(23342, 21125), Fraction(24092, 21125)]
    r = [Fraction(GM3_B[i]) - sum(Fraction(GM3_A[i][j]) * x[j] for j in range(3)) for i in range(3)]
    r2 = sum(v * v for v in r)
    # GM3_A and GM3_B hold binary roundings of thirds; the exact system's
    # value is the one stated above
    assert abs(float(r2) - 105339456 / 17850625) <= 1e-13
    ref = O.gmres_jacobi(A, jac, b, eps=1e-10, m=2, iter_max=0)
    assert ref["iters"] == 1 and ref["reason"] == 0 and ref["cycles"] == 3
    want = (105339456 / 17850625) ** 0.5
    assert abs(ref["hist"][0] - want) <= 1e-13 * want


def test_coo_accumulate_adds_entry_by_entry():
    """The COO product into a non-zero y is not the CSR one: y + (a1 x1 +
    a2 x2) rounds differently from (y + a1 x1) + a2 x2.  (Pins that the
    accumulating COO case of the GPU tests checks the entry order.)"""
    A = O.CSR(np.array([1, 3]), np.array([1, 2]), np.array([1.0, 1.0]))
    x = np.array([1.0, 1e-16])
    y0 = np.array([-1.0])
    coo = O.spmv_coo(np.array([1, 1]), A.cols, A.vals, 0, 0, x, y0)
    csr = O.spmv_csr(A, x, y0)
    assert coo[0] == 1e-16 and csr[0] == 0.0
