"""GPU parity of the COO ordering and COO -> CSR (cel_gpu_coo_sort,
cel_gpu_coo_row_offsets, cel_gpu_mat_create_coo_sorted) and of the Matrix
Market -> matrix pipeline, against the oracle's restatement of
cel_sp_mat_format_sparse_31_sort and cel_sp_mat_to_csr
(src/fspmt/cel_sp_mat_module.f90:518-585, :430-456).  The oracle itself is
pinned by hand traces and by the reference's own routine compiled from its
sources (tests/test_oracle.py).  Integer arrays and the carried values must
be bit-exact, and so must the status (the reference's result is undefined
where its loop reads past its arrays)."""
import numpy as np
import pytest

import oracle as O
from test_oracle import SORT_TRACES, random_coo_case

pytestmark = pytest.mark.gpu


def _check_same(got, want):
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)


@pytest.mark.parametrize("name", sorted(SORT_TRACES))
def test_hand_traces(gpu, name):
    nrows, rows, cols, st_want, r_want, c_want, p_want = SORT_TRACES[name]
    vals = np.arange(len(rows), dtype=np.float64)
    st, r, c, v = gpu.coo_sort(nrows, rows, cols, vals, ncols=max(cols, default=0) or 1,
                               allow_undefined=True)
    assert st == (0 if st_want == 0 else gpu.ERR_REF_UNDEFINED)
    if st_want == 0:
        _check_same((r, c, v), (r_want, c_want, p_want))
    else:
        order = np.argsort(np.asarray(rows), kind="stable")
        np.testing.assert_array_equal(v, vals[order])


def test_random_cases_both_modes(gpu):
    rng = np.random.default_rng(31)
    undefined = 0
    for t in range(300):
        n, rows, cols, vals = random_coo_case(rng, all_rows=t % 3 == 0)
        nc = n + 3
        print(f"case {t} n={n} nnz={rows.size}", flush=True)
        st0, r0, c0, v0 = O.coo_sort_cel(n, rows, cols, vals)
        st, r, c, v = gpu.coo_sort(n, rows, cols, vals, ncols=nc, allow_undefined=True)
        assert st == (0 if st0 == 0 else gpu.ERR_REF_UNDEFINED), (t, st, st0)
        undefined += st != 0
        _check_same((r, c, v), (r0, c0, v0))
        r1, c1, v1 = O.coo_sort_rowcol(rows, cols, vals)
        _, r, c, v = gpu.coo_sort(n, rows, cols, vals, ncols=nc, mode=gpu.SORT_ROW_COL)
        _check_same((r, c, v), (r1, c1, v1))
    assert undefined > 20


def test_offsets_base_zero_and_device_arrays(gpu):
    """Global row/column numbers of a block (row_offset, col_offset), 0-based
    input, arrays already in HBM: same order as the oracle on the 1-based
    local form.  A row offset makes the first row value differ from 1, so
    the reference's first-entry quirk applies."""
    rng = np.random.default_rng(32)
    n, nc, nnz = 40, 50, 400
    rows = np.concatenate([np.arange(n), rng.integers(0, n, nnz - n)]).astype(np.int32)
    rng.shuffle(rows)
    cols = rng.integers(0, nc, nnz).astype(np.int32)
    vals = rng.normal(size=nnz)
    roff, coff = 1000, 7
    g_rows, g_cols = rows + roff, cols + coff  # 0-based global
    d = [gpu.DeviceArray.from_host(g_rows), gpu.DeviceArray.from_host(g_cols),
         gpu.DeviceArray.from_host(vals)]
    st, *_ = gpu.coo_sort(n, *d, ncols=nc, index_base=0, row_offset=roff, col_offset=coff)
    got = [a.to_host() for a in d]
    # the oracle holds 1-based values: local row + row_offset + 1
    st1, r1, c1, v1 = O.coo_sort_cel(n, rows + 1 + roff, cols + 1 + coff, vals)
    assert st == 0 and st1 == 0
    _check_same(got, (r1 - 1, c1 - 1, v1))
    # and the quirk did fire: the first entry stayed out of its row's sort
    order = np.argsort(rows, kind="stable")
    assert got[1][0] == g_cols[order[0]]


def test_out_of_range_and_bad_arguments(gpu):
    with pytest.raises(gpu.CelError) as e:
        gpu.coo_sort(3, [1, 4], [1, 1], [1.0, 2.0], ncols=3)
    assert e.value.code == gpu.ERR_BAD_MATRIX
    with pytest.raises(gpu.CelError) as e:
        gpu.coo_sort(3, [1, 2], [0, 1], [1.0, 2.0], ncols=3)
    assert e.value.code == gpu.ERR_BAD_MATRIX
    with pytest.raises(gpu.CelError):
        gpu.coo_sort(3, [1, 2], [1, 1], [1.0, 2.0], mode=7)
    st, r, c, v = gpu.coo_sort(3, [], [], [], mode=gpu.SORT_ROW_COL)
    assert st == 0 and r.size == 0
    with pytest.raises(gpu.CelError) as e:
        gpu.coo_sort(3, [], [], [])
    assert e.value.code == gpu.ERR_REF_UNDEFINED


def test_row_offsets(gpu):
    rng = np.random.default_rng(33)
    for n, nnz in ((1, 0), (1, 5), (17, 0), (17, 100), (1000, 20000)):
        rows = np.sort(rng.integers(1, n + 1, nnz)).astype(np.int32)
        np.testing.assert_array_equal(gpu.coo_row_offsets(n, rows), O.coo_to_csr(n, rows))
        x0 = gpu.coo_row_offsets(n, rows - 1, index_base=0)
        np.testing.assert_array_equal(x0, O.coo_to_csr(n, rows) - 1)
    # the reference counts without moving entries: unsorted rows give counts
    rows = np.array([3, 1, 3, 2], np.int32)
    np.testing.assert_array_equal(gpu.coo_row_offsets(3, rows), [1, 2, 3, 5])
    with pytest.raises(gpu.CelError) as e:
        gpu.coo_row_offsets(3, [1, 4])
    assert e.value.code == gpu.ERR_BAD_MATRIX


def shuffled_coo(A, rng):
    rows = np.repeat(np.arange(1, A.n + 1), np.diff(A.xadj)).astype(np.int32)
    order = rng.permutation(A.nnz)
    return rows[order], A.cols[order].copy(), A.vals[order].copy()


def test_shuffled_poisson_full_size(gpu):
    """64^3 7-point matrix (1.8M entries) in random order: the (row, col)
    sort restores the generator's CSR exactly; the reference order equals
    the oracle's entry for entry."""
    A = O.gen_poisson7(64, 64, 64)
    rng = np.random.default_rng(34)
    vals_tag = np.arange(A.nnz, dtype=np.float64)  # carry positions
    rows, cols, perm = shuffled_coo(O.CSR(A.xadj, A.cols, vals_tag), rng)
    _, r, c, v = gpu.coo_sort(A.n, rows, cols, perm, mode=gpu.SORT_ROW_COL)
    want_rows = np.repeat(np.arange(1, A.n + 1), np.diff(A.xadj))
    _check_same((r, c, v), (want_rows, A.cols, vals_tag))
    st0, r0, c0, v0 = O.coo_sort_cel(A.n, rows, cols, perm)
    st, r, c, v = gpu.coo_sort(A.n, rows, cols, perm)
    assert st == st0 == 0
    _check_same((r, c, v), (r0, c0, v0))
    # every row but the last comes out column-sorted
    last = A.xadj[-2] - 1
    np.testing.assert_array_equal(c[:last], A.cols[:last])


def test_workspace_kept_between_sorts_of_changing_size(gpu):
    """The sort keeps its device workspace between calls: a large sort, a
    small one (the workspace holds the large one's leftovers), a larger one
    again (the workspace grows), and the small one repeated; each equals the
    oracle entry for entry, in both orders."""
    rng = np.random.default_rng(35)
    big = O.gen_poisson7(40, 40, 40)
    bigger = O.gen_poisson7(48, 48, 48)
    small = O.gen_poisson7(5, 6, 7)
    for A in (big, small, bigger, small):
        tag = np.arange(A.nnz, dtype=np.float64)
        rows, cols, perm = shuffled_coo(O.CSR(A.xadj, A.cols, tag), rng)
        print(f"n={A.n} nnz={A.nnz}", flush=True)
        st0, r0, c0, v0 = O.coo_sort_cel(A.n, rows, cols, perm)
        st, r, c, v = gpu.coo_sort(A.n, rows, cols, perm)
        assert st == st0 == 0
        _check_same((r, c, v), (r0, c0, v0))
        _, r, c, v = gpu.coo_sort(A.n, rows, cols, perm, mode=gpu.SORT_ROW_COL)
        want_rows = np.repeat(np.arange(1, A.n + 1), np.diff(A.xadj))
        _check_same((r, c, v), (want_rows, A.cols, tag))


def test_create_coo_sorted_matches_oracle_pipeline(gpu):
    """Sort + cel_sp_mat_to_csr + build in HBM, then SpMV, equals the
    oracle's sort + to_csr + CSR SpMV bitwise (both modes)."""
    rng = np.random.default_rng(35)
    kappa = np.exp(rng.normal(0.0, 0.5, 20 * 18 * 16))
    A = O.gen_diffusion7(20, 18, 16, kappa)
    rows, cols, vals = shuffled_coo(A, rng)
    x = rng.normal(size=A.n)
    for mode in (gpu.SORT_CEL_COMPAT, gpu.SORT_ROW_COL):
        if mode == gpu.SORT_CEL_COMPAT:
            st, r0, c0, v0 = O.coo_sort_cel(A.n, rows, cols, vals)
            assert st == 0
        else:
            r0, c0, v0 = O.coo_sort_rowcol(rows, cols, vals)
        B = O.CSR(O.coo_to_csr(A.n, r0), c0, v0)
        M = gpu.Matrix.from_coo_sorted(A.n, rows, cols, vals, mode=mode)
        xa, ca, va = M.get_csr()
        _check_same((xa, ca, va), (B.xadj, B.cols, B.vals))
        np.testing.assert_array_equal(M.spmv(x), O.spmv_csr(B, x))
        dev = [gpu.DeviceArray.from_host(a) for a in (rows, cols, vals)]
        M2 = gpu.Matrix.from_coo_sorted(A.n, *dev, mode=mode)
        _check_same(M2.get_csr(), (B.xadj, B.cols, B.vals))


def test_create_coo_sorted_refuses_what_cannot_be_a_csr(gpu):
    # the reference's sort moves the one entry of the last row among row 2
    n, rows, cols, *_ = SORT_TRACES["single_last_row_merged"]
    with pytest.raises(gpu.CelError) as e:
        gpu.Matrix.from_coo_sorted(n, rows, cols, np.ones(len(rows)))
    assert e.value.code == gpu.ERR_BAD_MATRIX
    M = gpu.Matrix.from_coo_sorted(n, rows, cols, np.ones(len(rows)), mode=gpu.SORT_ROW_COL)
    assert M.nnz == len(rows)
    n, rows, cols, *_ = SORT_TRACES["reads_past_end"]
    with pytest.raises(gpu.CelError) as e:
        gpu.Matrix.from_coo_sorted(n, rows, cols, np.ones(len(rows)))
    assert e.value.code == gpu.ERR_REF_UNDEFINED


def test_matrix_market_file_to_pcg(gpu, tmp_path):
    """A symmetric 24^3 diffusion matrix stored as its lower triangle in
    shuffled order, 17-digit values: read, expand, sort and build in HBM;
    the matrix is the generator's exactly and the PCG solve is bitwise the
    solve on the generated matrix."""
    shape = (24, 24, 24)
    n = int(np.prod(shape))
    rng = np.random.default_rng(36)
    kappa = np.exp(rng.normal(0.0, 0.5, n))
    A = O.gen_diffusion7(*shape, kappa)
    rows = np.repeat(np.arange(1, n + 1), np.diff(A.xadj))
    low = A.cols <= rows
    r, c, v = rows[low], A.cols[low], A.vals[low]
    order = rng.permutation(r.size)
    lines = "\n".join(f"{a} {b} {float(x)!r}" for a, b, x in zip(r[order], c[order], v[order]))
    path = tmp_path / "diff24.mtx"
    path.write_text("%%MatrixMarket matrix coordinate real symmetric\n% 7-point diffusion\n"
                    f"{n} {n} {r.size}\n{lines}\n")
    M = gpu.Matrix.from_mm(str(path), mode=gpu.SORT_ROW_COL, expand_symmetric=True)
    _check_same(M.get_csr(), (A.xadj, A.cols, A.vals))
    G = gpu.Matrix.diffusion7(*shape, kappa)
    b = rng.random(n)
    outs = []
    for X in (M, G):
        X.set_jacobi()
        outs.append(X.pcg(b, tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO))
    assert outs[0]["iters"] == outs[1]["iters"]
    np.testing.assert_array_equal(outs[0]["x"], outs[1]["x"])
    # the reference's order on the same file: as the oracle pipeline gives it
    m = gpu.mm_read(str(path), expand_symmetric=True)
    st, r0, c0, v0 = O.coo_sort_cel(n, m["rows"], m["cols"], m["vals"])
    Mc = gpu.Matrix.from_mm(str(path), expand_symmetric=True)
    _check_same(Mc.get_csr(), (O.coo_to_csr(n, r0), c0, v0))


def test_long_groups_take_the_lds_and_radix_paths(gpu):
    """Rows longer than one thread's sort (33..4096 entries: one workgroup
    each) and longer than the LDS sort (> 4096: one radix sort over all
    groups), with duplicate columns, against the oracle in both modes."""
    rng = np.random.default_rng(37)
    for lens in ((40, 3, 700, 5, 33, 4096, 2), (3, 5000, 7, 64), (9000,), (1, 1, 1, 35, 1)):
        n = len(lens)
        rows = np.repeat(np.arange(1, n + 1), lens).astype(np.int32)
        order = rng.permutation(rows.size)
        rows = rows[order]
        cols = rng.integers(1, 300, rows.size).astype(np.int32)
        vals = rng.normal(size=rows.size)
        print(f"lengths {lens}", flush=True)
        st0, r0, c0, v0 = O.coo_sort_cel(n, rows, cols, vals)
        st, r, c, v = gpu.coo_sort(n, rows, cols, vals, ncols=300, allow_undefined=True)
        assert st == (0 if st0 == 0 else gpu.ERR_REF_UNDEFINED)
        _check_same((r, c, v), (r0, c0, v0))
        r1, c1, v1 = O.coo_sort_rowcol(rows, cols, vals)
        _, r, c, v = gpu.coo_sort(n, rows, cols, vals, ncols=300, mode=gpu.SORT_ROW_COL)
        _check_same((r, c, v), (r1, c1, v1))


def test_dense_short_rows_overflow_the_lds_tile(gpu):
    """600 rows of 20..32 entries: 256 consecutive groups hold more entries
    than one workgroup's LDS tile, so the per-thread path in global memory
    sorts them; 200 rows of 3 entries around them take the tile path."""
    rng = np.random.default_rng(38)
    lens = np.concatenate([rng.integers(1, 4, 200), rng.integers(20, 33, 600),
                           rng.integers(1, 4, 200)])
    n = lens.size
    rows = np.repeat(np.arange(1, n + 1), lens).astype(np.int32)
    rows = rows[rng.permutation(rows.size)]
    cols = rng.integers(1, 50, rows.size).astype(np.int32)
    vals = rng.normal(size=rows.size)
    st0, r0, c0, v0 = O.coo_sort_cel(n, rows, cols, vals)
    st, r, c, v = gpu.coo_sort(n, rows, cols, vals, ncols=50)
    assert st == st0 == 0
    _check_same((r, c, v), (r0, c0, v0))
    r1, c1, v1 = O.coo_sort_rowcol(rows, cols, vals)
    _, r, c, v = gpu.coo_sort(n, rows, cols, vals, ncols=50, mode=gpu.SORT_ROW_COL)
    _check_same((r, c, v), (r1, c1, v1))
