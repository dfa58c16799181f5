"""DIA form (cel_gpu_mat_to_dia; no reference counterpart) of CSR blocks
whose entries lie on a few diagonals.  Its products must be bit-identical to
the CSR products -- hence to the oracle's restatement of the reference CSR
kernel (src/fop/cel_fop_mv_module.f90:215-229) -- and a PCG solve with it
bit-identical to the CSR solve, FP64 and mixed."""
import numpy as np
import pytest

import oracle as O
from test_oracle import convdiff_csr

pytestmark = pytest.mark.gpu


def cases():
    rng = np.random.default_rng(61)
    kappa = np.exp(rng.normal(0.0, 0.7, 14 * 11 * 9))
    gold = np.load(O.__file__.replace("oracle/__init__.py", "tests/golden/ref_generator.npz"))
    out = [("poisson20x17x13", O.gen_poisson7(20, 17, 13)),
           ("diffusion14x11x9", O.gen_diffusion7(14, 11, 9, kappa)),
           ("convdiff12x10x8", convdiff_csr(12, 10, 8)),
           ("poisson1x1x5", O.gen_poisson7(1, 1, 5))]
    for base in ("s27_5x5x5_r0", "s19_4x5x6_r0"):  # the reference generator's matrices
        n = int(gold[base + "_xt"].size)
        r, c, v = O.coo_sort_rowcol(gold[base + "_rows"] + 1, gold[base + "_cols"] + 1,
                                    gold[base + "_vals"])  # stored 0-based
        out.append((f"refgen_{base}", O.CSR(O.coo_to_csr(n, r), c, v)))
    return out


def dev(gpu, A):
    return gpu.Matrix.from_csr(A.xadj, A.cols, A.vals)


@pytest.mark.parametrize("name,A", cases(), ids=[c[0] for c in cases()])
def test_dia_spmv_is_the_csr_product(gpu, name, A):
    print(f"case {name} n={A.n} nnz={A.nnz}", flush=True)
    M = dev(gpu, A)
    offs = M.to_dia()
    assert offs == sorted(offs) and len(offs) <= 32
    rng = np.random.default_rng(62)
    x = rng.normal(size=A.n)
    np.testing.assert_array_equal(M.spmv(x, fmt=gpu.FMT_DIA), O.spmv_csr(A, x))
    y0 = rng.normal(size=A.n)
    np.testing.assert_array_equal(M.spmv(x, y0, fmt=gpu.FMT_DIA, accumulate=True), O.spmv_csr(A, x, y0))
    x32 = x.astype(np.float32)
    np.testing.assert_array_equal(M.spmv_f32(x32, fmt=gpu.FMT_DIA), O.spmv_csr_f32(A, x32))


def test_dia_skips_absent_entries_even_for_non_finite_x(gpu):
    """Boundary rows read a clamped x for their absent diagonals and must
    discard it: Inf and NaN in x reach exactly the rows CSR sends them to."""
    A = O.gen_poisson7(9, 8, 7)
    M = dev(gpu, A)
    M.to_dia()
    x = np.random.default_rng(63).normal(size=A.n)
    x[0], x[-1], x[9] = np.inf, np.nan, -np.inf
    np.testing.assert_array_equal(M.spmv(x, fmt=gpu.FMT_DIA), M.spmv(x))


def test_dia_refuses_what_it_cannot_hold(gpu):
    rng = np.random.default_rng(64)
    n = 200
    rows = np.repeat(np.arange(n), 5)
    cols = rng.integers(0, n, rows.size)
    keys = np.unique(rows.astype(np.int64) * n + cols)
    r, c = keys // n, keys % n
    xadj = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=n))]) + 1
    M = gpu.Matrix.from_csr(xadj, c + 1, rng.normal(size=c.size))
    with pytest.raises(gpu.CelError) as e:
        M.to_dia()
    assert e.value.code == gpu.ERR_BAD_MATRIX
    with pytest.raises(gpu.CelError):
        M.spmv(np.ones(n), fmt=gpu.FMT_DIA)
    # a rectangular 4 x 8 block (an entry in column 6): the DIA row sums
    # read x as n entries, so the block has no DIA form; its CSR product
    # stays available
    R = gpu.Matrix.from_csr([1, 3, 4, 6, 7], [1, 7, 2, 3, 4, 4], [1.0, 2.0, 3.0, 4.0, 5.0, 6.0],
                            ncols=8)
    with pytest.raises(gpu.CelError) as e:
        R.to_dia()
    assert e.value.code == gpu.ERR_BAD_MATRIX
    x = np.arange(1.0, 9.0)
    np.testing.assert_array_equal(R.spmv(x), [1 + 14, 6, 12 + 20, 24])


def test_scaling_drops_the_dia_form(gpu):
    A = O.gen_poisson7(10, 9, 8)
    M = dev(gpu, A)
    M.to_dia()
    M.scale_diagonal()
    x = np.random.default_rng(65).normal(size=A.n)
    with pytest.raises(gpu.CelError):
        M.spmv(x, fmt=gpu.FMT_DIA)
    M.to_dia()
    np.testing.assert_array_equal(M.spmv(x, fmt=gpu.FMT_DIA), M.spmv(x))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pcg_with_dia_is_bitwise_the_csr_solve(gpu, precision):
    rng = np.random.default_rng(66)
    kappa = np.exp(rng.normal(0.0, 0.5, 40 * 36 * 32))
    M = gpu.Matrix.diffusion7(40, 36, 32, kappa)
    M.set_jacobi()
    b = rng.random(M.n)
    prec = gpu.PREC_FP64 if precision == "fp64" else gpu.PREC_MIXED
    kw = dict(tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, precision=prec)
    a = M.pcg(b, **kw)
    M.to_dia()
    d = M.pcg(b, fmt=gpu.FMT_DIA, **kw)
    assert a["status"] == d["status"] == 0
    assert (a["iters"], a["outer_iters"]) == (d["iters"], d["outer_iters"])
    np.testing.assert_array_equal(a["x"], d["x"])
    assert a["final_norm"] == d["final_norm"]


def test_pcg_dia_with_halo_faces_is_bitwise_the_csr_solve(gpu):
    """The centre block of a 3x3x3 split (six halo faces, face packing every
    iteration; edge rows carry two halo entries, corner rows three, which
    take the compact lists): the DIA K1 adds the halo products of a row
    after its local sum, in block order, as the CSR K1 does.  The two halo
    solves, and the local solve (the ghost region of one subdomain is
    zero), agree bit for bit, and the plain DIA product is unchanged."""
    from test_gpu_distr import block_major_poisson

    m, d, rank = 12, 3, 13
    vtx, rows, cols, vals = block_major_poisson(m, d, rank)
    M = gpu.Matrix.distr_split(vtx, rank, rows, cols, vals)
    for blk_id in range(1, 7):
        blk = M.get_block(blk_id)
        M.set_send_list(blk["proc"], np.unique(blk["rows"] - vtx[rank] + 1))
    M.set_jacobi()
    assert len(M.to_dia()) == 7
    rng = np.random.default_rng(67)
    b = rng.random(M.n)
    kw = dict(tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=2000, history=True)
    c = M.pcg(b, use_halo=1, **kw)
    h = M.pcg(b, use_halo=1, fmt=gpu.FMT_DIA, **kw)
    loc = M.pcg(b, fmt=gpu.FMT_DIA, **kw)
    assert c["status"] == h["status"] == loc["status"] == 0
    assert c["iters"] == h["iters"] == loc["iters"]
    np.testing.assert_array_equal(c["hist"], h["hist"])
    np.testing.assert_array_equal(c["x"], h["x"])
    np.testing.assert_array_equal(h["x"], loc["x"])
    x = rng.normal(size=M.n)
    np.testing.assert_array_equal(M.spmv(x, fmt=gpu.FMT_DIA), M.spmv(x))
    # the halo path is FP64 (CSR or DIA form)
    with pytest.raises(gpu.CelError):
        M.pcg(b, use_halo=1, fmt=gpu.FMT_DIA, precision=gpu.PREC_MIXED, **kw)
