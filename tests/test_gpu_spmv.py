"""GPU parity of the SpMV kernels, the device generators, the CSR->JAD
converter and the Jacobi/scaling set-up, against the CPU oracle.  Integer
and index arrays must be bit-exact; FP64/FP32 SpMV results must be bit-exact
too, because every row is summed in the reference's column order with
products rounded before the add on both sides."""
import numpy as np
import pytest

import oracle as O
from helpers import spmv_matrices

pytestmark = pytest.mark.gpu

CASES = spmv_matrices()


def dev_matrix(cel, A: O.CSR):
    return cel.Matrix.from_csr(A.xadj, A.cols, A.vals, index_base=1)


@pytest.mark.parametrize("name,A", CASES, ids=[c[0] for c in CASES])
def test_csr_spmv_bitwise(gpu, name, A):
    print(f"case {name} n={A.n} nnz={A.nnz}", flush=True)
    M = dev_matrix(gpu, A)
    rng = np.random.default_rng(1)
    x = rng.normal(size=A.n)
    np.testing.assert_array_equal(M.spmv(x), O.spmv_csr(A, x))
    y0 = rng.normal(size=A.n)
    np.testing.assert_array_equal(M.spmv(x, y0, accumulate=True), O.spmv_csr(A, x, y0))
    xa, ca, va = M.get_csr()
    np.testing.assert_array_equal(xa, A.xadj)
    np.testing.assert_array_equal(ca, A.cols)
    np.testing.assert_array_equal(va, A.vals)


@pytest.mark.parametrize("name,A", CASES, ids=[c[0] for c in CASES])
def test_coo_spmv_bitwise(gpu, name, A):
    """CEL_FMT_COO (cel_fop_mv_module.f90:357-403): y(row) = y(row) + a x for
    every entry in the block's row-sorted order, from y = 0 or, when
    accumulating, from the caller's y; bitwise the oracle's entry loop."""
    M = dev_matrix(gpu, A)
    rows = np.repeat(np.arange(1, A.n + 1), np.diff(A.xadj))
    rng = np.random.default_rng(4)
    x = rng.normal(size=A.n)
    np.testing.assert_array_equal(M.spmv(x, fmt=gpu.FMT_COO),
                                  O.spmv_coo(rows, A.cols, A.vals, 0, 0, x, np.zeros(A.n)))
    y0 = 1e3 * rng.normal(size=A.n)
    yc = O.spmv_coo(rows, A.cols, A.vals, 0, 0, x, y0)
    np.testing.assert_array_equal(M.spmv(x, y0, fmt=gpu.FMT_COO, accumulate=True), yc)
    # from 0 the entry-by-entry sum is the CSR row sum; FP32 as well
    np.testing.assert_array_equal(M.spmv(x, fmt=gpu.FMT_COO), M.spmv(x))
    xf = x.astype(np.float32)
    np.testing.assert_array_equal(M.spmv_f32(xf, fmt=gpu.FMT_COO), O.spmv_csr_f32(A, xf))


@pytest.mark.parametrize("name,A", CASES, ids=[c[0] for c in CASES])
def test_csr_spmv_f32_bitwise(gpu, name, A):
    M = dev_matrix(gpu, A)
    x = np.random.default_rng(2).normal(size=A.n).astype(np.float32)
    np.testing.assert_array_equal(M.spmv_f32(x), O.spmv_csr_f32(A, x))


@pytest.mark.parametrize("name,A", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("mode", [O.JAD_SPARSKIT, O.JAD_CEL_COMPAT])
def test_jad_conversion_bit_exact(gpu, name, A, mode):
    M = dev_matrix(gpu, A)
    ref = O.jad_convert(A, mode)
    defined = ref["ndefined"] == ref["mmax"]
    rc = M.to_jad(mode, allow_undefined=True)
    assert (rc == 0) == defined
    j = M.get_jad()
    assert j["mmax"] == ref["mmax"] and j["ndefined"] == ref["ndefined"]
    np.testing.assert_array_equal(j["perm"], ref["perm"])
    np.testing.assert_array_equal(j["begin"], ref["begin"])
    np.testing.assert_array_equal(j["cols"], ref["cols"])
    np.testing.assert_array_equal(j["vals"], ref["vals"])
    if defined:
        x = np.random.default_rng(3).normal(size=A.n)
        y = M.spmv(x, fmt=gpu.FMT_JAD)
        np.testing.assert_array_equal(y, O.spmv_jad(ref, x))
        # the reference's own counts drop entries when the shortest row is
        # non-empty (a 1-row matrix: ilo = 1, diagonal 1 gets #{len > 1} = 0)
        if mode == O.JAD_SPARSKIT or ref["ilo"] == 0:
            np.testing.assert_array_equal(y, O.spmv_csr(A, x))
            xf = x.astype(np.float32)
            np.testing.assert_array_equal(M.spmv_f32(xf, fmt=gpu.FMT_JAD), O.spmv_csr_f32(A, xf))


def test_jad_undefined_reference_counts_refuse_spmv(gpu):
    A = O.gen_poisson7(4, 4, 4)
    M = dev_matrix(gpu, A)
    assert M.to_jad(O.JAD_CEL_COMPAT, allow_undefined=True) == -9
    j = M.get_jad()
    np.testing.assert_array_equal(j["begin"][:5] - j["begin"][0], [0, 56, 88, 96, 96])
    assert list(j["begin"][5:]) == [-1, -1, -1]
    with pytest.raises(gpu.CelError):
        M.spmv(np.ones(A.n), fmt=gpu.FMT_JAD)


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 1, 1), (5, 4, 3), (32, 32, 32), (31, 17, 65)])
def test_device_poisson7_generator(gpu, shape):
    A = O.gen_poisson7(*shape)
    M = gpu.Matrix.poisson7(*shape)
    assert (M.n, M.nnz) == (A.n, A.nnz)
    xa, ca, va = M.get_csr()
    np.testing.assert_array_equal(xa, A.xadj)
    np.testing.assert_array_equal(ca, A.cols)
    np.testing.assert_array_equal(va, A.vals)


@pytest.mark.parametrize("shape", [(2, 3, 4), (16, 9, 7)])
def test_device_diffusion7_generator(gpu, shape):
    k = np.exp(np.random.default_rng(4).normal(0, 1, int(np.prod(shape))))
    A = O.gen_diffusion7(*shape, k)
    M = gpu.Matrix.diffusion7(*shape, k)
    xa, ca, va = M.get_csr()
    np.testing.assert_array_equal(xa, A.xadj)
    np.testing.assert_array_equal(ca, A.cols)
    np.testing.assert_array_equal(va, A.vals)


def test_poisson7_256_properties(gpu):
    """Full-size matrix: nnz closed form and A*1 = 6 - #neighbours per row."""
    n = 256
    M = gpu.Matrix.poisson7(n, n, n)
    assert M.n == n ** 3 and M.nnz == 7 * n ** 3 - 6 * n ** 2
    x = gpu.DeviceArray.from_host(np.ones(M.n))
    y = gpu.DeviceArray(M.n)
    M.spmv(x, y)
    yh = y.to_host()
    idx = np.arange(M.n)
    xs, ys, zs = idx % n, (idx // n) % n, idx // (n * n)
    bnd = (xs == 0).astype(int) + (xs == n - 1) + (ys == 0) + (ys == n - 1) + (zs == 0) + (zs == n - 1)
    np.testing.assert_array_equal(yh, bnd.astype(np.float64))
    M.to_jad()
    j = M.get_jad()
    counts = np.diff(j["begin"])
    # diagonal d holds #{rows with >= d entries}: all rows have 4..7
    nint = (n - 2) ** 3
    nface = 6 * (n - 2) ** 2
    nedge = 12 * (n - 2)
    np.testing.assert_array_equal(counts, [n ** 3] * 4 + [n ** 3 - 8, nint + nface, nint])
    del nedge


@pytest.mark.parametrize("name,A", CASES[:4], ids=[c[0] for c in CASES[:4]])
@pytest.mark.parametrize("scaled", [False, True])
def test_jacobi_and_row_scaling(gpu, name, A, scaled):
    M = dev_matrix(gpu, A)
    d = O.csr_diagonal(A)
    np.testing.assert_array_equal(M.diagonal(), d)
    M.set_jacobi(row_scaled=scaled)
    np.testing.assert_array_equal(M.jacobi(), O.set_jacobi(d, scaled))
    if scaled:
        M.scale_diagonal()
        S = O.scale_csr(A, d)
        np.testing.assert_array_equal(M.get_csr()[2], S.vals)
        b = np.random.default_rng(5).normal(size=A.n)
        np.testing.assert_array_equal(M.scale_vec(b), O.scale_vec(b, d))


def test_zero_diagonal_is_reported(gpu):
    A = O.gen_poisson7(3, 3, 3)
    vals = A.vals.copy()
    vals[A.xadj[4] - 1 + 3] = 0.0  # diagonal of row 5 (interior-free corner rows have 4 entries)
    M = gpu.Matrix.from_csr(A.xadj, A.cols, vals)
    d = M.diagonal()
    if (d == 0).any():
        with pytest.raises(gpu.CelError) as e:
            M.set_jacobi()
        assert e.value.code == -5


def test_bad_matrix_inputs_are_rejected(gpu):
    A = O.gen_poisson7(3, 3, 3)
    cols = A.cols.copy()
    cols[5] = A.n + 7
    with pytest.raises(gpu.CelError) as e:
        gpu.Matrix.from_csr(A.xadj, cols, A.vals)
    assert e.value.code == -4
    with pytest.raises(gpu.CelError) as e:
        gpu.Matrix.from_coo(3, [1, 3, 2], [1, 3, 2], [1.0, 1.0, 1.0])
    assert e.value.code == -4


def test_coo_create_matches_to_csr(gpu):
    d = O.gen_poisson7(5, 4, 3)
    rows = np.repeat(np.arange(1, d.n + 1), np.diff(d.xadj))
    M = gpu.Matrix.from_coo(d.n, rows, d.cols, d.vals)
    xa, ca, va = M.get_csr()
    np.testing.assert_array_equal(xa, O.coo_to_csr(d.n, rows, 0))
    np.testing.assert_array_equal(ca, d.cols)


def test_jad_256_conversion_and_spmv(gpu):
    """The 256^3 block of the bench converted on the device: jad_perm is the
    stable descending sort of the row lengths (ties in ascending row order),
    diagonal d holds entry d of every row in perm order, and the JAD SpMV
    equals the CSR SpMV bit for bit (same per-row order)."""
    n = 256
    M = gpu.Matrix.poisson7(n, n, n)
    M.to_jad()
    j = M.get_jad()
    xadj, cols, vals = M.get_csr()
    lens = np.diff(xadj)
    perm0 = j["perm"] - 1
    np.testing.assert_array_equal(perm0, np.argsort(-lens, kind="stable"))
    begin0 = j["begin"] - 1
    for d in range(j["mmax"]):
        k = begin0[d + 1] - begin0[d]
        rows = perm0[:k]
        assert (lens[rows] > d).all() and (k == M.n or lens[perm0[k]] <= d)
        src = xadj[rows] - 1 + d
        np.testing.assert_array_equal(j["cols"][begin0[d]:begin0[d + 1]], cols[src])
        np.testing.assert_array_equal(j["vals"][begin0[d]:begin0[d + 1]], vals[src])
    x = np.random.default_rng(5).normal(size=M.n)
    np.testing.assert_array_equal(M.spmv(x, fmt=gpu.FMT_JAD), M.spmv(x))


@pytest.mark.parametrize("name,A", CASES, ids=[c[0] for c in CASES])
def test_fused_product_equals_the_oracle_and_refuses_bad_forms(gpu, name, A):
    """cel_gpu_spmv_dot (the PCG's fused K1 on caller vectors): y bitwise the
    oracle's row sums, the dot product x.y to rounding (its grouping is the
    kernel's block tree); a JAD form, ghost values without halo blocks and a
    DIA form that was never built are refused."""
    M = dev_matrix(gpu, A)
    x = np.random.default_rng(5).normal(size=A.n)
    y, dot = M.spmv_dot(x)
    want = O.spmv_csr(A, x)
    np.testing.assert_array_equal(y, want)
    np.testing.assert_allclose(dot, float(np.dot(x, want)), rtol=1e-12,
                               atol=1e-12 * float(np.abs(x * want).sum()))
    with pytest.raises(gpu.CelError):
        M.spmv_dot(x, fmt=gpu.FMT_JAD)
    with pytest.raises(gpu.CelError):
        M.spmv_dot(x, ghost=np.zeros(4))
    with pytest.raises(gpu.CelError):
        M.spmv_dot(x, fmt=gpu.FMT_DIA)


def test_stream_read_time_is_a_plausible_rate_and_checks_its_arguments(gpu):
    """cel_gpu_stream_read_time: a 256 MiB read in 1 and 4 arrays runs at
    between 1 and 9 TB/s (the part's peak is 8 TB/s); out-of-range arguments
    are refused."""
    nbytes = 256 << 20
    for streams in (1, 4):
        ms = gpu.stream_read_time(nbytes, streams=streams, blocks_per_cu=4, reps=3)
        rate = nbytes / (ms * 1e-3)
        print(f"{streams} arrays: {rate / 1e12:.2f} TB/s", flush=True)
        assert 1e12 < rate < 9e12
    for bad in (dict(streams=0), dict(streams=9), dict(blocks_per_cu=0), dict(blocks_per_cu=33),
                dict(reps=0)):
        kw = dict(streams=1, blocks_per_cu=4, reps=1)
        kw.update(bad)
        with pytest.raises(gpu.CelError):
            gpu.stream_read_time(nbytes, **kw)
    with pytest.raises(gpu.CelError):
        gpu.stream_read_time(8, streams=1)
