"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to check results against.  The product library
never imports this package.

Every index array that crosses this module is 1-based, exactly as in the
reference Fortran (xadj[0] == 1).  The C sources (cel_oracle.c) cite the
reference file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CEL_ORACLE_LIB names another build of the same source (the sanitizer run)
_LIB = os.environ.get("CEL_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libcel_base_ref.so")
_REF_MM = os.path.join(_HERE, "_ref", "libcel_mm_ref.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")

X0_ONE, X0_ZERO, X0_USER = 0, 1, 2
JAD_SPARSKIT, JAD_CEL_COMPAT = 0, 1


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is mounted)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if os.path.isdir("/root/reference/src/base"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


def _declare(L):
    L.orc_real_small.restype = C.c_double
    L.orc_work_distr.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_poisson7_nnz.restype = C.c_longlong
    L.orc_poisson7_nnz.argtypes = [C.c_int] * 3
    L.orc_gen_poisson7.argtypes = [C.c_int] * 3 + [_i32p, _i32p, _f64p]
    L.orc_gen_diffusion7.argtypes = [C.c_int] * 3 + [_f64p, _i32p, _i32p, _f64p]
    L.orc_coo_to_csr.argtypes = [C.c_int, C.c_int, _i32p, C.c_int, _i32p]
    L.orc_csr_diagonal.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int, _f64p]
    L.orc_spmv_csr.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int, _f64p, _f64p, C.c_int]
    L.orc_spmv_csr_f32.argtypes = [C.c_int, _i32p, _i32p, _f32p, C.c_int, _f32p, _f32p]
    L.orc_spmv_jad.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _i32p, _f64p, _f64p]
    L.orc_spmv_coo.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int, C.c_int, _f64p, _f64p]
    L.orc_jad_dcsort.argtypes = [_i32p, C.c_int, _i32p, _i32p, C.c_int, C.c_int]
    L.orc_jad_convert.restype = C.c_int
    L.orc_jad_convert.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int, C.c_int, _i32p, _i32p,
                                  _i32p, _f64p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_longlong)]
    L.orc_set_jacobi.argtypes = [C.c_int, _f64p, C.c_int, _f64p]
    L.orc_scale_csr.argtypes = [C.c_int, _i32p, _f64p, _f64p]
    L.orc_scale_coo.argtypes = [C.c_int, _i32p, C.c_int, _f64p, _f64p]
    L.orc_scale_vec.argtypes = [C.c_int, _f64p, _f64p]
    L.orc_pcg_jacobi.restype = C.c_int
    L.orc_pcg_jacobi.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int, _f64p, _f64p, _f64p,
                                 C.c_int, C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.POINTER(C.c_double)]
    L.orc_owner_of.restype = C.c_int
    L.orc_owner_of.argtypes = [C.c_int, _i32p, C.c_int]
    L.orc_distr_split.restype = C.c_int
    L.orc_distr_split.argtypes = [C.c_int, _i32p, C.c_int, C.c_int, _i32p, _i32p, _f64p,
                                  _i32p, _i32p, _i32p, _i32p, _f64p, C.POINTER(C.c_int), _i32p,
                                  _i32p, _i32p, _i32p, _f64p]
    L.orc_face_pack.argtypes = [C.c_int, _i32p, _f64p, _f64p]
    L.orc_gmres_jacobi.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int, _f64p, _f64p, _f64p,
                                   C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   _f64p, C.c_int]
    L.orc_coo_sort_cel.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p]
    L.orc_coo_sort_rowcol.argtypes = [C.c_int, _i32p, _i32p, _f64p]
    L.orc_sddmm_csr.argtypes = [C.c_int, _i32p, _i32p, C.c_int, _f64p, _f64p, _f64p]
    L.orc_diffusion7_kappa_grad.argtypes = [C.c_int] * 3 + [_f64p, _f64p, _f64p, _f64p]


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def real_small() -> float:
    return lib().orc_real_small()


def work_distr(length: int, workers: int, worker: int) -> tuple[int, int]:
    s, e = C.c_int(), C.c_int()
    lib().orc_work_distr(length, workers, worker, C.byref(s), C.byref(e))
    return s.value, e.value


class CSR:
    """1-based CSR block as the reference stores it."""

    def __init__(self, xadj, cols, vals, col_offset=0, row_offset=0):
        self.xadj = _i32(xadj)
        self.cols = _i32(cols)
        self.vals = _f64(vals)
        self.col_offset = int(col_offset)
        self.row_offset = int(row_offset)

    @property
    def n(self):
        return self.xadj.size - 1

    @property
    def nnz(self):
        return self.cols.size


def poisson7_nnz(nx, ny, nz) -> int:
    return int(lib().orc_poisson7_nnz(nx, ny, nz))


def gen_poisson7(nx, ny, nz) -> CSR:
    n = nx * ny * nz
    nnz = poisson7_nnz(nx, ny, nz)
    xadj = np.empty(n + 1, np.int32)
    cols = np.empty(nnz, np.int32)
    vals = np.empty(nnz, np.float64)
    lib().orc_gen_poisson7(nx, ny, nz, xadj, cols, vals)
    return CSR(xadj, cols, vals)


def gen_diffusion7(nx, ny, nz, kappa) -> CSR:
    n = nx * ny * nz
    nnz = poisson7_nnz(nx, ny, nz)
    xadj = np.empty(n + 1, np.int32)
    cols = np.empty(nnz, np.int32)
    vals = np.empty(nnz, np.float64)
    lib().orc_gen_diffusion7(nx, ny, nz, _f64(kappa), xadj, cols, vals)
    return CSR(xadj, cols, vals)


def coo_to_csr(nrows, rows, row_offset=0):
    rows = _i32(rows)
    xadj = np.empty(nrows + 1, np.int32)
    lib().orc_coo_to_csr(nrows, rows.size, rows, row_offset, xadj)
    return xadj


def csr_diagonal(A: CSR):
    d = np.zeros(A.n, np.float64)
    lib().orc_csr_diagonal(A.n, A.xadj, A.cols, A.vals, A.row_offset, d)
    return d


def spmv_csr(A: CSR, x, y=None, workers=1):
    """y += A x (y defaults to zeros)."""
    y = np.zeros(A.n) if y is None else _f64(y).copy()
    lib().orc_spmv_csr(A.n, A.xadj, A.cols, A.vals, A.col_offset, _f64(x), y, workers)
    return y


def spmv_csr_f32(A: CSR, x, vals32=None):
    vals32 = A.vals.astype(np.float32) if vals32 is None else np.ascontiguousarray(vals32, np.float32)
    y = np.zeros(A.n, np.float32)
    lib().orc_spmv_csr_f32(A.n, A.xadj, A.cols, vals32, A.col_offset,
                           np.ascontiguousarray(x, np.float32), y)
    return y


def spmv_jad(jad: dict, x, y=None):
    y = np.zeros(jad["perm"].size) if y is None else _f64(y).copy()
    lib().orc_spmv_jad(jad["perm"].size, jad["mmax"], jad["begin"], jad["cols"], jad["vals"],
                       jad["perm"], _f64(x), y)
    return y


def spmv_coo(rows, cols, vals, row_offset, col_offset, x, y):
    y = _f64(y).copy()
    rows, cols, vals = _i32(rows), _i32(cols), _f64(vals)
    lib().orc_spmv_coo(rows.size, rows, cols, vals, row_offset, col_offset, _f64(x), y)
    return y


def jad_dcsort(ival, ilo, ihi):
    ival = _i32(ival)
    icnt = np.zeros(max(ihi - ilo + 1, 1), np.int32)
    index = np.zeros(max(ival.size, 1), np.int32)
    lib().orc_jad_dcsort(ival, ival.size, icnt, index, ilo, ihi)
    return icnt, index[: ival.size]


def jad_convert(A: CSR, mode=JAD_SPARSKIT) -> dict:
    n = A.n
    lens = np.diff(A.xadj)
    mmax_ub = int(lens.max()) if n else 0
    perm = np.zeros(max(n, 1), np.int32)
    begin = np.zeros(mmax_ub + 1, np.int32)
    jcols = np.zeros(max(A.nnz, 1), np.int32)
    jvals = np.zeros(max(A.nnz, 1), np.float64)
    mmax, ilo, written = C.c_int(), C.c_int(), C.c_longlong()
    ndef = lib().orc_jad_convert(n, A.xadj, A.cols, A.vals, A.col_offset, mode, perm, begin,
                                 jcols, jvals, C.byref(mmax), C.byref(ilo), C.byref(written))
    w = written.value
    return dict(perm=perm[:n], begin=begin[: mmax.value + 1], cols=jcols[:w], vals=jvals[:w],
                mmax=mmax.value, ilo=ilo.value, ndefined=ndef, nmax=n)


def set_jacobi(diag, scaled=False):
    diag = _f64(diag)
    jac = np.empty_like(diag)
    lib().orc_set_jacobi(diag.size, diag, int(bool(scaled)), jac)
    return jac


def scale_csr(A: CSR, diag) -> CSR:
    vals = A.vals.copy()
    lib().orc_scale_csr(A.n, A.xadj, vals, _f64(diag))
    return CSR(A.xadj, A.cols, vals, A.col_offset, A.row_offset)


def scale_vec(b, diag):
    b = _f64(b).copy()
    lib().orc_scale_vec(b.size, b, _f64(diag))
    return b


def pcg_jacobi(A: CSR, jac, b, x0=None, x0_mode=X0_ONE, eps=1e-8, iter_max=1000, workers=1):
    """Reference-order Jacobi-PCG.  Returns dict(x, iters, norm, hist, alpha, beta)."""
    n = A.n
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    if x0 is not None:
        x0_mode = X0_USER
    hist = np.zeros(iter_max + 1)
    ah = np.zeros(iter_max + 1)
    sh = np.zeros(iter_max + 1)
    fn = C.c_double()
    it = lib().orc_pcg_jacobi(n, A.xadj, A.cols, A.vals, A.col_offset, _f64(jac), _f64(b), x,
                              x0_mode, eps, iter_max, workers, hist.ctypes.data,
                              ah.ctypes.data, sh.ctypes.data, C.byref(fn))
    return dict(x=x, iters=it, norm=fn.value, hist=hist[: it + 1], alpha=ah[: it + 1],
                beta=sh[: it + 1])


def gmres_jacobi(A: CSR, jac, b, x0=None, x0_mode=X0_ONE, eps=1e-8, iter_max=5000, m=3,
                 workers=1, hist_cap=100000):
    """cel_gmresalg_2jacobi restated (cel_oracle.h).  Returns dict(iters,
    x, norm, reason, cycles, hist)."""
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    if x0 is not None:
        x0_mode = X0_USER
    hist = np.zeros(max(hist_cap, 1))
    fn, rs, cy = C.c_double(), C.c_int(), C.c_int()
    it = lib().orc_gmres_jacobi(A.n, A.xadj, A.cols, A.vals, A.col_offset, _f64(jac), _f64(b), x,
                                x0_mode, eps, iter_max, m, workers, C.byref(fn), C.byref(rs),
                                C.byref(cy), hist, hist.size)
    return dict(iters=it, x=x, norm=fn.value, reason=rs.value, cycles=cy.value,
                hist=hist[:min(cy.value, hist.size)].copy())


def owner_of(vtxdist, col):
    vtxdist = _i32(vtxdist)
    return lib().orc_owner_of(vtxdist.size - 1, vtxdist, col)


def distr_split(vtxdist, my_rank, rows, cols, vals) -> dict:
    vtxdist, rows, cols, vals = _i32(vtxdist), _i32(rows), _i32(cols), _f64(vals)
    nranks = vtxdist.size - 1
    n = int(vtxdist[my_rank + 1] - vtxdist[my_rank])
    nnz = rows.size
    procs = np.zeros(nranks, np.int32)
    bnnz = np.zeros(nranks, np.int32)
    orow = np.zeros(max(nnz, 1), np.int32)
    ocol = np.zeros(max(nnz, 1), np.int32)
    oval = np.zeros(max(nnz, 1), np.float64)
    roff = C.c_int()
    coff = np.zeros(nranks, np.int32)
    xadj = np.zeros(nranks * (n + 1), np.int32)
    clen = np.zeros(nranks, np.int32)
    cperm = np.zeros(max(nnz, 1), np.int32)
    diag = np.zeros(max(n, 1), np.float64)
    nb = lib().orc_distr_split(nranks, vtxdist, my_rank, nnz, rows, cols, vals, procs, bnnz,
                               orow, ocol, oval, C.byref(roff), coff, xadj, clen, cperm, diag)
    if nb < 0:
        raise ValueError("column index out of range")
    blocks = []
    pos = 0
    ppos = 0
    for b in range(nb):
        k = int(bnnz[b])
        blk = dict(proc=int(procs[b]), rows=orow[pos:pos + k].copy(), cols=ocol[pos:pos + k].copy(),
                   vals=oval[pos:pos + k].copy(), col_offset=int(coff[b]),
                   xadj=xadj[b * (n + 1):(b + 1) * (n + 1)].copy(), comm_len=int(clen[b]))
        if b >= 1:
            blk["comm_perm"] = cperm[ppos:ppos + int(clen[b])].copy()
            ppos += int(clen[b])
        blocks.append(blk)
        pos += k
    return dict(blocks=blocks, row_offset=roff.value, diagonal=diag[:n], n=n)


def coo_sort_cel(nrows, rows, cols, vals):
    """cel_sp_mat_format_sparse_31_sort restated (cel_oracle.h).  Returns
    (status, rows, cols, vals) on copies: status 0, -1 (the reference reads
    past its arrays: undefined) or -2 (the reference stops: row below 1)."""
    r, c, v = _i32(rows).copy(), _i32(cols).copy(), _f64(vals).copy()
    st = lib().orc_coo_sort_cel(int(nrows), r.size, r, c, v)
    return st, r, c, v


def coo_sort_rowcol(rows, cols, vals):
    """Stable sort by (row, col) on copies."""
    r, c, v = _i32(rows).copy(), _i32(cols).copy(), _f64(vals).copy()
    lib().orc_coo_sort_rowcol(r.size, r, c, v)
    return r, c, v


def face_pack(perm, x):
    perm = _i32(perm)
    buf = np.empty(perm.size)
    lib().orc_face_pack(perm.size, perm, _f64(x), buf)
    return buf


def sddmm_csr(A: CSR, lam, x):
    g = np.empty(A.nnz)
    lib().orc_sddmm_csr(A.n, A.xadj, A.cols, A.col_offset, _f64(lam), _f64(x), g)
    return g


def diffusion7_kappa_grad(nx, ny, nz, kappa, x, lam):
    g = np.empty(nx * ny * nz)
    lib().orc_diffusion7_kappa_grad(nx, ny, nz, _f64(kappa), _f64(x), _f64(lam), g)
    return g


# ---------------------------------------------------------------- reference generator

class _Domain(C.Structure):
    _fields_ = [(f, C.c_int) for f in ("block_indx", "nx", "ny", "nz", "dx", "dy", "dz", "blocks",
                                       "bx", "by", "bz", "bxyz")]


class _Task(C.Structure):
    _fields_ = [("problem", C.c_int), ("stencil", C.c_int), ("boundary_cond", C.c_int)]


class _Coo(C.Structure):
    _fields_ = [("block_indx", C.c_int), ("elements_num", C.c_int), ("rows_num", C.c_int),
                ("cols_num", C.c_int), ("block_rows_num", C.c_int), ("block_cols_num", C.c_int),
                ("rows", C.POINTER(C.c_int)), ("cols", C.POINTER(C.c_int)),
                ("values", C.POINTER(C.c_double))]


class _Vec(C.Structure):
    _fields_ = [("length", C.c_int), ("values", C.POINTER(C.c_double))]


def ref_available() -> bool:
    return os.path.exists(_REF)


def ref_generate(nx, ny, nz, dx=1, dy=1, dz=1, rank=0, stencil=7):
    """Calls the reference's own generator (src/base, compiled into oracle/_ref):
    cel_sparce_matrix_coo_interface (cel_sparce_matrix_coo_interface.c:22-60).
    Returns 0-based COO rows/cols/vals of block `rank`, b and x_theory."""
    L = C.CDLL(_REF)
    dom = _Domain(0, nx, ny, nz, dx, dy, dz, 0, 0, 0, 0, 0)
    task = _Task(1, {7: 1, 19: 2, 27: 3}[stencil], 1)
    coo = _Coo()
    bvec = _Vec()
    xvec = _Vec()
    r = C.c_int(rank)
    verbose = C.c_int(0)
    err = L.cel_sparce_matrix_coo_interface(C.byref(r), C.byref(task), C.byref(dom), C.byref(coo),
                                           C.byref(bvec), C.byref(xvec), C.byref(verbose))
    if err != 0:
        raise RuntimeError(f"reference generator returned {err}")
    k = coo.elements_num
    rows = np.ctypeslib.as_array(coo.rows, (k,)).copy()
    cols = np.ctypeslib.as_array(coo.cols, (k,)).copy()
    vals = np.ctypeslib.as_array(coo.values, (k,)).copy()
    b = np.ctypeslib.as_array(bvec.values, (bvec.length,)).copy()
    xt = np.ctypeslib.as_array(xvec.values, (xvec.length,)).copy()
    return dict(rows=rows, cols=cols, vals=vals, b=b, x_theory=xt, rows_num=coo.rows_num,
                cols_num=coo.cols_num, block_rows_num=coo.block_rows_num, bxyz=dom.bxyz)


# ---------------------------------------------------------------- reference reader and sort
_mm = None


def ref_mm_available() -> bool:
    return os.path.exists(_REF_MM)


def _ref_mm():
    global _mm
    if _mm is None:
        L = C.CDLL(_REF_MM)
        L.ref_mm_info.argtypes = [C.c_char_p, C.c_int] + [C.POINTER(C.c_int)] * 3
        L.ref_mm_read.argtypes = [C.c_char_p, C.c_int, C.c_int, _i32p, _i32p, _f64p]
        L.ref_coo_sort.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p]
        _mm = L
    return _mm


def ref_mm_read(path):
    """The reference's mminfo + mmread on a file (oracle/_ref, built from the
    reference sources).  The reference STOPs the process on a malformed
    file: call it only on files it accepts, or from a child process."""
    b = os.fsencode(path)
    r, c, z = C.c_int(), C.c_int(), C.c_int()
    _ref_mm().ref_mm_info(b, len(b), C.byref(r), C.byref(c), C.byref(z))
    nnz = z.value
    rows = np.zeros(max(nnz, 1), np.int32)
    cols = np.zeros(max(nnz, 1), np.int32)
    vals = np.zeros(max(nnz, 1), np.float64)
    got = _ref_mm().ref_mm_read(b, len(b), max(nnz, 1), rows, cols, vals)
    if got < 0:
        raise RuntimeError("reference reader refused")
    return r.value, c.value, rows[:nnz], cols[:nnz], vals[:nnz]


def ref_coo_sort(nrows, rows, cols, vals):
    """The reference's cel_sp_mat_format_sparse_31_sort on copies.  Call it
    only where coo_sort_cel returns 0: elsewhere the reference reads past
    its arrays."""
    r, c, v = _i32(rows).copy(), _i32(cols).copy(), _f64(vals).copy()
    err = _ref_mm().ref_coo_sort(int(nrows), r.size, r, c, v)
    return err, r, c, v
