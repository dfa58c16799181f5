"""cel_v3_amd -- MI355X (gfx950) Jacobi-PCG hot path of the CEL solver.

The product is the C-ABI shared library cel_v3_amd/libcel_gpu.so
(include/cel_gpu.h).  This module is a thin ctypes front end over it, used by
the tests and bench.py; it has no compute path of its own and raises if the
library is missing or a call fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcel_gpu.so")

OK, NOT_CONVERGED = 0, 1
ERR = {-1: "ARG", -2: "HIP", -3: "NO_DEVICE", -4: "BAD_MATRIX", -5: "BAD_DIAG", -6: "BREAKDOWN",
       -7: "NONFINITE", -8: "OOM", -9: "REF_UNDEFINED", -10: "STATE", -11: "FILE"}
ERR_ARG, ERR_BAD_MATRIX, ERR_REF_UNDEFINED, ERR_FILE = -1, -4, -9, -11
FMT_CSR, FMT_JAD, FMT_COO, FMT_DIA = 1, 2, 3, 4
MEM_HOST, MEM_DEVICE = 0, 1
X0_ONE, X0_ZERO, X0_USER = 0, 1, 2
TOL_ABS, TOL_REL_B, TOL_REL_R0 = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
JAD_SPARSKIT, JAD_CEL_COMPAT = 0, 1
MM_COORDINATE, MM_ARRAY = 0, 1
MM_REAL, MM_COMPLEX, MM_PATTERN = 0, 1, 2
MM_GENERAL, MM_SYMMETRIC, MM_HERMITIAN, MM_SKEW_SYMMETRIC = 0, 1, 2, 3
MM_AS_STORED, MM_EXPAND_SYMMETRIC = 0, 1
SORT_ROW_COL, SORT_CEL_COMPAT = 0, 1
ELM_GMRES, ELM_PCG = 0, 1

# every entry point declared in include/cel_gpu.h
EXPORTS = [
    "cel_gpu_init", "cel_gpu_finalize", "cel_gpu_last_error", "cel_gpu_device_info",
    "cel_gpu_malloc", "cel_gpu_free", "cel_gpu_memcpy_h2d", "cel_gpu_memcpy_d2h",
    "cel_gpu_memcpy_d2d", "cel_gpu_synchronize", "cel_gpu_mat_create_csr",
    "cel_gpu_mat_create_coo", "cel_gpu_mat_gen_poisson7", "cel_gpu_mat_gen_diffusion7",
    "cel_gpu_mat_destroy", "cel_gpu_mat_info", "cel_gpu_mat_get_csr", "cel_gpu_mat_to_jad",
    "cel_gpu_mat_get_jad", "cel_gpu_mat_set_jacobi", "cel_gpu_mat_scale_diagonal",
    "cel_gpu_vec_scale_diagonal", "cel_gpu_mat_get_jacobi", "cel_gpu_mat_get_diagonal",
    "cel_gpu_spmv", "cel_gpu_spmv_f32", "cel_gpu_spmv_time", "cel_gpu_mat_distr_split",
    "cel_gpu_mat_get_block", "cel_gpu_mat_set_send_list", "cel_gpu_face_pack",
    "cel_gpu_spmv_distr", "cel_gpu_mat_ghost_len", "cel_gpu_pcg_params_default",
    "cel_gpu_pcg_solve", "cel_gpu_cg_adjoint", "cel_gpu_diffusion7_kappa_grad",
    "cel_gpu_mat_add_halo_block", "cel_gpu_stream_copy_time", "cel_gpu_stream_read_time",
    "cel_gpu_spmv_dot",
    "cel_gpu_batch_create",
    "cel_gpu_batch_destroy", "cel_gpu_batch_tune", "cel_gpu_batch_info", "cel_gpu_batch_solve",
    "cel_gpu_batch_get_config", "cel_gpu_batch_set_config",
    "cel_gpu_mm_info", "cel_gpu_mm_read", "cel_gpu_coo_sort", "cel_gpu_coo_row_offsets",
    "cel_gpu_mat_create_coo_sorted", "cel_gpu_gmres_params_default", "cel_gpu_gmres_solve",
    "cel_gpu_elm_create", "cel_gpu_elm_row_start", "cel_gpu_elm_insert", "cel_gpu_elm_reinsert",
    "cel_gpu_elm_row_completed", "cel_gpu_elm_matrix_completed", "cel_gpu_elm_set_solver",
    "cel_gpu_elm_solve", "cel_gpu_elm_resolve", "cel_gpu_elm_collect_solution", "cel_gpu_elm_stats",
    "cel_gpu_elm_matrix", "cel_gpu_elm_destroy", "cel_gpu_mat_to_dia", "cel_gpu_mat_dia_info",
    "cel_gpu_pcg_solve_noalloc",
]


class CelError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERR.get(code, code)}: {msg}")
        self.code = code


class PcgParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("tol_mode", C.c_int), ("iter_max", C.c_int),
                ("x0_mode", C.c_int), ("precision", C.c_int), ("format", C.c_int),
                ("memspace", C.c_int), ("use_halo", C.c_int), ("ru_delta", C.c_double),
                ("ru_restart", C.c_int), ("check_every", C.c_int), ("use_graph", C.c_int),
                ("profile", C.c_int), ("alpha_stop", C.c_int)]


class GmresParams(C.Structure):
    _fields_ = [("eps", C.c_double), ("iter_max", C.c_int), ("restart", C.c_int),
                ("x0_mode", C.c_int), ("memspace", C.c_int), ("format", C.c_int)]


class PcgStats(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_int), ("outer_iters", C.c_int),
                ("norm0", C.c_double), ("final_norm", C.c_double), ("eps_abs", C.c_double),
                ("alpha", C.c_double), ("solve_ms", C.c_double), ("setup_ms", C.c_double),
                ("spmv_ms", C.c_double), ("spmv_launches", C.c_int), ("spmv_bytes", C.c_double),
                ("iter_bytes", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


def lib():
    """Loads libcel_gpu.so.  Raises if it has not been built: there is no
    other implementation to fall back to."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing; run `python -m cel_v3_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.cel_gpu_last_error.restype = C.c_char_p
        L.cel_gpu_pcg_params_default.restype = None
        L.cel_gpu_pcg_params_default.argtypes = [C.POINTER(PcgParams)]
        L.cel_gpu_gmres_params_default.restype = None
        L.cel_gpu_gmres_params_default.argtypes = [C.POINTER(GmresParams)]
        L.cel_gpu_elm_insert.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.cel_gpu_elm_reinsert.argtypes = [C.c_void_p, C.c_int, C.c_double]
        for name in EXPORTS:
            getattr(L, name)  # every declared symbol must be exported
        _lib = L
    return _lib


def _check(rc, allow=(OK,)):
    if rc in allow:
        return rc
    msg = lib().cel_gpu_last_error().decode(errors="replace")
    raise CelError(rc, msg)


def init(device=0):
    _check(lib().cel_gpu_init(device))


def device_info():
    buf = C.create_string_buffer(256)
    cus = C.c_int()
    _check(lib().cel_gpu_device_info(buf, 256, C.byref(cus)))
    return buf.value.decode(), cus.value


def _ptr(a):
    return a.ctypes.data_as(_vp)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- device arrays
class DeviceArray:
    """A device allocation owned by the library's allocator."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.ptr = _vp()
        _check(lib().cel_gpu_malloc(C.byref(self.ptr), C.c_size_t(max(1, self.n * self.dtype.itemsize))))

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.size, a.dtype)
        _check(lib().cel_gpu_memcpy_h2d(d.ptr, _ptr(a), C.c_size_t(a.nbytes)))
        return d

    def to_host(self):
        out = np.empty(self.n, self.dtype)
        _check(lib().cel_gpu_memcpy_d2h(_ptr(out), self.ptr, C.c_size_t(out.nbytes)))
        return out

    def copy_from(self, other: "DeviceArray"):
        _check(lib().cel_gpu_memcpy_d2d(self.ptr, other.ptr, C.c_size_t(self.n * self.dtype.itemsize)))

    def free(self):
        if self.ptr:
            lib().cel_gpu_free(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- matrices
class Matrix:
    """Device-resident local block (+ halo blocks after a distributed split)."""

    def __init__(self, handle):
        self.h = handle
        n, nc, nnz, mm, nb = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().cel_gpu_mat_info(self.h, C.byref(n), C.byref(nc), C.byref(nnz), C.byref(mm),
                                      C.byref(nb)))
        self.n, self.ncols, self.nnz, self.nblocks = n.value, nc.value, nnz.value, nb.value

    # constructors -------------------------------------------------------
    @classmethod
    def from_csr(cls, xadj, cols, vals, ncols=None, index_base=1, row_offset=0, col_offset=0):
        xadj, cols, vals = _i32(xadj), _i32(cols), _f64(vals)
        n = xadj.size - 1
        h = _vp()
        _check(lib().cel_gpu_mat_create_csr(C.byref(h), n, n if ncols is None else ncols, cols.size,
                                            _ptr(xadj), _ptr(cols), _ptr(vals), index_base,
                                            row_offset, col_offset, MEM_HOST))
        return cls(h)

    @classmethod
    def from_coo(cls, nrows, rows, cols, vals, ncols=None, index_base=1, row_offset=0, col_offset=0):
        rows, cols, vals = _i32(rows), _i32(cols), _f64(vals)
        h = _vp()
        _check(lib().cel_gpu_mat_create_coo(C.byref(h), nrows, nrows if ncols is None else ncols,
                                            rows.size, _ptr(rows), _ptr(cols), _ptr(vals),
                                            index_base, row_offset, col_offset))
        return cls(h)

    @classmethod
    def from_coo_sorted(cls, nrows, rows, cols, vals, ncols=None, mode=SORT_CEL_COMPAT,
                        index_base=1, row_offset=0, col_offset=0):
        """Sort (as coo_sort), count rows and build the matrix in HBM.  rows,
        cols, vals are host arrays or DeviceArrays (all of one kind)."""
        h = _vp()
        nc = nrows if ncols is None else ncols
        if isinstance(rows, DeviceArray):
            _check(lib().cel_gpu_mat_create_coo_sorted(C.byref(h), nrows, nc, rows.n, rows.ptr,
                                                       cols.ptr, vals.ptr, index_base, row_offset,
                                                       col_offset, mode, MEM_DEVICE))
        else:
            rows, cols, vals = _i32(rows), _i32(cols), _f64(vals)
            _check(lib().cel_gpu_mat_create_coo_sorted(C.byref(h), nrows, nc, rows.size, _ptr(rows),
                                                       _ptr(cols), _ptr(vals), index_base, row_offset,
                                                       col_offset, mode, MEM_HOST))
        return cls(h)

    @classmethod
    def from_mm(cls, path, mode=SORT_CEL_COMPAT, expand_symmetric=False):
        """Matrix Market file -> matrix: read on the host (as the reference),
        then one upload, sort, row count and build in HBM."""
        m = mm_read(path, expand_symmetric=expand_symmetric)
        if m["field"] != MM_REAL:
            raise CelError(-1, "from_mm needs a real-valued file")
        d = [DeviceArray.from_host(m["rows"]), DeviceArray.from_host(m["cols"]),
             DeviceArray.from_host(m["vals"])]
        try:
            return cls.from_coo_sorted(m["nrows"], *d, ncols=m["ncols"], mode=mode)
        finally:
            for a in d:
                a.free()

    @classmethod
    def poisson7(cls, nx, ny, nz):
        h = _vp()
        _check(lib().cel_gpu_mat_gen_poisson7(C.byref(h), nx, ny, nz))
        return cls(h)

    @classmethod
    def diffusion7(cls, nx, ny, nz, kappa):
        h = _vp()
        if isinstance(kappa, DeviceArray):
            _check(lib().cel_gpu_mat_gen_diffusion7(C.byref(h), nx, ny, nz, kappa.ptr, MEM_DEVICE))
        else:
            k = _f64(kappa)
            _check(lib().cel_gpu_mat_gen_diffusion7(C.byref(h), nx, ny, nz, _ptr(k), MEM_HOST))
        return cls(h)

    @classmethod
    def distr_split(cls, vtxdist, my_rank, rows, cols, vals, index_base=1):
        vtxdist, rows, cols, vals = _i32(vtxdist), _i32(rows), _i32(cols), _f64(vals)
        h = _vp()
        _check(lib().cel_gpu_mat_distr_split(C.byref(h), vtxdist.size - 1, _ptr(vtxdist), my_rank,
                                             rows.size, _ptr(rows), _ptr(cols), _ptr(vals),
                                             index_base))
        return cls(h)

    def destroy(self):
        if self.h:
            lib().cel_gpu_mat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # getters ------------------------------------------------------------
    def get_csr(self):
        xadj = np.empty(self.n + 1, np.int32)
        cols = np.empty(max(self.nnz, 1), np.int32)
        vals = np.empty(max(self.nnz, 1), np.float64)
        _check(lib().cel_gpu_mat_get_csr(self.h, _ptr(xadj), _ptr(cols), _ptr(vals)))
        return xadj, cols[: self.nnz], vals[: self.nnz]

    def to_jad(self, mode=JAD_SPARSKIT, allow_undefined=False):
        rc = lib().cel_gpu_mat_to_jad(self.h, mode)
        _check(rc, allow=(OK, -9) if allow_undefined else (OK,))
        return rc

    def to_dia(self):
        """Diagonal storage for the CSR/PCG kernels (bit-identical products)."""
        _check(lib().cel_gpu_mat_to_dia(self.h))
        return self.dia_info()

    def dia_info(self):
        nd = C.c_int()
        offs = (C.c_int * 32)()
        _check(lib().cel_gpu_mat_dia_info(self.h, C.byref(nd), offs))
        return list(offs[: nd.value])

    def get_jad(self):
        mm, nd, nw = C.c_int(), C.c_int(), C.c_longlong()
        _check(lib().cel_gpu_mat_get_jad(self.h, None, None, None, None, C.byref(mm), C.byref(nd),
                                         C.byref(nw)))
        perm = np.empty(max(self.n, 1), np.int32)
        begin = np.empty(mm.value + 1, np.int32)
        cols = np.empty(max(nw.value, 1), np.int32)
        vals = np.empty(max(nw.value, 1), np.float64)
        _check(lib().cel_gpu_mat_get_jad(self.h, _ptr(perm), _ptr(begin), _ptr(cols), _ptr(vals),
                                         C.byref(mm), C.byref(nd), C.byref(nw)))
        return dict(perm=perm[: self.n], begin=begin, cols=cols[: nw.value], vals=vals[: nw.value],
                    mmax=mm.value, ndefined=nd.value)

    def set_jacobi(self, row_scaled=False):
        _check(lib().cel_gpu_mat_set_jacobi(self.h, int(bool(row_scaled))))

    def scale_diagonal(self):
        _check(lib().cel_gpu_mat_scale_diagonal(self.h))

    def scale_vec(self, b):
        b = _f64(b).copy()
        _check(lib().cel_gpu_vec_scale_diagonal(self.h, _ptr(b), MEM_HOST))
        return b

    def jacobi(self):
        out = np.empty(self.n)
        _check(lib().cel_gpu_mat_get_jacobi(self.h, _ptr(out)))
        return out

    def diagonal(self):
        out = np.empty(self.n)
        _check(lib().cel_gpu_mat_get_diagonal(self.h, _ptr(out)))
        return out

    # operators ----------------------------------------------------------
    def spmv(self, x, y=None, fmt=FMT_CSR, accumulate=False):
        if isinstance(x, DeviceArray):
            _check(lib().cel_gpu_spmv(self.h, fmt, x.ptr, y.ptr, int(accumulate), MEM_DEVICE))
            return y
        x = _f64(x)
        y = np.zeros(self.n) if y is None else _f64(y).copy()
        _check(lib().cel_gpu_spmv(self.h, fmt, _ptr(x), _ptr(y), int(accumulate), MEM_HOST))
        return y

    def spmv_f32(self, x, y=None, fmt=FMT_CSR, accumulate=False):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(self.n, np.float32) if y is None else np.ascontiguousarray(y, np.float32).copy()
        _check(lib().cel_gpu_spmv_f32(self.h, fmt, _ptr(x), _ptr(y), int(accumulate), MEM_HOST))
        return y

    def spmv_time(self, x: DeviceArray, y: DeviceArray, fmt=FMT_CSR, precision=PREC_FP64, reps=20):
        ms, nb = C.c_double(), C.c_double()
        _check(lib().cel_gpu_spmv_time(self.h, fmt, precision, x.ptr, y.ptr, reps, C.byref(ms),
                                       C.byref(nb)))
        return ms.value, nb.value

    # distributed --------------------------------------------------------
    def get_block(self, b):
        proc, nnz, rlen = C.c_int(), C.c_int(), C.c_int()
        _check(lib().cel_gpu_mat_get_block(self.h, b, C.byref(proc), C.byref(nnz), None, None, None,
                                           None, C.byref(rlen), None))
        k = nnz.value
        rows = np.empty(max(k, 1), np.int32)
        cols = np.empty(max(k, 1), np.int32)
        vals = np.empty(max(k, 1), np.float64)
        xadj = np.empty(self.n + 1, np.int32)
        perm = np.empty(max(rlen.value, 1), np.int32)
        _check(lib().cel_gpu_mat_get_block(self.h, b, C.byref(proc), C.byref(nnz), _ptr(rows),
                                           _ptr(cols), _ptr(vals), _ptr(xadj), C.byref(rlen),
                                           _ptr(perm)))
        return dict(proc=proc.value, rows=rows[:k], cols=cols[:k], vals=vals[:k], xadj=xadj,
                    recv_perm=perm[: rlen.value])

    def add_halo_block(self, proc, proc_first_row, proc_len, rows, cols, vals, index_base=1):
        rows, cols, vals = _i32(rows), _i32(cols), _f64(vals)
        _check(lib().cel_gpu_mat_add_halo_block(self.h, proc, proc_first_row, proc_len, rows.size,
                                                _ptr(rows), _ptr(cols), _ptr(vals), index_base))
        self.nblocks += 1

    def set_send_list(self, peer, perm):
        perm = _i32(perm)
        _check(lib().cel_gpu_mat_set_send_list(self.h, peer, perm.size, _ptr(perm)))

    def ghost_len(self):
        g, s = C.c_int(), C.c_int()
        _check(lib().cel_gpu_mat_ghost_len(self.h, C.byref(g), C.byref(s)))
        return g.value, s.value

    def face_pack(self, x):
        _, st = self.ghost_len()
        buf = np.empty(max(st, 1))
        _check(lib().cel_gpu_face_pack(self.h, _ptr(_f64(x)), _ptr(buf), MEM_HOST))
        return buf[:st]

    def spmv_distr(self, x, ghost, fmt=FMT_CSR):
        y = np.zeros(self.n)
        g = _f64(ghost) if len(ghost) else np.zeros(1)
        _check(lib().cel_gpu_spmv_distr(self.h, fmt, _ptr(_f64(x)), _ptr(g), _ptr(y), MEM_HOST))
        return y

    def spmv_dot(self, x, ghost=None, fmt=FMT_CSR):
        """The PCG's fused product (host arrays): returns (y, x.y); with ghost,
        the halo blocks are applied to those received values."""
        y = np.zeros(self.n)
        d = C.c_double()
        g = None if ghost is None else _ptr(_f64(ghost) if len(ghost) else np.zeros(1))
        _check(lib().cel_gpu_spmv_dot(self.h, fmt, _ptr(_f64(x)), g, _ptr(y), C.byref(d), MEM_HOST))
        return y, d.value

    # solvers ------------------------------------------------------------
    def pcg(self, b, x0=None, tol=1e-8, tol_mode=TOL_ABS, iter_max=1000, x0_mode=X0_ONE,
            precision=PREC_FP64, fmt=FMT_CSR, history=False, allow_not_converged=True, **kw):
        """Jacobi-PCG (cel_cgalg_jacobi).  b/x0 host arrays or DeviceArrays."""
        p = params(tol=tol, tol_mode=tol_mode, iter_max=iter_max, x0_mode=x0_mode,
                   precision=precision, format=fmt, **kw)
        st = PcgStats()
        hist = np.zeros(iter_max + 1) if history else None
        hp = _ptr(hist) if history else None
        if isinstance(b, DeviceArray):
            p.memspace = MEM_DEVICE
            x = x0 if x0 is not None else DeviceArray(self.n)
            if x0 is not None:
                p.x0_mode = X0_USER
            rc = lib().cel_gpu_pcg_solve(self.h, b.ptr, x.ptr, C.byref(p), C.byref(st), hp)
        else:
            p.memspace = MEM_HOST
            x = np.zeros(self.n) if x0 is None else _f64(x0).copy()
            if x0 is not None:
                p.x0_mode = X0_USER
            rc = lib().cel_gpu_pcg_solve(self.h, _ptr(_f64(b)), _ptr(x), C.byref(p), C.byref(st), hp)
        _check(rc, allow=(OK, NOT_CONVERGED) if allow_not_converged else (OK,))
        out = st.as_dict()
        out["x"] = x
        if history:
            out["hist"] = hist[: st.iters + 1]
        return out

    def pcg_noalloc(self, b: "DeviceArray", x: "DeviceArray", tol=1e-8, tol_mode=TOL_ABS,
                    iter_max=1000, x0_mode=X0_ONE, fmt=FMT_CSR, precision=PREC_FP64,
                    allow_not_converged=True, **kw):
        """cel_gpu_pcg_solve_noalloc: the FP64 solve without allocations, b and
        x device arrays (x holds x0 with X0_USER and receives the solution)."""
        p = params(tol=tol, tol_mode=tol_mode, iter_max=iter_max, x0_mode=x0_mode,
                   precision=precision, format=fmt, **kw)
        p.memspace = MEM_DEVICE
        st = PcgStats()
        rc = lib().cel_gpu_pcg_solve_noalloc(self.h, b.ptr, x.ptr, C.byref(p), C.byref(st))
        _check(rc, allow=(OK, NOT_CONVERGED) if allow_not_converged else (OK,))
        return st.as_dict()

    def gmres(self, b, x0=None, eps=1e-8, iter_max=5000, restart=3, x0_mode=X0_ONE, fmt=FMT_CSR,
              allow_not_converged=True):
        """Jacobi-GMRES (cel_gmresalg_2jacobi).  b/x0 host arrays or DeviceArrays."""
        p = GmresParams()
        lib().cel_gpu_gmres_params_default(C.byref(p))
        p.eps, p.iter_max, p.restart, p.x0_mode, p.format = eps, iter_max, restart, x0_mode, fmt
        st = PcgStats()
        if isinstance(b, DeviceArray):
            p.memspace = MEM_DEVICE
            x = x0 if x0 is not None else DeviceArray(self.n)
            if x0 is not None:
                p.x0_mode = X0_USER
            rc = lib().cel_gpu_gmres_solve(self.h, b.ptr, x.ptr, C.byref(p), C.byref(st))
        else:
            p.memspace = MEM_HOST
            x = np.zeros(self.n) if x0 is None else _f64(x0).copy()
            if x0 is not None:
                p.x0_mode = X0_USER
            rc = lib().cel_gpu_gmres_solve(self.h, _ptr(_f64(b)), _ptr(x), C.byref(p), C.byref(st))
        _check(rc, allow=(OK, NOT_CONVERGED) if allow_not_converged else (OK,))
        out = st.as_dict()
        out["x"] = x
        out["cycles"] = out["outer_iters"]
        return out

    def adjoint(self, x, dLdx, want_dLda=True, tol=1e-12, tol_mode=TOL_REL_B, iter_max=10000,
                alpha_stop=0, **kw):
        p = params(tol=tol, tol_mode=tol_mode, iter_max=iter_max, x0_mode=X0_ZERO,
                   alpha_stop=alpha_stop, **kw)
        p.memspace = MEM_HOST
        st = PcgStats()
        lam = np.zeros(self.n)
        g = np.zeros(max(self.nnz, 1)) if want_dLda else None
        rc = lib().cel_gpu_cg_adjoint(self.h, _ptr(_f64(x)), _ptr(_f64(dLdx)), _ptr(lam),
                                      _ptr(g) if want_dLda else None, C.byref(p), C.byref(st))
        _check(rc, allow=(OK, NOT_CONVERGED))
        return lam, (g[: self.nnz] if want_dLda else None), st.as_dict()


def params(**kw) -> PcgParams:
    p = PcgParams()
    lib().cel_gpu_pcg_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown PCG parameter {k}")
        setattr(p, k, v)
    return p


# ---------------------------------------------------------------- batches
class Batch:
    """Independent systems (and blocks of right-hand sides sharing one
    matrix) solved together: cel_gpu_batch_*.  b and x of a system with k
    right-hand sides are arrays of shape (k, n) (k = 1: shape (n,))."""

    def __init__(self, mats, nrhs=None):
        self.mats = list(mats)  # keeps the matrices alive
        self.nrhs = [1] * len(self.mats) if nrhs is None else [int(k) for k in nrhs]
        hs = (_vp * len(self.mats))(*[m.h for m in self.mats])
        ks = np.asarray(self.nrhs, dtype=np.int32)
        self.h = _vp()
        _check(lib().cel_gpu_batch_create(C.byref(self.h), len(self.mats), hs, _ptr(ks)))
        self.lanes = sum(self.nrhs)

    def _vecs(self, arrays):
        out = []
        for m, k, a in zip(self.mats, self.nrhs, arrays):
            a = _f64(a)
            if a.size != m.n * k:
                raise ValueError(f"expected {k} x {m.n} values, got {a.shape}")
            out.append(a)
        return out

    def tune(self, bs, iters=24):
        bs = self._vecs(bs)
        ptrs = (_vp * len(bs))(*[_ptr(b) for b in bs])
        _check(lib().cel_gpu_batch_tune(self.h, ptrs, MEM_HOST, int(iters)))
        return self.info()

    def info(self):
        lanes, grid, chunk, weight, reb = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        el, ms = C.c_longlong(), C.c_double()
        _check(lib().cel_gpu_batch_info(self.h, C.byref(lanes), C.byref(el), C.byref(grid),
                                        C.byref(chunk), C.byref(weight), C.byref(reb),
                                        C.byref(ms)))
        g23 = C.c_int()
        _check(lib().cel_gpu_batch_get_config(self.h, None, C.byref(g23), None, None))
        return {"lanes": lanes.value, "elements": el.value, "grid": grid.value,
                "grid23": g23.value, "chunk": chunk.value, "weight": weight.value,
                "rebalance": reb.value, "ms_per_iter": ms.value}

    def set_config(self, grid, grid23=0, chunk=32, weight=0):
        """Fixes the launch configuration of later solves (cel_gpu_batch_set_config)."""
        _check(lib().cel_gpu_batch_set_config(self.h, int(grid), int(grid23), int(chunk),
                                              int(weight)))

    def solve(self, bs, x0s=None, tol=1e-9, tol_mode=TOL_REL_B, iter_max=10000, x0_mode=X0_ZERO,
              allow_not_converged=True, **kw):
        """Returns {"x": [arrays shaped like b], "lanes": [stats dicts], "total": stats dict}."""
        p = params(tol=tol, tol_mode=tol_mode, iter_max=iter_max, x0_mode=x0_mode, **kw)
        p.memspace = MEM_HOST
        bs = self._vecs(bs)
        if x0s is not None:
            xs = [a.copy() for a in self._vecs(x0s)]
            p.x0_mode = X0_USER
        else:
            xs = [np.zeros(b.size) for b in bs]
        bp = (_vp * len(bs))(*[_ptr(b) for b in bs])
        xp = (_vp * len(xs))(*[_ptr(x) for x in xs])
        lane_st = (PcgStats * self.lanes)()
        tot = PcgStats()
        rc = lib().cel_gpu_batch_solve(self.h, bp, xp, C.byref(p), lane_st, C.byref(tot))
        # a lane that broke down or had a non-finite right-hand side is
        # reported per lane (the other lanes' results are valid); any other
        # error is raised
        lane_errors = (-6, -7)
        if not (allow_not_converged and rc in lane_errors):
            _check(rc, allow=(OK, NOT_CONVERGED) if allow_not_converged else (OK,))
        shaped = [x.reshape(k, m.n) if k > 1 else x for x, k, m in zip(xs, self.nrhs, self.mats)]
        return {"x": shaped, "lanes": [st.as_dict() for st in lane_st], "total": tot.as_dict(),
                "rc": rc}

    def close(self):
        if self.h:
            lib().cel_gpu_batch_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- Matrix Market and COO order
def mm_info(path):
    """Header and size line of a Matrix Market file (host only)."""
    v = [C.c_int() for _ in range(6)]
    _check(lib().cel_gpu_mm_info(os.fsencode(path), *[C.byref(x) for x in v]))
    return dict(zip(("rep", "field", "symm", "nrows", "ncols", "nnz"), (x.value for x in v)))


def mm_read(path, expand_symmetric=False):
    """Coordinate entries in file order, 1-based as written (host only).
    Pattern files give vals of zeros (the library leaves them untouched)."""
    info = mm_info(path)
    cap = max(info["nnz"] * (2 if expand_symmetric else 1), 1)
    rows, cols = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    vals = np.zeros(cap, np.float64)
    got = C.c_int()
    _check(lib().cel_gpu_mm_read(os.fsencode(path), MM_EXPAND_SYMMETRIC if expand_symmetric else
                                 MM_AS_STORED, cap, _ptr(rows), _ptr(cols), _ptr(vals),
                                 C.byref(got)))
    k = got.value
    return dict(info, rows=rows[:k].copy(), cols=cols[:k].copy(), vals=vals[:k].copy())


def coo_sort(nrows, rows, cols, vals, ncols=None, mode=SORT_CEL_COMPAT, index_base=1,
             row_offset=0, col_offset=0, allow_undefined=False):
    """Sorts in HBM.  Host arrays: returns (status, rows, cols, vals) as new
    arrays.  DeviceArrays: sorted in place, returns (status, rows, cols, vals)
    with the same DeviceArrays.  status is 0, or ERR_REF_UNDEFINED with
    allow_undefined (arrays then sorted by row only)."""
    nc = nrows if ncols is None else ncols
    allow = (OK, ERR_REF_UNDEFINED) if allow_undefined else (OK,)
    if isinstance(rows, DeviceArray):
        rc = _check(lib().cel_gpu_coo_sort(nrows, nc, rows.n, rows.ptr, cols.ptr, vals.ptr,
                                           index_base, row_offset, col_offset, mode, MEM_DEVICE),
                    allow)
        return rc, rows, cols, vals
    r, c, v = _i32(rows).copy(), _i32(cols).copy(), _f64(vals).copy()
    rc = _check(lib().cel_gpu_coo_sort(nrows, nc, r.size, _ptr(r), _ptr(c), _ptr(v), index_base,
                                       row_offset, col_offset, mode, MEM_HOST), allow)
    return rc, r, c, v


def coo_row_offsets(nrows, rows, index_base=1, row_offset=0):
    """cel_sp_mat_to_csr's xadj (nrows+1, starting at index_base)."""
    r = _i32(rows)
    xadj = np.zeros(nrows + 1, np.int32)
    _check(lib().cel_gpu_coo_row_offsets(nrows, r.size, _ptr(r), index_base, row_offset, _ptr(xadj),
                                         MEM_HOST))
    return xadj


# ---------------------------------------------------------------- element solver API
class Elm:
    """cel_elmSolver* as a handle (cel_gpu_elm_*): insert rows, complete the
    matrix, solve, collect the solution."""

    def __init__(self, n, nnz_per_row=(7, 0)):
        self.n = n
        self.h = _vp()
        nz = (C.c_int * 2)(*nnz_per_row)
        _check(lib().cel_gpu_elm_create(C.byref(self.h), n, n, 1, nz))

    def row_start(self, row):
        _check(lib().cel_gpu_elm_row_start(self.h, row))

    def insert(self, col, value):
        _check(lib().cel_gpu_elm_insert(self.h, col, value))

    def reinsert(self, col, value):
        _check(lib().cel_gpu_elm_reinsert(self.h, col, value))

    def row_completed(self):
        _check(lib().cel_gpu_elm_row_completed(self.h))

    def insert_rows(self, rows, cols, vals):
        """Entries in order; a row is started whenever it changes."""
        cur = None
        for r, c, v in zip(rows, cols, vals):
            if r != cur:
                if cur is not None:
                    self.row_completed()
                self.row_start(int(r))
                cur = r
            self.insert(int(c), float(v))
        if cur is not None:
            self.row_completed()

    def matrix_completed(self):
        _check(lib().cel_gpu_elm_matrix_completed(self.h))

    def set_solver(self, solver):
        _check(lib().cel_gpu_elm_set_solver(self.h, solver))

    def _run(self, fn, b):
        ierr, nit, ftol = C.c_int(), C.c_int(), C.c_double()
        rc = fn(self.h, _ptr(_f64(b)), C.byref(ierr), C.byref(nit), C.byref(ftol))
        _check(rc, allow=(OK, NOT_CONVERGED))
        return dict(rc=rc, iierr=ierr.value, niter=nit.value, ftol=ftol.value)

    def solve(self, b):
        return self._run(lib().cel_gpu_elm_solve, b)

    def resolve(self, b):
        return self._run(lib().cel_gpu_elm_resolve, b)

    def solution(self):
        x = np.empty(self.n)
        _check(lib().cel_gpu_elm_collect_solution(self.h, _ptr(x)))
        return x

    def stats(self):
        st = PcgStats()
        _check(lib().cel_gpu_elm_stats(self.h, C.byref(st)))
        return st.as_dict()

    def matrix_csr(self):
        """1-based (xadj, cols, vals) of the assembled (scaled) matrix."""
        a = _vp()
        _check(lib().cel_gpu_elm_matrix(self.h, C.byref(a)))
        m = Matrix.__new__(Matrix)
        m.h = a
        n, nc, nnz, mm, nb = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().cel_gpu_mat_info(a, C.byref(n), C.byref(nc), C.byref(nnz), C.byref(mm),
                                      C.byref(nb)))
        m.n, m.ncols, m.nnz, m.nblocks = n.value, nc.value, nnz.value, nb.value
        out = m.get_csr()
        m.h = None  # owned by the handle
        return out

    def destroy(self):
        if self.h:
            lib().cel_gpu_elm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def stream_copy_time(nbytes, reps=20):
    ms = C.c_double()
    _check(lib().cel_gpu_stream_copy_time(C.c_size_t(nbytes), reps, C.byref(ms)))
    return ms.value


def stream_read_time(nbytes, streams=1, blocks_per_cu=4, reps=10):
    ms = C.c_double()
    _check(lib().cel_gpu_stream_read_time(C.c_size_t(nbytes), streams, blocks_per_cu, reps,
                                          C.byref(ms)))
    return ms.value


def kappa_grad(nx, ny, nz, kappa, x, lam):
    g = np.empty(nx * ny * nz)
    _check(lib().cel_gpu_diffusion7_kappa_grad(nx, ny, nz, _ptr(_f64(kappa)), _ptr(_f64(x)),
                                               _ptr(_f64(lam)), _ptr(g), MEM_HOST))
    return g
