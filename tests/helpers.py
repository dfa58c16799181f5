"""Shared inputs for the parity tests (seeded, small enough for the oracle)."""
import os

import numpy as np

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_generator.npz")


def random_csr(n, m, per_row, rng, empty_every=0, long_rows=()):
    """1-based CSR with ~per_row sorted random columns per row; every
    `empty_every`-th row empty; rows in long_rows get m entries."""
    xadj = [1]
    cols, vals = [], []
    for i in range(n):
        if empty_every and i % empty_every == 1:
            k = 0
        elif i in long_rows:
            k = m
        else:
            k = int(rng.integers(0, 2 * per_row + 1))
            k = min(k, m)
        c = np.sort(rng.choice(m, size=k, replace=False)) + 1
        cols.extend(c)
        vals.extend(rng.normal(size=k))
        xadj.append(xadj[-1] + k)
    return O.CSR(np.array(xadj), np.array(cols, dtype=np.int64), np.array(vals))


def gold():
    return np.load(GOLD)


def gold_csr(name):
    """Reference-generator block 0 of a single-rank case as 1-based CSR."""
    d = gold()
    key = name + "_r0"
    rows, cols, vals = d[key + "_rows"], d[key + "_cols"], d[key + "_vals"]
    n = int(d[key + "_meta"][7])
    xadj = O.coo_to_csr(n, rows + 1, 0)
    return O.CSR(xadj, cols + 1, vals), d[key + "_b"], d[key + "_xt"]


def gold_ranks(name, nblocks):
    d = gold()
    parts = []
    for r in range(nblocks):
        key = f"{name}_r{r}"
        parts.append((d[key + "_rows"] + 1, d[key + "_cols"] + 1, d[key + "_vals"], d[key + "_b"]))
    sizes = [p[3].size for p in parts]
    vtx = np.concatenate([[1], 1 + np.cumsum(sizes)]).astype(np.int32)
    return vtx, parts


def spmv_matrices():
    """(name, 1-based CSR) cases covering ragged shapes, tiles that are not
    multiples of 256 rows, empty rows, rows past the LDS stage, explicit
    zeros and identity rows from the reference generator."""
    rng = np.random.default_rng(11)
    out = [
        ("p7_1", O.gen_poisson7(1, 1, 1)),
        ("p7_7x5x3", O.gen_poisson7(7, 5, 3)),
        ("p7_32", O.gen_poisson7(32, 32, 32)),
        ("p7_33x17x9", O.gen_poisson7(33, 17, 9)),
        ("rand_empty", random_csr(700, 700, 4, rng, empty_every=7)),
        # a 600-entry row pushes its tile past the LDS stage (fallback path)
        ("rand_long", random_csr(600, 600, 6, rng, long_rows=(3, 300))),
        ("rand_dense_tile", random_csr(300, 300, 12, rng)),
    ]
    if os.path.exists(GOLD):
        for name in ("s7_6x6x6", "s27_5x5x5", "s19_4x5x6"):
            out.append(("ref_" + name, gold_csr(name)[0]))
    return out


def face_coeffs(shape, kappa):
    """Interior face coefficients 2 k_i k_j / (k_i + k_j) of the diffusion7
    matrix along x, y and z (arrays indexed [z, y, x])."""
    nx, ny, nz = shape
    k3 = kappa.reshape(nz, ny, nx)
    out = []
    for ax, m in ((2, nx), (1, ny), (0, nz)):
        a = np.take(k3, range(0, m - 1), axis=ax)
        c = np.take(k3, range(1, m), axis=ax)
        out.append((2.0 * a * c) / (a + c))
    return out


def delta_apply(shape, kappa, kappa_p, f0, v):
    """(A(kappa_p) - A(kappa)) v from the changes of the face coefficients:
    row i gets sum_f da_f (v_i - v_j) over its interior faces and
    dkappa_i v_i per boundary face.  Nothing here is rounded to the size of
    a matrix entry, only to the size of the change."""
    nx, ny, nz = shape
    v3 = v.reshape(nz, ny, nx)
    out = np.zeros_like(v3)
    for ax, m, a0, a1 in zip((2, 1, 0), (nx, ny, nz), f0, face_coeffs(shape, kappa_p)):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, m - 1)
        hi[ax] = slice(1, m)
        t = (a1 - a0) * (v3[tuple(lo)] - v3[tuple(hi)])
        out[tuple(lo)] += t
        out[tuple(hi)] -= t
    dk = (kappa_p - kappa).reshape(nz, ny, nx)
    for ax, m in ((2, nx), (1, ny), (0, nz)):
        for end in (0, m - 1):
            sl = [slice(None)] * 3
            sl[ax] = end
            out[tuple(sl)] += dk[tuple(sl)] * v3[tuple(sl)]
    return out.ravel()


def stencil27(nx, ny, nz):
    """27 I - (all-ones 27-point stencil), Dirichlet outside, as a 0-based
    scipy CSR matrix: up to 27 entries per row, symmetric and diagonally
    dominant (SPD); 27 diagonals."""
    import scipy.sparse as sp

    def t(m):
        return sp.diags([np.ones(m - 1), np.ones(m), np.ones(m - 1)], [-1, 0, 1])

    K = sp.kron(t(nz), sp.kron(t(ny), t(nx)))
    A = (27.0 * sp.identity(nx * ny * nz) - K).tocsr()
    A.sort_indices()
    return A
