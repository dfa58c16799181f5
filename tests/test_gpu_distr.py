"""Distributed block split, face packing and halo SpMV on one GPU.

Every rank of a dx*dy*dz decomposition of the reference generator's matrix
(tests/golden) is built on the same GPU; the exchange that MPI would do is a
host copy here.  Checks:
  * the split (block order, entries, xadj, receive lists) is bit-exact with
    the oracle's restatement of cel_sp_mat_distr_from_sp_mat_local,
  * packed send buffers equal the receiving rank's ghost values,
  * y = A_loc x + halo blocks is bit-exact with the oracle (local CSR sum
    first, then the COO halo blocks in block order) and reassembles the
    global product, through the plain distributed SpMV (CSR, JAD, COO) and
    through the PCG's fused product kernel with the same ghost values (CSR
    and DIA; the 7-point rows have 1 or 2 halo entries, which the kernel's
    per-row record holds, the 27-point rows 4 to 15, which take the compact
    lists),
  * a one-subdomain PCG with its six halo faces and zero ghost values is
    bit-identical to the plain local solve.
"""
import numpy as np
import pytest

import oracle as O
from helpers import gold_ranks

pytestmark = pytest.mark.gpu

CASES = [("s7_8x8x4_d2x2x1", 4), ("s27_6x6x6_d2x1x3", 6)]


@pytest.mark.parametrize("name,nblocks", CASES)
def test_split_pack_and_distributed_spmv(gpu, name, nblocks):
    vtx, parts = gold_ranks(name, nblocks)
    N = int(vtx[-1] - 1)
    x = np.random.default_rng(7).normal(size=N)
    xs = [x[vtx[r] - 1:vtx[r + 1] - 1] for r in range(nblocks)]
    mats, refs = [], []
    for r, (rows, cols, vals, _) in enumerate(parts):
        print(f"{name} rank {r}", flush=True)
        M = gpu.Matrix.distr_split(vtx, r, rows, cols, vals)
        ref = O.distr_split(vtx, r, rows, cols, vals)
        assert M.nblocks == len(ref["blocks"])
        for b, rb in enumerate(ref["blocks"]):
            gb = M.get_block(b)
            assert gb["proc"] == rb["proc"]
            np.testing.assert_array_equal(gb["rows"], rb["rows"])
            np.testing.assert_array_equal(gb["cols"], rb["cols"])
            np.testing.assert_array_equal(gb["vals"], rb["vals"])
            np.testing.assert_array_equal(gb["xadj"], rb["xadj"])
            if b >= 1:
                np.testing.assert_array_equal(gb["recv_perm"], rb["comm_perm"])
        np.testing.assert_array_equal(M.diagonal(), ref["diagonal"])
        mats.append(M)
        refs.append(ref)
    # send lists: the owner packs what each receiver listed (cel_comm)
    for r, ref in enumerate(refs):
        for rb in ref["blocks"][1:]:
            mats[rb["proc"]].set_send_list(r, rb["comm_perm"])
    packed = [mats[q].face_pack(xs[q]) for q in range(nblocks)]
    for r, ref in enumerate(refs):
        ghost = []
        for rb in ref["blocks"][1:]:
            q = rb["proc"]
            g = O.face_pack(rb["comm_perm"], xs[q])
            # slice of q's packed buffer destined for r (peers ascending)
            peers = sorted(r2 for r2, rf in enumerate(refs) for b2 in rf["blocks"][1:] if b2["proc"] == q)
            off = 0
            for p in peers:
                if p == r:
                    break
                blk = next(b2 for b2 in refs[p]["blocks"][1:] if b2["proc"] == q)
                off += blk["comm_perm"].size
            np.testing.assert_array_equal(packed[q][off:off + g.size], g)
            ghost.append(g)
        ghost = np.concatenate(ghost) if ghost else np.zeros(0)
        y = mats[r].spmv_distr(xs[r], ghost)
        # oracle: local CSR, then halo blocks (COO) through unpacked copies
        b0 = ref["blocks"][0]
        yo = O.spmv_csr(O.CSR(b0["xadj"], b0["cols"], b0["vals"], b0["col_offset"]), xs[r])
        for rb in ref["blocks"][1:]:
            q = rb["proc"]
            xn = np.zeros(vtx[q + 1] - vtx[q])
            xn[rb["comm_perm"] - 1] = xs[q][rb["comm_perm"] - 1]
            yo = O.spmv_coo(rb["rows"], rb["cols"], rb["vals"], ref["row_offset"], rb["col_offset"], xn, yo)
        np.testing.assert_array_equal(y, yo)
        # the solve's fused product (cel_gpu_spmv_dot, the PCG K1) with these
        # ghost values: the same y bit for bit, x.y in its own order
        dref = float(np.dot(xs[r], yo))
        dtol = 1e-13 * float(np.abs(xs[r] * yo).sum())
        yf, dot = mats[r].spmv_dot(xs[r], ghost)
        np.testing.assert_array_equal(yf, yo)
        assert abs(dot - dref) <= dtol
        if name.startswith("s7"):  # (the small 27-point blocks have no DIA form)
            mats[r].to_dia()
            yf, dot = mats[r].spmv_dot(xs[r], ghost, fmt=gpu.FMT_DIA)
            np.testing.assert_array_equal(yf, yo)
            assert abs(dot - dref) <= dtol
        # without the halo blocks: the local product
        yl, _ = mats[r].spmv_dot(xs[r])
        np.testing.assert_array_equal(
            yl, O.spmv_csr(O.CSR(b0["xadj"], b0["cols"], b0["vals"], b0["col_offset"]), xs[r]))
        mats[r].to_jad()
        np.testing.assert_array_equal(mats[r].spmv_distr(xs[r], ghost, fmt=gpu.FMT_JAD), yo)
        # the local block as COO (mv_algorithm = MV_COO): every block's
        # entries added to y = 0 one by one, local block first
        yc = O.spmv_coo(b0["rows"], b0["cols"], b0["vals"], ref["row_offset"], b0["col_offset"],
                        xs[r], np.zeros(xs[r].size))
        for rb in ref["blocks"][1:]:
            q = rb["proc"]
            xn = np.zeros(vtx[q + 1] - vtx[q])
            xn[rb["comm_perm"] - 1] = xs[q][rb["comm_perm"] - 1]
            yc = O.spmv_coo(rb["rows"], rb["cols"], rb["vals"], ref["row_offset"], rb["col_offset"], xn, yc)
        np.testing.assert_array_equal(mats[r].spmv_distr(xs[r], ghost, fmt=gpu.FMT_COO), yc)


def block_major_poisson(m, d, rank):
    """Rows of `rank` in a d*d*d decomposition of a (d*m)^3 7-point Poisson
    grid with block-major numbering (each rank's rows contiguous, as the
    reference's cel_domain_3D_get_index orders them).  1-based global ids."""
    bx, by, bz = rank % d, (rank // d) % d, rank // (d * d)
    per = m ** 3
    rows, cols, vals = [], [], []

    def gid(X, Y, Z):
        ob = (X // m) + d * ((Y // m) + d * (Z // m))
        lx, ly, lz = X % m, Y % m, Z % m
        return ob * per + lx + m * (ly + m * lz) + 1

    G = d * m
    for lz in range(m):
        for ly in range(m):
            for lx in range(m):
                X, Y, Z = bx * m + lx, by * m + ly, bz * m + lz
                me = gid(X, Y, Z)
                ent = []
                for dx, dy, dz in ((0, 0, -1), (0, -1, 0), (-1, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    XX, YY, ZZ = X + dx, Y + dy, Z + dz
                    if 0 <= XX < G and 0 <= YY < G and 0 <= ZZ < G:
                        ent.append((gid(XX, YY, ZZ), -1.0))
                ent.append((me, 6.0))
                ent.sort()
                for c, v in ent:
                    rows.append(me)
                    cols.append(c)
                    vals.append(v)
    vtx = np.arange(d ** 3 + 1, dtype=np.int32) * per + 1
    return vtx, np.array(rows), np.array(cols), np.array(vals)


def test_subdomain_pcg_with_six_faces_equals_local_solve(gpu):
    m, d, rank = 12, 3, 13  # centre block: six neighbours
    vtx, rows, cols, vals = block_major_poisson(m, d, rank)
    M = gpu.Matrix.distr_split(vtx, rank, rows, cols, vals)
    assert M.nblocks == 7
    for b in range(1, 7):
        blk = M.get_block(b)
        assert blk["rows"].size == m * m and blk["recv_perm"].size == m * m
        # by symmetry of the stencil the neighbour needs my rows of this face
        M.set_send_list(blk["proc"], np.unique(blk["rows"] - vtx[rank] + 1))
    ghost, send = M.ghost_len()
    assert ghost == 6 * m * m and send == 6 * m * m
    M.set_jacobi()
    P = gpu.Matrix.poisson7(m, m, m)
    P.set_jacobi()
    # the local block is the Dirichlet-outside Poisson matrix of the subdomain
    xa, ca, va = M.get_csr()
    pa, pc, pv = P.get_csr()
    np.testing.assert_array_equal(xa, pa)
    np.testing.assert_array_equal(ca - (vtx[rank] - 1), pc)
    np.testing.assert_array_equal(va, pv)
    b = P.spmv(np.random.default_rng(1).random(P.n))
    kw = dict(tol=1e-10, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, iter_max=2000, history=True)
    r_halo = M.pcg(b, use_halo=1, **kw)
    r_loc = P.pcg(b, **kw)
    assert r_halo["iters"] == r_loc["iters"]
    np.testing.assert_array_equal(r_halo["hist"], r_loc["hist"])
    np.testing.assert_array_equal(r_halo["x"], r_loc["x"])
