"""cel_gpu_pcg_solve_noalloc (FP64 Jacobi-PCG without allocations, its
work vectors in a fixed device array) against cel_gpu_pcg_solve, whose
kernels it shares: the same iterations, norms and solution bits, on two
calls in a row, the second with a smaller system (the array keeps the first
call's contents), in the CSR, COO and DIA forms; what the array cannot hold
is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW = dict(tol=1e-10, iter_max=5000)


def _system(gpu, shape, seed):
    rng = np.random.default_rng(seed)
    kappa = np.exp(rng.normal(0.0, 0.5, shape[0] * shape[1] * shape[2]))
    M = gpu.Matrix.diffusion7(*shape, kappa)
    M.set_jacobi()
    b = rng.uniform(-1.0, 1.0, M.n)
    return M, gpu.DeviceArray.from_host(b)


@pytest.mark.parametrize("form", ["csr", "coo", "dia"])
def test_noalloc_solve_is_the_allocating_solve(gpu, form):
    fmt = {"csr": gpu.FMT_CSR, "coo": gpu.FMT_COO, "dia": gpu.FMT_DIA}[form]
    # 96 x 96 x 100 = 921,600 rows (close to the 2^20 limit), then 30 x 28 x 26
    for shape, seed in (((96, 96, 100), 5), ((30, 28, 26), 6)):
        M, d_b = _system(gpu, shape, seed)
        if form == "dia":
            M.to_dia()
        kw = dict(KW, tol_mode=gpu.TOL_REL_B, x0_mode=gpu.X0_ZERO, fmt=fmt)
        ref = M.pcg(d_b, use_graph=0, **kw)
        d_x = gpu.DeviceArray(M.n)
        got = M.pcg_noalloc(d_b, d_x, **kw)
        print(f"{form} {shape}: {got['iters']} iterations (allocating solve {ref['iters']}), "
              f"status {got['status']}", flush=True)
        assert ref["status"] == got["status"] == 0
        assert got["iters"] == ref["iters"]
        assert got["final_norm"] == ref["final_norm"] and got["norm0"] == ref["norm0"]
        np.testing.assert_array_equal(d_x.to_host(), ref["x"].to_host())
        M.destroy()


def test_noalloc_user_x0_and_absolute_tolerance(gpu):
    """The reference's absolute stop and a caller's x0 (x holds x0 on entry)."""
    M, d_b = _system(gpu, (20, 22, 24), 7)
    x0 = np.random.default_rng(8).normal(size=M.n)
    ref = M.pcg(d_b, x0=gpu.DeviceArray.from_host(x0), tol=1e-9, use_graph=0)
    d_x = gpu.DeviceArray.from_host(x0)
    got = M.pcg_noalloc(d_b, d_x, tol=1e-9, x0_mode=gpu.X0_USER)
    assert got["iters"] == ref["iters"]
    np.testing.assert_array_equal(d_x.to_host(), ref["x"].to_host())


def test_noalloc_refuses_what_it_cannot_hold(gpu):
    M = gpu.Matrix.poisson7(128, 128, 65)  # 1,064,960 rows > 2^20
    M.set_jacobi()
    d_b = gpu.DeviceArray(M.n)
    d_x = gpu.DeviceArray(M.n)
    with pytest.raises(gpu.CelError) as e:
        M.pcg_noalloc(d_b, d_x)
    assert e.value.code == gpu.ERR_ARG
    S, d_s = _system(gpu, (8, 8, 8), 9)
    d_y = gpu.DeviceArray(S.n)
    for bad in (dict(precision=gpu.PREC_MIXED), dict(use_halo=1)):
        with pytest.raises(gpu.CelError):
            S.pcg_noalloc(d_s, d_y, **bad)
    S.to_jad()
    with pytest.raises(gpu.CelError):
        S.pcg_noalloc(d_s, d_y, fmt=gpu.FMT_JAD)
