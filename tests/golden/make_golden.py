"""Writes tests/golden/ref_generator.npz from the reference's own C
generator (src/base/cel_discretisation_poisson_3d.c, compiled into
oracle/_ref by `make -C oracle ref`).

The fixtures are data only: COO rows/cols/values of each block, the
right-hand side b and the exact solution x_theory = x^2+y^2+z^2 that the
generator returns (0-based indices, as the C code emits them).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

CASES = [
    # (name, nx, ny, nz, dx, dy, dz, stencil)
    ("s7_5x4x3", 5, 4, 3, 1, 1, 1, 7),
    ("s7_6x6x6", 6, 6, 6, 1, 1, 1, 7),
    ("s27_5x5x5", 5, 5, 5, 1, 1, 1, 27),
    ("s19_4x5x6", 4, 5, 6, 1, 1, 1, 19),
    ("s7_8x8x4_d2x2x1", 8, 8, 4, 2, 2, 1, 7),
    ("s27_6x6x6_d2x1x3", 6, 6, 6, 2, 1, 3, 27),
]


def main():
    if not oracle.ref_available():
        oracle.build()
    out = {}
    for name, nx, ny, nz, dx, dy, dz, st in CASES:
        nblocks = dx * dy * dz
        for r in range(nblocks):
            g = oracle.ref_generate(nx, ny, nz, dx, dy, dz, rank=r, stencil=st)
            key = f"{name}_r{r}"
            out[key + "_rows"] = g["rows"].astype(np.int32)
            out[key + "_cols"] = g["cols"].astype(np.int32)
            out[key + "_vals"] = g["vals"]
            out[key + "_b"] = g["b"]
            out[key + "_xt"] = g["x_theory"]
            out[key + "_meta"] = np.array([nx, ny, nz, dx, dy, dz, st, g["rows_num"], g["cols_num"],
                                           g["block_rows_num"]], np.int64)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_generator.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
