"""Matrix Market reader (cel_gpu_mm_info / cel_gpu_mm_read; host code, no GPU).

Each fixture below is a small file with its entries written out by hand; the
expected arrays are what the file says.  Where the reference's own reader is
built (oracle/_ref/libcel_mm_ref.so: mminfo + mmread from
src/fbase/cel_mmio_module.f90, called as
src/fspmt/cel_sp_mat_generate_module.f90:219-273 calls them) every accepted
file is also read by it and must give the same arrays, and every file the
library refuses with CEL_ERR_FILE must make the reference stop (checked in a
child process, since the reference ends the process).  Files whose reference
result is undefined (array files, complex files of more than one entry) are
not given to the reference.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import cel_v3_amd as cel
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# name: (file text, expected (nrows, ncols, rows, cols, vals) or None for pattern values)
ACCEPTED = {
    "general": (
        "%%MatrixMarket matrix coordinate real general\n"
        "% a comment\n"
        "%another\n"
        "3 4 5\n"
        "1 1 4.0\n"
        "3 2 -1.5\n"
        "2 4 1e-3\n"
        "1 3 2.5E+2\n"
        "3 3 .125\n",
        (3, 4, [1, 3, 2, 1, 3], [1, 2, 4, 3, 3], [4.0, -1.5, 1e-3, 250.0, 0.125])),
    "case_and_width": (  # header words lowered; words cut to the reference's widths
        "%%MatrixMarket MATRIX Coordinates REAL General\n"
        "2 2 2\n"
        "1 1 1\n"
        "2 2 2\n",
        (2, 2, [1, 2], [1, 2], [1.0, 2.0])),
    "fortran_reals": (  # D and Q exponents, a signed exponent without a letter
        "%%MatrixMarket matrix coordinate real general\n"
        "2 2 4\n"
        "1 1 1.5D+2\n"
        "1 2 -2.0d-1\n"
        "2 1 3.25q0\n"
        "2 2 7.0+2\n",
        (2, 2, [1, 1, 2, 2], [1, 2, 1, 2], [150.0, -0.2, 3.25, 700.0])),
    "list_directed_layout": (  # commas, tabs, a READ spanning lines, extra values, blank lines
        "%%MatrixMarket matrix coordinate real general\n"
        "3 3 4\n"
        "1, 1, 2.0\n"
        "2\t2\t3.0\n"
        "\n"
        "3 3\n"
        "   4.0   99 98\n"
        "1 3 -5.0 % trailing text is dropped\n",
        (3, 3, [1, 2, 3, 1], [1, 2, 3, 3], [2.0, 3.0, 4.0, -5.0])),
    "crlf": (
        "%%MatrixMarket matrix coordinate real general\r\n"
        "2 2 2\r\n"
        "1 2 0.5\r\n"
        "2 1 0.25\r\n",
        (2, 2, [1, 2], [2, 1], [0.5, 0.25])),
    "symmetric_as_stored": (
        "%%MatrixMarket matrix coordinate real symmetric\n"
        "3 3 4\n"
        "1 1 2.0\n"
        "2 1 -1.0\n"
        "2 2 2.0\n"
        "3 2 -1.0\n",
        (3, 3, [1, 2, 2, 3], [1, 1, 2, 2], [2.0, -1.0, 2.0, -1.0])),
    "pattern": (
        "%%MatrixMarket matrix coordinate pattern general\n"
        "2 3 3\n"
        "1 3\n"
        "2 1\n"
        "2 2\n",
        (2, 3, [1, 2, 2], [3, 1, 2], None)),
    "patterns_truncated": (  # 'patterns' is cut to the 7-character field 'pattern'
        "%%MatrixMarket matrix coordinate patterns general\n"
        "1 1 1\n"
        "1 1\n",
        (1, 1, [1], [1], None)),
    "indices_not_checked": (  # the reader returns indices as written
        "%%MatrixMarket matrix coordinate real general\n"
        "2 2 2\n"
        "5 1 1.0\n"
        "0 -3 2.0\n",
        (2, 2, [5, 0], [1, -3], [1.0, 2.0])),
    "empty": (
        "%%MatrixMarket matrix coordinate real general\n"
        "4 4 0\n",
        (4, 4, [], [], [])),
}

# refused by both (CEL_ERR_FILE; the reference prints a message and stops)
REFUSED = {
    "integer_field": ("%%MatrixMarket matrix coordinate integer general\n2 2 1\n1 1 3\n",
                      "field is not recognized"),
    "reals_field": ("%%MatrixMarket matrix coordinate reals general\n1 1 1\n1 1 1.0\n",
                    "field is not recognized"),
    "bad_symmetry": ("%%MatrixMarket matrix coordinate real upper\n1 1 1\n1 1 1.0\n",
                     "symmetry is not recognized"),
    "bad_rep": ("%%MatrixMarket matrix dense real general\n1 1\n1.0\n",
                "representation not recognized"),
    "tabbed_size_line": ("%%MatrixMarket matrix coordinate real general\n2\t2\t1\n1 1 1.0\n",
                         "Size info inconsistant"),
    "blank_before_size": ("%%MatrixMarket matrix coordinate real general\n%c\n\n2 2 1\n1 1 1.0\n",
                          "Size info inconsistant"),
    "premature_eof": ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n2 2 2.0\n",
                      "Premature end-of-file"),
    "no_size_line": ("%%MatrixMarket matrix coordinate real general\n% only comments\n",
                     "Premature end-of-file"),
}

# refused by this reader; the reference interprets or crashes on them
REFUSED_HERE_ONLY = {
    "null_value": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1,,2.0\n",
    "repeat_count": "%%MatrixMarket matrix coordinate real general\n2 2 1\n2*1 1.0\n",
    "slash": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 /\n",
    "integer_row_as_real": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1.0 1 1.0\n",
    "not_a_header": "MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1.0\n",
    "missing_file": None,
}

UNDEFINED = {
    "array": "%%MatrixMarket matrix array real general\n2 2\n1.0\n2.0\n3.0\n4.0\n",
    "complex": "%%MatrixMarket matrix coordinate complex general\n2 2 2\n1 1 1.0 0.0\n2 2 1.0 2.0\n",
}


def write(tmp_path, name, text):
    p = tmp_path / f"{name}.mtx"
    p.write_bytes(text.encode())
    return str(p)


@pytest.mark.parametrize("name", sorted(ACCEPTED))
def test_reader_fixtures(tmp_path, name):
    text, (nr, nc, rows, cols, vals) = ACCEPTED[name]
    path = write(tmp_path, name, text)
    m = cel.mm_read(path)
    assert (m["nrows"], m["ncols"], m["nnz"]) == (nr, nc, len(rows))
    np.testing.assert_array_equal(m["rows"], rows)
    np.testing.assert_array_equal(m["cols"], cols)
    if vals is not None:
        np.testing.assert_array_equal(m["vals"], np.array(vals, np.float64))
    if O.ref_mm_available():
        rr, rc, rrow, rcol, rval = O.ref_mm_read(path)
        assert (rr, rc) == (nr, nc)
        np.testing.assert_array_equal(rrow, m["rows"])
        np.testing.assert_array_equal(rcol, m["cols"])
        if vals is not None:
            np.testing.assert_array_equal(rval, m["vals"])  # bit for bit: both round correctly


def test_header_fields(tmp_path):
    text = ACCEPTED["symmetric_as_stored"][0]
    info = cel.mm_info(write(tmp_path, "s", text))
    assert info == dict(rep=cel.MM_COORDINATE, field=cel.MM_REAL, symm=cel.MM_SYMMETRIC, nrows=3,
                        ncols=3, nnz=4)
    info = cel.mm_info(write(tmp_path, "a", UNDEFINED["array"]))
    assert info["rep"] == cel.MM_ARRAY and info["nnz"] == 4
    sym = "%%MatrixMarket matrix array real symmetric\n3 3\n1\n2\n3\n4\n5\n6\n"
    assert cel.mm_info(write(tmp_path, "as", sym))["nnz"] == 6        # (9-3)/2 + 3
    skew = "%%MatrixMarket matrix array real skew-symmetric\n3 3\n1\n2\n3\n"
    assert cel.mm_info(write(tmp_path, "ak", skew))["nnz"] == 3       # (9-3)/2


def test_symmetric_expansion(tmp_path):
    path = write(tmp_path, "s", ACCEPTED["symmetric_as_stored"][0])
    m = cel.mm_read(path, expand_symmetric=True)
    np.testing.assert_array_equal(m["rows"], [1, 2, 2, 3, 1, 2])
    np.testing.assert_array_equal(m["cols"], [1, 1, 2, 2, 2, 3])
    np.testing.assert_array_equal(m["vals"], [2.0, -1.0, 2.0, -1.0, -1.0, -1.0])
    skew = ("%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 1.5\n3 1 -2.0\n")
    m = cel.mm_read(write(tmp_path, "k", skew), expand_symmetric=True)
    np.testing.assert_array_equal(m["rows"], [2, 3, 1, 1])
    np.testing.assert_array_equal(m["cols"], [1, 1, 2, 3])
    np.testing.assert_array_equal(m["vals"], [1.5, -2.0, -1.5, 2.0])
    # general files are unchanged by the flag
    g = cel.mm_read(write(tmp_path, "g", ACCEPTED["general"][0]), expand_symmetric=True)
    assert g["nnz"] == 5 and g["rows"].size == 5


def _reference_stops(path):
    """Runs the reference reader on `path` in a child process; returns its
    output and whether the call returned (it should STOP instead)."""
    code = ("import sys; sys.path.insert(0, %r); import oracle as O\n"
            "O.ref_mm_read(%r)\nprint('RETURNED', flush=True)\n" % (ROOT, path))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    return r.stdout + r.stderr, "RETURNED" in r.stdout


@pytest.mark.parametrize("name", sorted(REFUSED))
def test_refused_like_the_reference(tmp_path, name):
    text, message = REFUSED[name]
    path = write(tmp_path, name, text)
    with pytest.raises(cel.CelError) as e:
        cel.mm_read(path)
    assert e.value.code == cel.ERR_FILE
    if O.ref_mm_available():
        out, returned = _reference_stops(path)
        assert not returned
        assert message in out, out


@pytest.mark.parametrize("name", sorted(REFUSED_HERE_ONLY))
def test_refused_forms(tmp_path, name):
    text = REFUSED_HERE_ONLY[name]
    path = str(tmp_path / "missing.mtx") if text is None else write(tmp_path, name, text)
    with pytest.raises(cel.CelError) as e:
        cel.mm_read(path)
    assert e.value.code == cel.ERR_FILE


@pytest.mark.parametrize("name", sorted(UNDEFINED))
def test_reference_undefined_files(tmp_path, name):
    with pytest.raises(cel.CelError) as e:
        cel.mm_read(write(tmp_path, name, UNDEFINED[name]))
    assert e.value.code == cel.ERR_REF_UNDEFINED


def test_single_complex_entry_reads_indices(tmp_path):
    """One complex entry fits the reference's one-element buffer: indices are
    read, values left unset (zeros here)."""
    text = "%%MatrixMarket matrix coordinate complex general\n3 3 1\n2 3 1.0 -1.0\n"
    m = cel.mm_read(write(tmp_path, "c1", text))
    assert m["field"] == cel.MM_COMPLEX
    np.testing.assert_array_equal(m["rows"], [2])
    np.testing.assert_array_equal(m["cols"], [3])


def test_capacity_is_checked(tmp_path):
    path = write(tmp_path, "g", ACCEPTED["general"][0])
    import ctypes as C
    rows = np.zeros(4, np.int32)
    cols = np.zeros(4, np.int32)
    vals = np.zeros(4)
    got = C.c_int()
    rc = cel.lib().cel_gpu_mm_read(os.fsencode(path), 0, 4, rows.ctypes.data_as(C.c_void_p),
                                   cols.ctypes.data_as(C.c_void_p),
                                   vals.ctypes.data_as(C.c_void_p), C.byref(got))
    assert rc == -1


def test_large_file_round_trip(tmp_path):
    """A 7-point Poisson matrix written in shuffled order with 17-digit
    values reads back exactly."""
    A = O.gen_poisson7(9, 8, 7)
    rows = np.repeat(np.arange(1, A.n + 1), np.diff(A.xadj))
    rng = np.random.default_rng(5)
    order = rng.permutation(A.nnz)
    vals = A.vals * rng.uniform(0.5, 2.0, A.nnz)
    lines = [f"{r} {c} {float(v)!r}" for r, c, v in zip(rows[order], A.cols[order], vals[order])]
    text = ("%%MatrixMarket matrix coordinate real general\n"
            f"{A.n} {A.n} {A.nnz}\n" + "\n".join(lines) + "\n")
    path = write(tmp_path, "p", text)
    m = cel.mm_read(path)
    np.testing.assert_array_equal(m["rows"], rows[order])
    np.testing.assert_array_equal(m["cols"], A.cols[order])
    np.testing.assert_array_equal(m["vals"], vals[order])
