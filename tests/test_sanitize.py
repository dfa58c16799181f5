"""Sanitizer runs of the host code (CPU only).

The Matrix Market reader (cel_v3_amd/csrc/mmio.cpp) and the C oracle
(oracle/cel_oracle.c) are built with AddressSanitizer and
UndefinedBehaviorSanitizer, aborting on the first report
(tests/sanitize/Makefile, gcc/g++):

* the reader reads every fixture of test_mmio.py and a seeded set of
  malformed files made from them (truncations, substituted bytes from the
  characters its grammar looks at, deleted / repeated / swapped lines,
  extreme size lines and indices, over-long lines and headers, odd numbers,
  binary noise), each with output arrays of exactly the needed length, one
  entry short and room for the symmetric expansion;
* the oracle's own tests (test_oracle.py) run in a child process against
  the sanitized oracle library with the ASan runtime preloaded.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from test_mmio import ACCEPTED, REFUSED, REFUSED_HERE_ONLY, UNDEFINED

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
BUILD = os.path.join(SAN, "_build")
CODES = {0, -1, -9, -11}  # OK, ARG (capacity), REF_UNDEFINED, FILE


def _asan_runtime():
    try:
        out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                             check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return out if os.path.isabs(out) and os.path.exists(out) else None


@pytest.fixture(scope="module")
def san_build():
    if shutil.which("g++") is None or shutil.which("make") is None or _asan_runtime() is None:
        pytest.skip("no gcc sanitizer runtime")
    subprocess.run(["make", "-s", "-C", SAN], check=True)
    return BUILD


SIZE_LINES = ["2147483647 2147483647 2147483647", "-1 -1 -1", "99999999999999999999 1 1",
              "0 0 0", "3 3 -5", "1 1 4194304", "2 2", "2 2 2 2", "1e3 1 1", "", "   ", "2,2,1",
              "4194304 4194304 4194304"]
NUMBERS = ["1e99999", "-1e99999", "nan", "NaN", "inf", "-Infinity", "1e-400", "-0", "0x10", "1.2.3",
           "++1", "1e", "e5", ".", "-", "1d", "1q+", "2147483648", "-2147483649", "99999999999999999999",
           "1.5", "3*", "*", ",", ",,", "/", "1/2", "\t", "1e+", "1E-", "0.0000000001e400"]
NOISE = " \t\n\r,.-+eEdDqQ%*/0123456789\x00\xff"


def corpus(seed=2024, per_base=60):
    """Malformed Matrix Market files derived from the test fixtures (bytes)."""
    rng = np.random.default_rng(seed)
    bases = [t[0] for t in ACCEPTED.values()] + [t[0] for t in REFUSED.values()] + \
            [t for t in REFUSED_HERE_ONLY.values() if t is not None] + list(UNDEFINED.values())
    bases = [b.encode() for b in bases]
    out = list(bases)
    for base in bases:
        lines = base.split(b"\n")
        for _ in range(per_base):
            kind = int(rng.integers(9))
            if kind == 0:  # truncation
                f = base[: int(rng.integers(len(base) + 1))]
            elif kind == 1:  # substituted bytes
                b = bytearray(base)
                for _ in range(int(rng.integers(1, 6))):
                    b[int(rng.integers(len(b)))] = ord(NOISE[int(rng.integers(len(NOISE)))])
                f = bytes(b)
            elif kind == 2:  # a line deleted, repeated or swapped
                ls = list(lines)
                i, j = int(rng.integers(len(ls))), int(rng.integers(len(ls)))
                op = int(rng.integers(3))
                if op == 0:
                    del ls[i]
                elif op == 1:
                    ls.insert(i, ls[j])
                else:
                    ls[i], ls[j] = ls[j], ls[i]
                f = b"\n".join(ls)
            elif kind == 3:  # an extreme size line
                ls = list(lines)
                k = next((i for i, l in enumerate(ls) if l and not l.startswith(b"%")), 1)
                if k < len(ls):
                    ls[k] = SIZE_LINES[int(rng.integers(len(SIZE_LINES)))].encode()
                f = b"\n".join(ls)
            elif kind == 4:  # odd numbers in place of an entry's words
                ls = list(lines)
                i = int(rng.integers(len(ls)))
                words = ls[i].split()
                if words:
                    words[int(rng.integers(len(words)))] = NUMBERS[int(rng.integers(len(NUMBERS)))].encode()
                    ls[i] = b" ".join(words)
                f = b"\n".join(ls)
            elif kind == 5:  # an over-long line (entry or header)
                ls = list(lines)
                i = int(rng.integers(len(ls)))
                fill = [b" ", b"1", b"1 ", b"%", b"x"][int(rng.integers(5))]
                ls[i] = ls[i] + fill * int(rng.integers(1000, 5000))
                f = b"\n".join(ls)
            elif kind == 6:  # the header cut or with its words repeated
                ls = list(lines)
                h = ls[0]
                ls[0] = h[: int(rng.integers(len(h) + 1))] if rng.integers(2) else h + b" " + h[15:]
                f = b"\n".join(ls)
            elif kind == 7:  # binary noise after the header
                f = lines[0] + b"\n" + rng.integers(0, 256, int(rng.integers(1, 200)), dtype=np.uint8).tobytes()
            else:  # no final newline, or CR-only line ends
                f = base.rstrip(b"\n") if rng.integers(2) else base.replace(b"\n", b"\r")
            out.append(f)
    return out


def test_reader_fuzz_under_sanitizers(san_build, tmp_path):
    files = corpus()
    paths = []
    for i, data in enumerate(files):
        p = tmp_path / f"f{i:04d}.mtx"
        p.write_bytes(data)
        paths.append(str(p))
    paths.append(str(tmp_path / "missing.mtx"))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(san_build, "mm_driver")] + paths, capture_output=True, text=True,
                       env=env, timeout=600)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    lines = r.stdout.strip().split("\n")
    assert len(lines) == len(paths)
    seen = set()
    for line in lines:
        codes = [int(w) for w in line.split()[2:] if w.lstrip("-").isdigit()]
        assert set(codes) <= CODES, line
        seen.add(codes[0])
    # the corpus reaches both outcomes of the header read
    assert {0, -11} <= seen


def test_oracle_tests_under_sanitizers(san_build):
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(), ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1",
               CEL_ORACLE_LIB=os.path.join(san_build, "liboracle_san.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(HERE, "test_oracle.py"), "-q",
                        "-p", "no:cacheprovider"], capture_output=True, text=True, env=env,
                       cwd=os.path.dirname(HERE), timeout=900)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout[-3000:]
    assert " passed" in r.stdout
