"""The host-only sources under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU): `make -C xtax_amd sanitize` builds, with g++, the two command lines and
a driver (tests/host_check.cpp) over the engine's host code (verify.cpp,
publish.cpp) and the CPU oracle. The driver runs the product's verifier,
reordering, unimodularity test and pivot clearing beside the oracle on seeded
random inputs, and the oracle's primitives and both trajectory-model modes; the
command lines run their input-checking paths (every error they report before
a device is opened); a third driver (tests/lane_eval_check.cpp) checks the
lane kernel's per-lane arithmetic against the oracle. Any sanitizer report
fails the test."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "xtax_amd", "_san")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "xtax_amd"), "sanitize"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + r.stderr[-3000:])
    return SAN


def clean(r):
    text = r.stdout + r.stderr
    return "AddressSanitizer" not in text and "runtime error" not in text and \
        "LeakSanitizer" not in text


def test_host_code_and_oracle(built):
    r = subprocess.run([os.path.join(built, "host_check")], capture_output=True, text=True,
                       env=ENV, timeout=600)
    assert r.returncode == 0 and clean(r), r.stderr[-4000:]
    assert "host_check: ok" in r.stdout


def test_lane_arithmetic_against_oracle(built):
    """The lane kernel's one-lane best shear and evaluation (lane_eval.h, both
    the int32 path for entries below 2^16 and the int64 path) equal the
    oracle's best_shear_l1 and try_add on every ordered pair of seeded random
    matrices of every order to 48 and at 64, entries past the bound included."""
    r = subprocess.run([os.path.join(built, "lane_eval_check")], capture_output=True, text=True,
                       env=ENV, timeout=900)
    assert r.returncode == 0 and clean(r), r.stderr[-4000:]
    assert " 0 failures" in r.stdout


def _write(path, rows):
    with open(path, "w") as f:
        for row in rows:
            f.write(",".join(str(v) for v in row) + "\n")


XTAX_CASES = [
    ([], "provide exactly one of -A or -L"),
    (["-A", "a.csv", "-L", "a.csv"], "provide exactly one"),
    (["-A", "a.csv", "--deflate", "--deflate-blocks"], "at most one of --deflate"),
    (["-A", "a.csv", "--gram-rows", "0,1", "--rcm"], "--gram-rows cannot be combined"),
    (["-A", "a.csv", "--lattice-dims", "0"], "--lattice-dims requires -L"),
    (["-A", "bad.csv"], "Bad CSV row length"),
    (["-A", "rect.csv"], "Matrix must be square"),
    (["-A", "empty.csv"], "Empty matrix file"),
    (["-A", "nope.csv"], "Cannot open nope.csv"),
    (["-A", "a.csv", "-X", "x3.csv"], "X dimensions do not match A"),
    (["-A", "a.csv", "--gram-rows", "0,5"], "out of range"),
    (["-A", "a.csv", "--gram-rows", "1..0"], "reversed range"),
    (["-A", "a.csv", "--gram-rows", "0"], "need at least 2 rows"),
    (["--bogus"], "not expected"),
]


@pytest.mark.parametrize("args,msg", XTAX_CASES)
def test_xtax_input_errors(built, tmp_path, args, msg):
    _write(tmp_path / "a.csv", [[2, 1], [1, 2]])
    _write(tmp_path / "x3.csv", [[1, 0, 0], [0, 1, 0], [0, 0, 1]])
    (tmp_path / "bad.csv").write_text("1,2\n3\n")
    (tmp_path / "rect.csv").write_text("1,2,3\n4,5,6\n")
    (tmp_path / "empty.csv").write_text("\n\n")
    r = subprocess.run([os.path.join(built, "xtax")] + args, cwd=tmp_path, capture_output=True,
                       text=True, env=ENV, timeout=120)
    assert r.returncode != 0 and msg in r.stderr, (r.returncode, r.stderr[-2000:])
    assert clean(r), r.stderr[-4000:]


@pytest.mark.parametrize("args,msg", [
    ([], "provide exactly one of -A or -L"),
    (["-A", "nope.csv"], "Cannot open nope.csv"),
    (["-L", "one.csv"], "at least 2 basis vectors"),
    (["--bogus"], "not expected"),
])
def test_xdual_input_errors(built, tmp_path, args, msg):
    _write(tmp_path / "one.csv", [[1, 2, 3]])
    r = subprocess.run([os.path.join(built, "xdual")] + args, cwd=tmp_path, capture_output=True,
                       text=True, env=ENV, timeout=120)
    assert r.returncode != 0 and msg in r.stderr, (r.returncode, r.stderr[-2000:])
    assert clean(r), r.stderr[-4000:]
