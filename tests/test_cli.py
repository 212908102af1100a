"""The xtax command-line drop-in: input errors and exit codes (CPU; they are
reported before any device is opened) and end-to-end runs on the GPU whose
output files are checked with the oracle's exact verifier."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
XTAX = os.path.join(ROOT, "xtax_amd", "xtax")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(args, cwd, timeout=120):
    return subprocess.run([XTAX] + args, cwd=cwd, capture_output=True, text=True,
                          timeout=timeout)


def write_csv(path, M, sep=","):
    with open(path, "w") as f:
        for row in M:
            f.write(sep.join(str(int(v)) for v in row) + "\n")


def read_csv(path):
    return np.array([[int(x) for x in r.split(",")] for r in open(path).read().split("\n")
                     if r.strip()], np.int64)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(XTAX):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "xtax_amd"), "xtax"], check=True)


def test_exactly_one_input(tmp_path):
    r = run([], tmp_path)
    assert r.returncode == 1 and "provide exactly one of -A or -L" in r.stderr
    write_csv(tmp_path / "a.csv", [[1, 0], [0, 1]])
    r = run(["-A", "a.csv", "-L", "a.csv"], tmp_path)
    assert r.returncode == 1


@pytest.mark.parametrize("args,msg", [
    (["--deflate", "--deflate-blocks"], "at most one of --deflate"),
    (["--gram-rows", "0,1", "--rcm"], "--gram-rows cannot be combined with --rcm"),
    (["--gram-rows", "0,1", "--deflate"], "--gram-rows cannot be combined with --deflate"),
    (["--lattice-dims", "0"], "--lattice-dims requires -L"),
])
def test_flag_conflicts(tmp_path, args, msg):
    write_csv(tmp_path / "a.csv", [[2, 1], [1, 2]])
    r = run(["-A", "a.csv"] + args, tmp_path)
    assert r.returncode == 1 and msg in r.stderr


@pytest.mark.parametrize("content,msg", [
    ("1,2\n3\n", "Bad CSV row length"),
    ("1,2,3\n4,5,6\n", "Matrix must be square"),
    ("\n\n", "Empty matrix file"),
])
def test_malformed_csv(tmp_path, content, msg):
    (tmp_path / "a.csv").write_text(content)
    r = run(["-A", "a.csv"], tmp_path)
    assert r.returncode == 1 and msg in r.stderr


def test_missing_file_and_x_mismatch(tmp_path):
    r = run(["-A", "nope.csv"], tmp_path)
    assert r.returncode == 1 and "Cannot open nope.csv" in r.stderr
    write_csv(tmp_path / "a.csv", [[2, 1], [1, 2]])
    write_csv(tmp_path / "x.csv", [[1, 0, 0], [0, 1, 0], [0, 0, 1]])
    r = run(["-A", "a.csv", "-X", "x.csv"], tmp_path)
    assert r.returncode == 1 and "X dimensions do not match A" in r.stderr


def test_bad_index_spec(tmp_path):
    write_csv(tmp_path / "a.csv", [[2, 1], [1, 2]])
    r = run(["-A", "a.csv", "--gram-rows", "0,5"], tmp_path)
    assert r.returncode == 1 and "out of range" in r.stderr
    r = run(["-A", "a.csv", "--gram-rows", "1..0"], tmp_path)
    assert r.returncode == 1 and "reversed range" in r.stderr
    r = run(["-A", "a.csv", "--gram-rows", "0"], tmp_path)
    assert r.returncode == 1 and "need at least 2 rows" in r.stderr


def test_unknown_flag(tmp_path):
    r = run(["--bogus"], tmp_path)
    assert r.returncode != 0 and "not expected" in r.stderr


# ------------------------------------------------------------------ GPU runs
def sample_matrix():
    return np.load(os.path.join(GOLD, "g2_end_to_end.npz"))["r0_A0"]


@pytest.mark.gpu
def test_matrix_mode_solves_sample(tmp_path, orc):
    A0 = sample_matrix()
    write_csv(tmp_path / "a.csv", A0, sep=", ")  # the sample's own spacing
    r = run(["-A", "a.csv", "-t", "16", "--seed", "3", "--t-init", "50", "--t-min", "0.5",
             "--max-seconds", "60"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "done moves=" in r.stdout and "Final best score:" in r.stdout
    assert "matrix is diagonal" in r.stdout
    # "new best" progress lines: strictly decreasing scores, below the start
    start = int(r.stdout.split("[t=0s] start score=")[1].split()[0])
    best = [int(l.split("new best score=")[1].split()[0]) for l in r.stdout.splitlines()
            if "] new best score=" in l]
    assert all(b < a for a, b in zip([start] + best, best))
    A = read_csv(tmp_path / "best_A.csv")
    X = read_csv(tmp_path / "best_X.csv")
    status, sign, _ = orc.verify(A0, X, A)
    assert status == 0 and sorted(np.diag(A).tolist()) == [-1] + [1] * 9


@pytest.mark.gpu
def test_resume_from_x(tmp_path, orc):
    from xtax_amd import workloads
    A0, D, U = workloads.planted(8, 9, 16, seed=21)
    # start from a partial transform: X0 = first column shears of U^{-1} are unknown,
    # so use a random unimodular X0 and check the invariant against the input
    rng = np.random.default_rng(2)
    X0 = np.eye(8, dtype=np.int64)
    for _ in range(6):
        i, j = rng.choice(8, 2, replace=False)
        X0[:, j] += X0[:, i]
    write_csv(tmp_path / "a.csv", A0)
    write_csv(tmp_path / "x.csv", X0)
    r = run(["-A", "a.csv", "-X", "x.csv", "-t", "16", "--seed", "5", "--t-init", "50",
             "--t-min", "0.5", "--max-seconds", "5"], tmp_path)
    assert r.returncode == 0, r.stderr
    A = read_csv(tmp_path / "best_A.csv")
    X = read_csv(tmp_path / "best_X.csv")
    status, _, _ = orc.verify(A0, X, A)
    assert status & ~2 == 0


@pytest.mark.gpu
def test_lattice_mode(tmp_path):
    from xtax_amd import workloads
    L, G = workloads.lattice(6, 6, seed=3)
    write_csv(tmp_path / "l.csv", L)
    r = run(["-L", "l.csv", "-t", "16", "--seed", "2", "--max-seconds", "2"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "[lattice] m=6 d=6" in r.stdout and "[warn]" not in r.stdout
    Lf = read_csv(tmp_path / "final_L.csv")
    A = read_csv(tmp_path / "best_A.csv")
    X = read_csv(tmp_path / "best_X.csv")
    assert (Lf @ Lf.T == A).all() and (X.T @ L == Lf).all()
    assert round(abs(np.linalg.det(X))) == 1


@pytest.mark.gpu
@pytest.mark.parametrize("spec,sel", [("0..4,6", [0, 1, 2, 3, 4, 6]), ("1,3..5", [1, 3, 4, 5])])
def test_lattice_dims_restricts_the_metric(tmp_path, spec, sel):
    """--lattice-dims (main.cpp:765-800): the Gram is built from the selected
    lattice columns only; the transform acts on the full lattice, and the
    annealed Gram equals the restricted Gram of final_L (main.cpp:849-857)."""
    from xtax_amd import workloads
    L = workloads.lattice(6, 8, seed=4)[0]
    write_csv(tmp_path / "l.csv", L)
    r = run(["-L", "l.csv", "--lattice-dims", spec, "-t", "16", "--seed", "3",
             "--max-seconds", "2"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert f"[lattice-dims] Gram built from {len(sel)} of 8 lattice dimensions" in r.stdout
    assert ("fewer selected dimensions" in r.stdout) == (len(sel) < 6)
    assert "[warn]" not in r.stdout
    Lf = read_csv(tmp_path / "final_L.csv")
    A = read_csv(tmp_path / "best_A.csv")
    X = read_csv(tmp_path / "best_X.csv")
    assert Lf.shape == L.shape and (X.T @ L == Lf).all()
    assert (Lf[:, sel] @ Lf[:, sel].T == A).all()
    assert round(abs(np.linalg.det(X))) == 1
    # the restricted metric is what was annealed: its score did not grow
    G0 = L[:, sel] @ L[:, sel].T
    score = lambda M: 2 * np.abs(M).sum() - np.abs(np.diag(M)).sum()
    assert score(A) <= score(G0)


@pytest.mark.gpu
def test_gram_rows_restricts_moves(tmp_path, orc):
    A0 = sample_matrix()
    write_csv(tmp_path / "a.csv", A0)
    r = run(["-A", "a.csv", "--gram-rows", "0..5", "-t", "8", "--seed", "1",
             "--max-seconds", "2"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "[gram-rows] restricting moves to 6 of 10 rows" in r.stdout
    X = read_csv(tmp_path / "best_X.csv")
    A = read_csv(tmp_path / "best_A.csv")
    # rows/columns outside the set are never pivot or target: X is the identity there
    assert (X[:, 6:] == np.eye(10, dtype=np.int64)[:, 6:]).all()
    assert (X[6:, :6] == 0).all()
    assert orc.verify(A0, X, A)[0] & ~2 == 0


@pytest.mark.gpu
def test_modulus(tmp_path, orc):
    A0 = np.array([[7, 12, 0], [12, -9, 5], [0, 5, 20]], np.int64)
    write_csv(tmp_path / "a.csv", A0)
    r = run(["-A", "a.csv", "--modulus", "5", "-t", "8", "--max-seconds", "2"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "[modulus] reduced entries mod 5" in r.stdout
    red = ((A0 % 5) + 5) % 5
    red = np.where(red > 2, red - 5, red)
    X = read_csv(tmp_path / "best_X.csv")
    A = read_csv(tmp_path / "best_A.csv")
    assert orc.verify(red, X, A)[0] & ~2 == 0


def test_deflate_needs_a_unimodular_matrix(tmp_path):
    write_csv(tmp_path / "a.csv", [[2, 1], [1, 2]])   # det 3
    r = run(["-A", "a.csv", "--deflate"], tmp_path)
    assert r.returncode == 1
    assert "[deflate] verifying the matrix is unimodular" in r.stdout
    assert "--deflate requires a unimodular matrix" in r.stderr


def _unimodular(n, shears, seed):
    rng = np.random.default_rng(seed)
    d = rng.choice([-1, 1], n)
    U = np.eye(n, dtype=np.int64)
    for _ in range(shears):
        i, j = rng.choice(n, 2, replace=False)
        U[:, j] += int(rng.choice([-1, 1])) * U[:, i]
    return np.asarray((U.T * d) @ U, dtype=np.int64), d


@pytest.mark.gpu
def test_strict_deflation_diagonalizes_a_unimodular_form(tmp_path, orc):
    A0, d = _unimodular(12, 30, seed=4)
    write_csv(tmp_path / "a.csv", A0)
    r = run(["-A", "a.csv", "--deflate", "-t", "16", "--seed", "3", "--deflate-slice", "0.3",
             "--max-seconds", "60"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "[deflate] start mode=strict" in r.stdout and "[deflate] done" in r.stdout
    A = read_csv(tmp_path / "best_A.csv")
    X = read_csv(tmp_path / "best_X.csv")
    status, _, _ = orc.verify(A0, X, A)
    assert status == 0   # X^T A0 X = A, A diagonal, |det X| = 1
    assert sorted(np.diag(A).tolist()) == sorted(d.tolist())   # Sylvester inertia


@pytest.mark.gpu
def test_relaxed_deflation_keeps_the_congruence(tmp_path, orc):
    from xtax_amd import workloads
    A0 = workloads.planted(10, 6, 20, seed=8)[0]
    write_csv(tmp_path / "a.csv", A0)
    r = run(["-A", "a.csv", "--deflate-blocks", "-t", "16", "--seed", "2", "--deflate-slice",
             "0.2", "--max-seconds", "3"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "[deflate] start mode=blocks" in r.stdout and "[deflate] done active=" in r.stdout
    A = read_csv(tmp_path / "best_A.csv")
    X = read_csv(tmp_path / "best_X.csv")
    assert orc.verify(A0, X, A)[0] & ~2 == 0


@pytest.mark.gpu
@pytest.mark.parametrize("flag,bits", [("--rcm", 1), ("--centroid", 2)])
def test_publish_reorder(tmp_path, flag, bits):
    """With a fixed epoch budget the run is deterministic: the reordered run's
    output is the plain run's best with the reference's permutation applied."""
    import ctypes as C
    from xtax_amd import _native as N
    from xtax_amd import workloads
    A0 = workloads.planted(14, 20, 30, seed=12)[0]
    write_csv(tmp_path / "a.csv", A0)
    args = ["-A", "a.csv", "-t", "8", "--seed", "5", "--max-epochs", "3",
            "--exchange-interval", "500"]
    plain = tmp_path / "plain"
    plain.mkdir()
    write_csv(plain / "a.csv", A0)
    assert run(args, plain).returncode == 0
    r = run(args + [flag], tmp_path)
    assert r.returncode == 0, r.stderr
    Ap, Xp = read_csv(plain / "best_A.csv"), read_csv(plain / "best_X.csv")
    A, X = read_csv(tmp_path / "best_A.csv"), read_csv(tmp_path / "best_X.csv")
    Ae, Xe = np.ascontiguousarray(Ap).copy(), np.ascontiguousarray(Xp).copy()
    perm = np.zeros(14, np.int32)
    assert N.lib().xtx_reorder(Ae.ctypes.data_as(N.I64P), Xe.ctypes.data_as(N.I64P), 14, bits,
                               perm.ctypes.data_as(N.I32P)) == 0
    assert np.array_equal(A, Ae) and np.array_equal(X, Xe)
    assert np.array_equal(A, Ap[np.ix_(perm, perm)]) and np.array_equal(X, Xp[:, perm])
    assert (X.T @ A0 @ X == A).all()


@pytest.mark.gpu
def test_lookahead_flag_changes_speed_only(tmp_path):
    """--lookahead: one lane group per replica, or the rule's 16 proposals at
    once, write the same best_A.csv and best_X.csv for a fixed epoch budget;
    an unsupported value is an input error."""
    from xtax_amd import workloads
    A0 = workloads.random_symmetric(12, -20, 20, seed=31)
    outs = []
    for look in ("1", "0", "4"):
        d = tmp_path / look
        d.mkdir()
        write_csv(d / "a.csv", A0)
        r = run(["-A", "a.csv", "-t", "6", "--seed", "7", "--max-epochs", "4",
                 "--exchange-interval", "400", "--lookahead", look], d)
        assert r.returncode == 0, r.stderr
        outs.append((read_csv(d / "best_A.csv"), read_csv(d / "best_X.csv")))
    for A, X in outs[1:]:
        assert np.array_equal(A, outs[0][0]) and np.array_equal(X, outs[0][1])
    write_csv(tmp_path / "a.csv", A0)
    r = run(["-A", "a.csv", "-t", "6", "--max-epochs", "1", "--lookahead", "3"], tmp_path)
    assert r.returncode == 1 and "lookahead" in r.stderr
