"""The xdual command-line drop-in (src/xdual_main.cpp:426-652 over the gfx950
engine): input errors and exit codes on the CPU (reported before any device is
opened), end-to-end runs on the GPU whose output files are checked exactly
(X^T A X == best_P, the lattice Gram of final_L == best_P) and numerically
(best_Q = best_P^-1 to rounding)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
XDUAL = os.path.join(ROOT, "xtax_amd", "xdual")


def run(args, cwd, timeout=120):
    return subprocess.run([XDUAL] + args, cwd=cwd, capture_output=True, text=True,
                          timeout=timeout)


def write_csv(path, M):
    with open(path, "w") as f:
        for row in M:
            f.write(",".join(str(int(v)) for v in row) + "\n")


def read_csv(path, dtype=np.int64):
    conv = int if dtype == np.int64 else float
    return np.array([[conv(x) for x in r.split(",")] for r in open(path).read().split("\n")
                     if r.strip()], dtype)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(XDUAL):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "xtax_amd"), "xdual"], check=True)


def gram(n, seed):
    rng = np.random.default_rng(seed)
    B = np.eye(n, dtype=np.int64)
    for _ in range(2 * n):
        a, b = rng.choice(n, 2, replace=False)
        B[a] += int(rng.integers(-1, 2)) * B[b]
    return (B @ np.diag(rng.integers(1, 3, n)) @ B.T).astype(np.int64)


def test_exactly_one_input(tmp_path):
    r = run([], tmp_path)
    assert r.returncode == 1 and "provide exactly one of -A or -L" in r.stderr


def test_unknown_flag_and_bad_values(tmp_path):
    write_csv(tmp_path / "a.csv", [[2, 1], [1, 2]])
    r = run(["-A", "a.csv", "--rcm"], tmp_path)          # an xtax flag xdual does not have
    assert r.returncode == 109 and "not expected: --rcm" in r.stderr
    r = run(["-A", "a.csv", "--lambda"], tmp_path)
    assert r.returncode == 109 and "requires an argument" in r.stderr
    r = run(["-A", "a.csv", "--gram-rows", "1"], tmp_path)
    assert r.returncode == 1 and "need at least 2 rows" in r.stderr


def test_lattice_needs_two_vectors(tmp_path):
    write_csv(tmp_path / "l.csv", [[1, 2, 3]])
    r = run(["-L", "l.csv"], tmp_path)
    assert r.returncode == 1 and "at least 2 basis vectors" in r.stderr


@pytest.mark.gpu
def test_matrix_mode(tmp_path):
    n = 10
    A = gram(n, 4)
    write_csv(tmp_path / "a.csv", A)
    r = run(["-A", "a.csv", "-t", "16", "--seed", "5", "--max-seconds", "20"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "[xdual] inverting the Gram" in r.stdout and "[xdual] lambda=1 c=" in r.stdout
    assert "Final best score:" in r.stdout and "P:\n" in r.stdout
    P = read_csv(tmp_path / "best_P.csv")
    X = read_csv(tmp_path / "best_X.csv")
    Q = read_csv(tmp_path / "best_Q.csv", np.float64)
    Xo = X.astype(object)
    assert np.array_equal((Xo.T @ A.astype(object) @ Xo).astype(np.int64), P)
    assert round(abs(np.linalg.det(X.astype(float)))) == 1
    assert np.abs(P.astype(float) @ Q - np.eye(n)).max() < 1e-6
    assert "matrix is diagonal" in r.stdout
    assert np.count_nonzero(P - np.diag(np.diag(P))) == 0


@pytest.mark.gpu
def test_resume_from_x_and_options(tmp_path):
    n = 8
    A = gram(n, 9)
    X0 = np.eye(n, dtype=np.int64)
    X0[0, 2] = 1
    write_csv(tmp_path / "a.csv", A)
    write_csv(tmp_path / "x.csv", X0)
    r = run(["-A", "a.csv", "-X", "x.csv", "-t", "4", "--max-epochs", "3", "--lambda", "0.5",
             "--dual-refresh", "500", "--dual-check", "100", "--dual-tol", "1e-9",
             "--lambda-ramp", "0.01", "--no-tempering", "--gram-rows", "0..5"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "[gram-rows] restricting moves to 6 of 8 rows" in r.stdout
    P = read_csv(tmp_path / "best_P.csv")
    X = read_csv(tmp_path / "best_X.csv")
    Xo = X.astype(object)
    assert np.array_equal((Xo.T @ A.astype(object) @ Xo).astype(np.int64), P)
    for k in (6, 7):   # columns outside --gram-rows keep X0's
        assert np.array_equal(X[:, k], X0[:, k])


@pytest.mark.gpu
def test_lattice_mode(tmp_path):
    rng = np.random.default_rng(3)
    L = rng.integers(-4, 5, (6, 8))
    write_csv(tmp_path / "l.csv", L)
    r = run(["-L", "l.csv", "-t", "8", "--max-epochs", "20"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "[lattice] m=6 d=8" in r.stdout and "Primal Gram P:" in r.stdout
    Lf = read_csv(tmp_path / "final_L.csv")
    P = read_csv(tmp_path / "best_P.csv")
    assert np.array_equal(Lf @ Lf.T, P)
    X = read_csv(tmp_path / "best_X.csv")
    assert np.array_equal(X.T @ L, Lf)
    assert "[warn]" not in r.stdout


@pytest.mark.gpu
def test_singular_gram_is_an_error(tmp_path):
    write_csv(tmp_path / "a.csv", [[1, 2, 0], [2, 4, 0], [0, 0, 1]])
    r = run(["-A", "a.csv"], tmp_path)
    assert r.returncode == 1 and "could not invert the Gram" in r.stderr
