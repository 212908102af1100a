"""Generate golden vectors from the reference implementation.

Run in the build container only (it needs the reference sources):

    make -C oracle ref            # builds oracle/_ref/{xtax_ref,xdual_ref,
                                  #   libref_harness.so,libref_dual.so}
    python tests/golden/gen_golden.py

G1 (per primitive): seeded matrices and (i, j, s) cases are pushed through the
reference's own try_add / add_transform_row / commit_add / compute_row_mass /
score_delta_l1 / best_shear_l1 / rounded_div / score / count_offdiag_nonzero /
congruence_of / gram_of / xt_times (src/main.cpp:46-242, src/mat_io.hpp) via
oracle/_ref/libref_harness.so, and inputs plus outputs are written to
g1_*.npz.

G2 (end to end): the reference binary xtax_ref is run on the reference's sample
input (10x10.csv, stored here as data) and on small planted forms, single
threaded with fixed seeds. Runs that stop on a diagonal are deterministic and
record moves and the final best_A/best_X; every run's (A0, X, A) triple is also
kept as a verifier fixture.

G3 (publish-time reordering and deflation): rcm_order, centroid_order and
reorder_for_publish (src/main.cpp:244-350, :407-411), det_mod_prime and
is_unimodular (src/mat_io.hpp:597-644), deflate_index (src/main.cpp:435-455),
through the same harness, into g3_publish_deflate.npz. `gen_golden.py g3`
regenerates only that file.

G4 (xdual): eval_add / eval_add_d, DualObjective::best_shear, invert_to and
the off-diagonal norms (src/xdual_main.cpp:67-364, src/mat_io.hpp:297-337)
through oracle/_ref/libref_dual.so into g4_xdual.npz (`gen_golden.py g4`).

G5 (the worker loop): 64 runs of the reference's own canneal::anneal_worker
(src/congruence_anneal.hpp:177-416) on the calling thread, recorded by the
harness's objective wrapper (ref_worker_trace: every accepted move, sweep
commit and reseed, the final state and the global best, capped at a move
count), over random configurations of both temperature modes; and
single-threaded runs of the reference binary with non-default flags that stop
at a diagonal (deterministic), into g5_worker.npz (`gen_golden.py g5`).
G5b: 21 more worker traces with tournaments of 17 to 257 samples, into
g5b_samples.npz (`gen_golden.py g5b`).

G6 (xdual end to end): the reference's xdual binary (oracle/_ref/xdual_ref),
single threaded with fixed seeds and non-default flags, on small planted
forms; runs that stop at a diagonal primal are deterministic and record the
"done moves=" count, best_P, best_Q (written with 17 significant digits, so
exactly) and best_X, into g6_xdual_runs.npz (`gen_golden.py g6`).
G6b: 48 runs of the reference's anneal_worker on xdual's objective
(oracle/_ref/libref_dual.so refd_worker_trace, the G5 recording), over random
configurations of both temperature modes, dual weights and maintenance
settings, into g6b_dual_worker.npz (`gen_golden.py g6b`).
"""
import ctypes as C
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
I64P = C.POINTER(C.c_int64)
LIM = 1 << 48


def load_harness():
    L = C.CDLL(os.path.join(REF_DIR, "libref_harness.so"))
    L.ref_rounded_div.restype = C.c_int64
    L.ref_rounded_div.argtypes = [C.c_int64, C.c_int64]
    L.ref_score.restype = C.c_int64
    L.ref_score.argtypes = [I64P, C.c_int]
    L.ref_offdiag_nonzero.restype = C.c_int64
    L.ref_offdiag_nonzero.argtypes = [I64P, C.c_int]
    L.ref_try_add.restype = C.c_int
    L.ref_try_add.argtypes = [I64P, I64P, C.c_int, C.c_int, C.c_int, C.c_int64, I64P, I64P,
                              I64P, I64P]
    L.ref_add_transform_row.restype = C.c_int
    L.ref_add_transform_row.argtypes = [I64P, C.c_int, C.c_int, C.c_int, C.c_int64, I64P]
    L.ref_commit_add.restype = None
    L.ref_commit_add.argtypes = [I64P, I64P, C.c_int, C.c_int, I64P, I64P, I64P]
    L.ref_compute_row_mass.restype = None
    L.ref_compute_row_mass.argtypes = [I64P, C.c_int, I64P]
    L.ref_score_delta_l1.restype = C.c_int64
    L.ref_score_delta_l1.argtypes = [I64P, C.c_int, C.c_int, C.c_int, C.c_int64]
    L.ref_best_shear_l1.restype = C.c_int64
    L.ref_best_shear_l1.argtypes = [I64P, C.c_int, C.c_int, C.c_int]
    L.ref_congruence_of.restype = None
    L.ref_congruence_of.argtypes = [I64P, I64P, C.c_int, I64P]
    L.ref_gram_of.restype = None
    L.ref_gram_of.argtypes = [I64P, C.c_int, C.c_int, I64P]
    L.ref_xt_times.restype = None
    L.ref_xt_times.argtypes = [I64P, I64P, C.c_int, C.c_int, I64P]
    I32P = C.POINTER(C.c_int)
    L.ref_rcm_order.restype = None
    L.ref_rcm_order.argtypes = [I64P, C.c_int, I32P]
    L.ref_centroid_order.restype = None
    L.ref_centroid_order.argtypes = [I64P, C.c_int, I32P]
    L.ref_reorder_publish.restype = None
    L.ref_reorder_publish.argtypes = [I64P, I64P, C.c_int, C.c_int, C.c_int]
    L.ref_is_unimodular.restype = C.c_int
    L.ref_is_unimodular.argtypes = [I64P, C.c_int]
    L.ref_det_mod_prime.restype = C.c_int64
    L.ref_det_mod_prime.argtypes = [I64P, C.c_int, C.c_int64]
    L.ref_deflate_index.restype = C.c_int
    L.ref_deflate_index.argtypes = [I64P, I64P, C.c_int, C.c_int, I32P, C.c_int]
    return L


def P(a):
    return a.ctypes.data_as(I64P)


def sym(M):
    U = np.triu(M)
    return U + np.triu(U, 1).T


def gen_matrix(rng, n, kind):
    """Symmetric int64 test matrices of several shapes."""
    if kind == "small":
        A = rng.integers(-20, 21, (n, n))
    elif kind == "sparse":
        A = rng.integers(-5, 6, (n, n)) * (rng.random((n, n)) < 0.3)
    elif kind == "big":
        A = rng.integers(-(1 << 47), 1 << 47, (n, n))
    elif kind == "edge":
        A = rng.integers(-(1 << 20), 1 << 20, (n, n))
        k = rng.integers(0, n, max(1, n // 3))
        A[k, :] = rng.choice([LIM - 1, LIM, -LIM, -(LIM - 1), LIM >> 1, 3], size=(len(k), n))
    elif kind == "median":
        # pivot rows of +/-1, +/-2, +/-3 and target rows with large multiples:
        # breakpoints spread far beyond the +/-64 shear cap, ties common.
        A = rng.integers(-3, 4, (n, n)) * rng.integers(1, 400, (n, n))
        A[rng.random((n, n)) < 0.2] = 0
    elif kind == "diag":
        A = np.diag(rng.integers(-50, 51, n))
    elif kind == "clamp":
        # Gram of rows that are large multiples of one another: breakpoints far
        # outside +/-64, so the best shear clamps to the cap.
        L = rng.integers(-2, 3, (n, n))
        mult = rng.integers(70, 300, n) * rng.choice([-1, 1], n)
        mult[0] = 1
        L = L + mult[:, None] * L[0][None, :] * (rng.random(n) < 0.7)[:, None]
        return np.asarray(L @ L.T, dtype=np.int64)
    else:
        raise ValueError(kind)
    return sym(np.asarray(A, dtype=np.int64))


def gen_xt(rng, n, kind):
    if kind == "eye":
        return np.eye(n, dtype=np.int64)
    if kind == "small":
        return rng.integers(-3, 4, (n, n)).astype(np.int64)
    X = rng.integers(-(1 << 46), 1 << 46, (n, n)).astype(np.int64)
    X[rng.integers(0, n), :] = LIM - 2
    return X


def checksum(M):
    """Weighted sum mod 2^61-1 of all entries, for large-n commit outputs."""
    p = (1 << 61) - 1
    acc = 0
    for t, v in enumerate(np.asarray(M).ravel().tolist()):
        acc = (acc + (t + 1) * (int(v) % p)) % p
    return acc


def g1(L):
    rng = np.random.default_rng(20261203)
    sizes = list(range(2, 13)) + [24, 48, 64, 80]
    kinds = ["clamp", "median", "small", "sparse", "big", "edge", "diag"]
    out = {}
    for n in sizes:
        nmat = 12 if n <= 12 else (6 if n <= 48 else 3)
        ncase = 6 if n <= 12 else 4
        As, Xts, cases = [], [], []
        res = {k: [] for k in ["ok", "delta", "dnz", "row", "okx", "deltax", "dnzx", "rowx",
                               "xrow", "okt", "trow", "sdl", "best", "mass0", "masspost",
                               "Apost", "Xtpost", "cks"]}
        for mi in range(nmat):
            kind = kinds[mi % len(kinds)]
            A = gen_matrix(rng, n, kind)
            Xt = gen_xt(rng, n, ["eye", "small", "big"][mi % 3])
            As.append(A)
            Xts.append(Xt)
            mass0 = np.zeros(n, np.int64)
            L.ref_compute_row_mass(P(A), n, P(mass0))
            for ci in range(ncase):
                i = int(rng.integers(0, n))
                j = int(rng.integers(0, n - 1))
                j += j >= i
                best = L.ref_best_shear_l1(P(A), n, i, j)
                choice = ci % 4
                if choice == 0:
                    s = best
                elif choice == 1:
                    s = int(rng.choice([-1, 1]))
                elif choice == 2:
                    s = int(rng.integers(-64, 65))
                else:
                    s = int(rng.choice([64, -64, 2, -2, 0]))
                cases.append((mi, i, j, s))
                row = np.zeros(n, np.int64)
                d = C.c_int64(0)
                dz = C.c_int64(0)
                ok = L.ref_try_add(P(A), P(Xt), n, i, j, s, P(row), None, C.byref(d), C.byref(dz))
                rowx = np.zeros(n, np.int64)
                xrow = np.zeros(n, np.int64)
                dx = C.c_int64(0)
                dzx = C.c_int64(0)
                okx = L.ref_try_add(P(A), P(Xt), n, i, j, s, P(rowx), P(xrow), C.byref(dx),
                                    C.byref(dzx))
                trow = np.zeros(n, np.int64)
                okt = L.ref_add_transform_row(P(Xt), n, i, j, s, P(trow))
                sdl = L.ref_score_delta_l1(P(A), n, i, j, s)
                Apost = A.copy()
                Xtpost = Xt.copy()
                mpost = mass0.copy()
                if okx:
                    L.ref_commit_add(P(Apost), P(Xtpost), n, j, P(rowx), P(xrow), P(mpost))
                for k, v in [("ok", ok), ("delta", d.value), ("dnz", dz.value), ("row", row),
                             ("okx", okx), ("deltax", dx.value), ("dnzx", dzx.value),
                             ("rowx", rowx), ("xrow", xrow), ("okt", okt), ("trow", trow),
                             ("sdl", sdl), ("best", best), ("mass0", mass0),
                             ("masspost", mpost)]:
                    res[k].append(v)
                if n <= 24:
                    res["Apost"].append(Apost)
                    res["Xtpost"].append(Xtpost)
                res["cks"].append([checksum(Apost), checksum(Xtpost)])
            # exhaustive best_shear over all (i, j) pairs for small n
        out[f"n{n}_A"] = np.array(As)
        out[f"n{n}_Xt"] = np.array(Xts)
        out[f"n{n}_cases"] = np.array(cases, np.int64)
        for k, v in res.items():
            if v:
                out[f"n{n}_{k}"] = np.array(v, dtype=np.int64)
        # best_shear for every ordered pair of the first two matrices
        allb = []
        for mi in range(min(2, nmat)):
            A = As[mi]
            allb.append([[L.ref_best_shear_l1(P(A), n, i, j) if i != j else 0
                          for j in range(n)] for i in range(n)])
        out[f"n{n}_best_all"] = np.array(allb, np.int64)
        # score and nonzero count of every matrix
        out[f"n{n}_score"] = np.array([L.ref_score(P(A), n) for A in As], np.int64)
        out[f"n{n}_nnz"] = np.array([L.ref_offdiag_nonzero(P(A), n) for A in As], np.int64)
    np.savez_compressed(os.path.join(OUT, "g1_primitives.npz"), **out)

    # rounded_div: signs, exact ties, remainders either side of a half.
    vals = [0, 1, 2, 3, 4, 5, 6, 7, 9, 10, 11, 13, 1000001, (1 << 48) - 1, 1 << 48]
    rd = []
    for a in vals:
        for b in vals[1:]:
            for sa in (1, -1):
                for sb in (1, -1):
                    rd.append((sa * a, sb * b, L.ref_rounded_div(sa * a, sb * b)))
    np.savez_compressed(os.path.join(OUT, "g1_rounded_div.npz"), cases=np.array(rd, np.int64))

    # lattice helpers and congruence_of
    lat = {}
    for t, (m, d) in enumerate([(3, 3), (4, 6), (6, 4), (24, 24)]):
        Lm = rng.integers(-9, 10, (m, d)).astype(np.int64)
        G = np.zeros((m, m), np.int64)
        L.ref_gram_of(P(Lm), m, d, P(G))
        X = rng.integers(-2, 3, (m, m)).astype(np.int64)
        R = np.zeros((m, d), np.int64)
        L.ref_xt_times(P(X), P(Lm), m, d, P(R))
        A = gen_matrix(rng, m, "small")
        XAX = np.zeros((m, m), np.int64)
        L.ref_congruence_of(P(A), P(X), m, P(XAX))
        lat.update({f"c{t}_L": Lm, f"c{t}_G": G, f"c{t}_X": X, f"c{t}_R": R, f"c{t}_A": A,
                    f"c{t}_XAX": XAX})
    np.savez_compressed(os.path.join(OUT, "g1_lattice.npz"), **lat)


def planted(seed, n, dmax, nshear):
    """Small planted form U^T D U (Python ints), as the fixtures use."""
    rng = np.random.default_rng(seed)
    d = [int(v) for v in rng.integers(1, dmax + 1, n)]
    d = [v if rng.integers(0, 2) else -v for v in d]
    U = [[int(r == c) for c in range(n)] for r in range(n)]
    for _ in range(nshear):
        i = int(rng.integers(0, n))
        j = int(rng.integers(0, n - 1))
        j += j >= i
        s = 1 if rng.integers(0, 2) else -1
        for r in range(n):
            U[r][j] += s * U[r][i]
    A = [[sum(U[k][r] * d[k] * U[k][c] for k in range(n)) for c in range(n)] for r in range(n)]
    return np.array(A, np.int64)


def read_csv(path):
    rows = [r for r in open(path).read().splitlines() if r.strip()]
    return np.array([[int(x) for x in r.split(",")] for r in rows], np.int64)


def run_ref(A, args, timeout=60):
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "A.csv")
        with open(path, "w") as f:
            for row in A:
                f.write(",".join(str(int(v)) for v in row) + "\n")
        cp = subprocess.run([os.path.join(REF_DIR, "xtax_ref"), "-A", path] + args, cwd=td,
                            capture_output=True, text=True, timeout=timeout)
        moves = None
        solved = "matrix is diagonal" in cp.stdout
        for line in cp.stdout.splitlines():
            if line.startswith("done moves="):
                moves = int(line.split()[1].split("=")[1])
        return (solved, moves, read_csv(os.path.join(td, "best_A.csv")),
                read_csv(os.path.join(td, "best_X.csv")), cp.stdout)


def g2():
    A10 = read_csv("/root/reference/10x10.csv")
    runs = []
    cases = [("10x10", A10, ["-t", "1", "--seed", str(s), "--max-seconds", "20"])
             for s in (1, 2, 3, 4)]
    for s in (11, 12, 13):
        for n in (4, 5, 6):
            cases.append((f"planted{n}_{s}", planted(s, n, 9, 2 * n),
                          ["-t", "1", "--seed", str(s), "--max-seconds", "5"]))
    cases.append(("10x10_t4", A10, ["-t", "4", "--seed", "5", "--max-seconds", "10"]))
    cases.append(("planted8_t4", planted(7, 8, 100, 24),
                  ["-t", "4", "--seed", "5", "--max-seconds", "5", "--t-init", "200",
                   "--t-min", "5"]))
    out = {}
    meta = []
    for t, (name, A0, args) in enumerate(cases):
        solved, moves, bA, bX, _ = run_ref(A0, args)
        det = solved and "-t" in args and args[args.index("-t") + 1] == "1"
        meta.append((t, int(solved), int(moves or -1), int(det)))
        out[f"r{t}_A0"] = A0
        out[f"r{t}_A"] = bA
        out[f"r{t}_X"] = bX
        print(f"G2 {name}: solved={solved} moves={moves} deterministic={det}", file=sys.stderr)
    out["meta"] = np.array(meta, np.int64)
    out["names"] = np.array([c[0] for c in cases])
    out["args"] = np.array([" ".join(c[2]) for c in cases])
    np.savez_compressed(os.path.join(OUT, "g2_end_to_end.npz"), **out)


def gen_pattern(rng, n, kind):
    """Symmetric matrices for the reordering fixtures: sparsity patterns that
    exercise RCM's root choice, degree ties and components, and magnitude
    patterns that exercise the centroid keys (ties included)."""
    if kind == "sparse":
        A = rng.integers(-9, 10, (n, n)) * (rng.random((n, n)) < rng.uniform(0.05, 0.5))
    elif kind == "blocks":  # disconnected components, isolated indices
        A = np.zeros((n, n), np.int64)
        cuts = sorted(set(int(c) for c in rng.integers(1, max(2, n), 3)))
        lo = 0
        for hi in cuts + [n]:
            if hi - lo >= 2 and rng.random() < 0.8:
                B = rng.integers(-4, 5, (hi - lo, hi - lo)) * (rng.random((hi - lo,) * 2) < 0.6)
                A[lo:hi, lo:hi] = B
            lo = hi
        perm = rng.permutation(n)
        A = A[perm][:, perm]
    elif kind == "path":  # a permuted path / band: RCM recovers a narrow band
        A = np.zeros((n, n), np.int64)
        for i in range(n - 1):
            A[i, i + 1] = int(rng.integers(1, 5))
        A += np.diag(rng.integers(1, 5, n))
        perm = rng.permutation(n)
        A = A[perm][:, perm]
    elif kind == "ties":  # equal magnitudes: equal degrees and equal centroid keys
        A = rng.choice([-1, 0, 1], (n, n), p=[0.25, 0.5, 0.25])
    elif kind == "dense":
        A = rng.integers(-1000, 1001, (n, n))
    elif kind == "star":
        A = np.zeros((n, n), np.int64)
        c = int(rng.integers(0, n))
        A[c, :] = rng.integers(1, 7, n)
    else:
        raise ValueError(kind)
    return sym(np.asarray(A, dtype=np.int64))


def unimodular(rng, n, shears, sign=None):
    """U^T D U with D = diag(+-1): det = +-1."""
    d = rng.choice([-1, 1], n) if sign is None else np.full(n, sign)
    U = np.eye(n, dtype=np.int64)
    for _ in range(shears if n >= 2 else 0):
        i, j = rng.choice(n, 2, replace=False)
        U[:, j] += int(rng.choice([-1, 1])) * U[:, i]
    return np.asarray((U.T * d) @ U, dtype=np.int64), U


def load_dual_harness():
    L = C.CDLL(os.path.join(REF_DIR, "libref_dual.so"))
    DP = C.POINTER(C.c_double)
    L.refd_eval.restype = C.c_int
    L.refd_eval.argtypes = [I64P, DP, C.c_int, C.c_int, C.c_int, C.c_int64, I64P, DP, DP, DP,
                            I64P]
    L.refd_best_shear.restype = C.c_int64
    L.refd_best_shear.argtypes = [I64P, DP, C.c_int, C.c_double, C.c_int, C.c_int]
    L.refd_invert.restype = C.c_int
    L.refd_invert.argtypes = [I64P, C.c_int, DP]
    L.refd_frob2_off_p.restype = C.c_double
    L.refd_frob2_off_p.argtypes = [I64P, C.c_int]
    L.refd_frob2_off_q.restype = C.c_double
    L.refd_frob2_off_q.argtypes = [DP, C.c_int]
    return L


def g4(D):
    """xdual: eval_add / eval_add_d, DualObjective::best_shear, invert_to and
    the off-diagonal norms (src/xdual_main.cpp:67-364, src/mat_io.hpp:297-337)
    on seeded full-rank Grams."""
    rng = np.random.default_rng(44)
    DP = C.POINTER(C.c_double)
    out = {}
    t = 0
    for n in [2, 3, 4, 6, 9, 12, 17, 24, 33, 48]:
        for kind in ["lattice", "planted", "big"]:
            if kind == "lattice":
                while True:
                    Lm = rng.integers(-6, 7, (n, n))
                    if abs(np.linalg.det(Lm)) > 0.5:
                        break
                P = np.asarray(Lm @ Lm.T, dtype=np.int64)
            elif kind == "planted":
                d = rng.integers(1, 30, n) * rng.choice([-1, 1], n)
                U = np.eye(n, dtype=np.int64)
                for _ in range(2 * n):
                    a, b = rng.choice(n, 2, replace=False)
                    U[:, b] += int(rng.choice([-1, 1])) * U[:, a]
                P = np.asarray((U.T * d) @ U, dtype=np.int64)
            else:
                while True:
                    Lm = rng.integers(-(1 << 20), 1 << 20, (n, n))
                    if abs(np.linalg.det(Lm.astype(float))) > 1.0:
                        break
                P = np.asarray(Lm.astype(object).dot(Lm.T.astype(object)), dtype=np.int64)
            Q = np.zeros((n, n))
            ok = L_inv = D.refd_invert(P.ctypes.data_as(I64P), n, Q.ctypes.data_as(DP))
            out[f"c{t}_P"] = P
            out[f"c{t}_ok"] = np.array(ok)
            out[f"c{t}_Q"] = Q
            out[f"c{t}_fp"] = np.array(D.refd_frob2_off_p(P.ctypes.data_as(I64P), n))
            out[f"c{t}_fq"] = np.array(D.refd_frob2_off_q(Q.ctypes.data_as(DP), n))
            c = float(rng.choice([0.0, 0.5, 1.0, 37.0, 1e-3])) * (
                out[f"c{t}_fp"] / out[f"c{t}_fq"] if out[f"c{t}_fq"] > 0 else 1.0)
            out[f"c{t}_c"] = np.array(c)
            qs, ev, evq, evp = [], [], [], []
            for _ in range(60):
                i = int(rng.integers(0, n)); j = int(rng.integers(0, n - 1)); j += j >= i
                s = int(rng.choice([-64, -3, -1, 1, 2, 64, int(rng.integers(-64, 65))]))
                prow = np.zeros(n, np.int64); qrow = np.zeros(n)
                dp = C.c_double(); dq = C.c_double(); dnz = C.c_int64()
                okm = D.refd_eval(P.ctypes.data_as(I64P), Q.ctypes.data_as(DP), n, i, j, s,
                                  prow.ctypes.data_as(I64P), qrow.ctypes.data_as(DP),
                                  C.byref(dp), C.byref(dq), C.byref(dnz))
                qs.append((i, j, s))
                ev.append((okm, dnz.value))
                evq.append((dp.value, dq.value))
                evp.append(np.concatenate([prow.astype(float), qrow]) if okm else np.zeros(2 * n))
            out[f"c{t}_q"] = np.array(qs, np.int64)
            out[f"c{t}_ev"] = np.array(ev, np.int64)
            out[f"c{t}_evd"] = np.array(evq)
            out[f"c{t}_rows_q"] = np.array([r[n:] for r in evp])
            best = np.zeros((n, n), np.int64)
            for i in range(n):
                for j in range(n):
                    if i != j:
                        best[i, j] = D.refd_best_shear(P.ctypes.data_as(I64P),
                                                       Q.ctypes.data_as(DP), n, c, i, j)
            out[f"c{t}_best"] = best
            t += 1
    out["cases"] = np.array(t)
    np.savez_compressed(os.path.join(OUT, "g4_xdual.npz"), **out)


def g3(L):
    """Publish-time reordering (rcm_order, centroid_order, reorder_for_publish),
    is_unimodular / det_mod_prime and deflate_index, from the reference."""
    rng = np.random.default_rng(20261203)
    out = {}
    t = 0
    for kind in ["sparse", "blocks", "path", "ties", "dense", "star"]:
        for n in [1, 2, 3, 5, 8, 13, 17, 24, 31, 40]:
            A = gen_pattern(rng, n, kind)
            Xt = rng.integers(-5, 6, (n, n)).astype(np.int64)
            rc = np.zeros(n, np.int32)
            ce = np.zeros(n, np.int32)
            L.ref_rcm_order(P(A), n, rc.ctypes.data_as(C.POINTER(C.c_int)))
            L.ref_centroid_order(P(A), n, ce.ctypes.data_as(C.POINTER(C.c_int)))
            out[f"r{t}_A"] = A
            out[f"r{t}_Xt"] = Xt
            out[f"r{t}_rcm"] = rc
            out[f"r{t}_centroid"] = ce
            for flags in (1, 2, 3):
                A2, X2 = A.copy(), Xt.copy()
                L.ref_reorder_publish(P(A2), P(X2), n, flags & 1, (flags >> 1) & 1)
                out[f"r{t}_A{flags}"] = A2
                out[f"r{t}_Xt{flags}"] = X2
            t += 1
    out["reorder_cases"] = np.array(t)
    # is_unimodular and det mod each of its primes
    u = 0
    for n in [1, 2, 3, 4, 6, 9, 14, 20, 33]:
        for variant in ["uni", "uni_neg", "det2", "det3", "singular", "random", "bigent"]:
            if variant == "uni":
                A, _ = unimodular(rng, n, 3 * n)
            elif variant == "uni_neg":
                A, _ = unimodular(rng, n, 3 * n, sign=-1)
            elif variant in ("det2", "det3"):
                _, U = unimodular(rng, n, 2 * n)
                k = int(rng.integers(0, n))
                d = rng.choice([-1, 1], n)
                d[k] *= int(variant[-1])
                A = np.asarray((U.T * d) @ U, dtype=np.int64)
            elif variant == "singular":
                B = rng.integers(-4, 5, (n, max(1, n - 1)))
                A = np.asarray(B @ B.T, dtype=np.int64)
            elif variant == "random":
                A = sym(rng.integers(-30, 31, (n, n)).astype(np.int64))
            else:
                A = sym(rng.integers(-(1 << 40), 1 << 40, (n, n)).astype(np.int64))
            dets = [L.ref_det_mod_prime(P(A), n, p) for p in (1000000007, 998244353)]
            out[f"u{u}_A"] = A
            out[f"u{u}_uni"] = np.array(L.ref_is_unimodular(P(A), n))
            out[f"u{u}_det"] = np.array(dets, np.int64)
            u += 1
    out["unimodular_cases"] = np.array(u)
    # deflate_index: a +/-1 pivot cleared over an active set, or rolled back
    d = 0
    for n in [2, 3, 5, 8, 12, 20, 32]:
        for variant in ["plain", "subset", "near_bound", "xt_bound", "neg", "m_bound"]:
            M = sym(rng.integers(-50, 51, (n, n)).astype(np.int64))
            i = int(rng.integers(0, n))
            M[i, i] = -1 if variant == "neg" else int(rng.choice([-1, 1]))
            Xt = np.eye(n, dtype=np.int64) if variant != "xt_bound" else \
                rng.integers(-3, 4, (n, n)).astype(np.int64)
            act = list(range(n))
            if variant == "subset":
                act = sorted(set(int(v) for v in rng.choice(n, max(2, n // 2), replace=False))
                             | {i})
            if variant == "near_bound" and n > 2:
                k = (i + 1) % n
                M[k, :] = M[:, k] = rng.integers(-(1 << 47), 1 << 47, n)
                M[k, i] = M[i, k] = int(rng.integers(1 << 20, 1 << 22))
                M[i, i] = 1
            if variant == "xt_bound":
                Xt[i, :] = (1 << 47) + int(rng.integers(0, 1 << 40))
            if variant == "m_bound":
                # earlier targets clear, a later one overflows: full rollback
                k = n - 1 if i != n - 1 else 0
                M[i, :] = M[:, i] = 1
                M[i, i] = 1
                for c in range(n):
                    if c not in (i, k):
                        M[k, c] = M[c, k] = LIM
                M[k, i] = M[i, k] = -1
            a = np.array(act, np.int32)
            M2, X2 = M.copy(), Xt.copy()
            ok = L.ref_deflate_index(P(M2), P(X2), n, i, a.ctypes.data_as(C.POINTER(C.c_int)),
                                     len(act))
            out[f"d{d}_M"] = M
            out[f"d{d}_Xt"] = Xt
            out[f"d{d}_i"] = np.array(i)
            out[f"d{d}_active"] = a
            out[f"d{d}_ok"] = np.array(ok)
            out[f"d{d}_M2"] = M2
            out[f"d{d}_Xt2"] = X2
            d += 1
    out["deflate_cases"] = np.array(d)
    np.savez_compressed(os.path.join(OUT, "g3_publish_deflate.npz"), **out)


def g5_configs(rng, count):
    """Random worker configurations for the trace fixtures: matrix families
    (random symmetric, planted, bound-edge magnitudes, lattice Grams), both
    temperature modes, every stall setting, active lists and start
    transforms."""
    from oracle import refharness
    sys.path.insert(0, ROOT)
    from xtax_amd import workloads
    out = []
    while len(out) < count:
        n = int(rng.integers(2, 33))
        fam = int(rng.integers(0, 4))
        if fam == 0:
            A = workloads.random_symmetric(n, -20, 20, seed=int(rng.integers(1 << 30)))
        elif fam == 1:
            A = workloads.planted(n, 100, 3 * n, seed=int(rng.integers(1 << 30)))[0]
        elif fam == 2:
            sc = int(rng.choice([3, 1000, 1 << 40, LIM - 1]))
            M = rng.integers(-sc, sc + 1, (n, n)) * (rng.random((n, n)) < 0.7)
            A = (np.triu(M) + np.triu(M, 1).T).astype(np.int64)
        else:
            A = workloads.lattice(min(n, 24), min(n, 24), -9, 9,
                                  seed=int(rng.integers(1 << 30)))[1]
            n = A.shape[0]
        threads = int(rng.choice([1, 2, 4, 8, 16]))
        kw = dict(tempering=int(rng.integers(0, 2)),
                  greedy_fraction=float(rng.choice([0, 0.3, 0.5, 1.0])),
                  target_fraction=float(rng.choice([0, 0.5, 1.0])),
                  target_samples=int(rng.integers(1, 17)),
                  worker_diversity=float(rng.choice([0, 0.4, 1.0])),
                  sweep_fraction=float(rng.choice([0, 0.05, 0.5, 1.0])),
                  stuck_threshold=int(rng.choice([20, 100, 1000, 20000])),
                  t_init=float(rng.choice([0, 1, 30, 500])),
                  t_min=float(rng.choice([1e-3, 0.5, 5])),
                  cooling=float(rng.choice([0.999, 0.95])),
                  moves_per_cool=int(rng.choice([1, 20, 200])),
                  reheat=float(rng.choice([1.0, 0.5])),
                  reseed_factor=float(rng.choice([1.0, 1.25, 3.0])),
                  adaptive_cooling=int(rng.integers(0, 2)))
        act = np.zeros(0, np.int32)
        if n >= 3 and rng.random() < 0.3:
            act = np.sort(rng.choice(n, int(rng.integers(2, n + 1)), replace=False)).astype(np.int32)
        X0 = np.eye(n, dtype=np.int64)
        if fam != 2 and rng.random() < 0.25:
            for _ in range(n):
                a, b = rng.choice(n, 2, replace=False)
                X0[:, b] += int(rng.choice([-1, 1])) * X0[:, a]
        A_start = (X0.T.astype(object) @ A.astype(object) @ X0.astype(object)).astype(np.int64)
        dp, ip = refharness.pack(kw, threads, int(rng.integers(0, threads)),
                                 int(rng.integers(0, 1 << 32)),
                                 int(rng.choice([200, 2000, 6000])),
                                 float(rng.choice([0.5, 5, 50, 1000])))
        out.append((A_start, X0, act, dp, ip))
    return out


def g5():
    """Worker traces of the reference's own anneal_worker (recorded by
    oracle/_ref/libref_harness.so) and single-threaded reference runs with
    non-default flags, for the model's reference-draw mode."""
    from oracle import refharness
    rng = np.random.default_rng(2025)
    out = {}
    for t, (A, X0, act, dp, ip) in enumerate(g5_configs(rng, 64)):
        r = refharness.trace(A, X0, dp, ip, act)
        out.update({f"t{t}_A": A, f"t{t}_X0": X0, f"t{t}_act": act, f"t{t}_dp": dp,
                    f"t{t}_ip": ip, f"t{t}_Af": r["A"], f"t{t}_Xtf": r["Xt"],
                    f"t{t}_gbA": r["gbA"], f"t{t}_gbXt": r["gbXt"],
                    f"t{t}_log": np.stack([r["log"]["move"], r["log"]["i"], r["log"]["j"],
                                           r["log"]["s"], r["log"]["flags"]], 1).astype(np.int64),
                    f"t{t}_sc": np.array([r["moves"], r["score"], r["gb_score"], r["solved"]])})
    out["traces"] = np.array(64)

    # the reference binary, one thread (deterministic until it stops at a
    # diagonal), with non-default flags
    A10 = read_csv("/root/reference/10x10.csv")
    P5 = planted(31, 5, 9, 10)
    P6 = planted(32, 6, 9, 12)
    P7 = planted(33, 7, 20, 14)
    # a planted form on rows {0, 1, 2, 4} of order 6, rows 3 and 5 already split off
    G = np.zeros((6, 6), np.int64)
    sub = planted(34, 4, 9, 8)
    idx = [0, 1, 2, 4]
    for a in range(4):
        for b in range(4):
            G[idx[a], idx[b]] = sub[a, b]
    G[3, 3], G[5, 5] = 7, -3
    # a start transform for P6 (unimodular: a product of shears)
    Xs = np.eye(6, dtype=np.int64)
    for (a, b, sgn) in [(0, 3, 1), (2, 5, -1), (4, 1, 1), (3, 2, -1)]:
        Xs[:, b] += sgn * Xs[:, a]
    cases = [
        ("10x10_sweep", A10, ["--seed", "21", "--sweep-fraction", "0.5", "--stuck-threshold", "300"]),
        ("10x10_adaptive", A10, ["--seed", "22", "--adaptive-cooling", "--moves-per-cool", "50"]),
        ("10x10_samples3", A10, ["--seed", "23", "--target-samples", "3", "--target-fraction", "1.0"]),
        ("10x10_greedy1", A10, ["--seed", "24", "--greedy-fraction", "1.0"]),
        ("p7_greedy0", P7, ["--seed", "25", "--greedy-fraction", "0.0", "--t-init", "10"]),
        ("p7_reseed", P7, ["--seed", "26", "--reheat", "0.5", "--reseed-factor", "1.0",
                           "--stuck-threshold", "200"]),
        ("p6_cool", P6, ["--seed", "27", "--cooling", "0.99", "--t-min", "0.5"]),
        ("p5_target0", P5, ["--seed", "28", "--target-fraction", "0.0"]),
        ("gram_rows", G, ["--seed", "29", "--gram-rows", "0..2,4"]),
        ("p6_start_X", P6, ["--seed", "30", "-X", "X.csv"]),
        ("10x10_no_tempering_t1", A10, ["--seed", "31", "--no-tempering", "--t-init", "5"]),
    ]
    k = 0
    for name, A0, args in cases:
        X0 = Xs if "-X" in args else None
        solved, moves, bA, bX = run_ref_x(A0, X0, ["-t", "1", "--max-seconds", "20"] + args)
        print(f"G5 cli {name}: solved={solved} moves={moves}", file=sys.stderr)
        if not solved:
            continue
        out.update({f"c{k}_A0": A0, f"c{k}_X0": X0 if X0 is not None else np.zeros((0, 0), np.int64),
                    f"c{k}_A": bA, f"c{k}_X": bX, f"c{k}_moves": np.array(moves)})
        out[f"c{k}_args"] = np.array(" ".join(args))
        out[f"c{k}_name"] = np.array(name)
        k += 1
    out["cli_runs"] = np.array(k)
    np.savez_compressed(os.path.join(OUT, "g5_worker.npz"), **out)


def g5b():
    """Worker traces of the reference's anneal_worker with tournaments of more
    than 16 samples (17 to 257), the G5 configurations otherwise."""
    from oracle import refharness
    rng = np.random.default_rng(2026)
    out = {}
    sizes = [17, 24, 31, 40, 64, 100, 257]
    for t, (A, X0, act, dp, ip) in enumerate(g5_configs(rng, 21)):
        ip = ip.copy()
        ip[2] = sizes[t % len(sizes)]
        r = refharness.trace(A, X0, dp, ip, act)
        out.update({f"t{t}_A": A, f"t{t}_X0": X0, f"t{t}_act": act, f"t{t}_dp": dp,
                    f"t{t}_ip": ip, f"t{t}_Af": r["A"], f"t{t}_Xtf": r["Xt"],
                    f"t{t}_gbA": r["gbA"], f"t{t}_gbXt": r["gbXt"],
                    f"t{t}_log": np.stack([r["log"]["move"], r["log"]["i"], r["log"]["j"],
                                           r["log"]["s"], r["log"]["flags"]], 1).astype(np.int64),
                    f"t{t}_sc": np.array([r["moves"], r["score"], r["gb_score"], r["solved"]])})
    out["traces"] = np.array(21)
    np.savez_compressed(os.path.join(OUT, "g5b_samples.npz"), **out)


def run_ref_x(A, X0, args, timeout=60):
    """run_ref with an optional -X start transform written as X.csv."""
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "A.csv")
        with open(path, "w") as f:
            for row in A:
                f.write(",".join(str(int(v)) for v in row) + "\n")
        if X0 is not None:
            with open(os.path.join(td, "X.csv"), "w") as f:
                for row in X0:
                    f.write(",".join(str(int(v)) for v in row) + "\n")
        cp = subprocess.run([os.path.join(REF_DIR, "xtax_ref"), "-A", path] + args, cwd=td,
                            capture_output=True, text=True, timeout=timeout)
        moves = None
        for line in cp.stdout.splitlines():
            if line.startswith("done moves="):
                moves = int(line.split()[1].split("=")[1])
        return ("matrix is diagonal" in cp.stdout, moves, read_csv(os.path.join(td, "best_A.csv")),
                read_csv(os.path.join(td, "best_X.csv")))


def run_dual_ref(A, X0, args, timeout=60):
    """xdual_ref -A on A (and -X X0) in a scratch directory: (solved, moves,
    best_P, best_Q, best_X)."""
    with tempfile.TemporaryDirectory() as td:
        for name, M in [("A.csv", A), ("X.csv", X0)]:
            if M is None:
                continue
            with open(os.path.join(td, name), "w") as f:
                for row in M:
                    f.write(",".join(str(int(v)) for v in row) + "\n")
        try:
            cp = subprocess.run([os.path.join(REF_DIR, "xdual_ref"), "-A", "A.csv"] + args,
                                cwd=td, capture_output=True, text=True, timeout=timeout)
        except subprocess.TimeoutExpired:
            return False, None, None, None, None
        moves = None
        for line in cp.stdout.splitlines():
            if line.startswith("done moves="):
                moves = int(line.split()[1].split("=")[1])
        solved = "matrix is diagonal" in cp.stdout
        Q = np.array([[float(x) for x in r.split(",")]
                      for r in open(os.path.join(td, "best_Q.csv")).read().splitlines() if r.strip()])
        return (solved, moves, read_csv(os.path.join(td, "best_P.csv")), Q,
                read_csv(os.path.join(td, "best_X.csv")))


def g6():
    """Single-threaded runs of the reference's xdual binary with fixed seeds,
    for the dual model's reference-draw mode."""
    from concurrent.futures import ThreadPoolExecutor
    P5 = planted(1, 5, 9, 8)
    P6 = planted(2, 6, 9, 10)
    P7 = planted(6, 7, 5, 12)
    P9 = planted(7, 9, 12, 14)
    # a planted form on rows {0, 1, 2, 4} of order 6, rows 3 and 5 already split off
    G = np.zeros((6, 6), np.int64)
    sub = planted(34, 4, 9, 8)
    idx = [0, 1, 2, 4]
    for a in range(4):
        for b in range(4):
            G[idx[a], idx[b]] = sub[a, b]
    G[3, 3], G[5, 5] = 7, -3
    Xs = np.eye(6, dtype=np.int64)
    for (a, b, sgn) in [(0, 3, 1), (2, 5, -1), (4, 1, 1), (3, 2, -1)]:
        Xs[:, b] += sgn * Xs[:, a]
    cases = [
        ("p5_default", P5, ["--seed", "7"]),
        ("p6_lambda0.1", P6, ["--seed", "11", "--lambda", "0.1"]),
        ("p6_sweep", P6, ["--seed", "11", "--sweep-fraction", "1"]),
        ("p7_lambda0.1", P7, ["--seed", "11", "--lambda", "0.1"]),
        ("p7_lambda0.01", P7, ["--seed", "11", "--lambda", "0.01"]),
        ("p7_lambda0", P7, ["--seed", "11", "--lambda", "0"]),
        ("p7_target0.9", P7, ["--seed", "11", "--lambda", "0.3", "--target-fraction", "0.9"]),
        ("p9_lambda0.01", P9, ["--seed", "11", "--lambda", "0.01"]),
        ("p9_target0.9", P9, ["--seed", "11", "--lambda", "0.3", "--target-fraction", "0.9"]),
        ("p7_refresh", P7, ["--seed", "32", "--lambda", "0.01", "--dual-refresh", "20000"]),
        ("p7_check", P7, ["--seed", "32", "--lambda", "0.01", "--dual-check", "5000",
                          "--dual-tol", "1e-15"]),
        ("p7_reseed", P7, ["--seed", "22", "--lambda", "0.1", "--reseed-factor", "1.0",
                           "--stuck-threshold", "2000"]),
        ("p7_adaptive", P7, ["--seed", "15", "--lambda", "0.1", "--adaptive-cooling",
                             "--moves-per-cool", "50"]),
        ("p7_greedy1", P7, ["--seed", "16", "--lambda", "0.1", "--greedy-fraction", "1.0"]),
        ("p7_greedy0", P7, ["--seed", "17", "--lambda", "0.1", "--greedy-fraction", "0.0",
                            "--t-init", "10"]),
        ("p7_samples3", P7, ["--seed", "18", "--lambda", "0.1", "--target-samples", "3",
                             "--target-fraction", "1.0"]),
        ("p6_cool", P6, ["--seed", "19", "--lambda", "0.1", "--cooling", "0.99", "--t-min", "0.5"]),
        ("gram_rows", G, ["--seed", "20", "--lambda", "0.1", "--gram-rows", "0..2,4"]),
        ("p6_start_X", P6, ["--seed", "12", "-X", "X.csv"]),
    ]

    def go(case):
        name, A0, args = case
        X0 = Xs if "-X" in args else None
        return case, run_dual_ref(A0, X0, ["-t", "1", "--max-seconds", "20"] + args, timeout=60)

    out = {}
    k = 0
    with ThreadPoolExecutor(4) as ex:
        for (name, A0, args), (solved, moves, bP, bQ, bX) in ex.map(go, cases):
            print(f"G6 {name}: solved={solved} moves={moves}", file=sys.stderr)
            if not solved or moves > 3_000_000:
                continue
            X0 = Xs if "-X" in args else None
            out.update({f"c{k}_A0": A0, f"c{k}_X0": X0 if X0 is not None else np.zeros((0, 0), np.int64),
                        f"c{k}_P": bP, f"c{k}_Q": bQ, f"c{k}_X": bX, f"c{k}_moves": np.array(moves)})
            out[f"c{k}_args"] = np.array(" ".join(args))
            out[f"c{k}_name"] = np.array(name)
            k += 1
    out["runs"] = np.array(k)
    np.savez_compressed(os.path.join(OUT, "g6_xdual_runs.npz"), **out)


def g6b_configs(rng, count):
    """Random worker configurations on xdual's objective: full-rank planted
    forms and lattice Grams, both temperature modes, every stall setting,
    active lists, start transforms, dual weights and maintenance periods."""
    from oracle import pyoracle as po
    from oracle import refharness
    sys.path.insert(0, ROOT)
    from xtax_amd import workloads
    out = []
    while len(out) < count:
        n = int(rng.integers(2, 13))
        fam = int(rng.integers(0, 3))
        if fam == 0:
            A = planted(int(rng.integers(1 << 30)), n, 20, 2 * n)
        elif fam == 1:
            A = workloads.lattice(n, n, -6, 6, seed=int(rng.integers(1 << 30)))[1]
        else:
            A = planted(int(rng.integers(1 << 30)), n, 3, 3 * n)
        Q = np.zeros((n, n))
        if not po.lib().orc_invert(A.ctypes.data_as(I64P), n,
                                   Q.ctypes.data_as(C.POINTER(C.c_double))):
            continue
        threads = int(rng.choice([1, 2, 4, 8]))
        kw = dict(tempering=int(rng.integers(0, 2)),
                  greedy_fraction=float(rng.choice([0, 0.3, 0.5, 1.0])),
                  target_fraction=float(rng.choice([0, 0.5, 1.0])),
                  target_samples=int(rng.integers(1, 17)),
                  worker_diversity=float(rng.choice([0, 0.4, 1.0])),
                  sweep_fraction=float(rng.choice([0, 0.05, 0.5, 1.0])),
                  stuck_threshold=int(rng.choice([20, 100, 1000, 20000])),
                  t_init=float(rng.choice([0, 1, 30, 500])),
                  t_min=float(rng.choice([1e-3, 0.5, 5])),
                  cooling=float(rng.choice([0.999, 0.95])),
                  moves_per_cool=int(rng.choice([1, 20, 200])),
                  reheat=float(rng.choice([1.0, 0.5])),
                  reseed_factor=float(rng.choice([1.0, 1.25, 3.0])),
                  adaptive_cooling=int(rng.integers(0, 2)))
        act = np.zeros(0, np.int32)
        if n >= 3 and rng.random() < 0.3:
            act = np.sort(rng.choice(n, int(rng.integers(2, n + 1)), replace=False)).astype(np.int32)
        X0 = np.eye(n, dtype=np.int64)
        if rng.random() < 0.25:
            for _ in range(n):
                a, b = rng.choice(n, 2, replace=False)
                X0[:, b] += int(rng.choice([-1, 1])) * X0[:, a]
        P0 = (X0.T.astype(object) @ A.astype(object) @ X0.astype(object)).astype(np.int64)
        dp, ip = refharness.dual_pack(
            kw, threads, int(rng.integers(0, threads)), int(rng.integers(0, 1 << 32)),
            int(rng.choice([2000, 8000, 20000])), float(rng.choice([0.5, 5, 50, 500])),
            float(rng.choice([0, 0.01, 0.1, 1.0])), int(rng.choice([0, 3000, 200000])),
            int(rng.choice([0, 500, 25000])), float(rng.choice([1e-6, 1e-15])))
        out.append((P0, X0, act, dp, ip))
    return out


def g6b():
    """Worker traces of the reference's anneal_worker on xdual's objective
    (recorded by oracle/_ref/libref_dual.so), for the dual model's reference
    mode."""
    from oracle import refharness
    rng = np.random.default_rng(2027)
    out = {}
    for t, (P0, X0, act, dp, ip) in enumerate(g6b_configs(rng, 48)):
        r = refharness.dual_trace(P0, X0, dp, ip, act)
        out.update({f"t{t}_P0": P0, f"t{t}_X0": X0, f"t{t}_act": act, f"t{t}_dp": dp,
                    f"t{t}_ip": ip, f"t{t}_Pf": r["Pf"], f"t{t}_Qf": r["Qf"],
                    f"t{t}_Xtf": r["Xtf"], f"t{t}_P": r["P"], f"t{t}_Q": r["Q"],
                    f"t{t}_Xt": r["Xt"],
                    f"t{t}_log": np.stack([r["log"]["move"], r["log"]["i"], r["log"]["j"],
                                           r["log"]["s"], r["log"]["flags"]], 1).astype(np.int64),
                    f"t{t}_sc": np.array([r["moves"], r["solved"]]),
                    f"t{t}_dsc": np.array([r["score"], r["gb_score"], r["c"]])})
    out["traces"] = np.array(48)
    np.savez_compressed(os.path.join(OUT, "g6b_dual_worker.npz"), **out)


if __name__ == "__main__":
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    L = load_harness()
    which = sys.argv[1:] or ["g1", "g2", "g3", "g4", "g5", "g6", "g6b"]
    if "g1" in which:
        g1(L)
    if "g2" in which:
        g2()
    if "g3" in which:
        g3(L)
    if "g4" in which:
        g4(load_dual_harness())
    if "g5" in which:
        g5()
    if "g5b" in which:
        g5b()
    if "g6" in which:
        g6()
    if "g6b" in which:
        g6b()
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)), file=sys.stderr)
