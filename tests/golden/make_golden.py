"""Generate the golden fixtures under tests/golden/ from the compiled reference.

Runs only where the reference sources exist (oracle/_ref/libflensref.so is
built from them by oracle/Makefile).  Inputs come from the build's own seeded
uniform generator (oracle.uniform); the fixtures store only seeds, shapes,
SHA-256 digests of the reference's output bits, pivot vectors, INFO values
and a few sampled entries -- numbers, never reference source.

    python tests/golden/make_golden.py            # small grids (seconds)
    python tests/golden/make_golden.py --big      # n=8192 LU per nb (~20 min)
    python tests/golden/make_golden.py --gemm2048 # the 2048^3 DGEMM (~1 min)
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def gemm_grid():
    """dblat3-style grid: N in {0,1,2,3,7,31,63}, alpha {0,1,0.7},
    beta {0,1,1.3}, all transposes (flens/blas/interface/test.tgz dblat3.dat)."""
    cases = []
    sizes = [0, 1, 2, 3, 7, 31, 63]
    seed = 100
    for ta in (0, 1):
        for tb in (0, 1):
            for (m, n, k) in [(s, s, s) for s in sizes] + [(7, 31, 3), (63, 2, 31), (31, 63, 7)]:
                for alpha in (0.0, 1.0, 0.7):
                    for beta in (0.0, 1.0, 1.3):
                        seed += 3
                        a = O.uniform(k if ta else m, m if ta else k, seed)
                        b = O.uniform(n if tb else k, k if tb else n, seed + 1)
                        c = O.uniform(m, n, seed + 2)
                        O.dgemm(ta, tb, alpha, a, b, beta, c, use_ref=True)
                        cases.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta,
                                          seed=seed, sha=sha(c)))
    return cases


def lu_grid():
    """dchkge-style shapes (M in {0,1,2,3,5,10,50,105,122}, N up to 132,
    NRHS {1,2,15,16}), uniform inputs, nb in {1,3,20,64}."""
    cases = []
    seed = 5000
    shapes = [(1, 1), (2, 2), (3, 3), (5, 5), (10, 10), (50, 50), (105, 105), (122, 122),
              (132, 132), (50, 10), (10, 50), (122, 105), (300, 300), (257, 257), (520, 520)]
    for (m, n) in shapes:
        for nb in (1, 3, 20, 64):
            seed += 2
            a = O.uniform(m, n, seed)
            piv, info = O.dgetrf(a, nb=nb, use_ref=True)
            case = dict(m=m, n=n, nb=nb, seed=seed, info=int(info), piv=piv.tolist(),
                        sha_lu=sha(a))
            if m == n:
                for nrhs in (1, 2, 15, 16):
                    for trans in (0, 1):
                        b = O.uniform(n, nrhs, seed + 1)
                        O.dgetrs(trans, a, piv, b, use_ref=True)
                        case[f"sha_x_{nrhs}_{trans}"] = sha(b)
            cases.append(case)
    return cases


def special_cases():
    """Closed-form / structural cases: the 4x4 example of
    flens/examples/lapack-getrf.cc:43-46, exact zeros, a tie, permutation
    and diagonal matrices."""
    out = {}
    ab = np.asfortranarray(np.array([[2, 3, -1, 0, 20], [-6, -5, 0, 2, -33],
                                     [2, -5, 6, -6, -43], [4, 6, 2, -3, 49]], float))
    piv, info = O.dgetrf(ab, use_ref=True)
    out["kat_getrf"] = dict(piv=piv.tolist(), info=int(info), lu=ab.T.ravel().tolist())
    # exact-zero pivots built from power-of-two pivots, so every operation
    # of the elimination is exact and INFO/pivots do not depend on rounding
    zero_col = np.asfortranarray(np.array([[2., 1., 0.], [4., 3., 0.], [8., 5., 0.]]))
    z = zero_col.copy(order="F")
    piv, info = O.dgetrf(z, use_ref=True)
    out["zero_col"] = dict(a=zero_col.T.ravel().tolist(), piv=piv.tolist(), info=int(info))
    rank_def = np.asfortranarray(np.array([[1., 2., 3.], [2., 4., 5.], [4., 8., 7.]]))
    z = rank_def.copy(order="F")
    piv, info = O.dgetrf(z, use_ref=True)
    out["rank_def"] = dict(a=rank_def.T.ravel().tolist(), piv=piv.tolist(), info=int(info))
    zero_row = np.asfortranarray(np.array([[0., 0., 0.], [1., 2., 4.], [2., 2., 1.]]))
    z = zero_row.copy(order="F")
    piv, info = O.dgetrf(z, use_ref=True)
    out["zero_row"] = dict(a=zero_row.T.ravel().tolist(), piv=piv.tolist(), info=int(info))
    tie = np.asfortranarray(np.array([[1., 2.], [-2., 1.]]))
    z = tie.copy(order="F")
    piv, info = O.dgetrf(z, use_ref=True)
    out["tie"] = dict(a=tie.T.ravel().tolist(), piv=piv.tolist(), info=int(info))
    # beta = 0 with NaN in C: the generic path computes 0*C
    a = O.uniform(3, 3, 1)
    b = O.uniform(3, 3, 2)
    c = np.asfortranarray(np.full((3, 3), np.nan))
    O.dgemm(0, 0, 1.0, a, b, 0.0, c, use_ref=True)
    out["beta0_nan"] = dict(all_nan=bool(np.isnan(c).all()))
    return out


def big_lu(n: int, nb: int, nrhs: int = 16, seed_a: int = 2, seed_b: int = 3):
    """The n=8192 LU + solve of the FP64 LU workload, by the reference, plus
    the oracle's per-column pivot gaps (which mark near-ties)."""
    t = time.time()
    a = O.uniform(n, n, seed_a)
    piv_r, info_r = O.dgetrf(a, nb=nb, use_ref=True)
    t_trf = time.time() - t
    b = O.uniform(n, nrhs, seed_b)
    t = time.time()
    O.dgetrs(0, a, piv_r, b, use_ref=True)
    t_trs = time.time() - t
    lu_sha = sha(a)
    x_sha = sha(b)
    x_rows = b[:8, :].copy()
    del a
    a2 = O.uniform(n, n, seed_a)
    piv_o, info_o, gap = O.dgetrf_gaps(a2, nb=nb)
    oracle_equal = bool(np.array_equal(piv_o, piv_r)) and sha(a2) == lu_sha
    np.savez_compressed(os.path.join(OUT, f"lu{n}_nb{nb}.npz"), piv=piv_r.astype(np.int32),
                        gap=gap.astype(np.float32), x_rows=x_rows)
    meta = dict(n=n, nb=nb, nrhs=nrhs, seed_a=seed_a, seed_b=seed_b, info=int(info_r),
                sha_lu=lu_sha, sha_x=x_sha, oracle_bit_equal=oracle_equal,
                ref_seconds_trf=t_trf, ref_seconds_trs=t_trs,
                flags="g++ -O2 -ffp-contract=off (oracle/Makefile)")
    with open(os.path.join(OUT, f"lu{n}_nb{nb}.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta))


def gemm2048():
    n = 2048
    a, b, c = O.uniform(n, n, 11), O.uniform(n, n, 12), O.uniform(n, n, 13)
    t = time.time()
    O.dgemm(0, 0, 1.5, a, b, -0.5, c, use_ref=True)
    dt = time.time() - t
    rng = np.random.default_rng(0)
    idx = rng.integers(0, n, size=(64, 2))
    samples = [[int(i), int(j), float(c[i, j])] for i, j in idx]
    meta = dict(m=n, n=n, k=n, alpha=1.5, beta=-0.5, seeds=[11, 12, 13], sha=sha(c),
                samples=samples, ref_seconds=dt)
    with open(os.path.join(OUT, "gemm2048.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("gemm2048", dt)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", type=int, default=0, help="nb for the n=8192 LU fixture")
    ap.add_argument("--gemm2048", action="store_true")
    args = ap.parse_args()
    if O.ref() is None:
        sys.exit("oracle/_ref/libflensref.so is missing (needs the reference sources)")
    if args.big:
        big_lu(8192, args.big)
        return
    if args.gemm2048:
        gemm2048()
        return
    data = dict(gemm=gemm_grid(), lu=lu_grid(), special=special_cases())
    with open(os.path.join(OUT, "small.json"), "w") as f:
        json.dump(data, f)
    print("wrote", len(data["gemm"]), "gemm cases,", len(data["lu"]), "lu cases")


if __name__ == "__main__":
    main()
