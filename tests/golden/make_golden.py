"""Writes tests/golden/ref_fixtures.npz from the compiled reference.

Runs the reference's own lobpcg_solve_double (oracle/_ref/libblopex_ref.so,
built by oracle/Makefile from /root/reference) on small seeded cases and
stores inputs' seeds and outputs (eigenvalues, iteration counts, eigenvalue
and residual histories, final block) as plain arrays.  Run from the repo root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from blopex_amd import problems as pb  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def cases():
    # config 1: 20^3 constant 7-point, k=10, no B/T, drand48 seed 1, tol 1e-6/1e-50, maxit 200
    A = pb.laplacian_7pt(20, 20, 20)
    yield "cfg1", dict(A=A, X0=po.drand48_block(A.n, 10, 1, which="ref"), maxit=200)
    # stock serial driver: diag(1..40), k=5, maxit 100 (serial_driver.c:44-175)
    A = pb.dense_to_csr(np.diag(np.arange(1.0, 41.0)))
    yield "diag40", dict(A=A, X0=po.drand48_block(40, 5, 1, which="ref"), maxit=100)
    # test_diag.m: diag(1..n), 3 constraints eye(n,3), T = inv(A), tol 1e-4, maxit 40
    n = 1000
    A = pb.dense_to_csr(np.diag(np.arange(1.0, n + 1)))
    yield "testdiag", dict(A=A, X0=po.drand48_block(n, 3, 1, which="ref"), Y=np.eye(n, 3),
                           tdiag=1.0 / np.arange(1.0, n + 1), atol=1e-4, maxit=40)
    # config 3: config 1's A, B = diag(uniform[0.5, 1.5]) (seed 7), k = 64, no T; capped at 120
    # iterations (the reference runs on and breaks down at iteration 158)
    A = pb.laplacian_7pt(20, 20, 20)
    yield "cfg3", dict(A=A, X0=po.drand48_block(A.n, 64, 1, which="ref"),
                       bdiag=pb.uniform(7, A.n, 0.5, 1.5), maxit=120)
    # generalized + Jacobi on a small variable-coefficient grid
    A = pb.laplacian_7pt(12, 12, 12, sigma=1.0, seed=3)
    yield "gen12", dict(A=A, X0=po.drand48_block(A.n, 8, 1, which="ref"),
                        bdiag=pb.uniform(5, A.n, 0.5, 1.5), tdiag=1.0 / A.diagonal(), maxit=300)


def main():
    out = {}
    for name, kw in cases():
        r = po.solve(which="ref", **kw)
        out[name + "_flag"] = np.array(r.flag)
        out[name + "_iters"] = np.array(r.iterations)
        out[name + "_lam"] = r.lam
        out[name + "_res"] = r.res
        out[name + "_lam_hist"] = r.lam_hist
        out[name + "_res_hist"] = r.res_hist
        if r.X.size <= 20000:
            out[name + "_X"] = r.X
        print(name, r.flag, r.iterations, r.lam)
    out["drand48_seed1_first64"] = po.drand48_block(64, 1, 1, which="ref")[:, 0]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_fixtures.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
