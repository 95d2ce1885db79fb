"""Writes tests/golden/ref_fixtures.npz from the compiled reference.

Runs the reference's own lobpcg_solve_double (oracle/_ref/libblopex_ref.so,
built by oracle/Makefile from /root/reference) on small seeded cases and
stores inputs' seeds and outputs (eigenvalues, iteration counts, eigenvalue
and residual histories, final block) as plain arrays; and its
lobpcg_solve_complex (oracle/_ref/libblopex_ref_complex.so) on the complex
driver's own data files, stored with the outputs as the inputs A and X0.  Run from the repo root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from blopex_amd import problems as pb  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def cases(which="ref"):
    """The fixture cases; X0 from the reference's drand48 (which="ref") or the
    oracle's restatement of it (which="oracle": the same stream, pinned by
    the fixture's first 64 values), so the GPU tests need no compiled reference."""
    # 20^3 constant 7-point, k=10, no B/T, drand48 seed 1, tol 1e-6/1e-50, maxit 200
    A = pb.laplacian_7pt(20, 20, 20)
    yield "cfg1", dict(A=A, X0=po.drand48_block(A.n, 10, 1, which=which), maxit=200)
    # stock serial driver: diag(1..40), k=5, maxit 100 (serial_driver.c:44-175)
    A = pb.dense_to_csr(np.diag(np.arange(1.0, 41.0)))
    yield "diag40", dict(A=A, X0=po.drand48_block(40, 5, 1, which=which), maxit=100)
    # test_diag.m: diag(1..n), 3 constraints eye(n,3), T = inv(A), tol 1e-4, maxit 40
    n = 1000
    A = pb.dense_to_csr(np.diag(np.arange(1.0, n + 1)))
    yield "testdiag", dict(A=A, X0=po.drand48_block(n, 3, 1, which=which), Y=np.eye(n, 3),
                           tdiag=1.0 / np.arange(1.0, n + 1), atol=1e-4, maxit=40)
    # generalized: the 20^3 A, B = diag(uniform[0.5, 1.5]) (seed 7), k = 64, no T; capped at 120
    # iterations (the reference runs on and breaks down at iteration 158)
    A = pb.laplacian_7pt(20, 20, 20)
    yield "cfg3", dict(A=A, X0=po.drand48_block(A.n, 64, 1, which=which),
                       bdiag=pb.uniform(7, A.n, 0.5, 1.5), maxit=120)
    # generalized + Jacobi on a small variable-coefficient grid
    A = pb.laplacian_7pt(12, 12, 12, sigma=1.0, seed=3)
    yield "gen12", dict(A=A, X0=po.drand48_block(A.n, 8, 1, which=which),
                        bdiag=pb.uniform(5, A.n, 0.5, 1.5), tdiag=1.0 / A.diagonal(), maxit=300)
    # the 128^3 workload's generator at 20^3: lognormal(0, 1) coefficients
    # (seed 42), Jacobi, k = 32, tol 1e-6 (at 16^3 the reference breaks down
    # at iteration 54)
    A = pb.laplacian_7pt(20, 20, 20, sigma=1.0, seed=42)
    yield "cfg2s", dict(A=A, X0=po.drand48_block(A.n, 32, 1, which=which),
                        tdiag=1.0 / A.diagonal(), maxit=1000)


REF = "/root/reference"


def load_complex_driver_file(path: str) -> np.ndarray:
    """A complex matrix file of the reference's complex driver (first line
    "rows cols", then "real imag" per entry, column by column), with the
    entries rounded to single precision as serial_Multi_VectorLoad reads
    them (fscanf "%e" into float, blopex_serial_complex/multivector/
    multi_vector.c:668-700)."""
    with open(path) as f:
        n, m = map(int, f.readline().split())
        v = np.loadtxt(f)[: n * m]
    f32 = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    return (f32(v[:, 0]) + 1j * f32(v[:, 1])).reshape((m, n)).T.copy(order="F")


def complex_cases():
    # the complex driver's own inputs: A (Hermitian) and X0 from
    # blopex_serial_complex/driver/test{1,2}{a,x}.txt, tol 1e-6/1e-50, maxit 1000
    # (serial_driver.c:100-140)
    d = os.path.join(REF, "blopex_serial_complex", "driver")
    for t in ("test1", "test2"):
        yield "z" + t, load_complex_driver_file(os.path.join(d, t + "a.txt")), \
            load_complex_driver_file(os.path.join(d, t + "x.txt"))


def main():
    out = {}
    for name, A, X0 in complex_cases():
        r = po.ref_solve_complex(A, X0, maxit=1000)
        out[name + "_A"] = A
        out[name + "_X0"] = X0
        out[name + "_flag"] = np.array(r.flag)
        out[name + "_iters"] = np.array(r.iterations)
        out[name + "_lam"] = r.lam.real.copy()
        out[name + "_res"] = r.res
        out[name + "_lam_hist"] = r.lam_hist.real.copy()
        out[name + "_res_hist"] = r.res_hist
        print(name, r.flag, r.iterations, r.lam.real)
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
