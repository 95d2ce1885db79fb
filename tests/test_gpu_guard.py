"""GPU checks of the solver options that have no counterpart in lobpcg.c, and
of the 128^3 workload itself.

* Guard columns (settings.guard_columns): k + g columns iterated, the first
  k tested and returned.  The k smallest eigenvalues of small operators
  against numpy eigvalsh (1e-9 relative at tol 1e-8), B-orthonormal X.
* The eigenvalue error bound (settings.eig_rtol): the returned eigenvalues'
  errors against eigvalsh stay below the reported bound, which meets
  eig_rtol.
* The condition-estimate restart of lobpcg.m:846-863 (settings.cond_restart):
  on the Rayleigh-Ritz kernel a pencil whose Cholesky succeeds but whose P
  block is nearly dependent on X is solved again without P; in a solve to a
  tight tolerance the residuals recomputed on the host meet it.
* The size guard of lobpcg.m:480 (settings.size_guard).
* The look-ahead product: the next iteration's A R is enqueued behind the
  read of the residual norms and used only if the active set is unchanged.
  With the class timers on (settings.profile), which turn it off, a solve in
  which columns leave the active set one by one gives the same bits; so does
  a solve whose operator A reads device memory back to the host inside its
  apply (it runs while the residual-norm read is in flight).
* The 128^3 workload's generator at 20^3 (k = 32, Jacobi) against the
  reference's fixture, FP64 and MIXED; and the 128^3 workload itself with
  the bench's settings: residuals recomputed on the host from the returned
  X, X'X = I, MIXED eigenvalues against the FP64 ones.
"""
import ctypes
import sys

import numpy as np
import pytest
import scipy.linalg as sl

from blopex_amd import _lib as L
from blopex_amd import lobpcg as bl
from blopex_amd import problems as pb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("k,g", [(8, 4), (12, 8), (32, 8), (5, 1)])
def test_guard_columns_return_the_smallest_pairs(mode, k, g):
    say(f"guard columns, 12^3 lognormal, k={k}, g={g}, {mode}")
    A = pb.laplacian_7pt(12, 12, 12, sigma=1.0, seed=5)
    S = A.to_scipy().toarray()
    ref = np.linalg.eigvalsh(S)[:k]
    X0 = po.drand48_block(A.n, k, 1)
    r = bl.lobpcg(X0, A, T="jacobi", tol=1e-8, maxit=1500, mode=mode, guard_columns=g)
    say(f"  flag {r.flag} it {r.iterations} guard {r.stats['guard_columns']}")
    assert r.flag == 0 and r.stats["guard_columns"] == g
    assert r.X.shape == (A.n, k) and r.lam_hist.shape[0] == k
    # at residual 1e-8 the eigenvalue errors are ~||r||^2 / gap, below 1e-9
    assert relerr(r.lam, ref) < 1e-9
    np.testing.assert_allclose(r.X.T @ r.X, np.eye(k), atol=1e-9)
    res = np.linalg.norm(S @ r.X - r.X * r.lam, axis=0)
    assert np.all(res < 1.5e-8)


def test_guard_columns_generalized_with_b():
    say("guard columns, generalized 10^3 (diagonal B), k=6, g=4")
    A = pb.laplacian_7pt(10, 10, 10, sigma=0.5, seed=2)
    Bd = pb.uniform(4, A.n, 0.5, 1.5)
    ref = sl.eigh(A.to_scipy().toarray(), np.diag(Bd), eigvals_only=True)[:6]
    r = bl.lobpcg(po.drand48_block(A.n, 6, 1), A, B=Bd, T="jacobi", tol=1e-8, maxit=800,
                  guard_columns=4)
    assert r.flag == 0
    assert relerr(r.lam, ref) < 1e-10
    np.testing.assert_allclose(r.X.T @ (Bd[:, None] * r.X), np.eye(6), atol=1e-9)


def test_guard_columns_too_wide_is_a_flag():
    say("k + g above 64")
    A = pb.laplacian_7pt(10, 10, 10)
    r = bl.lobpcg(None, A, k=60, maxit=5, guard_columns=5)
    assert r.flag == 2 and "guard" in L.last_error()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_eigenvalue_error_bound_holds(mode):
    """At residual tolerance 1e-5 the eigenvalues alone would be off by
    ~||r||^2 / gap; with eig_rtol 1e-12 the solve goes on until the bound
    meets it, and the true errors are below the bound."""
    say(f"eigenvalue error bound, 14^3 lognormal, k=10, g=4, {mode}")
    A = pb.laplacian_7pt(14, 14, 14, sigma=1.0, seed=8)
    S = A.to_scipy().toarray()
    ref = np.linalg.eigvalsh(S)[:10]
    X0 = po.drand48_block(A.n, 10, 1)
    loose = bl.lobpcg(X0, A, T="jacobi", tol=1e-5, maxit=2000, mode=mode, guard_columns=4)
    r = bl.lobpcg(X0, A, T="jacobi", tol=1e-5, maxit=2000, mode=mode, guard_columns=4,
                  eig_rtol=1e-12)
    bound = r.stats["eig_error_bound"]
    err = np.max(np.abs(r.lam - ref) / ref)
    say(f"  loose: it {loose.iterations} err {np.max(np.abs(loose.lam - ref) / ref):.2e}; "
        f"bounded: it {r.iterations} err {err:.2e} bound {bound:.2e}")
    assert r.flag == 0 and loose.flag == 0
    assert r.iterations > loose.iterations
    assert 0.0 < bound <= 1e-12
    assert err <= bound * 1.01 + 1e-15


@pytest.mark.parametrize("mode,case", [("fp64", "plain"), ("mixed", "plain"), ("fp64", "B"),
                                       ("mixed", "B"), ("fp64", "refine")])
def test_lookahead_product_leaves_the_results_unchanged(mode, case):
    say(f"look-ahead product, 16^3 lognormal, k=12, {mode}, {case}")
    A = pb.laplacian_7pt(16, 16, 16, sigma=1.0, seed=3)
    X0 = po.drand48_block(A.n, 12, 1)
    extra = {}
    if case == "B":  # B x = lambda ... with a diagonal B: A R and B R both enqueued ahead
        extra["B"] = np.random.default_rng(4).uniform(0.5, 1.5, A.n)
    if case == "refine":  # every column active again after convergence
        extra["refine_iterations"] = 6
    a = bl.lobpcg(X0, A, T="jacobi", tol=1e-7, maxit=800, mode=mode, **extra)
    b = bl.lobpcg(X0, A, T="jacobi", tol=1e-7, maxit=800, mode=mode, profile=True, **extra)
    # the iteration at which each column first met the tolerance: the
    # active set shrank at several iterations, each one discarding the
    # product enqueued ahead
    first = [int(np.argmax(a.res_hist[i, 1:] < 1e-7)) for i in range(12)]
    say(f"  flag {a.flag}/{b.flag} it {a.iterations}/{b.iterations} "
        f"active-set changes at {sorted(set(first))}")
    assert a.flag == 0 and b.flag == 0
    assert len(set(first)) >= 3
    assert a.iterations == b.iterations
    np.testing.assert_array_equal(a.lam, b.lam)
    np.testing.assert_array_equal(a.res, b.res)
    np.testing.assert_array_equal(a.X, b.X)
    np.testing.assert_array_equal(a.res_hist, b.res_hist)


def test_lookahead_with_an_operator_that_reads_back():
    """A caller's operator that copies device results to the host inside its
    apply (here the MFMA layout probe after the CSR product) runs between the
    start and the end of the look-ahead read of the residual norms: the solve
    gives the same bits as with the library's CSR operator."""
    say("look-ahead, operator with a device-to-host read, 16^3 lognormal, k=12, fp64")
    lib = L.load()
    A = pb.laplacian_7pt(16, 16, 16, sigma=1.0, seed=3)
    X0 = po.drand48_block(A.n, 12, 1)
    fn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
    csr = fn(L.symbol_address("HipCSRMatMultiVec"))
    rows, cols = np.zeros(256, np.int32), np.zeros(256, np.int32)
    calls = [0]

    def apply_and_read(a, x, y):
        csr(a, x, y)
        lib.blopex_amd_kernel_mfma_probe(L.HIPMV_FP64, rows.ctypes.data, cols.ctypes.data)
        calls[0] += 1

    cb = fn(apply_and_read)
    out = []
    for own in (False, True):
        dA = bl.DeviceCSR(A)
        dT = bl.DeviceDiag(jacobi_of=dA)
        X = bl.DeviceBlock.from_numpy(X0)
        try:
            if own:
                dA.op = ctypes.cast(cb, ctypes.c_void_p).value
            r = bl.solve_device(dA, X, None, dT, tol=1e-7, maxit=800)
            r.X = X.to_numpy()
            out.append(r)
        finally:
            for o in (X, dT, dA):
                o.close()
    a, b = out
    say(f"  flag {a.flag}/{b.flag} it {a.iterations}/{b.iterations} operator calls {calls[0]}")
    assert a.flag == 0 and b.flag == 0
    assert calls[0] >= b.iterations
    assert a.iterations == b.iterations
    np.testing.assert_array_equal(a.lam, b.lam)
    np.testing.assert_array_equal(a.res, b.res)
    np.testing.assert_array_equal(a.X, b.X)
    np.testing.assert_array_equal(a.res_hist, b.res_hist)


def test_condition_restart_on_the_rayleigh_ritz_kernel():
    """A P column nearly along an X column: B_n's Cholesky succeeds (a pivot
    ~1e-10) but its condition estimate is above the limit 8: the step is
    solved without P (p_info = -2) and equals the [X R] solution, and the
    estimate recorded is that of the pencil solved (as lobpcg.m records it);
    with a limit of 20 the full pencil is solved and its estimate is above 8."""
    say("condition-estimate restart, k=8 a=8 p=8")
    from tests.test_gpu_rr import make_case, reference

    A, B, X, R, P, lam = make_case(8, 8, 8, seed=31, bdiag=True)
    P[:, 2] = X[:, 2] + 1e-6 * np.random.default_rng(3).standard_normal(A.shape[0])
    k, a, p = 8, 8, 8
    Raw = np.hstack([X, R, P])
    n = k + a + p
    lda = n + 3
    ga = np.zeros((lda, n), order="F")
    gb = np.zeros((lda, n), order="F")
    ga[:n] = Raw.T @ A @ Raw
    gb[:n] = Raw.T @ B @ Raw
    lib = L.load()

    def run(limit):
        lo = np.zeros(k)
        coef = np.zeros((n, k), order="F")
        flags = np.zeros(4, dtype=np.int32)
        ce = np.zeros(1)
        lib.blopex_amd_kernel_rr_cond(k, a, p, ga.ctypes.data, gb.ctypes.data, lda,
                                      np.ascontiguousarray(lam).ctypes.data, 0, limit,
                                      lo.ctypes.data, coef.ctypes.data, flags.ctypes.data,
                                      ce.ctypes.data)
        return lo, coef, flags, ce[0]

    lo2, coef2, flags2, ce2 = run(20.0)
    say(f"  limit 20: condest {ce2:.2f}, flags {flags2}")
    assert flags2[0] == 0 and flags2[2] == 0 and 8.0 < ce2 < 20.0
    assert np.any(coef2[k + a:] != 0.0)
    lo, coef, flags, ce = run(8.0)
    say(f"  limit 8: condest {ce:.2f}, flags {flags}")
    assert flags[0] == 0 and flags[2] == -2 and ce < 8.0
    assert np.all(coef[k + a:] == 0.0)
    ev, _ = reference(A, B, X, R, P[:, :0], lam)
    assert np.max(np.abs(lo - ev) / np.abs(ev)) < 1e-12


def test_condition_restart_in_a_tight_solve():
    say("cond_restart solve, 10^3 lognormal, k=6, tol 1e-11")
    A = pb.laplacian_7pt(10, 10, 10, sigma=1.0, seed=4)
    S = A.to_scipy().toarray()
    r = bl.lobpcg(po.drand48_block(A.n, 6, 1), A, T="jacobi", tol=1e-11, maxit=1500,
                  cond_restart=True, final_rr=True)
    say(f"  flag {r.flag} it {r.iterations} restarts {r.stats['cond_restarts']}")
    assert r.flag == 0
    res = np.linalg.norm(S @ r.X - r.X * r.lam, axis=0)
    assert np.all(res < 2e-11)
    assert relerr(r.lam, np.linalg.eigvalsh(S)[:6]) < 1e-12


def test_size_guard():
    say("size guard: n = 40, k = 10")
    A = pb.dense_to_csr(np.diag(np.arange(1.0, 41.0)))
    r = bl.lobpcg(po.drand48_block(40, 10, 1), A, maxit=50, size_guard=True)
    assert r.flag == 1 and r.iterations == 0
    r = bl.lobpcg(po.drand48_block(40, 5, 1), A, maxit=100, size_guard=True)
    assert r.flag == 0


# ------------------------------------------------- the 128^3 workload family
def _cfg2s():
    from tests.golden.make_golden import cases

    return dict(cases(which="oracle"))["cfg2s"]


def test_cfg2_generator_fp64_matches_reference(golden):
    """20^3 lognormal(0, 1), Jacobi, k = 32, tol 1e-6: the reference's exit
    flag, eigenvalues to 1e-8, iterations within max(2, 5%), the same
    initial Rayleigh-Ritz."""
    say("20^3 lognormal, k=32, FP64 against the reference fixture")
    kw = _cfg2s()
    r = bl.lobpcg(kw["X0"], kw["A"], T=kw["tdiag"], tol=1e-6, maxit=kw["maxit"])
    ref_it = int(golden["cfg2s_iters"])
    say(f"  flag {r.flag} it {r.iterations} (reference {ref_it})")
    assert r.flag == int(golden["cfg2s_flag"]) == 0
    assert abs(r.iterations - ref_it) <= max(2, ref_it // 20)
    assert relerr(r.lam, golden["cfg2s_lam"]) < 1e-8
    np.testing.assert_allclose(r.lam_hist[:, 0], golden["cfg2s_lam_hist"][:, 0], rtol=1e-12)


@pytest.mark.parametrize("guard", [0, 8])
def test_cfg2_generator_mixed_matches_reference(golden, guard):
    say(f"20^3 lognormal, k=32, MIXED (guard {guard}) against the reference fixture")
    kw = _cfg2s()
    r = bl.lobpcg(kw["X0"], kw["A"], T=kw["tdiag"], tol=1e-6, maxit=kw["maxit"], mode="mixed",
                  guard_columns=guard)
    say(f"  flag {r.flag} it {r.iterations} fp32 {r.stats['iterations_fp32']}")
    assert r.flag == 0
    assert relerr(r.lam, golden["cfg2s_lam"]) < 1e-8
    S = kw["A"].to_scipy()
    res = np.linalg.norm(S @ r.X - r.X * r.lam, axis=0)
    assert np.all(res < 1.01e-6)


def test_128cube_workload_properties():
    """The 128^3 workload (n = 2,097,152, k = 32, Jacobi) with the bench's
    settings (8 guard columns, eigenvalue bound 1e-10, final Rayleigh-Ritz
    step): the residuals of the
    returned X recomputed on the host (scipy SpMM) meet the tolerance, X is
    orthonormal to 1e-12, the bound holds, and the MIXED eigenvalues agree
    with the FP64 ones to 1e-9."""
    say("128^3 workload, FP64 and MIXED, guard 8, eig_rtol 1e-10")
    A = pb.laplacian_7pt(128, 128, 128, sigma=1.0, seed=42)
    S = A.to_scipy()
    out = {}
    for mode in ("fp64", "mixed"):
        r = bl.lobpcg(None, A, T="jacobi", k=32, seed=1, tol=1e-6, maxit=4000, mode=mode,
                      guard_columns=8, eig_rtol=1e-10, final_rr=True)
        say(f"  {mode}: flag {r.flag} it {r.iterations} fp32 {r.stats['iterations_fp32']} "
            f"bound {r.stats['eig_error_bound']:.2e} {r.stats['seconds']:.2f}s")
        assert r.flag == 0
        assert r.stats["eig_error_bound"] <= 1e-10
        res = np.linalg.norm(S @ r.X - r.X * r.lam, axis=0)
        assert np.all(res < 1.01e-6), res.max()
        G = r.X.T @ r.X
        assert np.max(np.abs(G - np.eye(32))) < 1e-12
        out[mode] = r.lam
    say(f"  max rel diff {relerr(out['mixed'], out['fp64']):.2e}")
    assert relerr(out["mixed"], out["fp64"]) < 1e-9
