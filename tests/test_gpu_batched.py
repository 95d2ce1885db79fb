"""GPU checks of the batched solver (hip_LobpcgSolveBatched) through the C ABI.

* A batch of small 2D lognormal(0, 0.5) diffusion problems of different sizes
  and block sizes: every problem converges, its eigenvalues are the smallest
  of its matrix (dense numpy eigh) to 1e-9 relative, its eigenvectors are
  orthonormal (1e-11) with residuals (recomputed on the host) at the
  tolerance; the run replayed from hipGraphs equals the eager run bit for bit.
* The batched workload's generator at a reduced count (32 problems, sides 16
  to 48, k in {4, 8, 16}) against scipy's shift-invert eigsh per problem.
* Argument errors are reported, not run.
"""
import sys

import numpy as np
import pytest

from blopex_amd import lobpcg as bl
from blopex_amd import problems as pb

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _small_batch():
    sides = [6, 9, 12, 20, 7, 15, 11, 18, 8, 14, 10, 16]
    ks = [2, 4, 8, 8, 1, 4, 2, 8, 4, 2, 8, 4]
    return [(pb.laplacian_5pt(s, s, sigma=0.5, seed=3, stream=i), k)
            for i, (s, k) in enumerate(zip(sides, ks))]


def _check(problems, r, tol, lam_rtol):
    for i, (A, k) in enumerate(problems):
        S = A.to_scipy()
        lam, X = r.lam[i], r.X[i]
        assert r.flags[i] == 0, (i, r.flags[i])
        assert np.all(np.diff(lam) >= -1e-12)
        if A.n <= 500:
            ref = np.linalg.eigvalsh(S.toarray())[:k]
        else:
            import scipy.sparse.linalg as sla

            ref = np.sort(sla.eigsh(S, k=k, sigma=0.0, which="LM", return_eigenvectors=False))
        assert np.max(np.abs(lam - ref) / ref) < lam_rtol, (i, lam, ref)
        # X = [X R P] C with C orthonormal in the Gram metric of the basis:
        # orthonormal to the basis' conditioning times rounding
        assert np.max(np.abs(X.T @ X - np.eye(k))) < 1e-11
        res = np.linalg.norm(S @ X - X * lam, axis=0)
        assert np.all(res <= tol * 1.01), (i, res)
        np.testing.assert_allclose(res, r.res[i], rtol=1e-6, atol=1e-12)


def test_batched_small_problems_and_graph_replay():
    say("batched solve, 12 small problems (graph and eager)")
    problems = _small_batch()
    # a fixed chunk size: the autotuner's choice sets the summation order
    r = bl.lobpcg_batched(problems, tol=1e-8, maxit=500, want_X=True, autotune=False)
    say(f"  iterations {r.iterations.tolist()} stats {r.stats}")
    _check(problems, r, 1e-8, 1e-9)
    assert r.stats["graph_replays"] > 0
    e = bl.lobpcg_batched(problems, tol=1e-8, maxit=500, want_X=True, graph=False,
                          autotune=False)
    assert e.stats["graph_replays"] == 0
    np.testing.assert_array_equal(e.iterations, r.iterations)
    for a, b in zip(r.lam, e.lam):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(r.X, e.X):
        np.testing.assert_array_equal(a, b)


def test_batched_workload_generator_reduced():
    say("batched solve, 32 problems of the batched workload's generator (sides 16..48)")
    problems = pb.batched_problems(32, smin=16, smax=48)
    r = bl.lobpcg_batched(problems, tol=1e-6, maxit=1000, want_X=True)
    say(f"  max iterations {r.iterations.max()} stats {r.stats}")
    assert r.stats["chunk_rows"] in (512, 1024, 2048, 4096)
    assert r.stats["gram_rows"] == 32  # FP64 blocks: the VALU Gram, not tuned
    _check(problems, r, 1e-6, 1e-8)


def test_batched_mixed_tunes_the_gram_stage():
    """MIXED: the autotuner also picks the FP32 Gram's staged rows per step;
    every stage height gives the same eigenpairs to the tolerance."""
    say("batched solve, 24 problems, mixed, autotuned launch shapes")
    problems = pb.batched_problems(24, smin=16, smax=64)
    r = bl.lobpcg_batched(problems, tol=1e-6, maxit=1000, mode="mixed", want_X=True)
    say(f"  stats {r.stats}")
    assert r.stats["gram_rows"] in (16, 32, 64)
    _check(problems, r, 1e-6, 1e-7)


def test_batched_argument_errors():
    A = pb.laplacian_5pt(3, 3)
    with pytest.raises(RuntimeError, match="n >= 3k"):
        bl.lobpcg_batched([(A, 4)])
    with pytest.raises(RuntimeError, match="1 <= k"):
        bl.lobpcg_batched([(pb.laplacian_5pt(10, 10), 17)])


def test_batched_mixed_precision():
    """MIXED: the blocks in FP32 (FP64 Gram sums and Rayleigh-Ritz) to 1000 x tol,
    then FP64 from the FP32 blocks: the same eigenpairs as the FP64 solve
    (1e-9) and the residuals of the returned pairs at the tolerance."""
    say("batched solve, mixed precision (32 problems of the workload's generator)")
    problems = pb.batched_problems(32, smin=16, smax=48)
    r = bl.lobpcg_batched(problems, tol=1e-6, maxit=1000, want_X=True, mode="mixed")
    say(f"  max iterations {r.iterations.max()} stats {r.stats}")
    assert r.stats["iterations_fp32"] > 0
    _check(problems, r, 1e-6, 1e-8)
    f = bl.lobpcg_batched(problems, tol=1e-6, maxit=1000)
    for a, b in zip(r.lam, f.lam):
        assert np.max(np.abs(a - b) / b) < 1e-9


def test_batched_mixed_graph_matches_eager():
    """MIXED with graph replay (the convergence test and the end of the FP32
    phase, once fewer than 5% of the problems are left in it, decided on the
    device) against the eager path (both on the host): the same iterations,
    FP32 iterations and bits."""
    say("batched mixed: graph against eager (40 problems)")
    problems = pb.batched_problems(40, smin=16, smax=40)
    g = bl.lobpcg_batched(problems, tol=1e-6, maxit=1000, want_X=True, mode="mixed",
                          autotune=False)
    e = bl.lobpcg_batched(problems, tol=1e-6, maxit=1000, want_X=True, mode="mixed",
                          autotune=False, graph=False)
    say(f"  graph {g.stats['iterations']} it ({g.stats['iterations_fp32']} FP32, "
        f"{g.stats['graph_replays']} replays); eager {e.stats['iterations']} it")
    assert g.stats["graph_replays"] > 0 and e.stats["graph_replays"] == 0
    assert g.stats["iterations"] == e.stats["iterations"]
    assert g.stats["iterations_fp32"] == e.stats["iterations_fp32"]
    np.testing.assert_array_equal(g.iterations, e.iterations)
    np.testing.assert_array_equal(g.flags, e.flags)
    for a, b in zip(g.lam, e.lam):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(g.X, e.X):
        np.testing.assert_array_equal(a, b)
