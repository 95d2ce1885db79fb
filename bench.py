"""Benchmark of the device LOBPCG on one MI355X.

Workload (the 128^3 configuration): 3D 7-point variable-coefficient
diffusion operator on a 128^3 grid (n = 2,097,152, nnz = 14,581,760),
per-cell coefficients lognormal(0, 1) (seed 42, counter-based generator of
blopex_amd.problems), harmonic face averages, Dirichlet boundary; Jacobi
preconditioner; k = 32 smallest eigenpairs; X0 = 2*drand48()-1 after
srand48(1), generated on the device; absolute residual tolerance 1e-6
(relative 1e-50).

A step is one full solve in MIXED mode (FP32 blocks, FP64 Gram
accumulation / small dense / finish), starting from X0 already in HBM,
until every residual is below 1e-6 AND every eigenvalue's error bound
(settings.eig_rtol) is below 1e-9 relative, then the final Rayleigh-Ritz
step.  Steps cycle through six residual-replacement intervals, which perturb
the FP32 trajectory: `value` is the mean time over them (s, lower is
better).  After the steps: the FP64 oracle (FP64, error bound 1e-12), the
same solve in FP64 (speed-up), the eigenvalue errors of every run against
the oracle, the two modes stopped by the residual tolerance alone, and the
solves with 8 guard columns.  The roofline object is the SpMM kernel the
library picks for this matrix (the z-marching stencil kernel k_zm here; the
diagonal-copy kernel k_dia or the CSR kernel for other matrices; HIP events
around back-to-back launches on the library stream, k = 32 columns, FP32),
priced at the CSR algorithmic bytes, at the bytes of the format it reads
and at its measured traffic, with the rocprofv3 trace average beside the
event time.  Its `traffic` (HBM bytes per launch) is read from the
committed rocprofv3 counter summary of `bench.py --roofline-only`
(FETCH_SIZE and WRITE_SIZE passes scaled by a 512 MiB copy).  cpu_baseline
is the reference's own lobpcg_solve_double (oracle/_ref) on this host and
this operator, one thread, four iterations (three with the P block).

    python bench.py [--steps K] [--warmup W] [--n 128] [--k 32] [--guard G] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

COUNTERS = os.path.join(ROOT, "profiles", "spmm_roofline_128cube_k32_counters.txt")
TRACE = os.path.join(ROOT, "profiles", "spmm_roofline_128cube_k32_trace.txt")
# the general CSR SpMM kernel as rocprofv3 names it (FP32, up to 8 entries
# a row in registers)
CSR_KERNEL = "k_spmm<float, 8>"
EIG_RTOL = 1e-9  # eigenvalue error bound of the headline solve, relative
MFMA_SUMMARY = os.path.join(ROOT, "profiles", "lobpcg_mixed_128cube_k32_mfma.txt")
GRAM_KERNEL = "k_gram<float, 6, 3, 64>"  # the Gram plan of the FP32 iterations (up to 192 columns)
METRIC = ("time-to-residual 1e-6, k=32, 128³ Laplacian: mixed vs FP64 on GPU; SpMM GB/s")
HBM_PEAK_GBS = 8000.0


class HipEvents:
    """hipEvent timing on an explicit stream (ctypes on libamdhip64)."""

    def __init__(self):
        self.hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                                 ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value


def spmm_roofline(lib, dA, A, k, prec, reps=100, force_csr=False):
    """Average duration of the SpMM kernel (HipCSRMatMultiVec) over `reps` back-to-back launches
    after ten untimed ones;
    algorithmic bytes of the format the kernel reads (the z-marching and diagonal kernels read
    the diagonal copy, the CSR kernel values, columns and row pointers)."""
    from blopex_amd import _lib as L
    from blopex_amd.lobpcg import DeviceBlock

    X = DeviceBlock(A.n, k, prec)
    X.set_random(3)
    Y = DeviceBlock(A.n, k, prec)
    c = dA.ptr.contents
    nd = c.ndiag
    if force_csr:  # the general CSR kernel on the same matrix (no diagonal copy)
        c.ndiag = 0
    lib.HipCSRMatMultiVec.argtypes = [ctypes.c_void_p] * 3
    call = lambda: lib.HipCSRMatMultiVec(ctypes.cast(dA.ptr, ctypes.c_void_p),
                                         ctypes.cast(X.ptr, ctypes.c_void_p),
                                         ctypes.cast(Y.ptr, ctypes.c_void_p))
    for _ in range(10):
        call()
    lib.blopex_amd_synchronize()
    ev = HipEvents()
    s = lib.blopex_amd_stream()
    a, b = ev.event(), ev.event()
    ev.record(a, s)
    for _ in range(reps):
        call()
    ev.record(b, s)
    ms = ev.elapsed_ms(a, b) / reps
    c.ndiag = nd
    es = 4 if prec == L.HIPMV_FP32 else 8
    if nd and not force_csr:  # diagonal copy: nd value streams, no index stream
        algo = nd * dA.ptr.contents.diag_ld * es + 2 * A.n * k * es
    else:  # CSR: values, column indices, row pointers
        algo = A.nnz * (es + 4) + (A.n + 1) * 4 + 2 * A.n * k * es
    X.close()
    Y.close()
    return ms, algo


def cpu_baseline(A, k, seed_x=1, iterations=4):
    """The reference's own lobpcg_solve_double (oracle/_ref) on this workload
    (the same CSR operator, Jacobi T, k, X0), one thread, on a bounded sample:
    four iterations.  The harness stamps CLOCK_MONOTONIC at every operator-A
    call (the initial AX, then one per iteration).  Iteration 1 has no P
    block (its Rayleigh-Ritz basis is [X R], 2k columns); from iteration 2 on
    the basis is [X R P] with every column active (none converges that early),
    so the intervals between the calls of iterations 2..4 are full-width
    iterations, and their mean is the reported s/iteration."""
    from threadpoolctl import threadpool_limits

    from oracle import pyoracle as po

    if not po.ref_available():
        return None
    X0 = po.drand48_block(A.n, k, seed_x)
    t0 = time.perf_counter()
    with threadpool_limits(limits=1):
        po.solve(A=A, X0=X0, tdiag=1.0 / A.diagonal(), maxit=iterations, atol=1e-6, which="ref")
    wall = time.perf_counter() - t0
    st = po.ref_operator_stamps()
    if len(st) < 3:
        return None
    per = [float(st[i + 1] - st[i]) for i in range(1, len(st) - 1)]
    if len(per) < 2:
        return None
    return {"s_per_iteration": float(np.mean(per[1:])), "s_per_iteration_each": per[1:],
            "s_first_iteration_without_p": per[0],
            "setup_s": float(st[1] - st[0]), "sample_wall_s": wall,
            "sample": f"reference lobpcg_solve_double on the same 128^3 operator, Jacobi, k={k}, "
                      f"X0 = drand48 seed {seed_x}, maxit {iterations}, 1 thread ({wall:.1f} s "
                      f"of CPU work); s/iteration = mean time between the operator-A calls of "
                      f"iterations 2..{iterations} (basis [X R P], all columns active; "
                      f"iteration 1, without P, reported apart)"}


def run_batched(count, graph=True, maxit=2000, mode="mixed", compare_fp64=True):
    """The batched workload: count independent 2D 5-point lognormal(0, 0.5)
    diffusion problems (sides 32..256, k in {4, 8, 16}, Jacobi, tolerance
    1e-6) in one hip_LobpcgSolveBatched call: FP32 blocks with FP64 Gram sums
    and Rayleigh-Ritz, then FP64 (mode "mixed"), and the FP64 solve beside it."""
    from blopex_amd import problems as pb
    from blopex_amd.lobpcg import lobpcg_batched

    problems = pb.batched_problems(count)
    t = time.perf_counter()
    r = lobpcg_batched(problems, tol=1e-6, maxit=maxit, graph=graph, mode=mode)
    wall = time.perf_counter() - t
    conv = int(np.sum(r.flags == 0))
    pairs = sum(k for _, k in problems)
    out = {"workload": f"{count} 2D 5-point lognormal(0,0.5) problems, sides 32..256, "
                       f"k in {{4,8,16}}, Jacobi, {mode}, tol 1e-6 (batched)",
           "problems": count, "rows": int(r.stats["total_rows"]), "eigenpairs": pairs,
           "converged": conv, "maxit": maxit, "seconds": r.stats["seconds"],
           "wall_seconds": wall, "iterations_max": int(r.iterations.max()),
           "iterations_mean": float(r.iterations.mean()),
           "iterations_fp32": r.stats["iterations_fp32"],
           "graph_replays": r.stats["graph_replays"], "chunk_rows": r.stats["chunk_rows"],
           "gram_rows": r.stats["gram_rows"]}
    if compare_fp64 and mode != "fp64":
        f = lobpcg_batched(problems, tol=1e-6, maxit=maxit, graph=graph)
        out["fp64"] = {"seconds": f.stats["seconds"], "converged": int(np.sum(f.flags == 0)),
                       "iterations_mean": float(f.iterations.mean())}
        both = [i for i in range(count) if r.flags[i] == 0 and f.flags[i] == 0]
        out["max_rel_eig_diff_vs_fp64"] = float(max(
            np.max(np.abs(r.lam[i] - f.lam[i]) / np.abs(f.lam[i])) for i in both))
    return out


def run_config3():
    """The generalized workload: the 20^3 constant 7-point Laplacian (n =
    8000), B = diag(uniform[0.5, 1.5]) (seed 7), k = 64 (Rayleigh-Ritz
    pencils up to 192 x 192), no preconditioner, X0 = drand48 seed 1, tol
    1e-6, with the condition-estimate restart and the final Rayleigh-Ritz
    step; FP64 and MIXED, eigenvalues against scipy shift-invert."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    from blopex_amd import problems as pb
    from blopex_amd.lobpcg import lobpcg
    from oracle import pyoracle as po

    A = pb.laplacian_7pt(20, 20, 20)
    Bd = pb.uniform(7, A.n, 0.5, 1.5)
    X0 = po.drand48_block(A.n, 64, 1)
    true = np.sort(sla.eigsh(A.to_scipy(), k=70, M=sp.diags(Bd), sigma=0.0, which="LM",
                             return_eigenvectors=False))[:64]
    out = {"workload": "20^3 7-point Laplacian, B = diag(uniform[0.5,1.5]), k=64, no T, tol 1e-6, "
                       "condition-estimate restart, final Rayleigh-Ritz"}
    lam = {}
    for mode in ("fp64", "mixed"):
        lobpcg(X0, A, B=Bd, tol=1e-6, maxit=1000, mode=mode, final_rr=True, cond_restart=True)
        t = time.perf_counter()
        r = lobpcg(X0, A, B=Bd, tol=1e-6, maxit=1000, mode=mode, final_rr=True, cond_restart=True)
        dt = time.perf_counter() - t
        lam[mode] = np.sort(r.lam)
        out[mode] = {"seconds": dt, "solver_seconds": r.stats["seconds"], "flag": r.flag,
                     "iterations": r.iterations, "iterations_fp32": r.stats["iterations_fp32"],
                     "converged": int(np.sum(r.res < 1e-6)), "restarts": r.stats["cond_restarts"],
                     "max_rel_err_vs_shift_invert": float(np.max(np.abs(lam[mode] - true) / true))}
    out["speedup_mixed_vs_fp64"] = out["fp64"]["solver_seconds"] / out["mixed"]["solver_seconds"]
    out["max_rel_diff_mixed_vs_fp64"] = float(np.max(np.abs(lam["mixed"] - lam["fp64"]) /
                                                     lam["fp64"]))
    return out


def run_adjoint(N=64, mode="mixed"):
    """The differentiation workload: the 7-point lognormal(0, 1) operator on an
    N^3 grid, loss = sum of the 8 smallest eigenvalues + ||x_1||^2 on the
    central (N/4)^3 box; forward solve (11 columns, tol 1e-10, final
    Rayleigh-Ritz) and hip_LobpcgAdjoint (gradients of the CSR values)."""
    from blopex_amd import problems as pb
    from blopex_amd.lobpcg import adjoint, lobpcg

    c = pb.lognormal_coefficients(N ** 3, 1.0, 64)
    A = pb.diffusion_stencil((N, N, N), c)
    t = time.perf_counter()
    r = lobpcg(None, A, T="jacobi", k=11, tol=1e-10, maxit=4000, mode=mode, final_rr=True)
    t_fwd = time.perf_counter() - t
    lo, hi = 3 * N // 8, 5 * N // 8
    box = np.zeros((N, N, N), dtype=bool)
    box[lo:hi, lo:hi, lo:hi] = True
    box = box.reshape(-1)
    glam = np.r_[np.ones(8), np.zeros(3)]
    GX = np.zeros_like(r.X)
    GX[:, 0] = 2.0 * r.X[:, 0] * box
    t = time.perf_counter()
    ad = adjoint(A, r.X, r.lam, glam, GX, T="jacobi", mode=mode, tol=1e-10, maxit=5000)
    t_adj = time.perf_counter() - t
    return {"workload": f"{N}^3 7-point lognormal(0,1), loss sum(lambda_1..8) + "
                        f"||x_1 on the central {hi - lo}^3 box||^2, {mode}",
            "forward_flag": int(r.flag), "forward_iterations": int(r.iterations),
            "forward_seconds": t_fwd, "adjoint_flag": int(ad.flag),
            "adjoint_cg_iterations": int(ad.stats["iterations"]),
            "adjoint_corrections": int(ad.stats["corrections"]),
            "adjoint_seconds": ad.stats["seconds"], "adjoint_wall_seconds": t_adj,
            "gradient_entries": int(A.nnz)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--guard", type=int, default=0, help="guard columns of the headline solve")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="grid side")
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--maxit", type=int, default=6000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the SpMM roofline launches and a device copy of known size "
                         "(the command the committed rocprofv3 summaries and counters use)")
    ap.add_argument("--profile-only", action="store_true",
                    help="one mixed solve and the SpMM launches only (for rocprofv3)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched workload (512 independent 2D problems)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched workload (for rocprofv3)")
    ap.add_argument("--batched-count", type=int, default=512)
    ap.add_argument("--batched-no-graph", action="store_true",
                    help="batched workload with eager launches instead of graph replays")
    ap.add_argument("--batched-mode", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-config3", action="store_true",
                    help="skip the generalized k=64 workload (20^3, diagonal B)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiation workload (64^3 forward solve + adjoint)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run the differentiation workload alone and print its object")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from blopex_amd import _lib as L
    from blopex_amd import problems as pb
    from blopex_amd.lobpcg import DeviceBlock, DeviceCSR, DeviceDiag, solve_device

    lib = L.load()
    lib.blopex_amd_init()
    if args.adjoint_only:
        print(json.dumps(run_adjoint()), flush=True)
        return
    if args.batched_only:
        print(json.dumps(run_batched(args.batched_count, not args.batched_no_graph,
                                     mode=args.batched_mode, compare_fp64=False)), flush=True)
        return
    nside, k = args.n, args.k
    t = time.perf_counter()
    A = pb.laplacian_7pt(nside, nside, nside, sigma=1.0, seed=42)
    print(f"# matrix {nside}^3 n={A.n} nnz={A.nnz} built in {time.perf_counter() - t:.1f}s",
          file=sys.stderr, flush=True)
    dA = DeviceCSR(A)
    dT = DeviceDiag(jacobi_of=dA)
    X0 = DeviceBlock(A.n, k)
    X0.set_random(1)
    X = DeviceBlock(A.n, k)
    nbytes = A.n * k * 8
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                   ctypes.c_void_p]
    xdata = lambda b: L.HipMultiVector.from_address(ctypes.addressof(b.ptr.contents)).data

    def solve(mode, profile=False, **settings):
        hip.hipMemcpyAsync(xdata(X), xdata(X0), nbytes, 3, lib.blopex_amd_stream())
        return solve_device(dA, X, T=dT, tol=1e-6, maxit=args.maxit, mode=mode, profile=profile,
                            histories=False, **settings)

    def log(msg):
        print(f"# {msg}", file=sys.stderr, flush=True)

    if args.roofline_only:
        ms32, algo32 = spmm_roofline(lib, dA, A, k, L.HIPMV_FP32)
        ms64, algo64 = spmm_roofline(lib, dA, A, k, L.HIPMV_FP64)
        msc, algoc = spmm_roofline(lib, dA, A, k, L.HIPMV_FP32, force_csr=True)
        log(f"spmm fp32 {ms32 * 1e3:.1f} us ({algo32 / ms32 / 1e6:.0f} GB/s)  "
            f"fp64 {ms64 * 1e3:.1f} us ({algo64 / ms64 / 1e6:.0f} GB/s)  "
            f"general CSR fp32 {msc * 1e3:.1f} us ({algoc / msc / 1e6:.0f} GB/s)")
        # reference for the counters: one device-to-device copy of 512 MiB
        nb = 512 << 20
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nb)) == 0
        assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nb)) == 0
        for _ in range(3):
            hip.hipMemcpyAsync(dst, src, nb, 3, lib.blopex_amd_stream())
        lib.blopex_amd_synchronize()
        hip.hipFree(src)
        hip.hipFree(dst)
        log("copy 512 MiB x3 done")
        return

    if args.profile_only:
        r = solve("mixed", eig_rtol=EIG_RTOL, final_rr=True, guard_columns=args.guard)
        log(f"profile-only mixed solve: flag {r.flag} it {r.iterations} {r.stats['seconds']:.3f}s")
        ms32, _ = spmm_roofline(lib, dA, A, k, L.HIPMV_FP32, reps=10)
        ms64, _ = spmm_roofline(lib, dA, A, k, L.HIPMV_FP64, reps=10)
        log(f"spmm fp32 {ms32 * 1e3:.1f} us  fp64 {ms64 * 1e3:.1f} us")
        return

    # The headline solve: tolerance 1e-6 on every residual AND the
    # eigenvalue error bound (settings.eig_rtol) at most 1e-9 |lambda_i|,
    # then the final Rayleigh-Ritz step (X B-orthonormal, residuals of the
    # returned X).  Steps cycle through residual-replacement intervals, which
    # perturb the FP32 trajectory, so `value` is the mean over trajectories.
    head = dict(eig_rtol=EIG_RTOL, final_rr=True, guard_columns=args.guard)
    intervals = [20, 15, 25, 30, 35, 40]
    for w in range(args.warmup):
        r = solve("mixed", **head)
        log(f"warmup {w}: flag {r.flag} iterations {r.iterations} "
            f"(fp32 {r.stats['iterations_fp32']}, fp64 {r.stats['iterations_fp64']}) "
            f"{r.stats['seconds']:.3f}s")
    lib.blopex_amd_synchronize()
    times, results = [], []
    t_all = time.perf_counter()
    for s in range(args.steps):
        t0 = time.perf_counter()
        r = solve("mixed", refresh_interval=intervals[s % len(intervals)], **head)
        lib.blopex_amd_synchronize()
        times.append(time.perf_counter() - t0)
        results.append(r)
        log(f"step {s}: refresh {intervals[s % len(intervals)]} flag {r.flag} iterations "
            f"{r.iterations} (fp32 {r.stats['iterations_fp32']}, fp64 {r.stats['iterations_fp64']}) "
            f"{times[-1]:.3f}s")
    lib.blopex_amd_synchronize()
    total = time.perf_counter() - t_all
    rm = results[-1]

    out = {"metric": METRIC, "value": float(np.mean(times)), "unit": "s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * total / args.steps, "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic (seeded lognormal coefficients, drand48 X0)",
           "config": {"workload": f"LOBPCG k={k}, {nside}^3 7-point variable-coefficient "
                                  f"diffusion (lognormal sigma=1), Jacobi, residual tol 1e-6 and "
                                  f"eigenvalue error bound {EIG_RTOL:g} relative, mixed precision",
                      "n": A.n, "nnz": A.nnz, "k": k, "tol": 1e-6, "eig_rtol": EIG_RTOL,
                      "guard_columns": args.guard, "final_rr": True, "mode": "mixed"},
           "mixed": {"flag": [int(x.flag) for x in results],
                     "refresh_intervals": [intervals[s % len(intervals)] for s in range(args.steps)],
                     "seconds": times,
                     "iterations": [int(x.iterations) for x in results],
                     "iterations_fp32": [int(x.stats["iterations_fp32"]) for x in results],
                     "iterations_fp64": [int(x.stats["iterations_fp64"]) for x in results],
                     "ms_per_iteration": 1e3 * float(np.sum(times)) /
                                         float(sum(x.iterations for x in results)),
                     "max_residual": float(max(np.max(x.res) for x in results)),
                     "eig_error_bound": float(max(x.stats["eig_error_bound"] for x in results))}}

    if not args.no_fp64:
        # the FP64 oracle: FP64 with the eigenvalue error bound at 1e-12
        t0 = time.perf_counter()
        rref = solve("fp64", eig_rtol=1e-12, final_rr=True, guard_columns=args.guard)
        lib.blopex_amd_synchronize()
        tref = time.perf_counter() - t0
        ref = np.sort(rref.lam)
        log(f"fp64 oracle (eig bound 1e-12): flag {rref.flag} iterations {rref.iterations} "
            f"{tref:.3f}s bound {rref.stats['eig_error_bound']:.2e}")

        def accuracy(lam):
            lam = np.sort(lam)
            near = np.array([np.min(np.abs(ref - v)) / abs(v) for v in lam])
            return {"max_rel_err": float(np.max(np.abs(lam - ref) / np.abs(ref))),
                    "unmatched": int(np.sum(near > 1e-6))}

        t0 = time.perf_counter()
        r64 = solve("fp64", **head)
        lib.blopex_amd_synchronize()
        t64 = time.perf_counter() - t0
        log(f"fp64: flag {r64.flag} iterations {r64.iterations} {t64:.3f}s")
        acc_m = [accuracy(x.lam) for x in results]
        out["fp64"] = {"seconds": t64, "flag": r64.flag, "iterations": r64.iterations,
                       "ms_per_iteration": 1e3 * t64 / max(r64.iterations, 1),
                       "max_residual": float(np.max(r64.res)),
                       "eig_error_bound": r64.stats["eig_error_bound"]}
        out["speedup_mixed_vs_fp64"] = t64 / out["value"]
        out["accuracy_vs_fp64_oracle"] = {
            "oracle": {"mode": "fp64", "eig_rtol": 1e-12, "seconds": tref, "flag": rref.flag,
                       "iterations": rref.iterations,
                       "eig_error_bound": rref.stats["eig_error_bound"]},
            "mixed_max_rel_err": max(a["max_rel_err"] for a in acc_m),
            "mixed_unmatched": max(a["unmatched"] for a in acc_m),
            "fp64": accuracy(r64.lam)}
        # the same two modes stopped by the residual tolerance alone (no
        # eigenvalue bound, no final Rayleigh-Ritz): faster, and which 32
        # pairs they return depends on the trajectory
        plain = {}
        for mode in ("mixed", "fp64"):
            t0 = time.perf_counter()
            rp = solve(mode)
            lib.blopex_amd_synchronize()
            plain[mode] = dict(seconds=time.perf_counter() - t0, iterations=rp.iterations,
                               flag=rp.flag, **accuracy(rp.lam))
            log(f"tolerance only, {mode}: {plain[mode]}")
        out["tolerance_only"] = plain
        if args.guard == 0:
            # with 8 guard columns (40 iterated, 32 tested and returned)
            g8 = {}
            for mode in ("mixed", "fp64"):
                t0 = time.perf_counter()
                rg = solve(mode, eig_rtol=EIG_RTOL, final_rr=True, guard_columns=8)
                lib.blopex_amd_synchronize()
                g8[mode] = dict(seconds=time.perf_counter() - t0, iterations=rg.iterations,
                                flag=rg.flag, **accuracy(rg.lam))
                log(f"guard 8, {mode}: {g8[mode]}")
            out["guard8"] = g8

    # per-class device time of one profiled mixed solve
    rp = solve("mixed", profile=True, **head)
    st = rp.stats
    out["profile_mixed"] = {kk: st[kk] for kk in ("seconds", "seconds_spmm", "seconds_gram",
                                                  "seconds_update", "seconds_small_dense",
                                                  "seconds_other", "spmm_calls")}
    if st["seconds_spmm"] > 0:
        out["profile_mixed"]["spmm_gbps_in_solve"] = st["spmm_bytes"] / st["seconds_spmm"] / 1e9

    ms32, algo32 = spmm_roofline(lib, dA, A, k, L.HIPMV_FP32)
    ms64, algo64 = spmm_roofline(lib, dA, A, k, L.HIPMV_FP64)
    ach = algo32 / (ms32 * 1e-3) / 1e9
    nd = dA.ptr.contents.ndiag
    path = lib.hip_CSRMatrixProductKernel(dA.ptr, L.HIPMV_FP32, k, A.n)
    ksym = {2: f"k_zm<float, 32, {dA.ptr.contents.grid[0]}>", 1: f"k_dia<float, {nd}>"}.get(
        path, CSR_KERNEL)
    kname = ksym + {2: f" (z-marching 7-point stencil SpMM on the diagonal copy, k={k}, FP32)",
                    1: f" (SpMM on the diagonal copy of the CSR matrix, k={k}, FP32)"}.get(
        path, f" (CSR SpMM, k={k}, FP32)")
    from blopex_amd.profsum import read_traffic
    traffic = read_traffic(COUNTERS, ksym) if (nside, k) == (128, 32) else None
    # algorithmic bytes per launch as the metric defines them for a CSR SpMM:
    # values and int32 columns per nonzero, row pointers, X read once, Y
    # written once (661,913,604 B at 128^3, k = 32, FP32)
    csr_bytes = A.nnz * 8 + (A.n + 1) * 4 + 2 * A.n * k * 4
    ach_csr = csr_bytes / (ms32 * 1e-3) / 1e9
    out["roofline"] = {"kernel": kname, "bound": "hbm",
                       "achieved": ach_csr, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach_csr / HBM_PEAK_GBS, "traffic": traffic,
                       "traffic_source": os.path.relpath(COUNTERS, ROOT) if traffic else None,
                       "algorithmic_bytes": csr_bytes, "avg_launch_us": ms32 * 1e3}
    # the same launch priced with the bytes of the format the kernel reads
    # (the diagonal copy: no index streams), fewer than the CSR count
    out["roofline"]["format_bytes"] = {"bytes": algo32, "achieved": ach,
                                       "frac": ach / HBM_PEAK_GBS}
    # the same launch priced at the measured HBM traffic, and the rocprofv3
    # kernel-trace average of the committed summary beside the HIP-event time
    if traffic:
        out["roofline"]["traffic_rate"] = {"achieved": traffic / (ms32 * 1e-3) / 1e9,
                                           "frac": traffic / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS}
    from blopex_amd.profsum import read_trace_avg_us
    tr_us = read_trace_avg_us(TRACE, ksym) if (nside, k) == (128, 32) else None
    if tr_us:
        out["roofline"]["trace_avg_us"] = tr_us
        out["roofline"]["trace_frac_csr_bytes"] = csr_bytes / (tr_us * 1e-6) / 1e9 / HBM_PEAK_GBS
        out["roofline"]["trace_source"] = os.path.relpath(TRACE, ROOT)
    out["spmm_fp64"] = {"achieved": algo64 / (ms64 * 1e-3) / 1e9, "avg_launch_us": ms64 * 1e3,
                        "algorithmic_bytes": algo64}
    # the general CSR kernel (CSR_KERNEL: matrix slices staged through
    # LDS, thread per row) on the same matrix, for matrices without a
    # diagonal copy
    msc, algoc = spmm_roofline(lib, dA, A, k, L.HIPMV_FP32, force_csr=True)
    out["csr_general"] = {"kernel": CSR_KERNEL, "avg_launch_us": msc * 1e3,
                          "algorithmic_bytes": algoc,
                          "achieved": algoc / (msc * 1e-3) / 1e9,
                          "frac": algoc / (msc * 1e-3) / 1e9 / HBM_PEAK_GBS}
    trc = read_trace_avg_us(TRACE, CSR_KERNEL) if (nside, k) == (128, 32) else None
    if trc:
        out["csr_general"]["trace_avg_us"] = trc
    tfc = read_traffic(COUNTERS, CSR_KERNEL) if (nside, k) == (128, 32) else None
    if tfc:
        out["csr_general"]["traffic"] = tfc
    # MFMA rate of the FP32 Gram kernel in the mixed solve, from the committed
    # rocprofv3 summary (MOPS counters x 512 over the kernel-trace durations)
    from blopex_amd.profsum import read_mfma
    gm = read_mfma(MFMA_SUMMARY, GRAM_KERNEL) if (nside, k) == (128, 32) else None
    if gm:
        out["gram_mfma"] = {"kernel": GRAM_KERNEL, "bound": "mfma", "achieved": gm[0],
                            "peak": gm[1], "unit": "TFLOP/s", "frac": gm[0] / gm[1],
                            "source": os.path.relpath(MFMA_SUMMARY, ROOT)}

    if not args.no_cpu:
        cb = cpu_baseline(A, k)
        if cb is not None:
            its = out.get("fp64", {}).get("iterations", rm.iterations)
            gpu_it = out.get("fp64", {}).get("ms_per_iteration")
            out["cpu_baseline"] = {
                "value": cb["s_per_iteration"], "unit": "s per full-width iteration (k=32)",
                "iterations_per_s": 1.0 / cb["s_per_iteration"], "cores": 1, "kind": "reference",
                "s_per_iteration_each": cb["s_per_iteration_each"],
                "s_first_iteration_without_p": cb["s_first_iteration_without_p"],
                "setup_s": cb["setup_s"], "sample": cb["sample"],
                "gpu_fp64_ms_per_iteration": gpu_it,
                "extrapolated_time_to_tolerance_s": cb["s_per_iteration"] * its,
                "extrapolation": f"s per iteration x {its} iterations (the GPU FP64 count); "
                                 f"an upper bound, later iterations have fewer active columns"}
    for o in (X, X0, dT, dA):
        o.close()
    if not args.no_batched:
        out["batched"] = run_batched(args.batched_count, not args.batched_no_graph,
                                     mode=args.batched_mode)
        log(f"batched: {out['batched']}")
    if not args.no_config3:
        out["config3"] = run_config3()
        log(f"config3: {out['config3']}")
    if not args.no_adjoint:
        out["adjoint"] = run_adjoint()
        log(f"adjoint: {out['adjoint']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
