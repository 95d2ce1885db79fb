"""ctypes loader for the test oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so (the FP64 C restatement of lobpcg_solve) and, when
built, oracle/_ref/libblopex_ref.so (the reference's own lobpcg_solve_double
compiled from its sources plus a CSR/diagonal operator harness).  LAPACK
dpotrf/dsygv are handed to both as function pointers taken from scipy's
cython_lapack capsules, the way lobpcg_BLASLAPACKFunctions takes them
(blopex_abstract/krylov/lobpcg.h:25-46).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libblopex_ref.so")


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference where its sources are present)."""
    if force or not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)


def _capsule_ptr(cap) -> int:
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = ctypes.c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    return get_ptr(cap, get_name(cap))


def lapack_pointers() -> tuple[int, int]:
    from scipy.linalg import cython_lapack

    capi = cython_lapack.__pyx_capi__
    return _capsule_ptr(capi["dpotrf"]), _capsule_ptr(capi["dsygv"])


def zlapack_pointers() -> tuple[int, int]:
    from scipy.linalg import cython_lapack

    capi = cython_lapack.__pyx_capi__
    return _capsule_ptr(capi["zpotrf"]), _capsule_ptr(capi["zhegv"])


REF_COMPLEX_SO = os.path.join(HERE, "_ref", "libblopex_ref_complex.so")


def ref_solve_complex(A: np.ndarray, X0: np.ndarray, atol: float = 1e-6, rtol: float = 1e-50,
                      maxit: int = 1000) -> "SolveResult":
    """The reference's lobpcg_solve_complex (compiled in oracle/_ref) on a
    dense complex A, as its complex serial driver runs it."""
    lib = ctypes.CDLL(REF_COMPLEX_SO)
    fn = lib.ref_lobpcg_solve_complex
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                   ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_void_p]
    n, k = X0.shape
    Af = np.asfortranarray(A, dtype=np.complex128)
    X = np.asfortranarray(X0, dtype=np.complex128).copy(order="F")
    lam = np.zeros(k, dtype=np.complex128)
    lamh = np.zeros((k, maxit + 1), dtype=np.complex128, order="F")
    res = np.zeros(k)
    resh = np.zeros((k, maxit + 1), order="F")
    it = ctypes.c_int(0)
    zp, zh = zlapack_pointers()
    flag = fn(n, k, Af.ctypes.data, X.ctypes.data, atol, rtol, maxit, zp, zh, ctypes.byref(it),
              lam.ctypes.data, lamh.ctypes.data, res.ctypes.data, resh.ctypes.data)
    m = it.value
    return SolveResult(flag, m, X, lam, res, lamh[:, : m + 1], resh[:, : m + 1])


_SOLVE_ARGS = [
    ctypes.c_long, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
    ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
]

_libs: dict[str, ctypes.CDLL] = {}


def _load(path: str, prefix: str) -> ctypes.CDLL:
    if path in _libs:
        return _libs[path]
    lib = ctypes.CDLL(path)
    fn = getattr(lib, prefix + "_lobpcg_solve")
    fn.argtypes = _SOLVE_ARGS
    fn.restype = ctypes.c_int
    fill = getattr(lib, prefix + "_drand48_fill")
    fill.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_int, ctypes.c_int]
    fill.restype = None
    _libs[path] = lib
    return lib


def oracle_lib() -> ctypes.CDLL:
    build()
    lib = _load(ORACLE_SO, "oracle")
    lib.oracle_spmm.argtypes = [ctypes.c_long, ctypes.c_int] + [ctypes.c_void_p] * 5
    lib.oracle_upper_inv.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    return lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib() -> ctypes.CDLL:
    return _load(REF_SO, "ref")


def ref_operator_stamps() -> np.ndarray:
    """CLOCK_MONOTONIC at each operator-A call of the last reference solve
    (the initial AX, then one per iteration)."""
    lib = ref_lib()
    fn = lib.ref_operator_stamps
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int]
    fn.restype = ctypes.c_int
    out = np.zeros(64)
    m = fn(out.ctypes.data, 64)
    return out[:m]


@dataclass
class SolveResult:
    flag: int
    iterations: int
    X: np.ndarray
    lam: np.ndarray
    res: np.ndarray
    lam_hist: np.ndarray
    res_hist: np.ndarray


def _ptr(a):
    return None if a is None else a.ctypes.data


def solve(A, X0: np.ndarray, bdiag=None, tdiag=None, Y=None, atol=1e-6, rtol=1e-50, maxit=100,
          verbosity=0, which: str = "oracle") -> SolveResult:
    """Run LOBPCG in the oracle ("oracle") or the compiled reference ("ref")."""
    lib = oracle_lib() if which == "oracle" else ref_lib()
    fn = lib.oracle_lobpcg_solve if which == "oracle" else lib.ref_lobpcg_solve
    n, k = X0.shape
    X = np.asfortranarray(X0, dtype=np.float64).copy(order="F")
    Yf = None if Y is None else np.asfortranarray(Y, dtype=np.float64)
    ny = 0 if Y is None else Y.shape[1]
    bd = None if bdiag is None else np.ascontiguousarray(bdiag, dtype=np.float64)
    td = None if tdiag is None else np.ascontiguousarray(tdiag, dtype=np.float64)
    lam = np.zeros(k)
    res = np.zeros(k)
    lh = np.zeros((k, maxit + 1), order="F")
    rh = np.zeros((k, maxit + 1), order="F")
    it = ctypes.c_int(0)
    potrf, sygv = lapack_pointers()
    flag = fn(n, k, X.ctypes.data, A.rowptr.ctypes.data, A.colind.ctypes.data, A.val.ctypes.data,
              _ptr(bd), _ptr(td), ny, _ptr(Yf), atol, rtol, maxit, verbosity, potrf, sygv,
              lam.ctypes.data, lh.ctypes.data, k, res.ctypes.data, rh.ctypes.data, k,
              ctypes.byref(it))
    return SolveResult(flag, it.value, X, lam, res, lh[:, : it.value + 1], rh[:, : it.value + 1])


def drand48_block(n: int, k: int, seed: int, which: str = "oracle") -> np.ndarray:
    lib = oracle_lib() if which == "oracle" else ref_lib()
    fill = lib.oracle_drand48_fill if which == "oracle" else lib.ref_drand48_fill
    x = np.zeros((n, k), order="F")
    fill(x.ctypes.data, n, k, seed)
    return x


def spmm(A, X: np.ndarray) -> np.ndarray:
    lib = oracle_lib()
    Xf = np.asfortranarray(X, dtype=np.float64)
    Y = np.zeros_like(Xf, order="F")
    lib.oracle_spmm(A.n, Xf.shape[1], A.rowptr.ctypes.data, A.colind.ctypes.data,
                    A.val.ctypes.data, Xf.ctypes.data, Y.ctypes.data)
    return Y
