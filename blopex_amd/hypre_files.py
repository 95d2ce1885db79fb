"""HYPRE IJ matrix and ParVector file sets: the inputs of the reference's
hypre workflows (`ij -lobpcg -fromfile prefix`, X0 from ParVector files).

Reading goes through the library (`hip_IJReadCSR`, `hip_ParVectorsRead`:
host code, no GPU needed).  The writers produce the layout of the reference's
MATLAB tools, blopex_tools/matlab/matlab2hypre/matlab2hypreIJ.m:56-99 and
matlab2hypreParVectors.m:40-75: rows split over `num_procs` files, the first
part taking the remainder rows (part boundaries 0, rest + s, rest + 2 s, ...,
n with s = n // num_procs), IJ files `prefix.00000` ... holding
"ilower iupper jlower jupper" and 0-based "i j value" lines sorted by row and
column, ParVector files `prefix.<vector>.<part>` holding the part's row count
and one value per line, with `prefix.<vector>.INFO.<part>` holding n and the
part starts.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .problems import CSR


def parts(n: int, num_procs: int) -> list[int]:
    """Row boundaries of the parts (matlab2hypreIJ.m:64-67)."""
    s, rest = n // num_procs, n % num_procs
    return [0] + [rest + s * (i + 1) for i in range(num_procs)]


def write_ij(A, num_procs: int, prefix: str, precision: str = "16.15e") -> None:
    """A (scipy sparse or CSR) as an IJ file set."""
    S = A.to_scipy() if isinstance(A, CSR) else A
    S = S.tocoo()
    n = S.shape[0]
    order = np.lexsort((S.col, S.row))
    rows, cols, vals = S.row[order], S.col[order], S.data[order]
    bounds = parts(n, num_procs)
    for p in range(num_procs):
        lo, hi = bounds[p], bounds[p + 1]
        sel = (rows >= lo) & (rows < hi)
        with open(f"{prefix}.{p:05d}", "w") as f:
            f.write(f"{lo} {hi - 1} {lo} {hi - 1}\n")
            for i, j, v in zip(rows[sel], cols[sel], vals[sel]):
                f.write(f"{i} {j} {v:{precision}}\n")


def write_parvectors(V: np.ndarray, num_procs: int, prefix: str, precision: str = "16.15e") -> None:
    """The columns of V (n x m) as a ParVector file set."""
    V = np.asarray(V, dtype=np.float64).reshape(V.shape[0], -1)
    n, m = V.shape
    bounds = parts(n, num_procs)
    for p in range(num_procs):
        lo, hi = bounds[p], bounds[p + 1]
        for j in range(m):
            with open(f"{prefix}.{j}.{p}", "w") as f:
                f.write(f"{hi - lo}\n")
                for v in V[lo:hi, j]:
                    f.write(f"{v:{precision}}\n")
            with open(f"{prefix}.{j}.INFO.{p}", "w") as f:
                f.write("\n".join(str(x) for x in [n] + bounds[:num_procs]) + "\n")


def read_ij(prefix: str) -> CSR:
    """The matrix of an IJ file set (repeated entries summed)."""
    lib = L.load()
    n, nnz = ctypes.c_long(), ctypes.c_long()
    rp, ci = ctypes.POINTER(ctypes.c_int)(), ctypes.POINTER(ctypes.c_int)()
    va = ctypes.POINTER(ctypes.c_double)()
    if lib.hip_IJReadCSR(prefix.encode(), ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(rp),
                         ctypes.byref(ci), ctypes.byref(va)) != 0:
        raise ValueError(L.last_error())
    try:
        indptr = np.ctypeslib.as_array(rp, (n.value + 1,)).copy()
        indices = np.ctypeslib.as_array(ci, (max(nnz.value, 1),))[: nnz.value].copy()
        data = np.ctypeslib.as_array(va, (max(nnz.value, 1),))[: nnz.value].copy()
    finally:
        for ptr in (rp, ci, va):
            lib.hip_IJFree(ctypes.cast(ptr, ctypes.c_void_p))
    return CSR(indptr, indices, data)


def read_parvectors(prefix: str) -> np.ndarray:
    """The n x m block of a ParVector file set."""
    lib = L.load()
    n, m = ctypes.c_long(), ctypes.c_int()
    va = ctypes.POINTER(ctypes.c_double)()
    if lib.hip_ParVectorsRead(prefix.encode(), ctypes.byref(n), ctypes.byref(m),
                              ctypes.byref(va)) != 0:
        raise ValueError(L.last_error())
    try:
        V = np.ctypeslib.as_array(va, (max(n.value * m.value, 1),))[: n.value * m.value].copy()
    finally:
        lib.hip_IJFree(ctypes.cast(va, ctypes.c_void_p))
    return V.reshape(m.value, n.value).T.copy(order="F")
