"""ctypes binding of libblopex_amd.so (the C ABI declared in include/*.h).

The library must have been built (python -m blopex_amd.build); loading fails
with an error otherwise -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libblopex_amd.so")

c_int, c_long, c_double, c_void_p = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
OP = ctypes.CFUNCTYPE(None, c_void_p, c_void_p, c_void_p)

HIPMV_FP64, HIPMV_FP32, HIPMV_C64 = 0, 1, 2
HIPLOBPCG_FP64, HIPLOBPCG_MIXED = 0, 1


class Tolerance(ctypes.Structure):
    _fields_ = [("absolute", c_double), ("relative", c_double)]


class BlasLapack(ctypes.Structure):
    _fields_ = [("dpotrf", c_void_p), ("dsygv", c_void_p), ("zpotrf", c_void_p), ("zhegv", c_void_p)]


class Settings(ctypes.Structure):
    _fields_ = [("mode", c_int), ("switch_factor", c_double), ("stall_iterations", c_int),
                ("refine_iterations", c_int), ("profile", c_int), ("refresh_interval", c_int),
                ("final_rr", c_int), ("guard_columns", c_int), ("eig_rtol", c_double),
                ("cond_restart", c_int), ("size_guard", c_int), ("guard_seed", c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("seconds", c_double), ("iterations_fp32", c_int), ("iterations_fp64", c_int),
                ("refine_iterations", c_int), ("exit_flag", c_int), ("seconds_small_dense", c_double),
                ("seconds_spmm", c_double), ("seconds_gram", c_double), ("seconds_update", c_double),
                ("seconds_other", c_double), ("spmm_bytes", c_double), ("spmm_calls", c_int),
                ("fp32_exit", c_int), ("final_rr", c_int), ("guard_columns", c_int),
                ("cond_restarts", c_int), ("eig_error_bound", c_double)]


class AdjointSettings(ctypes.Structure):
    _fields_ = [("mode", c_int), ("tol", c_double), ("maxit", c_int)]


class AdjointStats(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("corrections", c_int), ("residual", c_double),
                ("seconds", c_double)]


class BatchProblem(ctypes.Structure):
    _fields_ = [("n", c_long), ("nnz", c_long), ("rowptr", c_void_p), ("colind", c_void_p),
                ("val", c_void_p), ("k", c_int)]


class BatchSettings(ctypes.Structure):
    _fields_ = [("tol", c_double), ("maxit", c_int), ("jacobi", c_int), ("seed", c_int),
                ("graph", c_int), ("guard", c_int), ("mode", c_int), ("switch_factor", c_double),
                ("autotune", c_int)]


class BatchStats(ctypes.Structure):
    _fields_ = [("seconds", c_double), ("iterations", c_int), ("graph_captures", c_int),
                ("graph_replays", c_int), ("total_rows", c_long), ("iterations_fp32", c_int),
                ("chunk_rows", c_int), ("gram_rows", c_int)]


class HipMultiVector(ctypes.Structure):
    _fields_ = [("data", c_void_p), ("size", c_long), ("ld", c_long), ("num_vectors", c_int),
                ("num_active_vectors", c_int), ("active_indices", ctypes.POINTER(c_int)),
                ("owns_data", c_int), ("precision", c_int), ("settings", Settings),
                ("stats", Stats)]


class Interpreter(ctypes.Structure):
    """mv_InterfaceInterpreter: 26 function pointers (interpreter.h)."""
    _fields_ = [(f"slot{i}", c_void_p) for i in range(26)]


class CSRMatrix(ctypes.Structure):
    _fields_ = [("n", c_long), ("nnz", c_long), ("rowptr", c_void_p), ("colind", c_void_p),
                ("val", c_void_p), ("val32", c_void_p), ("ndiag", c_int),
                ("diag_offset", c_int * 16), ("diag_ld", c_long), ("diag_val", c_void_p),
                ("diag_val32", c_void_p), ("grid", c_int * 3)]


class DiagMatrix(ctypes.Structure):
    _fields_ = [("n", c_long), ("d", c_void_p), ("d32", c_void_p)]


class FortranMatrix(ctypes.Structure):
    _fields_ = [("globalHeight", c_long), ("height", c_long), ("width", c_long),
                ("value", c_void_p), ("ownsValues", c_int)]


class LobpcgInterp(ctypes.Structure):
    _fields_ = [("chol", c_void_p), ("solveGEVP", c_void_p)]


class FortranInterp(ctypes.Structure):
    _fields_ = [(f"slot{i}", c_void_p) for i in range(28)]


P = ctypes.POINTER
_SIGS = {
    "blopex_amd_init": (c_int, []),
    "blopex_amd_stream": (c_void_p, []),
    "blopex_amd_synchronize": (c_int, []),
    "blopex_amd_last_error": (ctypes.c_char_p, []),
    "blopex_amd_release_workspace": (c_int, []),
    "hip_Multi_VectorCreate": (P(HipMultiVector), [c_long, c_int, c_int]),
    "hip_Multi_VectorWrapDevice": (P(HipMultiVector), [c_void_p, c_long, c_int, c_int]),
    "hip_Multi_VectorDestroy": (None, [P(HipMultiVector)]),
    "hip_Multi_VectorUpload": (c_int, [P(HipMultiVector), c_void_p]),
    "hip_Multi_VectorDownload": (c_int, [P(HipMultiVector), c_void_p]),
    "hip_Multi_VectorSetRandomValues": (c_int, [P(HipMultiVector), c_int]),
    "hip_Multi_VectorSetSettings": (c_int, [P(HipMultiVector), P(Settings)]),
    "hip_LobpcgSettingsDefault": (None, [P(Settings)]),
    "hip_Multi_VectorInnerProdNoAlloc": (c_int, [P(HipMultiVector), P(HipMultiVector), c_int,
                                                 c_int, c_int, c_void_p]),
    "hip_Multi_VectorGetStats": (c_int, [P(HipMultiVector), P(Stats)]),
    "HipSetupInterpreter": (c_int, [P(Interpreter)]),
    "hip_CSRMatrixCreate": (P(CSRMatrix), [c_long, c_long, c_void_p, c_void_p, c_void_p]),
    "hip_CSRMatrixDestroy": (None, [P(CSRMatrix)]),
    "hip_CSRMatrixProductKernel": (c_int, [P(CSRMatrix), c_int, c_int, c_long]),
    "hip_DiagMatrixCreate": (P(DiagMatrix), [c_long, c_void_p]),
    "hip_DiagMatrixCreateJacobi": (P(DiagMatrix), [P(CSRMatrix)]),
    "hip_DiagMatrixDestroy": (None, [P(DiagMatrix)]),
    "mv_MultiVectorWrap": (c_void_p, [P(Interpreter), c_void_p, c_int]),
    "mv_MultiVectorDestroy": (None, [c_void_p]),
    "mv_MultiVectorGetData": (c_void_p, [c_void_p]),
    "mv_MultiVectorWidth": (c_int, [c_void_p]),
    "mv_MultiVectorHeight": (c_int, [c_void_p]),
    "mv_MultiVectorSetMask": (None, [c_void_p, c_void_p]),
    "mv_MultiVectorCreateCopy": (c_void_p, [c_void_p, c_int]),
    "mv_MultiVectorClear": (None, [c_void_p]),
    "mv_MultiVectorSetRandom": (None, [c_void_p, c_int]),
    "mv_MultiVectorCopy": (None, [c_void_p, c_void_p]),
    "mv_MultiVectorAxpy": (None, [c_double, c_void_p, c_void_p]),
    "mv_MultiVectorByMultiVector": (None, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "mv_MultiVectorByMultiVectorDiag": (None, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "mv_MultiVectorByMatrix": (None, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "mv_MultiVectorXapy": (None, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "mv_MultiVectorByDiagonal": (None, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "mv_MultiVectorEval": (None, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "lobpcg_solve_double": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, BlasLapack, Tolerance, c_int, c_int,
                                    P(c_int), c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int]),
    "lobpcg_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, BlasLapack, Tolerance, c_int, c_int, P(c_int),
                             c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int,
                             P(LobpcgInterp), P(FortranInterp)]),
    "lobpcg_solve_complex": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, BlasLapack, Tolerance, c_int, c_int,
                                     P(c_int), c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                     c_int]),
    "blopex_amd_default_interpreters": (None, [P(LobpcgInterp), P(FortranInterp)]),
    "blopex_amd_complex_interpreters": (None, [P(LobpcgInterp), P(FortranInterp)]),
    "zutilities_FortranMatrixMultiply": (None, [P(FortranMatrix), c_int, P(FortranMatrix), c_int,
                                                P(FortranMatrix)]),
    "zutilities_FortranMatrixSymmetrize": (None, [P(FortranMatrix)]),
    "zutilities_FortranMatrixTransposeSquare": (None, [P(FortranMatrix)]),
    "zutilities_FortranMatrixUpperInv": (None, [P(FortranMatrix)]),
    "zutilities_FortranMatrixCopy": (None, [P(FortranMatrix), c_int, P(FortranMatrix)]),
    "zutilities_FortranMatrixIndexCopy": (None, [c_void_p, P(FortranMatrix), c_int,
                                                 P(FortranMatrix)]),
    "zutilities_FortranMatrixAbs": (c_double, [P(FortranMatrix), c_long, c_long]),
    "blopex_amd_kernel_update": (c_int, [c_int, c_long, c_int, c_int, c_int, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_void_p]),
    "blopex_amd_kernel_update_resid": (c_int, [c_int, c_long, c_int, c_int, c_int] + [c_void_p] * 8
                                       + [c_int] + [c_void_p] * 8),
    "blopex_amd_kernel_gram": (c_int, [c_int, c_long, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "blopex_amd_kernel_mfma_probe": (c_int, [c_int, c_void_p, c_void_p]),
    "blopex_amd_kernel_rr": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                     c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "blopex_amd_kernel_rr_cond": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_int,
                                          c_void_p, c_int, c_double, c_void_p, c_void_p,
                                          c_void_p, c_void_p]),
    "hip_IJReadCSR": (c_int, [ctypes.c_char_p, P(c_long), P(c_long), P(P(c_int)), P(P(c_int)),
                              P(P(c_double))]),
    "hip_ParVectorsRead": (c_int, [ctypes.c_char_p, P(c_long), P(c_int), P(P(c_double))]),
    "hip_IJFree": (None, [c_void_p]),
    "hip_CSRMatrixReadIJ": (P(CSRMatrix), [ctypes.c_char_p]),
    "hip_Multi_VectorReadParVectors": (P(HipMultiVector), [ctypes.c_char_p]),
    "hip_BatchSettingsDefault": (None, [P(BatchSettings)]),
    "hip_LobpcgSolveBatched": (c_int, [c_int, P(BatchProblem), P(BatchSettings), c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p, P(BatchStats)]),
    "hip_AdjointSettingsDefault": (None, [P(AdjointSettings)]),
    "hip_LobpcgAdjoint": (c_int, [P(CSRMatrix), P(DiagMatrix), P(DiagMatrix), P(HipMultiVector),
                                  c_void_p, c_void_p, P(HipMultiVector), c_void_p, c_void_p,
                                  P(HipMultiVector), P(AdjointSettings), P(AdjointStats)]),
    "utilities_FortranMatrixUpperInv": (None, [P(FortranMatrix)]),
    "utilities_FortranMatrixSymmetrize": (None, [P(FortranMatrix)]),
}

# every function include/*.h declares (checked against the built library by the CPU tests)
HEADER_FUNCTIONS = sorted(_SIGS) + [
    "HipCSRMatMultiVec", "HipDiagMatMultiVec", "HipDenseMatMultiVec",
    "mv_MultiVectorCreateFromSampleVector",
    "mv_MultiVectorPrint",
]

_lib = None


def load() -> ctypes.CDLL:
    """Load the library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it with `python -m blopex_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def symbol_address(name: str) -> int:
    return ctypes.cast(getattr(load(), name), c_void_p).value


def last_error() -> str:
    return (load().blopex_amd_last_error() or b"").decode()
