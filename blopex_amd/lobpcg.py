"""Python host side of the device LOBPCG, mirroring the BLOPEX harness contract.

lobpcg(X, A, B, T, Y, tol, maxit, verbosity) -> X, lambda, failureFlag,
lambdaHistory, residualNormsHistory, as blopex_matlab.m:279-287 returns them
(blopex_matlab/driver/blopex_matlab.m, gateway blopex_matlab_gateway.c:
64-67, 207-210), with the relative tolerance fixed at 1e-50 as the gateway
does (:66).  Everything runs through the C ABI of libblopex_amd.so:
lobpcg_solve_double on a hip_Multi_Vector wrapped by HipSetupInterpreter,
with HipCSRMatMultiVec / HipDiagMatMultiVec as the operator callbacks.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .problems import CSR


@dataclass
class Result:
    X: np.ndarray | None
    lam: np.ndarray
    flag: int
    iterations: int
    lam_hist: np.ndarray
    res_hist: np.ndarray
    res: np.ndarray
    stats: dict = field(default_factory=dict)


class DeviceCSR:
    def __init__(self, A: CSR):
        self.lib = L.load()
        self.n = A.n
        self.ptr = self.lib.hip_CSRMatrixCreate(A.n, A.nnz, A.rowptr.ctypes.data,
                                                A.colind.ctypes.data, A.val.ctypes.data)
        self.op = L.symbol_address("HipCSRMatMultiVec")

    def close(self):
        if self.ptr:
            self.lib.hip_CSRMatrixDestroy(self.ptr)
            self.ptr = None


class DeviceDiag:
    def __init__(self, d: np.ndarray | None = None, jacobi_of: DeviceCSR | None = None):
        self.lib = L.load()
        if jacobi_of is not None:
            self.ptr = self.lib.hip_DiagMatrixCreateJacobi(jacobi_of.ptr)
        else:
            d = np.ascontiguousarray(d, dtype=np.float64)
            self.ptr = self.lib.hip_DiagMatrixCreate(len(d), d.ctypes.data)
        self.op = L.symbol_address("HipDiagMatMultiVec")

    def close(self):
        if self.ptr:
            self.lib.hip_DiagMatrixDestroy(self.ptr)
            self.ptr = None


class DeviceBlock:
    """hip_Multi_Vector (n x k, FP64 unless precision says otherwise)."""

    def __init__(self, n: int, k: int, precision: int = L.HIPMV_FP64):
        self.lib = L.load()
        self.n, self.k = n, k
        self.ptr = self.lib.hip_Multi_VectorCreate(n, k, precision)

    @classmethod
    def from_numpy(cls, X: np.ndarray, precision: int | None = None) -> "DeviceBlock":
        """Upload X; complex arrays make HIPMV_C64 blocks."""
        if precision is None:
            precision = L.HIPMV_C64 if np.iscomplexobj(X) else L.HIPMV_FP64
        b = cls(X.shape[0], X.shape[1], precision)
        dt = np.complex128 if precision == L.HIPMV_C64 else np.float64
        Xf = np.asfortranarray(X, dtype=dt)
        b.lib.hip_Multi_VectorUpload(b.ptr, Xf.ctypes.data)
        return b

    def to_numpy(self) -> np.ndarray:
        cplx = self.ptr.contents.precision == L.HIPMV_C64
        out = np.zeros((self.n, self.k), dtype=np.complex128 if cplx else np.float64, order="F")
        self.lib.hip_Multi_VectorDownload(self.ptr, out.ctypes.data)
        return out

    def set_random(self, seed: int) -> None:
        self.lib.hip_Multi_VectorSetRandomValues(self.ptr, seed)

    def close(self):
        if self.ptr:
            self.lib.hip_Multi_VectorDestroy(self.ptr)
            self.ptr = None


class DeviceDense:
    """operatorA data for HipDenseMatMultiVec, the reference's MatMultiVec
    (blopex_serial_double/pcg_multi/pcg_multi.c:167-193): a dense n x n FP64
    (or complex) matrix on the device, or, with A None, no matrix, which
    makes the callback apply the 1D Laplacian (matmultivec.c:11-48)."""

    def __init__(self, A: np.ndarray | None, n: int | None = None):
        self.lib = L.load()
        self.block = None if A is None else DeviceBlock.from_numpy(np.asarray(A))
        self.n = n if A is None else A.shape[0]
        self.ptr = None if self.block is None else self.block.ptr
        self.op = L.symbol_address("HipDenseMatMultiVec")

    def close(self):
        if self.block is not None:
            self.block.close()
            self.block = self.ptr = None


def interpreter() -> L.Interpreter:
    ii = L.Interpreter()
    L.load().HipSetupInterpreter(ctypes.byref(ii))
    return ii


def solve_device(A: DeviceCSR | DeviceDense, X: DeviceBlock, B=None, T=None, Y: DeviceBlock | None = None,
                 tol: float = 1e-6, rtol: float = 1e-50, maxit: int = 200, verbosity: int = 0,
                 mode: str = "fp64", switch_factor: float = 8.0, stall_iterations: int = 50,
                 refine_iterations: int = 0, histories: bool = True, profile: bool = False,
                 refresh_interval: int = 20, final_rr: bool = False, guard_columns: int = 0,
                 eig_rtol: float = 0.0, cond_restart: bool = False, size_guard: bool = False,
                 guard_seed: int = 7) -> Result:
    """lobpcg_solve_double on device-resident operands; X is overwritten."""
    lib = L.load()
    k = X.k
    s = L.Settings()
    lib.hip_LobpcgSettingsDefault(ctypes.byref(s))
    s.mode = L.HIPLOBPCG_MIXED if mode == "mixed" else L.HIPLOBPCG_FP64
    s.switch_factor, s.stall_iterations = switch_factor, stall_iterations
    s.refine_iterations, s.profile = refine_iterations, int(profile)
    s.refresh_interval, s.final_rr = refresh_interval, int(final_rr)
    s.guard_columns, s.eig_rtol = guard_columns, eig_rtol
    s.cond_restart, s.size_guard, s.guard_seed = int(cond_restart), int(size_guard), guard_seed
    if lib.hip_Multi_VectorSetSettings(X.ptr, ctypes.byref(s)) != 0:
        raise ValueError(L.last_error())
    ii = interpreter()
    xx = lib.mv_MultiVectorWrap(ctypes.byref(ii), ctypes.cast(X.ptr, ctypes.c_void_p), 0)
    yy = None
    if Y is not None:
        yy = lib.mv_MultiVectorWrap(ctypes.byref(ii), ctypes.cast(Y.ptr, ctypes.c_void_p), 0)
    lam = np.zeros(k)
    res = np.zeros(k)
    lh = np.zeros((k, maxit + 1), order="F") if histories else None
    rh = np.zeros((k, maxit + 1), order="F") if histories else None
    it = ctypes.c_int(0)
    flag = lib.lobpcg_solve_double(
        xx, ctypes.cast(A.ptr, ctypes.c_void_p), A.op,
        None if B is None else ctypes.cast(B.ptr, ctypes.c_void_p), None if B is None else B.op,
        None if T is None else ctypes.cast(T.ptr, ctypes.c_void_p), None if T is None else T.op,
        yy, L.BlasLapack(), L.Tolerance(tol, rtol), maxit, verbosity, ctypes.byref(it),
        lam.ctypes.data, None if lh is None else lh.ctypes.data, k, res.ctypes.data,
        None if rh is None else rh.ctypes.data, k)
    st = L.Stats()
    lib.hip_Multi_VectorGetStats(X.ptr, ctypes.byref(st))
    lib.mv_MultiVectorDestroy(xx)
    if yy:
        lib.mv_MultiVectorDestroy(yy)
    n_it = it.value
    stats = {f: getattr(st, f) for f, _ in L.Stats._fields_}
    return Result(None, lam, flag, n_it,
                  np.zeros((k, 0)) if lh is None else lh[:, : n_it + 1],
                  np.zeros((k, 0)) if rh is None else rh[:, : n_it + 1], res, stats)


def lobpcg(X0: np.ndarray | None, A: CSR, B: np.ndarray | CSR | None = None, T=None,
           Y: np.ndarray | None = None, tol: float = 1e-6, maxit: int = 200, verbosity: int = 0,
           *, k: int | None = None, seed: int = 1, rtol: float = 1e-50, mode: str = "fp64",
           **settings) -> Result:
    """Smallest eigenpairs of A x = lambda B x.

    X0: n x k initial block (None: k columns of 2*drand48()-1 after srand48(seed),
    generated on the device).  B: None, a diagonal (1-D array) or a CSR matrix.
    T: None, "jacobi" (diag(A)^-1) or a 1-D array d (T = diag(d)).  Y: n x ny
    constraints.  mode: "fp64" or "mixed".
    """
    res = []
    dA = DeviceCSR(A)
    res.append(dA)
    try:
        dB = None
        if B is not None:
            dB = DeviceCSR(B) if isinstance(B, CSR) else DeviceDiag(B)
            res.append(dB)
        dT = None
        if isinstance(T, str):
            if T != "jacobi":
                raise ValueError(f"unknown preconditioner {T!r}")
            dT = DeviceDiag(jacobi_of=dA)
            res.append(dT)
        elif T is not None:
            dT = DeviceDiag(np.asarray(T))
            res.append(dT)
        if X0 is None:
            X = DeviceBlock(A.n, k)
            X.set_random(seed)
        else:
            X = DeviceBlock.from_numpy(X0)
        res.append(X)
        dY = None
        if Y is not None:
            dY = DeviceBlock.from_numpy(Y)
            res.append(dY)
        r = solve_device(dA, X, dB, dT, dY, tol=tol, rtol=rtol, maxit=maxit, verbosity=verbosity,
                         mode=mode, **settings)
        r.X = X.to_numpy()
        return r
    finally:
        for o in reversed(res):
            o.close()


@dataclass
class AdjointResult:
    gval: np.ndarray            # dL/d(stored values of A), CSR order
    gb: np.ndarray | None       # dL/d(diagonal of B)
    W: np.ndarray | None        # the adjoint vectors w_i
    flag: int                   # 0 converged, 1 CG short of tol
    stats: dict = field(default_factory=dict)


def adjoint(A: CSR, X: np.ndarray, lam: np.ndarray, glam: np.ndarray | None = None,
            GX: np.ndarray | None = None, B: np.ndarray | None = None, T=None,
            mode: str = "fp64", tol: float = 1e-10, maxit: int = 2000, want_gb: bool = False,
            want_W: bool = False) -> AdjointResult:
    """Gradients of a loss L(lambda, X) of the eigenpairs lobpcg returned, with
    respect to the stored values of A (and the diagonal of B), through
    hip_LobpcgAdjoint.  glam = dL/dlambda (k), GX = dL/dX (n x k); B: None or
    a diagonal; T as for lobpcg (preconditions the projected solves)."""
    lib = L.load()
    n, k = X.shape
    if B is not None and isinstance(B, CSR):
        raise ValueError("the adjoint takes B = I or a diagonal B")
    res = []
    try:
        dA = DeviceCSR(A)
        res.append(dA)
        dB = None
        if B is not None:
            dB = DeviceDiag(np.asarray(B))
            res.append(dB)
        dT = None
        if isinstance(T, str):
            if T != "jacobi":
                raise ValueError(f"unknown preconditioner {T!r}")
            dT = DeviceDiag(jacobi_of=dA)
            res.append(dT)
        elif T is not None:
            dT = DeviceDiag(np.asarray(T))
            res.append(dT)
        dX = DeviceBlock.from_numpy(X)
        res.append(dX)
        dG = None
        if GX is not None:
            dG = DeviceBlock.from_numpy(GX)
            res.append(dG)
        gv = DeviceBlock(A.nnz, 1)
        res.append(gv)
        gb = None
        if want_gb:
            gb = DeviceBlock(n, 1)
            res.append(gb)
        W = None
        if want_W:
            W = DeviceBlock(n, k)
            res.append(W)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        gl = None if glam is None else np.ascontiguousarray(glam, dtype=np.float64)
        s = L.AdjointSettings()
        lib.hip_AdjointSettingsDefault(ctypes.byref(s))
        s.mode = L.HIPLOBPCG_MIXED if mode == "mixed" else L.HIPLOBPCG_FP64
        s.tol, s.maxit = tol, maxit
        st = L.AdjointStats()
        flag = lib.hip_LobpcgAdjoint(
            dA.ptr, dB.ptr if dB else None, dT.ptr if dT else None, dX.ptr, lam.ctypes.data,
            None if gl is None else gl.ctypes.data, dG.ptr if dG else None,
            gv.ptr.contents.data, gb.ptr.contents.data if gb else None,
            W.ptr if W else None, ctypes.byref(s), ctypes.byref(st))
        if flag < 0:
            raise RuntimeError(L.last_error())
        return AdjointResult(gv.to_numpy()[:, 0], None if gb is None else gb.to_numpy()[:, 0],
                             None if W is None else W.to_numpy(), flag,
                             {f: getattr(st, f) for f, _ in L.AdjointStats._fields_})
    finally:
        for o in reversed(res):
            o.close()


def _zlapack() -> tuple[int, int]:
    """zpotrf / zhegv of scipy's LAPACK, as function pointers (the caller's
    LAPACK that lobpcg_BLASLAPACKFunctions carries)."""
    from scipy.linalg import cython_lapack

    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = ctypes.c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    capi = cython_lapack.__pyx_capi__
    return tuple(get_ptr(capi[f], get_name(capi[f])) for f in ("zpotrf", "zhegv"))


def lobpcg_complex(X0: np.ndarray, A: np.ndarray, tol: float = 1e-6, maxit: int = 1000,
                   verbosity: int = 0, rtol: float = 1e-50, B: np.ndarray | None = None,
                   Y: np.ndarray | None = None, T: np.ndarray | None = None,
                   mode: str = "fp64") -> Result:
    """Smallest eigenpairs of a dense complex Hermitian A (B: dense Hermitian
    positive definite, Y: constraints, T: dense preconditioner) through lobpcg_solve_complex on
    HIPMV_C64 blocks (the reference's complex driver,
    blopex_serial_complex/driver/serial_driver.c, on the device).  Blocks of
    up to 32 columns run on the device solver; wider ones on the generic
    algorithm with scipy's zpotrf/zhegv."""
    lib = L.load()
    n, k = X0.shape
    res_objs = []
    dense = L.symbol_address("HipDenseMatMultiVec")
    try:
        dA = DeviceBlock.from_numpy(np.asarray(A, dtype=np.complex128))
        res_objs.append(dA)
        dB = dT = None
        if B is not None:
            dB = DeviceBlock.from_numpy(np.asarray(B, dtype=np.complex128))
            res_objs.append(dB)
        if T is not None:
            dT = DeviceBlock.from_numpy(np.asarray(T, dtype=np.complex128))
            res_objs.append(dT)
        X = DeviceBlock.from_numpy(np.asarray(X0, dtype=np.complex128))
        res_objs.append(X)
        s = L.Settings()
        lib.hip_LobpcgSettingsDefault(ctypes.byref(s))
        s.mode = L.HIPLOBPCG_MIXED if mode == "mixed" else L.HIPLOBPCG_FP64
        if lib.hip_Multi_VectorSetSettings(X.ptr, ctypes.byref(s)) != 0:
            raise ValueError(L.last_error())
        ii = interpreter()
        xx = lib.mv_MultiVectorWrap(ctypes.byref(ii), ctypes.cast(X.ptr, ctypes.c_void_p), 0)
        yy = None
        if Y is not None:
            dY = DeviceBlock.from_numpy(np.asarray(Y, dtype=np.complex128))
            res_objs.append(dY)
            yy = lib.mv_MultiVectorWrap(ctypes.byref(ii), ctypes.cast(dY.ptr, ctypes.c_void_p), 0)
        zp, zh = _zlapack()
        blap = L.BlasLapack(None, None, zp, zh)
        lam = np.zeros(k, dtype=np.complex128)
        lh = np.zeros((k, maxit + 1), dtype=np.complex128, order="F")
        res = np.zeros(k)
        rh = np.zeros((k, maxit + 1), order="F")
        it = ctypes.c_int(0)
        flag = lib.lobpcg_solve_complex(
            xx, ctypes.cast(dA.ptr, ctypes.c_void_p), dense,
            None if dB is None else ctypes.cast(dB.ptr, ctypes.c_void_p),
            None if dB is None else dense,
            None if dT is None else ctypes.cast(dT.ptr, ctypes.c_void_p),
            None if dT is None else dense, yy, blap, L.Tolerance(tol, rtol), maxit,
            verbosity, ctypes.byref(it), lam.ctypes.data, lh.ctypes.data, k, res.ctypes.data,
            rh.ctypes.data, k)
        st = L.Stats()
        lib.hip_Multi_VectorGetStats(X.ptr, ctypes.byref(st))
        lib.mv_MultiVectorDestroy(xx)
        if yy:
            lib.mv_MultiVectorDestroy(yy)
        m = it.value
        stats = {f: getattr(st, f) for f, _ in L.Stats._fields_}
        return Result(X.to_numpy(), lam, flag, m, lh[:, : m + 1], rh[:, : m + 1], res, stats)
    finally:
        for o in reversed(res_objs):
            o.close()


@dataclass
class BatchResult:
    lam: list          # per problem, k eigenvalues (ascending)
    res: list          # per problem, k residual norms
    X: list | None     # per problem, n x k eigenvectors
    iterations: np.ndarray
    flags: np.ndarray
    stats: dict = field(default_factory=dict)


def lobpcg_batched(problems, tol: float = 1e-6, maxit: int = 500, jacobi: bool = True,
                   seed: int = 1, graph: bool = True, want_X: bool = False,
                   guard: int = 2, mode: str = "fp64",
                   switch_factor: float = 100.0, autotune: bool = True) -> BatchResult:
    """Smallest eigenpairs of many independent problems in one device solve
    (hip_LobpcgSolveBatched).  problems: list of (CSR, k)."""
    lib = L.load()
    arr = (L.BatchProblem * len(problems))()
    keep = []
    for i, (A, k) in enumerate(problems):
        keep.append(A)
        arr[i] = L.BatchProblem(A.n, A.nnz, A.rowptr.ctypes.data, A.colind.ctypes.data,
                                A.val.ctypes.data, k)
    s = L.BatchSettings()
    lib.hip_BatchSettingsDefault(ctypes.byref(s))
    s.tol, s.maxit, s.jacobi, s.seed, s.graph = tol, maxit, int(jacobi), seed, int(graph)
    s.guard = guard
    s.mode = L.HIPLOBPCG_MIXED if mode == "mixed" else L.HIPLOBPCG_FP64
    s.switch_factor = switch_factor
    s.autotune = int(autotune)
    ks = [k for _, k in problems]
    lam = np.zeros(sum(ks))
    res = np.zeros(sum(ks))
    its = np.zeros(len(problems), dtype=np.int32)
    flags = np.zeros(len(problems), dtype=np.int32)
    X = np.zeros(sum(A.n * k for A, k in problems)) if want_X else None
    st = L.BatchStats()
    rc = lib.hip_LobpcgSolveBatched(len(problems), arr, ctypes.byref(s), lam.ctypes.data,
                                    res.ctypes.data, its.ctypes.data, flags.ctypes.data,
                                    None if X is None else X.ctypes.data, ctypes.byref(st))
    if rc < 0 and np.all(flags == 0):  # an argument error (no problem was run)
        raise RuntimeError(L.last_error())
    lam_l, res_l, X_l = [], [], [] if want_X else None
    o = xo = 0
    for A, k in problems:
        lam_l.append(lam[o:o + k].copy())
        res_l.append(res[o:o + k].copy())
        if want_X:
            X_l.append(X[xo:xo + A.n * k].reshape((k, A.n)).T.copy())
            xo += A.n * k
        o += k
    return BatchResult(lam_l, res_l, X_l, its, flags,
                       {f: getattr(st, f) for f, _ in L.BatchStats._fields_})
