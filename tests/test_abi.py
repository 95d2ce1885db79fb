"""CPU checks of the C ABI: the built library loads and exports every function
include/*.h declares, and the struct layouts the bindings use match the
headers.  No GPU calls."""
import ctypes
import os
import re

import pytest

from blopex_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        # function declarations: identifier followed by '(' at top level ending in ');'
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof", "void", "double", "long", "int",
                        "char", "float", "BlopexInt", "komplex"):
                continue  # function-pointer members: `void (*Slot)(...)`
            names.add(name)
    # typedef'd function-pointer members are not exports
    return {n for n in names if not n[0].isupper() or n.startswith("Hip")}


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("library not built")
    return L.load()


def test_exports_every_declared_function(lib):
    declared = _declared_functions()
    assert "lobpcg_solve_double" in declared and "HipSetupInterpreter" in declared
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    for n in L.HEADER_FUNCTIONS:
        assert hasattr(lib, n), n


def test_struct_layouts():
    assert ctypes.sizeof(L.Interpreter) == 26 * 8
    assert ctypes.sizeof(L.FortranInterp) == 28 * 8
    assert ctypes.sizeof(L.FortranMatrix) == 40
    assert ctypes.sizeof(L.Tolerance) == 16
    assert ctypes.sizeof(L.BlasLapack) == 32
    # hip_Multi_Vector: data, size, ld, 2 ints, ptr, 2 ints, settings(40), stats(88)
    assert L.HipMultiVector.settings.offset == 48
    assert ctypes.sizeof(L.Settings) == 64 and ctypes.sizeof(L.Stats) == 104
    assert L.Settings.refresh_interval.offset == 28 and L.Settings.final_rr.offset == 32
    assert L.Settings.eig_rtol.offset == 40 and L.Settings.guard_seed.offset == 56
    assert L.Stats.eig_error_bound.offset == 96
    assert L.Stats.final_rr.offset == 80
    assert ctypes.sizeof(L.HipMultiVector) == 48 + 64 + 104
    # hip_CSRMatrix: 6 words, ndiag + 16 offsets (68 bytes, padded to 72), diag_ld, 2 pointers,
    # grid[3] (padded to 16)
    assert L.CSRMatrix.diag_ld.offset == 48 + 72 and L.CSRMatrix.grid.offset == 48 + 72 + 24
    assert ctypes.sizeof(L.CSRMatrix) == 48 + 72 + 24 + 16
    # hip_AdjointSettings {int, double, int}, hip_AdjointStats {int, int, double, double}
    assert ctypes.sizeof(L.AdjointSettings) == 24 and L.AdjointSettings.maxit.offset == 16
    assert ctypes.sizeof(L.AdjointStats) == 24 and L.AdjointStats.seconds.offset == 16
    # hip_BatchProblem {long, long, 3 pointers, int}, hip_BatchSettings {double, 5 ints},
    # hip_BatchStats {double, 3 ints, long, 3 ints}
    assert ctypes.sizeof(L.BatchProblem) == 48 and L.BatchProblem.k.offset == 40
    assert ctypes.sizeof(L.BatchSettings) == 48 and L.BatchSettings.autotune.offset == 40
    assert ctypes.sizeof(L.BatchStats) == 48 and L.BatchStats.chunk_rows.offset == 36
    assert L.BatchStats.gram_rows.offset == 40


def test_adjoint_defaults_and_argument_checks(lib):
    """Host-only paths of hip_LobpcgAdjoint: the defaults, and missing
    required arguments rejected before any device work."""
    s = L.AdjointSettings()
    lib.hip_AdjointSettingsDefault(ctypes.byref(s))
    assert (s.mode, s.tol, s.maxit) == (L.HIPLOBPCG_FP64, 1e-10, 2000)
    rc = lib.hip_LobpcgAdjoint(None, None, None, None, None, None, None, None, None, None, None,
                               None)
    assert rc == -1 and "required" in L.last_error()


def test_fortran_matrix_host_routines(lib):
    """utilities_FortranMatrixUpperInv / Symmetrize against numpy (host code only)."""
    import numpy as np

    rng = np.random.default_rng(3)
    n, gh = 7, 9
    U = np.triu(rng.standard_normal((n, n))) + 4 * np.eye(n)
    buf = np.zeros((gh, n), order="F")
    buf[:n, :] = U
    m = L.FortranMatrix(gh, n, n, buf.ctypes.data, 0)
    lib.utilities_FortranMatrixUpperInv(ctypes.byref(m))
    np.testing.assert_allclose(buf[:n, :] @ U, np.eye(n), atol=1e-12)
    S = rng.standard_normal((n, n))
    buf[:n, :] = S
    lib.utilities_FortranMatrixSymmetrize(ctypes.byref(m))
    np.testing.assert_array_equal(buf[:n, :], (S + S.T) * 0.5)


def test_complex_solver_requires_zlapack(lib):
    """lobpcg_solve_complex runs the generic algorithm, whose small dense work
    needs the caller's zpotrf / zhegv: without them it returns -1 at once."""
    it = ctypes.c_int(7)
    r = lib.lobpcg_solve_complex(None, None, None, None, None, None, None, None, L.BlasLapack(),
                                 L.Tolerance(1e-6, 0), 10, 0, ctypes.byref(it), None, None, 0,
                                 None, None, 0)
    assert r == -1 and it.value == 0
    assert "zpotrf" in L.last_error()


def test_c_driver_compiles_and_links(tmp_path):
    """The C caller (blopex_amd/driver/hip_driver.c, the serial driver's
    pattern) compiles as C99 against include/*.h with -Wall -Wextra -Werror
    and links to libblopex_amd.so, every symbol resolved (no GPU call)."""
    import subprocess

    from blopex_amd import build as b

    exe = b.build_driver(str(tmp_path / "hip_driver"))
    out = subprocess.run(["ldd", "-r", exe], capture_output=True, text=True)
    assert out.returncode == 0
    assert "libblopex_amd.so" in out.stdout
    assert "undefined symbol" not in out.stdout + out.stderr
