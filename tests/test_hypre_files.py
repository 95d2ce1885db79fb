"""HYPRE IJ matrix and ParVector file sets (the reference's hypre inputs):
reading through the library's host code, against files written by hand in
the layout of blopex_tools/matlab/matlab2hypre/matlab2hypreIJ.m and
matlab2hypreParVectors.m, and round trips through the writers.  CPU only,
except the last test (a device solve from files)."""
import ctypes
import os

import numpy as np
import pytest

from blopex_amd import _lib as L
from blopex_amd import hypre_files as hf
from blopex_amd import problems as pb

pytestmark = pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="library not built")


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def test_parts_match_the_matlab_partition():
    # matlab2hypreIJ.m:64-67: part = [0 (rest + s):s:n], s = floor(n / P)
    assert hf.parts(5, 2) == [0, 3, 5]
    assert hf.parts(10, 3) == [0, 4, 7, 10]
    assert hf.parts(8, 4) == [0, 2, 4, 6, 8]
    assert hf.parts(3, 1) == [0, 3]


def test_read_ij_hand_written(tmp_path):
    """5 x 5 over two parts (rows 0-2 and 3-4), one entry given twice (summed)."""
    pre = str(tmp_path / "A")
    _write(pre + ".00000", "0 2 0 2\n0 0 4.0\n0 1 -1.0\n1 0 -1.0\n1 1 4.0\n1 2 -1.0\n"
                           "2 1 -1.0\n2 2 3.5\n2 2 0.5\n")
    _write(pre + ".00001", "3 4 3 4\n3 3 4.0\n3 4 -1.0\n4 3 -1.0\n4 4 4.0\n4 0 2.5e-1\n")
    A = hf.read_ij(pre)
    D = np.zeros((5, 5))
    D[0, 0] = D[1, 1] = D[2, 2] = D[3, 3] = D[4, 4] = 4.0
    D[0, 1] = D[1, 0] = D[1, 2] = D[2, 1] = D[3, 4] = D[4, 3] = -1.0
    D[4, 0] = 0.25
    np.testing.assert_array_equal(A.to_scipy().toarray(), D)
    assert A.nnz == 12
    for r in range(5):  # columns ascending inside each row
        cols = A.colind[A.rowptr[r]:A.rowptr[r + 1]]
        assert np.all(np.diff(cols) > 0)


def test_ij_round_trip(tmp_path):
    pre = str(tmp_path / "L")
    A = pb.laplacian_7pt(6, 5, 4, sigma=1.0, seed=3)
    hf.write_ij(A, 3, pre, precision="17.16e")
    assert sorted(os.listdir(tmp_path)) == ["L.00000", "L.00001", "L.00002"]
    bounds = hf.parts(A.n, 3)
    for p in range(3):
        head = open(f"{pre}.{p:05d}").readline().split()
        assert [int(x) for x in head] == [bounds[p], bounds[p + 1] - 1] * 2
    B = hf.read_ij(pre)
    np.testing.assert_array_equal(B.rowptr, A.rowptr)
    np.testing.assert_array_equal(B.colind, A.colind)
    np.testing.assert_array_equal(B.val, A.val)  # 17 digits: exact


def test_read_parvectors_hand_written(tmp_path):
    pre = str(tmp_path / "x")
    _write(pre + ".0.0", "3\n1.0\n2.0\n3.0\n")
    _write(pre + ".0.1", "2\n4.0\n5.0\n")
    _write(pre + ".1.0", "3\n-1.0\n-2.0\n-3.0\n")
    _write(pre + ".1.1", "2\n-4.0\n-5.0e0\n")
    _write(pre + ".0.INFO.0", "5\n0\n3\n")
    V = hf.read_parvectors(pre)
    np.testing.assert_array_equal(V, np.array([[1, -1], [2, -2], [3, -3], [4, -4], [5, -5]], float))


def test_parvectors_round_trip(tmp_path):
    pre = str(tmp_path / "v")
    V = np.random.default_rng(1).standard_normal((23, 4))
    hf.write_parvectors(V, 5, pre, precision="17.16e")
    assert open(pre + ".2.INFO.3").read().split() == ["23"] + [str(b) for b in hf.parts(23, 5)[:5]]
    np.testing.assert_array_equal(hf.read_parvectors(pre), V)


def test_read_errors(tmp_path):
    with pytest.raises(ValueError, match="no file"):
        hf.read_ij(str(tmp_path / "none"))
    pre = str(tmp_path / "B")
    _write(pre + ".00000", "0 1 0 1\n0 0 1.0\n1 x 2.0\n")
    with pytest.raises(ValueError, match="line 3"):
        hf.read_ij(pre)
    _write(pre + ".00000", "0 1 0 1\n0 0 1.0\n2 1 2.0\n")
    with pytest.raises(ValueError, match="outside the part"):
        hf.read_ij(pre)
    _write(pre + ".00000", "0 1 0 1\n0 0 1.0\n1 7 2.0\n")
    with pytest.raises(ValueError, match="column 7"):
        hf.read_ij(pre)
    pv = str(tmp_path / "w")
    _write(pv + ".0.0", "3\n1.0\n2.0\n")
    with pytest.raises(ValueError, match="2 of 3"):
        hf.read_parvectors(pv)


@pytest.mark.gpu
def test_device_solve_from_files(tmp_path):
    """A and X0 from IJ / ParVector files through the device constructors give
    the same solve as the in-memory matrix and block."""
    from blopex_amd import lobpcg as bl
    from oracle import pyoracle as po

    A = pb.laplacian_7pt(9, 8, 7, sigma=1.0, seed=6)
    X0 = po.drand48_block(A.n, 4, 1)
    hf.write_ij(A, 4, str(tmp_path / "A"), precision="17.16e")
    hf.write_parvectors(X0, 2, str(tmp_path / "X"), precision="17.16e")
    lib = L.load()
    dA = lib.hip_CSRMatrixReadIJ(str(tmp_path / "A").encode())
    assert dA and dA.contents.n == A.n and dA.contents.nnz == A.nnz
    lib.hip_CSRMatrixDestroy(dA)
    x = lib.hip_Multi_VectorReadParVectors(str(tmp_path / "X").encode())
    assert x and x.contents.size == A.n and x.contents.num_vectors == 4
    back = np.zeros((A.n, 4), order="F")
    lib.hip_Multi_VectorDownload(x, back.ctypes.data)
    lib.hip_Multi_VectorDestroy(x)
    np.testing.assert_array_equal(back, X0)
    r0 = bl.lobpcg(X0, A, T="jacobi", maxit=200)
    r1 = bl.lobpcg(hf.read_parvectors(str(tmp_path / "X")), hf.read_ij(str(tmp_path / "A")),
                   T="jacobi", maxit=200)
    assert r0.iterations == r1.iterations
    np.testing.assert_array_equal(r0.lam, r1.lam)
