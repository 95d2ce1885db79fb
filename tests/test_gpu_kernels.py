"""GPU checks of the MFMA building blocks against numpy FP64 (host arrays
through include/blopex_amd_kernels.h).

Bars: FP64 kernels within 1e-13 of numpy relative to the sum of absolute
terms; FP32 kernels (FP32 inputs, FP32 MFMA, FP64 accumulation across 64-row
chunks for the Gram) within 64-96 FP32 roundings (4e-6 / 8e-6) of the same scale."""
import sys

import numpy as np
import pytest

from blopex_amd import _lib as L

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def f(a):
    return np.asfortranarray(a, dtype=np.float64)


@pytest.mark.parametrize("prec", [L.HIPMV_FP64, L.HIPMV_FP32])
def test_mfma_accumulator_layout(prec):
    """The (row, col) each lane/register holds is what the kernels assume:
    col = lane % 16; row = 4*(lane/16)+r (FP32) or lane/16+4*r (FP64)."""
    say(f"mfma layout probe prec={prec}")
    rows = np.zeros(256, dtype=np.int32)
    cols = np.zeros(256, dtype=np.int32)
    L.load().blopex_amd_kernel_mfma_probe(prec, rows.ctypes.data, cols.ctypes.data)
    lane = np.repeat(np.arange(64), 4)
    r = np.tile(np.arange(4), 64)
    np.testing.assert_array_equal(cols, lane % 16)
    want = 4 * (lane // 16) + r if prec == L.HIPMV_FP32 else lane // 16 + 4 * r
    np.testing.assert_array_equal(rows, want)


@pytest.mark.parametrize("prec", [L.HIPMV_FP64, L.HIPMV_FP32])
@pytest.mark.parametrize("n,k,a,p", [(1000, 20, 5, 13), (4096, 32, 32, 32), (777, 8, 8, 0),
                                     (130, 64, 64, 64), (64, 3, 2, 1)])
def test_block_update(prec, n, k, a, p):
    say(f"block update prec={prec} n={n} k={k} a={a} p={p}")
    rng = np.random.default_rng(n + k)
    X, R, P = (f(rng.standard_normal((n, w))) for w in (k, a, p))
    C = f(rng.standard_normal((k + a + p, k)))
    if prec == L.HIPMV_FP32:
        X, R, P = (f(v.astype(np.float32)) for v in (X, R, P))
    full = np.zeros((n, k), order="F")
    tail = np.zeros((n, k), order="F")
    L.load().blopex_amd_kernel_update(prec, n, k, a, p, X.ctypes.data, R.ctypes.data,
                                      P.ctypes.data, C.ctypes.data, full.ctypes.data,
                                      tail.ctypes.data)
    ref_tail = R @ C[k:k + a] + P @ C[k + a:]
    ref_full = ref_tail + X @ C[:k]
    scale = np.abs(X) @ np.abs(C[:k]) + np.abs(R) @ np.abs(C[k:k + a]) + np.abs(P) @ np.abs(C[k + a:])
    tol = 1e-13 if prec == L.HIPMV_FP64 else 8e-6
    if a + p:
        assert np.max(np.abs(tail - ref_tail) / (scale + 1e-300)) < tol
    assert np.max(np.abs(full - ref_full) / (scale + 1e-300)) < tol


@pytest.mark.parametrize("prec", [L.HIPMV_FP64, L.HIPMV_FP32])
@pytest.mark.parametrize("n,k,a,p,active,tdiag", [
    (1000, 20, 5, 13, "some", True), (4096, 32, 32, 32, "all", True), (777, 8, 8, 0, "all", False),
    (130, 32, 3, 3, "some", True), (64, 3, 2, 1, "all", False), (70000, 32, 16, 16, "some", True)])
def test_block_update_with_residual(prec, n, k, a, p, active, tdiag):
    """The fused update and residual against numpy: X', AX' (in place), P'
    and AP' at the compact tail positions of the active columns, and
    R = t .* (X' lam - AX') with its squared norms.  Row counts that are not
    a multiple of 64 leave padding rows, where the diagonal t is not read."""
    say(f"block update + residual prec={prec} n={n} k={k} a={a} p={p} {active} tdiag={tdiag}")
    rng = np.random.default_rng(n + 3 * k + a)
    X, AX, R, AR, P, AP = (f(rng.standard_normal((n, w))) for w in (k, k, a, a, p, p))
    C = f(rng.standard_normal((k + a + p, k)))
    lam = np.sort(rng.uniform(0.5, 2.0, k))
    t = rng.uniform(0.5, 1.5, n) if tdiag else None
    if prec == L.HIPMV_FP32:
        X, AX, R, AR, P, AP = (f(v.astype(np.float32)) for v in (X, AX, R, AR, P, AP))
        if t is not None:
            t = t.astype(np.float32).astype(np.float64)
    cols = list(range(k)) if active == "all" else [c for c in range(k) if c % 3 != 1]
    tail_pos = np.full(k, -1, dtype=np.int32)
    tail_pos[cols] = np.arange(len(cols), dtype=np.int32)
    nt = len(cols)
    Xo, AXo = (np.zeros((n, k), order="F") for _ in range(2))
    Po, APo, Ro = (np.zeros((n, nt), order="F") for _ in range(3))
    rn2 = np.zeros(nt)
    rc = L.load().blopex_amd_kernel_update_resid(
        prec, n, k, a, p, X.ctypes.data, AX.ctypes.data, R.ctypes.data, AR.ctypes.data,
        P.ctypes.data, AP.ctypes.data, C.ctypes.data, tail_pos.ctypes.data, nt, lam.ctypes.data,
        None if t is None else t.ctypes.data, Xo.ctypes.data, AXo.ctypes.data, Po.ctypes.data,
        APo.ctypes.data, Ro.ctypes.data, rn2.ctypes.data)
    assert rc == 0
    tol = 1e-13 if prec == L.HIPMV_FP64 else 8e-6
    for In, Rin, Pin, full, tail in ((X, R, P, Xo, Po), (AX, AR, AP, AXo, APo)):
        ref_tail = Rin @ C[k:k + a] + Pin @ C[k + a:]
        ref_full = ref_tail + In @ C[:k]
        scale = (np.abs(In) @ np.abs(C[:k]) + np.abs(Rin) @ np.abs(C[k:k + a])
                 + np.abs(Pin) @ np.abs(C[k + a:]))
        assert np.max(np.abs(full - ref_full) / (scale + 1e-300)) < tol
        if a + p:
            st = scale[:, cols]
            assert np.max(np.abs(tail - ref_tail[:, cols]) / (st + 1e-300)) < tol
    # the residual from the kernel's own X' and AX' (exact in either precision)
    v = Xo[:, cols] * lam[cols] - AXo[:, cols]
    vs = np.abs(Xo[:, cols] * lam[cols]) + np.abs(AXo[:, cols])
    rtol = 1e-15 if prec == L.HIPMV_FP64 else 2e-7
    ref_r = v * (t[:, None] if t is not None else 1.0)
    rs = vs * (t[:, None] if t is not None else 1.0)
    assert np.max(np.abs(Ro - ref_r) / (rs + 1e-300)) < 4 * rtol
    np.testing.assert_allclose(rn2, np.sum(v * v, axis=0), rtol=1e-12 if prec == L.HIPMV_FP64 else 1e-6)


@pytest.mark.parametrize("prec", [L.HIPMV_FP64, L.HIPMV_FP32])
@pytest.mark.parametrize("n,lw,qw", [(1000, 20, 13), (70000, 32, 32), (257, 64, 96), (64, 1, 1),
                                     (5000, 96, 64)])
def test_gram(prec, n, lw, qw):
    say(f"gram prec={prec} n={n} lw={lw} qw={qw}")
    rng = np.random.default_rng(n + lw)
    Lm, Qm = f(rng.standard_normal((n, lw))), f(rng.standard_normal((n, qw)))
    if prec == L.HIPMV_FP32:
        Lm, Qm = f(Lm.astype(np.float32)), f(Qm.astype(np.float32))
    G = np.zeros((lw, qw), order="F")
    L.load().blopex_amd_kernel_gram(prec, n, lw, qw, Lm.ctypes.data, Qm.ctypes.data, G.ctypes.data)
    ref = Lm.T @ Qm
    scale = np.abs(Lm).T @ np.abs(Qm)
    tol = 1e-14 if prec == L.HIPMV_FP64 else 4e-6
    assert np.max(np.abs(G - ref) / scale) < tol


def test_fp32_gram_error_does_not_grow_with_n():
    """The FP32 Gram (FP32 MFMA over one staged chunk of at most 256 rows,
    the chunk sums added in FP64): its error against the exact product of the
    same FP32 data is bounded by the chunk's rounding, c u32 |L|'|Q| with c
    <= 256 rows, whatever n -- a plain FP32 accumulation over n rows would
    carry n u32.  Measured at n = 2^16 and 2^20: both within 256 u32, and the
    larger not growing with n (at most 2x the smaller)."""
    u32 = 2.0 ** -24
    errs = []
    for n in (1 << 16, 1 << 20):
        say(f"fp32 gram error bound n={n}")
        rng = np.random.default_rng(3)
        Lm = f(rng.standard_normal((n, 16)).astype(np.float32))
        Qm = f(rng.standard_normal((n, 16)).astype(np.float32))
        G = np.zeros((16, 16), order="F")
        L.load().blopex_amd_kernel_gram(L.HIPMV_FP32, n, 16, 16, Lm.ctypes.data, Qm.ctypes.data,
                                        G.ctypes.data)
        exact = Lm.T @ Qm  # FP64 of FP32 data: exact to 1e-16 relative here
        scale = np.abs(Lm).T @ np.abs(Qm)
        errs.append(np.max(np.abs(G - exact) / scale))
        say(f"  max |G - exact| / (|L|'|Q|) = {errs[-1]:.2e} (256 u32 = {256 * u32:.2e}, "
            f"n u32 = {n * u32:.2e})")
    assert max(errs) < 256 * u32
    assert errs[1] < 2.0 * errs[0] + 4 * u32
