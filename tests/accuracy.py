"""Parity bars shared by the tests (stated once, used everywhere).

* counts: m must match exactly;
* eigenvalues: |W - W_ref| <= 4 n eps ||T||_inf;
* eigenvectors: resid and orth (definitions in mrrr_ctypes.resid_orth, in units of
  n eps) must be no worse than the reference's on the same input, with slack:
  <= 2 * max(ref, 1).
* below n = 500, where the reference's mrrr() takes LAPACK's dstemr route
  (SRC/mrrr.c:149-153), the library restates that route bit for bit: there the tests
  compare W, Z and Zsupp byte for byte (mrrr_ctypes.same_bits) on top of these bars.
"""
SAME_ALGO_FLOOR = 1.0
import numpy as np

from mrrr_ctypes import resid_orth, tnorm

EPS = np.finfo(np.float64).eps


def w_tol(D, E):
    return 4.0 * len(D) * EPS * max(tnorm(D, E), np.finfo(np.float64).tiny)


def vec_bound(ref_value, floor=SAME_ALGO_FLOOR):
    return 2.0 * max(ref_value, floor)


def check_pair(D, E, got, ref, label="", vectors=True, floor=SAME_ALGO_FLOOR):
    assert got["info"] == 0, f"{label}: info {got['info']}"
    assert got["m"] == ref["m"], f"{label}: m {got['m']} vs {ref['m']}"
    if got["m"] == 0:
        return
    dw = float(np.abs(np.asarray(got["W"]) - np.asarray(ref["W"])).max())
    assert dw <= w_tol(D, E), f"{label}: |dW| {dw:.3g} > {w_tol(D, E):.3g}"
    assert np.all(np.diff(got["W"]) >= 0), f"{label}: W not ascending"
    if vectors and got.get("Z") is not None:
        r1, o1 = resid_orth(D, E, got["W"], got["Z"])
        if ref.get("Z") is not None:
            r0, o0 = resid_orth(D, E, ref["W"], ref["Z"])
        else:
            r0, o0 = ref["resid"], ref["orth"]
        assert r1 <= vec_bound(r0, floor), f"{label}: resid {r1:.3g} vs ref {r0:.3g}"
        assert o1 <= vec_bound(o0, floor), f"{label}: orth {o1:.3g} vs ref {o0:.3g}"
        check_support(got["Z"], got["Zsupp"], label)


def check_support(Z, Zsupp, label=""):
    """Zsupp brackets the nonzeros of each column (1-based, inclusive)."""
    n, m = Z.shape
    for j in range(m):
        lo, hi = int(Zsupp[2 * j]), int(Zsupp[2 * j + 1])
        assert 1 <= lo <= hi <= n, f"{label}: support {lo},{hi}"
        col = Z[:, j]
        assert not np.any(col[: lo - 1]) and not np.any(col[hi:]), f"{label}: nonzero outside support col {j}"
