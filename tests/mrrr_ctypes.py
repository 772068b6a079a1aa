"""ctypes bindings used by the tests to call any library exporting the reference's
mrrr() prototype (INCLUDE/mrrr.h:62-65): the HIP product, the CPU oracle, and the
compiled reference.  Also holds the accuracy metrics used across the test-suite.
"""
import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libmrrr_ref.so")

_P = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def load(path, symbol):
    lib = ctypes.CDLL(path)
    fn = getattr(lib, symbol)
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_char_p, ctypes.c_char_p] + [_P] * 13
    return lib, fn


def call_mrrr(fn, jobz, rng, D, E, vl=0.0, vu=0.0, il=1, iu=1, tryrac=1,
              want_z=True, mcols=None):
    """Calls an mrrr()-compatible function; returns a dict of outputs."""
    n = len(D)
    D = np.ascontiguousarray(D, dtype=np.float64).copy()
    E = np.ascontiguousarray(E, dtype=np.float64).copy()
    if len(E) < n:
        E = np.concatenate([E, np.zeros(n - len(E))])
    W = np.zeros(max(n, 1))
    if mcols is None:
        mcols = n
    Z = np.zeros((max(mcols, 1), n)) if (want_z and jobz.upper() == "V") else None
    Zs = np.zeros(2 * max(n, 1), dtype=np.int32)
    ni = ctypes.c_int(n)
    m = ctypes.c_int(0)
    ldz = ctypes.c_int(n)
    cvl = ctypes.c_double(vl)
    cvu = ctypes.c_double(vu)
    cil = ctypes.c_int(il)
    ciu = ctypes.c_int(iu)
    tr = ctypes.c_int(tryrac)
    info = fn(jobz.encode(), rng.encode(), ctypes.byref(ni), _ptr(D), _ptr(E),
              ctypes.byref(cvl), ctypes.byref(cvu), ctypes.byref(cil), ctypes.byref(ciu),
              ctypes.byref(tr), ctypes.byref(m), _ptr(W), _ptr(Z), ctypes.byref(ldz), _ptr(Zs))
    mm = m.value
    out = dict(info=info, m=mm, W=W[:mm].copy(), Zsupp=Zs[:2 * mm].copy(),
               vl=cvl.value, vu=cvu.value, il=cil.value, iu=ciu.value, tryrac=tr.value)
    out["Z"] = Z[:mm].T if Z is not None else None   # n x m view, no copy
    return out


def call_dstemr(lib, jobz, rng, D, E, vl=0.0, vu=0.0, il=1, iu=1, tryrac=1):
    """Calls the oracle's restatement of the reference's dstemr (oracle_dstemr) as
    mrrr() does for its sequential route (NZC = n); same output dict as call_mrrr."""
    fn = lib.oracle_dstemr
    fn.restype = ctypes.c_int
    fn.argtypes = ([ctypes.c_char, ctypes.c_char, ctypes.c_int, _P, _P, ctypes.c_double,
                    ctypes.c_double, ctypes.c_int, ctypes.c_int, _P, _P, _P, ctypes.c_int,
                    ctypes.c_int, _P, _P])
    n = len(D)
    D = np.ascontiguousarray(D, dtype=np.float64).copy()
    E = np.ascontiguousarray(E, dtype=np.float64).copy()
    W = np.zeros(max(n, 1))
    Z = np.zeros((max(n, 1), n)) if jobz.upper() == "V" else np.zeros(1)
    Zs = np.zeros(2 * max(n, 1), dtype=np.int32)
    m = ctypes.c_int(0)
    tr = ctypes.c_int(tryrac)
    info = fn(jobz.encode(), rng.encode(), n, _ptr(D), _ptr(E), vl, vu, il, iu,
              ctypes.byref(m), _ptr(W), _ptr(Z), max(n, 1), n, _ptr(Zs), ctypes.byref(tr))
    mm = m.value
    return dict(info=info, m=mm, W=W[:mm].copy(), Zsupp=Zs[:2 * mm].copy(), tryrac=tr.value,
                Z=Z[:mm].T if jobz.upper() == "V" else None)


def same_bits(got, ref, vectors=True):
    """True when two solver outputs agree bit for bit (m, W and, for jobz V, Z and
    Zsupp)."""
    if got["m"] != ref["m"] or not np.array_equal(got["W"], ref["W"]):
        return False
    if vectors and got.get("Z") is not None and ref.get("Z") is not None:
        return (np.array_equal(got["Zsupp"], ref["Zsupp"]) and
                np.array_equal(got["Z"], ref["Z"]))
    return True


def tnorm(D, E):
    n = len(D)
    T1 = np.abs(D).copy()
    T1[:-1] += np.abs(E[: n - 1])
    T1[1:] += np.abs(E[: n - 1])
    return float(T1.max()) if n else 0.0


def tz(D, E, Z):
    """T @ Z for a tridiagonal T, Z of shape (n, m)."""
    n = len(D)
    out = D[:, None] * Z
    if n > 1:
        e = E[: n - 1][:, None]
        out[:-1] += e * Z[1:]
        out[1:] += e * Z[:-1]
    return out


def resid_orth(D, E, W, Z):
    """resid = max|TZ - ZW| / (n eps ||T||_inf), orth = max|Z^T Z - I| / (n eps)."""
    n = len(D)
    eps = np.finfo(np.float64).eps
    tn = max(tnorm(D, E), np.finfo(np.float64).tiny)
    R = tz(D, E, Z) - Z * W[None, :]
    resid = float(np.abs(R).max()) / (n * eps * tn) if Z.size else 0.0
    G = Z.T @ Z
    G[np.diag_indices_from(G)] -= 1.0
    orth = float(np.abs(G).max()) / (n * eps) if Z.size else 0.0
    return resid, orth


def resid_orth_device(D, E, W, Z, chunk=512):
    """resid and orth as in resid_orth, computed on the device: D, E, W float64 tensors,
    Z of shape (m, n) with eigenvector j in row j (the batch API's layout)."""
    import torch
    n, m = D.numel(), W.numel()
    if m == 0 or n == 0:
        return 0.0, 0.0
    eps = np.finfo(np.float64).eps
    e = E[: n - 1]
    t1 = D.abs().clone()
    if n > 1:
        t1[:-1] += e.abs()
        t1[1:] += e.abs()
    tn = max(float(t1.max()), np.finfo(np.float64).tiny)
    rmax = 0.0
    for a in range(0, m, chunk):
        Zc = Z[a:a + chunk]
        R = Zc * D[None, :] - Zc * W[a:a + chunk, None]
        if n > 1:
            R[:, :-1] += e[None, :] * Zc[:, 1:]
            R[:, 1:] += e[None, :] * Zc[:, :-1]
        rmax = max(rmax, float(R.abs().max()))
    G = torch.mm(Z, Z.t().contiguous())
    G.diagonal().sub_(1.0)
    return rmax / (n * eps * tn), float(G.abs().max()) / (n * eps)


def support_ok_device(Z, Zsupp, chunk=512):
    """Zsupp (1-based, inclusive, two per row of Z) brackets the nonzeros of every
    eigenvector; Z of shape (m, n) on the device."""
    import torch
    m, n = Z.shape
    lo = Zsupp[0:2 * m:2].long()
    hi = Zsupp[1:2 * m:2].long()
    if not bool(((lo >= 1) & (lo <= hi) & (hi <= n)).all()):
        return False
    idx = torch.arange(n, device=Z.device)
    for a in range(0, m, chunk):
        out = (idx[None, :] < lo[a:a + chunk, None] - 1) | (idx[None, :] >= hi[a:a + chunk, None])
        if bool(((Z[a:a + chunk] != 0) & out).any()):
            return False
    return True
