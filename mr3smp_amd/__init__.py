"""mr3smp_amd -- MI355X MRRR tridiagonal eigensolver.

The compute path is the HIP library ``libmr3hip.so`` built from ``csrc/`` for gfx950;
this module is a thin ctypes layer over its C ABI (``include/mr3hip.h``):

* :func:`mrrr` -- the drop-in routine with the reference's argument meaning
  (HPAC/mr3smp ``INCLUDE/mrrr.h:62-65``), on host numpy arrays;
* :func:`solve_batch` -- many problems in one pass of the device pipeline, on
  device tensors that stay in HBM;
* :class:`Server` -- the continuous-batching request loop.

There is no CPU fallback: if the library is missing or no HIP device is present the
calls raise.
"""
import ctypes
import os

import numpy as np

__all__ = ["lib", "mrrr", "solve_batch", "prepare_batch", "run_batch", "Problem", "Server", "Request", "build", "LIB_PATH"]

HERE = os.path.dirname(os.path.abspath(__file__))
# MR3_LIB selects another build of the library (a debug or profiling build)
LIB_PATH = os.environ.get("MR3_LIB") or os.path.join(HERE, "libmr3hip.so")
_lib = None
_P = ctypes.c_void_p


class Problem(ctypes.Structure):
    """Mirror of ``mr3_problem`` (include/mr3hip.h)."""
    _fields_ = [
        ("n", ctypes.c_int), ("jobz", ctypes.c_char), ("range", ctypes.c_char),
        ("pad0", ctypes.c_char), ("pad1", ctypes.c_char),
        ("vl", ctypes.c_double), ("vu", ctypes.c_double),
        ("il", ctypes.c_int), ("iu", ctypes.c_int), ("tryrac", ctypes.c_int),
        ("ldz", ctypes.c_int), ("mcols", ctypes.c_int),
        ("D", _P), ("E", _P), ("W", _P), ("Z", _P), ("Zsupp", _P),
        ("m", ctypes.c_int), ("info", ctypes.c_int),
        ("vl_out", ctypes.c_double), ("vu_out", ctypes.c_double),
        ("il_out", ctypes.c_int), ("iu_out", ctypes.c_int), ("tryrac_out", ctypes.c_int),
    ]


class Tuning(ctypes.Structure):
    """Mirror of ``mr3_tuning`` (include/mr3hip.h)."""
    _fields_ = [("side_streams", ctypes.c_int), ("tail_singletons", ctypes.c_int),
                ("group_lanes", ctypes.c_int), ("rrf_group_max", ctypes.c_int),
                ("singleton_waves", ctypes.c_int), ("small_first", ctypes.c_int)]


class Request(ctypes.Structure):
    """Mirror of ``mr3_request`` (include/mr3hip.h)."""
    _fields_ = [
        ("n", ctypes.c_int), ("jobz", ctypes.c_char), ("range", ctypes.c_char),
        ("pad0", ctypes.c_char), ("pad1", ctypes.c_char),
        ("vl", ctypes.c_double), ("vu", ctypes.c_double),
        ("il", ctypes.c_int), ("iu", ctypes.c_int), ("tryrac", ctypes.c_int),
        ("D", _P), ("E", _P), ("W", _P), ("Z", _P), ("Zsupp", _P),
        ("mcols", ctypes.c_int),
        ("m", ctypes.c_int), ("info", ctypes.c_int),
        ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
    ]


def build():
    """Compiles libmr3hip.so for gfx950 in place (hipcc)."""
    import subprocess
    subprocess.run(["bash", os.path.join(HERE, "csrc", "build.sh")], check=True)


def lib():
    """Loads libmr3hip.so; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run mr3smp_amd.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.mrrr.restype = ctypes.c_int
        L.mrrr.argtypes = [ctypes.c_char_p, ctypes.c_char_p] + [_P] * 13
        for name in ("dsyeig", "zheeig"):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = [ctypes.c_char_p] * 3 + [_P] * 11
        for name in ("dspeig", "zhpeig"):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = [ctypes.c_char_p] * 3 + [_P] * 10
        for name in ("dsygeig", "zhegeig", "dspgeig", "zhpgeig"):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = [_P] + [ctypes.c_char_p] * 3 + [_P] * 11
        L.mr3_dsyeig_device.restype = ctypes.c_int
        L.mr3_dsyeig_device.argtypes = ([ctypes.c_char] * 3 + [ctypes.c_int, _P, ctypes.c_int]
                                        + [_P] * 7 + [ctypes.c_int, ctypes.c_int, _P])
        L.mr3_get_tuning.restype = None
        L.mr3_get_tuning.argtypes = [ctypes.POINTER(Tuning)]
        L.mr3_set_tuning.restype = ctypes.c_int
        L.mr3_set_tuning.argtypes = [ctypes.POINTER(Tuning)]
        L.mr3_solve_batch.restype = ctypes.c_int
        L.mr3_solve_batch.argtypes = [ctypes.POINTER(Problem), ctypes.c_int, _P]
        L.mr3_solve_one.restype = ctypes.c_int
        L.mr3_solve_one.argtypes = [ctypes.POINTER(Problem), _P]
        L.mr3_solve_small_noalloc.restype = ctypes.c_int
        L.mr3_solve_small_noalloc.argtypes = [ctypes.POINTER(Problem), ctypes.c_int, _P]
        L.mr3_graph_create.restype = ctypes.c_int
        L.mr3_graph_create.argtypes = [ctypes.POINTER(Problem), ctypes.c_int, _P, ctypes.POINTER(_P)]
        L.mr3_graph_launch.restype = ctypes.c_int
        L.mr3_graph_launch.argtypes = [_P, ctypes.POINTER(Problem), _P]
        L.mr3_graph_destroy.restype = None
        L.mr3_graph_destroy.argtypes = [_P]
        L.mr3_server_create.restype = _P
        L.mr3_server_create.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_int]
        L.mr3_server_submit.restype = ctypes.c_longlong
        L.mr3_server_submit.argtypes = [_P, ctypes.POINTER(Request)]
        L.mr3_server_wait.restype = ctypes.c_int
        L.mr3_server_wait.argtypes = [_P, ctypes.c_longlong]
        L.mr3_server_drain.restype = None
        L.mr3_server_drain.argtypes = [_P]
        L.mr3_server_destroy.restype = None
        L.mr3_server_destroy.argtypes = [_P]
        L.mr3_now.restype = ctypes.c_double
        L.mr3_profile_enable.argtypes = [ctypes.c_int]
        L.mr3_profile_enable.restype = None
        L.mr3_profile_reset.restype = None
        L.mr3_profile_report.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.mr3_profile_report.restype = ctypes.c_int
        L.mr3_version.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def mrrr(jobz, rng, D, E, vl=0.0, vu=0.0, il=1, iu=1, tryrac=1, ldz=None, mcols=None):
    """The drop-in routine on host arrays.

    Returns a dict with info, m, W[m], Z (n x m or None), Zsupp[2m] and the in/out
    scalars vl, vu, il, iu, tryrac.  D and E are copied, never modified.
    """
    L = lib()
    n = int(len(D))
    D = np.ascontiguousarray(D, dtype=np.float64).copy()
    E = np.ascontiguousarray(E, dtype=np.float64).copy()
    if len(E) < n:
        E = np.concatenate([E, np.zeros(n - len(E))])
    ldz = n if ldz is None else int(ldz)
    if mcols is None:
        mcols = n
    wantz = jobz[0].upper() == "V"
    W = np.zeros(max(n, 1))
    Z = np.zeros((max(mcols, 1), max(ldz, 1))) if wantz else None
    Zs = np.zeros(2 * max(n, 1), dtype=np.int32)
    ni, m, cldz = ctypes.c_int(n), ctypes.c_int(0), ctypes.c_int(ldz)
    cvl, cvu = ctypes.c_double(vl), ctypes.c_double(vu)
    cil, ciu, ctr = ctypes.c_int(il), ctypes.c_int(iu), ctypes.c_int(tryrac)
    info = L.mrrr(jobz.encode(), rng.encode(), ctypes.byref(ni), _ptr(D), _ptr(E),
                  ctypes.byref(cvl), ctypes.byref(cvu), ctypes.byref(cil), ctypes.byref(ciu),
                  ctypes.byref(ctr), ctypes.byref(m), _ptr(W), _ptr(Z), ctypes.byref(cldz),
                  _ptr(Zs))
    mm = m.value
    out = dict(info=info, m=mm, W=W[:mm].copy(), Zsupp=Zs[:2 * mm].copy(), vl=cvl.value,
               vu=cvu.value, il=cil.value, iu=ciu.value, tryrac=ctr.value)
    # n x m view of the column-major output (column j is eigenvector j), no copy
    out["Z"] = Z[:mm, :n].T if wantz else None
    return out


def pack(A, uplo):
    """The uplo triangle of A packed by columns (LAPACK packed storage)."""
    A = np.asarray(A)
    n = A.shape[0]
    if uplo[0].upper() == "U":
        return np.concatenate([A[:j + 1, j] for j in range(n)]) if n else A.ravel()
    return np.concatenate([A[j:, j] for j in range(n)]) if n else A.ravel()


def unpack(AP, n, uplo):
    """Full matrix with the packed triangle of AP in place (the other triangle zero)."""
    AP = np.asarray(AP)
    A = np.zeros((n, n), dtype=AP.dtype)
    k = 0
    for j in range(n):
        if uplo[0].upper() == "U":
            A[:j + 1, j] = AP[k:k + j + 1]
            k += j + 1
        else:
            A[j:, j] = AP[k:k + n - j]
            k += n - j
    return A


def _dense_standard(name, cplx, packed, jobz, rng, uplo, A, n, vl, vu, il, iu):
    L = lib()
    dt = np.complex128 if cplx else np.float64
    # the standard drivers read A (or AP) and do not write it: no copy unless the layout
    # or type needs one
    if packed:
        Ah = np.ascontiguousarray(A, dtype=dt)
    else:
        Ah = np.asfortranarray(np.asarray(A, dtype=dt))
        n = int(Ah.shape[0])
    wantz = jobz[0].upper() == "V"
    mcols = (iu - il + 1) if rng[0].upper() == "I" else n
    W = np.zeros(max(n, 1))
    Z = np.zeros((max(n, 1), max(mcols, 1)), dtype=dt, order="F") if wantz else None
    ni, lda, ldz, m = ctypes.c_int(n), ctypes.c_int(max(n, 1)), ctypes.c_int(max(n, 1)), ctypes.c_int(0)
    cvl, cvu, cil, ciu = ctypes.c_double(vl), ctypes.c_double(vu), ctypes.c_int(il), ctypes.c_int(iu)
    head = [jobz.encode(), rng.encode(), uplo.encode(), ctypes.byref(ni), _ptr(Ah)]
    if not packed:
        head.append(ctypes.byref(lda))
    info = getattr(L, name)(*head, ctypes.byref(cvl), ctypes.byref(cvu), ctypes.byref(cil),
                            ctypes.byref(ciu), ctypes.byref(m), _ptr(W), _ptr(Z), ctypes.byref(ldz))
    mm = m.value
    return dict(info=info, m=mm, W=W[:mm].copy(), Z=Z[:, :mm] if wantz else None,
                vl=cvl.value, vu=cvu.value, il=cil.value, iu=ciu.value)


def _dense_general(name, cplx, packed, itype, jobz, rng, uplo, A, B, n, vl, vu, il, iu):
    L = lib()
    dt = np.complex128 if cplx else np.float64
    if packed:
        Ah = np.ascontiguousarray(A, dtype=dt).copy()
        Bh = np.ascontiguousarray(B, dtype=dt).copy()
    else:
        Ah = np.asfortranarray(np.asarray(A, dtype=dt)).copy(order="F")
        Bh = np.asfortranarray(np.asarray(B, dtype=dt)).copy(order="F")
        n = int(Ah.shape[0])
    wantz = jobz[0].upper() == "V"
    mcols = (iu - il + 1) if rng[0].upper() == "I" else n
    W = np.zeros(max(n, 1))
    Z = np.zeros((max(n, 1), max(mcols, 1)), dtype=dt, order="F") if (packed and wantz) else None
    it, ni, ld, m = ctypes.c_int(itype), ctypes.c_int(n), ctypes.c_int(max(n, 1)), ctypes.c_int(0)
    cvl, cvu, cil, ciu = ctypes.c_double(vl), ctypes.c_double(vu), ctypes.c_int(il), ctypes.c_int(iu)
    args = [ctypes.byref(it), jobz.encode(), rng.encode(), uplo.encode(), ctypes.byref(ni), _ptr(Ah)]
    if packed:
        args += [_ptr(Bh)]
    else:
        args += [ctypes.byref(ld), _ptr(Bh), ctypes.byref(ld)]
    args += [ctypes.byref(cvl), ctypes.byref(cvu), ctypes.byref(cil), ctypes.byref(ciu),
             ctypes.byref(m), _ptr(W)]
    if packed:
        args += [_ptr(Z), ctypes.byref(ld)]
    info = getattr(L, name)(*args)
    mm = m.value
    X = None
    if wantz:
        X = Z[:, :mm].copy() if packed else Ah[:, :mm].copy()
    return dict(info=info, m=mm, W=W[:mm].copy(), X=X, L=Bh)


def dsyeig(jobz, rng, uplo, A, vl=0.0, vu=0.0, il=1, iu=1):
    """Dense symmetric eigensolver on a host matrix (SRC/dense.c:151-230).

    A (n x n) is read through its uplo triangle and not modified.  Returns a dict
    with info, m, W[m], Z (n x m or None) and the in/out scalars vl, vu, il, iu.
    """
    return _dense_standard("dsyeig", False, False, jobz, rng, uplo, A, None, vl, vu, il, iu)


def dspeig(jobz, rng, uplo, AP, n, vl=0.0, vu=0.0, il=1, iu=1):
    """dsyeig on the packed triangle AP of order n (SRC/dense.c:319-398)."""
    return _dense_standard("dspeig", False, True, jobz, rng, uplo, AP, n, vl, vu, il, iu)


def zheeig(jobz, rng, uplo, A, vl=0.0, vu=0.0, il=1, iu=1):
    """Dense Hermitian eigensolver (SRC/dense.c:568-672); Z is complex."""
    return _dense_standard("zheeig", True, False, jobz, rng, uplo, A, None, vl, vu, il, iu)


def zhpeig(jobz, rng, uplo, AP, n, vl=0.0, vu=0.0, il=1, iu=1):
    """zheeig on the packed triangle AP of order n (SRC/dense.c:760-865)."""
    return _dense_standard("zhpeig", True, True, jobz, rng, uplo, AP, n, vl, vu, il, iu)


def dsygeig(itype, jobz, rng, uplo, A, B, vl=0.0, vu=0.0, il=1, iu=1):
    """Generalized symmetric-definite eigensolver on host matrices (SRC/dense.c:236-316).

    itype 1: A x = lambda B x, 2: A B x = lambda x, 3: B A x = lambda x.  A and B are
    not modified (the routine works on copies).  Returns info, m, W[m], X (n x m or
    None, the eigenvectors the reference leaves in A) and L (what the reference leaves
    in B: its Cholesky factor in the uplo triangle).
    """
    return _dense_general("dsygeig", False, False, itype, jobz, rng, uplo, A, B, None, vl, vu, il, iu)


def dspgeig(itype, jobz, rng, uplo, AP, BP, n, vl=0.0, vu=0.0, il=1, iu=1):
    """dsygeig on packed triangles (SRC/dense.c:403-471); L is the packed factor."""
    return _dense_general("dspgeig", False, True, itype, jobz, rng, uplo, AP, BP, n, vl, vu, il, iu)


def zhegeig(itype, jobz, rng, uplo, A, B, vl=0.0, vu=0.0, il=1, iu=1):
    """Generalized Hermitian-definite eigensolver (SRC/dense.c:676-757)."""
    return _dense_general("zhegeig", True, False, itype, jobz, rng, uplo, A, B, None, vl, vu, il, iu)


def zhpgeig(itype, jobz, rng, uplo, AP, BP, n, vl=0.0, vu=0.0, il=1, iu=1):
    """zhegeig on packed triangles (SRC/dense.c:869-935)."""
    return _dense_general("zhpgeig", True, True, itype, jobz, rng, uplo, AP, BP, n, vl, vu, il, iu)


def solve_batch(problems, stream=None):
    """Solves problems given as dicts with device tensors in one pipeline pass.

    Each dict holds ``D``, ``E`` (float64 tensors on the HIP device), ``jobz``,
    ``range`` and optional ``vl``, ``vu``, ``il``, ``iu``, ``tryrac``, ``W``, ``Z``
    (float64, shape (mcols, ldz): column j of the math is row j of the tensor),
    ``Zsupp`` (int32, 2*mcols).  Missing outputs are allocated.  Returns the
    Problem array (scalar outputs) and the list of dicts with the outputs filled.
    """
    arr, keep = prepare_batch(problems)
    return run_batch(arr, stream), arr, keep


def run_batch(arr, stream=None):
    """One mr3_solve_batch call on a Problem array from prepare_batch (the inputs
    it points at are read again, the outputs rewritten); returns the return code."""
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    return lib().mr3_solve_batch(arr, len(arr), st)


class BatchGraph:
    """A batched solve captured as one HIP graph (mr3_graph_create) over a Problem
    array from prepare_batch; launch() replays it (mr3_graph_launch) and returns the
    return code (4: the tree needed more levels or slots than the capture had)."""

    def __init__(self, arr, stream=None):
        self.arr = arr
        self.h = ctypes.c_void_p(None)
        st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        self.rc = lib().mr3_graph_create(arr, len(arr), st, ctypes.byref(self.h))
        if not self.h.value:
            raise RuntimeError(f"mr3_graph_create failed (rc {self.rc})")

    def launch(self, stream=None):
        st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        return lib().mr3_graph_launch(self.h, self.arr, st)

    def close(self):
        if self.h.value:
            lib().mr3_graph_destroy(self.h)
            self.h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_batch_noalloc(arr, stream=None):
    """One mr3_solve_small_noalloc call on a Problem array from prepare_batch (every
    problem of order n < 500); returns the return code."""
    st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
    return lib().mr3_solve_small_noalloc(arr, len(arr), st)


def prepare_batch(problems):
    """The Problem array of solve_batch, and the dicts with their outputs allocated,
    without solving."""
    import torch
    lib()
    arr = (Problem * len(problems))()
    keep = []
    for i, p in enumerate(problems):
        D, E = p["D"], p["E"]
        n = int(D.numel())
        dev = D.device
        jobz = p.get("jobz", "V")
        wantz = jobz.upper() == "V"
        mcols = int(p.get("mcols", n))
        if "W" not in p:
            p["W"] = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
        if wantz and "Z" not in p:
            p["Z"] = torch.empty((max(mcols, 1), n), dtype=torch.float64, device=dev)
        if wantz and "Zsupp" not in p:
            p["Zsupp"] = torch.empty(2 * max(mcols, 1), dtype=torch.int32, device=dev)
        q = arr[i]
        q.n = n
        q.jobz = jobz[0].encode()
        q.range = p.get("range", "A")[0].encode()
        q.vl, q.vu = float(p.get("vl", 0.0)), float(p.get("vu", 0.0))
        q.il, q.iu = int(p.get("il", 1)), int(p.get("iu", 1))
        q.tryrac = int(p.get("tryrac", 1))
        q.ldz = int(p["Z"].shape[1]) if wantz else n
        q.mcols = mcols if wantz else 0
        q.D, q.E = D.data_ptr(), E.data_ptr()
        q.W = p["W"].data_ptr()
        q.Z = p["Z"].data_ptr() if wantz else None
        q.Zsupp = p["Zsupp"].data_ptr() if wantz else None
        keep.append(p)
    return arr, keep


def get_tuning():
    """The process-wide launch-shape parameters (mr3_get_tuning) as a dict."""
    t = Tuning()
    lib().mr3_get_tuning(ctypes.byref(t))
    return {f: getattr(t, f) for f, _ in Tuning._fields_}


def set_tuning(**kw):
    """Changes some of the launch-shape parameters; raises ValueError when the library
    rejects a value (nothing is changed then).  Results never depend on them."""
    cur = get_tuning()
    unknown = set(kw) - set(cur)
    if unknown:
        raise ValueError(f"unknown tuning parameters {sorted(unknown)}")
    cur.update({k: int(v) for k, v in kw.items()})
    t = Tuning(**cur)
    if lib().mr3_set_tuning(ctypes.byref(t)) != 0:
        raise ValueError(f"tuning rejected: {cur}")
    return cur


def profile_enable(on=True, counters=True):
    """Per-kernel HIP-event timing for later solves, with the on-device work counters
    unless counters is False."""
    lib().mr3_profile_enable((1 if counters else 2) if on else 0)


def profile_reset():
    lib().mr3_profile_reset()


def profile_report():
    import json
    L = lib()
    size = 1 << 16
    while True:
        buf = ctypes.create_string_buffer(size)
        n = L.mr3_profile_report(buf, size)
        if n >= 0:
            return json.loads(buf.value.decode())
        size = -n + 16


class Server:
    """Continuous-batching request loop over host arrays (mr3_server_* in the C ABI)."""

    def __init__(self, budget_bytes=200 << 30, max_inflight=2, max_batch=256):
        self.L = lib()
        self.h = self.L.mr3_server_create(int(budget_bytes), int(max_inflight), int(max_batch))
        self._keep = {}

    def submit(self, jobz, rng, D, E, vl=0.0, vu=0.0, il=1, iu=1, tryrac=1, mcols=None):
        n = int(len(D))
        D = np.ascontiguousarray(D, dtype=np.float64)
        E = np.ascontiguousarray(E, dtype=np.float64)
        wantz = jobz[0].upper() == "V"
        mc = n if mcols is None else int(mcols)
        req = Request()
        req.n = n
        req.jobz, req.range = jobz[0].encode(), rng[0].encode()
        req.vl, req.vu, req.il, req.iu, req.tryrac = vl, vu, il, iu, tryrac
        W = np.zeros(max(n, 1))
        Z = np.zeros((max(mc, 1), n)) if wantz else None
        Zs = np.zeros(2 * max(mc, 1), dtype=np.int32) if wantz else None
        req.D, req.E, req.W = _ptr(D), _ptr(E), _ptr(W)
        req.Z, req.Zsupp = (_ptr(Z), _ptr(Zs)) if wantz else (None, None)
        # 0: no caller limit (Z holds n columns), so that the server may size a range-V
        # request by its interval's count
        req.mcols = (0 if mcols is None else mc) if wantz else 0
        rid = self.L.mr3_server_submit(self.h, ctypes.byref(req))
        if rid < 0:
            raise ValueError("request can never fit the memory budget")
        self._keep[rid] = (req, D, E, W, Z, Zs)
        return rid

    def wait(self, rid):
        info = self.L.mr3_server_wait(self.h, rid)
        req, D, E, W, Z, Zs = self._keep.pop(rid)
        m = req.m
        return dict(info=info, m=m, W=W[:m].copy(), Z=None if Z is None else Z[:m].T,
                    Zsupp=None if Zs is None else Zs[:2 * m].copy(),
                    t_submit=req.t_submit, t_start=req.t_start, t_done=req.t_done)

    def drain(self):
        self.L.mr3_server_drain(self.h)

    def close(self):
        if self.h:
            self.L.mr3_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
