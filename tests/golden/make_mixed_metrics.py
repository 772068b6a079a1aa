"""Writes tests/golden/mixed_ref.npz: the compiled reference (oracle/_ref/libmrrr_ref.so,
built from /root/reference by oracle/build_ref.sh) on every matrix of the batched mixed
workload (mr3smp_amd.workloads.batched_mixed(4096): n log-uniform in [32, 2048], 50%
U(-1,1), 25% perturbed 1-2-1, 25% glued W21+, jobz V, range A).

The reference runs with PMR_NUM_THREADS=16 (the core count of the host where the CPU
baseline is timed): matrices with n < 500 take its dstemr route (mrrr.c:149-153), larger
ones its task-parallel route with bisection for the root eigenvalues (more than 12
threads, mrrr_val.c:281-282).

Per matrix it stores n, m, resid = max|TZ - ZW| / (n eps ||T||), orth = max|Z^T Z - I| /
(n eps), the eigenvalues at positions 0, every max(1, m // 32)-th and m - 1
(concatenated, with offsets), and SHA-1 digests of the raw bytes of W[:m], of Z
(column-major, ldz = n) and of Zsupp (int32), so that a result can be compared with the
reference's bit for bit without storing it.  The inputs are not stored: the workload
generator rebuilds them from its fixed seed.  Takes about ten minutes on 8 cores.
Run from the repository root:  python tests/golden/make_mixed_metrics.py
"""
import hashlib
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from mrrr_ctypes import REF_SO, call_mrrr, load, resid_orth  # noqa: E402
from mr3smp_amd import workloads as WL  # noqa: E402

THREADS = 16
_MATS = None


def sample_positions(m):
    return sorted(set([0, m - 1] + list(range(0, m, max(1, m // 32)))))


def digests(W, Z, Zs):
    """SHA-1 of W, of Z as n x m column-major bytes, of Zsupp as int32 bytes."""
    w = hashlib.sha1(np.ascontiguousarray(W, dtype=np.float64).tobytes()).hexdigest()
    z = hashlib.sha1(np.ascontiguousarray(np.asarray(Z).T, dtype=np.float64).tobytes()).hexdigest()
    s = hashlib.sha1(np.ascontiguousarray(Zs, dtype=np.int32).tobytes()).hexdigest()
    return w, z, s


def one(i):
    global _MATS
    if _MATS is None:
        _MATS = WL.batched_mixed(4096)
    os.environ["PMR_NUM_THREADS"] = str(THREADS)
    _, fn = load(REF_SO, "mrrr")
    D, E, kind = _MATS[i]
    r = call_mrrr(fn, "V", "A", D, E)
    assert r["info"] == 0 and r["m"] == len(D), (i, r["info"], r["m"])
    res, orth = resid_orth(D, E, r["W"], r["Z"])
    return len(D), r["m"], res, orth, r["W"][sample_positions(r["m"])], digests(r["W"], r["Z"], r["Zsupp"])


def main():
    t0 = time.time()
    with Pool(8) as p:
        rows = p.map(one, range(4096), chunksize=8)
    n = np.array([r[0] for r in rows], dtype=np.int64)
    m = np.array([r[1] for r in rows], dtype=np.int64)
    woff = np.concatenate([[0], np.cumsum([len(r[4]) for r in rows])]).astype(np.int64)
    np.savez_compressed(os.path.join(HERE, "mixed_ref.npz"), n=n, m=m,
                        resid=np.array([r[2] for r in rows]), orth=np.array([r[3] for r in rows]),
                        W=np.concatenate([r[4] for r in rows]), Woff=woff,
                        hW=np.array([r[5][0] for r in rows]), hZ=np.array([r[5][1] for r in rows]),
                        hS=np.array([r[5][2] for r in rows]), threads=np.array(THREADS))
    print(f"wrote mixed_ref.npz ({time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
