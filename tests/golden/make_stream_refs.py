"""Writes the compiled reference's (oracle/_ref/libmrrr_ref.so, built from /root/reference
by oracle/build_ref.sh) figures for two workloads into tests/golden/:

* served_ref.npz -- every request of the served stream mr3smp_amd.workloads.
  served_stream(192, ...) (n log-uniform in [64, 16384], jobz 70% V / 30% N, range 60%
  A / 25% I / 15% V; the matrices do not depend on the arrival rate);
* glued20000_ref.npz -- the n = 20000 glued Wilkinson matrix (glue 1e-8), jobz V,
  range A (mr3smp_amd.workloads.single_glued_20000).

The reference runs with PMR_NUM_THREADS=16, as for mixed_ref.npz.  Per problem: n, m,
info, the eigenvalues at positions 0, every max(1, m // 32)-th and m - 1 (concatenated,
with offsets), and for jobz V resid = max|TZ - ZW| / (n eps ||T||) and orth =
max|Z^T Z - I| / (n eps) over all columns.  Inputs are not stored: the generators
rebuild them from their fixed seeds.  Takes several minutes (the Z^T Z products of the
largest problems).  Run from the repository root:  python tests/golden/make_stream_refs.py
"""
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
COUNT = 192


def sample_positions(m):
    return sorted(set([0, m - 1] + list(range(0, m, max(1, m // 32)))))


def solve(args):
    D, E, jobz, rng, kw = args
    os.environ["PMR_NUM_THREADS"] = str(THREADS)
    _, fn = load(REF_SO, "mrrr")
    r = call_mrrr(fn, jobz, rng, D, E, **kw)
    m = r["m"]
    res = orth = -1.0
    if jobz == "V" and r["info"] == 0 and m > 0:
        res, orth = resid_orth(D, E, r["W"], r["Z"])
    W = r["W"][sample_positions(m)] if m > 0 else np.zeros(0)
    return len(D), m, r["info"], res, orth, W


def pack(rows, path, **extra):
    woff = np.concatenate([[0], np.cumsum([len(r[5]) for r in rows])]).astype(np.int64)
    np.savez_compressed(path, n=np.array([r[0] for r in rows]), m=np.array([r[1] for r in rows]),
                        info=np.array([r[2] for r in rows]), resid=np.array([r[3] for r in rows]),
                        orth=np.array([r[4] for r in rows]),
                        W=np.concatenate([r[5] for r in rows]) if rows else np.zeros(0),
                        Woff=woff, threads=np.array(THREADS), **extra)


def main():
    t0 = time.time()
    reqs = WL.served_stream(COUNT, rate=1.0)
    jobs = []
    for r in reqs:
        kw = {k: r[k] for k in ("vl", "vu", "il", "iu") if k in r}
        jobs.append((r["D"], r["E"], r["jobz"], r["range"], kw))
    order = sorted(range(len(jobs)), key=lambda i: -len(jobs[i][0]))   # largest first
    with Pool(3) as p:
        out = dict(zip(order, p.map(solve, [jobs[i] for i in order], chunksize=1)))
    pack([out[i] for i in range(len(jobs))], os.path.join(HERE, "served_ref.npz"))
    print(f"wrote served_ref.npz ({time.time() - t0:.0f} s)", flush=True)
    D, E = WL.single_glued_20000()
    with Pool(1) as p:
        row = p.map(solve, [(D, E, "V", "A", {})])[0]
    pack([row], os.path.join(HERE, "glued20000_ref.npz"))
    print(f"wrote glued20000_ref.npz ({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
