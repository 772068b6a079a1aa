"""Writes tests/golden/subset_ref.npz: the compiled reference (oracle/_ref/libmrrr_ref.so,
built from /root/reference by oracle/build_ref.sh) on every matrix of the batched subset
workload (mr3smp_amd.workloads.batched_range_i(256): N(0,1) tridiagonals, n log-uniform
in [4096, 50000], jobz V, range I over a seeded window of ceil(0.1 n) indices).

Per matrix it stores n, il, iu, m, the reference's resid = max|TZ - ZW| / (n eps ||T||)
and orth = max|Z^T Z - I| / (n eps), and its eigenvalues at the sampled positions
j = 0, every max(1, m // 32)-th and m - 1 of the window (concatenated, with offsets).
The inputs are not stored: the workload generator rebuilds them from its fixed seed.
The reference runs its task-parallel route (PMR_NUM_THREADS=8).  Takes about twenty
minutes on 8 cores.  Run from the repository root:  python tests/golden/make_subset_metrics.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from mrrr_ctypes import REF_SO, call_mrrr, load, tnorm, tz  # noqa: E402
from mr3smp_amd import workloads as WL  # noqa: E402


def sample_positions(m):
    return sorted(set([0, m - 1] + list(range(0, m, max(1, m // 32)))))


def metrics(D, E, W, Z):
    """resid and orth of mrrr_ctypes.resid_orth, with Z^T Z by blocks to bound memory."""
    n = len(D)
    eps = np.finfo(np.float64).eps
    tn = max(tnorm(D, E), np.finfo(np.float64).tiny)
    resid = 0.0
    for a in range(0, Z.shape[1], 512):
        Zc = Z[:, a:a + 512]
        resid = max(resid, float(np.abs(tz(D, E, Zc) - Zc * W[None, a:a + 512]).max()))
    G = Z.T @ Z
    G[np.diag_indices_from(G)] -= 1.0
    return resid / (n * eps * tn), float(np.abs(G).max()) / (n * eps)


def main():
    os.environ["PMR_NUM_THREADS"] = "8"
    _, fn = load(REF_SO, "mrrr")
    mats = WL.batched_range_i(256)
    rows, wsamp, woff = [], [], [0]
    t0 = time.time()
    for i, (D, E, il, iu) in enumerate(mats):
        r = call_mrrr(fn, "V", "I", D, E, il=il, iu=iu, mcols=iu - il + 1)
        assert r["info"] == 0 and r["m"] == iu - il + 1, (i, r["info"], r["m"])
        res, orth = metrics(D, E, r["W"], r["Z"])
        rows.append((len(D), il, iu, r["m"], res, orth))
        wsamp.append(r["W"][sample_positions(r["m"])])
        woff.append(woff[-1] + len(wsamp[-1]))
        print(f"{i} n={len(D)} m={r['m']} resid={res:.3g} orth={orth:.3g} "
              f"({time.time() - t0:.0f} s)", flush=True)
    rows = np.array(rows)
    np.savez_compressed(os.path.join(HERE, "subset_ref.npz"),
                        n=rows[:, 0].astype(np.int64), il=rows[:, 1].astype(np.int64),
                        iu=rows[:, 2].astype(np.int64), m=rows[:, 3].astype(np.int64),
                        resid=rows[:, 4], orth=rows[:, 5], W=np.concatenate(wsamp),
                        Woff=np.array(woff, dtype=np.int64))


if __name__ == "__main__":
    main()
