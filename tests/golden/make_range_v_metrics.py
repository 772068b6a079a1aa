"""Writes tests/golden/range_v_ref.npz: the compiled reference (oracle/_ref/libmrrr_ref.so,
built from /root/reference by oracle/build_ref.sh) on one N(0,1) tridiagonal of order
30000 (mr3smp_amd.workloads.normal with the generator stream rng_for(6)), jobz V,
range V over three windows: (-0.3, 0.2], (1.0, 1.4] and (-4.5, -2.0], the last one
holding the lower end of the spectrum.

Per window it stores vl, vu, m, the reference's resid = max|TZ - ZW| / (n eps ||T||)
and orth = max|Z^T Z - I| / (n eps), and its eigenvalues at the positions j = 0, every
max(1, m // 32)-th and m - 1 (concatenated, with offsets).  The matrix is not stored:
the generator rebuilds it from its fixed seed.  The reference runs its task-parallel
route (PMR_NUM_THREADS=8).  Run from the repository root:
python tests/golden/make_range_v_metrics.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from make_subset_metrics import metrics, sample_positions  # noqa: E402
from mrrr_ctypes import REF_SO, call_mrrr, load  # noqa: E402
from mr3smp_amd import workloads as WL  # noqa: E402

N = 30000
WINDOWS = [(-0.3, 0.2), (1.0, 1.4), (-4.5, -2.0)]


def matrix():
    return WL.normal(N, WL.rng_for(6))


def main():
    os.environ["PMR_NUM_THREADS"] = "8"
    _, fn = load(REF_SO, "mrrr")
    D, E = matrix()
    rows, wsamp, woff = [], [], [0]
    for vl, vu in WINDOWS:
        r = call_mrrr(fn, "V", "V", D, E, vl=vl, vu=vu)
        assert r["info"] == 0, r["info"]
        res, orth = metrics(D, E, r["W"], r["Z"])
        rows.append((vl, vu, r["m"], res, orth))
        wsamp.append(r["W"][sample_positions(r["m"])])
        woff.append(woff[-1] + len(wsamp[-1]))
        print(f"({vl}, {vu}]: m={r['m']} resid={res:.3g} orth={orth:.3g}", flush=True)
    rows = np.array(rows)
    np.savez_compressed(os.path.join(HERE, "range_v_ref.npz"), vl=rows[:, 0], vu=rows[:, 1],
                        m=rows[:, 2].astype(np.int64), resid=rows[:, 3], orth=rows[:, 4],
                        W=np.concatenate(wsamp), Woff=np.array(woff, dtype=np.int64))


if __name__ == "__main__":
    main()
