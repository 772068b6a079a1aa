"""Writes tests/golden/ref_cases.npz from the compiled reference library
(oracle/_ref/libmrrr_ref.so, built from /root/reference by oracle/build_ref.sh).

Each case stores the inputs, the call parameters and the reference's outputs: info,
m, W, Zsupp and the accuracy figures resid = max|TZ - ZW| / (n eps ||T||) and
orth = max|Z^T Z - I| / (n eps) of its eigenvectors (Z itself is kept only for the
small cases).  Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from mrrr_ctypes import REF_SO, call_mrrr, load, resid_orth  # noqa: E402
from mr3smp_amd import workloads as WL  # noqa: E402

W21_DATA = "/root/reference/EXAMPLES/Tridiag/C/Wilkinson21.data"


def w21():
    D = np.abs(10.0 - np.arange(21.0))
    E = np.ones(21)
    E[20] = 0.0
    if os.path.exists(W21_DATA):   # the reference's own fixture holds the same matrix
        vals = np.loadtxt(W21_DATA, skiprows=1)
        assert np.array_equal(vals[:, 0], D) and np.array_equal(vals[:20, 1], E[:20]), "W21+ data"
    return D, E


def cases():
    rng = np.random.default_rng(20261203)
    D21, E21 = w21()
    out = [
        ("w21_VA", D21, E21, dict(jobz="V", rng="A")),
        ("w21_VI", D21, E21, dict(jobz="V", rng="I", il=3, iu=18)),
        ("w21_VV", D21, E21, dict(jobz="V", rng="V", vl=-1.22, vu=5.93)),
        ("w21_CV", D21, E21, dict(jobz="C", rng="V", vl=-1.22, vu=5.93)),
        ("w21_NA", D21, E21, dict(jobz="N", rng="A")),
    ]
    D, E = WL.uniform(200, rng); out.append(("uni200_VA", D, E, dict(jobz="V", rng="A")))
    D, E = WL.uniform(1000, WL.rng_for(1)); out.append(("u1000_VA", D, E, dict(jobz="V", rng="A")))
    out.append(("u1000_VA_threads", D, E, dict(jobz="V", rng="A", threads=16)))
    D, E = WL.uniform(500, rng); out.append(("uni500_NA", D, E, dict(jobz="N", rng="A")))
    D, E = WL.toeplitz_pert(300, rng); out.append(("toep300_VA", D, E, dict(jobz="V", rng="A")))
    D, E = WL.glued_wilkinson(420, rng, 1e-10); out.append(("glued420_VA", D, E, dict(jobz="V", rng="A")))
    D, E = WL.glued_wilkinson(840, rng, 1e-8); out.append(("glued840_VA_threads", D, E, dict(jobz="V", rng="A", threads=16)))
    D, E = WL.normal(600, rng); out.append(("normal600_VI", D, E, dict(jobz="V", rng="I", il=211, iu=270)))
    out.append(("normal600_VI_threads", D, E, dict(jobz="V", rng="I", il=211, iu=270, threads=4)))
    out.append(("normal600_VV", D, E, dict(jobz="V", rng="V", vl=-0.4, vu=0.9)))
    D, E = WL.uniform(60, rng); E[[9, 10, 30, 31, 32, 50]] = 0.0
    out.append(("split60_VA", D, E, dict(jobz="V", rng="A")))
    D = 4.0 + rng.uniform(-1, 1, 120); E = rng.uniform(-1, 1, 120); E[119] = 0
    out.append(("relacc120_VA", D, E, dict(jobz="V", rng="A", tryrac=1)))
    out.append(("relacc120_NA", D, E, dict(jobz="N", rng="A", tryrac=1)))
    D, E = WL.uniform(40, rng); out.append(("tiny40_VA", D * 1e-300, E * 1e-300, dict(jobz="V", rng="A")))
    D, E = WL.uniform(40, rng); out.append(("huge40_VA", D * 1e300, E * 1e300, dict(jobz="V", rng="A")))
    D = rng.uniform(-1, 1, 30); E = np.zeros(30); out.append(("diag30_VA", D, E, dict(jobz="V", rng="A")))
    out.append(("n1_VA", np.array([0.7]), np.array([0.0]), dict(jobz="V", rng="A")))
    out.append(("n2_VA", np.array([0.7, -0.2]), np.array([0.3, 0.0]), dict(jobz="V", rng="A")))
    return out


def kernel_vectors(refl, store):
    """Inputs/outputs of the reference's Fortran kernels (odneg, odr1v, odrrb, odrrf,
    odrnv) on a fixed L D L^T, for bit-level checks of the oracle's restatements."""
    import ctypes
    P = ctypes.c_void_p
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    Dd = lambda v: ctypes.byref(ctypes.c_double(v))
    ptr = lambda a: a.ctypes.data_as(P)
    rng = np.random.default_rng(7)
    n = 40
    Dt = rng.uniform(-1, 1, n)
    Et = rng.uniform(-1, 1, n)
    lam = np.linalg.eigvalsh(np.diag(Dt) + np.diag(Et[:-1], 1) + np.diag(Et[:-1], -1))
    sig = lam[0] - 0.01
    d = np.zeros(n); l = np.zeros(n)
    d[0] = Dt[0] - sig
    for i in range(n - 1):
        l[i] = Et[i] / d[i]
        d[i + 1] = (Dt[i + 1] - sig) - l[i] * Et[i]
    ld = d * l; lld = ld * l
    ev = lam - sig
    store["k/d"], store["k/l"], store["k/ld"], store["k/lld"], store["k/ev"] = d, l, ld, lld, ev
    pivmin = np.finfo(float).tiny
    refl.odneg_.restype = ctypes.c_int
    sigmas = np.concatenate([ev[::4] + 1e-9, ev[::4] - 1e-9, [0.0, -1.0, 10.0]])
    negs = []
    for s_ in sigmas:
        for r in (1, 17, n):
            negs.append(refl.odneg_(I(n), ptr(d), ptr(lld), Dd(s_), Dd(pivmin), I(r)))
    store["k/neg_sigmas"] = sigmas
    store["k/neg_counts"] = np.array(negs, dtype=np.int64)
    # dlar1v at a perturbed eigenvalue, twist free and fixed
    outs = []
    for k, rin, gaptol in ((5, 0, 0.0), (5, 23, 1e-14), (30, 0, 1e-12)):
        z = np.zeros(n); work = np.zeros(4 * n); isz = np.zeros(2, dtype=np.int32)
        negc = ctypes.c_int(0); ztz = ctypes.c_double(); ming = ctypes.c_double()
        rr = ctypes.c_int(rin); nrm = ctypes.c_double(); res = ctypes.c_double(); rq = ctypes.c_double()
        wantnc = ctypes.c_int(1)
        refl.odr1v_(I(n), I(1), I(n), Dd(ev[k] * (1 + 1e-10)), ptr(d), ptr(l), ptr(ld), ptr(lld),
                    Dd(pivmin), Dd(gaptol), ptr(z), ctypes.byref(wantnc), ctypes.byref(negc),
                    ctypes.byref(ztz), ctypes.byref(ming), ctypes.byref(rr), ptr(isz),
                    ctypes.byref(nrm), ctypes.byref(res), ctypes.byref(rq), ptr(work))
        outs.append(np.concatenate([[k, rin, gaptol, negc.value, ztz.value, ming.value, rr.value,
                                     isz[0], isz[1], nrm.value, res.value, rq.value], z]))
    store["k/r1v"] = np.array(outs)
    # dlarrb on rough brackets of all eigenvalues
    w = ev + 1e-6 * rng.uniform(-1, 1, n); werr = np.full(n, 3e-6)
    wgap = np.maximum(0, np.diff(np.concatenate([w, [w[-1] + 1]])) - 6e-6)
    store["k/rrb_in"] = np.stack([w, werr, wgap])
    wo, wgo, weo = w.copy(), wgap.copy(), werr.copy()
    work = np.zeros(2 * n); iwork = np.zeros(2 * n, dtype=np.int32); info = ctypes.c_int()
    refl.odrrb_(I(n), ptr(d), ptr(lld), I(1), I(n), Dd(1e-8), Dd(1e-11), I(0), ptr(wo), ptr(wgo),
                ptr(weo), ptr(work), ptr(iwork), Dd(pivmin), Dd(3.0), I(n), ctypes.byref(info))
    store["k/rrb_out"] = np.stack([wo, weo, wgo])
    # dlarrf for a fake cluster of eigenvalues 10..13
    cl = slice(10, 14)
    wc, gc, ec = ev[cl].copy(), np.full(4, 1e-3), np.full(4, 1e-12)
    dplus = np.zeros(n); lplus = np.zeros(n); work = np.zeros(2 * n); sig2 = ctypes.c_double()
    refl.odrrf_(I(n), ptr(d), ptr(l), ptr(ld), I(1), I(4), ptr(wc), ptr(gc), ptr(ec), Dd(3.0),
                Dd(2e-3), Dd(5e-3), Dd(pivmin), ctypes.byref(sig2), ptr(dplus), ptr(lplus),
                ptr(work), ctypes.byref(info))
    store["k/rrf_out"] = np.concatenate([[sig2.value, info.value], dplus, lplus[:n - 1]])
    # dlarnv IDIST=2 with ISEED = 1,1,1,1 (the root-perturbation stream)
    seed = (ctypes.c_int * 4)(1, 1, 1, 1)
    x = np.zeros(300)
    refl.odrnv_(I(2), seed, I(300), ptr(x))
    store["k/rnv300"] = x
    store["k/rnv_seed_after"] = np.array(list(seed), dtype=np.int64)


def main():
    refl, ref = load(REF_SO, "mrrr")
    store = {}
    names = []
    for name, D, E, kw in cases():
        kw = dict(kw)
        threads = kw.pop("threads", None)
        if threads:
            os.environ["PMR_NUM_THREADS"] = str(threads)
        else:
            os.environ.pop("PMR_NUM_THREADS", None)
        jobz, rg = kw.pop("jobz"), kw.pop("rng")
        r = call_mrrr(ref, jobz, rg, D, E, **kw)
        n = len(D)
        store[f"{name}/D"] = D
        store[f"{name}/E"] = E
        store[f"{name}/params"] = np.array([ord(jobz), ord(rg), kw.get("il", 1), kw.get("iu", 1),
                                            kw.get("tryrac", 1), threads or 1], dtype=np.int64)
        store[f"{name}/vlvu"] = np.array([kw.get("vl", 0.0), kw.get("vu", 0.0)])
        store[f"{name}/info_m"] = np.array([r["info"], r["m"]], dtype=np.int64)
        store[f"{name}/W"] = r["W"]
        if jobz == "V" and r["m"] > 0:
            res, orth = resid_orth(D, E, r["W"], r["Z"])
            store[f"{name}/resid_orth"] = np.array([res, orth])
            store[f"{name}/Zsupp"] = r["Zsupp"]
            if n <= 300:
                store[f"{name}/Z"] = r["Z"]
        names.append(name)
        print(name, n, jobz, rg, "info", r["info"], "m", r["m"], flush=True)
    kernel_vectors(refl, store)
    store["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "ref_cases.npz"), **store)
    print("wrote", os.path.join(HERE, "ref_cases.npz"))


if __name__ == "__main__":
    main()
