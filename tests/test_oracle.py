"""CPU tests of the oracle (oracle/mr3_oracle.c), the checker for the HIP path.

The oracle is pinned against (1) the reference's own outputs stored in
tests/golden/ref_cases.npz (written by tests/golden/make_golden.py from the compiled
reference), bit for bit at the kernel level (dlaneg, dlar1v, dlarrb, dlarrf, dlarnv)
and to the parity bars at the mrrr() level; (2) the compiled reference itself when
oracle/_ref is built; (3) what the mathematics fixes: closed-form spectra, dense
eigensolvers on small inputs, Sylvester inertia for Sturm counts.
"""
import ctypes
import os

import numpy as np
import pytest

from accuracy import EPS, check_pair, check_support, w_tol
from mrrr_ctypes import REF_SO, call_dstemr, call_mrrr, load, resid_orth, same_bits, tnorm

P = ctypes.c_void_p


def ptr(a):
    return a.ctypes.data_as(P)


def dense(D, E):
    n = len(D)
    return np.diag(D) + np.diag(E[: n - 1], 1) + np.diag(E[: n - 1], -1)


# ---------------------------------------------------------------------------------
# mrrr() level, against the reference's stored outputs
# ---------------------------------------------------------------------------------
def golden_names():
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_cases.npz"))
    return [str(x) for x in z["names"]]


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_outputs(name, golden, oracle):
    c = golden[name]
    _, fn = oracle
    D, E = c["D"], c["E"]
    got = call_mrrr(fn, c["jobz"], c["rng"], D, E, **c["kw"])
    info, m = (int(v) for v in c["info_m"])
    ref = dict(info=info, m=m, W=c["W"])
    if c["jobz"] == "C":
        assert got["info"] == 0 and got["m"] == m
        return
    if "resid_orth" in c:
        ref["resid"], ref["orth"] = (float(v) for v in c["resid_orth"])
    if "Z" in c:
        ref["Z"] = c["Z"]
        ref["Zsupp"] = c["Zsupp"]
    if len(D) < 500:
        # the reference took dstemr (mrrr.c:149-153); the oracle restates it bit for bit
        assert got["info"] == info and got["m"] == m, name
        assert same_bits(got, ref, vectors=("Z" in c)), name
        return
    check_pair(D, E, got, ref, name, vectors=(c["jobz"] == "V"))


@pytest.mark.parametrize("name", [n for n in golden_names()])
def test_dstemr_restatement_on_reference_outputs(name, golden, oracle_lib):
    """Every stored case that the reference solved through dstemr (n < 500, or one
    thread at any n: mrrr.c:149-153) is reproduced bit for bit by oracle_dstemr."""
    c = golden[name]
    D, E = c["D"], c["E"]
    if not (len(D) < 500 or c["threads"] == 1) or c["jobz"] == "C":
        pytest.skip("the reference took its task-parallel route")
    kw = c["kw"]
    got = call_dstemr(oracle_lib, c["jobz"], c["rng"], D, E, kw["vl"], kw["vu"], kw["il"],
                      kw["iu"], kw["tryrac"])
    info, m = (int(v) for v in c["info_m"])
    assert got["info"] == 0 and got["m"] == m
    ref = dict(m=m, W=c["W"], Z=c.get("Z"), Zsupp=c.get("Zsupp"))
    assert same_bits(got, ref, vectors=("Z" in c)), name


# ---------------------------------------------------------------------------------
# kernel level: bit-identical to the reference's Fortran on stored vectors
# ---------------------------------------------------------------------------------
def test_dlaneg_bit_exact(golden, oracle_lib):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_cases.npz"))
    d, lld = z["k/d"], z["k/lld"]
    n = len(d)
    f = oracle_lib.oracle_dlaneg
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    got = [f(n, ptr(d), ptr(lld), float(s), np.finfo(float).tiny, r)
           for s in z["k/neg_sigmas"] for r in (1, 17, n)]
    assert np.array_equal(np.array(got), z["k/neg_counts"])


def test_dlar1v_bit_exact(oracle_lib):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_cases.npz"))
    d, l, ld, lld, ev = z["k/d"], z["k/l"], z["k/ld"], z["k/lld"], z["k/ev"]
    n = len(d)
    f = oracle_lib.oracle_dlar1v
    f.restype = None
    f.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] + [P] * 4 + [ctypes.c_double] * 2 +
                  [P, ctypes.c_int] + [P] * 9)
    for row in z["k/r1v"]:
        k, rin, gaptol = int(row[0]), int(row[1]), float(row[2])
        zz = np.zeros(n)
        work = np.zeros(4 * n)
        isz = np.zeros(2, dtype=np.int32)
        negc, r = ctypes.c_int(), ctypes.c_int(rin)
        ztz, ming, nrm, res, rq = (ctypes.c_double() for _ in range(5))
        f(n, 1, n, float(ev[k] * (1 + 1e-10)), ptr(d), ptr(l), ptr(ld), ptr(lld),
          np.finfo(float).tiny, gaptol, ptr(zz), 1, ctypes.byref(negc), ctypes.byref(ztz),
          ctypes.byref(ming), ctypes.byref(r), ptr(isz), ctypes.byref(nrm), ctypes.byref(res),
          ctypes.byref(rq), ptr(work))
        got = np.concatenate([[k, rin, gaptol, negc.value, ztz.value, ming.value, r.value,
                               isz[0], isz[1], nrm.value, res.value, rq.value], zz])
        assert np.array_equal(got, row), f"dlar1v k={k} r={rin}"


def test_dlarrb_bit_exact(oracle_lib):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_cases.npz"))
    d, lld = z["k/d"], z["k/lld"]
    n = len(d)
    w, werr, wgap = (a.copy() for a in z["k/rrb_in"])
    f = oracle_lib.oracle_dlarrb
    f.restype = None
    f.argtypes = ([ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                   ctypes.c_double, ctypes.c_int, P, P, P, ctypes.c_double, ctypes.c_double,
                   ctypes.c_int])
    f(n, ptr(d), ptr(lld), 1, n, 1e-8, 1e-11, 0, ptr(w), ptr(wgap), ptr(werr),
      np.finfo(float).tiny, 3.0, n)
    wo, weo, wgo = z["k/rrb_out"]
    assert np.array_equal(w, wo) and np.array_equal(werr, weo) and np.array_equal(wgap, wgo)


def test_dlarrf_bit_exact(oracle_lib):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_cases.npz"))
    d, l, ld, ev = z["k/d"], z["k/l"], z["k/ld"], z["k/ev"]
    n = len(d)
    wc, gc, ec = ev[10:14].copy(), np.full(4, 1e-3), np.full(4, 1e-12)
    dplus, lplus, work = np.zeros(n), np.zeros(n), np.zeros(2 * n)
    sig = ctypes.c_double()
    f = oracle_lib.oracle_dlarrf
    f.restype = ctypes.c_int
    f.argtypes = ([ctypes.c_int, P, P, P, ctypes.c_int, ctypes.c_int, P, P, P] +
                  [ctypes.c_double] * 4 + [P] * 4)
    info = f(n, ptr(d), ptr(l), ptr(ld), 1, 4, ptr(wc), ptr(gc), ptr(ec), 3.0, 2e-3, 5e-3,
             np.finfo(float).tiny, ctypes.byref(sig), ptr(dplus), ptr(lplus), ptr(work))
    got = np.concatenate([[sig.value, info], dplus, lplus[: n - 1]])
    assert np.array_equal(got, z["k/rrf_out"])


def test_dlarnv_bit_exact(oracle_lib):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_cases.npz"))
    seed = (ctypes.c_int * 4)(1, 1, 1, 1)
    x = np.zeros(300)
    oracle_lib.oracle_dlarnv2(seed, 300, ptr(x))
    assert np.array_equal(x, z["k/rnv300"])
    assert list(seed) == list(z["k/rnv_seed_after"])


# ---------------------------------------------------------------------------------
# what the mathematics fixes
# ---------------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [(1, 1, 1, 1), (4095, 1, 77, 3001), (12, 345, 678, 911)])
def test_dlarnv_is_the_48bit_lcg(oracle_lib, seed):
    """dlaruv's 128 multipliers (odruv.f) are the powers a^i mod 2^48 of
    a = 33952834046453: entry i of the stream is seed * a^(i+1) mod 2^48 scaled to
    (0,1), and dlarnv IDIST=2 maps u to 2u - 1.  700 entries cross five 128-blocks."""
    a, mod = 33952834046453, 1 << 48
    x = ((seed[0] * 4096 + seed[1]) * 4096 + seed[2]) * 4096 + seed[3]
    want = []
    for _ in range(700):
        x = (x * a) % mod
        want.append(2.0 * (x / float(mod)) - 1.0)
    iseed = (ctypes.c_int * 4)(*seed)
    got = np.zeros(700)
    oracle_lib.oracle_dlarnv2(iseed, 700, ptr(got))
    assert np.array_equal(got, np.array(want))
    assert list(iseed) == [(x >> 36) & 4095, (x >> 24) & 4095, (x >> 12) & 4095, x & 4095]


@pytest.mark.parametrize("n", [3, 17, 64, 250])
def test_toeplitz_closed_form(n, oracle):
    _, fn = oracle
    D, E = np.full(n, 2.0), np.ones(n)
    got = call_mrrr(fn, "V", "A", D, E)
    k = np.arange(1, n + 1)
    lam = 2.0 - 2.0 * np.cos(k * np.pi / (n + 1))
    assert got["m"] == n
    assert np.abs(got["W"] - lam).max() <= w_tol(D, E)
    # for off-diagonals +1 the eigenvector of lambda_k is (-1)^j sin(j k pi / (n+1))
    j = np.arange(1, n + 1)[:, None]
    V = (-1.0) ** j * np.sin(j * k[None, :] * np.pi / (n + 1))
    V /= np.linalg.norm(V, axis=0)
    # eigenvectors are unique up to sign for a simple spectrum
    dots = np.abs(np.sum(V * got["Z"], axis=0))
    assert np.all(dots > 1 - 1e3 * n * EPS)
    check_support(got["Z"], got["Zsupp"])


@pytest.mark.parametrize("seed", range(6))
def test_bruteforce_small(seed, oracle):
    _, fn = oracle
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(3, 65))
    D, E = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    lam = np.linalg.eigvalsh(dense(D, E))
    got = call_mrrr(fn, "V", "A", D, E)
    assert got["m"] == n and np.abs(got["W"] - lam).max() <= w_tol(D, E)
    r, o = resid_orth(D, E, got["W"], got["Z"])
    assert r < 10 and o < 100   # units of n eps
    # range V counts equal the dense count when the ends are away from eigenvalues
    vl, vu = lam[n // 4] + 1e-6, lam[3 * n // 4] + 1e-6
    if np.min(np.abs(lam - vl)) > 1e-9 and np.min(np.abs(lam - vu)) > 1e-9:
        gv = call_mrrr(fn, "V", "V", D, E, vl=vl, vu=vu)
        assert gv["m"] == int(np.sum((lam > vl) & (lam <= vu)))
        cv = call_mrrr(fn, "C", "V", D, E, vl=vl, vu=vu)
        assert cv["m"] == gv["m"]
    # range I returns exactly the requested indices
    il, iu = 2, n - 1
    gi = call_mrrr(fn, "V", "I", D, E, il=il, iu=iu)
    assert gi["m"] == iu - il + 1
    assert np.abs(gi["W"] - lam[il - 1:iu]).max() <= w_tol(D, E)


def test_sturm_count_inertia(oracle_lib):
    """Sturm count on T equals the number of eigenvalues below x (Sylvester)."""
    f = oracle_lib.oracle_sturm_count_T
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, P, P, ctypes.c_double, ctypes.c_double]
    rng = np.random.default_rng(3)
    for n in (5, 40, 200):
        D, E = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        lam = np.linalg.eigvalsh(dense(D, E))
        e2 = E ** 2
        for x in rng.uniform(-3, 3, 25):
            if np.min(np.abs(lam - x)) < 1e-8:
                continue
            assert f(n, ptr(D), ptr(e2), float(x), np.finfo(float).tiny) == int(np.sum(lam < x))


def test_dlaneg_inertia_all_twists(oracle_lib):
    """dlaneg gives the inertia of L D L^T - sigma I for every twist index."""
    f = oracle_lib.oracle_dlaneg
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    rng = np.random.default_rng(4)
    n = 30
    d = rng.uniform(0.5, 2, n) * rng.choice([-1, 1], n)
    l = rng.uniform(-1, 1, n)
    Lm = np.eye(n) + np.diag(l[: n - 1], -1)
    lam = np.linalg.eigvalsh(Lm @ np.diag(d) @ Lm.T)
    lld = d * l * l
    for s in rng.uniform(lam[0] - 1, lam[-1] + 1, 20):
        if np.min(np.abs(lam - s)) < 1e-7:
            continue
        want = int(np.sum(lam < s))
        for r in range(1, n + 1):
            assert f(n, ptr(d), ptr(lld), float(s), 0.0, r) == want


def test_two_by_two_and_one(oracle):
    _, fn = oracle
    D, E = np.array([0.7, -0.2]), np.array([0.3, 0.0])
    got = call_mrrr(fn, "V", "A", D, E)
    assert got["m"] == 2
    assert np.abs(got["W"] - np.linalg.eigvalsh(dense(D, E))).max() <= w_tol(D, E)
    assert np.abs(got["Z"].T @ got["Z"] - np.eye(2)).max() < 4 * EPS
    one = call_mrrr(fn, "V", "A", np.array([3.0]), np.array([0.0]))
    assert one["m"] == 1 and one["W"][0] == 3.0 and one["Z"][0, 0] == 1.0


def test_split_matrix_blocks(oracle):
    _, fn = oracle
    rng = np.random.default_rng(9)
    n = 50
    D, E = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    E[[0, 7, 8, 20, 48]] = 0.0
    lam = np.linalg.eigvalsh(dense(D, E))
    got = call_mrrr(fn, "V", "A", D, E)
    assert got["m"] == n and np.abs(got["W"] - lam).max() <= w_tol(D, E)
    r, o = resid_orth(D, E, got["W"], got["Z"])
    assert r < 10 and o < 100   # units of n eps


def test_invalid_arguments(oracle):
    _, fn = oracle
    D, E = np.ones(4), np.ones(4)
    assert call_mrrr(fn, "X", "A", D, E)["info"] == 1
    assert call_mrrr(fn, "V", "Q", D, E)["info"] == 1
    assert call_mrrr(fn, "V", "V", D, E, vl=1.0, vu=1.0)["info"] == 1
    assert call_mrrr(fn, "V", "I", D, E, il=0, iu=2)["info"] == 1
    assert call_mrrr(fn, "V", "I", D, E, il=3, iu=2)["info"] == 1
    assert call_mrrr(fn, "V", "I", D, E, il=1, iu=5)["info"] == 1
    assert call_mrrr(fn, "C", "A", D, E)["m"] == 4
    assert call_mrrr(fn, "C", "I", D, E, il=2, iu=3)["m"] == 2


# ---------------------------------------------------------------------------------
# the compiled reference itself (only where oracle/_ref is built)
# ---------------------------------------------------------------------------------
needs_ref = pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("kind,n,threads", [("uniform", 150, 1), ("uniform", 700, 16),
                                            ("toeplitz", 400, 1), ("glued", 630, 16),
                                            ("normal", 520, 16), ("uniform", 640, 1)])
def test_oracle_vs_compiled_reference(kind, n, threads, oracle, oracle_lib):
    """The dstemr route (n < 500, or one thread) bit for bit; the task-parallel route
    (>= 13 threads: bisection for the root eigenvalues, mrrr_val.c:281-282) to the
    parity bars."""
    from mr3smp_amd import workloads as WL
    rng = np.random.default_rng(n)
    if kind == "uniform":
        D, E = WL.uniform(n, rng)
    elif kind == "toeplitz":
        D, E = WL.toeplitz_pert(n, rng)
    elif kind == "glued":
        D, E = WL.glued_wilkinson(n, rng, 1e-9)
    else:
        D, E = WL.normal(n, rng)
    _, ref = load(REF_SO, "mrrr")
    old = os.environ.get("PMR_NUM_THREADS")
    os.environ["PMR_NUM_THREADS"] = str(threads)
    try:
        for jobz, rg, kw in (("V", "A", {}), ("N", "A", {}), ("V", "I", dict(il=n // 3, iu=n // 3 + n // 10))):
            r = call_mrrr(ref, jobz, rg, D, E, **kw)
            label = f"{kind}{n}-{jobz}{rg}"
            if n < 500:
                g = call_mrrr(oracle[1], jobz, rg, D, E, **kw)
                assert g["info"] == r["info"] == 0 and same_bits(g, r, jobz == "V"), label
            elif threads == 1:
                g = call_dstemr(oracle_lib, jobz, rg, D, E, il=kw.get("il", 1), iu=kw.get("iu", 1))
                assert g["info"] == r["info"] == 0 and same_bits(g, r, jobz == "V"), label
            else:
                g = call_mrrr(oracle[1], jobz, rg, D, E, **kw)
                check_pair(D, E, g, r, label, vectors=(jobz == "V"))
    finally:
        if old is None:
            os.environ.pop("PMR_NUM_THREADS", None)
        else:
            os.environ["PMR_NUM_THREADS"] = old


# matrices of the batched mixed workload (n < 500) that exercise the dstemr route's
# clusters: glued W21+ copies with tight pairs (37.6-42.9 n eps orthogonality in the
# reference's task-parallel algorithm, 1.9-3.9 in dstemr), uniform and 1-2-1 ones
MIXED_SMALL = [937, 3716, 1668, 3883, 0, 1, 2, 3, 5, 8, 13, 21, 34, 55]


@needs_ref
@pytest.mark.parametrize("idx", MIXED_SMALL)
def test_dstemr_route_bit_exact_on_mixed_workload(idx, oracle):
    from mr3smp_amd import workloads as WL
    D, E, kind = WL.batched_mixed(4096)[idx]
    n = len(D)
    if n >= 500:
        pytest.skip("n >= 500")
    _, ref = load(REF_SO, "mrrr")
    old = os.environ.pop("PMR_NUM_THREADS", None)
    try:
        Wd = np.linalg.eigvalsh(np.diag(D) + np.diag(E[: n - 1], 1) + np.diag(E[: n - 1], -1))
        vl, vu = float(Wd[n // 4]) - 1e-3, float(Wd[n // 2]) + 1e-3
        calls = (("V", "A", {}), ("N", "A", {}), ("V", "I", dict(il=n // 5 + 1, iu=n // 2)),
                 ("V", "V", dict(vl=vl, vu=vu)), ("N", "V", dict(vl=vl, vu=vu)),
                 ("V", "A", dict(tryrac=0)))
        for jobz, rg, kw in calls:
            r = call_mrrr(ref, jobz, rg, D, E, **kw)
            g = call_mrrr(oracle[1], jobz, rg, D, E, **kw)
            assert g["info"] == r["info"] == 0, (idx, jobz, rg)
            assert same_bits(g, r, jobz == "V"), (idx, kind, jobz, rg)
            for k in ("vl", "vu", "il", "iu", "tryrac"):
                assert g[k] == r[k], (idx, jobz, rg, k)
        c = call_mrrr(oracle[1], "C", "V", D, E, vl=vl, vu=vu)
        cr = call_mrrr(ref, "C", "V", D, E, vl=vl, vu=vu)
        assert c["m"] == cr["m"]
    finally:
        if old is not None:
            os.environ["PMR_NUM_THREADS"] = old
