"""GPU tests of the dense drivers (SRC/dense.c:151-1114): dsyeig, dspeig, dsygeig,
dspgeig, their complex Hermitian twins zheeig, zhpeig, zhegeig, zhpgeig, and the
device-array variant of dsyeig.

The reference reduces with LAPACK's dsytrd/dormtr and dpotrf/dsygst/dtrsm/dtrmm,
which it links from the system and does not ship; here rocSOLVER and rocBLAS do
those steps, so the dense results agree with the reference to rounding, not bit for
bit.  The bars: eigenvalues within 4 n eps ||A|| of LAPACK's (numpy/scipy eigh),
residual ||A Z - Z W|| and orthogonality ||Z^H Z - I|| within 20 n eps (scaled by
||A||), counts exact.  Packed storage is unpacked on the device into the same
full-storage reduction, so every packed driver must agree bit for bit with its
full-storage twin.  The composition itself is pinned bit for bit on tridiagonal
input: there every Householder vector is zero, so dsyeig must return exactly what
the tridiagonal solver returns for the same D and E.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def M():
    import mr3smp_amd
    return mr3smp_amd


def sym(n, seed, spread=1.0, cplx=False):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) * spread
    if cplx:
        X = X + 1j * rng.standard_normal((n, n)) * spread
    return (X + X.conj().T) / 2


def spd(n, seed, cplx=False):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n))
    if cplx:
        X = X + 1j * rng.standard_normal((n, n))
    return X @ X.conj().T / n + np.eye(n)


def only_triangle(A, uplo):
    """A with the triangle the routine must not read overwritten by garbage."""
    B = A.copy()
    n = len(A)
    junk = np.full((n, n), np.nan, dtype=A.dtype)
    if uplo == "L":
        B[np.triu_indices(n, 1)] = junk[np.triu_indices(n, 1)]
    else:
        B[np.tril_indices(n, -1)] = junk[np.tril_indices(n, -1)]
    return B


def check_standard(A, out, Wref, label):
    n = len(A)
    nrm = max(np.linalg.norm(A, 2), np.finfo(np.float64).tiny)
    W, Z = out["W"], out["Z"]
    assert out["m"] == len(Wref), label
    if n == 2:
        # n = 2 goes through dstemr's 2x2 case (mrrr.c:150), which lists the
        # eigenvalue of smaller magnitude first
        order = np.argsort(W, kind="stable")
        W = W[order]
        Z = Z[:, order] if Z is not None else None
    assert np.all(np.diff(W) >= 0), label
    assert np.max(np.abs(W - Wref), initial=0.0) <= 4 * n * EPS * nrm, label
    if Z is not None and out["m"] > 0:
        R = A @ Z - Z * W
        res = np.max(np.abs(R)) / (n * EPS * nrm)
        orth = np.max(np.abs(Z.conj().T @ Z - np.eye(out["m"]))) / (n * EPS)
        assert res <= 20 and orth <= 20, (label, res, orth)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n", [1, 2, 17, 300, 1500])
def test_dsyeig_all(M, n, uplo):
    print(f"dsyeig A n={n} uplo={uplo}", flush=True)
    A = sym(n, 100 + n)
    out = M.dsyeig("V", "A", uplo, only_triangle(A, uplo))
    assert out["info"] == 0
    check_standard(A, out, np.linalg.eigvalsh(A), f"n={n} {uplo}")
    # values only: bisection values, not the Rayleigh-quotient refined ones of jobz V
    o2 = M.dsyeig("N", "A", uplo, only_triangle(A, uplo))
    assert o2["info"] == 0 and o2["m"] == n
    check_standard(A, dict(o2, Z=None), np.linalg.eigvalsh(A), f"N n={n} {uplo}")


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dsyeig_index_and_value_ranges(M, uplo):
    n = 800
    A = sym(n, 7)
    Wr = np.linalg.eigvalsh(A)
    print(f"dsyeig I/V n={n} uplo={uplo}", flush=True)
    for il, iu in [(1, 1), (1, 80), (391, 410), (n, n), (1, n)]:
        out = M.dsyeig("V", "I", uplo, only_triangle(A, uplo), il=il, iu=iu)
        assert out["info"] == 0
        check_standard(A, out, Wr[il - 1:iu], f"I {il}:{iu}")
    # value ranges with ends half-way between neighbouring eigenvalues
    for a, b in [(10, 90), (0, 400), (650, 799)]:
        vl = (Wr[a - 1] + Wr[a]) / 2 if a > 0 else Wr[0] - 1.0
        vu = (Wr[b] + Wr[b + 1]) / 2 if b + 1 < n else Wr[-1] + 1.0
        out = M.dsyeig("V", "V", uplo, only_triangle(A, uplo), vl=vl, vu=vu)
        assert out["info"] == 0
        check_standard(A, out, Wr[(Wr > vl) & (Wr <= vu)], f"V {a}:{b}")
        cnt = M.dsyeig("C", "V", uplo, only_triangle(A, uplo), vl=vl, vu=vu)
        assert cnt["info"] == 0 and cnt["m"] == np.count_nonzero((Wr > vl) & (Wr <= vu))
    cnt = M.dsyeig("C", "I", uplo, A, il=5, iu=44)
    assert cnt["info"] == 0 and cnt["m"] == 40


@pytest.mark.parametrize("s", [1e-160, 1e160])
def test_dsyeig_scales_into_the_safe_range(M, s):
    """Norms below sqrt(DBL_MIN/eps) or above the upper limit are scaled
    (dense.c:476-526) and the eigenvalues and interval scaled back."""
    n = 120
    A0 = sym(n, 3)
    Wr = np.linalg.eigvalsh(A0)
    A = A0 * s
    print(f"dsyeig scaled s={s:g}", flush=True)
    out = M.dsyeig("V", "A", "L", A)
    assert out["info"] == 0 and out["m"] == n
    nrm = np.linalg.norm(A0, 2)
    assert np.max(np.abs(out["W"] / s - Wr)) <= 4 * n * EPS * nrm
    Z = out["Z"]
    res = np.max(np.abs(A0 @ Z - Z * (out["W"] / s))) / (n * EPS * nrm)
    assert res <= 20
    vl, vu = (Wr[9] + Wr[10]) / 2 * s, (Wr[59] + Wr[60]) / 2 * s
    out = M.dsyeig("V", "V", "U", A, vl=vl, vu=vu)
    assert out["info"] == 0 and out["m"] == 50
    assert np.max(np.abs(out["W"] / s - Wr[10:60])) <= 4 * n * EPS * nrm
    assert abs(out["vl"] - vl) <= 4 * EPS * abs(vl) and abs(out["vu"] - vu) <= 4 * EPS * abs(vu)


def test_dsyeig_on_tridiagonal_input_is_the_tridiagonal_solver(M):
    """A tridiagonal A makes every Householder reflector the identity, so the dense
    route must hand the solver exactly D and E and return its W and Z unchanged."""
    from mr3smp_amd import workloads as WL
    rng = WL.rng_for(40)
    cases = [("glued", WL.glued_wilkinson(210, rng)), ("uniform", WL.uniform(300, rng)),
             ("toeplitz", WL.toeplitz_pert(257, rng))]
    for name, (D, E) in cases:
        n = len(D)
        A = np.diag(D) + np.diag(E[:n - 1], -1) + np.diag(E[:n - 1], 1)
        print(f"dsyeig tridiagonal {name} n={n}", flush=True)
        out = M.dsyeig("V", "A", "L", A)
        ref = M.mrrr("V", "A", D, E)
        assert out["info"] == 0 and ref["info"] == 0 and out["m"] == ref["m"] == n
        assert np.array_equal(out["W"], ref["W"]), name
        assert np.array_equal(np.abs(out["Z"]), np.abs(ref["Z"])), name


@pytest.mark.parametrize("itype", [1, 2, 3])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dsygeig(M, itype, uplo):
    import scipy.linalg as sl
    n = 400
    A, B = sym(n, 11), spd(n, 12)
    Wr = sl.eigh(A, B, type=itype, eigvals_only=True)
    print(f"dsygeig itype={itype} uplo={uplo} n={n}", flush=True)
    out = M.dsygeig(itype, "V", "A", uplo, only_triangle(A, uplo), only_triangle(B, uplo))
    assert out["info"] == 0 and out["m"] == n
    W, X = out["W"], out["X"]
    scale = np.linalg.norm(A, 2) * max(np.linalg.norm(B, 2), np.linalg.norm(np.linalg.inv(B), 2))
    assert np.max(np.abs(W - Wr)) <= 4 * n * EPS * scale * np.linalg.cond(B)
    if itype == 1:
        R = A @ X - (B @ X) * W
        G = X.T @ B @ X
    elif itype == 2:
        R = A @ (B @ X) - X * W
        G = X.T @ B @ X
    else:
        R = B @ (A @ X) - X * W
        G = X.T @ np.linalg.solve(B, X)
    xn = np.max(np.abs(X))
    assert np.max(np.abs(R)) <= 50 * n * EPS * scale * max(xn, 1.0)
    assert np.max(np.abs(G - np.eye(n))) <= 50 * n * EPS * np.linalg.cond(B)
    Lc = np.linalg.cholesky(B)
    if uplo == "L":
        assert np.allclose(np.tril(out["L"]), Lc, rtol=0, atol=1e-12 * np.abs(Lc).max())
    else:
        assert np.allclose(np.triu(out["L"]), Lc.T, rtol=0, atol=1e-12 * np.abs(Lc).max())
    # index subset and the values-only call
    o2 = M.dsygeig(itype, "N", "I", uplo, A, B, il=20, iu=60)
    assert o2["info"] == 0 and o2["m"] == 41
    assert np.max(np.abs(o2["W"] - Wr[19:60])) <= 4 * n * EPS * scale * np.linalg.cond(B)


def test_dsygeig_rejects_indefinite_B(M):
    n = 10
    A, B = sym(n, 1), -np.eye(n)
    out = M.dsygeig(1, "V", "A", "L", A, B)
    assert out["info"] == 2


def test_dsyeig_device_arrays(M):
    """The device-array entry point on torch-owned HBM, on a caller stream."""
    import torch
    n = 1000
    A0 = sym(n, 21)
    Wr = np.linalg.eigvalsh(A0)
    L = M.lib()
    dA = torch.tensor(A0, device="cuda:0")      # symmetric: row/column order agree
    dW = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    dZ = torch.zeros((n, n), dtype=torch.float64, device="cuda:0")
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    vl, vu = ctypes.c_double(0.0), ctypes.c_double(0.0)
    il, iu, m = ctypes.c_int(1), ctypes.c_int(n), ctypes.c_int(0)
    print(f"mr3_dsyeig_device n={n}", flush=True)
    info = L.mr3_dsyeig_device(b"V", b"A", b"L", n, ctypes.c_void_p(dA.data_ptr()), n,
                               ctypes.byref(vl), ctypes.byref(vu), ctypes.byref(il),
                               ctypes.byref(iu), ctypes.byref(m), ctypes.c_void_p(dW.data_ptr()),
                               ctypes.c_void_p(dZ.data_ptr()), n, n, ctypes.c_void_p(s.cuda_stream))
    assert info == 0 and m.value == n
    W = dW.cpu().numpy()
    Z = dZ.cpu().numpy().T       # torch row-major view of the column-major Z
    out = dict(m=n, W=W, Z=Z)
    check_standard(A0, out, Wr, "device")
    host = M.dsyeig("V", "A", "L", A0)
    assert np.array_equal(host["W"], W)


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_packed_drivers_equal_full_storage(M, uplo):
    """dspeig / zhpeig unpack AP into the full-storage reduction: bit-identical W and Z."""
    for cplx in (False, True):
        n = 257
        A = sym(n, 31, cplx=cplx)
        full, packed = (M.zheeig, M.zhpeig) if cplx else (M.dsyeig, M.dspeig)
        print(f"packed standard cplx={cplx} uplo={uplo}", flush=True)
        for kw in (dict(jobz="V", rng="A"), dict(jobz="V", rng="I", il=30, iu=90),
                   dict(jobz="N", rng="A")):
            Af = np.asfortranarray(only_triangle(A, uplo))
            AP = M.pack(A, uplo)
            Af0, AP0 = Af.copy(), AP.copy()
            a = full(kw["jobz"], kw["rng"], uplo, Af, il=kw.get("il", 1), iu=kw.get("iu", 1))
            b = packed(kw["jobz"], kw["rng"], uplo, AP, n, il=kw.get("il", 1), iu=kw.get("iu", 1))
            # the standard drivers leave A and AP as they were (the wrappers pass them
            # without a copy)
            assert np.array_equal(Af, Af0, equal_nan=True) and np.array_equal(AP, AP0, equal_nan=True), kw
            assert a["info"] == 0 and b["info"] == 0 and a["m"] == b["m"], kw
            assert np.array_equal(a["W"], b["W"]), kw
            if kw["jobz"] == "V":
                assert np.array_equal(a["Z"], b["Z"]), kw


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n", [1, 2, 33, 400])
def test_zheeig(M, n, uplo):
    print(f"zheeig n={n} uplo={uplo}", flush=True)
    A = sym(n, 200 + n, cplx=True)
    Wr = np.linalg.eigvalsh(A)
    out = M.zheeig("V", "A", uplo, only_triangle(A, uplo))
    assert out["info"] == 0 and out["Z"].dtype == np.complex128
    check_standard(A, out, Wr, f"zheeig n={n} {uplo}")
    if n >= 33:
        il, iu = n // 4, n // 2
        out = M.zheeig("V", "I", uplo, only_triangle(A, uplo), il=il, iu=iu)
        assert out["info"] == 0
        check_standard(A, out, Wr[il - 1:iu], f"zheeig I n={n} {uplo}")
        vl, vu = (Wr[4] + Wr[5]) / 2, (Wr[n - 6] + Wr[n - 5]) / 2
        out = M.zheeig("V", "V", uplo, only_triangle(A, uplo), vl=vl, vu=vu)
        assert out["info"] == 0
        check_standard(A, out, Wr[5:n - 5], f"zheeig V n={n} {uplo}")


def test_zheeig_scaled(M):
    n = 90
    A0 = sym(n, 5, cplx=True)
    Wr = np.linalg.eigvalsh(A0)
    for s in (1e-160, 1e160):
        out = M.zheeig("V", "A", "U", A0 * s)
        assert out["info"] == 0 and out["m"] == n
        nrm = np.linalg.norm(A0, 2)
        assert np.max(np.abs(out["W"] / s - Wr)) <= 4 * n * EPS * nrm
        Z = out["Z"]
        assert np.max(np.abs(A0 @ Z - Z * (out["W"] / s))) <= 20 * n * EPS * nrm


def check_general(itype, A, B, out, Wr, label):
    n = len(A)
    W, X = out["W"], out["X"]
    scale = np.linalg.norm(A, 2) * max(np.linalg.norm(B, 2), np.linalg.norm(np.linalg.inv(B), 2))
    cb = np.linalg.cond(B)
    assert out["m"] == len(Wr), label
    assert np.max(np.abs(W - Wr), initial=0.0) <= 4 * n * EPS * scale * cb, label
    if X is None:
        return
    if itype == 1:
        R, G = A @ X - (B @ X) * W, X.conj().T @ B @ X
    elif itype == 2:
        R, G = A @ (B @ X) - X * W, X.conj().T @ B @ X
    else:
        R, G = B @ (A @ X) - X * W, X.conj().T @ np.linalg.solve(B, X)
    xn = np.max(np.abs(X))
    assert np.max(np.abs(R)) <= 50 * n * EPS * scale * max(xn, 1.0), label
    assert np.max(np.abs(G - np.eye(out["m"]))) <= 50 * n * EPS * cb, label


@pytest.mark.parametrize("itype", [1, 2, 3])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_zhegeig_and_packed_generalized(M, itype, uplo):
    import scipy.linalg as sl
    n = 300
    for cplx in (True, False):
        A, B = sym(n, 41, cplx=cplx), spd(n, 42, cplx=cplx)
        Wr = sl.eigh(A, B, type=itype, eigvals_only=True)
        full, packed = (M.zhegeig, M.zhpgeig) if cplx else (M.dsygeig, M.dspgeig)
        print(f"generalized cplx={cplx} itype={itype} uplo={uplo}", flush=True)
        a = full(itype, "V", "A", uplo, only_triangle(A, uplo), only_triangle(B, uplo))
        assert a["info"] == 0
        check_general(itype, A, B, a, Wr, f"full cplx={cplx}")
        Lc = np.linalg.cholesky(B)
        tri = np.tril(a["L"]) if uplo == "L" else np.triu(a["L"])
        ref = Lc if uplo == "L" else Lc.conj().T
        assert np.max(np.abs(tri - ref)) <= 1e-12 * np.abs(Lc).max()
        b = packed(itype, "V", "A", uplo, M.pack(A, uplo), M.pack(B, uplo), n)
        assert b["info"] == 0 and b["m"] == n
        assert np.array_equal(a["W"], b["W"]) and np.array_equal(a["X"], b["X"])
        assert np.array_equal(b["L"], M.pack(a["L"], uplo))
        c = packed(itype, "N", "I", uplo, M.pack(A, uplo), M.pack(B, uplo), n, il=10, iu=50)
        assert c["info"] == 0 and c["m"] == 41
        check_general(itype, A, B, c, Wr[9:50], f"packed I cplx={cplx}")
