"""CPU checks of the product library: it is built, loads, and exports every entry
point declared in include/mr3hip.h (no compute call: there is no GPU here)."""
import ctypes
import os
import re

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(REPO, "include", "mr3hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_header_declares_the_dropin():
    names = declared_functions()
    assert "mrrr" in names and "mrrr_" in names and "mr3_solve_batch" in names


def test_library_exports_every_declared_symbol():
    import mr3smp_amd
    assert os.path.exists(mr3smp_amd.LIB_PATH), "libmr3hip.so not built"
    lib = ctypes.CDLL(mr3smp_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    import mr3smp_amd
    blob = open(mr3smp_amd.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_workload_generators():
    from mr3smp_amd import workloads as WL
    D, E = WL.single_glued_20000()
    assert len(D) == 20000 and E[-1] == 0.0
    assert np.count_nonzero(D[:21] - np.abs(10 - np.arange(21.0))) == 0
    glue = E[20:952 * 21:21]
    assert np.all((glue >= 0.5e-8) & (glue <= 1e-8))
    assert np.all(D[952 * 21:] == 2.0)
    mats = WL.batched_mixed(64)
    ns = [len(m[0]) for m in mats]
    assert min(ns) >= 32 and max(ns) <= 2048
    c4 = WL.batched_range_i(4)
    for D, E, il, iu in c4:
        n = len(D)
        assert 4096 <= n <= 50000 and iu - il + 1 == int(np.ceil(0.1 * n)) and 1 <= il <= iu <= n
    reqs = WL.served_stream(50, 10.0)
    assert all(np.diff([r["t"] for r in reqs]) >= 0)
    assert {r["range"] for r in reqs} <= {"A", "I", "V"}


def test_dense_drivers_reject_invalid_arguments():
    """dsyeig / dsygeig return 1 for the argument errors of dense.c:174-180 and
    :256-260 before touching the device (so this runs without a GPU)."""
    import mr3smp_amd
    L = mr3smp_amd.lib()
    P = ctypes.byref
    A = np.eye(4)
    W = np.zeros(4)
    Z = np.zeros((4, 4))
    a = A.ctypes.data_as(ctypes.c_void_p)

    def syeig(jobz, rng, uplo, n=4, lda=4, vl=0.0, vu=1.0, il=1, iu=4, ldz=4):
        args = [ctypes.c_int(n), ctypes.c_int(lda), ctypes.c_double(vl), ctypes.c_double(vu),
                ctypes.c_int(il), ctypes.c_int(iu), ctypes.c_int(0), ctypes.c_int(ldz)]
        return L.dsyeig(jobz, rng, uplo, P(args[0]), a, P(args[1]), P(args[2]), P(args[3]),
                        P(args[4]), P(args[5]), P(args[6]), W.ctypes.data_as(ctypes.c_void_p),
                        Z.ctypes.data_as(ctypes.c_void_p), P(args[7]))

    assert syeig(b"X", b"A", b"L") == 1
    assert syeig(b"V", b"X", b"L") == 1
    assert syeig(b"V", b"A", b"X") == 1
    assert syeig(b"V", b"A", b"L", n=0) == 1
    assert syeig(b"V", b"A", b"L", lda=3) == 1
    assert syeig(b"V", b"A", b"L", ldz=3) == 1
    assert syeig(b"V", b"V", b"L", vl=1.0, vu=1.0) == 1
    assert syeig(b"V", b"I", b"U", il=0) == 1
    assert syeig(b"V", b"I", b"U", il=3, iu=2) == 1
    assert syeig(b"V", b"I", b"U", il=1, iu=5) == 1

    def sygeig(itype, jobz, rng, uplo, n=4):
        args = [ctypes.c_int(itype), ctypes.c_int(n), ctypes.c_int(4), ctypes.c_int(4),
                ctypes.c_double(0.0), ctypes.c_double(1.0), ctypes.c_int(1), ctypes.c_int(4),
                ctypes.c_int(0)]
        B = np.eye(4)
        return L.dsygeig(P(args[0]), jobz, rng, uplo, P(args[1]), a, P(args[2]),
                         B.ctypes.data_as(ctypes.c_void_p), P(args[3]), P(args[4]), P(args[5]),
                         P(args[6]), P(args[7]), P(args[8]), W.ctypes.data_as(ctypes.c_void_p))

    assert sygeig(1, b"V", b"A", b"L", n=1) == 1
    assert sygeig(0, b"V", b"A", b"L") == 1
    assert sygeig(4, b"V", b"A", b"L") == 1
    assert sygeig(1, b"Q", b"A", b"L") == 1
    assert sygeig(2, b"V", b"Q", b"L") == 1
    assert sygeig(3, b"V", b"A", b"Q") == 1


def test_packed_and_complex_drivers_reject_invalid_arguments():
    """The packed and complex dense drivers share the argument checks (dense.c:341-347,
    :422-426, :596-602, :697-701, :788-794, :888-892) and return 1 without a device."""
    import mr3smp_amd
    L = mr3smp_amd.lib()
    P = ctypes.byref
    buf = np.zeros(64, dtype=np.complex128)
    W = np.zeros(4)
    b = buf.ctypes.data_as(ctypes.c_void_p)
    w = W.ctypes.data_as(ctypes.c_void_p)

    def scal():
        return [ctypes.c_double(0.0), ctypes.c_double(1.0), ctypes.c_int(1), ctypes.c_int(4),
                ctypes.c_int(0), ctypes.c_int(4)]

    for name in ("dspeig", "zhpeig"):
        for jobz, rng, uplo, n in [(b"X", b"A", b"L", 4), (b"V", b"X", b"L", 4), (b"V", b"A", b"X", 4),
                                   (b"V", b"A", b"L", 0)]:
            vl, vu, il, iu, m, ldz = scal()
            nn = ctypes.c_int(n)
            assert getattr(L, name)(jobz, rng, uplo, P(nn), b, P(vl), P(vu), P(il), P(iu), P(m), w, b,
                                    P(ldz)) == 1, (name, jobz, rng, uplo, n)
    for name in ("zheeig",):
        vl, vu, il, iu, m, ldz = scal()
        nn, lda = ctypes.c_int(4), ctypes.c_int(4)
        il.value = 3
        iu.value = 2
        assert L.zheeig(b"V", b"I", b"U", P(nn), b, P(lda), P(vl), P(vu), P(il), P(iu), P(m), w, b,
                        P(ldz)) == 1
    for name in ("zhegeig",):
        for itype, n in [(0, 4), (4, 4), (1, 1)]:
            vl, vu, il, iu, m, ldz = scal()
            it, nn, ld = ctypes.c_int(itype), ctypes.c_int(n), ctypes.c_int(4)
            assert L.zhegeig(P(it), b"V", b"A", b"L", P(nn), b, P(ld), b, P(ld), P(vl), P(vu), P(il),
                             P(iu), P(m), w) == 1
    for name, small in (("dspgeig", 0), ("zhpgeig", 1)):
        for itype, n in [(0, 4), (4, 4), (1, small)]:
            vl, vu, il, iu, m, ldz = scal()
            it, nn = ctypes.c_int(itype), ctypes.c_int(n)
            assert getattr(L, name)(P(it), b"V", b"A", b"L", P(nn), b, b, P(vl), P(vu), P(il), P(iu),
                                    P(m), w, b, P(ldz)) == 1, (name, itype, n)


def test_tuning_parameters_round_trip_and_reject_bad_values():
    """mr3_set_tuning / mr3_get_tuning are host-only: values round-trip, out-of-range
    values are rejected without changing anything."""
    import pytest
    import mr3smp_amd as M
    before = M.get_tuning()
    try:
        got = M.set_tuning(tail_singletons=4096, group_lanes=131072)
        assert got["tail_singletons"] == 4096 and M.get_tuning()["group_lanes"] == 131072
        for bad in (dict(side_streams=4), dict(group_lanes=100), dict(singleton_waves=1),
                    dict(tail_singletons=-1)):
            with pytest.raises(ValueError):
                M.set_tuning(**bad)
        assert M.get_tuning()["tail_singletons"] == 4096
        with pytest.raises(ValueError):
            M.set_tuning(no_such_knob=1)
    finally:
        M.set_tuning(**before)
    assert M.get_tuning() == before
