"""The drop-in boundary as a C program and a Fortran-style caller see it.

A C program (tests/c_caller.c, written against the mrrr() prototype of
include/mr3hip.h, which is the reference's INCLUDE/mrrr.h:62-65) is compiled with gcc
and linked with -lmr3hip, the way the reference's users link -lmrrr; it solves the
glued Wilkinson W21+ golden cases, whose outputs the compiled reference wrote (n < 500:
the reference's dstemr route, which the library restates bit for bit), and its W, Z
and Zsupp must equal them byte for byte.  The Fortran binding mrrr_ (SRC/mrrr.c:630-637:
every argument by reference, info as the last one) is called through ctypes on the same
cases."""
import ctypes
import os
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CASES = ["w21_VA", "w21_VI", "w21_VV", "w21_CV"]


@pytest.fixture(scope="module")
def c_caller(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("ccaller") / "c_caller")
    libdir = os.path.join(ROOT, "mr3smp_amd")
    subprocess.run(["gcc", "-O2", "-o", exe, os.path.join(HERE, "c_caller.c"),
                    "-I" + os.path.join(ROOT, "include"), "-L" + libdir, "-lmr3hip",
                    "-Wl,-rpath," + libdir], check=True)
    return exe


def _expect(c):
    info, m = (int(v) for v in c["info_m"])
    return info, m


@pytest.mark.parametrize("name", CASES)
def test_c_program_linked_with_the_library(name, golden, c_caller, tmp_path):
    c = golden[name]
    D, E = c["D"], c["E"]
    n = len(D)
    kw = c["kw"]
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(inp, "wb") as f:
        f.write(struct.pack("<icciiidd", n, c["jobz"].encode(), c["rng"].encode(), kw["il"], kw["iu"],
                            kw["tryrac"], kw["vl"], kw["vu"]))
        f.write(np.ascontiguousarray(D, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(E, dtype="<f8").tobytes())
    print(f"c caller {name} n={n}", flush=True)
    subprocess.run([c_caller, inp, out], check=True, timeout=120)
    raw = open(out, "rb").read()
    info, m, tryrac = struct.unpack_from("<iii", raw, 0)
    pos = 12 + 16
    info0, m0 = _expect(c)
    assert (info, m) == (info0, m0), name
    if c["jobz"] == "C" or m == 0:
        return
    W = np.frombuffer(raw, "<f8", m, pos)
    pos += 8 * m
    assert np.array_equal(W, c["W"]), name + ": W bits differ from the reference"
    if c["jobz"] == "V":
        Z = np.frombuffer(raw, "<f8", n * m, pos).reshape(m, n).T
        pos += 8 * n * m
        Zs = np.frombuffer(raw, "<i4", 2 * m, pos)
        assert np.array_equal(Z, c["Z"]), name + ": Z bits differ from the reference"
        assert np.array_equal(Zs, c["Zsupp"]), name + ": Zsupp differs from the reference"


@pytest.mark.parametrize("name", CASES)
def test_fortran_binding(name, golden, M):
    c = golden[name]
    D = np.ascontiguousarray(c["D"], dtype=np.float64).copy()
    E = np.ascontiguousarray(c["E"], dtype=np.float64).copy()
    n = len(D)
    kw = c["kw"]
    W = np.zeros(n)
    Z = np.zeros((n, n))   # column j of the math is row j here (ldz = n)
    Zs = np.zeros(2 * n, dtype=np.int32)
    ni, m, ldz, info = ctypes.c_int(n), ctypes.c_int(0), ctypes.c_int(n), ctypes.c_int(-7)
    vl, vu = ctypes.c_double(kw["vl"]), ctypes.c_double(kw["vu"])
    il, iu, tr = ctypes.c_int(kw["il"]), ctypes.c_int(kw["iu"]), ctypes.c_int(kw["tryrac"])
    P = ctypes.c_void_p
    f = M.lib().mrrr_
    f.restype = None
    f.argtypes = [ctypes.c_char_p, ctypes.c_char_p] + [P] * 14
    ptr = lambda a: a.ctypes.data_as(P)
    f(c["jobz"].encode(), c["rng"].encode(), ctypes.byref(ni), ptr(D), ptr(E), ctypes.byref(vl),
      ctypes.byref(vu), ctypes.byref(il), ctypes.byref(iu), ctypes.byref(tr), ctypes.byref(m), ptr(W),
      ptr(Z), ctypes.byref(ldz), ptr(Zs), ctypes.byref(info))
    info0, m0 = _expect(c)
    assert (info.value, m.value) == (info0, m0), name
    mm = m.value
    if c["jobz"] == "C" or mm == 0:
        return
    assert np.array_equal(W[:mm], c["W"]), name
    if c["jobz"] == "V":
        assert np.array_equal(Z[:mm].T, c["Z"]) and np.array_equal(Zs[:2 * mm], c["Zsupp"]), name


@pytest.fixture(scope="module")
def M():
    import mr3smp_amd
    return mr3smp_amd
