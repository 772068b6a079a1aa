"""CPU tests of the GPU library's small-matrix route (mr3smp_amd/csrc/mr3_small_core.h).

The route restates the reference's dstemr (odstmr.f, odrre.f with dqds, odrrv.f), which
mrrr() takes for every n < 500 (SRC/mrrr.c:149-153).  Its source is written for one
team of threads (a wave on the GPU) with barriers between phases; compiled with g++
(tests/small_route_cpu.cpp) it runs as a team of one thread, every phase in order, or
as 64 POSIX threads with a barrier per synchronisation, as the device's wave does.  So
the device algorithm and its work split can be checked bit for bit against the oracle's
restatement (which is pinned against the compiled reference in test_oracle.py) without
a GPU.  The harness poisons its work memory, as the device's starts uninitialised.
"""
import os
import subprocess

import numpy as np
import pytest

from mrrr_ctypes import call_mrrr, load, same_bits

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="session")
def small_cpu(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("smallcpu") / "libsmall_cpu.so")
    subprocess.run(["g++", "-O2", "-fPIC", "-ffp-contract=off", "-shared", "-pthread", "-o", out,
                    os.path.join(HERE, "small_route_cpu.cpp")], check=True)
    return load(out, "small_route_mrrr")


def _calls(D, E, rng):
    n = len(D)
    il = int(rng.integers(1, n + 1))
    iu = int(rng.integers(il, n + 1))
    Wd = np.linalg.eigvalsh(np.diag(D) + np.diag(E[: n - 1], 1) + np.diag(E[: n - 1], -1))
    vl, vu = sorted(rng.uniform(Wd[0] - 0.1, Wd[-1] + 0.1, 2))
    return (("V", "A", {}), ("N", "A", {}), ("V", "I", dict(il=il, iu=iu)),
            ("N", "I", dict(il=il, iu=iu)), ("V", "V", dict(vl=vl, vu=vu)),
            ("N", "V", dict(vl=vl, vu=vu)), ("C", "V", dict(vl=vl, vu=vu)),
            ("V", "A", dict(tryrac=0)))


def _check(small_cpu, oracle, D, E, label, rng, team=1, staged=0):
    lib, fn = small_cpu
    lib.small_route_set_team(team)
    lib.small_route_set_staged(staged)
    for jobz, rg, kw in _calls(D, E, rng):
        g = call_mrrr(fn, jobz, rg, D, E, **kw)
        o = call_mrrr(oracle[1], jobz, rg, D, E, **kw)
        assert g["info"] == o["info"] == 0, (label, jobz, rg)
        assert g["m"] == o["m"], (label, jobz, rg, g["m"], o["m"])
        if jobz == "C":
            continue
        assert same_bits(g, o, jobz == "V"), (label, jobz, rg)
        assert g["tryrac"] == o["tryrac"], (label, jobz, rg)


def test_small_route_mixed_workload(small_cpu, oracle):
    """Every twelfth small matrix of the batched mixed workload (about 225 matrices of
    the three kinds), eight calls each, bit for bit."""
    from mr3smp_amd import workloads as WL
    mats = WL.batched_mixed(4096)
    idx = [i for i, (D, _, _) in enumerate(mats) if len(D) < 500][::12]
    for i in idx:
        D, E, kind = mats[i]
        _check(small_cpu, oracle, D, E, f"{i} {kind} n={len(D)}", np.random.default_rng(i))


def test_small_route_mixed_workload_team64(small_cpu, oracle):
    """Every 96th small matrix of the mixed workload run by a team of 64 threads."""
    from mr3smp_amd import workloads as WL
    mats = WL.batched_mixed(4096)
    idx = [i for i, (D, _, _) in enumerate(mats) if len(D) < 500][5::96]
    for i in idx:
        D, E, kind = mats[i]
        _check(small_cpu, oracle, D, E, f"{i} {kind} n={len(D)}", np.random.default_rng(i), team=64)


def test_small_route_mixed_workload_team256(small_cpu, oracle):
    """The largest small matrices of the mixed workload (n >= 400, every 40th) run by a
    team of 256 threads: the work split of any team size gives the same bits."""
    from mr3smp_amd import workloads as WL
    mats = WL.batched_mixed(4096)
    idx = [i for i, (D, _, _) in enumerate(mats) if 400 <= len(D) < 500][::40]
    for i in idx:
        D, E, kind = mats[i]
        _check(small_cpu, oracle, D, E, f"{i} {kind} n={len(D)}", np.random.default_rng(i), team=256)


def test_small_route_staged_team64(small_cpu, oracle):
    """The two-stage form: the first stage defers every block's dqds (dlarre's
    root-finding of the block's representation), the team's LDS is lost, each deferred
    dqds runs on its own and the second stage resumes from the saved state -- the same
    bits as the one-launch form, on every 48th small matrix of the mixed workload."""
    from mr3smp_amd import workloads as WL
    mats = WL.batched_mixed(4096)
    idx = [i for i, (D, _, _) in enumerate(mats) if len(D) < 500][7::48]
    for i in idx:
        D, E, kind = mats[i]
        _check(small_cpu, oracle, D, E, f"{i} {kind} n={len(D)}", np.random.default_rng(i),
               team=64, staged=1)


@pytest.mark.parametrize("staged", [0, 1])
@pytest.mark.parametrize("team", [1, 64, 256])
@pytest.mark.parametrize("name", ["n1", "n2", "n3", "diag", "diag8", "diag16", "zerosD", "split",
                                  "tiny", "huge", "repeated", "relacc", "toeplitz499", "glued499"])
def test_small_route_edge_cases(name, team, staged, small_cpu, oracle):
    from mr3smp_amd import workloads as WL
    rng = np.random.default_rng(5)
    cases = {
        "n1": (np.array([3.5]), np.array([0.0])),
        "n2": (np.array([1.0, 2.0]), np.array([0.5, 0.0])),
        "n3": (rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)),
        "diag": (rng.uniform(-1, 1, 40), np.zeros(40)),
        "diag8": (np.linspace(1, -1, 8), np.zeros(8)),
        "diag16": (np.linspace(1, -1, 16), np.zeros(16)),
        "zerosD": (np.zeros(30), np.ones(30)),
        "split": (rng.uniform(-1, 1, 80), np.where(np.arange(80) % 9 == 0, 0.0, rng.uniform(-1, 1, 80))),
        "tiny": (1e-300 * rng.uniform(-1, 1, 25), 1e-300 * rng.uniform(-1, 1, 25)),
        "huge": (1e300 * rng.uniform(-1, 1, 25), 1e300 * rng.uniform(-1, 1, 25)),
        "repeated": (np.tile([1.0, 2.0, 3.0], 20), np.where(np.arange(60) % 3 == 2, 0.0, 1e-3)),
        "relacc": (4.0 + rng.uniform(-1, 1, 120), rng.uniform(-1, 1, 120)),
        "toeplitz499": WL.toeplitz_pert(499, rng),
        "glued499": WL.glued_wilkinson(499, rng, 1e-10),
    }
    D, E = cases[name]
    E = E.copy()
    E[-1] = 0.0
    _check(small_cpu, oracle, D, E, name, np.random.default_rng(len(D)), team=team, staged=staged)
