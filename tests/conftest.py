import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no HIP device")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def golden():
    z = np.load(os.path.join(HERE, "golden", "ref_cases.npz"), allow_pickle=False)
    cases = {}
    for name in z["names"]:
        name = str(name)
        c = {}
        for k in ("D", "E", "params", "vlvu", "info_m", "W", "resid_orth", "Zsupp", "Z"):
            key = f"{name}/{k}"
            if key in z.files:
                c[k] = z[key]
        p = c["params"]
        c["jobz"], c["rng"] = chr(p[0]), chr(p[1])
        c["kw"] = dict(il=int(p[2]), iu=int(p[3]), tryrac=int(p[4]),
                       vl=float(c["vlvu"][0]), vu=float(c["vlvu"][1]))
        c["threads"] = int(p[5])
        cases[name] = c
    return cases


@pytest.fixture(scope="session")
def oracle():
    from mrrr_ctypes import ORACLE_SO, load
    if not os.path.exists(ORACLE_SO):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(os.path.dirname(HERE), "oracle")], check=True)
    return load(ORACLE_SO, "oracle_mrrr")


@pytest.fixture(scope="session")
def oracle_lib(oracle):
    return oracle[0]
