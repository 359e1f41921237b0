import importlib.util
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "astrogenesis2.0_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (runs the HIP library)")


def load_pkg():
    if "ag2_amd" in sys.modules:
        return sys.modules["ag2_amd"]
    spec = importlib.util.spec_from_file_location("ag2_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ag2_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True,
                       stdout=subprocess.DEVNULL)
    return pyoracle


@pytest.fixture(scope="session")
def ag2():
    return load_pkg()


@pytest.fixture(scope="session")
def orc():
    return load_oracle()


def vec_rel_err(ax, ay, az, bx, by, bz):
    """|a - b| / |b| per particle (0 where both are zero)."""
    num = np.sqrt((ax - bx) ** 2 + (ay - by) ** 2 + (az - bz) ** 2)
    den = np.sqrt(bx ** 2 + by ** 2 + bz ** 2)
    out = np.zeros_like(num)
    nz = den > 0
    out[nz] = num[nz] / den[nz]
    out[~nz] = num[~nz]
    return out


def rel_err(a, b):
    den = np.abs(b)
    out = np.abs(a - b)
    nz = den > 0
    out[nz] = out[nz] / den[nz]
    return out
