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


def coincident_groups(ics, n=2500, seed=12):
    """Particles at identical positions (buckets) in the situations the
    bucket rule has to get right: an isolated triple; a pair with a distinct
    particle 1e-14 of R away inserted after it (the bucket is split again,
    deeper than 42 levels) and one inserted before it; five with a distinct
    particle 1e-10 of R away."""
    ps = ics.config1_plummer(n, seed=seed)
    R = float(np.sqrt(ps.x ** 2 + ps.y ** 2 + ps.z ** 2).max())
    P = np.stack([ps.x, ps.y, ps.z], axis=1)

    def put(i, src, dx=0.0):
        ps.x[i], ps.y[i], ps.z[i] = P[src, 0] + dx, P[src, 1], P[src, 2]

    for i in (200, 201, 202):
        put(i, 199)
    for i in (300, 301):
        put(i, 302)
    put(1500, 302, 1e-14 * R)
    for i in (450, 451):
        put(i, 452)
    put(50, 452, -3e-14 * R)
    for i in (700, 701, 702, 703):
        put(i, 704)
    put(2100, 704, 1e-10 * R)
    return ps


def fixture_input(ag2, name):
    z = np.load(os.path.join(ROOT, "tests", "golden", name + ".npz"), allow_pickle=False)
    ps = ag2.ics.ParticleSet(len(z["in_x"]))
    for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu", "type", "T"):
        getattr(ps, f)[:] = z["in_" + f]
    return ps, float(z["e0"]), float(z["massInH"])


def moment_bound_ratios(tree, ref, R):
    """Node moments of the engine against the oracle's (the reference's
    summation order): the rounding two fixed-order sums of the same k positive
    terms allow (each within (k - 1) u of the sum, u = 2^-53), and for a centre
    of mass 2 k u of the largest coordinate in the node (|com| + sqrt(3)
    half-width) with a factor 4 for the running average's two roundings per
    step.  Returns the largest error / bound of masses and of centres."""
    u = 2.0 ** -53
    k = tree["count"].astype(float)
    dm = np.abs(tree["mass"] - ref["mass"])
    bm = 2 * k * u * np.abs(ref["mass"])
    half = R * np.ldexp(1.0, -tree["level"])
    dc = np.abs(tree["com"] - ref["com"]).max(axis=1)
    bc = 8 * k * u * (np.abs(ref["com"]).max(axis=1) + np.sqrt(3) * half)
    return float((dm / np.where(bm > 0, bm, 1)).max()), float((dc / np.where(bc > 0, bc, 1)).max())
