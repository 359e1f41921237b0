"""The reference-side caller compiled against include/ag2.h: a C++ driver
with the reference's Particle / Simulation shape (Particle.h:9-57,
Simulation.cpp:121-139 and :166-364), whose tree calls are the engine's C
ABI, linked to libag2.so and run as its own process.  Its particles after
the initial force phase and three KDK steps equal the oracle's run."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR, rel_err

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(PKG_DIR, "sim_driver")


def write_input(path, ps):
    with open(path, "wb") as f:
        np.array([ps.n], np.int64).tofile(f)
        for fld in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu"):
            np.ascontiguousarray(getattr(ps, fld), np.float64).tofile(f)
        np.ascontiguousarray(ps.type, np.uint8).tofile(f)


def read_output(path, n):
    raw = np.fromfile(path, np.float64, offset=8)
    assert int(np.fromfile(path, np.int64, count=1)[0]) == n
    cols = raw.reshape(15, n)
    names = ("x", "y", "z", "vx", "vy", "vz", "U", "ax", "ay", "az", "h", "rho", "P", "T", "dUdt")
    return dict(zip(names, cols))


@pytest.mark.parametrize("precision", [0, 1])
def test_cxx_driver_matches_oracle_run(ag2, orc, tmp_path, precision):
    assert os.path.exists(DRIVER), "sim_driver is not built (run __graft_entry__.build())"
    U = ag2.units
    ps = ag2.ics.shuffled(ag2.ics.disk_galaxy(800, 600, 600, seed=81), 4)
    e0, mih, dt, steps = 0.1 * U.KPC, ag2.ics.mass_in_h(ps), 0.5 * U.MYR, 3
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(str(src), ps)
    r = subprocess.run([DRIVER, str(src), str(dst), str(steps), repr(0.5), repr(e0), repr(mih), repr(dt), "100",
                        str(precision)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = read_output(str(dst), ps.n)
    a = ps.copy()
    n = orc.run(a, orc.params(0.5, e0, mih, 100), 2.0, dt, dt, (steps - 0.5) * dt)
    assert n == steps
    pos_tol, vel_tol = (1e-11, 1e-9) if precision == 0 else (1e-8, 1e-5)
    for f in ("x", "y", "z"):
        assert np.abs(got[f] - getattr(a, f)).max() / np.abs(a.x).max() < pos_tol, f
    for f in ("vx", "vy", "vz"):
        assert np.abs(got[f] - getattr(a, f)).max() / np.abs(a.vx).max() < vel_tol, f
    gas = ps.type == U.TYPE_GAS
    assert rel_err(got["U"][gas], a.U[gas]).max() < (1e-9 if precision == 0 else 1e-5)
