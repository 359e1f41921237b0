"""Write the golden fixtures of tests/golden/ from the compiled reference.

Runs only where the reference sources exist (they are compiled by
oracle/Makefile into oracle/_ref/libag2ref.so).  The reference is pinned:
OMP_THREAD_LIMIT=1 makes every OpenMP region single-threaded, so the SPH
density sum (Node.cpp:770-779) is race-free and the octant partition order
is fixed, while `cores` still selects the split threshold 100*cores.

Each fixture stores the input arrays (so that the inputs are bit-identical
wherever the tests run) and the reference's outputs.

    python tests/golden/make_golden.py             # every fixture
    python tests/golden/make_golden.py visual3k    # only the named ones
"""
import os
import subprocess
import sys
import tempfile

os.environ["OMP_THREAD_LIMIT"] = "1"   # before libgomp is loaded

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import load_pkg  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

FIELDS_IN = ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu", "type", "T")
FIELDS_OUT = ("ax", "ay", "az", "h", "rho", "P", "T", "dUdt")


def _inputs(ps):
    return {"in_" + f: getattr(ps, f).copy() for f in FIELDS_IN}


def _outputs(ps, tag):
    return {"%s_%s" % (tag, f): getattr(ps, f).copy() for f in FIELDS_OUT}


def step_fixture(name, ps, theta, e0, mih, cores=(1, 8), tree=True):
    out = _inputs(ps)
    out.update(theta=theta, e0=e0, massInH=mih)
    for c in cores:
        b = ps.copy()
        R = O.ref_force_step(b, theta, e0, mih, cores=c)
        out["R"] = R
        out.update(_outputs(b, "c%d" % c))
        if tree:
            t = O.ref_tree_export(ps.copy(), cores=c)
            for k in ("level", "first", "count", "mass", "com", "gas", "perm"):
                out["c%d_tree_%s" % (c, k)] = t[k]
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, ps.n)


def run_fixture(name, ps, theta, e0, mih, dt, steps, cores=1):
    out = _inputs(ps)
    b = ps.copy()
    n = O.ref_run(b, theta, e0, mih, 2.0, dt, dt, (steps - 0.5) * dt, cores=cores)
    assert n == steps
    out.update(theta=theta, e0=e0, massInH=mih, dt=dt, steps=steps, cores=cores)
    for f in ("x", "y", "z", "vx", "vy", "vz", "U") + FIELDS_OUT:
        out["out_" + f] = getattr(b, f).copy()
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, ps.n)


def run_fixture_general(name, ps, theta, e0, mih, eta, min_dt, max_dt, end_time, H0, cores=1):
    """Simulation::run with block time steps (min_dt < max_dt) and / or Hubble
    expansion (H0 != 0): the step count is the reference's own."""
    out = _inputs(ps)
    b = ps.copy()
    n = O.ref_run(b, theta, e0, mih, eta, min_dt, max_dt, end_time, H0=H0, cores=cores)
    out.update(theta=theta, e0=e0, massInH=mih, eta=eta, min_dt=min_dt, max_dt=max_dt, end_time=end_time, H0=H0,
               steps=n, cores=cores)
    for f in ("x", "y", "z", "vx", "vy", "vz", "U") + FIELDS_OUT:
        out["out_" + f] = getattr(b, f).copy()
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, ps.n, "steps", n)


def syw_cluster(ics, U, n, seed):
    """Particles within 3e4 m of each other with e0 = 1e4 m: below
    e0 * 0.0001 < 2147483647 the softening-kernel branch of Node.cpp:302 can
    fire, and for pairs closer than about 2.8e4 m it does."""
    g = np.random.default_rng(seed)
    ps = ics.ParticleSet(n)
    r = 3e4 * g.random(n) ** (1.0 / 3.0)
    u = g.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    ps.x[:], ps.y[:], ps.z[:] = (r[:, None] * u).T
    ps.m[:] = g.uniform(0.5e20, 1.5e20, n)
    ps.type[:] = U.TYPE_STAR
    return ps


def _in_tree(ps):
    """Particles the compiled reference places in leaves (a particle in a
    rounding gap of a deep cell's bounds check is dropped, Node.cpp:606-612)."""
    return int(O.ref_tree_export(ps.copy(), cores=1)["n_in_tree"])


def close_pairs(ics, n, seed):
    """A Plummer sphere with twelve distinct pairs 1e-13 .. 3e-15 of the root
    half-width apart: the reference splits their cells 44 to 50 levels deep
    until they separate (Node.cpp:597-699).  Seeds whose deep cells drop a
    particle into a rounding gap are skipped."""
    for s in range(seed, seed + 50):
        ps = ics.config1_plummer(n, seed=s)
        g = np.random.default_rng(s + 7)
        r = np.sqrt(ps.x ** 2 + ps.y ** 2 + ps.z ** 2)
        R = float(O.tree_width(ps.x, ps.y, ps.z))
        inner = np.flatnonzero(r < 0.5 * np.median(r))
        k = g.choice(inner, 24, replace=False)
        for j, (a, b) in enumerate(zip(k[:12], k[12:])):
            eps = R * (1e-13, 3e-14, 1e-14, 3e-15)[j % 4]
            d = g.normal(size=3)
            d *= eps / np.abs(d).max()
            ps.x[b], ps.y[b], ps.z[b] = ps.x[a] + d[0], ps.y[a] + d[1], ps.z[a] + d[2]
        if _in_tree(ps) == n:
            return ps
    raise RuntimeError("no seed without a rounding gap")


def runaway_disk(ics, U, n_far, seed):
    """A 2,000-particle disk galaxy with gas plus n_far particles between 5e23
    and 1e24 kpc in random directions -- enough of them that mean + 10 sigma
    of |x| keeps them in the root cube (Tree.cpp:86-117), so R is about 1e24
    kpc and the galaxy sits some 75 levels down a chain of single-child
    cells, as in a run whose gas has escaped."""
    ds = ics.disk_galaxy(800, 600, 600, seed=seed)
    g = np.random.default_rng(seed + 1)
    far = ics.ParticleSet(n_far)
    u = g.normal(size=(n_far, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    rad = g.uniform(5e23, 1e24, n_far) * U.KPC
    far.x[:], far.y[:], far.z[:] = (rad[:, None] * u).T
    far.vx[:] = g.normal(0, 1e5, n_far)
    far.m[:] = ds.m[ds.type == U.TYPE_DM].mean()
    far.type[:] = U.TYPE_DM
    ps = ics.shuffled(ics.ParticleSet.concat([ds, far]), seed + 2)
    R = float(O.tree_width(ps.x, ps.y, ps.z))
    assert R > 5e23 * U.KPC, R
    assert _in_tree(ps) == ps.n
    return ps


def visual_fixture(name, ps, radii, cores=(1, 8)):
    """Tree::calcVisualDensity for each radius (0: the reference's R/1e5)."""
    out = _inputs(ps)
    out["radii"] = np.array(radii, dtype=np.float64)
    for c in cores:
        for k, r in enumerate(radii):
            vd, R = O.ref_visual_density(ps.copy(), radius=r, cores=c)
            out["c%d_vd%d" % (c, k)] = vd
        out["R"] = R
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, ps.n)


def clumped_plummer(ics, n, nclump, seed):
    """A Plummer sphere with a clump of nclump particles packed within 1e-7
    of the root half-width, so that leaves sit deeper than the visual-density
    level (about 17) and the climb is exercised at the default radius."""
    ps = ics.config1_plummer(n, seed=seed)
    g = np.random.default_rng(seed + 100)
    R = float(np.sqrt(ps.x ** 2 + ps.y ** 2 + ps.z ** 2).max())
    c = np.array([0.21, -0.13, 0.07]) * R
    k = g.choice(n, nclump, replace=False)
    d = g.uniform(-1.0, 1.0, (nclump, 3)) * 1e-7 * R
    ps.x[k], ps.y[k], ps.z[k] = c[0] + d[:, 0], c[1] + d[:, 1], c[2] + d[:, 2]
    return ps


def snapshot_fixture(name, ics, U):
    """DataManager::saveData output bytes for every format, from a mixed set
    with every galaxyPart/type pairing (including ones the Gadget writer only
    counts), written whole and cut to the first 150 particles."""
    g = np.random.default_rng(23)
    ps = ics.shuffled(ics.disk_galaxy(120, 100, 80, seed=9), 3)
    n = ps.n
    ps.T[:] = g.uniform(1e3, 1e6, n)
    ps.P[:] = g.uniform(0.0, 1e-12, n)
    ps.U[:] = np.where(ps.type == U.TYPE_GAS, g.uniform(1e7, 1e9, n), 0.0)
    ps.id[:] = g.permutation(n).astype(np.uint32)
    extras = {"galaxyPart": g.integers(0, 4, n).astype(np.uint8),
              "visualDensity": g.uniform(0.0, 1e-18, n), "sfr": g.uniform(0.0, 1e-3, n)}
    out = _inputs(ps)
    out["in_P"] = ps.P.copy()
    out["in_id"] = ps.id.copy()
    for k, v in extras.items():
        out["in_" + k] = v
    times = dict(delta_time=1.57788e13, end_time=3.15576e14, current_time=4.7e13)
    out.update(times)
    tmp = tempfile.mkdtemp(prefix="ag2snap")
    for fmt in ("ag", "agc", "age", "gadget", "hdf5"):
        for n_out in (n, 150):
            b = O.ref_save_data(ps, extras, tmp, fmt, 7, n_out, times["delta_time"], times["end_time"],
                                times["current_time"])
            out["%s_%d" % (fmt, n_out)] = np.frombuffer(b, np.uint8).copy()
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, ps.n)


def gas_sources_fixture(name, ics, U):
    """Cooling::coolingRoutine and SFR::sfrRoutine of the reference on gas
    spanning both thresholds (rho 1e-22 kg/m^3, T 1e4 K)."""
    g = np.random.default_rng(29)
    n = 4000
    ps = ics.ParticleSet(n)
    ps.type[:] = np.where(g.random(n) < 0.8, U.TYPE_GAS, U.TYPE_STAR).astype(np.uint8)
    ps.T[:] = 10.0 ** g.uniform(2.0, 7.0, n)
    ps.rho[:] = 10.0 ** g.uniform(-25.0, -19.0, n)
    ps.rho[::97] = 0.0
    ps.U[:] = g.uniform(1e6, 1e9, n)
    ps.dUdt[:] = g.normal(0.0, 1e-3, n)
    dt = 10.0 ** g.uniform(12.0, 15.0, n)
    out = {"in_type": ps.type.copy(), "in_T": ps.T.copy(), "in_rho": ps.rho.copy(), "in_U": ps.U.copy(),
           "in_dUdt": ps.dUdt.copy(), "in_dt": dt}
    for tag, cool, sf in (("cool", 1, 0), ("sfr", 0, 1), ("both", 1, 1)):
        b = ps.copy()
        sfr = O.ref_gas_sources(b, dt, cool, sf)
        out.update({"%s_dUdt" % tag: b.dUdt, "%s_sfr" % tag: sfr, "%s_type" % tag: b.type, "%s_U" % tag: b.U})
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, n)


def main(only=()):
    if not O.ref_available():
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    ag2 = load_pkg()
    U = ag2.units
    ics = ag2.ics
    if only:
        if "close_pairs2k" in only:
            step_fixture("close_pairs2k", close_pairs(ics, 2000, seed=61), 0.5, 0.05 * U.KPC, 1.0)
        if "runaway2k" in only:
            rw = runaway_disk(ics, U, 50, seed=67)
            step_fixture("runaway2k", rw, 0.5, 0.1 * U.KPC, ics.mass_in_h(rw))
        if "syw400" in only:
            step_fixture("syw400", syw_cluster(ics, U, 400, seed=51), 0.5, 1e4, 1.0)
        if "disk2k_block" in only:
            # block time steps: eta 0.1, four power-of-two bins from 0.125 to 1 Myr
            db = ics.shuffled(ics.disk_galaxy(700, 500, 800, seed=43), 23)
            run_fixture_general("disk2k_block", db, 0.5, 0.1 * U.KPC, ics.mass_in_h(db), 0.1, 0.125 * U.MYR,
                                1.0 * U.MYR, 2.0 * U.MYR, 0.0)
        if "disk2k_hubble" in only:
            # Hubble expansion of positions, H0 = 70 km/s/Mpc (Simulation.cpp:330-332)
            dh = ics.shuffled(ics.disk_galaxy(700, 500, 800, seed=47), 29)
            dt = 0.5 * U.MYR
            run_fixture_general("disk2k_hubble", dh, 0.5, 0.1 * U.KPC, ics.mass_in_h(dh), 2.0, dt, dt, 3.5 * dt,
                                70.0)
        if "gas_sources" in only:
            gas_sources_fixture("gas_sources", ics, U)
        if "snapshots" in only:
            snapshot_fixture("snapshots", ics, U)
        if "visual3k" in only:
            ps = clumped_plummer(ics, 3000, 200, seed=7)
            R = O.ref_visual_density(ps.copy())[1]
            visual_fixture("visual3k", ps, [0.0, R / 50, R / 3000])
        return
    # Plummer of the CPU-runnable configuration (smaller N), theta 0.5
    step_fixture("plummer2k", ics.config1_plummer(2000, seed=1), 0.5, 0.05 * U.KPC, 1.0)
    # the reference's own gas-sphere example file (1472 gas particles)
    gs = ag2.ageio.read_gadget("/root/reference/input_data/Example/gassphere_littleendian.dat")
    gs.vx[:] = np.random.default_rng(5).normal(0.0, 3e3, gs.n)
    step_fixture("gassphere", gs, 0.5, 0.05 * U.KPC, 64 * gs.m.mean())
    # mixed disk, shuffled array order (the reference shuffles at start-up)
    ds = ics.shuffled(ics.disk_galaxy(1000, 800, 1200, seed=31), 17)
    step_fixture("disk3k", ds, 0.5, 0.1 * U.KPC, ics.mass_in_h(ds))
    # five KDK steps of a mixed disk
    dr = ics.shuffled(ics.disk_galaxy(700, 500, 800, seed=41), 19)
    run_fixture("disk2k_run", dr, 0.5, 0.1 * U.KPC, ics.mass_in_h(dr), 0.5 * U.MYR, 5)
    main(("visual3k", "snapshots", "gas_sources", "syw400", "disk2k_block", "disk2k_hubble", "close_pairs2k",
          "runaway2k"))


if __name__ == "__main__":
    # python tests/golden/make_golden.py [fixture ...]: all, or only those named
    main(tuple(sys.argv[1:]))
