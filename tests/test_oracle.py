"""The oracle pinned against the reference's own outputs and against what the
mathematics fixes.  CPU only.

Golden fixtures (tests/golden/*.npz) hold outputs of the compiled reference
run single-threaded (make_golden.py); the oracle restates the same
arithmetic in the same order, so the comparison is bit-exact.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

from conftest import ROOT, vec_rel_err

GOLDEN = os.path.join(ROOT, "tests", "golden")
STEP_FIXTURES = ["plummer2k", "gassphere", "disk3k", "syw400", "close_pairs2k", "runaway2k"]


def load_fixture(ag2, name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    ps = ag2.ics.ParticleSet(len(z["in_x"]))
    for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu", "type", "T"):
        getattr(ps, f)[:] = z["in_" + f]
    return z, ps


@pytest.mark.parametrize("name", STEP_FIXTURES)
@pytest.mark.parametrize("cores", [1, 8])
def test_oracle_matches_reference_step(ag2, orc, name, cores):
    z, ps = load_fixture(ag2, name)
    a = ps.copy()
    st = orc.force_step(a, orc.params(float(z["theta"]), float(z["e0"]), float(z["massInH"]), 100 * cores))
    assert st["R"] == float(z["R"])
    for f in ("ax", "ay", "az", "h", "rho", "P", "T", "dUdt"):
        assert np.array_equal(getattr(a, f), z["c%d_%s" % (cores, f)]), f


@pytest.mark.parametrize("name", STEP_FIXTURES)
@pytest.mark.parametrize("cores", [1, 8])
def test_oracle_tree_matches_reference_tree(ag2, orc, name, cores):
    z, ps = load_fixture(ag2, name)
    t = orc.tree_export(ps, orc.params(0.5, 1.0, 1.0, 100 * cores))
    for f in ("level", "first", "count", "perm", "mass", "gas"):
        assert np.array_equal(t[f], z["c%d_tree_%s" % (cores, f)]), f
    assert np.array_equal(t["com"], z["c%d_tree_com" % cores])


def test_oracle_matches_reference_run(ag2, orc):
    z, ps = load_fixture(ag2, "disk2k_run")
    dt = float(z["dt"])
    n = orc.run(ps, orc.params(float(z["theta"]), float(z["e0"]), float(z["massInH"]), 100 * int(z["cores"])),
                2.0, dt, dt, (int(z["steps"]) - 0.5) * dt)
    assert n == int(z["steps"])
    for f in ("x", "y", "z", "vx", "vy", "vz", "U", "ax", "ay", "az", "h", "rho", "P", "dUdt"):
        assert np.array_equal(getattr(ps, f), z["out_" + f]), f


@pytest.mark.parametrize("name", ["disk2k_block", "disk2k_hubble"])
def test_oracle_matches_reference_run_block_hubble(ag2, orc, name):
    """Block time steps (eta 0.1, min_dt < max_dt: Simulation.cpp:199-254) and
    Hubble expansion (H0 = 70: Simulation.cpp:330-332) of the reference's step
    loop, bit for bit."""
    z, ps = load_fixture(ag2, name)
    n = orc.run(ps, orc.params(float(z["theta"]), float(z["e0"]), float(z["massInH"]), 100 * int(z["cores"])),
                float(z["eta"]), float(z["min_dt"]), float(z["max_dt"]), float(z["end_time"]), H0=float(z["H0"]))
    assert n == int(z["steps"])
    for f in ("x", "y", "z", "vx", "vy", "vz", "U", "ax", "ay", "az", "h", "rho", "P", "dUdt"):
        assert np.array_equal(getattr(ps, f), z["out_" + f]), f


def test_block_fixture_uses_several_bins(ag2):
    """The block-step fixture exercises more than one time-step bin: its step
    count exceeds end_time / max_dt."""
    z = np.load(os.path.join(GOLDEN, "disk2k_block.npz"), allow_pickle=False)
    assert int(z["steps"]) > float(z["end_time"]) / float(z["max_dt"]) * 4


def test_syw_branch_is_live_in_fixture(ag2, orc):
    """syw400 sits where the softening-kernel branch (Node.cpp:302) fires: at
    theta = 0 the tree walk differs from the e0-only direct sum for the
    closest pairs, and the reference fixture matches the oracle (above)."""
    z, ps = load_fixture(ag2, "syw400")
    e0 = float(z["e0"])
    a = ps.copy()
    orc.force_step(a, orc.params(0.0, e0, 1.0, 100))
    dx, dy, dz, _ = orc.direct_sum(ps.x, ps.y, ps.z, ps.m, e0)
    err = vec_rel_err(a.ax, a.ay, a.az, dx, dy, dz)
    assert (err > 1e-3).sum() > 10


def test_theta_zero_walk_is_direct_sum(ag2, orc):
    ps = ag2.ics.config1_plummer(2000, seed=5)
    e0 = 0.05 * ag2.units.KPC
    a = ps.copy()
    orc.force_step(a, orc.params(0.0, e0, 1.0, 100))
    ax, ay, az, _ = orc.direct_sum(ps.x, ps.y, ps.z, ps.m, e0)
    assert vec_rel_err(a.ax, a.ay, a.az, ax, ay, az).max() < 1e-12


def test_two_body_closed_form(orc):
    G, e0 = 6.67430e-11, 3.0e18
    x = np.array([0.0, 4.0e19]); y = np.array([0.0, 3.0e19]); z = np.zeros(2)
    m = np.array([2.0e40, 5.0e39])
    ax, ay, az, pot = orc.direct_sum(x, y, z, m, e0)
    r = 5.0e19
    f0 = G * m[1] / (r * r + e0 * e0)
    assert abs(ax[0] - f0 * 0.8) / (f0 * 0.8) < 1e-15 and abs(ay[0] - f0 * 0.6) / (f0 * 0.6) < 1e-15
    # equal and opposite forces (the softened law is antisymmetric)
    assert abs(m[0] * ax[0] + m[1] * ax[1]) < 1e-15 * abs(m[0] * ax[0])
    # potential -(G m / e0)(pi/2 - atan(r/e0)) and its derivative is the force
    p0 = -(G * m[1] / e0) * (np.pi / 2 - np.arctan(r / e0))
    assert abs(pot[0] - p0) / abs(p0) < 1e-14
    dr = r * 1e-6
    dphi = (-(G * m[1] / e0) * (np.pi / 2 - np.arctan((r + dr) / e0)) - p0) / dr
    assert abs(dphi - f0) / f0 < 1e-5


def test_tree_two_body_matches_closed_form(ag2, orc):
    ps = ag2.ics.ParticleSet(2)
    ps.x[:] = [1.0e19, -2.0e19]; ps.y[:] = [0.0, 1.0e19]; ps.m[:] = [1.0e40, 3.0e40]
    orc.force_step(ps, orc.params(0.5, 1.0e18, 1.0, 100))
    d = np.array([-3.0e19, 1.0e19, 0.0])
    r = np.linalg.norm(d)
    f = 6.67430e-11 * 3.0e40 / (r * r + 1.0e36)
    assert abs(ps.ax[0] - f * d[0] / r) / abs(f * d[0] / r) < 1e-15


def test_force_scale_closed_form(ag2, orc):
    """The contribution-magnitude sum: a particle midway between two equal
    masses has zero acceleration but a scale of 2 G m / (d^2 + e0^2); an end
    particle's scale is the sum of the magnitudes of its two terms; the
    accelerations equal those of the plain force step."""
    ps = ag2.ics.ParticleSet(3)
    d, e0, m = 2.0e19, 1.0e18, 5.0e39
    ps.x[:] = [-d, 0.0, d]; ps.m[:] = m
    scale = np.zeros(3)
    a = ps.copy()
    orc.force_step(a, orc.params(0.0, e0, 1.0, 100), scale=scale)
    G = 6.67430e-11
    assert a.ax[1] == 0.0
    assert abs(scale[1] - 2 * G * m / (d * d + e0 * e0)) < 1e-15 * scale[1]
    want0 = G * m / (d * d + e0 * e0) + G * m / (4 * d * d + e0 * e0)
    assert abs(scale[0] - want0) < 1e-15 * want0
    b = ps.copy()
    orc.force_step(b, orc.params(0.0, e0, 1.0, 100))
    assert np.array_equal(a.ax, b.ax)


def test_force_step_with_own_moments_is_identical(ag2, orc):
    """Feeding the oracle's exported node moments back reproduces its force
    step bit for bit (the export order is the override's order); moving one
    node's centre changes the result."""
    ps = ag2.ics.shuffled(ag2.ics.disk_galaxy(1500, 1000, 1500, seed=21), 2)
    e0, mih = 0.1 * ag2.units.KPC, ag2.ics.mass_in_h(ps)
    prm = orc.params(0.5, e0, mih, 100)
    a, b = ps.copy(), ps.copy()
    sa = orc.force_step(a, prm)
    t = orc.tree_export(ps.copy(), prm)
    mom = {k: t[k].copy() for k in ("mass", "com", "gas", "mvel")}
    sb = orc.force_step(b, prm, moments=mom)
    assert sa == sb
    for f in ("ax", "ay", "az", "h", "rho", "P", "T", "dUdt"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    mom["com"][1, 0] += 0.01 * ag2.units.KPC
    c = ps.copy()
    orc.force_step(c, prm, moments=mom)
    assert not np.array_equal(a.ax, c.ax)
    with pytest.raises(ValueError):
        orc.force_step(ps.copy(), prm, moments={k: v[:-1] for k, v in mom.items()})


def test_direct_sum_newton_third_law(ag2, orc):
    ps = ag2.ics.config1_plummer(1500, seed=8)
    ax, ay, az, _ = orc.direct_sum(ps.x, ps.y, ps.z, ps.m, 0.05 * ag2.units.KPC)
    tot = np.array([np.sum(ps.m * ax), np.sum(ps.m * ay), np.sum(ps.m * az)])
    scale = np.sum(ps.m * np.sqrt(ax ** 2 + ay ** 2 + az ** 2))
    assert np.abs(tot).max() / scale < 1e-13


def test_plummer_mean_field(ag2, orc):
    """Direct-sum field of a Plummer sphere against -G M r / (r^2 + a^2)^1.5."""
    U = ag2.units
    ps = ag2.ics.plummer(6000, 1e10, 1.0, 20.0, seed=2, zero_velocity=True)
    ax, ay, az, _ = orc.direct_sum(ps.x, ps.y, ps.z, ps.m, 0.005 * U.KPC)
    r = np.sqrt(ps.x ** 2 + ps.y ** 2 + ps.z ** 2)
    ar = (ax * ps.x + ay * ps.y + az * ps.z) / r
    a = U.KPC
    M = ps.m.sum() * (1 + 1 / 20.0 ** 2) ** 1.5   # untruncated mass of the truncated sample
    sel = (r > 0.5 * a) & (r < 2.0 * a)
    model = -U.G * M * r[sel] / (r[sel] ** 2 + a * a) ** 1.5
    assert abs(np.mean(ar[sel] / model) - 1.0) < 0.03


def exact_digits(x, R, levels):
    """Octant bits by exact rational arithmetic for R a power of two (all
    centres are then exact dyadic numbers): bit set iff x > centre."""
    c, r, bits = Fraction(0), Fraction(R), []
    for _ in range(levels):
        b = 1 if Fraction(x) > c else 0
        bits.append(b)
        c = c + r * (Fraction(1, 2) if b else Fraction(-1, 2))
        r = r / 2
    return bits


def test_morton_keys_exact(ag2, orc):
    g = np.random.default_rng(3)
    R = 2.0 ** 70
    n = 400
    x = g.uniform(-R, R, n); y = g.uniform(-R, R, n); z = g.uniform(-R, R, n)
    # points exactly on cell-centre planes (ties go to the lower octant)
    x[:50] = np.ldexp(np.round(np.ldexp(x[:50], -60)), 60)
    y[50:100] = 0.0
    z[100:120] = R
    z[120:140] = -R
    k1, k2, gap = orc.morton_keys(x, y, z, R)
    for i in range(n):
        bx, by, bz = (exact_digits(v, R, 42) for v in (x[i], y[i], z[i]))
        digits = [bx[L] | (by[L] << 1) | (bz[L] << 2) for L in range(42)]
        e1 = 0
        for d in digits[:21]:
            e1 = (e1 << 3) | d
        e2 = 0
        for d in digits[21:]:
            e2 = (e2 << 3) | d
        assert int(k1[i]) == e1 and int(k2[i]) == e2, i
    assert (gap == 0).all()


def test_morton_keys_match_tree_order(ag2, orc):
    ps = ag2.ics.config1_plummer(3000, seed=4)
    R = orc.tree_width(ps.x, ps.y, ps.z)
    k1, k2, gap = orc.morton_keys(ps.x, ps.y, ps.z, R)
    order = np.lexsort((np.arange(ps.n), k2, k1))
    t = orc.tree_export(ps, orc.params(0.5, 1.0, 1.0, 100))
    inside = gap >= 0
    assert np.array_equal(t["perm"], order[inside[order]])


def test_tree_width_closed_form(orc):
    x = np.array([3.0, 0.0, 0.0, 1.0]); y = np.array([4.0, 1.0, 0.0, 0.0]); z = np.zeros(4)
    assert orc.tree_width(x, y, z) == 5.0
    # a far outlier beyond mean + 10 sigma is cut
    n = 200
    x = np.ones(n); x[-1] = 1e6
    assert orc.tree_width(x, np.zeros(n), np.zeros(n)) == 1.0


# ---------------------------------------------------------------- visual density
@pytest.mark.parametrize("cores", [1, 8])
def test_oracle_visual_density_matches_reference(ag2, orc, cores):
    z, ps = load_fixture(ag2, "visual3k")
    for k, r in enumerate(z["radii"]):
        vd, R = orc.visual_density(ps, orc.params(0.5, 1.0, 1.0, 100 * cores), float(r))
        assert R == float(z["R"])
        assert np.array_equal(vd, z["c%d_vd%d" % (cores, k)]), k


def visual_density_from_export(t, R, r, n):
    """Visual density from the pre-order node export alone: the climb stops at
    the deepest level s <= leaf level whose half-width R/2^s is not farther
    from r than its parent's, and the value is that node's mass / (R/2^s)^3."""
    level, first, count, mass = t["level"], t["first"], t["count"], t["mass"]
    rad = lambda l: np.ldexp(R, -int(l))   # noqa: E731
    out = np.zeros(n)
    leaves = np.nonzero(count == 1)[0]
    for nd in leaves:
        k = first[nd]
        s = int(level[nd])
        while s > 0 and abs(r - rad(s)) > abs(r - rad(s - 1)):
            s -= 1
        if s == 0:
            continue
        # the level-s node whose leaf interval holds k
        cand = np.nonzero((level == s) & (first <= k) & (first + count > k))[0]
        assert len(cand) == 1
        out[t["perm"][k]] = mass[cand[0]] / (rad(s) * rad(s) * rad(s))
    return out


def test_visual_density_equals_export_restatement(ag2, orc):
    z, ps = load_fixture(ag2, "visual3k")
    prm = orc.params(0.5, 1.0, 1.0, 100)
    t = orc.tree_export(ps, prm)
    for r in (0.0, float(z["radii"][1]), float(z["radii"][2])):
        vd, R = orc.visual_density(ps, prm, r)
        want = visual_density_from_export(t, R, r if r > 0 else R / 100000, ps.n)
        assert np.allclose(vd, want, rtol=1e-14, atol=0), r
    # the clump's leaves lie deeper than level 17, so at the default radius
    # some particles take a node mass larger than their own
    vd, R = orc.visual_density(ps, prm, 0.0)
    leaf_only = np.isin(vd, ps.m / np.ldexp(R, -np.arange(1, 44))[:, None] ** 3)
    assert (~leaf_only).sum() >= 100


def test_visual_density_closed_form(ag2, orc):
    """Two particles in opposite octants: leaves at level 1, half-width R/2."""
    ps = ag2.ics.ParticleSet(2)
    ps.x[:] = [1.0, -2.0]; ps.y[:] = [1.0, -2.0]; ps.z[:] = [1.0, -2.0]
    ps.m[:] = [3.0, 5.0]
    prm = orc.params(0.5, 1.0, 1.0, 100)
    vd, R = orc.visual_density(ps, prm, 0.0)
    assert R == np.sqrt(12.0)
    assert np.array_equal(vd, ps.m / (R / 2) ** 3)
    # a radius nearer the root's half-width than the leaves': the climb ends at the root, which writes nothing
    vd, _ = orc.visual_density(ps, prm, 5 * R)
    assert np.array_equal(vd, np.zeros(2))
    # a radius exactly the leaves' half-width stays at the leaf
    vd, _ = orc.visual_density(ps, prm, R / 2)
    assert np.array_equal(vd, ps.m / (R / 2) ** 3)


# ---------------------------------------------------------------- cooling and star formation
def load_gas_sources(ag2):
    z = np.load(os.path.join(GOLDEN, "gas_sources.npz"), allow_pickle=False)
    ps = ag2.ics.ParticleSet(len(z["in_T"]))
    for f in ("type", "T", "rho", "U", "dUdt"):
        getattr(ps, f)[:] = z["in_" + f]
    return z, ps


@pytest.mark.parametrize("tag,cool,sf", [("cool", 1, 0), ("sfr", 0, 1), ("both", 1, 1)])
def test_oracle_gas_sources_match_reference(ag2, orc, tag, cool, sf):
    z, ps = load_gas_sources(ag2)
    sfr = np.zeros(ps.n)
    orc.gas_sources(ps, z["in_dt"], cool, sf, seed=3, step=0, sfr=sfr)
    assert np.array_equal(ps.dUdt, z[tag + "_dUdt"])
    assert np.array_equal(sfr, z[tag + "_sfr"])
    if sf:
        eligible = z[tag + "_sfr"] > 0
        assert eligible.sum() > 100
        # the draws differ from rand(); conversions only among eligible gas, U zeroed
        conv = ps.type != z["in_type"]
        assert not (conv & ~eligible).any()
        assert np.array_equal(ps.U[conv], np.zeros(conv.sum()))
        ref_conv = z[tag + "_type"] != z["in_type"]
        assert not (ref_conv & ~eligible).any()


def test_cooling_closed_form(ag2, orc):
    ps = ag2.ics.ParticleSet(3)
    ps.type[:] = 2
    ps.T[:] = [1e4, 4e6, 1e4]
    ps.rho[:] = [1e-21, 2e-21, 0.0]
    ps.dUdt[:] = 0.0
    orc.gas_sources(ps, np.full(3, 1e13), 1, 0)
    # 1.42e-27 erg s^-1 cm^3 * g_ff 1.1 * sqrt(T) * (1e6 cm^-3)^2, erg/s/cm^3 -> W/m^3 is x 1e-7
    want = -np.array([1.42e-27 * 1.1 * 100 * 1e12 * 1e-7 / 1e-21, 1.42e-27 * 1.1 * 2000 * 1e12 * 1e-7 / 2e-21, 0.0])
    assert np.allclose(ps.dUdt, want, rtol=1e-15, atol=0)


def test_sfr_draws_uniform_and_conversion_rate(ag2, orc):
    u = orc.sfr_uniform(7, np.arange(20000), 3)
    assert ((u >= 0) & (u < 1)).all()
    assert abs(u.mean() - 0.5) < 0.01 and abs((u < 0.1).mean() - 0.1) < 0.01
    # different steps and seeds give independent draws
    assert np.corrcoef(u, orc.sfr_uniform(7, np.arange(20000), 4))[0, 1] < 0.03
    assert not np.array_equal(u, orc.sfr_uniform(8, np.arange(20000), 3))
    # eligible gas converts with probability p = 1 - exp(-0.1 dt / 1e15)
    n = 20000
    ps = ag2.ics.ParticleSet(n)
    ps.type[:] = 2
    ps.T[:] = 5e3
    ps.rho[:] = 1e-21
    dt = np.full(n, 1e15)
    sfr = np.zeros(n)
    orc.gas_sources(ps, dt, 0, 1, seed=11, step=0, sfr=sfr)
    p = 1 - np.exp(-0.1)
    assert np.array_equal(sfr, np.full(n, p))
    k = (ps.type == 1).sum()
    assert abs(k - n * p) < 5 * np.sqrt(n * p * (1 - p))
