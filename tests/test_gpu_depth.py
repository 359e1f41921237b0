"""GPU tests of the tree at depth: the reference splits a cell until its
distinct particles separate (Node.cpp:597-699), whatever the depth, and the
engine follows it to 126 levels with key words computed only where
particles still share the shallower ones.

- a run of escaped gas particles (the 2 x 20k merger with the reference's
  cooling and star-formation rules) pushes the root half-width to 5e22 kpc
  within six steps; one force step on that state equals the oracle's;
- distinct particles closer than 2^-126 of the root half-width fail the build
  with AG2_ERR_CAPACITY, as the oracle reports a depth error.
"""
import time

import numpy as np
import pytest

from conftest import moment_bound_ratios, rel_err, vec_rel_err

pytestmark = pytest.mark.gpu


def engine_for(ag2, ps, theta, e0, mih, thr, precision=0, **kw):
    E = ag2.engine
    prm = E.make_params(theta=theta, e0=e0, massInH=mih, split_threshold=thr, precision=precision, **kw)
    eng = E.Engine(prm, ps.n)
    eng.upload(ps)
    return eng


@pytest.fixture(scope="module")
def runaway_state(ag2, orc):
    """The 2 x 20k disk merger after six KDK steps of the oracle with cooling
    and star formation on (Cooling.cpp:6-26, SFR.cpp:12-28): the free-free
    rule drives gas U below zero, negative pressure flings gas outwards, and
    R reaches about 5e22 kpc.  Most of the system then shares the first 42
    octant levels (some 7,000 particles in one run)."""
    U = ag2.units
    ps = ag2.ics.merger(20000)
    e0, mih, dt = 0.1 * U.KPC, ag2.ics.mass_in_h(ps), 0.5 * U.MYR
    orc.run(ps, orc.params(0.5, e0, mih, 800, nthreads=16), 2.0, dt, dt, 5.5 * dt, cooling=1, star_formation=1,
            sfr_seed=7)
    return ps, e0, mih


@pytest.mark.parametrize("precision", [0, 1])
def test_runaway_merger_force_step_matches_oracle(ag2, orc, runaway_state, precision):
    """One force phase of the degenerate state: the tree (83 levels deep) is
    the oracle's, node for node; the node moments agree within the rounding
    two summation orders allow; and on the engine's moments the oracle's
    forces, densities and counts equal the engine's, within the rounding of
    two summation orders.  (Against the oracle's own moments the last bits
    of the centres of mass of cells 80 levels down a 5e22 kpc root move
    forces by up to 2e-10: the geometry is ill-conditioned, as in the clump
    case of the parity tests.)"""
    ps, e0, mih = runaway_state
    R = orc.tree_width(ps.x, ps.y, ps.z)
    assert R > 1e20 * ag2.units.KPC
    prm = orc.params(0.5, e0, mih, 800, nthreads=16)
    ref = orc.tree_export(ps.copy(), prm)
    assert ref["level"].max() > 60
    eng = engine_for(ag2, ps, 0.5, e0, mih, 800, precision=precision)
    t0 = time.perf_counter()
    eng.init_forces()
    dt = time.perf_counter() - t0
    d = eng.diagnostics()
    b = eng.download(ps.copy())
    t = eng.tree()
    eng.close()
    assert dt < 10.0, dt   # bounded: the deep refinement takes a few host reads per key word
    assert d["R"] == R
    assert d["max_level"] == ref["level"].max()
    for f in ("level", "first", "count"):
        assert np.array_equal(t[f], ref[f]), f
    rm, rc = moment_bound_ratios(t, ref, R)
    assert rm <= 1 and rc <= 1, (rm, rc)
    a = ps.copy()
    scale = np.zeros(ps.n)
    st = orc.force_step(a, prm, scale=scale, moments={k: t[k] for k in ("mass", "com", "gas", "mvel")})
    assert d["interactions"] == st["interactions"]
    assert d["sph_pairs"] == st["sph_interactions"]
    assert d["fp64_fallback"] == precision   # e0 / R far below 2^-24
    assert d["gap_particles"] == 0
    # Summation order differs (the block walk against the recursion), so
    # the error is bounded by the sum of contribution magnitudes: gas flung
    # out by negative pressure meets forces that cancel to a part in 1e6
    da = np.sqrt((b.ax - a.ax) ** 2 + (b.ay - a.ay) ** 2 + (b.az - a.az) ** 2)
    amag = np.sqrt(a.ax ** 2 + a.ay ** 2 + a.az ** 2)
    live = scale > 0
    es = da[live] / scale[live]
    cond = live & (amag > 1e-3 * scale)
    err = vec_rel_err(b.ax, b.ay, b.az, a.ax, a.ay, a.az)
    w = int(np.argmax(err))
    print("error / scale max %.3g; relative error max %.3g (particle %d, type %d, |a| / scale %.3g), "
          "%.3g where |a| > 1e-3 scale; %d particles without contributions"
          % (es.max(), err[w], w, ps.type[w], amag[w] / scale[w] if scale[w] > 0 else 0.0, err[cond].max(),
             (~live).sum()), flush=True)
    assert (da[~live] == 0).all()
    assert es.max() < 1e-13
    assert err[cond].max() < 1e-10
    assert np.array_equal(b.h, a.h)
    gas = ps.type == 2
    for f in ("rho", "P", "T"):
        assert (rel_err(getattr(b, f)[gas], getattr(a, f)[gas]) < 1e-12).all(), f


def test_runaway_merger_steps_bounded(ag2, orc, runaway_state):
    """Stepping the degenerate state on: within one or two steps the escaped
    particles stretch the root so far that distinct particles near the
    origin share all 126 levels, and the oracle's run stops with a depth
    error.  The engine's steps end the same way or succeed, in bounded time,
    and a failed step says why (the reference would recurse deeper)."""
    ps, e0, mih = runaway_state
    E = ag2.engine
    dt = 0.5 * ag2.units.MYR
    eng = engine_for(ag2, ps, 0.5, e0, mih, 800, precision=1, eta=2.0, min_dt=dt, max_dt=dt, cooling=1,
                     star_formation=1, sfr_seed=7)
    eng.init_forces()
    t0 = time.perf_counter()
    try:
        for _ in range(2):
            eng.step()
        eng.sync()
        assert eng.diagnostics()["max_level"] > 42
    except E.EngineError as e:
        assert e.code == E.AG2_ERR_CAPACITY
        assert "126 octant levels" in str(e)
    assert time.perf_counter() - t0 < 30.0
    eng.close()


def test_depth_limit_is_capacity_error(ag2, orc):
    """Distinct particles closer than 2^-126 of R (near the origin, where the
    cell centres stay exact that deep): AG2_ERR_CAPACITY from the build, a
    depth error from the oracle."""
    E = ag2.engine
    ps = ag2.ics.config1_plummer(500, seed=3)
    ps.x[10], ps.y[10], ps.z[10] = 1e-200, 1e-200, 1e-200
    ps.x[11], ps.y[11], ps.z[11] = 2e-200, 1e-200, 1e-200
    with pytest.raises(orc.DepthError):
        orc.force_step(ps.copy(), orc.params(0.5, 0.05 * ag2.units.KPC, 1.0, 100))
    eng = engine_for(ag2, ps, 0.5, 0.05 * ag2.units.KPC, 1.0, 100)
    with pytest.raises(E.EngineError) as ei:
        eng.build_tree()
    assert ei.value.code == E.AG2_ERR_CAPACITY
    # the context stays usable: without the pair the build succeeds
    ps.x[11] = 3e-1 * ps.x[12]
    eng.upload(ps)
    eng.build_tree()
    eng.close()
