"""Edge cases of the force phase against the oracle, and the C-ABI's error
behaviour: tiny and ragged particle counts (one particle, partial
wavefronts), particles outside the root cube (beyond mean + 10 sigma,
Tree.cpp:86-117), zero-mass particles (skipped as sources and targets,
Node.cpp:249-257), an all-gas set, and calls out of order."""
import numpy as np
import pytest

from conftest import vec_rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def compare(ag2, orc, ps, e0, mih, thr):
    E = ag2.engine
    eng = E.Engine(E.make_params(theta=0.5, e0=e0, massInH=mih, split_threshold=thr), ps.n)
    eng.upload(ps)
    eng.init_forces()
    got = eng.download(ps.copy())
    d = eng.diagnostics()
    eng.close()
    want = ps.copy()
    st = orc.force_step(want, orc.params(0.5, e0, mih, thr))
    assert d["interactions"] == st["interactions"]
    assert d["sph_pairs"] == st["sph_interactions"]
    assert d["n_in_tree"] == st["n_in_tree"]
    nz = (want.ax != 0) | (want.ay != 0) | (want.az != 0)
    assert np.array_equal(nz, (got.ax != 0) | (got.ay != 0) | (got.az != 0))
    if nz.any():
        err = vec_rel_err(got.ax[nz], got.ay[nz], got.az[nz], want.ax[nz], want.ay[nz], want.az[nz])
        assert err.max() < TOL
    assert np.array_equal(got.h, want.h)
    gas = want.type == 2
    if gas.any():
        s = np.abs(want.rho[gas]).max()
        if s > 0:
            assert np.abs(got.rho[gas] - want.rho[gas]).max() / s < 1e-12
    return d


@pytest.mark.parametrize("n", [1, 2, 3, 7, 63, 64, 65, 130])
@pytest.mark.parametrize("thr", [1, 100])
def test_tiny_and_ragged_counts(ag2, orc, n, thr):
    ps = ag2.ics.config1_plummer(n, seed=100 + n)
    d = compare(ag2, orc, ps, 0.05 * ag2.units.KPC, 1.0, thr)
    assert d["n_in_tree"] <= n


@pytest.mark.parametrize("thr", [100, 800])
def test_particles_outside_root(ag2, orc, thr):
    """Three far outliers lie beyond mean + 10 sigma: they stay out of the
    tree (and, on the parallel root, still count in its mass) but receive
    forces from it."""
    ps = ag2.ics.config1_plummer(2000, seed=7)
    R = np.sqrt(ps.x ** 2 + ps.y ** 2 + ps.z ** 2).max()
    for k, f in enumerate((300.0, -500.0, 800.0)):
        ps.x[k * 97] = f * R
    d = compare(ag2, orc, ps, 0.05 * ag2.units.KPC, 1.0, thr)
    assert d["n_in_tree"] == 2000 - 3


def test_zero_mass_particles(ag2, orc):
    ps = ag2.ics.config1_plummer(1500, seed=9)
    ps.m[::11] = 0.0
    compare(ag2, orc, ps, 0.05 * ag2.units.KPC, 1.0, 100)


@pytest.mark.parametrize("n", [5, 300])
def test_all_gas(ag2, orc, n):
    ps = ag2.ics.gas_sphere(n, seed=3)
    ps.vx[:] = np.random.default_rng(2).normal(0, 2e4, n)
    compare(ag2, orc, ps, 0.02 * ag2.units.KPC, 8 * ps.m[0], 100)


def test_api_errors(ag2):
    E = ag2.engine
    ps = ag2.ics.config1_plummer(200, seed=1)
    eng = E.Engine(E.make_params(), 100)
    with pytest.raises(E.EngineError) as e:
        eng.upload(ps)                                    # above the capacity
    assert e.value.code == E.AG2_ERR_CAPACITY
    eng.close()
    eng = E.Engine(E.make_params(), ps.n)
    eng.upload(ps)
    for call in (eng.forces, eng.gas_density, lambda: eng.visual_density()):
        with pytest.raises(E.EngineError) as e:
            call()                                        # before any tree
        assert e.value.code == E.AG2_ERR_STATE
    eng.build_tree()
    with pytest.raises(E.EngineError) as e:
        eng.download(ag2.ics.ParticleSet(ps.n - 1))      # count differs
    assert e.value.code == E.AG2_ERR_ARG
    eng.close()
    with pytest.raises(E.EngineError) as e:
        E.Engine(E.make_params(theta=-1.0), 10)
    assert e.value.code == E.AG2_ERR_ARG
    with pytest.raises(E.EngineError) as e:
        E.Engine(E.make_params(min_dt=2.0, max_dt=1.0), 10)
    assert e.value.code == E.AG2_ERR_ARG


def test_hubble_expansion_run(ag2, orc):
    """Simulation.cpp:330-332: positions scaled by exp(H0 dt) each step."""
    U = ag2.units
    ps = ag2.ics.shuffled(ag2.ics.disk_galaxy(800, 600, 600, seed=33), 5)
    e0, mih, dt, H0 = 0.1 * U.KPC, ag2.ics.mass_in_h(ps), 0.5 * U.MYR, 70.0
    a = ps.copy()
    steps = orc.run(a, orc.params(0.5, e0, mih, 100), 2.0, dt, dt, 3.5 * dt, H0=H0)
    E = ag2.engine
    eng = E.Engine(E.make_params(theta=0.5, e0=e0, massInH=mih, split_threshold=100, eta=2.0, min_dt=dt, max_dt=dt,
                                 H0=H0), ps.n)
    eng.upload(ps)
    eng.init_forces()
    for _ in range(steps):
        eng.step()
    eng.sync()
    b = eng.download(ps.copy())
    eng.close()
    for f in ("x", "y", "z"):
        assert np.abs(getattr(b, f) - getattr(a, f)).max() / np.abs(a.x).max() < 1e-11, f
    for f in ("vx", "vy", "vz"):
        assert np.abs(getattr(b, f) - getattr(a, f)).max() / np.abs(a.vx).max() < 1e-9, f


@pytest.mark.parametrize("level", [0, 1, 2])
def test_timing_event_levels(ag2, level):
    """ag2_params.timing_events: 0 records the phase boundaries and the walk
    kernel, 1 the walk kernel only, 2 nothing; results do not depend on it."""
    E, U = ag2.engine, ag2.units
    ps = ag2.ics.shuffled(ag2.ics.disk_galaxy(3000, 2000, 2000, seed=61), 4)
    dt = 0.5 * U.MYR
    out = []
    for lv in (0, level):
        eng = E.Engine(E.make_params(theta=0.5, e0=0.1 * U.KPC, massInH=ag2.ics.mass_in_h(ps), split_threshold=800,
                                     precision=E.AG2_FP32_SPLIT, min_dt=dt, max_dt=dt, timing_events=lv), ps.n)
        eng.upload(ps)
        eng.init_forces()
        eng.step()
        eng.sync()
        ph, kern = eng.phase_times(), eng.walk_kernel_ms()
        out.append(eng.download(ps.copy()))
        eng.close()
        if lv == 0:
            assert (ph[:5] > 0).all() and 0 < kern < ph[4]
        elif lv == 1:
            assert (ph == 0).all() and kern > 0
        else:
            assert (ph == 0).all() and kern == 0
    for f in ("x", "vx", "ax", "rho", "U"):
        assert np.array_equal(getattr(out[0], f), getattr(out[1], f)), f
