"""Serving pool (ag2_pool_*): continuous batching of independent simulations.

Each job's result in the mix must equal the same job run alone on a context
of its own (same steps, same parameters): the batched walk gives every job
the arithmetic of its own walk, so the tolerance below is only a guard
against summation-order changes (it is met bit for bit in practice).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "U", "h", "rho", "dUdt")
TOL = 1e-12


def small_jobs(ag2, count, seed=11):
    """Serving-stream jobs shrunk to test size (same recipe, fewer particles)."""
    U = ag2.units
    out = []
    for j in range(count):
        g = ag2.ics.rng(seed, 4, 9, j)
        n = int(g.integers(1500, 5000))
        if j % 2 == 0:
            ps = ag2.ics.plummer(n, 1e10, 1.0, 20.0, seed=seed, config=4, component=1, job=j)
        else:
            ps = ag2.ics.hernquist(n, 5e9, 0.8, 50.0, seed=seed, config=4, component=2, job=j)
        if j % 3 == 1:
            k = g.choice(n, n // 5, replace=False)
            ps.type[k] = U.TYPE_GAS
            ps.U[k] = U.u_from_temperature(1e4)
            ps.init_temperature()
        prec = ag2.engine.AG2_FP32_SPLIT if j % 4 != 3 else ag2.engine.AG2_FP64
        dt = 0.1 * U.MYR
        prm = ag2.engine.make_params(theta=0.5, e0=0.05 * U.KPC, massInH=ag2.ics.mass_in_h(ps),
                                     split_threshold=800, precision=prec, eta=2.0, min_dt=dt, max_dt=dt)
        out.append((ps, prm, 3 + (j % 4)))
    return out


def solo(ag2, ps, prm, steps):
    eng = ag2.engine.Engine(prm, ps.n)
    eng.upload(ps)
    eng.init_forces()
    for _ in range(steps):
        eng.step()
    eng.sync()
    out = eng.download(ps.copy())
    eng.close()
    return out


def close(a, b):
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        scale = np.abs(y).max()
        if scale > 0 and np.abs(x - y).max() / scale > TOL:
            return f
    return None


def run_pool(ag2, jobs, max_jobs, budget=0, trace=None):
    pool = ag2.engine.Pool(max_jobs=max_jobs, hbm_budget=budget)
    ids = [pool.submit(prm, ps, steps) for ps, prm, steps in jobs]
    res = {}
    for _ in range(1000):
        n = pool.step()
        st = pool.stats()
        if trace is not None:
            trace.append((n, st["running"], st["queued"], st["done"]))
        for i, jid in enumerate(ids):
            if i not in res and pool.status(jid)[0] == ag2.engine.JOB_DONE:
                res[i] = pool.result(jid, jobs[i][0].copy())
        if len(res) == len(jobs):
            break
    stats = pool.stats()
    pool.close()
    return res, stats


def test_pool_results_equal_solo(ag2):
    jobs = small_jobs(ag2, 6)
    res, st = run_pool(ag2, jobs, max_jobs=16)
    assert len(res) == len(jobs)
    assert st["job_steps"] == sum(s for _, _, s in jobs)
    for i, (ps, prm, steps) in enumerate(jobs):
        bad = close(res[i], solo(ag2, ps, prm, steps))
        assert bad is None, (i, bad)


def test_pool_continuous_admission(ag2):
    """Two slots, five jobs of different lengths: later jobs take the slots
    that finished jobs free, while the others keep running."""
    jobs = small_jobs(ag2, 5, seed=12)
    trace = []
    res, st = run_pool(ag2, jobs, max_jobs=2, trace=trace)
    assert len(res) == 5
    assert all(n <= 2 for n, _, _, _ in trace)
    # a round in which a job was still queued while two ran: batching, not a queue alone
    assert any(n == 2 and q > 0 for n, _, q, _ in trace)
    # and the last rounds ran with fewer jobs once the queue emptied
    assert st["done"] == 5 and st["queued"] == 0 and st["running"] == 0
    for i, (ps, prm, steps) in enumerate(jobs):
        assert close(res[i], solo(ag2, ps, prm, steps)) is None, i


def test_pool_memory_budget_admits_one_at_a_time(ag2):
    ps, prm, _ = small_jobs(ag2, 1, seed=13)[0]
    jobs = [(ps, prm, 2), (ps, prm, 3), (ps, prm, 2)]
    # room for one such job, not for two
    budget = int(1.5 * ag2.engine.context_bytes(ps.n, int((ps.type == 2).sum())))
    trace = []
    res, st = run_pool(ag2, jobs, max_jobs=8, budget=budget, trace=trace)
    assert len(res) == 3
    assert max(n for n, _, _, _ in trace) == 1
    assert st["hbm_used"] <= st["hbm_budget"]


def test_pool_rejects_job_over_budget(ag2):
    ps, prm, steps = small_jobs(ag2, 1)[0]
    pool = ag2.engine.Pool(max_jobs=2, hbm_budget=1 << 20)
    with pytest.raises(ag2.engine.EngineError):
        pool.submit(prm, ps, steps)
    pool.close()


def test_pool_ragged_sizes_equal_solo(ag2):
    """Jobs at and around the sort's tile size (2,048 keys) and the smallest
    sets, resident together: the batched sort maps each workgroup to its job
    by the tile offsets, so a job of one tile, one key or a partial last tile
    must come out exactly as alone."""
    U = ag2.units
    jobs = []
    for j, n in enumerate((1, 2, 5, 2047, 2048, 2049, 4097)):
        ps = ag2.ics.plummer(n, 1e10, 1.0, 20.0, seed=31, config=4, component=1, job=j)
        dt = 0.1 * U.MYR
        prec = ag2.engine.AG2_FP32_SPLIT if j % 2 == 0 else ag2.engine.AG2_FP64
        prm = ag2.engine.make_params(theta=0.5, e0=0.05 * U.KPC, massInH=1.0, split_threshold=800,
                                     precision=prec, eta=2.0, min_dt=dt, max_dt=dt)
        jobs.append((ps, prm, 3))
    res, st = run_pool(ag2, jobs, max_jobs=16)
    assert len(res) == len(jobs)
    for i, (ps, prm, steps) in enumerate(jobs):
        want = solo(ag2, ps, prm, steps)
        for f in FIELDS:
            assert np.array_equal(getattr(res[i], f), getattr(want, f)), (i, f)


def test_pool_jobs_match_oracle_runs(ag2, orc):
    """Jobs of the serving stream's recipe at its ends -- a 10,000-particle
    Plummer sphere with 20 % gas and a 200,000-particle Hernquist sphere --
    and a 40,000-particle gas-free Plummer sphere, resident together, in both
    precision forms: each job's final state equals the oracle's run of
    Simulation::run (Simulation.cpp:121-139, :166-356) over the same steps,
    at the FP64 run bar for FP64 jobs and at the split form's precision bar
    for FP32 jobs."""
    U = ag2.units
    E = ag2.engine
    specs = [("plummer", 10_000, True, E.AG2_FP64), ("hernquist", 200_000, False, E.AG2_FP32_SPLIT),
             ("plummer", 40_000, False, E.AG2_FP32_SPLIT), ("plummer", 10_000, True, E.AG2_FP32_SPLIT)]
    jobs = []
    for j, (kind, n, gas, prec) in enumerate(specs):
        if kind == "plummer":
            ps = ag2.ics.plummer(n, 3e9, 0.7, 20.0, seed=41, config=4, component=1, job=j)
        else:
            ps = ag2.ics.hernquist(n, 5e10, 1.5, 50.0, seed=41, config=4, component=2, job=j)
        if gas:
            k = np.random.default_rng(j).choice(n, n // 5, replace=False)
            ps.type[k] = U.TYPE_GAS
            ps.U[k] = U.u_from_temperature(1e4)
            ps.init_temperature()
        dt = 0.1 * U.MYR
        e0 = 0.05 * U.KPC
        prm = E.make_params(theta=0.5, e0=e0, massInH=ag2.ics.mass_in_h(ps), split_threshold=800, precision=prec,
                            eta=2.0, min_dt=dt, max_dt=dt)
        jobs.append((ps, prm, 3))
    res, st = run_pool(ag2, jobs, max_jobs=16)
    assert len(res) == len(jobs)
    assert st["max_resident"] == len(jobs)
    for i, (ps, prm, steps) in enumerate(jobs):
        a = ps.copy()
        n = orc.run(a, orc.params(prm.theta, prm.e0, prm.massInH, prm.split_threshold, nthreads=16), 2.0,
                    prm.min_dt, prm.max_dt, (steps - 0.5) * prm.min_dt)
        assert n == steps
        b = res[i]
        fp64 = prm.precision == E.AG2_FP64
        pos_tol, vel_tol = (1e-11, 1e-9) if fp64 else (1e-8, 1e-5)
        for f in ("x", "y", "z"):
            assert np.abs(getattr(b, f) - getattr(a, f)).max() / np.abs(a.x).max() < pos_tol, (i, f)
        for f in ("vx", "vy", "vz"):
            assert np.abs(getattr(b, f) - getattr(a, f)).max() / np.abs(a.vx).max() < vel_tol, (i, f)
        gas = ps.type == U.TYPE_GAS
        if gas.any():
            assert (np.abs(b.U[gas] - a.U[gas]) <= (1e-9 if fp64 else 1e-5) * np.abs(a.U[gas]).max()).all(), i


def test_pool_admits_and_retires_beside_the_round(ag2):
    """One slot, three jobs of two steps: while a job runs its last round the
    next one is admitted beside it (or, if that admission has not started by
    the time the round drains, at the start of the next call, into the then
    empty pool); either way it steps from the next round.  A finished job is
    downloaded during the next call, so it reads as done one call after its
    last step.  Every round steps exactly one job, and each result equals the
    job run alone."""
    E = ag2.engine
    jobs = [(ps, prm, 2) for ps, prm, _ in small_jobs(ag2, 3, seed=17)]
    pool = E.Pool(max_jobs=1)
    ids = [pool.submit(prm, ps, steps) for ps, prm, steps in jobs]
    stepped, states = [], []
    res = {}
    for _ in range(20):
        stepped.append(pool.step())
        # (a collected job is forgotten by the pool)
        st = [E.JOB_DONE if i in res else pool.status(jid)[0] for i, jid in enumerate(ids)]
        states.append(st)
        for i, jid in enumerate(ids):
            if i not in res and st[i] == E.JOB_DONE:
                res[i] = pool.result(jid, jobs[i][0].copy())
        if len(res) == len(jobs):
            break
    stats = pool.stats()
    pool.close()
    # rounds 1-2 step job 0, rounds 3-4 job 1, rounds 5-6 job 2; job 0 is
    # downloaded in round 3; a last call with nothing resident downloads job 2
    assert stepped == [1, 1, 1, 1, 1, 1, 0], stepped
    assert states[0] == [E.JOB_RUNNING, E.JOB_QUEUED, E.JOB_QUEUED]
    assert states[1][0] == E.JOB_RUNNING and states[1][2] == E.JOB_QUEUED
    assert states[2][:2] == [E.JOB_DONE, E.JOB_RUNNING]
    assert states[5][2] == E.JOB_RUNNING and states[6][2] == E.JOB_DONE
    assert stats["job_steps"] == 6 and stats["max_resident"] == 1
    for i, (ps, prm, steps) in enumerate(jobs):
        assert close(res[i], solo(ag2, ps, prm, steps)) is None, i
