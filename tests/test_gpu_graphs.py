"""Replayed build segments (HIP graphs, ag2_params.graph_launch = 1) against
direct launches (the default): every step of a run must leave bit-identical
state, across re-uploads with another particle count, a parameter change
and a node-capacity growth, each of which drops the captured segments; and
the serving pool, whose jobs replay their segments on shared streams, must
match solo runs with direct launches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "h", "rho", "P", "U", "dUdt")


def same(a, b):
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f), equal_nan=True), f


def params(ag2, ps, graphs, thr=100, precision=0):
    U = ag2.units
    dt = 0.5 * U.MYR
    return ag2.engine.make_params(theta=0.5, e0=0.1 * U.KPC, massInH=ag2.ics.mass_in_h(ps), split_threshold=thr,
                                  precision=precision, eta=2.0, min_dt=dt, max_dt=dt, graph_launch=graphs)


def trajectory(ag2, sets, graphs, precision=0):
    """Steps over a sequence of (particles, steps, split_threshold) on one
    context: each entry re-uploads (or, with particles None, changes the
    split threshold in place); the state after every step is returned."""
    E = ag2.engine
    cap = max(ps.n for ps, _, _ in sets if ps is not None)
    first = sets[0][0]
    eng = E.Engine(params(ag2, first, graphs, sets[0][2], precision), cap)
    out = []
    last = first
    for ps, steps, thr in sets:
        if ps is not None:
            eng.upload(ps)
            eng.init_forces()
            last = ps
        else:
            eng.set_params(params(ag2, last, graphs, thr, precision))
        for _ in range(steps):
            eng.step()
            eng.sync()
            out.append(eng.download(last.copy()))
    eng.close()
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_steps_equal_direct(ag2, precision):
    ics = ag2.ics
    a = ics.shuffled(ics.disk_galaxy(3000, 2000, 3000, seed=41), 1)
    b = ics.shuffled(ics.disk_galaxy(1000, 900, 700, seed=42), 2)
    sets = [(a, 4, 100), (None, 3, 800), (b, 3, 800), (a, 2, 800)]
    g = trajectory(ag2, sets, 1, precision)
    e = trajectory(ag2, sets, 0, precision)
    assert len(g) == len(e) == 12
    for k, (x, y) in enumerate(zip(g, e)):
        same(x, y)


def test_graph_steps_after_node_growth(ag2):
    """A context created for a small set, then given a set whose tree needs
    more nodes than the first allocation: the build grows the node arrays
    (and drops the captured segments) and the run still equals direct launches."""
    ics = ag2.ics
    small = ics.config1_plummer(64, seed=5)
    big = ics.shuffled(ics.disk_galaxy(4000, 3000, 3000, seed=43), 3)
    sets = [(small, 2, 100), (big, 3, 100)]
    # capacity is the larger count; the node arrays start from 2 * capacity
    same_len = trajectory(ag2, sets, 1), trajectory(ag2, sets, 0)
    for x, y in zip(*same_len):
        same(x, y)


def test_pool_with_graphs_equals_direct_solo(ag2):
    E = ag2.engine
    ics = ag2.ics
    jobs = [ics.shuffled(ics.disk_galaxy(800 + 100 * k, 500, 400 + 50 * k, seed=50 + k), k) for k in range(4)]
    pool = E.Pool(max_jobs=3)
    ids = [pool.submit(params(ag2, ps, 1, 800, 1), ps, 5) for ps in jobs]
    res = {}
    for _ in range(100):
        pool.step()
        for k, jid in enumerate(ids):
            if k not in res and pool.status(jid)[0] == E.JOB_DONE:
                res[k] = pool.result(jid, jobs[k].copy())
        if len(res) == len(jobs):
            break
    pool.close()
    assert len(res) == len(jobs)
    for k, ps in enumerate(jobs):
        solo = trajectory(ag2, [(ps, 5, 800)], 0, 1)[-1]
        same(res[k], solo)


def test_sort_byte_range_leaves_steps_unchanged(ag2, orc):
    """Direct launches sort six bytes of word 0 and rank the runs sharing them;
    a run too long for the ranking sends the build back to all eight bytes,
    and a build with long runs has the next one sort all eight.  Graph
    launches always sort all eight.  A Plummer sphere with 1,500 particles
    inside one level-16 cell (one such long run), stepped both ways, leaves
    bit-identical state after every step."""
    import test_gpu_parity as tp
    ps = tp.prefix_clump(ag2.ics, orc, 4000, 1500, seed=73)
    sets = [(ps, 4, 100), (None, 3, 800)]
    g = trajectory(ag2, sets, 1)
    e = trajectory(ag2, sets, 0)
    assert len(g) == len(e) == 7
    for x, y in zip(g, e):
        same(x, y)
