"""ag2_direct_sum_noalloc against ag2_direct_sum through the C ABI.

The no-allocation entry point runs the direct-sum kernel on a fixed
per-device scratch array instead of per-call device allocations, so its
results must equal ag2_direct_sum's bit for bit: two calls in a row (the
second with a smaller particle set and fewer targets, so that stale scratch
from the first call would show), with and without the potential, on two
contexts that share the device's scratch, and a target count above the limit
is refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _engine(ag2, ps):
    E = ag2.engine
    prm = E.make_params(theta=0.5, e0=0.05 * ag2.units.KPC, massInH=1.0, split_threshold=100)
    eng = E.Engine(prm, ps.n)
    eng.upload(ps)
    return eng


def _both(eng, tg, pot):
    a = eng.direct_sum(tg, with_potential=pot)
    b = eng.direct_sum_noalloc(tg, with_potential=pot)
    for u, v in zip(a, b):
        if u is None:
            assert v is None
        else:
            assert np.array_equal(u, v)
    return b


def test_noalloc_equals_direct_sum_twice(ag2, orc):
    big = ag2.ics.config1_plummer(6000, seed=21)
    small = ag2.ics.config1_plummer(2500, seed=22)
    g = np.random.default_rng(5)
    e_big = _engine(ag2, big)
    tg1 = np.sort(g.choice(big.n, 5000, replace=False)).astype(np.int64)
    ax, ay, az, ph = _both(e_big, tg1, True)
    # a second call with a smaller input, on another context of the same device
    e_small = _engine(ag2, small)
    tg2 = np.sort(g.choice(small.n, 700, replace=False)).astype(np.int64)
    bx, by, bz, _ = _both(e_small, tg2, False)
    # and once more on the first context, fewer targets than its last call
    _both(e_big, tg1[:300], True)
    # the oracle's direct sum agrees to FP64 rounding (different source order)
    ox, oy, oz, op = orc.direct_sum(big.x, big.y, big.z, big.m, 0.05 * ag2.units.KPC, targets=tg1)
    den = np.sqrt(ox ** 2 + oy ** 2 + oz ** 2)
    assert (np.sqrt((ax - ox) ** 2 + (ay - oy) ** 2 + (az - oz) ** 2) / den).max() < 1e-12
    assert (np.abs(ph - op) / np.abs(op)).max() < 1e-12
    ox, oy, oz, _ = orc.direct_sum(small.x, small.y, small.z, small.m, 0.05 * ag2.units.KPC, targets=tg2)
    den = np.sqrt(ox ** 2 + oy ** 2 + oz ** 2)
    assert (np.sqrt((bx - ox) ** 2 + (by - oy) ** 2 + (bz - oz) ** 2) / den).max() < 1e-12
    e_big.close()
    e_small.close()


def test_noalloc_refuses_above_limit(ag2):
    E = ag2.engine
    ps = ag2.ics.config1_plummer(E.DIRECT_NOALLOC_MAX + 10, seed=23)
    eng = _engine(ag2, ps)
    tg = np.arange(E.DIRECT_NOALLOC_MAX + 1, dtype=np.int64)
    with pytest.raises(E.EngineError) as ei:
        eng.direct_sum_noalloc(tg)
    assert ei.value.code == E.AG2_ERR_CAPACITY
    # the limit itself is accepted and equals the allocating entry point
    _both(eng, tg[:E.DIRECT_NOALLOC_MAX], False)
    eng.close()
