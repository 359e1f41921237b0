"""Snapshot writers (DataManager::saveData, DataManager.cpp:86-424) against the
bytes the compiled reference writes (tests/golden/snapshots.npz), and the
readers against the writers."""
import os

import numpy as np
import pytest

from conftest import ROOT

FORMATS = ("ag", "agc", "age", "gadget", "hdf5")


def load(ag2):
    z = np.load(os.path.join(ROOT, "tests", "golden", "snapshots.npz"), allow_pickle=False)
    ps = ag2.ics.ParticleSet(len(z["in_x"]))
    for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu", "type", "T", "P", "id"):
        getattr(ps, f)[:] = z["in_" + f]
    extras = {k: z["in_" + k] for k in ("galaxyPart", "visualDensity", "sfr")}
    return z, ps, extras


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("cut", [False, True])
def test_save_data_matches_reference_bytes(ag2, tmp_path, fmt, cut):
    z, ps, ex = load(ag2)
    n_out = 150 if cut else ps.n
    path = ag2.ageio.save_data(str(tmp_path), ps, fmt, 7, n_out, float(z["delta_time"]), float(z["end_time"]),
                               float(z["current_time"]), visual_density=ex["visualDensity"], sfr=ex["sfr"],
                               galaxy_part=ex["galaxyPart"])
    assert os.path.basename(path) == "7." + fmt
    got = np.fromfile(path, np.uint8)
    want = z["%s_%d" % (fmt, n_out)]
    assert got.size == want.size
    assert np.array_equal(got, want)


def test_readers_round_trip(ag2, tmp_path):
    z, ps, ex = load(ag2)
    io = ag2.ageio
    kw = dict(visual_density=ex["visualDensity"], sfr=ex["sfr"], galaxy_part=ex["galaxyPart"])
    a, hdr, e = io.read_ag(io.save_data(str(tmp_path), ps, "ag", 1, None, 1.0, 2.0, 3.0, **kw))
    assert hdr == {"deltaTime": 1.0, "endTime": 2.0, "currentTime": 3.0}
    for f in ("x", "y", "z", "m", "T", "type", "id"):
        assert np.array_equal(getattr(a, f), getattr(ps, f)), f
    for k in ("visualDensity", "sfr", "galaxyPart"):
        assert np.array_equal(e[k], ex[k]), k
    c, _, e = io.read_agc(io.save_data(str(tmp_path), ps, "agc", 2, None, **kw))
    assert np.array_equal(c.x, ps.x.astype(np.float32).astype(np.float64))
    assert np.array_equal(e["visualDensity"], ex["visualDensity"].astype(np.float32).astype(np.float64))
    b, _ = io.read_age(io.save_data(str(tmp_path), ps, "age", 3, None, **kw))
    for f in ("x", "vx", "m", "T", "P", "U", "type", "id"):
        assert np.array_equal(getattr(b, f), getattr(ps, f)), f


def test_gadget_writer_reader_round_trip(ag2, tmp_path):
    """Stars as bulge and dark matter as halo (the pairings whose blocks the
    reference's reader accepts back): every particle returns, to float32
    precision, with its id."""
    z, ps, ex = load(ag2)
    U = ag2.units
    gp = np.where(ps.type == U.TYPE_DM, 3, 2).astype(np.uint8)
    path = ag2.ageio.save_data(str(tmp_path), ps, "gadget", 0, None, galaxy_part=gp)
    g = ag2.ageio.read_gadget(path)
    assert g.n == ps.n
    order = np.argsort(ps.id)
    back = np.argsort(g.id)
    assert np.array_equal(np.sort(g.id), np.sort(ps.id))
    assert np.allclose(g.x[back], ps.x[order], rtol=1e-6, atol=1e-6 * np.abs(ps.x).max())
    assert np.allclose(g.m[back], ps.m[order], rtol=1e-6)


def test_unknown_format_rejected(ag2, tmp_path):
    _, ps, _ = load(ag2)
    with pytest.raises(ValueError):
        ag2.ageio.save_data(str(tmp_path), ps, "vtk", 0)
