"""Particle file formats of the reference: `.ag`, `.agc`, `.age` (read and
write), Gadget-2 format-1 snapshots (read and write), and the snapshot writer
`save_data` (DataManager::saveData, DataManager.cpp:86-424).

.ag and .agc (DataManager.cpp:119-215 writers, :446-530 readers) share the
40-byte header of .age; .ag records are 62 bytes {double pos[3], mass, T,
visualDensity, sfr; uint8 type, galaxyPart; uint32 id}, .agc records 26 bytes
{float pos[3], visualDensity, sfr, T; uint8 type, galaxyPart}.

.age layout (DataManager.cpp:216-258 writer, :534-574 reader): a 40-byte
header {int32 numParticles[3] (stars, gas, dark matter), 4 pad bytes,
double deltaTime, endTime, currentTime} followed by one 94-byte record per
particle {double pos[3], vel[3], mass, T, P, visualDensity, U; uint8 type,
galaxyPart; uint32 id}, little-endian, in array order.

Gadget-2 (DataManager.cpp:811-1324): header block, float32 POS and VEL, uint32
ID, float32 MASS if any populated type has massarr == 0, float32 U for gas;
units kpc, km/s, 1e10 Msun and (km/s)^2, converted to SI with Units.h.
"""
import os
import struct

import numpy as np

from . import units as U
from .ics import ParticleSet

AGE_HEADER = struct.Struct("<3i4x3d")
AGE_RECORD = np.dtype([("pos", "<f8", 3), ("vel", "<f8", 3), ("mass", "<f8"), ("T", "<f8"),
                       ("P", "<f8"), ("visualDensity", "<f8"), ("U", "<f8"), ("type", "u1"),
                       ("galaxyPart", "u1"), ("id", "<u4")])
AG_RECORD = np.dtype([("pos", "<f8", 3), ("mass", "<f8"), ("T", "<f8"), ("visualDensity", "<f8"),
                      ("sfr", "<f8"), ("type", "u1"), ("galaxyPart", "u1"), ("id", "<u4")])
AGC_RECORD = np.dtype([("pos", "<f4", 3), ("visualDensity", "<f4"), ("sfr", "<f4"), ("T", "<f4"),
                       ("type", "u1"), ("galaxyPart", "u1")])
GADGET_HEADER = struct.Struct("<6I6d2d2i6I2i4d2i6Ii60x")
assert AGE_HEADER.size == 40 and AGE_RECORD.itemsize == 94
assert AG_RECORD.itemsize == 62 and AGC_RECORD.itemsize == 26 and GADGET_HEADER.size == 256


def _header(ps, delta_time, end_time, current_time):
    counts = [int(np.count_nonzero(ps.type == t)) for t in (U.TYPE_STAR, U.TYPE_GAS, U.TYPE_DM)]
    return AGE_HEADER.pack(*counts, delta_time, end_time, current_time)


def _extra(ps, a, default):
    return np.full(ps.n, default) if a is None else np.asarray(a)


def write_ag(path, ps, delta_time=0.0, end_time=0.0, current_time=0.0, visual_density=None, sfr=None,
             galaxy_part=None):
    rec = np.zeros(ps.n, dtype=AG_RECORD)
    rec["pos"] = np.stack([ps.x, ps.y, ps.z], axis=1)
    rec["mass"], rec["T"] = ps.m, ps.T
    rec["visualDensity"] = _extra(ps, visual_density, 0.0)
    rec["sfr"] = _extra(ps, sfr, 0.0)
    rec["type"] = ps.type
    rec["galaxyPart"] = _extra(ps, galaxy_part, 1)
    rec["id"] = ps.id
    with open(path, "wb") as f:
        f.write(_header(ps, delta_time, end_time, current_time))
        f.write(rec.tobytes())


def write_agc(path, ps, delta_time=0.0, end_time=0.0, current_time=0.0, visual_density=None, sfr=None,
              galaxy_part=None):
    rec = np.zeros(ps.n, dtype=AGC_RECORD)
    rec["pos"] = np.stack([ps.x, ps.y, ps.z], axis=1).astype(np.float32)
    rec["visualDensity"] = _extra(ps, visual_density, 0.0).astype(np.float32)
    rec["sfr"] = _extra(ps, sfr, 0.0).astype(np.float32)
    rec["T"] = ps.T.astype(np.float32)
    rec["type"] = ps.type
    rec["galaxyPart"] = _extra(ps, galaxy_part, 1)
    with open(path, "wb") as f:
        f.write(_header(ps, delta_time, end_time, current_time))
        f.write(rec.tobytes())


def _read_records(path, dtype):
    with open(path, "rb") as f:
        hdr = AGE_HEADER.unpack(f.read(AGE_HEADER.size))
        n = hdr[0] + hdr[1] + hdr[2]
        rec = np.frombuffer(f.read(n * dtype.itemsize), dtype=dtype, count=n)
    return rec, {"deltaTime": hdr[3], "endTime": hdr[4], "currentTime": hdr[5]}


def read_ag(path):
    """Returns (ParticleSet, header, extras) with extras visualDensity, sfr, galaxyPart."""
    rec, hdr = _read_records(path, AG_RECORD)
    ps = ParticleSet(len(rec))
    ps.x[:], ps.y[:], ps.z[:] = rec["pos"].T
    ps.m[:], ps.T[:] = rec["mass"], rec["T"]
    ps.type[:] = rec["type"]
    ps.id[:] = rec["id"]
    extras = {k: rec[k].copy() for k in ("visualDensity", "sfr", "galaxyPart")}
    return ps, hdr, extras


def read_agc(path):
    """Returns (ParticleSet, header, extras); FP32 fields widened to FP64."""
    rec, hdr = _read_records(path, AGC_RECORD)
    ps = ParticleSet(len(rec))
    ps.x[:], ps.y[:], ps.z[:] = rec["pos"].astype(np.float64).T
    ps.T[:] = rec["T"]
    ps.type[:] = rec["type"]
    extras = {"visualDensity": rec["visualDensity"].astype(np.float64), "sfr": rec["sfr"].astype(np.float64),
              "galaxyPart": rec["galaxyPart"].copy()}
    return ps, hdr, extras


def write_age(path, ps, delta_time=0.0, end_time=0.0, current_time=0.0, visual_density=None,
              galaxy_part=None):
    rec = np.zeros(ps.n, dtype=AGE_RECORD)
    rec["pos"] = np.stack([ps.x, ps.y, ps.z], axis=1)
    rec["vel"] = np.stack([ps.vx, ps.vy, ps.vz], axis=1)
    rec["mass"], rec["T"], rec["P"], rec["U"] = ps.m, ps.T, ps.P, ps.U
    rec["visualDensity"] = _extra(ps, visual_density, 0.0)
    rec["type"] = ps.type
    rec["galaxyPart"] = _extra(ps, galaxy_part, 1)
    rec["id"] = ps.id
    with open(path, "wb") as f:
        f.write(_header(ps, delta_time, end_time, current_time))
        f.write(rec.tobytes())


def read_age(path):
    with open(path, "rb") as f:
        hdr = AGE_HEADER.unpack(f.read(AGE_HEADER.size))
        n = hdr[0] + hdr[1] + hdr[2]
        rec = np.frombuffer(f.read(n * AGE_RECORD.itemsize), dtype=AGE_RECORD, count=n)
    ps = ParticleSet(n)
    ps.x[:], ps.y[:], ps.z[:] = rec["pos"].T
    ps.vx[:], ps.vy[:], ps.vz[:] = rec["vel"].T
    ps.m[:], ps.T[:], ps.P[:], ps.U[:] = rec["mass"], rec["T"], rec["P"], rec["U"]
    ps.type[:] = rec["type"]
    ps.id[:] = rec["id"]
    return ps, {"deltaTime": hdr[3], "endTime": hdr[4], "currentTime": hdr[5]}


def _block(f, dtype, count):
    (size,) = struct.unpack("<I", f.read(4))
    data = np.frombuffer(f.read(size), dtype=dtype)
    (size2,) = struct.unpack("<I", f.read(4))
    if size != size2:
        raise ValueError("gadget block markers differ")
    if count is not None and data.size != count:
        raise ValueError("gadget block has %d entries, expected %d" % (data.size, count))
    return data


def read_gadget(path):
    with open(path, "rb") as f:
        (size,) = struct.unpack("<I", f.read(4))
        if size != 256:
            raise ValueError("not a gadget-2 format-1 file")
        hdr = f.read(256)
        f.read(4)
        npart = struct.unpack("<6I", hdr[0:24])
        massarr = struct.unpack("<6d", hdr[24:72])
        n = sum(npart)
        pos = _block(f, "<f4", 3 * n).reshape(n, 3)
        vel = _block(f, "<f4", 3 * n).reshape(n, 3)
        ids = _block(f, "<u4", n)
        individual = any(massarr[t] < 1e-10 and npart[t] != 0 for t in range(6))
        masses = _block(f, "<f4", n) if individual else None
        u = _block(f, "<f4", npart[0]) if npart[0] > 0 else None
    ps = ParticleSet(n)
    k = 0
    kind = {0: U.TYPE_GAS, 1: U.TYPE_DM, 2: U.TYPE_STAR, 3: U.TYPE_STAR, 4: U.TYPE_STAR, 5: U.TYPE_STAR}
    for t in range(6):
        sl = slice(k, k + npart[t])
        ps.type[sl] = kind[t]
        if individual:
            ps.m[sl] = masses[sl].astype(np.float64) * U.MSUN * 1e10
        else:
            ps.m[sl] = massarr[t] * U.MSUN * 1e10
        k += npart[t]
    ps.x[:] = pos[:, 0].astype(np.float64) * U.KPC
    ps.y[:] = pos[:, 1].astype(np.float64) * U.KPC
    ps.z[:] = pos[:, 2].astype(np.float64) * U.KPC
    ps.vx[:] = vel[:, 0].astype(np.float64) * U.KMS
    ps.vy[:] = vel[:, 1].astype(np.float64) * U.KMS
    ps.vz[:] = vel[:, 2].astype(np.float64) * U.KMS
    ps.id[:] = ids
    if u is not None:
        ps.U[:npart[0]] = u.astype(np.float64) * 1e6
    return ps.init_temperature()


def write_gadget(path, ps, current_time=0.0, galaxy_part=None):
    """Gadget-2 format-1 snapshot as the reference writes it (DataManager.cpp:
    263-415), including its list assignment: disk stars (galaxyPart 1, type 1)
    and gas go to the first list, bulge stars (galaxyPart 2, type 1) to the
    second, halo dark matter (galaxyPart 3, type 3) to the third; npart counts
    them as gadget types 2/0, 3 and 1, and any other particle as type 0
    without data.  POS (kpc), VEL, ID, MASS (1e10 Msun) as float32 / uint32
    in list order, then U / 1e6 of every first-list particle.  VEL holds the
    positions in units of 1000 m, as the reference writes it.  When
    disk stars are present that last block is longer than npart[0], so the
    reference's reader (and read_gadget) rejects such a file."""
    gp = _extra(ps, galaxy_part, 1)
    first, second, third = [], [], []
    npart = [0] * 6
    for i in range(ps.n):
        g = 0
        t = int(ps.type[i])
        if gp[i] == 1 and t == U.TYPE_STAR:
            g = 2
            first.append(i)
        if gp[i] == 3 and t == U.TYPE_DM:
            g = 1
            third.append(i)
        if gp[i] == 2 and t == U.TYPE_STAR:
            g = 3
            second.append(i)
        if t == U.TYPE_GAS:
            g = 0
            first.append(i)
        npart[g] += 1
    order = np.array(first + second + third, dtype=np.int64)
    # the reference passes the unit as a float parameter (DataManager.cpp:333), so kpc is rounded to FP32 first
    kpc, kms = float(np.float32(U.KPC)), float(np.float32(U.KMS))
    hdr = GADGET_HEADER.pack(*npart, *([0.0] * 6), current_time, 0.0, 0, 0, *npart, 0, 1, 0.0, 0.0, 0.0, 1.0,
                             0, 0, *([0] * 6), 0)

    def block(f, a):
        b = a.tobytes()
        f.write(struct.pack("<I", len(b)))
        f.write(b)
        f.write(struct.pack("<I", len(b)))

    with open(path, "wb") as f:
        block(f, np.frombuffer(hdr, np.uint8))
        block(f, np.stack([ps.x[order] / kpc, ps.y[order] / kpc, ps.z[order] / kpc], 1).astype(np.float32))
        # the reference's VEL block goes through the same position helper (DataManager.cpp:332-339, :362-369):
        # it holds positions divided by 1000
        block(f, np.stack([ps.x[order] / kms, ps.y[order] / kms, ps.z[order] / kms], 1).astype(np.float32))
        block(f, ps.id[order].astype(np.uint32))
        block(f, (ps.m[order] / (U.MSUN * 1e10)).astype(np.float32))
        if first:
            block(f, (ps.U[np.array(first)] / 1e6).astype(np.float32))


FORMATS = {"ag": ".ag", "agc": ".agc", "age": ".age", "hdf5": ".hdf5", "gadget": ".gadget"}


def save_data(out_dir, ps, fmt, time_step, n_output=None, delta_time=0.0, end_time=0.0, current_time=0.0,
              visual_density=None, sfr=None, galaxy_part=None):
    """DataManager::saveData (DataManager.cpp:86-424): the first n_output
    particles (the reference shuffles its array once, so they are a random
    subset) into <out_dir>/<time_step><ending>.  The reference's hdf5 branch
    writes nothing into the file it opens (:259-262), and neither does this
    one.  Returns the file path."""
    if fmt not in FORMATS:
        raise ValueError("unknown output data format: %s" % fmt)
    os.makedirs(out_dir, exist_ok=True)
    n = ps.n if n_output is None else min(int(n_output), ps.n)
    sub = ps.take(np.arange(n)) if n < ps.n else ps
    cut = lambda a: None if a is None else np.asarray(a)[:n]   # noqa: E731
    vd, sf, gp = cut(visual_density), cut(sfr), cut(galaxy_part)
    path = os.path.join(out_dir, "%d%s" % (int(time_step), FORMATS[fmt]))
    if fmt == "ag":
        write_ag(path, sub, delta_time, end_time, current_time, vd, sf, gp)
    elif fmt == "agc":
        write_agc(path, sub, delta_time, end_time, current_time, vd, sf, gp)
    elif fmt == "age":
        write_age(path, sub, delta_time, end_time, current_time, vd, gp)
    elif fmt == "gadget":
        write_gadget(path, sub, current_time, gp)
    else:
        open(path, "wb").close()
    return path
