"""Particle file formats of the reference: `.age` (read and write) and Gadget-2
format-1 snapshots (read).

.age layout (DataManager.cpp:216-258 writer, :534-574 reader): a 40-byte
header {int32 numParticles[3] (stars, gas, dark matter), 4 pad bytes,
double deltaTime, endTime, currentTime} followed by one 94-byte record per
particle {double pos[3], vel[3], mass, T, P, visualDensity, U; uint8 type,
galaxyPart; uint32 id}, little-endian, in array order.

Gadget-2 (DataManager.cpp:811-1324): header block, float32 POS and VEL, uint32
ID, float32 MASS if any populated type has massarr == 0, float32 U for gas;
units kpc, km/s, 1e10 Msun and (km/s)^2, converted to SI with Units.h.
"""
import struct

import numpy as np

from . import units as U
from .ics import ParticleSet

AGE_HEADER = struct.Struct("<3i4x3d")
AGE_RECORD = np.dtype([("pos", "<f8", 3), ("vel", "<f8", 3), ("mass", "<f8"), ("T", "<f8"),
                       ("P", "<f8"), ("visualDensity", "<f8"), ("U", "<f8"), ("type", "u1"),
                       ("galaxyPart", "u1"), ("id", "<u4")])
assert AGE_HEADER.size == 40 and AGE_RECORD.itemsize == 94


def write_age(path, ps, delta_time=0.0, end_time=0.0, current_time=0.0, visual_density=None,
              galaxy_part=None):
    counts = [int(np.count_nonzero(ps.type == t)) for t in (U.TYPE_STAR, U.TYPE_GAS, U.TYPE_DM)]
    rec = np.zeros(ps.n, dtype=AGE_RECORD)
    rec["pos"] = np.stack([ps.x, ps.y, ps.z], axis=1)
    rec["vel"] = np.stack([ps.vx, ps.vy, ps.vz], axis=1)
    rec["mass"], rec["T"], rec["P"], rec["U"] = ps.m, ps.T, ps.P, ps.U
    rec["visualDensity"] = 0.0 if visual_density is None else visual_density
    rec["type"] = ps.type
    rec["galaxyPart"] = 1 if galaxy_part is None else galaxy_part
    rec["id"] = ps.id
    with open(path, "wb") as f:
        f.write(AGE_HEADER.pack(*counts, delta_time, end_time, current_time))
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
