"""ctypes access to the CPU oracle (liboracle.so) and, where it was built in
this container, to the compiled reference (_ref/libag2ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always to check results or time the CPU path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libag2ref.so")
G_SI = 6.67430e-11

_d = C.POINTER(C.c_double)
_u8 = C.POINTER(C.c_uint8)
_i64 = C.POINTER(C.c_int64)
_i32 = C.POINTER(C.c_int32)
_u64 = C.POINTER(C.c_uint64)


class Params(C.Structure):
    _fields_ = [("theta", C.c_double), ("e0", C.c_double), ("massInH", C.c_double), ("G", C.c_double),
                ("split_threshold", C.c_int64), ("max_level", C.c_int32), ("nthreads", C.c_int32)]


class Particles(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(f, _d) for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu")] + \
               [("type", _u8)] + [(f, _d) for f in ("ax", "ay", "az", "h", "rho", "P", "T", "dUdt")] + \
               [("active", _u8)]


class Stats(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_in_tree", C.c_int64), ("interactions", C.c_int64),
                ("sph_interactions", C.c_int64), ("groups_fired", C.c_int64), ("R", C.c_double)]


class RunParams(C.Structure):
    _fields_ = [("eta", C.c_double), ("min_dt", C.c_double), ("max_dt", C.c_double),
                ("end_time", C.c_double), ("H0", C.c_double), ("max_steps", C.c_int32),
                ("cooling", C.c_int32), ("star_formation", C.c_int32), ("sfr_seed", C.c_uint64), ("sfr", _d)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_tree_width.restype = C.c_double
        L.orc_tree_width.argtypes = [C.c_int64, _d, _d, _d]
        L.orc_force_step.argtypes = [C.POINTER(Particles), C.POINTER(Params), C.POINTER(Stats)]
        L.orc_force_step_scale.argtypes = [C.POINTER(Particles), C.POINTER(Params), C.POINTER(Stats), _d]
        L.orc_force_step_moments.argtypes = [C.POINTER(Particles), C.POINTER(Params), C.POINTER(Stats), _d,
                                             C.c_int64, _d, _d, _d, _d]
        L.orc_tree_export.restype = C.c_int64
        L.orc_tree_export.argtypes = [C.POINTER(Particles), C.POINTER(Params), C.c_int64, _i32, _i64, _i64,
                                      _d, _d, _d, _d, _i64, _i64]
        L.orc_morton_keys.argtypes = [C.c_int64, _d, _d, _d, C.c_double, _u64, _u64, _i32]
        L.orc_direct_sum.argtypes = [C.c_int64, _d, _d, _d, _d, C.c_int64, _i64, C.c_double, C.c_double,
                                     C.c_int32, _d, _d, _d, _d]
        L.orc_run.argtypes = [C.POINTER(Particles), C.POINTER(Params), C.POINTER(RunParams)]
        L.orc_visual_density.argtypes = [C.POINTER(Particles), C.POINTER(Params), C.c_double, _d, _d]
        L.orc_sfr_uniform.restype = C.c_double
        L.orc_sfr_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_int64]
        L.orc_gas_sources.argtypes = [C.c_int64, _u8, _d, _d, _d, C.c_int32, C.c_int32, C.c_uint64, C.c_int64,
                                      C.c_int32, _u8, _d, _d, _d]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        arr = [C.c_int64] + [_d] * 9 + [_u8] + [_d] * 8
        L.ref_force_step.argtypes = arr + [C.c_double, C.c_double, C.c_double, C.c_int, _d]
        L.ref_tree_export.restype = C.c_int64
        L.ref_tree_export.argtypes = [C.c_int64] + [_d] * 9 + [_u8, C.c_int, C.c_int64, _i32, _i64, _i64,
                                                              _d, _d, _d, _d, _i64, _d]
        L.ref_run.argtypes = arr + [C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_int]
        L.ref_gas_sources.argtypes = [C.c_int64, _d, _d, _d, _u8, _d, _d, C.c_int, C.c_int, C.c_uint, _d]
        L.ref_visual_density.argtypes = [C.c_int64, _d, _d, _d, _d, _u8, C.c_int, C.c_double, _d, _d]
        L.ref_save_data.argtypes = [C.c_int64] + [_d] * 10 + [_u8, _u8, C.POINTER(C.c_uint32), _d, _d,
                                                              C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_double,
                                                              C.c_double, C.c_double]
        _ref = L
    return _ref


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _c(a, dt=np.float64):
    a = np.ascontiguousarray(a, dtype=dt)
    return a


def _particles(ps, active=None):
    """Build the orc_particles view; arrays must stay alive (kept on the struct)."""
    keep = {}
    for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu", "ax", "ay", "az", "h", "rho", "P", "T", "dUdt"):
        a = getattr(ps, f)
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise ValueError("field %s must be contiguous float64" % f)
        keep[f] = a
    if ps.type.dtype != np.uint8:
        raise ValueError("type must be uint8")
    keep["type"] = ps.type
    P = Particles()
    P.n = ps.n
    for f, a in keep.items():
        setattr(P, f, _p(a, _u8 if f == "type" else _d))
    if active is not None:
        act = _c(active, np.uint8)
        keep["active"] = act
        P.active = _p(act, _u8)
    else:
        P.active = C.cast(None, _u8)
    P._keep = keep
    return P


def params(theta, e0, massInH=1.0, split_threshold=100, G=G_SI, max_level=126, nthreads=0):
    return Params(theta, e0, massInH, G, split_threshold, max_level, nthreads)


def tree_width(x, y, z):
    x, y, z = _c(x), _c(y), _c(z)
    return lib().orc_tree_width(len(x), _p(x), _p(y), _p(z))


class DepthError(RuntimeError):
    """Distinct particles share more levels than prm.max_level (orc_* -2)."""


def force_step(ps, prm, active=None, scale=None, moments=None):
    """Build + SPH density + forces on ps in place; returns stats dict.
    scale (n doubles, optional) receives each particle's sum of contribution
    magnitudes (orc_force_step_scale); moments (optional: dict of mass, com,
    gas, mvel per non-empty node in export order, com and mvel as (nn, 3))
    replace the tree's own node moments (orc_force_step_moments)."""
    P = _particles(ps, active)
    st = Stats()
    if scale is not None:
        assert scale.dtype == np.float64 and scale.flags.c_contiguous and len(scale) == ps.n
    if moments is not None:
        nn = len(moments["mass"])
        arrs = [_c(moments["mass"]), _c(np.asarray(moments["com"]).reshape(-1)), _c(moments["gas"]),
                _c(np.asarray(moments["mvel"]).reshape(-1))]
        assert all(len(a) == k * nn for a, k in zip(arrs, (1, 3, 1, 3)))
        rc = lib().orc_force_step_moments(C.byref(P), C.byref(prm), C.byref(st),
                                          None if scale is None else _p(scale), nn, *[_p(a) for a in arrs])
        if rc == -1:
            raise ValueError("node count differs from the oracle's tree")
    elif scale is None:
        rc = lib().orc_force_step(C.byref(P), C.byref(prm), C.byref(st))
    else:
        rc = lib().orc_force_step_scale(C.byref(P), C.byref(prm), C.byref(st), _p(scale))
    if rc == -2:
        raise DepthError("distinct particles share more than %d levels" % prm.max_level)
    return {f: getattr(st, f) for f, _ in Stats._fields_}


def tree_export(ps, prm, cap=None):
    n = ps.n
    cap = 4 * n + 64 if cap is None else cap
    out = dict(level=np.zeros(cap, np.int32), first=np.zeros(cap, np.int64), count=np.zeros(cap, np.int64),
               mass=np.zeros(cap), com=np.zeros(3 * cap), gas=np.zeros(cap), mvel=np.zeros(3 * cap),
               perm=np.full(n, -1, np.int64))
    nin = C.c_int64(0)
    P = _particles(ps)
    k = lib().orc_tree_export(C.byref(P), C.byref(prm), cap, _p(out["level"], _i32), _p(out["first"], _i64),
                              _p(out["count"], _i64), _p(out["mass"]), _p(out["com"]), _p(out["gas"]),
                              _p(out["mvel"]), _p(out["perm"], _i64), C.byref(nin))
    if k == -2:
        raise DepthError("distinct particles share more than %d levels" % prm.max_level)
    if k < 0:   # deep trees: chains of single-child cells
        return tree_export(ps, prm, 4 * cap)
    res = {f: v[:k] for f, v in out.items() if f not in ("com", "mvel", "perm")}
    res["com"] = out["com"][:3 * k].reshape(k, 3)
    res["mvel"] = out["mvel"][:3 * k].reshape(k, 3)
    res["perm"] = out["perm"][:nin.value]
    res["n_in_tree"] = nin.value
    return res


def morton_keys(x, y, z, R):
    x, y, z = _c(x), _c(y), _c(z)
    n = len(x)
    k1 = np.zeros(n, np.uint64)
    k2 = np.zeros(n, np.uint64)
    gap = np.zeros(n, np.int32)
    lib().orc_morton_keys(n, _p(x), _p(y), _p(z), R, _p(k1, _u64), _p(k2, _u64), _p(gap, _i32))
    return k1, k2, gap


def direct_sum(x, y, z, m, e0, targets=None, G=G_SI, nthreads=0):
    x, y, z, m = _c(x), _c(y), _c(z), _c(m)
    n = len(x)
    t = None if targets is None else _c(targets, np.int64)
    nt = n if t is None else len(t)
    ax, ay, az, pot = (np.zeros(nt) for _ in range(4))
    lib().orc_direct_sum(n, _p(x), _p(y), _p(z), _p(m), nt, None if t is None else _p(t, _i64), e0, G,
                         nthreads, _p(ax), _p(ay), _p(az), _p(pot))
    return ax, ay, az, pot


def run(ps, prm, eta, min_dt, max_dt, end_time, H0=0.0, max_steps=1 << 30, cooling=0, star_formation=0,
        sfr_seed=0, sfr=None):
    """Simulation::init + run on ps in place; sfr (n doubles) receives Particle::sfr."""
    P = _particles(ps)
    rp = RunParams(eta, min_dt, max_dt, end_time, H0, max_steps, int(cooling), int(star_formation), int(sfr_seed),
                   None if sfr is None else _p(sfr))
    n = lib().orc_run(C.byref(P), C.byref(prm), C.byref(rp))
    if n == -2:
        raise DepthError("distinct particles share more than %d levels" % prm.max_level)
    return n


def gas_sources(ps, dt, cooling, star_formation, seed=0, step=0, ueuler=False, active=None, sfr=None):
    """Cooling / SFR / Ueuler of the second-kick loop on ps in place."""
    dt = _c(dt)
    act = None if active is None else np.ascontiguousarray(active, np.uint8)
    lib().orc_gas_sources(ps.n, None if act is None else _p(act, _u8), _p(dt), _p(ps.T), _p(ps.rho), int(cooling),
                          int(star_formation), int(seed), int(step), int(ueuler), _p(ps.type, _u8), _p(ps.U),
                          _p(ps.dUdt), None if sfr is None else _p(sfr))


def sfr_uniform(seed, ids, step):
    return np.array([lib().orc_sfr_uniform(int(seed), int(i), int(step)) for i in ids])


def visual_density(ps, prm, radius=0.0):
    """Tree::calcVisualDensity on a fresh tree; returns (vd, R)."""
    P = _particles(ps)
    vd = np.zeros(ps.n)
    R = C.c_double(0)
    lib().orc_visual_density(C.byref(P), C.byref(prm), float(radius), _p(vd), C.byref(R))
    return vd, R.value


# ---------------------------------------------------------------- reference
def _ref_arrays(ps):
    fs = ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu")
    out = [_p(getattr(ps, f)) for f in fs] + [_p(ps.type, _u8)]
    out += [_p(getattr(ps, f)) for f in ("ax", "ay", "az", "h", "rho", "P", "T", "dUdt")]
    return out


def ref_force_step(ps, theta, e0, massInH, cores=1):
    R = C.c_double(0)
    ref().ref_force_step(ps.n, *_ref_arrays(ps), theta, e0, massInH, cores, C.byref(R))
    return R.value


def ref_tree_export(ps, cores=1, cap=None):
    n = ps.n
    cap = 4 * n + 64 if cap is None else cap
    out = dict(level=np.zeros(cap, np.int32), first=np.zeros(cap, np.int64), count=np.zeros(cap, np.int64),
               mass=np.zeros(cap), com=np.zeros(3 * cap), gas=np.zeros(cap), mvel=np.zeros(3 * cap),
               perm=np.full(n, -1, np.int64))
    R = C.c_double(0)
    fs = [_p(getattr(ps, f)) for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu")]
    k = ref().ref_tree_export(n, *fs, _p(ps.type, _u8), cores, cap, _p(out["level"], _i32),
                              _p(out["first"], _i64), _p(out["count"], _i64), _p(out["mass"]), _p(out["com"]),
                              _p(out["gas"]), _p(out["mvel"]), _p(out["perm"], _i64), C.byref(R))
    if k < 0:   # deep trees: chains of single-child cells
        return ref_tree_export(ps, cores, 4 * cap)
    res = {f: v[:k] for f, v in out.items() if f not in ("com", "mvel", "perm")}
    res["com"] = out["com"][:3 * k].reshape(k, 3)
    res["mvel"] = out["mvel"][:3 * k].reshape(k, 3)
    nin = int(res["count"][0]) if k else 0
    res["perm"] = out["perm"][:nin]
    res["n_in_tree"] = nin
    res["R"] = R.value
    return res


def ref_run(ps, theta, e0, massInH, eta, min_dt, max_dt, end_time, H0=0.0, cores=1, max_steps=1 << 30):
    return ref().ref_run(ps.n, *_ref_arrays(ps), theta, e0, massInH, cores, eta, min_dt, max_dt, end_time, H0,
                         max_steps)


def ref_visual_density(ps, radius=0.0, cores=1):
    vd = np.zeros(ps.n)
    R = C.c_double(0)
    ref().ref_visual_density(ps.n, _p(ps.x), _p(ps.y), _p(ps.z), _p(ps.m), _p(ps.type, _u8), cores,
                             C.c_double(radius), _p(vd), C.byref(R))
    return vd, R.value


def ref_save_data(ps, extras, directory, fmt, time_step, n_output, delta_time, end_time, current_time):
    """DataManager::saveData of the compiled reference; returns the file's bytes."""
    import os
    d = directory.rstrip("/") + "/"
    ids = np.ascontiguousarray(ps.id, dtype=np.uint32)
    gp = np.ascontiguousarray(extras["galaxyPart"], dtype=np.uint8)
    vd = np.ascontiguousarray(extras["visualDensity"], dtype=np.float64)
    sf = np.ascontiguousarray(extras["sfr"], dtype=np.float64)
    fs = [_p(getattr(ps, f)) for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "T", "P")]
    ref().ref_save_data(ps.n, *fs, _p(ps.type, _u8), _p(gp, _u8), ids.ctypes.data_as(C.POINTER(C.c_uint32)),
                        _p(vd), _p(sf), d.encode(), fmt.encode(), time_step, n_output, delta_time, end_time,
                        current_time)
    ending = {"ag": ".ag", "agc": ".agc", "age": ".age", "hdf5": ".hdf5", "gadget": ".gadget"}[fmt]
    with open(os.path.join(d, "%d%s" % (time_step, ending)), "rb") as f:
        return f.read()


def ref_gas_sources(ps, dt, cooling, star_formation, srand_seed=1):
    """Cooling::coolingRoutine then SFR::sfrRoutine of the compiled reference on
    every gas particle (rand() seeded with srand_seed); returns sfr."""
    sfr = np.zeros(ps.n)
    dt = _c(dt)
    ref().ref_gas_sources(ps.n, _p(ps.T), _p(ps.rho), _p(dt), _p(ps.type, _u8), _p(ps.U), _p(ps.dUdt), int(cooling),
                          int(star_formation), int(srand_seed), _p(sfr))
    return sfr
