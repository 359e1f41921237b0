"""ctypes binding of the C ABI in include/ag2.h, with a `Tree` facade that keeps
the reference's method names (simulation/src/Physics/Tree/Tree.h:13-27).

The engine runs only on the HIP library: constructing an Engine without the
built libag2.so or without a GPU raises; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import units as U

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("AG2_LIBRARY", os.path.join(PKG_DIR, "libag2.so"))

AG2_OK, AG2_ERR_ARG, AG2_ERR_HIP, AG2_ERR_CAPACITY, AG2_ERR_STATE, AG2_ERR_NODEV = range(6)
AG2_FP64, AG2_FP32_SPLIT = 0, 1

_d = C.POINTER(C.c_double)
_u8 = C.POINTER(C.c_uint8)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_u64 = C.POINTER(C.c_uint64)


class Params(C.Structure):
    _fields_ = [("theta", C.c_double), ("e0", C.c_double), ("massInH", C.c_double), ("G", C.c_double),
                ("split_threshold", C.c_int64), ("precision", C.c_int32), ("device", C.c_int32),
                ("eta", C.c_double), ("min_dt", C.c_double), ("max_dt", C.c_double), ("H0", C.c_double),
                ("cooling", C.c_int32), ("star_formation", C.c_int32), ("sfr_seed", C.c_uint64),
                ("graph_launch", C.c_int32), ("timing_events", C.c_int32)]


class HostParticles(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(f, _d) for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu")] + \
               [("type", _u8)] + [(f, _d) for f in ("ax", "ay", "az", "h", "rho", "P", "T", "dUdt")]


class Diag(C.Structure):
    _fields_ = [("global_time", C.c_double), ("steps", C.c_int64), ("R", C.c_double), ("n_in_tree", C.c_int64),
                ("n_nodes", C.c_int64), ("max_level", C.c_int64), ("interactions", C.c_int64),
                ("sph_pairs", C.c_int64), ("groups", C.c_int64), ("gap_particles", C.c_int64),
                ("nan_kicks", C.c_int64), ("kinetic", C.c_double), ("potential", C.c_double),
                ("thermal", C.c_double), ("momentum", C.c_double * 3), ("walk_entries", C.c_int64),
                ("walk_tests", C.c_int64), ("walk_mixed", C.c_int64),
                ("interactions_total", C.c_int64), ("sph_overflow", C.c_int64), ("fp64_fallback", C.c_int64),
                ("near_pairs", C.c_int64), ("near_redo_groups", C.c_int64)]


class PoolParams(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_jobs", C.c_int32), ("hbm_budget", C.c_int64)]


class PoolStats(C.Structure):
    _fields_ = [("rounds", C.c_int64), ("job_steps", C.c_int64), ("particle_steps", C.c_int64),
                ("queued", C.c_int64), ("running", C.c_int64), ("done", C.c_int64),
                ("p50_step_ms", C.c_double), ("p99_step_ms", C.c_double), ("max_step_ms", C.c_double),
                ("busy_s", C.c_double), ("hbm_used", C.c_int64), ("hbm_budget", C.c_int64),
                ("mean_resident", C.c_double), ("max_resident", C.c_int64)]


JOB_QUEUED, JOB_RUNNING, JOB_DONE = 0, 1, 2
DIRECT_NOALLOC_MAX = 65536   # AG2_DIRECT_NOALLOC_MAX

EXPORTS = ("ag2_create", "ag2_destroy", "ag2_last_error", "ag2_set_params", "ag2_upload", "ag2_download",
           "ag2_build_tree", "ag2_gas_density", "ag2_forces", "ag2_init_forces", "ag2_step", "ag2_sync",
           "ag2_energy", "ag2_diagnostics", "ag2_phase_times", "ag2_walk_kernel_ms", "ag2_get_order", "ag2_get_tree",
           "ag2_direct_sum", "ag2_direct_sum_noalloc", "ag2_visual_density", "ag2_get_sfr", "ag2_pool_create", "ag2_pool_destroy", "ag2_pool_last_error",
           "ag2_job_submit",           "ag2_pool_step", "ag2_job_status", "ag2_job_result", "ag2_pool_get_stats", "ag2_context_bytes")

_lib = None


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("ag2 error %d: %s" % (code, msg))
        self.code = code


def load_library(path=LIB_PATH):
    """Load libag2.so and declare its signatures; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libag2.so is not built (run __graft_entry__.build())")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.ag2_create.argtypes = [C.POINTER(Params), C.c_int64, C.POINTER(vp)]
    L.ag2_destroy.argtypes = [vp]
    L.ag2_last_error.argtypes = [vp]
    L.ag2_last_error.restype = C.c_char_p
    L.ag2_set_params.argtypes = [vp, C.POINTER(Params)]
    L.ag2_upload.argtypes = [vp, C.POINTER(HostParticles)]
    L.ag2_download.argtypes = [vp, C.POINTER(HostParticles)]
    for f in ("ag2_build_tree", "ag2_gas_density", "ag2_forces", "ag2_init_forces", "ag2_step", "ag2_sync"):
        getattr(L, f).argtypes = [vp]
    L.ag2_energy.argtypes = [vp, C.c_int, C.POINTER(Diag)]
    L.ag2_diagnostics.argtypes = [vp, C.POINTER(Diag)]
    L.ag2_phase_times.argtypes = [vp, _d]
    L.ag2_walk_kernel_ms.argtypes = [vp, _d]
    L.ag2_get_order.argtypes = [vp, _i64, _u64, _u64]
    L.ag2_get_tree.argtypes = [vp, C.c_int64, _i32, _i64, _i64, _d, _d, _d, _d, _i64]
    L.ag2_direct_sum.argtypes = [vp, C.c_int64, _i64, _d, _d, _d, _d]
    L.ag2_direct_sum_noalloc.argtypes = [vp, C.c_int64, _i64, _d, _d, _d, _d]
    L.ag2_visual_density.argtypes = [vp, C.c_double, _d]
    L.ag2_get_sfr.argtypes = [vp, _d]
    L.ag2_pool_create.argtypes = [C.POINTER(PoolParams), C.POINTER(vp)]
    L.ag2_pool_destroy.argtypes = [vp]
    L.ag2_pool_last_error.argtypes = [vp]
    L.ag2_pool_last_error.restype = C.c_char_p
    L.ag2_job_submit.argtypes = [vp, C.POINTER(Params), C.POINTER(HostParticles), C.c_int64, _i64]
    L.ag2_pool_step.argtypes = [vp, _i32]
    L.ag2_job_status.argtypes = [vp, C.c_int64, _i32, _i64]
    L.ag2_job_result.argtypes = [vp, C.c_int64, C.POINTER(HostParticles)]
    L.ag2_pool_get_stats.argtypes = [vp, C.POINTER(PoolStats)]
    for f in EXPORTS:
        if f not in ("ag2_last_error", "ag2_pool_last_error"):
            getattr(L, f).restype = C.c_int
    L.ag2_context_bytes.argtypes = [C.c_int64, C.c_int64]
    L.ag2_context_bytes.restype = C.c_int64
    _lib = L
    return L


def context_bytes(capacity, n_gas=0):
    """Device bytes of a context (or pool job) of this capacity and gas count."""
    return int(load_library().ag2_context_bytes(int(capacity), int(n_gas)))


def _ptr(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


def make_params(theta=0.5, e0=0.05 * U.KPC, massInH=1.0, G=U.G, split_threshold=100, precision=AG2_FP64,
                device=0, eta=2.0, min_dt=1e13, max_dt=1e13, H0=0.0, cooling=0, star_formation=0, sfr_seed=0,
                graph_launch=0, timing_events=0):
    return Params(theta, e0, massInH, G, split_threshold, precision, device, eta, min_dt, max_dt, H0,
                  int(cooling), int(star_formation), int(sfr_seed), int(graph_launch), int(timing_events))


class Engine:
    """One device context (ag2_ctx) holding a particle set resident in HBM."""

    def __init__(self, params, capacity):
        self.lib = load_library()
        self.params = params
        self.ctx = C.c_void_p()
        rc = self.lib.ag2_create(C.byref(params), int(capacity), C.byref(self.ctx))
        if rc != AG2_OK:
            raise EngineError(rc, "ag2_create failed (no GPU or out of memory)")
        self.capacity = int(capacity)
        self.n = 0

    def _check(self, rc):
        if rc != AG2_OK:
            raise EngineError(rc, self.lib.ag2_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.lib.ag2_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        self._check(self.lib.ag2_set_params(self.ctx, C.byref(params)))
        self.params = params

    @staticmethod
    def _host(ps, fields):
        hp = HostParticles()
        hp.n = ps.n
        keep = []
        for f, _ in HostParticles._fields_[1:]:
            a = getattr(ps, f, None) if f in fields else None
            if a is not None:
                want = np.uint8 if f == "type" else np.float64
                if a.dtype != want or not a.flags.c_contiguous:
                    raise ValueError("field %s must be contiguous %s" % (f, want.__name__))
                keep.append(a)
                setattr(hp, f, _ptr(a, _u8 if f == "type" else _d))
        hp._keep = keep
        return hp

    ALL = ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu", "type", "ax", "ay", "az", "h", "rho", "P", "T", "dUdt")

    def upload(self, ps, fields=ALL):
        self._check(self.lib.ag2_upload(self.ctx, C.byref(self._host(ps, fields))))
        self.n = ps.n

    def download(self, ps, fields=ALL):
        self._check(self.lib.ag2_download(self.ctx, C.byref(self._host(ps, fields))))
        return ps

    def build_tree(self):
        self._check(self.lib.ag2_build_tree(self.ctx))

    def gas_density(self):
        self._check(self.lib.ag2_gas_density(self.ctx))

    def forces(self):
        self._check(self.lib.ag2_forces(self.ctx))

    def init_forces(self):
        self._check(self.lib.ag2_init_forces(self.ctx))

    def step(self):
        self._check(self.lib.ag2_step(self.ctx))

    def sync(self):
        self._check(self.lib.ag2_sync(self.ctx))

    def diagnostics(self):
        d = Diag()
        self._check(self.lib.ag2_diagnostics(self.ctx, C.byref(d)))
        return {f: (list(getattr(d, f)) if f == "momentum" else getattr(d, f)) for f, _ in Diag._fields_}

    def energy(self, with_potential=True):
        d = Diag()
        self._check(self.lib.ag2_energy(self.ctx, 1 if with_potential else 0, C.byref(d)))
        return dict(kinetic=d.kinetic, potential=d.potential, thermal=d.thermal, momentum=list(d.momentum))

    def phase_times(self):
        ms = np.zeros(6)
        self._check(self.lib.ag2_phase_times(self.ctx, _ptr(ms)))
        return ms

    def walk_kernel_ms(self):
        ms = np.zeros(1)
        self._check(self.lib.ag2_walk_kernel_ms(self.ctx, _ptr(ms)))
        return float(ms[0])

    def order(self):
        nin = self.diagnostics()["n_in_tree"]
        perm = np.zeros(max(nin, 1), np.int64)
        k1 = np.zeros(max(nin, 1), np.uint64)
        k2 = np.zeros(max(nin, 1), np.uint64)
        self._check(self.lib.ag2_get_order(self.ctx, _ptr(perm, _i64), _ptr(k1, _u64), _ptr(k2, _u64)))
        return perm[:nin], k1[:nin], k2[:nin]

    def direct_sum(self, targets, with_potential=False):
        """FP64 direct-sum accelerations (and potential) of the listed targets."""
        t = np.ascontiguousarray(targets, dtype=np.int64)
        nt = len(t)
        ax, ay, az = np.zeros(nt), np.zeros(nt), np.zeros(nt)
        pot = np.zeros(nt) if with_potential else None
        self._check(self.lib.ag2_direct_sum(self.ctx, nt, _ptr(t, _i64), _ptr(ax), _ptr(ay), _ptr(az), _ptr(pot)))
        return ax, ay, az, pot

    def direct_sum_noalloc(self, targets, with_potential=False):
        """direct_sum through ag2_direct_sum_noalloc (no device allocation; at
        most DIRECT_NOALLOC_MAX targets)."""
        t = np.ascontiguousarray(targets, dtype=np.int64)
        nt = len(t)
        ax, ay, az = np.zeros(nt), np.zeros(nt), np.zeros(nt)
        pot = np.zeros(nt) if with_potential else None
        self._check(self.lib.ag2_direct_sum_noalloc(self.ctx, nt, _ptr(t, _i64), _ptr(ax), _ptr(ay), _ptr(az),
                                                    _ptr(pot)))
        return ax, ay, az, pot

    def visual_density(self, radius=0.0):
        """Visual density of the last tree, caller order (Tree::calcVisualDensity).
        radius 0 takes the root half-width of the first tree after upload / 1e5."""
        out = np.zeros(self.n)
        self._check(self.lib.ag2_visual_density(self.ctx, float(radius), _ptr(out)))
        return out

    def sfr(self):
        """Particle::sfr of every particle, caller order."""
        out = np.zeros(self.n)
        self._check(self.lib.ag2_get_sfr(self.ctx, _ptr(out)))
        return out

    def save_snapshot(self, out_dir, ps, fmt, time_step, n_output=None, delta_time=0.0, end_time=0.0,
                      galaxy_part=None, sfr=None, visual_radius=0.0):
        """The run loop's snapshot (Simulation.cpp:348-353): download the state
        into ps (caller order), take the last tree's visual density and write
        DataManager::saveData's file for fmt at the current global time."""
        from . import ageio
        self.sync()
        self.download(ps)
        vd = self.visual_density(visual_radius)
        if sfr is None:
            sfr = self.sfr()
        return ageio.save_data(out_dir, ps, fmt, time_step, n_output, delta_time, end_time,
                               self.diagnostics()["global_time"], visual_density=vd, sfr=sfr,
                               galaxy_part=galaxy_part)

    def tree(self):
        nn = C.c_int64(0)
        self.lib.ag2_get_tree(self.ctx, 0, None, None, None, None, None, None, None, C.byref(nn))
        k = nn.value
        out = dict(level=np.zeros(k, np.int32), first=np.zeros(k, np.int64), count=np.zeros(k, np.int64),
                   mass=np.zeros(k), com=np.zeros(3 * k), gas=np.zeros(k), mvel=np.zeros(3 * k))
        self._check(self.lib.ag2_get_tree(self.ctx, k, _ptr(out["level"], _i32), _ptr(out["first"], _i64),
                                          _ptr(out["count"], _i64), _ptr(out["mass"]), _ptr(out["com"]),
                                          _ptr(out["gas"]), _ptr(out["mvel"]), C.byref(nn)))
        out["com"] = out["com"].reshape(k, 3)
        out["mvel"] = out["mvel"].reshape(k, 3)
        return out


class Pool:
    """A serving pool (ag2_pool_*): independent jobs resident on one device,
    advanced one step per round with continuous batching."""

    def __init__(self, max_jobs=16, hbm_budget=0, device=0):
        self.lib = load_library()
        self.pool = C.c_void_p()
        rc = self.lib.ag2_pool_create(C.byref(PoolParams(device, max_jobs, int(hbm_budget))), C.byref(self.pool))
        if rc != AG2_OK:
            raise EngineError(rc, "ag2_pool_create failed (no GPU or out of memory)")
        self.device = device

    def _check(self, rc):
        if rc != AG2_OK:
            raise EngineError(rc, self.lib.ag2_pool_last_error(self.pool).decode())

    def close(self):
        if self.pool:
            self.lib.ag2_pool_destroy(self.pool)
            self.pool = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, params, ps, steps):
        hp = Engine._host(ps, ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu", "type"))
        jid = C.c_int64(0)
        self._check(self.lib.ag2_job_submit(self.pool, C.byref(params), C.byref(hp), int(steps), C.byref(jid)))
        return jid.value

    def step(self):
        n = C.c_int32(0)
        self._check(self.lib.ag2_pool_step(self.pool, C.byref(n)))
        return n.value

    def status(self, jid):
        st, k = C.c_int32(0), C.c_int64(0)
        self._check(self.lib.ag2_job_status(self.pool, int(jid), C.byref(st), C.byref(k)))
        return st.value, k.value

    def result(self, jid, ps):
        hp = Engine._host(ps, Engine.ALL)
        self._check(self.lib.ag2_job_result(self.pool, int(jid), C.byref(hp)))
        return ps

    def stats(self):
        s = PoolStats()
        self._check(self.lib.ag2_pool_get_stats(self.pool, C.byref(s)))
        return {f: getattr(s, f) for f, _ in PoolStats._fields_}


class Tree:
    """The reference's per-step tree phases (Tree.h:13-27) on a resident
    Engine: buildTree, calcVisualDensity, calcGasDensity, calculateForces."""

    def __init__(self, engine):
        self.engine = engine

    def buildTree(self):
        self.engine.build_tree()

    def calcGasDensity(self):
        self.engine.gas_density()

    def calculateForces(self):
        self.engine.forces()

    def calcTreeWidth(self):
        return self.engine.diagnostics()["R"]

    def calcVisualDensity(self, radius=0.0):
        """Particle::visualDensity of every particle, caller order."""
        return self.engine.visual_density(radius)
