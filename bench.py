"""Benchmark of the MI355X force engine.

A step is one iteration of the reference's Simulation::run on the device:
block time steps, kick + drift, tree build, SPH group density, gravity walk
with embedded SPH force, U update and second kick.  Inputs are resident in
HBM before timing starts.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision both|fp64|fp32]

The headline is the 1M-particle isolated disk galaxy: `value` is the
split-precision FP32 form's throughput and the plain FP64 form is timed beside
it.  With it: accuracy against an FP64 direct sum on a 65,536-particle sample
(and whether that meets the 1e-3 gate, which the reference's opening rule at
theta 0.5 cannot), the FP32 form at the largest theta that meets the gate,
the total energy after every step with runaway indicators, the other
configurations (10k Plummer, 4M disk, 16M merger), the serving stream, and the
reference's CPU path on the host cores.
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import time

import numpy as np


def _vec_err(ax, ay, az, bx, by, bz):
    num = np.sqrt((ax - bx) ** 2 + (ay - by) ** 2 + (az - bz) ** 2)
    den = np.sqrt(bx ** 2 + by ** 2 + bz ** 2)
    return num / np.where(den > 0, den, 1.0)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "astrogenesis2.0_amd")
METRIC = "particle-steps/s FP32 vs FP64 at ≤1e-3 p99 accel err; p50/p99 step latency"
FLOPS_PER_INTERACTION = 19          # 3 sub, 5 r^2, 1 add, rsqrt, rcp, 2 mul, 3 fma
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
GATE = 1e-3                         # p99 acceleration error bound of the metric
ENERGY_FLAG = 0.10                  # a step whose total energy moved by more than this fraction is flagged


def host_cpu():
    """Host cores for the CPU baseline: the box's CPU share (16 per GPU), at
    most what the OS reports; and the CPU model."""
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": n, "threads": min(16, n), "model": model}


def load_pkg():
    spec = importlib.util.spec_from_file_location("ag2_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ag2_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def workload(ag2, n_total, kind="disk"):
    """Disk galaxy: 50 % NFW halo, 30 % stellar disk, 20 % gas disk (or a
    Plummer sphere of the CPU-runnable configuration's recipe, or two disk
    galaxies of n_total / 2 each on a parabolic orbit)."""
    U = ag2.units
    if kind == "merger":
        ps = ag2.ics.merger(n_total // 2)
        n_each = n_total // 2
        nh, nd = n_each // 2, (3 * n_each) // 10
        prm = dict(theta=0.5, e0=0.1 * U.KPC, massInH=ag2.ics.mass_in_h(ps), split_threshold=800,
                   dt=0.5 * U.MYR)
        return ps, prm, (2 * nh, 2 * nd, ps.n - 2 * nh - 2 * nd)
    if kind == "plummer":
        ps = ag2.ics.config1_plummer(n_total, seed=1)
        return ps, dict(theta=0.5, e0=0.05 * U.KPC, massInH=1.0, split_threshold=800, dt=0.1 * U.MYR), (n_total, 0, 0)
    nh, nd = n_total // 2, (3 * n_total) // 10
    ng = n_total - nh - nd
    ps = ag2.ics.disk_galaxy(nh, nd, ng, seed=2, config=2 if n_total <= 1_000_000 else 3)
    prm = dict(theta=0.5, e0=0.1 * U.KPC, massInH=ag2.ics.mass_in_h(ps), split_threshold=800,
               dt=0.5 * U.MYR)
    return ps, prm, (nh, nd, ng)


def describe(kind, n, split):
    if kind == "disk":
        return ("isolated disk galaxy N=%d (%d NFW halo, %d stellar disk, %d SPH gas), theta=0.5, e0=0.1 kpc, "
                "dt=0.5 Myr KDK, massInH=64 m_gas, split_threshold=800" % ((n,) + split))
    if kind == "merger":
        return ("merger of two disk galaxies, N=%d (%d NFW halo, %d stellar disk, %d SPH gas), parabolic orbit from "
                "50 kpc, pericentre 8 kpc, second disk inclined 60 deg, theta=0.5, e0=0.1 kpc, dt=0.5 Myr KDK, "
                "massInH=64 m_gas, split_threshold=800" % ((n,) + split))
    return "Plummer sphere N=%d, theta=0.5, e0=0.05 kpc, dt=0.1 Myr KDK" % n


# ------------------------------------------------------------------ CPU side
def cpu_child(args):
    """In a subprocess with the OpenMP thread count set: time the compiled
    reference (or, where it is absent, the oracle) on a workload.  mode
    'force': one force evaluation (the reference's per-step tree phases);
    'run': the reference's step loop (init force phase + steps)."""
    ag2 = load_pkg()
    sys.path.insert(0, ROOT)
    from oracle import pyoracle as O
    ps, prm, _ = workload(ag2, args.n, args.workload)
    if args.subsample and args.subsample < ps.n:
        g = ag2.ics.rng(11, 0, 0)
        ps = ps.take(np.sort(g.choice(ps.n, args.subsample, replace=False)))
    threads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    if args.cpu_mode == "run":
        steps = args.cpu_steps
        if O.ref_available():
            t0 = time.perf_counter()
            n = O.ref_run(ps, prm["theta"], prm["e0"], prm["massInH"], 2.0, prm["dt"], prm["dt"],
                          (steps - 0.5) * prm["dt"], cores=threads)
            kind = "reference"
        else:
            t0 = time.perf_counter()
            n = O.run(ps, O.params(prm["theta"], prm["e0"], prm["massInH"], prm["split_threshold"], nthreads=threads),
                      2.0, prm["dt"], prm["dt"], (steps - 0.5) * prm["dt"])
            kind = "port"
        total = time.perf_counter() - t0
        print(json.dumps({"kind": kind, "seconds": total, "n": ps.n, "steps": int(n), "phases": {}}))
        return
    if O.ref_available():
        import ctypes as C
        L = O.ref()
        d = C.POINTER(C.c_double)
        L.ref_time_step.argtypes = [C.c_int64] + [d] * 9 + [C.POINTER(C.c_uint8), C.c_double, C.c_double,
                                                             C.c_double, d]
        sec = np.zeros(4)
        arrs = [getattr(ps, f).ctypes.data_as(d) for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu")]
        L.ref_time_step(ps.n, *arrs, ps.type.ctypes.data_as(C.POINTER(C.c_uint8)), prm["theta"], prm["e0"],
                        prm["massInH"], sec.ctypes.data_as(d))
        kind, total = "reference", float(sec.sum())
        phases = dict(build_tree=sec[0], sph_density=sec[1], forces=sec[2], delete_tree=sec[3])
    else:
        a = ps.copy()
        t0 = time.perf_counter()
        O.force_step(a, O.params(prm["theta"], prm["e0"], prm["massInH"], prm["split_threshold"], nthreads=threads))
        kind, total = "port", time.perf_counter() - t0
        phases = {}
    print(json.dumps({"kind": kind, "seconds": total, "n": ps.n, "steps": 1,
                      "phases": {k: float(v) for k, v in phases.items()}}))


def cpu_baseline(kind, n, mode="force", steps=1, subsample=0):
    hc = host_cpu()
    th = hc["threads"]
    env = dict(os.environ, OMP_NUM_THREADS=str(th), OMP_THREAD_LIMIT=str(th), OMP_PROC_BIND="close")
    cmd = [sys.executable, os.path.abspath(__file__), "--_cpu_child", "--n", str(n), "--workload", kind,
           "--cpu-mode", mode, "--cpu-steps", str(steps), "--subsample", str(subsample)]
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=900)
    if r.returncode != 0:
        return {"value": None, "error": r.stderr[-400:]}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    if mode == "run":
        what = ("the reference's step loop (Simulation::init force phase and %d KDK steps, Simulation.cpp:121-356) "
                "on the full %d-particle workload" % (res["steps"], res["n"]))
        value = res["n"] * res["steps"] / res["seconds"]
    else:
        what = ("one force evaluation (build tree + SPH density + forces + delete tree, the reference's per-step tree "
                "phases) of %s" % ("the full %d-particle workload" % res["n"] if not subsample else
                                   "a seeded random %d of the workload's %d particles" % (res["n"], n)))
        value = res["n"] / res["seconds"]
    return {"value": value, "unit": "particle-steps/s", "cores": th, "kind": res["kind"],
            "sample": what + ", OpenMP on %d host threads (%s; nproc %d)" % (th, hc["model"], hc["nproc"]),
            "seconds": res["seconds"], "phases_s": res["phases"]}


def copy_probe():
    """A plain device copy of 1 GiB, run three times: under a counter profiler
    its FETCH_SIZE / WRITE_SIZE per dispatch calibrate the HBM-traffic
    counters against a known byte count."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4}))


# ------------------------------------------------------------------ GPU side
def engine_params(ag2, prm, precision, timing_events=0, theta=None):
    return ag2.engine.make_params(theta=prm["theta"] if theta is None else theta, e0=prm["e0"],
                                  massInH=prm["massInH"], split_threshold=prm["split_threshold"], precision=precision,
                                  eta=2.0, min_dt=prm["dt"], max_dt=prm["dt"], H0=0.0, cooling=prm.get("cooling", 0),
                                  star_formation=prm.get("star_formation", 0), timing_events=timing_events)


def make_engine(ag2, ps, prm, precision, timing_events=0, theta=None):
    eng = ag2.engine.Engine(engine_params(ag2, prm, precision, timing_events, theta), ps.n)
    eng.upload(ps)
    return eng


PHASE_STEPS = 3


def run_gpu(steps, warmup, ag2, ps, prm, precision, theta=None, phases=True):
    """Time `steps` steps of one precision form after `warmup`.  The timed
    steps record HIP events around the walk kernel only (its duration is the
    roofline's); the per-phase breakdown comes from PHASE_STEPS more steps
    afterwards with every phase boundary recorded."""
    eng = make_engine(ag2, ps, prm, precision, timing_events=1, theta=theta)
    eng.init_forces()
    for _ in range(warmup):
        eng.step()
    eng.sync()
    lat, kern_ms = [], []
    # interactions come from the engine's running total, read once on each
    # side of the timed loop; inside it only host-side event queries remain
    inter0 = eng.diagnostics()["interactions_total"]
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        eng.step()
        eng.sync()
        lat.append(time.perf_counter() - ts)
        kern_ms.append(eng.walk_kernel_ms())
    elapsed = time.perf_counter() - t0
    d = eng.diagnostics()
    inter = [(d["interactions_total"] - inter0) / steps] * steps
    out = dict(elapsed=elapsed, lat=np.array(lat), kern_ms=np.array(kern_ms), inter=np.array(inter, float), diag=d,
               mask_density=d["interactions"] / max(1, 64 * d["walk_entries"]))
    if phases:
        # per-phase device times, outside the timed steps
        eng.set_params(engine_params(ag2, prm, precision, timing_events=0, theta=theta))
        ph = np.zeros(6)
        for _ in range(PHASE_STEPS):
            eng.step()
            eng.sync()
            ph += eng.phase_times()
        out["phases"] = ph / PHASE_STEPS
        # visual density of the last tree, as a snapshot writer asks for it (host clock, including the
        # caller-order download of n doubles)
        tv = time.perf_counter()
        vd = eng.visual_density()
        out["visual_density_ms"] = 1e3 * (time.perf_counter() - tv)
        out["visual_density_nonzero"] = int((vd > 0).sum())
    eng.close()
    return out


def form_summary(n, rr, name):
    fl = rr["inter"] * FLOPS_PER_INTERACTION
    ach = float(fl.sum() / (rr["kern_ms"].sum() * 1e-3) / 1e12)
    s = {"particle_steps_per_s": n * len(rr["lat"]) / rr["elapsed"],
         "ms_per_step": 1e3 * rr["elapsed"] / len(rr["lat"]),
         "p50_step_ms": pct(rr["lat"], 50) * 1e3, "p99_step_ms": pct(rr["lat"], 99) * 1e3,
         "interactions_per_step": float(rr["inter"].mean()),
         "interactions_per_s": float(rr["inter"].sum() / rr["elapsed"]),
         "walk_kernel_ms": float(rr["kern_ms"].mean()),
         "walk_tflops": ach, "walk_frac_of_vector_peak": ach / PEAK_TFLOPS[name],
         "walk_mask_density": rr["mask_density"],
         "walk_node_tests_per_step": rr["diag"]["walk_tests"],
         "walk_batch_entries_per_step": rr["diag"]["walk_entries"],
         "walk_mixed_tests_per_step": rr["diag"]["walk_mixed"],
         "sph_pairs_per_step": rr["diag"]["sph_pairs"],
         "sph_overflow_targets": rr["diag"]["sph_overflow"],
         "fp64_fallback": rr["diag"]["fp64_fallback"],
         "n_in_tree_last_step": rr["diag"]["n_in_tree"],
         "max_level": rr["diag"]["max_level"]}
    if "phases" in rr:
        s["phases_ms"] = dict(zip(["keys", "sort", "topology_moments", "sph_density", "walk", "kdk"],
                                  [float(v) for v in rr["phases"]]))
        s["visual_density_call_ms"] = rr["visual_density_ms"]
    return s


def total_energy(eng):
    e = eng.energy(with_potential=True)
    return e["kinetic"] + e["potential"] + e["thermal"], e


def energy_series(ag2, ps, prm, precision, steps):
    """Total energy (kinetic + thermal + tree potential) after the initial
    force phase and after every step, in a run of its own (untimed), with the
    steps whose energy moved by more than ENERGY_FLAG of the initial magnitude
    flagged, and runaway indicators at the end: the largest speed, gas with
    U < 0 (the reference's SPH lets U go negative, TimeIntegration.cpp:28-41,
    and negative pressure then accelerates gas without bound) and kicks
    skipped for a NaN acceleration."""
    eng = make_engine(ag2, ps, prm, precision, timing_events=2)
    eng.init_forces()
    e0, parts0 = total_energy(eng)
    series = [e0]
    for _ in range(steps):
        eng.step()
        series.append(total_energy(eng)[0])
    b = eng.download(ps.copy(), fields=("vx", "vy", "vz", "U", "type"))
    d = eng.diagnostics()
    eng.close()
    rel = [(e - e0) / abs(e0) for e in series]
    # a runaway can overflow speeds and energies: such steps are flagged and
    # their values reported as null (JSON has no infinity)
    bad = lambda x: not np.isfinite(x)  # noqa: E731
    flagged = [k for k in range(1, len(rel))
               if bad(rel[k]) or bad(rel[k - 1]) or abs(rel[k] - rel[k - 1]) > ENERGY_FLAG or abs(rel[k]) > ENERGY_FLAG]
    fin = lambda x: float(x) if np.isfinite(x) else None  # noqa: E731
    v = np.sqrt(b.vx ** 2 + b.vy ** 2 + b.vz ** 2)
    gas = b.type == ag2.units.TYPE_GAS
    return {"steps": steps, "initial_J": {k: parts0[k] for k in ("kinetic", "potential", "thermal")},
            "relative_change_per_step": [fin(x) for x in rel],
            "drift_over_run": fin(rel[-1]), "flag_bound": ENERGY_FLAG,
            "flagged_steps": flagged, "flagged": bool(flagged),
            "max_speed_km_s": fin(np.nanmax(v) / ag2.units.KMS) if np.isfinite(v).any() else None,
            "non_finite_speeds": int((~np.isfinite(v)).sum()),
            "gas_with_negative_U": int((gas & (b.U < 0)).sum()), "gas": int(gas.sum()),
            "nan_kicks": int(d["nan_kicks"])}


def pct(a, q):
    return float(np.percentile(a, q))


def finite_json(o):
    """The bench line as strict JSON: a non-finite number (a runaway's
    overflowed speed or energy) becomes null."""
    if isinstance(o, dict):
        return {k: finite_json(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [finite_json(v) for v in o]
    if isinstance(o, float) and not np.isfinite(o):
        return None
    return o


def accuracy(args, ag2, ps, prm, sample, cpu_check):
    """Acceleration errors on the initial state: FP32 split vs FP64 tree (all
    particles), and both tree forms vs an FP64 direct sum on a seeded sample
    of non-gas particles (SPH forces are not part of the direct sum); the GPU
    direct sum is checked against the CPU oracle on a sub-sample.  Returns
    the summary and the direct-sum sample (targets, accelerations) for reuse."""
    vec_err = _vec_err
    f64 = make_engine(ag2, ps, prm, 0)
    f64.init_forces()
    a64 = f64.download(ps.copy(), fields=("ax", "ay", "az"))
    f32 = make_engine(ag2, ps, prm, 1)
    f32.init_forces()
    a32 = f32.download(ps.copy(), fields=("ax", "ay", "az"))
    f32.close()
    prec = vec_err(a32.ax, a32.ay, a32.az, a64.ax, a64.ay, a64.az)
    g = ag2.ics.rng(7, 2, 9)
    pool = np.flatnonzero(ps.type != ag2.units.TYPE_GAS)
    tg = np.sort(g.choice(pool, min(sample, len(pool)), replace=False)).astype(np.int64)
    dx, dy, dz, _ = f64.direct_sum(tg)
    f64.close()
    e64 = vec_err(a64.ax[tg], a64.ay[tg], a64.az[tg], dx, dy, dz)
    e32 = vec_err(a32.ax[tg], a32.ay[tg], a32.az[tg], dx, dy, dz)
    out = {"theta": prm["theta"],
           "precision_fp32_vs_fp64_tree": {"p50": pct(prec, 50), "p99": pct(prec, 99), "max": float(prec.max()),
                                           "n": int(ps.n)},
           "tree_fp64_vs_direct": {"p50": pct(e64, 50), "p99": pct(e64, 99), "n": int(len(tg))},
           "tree_fp32_vs_direct": {"p50": pct(e32, 50), "p99": pct(e32, 99), "n": int(len(tg))}}
    try:
        sys.path.insert(0, ROOT)
        from oracle import pyoracle as O
        sub = tg[:: max(1, len(tg) // cpu_check)]
        ox, oy, oz, _ = O.direct_sum(ps.x, ps.y, ps.z, ps.m, prm["e0"], targets=sub, nthreads=host_cpu()["threads"])
        k = np.searchsorted(tg, sub)
        chk = vec_err(dx[k], dy[k], dz[k], ox, oy, oz)
        out["direct_sum_gpu_vs_cpu_oracle_max"] = float(chk.max())
        out["direct_sum_cpu_check_n"] = int(len(sub))
    except Exception as e:  # the oracle is absent only if it was never built
        out["direct_sum_cpu_check"] = "skipped: %s" % e
    return out, (tg, dx, dy, dz)


def accurate_fp32(args, ag2, ps, prm, direct, fp64_ref):
    """The split-precision form at the largest theta of a short list that
    meets the gate (p99 acceleration error <= 1e-3 against the FP64 direct
    sum on the accuracy sample), under the reference's own per-particle
    opening rule (Node.cpp:331-334): its error, throughput, and its ratio to
    the plain FP64 form at theta 0.5."""
    tg, dx, dy, dz = direct
    tried = []
    chosen = None
    for theta in (0.12, 0.1, 0.09, 0.08, 0.07, 0.06):
        eng = make_engine(ag2, ps, prm, 1, theta=theta)
        eng.init_forces()
        a = eng.download(ps.copy(), fields=("ax", "ay", "az"))
        eng.close()
        e = _vec_err(a.ax[tg], a.ay[tg], a.az[tg], dx, dy, dz)
        tried.append({"theta": theta, "p50": pct(e, 50), "p99": pct(e, 99)})
        if pct(e, 99) <= GATE:
            chosen = theta
            break
    out = {"rule": "the reference's per-particle opening test radius / r < theta (Node.cpp:331-334)",
           "tried": tried, "theta": chosen, "gate_met": chosen is not None}
    if chosen is None:
        return out
    rr = run_gpu(args.accurate_steps, 1, ag2, ps, prm, 1, theta=chosen, phases=False)
    s = form_summary(ps.n, rr, "fp32")
    out.update(particle_steps_per_s=s["particle_steps_per_s"], ms_per_step=s["ms_per_step"],
               interactions_per_step=s["interactions_per_step"], walk_tflops=s["walk_tflops"],
               walk_frac_of_vector_peak=s["walk_frac_of_vector_peak"], steps=args.accurate_steps,
               ratio_to_fp64_theta05=s["particle_steps_per_s"] / fp64_ref if fp64_ref else None,
               p99_vs_direct=tried[-1]["p99"])
    return out


# ------------------------------------------------------------------ serving
def serving(args, ag2):
    """The serving stream: independent jobs (Plummer / Hernquist spheres, N
    log-uniform in [1e4, 2e5], 30 % with gas) arriving at a Poisson rate
    above the pool's service rate, so every slot stays busy; 50
    split-precision FP32 steps each; a job arriving while others run is
    admitted at the next round with a free slot.  Latency is per job-step:
    from the end of the job's previous step (or the start of the round that
    admitted it) to the end of the round completing this step, admissions and
    retirements included."""
    E = ag2.engine
    jobs = []
    for j in range(args.serving_jobs):
        ps, prm = ag2.ics.serving_job(j)
        p = E.make_params(theta=prm["theta"], e0=prm["e0"], massInH=prm["massInH"],
                          split_threshold=prm["split_threshold"], precision=E.AG2_FP32_SPLIT, eta=2.0,
                          min_dt=prm["dt"], max_dt=prm["dt"], H0=0.0, cooling=prm.get("cooling", 0),
                          star_formation=prm.get("star_formation", 0))
        jobs.append((ps, p, prm["steps"]))
    g = ag2.ics.rng(4, 4, 99)
    arrive = np.cumsum(g.exponential(1.0 / args.serving_rate, len(jobs)))
    log("serving: %d jobs, %d particles, rate %.0f/s" % (len(jobs), sum(j[0].n for j in jobs), args.serving_rate))
    pool = E.Pool(max_jobs=args.serving_slots)
    ids, nxt, finished = {}, 0, 0
    t0 = time.perf_counter()
    while finished < len(jobs):
        now = time.perf_counter() - t0
        while nxt < len(jobs) and arrive[nxt] <= now:
            ps, p, steps = jobs[nxt]
            ids[pool.submit(p, ps, steps)] = nxt
            nxt += 1
        n = pool.step()
        if n == 0 and nxt < len(jobs):
            time.sleep(max(0.0, arrive[nxt] - (time.perf_counter() - t0)))
            continue
        for jid in list(ids):
            if pool.status(jid)[0] == E.JOB_DONE:
                i = ids.pop(jid)
                pool.result(jid, jobs[i][0].copy())
                finished += 1
    wall = time.perf_counter() - t0
    st = pool.stats()
    pool.close()
    log("serving done: %.1f s" % wall)
    return {"jobs": len(jobs), "slots": args.serving_slots, "arrival_rate_per_s": args.serving_rate,
            "steps_per_job": 50, "form": "split-precision FP32",
            "p50_step_ms": st["p50_step_ms"], "p99_step_ms": st["p99_step_ms"], "max_step_ms": st["max_step_ms"],
            "latency": "per job-step interval: end of the job's previous step (or start of its admitting round) to "
                       "the end of this step's round, admissions and retirements included",
            "mean_resident_jobs": st["mean_resident"], "max_resident_jobs": st["max_resident"],
            "particle_steps_per_s_busy": st["particle_steps"] / st["busy_s"] if st["busy_s"] > 0 else None,
            "particle_steps_per_s_wall": st["particle_steps"] / wall, "wall_s": wall, "rounds": st["rounds"],
            "job_steps": st["job_steps"], "particle_steps": st["particle_steps"],
            "data": "synthetic seeded jobs (Plummer / Hernquist, N log-uniform 1e4..2e5, 30 % with 20 % gas)"}


# ------------------------------------------------------------------ configs
def other_config(args, ag2, kind, n, steps, warmup, energy_steps, sample, cpu, gas_sources=False):
    """One more configuration: both forms timed, accuracy (sample > 0),
    energy series and the CPU baseline (cpu = (mode, steps, subsample), or
    None); gas_sources: the reference's cooling and star formation on."""
    ps, prm, split = workload(ag2, n, kind)
    if gas_sources:
        prm.update(cooling=1, star_formation=1)
    log("%s n=%d ready" % (kind, ps.n))
    forms = {}
    for name, prec in (("fp64", 0), ("fp32", 1)):
        rr = run_gpu(steps, warmup, ag2, ps, prm, prec, phases=False)
        forms[name] = form_summary(ps.n, rr, name)
    out = {"workload": describe(kind, ps.n, split), "steps": steps, "warmup": warmup, "forms": forms,
           "gas_sources": "cooling and star formation on" if gas_sources else "off",
           "fp32_over_fp64": forms["fp32"]["particle_steps_per_s"] / forms["fp64"]["particle_steps_per_s"]}
    if sample:
        acc, _ = accuracy(args, ag2, ps, prm, sample, 256)
        out["accuracy"] = acc
    if energy_steps:
        out["energy"] = energy_series(ag2, ps, prm, 1, energy_steps)
    if cpu is not None and not args.no_cpu_baseline:
        mode, csteps, sub = cpu
        out["cpu_baseline"] = cpu_baseline(kind, n, mode, csteps, sub)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--precision", choices=["fp64", "fp32", "both"], default="both")
    ap.add_argument("--workload", choices=["disk", "plummer", "merger"], default="disk")
    ap.add_argument("--gas-sources", action="store_true",
                    help="cooling and star formation on (the reference's per-particle rules)")
    ap.add_argument("--sample", type=int, default=65536)
    ap.add_argument("--cpu-check", type=int, default=2048)
    ap.add_argument("--accurate-steps", type=int, default=3)
    ap.add_argument("--energy-steps", type=int, default=25)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="only the headline workload")
    ap.add_argument("--serving-jobs", type=int, default=256)
    ap.add_argument("--serving-slots", type=int, default=16)
    ap.add_argument("--serving-rate", type=float, default=400.0,
                    help="job arrivals per second (Poisson); above the pool's service rate")
    ap.add_argument("--_cpu_child", action="store_true")
    ap.add_argument("--cpu-mode", default="force")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--subsample", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args._cpu_child:
        return cpu_child(args)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (a single simulation is not sharded "
              "and serving jobs are placed per GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    ag2 = load_pkg()
    ps, prm, split = workload(ag2, args.n, args.workload)
    if args.gas_sources:
        prm.update(cooling=1, star_formation=1)
    n = ps.n
    forms = {"fp64": 0, "fp32": 1} if args.precision == "both" else {args.precision: 0 if args.precision == "fp64" else 1}
    log("workload ready: n=%d" % n)
    runs = {}
    for name, prec in forms.items():
        runs[name] = run_gpu(args.steps, args.warmup, ag2, ps, prm, prec)
        log("%s: %.3f ms/step" % (name, 1e3 * runs[name]["elapsed"] / args.steps))
    head = "fp32" if "fp32" in runs else "fp64"
    r = runs[head]
    value = n * args.steps / r["elapsed"]
    flops = r["inter"] * FLOPS_PER_INTERACTION
    achieved = float(flops.sum() / (r["kern_ms"].sum() * 1e-3) / 1e12)
    peak = PEAK_TFLOPS[head]
    per_form = {name: form_summary(n, rr, name) for name, rr in runs.items()}
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * r["elapsed"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if head == "fp32" else "f64",
        "data": "synthetic (seeded initial conditions)",
        "config": {"workload": describe(args.workload, n, split),
                   "gas_sources": "cooling and star formation on" if args.gas_sources else "off",
                   "form": "split-precision FP32 forces (FP64 positions, tree, integrator)" if head == "fp32"
                   else "plain FP64"},
        "p50_step_ms": pct(r["lat"], 50) * 1e3,
        "p99_step_ms": pct(r["lat"], 99) * 1e3,
        "interactions_per_s": float(r["inter"].sum() / r["elapsed"]),
        "interactions_per_step": float(r["inter"].mean()),
        "forms": per_form,
        "roofline": {"bound": "valu", "kernel": "k_walk_block32" if head == "fp32" else "k_walk_block64",
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": profiled_traffic("k_walk_block32" if head == "fp32" else "k_walk_block64", head),
                     "traffic_source": "profiles/hbm_traffic_disk1m_%s.md (FETCH_SIZE+WRITE_SIZE, copy-scaled)" % head,
                     "kernel_ms": float(r["kern_ms"].mean()),
                     "work": "%d flops per gravity interaction x interactions of the launch (device counter), "
                             "the launch's duration from HIP events around it on the engine's stream"
                             % FLOPS_PER_INTERACTION},
    }
    if "fp32" in runs and "fp64" in runs:
        out["fp32_over_fp64"] = runs["fp64"]["elapsed"] / runs["fp32"]["elapsed"]
    if args.workload == "disk" and args.n == 1_000_000 and head == "fp32":
        out["hbm_kernels"] = {"peak_GB_s": HBM_PEAK_GBS, "kernels": profiled_hbm("fp32"),
                              "source": "profiles/hbm_traffic_disk1m_fp32.md (bytes) and "
                                        "profiles/kernel_trace_disk1m_fp32.md (average duration)"}
    if not args.no_accuracy:
        acc, direct = accuracy(args, ag2, ps, prm, args.sample, args.cpu_check)
        out["accuracy"] = acc
        p_dir = acc["tree_%s_vs_direct" % head]["p99"]
        p_prec = acc["precision_fp32_vs_fp64_tree"]["p99"]
        # The metric's gate: p99 acceleration error <= 1e-3.  Against the FP64
        # direct sum the reference's own opening rule at theta 0.5 is far
        # above it (both forms); the split form's precision error against the
        # FP64 tree is far below it.  accurate_fp32 is the form that meets it.
        out["gate_met"] = bool(p_dir <= GATE)
        out["gate"] = {"bound": GATE, "value_form_p99_vs_direct": p_dir, "met_vs_direct": bool(p_dir <= GATE),
                       "precision_p99_fp32_vs_fp64_tree": p_prec, "met_precision": bool(p_prec <= GATE)}
        log("accuracy done")
        if head == "fp32" and "fp64" in per_form:
            out["accurate_fp32"] = accurate_fp32(args, ag2, ps, prm, direct, per_form["fp64"]["particle_steps_per_s"])
            log("accurate mode done")
    if args.energy_steps:
        out["energy"] = {name: energy_series(ag2, ps, prm, prec, args.energy_steps) for name, prec in forms.items()}
        out["energy_flagged"] = any(e["flagged"] for e in out["energy"].values())
        log("energy series done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload, args.n)
        log("cpu baseline done")
    if not args.no_configs and args.workload == "disk" and args.n == 1_000_000:
        cfg = {}
        cfg["plummer10k"] = other_config(args, ag2, "plummer", 10_000, 10, 2, 10, 10_000, ("run", 10, 0))
        log("plummer done")
        cfg["disk4m"] = other_config(args, ag2, "disk", 4_000_000, 10, 2, 50, args.sample, ("force", 1, 0))
        log("disk4m done")
        cfg["merger16m"] = other_config(args, ag2, "merger", 16_000_000, 5, 1, 0, args.sample,
                                        ("force", 1, 4_000_000))
        log("merger16m done")
        # the same merger as specified, with the reference's gas rules (the
        # escaping gas deepens the tree to some 50 levels within these steps)
        cfg["merger16m_gas_sources"] = other_config(args, ag2, "merger", 16_000_000, 5, 1, 0, 0, None, gas_sources=True)
        log("merger16m with gas sources done")
        out["configs"] = cfg
    if not args.no_serving:
        out["serving"] = serving(args, ag2)
    print(json.dumps(finite_json(out)))


# ------------------------------------------------------------------ profiles
def profiled_traffic(kernel, form):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by a
    plain copy) from the committed counter summary of this workload, or None
    when no summary holds the kernel."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_disk1m_%s.md" % form)
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 3 and cells[0] == kernel:
            return (float(cells[1]) + float(cells[2])) * 1e6
    return None


# HBM-bound kernels of the build, density and kick phases (k_permute is left
# out: its averages are dominated by the first build after upload, which
# gathers from caller order, and the two summaries hold different numbers
# of such builds).
HBM_KERNELS = ("k_node_moments", "k_sums_direct", "k_tile_sums", "k_make_leaf", "k_make_internal", "k_keygen",
               "k_rs_scatter", "k_kick_drift", "k_kick2", "k_sph_pairs", "k_group_rho")
HBM_PEAK_GBS = 8000.0


def _md_table(path):
    rows = {}
    if os.path.exists(path):
        for line in open(path):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 3 and cells[0].startswith("k_"):
                rows[cells[0]] = cells[1:]
    return rows


def profiled_hbm(form):
    """Achieved HBM bandwidth of the HBM-bound kernels: bytes per launch from
    the committed counter summary over the average launch duration from the
    committed kernel-trace summary of the same workload."""
    traffic = _md_table(os.path.join(ROOT, "profiles", "hbm_traffic_disk1m_%s.md" % form))
    trace = _md_table(os.path.join(ROOT, "profiles", "kernel_trace_disk1m_%s.md" % form))
    out = {}
    for k in HBM_KERNELS:
        if k in traffic and k in trace:
            mb = float(traffic[k][0]) + float(traffic[k][1])
            us = float(trace[k][1])
            gbs = mb * 1e6 / (us * 1e-6) / 1e9
            out[k] = {"MB": round(mb, 1), "avg_us": us, "GB_s": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 3)}
    return out


if __name__ == "__main__":
    main()
