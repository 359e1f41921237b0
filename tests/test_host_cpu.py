"""CPU tests of the engine's host side: the CPU build of the shared per-thread
core against the oracle, the file formats, and the C ABI's exports."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, PKG_DIR, coincident_groups, fixture_input, vec_rel_err

CORE_SO = os.path.join(PKG_DIR, "libag2core_cpu.so")
HEADER = os.path.join(ROOT, "include", "ag2.h")


@pytest.fixture(scope="module")
def core_lib():
    if not os.path.exists(CORE_SO):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", CORE_SO,
                        os.path.join(PKG_DIR, "csrc", "core_cpu.cpp")], check=True)
    L = C.CDLL(CORE_SO)
    L.core_cpu_force_step.restype = C.c_int64
    return L


def core_step(L, ps, theta, e0, mih, thr):
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    cap = 8 * ps.n + 256   # deep trees add chains of single-child cells
    out = dict(level=np.zeros(cap, np.int32), first=np.zeros(cap, np.int64), count=np.zeros(cap, np.int64),
               mass=np.zeros(cap), com=np.zeros(3 * cap), gas=np.zeros(cap), perm=np.zeros(ps.n, np.int64),
               R=np.zeros(1))
    arrs = [P(getattr(ps, f)) for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu", "type", "ax", "ay", "az",
                                        "h", "rho", "P", "T", "dUdt")]
    nn = L.core_cpu_force_step(C.c_int64(ps.n), *arrs, C.c_double(theta), C.c_double(e0), C.c_double(mih),
                               C.c_double(6.67430e-11), C.c_int64(thr), C.c_int64(cap), P(out["level"]),
                               P(out["first"]), P(out["count"]), P(out["mass"]), P(out["com"]), P(out["gas"]),
                               P(out["perm"]), P(out["R"]))
    if nn == -2:
        raise DepthError("distinct particles share all 126 levels")
    assert nn > 0
    for k in ("level", "first", "count", "mass", "gas"):
        out[k] = out[k][:nn]
    out["com"] = out["com"][:3 * nn].reshape(nn, 3)
    return out


class DepthError(RuntimeError):
    pass


def cpu_cases(ag2):
    U = ag2.units
    ics = ag2.ics
    pl = ics.config1_plummer(2500, seed=12)
    ties = pl.copy()
    ties.x[100:120] = ties.x[:20]
    ties.y[100:120] = ties.y[:20]
    ties.z[100:120] = ties.z[:20] * (1 + 1e-13)
    ds = ics.shuffled(ics.disk_galaxy(900, 700, 1400, seed=13), 3)
    cp, cp_e0, cp_m = fixture_input(ag2, "close_pairs2k")
    rw, rw_e0, rw_m = fixture_input(ag2, "runaway2k")
    return [("plummer", pl, 0.05 * U.KPC, 1.0), ("ties", ties, 0.05 * U.KPC, 1.0),
            ("disk", ds, 0.1 * U.KPC, ics.mass_in_h(ds)), ("close_pairs", cp, cp_e0, cp_m),
            ("runaway", rw, rw_e0, rw_m), ("coincident", coincident_groups(ics), 0.05 * U.KPC, 1.0)]


CPU_CASE_NAMES = ["plummer", "ties", "disk", "close_pairs", "runaway", "coincident"]


@pytest.mark.parametrize("case", range(6), ids=CPU_CASE_NAMES)
@pytest.mark.parametrize("thr", [100, 800])
def test_core_cpu_matches_oracle(ag2, orc, core_lib, case, thr):
    name, ps, e0, mih = cpu_cases(ag2)[case]
    a, b = ps.copy(), ps.copy()
    orc.force_step(a, orc.params(0.5, e0, mih, thr))
    t = core_step(core_lib, b, 0.5, e0, mih, thr)
    ref = orc.tree_export(ps.copy(), orc.params(0.5, e0, mih, thr))
    for f in ("level", "first", "count"):
        assert np.array_equal(t[f], ref[f]), (name, f)
    assert np.array_equal(t["perm"][:ref["n_in_tree"]], ref["perm"])
    assert vec_rel_err(b.ax, b.ay, b.az, a.ax, a.ay, a.az).max() < 1e-10
    assert np.array_equal(b.h, a.h)
    gas = ps.type == 2
    if gas.any():
        assert (np.abs(b.rho - a.rho)[gas] <= 1e-12 * np.abs(a.rho)[gas]).all()


def test_age_roundtrip(ag2, tmp_path):
    ps = ag2.ics.shuffled(ag2.ics.disk_galaxy(300, 200, 100, seed=3), 1)
    p = tmp_path / "s.age"
    ag2.ageio.write_age(str(p), ps, 1.0, 2.0, 3.0)
    assert os.path.getsize(p) == 40 + 94 * ps.n
    q, hdr = ag2.ageio.read_age(str(p))
    assert hdr == {"deltaTime": 1.0, "endTime": 2.0, "currentTime": 3.0}
    for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "T", "type", "id"):
        assert np.array_equal(getattr(q, f), getattr(ps, f)), f


@pytest.mark.skipif(not os.path.exists("/root/reference/input_data/Example/gassphere_littleendian.dat"),
                    reason="reference example files not present")
def test_read_reference_gadget_example(ag2):
    ps = ag2.ageio.read_gadget("/root/reference/input_data/Example/gassphere_littleendian.dat")
    assert ps.n == 1472 and (ps.type == 2).all()
    assert np.all(ps.U > 0) and np.all(ps.m > 0)


def test_ic_generators_are_seeded(ag2):
    a = ag2.ics.config1_plummer(500, seed=7)
    b = ag2.ics.config1_plummer(500, seed=7)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.vz, b.vz)
    g = ag2.ics.disk_galaxy(200, 150, 100, seed=2)
    assert (g.type[:200] == 3).all() and (g.type[200:350] == 1).all() and (g.type[350:] == 2).all()
    assert abs(g.U[-1] - ag2.units.u_from_temperature(1e4)) < 1e-6 * g.U[-1]
    h = ag2.ics.hernquist(300, seed=4)
    assert h.n == 300 and np.isfinite(h.vx).all()


@pytest.mark.parametrize("sep,peri", [(50.0, 8.0), (20.0, 20.0), (300.0, 1.0)])
def test_parabolic_orbit_closed_form(ag2, sep, peri):
    """Zero orbital energy, pericentre L^2 / (2 G M), separation as asked and
    on the way in."""
    G, M = 4.30091e-6, 2.1e12
    x, v = ag2.ics.parabolic_orbit(M, sep, peri, G)
    r = np.linalg.norm(x)
    assert abs(r - sep) < 1e-12 * sep
    assert abs(0.5 * v @ v - G * M / r) < 1e-12 * G * M / r
    L = np.cross(x, v)
    assert abs(L @ L / (2 * G * M) - peri) < 1e-12 * peri
    assert x @ v <= 1e-9 * r * np.linalg.norm(v) and x[2] == 0 and v[2] == 0
    with pytest.raises(ValueError):
        ag2.ics.parabolic_orbit(M, peri * 0.5, peri, G)


def test_merger_ics(ag2):
    """Two seeded disks: own streams, the second inclined by 60 degrees about
    x, relative centre-of-mass motion equal to the parabolic orbit."""
    ics, U = ag2.ics, ag2.units
    n = 2000
    ps = ics.merger(n)
    assert ps.n == 2 * n and np.array_equal(ps.id, np.arange(2 * n))
    one = ics.disk_galaxy(n // 2, 3 * n // 10, n - n // 2 - 3 * n // 10, seed=5, config=5, job=1)
    two = ps.take(np.arange(n, 2 * n))
    assert np.array_equal(two.type, one.type) and np.array_equal(two.m, one.m)
    assert not np.array_equal(ps.x[:n] - ps.x[:n].mean(), two.x - two.x.mean())
    a, b = ps.take(np.arange(n)), two
    cm = lambda p, f: float((p.m * getattr(p, f)).sum() / p.m.sum())
    ma, mb = a.m.sum(), b.m.sum()
    rel_x, rel_v = ics.parabolic_orbit((ma + mb) / U.MSUN, 50.0, 8.0)
    # the first disk is only shifted; the second is rotated then shifted
    base = ics.disk_galaxy(n // 2, 3 * n // 10, n - n // 2 - 3 * n // 10, seed=5, config=5, job=0)
    assert np.allclose(a.x - base.x, -mb / (ma + mb) * rel_x[0] * U.KPC, rtol=0, atol=1e-6 * U.KPC)
    c, s = np.cos(np.pi / 3), np.sin(np.pi / 3)
    assert rel_x[2] == 0.0
    assert np.allclose(b.z, s * one.y + c * one.z, rtol=1e-12, atol=1e-9 * U.KPC)
    for k, f in enumerate("xyz"):
        d = (cm(b, f) - cm(a, f)) - ((cm(one, f) if f == "x" else 0.0) - cm(base, f))
        if f == "x":
            assert abs(d - rel_x[k] * U.KPC) < 1e-9 * 50 * U.KPC
        dv = cm(b, "v" + f) - cm(a, "v" + f)
        want = rel_v[k] * U.KMS + (cm(one, "vx") - cm(base, "vx") if f == "x" else
                                   (c * cm(one, "vy") - s * cm(one, "vz") - cm(base, "vy")) if f == "y" else
                                   (s * cm(one, "vy") + c * cm(one, "vz") - cm(base, "vz")))
        assert abs(dv - want) < 1e-9 * abs(rel_v).max() * U.KMS


def test_bench_reads_committed_profiles():
    """bench.py's roofline traffic and HBM fractions come from the committed
    counter and kernel-trace summaries of the 1M disk: both must hold the
    walk kernel and the HBM-bound kernels it reports."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    t = b.profiled_traffic("k_walk_block32", "fp32")
    assert t is not None and t > 0
    h = b.profiled_hbm("fp32")
    assert set(h) == set(b.HBM_KERNELS)
    for k, v in h.items():
        assert 0 < v["frac_of_peak"] < 1, k


def header_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(ag2_\w+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol(ag2):
    from importlib import import_module  # noqa: F401
    E = ag2.engine
    L = E.load_library()
    syms = header_symbols()
    assert len(syms) >= 17
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) <= set(E.EXPORTS) | {"ag2_run_nocopy"}


def test_params_struct_matches_header(ag2, tmp_path):
    """ctypes mirrors against the C compiler's layout of include/ag2.h."""
    E = ag2.engine
    structs = {"ag2_params": E.Params, "ag2_particles": E.HostParticles, "ag2_diag": E.Diag,
               "ag2_pool_params": E.PoolParams, "ag2_pool_stats": E.PoolStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ag2.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append('printf("%%s %%zu\\n", "%s", sizeof(%s));' % (name, name))
        for f, _ in cls._fields_:
            lines.append('printf("%%s.%%s %%zu\\n", "%s", "%s", offsetof(%s, %s));' % (name, f, name, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.split("\n") if l)
    for name, cls in structs.items():
        assert int(out[name]) == C.sizeof(cls), name
        for f, _ in cls._fields_:
            assert int(out["%s.%s" % (name, f)]) == getattr(cls, f).offset, (name, f)


def test_core_cpu_gas_mass_ties(ag2, orc, core_lib):
    """Equal-mass gas: group selection compares node gas masses that can be
    exact ties in real arithmetic (a child of (64-j) m against a parent of
    (64+j) m); the reference breaks them by the rounding of its running sums,
    which the engine reproduces.  Positions after one step of this disk hit
    such ties."""
    U = ag2.units
    ps = ag2.ics.shuffled(ag2.ics.disk_galaxy(1500, 1000, 1500, seed=21), 2)
    e0, mih, dt = 0.1 * U.KPC, ag2.ics.mass_in_h(ps), 0.5 * U.MYR
    a = ps.copy()
    orc.run(a, orc.params(0.5, e0, mih, 100), 2.0, dt, dt, 0.5 * dt)
    b, c = a.copy(), a.copy()
    b.dUdt[:] = 0.0
    c.dUdt[:] = 0.0
    orc.force_step(b, orc.params(0.5, e0, mih, 100))
    core_step(core_lib, c, 0.5, e0, mih, 100)
    assert np.array_equal(b.h, c.h)
    gas = ps.type == 2
    assert (np.abs(b.rho - c.rho)[gas] <= 1e-12 * np.abs(b.rho)[gas]).all()


def test_split_precision_refuses_live_softening_branch(ag2):
    """The FP32 sweep evaluates only G m / (r^2 + e0^2): where the reference's
    softening-kernel branch can fire (e0 * 0.0001 < 2147483647, Node.cpp:302)
    ag2_create refuses the split-precision form before touching a device; the
    bound is the reference's own comparison."""
    E = ag2.engine
    L = E.load_library()
    for e0, want in ((1e4, E.AG2_ERR_ARG), (2.147483646e13, E.AG2_ERR_ARG)):
        ctx = C.c_void_p()
        rc = L.ag2_create(C.byref(E.make_params(e0=e0, precision=E.AG2_FP32_SPLIT)), 10, C.byref(ctx))
        assert rc == want, e0
    # at and above the bound the parameters are valid (creation then needs a device)
    ctx = C.c_void_p()
    rc = L.ag2_create(C.byref(E.make_params(e0=2.147483647e13, precision=E.AG2_FP32_SPLIT)), 10, C.byref(ctx))
    assert rc != E.AG2_ERR_ARG
    if rc == E.AG2_OK:
        L.ag2_destroy(ctx)


def test_core_cpu_depth_limit(ag2, orc, core_lib):
    """Distinct particles closer than 2^-126 of the root half-width (near the
    origin, where the cell centres stay exact that deep): the engine's core
    refuses the tree and the oracle reports the same depth error."""
    ps = ag2.ics.config1_plummer(500, seed=3)
    ps.x[10], ps.y[10], ps.z[10] = 1e-200, 1e-200, 1e-200
    ps.x[11], ps.y[11], ps.z[11] = 2e-200, 1e-200, 1e-200
    with pytest.raises(DepthError):
        core_step(core_lib, ps.copy(), 0.5, 0.05 * ag2.units.KPC, 1.0, 100)
    with pytest.raises(orc.DepthError):
        orc.force_step(ps.copy(), orc.params(0.5, 0.05 * ag2.units.KPC, 1.0, 100))


def test_hoisted_sph_term_is_the_reference_form(core_lib):
    """The force walk's SPH pair term (per-target constants hoisted) equals
    the term as Node::calcSPHForce writes it (Node.cpp:88-172) bit for bit,
    also where a last-bit change would grow: pairs just inside the kernel's
    edge (q = r / h -> 2, where the term is (2 - q)^2), at the join of its
    two pieces (q = 1), approaching and receding, with negative pressure,
    and at the scale of gas flung to 1e42 m (h 8e41 m, rho 3e-90)."""
    rng = np.random.default_rng(5)
    n = 20000
    h = 10.0 ** rng.uniform(-3, 3, n) * 3.0857e19
    h[:2000] = 7.992066773505963e41
    rho = 10.0 ** rng.uniform(-24, -18, n)
    rho[:2000] = 10.0 ** rng.uniform(-92, -88, 2000)
    P = 10.0 ** rng.uniform(-14, -8, n) * np.where(rng.random(n) < 0.25, -1.0, 1.0)
    P[:2000] = -1677319609.610992 * rng.uniform(0.5, 2.0, 2000)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    eps = 10.0 ** rng.uniform(-16, 0, n)
    q = np.where(rng.random(n) < 0.7, 2.0 * (1.0 - eps), 1.0 + np.where(rng.random(n) < 0.5, -1, 1) * eps * 0.5)
    q = np.clip(q, 1e-6, np.nextafter(2.0, 0.0))
    d = u * (q * h)[:, None]
    # the walk's own support test (r < 2h, r the rounded |d|) picks the pairs
    inside = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) < h * 2
    assert inside.sum() > 0.95 * n
    n = int(inside.sum())
    d, h, rho, P = d[inside], h[inside], rho[inside], P[inside]
    v = rng.normal(size=(n, 3)) * 1e4
    gas = 10.0 ** rng.uniform(30, 40, n)
    plain = np.zeros(4 * n)
    hoisted = np.zeros(4 * n)
    P_ = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.c_void_p)  # noqa: E731
    cols = [np.ascontiguousarray(a, dtype=np.float64) for a in (d[:, 0], d[:, 1], d[:, 2], gas, v[:, 0], v[:, 1],
                                                              v[:, 2], h, rho, P)]
    core_lib.core_cpu_sph_terms(C.c_int64(n), *[P_(a) for a in cols], P_(plain), P_(hoisted))
    assert np.isfinite(plain[0::4]).sum() > n // 2
    assert np.array_equal(plain.view(np.uint64)[np.isfinite(plain)], hoisted.view(np.uint64)[np.isfinite(plain)])
    assert np.array_equal(np.isnan(plain), np.isnan(hoisted))
