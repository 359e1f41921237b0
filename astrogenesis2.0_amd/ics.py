"""Seeded initial-condition generators and the structure-of-arrays particle set.

The reference reads its initial conditions from files only
(simulation/src/File/DataManager.cpp:434-1331); these generators produce the
synthetic workloads the engine is measured on, converted to the reference's SI
units with its own constants (Units.h) so that every implementation sees the
same FP64 bits.

Random streams: numpy's PCG64 seeded through SeedSequence((seed, config,
component, job)), so that every component of every configuration and job has
its own independent, reproducible stream.
"""
import math

import numpy as np

from . import units as U

FIELDS_F64 = ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu",
              "ax", "ay", "az", "h", "rho", "P", "T", "dUdt")


class ParticleSet:
    """Structure-of-arrays particle state (the hot subset of Particle.h:9-57).

    type: 1 star, 2 gas, 3 dark matter (Particle.h:32).  Units are SI.
    """

    def __init__(self, n):
        self.n = int(n)
        for f in FIELDS_F64:
            setattr(self, f, np.zeros(self.n, dtype=np.float64))
        self.mu[:] = U.MU_DEFAULT
        self.type = np.ones(self.n, dtype=np.uint8)
        self.id = np.arange(self.n, dtype=np.uint32)

    @classmethod
    def concat(cls, parts):
        out = cls(sum(p.n for p in parts))
        k = 0
        for p in parts:
            for f in FIELDS_F64 + ("type", "id"):
                getattr(out, f)[k:k + p.n] = getattr(p, f)
            k += p.n
        out.id[:] = np.arange(out.n, dtype=np.uint32)
        return out

    def copy(self):
        out = ParticleSet(self.n)
        for f in FIELDS_F64 + ("type", "id"):
            setattr(out, f, getattr(self, f).copy())
        return out

    def take(self, idx):
        idx = np.asarray(idx)
        out = ParticleSet(len(idx))
        for f in FIELDS_F64 + ("type", "id"):
            setattr(out, f, getattr(self, f)[idx].copy())
        return out

    def init_temperature(self):
        """T from U for gas, as Simulation::init does (Simulation.cpp:101-113)."""
        g = self.type == U.TYPE_GAS
        self.T[g] = (U.GAMMA - 1.0) * self.U[g] * U.M_PROTON * self.mu[g] / U.K_B
        return self

    def arrays(self):
        return {f: getattr(self, f) for f in FIELDS_F64 + ("type", "id")}


def rng(seed, config=0, component=0, job=0):
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, config, component, job])))


def _isotropic(g, n):
    cz = g.uniform(-1.0, 1.0, n)
    ph = g.uniform(0.0, 2.0 * math.pi, n)
    s = np.sqrt(np.maximum(0.0, 1.0 - cz * cz))
    return s * np.cos(ph), s * np.sin(ph), cz


def _to_si(ps, pos_kpc, vel_kms, mass_msun):
    ps.x[:], ps.y[:], ps.z[:] = (pos_kpc[0] * U.KPC, pos_kpc[1] * U.KPC, pos_kpc[2] * U.KPC)
    ps.vx[:], ps.vy[:], ps.vz[:] = (vel_kms[0] * U.KMS, vel_kms[1] * U.KMS, vel_kms[2] * U.KMS)
    ps.m[:] = mass_msun * U.MSUN
    return ps


def plummer(n, mass_msun=1e10, a_kpc=1.0, rmax_a=20.0, seed=1, config=1, component=0, job=0,
            G=4.30091e-6, zero_velocity=False, ptype=U.TYPE_STAR):
    """Plummer sphere: radii from the inverse cumulative mass truncated at
    rmax_a * a, isotropic directions, speeds by Aarseth, Henon & Wielen (1974)
    rejection from the isotropic distribution function."""
    g = rng(seed, config, component, job)
    xmax = rmax_a ** 3 / (1.0 + rmax_a ** 2) ** 1.5
    X = g.uniform(0.0, xmax, n)
    X = np.maximum(X, 1e-300)
    r = a_kpc / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
    ux, uy, uz = _isotropic(g, n)
    pos = (r * ux, r * uy, r * uz)
    vel = (np.zeros(n),) * 3
    if not zero_velocity:
        q = np.empty(n)
        todo = np.arange(n)
        while todo.size:
            x1 = g.uniform(0.0, 1.0, todo.size)
            y1 = g.uniform(0.0, 0.1, todo.size)
            ok = y1 < x1 * x1 * (1.0 - x1 * x1) ** 3.5
            q[todo[ok]] = x1[ok]
            todo = todo[~ok]
        vesc = np.sqrt(2.0 * G * mass_msun / np.sqrt(r * r + a_kpc * a_kpc))
        v = q * vesc
        vx, vy, vz = _isotropic(g, n)
        vel = (v * vx, v * vy, v * vz)
    ps = ParticleSet(n)
    ps.type[:] = ptype
    return _to_si(ps, pos, vel, mass_msun / n)


def _jeans_sigma(r_eval, r_grid, rho_grid, menc_grid, G):
    """Isotropic Jeans equation sigma_r(r) = sqrt(1/rho * int_r^rmax rho G M / r'^2 dr')."""
    integrand = rho_grid * G * menc_grid / r_grid ** 2
    seg = 0.5 * (integrand[1:] + integrand[:-1]) * np.diff(r_grid)
    tail = np.concatenate([np.cumsum(seg[::-1])[::-1], [0.0]])
    s2 = np.interp(r_eval, r_grid, tail / np.maximum(rho_grid, 1e-300))
    return np.sqrt(np.maximum(s2, 0.0))


def hernquist(n, mass_msun=1e10, a_kpc=1.0, rmax_a=50.0, seed=1, config=4, component=0, job=0,
              G=4.30091e-6, ptype=U.TYPE_STAR):
    """Hernquist sphere: inverse cumulative mass M r^2/(r+a)^2 truncated at
    rmax_a * a, Gaussian velocities from the isotropic Jeans equation."""
    g = rng(seed, config, component, job)
    xmax = (rmax_a / (1.0 + rmax_a)) ** 2
    s = np.sqrt(g.uniform(1e-12, xmax, n))
    r = a_kpc * s / (1.0 - s)
    ux, uy, uz = _isotropic(g, n)
    rg = np.geomspace(1e-4 * a_kpc, rmax_a * a_kpc, 2048)
    rho = mass_msun * a_kpc / (2 * math.pi * rg * (rg + a_kpc) ** 3)
    menc = mass_msun * rg ** 2 / (rg + a_kpc) ** 2
    sig = _jeans_sigma(r, rg, rho, menc, G)
    ps = ParticleSet(n)
    ps.type[:] = ptype
    vel = (sig * g.standard_normal(n), sig * g.standard_normal(n), sig * g.standard_normal(n))
    return _to_si(ps, (r * ux, r * uy, r * uz), vel, mass_msun / n)


def _nfw_mu(x):
    return np.log1p(x) - x / (1.0 + x)


def disk_galaxy(n_halo=500_000, n_disk=300_000, n_gas=200_000, seed=2, config=2,
                m200=1e12, conc=10.0, md=4e10, mgas=1e10, rd=3.0, z0=0.3,
                T_gas=1e4, H0_kms_mpc=70.0, G=4.30091e-6, job=0):
    """Isolated disk galaxy: NFW halo (truncated at r200, Jeans velocities with
    Gaussian components), exponential stellar disk with sech^2 vertical profile
    (circular velocity plus 10 % Gaussian dispersion) and a gas disk of the
    same profile at temperature T_gas.  Particle order: halo, disk, gas."""
    H = H0_kms_mpc / 1000.0  # km/s/kpc
    r200 = (G * m200 / (100.0 * H * H)) ** (1.0 / 3.0)
    rs = r200 / conc
    # halo radii: inverse CDF of mu(r/rs)/mu(c)
    gh = rng(seed, config, 0, job)
    xg = np.geomspace(1e-6, conc, 8192)
    cdf = _nfw_mu(xg) / _nfw_mu(conc)
    r_h = rs * np.interp(gh.uniform(cdf[0], 1.0, n_halo), cdf, xg)
    ux, uy, uz = _isotropic(gh, n_halo)
    # enclosed mass of the baryonic disk (spherical approximation)
    mb = md + mgas
    rg = np.geomspace(1e-4, r200, 4096)
    mdisk_enc = mb * (1.0 - (1.0 + rg / rd) * np.exp(-rg / rd))
    mhalo_enc = m200 * _nfw_mu(rg / rs) / _nfw_mu(conc)
    rho_h = m200 / (4 * math.pi * rs ** 3 * _nfw_mu(conc)) / ((rg / rs) * (1 + rg / rs) ** 2)
    sig = _jeans_sigma(r_h, rg, rho_h, mhalo_enc + mdisk_enc, G)
    halo = ParticleSet(n_halo)
    halo.type[:] = U.TYPE_DM
    _to_si(halo, (r_h * ux, r_h * uy, r_h * uz),
           (sig * gh.standard_normal(n_halo), sig * gh.standard_normal(n_halo), sig * gh.standard_normal(n_halo)),
           m200 / n_halo)

    def exp_disk(nn, mtot, comp, ptype):
        gd = rng(seed, config, comp, job)
        Rg = np.linspace(0.0, 10.0 * rd, 8192)
        cdfR = 1.0 - (1.0 + Rg / rd) * np.exp(-Rg / rd)
        cdfR /= cdfR[-1]
        R = np.interp(gd.uniform(0.0, 1.0, nn), cdfR, Rg)
        ph = gd.uniform(0.0, 2.0 * math.pi, nn)
        z = z0 * np.arctanh(np.clip(2.0 * gd.uniform(0.0, 1.0, nn) - 1.0, -0.999999, 0.999999))
        Rs = np.maximum(R, 1e-3)
        menc = np.interp(Rs, rg, mhalo_enc + mdisk_enc)
        vc = np.sqrt(G * menc / Rs)
        sd = 0.1 * vc
        vR = sd * gd.standard_normal(nn)
        vp = vc + sd * gd.standard_normal(nn)
        vz = sd * gd.standard_normal(nn)
        c, s = np.cos(ph), np.sin(ph)
        ps = ParticleSet(nn)
        ps.type[:] = ptype
        _to_si(ps, (R * c, R * s, z), (vR * c - vp * s, vR * s + vp * c, vz), mtot / nn)
        return ps

    disk = exp_disk(n_disk, md, 1, U.TYPE_STAR)
    gas = exp_disk(n_gas, mgas, 2, U.TYPE_GAS)
    gas.U[:] = U.u_from_temperature(T_gas)
    out = ParticleSet.concat([halo, disk, gas])
    out.init_temperature()
    return out


def parabolic_orbit(m_total_msun, sep_kpc, peri_kpc, G=4.30091e-6):
    """Relative position (kpc) and velocity (km/s) of two point masses of
    total mass m_total on a parabolic (e = 1) orbit with pericentre peri_kpc,
    at separation sep_kpc on the way in.  The orbit lies in the x-y plane with
    the pericentre on +x: r = p / (1 + cos f) with p = 2 q, radial speed
    sqrt(G M / p) sin f (negative: approaching), tangential speed
    sqrt(G M / p) (1 + cos f); |v|^2 = 2 G M / r."""
    p = 2.0 * peri_kpc
    cf = p / sep_kpc - 1.0
    if not -1.0 < cf <= 1.0:
        raise ValueError("separation must be at least the pericentre")
    sf = -math.sqrt(max(0.0, 1.0 - cf * cf))
    k = math.sqrt(G * m_total_msun / p)
    vr, vt = k * sf, k * (1.0 + cf)
    pos = np.array([sep_kpc * cf, sep_kpc * sf, 0.0])
    vel = np.array([vr * cf - vt * sf, vr * sf + vt * cf, 0.0])
    return pos, vel


def merger(n_each=8_000_000, sep_kpc=50.0, peri_kpc=8.0, incl_deg=60.0, seed=5, config=5, **disk):
    """Two disk galaxies of the disk_galaxy recipe, each of n_each particles
    split 50 % halo / 30 % stars / 20 % gas, on a parabolic orbit
    (parabolic_orbit) about their common centre of mass; the second disk is
    inclined by incl_deg about the x axis (positions and velocities).  Each
    galaxy has its own random streams (job 0 and 1).  Particle order: the first
    galaxy's halo, disk, gas, then the second's.  The reference has the same
    placement step commented out in Simulation::init (Simulation.cpp:57-90),
    with a fixed rotation and offset instead of an orbit."""
    nh, nd = n_each // 2, (3 * n_each) // 10
    ng = n_each - nh - nd
    gals = [disk_galaxy(nh, nd, ng, seed=seed, config=config, job=j, **disk) for j in (0, 1)]
    m1, m2 = (float(g.m.sum()) / U.MSUN for g in gals)
    rel_x, rel_v = parabolic_orbit(m1 + m2, sep_kpc, peri_kpc, disk.get("G", 4.30091e-6))
    c, s = math.cos(math.radians(incl_deg)), math.sin(math.radians(incl_deg))
    b = gals[1]
    for px, py in (("y", "z"), ("vy", "vz")):
        y, z = getattr(b, px).copy(), getattr(b, py).copy()
        setattr(b, px, c * y - s * z)
        setattr(b, py, s * y + c * z)
    for g, f in ((gals[0], -m2 / (m1 + m2)), (gals[1], m1 / (m1 + m2))):
        for k, ax in enumerate("xyz"):
            getattr(g, ax)[:] += f * rel_x[k] * U.KPC
            getattr(g, "v" + ax)[:] += f * rel_v[k] * U.KMS
    return ParticleSet.concat(gals)


def gas_sphere(n, mass_msun=1e9, radius_kpc=1.0, T=1e4, seed=3, config=0, component=0, job=0):
    """Uniform-density gas sphere at rest (test workload for the SPH path)."""
    g = rng(seed, config, component, job)
    r = radius_kpc * g.uniform(0.0, 1.0, n) ** (1.0 / 3.0)
    ux, uy, uz = _isotropic(g, n)
    ps = ParticleSet(n)
    ps.type[:] = U.TYPE_GAS
    _to_si(ps, (r * ux, r * uy, r * uz), (np.zeros(n),) * 3, mass_msun / n)
    ps.U[:] = U.u_from_temperature(T)
    return ps.init_temperature()


def shuffled(ps, seed):
    """The reference shuffles its particle array once at start-up
    (Simulation.cpp:114-117); a seeded permutation stands in for it."""
    return ps.take(rng(seed, 99).permutation(ps.n))


def config1_plummer(n=10_000, seed=1):
    """Plummer sphere of the CPU-runnable configuration: 1e10 Msun, a = 1 kpc,
    truncated at 20 a, Aarseth velocities."""
    return plummer(n, 1e10, 1.0, 20.0, seed=seed, config=1)


# Run parameters of the configurations, SI (theta, e0, dt, steps).
CONFIG1 = dict(theta=0.5, e0=0.05 * U.KPC, dt=0.1 * U.MYR, steps=10)
CONFIG2 = dict(theta=0.5, e0=0.1 * U.KPC, dt=0.5 * U.MYR, steps=20)
CONFIG3 = dict(theta=0.5, e0=0.1 * U.KPC, dt=0.5 * U.MYR, steps=50)


def serving_job(job, seed=4):
    """One job of the serving stream: a Plummer or Hernquist sphere (equal
    odds) with N log-uniform in [1e4, 2e5] and total mass log-uniform in
    [1e9, 1e11] Msun; 30 % of jobs turn a random 20 % of their particles into
    isothermal 1e4 K gas.  Scale radius a = 1 kpc (M / 1e10 Msun)^(1/3).
    Returns (particles, run parameters)."""
    g = rng(seed, 4, 0, job)
    n = int(round(math.exp(g.uniform(math.log(1e4), math.log(2e5)))))
    mass = math.exp(g.uniform(math.log(1e9), math.log(1e11)))
    a = (mass / 1e10) ** (1.0 / 3.0)
    kind = "plummer" if g.uniform() < 0.5 else "hernquist"
    with_gas = g.uniform() < 0.3
    if kind == "plummer":
        ps = plummer(n, mass, a, 20.0, seed=seed, config=4, component=1, job=job)
    else:
        ps = hernquist(n, mass, a, 50.0, seed=seed, config=4, component=2, job=job)
    if with_gas:
        k = g.choice(n, max(1, n // 5), replace=False)
        ps.type[k] = U.TYPE_GAS
        ps.U[k] = U.u_from_temperature(1e4)
        ps.init_temperature()
    e0 = 0.05 * U.KPC * a
    prm = dict(theta=0.5, e0=e0, massInH=mass_in_h(ps), split_threshold=800, dt=0.1 * U.MYR, steps=50,
               kind=kind, gas=with_gas)
    return ps, prm


def mass_in_h(ps, k=64):
    """SPH group target mass: k gas particle masses (64-neighbour target)."""
    g = ps.type == U.TYPE_GAS
    return float(k * ps.m[g].mean()) if g.any() else 1.0
