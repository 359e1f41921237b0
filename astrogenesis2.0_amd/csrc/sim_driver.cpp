// The reference's driver with its tree calls replaced by the engine's C ABI:
// a Particle array of structures and a Simulation with init() and run(), as
// in simulation/src/Physics/Particle.h:9-57 and Simulation.cpp:25-364, where
// `new Tree(this); buildTree(); calcGasDensity(); calculateForces();` and the
// integrator loop become ag2_upload, ag2_init_forces, ag2_step and
// ag2_download (include/ag2.h).  Compiled with the host compiler against
// include/ag2.h and linked to libag2.so; no HIP header is involved.
//
//   sim_driver <in.bin> <out.bin> <steps> <theta> <e0> <massInH> <dt> <split_threshold> <precision>
//
// in.bin: int64 n, then n doubles each of x y z vx vy vz m U mu, then n uint8
// types.  out.bin: int64 n, then n doubles each of x y z vx vy vz U ax ay az h
// rho P T dUdt, in the input order.  Exit status 0 on success.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/ag2.h"

namespace {

const double kG = 6.67430e-11;   // Constants::G (Constants.h:7)

struct Vec3 {
    double x = 0, y = 0, z = 0;
};

// The fields of Particle.h:18-49 the tree phases and the integrator use.
struct Particle {
    Vec3 position, velocity, acc;
    double mass = 0;
    uint8_t type = 1;
    double h = 0, rho = 0, P = 0, T = 0, mu = 0.58, dUdt = 0, U = 0;
};

struct Simulation {
    std::vector<Particle> particles;
    double theta = 0.5, e0 = 0, massInH = 1, minTimeStep = 0, maxTimeStep = 0, eta = 2.0, H0 = 0;
    int64_t splitThreshold = 800;
    int precision = AG2_FP64;
    void *ag2 = nullptr;
    // structure-of-arrays staging, filled at upload and download only
    std::vector<double> sx, sy, sz, svx, svy, svz, sm, sU, smu, sax, say, saz, sh, srho, sP, sT, sdUdt;
    std::vector<uint8_t> stype;

    ag2_particles soa() {
        const size_t n = particles.size();
        for (auto *v : {&sx, &sy, &sz, &svx, &svy, &svz, &sm, &sU, &smu, &sax, &say, &saz, &sh, &srho, &sP, &sT,
                        &sdUdt})
            v->resize(n);
        stype.resize(n);
        ag2_particles h{};
        h.n = (int64_t)n;
        h.x = sx.data(); h.y = sy.data(); h.z = sz.data();
        h.vx = svx.data(); h.vy = svy.data(); h.vz = svz.data();
        h.m = sm.data(); h.U = sU.data(); h.mu = smu.data(); h.type = stype.data();
        h.ax = sax.data(); h.ay = say.data(); h.az = saz.data();
        h.h = sh.data(); h.rho = srho.data(); h.P = sP.data(); h.T = sT.data(); h.dUdt = sdUdt.data();
        return h;
    }

    void upload() {
        ag2_particles h = soa();
        for (size_t i = 0; i < particles.size(); i++) {
            const Particle &p = particles[i];
            sx[i] = p.position.x; sy[i] = p.position.y; sz[i] = p.position.z;
            svx[i] = p.velocity.x; svy[i] = p.velocity.y; svz[i] = p.velocity.z;
            sax[i] = p.acc.x; say[i] = p.acc.y; saz[i] = p.acc.z;
            sm[i] = p.mass; sU[i] = p.U; smu[i] = p.mu; stype[i] = p.type;
            sh[i] = p.h; srho[i] = p.rho; sP[i] = p.P; sT[i] = p.T; sdUdt[i] = p.dUdt;
        }
        check(ag2_upload(ag2, &h), "ag2_upload");
    }

    void download() {
        ag2_particles h = soa();
        check(ag2_download(ag2, &h), "ag2_download");
        for (size_t i = 0; i < particles.size(); i++) {
            Particle &p = particles[i];
            p.position = {sx[i], sy[i], sz[i]};
            p.velocity = {svx[i], svy[i], svz[i]};
            p.acc = {sax[i], say[i], saz[i]};
            p.U = sU[i]; p.h = sh[i]; p.rho = srho[i]; p.P = sP[i]; p.T = sT[i]; p.dUdt = sdUdt[i];
        }
    }

    void check(int rc, const char *what) {
        if (rc != AG2_OK) {
            std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, ag2 ? ag2_last_error(ag2) : "");
            std::exit(2);
        }
    }

    // Simulation::init (Simulation.cpp:121-139): the first tree and forces.
    void init() {
        ag2_params p{};
        p.theta = theta; p.e0 = e0; p.massInH = massInH; p.G = kG;
        p.split_threshold = splitThreshold;   // 100 * cores in the reference (Node.cpp:420)
        p.precision = precision; p.device = 0;
        p.eta = eta; p.min_dt = minTimeStep; p.max_dt = maxTimeStep; p.H0 = H0;
        p.cooling = 0; p.star_formation = 0;  // commented out in the reference (Simulation.cpp:309-317)
        p.timing_events = 2;
        int rc = ag2_create(&p, (int64_t)particles.size(), &ag2);
        if (rc != AG2_OK) {
            std::fprintf(stderr, "ag2_create failed (%d)\n", rc);
            std::exit(2);
        }
        upload();
        check(ag2_init_forces(ag2), "ag2_init_forces");
    }

    // Simulation::run (Simulation.cpp:166-364): each iteration of its loop is
    // one ag2_step; the particles come back for the snapshot at the end.
    void run(int steps) {
        for (int s = 0; s < steps; s++) check(ag2_step(ag2), "ag2_step");
        check(ag2_sync(ag2), "ag2_sync");
        download();
    }

    ~Simulation() {
        if (ag2) ag2_destroy(ag2);
    }
};

bool read_all(FILE *f, void *p, size_t bytes) { return std::fread(p, 1, bytes, f) == bytes; }

}  // namespace

int main(int argc, char **argv) {
    if (argc != 10) {
        std::fprintf(stderr, "usage: %s in.bin out.bin steps theta e0 massInH dt split_threshold precision\n", argv[0]);
        return 1;
    }
    Simulation sim;
    const int steps = std::atoi(argv[3]);
    sim.theta = std::atof(argv[4]);
    sim.e0 = std::atof(argv[5]);
    sim.massInH = std::atof(argv[6]);
    sim.minTimeStep = sim.maxTimeStep = std::atof(argv[7]);
    sim.splitThreshold = std::atoll(argv[8]);
    sim.precision = std::atoi(argv[9]);
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 1;
    int64_t n = 0;
    if (!read_all(f, &n, 8) || n < 1) return 1;
    std::vector<double> col((size_t)n);
    sim.particles.resize((size_t)n);
    for (int k = 0; k < 9; k++) {
        if (!read_all(f, col.data(), (size_t)n * 8)) return 1;
        for (int64_t i = 0; i < n; i++) {
            Particle &p = sim.particles[(size_t)i];
            double v = col[(size_t)i];
            switch (k) {
                case 0: p.position.x = v; break;
                case 1: p.position.y = v; break;
                case 2: p.position.z = v; break;
                case 3: p.velocity.x = v; break;
                case 4: p.velocity.y = v; break;
                case 5: p.velocity.z = v; break;
                case 6: p.mass = v; break;
                case 7: p.U = v; break;
                default: p.mu = v; break;
            }
        }
    }
    std::vector<uint8_t> types((size_t)n);
    if (!read_all(f, types.data(), (size_t)n)) return 1;
    std::fclose(f);
    for (int64_t i = 0; i < n; i++) sim.particles[(size_t)i].type = types[(size_t)i];

    sim.init();
    sim.run(steps);

    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 1;
    std::fwrite(&n, 8, 1, o);
    for (int k = 0; k < 15; k++) {
        for (int64_t i = 0; i < n; i++) {
            const Particle &p = sim.particles[(size_t)i];
            const double v[15] = {p.position.x, p.position.y, p.position.z, p.velocity.x, p.velocity.y, p.velocity.z,
                                  p.U, p.acc.x, p.acc.y, p.acc.z, p.h, p.rho, p.P, p.T, p.dUdt};
            col[(size_t)i] = v[k];
        }
        std::fwrite(col.data(), 8, (size_t)n, o);
    }
    std::fclose(o);
    return 0;
}
