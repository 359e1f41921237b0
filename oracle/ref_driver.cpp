/*
 * Driver for the compiled reference (test infrastructure only).
 *
 * Built by oracle/Makefile together with the reference's own sources under
 * /root/reference/simulation/src (everything except main.cpp) into
 * oracle/_ref/libag2ref.so.  It calls the reference's Tree / Node /
 * TimeIntegration code unchanged and only replaces what main.cpp and
 * Simulation::init/run do around it (config file, IC file, shuffle, snapshot
 * output), so that tests can feed arrays in and read arrays back.
 *
 * Pinned mode: run the host process with OMP_THREAD_LIMIT=1.  Every OpenMP
 * region of the reference then runs on one thread in list order, which makes
 * the SPH density sum race-free (Node.cpp:770-779) and the octant partition
 * order fixed (Node.cpp:465-522), while `cores` still sets the
 * serial/parallel split threshold 100*cores (Node.cpp:420).
 */
#include "Simulation.h"
#include "Tree.h"
#include "Node.h"
#include "TimeIntegration.h"
#include "DataManager.h"
#include "Cooling.h"
#include "SFR.h"

#include <omp.h>
#include <cstdint>
#include <vector>
#include <limits>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <algorithm>

namespace {

struct Arrays {
    int64_t n;
    double *x, *y, *z, *vx, *vy, *vz, *m, *U, *mu;
    uint8_t *type;
    double *ax, *ay, *az, *h, *rho, *P, *T, *dUdt;
};

Simulation *make_sim(const Arrays &a, double theta, double e0, double mih) {
    Simulation *sim = new Simulation();
    sim->numberOfParticles = (int)a.n;
    sim->theta = theta;
    sim->e0 = e0;
    sim->massInH = mih;
    sim->globalTime = 0.0;
    sim->particles.reserve((size_t)a.n);
    for (int64_t i = 0; i < a.n; i++) {
        Particle *p = new Particle();
        p->position = vec3(a.x[i], a.y[i], a.z[i]);
        p->velocity = vec3(a.vx[i], a.vy[i], a.vz[i]);
        p->acc = vec3(a.ax[i], a.ay[i], a.az[i]);
        p->mass = a.m[i];
        p->U = a.U[i];
        p->mu = a.mu[i];
        p->type = a.type[i];
        p->h = a.h[i];
        p->rho = a.rho[i];
        p->P = a.P[i];
        p->T = a.T[i];
        p->dUdt = a.dUdt[i];
        p->id = (unsigned int)i;
        p->node = nullptr;
        p->nextIntegrationTime = 0.0;
        sim->particles.push_back(p);
    }
    return sim;
}

void read_back(const Simulation *sim, Arrays &a) {
    for (int64_t i = 0; i < a.n; i++) {
        const Particle *p = sim->particles[(size_t)i];
        a.x[i] = p->position.x; a.y[i] = p->position.y; a.z[i] = p->position.z;
        a.vx[i] = p->velocity.x; a.vy[i] = p->velocity.y; a.vz[i] = p->velocity.z;
        a.ax[i] = p->acc.x; a.ay[i] = p->acc.y; a.az[i] = p->acc.z;
        a.U[i] = p->U;
        a.h[i] = p->h; a.rho[i] = p->rho; a.P[i] = p->P; a.T[i] = p->T;
        a.dUdt[i] = p->dUdt;
    }
}

void free_sim(Simulation *sim) {
    for (Particle *p : sim->particles) delete p;
    sim->particles.clear();
    delete sim;
}

/* The reference's Tree::buildTree with the core count passed in. */
Tree *build(Simulation *sim, int cores) {
    for (Particle *p : sim->particles) p->node = nullptr;
    Tree *t = new Tree(sim);
    if (cores <= 0) {
        t->buildTree();
        return t;
    }
    t->root = new Node();
    t->root->position = vec3(0.0, 0.0, 0.0);
    t->root->radius = t->calcTreeWidth();
    t->root->depth = 0;
    t->root->insert(sim->particles, cores);
    return t;
}

int64_t count_leaves(const Node *n) {
    if (!n) return 0;
    if (n->isLeaf) return n->particle ? 1 : 0;
    int64_t c = 0;
    for (int i = 0; i < 8; i++) c += count_leaves(n->children[i]);
    return c;
}

struct Dump {
    int64_t cap, k = 0, leafpos = 0;
    int32_t *level; int64_t *first, *count;
    double *mass, *com, *gas, *mvel;
    int64_t *perm;
    bool overflow = false;
    void visit(const Node *n) {
        if (!n) return;
        int64_t c = count_leaves(n);
        if (c == 0) return;
        if (k >= cap) { overflow = true; return; }
        level[k] = n->depth; first[k] = leafpos; count[k] = c;
        mass[k] = n->mass;
        com[3 * k] = n->centerOfMass.x; com[3 * k + 1] = n->centerOfMass.y; com[3 * k + 2] = n->centerOfMass.z;
        gas[k] = n->gasMass;
        mvel[3 * k] = n->mVel.x; mvel[3 * k + 1] = n->mVel.y; mvel[3 * k + 2] = n->mVel.z;
        k++;
        if (n->isLeaf) { perm[leafpos++] = n->particle->id; return; }
        for (int i = 0; i < 8; i++) visit(n->children[i]);
    }
};

}  // namespace

extern "C" {

/* Build + calcGasDensity + calculateForces on the given arrays (all active).
 * cores <= 0 uses Tree::buildTree (cores = omp_get_max_threads()). */
int ref_force_step(int64_t n, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                   double *m, double *U, double *mu, uint8_t *type,
                   double *ax, double *ay, double *az, double *h, double *rho, double *P,
                   double *T, double *dUdt, double theta, double e0, double mih, int cores,
                   double *R_out) {
    Arrays a{n, x, y, z, vx, vy, vz, m, U, mu, type, ax, ay, az, h, rho, P, T, dUdt};
    Simulation *sim = make_sim(a, theta, e0, mih);
    Tree *t = build(sim, cores);
    if (R_out) *R_out = t->root->radius;
    t->calcGasDensity();
    t->calculateForces();
    read_back(sim, a);
    delete t;
    free_sim(sim);
    return 0;
}

/* The same phases timed like Simulation::run's phase log (Simulation.cpp:
 * 275-345): sec[0] build tree, [1] SPH density, [2] forces, [3] delete tree. */
int ref_time_step(int64_t n, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                  double *m, double *U, double *mu, uint8_t *type, double theta, double e0, double mih,
                  double *sec) {
    std::vector<double> zero((size_t)n, 0.0), ax((size_t)n), ay((size_t)n), az((size_t)n), h((size_t)n),
        rho((size_t)n), P((size_t)n), T((size_t)n), du((size_t)n, 0.0);
    Arrays a{n, x, y, z, vx, vy, vz, m, U, mu, type, ax.data(), ay.data(), az.data(), h.data(), rho.data(),
             P.data(), T.data(), du.data()};
    Simulation *sim = make_sim(a, theta, e0, mih);
    double t0 = omp_get_wtime();
    Tree *t = build(sim, 0);
    double t1 = omp_get_wtime();
    t->calcGasDensity();
    double t2 = omp_get_wtime();
    t->calculateForces();
    double t3 = omp_get_wtime();
    delete t;
    double t4 = omp_get_wtime();
    sec[0] = t1 - t0; sec[1] = t2 - t1; sec[2] = t3 - t2; sec[3] = t4 - t3;
    free_sim(sim);
    return 0;
}

/* Pre-order dump of the reference octree's non-empty nodes. */
int64_t ref_tree_export(int64_t n, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                        double *m, double *U, double *mu, uint8_t *type, int cores, int64_t cap,
                        int32_t *level, int64_t *first, int64_t *count, double *mass, double *com,
                        double *gas, double *mvel, int64_t *perm, double *R_out) {
    std::vector<double> zero((size_t)n, 0.0);
    Arrays a{n, x, y, z, vx, vy, vz, m, U, mu, type, zero.data(), zero.data(), zero.data(),
             zero.data(), zero.data(), zero.data(), zero.data(), zero.data()};
    Simulation *sim = make_sim(a, 0.5, 1.0, 1.0);
    Tree *t = build(sim, cores);
    if (R_out) *R_out = t->root->radius;
    Dump d;
    d.cap = cap; d.level = level; d.first = first; d.count = count;
    d.mass = mass; d.com = com; d.gas = gas; d.mvel = mvel; d.perm = perm;
    d.visit(t->root);
    int64_t k = d.overflow ? -1 : d.k;
    delete t;
    free_sim(sim);
    return k;
}

/* Tree::buildTree then Tree::calcVisualDensity (Simulation.cpp:121-128).
 * radius <= 0 takes root->radius / 100000 (Simulation.cpp:126). */
int ref_visual_density(int64_t n, double *x, double *y, double *z, double *m, uint8_t *type, int cores,
                       double radius, double *vd, double *R_out) {
    std::vector<double> zero((size_t)n, 0.0), mu((size_t)n, 0.58);
    Arrays a{n, x, y, z, zero.data(), zero.data(), zero.data(), m, zero.data(), mu.data(), type,
             zero.data(), zero.data(), zero.data(), zero.data(), zero.data(), zero.data(), zero.data(),
             zero.data()};
    Simulation *sim = make_sim(a, 0.5, 1.0, 1.0);
    Tree *t = build(sim, cores);
    if (R_out) *R_out = t->root->radius;
    sim->visualDensityRadius = radius > 0 ? radius : t->root->radius / 100000;
    t->calcVisualDensity();
    for (int64_t i = 0; i < n; i++) vd[i] = sim->particles[(size_t)i]->visualDensity;
    delete t;
    free_sim(sim);
    return 0;
}

/* DataManager::saveData (DataManager.cpp:86-424) of the given particles into
 * dir (which must end in '/'), file name <timeStep><ending>. */
int ref_save_data(int64_t n, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                  double *m, double *U, double *T, double *P, uint8_t *type, uint8_t *galaxy_part,
                  uint32_t *id, double *vd, double *sfr, const char *dir, const char *format, int time_step,
                  int n_output, double delta_time, double end_time, double current_time) {
    std::vector<double> zero((size_t)n, 0.0), mu((size_t)n, 0.58);
    Arrays a{n, x, y, z, vx, vy, vz, m, U, mu.data(), type, zero.data(), zero.data(), zero.data(),
             zero.data(), zero.data(), P, T, zero.data()};
    Simulation *sim = make_sim(a, 0.5, 1.0, 1.0);
    for (int64_t i = 0; i < n; i++) {
        Particle *p = sim->particles[(size_t)i];
        p->galaxyPart = galaxy_part[i];
        p->id = id[i];
        p->visualDensity = vd[i];
        p->sfr = sfr[i];
    }
    DataManager dm(dir);
    dm.outputFormat = format;
    dm.saveData(sim->particles, time_step, 0, n_output, delta_time, end_time, current_time);
    free_sim(sim);
    return 0;
}

/* Cooling::coolingRoutine and SFR::sfrRoutine (Cooling.cpp:6-26, SFR.cpp:12-28)
 * on every gas particle, in that order, as Simulation::run would call them
 * (Simulation.cpp:306-317); rand() is seeded with srand_seed. */
int ref_gas_sources(int64_t n, double *T, double *rho, double *dt, uint8_t *type, double *U, double *dUdt,
                    int cooling, int star_formation, unsigned srand_seed, double *sfr) {
    Cooling cool;
    SFR sf;
    srand(srand_seed);
    for (int64_t i = 0; i < n; i++) {
        if (type[i] != 2) continue;
        Particle part;
        part.T = T[i]; part.rho = rho[i]; part.timeStep = dt[i]; part.type = type[i];
        part.U = U[i]; part.dUdt = dUdt[i]; part.sfr = 0.0;
        std::shared_ptr<Particle> sp(&part, [](Particle *) {});
        if (cooling) cool.coolingRoutine(sp);
        if (star_formation) sf.sfrRoutine(sp);
        type[i] = part.type; U[i] = part.U; dUdt[i] = part.dUdt; sfr[i] = part.sfr;
    }
    return 0;
}

/* Simulation::init's force phase and Simulation::run's step loop
 * (Simulation.cpp:121-139, :166-356) with the reference's Tree and
 * TimeIntegration, without I/O, shuffle or Hubble-free shortcuts. */
int ref_run(int64_t n, double *x, double *y, double *z, double *vx, double *vy, double *vz,
            double *m, double *U, double *mu, uint8_t *type,
            double *ax, double *ay, double *az, double *h, double *rho, double *P,
            double *T, double *dUdt, double theta, double e0, double mih, int cores,
            double eta, double min_dt, double max_dt, double end_time, double H0, int max_steps) {
    Arrays a{n, x, y, z, vx, vy, vz, m, U, mu, type, ax, ay, az, h, rho, P, T, dUdt};
    Simulation *sim = make_sim(a, theta, e0, mih);
    TimeIntegration ti;
    sim->eta = eta; sim->minTimeStep = min_dt; sim->maxTimeStep = max_dt; sim->endTime = end_time;
    sim->H0 = H0;
    std::vector<Particle *> &ps = sim->particles;
    {
        Tree *t = build(sim, cores);
        t->calcGasDensity();
        t->calculateForces();
        delete t;
    }
    double &gt = sim->globalTime;
    gt = 0.0;
    auto step_of = [&](Particle *p) {
        double am = p->acc.length();
        if (am > 0) {
            double ts = eta * std::sqrt(e0 / am);
            p->timeStep = std::clamp(ts, min_dt, max_dt);
            p->timeStep = std::max(std::pow(2, std::floor(std::log2(p->timeStep))), min_dt);
        } else {
            p->timeStep = min_dt;
        }
        p->nextIntegrationTime = gt + p->timeStep;
    };
    for (Particle *p : ps) p->nextIntegrationTime = 0.0;
    for (Particle *p : ps) step_of(p);
    int steps = 0;
    while (gt < end_time && steps < max_steps) {
        for (Particle *p : ps)
            if (gt >= p->nextIntegrationTime) step_of(p);
        double mn = std::numeric_limits<double>::max();
        for (Particle *p : ps) mn = std::min(mn, p->nextIntegrationTime);
        gt = mn;
        for (Particle *p : ps)
            if (gt == p->nextIntegrationTime) {
                ti.Kick(p, p->timeStep);
                ti.Drift(p, p->timeStep);
            }
        Tree *t = build(sim, cores);
        t->calcGasDensity();
        t->calculateForces();
        for (Particle *p : ps)
            if (gt == p->nextIntegrationTime) {
                if (p->type == 2) ti.Ueuler(p, p->timeStep);
                double H0SI = (H0 * 1.0e3) / 3.08567758149137e22;
                double sf = std::exp(H0SI * p->timeStep);
                p->position *= sf;
                ti.Kick(p, p->timeStep);
                p->nextIntegrationTime += p->timeStep;
            }
        delete t;
        steps++;
    }
    read_back(sim, a);
    free_sim(sim);
    return steps;
}

}  // extern "C"
