/*
 * CPU oracle for the AstroGen2 force engine -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch CPU restatement of what the reference's per-step tree code
 * computes (Philip-Spaeth/AstroGenesis2.0: simulation/src/Physics/Tree/Node.cpp
 * and Tree.cpp, Math/kernel.cpp, Physics/TimeIntegration.cpp,
 * Physics/Simulation.cpp).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only to check results.  The product library never links or calls it.
 *
 * Units are SI (metres, kilograms, seconds), like the reference.
 * Every function states the reference lines it restates.
 *
 * Pinning (what checks each function against something independent):
 *   orc_tree_width        golden vectors from the compiled reference (R is part
 *                         of every tree fixture); closed form on tiny inputs.
 *   orc_force_step        golden vectors from the compiled reference (1 thread,
 *                         fixed order): accelerations, h, rho, P, T, dU/dt on
 *                         Plummer, gas-sphere and mixed-disk fixtures; theta=0
 *                         equals orc_direct_sum; two-body closed form.
 *   orc_tree_export       reference tree dump (pre-order node list) fixtures.
 *   orc_morton_keys       brute-force pointer descent (orc_tree_export order)
 *                         and per-axis dyadic comparisons in tests.
 *   orc_direct_sum        Newton's third law (sum m a = 0), two-body closed form,
 *                         Plummer analytic field in the large-N mean.
 *   orc_run               golden multi-step trajectories from the compiled
 *                         reference; time reversibility.
 *   orc_gas_sources       golden vectors of the compiled reference's
 *                         Cooling::coolingRoutine (dU/dt bit-exact) and
 *                         SFR::sfrRoutine (sfr and eligibility bit-exact; its
 *                         rand() draws are not reproducible, so conversions are
 *                         checked as a subset of the eligible particles and, for
 *                         orc_sfr_uniform, by their count against sum p);
 *                         closed forms of the free-free rate.
 *   orc_visual_density    golden vectors from the compiled reference's
 *                         Tree::calcVisualDensity; closed forms (one particle per
 *                         leaf: m / r_leaf^3; a climb to a level-k node:
 *                         M_k / (R/2^k)^3); radius 0 equals R/1e5 passed in.
 */
#ifndef AG2_ORACLE_H
#define AG2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    double theta;            /* opening angle on the half-width (Node.cpp:331-334) */
    double e0;               /* softening length, m (Simulation.h:52)              */
    double massInH;          /* target gas mass of an SPH group, kg (Simulation.h:55) */
    double G;                /* gravitational constant (Constants.h:7)             */
    int64_t split_threshold; /* 100*cores in the reference (Node.cpp:420)          */
    int32_t max_level;       /* deepest level at which distinct particles may still
                                be split apart; 126 in the engine (deeper: the
                                force step returns -2)                            */
    int32_t nthreads;        /* OpenMP threads for the per-particle walk (results
                                do not depend on it)                               */
} orc_params;

/* Structure-of-arrays particle state owned by the caller. */
typedef struct orc_particles {
    int64_t n;
    double *x, *y, *z;
    double *vx, *vy, *vz;
    double *m;
    double *U;       /* internal energy per mass, J/kg                             */
    double *mu;      /* mean molecular weight (Particle.h:41, default 0.58)        */
    uint8_t *type;   /* 1 star, 2 gas, 3 dark matter (Particle.h:32)               */
    double *ax, *ay, *az;
    double *h, *rho, *P, *T, *dUdt;
    uint8_t *active; /* NULL: every particle is active                             */
} orc_particles;

typedef struct orc_stats {
    int64_t n_nodes;          /* allocated nodes, including empty octants          */
    int64_t n_in_tree;        /* particles that ended in a leaf                    */
    int64_t interactions;     /* accepted (particle, node) gravity interactions    */
    int64_t sph_interactions; /* accepted pairs that also got the SPH term         */
    int64_t groups_fired;     /* SPH density groups assigned                       */
    double  R;                /* root half-width                                   */
} orc_stats;

/* Root half-width (Tree.cpp:86-117 / Tree::calcTreeWidth). */
double orc_tree_width(int64_t n, const double *x, const double *y, const double *z);

/* One force evaluation in the reference's phase order: build tree
 * (Tree::buildTree), SPH group density (Tree::calcGasDensity), forces
 * (Tree::calculateForces).  acc is zeroed for active particles; dUdt
 * accumulates (the reference never resets it inside the force phase).
 *
 * Depth: like the reference, a leaf splits until its particles separate
 * (Node.cpp:597-699), at any depth up to max_level; particles at one exact
 * position, which the reference never separates, form a bucket of member
 * leaves at the level where they become alone.  Returns -2 (after
 * computing) when distinct particles would need a split below max_level.
 * orc_tree_export returns -2 and orc_run -2 in the same case. */
int orc_force_step(orc_particles *p, const orc_params *prm, orc_stats *st);
/* The same, and scale[i] = the sum over the contributions added to particle
 * i's acceleration (gravity of each accepted node, SPH term of each accepted
 * gas node) of their magnitudes: two sums of the same terms in different
 * orders differ by a few units of 1e-16 * scale, whatever the cancellation
 * in the result.  Inactive particles keep their scale entry. */
int orc_force_step_scale(orc_particles *p, const orc_params *prm, orc_stats *st, double *scale);
/* The same on a tree whose node moments are replaced by the given ones (nn
 * non-empty nodes in the order of orc_tree_export: mass, centre of mass xyz
 * interleaved, gas mass, gas-mass-weighted velocity xyz interleaved) before
 * the density and force phases; -1 if the tree has another node count.  It
 * separates the force and density arithmetic from the rounding of the
 * moments, which follows the reference's summation order here and differs
 * in the last bits from a correctly rounded sum. */
int orc_force_step_moments(orc_particles *p, const orc_params *prm, orc_stats *st, double *scale, int64_t nn,
                           const double *mass, const double *com, const double *gas, const double *mvel);

/* Build the tree, then Tree::calcVisualDensity (Tree.cpp:152-176, Node.cpp:
 * 832-877): vd[i] = mass / radius^3 of the node where particle i's climb from
 * its leaf stops (0 if it reaches the root or i is outside the root cube).
 * radius <= 0 takes root half-width / 100000 (Simulation.cpp:126).  R_out
 * (may be NULL) receives the root half-width. */
int orc_visual_density(const orc_particles *p, const orc_params *prm, double radius, double *vd, double *R_out);

/* Build the pointer octree and export its non-empty nodes in pre-order
 * (children in octant order 0..7).  Per node: level, first leaf position in
 * DFS leaf order, particle count, mass, centre of mass, gas mass, gas-weighted
 * mean velocity.  perm[k] = particle index of the k-th leaf in DFS order.
 * Returns the node count, or -1 if cap is too small. */
int64_t orc_tree_export(const orc_particles *p, const orc_params *prm, int64_t cap,
                        int32_t *level, int64_t *first, int64_t *count,
                        double *mass, double *com, double *gas, double *mvel,
                        int64_t *perm, int64_t *n_in_tree);

/* Octant path of every particle from the root (centre 0, half-width R) down to
 * level 42: k1 holds levels 1..21, k2 levels 22..42, three bits per level in
 * the order x | y<<1 | z<<2, level 1 in the top bits.  Centres are formed
 * exactly like Node.cpp:436-440 and the bit is set iff x > centre
 * (Node.cpp:713-716).  gap[i] = first level whose bounds check
 * (Node.cpp:706-711) fails, 0 if none, -1 if outside the root. */
int orc_morton_keys(int64_t n, const double *x, const double *y, const double *z,
                    double R, uint64_t *k1, uint64_t *k2, int32_t *gap);

/* FP64 direct summation with the reference's softened law
 * a = G m d / ((r^2 + e0^2) r) and potential -(G m/e0)(pi/2 - atan(r/e0)),
 * for the listed targets (r == 0 pairs skipped). */
int orc_direct_sum(int64_t n, const double *x, const double *y, const double *z,
                   const double *m, int64_t nt, const int64_t *targets,
                   double e0, double G, int32_t nthreads,
                   double *ax, double *ay, double *az, double *pot);

typedef struct orc_run_params {
    double eta, min_dt, max_dt;   /* Config.ini keys (DataManager.cpp:1385-1402) */
    double end_time;
    double H0;                    /* km/s/Mpc                                    */
    int32_t max_steps;            /* safety stop                                 */
    int32_t cooling;              /* Cooling::coolingRoutine in the second kick  */
    int32_t star_formation;       /* SFR::sfrRoutine in the second kick          */
    uint64_t sfr_seed;            /* seed of orc_sfr_uniform                     */
    double *sfr;                  /* n values, Particle::sfr (NULL: not kept)    */
} orc_run_params;

/* Uniform in [0, 1) for (seed, particle index, step) that takes the place of
 * rand() / RAND_MAX in SFR.cpp:19: two rounds of the SplitMix64 output
 * function (Steele, Lea, Flood 2014), top 53 bits. */
double orc_sfr_uniform(uint64_t seed, uint32_t id, int64_t step);

/* The gas source terms of Simulation::run's second-kick loop
 * (Simulation.cpp:306-325) for every particle with active[i] (NULL: all):
 * Cooling::coolingRoutine (Cooling.cpp:6-26) if cooling, then
 * SFR::sfrRoutine (SFR.cpp:12-28) if star_formation with the draw
 * orc_sfr_uniform(seed, i, step), then TimeIntegration::Ueuler for particles
 * still gas when do_ueuler.  dt[i] is the particle's time step. */
void orc_gas_sources(int64_t n, const uint8_t *active, const double *dt, const double *T, const double *rho,
                     int32_t cooling, int32_t star_formation, uint64_t seed, int64_t step, int32_t do_ueuler,
                     uint8_t *type, double *U, double *dUdt, double *sfr);

/* Simulation::init force phase followed by Simulation::run's step loop
 * (Simulation.cpp:121-139, :166-356) on the caller's arrays.  Returns the
 * number of steps taken. */
int orc_run(orc_particles *p, const orc_params *prm, const orc_run_params *rp);

#ifdef __cplusplus
}
#endif
#endif
