/*
 * ag2 -- MI355X (gfx950) force engine for AstroGen2, C ABI.
 *
 * This is the drop-in boundary at the reference's `class Tree` seam
 * (simulation/src/Physics/Tree/Tree.h:13-27).  The reference's driver calls
 *     new Tree(this); tree->buildTree(); tree->calcGasDensity();
 *     tree->calculateForces(); delete tree;
 * once per step (Simulation.cpp:121-139 in init, :276-285 and :345 in run)
 * and the tree reads and writes Particle fields through Simulation*.  Here the
 * context owns device-resident structure-of-arrays particle buffers, the tree
 * and all scratch memory; the caller owns host arrays and copies only at
 * upload / download.  Particle order is the caller's array order throughout
 * (the device keeps the sort permutation).
 *
 * Conventions: SI units (like the reference); every entry point returns 0 on
 * success or an AG2_ERR_* code and never throws; ag2_last_error() gives the
 * message.  One host thread per context; contexts are independent.
 */
#ifndef AG2_H
#define AG2_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    AG2_OK = 0,
    AG2_ERR_ARG = 1,       /* invalid argument                                   */
    AG2_ERR_HIP = 2,       /* HIP runtime error                                  */
    AG2_ERR_CAPACITY = 3,  /* n above the context capacity / tree overflow       */
    AG2_ERR_STATE = 4,     /* call order (e.g. forces before build)              */
    AG2_ERR_NODEV = 5      /* no HIP device                                      */
};

enum { AG2_FP64 = 0, AG2_FP32_SPLIT = 1 };

typedef struct ag2_params {
    double theta;            /* opening angle on the half-width: Simulation::theta,
                                Node.cpp:331-334                                   */
    double e0;               /* softening length, m: Simulation::e0              */
    double massInH;          /* SPH group target gas mass, kg: Simulation::massInH */
    double G;                /* Constants::G (Constants.h:7)                     */
    int64_t split_threshold; /* 100 * cores (Node.cpp:420); the reference uses
                                cores = omp_get_max_threads() (Tree.cpp:44-48)    */
    int32_t precision;       /* AG2_FP64 (reference arithmetic) or AG2_FP32_SPLIT */
    int32_t device;          /* HIP device ordinal                               */
    double eta;              /* block time-step accuracy (Simulation.cpp:199)     */
    double min_dt, max_dt;   /* Simulation::minTimeStep / maxTimeStep, s          */
    double H0;               /* km/s/Mpc (Simulation.cpp:330-332)                 */
    int32_t cooling;         /* 1: Cooling::coolingRoutine (Cooling.cpp:6-26) on every
                                active gas particle before its U update, where
                                Simulation::run has the call (Simulation.cpp:309-312,
                                commented out there); 0: off, as in the reference   */
    int32_t star_formation;  /* 1: SFR::sfrRoutine (SFR.cpp:12-28) at the same place
                                (Simulation.cpp:313-317); 0: off                    */
    uint64_t sfr_seed;       /* seed of the counter-based uniform draw that takes the
                                place of rand() / RAND_MAX in SFR.cpp:19            */
    int32_t graph_launch;    /* 1: the first half of each tree build (keys, sort,
                                state permutation, node count) is captured once as
                                three HIP graphs and replayed; 0 (default): every
                                kernel is launched on its own, which measured about
                                0.02 ms per step faster on the 1 M disk.  Results
                                are identical.                                     */
    int32_t timing_events;   /* HIP events recorded in each step: 0 (default) the phase
                                boundaries and the gravity walk kernel
                                (ag2_phase_times, ag2_walk_kernel_ms); 1 the walk
                                kernel only; 2 none.  Each event is a marker in the
                                stream that the device waits on.                    */
} ag2_params;

/* Caller-owned host arrays in the caller's particle order.  For upload a NULL
 * field takes its default (mu 0.58, everything else 0); for download a NULL
 * field is skipped.  Fields follow Particle.h:18-57. */
typedef struct ag2_particles {
    int64_t n;
    double *x, *y, *z;
    double *vx, *vy, *vz;
    double *m;
    double *U, *mu;
    uint8_t *type;           /* 1 star, 2 gas, 3 dark matter                    */
    double *ax, *ay, *az;
    double *h, *rho, *P, *T, *dUdt;
} ag2_particles;

typedef struct ag2_diag {
    double global_time;      /* Simulation::globalTime after the last step       */
    int64_t steps;           /* steps taken by ag2_step                          */
    double R;                /* root half-width of the last tree                 */
    int64_t n_in_tree;       /* particles inside the root cube                   */
    int64_t n_nodes;         /* non-empty nodes of the last tree                 */
    int64_t max_level;       /* deepest node level                               */
    int64_t interactions;    /* accepted gravity (particle, node) pairs, last forces */
    int64_t sph_pairs;       /* of which also got the SPH term                   */
    int64_t groups;          /* SPH density groups in effect                     */
    int64_t gap_particles;   /* particles in a node-boundary rounding gap        */
    int64_t nan_kicks;       /* kicks skipped because acc was NaN (cumulative)   */
    double kinetic, potential, thermal;  /* J; potential from ag2_energy()       */
    double momentum[3];      /* kg m / s                                         */
    int64_t walk_entries;    /* (node, lane mask) entries swept by the last walk */
    int64_t walk_tests;      /* node records evaluated by the last walk          */
    int64_t walk_mixed;      /* of which needed per-target opening tests         */
    int64_t interactions_total; /* interactions summed over all force phases since
                                the last upload (a caller can time many steps and
                                read the count once)                             */
    int64_t sph_overflow;    /* gas targets of the last forces with more SPH
                                candidates than their list holds (256): their SPH
                                term came from the overflow walk                 */
    int64_t fp64_fallback;   /* 1: the last forces of an AG2_FP32_SPLIT context ran
                                the FP64 sweep, because the softening length was
                                below 2^-24 of the root half-width (FP32 cannot
                                hold (e0 / R)^4 there)                          */
    int64_t near_pairs;      /* AG2_FP32_SPLIT: (target, node) pairs of the last
                                forces closer than the FP32 offsets resolve,
                                evaluated in FP64 instead                        */
    int64_t near_redo_groups; /* AG2_FP32_SPLIT: groups of 64 targets with more
                                such pairs than recorded (32 nodes), whose
                                gravity came from an FP64 walk instead           */
} ag2_diag;

/* Context lifetime.  capacity: largest particle count this context will hold. */
int ag2_create(const ag2_params *params, int64_t capacity, void **ctx);
int ag2_destroy(void *ctx);
const char *ag2_last_error(const void *ctx);
int ag2_set_params(void *ctx, const ag2_params *params);

/* Host <-> device copies (synchronous).  Upload sets every particle's next
 * integration time to 0, i.e. active for the next force phase. */
int ag2_upload(void *ctx, const ag2_particles *host);
int ag2_download(void *ctx, ag2_particles *host);

/* Tree::buildTree (Tree.cpp:24-55): root half-width (Tree::calcTreeWidth,
 * Tree.cpp:86-117), 42-level Morton keys, LSD radix sort, linear octree and
 * bottom-up moments. */
int ag2_build_tree(void *ctx);
/* Tree::calcGasDensity (Tree.cpp:119-150, Node.cpp:722-796). */
int ag2_gas_density(void *ctx);
/* Tree::calcVisualDensity (Tree.cpp:152-176, Node.cpp:832-877) on the last
 * tree: from each particle's leaf, climb while the parent's half-width is
 * closer to radius, then mass / half-width^3 of that node; 0 where the climb
 * reaches the root and for particles outside the root cube.  radius is
 * Simulation::visualDensityRadius in m; 0 takes the reference's choice, the
 * root half-width of the first tree after the last upload / 100000
 * (Simulation.cpp:126).  out: n values in caller order (NULL: compute only).
 * The reference runs it after every tree build (Simulation.cpp:128, :280);
 * the tree stays valid after ag2_step, so calling this before a snapshot
 * gives that step's values.  Synchronous. */
int ag2_visual_density(void *ctx, double radius, double *out);
/* Particle::sfr (Particle.h:52) in caller order: the conversion probability
 * SFR::sfrRoutine last wrote for the particle, 0 if never (star_formation on). */
int ag2_get_sfr(void *ctx, double *out);
/* Tree::calculateForces (Tree.cpp:57-83, Node.cpp:247-399, SPH Node.cpp:88-172)
 * for active particles (globalTime == nextIntegrationTime). */
int ag2_forces(void *ctx);
/* Simulation::init force phase (Simulation.cpp:121-139): build, density,
 * forces with every particle active. */
int ag2_init_forces(void *ctx);
/* One iteration of Simulation::run (Simulation.cpp:210-356) on the device:
 * block time steps, first kick + drift, tree phases, U update, Hubble scaling,
 * second kick.  Returns once the whole step is queued: the host waits midway
 * for the tree's node counts (they size the launches of the second half),
 * while the particle prefix sums keep the device busy.  ag2_sync waits for
 * the step to finish. */
int ag2_step(void *ctx);
int ag2_sync(void *ctx);

/* Diagnostics: kinetic, thermal, momentum and (if with_potential) the tree
 * potential -(G m/e0)(pi/2 - atan(r/e0)) summed over accepted nodes. */
int ag2_energy(void *ctx, int with_potential, ag2_diag *out);
int ag2_diagnostics(void *ctx, ag2_diag *out);

/* Per-phase device times of the last step, ms: [0] tree width + keys,
 * [1] sort, [2] topology + moments, [3] density, [4] forces, [5] kicks. */
int ag2_phase_times(void *ctx, double *ms6);
/* Device time of the last force phase's gravity walk kernel alone (the
 * dominant kernel of a step, inside phase [4]), ms. */
int ag2_walk_kernel_ms(void *ctx, double *ms);

/* FP64 direct summation over all particles with the same softened law, for
 * the listed targets (caller indices): accuracy diagnostic of the tree forces.
 * pot may be NULL.  Synchronous. */
int ag2_direct_sum(void *ctx, int64_t nt, const int64_t *targets, double *ax, double *ay, double *az,
                   double *pot);

/* ag2_direct_sum without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, code that forbids hipMalloc).  Same arguments, result and
 * kernel as ag2_direct_sum; its device staging (the target list and the
 * output columns) is a fixed per-device array of the library, sized for at
 * most AG2_DIRECT_NOALLOC_MAX targets (2.5 MiB, held until the process
 * ends); a larger nt returns AG2_ERR_CAPACITY.  Every context on one device
 * shares that array: calls that overlap on one device race (serialise them).
 * Synchronous. */
#define AG2_DIRECT_NOALLOC_MAX 65536
int ag2_direct_sum_noalloc(void *ctx, int64_t nt, const int64_t *targets, double *ax, double *ay, double *az,
                           double *pot);

/* Sorted order of the last tree: perm[k] = caller index of the k-th particle
 * in key order (n_in_tree entries), and the 42-level keys in that order. */
int ag2_get_order(void *ctx, int64_t *perm, uint64_t *key_hi, uint64_t *key_lo);
/* Non-empty nodes of the last tree in pre-order: level, first sorted particle,
 * particle count, mass, centre of mass (xyz interleaved), gas mass,
 * gas-weighted velocity (xyz).  cap: room in the arrays. */
int ag2_get_tree(void *ctx, int64_t cap, int32_t *level, int64_t *first, int64_t *count,
                 double *mass, double *com, double *gas, double *mvel, int64_t *n_nodes);

/* ------------------------------------------------------------------ serving
 * A pool holds many independent simulations (jobs) on one device and
 * advances them together: each ag2_pool_step is one round in which every
 * resident job takes one step of Simulation::run (Simulation.cpp:210-356),
 * the walks of all resident jobs of a precision form run as one batched
 * launch, and jobs that reach their step count leave.  Beside the round, a
 * host thread downloads the jobs that finished in the last round (they are
 * AG2_JOB_DONE when the round returns) and admits queued jobs into the slots
 * the round frees (first come first served, while the job limit and the
 * device-memory budget allow); those step from the next round.  A pool with
 * no resident job admits first and steps the admitted jobs at once.
 * The reference has no counterpart: it runs one simulation per process
 * (main.cpp:15-24).  A job's result equals that of the same steps on a
 * context of its own.  All pool calls are serialised by the pool's lock. */
enum { AG2_JOB_QUEUED = 0, AG2_JOB_RUNNING = 1, AG2_JOB_DONE = 2 };

typedef struct ag2_pool_params {
    int32_t device;          /* HIP device ordinal                                */
    int32_t max_jobs;        /* resident jobs at most (slots)                     */
    int64_t hbm_budget;      /* device bytes for job contexts; 0: 90 % of free    */
} ag2_pool_params;

typedef struct ag2_pool_stats {
    int64_t rounds, job_steps, particle_steps;
    int64_t queued, running, done;
    double p50_step_ms, p99_step_ms, max_step_ms;   /* per job-step: the interval from the
                                                       end of the job's previous step (or
                                                       the start of the round admitting
                                                       it) to the end of this step's
                                                       round, the round's admissions and
                                                       downloads included                */
    double busy_s;                                  /* wall time of all rounds          */
    int64_t hbm_used, hbm_budget;
    double mean_resident;                           /* jobs stepped per round, mean      */
    int64_t max_resident;                           /* and most                          */
} ag2_pool_stats;

/* Device bytes a context (or a pool job) of this capacity with n_gas gas
 * particles takes (each gas particle adds its SPH candidate list). */
int64_t ag2_context_bytes(int64_t capacity, int64_t n_gas);
int ag2_pool_create(const ag2_pool_params *params, void **pool);
int ag2_pool_destroy(void *pool);
const char *ag2_pool_last_error(const void *pool);
/* Queue a job: its parameters, its initial particles (copied) and the number
 * of steps to run.  *id identifies it in later calls. */
int ag2_job_submit(void *pool, const ag2_params *params, const ag2_particles *initial, int64_t steps,
                   int64_t *id);
/* One round; *n_stepped = jobs that took a step in it.  A job that takes its
 * last step in a round is AG2_JOB_DONE after the next call. */
int ag2_pool_step(void *pool, int32_t *n_stepped);
int ag2_job_status(void *pool, int64_t id, int32_t *state, int64_t *steps_done);
/* Final state of a finished job in its caller order; the job is then forgotten. */
int ag2_job_result(void *pool, int64_t id, ag2_particles *out);
int ag2_pool_get_stats(void *pool, ag2_pool_stats *out);

#ifdef __cplusplus
}
#endif
#endif
