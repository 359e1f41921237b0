// Internal declarations of the ag2 engine: context layout, device-side state,
// shared device helpers and the host-side launch wrappers of every kernel file.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/ag2.h"
#include "tree_core.h"

namespace ag2 {

// Levels: up to core::kDeepest = 126 separate distinct particles (six key
// words of 21 levels; most builds use two), and bucket members sit one below.
constexpr int kNearRecs = 32;   // near-pair records per group of 64 targets (split-precision walk)
constexpr int kWalkStackCap = 384;                 // pending child blocks per wave of the block walk
constexpr int kWalkGrowthMax = kWalkStackCap - 1;  // stack growth a tree may need below any point
constexpr int kSortTile = 2048;   // keys per radix-sort workgroup (256 threads x 8)
constexpr int kSortLoMax = 2;     // low bytes of word 0 the radix passes may leave to the run ranking
constexpr int kSortRunLong = 64;  // a longest run above this sorts every byte in the next build
constexpr int kStateD = 19;        // FP64 per-particle state arrays carried across steps
constexpr int kSphListLen = 256;   // SPH candidates recorded per gas target (gravity.hip kSphK)
constexpr size_t kCounterBinBytes = 8 * 256 * 64;   // counters x bins x one 64-byte line (gravity.hip)

// Reference constants (Constants.h:7-18, Units.h).
constexpr double kPI = 3.14159265358979323846;
constexpr double kGAMMA = 5.0 / 3.0;
constexpr double kKB = 1.38064852e-23;
constexpr double kPRTN = 1.6726219e-27;
constexpr double kKMS = 1.0e3;
constexpr double kMPC = 3.08567758149137e22;

using core::WalkNode;

// Sums of a node (tree_build.hip k_moments_up): mass, mass-weighted position,
// gas mass, gas-mass-weighted velocity.
struct alignas(64) NodeSums {
    double m, sx, sy, sz, g, gx, gy, gz;
};
static_assert(sizeof(WalkNode) == 64, "WalkNode must be 64 bytes");

// Scalars that live in device memory so that a step needs no host round trip.
struct DevState {
    double sum, sumsq, cut, R;
    unsigned long long Rbits;
    unsigned long long min_next_bits;
    double gtime;
    int32_t nin, n_int, n_nodes, err;
    int32_t max_level, final_count, n_sph;   // n_sph: active SPH targets of the last forces
    int32_t max_prefix_run;   // longest run of particles sharing the radix-sorted key bits (this build)
    unsigned long long interactions, sph_pairs, gap, nan_kicks;
    unsigned long long walk_entries, walk_tests;   // sweep entries and node evaluations of the block walk
    unsigned long long walk_mixed;                 // node evaluations needing per-lane opening tests
    unsigned long long interactions_total;         // interactions summed over every force phase since upload
    unsigned long long sph_overflow;               // SPH targets of the last forces whose candidate list overflowed
    unsigned long long near_pairs, near_redo;      // split-precision walk: FP64 near pairs, groups walked again
    int32_t tie_big;      // this build has a tie run longer than the ranking kernel takes
    int32_t tie_deep;     // this build has distinct adjacent particles agreeing on every key word
    int32_t walk_depth;   // blocks the walk's stack may grow by below any point: 7 (deepest level + 3), or
                          // in a deep tree the largest sum of (internal children - 1) along a path, + 8
    int32_t deep_m;       // particles in tie runs at the current word of the deep refinement
    int32_t walk_err;     // a walk's stack overflowed (never expected; kept until the next upload)
    int32_t any_bucket;   // this build has adjacent particles at one position (bucket nodes)
    double red[8];   // energy reductions
};

struct Ctx {
    ag2_params p{};
    int64_t cap = 0, n = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // Particle state in resident order: the key order of the last tree
    // build (caller order right after an upload).  rid[r] is the caller
    // index of resident slot r.  Each build permutes the whole state into the
    // new key order (double-buffered: the *_alt / alt buffers), so every
    // kernel after it reads and writes particles in key order.
    double *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *m = nullptr, *U = nullptr, *mu = nullptr;
    double *ax = nullptr, *ay = nullptr, *az = nullptr;
    double *h = nullptr, *rho = nullptr, *P = nullptr, *T = nullptr, *dUdt = nullptr;
    double *ts = nullptr, *next = nullptr, *pot = nullptr;
    uint8_t *type = nullptr, *active = nullptr, *gapl = nullptr;
    uint32_t *rid = nullptr;
    double *alt[kStateD] = {};
    uint8_t *type_alt = nullptr, *active_alt = nullptr;
    uint32_t *rid_alt = nullptr;
    double *stage = nullptr;     // n values: caller-order staging for download
    int32_t *inv = nullptr;      // caller index -> resident slot, built on demand
    double *vd = nullptr;        // visual density in key order, allocated on first use
    double *sfr = nullptr;       // Particle::sfr in caller order (written by the second kick)
    double vd_radius0 = 0.0;     // root half-width / 1e5 of the first tree after the last upload

    // sort: word 0 of the octant keys sorts in key / idx (key_alt, idx_alt
    // ping-pong) on its bits from 8 sort_lo up; the run ranking then leaves
    // word 0 in w0, word 1 in key_alt and the final permutation in idx_alt
    // (key2 holds word 1 in sorted order before it)
    uint64_t *key = nullptr, *key_alt = nullptr, *key2 = nullptr, *w0 = nullptr;
    uint32_t *idx = nullptr, *idx_alt = nullptr;
    // Low bytes of word 0 the radix passes leave out (0 or 2).  Particles
    // sharing the sorted bits are few and close together (a cell a few
    // levels above the deepest of word 0), and the run ranking orders them
    // by the whole key.  A build whose runs are long sorts every byte again.
    int sort_lo = kSortLoMax;
    // deep refinement (allocated on first use): key words 2..5, and the
    // compacted sort keys, values, positions and run starts
    uint64_t *kdeep[core::kKeyWords - 2] = {};
    uint64_t *sk = nullptr, *sk_alt = nullptr;
    uint32_t *sv = nullptr, *sv_alt = nullptr, *spos = nullptr, *srun = nullptr;
    int kw_used = 2;             // key words of the last build
    bool fp64_fallback = false;  // the last forces of a split-precision context ran the FP64 sweep
    uint32_t *hist = nullptr, *dtotal = nullptr;
    // topology over particles in key order
    int8_t *lcp = nullptr;
    int8_t *lk = nullptr;   // key levels particles i and i + 1 share (-1 after the last one inside)
    int32_t *cnt = nullptr, *base = nullptr, *bsum = nullptr;

    // nodes (pre-order)
    int64_t node_cap = 0;
    int32_t *n_first = nullptr, *n_count = nullptr, *n_skip = nullptr, *n_parent = nullptr;
    int32_t *n_pidx = nullptr;
    int8_t *n_level = nullptr;
    double *n_mass = nullptr;
    double *n_gas = nullptr, *n_gvx = nullptr, *n_gvy = nullptr, *n_gvz = nullptr;
    double *n_cx = nullptr, *n_cy = nullptr, *n_cz = nullptr;
    WalkNode *wnode = nullptr;   // pre-order records of the potential walk, built on demand
    bool wnode_valid = false;
    // child-contiguous layout for the block walk
    core::TreeNode *tnode = nullptr;
    core::SphNode *tsph = nullptr;   // SPH records in the walk layout (written when gas is present)
    int32_t *n_nchild = nullptr, *n_rank = nullptr, *n_cstart = nullptr, *n_bsum = nullptr;

    // SPH density: per node the first selecting caller index and the fired
    // flag; groups in effect listed 256 per workgroup (g_list) with the
    // per-workgroup counts in g_final
    int32_t *g_first = nullptr, *g_list = nullptr;
    uint8_t *g_fired = nullptr;
    int32_t *g_final = nullptr;
    double *g_rho = nullptr;

    // moments: each node's sums, built bottom-up (one arrival counter per
    // node; internal-children counts bound the walk's stack in deep trees)
    NodeSums *nsum = nullptr;
    int32_t *n_ichild = nullptr;
    NodeSums *nsum_tile = nullptr;   // sums of aligned 64- and 4096-particle blocks (node sums)
    // The reference sums a node's gas mass one particle at a time in array
    // order (Node.cpp:479, :681).  When every gas particle has the same mass m
    // that sum depends only on the count k: gas_table[k] = fl(gas_table[k-1] + m).
    double *gas_table = nullptr;
    int gas_equal = 0;
    bool sph_active = false;   // any gas particle uploaded: forces record and evaluate SPH candidates
    int64_t gas_count = 0;     // gas particles uploaded
    int64_t zero_mass = 0;     // particles uploaded with m == 0
    int64_t sph_cap = 0;       // targets the SPH list buffers hold
    int64_t sph_reserve = 0;   // the SPH lists are allocated for at least this many targets
    int32_t *gslot = nullptr, *gidx = nullptr, *gtiles = nullptr;   // target -> list slot, slot -> target
    int32_t *sph_list = nullptr, *sph_cnt = nullptr;                 // candidate lists (kSphListLen per slot)
    uint8_t *gover = nullptr;                                        // per group: a list overflowed
    // split-precision walk: (node, lane mask) pairs whose targets lie within
    // the FP32 offsets' resolution, evaluated in FP64 after the walk
    // (kNearRecs per group; a group with more is walked again in FP64)
    int32_t *near_cnt = nullptr, *near_node = nullptr;
    unsigned long long *near_mask = nullptr;

    // reductions
    double *partials = nullptr;
    // the particle buffers and the node buffers are each carved from one
    // device allocation (Carve)
    void *slab = nullptr, *node_slab = nullptr;
    DevState *ds = nullptr;
    unsigned long long *cbins = nullptr;   // walk counters spread over bins (gravity.hip bin_add)
    DevState hs{};  // host copy of the last read
    DevState *hs_pinned = nullptr;   // pinned landing buffer of the build's one device->host read
    bool own_stream = true;          // false inside a serving pool (the pool's stream)

    // bookkeeping
    bool tree_valid = false, started = false;
    // Set by callers whose density and force phases follow this build
    // (a step, the initial forces): the build's permutation then writes h = 0
    // for the gas (the density's first act, Tree.cpp:121-130; h_cleared
    // records that it did) and, with every particle active, leaves the
    // accelerations out, since the walk writes every one of them.
    bool density_follows = false, acc_overwritten = false, h_cleared = false;
    // the build cleared the nodes' group state (k_node_moments); the next
    // density phase uses it and clears it again if called once more
    bool groups_cleared = false;
    int64_t steps = 0;
    int nin = 0, n_int = 0, n_nodes = 0;
    hipEvent_t ev[10] = {};          // phase boundaries [0..7], the gravity walk kernel [8, 9]
    hipEvent_t ev_counts = nullptr;   // recorded when the node counts have reached hs_pinned
    // Captured launch sequences of the first build half (build_tree_a):
    // segment s (0 tree width and keys, 1 sort, 2 ordering, state permutation
    // and node count) for state parity q (the double-buffered particle state
    // swaps its buffers with every permutation).  Replayed while the particle
    // count, the node capacity and the parameters stay the same
    // (reset_graphs drops them).
    hipGraphExec_t seg_exec[3][2] = {};
    int parity = 0;
    int timing = 2;   // events: 2 phases and walk kernel, 1 walk kernel, 0 none (2 - timing_events)
    double phase_ms[6] = {0, 0, 0, 0, 0, 0};
};

// Scope of a build whose density and force phases follow
// (Ctx::density_follows, Ctx::acc_overwritten), cleared again on the way
// out, errors included.
struct StepBuild {
    Ctx &c;
    StepBuild(Ctx &ctx, bool all_active) : c(ctx) {
        c.density_follows = true;
        c.acc_overwritten = all_active;
    }
    ~StepBuild() { c.density_follows = c.acc_overwritten = false; }
    StepBuild(const StepBuild &) = delete;
    StepBuild &operator=(const StepBuild &) = delete;
};

constexpr int32_t kGroupNone = 0x7fffffff;   // no gas particle selects this node as its group

// ---------------------------------------------------------------- helpers
#define AG2_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            throw ::ag2::HipError(std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                          \
    } while (0)

struct HipError {
    std::string msg;
    explicit HipError(std::string m) : msg(std::move(m)) {}
};

// A build or walk that exceeds a stated bound (AG2_ERR_CAPACITY at the ABI).
struct CapacityError {
    std::string msg;
    explicit CapacityError(std::string m) : msg(std::move(m)) {}
};

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// Sub-buffers of one device allocation, each 256-byte aligned and followed by
// 4 KiB of slack (as much as a separate allocation's page rounding leaves at
// least): a pass with no base sums the sizes, a pass with the allocation's
// base hands out addresses.  One hipMalloc per buffer set instead of one per
// buffer (a context holds about a hundred).
struct Carve {
    char *base = nullptr;
    size_t off = 0;
    template <class T>
    void operator()(T **p, size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        if (base) *p = (T *)(base + off);
        off += bytes + 4096;
    }
};

__device__ inline int lane_id() { return (int)__lane_id(); }
// Workgroups are handed to the 8 XCDs round robin (block b on XCD b % 8), and
// each XCD has its own L2.  xcd_block renumbers them so that XCD x runs one
// contiguous range of the nb work items: neighbouring items share one L2
// (reads of the same lines, writes that fill the same lines).  A bijection on
// [0, nb).
__device__ inline int xcd_block(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}
__device__ inline unsigned long long lanes_below() {
    int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive wave scan of an int (64 lanes).
__device__ inline int wave_incl_scan(int v) {
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(v, d, 64);
        if (lane_id() >= d) v += t;
    }
    return v;
}

// Exclusive scan over a 256-thread workgroup; *total gets the sum.
// scratch: at least 4 ints of LDS.
__device__ inline int block256_excl_scan(int v, int *scratch, int *total) {
    int w = threadIdx.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane_id() == 63) scratch[w] = inc;
    __syncthreads();
    int off = 0, tot = 0;
    for (int k = 0; k < 4; k++) {
        int s = scratch[k];
        if (k < w) off += s;
        tot += s;
    }
    __syncthreads();
    if (total) *total = tot;
    return off + inc - v;
}

// Offset of tile `tile` in a device-wide exclusive scan, from the per-tile
// sums bsum[0 .. nb) (one workgroup of 256 reads them all: a few thousand
// values at most), and optionally the total.  The scan's apply kernel forms
// its own offset, so no kernel scans the tile sums in between.
__device__ inline int tile_offset(const int32_t *bsum, int nb, int tile, int *scratch, int *total) {
    int before = 0, all = 0;
    for (int k = threadIdx.x; k < nb; k += 256) {
        const int v = bsum[k];
        all += v;
        before += k < tile ? v : 0;
    }
    int b = 0, t = 0;
    block256_excl_scan(before, scratch, &b);
    block256_excl_scan(all, scratch, &t);
    if (total) *total = t;
    return b;
}

// Sum of a double over a 64-lane wave in a fixed butterfly order.
__device__ inline double wave_sum(double v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// --------------------------------------------------------- host wrappers
void build_tree(Ctx &c);                 // tree_build.hip: build_tree_a, stream sync, build_tree_b
void build_tree_a(Ctx &c);               // tree_build.hip: up to the node count (async)
// build_tree_a in three parts, for callers that sort several contexts in one
// batch: tree width and keys; sort_batch over contexts on one stream (the
// caller orders it after every context's keys); then ordering, state
// permutation and node count.
void build_keys(Ctx &c);
void build_order(Ctx &c);
size_t sort_job_bytes();
void sort_batch(Ctx *const *cs, int nc, hipStream_t st, void *scratch_dev, void *scratch_host);
void build_tree_b(Ctx &c);               // tree_build.hip: nodes, moments, layout (after a sync)
core::KeyWords key_words(const Ctx &c);  // tree_build.hip: the key words of the last build
void gas_density(Ctx &c);                // sph_density.hip
void visual_density(Ctx &c, double rvd, double *vd);   // visual_density.hip: key order into vd
void forces(Ctx &c, int mode);           // gravity.hip  (mode 0 forces, 1 potential)
void forces_fp32(Ctx &c, int mode);      // gravity.hip  (split-precision FP32 forces)
bool fp32_usable(const Ctx &c);          // gravity.hip: e0 / R of the last build suits the FP32 sweep
void sph_prepare(Ctx &c);                // gravity.hip: SPH list slots before a gravity walk
void sph_finish(Ctx &c);                 // gravity.hip: SPH candidate evaluation after it
void near_finish(Ctx &c);                // gravity.hip: FP64 near pairs of a split-precision walk
void fold_counters(Ctx &c);              // gravity.hip: walk counters' bins into the DevState
void kdk_begin(Ctx &c);                  // kdk.hip: time steps, min, first kick + drift
void kdk_end(Ctx &c);                    // kdk.hip: U update, Hubble, second kick
void kdk_init_steps(Ctx &c);
void energies(Ctx &c, int with_pot, double out[6]);
void fill_active(Ctx &c, uint8_t v);
void alloc_nodes(Ctx &c, int64_t node_cap);
void pack_walknodes(Ctx &c);             // tree_build.hip: pre-order records for the potential walk
void direct_sum(Ctx &c, int nt, const int64_t *d_targets, double *d_out, int with_pot);   // direct.hip
// direct.hip: the device address of the no-allocation scratch on `device`
// (current), looked up once per device; then the direct sum of host-listed
// targets through that scratch, with no device allocation (synchronous).
void direct_scratch_init(int device);
void direct_sum_noalloc(Ctx &c, int nt, const int64_t *h_targets, double *ax, double *ay, double *az, double *pot);
void state_fields(Ctx &c, double **f[kStateD]);   // tree_build.hip: the kStateD arrays, in a fixed order
void build_inverse(Ctx &c);              // tree_build.hip: inv[rid[r]] = r
void reset_graphs(Ctx &c);               // tree_build.hip: drop the captured build segments

// Contexts (api.hip).  ctx_new: stream == nullptr creates the context's own
// stream, otherwise the context runs on the given one (a serving pool's).
int check_params(const ag2_params *p, std::string &why);   // 0 or AG2_ERR_ARG with the reason
Ctx *ctx_new(const ag2_params &p, int64_t capacity, hipStream_t stream);
void ctx_delete(Ctx *c);
int64_t ctx_footprint(int64_t capacity, int64_t n_gas);   // device bytes of a context
void ctx_upload(Ctx &c, const ag2_particles &hp);
void ctx_download(Ctx &c, ag2_particles &hp);
// One batched gravity walk (plus SPH) over several contexts on one stream
// (gravity.hip); all contexts share the precision form given.
void forces_batch(Ctx *const *cs, int nc, hipStream_t stream, int f32, void *scratch_dev, void *scratch_host);
size_t job_walk_bytes();                 // gravity.hip: size of one job record of forces_batch

}  // namespace ag2
