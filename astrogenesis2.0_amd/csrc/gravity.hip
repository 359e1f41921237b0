// Barnes-Hut gravity walk with the embedded SPH force:
// Tree::calculateForces (Tree.cpp:57-83) -> Node::calculateGravityForce
// (Node.cpp:247-399) -> Node::calcSPHForce (Node.cpp:88-172).
//
// k_walk_block: one wavefront per group of 64 targets that are consecutive in
// key order (spatial neighbours).  The wave walks a child-contiguous copy of
// the tree together.  Each iteration pops up to eight sibling blocks (64
// nodes, one per lane) from an LDS stack of (first child, count, lane mask)
// and classifies every node against the group's bounding box:
//   - accepted by every lane of the block's mask -> one batch entry;
//   - opened by every lane of the mask -> its child block is pushed;
//   - mixed -> each target lane runs the reference's own opening test, and
//     ballots give the entry mask (lanes that accept) and the pushed mask
//     (lanes that open).
// Entries and pushes are appended with ballot + mbcnt, so every lane sees
// exactly the node set of its own recursive walk in the reference
// (Node.cpp:382-396 visits the children of an opened node; the visiting order
// differs, which only reorders the sum).  Entries are swept 64 at a time:
// lane j stages entry j in LDS and every lane computes its interaction with
// every staged entry it belongs to.
// The opening test decides radius^2 against theta^2 r^2 with a 1e-10 margin
// and falls back to the reference's correctly rounded radius / sqrt(r^2)
// inside the margin, so the decision equals the reference's bit for bit.
//
// SPH (the term the reference adds at every accepted gas node within 2 h of
// a gas target): the gravity walk (KIND 1) records, per active gas target,
// the accepted gas nodes within a loose margin of 2 h in a list of kSphK
// node records; k_sph_pairs then evaluates the lists, eight lanes per
// target, with the exact r < 2 h test and the reference's self skip.  A
// group whose list filled sets a flag, and the overflow walk (KIND 2) walks
// only those groups again with the SPH term inline.  Gas-free uploads use
// the KIND 0 kernels, which carry no SPH code.
//
// FP64 sweep: the reference's softened law G m / (r^2 + e0^2) d / r.
// FP32 sweep (split precision): node centres and targets become FP32 offsets
// from the group's bounding-box centre in units of the root half-width
// (FP64 subtraction first), masses are scaled by the root mass, the force
// is m' / sqrt(r'^2 (r'^2 + e'^2)^2) with one v_rsq_f32, each 64-entry chunk
// is summed plainly and added to the running total with Kahan compensation.
// The SPH term is FP64 in both forms.
//
// k_walk64 (one lane per target, stackless) is kept for the potential.
#include "ag2_internal.h"

#include <type_traits>

namespace ag2 {
namespace {

constexpr int kWaves = 4;
constexpr int kStackCap = kWalkStackCap;   // pending child blocks per wave
constexpr int kEnt = 128;        // batch entries: < 64 left after a sweep + <= 64 added per iteration
constexpr int kSphK = kSphListLen;   // recorded SPH candidates per gas target (more: the overflow walk)

// A pending block of siblings: nodes first .. first+count-1, opened by the
// lanes in mask.
struct alignas(16) StackEnt {
    int32_t first, count;
    unsigned long long mask;
};

typedef unsigned u32_alias __attribute__((may_alias));

struct BlockLds {
    StackEnt stk[kStackCap];
    int32_t ent_node[kEnt];
    unsigned long long ent_mask[kEnt];
    union {
        struct { double cx[64], cy[64], cz[64], m[64]; } d;   // FP64 sweep
        struct { float4 f[64], lo[64]; } s;                   // FP32 sweep: offsets, their remainders
        double acc[4][64];                                    // SPH sums per target lane; mixed-node staging
    } e;
    uint16_t pairs[128];   // pending SPH pairs: target lane | entry << 8
};

struct WalkArgs {
    int n, nin, nn;
    const core::TreeNode *__restrict__ tree;
    const core::SphNode *__restrict__ tsph;
    // SPH candidate lists: gslot[t] = list slot of key-order target t (-1:
    // none), sph_list[slot * kSphK + k], sph_cnt[slot], gover[group] = 1 if a
    // list of the group overflowed
    const int32_t *__restrict__ gslot;
    int32_t *sph_list, *sph_cnt;
    uint8_t *gover;
    const double *__restrict__ sx, *__restrict__ sy, *__restrict__ sz, *__restrict__ sm;
    const uint8_t *__restrict__ stype;
    const double *__restrict__ svx, *__restrict__ svy, *__restrict__ svz;
    const double *__restrict__ mvx, *__restrict__ mvy, *__restrict__ mvz;
    const double *__restrict__ h, *__restrict__ rho, *__restrict__ P;
    const uint8_t *__restrict__ active;
    const DevState *ds;
    unsigned long long *cbins;   // spread counters (bin_add), folded into the DevState after the forces
    core::WalkParams wp;
    double *ax, *ay, *az, *dUdt;
    DevState *dsw;
    // split-precision walk: near pairs per group (count, then kNearRecs
    // records of node and lane mask), evaluated in FP64 by k_near_pass
    int32_t *near_cnt, *near_node;
    unsigned long long *near_mask;
};

// FP64 1/sqrt(x) and 1/x: the hardware estimates (v_rsq_f64, v_rcp_f64; at
// most 5.3e-8 and 4.5e-8 relative error over 1e-60 .. 1e60, measured on
// gfx950) and one Newton step, which leaves at most 4.2e-15 and 2.2e-15 --
// four orders of magnitude inside the 1e-10 parity bound of the force sums.
__device__ inline double rsq64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double hx = 0.5 * x;
    y = y * fma(-hx * y, y, 1.5);
    return y;
}

__device__ inline double rcp64(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = fma(y, fma(-x, y, 1.0), y);
    return y;
}

__device__ inline int rl(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ inline float rl(float v, int j) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j)); }
__device__ inline double uni(double v) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ inline double rl(double v, int j) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), j);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline unsigned long long rl(unsigned long long v, int j) {
    int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffull), j);
    int hi = __builtin_amdgcn_readlane((int)(v >> 32), j);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
// v_writelane_b32 (the compiler's own intrinsic; clang has no builtin for
// it): lane j of the result is v (uniform), the other lanes keep old.  One
// instruction per dword where a select on lane == j costs a compare, moves
// and selects.
extern "C" __device__ int ag2_writelane_i32(int v, int j, int old) __asm("llvm.amdgcn.writelane.i32");
__device__ inline unsigned long long wl(unsigned long long old, unsigned long long v, int j) {
    const unsigned lo = (unsigned)ag2_writelane_i32((int)(v & 0xffffffffull), j, (int)(old & 0xffffffffull));
    const unsigned hi = (unsigned)ag2_writelane_i32((int)(v >> 32), j, (int)(old >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
__device__ inline int shfl_i(int v, int src) { return __shfl(v, src, 64); }
__device__ inline unsigned long long shfl_u64(unsigned long long v, int src) {
    unsigned lo = (unsigned)__shfl((int)(v & 0xffffffffull), src, 64);
    unsigned hi = (unsigned)__shfl((int)(v >> 32), src, 64);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ inline double shfl_d(double v, int src) {
    long long b = __double_as_longlong(v);
    unsigned lo = (unsigned)__shfl((int)(b & 0xffffffffll), src, 64);
    unsigned hi = (unsigned)__shfl((int)(b >> 32), src, 64);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Counters that every wave adds to go through kCountBins bins, each on its
// own 64-byte line, picked by the wave's group: a single address would
// serialise tens of thousands of atomics.  k_fold_counters sums the bins.
enum { kCntInteractions = 0, kCntSphPairs, kCntEntries, kCntTests, kCntMixed, kCntSphOver, kCntNear, kCntNearRedo,
       kCounters };
constexpr int kCountBins = 256;
static_assert((size_t)kCounters * kCountBins * 64 <= kCounterBinBytes, "counter bins exceed their buffer");
__device__ inline void bin_add(unsigned long long *bins, int counter, int bin, unsigned long long v) {
    if (v) atomicAdd(&bins[((size_t)counter * kCountBins + (bin & (kCountBins - 1))) * 8], v);
}

// Number of set bits of b below this lane.
__device__ inline int mbcnt(unsigned long long b) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
}
// Lane-mask -> per-lane bool with no vector work (the mask is the condition).
__device__ inline bool in_mask(unsigned long long m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
__device__ inline void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The SPH term with the target's constants hoisted (core::sph_term_hoisted:
// the reference's arithmetic bit for bit).
__device__ inline void sph_pair(double dx, double dy, double dz, double r2, double gas, double vx, double vy,
                                double vz, const core::SphConst &k, core::WalkSums *s) {
    core::sph_term_hoisted(dx, dy, dz, r2, gas, vx, vy, vz, k, s);
}

// xcd_block (ag2_internal.h) keeps neighbouring groups, which read mostly
// the same tree records, on the same L2.


// Group g of 64 targets: particles 64 g .. 64 g + 63 in key order.
// KIND 0: gravity.  KIND 1: gravity, and every accepted gas node within a
// margin of 2h of a gas target goes on that target's SPH candidate list
// (evaluated by k_sph_pairs).  KIND 3: FP64 gravity again for the groups whose
// split-precision walk found more near pairs than kNearRecs (their results
// replace that walk's).  KIND 2: the SPH term, inline, for the targets
// whose list overflowed (groups flagged in gover only; no gravity).
template <bool F32, int KIND>
__device__ __forceinline__ void walk_block(const WalkArgs &a, int g) {
    constexpr bool GAS = KIND == 1 || KIND == 2;
    if (KIND == 2 && (g >= (a.n + 63) / 64 || !a.gover[g])) return;
    if (KIND == 3 && (g >= (a.n + 63) / 64 || a.near_cnt[g] <= kNearRecs)) return;
    __shared__ BlockLds lds[kWaves];
    const int wv = threadIdx.x >> 6;
    const int lane = lane_id();
    BlockLds &S = lds[wv];
    const int t = g * 64 + lane;
    const bool valid_t = t < a.n;
    const int q = t;   // particle state is in key order
    const bool act = valid_t && a.active[q] != 0;
    double px = 0.0, py = 0.0, pz = 0.0, pm = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    if (valid_t) {
        px = a.sx[t]; py = a.sy[t]; pz = a.sz[t]; pm = a.sm[t];
        vx = a.svx[t]; vy = a.svy[t]; vz = a.svz[t];
    }
    core::Target tg;
    tg.x = px; tg.y = py; tg.z = pz; tg.m = pm; tg.vx = vx; tg.vy = vy; tg.vz = vz;
    tg.gas = valid_t && a.stype[t] == 2;
    tg.h = act ? a.h[q] : 0.0;
    tg.rho = act ? a.rho[q] : 0.0;
    tg.P = act ? a.P[q] : 0.0;
    tg.self = t;
    const unsigned long long amask = __ballot(act && pm != 0.0);
    bool sph_lane = GAS && ((amask >> lane) & 1ull) && tg.gas && tg.h > 0.0;
    int slot = -1;
    if (GAS && sph_lane) {
        slot = a.gslot[t];
        if (KIND == 2) sph_lane = a.sph_cnt[slot] > kSphK;
    }
    const unsigned long long sphB = __ballot(sph_lane);
    const double h2x4 = GAS ? (2.0 * tg.h) * (2.0 * tg.h) : 0.0;   // SPH support (2h)^2
    int lcnt = 0;   // KIND 1: candidates recorded for this target

    core::WalkSums s{0, 0, 0, 0, 0, 0, 0};
    unsigned long long nint = 0ull;   // uniform: interactions of the whole wave
    double invL = 0.0, invMu = 0.0;
    float xr = 0.f, yr = 0.f, zr = 0.f, e2f = 0.f;
    float xlo = 0.f, ylo = 0.f, zlo = 0.f;   // FP32 remainders of the target offsets (near chunks)
    float r2df = 0.f;                         // near chunks: r^2 below which the pair goes to the FP64 pass
    // FP32 form: entries closer than r2min (scaled units) to this lane's
    // target are too close for FP32 offsets to resolve their direction; they
    // are recorded and evaluated in FP64 after the walk (k_near_pass)
    float r2min = 0.f;
    double near_a = 0.0;    // uniform: see the near-entry test at staging
    unsigned lnint = 0u;    // interactions of the fast chunks, counted per staging lane
    int nrec = 0;           // uniform: near pairs recorded for k_near_pass
    bool dfmode = false;    // uniform: this group's sweep carries the offsets' remainders
    float sumx = 0.f, sumy = 0.f, sumz = 0.f, cmpx = 0.f, cmpy = 0.f, cmpz = 0.f;
    const double theta = a.wp.theta, th2 = a.wp.theta * a.wp.theta;
    const double th2lo = th2 * (1.0 - 1e-10), th2hi = th2 * (1.0 + 1e-10);   // group-box accept / open margins
    const double inv_th2m = 1.0 / (th2 * (1.0 - 1e-10));         // mixed-node accept threshold factor
    const double open_ratio = (1.0 - 1e-10) / (1.0 + 1e-10);     // open threshold / accept threshold
    const double e02 = a.wp.e0 * a.wp.e0;
    const double R = __longlong_as_double((long long)a.ds->Rbits);

    if (amask != 0ull && a.nn > 0) {
        // bounding box of the group's active targets (decides whole-group
        // accept / open for most nodes without per-lane tests)
        double lo[3] = {px, py, pz}, hi[3] = {px, py, pz};
        if (!((amask >> lane) & 1ull)) { lo[0] = lo[1] = lo[2] = 1e308; hi[0] = hi[1] = hi[2] = -1e308; }
        // box of the SPH targets and their largest support 2h
        double glo[3] = {px, py, pz}, ghi[3] = {px, py, pz}, h2 = sph_lane ? 2.0 * tg.h : 0.0;
        if (!sph_lane) { glo[0] = glo[1] = glo[2] = 1e308; ghi[0] = ghi[1] = ghi[2] = -1e308; }
        for (int d = 32; d >= 1; d >>= 1) {
            for (int k = 0; k < 3; k++) {
                double a1 = __shfl_xor(lo[k], d, 64), b1 = __shfl_xor(hi[k], d, 64);
                lo[k] = a1 < lo[k] ? a1 : lo[k];
                hi[k] = b1 > hi[k] ? b1 : hi[k];
            }
            if (sphB) {
                for (int k = 0; k < 3; k++) {
                    double a1 = __shfl_xor(glo[k], d, 64), b1 = __shfl_xor(ghi[k], d, 64);
                    glo[k] = a1 < glo[k] ? a1 : glo[k];
                    ghi[k] = b1 > ghi[k] ? b1 : ghi[k];
                }
                double c1 = __shfl_xor(h2, d, 64);
                h2 = c1 > h2 ? c1 : h2;
            }
        }
        for (int k = 0; k < 3; k++) {
            lo[k] = uni(lo[k]); hi[k] = uni(hi[k]); glo[k] = uni(glo[k]); ghi[k] = uni(ghi[k]);
        }
        const double reach2 = uni(h2 * h2) * (1.0 + 1e-9);
        const double gcx = 0.5 * (lo[0] + hi[0]), gcy = 0.5 * (lo[1] + hi[1]), gcz = 0.5 * (lo[2] + hi[2]);
        if (F32) {
            double Mu = a.tree[0].mass;
            invL = uni(R > 0 ? 1.0 / R : 1.0);
            invMu = uni(Mu > 0 ? 1.0 / Mu : 1.0);
            const double oxd = (px - gcx) * invL, oyd = (py - gcy) * invL, ozd = (pz - gcz) * invL;
            xr = (float)oxd; yr = (float)oyd; zr = (float)ozd;
            xlo = (float)(oxd - (double)xr); ylo = (float)(oyd - (double)yr); zlo = (float)(ozd - (double)zr);
            double el = a.wp.e0 * invL;
            e2f = uni((float)(el * el));
            // each offset is rounded to 2^-24 of itself, so d carries an error
            // of about 2^-24 (|target| + |entry|) per component; below 2^-10
            // of 2 (|x| + |y| + |z|) of the target offset the FP32 direction
            // could be off by more than 2^-13, and FP32 may round d to zero
            // where the FP64 positions differ (the reference's r == 0 skip is
            // on FP64 r)
            const float ao = 2.f * (fabsf(xr) + fabsf(yr) + fabsf(zr)) * 0x1p-10f;
            r2min = ao * ao;
            // in a chunk with such an entry the sweep forms d from the offsets
            // and their remainders (48 bits: exact difference of the high parts
            // for close pairs by Sterbenz), which resolves the direction down
            // to 2^-34 of the offsets; closer pairs go to the FP64 pass
            const float adf = ao * 0x1p-24f;
            r2df = adf * adf;
            // an entry farther from the targets' box than near_a + near_b
            // (its offset sum) is beyond every lane's r2min with room for the
            // FP32 offsets' rounding (2^-23 of target and entry offsets), so
            // its lanes need no per-lane test
            float aomax = (amask >> lane) & 1ull ? ao : 0.f;
            for (int d = 32; d >= 1; d >>= 1) aomax = fmaxf(aomax, __shfl_xor(aomax, d, 64));
            const double ext = 0.5 * ((hi[0] - lo[0]) + (hi[1] - lo[1]) + (hi[2] - lo[2])) * invL;
            near_a = uni(4.0 * (double)aomax + 0x1p-20 * ext);
        }
        // KIND 1: the candidate test on the sweep's own r^2, loose by a
        // margin; k_sph_pairs decides exactly.  In the split form r is
        // computed from FP32 offsets from the group centre, each rounded by at
        // most 2^-24 of its size: for a node within 2h (scaled: D) of a target
        // with offset sum A, the computed r exceeds the true one by less than
        // 2^-23 (A + D) + rounding of r^2 itself.
        const double lim64 = h2x4 * (1.0 + 1e-9);
        float lim32 = 0.f;
        if (F32 && KIND == 1) {
            const double D = 2.0 * tg.h * invL;
            const double A = fabs((double)xr) + fabs((double)yr) + fabs((double)zr);
            const double rl32 = D * (1.0 + 1e-6) + 2.4e-7 * (A + 2.0 * D);
            lim32 = (float)(rl32 * rl32 * (1.0 + 1e-5));
        }
        int32_t *lbase = (KIND == 1 && slot >= 0) ? a.sph_list + (size_t)slot * kSphK : nullptr;
        // Multi-block pops keep the stack within kStackCap: a pop of up to
        // 64 nodes pushes at most 64 blocks, and above `single_above` the walk
        // pops one block at a time; from any point the stack then grows by at
        // most walk_depth blocks (7 per level below, or in a deep tree the
        // sum of (internal children - 1) along a path -- leaves and bucket
        // members push nothing).  The build refuses trees where that bound
        // exceeds kStackCap - 1.
        const int growth = __builtin_amdgcn_readfirstlane(a.ds->walk_depth);
        const int single_above = kStackCap - 64 - growth;
        if (lane == 0) { S.stk[0].first = 0; S.stk[0].count = 1; S.stk[0].mask = amask; }
        int sp = 1;
        int ni = 0;
        unsigned long long n_entries = 0ull, n_tests = 0ull, n_mixed = 0ull;
        while (true) {
            wave_sync_lds();
            if (sp > 0) {
                // ------------------------------------------- pop blocks
                int k = sp > single_above ? 1 : (sp < 8 ? sp : 8);
                StackEnt se{0, 0, 0ull};
                if (lane < k) se = S.stk[sp - 1 - lane];
                // (lanes at or past k hold count 0.)  The ballots of the pop
                // are of plain compares, combined on the scalar unit: a ballot
                // of a bool formed across a divergent branch costs a select
                // and a compare to rebuild its mask.
                const unsigned long long big = __ballot(se.count > 8);
                if (big) k = __builtin_ctzll(big) + 1;   // a bucket block is taken 8 at a time
                const int b = lane >> 3, sl = lane & 7;
                const int fb = shfl_i(se.first, b), cb = shfl_i(se.count, b);
                const unsigned long long Mb = shfl_u64(se.mask, b);
                sp -= k;
                if (big) {
                    if (lane == k - 1) { S.stk[sp].first = se.first + 8; S.stk[sp].count = se.count - 8; S.stk[sp].mask = se.mask; }
                    sp++;
                }
                const unsigned long long validB = __ballot(b < k) & __ballot(sl < cb);
                const int node = fb + sl;
                // lane classifies its node against the group box: 1 accept
                // for every lane, 2 open for every lane, 3 decide per lane,
                // 0 nothing (no node or mass 0)
                core::TreeNode r;
                int cls = 0, rnch = 0, rpidx = -1;
                double rad = 0.0;
                if (in_mask(validB)) {
                    r = a.tree[node];
                    rnch = r.nchild;
                    rpidx = r.pidx;
                    rad = ldexp(R, -(int)r.level);
                    if (r.mass == 0.0) cls = 0;
                    else if (r.pidx >= 0) cls = 1;
                    else {
                        // nearest and farthest box points from the centre;
                        // the 1e-10 margins absorb the fused roundings
                        const double ax = lo[0] - r.cx, bx = r.cx - hi[0];
                        const double ay = lo[1] - r.cy, by = r.cy - hi[1];
                        const double az = lo[2] - r.cz, bz = r.cz - hi[2];
                        const double ex = fmax(fmax(ax, bx), 0.0), ey = fmax(fmax(ay, by), 0.0), ez = fmax(fmax(az, bz), 0.0);
                        const double rmin2 = __builtin_fma(ex, ex, __builtin_fma(ey, ey, ez * ez));
                        const double fx = fmax(fabs(ax), fabs(bx)), fy = fmax(fabs(ay), fabs(by)), fz = fmax(fabs(az), fabs(bz));
                        const double rmax2 = __builtin_fma(fx, fx, __builtin_fma(fy, fy, fz * fz));
                        const double rad2 = rad * rad;
                        if (rmin2 > 0.0 && rad2 < th2lo * rmin2) cls = 1;
                        else if (rmin2 > 0.0 && rad2 > th2hi * rmax2) cls = 2;
                        else cls = 3;
                    }
                }
                n_tests += (unsigned long long)__builtin_popcountll(validB);
                // whole-mask accepts -> entries (a leaf never acts on itself)
                const unsigned long long c1B = __ballot(cls == 1);
                const int self = rpidx - g * 64;
                const bool drop = rpidx >= 0 && (unsigned)self < 64u;   // (its own leaf: a leaf is class 1 or 0)
                const unsigned long long A = drop ? Mb & ~(1ull << (self & 63)) : Mb;
                const unsigned long long appB = c1B & __ballot(A != 0ull);
                if (in_mask(appB)) {
                    const int slot = ni + mbcnt(appB);
                    S.ent_node[slot] = node;
                    S.ent_mask[slot] = A;
                }
                ni += __builtin_popcountll(appB);
                // whole-mask opens -> child blocks
                const unsigned long long nchB = __ballot(rnch > 0);
                const unsigned long long pB = __ballot(cls == 2) & nchB;
                if (in_mask(pB)) {
                    const int slot = sp + mbcnt(pB);
                    if (slot < kStackCap) {
                        S.stk[slot].first = r.first_child; S.stk[slot].count = r.nchild; S.stk[slot].mask = Mb;
                    } else {
                        a.dsw->walk_err = 1;   // unreachable: see single_above
                    }
                }
                sp += __builtin_popcountll(pB);
                if (sp > kStackCap) sp = kStackCap;
                // per-lane decisions, exactly the reference's test
                const unsigned long long mixB = __ballot(cls == 3);
                n_mixed += (unsigned long long)__builtin_popcountll(mixB);
                // The mixed nodes' centres and thresholds go to the staging
                // area (free outside the sweep) and are read back at a
                // uniform address (KIND 2 broadcasts them lane by lane).
                // A lane accepts a mixed node when r^2 > rad^2 / (theta^2 (1 - 1e-10))
                // and opens it when r^2 < rad^2 / (theta^2 (1 + 1e-10)): outside
                // that band the reference's rounded radius / sqrt(r^2) < theta
                // cannot decide otherwise; inside it, its own expression
                // decides on its own unfused r^2.  Each node's accept and
                // open masks are selected into the lane that holds the node,
                // and its entry and child block are appended after the loop
                // in lane order (the order the nodes were decided in).
                const double tacc = rad * rad * inv_th2m;
                if (KIND != 2 && cls == 3) {
                    const int slot = mbcnt(mixB);
                    S.e.acc[0][slot] = r.cx; S.e.acc[1][slot] = r.cy; S.e.acc[2][slot] = r.cz; S.e.acc[3][slot] = tacc;
                }
                if (KIND != 2 && mixB) wave_sync_lds();
                unsigned long long mAA = 0ull, mOO = 0ull;   // decisions on this lane's mixed node
                int mslot = 0;
                for (unsigned long long bb = mixB; bb; bb &= bb - 1, mslot++) {
                    const int kk = __builtin_ctzll(bb);
                    double ncx, ncy, ncz, Tacc;
                    if (KIND == 2) {
                        ncx = rl(r.cx, kk); ncy = rl(r.cy, kk); ncz = rl(r.cz, kk); Tacc = rl(tacc, kk);
                    } else {
                        ncx = S.e.acc[0][mslot]; ncy = S.e.acc[1][mslot]; ncz = S.e.acc[2][mslot]; Tacc = S.e.acc[3][mslot];
                    }
                    const double Topn = Tacc * open_ratio;
                    const unsigned long long Mk = rl(Mb, kk);
                    const double dx = ncx - px, dy = ncy - py, dz = ncz - pz;
                    const double r2f = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
                    const unsigned long long live = Mk & __ballot(r2f != 0.0);
                    unsigned long long AA = live & __ballot(r2f > Tacc);
                    unsigned long long OO = live & __ballot(r2f < Topn);
                    const unsigned long long band = live & ~(AA | OO);
                    if (band) {
                        const double nrad = rl(rad, kk);
                        bool a2 = false;
                        if (in_mask(band)) {
                            const double r2 = dx * dx + dy * dy + dz * dz;
                            a2 = (nrad / sqrt(r2)) < theta;
                        }
                        const unsigned long long BA = band & __ballot(a2);
                        AA |= BA;
                        OO |= band & ~BA;
                    }
                    mAA = wl(mAA, AA, kk);
                    mOO = wl(mOO, OO, kk);
                }
                if (mixB) {
                    const unsigned long long eB = __ballot(mAA != 0ull);
                    if (in_mask(eB)) {
                        const int slot = ni + mbcnt(eB);
                        S.ent_node[slot] = node;
                        S.ent_mask[slot] = mAA;
                    }
                    ni += __builtin_popcountll(eB);
                    const unsigned long long qB = __ballot(mOO != 0ull) & nchB;
                    if (in_mask(qB)) {
                        const int slot = sp + mbcnt(qB);
                        if (slot < kStackCap) {
                            S.stk[slot].first = r.first_child; S.stk[slot].count = r.nchild; S.stk[slot].mask = mOO;
                        } else {
                            a.dsw->walk_err = 1;
                        }
                    }
                    sp += __builtin_popcountll(qB);
                    if (sp > kStackCap) sp = kStackCap;
                }
            }
            // ------------------------------------------------- sweep
            while (ni >= 64 || (sp == 0 && ni > 0)) {
                const int cnt = ni >= 64 ? 64 : ni;
                const int e0 = ni - cnt;
                wave_sync_lds();
                double ecx = 0.0, ecy = 0.0, ecz = 0.0, exd = 0.0, eyd = 0.0, ezd = 0.0;
                unsigned long long emask = 0ull;
                bool cand = false, near = false;
                if (lane < cnt) {
                    const core::TreeNode w = a.tree[S.ent_node[e0 + lane]];
                    emask = S.ent_mask[e0 + lane];
                    // a target never meets its own leaf (r == 0, Node.cpp:274-276):
                    // dropped here, so only distinct particles reach the sweep's
                    // r == 0 and near-pair tests
                    const int own = w.pidx - g * 64;
                    if (w.pidx >= 0 && own >= 0 && own < 64) emask &= ~(1ull << own);
                    ecx = w.cx; ecy = w.cy; ecz = w.cz;
                    if (F32) {
                        S.ent_mask[e0 + lane] = emask;   // the fast sweep reads its lane's bit from here
                        exd = (w.cx - gcx) * invL; eyd = (w.cy - gcy) * invL; ezd = (w.cz - gcz) * invL;
                        S.e.s.f[lane] = make_float4((float)exd, (float)eyd, (float)ezd, (float)(w.mass * invMu));
                        lnint += (unsigned)__builtin_popcountll(emask);   // less the lanes excluded below
                        const double bx = fmax(fmax(lo[0] - w.cx, w.cx - hi[0]), 0.0);
                        const double by = fmax(fmax(lo[1] - w.cy, w.cy - hi[1]), 0.0);
                        const double bz = fmax(fmax(lo[2] - w.cz, w.cz - hi[2]), 0.0);
                        const double eoff = (fabs(w.cx - gcx) + fabs(w.cy - gcy) + fabs(w.cz - gcz)) * invL;
                        const double lim = near_a + 0x1p-20 * eoff;
                        near = (bx * bx + by * by + bz * bz) * (invL * invL) <= lim * lim;
                    } else {
                        // G m staged once per entry (the sweep forms ((G m) rcp) ri, as before)
                        S.e.d.cx[lane] = w.cx; S.e.d.cy[lane] = w.cy; S.e.d.cz[lane] = w.cz;
                        S.e.d.m[lane] = a.wp.soft_live ? w.mass : a.wp.G * w.mass;
                    }
                    if (GAS && sphB && (emask & sphB)) {
                        // a leaf's gas mass is its mass when it is gas; nodes of
                        // mass 0 never become entries
                        const bool gasn = w.gas > 0.0;
                        double ex = fmax(fmax(glo[0] - w.cx, w.cx - ghi[0]), 0.0);
                        double ey = fmax(fmax(glo[1] - w.cy, w.cy - ghi[1]), 0.0);
                        double ez = fmax(fmax(glo[2] - w.cz, w.cz - ghi[2]), 0.0);
                        cand = gasn && ex * ex + ey * ey + ez * ez < reach2;
                    }
                }
                const unsigned long long candB = GAS ? __ballot(cand) : 0ull;
                const unsigned long long nearB = F32 ? __ballot(near) : 0ull;
                // remainders of the entry offsets: in LDS for a dfmode sweep, in
                // the staging lanes' registers for a near chunk's rescan
                float elx = 0.f, ely = 0.f, elz = 0.f;
                if (F32 && nearB && !dfmode) {
                    elx = (float)(exd - (double)(float)exd);
                    ely = (float)(eyd - (double)(float)eyd);
                    elz = (float)(ezd - (double)(float)ezd);
                }
                if (F32 && dfmode && lane < cnt)
                    S.e.s.lo[lane] = make_float4((float)(exd - (double)(float)exd), (float)(eyd - (double)(float)eyd),
                                                 (float)(ezd - (double)(float)ezd), 0.f);
                wave_sync_lds();
                if (KIND == 2) {
                    // no gravity
                } else if (F32) {
                    // Three sweeps, counted from the masks at staging:
                    //  - no entry near a target: every lane of an entry's mask
                    //    interacts (r2 > r2min > 0 for each);
                    //  - near entries: lanes within r2min of an entry (FP32
                    //    offsets cannot resolve the direction) are left out;
                    //  - dfmode (a group that has met such lanes): offsets with
                    //    their FP32 remainders (48 bits; the difference of the
                    //    high parts is exact for close pairs), down to r2df.
                    // Left-out lanes are found again after the sweep: resolved
                    // with the remainders where they can be, else recorded for
                    // the FP64 pass (and dropped from the count).
                    const float4 *fs = S.e.s.f, *fl = S.e.s.lo;
                    const bool dfc = dfmode;
                    float cx = 0.f, cy = 0.f, cz = 0.f;
                    unsigned long long exz = 0ull;   // lanes left out of some entry of this chunk
                    if (dfc) {
                        for (int j = 0; j < cnt; j++) {
                            const unsigned long long Mj = rl(emask, j);
                            const float4 e = fs[j], l = fl[j];
                            const float dx = (e.x - xr) + (l.x - xlo), dy = (e.y - yr) + (l.y - ylo);
                            const float dz = (e.z - zr) + (l.z - zlo);
                            const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                            const float s2 = r2 + e2f;
                            float f = e.w * __builtin_amdgcn_rsqf(r2 * s2 * s2);
                            const unsigned long long okB = Mj & __ballot(r2 > r2df);
                            exz |= Mj & ~okB;
                            f = in_mask(okB) ? f : 0.f;
                            cx = __builtin_fmaf(f, dx, cx);
                            cy = __builtin_fmaf(f, dy, cy);
                            cz = __builtin_fmaf(f, dz, cz);
                        }
                    } else if (!nearB) {
                        // each lane reads the 32-bit half of entry j's mask that
                        // holds its bit and sign-extends the bit into an AND
                        // mask for f (two instructions, against two readlanes
                        // and a select)
                        const u32_alias *mh = reinterpret_cast<const u32_alias *>(S.ent_mask + e0) + (lane >> 5);
                        const int lb = lane & 31;
                        auto one = [&](int j) {
                            const float4 e = fs[j];
                            const float dx = e.x - xr, dy = e.y - yr, dz = e.z - zr;
                            const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                            const float s2 = r2 + e2f;
                            float f = e.w * __builtin_amdgcn_rsqf(r2 * s2 * s2);
                            const int keep = __builtin_amdgcn_sbfe((int)mh[2 * j], lb, 1);
                            f = __int_as_float(__float_as_int(f) & keep);
                            cx = __builtin_fmaf(f, dx, cx);
                            cy = __builtin_fmaf(f, dy, cy);
                            cz = __builtin_fmaf(f, dz, cz);
                        };
                        int j = 0;
                        for (; j + 4 <= cnt; j += 4) { one(j); one(j + 1); one(j + 2); one(j + 3); }
                        for (; j < cnt; j++) one(j);
                    } else {
                        auto one = [&](int j) {
                            const unsigned long long Mj = rl(emask, j);
                            const float4 e = fs[j];
                            const float dx = e.x - xr, dy = e.y - yr, dz = e.z - zr;
                            const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                            const float s2 = r2 + e2f;
                            float f = e.w * __builtin_amdgcn_rsqf(r2 * s2 * s2);
                            const unsigned long long okB = Mj & __ballot(r2 > r2min);
                            exz |= Mj & ~okB;
                            f = in_mask(okB) ? f : 0.f;
                            cx = __builtin_fmaf(f, dx, cx);
                            cy = __builtin_fmaf(f, dy, cy);
                            cz = __builtin_fmaf(f, dz, cz);
                        };
                        int j = 0;
                        for (; j + 4 <= cnt; j += 4) { one(j); one(j + 1); one(j + 2); one(j + 3); }
                        for (; j < cnt; j++) one(j);
                    }
                    unsigned long long hb = 0ull;   // KIND 1: candidate entries resolved or recorded below
                    if (exz) {
                        // the left-out lanes, entry by entry (rare: close pairs, or
                        // a group whose box is far wider than its neighbours' gaps)
                        for (int jj = 0; jj < cnt; jj++) {
                            const unsigned long long Mj = rl(emask, jj);
                            if (!(Mj & exz)) continue;
                            const float4 e = fs[jj];
                            float dx = e.x - xr, dy = e.y - yr, dz = e.z - zr;
                            float lx, ly, lz;
                            if (dfc) { const float4 l = fl[jj]; lx = l.x; ly = l.y; lz = l.z; }
                            else { lx = rl(elx, jj); ly = rl(ely, jj); lz = rl(elz, jj); }
                            unsigned long long Z1;
                            if (dfc) {
                                Z1 = Mj;   // every lane of the entry is tested again below
                            } else {
                                const float r2p = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                                Z1 = Mj & ~__ballot(r2p > r2min);
                            }
                            if (!Z1) continue;
                            dx += lx - xlo; dy += ly - ylo; dz += lz - zlo;
                            const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                            const unsigned long long Z = Z1 & ~__ballot(r2 > r2df);
                            if (!dfc) {
                                const unsigned long long okD = Z1 & ~Z;
                                if (okD) {
                                    const float s2 = r2 + e2f;
                                    float f = e.w * __builtin_amdgcn_rsqf(r2 * s2 * s2);
                                    f = in_mask(okD) ? f : 0.f;
                                    cx = __builtin_fmaf(f, dx, cx);
                                    cy = __builtin_fmaf(f, dy, cy);
                                    cz = __builtin_fmaf(f, dz, cz);
                                    if (KIND == 1 && ((candB >> jj) & 1ull) && in_mask(okD & sphB & __ballot(r2 < lim32)))
                                        hb |= 1ull << jj;
                                }
                            }
                            if (Z) {
                                if (lane == 0 && nrec < kNearRecs) {
                                    a.near_node[(size_t)g * kNearRecs + nrec] = S.ent_node[e0 + jj];
                                    a.near_mask[(size_t)g * kNearRecs + nrec] = Z;
                                }
                                nrec++;
                                nint -= (unsigned long long)__builtin_popcountll(Z);
                                // as SPH candidates they go on the lists as they are
                                // (the pair kernel skips r == 0)
                                if (KIND == 1 && ((candB >> jj) & 1ull) && in_mask(Z & sphB)) hb |= 1ull << jj;
                            }
                        }
                        dfmode = true;   // the rest of this group's sweep carries the remainders
                    }
                    // SPH candidates (KIND 1): the candidate entries within a
                    // lane's margin on the sweep's own r^2 (and those resolved
                    // or recorded above), in entry order
                    if (KIND == 1) {
                        for (unsigned long long cb = candB; cb; cb &= cb - 1) {
                            const int j = __builtin_ctzll(cb);
                            const unsigned long long Mj = rl(emask, j);
                            const float4 e = fs[j];
                            float dx = e.x - xr, dy = e.y - yr, dz = e.z - zr;
                            if (dfc) {
                                const float4 l = fl[j];
                                dx += l.x - xlo; dy += l.y - ylo; dz += l.z - zlo;
                            }
                            const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                            const float cut = dfc ? r2df : (nearB ? r2min : -1.f);
                            const unsigned long long hit = Mj & sphB & __ballot(r2 > cut && r2 < lim32);
                            if (in_mask(hit) || ((hb >> j) & 1ull)) {
                                if (lcnt < kSphK) lbase[lcnt] = S.ent_node[e0 + j];
                                lcnt++;
                            }
                        }
                    }
                    float y = cx - cmpx, tt = sumx + y; cmpx = (tt - sumx) - y; sumx = tt;
                    y = cy - cmpy; tt = sumy + y; cmpy = (tt - sumy) - y; sumy = tt;
                    y = cz - cmpz; tt = sumz + y; cmpz = (tt - sumz) - y; sumz = tt;
                } else {
                    for (int j = 0; j < cnt; j++) {
                        const unsigned long long Mj = rl(emask, j);
                        double dx = S.e.d.cx[j] - px, dy = S.e.d.cy[j] - py, dz = S.e.d.cz[j] - pz;
                        double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
                        bool ok = in_mask(Mj) && r2 != 0.0;
                        double gx = 0.0, gy = 0.0, gz = 0.0;
                        if (ok) {
                            if (a.wp.soft_live) {
                                ok = core::soft_force(S.e.d.m[j], dx, dy, dz, sqrt(r2), a.wp, &gx, &gy, &gz);
                            } else {
                                double ri = rsq64(r2);
                                double rr = r2 * ri;
                                double f = S.e.d.m[j] * rcp64(rr * rr + e02) * ri;
                                gx = f * dx; gy = f * dy; gz = f * dz;
                                ok = !(gx != gx || gy != gy || gz != gz);
                            }
                        }
                        if (ok) { s.ax += gx; s.ay += gy; s.az += gz; }
                        nint += (unsigned long long)__builtin_popcountll(__ballot(ok));
                        if (KIND == 1 && ((candB >> j) & 1ull)) {
                            const unsigned long long hit = __ballot(ok && sph_lane && r2 < lim64);
                            if (hit) {
                                const int nd = S.ent_node[e0 + j];
                                if (in_mask(hit)) {
                                    if (lcnt < kSphK) lbase[lcnt] = nd;
                                    lcnt++;
                                }
                            }
                        }
                    }
                }
                // SPH term.  Pairs (target lane t, entry j) with r < 2 h_t are
                // queued in LDS as they are found and evaluated 64 at a time,
                // one pair per lane: the lane loads t's state (key order) and
                // j's node record again rather than keeping them in
                // registers through the walk, and stages the pair's result in
                // LDS by its queue slot (the staging area, free once the
                // chunk's gravity loop is done).  Each target lane then adds
                // its own pairs' results in slot order, so the sums do not
                // depend on timing.
                if (KIND == 2 && sphB) {
                    if (candB) {
                        double *acc = &S.e.acc[0][0];
                        int np = 0;
                        auto run_pairs = [&](int base, int cntp) {
                            wave_sync_lds();
                            const int pr = lane < cntp ? (int)S.pairs[base + lane] : 0;
                            const int jl = pr >> 8;
                            double wax = 0.0, way = 0.0, waz = 0.0, wdu = 0.0;
                            if (lane < cntp) {
                                const int tl = pr & 63;
                                const int ti = g * 64 + tl;
                                const double tx = a.sx[ti], ty = a.sy[ti], tz = a.sz[ti];
                                const double tvx = a.svx[ti], tvy = a.svy[ti], tvz = a.svz[ti];
                                core::Target tt;
                                tt.h = a.h[ti]; tt.rho = a.rho[ti]; tt.P = a.P[ti];
                                const core::TreeNode nj = a.tree[S.ent_node[e0 + jl]];
                                const double jx = nj.cx, jy = nj.cy, jz = nj.cz, jg = nj.gas;
                                double jvx, jvy, jvz;
                                if (nj.pidx >= 0) { jvx = a.svx[nj.pidx]; jvy = a.svy[nj.pidx]; jvz = a.svz[nj.pidx]; }
                                else { jvx = a.mvx[nj.pre]; jvy = a.mvy[nj.pre]; jvz = a.mvz[nj.pre]; }
                                const core::SphConst k = core::sph_const(tt);
                                core::WalkSums w{0, 0, 0, 0, 0, 0, 0};
                                const double ddx = tx - jx, ddy = ty - jy, ddz = tz - jz;
                                const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                                sph_pair(ddx, ddy, ddz, r2, jg, tvx - jvx, tvy - jvy, tvz - jvz, k, &w);
                                wax = w.ax; way = w.ay; waz = w.az; wdu = w.du;
                                s.nsph += w.nsph;
                            }
                            acc[lane] = wax; acc[64 + lane] = way; acc[128 + lane] = waz; acc[192 + lane] = wdu;
                            wave_sync_lds();
                            for (int k = 0; k < cntp; k++) {
                                if (((int)S.pairs[base + k] & 63) == lane) {
                                    s.ax += acc[k]; s.ay += acc[64 + k]; s.az += acc[128 + k]; s.du += acc[192 + k];
                                }
                            }
                        };
                        for (unsigned long long cb = candB; cb; cb &= cb - 1) {
                            const int j = __builtin_ctzll(cb);
                            const unsigned long long W = rl(emask, j) & sphB;
                            if (!W) continue;
                            const double jx = rl(ecx, j), jy = rl(ecy, j), jz = rl(ecz, j);
                            bool inside = false;
                            if (in_mask(W)) {
                                const double ddx = px - jx, ddy = py - jy, ddz = pz - jz;
                                const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                                if (r2 != 0.0) {
                                    if (r2 < h2x4 * (1.0 - 1e-12)) inside = true;
                                    else if (!(r2 > h2x4 * (1.0 + 1e-12))) inside = sqrt(r2) < a.h[q] * 2;
                                }
                            }
                            const unsigned long long I = __ballot(inside);
                            if (inside) S.pairs[np + mbcnt(I)] = (uint16_t)(lane | (j << 8));
                            np += __builtin_popcountll(I);
                            if (np >= 64) {
                                run_pairs(np - 64, 64);
                                np -= 64;
                            }
                        }
                        if (np > 0) run_pairs(0, np);
                    }
                }
                n_entries += (unsigned long long)cnt;
                ni -= cnt;
            }
            if (sp == 0 && ni == 0) break;
        }
        if (KIND != 2 && lane == 0) {
            bin_add(a.cbins, kCntEntries, g, n_entries);
            bin_add(a.cbins, kCntTests, g, n_tests);
            bin_add(a.cbins, kCntMixed, g, n_mixed);
        }
    }
    if (KIND == 2) {
        // SPH of overflowed targets, added to their gravity
        if (sph_lane) {
            a.ax[q] += s.ax; a.ay[q] += s.ay; a.az[q] += s.az;
            a.dUdt[q] += s.du;
        }
        unsigned long long cs = wave_sum_u64((unsigned long long)s.nsph);
        if (lane == 0) bin_add(a.cbins, kCntSphPairs, g, cs);
        return;
    }
    if (act) {
        double ox, oy, oz;
        if (F32) {
            double scale = a.nn > 0 ? a.wp.G * a.tree[0].mass * invL * invL : 0.0;
            ox = (double)sumx * scale;
            oy = (double)sumy * scale;
            oz = (double)sumz * scale;
        } else {
            ox = s.ax; oy = s.ay; oz = s.az;
        }
        a.ax[q] = ox; a.ay[q] = oy; a.az[q] = oz;
    }
    if (KIND == 1 && slot >= 0) {
        a.sph_cnt[slot] = lcnt;
        if (lcnt > kSphK) a.gover[g] = 1;
    }
    if (KIND == 1) {
        // targets whose candidate list overflowed (their SPH term comes from
        // the overflow walk)
        const unsigned long long ovB = __ballot(slot >= 0 && lcnt > kSphK);
        if (lane == 0) bin_add(a.cbins, kCntSphOver, g, (unsigned long long)__builtin_popcountll(ovB));
    }
    if (F32) {
        nint += wave_sum_u64((unsigned long long)lnint);
        if (lane == 0 && g < (a.n + 63) / 64) a.near_cnt[g] = nrec;
        // a group walked again in FP64 (KIND 3) counts its interactions there
        if (nrec > kNearRecs) nint = 0ull;
    }
    if (lane == 0) bin_add(a.cbins, kCntInteractions, g, nint);
}

__device__ inline int wave_group() {
    return xcd_block(blockIdx.x, gridDim.x) * kWaves + (int)(threadIdx.x >> 6);
}

// Kernels per precision form: gravity with SPH candidate recording (any gas
// uploaded) or plain gravity, and the overflow SPH walk (FP64 for both forms).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_block64(WalkArgs a) {
    walk_block<false, 1>(a, wave_group());
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_block64_nogas(WalkArgs a) {
    walk_block<false, 0>(a, wave_group());
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_block32(WalkArgs a) {
    walk_block<true, 1>(a, wave_group());
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_block32_nogas(WalkArgs a) {
    walk_block<true, 0>(a, wave_group());
}

// The overflow walk and the near pass usually touch few groups, but a
// small opening angle can flag thousands (every SPH list overflowing, many
// near pairs): one wave per group, exiting at once when its group is not
// flagged, keeps those walks in parallel.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_sph_overflow(WalkArgs a) {
    walk_block<false, 2>(a, wave_group());   // exits unless the group is flagged in gover
}

// The near pairs of a split-precision walk in FP64 on the FP64 positions,
// the reference's law (Node.cpp:337-365; r == 0 skipped, Node.cpp:274-276),
// added to the walk's results in record order.  Groups with more records
// than kept are walked again in FP64 instead (k_near_pass).
__device__ inline void near_fp64_group(const WalkArgs &a, int g, int cnt) {
    const int lane = lane_id();
    const int t = g * 64 + lane;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (t < a.n) { px = a.sx[t]; py = a.sy[t]; pz = a.sz[t]; }
    double fx = 0.0, fy = 0.0, fz = 0.0;
    unsigned long long nint = 0ull, any = 0ull, npairs = 0ull;
    for (int k = 0; k < cnt; k++) {
        const int node = a.near_node[(size_t)g * kNearRecs + k];
        const unsigned long long M = a.near_mask[(size_t)g * kNearRecs + k];
        any |= M;
        npairs += (unsigned long long)__builtin_popcountll(M);
        bool live = false;
        if (in_mask(M)) {
            const core::TreeNode w = a.tree[node];
            const double dx = w.cx - px, dy = w.cy - py, dz = w.cz - pz;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 != 0.0) {
                live = true;
                const double r = sqrt(r2);
                const double f = a.wp.G * w.mass / (r * r + a.wp.e0 * a.wp.e0);
                fx += f * (dx / r); fy += f * (dy / r); fz += f * (dz / r);
            }
        }
        nint += (unsigned long long)__builtin_popcountll(__ballot(live));
    }
    if (in_mask(any)) { a.ax[t] += fx; a.ay[t] += fy; a.az[t] += fz; }
    if (lane == 0) {
        bin_add(a.cbins, kCntInteractions, g, nint);
        bin_add(a.cbins, kCntNear, g, npairs);
    }
}

// The near pass: groups with at most kNearRecs recorded pairs evaluate them
// in FP64 (near_fp64_group); groups with more are walked again in FP64.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_near_pass(WalkArgs a) {
    const int g = wave_group();
    if (g >= (a.n + 63) / 64) return;
    const int cnt = a.near_cnt[g];
    if (cnt > kNearRecs) {
        if (lane_id() == 0) bin_add(a.cbins, kCntNearRedo, g, 1ull);
        walk_block<false, 3>(a, g);
    } else if (cnt > 0) {
        near_fp64_group(a, g, cnt);
    }
}

// Batched walk of several independent simulations (a serving round): job j
// owns the global groups [group0[j], group0[j + 1]); a wave finds its job by
// binary search and walks that job's tree for its own targets, so each job's
// arithmetic is the same as in a walk of its own.
struct JobWalk {
    WalkArgs a;
    int group0;
    int pad;
};

__device__ inline int find_job(const JobWalk *jobs, int nj, int G) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].group0 <= G) lo = mid; else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

template <bool F32, int KIND>
__device__ __forceinline__ void walk_jobs(const JobWalk *jobs, int nj, int total) {
    const int G = wave_group();
    if (G >= total) return;   // whole waves only: nothing below synchronises across waves
    const int j = find_job(jobs, nj, G);
    walk_block<F32, KIND>(jobs[j].a, G - jobs[j].group0);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_jobs64(const JobWalk *jobs, int nj, int total) {
    walk_jobs<false, 1>(jobs, nj, total);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_jobs64_nogas(const JobWalk *jobs, int nj, int total) {
    walk_jobs<false, 0>(jobs, nj, total);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_jobs32(const JobWalk *jobs, int nj, int total) {
    walk_jobs<true, 1>(jobs, nj, total);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_jobs32_nogas(const JobWalk *jobs, int nj, int total) {
    walk_jobs<true, 0>(jobs, nj, total);
}


// Active SPH targets (gas, h > 0, due this step) in key order get list
// slots: per-tile counts, a one-workgroup scan, the scatter.  The count
// pass also clears gslot and the groups' overflow flags.
__global__ __launch_bounds__(256) void k_sph_count(int n, const uint8_t *type, const double *h,
                                                   const uint8_t *active, const double *m, int32_t *tiles,
                                                   int32_t *gslot, uint8_t *gover) {
    __shared__ int scr[4];
    int b0 = blockIdx.x * 1024 + threadIdx.x * 4;
    int c = 0;
    for (int k = 0; k < 4; k++) {
        int i = b0 + k;
        if (i < n) {
            gslot[i] = -1;
            if ((i & 63) == 0) gover[i >> 6] = 0;
            c += (type[i] == 2 && h[i] > 0.0 && active[i] && m[i] != 0.0) ? 1 : 0;
        }
    }
    int tot;
    block256_excl_scan(c, scr, &tot);
    if (threadIdx.x == 0) tiles[blockIdx.x] = tot;
}

// The active SPH targets in key order: each tile forms its offset from the
// tile counts; tile 0 records their number.
__global__ __launch_bounds__(256) void k_sph_scatter(int n, const uint8_t *type, const double *h,
                                                     const uint8_t *active, const double *m,
                                                     const int32_t *tiles, int32_t *gidx, int32_t *gslot,
                                                     DevState *ds) {
    __shared__ int scr[4];
    int b0 = blockIdx.x * 1024 + threadIdx.x * 4;
    int f[4], c = 0;
    for (int k = 0; k < 4; k++) {
        int i = b0 + k;
        f[k] = (i < n && type[i] == 2 && h[i] > 0.0 && active[i] && m[i] != 0.0) ? 1 : 0;
        c += f[k];
    }
    int tot;
    const int pre = tile_offset(tiles, (int)gridDim.x, (int)blockIdx.x, scr, &tot);
    if (blockIdx.x == 0 && threadIdx.x == 0) ds->n_sph = tot;
    int off = block256_excl_scan(c, scr, nullptr) + pre;
    for (int k = 0; k < 4; k++)
        if (f[k]) { gidx[off] = b0 + k; gslot[b0 + k] = off; off++; }
}

// The SPH term (Node::calcSPHForce, Node.cpp:88-172) of every recorded
// (target, node) candidate.  Eight lanes per target stride through its
// list; the four sums are then reduced over the eight lanes in a fixed
// butterfly order.  A candidate counts when r < 2 h (decided like the
// reference: on r^2 with a margin, exactly inside it), r != 0 and the node is
// not the target's own leaf.  Lists that overflowed are left to the overflow
// walk.
constexpr int kPairLanes = 8;

__global__ __launch_bounds__(256) void k_sph_pairs(WalkArgs a, const int32_t *gidx, const int32_t *cnts) {
    const int lane = lane_id();
    const int sub = lane & (kPairLanes - 1);
    // targets in key order, in XCD ranges: neighbouring targets share most
    // of their candidate records, which then come from one L2
    const int slot = (int)((xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x) / kPairLanes);
    const int nsph = a.ds->n_sph;
    core::WalkSums s{0, 0, 0, 0, 0, 0, 0};
    int t = -1, c = 0;
    if (slot < nsph) {
        c = cnts[slot];
        if (c <= kSphK) t = gidx[slot];
        else c = 0;
    }
    if (t >= 0) {
        core::Target tg;
        tg.x = a.sx[t]; tg.y = a.sy[t]; tg.z = a.sz[t];
        tg.h = a.h[t]; tg.rho = a.rho[t]; tg.P = a.P[t];
        const double vx = a.svx[t], vy = a.svy[t], vz = a.svz[t];
        const core::SphConst sk = core::sph_const(tg);
        const int32_t *lst = a.sph_list + (size_t)slot * kSphK;
        // one candidate ahead: the next record's load overlaps this pair
        core::SphNode wn{};
        if (sub < c) wn = a.tsph[lst[sub]];
        for (int k = sub; k < c; k += kPairLanes) {
            const core::SphNode w = wn;
            if (k + kPairLanes < c) wn = a.tsph[lst[k + kPairLanes]];
            if (w.pidx == t) continue;
            const double dx = tg.x - w.cx, dy = tg.y - w.cy, dz = tg.z - w.cz;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0) continue;
            bool inside = r2 < sk.h2x4 * (1.0 - 1e-12);
            if (!inside && !(r2 > sk.h2x4 * (1.0 + 1e-12))) inside = sqrt(r2) < sk.h * 2;
            if (!inside) continue;
            sph_pair(dx, dy, dz, r2, w.gas, vx - w.vx, vy - w.vy, vz - w.vz, sk, &s);
        }
    }
    for (int d = kPairLanes / 2; d >= 1; d >>= 1) {
        s.ax += __shfl_xor(s.ax, d, 64);
        s.ay += __shfl_xor(s.ay, d, 64);
        s.az += __shfl_xor(s.az, d, 64);
        s.du += __shfl_xor(s.du, d, 64);
    }
    if (t >= 0 && sub == 0) {
        a.ax[t] += s.ax; a.ay[t] += s.ay; a.az[t] += s.az;
        a.dUdt[t] += s.du;
    }
    unsigned long long cs = wave_sum_u64((unsigned long long)s.nsph);
    if (lane == 0) bin_add(a.cbins, kCntSphPairs, (int)(blockIdx.x * 4 + (threadIdx.x >> 6)), cs);
}

// Each counter of the force phase is the sum of its bins (set, not added, so
// the DevState needs no clearing beforehand); the bins are cleared.
__global__ __launch_bounds__(256) void k_fold_counters(unsigned long long *bins, DevState *ds) {
    __shared__ unsigned long long part[kCounters][4];
    // every bin's loads first: one round trip for all counters
    unsigned long long v[kCounters];
    for (int c = 0; c < kCounters; c++) v[c] = bins[((size_t)c * kCountBins + threadIdx.x) * 8];
    for (int c = 0; c < kCounters; c++) {
        bins[((size_t)c * kCountBins + threadIdx.x) * 8] = 0ull;
        const unsigned long long w = wave_sum_u64(v[c]);
        if (lane_id() == 0) part[c][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t[kCounters];
        for (int c = 0; c < kCounters; c++) t[c] = part[c][0] + part[c][1] + part[c][2] + part[c][3];
        ds->interactions = t[kCntInteractions];
        ds->interactions_total += t[kCntInteractions];
        ds->sph_pairs = t[kCntSphPairs];
        ds->walk_entries = t[kCntEntries];
        ds->walk_tests = t[kCntTests];
        ds->walk_mixed = t[kCntMixed];
        ds->sph_overflow = t[kCntSphOver];
        ds->near_pairs = t[kCntNear];
        ds->near_redo = t[kCntNearRedo];
    }
}

__global__ void k_zero_job_counters(const JobWalk *jobs, int nj) {
    int j = blockIdx.x * 64 + threadIdx.x;
    if (j >= nj) return;
    DevState *ds = jobs[j].a.dsw;
    ds->interactions = 0ull;
    ds->sph_pairs = 0ull;
    ds->walk_entries = 0ull;
    ds->walk_tests = 0ull;
    ds->walk_mixed = 0ull;
    ds->sph_overflow = 0ull;
    ds->near_pairs = 0ull;
    ds->near_redo = 0ull;
}

// One lane per target, stackless pre-order walk (core::walk_fp64); used for
// the potential diagnostic.
__global__ __launch_bounds__(256) void k_walk64(int n, int nin, int nn, const WalkNode *__restrict__ wn,
                                                const double *__restrict__ sx,
                                                const double *__restrict__ sy, const double *__restrict__ sz,
                                                const double *__restrict__ sm, const uint8_t *__restrict__ stype,
                                                const double *__restrict__ svx, const double *__restrict__ svy,
                                                const double *__restrict__ svz, const double *__restrict__ mvx,
                                                const double *__restrict__ mvy, const double *__restrict__ mvz,
                                                const double *__restrict__ h, const double *__restrict__ rho,
                                                const double *__restrict__ P, const uint8_t *__restrict__ active,
                                                core::WalkParams wp, int mode, double *ax, double *ay, double *az,
                                                double *dUdt, double *pot, DevState *ds) {
    int t = blockIdx.x * 256 + threadIdx.x;
    core::WalkSums s{0, 0, 0, 0, 0, 0, 0};
    if (t < n) {
        int q = t;
        bool act = mode == 1 || active[q];
        if (act) {
            core::Target tg;
            tg.x = sx[t]; tg.y = sy[t]; tg.z = sz[t]; tg.m = sm[t];
            tg.vx = svx[t]; tg.vy = svy[t]; tg.vz = svz[t];
            tg.h = h[q]; tg.rho = rho[q]; tg.P = P[q];
            tg.gas = stype[t] == 2;
            tg.self = t < nin ? t : -1;
            core::walk_fp64(wn, nn, tg, svx, svy, svz, stype, mvx, mvy, mvz, wp, mode, &s);
            if (mode == 0) {
                ax[q] = s.ax; ay[q] = s.ay; az[q] = s.az;
                dUdt[q] += s.du;
            } else {
                pot[q] = s.pot;
            }
        }
    }
    unsigned long long a = wave_sum_u64((unsigned long long)s.nint);
    unsigned long long b = wave_sum_u64((unsigned long long)s.nsph);
    if (lane_id() == 0 && mode == 0) {
        if (a) atomicAdd(&ds->interactions, a);
        if (b) atomicAdd(&ds->sph_pairs, b);
    }
}

WalkArgs walk_args(Ctx &c) {
    WalkArgs a;
    a.n = (int)c.n; a.nin = c.nin; a.nn = c.n_nodes;
    a.tree = c.tnode;
    a.tsph = c.tsph;
    a.gslot = c.gslot; a.sph_list = c.sph_list; a.sph_cnt = c.sph_cnt; a.gover = c.gover;
    a.sx = c.x; a.sy = c.y; a.sz = c.z; a.sm = c.m; a.stype = c.type;
    a.svx = c.vx; a.svy = c.vy; a.svz = c.vz;
    a.mvx = c.n_gvx; a.mvy = c.n_gvy; a.mvz = c.n_gvz;
    a.h = c.h; a.rho = c.rho; a.P = c.P; a.active = c.active;
    a.ds = c.ds;
    a.wp = core::WalkParams{c.p.theta, c.p.e0, c.p.G, core::soft_branch_live(c.p.e0)};
    a.ax = c.ax; a.ay = c.ay; a.az = c.az; a.dUdt = c.dUdt;
    a.dsw = c.ds;
    a.cbins = c.cbins;
    a.near_cnt = c.near_cnt; a.near_node = c.near_node; a.near_mask = c.near_mask;
    return a;
}

}  // namespace

size_t job_walk_bytes() { return sizeof(JobWalk); }

void forces_batch(Ctx *const *cs, int nc, hipStream_t st, int f32, void *scratch_dev, void *scratch_host) {
    if (nc <= 0) return;
    JobWalk *hj = (JobWalk *)scratch_host;
    bool gas = false;
    int total = 0;
    for (int k = 0; k < nc; k++) {
        hj[k].a = walk_args(*cs[k]);
        hj[k].group0 = total;
        hj[k].pad = 0;
        total += (int)((cs[k]->n + 63) / 64);
        gas = gas || cs[k]->sph_active;
    }
    JobWalk *dj = (JobWalk *)scratch_dev;
    AG2_HIP(hipMemcpyAsync(dj, hj, sizeof(JobWalk) * (size_t)nc, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_zero_job_counters, dim3((nc + 63) / 64), dim3(64), 0, st, dj, nc);
    dim3 grid((total + kWaves - 1) / kWaves), block(64 * kWaves);
    if (f32) {
        if (gas) hipLaunchKernelGGL(k_walk_jobs32, grid, block, 0, st, dj, nc, total);
        else hipLaunchKernelGGL(k_walk_jobs32_nogas, grid, block, 0, st, dj, nc, total);
    } else {
        if (gas) hipLaunchKernelGGL(k_walk_jobs64, grid, block, 0, st, dj, nc, total);
        else hipLaunchKernelGGL(k_walk_jobs64_nogas, grid, block, 0, st, dj, nc, total);
    }
    AG2_HIP(hipGetLastError());
}

void fold_counters(Ctx &c) {
    static_assert(kCountBins == 256, "k_fold_counters runs one thread per bin");
    hipLaunchKernelGGL(k_fold_counters, dim3(1), dim3(256), 0, c.stream, c.cbins, c.ds);
    AG2_HIP(hipGetLastError());
}

// SPH candidate lists for a step: slots for the active SPH targets (before
// the gravity walk, which fills the lists) ...
void sph_prepare(Ctx &c) {
    if (!c.sph_active) return;
    hipStream_t st = c.stream;
    int n = (int)c.n;
    int nt = (n + 1023) / 1024;
    hipLaunchKernelGGL(k_sph_count, dim3(nt), dim3(256), 0, st, n, c.type, c.h, c.active, c.m, c.gtiles, c.gslot,
                       c.gover);
    hipLaunchKernelGGL(k_sph_scatter, dim3(nt), dim3(256), 0, st, n, c.type, c.h, c.active, c.m, c.gtiles, c.gidx,
                       c.gslot, c.ds);
    AG2_HIP(hipGetLastError());
}

// ... and their evaluation after it, with the overflow walk for lists that
// did not fit (it exits at once for groups without one).
void sph_finish(Ctx &c) {
    if (!c.sph_active) return;
    hipStream_t st = c.stream;
    WalkArgs a = walk_args(c);
    int64_t lanes = c.gas_count * kPairLanes;
    int nb = (int)((lanes + 255) / 256);
    if (nb > 0) hipLaunchKernelGGL(k_sph_pairs, dim3(nb), dim3(256), 0, st, a, c.gidx, c.sph_cnt);
    const int groups = (int)((c.n + 63) / 64);
    hipLaunchKernelGGL(k_walk_sph_overflow, dim3((groups + kWaves - 1) / kWaves), dim3(64 * kWaves), 0, st, a);
    AG2_HIP(hipGetLastError());
}

// After a split-precision walk: the near pairs in FP64, then the groups with
// too many of them walked again in FP64 (both exit at once for other groups).
void near_finish(Ctx &c) {
    WalkArgs a = walk_args(c);
    const int groups = (int)((c.n + 63) / 64);
    hipLaunchKernelGGL(k_near_pass, dim3((groups + kWaves - 1) / kWaves), dim3(64 * kWaves), 0, c.stream, a);
    AG2_HIP(hipGetLastError());
}

void forces(Ctx &c, int mode) {
    hipStream_t st = c.stream;
    int n = (int)c.n;
    if (c.timing >= 2) AG2_HIP(hipEventRecord(c.ev[4], st));
    if (mode == 0) {
        sph_prepare(c);
        WalkArgs a = walk_args(c);
        int groups = (n + 63) / 64;
        if (c.timing >= 1) AG2_HIP(hipEventRecord(c.ev[8], st));
        if (c.sph_active)
            hipLaunchKernelGGL(k_walk_block64, dim3((groups + kWaves - 1) / kWaves), dim3(64 * kWaves), 0, st, a);
        else
            hipLaunchKernelGGL(k_walk_block64_nogas, dim3((groups + kWaves - 1) / kWaves), dim3(64 * kWaves), 0, st, a);
        if (c.timing >= 1) AG2_HIP(hipEventRecord(c.ev[9], st));
        sph_finish(c);
        fold_counters(c);
    } else {
        core::WalkParams wp{c.p.theta, c.p.e0, c.p.G, core::soft_branch_live(c.p.e0)};
        pack_walknodes(c);
        hipLaunchKernelGGL(k_walk64, dim3(grid_for(n, 256)), dim3(256), 0, st, n, c.nin, c.n_nodes, c.wnode,
                           c.x, c.y, c.z, c.m, c.type, c.vx, c.vy, c.vz, c.n_gvx, c.n_gvy, c.n_gvz, c.h,
                           c.rho, c.P, c.active, wp, mode, c.ax, c.ay, c.az, c.dUdt, c.pot, c.ds);
    }
    if (c.timing >= 2) AG2_HIP(hipEventRecord(c.ev[5], st));
    AG2_HIP(hipGetLastError());
}

// The split-precision sweep works in units of the root half-width R, and its
// force m' / sqrt(r'^2 (r'^2 + e'^2)^2) needs (e0 / R)^4 well inside the FP32
// exponent range.  When R has grown far beyond the softening length (escaped
// particles can push it to 1e24 kpc), the FP64 sweep runs instead.
bool fp32_usable(const Ctx &c) {
    const double R = __builtin_bit_cast(double, c.hs.Rbits);
    return !(R > 0.0) || c.p.e0 >= ldexp(R, -24);
}

void forces_fp32(Ctx &c, int mode) {
    if (mode != 0) return forces(c, mode);
    hipStream_t st = c.stream;
    int n = (int)c.n;
    if (c.timing >= 2) AG2_HIP(hipEventRecord(c.ev[4], st));
    sph_prepare(c);
    WalkArgs a = walk_args(c);
    int groups = (n + 63) / 64;
    if (c.timing >= 1) AG2_HIP(hipEventRecord(c.ev[8], st));
    if (c.sph_active)
        hipLaunchKernelGGL(k_walk_block32, dim3((groups + kWaves - 1) / kWaves), dim3(64 * kWaves), 0, st, a);
    else
        hipLaunchKernelGGL(k_walk_block32_nogas, dim3((groups + kWaves - 1) / kWaves), dim3(64 * kWaves), 0, st, a);
    if (c.timing >= 1) AG2_HIP(hipEventRecord(c.ev[9], st));
    near_finish(c);
    sph_finish(c);
    fold_counters(c);
    if (c.timing >= 2) AG2_HIP(hipEventRecord(c.ev[5], st));
    AG2_HIP(hipGetLastError());
}

}  // namespace ag2
