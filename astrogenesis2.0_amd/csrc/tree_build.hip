// Octree build on the device: root half-width, 42-level octant keys,
// LSD radix sort, tie refinement, linear-octree topology and bottom-up
// moments.  Replaces Tree::buildTree / Node::insert (Tree.cpp:24-55,
// Node.cpp:405-699).  All passes are HBM-bound integer/FP64 streaming work.
#include "ag2_internal.h"

namespace ag2 {
namespace {

constexpr int kTwBlocks = 512;   // fixed grid: fixed reduction order for R
constexpr unsigned long long kOut = ~0ull;

// ------------------------------------------------------------ tree width
__global__ __launch_bounds__(256) void k_tw_partial(int n, const double *__restrict__ x,
                                                    const double *__restrict__ y,
                                                    const double *__restrict__ z, double *part) {
    double s = 0.0, q = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double d = sqrt(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
        s += d;
        q += d * d;
    }
    s = wave_sum(s);
    q = wave_sum(q);
    __shared__ double ls[4], lq[4];
    int w = threadIdx.x >> 6;
    if (lane_id() == 0) { ls[w] = s; lq[w] = q; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ((ls[0] + ls[1]) + ls[2]) + ls[3];
        part[2 * blockIdx.x + 1] = ((lq[0] + lq[1]) + lq[2]) + lq[3];
    }
}

// Tree::calcTreeWidth (Tree.cpp:86-117): cut = mean + 10 sigma of |x|,
// then the largest |x| within the cut.  Every workgroup folds the sum
// partials itself in the same fixed order (so all get the same cut, and
// workgroup 0 records it), then writes the maximum of its particles.
__global__ __launch_bounds__(256) void k_tw_max(int n, const double *__restrict__ x, const double *__restrict__ y,
                                                const double *__restrict__ z, int nb, const double *tw,
                                                DevState *ds, unsigned long long *part) {
    __shared__ unsigned long long wmax[4];
    __shared__ double ls[4], lq[4], scut;
    {
        double s = 0.0, q = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) { s += tw[2 * b]; q += tw[2 * b + 1]; }
        s = wave_sum(s);
        q = wave_sum(q);
        int w = threadIdx.x >> 6;
        if (lane_id() == 0) { ls[w] = s; lq[w] = q; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double sum = ((ls[0] + ls[1]) + ls[2]) + ls[3];
            double sq = ((lq[0] + lq[1]) + lq[2]) + lq[3];
            double mean = sum / (double)n;
            double sd = sqrt(sq / (double)n - mean * mean);
            scut = mean + 10.0 * sd;
            if (blockIdx.x == 0) {
                ds->sum = sum;
                ds->sumsq = sq;
                ds->cut = scut;
                ds->gap = 0ull;
                ds->err = 0;
                ds->max_level = 0;
                ds->tie_big = 0;
                ds->tie_deep = 0;
                ds->any_bucket = 0;
                ds->max_prefix_run = 0;
            }
        }
        __syncthreads();
    }
    double cut = scut;
    unsigned long long best = 0ull;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double d = sqrt(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
        if (d <= cut) {
            unsigned long long b = (unsigned long long)__double_as_longlong(d);
            best = b > best ? b : best;
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        unsigned long long t = __shfl_xor(best, o, 64);
        best = t > best ? t : best;
    }
    if (lane_id() == 0) wmax[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = wmax[0];
        for (int w = 1; w < 4; w++) b = wmax[w] > b ? wmax[w] : b;
        part[blockIdx.x] = b;
    }
}

// ------------------------------------------------------------ keys
// Every workgroup first folds the per-workgroup maxima of k_tw_max into R
// (workgroup 0 records it).
__global__ __launch_bounds__(256) void k_keygen(int n, const double *__restrict__ x, const double *__restrict__ y,
                                                const double *__restrict__ z, int nb,
                                                const unsigned long long *rmax, DevState *ds,
                                                uint64_t *key, uint32_t *idx, uint8_t *gapl) {
    __shared__ unsigned long long w4[4];
    // the position is loaded first: its loads are in flight while R is folded
    const int i = blockIdx.x * 256 + threadIdx.x;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (i < n) { px = x[i]; py = y[i]; pz = z[i]; }
    unsigned long long v = 0ull;
    for (int k = threadIdx.x; k < nb; k += 256) v = rmax[k] > v ? rmax[k] : v;
    for (int o = 32; o >= 1; o >>= 1) {
        unsigned long long q = __shfl_xor(v, o, 64);
        v = q > v ? q : v;
    }
    if (lane_id() == 0) w4[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long rb = w4[0];
    for (int w = 1; w < 4; w++) rb = w4[w] > rb ? w4[w] : rb;
    if (blockIdx.x == 0 && threadIdx.x == 0) ds->Rbits = rb;
    if (i >= n) return;
    double R = __longlong_as_double((long long)rb);
    uint64_t k1, k2;
    int g;
    bool in = core::octant_path(px, py, pz, R, 21, &k1, &k2, &g);
    key[i] = in ? k1 : kOut;
    idx[i] = (uint32_t)i;
    gapl[i] = in ? (uint8_t)g : (uint8_t)255;
}

// ------------------------------------------------------------ radix sort
// Workgroup tile: kSortTile keys, kSortItems per thread; wave w owns tile
// positions [64 kSortItems w, 64 kSortItems (w + 1)), item j of lane l at
// 64 kSortItems w + 64 j + l, so (wave, item, lane) is tile order and the
// scatter is stable.
constexpr int kSortItems = kSortTile / 256;
constexpr int kSortSmallTiles = 16;   // up to this many tiles the scatter forms its own offsets
__device__ __forceinline__ void rs_hist_tile(int n, const uint64_t *__restrict__ key, int shift, uint32_t *hist,
                                             int nblk, int tile) {
    __shared__ uint32_t cnt[256];
    int w = threadIdx.x >> 6;
    int base = tile * kSortTile + w * (64 * kSortItems) + lane_id();
    // the keys are loaded before the counters are cleared: the loads are in
    // flight across the barrier
    int dg[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        int pos = base + j * 64;
        dg[j] = pos < n ? (int)((key[pos] >> shift) & 255) : -1;
    }
    cnt[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; j++)
        if (dg[j] >= 0) atomicAdd(&cnt[dg[j]], 1u);
    __syncthreads();
    hist[threadIdx.x * nblk + tile] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_rs_hist(int n, const uint64_t *__restrict__ key, int shift,
                                                 uint32_t *hist, int nblk) {
    rs_hist_tile(n, key, shift, hist, nblk, blockIdx.x);
}

// One workgroup per digit: exclusive scan of that digit's per-tile counts.
__device__ __forceinline__ void rs_scan_digit(int nblk, uint32_t *hist, uint32_t *dtotal, int digit) {
    __shared__ int scr[4];
    uint32_t *h = hist + (size_t)digit * nblk;
    int per = (nblk + 255) / 256;
    int s0 = threadIdx.x * per;
    int s1 = s0 + per < nblk ? s0 + per : nblk;
    int local = 0;
    for (int k = s0; k < s1; k++) local += (int)h[k];
    int tot;
    int run = block256_excl_scan(local, scr, &tot);
    for (int k = s0; k < s1; k++) {
        int t = (int)h[k];
        h[k] = (uint32_t)run;
        run += t;
    }
    if (threadIdx.x == 0) dtotal[digit] = (uint32_t)tot;
}

__global__ __launch_bounds__(256) void k_rs_scan(int nblk, uint32_t *hist, uint32_t *dtotal) {
    rs_scan_digit(nblk, hist, dtotal, blockIdx.x);
}

// SMALL: hist holds the raw per-tile counts (no k_rs_scan pass) and each
// workgroup forms its own offsets from them -- nblk reads per digit, cheap
// for the few tiles of a small particle set, and one launch fewer per pass.
template <bool SMALL>
__device__ __forceinline__ void rs_scatter_tile(int n, const uint64_t *__restrict__ kin,
                                                const uint32_t *__restrict__ vin, uint64_t *__restrict__ kout,
                                                uint32_t *__restrict__ vout, int shift,
                                                const uint32_t *__restrict__ hist,
                                                const uint32_t *__restrict__ dtotal, int nblk, int tile) {
    __shared__ int sbase[256];
    __shared__ int wcnt[4][256];
    __shared__ int scr[4];
    int tid = threadIdx.x;
    int tot;
    int w = tid >> 6;
    int base = tile * kSortTile + w * (64 * kSortItems) + lane_id();
    // the tile's keys are loaded first: the offsets below do not need them,
    // so the loads are in flight while the offsets are formed
    uint64_t keys[kSortItems];
    uint32_t vals[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const int pos = base + j * 64;
        keys[j] = pos < n ? kin[pos] : 0ull;
        vals[j] = pos < n ? vin[pos] : 0u;
    }
    if (SMALL) {
        const uint32_t *h = hist + (size_t)tid * nblk;
        int before = 0, all = 0;
        for (int t = 0; t < nblk; t++) {
            const int v = (int)h[t];
            before += t < tile ? v : 0;
            all += v;
        }
        int ex = block256_excl_scan(all, scr, &tot);
        sbase[tid] = ex + before;
    } else {
        int ex = block256_excl_scan((int)dtotal[tid], scr, &tot);
        sbase[tid] = ex + (int)hist[tid * nblk + tile];
    }
    wcnt[0][tid] = 0; wcnt[1][tid] = 0; wcnt[2][tid] = 0; wcnt[3][tid] = 0;
    __syncthreads();
    unsigned long long below = lanes_below();
    int loc[kSortItems], dig[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        int pos = base + j * 64;
        bool valid = pos < n;
        const uint64_t k = keys[j];
        int d = (int)((k >> shift) & 255);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            int bit = (d >> b) & 1;
            unsigned long long bm = __ballot(bit);
            peers &= bit ? bm : ~bm;
        }
        if (!valid) peers = 0ull;
        int rank = __popcll(peers & below);
        int cur = valid ? wcnt[w][d] : 0;
        if (valid && (peers & below) == 0ull) wcnt[w][d] = cur + __popcll(peers);
        loc[j] = valid ? cur + rank : -1; dig[j] = d;
    }
    __syncthreads();
    {
        int s = 0;
        for (int k = 0; k < 4; k++) { int t = wcnt[k][tid]; wcnt[k][tid] = s; s += t; }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        if (loc[j] >= 0) {
            int dst = sbase[dig[j]] + wcnt[w][dig[j]] + loc[j];
            kout[dst] = keys[j];
            vout[dst] = vals[j];
        }
    }
}

template <bool SMALL>
__global__ __launch_bounds__(256) void k_rs_scatter(int n, const uint64_t *__restrict__ kin,
                                                    const uint32_t *__restrict__ vin, uint64_t *__restrict__ kout,
                                                    uint32_t *__restrict__ vout, int shift,
                                                    const uint32_t *__restrict__ hist,
                                                    const uint32_t *__restrict__ dtotal, int nblk) {
    rs_scatter_tile<SMALL>(n, kin, vin, kout, vout, shift, hist, dtotal, nblk, blockIdx.x);
}

// Batched sort of several contexts (a serving round): job j's tiles are the
// workgroups [tile0, tile0 + nblk) of one launch per kernel and pass, found
// by binary search; each job keeps its own buffers, so its result is that of
// its own sort.
struct SortJob {
    int n, nblk, tile0, pad;
    uint64_t *key, *key_alt;
    uint32_t *idx, *idx_alt;
    uint32_t *hist, *dtotal;
};

__device__ inline int find_sort_job(const SortJob *jobs, int nj, int b) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_rsb_hist(const SortJob *jobs, int nj, int pass) {
    const SortJob &J = jobs[find_sort_job(jobs, nj, blockIdx.x)];
    rs_hist_tile(J.n, (pass & 1) ? J.key_alt : J.key, 8 * pass, J.hist, J.nblk, (int)blockIdx.x - J.tile0);
}

__global__ __launch_bounds__(256) void k_rsb_scan(const SortJob *jobs) {
    const SortJob &J = jobs[blockIdx.x >> 8];
    rs_scan_digit(J.nblk, J.hist, J.dtotal, blockIdx.x & 255);
}

__global__ __launch_bounds__(256) void k_rsb_scatter(const SortJob *jobs, int nj, int pass) {
    const SortJob &J = jobs[find_sort_job(jobs, nj, blockIdx.x)];
    const bool odd = pass & 1;
    rs_scatter_tile<false>(J.n, odd ? J.key_alt : J.key, odd ? J.idx_alt : J.idx, odd ? J.key : J.key_alt,
                           odd ? J.idx : J.idx_alt, 8 * pass, J.hist, J.dtotal, J.nblk, (int)blockIdx.x - J.tile0);
}

// ------------------------------------------------------------ after sort

// Tie runs (equal word 0) in key order get their order by (word 1, caller
// index): the reference inserts a run's particles in array order
// (Tree.cpp:44-48) whatever resident order they were sorted from.  Each
// particle of a run of k counts the members before it (k steps per thread),
// and everything goes to the alternate buffers: word 1 to w1, the sort
// permutation to ord (particles outside runs keep their place).  Runs longer
// than kTieRankMax are left in order and flagged in ds->tie_big; the build
// then sorts them with the general segmented sort (refine_ties).
constexpr int kTieRankMax = 1024;

// First and last particle of the run sharing k[i] >> shift (galloping).
__device__ inline int word_run_first(const uint64_t *k, int i, int shift = 0) {
    const uint64_t pi = k[i] >> shift;
    int hi = i, lo = -1;
    for (int step = 1;; step <<= 1) {
        const int probe = i - step;
        if (probe < 0) break;
        if ((k[probe] >> shift) != pi) { lo = probe; break; }
        hi = probe;
        if (step > (1 << 29)) break;
    }
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if ((k[mid] >> shift) == pi) hi = mid; else lo = mid;
    }
    return hi;
}

__device__ inline int word_run_last(const uint64_t *k, int i, int nin, int shift = 0) {
    const uint64_t pi = k[i] >> shift;
    int lo = i, hi = nin;
    for (int step = 1;; step <<= 1) {
        const int probe = i + step;
        if (probe >= nin) break;
        if ((k[probe] >> shift) != pi) { hi = probe; break; }
        lo = probe;
        if (step > (1 << 29)) break;
    }
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if ((k[mid] >> shift) == pi) lo = mid; else hi = mid;
    }
    return lo;
}

// The particles in the root cube (sorted before the others, whose key is
// kOut) -> ds->nin.  Runs of particles sharing the sorted bits of word 0
// (key >> shift) are ordered by k_tie_rank; a particle whose whole word 0
// another particle of its run has gets its levels 22..42 (word 1) here.
// kOut shares no prefix with a key in the cube (bit 63 is clear in those),
// so runs stay below nin.
__global__ __launch_bounds__(256) void k_tie_key2(int n, const uint64_t *key, const uint32_t *idx,
                                                  const double *x, const double *y, const double *z,
                                                  DevState *ds, uint64_t *key2, uint8_t *gapl, int shift) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint64_t ki = key[i];
    const bool in = ki != kOut;
    if (in && (i + 1 == n || key[i + 1] == kOut)) ds->nin = i + 1;
    if (i == 0 && !in) ds->nin = 0;
    if (!in) return;
    const uint64_t pi = ki >> shift;
    const bool run = (i > 0 && (key[i - 1] >> shift) == pi) || (i + 1 < n && (key[i + 1] >> shift) == pi);
    if (!run) return;   // k_tie_rank reads word 1 of run members only
    bool tie = true;
    if (shift > 0) {
        // short runs: is the whole key shared?  (a run too long for the
        // ranking sends the build back to a full sort)
        tie = false;
        const int s = word_run_first(key, i, shift), e = word_run_last(key, i, n, shift);
        if (e - s + 1 <= kTieRankMax)
            for (int j = s; j <= e && !tie; j++) tie = j != i && key[j] == ki;
    }
    if (!tie) { key2[i] = 0ull; return; }
    double R = __longlong_as_double((long long)ds->Rbits);
    uint32_t q = idx[i];
    uint64_t a, b;
    int g;
    core::octant_path(x[q], y[q], z[q], R, 42, &a, &b, &g);
    key2[i] = b;
    if (gapl[q] == 0 && g > 0) gapl[q] = (uint8_t)(g > 254 ? 254 : g);
}

// Runs of particles sharing key >> shift (the radix-sorted bits) take their
// order by (word 0, word 1, caller index): the reference inserts particles
// of one cell in array order (Tree.cpp:44-48) whatever resident order they
// were sorted from.  Each particle of a run of k counts the members before it
// (k steps per thread), and everything goes to the output buffers: word 0 to
// w0, word 1 to w1, the sort permutation to ord (particles outside runs keep
// their place).  Runs longer than kTieRankMax are left in order and flagged
// in ds->tie_big: the build then sorts every byte of word 0 again (shift >
// 0) or refines them with the general segmented sort (refine_ties).
__global__ __launch_bounds__(256) void k_tie_rank(int n, const uint64_t *key, const uint64_t *key2,
                                                  const uint32_t *idx, const uint32_t *rid, DevState *ds,
                                                  uint64_t *w0, uint64_t *w1, uint32_t *ord, int shift) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int nin = ds->nin;
    const uint64_t k0 = key[i];
    if (i >= nin) { ord[i] = idx[i]; w0[i] = k0; return; }
    const uint64_t p0 = k0 >> shift;
    const bool run = (i > 0 && (key[i - 1] >> shift) == p0) || (i + 1 < nin && (key[i + 1] >> shift) == p0);
    if (!run) { ord[i] = idx[i]; w0[i] = k0; w1[i] = 0ull; return; }
    const int s = word_run_first(key, i, shift), e = word_run_last(key, i, nin, shift);
    if (i == s) atomicMax(&ds->max_prefix_run, e - s + 1);
    if (e - s + 1 > kTieRankMax) {
        ord[i] = idx[i];
        w0[i] = k0;
        w1[i] = key2[i];
        if (i == s) ds->tie_big = 1;
        return;
    }
    const uint64_t ki = key2[i];
    const uint32_t ci = rid[idx[i]];
    int rank = 0;
    for (int j = s; j <= e; j++) {
        const uint64_t kj0 = key[j];
        if (kj0 < k0) {
            rank++;
        } else if (kj0 == k0) {
            const uint64_t kj = key2[j];
            rank += (kj < ki || (kj == ki && rid[idx[j]] < ci)) ? 1 : 0;
        }
    }
    ord[s + rank] = idx[i];
    w0[s + rank] = k0;
    w1[s + rank] = ki;
}

// The whole particle state moves into the new key order: out[k] = in[idx[k]].
// Particles move little between steps, so idx is close to the identity and
// the reads stay nearly coalesced.
struct StateSet {
    double *d[kStateD];
    uint8_t *type, *active;
    uint32_t *rid;
};

// Output slots in XCD ranges: neighbouring slots gather from neighbouring
// source lines, which then share one L2.
// skip_acc: fields kAccField .. kAccField + 2 (the accelerations) are left
// out, for a build whose force phase writes every one of them.  zero_gas_h:
// field kHField (h) becomes 0 for gas, for a build whose density phase
// follows (it starts so, Tree.cpp:121-130).
constexpr int kAccField = 9;
constexpr int kHField = 12;

__global__ __launch_bounds__(256) void k_permute(int n, const uint32_t *__restrict__ idx, StateSet in, StateSet out,
                                                 int skip_acc, int zero_gas_h) {
    int k = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (k >= n) return;
    uint32_t q = idx[k];
    const uint8_t ty = in.type[q];
#pragma unroll
    for (int j = 0; j < kStateD; j++) {
        if (skip_acc && j >= kAccField && j < kAccField + 3) continue;
        out.d[j][k] = (j == kHField && zero_gas_h && ty == 2) ? 0.0 : in.d[j][q];
    }
    out.type[k] = ty;
    out.active[k] = in.active[q];
    out.rid[k] = in.rid[q];
}

__global__ __launch_bounds__(256) void k_inverse(int n, const uint32_t *rid, int32_t *inv) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) inv[rid[r]] = r;
}

// L[i] and c[i] = max(0, L[i] - L[i-1]), over the key words of this build
// (nw of them) and the positions in key order (runs at one position are
// buckets).  A distinct pair that agrees on every word sets ds->tie_deep:
// the build refines it with deeper words (refine_ties), or fails past
// kDeepest levels.
// Four particles per thread (a 1024-particle scan tile per workgroup); the
// tile's sum of c goes to bsum for k_scan_apply.
__global__ __launch_bounds__(256) void k_lcp(int n, core::KeyWords kw, int nw, core::Pos pos, DevState *ds,
                                             int parallel_root, int8_t *L, int8_t *Lk, int32_t *c, int32_t *bsum) {
    __shared__ int scr[4];
    const int nin = ds->nin;
    const bool proot = parallel_root != 0, splits = proot && nin < n;
    int tile = 0;
    for (int k = 0; k < 4; k++) {
        const int i = blockIdx.x * 1024 + k * 256 + threadIdx.x;
        if (i >= nin) continue;
        int li = core::pair_lcp(kw, pos, i, nin, proot, splits);
        int lp = i == 0 ? -1 : core::pair_lcp(kw, pos, i - 1, nin, proot, splits);
        // a coincident pair's level is its bucket's (pair_lcp), whatever the key words say
        const bool coincident = i + 1 < nin && core::same_pos(pos, i, i + 1);
        if (coincident) ds->any_bucket = 1;
        if (i + 1 < nin && li >= 21 * nw && !coincident) ds->tie_deep = 1;
        if (li > core::kDeepest) { li = core::kDeepest; ds->tie_deep = 1; }
        if (lp > core::kDeepest) lp = core::kDeepest;
        L[i] = (int8_t)li;
        // the keys' own adjacent levels, for the cell searches (buckets share every level)
        Lk[i] = i + 1 < nin ? (int8_t)core::shared_levels(kw, i, i + 1) : (int8_t)-1;
        const int v = li - lp > 0 ? li - lp : 0;
        c[i] = v;
        tile += v;
    }
    int tot;
    block256_excl_scan(tile, scr, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// The exclusive scan of c into base over tiles of 1024 (256 threads x 4):
// each tile forms its own offset from the tile sums k_lcp left; tile 0 also
// records the node counts.
__global__ __launch_bounds__(256) void k_scan_apply(const int32_t *c, const int32_t *bsum, DevState *ds,
                                                    int parallel_root, int64_t node_cap, int32_t *base) {
    __shared__ int scr[4];
    int nin = ds->nin;
    int b0 = blockIdx.x * 1024 + threadIdx.x * 4;
    int v[4];
    int s = 0;
    for (int k = 0; k < 4; k++) { v[k] = (b0 + k < nin) ? c[b0 + k] : 0; s += v[k]; }
    int tot;
    const int pre = tile_offset(bsum, (int)gridDim.x, (int)blockIdx.x, scr, &tot);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int n_int = (nin == 0) ? (parallel_root ? 1 : 0) : tot;
        ds->n_int = n_int;
        ds->n_nodes = n_int + nin;
        if ((int64_t)n_int + nin > node_cap) ds->err = 1;
    }
    int ex = block256_excl_scan(s, scr, nullptr) + pre;
    for (int k = 0; k < 4; k++) {
        if (b0 + k <= nin) base[b0 + k] = ex;
        ex += v[k];
    }
}

// ------------------------------------------------------------ nodes
// One thread per internal node.  The internal nodes, numbered q = 0 .. n_int
// - 1 in pre-order, are those of particle i for base[i] <= q < base[i] + c[i]
// (k_make_leaf lists own[q] = i); node q of particle i sits at pre-order
// index i + q, at level L[i-1] + 1 + (q - base[i]).  A particle can start a
// chain of many nested cells, each needing a galloping search for its last
// particle; giving every cell its own thread keeps those searches from
// running one after another in a single lane, and listing only internal
// nodes keeps every lane of a wave searching.
__global__ __launch_bounds__(256) void k_make_internal(int nin, int nn, int n_int, core::KeyWords kw,
                                                       const int8_t *L, const int32_t *base, const int32_t *own,
                                                       core::NodeOut o) {
    int q = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (q >= n_int) return;
    int i = own[q];
    int l = core::lcp_prev(L, i) + 1 + (q - base[i]);
    core::make_internal(kw, L, base, i, l, nin, nn, o);
}

__global__ __launch_bounds__(256) void k_make_root_only(core::NodeOut o) {
    o.first[0] = 0; o.count[0] = 0; o.skip[0] = 1; o.parent[0] = -1; o.pidx[0] = -1; o.level[0] = 0;
}

// Grid-stride over a fixed grid (at most kLeafBlocks workgroups), so the
// deepest level and the gap count cost one pair of atomics per workgroup.
constexpr int kLeafBlocks = 1024;

__global__ __launch_bounds__(256) void k_make_leaf(int nin, core::KeyWords kw,
                                                   const int8_t *L, const int8_t *Lk, const int32_t *base,
                                                   const int32_t *c,
                                                   core::NodeOut o, const uint32_t *idx, const uint8_t *gapl,
                                                   int32_t *own, unsigned long long *part_lv,
                                                   unsigned long long *part_gap) {
    __shared__ int wl[4], wg[4];
    int lv = 0, gp = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nin; i += gridDim.x * 256) {
        int li = 0;
        core::make_leaf(kw, L, Lk, base, c, i, o, &li);
        // particle i starts internal nodes base[i] .. base[i] + c[i] - 1
        const int b0 = base[i], ci = c[i];
        for (int k = 0; k < ci; k++) own[b0 + k] = i;
        lv = li > lv ? li : lv;
        int g = gapl[idx[i]];
        gp += (g > 0 && g <= li) ? 1 : 0;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        int t = __shfl_xor(lv, d, 64);
        lv = t > lv ? t : lv;
        gp += __shfl_xor(gp, d, 64);
    }
    if (lane_id() == 0) { wl[threadIdx.x >> 6] = lv; wg[threadIdx.x >> 6] = gp; }
    __syncthreads();
    if (threadIdx.x == 0) {
        int bl = wl[0], bg = wg[0];
        for (int w = 1; w < 4; w++) { bl = wl[w] > bl ? wl[w] : bl; bg += wg[w]; }
        part_lv[blockIdx.x] = (unsigned long long)bl;
        part_gap[blockIdx.x] = (unsigned long long)bg;
    }
}

// Deepest leaf level and gap count from make_leaf's per-block partials
// (one workgroup of 256; run by the first workgroup of k_child_count).
__device__ inline void leaf_fold(int nb, const unsigned long long *part_lv, const unsigned long long *part_gap,
                                 DevState *ds) {
    __shared__ unsigned long long wl[4], wg[4];
    unsigned long long lv = 0ull, gp = 0ull;
    for (int k = threadIdx.x; k < nb; k += 256) {
        lv = part_lv[k] > lv ? part_lv[k] : lv;
        gp += part_gap[k];
    }
    for (int o = 32; o >= 1; o >>= 1) {
        unsigned long long q = __shfl_xor(lv, o, 64);
        lv = q > lv ? q : lv;
        gp += __shfl_xor(gp, o, 64);
    }
    if (lane_id() == 0) { wl[threadIdx.x >> 6] = lv; wg[threadIdx.x >> 6] = gp; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long l = wl[0], g = wg[0];
        for (int w = 1; w < 4; w++) { l = wl[w] > l ? wl[w] : l; g += wg[w]; }
        ds->max_level = (int32_t)l;
        // stack growth bound: at most 7 blocks per level below
        ds->walk_depth = 7 * ((int32_t)l + 3);
        ds->gap = g;
    }
}

// ------------------------------------------------------------ moments
// Node sums (mass, mass-weighted position, gas mass, gas-mass-weighted
// velocity) straight from each internal node's particle range in key order,
// in a fixed order that depends on the range alone, so they are
// deterministic and round relative to the node's own terms wherever the
// rest of the system lies.  No node waits for another: a bottom-up pass
// would chain the deepest leaf's latency through every level.
//
// k_tile_sums forms the sums of aligned blocks of kBlk1 = 64 particles (a
// butterfly over one wave) and of kBlk2 = 4096 (a butterfly over its 64
// blocks).  A node of up to kSumSmall particles is summed by one lane in
// index order from particles its workgroup staged in LDS; a larger one by a
// wave, over the greedy ascending cover of its range by particles, 64-blocks
// and 4096-blocks (lane j taking pieces j, j + 64, ...; then a butterfly),
// the parallel root adding the particles outside its cube (Node.cpp:477-499)
// as further pieces.
constexpr int kSumSmall = 32;
constexpr int kBlk1 = 64;
constexpr int kBlk2 = 4096;
constexpr int kStage = 256 + kSumSmall;   // particles a workgroup of 256 nodes stages

struct PVals {
    const double *sx, *sy, *sz, *sm, *svx, *svy, *svz;
    const uint8_t *stype;
};

__device__ inline void add_particle(NodeSums &s, int p, const PVals &v) {
    const double m = v.sm[p];
    s.m += m;
    s.sx += v.sx[p] * m;
    s.sy += v.sy[p] * m;
    s.sz += v.sz[p] * m;
    if (v.stype[p] == 2) {
        s.g += m;
        s.gx += v.svx[p] * m;
        s.gy += v.svy[p] * m;
        s.gz += v.svz[p] * m;
    }
}

__device__ inline void add_sums(NodeSums &s, const NodeSums &t) {
    s.m += t.m; s.sx += t.sx; s.sy += t.sy; s.sz += t.sz;
    s.g += t.g; s.gx += t.gx; s.gy += t.gy; s.gz += t.gz;
}

__device__ inline NodeSums shfl_sums(const NodeSums &s, int d) {
    return NodeSums{__shfl_xor(s.m, d, 64), __shfl_xor(s.sx, d, 64), __shfl_xor(s.sy, d, 64),
                    __shfl_xor(s.sz, d, 64), __shfl_xor(s.g, d, 64), __shfl_xor(s.gx, d, 64),
                    __shfl_xor(s.gy, d, 64), __shfl_xor(s.gz, d, 64)};
}

// Butterfly over the wave; every lane ends with a sum, lane 0's is the one used.
__device__ inline NodeSums wave_sums(NodeSums s) {
    for (int d = 32; d >= 1; d >>= 1) add_sums(s, shfl_sums(s, d));
    return s;
}

// One workgroup of 16 waves per 4096 particles: wave w forms blocks 4 w ..
// 4 w + 3 of 64, then wave 0 the 4096-block from the 64.  Blocks past nin
// hold zeros.  The grid covers all n particles: the launch does not wait
// for the host to learn nin.
__global__ __launch_bounds__(1024) void k_tile_sums(const DevState *ds, PVals v, NodeSums *blk1, NodeSums *blk2) {
    __shared__ NodeSums sb[64];
    const int nin = ds->nin;
    const int w = threadIdx.x >> 6, lane = lane_id();
    const int b0 = blockIdx.x * 64 + w * 4;   // first 64-block of this wave
    NodeSums t[4];
    for (int u = 0; u < 4; u++) {
        t[u] = NodeSums{0, 0, 0, 0, 0, 0, 0, 0};
        const int p = (b0 + u) * kBlk1 + lane;
        if (p < nin) add_particle(t[u], p, v);
    }
    for (int u = 0; u < 4; u++) {
        const NodeSums r = wave_sums(t[u]);
        if (lane == 0) {
            blk1[b0 + u] = r;
            sb[w * 4 + u] = r;
        }
    }
    __syncthreads();
    if (w == 0) {
        const NodeSums r = wave_sums(sb[lane]);
        if (lane == 0) blk2[blockIdx.x] = r;
    }
}

// The greedy ascending cover of [s, e): head particles up to a 64-boundary,
// 64-blocks up to a 4096-boundary, 4096-blocks, 64-blocks, tail particles;
// then, for the parallel root, the particles [nin, n).
struct Cover {
    int s, h1, t2, t1, tail, p1, p2, p3, p4, nout, nin;
    int k1, k2, k3, k4, k5, total;
};

__device__ inline Cover make_cover(int s, int e, int nin, int nout) {
    Cover c;
    c.s = s;
    c.nin = nin;
    int a = (s + kBlk1 - 1) / kBlk1 * kBlk1;
    if (a > e) a = e;
    const int h0 = a - s;
    c.p1 = a;
    const int e1 = e / kBlk1 * kBlk1, e2 = e / kBlk2 * kBlk2;
    int b = (a + kBlk2 - 1) / kBlk2 * kBlk2;
    if (b > e1) b = e1 > a ? e1 : a;
    c.h1 = (b - a) / kBlk1;
    c.p2 = a + c.h1 * kBlk1;
    c.t2 = (c.p2 % kBlk2 == 0 && e2 > c.p2) ? (e2 - c.p2) / kBlk2 : 0;
    c.p3 = c.p2 + c.t2 * kBlk2;
    c.t1 = e1 > c.p3 ? (e1 - c.p3) / kBlk1 : 0;
    c.p4 = c.p3 + c.t1 * kBlk1;
    c.tail = e - c.p4;
    c.nout = nout;
    c.k1 = h0;
    c.k2 = c.k1 + c.h1;
    c.k3 = c.k2 + c.t2;
    c.k4 = c.k3 + c.t1;
    c.k5 = c.k4 + c.tail;
    c.total = c.k5 + nout;
    return c;
}

__device__ inline void add_piece(NodeSums &acc, const Cover &c, int k, const PVals &v, const NodeSums *blk1,
                                 const NodeSums *blk2) {
    if (k < c.k1) add_particle(acc, c.s + k, v);
    else if (k < c.k2) add_sums(acc, blk1[c.p1 / kBlk1 + (k - c.k1)]);
    else if (k < c.k3) add_sums(acc, blk2[c.p2 / kBlk2 + (k - c.k2)]);
    else if (k < c.k4) add_sums(acc, blk1[c.p3 / kBlk1 + (k - c.k3)]);
    else if (k < c.k5) add_particle(acc, c.p4 + (k - c.k4), v);
    else add_particle(acc, c.nin + (k - c.k5), v);
}

struct StagedP {
    double m, x, y, z, vx, vy, vz;
    int gas;
};

__global__ __launch_bounds__(256) void k_sums_direct(int nn, int nin, int n, int proot, const int32_t *first,
                                                     const int32_t *count, const int32_t *pidx, PVals v,
                                                     const NodeSums *blk1, const NodeSums *blk2,
                                                     NodeSums *nsum) {
    __shared__ double st_m[kStage], st_x[kStage], st_y[kStage], st_z[kStage];
    __shared__ double st_vx[kStage], st_vy[kStage], st_vz[kStage];
    __shared__ uint8_t st_g[kStage];
    // nodes in XCD ranges: neighbours in pre-order share particles
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int nd0 = b * 256;
    const int last = nd0 + 255 < nn ? nd0 + 255 : nn - 1;
    // every particle between the first particles of two nodes has its leaf
    // between them in pre-order, so the span is at most 255 + kSumSmall
    const int P0 = first[nd0];
    int P1 = first[last] + kSumSmall;
    P1 = P1 < nin ? P1 : nin;
    for (int k = threadIdx.x; k < P1 - P0; k += 256) {
        const int p = P0 + k;
        const double m = v.sm[p];
        st_m[k] = m; st_x[k] = v.sx[p]; st_y[k] = v.sy[p]; st_z[k] = v.sz[p];
        const bool g = v.stype[p] == 2;
        st_g[k] = g ? 1 : 0;
        if (g) { st_vx[k] = v.svx[p]; st_vy[k] = v.svy[p]; st_vz[k] = v.svz[p]; }
    }
    __syncthreads();
    const int nd = nd0 + (int)threadIdx.x;
    int s = 0, c = 0;
    bool small = false, wide = false;
    if (nd < nn && pidx[nd] < 0) {
        s = first[nd];
        c = count[nd];
        wide = (nd == 0 && proot) || c > kSumSmall;
        small = !wide;
    }
    if (small) {
        NodeSums t{0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = s - P0; k < s - P0 + c; k++) {
            const double m = st_m[k];
            t.m += m;
            t.sx += st_x[k] * m;
            t.sy += st_y[k] * m;
            t.sz += st_z[k] * m;
            if (st_g[k]) {
                t.g += m;
                t.gx += st_vx[k] * m;
                t.gy += st_vy[k] * m;
                t.gz += st_vz[k] * m;
            }
        }
        nsum[nd] = t;
    }
    unsigned long long wb = __ballot(wide);
    while (wb) {
        const int l = __builtin_ctzll(wb);
        wb &= wb - 1;
        const int ws = __shfl(s, l, 64), wc = __shfl(c, l, 64), wn = __shfl(nd, l, 64);
        const Cover cv = make_cover(ws, ws + wc, nin, (wn == 0 && proot) ? n - nin : 0);
        NodeSums t{0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = lane_id(); k < cv.total; k += 64) add_piece(t, cv, k, v, blk1, blk2);
        t = wave_sums(t);
        if (lane_id() == 0) nsum[wn] = t;
    }
}

// A tree with no particle inside the root cube: the parallel root's sums are
// those of the particles outside it.
__global__ void k_moments_root_only(int n, PVals v, NodeSums *nsum) {
    NodeSums s{0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < n; q++) add_particle(s, q, v);
    nsum[0] = s;
}

struct NodeMoments {
    double *m, *gas, *cx, *cy, *cz, *mvx, *mvy, *mvz;
};

// Child-contiguous layout for the block walk (core::TreeNode): the root at 0,
// the children of node p at [1 + cstart[p], 1 + cstart[p] + nchild[p]) in
// octant order, cstart being the exclusive scan of nchild in pre-order.  Every node
// follows its children's skip links (pull, no atomics): it counts them and
// gives each its rank among its siblings, which in pre-order is its octant
// order.  A bucket (particles at one position) holds its member leaves, the
// next count nodes, as children.
__global__ __launch_bounds__(256) void k_child_count(int nn, const int32_t *pidx, const int8_t *level,
                                                     const int32_t *count, const int32_t *skip, int32_t *nchild,
                                                     int32_t *rank, int32_t *ichild, int nleaf_blocks,
                                                     const unsigned long long *part_lv,
                                                     const unsigned long long *part_gap, DevState *ds,
                                                     int32_t *bsum) {
    __shared__ int scr[4];
    if (blockIdx.x == 0) leaf_fold(nleaf_blocks, part_lv, part_gap, ds);
    // four nodes per thread (a 1024-node scan tile per workgroup); the tile's
    // sum of nchild goes to bsum for k_iscan_apply
    // The four nodes' child chains are followed side by side, so their skip
    // loads are in flight together.  A bucket's members are consecutive
    // leaves: the same loop ranks them.
    int ch[4], end[4], nc[4], ni[4];
    for (int k = 0; k < 4; k++) {
        const int p = blockIdx.x * 1024 + k * 256 + threadIdx.x;
        nc[k] = ni[k] = 0;
        ch[k] = p + 1;
        end[k] = (p < nn && pidx[p] < 0) ? skip[p] : 0;
    }
    while (ch[0] < end[0] || ch[1] < end[1] || ch[2] < end[2] || ch[3] < end[3]) {
        for (int k = 0; k < 4; k++) {
            if (ch[k] < end[k]) {
                rank[ch[k]] = nc[k]++;
                ni[k] += pidx[ch[k]] < 0 ? 1 : 0;
                ch[k] = skip[ch[k]];
            }
        }
    }
    int tile = 0;
    for (int k = 0; k < 4; k++) {
        const int p = blockIdx.x * 1024 + k * 256 + threadIdx.x;
        if (p >= nn) continue;
        nchild[p] = nc[k];
        ichild[p] = ni[k];
        tile += nc[k];
    }
    int tot;
    block256_excl_scan(tile, scr, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Every node at once: sums over its key range (the parallel root also
// counts the particles outside its cube, Node.cpp:477-499), centre of mass,
// gas-weighted velocity, and the walk record at the node's layout place.
//
// Zero masses (zm, set when any particle has m == 0): the serial insert's
// running centre of mass (Node.cpp:692-697) divides by the node's mass after
// each particle; when the first particle to reach a node (the lowest caller
// index in it, since particles are inserted in array order) has m == 0, that
// is 0/0 and the node's centre stays NaN for good.  Such a node is never
// accepted (radius / NaN < theta is false), so the walk opens it.
__global__ __launch_bounds__(256) void k_node_moments(int nn, int n, int proot, int64_t thr, int zm,
                                                      const uint32_t *rid, const DevState *ds,
                                                      const int32_t *first, const int32_t *count,
                                                      const int32_t *parent, const int32_t *pidx,
                                                      const int8_t *level, const NodeSums *nsum,
                                                      const double *gas_table, int gas_equal, int64_t ngas, PVals v,
                                                      const int32_t *nchild, const int32_t *rank,
                                                      const int32_t *cstart, NodeMoments o,
                                                      core::TreeNode *tree, core::SphNode *tsph,
                                                      int32_t *group_first, uint8_t *group_fired) {
    // nodes in XCD ranges: neighbours in pre-order read neighbouring prefix
    // sums and particles
    int nd = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (nd >= nn) return;
    // cleared group state for the density phase (Node.cpp:722-796)
    group_first[nd] = kGroupNone;
    group_fired[nd] = 0;
    int p = pidx[nd];
    int par = parent[nd];
    double M, g, cx, cy, cz, mvx = 0.0, mvy = 0.0, mvz = 0.0;
    const core::Pos pos{v.sx, v.sy, v.sz};
    // buckets exist only where k_lcp met adjacent particles at one position
    const bool buckets = ds->any_bucket != 0;
    if (p >= 0) {
        M = v.sm[p];
        g = v.stype[p] == 2 ? M : 0.0;
        bool vec = par < 0 ? (n == 1) : core::parallel_node(count, parent, par, n, thr);
        if (par >= 0 && !vec && buckets) {
            // members of a bucket sit exactly at their position
            const int ps = (par == 0 && proot) ? 0 : first[par];
            const int pe = (par == 0 && proot) ? n : first[par] + count[par];
            vec = core::is_bucket(pos, ps, pe);
        }
        cx = core::leaf_com(v.sx[p], M, vec);
        cy = core::leaf_com(v.sy[p], M, vec);
        cz = core::leaf_com(v.sz[p], M, vec);
    } else {
        int s = first[nd], e = first[nd] + count[nd];
        if (nd == 0 && proot) { s = 0; e = n; }
        const NodeSums ns = nsum[nd];
        const double S[8] = {ns.m, ns.sx, ns.sy, ns.sz, ns.g, ns.gx, ns.gy, ns.gz};
        M = S[0];
        // equal gas masses: the reference's running sum, bit for bit (its
        // group selection compares gas masses and can sit on exact ties).
        // The node's gas count is its exactly rounded gas mass over the
        // common mass, rounded to the nearest integer (exact below 2^51).
        g = S[4];
        if (gas_equal && g > 0.0) {
            double q = rint(g / gas_table[1]);
            int64_t k = q < 0.0 ? 0 : (q > (double)ngas ? ngas : (int64_t)q);
            g = gas_table[k];
        }
        if (M > 0) { cx = S[1] / M; cy = S[2] / M; cz = S[3] / M; }
        else { cx = cy = cz = 0.0; }
        if (zm && !(nd == 0 && proot) && !core::parallel_node(count, parent, nd, n, thr)) {
            uint32_t first_any = 0xffffffffu, first_pos = 0xffffffffu;
            for (int j = first[nd]; j < first[nd] + count[nd]; j++) {
                const uint32_t q = rid[j];
                first_any = q < first_any ? q : first_any;
                if (v.sm[j] != 0.0) first_pos = q < first_pos ? q : first_pos;
            }
            if (first_any < first_pos) cx = cy = cz = __builtin_nan("");
        }
        if (g > 0) { mvx = S[5] / g; mvy = S[6] / g; mvz = S[7] / g; }
        if (buckets && core::is_bucket(pos, s, e)) { cx = v.sx[s]; cy = v.sy[s]; cz = v.sz[s]; }
    }
    o.m[nd] = M; o.gas[nd] = g;
    o.cx[nd] = cx; o.cy[nd] = cy; o.cz[nd] = cz;
    o.mvx[nd] = mvx; o.mvy[nd] = mvy; o.mvz[nd] = mvz;
    const int place = par < 0 ? 0 : 1 + cstart[par] + rank[nd];
    const int kc = nchild[nd];
    core::TreeNode t;
    t.cx = cx; t.cy = cy; t.cz = cz; t.mass = M; t.gas = g;
    t.first_child = kc > 0 ? 1 + cstart[nd] : -1;
    t.nchild = kc;
    t.pidx = p;
    t.pre = nd;
    t.level = level[nd];
    t.flags = 0;
    t.pad0 = 0;
    t.pad1 = 0;
    tree[place] = t;
    if (tsph && g > 0.0) {   // only gas nodes are ever SPH candidates
        core::SphNode w;
        w.cx = cx; w.cy = cy; w.cz = cz; w.gas = g;
        if (p >= 0) { w.vx = v.svx[p]; w.vy = v.svy[p]; w.vz = v.svz[p]; }
        else { w.vx = mvx; w.vy = mvy; w.vz = mvz; }
        w.pidx = p;
        w.pad = 0;
        tsph[place] = w;
    }
}

// Exclusive scan of an int array of host-known length (tiles of 1024).
__global__ __launch_bounds__(256) void k_iscan_tiles(int n, const int32_t *in, int32_t *bsum) {
    __shared__ int scr[4];
    int b0 = blockIdx.x * 1024 + threadIdx.x * 4;
    int s = 0;
    for (int k = 0; k < 4; k++) s += (b0 + k < n) ? in[b0 + k] : 0;
    int tot;
    block256_excl_scan(s, scr, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_iscan_apply(int n, const int32_t *in, const int32_t *bsum, int32_t *out) {
    __shared__ int scr[4];
    int b0 = blockIdx.x * 1024 + threadIdx.x * 4;
    int v[4];
    int s = 0;
    for (int k = 0; k < 4; k++) { v[k] = (b0 + k < n) ? in[b0 + k] : 0; s += v[k]; }
    const int pre = tile_offset(bsum, (int)gridDim.x, (int)blockIdx.x, scr, nullptr);
    int ex = block256_excl_scan(s, scr, nullptr) + pre;
    for (int k = 0; k < 4; k++) {
        if (b0 + k < n) out[b0 + k] = ex;
        ex += v[k];
    }
}

// Pre-order walk records (core::WalkNode) for the one-lane-per-target
// potential walk; built only when the potential is asked for.
__global__ __launch_bounds__(256) void k_pack_walknode(int nn, const DevState *ds, const int32_t *skip,
                                                       const int32_t *pidx, const int8_t *level, const double *m,
                                                       const double *gas, const double *cx, const double *cy,
                                                       const double *cz, WalkNode *wn) {
    int nd = blockIdx.x * 256 + threadIdx.x;
    if (nd >= nn) return;
    double R = __longlong_as_double((long long)ds->Rbits);
    WalkNode w;
    w.cx = cx[nd]; w.cy = cy[nd]; w.cz = cz[nd]; w.mass = m[nd];
    w.radius = ldexp(R, -(int)level[nd]);
    w.gas = gas[nd];
    w.skip = skip[nd]; w.pidx = pidx[nd]; w.level = level[nd]; w.flags = 0;
    wn[nd] = w;
}

// ------------------------------------------------------------ deep refinement
// Distinct particles that agree on every key word so far, and tie runs too
// long for k_tie_rank, are refined word by word: the particles tied on words
// 0..w-1 get word w and each tie run is sorted by (word w, caller index)
// with a general segmented LSD radix sort over just those particles
// (compacted; run start, word and caller index as three stable sorts).
// Runs sit at contiguous positions, so the k-th particle of the sorted
// compact list takes the k-th compact position.  All of it works on the
// order before the state permutation (ord = the sort permutation).
__device__ inline bool words_equal(const core::KeyWords &kw, int a, int b, int w) {
    for (int j = 0; j < w; j++)
        if (kw.w[j][a] != kw.w[j][b]) return false;
    return true;
}

__device__ inline int tie_first(const core::KeyWords &kw, int w, int i) {
    int hi = i, lo = -1;
    for (int step = 1;; step <<= 1) {
        const int probe = i - step;
        if (probe < 0) break;
        if (!words_equal(kw, probe, i, w)) { lo = probe; break; }
        hi = probe;
        if (step > (1 << 29)) break;
    }
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if (words_equal(kw, mid, i, w)) hi = mid; else lo = mid;
    }
    return hi;
}

// flag[i] = 1 if particle i shares words 0..w-1 with a neighbour
__global__ __launch_bounds__(256) void k_deep_flag(const DevState *ds, core::KeyWords kw, int w, int32_t *flag) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int nin = ds->nin;
    if (i >= nin) return;
    const bool t = (i > 0 && words_equal(kw, i - 1, i, w)) || (i + 1 < nin && words_equal(kw, i, i + 1, w));
    flag[i] = t ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_deep_compact(const DevState *ds, const int32_t *flag, const int32_t *offs,
                                                      uint32_t *spos, int32_t *m_out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int nin = ds->nin;
    if (i >= nin) return;
    if (flag[i]) spos[offs[i]] = (uint32_t)i;
    if (i == nin - 1) *m_out = offs[i] + flag[i];
}

// Word w of every flagged particle (w >= 2; word 1 exists already), its
// run start, and the first sort key (caller index).
__global__ __launch_bounds__(256) void k_deep_word(int m, const uint32_t *spos, const uint32_t *ord,
                                                   const double *x, const double *y, const double *z,
                                                   const uint32_t *rid, const DevState *ds, core::KeyWords kw,
                                                   int w, uint64_t *word_w, uint8_t *gapl, uint32_t *srun,
                                                   uint64_t *sk, uint32_t *sv) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= m) return;
    const int i = (int)spos[k];
    const uint32_t q = ord[i];
    if (w >= 2) {
        const double R = __longlong_as_double((long long)ds->Rbits);
        int g = 0;
        word_w[i] = core::octant_word(x[q], y[q], z[q], R, w, &g);
        if (gapl[q] == 0 && g > 0) gapl[q] = (uint8_t)(g > 254 ? 254 : g);
    }
    srun[k] = (uint32_t)tie_first(kw, w, i);
    sk[k] = (uint64_t)rid[q];
    sv[k] = (uint32_t)k;
}

// sk[j] = word w of the compact element sv[j] (second sort key)
__global__ __launch_bounds__(256) void k_deep_key_word(int m, const uint32_t *sv, const uint32_t *spos,
                                                       const uint64_t *word_w, uint64_t *sk) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < m) sk[j] = word_w[spos[sv[j]]];
}

// sk[j] = run start of the compact element sv[j] (third sort key)
__global__ __launch_bounds__(256) void k_deep_key_run(int m, const uint32_t *sv, const uint32_t *srun, uint64_t *sk) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < m) sk[j] = (uint64_t)srun[sv[j]];
}

// The j-th sorted element moves to the j-th compact position.
__global__ __launch_bounds__(256) void k_deep_scatter(int m, const uint32_t *spos, const uint32_t *sv,
                                                      const uint32_t *ord, const uint64_t *word_w, uint32_t *ord_tmp,
                                                      uint64_t *word_tmp) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m) return;
    const uint32_t dst = spos[j], src = spos[sv[j]];
    ord_tmp[dst] = ord[src];
    word_tmp[dst] = word_w[src];
}

__global__ __launch_bounds__(256) void k_deep_back(int m, const uint32_t *spos, const uint32_t *ord_tmp,
                                                   const uint64_t *word_tmp, uint32_t *ord, uint64_t *word_w) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m) return;
    const uint32_t p = spos[j];
    ord[p] = ord_tmp[p];
    word_w[p] = word_tmp[p];
}

__global__ void k_reset_tie_flags(DevState *ds) {
    ds->tie_big = 0;
    ds->tie_deep = 0;
}

__global__ void k_reset_walk_depth(DevState *ds) { ds->walk_depth = 0; }

// Stack growth bound of the block walk in a deep tree.  Taking one block of
// siblings off the stack pushes one block per sibling opened, and only
// internal nodes are opened, so the stack grows by at most (internal
// children - 1) at a node; along any root-to-leaf path the sum of those is
// the most the stack can grow below a point (the walk adds 8 of slack).
// Chains of single-internal-child cells (a galaxy far down a root cube
// stretched by escaped particles, with a stray particle beside each level)
// add nothing.
__global__ __launch_bounds__(256) void k_branch_depth(int nin, const int32_t *base, const int32_t *c,
                                                      const int32_t *parent, const int32_t *ichild, DevState *ds) {
    __shared__ int wmax[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    int d = 0;
    if (i < nin) {
        const int leaf = i + base[i] + c[i];
        for (int a = parent[leaf]; a >= 0; a = parent[a]) d += ichild[a] > 1 ? ichild[a] - 1 : 0;
        d += 8;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const int t = __shfl_xor(d, o, 64);
        d = t > d ? t : d;
    }
    if (lane_id() == 0) wmax[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        int b = wmax[0];
        for (int k = 1; k < 4; k++) b = wmax[k] > b ? wmax[k] : b;
        atomicMax(&ds->walk_depth, b);
    }
}

}  // namespace

void alloc_nodes(Ctx &c, int64_t cap) {
    reset_graphs(c);   // the node capacity is an argument of the captured segments
    if (c.node_slab) {
        (void)hipFree(c.node_slab);
        c.node_slab = nullptr;
    }
    c.node_cap = cap;
    const size_t n = (size_t)cap;
    // block sums for every particle slot (launch_tile_sums covers all c.n <= c.cap)
    const size_t pcap = (size_t)c.cap > n ? (size_t)c.cap : n;
    auto plan = [&](Carve &cv) {
        cv(&c.n_first, n * 4); cv(&c.n_count, n * 4); cv(&c.n_skip, n * 4);
        cv(&c.n_parent, n * 4);
        cv(&c.n_pidx, n * 4); cv(&c.n_level, n);
        double **d8[] = {&c.n_mass, &c.n_gas, &c.n_gvx, &c.n_gvy, &c.n_gvz, &c.n_cx, &c.n_cy, &c.n_cz, &c.g_rho};
        for (double **p : d8) cv(p, n * 8);
        cv(&c.wnode, n * sizeof(WalkNode));
        cv(&c.g_first, n * 4); cv(&c.g_list, (n + 256) * 4);
        cv(&c.g_fired, n); cv(&c.g_final, n * 4);
        cv(&c.tnode, n * sizeof(core::TreeNode));
        cv(&c.tsph, n * sizeof(core::SphNode));
        cv(&c.n_nchild, n * 4); cv(&c.n_rank, n * 4); cv(&c.n_cstart, n * 4);
        cv(&c.n_ichild, n * 4);
        cv(&c.nsum_tile, (pcap / kBlk2 + 1) * (kBlk2 / kBlk1 + 1) * sizeof(NodeSums));
        cv(&c.nsum, n * sizeof(NodeSums));
        cv(&c.n_bsum, (n / 1024 + 2) * 4);
    };
    Carve size;
    plan(size);
    AG2_HIP(hipMalloc(&c.node_slab, size.off));
    Carve at;
    at.base = (char *)c.node_slab;
    plan(at);
}

void state_fields(Ctx &c, double **f[kStateD]) {
    double **l[kStateD] = {&c.x, &c.y, &c.z, &c.vx, &c.vy, &c.vz, &c.m, &c.U, &c.mu, &c.ax,
                           &c.ay, &c.az, &c.h, &c.rho, &c.P, &c.T, &c.dUdt, &c.ts, &c.next};
    static_assert(kAccField == 9 && kHField == 12, "field order: accelerations after nine fields, then h");
    for (int j = 0; j < kStateD; j++) f[j] = l[j];
}

// The permutation into key order writes the alternate buffers; swap_state
// then makes them current on the host side (also when the launch came from
// a replayed graph).
static void launch_permute(Ctx &c, int n) {
    double **f[kStateD];
    state_fields(c, f);
    StateSet in, out;
    for (int j = 0; j < kStateD; j++) { in.d[j] = *f[j]; out.d[j] = c.alt[j]; }
    in.type = c.type; in.active = c.active; in.rid = c.rid;
    out.type = c.type_alt; out.active = c.active_alt; out.rid = c.rid_alt;
    // idx_alt: the sort permutation after the tie runs' order (k_tie_rank).
    // A captured segment keeps its arguments for every replay, so graph
    // launches always move the accelerations.
    const int skip_acc = c.acc_overwritten && !c.p.graph_launch ? 1 : 0;
    const int zero_gas_h = c.density_follows && !c.p.graph_launch ? 1 : 0;
    c.h_cleared = zero_gas_h != 0;
    hipLaunchKernelGGL(k_permute, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, n, c.idx_alt, in, out, skip_acc,
                       zero_gas_h);
}

static void swap_state(Ctx &c) {
    double **f[kStateD];
    state_fields(c, f);
    for (int j = 0; j < kStateD; j++) std::swap(*f[j], c.alt[j]);
    std::swap(c.type, c.type_alt);
    std::swap(c.active, c.active_alt);
    std::swap(c.rid, c.rid_alt);
    c.parity ^= 1;
}

void build_inverse(Ctx &c) {
    hipLaunchKernelGGL(k_inverse, dim3(grid_for(c.n, 256)), dim3(256), 0, c.stream, (int)c.n, c.rid, c.inv);
    AG2_HIP(hipGetLastError());
}

// LSD radix passes over (key, value) pairs, 8-bit digits from bit shift0:
// an even number of passes leaves the result in kin / vin.
static void radix_pairs(Ctx &c, int n, uint64_t *kin, uint64_t *kout, uint32_t *vin, uint32_t *vout, int shift0,
                        int passes) {
    const int nblk = (n + kSortTile - 1) / kSortTile;
    if (nblk < 1) return;
    const bool small = nblk <= kSortSmallTiles;
    for (int pass = 0; pass < passes; pass++) {
        const int shift = shift0 + pass * 8;
        hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(256), 0, c.stream, n, kin, shift, c.hist, nblk);
        if (small) {
            hipLaunchKernelGGL(k_rs_scatter<true>, dim3(nblk), dim3(256), 0, c.stream, n, kin, vin, kout, vout, shift,
                               c.hist, c.dtotal, nblk);
        } else {
            hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(256), 0, c.stream, nblk, c.hist, c.dtotal);
            hipLaunchKernelGGL(k_rs_scatter<false>, dim3(nblk), dim3(256), 0, c.stream, n, kin, vin, kout, vout, shift,
                               c.hist, c.dtotal, nblk);
        }
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
}

static void radix_sort(Ctx &c, int n) {
    int nblk = (n + kSortTile - 1) / kSortTile;
    uint64_t *kin = c.key, *kout = c.key_alt;
    uint32_t *vin = c.idx, *vout = c.idx_alt;
    const bool small = nblk <= kSortSmallTiles;
    // bytes sort_lo .. 7 (an even number of passes)
    for (int pass = c.sort_lo; pass < 8; pass++) {
        int shift = pass * 8;
        hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(256), 0, c.stream, n, kin, shift, c.hist, nblk);
        if (small) {
            hipLaunchKernelGGL(k_rs_scatter<true>, dim3(nblk), dim3(256), 0, c.stream, n, kin, vin, kout, vout, shift,
                               c.hist, c.dtotal, nblk);
        } else {
            hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(256), 0, c.stream, nblk, c.hist, c.dtotal);
            hipLaunchKernelGGL(k_rs_scatter<false>, dim3(nblk), dim3(256), 0, c.stream, n, kin, vin, kout, vout, shift,
                               c.hist, c.dtotal, nblk);
        }
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // an even number of passes: the result is back in c.key / c.idx
    static_assert(kSortLoMax % 2 == 0, "the sort must end in its input buffers");
}

// First half of the build in three launch segments.  Their kernel
// arguments depend only on the particle count, the node capacity, the
// parameters and the state parity, so with ag2_params.graph_launch each
// segment is captured once per parity as a HIP graph and replayed; by
// default every kernel is launched on its own.  No segment holds a memset.
static void seg_keys(Ctx &c) {
    int n = (int)c.n;
    hipStream_t st = c.stream;
    int g = grid_for(n, 256);
    int twb = g < kTwBlocks ? (g > 0 ? g : 1) : kTwBlocks;
    // partials: the sums at [0, 2 twb), the maxima after them
    unsigned long long *rmax = (unsigned long long *)(c.partials + 2 * kTwBlocks);
    hipLaunchKernelGGL(k_tw_partial, dim3(twb), dim3(256), 0, st, n, c.x, c.y, c.z, c.partials);
    hipLaunchKernelGGL(k_tw_max, dim3(twb), dim3(256), 0, st, n, c.x, c.y, c.z, twb, c.partials, c.ds, rmax);
    hipLaunchKernelGGL(k_keygen, dim3(g), dim3(256), 0, st, n, c.x, c.y, c.z, twb, rmax, c.ds, c.key, c.idx,
                       c.gapl);
}

static void seg_sort(Ctx &c) { radix_sort(c, (int)c.n); }

static void seg_order(Ctx &c) {
    int n = (int)c.n;
    hipStream_t st = c.stream;
    int g = grid_for(n, 256);
    const int shift = 8 * c.sort_lo;
    hipLaunchKernelGGL(k_tie_key2, dim3(g), dim3(256), 0, st, n, c.key, c.idx, c.x, c.y, c.z, c.ds, c.key2, c.gapl,
                       shift);
    // word 1 in run order -> key_alt, the permutation -> idx_alt
    hipLaunchKernelGGL(k_tie_rank, dim3(g), dim3(256), 0, st, n, c.key, c.key2, c.idx, c.rid, c.ds, c.w0, c.key_alt,
                       c.idx_alt, shift);
    launch_permute(c, n);
    int proot = (c.n >= 2 && c.n >= c.p.split_threshold) ? 1 : 0;
    // from here on the state is the permuted one (alternate buffers)
    core::KeyWords kw{{c.w0, c.key_alt, nullptr, nullptr, nullptr, nullptr}};
    core::Pos pos{c.alt[0], c.alt[1], c.alt[2]};
    int nb = (n + 1023) / 1024;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_lcp, dim3(nb), dim3(256), 0, st, n, kw, 2, pos, c.ds, proot, c.lcp, c.lk, c.cnt, c.bsum);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, st, c.cnt, c.bsum, c.ds, proot, c.node_cap, c.base);
    AG2_HIP(hipMemcpyAsync(c.hs_pinned, c.ds, sizeof(DevState), hipMemcpyDeviceToHost, st));
}

void reset_graphs(Ctx &c) {
    for (auto &row : c.seg_exec)
        for (auto &e : row)
            if (e) {
                (void)hipGraphExecDestroy(e);
                e = nullptr;
            }
}

static void run_segment(Ctx &c, int s, void (*fn)(Ctx &)) {
    if (!c.p.graph_launch) {
        fn(c);
        AG2_HIP(hipGetLastError());
        return;
    }
    hipGraphExec_t &ex = c.seg_exec[s][s == 1 ? 0 : c.parity];
    if (!ex) {
        hipGraph_t gr = nullptr;
        AG2_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        try {
            fn(c);
        } catch (...) {
            // leave the stream out of capture mode before the error goes up
            hipGraph_t dead = nullptr;
            (void)hipStreamEndCapture(c.stream, &dead);
            if (dead) (void)hipGraphDestroy(dead);
            throw;
        }
        hipError_t le = hipGetLastError();
        hipError_t ce = hipStreamEndCapture(c.stream, &gr);
        AG2_HIP(le);
        AG2_HIP(ce);
        hipError_t ie = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ie != hipSuccess) ex = nullptr;
        AG2_HIP(ie);
    }
    AG2_HIP(hipGraphLaunch(ex, c.stream));
}

// First half of the build, up to the node count, and the particle prefix
// sums: everything here is asynchronous; the state it leaves in c.ds (nin,
// n_int, n_nodes, err) is copied to c.hs_pinned, and c.ev_counts completes
// when it is there -- build_tree_b must not start before.
void build_keys(Ctx &c) {
    if (c.timing >= 2) AG2_HIP(hipEventRecord(c.ev[0], c.stream));
    run_segment(c, 0, seg_keys);
}

void build_tree_a(Ctx &c) {
    build_keys(c);
    if (c.timing >= 2) AG2_HIP(hipEventRecord(c.ev[1], c.stream));
    run_segment(c, 1, seg_sort);
    build_order(c);
}

size_t sort_job_bytes() { return sizeof(SortJob); }

void sort_batch(Ctx *const *cs, int nc, hipStream_t st, void *scratch_dev, void *scratch_host) {
    if (nc <= 0) return;
    SortJob *hj = (SortJob *)scratch_host;
    int total = 0;
    for (int k = 0; k < nc; k++) {
        Ctx &c = *cs[k];
        const int n = (int)c.n;
        const int nblk = (n + kSortTile - 1) / kSortTile;
        hj[k] = SortJob{n, nblk, total, 0, c.key, c.key_alt, c.idx, c.idx_alt, c.hist, c.dtotal};
        total += nblk;
    }
    if (total < 1) return;
    SortJob *dj = (SortJob *)scratch_dev;
    AG2_HIP(hipMemcpyAsync(dj, hj, sizeof(SortJob) * (size_t)nc, hipMemcpyHostToDevice, st));
    // the bytes every job sorts (an even number of passes ends in key / idx)
    int lo = kSortLoMax;
    for (int k = 0; k < nc; k++) lo = cs[k]->sort_lo < lo ? cs[k]->sort_lo : lo;
    for (int pass = lo; pass < 8; pass++) {
        hipLaunchKernelGGL(k_rsb_hist, dim3(total), dim3(256), 0, st, dj, nc, pass);
        hipLaunchKernelGGL(k_rsb_scan, dim3(256 * nc), dim3(256), 0, st, dj);
        hipLaunchKernelGGL(k_rsb_scatter, dim3(total), dim3(256), 0, st, dj, nc, pass);
    }
    // every job's result is back in its key / idx
    AG2_HIP(hipGetLastError());
}

// The block sums of the node-sum cover (k_tile_sums), laid out for all n
// particles: 64-blocks, then 4096-blocks.
static void tile_sum_buffers(const Ctx &c, NodeSums **blk1, NodeSums **blk2) {
    const int64_t ntile = (c.n + kBlk2 - 1) / kBlk2;
    *blk1 = c.nsum_tile;
    *blk2 = c.nsum_tile + (size_t)ntile * (kBlk2 / kBlk1);
}

static void launch_tile_sums(Ctx &c) {
    const int ntile = (int)((c.n + kBlk2 - 1) / kBlk2);
    if (ntile < 1) return;
    NodeSums *blk1, *blk2;
    tile_sum_buffers(c, &blk1, &blk2);
    PVals pv{c.x, c.y, c.z, c.m, c.vx, c.vy, c.vz, c.type};
    hipLaunchKernelGGL(k_tile_sums, dim3(ntile), dim3(1024), 0, c.stream, c.ds, pv, blk1, blk2);
}

void build_order(Ctx &c) {
    int n = (int)c.n;
    hipStream_t st = c.stream;
    if (c.timing >= 2) AG2_HIP(hipEventRecord(c.ev[2], st));
    // poisoned until this build's counts overwrite it (checked in build_tree_b)
    c.hs_pinned->n_nodes = -1;
    run_segment(c, 2, seg_order);
    swap_state(c);
    AG2_HIP(hipEventRecord(c.ev_counts, st));
    // the block sums need only the particles in key order: they run while
    // the host waits for the counts
    launch_tile_sums(c);
    (void)n;
}

core::KeyWords key_words(const Ctx &c) {
    core::KeyWords kw{{c.w0, c.key_alt, nullptr, nullptr, nullptr, nullptr}};
    for (int w = 2; w < c.kw_used; w++) kw.w[w] = c.kdeep[w - 2];
    return kw;
}

// Buffers of the deep refinement, allocated on its first use.
static void alloc_deep(Ctx &c) {
    if (c.sk) return;
    const size_t n = (size_t)c.cap;
    for (uint64_t *&p : c.kdeep) AG2_HIP(hipMalloc(&p, n * 8));
    AG2_HIP(hipMalloc(&c.sk, n * 8));
    AG2_HIP(hipMalloc(&c.sk_alt, n * 8));
    AG2_HIP(hipMalloc(&c.sv, n * 4));
    AG2_HIP(hipMalloc(&c.sv_alt, n * 4));
    AG2_HIP(hipMalloc(&c.spos, n * 4));
    AG2_HIP(hipMalloc(&c.srun, n * 4));
}


// The slow path of a build, taken only when the first half flagged a tie
// run longer than k_tie_rank's bound or distinct particles that agree on
// both key words (the reference splits such cells until they separate,
// Node.cpp:597-699).  The first half's permutation read the state without
// changing it, so the state goes back to its input buffers; the order is
// refined word by word (deep refinement kernels), and the state permuted,
// the topology counted and the prefix sums formed again.  Each word costs a
// few host reads; distinct particles still together after kDeepest levels
// fail the build.
static void refine_ties(Ctx &c) {
    hipStream_t st = c.stream;
    const int n = (int)c.n;
    const int nin = c.hs.nin;
    alloc_deep(c);
    swap_state(c);   // the state before this build's permutation
    const int g = grid_for(nin, 256);
    const int nb = (nin + 1023) / 1024;
    c.kw_used = 2;
    for (int w = c.hs.tie_big ? 1 : 2; w < core::kKeyWords; w++) {
        const core::KeyWords kw = key_words(c);   // words 0 .. w-1 are present
        hipLaunchKernelGGL(k_deep_flag, dim3(g), dim3(256), 0, st, c.ds, kw, w, c.cnt);
        hipLaunchKernelGGL(k_iscan_tiles, dim3(nb), dim3(256), 0, st, nin, c.cnt, c.bsum);
        hipLaunchKernelGGL(k_iscan_apply, dim3(nb), dim3(256), 0, st, nin, c.cnt, c.bsum, c.base);
        hipLaunchKernelGGL(k_deep_compact, dim3(g), dim3(256), 0, st, c.ds, c.cnt, c.base, c.spos, &c.ds->deep_m);
        AG2_HIP(hipMemcpyAsync(&c.hs_pinned->deep_m, &c.ds->deep_m, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        AG2_HIP(hipStreamSynchronize(st));
        const int m = c.hs_pinned->deep_m;
        if (m <= 0) break;
        uint64_t *word_w = w == 1 ? c.key_alt : c.kdeep[w - 2];
        const unsigned gm = grid_for(m, 256);
        hipLaunchKernelGGL(k_deep_word, dim3(gm), dim3(256), 0, st, m, c.spos, c.idx_alt, c.x, c.y, c.z, c.rid, c.ds,
                           kw, w, word_w, c.gapl, c.srun, c.sk, c.sv);
        radix_pairs(c, m, c.sk, c.sk_alt, c.sv, c.sv_alt, 0, 4);   // caller index
        hipLaunchKernelGGL(k_deep_key_word, dim3(gm), dim3(256), 0, st, m, c.sv, c.spos, word_w, c.sk);
        radix_pairs(c, m, c.sk, c.sk_alt, c.sv, c.sv_alt, 0, 8);   // word w
        hipLaunchKernelGGL(k_deep_key_run, dim3(gm), dim3(256), 0, st, m, c.sv, c.srun, c.sk);
        radix_pairs(c, m, c.sk, c.sk_alt, c.sv, c.sv_alt, 0, 4);   // run start
        hipLaunchKernelGGL(k_deep_scatter, dim3(gm), dim3(256), 0, st, m, c.spos, c.sv, c.idx_alt, word_w, c.idx,
                           c.key2);
        hipLaunchKernelGGL(k_deep_back, dim3(gm), dim3(256), 0, st, m, c.spos, c.idx, c.key2, c.idx_alt, word_w);
        c.kw_used = w + 1 > c.kw_used ? w + 1 : c.kw_used;
    }
    hipLaunchKernelGGL(k_reset_tie_flags, dim3(1), dim3(1), 0, st, c.ds);
    launch_permute(c, n);
    const int proot = (c.n >= 2 && c.n >= c.p.split_threshold) ? 1 : 0;
    const core::Pos pos{c.alt[0], c.alt[1], c.alt[2]};
    const int nbn = (n + 1023) / 1024 < 1 ? 1 : (n + 1023) / 1024;
    hipLaunchKernelGGL(k_lcp, dim3(nbn), dim3(256), 0, st, n, key_words(c), c.kw_used, pos, c.ds,
                       proot, c.lcp, c.lk, c.cnt, c.bsum);
    hipLaunchKernelGGL(k_scan_apply, dim3(nbn), dim3(256), 0, st, c.cnt, c.bsum, c.ds, proot, c.node_cap, c.base);
    swap_state(c);
    AG2_HIP(hipMemcpyAsync(c.hs_pinned, c.ds, sizeof(DevState), hipMemcpyDeviceToHost, st));
    AG2_HIP(hipStreamSynchronize(st));
    c.hs = *c.hs_pinned;
    if (c.hs.tie_deep)
        throw CapacityError("distinct particles share all 126 octant levels of the root cube (closer than 2^-126 "
                            "of its half-width); the reference would keep splitting them (Node.cpp:597-699)");
}

// Second half: nodes, moments and the walk layout, sized by the counts the
// first half left in c.hs_pinned (c.ev_counts has completed since; the
// prefix sums queued after it may still run).
void build_tree_b(Ctx &c) {
    int n = (int)c.n;
    hipStream_t st = c.stream;
    int proot = (c.n >= 2 && c.n >= c.p.split_threshold) ? 1 : 0;
    c.hs = *c.hs_pinned;
    if (c.hs.n_nodes < 0) throw HipError("the node counts of this build did not reach the host");
    if (c.hs.tie_big && c.sort_lo > 0) {
        // a run of particles sharing the sorted bits too long for the
        // ranking: the first half again, sorting every byte of word 0
        c.sort_lo = 0;
        build_tree_a(c);
        AG2_HIP(hipEventSynchronize(c.ev_counts));
        c.hs = *c.hs_pinned;
        if (c.hs.n_nodes < 0) throw HipError("the node counts of this build did not reach the host");
    } else if (!c.p.graph_launch) {
        // the next build leaves the low bytes out again unless its runs got long
        c.sort_lo = c.hs.max_prefix_run > kSortRunLong ? 0 : kSortLoMax;
    }
    c.kw_used = 2;
    const bool deep = c.hs.tie_big || c.hs.tie_deep;
    const bool tiles_current = !deep && !c.hs.err;
    if (deep) refine_ties(c);
    if (c.hs.err) {
        // more nodes than room: grow before the node phase
        int64_t need = (int64_t)c.hs.n_int + c.hs.nin;
        alloc_nodes(c, need + need / 4 + 64);
        int zero = 0;
        AG2_HIP(hipMemcpy(&c.ds->err, &zero, sizeof(int), hipMemcpyHostToDevice));
    }
    c.nin = c.hs.nin;
    c.n_int = c.hs.n_int;
    // Simulation.cpp:126: the visual-density radius comes from the first tree
    if (c.vd_radius0 == 0.0) c.vd_radius0 = __builtin_bit_cast(double, c.hs.Rbits) / 100000;
    c.n_nodes = c.hs.n_nodes;
    int nin = c.nin, nn = c.n_nodes;
    const core::KeyWords kw = key_words(c);
    core::NodeOut o{c.n_first, c.n_count, c.n_skip, c.n_parent, c.n_pidx, c.n_level};
    if (nn > 0) {
        // make_leaf's per-block partials (deepest level, gap count), folded by k_child_count
        unsigned long long *part = (unsigned long long *)c.partials;
        int gl = 0;
        if (nin == 0) {
            hipLaunchKernelGGL(k_make_root_only, dim3(1), dim3(1), 0, st, o);
        } else {
            int gi = grid_for(nin, 256);
            gl = gi < kLeafBlocks ? gi : kLeafBlocks;
            // n_cstart lists the internal nodes' particles until the child-place scan overwrites it
            hipLaunchKernelGGL(k_make_leaf, dim3(gl), dim3(256), 0, st, nin, kw, c.lcp, c.lk, c.base, c.cnt, o,
                               c.idx_alt, c.gapl, c.n_cstart, part, part + kLeafBlocks);
            if (c.n_int > 0)
                hipLaunchKernelGGL(k_make_internal, dim3(grid_for(c.n_int, 256)), dim3(256), 0, st, nin, nn,
                                   c.n_int, kw, c.lcp, c.base, c.n_cstart, o);
        }
        PVals pv{c.x, c.y, c.z, c.m, c.vx, c.vy, c.vz, c.type};
        // child counts, node sums bottom-up, the child-place scan, then
        // moments and walk records
        int gn = grid_for(nn, 256);
        int nbn = (nn + 1023) / 1024;
        hipLaunchKernelGGL(k_child_count, dim3(nbn), dim3(256), 0, st, nn, c.n_pidx, c.n_level, c.n_count, c.n_skip,
                           c.n_nchild, c.n_rank, c.n_ichild, gl, part, part + kLeafBlocks, c.ds, c.n_bsum);
        if (nin > 0) {
            // the block sums from build_order are stale after a refinement
            // re-ordered the particles or the node arrays grew
            if (!tiles_current) launch_tile_sums(c);
            NodeSums *blk1, *blk2;
            tile_sum_buffers(c, &blk1, &blk2);
            hipLaunchKernelGGL(k_sums_direct, dim3(gn), dim3(256), 0, st, nn, nin, n, proot, c.n_first, c.n_count,
                               c.n_pidx, pv, blk1, blk2, c.nsum);
        } else if (proot)
            hipLaunchKernelGGL(k_moments_root_only, dim3(1), dim3(1), 0, st, n, pv, c.nsum);
        hipLaunchKernelGGL(k_iscan_apply, dim3(nbn), dim3(256), 0, st, nn, c.n_nchild, c.n_bsum, c.n_cstart);
        NodeMoments nm{c.n_mass, c.n_gas, c.n_cx, c.n_cy, c.n_cz, c.n_gvx, c.n_gvy, c.n_gvz};
        hipLaunchKernelGGL(k_node_moments, dim3(gn), dim3(256), 0, st, nn, n, proot, c.p.split_threshold,
                           c.zero_mass > 0 ? 1 : 0, c.rid, c.ds,
                           c.n_first, c.n_count, c.n_parent, c.n_pidx, c.n_level, c.nsum, c.gas_table, c.gas_equal,
                           c.gas_count, pv,
                           c.n_nchild, c.n_rank, c.n_cstart, nm, c.tnode,
                           c.sph_active ? c.tsph : nullptr, c.g_first, c.g_fired);
        c.groups_cleared = true;
        if (deep && nin > 0) {
            // the walk's stack bound comes from the branching depth here
            hipLaunchKernelGGL(k_reset_walk_depth, dim3(1), dim3(1), 0, st, c.ds);
            hipLaunchKernelGGL(k_branch_depth, dim3(grid_for(nin, 256)), dim3(256), 0, st, nin, c.base, c.cnt,
                               c.n_parent, c.n_ichild, c.ds);
            AG2_HIP(hipMemcpyAsync(&c.hs_pinned->walk_depth, &c.ds->walk_depth, sizeof(int32_t),
                                   hipMemcpyDeviceToHost, st));
            AG2_HIP(hipStreamSynchronize(st));
            if (c.hs_pinned->walk_depth > kWalkGrowthMax) {
                c.tree_valid = false;
                throw CapacityError("the walk's stack could grow by " + std::to_string(c.hs_pinned->walk_depth) +
                                    " blocks along one path of this tree, more than it holds (" +
                                    std::to_string(kWalkGrowthMax) + ")");
            }
        }
    }
    c.wnode_valid = false;
    if (c.timing >= 2) AG2_HIP(hipEventRecord(c.ev[3], st));
    AG2_HIP(hipGetLastError());
    c.tree_valid = true;
}

void build_tree(Ctx &c) {
    build_tree_a(c);
    AG2_HIP(hipEventSynchronize(c.ev_counts));
    build_tree_b(c);
}

void pack_walknodes(Ctx &c) {
    if (c.wnode_valid || c.n_nodes <= 0) return;
    hipLaunchKernelGGL(k_pack_walknode, dim3(grid_for(c.n_nodes, 256)), dim3(256), 0, c.stream, c.n_nodes, c.ds,
                       c.n_skip, c.n_pidx, c.n_level, c.n_mass, c.n_gas, c.n_cx, c.n_cy, c.n_cz, c.wnode);
    AG2_HIP(hipGetLastError());
    c.wnode_valid = true;
}

}  // namespace ag2
