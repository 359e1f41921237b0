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
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double R = __longlong_as_double((long long)rb);
    uint64_t k1, k2;
    int g;
    bool in = core::octant_path(x[i], y[i], z[i], R, 21, &k1, &k2, &g);
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
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int w = threadIdx.x >> 6;
    int base = tile * kSortTile + w * (64 * kSortItems) + lane_id();
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        int pos = base + j * 64;
        if (pos < n) atomicAdd(&cnt[(key[pos] >> shift) & 255], 1u);
    }
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
    int w = tid >> 6;
    unsigned long long below = lanes_below();
    int base = tile * kSortTile + w * (64 * kSortItems) + lane_id();
    uint64_t keys[kSortItems];
    uint32_t vals[kSortItems];
    int loc[kSortItems], dig[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        int pos = base + j * 64;
        bool valid = pos < n;
        uint64_t k = valid ? kin[pos] : 0ull;
        uint32_t v = valid ? vin[pos] : 0u;
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
        keys[j] = k; vals[j] = v; loc[j] = valid ? cur + rank : -1; dig[j] = d;
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
__global__ __launch_bounds__(256) void k_count_in(int n, const uint64_t *key, DevState *ds) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    bool in = key[i] != kOut;
    if (in && (i + 1 == n || key[i + 1] == kOut)) ds->nin = i + 1;
    if (i == 0 && !in) ds->nin = 0;
}

// Particles whose 21-level key equals a neighbour's get their levels 22..42.
__global__ __launch_bounds__(256) void k_tie_key2(int n, const uint64_t *key, const uint32_t *idx,
                                                  const double *x, const double *y, const double *z,
                                                  DevState *ds, uint64_t *key2, uint8_t *gapl) {
    int i = blockIdx.x * 256 + threadIdx.x;
    int nin = ds->nin;
    if (i >= nin) return;
    bool tie = (i > 0 && key[i - 1] == key[i]) || (i + 1 < nin && key[i + 1] == key[i]);
    if (!tie) { key2[i] = 0ull; return; }
    double R = __longlong_as_double((long long)ds->Rbits);
    uint32_t q = idx[i];
    uint64_t a, b;
    int g;
    core::octant_path(x[q], y[q], z[q], R, 42, &a, &b, &g);
    key2[i] = b;
    if (gapl[q] == 0 && g > 0) gapl[q] = (uint8_t)(g > 254 ? 254 : g);
}

// Insertion sort of each tie run by (second key word, caller index): the
// run's particles end in the order the reference inserts them (array order,
// Tree.cpp:44-48) whatever the resident order they were sorted from.
__global__ __launch_bounds__(256) void k_tie_sort(const uint64_t *key, uint64_t *key2, uint32_t *idx,
                                                  const uint32_t *rid, const DevState *ds) {
    int i = blockIdx.x * 256 + threadIdx.x;
    int nin = ds->nin;
    if (i + 1 >= nin) return;
    if (key[i + 1] != key[i] || (i > 0 && key[i - 1] == key[i])) return;
    int e = i + 1;
    while (e + 1 < nin && key[e + 1] == key[i]) e++;
    for (int a = i + 1; a <= e; a++) {
        uint64_t k = key2[a];
        uint32_t v = idx[a];
        uint32_t cv = rid[v];
        int b = a - 1;
        while (b >= i && (key2[b] > k || (key2[b] == k && rid[idx[b]] > cv))) {
            key2[b + 1] = key2[b];
            idx[b + 1] = idx[b];
            b--;
        }
        key2[b + 1] = k;
        idx[b + 1] = v;
    }
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
__global__ __launch_bounds__(256) void k_permute(int n, const uint32_t *__restrict__ idx, StateSet in, StateSet out) {
    int k = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (k >= n) return;
    uint32_t q = idx[k];
#pragma unroll
    for (int j = 0; j < kStateD; j++) out.d[j][k] = in.d[j][q];
    out.type[k] = in.type[q];
    out.active[k] = in.active[q];
    out.rid[k] = in.rid[q];
}

__global__ __launch_bounds__(256) void k_inverse(int n, const uint32_t *rid, int32_t *inv) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) inv[rid[r]] = r;
}

// L[i] and c[i] = max(0, L[i] - L[i-1]).
__global__ __launch_bounds__(256) void k_lcp(const uint64_t *key, const uint64_t *key2, const DevState *ds,
                                             int parallel_root, int8_t *L, int32_t *c) {
    int i = blockIdx.x * 256 + threadIdx.x;
    int nin = ds->nin;
    if (i >= nin) return;
    int li = core::pair_lcp(key, key2, i, nin, parallel_root != 0);
    int lp = i == 0 ? -1 : core::pair_lcp(key, key2, i - 1, nin, parallel_root != 0);
    L[i] = (int8_t)li;
    int v = li - lp;
    c[i] = v > 0 ? v : 0;
}

// Device-wide exclusive scan of c over [0, nin): per-tile sums, tile-sum scan,
// tile scan.  Tiles of 1024 (256 threads x 4).
__global__ __launch_bounds__(256) void k_scan_tiles(const int32_t *c, const DevState *ds, int32_t *bsum) {
    __shared__ int scr[4];
    int nin = ds->nin;
    int b0 = blockIdx.x * 1024 + threadIdx.x * 4;
    int s = 0;
    for (int k = 0; k < 4; k++) s += (b0 + k < nin) ? c[b0 + k] : 0;
    int tot;
    block256_excl_scan(s, scr, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_sums(int nb, int32_t *bsum, DevState *ds, int parallel_root,
                                                   int64_t node_cap) {
    __shared__ int scr[4];
    int per = (nb + 255) / 256;
    int s0 = threadIdx.x * per;
    int s1 = s0 + per < nb ? s0 + per : nb;
    int local = 0;
    for (int k = s0; k < s1; k++) local += bsum[k];
    int tot;
    int run = block256_excl_scan(local, scr, &tot);
    for (int k = s0; k < s1; k++) { int t = bsum[k]; bsum[k] = run; run += t; }
    if (threadIdx.x == 0) {
        int nin = ds->nin;
        int n_int = (nin == 0) ? (parallel_root ? 1 : 0) : tot;
        ds->n_int = n_int;
        ds->n_nodes = n_int + nin;
        if ((int64_t)n_int + nin > node_cap) ds->err = 1;
    }
}

__global__ __launch_bounds__(256) void k_scan_apply(const int32_t *c, const int32_t *bsum, const DevState *ds,
                                                    int32_t *base) {
    __shared__ int scr[4];
    int nin = ds->nin;
    int b0 = blockIdx.x * 1024 + threadIdx.x * 4;
    int v[4];
    int s = 0;
    for (int k = 0; k < 4; k++) { v[k] = (b0 + k < nin) ? c[b0 + k] : 0; s += v[k]; }
    int ex = block256_excl_scan(s, scr, nullptr) + bsum[blockIdx.x];
    for (int k = 0; k < 4; k++) {
        if (b0 + k <= nin) base[b0 + k] = ex;
        ex += v[k];
    }
}

// ------------------------------------------------------------ nodes
// One thread per node: own[nd] (written by k_make_leaf) is the particle
// whose pair starts internal node nd, or -1 for a leaf.  A particle can start
// a chain of up to 42 nested cells, each needing a galloping search for its
// last particle; giving every cell its own thread keeps those searches from
// running one after another in a single lane.
__global__ __launch_bounds__(256) void k_make_internal(int nin, int nn, const uint64_t *key, const uint64_t *key2,
                                                       const int8_t *L, const int32_t *base, const int32_t *own,
                                                       core::NodeOut o) {
    int nd = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (nd >= nn) return;
    int i = own[nd];
    if (i < 0) return;
    int l = core::lcp_prev(L, i) + 1 + (nd - i - base[i]);
    core::make_internal(key, key2, L, base, i, l, nin, nn, o);
}

__global__ __launch_bounds__(256) void k_make_root_only(core::NodeOut o) {
    o.first[0] = 0; o.count[0] = 0; o.skip[0] = 1; o.parent[0] = -1; o.pidx[0] = -1; o.level[0] = 0;
}

// Grid-stride over a fixed grid (at most kLeafBlocks workgroups), so the
// deepest level and the gap count cost one pair of atomics per workgroup.
constexpr int kLeafBlocks = 1024;

__global__ __launch_bounds__(256) void k_make_leaf(int nin, const uint64_t *key, const uint64_t *key2,
                                                   const int8_t *L, const int32_t *base, const int32_t *c,
                                                   core::NodeOut o, const uint32_t *idx, const uint8_t *gapl,
                                                   int32_t *own, unsigned long long *part_lv,
                                                   unsigned long long *part_gap) {
    __shared__ int wl[4], wg[4];
    int lv = 0, gp = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nin; i += gridDim.x * 256) {
        int li = 0;
        core::make_leaf(key, key2, L, base, c, i, o, &li);
        // owners of the node indices of particle i: c[i] internal nodes, then its leaf
        int nd0 = i + base[i], ci = c[i];
        for (int k = 0; k < ci; k++) own[nd0 + k] = i;
        own[nd0 + ci] = -1;
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

// Deepest leaf level and gap count from make_leaf's per-block partials.
__global__ __launch_bounds__(256) void k_leaf_fold(int nb, const unsigned long long *part_lv,
                                                   const unsigned long long *part_gap, DevState *ds) {
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
        ds->gap = g;
    }
}

// ------------------------------------------------------------ moments
// Per-particle quantities in key order: m, m x, m y, m z, and for gas m,
// m vx, m vy, m vz.  Their exclusive prefix sums (double-double) give any
// node's sums as P[first + count] - P[first], exactly rounded.
__device__ inline double pval(int k, int i, const double *sx, const double *sy, const double *sz,
                              const double *sm, const uint8_t *stype, const double *svx, const double *svy,
                              const double *svz) {
    double m = sm[i];
    bool gas = stype[i] == 2;
    switch (k) {
        case 0: return m;
        case 1: return sx[i] * m;
        case 2: return sy[i] * m;
        case 3: return sz[i] * m;
        case 4: return gas ? m : 0.0;
        case 5: return gas ? svx[i] * m : 0.0;
        case 6: return gas ? svy[i] * m : 0.0;
        default: return gas ? svz[i] * m : 0.0;
    }
}

__device__ inline core::DD dd_shfl_up(core::DD a, int d) {
    return core::DD{__shfl_up(a.hi, d, 64), __shfl_up(a.lo, d, 64)};
}

__device__ inline core::DD wave_dd_incl(core::DD v) {
    for (int d = 1; d < 64; d <<= 1) {
        core::DD t = dd_shfl_up(v, d);
        if (lane_id() >= d) v = core::dd_add(t, v);
    }
    return v;
}

struct PVals {
    const double *sx, *sy, *sz, *sm, *svx, *svy, *svz;
    const uint8_t *stype;
};

// The kPrefixQ values of particles b0..b0+3 (0 past n): every quantity from
// one read of the particle.
__device__ inline void pvals4(int n, int b0, const PVals &v, double (&val)[kPrefixQ][4]) {
    for (int j = 0; j < 4; j++) {
        int i = b0 + j;
        if (i < n) {
            for (int k = 0; k < kPrefixQ; k++) val[k][j] = pval(k, i, v.sx, v.sy, v.sz, v.sm, v.stype, v.svx, v.svy, v.svz);
        } else {
            for (int k = 0; k < kPrefixQ; k++) val[k][j] = 0.0;
        }
    }
}

// Tiles of 1024 particles (256 threads x 4 consecutive): tile totals.  All
// quantities are scanned from one read of the tile, with one barrier.
__global__ __launch_bounds__(256) void k_dd_tile(int n, PVals v, core::DD *tsum, int ntiles) {
    __shared__ core::DD wt[kPrefixQ][4];
    int w = threadIdx.x >> 6;
    int b0 = blockIdx.x * 1024 + threadIdx.x * 4;
    double val[kPrefixQ][4];
    pvals4(n, b0, v, val);
    for (int k = 0; k < kPrefixQ; k++) {
        core::DD loc{0.0, 0.0};
        for (int j = 0; j < 4; j++) loc = core::dd_add(loc, core::DD{val[k][j], 0.0});
        core::DD inc = wave_dd_incl(loc);
        if (lane_id() == 63) wt[k][w] = inc;
    }
    __syncthreads();
    if (threadIdx.x < kPrefixQ) {
        int k = threadIdx.x;
        tsum[(size_t)k * ntiles + blockIdx.x] =
            core::dd_add(core::dd_add(core::dd_add(wt[k][0], wt[k][1]), wt[k][2]), wt[k][3]);
    }
}

// Exclusive scan of the tile totals, one workgroup per quantity; writes the
// grand totals to P[n].
__global__ __launch_bounds__(256) void k_dd_tsum(int n, int ntiles, core::DD *tsum, core::DD *pre) {
    __shared__ core::DD wt[4];
    int w = threadIdx.x >> 6;
    int per = (ntiles + 255) / 256;
    int s0 = threadIdx.x * per;
    int s1 = s0 + per < ntiles ? s0 + per : ntiles;
    {
        int k = blockIdx.x;
        core::DD *t = tsum + (size_t)k * ntiles;
        core::DD loc{0.0, 0.0};
        for (int j = s0; j < s1; j++) loc = core::dd_add(loc, t[j]);
        core::DD inc = wave_dd_incl(loc);
        core::DD prev = dd_shfl_up(inc, 1);
        if (lane_id() == 0) prev = core::DD{0.0, 0.0};
        if (lane_id() == 63) wt[w] = inc;
        __syncthreads();
        core::DD off{0.0, 0.0}, tot{0.0, 0.0};
        for (int q = 0; q < 4; q++) {
            if (q < w) off = core::dd_add(off, wt[q]);
            tot = core::dd_add(tot, wt[q]);
        }
        core::DD run = core::dd_add(off, prev);
        for (int j = s0; j < s1; j++) {
            core::DD x = t[j];
            t[j] = run;
            run = core::dd_add(run, x);
        }
        if (threadIdx.x == 0) pre[(size_t)k * (n + 1) + n] = tot;
    }
}

__global__ __launch_bounds__(256) void k_dd_apply(int n, PVals v, const core::DD *toff, int ntiles, core::DD *pre) {
    __shared__ core::DD wt[kPrefixQ][4];
    int w = threadIdx.x >> 6;
    int b0 = blockIdx.x * 1024 + threadIdx.x * 4;
    double val[kPrefixQ][4];
    pvals4(n, b0, v, val);
    core::DD prev[kPrefixQ];
    for (int k = 0; k < kPrefixQ; k++) {
        core::DD loc{0.0, 0.0};
        for (int j = 0; j < 4; j++) loc = core::dd_add(loc, core::DD{val[k][j], 0.0});
        core::DD inc = wave_dd_incl(loc);
        prev[k] = dd_shfl_up(inc, 1);
        if (lane_id() == 0) prev[k] = core::DD{0.0, 0.0};
        if (lane_id() == 63) wt[k][w] = inc;
    }
    __syncthreads();
    for (int k = 0; k < kPrefixQ; k++) {
        core::DD off = toff[(size_t)k * ntiles + blockIdx.x];
        for (int q = 0; q < w; q++) off = core::dd_add(off, wt[k][q]);
        core::DD run = core::dd_add(off, prev[k]);
        core::DD *P = pre + (size_t)k * (n + 1);
        for (int j = 0; j < 4; j++) {
            if (b0 + j < n) P[b0 + j] = run;
            run = core::dd_add(run, core::DD{val[k][j], 0.0});
        }
    }
}

struct NodeMoments {
    double *m, *gas, *cx, *cy, *cz, *mvx, *mvy, *mvz;
};

// Child-contiguous layout for the block walk (core::TreeNode): the root at 0,
// the children of node p at [1 + cstart[p], 1 + cstart[p] + nchild[p]) in
// octant order, cstart being the exclusive scan of nchild in pre-order.  Every node
// follows its children's skip links (pull, no atomics): it counts them and
// gives each its rank among its siblings, which in pre-order is its octant
// order.  Bucket nodes at the last key level hold their member leaves, the
// next count nodes, as children.
__global__ __launch_bounds__(256) void k_child_count(int nn, const int32_t *pidx, const int8_t *level,
                                                     const int32_t *count, const int32_t *skip, int32_t *nchild,
                                                     int32_t *rank) {
    int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nn) return;
    int nc = 0;
    if (pidx[p] < 0) {
        if (level[p] >= kMaxLevel) {
            nc = count[p];
            for (int k = 0; k < nc; k++) rank[p + 1 + k] = k;
        } else {
            const int end = skip[p];
            for (int ch = p + 1; ch < end; ch = skip[ch]) rank[ch] = nc++;
        }
    }
    nchild[p] = nc;
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
                                                      const int8_t *level, const core::DD *pre,
                                                      const double *gas_table, int gas_equal, int64_t ngas, PVals v,
                                                      const int32_t *nchild, const int32_t *rank,
                                                      const int32_t *cstart, NodeMoments o,
                                                      core::TreeNode *tree, core::SphNode *tsph) {
    // nodes in XCD ranges: neighbours in pre-order read neighbouring prefix
    // sums and particles
    int nd = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (nd >= nn) return;
    int p = pidx[nd];
    int par = parent[nd];
    double M, g, cx, cy, cz, mvx = 0.0, mvy = 0.0, mvz = 0.0;
    if (p >= 0) {
        M = v.sm[p];
        g = v.stype[p] == 2 ? M : 0.0;
        bool vec = par < 0 ? (n == 1) : core::parallel_node(count, parent, par, n, thr);
        cx = core::leaf_com(v.sx[p], M, vec);
        cy = core::leaf_com(v.sy[p], M, vec);
        cz = core::leaf_com(v.sz[p], M, vec);
    } else {
        int s = first[nd], e = first[nd] + count[nd];
        if (nd == 0 && proot) { s = 0; e = n; }
        double S[kPrefixQ];
        for (int k = 0; k < kPrefixQ; k++) {
            const core::DD *P = pre + (size_t)k * (n + 1);
            S[k] = core::dd_val(core::dd_add(P[e], core::dd_neg(P[s])));
        }
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

__global__ __launch_bounds__(256) void k_iscan_sums(int nb, int32_t *bsum) {
    __shared__ int scr[4];
    int per = (nb + 255) / 256;
    int s0 = threadIdx.x * per;
    int s1 = s0 + per < nb ? s0 + per : nb;
    int local = 0;
    for (int k = s0; k < s1; k++) local += bsum[k];
    int run = block256_excl_scan(local, scr, nullptr);
    for (int k = s0; k < s1; k++) { int t = bsum[k]; bsum[k] = run; run += t; }
}

__global__ __launch_bounds__(256) void k_iscan_apply(int n, const int32_t *in, const int32_t *bsum, int32_t *out) {
    __shared__ int scr[4];
    int b0 = blockIdx.x * 1024 + threadIdx.x * 4;
    int v[4];
    int s = 0;
    for (int k = 0; k < 4; k++) { v[k] = (b0 + k < n) ? in[b0 + k] : 0; s += v[k]; }
    int ex = block256_excl_scan(s, scr, nullptr) + bsum[blockIdx.x];
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

}  // namespace

void alloc_nodes(Ctx &c, int64_t cap) {
    reset_graphs(c);   // the node capacity is an argument of the captured segments
    auto fr = [](void *p) { if (p) (void)hipFree(p); };
    fr(c.n_first); fr(c.n_count); fr(c.n_skip); fr(c.n_parent); fr(c.n_pidx);
    fr(c.n_level); fr(c.n_mass); fr(c.n_gas); fr(c.n_gvx); fr(c.n_gvy);
    fr(c.n_gvz); fr(c.n_cx); fr(c.n_cy); fr(c.n_cz); fr(c.wnode); fr(c.g_first); fr(c.g_fired);
    fr(c.g_final); fr(c.g_rho); fr(c.g_list); fr(c.tnode); fr(c.tsph); fr(c.n_nchild); fr(c.n_rank); fr(c.n_cstart);
    fr(c.n_bsum);
    c.node_cap = cap;
    size_t n = (size_t)cap;
    AG2_HIP(hipMalloc(&c.n_first, n * 4)); AG2_HIP(hipMalloc(&c.n_count, n * 4)); AG2_HIP(hipMalloc(&c.n_skip, n * 4));
    AG2_HIP(hipMalloc(&c.n_parent, n * 4));
    AG2_HIP(hipMalloc(&c.n_pidx, n * 4)); AG2_HIP(hipMalloc(&c.n_level, n));
    double **d8[] = {&c.n_mass, &c.n_gas, &c.n_gvx, &c.n_gvy, &c.n_gvz, &c.n_cx, &c.n_cy, &c.n_cz, &c.g_rho};
    for (double **p : d8) AG2_HIP(hipMalloc(p, n * 8));
    AG2_HIP(hipMalloc(&c.wnode, n * sizeof(WalkNode)));
    AG2_HIP(hipMalloc(&c.g_first, n * 4)); AG2_HIP(hipMalloc(&c.g_list, (n + 256) * 4));
    AG2_HIP(hipMalloc(&c.g_fired, n)); AG2_HIP(hipMalloc(&c.g_final, n * 4));
    AG2_HIP(hipMalloc(&c.tnode, n * sizeof(core::TreeNode)));
    AG2_HIP(hipMalloc(&c.tsph, n * sizeof(core::SphNode)));
    AG2_HIP(hipMalloc(&c.n_nchild, n * 4)); AG2_HIP(hipMalloc(&c.n_rank, n * 4)); AG2_HIP(hipMalloc(&c.n_cstart, n * 4));
    AG2_HIP(hipMalloc(&c.n_bsum, (n / 1024 + 2) * 4));
}

void state_fields(Ctx &c, double **f[kStateD]) {
    double **l[kStateD] = {&c.x, &c.y, &c.z, &c.vx, &c.vy, &c.vz, &c.m, &c.U, &c.mu, &c.ax,
                           &c.ay, &c.az, &c.h, &c.rho, &c.P, &c.T, &c.dUdt, &c.ts, &c.next};
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
    hipLaunchKernelGGL(k_permute, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, n, c.idx, in, out);
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

static void radix_sort(Ctx &c, int n) {
    int nblk = (n + kSortTile - 1) / kSortTile;
    uint64_t *kin = c.key, *kout = c.key_alt;
    uint32_t *vin = c.idx, *vout = c.idx_alt;
    const bool small = nblk <= kSortSmallTiles;
    for (int pass = 0; pass < 8; pass++) {
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
    // eight passes: the result is back in c.key / c.idx
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
    hipLaunchKernelGGL(k_count_in, dim3(g), dim3(256), 0, st, n, c.key, c.ds);
    hipLaunchKernelGGL(k_tie_key2, dim3(g), dim3(256), 0, st, n, c.key, c.idx, c.x, c.y, c.z, c.ds, c.key2, c.gapl);
    hipLaunchKernelGGL(k_tie_sort, dim3(g), dim3(256), 0, st, c.key, c.key2, c.idx, c.rid, c.ds);
    launch_permute(c, n);
    int proot = (c.n >= 2 && c.n >= c.p.split_threshold) ? 1 : 0;
    // from here on the state is the permuted one (alternate buffers)
    hipLaunchKernelGGL(k_lcp, dim3(g), dim3(256), 0, st, c.key, c.key2, c.ds, proot, c.lcp, c.cnt);
    int nb = (n + 1023) / 1024;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(256), 0, st, c.cnt, c.ds, c.bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, st, nb, c.bsum, c.ds, proot, c.node_cap);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, st, c.cnt, c.bsum, c.ds, c.base);
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
    for (int pass = 0; pass < 8; pass++) {
        hipLaunchKernelGGL(k_rsb_hist, dim3(total), dim3(256), 0, st, dj, nc, pass);
        hipLaunchKernelGGL(k_rsb_scan, dim3(256 * nc), dim3(256), 0, st, dj);
        hipLaunchKernelGGL(k_rsb_scatter, dim3(total), dim3(256), 0, st, dj, nc, pass);
    }
    // eight passes: every job's result is back in its key / idx
    AG2_HIP(hipGetLastError());
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
    // The particle prefix sums need only the key-ordered state, so they run
    // while the host waits for the counts and launches the node phase.
    if (n > 0) {
        PVals pv{c.x, c.y, c.z, c.m, c.vx, c.vy, c.vz, c.type};
        int ntiles = (n + 1023) / 1024;
        hipLaunchKernelGGL(k_dd_tile, dim3(ntiles), dim3(256), 0, st, n, pv, c.tsum, ntiles);
        hipLaunchKernelGGL(k_dd_tsum, dim3(kPrefixQ), dim3(256), 0, st, n, ntiles, c.tsum, c.pre);
        hipLaunchKernelGGL(k_dd_apply, dim3(ntiles), dim3(256), 0, st, n, pv, c.tsum, ntiles, c.pre);
    }
}

// Second half: nodes, moments and the walk layout, sized by the counts the
// first half left in c.hs_pinned (c.ev_counts has completed since; the
// prefix sums queued after it may still run).
void build_tree_b(Ctx &c) {
    int n = (int)c.n;
    hipStream_t st = c.stream;
    int g = grid_for(n, 256);
    int proot = (c.n >= 2 && c.n >= c.p.split_threshold) ? 1 : 0;
    c.hs = *c.hs_pinned;
    if (c.hs.n_nodes < 0) throw HipError("the node counts of this build did not reach the host");
    if (c.hs.err) {
        // more nodes than room: grow and redo the node phase
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
    core::NodeOut o{c.n_first, c.n_count, c.n_skip, c.n_parent, c.n_pidx, c.n_level};
    if (nn > 0) {
        if (nin == 0) {
            hipLaunchKernelGGL(k_make_root_only, dim3(1), dim3(1), 0, st, o);
        } else {
            int gi = grid_for(nin, 256);
            int gl = gi < kLeafBlocks ? gi : kLeafBlocks;
            unsigned long long *part = (unsigned long long *)c.partials;
            // n_cstart holds the node owners until the child-place scan overwrites it
            hipLaunchKernelGGL(k_make_leaf, dim3(gl), dim3(256), 0, st, nin, c.key, c.key2, c.lcp, c.base, c.cnt, o,
                               c.idx, c.gapl, c.n_cstart, part, part + kLeafBlocks);
            hipLaunchKernelGGL(k_make_internal, dim3(grid_for(nn, 256)), dim3(256), 0, st, nin, nn, c.key, c.key2,
                               c.lcp, c.base, c.n_cstart, o);
            hipLaunchKernelGGL(k_leaf_fold, dim3(1), dim3(256), 0, st, gl, part, part + kLeafBlocks, c.ds);
        }
        PVals pv{c.x, c.y, c.z, c.m, c.vx, c.vy, c.vz, c.type};
        // child counts, their scan (first-child places), then moments and
        // walk records
        int gn = grid_for(nn, 256);
        int nbn = (nn + 1023) / 1024;
        hipLaunchKernelGGL(k_child_count, dim3(gn), dim3(256), 0, st, nn, c.n_pidx, c.n_level, c.n_count, c.n_skip,
                           c.n_nchild, c.n_rank);
        hipLaunchKernelGGL(k_iscan_tiles, dim3(nbn), dim3(256), 0, st, nn, c.n_nchild, c.n_bsum);
        hipLaunchKernelGGL(k_iscan_sums, dim3(1), dim3(256), 0, st, nbn, c.n_bsum);
        hipLaunchKernelGGL(k_iscan_apply, dim3(nbn), dim3(256), 0, st, nn, c.n_nchild, c.n_bsum, c.n_cstart);
        NodeMoments nm{c.n_mass, c.n_gas, c.n_cx, c.n_cy, c.n_cz, c.n_gvx, c.n_gvy, c.n_gvz};
        hipLaunchKernelGGL(k_node_moments, dim3(gn), dim3(256), 0, st, nn, n, proot, c.p.split_threshold,
                           c.zero_mass > 0 ? 1 : 0, c.rid, c.ds,
                           c.n_first, c.n_count, c.n_parent, c.n_pidx, c.n_level, c.pre, c.gas_table, c.gas_equal,
                           c.gas_count, pv,
                           c.n_nchild, c.n_rank, c.n_cstart, nm, c.tnode,
                           c.sph_active ? c.tsph : nullptr);
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
