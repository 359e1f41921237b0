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

// Tree::calcTreeWidth (Tree.cpp:86-117): cut = mean + 10 sigma of |x|.
__global__ __launch_bounds__(256) void k_tw_final(int n, int nb, const double *part, DevState *ds) {
    double s = 0.0, q = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) { s += part[2 * b]; q += part[2 * b + 1]; }
    s = wave_sum(s);
    q = wave_sum(q);
    __shared__ double ls[4], lq[4];
    int w = threadIdx.x >> 6;
    if (lane_id() == 0) { ls[w] = s; lq[w] = q; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = ((ls[0] + ls[1]) + ls[2]) + ls[3];
        double sq = ((lq[0] + lq[1]) + lq[2]) + lq[3];
        double mean = sum / (double)n;
        double sd = sqrt(sq / (double)n - mean * mean);
        ds->sum = sum;
        ds->sumsq = sq;
        ds->cut = mean + 10.0 * sd;
        ds->Rbits = 0ull;
        ds->gap = 0ull;
        ds->err = 0;
        ds->max_level = 0;
    }
}

__global__ __launch_bounds__(256) void k_tw_max(int n, const double *__restrict__ x, const double *__restrict__ y,
                                                const double *__restrict__ z, DevState *ds) {
    double cut = ds->cut;
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
    if (lane_id() == 0 && best) atomicMax(&ds->Rbits, best);
}

// ------------------------------------------------------------ keys
__global__ __launch_bounds__(256) void k_keygen(int n, const double *__restrict__ x, const double *__restrict__ y,
                                                const double *__restrict__ z, const DevState *ds,
                                                uint64_t *key, uint32_t *idx, uint8_t *gapl) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double R = __longlong_as_double((long long)ds->Rbits);
    uint64_t k1, k2;
    int g;
    bool in = core::octant_path(x[i], y[i], z[i], R, 21, &k1, &k2, &g);
    key[i] = in ? k1 : kOut;
    idx[i] = (uint32_t)i;
    gapl[i] = in ? (uint8_t)g : (uint8_t)255;
}

// ------------------------------------------------------------ radix sort
// Workgroup tile: 2048 keys, wave w owns tile positions [512 w, 512 w + 512),
// item j of lane l at 512 w + 64 j + l, so (wave, item, lane) is tile order
// and the scatter is stable.
__global__ __launch_bounds__(256) void k_rs_hist(int n, const uint64_t *__restrict__ key, int shift,
                                                 uint32_t *hist, int nblk) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int w = threadIdx.x >> 6;
    int base = blockIdx.x * kSortTile + w * 512 + lane_id();
#pragma unroll
    for (int j = 0; j < 8; j++) {
        int pos = base + j * 64;
        if (pos < n) atomicAdd(&cnt[(key[pos] >> shift) & 255], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

// One workgroup per digit: exclusive scan of that digit's per-tile counts.
__global__ __launch_bounds__(256) void k_rs_scan(int nblk, uint32_t *hist, uint32_t *dtotal) {
    __shared__ int scr[4];
    uint32_t *h = hist + (size_t)blockIdx.x * nblk;
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
    if (threadIdx.x == 0) dtotal[blockIdx.x] = (uint32_t)tot;
}

__global__ __launch_bounds__(256) void k_rs_scatter(int n, const uint64_t *__restrict__ kin,
                                                    const uint32_t *__restrict__ vin, uint64_t *__restrict__ kout,
                                                    uint32_t *__restrict__ vout, int shift,
                                                    const uint32_t *__restrict__ hist,
                                                    const uint32_t *__restrict__ dtotal, int nblk) {
    __shared__ int sbase[256];
    __shared__ int wcnt[4][256];
    __shared__ int scr[4];
    int tid = threadIdx.x;
    int tot;
    int ex = block256_excl_scan((int)dtotal[tid], scr, &tot);
    sbase[tid] = ex + (int)hist[tid * nblk + blockIdx.x];
    wcnt[0][tid] = 0; wcnt[1][tid] = 0; wcnt[2][tid] = 0; wcnt[3][tid] = 0;
    __syncthreads();
    int w = tid >> 6;
    unsigned long long below = lanes_below();
    int base = blockIdx.x * kSortTile + w * 512 + lane_id();
    uint64_t keys[8];
    uint32_t vals[8];
    int loc[8], dig[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
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
    for (int j = 0; j < 8; j++) {
        if (loc[j] >= 0) {
            int dst = sbase[dig[j]] + wcnt[w][dig[j]] + loc[j];
            kout[dst] = keys[j];
            vout[dst] = vals[j];
        }
    }
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

// Stable insertion sort of each tie run by its second key word.
__global__ __launch_bounds__(256) void k_tie_sort(const uint64_t *key, uint64_t *key2, uint32_t *idx,
                                                  const DevState *ds) {
    int i = blockIdx.x * 256 + threadIdx.x;
    int nin = ds->nin;
    if (i + 1 >= nin) return;
    if (key[i + 1] != key[i] || (i > 0 && key[i - 1] == key[i])) return;
    int e = i + 1;
    while (e + 1 < nin && key[e + 1] == key[i]) e++;
    for (int a = i + 1; a <= e; a++) {
        uint64_t k = key2[a];
        uint32_t v = idx[a];
        int b = a - 1;
        while (b >= i && key2[b] > k) { key2[b + 1] = key2[b]; idx[b + 1] = idx[b]; b--; }
        key2[b + 1] = k;
        idx[b + 1] = v;
    }
}

__global__ __launch_bounds__(256) void k_gather(int n, const uint32_t *__restrict__ idx,
                                                const double *__restrict__ x, const double *__restrict__ y,
                                                const double *__restrict__ z, const double *__restrict__ m,
                                                const uint8_t *__restrict__ type, const double *__restrict__ vx,
                                                const double *__restrict__ vy, const double *__restrict__ vz,
                                                double *sx, double *sy, double *sz, double *sm, uint8_t *stype,
                                                double *svx, double *svy, double *svz) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t q = idx[i];
    sx[i] = x[q]; sy[i] = y[q]; sz[i] = z[q]; sm[i] = m[q]; stype[i] = type[q];
    svx[i] = vx[q]; svy[i] = vy[q]; svz[i] = vz[q];
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
__global__ __launch_bounds__(256) void k_zero_nodes(int nn, int32_t *nchild, int32_t *arrive) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < nn) { nchild[i] = 0; arrive[i] = 0; }
}

__global__ __launch_bounds__(256) void k_make_internal(int nin, int nn, const uint64_t *key, const uint64_t *key2,
                                                       const int8_t *L, const int32_t *base, core::NodeOut o,
                                                       int32_t *nchild) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nin) return;
    int lp = core::lcp_prev(L, i);
    for (int l = lp + 1; l <= L[i]; l++) {
        int par = core::make_internal(key, key2, L, base, i, l, nin, nn, o);
        if (par >= 0) atomicAdd(&nchild[par], 1);
    }
}

__global__ __launch_bounds__(256) void k_make_root_only(core::NodeOut o) {
    o.first[0] = 0; o.count[0] = 0; o.skip[0] = 1; o.parent[0] = -1; o.pidx[0] = -1; o.level[0] = 0;
}

__global__ __launch_bounds__(256) void k_make_leaf(int nin, const uint64_t *key, const uint64_t *key2,
                                                   const int8_t *L, const int32_t *base, const int32_t *c,
                                                   core::NodeOut o, int32_t *nchild, const uint32_t *idx,
                                                   const uint8_t *gapl, DevState *ds) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nin) return;
    int lv;
    int par = core::make_leaf(key, key2, L, base, c, i, o, &lv);
    if (par >= 0) atomicAdd(&nchild[par], 1);
    int g = gapl[idx[i]];
    if (g > 0 && g <= lv) atomicAdd(&ds->gap, 1ull);
    atomicMax(&ds->max_level, lv);
}

struct NodeSums {
    double *m, *sx, *sy, *sz, *g, *gx, *gy, *gz, *cx, *cy, *cz;
};

// Bottom-up moments: every leaf climbs; the last child to arrive at a node
// sums all of that node's children in octant order and climbs on.
__global__ __launch_bounds__(256) void k_moments_up(int nin, int64_t N, int64_t thr, const int32_t *base,
                                                    const int32_t *c, const int32_t *count, const int32_t *parent,
                                                    const int32_t *skip, const int32_t *nchild, int32_t *arrive,
                                                    const double *sx, const double *sy, const double *sz,
                                                    const double *sm, const uint8_t *stype, const uint32_t *idx,
                                                    const double *vx, const double *vy, const double *vz,
                                                    NodeSums s) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nin) return;
    int nd = i + base[i] + c[i];
    double m = sm[i];
    bool gas = stype[i] == 2;
    uint32_t q = idx[i];
    s.m[nd] = m;
    s.sx[nd] = sx[i] * m; s.sy[nd] = sy[i] * m; s.sz[nd] = sz[i] * m;
    s.g[nd] = gas ? m : 0.0;
    s.gx[nd] = gas ? vx[q] * m : 0.0; s.gy[nd] = gas ? vy[q] * m : 0.0; s.gz[nd] = gas ? vz[q] * m : 0.0;
    int par = parent[nd];
    bool vec = par < 0 ? (N == 1) : core::parallel_node(count, parent, par, N, thr);
    s.cx[nd] = core::leaf_com(sx[i], m, vec);
    s.cy[nd] = core::leaf_com(sy[i], m, vec);
    s.cz[nd] = core::leaf_com(sz[i], m, vec);
    int cur = par;
    while (cur >= 0) {
        __threadfence();
        int old = atomicAdd(&arrive[cur], 1);
        if (old != nchild[cur] - 1) return;
        __threadfence();
        core::Sums t = core::sum_children(cur, skip, s.m, s.sx, s.sy, s.sz, s.g, s.gx, s.gy, s.gz);
        s.m[cur] = t.m; s.sx[cur] = t.sx; s.sy[cur] = t.sy; s.sz[cur] = t.sz;
        s.g[cur] = t.g; s.gx[cur] = t.gx; s.gy[cur] = t.gy; s.gz[cur] = t.gz;
        cur = parent[cur];
    }
}

// The parallel root also sums the particles outside its cube
// (Node.cpp:477-499 adds every listed particle before getOctant drops it).
__global__ __launch_bounds__(256) void k_root_outside(int n, const DevState *ds, const double *sx,
                                                      const double *sy, const double *sz, const double *sm,
                                                      const uint8_t *stype, const uint32_t *idx,
                                                      const double *vx, const double *vy, const double *vz,
                                                      NodeSums s, int empty_root) {
    int nin = ds->nin;
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = nin + threadIdx.x; k < n; k += 256) {
        double m = sm[k];
        a[0] += m; a[1] += sx[k] * m; a[2] += sy[k] * m; a[3] += sz[k] * m;
        if (stype[k] == 2) {
            uint32_t q = idx[k];
            a[4] += m; a[5] += vx[q] * m; a[6] += vy[q] * m; a[7] += vz[q] * m;
        }
    }
    __shared__ double red[4][8];
    int w = threadIdx.x >> 6;
    for (int j = 0; j < 8; j++) {
        double v = wave_sum(a[j]);
        if (lane_id() == 0) red[w][j] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[8];
        for (int j = 0; j < 8; j++) t[j] = ((red[0][j] + red[1][j]) + red[2][j]) + red[3][j];
        double b[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (!empty_root) { b[0] = s.m[0]; b[1] = s.sx[0]; b[2] = s.sy[0]; b[3] = s.sz[0]; b[4] = s.g[0]; b[5] = s.gx[0]; b[6] = s.gy[0]; b[7] = s.gz[0]; }
        s.m[0] = b[0] + t[0]; s.sx[0] = b[1] + t[1]; s.sy[0] = b[2] + t[2]; s.sz[0] = b[3] + t[3];
        s.g[0] = b[4] + t[4]; s.gx[0] = b[5] + t[5]; s.gy[0] = b[6] + t[6]; s.gz[0] = b[7] + t[7];
    }
}

// Centre of mass, gas-weighted velocity and the walk records.
__global__ __launch_bounds__(256) void k_finalize(int nn, const DevState *ds, const int8_t *level,
                                                  const int32_t *skip, const int32_t *pidx, NodeSums s,
                                                  WalkNode *wn) {
    int nd = blockIdx.x * 256 + threadIdx.x;
    if (nd >= nn) return;
    double R = __longlong_as_double((long long)ds->Rbits);
    double M = s.m[nd];
    int p = pidx[nd];
    double cx, cy, cz;
    if (p >= 0) { cx = s.cx[nd]; cy = s.cy[nd]; cz = s.cz[nd]; }
    else if (M > 0) { cx = s.sx[nd] / M; cy = s.sy[nd] / M; cz = s.sz[nd] / M; s.cx[nd] = cx; s.cy[nd] = cy; s.cz[nd] = cz; }
    else { cx = cy = cz = 0.0; s.cx[nd] = 0.0; s.cy[nd] = 0.0; s.cz[nd] = 0.0; }
    double g = s.g[nd];
    double mvx = 0.0, mvy = 0.0, mvz = 0.0;
    if (g > 0) { mvx = s.gx[nd] / g; mvy = s.gy[nd] / g; mvz = s.gz[nd] / g; }
    s.gx[nd] = mvx; s.gy[nd] = mvy; s.gz[nd] = mvz;
    WalkNode w;
    w.cx = cx; w.cy = cy; w.cz = cz; w.mass = M;
    w.radius = ldexp(R, -(int)level[nd]);
    w.gas = g;
    w.skip = skip[nd]; w.pidx = p; w.level = level[nd]; w.flags = 0;
    wn[nd] = w;
}

}  // namespace

void alloc_nodes(Ctx &c, int64_t cap) {
    auto fr = [](void *p) { if (p) (void)hipFree(p); };
    fr(c.n_first); fr(c.n_count); fr(c.n_skip); fr(c.n_parent); fr(c.n_nchild); fr(c.n_arrive); fr(c.n_pidx);
    fr(c.n_level); fr(c.n_mass); fr(c.n_sx); fr(c.n_sy); fr(c.n_sz); fr(c.n_gas); fr(c.n_gvx); fr(c.n_gvy);
    fr(c.n_gvz); fr(c.n_cx); fr(c.n_cy); fr(c.n_cz); fr(c.wnode); fr(c.wnode32); fr(c.g_first); fr(c.g_fired);
    fr(c.g_final); fr(c.g_rho); fr(c.g_list);
    c.node_cap = cap;
    size_t n = (size_t)cap;
    AG2_HIP(hipMalloc(&c.n_first, n * 4)); AG2_HIP(hipMalloc(&c.n_count, n * 4)); AG2_HIP(hipMalloc(&c.n_skip, n * 4));
    AG2_HIP(hipMalloc(&c.n_parent, n * 4)); AG2_HIP(hipMalloc(&c.n_nchild, n * 4)); AG2_HIP(hipMalloc(&c.n_arrive, n * 4));
    AG2_HIP(hipMalloc(&c.n_pidx, n * 4)); AG2_HIP(hipMalloc(&c.n_level, n));
    double **d8[] = {&c.n_mass, &c.n_sx, &c.n_sy, &c.n_sz, &c.n_gas, &c.n_gvx, &c.n_gvy, &c.n_gvz, &c.n_cx, &c.n_cy, &c.n_cz, &c.g_rho};
    for (double **p : d8) AG2_HIP(hipMalloc(p, n * 8));
    AG2_HIP(hipMalloc(&c.wnode, n * sizeof(WalkNode)));
    AG2_HIP(hipMalloc(&c.wnode32, n * sizeof(WalkNode32)));
    AG2_HIP(hipMalloc(&c.g_first, n * 4)); AG2_HIP(hipMalloc(&c.g_list, n * 4));
    AG2_HIP(hipMalloc(&c.g_fired, n)); AG2_HIP(hipMalloc(&c.g_final, n * 4));
}

static void radix_sort(Ctx &c, int n) {
    int nblk = (n + kSortTile - 1) / kSortTile;
    uint64_t *kin = c.key, *kout = c.key_alt;
    uint32_t *vin = c.idx, *vout = c.idx_alt;
    for (int pass = 0; pass < 8; pass++) {
        int shift = pass * 8;
        hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(256), 0, c.stream, n, kin, shift, c.hist, nblk);
        hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(256), 0, c.stream, nblk, c.hist, c.dtotal);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblk), dim3(256), 0, c.stream, n, kin, vin, kout, vout, shift, c.hist,
                           c.dtotal, nblk);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // eight passes: the result is back in c.key / c.idx
}

void build_tree(Ctx &c) {
    int n = (int)c.n;
    hipStream_t st = c.stream;
    if (c.timing) AG2_HIP(hipEventRecord(c.ev[0], st));
    int g = grid_for(n, 256);
    int twb = g < kTwBlocks ? (g > 0 ? g : 1) : kTwBlocks;
    hipLaunchKernelGGL(k_tw_partial, dim3(twb), dim3(256), 0, st, n, c.x, c.y, c.z, c.partials);
    hipLaunchKernelGGL(k_tw_final, dim3(1), dim3(256), 0, st, n, twb, c.partials, c.ds);
    hipLaunchKernelGGL(k_tw_max, dim3(twb), dim3(256), 0, st, n, c.x, c.y, c.z, c.ds);
    hipLaunchKernelGGL(k_keygen, dim3(g), dim3(256), 0, st, n, c.x, c.y, c.z, c.ds, c.key, c.idx, c.gapl);
    if (c.timing) AG2_HIP(hipEventRecord(c.ev[1], st));
    radix_sort(c, n);
    if (c.timing) AG2_HIP(hipEventRecord(c.ev[2], st));
    hipLaunchKernelGGL(k_count_in, dim3(g), dim3(256), 0, st, n, c.key, c.ds);
    hipLaunchKernelGGL(k_tie_key2, dim3(g), dim3(256), 0, st, n, c.key, c.idx, c.x, c.y, c.z, c.ds, c.key2, c.gapl);
    hipLaunchKernelGGL(k_tie_sort, dim3(g), dim3(256), 0, st, c.key, c.key2, c.idx, c.ds);
    hipLaunchKernelGGL(k_gather, dim3(g), dim3(256), 0, st, n, c.idx, c.x, c.y, c.z, c.m, c.type, c.vx, c.vy, c.vz,
                       c.sx, c.sy, c.sz, c.sm, c.stype, c.svx, c.svy, c.svz);
    int proot = (c.n >= 2 && c.n >= c.p.split_threshold) ? 1 : 0;
    hipLaunchKernelGGL(k_lcp, dim3(g), dim3(256), 0, st, c.key, c.key2, c.ds, proot, c.lcp, c.cnt);
    int nb = (n + 1023) / 1024;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(256), 0, st, c.cnt, c.ds, c.bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, st, nb, c.bsum, c.ds, proot, c.node_cap);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, st, c.cnt, c.bsum, c.ds, c.base);
    AG2_HIP(hipMemcpyAsync(&c.hs, c.ds, sizeof(DevState), hipMemcpyDeviceToHost, st));
    AG2_HIP(hipStreamSynchronize(st));
    if (c.hs.err) {
        // more nodes than room: grow and redo the node phase
        int64_t need = (int64_t)c.hs.n_int + c.hs.nin;
        alloc_nodes(c, need + need / 4 + 64);
        int zero = 0;
        AG2_HIP(hipMemcpy(&c.ds->err, &zero, sizeof(int), hipMemcpyHostToDevice));
    }
    c.nin = c.hs.nin;
    c.n_int = c.hs.n_int;
    c.n_nodes = c.hs.n_nodes;
    int nin = c.nin, nn = c.n_nodes;
    core::NodeOut o{c.n_first, c.n_count, c.n_skip, c.n_parent, c.n_pidx, c.n_level};
    NodeSums s{c.n_mass, c.n_sx, c.n_sy, c.n_sz, c.n_gas, c.n_gvx, c.n_gvy, c.n_gvz, c.n_cx, c.n_cy, c.n_cz};
    if (nn > 0) {
        hipLaunchKernelGGL(k_zero_nodes, dim3(grid_for(nn, 256)), dim3(256), 0, st, nn, c.n_nchild, c.n_arrive);
        if (nin == 0) {
            hipLaunchKernelGGL(k_make_root_only, dim3(1), dim3(1), 0, st, o);
        } else {
            int gi = grid_for(nin, 256);
            hipLaunchKernelGGL(k_make_internal, dim3(gi), dim3(256), 0, st, nin, nn, c.key, c.key2, c.lcp, c.base, o,
                               c.n_nchild);
            hipLaunchKernelGGL(k_make_leaf, dim3(gi), dim3(256), 0, st, nin, c.key, c.key2, c.lcp, c.base, c.cnt, o,
                               c.n_nchild, c.idx, c.gapl, c.ds);
            hipLaunchKernelGGL(k_moments_up, dim3(gi), dim3(256), 0, st, nin, c.n, c.p.split_threshold, c.base,
                               c.cnt, c.n_count, c.n_parent, c.n_skip, c.n_nchild, c.n_arrive, c.sx, c.sy, c.sz,
                               c.sm, c.stype, c.idx, c.vx, c.vy, c.vz, s);
        }
        if (proot && nin < n)
            hipLaunchKernelGGL(k_root_outside, dim3(1), dim3(256), 0, st, n, c.ds, c.sx, c.sy, c.sz, c.sm, c.stype,
                               c.idx, c.vx, c.vy, c.vz, s, nin == 0 ? 1 : 0);
        hipLaunchKernelGGL(k_finalize, dim3(grid_for(nn, 256)), dim3(256), 0, st, nn, c.ds, c.n_level, c.n_skip,
                           c.n_pidx, s, c.wnode);
    }
    if (c.timing) AG2_HIP(hipEventRecord(c.ev[3], st));
    AG2_HIP(hipGetLastError());
    c.tree_valid = true;
}

}  // namespace ag2
