// Per-thread logic of the linear octree, shared by the HIP kernels and the
// CPU self-check build (csrc/core_cpu.cpp, compiled with g++).  Everything
// here is plain arithmetic on arrays in key order; no GPU intrinsics.
//
// Tree model (the reference's pointer octree, Node.cpp:405-699, restated on
// sorted keys): particle i has an octant path of up to kDeepest = 126 levels
// in kKeyWords words of 21 levels (3 bits per level, the word's first level in
// bits 62..60); word w holds levels 21 w + 1 .. 21 w + 21.  Word 0 is sorted
// for every particle; a deeper word exists only for particles that share all
// shallower words with a neighbour (the reference keeps splitting until
// distinct particles separate, Node.cpp:597-699).  For adjacent particles in
// key order L[i] = number of shared leading levels of i and i+1.  An
// internal node (a cell holding >= 2 particles) at level l starting at
// particle s exists iff L[s-1] < l <= L[s]; pair i therefore owns
// c[i] = max(0, L[i] - L[i-1]) internal nodes (L[-1] = -1, L[nin-1] = -1).
// Numbering nodes in pre-order (children in octant order 0..7 -- the
// reference's walk order, Node.cpp:382-396) gives
//   internal (s, l): s + base[s] + (l - L[s-1] - 1)
//   leaf of particle i: i + base[i] + c[i]
// with base = exclusive scan of c.
//
// Particles at identical positions cannot be separated at any depth (the
// reference recurses until its bounds check drops them, Node.cpp:606-612).
// A run of them (contiguous in key order) is kept as a bucket: the cell at
// the level where the run becomes alone, one below its deepest level shared
// with a neighbouring particle, holds them as member leaves one level down,
// in caller order.  Pairs inside such a run get L = that bucket level.
#pragma once

#include <cstdint>
#include <cmath>

#ifdef __HIPCC__
#define AG2_HD __host__ __device__ inline
#else
#define AG2_HD inline
#endif

namespace ag2 {
namespace core {

constexpr int kKeyWords = 6;                   // words of 21 levels each
constexpr int kDeepest = 21 * kKeyWords;       // 126: deepest level that separates distinct particles

// The key words of the particles in key order; words this build did not
// compute are null (the particles that need them have none).
struct KeyWords {
    const uint64_t *w[kKeyWords];
};

// Positions in key order (identical positions form buckets).
struct Pos {
    const double *x, *y, *z;
};

AG2_HD bool same_pos(const Pos &p, int a, int b) {
    return p.x[a] == p.x[b] && p.y[a] == p.y[b] && p.z[a] == p.z[b];
}

// Levels shared by the paths of a and b, over the words present (21 times
// their count when all of them agree).
AG2_HD int shared_levels(const KeyWords &k, int a, int b) {
    int w = 0;
    for (; w < kKeyWords && k.w[w]; w++) {
        const uint64_t d = k.w[w][a] ^ k.w[w][b];
        if (d) return 21 * w + (__builtin_clzll(d) - 1) / 3;
    }
    return 21 * w;
}

// Do particles a and b (key order) share their first l levels (l <= 126)?
AG2_HD bool shares(const KeyWords &k, int a, int b, int l) {
    if (l <= 0) return true;
    const int full = (l - 1) / 21;        // words compared whole
    for (int w = 0; w < full; w++)
        if (k.w[w][a] != k.w[w][b]) return false;
    const int rem = l - 21 * full;        // 1..21 levels of word `full`
    return ((k.w[full][a] ^ k.w[full][b]) >> (63 - 3 * rem)) == 0;
}

// First / last particle of the run at the same position as particle i
// (galloping: such runs are contiguous in key order and usually short).
AG2_HD int same_pos_first(const Pos &p, int i) {
    int hi = i, lo = -1;   // same(hi, i); lo < answer
    for (int step = 1;; step <<= 1) {
        const int probe = i - step;
        if (probe < 0) break;
        if (!same_pos(p, probe, i)) { lo = probe; break; }
        hi = probe;
        if (step > (1 << 29)) break;
    }
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if (same_pos(p, mid, i)) hi = mid; else lo = mid;
    }
    return hi;
}

AG2_HD int same_pos_last(const Pos &p, int i, int nin) {
    int lo = i, hi = nin;   // same(i, lo); answer < hi
    for (int step = 1;; step <<= 1) {
        const int probe = i + step;
        if (probe >= nin) break;
        if (!same_pos(p, i, probe)) { hi = probe; break; }
        lo = probe;
        if (step > (1 << 29)) break;
    }
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if (same_pos(p, i, mid)) lo = mid; else hi = mid;
    }
    return lo;
}

// L[i]; nin == 1 uses the root rule: a root on the parallel path
// (Node.cpp:430-532) always splits, so its single particle sits at level 1.
// Inside a run of identical positions L is the run's bucket level; when the
// run is every particle in the tree, the root is the bucket unless it is on
// the parallel path with particles outside its cube (root_splits): it then
// splits first, as with a single particle.  A distinct pair that agrees on
// every word present returns 21 * words (the build refines such pairs with
// further words, or fails past kDeepest).
AG2_HD int pair_lcp(const KeyWords &k, const Pos &p, int i, int nin, bool parallel_root, bool root_splits) {
    if (i == nin - 1) return (nin == 1 && parallel_root) ? 0 : -1;
    if (same_pos(p, i, i + 1)) {
        const int s = same_pos_first(p, i), e = same_pos_last(p, i + 1, nin);
        const int lb = s > 0 ? shared_levels(k, s - 1, s) : -1;
        const int la = e + 1 < nin ? shared_levels(k, e, e + 1) : -1;
        const int b = (lb > la ? lb : la) + 1;
        return (b == 0 && root_splits) ? 1 : b;
    }
    return shared_levels(k, i, i + 1);
}

// Is the node over particles [s, e) a bucket (two or more particles, all at
// one position)?  Particles at one position are contiguous in key order, so
// the ends decide.  A bucket's centre of mass is that position exactly, and
// so is each member leaf's: a member then meets its bucket at r == 0 and
// skips it like the reference skips r == 0 (Node.cpp:274-276), so members
// exert no force on each other -- in the reference their cell's children are
// empty once its bounds check has dropped them -- and no rounding of the
// sums can turn r == 0 into a direction.
AG2_HD bool is_bucket(const Pos &p, int s, int e) { return e - s >= 2 && same_pos(p, s, e - 1); }

AG2_HD int lcp_prev(const int8_t *L, int i) { return i == 0 ? -1 : (int)L[i - 1]; }

// First particle of the level-l cell holding particle i.  Galloping search
// from i: cells are contiguous in key order and most are small, so the
// probes stay near i (O(log cell size) instead of O(log nin)).  hi0: a
// particle at or before i known to share the cell.
AG2_HD int cell_start(const KeyWords &k, int i, int l, int hi0 = -1) {
    if (l <= 0) return 0;
    if (hi0 < 0) hi0 = i;
    int hi = hi0, lo = -1;   // shares(hi, i); lo < answer
    for (int step = 1;; step <<= 1) {
        int probe = hi0 - step;
        if (probe < 0) break;
        if (!shares(k, probe, i, l)) { lo = probe; break; }
        hi = probe;
        if (step > (1 << 29)) break;
    }
    while (hi - lo > 1) {
        int mid = lo + ((hi - lo) >> 1);
        if (shares(k, mid, i, l)) hi = mid; else lo = mid;
    }
    return hi;
}

// Last particle of the level-l cell holding particle i (galloping, as above).
// lo0: a particle at or after i known to share the cell.
AG2_HD int cell_last(const KeyWords &k, int i, int l, int nin, int lo0 = -1) {
    if (l <= 0) return nin - 1;
    if (lo0 < 0) lo0 = i;
    int lo = lo0, hi = nin;  // shares(i, lo); answer < hi
    for (int step = 1;; step <<= 1) {
        int probe = lo0 + step;
        if (probe >= nin) break;
        if (!shares(k, i, probe, l)) { hi = probe; break; }
        lo = probe;
        if (step > (1 << 29)) break;
    }
    while (hi - lo > 1) {
        int mid = lo + ((hi - lo) >> 1);
        if (shares(k, i, mid, l)) lo = mid; else hi = mid;
    }
    return lo;
}

// The start search on the adjacent key levels Lk (Lk[j]: levels j and j + 1
// share, -1 after the last particle): keys in order share the minimum of the
// adjacent levels between them, so the level-l cell of i starts after the
// last j < i with Lk[j] < l.  A leaf's parent usually starts a few particles
// back: sixteen neighbours are read at once, and larger cells go on
// galloping over the keys from there.
constexpr int kNearScan = 32;

AG2_HD int cell_start_lk(const int8_t *Lk, const KeyWords &k, int i, int l) {
    if (l <= 0) return 0;
    int8_t v[kNearScan];
    for (int u = 0; u < kNearScan; u++) v[u] = i - 1 - u >= 0 ? Lk[i - 1 - u] : (int8_t)-1;
    for (int u = 0; u < kNearScan; u++)
        if (v[u] < l) return i - u;
    return cell_start(k, i, l, i - kNearScan);
}

AG2_HD int internal_index(const int8_t *L, const int32_t *base, int s, int l) {
    return s + base[s] + (l - lcp_prev(L, s) - 1);
}

struct NodeOut {
    int32_t *first, *count, *skip, *parent, *pidx;
    int8_t *level;
};

// Internal node of pair i at level l (L[i-1] < l <= L[i]); returns its parent.
AG2_HD int make_internal(const KeyWords &k, const int8_t *L, const int32_t *base,
                         int i, int l, int nin, int n_nodes, NodeOut o) {
    int lp = lcp_prev(L, i);
    int node = i + base[i] + (l - lp - 1);
    int last = cell_last(k, i, l, nin);
    int par;
    if (l == 0) par = -1;
    else if (l > lp + 1) par = node - 1;
    else par = internal_index(L, base, cell_start(k, i, l - 1), l - 1);
    o.first[node] = i;
    o.count[node] = last - i + 1;
    o.skip[node] = (last + 1 < nin) ? (last + 1 + base[last + 1]) : n_nodes;
    o.parent[node] = par;
    o.pidx[node] = -1;
    o.level[node] = (int8_t)l;
    return par;
}

// Leaf node of particle i; returns its parent.
AG2_HD int make_leaf(const KeyWords &k, const int8_t *L, const int8_t *Lk, const int32_t *base,
                     const int32_t *c, int i, NodeOut o, int *leaf_level) {
    int lp = lcp_prev(L, i), li = L[i];
    int lv = (lp > li ? lp : li) + 1;
    int node = i + base[i] + c[i];
    int par;
    if (lv == 0) par = -1;
    else if (c[i] > 0) par = node - 1;
    else par = internal_index(L, base, cell_start_lk(Lk, k, i, lv - 1), lv - 1);
    o.first[node] = i;
    o.count[node] = 1;
    o.skip[node] = node + 1;
    o.parent[node] = par;
    o.pidx[node] = i;
    o.level[node] = (int8_t)lv;
    *leaf_level = lv;
    return par;
}

// Node record of the block walk (child-contiguous layout), 64 bytes.
struct alignas(64) TreeNode {
    double cx, cy, cz, mass, gas;
    int32_t first_child;  // layout index of the first child, -1 for leaves
    int32_t pidx;         // sorted particle index of a leaf, -1 for internal nodes
    int32_t pre;          // pre-order index (moment arrays)
    int32_t nchild;
    int8_t level, flags;
    int16_t pad0;
    int32_t pad1;
};

// SPH view of a node in the walk layout (same index as its TreeNode): centre
// of mass, gas mass, gas velocity (a leaf: its particle's velocity; a node:
// the gas-mass-weighted mean), particle of a leaf.  64 bytes.
struct alignas(64) SphNode {
    double cx, cy, cz, gas, vx, vy, vz;
    int32_t pidx, pad;
};

// Is node `nd` (with parent pointer) on the reference's parallel insertion
// path?  The root sees all N particles (Tree.cpp:44-48); other nodes see the
// particles routed to them.  Node.cpp:409 (size 1 -> leaf) precedes :420.
AG2_HD bool parallel_node(const int32_t *count, const int32_t *parent, int nd, int64_t N, int64_t thr) {
    if (parent[nd] < 0) return N >= 2 && N >= thr;
    return count[nd] >= 2 && (int64_t)count[nd] >= thr;
}

// Centre of mass of a leaf: the vector path's size-1 branch copies the
// position (Node.cpp:409-417); the serial insert forms (0*(m-m) + x*m)/m
// (Node.cpp:698).
AG2_HD double leaf_com(double x, double m, bool vector_leaf) {
    return vector_leaf ? x : (x * m) / m;
}

// Node sums (mass, mass-weighted position, gas mass, gas-mass-weighted
// velocity) of an internal node from its children in octant order.
struct Sums { double m, sx, sy, sz, g, gx, gy, gz; };

AG2_HD Sums sum_children(int nd, const int32_t *skip, const double *m, const double *sx, const double *sy,
                         const double *sz, const double *g, const double *gx, const double *gy,
                         const double *gz) {
    Sums s{0, 0, 0, 0, 0, 0, 0, 0};
    int end = skip[nd];
    for (int c = nd + 1; c < end; c = skip[c]) {
        s.m += m[c]; s.sx += sx[c]; s.sy += sy[c]; s.sz += sz[c];
        s.g += g[c]; s.gx += gx[c]; s.gy += gy[c]; s.gz += gz[c];
    }
    return s;
}

// SPH group selection of Node::calcGasDensity (Node.cpp:722-796) from the
// leaf of a gas particle: climb while the parent's gas mass is strictly
// closer to massInH; the stop node is the group if it is strictly closer than
// its parent.  Returns the node or -1.
AG2_HD int gas_group(int leaf, const int32_t *parent, const double *gas, double mih) {
    int x = leaf;
    for (int guard = 0; guard <= kDeepest + 1; guard++) {
        double gx = gas[x];
        if (gx == 0) return -1;
        int p = parent[x];
        if (p < 0) return -1;
        double gp = gas[p];
        if (gx < mih && fabs(mih - gx) > fabs(mih - gp)) { x = p; continue; }
        return (fabs(mih - gx) < fabs(mih - gp) && gx != 0) ? x : -1;
    }
    return -1;
}

// A group fires when the first particle (in caller order) that selects it is
// not yet covered by a fired group above it (Tree.cpp:134-149 processes
// particles in array order and skips those whose h is already set).  Going
// down the selected ancestors, each fires iff its first selector precedes
// every one above it, and the covering index is the running minimum; so g
// fires iff its first selector precedes that of every ancestor.
AG2_HD bool group_fires(int g, const int32_t *parent, const int32_t *first_q) {
    int cover = 0x7fffffff;
    for (int a = parent[g]; a >= 0; a = parent[a]) cover = first_q[a] < cover ? first_q[a] : cover;
    return first_q[g] < cover;
}

// Cubic spline W (kernel.cpp:4-16) and its gradient scalar g/h/r
// (kernel.cpp:18-39), FP64.
AG2_HD double spline_w(double r, double h) {
    double a = 1.0 / (3.14159265358979323846 * h * h * h);
    double q = r / h;
    if (q < 1.0) return a * (1 - 1.5 * q * q + 0.75 * q * q * q);
    if (q < 2.0) { double t = 2 - q; return a * 0.25 * (t * t * t); }
    return 0.0;
}

AG2_HD double spline_grad_over_r(double r, double h) {
    if (r == 0.0) return 0.0;
    double a = 1.0 / (3.14159265358979323846 * h * h * h);
    double q = r / h;
    double g;
    if (q < 1.0) g = a * (-3.0 * q + 2.25 * q * q);
    else if (q < 2.0) { double t = 2 - q; g = a * (-0.75 * (t * t)); }
    else return 0.0;
    g /= h;
    return g / r;
}


// ------------------------------------------------------------ octant path
// Node::getOctant descent (Node.cpp:702-719) from the root (centre 0,
// half-width R) through `levels` levels (21 or 42).  Child centres are formed
// as c + r*(+-0.5) with r the parent half-width (Node.cpp:436-440); the bit is
// set iff the coordinate is strictly greater than the centre.  Returns false
// outside the root cube.  *gap = first level (>= 1) whose node bounds check
// (Node.cpp:706-711) fails, 0 if none.
AG2_HD bool octant_path(double px, double py, double pz, double R, int levels, uint64_t *k1, uint64_t *k2,
                        int *gap) {
    double cx = 0.0, cy = 0.0, cz = 0.0, r = R;
    *k1 = 0; *k2 = 0; *gap = 0;
    if (px < cx - r || px > cx + r || py < cy - r || py > cy + r || pz < cz - r || pz > cz + r) return false;
    uint64_t a = 0, b = 0;
    int g = 0;
    for (int L = 1; L <= levels; L++) {
        if (g == 0 && L > 1 &&
            (px < cx - r || px > cx + r || py < cy - r || py > cy + r || pz < cz - r || pz > cz + r))
            g = L - 1;
        int o = (px > cx ? 1 : 0) | (py > cy ? 2 : 0) | (pz > cz ? 4 : 0);
        if (L <= 21) a = (a << 3) | (uint64_t)o;
        else b = (b << 3) | (uint64_t)o;
        cx = cx + r * ((o & 1) ? 0.5 : -0.5);
        cy = cy + r * ((o & 2) ? 0.5 : -0.5);
        cz = cz + r * ((o & 4) ? 0.5 : -0.5);
        r = r / 2;
    }
    *k1 = a; *k2 = b; *gap = g;
    return true;
}

// Word w of the octant path (levels 21 w + 1 .. 21 w + 21) of a particle
// inside the root cube, by the descent of octant_path; *gap as there, over
// levels 1 .. 21 w + 21.
AG2_HD uint64_t octant_word(double px, double py, double pz, double R, int w, int *gap) {
    double cx = 0.0, cy = 0.0, cz = 0.0, r = R;
    uint64_t a = 0;
    int g = 0;
    const int last = 21 * (w + 1);
    for (int L = 1; L <= last; L++) {
        if (g == 0 && L > 1 &&
            (px < cx - r || px > cx + r || py < cy - r || py > cy + r || pz < cz - r || pz > cz + r))
            g = L - 1;
        int o = (px > cx ? 1 : 0) | (py > cy ? 2 : 0) | (pz > cz ? 4 : 0);
        if (L > 21 * w) a = (a << 3) | (uint64_t)o;
        cx = cx + r * ((o & 1) ? 0.5 : -0.5);
        cy = cy + r * ((o & 2) ? 0.5 : -0.5);
        cz = cz + r * ((o & 4) ? 0.5 : -0.5);
        r = r / 2;
    }
    *gap = g;
    return a;
}

// ------------------------------------------------------------ FP64 walk
// Node record read by the gravity walk: 64 bytes.
struct alignas(64) WalkNode {
    double cx, cy, cz;   // centre of mass
    double mass;
    double radius;       // half-width R / 2^level
    double gas;          // gas mass
    int32_t skip;        // pre-order index of the first node after this subtree
    int32_t pidx;        // sorted particle index of a leaf, -1 for internal nodes
    int32_t level;
    int32_t flags;
};

struct WalkParams {
    double theta, e0, G;
    int soft_live;       // 1 when the SYW softening branch can fire (see soft_force)
};

struct Target {
    double x, y, z, m;
    double vx, vy, vz;
    double h, rho, P;
    int gas;             // target is a gas particle
    int self;            // sorted index of the target's own leaf, -1 if none
};

struct WalkSums {
    double ax, ay, az, du, pot;
    long long nint, nsph;
};

// Springel-Yoshida-White kernel (kernel.cpp:41-56).
AG2_HD double soft_kernel(double u) {
    if (u >= 0 && u < 0.5)
        return (16.0 / 3.0) * (u * u) - (48.0 / 5.0) * (u * u * u * u) + (32.0 / 5.0) * (u * u * u * u * u) - (14.0 / 5.0);
    if (u >= 0.5 && u < 1.0)
        return (1.0 / (15.0 * u)) + (32.0 / 3.0) * (u * u) - (16.0) * (u * u * u) - (48.0 / 5.0) * (u * u * u * u) -
               (32.0 / 15.0) * (u * u * u * u * u) - (16.0 / 5.0);
    if (u >= 1.0) return -1.0 / u;
    return 0.0;
}

// The reference's abs(e) binds to int abs(int) (Node.cpp:302, :354): e is
// truncated to int; out-of-range conversions give INT_MIN on x86-64.
AG2_HD double int_abs(double e) {
    if (!(e > -2147483649.0 && e < 2147483648.0)) return -2147483648.0;
    long long t = (long long)e;
    if (t == -2147483648LL) return -2147483648.0;
    return (double)(t < 0 ? -t : t);
}

// Can the SYW branch of Node.cpp:302 / :354 fire for this e0?  It needs
// int_abs(e) > e0 * 0.0001, and int_abs is at most 2147483647 (INT_MIN maps
// to a negative value), so the branch is dead exactly when
// e0 * 0.0001 >= 2147483647 -- the reference's own product, not a rounded bound.
AG2_HD int soft_branch_live(double e0) { return (e0 * 0.0001) < 2147483647.0 ? 1 : 0; }

// Gravity of an accepted node (Node.cpp:285-313 / :337-365); false where the
// reference returns early.  Where soft_branch_live(e0) is 0 the SYW branch is
// dead and only G m/(r^2+e0^2) remains.
AG2_HD bool soft_force(double mass, double dx, double dy, double dz, double r, const WalkParams &wp,
                       double *gx, double *gy, double *gz) {
    double e0 = wp.e0;
    double f = wp.G * mass / (r * r + e0 * e0);
    if (wp.soft_live) {
        double kv = soft_kernel(r / (2.8 * e0));
        if ((kv - r) == 0) return false;
        double e = -(2.8 * e0) / (kv - r);
        if (e != e) return false;
        double ae = int_abs(e);
        if (ae > (e0 * 0.0001) && ae < 1e30) f = wp.G * mass / (r * r + e * e);
    }
    double l = sqrt(dx * dx + dy * dy + dz * dz);
    double nx = 0, ny = 0, nz = 0;
    if (l > 0) { nx = dx / l; ny = dy / l; nz = dz / l; }
    *gx = f * nx; *gy = f * ny; *gz = f * nz;
    return !(*gx != *gx || *gy != *gy || *gz != *gz);
}

// Node::calcSPHForce (Node.cpp:88-172): d = x_i - COM, h_j = h_i, rho_j = rho_i,
// P_j = P_i, alpha 0.5, beta 1, eta 0.01; v_j is the leaf particle's velocity
// or the node's gas-weighted mean velocity.
AG2_HD void sph_term(double dx, double dy, double dz, double gas, double vx, double vy, double vz,
                     const Target &t, WalkSums *s) {
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double h = t.h, rho = t.rho, P = t.P;
    double c = sqrt((5.0 / 3.0) * P / rho);
    double cij = (c + c) / 2.0;
    double hij = (h + h) / 2.0;
    double pr = P / (rho * rho) + P / (rho * rho);
    double gsc = spline_grad_over_r(r, h);
    double gx = dx * gsc, gy = dy * gsc, gz = dz * gsc;
    double ax = -gas * pr * gx, ay = -gas * pr * gy, az = -gas * pr * gz;
    double vd = vx * dx + vy * dy + vz * dz;
    double mu = hij * vd / (r * r + 0.01 * (hij * hij));
    double MU = 0.0;
    if (vd < 0) MU = -0.5 * cij * mu + 1.0 * (mu * mu);
    double gv = spline_grad_over_r(r, hij);
    ax += -gas * MU * (dx * gv);
    ay += -gas * MU * (dy * gv);
    az += -gas * MU * (dz * gv);
    s->du += 1.0 / 2.0 * gas * (pr + MU) * (vx * gx + vy * gy + vz * gz);
    if (ax != ax || ay != ay || az != az) return;
    s->ax += ax; s->ay += ay; s->az += az;
}

// Per-target constants of the SPH term (h_j = h_i, rho_j = rho_i, P_j = P_i in
// the reference, so everything except the pair geometry is per target).
struct SphConst {
    double h, a3, pr, cij, eta_h2, h2x4;
};

AG2_HD SphConst sph_const(const Target &t) {
    SphConst k;
    k.h = t.h;
    k.a3 = 1.0 / (3.14159265358979323846 * t.h * t.h * t.h);
    k.pr = t.P / (t.rho * t.rho) + t.P / (t.rho * t.rho);
    double c = sqrt((5.0 / 3.0) * t.P / t.rho);
    k.cij = (c + c) / 2.0;
    k.eta_h2 = 0.01 * (t.h * t.h);
    k.h2x4 = (t.h * 2) * (t.h * 2);
    return k;
}

// sph_term with the per-target constants hoisted (the same expressions, so
// the same values) and every other step in the reference's own arithmetic:
// a correctly rounded square root and quotients, h_ij = h_i (exactly: (h +
// h) / 2 == h), the pressure and viscosity parts added as two products.
// Near the kernel's edge (q -> 2) the term is (2 - q)^2, and with negative
// pressure p / rho^2 + MU can cancel, so a last-bit change in r or q would
// grow many times over; in this form the term equals sph_term bit for bit.
// d = x_i - COM, r2 = |d|^2 (0 < r < 2h), v = v_i - v_j; counts the pair.
AG2_HD void sph_term_hoisted(double dx, double dy, double dz, double r2, double gas, double vx, double vy,
                             double vz, const SphConst &k, WalkSums *s) {
    s->nsph++;
    const double r = sqrt(r2);
    const double q = r / k.h;
    const double poly = q < 1.0 ? (-3.0 * q + 2.25 * q * q) : -0.75 * ((2 - q) * (2 - q));
    const double gsc = ((k.a3 * poly) / k.h) / r;
    const double gx = dx * gsc, gy = dy * gsc, gz = dz * gsc;
    const double vd = vx * dx + vy * dy + vz * dz;
    const double mu = k.h * vd / (r * r + k.eta_h2);
    const double MU = vd < 0 ? -0.5 * k.cij * mu + 1.0 * (mu * mu) : 0.0;
    const double fp = -gas * k.pr, fv = -gas * MU;
    double ax = fp * gx, ay = fp * gy, az = fp * gz;
    ax += fv * gx;
    ay += fv * gy;
    az += fv * gz;
    s->du += 0.5 * gas * (k.pr + MU) * (vx * gx + vy * gy + vz * gz);
    if (ax != ax || ay != ay || az != az) return;
    s->ax += ax; s->ay += ay; s->az += az;
}

// Node::calculateGravityForce (Node.cpp:247-399) for one target as a
// stackless pre-order traversal: opening a node moves to the next node in
// pre-order (its first child), accepting or skipping moves past its subtree.
// mode 0: forces (+ SPH); mode 1: potential -(G m/e0)(pi/2 - atan(r/e0)).
AG2_HD void walk_fp64(const WalkNode *nodes, int n_nodes, const Target &t, const double *svx,
                      const double *svy, const double *svz, const uint8_t *stype, const double *mvx,
                      const double *mvy, const double *mvz, const WalkParams &wp, int mode, WalkSums *s) {
    if (t.m == 0) return;
    int n = 0;
    while (n < n_nodes) {
        const WalkNode &nd = nodes[n];
        int nxt = nd.skip > n ? nd.skip : n + 1;
        if (nd.mass == 0 || (nd.pidx >= 0 && nd.pidx == t.self)) { n = nxt; continue; }
        double dx = nd.cx - t.x, dy = nd.cy - t.y, dz = nd.cz - t.z;
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (r == 0) { n = nxt; continue; }
        bool leaf = nd.pidx >= 0;
        if (!leaf) {
            double sopen = nd.radius / r;
            if (!(sopen < wp.theta)) { n = n + 1; continue; }
        }
        int cur = n;
        n = nxt;
        if (mode == 1) {
            s->pot += -(wp.G * nd.mass / wp.e0) * (1.5707963267948966 - atan(r / wp.e0));
            s->nint++;
            continue;
        }
        double gx, gy, gz;
        if (!soft_force(nd.mass, dx, dy, dz, r, wp, &gx, &gy, &gz)) continue;
        s->ax += gx; s->ay += gy; s->az += gz;
        s->nint++;
        if (r < t.h * 2 && t.gas) {
            if (leaf) {
                int j = nd.pidx;
                if (stype[j] == 2) {
                    s->nsph++;
                    sph_term(-dx, -dy, -dz, nd.gas, t.vx - svx[j], t.vy - svy[j], t.vz - svz[j], t, s);
                }
            } else if (nd.gas > 0) {
                s->nsph++;
                sph_term(-dx, -dy, -dz, nd.gas, t.vx - mvx[cur], t.vy - mvy[cur], t.vz - mvz[cur], t, s);
            }
        }
    }
}

}  // namespace core
}  // namespace ag2
