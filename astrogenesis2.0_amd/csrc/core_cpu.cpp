// Sequential CPU build of the engine's per-thread core (tree_core.h), used by
// the CPU test suite to check the linear-octree construction, moments, SPH
// group selection and the FP64 walk against the oracle without a GPU.  The
// GPU path runs the same core functions inside its kernels; only the sort,
// scans and reductions differ (std::stable_sort and plain loops here).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "tree_core.h"

using namespace ag2::core;

extern "C" int64_t core_cpu_force_step(
    int64_t n, const double *x, const double *y, const double *z, const double *vx, const double *vy,
    const double *vz, const double *m, const double *U, const double *mu, const uint8_t *type, double *ax,
    double *ay, double *az, double *h, double *rho, double *P, double *T, double *dUdt, double theta,
    double e0, double mih, double G, int64_t thr, int64_t cap, int32_t *o_level, int64_t *o_first,
    int64_t *o_count, double *o_mass, double *o_com, double *o_gas, int64_t *o_perm, double *o_R) {
    // root half-width (Tree.cpp:86-117)
    std::vector<double> d((size_t)n);
    double sum = 0, sq = 0;
    for (int64_t i = 0; i < n; i++) { d[(size_t)i] = std::sqrt(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]); sum = sum + d[(size_t)i]; }
    for (int64_t i = 0; i < n; i++) sq = sq + d[(size_t)i] * d[(size_t)i];
    double mean = sum / (double)n;
    double cut = mean + 10 * std::sqrt(sq / (double)n - mean * mean);
    double R = 0;
    for (int64_t i = 0; i < n; i++) if (d[(size_t)i] <= cut && d[(size_t)i] > R) R = d[(size_t)i];
    if (o_R) *o_R = R;

    // keys (all six words for every particle here; the GPU computes deeper
    // words only where particles still share the shallower ones) and order
    std::vector<std::vector<uint64_t>> W(kKeyWords, std::vector<uint64_t>((size_t)n, 0));
    std::vector<char> in((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        double cx = 0.0, cy = 0.0, cz = 0.0, r = R;
        in[(size_t)i] = !(x[i] < -r || x[i] > r || y[i] < -r || y[i] > r || z[i] < -r || z[i] > r);
        if (!in[(size_t)i]) continue;
        for (int L = 1; L <= kDeepest; L++) {   // the descent of octant_path (Node.cpp:436-440, :713-716)
            int o = (x[i] > cx ? 1 : 0) | (y[i] > cy ? 2 : 0) | (z[i] > cz ? 4 : 0);
            uint64_t &w = W[(size_t)(L - 1) / 21][(size_t)i];
            w = (w << 3) | (uint64_t)o;
            cx = cx + r * ((o & 1) ? 0.5 : -0.5);
            cy = cy + r * ((o & 2) ? 0.5 : -0.5);
            cz = cz + r * ((o & 4) ? 0.5 : -0.5);
            r = r / 2;
        }
    }
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (in[(size_t)a] != in[(size_t)b]) return in[(size_t)a] > in[(size_t)b];
        if (!in[(size_t)a]) return false;
        for (int w = 0; w < kKeyWords; w++)
            if (W[(size_t)w][(size_t)a] != W[(size_t)w][(size_t)b]) return W[(size_t)w][(size_t)a] < W[(size_t)w][(size_t)b];
        return false;
    });
    int nin = 0;
    for (int64_t i = 0; i < n; i++) nin += in[(size_t)i];
    std::vector<std::vector<uint64_t>> kws(kKeyWords, std::vector<uint64_t>((size_t)n));
    std::vector<double> sx((size_t)n), sy((size_t)n), sz((size_t)n), sm((size_t)n), svx((size_t)n), svy((size_t)n), svz((size_t)n);
    std::vector<uint8_t> st((size_t)n);
    for (int64_t k = 0; k < n; k++) {
        int i = order[(size_t)k];
        for (int w = 0; w < kKeyWords; w++) kws[(size_t)w][(size_t)k] = W[(size_t)w][(size_t)i];
        sx[(size_t)k] = x[i]; sy[(size_t)k] = y[i]; sz[(size_t)k] = z[i]; sm[(size_t)k] = m[i];
        svx[(size_t)k] = vx[i]; svy[(size_t)k] = vy[i]; svz[(size_t)k] = vz[i]; st[(size_t)k] = type[i];
    }
    KeyWords kw;
    for (int w = 0; w < kKeyWords; w++) kw.w[w] = kws[(size_t)w].data();
    const Pos pos{sx.data(), sy.data(), sz.data()};
    bool proot = n >= 2 && n >= thr;

    // topology
    std::vector<int8_t> L((size_t)std::max(nin, 1));
    std::vector<int32_t> c((size_t)std::max(nin, 1) + 1), base((size_t)std::max(nin, 1) + 1);
    int n_int = 0;
    for (int i = 0; i < nin; i++) {
        int li = pair_lcp(kw, pos, i, nin, proot, proot && nin < n);
        // distinct particles closer than 2^-126 of the root half-width: the
        // engine refuses the build
        if (i + 1 < nin && li >= kDeepest && !same_pos(pos, i, i + 1)) return -2;
        L[(size_t)i] = (int8_t)li;
    }
    // adjacent key levels for the cell searches (as the GPU build forms them)
    std::vector<int8_t> Lk((size_t)std::max(nin, 1));
    for (int i = 0; i < nin; i++) Lk[(size_t)i] = i + 1 < nin ? (int8_t)shared_levels(kw, i, i + 1) : (int8_t)-1;
    for (int i = 0; i < nin; i++) {
        int v = (int)L[(size_t)i] - lcp_prev(L.data(), i);
        c[(size_t)i] = v > 0 ? v : 0;
        base[(size_t)i] = n_int;
        n_int += c[(size_t)i];
    }
    base[(size_t)nin] = n_int;
    if (nin == 0 && proot) n_int = 1;
    int nn = n_int + nin;
    std::vector<int32_t> first((size_t)nn + 1), count((size_t)nn + 1), skip((size_t)nn + 1), parent((size_t)nn + 1), pidx((size_t)nn + 1);
    std::vector<int8_t> level((size_t)nn + 1);
    NodeOut o{first.data(), count.data(), skip.data(), parent.data(), pidx.data(), level.data()};
    if (nin == 0 && proot) { first[0] = 0; count[0] = 0; skip[0] = 1; parent[0] = -1; pidx[0] = -1; level[0] = 0; }
    for (int i = 0; i < nin; i++)
        for (int l = lcp_prev(L.data(), i) + 1; l <= L[(size_t)i]; l++)
            make_internal(kw, L.data(), base.data(), i, l, nin, nn, o);
    std::vector<int> leaf_of((size_t)std::max(nin, 1));
    for (int i = 0; i < nin; i++) {
        int lv;
        make_leaf(kw, L.data(), Lk.data(), base.data(), c.data(), i, o, &lv);
        leaf_of[(size_t)i] = i + base[(size_t)i] + c[(size_t)i];
    }

    // moments: leaves, then reverse pre-order (children follow parents)
    std::vector<double> M((size_t)nn + 1, 0), Sx((size_t)nn + 1, 0), Sy((size_t)nn + 1, 0), Sz((size_t)nn + 1, 0), Gm((size_t)nn + 1, 0),
        Gx((size_t)nn + 1, 0), Gy((size_t)nn + 1, 0), Gz((size_t)nn + 1, 0), Cx((size_t)nn + 1, 0), Cy((size_t)nn + 1, 0), Cz((size_t)nn + 1, 0);
    for (int i = 0; i < nin; i++) {
        int nd = leaf_of[(size_t)i];
        double mm = sm[(size_t)i];
        bool gas = st[(size_t)i] == 2;
        M[(size_t)nd] = mm; Sx[(size_t)nd] = sx[(size_t)i] * mm; Sy[(size_t)nd] = sy[(size_t)i] * mm; Sz[(size_t)nd] = sz[(size_t)i] * mm;
        Gm[(size_t)nd] = gas ? mm : 0; Gx[(size_t)nd] = gas ? svx[(size_t)i] * mm : 0; Gy[(size_t)nd] = gas ? svy[(size_t)i] * mm : 0; Gz[(size_t)nd] = gas ? svz[(size_t)i] * mm : 0;
        int par = parent[(size_t)nd];
        bool vec = par < 0 ? (n == 1) : parallel_node(count.data(), parent.data(), par, n, thr);
        if (par >= 0 && !vec) {   // bucket members sit exactly at their position
            const bool whole = par == 0 && proot;
            vec = is_bucket(pos, whole ? 0 : first[(size_t)par], whole ? (int)n : first[(size_t)par] + count[(size_t)par]);
        }
        Cx[(size_t)nd] = leaf_com(sx[(size_t)i], mm, vec); Cy[(size_t)nd] = leaf_com(sy[(size_t)i], mm, vec); Cz[(size_t)nd] = leaf_com(sz[(size_t)i], mm, vec);
    }
    for (int nd = nn - 1; nd >= 0; nd--) {
        if (pidx[(size_t)nd] >= 0) continue;
        Sums s = sum_children(nd, skip.data(), M.data(), Sx.data(), Sy.data(), Sz.data(), Gm.data(), Gx.data(), Gy.data(), Gz.data());
        M[(size_t)nd] = s.m; Sx[(size_t)nd] = s.sx; Sy[(size_t)nd] = s.sy; Sz[(size_t)nd] = s.sz;
        Gm[(size_t)nd] = s.g; Gx[(size_t)nd] = s.gx; Gy[(size_t)nd] = s.gy; Gz[(size_t)nd] = s.gz;
    }
    // equal gas masses: the reference's running sum depends only on the count
    bool gas_tab_used = false;
    {
        int64_t ng = 0;
        double mg = 0.0;
        bool eq = true;
        for (int64_t i = 0; i < n; i++)
            if (type[i] == 2) { if (ng == 0) mg = m[i]; else if (m[i] != mg) eq = false; ng++; }
        if (ng > 0 && eq) {
            gas_tab_used = true;
            std::vector<double> tab((size_t)ng + 1, 0.0);
            for (int64_t k = 1; k <= ng; k++) tab[(size_t)k] = tab[(size_t)k - 1] + mg;
            std::vector<int64_t> gcount((size_t)nin + 1, 0);
            for (int i = 0; i < nin; i++) gcount[(size_t)i + 1] = gcount[(size_t)i] + (st[(size_t)i] == 2 ? 1 : 0);
            for (int nd = 0; nd < nn; nd++) {
                if (pidx[(size_t)nd] >= 0) continue;
                int64_t k = gcount[(size_t)(first[(size_t)nd] + count[(size_t)nd])] - gcount[(size_t)first[(size_t)nd]];
                if (nd == 0 && proot) k = ng;
                Gm[(size_t)nd] = tab[(size_t)k];
            }
        }
    }
    if (proot && nn > 0) {
        for (int64_t k = nin; k < n; k++) {
            double mm = sm[(size_t)k];
            bool gas = st[(size_t)k] == 2;
            M[0] += mm; Sx[0] += sx[(size_t)k] * mm; Sy[0] += sy[(size_t)k] * mm; Sz[0] += sz[(size_t)k] * mm;
            if (gas) { if (!gas_tab_used) Gm[0] += mm; Gx[0] += svx[(size_t)k] * mm; Gy[0] += svy[(size_t)k] * mm; Gz[0] += svz[(size_t)k] * mm; }
        }
    }
    std::vector<WalkNode> wn((size_t)nn + 1);
    std::vector<double> mvx((size_t)nn + 1), mvy((size_t)nn + 1), mvz((size_t)nn + 1);
    for (int nd = 0; nd < nn; nd++) {
        if (pidx[(size_t)nd] < 0 && M[(size_t)nd] > 0) { Cx[(size_t)nd] = Sx[(size_t)nd] / M[(size_t)nd]; Cy[(size_t)nd] = Sy[(size_t)nd] / M[(size_t)nd]; Cz[(size_t)nd] = Sz[(size_t)nd] / M[(size_t)nd]; }
        if (pidx[(size_t)nd] < 0) {   // a bucket sits exactly at its members' position
            const bool whole = nd == 0 && proot;
            const int bs = whole ? 0 : first[(size_t)nd], be = whole ? (int)n : first[(size_t)nd] + count[(size_t)nd];
            if (is_bucket(pos, bs, be)) { Cx[(size_t)nd] = sx[(size_t)bs]; Cy[(size_t)nd] = sy[(size_t)bs]; Cz[(size_t)nd] = sz[(size_t)bs]; }
        }
        if (Gm[(size_t)nd] > 0) { mvx[(size_t)nd] = Gx[(size_t)nd] / Gm[(size_t)nd]; mvy[(size_t)nd] = Gy[(size_t)nd] / Gm[(size_t)nd]; mvz[(size_t)nd] = Gz[(size_t)nd] / Gm[(size_t)nd]; }
        WalkNode &w = wn[(size_t)nd];
        w.cx = Cx[(size_t)nd]; w.cy = Cy[(size_t)nd]; w.cz = Cz[(size_t)nd]; w.mass = M[(size_t)nd];
        w.radius = std::ldexp(R, -level[(size_t)nd]); w.gas = Gm[(size_t)nd];
        w.skip = skip[(size_t)nd]; w.pidx = pidx[(size_t)nd]; w.level = level[(size_t)nd]; w.flags = 0;
    }

    // SPH group density (Tree.cpp:119-150)
    for (int64_t i = 0; i < n; i++) if (type[i] == 2) h[i] = 0;
    std::vector<int32_t> firstq((size_t)nn + 1, 0x7fffffff);
    for (int i = 0; i < nin; i++) {
        if (st[(size_t)i] != 2) continue;
        int g = gas_group(leaf_of[(size_t)i], parent.data(), Gm.data(), mih);
        if (g >= 0) firstq[(size_t)g] = std::min(firstq[(size_t)g], order[(size_t)i]);
    }
    std::vector<char> fired((size_t)nn + 1, 0);
    for (int nd = 0; nd < nn; nd++)
        if (firstq[(size_t)nd] != 0x7fffffff) fired[(size_t)nd] = group_fires(nd, parent.data(), firstq.data());
    std::vector<double> grho((size_t)nn + 1, -1.0);
    for (int i = 0; i < nin; i++) {
        if (st[(size_t)i] != 2) continue;
        int F = -1;
        for (int a = leaf_of[(size_t)i]; a >= 0; a = parent[(size_t)a]) if (fired[(size_t)a]) F = a;
        if (F < 0) continue;
        double hh = std::ldexp(R, -level[(size_t)F]) * 2;
        if (grho[(size_t)F] < 0) {
            double r = 0;
            for (int j = first[(size_t)F]; j < first[(size_t)F] + count[(size_t)F]; j++)
                if (st[(size_t)j] == 2) {
                    double dx = sx[(size_t)j] - Cx[(size_t)F], dy = sy[(size_t)j] - Cy[(size_t)F], dz = sz[(size_t)j] - Cz[(size_t)F];
                    r += sm[(size_t)j] * spline_w(std::sqrt(dx * dx + dy * dy + dz * dz), hh);
                }
            bool dup = count[(size_t)F] >= 2 && count[(size_t)F] < thr && parent[(size_t)F] >= 0 &&
                       parallel_node(count.data(), parent.data(), parent[(size_t)F], n, thr);
            grho[(size_t)F] = dup ? 2 * r : r;
        }
        int q = order[(size_t)i];
        h[q] = hh;
        rho[q] = grho[(size_t)F];
        P[q] = (5.0 / 3.0 - 1.0) * U[q] * rho[q];
        T[q] = (5.0 / 3.0 - 1.0) * U[q] * 1.6726219e-27 * mu[q] / (1.38064852e-23);
    }

    // forces (Tree::calculateForces), every particle active
    WalkParams wp{theta, e0, G, soft_branch_live(e0)};
    for (int64_t k = 0; k < n; k++) {
        int q = order[(size_t)k];
        Target t{x[q], y[q], z[q], m[q], vx[q], vy[q], vz[q], h[q], rho[q], P[q], type[q] == 2, k < nin ? (int)k : -1};
        WalkSums s{0, 0, 0, 0, 0, 0, 0};
        walk_fp64(wn.data(), nn, t, svx.data(), svy.data(), svz.data(), st.data(), mvx.data(), mvy.data(), mvz.data(), wp, 0, &s);
        ax[q] = s.ax; ay[q] = s.ay; az[q] = s.az;
        dUdt[q] += s.du;
    }

    // export
    if (nn > cap) return -1;
    for (int nd = 0; nd < nn; nd++) {
        o_level[nd] = level[(size_t)nd]; o_first[nd] = first[(size_t)nd]; o_count[nd] = count[(size_t)nd];
        o_mass[nd] = M[(size_t)nd]; o_com[3 * nd] = Cx[(size_t)nd]; o_com[3 * nd + 1] = Cy[(size_t)nd]; o_com[3 * nd + 2] = Cz[(size_t)nd];
        o_gas[nd] = Gm[(size_t)nd];
    }
    for (int i = 0; i < nin; i++) o_perm[i] = order[(size_t)i];
    return nn;
}

// One SPH pair term per case in both forms the engine has: sph_term, as the
// reference writes Node::calcSPHForce (Node.cpp:88-172), and the hoisted form
// the force walk evaluates on the GPU.  Each output holds ax ay az du per
// case; the pair is inside the support (0 < r < 2h) by the caller's choice.
extern "C" void core_cpu_sph_terms(int64_t n, const double *dx, const double *dy, const double *dz,
                                   const double *gas, const double *vx, const double *vy, const double *vz,
                                   const double *h, const double *rho, const double *P, double *plain,
                                   double *hoisted) {
    for (int64_t i = 0; i < n; i++) {
        Target t{};
        t.h = h[i]; t.rho = rho[i]; t.P = P[i];
        WalkSums a{0, 0, 0, 0, 0, 0, 0}, b{0, 0, 0, 0, 0, 0, 0};
        sph_term(dx[i], dy[i], dz[i], gas[i], vx[i], vy[i], vz[i], t, &a);
        const SphConst k = sph_const(t);
        sph_term_hoisted(dx[i], dy[i], dz[i], dx[i] * dx[i] + dy[i] * dy[i] + dz[i] * dz[i], gas[i], vx[i], vy[i],
                         vz[i], k, &b);
        const double pa[4] = {a.ax, a.ay, a.az, a.du}, pb[4] = {b.ax, b.ay, b.az, b.du};
        for (int j = 0; j < 4; j++) { plain[4 * i + j] = pa[j]; hoisted[4 * i + j] = pb[j]; }
    }
}
