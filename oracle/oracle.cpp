/*
 * CPU oracle for the AstroGen2 force engine -- TEST INFRASTRUCTURE ONLY.
 * See oracle.h for the pinning of each function.  Built with
 * -ffp-contract=off so that every product and sum rounds on its own, as the
 * reference's x86-64 build does.
 */
#include "oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* Constants.h:7, :10, :15-18 and Units.h:10 (km/s), Units.h:8 (Mpc). */
const double kPI = 3.14159265358979323846;
const double kGAMMA = 5.0 / 3.0;
const double kKB = 1.38064852e-23;
const double kPRTN = 1.6726219e-27;
const double kKMS = 1.0e3;
const double kMPC = 3.08567758149137e22;

/* vec3::length (vec3.cpp:76-78): sqrt(x*x + y*y + z*z), left to right. */
inline double len3(double x, double y, double z) { return std::sqrt(x * x + y * y + z * z); }

/* kernel::cubicSplineKernel (kernel.cpp:4-16). */
double spline_w(double r, double h) {
    double a = 1.0 / (kPI * h * h * h);
    double q = r / h;
    if (q < 1.0) return a * (1 - 1.5 * q * q + 0.75 * q * q * q);
    if (q < 2.0) return a * 0.25 * std::pow(2 - q, 3);
    return 0.0;
}

/* kernel::gradientCubicSplineKernel (kernel.cpp:18-39). */
void spline_grad(double dx, double dy, double dz, double h, double out[3]) {
    double rn = len3(dx, dy, dz);
    out[0] = out[1] = out[2] = 0.0;
    if (rn == 0.0) return;
    const double a = 1.0 / (kPI * h * h * h);
    double q = rn / h;
    double g;
    if (q < 1.0) g = a * (-3.0 * q + 2.25 * q * q);
    else if (q < 2.0) g = a * (-0.75 * std::pow(2 - q, 2));
    else return;
    g /= h;
    double s = g / rn;
    out[0] = dx * s; out[1] = dy * s; out[2] = dz * s;
}

/* kernel::softeningKernel (kernel.cpp:41-56). */
double soft_kernel(double u) {
    if (u >= 0 && u < 1.0 / 2.0)
        return (16.0 / 3.0) * std::pow(u, 2) - (48.0 / 5.0) * std::pow(u, 4) + (32.0 / 5.0) * std::pow(u, 5) - (14.0 / 5.0);
    if (u >= 1.0 / 2.0 && u < 1.0)
        return (1.0 / (15.0 * u)) + (32.0 / 3.0) * std::pow(u, 2) - (16.0) * std::pow(u, 3) - (48.0 / 5.0) * std::pow(u, 4) - (32.0 / 15.0) * std::pow(u, 5) - (16.0 / 5.0);
    if (u >= 1.0) return -1.0 / u;
    return 0.0;
}

/* Node.cpp:302 and :354 call abs() on a double with only the C library's
 * int abs(int) in scope, so e is truncated to int first; an out-of-range
 * conversion yields INT_MIN on x86-64, whose abs() is still INT_MIN. */
double int_abs(double e) {
    if (!(e > -2147483649.0 && e < 2147483648.0)) return -2147483648.0;
    long long t = (long long)e;
    if (t == -2147483648LL) return -2147483648.0;
    return (double)(t < 0 ? -t : t);
}

struct Node {
    double cx = 0, cy = 0, cz = 0, radius = 0;
    int depth = 0;
    double mass = 0, comx = 0, comy = 0, comz = 0;
    double gas = 0, mvx = 0, mvy = 0, mvz = 0;
    int child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    std::vector<int> kids;  /* member leaves of a bucket (particles at one position) */
    int parent = -1;
    int particle = -1;
    bool leaf = true;
    std::vector<int> list;  /* Node::childParticles */
};

struct Tree {
    const orc_particles *p;
    const orc_params *prm;
    std::vector<Node> nodes;
    std::vector<int> leafOf;  /* Particle::node; -1 == nullptr */
    double R = 0;
    bool err_depth = false;   /* distinct particles share more than max_level levels */

    int new_node(double cx, double cy, double cz, double radius, int depth, int parent) {
        Node n;
        n.cx = cx; n.cy = cy; n.cz = cz; n.radius = radius; n.depth = depth; n.parent = parent;
        nodes.push_back(std::move(n));
        return (int)nodes.size() - 1;
    }

    /* Node.cpp:433-443 / :630-642: all eight children, centres c + r*(+-0.5). */
    void make_children(int ni) {
        for (int i = 0; i < 8; i++) {
            double r = nodes[ni].radius;
            int c = new_node(nodes[ni].cx + r * ((i & 1) ? 0.5 : -0.5),
                             nodes[ni].cy + r * ((i & 2) ? 0.5 : -0.5),
                             nodes[ni].cz + r * ((i & 4) ? 0.5 : -0.5),
                             r / 2, nodes[ni].depth + 1, ni);
            nodes[ni].child[i] = c;
        }
    }

    bool outside(int ni, int q) const {
        const Node &n = nodes[ni];
        double x = p->x[q], y = p->y[q], z = p->z[q];
        return x < n.cx - n.radius || x > n.cx + n.radius ||
               y < n.cy - n.radius || y > n.cy + n.radius ||
               z < n.cz - n.radius || z > n.cz + n.radius;
    }

    /* Node::getOctant (Node.cpp:702-719). */
    int octant(int ni, int q) const {
        if (outside(ni, q)) return -1;
        const Node &n = nodes[ni];
        int o = 0;
        if (p->x[q] > n.cx) o |= 1;
        if (p->y[q] > n.cy) o |= 2;
        if (p->z[q] > n.cz) o |= 4;
        return o;
    }

    /* Running moment update of the serial insert (Node.cpp:678-698). */
    void serial_moments(int ni, int q) {
        Node &n = nodes[ni];
        double m = p->m[q];
        n.mass += m;
        if (p->type[q] == 2) {
            n.gas += m;
            if (n.gas > 0) {
                n.mvx = (n.mvx * (n.gas - m) + p->vx[q] * m) / n.gas;
                n.mvy = (n.mvy * (n.gas - m) + p->vy[q] * m) / n.gas;
                n.mvz = (n.mvz * (n.gas - m) + p->vz[q] * m) / n.gas;
            }
        }
        n.comx = (n.comx * (n.mass - m) + p->x[q] * m) / n.mass;
        n.comy = (n.comy * (n.mass - m) + p->y[q] * m) / n.mass;
        n.comz = (n.comz * (n.mass - m) + p->z[q] * m) / n.mass;
    }

    bool same_pos(int a, int b) const {
        return p->x[a] == p->x[b] && p->y[a] == p->y[b] && p->z[a] == p->z[b];
    }

    /* Node::insert(Particle*) (Node.cpp:597-699).  The reference splits a
     * leaf until its two particles fall into different octants; for
     * particles at the same position that never happens (it recurses until
     * its bounds check drops them, Node.cpp:606-612), so here:
     *  - a particle at the exact position of a leaf's particle makes the leaf
     *    a bucket: both become member leaves one level below;
     *  - a bucket receiving a particle at another position becomes an
     *    ordinary node: its members, then the new particle, are inserted into
     *    their octant children in that order -- the order in which the
     *    reference's split routes a leaf's particle and then the newcomer
     *    (Node.cpp:646-662);
     *  - distinct particles that would need a split below max_level set
     *    err_depth (the engine fails with a capacity error there). */
    void insert_one(int ni, int q) {
        if (outside(ni, q)) return;
        nodes[ni].list.push_back(q);
        if (!nodes[ni].kids.empty()) {
            const int first = nodes[(size_t)nodes[ni].kids[0]].particle;
            if (same_pos(first, q)) add_member(ni, q);
            else unbucket(ni, q);
        } else if (nodes[ni].leaf) {
            if (nodes[ni].particle < 0) {
                nodes[ni].particle = q;
                leafOf[q] = ni;
            } else if (same_pos(nodes[ni].particle, q) || nodes[ni].depth >= prm->max_level) {
                if (!same_pos(nodes[ni].particle, q)) err_depth = true;
                const int old = nodes[ni].particle;
                nodes[ni].leaf = false;
                nodes[ni].particle = -1;
                add_member(ni, old);
                add_member(ni, q);
            } else {
                make_children(ni);
                int old = nodes[ni].particle;
                int o = octant(ni, old);
                if (o != -1) insert_one(nodes[ni].child[o], old);
                o = octant(ni, q);
                if (o != -1) insert_one(nodes[ni].child[o], q);
                nodes[ni].leaf = false;
                nodes[ni].particle = -1;
            }
        } else {
            int o = octant(ni, q);
            if (o != -1) insert_one(nodes[ni].child[o], q);
        }
        serial_moments(ni, q);
    }

    /* A bucket meets a particle at another position: split it like a leaf. */
    void unbucket(int ni, int q) {
        if (nodes[ni].depth >= prm->max_level) {
            err_depth = true;
            add_member(ni, q);
            return;
        }
        std::vector<int> mem;
        for (int c : nodes[ni].kids) mem.push_back(nodes[(size_t)c].particle);
        nodes[ni].kids.clear();
        make_children(ni);
        for (int m : mem) {
            leafOf[m] = -1;
            int o = octant(ni, m);
            if (o != -1) insert_one(nodes[ni].child[o], m);
        }
        int o = octant(ni, q);
        if (o != -1) insert_one(nodes[ni].child[o], q);
    }

    void add_member(int ni, int q) {
        double r = nodes[ni].radius;
        int c = new_node(nodes[ni].cx, nodes[ni].cy, nodes[ni].cz, r / 2, nodes[ni].depth + 1, ni);
        nodes[ni].kids.push_back(c);
        nodes[c].list.push_back(q);
        nodes[c].particle = q;
        leafOf[q] = c;
        serial_moments(c, q);
    }

    /* A member of a bucket on the parallel path: the size-1 branch of
     * Node::insert(vector) (Node.cpp:409-417) -- centre of mass = position. */
    void add_vector_member(int ni, int q) {
        double r = nodes[ni].radius;
        int c = new_node(nodes[ni].cx, nodes[ni].cy, nodes[ni].cz, r / 2, nodes[ni].depth + 1, ni);
        nodes[ni].kids.push_back(c);
        nodes[c].list.push_back(q);
        Node &n = nodes[c];
        n.particle = q;
        n.comx = p->x[q]; n.comy = p->y[q]; n.comz = p->z[q];
        n.mass = p->m[q];
        n.gas = (p->type[q] == 2) ? p->m[q] : 0.0;
        leafOf[q] = c;
    }

    /* Node::insert(vector, cores) (Node.cpp:405-534), executed by one thread
     * in list order (the reference's OpenMP partition is order-free only with
     * one thread). */
    void insert_vec(int ni, const std::vector<int> &plist) {
        if (plist.empty()) return;
        if (plist.size() == 1) {
            Node &n = nodes[ni];
            int q = plist[0];
            n.leaf = true;
            n.particle = q;
            leafOf[q] = ni;
            n.comx = p->x[q]; n.comy = p->y[q]; n.comz = p->z[q];
            n.mass = p->m[q];
            n.gas = (p->type[q] == 2) ? p->m[q] : 0.0;
            return;
        }
        if ((int64_t)plist.size() < prm->split_threshold || nodes[ni].depth >= prm->max_level) {
            for (int q : plist) insert_one(ni, q);
            return;
        }
        bool all_same = true;
        for (int q : plist) all_same = all_same && same_pos(plist[0], q);
        nodes[ni].leaf = false;
        if (!all_same) make_children(ni);
        double tm = 0, tg = 0, sx = 0, sy = 0, sz = 0;
        std::vector<int> sub[8];
        for (int q : plist) {
            if (nodes[ni].depth == 0) leafOf[q] = -1;
            double m = p->m[q];
            tm += m;
            if (p->type[q] == 2) {
                tg += m;
                Node &n = nodes[ni];
                n.gas += m;
                if (n.gas > 0) {
                    n.mvx = (n.mvx * (n.gas - m) + p->vx[q] * m) / n.gas;
                    n.mvy = (n.mvy * (n.gas - m) + p->vy[q] * m) / n.gas;
                    n.mvz = (n.mvz * (n.gas - m) + p->vz[q] * m) / n.gas;
                }
            }
            sx += p->x[q] * m;
            sy += p->y[q] * m;
            sz += p->z[q] * m;
            if (all_same) continue;
            int o = octant(ni, q);
            if (o != -1) sub[o].push_back(q);
        }
        Node &n = nodes[ni];
        n.mass = tm;
        n.gas = tg;
        if (n.mass > 0.0) { n.comx = sx / n.mass; n.comy = sy / n.mass; n.comz = sz / n.mass; }
        if (all_same) {
            /* particles at one position: a bucket of vector-path members */
            for (int q : plist) add_vector_member(ni, q);
            return;
        }
        for (int o = 0; o < 8; o++) {
            int c = nodes[ni].child[o];
            nodes[c].list.insert(nodes[c].list.end(), sub[o].begin(), sub[o].end());
        }
        for (int o = 0; o < 8; o++) {
            int c = nodes[ni].child[o];
            if (!nodes[c].list.empty()) {
                std::vector<int> l = nodes[c].list;
                insert_vec(c, l);
            }
        }
    }

    /* Tree::buildTree (Tree.cpp:24-55): root at the origin, half-width R. */
    void build() {
        int64_t n = p->n;
        nodes.clear();
        nodes.reserve((size_t)(4 * n + 16));
        leafOf.assign((size_t)n, -1);
        R = orc_tree_width(n, p->x, p->y, p->z);
        new_node(0.0, 0.0, 0.0, R, 0, -1);
        std::vector<int> all((size_t)n);
        for (int64_t i = 0; i < n; i++) all[(size_t)i] = (int)i;
        insert_vec(0, all);
        /* A bucket and its member leaves sit exactly at the members' common
         * position (whatever the running average left): a member meets its
         * bucket at r == 0 and skips it (Node.cpp:274-276), as the reference's
         * dropped coincident particles meet only empty children. */
        for (Node &b : nodes) {
            if (b.kids.empty()) continue;
            const int q = nodes[(size_t)b.kids[0]].particle;
            b.comx = p->x[q]; b.comy = p->y[q]; b.comz = p->z[q];
            for (int k : b.kids) {
                Node &mb = nodes[(size_t)k];
                mb.comx = p->x[q]; mb.comy = p->y[q]; mb.comz = p->z[q];
            }
        }
    }

    /* Node::calcGasDensity (Node.cpp:722-796), single-threaded. */
    void gas_density(int ni, double *h, double *rho, double *P, double *T, int64_t *fired) {
        const Node &n = nodes[ni];
        double mih = prm->massInH;
        if (n.gas == 0) return;
        if (n.parent < 0) return;
        const Node &par = nodes[n.parent];
        if (n.gas < mih) {
            double md = mih - n.gas;
            double pmd = mih - par.gas;
            if (std::abs(md) > std::abs(pmd)) gas_density(n.parent, h, rho, P, T, fired);
        }
        double md = mih - n.gas;
        double pmd = mih - par.gas;
        if (std::abs(md) < std::abs(pmd) && n.gas != 0) {
            (*fired)++;
            for (int q : n.list)
                if (p->type[q] == 2) h[q] = n.radius * 2;
            double r = 0;
            for (int q : n.list)
                if (p->type[q] == 2)
                    r += p->m[q] * spline_w(len3(p->x[q] - n.comx, p->y[q] - n.comy, p->z[q] - n.comz), h[q]);
            for (int q : n.list)
                if (p->type[q] == 2) {
                    rho[q] = r;
                    P[q] = (kGAMMA - 1.0) * p->U[q] * rho[q];
                    T[q] = (kGAMMA - 1.0) * p->U[q] * kPRTN * p->mu[q] / (kKB);
                }
        }
    }

    /* Node::calcSPHForce (Node.cpp:88-172) for target q against node ni. */
    /* acc[0..2]: the acceleration; acc[3]: the sum of the magnitudes of the
     * contributions added (the scale of the rounding of the sum). */
    void sph_force(int ni, int q, double acc[4], double *dudt) const {
        const Node &n = nodes[ni];
        double h_i = p->h[q];
        double rho_i = p->rho[q];
        double P_i = p->P[q];
        double vjx = n.mvx, vjy = n.mvy, vjz = n.mvz;
        if (n.leaf) { int j = n.particle; vjx = p->vx[j]; vjy = p->vy[j]; vjz = p->vz[j]; }
        double vx = p->vx[q] - vjx, vy = p->vy[q] - vjy, vz = p->vz[q] - vjz;
        double dx = p->x[q] - n.comx, dy = p->y[q] - n.comy, dz = p->z[q] - n.comz;
        double r = len3(dx, dy, dz);
        double c = std::sqrt(kGAMMA * P_i / rho_i);
        double c_ij = (c + c) / 2.0;
        double h_ij = (h_i + h_i) / 2.0;
        double pr = P_i / (rho_i * rho_i) + P_i / (rho_i * rho_i);
        double g[3];
        spline_grad(dx, dy, dz, h_i, g);
        double ax = -n.gas * pr * g[0], ay = -n.gas * pr * g[1], az = -n.gas * pr * g[2];
        double vd = vx * dx + vy * dy + vz * dz;
        double mu = h_ij * vd / (r * r + 0.01 * (h_ij * h_ij));
        double MU = 0.0;
        if (vd < 0) MU = -0.5 * c_ij * mu + 1.0 * (mu * mu);
        double gv[3];
        spline_grad(dx, dy, dz, h_ij, gv);
        ax += -n.gas * MU * gv[0];
        ay += -n.gas * MU * gv[1];
        az += -n.gas * MU * gv[2];
        *dudt += 1.0 / 2.0 * n.gas * (pr + MU) * (vx * g[0] + vy * g[1] + vz * g[2]);
        if (std::isnan(ax) || std::isnan(ay) || std::isnan(az)) return;
        acc[0] += ax; acc[1] += ay; acc[2] += az;
        acc[3] += len3(ax, ay, az);
    }

    /* Gravity of an accepted node (Node.cpp:285-313 / :337-365). Returns false
     * when the reference returns early (no gravity, no SPH). */
    bool gravity(const Node &n, double dx, double dy, double dz, double r, double acc[4]) const {
        double e0 = prm->e0;
        double u = r / (2.8 * e0);
        double kv = soft_kernel(u);
        if ((kv - r) == 0) return false;
        double e = -(2.8 * e0) / (kv - r);
        if (std::isnan(e)) return false;
        double f = prm->G * n.mass / (r * r + e0 * e0);
        double ae = int_abs(e);
        if (ae > (e0 * 0.0001) && ae < 1e30) f = prm->G * n.mass / (r * r + e * e);
        double l = len3(dx, dy, dz);
        double nx = 0, ny = 0, nz = 0;
        if (l > 0) { nx = dx / l; ny = dy / l; nz = dz / l; }
        double gx = f * nx, gy = f * ny, gz = f * nz;
        if (std::isnan(gx) || std::isnan(gy) || std::isnan(gz)) return false;
        acc[0] += gx; acc[1] += gy; acc[2] += gz;
        acc[3] += len3(gx, gy, gz);
        return true;
    }

    /* Node::calculateGravityForce (Node.cpp:247-399). */
    void walk(int ni, int q, double acc[4], double *dudt, int64_t *nint, int64_t *nsph) const {
        const Node &n = nodes[ni];
        if (n.mass == 0) return;
        if (n.particle == q && n.leaf) return;
        if (p->m[q] == 0) return;
        double dx = n.comx - p->x[q], dy = n.comy - p->y[q], dz = n.comz - p->z[q];
        double r = len3(dx, dy, dz);
        if (r == 0) return;
        if (n.leaf) {
            if (n.particle >= 0) {
                if (!gravity(n, dx, dy, dz, r, acc)) return;
                (*nint)++;
                if (r < p->h[q] * 2 && p->type[n.particle] == 2 && p->type[q] == 2) {
                    (*nsph)++;
                    sph_force(ni, q, acc, dudt);
                }
            }
            return;
        }
        double s = n.radius / r;
        if (s < prm->theta) {
            if (!gravity(n, dx, dy, dz, r, acc)) return;
            (*nint)++;
            if (r < p->h[q] * 2 && p->type[q] == 2 && n.gas > 0) {
                (*nsph)++;
                sph_force(ni, q, acc, dudt);
            }
            return;
        }
        if (!n.kids.empty()) {
            for (int c : n.kids)
                if (nodes[c].mass != 0) walk(c, q, acc, dudt, nint, nsph);
            return;
        }
        for (int o = 0; o < 8; o++) {
            int c = n.child[o];
            if (c < 0) continue;
            if (nodes[c].mass == 0) continue;
            walk(c, q, acc, dudt, nint, nsph);
        }
    }

    /* Node::calcVisualDensity (Node.cpp:832-877): climb while the parent's
     * half-width is strictly closer to r, then mass / radius^3 into every
     * particle listed under the node (nothing at the root, nothing for a zero
     * or infinite density). */
    void visual_density(int ni, double r, double *vd) const {
        const Node &n = nodes[ni];
        if (n.parent < 0) return;
        const Node &par = nodes[n.parent];
        if (std::abs(r - n.radius) > std::abs(r - par.radius)) {
            visual_density(n.parent, r, vd);
            return;
        }
        double vol = n.radius * n.radius * n.radius;
        if (vol == 0 || n.mass == 0) return;
        double d = n.mass / vol;
        if (d == 0 || d == std::numeric_limits<double>::infinity()) return;
        for (int q : n.list) vd[q] = d;
    }

    /* Pre-order export of non-empty nodes (children 0..7, bucket members in
     * index order).  Returns the number of in-tree particles below ni. */
    int64_t count_below(int ni, std::vector<int64_t> &cnt) const {
        const Node &n = nodes[ni];
        int64_t c = 0;
        if (n.leaf) c = (n.particle >= 0) ? 1 : 0;
        else if (!n.kids.empty()) { for (int k : n.kids) c += count_below(k, cnt); }
        else for (int o = 0; o < 8; o++) if (n.child[o] >= 0) c += count_below(n.child[o], cnt);
        cnt[(size_t)ni] = c;
        return c;
    }
};

}  // namespace

extern "C" {

double orc_tree_width(int64_t n, const double *x, const double *y, const double *z) {
    /* Tree.cpp:86-117: std::accumulate and std::inner_product run in array
     * order; mean + 10 sigma cut; largest accepted distance. */
    const double acceptance = 10;
    if (n == 0) return 0;
    std::vector<double> d((size_t)n);
    for (int64_t i = 0; i < n; i++) d[(size_t)i] = len3(x[i], y[i], z[i]);
    double sum = 0.0;
    for (int64_t i = 0; i < n; i++) sum = sum + d[(size_t)i];
    double mean = sum / (double)(int)n;
    double sq = 0.0;
    for (int64_t i = 0; i < n; i++) sq = sq + d[(size_t)i] * d[(size_t)i];
    double sd = std::sqrt(sq / (double)(int)n - mean * mean);
    double cut = mean + acceptance * sd;
    double mx = 0;
    for (int64_t i = 0; i < n; i++)
        if (d[(size_t)i] <= cut && d[(size_t)i] > mx) mx = d[(size_t)i];
    return mx;
}

int orc_force_step(orc_particles *p, const orc_params *prm, orc_stats *st) {
    return orc_force_step_scale(p, prm, st, nullptr);
}

/* Non-empty nodes of a built tree in the pre-order of orc_tree_export (and
 * of the engine's node arrays): children in octant order, bucket members by
 * particle index. */
static std::vector<int> export_order(const Tree &t, std::vector<int64_t> &cnt, int64_t *total) {
    cnt.assign(t.nodes.size(), 0);
    int64_t tot = t.count_below(0, cnt);
    if (total) *total = tot;
    std::vector<int> order, stack;
    stack.push_back(0);
    while (!stack.empty()) {
        int ni = stack.back();
        stack.pop_back();
        const Node &n = t.nodes[(size_t)ni];
        if (cnt[(size_t)ni] == 0) continue;
        order.push_back(ni);
        if (n.leaf) continue;
        if (!n.kids.empty()) {
            std::vector<int> ks = n.kids;
            std::sort(ks.begin(), ks.end(), [&](int a, int b) { return t.nodes[a].particle < t.nodes[b].particle; });
            for (int i = (int)ks.size() - 1; i >= 0; i--) stack.push_back(ks[(size_t)i]);
        } else {
            for (int o = 7; o >= 0; o--)
                if (n.child[o] >= 0) stack.push_back(n.child[o]);
        }
    }
    return order;
}

static int force_phase(Tree &t, orc_particles *p, const orc_params *prm, orc_stats *st, double *scale);

int orc_force_step_scale(orc_particles *p, const orc_params *prm, orc_stats *st, double *scale) {
    Tree t;
    t.p = p;
    t.prm = prm;
    t.build();
    int rc = force_phase(t, p, prm, st, scale);
    return t.err_depth ? -2 : rc;
}

int orc_force_step_moments(orc_particles *p, const orc_params *prm, orc_stats *st, double *scale, int64_t nn,
                           const double *mass, const double *com, const double *gas, const double *mvel) {
    Tree t;
    t.p = p;
    t.prm = prm;
    t.build();
    std::vector<int64_t> cnt;
    std::vector<int> order = export_order(t, cnt, nullptr);
    if ((int64_t)order.size() != nn) return -1;
    for (int64_t k = 0; k < nn; k++) {
        Node &n = t.nodes[(size_t)order[(size_t)k]];
        n.mass = mass[k];
        n.comx = com[3 * k]; n.comy = com[3 * k + 1]; n.comz = com[3 * k + 2];
        n.gas = gas[k];
        n.mvx = mvel[3 * k]; n.mvy = mvel[3 * k + 1]; n.mvz = mvel[3 * k + 2];
    }
    int rc = force_phase(t, p, prm, st, scale);
    return t.err_depth ? -2 : rc;
}

static int force_phase(Tree &t, orc_particles *p, const orc_params *prm, orc_stats *st, double *scale) {
    int64_t n = p->n;
    /* Tree::calcGasDensity (Tree.cpp:119-150): reset h, then array order. */
    for (int64_t i = 0; i < n; i++)
        if (p->type[i] == 2) p->h[i] = 0;
    int64_t fired = 0;
    for (int64_t i = 0; i < n; i++) {
        if (p->type[i] != 2) continue;
        if (t.leafOf[(size_t)i] < 0) continue;
        if (p->h[i] != 0) continue;
        t.gas_density(t.leafOf[(size_t)i], p->h, p->rho, p->P, p->T, &fired);
    }
    /* Tree::calculateForces (Tree.cpp:57-83). */
    int64_t nint = 0, nsph = 0;
#ifdef _OPENMP
    int nth = prm->nthreads > 0 ? prm->nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nth) reduction(+ : nint, nsph)
#endif
    for (int64_t i = 0; i < n; i++) {
        if (p->active && !p->active[i]) continue;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        double du = p->dUdt[i];
        int64_t a = 0, b = 0;
        t.walk(0, (int)i, acc, &du, &a, &b);
        p->ax[i] = acc[0]; p->ay[i] = acc[1]; p->az[i] = acc[2];
        if (scale) scale[i] = acc[3];
        p->dUdt[i] = du;
        nint += a; nsph += b;
    }
    if (st) {
        st->n_nodes = (int64_t)t.nodes.size();
        int64_t k = 0;
        for (int64_t i = 0; i < n; i++) k += t.leafOf[(size_t)i] >= 0;
        st->n_in_tree = k;
        st->interactions = nint;
        st->sph_interactions = nsph;
        st->groups_fired = fired;
        st->R = t.R;
    }
    return 0;
}

int orc_visual_density(const orc_particles *p, const orc_params *prm, double radius, double *vd, double *R_out) {
    Tree t;
    t.p = p;
    t.prm = prm;
    t.build();
    /* Simulation.cpp:126: r = root half-width / 100000 unless given. */
    double r = radius > 0 ? radius : t.R / 100000;
    /* Tree::calcVisualDensity (Tree.cpp:152-176): reset, then array order,
     * skipping particles already written. */
    for (int64_t i = 0; i < p->n; i++) vd[i] = 0;
    for (int64_t i = 0; i < p->n; i++) {
        if (vd[i] != 0) continue;
        int leaf = t.leafOf[(size_t)i];
        if (leaf >= 0) t.visual_density(leaf, r, vd);
    }
    if (R_out) *R_out = t.R;
    return 0;
}

int64_t orc_tree_export(const orc_particles *p, const orc_params *prm, int64_t cap,
                        int32_t *level, int64_t *first, int64_t *count,
                        double *mass, double *com, double *gas, double *mvel,
                        int64_t *perm, int64_t *n_in_tree) {
    Tree t;
    t.p = p;
    t.prm = prm;
    t.build();
    std::vector<int64_t> cnt;
    int64_t total = 0;
    if (t.err_depth) return -2;
    std::vector<int> order = export_order(t, cnt, &total);
    if (n_in_tree) *n_in_tree = total;
    int64_t k = 0, leafpos = 0;
    for (int ni : order) {
        const Node &n = t.nodes[(size_t)ni];
        if (k >= cap) return -1;
        level[k] = n.depth;
        first[k] = leafpos;
        count[k] = cnt[(size_t)ni];
        mass[k] = n.mass;
        com[3 * k] = n.comx; com[3 * k + 1] = n.comy; com[3 * k + 2] = n.comz;
        gas[k] = n.gas;
        mvel[3 * k] = n.mvx; mvel[3 * k + 1] = n.mvy; mvel[3 * k + 2] = n.mvz;
        k++;
        if (n.leaf) perm[leafpos++] = n.particle;
    }
    return k;
}

int orc_morton_keys(int64_t n, const double *x, const double *y, const double *z,
                    double R, uint64_t *k1, uint64_t *k2, int32_t *gap) {
    for (int64_t i = 0; i < n; i++) {
        double px = x[i], py = y[i], pz = z[i];
        double cx = 0.0, cy = 0.0, cz = 0.0, r = R;
        uint64_t a = 0, b = 0;
        int32_t g = 0;
        if (px < cx - r || px > cx + r || py < cy - r || py > cy + r || pz < cz - r || pz > cz + r) {
            k1[i] = 0; k2[i] = 0; gap[i] = -1;
            continue;
        }
        for (int L = 1; L <= 42; L++) {
            /* the node at level L-1 routes the particle (getOctant) */
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
        k1[i] = a; k2[i] = b; gap[i] = g;
    }
    return 0;
}

int orc_direct_sum(int64_t n, const double *x, const double *y, const double *z,
                   const double *m, int64_t nt, const int64_t *targets,
                   double e0, double G, int32_t nthreads,
                   double *ax, double *ay, double *az, double *pot) {
#ifdef _OPENMP
    int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nth)
#endif
    for (int64_t k = 0; k < nt; k++) {
        int64_t i = targets ? targets[k] : k;
        double sx = 0, sy = 0, sz = 0, sp = 0;
        for (int64_t j = 0; j < n; j++) {
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double r = len3(dx, dy, dz);
            if (r == 0) continue;
            double f = G * m[j] / (r * r + e0 * e0);
            sx += f * (dx / r); sy += f * (dy / r); sz += f * (dz / r);
            if (e0 > 0) sp += -(G * m[j] / e0) * (kPI / 2 - std::atan(r / e0));
            else sp += -G * m[j] / r;
        }
        ax[k] = sx; ay[k] = sy; az[k] = sz;
        if (pot) pot[k] = sp;
    }
    return 0;
}

static uint64_t splitmix_out(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double orc_sfr_uniform(uint64_t seed, uint32_t id, int64_t step) {
    uint64_t k = splitmix_out(seed + 0x9E3779B97F4A7C15ull * ((uint64_t)id + 1));
    k = splitmix_out(k + 0x9E3779B97F4A7C15ull * ((uint64_t)step + 1));
    return (double)(k >> 11) * (1.0 / 9007199254740992.0);
}

void orc_gas_sources(int64_t n, const uint8_t *active, const double *dt, const double *T, const double *rho,
                     int32_t cooling, int32_t star_formation, uint64_t seed, int64_t step, int32_t do_ueuler,
                     uint8_t *type, double *U, double *dUdt, double *sfr) {
    for (int64_t i = 0; i < n; i++) {
        if (active && !active[i]) continue;
        if (type[i] != 2) continue;
        if (cooling) {
            /* Cooling.cpp:8-25: free-free emission 1.42e-27 g_ff sqrt(T) n_e n_i
             * erg/s/cm^3 with g_ff = 1.1, n_e = n_i = 1e6 cm^-3; x 1e-7 to SI */
            double Ti = T[i];
            double gff = 1.1, ne = 1e6, ni = 1e6;
            double ff = 1.42e-27 * gff * std::sqrt(Ti) * ne * ni;
            double rate = ff * 1e-7;
            if (rate > 0 && rho[i] > 0) dUdt[i] -= rate / rho[i];
        }
        if (star_formation) {
            /* SFR.cpp:8-27: thresholds rho > 1e-22 kg/m^3, T < 1e4 K;
             * t_star = 1e15 s; epsilon = 0.1 */
            if (rho[i] > 1e-22 && T[i] < 1e4) {
                double eps = 0.1, t_star = 1e15;
                double pr = 1 - std::exp(-eps * dt[i] / t_star);
                if (sfr) sfr[i] = pr;
                if (orc_sfr_uniform(seed, (uint32_t)i, step) < pr) {
                    type[i] = 1;
                    U[i] = 0.0;
                }
            }
        }
        if (do_ueuler && type[i] == 2) {
            if (!std::isnan(dUdt[i])) U[i] += dUdt[i] * dt[i];
            dUdt[i] = 0;
        }
    }
}

int orc_run(orc_particles *p, const orc_params *prm, const orc_run_params *rp) {
    int64_t n = p->n;
    std::vector<double> ts((size_t)n, 0.0), next((size_t)n, 0.0);
    std::vector<uint8_t> act((size_t)n, 1);
    uint8_t *saved_active = p->active;
    /* Simulation::init force phase: globalTime == nextIntegrationTime == 0. */
    p->active = act.data();
    bool deep = orc_force_step(p, prm, nullptr) == -2;
    double gt = 0.0;
    auto step_of = [&](int64_t i) {
        double am = len3(p->ax[i], p->ay[i], p->az[i]);
        if (am > 0) {
            double t = rp->eta * std::sqrt(prm->e0 / am);
            t = std::clamp(t, rp->min_dt, rp->max_dt);
            t = std::max(std::pow(2, std::floor(std::log2(t))), rp->min_dt);
            ts[(size_t)i] = t;
        } else {
            ts[(size_t)i] = rp->min_dt;
        }
        next[(size_t)i] = gt + ts[(size_t)i];
    };
    for (int64_t i = 0; i < n; i++) step_of(i);
    int steps = 0;
    double h0 = (rp->H0 * kKMS) / kMPC;
    while (gt < rp->end_time && steps < rp->max_steps) {
        for (int64_t i = 0; i < n; i++)
            if (gt >= next[(size_t)i]) step_of(i);
        double mn = std::numeric_limits<double>::max();
        for (int64_t i = 0; i < n; i++) mn = std::min(mn, next[(size_t)i]);
        gt = mn;
        for (int64_t i = 0; i < n; i++) {
            act[(size_t)i] = (gt == next[(size_t)i]);
            if (!act[(size_t)i]) continue;
            double dt = ts[(size_t)i];
            /* TimeIntegration::Kick / Drift (TimeIntegration.cpp:10-25) */
            if (!(std::isnan(p->ax[i]) || std::isnan(p->ay[i]) || std::isnan(p->az[i]))) {
                p->vx[i] = p->vx[i] + p->ax[i] * dt / 2;
                p->vy[i] = p->vy[i] + p->ay[i] * dt / 2;
                p->vz[i] = p->vz[i] + p->az[i] * dt / 2;
            }
            p->x[i] = p->x[i] + p->vx[i] * dt;
            p->y[i] = p->y[i] + p->vy[i] * dt;
            p->z[i] = p->z[i] + p->vz[i] * dt;
        }
        deep = (orc_force_step(p, prm, nullptr) == -2) || deep;
        /* cooling, star formation and TimeIntegration::Ueuler
         * (TimeIntegration.cpp:28-41) of active gas (Simulation.cpp:306-325) */
        orc_gas_sources(n, act.data(), ts.data(), p->T, p->rho, rp->cooling, rp->star_formation, rp->sfr_seed,
                        steps, 1, p->type, p->U, p->dUdt, rp->sfr);
        for (int64_t i = 0; i < n; i++) {
            if (!act[(size_t)i]) continue;
            double dt = ts[(size_t)i];
            double sf = std::exp(h0 * dt);
            p->x[i] *= sf; p->y[i] *= sf; p->z[i] *= sf;
            if (!(std::isnan(p->ax[i]) || std::isnan(p->ay[i]) || std::isnan(p->az[i]))) {
                p->vx[i] = p->vx[i] + p->ax[i] * dt / 2;
                p->vy[i] = p->vy[i] + p->ay[i] * dt / 2;
                p->vz[i] = p->vz[i] + p->az[i] * dt / 2;
            }
            next[(size_t)i] += dt;
        }
        steps++;
    }
    p->active = saved_active;
    return deep ? -2 : steps;
}

}  // extern "C"
