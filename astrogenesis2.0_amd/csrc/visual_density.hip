// Visual density on the device: Tree::calcVisualDensity (Tree.cpp:152-176)
// and Node::calcVisualDensity (Node.cpp:832-877).
//
// The reference starts from each particle's leaf (Particle::node), climbs to
// the parent while the parent's half-width is strictly closer to the
// estimation radius r_vd, and writes mass / half-width^3 of the node it stops
// at into every particle listed under that node; a climb that reaches the
// root writes nothing (the root has no parent), and nor does a zero or
// infinite density.  Particles already written are skipped.
//
// Half-widths are R / 2^level, so whether a node climbs depends only on its
// level: the levels that climb are exactly those deeper than one stop level.
// Every particle below a stop node therefore stops at that same node, and
// the per-particle climb below gives each particle the value the reference's
// sequential loop leaves in it, in any order.  One thread per particle, key
// order; each reads its leaf index (i + base + cnt from the topology pass)
// and walks parent / level / mass of the pre-order node arrays.
#include "ag2_internal.h"

#include <cmath>

namespace ag2 {
namespace {

__global__ __launch_bounds__(256) void k_visual_density(int n, int nin, const int32_t *__restrict__ base,
                                                        const int32_t *__restrict__ cnt,
                                                        const int32_t *__restrict__ parent,
                                                        const int8_t *__restrict__ level,
                                                        const double *__restrict__ nmass, const DevState *ds,
                                                        double rvd, double *__restrict__ vd) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double out = 0.0;
    if (i < nin) {
        double R = __longlong_as_double((long long)ds->Rbits);
        int a = i + base[i] + cnt[i];
        // Node.cpp:848-853: climb while |r_vd - r| > |r_vd - r_parent|
        for (;;) {
            int p = parent[a];
            if (p < 0) { a = -1; break; }   // Node.cpp:834: the root assigns nothing
            double ra = ldexp(R, -(int)level[a]);
            double rp = ldexp(R, -(int)level[p]);
            if (fabs(rvd - ra) > fabs(rvd - rp)) a = p;
            else break;
        }
        if (a >= 0) {
            // Node.cpp:855-860
            double r = ldexp(R, -(int)level[a]);
            double vol = r * r * r;
            double m = nmass[a];
            if (vol != 0.0 && m != 0.0) {
                double d = m / vol;
                if (d != 0.0 && d != INFINITY) out = d;
            }
        }
    }
    vd[i] = out;   // Tree.cpp:155-161: particles nobody writes keep 0
}

}  // namespace

void visual_density(Ctx &c, double rvd, double *vd) {
    int n = (int)c.n;
    if (n < 1) return;
    hipLaunchKernelGGL(k_visual_density, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, n, c.nin, c.base, c.cnt,
                       c.n_parent, c.n_level, c.n_mass, c.ds, rvd, vd);
    AG2_HIP(hipGetLastError());
}

}  // namespace ag2
