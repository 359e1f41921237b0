// Barnes-Hut gravity walk with the embedded SPH force, plain FP64 form:
// Tree::calculateForces (Tree.cpp:57-83) -> Node::calculateGravityForce
// (Node.cpp:247-399) -> Node::calcSPHForce (Node.cpp:88-172).
//
// One lane per target particle; targets run in key order so that the 64 lanes
// of a wavefront are spatial neighbours and walk nearly the same nodes (the
// node records they load coalesce in L1/L2).  Each lane walks the pre-order
// node array stacklessly (core::walk_fp64): open -> next node, accept or
// skip -> the node's skip pointer, so the walk is a forward scan that always
// terminates.  The opening test uses correctly rounded sqrt and division,
// like the reference, so every particle sees exactly the reference's accepted
// node set.
#include "ag2_internal.h"

namespace ag2 {
namespace {

__global__ __launch_bounds__(256) void k_walk64(int n, int nin, int nn, const WalkNode *__restrict__ wn,
                                                const uint32_t *__restrict__ idx, const double *__restrict__ sx,
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
        uint32_t q = idx[t];
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

__global__ void k_zero_counters(DevState *ds) {
    ds->interactions = 0ull;
    ds->sph_pairs = 0ull;
}

}  // namespace

void forces(Ctx &c, int mode) {
    hipStream_t st = c.stream;
    int n = (int)c.n;
    if (c.timing) AG2_HIP(hipEventRecord(c.ev[4], st));
    if (mode == 0) hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(1), 0, st, c.ds);
    core::WalkParams wp{c.p.theta, c.p.e0, c.p.G, c.p.e0 < 2.147e13 ? 1 : 0};
    hipLaunchKernelGGL(k_walk64, dim3(grid_for(n, 256)), dim3(256), 0, st, n, c.nin, c.n_nodes, c.wnode, c.idx, c.sx,
                       c.sy, c.sz, c.sm, c.stype, c.svx, c.svy, c.svz, c.n_gvx, c.n_gvy, c.n_gvz, c.h, c.rho, c.P,
                       c.active, wp, mode, c.ax, c.ay, c.az, c.dUdt, c.pot, c.ds);
    if (c.timing) AG2_HIP(hipEventRecord(c.ev[5], st));
    AG2_HIP(hipGetLastError());
}

}  // namespace ag2
