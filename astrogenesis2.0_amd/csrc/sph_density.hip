// SPH group density on the device: Tree::calcGasDensity (Tree.cpp:119-150)
// and Node::calcGasDensity (Node.cpp:722-796).
//
// The reference walks gas particles serially in array order; a particle whose
// h is already set is skipped, and a group assignment overwrites h, rho, P, T
// of every gas particle listed in the group node.  Restated in parallel:
//  1. every gas particle climbs from its leaf to its group node G(q)
//     (core::gas_group); each group keeps the smallest caller index that
//     selects it (first_q);
//  2. a group fires iff that first particle is not covered by a fired group
//     above it (core::group_fires, top-down along the ancestor chain);
//  3. later firing groups that contain a particle are supersets of earlier
//     ones, so each particle takes the values of the largest fired group
//     above its leaf: the fired nodes with no fired ancestor are the groups
//     in effect, and their particle ranges are disjoint;
//  4. rho of a group = sum over its gas particles of m W(|x - COM|, 2 r_node),
//     doubled at nodes where the serial insert lists every particle twice
//     (Node.cpp:518 + :615); the same wave then writes h, rho, P, T of the
//     gas particles in the group's range.
#include "ag2_internal.h"

namespace ag2 {
namespace {

constexpr int kNone = kGroupNone;

__global__ void k_zero_groups(DevState *ds) { ds->final_count = 0; }

// h = 0 for every gas particle (Tree.cpp:121-130) and cleared group state of
// every node, in one pass over max(n, nodes) threads.
__global__ __launch_bounds__(256) void k_group_init(int n, const uint8_t *type, double *h, int nn, int32_t *first_q,
                                                    uint8_t *fired) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n && type[i] == 2) h[i] = 0.0;
    if (i < nn) { first_q[i] = kNone; fired[i] = 0; }
}

__global__ __launch_bounds__(256) void k_group_select(int nin, const uint8_t *stype, const int32_t *base,
                                                      const int32_t *c, const int32_t *parent, const double *gas,
                                                      double mih, const uint32_t *rid, int32_t *first_q) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nin || stype[i] != 2) return;
    int leaf = i + base[i] + c[i];
    int g = core::gas_group(leaf, parent, gas, mih);
    if (g >= 0) atomicMin(&first_q[g], (int)rid[i]);   // caller (array) order
}

__global__ __launch_bounds__(256) void k_group_fire(int nn, const int32_t *parent, const int32_t *first_q,
                                                    uint8_t *fired) {
    int nd = blockIdx.x * 256 + threadIdx.x;
    if (nd >= nn || first_q[nd] == kNone) return;
    fired[nd] = core::group_fires(nd, parent, first_q) ? 1 : 0;
}

// Groups in effect: fired nodes with no fired ancestor.  Each workgroup
// lists its own in LDS and writes them to its slice of the list (256 entries
// per workgroup, at most one per node) with their count, so no counter is
// shared between workgroups.
__global__ __launch_bounds__(256) void k_group_top(int nn, const int32_t *parent, const uint8_t *fired,
                                                   int32_t *list, int32_t *bcount) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int nd = blockIdx.x * 256 + threadIdx.x;
    if (nd < nn && fired[nd]) {
        bool top = true;
        for (int a = parent[nd]; a >= 0; a = parent[a])
            if (fired[a]) { top = false; break; }
        if (top) list[blockIdx.x * 256 + atomicAdd(&cnt, 1)] = nd;
    }
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = cnt;
}

// Number of groups in effect (diagnostics), by the first workgroup of k_group_rho.
__device__ inline void group_count(int nb, const int32_t *bcount, DevState *ds) {
    __shared__ int scr[4];
    int local = 0;
    for (int k = threadIdx.x; k < nb; k += 256) local += bcount[k];
    int tot;
    block256_excl_scan(local, scr, &tot);
    if (threadIdx.x == 0) ds->final_count = tot;
}

// One wave per group in effect, in a fixed lane-strided order with a
// butterfly sum (deterministic); the wave then writes h, rho, P, T of every
// gas particle in the group's range (Node.cpp:757-795).
__global__ __launch_bounds__(256) void k_group_rho(int nb, const int32_t *list, const int32_t *bcount,
                                                   const int32_t *first, const int32_t *count, const int8_t *level,
                                                   const int32_t *parent, const double *cx, const double *cy,
                                                   const double *cz, const double *sx, const double *sy,
                                                   const double *sz, const double *sm, const uint8_t *stype,
                                                   const double *U, const double *mu, int64_t N, int64_t thr,
                                                   DevState *ds, double *g_rho, double *h, double *rho,
                                                   double *P, double *T) {
    if (blockIdx.x == 0) group_count(nb, bcount, ds);
    // one workgroup per slice of the list, its four waves taking the slice's
    // groups in turn
    const int w = threadIdx.x >> 6;
    double R = __longlong_as_double((long long)ds->Rbits);
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        int nf = bcount[b];
        for (int k = w; k < nf; k += 4) {
            int F = list[b * 256 + k];
            double hh = ldexp(R, -(int)level[F]) * 2;
            double gx = cx[F], gy = cy[F], gz = cz[F];
            int j0 = first[F], j1 = first[F] + count[F];
            double s = 0.0;
            for (int j = j0 + lane_id(); j < j1; j += 64) {
                if (stype[j] == 2) {
                    double dx = sx[j] - gx, dy = sy[j] - gy, dz = sz[j] - gz;
                    s += sm[j] * core::spline_w(sqrt(dx * dx + dy * dy + dz * dz), hh);
                }
            }
            s = wave_sum(s);
            bool dup = count[F] >= 2 && count[F] < thr && parent[F] >= 0 &&
                       core::parallel_node(count, parent, parent[F], N, thr);
            double r = dup ? 2 * s : s;
            if (lane_id() == 0) g_rho[F] = r;
            for (int j = j0 + lane_id(); j < j1; j += 64) {
                if (stype[j] == 2) {
                    h[j] = hh;
                    rho[j] = r;
                    P[j] = (kGAMMA - 1.0) * U[j] * r;
                    T[j] = (kGAMMA - 1.0) * U[j] * kPRTN * mu[j] / (kKB);
                }
            }
        }
    }
}

}  // namespace

void gas_density(Ctx &c) {
    hipStream_t st = c.stream;
    int n = (int)c.n, nin = c.nin, nn = c.n_nodes;
    if (c.timing >= 2) AG2_HIP(hipEventRecord(c.ev[3], st));
    // an upload without gas never gains any (star formation only turns gas
    // into stars): no group can fire, so only the count is cleared
    if (nin > 0 && nn > 0 && c.sph_active) {
        // the build's permutation set h = 0 for the gas when a density phase
        // was to follow it, and the build cleared the nodes' group state;
        // what this call finds not done, it does
        const int nh = c.h_cleared ? 0 : n, ng = c.groups_cleared ? 0 : nn;
        if (nh > 0 || ng > 0)
            hipLaunchKernelGGL(k_group_init, dim3(grid_for(nh > ng ? nh : ng, 256)), dim3(256), 0, st, nh, c.type,
                               c.h, ng, c.g_first, c.g_fired);
        int gi = grid_for(nin, 256);
        hipLaunchKernelGGL(k_group_select, dim3(gi), dim3(256), 0, st, nin, c.type, c.base, c.cnt, c.n_parent,
                           c.n_gas, c.p.massInH, c.rid, c.g_first);
        hipLaunchKernelGGL(k_group_fire, dim3(grid_for(nn, 256)), dim3(256), 0, st, nn, c.n_parent, c.g_first,
                           c.g_fired);
        // g_final holds the per-workgroup group counts
        int nb = grid_for(nn, 256);
        hipLaunchKernelGGL(k_group_top, dim3(nb), dim3(256), 0, st, nn, c.n_parent, c.g_fired, c.g_list, c.g_final);
        hipLaunchKernelGGL(k_group_rho, dim3(nb), dim3(256), 0, st, nb, c.g_list,
                           c.g_final, c.n_first, c.n_count, c.n_level, c.n_parent, c.n_cx, c.n_cy, c.n_cz, c.x, c.y,
                           c.z, c.m, c.type, c.U, c.mu, c.n, c.p.split_threshold, c.ds, c.g_rho, c.h, c.rho, c.P,
                           c.T);
    } else {
        if (n > 0 && c.sph_active)
            hipLaunchKernelGGL(k_group_init, dim3(grid_for(n, 256)), dim3(256), 0, st, n, c.type, c.h, 0, c.g_first,
                               c.g_fired);
        hipLaunchKernelGGL(k_zero_groups, dim3(1), dim3(1), 0, st, c.ds);
    }
    c.h_cleared = c.groups_cleared = false;
    if (c.timing >= 2) AG2_HIP(hipEventRecord(c.ev[4], st));
    AG2_HIP(hipGetLastError());
}

}  // namespace ag2
