// FP64 kick-drift-kick integrator and diagnostics on the device:
// Simulation::run's per-particle loops (Simulation.cpp:166-356) with
// TimeIntegration::Kick / Drift / Ueuler (TimeIntegration.cpp:10-41).
// Elementwise and HBM-bound: ~120 bytes per particle per step.
#include "ag2_internal.h"

namespace ag2 {
namespace {

__device__ inline double step_rule(double ax, double ay, double az, double eta, double e0, double mn, double mx) {
    double am = sqrt(ax * ax + ay * ay + az * az);
    if (am > 0) {
        double t = eta * sqrt(e0 / am);
        t = t < mn ? mn : (mx < t ? mx : t);               // std::clamp
        double p2 = ldexp(1.0, (int)floor(log2(t)));          // pow(2, floor(log2 t))
        return p2 < mn ? mn : p2;                          // std::max(p2, min)
    }
    return mn;
}

// Simulation.cpp:174-207: next = 0, then every particle gets a step.
__global__ __launch_bounds__(256) void k_init_steps(int n, const double *ax, const double *ay, const double *az,
                                                    double eta, double e0, double mn, double mx, double *ts,
                                                    double *next, DevState *ds) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) ds->gtime = 0.0;
    if (i >= n) return;
    double t = step_rule(ax[i], ay[i], az[i], eta, e0, mn, mx);
    ts[i] = t;
    next[i] = 0.0 + t;
}

// Simulation.cpp:213-254: due particles get a new step; global min of next.
// Grid-stride over a fixed grid: each workgroup reduces its minimum in LDS and
// issues one atomic, so the single address sees at most kStepBlocks atomics.
constexpr int kStepBlocks = 1024;

__global__ __launch_bounds__(256) void k_steps(int n, const double *ax, const double *ay, const double *az, double eta,
                                               double e0, double mn, double mx, double *ts, double *next,
                                               const DevState *ds, unsigned long long *part) {
    __shared__ unsigned long long wmin[4];
    double gt = ds->gtime;
    unsigned long long best = 0x7ff0000000000000ull;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double nx = next[i];
        if (gt >= nx) {
            double t = step_rule(ax[i], ay[i], az[i], eta, e0, mn, mx);
            ts[i] = t;
            nx = gt + t;
            next[i] = nx;
        }
        unsigned long long b = (unsigned long long)__double_as_longlong(nx);
        best = b < best ? b : best;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        unsigned long long t = __shfl_xor(best, o, 64);
        best = t < best ? t : best;
    }
    if (lane_id() == 0) wmin[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = wmin[0];
        for (int w = 1; w < 4; w++) b = wmin[w] < b ? wmin[w] : b;
        part[blockIdx.x] = b;
    }
}

// Global minimum of k_steps' per-block partials, formed by every workgroup
// of k_kick_drift for itself (a thousand values at most, all in the same
// order): no kernel between the two.
__device__ inline unsigned long long min_next_fold(int nb, const unsigned long long *part) {
    __shared__ unsigned long long w4[4];
    unsigned long long v = 0x7ff0000000000000ull;
    for (int k = threadIdx.x; k < nb; k += 256) v = part[k] < v ? part[k] : v;
    for (int o = 32; o >= 1; o >>= 1) {
        unsigned long long q = __shfl_xor(v, o, 64);
        v = q < v ? q : v;
    }
    if (lane_id() == 0) w4[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long r = w4[0];
    for (int w = 1; w < 4; w++) r = w4[w] < r ? w4[w] : r;
    return r;
}

// Simulation.cpp:256-272: globalTime = min next; active particles kick + drift.
__global__ __launch_bounds__(256) void k_kick_drift(int n, int nb, const unsigned long long *part, DevState *ds,
                                                    const double *ts, const double *next,
                                                    double *x, double *y, double *z, double *vx, double *vy,
                                                    double *vz, const double *ax, const double *ay, const double *az,
                                                    uint8_t *active) {
    const unsigned long long mb = min_next_fold(nb, part);
    int i = blockIdx.x * 256 + threadIdx.x;
    double gt = __longlong_as_double((long long)mb);
    if (i == 0) { ds->min_next_bits = mb; ds->gtime = gt; }
    if (i >= n) return;
    bool a = gt == next[i];
    active[i] = a ? 1 : 0;
    if (!a) return;
    double dt = ts[i];
    double gx = ax[i], gy = ay[i], gz = az[i];
    if (!(gx != gx || gy != gy || gz != gz)) {
        vx[i] = vx[i] + gx * dt / 2;
        vy[i] = vy[i] + gy * dt / 2;
        vz[i] = vz[i] + gz * dt / 2;
    } else {
        atomicAdd(&ds->nan_kicks, 1ull);
    }
    x[i] = x[i] + vx[i] * dt;
    y[i] = y[i] + vy[i] * dt;
    z[i] = z[i] + vz[i] * dt;
}

// Simulation.cpp:295-341: U update for gas, Hubble scaling, second kick.
// SplitMix64 output function (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform in [0, 1) for (seed, caller index, step): a counter-based draw, so
// the value does not depend on launch order or on which particles are active.
__device__ __forceinline__ double sfr_uniform(uint64_t seed, uint32_t id, int64_t step) {
    uint64_t k = mix64(seed + 0x9E3779B97F4A7C15ull * ((uint64_t)id + 1));
    k = mix64(k + 0x9E3779B97F4A7C15ull * ((uint64_t)step + 1));
    return (double)(k >> 11) * 0x1.0p-53;
}

struct GasSources {
    int cooling, star_formation;
    uint64_t seed;
    int64_t step;
    const uint32_t *rid;   // caller index of each resident slot
    const double *rho, *T;
    double *sfr;           // caller order
};

__global__ __launch_bounds__(256) void k_kick2(int n, const uint8_t *active, uint8_t *type, const double *ts,
                                               double *next, double *x, double *y, double *z, double *vx,
                                               double *vy, double *vz, const double *ax, const double *ay,
                                               const double *az, double *U, double *dUdt, double H0, DevState *ds,
                                               GasSources gs) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || !active[i]) return;
    double dt = ts[i];
    if (type[i] == 2) {
        if (gs.cooling) {
            // Cooling::coolingRoutine (Cooling.cpp:6-26): free-free emission at
            // n_e = n_i = 1e6 cm^-3, g_ff = 1.1, erg/s/cm^3 -> W/m^3
            double ff = 1.42e-27 * 1.1 * sqrt(gs.T[i]) * 1e6 * 1e6;
            double rate = ff * 1e-7;
            double r = gs.rho[i];
            if (rate > 0 && r > 0) dUdt[i] -= rate / r;
        }
        if (gs.star_formation) {
            // SFR::sfrRoutine (SFR.cpp:12-28): rho > 1e-22 kg/m^3 and T < 1e4 K,
            // t_star = 1e15 s, efficiency 0.1
            double r = gs.rho[i];
            if (r > 1e-22 && gs.T[i] < 1e4) {
                double p = 1 - exp(-0.1 * dt / 1e15);
                uint32_t id = gs.rid[i];
                gs.sfr[id] = p;
                if (sfr_uniform(gs.seed, id, gs.step) < p) {
                    type[i] = 1;
                    U[i] = 0.0;
                }
            }
        }
        if (type[i] == 2) {
            double du = dUdt[i];
            if (!(du != du)) U[i] += du * dt;
            dUdt[i] = 0;
        }
    }
    if (H0 != 0.0) {   // exp(0) = 1 leaves the positions as they are
        double h0 = (H0 * kKMS) / kMPC;
        double sf = exp(h0 * dt);
        x[i] *= sf; y[i] *= sf; z[i] *= sf;
    }
    double gx = ax[i], gy = ay[i], gz = az[i];
    if (!(gx != gx || gy != gy || gz != gz)) {
        vx[i] = vx[i] + gx * dt / 2;
        vy[i] = vy[i] + gy * dt / 2;
        vz[i] = vz[i] + gz * dt / 2;
    } else {
        atomicAdd(&ds->nan_kicks, 1ull);
    }
    next[i] += dt;
}

__global__ __launch_bounds__(256) void k_fill_u8(int n, uint8_t *a, uint8_t v) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = v;
}

// Energies and momentum: per-block partials in fixed order.
__global__ __launch_bounds__(256) void k_energy_part(int n, const double *m, const double *vx, const double *vy,
                                                     const double *vz, const double *U, const uint8_t *type,
                                                     const double *pot, int with_pot, double *part) {
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double mm = m[i];
        a[0] += 0.5 * mm * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
        if (type[i] == 2) a[1] += mm * U[i];
        a[2] += mm * vx[i]; a[3] += mm * vy[i]; a[4] += mm * vz[i];
        if (with_pot) a[5] += 0.5 * mm * pot[i];
    }
    __shared__ double red[4][6];
    int w = threadIdx.x >> 6;
    for (int j = 0; j < 6; j++) {
        double v = wave_sum(a[j]);
        if (lane_id() == 0) red[w][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int j = threadIdx.x;
        part[6 * blockIdx.x + j] = ((red[0][j] + red[1][j]) + red[2][j]) + red[3][j];
    }
}

__global__ void k_energy_final(int nb, const double *part, DevState *ds) {
    if (threadIdx.x != 0) return;
    for (int j = 0; j < 6; j++) {
        double s = 0;
        for (int b = 0; b < nb; b++) s += part[6 * b + j];
        ds->red[j] = s;
    }
}

}  // namespace

void fill_active(Ctx &c, uint8_t v) {
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(c.n, 256)), dim3(256), 0, c.stream, (int)c.n, c.active, v);
}

void kdk_init_steps(Ctx &c) {
    int n = (int)c.n;
    hipLaunchKernelGGL(k_init_steps, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, n, c.ax, c.ay, c.az, c.p.eta,
                       c.p.e0, c.p.min_dt, c.p.max_dt, c.ts, c.next, c.ds);
}

void kdk_begin(Ctx &c) {
    int n = (int)c.n;
    hipStream_t st = c.stream;
    unsigned sb = grid_for(n, 256);
    sb = sb < (unsigned)kStepBlocks ? (sb ? sb : 1u) : (unsigned)kStepBlocks;
    unsigned long long *part = (unsigned long long *)c.partials;
    hipLaunchKernelGGL(k_steps, dim3(sb), dim3(256), 0, st, n, c.ax, c.ay, c.az, c.p.eta, c.p.e0,
                       c.p.min_dt, c.p.max_dt, c.ts, c.next, c.ds, part);
    hipLaunchKernelGGL(k_kick_drift, dim3(grid_for(n, 256)), dim3(256), 0, st, n, (int)sb, part, c.ds, c.ts, c.next,
                       c.x, c.y, c.z, c.vx, c.vy, c.vz, c.ax, c.ay, c.az, c.active);
    AG2_HIP(hipGetLastError());
}

void kdk_end(Ctx &c) {
    int n = (int)c.n;
    GasSources gs{c.p.cooling != 0, c.p.star_formation != 0, c.p.sfr_seed, c.steps, c.rid, c.rho, c.T, c.sfr};
    hipLaunchKernelGGL(k_kick2, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, n, c.active, c.type, c.ts, c.next,
                       c.x, c.y, c.z, c.vx, c.vy, c.vz, c.ax, c.ay, c.az, c.U, c.dUdt, c.p.H0, c.ds, gs);
    AG2_HIP(hipGetLastError());
}

void energies(Ctx &c, int with_pot, double out[6]) {
    int n = (int)c.n;
    int nb = grid_for(n, 256);
    if (nb > 512) nb = 512;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_energy_part, dim3(nb), dim3(256), 0, c.stream, n, c.m, c.vx, c.vy, c.vz, c.U, c.type, c.pot,
                       with_pot, c.partials);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, c.stream, nb, c.partials, c.ds);
    AG2_HIP(hipMemcpyAsync(&c.hs, c.ds, sizeof(DevState), hipMemcpyDeviceToHost, c.stream));
    AG2_HIP(hipStreamSynchronize(c.stream));
    for (int j = 0; j < 6; j++) out[j] = c.hs.red[j];
}

}  // namespace ag2
