// FP64 direct summation with the reference's softened law, for accuracy
// diagnostics of the tree forces: a_i = sum_j G m_j d / ((r^2 + e0^2) r),
// phi_i = sum_j -(G m_j / e0)(pi/2 - atan(r / e0)), r == 0 pairs skipped
// (Node.cpp:274-276, :301).  Each workgroup takes 256 targets; all sources
// stream through LDS in tiles of 256, so every source record is read once
// per workgroup.
#include <mutex>

#include "ag2_internal.h"

namespace ag2 {
namespace {

constexpr int kTile = 256;

__global__ __launch_bounds__(256) void k_direct(int n, const double *__restrict__ x, const double *__restrict__ y,
                                                const double *__restrict__ z, const double *__restrict__ m, int nt,
                                                const int64_t *__restrict__ targets, const int32_t *__restrict__ inv,
                                                double G, double e0,
                                                int with_pot, double *ax, double *ay, double *az, double *pot) {
    __shared__ double sx[kTile], sy[kTile], sz[kTile], sm[kTile];
    int k = blockIdx.x * 256 + threadIdx.x;
    bool valid = k < nt;
    int64_t i = valid ? inv[targets[k]] : 0;   // caller index -> resident slot
    double px = valid ? x[i] : 0.0, py = valid ? y[i] : 0.0, pz = valid ? z[i] : 0.0;
    double e02 = e0 * e0;
    double gx = 0.0, gy = 0.0, gz = 0.0, ph = 0.0;
    for (int base = 0; base < n; base += kTile) {
        int j = base + threadIdx.x;
        __syncthreads();
        sx[threadIdx.x] = j < n ? x[j] : 0.0;
        sy[threadIdx.x] = j < n ? y[j] : 0.0;
        sz[threadIdx.x] = j < n ? z[j] : 0.0;
        sm[threadIdx.x] = j < n ? m[j] : 0.0;
        __syncthreads();
        int cnt = n - base < kTile ? n - base : kTile;
        for (int s = 0; s < cnt; s++) {
            double dx = sx[s] - px, dy = sy[s] - py, dz = sz[s] - pz;
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0) continue;
            double r = sqrt(r2);
            double f = G * sm[s] / ((r2 + e02) * r);
            gx += f * dx; gy += f * dy; gz += f * dz;
            if (with_pot) ph += -(G * sm[s] / e0) * (1.5707963267948966 - atan(r / e0));
        }
    }
    if (valid) {
        ax[k] = gx; ay[k] = gy; az[k] = gz;
        if (with_pot) pot[k] = ph;
    }
}

}  // namespace

void direct_sum(Ctx &c, int nt, const int64_t *d_targets, double *d_out, int with_pot) {
    build_inverse(c);
    hipLaunchKernelGGL(k_direct, dim3(grid_for(nt, 256)), dim3(256), 0, c.stream, (int)c.n, c.x, c.y, c.z, c.m, nt,
                       d_targets, c.inv, c.p.G, c.p.e0, with_pot, d_out, d_out + nt, d_out + 2 * (size_t)nt,
                       d_out + 3 * (size_t)nt);
    AG2_HIP(hipGetLastError());
}

// ------------------------------------------------ direct sum without allocation
// Scratch of ag2_direct_sum_noalloc: the device copies of the target list
// (int64) and of the four output columns (ax, ay, az, potential), for at most
// AG2_DIRECT_NOALLOC_MAX targets: 5 x 8 B x 65536 = 2.5 MiB per device, held
// until the process ends.  In host code `scratch` is only a stand-in for the
// symbol: its device address, per device, comes from hipGetSymbolAddress.
namespace {
constexpr int64_t kNoAllocTargets = AG2_DIRECT_NOALLOC_MAX;
static __device__ __align__(256) int64_t scratch[kNoAllocTargets * 5];
constexpr int kMaxDevices = 64;
void *g_scratch_addr[kMaxDevices] = {};
std::mutex g_scratch_lock;
}  // namespace

void direct_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) throw HipError("device ordinal out of range for the direct-sum scratch");
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (g_scratch_addr[device]) return;
    // the caller has made `device` current; each device holds its own copy
    void *p = nullptr;
    AG2_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_addr[device] = p;
}

void direct_sum_noalloc(Ctx &c, int nt, const int64_t *h_targets, double *ax, double *ay, double *az, double *pot) {
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> hold(g_scratch_lock);
        base = (c.device >= 0 && c.device < kMaxDevices) ? g_scratch_addr[c.device] : nullptr;
    }
    if (!base) throw HipError("direct-sum scratch of this device was not set up at ag2_create");
    int64_t *d_targets = (int64_t *)base;
    double *d_out = (double *)(d_targets + kNoAllocTargets);
    const size_t bytes = (size_t)nt * 8;
    AG2_HIP(hipMemcpyAsync(d_targets, h_targets, bytes, hipMemcpyHostToDevice, c.stream));
    // the same launch as direct_sum, with the columns kNoAllocTargets apart
    build_inverse(c);
    hipLaunchKernelGGL(k_direct, dim3(grid_for(nt, 256)), dim3(256), 0, c.stream, (int)c.n, c.x, c.y, c.z, c.m, nt,
                       d_targets, c.inv, c.p.G, c.p.e0, pot ? 1 : 0, d_out, d_out + kNoAllocTargets,
                       d_out + 2 * kNoAllocTargets, d_out + 3 * kNoAllocTargets);
    AG2_HIP(hipGetLastError());
    AG2_HIP(hipMemcpyAsync(ax, d_out, bytes, hipMemcpyDeviceToHost, c.stream));
    AG2_HIP(hipMemcpyAsync(ay, d_out + kNoAllocTargets, bytes, hipMemcpyDeviceToHost, c.stream));
    AG2_HIP(hipMemcpyAsync(az, d_out + 2 * kNoAllocTargets, bytes, hipMemcpyDeviceToHost, c.stream));
    if (pot) AG2_HIP(hipMemcpyAsync(pot, d_out + 3 * kNoAllocTargets, bytes, hipMemcpyDeviceToHost, c.stream));
    AG2_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace ag2
