// Barnes-Hut gravity walk with the embedded SPH force:
// Tree::calculateForces (Tree.cpp:57-83) -> Node::calculateGravityForce
// (Node.cpp:247-399) -> Node::calcSPHForce (Node.cpp:88-172).
//
// k_walk_block: one wavefront per group of 64 targets that are consecutive in
// key order (spatial neighbours).  The wave walks a child-contiguous copy of
// the tree together, a block of up to eight siblings per step; each lane
// evaluates its own opening test and ballots decide what happens:
//   - lanes that accept a node (leaf, or radius/r < theta) append
//     (node, lane mask) to an interaction batch;
//   - if any lane opens it, its child block is pushed with the opening lanes
//     as its mask;
// so every lane sees exactly the node set of its own recursive walk in the
// reference.  The batch is swept when full: lane j stages entry j in LDS and
// every lane computes its interactions with all staged entries.
// The opening test decides radius^2 against theta^2 r^2 with a 1e-10 margin
// and falls back to the reference's correctly rounded radius / sqrt(r^2)
// inside the margin, so the decision equals the reference's bit for bit.
//
// FP64 sweep: the reference's softened law G m / (r^2 + e0^2) d / r.
// FP32 sweep (split precision): node centres and targets become FP32 offsets
// from the group's bounding-box centre in units of the root half-width
// (FP64 subtraction first), masses are scaled by the root mass, the force
// is m' / sqrt(r'^2 (r'^2 + e'^2)^2) with one v_rsq_f32, each 64-entry chunk
// is summed plainly and added to the running total with Kahan compensation.
// The SPH term stays FP64 in both forms.
//
// k_walk64 (one lane per target, stackless) is kept for the potential.
#include "ag2_internal.h"

namespace ag2 {
namespace {

constexpr int kWaves = 4;
constexpr int kBatch = 64;
constexpr int kStack = 320;   // pending child blocks: <= 7 per level x 44 levels + a bucket remainder

struct BlockLds {
    int32_t sfirst[kStack];
    int32_t scount[kStack];
    unsigned long long smask[kStack];
    unsigned long long emask[kBatch];
    int32_t eflag[kBatch];
    union {
        struct { double cx[kBatch], cy[kBatch], cz[kBatch], m[kBatch]; } d;   // FP64 sweep
        float4 f[kBatch];                                                     // FP32 sweep
    } e;
};

struct WalkArgs {
    int n, nin, nn;
    const core::TreeNode *__restrict__ tree;
    const uint32_t *__restrict__ idx;
    const double *__restrict__ sx, *__restrict__ sy, *__restrict__ sz, *__restrict__ sm;
    const uint8_t *__restrict__ stype;
    const double *__restrict__ svx, *__restrict__ svy, *__restrict__ svz;
    const double *__restrict__ mvx, *__restrict__ mvy, *__restrict__ mvz;
    const double *__restrict__ h, *__restrict__ rho, *__restrict__ P;
    const uint8_t *__restrict__ active;
    const DevState *ds;
    core::WalkParams wp;
    double *ax, *ay, *az, *dUdt;
    DevState *dsw;
};

__device__ inline double rsq64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double hx = 0.5 * x;
    y = y * fma(-hx * y, y, 1.5);
    y = y * fma(-hx * y, y, 1.5);
    return y;
}

__device__ inline double rcp64(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = fma(y, fma(-x, y, 1.0), y);
    y = fma(y, fma(-x, y, 1.0), y);
    return y;
}

__device__ inline int rl(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ inline double uni(double v) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ inline double rl(double v, int j) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), j);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline unsigned long long rl(unsigned long long v, int j) {
    int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffull), j);
    int hi = __builtin_amdgcn_readlane((int)(v >> 32), j);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// Block walk: one wavefront per group of 64 targets consecutive in key order.
// The tree is stored child-contiguous (core::TreeNode), and the wave keeps
// an LDS stack of pending child blocks (first, count, lane mask).  A block of
// up to 8 siblings is loaded in one go (lane k loads sibling k) and its
// records are broadcast with readlane; every lane tests every sibling, and
// ballots turn the results into
//   - accepted (node, mask) entries of the interaction batch, and
//   - pushed child blocks carrying the mask of the lanes that opened.
// Each lane thus meets exactly the nodes of its own recursive walk in the
// reference (Node.cpp:382-396 visits the children of an opened node; the
// visiting order differs, which only reorders the sum).  The batch (one
// entry per lane register) is swept when it may overflow: lane j stages its
// entry in LDS, then every lane computes its interaction with every entry.
#ifndef AG2_WALK32_WAVES
#define AG2_WALK32_WAVES 4
#endif

template <bool F32>
__device__ __forceinline__ void walk_block(const WalkArgs &a) {
    __shared__ BlockLds lds[kWaves];
    const int wv = threadIdx.x >> 6;
    const int lane = lane_id();
    BlockLds &S = lds[wv];
    const int g = blockIdx.x * kWaves + wv;
    const int t = g * 64 + lane;
    const bool valid = t < a.n;
    const uint32_t q = valid ? a.idx[t] : 0u;
    const bool act = valid && a.active[q] != 0;
    double px = 0.0, py = 0.0, pz = 0.0, pm = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    if (valid) {
        px = a.sx[t]; py = a.sy[t]; pz = a.sz[t]; pm = a.sm[t];
        vx = a.svx[t]; vy = a.svy[t]; vz = a.svz[t];
    }
    core::Target tg;
    tg.x = px; tg.y = py; tg.z = pz; tg.m = pm; tg.vx = vx; tg.vy = vy; tg.vz = vz;
    tg.gas = valid && a.stype[t] == 2;
    tg.h = act ? a.h[q] : 0.0;
    tg.rho = act ? a.rho[q] : 0.0;
    tg.P = act ? a.P[q] : 0.0;
    tg.self = t;
    const unsigned long long amask = __ballot(act && pm != 0.0);
    const bool sph_lane = act && tg.gas && tg.h > 0.0;
    const bool any_gas = __ballot(sph_lane) != 0ull;

    core::WalkSums s{0, 0, 0, 0, 0, 0, 0};
    unsigned nint32 = 0u;
    double invL = 0.0, invMu = 0.0;
    float xr = 0.f, yr = 0.f, zr = 0.f, e2f = 0.f;
    float sumx = 0.f, sumy = 0.f, sumz = 0.f, cmpx = 0.f, cmpy = 0.f, cmpz = 0.f;
    const double theta = a.wp.theta, th2 = a.wp.theta * a.wp.theta;
    const double e02 = a.wp.e0 * a.wp.e0;
    const double R = __longlong_as_double((long long)a.ds->Rbits);

    if (amask != 0ull && a.nn > 0) {
        // bounding box of the group's active targets (decides whole-group
        // accept / open for most nodes without per-lane tests)
        double lo[3] = {px, py, pz}, hi[3] = {px, py, pz};
        if (!((amask >> lane) & 1ull)) { lo[0] = lo[1] = lo[2] = 1e308; hi[0] = hi[1] = hi[2] = -1e308; }
        for (int d = 32; d >= 1; d >>= 1)
            for (int k = 0; k < 3; k++) {
                double a1 = __shfl_xor(lo[k], d, 64), b1 = __shfl_xor(hi[k], d, 64);
                lo[k] = a1 < lo[k] ? a1 : lo[k];
                hi[k] = b1 > hi[k] ? b1 : hi[k];
            }
        for (int k = 0; k < 3; k++) { lo[k] = uni(lo[k]); hi[k] = uni(hi[k]); }
        const double gcx = 0.5 * (lo[0] + hi[0]), gcy = 0.5 * (lo[1] + hi[1]), gcz = 0.5 * (lo[2] + hi[2]);
        if (F32) {
            double Mu = a.tree[0].mass;
            invL = uni(R > 0 ? 1.0 / R : 1.0);
            invMu = uni(Mu > 0 ? 1.0 / Mu : 1.0);
            xr = (float)((px - gcx) * invL); yr = (float)((py - gcy) * invL); zr = (float)((pz - gcz) * invL);
            double el = a.wp.e0 * invL;
            e2f = uni((float)(el * el));
        }
        if (lane == 0) { S.sfirst[0] = 0; S.scount[0] = 1; S.smask[0] = amask; }
        int sp = 1;
        int ni = 0;
        int myNode = 0;
        unsigned long long myMask = 0ull;
        unsigned long long n_entries = 0ull, n_tests = 0ull;
        while (true) {
            if (sp > 0) {
                // ------------------------------------------- pop a block
                sp--;
                const int f = __builtin_amdgcn_readfirstlane(S.sfirst[sp]);
                int cnt = __builtin_amdgcn_readfirstlane(S.scount[sp]);
                unsigned long long M = S.smask[sp];
                M = ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(M >> 32)) << 32) |
                    (unsigned)__builtin_amdgcn_readfirstlane((int)(M & 0xffffffffull));
                if (cnt > 8) {
                    if (lane == 0) { S.sfirst[sp] = f + 8; S.scount[sp] = cnt - 8; S.smask[sp] = M; }
                    sp++;
                    cnt = 8;
                }
                n_tests += (unsigned long long)cnt;
                // lane k < cnt holds sibling k and classifies it against the
                // group box: 1 accept for every lane, 2 open for every lane,
                // 3 decide per lane, 0 skip (mass 0)
                core::TreeNode r;
                int cls = 0;
                double rad = 0.0;
                if (lane < cnt) {
                    r = a.tree[f + lane];
                    rad = ldexp(R, -(int)r.level);
                    if (r.mass == 0.0) cls = 0;
                    else if (r.pidx >= 0) cls = 1;
                    else {
                        double ex = fmax(fmax(lo[0] - r.cx, r.cx - hi[0]), 0.0);
                        double ey = fmax(fmax(lo[1] - r.cy, r.cy - hi[1]), 0.0);
                        double ez = fmax(fmax(lo[2] - r.cz, r.cz - hi[2]), 0.0);
                        double rmin2 = ex * ex + ey * ey + ez * ez;
                        double fx = fmax(fabs(r.cx - lo[0]), fabs(r.cx - hi[0]));
                        double fy = fmax(fabs(r.cy - lo[1]), fabs(r.cy - hi[1]));
                        double fz = fmax(fabs(r.cz - lo[2]), fabs(r.cz - hi[2]));
                        double rmax2 = fx * fx + fy * fy + fz * fz;
                        double rad2 = rad * rad;
                        if (rmin2 > 0.0 && rad2 < th2 * rmin2 * (1.0 - 1e-10)) cls = 1;
                        else if (rmin2 > 0.0 && rad2 > th2 * rmax2 * (1.0 + 1e-10)) cls = 2;
                        else cls = 3;
                    }
                }
                const unsigned long long accB = __ballot(cls == 1);
                const unsigned long long opnB = __ballot(cls == 2);
                const unsigned long long mixB = __ballot(cls == 3);
                // whole-group accepts
                for (unsigned long long b = accB; b; b &= b - 1) {
                    int k = __builtin_ctzll(b);
                    unsigned long long A = M;
                    int pidx = rl(r.pidx, k);
                    if (pidx >= 0) {
                        int self = pidx - g * 64;
                        if ((unsigned)self < 64u) A &= ~(1ull << self);
                    }
                    if (A) {
                        if (lane == ni) { myNode = f + k; myMask = A; }
                        ni++;
                    }
                }
                // whole-group opens
                for (unsigned long long b = opnB; b; b &= b - 1) {
                    int k = __builtin_ctzll(b);
                    int fc = rl(r.first_child, k), nc = rl(r.nchild, k);
                    if (nc > 0) {
                        if (sp < kStack) {
                            if (lane == 0) { S.sfirst[sp] = fc; S.scount[sp] = nc; S.smask[sp] = M; }
                            sp++;
                        } else if (lane == 0) {
                            a.dsw->err = 3;   // unreachable: depth <= 44 levels x 7 pending siblings
                        }
                    }
                }
                // per-lane decisions, exactly the reference's test
                for (unsigned long long b = mixB; b; b &= b - 1) {
                    int k = __builtin_ctzll(b);
                    double ncx = rl(r.cx, k), ncy = rl(r.cy, k), ncz = rl(r.cz, k), nrad = rl(rad, k);
                    double nrad2 = nrad * nrad;
                    bool acc = false, opn = false;
                    if ((M >> lane) & 1ull) {
                        double dx = ncx - px, dy = ncy - py, dz = ncz - pz;
                        double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 != 0.0) {
                            double rhs = th2 * r2;
                            if (nrad2 < rhs * (1.0 - 1e-10)) acc = true;
                            else if (nrad2 > rhs * (1.0 + 1e-10)) opn = true;
                            else {
                                double rr = sqrt(r2);
                                acc = (nrad / rr) < theta;
                                opn = !acc;
                            }
                        }
                    }
                    const unsigned long long A = __ballot(acc), O = __ballot(opn);
                    if (A) {
                        if (lane == ni) { myNode = f + k; myMask = A; }
                        ni++;
                    }
                    if (O) {
                        int fc = rl(r.first_child, k), nc = rl(r.nchild, k);
                        if (nc > 0) {
                            if (sp < kStack) {
                                if (lane == 0) { S.sfirst[sp] = fc; S.scount[sp] = nc; S.smask[sp] = O; }
                                sp++;
                            } else if (lane == 0) {
                                a.dsw->err = 3;
                            }
                        }
                    }
                }
            }
            if (ni > kBatch - 8 || (sp == 0 && ni > 0)) {
                // ---------------------------------------------- sweep
                double ecx = 0.0, ecy = 0.0, ecz = 0.0, egas = 0.0, evx = 0.0, evy = 0.0, evz = 0.0;
                if (lane < ni) {
                    const core::TreeNode w = a.tree[myNode];
                    ecx = w.cx; ecy = w.cy; ecz = w.cz;
                    S.emask[lane] = myMask;
                    if (F32) {
                        S.e.f[lane] = make_float4((float)((w.cx - gcx) * invL), (float)((w.cy - gcy) * invL),
                                                  (float)((w.cz - gcz) * invL), (float)(w.mass * invMu));
                    } else {
                        S.e.d.cx[lane] = w.cx; S.e.d.cy[lane] = w.cy; S.e.d.cz[lane] = w.cz; S.e.d.m[lane] = w.mass;
                    }
                    if (any_gas) {
                        egas = w.gas;
                        int fl;
                        if (w.pidx >= 0) {
                            evx = a.svx[w.pidx]; evy = a.svy[w.pidx]; evz = a.svz[w.pidx];
                            fl = a.stype[w.pidx] == 2 ? 1 : 0;
                        } else {
                            evx = a.mvx[w.pre]; evy = a.mvy[w.pre]; evz = a.mvz[w.pre];
                            fl = w.gas > 0.0 ? 1 : 0;
                        }
                        S.eflag[lane] = fl;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (F32) {
                    const float4 *fs = S.e.f;
                    float cx = 0.f, cy = 0.f, cz = 0.f;
                    for (int j = 0; j < ni; j++) {
                        const unsigned long long Mj = S.emask[j];
                        const float4 e = fs[j];
                        bool on = (Mj >> lane) & 1ull;
                        float dx = e.x - xr, dy = e.y - yr, dz = e.z - zr;
                        float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                        float s2 = r2 + e2f;
                        float f = e.w * __builtin_amdgcn_rsqf(r2 * s2 * s2);
                        bool ok = on && r2 != 0.f;
                        f = ok ? f : 0.f;
                        cx = __builtin_fmaf(f, dx, cx);
                        cy = __builtin_fmaf(f, dy, cy);
                        cz = __builtin_fmaf(f, dz, cz);
                        nint32 += ok ? 1u : 0u;
                        if (any_gas) {
                            bool want = ok && sph_lane && S.eflag[j];
                            if (__ballot(want)) {
                                double jx = rl(ecx, j), jy = rl(ecy, j), jz = rl(ecz, j);
                                double jg = rl(egas, j), jvx = rl(evx, j), jvy = rl(evy, j), jvz = rl(evz, j);
                                if (want) {
                                    const core::SphConst sk = core::sph_const(tg);
                                    double ddx = px - jx, ddy = py - jy, ddz = pz - jz;
                                    double rr;
                                    if (core::sph_inside(ddx * ddx + ddy * ddy + ddz * ddz, sk, &rr)) {
                                        s.nsph++;
                                        core::sph_term_fast(ddx, ddy, ddz, rr, jg, vx - jvx, vy - jvy, vz - jvz, sk, &s);
                                    }
                                }
                            }
                        }
                    }
                    float y = cx - cmpx, tt = sumx + y; cmpx = (tt - sumx) - y; sumx = tt;
                    y = cy - cmpy; tt = sumy + y; cmpy = (tt - sumy) - y; sumy = tt;
                    y = cz - cmpz; tt = sumz + y; cmpz = (tt - sumz) - y; sumz = tt;
                } else {
                    for (int j = 0; j < ni; j++) {
                        const unsigned long long Mj = S.emask[j];
                        bool on = (Mj >> lane) & 1ull;
                        double dx = S.e.d.cx[j] - px, dy = S.e.d.cy[j] - py, dz = S.e.d.cz[j] - pz;
                        double r2 = dx * dx + dy * dy + dz * dz;
                        bool ok = on && r2 != 0.0;
                        double gx = 0.0, gy = 0.0, gz = 0.0;
                        if (ok) {
                            if (a.wp.soft_live) {
                                ok = core::soft_force(S.e.d.m[j], dx, dy, dz, sqrt(r2), a.wp, &gx, &gy, &gz);
                            } else {
                                double ri = rsq64(r2);
                                double rr = r2 * ri;
                                double f = a.wp.G * S.e.d.m[j] * rcp64(rr * rr + e02) * ri;
                                gx = f * dx; gy = f * dy; gz = f * dz;
                                ok = !(gx != gx || gy != gy || gz != gz);
                            }
                        }
                        if (ok) {
                            s.ax += gx; s.ay += gy; s.az += gz;
                            nint32++;
                        }
                        if (any_gas) {
                            bool want = ok && sph_lane && S.eflag[j];
                            if (__ballot(want)) {
                                double jg = rl(egas, j), jvx = rl(evx, j), jvy = rl(evy, j), jvz = rl(evz, j);
                                if (want) {
                                    const core::SphConst sk = core::sph_const(tg);
                                    double rr;
                                    if (core::sph_inside(r2, sk, &rr)) {
                                        s.nsph++;
                                        core::sph_term_fast(-dx, -dy, -dz, rr, jg, vx - jvx, vy - jvy, vz - jvz, sk, &s);
                                    }
                                }
                            }
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                n_entries += (unsigned long long)ni;
                ni = 0;
            }
            if (sp == 0 && ni == 0) break;
        }
        if (lane == 0) {
            atomicAdd(&a.dsw->walk_entries, n_entries);
            atomicAdd(&a.dsw->walk_tests, n_tests);
        }
    }
    if (act) {
        double ox, oy, oz;
        if (F32) {
            double scale = a.nn > 0 ? a.wp.G * a.tree[0].mass * invL * invL : 0.0;
            ox = (double)sumx * scale + s.ax;
            oy = (double)sumy * scale + s.ay;
            oz = (double)sumz * scale + s.az;
        } else {
            ox = s.ax; oy = s.ay; oz = s.az;
        }
        a.ax[q] = ox; a.ay[q] = oy; a.az[q] = oz;
        a.dUdt[q] += s.du;
    }
    unsigned long long ci = wave_sum_u64((unsigned long long)nint32);
    unsigned long long cs = wave_sum_u64((unsigned long long)s.nsph);
    if (lane == 0) {
        if (ci) atomicAdd(&a.dsw->interactions, ci);
        if (cs) atomicAdd(&a.dsw->sph_pairs, cs);
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_block64(WalkArgs a) {
    walk_block<false>(a);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(AG2_WALK32_WAVES))) void k_walk_block32(WalkArgs a) {
    walk_block<true>(a);
}

// One lane per target, stackless pre-order walk (core::walk_fp64); used for
// the potential diagnostic.
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
    ds->walk_entries = 0ull;
    ds->walk_tests = 0ull;
}

WalkArgs walk_args(Ctx &c) {
    WalkArgs a;
    a.n = (int)c.n; a.nin = c.nin; a.nn = c.n_nodes;
    a.tree = c.tnode; a.idx = c.idx;
    a.sx = c.sx; a.sy = c.sy; a.sz = c.sz; a.sm = c.sm; a.stype = c.stype;
    a.svx = c.svx; a.svy = c.svy; a.svz = c.svz;
    a.mvx = c.n_gvx; a.mvy = c.n_gvy; a.mvz = c.n_gvz;
    a.h = c.h; a.rho = c.rho; a.P = c.P; a.active = c.active;
    a.ds = c.ds;
    a.wp = core::WalkParams{c.p.theta, c.p.e0, c.p.G, c.p.e0 < 2.147e13 ? 1 : 0};
    a.ax = c.ax; a.ay = c.ay; a.az = c.az; a.dUdt = c.dUdt;
    a.dsw = c.ds;
    return a;
}

}  // namespace

void forces(Ctx &c, int mode) {
    hipStream_t st = c.stream;
    int n = (int)c.n;
    if (c.timing) AG2_HIP(hipEventRecord(c.ev[4], st));
    if (mode == 0) {
        hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(1), 0, st, c.ds);
        WalkArgs a = walk_args(c);
        int groups = (n + 63) / 64;
        hipLaunchKernelGGL(k_walk_block64, dim3((groups + kWaves - 1) / kWaves), dim3(64 * kWaves), 0, st, a);
    } else {
        core::WalkParams wp{c.p.theta, c.p.e0, c.p.G, c.p.e0 < 2.147e13 ? 1 : 0};
        hipLaunchKernelGGL(k_walk64, dim3(grid_for(n, 256)), dim3(256), 0, st, n, c.nin, c.n_nodes, c.wnode, c.idx,
                           c.sx, c.sy, c.sz, c.sm, c.stype, c.svx, c.svy, c.svz, c.n_gvx, c.n_gvy, c.n_gvz, c.h,
                           c.rho, c.P, c.active, wp, mode, c.ax, c.ay, c.az, c.dUdt, c.pot, c.ds);
    }
    if (c.timing) AG2_HIP(hipEventRecord(c.ev[5], st));
    AG2_HIP(hipGetLastError());
}

void forces_fp32(Ctx &c, int mode) {
    if (mode != 0) return forces(c, mode);
    hipStream_t st = c.stream;
    int n = (int)c.n;
    if (c.timing) AG2_HIP(hipEventRecord(c.ev[4], st));
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(1), 0, st, c.ds);
    WalkArgs a = walk_args(c);
    int groups = (n + 63) / 64;
    hipLaunchKernelGGL(k_walk_block32, dim3((groups + kWaves - 1) / kWaves), dim3(64 * kWaves), 0, st, a);
    if (c.timing) AG2_HIP(hipEventRecord(c.ev[5], st));
    AG2_HIP(hipGetLastError());
}

}  // namespace ag2
