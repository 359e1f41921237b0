// C ABI of the engine (include/ag2.h): context lifetime, host<->device copies
// and the Tree-seam entry points.  No exception crosses the ABI.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "ag2_internal.h"

using ag2::Ctx;

namespace {

__global__ __launch_bounds__(256) void k_fill_f64(int n, double *a, double v) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ __launch_bounds__(256) void k_iota_u32(int n, uint32_t *a) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = (uint32_t)i;
}

// Resident slot r holds caller index rid[r].
__global__ __launch_bounds__(256) void k_to_caller_f64(int n, const uint32_t *rid, const double *src, double *dst) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) dst[rid[r]] = src[r];
}

__global__ __launch_bounds__(256) void k_to_caller_u8(int n, const uint32_t *rid, const uint8_t *src, uint8_t *dst) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) dst[rid[r]] = src[r];
}

const char *kNullCtx = "null context";

int fail(Ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

template <class F>
int guarded(Ctx *c, F &&f) {
    if (!c) return AG2_ERR_ARG;
    try {
        AG2_HIP(hipSetDevice(c->device));
        return f();
    } catch (const ag2::HipError &e) {
        return fail(c, AG2_ERR_HIP, e.msg);
    } catch (const ag2::CapacityError &e) {
        return fail(c, AG2_ERR_CAPACITY, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(c, AG2_ERR_CAPACITY, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(c, AG2_ERR_HIP, e.what());
    }
}


void free_all(Ctx &c) {
    ag2::reset_graphs(c);
    if (c.slab) (void)hipFree(c.slab);
    if (c.node_slab) (void)hipFree(c.node_slab);
    // allocated on demand
    void *ptrs[] = {c.sph_list, c.sph_cnt, c.gas_table, c.vd};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (uint64_t *p : c.kdeep)
        if (p) (void)hipFree(p);
    void *deep[] = {c.sk, c.sk_alt, c.sv, c.sv_alt, c.spos, c.srun};
    for (void *p : deep)
        if (p) (void)hipFree(p);
    if (c.hs_pinned) (void)hipHostFree(c.hs_pinned);
    for (auto &e : c.ev)
        if (e) (void)hipEventDestroy(e);
    if (c.ev_counts) (void)hipEventDestroy(c.ev_counts);
    if (c.stream && c.own_stream) (void)hipStreamDestroy(c.stream);
}

void alloc_all(Ctx &c, int64_t cap) {
    const size_t n = (size_t)cap;
    auto plan = [&](ag2::Carve &cv) {
        double **d[] = {&c.x, &c.y, &c.z, &c.vx, &c.vy, &c.vz, &c.m, &c.U, &c.mu, &c.ax, &c.ay, &c.az, &c.h, &c.rho,
                        &c.P, &c.T, &c.dUdt, &c.ts, &c.next, &c.pot, &c.stage, &c.sfr};
        for (double **p : d) cv(p, n * 8);
        for (double *&p : c.alt) cv(&p, n * 8);
        cv(&c.type, n); cv(&c.active, n); cv(&c.gapl, n);
        cv(&c.type_alt, n); cv(&c.active_alt, n); cv(&c.lcp, n); cv(&c.lk, n + 1);
        cv(&c.rid, n * 4); cv(&c.rid_alt, n * 4); cv(&c.inv, n * 4);
        cv(&c.gslot, n * 4); cv(&c.gidx, n * 4);
        cv(&c.gtiles, ((n + 1023) / 1024 + 1) * 4); cv(&c.gover, (n + 63) / 64 + 1);
        const size_t groups = (size_t)(n + 63) / 64 + 1;
        cv(&c.near_cnt, groups * 4);
        cv(&c.near_node, groups * ag2::kNearRecs * 4);
        cv(&c.near_mask, groups * ag2::kNearRecs * 8);
        cv(&c.key, n * 8); cv(&c.key_alt, n * 8); cv(&c.key2, n * 8);
        cv(&c.w0, n * 8);
        cv(&c.idx, n * 4); cv(&c.idx_alt, n * 4);
        const size_t nblk = (n + ag2::kSortTile - 1) / ag2::kSortTile + 1;
        cv(&c.hist, 256 * nblk * 4); cv(&c.dtotal, 256 * 4);
        cv(&c.cnt, (n + 1) * 4); cv(&c.base, (n + 1) * 4);
        cv(&c.bsum, ((n + 1023) / 1024 + 1) * 4);
        cv(&c.partials, 8 * 1024 * 8);
        cv(&c.ds, sizeof(ag2::DevState));
        cv(&c.cbins, ag2::kCounterBinBytes);
    };
    ag2::Carve size;
    plan(size);
    AG2_HIP(hipMalloc(&c.slab, size.off));
    ag2::Carve at;
    at.base = (char *)c.slab;
    plan(at);
    // on the context's stream: a memset on the null stream would wait for
    // every other stream's work (a pool admits while its round runs)
    AG2_HIP(hipMemsetAsync(c.cbins, 0, ag2::kCounterBinBytes, c.stream));
    AG2_HIP(hipHostMalloc(&c.hs_pinned, sizeof(ag2::DevState)));
    AG2_HIP(hipMemsetAsync(c.ds, 0, sizeof(ag2::DevState), c.stream));
    ag2::alloc_nodes(c, 2 * cap + 64);
    for (auto &e : c.ev) AG2_HIP(hipEventCreate(&e));
    AG2_HIP(hipEventCreateWithFlags(&c.ev_counts, hipEventDisableTiming));
}

void upload_field(Ctx &c, double *dst, const double *src, double dflt) {
    if (src) AG2_HIP(hipMemcpyAsync(dst, src, (size_t)c.n * 8, hipMemcpyHostToDevice, c.stream));
    else hipLaunchKernelGGL(k_fill_f64, dim3(ag2::grid_for(c.n, 256)), dim3(256), 0, c.stream, (int)c.n, dst, dflt);
}

// 0 when either event was never recorded (timing_events above 0); the
// failed query's error is cleared so that later launches do not report it.
float ev_ms(Ctx &c, int a, int b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c.ev[a], c.ev[b]) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return ms;
}

}  // namespace

namespace ag2 {

int check_params(const ag2_params *p, std::string &why) {
    if (!p) { why = "params is NULL"; return AG2_ERR_ARG; }
    if (!(p->theta >= 0)) { why = "theta must be >= 0"; return AG2_ERR_ARG; }
    if (!(p->e0 >= 0)) { why = "e0 must be >= 0"; return AG2_ERR_ARG; }
    if (p->split_threshold < 1) { why = "split_threshold must be >= 1"; return AG2_ERR_ARG; }
    if (p->precision != AG2_FP64 && p->precision != AG2_FP32_SPLIT) { why = "unknown precision"; return AG2_ERR_ARG; }
    if (p->min_dt > p->max_dt) { why = "min_dt > max_dt (Simulation.cpp:40-44)"; return AG2_ERR_ARG; }
    if (p->precision == AG2_FP32_SPLIT && core::soft_branch_live(p->e0)) {
        // the FP32 sweep evaluates only G m / (r^2 + e0^2); below this e0 the
        // reference's softening-kernel branch (Node.cpp:302, :354) can fire
        why = "the split-precision form needs e0 * 0.0001 >= 2147483647 m (below it the reference's "
              "softening-kernel branch, Node.cpp:302, can fire); use AG2_FP64";
        return AG2_ERR_ARG;
    }
    return AG2_OK;
}

// Device bytes alloc_all takes for a capacity (per-particle arrays, node
// arrays at 2 capacity + 64, sort histograms and tiles).
int64_t ctx_footprint(int64_t cap, int64_t n_gas) {
    int64_t n = cap, nodes = 2 * cap + 64;
    int64_t per_p = 22 * 8            // state, potential, staging, star-formation rate
                    + kStateD * 8     // permute buffers
                    + 7               // type, active, gap level, their copies, lcp, adjacent key levels
                    + 4 * 8 + 2 * 4   // keys, sort indices
                    + 3 * 4           // rid, rid copy, inverse
                    + 2 * 4           // c, base
                    + 2 * 4 + 1       // SPH list slots, targets, group flags
                    + (4 + ag2::kNearRecs * 12) / 64 + 1;   // near-pair records
    int64_t per_node = 5 * 4 + 1 + 9 * 8 + (int64_t)sizeof(WalkNode) + 2 * 4 + 1 + 4 +
                       2 * (int64_t)sizeof(core::TreeNode) + 3 * 4 +   // walk and SPH records
                       (int64_t)sizeof(ag2::NodeSums) + 4 +            // node sums, internal-child counts
                       (int64_t)sizeof(ag2::NodeSums) / 64 + 1;        // block sums
    int64_t fixed = 256 * ((n + kSortTile - 1) / kSortTile + 1) * 4 + 8 * 1024 * 8 + (1 << 20) +
                    128 * (4096 + 256);   // the carved buffers' slack and alignment
    return n * per_p + nodes * per_node + fixed + n_gas * (kSphListLen * 4 + 4);
}

Ctx *ctx_new(const ag2_params &p, int64_t capacity, hipStream_t stream) {
    Ctx *c = new Ctx();
    c->p = p;
    c->timing = 2 - (p.timing_events < 0 ? 0 : (p.timing_events > 2 ? 2 : p.timing_events));
    c->device = p.device;
    c->cap = capacity;
    // captured sort segments keep their passes: graph launches sort every byte
    c->sort_lo = p.graph_launch ? 0 : ag2::kSortLoMax;
    try {
        AG2_HIP(hipSetDevice(c->device));
        // once per device, with the device current: the address of the
        // no-allocation direct-sum scratch
        direct_scratch_init(c->device);
        if (stream) {
            c->stream = stream;
            c->own_stream = false;
        } else {
            AG2_HIP(hipStreamCreate(&c->stream));
        }
        alloc_all(*c, capacity);
    } catch (...) {
        free_all(*c);
        delete c;
        throw;
    }
    return c;
}

void ctx_delete(Ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_all(*c);
    delete c;
}

void ctx_upload(Ctx &ctx, const ag2_particles &h) {
    const ag2_particles *hp = &h;
    Ctx *c = &ctx;
    {
        ag2::reset_graphs(*c);   // the particle count is an argument of the captured segments
        c->n = hp->n;
        upload_field(*c, c->x, hp->x, 0); upload_field(*c, c->y, hp->y, 0); upload_field(*c, c->z, hp->z, 0);
        upload_field(*c, c->vx, hp->vx, 0); upload_field(*c, c->vy, hp->vy, 0); upload_field(*c, c->vz, hp->vz, 0);
        upload_field(*c, c->m, hp->m, 0); upload_field(*c, c->U, hp->U, 0); upload_field(*c, c->mu, hp->mu, 0.58);
        upload_field(*c, c->ax, hp->ax, 0); upload_field(*c, c->ay, hp->ay, 0); upload_field(*c, c->az, hp->az, 0);
        upload_field(*c, c->h, hp->h, 0); upload_field(*c, c->rho, hp->rho, 0); upload_field(*c, c->P, hp->P, 0);
        upload_field(*c, c->T, hp->T, 0); upload_field(*c, c->dUdt, hp->dUdt, 0);
        upload_field(*c, c->ts, nullptr, 0); upload_field(*c, c->next, nullptr, 0);
        upload_field(*c, c->sfr, nullptr, 0);
        if (hp->type) AG2_HIP(hipMemcpyAsync(c->type, hp->type, (size_t)c->n, hipMemcpyHostToDevice, c->stream));
        else AG2_HIP(hipMemsetAsync(c->type, 1, (size_t)c->n, c->stream));
        ag2::fill_active(*c, 1);
        hipLaunchKernelGGL(k_iota_u32, dim3(ag2::grid_for(c->n, 256)), dim3(256), 0, c->stream, (int)c->n, c->rid);
        AG2_HIP(hipMemsetAsync(c->ds, 0, sizeof(ag2::DevState), c->stream));
        // running gas-mass sums of equal-mass gas (see Ctx::gas_table)
        int64_t ng = 0;
        double mg = 0.0;
        bool eq = true;
        if (hp->type)
            for (int64_t i = 0; i < hp->n; i++)
                if (hp->type[i] == 2) {
                    if (ng == 0) mg = hp->m[i];
                    else if (hp->m[i] != mg) eq = false;
                    ng++;
                }
        int64_t nz = 0;
        for (int64_t i = 0; i < hp->n; i++) nz += hp->m[i] == 0.0 ? 1 : 0;
        c->zero_mass = nz;
        c->gas_equal = (ng > 0 && eq) ? 1 : 0;
        c->sph_active = ng > 0;
        c->gas_count = ng;
        if (ng > c->sph_cap) {
            // SPH candidate lists, one per gas particle
            if (c->sph_list) { AG2_HIP(hipFree(c->sph_list)); c->sph_list = nullptr; }
            if (c->sph_cnt) { AG2_HIP(hipFree(c->sph_cnt)); c->sph_cnt = nullptr; }
            c->sph_cap = 0;
            const int64_t cap = ng > c->sph_reserve ? ng : c->sph_reserve;
            AG2_HIP(hipMalloc(&c->sph_list, (size_t)cap * kSphListLen * 4));
            AG2_HIP(hipMalloc(&c->sph_cnt, (size_t)cap * 4));
            c->sph_cap = cap;
        }
        if (c->gas_table) { AG2_HIP(hipFree(c->gas_table)); c->gas_table = nullptr; }
        if (c->gas_equal) {
            std::vector<double> tab((size_t)ng + 1);
            tab[0] = 0.0;
            for (int64_t k = 1; k <= ng; k++) tab[(size_t)k] = tab[(size_t)k - 1] + mg;
            AG2_HIP(hipMalloc(&c->gas_table, ((size_t)ng + 1) * 8));
            AG2_HIP(hipMemcpyAsync(c->gas_table, tab.data(), ((size_t)ng + 1) * 8, hipMemcpyHostToDevice, c->stream));
            AG2_HIP(hipStreamSynchronize(c->stream));
        }
        AG2_HIP(hipStreamSynchronize(c->stream));
        c->tree_valid = false;
        c->started = false;
        c->steps = 0;
        c->vd_radius0 = 0.0;
    }
}

void ctx_download(Ctx &ctx, ag2_particles &h) {
    ag2_particles *hp = &h;
    Ctx *c = &ctx;
    {
        struct F { double *h; double *d; } fs[] = {{hp->x, c->x}, {hp->y, c->y}, {hp->z, c->z}, {hp->vx, c->vx},
            {hp->vy, c->vy}, {hp->vz, c->vz}, {hp->m, c->m}, {hp->U, c->U}, {hp->mu, c->mu}, {hp->ax, c->ax},
            {hp->ay, c->ay}, {hp->az, c->az}, {hp->h, c->h}, {hp->rho, c->rho}, {hp->P, c->P}, {hp->T, c->T},
            {hp->dUdt, c->dUdt}};
        // resident (key) order -> caller order through rid, one field at a
        // time via the staging buffer (stream order keeps the copies apart)
        unsigned gb = ag2::grid_for(c->n, 256);
        for (auto &f : fs)
            if (f.h) {
                hipLaunchKernelGGL(k_to_caller_f64, dim3(gb), dim3(256), 0, c->stream, (int)c->n, c->rid, f.d, c->stage);
                AG2_HIP(hipMemcpyAsync(f.h, c->stage, (size_t)c->n * 8, hipMemcpyDeviceToHost, c->stream));
            }
        if (hp->type) {
            uint8_t *st8 = (uint8_t *)c->stage;
            hipLaunchKernelGGL(k_to_caller_u8, dim3(gb), dim3(256), 0, c->stream, (int)c->n, c->rid, c->type, st8);
            AG2_HIP(hipMemcpyAsync(hp->type, st8, (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
        }
        AG2_HIP(hipStreamSynchronize(c->stream));
    }
}

}  // namespace ag2

extern "C" {

int ag2_create(const ag2_params *params, int64_t capacity, void **out) {
    if (!out) return AG2_ERR_ARG;
    *out = nullptr;
    std::string why;
    int rc = ag2::check_params(params, why);
    if (rc) return rc;
    if (capacity < 1 || capacity > 0x7fffff00LL / 2) return AG2_ERR_CAPACITY;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return AG2_ERR_NODEV;
    if (params->device < 0 || params->device >= ndev) return AG2_ERR_NODEV;
    try {
        *out = ag2::ctx_new(*params, capacity, nullptr);
    } catch (const ag2::HipError &) {
        return AG2_ERR_HIP;
    } catch (const std::bad_alloc &) {
        return AG2_ERR_CAPACITY;
    }
    return AG2_OK;
}

int ag2_destroy(void *ctx) {
    Ctx *c = (Ctx *)ctx;
    if (!c) return AG2_ERR_ARG;
    ag2::ctx_delete(c);
    return AG2_OK;
}

const char *ag2_last_error(const void *ctx) {
    const Ctx *c = (const Ctx *)ctx;
    return c ? c->err.c_str() : kNullCtx;
}

int ag2_set_params(void *ctx, const ag2_params *params) {
    Ctx *c = (Ctx *)ctx;
    if (!c) return AG2_ERR_ARG;
    std::string why;
    int rc = ag2::check_params(params, why);
    if (rc) return fail(c, rc, why);
    if (params->device != c->device) return fail(c, AG2_ERR_ARG, "device cannot change");
    c->p = *params;
    c->timing = 2 - (params->timing_events < 0 ? 0 : (params->timing_events > 2 ? 2 : params->timing_events));
    ag2::reset_graphs(*c);
    return AG2_OK;
}

int ag2_upload(void *ctx, const ag2_particles *hp) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!hp || !hp->x || !hp->y || !hp->z || !hp->m) return fail(c, AG2_ERR_ARG, "x, y, z and m are required");
        if (hp->n < 1) return fail(c, AG2_ERR_ARG, "n must be >= 1");
        if (hp->n > c->cap) return fail(c, AG2_ERR_CAPACITY, "n exceeds the context capacity");
        ag2::ctx_upload(*c, *hp);
        return AG2_OK;
    });
}

int ag2_download(void *ctx, ag2_particles *hp) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!hp) return fail(c, AG2_ERR_ARG, "host particles is NULL");
        if (hp->n != c->n) return fail(c, AG2_ERR_ARG, "n differs from the uploaded count");
        ag2::ctx_download(*c, *hp);
        return AG2_OK;
    });
}

int ag2_build_tree(void *ctx) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (c->n < 1) return fail(c, AG2_ERR_STATE, "no particles uploaded");
        c->density_follows = c->acc_overwritten = false;   // the state stays with its particles
        ag2::build_tree(*c);
        return AG2_OK;
    });
}

int ag2_gas_density(void *ctx) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!c->tree_valid) return fail(c, AG2_ERR_STATE, "build the tree first");
        ag2::gas_density(*c);
        return AG2_OK;
    });
}

static void run_forces(Ctx &c, int mode) {
    c.fp64_fallback = c.p.precision == AG2_FP32_SPLIT && !ag2::fp32_usable(c);
    if (c.p.precision == AG2_FP32_SPLIT && !c.fp64_fallback) ag2::forces_fp32(c, mode);
    else ag2::forces(c, mode);
}

int ag2_forces(void *ctx) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!c->tree_valid) return fail(c, AG2_ERR_STATE, "build the tree first");
        run_forces(*c, 0);
        AG2_HIP(hipStreamSynchronize(c->stream));
        return AG2_OK;
    });
}

int ag2_init_forces(void *ctx) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (c->n < 1) return fail(c, AG2_ERR_STATE, "no particles uploaded");
        ag2::fill_active(*c, 1);
        {
            ag2::StepBuild sb(*c, true);   // every particle active
            ag2::build_tree(*c);
        }
        ag2::gas_density(*c);
        run_forces(*c, 0);
        AG2_HIP(hipStreamSynchronize(c->stream));
        return AG2_OK;
    });
}

int ag2_step(void *ctx) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (c->n < 1) return fail(c, AG2_ERR_STATE, "no particles uploaded");
        if (!c->started) {
            ag2::kdk_init_steps(*c);
            c->started = true;
        }
        if (c->timing >= 2) AG2_HIP(hipEventRecord(c->ev[6], c->stream));
        ag2::kdk_begin(*c);
        {
            // one step size for all: every particle is active in every step
            ag2::StepBuild sb(*c, c->p.min_dt >= c->p.max_dt);
            ag2::build_tree(*c);
        }
        ag2::gas_density(*c);
        run_forces(*c, 0);
        ag2::kdk_end(*c);
        if (c->timing >= 2) AG2_HIP(hipEventRecord(c->ev[7], c->stream));
        c->steps++;
        return AG2_OK;
    });
}

// The walk's stack-overflow flag (ds->walk_err): the build bounds the
// stack for every tree it accepts, so the flag is never expected; if a walk
// ever sets it, its forces are incomplete and the caller hears of it at the
// next sync or diagnostics read.
static int walk_error(Ctx *c) {
    if (c->hs_pinned->walk_err)
        return fail(c, AG2_ERR_CAPACITY, "a gravity walk overflowed its stack: forces of that step are incomplete");
    return AG2_OK;
}

int ag2_sync(void *ctx) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        AG2_HIP(hipMemcpyAsync(&c->hs_pinned->walk_err, &c->ds->walk_err, sizeof(int32_t), hipMemcpyDeviceToHost,
                               c->stream));
        AG2_HIP(hipStreamSynchronize(c->stream));
        return walk_error(c);
    });
}

int ag2_energy(void *ctx, int with_potential, ag2_diag *out) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!out) return fail(c, AG2_ERR_ARG, "out is NULL");
        if (with_potential) {
            if (!c->tree_valid) return fail(c, AG2_ERR_STATE, "build the tree first");
            ag2::forces(*c, 1);
        }
        double e[6];
        ag2::energies(*c, with_potential ? 1 : 0, e);
        out->kinetic = e[0];
        out->thermal = e[1];
        out->momentum[0] = e[2]; out->momentum[1] = e[3]; out->momentum[2] = e[4];
        out->potential = with_potential ? e[5] : 0.0;
        return AG2_OK;
    });
}

int ag2_diagnostics(void *ctx, ag2_diag *out) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!out) return fail(c, AG2_ERR_ARG, "out is NULL");
        AG2_HIP(hipMemcpyAsync(&c->hs, c->ds, sizeof(ag2::DevState), hipMemcpyDeviceToHost, c->stream));
        AG2_HIP(hipStreamSynchronize(c->stream));
        const ag2::DevState &s = c->hs;
        out->global_time = s.gtime;
        out->steps = c->steps;
        out->R = __builtin_bit_cast(double, s.Rbits);
        out->n_in_tree = s.nin;
        out->n_nodes = s.n_nodes;
        out->max_level = s.max_level;
        out->interactions = (int64_t)s.interactions;
        out->sph_pairs = (int64_t)s.sph_pairs;
        out->groups = s.final_count;
        out->gap_particles = (int64_t)s.gap;
        out->nan_kicks = (int64_t)s.nan_kicks;
        out->walk_entries = (int64_t)s.walk_entries;
        out->walk_tests = (int64_t)s.walk_tests;
        out->walk_mixed = (int64_t)s.walk_mixed;
        out->interactions_total = (int64_t)s.interactions_total;
        out->sph_overflow = (int64_t)s.sph_overflow;
        out->fp64_fallback = c->fp64_fallback ? 1 : 0;
        out->near_pairs = (int64_t)s.near_pairs;
        out->near_redo_groups = (int64_t)s.near_redo;
        c->hs_pinned->walk_err = s.walk_err;
        return walk_error(c);
    });
}

int ag2_phase_times(void *ctx, double *ms) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!ms) return fail(c, AG2_ERR_ARG, "ms is NULL");
        AG2_HIP(hipStreamSynchronize(c->stream));
        ms[0] = ev_ms(*c, 0, 1);
        ms[1] = ev_ms(*c, 1, 2);
        ms[2] = ev_ms(*c, 2, 3);
        ms[3] = ev_ms(*c, 3, 4);
        ms[4] = ev_ms(*c, 4, 5);
        ms[5] = c->steps > 0 ? (double)ev_ms(*c, 6, 7) - (ms[0] + ms[1] + ms[2] + ms[3] + ms[4]) : 0.0;
        return AG2_OK;
    });
}

int ag2_walk_kernel_ms(void *ctx, double *ms) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!ms) return fail(c, AG2_ERR_ARG, "ms is NULL");
        AG2_HIP(hipStreamSynchronize(c->stream));
        *ms = ev_ms(*c, 8, 9);
        return AG2_OK;
    });
}

int ag2_direct_sum(void *ctx, int64_t nt, const int64_t *targets, double *ax, double *ay, double *az, double *pot) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (nt < 1 || !targets || !ax || !ay || !az) return fail(c, AG2_ERR_ARG, "targets and outputs are required");
        if (c->n < 1) return fail(c, AG2_ERR_STATE, "no particles uploaded");
        for (int64_t k = 0; k < nt; k++)
            if (targets[k] < 0 || targets[k] >= c->n) return fail(c, AG2_ERR_ARG, "target index out of range");
        int64_t *dt = nullptr;
        double *dout = nullptr;
        AG2_HIP(hipMalloc(&dt, (size_t)nt * 8));
        hipError_t e = hipMalloc(&dout, (size_t)nt * 4 * 8);
        if (e != hipSuccess) { (void)hipFree(dt); AG2_HIP(e); }
        try {
            AG2_HIP(hipMemcpyAsync(dt, targets, (size_t)nt * 8, hipMemcpyHostToDevice, c->stream));
            ag2::direct_sum(*c, (int)nt, dt, dout, pot ? 1 : 0);
            AG2_HIP(hipMemcpyAsync(ax, dout, (size_t)nt * 8, hipMemcpyDeviceToHost, c->stream));
            AG2_HIP(hipMemcpyAsync(ay, dout + nt, (size_t)nt * 8, hipMemcpyDeviceToHost, c->stream));
            AG2_HIP(hipMemcpyAsync(az, dout + 2 * nt, (size_t)nt * 8, hipMemcpyDeviceToHost, c->stream));
            if (pot) AG2_HIP(hipMemcpyAsync(pot, dout + 3 * nt, (size_t)nt * 8, hipMemcpyDeviceToHost, c->stream));
            AG2_HIP(hipStreamSynchronize(c->stream));
        } catch (...) {
            (void)hipFree(dt);
            (void)hipFree(dout);
            throw;
        }
        (void)hipFree(dt);
        (void)hipFree(dout);
        return AG2_OK;
    });
}

int ag2_direct_sum_noalloc(void *ctx, int64_t nt, const int64_t *targets, double *ax, double *ay, double *az,
                           double *pot) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (nt < 1 || !targets || !ax || !ay || !az) return fail(c, AG2_ERR_ARG, "targets and outputs are required");
        if (nt > AG2_DIRECT_NOALLOC_MAX)
            return fail(c, AG2_ERR_CAPACITY, "ag2_direct_sum_noalloc takes at most AG2_DIRECT_NOALLOC_MAX (65536) "
                                             "targets; use ag2_direct_sum for more");
        if (c->n < 1) return fail(c, AG2_ERR_STATE, "no particles uploaded");
        for (int64_t k = 0; k < nt; k++)
            if (targets[k] < 0 || targets[k] >= c->n) return fail(c, AG2_ERR_ARG, "target index out of range");
        ag2::direct_sum_noalloc(*c, (int)nt, targets, ax, ay, az, pot);
        return AG2_OK;
    });
}

int ag2_visual_density(void *ctx, double radius, double *out) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!c->tree_valid) return fail(c, AG2_ERR_STATE, "build the tree first");
        if (!(radius >= 0.0) || radius == INFINITY) return fail(c, AG2_ERR_ARG, "radius must be finite and >= 0");
        if (!c->vd) AG2_HIP(hipMalloc(&c->vd, (size_t)c->cap * 8));
        ag2::visual_density(*c, radius > 0.0 ? radius : c->vd_radius0, c->vd);
        if (out) {
            hipLaunchKernelGGL(k_to_caller_f64, dim3(ag2::grid_for(c->n, 256)), dim3(256), 0, c->stream, (int)c->n,
                               c->rid, c->vd, c->stage);
            AG2_HIP(hipMemcpyAsync(out, c->stage, (size_t)c->n * 8, hipMemcpyDeviceToHost, c->stream));
        }
        AG2_HIP(hipStreamSynchronize(c->stream));
        return AG2_OK;
    });
}

int ag2_get_sfr(void *ctx, double *out) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!out) return fail(c, AG2_ERR_ARG, "out is required");
        if (c->n < 1) return fail(c, AG2_ERR_STATE, "no particles uploaded");
        // kept in caller order by the second kick
        AG2_HIP(hipMemcpyAsync(out, c->sfr, (size_t)c->n * 8, hipMemcpyDeviceToHost, c->stream));
        AG2_HIP(hipStreamSynchronize(c->stream));
        return AG2_OK;
    });
}

int ag2_get_order(void *ctx, int64_t *perm, uint64_t *key_hi, uint64_t *key_lo) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!c->tree_valid) return fail(c, AG2_ERR_STATE, "build the tree first");
        AG2_HIP(hipStreamSynchronize(c->stream));
        size_t nin = (size_t)c->nin;
        std::vector<uint32_t> ix(nin);
        if (nin) AG2_HIP(hipMemcpy(ix.data(), c->rid, nin * 4, hipMemcpyDeviceToHost));
        if (perm)
            for (size_t i = 0; i < nin; i++) perm[i] = ix[i];
        if (key_hi && nin) AG2_HIP(hipMemcpy(key_hi, c->w0, nin * 8, hipMemcpyDeviceToHost));
        if (key_lo && nin) AG2_HIP(hipMemcpy(key_lo, c->key_alt, nin * 8, hipMemcpyDeviceToHost));
        return AG2_OK;
    });
}

int ag2_get_tree(void *ctx, int64_t cap, int32_t *level, int64_t *first, int64_t *count, double *mass, double *com,
                 double *gas, double *mvel, int64_t *n_nodes) {
    Ctx *c = (Ctx *)ctx;
    return guarded(c, [&]() -> int {
        if (!c->tree_valid) return fail(c, AG2_ERR_STATE, "build the tree first");
        AG2_HIP(hipStreamSynchronize(c->stream));
        size_t nn = (size_t)c->n_nodes;
        if (n_nodes) *n_nodes = (int64_t)nn;
        if ((int64_t)nn > cap) return fail(c, AG2_ERR_CAPACITY, "cap is smaller than the node count");
        std::vector<int8_t> lv(nn);
        std::vector<int32_t> f(nn), k(nn);
        std::vector<double> cx(nn), cy(nn), cz(nn), vx(nn), vy(nn), vz(nn);
        if (nn) {
            AG2_HIP(hipMemcpy(lv.data(), c->n_level, nn, hipMemcpyDeviceToHost));
            AG2_HIP(hipMemcpy(f.data(), c->n_first, nn * 4, hipMemcpyDeviceToHost));
            AG2_HIP(hipMemcpy(k.data(), c->n_count, nn * 4, hipMemcpyDeviceToHost));
            if (mass) AG2_HIP(hipMemcpy(mass, c->n_mass, nn * 8, hipMemcpyDeviceToHost));
            if (gas) AG2_HIP(hipMemcpy(gas, c->n_gas, nn * 8, hipMemcpyDeviceToHost));
            AG2_HIP(hipMemcpy(cx.data(), c->n_cx, nn * 8, hipMemcpyDeviceToHost));
            AG2_HIP(hipMemcpy(cy.data(), c->n_cy, nn * 8, hipMemcpyDeviceToHost));
            AG2_HIP(hipMemcpy(cz.data(), c->n_cz, nn * 8, hipMemcpyDeviceToHost));
            AG2_HIP(hipMemcpy(vx.data(), c->n_gvx, nn * 8, hipMemcpyDeviceToHost));
            AG2_HIP(hipMemcpy(vy.data(), c->n_gvy, nn * 8, hipMemcpyDeviceToHost));
            AG2_HIP(hipMemcpy(vz.data(), c->n_gvz, nn * 8, hipMemcpyDeviceToHost));
        }
        for (size_t i = 0; i < nn; i++) {
            if (level) level[i] = lv[i];
            if (first) first[i] = f[i];
            if (count) count[i] = k[i];
            if (com) { com[3 * i] = cx[i]; com[3 * i + 1] = cy[i]; com[3 * i + 2] = cz[i]; }
            if (mvel) { mvel[3 * i] = vx[i]; mvel[3 * i + 1] = vy[i]; mvel[3 * i + 2] = vz[i]; }
        }
        return AG2_OK;
    });
}

}  // extern "C"
