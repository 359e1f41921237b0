// Serving pool: many independent simulations (jobs) resident on one device,
// advanced together with continuous batching.  The reference runs one
// simulation per process (main.cpp:15-24); a job here is one such run (its
// particles, tree and integrator state), and a round advances every resident
// job by one step of Simulation::run (Simulation.cpp:210-356).
//
// A round:
//   1. on a host thread of its own, alongside phases 2-6: the jobs that
//      finished in the last round are downloaded and leave (their contexts
//      become spare), then queued jobs, first come first served, take the
//      slots this round frees while the job count and the device-memory
//      budget allow (a spare context is reused when it is large enough); an
//      admitted job uploads and runs its initial force phase on the
//      admission stream, and steps from the next round.  A pool with no
//      running job does this first, and its admitted jobs step at once;
//   2. every resident job: time steps, first kick and drift, tree width and
//      keys, on its side stream (job k on stream k mod 8); one batched radix
//      sort of all jobs' keys on the pool stream; then each job's ordering,
//      state permutation and node count on its side stream;
//   3. each job's node counts reach the host (an event per job);
//   4. every job: the rest of its tree and its SPH density;
//   5. one batched walk launch over all jobs of a precision form;
//   6. every job: U update and second kick; finished jobs leave the round.
// Phases 2-4 and 6 are issued by one host thread per side stream, so the
// sixty-odd launches per job of a round are enqueued in parallel.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <functional>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "ag2_internal.h"

namespace {

using ag2::Ctx;

constexpr int kFields = 17;   // double fields of ag2_particles, x .. dUdt

double *field(ag2_particles &h, int k) {
    double **f[kFields] = {&h.x, &h.y, &h.z, &h.vx, &h.vy, &h.vz, &h.m, &h.U, &h.mu,
                           &h.ax, &h.ay, &h.az, &h.h, &h.rho, &h.P, &h.T, &h.dUdt};
    return *f[k];
}
const double *field(const ag2_particles &h, int k) { return field(const_cast<ag2_particles &>(h), k); }
void set_field(ag2_particles &h, int k, double *p) {
    double **f[kFields] = {&h.x, &h.y, &h.z, &h.vx, &h.vy, &h.vz, &h.m, &h.U, &h.mu,
                           &h.ax, &h.ay, &h.az, &h.h, &h.rho, &h.P, &h.T, &h.dUdt};
    *f[k] = p;
}

struct Job {
    int64_t id = 0;
    ag2_params p{};
    int64_t n = 0, n_gas = 0, steps_total = 0, steps_done = 0;
    std::vector<double> f[kFields];   // input, then the final state (caller order)
    std::vector<uint8_t> type;
    int state = AG2_JOB_QUEUED;
    Ctx *ctx = nullptr;
    std::string err;
    // end of the job's last completed step (or the start of the round that
    // admitted it): its next step interval runs from here
    std::chrono::steady_clock::time_point t_prev{};
    bool timed = false;
};

// Host threads that issue per-stream work: run(f) calls f(k) on worker k for
// every k, waits for all of them and rethrows the first error.
class Issuers {
  public:
    void start(int n, int device) {
        for (int k = 0; k < n; k++) th_.emplace_back([this, k, device] { loop(k, device); });
    }
    void run(const std::function<void(int)> &f) {
        {
            std::lock_guard<std::mutex> lk(m_);
            task_ = &f;
            pending_ = (int)th_.size();
            err_ = nullptr;
            gen_++;
        }
        go_.notify_all();
        std::exception_ptr e;
        {
            std::unique_lock<std::mutex> lk(m_);
            done_.wait(lk, [&] { return pending_ == 0; });
            e = err_;
        }
        if (e) std::rethrow_exception(e);
    }
    void stop() {
        {
            std::lock_guard<std::mutex> lk(m_);
            quit_ = true;
        }
        go_.notify_all();
        for (auto &t : th_) t.join();
        th_.clear();
    }
    ~Issuers() { stop(); }

  private:
    void loop(int k, int device) {
        (void)hipSetDevice(device);   // the current device is per host thread
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)> *t;
            {
                std::unique_lock<std::mutex> lk(m_);
                go_.wait(lk, [&] { return quit_ || gen_ != seen; });
                if (quit_) return;
                seen = gen_;
                t = task_;
            }
            std::exception_ptr e;
            try {
                (*t)(k);
            } catch (...) {
                e = std::current_exception();
            }
            std::lock_guard<std::mutex> lk(m_);
            if (e && !err_) err_ = e;
            if (--pending_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable go_, done_;
    const std::function<void(int)> *task_ = nullptr;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool quit_ = false;
    std::exception_ptr err_;
};

struct Pool {
    int device = 0;
    int max_jobs = 16;
    int64_t budget = 0, used = 0;
    hipStream_t stream = nullptr;                 // admission into an idle pool, the batched sort and walk
    hipStream_t astream = nullptr;                // admission while a round runs
    static constexpr int kSide = 8;
    hipStream_t side[kSide] = {};                 // per-job phases, spread round robin
    hipEvent_t ev_side[kSide] = {}, ev_walk = nullptr;
    Issuers issuers;                              // one host thread per side stream
    void *jw_dev = nullptr, *jw_host = nullptr;   // job records of the batched walk
    void *sj_dev = nullptr, *sj_host = nullptr;   // job records of the batched sort
    std::mutex mu;
    std::string err;
    int64_t next_id = 1;
    std::deque<Job *> queue;
    std::vector<Job *> running;
    std::vector<Job *> retiring;                  // finished in the last round, downloaded during the next
    std::map<int64_t, Job *> jobs;
    std::vector<Ctx *> spare;                     // contexts of finished jobs
    std::vector<double> step_ms;                  // one sample per job-step: its step interval
    int64_t rounds = 0, job_steps = 0, particle_steps = 0, done = 0;
    int64_t resident_sum = 0, resident_max = 0;   // resident jobs over the rounds
    double busy_s = 0.0;
};

int pfail(Pool *p, int code, const std::string &m) {
    p->err = m;
    return code;
}

template <class F>
int pguard(Pool *p, F &&f) {
    if (!p) return AG2_ERR_ARG;
    std::lock_guard<std::mutex> lk(p->mu);
    try {
        AG2_HIP(hipSetDevice(p->device));
        return f();
    } catch (const ag2::HipError &e) {
        return pfail(p, AG2_ERR_HIP, e.msg);
    } catch (const ag2::CapacityError &e) {
        return pfail(p, AG2_ERR_CAPACITY, e.msg);
    } catch (const std::bad_alloc &) {
        return pfail(p, AG2_ERR_CAPACITY, "host allocation failed");
    } catch (const std::exception &e) {   // host threads and the like
        return pfail(p, AG2_ERR_HIP, e.what());
    }
}

int64_t ngas_of(const ag2_particles &h) {
    int64_t k = 0;
    if (h.type)
        for (int64_t i = 0; i < h.n; i++) k += h.type[i] == 2 ? 1 : 0;
    return k;
}

ag2_particles view(Job &j) {
    ag2_particles h{};
    h.n = j.n;
    for (int k = 0; k < kFields; k++) set_field(h, k, j.f[k].empty() ? nullptr : j.f[k].data());
    h.type = j.type.data();
    return h;
}

int64_t pow2_at_least(int64_t v) {
    int64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// A context for job j: the smallest spare one that is large enough, else a
// new one if the budget allows (dropping spare contexts to make room).  New
// contexts are sized in classes (particles and gas targets rounded up to a
// power of two, from 2^14 and 2^12), so that later jobs of up to that size
// reuse them instead of allocating; exactly when the class does not fit the
// budget.
Ctx *take_context(Pool &P, Job &j, hipStream_t st) {
    Ctx *best = nullptr;
    size_t bi = 0;
    for (size_t i = 0; i < P.spare.size(); i++)
        if (P.spare[i]->cap >= j.n && P.spare[i]->sph_cap >= j.n_gas && (!best || P.spare[i]->cap < best->cap)) {
            best = P.spare[i];
            bi = i;
        }
    if (best) {
        P.spare.erase(P.spare.begin() + (ptrdiff_t)bi);
        best->p = j.p;
        return best;
    }
    int64_t cap = pow2_at_least(j.n > (1 << 14) ? j.n : (1 << 14));
    int64_t gcap = j.n_gas > 0 ? pow2_at_least(j.n_gas > (1 << 12) ? j.n_gas : (1 << 12)) : 0;
    int64_t need = ag2::ctx_footprint(cap, gcap);
    if (P.used + need > P.budget) {
        cap = j.n;
        gcap = j.n_gas;
        need = ag2::ctx_footprint(cap, gcap);
    }
    while (P.used + need > P.budget && !P.spare.empty()) {
        Ctx *c = P.spare.back();
        P.spare.pop_back();
        P.used -= ag2::ctx_footprint(c->cap, c->sph_cap);
        ag2::ctx_delete(c);
    }
    if (P.used + need > P.budget) return nullptr;
    Ctx *c = ag2::ctx_new(j.p, cap, st);
    c->sph_reserve = gcap;
    c->timing = 0;
    P.used += need;
    return c;
}

// Up to `slots` queued jobs, first come first served, take a context, upload
// and run their initial forces on stream st; they join `into`.  No further job
// is started once *stop is set.
void admit(Pool &P, int slots, hipStream_t st, std::vector<Job *> &into, const std::atomic<bool> *stop = nullptr) {
    while (!P.queue.empty() && slots > 0 && !(stop && stop->load())) {
        Job *j = P.queue.front();
        Ctx *c = take_context(P, *j, st);
        if (!c) break;   // first come first served: wait for memory
        P.queue.pop_front();
        j->ctx = c;
        c->stream = st;
        ag2_particles h = view(*j);
        ag2::ctx_upload(*c, h);
        ag2::fill_active(*c, 1);
        {
            ag2::StepBuild sb(*c, true);   // every particle active
            ag2::build_tree(*c);
        }
        ag2::gas_density(*c);
        c->fp64_fallback = c->p.precision == AG2_FP32_SPLIT && !ag2::fp32_usable(*c);
        if (c->p.precision == AG2_FP32_SPLIT && !c->fp64_fallback) ag2::forces_fp32(*c, 0);
        else ag2::forces(*c, 0);
        j->state = AG2_JOB_RUNNING;
        into.push_back(j);
        slots--;
    }
}

// The admission of a round, on a host thread of its own; join() waits for it
// and rethrows its error, and the destructor waits for it on every path.
class Admission {
  public:
    template <class F>
    void start(F &&f) {
        th_ = std::thread([this, f]() {
            try {
                f();
            } catch (...) {
                err_ = std::current_exception();
            }
        });
    }
    void join() {
        if (th_.joinable()) th_.join();
        if (err_) std::rethrow_exception(err_);
    }
    ~Admission() {
        if (th_.joinable()) th_.join();
    }

  private:
    std::thread th_;
    std::exception_ptr err_;
};

// A finished job's final state comes home (on stream st) and its context
// becomes spare.  The job's own work has completed: the round that finished
// it ended by waiting for every stream.
void retire(Pool &P, Job *j, hipStream_t st) {
    j->ctx->stream = st;
    for (int k = 0; k < kFields; k++) j->f[k].assign((size_t)j->n, 0.0);
    j->type.assign((size_t)j->n, 0);
    ag2_particles h = view(*j);
    ag2::ctx_download(*j->ctx, h);
    P.spare.push_back(j->ctx);
    j->ctx = nullptr;
    j->state = AG2_JOB_DONE;
    P.done++;
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    double pos = q * (double)(v.size() - 1);
    size_t lo = (size_t)pos;
    size_t hi = lo + 1 < v.size() ? lo + 1 : lo;
    double f = pos - (double)lo;
    return v[lo] * (1.0 - f) + v[hi] * f;
}

}  // namespace

extern "C" {

int64_t ag2_context_bytes(int64_t capacity, int64_t n_gas) {
    return capacity < 1 ? 0 : ag2::ctx_footprint(capacity, n_gas < 0 ? 0 : n_gas);
}

int ag2_pool_create(const ag2_pool_params *pp, void **out) {
    if (!pp || !out) return AG2_ERR_ARG;
    *out = nullptr;
    if (pp->max_jobs < 1) return AG2_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return AG2_ERR_NODEV;
    if (pp->device < 0 || pp->device >= ndev) return AG2_ERR_NODEV;
    Pool *P = new (std::nothrow) Pool();
    if (!P) return AG2_ERR_CAPACITY;
    P->device = pp->device;
    P->max_jobs = pp->max_jobs;
    try {
        AG2_HIP(hipSetDevice(P->device));
        AG2_HIP(hipStreamCreate(&P->stream));
        AG2_HIP(hipStreamCreate(&P->astream));
        for (int k = 0; k < Pool::kSide; k++) {
            AG2_HIP(hipStreamCreate(&P->side[k]));
            AG2_HIP(hipEventCreateWithFlags(&P->ev_side[k], hipEventDisableTiming));
        }
        AG2_HIP(hipEventCreateWithFlags(&P->ev_walk, hipEventDisableTiming));
        size_t fr = 0, tot = 0;
        AG2_HIP(hipMemGetInfo(&fr, &tot));
        P->budget = pp->hbm_budget > 0 ? pp->hbm_budget : (int64_t)(0.9 * (double)fr);
        size_t jb = ag2::job_walk_bytes() * (size_t)pp->max_jobs;
        AG2_HIP(hipMalloc(&P->jw_dev, jb));
        AG2_HIP(hipHostMalloc(&P->jw_host, jb));
        size_t sb = ag2::sort_job_bytes() * (size_t)pp->max_jobs;
        AG2_HIP(hipMalloc(&P->sj_dev, sb));
        AG2_HIP(hipHostMalloc(&P->sj_host, sb));
        P->issuers.start(Pool::kSide, P->device);
    } catch (...) {   // HIP errors, or no host threads
        P->issuers.stop();
        if (P->jw_dev) (void)hipFree(P->jw_dev);
        if (P->jw_host) (void)hipHostFree(P->jw_host);
        if (P->sj_dev) (void)hipFree(P->sj_dev);
        if (P->sj_host) (void)hipHostFree(P->sj_host);
        for (int k = 0; k < Pool::kSide; k++) {
            if (P->side[k]) (void)hipStreamDestroy(P->side[k]);
            if (P->ev_side[k]) (void)hipEventDestroy(P->ev_side[k]);
        }
        if (P->ev_walk) (void)hipEventDestroy(P->ev_walk);
        if (P->astream) (void)hipStreamDestroy(P->astream);
        if (P->stream) (void)hipStreamDestroy(P->stream);
        delete P;
        return AG2_ERR_HIP;
    }
    *out = P;
    return AG2_OK;
}

int ag2_pool_destroy(void *pool) {
    Pool *P = (Pool *)pool;
    if (!P) return AG2_ERR_ARG;
    (void)hipSetDevice(P->device);
    P->issuers.stop();
    (void)hipStreamSynchronize(P->stream);
    (void)hipStreamSynchronize(P->astream);
    for (auto &kv : P->jobs) {
        if (kv.second->ctx) ag2::ctx_delete(kv.second->ctx);
        delete kv.second;
    }
    for (Ctx *c : P->spare) ag2::ctx_delete(c);
    (void)hipFree(P->jw_dev);
    (void)hipHostFree(P->jw_host);
    (void)hipFree(P->sj_dev);
    (void)hipHostFree(P->sj_host);
    for (int k = 0; k < Pool::kSide; k++) {
        if (P->side[k]) { (void)hipStreamSynchronize(P->side[k]); (void)hipStreamDestroy(P->side[k]); }
        if (P->ev_side[k]) (void)hipEventDestroy(P->ev_side[k]);
    }
    if (P->ev_walk) (void)hipEventDestroy(P->ev_walk);
    (void)hipStreamDestroy(P->astream);
    (void)hipStreamDestroy(P->stream);
    delete P;
    return AG2_OK;
}

const char *ag2_pool_last_error(const void *pool) {
    const Pool *P = (const Pool *)pool;
    return P ? P->err.c_str() : "null pool";
}

int ag2_job_submit(void *pool, const ag2_params *params, const ag2_particles *h, int64_t steps, int64_t *id) {
    Pool *P = (Pool *)pool;
    return pguard(P, [&]() -> int {
        std::string why;
        int rc = ag2::check_params(params, why);
        if (rc) return pfail(P, rc, why);
        if (params->device != P->device) return pfail(P, AG2_ERR_ARG, "job device differs from the pool's");
        if (!h || !h->x || !h->y || !h->z || !h->m) return pfail(P, AG2_ERR_ARG, "x, y, z and m are required");
        if (h->n < 1) return pfail(P, AG2_ERR_ARG, "n must be >= 1");
        if (h->n > 0x7fffff00LL / 2) return pfail(P, AG2_ERR_CAPACITY, "n too large");
        if (steps < 0) return pfail(P, AG2_ERR_ARG, "steps must be >= 0");
        if (ag2::ctx_footprint(h->n, ngas_of(*h)) > P->budget) return pfail(P, AG2_ERR_CAPACITY, "job exceeds the pool's memory budget");
        Job *j = new Job();
        j->id = P->next_id++;
        j->p = *params;
        j->n = h->n;
        j->n_gas = ngas_of(*h);
        j->steps_total = steps;
        for (int k = 0; k < kFields; k++) {
            const double *src = field(*h, k);
            if (src) j->f[k].assign(src, src + h->n);
        }
        if (h->type) j->type.assign(h->type, h->type + h->n);
        else j->type.assign((size_t)h->n, 1);
        P->jobs[j->id] = j;
        P->queue.push_back(j);
        if (id) *id = j->id;
        return AG2_OK;
    });
}

// A job's step interval runs from the end of its previous step (for its first
// step: from the start of the round that admitted it) to the end of the round
// that completes this one, when every stream of the round has drained: the
// round's admissions and downloads (phase 1) and every phase count.
int ag2_pool_step(void *pool, int32_t *n_stepped) {
    Pool *P = (Pool *)pool;
    return pguard(P, [&]() -> int {
        const auto t_start = std::chrono::steady_clock::now();
        const bool idle = P->running.empty();
        if (idle) {
            for (Job *j : P->retiring) retire(*P, j, P->stream);
            P->retiring.clear();
            admit(*P, P->max_jobs, P->stream, P->running);
        }
        for (Job *j : P->running)
            if (!j->timed) { j->t_prev = t_start; j->timed = true; }
        // jobs with no steps left finish at once
        for (size_t i = 0; i < P->running.size();) {
            Job *j = P->running[i];
            if (j->steps_done >= j->steps_total) {
                retire(*P, j, P->stream);
                P->running.erase(P->running.begin() + (ptrdiff_t)i);
            } else {
                i++;
            }
        }
        const int nj = (int)P->running.size();
        if (n_stepped) *n_stepped = nj;
        if (nj == 0) return AG2_OK;
        // While the round runs, a host thread downloads the jobs that finished
        // in the last round (their contexts become spare), then admits queued
        // jobs into the slots this round frees (context, upload, initial
        // forces), all on the admission stream; the admitted jobs step from
        // the next round.  The round's own work touches neither those jobs,
        // the queue nor the spare contexts.
        int finishing = 0;
        for (Job *j : P->running) finishing += j->steps_done + 1 >= j->steps_total ? 1 : 0;
        const int slots = P->max_jobs - (nj - finishing);
        // The thread starts no new admission once the round's work has
        // drained, so the round waits for one admission at most; the rest stay
        // queued for the next rounds.
        std::vector<Job *> admitted, done;
        done.swap(P->retiring);
        std::atomic<bool> stop{false};
        Admission adm;
        if (!done.empty() || (slots > 0 && !P->queue.empty()))
            adm.start([P, slots, &admitted, &done, &stop]() {
                AG2_HIP(hipSetDevice(P->device));
                for (Job *j : done) retire(*P, j, P->astream);
                admit(*P, slots, P->astream, admitted, &stop);
            });
        // admitted jobs ran their initial forces on the main stream
        AG2_HIP(hipEventRecord(P->ev_walk, P->stream));
        for (int k = 0; k < Pool::kSide; k++) AG2_HIP(hipStreamWaitEvent(P->side[k], P->ev_walk, 0));
        const std::vector<Job *> stepped = P->running;
        // per-job phases run on the side streams, job k on stream k mod kSide,
        // each stream's jobs issued by its own host thread
        std::vector<Job *> per[Pool::kSide];
        for (int k = 0; k < nj; k++) {
            Job *j = P->running[(size_t)k];
            j->ctx->stream = P->side[k % Pool::kSide];
            per[k % Pool::kSide].push_back(j);
        }
        P->issuers.run([&](int s) {
            for (Job *j : per[s]) {
                Ctx &c = *j->ctx;
                if (!c.started) {
                    ag2::kdk_init_steps(c);
                    c.started = true;
                }
                ag2::kdk_begin(c);
                ag2::build_keys(c);
            }
        });
        // one batched radix sort of every job's keys on the pool stream
        for (int k = 0; k < Pool::kSide; k++) {
            AG2_HIP(hipEventRecord(P->ev_side[k], P->side[k]));
            AG2_HIP(hipStreamWaitEvent(P->stream, P->ev_side[k], 0));
        }
        {
            std::vector<Ctx *> all;
            for (Job *j : P->running) all.push_back(j->ctx);
            ag2::sort_batch(all.data(), (int)all.size(), P->stream, P->sj_dev, P->sj_host);
        }
        AG2_HIP(hipEventRecord(P->ev_walk, P->stream));
        for (int k = 0; k < Pool::kSide; k++) AG2_HIP(hipStreamWaitEvent(P->side[k], P->ev_walk, 0));
        P->issuers.run([&](int s) {
            for (Job *j : per[s]) {
                Ctx &c = *j->ctx;
                // density and forces follow; with one step size for all,
                // this round's walk writes every acceleration
                c.density_follows = true;
                c.acc_overwritten = c.p.min_dt >= c.p.max_dt;
                ag2::build_order(c);
            }
            for (Job *j : per[s]) {
                Ctx &c = *j->ctx;
                AG2_HIP(hipEventSynchronize(c.ev_counts));
                ag2::build_tree_b(c);
                c.density_follows = c.acc_overwritten = false;
                ag2::gas_density(c);
                ag2::sph_prepare(c);
            }
        });
        std::vector<Ctx *> f32, f64;
        for (Job *j : P->running) {
            Ctx *c = j->ctx;
            c->fp64_fallback = c->p.precision == AG2_FP32_SPLIT && !ag2::fp32_usable(*c);
            (c->p.precision == AG2_FP32_SPLIT && !c->fp64_fallback ? f32 : f64).push_back(c);
        }
        // the two forms use separate halves of the job-record buffers
        // the walk waits for every job's tree and density, the second kicks
        // for the walk
        for (int k = 0; k < Pool::kSide; k++) {
            AG2_HIP(hipEventRecord(P->ev_side[k], P->side[k]));
            AG2_HIP(hipStreamWaitEvent(P->stream, P->ev_side[k], 0));
        }
        char *hd = (char *)P->jw_dev, *hh = (char *)P->jw_host;
        size_t off = ag2::job_walk_bytes() * f32.size();
        ag2::forces_batch(f32.data(), (int)f32.size(), P->stream, 1, hd, hh);
        ag2::forces_batch(f64.data(), (int)f64.size(), P->stream, 0, hd + off, hh + off);
        AG2_HIP(hipEventRecord(P->ev_walk, P->stream));
        for (int k = 0; k < Pool::kSide; k++) AG2_HIP(hipStreamWaitEvent(P->side[k], P->ev_walk, 0));
        int64_t ps = 0;
        P->issuers.run([&](int s) {
            for (Job *j : per[s]) {
                Ctx &c = *j->ctx;
                if (c.p.precision == AG2_FP32_SPLIT && !c.fp64_fallback) ag2::near_finish(c);
                ag2::sph_finish(*j->ctx);
                ag2::fold_counters(*j->ctx);
                ag2::kdk_end(*j->ctx);
            }
        });
        for (Job *j : P->running) {
            j->ctx->steps++;
            j->steps_done++;
            ps += j->n;
        }
        // admissions go on while the device finishes the round
        for (int k = 0; k < Pool::kSide; k++) AG2_HIP(hipStreamSynchronize(P->side[k]));
        stop = true;
        adm.join();
        for (Job *j : admitted)
            if (!j->timed) { j->t_prev = t_start; j->timed = true; }
        AG2_HIP(hipStreamSynchronize(P->astream));
        for (Job *j : P->running) j->ctx->stream = P->stream;
        for (size_t i = 0; i < P->running.size();) {
            Job *j = P->running[i];
            if (j->steps_done >= j->steps_total) {
                P->retiring.push_back(j);
                P->running.erase(P->running.begin() + (ptrdiff_t)i);
            } else {
                i++;
            }
        }
        for (Job *j : admitted) {
            j->ctx->stream = P->stream;
            P->running.push_back(j);
        }
        const auto t_end = std::chrono::steady_clock::now();
        P->rounds++;
        P->job_steps += nj;
        P->particle_steps += ps;
        P->busy_s += std::chrono::duration<double>(t_end - t_start).count();
        P->resident_sum += nj;
        P->resident_max = nj > P->resident_max ? nj : P->resident_max;
        for (Job *j : stepped) {
            P->step_ms.push_back(std::chrono::duration<double, std::milli>(t_end - j->t_prev).count());
            j->t_prev = t_end;
        }
        return AG2_OK;
    });
}

int ag2_job_status(void *pool, int64_t id, int32_t *state, int64_t *steps_done) {
    Pool *P = (Pool *)pool;
    return pguard(P, [&]() -> int {
        auto it = P->jobs.find(id);
        if (it == P->jobs.end()) return pfail(P, AG2_ERR_ARG, "unknown job");
        if (state) *state = it->second->state;
        if (steps_done) *steps_done = it->second->steps_done;
        return AG2_OK;
    });
}

int ag2_job_result(void *pool, int64_t id, ag2_particles *out) {
    Pool *P = (Pool *)pool;
    return pguard(P, [&]() -> int {
        auto it = P->jobs.find(id);
        if (it == P->jobs.end()) return pfail(P, AG2_ERR_ARG, "unknown job");
        Job *j = it->second;
        if (j->state != AG2_JOB_DONE) return pfail(P, AG2_ERR_STATE, "job not finished");
        if (!out || out->n != j->n) return pfail(P, AG2_ERR_ARG, "n differs from the job's");
        for (int k = 0; k < kFields; k++) {
            double *dst = field(*out, k);
            if (dst) std::copy(j->f[k].begin(), j->f[k].end(), dst);
        }
        if (out->type) std::copy(j->type.begin(), j->type.end(), out->type);
        P->jobs.erase(it);
        delete j;
        return AG2_OK;
    });
}

int ag2_pool_get_stats(void *pool, ag2_pool_stats *out) {
    Pool *P = (Pool *)pool;
    return pguard(P, [&]() -> int {
        if (!out) return pfail(P, AG2_ERR_ARG, "out is NULL");
        out->rounds = P->rounds;
        out->job_steps = P->job_steps;
        out->particle_steps = P->particle_steps;
        out->queued = (int64_t)P->queue.size();
        out->running = (int64_t)P->running.size();
        out->done = P->done;
        out->p50_step_ms = percentile(P->step_ms, 0.50);
        out->p99_step_ms = percentile(P->step_ms, 0.99);
        out->max_step_ms = P->step_ms.empty() ? 0.0 : *std::max_element(P->step_ms.begin(), P->step_ms.end());
        out->busy_s = P->busy_s;
        out->mean_resident = P->rounds > 0 ? (double)P->resident_sum / (double)P->rounds : 0.0;
        out->max_resident = P->resident_max;
        out->hbm_used = P->used;
        out->hbm_budget = P->budget;
        return AG2_OK;
    });
}

}  // extern "C"
