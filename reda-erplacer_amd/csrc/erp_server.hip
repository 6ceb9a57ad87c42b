// Serving form: a persistent request loop that runs many placements at once on
// one GPU with continuous batching.
//
// Requests (designs) are submitted from any thread and queue up.  Admission
// threads take queued jobs in arrival order while the job slots and the HBM
// budget allow (a job that does not fit waits, later smaller ones may pass it),
// create their contexts (state stays resident in device memory) and run each
// new job's first `check_every` iterations on the job's own stream, while the
// batch worker keeps the running batches going on their streams; the worker
// then adds the prepared job to the batch of its bin grid.  Each batch is one
// hipGraph that holds `check_every` iterations of a fixed number of slots,
// every kernel launched once for all slots (job = blockIdx.y; an empty slot
// returns at once).  After each round the worker reads every job's scalars; a
// job whose overflow fell below its stop threshold, that reached its iteration
// limit, or whose scalars became non-finite is retired: its positions and exact
// HPWL are copied out, its device memory is freed and queued jobs are admitted
// into the freed room, taking the emptied slots.
//
// A job's kernels only read and write its own buffers and reduce in a fixed
// order, so its result is bit-identical alone (erp_run_until with the same
// check interval) and in any mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "erp_ctx.h"

namespace {

using Clock = std::chrono::steady_clock;

// Host copy of a submitted design (the caller's arrays may go away).
struct HostDesign {
    std::vector<float> x, y, w, h, npins, pox, poy;
    std::vector<int32_t> pin_node, net_start, net_pins, node_pin_start, node_pins;
    erp_design d{};
    void copy_from(const erp_design *s) {
        auto cp = [](auto &v, const auto *p, int64_t n) { v.assign(p, p + n); };
        cp(x, s->x, s->num_nodes);
        cp(y, s->y, s->num_nodes);
        cp(w, s->w, s->num_nodes);
        cp(h, s->h, s->num_nodes);
        cp(npins, s->num_pins, s->num_nodes);
        cp(pox, s->pin_offset_x, s->num_pins_total);
        cp(poy, s->pin_offset_y, s->num_pins_total);
        cp(pin_node, s->pin_node, s->num_pins_total);
        cp(net_start, s->net_start, s->num_nets + 1);
        cp(net_pins, s->net_pins, s->num_nets > 0 ? s->net_start[s->num_nets] : 0);
        cp(node_pin_start, s->node_pin_start, s->num_nodes + 1);
        cp(node_pins, s->node_pins, s->node_pin_start[s->num_nodes]);
        d = *s;
        d.x = x.data(); d.y = y.data(); d.w = w.data(); d.h = h.data(); d.num_pins = npins.data();
        d.pin_offset_x = pox.data(); d.pin_offset_y = poy.data(); d.pin_node = pin_node.data();
        d.net_start = net_start.data(); d.net_pins = net_pins.data();
        d.node_pin_start = node_pin_start.data(); d.node_pins = node_pins.data();
    }
};

struct Job {
    int64_t id = 0;
    HostDesign design;
    erp_config cfg{};
    int64_t max_iterations = 0;
    double stop_overflow = 0.0;
    size_t bytes = 0;
    erp_ctx *ctx = nullptr;
    int64_t iterations = 0;
    erp_job_result result{};
    std::vector<float> px, py;      // final positions
    bool done = false;
};

// The jobs of one bin grid.  Its graph is captured for a fixed number of slots
// (a power of two) and a launch geometry; a slot holds a job or is empty (its
// table entry inactive: every kernel returns at once for it).  A job that
// joins takes an empty slot and one that leaves empties its slot: only the
// table is uploaded again.  The graph is captured again when no slot is free,
// when a job needs a wider launch geometry, or when a quarter or less of the
// slots are used.
struct Group {
    int m = 0, n = 0;
    std::vector<Job *> slots;       // the graph's slots: a job or nullptr
    std::vector<Job *> joining;     // jobs the next capture places into slots
    std::vector<JobDesc> tab;       // host copy of the slot table
    int members = 0;                // jobs in slots or joining
    bool recapture = true;
    bool table_dirty = false;
    DevBuf table;                   // JobDesc[slots]
    BatchDims dims;                 // launch geometry of the graph
    hipGraphExec_t graph = nullptr;
    bool ready = false;             // eager mode: slots and table set, launched each round
    DevBuf state_dev;               // (hpwl, overflow, nonfinite) per member, written by the graph
    double *state_host = nullptr;   // pinned copy, read after each round
    int capacity = 0;
    // the group's own stream: batches of different grid sizes are small and run
    // side by side (non-blocking, default priority)
    hipStream_t stream = nullptr;
    Group() = default;
    Group(const Group &) = delete;
    Group &operator=(const Group &) = delete;
    ~Group() {
        drop_graph();
        if (stream) (void)hipStreamDestroy(stream);
        if (state_host) (void)hipHostFree(state_host);
    }
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        ready = false;
    }
};

}  // namespace

struct erp_server {
    erp_server_config cfg{};
    hipStream_t stream = nullptr;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::condition_variable cv_admit;                   // wakes the admission thread
    std::deque<std::unique_ptr<Job>> pending;
    std::map<int64_t, std::unique_ptr<Job>> preparing;  // context being created / first chunk
    std::vector<std::unique_ptr<Job>> ready;            // prepared, not yet in a batch
    std::map<int64_t, std::unique_ptr<Job>> active;     // in a batch, not finished
    std::map<int64_t, std::unique_ptr<Job>> finished;   // results not yet released
    std::vector<int64_t> unpolled;
    std::map<std::pair<int, int>, Group> groups;
    int64_t next_id = 1;
    size_t bytes_in_use = 0;
    int64_t rounds = 0, batched_launches = 0, max_batch = 0;
    double admit_s = 0, round_s = 0, retire_s = 0;
    bool stop = false;
    int worker_rc = ERP_OK;
    std::string worker_error;
    std::thread worker;
    // context creation is mostly host work (net ordering, CSR, uploads): several
    // jobs are prepared at once
    static constexpr int kAdmitters = 8;
    std::vector<std::thread> admitters;
    // jobs holding a slot (admitted and not finished)
    size_t resident() const { return preparing.size() + ready.size() + active.size(); }
    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void fail_job(erp_server *s, Job *j, int rc) {
    j->result.status = rc;
    j->result.t_done = s->now();
    j->done = true;
}

// Retire a finished job: copy results out and free its device memory.  Returns
// the seconds it took (the caller adds them to the stats under the lock).
double retire(erp_server *s, Job *j, const erp_step_info &info) {
    const double t0 = s->now();
    erp_ctx *c = j->ctx;
    j->result.iterations = j->iterations;
    j->result.hpwl = info.hpwl;
    j->result.overflow = info.overflow;
    j->result.converged = info.overflow < j->stop_overflow;
    j->px.resize((size_t)c->N);
    j->py.resize((size_t)c->N);
    int rc = erp_get_positions(c, j->px.data(), j->py.data());
    int64_t dbu = 0;
    if (!rc) rc = erp_final_hpwl_dbu(c, &dbu);
    j->result.hpwl_dbu = dbu;
    j->result.status = rc;
    j->result.t_done = s->now();
    erp_ctx_destroy(c);
    j->ctx = nullptr;
    j->done = true;
    return s->now() - t0;
}

bool finished_now(const Job *j, const erp_step_info &info) {
    return info.nonfinite || info.overflow < j->stop_overflow || j->iterations >= j->max_iterations;
}

// Admission thread: take a queued job that fits, prepare it outside the lock
// (context + first chunk on the job's own stream), hand it to the worker.
void admitter_main(erp_server *s) {
    const bool dev_ok = hipSetDevice(s->cfg.device) == hipSuccess;
    for (;;) {
        std::unique_ptr<Job> job;
        Job *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_admit.wait(lk, [&] {
                if (s->stop || s->worker_rc != ERP_OK) return true;
                if ((int)s->resident() >= s->cfg.max_active) return false;
                return std::any_of(s->pending.begin(), s->pending.end(), [&](const std::unique_ptr<Job> &q) {
                    return s->bytes_in_use + q->bytes <= (size_t)s->cfg.hbm_budget_bytes;
                });
            });
            if (s->stop || s->worker_rc != ERP_OK) return;
            auto it = std::find_if(s->pending.begin(), s->pending.end(), [&](const std::unique_ptr<Job> &q) {
                return s->bytes_in_use + q->bytes <= (size_t)s->cfg.hbm_budget_bytes;
            });
            j = it->get();
            s->bytes_in_use += j->bytes;
            s->preparing[j->id] = std::move(*it);
            s->pending.erase(it);
        }
        const double ta = s->now();
        j->result.t_admit = ta;
        int rc = dev_ok ? erp_ctx_create(&j->design.d, &j->cfg, &j->ctx)
                        : fail(ERP_ERR_HIP, "hipSetDevice failed in the admission thread");
        erp_step_info info{};
        if (!rc) {
            erp_set_store_gradients(j->ctx, 0);
            // a job's result holds no energy: the potential map is not computed
            // (the trajectory is unchanged, placement.rs:185 discards it too)
            erp_set_compute_potential(j->ctx, 0);
            // the first chunk runs alone with eager launches (no per-job graph
            // capture; results are bit-identical to the graph-replayed loop)
            const int64_t first = std::min<int64_t>(s->cfg.check_every, j->max_iterations);
            rc = run_eager(j->ctx, first, &info);
            j->iterations = first;
            const int64_t left = j->max_iterations - j->iterations;
            if (!rc && !finished_now(j, info) && left < s->cfg.check_every) {
                rc = run_eager(j->ctx, left, &info);
                j->iterations += left;
            }
        }
        double dt_retire = 0.0;
        if (rc) {
            if (j->ctx) erp_ctx_destroy(j->ctx);
            j->ctx = nullptr;
            fail_job(s, j, rc);
        } else if (finished_now(j, info)) {
            dt_retire = retire(s, j, info);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        job = std::move(s->preparing[j->id]);
        s->preparing.erase(j->id);
        s->admit_s += s->now() - ta;
        s->retire_s += dt_retire;
        if (!j->done && s->worker_rc != ERP_OK) {    // the worker stopped meanwhile
            erp_ctx_destroy(j->ctx);
            j->ctx = nullptr;
            fail_job(s, j, s->worker_rc);
        }
        if (j->done) {
            s->bytes_in_use -= j->bytes;
            s->unpolled.push_back(j->id);
            s->finished[j->id] = std::move(job);
            s->cv_done.notify_all();
            s->cv_admit.notify_all();
        } else {
            s->ready.push_back(std::move(job));
            s->cv.notify_all();
        }
    }
}

// b fits inside the launch geometry a
bool covers(const BatchDims &a, const BatchDims &b) {
    if (b.max_cells > a.max_cells || b.max_n2 > a.max_n2 || b.max_reach > a.max_reach) return false;
    for (int k = 0; k < kWaBuckets; ++k)
        if (b.max_bucket[k] > a.max_bucket[k]) return false;
    return true;
}

void join(Group &g, Job *j) {
    ++g.members;
    BatchDims one;
    one.add(j->ctx->job);
    if (!g.recapture && (g.graph || g.ready) && covers(g.dims, one)) {
        auto it = std::find(g.slots.begin(), g.slots.end(), nullptr);
        if (it != g.slots.end()) {
            *it = j;
            g.tab[(size_t)(it - g.slots.begin())] = j->ctx->job;
            g.table_dirty = true;
            return;
        }
    }
    g.joining.push_back(j);
    g.recapture = true;
}

// Removes j if it is a member; returns whether it was.
bool leave(Group &g, Job *j) {
    auto it = std::find(g.slots.begin(), g.slots.end(), j);
    if (it != g.slots.end()) {
        *it = nullptr;
        g.tab[(size_t)(it - g.slots.begin())].active = 0;
        g.table_dirty = true;
    } else {
        auto jt = std::find(g.joining.begin(), g.joining.end(), j);
        if (jt == g.joining.end()) return false;
        g.joining.erase(jt);
    }
    --g.members;
    if (g.members > 0 && 4 * g.members <= (int)g.slots.size()) g.recapture = true;
    return true;
}

int upload_table(erp_server *s, Group &g) {
    g.table_dirty = false;
    ERP_HIP_TRY(hipMemcpyAsync(g.table.p, g.tab.data(), sizeof(JobDesc) * g.tab.size(), hipMemcpyHostToDevice,
                               s->stream));
    ERP_HIP_TRY(hipStreamSynchronize(s->stream));
    return ERP_OK;
}

// Captures the group's graph again: the members are packed into the first
// slots of a power-of-two slot count, the launch geometry is theirs.
int rebuild(erp_server *s, Group &g) {
    g.drop_graph();
    g.recapture = false;
    std::vector<Job *> live;
    for (Job *j : g.slots)
        if (j) live.push_back(j);
    live.insert(live.end(), g.joining.begin(), g.joining.end());
    g.joining.clear();
    g.slots.clear();
    g.tab.clear();
    if (live.empty()) return ERP_OK;
    int nj = 1;
    while (nj < (int)live.size()) nj <<= 1;
    g.slots.assign((size_t)nj, nullptr);
    g.tab.assign((size_t)nj, JobDesc{});
    g.dims = BatchDims();
    for (size_t k = 0; k < live.size(); ++k) {
        g.slots[k] = live[k];
        g.tab[k] = live[k]->ctx->job;
        g.dims.add(live[k]->ctx->job);
    }
    for (size_t k = live.size(); k < (size_t)nj; ++k) g.tab[k].active = 0;
    int rc = ERP_OK;
    if (g.table.bytes < sizeof(JobDesc) * (size_t)nj &&
        (rc = g.table.alloc(sizeof(JobDesc) * (size_t)std::max(nj, s->cfg.max_active))))
        return rc;
    if (g.capacity < nj) {
        const int cap = std::max(nj, s->cfg.max_active);
        if ((rc = g.state_dev.alloc(sizeof(double) * 3 * (size_t)cap))) return rc;
        if (g.state_host) (void)hipHostFree(g.state_host);
        g.state_host = nullptr;
        ERP_HIP_TRY(hipHostMalloc((void **)&g.state_host, sizeof(double) * 3 * (size_t)cap, hipHostMallocDefault));
        g.capacity = cap;
    }
    if ((rc = upload_table(s, g))) return rc;
    if (s->cfg.eager) {             // launched directly by each round
        g.ready = true;
        return ERP_OK;
    }
    ERP_HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < s->cfg.check_every && !rc; ++k)
        rc = launch_batched_iteration(g.table.as<JobDesc>(), g.dims, nj, s->stream);
    if (!rc) rc = launch_gather_state(g.table.as<JobDesc>(), nj, g.state_dev.as<double>(), s->stream);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &graph);
    if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return fail(ERP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&g.graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(ERP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return ERP_OK;
}

// One round: every batch advances check_every iterations, then every active
// job is checked.
int round(erp_server *s) {
    std::vector<Job *> running;
    std::vector<const double *> state;      // (hpwl, overflow, nonfinite) of running[k]
    // captures and table uploads first, so that no batch is running while
    // another group's buffers are reallocated or its table is written
    for (auto &kv : s->groups) {
        Group &g = kv.second;
        int rc = ERP_OK;
        if (g.recapture) rc = rebuild(s, g);
        else if (g.table_dirty) rc = upload_table(s, g);
        if (rc) return rc;
    }
    std::vector<hipStream_t> used;
    for (auto &kv : s->groups) {
        Group &g = kv.second;
        if (!g.graph && !g.ready) continue;
        if (!g.stream) ERP_HIP_TRY(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
        if (g.graph) {
            ERP_HIP_TRY(hipGraphLaunch(g.graph, g.stream));
        } else {                    // eager: the kernels the graph would hold, in its order
            const int nj = (int)g.slots.size();
            int rc = ERP_OK;
            for (int k = 0; k < s->cfg.check_every && !rc; ++k)
                rc = launch_batched_iteration(g.table.as<JobDesc>(), g.dims, nj, g.stream);
            if (!rc) rc = launch_gather_state(g.table.as<JobDesc>(), nj, g.state_dev.as<double>(), g.stream);
            if (rc) return rc;
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            ++s->batched_launches;
            s->max_batch = std::max<int64_t>(s->max_batch, g.members);
        }
        ERP_HIP_TRY(hipMemcpyAsync(g.state_host, g.state_dev.p, sizeof(double) * 3 * g.slots.size(),
                                   hipMemcpyDeviceToHost, g.stream));
        used.push_back(g.stream);
        for (size_t k = 0; k < g.slots.size(); ++k) {
            if (!g.slots[k]) continue;
            running.push_back(g.slots[k]);
            state.push_back(g.state_host + 3 * k);
        }
    }
    if (running.empty()) return ERP_OK;
    for (hipStream_t st : used) ERP_HIP_TRY(hipStreamSynchronize(st));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        ++s->rounds;
    }
    for (size_t k = 0; k < running.size(); ++k) {
        Job *j = running[k];
        j->iterations += s->cfg.check_every;
        j->ctx->iter_host += s->cfg.check_every;
        erp_step_info info{};
        info.hpwl = state[k][0];
        info.overflow = state[k][1];
        info.nonfinite = (int32_t)state[k][2];
        const int64_t left = j->max_iterations - j->iterations;
        if (!finished_now(j, info) && left < s->cfg.check_every) {
            // the last, shorter chunk runs alone (as erp_run_until would)
            int rc = run_eager(j->ctx, left, &info);
            if (rc) return rc;
            j->iterations += left;
        }
        if (!finished_now(j, info)) continue;
        const double dt = retire(s, j, info);
        std::lock_guard<std::mutex> lk(s->mu);
        s->retire_s += dt;
        for (auto &kv : s->groups)
            if (leave(kv.second, j)) break;
        s->bytes_in_use -= j->bytes;
        s->unpolled.push_back(j->id);
        s->finished[j->id] = std::move(s->active[j->id]);
        s->active.erase(j->id);
        s->cv_done.notify_all();
        s->cv_admit.notify_all();           // a slot and memory are free
    }
    return ERP_OK;
}

// Fail every job that has not finished (after a HIP error in the worker).
void fail_all(erp_server *s, int rc) {
    auto finish = [&](std::unique_ptr<Job> &j) {
        if (j->ctx) erp_ctx_destroy(j->ctx);
        j->ctx = nullptr;
        fail_job(s, j.get(), rc);
        s->bytes_in_use -= j->bytes;
        s->unpolled.push_back(j->id);
        const int64_t id = j->id;
        s->finished[id] = std::move(j);
    };
    for (auto &kv : s->active) finish(kv.second);
    s->active.clear();
    for (auto &j : s->ready) finish(j);
    s->ready.clear();
    for (auto &j : s->pending) {
        s->bytes_in_use += j->bytes;       // finish() subtracts it
        finish(j);
    }
    s->pending.clear();
    for (auto &kv : s->groups) kv.second.drop_graph();
    s->groups.clear();
}

void worker_main(erp_server *s) {
    if (hipSetDevice(s->cfg.device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->worker_rc = ERP_ERR_HIP;
        s->worker_error = "hipSetDevice failed in the serving worker";
        s->cv_done.notify_all();
        s->cv_admit.notify_all();
        return;
    }
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->ready.empty() || !s->active.empty(); });
            if (s->stop) break;
            for (auto &job : s->ready) {              // prepared jobs join their batches
                Job *j = job.get();
                Group &g = s->groups[{j->ctx->m, j->ctx->n}];
                g.m = j->ctx->m;
                g.n = j->ctx->n;
                join(g, j);
                s->active[j->id] = std::move(job);
            }
            s->ready.clear();
        }
        const double tb = s->now();
        int rc = round(s);
        if (rc) {   // batches already launched finish before their contexts are freed
            for (auto &kv : s->groups)
                if (kv.second.stream) (void)hipStreamSynchronize(kv.second.stream);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        s->round_s += s->now() - tb;
        if (rc) {
            s->worker_rc = rc;
            s->worker_error = erp_last_error();
            fail_all(s, rc);                          // the server stops serving
            s->cv_done.notify_all();
            s->cv_admit.notify_all();
            break;
        }
        for (auto it = s->groups.begin(); it != s->groups.end();) {
            if (it->second.members == 0) {
                it->second.drop_graph();
                it = s->groups.erase(it);
            } else {
                ++it;
            }
        }
        if (s->stop) break;
    }
}

}  // namespace

extern "C" {

void erp_server_config_default(erp_server_config *cfg) {
    if (!cfg) return;
    cfg->hbm_budget_bytes = (int64_t)32 << 30;
    cfg->max_active = 64;
    cfg->check_every = 10;
    cfg->device = 0;
    cfg->eager = 0;
}

int erp_server_create(const erp_server_config *cfg, erp_server **out) {
    if (!cfg || !out) return fail(ERP_ERR_INVALID, "null argument");
    if (cfg->max_active <= 0 || cfg->check_every <= 0 || cfg->hbm_budget_bytes <= 0)
        return fail(ERP_ERR_INVALID, "bad server config");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(ERP_ERR_NO_DEVICE, "no HIP device");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(ERP_ERR_INVALID, "bad device ordinal");
    ERP_HIP_TRY(hipSetDevice(cfg->device));
    erp_server *s = new erp_server();
    s->cfg = *cfg;
    s->t0 = Clock::now();
    if (hipStreamCreate(&s->stream) != hipSuccess) {
        delete s;
        return fail(ERP_ERR_HIP, "hipStreamCreate failed");
    }
    s->worker = std::thread(worker_main, s);
    for (int k = 0; k < erp_server::kAdmitters; ++k) s->admitters.emplace_back(admitter_main, s);
    *out = s;
    return ERP_OK;
}

int erp_server_submit(erp_server *s, const erp_design *d, const erp_config *cfg,
                      int64_t max_iterations, double stop_overflow, int64_t *job_id) {
    if (!s || !d || !cfg || !job_id) return fail(ERP_ERR_INVALID, "null argument");
    if (max_iterations <= 0) return fail(ERP_ERR_INVALID, "max_iterations must be positive");
    if (cfg->spectral_mode != ERP_SPECTRAL_FFT)
        return fail(ERP_ERR_UNSUPPORTED, "the serving loop batches the FFT spectral form");
    auto job = std::make_unique<Job>();
    job->design.copy_from(d);
    job->cfg = *cfg;
    job->cfg.device = s->cfg.device;
    job->max_iterations = max_iterations;
    job->stop_overflow = stop_overflow;
    job->bytes = estimate_ctx_bytes(d, cfg);
    if (job->bytes > (size_t)s->cfg.hbm_budget_bytes)
        return fail(ERP_ERR_INVALID, "design needs more device memory than the server budget");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop || s->worker_rc) return fail(ERP_ERR_INVALID, "server is stopping or failed");
    job->id = s->next_id++;
    job->result.job_id = job->id;
    job->result.t_submit = s->now();
    *job_id = job->id;
    s->pending.push_back(std::move(job));
    s->cv_admit.notify_all();
    return ERP_OK;
}

int erp_server_poll(erp_server *s, erp_job_result *out, int32_t capacity, int32_t *count) {
    if (!s || !count || (capacity > 0 && !out)) return fail(ERP_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    int32_t k = 0;
    while (k < capacity && !s->unpolled.empty()) {
        const int64_t id = s->unpolled.front();
        s->unpolled.erase(s->unpolled.begin());
        out[k++] = s->finished[id]->result;
    }
    *count = k;
    return ERP_OK;
}

int erp_server_wait(erp_server *s, int64_t job_id, erp_job_result *out) {
    if (!s) return fail(ERP_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    const bool known = s->finished.count(job_id) || s->active.count(job_id) || s->preparing.count(job_id) ||
                       std::any_of(s->ready.begin(), s->ready.end(),
                                   [&](const std::unique_ptr<Job> &j) { return j->id == job_id; }) ||
                       std::any_of(s->pending.begin(), s->pending.end(),
                                   [&](const std::unique_ptr<Job> &j) { return j->id == job_id; });
    if (!known) return fail(ERP_ERR_INVALID, "unknown or released job id");
    s->cv_done.wait(lk, [&] {
        auto it = s->finished.find(job_id);
        return (it != s->finished.end() && it->second->done) || s->worker_rc != ERP_OK;
    });
    auto it = s->finished.find(job_id);
    if (it == s->finished.end()) return fail(s->worker_rc, s->worker_error);
    if (out) *out = it->second->result;
    return ERP_OK;
}

int erp_server_drain(erp_server *s) {
    if (!s) return fail(ERP_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        return (s->pending.empty() && s->resident() == 0) || s->worker_rc != ERP_OK;
    });
    if (s->worker_rc) return fail(s->worker_rc, s->worker_error);
    return ERP_OK;
}

int erp_server_get_positions(erp_server *s, int64_t job_id, float *x, float *y) {
    if (!s || !x || !y) return fail(ERP_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->finished.find(job_id);
    if (it == s->finished.end() || !it->second->done) return fail(ERP_ERR_INVALID, "job not finished");
    const Job &j = *it->second;
    std::copy(j.px.begin(), j.px.end(), x);
    std::copy(j.py.begin(), j.py.end(), y);
    return ERP_OK;
}

int erp_server_release(erp_server *s, int64_t job_id) {
    if (!s) return fail(ERP_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->finished.erase(job_id);
    return ERP_OK;
}

int erp_server_stats(erp_server *s, erp_server_stat *st) {
    if (!s || !st) return fail(ERP_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    st->pending = (int64_t)s->pending.size();
    st->active = (int64_t)s->resident();
    st->finished = (int64_t)s->finished.size();
    st->bytes_in_use = (int64_t)s->bytes_in_use;
    st->rounds = s->rounds;
    st->batched_launches = s->batched_launches;
    st->max_batch = s->max_batch;
    st->clock = s->now();
    st->admit_seconds = s->admit_s;
    st->round_seconds = s->round_s;
    st->retire_seconds = s->retire_s;
    return ERP_OK;
}

int erp_server_destroy(erp_server *s) {
    if (!s) return ERP_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->pending.clear();          // queued jobs are dropped
        s->cv.notify_all();
        s->cv_admit.notify_all();
    }
    for (auto &t : s->admitters)      // each finishes the job it is preparing
        if (t.joinable()) t.join();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->cfg.device);
    for (auto &kv : s->active)
        if (kv.second->ctx) erp_ctx_destroy(kv.second->ctx);
    for (auto &j : s->ready)
        if (j->ctx) erp_ctx_destroy(j->ctx);
    for (auto &kv : s->groups) kv.second.drop_graph();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return ERP_OK;
}

}  // extern "C"
