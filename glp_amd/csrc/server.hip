// Serving: a request loop with continuous batching under an HBM budget.
//
// Jobs (a graph, classic or layered LP, a step count) are queued by
// glp_server_submit and admitted in arrival order as soon as their device
// footprint fits the budget and a slot is free; admission uploads the CSR and
// creates the job's label state, all of it device-resident until the job
// ends.  One worker thread owns the device: every tick it advances every
// running job by one step with a single launch per tier, the jobs folded
// together through a device-side job table (k_multi), so requests that
// arrive while others run join the next tick and share its launches.  A job
// that completes copies its labels out and returns its memory at once,
// letting the next queued job in.  Classic jobs run the layered hooks with
// gamma = 0, which computes classic LP, so one kernel set serves both kinds.
// Results are independent of what else runs: each job's arithmetic touches
// only its own arrays.

#include "device.h"
#include "session.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace glp {

struct ServerJob {
    int64_t id = 0;
    std::unique_ptr<glp_graph> g;  // private copy of the caller's graph
    glp_params p{};
    Plan plan;
    int32_t steps = 0;
    size_t bytes = 0;
    int32_t state = 0;  // 0 queued, 1 running, 2 done, 3 cancelled, <0 error code
    bool cancel = false;
    double submit_ms = 0, admit_ms = 0, done_ms = 0;
    // device state while running
    DeviceGraph* dg = nullptr;
    int32_t* L[2] = {nullptr, nullptr};
    int32_t* V[2] = {nullptr, nullptr};
    HubScratch hs;
    unsigned long long* changes = nullptr;
    int cur = 0;
    int32_t done_steps = 0;
    std::vector<int32_t> result;
    void release() {
        delete dg;
        dg = nullptr;
        for (int i = 0; i < 2; ++i) {
            (void)hipFree(L[i]);
            (void)hipFree(V[i]);
            L[i] = V[i] = nullptr;
        }
        (void)hipFree(hs.pairs);
        (void)hipFree(hs.sorted);
        (void)hipFree(hs.hist);
        (void)hipFree(hs.total);
        (void)hipFree(hs.chunk_cnt);
        (void)hipFree(hs.best);
        (void)hipFree(changes);
        hs = HubScratch();
        changes = nullptr;
    }
    ~ServerJob() { release(); }
};

}  // namespace glp

using namespace glp;

struct glp_server {
    uint64_t budget = 0;
    int32_t max_slots = 0;
    Tuning tune;
    int32_t class_cap[kMaxBlockClasses] = {};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServerJob*> queue;
    std::vector<ServerJob*> running;
    std::map<int64_t, std::unique_ptr<ServerJob>> jobs;
    uint64_t used = 0;
    int64_t next_id = 1;
    bool stop = false;
    int worker_error = 0;
    std::string worker_msg;
    std::thread worker;
    hipStream_t stream = nullptr;
    JobDesc* d_jobs = nullptr;
    int32_t* d_cum = nullptr;
    JobDesc* h_jobs = nullptr;  // pinned staging
    int32_t* h_cum = nullptr;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void run();
    int admit(ServerJob* j);
    int tick();
};

static size_t job_bytes(const HostCSR& g, const Plan& plan, const glp_params& p) {
    size_t b = device_graph_bytes(g, plan);
    const size_t n = std::max<size_t>((size_t)g.n, 1);
    b += 2 * n * 4;                                  // labels
    if (p.method == GLP_LAYERED) b += 2 * n * 4;     // volumes
    b += (size_t)plan.pair_slots * 16 + (size_t)plan.hist_slots * 4 + plan.hub_chunks.size() * 4 +
         plan.hubs.size() * 12;
    b += 8 * kChangeSlots;
    return b;
}

int glp_server::admit(ServerJob* j) {
    const HostCSR& g = j->g->csr;
    int rc = device_graph_upload(g, std::move(j->plan), stream, &j->dg);
    if (rc != GLP_OK) return rc;
    const Plan& p = j->dg->plan;
    const size_t n = std::max<size_t>((size_t)g.n, 1);
    GLP_HIP(hipMalloc(&j->L[0], n * 4));
    GLP_HIP(hipMalloc(&j->L[1], n * 4));
    if ((rc = launch_iota(j->L[0], g.n, stream))) return rc;
    if ((rc = launch_iota(j->L[1], g.n, stream))) return rc;
    if (j->p.method == GLP_LAYERED) {
        GLP_HIP(hipMalloc(&j->V[0], n * 4));
        GLP_HIP(hipMalloc(&j->V[1], n * 4));
        if ((rc = launch_fill_i32(j->V[0], 1, g.n, stream))) return rc;  // V_0 = 1
    }
    if (!p.hubs.empty()) {
        GLP_HIP(hipMalloc(&j->hs.pairs, (size_t)p.pair_slots * 8));
        GLP_HIP(hipMalloc(&j->hs.sorted, (size_t)p.pair_slots * 8));
        GLP_HIP(hipMalloc(&j->hs.hist, (size_t)p.hist_slots * 4));
        GLP_HIP(hipMalloc(&j->hs.total, p.hubs.size() * 4));
        GLP_HIP(hipMalloc(&j->hs.chunk_cnt, p.hub_chunks.size() * 4));
        GLP_HIP(hipMalloc(&j->hs.best, p.hubs.size() * 8));
        GLP_HIP(hipMemsetAsync(j->hs.best, 0, p.hubs.size() * 8, stream));
    }
    GLP_HIP(hipMalloc(&j->changes, 8 * kChangeSlots));
    GLP_HIP(hipMemsetAsync(j->changes, 0, 8 * kChangeSlots, stream));
    GLP_HIP(hipStreamSynchronize(stream));
    j->cur = 0;
    j->done_steps = 0;
    return GLP_OK;
}

// Advances every running job by one step with folded launches.
int glp_server::tick() {
    std::vector<ServerJob*> act;
    for (auto* j : running)
        if (j->state == 1 && j->done_steps < j->steps && !j->cancel) act.push_back(j);
    const int32_t nj = (int32_t)act.size();
    if (nj == 0) return GLP_OK;
    bool weighted = false;
    for (int32_t k = 0; k < nj; ++k) {
        ServerJob* j = act[k];
        const DeviceGraph& dg = *j->dg;
        const Plan& p = dg.plan;
        JobDesc& d = h_jobs[k];
        std::memset(&d, 0, sizeof d);
        StepArgs& a = d.a;
        a.off = dg.off;
        a.nbr = dg.nbr;
        a.w = j->p.weighted ? dg.w : nullptr;
        a.L = j->L[j->cur];
        a.keep = j->L[j->cur];
        a.Lw = j->L[j->cur ^ 1];
        a.changes = j->changes;
        a.n = dg.n;
        a.t = j->done_steps + 1;
        if (j->p.method == GLP_LAYERED) {
            a.Vr = j->V[j->cur];
            a.Vw = j->V[j->cur ^ 1];
            a.score_mul = j->p.gamma_den + j->p.gamma_num;
            a.score_vmul = j->p.gamma_num;
            int rc = launch_fill_i32(j->V[j->cur ^ 1], 0, dg.n, stream);
            if (rc != GLP_OK) return rc;
        } else {
            a.score_mul = 1;
            a.score_vmul = 0;
        }
        weighted |= j->p.weighted != 0;
        d.groups = dg.d_groups;
        d.ngroups = (int32_t)p.groups_with_edges;  // layered hooks skip edge-less groups
        d.ngroups_e = (int32_t)p.groups_with_edges;
        d.half = dg.d_half;
        d.nhalf = (int32_t)p.half.size();
        d.wave = dg.d_wave;
        d.nwave = (int32_t)p.wave.size();
        for (int c = 0; c < kMaxBlockClasses; ++c) d.block[c] = dg.d_block[c];
        HubArgs& h = d.hub;
        h.hubs = dg.d_hubs;
        h.chunks = dg.d_hub_chunks;
        h.parts = dg.d_hub_parts;
        h.chunk0 = dg.d_hub_chunk0;
        h.base = dg.d_hub_base;
        h.hist_base = dg.d_hub_hist_base;
        h.chunk = tune.hub_chunk;
        h.part_keys = tune.part_keys;
        h.part_limit = tune.part_limit;
        h.nhubs = (int32_t)p.hubs.size();
        h.nparts = (int32_t)p.hub_parts.size();
        h.hs = j->hs;
        job_blocks(p, d.blocks);
    }
    int32_t grid[kLaunches];
    for (int L = 0; L < kLaunches; ++L) {
        int32_t* row = h_cum + (size_t)L * (nj + 1);
        row[0] = 0;
        for (int32_t k = 0; k < nj; ++k) row[k + 1] = row[k] + h_jobs[k].blocks[L];
        grid[L] = row[nj];
    }
    GLP_HIP(hipMemcpyAsync(d_jobs, h_jobs, sizeof(JobDesc) * nj, hipMemcpyHostToDevice, stream));
    GLP_HIP(hipMemcpyAsync(d_cum, h_cum, sizeof(int32_t) * kLaunches * (nj + 1), hipMemcpyHostToDevice,
                           stream));
    int rc = launch_multi(weighted, d_jobs, d_cum, nj, grid, tune, class_cap, stream);
    if (rc != GLP_OK) return rc;
    for (int32_t k = 0; k < nj; ++k)
        if (h_jobs[k].a.Vw && (rc = launch_volume(h_jobs[k].a.Lw, h_jobs[k].a.Vw, h_jobs[k].a.n, stream)))
            return rc;
    GLP_HIP(hipStreamSynchronize(stream));
    for (auto* j : act) {
        j->cur ^= 1;
        j->done_steps += 1;
    }
    return GLP_OK;
}

void glp_server::run() {
    (void)hipSetDevice(0);
    for (;;) {
        std::vector<ServerJob*> admitted, finished, cancelled;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty() || !running.empty(); });
            if (stop && queue.empty() && running.empty()) return;
            // drop cancelled queued jobs; admit in arrival order while memory and slots allow
            for (auto it = queue.begin(); it != queue.end();) {
                if ((*it)->cancel) {
                    (*it)->state = 3;
                    (*it)->done_ms = now();
                    it = queue.erase(it);
                } else {
                    ++it;
                }
            }
            while (!queue.empty() && (int32_t)running.size() < max_slots) {
                ServerJob* j = queue.front();
                if (used + j->bytes > budget) break;
                queue.pop_front();
                used += j->bytes;
                running.push_back(j);
                admitted.push_back(j);
            }
        }
        for (auto* j : admitted) {
            int rc = admit(j);
            std::lock_guard<std::mutex> lk(mu);
            j->admit_ms = now();
            j->state = rc == GLP_OK ? 1 : rc;
            if (rc != GLP_OK) {
                worker_msg = glp_last_error();
                j->release();
            }
        }
        int rc = tick();
        std::lock_guard<std::mutex> lk(mu);
        if (rc != GLP_OK) {
            worker_error = rc;
            worker_msg = glp_last_error();
            for (auto* j : running) {
                j->state = rc;
                j->done_ms = now();
                j->release();
            }
            used = 0;
            running.clear();
            cv_done.notify_all();
            continue;
        }
        for (auto it = running.begin(); it != running.end();) {
            ServerJob* j = *it;
            const bool failed = j->state < 0;
            if (failed || j->cancel || j->done_steps >= j->steps) {
                if (!failed && !j->cancel) {
                    j->result.resize(j->g->csr.n);
                    if (j->g->csr.n &&
                        hipMemcpy(j->result.data(), j->L[j->cur], sizeof(int32_t) * j->g->csr.n,
                                  hipMemcpyDeviceToHost) != hipSuccess)
                        j->state = GLP_E_HIP;
                    else
                        j->state = 2;
                } else if (j->cancel) {
                    j->state = 3;
                }
                j->done_ms = now();
                j->release();
                used -= j->bytes;
                it = running.erase(it);
            } else {
                ++it;
            }
        }
        cv_done.notify_all();
    }
}

extern "C" int glp_server_create(uint64_t hbm_budget_bytes, int32_t max_slots, glp_server** out) {
    clear_error();
    if (!out || hbm_budget_bytes == 0 || max_slots < 1 || max_slots > 65536)
        return set_error(GLP_E_ARG, "bad server arguments");
    *out = nullptr;
    int rc = ensure_device();
    if (rc != GLP_OK) return rc;
    init_kernels();
    auto sv = std::make_unique<glp_server>();
    sv->budget = hbm_budget_bytes;
    sv->max_slots = max_slots;
    sv->tune = default_tuning();
    {
        // block-class LDS capacities follow the snapshot of the tuning
        HostCSR empty;
        empty.n = 0;
        empty.offsets.assign(1, 0);
        Plan p;
        if ((rc = build_plan(empty, sv->tune, p)) != GLP_OK) return rc;
        for (int c = 0; c < kMaxBlockClasses; ++c) sv->class_cap[c] = p.class_cap[c];
    }
    GLP_HIP(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking));
    GLP_HIP(hipMalloc(&sv->d_jobs, sizeof(JobDesc) * max_slots));
    GLP_HIP(hipMalloc(&sv->d_cum, sizeof(int32_t) * kLaunches * (max_slots + 1)));
    GLP_HIP(hipHostMalloc((void**)&sv->h_jobs, sizeof(JobDesc) * max_slots, hipHostMallocDefault));
    GLP_HIP(hipHostMalloc((void**)&sv->h_cum, sizeof(int32_t) * kLaunches * (max_slots + 1),
                          hipHostMallocDefault));
    glp_server* raw = sv.get();
    sv->worker = std::thread([raw] { raw->run(); });
    *out = sv.release();
    return GLP_OK;
}

extern "C" int glp_server_submit(glp_server* sv, const glp_graph* g, const glp_params* p, int32_t steps,
                                 int64_t* job_id) {
    clear_error();
    if (!sv || !g || !p || !job_id || steps < 0) return set_error(GLP_E_ARG, "bad submit arguments");
    if (p->method != GLP_CLASSIC && p->method != GLP_LAYERED)
        return set_error(GLP_E_UNSUPPORTED, "the server runs classic and layered LP");
    if (p->gather) return set_error(GLP_E_UNSUPPORTED, "the server fuses the gather");
    int rc = check_params(g->csr, p);
    if (rc != GLP_OK) return rc;
    auto j = std::make_unique<ServerJob>();
    j->g = std::make_unique<glp_graph>();
    j->g->csr = g->csr;  // private copy: the caller may free its graph
    j->p = *p;
    j->steps = steps;
    if ((rc = build_plan(j->g->csr, sv->tune, j->plan)) != GLP_OK) return rc;
    j->bytes = job_bytes(j->g->csr, j->plan, *p);
    if (j->bytes > sv->budget)
        return set_error(GLP_E_BUDGET, "job needs %zu bytes, budget is %llu", j->bytes,
                         (unsigned long long)sv->budget);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return set_error(GLP_E_STATE, "server is shutting down");
    j->id = sv->next_id++;
    j->submit_ms = sv->now();
    *job_id = j->id;
    sv->queue.push_back(j.get());
    sv->jobs[j->id] = std::move(j);
    sv->cv_work.notify_one();
    return GLP_OK;
}

static void fill_stats(const ServerJob& j, glp_job_stats* st) {
    if (!st) return;
    st->submit_ms = j.submit_ms;
    st->admit_ms = j.admit_ms;
    st->done_ms = j.done_ms;
    st->bytes = (int64_t)j.bytes;
    st->steps = j.done_steps;
    st->state = j.state;
}

extern "C" int glp_server_poll(glp_server* sv, int64_t job_id, glp_job_stats* stats) {
    clear_error();
    if (!sv) return set_error(GLP_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(job_id);
    if (it == sv->jobs.end()) return set_error(GLP_E_ARG, "unknown job %lld", (long long)job_id);
    fill_stats(*it->second, stats);
    return GLP_OK;
}

extern "C" int glp_server_wait(glp_server* sv, int64_t job_id, int32_t* labels_out, glp_job_stats* stats) {
    clear_error();
    if (!sv) return set_error(GLP_E_ARG, "null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(job_id);
    if (it == sv->jobs.end()) return set_error(GLP_E_ARG, "unknown job %lld", (long long)job_id);
    ServerJob* j = it->second.get();
    sv->cv_done.wait(lk, [&] { return j->state >= 2 || j->state < 0; });
    fill_stats(*j, stats);
    if (j->state < 0) return set_error(j->state, "job failed: %s", sv->worker_msg.c_str());
    if (j->state == 3) return set_error(GLP_E_STATE, "job was cancelled");
    if (labels_out && !j->result.empty())
        std::memcpy(labels_out, j->result.data(), sizeof(int32_t) * j->result.size());
    return GLP_OK;
}

extern "C" int glp_server_cancel(glp_server* sv, int64_t job_id) {
    clear_error();
    if (!sv) return set_error(GLP_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(job_id);
    if (it == sv->jobs.end()) return set_error(GLP_E_ARG, "unknown job %lld", (long long)job_id);
    it->second->cancel = true;
    sv->cv_work.notify_one();
    return GLP_OK;
}

extern "C" int glp_server_free(glp_server* sv) {
    if (!sv) return GLP_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        sv->cv_work.notify_all();
    }
    if (sv->worker.joinable()) sv->worker.join();
    (void)hipFree(sv->d_jobs);
    (void)hipFree(sv->d_cum);
    (void)hipHostFree(sv->h_jobs);
    (void)hipHostFree(sv->h_cum);
    if (sv->stream) (void)hipStreamDestroy(sv->stream);
    delete sv;
    return GLP_OK;
}
