// Serving: a request loop with continuous batching under an HBM budget.
//
// Jobs (a graph, classic or layered LP, a step count) are queued by
// glp_server_submit* and admitted in arrival order as soon as their device
// footprint fits the budget and a slot is free.  Graphs are device-resident:
// glp_server_add_graph uploads one and glp_server_add_rmat generates one in
// HBM, once, for every job that names it; glp_server_submit copies a caller's
// graph for one job, uploaded by the upload thread when the job is admitted.
//
// Three threads share the work, so the device does not wait on the host:
//  * the worker owns the tick stream.  Every tick advances every running job
//    by one step with one launch per tier, the jobs folded together through
//    a device-side job table (k_multi), plus one folded launch that zeroes
//    and one that fills the layered jobs' volumes.  Layered jobs with
//    gamma > 0 run the slot-score hooks from a second table (one more set of
//    tier launches), after a folded launch writes their {label, volume}
//    records.  A step's tables go up in one copy from a ring of three
//    pinned staging slots; the worker waits only for the step that last
//    used a slot, so two steps stay queued on the device while it prepares
//    the next.  Jobs admitted meanwhile start (their state initialised on a
//    second stream, lstream) and join the next tick; those far smaller than
//    the tick's work (light jobs) run all their steps at once on lstream,
//    beside the heavy stream, instead of one step per tick.
//  * the upload thread copies private graphs to the device on its own
//    stream while the ticks run.
//  * the completion thread waits for a finished job's last tick, copies its
//    labels out on its own stream and returns its memory.
// Job state (labels, volumes, hub scratch) comes from a caching pool, so
// admission and completion make no hipMalloc / hipFree once the pool has
// warmed up (hipFree waits for the whole device).  Classic jobs run the
// layered hooks with gamma = 0, which is classic LP, so one kernel set
// serves both kinds (and gamma = 0 layered jobs).  Results do not depend on what else runs: each job's
// arithmetic touches only its own arrays.  Every field the threads share is
// read and written under `mu` or is atomic.

#include "device.h"
#include "session.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace glp {

// Caching device allocator: a freed block is kept and handed to a later
// request of at most 1.25x less size (plus 1 MiB).
class DevicePool {
public:
    ~DevicePool() { trim(0); }
    int alloc(size_t bytes, void** out, size_t* cap) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.lower_bound(bytes);
            if (it != free_.end() && it->first <= bytes + bytes / 4 + (1u << 20)) {
                *out = it->second;
                *cap = it->first;
                cached_ -= it->first;
                free_.erase(it);
                return GLP_OK;
            }
        }
        const size_t c = (std::max<size_t>(bytes, 256) + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
        GLP_HIP(hipMalloc(out, c));
        *cap = c;
        return GLP_OK;
    }
    void release(void* p, size_t cap) {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(cap, p);
        cached_ += cap;
    }
    // Frees cached blocks, largest first, until at most `keep` bytes stay.
    void trim(size_t keep) {
        std::lock_guard<std::mutex> lk(mu_);
        while (cached_ > keep && !free_.empty()) {
            auto it = std::prev(free_.end());
            (void)hipFree(it->second);
            cached_ -= it->first;
            free_.erase(it);
        }
    }

private:
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    size_t cached_ = 0;
};

// A device-resident graph and its tier plan.
struct ResidentGraph {
    int32_t id = 0;
    std::unique_ptr<DeviceGraph> dg;
    int32_t n = 0;
    int64_t m = 0;
    int32_t max_degree = 0;
    uint32_t wmax = 0;
    bool weighted = false;
    size_t bytes = 0;
    int32_t refs = 0;          // jobs queued or running on it, and copies in progress
    bool removed = false;      // takes no new jobs; freed when refs reaches 0
    bool ready = false;        // the device copy is complete
    bool private_job = false;  // made by glp_server_submit for one job
    std::unique_ptr<glp_graph> host;  // private: the copy waiting for upload
    Plan plan;                        // private: its plan, until uploaded
};

struct ServerJob {
    int64_t id = 0;
    ResidentGraph* rg = nullptr;
    glp_params p{};
    int32_t steps = 0;
    size_t bytes = 0;  // charged against the budget (job state, plus a private graph)
    // 0 queued, 4 admitted (uploading or starting), 1 running, 2 done,
    // 3 cancelled, < 0 error code
    std::atomic<int32_t> state{0};
    std::atomic<bool> cancel{false};
    std::atomic<int32_t> done_steps{0};
    std::atomic<double> start_ms{0};
    int32_t error = 0;
    double submit_ms = 0, admit_ms = 0, done_ms = 0;
    // device state (the worker's while running)
    void* block = nullptr;
    size_t block_cap = 0;
    int32_t* L[2] = {nullptr, nullptr};
    int32_t* V[2] = {nullptr, nullptr};
    int2* LV = nullptr;  // layered, gamma > 0: per-step {label, volume of the label} records
    HubScratch hs;
    unsigned long long* changes = nullptr;
    int cur = 0;
    hipEvent_t done_ev = nullptr;
    std::vector<int32_t> result;
};

// Sub-allocation of one job's state in one pooled block.
struct JobLayout {
    size_t L[2], V[2], LV, pairs, hist, lstart, best, changes, bytes;
};

// Slot-score (gamma > 0) layered jobs keep per-vertex {label, volume} records.
static bool job_records(const glp_params& p) { return p.method == GLP_LAYERED && p.gamma_num > 0; }

static JobLayout job_layout(int32_t n, const Plan& p, bool layered, bool records) {
    JobLayout o{};
    size_t at = 0;
    auto take = [&](size_t b) {
        const size_t here = at;
        at += (b + 255) & ~(size_t)255;
        return here;
    };
    const size_t nn = std::max<size_t>((size_t)n, 1);
    o.L[0] = take(nn * 4);
    o.L[1] = take(nn * 4);
    o.V[0] = layered ? take(nn * 4) : 0;
    o.V[1] = layered ? take(nn * 4) : 0;
    o.LV = records ? take(nn * 8) : 0;
    o.pairs = take((size_t)p.pair_slots * 8);
    o.hist = take((size_t)p.hist_slots * 4);
    o.lstart = take((size_t)p.hist_slots * 4);
    o.best = take(p.hubs.size() * 8);
    o.changes = take(8 * kChangeSlots);
    o.bytes = at;
    return o;
}

// One staging slot of the tick ring: the job table, its launch prefix sums
// and the layered jobs' volume table, packed in one pinned buffer and
// copied to its device twin in one transfer.
struct TickSlot {
    char* host = nullptr;
    char* dev = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool used = false;
};
constexpr int kTickSlots = 3;

}  // namespace glp

using namespace glp;

struct glp_server {
    uint64_t budget = 0;
    int32_t max_slots = 0;
    Tuning tune;
    int32_t class_cap[kMaxBlockClasses] = {};
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_upload, cv_finish;
    std::deque<ServerJob*> queue;     // waiting for admission
    std::deque<ServerJob*> upload_q;  // admitted, private graph to upload
    std::vector<ServerJob*> uploaded; // uploads finished, to start
    std::deque<ServerJob*> finish_q;  // last tick enqueued, to complete
    std::map<int64_t, std::unique_ptr<ServerJob>> jobs;
    std::map<int32_t, std::unique_ptr<ResidentGraph>> graphs;
    uint64_t used = 0;   // budget charged: resident graphs + admitted jobs
    int32_t active = 0;  // admitted jobs not yet completed
    int64_t next_id = 1;
    int32_t next_graph = 1;
    bool stop = false;
    bool worker_done = false;
    std::string worker_msg;
    std::vector<hipEvent_t> free_events;
    std::thread worker, uploader, completer;
    hipStream_t stream = nullptr;      // ticks: one step of every heavy job
    hipStream_t lstream = nullptr;     // job initialisation and light jobs' steps
    hipEvent_t init_ev = nullptr;      // after the latest initialisations on lstream
    hipStream_t up_stream = nullptr;   // private graph uploads
    hipStream_t dl_stream = nullptr;   // result copies
    hipStream_t add_stream = nullptr;  // glp_server_add_* and graph copies (callers' threads)
    std::mutex add_mu;
    DevicePool pool;
    TickSlot slots[kTickSlots];   // staging of the heavy steps (stream)
    int next_slot = 0;
    TickSlot lslots[kTickSlots];  // staging of the light steps (lstream)
    int next_lslot = 0;
    std::vector<ServerJob*> running;  // the worker's
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    bool can_admit_front() const {
        return !queue.empty() && active < max_slots && used + queue.front()->bytes <= budget;
    }
    // The first queued job cannot fit even with no job running (resident
    // graphs hold the rest of the budget): it fails instead of waiting.
    bool front_never_fits() const {
        return !queue.empty() && active == 0 && used + queue.front()->bytes > budget;
    }
    void run();
    void upload_loop();
    void complete_loop();
    int start_job(ServerJob* j);
    int tick();
    int step_jobs(const std::vector<ServerJob*>& jobs, hipStream_t s, TickSlot* ring, int& next);
    void hand_over(ServerJob* j, hipStream_t s);
    std::unique_ptr<ResidentGraph> release_graph_locked(ResidentGraph* rg);
    hipEvent_t take_event();
};

hipEvent_t glp_server::take_event() {
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!free_events.empty()) {
            hipEvent_t e = free_events.back();
            free_events.pop_back();
            return e;
        }
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Drops a graph reference (under mu).  A removed or private graph that no
// job uses any more leaves the registry; the caller destroys the returned
// graph after unlocking (its hipFree waits for the device).
std::unique_ptr<ResidentGraph> glp_server::release_graph_locked(ResidentGraph* rg) {
    rg->refs -= 1;
    if (rg->refs > 0 || !(rg->removed || rg->private_job)) return nullptr;
    if (!rg->private_job) used -= rg->bytes;  // a private graph's bytes are in its job's
    auto it = graphs.find(rg->id);
    std::unique_ptr<ResidentGraph> out = std::move(it->second);
    graphs.erase(it);
    return out;
}

static size_t resident_bytes(const DeviceGraph& dg, bool weighted) {
    return dg.bytes + ((size_t)dg.n + 1) * 4 + (size_t)dg.m * 4 * (weighted ? 2 : 1);
}

// Allocates and initialises a job's state on lstream (the heavy stream waits
// for init_ev before its next step).
int glp_server::start_job(ServerJob* j) {
    const DeviceGraph& dg = *j->rg->dg;
    const Plan& p = dg.plan;
    const bool layered = j->p.method == GLP_LAYERED;
    const bool records = job_records(j->p);
    const JobLayout lay = job_layout(dg.n, p, layered, records);
    int rc = pool.alloc(lay.bytes, &j->block, &j->block_cap);
    if (rc != GLP_OK) return rc;
    char* base = static_cast<char*>(j->block);
    j->L[0] = reinterpret_cast<int32_t*>(base + lay.L[0]);
    j->L[1] = reinterpret_cast<int32_t*>(base + lay.L[1]);
    j->V[0] = layered ? reinterpret_cast<int32_t*>(base + lay.V[0]) : nullptr;
    j->V[1] = layered ? reinterpret_cast<int32_t*>(base + lay.V[1]) : nullptr;
    j->LV = records ? reinterpret_cast<int2*>(base + lay.LV) : nullptr;
    if (!p.hubs.empty()) {
        j->hs.pairs = reinterpret_cast<unsigned long long*>(base + lay.pairs);
        j->hs.hist = reinterpret_cast<uint32_t*>(base + lay.hist);
        j->hs.lstart = reinterpret_cast<uint32_t*>(base + lay.lstart);
        j->hs.best = reinterpret_cast<unsigned long long*>(base + lay.best);
        if ((rc = launch_zero_bytes(j->hs.best, (int64_t)p.hubs.size() * 8, lstream))) return rc;
    }
    j->changes = reinterpret_cast<unsigned long long*>(base + lay.changes);
    if ((rc = launch_zero_bytes(j->changes, 8 * kChangeSlots, lstream))) return rc;
    if ((rc = launch_iota(j->L[0], dg.n, lstream))) return rc;
    if ((rc = launch_iota(j->L[1], dg.n, lstream))) return rc;
    if (layered && (rc = launch_fill_i32(j->V[0], 1, dg.n, lstream))) return rc;  // V_0 = 1
    j->cur = 0;
    j->done_steps = 0;
    j->start_ms = now();
    j->state = 1;
    return GLP_OK;
}

// Hands a job whose last step is enqueued on s (or that ends without one) to
// the completion thread.
void glp_server::hand_over(ServerJob* j, hipStream_t s) {
    j->done_ev = take_event();
    if (j->done_ev) (void)hipEventRecord(j->done_ev, s);
    std::lock_guard<std::mutex> lk(mu);
    finish_q.push_back(j);
    cv_finish.notify_one();
}

// The light jobs of a tick together cost at most 1/kLightShare of one step
// of every running job (cost: vertices + edges per step).
constexpr int64_t kLightShare = 2;

// One tick.  Jobs that have not stepped yet and are far smaller than the
// tick's work (light: taken smallest first while their whole runs sum to at
// most 1/kLightShare of one step of every running job) run all their steps
// now on lstream, beside the heavy stream, and are handed over at once;
// every other job advances one step on the heavy stream.  A small job then
// finishes within one tick instead of waiting for ten ticks whose length the
// largest jobs set, and the heavy ticks do not grow by its launches.  A job
// that has stepped on the heavy stream stays there (its steps are ordered
// by that stream).
int glp_server::tick() {
    std::vector<ServerJob*> light, heavy;
    std::vector<std::pair<int64_t, ServerJob*>> fresh;
    int64_t total = 0;
    for (auto* j : running) {
        if (j->done_steps >= j->steps || j->cancel) continue;
        const int64_t per = (int64_t)j->rg->n + j->rg->m;
        total += per;
        if (j->done_steps == 0) fresh.emplace_back((int64_t)j->steps * per, j);
        else heavy.push_back(j);
    }
    std::stable_sort(fresh.begin(), fresh.end(),
                     [](const std::pair<int64_t, ServerJob*>& x, const std::pair<int64_t, ServerJob*>& y) {
                         return x.first < y.first;
                     });
    int64_t taken = 0;
    for (const auto& f : fresh) {
        if ((taken + f.first) * kLightShare <= total) {
            taken += f.first;
            light.push_back(f.second);
        } else {
            heavy.push_back(f.second);
        }
    }
    while (!light.empty()) {
        int rc = step_jobs(light, lstream, lslots, next_lslot);
        if (rc != GLP_OK) return rc;
        for (auto it = light.begin(); it != light.end();) {
            ServerJob* j = *it;
            if (j->cancel || j->done_steps >= j->steps) {
                hand_over(j, lstream);
                running.erase(std::find(running.begin(), running.end(), j));
                it = light.erase(it);
            } else {
                ++it;
            }
        }
    }
    return heavy.empty() ? GLP_OK : step_jobs(heavy, stream, slots, next_slot);
}

// Enqueues one step of `jobs` on stream s with folded launches, staging the
// tables through `ring`.
//
// Two job tables: classic and gamma = 0 layered jobs run the layered hooks
// (gamma = 0 reads no volume), layered jobs with gamma > 0 the slot-score
// hooks, which read each neighbour's {label, volume} record (written for
// those jobs by one folded launch before the tiers) and score in the table
// slots.
int glp_server::step_jobs(const std::vector<ServerJob*>& jobs, hipStream_t s, TickSlot* ring, int& next) {
    std::vector<ServerJob*> act = jobs;
    if (act.empty()) return GLP_OK;
    // slot-score jobs last: jobs [0, na) form table A, [na, nj) table B
    std::stable_partition(act.begin(), act.end(), [](const ServerJob* j) { return j->LV == nullptr; });
    const int32_t nj = (int32_t)act.size();
    int32_t na = 0, nl = 0;
    for (auto* j : act) {
        na += j->LV == nullptr ? 1 : 0;
        nl += j->p.method == GLP_LAYERED ? 1 : 0;
    }
    const int32_t nb = nj - na;
    // packed layout of this tick's tables
    auto up16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
    const size_t o_jobs = 0;
    const size_t o_cum_a = up16(o_jobs + sizeof(JobDesc) * nj);
    const size_t o_cum_b = up16(o_cum_a + sizeof(int32_t) * kLaunches * (na + 1));
    const size_t o_vol = up16(o_cum_b + sizeof(int32_t) * kLaunches * (nb + 1));
    const size_t o_zc = up16(o_vol + sizeof(VolumeJob) * std::max(nl, 1));
    const size_t o_vc = up16(o_zc + sizeof(int32_t) * (nl + 1));
    const size_t o_rec = up16(o_vc + sizeof(int32_t) * (nl + 1));
    const size_t o_rc = up16(o_rec + sizeof(VolumeJob) * std::max(nb, 1));
    const size_t total = up16(o_rc + sizeof(int32_t) * (nb + 1));
    TickSlot& sl = ring[next];
    next = (next + 1) % kTickSlots;
    if (sl.used) GLP_HIP(hipEventSynchronize(sl.ev));  // the step that last used this slot is done
    if (total > sl.cap) {
        (void)hipHostFree(sl.host);
        (void)hipFree(sl.dev);
        sl.host = nullptr;
        sl.dev = nullptr;
        sl.cap = 0;
        const size_t cap = std::max(total * 2, (size_t)64 << 10);
        GLP_HIP(hipHostMalloc((void**)&sl.host, cap, hipHostMallocDefault));
        GLP_HIP(hipMalloc((void**)&sl.dev, cap));
        sl.cap = cap;
    }
    JobDesc* h_jobs = reinterpret_cast<JobDesc*>(sl.host + o_jobs);
    VolumeJob* h_vol = reinterpret_cast<VolumeJob*>(sl.host + o_vol);
    int32_t* h_zc = reinterpret_cast<int32_t*>(sl.host + o_zc);
    int32_t* h_vc = reinterpret_cast<int32_t*>(sl.host + o_vc);
    VolumeJob* h_rec = reinterpret_cast<VolumeJob*>(sl.host + o_rec);
    int32_t* h_rc = reinterpret_cast<int32_t*>(sl.host + o_rc);
    bool weighted_a = false, weighted_b = false;
    int32_t li = 0, ri = 0;
    h_zc[0] = 0;
    h_vc[0] = 0;
    h_rc[0] = 0;
    for (int32_t k = 0; k < nj; ++k) {
        ServerJob* j = act[k];
        const DeviceGraph& dg = *j->rg->dg;
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
            h_vol[li] = VolumeJob{a.Lw, a.Vw, dg.n, 0, nullptr};
            int32_t zb = 0, vb = 0;
            volume_job_blocks(dg.n, &zb, &vb);
            h_zc[li + 1] = h_zc[li] + zb;
            h_vc[li + 1] = h_vc[li] + vb;
            ++li;
            if (j->LV) {
                a.LV = j->LV;
                h_rec[ri] = VolumeJob{a.L, j->V[j->cur], dg.n, 0, j->LV};  // read-side volumes
                h_rc[ri + 1] = h_rc[ri] + zb;  // the records take the zero launch's rows per workgroup
                ++ri;
            }
        } else {
            a.score_mul = 1;
            a.score_vmul = 0;
        }
        (k < na ? weighted_a : weighted_b) |= j->p.weighted != 0;
        d.groups = dg.d_groups;
        d.ngroups = (int32_t)p.groups_with_edges;  // both hook sets skip edge-less groups
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
        h.chunk = p.tune.hub_chunk;
        h.part_keys = p.tune.part_keys;
        h.part_limit = p.tune.part_limit;
        h.nhubs = (int32_t)p.hubs.size();
        h.nparts = (int32_t)p.hub_parts.size();
        h.hs = j->hs;
        job_blocks(p, d.blocks);
    }
    // per launch, the prefix sums of each table's workgroups
    int32_t grid_a[kLaunches], grid_b[kLaunches];
    auto prefix = [&](size_t off, int32_t first, int32_t cnt, int32_t* grid) {
        int32_t* rows = reinterpret_cast<int32_t*>(sl.host + off);
        for (int L = 0; L < kLaunches; ++L) {
            int32_t* row = rows + (size_t)L * (cnt + 1);
            row[0] = 0;
            for (int32_t k = 0; k < cnt; ++k) row[k + 1] = row[k] + h_jobs[first + k].blocks[L];
            grid[L] = row[cnt];
        }
    };
    prefix(o_cum_a, 0, na, grid_a);
    prefix(o_cum_b, na, nb, grid_b);
    GLP_HIP(hipMemcpyAsync(sl.dev, sl.host, total, hipMemcpyHostToDevice, s));
    const JobDesc* d_jobs = reinterpret_cast<const JobDesc*>(sl.dev + o_jobs);
    const VolumeJob* d_vol = reinterpret_cast<const VolumeJob*>(sl.dev + o_vol);
    auto dev_i32 = [&](size_t off) { return reinterpret_cast<const int32_t*>(sl.dev + off); };
    int rc = launch_multi_volume(kVolumeZero, d_vol, dev_i32(o_zc), nl, h_zc[nl], s);
    if (rc != GLP_OK) return rc;
    if ((rc = launch_multi_volume(kVolumeRecords, reinterpret_cast<const VolumeJob*>(sl.dev + o_rec),
                                  dev_i32(o_rc), nb, h_rc[nb], s)) != GLP_OK)
        return rc;
    if (na > 0 && (rc = launch_multi(weighted_a, false, d_jobs, dev_i32(o_cum_a), na, grid_a, tune, class_cap,
                                     s)) != GLP_OK)
        return rc;
    if (nb > 0 && (rc = launch_multi(weighted_b, true, d_jobs + na, dev_i32(o_cum_b), nb, grid_b, tune,
                                     class_cap, s)) != GLP_OK)
        return rc;
    if ((rc = launch_multi_volume(kVolumeFill, d_vol, dev_i32(o_vc), nl, h_vc[nl], s)) != GLP_OK) return rc;
    GLP_HIP(hipEventRecord(sl.ev, s));
    sl.used = true;
    for (auto* j : act) {
        j->cur ^= 1;
        j->done_steps += 1;
    }
    return GLP_OK;
}

void glp_server::run() {
    (void)hipSetDevice(0);
    for (;;) {
        std::vector<ServerJob*> starting;
        std::vector<std::unique_ptr<ResidentGraph>> dead;  // destroyed after unlocking
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] {
                return (stop && active == 0 && queue.empty()) || !uploaded.empty() || !running.empty() ||
                       can_admit_front() || front_never_fits() || (!queue.empty() && queue.front()->cancel);
            });
            if (stop && queue.empty() && running.empty() && active == 0) break;
            // cancelled queued jobs end at once
            bool ended = false;
            for (auto it = queue.begin(); it != queue.end();) {
                if ((*it)->cancel) {
                    ServerJob* j = *it;
                    j->done_ms = now();
                    dead.push_back(release_graph_locked(j->rg));
                    j->rg = nullptr;
                    j->state = 3;
                    it = queue.erase(it);
                    ended = true;
                } else {
                    ++it;
                }
            }
            while (front_never_fits()) {
                ServerJob* j = queue.front();
                queue.pop_front();
                j->done_ms = now();
                dead.push_back(release_graph_locked(j->rg));
                j->rg = nullptr;
                worker_msg = "the job does not fit the budget next to the resident graphs";
                j->state = GLP_E_BUDGET;
                ended = true;
            }
            if (ended) cv_done.notify_all();
            // admission in arrival order while slots and memory allow
            while (can_admit_front()) {
                ServerJob* j = queue.front();
                queue.pop_front();
                used += j->bytes;
                active += 1;
                j->admit_ms = now();
                j->state = 4;
                if (j->rg->ready) {
                    starting.push_back(j);
                } else {
                    upload_q.push_back(j);
                    cv_upload.notify_one();
                }
            }
            for (auto* j : uploaded) starting.push_back(j);
            uploaded.clear();
        }
        dead.clear();
        bool started = false;
        for (auto* j : starting) {
            int rc = j->error ? j->error : start_job(j);
            if (rc != GLP_OK) {
                j->error = rc;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    worker_msg = glp_last_error();
                }
                hand_over(j, lstream);
                continue;
            }
            running.push_back(j);
            started = true;
        }
        if (started) {
            // the heavy stream's next step may include the jobs just
            // initialised on lstream
            if (hipEventRecord(init_ev, lstream) != hipSuccess ||
                hipStreamWaitEvent(stream, init_ev, 0) != hipSuccess) {
                (void)hipGetLastError();
                (void)hipStreamSynchronize(lstream);  // order it from the host instead
            }
        }
        int rc = tick();
        if (rc != GLP_OK) {
            {
                std::lock_guard<std::mutex> lk(mu);
                worker_msg = glp_last_error();
            }
            (void)hipStreamSynchronize(stream);
            (void)hipStreamSynchronize(lstream);
            for (auto* j : running) {
                j->error = rc;
                hand_over(j, stream);
            }
            running.clear();
            continue;
        }
        for (auto it = running.begin(); it != running.end();) {
            ServerJob* j = *it;
            if (j->cancel || j->done_steps >= j->steps) {
                hand_over(j, stream);
                it = running.erase(it);
            } else {
                ++it;
            }
        }
    }
    std::lock_guard<std::mutex> lk(mu);
    worker_done = true;
    cv_finish.notify_all();
    cv_upload.notify_all();
}

void glp_server::upload_loop() {
    (void)hipSetDevice(0);
    for (;;) {
        ServerJob* j = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_upload.wait(lk, [&] { return worker_done || !upload_q.empty(); });
            if (upload_q.empty()) break;  // the worker has ended
            j = upload_q.front();
            upload_q.pop_front();
        }
        ResidentGraph* rg = j->rg;
        DeviceGraph* dg = nullptr;
        int rc = device_graph_upload(rg->host->csr, std::move(rg->plan), up_stream, &dg);
        // the copies complete before the host copy is freed and before any
        // other stream runs the job
        if (rc == GLP_OK && hipStreamSynchronize(up_stream) != hipSuccess)
            rc = set_error(GLP_E_HIP, "graph upload: %s", hipGetErrorString(hipGetLastError()));
        std::lock_guard<std::mutex> lk(mu);
        if (rc == GLP_OK) {
            rg->dg.reset(dg);
            rg->host.reset();
            rg->ready = true;
        } else {
            j->error = rc;
            worker_msg = glp_last_error();
        }
        uploaded.push_back(j);
        cv_work.notify_one();
    }
}

void glp_server::complete_loop() {
    (void)hipSetDevice(0);
    for (;;) {
        ServerJob* j = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_finish.wait(lk, [&] { return !finish_q.empty() || worker_done; });
            if (finish_q.empty()) break;
            j = finish_q.front();
            finish_q.pop_front();
        }
        if (j->done_ev) (void)hipEventSynchronize(j->done_ev);
        int32_t st = 2;
        if (j->error) {
            st = j->error;
        } else if (j->cancel) {
            st = 3;
        } else {
            const int32_t n = j->rg->dg->n;
            j->result.resize((size_t)n);
            if (n && (hipMemcpyAsync(j->result.data(), j->L[j->cur], sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                                     dl_stream) != hipSuccess ||
                      hipStreamSynchronize(dl_stream) != hipSuccess))
                st = GLP_E_HIP;
        }
        pool.release(j->block, j->block_cap);
        j->block = nullptr;
        std::unique_ptr<ResidentGraph> dead;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (j->done_ev) free_events.push_back(j->done_ev);
            j->done_ev = nullptr;
            used -= j->bytes;
            active -= 1;
            dead = release_graph_locked(j->rg);
            j->rg = nullptr;
            j->done_ms = now();
            j->state = st;
            cv_done.notify_all();
            cv_work.notify_one();
        }
    }
}

// Bytes of a job's state (labels, volumes, hub scratch).
static size_t job_state_bytes(const ResidentGraph& rg, const Plan& plan, const glp_params& p) {
    return job_layout(rg.n, plan, p.method == GLP_LAYERED, job_records(p)).bytes;
}

static int check_job_params(const ResidentGraph& rg, const glp_params* p) {
    if (p->method != GLP_CLASSIC && p->method != GLP_LAYERED)
        return set_error(GLP_E_UNSUPPORTED, "the server runs classic and layered LP");
    if (p->gather) return set_error(GLP_E_UNSUPPORTED, "the server fuses the gather");
    if (p->frontier) return set_error(GLP_E_UNSUPPORTED, "the server steps every vertex (no frontier)");
    return check_params_shape(rg.n, rg.max_degree, rg.wmax, rg.weighted, p);
}

// Queues a job (takes mu).
static int enqueue_job(glp_server* sv, std::unique_ptr<ServerJob> j, int64_t* job_id) {
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
    {
        // The light jobs overlap the heavy ticks only if lstream does not
        // share the tick stream's hardware queue: streams share the
        // process's few queues (GPU_MAX_HW_QUEUES, 4 by default), and after
        // a process has created and destroyed many streams two new ones
        // can land on one queue.  A stream created with a CU mask (here
        // every CU) gets a queue of its own.
        int dev = 0, cus = 0;
        GLP_HIP(hipGetDevice(&dev));
        GLP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0xffffffffu);
        if (hipExtStreamCreateWithCUMask(&sv->lstream, (uint32_t)mask.size(), mask.data()) != hipSuccess) {
            (void)hipGetLastError();
            sv->lstream = nullptr;
            GLP_HIP(hipStreamCreateWithFlags(&sv->lstream, hipStreamNonBlocking));
        }
    }
    GLP_HIP(hipEventCreateWithFlags(&sv->init_ev, hipEventDisableTiming));
    GLP_HIP(hipStreamCreateWithFlags(&sv->up_stream, hipStreamNonBlocking));
    GLP_HIP(hipStreamCreateWithFlags(&sv->dl_stream, hipStreamNonBlocking));
    GLP_HIP(hipStreamCreateWithFlags(&sv->add_stream, hipStreamNonBlocking));
    for (auto& sl : sv->slots) GLP_HIP(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
    for (auto& sl : sv->lslots) GLP_HIP(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
    glp_server* raw = sv.get();
    sv->worker = std::thread([raw] { raw->run(); });
    sv->uploader = std::thread([raw] { raw->upload_loop(); });
    sv->completer = std::thread([raw] { raw->complete_loop(); });
    *out = sv.release();
    return GLP_OK;
}

// Registers an uploaded or generated graph, charging its bytes.
static int register_graph(glp_server* sv, std::unique_ptr<ResidentGraph> rg, int32_t* graph_id) {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return set_error(GLP_E_STATE, "server is shutting down");
    if (sv->used + rg->bytes > sv->budget)
        return set_error(GLP_E_BUDGET, "graph needs %zu bytes, %llu of the budget %llu are in use", rg->bytes,
                         (unsigned long long)sv->used, (unsigned long long)sv->budget);
    sv->used += rg->bytes;
    rg->id = sv->next_graph++;
    rg->ready = true;
    *graph_id = rg->id;
    sv->graphs[rg->id] = std::move(rg);
    return GLP_OK;
}

extern "C" int glp_server_add_graph(glp_server* sv, const glp_graph* g, int32_t* graph_id) {
    clear_error();
    if (!sv || !g || !graph_id) return set_error(GLP_E_ARG, "bad add_graph arguments");
    auto rg = std::make_unique<ResidentGraph>();
    const HostCSR& c = g->csr;
    rg->n = c.n;
    rg->m = c.m;
    rg->max_degree = c.max_degree;
    rg->weighted = c.weighted();
    for (uint32_t w : c.weights) rg->wmax = std::max(rg->wmax, w);
    Plan plan;
    int rc = build_plan(c, sv->tune, plan);
    if (rc != GLP_OK) return rc;
    DeviceGraph* dg = nullptr;
    {
        std::lock_guard<std::mutex> lk(sv->add_mu);
        if ((rc = device_graph_upload(c, std::move(plan), sv->add_stream, &dg)) != GLP_OK) return rc;
        // complete before any job stream reads it
        if (hipStreamSynchronize(sv->add_stream) != hipSuccess) {
            delete dg;
            return set_error(GLP_E_HIP, "graph upload: %s", hipGetErrorString(hipGetLastError()));
        }
    }
    rg->dg.reset(dg);
    rg->bytes = dg->bytes;
    return register_graph(sv, std::move(rg), graph_id);
}

extern "C" int glp_server_add_rmat(glp_server* sv, int32_t scale, int32_t edge_factor, double a, double b,
                                   double c, uint64_t seed, int32_t wmax, int32_t* graph_id) {
    clear_error();
    if (!sv || !graph_id) return set_error(GLP_E_ARG, "bad add_rmat arguments");
    std::lock_guard<std::mutex> lk(sv->add_mu);
    DeviceCSR dc;
    int rc = gen_rmat_device(scale, edge_factor, a, b, c, seed, wmax, sv->add_stream, &dc);
    if (rc != GLP_OK) return rc;
    auto dg = std::make_unique<DeviceGraph>();
    dg->n = dc.n;
    dg->m = dc.m;
    dg->off = dc.off;
    dg->nbr = dc.nbr;
    dg->w = dc.w;
    // the planner reads the degrees only: the offsets come back to the host
    HostCSR meta;
    meta.n = dc.n;
    meta.m = dc.m;
    meta.offsets.resize((size_t)dc.n + 1);
    GLP_HIP(hipMemcpyAsync(meta.offsets.data(), dc.off, ((size_t)dc.n + 1) * 4, hipMemcpyDeviceToHost,
                           sv->add_stream));
    GLP_HIP(hipStreamSynchronize(sv->add_stream));
    if ((rc = build_plan(meta, sv->tune, dg->plan)) != GLP_OK) return rc;
    if ((rc = device_plan_upload(dg.get(), sv->add_stream)) != GLP_OK) return rc;
    auto rg = std::make_unique<ResidentGraph>();
    rg->n = dc.n;
    rg->m = dc.m;
    rg->max_degree = dc.max_degree;
    rg->weighted = wmax > 0;
    rg->wmax = dc.wmax;
    rg->bytes = resident_bytes(*dg, rg->weighted);
    rg->dg = std::move(dg);
    return register_graph(sv, std::move(rg), graph_id);
}

static ResidentGraph* find_graph_locked(glp_server* sv, int32_t graph_id) {
    auto it = sv->graphs.find(graph_id);
    if (it == sv->graphs.end() || it->second->private_job || it->second->removed) return nullptr;
    return it->second.get();
}

extern "C" int glp_server_graph_info(glp_server* sv, int32_t graph_id, int32_t* n, int64_t* m,
                                     int32_t* max_degree, int64_t* bytes) {
    clear_error();
    if (!sv) return set_error(GLP_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    ResidentGraph* rg = find_graph_locked(sv, graph_id);
    if (!rg) return set_error(GLP_E_ARG, "unknown graph %d", graph_id);
    if (n) *n = rg->n;
    if (m) *m = rg->m;
    if (max_degree) *max_degree = rg->max_degree;
    if (bytes) *bytes = (int64_t)rg->bytes;
    return GLP_OK;
}

extern "C" int glp_server_graph_csr(glp_server* sv, int32_t graph_id, int32_t* offsets, int32_t* neighbors,
                                    uint32_t* weights) {
    clear_error();
    if (!sv) return set_error(GLP_E_ARG, "null server");
    ResidentGraph* rg = nullptr;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        rg = find_graph_locked(sv, graph_id);
        if (!rg) return set_error(GLP_E_ARG, "unknown graph %d", graph_id);
        if (weights && !rg->weighted) return set_error(GLP_E_ARG, "graph %d has no weights", graph_id);
        rg->refs += 1;  // kept alive while copying
    }
    const DeviceGraph* dg = rg->dg.get();
    int rc = GLP_OK;
    {
        std::lock_guard<std::mutex> lk(sv->add_mu);
        hipError_t e = hipSuccess;
        if (offsets)
            e = hipMemcpyAsync(offsets, dg->off, ((size_t)dg->n + 1) * 4, hipMemcpyDeviceToHost, sv->add_stream);
        if (e == hipSuccess && neighbors && dg->m)
            e = hipMemcpyAsync(neighbors, dg->nbr, (size_t)dg->m * 4, hipMemcpyDeviceToHost, sv->add_stream);
        if (e == hipSuccess && weights && dg->m)
            e = hipMemcpyAsync(weights, dg->w, (size_t)dg->m * 4, hipMemcpyDeviceToHost, sv->add_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(sv->add_stream);
        if (e != hipSuccess) rc = set_error(GLP_E_HIP, "graph copy: %s", hipGetErrorString(e));
    }
    std::unique_ptr<ResidentGraph> dead;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        dead = sv->release_graph_locked(rg);
    }
    return rc;
}

extern "C" int glp_server_remove_graph(glp_server* sv, int32_t graph_id) {
    clear_error();
    if (!sv) return set_error(GLP_E_ARG, "null server");
    std::unique_ptr<ResidentGraph> dead;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        ResidentGraph* rg = find_graph_locked(sv, graph_id);
        if (!rg) return set_error(GLP_E_ARG, "unknown graph %d", graph_id);
        rg->removed = true;
        rg->refs += 1;
        dead = sv->release_graph_locked(rg);  // leaves now unless jobs still use it
    }
    return GLP_OK;
}

extern "C" int glp_server_submit_resident(glp_server* sv, int32_t graph_id, const glp_params* p, int32_t steps,
                                          int64_t* job_id) {
    clear_error();
    if (!sv || !p || !job_id || steps < 0) return set_error(GLP_E_ARG, "bad submit arguments");
    auto j = std::make_unique<ServerJob>();
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        ResidentGraph* rg = find_graph_locked(sv, graph_id);
        if (!rg) return set_error(GLP_E_ARG, "unknown graph %d", graph_id);
        int rc = check_job_params(*rg, p);
        if (rc != GLP_OK) return rc;
        rg->refs += 1;
        j->rg = rg;
        j->bytes = job_state_bytes(*rg, rg->dg->plan, *p);
    }
    j->p = *p;
    j->steps = steps;
    ResidentGraph* rg = j->rg;
    int rc = enqueue_job(sv, std::move(j), job_id);
    if (rc != GLP_OK) {
        std::unique_ptr<ResidentGraph> dead;
        std::lock_guard<std::mutex> lk(sv->mu);
        dead = sv->release_graph_locked(rg);
    }
    return rc;
}

extern "C" int glp_server_submit(glp_server* sv, const glp_graph* g, const glp_params* p, int32_t steps,
                                 int64_t* job_id) {
    clear_error();
    if (!sv || !g || !p || !job_id || steps < 0) return set_error(GLP_E_ARG, "bad submit arguments");
    auto rg = std::make_unique<ResidentGraph>();
    rg->private_job = true;
    rg->host = std::make_unique<glp_graph>();
    rg->host->csr = g->csr;  // private copy: the caller may free its graph
    const HostCSR& c = rg->host->csr;
    rg->n = c.n;
    rg->m = c.m;
    rg->max_degree = c.max_degree;
    rg->weighted = c.weighted();
    for (uint32_t w : c.weights) rg->wmax = std::max(rg->wmax, w);
    int rc = check_job_params(*rg, p);
    if (rc != GLP_OK) return rc;
    if ((rc = build_plan(c, sv->tune, rg->plan)) != GLP_OK) return rc;
    auto j = std::make_unique<ServerJob>();
    j->p = *p;
    j->steps = steps;
    j->bytes = device_graph_bytes(c, rg->plan) + job_state_bytes(*rg, rg->plan, *p);
    rg->refs = 1;
    j->rg = rg.get();
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        rg->id = -(sv->next_graph++);  // private graphs are not visible to callers
        sv->graphs[rg->id] = std::move(rg);
    }
    ResidentGraph* raw = j->rg;
    rc = enqueue_job(sv, std::move(j), job_id);
    if (rc != GLP_OK) {
        std::unique_ptr<ResidentGraph> dead;
        std::lock_guard<std::mutex> lk(sv->mu);
        dead = sv->release_graph_locked(raw);
    }
    return rc;
}

static void fill_stats(const ServerJob& j, glp_job_stats* st) {
    if (!st) return;
    st->submit_ms = j.submit_ms;
    st->admit_ms = j.admit_ms;
    st->done_ms = j.done_ms;
    st->bytes = (int64_t)j.bytes;
    st->steps = j.done_steps;
    st->state = j.state;
    st->start_ms = j.start_ms;
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
    sv->cv_done.wait(lk, [&] {
        const int32_t s = j->state;
        return s == 2 || s == 3 || s < 0;
    });
    fill_stats(*j, stats);
    if (j->state < 0) return set_error(j->state, "job failed: %s", sv->worker_msg.c_str());
    if (j->state == 3) return set_error(GLP_E_STATE, "job was cancelled");
    if (labels_out && !j->result.empty())
        std::memcpy(labels_out, j->result.data(), sizeof(int32_t) * j->result.size());
    return GLP_OK;
}

extern "C" int glp_server_release(glp_server* sv, int64_t job_id) {
    clear_error();
    if (!sv) return set_error(GLP_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(job_id);
    if (it == sv->jobs.end()) return set_error(GLP_E_ARG, "unknown job %lld", (long long)job_id);
    const int32_t s = it->second->state;
    if (!(s == 2 || s == 3 || s < 0)) return set_error(GLP_E_STATE, "job %lld has not ended", (long long)job_id);
    sv->jobs.erase(it);
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
    if (sv->uploader.joinable()) sv->uploader.join();
    if (sv->completer.joinable()) sv->completer.join();
    (void)hipStreamSynchronize(sv->stream);
    (void)hipStreamSynchronize(sv->lstream);
    for (TickSlot* ring : {sv->slots, sv->lslots})
        for (int k = 0; k < kTickSlots; ++k) {
            TickSlot& sl = ring[k];
            (void)hipHostFree(sl.host);
            (void)hipFree(sl.dev);
            if (sl.ev) (void)hipEventDestroy(sl.ev);
        }
    if (sv->init_ev) (void)hipEventDestroy(sv->init_ev);
    for (hipEvent_t e : sv->free_events) (void)hipEventDestroy(e);
    sv->graphs.clear();
    sv->pool.trim(0);
    for (hipStream_t s : {sv->stream, sv->lstream, sv->up_stream, sv->dl_stream, sv->add_stream})
        if (s) (void)hipStreamDestroy(s);
    delete sv;
    return GLP_OK;
}
