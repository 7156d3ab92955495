// Sessions: device-resident LP state of one job, and the C ABI around them.
//
// A session owns one HIP stream and double-buffered labels.  A step swaps
// the read and write buffers instead of copying (the reference copies
// d_labels_write into d_labels, prop/classicLprop.cuh:371) and never clears
// a hash table with a memset (the reference clears 2*4*offsets[big] bytes
// per step, :377-380): the hub tables are cleared by the pass that reads them.

#include "device.h"
#include "device_prims.h"
#include "session.h"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace glp {

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

DeviceGraph::~DeviceGraph() {
    if (owns_csr) {
        (void)hipFree(off);
        (void)hipFree(nbr);
        (void)hipFree(w);
    }
    if (!owns_plan) return;
    (void)hipFree(d_groups);
    (void)hipFree(d_half);
    (void)hipFree(d_wave);
    for (int c = 0; c < kMaxBlockClasses; ++c) (void)hipFree(d_block[c]);
    (void)hipFree(d_hubs);
    (void)hipFree(d_hub_chunks);
    (void)hipFree(d_hub_chunk0);
    (void)hipFree(d_hub_base);
    (void)hipFree(d_hub_parts);
    (void)hipFree(d_hub_hist_base);
    (void)hipFree(inv.off);
    (void)hipFree(inv.pos);
}

template <class T>
static int upload(T** dst, const T* src, size_t count, hipStream_t s, size_t& bytes) {
    if (count == 0) return GLP_OK;
    GLP_HIP(hipMalloc((void**)dst, sizeof(T) * count));
    bytes += sizeof(T) * count;
    GLP_HIP(hipMemcpyAsync(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
    return GLP_OK;
}

size_t device_graph_bytes(const HostCSR& g, const Plan& p) {
    size_t b = sizeof(int32_t) * ((size_t)g.n + 1) + sizeof(int32_t) * (size_t)g.m;
    if (g.weighted()) b += sizeof(uint32_t) * (size_t)g.m;
    b += sizeof(int2) * p.groups.size() + sizeof(int4) * (p.half.size() + p.wave.size());
    for (int c = 0; c < p.nclasses; ++c) b += sizeof(int32_t) * p.block[c].size();
    b += sizeof(int32_t) * p.hubs.size() + sizeof(int2) * p.hub_chunks.size();
    b += (sizeof(int64_t) * 2 + sizeof(uint32_t)) * p.hubs.size();
    return b;
}

int device_graph_create(const HostCSR& g, const Tuning& t, hipStream_t s, DeviceGraph** out) {
    *out = nullptr;
    double t0 = now_ms();
    Plan plan;
    int rc = build_plan(g, t, plan);
    if (rc != GLP_OK) return rc;
    double t1 = now_ms();
    rc = device_graph_upload(g, std::move(plan), s, out);
    if (rc == GLP_OK) (*out)->plan_ms = t1 - t0;
    return rc;
}

int device_graph_upload(const HostCSR& g, Plan&& plan, hipStream_t s, DeviceGraph** out) {
    *out = nullptr;
    auto dg = std::make_unique<DeviceGraph>();
    dg->n = g.n;
    dg->m = g.m;
    dg->plan = std::move(plan);
    int rc;
    double t1 = now_ms();
    size_t bytes = 0;
    if ((rc = upload(&dg->off, g.offsets.data(), (size_t)g.n + 1, s, bytes))) return rc;
    if ((rc = upload(&dg->nbr, g.neighbors.data(), (size_t)g.m, s, bytes))) return rc;
    if (g.weighted() && (rc = upload(&dg->w, g.weights.data(), (size_t)g.m, s, bytes))) return rc;
    if ((rc = device_plan_upload(dg.get(), s))) return rc;
    double t2 = now_ms();
    dg->upload_ms = t2 - t1;
    dg->bytes += bytes;
    *out = dg.release();
    return GLP_OK;
}

int device_plan_upload(DeviceGraph* dg, hipStream_t s) {
    int rc;
    size_t bytes = 0;
    const Plan& p = dg->plan;
    if ((rc = upload(&dg->d_groups, p.groups.data(), p.groups.size(), s, bytes))) return rc;
    if ((rc = upload(&dg->d_half, p.half.data(), p.half.size(), s, bytes))) return rc;
    if ((rc = upload(&dg->d_wave, p.wave.data(), p.wave.size(), s, bytes))) return rc;
    for (int c = 0; c < p.nclasses; ++c)
        if ((rc = upload(&dg->d_block[c], p.block[c].data(), p.block[c].size(), s, bytes))) return rc;
    if ((rc = upload(&dg->d_hubs, p.hubs.data(), p.hubs.size(), s, bytes))) return rc;
    if ((rc = upload(&dg->d_hub_chunks, p.hub_chunks.data(), p.hub_chunks.size(), s, bytes))) return rc;
    if ((rc = upload(&dg->d_hub_chunk0, p.hub_chunk0.data(), p.hub_chunk0.size(), s, bytes))) return rc;
    if ((rc = upload(&dg->d_hub_base, p.hub_base.data(), p.hub_base.size(), s, bytes))) return rc;
    if ((rc = upload(&dg->d_hub_parts, p.hub_parts.data(), p.hub_parts.size(), s, bytes))) return rc;
    if ((rc = upload(&dg->d_hub_hist_base, p.hub_hist_base.data(), p.hub_hist_base.size(), s, bytes)))
        return rc;
    GLP_HIP(hipStreamSynchronize(s));
    dg->bytes += bytes;
    return GLP_OK;
}

int ensure_device() {
    static int ok = -1;
    if (ok == 1) return GLP_OK;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count < 1) return set_error(GLP_E_NODEVICE, "no HIP device visible");
    ok = 1;
    return GLP_OK;
}

static Method method_of(int32_t m) {
    switch (m) {
        case GLP_LAYERED: return Method::Layered;
        case GLP_SLPA: return Method::Slpa;
        case GLP_USER: return Method::User;
        case GLP_SEEDED: return Method::Seeded;
        default: return Method::Classic;
    }
}

// Methods run by the sketch / table counting kernels (counting.hip).
static bool counting_method(int32_t m) {
    return m == GLP_COUNTMIN || m == GLP_FREQUENT || m == GLP_BUFFER;
}

int check_params(const HostCSR& g, const glp_params* p) {
    uint32_t wmax = 1;
    if (p && p->weighted)
        for (uint32_t w : g.weights) wmax = std::max(wmax, w);
    return check_params_shape(g.n, g.max_degree, wmax, g.weighted(), p);
}

int check_params_shape(int32_t n, int32_t max_degree, uint32_t wmax_in, bool weighted, const glp_params* p) {
    if (!p) return set_error(GLP_E_ARG, "null params");
    for (int i = 0; i < 5; ++i)
        if (p->reserved[i]) return set_error(GLP_E_ARG, "reserved params fields must be zero");
    if (p->method < GLP_CLASSIC || p->method > GLP_SEEDED)
        return set_error(GLP_E_ARG, "unknown method %d", p->method);
    if (p->frontier < 0 || p->frontier > 2)
        return set_error(GLP_E_ARG, "frontier must be 0, 1 or 2 (2: with the gather's two-step lookahead)");
    if (p->frontier && p->method != GLP_CLASSIC)
        return set_error(GLP_E_UNSUPPORTED,
                         "the frontier is exact for classic LP only (layered volumes and SLPA speakers "
                         "change every label's score each step)");
    if (p->weighted && !weighted) return set_error(GLP_E_ARG, "weighted run on an unweighted graph");
    // frequencies are packed as non-negative int32 scores
    const uint64_t wmax = p->weighted ? std::max<uint64_t>(wmax_in, 1) : 1;
    const uint64_t max_freq = wmax * (uint64_t)std::max(max_degree, 1);
    if (max_freq > (uint64_t)INT32_MAX)
        return set_error(GLP_E_RANGE, "largest frequency sum %llu exceeds int32", (unsigned long long)max_freq);
    if (p->method == GLP_LAYERED) {
        const int64_t q = p->gamma_den, pn = p->gamma_num;
        if (q < 1 || (q & (q - 1)) || pn < 0)
            return set_error(GLP_E_ARG, "layered gamma must be p/q with q a power of two and p >= 0");
        // score (q+p)*C - p*V must stay inside int32
        if ((q + pn) * (int64_t)max_freq > INT32_MAX || pn * (int64_t)n > INT32_MAX)
            return set_error(GLP_E_RANGE, "layered scores exceed int32 for gamma=%lld/%lld", (long long)pn, (long long)q);
    }
    if (p->method == GLP_SLPA) {
        if (p->slpa_T < 1 || p->slpa_T > 255) return set_error(GLP_E_ARG, "slpa_T must be in [1,255]");
        if (p->slpa_r_den < 1 || p->slpa_r_num < 0) return set_error(GLP_E_ARG, "bad SLPA threshold");
    }
    if (p->method == GLP_COUNTMIN) {
        if (p->cm_rows < 1 || p->cm_rows > 5 || p->cm_width < 1 || p->cm_width > 4096)
            return set_error(GLP_E_ARG, "count-min needs 1<=rows<=5, 1<=width<=4096");
        if (p->cm_aliased && p->cm_rows < 3) return set_error(GLP_E_ARG, "aliased count-min needs rows >= 3");
    }
    if (p->method == GLP_FREQUENT || p->method == GLP_BUFFER) {
        if (p->table_cap < 1 || p->table_cap > 64) return set_error(GLP_E_ARG, "table_cap must be in [1,64]");
    }
    if (p->gather && p->method != GLP_CLASSIC && p->method != GLP_LAYERED && p->method != GLP_USER &&
        p->method != GLP_SEEDED)
        return set_error(GLP_E_UNSUPPORTED,
                         "separate gather is implemented for classic, layered, user and seeded LP");
    return GLP_OK;
}

}  // namespace glp

using namespace glp;

glp_session::~glp_session() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& ge : graph_exec)
        if (ge) (void)hipGraphExecDestroy(ge);
    for (int i = 0; i < 2; ++i) {
        (void)hipFree(L[i]);
        (void)hipFree(V[i]);
    }
    (void)hipFree(LV);
    (void)hipFree(shown);
    (void)hipFree(adj);
    (void)hipFree(adj_alt);
    (void)hipFree(mem);
    (void)hipFree(hs.pairs);
    (void)hipFree(hs.hist);
    (void)hipFree(hs.lstart);
    (void)hipFree(hs.changed);
    (void)hipFree(act[0]);
    (void)hipFree(act[1]);
    (void)hipFree(chg[0]);
    (void)hipFree(chg[1]);
    (void)hipFree(push);
    (void)hipFree(seeds);
    (void)hipFree(fctl);
    (void)hipFree(hs.best);
    (void)hipFree(hs.work);
    (void)hipFree(d_acc);
    (void)hipFree(d_ring);
    (void)hipFree(d_ctr);
    (void)hipFree(wide);
    (void)hipFree(cm_chunks);
    (void)hipFree(cm_sketch);
    (void)hipFree(cm_best);
    (void)hipFree(cm_keys);
    for (auto& set : ev_pool)
        for (auto& e : set)
            if (e) (void)hipEventDestroy(e);
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace glp {

size_t session_bytes(const HostCSR& g, const Plan& p, const glp_params& prm) {
    size_t n = (size_t)g.n;
    size_t b = 2 * n * 4;
    if (prm.method == GLP_LAYERED) b += 2 * n * 4;
    if (prm.method == GLP_LAYERED && !prm.gather) b += n * 8;  // LV records
    if (prm.method == GLP_SLPA) b += n * 4 + (size_t)(prm.slpa_T + 1) * n * 4;
    if (prm.method == GLP_USER) b += n * 4;
    if (prm.gather) b += (size_t)g.m * 4;
    if (prm.frontier) b += 2 * (n + 4);
    if (prm.frontier && prm.gather) b += 2 * ((n + 31) / 32) * 4 + push_entry_bytes(g.n, g.m);
    if (prm.frontier == 2 && prm.gather) b += (size_t)g.m * 4;  // the lookahead's second adj
    b += hub_scratch_bytes(p, prm.frontier != 0);
    b += 8 * kChangeSlots + 8 * kChangeRing + 4 + 4;  // change counters and ring, step and unit counters
    return b;
}

int session_create(glp_graph* g, const glp_params* prm, glp_session** out) {
    int rc = ensure_device();
    if (rc != GLP_OK) return rc;
    if ((rc = check_params(g->csr, prm)) != GLP_OK) return rc;
    if (prm->frontier && !csr_is_symmetric(g->csr))
        return set_error(GLP_E_UNSUPPORTED,
                         "the frontier needs a symmetric graph (every edge u->v with an edge v->u): "
                         "it marks a changed vertex's out-neighbours for recomputation");
    init_kernels();
    auto s = std::make_unique<glp_session>();
    s->g = g;
    s->p = *prm;
    s->method = method_of(prm->method);
    s->n = g->csr.n;
    GLP_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    if (!g->dev) {
        DeviceGraph* dg = nullptr;
        if ((rc = device_graph_create(g->csr, default_tuning(), s->stream, &dg)) != GLP_OK) return rc;
        g->dev = dg;
        s->timing.upload_ms = dg->upload_ms;
        s->timing.plan_ms = dg->plan_ms;
    }
    const DeviceGraph& dg = *g->dev;
    const Plan& p = dg.plan;
    const size_t n = (size_t)s->n;
    const size_t nn = std::max<size_t>(n, 1);
    GLP_HIP(hipMalloc(&s->L[0], nn * 4));
    GLP_HIP(hipMalloc(&s->L[1], nn * 4));
    if (prm->method == GLP_LAYERED) {
        GLP_HIP(hipMalloc(&s->V[0], nn * 4));
        GLP_HIP(hipMalloc(&s->V[1], nn * 4));
        if (!prm->gather) GLP_HIP(hipMalloc(&s->LV, nn * 8));
    }
    if (prm->method == GLP_SLPA || prm->method == GLP_USER) GLP_HIP(hipMalloc(&s->shown, nn * 4));
    if (prm->method == GLP_SLPA)
        GLP_HIP(hipMalloc(&s->mem, (size_t)(prm->slpa_T + 1) * nn * 4));
    if (prm->gather) GLP_HIP(hipMalloc(&s->adj, std::max<size_t>((size_t)g->csr.m, 1) * 4));
    if (prm->gather && prm->frontier == 2)
        GLP_HIP(hipMalloc(&s->adj_alt, std::max<size_t>((size_t)g->csr.m, 1) * 4));
    if (!p.hubs.empty()) {
        GLP_HIP(hipMalloc(&s->hs.pairs, (size_t)p.pair_slots * 8));
        GLP_HIP(hipMalloc(&s->hs.hist, (size_t)p.hist_slots * 4));
        GLP_HIP(hipMalloc(&s->hs.lstart, (size_t)p.hist_slots * 4));
        GLP_HIP(hipMalloc(&s->hs.best, p.hubs.size() * 8));
        GLP_HIP(hipMemsetAsync(s->hs.best, 0, p.hubs.size() * 8, s->stream));
        if (prm->frontier) GLP_HIP(hipMalloc(&s->hs.changed, p.hubs.size() * 4));
    }
    if (prm->frontier) {
        s->act_bytes = ((int64_t)nn + 4) & ~(int64_t)3;
        GLP_HIP(hipMalloc(&s->act[0], (size_t)s->act_bytes));
        GLP_HIP(hipMalloc(&s->act[1], (size_t)s->act_bytes));
        GLP_HIP(hipMalloc(&s->fctl, 8 * sizeof(int32_t)));
        if (prm->gather) {
            s->chg_bytes = (((int64_t)nn + 31) / 32) * 4;
            GLP_HIP(hipMalloc(&s->chg[0], (size_t)s->chg_bytes));
            GLP_HIP(hipMalloc(&s->chg[1], (size_t)s->chg_bytes));
            GLP_HIP(hipMalloc(&s->push, push_entry_bytes(s->n, g->csr.m)));
            DeviceGraph* dgm = g->dev;
            if (!dgm->inv.off) {
                // the push list doubles as the build's cursor scratch
                const size_t m = std::max<size_t>((size_t)g->csr.m, 1);
                GLP_HIP(hipMalloc(&dgm->inv.off, (nn + 1) * 4));
                GLP_HIP(hipMalloc(&dgm->inv.pos, m * 4));
                if ((rc = inverse_index_build(dgm->nbr, s->n, g->csr.m, static_cast<int32_t*>(s->push),
                                              dgm->inv, s->stream)) != GLP_OK)
                    return rc;
                dgm->bytes += (nn + 1) * 4 + m * 4;
            }
        }
    }
    GLP_HIP(hipMalloc(&s->hs.work, 4));  // the persistent small-tier launch's counter
    GLP_HIP(hipMalloc(&s->d_acc, 8 * kChangeSlots));
    GLP_HIP(hipMalloc(&s->d_ring, 8 * kChangeRing));
    GLP_HIP(hipMalloc(&s->d_ctr, 4));
    GLP_HIP(hipEventCreate(&s->ev_start));
    GLP_HIP(hipEventCreate(&s->ev_stop));
    s->bytes = session_bytes(g->csr, p, *prm);
    if (counting_method(prm->method) && (rc = counting_wide_list(s.get())) != GLP_OK) return rc;
    for (int t = 0; t < 4; ++t) {
        s->timing.tier_vertices[t] = p.verts[t];
        s->timing.tier_edges[t] = p.edges[t];
    }
    if ((rc = session_reset(s.get())) != GLP_OK) return rc;
    *out = s.release();
    return GLP_OK;
}

int session_reset(glp_session* s) {
    const int64_t n = s->n;
    s->cur = 0;
    s->t = 0;
    s->timing.steps = 0;
    s->timing.steps_ms = 0;
    for (int i = 0; i < GLP_NTIERS; ++i) {
        s->timing.tier_ms[i] = 0;
        s->timing.tier_launches[i] = 0;
    }
    s->host_changes.clear();
    int rc;
    // L_0 in both buffers (RS: both label buffers start equal): v, or for
    // seeded LP v on seeds and -1 elsewhere (all -1 until seeds are set)
    for (int b = 0; b < 2; ++b) {
        if (s->method != Method::Seeded) rc = launch_iota(s->L[b], n, s->stream);
        else if (s->seeds_set) rc = launch_seed_labels(s->L[b], s->seeds, n, s->stream);
        else rc = launch_fill_i32(s->L[b], -1, n, s->stream);
        if (rc) return rc;
    }
    if (s->V[0]) {
        // V_0(l) = #{u : L_0[u] = l} = 1
        if ((rc = launch_fill_i32(s->V[0], 1, n, s->stream))) return rc;
        if ((rc = launch_fill_i32(s->V[1], 0, n, s->stream))) return rc;
    }
    if (s->mem && (rc = launch_iota(s->mem, n, s->stream))) return rc;  // M_v[0] = v
    if (s->fctl) {
        // the first step recomputes every vertex and does not mark; marking
        // starts after a step that changed at most n/32 labels
        const int32_t ctl[8] = {1, 0, std::max<int32_t>(1, (int32_t)(n / 32)), 0, 0, 0, 0,
                                push_edge_limit(s->g->csr.m)};
        GLP_HIP(hipMemcpyAsync(s->fctl, ctl, sizeof ctl, hipMemcpyHostToDevice, s->stream));
        if ((rc = launch_zero_bytes(s->act[0], s->act_bytes, s->stream))) return rc;
        if ((rc = launch_zero_bytes(s->act[1], s->act_bytes, s->stream))) return rc;
    }
    for (uint32_t* c : s->chg)
        if (c && (rc = launch_zero_bytes(c, s->chg_bytes, s->stream))) return rc;
    GLP_HIP(hipMemsetAsync(s->d_acc, 0, 8 * kChangeSlots, s->stream));
    GLP_HIP(hipMemsetAsync(s->d_ring, 0, 8 * kChangeRing, s->stream));
    GLP_HIP(hipMemsetAsync(s->d_ctr, 0, 4, s->stream));
    GLP_HIP(hipStreamSynchronize(s->stream));
    return GLP_OK;
}

// Builds the argument block of the next step (step number s->t + 1).
static StepArgs step_args(glp_session* s) {
    const DeviceGraph& dg = *s->g->dev;
    StepArgs a;
    std::memset(&a, 0, sizeof a);
    a.off = dg.off;
    a.nbr = dg.nbr;
    a.w = dg.w;
    a.n = s->n;
    a.t = s->t + 1;
    a.changes = s->d_acc;
    a.seed = s->p.seed;
    const int c = s->cur;
    switch (s->method) {
        case Method::Classic:
            a.L = s->L[c];
            a.keep = s->L[c];
            a.Lw = s->L[c ^ 1];
            break;
        case Method::Layered:
            a.L = s->L[c];
            a.keep = s->L[c];
            a.Lw = s->L[c ^ 1];
            a.Vr = s->V[c];
            a.Vw = s->V[c ^ 1];
            a.score_mul = s->p.gamma_den + s->p.gamma_num;
            a.score_vmul = s->p.gamma_num;
            break;
        case Method::Seeded:
            a.L = s->L[c];
            a.keep = s->L[c];
            a.Lw = s->L[c ^ 1];
            a.seeds = s->seeds;
            break;
        case Method::User:
            a.L = s->shown;
            a.keep = s->L[c];
            a.Lw = s->L[c ^ 1];
            break;
        case Method::Slpa:
            a.mem = s->mem;
            a.L = s->shown;
            a.keep = s->mem + (size_t)s->t * s->n;
            a.Lw = s->mem + (size_t)(s->t + 1) * s->n;
            break;
    }
    // the lookahead's two adj buffers alternate with the label buffers, so
    // each captured step graph keeps its own pair
    a.adj = s->adj_alt && c ? s->adj_alt : s->adj;
    if (s->fctl) {
        a.act_r = s->act[c];
        a.act_w = s->act[c ^ 1];
        a.fctl = s->fctl;
    }
    if (s->chg[0]) {
        a.chg_r = s->chg[c];
        a.chg_w = s->chg[c ^ 1];
        a.lookahead = s->adj_alt ? 1 : 0;
    }
    return a;
}

// Enqueues one step.  ev: per-tier events or null.
int enqueue_step(glp_session* s, hipEvent_t* ev) {
    if (counting_method(s->p.method)) return enqueue_counting_step(s);
    if (s->method == Method::Seeded && !s->seeds_set)
        return set_error(GLP_E_STATE, "a seeded session needs its seeds (glp_session_set_seeds)");
    if (s->method == Method::Slpa && s->t >= s->p.slpa_T)
        return set_error(GLP_E_STATE, "SLPA memory holds %d steps", s->p.slpa_T);
    StepArgs a = step_args(s);
    int rc;
    if (s->method == Method::Layered && (rc = launch_fill_i32(s->V[s->cur ^ 1], 0, s->n, s->stream)))
        return rc;
    // layered with gamma > 0: each neighbour's label travels with its
    // label's volume, and the tables accumulate scores (LayeredVolHooks)
    Method method = s->method;
    if (s->method == Method::Layered && s->LV && s->p.gamma_num > 0) {
        method = Method::LayeredVol;
        a.LV = s->LV;
        if ((rc = launch_label_volumes(a.L, a.Vr, s->LV, s->n, s->stream))) return rc;
    }
    // frontier flags for the next step start clear (they held the last step's)
    if (s->fctl && (rc = launch_zero_bytes(s->act[s->cur ^ 1], s->act_bytes, s->stream))) return rc;
    // this step's changed-vertex bits start clear (they held two steps back)
    if (s->chg[0] && (rc = launch_zero_bytes(s->chg[s->cur ^ 1], s->chg_bytes, s->stream))) return rc;
    if (ev) (void)hipEventRecord(ev[2 * GLP_TIER_PICK], s->stream);
    if (s->method == Method::Slpa || s->method == Method::User) {
        if ((rc = launch_pick(s->method, a, s->shown, s->stream))) return rc;
    }
    if (ev) (void)hipEventRecord(ev[2 * GLP_TIER_PICK + 1], s->stream);
    if (ev) (void)hipEventRecord(ev[2 * GLP_TIER_GATHER], s->stream);
    if (s->p.gather) {
        int32_t* adj = const_cast<int32_t*>(a.adj);
        if (s->chg[0]) rc = launch_frontier_gather(a, s->g->csr.m, adj, s->g->dev->inv, s->push, s->stream);
        else rc = launch_gather(a, s->g->csr.m, adj, s->stream);
        if (rc) return rc;
    }
    if (ev) (void)hipEventRecord(ev[2 * GLP_TIER_GATHER + 1], s->stream);
    if ((rc = launch_step(method, s->p.weighted != 0, s->p.gather != 0, *s->g->dev, s->hs, a, s->stream, ev)))
        return rc;
    if (s->method == Method::Layered && (rc = launch_volume(a.Lw, a.Vw, s->n, s->stream))) return rc;
    if ((rc = launch_end_step(s->d_acc, s->d_ring, s->d_ctr, kChangeRing, s->stream, s->fctl))) return rc;
    if (s->method != Method::Slpa) s->cur ^= 1;
    s->t += 1;
    return GLP_OK;
}

// Adds the per-tier event spans of every profiled step since the last sync.
static void accumulate_profile(glp_session* s) {
    static const int tiers[] = {GLP_TIER_PACKED, GLP_TIER_WAVE, GLP_TIER_BLOCK, GLP_TIER_HUB,
                                GLP_TIER_GATHER, GLP_TIER_PICK, GLP_TIER_AUX};
    for (size_t k = 0; k < s->ev_used; ++k) {
        hipEvent_t* ev = s->ev_pool[k].data();
        for (int t : tiers) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, ev[2 * t], ev[2 * t + 1]) == hipSuccess) {
                s->timing.tier_ms[t] += ms;
                s->timing.tier_launches[t] += 1;
            }
        }
    }
    (void)hipGetLastError();
    s->ev_used = 0;
}

int session_step(glp_session* s, int32_t steps) {
    if (steps < 0) return set_error(GLP_E_ARG, "negative step count");
    if (steps == 0) return GLP_OK;
    if (s->pending_steps == 0) GLP_HIP(hipEventRecord(s->ev_start, s->stream));
    for (int32_t i = 0; i < steps; ++i) {
        hipEvent_t* ev = nullptr;
        if (s->p.profile) {
            // one event set per step, recorded between launches without host syncs
            if (s->ev_used == s->ev_pool.size()) {
                std::array<hipEvent_t, 2 * GLP_NTIERS> set{};
                for (auto& e : set) GLP_HIP(hipEventCreate(&e));
                s->ev_pool.push_back(set);
            }
            ev = s->ev_pool[s->ev_used++].data();
        }
        int rc = enqueue_step(s, ev);
        if (rc != GLP_OK) return rc;
    }
    GLP_HIP(hipEventRecord(s->ev_stop, s->stream));
    s->pending_steps += steps;
    return GLP_OK;
}

int session_sync(glp_session* s) {
    GLP_HIP(hipStreamSynchronize(s->stream));
    if (s->pending_steps) {
        float ms = 0;
        GLP_HIP(hipEventElapsedTime(&ms, s->ev_start, s->ev_stop));
        s->timing.steps_ms += ms;
        s->timing.steps += s->pending_steps;
        s->pending_steps = 0;
    }
    accumulate_profile(s);
    return GLP_OK;
}

int session_step_graph(glp_session* s, int32_t steps) {
    if (s->method == Method::Slpa || counting_method(s->p.method))
        return session_step(s, steps);  // step-dependent pointers: plain launches
    if (steps <= 0) return steps == 0 ? GLP_OK : set_error(GLP_E_ARG, "negative step count");
    // one captured graph per parity of the label buffers
    for (int parity = 0; parity < 2; ++parity) {
        if (s->graph_exec[parity]) continue;
        const int save_cur = s->cur, save_t = s->t;
        s->cur = parity;
        GLP_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_step(s, nullptr);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s->stream, &graph);
        s->cur = save_cur;
        s->t = save_t;
        if (rc != GLP_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return set_error(GLP_E_HIP, "capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&s->graph_exec[parity], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return set_error(GLP_E_HIP, "instantiate: %s", hipGetErrorString(e));
    }
    if (s->pending_steps == 0) GLP_HIP(hipEventRecord(s->ev_start, s->stream));
    for (int32_t i = 0; i < steps; ++i) {
        GLP_HIP(hipGraphLaunch(s->graph_exec[s->cur], s->stream));
        s->cur ^= 1;
        s->t += 1;
    }
    GLP_HIP(hipEventRecord(s->ev_stop, s->stream));
    s->pending_steps += steps;
    return GLP_OK;
}

int session_labels(glp_session* s, int32_t* out) {
    int rc = session_sync(s);
    if (rc != GLP_OK) return rc;
    const int32_t* src = s->method == Method::Slpa ? s->mem + (size_t)s->t * s->n : s->L[s->cur];
    if (s->n) GLP_HIP(hipMemcpy(out, src, sizeof(int32_t) * s->n, hipMemcpyDeviceToHost));
    return GLP_OK;
}

int session_changes(glp_session* s, int64_t* out, int32_t cap, int32_t* nsteps) {
    int rc = session_sync(s);
    if (rc != GLP_OK) return rc;
    std::vector<unsigned long long> ring(kChangeRing);
    GLP_HIP(hipMemcpy(ring.data(), s->d_ring, 8 * kChangeRing, hipMemcpyDeviceToHost));
    const int32_t t = s->t;
    const int32_t first = std::max(0, t - kChangeRing);
    int32_t k = 0;
    for (int32_t i = first; i < t && k < cap; ++i) out[k++] = (int64_t)ring[i % kChangeRing];
    if (nsteps) *nsteps = k;
    return GLP_OK;
}

}  // namespace glp

// ------------------------------------------------------------------ C ABI ----

#define GLP_GUARD(ptr)                                                   \
    do {                                                                 \
        clear_error();                                                   \
        if (!(ptr)) return set_error(GLP_E_ARG, "null handle");         \
    } while (0)

extern "C" void glp_params_default(glp_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->method = GLP_CLASSIC;
    p->gamma_num = 0;
    p->gamma_den = 1;
    p->seed = 1;
    p->slpa_T = 20;
    p->slpa_r_num = 1;
    p->slpa_r_den = 10;
    p->cm_rows = 3;
    p->cm_width = 200;
    p->table_cap = 16;
}

extern "C" int glp_device_count(int32_t* count) {
    clear_error();
    if (!count) return set_error(GLP_E_ARG, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *count = e == hipSuccess ? c : 0;
    return GLP_OK;
}

extern "C" int glp_calib_random_read(int64_t table_entries, int64_t reads, int32_t repeats,
                                     double* ms_best) {
    clear_error();
    if (!ms_best || reads <= 0 || repeats <= 0) return set_error(GLP_E_ARG, "bad calibration arguments");
    if (table_entries <= 0 || (table_entries & (table_entries - 1)) || table_entries > (int64_t)1 << 31)
        return set_error(GLP_E_ARG, "table entries must be a power of two <= 2^31");
    int rc = ensure_device();
    if (rc != GLP_OK) return rc;
    const int64_t out_entries = 256 * 32 * 256;
    int32_t* T = nullptr;
    int32_t* out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    float best = 0.f;
    auto fail = [&](hipError_t e) { return set_error(GLP_E_HIP, "calibration: %s", hipGetErrorString(e)); };
    hipError_t e = hipMalloc(&T, (size_t)std::max<int64_t>(table_entries, 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&out, (size_t)out_entries * 4);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e != hipSuccess) rc = fail(e);
    if (!rc) rc = launch_fill_i32(T, 1, table_entries, st);
    // one untimed launch, then the best of `repeats`
    for (int32_t r = 0; !rc && r <= repeats; ++r) {
        if ((e = hipEventRecord(a, st)) != hipSuccess) { rc = fail(e); break; }
        if ((rc = launch_calib_read(T, table_entries, reads, 11u + (uint32_t)r, out, out_entries, st))) break;
        if ((e = hipEventRecord(b, st)) != hipSuccess || (e = hipEventSynchronize(b)) != hipSuccess) {
            rc = fail(e);
            break;
        }
        float ms = 0.f;
        if ((e = hipEventElapsedTime(&ms, a, b)) != hipSuccess) { rc = fail(e); break; }
        if (r > 0 && (best == 0.f || ms < best)) best = ms;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(T);
    (void)hipFree(out);
    if (!rc) *ms_best = best;
    return rc;
}

extern "C" int glp_session_create(glp_graph* g, const glp_params* p, glp_session** out) {
    clear_error();
    if (!g || !p || !out) return set_error(GLP_E_ARG, "null argument");
    *out = nullptr;
    return session_create(g, p, out);
}
extern "C" int glp_session_reset(glp_session* s) {
    GLP_GUARD(s);
    return session_reset(s);
}
extern "C" int glp_session_step(glp_session* s, int32_t steps) {
    GLP_GUARD(s);
    return session_step(s, steps);
}
extern "C" int glp_session_step_graph(glp_session* s, int32_t steps) {
    GLP_GUARD(s);
    return session_step_graph(s, steps);
}
extern "C" int glp_session_set_gamma(glp_session* s, int32_t gamma_num, int32_t gamma_den) {
    GLP_GUARD(s);
    if (s->method != Method::Layered) return set_error(GLP_E_STATE, "gamma exists for layered sessions only");
    glp_params q = s->p;
    q.gamma_num = gamma_num;
    q.gamma_den = gamma_den;
    int rc = check_params(s->g->csr, &q);
    if (rc != GLP_OK) return rc;
    s->p = q;
    // captured step graphs baked the old scores in
    for (auto& ge : s->graph_exec)
        if (ge) {
            (void)hipGraphExecDestroy(ge);
            ge = nullptr;
        }
    return GLP_OK;
}

extern "C" int glp_session_sync(glp_session* s) {
    GLP_GUARD(s);
    return session_sync(s);
}
extern "C" int glp_session_labels(glp_session* s, int32_t* out) {
    GLP_GUARD(s);
    if (!out && s->n) return set_error(GLP_E_ARG, "null output");
    return session_labels(s, out);
}
extern "C" int glp_session_changes(glp_session* s, int64_t* out, int32_t cap, int32_t* nsteps) {
    GLP_GUARD(s);
    if (!out || cap < 0) return set_error(GLP_E_ARG, "bad output");
    return session_changes(s, out, cap, nsteps);
}
extern "C" int glp_session_timing(glp_session* s, glp_timing* out) {
    GLP_GUARD(s);
    if (!out) return set_error(GLP_E_ARG, "null output");
    int rc = session_sync(s);
    if (rc != GLP_OK) return rc;
    *out = s->timing;
    return GLP_OK;
}
extern "C" int glp_session_volume(glp_session* s, int32_t* out) {
    GLP_GUARD(s);
    if (!s->V[0]) return set_error(GLP_E_STATE, "volume exists for layered sessions only");
    int rc = session_sync(s);
    if (rc != GLP_OK) return rc;
    if (s->n) GLP_HIP(hipMemcpy(out, s->V[s->cur], sizeof(int32_t) * s->n, hipMemcpyDeviceToHost));
    return GLP_OK;
}
extern "C" int glp_session_slpa_memory(glp_session* s, int32_t* out) {
    GLP_GUARD(s);
    if (!s->mem) return set_error(GLP_E_STATE, "memory exists for SLPA sessions only");
    int rc = session_sync(s);
    if (rc != GLP_OK) return rc;
    GLP_HIP(hipMemcpy(out, s->mem, sizeof(int32_t) * (size_t)(s->t + 1) * s->n, hipMemcpyDeviceToHost));
    return GLP_OK;
}

struct StateHeader {
    uint32_t magic;
    int32_t method;
    int32_t n;
    int32_t t;
    int32_t cur_unused;
    int32_t words;  // int32 words that follow
};
static const uint32_t kStateMagic = 0x474c5053u;  // "GLPS"

extern "C" int glp_session_save(glp_session* s, void* state, size_t* state_bytes) {
    GLP_GUARD(s);
    if (!state_bytes) return set_error(GLP_E_ARG, "null size");
    int rc = session_sync(s);
    if (rc != GLP_OK) return rc;
    const size_t n = (size_t)s->n;
    size_t words = 0;
    if (s->mem) words = (size_t)(s->t + 1) * n;
    else words = n + (s->V[0] ? n : 0);
    const size_t need = sizeof(StateHeader) + words * 4;
    if (!state) {
        *state_bytes = need;
        return GLP_OK;
    }
    if (*state_bytes < need) return set_error(GLP_E_ARG, "state buffer too small (%zu < %zu)", *state_bytes, need);
    StateHeader h{kStateMagic, s->p.method, s->n, s->t, 0, (int32_t)words};
    std::memcpy(state, &h, sizeof h);
    int32_t* body = reinterpret_cast<int32_t*>(static_cast<char*>(state) + sizeof h);
    if (s->mem) {
        GLP_HIP(hipMemcpy(body, s->mem, words * 4, hipMemcpyDeviceToHost));
    } else {
        GLP_HIP(hipMemcpy(body, s->L[s->cur], n * 4, hipMemcpyDeviceToHost));
        if (s->V[0]) GLP_HIP(hipMemcpy(body + n, s->V[s->cur], n * 4, hipMemcpyDeviceToHost));
    }
    *state_bytes = need;
    return GLP_OK;
}

extern "C" int glp_session_load(glp_session* s, const void* state, size_t state_bytes) {
    GLP_GUARD(s);
    if (!state || state_bytes < sizeof(StateHeader)) return set_error(GLP_E_ARG, "bad state");
    StateHeader h;
    std::memcpy(&h, state, sizeof h);
    if (h.magic != kStateMagic || h.method != s->p.method || h.n != s->n)
        return set_error(GLP_E_FORMAT, "state does not belong to this session's graph/method");
    if (state_bytes < sizeof h + (size_t)h.words * 4) return set_error(GLP_E_FORMAT, "truncated state");
    if (s->mem && h.t > s->p.slpa_T) return set_error(GLP_E_FORMAT, "state beyond SLPA memory");
    int rc = session_reset(s);
    if (rc != GLP_OK) return rc;
    const int32_t* body = reinterpret_cast<const int32_t*>(static_cast<const char*>(state) + sizeof h);
    const size_t n = (size_t)s->n;
    if (s->mem) {
        GLP_HIP(hipMemcpy(s->mem, body, (size_t)h.words * 4, hipMemcpyHostToDevice));
    } else {
        GLP_HIP(hipMemcpy(s->L[0], body, n * 4, hipMemcpyHostToDevice));
        GLP_HIP(hipMemcpy(s->L[1], body, n * 4, hipMemcpyHostToDevice));
        if (s->V[0]) GLP_HIP(hipMemcpy(s->V[0], body + n, n * 4, hipMemcpyHostToDevice));
    }
    s->cur = 0;
    s->t = h.t;
    GLP_HIP(hipMemcpy(s->d_ctr, &h.t, 4, hipMemcpyHostToDevice));
    return GLP_OK;
}

extern "C" int glp_session_slpa_post(glp_session* s, int32_t kmax, int32_t* dominant, int32_t* nsets,
                                     int32_t* sets) {
    GLP_GUARD(s);
    if (!s->mem) return set_error(GLP_E_STATE, "post-processing exists for SLPA sessions only");
    if (kmax < 0 || (kmax > 0 && !sets)) return set_error(GLP_E_ARG, "bad set output");
    int rc = session_sync(s);
    if (rc != GLP_OK) return rc;
    return slpa_post(s, kmax, dominant, nsets, sets);
}

extern "C" int glp_session_free(glp_session* s) {
    if (!s) return GLP_OK;
    delete s;
    return GLP_OK;
}

static int run_impl(glp_graph* g, const glp_params* p, const uint8_t* seeds, int32_t steps,
                    int32_t* labels_out, int32_t* step_dump, glp_timing* timing) {
    if (!g || !p || steps < 0 || (!labels_out && g->csr.n)) return set_error(GLP_E_ARG, "bad run arguments");
    glp_params q = *p;
    if (q.method == GLP_SLPA && q.slpa_T < steps) q.slpa_T = steps;
    glp_session* s = nullptr;
    int rc = session_create(g, &q, &s);
    if (rc != GLP_OK) return rc;
    std::unique_ptr<glp_session> guard(s);
    if (q.method == GLP_SEEDED && (rc = glp_session_set_seeds(s, seeds, s->n)) != GLP_OK) return rc;
    for (int32_t i = 0; i < steps; ++i) {
        if ((rc = session_step(s, 1)) != GLP_OK) return rc;
        if (step_dump) {
            if ((rc = session_labels(s, step_dump + (size_t)i * s->n)) != GLP_OK) return rc;
        }
    }
    if ((rc = session_sync(s)) != GLP_OK) return rc;
    if (q.method == GLP_SLPA) {
        if ((rc = slpa_post(s, 0, labels_out, nullptr, nullptr)) != GLP_OK) return rc;
    } else if ((rc = session_labels(s, labels_out)) != GLP_OK) {
        return rc;
    }
    if (timing) *timing = s->timing;
    return GLP_OK;
}

extern "C" int glp_run(glp_graph* g, const glp_params* p, int32_t steps, int32_t* labels_out,
                       int32_t* step_dump, glp_timing* timing) {
    clear_error();
    if (p && p->method == GLP_SEEDED) return set_error(GLP_E_ARG, "seeded LP runs through glp_run_seeded");
    return run_impl(g, p, nullptr, steps, labels_out, step_dump, timing);
}

extern "C" int glp_session_set_seeds(glp_session* s, const uint8_t* seeds, int32_t n) {
    clear_error();
    if (!s || (!seeds && n > 0)) return set_error(GLP_E_ARG, "bad seed arguments");
    if (s->method != Method::Seeded) return set_error(GLP_E_STATE, "seeds exist for seeded sessions only");
    if (n != s->n) return set_error(GLP_E_ARG, "seeds hold %d entries, the graph has %d vertices", n, s->n);
    if (!s->seeds) GLP_HIP(hipMalloc(&s->seeds, std::max<size_t>((size_t)n, 1)));
    if (n > 0) GLP_HIP(hipMemcpyAsync(s->seeds, seeds, (size_t)n, hipMemcpyHostToDevice, s->stream));
    s->seeds_set = true;
    return session_reset(s);  // synchronises the stream (the host buffer may go)
}

extern "C" int glp_run_seeded(glp_graph* g, const glp_params* p, const uint8_t* seeds, int32_t steps,
                              int32_t* labels_out, int32_t* step_dump, glp_timing* timing) {
    clear_error();
    if (!g || !p || p->method != GLP_SEEDED) return set_error(GLP_E_ARG, "glp_run_seeded needs method GLP_SEEDED");
    return run_impl(g, p, seeds, steps, labels_out, step_dump, timing);
}

extern "C" void glp_graph_free(glp_graph* g) {
    if (!g) return;
    delete g->dev;
    g->dev = nullptr;
    delete g;
}
