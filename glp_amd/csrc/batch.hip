// Batched LP: many independent graphs stepped together.
//
// The graphs are packed into one disjoint-union CSR (graph g's vertices are
// renumbered base_g + v, neighbour ids likewise), and one session steps the
// union: every tier launch of a step covers all graphs at once, and mode 1
// replays the step from a captured HIP graph.  LP never crosses between
// components, and renumbering adds the same base_g to every label of graph
// g, an order-preserving shift, so argmax ties resolve exactly as they do
// for the graph alone: labels - base_g are bit-identical to a solo run.
// Counting methods whose result depends on label values beyond their order
// (count-min hashes them, SLPA draws from a stream indexed by vertex id) are
// not batched this way and are rejected.
//
// Mode 2 is the comparison form: one plain session per graph, stepped one
// graph after another (the reference's one-graph-per-run shape, main.cu).

#include "device.h"
#include "session.h"

#include <chrono>
#include <cstring>
#include <memory>

struct glp_batch {
    std::vector<glp_graph*> graphs;  // caller-owned
    glp_params p{};
    std::unique_ptr<glp_graph> uni;  // union graph
    glp_session* sess = nullptr;     // union session
    std::vector<int64_t> base;       // vertex offset of each graph in the union
    std::vector<glp_session*> plain; // per-graph sessions (mode 2)
    int last_mode = 0;
    ~glp_batch() {
        if (sess) delete sess;
        for (auto* s : plain) delete s;
        if (uni) {
            delete uni->dev;
            uni->dev = nullptr;
        }
    }
};

using namespace glp;

extern "C" int glp_batch_create(int32_t ngraphs, glp_graph* const* graphs, const glp_params* p,
                                glp_batch** out) {
    clear_error();
    if (!out || !p || ngraphs < 1 || !graphs) return set_error(GLP_E_ARG, "bad batch arguments");
    *out = nullptr;
    if (p->method == GLP_SLPA || p->method == GLP_COUNTMIN)
        return set_error(GLP_E_UNSUPPORTED,
                         "batched SLPA / count-min would change results (stream and hash depend on ids)");
    if (p->method == GLP_SEEDED)
        return set_error(GLP_E_UNSUPPORTED, "seeded LP has per-graph seeds: run each graph on its own");
    auto b = std::make_unique<glp_batch>();
    b->p = *p;
    b->graphs.assign(graphs, graphs + ngraphs);
    int64_t n = 0, m = 0;
    bool weighted = false;
    for (int32_t i = 0; i < ngraphs; ++i) {
        if (!graphs[i]) return set_error(GLP_E_ARG, "null graph %d", i);
        b->base.push_back(n);
        n += graphs[i]->csr.n;
        m += graphs[i]->csr.m;
        weighted |= graphs[i]->csr.weighted();
    }
    if (n > INT32_MAX - 1 || m > INT32_MAX)
        return set_error(GLP_E_RANGE, "batch of %lld vertices / %lld edges exceeds int32 CSR", (long long)n,
                         (long long)m);
    if (p->weighted && !weighted) return set_error(GLP_E_ARG, "weighted batch of unweighted graphs");
    b->uni = std::make_unique<glp_graph>();
    HostCSR& u = b->uni->csr;
    u.n = (int32_t)n;
    u.m = m;
    u.offsets.resize((size_t)n + 1);
    u.neighbors.resize((size_t)m);
    if (weighted) u.weights.resize((size_t)m);
    int64_t eo = 0;
    u.offsets[0] = 0;
    for (int32_t i = 0; i < ngraphs; ++i) {
        const HostCSR& g = graphs[i]->csr;
        const int32_t bs = (int32_t)b->base[i];
        for (int32_t v = 0; v < g.n; ++v) u.offsets[bs + v + 1] = (int32_t)(eo + g.offsets[v + 1]);
        for (int64_t j = 0; j < g.m; ++j) u.neighbors[eo + j] = g.neighbors[j] + bs;
        if (weighted) {
            for (int64_t j = 0; j < g.m; ++j) u.weights[eo + j] = g.weighted() ? g.weights[j] : 1u;
        }
        eo += g.m;
    }
    u.finish();
    // a disjoint union is symmetric when every part is
    u.symmetric = 1;
    for (int32_t i = 0; i < ngraphs; ++i) u.symmetric = std::min(u.symmetric, graphs[i]->csr.symmetric);
    int rc = session_create(b->uni.get(), p, &b->sess);
    if (rc != GLP_OK) return rc;
    *out = b.release();
    return GLP_OK;
}

extern "C" int glp_batch_run(glp_batch* b, int32_t steps, int32_t mode, double* ms) {
    clear_error();
    if (!b || steps < 0 || mode < 0 || mode > 2) return set_error(GLP_E_ARG, "bad batch run arguments");
    int rc;
    auto t0 = std::chrono::steady_clock::now();
    if (mode == 2) {
        if (b->plain.empty()) {
            b->plain.resize(b->graphs.size(), nullptr);
            for (size_t i = 0; i < b->graphs.size(); ++i)
                if ((rc = session_create(b->graphs[i], &b->p, &b->plain[i])) != GLP_OK) return rc;
            t0 = std::chrono::steady_clock::now();
        }
        for (auto* s : b->plain) {
            if ((rc = session_step(s, steps)) != GLP_OK) return rc;
            if ((rc = session_sync(s)) != GLP_OK) return rc;
        }
    } else {
        if (!b->sess) return set_error(GLP_E_STATE, "the batch has no union session (autotune failed)");
        rc = mode == 1 ? session_step_graph(b->sess, steps) : session_step(b->sess, steps);
        if (rc != GLP_OK) return rc;
        if ((rc = session_sync(b->sess)) != GLP_OK) return rc;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    b->last_mode = mode;
    return GLP_OK;
}

extern "C" int glp_batch_labels(glp_batch* b, int32_t graph, int32_t* out) {
    clear_error();
    if (!b || graph < 0 || graph >= (int32_t)b->graphs.size())
        return set_error(GLP_E_ARG, "bad batch graph index");
    const int32_t n = b->graphs[graph]->csr.n;
    if (n && !out) return set_error(GLP_E_ARG, "null output");
    if (b->last_mode == 2 && !b->plain.empty()) return session_labels(b->plain[graph], out);
    if (!b->sess) return set_error(GLP_E_STATE, "the batch has no union session (autotune failed)");
    int rc = session_sync(b->sess);
    if (rc != GLP_OK) return rc;
    if (n == 0) return GLP_OK;
    const int32_t bs = (int32_t)b->base[graph];
    GLP_HIP(hipMemcpy(out, b->sess->L[b->sess->cur] + bs, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    for (int32_t v = 0; v < n; ++v) out[v] -= bs;
    return GLP_OK;
}

extern "C" int glp_batch_reset(glp_batch* b) {
    clear_error();
    if (!b) return set_error(GLP_E_ARG, "null batch");
    if (!b->sess) return set_error(GLP_E_STATE, "the batch has no union session (autotune failed)");
    int rc = session_reset(b->sess);
    if (rc != GLP_OK) return rc;
    for (auto* s : b->plain)
        if ((rc = session_reset(s)) != GLP_OK) return rc;
    return GLP_OK;
}

extern "C" int glp_batch_autotune(glp_batch* b, int32_t steps, double* best_ms) {
    clear_error();
    if (!b || steps < 1) return set_error(GLP_E_ARG, "bad batch autotune arguments");
    // on failure the union keeps its plan, which the current session fits
    int rc = glp_autotune(b->uni.get(), &b->p, steps, nullptr, best_ms);
    if (rc != GLP_OK) return rc;
    // the union session's scratch was sized by the old plan: replace it (a
    // failed creation leaves no session, which run/labels/reset report)
    glp_session* s = nullptr;
    rc = session_create(b->uni.get(), &b->p, &s);
    delete b->sess;
    b->sess = s;
    return rc;
}

extern "C" void glp_batch_free(glp_batch* b) { delete b; }
