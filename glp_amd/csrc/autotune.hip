// Autotuner: picks the tier thresholds, workgroup size and LDS table sizes
// by timing real steps on the caller's graph.
//
// Coordinate descent: starting from the current defaults, each key in turn
// tries its candidate values with the others fixed and keeps a value that is
// at least 2% faster; two sweeps.  A candidate reuses the graph's uploaded
// CSR and only re-plans and re-uploads the tier lists.  Every tuning computes
// the same labels (tiers differ only in how a vertex's count is organised),
// so tuning never changes results.

#include "device.h"
#include "session.h"

#include <cstdio>
#include <fstream>
#include <memory>

namespace {

struct Candidates {
    const char* key;
    std::vector<int32_t> values;
};

const Candidates kSpace[] = {
    {"pack_max_deg", {16, 24, 32, 48, 64}},
    {"half_max_deg", {0, 48, 64}},
    {"half_lanes", {0, 16, 32}},
    {"wave_max_deg", {64, 128, 192, 256}},
    {"block_max_deg", {512, 1024, 2048, 4096, 8192}},
    {"block_threads", {256, 512, 1024}},
    {"hub_chunk", {1024, 2048, 4096}},
    {"part_keys", {1024, 2048, 4096}},
    {"fused_small", {0, 1}},
};

}  // namespace

using namespace glp;

extern "C" int glp_autotune(glp_graph* g, const glp_params* p, int32_t steps, const char* out_path,
                            double* best_ms) {
    clear_error();
    if (!g || !p || steps < 1) return set_error(GLP_E_ARG, "bad autotune arguments");
    int rc = ensure_device();
    if (rc != GLP_OK) return rc;
    if ((rc = check_params(g->csr, p)) != GLP_OK) return rc;
    hipStream_t stream;
    GLP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{stream};
    if (!g->dev) {
        DeviceGraph* dg = nullptr;
        if ((rc = device_graph_create(g->csr, default_tuning(), stream, &dg)) != GLP_OK) return rc;
        g->dev = dg;
    }
    DeviceGraph* base = g->dev;

    // Times `steps` steps from L_0 with tuning t; ms < 0 when t is invalid.
    auto time_tuning = [&](const Tuning& t, double* ms) -> int {
        *ms = -1.0;
        auto dg = std::make_unique<DeviceGraph>();
        if (build_plan(g->csr, t, dg->plan) != GLP_OK) {
            clear_error();
            return GLP_OK;
        }
        dg->n = base->n;
        dg->m = base->m;
        dg->off = base->off;
        dg->nbr = base->nbr;
        dg->w = base->w;
        dg->owns_csr = false;
        int r = device_plan_upload(dg.get(), stream);
        if (r != GLP_OK) return r;
        g->dev = dg.get();
        glp_session* s = nullptr;
        r = session_create(g, p, &s);
        if (r == GLP_OK) {
            r = session_step(s, steps);
            if (r == GLP_OK) r = session_sync(s);
            if (r == GLP_OK) *ms = s->timing.steps_ms;
            delete s;
        }
        g->dev = base;
        return r;
    };

    // from the graph's current tuning
    Tuning cur = base->plan.tune;
    double best = 0;
    if ((rc = time_tuning(cur, &best)) != GLP_OK) return rc;
    if (best < 0) return set_error(GLP_E_ARG, "current tuning is invalid for this graph");
    for (int sweep = 0; sweep < 2; ++sweep) {
        bool improved = false;
        for (const auto& c : kSpace) {
            int32_t curv = 0;
            tuning_get(cur, c.key, &curv);
            for (int32_t v : c.values) {
                if (v == curv) continue;
                Tuning t = cur;
                tuning_set(t, c.key, v);
                double ms = 0;
                if ((rc = time_tuning(t, &ms)) != GLP_OK) return rc;
                if (ms >= 0 && ms < best * 0.98) {
                    best = ms;
                    cur = t;
                    curv = v;
                    improved = true;
                }
            }
        }
        if (!improved) break;
    }
    default_tuning() = cur;
    // re-plan the graph's own device copy with the winner (CSR arrays kept)
    {
        auto dg = std::make_unique<DeviceGraph>();
        if ((rc = build_plan(g->csr, cur, dg->plan)) != GLP_OK) return rc;
        dg->n = base->n;
        dg->m = base->m;
        dg->off = base->off;
        dg->nbr = base->nbr;
        dg->w = base->w;
        if ((rc = device_plan_upload(dg.get(), stream)) != GLP_OK) return rc;
        dg->owns_csr = true;
        base->owns_csr = false;
        delete base;
        g->dev = dg.release();
    }
    if (out_path) {
        std::ofstream out(out_path);
        if (!out) return set_error(GLP_E_IO, "cannot write '%s'", out_path);
        for (int i = 0; const char* k = tuning_key(i); ++i) {
            int32_t v = 0;
            tuning_get(cur, k, &v);
            out << k << "=" << v << "\n";
        }
    }
    if (best_ms) *best_ms = best;
    return GLP_OK;
}
