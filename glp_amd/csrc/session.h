// Session state shared by the session, counting, batch and server units.
#pragma once

#include "device.h"

#include <array>
#include <vector>

namespace glp {
constexpr int32_t kChangeRing = 4096;  // per-step change counts kept on the device
}

struct glp_session {
    glp_graph* g = nullptr;
    glp_params p{};
    glp::Method method = glp::Method::Classic;
    int32_t n = 0;
    hipStream_t stream = nullptr;
    int32_t* L[2] = {nullptr, nullptr};  // labels, double-buffered
    int32_t* V[2] = {nullptr, nullptr};  // layered volume histogram, double-buffered
    int2* LV = nullptr;                  // layered, gamma > 0: {label, its volume} per vertex
    int32_t* shown = nullptr;            // PickLabel output (SLPA speakers, user hooks)
    int32_t* adj = nullptr;              // pre-gathered neighbour labels (gather mode)
    int32_t* adj_alt = nullptr;          // the second adj of the two-step lookahead
    int32_t* mem = nullptr;              // SLPA memory (T+1) x n
    int32_t* wide = nullptr;             // counting methods' vertex lists (count-min: nmid, then nwide)
    int32_t nmid = 0, nwide = 0;
    int32_t ncmhub = 0, ncmchunk = 0;    // count-min hubs (after nwide in `wide`) and their chunks
    int32_t ncm256 = 0;                  // count-min: the first ncm256 of nwide take 256-thread workgroups
    int32_t nsmall = 0, niso = 0;        // count-min degrees 1..16 and 0 (after the hubs in `wide`)
    int32_t nsmall_g[5] = {0, 0, 0, 0, 0};  // the small list's degree classes 1, 2, 3-4, 5-8, 9-16
    int2* cm_chunks = nullptr;           // (hub slot, first edge) per chunk
    uint32_t* cm_sketch = nullptr;       // rows x width cells per hub, zero between steps
    unsigned long long* cm_best = nullptr;  // packed maximum per hub, zero between steps
    uint32_t* cm_keys = nullptr;         // each chunk's edge keys, from the sketch pass to the estimate pass
    uint8_t* act[2] = {nullptr, nullptr};  // frontier flags, double-buffered like L
    int32_t* fctl = nullptr;             // frontier control {all_active, mark, threshold, gather words}
    int64_t act_bytes = 0;
    uint32_t* chg[2] = {nullptr, nullptr};  // frontier gather: changed-vertex bitmaps
    void* push = nullptr;                   // frontier gather: push entries of a step
    uint8_t* seeds = nullptr;              // seeded LP: n bytes, nonzero = seed
    bool seeds_set = false;
    int64_t chg_bytes = 0;
    glp::HubScratch hs;
    unsigned long long* d_acc = nullptr;   // change counter of the running step
    unsigned long long* d_ring = nullptr;  // change count of each finished step
    int32_t* d_ctr = nullptr;              // device-side step counter
    int cur = 0;                           // index of the current label buffer
    int32_t t = 0;                         // steps enqueued
    int32_t pending_steps = 0;
    size_t bytes = 0;                      // device bytes owned (graph excluded)
    std::vector<std::array<hipEvent_t, 2 * GLP_NTIERS>> ev_pool;  // profiled steps' events
    size_t ev_used = 0;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    glp_timing timing{};
    std::vector<int64_t> host_changes;
    ~glp_session();
};

namespace glp {
int ensure_device();
int check_params(const HostCSR& g, const glp_params* p);
// The same from the graph's shape: n, largest degree, largest weight.
int check_params_shape(int32_t n, int32_t max_degree, uint32_t wmax, bool weighted, const glp_params* p);
size_t session_bytes(const HostCSR& g, const Plan& p, const glp_params& prm);
int session_create(glp_graph* g, const glp_params* p, glp_session** out);
int session_reset(glp_session* s);
int session_step(glp_session* s, int32_t steps);
int session_step_graph(glp_session* s, int32_t steps);
int session_sync(glp_session* s);
int session_labels(glp_session* s, int32_t* out);
int enqueue_step(glp_session* s, hipEvent_t* ev);
// count-min / frequent / buffer steps (counting.hip)
int enqueue_counting_step(glp_session* s);
int counting_wide_list(glp_session* s);
// SLPA post-processing (counting.hip)
int slpa_post(glp_session* s, int32_t kmax, int32_t* dominant, int32_t* nsets, int32_t* sets);
}  // namespace glp
