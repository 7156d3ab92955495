// Device-side graph copy, tier plan and kernel launch interface.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "internal.h"
#include "hooks.h"

namespace glp {

#define GLP_HIP(call)                                                               \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            return set_error(GLP_E_HIP, "%s failed: %s (%s:%d)", #call,             \
                             hipGetErrorString(e_), __FILE__, __LINE__);            \
    } while (0)

// Degree thresholds and launch shapes; the autotuner writes these.
struct Tuning {
    int32_t pack_max_deg = 32;     // deg <= this: lane-packed groups of <= 64 edges
    int32_t wave_max_deg = 256;    // deg <= this: one wavefront per vertex (LDS table)
    int32_t block_max_deg = 4096;  // deg <= this: one workgroup per vertex (LDS table)
    int32_t hub_chunk = 4096;      // edges per workgroup for hubs (deg > block_max_deg)
    int32_t block_threads = 256;   // workgroup size of the block tier
};
Tuning& default_tuning();  // process-wide defaults (GLP_TUNING env file overrides)

constexpr int kWaveCap = 512;      // LDS slots per wavefront in the wave tier
constexpr int kHubLdsCap = 8192;   // LDS slots per workgroup in the hub count pass
constexpr int kMaxBlockClasses = 5;

// Host-built tier plan.  Vertex lists hold ids; groups hold (first id, count).
struct Plan {
    Tuning tune;
    std::vector<int2> groups;                 // packed tier
    std::vector<int32_t> wave;                // wave tier
    int32_t nclasses = 0;
    int32_t class_cap[kMaxBlockClasses] = {};  // LDS slots of each block class
    std::vector<int32_t> block[kMaxBlockClasses];
    std::vector<int32_t> hubs;                // hub tier
    std::vector<int2> hub_chunks;             // (hub index, chunk index)
    std::vector<int64_t> hub_tbl_off;         // slot offset of each hub's HBM table
    std::vector<uint32_t> hub_cap;            // table slots (power of two)
    std::vector<int64_t> hub_occ_off;         // offset of each hub's occupied-slot list
    int64_t tbl_slots = 0, occ_slots = 0;
    int64_t edges[4] = {}, verts[4] = {};     // per tier: packed, wave, block, hub
};
int build_plan(const HostCSR& g, const Tuning& t, Plan& p);

// HBM copy of a graph plus its plan (shared by all sessions on the graph).
struct DeviceGraph {
    int32_t n = 0;
    int64_t m = 0;
    int32_t* off = nullptr;
    int32_t* nbr = nullptr;
    uint32_t* w = nullptr;
    Plan plan;
    int2* d_groups = nullptr;
    int32_t* d_wave = nullptr;
    int32_t* d_block[kMaxBlockClasses] = {};
    int32_t* d_hubs = nullptr;
    int2* d_hub_chunks = nullptr;
    int64_t* d_hub_tbl_off = nullptr;
    uint32_t* d_hub_cap = nullptr;
    int64_t* d_hub_occ_off = nullptr;
    double upload_ms = 0, plan_ms = 0;
    size_t bytes = 0;
    ~DeviceGraph();
};
int device_graph_create(const HostCSR& g, const Tuning& t, hipStream_t s, DeviceGraph** out);
// HBM bytes a device graph of this shape needs (planning estimate).
size_t device_graph_bytes(const HostCSR& g, const Plan& p);

// Per-session hub scratch: HBM hash tables (zero between steps), occupied
// lists, per-hub counters and winners.
struct HubScratch {
    unsigned long long* tbl = nullptr;
    uint32_t* occ = nullptr;
    uint32_t* occ_cnt = nullptr;
    unsigned long long* best = nullptr;
};

enum class Method { Classic, Layered, Slpa, User };

// One-time kernel attribute set-up (LDS limits); call before any capture.
void init_kernels();

// Enqueues the counting part of one step (all tiers) on stream s.
// ev (may be null) holds 2*GLP_NTIERS events for per-tier timing.
int launch_step(Method method, bool weighted, bool gathered, const DeviceGraph& g,
                const HubScratch& hs, const StepArgs& a, hipStream_t s, hipEvent_t* ev);
int launch_pick(Method method, const StepArgs& a, int32_t* shown, hipStream_t s);
int launch_gather(const StepArgs& a, int64_t m, int32_t* adj, hipStream_t s);
int launch_fill_i32(int32_t* p, int32_t value, int64_t count, hipStream_t s);
int launch_iota(int32_t* p, int64_t count, hipStream_t s);
int launch_end_step(unsigned long long* acc, unsigned long long* ring, int32_t* step_ctr,
                    int32_t ring_size, hipStream_t s);

}  // namespace glp
