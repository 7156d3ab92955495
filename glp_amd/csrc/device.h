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
    int32_t half_max_deg = 64;     // deg <= this (<= 64): 64/half_lanes vertices per wavefront (0: off)
    int32_t half_lanes = 0;        // lanes per vertex there: 16, 32, or 0 = chosen per launch
    int32_t wave_max_deg = 256;    // deg <= this: one wavefront per vertex (LDS table)
    int32_t block_max_deg = 4096;  // deg <= this: one workgroup per vertex (LDS table)
    int32_t hub_chunk = 2048;      // edges per aggregation workgroup for hubs (deg > block_max_deg)
    int32_t part_keys = 2048;      // expected distinct labels per hub counting workgroup
    int32_t part_limit = 0;        // distinct labels before a hub range is halved (0: table/2)
    int32_t block_threads = 256;   // block-tier workgroup size (256: by table size, up to 1024)
    // 1: the packed and half tiers run as one persistent launch whose waves
    // take units from a counter; 0: one launch per tier and shape
    int32_t fused_small = 0;
};
int tuning_set(Tuning& t, const char* key, int32_t value);  // GLP_E_ARG on unknown key
int tuning_get(const Tuning& t, const char* key, int32_t* value);
const char* tuning_key(int i);  // i-th tuning key, nullptr past the last
Tuning& default_tuning();  // process-wide defaults (GLP_TUNING env file overrides)

constexpr int kChangeSlots = 1024;  // change counters, spread to avoid one hot address
constexpr int kWaveCap = 512;      // LDS slots per wavefront in the wave tier
constexpr int kHalfMaxDeg = 64;    // largest degree of the half-wave part of the wave tier
constexpr int kMaxBlockClasses = 5;

// Host-built tier plan.  Vertex lists hold ids; groups hold (first id, count).
struct Plan {
    Tuning tune;
    std::vector<int2> groups;                 // packed tier (edge-less groups last)
    int64_t groups_with_edges = 0;            // groups before the edge-less ones
    // wave-tier vertices as {vertex, first edge, end of edges, 0}: the edge
    // range travels with the vertex, so a wave's loads skip the offsets
    std::vector<int4> half;                   // deg <= half_max_deg: <= 32 first, then the rest
    int64_t half_short = 0;                   // half entries of degree <= 32 (the first ones)
    std::vector<int4> wave;                   // larger degrees: <= 128 first, then the rest
    int64_t wave_short = 0;                   // wave entries of degree <= 128 (the first ones)
    int32_t nclasses = 0;
    int32_t class_cap[kMaxBlockClasses] = {};  // LDS slots of each block class
    std::vector<int32_t> block[kMaxBlockClasses];
    std::vector<int32_t> hubs;                // hub tier
    std::vector<int2> hub_chunks;             // (hub index, chunk index): aggregation workgroups
    std::vector<int32_t> hub_chunk0;          // index of each hub's first chunk
    std::vector<int64_t> hub_base;            // first pair slot of each hub (chunk regions)
    std::vector<int2> hub_parts;              // (hub index, bucket): counting workgroups
    std::vector<int64_t> hub_hist_base;       // first histogram entry of each hub
    int64_t pair_slots = 0;                   // sum over hubs of chunks * hub_chunk
    int64_t hist_slots = 0;                   // sum over hubs of buckets * chunks
    int32_t hub_max_buckets = 0;              // most buckets of any hub (sizes k_hub_agg's LDS)
    int64_t edges[4] = {}, verts[4] = {};     // per tier: packed, wave, block, hub
};
int build_plan(const HostCSR& g, const Tuning& t, Plan& p);

// HBM copy of a graph plus its plan (shared by all sessions on the graph).
// For each vertex u, the CSR positions j with nbr[j] = u: off[n+1], pos[m]
// (built on the device when a frontier gather first needs it).
struct InverseIndex {
    int32_t* off = nullptr;
    int32_t* pos = nullptr;
};

struct DeviceGraph {
    int32_t n = 0;
    int64_t m = 0;
    int32_t* off = nullptr;
    int32_t* nbr = nullptr;
    uint32_t* w = nullptr;
    Plan plan;
    int2* d_groups = nullptr;
    int4* d_half = nullptr;
    int4* d_wave = nullptr;
    int32_t* d_block[kMaxBlockClasses] = {};
    int32_t* d_hubs = nullptr;
    int2* d_hub_chunks = nullptr;
    int32_t* d_hub_chunk0 = nullptr;
    int64_t* d_hub_base = nullptr;
    int2* d_hub_parts = nullptr;
    int64_t* d_hub_hist_base = nullptr;
    InverseIndex inv;
    double upload_ms = 0, plan_ms = 0;
    size_t bytes = 0;
    bool owns_csr = true;   // false: off/nbr/w borrowed from another DeviceGraph
    bool owns_plan = true;  // false: the d_* plan arrays live in memory the graph does not own
    ~DeviceGraph();
};
int device_graph_create(const HostCSR& g, const Tuning& t, hipStream_t s, DeviceGraph** out);
// Uploads a graph with an already built plan (synchronises stream s).
int device_graph_upload(const HostCSR& g, Plan&& plan, hipStream_t s, DeviceGraph** out);
// Uploads dg->plan's lists into dg's device arrays (synchronises stream s).
int device_plan_upload(DeviceGraph* dg, hipStream_t s);
// HBM bytes a device graph of this shape needs (planning estimate).
size_t device_graph_bytes(const HostCSR& g, const Plan& p);

// Per-session hub scratch: each chunk's aggregated (label+1, frequency)
// pairs grouped by bucket, the (bucket, chunk) runs they form, and per-hub
// winners (zero between steps).
struct HubScratch {
    unsigned long long* pairs = nullptr;  // per chunk region: its distinct pairs, by bucket
    uint32_t* hist = nullptr;             // per hub: P x chunks run lengths (bucket-major)
    uint32_t* lstart = nullptr;           // per hub: P x chunks run starts in the chunk regions
    unsigned long long* best = nullptr;
    uint32_t* changed = nullptr;  // frontier: hub's label changed this step
    uint32_t* work = nullptr;     // the persistent small-tier launch's unit counter
};
// Device bytes of a plan's hub scratch (frontier: the changed flags too).
inline size_t hub_scratch_bytes(const Plan& p, bool frontier) {
    return (size_t)p.pair_slots * 8 + (size_t)p.hist_slots * 8 + p.hubs.size() * (frontier ? 12 : 8);
}

// Buckets of a hub of degree d (host and device agree): enough that a bucket
// expects at most 3/4 part_keys distinct labels, at most kMaxBuckets.
constexpr int kMaxBuckets = 4096;
__host__ __device__ __forceinline__ uint32_t hub_buckets(int32_t d, int32_t part_keys) {
    const int64_t per = 3 * (int64_t)part_keys;  // buckets = ceil(4 d / (3 part_keys))
    const uint32_t p = (uint32_t)((4 * (int64_t)d + per - 1) / per);
    return p < 1u ? 1u : (p > (uint32_t)kMaxBuckets ? (uint32_t)kMaxBuckets : p);
}

// Hub-tier launch arguments (plan lists plus session scratch).
struct HubArgs {
    const int32_t* hubs;
    const int2* chunks;       // (hub, chunk) per aggregation workgroup
    const int2* parts;        // (hub, part) per counting workgroup
    const int32_t* chunk0;    // first chunk index of each hub
    const int64_t* base;      // first pair slot of each hub
    const int64_t* hist_base; // first histogram entry of each hub
    int32_t chunk;            // edges per chunk = pair slots per chunk region
    int32_t part_keys;        // expected distinct labels per counting workgroup
    int32_t part_limit;       // distinct labels that trigger halving (0: table/2)
    int32_t nhubs;
    int32_t nparts;           // (hub, bucket) parts; k_hub_part takes them in batches
    HubScratch hs;
};

// One job's share of a folded (multi-job) step: its step arguments and
// tier work lists.  blocks[] holds the workgroups it needs in each launch,
// in the order LAUNCH_* below.
enum { kLaunchPacked = 0, kLaunchHalf = 1, kLaunchWave = 2, kLaunchBlock0 = 3,
       kLaunchHubAgg = 3 + kMaxBlockClasses, kLaunchHubPart, kLaunchHubFinal, kLaunches };
struct JobDesc {
    StepArgs a;
    const int2* groups;
    const int4* half;
    const int4* wave;
    const int32_t* block[kMaxBlockClasses];
    HubArgs hub;
    int32_t ngroups, nhalf, nwave;
    int32_t ngroups_e;  // groups with edges (the first ngroups_e of the list)
    int32_t blocks[kLaunches];
};

// LayeredVol: layered LP with gamma > 0 scored in the table slots from the
// (label, volume) records of StepArgs::LV (LayeredVolHooks).
enum class Method { Classic, Layered, LayeredVol, Slpa, User, Seeded };

// One-time kernel attribute set-up (LDS limits); call before any capture.
void init_kernels();
// The same for the count-min and table kernels (counting.hip); init_kernels
// calls it.
void init_counting_kernels();

// Enqueues the counting part of one step (all tiers) on stream s.
// ev (may be null) holds 2*GLP_NTIERS events for per-tier timing.
int launch_step(Method method, bool weighted, bool gathered, const DeviceGraph& g,
                const HubScratch& hs, const StepArgs& a, hipStream_t s, hipEvent_t* ev);
int launch_pick(Method method, const StepArgs& a, int32_t* shown, hipStream_t s);
int launch_gather(const StepArgs& a, int64_t m, int32_t* adj, hipStream_t s);
// Builds inv (arrays allocated by the caller) from nbr; cursor is n+1 int32
// of scratch.  Synchronises s.
int inverse_index_build(const int32_t* nbr, int32_t n, int64_t m, int32_t* cursor, InverseIndex& inv,
                        hipStream_t s);
// Most in-edges of changed sources for which a frontier gather pushes
// (frontier control word 7); above it the gather rewrites every edge.
int32_t push_edge_limit(int64_t m);
// Random-read calibration: `reads` 4-byte reads of T (entries a power of
// two) at uniform hashed indices; out holds out_entries int32 sums.
int launch_calib_read(const int32_t* T, int64_t entries, int64_t reads, uint32_t seed, int32_t* out,
                      int64_t out_entries, hipStream_t s);
// Bytes of the push-entry list a frontier gather over (n, m) needs.
size_t push_entry_bytes(int32_t n, int64_t m);
// The gather of a frontier step (a.chg_r, a.fctl set): pushes the labels of
// the sources that changed in the previous step through the inverse index,
// or rewrites every edge after a step with many changes.
int launch_frontier_gather(const StepArgs& a, int64_t m, int32_t* adj, const InverseIndex& inv,
                           void* push, hipStream_t s);
int launch_fill_i32(int32_t* p, int32_t value, int64_t count, hipStream_t s);
int launch_iota(int32_t* p, int64_t count, hipStream_t s);
// LV[u] = {L[u], V[L[u]]} for u < n (layered LP's records for LayeredVol).
int launch_label_volumes(const int32_t* L, const int32_t* V, int2* LV, int32_t n, hipStream_t s);
// V[l] += #{v < n : L[v] = l} (layered LP's volume of the written labels;
// V zeroed first).
int launch_volume(const int32_t* L, int32_t* V, int32_t n, hipStream_t s);
// p[i] = seeds[i] ? i : -1 (seeded LP's initial labels).
int launch_seed_labels(int32_t* p, const uint8_t* seeds, int64_t count, hipStream_t s);
int launch_end_step(unsigned long long* acc, unsigned long long* ring, int32_t* step_ctr,
                    int32_t ring_size, hipStream_t s, int32_t* fctl = nullptr);
// Fills count bytes (count a multiple of 4, p 4-byte aligned) with zero.
inline int launch_zero_bytes(void* p, int64_t count, hipStream_t s) {
    return launch_fill_i32(static_cast<int32_t*>(p), 0, count / 4, s);
}

// One layered job's labels and volumes for the folded volume launches of
// the serving path: the written labels and write-side volumes (zero, fill),
// or the read labels, read-side volumes and the records to write (records).
struct VolumeJob {
    const int32_t* L;
    int32_t* V;
    int32_t n;
    int32_t pad;
    int2* LV;
};
// Workgroups a job of n vertices needs in the folded zero (and records) and
// volume launches.
void volume_job_blocks(int32_t n, int32_t* zero_blocks, int32_t* volume_blocks);
// kVolumeZero: V = 0 for every job (before the tiers); kVolumeRecords:
// LV[u] = {L[u], V[L[u]]} (before the tiers); kVolumeFill: V += histogram
// of L (after them).  cum: nj + 1 prefix sums of the job's workgroups;
// grid = cum[nj].
enum VolumeMode { kVolumeZero, kVolumeRecords, kVolumeFill };
int launch_multi_volume(VolumeMode mode, const VolumeJob* jobs, const int32_t* cum, int32_t nj, int32_t grid,
                        hipStream_t s);

// Device CSR built on the device (gen_device.hip); n + 1 offsets, m
// neighbours and m weights (null when unweighted), all hipMalloc'ed.
struct DeviceCSR {
    int32_t n = 0;
    int64_t m = 0;
    int32_t* off = nullptr;
    int32_t* nbr = nullptr;
    uint32_t* w = nullptr;
    int32_t max_degree = 0;
    uint32_t wmax = 0;
};
// The graph glp_gen_rmat(scale, edge_factor, a, b, c, seed, degree_sort = 0,
// wmax) builds on the host (same draws, same CSR), built in HBM on stream s
// (synchronises s).
int gen_rmat_device(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                    int32_t wmax, hipStream_t s, DeviceCSR* out);

// Folded step for nj jobs with the layered hooks (gamma 0 = classic) or,
// slot_scores, the slot-score hooks (layered, gamma > 0: reads a.LV).
// jobs and cum are device arrays; grid (host) holds each launch's size.
int launch_multi(bool weighted, bool slot_scores, const JobDesc* jobs, const int32_t* cum, int32_t nj,
                 const int32_t* grid, const Tuning& t, const int32_t* class_cap, hipStream_t s);
// Workgroups each folded launch needs for a job with plan p.
void job_blocks(const Plan& p, int32_t* blocks);

}  // namespace glp
