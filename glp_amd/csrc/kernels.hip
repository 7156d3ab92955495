// Degree-tiered LP counting kernels for gfx950 (wave64, 160 KB LDS per CU).
//
// One step computes, for every vertex v, the packed maximum over the labels
// l present among its neighbours of LabelScore(v, l, C_v(l)) with
// C_v(l) = sum of the frequencies LoadNeighbor assigns to v's edges showing l
// (the reference's per-tier kernels header/kernel.cuh:864-911 (huge),
// :1655-1739 (big), :3124-3222 (medium), :4365-4426 (small), :5187-5234
// (tiny)).  The work is split by degree:
//   packed  deg <= pack_max   several vertices share one wavefront: <= 64
//                             edges, counted by an all-pairs compare inside
//                             each vertex's lane segment, segmented max
//   wave    deg <= wave_max   one wavefront per vertex, LDS hash table
//   block   deg <= block_max  one workgroup per vertex, LDS hash table sized
//                             by the vertex's degree class
//   hub     deg >  block_max  several workgroups per vertex: each counts a
//                             chunk in LDS and flushes (label,count) pairs
//                             into the vertex's HBM hash table; a second pass
//                             walks the occupied slots, takes the maximum and
//                             clears them, so no table memset is ever needed
// Counts are exact integer sums and the winner is a maximum over packed
// (score,label) words, so every tier computes the same function and the
// result does not depend on scheduling.

#include "device.h"
#include "device_prims.h"

#include <algorithm>

namespace glp {
using namespace dev;

template <class T>
__device__ __forceinline__ T lds_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Inserts (key, f) into an LDS open-addressing table of `mask+1` slots.
// The table always has a free slot: its size is at least twice the number of
// edges inserted into it.
__device__ __forceinline__ void lds_insert(uint32_t* K, uint32_t* C, uint32_t mask, uint32_t key,
                                           uint32_t f) {
    uint32_t h = fmix32(key) & mask;
    for (;;) {
        uint32_t k = lds_load(&K[h]);
        if (k == key) {
            atomicAdd(&C[h], f);
            return;
        }
        if (k == 0u) {
            uint32_t prev = atomicCAS(&K[h], 0u, key);
            if (prev == 0u || prev == key) {
                atomicAdd(&C[h], f);
                return;
            }
        }
        h = (h + 1u) & mask;
    }
}

__device__ __forceinline__ void count_changes(const StepArgs& a, bool changed) {
    unsigned long long bal = __ballot(changed);
    if (lane_id() == 0 && bal) atomicAdd(a.changes, (unsigned long long)__popcll(bal));
}

// ---------------------------------------------------------------- packed ----
template <class H, bool W, bool G>
__global__ __launch_bounds__(256) void k_packed(StepArgs a, const int2* __restrict__ groups,
                                                int32_t ngroups) {
    const int lane = lane_id();
    const int32_t gid = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (gid >= ngroups) return;  // wave-uniform
    const int2 gr = groups[gid];
    const int32_t v0 = gr.x, vc = gr.y;
    const int32_t base = a.off[v0];
    int32_t vb = 0, ve = 0;
    if (lane < vc) {
        vb = a.off[v0 + lane] - base;
        ve = a.off[v0 + lane + 1] - base;
    }
    const int32_t nE = shfl_i32(ve, vc - 1);
    // owner of edge lane `lane`: smallest vertex slot i with ve_i > lane
    int lo = 0, hi = vc - 1;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int mid = (lo + hi) >> 1;
        const int32_t vm = shfl_i32(ve, mid);
        if (vm <= lane) lo = min(mid + 1, hi);
        else hi = mid;
    }
    const int own = lo;
    const int32_t sb = shfl_i32(vb, own), se = shfl_i32(ve, own);
    const int32_t vid = v0 + own;
    const bool live = lane < nE;
    uint32_t key = 0u, fr = 0u;
    if (live) key = H::template load_neighbor<W, G>(a, vid, base + lane, &fr);
    // frequency of my label inside my vertex's segment
    const int32_t len = live ? se - sb : 0;
    uint32_t cnt = 0u;
    for (int k = 0; __ballot(k < len) != 0ull; ++k) {
        const int src = min(sb + k, 63);
        const uint32_t ok = shfl_u32(key, src);
        const uint32_t of = W ? shfl_u32(fr, src) : 1u;
        if (k < len && ok == key) cnt += of;
    }
    uint64_t p = live ? H::label_score(a, vid, key, cnt) : 0ull;
    // segmented max: after the sweep the last lane of a segment holds its max
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t o = shfl_u64(p, lane >= d ? lane - d : lane);
        if (lane - d >= sb) p = max_u64(p, o);
    }
    const bool has = lane < vc && ve > vb;
    const uint64_t best = shfl_u64(p, has ? ve - 1 : lane);
    bool changed = false;
    if (lane < vc) {
        const int32_t v = v0 + lane;
        const int32_t old = a.keep[v];
        const int32_t nl = has ? packed_label(best) : old;
        H::update_vertex(a, v, nl, has ? best : 0ull);
        changed = nl != old;
    }
    count_changes(a, changed);
}

// ------------------------------------------------------------------ wave ----
template <class H, bool W, bool G>
__global__ __launch_bounds__(256) void k_wave(StepArgs a, const int32_t* __restrict__ verts,
                                              int32_t nv) {
    __shared__ uint32_t s_key[4][kWaveCap];
    __shared__ uint32_t s_cnt[4][kWaveCap];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int32_t idx = (int32_t)blockIdx.x * 4 + wv;
    if (idx >= nv) return;  // wave-uniform
    const int32_t v = verts[idx];
    const int32_t b = a.off[v], e = a.off[v + 1];
    const uint32_t cap = min(max(next_pow2(2u * (uint32_t)(e - b)), 64u), (uint32_t)kWaveCap);
    const uint32_t mask = cap - 1u;
    uint32_t* K = s_key[wv];
    uint32_t* C = s_cnt[wv];
    for (uint32_t i = lane; i < cap; i += 64) {
        K[i] = 0u;
        C[i] = 0u;
    }
    wave_lds_sync();
    for (int32_t j = b + lane; j < e; j += 64) {
        uint32_t fr;
        const uint32_t key = H::template load_neighbor<W, G>(a, v, j, &fr);
        lds_insert(K, C, mask, key, fr);
    }
    wave_lds_sync();
    uint64_t best = 0ull;
    for (uint32_t i = lane; i < cap; i += 64) {
        const uint32_t k = K[i];
        if (k) best = max_u64(best, H::label_score(a, v, k, C[i]));
    }
    best = wave_max_u64(best);
    bool changed = false;
    if (lane == 0) {
        const int32_t old = a.keep[v];
        const int32_t nl = packed_label(best);
        H::update_vertex(a, v, nl, best);
        changed = nl != old;
    }
    count_changes(a, changed);
}

// ----------------------------------------------------------------- block ----
template <int NT>
__device__ __forceinline__ uint64_t block_max_u64(uint64_t x, unsigned long long* s_red) {
    x = wave_max_u64(x);
    const int wv = threadIdx.x >> 6;
    if (lane_id() == 0) s_red[wv] = x;
    __syncthreads();
    if (threadIdx.x < 64) {
        x = threadIdx.x < NT / 64 ? s_red[threadIdx.x] : 0ull;
        x = wave_max_u64(x);
    }
    return x;  // valid in wave 0
}

template <class H, bool W, bool G, int NT>
__global__ __launch_bounds__(NT) void k_block(StepArgs a, const int32_t* __restrict__ verts,
                                              uint32_t cap_max) {
    extern __shared__ uint32_t s_dyn[];
    __shared__ unsigned long long s_red[NT / 64];
    uint32_t* K = s_dyn;
    uint32_t* C = s_dyn + cap_max;
    const int32_t v = verts[blockIdx.x];
    const int32_t b = a.off[v], e = a.off[v + 1];
    const uint32_t cap = min(max(next_pow2(2u * (uint32_t)(e - b)), 64u), cap_max);
    const uint32_t mask = cap - 1u;
    for (uint32_t i = threadIdx.x; i < cap; i += NT) {
        K[i] = 0u;
        C[i] = 0u;
    }
    __syncthreads();
    for (int32_t j = b + (int32_t)threadIdx.x; j < e; j += NT) {
        uint32_t fr;
        const uint32_t key = H::template load_neighbor<W, G>(a, v, j, &fr);
        lds_insert(K, C, mask, key, fr);
    }
    __syncthreads();
    uint64_t best = 0ull;
    for (uint32_t i = threadIdx.x; i < cap; i += NT) {
        const uint32_t k = K[i];
        if (k) best = max_u64(best, H::label_score(a, v, k, C[i]));
    }
    best = block_max_u64<NT>(best, s_red);
    if (threadIdx.x < 64) {
        bool changed = false;
        if (threadIdx.x == 0) {
            const int32_t old = a.keep[v];
            const int32_t nl = packed_label(best);
            H::update_vertex(a, v, nl, best);
            changed = nl != old;
        }
        count_changes(a, changed);
    }
}

// ------------------------------------------------------------------- hub ----
struct HubArgs {
    const int32_t* hubs;
    const int2* chunks;
    const int64_t* tbl_off;
    const uint32_t* cap;
    const int64_t* occ_off;
    int32_t chunk;
    int32_t nhubs;
    HubScratch hs;
};

template <class H, bool W, bool G>
__global__ __launch_bounds__(256) void k_hub_count(StepArgs a, HubArgs h) {
    __shared__ uint32_t K[kHubLdsCap];
    __shared__ uint32_t C[kHubLdsCap];
    const int2 hc = h.chunks[blockIdx.x];
    const int32_t v = h.hubs[hc.x];
    const int32_t vb = a.off[v], ve = a.off[v + 1];
    const int32_t b = vb + hc.y * h.chunk;
    const int32_t e = min(b + h.chunk, ve);
    const uint32_t cap = min(max(next_pow2(2u * (uint32_t)(e - b)), 64u), (uint32_t)kHubLdsCap);
    const uint32_t mask = cap - 1u;
    for (uint32_t i = threadIdx.x; i < cap; i += 256) {
        K[i] = 0u;
        C[i] = 0u;
    }
    __syncthreads();
    for (int32_t j = b + (int32_t)threadIdx.x; j < e; j += 256) {
        uint32_t fr;
        const uint32_t key = H::template load_neighbor<W, G>(a, v, j, &fr);
        lds_insert(K, C, mask, key, fr);
    }
    __syncthreads();
    unsigned long long* T = h.hs.tbl + h.tbl_off[hc.x];
    const uint32_t gmask = h.cap[hc.x] - 1u;
    for (uint32_t i = threadIdx.x; i < cap; i += 256) {
        const uint32_t k = K[i];
        if (!k) continue;
        const uint32_t c = C[i];
        uint32_t pos = fmix32(k) & gmask;
        for (;;) {
            unsigned long long cur =
                __hip_atomic_load(&T[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == 0ull) {
                const unsigned long long prev =
                    atomicCAS(&T[pos], 0ull, ((unsigned long long)k << 32) | c);
                if (prev == 0ull) {
                    const uint32_t slot = atomicAdd(&h.hs.occ_cnt[hc.x], 1u);
                    h.hs.occ[h.occ_off[hc.x] + slot] = pos;
                    break;
                }
                cur = prev;
            }
            if ((uint32_t)(cur >> 32) == k) {
                atomicAdd(&T[pos], (unsigned long long)c);
                break;
            }
            pos = (pos + 1u) & gmask;
        }
    }
}

template <class H>
__global__ __launch_bounds__(256) void k_hub_reduce(StepArgs a, HubArgs h) {
    __shared__ unsigned long long s_red[4];
    const int2 hc = h.chunks[blockIdx.x];
    const int32_t v = h.hubs[hc.x];
    const uint32_t n_occ = h.hs.occ_cnt[hc.x];
    const uint32_t lo = (uint32_t)hc.y * (uint32_t)h.chunk;
    const uint32_t hi = min(lo + (uint32_t)h.chunk, n_occ);
    unsigned long long* T = h.hs.tbl + h.tbl_off[hc.x];
    const uint32_t* occ = h.hs.occ + h.occ_off[hc.x];
    uint64_t best = 0ull;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) {
        const uint32_t pos = occ[i];
        const unsigned long long s = T[pos];
        T[pos] = 0ull;
        best = max_u64(best, H::label_score(a, v, (uint32_t)(s >> 32), (uint32_t)s));
    }
    best = block_max_u64<256>(best, s_red);
    if (threadIdx.x == 0 && best) atomicMax(&h.hs.best[hc.x], (unsigned long long)best);
}

template <class H>
__global__ __launch_bounds__(256) void k_hub_final(StepArgs a, HubArgs h) {
    const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    bool changed = false;
    if (i < h.nhubs) {
        const int32_t v = h.hubs[i];
        const unsigned long long best = h.hs.best[i];
        h.hs.best[i] = 0ull;
        h.hs.occ_cnt[i] = 0u;
        const int32_t old = a.keep[v];
        const int32_t nl = packed_label(best);
        H::update_vertex(a, v, nl, best);
        changed = nl != old;
    }
    count_changes(a, changed);
}

// ------------------------------------------------------------ auxiliaries ----
template <class H>
__global__ __launch_bounds__(256) void k_pick(StepArgs a, int32_t* __restrict__ shown) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < a.n) shown[v] = H::pick_label(a, (int32_t)v);
}

// adj[j] = L[nbr[j]] (header/kernel.cuh:225-256 gather_labels), four
// neighbours per lane with 16-byte streaming loads and stores.
typedef int32_t v4i32 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_gather(const int32_t* __restrict__ nbr,
                                                const int32_t* __restrict__ L,
                                                int32_t* __restrict__ adj, int64_t m) {
    const int64_t nvec = m >> 2;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const v4i32* nb4 = reinterpret_cast<const v4i32*>(nbr);
    v4i32* ad4 = reinterpret_cast<v4i32*>(adj);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride) {
        const v4i32 u = __builtin_nontemporal_load(&nb4[i]);
        v4i32 r;
        r.x = L[u.x];
        r.y = L[u.y];
        r.z = L[u.z];
        r.w = L[u.w];
        __builtin_nontemporal_store(r, &ad4[i]);
    }
    const int64_t t = (nvec << 2) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && t < m) adj[t] = L[nbr[t]];
}

__global__ void k_fill_i32(int32_t* p, int32_t value, int64_t count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        p[i] = value;
}
__global__ void k_iota(int32_t* p, int64_t count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        p[i] = (int32_t)i;
}
// Moves the step's change count into the ring and advances the step counter.
__global__ void k_end_step(unsigned long long* acc, unsigned long long* ring, int32_t* ctr,
                           int32_t ring_size) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int32_t t = *ctr;
        ring[t % ring_size] = *acc;
        *acc = 0ull;
        *ctr = t + 1;
    }
}

// ---------------------------------------------------------------- launch ----
static inline unsigned grid_for(int64_t work, int per_block) {
    int64_t g = (work + per_block - 1) / per_block;
    return (unsigned)std::max<int64_t>(1, g);
}

template <class H, bool W, bool G>
static int launch_tiers(const DeviceGraph& g, const HubScratch& hs, const StepArgs& a,
                        hipStream_t s, hipEvent_t* ev) {
    const Plan& p = g.plan;
    auto mark = [&](int tier, int edge) {
        if (ev) (void)hipEventRecord(ev[2 * tier + edge], s);
    };
    mark(GLP_TIER_PACKED, 0);
    if (!p.groups.empty()) {
        const int32_t ng = (int32_t)p.groups.size();
        hipLaunchKernelGGL((k_packed<H, W, G>), dim3(grid_for(ng, 4)), dim3(256), 0, s, a,
                           g.d_groups, ng);
    }
    mark(GLP_TIER_PACKED, 1);
    mark(GLP_TIER_WAVE, 0);
    if (!p.wave.empty()) {
        const int32_t nw = (int32_t)p.wave.size();
        hipLaunchKernelGGL((k_wave<H, W, G>), dim3(grid_for(nw, 4)), dim3(256), 0, s, a, g.d_wave,
                           nw);
    }
    mark(GLP_TIER_WAVE, 1);
    mark(GLP_TIER_BLOCK, 0);
    for (int c = 0; c < p.nclasses; ++c) {
        if (p.block[c].empty()) continue;
        const uint32_t cap = (uint32_t)p.class_cap[c];
        const size_t lds = (size_t)cap * 2 * sizeof(uint32_t);
        const unsigned nb = (unsigned)p.block[c].size();
        if (p.tune.block_threads == 512) {
            hipLaunchKernelGGL((k_block<H, W, G, 512>), dim3(nb), dim3(512), lds, s, a,
                               g.d_block[c], cap);
        } else if (p.tune.block_threads == 1024) {
            hipLaunchKernelGGL((k_block<H, W, G, 1024>), dim3(nb), dim3(1024), lds, s, a,
                               g.d_block[c], cap);
        } else {
            hipLaunchKernelGGL((k_block<H, W, G, 256>), dim3(nb), dim3(256), lds, s, a,
                               g.d_block[c], cap);
        }
    }
    mark(GLP_TIER_BLOCK, 1);
    mark(GLP_TIER_HUB, 0);
    if (!p.hubs.empty()) {
        HubArgs h;
        h.hubs = g.d_hubs;
        h.chunks = g.d_hub_chunks;
        h.tbl_off = g.d_hub_tbl_off;
        h.cap = g.d_hub_cap;
        h.occ_off = g.d_hub_occ_off;
        h.chunk = p.tune.hub_chunk;
        h.nhubs = (int32_t)p.hubs.size();
        h.hs = hs;
        const unsigned nc = (unsigned)p.hub_chunks.size();
        hipLaunchKernelGGL((k_hub_count<H, W, G>), dim3(nc), dim3(256), 0, s, a, h);
        mark(GLP_TIER_HUB, 1);
        mark(GLP_TIER_AUX, 0);
        hipLaunchKernelGGL((k_hub_reduce<H>), dim3(nc), dim3(256), 0, s, a, h);
        hipLaunchKernelGGL((k_hub_final<H>), dim3(grid_for(h.nhubs, 256)), dim3(256), 0, s, a, h);
        mark(GLP_TIER_AUX, 1);
    } else {
        mark(GLP_TIER_HUB, 1);
        mark(GLP_TIER_AUX, 0);
        mark(GLP_TIER_AUX, 1);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(GLP_E_HIP, "tier launch failed: %s", hipGetErrorString(e));
    return GLP_OK;
}

template <class H>
static int launch_hooks(bool weighted, bool gathered, const DeviceGraph& g, const HubScratch& hs,
                        const StepArgs& a, hipStream_t s, hipEvent_t* ev) {
    if (weighted) {
        if (gathered) return launch_tiers<H, true, true>(g, hs, a, s, ev);
        return launch_tiers<H, true, false>(g, hs, a, s, ev);
    }
    if (gathered) return launch_tiers<H, false, true>(g, hs, a, s, ev);
    return launch_tiers<H, false, false>(g, hs, a, s, ev);
}

void init_kernels() {
    static bool done = false;
    if (done) return;
    done = true;
    const int lim = 160 * 1024 - 1024;
#define GLP_SETLDS(HK, W, G)                                                                    \
    (void)hipFuncSetAttribute((const void*)k_block<HK, W, G, 256>,                              \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);                 \
    (void)hipFuncSetAttribute((const void*)k_block<HK, W, G, 512>,                              \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);                 \
    (void)hipFuncSetAttribute((const void*)k_block<HK, W, G, 1024>,                             \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);
#define GLP_SETLDS4(HK) GLP_SETLDS(HK, false, false) GLP_SETLDS(HK, false, true) \
                        GLP_SETLDS(HK, true, false) GLP_SETLDS(HK, true, true)
    GLP_SETLDS4(ClassicHooks)
    GLP_SETLDS4(LayeredHooks)
    GLP_SETLDS4(SlpaHooks)
    GLP_SETLDS4(UserHooks)
#undef GLP_SETLDS4
#undef GLP_SETLDS
    (void)hipGetLastError();
}

int launch_step(Method method, bool weighted, bool gathered, const DeviceGraph& g,
                const HubScratch& hs, const StepArgs& a, hipStream_t s, hipEvent_t* ev) {
    switch (method) {
        case Method::Classic: return launch_hooks<ClassicHooks>(weighted, gathered, g, hs, a, s, ev);
        case Method::Layered: return launch_hooks<LayeredHooks>(weighted, gathered, g, hs, a, s, ev);
        case Method::Slpa: return launch_hooks<SlpaHooks>(weighted, gathered, g, hs, a, s, ev);
        case Method::User: return launch_hooks<UserHooks>(weighted, gathered, g, hs, a, s, ev);
    }
    return set_error(GLP_E_ARG, "unknown method");
}

int launch_pick(Method method, const StepArgs& a, int32_t* shown, hipStream_t s) {
    const unsigned grid = grid_for(a.n, 256);
    switch (method) {
        case Method::Slpa:
            hipLaunchKernelGGL(k_pick<SlpaHooks>, dim3(grid), dim3(256), 0, s, a, shown);
            break;
        case Method::User:
            hipLaunchKernelGGL(k_pick<UserHooks>, dim3(grid), dim3(256), 0, s, a, shown);
            break;
        default:
            return set_error(GLP_E_ARG, "method has no pick kernel");
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "pick launch: %s", hipGetErrorString(e));
}

int launch_gather(const StepArgs& a, int64_t m, int32_t* adj, hipStream_t s) {
    if (m == 0) return GLP_OK;
    const unsigned grid = (unsigned)std::min<int64_t>(grid_for(m / 4 + 1, 256), 256 * 64);
    hipLaunchKernelGGL(k_gather, dim3(grid), dim3(256), 0, s, a.nbr, a.L, adj, m);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "gather launch: %s", hipGetErrorString(e));
}

int launch_fill_i32(int32_t* p, int32_t value, int64_t count, hipStream_t s) {
    if (count <= 0) return GLP_OK;
    const unsigned grid = (unsigned)std::min<int64_t>(grid_for(count, 256), 256 * 32);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid), dim3(256), 0, s, p, value, count);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "fill launch: %s", hipGetErrorString(e));
}

int launch_iota(int32_t* p, int64_t count, hipStream_t s) {
    if (count <= 0) return GLP_OK;
    const unsigned grid = (unsigned)std::min<int64_t>(grid_for(count, 256), 256 * 32);
    hipLaunchKernelGGL(k_iota, dim3(grid), dim3(256), 0, s, p, count);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "iota launch: %s", hipGetErrorString(e));
}

int launch_end_step(unsigned long long* acc, unsigned long long* ring, int32_t* step_ctr,
                    int32_t ring_size, hipStream_t s) {
    hipLaunchKernelGGL(k_end_step, dim3(1), dim3(64), 0, s, acc, ring, step_ctr, ring_size);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "end_step launch: %s", hipGetErrorString(e));
}

}  // namespace glp
