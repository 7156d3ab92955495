// Approximate and bounded-memory counting methods, and SLPA post-processing.
//
//   count-min  (method/countmin.cuh:10-263, used as a filter by
//               l_big_update_syn3, header/kernel.cuh:1420-1566): per vertex a
//               d x W sketch in LDS; every edge adds its frequency to one cell
//               per row; a label's estimate is the minimum of its cells after
//               all edges are in; the winner is the packed maximum of
//               (estimate, label).  Integer sums commute, so the result does
//               not depend on the order edges are added.  cm_aliased=1 keeps
//               the rows exactly as countmin.cuh writes them (d=3 adds row 3's
//               hash into row 2, :185-191; d=4 adds h3 into row 2 twice,
//               :197-221); cm_aliased=0 uses d independent rows.
//   frequent   (method/frequent.cuh SharedFA, a Misra-Gries table of C
//               slots): each vertex's edges are taken in CSR order; a present
//               label gains its frequency, an absent one takes a free slot,
//               and with no free slot every counter drops by
//               min(frequency, smallest counter), emptied slots are freed
//               and the label enters with what is left (the unit-step
//               Misra-Gries rule applied frequency times).
//   buffer     (method/buffer.cuh Buffer): as frequent, except that with no
//               free slot every counter drops by min(counter, frequency),
//               emptied slots are freed and the incoming label is dropped
//               (buffer.cuh:161-172).
// For frequent and buffer the table is processed sequentially per vertex,
// the order-dependent semantics the reference's concurrent tables only
// approximate.  The winner is the largest (count, label) left in the table;
// a vertex whose table ends empty keeps its label.
//
// Degree tiers.  Count-min: 16 lanes per vertex up to kCountminSmallDeg
// edges, with no sketch (each edge's estimate is summed over the vertex's
// edges whose cells coincide with its own, which is what the sketch cell
// would hold); one wavefront per vertex up to kCountminWaveMaxDeg edges and
// one 1024-thread workgroup per vertex up to kCountminHubDeg, with the sketch
// in LDS; above that, chunks of kCountminChunk edges on workgroups of their
// own (one pass adds each chunk's LDS sketch into the vertex's sketch in HBM,
// a second estimates every edge's label against the whole sketch and keeps
// the packed maximum with a 64-bit atomic max, a third commits).  Frequent and buffer: one lane per
// vertex up to kTableLaneMaxDeg edges (the lane walks its table slot by
// slot), one wavefront per vertex above, with table slot i held by lane i
// and each edge applied by the whole wavefront in CSR order: the order the
// semantics fix, so both forms give the same table.  The vertices above the
// thresholds are listed once per session.

#include "device.h"
#include "device_prims.h"
#include "session.h"

#include <algorithm>
#include <memory>

namespace glp {
using namespace dev;

// Row hashes of method/countmin.cuh:97-142 (the constants must match for
// count-min results to match).
__device__ __forceinline__ uint32_t cm_hash(int r, uint32_t k) {
    switch (r) {
        case 0:
            k ^= k >> 16; k *= 0x12ebcb3bu; k ^= k >> 13; k *= 0xc123ae53u; k ^= k >> 16;
            return k;
        case 1:
            k ^= k >> 16; k *= 0x21eb21bbu; k ^= k >> 13; k *= 0x2cb2ea35u; k ^= k >> 16;
            return k;
        case 2:
            k ^= k >> 16; k *= 0x8febce61u; k ^= k >> 13; k *= 0xa2b2be25u; k ^= k >> 16;
            return k;
        case 3:
            k ^= k >> 12; k *= 0x13ebdb3au; k ^= k >> 12; k *= 0xc123ae53u; k ^= k >> 14;
            return k;
        default:
            k ^= k >> 13; k *= 0x312b521bu; k ^= k >> 5; k *= 0x1cb3ea22u; k ^= k >> 17;
            return k;
    }
}

// Cells (row*W + column) a key touches, one entry per increment; the first
// `rows` entries are used.  All five are computed with constant indices so
// the array stays in registers.  Aliased rows (d = 3 or 4): row 3's hash
// lands in row 2, and for d = 4 h3 goes into row 2 again.
// x mod W by Lemire, Kaser and Kurz's direct remainder ("Faster remainder
// by direct computation", 2019): with M = floor((2^64 - 1) / W) + 1, the
// remainder of any 32-bit x is the high 64 bits of (M x mod 2^64) * W.
__device__ __forceinline__ uint32_t cm_mod(uint32_t x, uint64_t M, uint32_t W) {
    return (uint32_t)__umul64hi(M * (uint64_t)x, (uint64_t)W);
}
__host__ __device__ inline uint64_t cm_mod_magic(uint32_t W) { return ~0ull / W + 1ull; }

__device__ __forceinline__ int cm_cells(uint32_t key, int rows, int W, int aliased, uint64_t M, uint32_t* cell) {
#pragma unroll
    for (int r = 0; r < 5; ++r) cell[r] = (uint32_t)r * W + cm_mod(cm_hash(r, key), M, (uint32_t)W);
    if (aliased && (rows == 3 || rows == 4)) {
        cell[2] -= (uint32_t)W;
        cell[3] = cell[2];
    }
    return rows;
}

constexpr int32_t kCountminSmallDeg = 16;
constexpr int kCmSmallClasses = 5;  // lane groups of 1, 2, 4, 8, 16 for degrees 1, 2, 3-4, 5-8, 9-16
static_assert(1 << (kCmSmallClasses - 1) == kCountminSmallDeg, "the last class reaches kCountminSmallDeg");
constexpr int32_t kCountminWaveMaxDeg = 1024;
constexpr int32_t kCountminMidDeg = 4096;  // up to here 256 threads per vertex, then 1024
constexpr int32_t kCountminHubDeg = 8192;
constexpr int32_t kCountminChunk = 8192;  // 1024 threads x 8 edges
constexpr int32_t kTableLaneMaxDeg = 32;

// The vertices of `verts` (degrees 1..G), G lanes each (lane i holds edge
// i; G = 1, 2, 4, 8 or 16, so 64 / G vertices share a wavefront).  The
// sketch cell an entry c of edge i lands in ends up holding the sum, over
// the vertex's edges j and their entries, of f_j where the entry is c; the
// estimate is the minimum of that over edge i's entries.
template <bool W, int G>
__global__ __launch_bounds__(256) void k_countmin_small(StepArgs a, const int32_t* __restrict__ verts, int32_t nv,
                                                       int32_t rows, int32_t width, int32_t aliased, uint64_t M) {
    constexpr int kShift = G == 1 ? 0 : G == 2 ? 1 : G == 4 ? 2 : G == 8 ? 3 : 4;
    static_assert((1 << kShift) == G && G <= kCountminSmallDeg, "group of 1, 2, 4, 8 or 16 lanes");
    const int lane = lane_id(), ql = lane & (G - 1), g0 = lane & ~(G - 1);
    const int64_t vv = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> kShift;
    const bool in = vv < nv;
    const int32_t v = in ? verts[vv] : 0;
    int32_t b = 0, e = 0;
    if (in) {
        b = a.off[v];
        e = a.off[v + 1];
    }
    const bool live = in && ql < e - b;
    uint32_t key = 0u, f = 0u;  // a lane without an edge adds nothing (f = 0)
    if (live) {
        key = (uint32_t)a.L[a.nbr[b + ql]] + 1u;
        f = W ? a.w[b + ql] : 1u;
    }
    uint32_t c[5];
    const int k = cm_cells(key, rows, width, aliased, M, c);  // k depends on rows and aliased only
    // Proper rows occupy disjoint bands [r W, (r+1) W), so entry t of one
    // edge can only meet entry t of another; aliased rows put entries 1..3
    // in one band, so those are compared crosswise.  (Loops unrolled to the
    // 5 entries and guarded by k, so the arrays stay in registers.)
    const bool cross = aliased && (rows == 3 || rows == 4);
    // edges of the wave's groups: the loop runs to the largest degree among them
    const int32_t dmax = __builtin_amdgcn_readfirstlane(
        (int32_t)wave_max_u64(live ? (uint64_t)(e - b) : 0ull));
    uint32_t sum[5] = {0u, 0u, 0u, 0u, 0u};
    for (int jj = 0; jj < dmax; ++jj) {
        const uint32_t fj = shfl_u32(f, g0 + jj);
        uint32_t cj[5];
#pragma unroll
        for (int t = 0; t < 5; ++t) cj[t] = t < k ? shfl_u32(c[t], g0 + jj) : 0xffffffffu;
        if (!cross) {
#pragma unroll
            for (int t = 0; t < 5; ++t) sum[t] += t < k && cj[t] == c[t] ? fj : 0u;
        } else {
#pragma unroll
            for (int t = 0; t < 5; ++t)
#pragma unroll
                for (int t2 = 0; t2 < 5; ++t2) sum[t] += t < k && t2 < k && cj[t2] == c[t] ? fj : 0u;
        }
    }
    uint32_t est = 0xffffffffu;
#pragma unroll
    for (int t = 0; t < 5; ++t)
        if (t < k) est = min(est, sum[t]);
    uint64_t best = live ? pack_score((int32_t)min(est, 0x7fffffffu), key) : 0ull;
#pragma unroll
    for (int d = G / 2; d >= 1; d >>= 1) best = max_u64(best, shfl_u64(best, lane ^ d));
    bool changed = false;
    if (in && ql == 0) {
        const int32_t old = a.keep[v];
        const int32_t nl = best ? packed_label(best) : old;
        a.Lw[v] = nl;
        changed = nl != old;
    }
    const unsigned long long bal = __ballot(changed);
    if (lane == 0 && bal)
        atomicAdd(&a.changes[blockIdx.x & (kChangeSlots - 1)], (unsigned long long)__popcll(bal));
}

// One launch of k_countmin_small over n listed vertices of lane-group class
// gi (groups of 1 << gi lanes).
template <int G>
static void launch_countmin_small_g(bool w, const StepArgs& a, const int32_t* verts, int32_t n, int32_t rows,
                                    int32_t width, int32_t aliased, uint64_t M, hipStream_t s) {
    const unsigned grid = (unsigned)(((int64_t)n * G + 255) / 256);
    if (w) hipLaunchKernelGGL((k_countmin_small<true, G>), dim3(grid), dim3(256), 0, s, a, verts, n, rows, width,
                              aliased, M);
    else hipLaunchKernelGGL((k_countmin_small<false, G>), dim3(grid), dim3(256), 0, s, a, verts, n, rows, width,
                            aliased, M);
}
static void launch_countmin_small(int gi, bool w, const StepArgs& a, const int32_t* verts, int32_t n, int32_t rows,
                                  int32_t width, int32_t aliased, uint64_t M, hipStream_t s) {
    switch (gi) {
        case 0: launch_countmin_small_g<1>(w, a, verts, n, rows, width, aliased, M, s); break;
        case 1: launch_countmin_small_g<2>(w, a, verts, n, rows, width, aliased, M, s); break;
        case 2: launch_countmin_small_g<4>(w, a, verts, n, rows, width, aliased, M, s); break;
        case 3: launch_countmin_small_g<8>(w, a, verts, n, rows, width, aliased, M, s); break;
        default: launch_countmin_small_g<16>(w, a, verts, n, rows, width, aliased, M, s); break;
    }
}

// Vertices without edges keep their label.
__global__ __launch_bounds__(256) void k_countmin_keep(StepArgs a, const int32_t* __restrict__ verts, int32_t nv) {
    const int32_t i = (int32_t)(blockIdx.x * 256 + threadIdx.x);
    if (i < nv) {
        const int32_t v = verts[i];
        a.Lw[v] = a.keep[v];
    }
}

// Up to kCmBatch edges of a thread (j, j + NT, ...) loaded at once: the
// indices past the last edge are clamped to it (every vertex here has edges)
// and their keys never applied, so the loads need no branches and all are in
// flight together.
constexpr int kCmBatch = 4;

template <bool W, int NT>
__device__ __forceinline__ void cm_load_batch(const StepArgs& a, int32_t j, int32_t last, uint32_t* key,
                                              uint32_t* f) {
    int32_t id[kCmBatch];
#pragma unroll
    for (int u = 0; u < kCmBatch; ++u) id[u] = a.nbr[min(j + u * NT, last)];
#pragma unroll
    for (int u = 0; u < kCmBatch; ++u) key[u] = (uint32_t)a.L[id[u]] + 1u;
#pragma unroll
    for (int u = 0; u < kCmBatch; ++u) f[u] = W ? a.w[min(j + u * NT, last)] : 1u;
}

// Vertex verts[blockIdx.x], the sketch in LDS (NT = 64: degrees up to
// kCountminWaveMaxDeg; NT = 256: up to kCountminMidDeg; NT = 1024: up to
// kCountminHubDeg).  A thread's keys (at most KB of them) stay in registers
// from the sketch pass to the estimate pass, so each neighbour label is read
// once.
template <bool W, int NT>
__global__ __launch_bounds__(NT) void k_countmin(StepArgs a, const int32_t* __restrict__ verts, int32_t rows,
                                                 int32_t width, int32_t aliased, uint64_t M) {
    constexpr int KB =
        (NT == 64 ? kCountminWaveMaxDeg : NT == 256 ? kCountminMidDeg : kCountminHubDeg) / NT;
    static_assert(KB % kCmBatch == 0, "whole batches");
    extern __shared__ uint32_t cells[];
    __shared__ unsigned long long s_best[NT / 64];
    const int lane = lane_id();
    const int32_t v = verts[blockIdx.x];
    const int32_t b = a.off[v], e = a.off[v + 1];
    const int32_t last = e - 1;
    const int32_t ncell = rows * width;
    for (int32_t i = threadIdx.x; i < ncell; i += NT) cells[i] = 0u;
    __syncthreads();
    uint32_t key[KB];
#pragma unroll
    for (int q = 0; q < KB; q += kCmBatch) {
        if (b + q * NT > last) break;  // uniform: no edges left for any thread
        const int32_t j = b + (int32_t)threadIdx.x + q * NT;
        uint32_t f[kCmBatch];
        cm_load_batch<W, NT>(a, j, last, key + q, f);
#pragma unroll
        for (int u = 0; u < kCmBatch; ++u) {
            if (j + u * NT > last) break;
            uint32_t c[5];
            const int k = cm_cells(key[q + u], rows, width, aliased, M, c);
#pragma unroll
            for (int i = 0; i < 5; ++i)
                if (i < k) atomicAdd(&cells[c[i]], f[u]);
        }
    }
    __syncthreads();
    uint64_t best = 0ull;
#pragma unroll
    for (int q = 0; q < KB; q += kCmBatch) {
        if (b + q * NT > last) break;
        const int32_t j = b + (int32_t)threadIdx.x + q * NT;
#pragma unroll
        for (int u = 0; u < kCmBatch; ++u) {
            if (j + u * NT > last) break;
            uint32_t c[5];
            const int k = cm_cells(key[q + u], rows, width, aliased, M, c);
            uint32_t est = 0xffffffffu;
#pragma unroll
            for (int i = 0; i < 5; ++i)
                if (i < k) est = min(est, cells[c[i]]);
            best = max_u64(best, pack_score((int32_t)min(est, 0x7fffffffu), key[q + u]));
        }
    }
    best = wave_max_u64(best);
    if (NT > 64) {
        if (lane == 0) s_best[threadIdx.x >> 6] = best;
        __syncthreads();
        best = lane < NT / 64 ? s_best[lane] : 0ull;
        best = wave_max_u64(best);
    }
    if (threadIdx.x == 0) {
        const int32_t old = a.keep[v];
        const int32_t nl = best ? packed_label(best) : old;
        a.Lw[v] = nl;
        if (nl != old) atomicAdd(&a.changes[blockIdx.x & (kChangeSlots - 1)], 1ull);
    }
}

// Count-min hubs (degree above kCountminHubDeg): workgroup i takes chunk
// chunks[i] = (hub slot h, first edge of the chunk within the hub's edges),
// kCountminChunk edges.  `sketch` holds rows * width cells per hub and
// `best` one packed maximum per hub; both are zero between steps
// (k_countmin_hub_final clears them).  Sums of cells wrap modulo 2^32 as the
// single-workgroup form's LDS sums do, so the estimates are the same.  The
// sketch pass leaves each edge's key in `keys` (kCountminChunk slots per
// chunk, in edge order), which the estimate pass reads back in order instead
// of gathering the labels again.
template <bool W>
__global__ __launch_bounds__(1024) void k_countmin_hub_sketch(StepArgs a, const int32_t* __restrict__ hubs,
                                                              const int2* __restrict__ chunks,
                                                              uint32_t* __restrict__ sketch,
                                                              uint32_t* __restrict__ keys, int32_t rows,
                                                              int32_t width, int32_t aliased, uint64_t M) {
    extern __shared__ uint32_t cells[];
    const int2 ch = chunks[blockIdx.x];
    const int32_t v = hubs[ch.x];
    const int32_t b = a.off[v] + ch.y, e = min(a.off[v + 1], b + kCountminChunk);
    const int32_t last = e - 1;
    const int32_t ncell = rows * width;
    for (int32_t i = threadIdx.x; i < ncell; i += 1024) cells[i] = 0u;
    __syncthreads();
    uint32_t* ck = keys + (size_t)blockIdx.x * kCountminChunk;  // ck[j - b]: edge j's key
    for (int32_t j = b + (int32_t)threadIdx.x; j < e; j += 1024 * kCmBatch) {
        uint32_t key[kCmBatch], f[kCmBatch];
        cm_load_batch<W, 1024>(a, j, last, key, f);
#pragma unroll
        for (int u = 0; u < kCmBatch; ++u) {
            if (j + u * 1024 > last) break;
            ck[j - b + u * 1024] = key[u];
            uint32_t c[5];
            const int k = cm_cells(key[u], rows, width, aliased, M, c);
#pragma unroll
            for (int i = 0; i < 5; ++i)
                if (i < k) atomicAdd(&cells[c[i]], f[u]);
        }
    }
    __syncthreads();
    uint32_t* hs = sketch + (size_t)ch.x * ncell;
    for (int32_t i = threadIdx.x; i < ncell; i += 1024) {
        const uint32_t c = cells[i];
        if (c) atomicAdd(&hs[i], c);
    }
}

__global__ __launch_bounds__(1024) void k_countmin_hub_est(StepArgs a, const int32_t* __restrict__ hubs,
                                                           const int2* __restrict__ chunks,
                                                           const uint32_t* __restrict__ sketch,
                                                           const uint32_t* __restrict__ keys,
                                                           unsigned long long* __restrict__ best_out, int32_t rows,
                                                           int32_t width, int32_t aliased, uint64_t M) {
    extern __shared__ uint32_t cells[];
    __shared__ unsigned long long s_best[16];
    const int lane = lane_id();
    const int2 ch = chunks[blockIdx.x];
    const int32_t v = hubs[ch.x];
    const int32_t b = a.off[v] + ch.y, e = min(a.off[v + 1], b + kCountminChunk);
    const int32_t last = e - 1;
    const int32_t ncell = rows * width;
    const uint32_t* hs = sketch + (size_t)ch.x * ncell;
    for (int32_t i = threadIdx.x; i < ncell; i += 1024) cells[i] = hs[i];
    __syncthreads();
    const uint32_t* ck = keys + (size_t)blockIdx.x * kCountminChunk;
    uint64_t best = 0ull;
    for (int32_t j = b + (int32_t)threadIdx.x; j < e; j += 1024 * kCmBatch) {
        uint32_t key[kCmBatch];
#pragma unroll
        for (int u = 0; u < kCmBatch; ++u) key[u] = ck[min(j + u * 1024, last) - b];
#pragma unroll
        for (int u = 0; u < kCmBatch; ++u) {
            if (j + u * 1024 > last) break;
            uint32_t c[5];
            const int k = cm_cells(key[u], rows, width, aliased, M, c);
            uint32_t est = 0xffffffffu;
#pragma unroll
            for (int i = 0; i < 5; ++i)
                if (i < k) est = min(est, cells[c[i]]);
            best = max_u64(best, pack_score((int32_t)min(est, 0x7fffffffu), key[u]));
        }
    }
    best = wave_max_u64(best);
    if (lane == 0) s_best[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x < 64) {
        best = wave_max_u64(lane < 16 ? s_best[lane] : 0ull);
        if (lane == 0 && best) atomicMax(&best_out[ch.x], (unsigned long long)best);
    }
}

// One workgroup per count-min hub: commits its winner, then clears its
// sketch and maximum for the next step.
__global__ __launch_bounds__(256) void k_countmin_hub_final(StepArgs a, const int32_t* __restrict__ hubs,
                                                            uint32_t* __restrict__ sketch,
                                                            unsigned long long* __restrict__ best, int32_t ncell) {
    const int32_t h = (int32_t)blockIdx.x;
    uint32_t* hs = sketch + (size_t)h * ncell;
    for (int32_t i = threadIdx.x; i < ncell; i += 256) hs[i] = 0u;
    if (threadIdx.x == 0) {
        const int32_t v = hubs[h];
        const unsigned long long bb = best[h];
        const int32_t old = a.keep[v];
        const int32_t nl = bb ? packed_label(bb) : old;
        a.Lw[v] = nl;
        if (nl != old) atomicAdd(&a.changes[(uint32_t)h & (kChangeSlots - 1)], 1ull);
        best[h] = 0ull;
    }
}

// Min over the 64 lanes, every lane active, without the LDS pipe: DPP moves
// within each row of 16 (quad pairs, quad with mirror quad, half-row with
// half-row), then the four rows' minima by readlane.  A lane whose DPP
// source is unavailable keeps its own value (old = x).  The full-table
// case of k_table_wave takes one per edge, in series, so its latency sets
// the wide tables' speed.
__device__ __forceinline__ uint32_t wave_min_dpp(uint32_t x) {
    x = min(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0xB1, 0xf, 0xf, false));   // quad_perm [1,0,3,2]
    x = min(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x4E, 0xf, 0xf, false));   // quad_perm [2,3,0,1]
    x = min(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x141, 0xf, 0xf, false));  // row half-mirror
    x = min(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x140, 0xf, 0xf, false));  // row mirror
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)x, 0);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)x, 16);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)x, 32);
    const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)x, 48);
    return min(min(r0, r1), min(r2, r3));
}

// One wavefront per vertex of `verts` (degrees above kTableLaneMaxDeg):
// slot i of the vertex's table is lane i's (K, C); lanes >= cap hold the key
// 0xffffffff, which no edge carries (keys are labels + 1 <= 2^31), so they
// never match and are never free.  Edges are applied one after another in
// CSR order; a table's keys are distinct, so at most one lane hits.  The
// vertex and its edge range are wave-uniform (scalar loop control), and
// edges come in chunks of 64, one per lane: while chunk c is applied, chunk
// c+1's labels and chunk c+2's neighbour ids are in flight.  Lanes past the
// last edge load the last edge again (in bounds, never applied), so the
// loads need no branches and stay in flight across the chunk.
template <bool W, bool BUF>
__global__ __launch_bounds__(256) void k_table_wave(StepArgs a, const int32_t* __restrict__ verts, int32_t nv,
                                                    int32_t cap) {
    const int lane = lane_id();
    const int32_t i = (int32_t)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= nv) return;  // wave-uniform
    const int32_t v = __builtin_amdgcn_readfirstlane(verts[i]);
    const int32_t b = __builtin_amdgcn_readfirstlane(a.off[v]);
    const int32_t e = __builtin_amdgcn_readfirstlane(a.off[v + 1]);
    const bool slot = lane < cap;
    const uint64_t slots = __ballot(slot);
    uint32_t K = slot ? 0u : 0xffffffffu, C = 0u;
    const int32_t last = e - 1;  // e > b: these vertices have more than 32 edges
    int32_t jn = b + lane;
    int32_t id = a.nbr[min(jn, last)];
    uint32_t lab = (uint32_t)a.L[id];
    uint32_t f = W ? a.w[min(jn, last)] : 1u;
    jn += 64;
    id = a.nbr[min(jn, last)];
    for (int32_t j0 = b; j0 < e; j0 += 64) {
        const uint32_t nlab = (uint32_t)a.L[id];
        const uint32_t nf = W ? a.w[min(jn, last)] : 1u;
        jn += 64;
        id = a.nbr[min(jn, last)];
        const int32_t cnt = __builtin_amdgcn_readfirstlane(min(64, e - j0));  // held in a scalar register
        if constexpr (!W) {
            // Unweighted, without branches: the key's slot (the one holding
            // it, else the lowest free one) gains one and holds it, else
            // (full) every slot drops by one and the emptied ones are freed.
            // Frequent and buffer agree here (every count >= 1, frequency 1).
            // The two slot sets are wave-uniform masks, read back per lane.
            // A full chunk runs unrolled (no loop control per edge).
            const uint32_t key = lab + 1u;
            auto apply = [&](uint32_t kk) {
                const uint64_t hm = __ballot(K == kk), fm = __ballot(K == 0u);
                const uint64_t upm = hm ? hm : fm & (0ull - fm);
                const uint64_t dnm = hm | fm ? 0ull : slots;
                const bool up = __builtin_amdgcn_inverse_ballot_w64(upm);
                const bool dn = __builtin_amdgcn_inverse_ballot_w64(dnm);
                C = C + (uint32_t)up - (uint32_t)dn;
                K = up ? kk : K;
                K = dn && C == 0u ? 0u : K;
            };
            if (cnt == 64) {
#pragma unroll
                for (int k = 0; k < 64; ++k) apply((uint32_t)__builtin_amdgcn_readlane((int)key, k));
            } else {
                for (int32_t k = 0; k < cnt; ++k) apply((uint32_t)__builtin_amdgcn_readlane((int)key, k));
            }
        } else {
            for (int32_t k = 0; k < cnt; ++k) {
                const uint32_t kk = (uint32_t)__builtin_amdgcn_readlane((int)lab, k) + 1u;
                const uint32_t ff = (uint32_t)__builtin_amdgcn_readlane((int)f, k);
                const bool hit = K == kk;
                if (__ballot(hit)) {  // present: it gains the frequency
                    if (hit) C += ff;
                    continue;
                }
                const uint64_t fr = __ballot(K == 0u);
                if (fr) {  // the first free slot takes it
                    if (lane == __builtin_ctzll(fr)) {
                        K = kk;
                        C = ff;
                    }
                    continue;
                }
                // full: every counter drops (buffer: by the frequency;
                // frequent: by min(frequency, smallest counter)), emptied
                // slots are freed
                uint32_t d = ff;
                if (!BUF) d = min(ff, wave_min_dpp(slot ? C : 0xffffffffu));
                if (slot) C = BUF ? (C > ff ? C - ff : 0u) : C - d;
                const bool emptied = slot && C == 0u;
                if (emptied) K = 0u;
                if (!BUF && ff > d) {  // frequent: what is left enters the first freed slot
                    if (lane == __builtin_ctzll(__ballot(emptied))) {
                        K = kk;
                        C = ff - d;
                    }
                }
            }
        }
        lab = nlab;
        f = nf;
    }
    const uint64_t best = wave_max_u64(slot && K ? pack_score((int32_t)min(C, 0x7fffffffu), K) : 0ull);
    if (lane == 0) {
        const int32_t old = a.keep[v];
        const int32_t nl = best ? packed_label(best) : old;
        a.Lw[v] = nl;
        if (nl != old) atomicAdd(&a.changes[(uint32_t)i & (kChangeSlots - 1)], 1ull);
    }
}

// One lane per vertex of degree <= kTableLaneMaxDeg; the lane's table lives
// in LDS (cap key and cap count slots per lane), slot-major (slot i of lane l
// at word i * 64 + l), so the lanes' accesses to one slot fall in different
// banks.  The next edge's label loads while an edge is applied.
template <bool W, bool BUF>
__global__ __launch_bounds__(64) void k_table(StepArgs a, int32_t cap) {
    extern __shared__ uint32_t tbl[];
    const int lane = lane_id();
    const int32_t v = (int32_t)(blockIdx.x * 64 + lane);
    uint32_t* const K = tbl + lane;             // K[i * 64]: slot i's key
    uint32_t* const C = tbl + cap * 64 + lane;  // C[i * 64]: slot i's count
    bool changed = false;
    if (v < a.n && a.off[v + 1] - a.off[v] <= kTableLaneMaxDeg) {
        for (int i = 0; i < cap; ++i) { K[i * 64] = 0u; C[i * 64] = 0u; }
        const int32_t b = a.off[v], e = a.off[v + 1];
        uint32_t nkey = b < e ? (uint32_t)a.L[a.nbr[b]] + 1u : 0u;
        for (int32_t j = b; j < e; ++j) {
            const uint32_t key = nkey;
            if (j + 1 < e) nkey = (uint32_t)a.L[a.nbr[j + 1]] + 1u;
            const uint32_t f = W ? a.w[j] : 1u;
            int hit = -1, fr = -1;
            uint32_t cmin = 0xffffffffu;
            for (int i = 0; i < cap; ++i) {
                const uint32_t k = K[i * 64];
                if (k == key) hit = i;
                if (k == 0u) { if (fr < 0) fr = i; }
                else cmin = min(cmin, C[i * 64]);
            }
            if (hit >= 0) { C[hit * 64] += f; continue; }
            if (fr >= 0) { K[fr * 64] = key; C[fr * 64] = f; continue; }
            // table full
            const uint32_t d = min(f, cmin);
            int freed = -1;
            for (int i = 0; i < cap; ++i) {
                const uint32_t c = C[i * 64];
                const uint32_t nc = BUF ? (c > f ? c - f : 0u) : c - d;
                C[i * 64] = nc;
                if (nc == 0u) { K[i * 64] = 0u; if (freed < 0) freed = i; }
            }
            if (!BUF && f > d) { K[freed * 64] = key; C[freed * 64] = f - d; }
        }
        uint64_t best = 0ull;
        for (int i = 0; i < cap; ++i)
            if (K[i * 64]) best = max_u64(best, pack_score((int32_t)min(C[i * 64], 0x7fffffffu), K[i * 64]));
        const int32_t old = a.keep[v];
        const int32_t nl = best ? packed_label(best) : old;
        a.Lw[v] = nl;
        changed = nl != old;
    }
    unsigned long long bal = __ballot(changed);
    if (lane == 0 && bal)
        atomicAdd(&a.changes[blockIdx.x & (kChangeSlots - 1)], (unsigned long long)__popcll(bal));
}

// k_countmin takes rows * width * 4 bytes of dynamic LDS, up to 5 * 4096 * 4
// = 80 KB (the limits check_params enforces), above the 64 KB a launch gets
// without the attribute; k_table takes at most 64 * 2 * 64 * 4 = 32 KB.
void init_counting_kernels() {
    const int lim = 5 * 4096 * 4;
    (void)hipFuncSetAttribute((const void*)k_countmin<false, 64>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void*)k_countmin<true, 64>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void*)k_countmin<false, 256>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lim);
    (void)hipFuncSetAttribute((const void*)k_countmin<true, 256>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void*)k_countmin<false, 1024>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lim);
    (void)hipFuncSetAttribute((const void*)k_countmin<true, 1024>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lim);
    (void)hipFuncSetAttribute((const void*)k_countmin_hub_sketch<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lim);
    (void)hipFuncSetAttribute((const void*)k_countmin_hub_sketch<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lim);
    (void)hipFuncSetAttribute((const void*)k_countmin_hub_est, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
}

// The counting methods' vertex lists (built once per session), in one
// array: count-min, the vertices of degree in (kCountminSmallDeg,
// kCountminWaveMaxDeg] (nmid of them), then those up to kCountminHubDeg
// (nwide: the ncm256 up to kCountminMidDeg first), then the hubs above (ncmhub, with their chunk list, sketches and
// maxima), then degrees 1..kCountminSmallDeg (nsmall) and 0 (niso);
// frequent and buffer, the vertices above kTableLaneMaxDeg (nwide).  The
// lists of one workgroup or wavefront per vertex, and the hubs, run largest
// degree first, so the longest counts start at once whatever the vertex ids.
int counting_wide_list(glp_session* s) {
    const HostCSR& g = s->g->csr;
    const bool cm = s->p.method == GLP_COUNTMIN;
    const int32_t lim = cm ? kCountminWaveMaxDeg : kTableLaneMaxDeg;
    std::vector<int32_t> wide, big, big1024, hubs, iso;
    std::vector<int32_t> small[kCountminSmallDeg + 1];  // by degree, so a wavefront's groups have alike degrees
    std::vector<int2> chunks;
    for (int32_t v = 0; v < g.n; ++v) {
        const int32_t d = g.offsets[v + 1] - g.offsets[v];
        if (cm && d > kCountminHubDeg) hubs.push_back(v);
        else if (cm && d > kCountminMidDeg) big1024.push_back(v);
        else if (d > lim) big.push_back(v);
        else if (cm && d > kCountminSmallDeg) wide.push_back(v);
        else if (cm && d > 0) small[d].push_back(v);
        else if (cm) iso.push_back(v);
    }
    auto deg = [&g](int32_t v) { return g.offsets[v + 1] - g.offsets[v]; };
    for (auto* list : {&wide, &big, &big1024, &hubs})
        std::stable_sort(list->begin(), list->end(), [&](int32_t x, int32_t y) { return deg(x) > deg(y); });
    for (int32_t h = 0; h < (int32_t)hubs.size(); ++h)
        for (int32_t c0 = 0; c0 < deg(hubs[h]); c0 += kCountminChunk) chunks.push_back(make_int2(h, c0));
    s->nmid = (int32_t)wide.size();
    s->ncm256 = (int32_t)big.size();
    big.insert(big.end(), big1024.begin(), big1024.end());
    s->nwide = (int32_t)big.size();
    s->ncmhub = (int32_t)hubs.size();
    s->ncmchunk = (int32_t)chunks.size();
    s->niso = (int32_t)iso.size();
    wide.insert(wide.end(), big.begin(), big.end());
    wide.insert(wide.end(), hubs.begin(), hubs.end());
    for (const auto& sd : small) wide.insert(wide.end(), sd.begin(), sd.end());
    s->nsmall = (int32_t)(wide.size() - s->nmid - s->nwide - s->ncmhub);
    // the small list's lane-group classes: degree 1, 2, 3-4, 5-8, 9-16
    for (int gi = 0, d = 1; gi < kCmSmallClasses; ++gi) {
        s->nsmall_g[gi] = 0;
        for (; d <= (1 << gi); ++d) s->nsmall_g[gi] += (int32_t)small[d].size();
    }
    wide.insert(wide.end(), iso.begin(), iso.end());
    if (wide.empty()) return GLP_OK;
    GLP_HIP(hipMalloc(&s->wide, wide.size() * 4));
    GLP_HIP(hipMemcpyAsync(s->wide, wide.data(), wide.size() * 4, hipMemcpyHostToDevice, s->stream));
    s->bytes += wide.size() * 4;
    if (!hubs.empty()) {
        const size_t sk = (size_t)hubs.size() * s->p.cm_rows * s->p.cm_width * 4;
        GLP_HIP(hipMalloc(&s->cm_chunks, chunks.size() * sizeof(int2)));
        GLP_HIP(hipMalloc(&s->cm_sketch, sk));
        GLP_HIP(hipMalloc(&s->cm_best, hubs.size() * 8));
        GLP_HIP(hipMalloc(&s->cm_keys, chunks.size() * (size_t)kCountminChunk * 4));
        GLP_HIP(hipMemcpyAsync(s->cm_chunks, chunks.data(), chunks.size() * sizeof(int2), hipMemcpyHostToDevice,
                               s->stream));
        GLP_HIP(hipMemsetAsync(s->cm_sketch, 0, sk, s->stream));
        GLP_HIP(hipMemsetAsync(s->cm_best, 0, hubs.size() * 8, s->stream));
        s->bytes += chunks.size() * (sizeof(int2) + (size_t)kCountminChunk * 4) + sk + hubs.size() * 8;
    }
    GLP_HIP(hipStreamSynchronize(s->stream));  // the host vectors are released on return
    return GLP_OK;
}

int enqueue_counting_step(glp_session* s) {
    const DeviceGraph& dg = *s->g->dev;
    StepArgs a{};
    a.off = dg.off;
    a.nbr = dg.nbr;
    a.w = dg.w;
    a.n = s->n;
    a.t = s->t + 1;
    a.changes = s->d_acc;
    a.L = s->L[s->cur];
    a.keep = s->L[s->cur];
    a.Lw = s->L[s->cur ^ 1];
    const bool w = s->p.weighted != 0;
    if (s->n > 0) {
        if (s->p.method == GLP_COUNTMIN) {
            const size_t lds = (size_t)s->p.cm_rows * s->p.cm_width * 4;
            const int32_t r = s->p.cm_rows, wd = s->p.cm_width, al = s->p.cm_aliased;
            const uint64_t M = cm_mod_magic((uint32_t)wd);
            const int32_t* small = s->wide + s->nmid + s->nwide + s->ncmhub;
            for (int gi = 0, at = 0; gi < kCmSmallClasses; at += s->nsmall_g[gi], ++gi)
                if (s->nsmall_g[gi] > 0)
                    launch_countmin_small(gi, w, a, small + at, s->nsmall_g[gi], r, wd, al, M, s->stream);
            if (s->niso > 0)
                hipLaunchKernelGGL(k_countmin_keep, dim3((unsigned)((s->niso + 255) / 256)), dim3(256), 0, s->stream,
                                   a, small + s->nsmall, s->niso);
            if (s->nmid > 0) {
                if (w) hipLaunchKernelGGL((k_countmin<true, 64>), dim3(s->nmid), dim3(64), lds, s->stream, a,
                                          s->wide, r, wd, al, M);
                else hipLaunchKernelGGL((k_countmin<false, 64>), dim3(s->nmid), dim3(64), lds, s->stream, a,
                                        s->wide, r, wd, al, M);
            }
            if (s->ncm256 > 0) {
                if (w) hipLaunchKernelGGL((k_countmin<true, 256>), dim3(s->ncm256), dim3(256), lds, s->stream,
                                          a, s->wide + s->nmid, r, wd, al, M);
                else hipLaunchKernelGGL((k_countmin<false, 256>), dim3(s->ncm256), dim3(256), lds, s->stream,
                                        a, s->wide + s->nmid, r, wd, al, M);
            }
            if (s->nwide > s->ncm256) {
                const unsigned g1024 = (unsigned)(s->nwide - s->ncm256);
                const int32_t* v1024 = s->wide + s->nmid + s->ncm256;
                if (w) hipLaunchKernelGGL((k_countmin<true, 1024>), dim3(g1024), dim3(1024), lds, s->stream,
                                          a, v1024, r, wd, al, M);
                else hipLaunchKernelGGL((k_countmin<false, 1024>), dim3(g1024), dim3(1024), lds, s->stream,
                                        a, v1024, r, wd, al, M);
            }
            if (s->ncmhub > 0) {
                const int32_t* hubs = s->wide + s->nmid + s->nwide;
                const unsigned cgrid = (unsigned)s->ncmchunk;
                if (w)
                    hipLaunchKernelGGL((k_countmin_hub_sketch<true>), dim3(cgrid), dim3(1024), lds, s->stream, a,
                                       hubs, s->cm_chunks, s->cm_sketch, s->cm_keys, r, wd, al, M);
                else
                    hipLaunchKernelGGL((k_countmin_hub_sketch<false>), dim3(cgrid), dim3(1024), lds, s->stream, a,
                                       hubs, s->cm_chunks, s->cm_sketch, s->cm_keys, r, wd, al, M);
                hipLaunchKernelGGL(k_countmin_hub_est, dim3(cgrid), dim3(1024), lds, s->stream, a, hubs,
                                   s->cm_chunks, s->cm_sketch, s->cm_keys, s->cm_best, r, wd, al, M);
                hipLaunchKernelGGL(k_countmin_hub_final, dim3(s->ncmhub), dim3(256), 0, s->stream, a, hubs,
                                   s->cm_sketch, s->cm_best, r * wd);
            }
        } else {
            const size_t lds = (size_t)64 * 2 * s->p.table_cap * 4;
            const unsigned grid = (unsigned)((s->n + 63) / 64);
            const bool buf = s->p.method == GLP_BUFFER;
            const int32_t cap = s->p.table_cap;
            const unsigned wgrid = (unsigned)((s->nwide + 3) / 4);
            if (buf) {
                if (w) hipLaunchKernelGGL((k_table<true, true>), dim3(grid), dim3(64), lds, s->stream, a, cap);
                else hipLaunchKernelGGL((k_table<false, true>), dim3(grid), dim3(64), lds, s->stream, a, cap);
                if (s->nwide > 0) {
                    if (w) hipLaunchKernelGGL((k_table_wave<true, true>), dim3(wgrid), dim3(256), 0, s->stream, a,
                                              s->wide, s->nwide, cap);
                    else hipLaunchKernelGGL((k_table_wave<false, true>), dim3(wgrid), dim3(256), 0, s->stream, a,
                                            s->wide, s->nwide, cap);
                }
            } else {
                if (w) hipLaunchKernelGGL((k_table<true, false>), dim3(grid), dim3(64), lds, s->stream, a, cap);
                else hipLaunchKernelGGL((k_table<false, false>), dim3(grid), dim3(64), lds, s->stream, a, cap);
                if (s->nwide > 0) {
                    if (w) hipLaunchKernelGGL((k_table_wave<true, false>), dim3(wgrid), dim3(256), 0, s->stream,
                                              a, s->wide, s->nwide, cap);
                    else hipLaunchKernelGGL((k_table_wave<false, false>), dim3(wgrid), dim3(256), 0, s->stream,
                                            a, s->wide, s->nwide, cap);
                }
            }
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return set_error(GLP_E_HIP, "counting launch: %s", hipGetErrorString(e));
    }
    int rc = launch_end_step(s->d_acc, s->d_ring, s->d_ctr, kChangeRing, s->stream);
    if (rc != GLP_OK) return rc;
    s->cur ^= 1;
    s->t += 1;
    return GLP_OK;
}

// SLPA post-processing (builder-defined; the reference's SLPA keeps one
// memory slot and has no threshold): for vertex v the labels whose count in
// M_v[0..t] is >= r*(t+1), ascending, and the dominant label (highest count,
// then highest label).
__global__ __launch_bounds__(64) void k_slpa_post(const int32_t* __restrict__ mem, int32_t n,
                                                  int32_t rows, int32_t r_num, int32_t r_den,
                                                  int32_t kmax, int32_t* dominant, int32_t* nsets,
                                                  int32_t* sets) {
    const int32_t v = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    int32_t vals[256];
    for (int32_t i = 0; i < rows; ++i) vals[i] = mem[(size_t)i * n + v];
    for (int32_t i = 1; i < rows; ++i) {  // insertion sort, ascending
        const int32_t x = vals[i];
        int32_t j = i - 1;
        while (j >= 0 && vals[j] > x) { vals[j + 1] = vals[j]; --j; }
        vals[j + 1] = x;
    }
    uint64_t best = 0ull;
    int32_t kept = 0;
    for (int32_t i = 0; i < rows;) {
        int32_t j = i;
        while (j < rows && vals[j] == vals[i]) ++j;
        const int32_t c = j - i;
        best = max(best, ((uint64_t)(uint32_t)c << 32) | ((uint32_t)vals[i] + 1u));
        if ((int64_t)c * r_den >= (int64_t)r_num * rows) {
            if (kept < kmax) sets[(size_t)v * kmax + kept] = vals[i];
            ++kept;
        }
        i = j;
    }
    for (int32_t k = kept; k < kmax; ++k) sets[(size_t)v * kmax + k] = -1;
    if (dominant) dominant[v] = packed_label(best);
    if (nsets) nsets[v] = kept;
}

int slpa_post(glp_session* s, int32_t kmax, int32_t* dominant, int32_t* nsets, int32_t* sets) {
    const int32_t n = s->n;
    if (n == 0) return GLP_OK;
    const int32_t rows = s->t + 1;
    int32_t *d_dom = nullptr, *d_ns = nullptr, *d_sets = nullptr;
    struct Guard {
        int32_t** p[3];
        ~Guard() { for (auto q : p) (void)hipFree(*q); }
    } guard{{&d_dom, &d_ns, &d_sets}};
    GLP_HIP(hipMalloc(&d_dom, (size_t)n * 4));
    GLP_HIP(hipMalloc(&d_ns, (size_t)n * 4));
    if (kmax > 0) GLP_HIP(hipMalloc(&d_sets, (size_t)n * kmax * 4));
    hipLaunchKernelGGL(k_slpa_post, dim3((n + 63) / 64), dim3(64), 0, s->stream, s->mem, n, rows,
                       s->p.slpa_r_num, s->p.slpa_r_den, kmax, d_dom, d_ns, d_sets);
    GLP_HIP(hipGetLastError());
    GLP_HIP(hipStreamSynchronize(s->stream));
    if (dominant) GLP_HIP(hipMemcpy(dominant, d_dom, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (nsets) GLP_HIP(hipMemcpy(nsets, d_ns, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (kmax > 0 && sets) GLP_HIP(hipMemcpy(sets, d_sets, (size_t)n * kmax * 4, hipMemcpyDeviceToHost));
    return GLP_OK;
}

}  // namespace glp
