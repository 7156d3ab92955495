// Degree-tiered LP counting kernels for gfx950 (wave64, 160 KB LDS per CU).
//
// One step computes, for every vertex v, the packed maximum over the labels
// l present among its neighbours of LabelScore(v, l, C_v(l)) with
// C_v(l) = sum of the frequencies LoadNeighbor assigns to v's edges showing l
// (the reference's per-tier kernels header/kernel.cuh:864-911 (huge),
// :1655-1739 (big), :3124-3222 (medium), :4365-4426 (small), :5187-5234
// (tiny)).  The work is split by degree:
//   packed  deg <= pack_max   several vertices share one wavefront: <= 64
//                             edges, a two-candidate majority test per
//                             vertex, else an (owner, label) LDS table and
//                             a per-owner max
//   half    deg <= half_max   16 or 32 lanes per vertex (<= 64 edges), the
//                             same majority test, else an LDS table per lane
//                             group
//   wave    deg <= wave_max   one wavefront per vertex, LDS hash table
//   block   deg <= block_max  one workgroup per vertex, LDS hash table sized
//                             by the vertex's degree class
//   hub     deg >  block_max  several workgroups per vertex: each counts a
//                             chunk in LDS and writes its distinct
//                             (label,count) pairs; a counting sort groups the
//                             hub's pairs by label-hash bucket and one
//                             workgroup per bucket sums them in LDS, so no
//                             HBM hash table (and no memset) is needed
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

// Inserts key into an LDS table of (key << 32 | sum) slots and returns the
// slot, with kClaimed set when this insert created it.  A key claims a free
// slot with `first` already in it; a key that finds its slot adds `more` to
// the low word (frequency sums stay below 2^31, and slot-scored sums below
// 2^32 from a start above 0, so they never carry into the key).  The table
// always has a free slot.
constexpr uint32_t kClaimed = 0x80000000u;
__device__ __forceinline__ uint32_t kc_insert(unsigned long long* T, uint32_t mask, uint32_t key,
                                              uint32_t first, uint32_t more) {
    uint32_t h = fmix32(key) & mask;
    const unsigned long long mine = (unsigned long long)key << 32;
    // the claim is tried without a read first: one LDS operation when the
    // slot is free, and no dearer than the read when it is not (measured
    // faster on diverse and on converged steps alike)
    for (;;) {
        const unsigned long long cur = atomicCAS(&T[h], 0ull, mine | first);
        if (cur == 0ull) return h | kClaimed;
        if ((uint32_t)(cur >> 32) == key) {
            atomicAdd(&T[h], (unsigned long long)more);
            return h;
        }
        h = (h + 1u) & mask;
    }
}
__device__ __forceinline__ uint32_t kc_insert(unsigned long long* T, uint32_t mask, uint32_t key, uint32_t f) {
    return kc_insert(T, mask, key, f, f);
}

// Inserts one (key, f) per lane (key 0: nothing) into an LDS table; all 64
// lanes of the wave must be active.  Once labels converge most of a vertex's
// edges show the same label and 64 atomics on one LDS address serialise, so
// up to kPeels times the lanes holding the first pending lane's key are
// merged into one insert of their summed frequency (sums wrap mod 2^32 as
// the table's atomics do, so the counts are unchanged).
constexpr int kPeels = 2;
template <class H, bool W, bool S>
__device__ __forceinline__ void wave_insert(const StepArgs& a, unsigned long long* T, uint32_t mask, uint32_t key,
                                            uint32_t f, int32_t vol) {
    const int lane = lane_id();
    uint32_t ik = key, iff = f;  // what this lane finally inserts
#pragma unroll
    for (int r = 0; r < kPeels; ++r) {
        const unsigned long long rest = __ballot(key != 0u);
        if (rest == 0ull) break;
        const int first = (int)__builtin_ctzll(rest);
        const uint32_t lead = (uint32_t)__builtin_amdgcn_readlane((int)key, first);
        const bool same = key == lead;
        const unsigned long long sm = __ballot(same);
        if (__popcll(sm) < 2) break;  // a singleton gains nothing
        const uint32_t tot = W ? wave_sum_u32(same ? f : 0u) : (uint32_t)__popcll(sm);
        if (same) {
            ik = lane == first ? lead : 0u;
            iff = tot;
            key = 0u;
        }
    }
    // merged lanes show one label, so they share its volume
    if (ik) kc_insert(T, mask, ik, slot_first<H, S>(a, iff, vol), slot_more<H, S>(a, iff));
}

// Best packed score over the occupied slots of an LDS table of `cap` slots,
// `stride` threads cooperating (thread tid takes slots tid, tid + stride,
// ...); `clear` empties the slots read.  For hooks whose score reads a
// per-label volume from HBM the volumes of kScanU slots are loaded first,
// branch-free (an empty slot reads V[0]), so those loads are in flight
// together instead of one after another.
constexpr int kScanU = 8;
template <class H, bool S = false>
__device__ __forceinline__ uint64_t table_best(const StepArgs& a, int32_t v, unsigned long long* T,
                                               uint32_t cap, uint32_t tid, uint32_t stride,
                                               bool clear) {
    uint64_t best = 0ull;
    if constexpr (S) {
        // slots hold the scores themselves
        for (uint32_t i = tid; i < cap; i += stride) {
            const unsigned long long x = T[i];
            if (!x) continue;
            best = max_u64(best, H::slot_score((uint32_t)(x >> 32), (uint32_t)x));
            if (clear) T[i] = 0ull;
        }
    } else if constexpr (volume_of<H>::value) {
        for (uint32_t i0 = tid; i0 < cap; i0 += stride * kScanU) {
            unsigned long long x[kScanU];
            int32_t vol[kScanU];
#pragma unroll
            for (int u = 0; u < kScanU; ++u) {
                const uint32_t i = i0 + u * stride;
                x[u] = i < cap ? T[i] : 0ull;
            }
#pragma unroll
            for (int u = 0; u < kScanU; ++u) vol[u] = H::label_volume(a, x[u] ? (uint32_t)(x[u] >> 32) : 1u);
#pragma unroll
            for (int u = 0; u < kScanU; ++u) {
                if (!x[u]) continue;
                best = max_u64(best, H::label_score_vol(a, (uint32_t)(x[u] >> 32), (uint32_t)x[u], vol[u]));
                if (clear) T[i0 + u * stride] = 0ull;
            }
        }
    } else {
        for (uint32_t i = tid; i < cap; i += stride) {
            const unsigned long long x = T[i];
            if (!x) continue;
            best = max_u64(best, H::label_score(a, v, (uint32_t)(x >> 32), (uint32_t)x));
            if (clear) T[i] = 0ull;
        }
    }
    return best;
}

// Counts edges [b, e) of vertex v into an LDS table, STRIDE threads
// cooperating (thread `tid`; STRIDE a multiple of 64, whole waves).  Each
// thread first issues the loads of K edges (neighbour ids, then their
// labels) and only then inserts them, so K gathers are in flight per thread
// instead of one.  The loop is wave-uniform (wave_insert needs every lane).
// S: the slots accumulate scores (a slot-scoring hook set) instead of
// frequencies.
template <class H, bool W, bool G, int K, int STRIDE, bool S = false>
__device__ __forceinline__ void count_edges(const StepArgs& a, int32_t v, int32_t b, int32_t e,
                                            int tid, unsigned long long* T, uint32_t mask) {
    static_assert(STRIDE % 64 == 0, "whole waves");
    const int lane = tid & 63;
    for (int32_t jw = b + (tid & ~63); jw < e; jw += STRIDE * K) {
        uint32_t key[K], fr[K];
        int32_t vol[K];
        // branch-free: an index past the end is clamped to the last edge (a
        // valid address) and its key dropped afterwards, so the K neighbour
        // loads and then the K label loads issue back to back; a guarded load
        // would wait for every older load before issuing the next
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const int32_t j = jw + lane + i * STRIDE;
            key[i] = load_edge<H, W, G>(a, v, min(j, e - 1), &fr[i], &vol[i]);
        }
#pragma unroll
        for (int i = 0; i < K; ++i)
            if (jw + lane + i * STRIDE >= e) key[i] = 0u;
#pragma unroll
        for (int i = 0; i < K; ++i) wave_insert<H, W, S>(a, T, mask, key[i], fr[i], vol[i]);
    }
}

// Whether v sets its bit in the changed-vertex bitmap (frontier gather
// only): its new label differs from the previous step's, or, with the
// two-step lookahead, from the label of two steps back, which the buffer
// this step writes still holds (so this is called before the write).
__device__ __forceinline__ bool gather_changed(const StepArgs& a, int32_t v, int32_t nl, int32_t old) {
    if (!a.chg_w) return false;
    return a.lookahead ? nl != a.Lw[v] : nl != old;
}
// Sets v's bit in the changed-vertex bitmap (frontier gather only).
__device__ __forceinline__ void note_changed(const StepArgs& a, int32_t v) {
    if (a.chg_w) atomicOr(&a.chg_w[v >> 5], 1u << (v & 31));
}
// The same for vertices v0 + i, i a set bit of `bal` (one wave's ballot);
// one lane ORs the <= 3 words they span.
__device__ __forceinline__ void note_changed_run(const StepArgs& a, int32_t v0,
                                                 unsigned long long bal) {
    if (!a.chg_w || bal == 0ull || lane_id() != 0) return;
    const int32_t w0 = v0 >> 5;
    const uint32_t sh = (uint32_t)v0 & 31u;
    const unsigned long long lo = bal << sh;
    const uint32_t hi = sh ? (uint32_t)(bal >> (64u - sh)) : 0u;
    if ((uint32_t)lo) atomicOr(&a.chg_w[w0], (uint32_t)lo);
    if ((uint32_t)(lo >> 32)) atomicOr(&a.chg_w[w0 + 1], (uint32_t)(lo >> 32));
    if (hi) atomicOr(&a.chg_w[w0 + 2], hi);
}

// Adds n changed labels (lane 0's n) to one of kChangeSlots counters chosen
// by workgroup and wave, so no single address takes every wave's atomic.
__device__ __forceinline__ void add_changes(const StepArgs& a, uint32_t n) {
    if (lane_id() == 0 && n)
        atomicAdd(&a.changes[(blockIdx.x * 16u + (threadIdx.x >> 6)) & (kChangeSlots - 1)],
                  (unsigned long long)n);
}
__device__ __forceinline__ void count_changes(const StepArgs& a, bool changed) {
    add_changes(a, (uint32_t)__popcll(__ballot(changed)));
}

// Frontier helpers (classic LP; header/kernel.cuh:21-35's if_update idea):
// with the frontier on, a vertex whose neighbours all kept their labels in
// the previous step keeps its own and is not recomputed; its buffers already
// agree (a vertex that changed marks itself).  A changed vertex marks itself
// and its neighbours for the next step.
template <class H>
__device__ __forceinline__ bool frontier_on(const StepArgs& a) {
    return H::kFrontier && a.act_r != nullptr && a.fctl[0] == 0;
}
template <class H>
__device__ __forceinline__ bool is_active(const StepArgs& a, int32_t v) {
    return !frontier_on<H>(a) || a.act_r[v] != 0;
}
// Seeded LP: a seed keeps its own label and is never recomputed.
template <class H>
__device__ __forceinline__ bool is_fixed(const StepArgs& a, int32_t v) {
    if constexpr (seeded_of<H>::value) return a.seeds[v] != 0;
    else return false;
}
template <class H>
__device__ __forceinline__ bool marking(const StepArgs& a) {
    return H::kFrontier && a.act_w != nullptr && a.fctl[1] != 0;
}
// Marks v and its neighbours [b, e), `stride` threads cooperating.
__device__ __forceinline__ void mark_vertex(const StepArgs& a, int32_t v, int32_t b, int32_t e,
                                            int tid, int stride) {
    if (tid == 0) a.act_w[v] = 1;
    for (int32_t j = b + tid; j < e; j += stride) a.act_w[a.nbr[j]] = 1;
}

// Each tier is a __device__ body taking its workgroup index `blk`; the
// __global__ wrappers below run it for one job (blk = blockIdx.x) or for a
// table of jobs folded into one launch (blk = index inside the owning job).

// ---------------------------------------------------------------- packed ----
// A wave takes NG consecutive groups (NG chosen per launch from the amount of
// work: more groups per wave give each wave more loads in flight, fewer
// leave small launches more waves).  When all of them have edges (the
// planner puts the edge-less groups last) their loads are issued branch-free,
// every group's before any group is counted, so the dependent loads of the
// groups (offsets, neighbour ids, labels) overlap; otherwise, and with the
// frontier on, the groups go one by one through the guarded path.
//
// Counting a group: every edge lane inserts (owner slot, label) into a
// 128-slot LDS table (<= 64 keys, so a slot is always free) and adds its
// frequency; it reads back the sum for its pair, scores it, and an LDS
// 64-bit max per owner slot keeps each vertex's best packed score.  Sums and
// maxima are order-free integers, as in every tier.
constexpr int kFoldNG = 2;  // groups per wave in folded (multi-job) launches
constexpr uint32_t kPackSlots = 128;

// One group's state between its loads and its count.  Vertex lane i owns
// edge lanes [vb, ve) of the group; edge lane `lane` belongs to vertex slot
// `own`.
struct PackedGroup {
    int32_t v0, vc, base, ve, vb, nE, own, old;
    uint32_t key, fr;
    int32_t vol;  // the label's volume (slot-scoring hook sets)
};

// Per-wave LDS of the packed count.
struct PackedLds {
    unsigned long long key[kPackSlots];  // (owner slot + 1) << 32 | label key; 0 = free
    uint32_t cnt[kPackSlots];
    unsigned long long best[64];         // per vertex slot: best packed score
};

__device__ __forceinline__ void packed_segments(PackedGroup& s, int lane) {
    s.vb = shfl_i32(s.ve, lane > 0 ? lane - 1 : 0);
    if (lane == 0) s.vb = 0;
    s.nE = shfl_i32(s.ve, s.vc - 1);
    // owner: smallest vertex slot i with ve_i > lane
    int lo = 0, hi = s.vc - 1;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int mid = (lo + hi) >> 1;
        const int32_t vm = shfl_i32(s.ve, mid);
        if (vm <= lane) lo = min(mid + 1, hi);
        else hi = mid;
    }
    s.own = lo;
}

// Loads of a group with vc >= 1 vertices and nE >= 1 edges: indices past the
// group are clamped to its last vertex / edge and the values dropped.
template <class H, bool W, bool G>
__device__ __forceinline__ void packed_load_fast(const StepArgs& a, int2 gr, PackedGroup& s) {
    const int lane = lane_id();
    s.v0 = gr.x;
    s.vc = gr.y;
    s.base = a.off[s.v0];
    const int32_t li = min(lane, s.vc - 1);
    const int32_t end = a.off[s.v0 + li + 1];
    s.old = a.keep[s.v0 + li];
    s.ve = lane < s.vc ? end - s.base : 0;
    packed_segments(s, lane);
    s.key = load_edge<H, W, G>(a, s.v0 + s.own, s.base + min(lane, s.nE - 1), &s.fr, &s.vol);
    if (lane >= s.nE) {
        s.key = 0u;
        s.fr = 0u;
    }
}

// Guarded loads for any group; false when the frontier skips the group.
template <class H, bool W, bool G>
__device__ __forceinline__ bool packed_load_safe(const StepArgs& a, int2 gr, PackedGroup& s) {
    const int lane = lane_id();
    s.v0 = gr.x;
    s.vc = gr.y;
    if (frontier_on<H>(a)) {
        const bool act = lane < s.vc && a.act_r[s.v0 + lane] != 0;
        if (__ballot(act) == 0ull) return false;  // no vertex of the group needs work
    }
    s.base = a.off[s.v0];
    s.ve = 0;
    s.old = 0;
    if (lane < s.vc) {
        s.ve = a.off[s.v0 + lane + 1] - s.base;
        s.old = a.keep[s.v0 + lane];
    }
    packed_segments(s, lane);
    s.key = 0u;
    s.fr = 0u;
    s.vol = 0;
    if (lane < s.nE) s.key = load_edge<H, W, G>(a, s.v0 + s.own, s.base + lane, &s.fr, &s.vol);
    return true;
}

// Counts, picks and writes the group's labels (all lanes active); returns
// how many of them changed.
template <class H, bool W, bool G>
__device__ __forceinline__ uint32_t packed_finish(const StepArgs& a, const PackedGroup& s, PackedLds& t,
                                                  bool& cleared) {
    const int lane = lane_id();
    // key 0 is an unlabeled neighbour (seeded LP): it never counts
    const bool live = lane < s.nE && s.key != 0u;
    uint64_t best = 0ull;  // on vertex lanes: the vertex's best packed score
    bool decided = false;  // the whole group was decided without the table
    if constexpr (count_score_of<H>::value && !W) {
        // majority shortcut: a label on more than half of a vertex's counted
        // edges wins outright.  Candidates per vertex: the labels of its
        // first and of its last edge (one of them is the majority label
        // unless both edges carry minority labels).
        const unsigned long long livem = __ballot(live);
        const int32_t sb = shfl_i32(s.vb, s.own), se = shfl_i32(s.ve, s.own);
        const uint32_t o1 = shfl_u32(s.key, sb), o2 = shfl_u32(s.key, max(se - 1, 0));
        const unsigned long long m1 = __ballot(live && s.key == o1);
        const unsigned long long m2 = __ballot(live && s.key == o2);
        const uint32_t w1 = shfl_u32(s.key, min(s.vb, 63)), w2 = shfl_u32(s.key, max(s.ve - 1, 0));
        bool open = false;
        if (lane < s.vc && !is_fixed<H>(a, s.v0 + lane)) {
            const unsigned long long below = s.ve >= 64 ? ~0ull : ((1ull << s.ve) - 1ull);
            const unsigned long long seg = s.vb >= 64 ? 0ull : below & ~((1ull << s.vb) - 1ull);
            const uint32_t len = (uint32_t)__popcll(livem & seg);
            const uint32_t c1 = (uint32_t)__popcll(m1 & seg), c2 = (uint32_t)__popcll(m2 & seg);
            if (len == 0u) best = 0ull;  // no counted edge: the vertex keeps its label
            else if (2u * c1 > len) best = H::label_score(a, s.v0 + lane, w1, c1);
            else if (2u * c2 > len) best = H::label_score(a, s.v0 + lane, w2, c2);
            else open = true;
        }
        decided = __ballot(open) == 0ull;  // wave-uniform
    }
    if (!decided) {
        // the table is cleared before its first use in the wave, and each
        // group leaves it empty: a slot's creator clears its slot once it
        // has scored it, and each vertex lane its best word once read
        if (!cleared) {  // wave-uniform
            t.key[lane] = 0ull;
            t.key[lane + 64] = 0ull;
            t.cnt[lane] = 0u;
            t.cnt[lane + 64] = 0u;
            t.best[lane] = 0ull;
            cleared = true;
        }
        wave_lds_sync();  // the clear, or the previous group's clears, are done
        uint32_t h = 0u;
        bool claimed = false;  // this lane created the (owner, label) slot
        if (live) {
            const unsigned long long k64 = ((unsigned long long)(uint32_t)(s.own + 1) << 32) | s.key;
            h = fmix32(s.key ^ ((uint32_t)s.own * 0x9e3779b9u)) & (kPackSlots - 1u);
            for (;;) {
                unsigned long long cur = lds_load(&t.key[h]);
                if (cur == 0ull) {
                    cur = atomicCAS(&t.key[h], 0ull, k64);
                    if (cur == 0ull) {
                        claimed = true;
                        break;
                    }
                }
                if (cur == k64) break;
                h = (h + 1u) & (kPackSlots - 1u);
            }
            // a slot-scored count: every edge adds its share, the creator
            // also the label's first-edge term (sums mod 2^32 are exact)
            constexpr bool S = slot_score_of<H>::value;
            const uint32_t f = W ? s.fr : 1u;
            uint32_t add = slot_more<H, S>(a, f);
            if (claimed) add += slot_first<H, S>(a, f, s.vol) - add;
            atomicAdd(&t.cnt[h], add);
        }
        wave_lds_sync();
        // one score per distinct (owner, label): the slot's creator
        if (claimed) {
            const uint64_t p =
                slot_value_score<H, slot_score_of<H>::value>(a, s.v0 + s.own, s.key, t.cnt[h]);
            atomicMax(&t.best[s.own], (unsigned long long)p);
            t.key[h] = 0ull;
            t.cnt[h] = 0u;
        }
        wave_lds_sync();
        best = t.best[lane];
        t.best[lane] = 0ull;
    }
    bool changed = false, gch = false;
    if (lane < s.vc && !is_fixed<H>(a, s.v0 + lane)) {
        const int32_t v = s.v0 + lane;
        // no edge, or no counted label: the vertex keeps its label
        const int32_t nl = best ? packed_label(best) : s.old;
        gch = gather_changed(a, v, nl, s.old);
        H::update_vertex(a, v, nl, best);
        changed = nl != s.old;
    }
    if (marking<H>(a)) {
        if (changed) a.act_w[s.v0 + lane] = 1;
        const bool owner_changed = shfl_i32(changed ? 1 : 0, s.own) != 0;
        if (lane < s.nE && owner_changed) a.act_w[a.nbr[s.base + lane]] = 1;
    }
    note_changed_run(a, s.v0, __ballot(gch));
    return (uint32_t)__popcll(__ballot(changed));
}

// ngroups: groups to run; groups [0, nge) have edges.  The wave runs unit
// `wunit` (groups [wunit NG, wunit NG + NG)) in its LDS area t.
template <class H, bool W, bool G, int NG>
__device__ __forceinline__ void packed_body(const StepArgs& a, const int2* __restrict__ groups,
                                            int32_t ngroups, int32_t nge, uint32_t wunit, PackedLds& t) {
    // wave-uniform first group (readfirstlane keeps the group loads scalar)
    const int32_t g0 = __builtin_amdgcn_readfirstlane((int32_t)wunit) * NG;
    if (g0 >= ngroups) return;
    bool cleared = false;  // the wave's table (wave-uniform)
    uint32_t nch = 0u;
    if (g0 + NG <= nge && !frontier_on<H>(a)) {
        PackedGroup s[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) packed_load_fast<H, W, G>(a, groups[g0 + g], s[g]);
#pragma unroll
        for (int g = 0; g < NG; ++g) nch += packed_finish<H, W, G>(a, s[g], t, cleared);
    } else {
        for (int g = 0; g < NG && g0 + g < ngroups; ++g) {
            PackedGroup s;
            if (packed_load_safe<H, W, G>(a, groups[g0 + g], s)) nch += packed_finish<H, W, G>(a, s, t, cleared);
        }
    }
    add_changes(a, nch);
}

// ------------------------------------------------------------------ half ----
// Wave-tier vertices of degree <= half_max_deg (<= kHalfMaxDeg = 64): S lanes
// count one vertex, so a wave counts 64/S at a time (S = 32: halves, two
// edges per lane; S = 16: quarters, four edges per lane), each lane group in
// its own LDS table of kHalfCap slots.  A wave takes NP such units (chosen
// per launch like the packed tier's NG), their loads issued before any is
// counted.
constexpr uint32_t kHalfCap = 2 * kHalfMaxDeg;

// Per-wave LDS of the half tier: one (key << 32 | sum) table per lane group.
template <int S>
struct HalfLds {
    unsigned long long slot[64 / S][kHalfCap];
};

// EPL: edges loaded per lane (S * EPL >= the launch's largest degree).
// The wave runs unit `wunit` (vertices [wunit V NP, ...)) in its LDS area.
template <class H, bool W, bool G, int NP, int S = 32, int EPL = kHalfMaxDeg / S>
__device__ __forceinline__ void half_body(const StepArgs& a, const int4* __restrict__ verts,
                                          int32_t nv, uint32_t wunit, HalfLds<S>& lds) {
    static_assert(S == 16 || S == 32, "16 or 32 lanes per vertex");
    constexpr int V = 64 / S;               // vertices per unit
    const int lane = lane_id(), q = lane / S, ql = lane % S;
    const int g0 = q * S;  // the group's first lane
    const unsigned long long gm = ((1ull << S) - 1ull) << g0;  // the group's lanes
    const int32_t i0 = (int32_t)wunit * (V * NP);
    if (i0 >= nv) return;  // wave-uniform
    int32_t v[NP], b[NP], e[NP], old[NP];
    bool run[NP];
    uint32_t key[NP][EPL], fr[NP][EPL];
    int32_t vol[NP][EPL];
    constexpr bool S_ = slot_score_of<H>::value;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int32_t idx = i0 + V * p + q;
        const int4 ent = verts[min(idx, nv - 1)];
        v[p] = ent.x;
        b[p] = ent.y;
        e[p] = ent.z;
        run[p] = idx < nv;
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        run[p] = run[p] && is_active<H>(a, v[p]) && !is_fixed<H>(a, v[p]);
        old[p] = a.keep[v[p]];
    }
    // every tier vertex has >= 1 edge, so e-1 is a valid index; a vertex not
    // run loads its first edge only
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int i = 0; i < EPL; ++i) {
            const int32_t j = b[p] + ql + S * i;
            key[p][i] = load_edge<H, W, G>(a, v[p], run[p] ? min(j, e[p] - 1) : b[p], &fr[p][i], &vol[p][i]);
            if (!run[p] || j >= e[p]) key[p][i] = 0u;
        }
    unsigned long long* T = lds.slot[q];
    // the tables are cleared before their first use, and each unit leaves
    // them empty: a slot's creator clears it once it has scored it
    bool cleared = false;  // wave-uniform
    // lanes of this group where x holds
    auto group_count = [&](bool x) { return (uint32_t)__popcll(__ballot(x) & gm); };
    // sum over this group's lanes (xor distances below S stay in the group)
    auto group_sum = [&](uint32_t x) {
#pragma unroll
        for (int d = S / 2; d >= 1; d >>= 1) x += shfl_u32(x, lane ^ d);
        return x;
    };
    uint32_t nch = 0u;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        if (__ballot(run[p]) == 0ull) continue;  // wave-uniform: every group idle
        // majority shortcut (scores that are frequency sums): when the label
        // of the vertex's first edge, or of its (S/2+1)-th, holds more than
        // half of the vertex's counted frequency, no other label can reach
        // its score
        uint64_t fast = 0ull;
        bool peel = true;  // merge lanes of a shared label before the table
        if constexpr (count_score_of<H>::value) {
            const uint32_t lead = shfl_u32(key[p][0], g0);
            const uint32_t lead2 = shfl_u32(key[p][0], g0 + S / 2);
            uint32_t c = 0u, c2 = 0u, tot = 0u;
            if (W) {
#pragma unroll
                for (int i = 0; i < EPL; ++i) {
                    const uint32_t k = key[p][i];
                    c += k && k == lead ? fr[p][i] : 0u;
                    c2 += k && k == lead2 ? fr[p][i] : 0u;
                    tot += k ? fr[p][i] : 0u;
                }
                c = group_sum(c);
                c2 = group_sum(c2);
                tot = group_sum(tot);
            } else {
                // only seeded LP has unlabeled neighbours (key 0); otherwise
                // every edge of the vertex counts once
                if constexpr (seeded_of<H>::value) {
#pragma unroll
                    for (int i = 0; i < EPL; ++i) tot += group_count(key[p][i] != 0u);
                } else {
                    tot = run[p] ? (uint32_t)(e[p] - b[p]) : 0u;
                }
#pragma unroll
                for (int i = 0; i < EPL; ++i) c += group_count(key[p][i] != 0u && key[p][i] == lead);
                // the second candidate only when some group still needs one
                if (__ballot(run[p] && !(lead != 0u && 2ull * c > (unsigned long long)tot)) != 0ull) {
#pragma unroll
                    for (int i = 0; i < EPL; ++i)
                        c2 += group_count(key[p][i] != 0u && key[p][i] == lead2);
                }
            }
            if (run[p] && lead != 0u && 2ull * c > (unsigned long long)tot)
                fast = H::label_score(a, v[p], lead, c);
            else if (run[p] && lead2 != 0u && 2ull * c2 > (unsigned long long)tot)
                fast = H::label_score(a, v[p], lead2, c2);
            // neither candidate repeats much: the labels are diverse and
            // merging lanes would find nothing to merge (wave-uniform)
            peel = __ballot(run[p] && fast == 0ull && max(c, c2) >= 4u) != 0ull;
        }
        uint64_t best = fast;
        if (__ballot(run[p] && fast == 0ull) != 0ull) {  // wave-uniform: some group needs the table
            if (!cleared) {
#pragma unroll
                for (uint32_t i = 0; i < kHalfCap; i += S) T[i + ql] = 0ull;
                cleared = true;
            }
            wave_lds_sync();  // the clear, or the previous unit's clears, are done
            // per group, the lanes holding the group's first pending label
            // merge into one insert by the lowest of them (a converged vertex
            // shows one label on most edges, and same-address LDS atomics
            // serialise); the merged lanes neither insert nor score (the
            // leader's slot holds the same count).  A vertex decided by the
            // shortcut inserts nothing.
            uint32_t hs[EPL], sk[EPL];
#pragma unroll
            for (int i = 0; i < EPL; ++i) {
                uint32_t k = fast ? 0u : key[p][i], f = W ? fr[p][i] : 1u;
                if (peel) {
                    const unsigned long long rest = __ballot(k != 0u) & gm;
                    const int first = rest ? (int)__builtin_ctzll(rest) : g0;
                    const uint32_t lead = shfl_u32(k, first);
                    const bool same = k != 0u && k == lead;
                    const uint32_t cnt = group_count(same);
                    if (W) {
                        // weighted: the group's sum of the merged lanes' frequencies
                        const uint32_t x = group_sum(same ? f : 0u);
                        if (same && cnt >= 2u) f = x;
                    } else if (same && cnt >= 2u) {
                        f = cnt;
                    }
                    if (same && cnt >= 2u && lane != first) k = 0u;  // merged into the leader
                }
                sk[i] = k;
                // a merged lane shows the leader's label, so the volumes agree
                hs[i] = k ? kc_insert(T, kHalfCap - 1u, k, slot_first<H, S_>(a, f, vol[p][i]),
                                      slot_more<H, S_>(a, f))
                          : 0u;
            }
            wave_lds_sync();
            // one score per distinct label: the slot's creator
            uint64_t pm = 0ull;
#pragma unroll
            for (int i = 0; i < EPL; ++i)
                if (sk[i] && (hs[i] & kClaimed)) {
                    unsigned long long* slot = &T[hs[i] & ~kClaimed];
                    pm = max_u64(pm, slot_value_score<H, S_>(a, v[p], sk[i], (uint32_t)*slot));
                    *slot = 0ull;
                }
            // max over the group (xor distances below S stay in the group)
#pragma unroll
            for (int d = S / 2; d >= 1; d >>= 1) pm = max_u64(pm, shfl_u64(pm, lane ^ d));
            if (!fast) best = pm;
        }
        bool changed = false;
        if (ql == 0 && run[p]) {
            const int32_t nl = best ? packed_label(best) : old[p];
            const bool gch = gather_changed(a, v[p], nl, old[p]);
            H::update_vertex(a, v[p], nl, best);
            changed = nl != old[p];
            if (gch) note_changed(a, v[p]);
        }
        nch += (uint32_t)__popcll(__ballot(changed));
        if (marking<H>(a) && shfl_i32(changed ? 1 : 0, g0) != 0)
            mark_vertex(a, v[p], b[p], e[p], ql, S);
    }
    add_changes(a, nch);
}

// ------------------------------------------------------------------ wave ----
// A wave takes kWaveNV vertices of the tier.  A vertex has at most 256 edges
// (4 per lane), so one round of branch-free loads covers it: the loads of all
// kWaveNV vertices are issued first, then each vertex is counted in the
// wave's LDS table in turn.
constexpr int kWaveNV = 4;
// EPL: edges loaded per lane (4 covers 256 edges; 2 serves vertices of at
// most 128 edges without clamped duplicate loads)
// The wave runs unit `wunit` (vertices [wunit kWaveNV, ...)) with its LDS
// table T of kWaveCap (label key << 32 | count) slots.
template <class H, bool W, bool G, int EPL = 4>
__device__ __forceinline__ void wave_body(const StepArgs& a, const int4* __restrict__ verts,
                                          int32_t nv, uint32_t wunit, unsigned long long* T) {
    const int lane = lane_id();
    const int32_t i0 = (int32_t)wunit * kWaveNV;
    if (i0 >= nv) return;  // wave-uniform
    int32_t v[kWaveNV], b[kWaveNV], e[kWaveNV], old[kWaveNV];
    bool run[kWaveNV];
    uint32_t key[kWaveNV][EPL], fr[kWaveNV][EPL];
    int32_t vol[kWaveNV][EPL];
    constexpr bool S_ = slot_score_of<H>::value;
#pragma unroll
    for (int k = 0; k < kWaveNV; ++k) {
        const int4 ent = verts[min(i0 + k, nv - 1)];
        v[k] = ent.x;
        b[k] = ent.y;
        e[k] = ent.z;
    }
#pragma unroll
    for (int k = 0; k < kWaveNV; ++k) {
        run[k] = i0 + k < nv && is_active<H>(a, v[k]) && !is_fixed<H>(a, v[k]);
        old[k] = a.keep[v[k]];
    }
    // every tier vertex has >= 1 edge, so e-1 is a valid index; a vertex not
    // run loads its first edge only (one line for the whole wave)
#pragma unroll
    for (int k = 0; k < kWaveNV; ++k)
#pragma unroll
        for (int i = 0; i < EPL; ++i) {
            const int32_t j = b[k] + lane + 64 * i;
            key[k][i] = load_edge<H, W, G>(a, v[k], run[k] ? min(j, e[k] - 1) : b[k], &fr[k][i], &vol[k][i]);
            if (j >= e[k]) key[k][i] = 0u;
        }
    uint32_t nch = 0u;  // changed labels (lane 0)
#pragma unroll
    for (int k = 0; k < kWaveNV; ++k) {
        if (!run[k]) continue;  // wave-uniform
        uint64_t best = 0ull;
        bool decided = false;  // wave-uniform
        bool peel = true;      // merge lanes of a shared label before the table
        if constexpr (count_score_of<H>::value) {
            // majority shortcut: the label of the first edge, or of the
            // 33rd, wins outright when it holds more than half of the
            // counted frequency
            const uint32_t lead = (uint32_t)__builtin_amdgcn_readlane((int)key[k][0], 0);
            const uint32_t lead2 = (uint32_t)__builtin_amdgcn_readlane((int)key[k][0], 32);
            uint32_t c = 0u, c2 = 0u, tot = 0u;
            if (W) {
#pragma unroll
                for (int i = 0; i < EPL; ++i) {
                    c += key[k][i] && key[k][i] == lead ? fr[k][i] : 0u;
                    c2 += key[k][i] && key[k][i] == lead2 ? fr[k][i] : 0u;
                    tot += key[k][i] ? fr[k][i] : 0u;
                }
                c = wave_sum_u32(c);
                c2 = wave_sum_u32(c2);
                tot = wave_sum_u32(tot);
            } else {
                // only seeded LP has unlabeled neighbours (key 0); otherwise
                // every edge of the vertex counts once
                if constexpr (seeded_of<H>::value) {
#pragma unroll
                    for (int i = 0; i < EPL; ++i) tot += (uint32_t)__popcll(__ballot(key[k][i] != 0u));
                } else {
                    tot = (uint32_t)(e[k] - b[k]);
                }
#pragma unroll
                for (int i = 0; i < EPL; ++i)
                    c += (uint32_t)__popcll(__ballot(key[k][i] != 0u && key[k][i] == lead));
                if (!(lead != 0u && 2ull * c > (unsigned long long)tot)) {  // wave-uniform
#pragma unroll
                    for (int i = 0; i < EPL; ++i)
                        c2 += (uint32_t)__popcll(__ballot(key[k][i] != 0u && key[k][i] == lead2));
                }
            }
            if (lead != 0u && 2ull * c > (unsigned long long)tot) {
                best = H::label_score(a, v[k], lead, c);
                decided = true;
            } else if (lead2 != 0u && 2ull * c2 > (unsigned long long)tot) {
                best = H::label_score(a, v[k], lead2, c2);
                decided = true;
            }
            // neither candidate repeats much: the labels are diverse and
            // merging lanes would find nothing to merge
            peel = max(c, c2) >= 4u;
        }
        if (!decided) {
            const uint32_t cap = min(max(next_pow2(2u * (uint32_t)(e[k] - b[k])), 64u), (uint32_t)kWaveCap);
            const uint32_t mask = cap - 1u;
            for (uint32_t i = lane; i < cap; i += 64) T[i] = 0ull;
            wave_lds_sync();
#pragma unroll
            for (int i = 0; i < EPL; ++i) {
                if (peel) wave_insert<H, W, S_>(a, T, mask, key[k][i], fr[k][i], vol[k][i]);
                else if (key[k][i])
                    kc_insert(T, mask, key[k][i], slot_first<H, S_>(a, fr[k][i], vol[k][i]),
                              slot_more<H, S_>(a, fr[k][i]));
            }
            wave_lds_sync();
            best = wave_max_u64(table_best<H, S_>(a, v[k], T, cap, (uint32_t)lane, 64u, false));
            wave_lds_sync();  // the next vertex clears what this scan read
        }
        bool changed = false;
        if (lane == 0) {
            const int32_t nl = best ? packed_label(best) : old[k];
            const bool gch = gather_changed(a, v[k], nl, old[k]);
            H::update_vertex(a, v[k], nl, best);
            changed = nl != old[k];
            if (gch) note_changed(a, v[k]);
            nch += changed ? 1u : 0u;
        }
        if (marking<H>(a) && shfl_i32(changed ? 1 : 0, 0) != 0) mark_vertex(a, v[k], b[k], e[k], lane, 64);
    }
    add_changes(a, nch);
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

template <class H, bool W, bool G, int NT, int K = 8>
__device__ __forceinline__ void block_body(const StepArgs& a, const int32_t* __restrict__ verts,
                                           uint32_t cap_max, uint32_t blk) {
    extern __shared__ __align__(16) uint32_t s_dyn[];
    __shared__ unsigned long long s_red[NT / 64];
    __shared__ uint32_t s_changed;
    unsigned long long* T = reinterpret_cast<unsigned long long*>(s_dyn);  // cap_max slots
    const int32_t v = verts[blk];
    if (!is_active<H>(a, v) || is_fixed<H>(a, v)) return;  // block-uniform
    const int32_t b = a.off[v], e = a.off[v + 1];
    const uint32_t cap = min(max(next_pow2(2u * (uint32_t)(e - b)), 64u), cap_max);
    const uint32_t mask = cap - 1u;
    for (uint32_t i = threadIdx.x; i < cap; i += NT) T[i] = 0ull;
    __syncthreads();
    constexpr bool S = slot_score_of<H>::value;
    count_edges<H, W, G, K, NT, S>(a, v, b, e, threadIdx.x, T, mask);
    __syncthreads();
    uint64_t best = table_best<H, S>(a, v, T, cap, threadIdx.x, NT, false);
    best = block_max_u64<NT>(best, s_red);
    if (threadIdx.x < 64) {
        bool changed = false;
        if (threadIdx.x == 0) {
            const int32_t old = a.keep[v];
            const int32_t nl = best ? packed_label(best) : old;
            const bool gch = gather_changed(a, v, nl, old);
            H::update_vertex(a, v, nl, best);
            changed = nl != old;
            if (gch) note_changed(a, v);
            s_changed = changed ? 1u : 0u;
        }
        count_changes(a, changed);
    }
    if (marking<H>(a)) {
        __syncthreads();
        if (s_changed) mark_vertex(a, v, b, e, threadIdx.x, NT);
    }
}

// ------------------------------------------------------------------- hub ----
// Hubs are counted without any HBM hash table, in three passes:
//  k_hub_agg     one workgroup per chunk of `chunk` edges: gathers the labels
//                and counts them in an LDS table, then writes the chunk's
//                distinct (label+1, frequency) pairs to its own pair region
//                grouped by the hub's P label-hash buckets (a pair's bucket is
//                fmix32(label+1) * P >> 32): a histogram over the buckets, its
//                exclusive scan in LDS, then each pair to its bucket's place.
//                Per (bucket, chunk) it records the run: how many pairs and
//                where they start in the chunk's region.
//  k_hub_part    one workgroup per batch of (hub, bucket) parts: a part sums
//                its bucket's runs, read from every chunk region, in an LDS
//                table and keeps the best; if the bucket holds more labels
//                than the table, it counts the lower half of its fmix32 range,
//                then the rest, halving as needed (fmix32 is a bijection, so a
//                range no wider than the table cannot overflow)
//  k_hub_final   one thread per hub: commits the winner
// P = ceil(4 degree / (3 part_keys)), at most kMaxBuckets, so a bucket expects
// at most 3/4 part_keys distinct labels (its table of 2 part_keys slots takes
// part_keys before halving) even when every neighbour shows a different
// label; once labels converge most buckets are empty and exit at once.
// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ int32_t wave_incl_sum(int32_t x) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t y = shfl_i32(x, lane >= d ? lane - d : lane);
        if (lane >= d) x += y;
    }
    return x;
}
// Exclusive prefix of x over a workgroup of NT threads; *total = the sum.
template <int NT = 256>
__device__ __forceinline__ int32_t block_excl_sum(int32_t x, int32_t* total) {
    __shared__ int32_t s_w[NT / 64];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int32_t incl = wave_incl_sum(x);
    if (lane == 63) s_w[wv] = incl;
    __syncthreads();
    int32_t base = 0, all = 0;
    for (int i = 0; i < NT / 64; ++i) {
        if (i < wv) base += s_w[i];
        all += s_w[i];
    }
    __syncthreads();
    *total = all;
    return base + incl - x;
}
constexpr int kHubThreads = 256;
// k_hub_agg: a chunk's table takes 34 KB of LDS, so four workgroups share a
// CU; 512 threads of 4 gathers each keep its waves full
constexpr int kHubAggThreads = 512;

// dynamic LDS of k_hub_agg: the chunk's table plus the bucket histogram of
// the launch's largest hub (buckets <= kMaxBuckets); the smaller the
// histogram, the more workgroups share a CU
static inline size_t hub_agg_lds(int32_t chunk, int32_t buckets = kMaxBuckets) {
    return (size_t)next_pow2(2u * (uint32_t)chunk) * 2 * sizeof(uint32_t) +
           (size_t)((buckets + 63) & ~63) * sizeof(uint32_t);
}

__device__ __forceinline__ uint32_t bucket_of(uint32_t hk, uint32_t P) {
    return (uint32_t)(((uint64_t)hk * P) >> 32);
}
// first fmix32 value of bucket p: ceil(p * 2^32 / P)
__device__ __forceinline__ uint64_t bucket_lo(uint32_t p, uint32_t P) {
    return (((uint64_t)p << 32) + P - 1) / P;
}

// NT threads per workgroup, K gathers in flight per thread
template <class H, bool W, bool G, int NT = kHubThreads, int K = 8>
__device__ __forceinline__ void hub_agg_body(const StepArgs& a, const HubArgs& h, uint32_t blk) {
    // dynamic LDS: table of next_pow2(2*chunk) keys and counts, then the
    // bucket histogram (at least P counters for every hub of the launch)
    extern __shared__ __align__(16) uint32_t s_dyn[];
    const uint32_t cap_max = next_pow2(2u * (uint32_t)h.chunk);
    unsigned long long* T = reinterpret_cast<unsigned long long*>(s_dyn);  // cap_max slots
    uint32_t* s_hist = s_dyn + 2 * cap_max;
    const int2 hc = h.chunks[blk];
    const int32_t v = h.hubs[hc.x];
    const int32_t vb = a.off[v], ve = a.off[v + 1];
    const int32_t nch = (ve - vb + h.chunk - 1) / h.chunk;
    const uint32_t P = hub_buckets(ve - vb, h.part_keys);
    uint32_t* hist = h.hs.hist + h.hist_base[hc.x];      // [bucket][chunk]: run lengths
    uint32_t* lstart = h.hs.lstart + h.hist_base[hc.x];  // [bucket][chunk]: run starts in the chunk region
    if (!is_active<H>(a, v) || is_fixed<H>(a, v)) {  // block-uniform: the hub keeps its label
        for (uint32_t q = threadIdx.x; q < P; q += NT) hist[(int64_t)q * nch + hc.y] = 0u;
        return;
    }
    const int32_t b = vb + hc.y * h.chunk;
    const int32_t e = min(b + h.chunk, ve);
    const uint32_t cap = min(max(next_pow2(2u * (uint32_t)(e - b)), 64u), cap_max);
    const uint32_t mask = cap - 1u;
    for (uint32_t i = threadIdx.x; i < cap; i += NT) T[i] = 0ull;
    for (uint32_t q = threadIdx.x; q < P; q += NT) s_hist[q] = 0u;
    __syncthreads();
    count_edges<H, W, G, K, NT>(a, v, b, e, threadIdx.x, T, mask);
    __syncthreads();
    // the chunk's distinct pairs per bucket
    for (uint32_t i = threadIdx.x; i < cap; i += NT) {
        const uint32_t k = (uint32_t)(T[i] >> 32);
        if (k) atomicAdd(&s_hist[bucket_of(fmix32(k), P)], 1u);
    }
    __syncthreads();
    // run lengths out; their exclusive scan in place (each thread a run of
    // consecutive buckets) gives the runs' starts, kept as bucket cursors
    const uint32_t per = (P + NT - 1) / NT;
    const uint32_t q0 = min(threadIdx.x * per, P), q1 = min(q0 + per, P);
    uint32_t run = 0u;
    for (uint32_t q = q0; q < q1; ++q) {
        const uint32_t c = s_hist[q];
        hist[(int64_t)q * nch + hc.y] = c;
        run += c;
    }
    int32_t total;
    uint32_t at = (uint32_t)block_excl_sum<NT>((int32_t)run, &total);
    for (uint32_t q = q0; q < q1; ++q) {
        const uint32_t c = s_hist[q];
        lstart[(int64_t)q * nch + hc.y] = at;
        s_hist[q] = at;
        at += c;
    }
    __syncthreads();
    // each pair (label key << 32 | count: the pair itself) to its bucket's run
    unsigned long long* out = h.hs.pairs + h.base[hc.x] + (int64_t)hc.y * h.chunk;
    for (uint32_t i = threadIdx.x; i < cap; i += NT) {
        const unsigned long long x = T[i];
        const uint32_t k = (uint32_t)(x >> 32);
        if (k) out[atomicAdd(&s_hist[bucket_of(fmix32(k), P)], 1u)] = x;
    }
}

// A (hub, bucket) part: the hub, its vertex, the bucket, the hub's buckets
// and chunks, and the bucket's pairs over all chunks.
struct PartJob {
    int32_t hub, v;
    uint32_t q, P, nch, total;
};

// Sums the part's pairs whose fmix32 image lies in [lo, hi) into the LDS
// table.  The pairs lie in one run per chunk region; the workgroup walks the
// chunks 64 at a time (every wave holds the 64 runs' lengths and starts, one
// per lane) and spreads each window's pairs over its lanes, a pair's run
// found by a binary search over the lanes' inclusive sums.  Stops early (the
// caller halves the range) once more than `limit` distinct keys arrived;
// the table has room for limit + kHubThreads.
__device__ __forceinline__ void part_pass(const HubArgs& h, const PartJob& j, uint64_t lo, uint64_t hi,
                                          unsigned long long* T, uint32_t shift, uint32_t limit,
                                          uint32_t* s_distinct, uint32_t* s_over) {
    const uint32_t mask = (1u << (32 - shift)) - 1u;
    constexpr int kU = 4;  // pairs loaded per thread before any is inserted
    constexpr uint32_t kWaves = kHubThreads / 64;
    const int lane = lane_id();
    const uint32_t wv = threadIdx.x >> 6;
    const int64_t hb = h.hist_base[j.hub] + (int64_t)j.q * j.nch;
    const uint32_t* cnt = h.hs.hist + hb;
    const uint32_t* st = h.hs.lstart + hb;
    const unsigned long long* pr = h.hs.pairs + h.base[j.hub];
    for (uint32_t c0 = 0; c0 < j.nch; c0 += 64) {  // every wave walks the same windows
        const uint32_t c = c0 + (uint32_t)lane;
        const uint32_t n_c = c < j.nch ? cnt[c] : 0u;
        const uint32_t s_c = c < j.nch ? c * (uint32_t)h.chunk + st[c] : 0u;
        const int32_t incl = wave_incl_sum((int32_t)n_c);
        const uint32_t wtot = (uint32_t)__builtin_amdgcn_readfirstlane(shfl_i32(incl, 63));
        // the window's runs: lanes [0, nr); the owner search needs
        // ceil(log2(nr)) halvings (most hubs have a few chunks)
        const int nr = (int)min(64u, j.nch - c0);
        const int depth = nr > 1 ? 32 - __builtin_clz((uint32_t)(nr - 1)) : 0;
        // the trip count is wave-uniform, so the wave reductions below run
        // with every lane; lanes past the window's pairs load a clamped pair
        // and skip it
        for (uint32_t f0 = wv * 64u; f0 < wtot; f0 += kWaves * 64u * kU) {
            if (lds_load(s_over)) return;  // one LDS read for the wave: uniform
            unsigned long long xs[kU];
            bool in[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const uint32_t f = f0 + (uint32_t)u * kWaves * 64u + (uint32_t)lane;
                in[u] = f < wtot;
                const int32_t fc = (int32_t)min(f, wtot - 1u);
                // owner: the first lane whose inclusive sum exceeds fc
                // (lanes >= nr hold no pairs, so it lies in [0, nr))
                int o = 0, top = nr - 1;
                for (int k = 0; k < depth; ++k) {  // wave-uniform trip count
                    const int mid = (o + top) >> 1;
                    if (shfl_i32(incl, mid) <= fc) o = mid + 1;
                    else top = mid;
                }
                const int32_t oi = shfl_i32(incl, o);
                const uint32_t on = shfl_u32(n_c, o), os = shfl_u32(s_c, o);
                xs[u] = pr[os + (uint32_t)(fc - (oi - (int32_t)on))];
            }
            uint32_t claimed = 0u;
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const uint32_t key = (uint32_t)(xs[u] >> 32);
                const uint32_t hk = fmix32(key);
                if (!in[u] || hk < lo || hk >= hi) continue;
                // a pair is already a (key << 32 | count) slot value; most pairs
                // of a bucket bring a new label while labels are diverse, so the
                // claim is tried first
                uint32_t s = (hk * 0x9e3779b1u) >> shift;
                for (uint32_t probes = 0;; ++probes) {
                    if (probes > mask) {
                        // every slot is taken (waves that have not yet seen the
                        // flag overshoot the limit): give the range up
                        atomicExch(s_over, 1u);
                        break;
                    }
                    const unsigned long long cur = atomicCAS(&T[s], 0ull, xs[u]);
                    if (cur == 0ull) {
                        ++claimed;
                        break;
                    }
                    if ((uint32_t)(cur >> 32) == key) {
                        atomicAdd(&T[s], xs[u] & 0xffffffffull);
                        break;
                    }
                    s = (s + 1u) & mask;
                }
            }
            // one atomic per wave on the distinct-label counter
            const int32_t wave_claimed = __builtin_amdgcn_readfirstlane(
                shfl_i32(wave_incl_sum((int32_t)claimed), 63));
            if (lane == 0 && wave_claimed &&
                atomicAdd(s_distinct, (uint32_t)wave_claimed) + (uint32_t)wave_claimed > limit)
                atomicExch(s_over, 1u);
        }
    }
}

// Counts one non-empty part in the LDS table at s_dyn, which is all zero on
// entry and is left all zero (the scan clears what it reads).
template <class H>
__device__ __forceinline__ void hub_part_count(const StepArgs& a, const HubArgs& h, const PartJob& j,
                                               unsigned long long* s_red, uint32_t* s_distinct,
                                               uint32_t* s_over) {
    extern __shared__ __align__(16) uint32_t s_dyn[];
    const uint32_t n = j.total;
    const uint32_t cap_max = next_pow2(2u * (uint32_t)h.part_keys);
    // a bucket of n <= cap_max/2 pairs gets a table of >= 2n slots: it holds
    // at most n distinct labels, so it cannot overflow, and converged buckets
    // (a few pairs) scan a few slots instead of the whole table
    uint32_t cap = cap_max;
    if (h.part_limit == 0 && n <= cap_max / 2u) cap = max(next_pow2(2u * n), 64u);
    const uint32_t shift = 32u - (31u - __builtin_clz(cap));
    unsigned long long* T = reinterpret_cast<unsigned long long*>(s_dyn);  // cap slots
    const uint64_t end = bucket_lo(j.q + 1, j.P);
    uint64_t lo = bucket_lo(j.q, j.P), hi = end;
    const uint32_t limit = h.part_limit > 0 ? (uint32_t)h.part_limit : cap / 2u;
    uint64_t best = 0ull;
    while (lo < end) {
        if (threadIdx.x == 0) {
            *s_distinct = 0u;
            *s_over = 0u;
        }
        __syncthreads();
        part_pass(h, j, lo, hi, T, shift, limit, s_distinct, s_over);
        __syncthreads();
        if (*s_over != 0u) {
            // the pass stopped part-way: clear the table and retry the lower
            // half of the range (a range no wider than `limit` cannot
            // overflow, so the width stays >= 1)
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < 2u * cap; i += kHubThreads) s_dyn[i] = 0u;
            hi = lo + max((hi - lo) / 2, (uint64_t)1);
            continue;
        }
        best = max_u64(best, table_best<H>(a, j.v, T, cap, threadIdx.x, kHubThreads, true));
        lo = hi;
        hi = end;
    }
    best = block_max_u64<kHubThreads>(best, s_red);  // its barrier also ends the scan
    if (threadIdx.x == 0 && best) atomicMax(&h.hs.best[j.hub], (unsigned long long)best);
}

// One workgroup per batch of kPartBatch parts: each wave sums the run
// lengths of a quarter of them, the non-empty ones are listed in LDS and
// counted one after another.  Once labels converge nearly every bucket is
// empty, and a launch of one workgroup per part would spend its time
// starting and ending them; while labels are diverse every bucket is full,
// and a long batch would count its parts in series on one CU.  16 parts per
// workgroup measured best over both (hub_part on RMAT-25: diverse steps 2.8
// -> 1.8 ms against 256 parts, converged 0.11 -> 0.085 ms).
constexpr int kPartBatch = 16;

template <class H>
__device__ __forceinline__ void hub_part_batch_body(const StepArgs& a, const HubArgs& h,
                                                    uint32_t blk) {
    extern __shared__ __align__(16) uint32_t s_dyn[];
    __shared__ PartJob s_jobs[kPartBatch];
    __shared__ unsigned long long s_red[kHubThreads / 64];
    __shared__ uint32_t s_n, s_distinct, s_over;
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    const int lane = lane_id();
    for (int k = threadIdx.x >> 6; k < kPartBatch; k += kHubThreads / 64) {
        const int32_t p = (int32_t)blk * kPartBatch + k;
        if (p >= h.nparts) break;  // wave-uniform
        const int2 hp = h.parts[p];
        const int32_t v = h.hubs[hp.x];
        const int32_t d = a.off[v + 1] - a.off[v];
        const uint32_t nch = (uint32_t)((d + h.chunk - 1) / h.chunk);
        const uint32_t* cnt = h.hs.hist + h.hist_base[hp.x] + (int64_t)hp.y * nch;
        uint32_t sum = 0u;
        for (uint32_t c = (uint32_t)lane; c < nch; c += 64) sum += cnt[c];
        sum = wave_sum_u32(sum);
        if (lane == 0 && sum)
            s_jobs[atomicAdd(&s_n, 1u)] = PartJob{hp.x, v, (uint32_t)hp.y, hub_buckets(d, h.part_keys), nch, sum};
    }
    __syncthreads();
    const uint32_t n = s_n;
    if (n == 0u) return;  // block-uniform
    const uint32_t words = 2u * next_pow2(2u * (uint32_t)h.part_keys);
    for (uint32_t i = threadIdx.x; i < words; i += kHubThreads) s_dyn[i] = 0u;
    // (the first count's barrier orders this clear before its inserts)
    for (uint32_t i = 0; i < n; ++i) hub_part_count<H>(a, h, s_jobs[i], s_red, &s_distinct, &s_over);
}

template <class H>
__device__ __forceinline__ void hub_final_body(const StepArgs& a, const HubArgs& h, uint32_t blk) {
    const int32_t i = (int32_t)(blk * 256u + threadIdx.x);
    bool changed = false;
    if (i < h.nhubs) {
        const int32_t v = h.hubs[i];
        const unsigned long long best = h.hs.best[i];
        h.hs.best[i] = 0ull;
        if (is_active<H>(a, v) && !is_fixed<H>(a, v)) {
            const int32_t old = a.keep[v];
            const int32_t nl = best ? packed_label(best) : old;
            const bool gch = gather_changed(a, v, nl, old);
            H::update_vertex(a, v, nl, best);
            changed = nl != old;
            if (gch) note_changed(a, v);
        }
        if (marking<H>(a)) {
            h.hs.changed[i] = changed ? 1u : 0u;
            if (changed) a.act_w[v] = 1;
        }
    }
    count_changes(a, changed);
}

// Frontier: a changed hub marks its neighbours, one workgroup per chunk.
template <class H>
__device__ __forceinline__ void hub_mark_body(const StepArgs& a, const HubArgs& h, uint32_t blk) {
    if (!marking<H>(a)) return;
    const int2 hc = h.chunks[blk];
    if (!h.hs.changed[hc.x]) return;
    const int32_t v = h.hubs[hc.x];
    const int32_t vb = a.off[v], ve = a.off[v + 1];
    const int32_t b = vb + hc.y * h.chunk;
    const int32_t e = min(b + h.chunk, ve);
    for (int32_t j = b + (int32_t)threadIdx.x; j < e; j += 256) a.act_w[a.nbr[j]] = 1;
}

// ------------------------------------------------------------- wrappers ----
// Single job: workgroup index = blockIdx.x.

// One persistent small-tier launch's work: the unit ranges, heaviest first
// (half tier with 16 lanes per vertex and 4, then 2 edges per lane; packed
// groups); end[] holds the cumulative unit counts.
constexpr int kSmallNG = 4;  // packed groups per unit
constexpr int kSmallNP = 2;  // half-tier quarter-wave units per unit (8 vertices)
struct SmallTiers {
    const int4* half_long;
    const int4* half_short;
    const int2* groups;
    int32_t n_half_long, n_half_short, ngroups, nge;
    uint32_t end[3];
    uint32_t* counter;  // zero at launch; a wave takes kSmallGrab units per atomic
};
constexpr uint32_t kSmallGrab = 16;

// A wave's unit in a launch of 256-thread workgroups.
__device__ __forceinline__ uint32_t wave_unit(uint32_t blk) { return blk * 4u + (threadIdx.x >> 6); }

template <class H, bool W, bool G, int NG>
__global__ __launch_bounds__(256) void k_packed(StepArgs a, const int2* __restrict__ groups,
                                                int32_t ngroups, int32_t nge) {
    __shared__ PackedLds s_pk[4];
    packed_body<H, W, G, NG>(a, groups, ngroups, nge, wave_unit(blockIdx.x), s_pk[threadIdx.x >> 6]);
}
template <class H, bool W, bool G, int NP, int S, int EPL>
__global__ __launch_bounds__(256) void k_half(StepArgs a, const int4* __restrict__ verts,
                                              int32_t nv) {
    __shared__ HalfLds<S> s_half[4];
    half_body<H, W, G, NP, S, EPL>(a, verts, nv, wave_unit(blockIdx.x), s_half[threadIdx.x >> 6]);
}
template <class H, bool W, bool G, int EPL>
__global__ __launch_bounds__(256) void k_wave(StepArgs a, const int4* __restrict__ verts,
                                              int32_t nv) {
    __shared__ unsigned long long s_tab[4][kWaveCap];
    wave_body<H, W, G, EPL>(a, verts, nv, wave_unit(blockIdx.x), s_tab[threadIdx.x >> 6]);
}

// The packed and half tiers of a step in one persistent launch (tuning
// fused_small=1): each wave takes kSmallGrab consecutive units per atomic on
// a counter until none are left, the heavier half-tier units first.  (One
// atomic per unit serialises millions of them on one address.)  The two
// tiers' per-wave LDS areas overlap in a union.  Two half units per wave
// unit keep the kernel at 7 waves per SIMD; four would take 92 VGPRs (5
// waves).  Measured slower than one launch per tier replayed from a HIP
// graph (the batched planted mix: 7.8 against 6.8 ms per step), because the
// separate half launch keeps four units in flight per wave at 6 waves per
// SIMD; it is kept as an option the autotuner tries.
union SmallLds {
    PackedLds pk;
    HalfLds<16> hf;
};
template <class H, bool W, bool G>
__global__ __launch_bounds__(256) void k_small_tiers(StepArgs a, SmallTiers st) {
    __shared__ SmallLds s_lds[4];
    SmallLds& my = s_lds[threadIdx.x >> 6];
    for (;;) {
        uint32_t u0 = 0u;
        if (lane_id() == 0) u0 = atomicAdd(st.counter, kSmallGrab);
        u0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)u0);
        if (u0 >= st.end[2]) break;  // wave-uniform
        const uint32_t u1 = min(u0 + kSmallGrab, st.end[2]);
        for (uint32_t u = u0; u < u1; ++u) {
            if (u < st.end[0])
                half_body<H, W, G, kSmallNP, 16, 4>(a, st.half_long, st.n_half_long, u, my.hf);
            else if (u < st.end[1])
                half_body<H, W, G, kSmallNP, 16, 2>(a, st.half_short, st.n_half_short, u - st.end[0], my.hf);
            else
                packed_body<H, W, G, kSmallNG>(a, st.groups, st.ngroups, st.nge, u - st.end[1], my.pk);
            wave_lds_sync();  // the next unit's tier reuses the area
        }
    }
}
template <class H, bool W, bool G, int NT, int K = 8>
__global__ __launch_bounds__(NT) void k_block(StepArgs a, const int32_t* __restrict__ verts,
                                              uint32_t cap_max) {
    block_body<H, W, G, NT, K>(a, verts, cap_max, blockIdx.x);
}
template <class H, bool W, bool G>
__global__ __launch_bounds__(kHubAggThreads) void k_hub_agg(StepArgs a, HubArgs h) {
    hub_agg_body<H, W, G, kHubAggThreads, 4>(a, h, blockIdx.x);
}
template <class H>
__global__ __launch_bounds__(kHubThreads) void k_hub_part(StepArgs a, HubArgs h) {
    hub_part_batch_body<H>(a, h, blockIdx.x);
}
template <class H>
__global__ __launch_bounds__(256) void k_hub_final(StepArgs a, HubArgs h) {
    hub_final_body<H>(a, h, blockIdx.x);
}
template <class H>
__global__ __launch_bounds__(256) void k_hub_mark(StepArgs a, HubArgs h) {
    hub_mark_body<H>(a, h, blockIdx.x);
}

// Folded jobs: cum[j] is the first workgroup of job j in this launch
// (cum[njobs] = grid size); a workgroup finds its job by binary search.
__device__ __forceinline__ int32_t job_of(const int32_t* __restrict__ cum, int32_t njobs,
                                          uint32_t b) {
    int32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {  // largest j with cum[j] <= b
        const int32_t mid = (lo + hi + 1) >> 1;
        if ((uint32_t)cum[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class H, bool W, bool G, int L, int NT>
__global__ __launch_bounds__(NT) void k_multi(const JobDesc* __restrict__ jobs,
                                              const int32_t* __restrict__ cum, int32_t njobs,
                                              uint32_t cap_max) {
    const int32_t j = job_of(cum, njobs, blockIdx.x);
    const uint32_t blk = blockIdx.x - (uint32_t)cum[j];
    const JobDesc& jd = jobs[j];
    if constexpr (L == kLaunchPacked) {
        __shared__ PackedLds s_pk[4];
        packed_body<H, W, G, kFoldNG>(jd.a, jd.groups, jd.ngroups, jd.ngroups_e, wave_unit(blk),
                                      s_pk[threadIdx.x >> 6]);
    } else if constexpr (L == kLaunchHalf) {
        __shared__ HalfLds<32> s_half[4];
        half_body<H, W, G, kFoldNG>(jd.a, jd.half, jd.nhalf, wave_unit(blk), s_half[threadIdx.x >> 6]);
    } else if constexpr (L == kLaunchWave) {
        __shared__ unsigned long long s_tab[4][kWaveCap];
        wave_body<H, W, G>(jd.a, jd.wave, jd.nwave, wave_unit(blk), s_tab[threadIdx.x >> 6]);
    } else if constexpr (L >= kLaunchBlock0 && L < kLaunchBlock0 + kMaxBlockClasses) {
        block_body<H, W, G, NT>(jd.a, jd.block[L - kLaunchBlock0], cap_max, blk);
    } else if constexpr (L == kLaunchHubAgg) {
        hub_agg_body<H, W, G>(jd.a, jd.hub, blk);
    } else if constexpr (L == kLaunchHubPart) {
        hub_part_batch_body<H>(jd.a, jd.hub, blk);
    } else {
        hub_final_body<H>(jd.a, jd.hub, blk);
    }
}

// Folded launches of one step for `nj` jobs (the serving path).  cum holds
// kLaunches rows of nj+1 prefix sums; grid[L] = cum row L's last entry.
template <class H, bool W, int C>
static void launch_multi_block(const JobDesc* jobs, const int32_t* cum, int32_t nj,
                               const int32_t* grid, const Tuning& t, uint32_t cap, hipStream_t s) {
    constexpr int L = kLaunchBlock0 + C;
    if (!grid[L]) return;
    const size_t lds = (size_t)cap * 2 * sizeof(uint32_t);
    const int32_t* row = cum + (size_t)L * (nj + 1);
    if (t.block_threads == 1024)
        hipLaunchKernelGGL((k_multi<H, W, false, L, 1024>), dim3(grid[L]), dim3(1024), lds, s, jobs, row, nj, cap);
    else if (t.block_threads == 512)
        hipLaunchKernelGGL((k_multi<H, W, false, L, 512>), dim3(grid[L]), dim3(512), lds, s, jobs, row, nj, cap);
    else
        hipLaunchKernelGGL((k_multi<H, W, false, L, 256>), dim3(grid[L]), dim3(256), lds, s, jobs, row, nj, cap);
}

template <class H, bool W>
static int launch_multi_t(const JobDesc* jobs, const int32_t* cum, int32_t nj, const int32_t* grid,
                          const Tuning& t, const int32_t* class_cap, hipStream_t s) {
    auto row = [&](int L) { return cum + (size_t)L * (nj + 1); };
    if (grid[kLaunchPacked])
        hipLaunchKernelGGL((k_multi<H, W, false, kLaunchPacked, 256>), dim3(grid[kLaunchPacked]),
                           dim3(256), 0, s, jobs, row(kLaunchPacked), nj, 0u);
    if (grid[kLaunchHalf])
        hipLaunchKernelGGL((k_multi<H, W, false, kLaunchHalf, 256>), dim3(grid[kLaunchHalf]),
                           dim3(256), 0, s, jobs, row(kLaunchHalf), nj, 0u);
    if (grid[kLaunchWave])
        hipLaunchKernelGGL((k_multi<H, W, false, kLaunchWave, 256>), dim3(grid[kLaunchWave]),
                           dim3(256), 0, s, jobs, row(kLaunchWave), nj, 0u);
    launch_multi_block<H, W, 0>(jobs, cum, nj, grid, t, (uint32_t)class_cap[0], s);
    launch_multi_block<H, W, 1>(jobs, cum, nj, grid, t, (uint32_t)class_cap[1], s);
    launch_multi_block<H, W, 2>(jobs, cum, nj, grid, t, (uint32_t)class_cap[2], s);
    launch_multi_block<H, W, 3>(jobs, cum, nj, grid, t, (uint32_t)class_cap[3], s);
    launch_multi_block<H, W, 4>(jobs, cum, nj, grid, t, (uint32_t)class_cap[4], s);
    if (grid[kLaunchHubAgg]) {
        const size_t lds = hub_agg_lds(t.hub_chunk);
        hipLaunchKernelGGL((k_multi<H, W, false, kLaunchHubAgg, kHubThreads>), dim3(grid[kLaunchHubAgg]),
                           dim3(kHubThreads), lds, s, jobs, row(kLaunchHubAgg), nj, 0u);
    }
    if (grid[kLaunchHubPart]) {
        const size_t lds = (size_t)next_pow2(2u * (uint32_t)t.part_keys) * 2 * sizeof(uint32_t);
        hipLaunchKernelGGL((k_multi<H, W, false, kLaunchHubPart, kHubThreads>), dim3(grid[kLaunchHubPart]),
                           dim3(kHubThreads), lds, s, jobs, row(kLaunchHubPart), nj, 0u);
    }
    if (grid[kLaunchHubFinal])
        hipLaunchKernelGGL((k_multi<H, W, false, kLaunchHubFinal, 256>), dim3(grid[kLaunchHubFinal]),
                           dim3(256), 0, s, jobs, row(kLaunchHubFinal), nj, 0u);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "folded launch: %s", hipGetErrorString(e));
}

int launch_multi(bool weighted, bool slot_scores, const JobDesc* jobs, const int32_t* cum, int32_t nj,
                 const int32_t* grid, const Tuning& t, const int32_t* class_cap, hipStream_t s) {
    if (slot_scores) {
        if (weighted) return launch_multi_t<LayeredVolHooks, true>(jobs, cum, nj, grid, t, class_cap, s);
        return launch_multi_t<LayeredVolHooks, false>(jobs, cum, nj, grid, t, class_cap, s);
    }
    if (weighted) return launch_multi_t<LayeredHooks, true>(jobs, cum, nj, grid, t, class_cap, s);
    return launch_multi_t<LayeredHooks, false>(jobs, cum, nj, grid, t, class_cap, s);
}

// Workgroups each launch of a folded step needs for one job.
void job_blocks(const Plan& p, int32_t* blocks) {
    for (int L = 0; L < kLaunches; ++L) blocks[L] = 0;
    // folded jobs run the layered or slot-score hooks, which skip edge-less groups
    static_assert(LayeredHooks::kSkipIsolated && LayeredVolHooks::kSkipIsolated,
                  "folded launches run groups with edges only");
    blocks[kLaunchPacked] = (int32_t)((p.groups_with_edges + 4 * kFoldNG - 1) / (4 * kFoldNG));
    blocks[kLaunchHalf] = (int32_t)((p.half.size() + 8 * kFoldNG - 1) / (8 * kFoldNG));
    blocks[kLaunchWave] = (int32_t)((p.wave.size() + 4 * kWaveNV - 1) / (4 * kWaveNV));
    for (int c = 0; c < p.nclasses; ++c) blocks[kLaunchBlock0 + c] = (int32_t)p.block[c].size();
    blocks[kLaunchHubAgg] = (int32_t)p.hub_chunks.size();
    blocks[kLaunchHubPart] = (int32_t)((p.hub_parts.size() + kPartBatch - 1) / kPartBatch);
    blocks[kLaunchHubFinal] = (int32_t)((p.hubs.size() + 255) / 256);
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
__device__ __forceinline__ void gather_all(const int32_t* __restrict__ nbr,
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

__global__ __launch_bounds__(256) void k_gather(const int32_t* __restrict__ nbr,
                                                const int32_t* __restrict__ L,
                                                int32_t* __restrict__ adj, int64_t m) {
    gather_all(nbr, L, adj, m);
}

// ---------------------------------------------------- frontier gather ----
// The frontier form of the separate gather (header/kernel.cuh:21-78's
// if_update gathers) pushes labels from sources instead of pulling them per
// edge.  An inverse index lists, for each vertex u, the CSR positions j with
// nbr[j] = u (the reference gets the same lists by sorting (nbr[j], j) pairs
// with thrust::sort_by_key, prop/layeredLprop.cuh:228-237, and reads them in
// gather_labels_inverse, header/kernel.cuh:64-78; a counting pass builds
// them here in O(m)).  adj holds the labels of two steps back, so writing
// L[u] to the positions of every u that changed in the previous step makes
// it equal L[nbr]: the work is the changed sources' in-edges, not m.
//
// With the two-step lookahead (the reference's onestep = false,
// tutorial/Lprop.cuh:383-461) adj alternates between two buffers, each
// refreshed every other step, and a vertex's bit records a change against
// its label of two steps back: a vertex that flips between two labels then
// pushes nothing.
//
// Frontier control words used here (the rest of fctl is the tiers'):
//   fctl[3] = steps adj has been gathered in, at most 2 (set by k_end_step,
//             0 after a reset), fctl[4] = push entries listed for this step,
//   fctl[5] = in-edges of the sources that changed in the previous step,
//   fctl[6] = 1 when that step changed at most fctl[7] labels (a changed
//             vertex has an in-edge, so more changes than that rule the push
//             out before its sources are counted),
//   fctl[7] = m / kPushRatio, the most in-edges a push takes (set at reset).
// A push writes each in-edge's label to a random position; measured on
// RMAT-25 that costs about five times a streamed edge of the full gather, so
// the push is taken only below m / 8 in-edges.
constexpr int32_t kPushChunk = 256;   // in-edges per push entry at most
constexpr int64_t kPushRatio = 8;

__global__ __launch_bounds__(256) void k_inv_count(const int32_t* __restrict__ nbr, int64_t m,
                                                   int32_t* __restrict__ cnt) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride)
        atomicAdd(&cnt[nbr[j]], 1);
}
// cursor[u] starts at inv_off[u]; each edge claims the next slot of its
// source's list (the order inside a list does not matter).
__global__ __launch_bounds__(256) void k_inv_fill(const int32_t* __restrict__ nbr, int64_t m,
                                                  int32_t* __restrict__ cursor,
                                                  int32_t* __restrict__ inv_pos) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride)
        inv_pos[atomicAdd(&cursor[nbr[j]], 1)] = (int32_t)j;
}

// Exclusive scan of int32 counts in place (sums < 2^31): block sums, a scan
// of the sums in one workgroup, then each block rescans its run.
constexpr int kScanPer = 16;                       // elements per thread
constexpr int kScanBlock = 256 * kScanPer;         // elements per workgroup
__global__ __launch_bounds__(256) void k_scan_sums(const int32_t* __restrict__ x, int64_t count,
                                                   int32_t* __restrict__ sums) {
    const int64_t b0 = (int64_t)blockIdx.x * kScanBlock + (int64_t)threadIdx.x * kScanPer;
    int32_t s = 0;
    for (int i = 0; i < kScanPer; ++i)
        if (b0 + i < count) s += x[b0 + i];
    int32_t total;
    (void)block_excl_sum(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}
__global__ __launch_bounds__(256) void k_scan_top(int32_t* __restrict__ sums, int64_t nsums) {
    int32_t carry = 0;
    for (int64_t b = 0; b < nsums; b += 256) {
        const int64_t i = b + threadIdx.x;
        const int32_t x = i < nsums ? sums[i] : 0;
        int32_t total;
        const int32_t ex = block_excl_sum(x, &total);
        if (i < nsums) sums[i] = carry + ex;
        carry += total;
    }
}
__global__ __launch_bounds__(256) void k_scan_apply(int32_t* __restrict__ x, int64_t count,
                                                    const int32_t* __restrict__ sums) {
    const int64_t b0 = (int64_t)blockIdx.x * kScanBlock + (int64_t)threadIdx.x * kScanPer;
    int32_t v[kScanPer];
    int32_t s = 0;
    for (int i = 0; i < kScanPer; ++i) {
        v[i] = b0 + i < count ? x[b0 + i] : 0;
        s += v[i];
    }
    int32_t total;
    int32_t run = sums[blockIdx.x] + block_excl_sum(s, &total);
    for (int i = 0; i < kScanPer; ++i) {
        if (b0 + i < count) x[b0 + i] = run;
        run += v[i];
    }
}

// One push entry: label to write, first in-edge position, in-edge count.
struct PushEntry {
    int32_t label, begin, count, pad;
};

// Push entries of the sources whose bit is set in chg (the previous step's
// changes), one per kPushChunk in-edges, in vertex order: each thread takes
// one bitmap word; k_push_count sums every workgroup's entries and in-edges,
// k_push_scan turns the sums into offsets and totals (fctl[4], fctl[5]), and
// k_push_write lists the entries when the gather is going to push them.
// No atomics: the totals of a step with many changes come at the cost of one
// read of the bitmap.
__device__ __forceinline__ void push_word_counts(const uint32_t* __restrict__ chg, int64_t w,
                                                 int64_t nwords, const int32_t* __restrict__ inv_off,
                                                 int32_t* cnt, int32_t* edges) {
    const uint32_t bits = w < nwords ? chg[w] : 0u;
    int32_t c = 0, e = 0;
    if (__builtin_popcount(bits) >= 8) {
        // many changed: read the word's 33 offsets at once (inv_off + 32w is
        // 128-byte aligned, and inv_off has n + 1 >= 32w + 33 entries when
        // bit 31 can be set; the last word reads up to its highest bit + 1)
        const int32_t* o = inv_off + w * 32;
        const int hi = 32 - __builtin_clz(bits);  // highest set bit + 1
        int32_t prev = o[0];
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            if (i < hi) {
                const int32_t next = o[i + 1];
                if ((bits >> i) & 1u) {
                    const int32_t d = next - prev;
                    c += (d + kPushChunk - 1) / kPushChunk;
                    e += d;
                }
                prev = next;
            }
        }
    } else {
        for (uint32_t b = bits; b; b &= b - 1u) {
            const int32_t u = (int32_t)(w * 32 + __builtin_ctz(b));
            const int32_t d = inv_off[u + 1] - inv_off[u];
            c += (d + kPushChunk - 1) / kPushChunk;
            e += d;
        }
    }
    *cnt = c;
    *edges = e;
}
// need: steps adj must have been gathered in (1; 2 with the lookahead,
// whose two buffers are each first filled by a full gather).  The rule-out
// by the previous step's change count holds for one-step changes only.
__device__ __forceinline__ bool push_possible(const int32_t* fctl, int32_t need) {
    return fctl[3] >= need && (need == 2 || fctl[6] != 0);
}
__global__ __launch_bounds__(256) void k_push_count(const uint32_t* __restrict__ chg, int64_t nwords,
                                                    const int32_t* __restrict__ inv_off,
                                                    int2* __restrict__ psum,
                                                    const int32_t* __restrict__ fctl, int32_t need) {
    if (!push_possible(fctl, need)) return;  // the gather rewrites every edge this step
    int32_t c, e, tc, te;
    push_word_counts(chg, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, nwords, inv_off, &c, &e);
    (void)block_excl_sum(c, &tc);
    (void)block_excl_sum(e, &te);
    if (threadIdx.x == 0) psum[blockIdx.x] = make_int2(tc, te);
}
__global__ __launch_bounds__(256) void k_push_scan(int2* __restrict__ psum, int64_t nblocks,
                                                   int32_t* __restrict__ fctl, int32_t need) {
    if (!push_possible(fctl, need)) return;
    int32_t carry = 0, edges = 0;
    for (int64_t b = 0; b < nblocks; b += 256) {
        const int64_t i = b + threadIdx.x;
        const int2 x = i < nblocks ? psum[i] : make_int2(0, 0);
        int32_t tc, te;
        const int32_t ex = block_excl_sum(x.x, &tc);
        (void)block_excl_sum(x.y, &te);
        if (i < nblocks) psum[i].x = carry + ex;
        carry += tc;
        edges += te;
    }
    if (threadIdx.x == 0) {
        fctl[4] = carry;
        fctl[5] = edges;
    }
}
__device__ __forceinline__ bool push_taken(const int32_t* fctl, int32_t need) {
    return push_possible(fctl, need) && fctl[5] <= fctl[7];
}
__global__ __launch_bounds__(256) void k_push_write(const uint32_t* __restrict__ chg, int64_t nwords,
                                                    const int32_t* __restrict__ inv_off,
                                                    const int32_t* __restrict__ L,
                                                    const int2* __restrict__ psum,
                                                    PushEntry* __restrict__ ent,
                                                    const int32_t* __restrict__ fctl, int32_t need) {
    if (!push_taken(fctl, need)) return;
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t c, e, tc;
    push_word_counts(chg, w, nwords, inv_off, &c, &e);
    int32_t k = psum[blockIdx.x].x + block_excl_sum(c, &tc);
    if (c == 0) return;
    for (uint32_t b = chg[w]; b; b &= b - 1u) {
        const int32_t u = (int32_t)(w * 32 + __builtin_ctz(b));
        const int32_t end = inv_off[u + 1];
        const int32_t lab = L[u];
        for (int32_t p = inv_off[u]; p < end; p += kPushChunk)
            ent[k++] = PushEntry{lab, p, min(kPushChunk, end - p), 0};
    }
}

// The gather of a frontier step: the first step, and any step whose changed
// sources have more than m / kPushRatio in-edges, rewrite every edge
// (gather_all); otherwise each wave takes 64 push entries at a time and
// spreads their in-edges evenly over its lanes.
__global__ __launch_bounds__(256) void k_gather_frontier(const int32_t* __restrict__ nbr,
                                                         const int32_t* __restrict__ L,
                                                         int32_t* __restrict__ adj, int64_t m,
                                                         const int32_t* __restrict__ inv_pos,
                                                         const PushEntry* __restrict__ ent,
                                                         const int32_t* __restrict__ fctl, int32_t need) {
    if (!push_taken(fctl, need)) {
        gather_all(nbr, L, adj, m);
        return;
    }
    const int64_t nent = fctl[4];
    const int lane = lane_id();
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    for (int64_t b0 = wave * 64; b0 < nent; b0 += nwaves * 64) {
        PushEntry e{0, 0, 0, 0};
        if (b0 + lane < nent) e = ent[b0 + lane];
        const int32_t incl = wave_incl_sum(e.count);
        const int32_t total = shfl_i32(incl, 63);
        for (int32_t f0 = 0; f0 < total; f0 += 64) {
            const int32_t f = f0 + lane;
            // owner: the smallest lane whose inclusive sum exceeds f
            int lo = 0, hi = 63;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int mid = (lo + hi) >> 1;
                if (shfl_i32(incl, mid) <= f) lo = mid + 1;
                else hi = mid;
            }
            const int32_t o_incl = shfl_i32(incl, lo);
            const int32_t o_cnt = shfl_i32(e.count, lo);
            const int32_t o_beg = shfl_i32(e.begin, lo);
            const int32_t o_lab = shfl_i32(e.label, lo);
            if (f < total) adj[inv_pos[o_beg + f - (o_incl - o_cnt)]] = o_lab;
        }
    }
}

__global__ void k_fill_i32(int32_t* p, int32_t value, int64_t count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        p[i] = value;
}
// Layered LP's volume V(l) = #{v : L[v] = l} of the labels a step wrote
// (header/kernel.cuh:2626's gcounter_temp, counted in a pass of its own).
// While labels consolidate, a few labels each hold millions of vertices
// spread over the id range, and one global atomic per vertex on their
// counters serialises in L2.  A workgroup counts a chunk of kVolChunk
// vertices in an LDS table of (label+1, count) slots (each wave first merges
// its row's first label) and flushes one atomic per distinct label; a label
// that finds no slot within kVolProbe probes goes to global memory directly.
constexpr int kVolThreads = 256;
constexpr int kVolChunk = 16384;
constexpr uint32_t kVolSlots = 2048;
constexpr int kVolProbe = 8;
__device__ __forceinline__ void volume_body(const int32_t* __restrict__ L, int32_t* __restrict__ V,
                                            int32_t n, uint32_t blk) {
    __shared__ unsigned long long T[kVolSlots];
    const int lane = lane_id();
    for (uint32_t i = threadIdx.x; i < kVolSlots; i += kVolThreads) T[i] = 0ull;
    __syncthreads();
    const int64_t c0 = (int64_t)blk * kVolChunk;
    const int64_t c1 = min((int64_t)n, c0 + kVolChunk);
    for (int64_t r = c0 + (threadIdx.x & ~63); r < c1; r += kVolThreads) {  // wave-uniform rows
        const int64_t v = r + lane;
        int32_t l = v < c1 ? L[v] : -1;
        uint32_t c = 1u;
        const unsigned long long act = __ballot(l >= 0);
        const int first = (int)__builtin_ctzll(act);
        const int32_t lead = __builtin_amdgcn_readlane(l, first);
        const unsigned long long sm = __ballot(l == lead);
        if (l == lead) {
            if (lane == first) c = (uint32_t)__popcll(sm);
            else l = -1;  // merged into the first lane
        }
        if (l < 0) continue;
        const uint32_t key = (uint32_t)l + 1u;
        uint32_t h = fmix32(key) & (kVolSlots - 1u);
        bool done = false;
        for (int k = 0; k < kVolProbe && !done; ++k) {
            unsigned long long cur = lds_load(&T[h]);
            if (cur == 0ull) {
                cur = atomicCAS(&T[h], 0ull, ((unsigned long long)key << 32) | c);
                if (cur == 0ull) done = true;
            }
            if (!done && (uint32_t)(cur >> 32) == key) {
                atomicAdd(&T[h], (unsigned long long)c);
                done = true;
            }
            h = (h + 1u) & (kVolSlots - 1u);
        }
        if (!done) atomicAdd(&V[l], (int32_t)c);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kVolSlots; i += kVolThreads) {
        const unsigned long long x = T[i];
        if (x) atomicAdd(&V[(uint32_t)(x >> 32) - 1u], (int32_t)(uint32_t)x);
    }
}
__global__ __launch_bounds__(kVolThreads) void k_volume(const int32_t* __restrict__ L,
                                                        int32_t* __restrict__ V, int32_t n) {
    volume_body(L, V, n, blockIdx.x);
}

// Folded forms for the serving path: one launch zeroes (before the tiers)
// or fills (after them) the write-side volumes of every layered job of a
// step; cum[j] is job j's first workgroup, as for k_multi.
constexpr int kZeroPerBlock = 256 * 16;
__global__ __launch_bounds__(256) void k_multi_zero(const VolumeJob* __restrict__ jobs,
                                                    const int32_t* __restrict__ cum, int32_t nj) {
    const int32_t j = job_of(cum, nj, blockIdx.x);
    const VolumeJob& jb = jobs[j];
    const int64_t b0 = (int64_t)(blockIdx.x - (uint32_t)cum[j]) * kZeroPerBlock;
    const int64_t b1 = min((int64_t)jb.n, b0 + kZeroPerBlock);
    for (int64_t i = b0 + threadIdx.x; i < b1; i += 256) jb.V[i] = 0;
}
__global__ __launch_bounds__(kVolThreads) void k_multi_volume(const VolumeJob* __restrict__ jobs,
                                                              const int32_t* __restrict__ cum, int32_t nj) {
    const int32_t j = job_of(cum, nj, blockIdx.x);
    volume_body(jobs[j].L, jobs[j].V, jobs[j].n, blockIdx.x - (uint32_t)cum[j]);
}
// The slot-score jobs' records LV[u] = {L[u], V[L[u]]}, kZeroPerBlock rows
// per workgroup.
__global__ __launch_bounds__(256) void k_multi_records(const VolumeJob* __restrict__ jobs,
                                                       const int32_t* __restrict__ cum, int32_t nj) {
    const int32_t j = job_of(cum, nj, blockIdx.x);
    const VolumeJob& jb = jobs[j];
    const int64_t b0 = (int64_t)(blockIdx.x - (uint32_t)cum[j]) * kZeroPerBlock;
    const int64_t b1 = min((int64_t)jb.n, b0 + kZeroPerBlock);
    for (int64_t i = b0 + threadIdx.x; i < b1; i += 256) {
        const int32_t l = jb.L[i];
        jb.LV[i] = make_int2(l, jb.V[l]);
    }
}

void volume_job_blocks(int32_t n, int32_t* zero_blocks, int32_t* volume_blocks) {
    *zero_blocks = (int32_t)(((int64_t)n + kZeroPerBlock - 1) / kZeroPerBlock);
    *volume_blocks = (int32_t)(((int64_t)n + kVolChunk - 1) / kVolChunk);
}

int launch_multi_volume(VolumeMode mode, const VolumeJob* jobs, const int32_t* cum, int32_t nj, int32_t grid,
                        hipStream_t s) {
    if (nj == 0 || grid == 0) return GLP_OK;
    if (mode == kVolumeZero) hipLaunchKernelGGL(k_multi_zero, dim3(grid), dim3(256), 0, s, jobs, cum, nj);
    else if (mode == kVolumeRecords) hipLaunchKernelGGL(k_multi_records, dim3(grid), dim3(256), 0, s, jobs, cum, nj);
    else hipLaunchKernelGGL(k_multi_volume, dim3(grid), dim3(kVolThreads), 0, s, jobs, cum, nj);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "folded volume launch: %s", hipGetErrorString(e));
}

// Seeded LP's L_0 (header/kernel.cuh:453-466): v for a seed, -1 otherwise.
__global__ void k_seed_labels(int32_t* p, const uint8_t* __restrict__ seeds, int64_t count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        p[i] = seeds[i] ? (int32_t)i : -1;
}
__global__ void k_label_volumes(const int32_t* __restrict__ L, const int32_t* __restrict__ V,
                                int2* __restrict__ LV, int32_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int32_t l = L[i];
        LV[i] = make_int2(l, V[l]);
    }
}
__global__ void k_iota(int32_t* p, int64_t count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        p[i] = (int32_t)i;
}
// Sums and clears the step's change counters into the ring and advances
// the step counter.  One workgroup of kChangeSlots threads.
// With a frontier control block fctl = {all_active, mark, threshold, ...}:
// the next step recomputes everything unless this step marked, and marks
// when this step changed at most `threshold` labels; the frontier gather's
// words 3-5 are set up for the next step.
__global__ __launch_bounds__(kChangeSlots) void k_end_step(unsigned long long* acc,
                                                           unsigned long long* ring, int32_t* ctr,
                                                           int32_t ring_size, int32_t* fctl) {
    __shared__ unsigned long long part[kChangeSlots / 64];
    unsigned long long x = acc[threadIdx.x];
    acc[threadIdx.x] = 0ull;
    const int lane = lane_id();
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += shfl_u64(x, lane ^ d);
    if (lane == 0) part[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long sum = 0ull;
        for (int i = 0; i < kChangeSlots / 64; ++i) sum += part[i];
        const int32_t t = *ctr;
        ring[t % ring_size] = sum;
        *ctr = t + 1;
        if (fctl) {
            const int32_t marked = fctl[1];
            fctl[0] = marked ? 0 : 1;
            fctl[1] = sum <= (unsigned long long)fctl[2] ? 1 : 0;
            fctl[3] = min(fctl[3] + 1, 2);  // steps a frontier gather has filled adj in
            fctl[4] = 0;
            fctl[5] = 0;
            fctl[6] = sum <= (unsigned long long)fctl[7] ? 1 : 0;
        }
    }
}

// ---------------------------------------------------------------- launch ----
// Units per wave for a launch of `units` work units: 4 when that still
// leaves >= kFillWaves waves, else 2 (>= kFillWaves / 4 waves), else 1.
constexpr int64_t kFillWaves = 16384;
static inline int units_per_wave(int64_t units) {
    if (units / 4 >= kFillWaves) return 4;
    if (units / 2 >= kFillWaves / 4) return 2;
    return 1;
}

static inline unsigned grid_for(int64_t work, int per_block) {
    int64_t g = (work + per_block - 1) / per_block;
    return (unsigned)std::max<int64_t>(1, g);
}

// The half tier with S lanes per vertex: a unit is 64/S vertices, and a
// wave takes 4, 2 or 1 units (units_per_wave).
template <class H, bool W, bool G, int S, int EPL = kHalfMaxDeg / S>
static void launch_half(const StepArgs& a, const int4* verts, int32_t nh, hipStream_t s) {
    constexpr int V = 64 / S;
    if (nh <= 0) return;
    switch (units_per_wave((nh + V - 1) / V)) {
        case 4:
            hipLaunchKernelGGL((k_half<H, W, G, 4, S, EPL>), dim3(grid_for(nh, 16 * V)), dim3(256), 0, s,
                               a, verts, nh);
            break;
        case 2:
            hipLaunchKernelGGL((k_half<H, W, G, 2, S, EPL>), dim3(grid_for(nh, 8 * V)), dim3(256), 0, s,
                               a, verts, nh);
            break;
        default:
            hipLaunchKernelGGL((k_half<H, W, G, 1, S, EPL>), dim3(grid_for(nh, 4 * V)), dim3(256), 0, s,
                               a, verts, nh);
    }
}

// Workgroups of `threads` threads of `fn` that the device holds at once.
static int64_t resident_blocks(const void* fn, int threads) {
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess || per_cu < 1 ||
        cus < 1)
        return 1024;
    return (int64_t)per_cu * cus;
}

template <class H, bool W, bool G>
static int launch_tiers(const DeviceGraph& g, const HubScratch& hs, const StepArgs& a,
                        hipStream_t s, hipEvent_t* ev) {
    const Plan& p = g.plan;
    auto mark = [&](int tier, int edge) {
        if (ev) (void)hipEventRecord(ev[2 * tier + edge], s);
    };
    mark(GLP_TIER_PACKED, 0);
    const int32_t ng = (int32_t)(H::kSkipIsolated ? p.groups_with_edges : (int64_t)p.groups.size());
    if (p.tune.fused_small > 0 && hs.work) {
        // packed and half tiers in one persistent launch
        SmallTiers st;
        const int32_t hs_n = (int32_t)p.half_short;
        st.half_short = g.d_half;
        st.n_half_short = hs_n;
        st.half_long = g.d_half + hs_n;
        st.n_half_long = (int32_t)p.half.size() - hs_n;
        st.groups = g.d_groups;
        st.ngroups = ng;
        st.nge = (int32_t)p.groups_with_edges;
        auto units = [](int64_t items, int per) { return (uint32_t)((items + per - 1) / per); };
        st.end[0] = units(st.n_half_long, 4 * kSmallNP);
        st.end[1] = st.end[0] + units(st.n_half_short, 4 * kSmallNP);
        st.end[2] = st.end[1] + units(ng, kSmallNG);
        st.counter = hs.work;
        if (st.end[2] > 0) {
            int rc = launch_zero_bytes(hs.work, 4, s);
            if (rc != GLP_OK) return rc;
            // twice what the chip holds at once: workgroups that find the
            // units taken exit at once
            static const int64_t resident = resident_blocks((const void*)k_small_tiers<H, W, G>, 256);
            const unsigned grid =
                (unsigned)std::min<int64_t>((st.end[2] + 4 * kSmallGrab - 1) / (4 * kSmallGrab), 2 * resident);
            hipLaunchKernelGGL((k_small_tiers<H, W, G>), dim3(grid), dim3(256), 0, s, a, st);
        }
    } else {
        if (ng > 0) {
            const int32_t nge = (int32_t)p.groups_with_edges;
            switch (units_per_wave(ng)) {
                case 4:
                    hipLaunchKernelGGL((k_packed<H, W, G, 4>), dim3(grid_for(ng, 16)), dim3(256), 0, s, a,
                                       g.d_groups, ng, nge);
                    break;
                case 2:
                    hipLaunchKernelGGL((k_packed<H, W, G, 2>), dim3(grid_for(ng, 8)), dim3(256), 0, s, a,
                                       g.d_groups, ng, nge);
                    break;
                default:
                    hipLaunchKernelGGL((k_packed<H, W, G, 1>), dim3(grid_for(ng, 4)), dim3(256), 0, s, a,
                                       g.d_groups, ng, nge);
            }
        }
    }
    mark(GLP_TIER_PACKED, 1);
    mark(GLP_TIER_WAVE, 0);
    if (!p.half.empty() && !(p.tune.fused_small > 0 && hs.work)) {
        const int32_t nh = (int32_t)p.half.size();
        // 16 lanes per vertex halve the instructions per vertex and double
        // the vertices in flight per wave; they pay off once the launch has
        // enough vertices to fill the chip with waves of four units
        const bool quarter = p.tune.half_lanes == 16 ||
                             (p.tune.half_lanes == 0 && nh / 16 >= kFillWaves);
        if (quarter) {
            // with 16 lanes, degrees <= 32 (listed first) take two edges per lane
            const int32_t ns = (int32_t)p.half_short;
            launch_half<H, W, G, 16, 2>(a, g.d_half, ns, s);
            launch_half<H, W, G, 16, 4>(a, g.d_half + ns, nh - ns, s);
        } else {
            launch_half<H, W, G, 32>(a, g.d_half, nh, s);
        }
    }
    if (!p.wave.empty()) {
        // degrees <= 128 (the list's first wave_short entries) load two edges
        // per lane, the rest four
        const int32_t ns = (int32_t)p.wave_short, nw = (int32_t)p.wave.size();
        if (ns > 0)
            hipLaunchKernelGGL((k_wave<H, W, G, 2>), dim3(grid_for(ns, 4 * kWaveNV)), dim3(256), 0, s, a,
                               g.d_wave, ns);
        if (nw > ns)
            hipLaunchKernelGGL((k_wave<H, W, G, 4>), dim3(grid_for(nw - ns, 4 * kWaveNV)), dim3(256), 0,
                               s, a, g.d_wave + ns, nw - ns);
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
        } else if (cap <= 1024) {
            // gathers in flight per thread sized to the class's largest
            // degree (cap / 2 edges over 256 threads)
            hipLaunchKernelGGL((k_block<H, W, G, 256, 2>), dim3(nb), dim3(256), lds, s, a,
                               g.d_block[c], cap);
        } else if (cap <= 2048) {
            hipLaunchKernelGGL((k_block<H, W, G, 256, 4>), dim3(nb), dim3(256), lds, s, a,
                               g.d_block[c], cap);
        } else if (cap <= 4096) {
            // 32 KB tables: four workgroups per CU, of 512 threads of 4
            // gathers each
            hipLaunchKernelGGL((k_block<H, W, G, 512, 4>), dim3(nb), dim3(512), lds, s, a,
                               g.d_block[c], cap);
        } else {
            // tables of 64 KB or more leave room for two workgroups per CU:
            // 1024 threads of 4 gathers each keep the CU's waves full
            hipLaunchKernelGGL((k_block<H, W, G, 1024, 4>), dim3(nb), dim3(1024), lds, s, a,
                               g.d_block[c], cap);
        }
    }
    mark(GLP_TIER_BLOCK, 1);
    mark(GLP_TIER_HUB, 0);
    if (!p.hubs.empty()) {
        HubArgs h;
        h.hubs = g.d_hubs;
        h.chunks = g.d_hub_chunks;
        h.parts = g.d_hub_parts;
        h.chunk0 = g.d_hub_chunk0;
        h.base = g.d_hub_base;
        h.hist_base = g.d_hub_hist_base;
        h.chunk = p.tune.hub_chunk;
        h.part_keys = p.tune.part_keys;
        h.part_limit = p.tune.part_limit;
        h.nhubs = (int32_t)p.hubs.size();
        h.nparts = (int32_t)p.hub_parts.size();
        h.hs = hs;
        const unsigned nc = (unsigned)p.hub_chunks.size();
        const unsigned np = (unsigned)p.hub_parts.size();
        const size_t lds_part = (size_t)next_pow2(2u * (uint32_t)h.part_keys) * 2 * sizeof(uint32_t);
        hipLaunchKernelGGL((k_hub_agg<H, W, G>), dim3(nc), dim3(kHubAggThreads),
                           hub_agg_lds(h.chunk, p.hub_max_buckets), s, a, h);
        mark(GLP_TIER_HUB, 1);
        mark(GLP_TIER_AUX, 0);
        hipLaunchKernelGGL((k_hub_part<H>), dim3((np + kPartBatch - 1) / kPartBatch), dim3(kHubThreads),
                           lds_part, s, a, h);
        hipLaunchKernelGGL((k_hub_final<H>), dim3(grid_for(h.nhubs, 256)), dim3(256), 0, s, a, h);
        if (H::kFrontier && a.act_w)
            hipLaunchKernelGGL((k_hub_mark<H>), dim3(nc), dim3(256), 0, s, a, h);
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
    init_counting_kernels();
    const int lim = 160 * 1024 - 1024;
#define GLP_SETLDS(HK, W, G)                                                                    \
    (void)hipFuncSetAttribute((const void*)k_block<HK, W, G, 256>,                              \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);                 \
    (void)hipFuncSetAttribute((const void*)k_block<HK, W, G, 256, 2>,                           \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);                 \
    (void)hipFuncSetAttribute((const void*)k_block<HK, W, G, 256, 4>,                           \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);                 \
    (void)hipFuncSetAttribute((const void*)k_block<HK, W, G, 512>,                              \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);                 \
    (void)hipFuncSetAttribute((const void*)k_block<HK, W, G, 1024>,                             \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);                 \
    (void)hipFuncSetAttribute((const void*)k_block<HK, W, G, 1024, 4>,                          \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);                 \
    (void)hipFuncSetAttribute((const void*)k_block<HK, W, G, 512, 4>,                           \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);
#define GLP_SETLDS_HUB(HK, W, G)                                                                \
    (void)hipFuncSetAttribute((const void*)k_hub_agg<HK, W, G>,                                 \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);
#define GLP_SETLDS4(HK) GLP_SETLDS(HK, false, false) GLP_SETLDS(HK, false, true) \
                        GLP_SETLDS(HK, true, false) GLP_SETLDS(HK, true, true)   \
                        GLP_SETLDS_HUB(HK, false, false) GLP_SETLDS_HUB(HK, false, true) \
                        GLP_SETLDS_HUB(HK, true, false) GLP_SETLDS_HUB(HK, true, true)   \
    (void)hipFuncSetAttribute((const void*)k_hub_part<HK>,                                      \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    GLP_SETLDS4(ClassicHooks)
    GLP_SETLDS4(LayeredHooks)
    GLP_SETLDS4(SlpaHooks)
    GLP_SETLDS4(UserHooks)
    GLP_SETLDS4(SeededHooks)
    // slot-scored layered LP runs fused only (no gathered instantiation)
    GLP_SETLDS(LayeredVolHooks, false, false)
    GLP_SETLDS(LayeredVolHooks, true, false)
    GLP_SETLDS_HUB(LayeredVolHooks, false, false)
    GLP_SETLDS_HUB(LayeredVolHooks, true, false)
    (void)hipFuncSetAttribute((const void*)k_hub_part<LayeredVolHooks>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lim);
#undef GLP_SETLDS4
#undef GLP_SETLDS_HUB
#undef GLP_SETLDS
#define GLP_SETLDS_MULTI(W, L, NT)                                                              \
    (void)hipFuncSetAttribute((const void*)k_multi<LayeredHooks, W, false, L, NT>,              \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);                 \
    (void)hipFuncSetAttribute((const void*)k_multi<LayeredVolHooks, W, false, L, NT>,           \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim);
#define GLP_SETLDS_MULTI_C(W, C) GLP_SETLDS_MULTI(W, kLaunchBlock0 + C, 256)                    \
    GLP_SETLDS_MULTI(W, kLaunchBlock0 + C, 512) GLP_SETLDS_MULTI(W, kLaunchBlock0 + C, 1024)
#define GLP_SETLDS_MULTI_W(W) GLP_SETLDS_MULTI_C(W, 0) GLP_SETLDS_MULTI_C(W, 1)                 \
    GLP_SETLDS_MULTI_C(W, 2) GLP_SETLDS_MULTI_C(W, 3) GLP_SETLDS_MULTI_C(W, 4)                   \
    GLP_SETLDS_MULTI(W, kLaunchHubAgg, kHubThreads) GLP_SETLDS_MULTI(W, kLaunchHubPart, kHubThreads)
    GLP_SETLDS_MULTI_W(false)
    GLP_SETLDS_MULTI_W(true)
#undef GLP_SETLDS_MULTI_W
#undef GLP_SETLDS_MULTI_C
#undef GLP_SETLDS_MULTI
    (void)hipGetLastError();
}

int launch_step(Method method, bool weighted, bool gathered, const DeviceGraph& g,
                const HubScratch& hs, const StepArgs& a, hipStream_t s, hipEvent_t* ev) {
    switch (method) {
        case Method::Classic: return launch_hooks<ClassicHooks>(weighted, gathered, g, hs, a, s, ev);
        case Method::Layered: return launch_hooks<LayeredHooks>(weighted, gathered, g, hs, a, s, ev);
        case Method::LayeredVol:
            if (gathered || !a.LV)
                return set_error(GLP_E_UNSUPPORTED, "slot-scored layered steps need the (label, volume) records");
            return weighted ? launch_tiers<LayeredVolHooks, true, false>(g, hs, a, s, ev)
                            : launch_tiers<LayeredVolHooks, false, false>(g, hs, a, s, ev);
        case Method::Slpa: return launch_hooks<SlpaHooks>(weighted, gathered, g, hs, a, s, ev);
        case Method::User: return launch_hooks<UserHooks>(weighted, gathered, g, hs, a, s, ev);
        case Method::Seeded: return launch_hooks<SeededHooks>(weighted, gathered, g, hs, a, s, ev);
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

int launch_frontier_gather(const StepArgs& a, int64_t m, int32_t* adj, const InverseIndex& inv,
                           void* push, hipStream_t s) {
    if (m == 0) return GLP_OK;
    if (!a.chg_r || !a.fctl || !inv.off || !inv.pos || !push)
        return set_error(GLP_E_STATE, "frontier gather without its inverse index");
    const int64_t nwords = ((int64_t)a.n + 31) / 32;
    const int64_t nblocks = (nwords + 255) / 256;
    auto* psum = static_cast<int2*>(push);
    auto* ent = reinterpret_cast<PushEntry*>(psum + ((nblocks + 1) & ~(int64_t)1));
    int32_t* fctl = const_cast<int32_t*>(a.fctl);
    const int32_t need = a.lookahead ? 2 : 1;
    hipLaunchKernelGGL(k_push_count, dim3((unsigned)nblocks), dim3(256), 0, s, a.chg_r, nwords, inv.off,
                       psum, a.fctl, need);
    hipLaunchKernelGGL(k_push_scan, dim3(1), dim3(256), 0, s, psum, nblocks, fctl, need);
    hipLaunchKernelGGL(k_push_write, dim3((unsigned)nblocks), dim3(256), 0, s, a.chg_r, nwords, inv.off,
                       a.L, psum, ent, a.fctl, need);
    const unsigned grid = (unsigned)std::min<int64_t>(grid_for(m / 4 + 1, 256), 256 * 64);
    hipLaunchKernelGGL(k_gather_frontier, dim3(grid), dim3(256), 0, s, a.nbr, a.L, adj, m, inv.pos,
                       ent, a.fctl, need);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "gather launch: %s", hipGetErrorString(e));
}

int32_t push_edge_limit(int64_t m) {
    return (int32_t)std::min<int64_t>(m / kPushRatio, INT32_MAX);
}

size_t push_entry_bytes(int32_t n, int64_t m) {
    // per-workgroup sums (int2, padded to 16 bytes), then the entries
    const size_t nblocks = ((size_t)n / 32 + 1 + 255) / 256;
    return 8 * ((nblocks + 1) & ~(size_t)1) + sizeof(PushEntry) * ((size_t)n + (size_t)(m / kPushChunk) + 64);
}

int inverse_index_build(const int32_t* nbr, int32_t n, int64_t m, int32_t* cursor, InverseIndex& inv,
                        hipStream_t s) {
    // counts into inv.off[0..n], exclusive scan, then the fill (cursor = copy
    // of the offsets; n + 1 int32 of caller scratch)
    const int64_t cnt = (int64_t)n + 1;
    const int64_t nsums = (cnt + kScanBlock - 1) / kScanBlock;
    int32_t* sums = nullptr;
    GLP_HIP(hipMalloc(&sums, (size_t)nsums * 4));
    int rc = launch_fill_i32(inv.off, 0, cnt, s);
    const unsigned grid = (unsigned)std::min<int64_t>(grid_for(m, 256), 256 * 64);
    if (!rc && m > 0) hipLaunchKernelGGL(k_inv_count, dim3(grid), dim3(256), 0, s, nbr, m, inv.off);
    if (!rc) {
        hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nsums), dim3(256), 0, s, inv.off, cnt, sums);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, s, sums, nsums);
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nsums), dim3(256), 0, s, inv.off, cnt, sums);
        hipError_t e = hipMemcpyAsync(cursor, inv.off, (size_t)cnt * 4, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) rc = set_error(GLP_E_HIP, "inverse index: %s", hipGetErrorString(e));
    }
    if (!rc && m > 0) hipLaunchKernelGGL(k_inv_fill, dim3(grid), dim3(256), 0, s, nbr, m, cursor, inv.pos);
    if (!rc) {
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = set_error(GLP_E_HIP, "inverse index: %s", hipGetErrorString(e));
    }
    (void)hipFree(sums);
    return rc;
}

int launch_volume(const int32_t* L, int32_t* V, int32_t n, hipStream_t s) {
    if (n <= 0) return GLP_OK;
    hipLaunchKernelGGL(k_volume, dim3(grid_for(n, kVolChunk)), dim3(kVolThreads), 0, s, L, V, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "volume launch: %s", hipGetErrorString(e));
}

int launch_seed_labels(int32_t* p, const uint8_t* seeds, int64_t count, hipStream_t s) {
    if (count <= 0) return GLP_OK;
    const unsigned grid = (unsigned)std::min<int64_t>(grid_for(count, 256), 256 * 32);
    hipLaunchKernelGGL(k_seed_labels, dim3(grid), dim3(256), 0, s, p, seeds, count);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "seed label launch: %s", hipGetErrorString(e));
}

int launch_fill_i32(int32_t* p, int32_t value, int64_t count, hipStream_t s) {
    if (count <= 0) return GLP_OK;
    const unsigned grid = (unsigned)std::min<int64_t>(grid_for(count, 256), 256 * 32);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid), dim3(256), 0, s, p, value, count);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "fill launch: %s", hipGetErrorString(e));
}

int launch_label_volumes(const int32_t* L, const int32_t* V, int2* LV, int32_t n, hipStream_t s) {
    if (n <= 0) return GLP_OK;
    const unsigned grid = (unsigned)std::min<int64_t>(grid_for(n, 256), 256 * 64);
    hipLaunchKernelGGL(k_label_volumes, dim3(grid), dim3(256), 0, s, L, V, LV, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "label-volume launch: %s", hipGetErrorString(e));
}

int launch_iota(int32_t* p, int64_t count, hipStream_t s) {
    if (count <= 0) return GLP_OK;
    const unsigned grid = (unsigned)std::min<int64_t>(grid_for(count, 256), 256 * 32);
    hipLaunchKernelGGL(k_iota, dim3(grid), dim3(256), 0, s, p, count);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "iota launch: %s", hipGetErrorString(e));
}

// Random 4-byte reads of a table with uniform fmix32 indices (the label
// gather's access pattern without its CSR stream): four independent reads
// per lane per round, summed so the loads are kept.
__global__ __launch_bounds__(256) void k_calib_read(const int32_t* __restrict__ T, uint32_t mask,
                                                    int64_t reads, uint32_t seed,
                                                    int32_t* __restrict__ out) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int32_t acc = 0;
    for (int64_t j = tid * 4; j < reads; j += stride * 4) {
        const uint32_t b = (uint32_t)j ^ seed;
        acc += T[fmix32(b) & mask] + T[fmix32(b + 1u) & mask] + T[fmix32(b + 2u) & mask] +
               T[fmix32(b + 3u) & mask];
    }
    out[tid] = acc;
}

int launch_calib_read(const int32_t* T, int64_t entries, int64_t reads, uint32_t seed, int32_t* out,
                      int64_t out_entries, hipStream_t s) {
    if (entries <= 0 || (entries & (entries - 1)) || entries > (int64_t)1 << 31)
        return set_error(GLP_E_ARG, "table entries must be a power of two <= 2^31");
    const int64_t grid = std::min<int64_t>(256 * 32, out_entries / 256);
    if (grid < 1) return set_error(GLP_E_ARG, "calibration output too small");
    hipLaunchKernelGGL(k_calib_read, dim3((unsigned)grid), dim3(256), 0, s, T, (uint32_t)(entries - 1),
                       reads, seed, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "calibration launch: %s", hipGetErrorString(e));
}

int launch_end_step(unsigned long long* acc, unsigned long long* ring, int32_t* step_ctr,
                    int32_t ring_size, hipStream_t s, int32_t* fctl) {
    hipLaunchKernelGGL(k_end_step, dim3(1), dim3(kChangeSlots), 0, s, acc, ring, step_ctr, ring_size,
                       fctl);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GLP_OK : set_error(GLP_E_HIP, "end_step launch: %s", hipGetErrorString(e));
}

}  // namespace glp
