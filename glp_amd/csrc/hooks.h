// The four user-facing LP hooks as HIP device functors.
//
// The reference specifies them in README.md:72-103 and stubs them as empty
// kernels in header/kernel.cuh:6071-6085:
//   PickLabel(vid)                    - decide vid's label for this step
//   LoadNeighbor(vid, sid) -> (l, f)  - label and frequency an edge contributes
//   LabelScore(vid, l, f) -> score    - score of candidate label l
//   UpdateVertex(vid, l, score)       - commit the winner
// Here a hook set is a struct whose static __device__ members the tier
// kernels are instantiated over.  Frequencies are exact uint32 sums, and a
// score is returned already packed into a uint64 whose unsigned order is the
// lexicographic order (score, label): the high word is the order-preserving
// image of a signed 32-bit score, the low word is label+1.  The tier kernels
// take the maximum packed value over all candidate labels, which makes the
// result independent of the order edges are counted in, and ties go to the
// highest label id (the packed maximum of header/kernel.cuh:901-903).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

#include "device_prims.h"

namespace glp {

// Pointers and scalars one step needs; passed by value to every kernel.
struct StepArgs {
    const int32_t* off;      // CSR offsets, n+1
    const int32_t* nbr;      // CSR neighbours, m
    const uint32_t* w;       // edge weights, m (weighted runs only)
    const int32_t* L;        // labels neighbours show this step (PickLabel output)
    const int32_t* adj;      // pre-gathered neighbour labels, m (separate gather)
    const int32_t* keep;     // label a vertex keeps when it has no neighbours
    int32_t* Lw;             // labels written by this step
    const int32_t* Vr;       // layered: volume of each label before the step
    int32_t* Vw;             // layered: volume after the step (zeroed first)
    unsigned long long* changes;  // kChangeSlots counters of changed labels (summed per step)
    int32_t score_mul;       // layered: q+p for gamma = p/q
    int32_t score_vmul;      // layered: p
    int32_t n;
    int32_t t;               // step number (1-based)
    const int32_t* mem;      // SLPA memory, (T+1) x n
    uint64_t seed;           // SLPA stream seed
    // frontier (classic only; null when off): act_r[v] = some neighbour of v
    // changed in the previous step; act_w collects that for the next step;
    // fctl[0] = 1: this step recomputes every vertex, fctl[1] = 1: this step
    // marks act_w (both set on the device by k_end_step)
    const uint8_t* act_r;
    uint8_t* act_w;
    const int32_t* fctl;
    // frontier gather (separate gather with the frontier on; null otherwise):
    // bit v of chg_w is set when v's label changes in this step; chg_r holds
    // the previous step's bits, whose sources' labels the gather pushes to
    // their in-edges when they have few
    const uint32_t* chg_r;
    uint32_t* chg_w;
    // two-step lookahead of the frontier gather: adj alternates between two
    // buffers, each refreshed every other step, so the bit records a change
    // against the label of two steps back (which Lw still holds)
    int32_t lookahead;
    // seeded LP: seeds[v] != 0 marks a seed (its label stays v)
    const uint8_t* seeds;
    // layered LP with gamma > 0: LV[u] = {L[u], Vr[L[u]]}, a neighbour's label
    // and its label's volume in one 8-byte record (built before the tiers)
    const int2* LV;
};

__host__ __device__ __forceinline__ uint32_t order_i32(int32_t s) {
    return (uint32_t)s ^ 0x80000000u;
}
__host__ __device__ __forceinline__ uint64_t pack_score(int32_t score, uint32_t key) {
    return ((uint64_t)order_i32(score) << 32) | key;
}
__host__ __device__ __forceinline__ int32_t packed_label(uint64_t p) {
    return (int32_t)((uint32_t)p) - 1;
}

// ---- classic LP: score = frequency (prop/classicLprop.cuh,
// header/kernel.cuh:1724-1729 packed (count<<32)+key) --------------------
struct ClassicHooks {
    static constexpr bool kPick = false;
    // an isolated vertex keeps its label and has no side effects, so groups of
    // isolated vertices need no work once both label buffers hold the label
    static constexpr bool kSkipIsolated = true;
    // a vertex whose neighbours all kept their labels keeps its own: exact
    static constexpr bool kFrontier = true;
    // the score is the frequency sum, so a label holding more than half of a
    // vertex's counted frequency wins outright
    static constexpr bool kScoreIsCount = true;
    __device__ static __forceinline__ int32_t pick_label(const StepArgs& a, int32_t vid) {
        return a.keep[vid];
    }
    template <bool W, bool G>
    __device__ static __forceinline__ uint32_t load_neighbor(const StepArgs& a, int32_t vid,
                                                             int32_t j, uint32_t* freq) {
        (void)vid;
        // a weighted launch may fold in unweighted jobs (w == nullptr)
        *freq = W ? (a.w ? dev::ld_stream(&a.w[j]) : 1u) : 1u;
        return (uint32_t)(G ? dev::ld_stream(&a.adj[j]) : a.L[dev::ld_stream(&a.nbr[j])]) + 1u;
    }
    __device__ static __forceinline__ uint64_t label_score(const StepArgs& a, int32_t vid,
                                                           uint32_t key, uint32_t freq) {
        (void)a; (void)vid;
        // freq < 2^31 (checked on the host), so order_i32 keeps the order
        return pack_score((int32_t)freq, key);
    }
    __device__ static __forceinline__ void update_vertex(const StepArgs& a, int32_t vid,
                                                         int32_t label, uint64_t score) {
        (void)score;
        a.Lw[vid] = label;
    }
};

// ---- layered LP: score = (q+p)*C - p*V(l) for gamma = p/q
// (prop/MultiLprop.cuh:313-314, header/kernel.cuh:2579 delta*count - gcounter)
struct LayeredHooks {
    static constexpr bool kPick = false;
    // an isolated vertex keeps its label in both buffers; the volume pass
    // counts it from the written buffer, so it needs no work
    static constexpr bool kSkipIsolated = true;
    static constexpr bool kFrontier = false;      // V changes every label's score
    __device__ static __forceinline__ int32_t pick_label(const StepArgs& a, int32_t vid) {
        return a.keep[vid];
    }
    template <bool W, bool G>
    __device__ static __forceinline__ uint32_t load_neighbor(const StepArgs& a, int32_t vid,
                                                             int32_t j, uint32_t* freq) {
        return ClassicHooks::load_neighbor<W, G>(a, vid, j, freq);
    }
    __device__ static __forceinline__ uint64_t label_score(const StepArgs& a, int32_t vid,
                                                           uint32_t key, uint32_t freq) {
        (void)vid;
        return label_score_vol(a, key, freq, label_volume(a, key));
    }
    // The score reads V(l) from HBM, so table scans load it for several
    // labels before scoring them (the loads are then in flight together).
    static constexpr bool kScoreReadsVolume = true;
    __device__ static __forceinline__ int32_t label_volume(const StepArgs& a, uint32_t key) {
        // with score_vmul = 0 (gamma = 0) this is classic LP and Vr may be null
        return a.score_vmul ? a.Vr[key - 1] : 0;
    }
    __device__ static __forceinline__ uint64_t label_score_vol(const StepArgs& a, uint32_t key,
                                                               uint32_t freq, int32_t vol) {
        // both terms are range-checked on the host to fit int32
        return pack_score(a.score_mul * (int32_t)freq - a.score_vmul * vol, key);
    }
    // the write-side volume V_t (header/kernel.cuh:2626 gcounter_temp) is the
    // histogram of the labels written; k_volume computes it after the tiers
    __device__ static __forceinline__ void update_vertex(const StepArgs& a, int32_t vid,
                                                         int32_t label, uint64_t score) {
        (void)score;
        a.Lw[vid] = label;
    }
};

// ---- layered LP with gamma > 0, scored in the table slot ---------------
// The same scores as LayeredHooks, computed without a volume read per
// candidate: each edge reads its neighbour's label together with that
// label's volume (StepArgs::LV), and a table slot accumulates the score
// itself, biased by 2^31 so that it stays a non-negative 32-bit sum: the
// edge that creates the slot adds (q+p)f - pV(l) + 2^31, every later edge
// of the label (q+p)f.  The sums only grow from a value >= 2^31 - p n > 0
// to the final score + 2^31 < 2^32 (both bounds are checked on the host),
// so a slot never carries into its key, and the slot word is the packed
// score's high word as it stands.  Tables whose slots hold frequencies (the
// hub tier's chunk tables, whose sums cross chunks) use label_score.
struct LayeredVolHooks {
    static constexpr bool kPick = false;
    static constexpr bool kSkipIsolated = true;
    static constexpr bool kFrontier = false;
    static constexpr bool kScoreReadsVolume = true;  // for frequency-slot tables
    static constexpr bool kSlotScore = true;
    template <bool W, bool G>
    __device__ static __forceinline__ uint32_t load_neighbor_vol(const StepArgs& a, int32_t vid, int32_t j,
                                                                 uint32_t* freq, int32_t* vol) {
        (void)vid;
        *freq = W ? (a.w ? dev::ld_stream(&a.w[j]) : 1u) : 1u;
        const int2 lv = a.LV[dev::ld_stream(&a.nbr[j])];
        *vol = lv.y;
        return (uint32_t)lv.x + 1u;
    }
    template <bool W, bool G>
    __device__ static __forceinline__ uint32_t load_neighbor(const StepArgs& a, int32_t vid, int32_t j,
                                                             uint32_t* freq) {
        int32_t vol;
        return load_neighbor_vol<W, G>(a, vid, j, freq, &vol);
    }
    __device__ static __forceinline__ uint32_t slot_more(const StepArgs& a, uint32_t f) {
        return (uint32_t)a.score_mul * f;
    }
    __device__ static __forceinline__ uint32_t slot_first(const StepArgs& a, uint32_t f, int32_t vol) {
        return slot_more(a, f) - (uint32_t)a.score_vmul * (uint32_t)vol + 0x80000000u;
    }
    __device__ static __forceinline__ uint64_t slot_score(uint32_t key, uint32_t slot) {
        return ((uint64_t)slot << 32) | key;
    }
    __device__ static __forceinline__ int32_t label_volume(const StepArgs& a, uint32_t key) {
        return LayeredHooks::label_volume(a, key);
    }
    __device__ static __forceinline__ uint64_t label_score_vol(const StepArgs& a, uint32_t key, uint32_t freq,
                                                               int32_t vol) {
        return LayeredHooks::label_score_vol(a, key, freq, vol);
    }
    __device__ static __forceinline__ uint64_t label_score(const StepArgs& a, int32_t vid, uint32_t key,
                                                           uint32_t freq) {
        return LayeredHooks::label_score(a, vid, key, freq);
    }
    __device__ static __forceinline__ void update_vertex(const StepArgs& a, int32_t vid, int32_t label,
                                                         uint64_t score) {
        (void)score;
        a.Lw[vid] = label;
    }
};

// ---- SLPA listener: classic counting over the speakers' labels; the
// speaker choice is the pick kernel (slpa_speak).  prop/SLPAprop.cuh:337-398.
struct SlpaHooks {
    static constexpr bool kPick = true;
    static constexpr bool kSkipIsolated = false;  // every memory row is written
    static constexpr bool kFrontier = false;      // speakers change every step
    static constexpr bool kScoreIsCount = true;
    // Speaker rule: u shows M_u[R(seed,t,u) mod t] (memory rows 0..t-1).
    __device__ static __forceinline__ int32_t pick_label(const StepArgs& a, int32_t vid) {
        uint64_t x = (a.seed << 40) ^ ((uint64_t)(uint32_t)a.t << 32) ^ (uint64_t)(uint32_t)vid;
        x += 0x9e3779b97f4a7c15ULL;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
        x ^= x >> 31;
        uint32_t r = (uint32_t)x;
        uint32_t row = r % (uint32_t)a.t;
        return a.mem[(size_t)row * (size_t)a.n + (size_t)vid];
    }
    template <bool W, bool G>
    __device__ static __forceinline__ uint32_t load_neighbor(const StepArgs& a, int32_t vid,
                                                             int32_t j, uint32_t* freq) {
        return ClassicHooks::load_neighbor<W, G>(a, vid, j, freq);
    }
    __device__ static __forceinline__ uint64_t label_score(const StepArgs& a, int32_t vid,
                                                           uint32_t key, uint32_t freq) {
        return ClassicHooks::label_score(a, vid, key, freq);
    }
    __device__ static __forceinline__ void update_vertex(const StepArgs& a, int32_t vid,
                                                         int32_t label, uint64_t score) {
        (void)score;
        a.Lw[vid] = label;
    }
};

// ---- seeded (semi-supervised) LP (prop/layeredLprop.cuh:161-180,222-224;
// header/kernel.cuh:453-466 and the *_seed tier kernels): seeds start with
// and keep their own id, every other vertex starts unlabeled (-1).  A
// non-seed takes the classic argmax over its labeled neighbours only (an
// unlabeled neighbour shows key 0, which no tier counts) and keeps its
// label when no neighbour is labeled.
struct SeededHooks {
    static constexpr bool kPick = false;
    static constexpr bool kSkipIsolated = true;  // isolated vertices keep their label
    static constexpr bool kFrontier = false;
    static constexpr bool kSeeded = true;        // seeds are never recomputed
    static constexpr bool kScoreIsCount = true;
    template <bool W, bool G>
    __device__ static __forceinline__ uint32_t load_neighbor(const StepArgs& a, int32_t vid,
                                                             int32_t j, uint32_t* freq) {
        return ClassicHooks::load_neighbor<W, G>(a, vid, j, freq);  // label -1 gives key 0
    }
    __device__ static __forceinline__ uint64_t label_score(const StepArgs& a, int32_t vid,
                                                           uint32_t key, uint32_t freq) {
        return ClassicHooks::label_score(a, vid, key, freq);
    }
    __device__ static __forceinline__ void update_vertex(const StepArgs& a, int32_t vid,
                                                         int32_t label, uint64_t score) {
        (void)score;
        a.Lw[vid] = label;
    }
};

// Hook sets without kScoreIsCount = true get no majority shortcut.
template <class H, class = void>
struct count_score_of : std::false_type {};
template <class H>
struct count_score_of<H, std::void_t<decltype(H::kScoreIsCount)>>
    : std::bool_constant<H::kScoreIsCount> {};

// Hook sets without kScoreReadsVolume = true score without a per-label load.
template <class H, class = void>
struct volume_of : std::false_type {};
template <class H>
struct volume_of<H, std::void_t<decltype(H::kScoreReadsVolume)>>
    : std::bool_constant<H::kScoreReadsVolume> {};

// Hook sets with kSlotScore = true accumulate scores in table slots
// (LayeredVolHooks); the others accumulate frequencies.
template <class H, class = void>
struct slot_score_of : std::false_type {};
template <class H>
struct slot_score_of<H, std::void_t<decltype(H::kSlotScore)>> : std::bool_constant<H::kSlotScore> {};

// An edge's key, frequency and (slot-scoring hook sets) its label's volume.
template <class H, bool W, bool G>
__device__ __forceinline__ uint32_t load_edge(const StepArgs& a, int32_t v, int32_t j, uint32_t* fr,
                                              int32_t* vol) {
    if constexpr (slot_score_of<H>::value) {
        return H::template load_neighbor_vol<W, G>(a, v, j, fr, vol);
    } else {
        *vol = 0;
        return H::template load_neighbor<W, G>(a, v, j, fr);
    }
}
// What an edge of frequency f adds to its label's slot when it creates the
// slot (first) or finds it (more), and the packed score of a slot; S picks
// the slot-score arithmetic (a hook set with kSlotScore in a table whose
// slots hold scores) or plain frequency sums.
template <class H, bool S>
__device__ __forceinline__ uint32_t slot_first(const StepArgs& a, uint32_t f, int32_t vol) {
    if constexpr (S) return H::slot_first(a, f, vol);
    else { (void)a; (void)vol; return f; }
}
template <class H, bool S>
__device__ __forceinline__ uint32_t slot_more(const StepArgs& a, uint32_t f) {
    if constexpr (S) return H::slot_more(a, f);
    else { (void)a; return f; }
}
template <class H, bool S>
__device__ __forceinline__ uint64_t slot_value_score(const StepArgs& a, int32_t v, uint32_t key, uint32_t slot) {
    if constexpr (S) { (void)a; (void)v; return H::slot_score(key, slot); }
    else return H::label_score(a, v, key, slot);
}

// Hook sets without a kSeeded member have no fixed vertices.
template <class H, class = void>
struct seeded_of : std::false_type {};
template <class H>
struct seeded_of<H, std::void_t<decltype(H::kSeeded)>> : std::bool_constant<H::kSeeded> {};

}  // namespace glp

// The user-overridable hook set (method 4).
#include "user_hooks.h"
