// User-defined LP hooks, selected by method 4 (GLP_USER).
//
// This is the one file a user edits to run a customised LP, the counterpart
// of overwriting the API stubs in the reference's header/kernel.cuh:6071-6085
// as README.md:70-103 describes.  Rebuild the library after editing.
//
// Contract (all members static __device__):
//   kPick          true if pick_label must run for every vertex before the
//                  neighbour pass (its result becomes the label neighbours see)
//   pick_label(a, vid) -> int32 label shown to neighbours this step
//   load_neighbor<W,G>(a, vid, j, &freq) -> label+1 of edge j (j indexes the
//                  CSR neighbour array; a.nbr[j] is the neighbour id); sets the
//                  exact uint32 frequency the edge contributes
//   label_score(a, vid, key, freq) -> packed score (use pack_score(s, key);
//                  s must fit int32); the highest packed value wins
//   update_vertex(a, vid, label, packed) commits the winner (must write
//                  a.Lw[vid]); vertices with no neighbours receive a.keep[vid]
//
// The shipped example is the README's classic LP (README.md:83-103):
// PickLabel copies Lnext into L, LoadNeighbor returns (L[sid], 1), LabelScore
// returns the frequency and UpdateVertex writes the label.
#pragma once

namespace glp {

struct UserHooks {
    static constexpr bool kPick = true;
    static constexpr bool kSkipIsolated = false;  // update_vertex may have side effects
    static constexpr bool kFrontier = false;      // scores may depend on more than neighbours
    __device__ static __forceinline__ int32_t pick_label(const StepArgs& a, int32_t vid) {
        return a.keep[vid];  // L[vid] = Lnext[vid]
    }
    template <bool W, bool G>
    __device__ static __forceinline__ uint32_t load_neighbor(const StepArgs& a, int32_t vid,
                                                             int32_t j, uint32_t* freq) {
        (void)vid;
        *freq = W ? a.w[j] : 1u;
        return (uint32_t)(G ? a.adj[j] : a.L[a.nbr[j]]) + 1u;
    }
    __device__ static __forceinline__ uint64_t label_score(const StepArgs& a, int32_t vid,
                                                           uint32_t key, uint32_t freq) {
        (void)a; (void)vid;
        return pack_score((int32_t)freq, key);
    }
    __device__ static __forceinline__ void update_vertex(const StepArgs& a, int32_t vid,
                                                         int32_t label, uint64_t score) {
        (void)score;
        a.Lw[vid] = label;
    }
};

}  // namespace glp
