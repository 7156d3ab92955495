/*
 * glp_tools.h - C ABI of the data-preparation side of libglp.so: the
 * edge-list .bin writers (the reference's tools/ programs) and the seeded
 * synthetic graph generators used for the benchmark workloads.
 */
#ifndef GLP_TOOLS_H
#define GLP_TOOLS_H

#include <stdint.h>
#include "glp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces tools/edgelist2bin.cpp:13-73 (convert_edgelist_txt): text lines
 * "u<ws>v" ('#' lines skipped, self-loops dropped) -> ids compacted by
 * ascending original id -> (x,y) int32 pairs with x<y, grouped by ascending
 * x, y in insertion order.  Writes to out_path (the reference writes
 * "<basename>.bin" into the working directory; the CLI tool keeps that).    */
int glp_tool_edgelist2bin(const char* txt_path, const char* out_path);

/* Replaces tools/normalize.cpp:262-348 (normalize_sort_save): reads .bin
 * pairs (pairs whose first int is 35 skipped, abs() of ids, ids registered
 * before self-loops are dropped), orders vertices by non-increasing degree
 * with the same unstable std::sort and comparator, remaps ids and writes
 * (x,y) x<y pairs.  The reference writes "sorted_<basename>.bin".            */
int glp_tool_normalize(const char* bin_path, const char* out_path);

/* Replaces tools/direct_edgelist2bin.cpp:113-205 (buildfrombin, the path
 * its main runs): reads .bin pairs (u,v), drops self-loops, keeps an edge
 * only if u or v is below id_cutoff (the reference hard-codes 5013066 at
 * :137; id_cutoff <= 0 keeps every edge), keeps u -> v only (directed,
 * duplicates kept), orders the vertices seen by non-increasing out-degree
 * with the same unstable std::sort and comparator on the ascending-id list,
 * and writes each vertex's (x,y) pairs with neighbours ascending by original
 * id.  The reference writes "sorted_direct_<basename>.bin".                  */
int glp_tool_direct_edgelist2bin(const char* bin_path, const char* out_path,
                                 int32_t id_cutoff);

/* No counterpart in the reference: measures the device's ceiling for the
 * label gather's access pattern.  Issues `reads` random 4-byte reads of a
 * table of table_entries int32 (a power of two <= 2^31; uniform fmix32
 * indices, four independent reads per lane) in one launch, repeats the
 * launch `repeats` times after one untimed launch, and stores the fastest
 * launch's time in *ms_best.                                                 */
int glp_calib_random_read(int64_t table_entries, int64_t reads, int32_t repeats, double* ms_best);

/* Writes a graph back as a .bin: for every vertex u in id order, each
 * neighbour v > u in CSR order as an (u,v) pair (the inverse of
 * glp_graph_load_bin for undirected graphs without duplicates); weighted
 * graphs are written as (u,v,w) triples.                                     */
int glp_graph_save_bin(const glp_graph* g, const char* path);

/* Relabels vertices by non-increasing degree, ties by ascending old id (a
 * stable order, unlike normalize), keeping isolated vertices (at the end).
 * perm_out (n int32, may be NULL) receives new id -> old id.                 */
int glp_graph_degree_sort(const glp_graph* g, glp_graph** out, int32_t* perm_out);

/* ---- seeded generators (no reference counterpart: the reference ships no
 * generator and its default input datasets/test.bin is absent) ------------ */

/* R-MAT: edge_factor * 2^scale directed draws with quadrant probabilities
 * (a,b,c,1-a-b-c), self-loops dropped, symmetrized, deduplicated, int32 CSR
 * with neighbours ascending.  degree_sort=1 relabels as glp_graph_degree_sort.
 * wmax>0 attaches uint32 weights uniform in [1,wmax] (symmetric).  Threads:
 * 0 = all host cores.  The output is a pure function of the arguments.       */
int glp_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                 double c, uint64_t seed, int32_t degree_sort, int32_t wmax,
                 int32_t threads, glp_graph** out);

/* Planted partition (stochastic block model): n vertices in k near-equal
 * communities, every intra-community pair an edge with probability
 * deg_in/(size-1), every inter-community pair with probability
 * deg_out/(n-size), sampled by geometric skipping; no self-loops or
 * duplicates; neighbours ascending.  community_out (n int32) may be NULL.    */
int glp_gen_planted(int32_t n, int32_t k, double deg_in, double deg_out,
                    uint64_t seed, int32_t threads, int32_t* community_out,
                    glp_graph** out);

/* LFR-style benchmark graph: degrees power law (exponent tau1) on
 * [kmin,kmax], community sizes power law (tau2) on [cmin,cmax], each vertex
 * with a fraction mu of its stubs wired outside its community (configuration
 * model inside and across communities, self-loops and duplicates dropped).  */
int glp_gen_lfr(int32_t n, double tau1, int32_t kmin, int32_t kmax,
                double tau2, int32_t cmin, int32_t cmax, double mu,
                uint64_t seed, int32_t* community_out, glp_graph** out);

/* Parameters of graph i of the batched planted-partition mix: n log-uniform
 * in [nmin,nmax], ceil(n/1000) communities, average degree uniform in
 * [dmin,dmax], inter-community fraction uniform in [fmin,fmax].             */
int glp_gen_mix_params(uint64_t seed, int32_t i, int32_t nmin, int32_t nmax,
                       double dmin, double dmax, double fmin, double fmax,
                       int32_t* n, int32_t* k, double* deg_in, double* deg_out,
                       uint64_t* graph_seed);

/* The SLPA speaker stream: a counter-based uint32 R(seed, t, u) (SplitMix64
 * finaliser of (seed<<40) ^ (t<<32) ^ u, low 32 bits).  Shared by the device
 * kernels and any CPU check.                                                 */
uint32_t glp_stream_u32(uint64_t seed, int32_t t, int32_t u);
/* Fills out[0..count) with R(seed, t, u0 + i).                               */
int glp_stream_fill(uint64_t seed, int32_t t, int32_t u0, int32_t count,
                    uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif /* GLP_TOOLS_H */
