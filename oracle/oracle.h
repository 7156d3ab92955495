/*
 * oracle.h - CPU restatement of the reference's label-propagation semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libglp.so, the GPULP
 * CLI, bench.py's measured legs) may link, load or call this library; it is
 * the checker the GPU results are compared against, and bench.py's
 * cpu_baseline leg times orc_cpulp_run on the host.
 *
 * Each function names the reference file:line it restates.  Plain,
 * sequential, obviously-correct code: std::vector adjacency lists, std::map
 * counts, no hashing tricks.
 */
#ifndef GLP_ORACLE_H
#define GLP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* header/graph.h:69-147 + :227-258: size of the CSR a .bin file yields.
 * weighted=1 reads (u,v,w) triples.  Returns 0, or -1 on I/O or a negative id. */
int orc_load_bin_size(const char* path, int directed, int weighted, int32_t* n, int64_t* m);
/* Fills offsets (n+1), neighbors (m), weights (m, may be NULL). */
int orc_load_bin(const char* path, int directed, int weighted, int32_t* offsets,
                 int32_t* neighbors, uint32_t* weights);

/* RS classic step (prop/classicLprop.cuh:361-443 loop body; tie rule of the
 * packed max header/kernel.cuh:901-903): highest frequency, then highest
 * label; isolated vertices keep their label.  w may be NULL (all ones). */
void orc_classic_step(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                      const int32_t* prev, int32_t* next);

/* RS layered step (prop/MultiLprop.cuh:313-314, header/kernel.cuh:2579,
 * 2626): score (q+p)*C - p*V_prev(l) in int64 for gamma = p/q; vnext is the
 * histogram of next. */
void orc_layered_step(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                      int32_t p, int32_t q, const int32_t* prev, const int32_t* vprev,
                      int32_t* next, int32_t* vnext);

/* RS seeded step (prop/layeredLprop.cuh seed-constrained LP; seeds fixed as
 * in header/kernel.cuh:4490, unlabeled neighbours never counted as in
 * :4468,3272,1784): a seed keeps its own id; another vertex
 * takes the highest (frequency, label) over neighbours whose label is >= 0
 * and keeps its label when it has none.  L_0 = v on seeds, -1 elsewhere. */
void orc_seeded_step(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                     const uint8_t* seeds, const int32_t* prev, int32_t* next);

/* SLPA speaker stream R(seed,t,u). */
uint32_t orc_stream_u32(uint64_t seed, int32_t t, int32_t u);
/* SLPA (prop/SLPAprop.cuh:337-398 structure, speaker/listener rule defined
 * by this project): fills mem ((T+1) x n, row t = M_v[t]). */
void orc_slpa(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
              uint64_t seed, int32_t T, int32_t* mem);
/* Threshold r = r_num/r_den over rows 0..rows-1 of mem. */
void orc_slpa_post(int32_t n, int32_t rows, const int32_t* mem, int32_t r_num, int32_t r_den,
                   int32_t kmax, int32_t* dominant, int32_t* nsets, int32_t* sets);

/* method/countmin.cuh:97-142 row hashes (row 0..4). */
uint32_t orc_cm_hash(int32_t row, uint32_t key);
/* Count-min step: estimates from a per-vertex d x W sketch
 * (method/countmin.cuh:145-260); aliased=1 keeps the row aliasing of
 * :185-191 (d=3) and :197-221 (d=4). */
void orc_countmin_step(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                       int32_t rows, int32_t width, int32_t aliased, const int32_t* prev,
                       int32_t* next);
/* Sequential bounded tables: buffer=0 Misra-Gries (method/frequent.cuh),
 * buffer=1 decrement-and-drop (method/buffer.cuh:142-177). */
void orc_table_step(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                    int32_t cap, int32_t buffer, const int32_t* prev, int32_t* next);

/* tutorial/CPUlp.cuh:28-92 (sync_run_1) with OpenMP: niter steps from
 * L[v]=v.  mode 0 = verbatim (reads the live labels[], strict '>' first
 * max; timing only, racy under >1 thread); mode 1 = synchronous with the RS
 * tie rule (bit-identical to orc_classic_step).  Returns seconds. */
double orc_cpulp_run(int32_t n, const int32_t* off, const int32_t* nbr, int32_t niter,
                     int32_t threads, int32_t mode, int32_t* labels_out);

/* orc_classic_step for the vertices verts[0..nverts) only (next[k] is the
 * label of verts[k]), threads in parallel: the check of single steps of a
 * large graph on a sample of its vertices.  Neighbour labels are sorted per
 * vertex instead of counted in a map; same argmax and tie rule. */
void orc_classic_subset(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                        const int32_t* prev, const int32_t* verts, int64_t nverts, int32_t threads,
                        int32_t* next);

/* One step of tutorial/CPUlp.cuh:169-262 (layered_run) over verts[0..nverts),
 * for timing: live labels[] read and updated in place, a std::map count per
 * vertex, the running score (q+p)*count - p*Vr[label] with the strict '>'
 * first maximum, and Vw[winner] incremented; gamma = p/q.  Returns seconds. */
double orc_cpulp_layered_sample(int32_t n, const int32_t* off, const int32_t* nbr, int32_t* labels,
                                const int32_t* Vr, int32_t* Vw, const int32_t* verts, int64_t nverts,
                                int32_t threads, int32_t p, int32_t q);

/* One verbatim CPUlp step (tutorial/CPUlp.cuh:47-81) restricted to the
 * vertices in verts[0..nverts): the bounded CPU-baseline sample.  labels is
 * read and updated in place as the reference does; balanced = 0 keeps the
 * reference's static OpenMP schedule, 1 uses dynamic chunks; returns
 * seconds. */
double orc_cpulp_sample(int32_t n, const int32_t* off, const int32_t* nbr, int32_t* labels,
                        const int32_t* verts, int64_t nverts, int32_t threads, int32_t balanced);

#ifdef __cplusplus
}
#endif
#endif
