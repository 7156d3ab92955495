/*
 * glp.h - C ABI of libglp.so, the MI355X (gfx950) label-propagation library.
 *
 * Every entry point takes plain pointers and sizes, returns GLP_OK (0) or a
 * negative GLP_E_* code, and leaves a thread-local message behind for
 * glp_last_error().  The caller owns every host buffer it passes; the library
 * copies inputs and writes outputs.  Nothing here allocates on behalf of the
 * caller except the opaque handles, which the matching *_free releases.
 *
 * Each declaration names the reference interface (Gibyeng/GLP) it replaces.
 */
#ifndef GLP_H
#define GLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ----------------------------------------------------- */
enum {
    GLP_OK = 0,
    GLP_E_ARG = -1,         /* bad argument (null pointer, out-of-range field) */
    GLP_E_IO = -2,          /* file missing / unreadable / unwritable          */
    GLP_E_FORMAT = -3,      /* malformed input (negative id, odd-length file)  */
    GLP_E_HIP = -4,         /* a HIP runtime call failed                       */
    GLP_E_NOMEM = -5,       /* host or device allocation failed                */
    GLP_E_RANGE = -6,       /* sizes beyond int32 CSR / 32-bit score limits    */
    GLP_E_UNSUPPORTED = -7, /* parameter combination not implemented           */
    GLP_E_BUDGET = -8,      /* job larger than the server's HBM budget         */
    GLP_E_STATE = -9,       /* call not valid in the handle's current state    */
    GLP_E_NODEVICE = -10    /* no HIP device visible                           */
};

/* ---- LP methods (reference main.cu:63-101 "case 1..4", plus counting kinds) */
enum {
    GLP_CLASSIC = 1,   /* prop/classicLprop.cuh                                    */
    GLP_LAYERED = 2,   /* prop/MultiLprop.cuh (score (1+gamma)C - gamma*V)          */
    GLP_SLPA = 3,      /* prop/SLPAprop.cuh (speaker/listener memory)              */
    GLP_USER = 4,      /* prop/UserdefinedLprop.cuh: the compiled-in user hooks    */
    GLP_COUNTMIN = 5,  /* method/countmin.cuh SharedCM sketch counting             */
    GLP_FREQUENT = 6,  /* method/frequent.cuh SharedFA heavy-hitter table          */
    GLP_BUFFER = 7,    /* method/buffer.cuh Buffer decrement-by-x table            */
    GLP_SEEDED = 8     /* prop/layeredLprop.cuh seed-constrained LP (seeds fixed)  */
};

/* ---- opaque handles --------------------------------------------------- */
typedef struct glp_graph glp_graph;       /* host CSR (+ device copy once used) */
typedef struct glp_session glp_session;   /* device-resident LP state of one job */
typedef struct glp_batch glp_batch;       /* many graphs stepped as one union    */
typedef struct glp_server glp_server;     /* continuous-batching request loop    */

/* ---- run parameters --------------------------------------------------- */
typedef struct glp_params {
    int32_t method;       /* GLP_CLASSIC ... GLP_SEEDED                          */
    int32_t weighted;     /* 1: frequency = sum of uint32 edge weights           */
    int32_t gamma_num;    /* layered: gamma = gamma_num / gamma_den,             */
    int32_t gamma_den;    /*   gamma_den a power of two (dyadic, exact scores)   */
    uint64_t seed;        /* SLPA speaker-stream seed                            */
    int32_t slpa_T;       /* SLPA memory length = number of listener steps       */
    int32_t slpa_r_num;   /* SLPA post-processing threshold r = num / den        */
    int32_t slpa_r_den;
    int32_t cm_rows;      /* count-min rows d (1..5)                             */
    int32_t cm_width;     /* count-min row width W (1..4096)                     */
    int32_t cm_aliased;   /* 1: rows 2 and 3 share storage as countmin.cuh:190   */
    int32_t table_cap;    /* frequent / buffer table capacity C (1..64)          */
    int32_t gather;       /* 0: gather fused into counting; 1: separate gather   */
    int32_t profile;      /* 1: per-tier HIP-event timing inside each step       */
    int32_t frontier;     /* classic only: 1 = recompute only vertices with a
                             changed neighbour once few labels change (the
                             reference's if_update frontier,
                             header/kernel.cuh:21-35); labels are identical.
                             2 = the same, and with gather = 1 the gather's
                             two-step lookahead (tutorial/Lprop.cuh:383-461:
                             two adj buffers, changes against two steps
                             back); needs a symmetric graph              */
    int32_t reserved[5];  /* must be zero                                        */
} glp_params;

/* Per-tier timing and shape record.  Replaces header/test_result.h:5-23
 * (Test_result): times are milliseconds from HIP events, not host clocks. */
enum { GLP_TIER_PACKED = 0, GLP_TIER_WAVE = 1, GLP_TIER_BLOCK = 2, GLP_TIER_HUB = 3,
       GLP_TIER_GATHER = 4, GLP_TIER_PICK = 5, GLP_TIER_AUX = 6, GLP_NTIERS = 8 };
typedef struct glp_timing {
    double upload_ms;                /* H2D of CSR (load_graph_to_gpu)              */
    double plan_ms;                  /* host tier planning (other_proprocess)       */
    double steps_ms;                 /* sum over executed steps (lp)                */
    double tier_ms[GLP_NTIERS];      /* per-tier sums when params.profile = 1       */
    int64_t tier_launches[GLP_NTIERS];
    int64_t tier_vertices[GLP_NTIERS];
    int64_t tier_edges[GLP_NTIERS];
    int32_t steps;                   /* steps executed so far                       */
    int32_t reserved;
} glp_timing;

/* ---- library ---------------------------------------------------------- */
const char* glp_last_error(void);                 /* thread-local message          */
const char* glp_version(void);
void glp_params_default(glp_params* p);           /* classic, unweighted, gamma 0  */
int glp_device_count(int32_t* count);

/* Process-wide tier thresholds and launch shapes applied to graphs whose
 * device copy is created afterwards (the autotuner writes these; the
 * GLP_TUNING environment variable names a key=value file read at first use).
 * Replaces the compile-time constants of prop/classicLprop.cuh:178-198,
 * 235, 302, 311-316, 381 (tier thresholds, block sizes, LDS table sizes).
 * Keys: pack_max_deg, half_max_deg (0 turns the several-vertices-per-
 * wavefront part of the wave tier off), half_lanes (16 or 32 lanes per
 * vertex there; 0 chooses per launch from the amount of work),
 * wave_max_deg, block_max_deg, hub_chunk, part_keys, part_limit,
 * block_threads (256, the default: a single graph's block tier picks 256,
 * 512 or 1024 threads per workgroup by LDS table size, up to 16, 32 or
 * 64 KB; 512 or 1024: that many for every table size).                     */
int glp_tuning_set(const char* key, int32_t value);
int glp_tuning_get(const char* key, int32_t* value);
/* Autotuner (no reference counterpart; replaces hand-picked constants):
 * times `steps` steps from L_0 on g under candidate tunings (coordinate
 * descent over the keys above), makes the fastest the process default and,
 * when out_path is not NULL, writes it as key=value lines.  best_ms receives
 * the winner's time for the timed steps.                                    */
int glp_autotune(glp_graph* g, const glp_params* p, int32_t steps, const char* out_path,
                 double* best_ms);

/* ---- graphs ----------------------------------------------------------- */
/* Replaces header/graph.h:69-147 (CSRGraph ctor) with :227-258
 * (load_edgelist_bin): headerless little-endian int32 (u,v) pairs, trailing
 * partial pair dropped, self-loops dropped, symmetrized unless directed,
 * duplicates and file order kept.  weighted=1 reads (u,v,w) int32 triples
 * (no reference counterpart: README.md:54-63 names weighted input but no
 * binary form exists).  Negative ids are GLP_E_FORMAT.                       */
int glp_graph_load_bin(const char* path, int32_t directed, int32_t weighted,
                       glp_graph** out);
/* Builds a graph from caller CSR arrays (copied).  weights may be NULL.
 * Replaces the CSRGraph members offsets/neighbors (header/graph.h:195-196).  */
int glp_graph_from_csr(int32_t n, int64_t m, const int32_t* offsets,
                       const int32_t* neighbors, const uint32_t* weights,
                       glp_graph** out);
/* Shape of a graph: n (CSRGraph::n), m (CSRGraph::m), max degree, weighted. */
int glp_graph_info(const glp_graph* g, int32_t* n, int64_t* m,
                   int32_t* max_degree, int32_t* weighted);
/* No counterpart in the reference (its frontier code assumes undirected
 * input): *symmetric = 1 when every edge u->v has an edge v->u (as sets),
 * else 0.  The frontier modes need it; loaded undirected graphs and the
 * generators' graphs are symmetric by construction, others are checked
 * once (O(m log deg) on the host).                                           */
int glp_graph_symmetric(glp_graph* g, int32_t* symmetric);
/* Copies the host CSR out (offsets n+1, neighbors m, weights m or NULL).    */
int glp_graph_csr(const glp_graph* g, int32_t* offsets, int32_t* neighbors,
                  uint32_t* weights);
/* Replaces header/graph.h:149 (~CSRGraph); also frees the device copy.      */
void glp_graph_free(glp_graph* g);

/* ---- one-shot run ----------------------------------------------------- */
/* Replaces classicLprop/MultiLprop/SLPAprop/UserdefinedLprop::run(niter)
 * (prop/classicLprop.cuh:454-462, prop/MultiLprop.cuh:383-391,
 * prop/SLPAprop.cuh:407, main.cu:63-101).  Executes exactly `steps`
 * propagation steps from L_0[v] = v (the reference CLI runs niter-1 steps,
 * prop/classicLprop.cuh:361; the CLI applies that convention itself).
 * labels_out: n int32 (final labels; for SLPA the dominant label).
 * step_dump: steps*n int32 (labels after each step) or NULL.
 * timing: may be NULL.                                                       */
int glp_run(glp_graph* g, const glp_params* p, int32_t steps,
            int32_t* labels_out, int32_t* step_dump, glp_timing* timing);
/* glp_run for seeded LP (p->method == GLP_SEEDED): seeds as in
 * glp_session_set_seeds (n bytes, nonzero = seed).                          */
int glp_run_seeded(glp_graph* g, const glp_params* p, const uint8_t* seeds, int32_t steps,
                   int32_t* labels_out, int32_t* step_dump, glp_timing* timing);

/* ---- sessions: device-resident state for repeated steps --------------- */
/* Uploads the CSR (once per graph), plans tiers and allocates label state;
 * labels start at L_0[v] = v (header/kernel.cuh:444-451 initialize_labels). */
int glp_session_create(glp_graph* g, const glp_params* p, glp_session** out);
/* Resets labels to L_0 and the step counter to 0.                           */
int glp_session_reset(glp_session* s);
/* Enqueues `steps` steps on the session's stream (asynchronous).  One step is
 * the body of the reference loop prop/classicLprop.cuh:361-443: PickLabel,
 * LoadNeighbor + count per tier, LabelScore, UpdateVertex.                    */
int glp_session_step(glp_session* s, int32_t steps);
/* Captures one step in a HIP graph and replays it `steps` times.            */
int glp_session_step_graph(glp_session* s, int32_t steps);
/* Layered: changes gamma for the following steps; labels and volumes carry
 * over (the reference's per-round delta schedule, prop/MultiLprop.cuh:308-314,
 * with gamma = 1/delta).                                                     */
int glp_session_set_gamma(glp_session* s, int32_t gamma_num, int32_t gamma_den);
/* Seeded LP (method GLP_SEEDED): replaces layeredLprop::assign_seeds_for_device
 * and initialize_labels_by_seeds (prop/layeredLprop.cuh:161-180,222-224;
 * header/kernel.cuh:453-466).  seeds: n bytes, nonzero marks a seed.  Copies
 * them to the device and resets the session: L_0[v] = v for a seed, -1
 * otherwise.  A seeded session must have its seeds set before its first
 * step.                                                                      */
int glp_session_set_seeds(glp_session* s, const uint8_t* seeds, int32_t n);
/* Waits for the session's stream.                                           */
int glp_session_sync(glp_session* s);
/* Copies the current labels (n int32) to host; blocks.                      */
int glp_session_labels(glp_session* s, int32_t* out);
/* Number of labels changed by each executed step (the reference's
 * "node update" counter, prop/classicLprop.cuh:438-440); fills up to cap.   */
int glp_session_changes(glp_session* s, int64_t* out, int32_t cap,
                        int32_t* nsteps);
/* Timing record (Test_result replacement).                                  */
int glp_session_timing(glp_session* s, glp_timing* out);
/* Layered: current volume histogram V(l) (n int32).                         */
int glp_session_volume(glp_session* s, int32_t* out);
/* Checkpoint / resume: step counter plus label (and volume / SLPA memory)
 * state.  state_bytes reports the size needed; state may be NULL to query.  */
int glp_session_save(glp_session* s, void* state, size_t* state_bytes);
int glp_session_load(glp_session* s, const void* state, size_t state_bytes);
/* SLPA: memory rows M[0..t] (t+1 rows of n int32, row-major by step).       */
int glp_session_slpa_memory(glp_session* s, int32_t* out);
/* SLPA post-processing (threshold r): per vertex the labels whose frequency
 * in M_v[0..T] is >= r*(T+1), ascending, padded with -1 to kmax, the count
 * of such labels, and the dominant label (highest count, then highest id).  */
int glp_session_slpa_post(glp_session* s, int32_t kmax, int32_t* dominant,
                          int32_t* nsets, int32_t* sets);
/* Waits for the session's queued work and frees its device state (the
 * reference's free_gmem, prop/classicLprop.cuh:107-116).                    */
int glp_session_free(glp_session* s);

/* ---- allocation-free run ----------------------------------------------- */
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  glp_run for classic and layered LP (fused gather, no
 * frontier, no timing record) with the graph's CSR, its tier plan and all
 * LP state in a static device array of GLP_NOALLOC_BYTES per GPU, held from
 * the library's load until the process ends; a graph whose CSR, plan and
 * state do not fit fails with GLP_E_RANGE.  Copies use the null stream.
 * Calls on one GPU share its array: the library serialises them.           */
#define GLP_NOALLOC_BYTES (128u << 20)
int glp_run_noalloc(const glp_graph* g, const glp_params* p, int32_t steps,
                    int32_t* labels_out, int32_t* step_dump);

/* ---- batched: many graphs stepped as one disjoint union --------------- */
/* No single reference counterpart: the reference runs one graph per process
 * (main.cu:18-123).  All graphs use the same params; labels are identical to
 * running each graph alone.  Every launch of a step covers all graphs; with
 * the tuning key fused_small=1 the packed and half tiers are one persistent
 * launch.                                                                    */
int glp_batch_create(int32_t ngraphs, glp_graph* const* graphs,
                     const glp_params* p, glp_batch** out);
/* mode 0: the union's launches, enqueued per step; mode 1: the same steps
 * replayed from a captured HIP graph; mode 2: plain per-graph session
 * launches (the comparison form).                                          */
int glp_batch_run(glp_batch* b, int32_t steps, int32_t mode, double* ms);
int glp_batch_labels(glp_batch* b, int32_t graph, int32_t* out);
int glp_batch_reset(glp_batch* b);
/* Autotunes tier thresholds and LDS sizes on the packed batch (glp_autotune)
 * and resets the batch to L_0.                                               */
int glp_batch_autotune(glp_batch* b, int32_t steps, double* best_ms);
void glp_batch_free(glp_batch* b);

/* ---- serving: continuous batching under an HBM budget ------------------ */
/* No reference counterpart (the reference runs one graph per process,
 * main.cu:18-123).  Classic and layered LP only; results equal a solo run. */
typedef struct glp_job_stats {
    double submit_ms;    /* host clock at submit, relative to server start      */
    double admit_ms;     /* when its memory was reserved against the budget     */
    double done_ms;      /* when its labels were copied out                     */
    int64_t bytes;       /* device footprint admitted against the budget        */
    int32_t steps;       /* steps enqueued so far                               */
    int32_t state;       /* 0 queued, 4 admitted, 1 running, 2 done,
                            3 cancelled, <0 error                               */
    double start_ms;     /* when its first step could be enqueued               */
} glp_job_stats;
int glp_server_create(uint64_t hbm_budget_bytes, int32_t max_slots,
                      glp_server** out);
/* Resident graphs: on the device once, shared by every job that names them,
 * charged against the budget until removed.  add_graph uploads a caller's
 * graph; add_rmat builds glp_gen_rmat(scale, edge_factor, a, b, c, seed,
 * degree_sort 0, wmax)'s graph directly in HBM (same CSR).                 */
int glp_server_add_graph(glp_server* sv, const glp_graph* g, int32_t* graph_id);
int glp_server_add_rmat(glp_server* sv, int32_t scale, int32_t edge_factor,
                        double a, double b, double c, uint64_t seed,
                        int32_t wmax, int32_t* graph_id);
int glp_server_graph_info(glp_server* sv, int32_t graph_id, int32_t* n,
                          int64_t* m, int32_t* max_degree, int64_t* bytes);
/* Copies a resident graph's CSR out (offsets n+1, neighbours m, weights m;
 * any may be NULL).                                                         */
int glp_server_graph_csr(glp_server* sv, int32_t graph_id, int32_t* offsets,
                         int32_t* neighbors, uint32_t* weights);
/* The graph takes no new jobs and leaves once its running jobs end.         */
int glp_server_remove_graph(glp_server* sv, int32_t graph_id);
/* Queues a job on a resident graph.                                         */
int glp_server_submit_resident(glp_server* sv, int32_t graph_id,
                               const glp_params* p, int32_t steps,
                               int64_t* job_id);
/* Queues a job (graph is copied into the server; the caller may free it;
 * it is uploaded when the job is admitted and freed when it ends).          */
int glp_server_submit(glp_server* sv, const glp_graph* g, const glp_params* p,
                      int32_t steps, int64_t* job_id);
int glp_server_poll(glp_server* sv, int64_t job_id, glp_job_stats* stats);
/* Blocks until the job finishes, then copies its labels (n int32).          */
int glp_server_wait(glp_server* sv, int64_t job_id, int32_t* labels_out,
                    glp_job_stats* stats);
/* Forgets an ended job (its labels and stats); later calls on it fail.      */
int glp_server_release(glp_server* sv, int64_t job_id);
int glp_server_cancel(glp_server* sv, int64_t job_id);
/* Drains all queued jobs and stops the request loop.                        */
int glp_server_free(glp_server* sv);

#ifdef __cplusplus
}
#endif
#endif /* GLP_H */
