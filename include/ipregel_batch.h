/*
 * ipregel_batch.h - batched and served multi-query execution on resident
 * graphs (libipregel_batch.so). No reference counterpart: iPregel runs one
 * application per process (the main() of each benchmarks/ program); this layer keeps graphs in
 * HBM and runs many queries of three kinds over them at once:
 *
 *   SSSP   distances from a source over uint32 weights (1 when the graph has
 *          none); exact, identical to the single-query engine.
 *   PPR    personalised PageRank from a seed: x0 = e_seed,
 *          y = 0.85 (P^T x + D e_seed) + 0.15 e_seed, x' = y / sum(y),
 *          D = mass on vertices without out-arcs; `iters` applications.
 *          Values stored in FP32 with a Kahan-compensated pull sum; dangling
 *          mass, normalisation and the L1 residual in FP64.
 *   CC     connected-component labels of a symmetric graph (the smallest
 *          vertex id of each component), by union-find; in serving, CC queries
 *          of one graph that wait together share one computation.
 *
 * Batches hold up to 64 queries per graph and kind; a query's lane computes
 * only from its own column and every reduction runs in a fixed order, so a
 * result does not depend on which other queries share the batch.
 * All functions return 0 on success or a negative error code;
 * ipr_batch_last_error() describes the last failure.
 */
#ifndef IPREGEL_BATCH_H
#define IPREGEL_BATCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPR_BATCH_SLOTS 64

enum { IPR_Q_SSSP = 0, IPR_Q_PPR = 1, IPR_Q_CC = 2 };

typedef struct ipr_bgraph ipr_bgraph;   /* a graph resident in HBM */

const char* ipr_batch_last_error(void);

/* Upload a CSR (V+1 offsets, E arcs, optional weights) and build its
 * transpose on the device. */
int ipr_batch_graph_create(uint64_t V, uint64_t E, const uint64_t* idx, const uint32_t* adj,
                           const uint32_t* wgt, int device, ipr_bgraph** out);
int ipr_batch_graph_destroy(ipr_bgraph* g);
/* HBM held by the graph and by each kind's 64-slot batch state */
uint64_t ipr_batch_graph_bytes(const ipr_bgraph* g);
uint64_t ipr_batch_slot_bytes(const ipr_bgraph* g, int kind);

/* Launch shapes and scheduling of the push kernels of one graph (used by
 * states created afterwards). None of them changes a result. */
typedef struct ipr_batch_tuning {
    uint32_t large;   /* out-degree above which a frontier vertex's arcs are split into chunks (1024) */
    uint32_t chunk;   /* arcs per chunk of such a vertex (1024) */
    uint32_t grid;    /* workgroups of the grid-stride kernels (2048) */
    uint32_t delta;   /* SSSP near/far step: a vertex relaxes while its distance is below the
                         query's threshold, which then moves to the smallest pending distance
                         + delta; 0xFFFFFFFF = every superstep relaxes every improved vertex
                         (default: 4 x mean arc weight, or 0xFFFFFFFF without weights); used by
                         64-slot states */
    uint32_t delta1;  /* the same for one-query states (the plain loop) */
    uint32_t cc_mode; /* CC algorithm: 0 label propagation supersteps, 1 union-find (default) */
    uint32_t sssp_states; /* ipr_serve: 64-slot SSSP states this graph's queries are spread
                             over (each arrival joins the least loaded one), run concurrently on
                             their own streams; 1..8 (default 1) */
} ipr_batch_tuning;
int ipr_batch_graph_get_tuning(const ipr_bgraph* g, ipr_batch_tuning* t);
int ipr_batch_graph_set_tuning(ipr_bgraph* g, const ipr_batch_tuning* t);

typedef struct ipr_batch_stats {
    uint64_t supersteps;        /* batch supersteps (SSSP) or iterations (PPR) run */
    uint64_t relaxations;       /* SSSP: (query, arc) relaxations; PPR: query-arc products */
    double ms;                  /* device time of the batch (HIP events) */
} ipr_batch_stats;

/* Run nq (<= 64) queries of one kind to completion as one batch.
 * SSSP: out is nq x V uint32 (query-major). PPR: out is nq x V float,
 * residual (may be NULL) nq doubles (L1 change of the last iteration). */
int ipr_batch_sssp(ipr_bgraph* g, const uint32_t* sources, int nq, uint32_t* out,
                   ipr_batch_stats* st);
int ipr_batch_ppr(ipr_bgraph* g, const uint32_t* seeds, int nq, int iters, float* out,
                  double* residual, ipr_batch_stats* st);
/* the same computations for one query with scalar (not batched) layouts:
 * the plain per-query kernels the batches are compared against */
int ipr_single_sssp(ipr_bgraph* g, uint32_t source, uint32_t* out, ipr_batch_stats* st);
int ipr_single_ppr(ipr_bgraph* g, uint32_t seed, int iters, float* out, double* residual,
                   ipr_batch_stats* st);
/* CC labels of the graph (V uint32) */
int ipr_batch_cc(ipr_bgraph* g, uint32_t* out, ipr_batch_stats* st);

/* ---- serving ------------------------------------------------------------ */

typedef struct ipr_query {
    double t_arrival;           /* seconds after the start of the trace */
    int32_t kind;               /* IPR_Q_* */
    int32_t graph;              /* index into the graph array */
    uint32_t param;             /* SSSP source, PPR seed, CC: vertex queried */
    uint32_t pad;
} ipr_query;

typedef struct ipr_serve_opts {
    int32_t sequential;         /* 1: plain per-query loop (no batching), in arrival order */
    int32_t ppr_iters;          /* PPR applications per query (20) */
    uint64_t hbm_budget;        /* bytes the batch states may hold (0: no limit) */
    int32_t use_graph;          /* 1: replay each tick as HIP graphs (default); 2: plain launches */
    int32_t pad;
} ipr_serve_opts;

typedef struct ipr_serve_stats {
    double wall_s;              /* first arrival to last completion */
    uint64_t ticks;             /* batched supersteps */
    uint64_t admitted_waits;    /* admissions delayed by a full batch or the HBM budget */
    uint64_t peak_slots;        /* most queries in flight at once */
    uint64_t hbm_peak;          /* most HBM held by batch states */
} ipr_serve_stats;

/* Serve a trace of n queries over ng resident graphs, arrival times honoured
 * against the wall clock. For each query: admission and completion times (s)
 * and a 64-bit digest of its result (FNV-1a over the result bytes; PPR results
 * are FP32 values, SSSP uint32 distances, CC the label of `param`).
 * results[i], when non-NULL, receives the full result (V elements; CC: V labels). */
int ipr_serve(ipr_bgraph** graphs, int ng, const ipr_query* q, int n, const ipr_serve_opts* opt,
              double* t_admit, double* t_done, uint64_t* digest, void** results,
              ipr_serve_stats* st);

#ifdef __cplusplus
}
#endif
#endif
