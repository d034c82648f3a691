/*
 * ipr_oracle.h - CPU oracle for the iPregel superstep hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product path (ipregel_amd/) never links, loads or calls it.
 *
 * It restates, single-threaded and in plain C, the BSP semantics of the four
 * engine variants of iPregel and of its three benchmark applications:
 *   variant PUSH         src/combiner_postamble.h:123-199
 *   variant PULL         src/combiner_single_broadcast_postamble.h:125-259
 *   variant SPREAD       src/combiner_spread_postamble.h:160-431
 *   variant SPREAD_PULL  src/combiner_spread_single_broadcast_postamble.h (ip_run)
 *   app PR               benchmarks/pagerank.c:36-71
 *   app CC               benchmarks/cc.c:31-63
 *   app SSSP             benchmarks/sssp.c:38-79
 * plus two applications of this project that have no reference counterpart:
 *   app SSSP_W           weighted SSSP (min combiner, message = dist + w(e))
 *   app PPR              personalised PageRank (restart at one seed vertex)
 * and closed forms used to pin it (min-ancestor labels, BFS, Dijkstra,
 * dense power iteration).
 */
#ifndef IPR_ORACLE_H
#define IPR_ORACLE_H

#include <stdint.h>
#include "../include/ipregel_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { IPR_O_PUSH = 0, IPR_O_PULL = 1, IPR_O_SPREAD = 2, IPR_O_SPREAD_PULL = 3 };
enum { IPR_O_PR = 0, IPR_O_CC = 1, IPR_O_SSSP = 2, IPR_O_SSSP_W = 3, IPR_O_PPR = 4 };

typedef struct ipr_oracle_params {
    uint32_t source;        /* SSSP / SSSP_W / PPR seed */
    uint32_t rounds;        /* PR / PPR broadcasting supersteps (PR: 10) */
    double   damping;       /* PR: 0.85 */
    uint64_t max_supersteps;/* capacity of active_counts (0: unlimited, no trace) */
    uint64_t stop_after;    /* end after this many supersteps, a bounded sample (0: run to the end) */
    double   stop_seconds;  /* end at the first superstep boundary after this much run time (0: none) */
} ipr_oracle_params;

typedef struct ipr_oracle_result {
    uint64_t supersteps;    /* supersteps executed */
    uint64_t work_edges;    /* W: push = sum of out-degrees of broadcasters;
                               pull = in-arcs scanned in supersteps with a broadcast */
    uint64_t work_vertices; /* sum over supersteps of vertices computed */
    uint64_t messages;      /* messages delivered (push) */
} ipr_oracle_result;

/* Run one application under one engine variant.
 * g: out-edge CSR (weights used by SSSP_W). directed: as passed to ip_init;
 * when 0, in-lists alias out-lists (src/iPregel_postamble.h:368-396).
 * values: V entries, double for PR/PPR, uint32 for the others.
 * active_counts: per-superstep ActiveVertexCount (may be NULL). */
int ipr_oracle_run(const ipr_csr* g, int directed, int variant, int app,
                   const ipr_oracle_params* p, void* values,
                   uint64_t* active_counts, ipr_oracle_result* res);

/* Write values in ip_dump's text format (benchmarks/*.c ip_serialise_vertex). */
int ipr_oracle_dump(int app, uint64_t V, const void* values, const char* path);

/* Personalised PageRank (project definition, no reference counterpart), FP64:
 *   x_0 = e_seed;  y = 0.85 * (P^T x_t + D_t e_seed) + 0.15 e_seed,
 *   (P^T x)(v) = sum over in-neighbours u of x(u) / outdeg(u),
 *   D_t = sum of x_t over vertices without out-arcs,  x_{t+1} = y / sum(y).
 * g is the out-edge CSR; in-sums run over the exact transpose in ascending
 * source order. `iters` applications; x gets V doubles. */
int ipr_oracle_ppr(const ipr_csr* g, uint32_t seed, uint32_t iters, double* x);
/* n seeds, bit-identical to n calls of ipr_oracle_ppr, OpenMP over vertices;
 * x gets n rows of V doubles. */
int ipr_oracle_ppr_many(const ipr_csr* g, const uint32_t* seeds, uint32_t n, uint32_t iters, double* x);

/* Closed forms. */
/* label[v] = min{u : u reaches v}; on a symmetric graph the component minimum. */
int ipr_oracle_min_ancestor(const ipr_csr* g, uint32_t* label);
/* hop distances over out-arcs, UINT32_MAX if unreachable. */
int ipr_oracle_bfs(const ipr_csr* g, uint32_t source, uint32_t* dist);
/* shortest paths with g->wgt (uint32), UINT32_MAX if unreachable. */
int ipr_oracle_dijkstra(const ipr_csr* g, uint32_t source, uint32_t* dist);

#ifdef __cplusplus
}
#endif
#endif
