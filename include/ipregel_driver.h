/*
 * ipregel_driver.h - the C host driver behind ip_init / ip_run / ip_dump,
 * and the harness entry points that tests and the benchmark bind over ctypes.
 *
 * Reference interfaces replaced (file:line in capellil/iPregel):
 *   ipr_drv_init   ip_init         src/iPregel_postamble.h:195-242
 *                  ip_load_graph   src/iPregel_postamble.h:508-546
 *   ipr_drv_run    ip_run          e.g. src/combiner_spread_postamble.h:160-431
 *   ipr_drv_dump   ip_dump         src/iPregel_postamble.h:103-115
 * Error behaviour is the reference's: a message on stdout and exit(-1).
 */
#ifndef IPREGEL_DRIVER_H
#define IPREGEL_DRIVER_H

#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include "ipregel_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

void ipr_drv_init(const ipr_app_desc* app, const char* file_path, int number_of_threads,
                  const char* schedule, int chunk_size, bool directed, bool weighted);
int ipr_drv_run(void);
void ipr_drv_dump(FILE* f);
uint64_t ipr_drv_vertices_count(void);
uint64_t ipr_drv_edges_count(void);
uint64_t ipr_drv_superstep(void);

/* Host view of the vertices after ip_run, behind ip_get_vertex_by_id and
 * ip_get_vertex_by_location (src/iPregel_postamble.h:78-101), which iPregel.h
 * builds from these: */
uint64_t ipr_drv_runs(void);                 /* superstep loops run so far */
int ipr_drv_directed(void);                  /* ip_init's `directed` */
const uint64_t* ipr_drv_host_out_idx(void);  /* V+1 offsets (host copy) */
const uint32_t* ipr_drv_host_out_adj(void);  /* E out-neighbour ids (host copy) */
const uint32_t* ipr_drv_host_out_wgt(void);  /* E weights or NULL */
/* copy the device's values, active and has_message flags and messages (any
 * may be NULL); 0 on success */
int ipr_drv_vertex_state(void* val, uint8_t* act, uint8_t* hm, void* msg);
/* copy the device's in-lists (V+1 offsets, E ids); -ENOENT without them */
int ipr_drv_in_lists(uint64_t* idx, uint32_t* adj);

/* ---- harness (not part of the reference surface) ---------------------- */

typedef struct ipr_harness_cfg {
    /* graph from host arrays instead of <file_path>.{config,idx,adj,wgt} */
    uint64_t V, E;
    const uint64_t* idx;        /* V+1 offsets; NULL: read files */
    const uint32_t* adj;
    const uint32_t* wgt;        /* NULL unless weighted */
    int32_t skip_dump;          /* ip_dump writes nothing */
    int32_t repeat;             /* ip_run executes the superstep loop this many times */
    int32_t quiet;              /* no per-superstep log lines */
    int32_t device;             /* HIP device ordinal */
    void* stream;               /* hipStream_t to run on; NULL: driver's own */
    /* tuning; 0 keeps the default */
    int32_t pull_ordered;       /* 1 ordered, 2 reassociating */
    int32_t use_graph;          /* 1 replay supersteps as HIP graphs, 2 plain launches */
    int32_t dir_opt;            /* 1 on, 2 off */
    uint32_t dir_alpha;
    uint32_t bin_small_max;     /* arcs per chunk of a large broadcaster */
    uint32_t bin_medium_max;    /* out-degree above which a broadcaster is split in chunks */
    uint32_t grid_blocks;
    uint32_t pull_small_max, pull_medium_max;
    uint32_t max_chunk;
    int32_t defer_run;          /* ip_run returns at once; runs go through ipr_harness_run */
    int32_t kernel_events;      /* HIP events around every launch of the dominant kernel */
} ipr_harness_cfg;

typedef struct ipr_run_stats {
    uint64_t V, E;
    uint64_t supersteps;
    uint64_t work_edges;        /* W (GTEPS numerator) of the last run */
    uint64_t work_vertices;     /* vertices computed in the last run */
    uint64_t gathered;          /* supersteps delivered by gather */
    uint64_t runs;
    uint64_t launches;          /* superstep bodies enqueued (incl. no-ops after halting) */
    double last_ms;             /* host events around the last ip_run loop */
    double total_ms;            /* sum over runs */
    double min_ms;
    double device_ms;           /* device timestamps: first to last superstep */
    double load_s;              /* graph load + upload */
    double kernel_ms;           /* dominant kernel: summed event time of live launches */
    uint64_t kernel_launches;   /* dominant kernel: live launches timed */
    char device_name[128];
    uint64_t fused_steps;       /* supersteps of the last run run by the fused sparse kernel */
    uint64_t gather_tiles;      /* tiles of the receiver-blocked gather (0: the arc-balanced gather) */
} ipr_run_stats;

void ipr_harness_configure(const ipr_harness_cfg* cfg);
void ipr_harness_reset(void);
int ipr_harness_stats(ipr_run_stats* out);
/* copy the value array (V * sizeof(IP_VALUE_TYPE) bytes) to host */
int ipr_harness_values(void* out, uint64_t bytes);
/* per-superstep ActiveVertexCount, W and duration (ns) of the last run */
int ipr_harness_trace(uint64_t* active, uint64_t* edges, double* ns, uint64_t cap);
/* run the superstep loop `repeat` more times on the resident graph */
int ipr_harness_run(int repeat);
/* 1: HIP events around every launch of the dominant kernel (plain launches);
 * 0: off (chunks replayed as HIP graphs) */
int ipr_harness_kernel_events(int on);
/* with kernel events on: per superstep of the last run, the milliseconds of
 * the gather (or the fused pull kernel) and of the expansion that follows */
int ipr_harness_kernel_times(double* first, double* second, uint64_t cap);
/* the hipStream_t the driver launches on */
void* ipr_harness_stream(void);
/* release all device state of the current graph */
void ipr_harness_release(void);

#ifdef __cplusplus
}
#endif
#endif
