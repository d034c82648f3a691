/*
 * ipregel_hip.h - the thin C-ABI between the C host driver and HIP
 * (libipregel_hip.so). Plain pointers, sizes and opaque handles; no HIP or
 * torch types. Every function returns 0 on success or a negative hipError_t;
 * ipr_hip_last_error() gives the message of the last failure.
 *
 * The reference has no device boundary (it is one OpenMP process): these
 * entry points replace what its host code does with memory, files and the
 * OpenMP runtime on the superstep path:
 *   ipr_hip_malloc/free/memcpy     malloc/realloc of vertex and engine state
 *                                  (src/iPregel_postamble.h:117-146,520-533)
 *   ipr_hip_build_csc              in-neighbour mirroring
 *                                  (src/iPregel_postamble.h:429-487)
 *   ipr_hip_begin / ipr_hip_finalize_*
 *                                  the superstep loop control of ip_run: counter
 *                                  reset, active count, halting test, superstep
 *                                  increment (e.g. src/combiner_spread_postamble.h
 *                                  :212, :221-225, :278-285, :347-403)
 *   ipr_hip_build_pull_bins        (new) in-degree bins of the pull gather
 *   ipr_hip_build_gather           (new) arc-balanced index of the push
 *                                  applications' dense-superstep gather
 *   ipr_hip_build_gather_tiles     (new) receiver-blocked, source-sorted index
 *                                  of the same gather
 *   ipr_hip_graph_*                (new) replay of a chunk of supersteps
 */
#ifndef IPREGEL_HIP_H
#define IPREGEL_HIP_H

#include <stddef.h>
#include <stdint.h>
#include "ipregel_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* ipr_hip_last_error(void);

int ipr_hip_init(int device, char* name, size_t name_len, uint64_t* hbm_bytes);
int ipr_hip_wallclock_hz(uint64_t* hz);
/* compute units of the current device */
int ipr_hip_cu_count(uint32_t* cus);

int ipr_hip_malloc(void** p, size_t bytes);
int ipr_hip_free(void* p);
int ipr_hip_host_alloc(void** p, size_t bytes);      /* pinned */
int ipr_hip_host_free(void* p);
int ipr_hip_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int ipr_hip_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int ipr_hip_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
/* fill with a byte value by kernel (safe to record in a HIP graph) */
int ipr_hip_fill(void* dst, int byte, size_t bytes, void* stream);
/* store one 64-bit value in device memory, in stream order (a kernel, so it
 * may also be recorded into a HIP graph) */
int ipr_hip_set_u64(void* dst, uint64_t value, void* stream);

int ipr_hip_stream_create(void** stream);
int ipr_hip_stream_destroy(void* stream);
int ipr_hip_stream_sync(void* stream);
int ipr_hip_event_create(void** ev);
int ipr_hip_event_destroy(void* ev);
int ipr_hip_event_record(void* ev, void* stream);
int ipr_hip_event_elapsed_ms(void* start, void* stop, float* ms);

/* capture everything enqueued on `stream` between begin and end into an
 * executable graph */
int ipr_hip_graph_begin(void* stream);
int ipr_hip_graph_end(void* stream, void** exec);
int ipr_hip_graph_launch(void* exec, void* stream);
int ipr_hip_graph_destroy(void* exec);

/* CSC from a device CSR: in-rows ascending by source id, weights carried */
int ipr_hip_build_csc(uint64_t V, uint64_t E, const uint64_t* out_idx,
                      const uint32_t* out_adj, const uint32_t* out_wgt,
                      uint64_t* in_idx, uint32_t* in_adj, uint32_t* in_wgt,
                      void* stream);

/* dst[i] = src[i] widened to 64 bits (the neighbour-id arrays that _64
 * applications read through ip_vertex_t, src/iPregel_postamble.h:361-381) */
int ipr_hip_widen_ids(const uint32_t* src, uint64_t* dst, uint64_t n, void* stream);

/* static in-degree bins for the pull gather; sizes returned in dev->pbin_size */
int ipr_hip_build_pull_bins(ipr_dev* dev, uint32_t small_max, uint32_t medium_max,
                            void* stream);

/* record order: bpos[v] = rank of v by descending out-degree class
 * [2^k, 2^(k+1)), ids ascending within a class; no out-arcs last. With bvert
 * (may be NULL), also its inverse: bvert[bpos[v]] = v. */
int ipr_hip_build_record_order(uint64_t V, const uint64_t* out_idx, uint32_t* bpos, uint32_t* bvert,
                               void* stream);

/* index of the arc-balanced gather over a device CSC (in-lists):
 * g_arc[E] = the sources (their record positions bpos[u] when bpos is not
 * NULL) with IPR_GATHER_LAST on the last arc of every non-empty list,
 * g_dst[*n_dst] = the vertices with non-empty lists (ascending),
 * g_own[ceil(E/K)] = position in g_dst of the list holding arc c*K.
 * Needs V < 2^31 and K a multiple of 4. */
int ipr_hip_build_gather(uint64_t V, uint64_t E, const uint64_t* t_idx, const uint32_t* t_adj,
                         const uint32_t* bpos, uint32_t K, uint32_t* g_arc, uint32_t* g_own,
                         uint32_t* g_dst, uint64_t* n_dst, void* stream);

/* the same index split into IPR_GATHER_CLASSES classes by the 128-byte line of
 * the source's 4-byte record (line mod classes); per class c: arcs at
 * g_arc[cbeg[c], cend[c]) with cbeg[c] a multiple of K, chunk owners at
 * g_own[cbeg[c]/K ...] indexing g_dst + cdst[c]; with t_wgt and g_wgt, the
 * arcs' weights in g_wgt at the same positions as g_arc. g_arc and g_wgt need
 * E + classes*K entries, g_own (E + classes*K)/K, g_dst min(E, classes*V). */
int ipr_hip_build_gather_classes(uint64_t V, uint64_t E, const uint64_t* t_idx, const uint32_t* t_adj,
                                 const uint32_t* t_wgt, const uint32_t* bpos, uint32_t K, uint32_t* g_arc,
                                 uint32_t* g_wgt, uint32_t* g_own, uint32_t* g_dst, uint64_t* cbeg, uint64_t* cend,
                                 uint64_t* cdst, void* stream);

/* receiver-blocked gather index over the transpose (t_idx/t_adj/t_wgt).
 * Receivers split into blocks of consecutive ids with at most max_recv
 * (< 65535) receivers and at most max_arcs in-arcs; a receiver with more
 * in-arcs forms a block alone, split into tiles of whole 64-slot groups whose
 * deliveries are atomic (shared). Within a block the arcs are sorted by the
 * source's record position (bpos, or the id when bpos is NULL) and laid out
 * in groups of 64 slots as IPR_TQ_ESC describes; weights (when t_wgt is
 * given) at 1, 2 or 4 bytes, the narrowest that holds the largest. All of
 * *out is allocated here (device memory; free each array with ipr_hip_free). */
int ipr_hip_build_gather_tiles(uint64_t V, uint64_t E, const uint64_t* t_idx, const uint32_t* t_adj,
                               const uint32_t* t_wgt, const uint32_t* bpos, uint32_t max_recv, uint64_t max_arcs,
                               ipr_gather_tiles* out, void* stream);

/* byte-wise equality of two device buffers (multiple of 4 bytes) */
int ipr_hip_equal(const void* a, const void* b, uint64_t bytes, int* equal, void* stream);

/* ipr_hip_equal, allocating nothing: no counterpart in the reference. For
 * callers that cannot allocate device memory (code in a tight serving loop,
 * harnesses that forbid hipMalloc). Its scratch is a 256-byte device array per
 * GPU, held until the process ends (one 32-bit difference flag is used, so
 * the input size is not limited by it); bytes must be a multiple of 4, else
 * it returns an error. Calls that overlap on one GPU share the flag and race.
 * The same result as ipr_hip_equal for every input both accept. */
int ipr_hip_equal_noalloc(const void* a, const void* b, uint64_t bytes, int* equal, void* stream);

/* push variants: the next frontier (claim bitmap of superstep s+1) compacted
 * into front[(s+1)&1], ascending within runs; the bitmap is cleared */
/* finalize_variant >= 0: the compaction's last workgroup also ends the
 * superstep of that variant (ipregel_control.h), with no finalize launch.
 * relist != 0: the current superstep's frontier instead (claim bitmap of
 * superstep s, left by a fused sparse kernel that stopped) into front[s&1],
 * counted from fsize[s&1], which the caller zeroes first; no finalize */
int ipr_hip_compact_frontier(const ipr_dev* dev, int finalize_variant, int relist, void* stream);

/* superstep control (kernels on `stream`, graph-capturable) */
int ipr_hip_begin(const ipr_dev* dev, void* stream);
int ipr_hip_finalize(const ipr_dev* dev, int variant, void* stream);

#ifdef __cplusplus
}
#endif
#endif
