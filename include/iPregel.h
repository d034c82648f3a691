/*
 * iPregel.h - MI355X (gfx950) engine behind iPregel's programming surface.
 *
 * An application written for iPregel (src/iPregel.h, src/iPregel_preamble.h)
 * includes this header instead and is compiled by hipcc as HIP. It keeps:
 *   - the type macros IP_VERTEX_ID_TYPE (uint32_t), IP_NEIGHBOUR_COUNT_TYPE,
 *     IP_MESSAGE_TYPE, IP_VALUE_TYPE (defined by the app before inclusion),
 *   - the variant switches IP_USE_SPREAD / IP_USE_SINGLE_BROADCAST
 *     (src/iPregel_preamble.h:338-350),
 *   - the hooks ip_compute / ip_combine (now __device__) and
 *     ip_serialise_vertex (host) (src/iPregel_preamble.h:281,294,302),
 *   - ip_send_message, ip_broadcast, ip_get_next_message, ip_has_message,
 *     ip_vote_to_halt (device side, src/iPregel_preamble.h:141-185),
 *   - ip_get_superstep, ip_is_first_superstep, ip_get_vertices_count,
 *     ip_get_edges_count (host and device, src/iPregel_preamble.h:85-115),
 *   - ip_init, ip_run, ip_dump (host, src/iPregel_preamble.h:317,328,336).
 * Messages are passed by value, as in the reference code (not its README).
 *
 * The edits an application needs are device qualifiers: __device__ on
 * ip_compute and ip_combine, and __device__ on globals that ip_compute reads
 * after main() sets them (e.g. benchmarks/pagerank.c:32-33), followed by one
 * IP_PUBLISH(name) after the host assignment: host code writes the host copy
 * of a __device__ variable and IP_PUBLISH copies it to the device.
 * (__managed__ globals are not used: managed memory is HMM-backed, and GPU
 * access to it relies on page-fault support that these GPUs run without.)
 *
 * Optional compile-line declarations:
 *   -DIP_COMBINER_MIN / _MAX / _SUM  the combiner is min / max / sum, so
 *        mailboxes use hardware atomics and pull gathers may reassociate;
 *        checked against ip_combine at ip_init. Without one, ip_combine runs
 *        inside a compare-and-swap loop (32-bit messages).
 *   -DIP_EDGE_WEIGHT_TYPE=uint32_t and IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS enable
 *        ip_broadcast_weighted (project extension: each out-neighbour u
 *        receives message + w(v,u)); weights come from <graph>.wgt.
 *
 * The host driver (ipregel_amd/csrc/driver/ip_driver.c) is C; it loads the
 * graph, owns device memory through libipregel_hip.so and runs the superstep
 * loop; the kernels that call the hooks are instantiated here, in the
 * application's translation unit, and handed to the driver as a table.
 */
#ifndef IPREGEL_AMD_IPREGEL_H
#define IPREGEL_AMD_IPREGEL_H

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdbool.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

#include "ipregel_engine.h"
#include "ipregel_driver.h"

#ifndef IP_VERTEX_ID_TYPE
#error "IP_VERTEX_ID_TYPE must be defined (e.g. -DIP_VERTEX_ID_TYPE=uint32_t)"
#endif

/* implied needs, src/iPregel_preamble.h:32-38 */
#if defined(IP_NEEDS_OUT_NEIGHBOUR_IDS) && !defined(IP_NEEDS_OUT_NEIGHBOUR_COUNT)
#define IP_NEEDS_OUT_NEIGHBOUR_COUNT
#endif
#if defined(IP_NEEDS_IN_NEIGHBOUR_IDS) && !defined(IP_NEEDS_IN_NEIGHBOUR_COUNT)
#define IP_NEEDS_IN_NEIGHBOUR_COUNT
#endif

/* numeric, so that the preprocessor can test them; the values are those of
 * the IPR_VARIANT_* enumerators of ipregel_engine.h */
#if defined(IP_USE_SPREAD) && defined(IP_USE_SINGLE_BROADCAST)
#define IPR_VARIANT 3
#elif defined(IP_USE_SPREAD)
#define IPR_VARIANT 2
#elif defined(IP_USE_SINGLE_BROADCAST)
#define IPR_VARIANT 1
#else
#define IPR_VARIANT 0
#endif
#define IPR_IS_PULL (IPR_VARIANT == 1 || IPR_VARIANT == 3)

#if defined(IP_COMBINER_MIN)
#define IPR_COMBINER IPR_COMBINER_MIN
#elif defined(IP_COMBINER_MAX)
#define IPR_COMBINER IPR_COMBINER_MAX
#elif defined(IP_COMBINER_SUM)
#define IPR_COMBINER IPR_COMBINER_SUM
#else
#define IPR_COMBINER IPR_COMBINER_GENERIC
#endif

#ifndef VERSION
#define VERSION "ipregel-amd"
#endif
#ifndef IP_MACHINE
#define IP_MACHINE "MI355X"
#endif
#ifndef COMMITS
#define COMMITS "none"
#endif
#ifndef IP_APPLICATION
#define IP_APPLICATION "app"
#endif
#ifndef COMPILATION_FLAGS
#define COMPILATION_FLAGS ""
#endif
#define IPR_STRINGIFY(x) IPR_STRINGIFY_LITERAL(x)
#define IPR_STRINGIFY_LITERAL(x) #x

/* Host globals of the reference preamble (src/iPregel_preamble.h:63-76),
 * mirrored from the driver after ip_init and ip_run. */
size_t ip_superstep = 0;
size_t ip_edges_count = 0;
size_t ip_vertices_count = 0;
size_t ip_active_vertices = 0;
int ip_thread_count = 0;

/* Vertex seen by ip_compute: a register-resident proxy filled from the
 * structure-of-arrays state and written back after the hook returns.
 * Field names follow src/combiner_*_preamble.h. */
struct ip_vertex_t {
#ifdef IP_NEEDS_OUT_NEIGHBOUR_IDS
    IP_VERTEX_ID_TYPE* out_neighbours;
#endif
#ifdef IP_NEEDS_IN_NEIGHBOUR_IDS
    IP_VERTEX_ID_TYPE* in_neighbours;
#endif
#ifdef IP_NEEDS_OUT_NEIGHBOUR_COUNT
    IP_NEIGHBOUR_COUNT_TYPE out_neighbour_count;
#endif
#ifdef IP_NEEDS_IN_NEIGHBOUR_COUNT
    IP_NEIGHBOUR_COUNT_TYPE in_neighbour_count;
#endif
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
    IP_EDGE_WEIGHT_TYPE* out_neighbour_weights;
#endif
    bool active;
    bool has_message;
    IP_VERTEX_ID_TYPE id;
    IP_MESSAGE_TYPE message;
    IP_VALUE_TYPE value;
    /* engine-private: broadcast recorded during ip_compute */
    bool ipr_bcast;
    bool ipr_bcast_weighted;
    IP_MESSAGE_TYPE ipr_bcast_msg;
};

extern "C" {
/* user hooks */
__device__ void ip_compute(struct ip_vertex_t* v);
__device__ void ip_combine(IP_MESSAGE_TYPE* message_a, IP_MESSAGE_TYPE message_b);
void ip_serialise_vertex(FILE* f, struct ip_vertex_t* v);

/* run API */
void ip_init(const char* file_path, int number_of_threads, const char* schedule,
             int chunk_size, bool directed, bool weighted);
int ip_run(void);
void ip_dump(FILE* f);
}

#include "ipregel_kernels.h"

/* Copy the host copy of a __device__ application global to the device, e.g.
 *   __device__ double ratio;  ...  ratio = 0.15 / n; IP_PUBLISH(ratio);
 * Exits like the reference's safe functions on failure. */
template <typename T> static inline void ipr_publish(const T& var, const char* name)
{
    if (hipMemcpyToSymbol(HIP_SYMBOL(var), &var, sizeof(T), 0, hipMemcpyHostToDevice) != hipSuccess) {
        printf("Failed to publish the application global %s to the device.\n", name);
        exit(-1);
    }
}
#define IP_PUBLISH(var) ipr_publish((var), #var)

extern "C" {

__host__ __device__ inline size_t ip_get_superstep(void)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (size_t)ipr_sm_superstep;
#else
    return ip_superstep;
#endif
}

__host__ __device__ inline bool ip_is_first_superstep(void)
{
    return ip_get_superstep() == 0;
}

__host__ __device__ inline size_t ip_get_vertices_count(void)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (size_t)ipr_d_dev.V;
#else
    return ip_vertices_count;
#endif
}

__host__ __device__ inline size_t ip_get_edges_count(void)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (size_t)ipr_d_dev.E;
#else
    return ip_edges_count;
#endif
}


__device__ inline bool ip_has_message(struct ip_vertex_t* v)
{
    return v->has_message;
}

/* src/combiner_postamble.h:124-134 (identical in all variants) */
__device__ inline bool ip_get_next_message(struct ip_vertex_t* v, IP_MESSAGE_TYPE* message_value)
{
    if (v->has_message) {
        *message_value = v->message;
        v->has_message = false;
        return true;
    }
    return false;
}

/* Push variants record the broadcast; the expansion kernel (or a gather in a
 * dense superstep) delivers it to every out-neighbour. Two broadcasts in one
 * ip_compute reach the same neighbours, so they are combined here.
 * Pull variants write the vertex's broadcast slot (last one wins),
 * src/combiner_single_broadcast_postamble.h:155-159. */
__device__ inline void ip_broadcast(struct ip_vertex_t* v, IP_MESSAGE_TYPE message)
{
#if IPR_IS_PULL
    v->ipr_bcast = true;
    v->ipr_bcast_msg = message;
#else
    if (v->ipr_bcast) {
        if (v->ipr_bcast_weighted) ipr_raise_error(IPR_ERR_MIXED_BROADCAST);
        ip_combine(&v->ipr_bcast_msg, message);
    } else {
        v->ipr_bcast = true;
        v->ipr_bcast_msg = message;
    }
#endif
}

#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
/* Project extension (the reference never loads weights): every out-neighbour
 * u receives message + w(v,u). Push variants only, once per superstep. */
__device__ inline void ip_broadcast_weighted(struct ip_vertex_t* v, IP_MESSAGE_TYPE message)
{
#if IPR_IS_PULL
    ipr_raise_error(IPR_ERR_SEND_IN_PULL);
#else
    if (v->ipr_bcast) ipr_raise_error(IPR_ERR_MIXED_BROADCAST);
    v->ipr_bcast = true;
    v->ipr_bcast_weighted = true;
    v->ipr_bcast_msg = message;
#endif
}
#endif

/* Push variants: immediate delivery into the next superstep's mailbox
 * (src/combiner_postamble.h:149-174, src/combiner_spread_postamble.h:202-227).
 * Pull variants: the reference prints and exits
 * (src/combiner_single_broadcast_postamble.h:146-153); here the kernel flags
 * the error and the driver prints the same text and exits after the launch. */
__device__ inline void ip_send_message(IP_VERTEX_ID_TYPE id, IP_MESSAGE_TYPE message)
{
#if IPR_IS_PULL
    (void)id; (void)message;
    ipr_raise_error(IPR_ERR_SEND_IN_PULL);
#else
    ipr_send_immediate(id, message);
#endif
}

__device__ inline void ip_vote_to_halt(struct ip_vertex_t* v)
{
    /* a no-op in the spread variants (src/combiner_spread_postamble.h:562-565);
     * the spread kernels ignore the flag */
    v->active = false;
}

/* ------------------------------------------------------------------------ */
/* host glue                                                                 */
/* ------------------------------------------------------------------------ */

static void ipr_app_serialise(void* file, const void* val, const uint64_t* out_idx,
                              uint64_t first, uint64_t count)
{
    const IP_VALUE_TYPE* values = (const IP_VALUE_TYPE*)val;
    for (uint64_t i = first; i < first + count; i++) {
        struct ip_vertex_t v;
        memset(&v, 0, sizeof(v));
        v.id = (IP_VERTEX_ID_TYPE)i;
        v.value = values[i];
#ifdef IP_NEEDS_OUT_NEIGHBOUR_COUNT
        v.out_neighbour_count = (IP_NEIGHBOUR_COUNT_TYPE)(out_idx[i + 1] - out_idx[i]);
#else
        (void)out_idx;
#endif
        ip_serialise_vertex((FILE*)file, &v);
    }
}

void ip_init(const char* file_path, int number_of_threads, const char* schedule,
             int chunk_size, bool directed, bool weighted)
{
    ipr_app_desc d;
    memset(&d, 0, sizeof(d));
    d.variant = IPR_VARIANT;
    d.combiner = IPR_COMBINER;
    d.id_size = sizeof(IP_VERTEX_ID_TYPE);
    d.val_size = sizeof(IP_VALUE_TYPE);
    d.msg_size = sizeof(IP_MESSAGE_TYPE);
    d.mbox_size = sizeof(ipr_policy::slot_t);
    d.bslot_size = sizeof(ipr_bslot_t);
#if IPR_IS_PULL
    d.needs_in_lists = 1;
#endif
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
    d.needs_weights = 1;
#endif
    d.can_gather = ipr_policy::exact_gather;
    d.application = IP_APPLICATION;
    d.version = VERSION;
    d.machine = IP_MACHINE;
    d.commits = COMMITS;
    d.compilation_flags = IPR_STRINGIFY(COMPILATION_FLAGS);
    d.launch = ipr_app_launch;
    d.bind = ipr_app_bind;
    d.serialise = ipr_app_serialise;
    d.check_combiner = ipr_app_check_combiner;
#ifdef IP_ENABLE_THREAD_PROFILING
    d.thread_profiling = 1;
#endif
    ipr_drv_init(&d, file_path, number_of_threads, schedule, chunk_size, directed, weighted);
    ip_vertices_count = (size_t)ipr_drv_vertices_count();
    ip_edges_count = (size_t)ipr_drv_edges_count();
    ip_active_vertices = ip_vertices_count;
    ip_thread_count = number_of_threads;
    ip_superstep = 0;
}

int ip_run(void)
{
    int rc = ipr_drv_run();
    ip_superstep = (size_t)ipr_drv_superstep();
    ip_active_vertices = 0;
    return rc;
}

void ip_dump(FILE* f)
{
    ipr_drv_dump(f);
}

} /* extern "C" */

#endif /* IPREGEL_AMD_IPREGEL_H */
