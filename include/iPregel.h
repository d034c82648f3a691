/*
 * iPregel.h - MI355X (gfx950) engine behind iPregel's programming surface.
 *
 * An application written for iPregel (src/iPregel.h, src/iPregel_preamble.h)
 * includes this header instead and is compiled by hipcc as HIP. It keeps:
 *   - the type macros IP_VERTEX_ID_TYPE (uint32_t), IP_NEIGHBOUR_COUNT_TYPE,
 *     IP_MESSAGE_TYPE, IP_VALUE_TYPE (defined by the app before inclusion),
 *   - the variant switches IP_USE_SPREAD / IP_USE_SINGLE_BROADCAST
 *     (src/iPregel_preamble.h:319-331),
 *   - the hooks ip_combine / ip_compute (now __device__) and
 *     ip_serialise_vertex (host) (src/iPregel_preamble.h:262,283,275),
 *   - ip_send_message, ip_broadcast, ip_get_next_message, ip_has_message,
 *     ip_vote_to_halt (device side, src/iPregel_preamble.h:122-166),
 *   - ip_get_superstep, ip_is_first_superstep, ip_get_vertices_count,
 *     ip_get_edges_count (host and device, src/iPregel_preamble.h:66-96),
 *   - ip_get_vertex_by_location, ip_get_vertex_by_id (host, after ip_run,
 *     src/iPregel_preamble.h:104,110),
 *   - ip_init, ip_run, ip_dump (host, src/iPregel_preamble.h:298,309,317).
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
static_assert(sizeof(IP_VERTEX_ID_TYPE) == 4 || sizeof(IP_VERTEX_ID_TYPE) == 8,
              "IP_VERTEX_ID_TYPE must be a 4- or 8-byte integer (the _32 and _64 builds)");
static_assert(sizeof(IP_NEIGHBOUR_COUNT_TYPE) == 4 || sizeof(IP_NEIGHBOUR_COUNT_TYPE) == 8,
              "IP_NEIGHBOUR_COUNT_TYPE must be a 4- or 8-byte integer (the width of the .idx offsets)");

/* fields every vertex of a variant carries: out-neighbour ids and count in
 * the push variants (src/combiner_preamble.h:21-27,
 * src/combiner_spread_preamble.h:28-34), in-neighbour ids and count in the
 * single-broadcast ones (src/combiner_single_broadcast_preamble.h:32-38),
 * both in spread + single broadcast
 * (src/combiner_spread_single_broadcast_preamble.h:21-35). Unused fields cost
 * nothing on the device: the proxy's loads are dropped when nothing reads
 * them. */
#if !defined(IP_USE_SINGLE_BROADCAST) || defined(IP_USE_SPREAD)
#ifndef IP_NEEDS_OUT_NEIGHBOUR_IDS
#define IP_NEEDS_OUT_NEIGHBOUR_IDS
#endif
#ifndef IP_NEEDS_OUT_NEIGHBOUR_COUNT
#define IP_NEEDS_OUT_NEIGHBOUR_COUNT
#endif
#endif
#if defined(IP_USE_SINGLE_BROADCAST)
#ifndef IP_NEEDS_IN_NEIGHBOUR_IDS
#define IP_NEEDS_IN_NEIGHBOUR_IDS
#endif
#ifndef IP_NEEDS_IN_NEIGHBOUR_COUNT
#define IP_NEEDS_IN_NEIGHBOUR_COUNT
#endif
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

/* Host globals of the reference preamble (src/iPregel_preamble.h:45-57),
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

/* host-side vertex lookup after ip_run (src/iPregel_preamble.h:104,110) */
struct ip_vertex_t* ip_get_vertex_by_location(size_t location);
struct ip_vertex_t* ip_get_vertex_by_id(IP_VERTEX_ID_TYPE id);
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

/* src/combiner_postamble.h:29-39 (identical in all variants) */
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
 * src/combiner_single_broadcast_postamble.h:51-55. */
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
    (void)v; (void)message;
    ipr_raise_error(IPR_ERR_WEIGHTED_IN_PULL);
#else
    if (v->ipr_bcast) ipr_raise_error(IPR_ERR_MIXED_BROADCAST);
    v->ipr_bcast = true;
    v->ipr_bcast_weighted = true;
    v->ipr_bcast_msg = message;
#endif
}
#endif

/* Push variants: immediate delivery into the next superstep's mailbox
 * (src/combiner_postamble.h:54-79, src/combiner_spread_postamble.h:73-98).
 * Pull variants: the reference prints and exits
 * (src/combiner_single_broadcast_postamble.h:42-49,
 * src/combiner_spread_single_broadcast_postamble.h:55-62); here the kernel
 * flags the error and the driver prints the same text and exits after the
 * launch. An id outside [0, V) is flagged too (the reference would write out
 * of bounds) and the message is dropped. */
__device__ inline void ip_send_message(IP_VERTEX_ID_TYPE id, IP_MESSAGE_TYPE message)
{
#if IPR_IS_PULL
    (void)id; (void)message;
    ipr_raise_error(IPR_ERR_SEND_IN_PULL);
#else
    if ((uint64_t)id >= ipr_d_dev.V) {
        ipr_raise_error(IPR_ERR_SEND_OUT_OF_RANGE);
        return;
    }
    ipr_send_immediate(id, message);
#endif
}

__device__ inline void ip_vote_to_halt(struct ip_vertex_t* v)
{
    /* a no-op in the spread variants (src/combiner_spread_postamble.h:433-436);
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

static void ipr_host_view_free(void);

void ip_init(const char* file_path, int number_of_threads, const char* schedule,
             int chunk_size, bool directed, bool weighted)
{
    ipr_host_view_free();
    ipr_app_desc d;
    memset(&d, 0, sizeof(d));
    d.variant = IPR_VARIANT;
    d.combiner = IPR_COMBINER;
    d.id_size = sizeof(IP_VERTEX_ID_TYPE);
    d.val_size = sizeof(IP_VALUE_TYPE);
    d.msg_size = sizeof(IP_MESSAGE_TYPE);
    d.mbox_size = sizeof(ipr_policy::slot_t);
    d.bslot_size = sizeof(ipr_bslot_t);
    d.idx_size = sizeof(IP_NEIGHBOUR_COUNT_TYPE);
#if IPR_IS_PULL || defined(IP_NEEDS_IN_NEIGHBOUR_IDS) || defined(IP_NEEDS_IN_NEIGHBOUR_COUNT)
    /* in-lists are mirrored from the out-lists when the graph is directed
     * (src/iPregel_postamble.h:429-487) */
    d.needs_in_lists = 1;
#endif
#ifdef IP_NEEDS_OUT_NEIGHBOUR_IDS
    d.needs_out_ids = 1;
#endif
#ifdef IP_NEEDS_IN_NEIGHBOUR_IDS
    d.needs_in_ids = 1;
#endif
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
    d.needs_weights = 1;
#endif
    d.can_gather = ipr_policy::exact_gather;
    d.can_tile = ipr_tile_ok;
    d.tile_receivers = ipr_tile_r;
    d.tile_wgs_per_cu = ipr_tile_wgs_per_cu;
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

/* ------------------------------------------------------------------------ */
/* host view of the vertices: ip_get_vertex_by_location / ip_get_vertex_by_id */
/* ------------------------------------------------------------------------ */

/* The reference keeps every vertex in one host array and returns pointers
 * into it (src/iPregel_postamble.h:78-101). Here the vertex state lives in
 * HBM as structure of arrays, so the first lookup after ip_run builds a host
 * array of ip_vertex_t from it: id, value, active, has_message, message,
 * neighbour counts and neighbour id / weight pointers into host copies of the
 * adjacency. The pointers stay valid until the next ip_init or ip_run; ip_dump
 * serialises this array while it is current, so edits made through the
 * pointers show in the dump as they would in the reference. Host code only:
 * device code has no vertex array to point into. */
static struct ip_vertex_t* ipr_host_vertices = NULL;
static uint64_t ipr_host_vertices_run = ~0ull;
static IP_VERTEX_ID_TYPE* ipr_host_ids[2] = {NULL, NULL};   /* out, in: IP_VERTEX_ID_TYPE copies */
static uint64_t* ipr_host_in_idx = NULL;

static void ipr_host_view_free(void)
{
    free(ipr_host_vertices);
    ipr_host_vertices = NULL;
    if (ipr_host_ids[1] != ipr_host_ids[0]) free(ipr_host_ids[1]);
    free(ipr_host_ids[0]);
    ipr_host_ids[0] = ipr_host_ids[1] = NULL;
    free(ipr_host_in_idx);
    ipr_host_in_idx = NULL;
    ipr_host_vertices_run = ~0ull;
}

static bool ipr_host_view_current(void)
{
    return ipr_host_vertices != NULL && ipr_host_vertices_run == ipr_drv_runs();
}

static void* ipr_host_alloc(size_t n)
{
    void* p = malloc(n ? n : 1);
    if (!p) {
        printf("Failed to allocate %zu bytes for the host view of the vertices.\n", n);
        exit(-1);
    }
    return p;
}

static void ipr_host_view_build(void)
{
    ipr_host_view_free();
    const uint64_t V = ipr_drv_vertices_count(), E = ipr_drv_edges_count();
    const uint64_t* idx = ipr_drv_host_out_idx();
    IP_VALUE_TYPE* val = (IP_VALUE_TYPE*)ipr_host_alloc(V * sizeof(IP_VALUE_TYPE));
    IP_MESSAGE_TYPE* msg = (IP_MESSAGE_TYPE*)ipr_host_alloc(V * sizeof(IP_MESSAGE_TYPE));
    uint8_t* act = (uint8_t*)ipr_host_alloc(V);
    uint8_t* hm = (uint8_t*)ipr_host_alloc(V);
    if (ipr_drv_vertex_state(val, act, hm, msg)) {
        printf("Failed to read the vertex state back from the device.\n");
        exit(-1);
    }
#if defined(IP_NEEDS_OUT_NEIGHBOUR_IDS) || defined(IP_NEEDS_IN_NEIGHBOUR_IDS)
    /* out-neighbour ids in the application's id type */
    {
        const uint32_t* adj = ipr_drv_host_out_adj();
        ipr_host_ids[0] = (IP_VERTEX_ID_TYPE*)ipr_host_alloc(E * sizeof(IP_VERTEX_ID_TYPE));
        for (uint64_t e = 0; e < E; e++) ipr_host_ids[0][e] = (IP_VERTEX_ID_TYPE)adj[e];
    }
#endif
#if defined(IP_NEEDS_IN_NEIGHBOUR_IDS) || defined(IP_NEEDS_IN_NEIGHBOUR_COUNT)
    /* in-lists: the out-lists when undirected, the device's mirror otherwise */
    if (ipr_drv_directed()) {
        ipr_host_in_idx = (uint64_t*)ipr_host_alloc((V + 1) * sizeof(uint64_t));
        uint32_t* in_adj = (uint32_t*)ipr_host_alloc(E * sizeof(uint32_t));
        if (ipr_drv_in_lists(ipr_host_in_idx, in_adj)) {
            printf("Failed to read the in-neighbour lists back from the device.\n");
            exit(-1);
        }
#ifdef IP_NEEDS_IN_NEIGHBOUR_IDS
        ipr_host_ids[1] = (IP_VERTEX_ID_TYPE*)ipr_host_alloc(E * sizeof(IP_VERTEX_ID_TYPE));
        for (uint64_t e = 0; e < E; e++) ipr_host_ids[1][e] = (IP_VERTEX_ID_TYPE)in_adj[e];
#endif
        free(in_adj);
    } else {
        ipr_host_ids[1] = ipr_host_ids[0];
    }
#endif
    const bool spread = IPR_VARIANT == IPR_VARIANT_SPREAD || IPR_VARIANT == IPR_VARIANT_SPREAD_PULL;
    ipr_host_vertices = (struct ip_vertex_t*)ipr_host_alloc(V * sizeof(struct ip_vertex_t));
    for (uint64_t v = 0; v < V; v++) {
        struct ip_vertex_t* x = &ipr_host_vertices[v];
        memset(x, 0, sizeof(*x));
        x->id = (IP_VERTEX_ID_TYPE)v;
        x->value = val[v];
        /* the spread variants keep no active flags (ip_vote_to_halt is a no-op
         * there) and consume every message within its superstep */
        x->active = !spread && act[v] != 0;
        x->has_message = !spread && hm[v] != 0;
        if (x->has_message) x->message = msg[v];
#ifdef IP_NEEDS_OUT_NEIGHBOUR_COUNT
        x->out_neighbour_count = (IP_NEIGHBOUR_COUNT_TYPE)(idx[v + 1] - idx[v]);
#endif
#ifdef IP_NEEDS_OUT_NEIGHBOUR_IDS
        x->out_neighbours = ipr_host_ids[0] + idx[v];
#endif
#ifdef IP_NEEDS_IN_NEIGHBOUR_COUNT
        {
            const uint64_t* in_idx = ipr_host_in_idx ? ipr_host_in_idx : idx;
            x->in_neighbour_count = (IP_NEIGHBOUR_COUNT_TYPE)(in_idx[v + 1] - in_idx[v]);
        }
#endif
#ifdef IP_NEEDS_IN_NEIGHBOUR_IDS
        x->in_neighbours = ipr_host_ids[1] + (ipr_host_in_idx ? ipr_host_in_idx : idx)[v];
#endif
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
        x->out_neighbour_weights = (IP_EDGE_WEIGHT_TYPE*)ipr_drv_host_out_wgt() + idx[v];
#endif
    }
    (void)E;
    free(val);
    free(msg);
    free(act);
    free(hm);
    ipr_host_vertices_run = ipr_drv_runs();
}

/* src/iPregel_postamble.h:78-81; the location is the id here (the reference's
 * IP_FORCE_DIRECT_MAPPING / IP_ID_OFFSET == 0 case, :85-93) */
struct ip_vertex_t* ip_get_vertex_by_location(size_t location)
{
    if (!ipr_host_view_current()) ipr_host_view_build();
    return &ipr_host_vertices[location];
}

struct ip_vertex_t* ip_get_vertex_by_id(IP_VERTEX_ID_TYPE id)
{
    return ip_get_vertex_by_location((size_t)id);
}

int ip_run(void)
{
    ipr_host_view_free();
    int rc = ipr_drv_run();
    ip_superstep = (size_t)ipr_drv_superstep();
    ip_active_vertices = 0;
    return rc;
}

/* src/iPregel_postamble.h:103-115 */
void ip_dump(FILE* f)
{
    if (!ipr_host_view_current()) {
        ipr_drv_dump(f);
        return;
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    const uint64_t V = ipr_drv_vertices_count();
    for (uint64_t v = 0; v < V; v++) ip_serialise_vertex(f, &ipr_host_vertices[v]);
    fflush(f);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    printf("DumpingTime:%f\n", (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec));
}

} /* extern "C" */

#endif /* IPREGEL_AMD_IPREGEL_H */
