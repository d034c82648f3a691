/*
 * cc_send.c - hashmin connected components written with ip_send_message to
 * each out-neighbour instead of ip_broadcast. An API-coverage application of
 * the MI355X engine: it exercises
 *   - ip_send_message (src/combiner_postamble.h:54-79 in the full-sweep push
 *     variant, src/combiner_spread_postamble.h:73-98 in the spread one), whose
 *     deliveries go straight into the next superstep's mailboxes;
 *   - v->out_neighbours / v->out_neighbour_count, which the push variants give
 *     every vertex (src/combiner_preamble.h:21-27), in the application's id
 *     width (_32 and _64 builds);
 *   - IP_NEIGHBOUR_COUNT_TYPE = IP_VERTEX_ID_TYPE, as in
 *     benchmarks/template.c:2-3, so a _32 build reads 4-byte .idx offsets;
 *   - in a single-broadcast build, the reference's abort of ip_send_message
 *     (src/combiner_single_broadcast_postamble.h:42-49).
 * Sending the label to every out-neighbour is what ip_broadcast does, so the
 * labels and the per-superstep active counts are those of benchmarks/cc.c.
 * Command line and output format: benchmarks/cc.c:67,85-104.
 */
#include <stdlib.h>
#include <inttypes.h>

typedef IP_VERTEX_ID_TYPE IP_NEIGHBOUR_COUNT_TYPE;
typedef IP_VERTEX_ID_TYPE IP_MESSAGE_TYPE;
typedef IP_VERTEX_ID_TYPE IP_VALUE_TYPE;
#include "iPregel.h"

__device__ static void send_to_out_neighbours(struct ip_vertex_t* v, IP_MESSAGE_TYPE m)
{
#if defined(IP_USE_SINGLE_BROADCAST) && !defined(IP_USE_SPREAD)
    /* the single-broadcast variant has no out-neighbour list: one send is
     * enough to reach the reference's abort */
    ip_send_message(v->id, m);
#else
    for (IP_NEIGHBOUR_COUNT_TYPE i = 0; i < v->out_neighbour_count; i++)
        ip_send_message(v->out_neighbours[i], m);
#endif
}

__device__ void ip_compute(struct ip_vertex_t* v)
{
    if (ip_is_first_superstep()) {
        v->value = v->id;
        send_to_out_neighbours(v, v->value);
    } else {
        IP_VALUE_TYPE label = v->value;
        IP_MESSAGE_TYPE incoming;
        while (ip_get_next_message(v, &incoming))
            if (incoming < label) label = incoming;
        if (label != v->value) {
            v->value = label;
            send_to_out_neighbours(v, label);
        }
    }
    ip_vote_to_halt(v);
}

__device__ void ip_combine(IP_MESSAGE_TYPE* a, IP_MESSAGE_TYPE b)
{
    if (b < *a) *a = b;
}

void ip_serialise_vertex(FILE* f, struct ip_vertex_t* v)
{
    fprintf(f, "%llu: %llu\n", (unsigned long long)v->id, (unsigned long long)v->value);
}

int main(int argc, char* argv[])
{
    if (argc != 6) {
        printf("Incorrect number of parameters, expecting: %s <inputFile> <outputFile> "
               "<number_of_threads> <schedule> <chunk_size>.\n", argv[0]);
        return -1;
    }
    printf("ApplicationConfiguration:\n");
    ip_init(argv[1], atoi(argv[3]), argv[4], atoi(argv[5]), false, false);
    ip_run();
    FILE* out = fopen(argv[2], "w");
    if (!out) {
        perror("File opening failed.");
        return -1;
    }
    ip_dump(out);
    fclose(out);
    return EXIT_SUCCESS;
}
