/*
 * sssp.c - unit-weight single-source shortest paths as an iPregel application
 * for the MI355X engine. Semantics and output format of the reference
 * application benchmarks/sssp.c:38-84: the source starts at 0 and broadcasts
 * 1; every other vertex starts at UINT_MAX; a vertex whose smallest message
 * improves its distance takes it and broadcasts distance + 1; all vote to
 * halt. Combiner: minimum (-DIP_COMBINER_MIN).
 *
 * Command line: <in> <out> <threads> <schedule> <chunk> <source>. The source
 * is read from the sixth argument; the reference reads atoi(argv[4]) (the
 * schedule string), which always yields vertex 0 (benchmarks/sssp.c:101).
 */
#include <stdlib.h>
#include <limits.h>
#include <inttypes.h>

typedef uint64_t IP_NEIGHBOUR_COUNT_TYPE;
typedef IP_VERTEX_ID_TYPE IP_MESSAGE_TYPE;
typedef IP_VERTEX_ID_TYPE IP_VALUE_TYPE;
#include "iPregel.h"

__device__ IP_VERTEX_ID_TYPE sssp_source;

__device__ void ip_compute(struct ip_vertex_t* v)
{
    if (ip_is_first_superstep()) {
        if (v->id == sssp_source) {
            v->value = 0;
            ip_broadcast(v, 1);
        } else {
            v->value = UINT_MAX;
        }
    } else {
        IP_MESSAGE_TYPE best = UINT_MAX;
        IP_MESSAGE_TYPE incoming;
        while (ip_get_next_message(v, &incoming))
            if (incoming < best) best = incoming;
        if (best < v->value) {
            v->value = best;
            ip_broadcast(v, best + 1);
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
    fprintf(f, "%u: %u\n", (unsigned)v->id, (unsigned)v->value);
}

int main(int argc, char* argv[])
{
    if (argc != 7) {
        printf("Incorrect number of parameters, expecting: %s <inputFile> <outputFile> "
               "<number_of_threads> <schedule> <chunk_size> <SSSP_source_vertex>.\n", argv[0]);
        return -1;
    }
    sssp_source = (IP_VERTEX_ID_TYPE)strtoul(argv[6], NULL, 10);
    printf("ApplicationConfiguration:startVertex=%u\n", (unsigned)sssp_source);
    ip_init(argv[1], atoi(argv[3]), argv[4], atoi(argv[5]), false, false);
    IP_PUBLISH(sssp_source);
    ip_run();
    FILE* out = fopen(argv[2], "w");
    if (!out) {
        perror("File opening failed.");
        return -1;
    }
    ip_dump(out);
    fclose(out);
    return 0;
}
