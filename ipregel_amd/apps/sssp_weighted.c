/*
 * sssp_weighted.c - weighted single-source shortest paths on the MI355X
 * iPregel engine. The reference has no weighted application and its loader
 * reads no weights (benchmarks/sssp.c:66 sends distance + 1); this one reads
 * <graph>.wgt (uint32 per arc, aligned with <graph>.adj) and relaxes
 * distance + w(v,u) through ip_broadcast_weighted. Min combiner.
 * Command line: <in> <out> <threads> <schedule> <chunk> <source>.
 */
#include <stdlib.h>
#include <limits.h>
#include <inttypes.h>

typedef uint64_t IP_NEIGHBOUR_COUNT_TYPE;
typedef IP_VERTEX_ID_TYPE IP_MESSAGE_TYPE;
typedef IP_VERTEX_ID_TYPE IP_VALUE_TYPE;
#define IP_EDGE_WEIGHT_TYPE uint32_t
#define IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
#include "iPregel.h"

__device__ IP_VERTEX_ID_TYPE sssp_source;

__device__ void ip_compute(struct ip_vertex_t* v)
{
    if (ip_is_first_superstep()) {
        if (v->id == sssp_source) {
            v->value = 0;
            ip_broadcast_weighted(v, 0);
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
            ip_broadcast_weighted(v, best);
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
    printf("ApplicationConfiguration:startVertex=%u,weighted=1\n", (unsigned)sssp_source);
    ip_init(argv[1], atoi(argv[3]), argv[4], atoi(argv[5]), false, true);
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
