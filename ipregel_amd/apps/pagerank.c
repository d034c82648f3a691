/*
 * pagerank.c - PageRank as an iPregel application for the MI355X engine.
 *
 * Same recurrence, rounds, output format and command line as the reference
 * application benchmarks/pagerank.c:36-115: x0 = 1/V; then
 * x(v) = 0.15/V + 0.85 * (combined message); a vertex with out-degree > 0
 * broadcasts x/outdeg while the superstep is below 10, and halts at 10.
 * Build with -ffp-contract=off so 0.15/V + 0.85*sum rounds as on x86-64.
 */
#include <stdlib.h>
#include <inttypes.h>

typedef uint64_t IP_NEIGHBOUR_COUNT_TYPE;
typedef double IP_MESSAGE_TYPE;
typedef IP_MESSAGE_TYPE IP_VALUE_TYPE;
#define IP_NEEDS_OUT_NEIGHBOUR_COUNT
#include "iPregel.h"

#define PR_ROUNDS 10u

/* The reference application loads the graph with directed=false
 * (benchmarks/pagerank.c:91), so every vertex pulls over its own out-list. A
 * build with -DPAGERANK_DIRECTED passes directed=true instead: the in-lists
 * are mirrored from the out-lists in ascending source order
 * (src/iPregel_postamble.h:429-487) and the result is PageRank of G. */
#ifdef PAGERANK_DIRECTED
#define PR_DIRECTED true
#else
#define PR_DIRECTED false
#endif

/* set by main after ip_init, read by ip_compute on the device */
__device__ double pr_teleport;
__device__ double pr_start;

__device__ void ip_compute(struct ip_vertex_t* v)
{
    if (ip_is_first_superstep()) {
        v->value = pr_start;
    } else {
        IP_MESSAGE_TYPE total = 0.0;
        IP_MESSAGE_TYPE incoming;
        while (ip_get_next_message(v, &incoming)) total += incoming;
        v->value = pr_teleport + 0.85 * total;
    }
    if (ip_get_superstep() < PR_ROUNDS) {
        if (v->out_neighbour_count > 0) ip_broadcast(v, v->value / v->out_neighbour_count);
    } else {
        ip_vote_to_halt(v);
    }
}

__device__ void ip_combine(IP_MESSAGE_TYPE* a, IP_MESSAGE_TYPE b)
{
    *a += b;
}

void ip_serialise_vertex(FILE* f, struct ip_vertex_t* v)
{
    fprintf(f, "%lu: %0.20f\n", (unsigned long)v->id, v->value);
}

int main(int argc, char* argv[])
{
    if (argc != 6) {
        printf("Incorrect number of parameters, expecting: %s <inputFile> <outputFile> "
               "<number_of_threads> <schedule> <chunk_size>.\n", argv[0]);
        return -1;
    }
    printf("ApplicationConfiguration:maxSuperstepCount=10\n");
    ip_init(argv[1], atoi(argv[3]), argv[4], atoi(argv[5]), PR_DIRECTED, false);
    pr_teleport = 0.15 / ip_get_vertices_count();
    pr_start = 1.0 / ip_get_vertices_count();
    IP_PUBLISH(pr_teleport);
    IP_PUBLISH(pr_start);
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
