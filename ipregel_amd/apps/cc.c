/*
 * cc.c - connected components by label propagation (hashmin) as an iPregel
 * application for the MI355X engine. Same semantics, output format and
 * command line as the reference application benchmarks/cc.c:31-104: every
 * vertex starts with its own id and broadcasts it; afterwards it takes the
 * minimum of its label and its messages and broadcasts again when the label
 * dropped; it always votes to halt. Combiner: minimum (-DIP_COMBINER_MIN).
 */
#include <stdlib.h>
#include <inttypes.h>

typedef uint64_t IP_NEIGHBOUR_COUNT_TYPE;
typedef IP_VERTEX_ID_TYPE IP_MESSAGE_TYPE;
typedef IP_VERTEX_ID_TYPE IP_VALUE_TYPE;
#include "iPregel.h"

__device__ void ip_compute(struct ip_vertex_t* v)
{
    if (ip_is_first_superstep()) {
        v->value = v->id;
        ip_broadcast(v, v->value);
    } else {
        IP_VALUE_TYPE label = v->value;
        IP_MESSAGE_TYPE incoming;
        while (ip_get_next_message(v, &incoming))
            if (incoming < label) label = incoming;
        if (label != v->value) {
            v->value = label;
            ip_broadcast(v, label);
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
