/*
 * nbr_view.c - an API-coverage application of the MI355X engine for the
 * neighbour lists and the host-side vertex lookup:
 *   - ip_compute (superstep 0) reads v->in_neighbours and
 *     v->in_neighbour_count, in the application's id width, and stores an
 *     order-sensitive checksum of the in-list: sum over i of
 *     (in_neighbours[i] + 1) * (i + 1), modulo 2^64. With directed=true the
 *     engine mirrors the in-lists in ascending source order
 *     (src/iPregel_postamble.h:429-487); with directed=false they are the
 *     out-lists (:361-381);
 *   - after ip_run, main walks ip_get_vertex_by_id (src/iPregel_postamble.h:
 *     83-101) and folds the same checksum of v->out_neighbours into the value
 *     through the returned pointer; ip_dump then serialises those edited
 *     vertices, as the reference's ip_dump serialises its vertex array
 *     (:103-115).
 * Every vertex votes to halt in superstep 0 and sends nothing.
 * Command line: <in> <out> <threads> <schedule> <chunk> <directed 0|1>.
 * Output line per vertex: "id: value in_count out_count".
 */
#include <stdlib.h>
#include <inttypes.h>

#define IP_NEEDS_OUT_NEIGHBOUR_IDS
#define IP_NEEDS_IN_NEIGHBOUR_IDS
typedef uint64_t IP_NEIGHBOUR_COUNT_TYPE;
typedef IP_VERTEX_ID_TYPE IP_MESSAGE_TYPE;
typedef uint64_t IP_VALUE_TYPE;
#include "iPregel.h"

__host__ __device__ static uint64_t list_checksum(const IP_VERTEX_ID_TYPE* ids, IP_NEIGHBOUR_COUNT_TYPE n)
{
    uint64_t h = 0;
    for (IP_NEIGHBOUR_COUNT_TYPE i = 0; i < n; i++) h += ((uint64_t)ids[i] + 1u) * (uint64_t)(i + 1);
    return h;
}

__device__ void ip_compute(struct ip_vertex_t* v)
{
    if (ip_is_first_superstep()) v->value = list_checksum(v->in_neighbours, v->in_neighbour_count);
    ip_vote_to_halt(v);
}

__device__ void ip_combine(IP_MESSAGE_TYPE* a, IP_MESSAGE_TYPE b)
{
    if (b < *a) *a = b;
}

void ip_serialise_vertex(FILE* f, struct ip_vertex_t* v)
{
    fprintf(f, "%llu: %llu %llu %llu\n", (unsigned long long)v->id, (unsigned long long)v->value,
            (unsigned long long)v->in_neighbour_count, (unsigned long long)v->out_neighbour_count);
}

int main(int argc, char* argv[])
{
    if (argc != 7) {
        printf("Incorrect number of parameters, expecting: %s <inputFile> <outputFile> "
               "<number_of_threads> <schedule> <chunk_size> <directed>.\n", argv[0]);
        return -1;
    }
    const bool directed = atoi(argv[6]) != 0;
    printf("ApplicationConfiguration:directed=%d\n", (int)directed);
    ip_init(argv[1], atoi(argv[3]), argv[4], atoi(argv[5]), directed, false);
    ip_run();
    for (size_t i = 0; i < ip_get_vertices_count(); i++) {
        struct ip_vertex_t* v = ip_get_vertex_by_id((IP_VERTEX_ID_TYPE)i);
        v->value = v->value * 31u + list_checksum(v->out_neighbours, v->out_neighbour_count);
    }
    FILE* out = fopen(argv[2], "w");
    if (!out) {
        perror("File opening failed.");
        return -1;
    }
    ip_dump(out);
    fclose(out);
    return EXIT_SUCCESS;
}
