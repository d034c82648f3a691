/*
 * ipregel_graph.h - synthetic graph generation and the on-disk graph triple.
 *
 * The engine reads the graph format of iPregel's loader
 * (src/iPregel_postamble.h:244-267 config, :294-323 offsets, :325-360 edges):
 *   <root>.config  text "V E"
 *   <root>.idx     V start offsets, uint64 each, no trailing sentinel
 *   <root>.adj     E destination ids, uint32 each, rows in vertex order
 * plus one file of its own for weighted graphs (the reference loads no weights):
 *   <root>.wgt     E uint32 weights, aligned with <root>.adj
 *
 * In memory a graph is a CSR with a V+1 offset array.
 * All functions are plain C; ids are uint32.
 */
#ifndef IPREGEL_GRAPH_H
#define IPREGEL_GRAPH_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ipr_csr {
    uint64_t  V;     /* vertex count */
    uint64_t  E;     /* arc count */
    uint64_t* idx;   /* V+1 offsets */
    uint32_t* adj;   /* E destination ids, each row ascending */
    uint32_t* wgt;   /* E weights or NULL */
} ipr_csr;

/* R-MAT: 2^scale vertices, edge_factor*2^scale draws. Every draw descends
 * `scale` levels; level l of draw i picks a quadrant from the top 32 bits of
 * ipr_hash3(seed, i, l) against the cumulative thresholds floor(a*2^32),
 * floor((a+b)*2^32), floor((a+b+c)*2^32). No vertex permutation, no noise.
 * Self-loops and duplicate arcs are removed; each row is sorted ascending.
 * symmetrise != 0 adds the reverse of every arc before deduplication. */
int ipr_graph_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                   double c, uint64_t seed, int symmetrise, ipr_csr* out);

/* rows x cols 4-neighbour grid, id = r*cols + c, arcs in both directions,
 * each row ascending (up, left, right, down). */
int ipr_graph_grid(uint32_t rows, uint32_t cols, ipr_csr* out);

/* Fill g->wgt (allocating it) with weights uniform in [wmin, wmax]:
 * w(e) = wmin + ((top32(ipr_hash3(seed, e, 0x57)) * (wmax-wmin+1)) >> 32),
 * e being the CSR arc index. */
int ipr_graph_weights(ipr_csr* g, uint64_t seed, uint32_t wmin, uint32_t wmax);

/* Build a graph from an explicit arc list (deduplicated, self-loops removed
 * unless keep_self_loops, rows sorted). Used for hand-made fixtures. */
int ipr_graph_from_arcs(uint64_t V, uint64_t n, const uint32_t* src,
                        const uint32_t* dst, int keep_self_loops, ipr_csr* out);

/* Transpose: in-lists, each row ascending by source id. */
int ipr_graph_transpose(const ipr_csr* g, ipr_csr* out);

int  ipr_graph_write_triple(const char* root, const ipr_csr* g);
int  ipr_graph_read_triple(const char* root, int with_weights, ipr_csr* out);
void ipr_graph_free(ipr_csr* g);

/* Highest out-degree vertex, ties to the lowest id. */
uint64_t ipr_graph_max_degree_vertex(const ipr_csr* g);

/* Counter-based hash used by every generator (and by the device generator).
 * splitmix64 finaliser over a mixed (seed, i, l) key. */
static inline uint64_t ipr_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static inline uint64_t ipr_hash3(uint64_t seed, uint64_t i, uint64_t l)
{
    return ipr_mix64(ipr_mix64(seed + 0x9e3779b97f4a7c15ULL) ^ (i * 64u + l));
}

#ifdef __cplusplus
}
#endif
#endif
