/*
 * ipr_graph.c - synthetic graph generators (R-MAT, 2D grid), arc-list to CSR,
 * transpose, weights, and the .config/.idx/.adj(/.wgt) file triple.
 *
 * The triple layout follows iPregel's loader: src/iPregel_postamble.h:244-267
 * (config "V E"), :294-323 (V uint64 offsets, no sentinel), :325-360 (E ids).
 * The reference's own generator (src/graph_generators/graph_generator_ligra.cpp)
 * is uniform-random and writes 32-bit offsets, which its loader cannot read,
 * so this file defines the generators the engine is measured on.
 */
#define _GNU_SOURCE
#include "ipregel_graph.h"

#include <errno.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* helpers                                                                   */
/* ------------------------------------------------------------------------ */

static int cmp_u32(const void* a, const void* b)
{
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return (x > y) - (x < y);
}

static void sort_u32(uint32_t* p, uint64_t n)
{
    if (n < 2) return;
    if (n <= 24) {
        for (uint64_t i = 1; i < n; i++) {
            uint32_t x = p[i];
            uint64_t j = i;
            while (j > 0 && p[j - 1] > x) { p[j] = p[j - 1]; j--; }
            p[j] = x;
        }
        return;
    }
    qsort(p, n, sizeof(uint32_t), cmp_u32);
}

/* Exclusive scan of n counts into out[0..n] (out[n] = total), in parallel. */
static void scan_counts(const uint64_t* cnt, uint64_t* out, uint64_t n)
{
    int nt = omp_get_max_threads();
    uint64_t* part = calloc((size_t)nt + 1, sizeof(uint64_t));
    #pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num();
        uint64_t lo = n * t / nt, hi = n * (t + 1) / nt, s = 0;
        for (uint64_t i = lo; i < hi; i++) s += cnt[i];
        part[t + 1] = s;
        #pragma omp barrier
        #pragma omp single
        for (int k = 1; k <= nt; k++) part[k] += part[k - 1];
        s = part[t];
        for (uint64_t i = lo; i < hi; i++) { out[i] = s; s += cnt[i]; }
    }
    out[n] = part[nt];
    free(part);
}

/* Turn per-row buckets (rows[start[v] .. start[v]+fill[v]) unsorted) into a
 * compact CSR with sorted, deduplicated rows. Consumes `rows` and `start`. */
static int buckets_to_csr(uint64_t V, uint64_t* start, uint32_t* rows, ipr_csr* out)
{
    uint64_t* len = malloc(sizeof(uint64_t) * (V + 1));
    if (!len) return -ENOMEM;
    #pragma omp parallel for schedule(dynamic, 1024)
    for (uint64_t v = 0; v < V; v++) {
        uint32_t* r = rows + start[v];
        uint64_t n = start[v + 1] - start[v];
        sort_u32(r, n);
        uint64_t k = 0;
        for (uint64_t i = 0; i < n; i++)
            if (k == 0 || r[k - 1] != r[i]) r[k++] = r[i];
        len[v] = k;
    }
    uint64_t* idx = malloc(sizeof(uint64_t) * (V + 1));
    if (!idx) { free(len); return -ENOMEM; }
    scan_counts(len, idx, V);
    uint64_t E = idx[V];
    uint32_t* adj = malloc(sizeof(uint32_t) * (E ? E : 1));
    if (!adj) { free(len); free(idx); return -ENOMEM; }
    #pragma omp parallel for schedule(dynamic, 1024)
    for (uint64_t v = 0; v < V; v++)
        memcpy(adj + idx[v], rows + start[v], sizeof(uint32_t) * len[v]);
    free(len);
    out->V = V; out->E = E; out->idx = idx; out->adj = adj; out->wgt = NULL;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* R-MAT                                                                     */
/* ------------------------------------------------------------------------ */

typedef struct rmat_params {
    uint32_t scale;
    uint64_t seed;
    uint64_t t1, t2, t3; /* cumulative quadrant thresholds on a 32-bit draw */
} rmat_params;

static inline void rmat_draw(const rmat_params* p, uint64_t i, uint32_t* s, uint32_t* d)
{
    uint32_t src = 0, dst = 0;
    for (uint32_t l = 0; l < p->scale; l++) {
        uint64_t r = ipr_hash3(p->seed, i, l) >> 32;
        uint32_t rb = r >= p->t2;                    /* quadrants c, d: lower half */
        uint32_t cb = (r >= p->t1 && r < p->t2) || r >= p->t3; /* b or d: right half */
        src = (src << 1) | rb;
        dst = (dst << 1) | cb;
    }
    *s = src; *d = dst;
}

int ipr_graph_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                   double c, uint64_t seed, int symmetrise, ipr_csr* out)
{
    if (scale == 0 || scale > 31 || !out) return -EINVAL;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return -EINVAL;
    rmat_params p;
    p.scale = scale; p.seed = seed;
    p.t1 = (uint64_t)(a * 4294967296.0);
    p.t2 = (uint64_t)((a + b) * 4294967296.0);
    p.t3 = (uint64_t)((a + b + c) * 4294967296.0);
    uint64_t V = 1ull << scale;
    uint64_t M = (uint64_t)edge_factor << scale;

    uint64_t* cnt = calloc(V + 1, sizeof(uint64_t));
    uint64_t* start = malloc(sizeof(uint64_t) * (V + 1));
    if (!cnt || !start) { free(cnt); free(start); return -ENOMEM; }
    /* pass 1: row counts (self-loops dropped here) */
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < M; i++) {
        uint32_t s, d;
        rmat_draw(&p, i, &s, &d);
        if (s == d) continue;
        __atomic_fetch_add(&cnt[s], 1, __ATOMIC_RELAXED);
        if (symmetrise) __atomic_fetch_add(&cnt[d], 1, __ATOMIC_RELAXED);
    }
    scan_counts(cnt, start, V);
    uint64_t total = start[V];
    uint32_t* rows = malloc(sizeof(uint32_t) * (total ? total : 1));
    if (!rows) { free(cnt); free(start); return -ENOMEM; }
    memset(cnt, 0, sizeof(uint64_t) * (V + 1));
    /* pass 2: scatter (row order is fixed later by sorting) */
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < M; i++) {
        uint32_t s, d;
        rmat_draw(&p, i, &s, &d);
        if (s == d) continue;
        uint64_t k = __atomic_fetch_add(&cnt[s], 1, __ATOMIC_RELAXED);
        rows[start[s] + k] = d;
        if (symmetrise) {
            k = __atomic_fetch_add(&cnt[d], 1, __ATOMIC_RELAXED);
            rows[start[d] + k] = s;
        }
    }
    free(cnt);
    int rc = buckets_to_csr(V, start, rows, out);
    free(rows);
    free(start);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* grid                                                                      */
/* ------------------------------------------------------------------------ */

int ipr_graph_grid(uint32_t rows, uint32_t cols, ipr_csr* out)
{
    if (rows == 0 || cols == 0 || !out) return -EINVAL;
    uint64_t V = (uint64_t)rows * cols;
    if (V > 0xffffffffull) return -EINVAL;
    uint64_t* idx = malloc(sizeof(uint64_t) * (V + 1));
    if (!idx) return -ENOMEM;
    /* degree of (r, c) = [r>0] + [c>0] + [c<cols-1] + [r<rows-1] */
    #pragma omp parallel for schedule(static)
    for (uint64_t r = 0; r < rows; r++) {
        uint64_t base = r * cols;
        /* arcs in full rows before r */
        uint64_t before = 0;
        for (uint64_t rr = 0; rr < r; rr++)
            before += (uint64_t)cols * ((rr > 0) + (rr < rows - 1)) + 2ull * (cols - 1);
        uint64_t o = before;
        for (uint64_t c = 0; c < cols; c++) {
            idx[base + c] = o;
            o += (r > 0) + (c > 0) + (c + 1 < cols) + (r + 1 < rows);
        }
    }
    uint64_t E = 2ull * ((uint64_t)rows * (cols - 1) + (uint64_t)cols * (rows - 1));
    idx[V] = E;
    uint32_t* adj = malloc(sizeof(uint32_t) * (E ? E : 1));
    if (!adj) { free(idx); return -ENOMEM; }
    #pragma omp parallel for schedule(static)
    for (uint64_t r = 0; r < rows; r++) {
        for (uint64_t c = 0; c < cols; c++) {
            uint64_t v = r * cols + c, o = idx[v];
            if (r > 0) adj[o++] = (uint32_t)(v - cols);
            if (c > 0) adj[o++] = (uint32_t)(v - 1);
            if (c + 1 < cols) adj[o++] = (uint32_t)(v + 1);
            if (r + 1 < rows) adj[o++] = (uint32_t)(v + cols);
        }
    }
    out->V = V; out->E = E; out->idx = idx; out->adj = adj; out->wgt = NULL;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* weights, explicit arcs, transpose                                        */
/* ------------------------------------------------------------------------ */

int ipr_graph_weights(ipr_csr* g, uint64_t seed, uint32_t wmin, uint32_t wmax)
{
    if (!g || wmax < wmin) return -EINVAL;
    free(g->wgt);
    g->wgt = malloc(sizeof(uint32_t) * (g->E ? g->E : 1));
    if (!g->wgt) return -ENOMEM;
    uint64_t span = (uint64_t)wmax - wmin + 1;
    #pragma omp parallel for schedule(static)
    for (uint64_t e = 0; e < g->E; e++) {
        uint64_t r = ipr_hash3(seed, e, 0x57) >> 32;
        g->wgt[e] = wmin + (uint32_t)((r * span) >> 32);
    }
    return 0;
}

int ipr_graph_from_arcs(uint64_t V, uint64_t n, const uint32_t* src,
                        const uint32_t* dst, int keep_self_loops, ipr_csr* out)
{
    if (!out || (n && (!src || !dst))) return -EINVAL;
    uint64_t* cnt = calloc(V + 1, sizeof(uint64_t));
    uint64_t* start = malloc(sizeof(uint64_t) * (V + 1));
    if (!cnt || !start) { free(cnt); free(start); return -ENOMEM; }
    for (uint64_t i = 0; i < n; i++) {
        if (src[i] >= V || dst[i] >= V) { free(cnt); free(start); return -EINVAL; }
        if (src[i] == dst[i] && !keep_self_loops) continue;
        cnt[src[i]]++;
    }
    scan_counts(cnt, start, V);
    uint32_t* rows = malloc(sizeof(uint32_t) * (start[V] ? start[V] : 1));
    if (!rows) { free(cnt); free(start); return -ENOMEM; }
    memset(cnt, 0, sizeof(uint64_t) * (V + 1));
    for (uint64_t i = 0; i < n; i++) {
        if (src[i] == dst[i] && !keep_self_loops) continue;
        rows[start[src[i]] + cnt[src[i]]++] = dst[i];
    }
    free(cnt);
    int rc = buckets_to_csr(V, start, rows, out);
    free(rows);
    free(start);
    return rc;
}

int ipr_graph_transpose(const ipr_csr* g, ipr_csr* out)
{
    if (!g || !out) return -EINVAL;
    uint64_t V = g->V, E = g->E;
    uint64_t* cnt = calloc(V + 1, sizeof(uint64_t));
    uint64_t* idx = malloc(sizeof(uint64_t) * (V + 1));
    uint32_t* adj = malloc(sizeof(uint32_t) * (E ? E : 1));
    uint32_t* wgt = g->wgt ? malloc(sizeof(uint32_t) * (E ? E : 1)) : NULL;
    if (!cnt || !idx || !adj || (g->wgt && !wgt)) {
        free(cnt); free(idx); free(adj); free(wgt); return -ENOMEM;
    }
    for (uint64_t e = 0; e < E; e++) cnt[g->adj[e]]++;
    scan_counts(cnt, idx, V);
    memset(cnt, 0, sizeof(uint64_t) * (V + 1));
    /* sources visited in ascending order, so each in-row ends up ascending */
    for (uint64_t u = 0; u < V; u++)
        for (uint64_t e = g->idx[u]; e < g->idx[u + 1]; e++) {
            uint32_t v = g->adj[e];
            uint64_t pos = idx[v] + cnt[v]++;
            adj[pos] = (uint32_t)u;
            if (wgt) wgt[pos] = g->wgt[e];
        }
    free(cnt);
    out->V = V; out->E = E; out->idx = idx; out->adj = adj; out->wgt = wgt;
    return 0;
}

uint64_t ipr_graph_max_degree_vertex(const ipr_csr* g)
{
    uint64_t best = 0, bd = 0;
    for (uint64_t v = 0; v < g->V; v++) {
        uint64_t d = g->idx[v + 1] - g->idx[v];
        if (d > bd) { bd = d; best = v; }
    }
    return best;
}

/* ------------------------------------------------------------------------ */
/* file triple                                                               */
/* ------------------------------------------------------------------------ */

static FILE* open_ext(const char* root, const char* ext, const char* mode)
{
    size_t n = strlen(root) + strlen(ext) + 1;
    char* p = malloc(n);
    if (!p) return NULL;
    snprintf(p, n, "%s%s", root, ext);
    FILE* f = fopen(p, mode);
    free(p);
    return f;
}

int ipr_graph_write_triple(const char* root, const ipr_csr* g)
{
    FILE* f = open_ext(root, ".config", "w");
    if (!f) return -errno;
    fprintf(f, "%llu %llu\n", (unsigned long long)g->V, (unsigned long long)g->E);
    fclose(f);
    f = open_ext(root, ".idx", "wb");
    if (!f) return -errno;
    if (g->V && fwrite(g->idx, sizeof(uint64_t), g->V, f) != g->V) { fclose(f); return -EIO; }
    fclose(f);
    f = open_ext(root, ".adj", "wb");
    if (!f) return -errno;
    if (g->E && fwrite(g->adj, sizeof(uint32_t), g->E, f) != g->E) { fclose(f); return -EIO; }
    fclose(f);
    if (g->wgt) {
        f = open_ext(root, ".wgt", "wb");
        if (!f) return -errno;
        if (g->E && fwrite(g->wgt, sizeof(uint32_t), g->E, f) != g->E) { fclose(f); return -EIO; }
        fclose(f);
    }
    return 0;
}

int ipr_graph_read_triple(const char* root, int with_weights, ipr_csr* out)
{
    unsigned long long V = 0, E = 0;
    FILE* f = open_ext(root, ".config", "r");
    if (!f) return -errno;
    int ok = fscanf(f, "%llu %llu", &V, &E) == 2;
    fclose(f);
    if (!ok) return -EINVAL;
    uint64_t* idx = malloc(sizeof(uint64_t) * (V + 1));
    uint32_t* adj = malloc(sizeof(uint32_t) * (E ? E : 1));
    uint32_t* wgt = with_weights ? malloc(sizeof(uint32_t) * (E ? E : 1)) : NULL;
    if (!idx || !adj || (with_weights && !wgt)) { free(idx); free(adj); free(wgt); return -ENOMEM; }
    int rc = 0;
    f = open_ext(root, ".idx", "rb");
    if (!f) { rc = -errno; goto fail; }
    if (V && fread(idx, sizeof(uint64_t), V, f) != V) rc = -EIO;
    fclose(f);
    if (rc) goto fail;
    idx[V] = E;
    f = open_ext(root, ".adj", "rb");
    if (!f) { rc = -errno; goto fail; }
    if (E && fread(adj, sizeof(uint32_t), E, f) != E) rc = -EIO;
    fclose(f);
    if (rc) goto fail;
    if (wgt) {
        f = open_ext(root, ".wgt", "rb");
        if (!f) { rc = -errno; goto fail; }
        if (E && fread(wgt, sizeof(uint32_t), E, f) != E) rc = -EIO;
        fclose(f);
        if (rc) goto fail;
    }
    out->V = V; out->E = E; out->idx = idx; out->adj = adj; out->wgt = wgt;
    return 0;
fail:
    free(idx); free(adj); free(wgt);
    return rc;
}

void ipr_graph_free(ipr_csr* g)
{
    if (!g) return;
    free(g->idx); free(g->adj); free(g->wgt);
    g->idx = NULL; g->adj = NULL; g->wgt = NULL;
    g->V = g->E = 0;
}
