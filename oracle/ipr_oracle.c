/*
 * ipr_oracle.c - CPU oracle for the iPregel superstep hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never by the product.
 *
 * A single-threaded restatement of the BSP semantics of iPregel's four engine
 * variants (citations per function) and of its applications. Sends are
 * processed in program order, so for the exact combiners used here (integer
 * min, and FP64 sum only in the pull variant whose fold order is the in-list
 * order) the results do not depend on thread schedules, matching the reference.
 *
 * Pinned by: tests/test_oracle.py against golden outputs of the reference's
 * own binaries (tests/golden/), against min-ancestor labels / BFS / Dijkstra
 * (closed forms below), and against dense power iteration for PR.
 */
#include "ipr_oracle.h"

#include <errno.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef union msg { double f; uint32_t u; uint64_t bits; } msg_t;

typedef struct o_vertex {
    uint32_t id;
    uint64_t out_count;
    const uint32_t* out;
    const uint32_t* out_w;
    int active;
    int has_message;
    msg_t message;
    msg_t value;
} o_vertex;

typedef struct o_engine o_engine;
typedef void (*compute_fn)(o_engine*, o_vertex*);
typedef void (*combine_fn)(msg_t*, msg_t);

struct o_engine {
    const ipr_csr* g;
    uint64_t V;
    /* in-lists */
    const uint64_t* in_idx;
    const uint32_t* in_adj;
    int variant;
    uint64_t superstep;
    const ipr_oracle_params* p;
    compute_fn compute;
    combine_fn combine;
    o_vertex* vx;
    /* push mailboxes (message_next / has_message_next) */
    msg_t* nxt;
    unsigned char* has_nxt;
    /* spread: next frontier being built */
    uint32_t* nf;
    uint64_t nf_size;
    /* pull: broadcast slots */
    unsigned char* bcf;
    msg_t* bcm;
    /* spread_pull: broadcast targets */
    unsigned char* tgt;
    ipr_oracle_result* res;
    double t0;              /* start of the run, for p->stop_seconds */
};

/* ---------------------------------------------------------------- API ---- */

static int o_get_next_message(o_engine* e, o_vertex* v, msg_t* m)
{
    (void)e;
    /* every variant: src/combiner_postamble.h:29-39 and its three twins */
    if (v->has_message) { *m = v->message; v->has_message = 0; return 1; }
    return 0;
}

static void o_vote_to_halt(o_engine* e, o_vertex* v)
{
    /* no-op in the two spread variants (combiner_spread_postamble.h:433-436) */
    if (e->variant == IPR_O_PUSH || e->variant == IPR_O_PULL) v->active = 0;
}

/* Push delivery: first message of the superstep is stored, the rest combined
 * (combiner_postamble.h:41-79; spread version also appends the receiver to
 * the next frontier, combiner_spread_postamble.h:46-98). */
static void o_send(o_engine* e, uint32_t id, msg_t m)
{
    if (e->variant == IPR_O_PULL || e->variant == IPR_O_SPREAD_PULL) {
        fprintf(stderr, "oracle: send_message in a broadcast-only variant\n");
        exit(-1);
    }
    e->res->messages++;
    if (e->has_nxt[id]) {
        e->combine(&e->nxt[id], m);
    } else {
        e->nxt[id] = m;
        e->has_nxt[id] = 1;
        if (e->variant == IPR_O_SPREAD) e->nf[e->nf_size++] = id;
    }
}

static void o_broadcast(o_engine* e, o_vertex* v, msg_t m)
{
    e->res->work_edges += (e->variant == IPR_O_PULL) ? 0 : v->out_count;
    if (e->variant == IPR_O_PULL) {
        /* combiner_single_broadcast_postamble.h:51-55 */
        e->bcf[v->id] = 1; e->bcm[v->id] = m;
    } else if (e->variant == IPR_O_SPREAD_PULL) {
        /* combiner_spread_single_broadcast_postamble.h ip_broadcast */
        e->bcf[v->id] = 1; e->bcm[v->id] = m;
        for (uint64_t i = 0; i < v->out_count; i++) e->tgt[v->out[i]] = 1;
    } else {
        for (uint64_t i = 0; i < v->out_count; i++) o_send(e, v->out[i], m);
    }
}

/* Project extension with no reference counterpart: every out-neighbour u
 * receives m + w(v,u). Push variants only. */
static void o_broadcast_weighted(o_engine* e, o_vertex* v, msg_t m)
{
    e->res->work_edges += v->out_count;
    for (uint64_t i = 0; i < v->out_count; i++) {
        msg_t mm; mm.bits = 0; mm.u = m.u + v->out_w[i];
        o_send(e, v->out[i], mm);
    }
}

/* Pull fetch: first broadcasting in-neighbour's message, then the others
 * combined in in-list order (combiner_single_broadcast_postamble.h:57-89).
 * Returns 1 when a message was found. */
static int o_fetch(o_engine* e, uint32_t v)
{
    o_vertex* x = &e->vx[v];
    uint64_t i = e->in_idx[v], end = e->in_idx[v + 1];
    while (i < end && !e->bcf[e->in_adj[i]]) i++;
    if (i >= end) { x->has_message = 0; return 0; }
    x->has_message = 1;
    x->message = e->bcm[e->in_adj[i]];
    for (i++; i < end; i++)
        if (e->bcf[e->in_adj[i]]) e->combine(&x->message, e->bcm[e->in_adj[i]]);
    return 1;
}

/* --------------------------------------------------------------- apps ---- */

/* benchmarks/pagerank.c:36-66 with ratio = 0.15/V, initial = 1/V
 * (pagerank.c:98-99), ROUND = 10 (pagerank.c:34). No FMA: the reference is
 * compiled for plain x86-64 where contraction is unavailable. */
static void app_pr(o_engine* e, o_vertex* v)
{
    double n = (double)e->V;
    double ratio = 0.15 / n;
    if (e->superstep == 0) {
        v->value.f = 1.0 / n;
    } else {
        double sum = 0.0;
        msg_t m;
        while (o_get_next_message(e, v, &m)) sum += m.f;
        volatile double prod = 0.85 * sum; /* keep mul and add separate */
        v->value.f = ratio + prod;
    }
    if (e->superstep < e->p->rounds) {
        if (v->out_count > 0) {
            msg_t m; m.f = v->value.f / (double)v->out_count;
            o_broadcast(e, v, m);
        }
    } else {
        o_vote_to_halt(e, v);
    }
}
static void comb_sum_f64(msg_t* a, msg_t b) { a->f += b.f; }

/* benchmarks/cc.c:31-63 */
static void app_cc(o_engine* e, o_vertex* v)
{
    if (e->superstep == 0) {
        v->value.u = v->id;
        o_broadcast(e, v, v->value);
    } else {
        uint32_t before = v->value.u;
        msg_t m;
        while (o_get_next_message(e, v, &m))
            if (v->value.u > m.u) v->value.u = m.u;
        if (before != v->value.u) o_broadcast(e, v, v->value);
    }
    o_vote_to_halt(e, v);
}
static void comb_min_u32(msg_t* a, msg_t b) { if (a->u > b.u) a->u = b.u; }

/* benchmarks/sssp.c:38-71 (message = distance + 1) */
static void app_sssp(o_engine* e, o_vertex* v)
{
    if (e->superstep == 0) {
        if (v->id == e->p->source) {
            v->value.u = 0;
            msg_t m; m.bits = 0; m.u = 1;
            o_broadcast(e, v, m);
        } else {
            v->value.u = UINT_MAX;
        }
    } else {
        uint32_t best = UINT_MAX;
        msg_t m;
        while (o_get_next_message(e, v, &m))
            if (best > m.u) best = m.u;
        if (best < v->value.u) {
            v->value.u = best;
            msg_t b; b.bits = 0; b.u = best + 1;
            o_broadcast(e, v, b);
        }
    }
    o_vote_to_halt(e, v);
}

/* Weighted SSSP (project application, no reference counterpart). */
static void app_sssp_w(o_engine* e, o_vertex* v)
{
    if (e->superstep == 0) {
        if (v->id == e->p->source) {
            v->value.u = 0;
            o_broadcast_weighted(e, v, v->value);
        } else {
            v->value.u = UINT_MAX;
        }
    } else {
        uint32_t best = UINT_MAX;
        msg_t m;
        while (o_get_next_message(e, v, &m))
            if (best > m.u) best = m.u;
        if (best < v->value.u) {
            v->value.u = best;
            o_broadcast_weighted(e, v, v->value);
        }
    }
    o_vote_to_halt(e, v);
}

/* ------------------------------------------------------------ engines ---- */

static void record(o_engine* e, uint64_t* trace, uint64_t active)
{
    if (trace && (e->p->max_supersteps == 0 || e->superstep < e->p->max_supersteps))
        trace[e->superstep] = active;
}

/* Push, full sweep: src/combiner_postamble.h:123-199. */
static double o_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* a bounded sample: the loop ends after p->stop_after supersteps, or at the
 * first superstep boundary after p->stop_seconds of run time */
static int o_stopped(const o_engine* e)
{
    if (e->p->stop_after && e->superstep >= e->p->stop_after) return 1;
    return e->p->stop_seconds > 0.0 && e->superstep > 0 && o_now() - e->t0 >= e->p->stop_seconds;
}

static void run_push(o_engine* e, uint64_t* trace)
{
    uint64_t active = e->V; /* ip_active_vertices = V at load (iPregel_postamble.h:519) */
    while (active != 0 && !o_stopped(e)) {
        active = 0;
        for (uint64_t i = 0; i < e->V; i++) {
            o_vertex* v = &e->vx[i];
            if (v->active || v->has_message) {
                v->active = 1;
                e->res->work_vertices++;
                e->compute(e, v);
                if (v->active) active++;
            }
        }
        for (uint64_t i = 0; i < e->V; i++) {
            o_vertex* v = &e->vx[i];
            if (e->has_nxt[i]) {
                v->has_message = 1;
                v->message = e->nxt[i];
                e->has_nxt[i] = 0;
                if (!v->active) { v->active = 1; active++; }
            }
        }
        record(e, trace, active);
        e->superstep++;
    }
}

/* Pull (single broadcast): src/combiner_single_broadcast_postamble.h:125-259. */
static void run_pull(o_engine* e, uint64_t* trace)
{
    uint64_t active = e->V;
    while (active != 0 && !o_stopped(e)) {
        active = 0;
        int any_bcast = 0;
        for (uint64_t i = 0; i < e->V; i++) {
            o_vertex* v = &e->vx[i];
            e->bcf[i] = 0;
            if (v->active) {
                e->res->work_vertices++;
                e->compute(e, v);
                if (v->active) active++;
            }
            any_bcast |= e->bcf[i];
        }
        if (any_bcast) e->res->work_edges += e->in_idx[e->V];
        for (uint64_t i = 0; i < e->V; i++) {
            o_vertex* v = &e->vx[i];
            if (o_fetch(e, (uint32_t)i) && !v->active) { v->active = 1; active++; }
        }
        record(e, trace, active);
        e->superstep++;
    }
}

/* Push with frontier (spread): src/combiner_spread_postamble.h:160-431.
 * The frontier order of the reference depends on threads; it changes nothing
 * observable, so here it is the order of first delivery. */
static void run_spread(o_engine* e, uint64_t* trace)
{
    uint32_t* f = malloc(sizeof(uint32_t) * (e->V ? e->V : 1));
    uint64_t fsize = 0, active = 0;
    while ((e->superstep == 0 || active > 0) && !o_stopped(e)) {
        e->nf_size = 0;
        if (e->superstep == 0) {
            for (uint64_t i = 0; i < e->V; i++) { e->res->work_vertices++; e->compute(e, &e->vx[i]); }
        } else {
            for (uint64_t k = 0; k < fsize; k++) { e->res->work_vertices++; e->compute(e, &e->vx[f[k]]); }
        }
        active = e->nf_size;
        memcpy(f, e->nf, sizeof(uint32_t) * e->nf_size);
        fsize = e->nf_size;
        for (uint64_t k = 0; k < fsize; k++) {
            uint32_t u = f[k];
            e->vx[u].has_message = 1;
            e->vx[u].message = e->nxt[u];
            e->has_nxt[u] = 0;
        }
        record(e, trace, active);
        e->superstep++;
    }
    free(f);
}

/* Pull with broadcast targets (spread + single broadcast):
 * src/combiner_spread_single_broadcast_postamble.h ip_run. */
static void run_spread_pull(o_engine* e, uint64_t* trace)
{
    uint32_t* t = malloc(sizeof(uint32_t) * (e->V ? e->V : 1));
    uint64_t tsize = e->V;
    for (uint64_t i = 0; i < e->V; i++) t[i] = (uint32_t)i; /* targets = all V at start */
    while ((e->superstep == 0 || tsize > 0) && !o_stopped(e)) {
        int any_bcast = 0;
        for (uint64_t k = 0; k < tsize; k++) { e->res->work_vertices++; e->compute(e, &e->vx[t[k]]); }
        tsize = 0;
        for (uint64_t i = 0; i < e->V; i++) if (e->tgt[i]) t[tsize++] = (uint32_t)i;
        for (uint64_t k = 0; k < tsize; k++) {
            uint32_t v = t[k];
            if (e->tgt[v]) {
                e->res->work_edges += e->in_idx[v + 1] - e->in_idx[v];
                o_fetch(e, v);
                e->tgt[v] = 0;
            }
        }
        for (uint64_t i = 0; i < e->V; i++) { any_bcast |= e->bcf[i]; e->bcf[i] = 0; }
        (void)any_bcast;
        record(e, trace, tsize);
        e->superstep++;
    }
    free(t);
}

int ipr_oracle_run(const ipr_csr* g, int directed, int variant, int app,
                   const ipr_oracle_params* p, void* values,
                   uint64_t* active_counts, ipr_oracle_result* res)
{
    if (!g || !p || !values || !res) return -EINVAL;
    if (variant < IPR_O_PUSH || variant > IPR_O_SPREAD_PULL) return -EINVAL;
    if (app == IPR_O_SSSP_W && (!g->wgt || variant == IPR_O_PULL || variant == IPR_O_SPREAD_PULL))
        return -EINVAL;
    memset(res, 0, sizeof(*res));
    o_engine e;
    memset(&e, 0, sizeof(e));
    e.g = g; e.V = g->V; e.variant = variant; e.p = p; e.res = res;
    switch (app) {
    case IPR_O_PR:     e.compute = app_pr;     e.combine = comb_sum_f64; break;
    case IPR_O_CC:     e.compute = app_cc;     e.combine = comb_min_u32; break;
    case IPR_O_SSSP:   e.compute = app_sssp;   e.combine = comb_min_u32; break;
    case IPR_O_SSSP_W: e.compute = app_sssp_w; e.combine = comb_min_u32; break;
    default: return -EINVAL;
    }
    ipr_csr tr;
    memset(&tr, 0, sizeof(tr));
    if (directed && (variant == IPR_O_PULL || variant == IPR_O_SPREAD_PULL)) {
        /* mirrored in-lists, ascending source order (iPregel_postamble.h:429-487) */
        int rc = ipr_graph_transpose(g, &tr);
        if (rc) return rc;
        e.in_idx = tr.idx; e.in_adj = tr.adj;
    } else {
        e.in_idx = g->idx; e.in_adj = g->adj;
    }
    uint64_t V = g->V;
    e.vx = calloc(V ? V : 1, sizeof(o_vertex));
    e.nxt = calloc(V ? V : 1, sizeof(msg_t));
    e.has_nxt = calloc(V ? V : 1, 1);
    e.nf = malloc(sizeof(uint32_t) * (V ? V : 1));
    e.bcf = calloc(V ? V : 1, 1);
    e.bcm = calloc(V ? V : 1, sizeof(msg_t));
    e.tgt = calloc(V ? V : 1, 1);
    if (!e.vx || !e.nxt || !e.has_nxt || !e.nf || !e.bcf || !e.bcm || !e.tgt) {
        free(e.vx); free(e.nxt); free(e.has_nxt); free(e.nf); free(e.bcf); free(e.bcm); free(e.tgt);
        ipr_graph_free(&tr);
        return -ENOMEM;
    }
    for (uint64_t i = 0; i < V; i++) {
        o_vertex* v = &e.vx[i];
        v->id = (uint32_t)i;
        v->out_count = g->idx[i + 1] - g->idx[i];
        v->out = g->adj + g->idx[i];
        v->out_w = g->wgt ? g->wgt + g->idx[i] : NULL;
        v->active = 1;       /* ip_init_vertex_range, e.g. combiner_postamble.h:94 */
        v->has_message = 0;
    }
    if (V > 0) {
        e.t0 = o_now();
        switch (variant) {
        case IPR_O_PUSH: run_push(&e, active_counts); break;
        case IPR_O_PULL: run_pull(&e, active_counts); break;
        case IPR_O_SPREAD: run_spread(&e, active_counts); break;
        case IPR_O_SPREAD_PULL: run_spread_pull(&e, active_counts); break;
        }
    }
    res->supersteps = e.superstep;
    if (app == IPR_O_PR) {
        double* out = values;
        for (uint64_t i = 0; i < V; i++) out[i] = e.vx[i].value.f;
    } else {
        uint32_t* out = values;
        for (uint64_t i = 0; i < V; i++) out[i] = e.vx[i].value.u;
    }
    free(e.vx); free(e.nxt); free(e.has_nxt); free(e.nf); free(e.bcf); free(e.bcm); free(e.tgt);
    ipr_graph_free(&tr);
    return 0;
}

int ipr_oracle_dump(int app, uint64_t V, const void* values, const char* path)
{
    FILE* f = fopen(path, "w");
    if (!f) return -errno;
    for (uint64_t i = 0; i < V; i++) {
        if (app == IPR_O_PR || app == IPR_O_PPR)
            fprintf(f, "%lu: %0.20f\n", (unsigned long)i, ((const double*)values)[i]);
        else
            fprintf(f, "%u: %u\n", (unsigned)i, ((const uint32_t*)values)[i]);
    }
    fclose(f);
    return 0;
}

/* ---------------------------------------------------- personalised PR ---- */

int ipr_oracle_ppr(const ipr_csr* g, uint32_t seed, uint32_t iters, double* x)
{
    uint64_t V = g->V;
    if (seed >= V) return -EINVAL;
    ipr_csr t;
    int rc = ipr_graph_transpose(g, &t);
    if (rc) return rc;
    double* y = malloc(sizeof(double) * V);
    double* inv = malloc(sizeof(double) * V);
    if (!y || !inv) { free(y); free(inv); ipr_graph_free(&t); return -ENOMEM; }
    for (uint64_t v = 0; v < V; v++) {
        uint64_t d = g->idx[v + 1] - g->idx[v];
        inv[v] = d ? 1.0 / (double)d : 0.0;
        x[v] = v == seed ? 1.0 : 0.0;
    }
    for (uint32_t it = 0; it < iters; it++) {
        double dangling = 0.0;
        for (uint64_t v = 0; v < V; v++)
            if (g->idx[v + 1] == g->idx[v]) dangling += x[v];
        double total = 0.0;
        for (uint64_t v = 0; v < V; v++) {
            double s = 0.0;
            for (uint64_t k = t.idx[v]; k < t.idx[v + 1]; k++) s += x[t.adj[k]] * inv[t.adj[k]];
            double r = 0.85 * s;
            if (v == seed) r += 0.85 * dangling + 0.15;
            y[v] = r;
            total += r;
        }
        for (uint64_t v = 0; v < V; v++) x[v] = y[v] / total;
    }
    free(y); free(inv);
    ipr_graph_free(&t);
    return 0;
}

/* The same for n seeds at once, bit for bit what ipr_oracle_ppr gives for
 * each: the transpose is built once, eight seeds share a cache line per
 * vertex, and the vertices of an iteration are spread over the OpenMP
 * threads. Each y(v) is the same in-list-order sum as above; the dangling and
 * total sums stay serial in vertex order. x gets n rows of V doubles. */
int ipr_oracle_ppr_many(const ipr_csr* g, const uint32_t* seeds, uint32_t n, uint32_t iters, double* x)
{
    const uint64_t V = g->V;
    for (uint32_t j = 0; j < n; j++)
        if (seeds[j] >= V) return -EINVAL;
    ipr_csr t;
    int rc = ipr_graph_transpose(g, &t);
    if (rc) return rc;
    enum { B = 8 };
    double* xs = malloc(sizeof(double) * V * B);
    double* ys = malloc(sizeof(double) * V * B);
    double* inv = malloc(sizeof(double) * (V ? V : 1));
    if (!xs || !ys || !inv) { free(xs); free(ys); free(inv); ipr_graph_free(&t); return -ENOMEM; }
    for (uint64_t v = 0; v < V; v++) {
        uint64_t d = g->idx[v + 1] - g->idx[v];
        inv[v] = d ? 1.0 / (double)d : 0.0;
    }
    for (uint32_t j0 = 0; j0 < n; j0 += B) {
        const uint32_t nb = n - j0 < B ? n - j0 : B;
        for (uint64_t v = 0; v < V; v++)
            for (uint32_t j = 0; j < B; j++) xs[v * B + j] = j < nb && v == seeds[j0 + j] ? 1.0 : 0.0;
        for (uint32_t it = 0; it < iters; it++) {
            double dangling[B] = {0}, total[B] = {0};
            for (uint64_t v = 0; v < V; v++)
                if (g->idx[v + 1] == g->idx[v])
                    for (uint32_t j = 0; j < B; j++) dangling[j] += xs[v * B + j];
            #pragma omp parallel for schedule(dynamic, 4096)
            for (int64_t vi = 0; vi < (int64_t)V; vi++) {
                const uint64_t v = (uint64_t)vi;
                double s[B] = {0};
                for (uint64_t k = t.idx[v]; k < t.idx[v + 1]; k++) {
                    const uint32_t u = t.adj[k];
                    for (uint32_t j = 0; j < B; j++) s[j] += xs[(uint64_t)u * B + j] * inv[u];
                }
                for (uint32_t j = 0; j < B; j++) {
                    double r = 0.85 * s[j];
                    if (j < nb && v == seeds[j0 + j]) r += 0.85 * dangling[j] + 0.15;
                    ys[v * B + j] = r;
                }
            }
            for (uint64_t v = 0; v < V; v++)
                for (uint32_t j = 0; j < B; j++) total[j] += ys[v * B + j];
            for (uint64_t v = 0; v < V; v++)
                for (uint32_t j = 0; j < B; j++) xs[v * B + j] = total[j] != 0.0 ? ys[v * B + j] / total[j] : 0.0;
        }
        for (uint32_t j = 0; j < nb; j++)
            for (uint64_t v = 0; v < V; v++) x[(uint64_t)(j0 + j) * V + v] = xs[v * B + j];
    }
    free(xs); free(ys); free(inv);
    ipr_graph_free(&t);
    return 0;
}

/* ------------------------------------------------------- closed forms ---- */

int ipr_oracle_min_ancestor(const ipr_csr* g, uint32_t* label)
{
    uint64_t V = g->V;
    uint32_t* q = malloc(sizeof(uint32_t) * (V ? V : 1));
    if (!q) return -ENOMEM;
    for (uint64_t i = 0; i < V; i++) label[i] = UINT32_MAX;
    /* ascending roots: a vertex first reached from root r has min ancestor r,
       and anything reachable from an already-labelled vertex is labelled. */
    for (uint64_t r = 0; r < V; r++) {
        if (label[r] != UINT32_MAX) continue;
        uint64_t h = 0, t = 0;
        label[r] = (uint32_t)r; q[t++] = (uint32_t)r;
        while (h < t) {
            uint32_t u = q[h++];
            for (uint64_t k = g->idx[u]; k < g->idx[u + 1]; k++) {
                uint32_t w = g->adj[k];
                if (label[w] == UINT32_MAX) { label[w] = (uint32_t)r; q[t++] = w; }
            }
        }
    }
    free(q);
    return 0;
}

int ipr_oracle_bfs(const ipr_csr* g, uint32_t source, uint32_t* dist)
{
    uint64_t V = g->V;
    if (source >= V) return -EINVAL;
    uint32_t* q = malloc(sizeof(uint32_t) * V);
    if (!q) return -ENOMEM;
    for (uint64_t i = 0; i < V; i++) dist[i] = UINT32_MAX;
    uint64_t h = 0, t = 0;
    dist[source] = 0; q[t++] = source;
    while (h < t) {
        uint32_t u = q[h++];
        for (uint64_t k = g->idx[u]; k < g->idx[u + 1]; k++) {
            uint32_t w = g->adj[k];
            if (dist[w] == UINT32_MAX) { dist[w] = dist[u] + 1; q[t++] = w; }
        }
    }
    free(q);
    return 0;
}

/* binary heap of (dist, vertex) with lazy deletion */
typedef struct hp { uint64_t d; uint32_t v; } hp;
static int hp_less(hp a, hp b) { return a.d < b.d || (a.d == b.d && a.v < b.v); }
static void heap_push(hp* h, uint64_t* n, hp x)
{
    uint64_t i = (*n)++;
    while (i > 0) {
        uint64_t p = (i - 1) / 2;
        if (!hp_less(x, h[p])) break;
        h[i] = h[p]; i = p;
    }
    h[i] = x;
}
static hp heap_pop(hp* h, uint64_t* n)
{
    hp top = h[0], last = h[--(*n)];
    uint64_t i = 0;
    for (;;) {
        uint64_t l = 2 * i + 1, r = l + 1, m = i;
        hp mk = last;
        if (l < *n && hp_less(h[l], mk)) { m = l; mk = h[l]; }
        if (r < *n && hp_less(h[r], mk)) { m = r; }
        if (m == i) break;
        h[i] = h[m]; i = m;
    }
    if (*n) h[i] = last;
    return top;
}

int ipr_oracle_dijkstra(const ipr_csr* g, uint32_t source, uint32_t* dist)
{
    uint64_t V = g->V;
    if (source >= V || !g->wgt) return -EINVAL;
    hp* h = malloc(sizeof(hp) * (g->E + 2));
    if (!h) return -ENOMEM;
    uint64_t* d = malloc(sizeof(uint64_t) * V);
    if (!d) { free(h); return -ENOMEM; }
    for (uint64_t i = 0; i < V; i++) d[i] = UINT64_MAX;
    uint64_t n = 0;
    d[source] = 0;
    hp s0 = { 0, source };
    heap_push(h, &n, s0);
    while (n) {
        hp k = heap_pop(h, &n);
        uint32_t u = k.v;
        uint64_t du = k.d;
        if (du != d[u]) continue;
        for (uint64_t e = g->idx[u]; e < g->idx[u + 1]; e++) {
            uint32_t w = g->adj[e];
            uint64_t nd = du + g->wgt[e];
            if (nd < d[w]) { d[w] = nd; hp x = { nd, w }; heap_push(h, &n, x); }
        }
    }
    for (uint64_t i = 0; i < V; i++) dist[i] = d[i] >= UINT32_MAX ? UINT32_MAX : (uint32_t)d[i];
    free(d); free(h);
    return 0;
}
