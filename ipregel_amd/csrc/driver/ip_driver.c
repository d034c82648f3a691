/*
 * ip_driver.c - C host driver of the MI355X iPregel engine.
 *
 * Owns what the reference's shared runtime (src/iPregel_postamble.h) and the
 * superstep loops of its engine variants do on the host:
 *   ip_init        schedule check, provenance banner, graph load   (:176-242)
 *   loader         .config / .idx / .adj (+ .wgt)                    (:244-546)
 *   ip_run         the superstep loop; per-superstep log lines       (e.g.
 *                  src/combiner_spread_postamble.h:160-431)
 *   ip_dump        values in id order through ip_serialise_vertex    (:103-115)
 * All device work goes through libipregel_hip.so (include/ipregel_hip.h) and
 * the kernel table of the application (ipr_app_desc.launch). No superstep
 * needs a host decision: kernels read the superstep and the halting flag from
 * the device control block, and the host replays chunks of supersteps (as
 * HIP graphs) and reads the control block once per chunk.
 */
#define _GNU_SOURCE
#include "ipregel_driver.h"
#include "ipregel_hip.h"

#include <errno.h>
#include <fcntl.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <time.h>

#define IPR_TRACE_CAP (1u << 20)
#define IPR_MAX_GRAPHS 65   /* graph cache: one per chunk length 1..64 */

static struct {
    ipr_app_desc app;
    ipr_harness_cfg cfg;
    int have_cfg;
    int initialised;
    int directed, weighted;
    uint64_t V, E;
    uint64_t* h_idx;         /* V+1 */
    uint32_t* h_adj;
    uint32_t* h_wgt;
    int own_host;
    ipr_dev dev;
    void* stream;
    int own_stream;
    ipr_ctl* h_ctl;          /* pinned */
    void* ev0;
    void* ev1;
    void* graphs[IPR_MAX_GRAPHS];
    void* fgraphs[IPR_MAX_GRAPHS];   /* chunks of fused sparse supersteps */
    uint64_t wall_hz;
    ipr_run_stats stats;
    int use_graph;
    uint32_t max_chunk;
    uint32_t first_chunk;
    /* the launches the previous run of this graph needed, and whether it ran
     * fused sparse supersteps: a run that repeats one without them starts
     * with a chunk of exactly that many launches (no launches past the
     * halting superstep, one host check) */
    int adapt_chunk;
    uint32_t prev_launches;
    int prev_fused;
    /* the broadcasters' bitmap kernel is launched while runs use it
     * (IPR_GATHER_LIVE_LAUNCH: 2 adaptive, the default; 1 always; 0 never):
     * a run that built no bitmap and missed none turns it off, a gather that
     * wanted one and found none turns it back on; a gather without the
     * bitmap reads every record, with the same result */
    int live_launch;
    uint32_t live_mode;
    /* launch plan (push apps, IPR_LAUNCH_PLAN=1 by default): a run that
     * repeats the previous one (same launches, no fused sparse supersteps)
     * launches in each of its first 64 supersteps only the delivery kernel
     * that superstep used, gather or expansion. The kernel launched alone
     * stops the run if the superstep needs the other one, and the run is
     * done again without the plan. The plan is a chunk starting at superstep
     * 0, captured in a graph of its own; it needs the receiver-blocked
     * gather (the arc-balanced one has no room for the check). */
    uint32_t plan_mode;
    int plan_on;
    uint64_t plan_mask;
    uint32_t plan_steps;
    int planning;            /* the chunk being enqueued follows the plan */
    uint32_t chunk_no;       /* chunks enqueued so far in this run */
    void* plan_graph;
    int quiet;
    int kev_on;              /* events around the dominant kernel */
    void* kev[3][64];        /* start / split / stop per superstep slot of a chunk */
    uint32_t kev_slot;       /* slot of the superstep being enqueued */
    double* kev_ms[2];       /* per superstep of the last timed run: before / after the split */
    uint64_t kev_n;          /* supersteps in kev_ms */
    /* phase profile (IP_ENABLE_THREAD_PROFILING or IPR_PHASE_PROFILE=1):
     * events before compute, after compute, after delivery, after merging */
    int prof_on;
    void* pev[4][64];
    double* prof_ms[3];      /* per superstep: compute, delivery (mailbox), merging */
} D;

/* ------------------------------------------------------------------------ */
/* helpers                                                                   */
/* ------------------------------------------------------------------------ */

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void die(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    fflush(stdout);
    exit(-1);
}

static void hip_check(int rc, const char* what)
{
    if (rc) die("HIP failure in %s: %s\n", what, ipr_hip_last_error());
}

static void* safe_malloc(size_t n)
{
    void* p = malloc(n ? n : 1);
    if (!p) die("Failed to allocate %zu bytes.\n", n);
    return p;
}

static FILE* safe_fopen(const char* path, const char* mode)
{
    FILE* f = fopen(path, mode);
    if (!f) die("Cannot open the file \"%s\" in mode \"%s\".\n", path, mode);
    return f;
}

/* count elements of size bytes from the start of a file, read in 64 MiB
 * pieces by all threads (positional reads; the reference likewise reads its
 * offset file in per-thread chunks, src/iPregel_postamble.h:295-321) */
static void read_elements(const char* path, void* dst, size_t size, uint64_t count)
{
    int fd = open(path, O_RDONLY);
    if (fd < 0) die("Cannot open the file \"%s\" in mode \"%s\".\n", path, "rb");
    const uint64_t bytes = (uint64_t)size * count, piece = 64ull << 20;
    const int64_t pieces = (int64_t)((bytes + piece - 1) / piece);
    int bad = 0;
    #pragma omp parallel for schedule(dynamic, 1) reduction(|:bad)
    for (int64_t i = 0; i < pieces; i++) {
        const uint64_t off = (uint64_t)i * piece;
        const uint64_t len = bytes - off < piece ? bytes - off : piece;
        uint64_t done = 0;
        while (done < len) {
            const ssize_t r = pread(fd, (char*)dst + off + done, len - done, (off_t)(off + done));
            if (r <= 0) { bad = 1; break; }
            done += (uint64_t)r;
        }
    }
    close(fd);
    if (bad) die("Failed to read %zu elements of %zu bytes each.\n", (size_t)count, size);
}

static char* with_ext(const char* root, const char* ext)
{
    size_t n = strlen(root) + strlen(ext) + 1;
    char* p = safe_malloc(n);
    snprintf(p, n, "%s%s", root, ext);
    return p;
}

static uint32_t env_u32(const char* name, uint32_t def)
{
    const char* v = getenv(name);
    if (!v || !*v) return def;
    return (uint32_t)strtoul(v, NULL, 10);
}

/* ------------------------------------------------------------------------ */
/* device state                                                              */
/* ------------------------------------------------------------------------ */

static void dev_free_all(void)
{
    for (int i = 0; i < IPR_MAX_GRAPHS; i++) {
        if (D.graphs[i]) ipr_hip_graph_destroy(D.graphs[i]);
        if (D.fgraphs[i]) ipr_hip_graph_destroy(D.fgraphs[i]);
        D.graphs[i] = D.fgraphs[i] = NULL;
    }
    if (D.plan_graph) ipr_hip_graph_destroy(D.plan_graph);
    D.plan_graph = NULL;
    ipr_dev* d = &D.dev;
    if (d->in_ids && d->in_ids != d->in_adj && d->in_ids != d->out_ids) ipr_hip_free((void*)d->in_ids);
    if (d->out_ids && d->out_ids != d->out_adj) ipr_hip_free((void*)d->out_ids);
    if (d->t_idx && d->t_idx != d->out_idx) ipr_hip_free((void*)d->t_idx);
    if (d->t_adj && d->t_adj != d->out_adj) ipr_hip_free((void*)d->t_adj);
    if (d->t_wgt && d->t_wgt != d->out_wgt) ipr_hip_free((void*)d->t_wgt);
    ipr_hip_free((void*)d->g_arc);
    ipr_hip_free((void*)d->g_own);
    ipr_hip_free((void*)d->g_dst);
    ipr_hip_free((void*)d->g_wgt);
    ipr_hip_free((void*)d->g_tq);
    ipr_hip_free((void*)d->g_thead);
    ipr_hip_free((void*)d->g_tesc);
    ipr_hip_free((void*)d->g_twgt);
    ipr_hip_free((void*)d->g_tblk);
    ipr_hip_free((void*)d->bpos);
    ipr_hip_free((void*)d->bvert);
    ipr_hip_free(d->g_live);
    ipr_hip_free(d->brec4);
    ipr_hip_free(d->bslot);
    ipr_hip_free(d->lbase);
    if (d->in_idx && d->in_idx != d->out_idx) ipr_hip_free((void*)d->in_idx);
    if (d->in_adj && d->in_adj != d->out_adj) ipr_hip_free((void*)d->in_adj);
    if (d->in_wgt && d->in_wgt != d->out_wgt) ipr_hip_free((void*)d->in_wgt);
    ipr_hip_free((void*)d->out_idx);
    ipr_hip_free((void*)d->out_adj);
    ipr_hip_free((void*)d->out_wgt);
    ipr_hip_free(d->val);
    ipr_hip_free(d->msg);
    ipr_hip_free(d->act);
    ipr_hip_free(d->hm);
    for (int i = 0; i < 2; i++) {
        ipr_hip_free(d->bcm[i]);
        ipr_hip_free(d->mbox[i]);
        ipr_hip_free(d->recv[i]);
        ipr_hip_free(d->front[i]);
    }
    for (int b = 0; b < IPR_NBINS; b++) ipr_hip_free(d->pbins[b]);
    ipr_hip_free(d->lbins);
    ipr_hip_free(d->cnt);
    ipr_hip_free(d->trace_active);
    ipr_hip_free(d->trace_ts);
    ipr_hip_free(d->trace_edges);
    ipr_hip_free(d->ctl);
    memset(d, 0, sizeof(*d));
    if (D.h_ctl) ipr_hip_host_free(D.h_ctl);
    D.h_ctl = NULL;
    for (int i = 0; i < 64; i++) {
        for (int k = 0; k < 3; k++) {
            if (D.kev[k][i]) ipr_hip_event_destroy(D.kev[k][i]);
            D.kev[k][i] = NULL;
        }
    }
    for (int k = 0; k < 2; k++) {
        free(D.kev_ms[k]);
        D.kev_ms[k] = NULL;
    }
    D.kev_n = 0;
    D.kev_on = 0;
    for (int i = 0; i < 64; i++) {
        for (int k = 0; k < 4; k++) {
            if (D.pev[k][i]) ipr_hip_event_destroy(D.pev[k][i]);
            D.pev[k][i] = NULL;
        }
    }
    for (int k = 0; k < 3; k++) {
        free(D.prof_ms[k]);
        D.prof_ms[k] = NULL;
    }
    D.prof_on = 0;
    if (D.ev0) ipr_hip_event_destroy(D.ev0);
    if (D.ev1) ipr_hip_event_destroy(D.ev1);
    D.ev0 = D.ev1 = NULL;
    if (D.own_stream && D.stream) ipr_hip_stream_destroy(D.stream);
    D.stream = NULL;
    D.own_stream = 0;
}

static void* dalloc(size_t bytes, const char* what)
{
    void* p = NULL;
    hip_check(ipr_hip_malloc(&p, bytes), what);
    return p;
}

static void dev_setup(void)
{
    ipr_dev* d = &D.dev;
    const ipr_app_desc* a = &D.app;
    uint64_t V = D.V, E = D.E;
    void* st = D.stream;
    memset(d, 0, sizeof(*d));
    d->V = V;
    d->E = E;
    uint64_t vv = V ? V : 1;

    d->out_idx = dalloc((V + 1) * 8, "out_idx");
    d->out_adj = dalloc(E * 4, "out_adj");
    hip_check(ipr_hip_memcpy_h2d((void*)d->out_idx, D.h_idx, (V + 1) * 8, st), "upload idx");
    hip_check(ipr_hip_memcpy_h2d((void*)d->out_adj, D.h_adj, E * 4, st), "upload adj");
    if (D.h_wgt) {
        d->out_wgt = dalloc(E * 4, "out_wgt");
        hip_check(ipr_hip_memcpy_h2d((void*)d->out_wgt, D.h_wgt, E * 4, st), "upload wgt");
    }
    /* in-lists: aliased when undirected (src/iPregel_postamble.h:368-396),
     * mirrored in ascending source order when directed (:429-487) */
    if (a->needs_in_lists && D.directed) {
        d->in_idx = dalloc((V + 1) * 8, "in_idx");
        d->in_adj = dalloc(E * 4, "in_adj");
        if (d->out_wgt) d->in_wgt = dalloc(E * 4, "in_wgt");
        hip_check(ipr_hip_build_csc(V, E, d->out_idx, d->out_adj, d->out_wgt, (uint64_t*)d->in_idx,
                                    (uint32_t*)d->in_adj, (uint32_t*)d->in_wgt, st), "build CSC");
    } else {
        d->in_idx = d->out_idx;
        d->in_adj = d->out_adj;
        d->in_wgt = d->out_wgt;
    }
    /* neighbour ids as the application reads them (IP_VERTEX_ID_TYPE): the
     * 32-bit arrays in _32 builds, widened copies in _64 builds */
    if (a->needs_out_ids) {
        if (a->id_size == 8) {
            uint64_t* w = dalloc(E * 8, "out-neighbour ids");
            hip_check(ipr_hip_widen_ids(d->out_adj, w, E, st), "widen out-neighbour ids");
            d->out_ids = w;
        } else {
            d->out_ids = d->out_adj;
        }
    }
    if (a->needs_in_ids) {
        if (d->in_adj == d->out_adj && d->out_ids) {
            d->in_ids = d->out_ids;
        } else if (a->id_size == 8) {
            uint64_t* w = dalloc(E * 8, "in-neighbour ids");
            hip_check(ipr_hip_widen_ids(d->in_adj, w, E, st), "widen in-neighbour ids");
            d->in_ids = w;
        } else {
            d->in_ids = d->in_adj;
        }
    }

    d->val = dalloc(vv * a->val_size, "values");
    d->msg = dalloc(vv * a->msg_size, "messages");
    d->act = dalloc(vv, "active");
    d->hm = dalloc(vv, "has_message");
    /* PULL: broadcast slots by parity. PUSH / SPREAD: mailboxes, received
     * flags, broadcast records. SPREAD_PULL: target flags and broadcast
     * records (the fetch recognises this superstep's records by epoch). */
    const int var = a->variant;
    const int pull = var == IPR_VARIANT_PULL;
    const int records = !pull;
    for (int i = 0; i < 2; i++) {
        if (pull) {
            d->bcm[i] = dalloc(vv * a->bslot_size, "broadcast records");
        } else {
            if (var != IPR_VARIANT_SPREAD_PULL) d->mbox[i] = dalloc(vv * a->mbox_size, "mailboxes");
            d->recv[i] = dalloc(((vv + 15) / 16) * 16 + 16, "received flags");
            hip_check(ipr_hip_fill(d->recv[i], 0, ((vv + 15) / 16) * 16 + 16, st), "flags init");
        }
        if (var == IPR_VARIANT_SPREAD || var == IPR_VARIANT_SPREAD_PULL)
            d->front[i] = dalloc(vv * 4, "frontier");
    }
    if (records) d->bslot = dalloc(vv * a->bslot_size, "broadcast records");
    d->ctl = dalloc(sizeof(ipr_ctl), "control block");
    d->cnt = dalloc(IPR_NCNT * IPR_CSLOTS * IPR_CPAD * 8, "counters");
    hip_check(ipr_hip_fill(d->cnt, 0, IPR_NCNT * IPR_CSLOTS * IPR_CPAD * 8, st), "counters init");
    d->trace_active = dalloc(IPR_TRACE_CAP * 8, "trace");
    d->trace_ts = dalloc((IPR_TRACE_CAP + 1) * 8, "trace");
    d->trace_edges = dalloc(IPR_TRACE_CAP * 8, "trace");
    hip_check(ipr_hip_fill(d->ctl, 0, sizeof(ipr_ctl), st), "ctl init");
    uint64_t cap = IPR_TRACE_CAP;
    hip_check(ipr_hip_memcpy_h2d(&d->ctl->trace_cap, &cap, 8, st), "ctl init");
    hip_check(ipr_hip_stream_sync(st), "ctl init");

    /* tuning */
    const ipr_harness_cfg* c = D.have_cfg ? &D.cfg : NULL;
    d->large_min = c && c->bin_medium_max ? c->bin_medium_max : env_u32("IPR_LARGE_MIN", 2048);
    if (records) {
        /* broadcasters above large_min out-arcs register in one of IPR_CSLOTS
         * slots; each slot has room for all of them (a vertex computes at most
         * once per superstep) */
        uint64_t nl = 0;
        for (uint64_t v = 0; v < V; v++) nl += D.h_idx[v + 1] - D.h_idx[v] > d->large_min;
        d->lcap = nl;
        if (nl) {
            d->lbins = dalloc(nl * IPR_CSLOTS * 4, "large broadcasters");
            d->lbase = dalloc(nl * IPR_CSLOTS * 8, "large-bin chunk bases");
        }
    }
    d->grid_blocks = c && c->grid_blocks ? c->grid_blocks : env_u32("IPR_GRID_BLOCKS", 2048);
    {
        uint32_t cus = 256;
        hip_check(ipr_hip_cu_count(&cus), "CU count");
        d->pull_wave_blocks = env_u32("IPR_PULL_WAVE_BLOCKS", 3 * cus);
    }
    uint32_t po = c && c->pull_ordered ? (uint32_t)c->pull_ordered : env_u32("IPR_PULL_ORDERED", 1);
    d->pull_ordered = po == 1;
    uint32_t dopt = c && c->dir_opt ? (uint32_t)c->dir_opt : env_u32("IPR_DIR_OPT", 1);
    /* the gather index marks list ends in bit 31 of the source ids */
    d->dir_opt = dopt == 1 && a->can_gather && (var == IPR_VARIANT_PUSH || var == IPR_VARIANT_SPREAD) &&
                 V < (1ull << 31);
    d->dir_alpha = c && c->dir_alpha ? c->dir_alpha : env_u32("IPR_DIR_ALPHA", 6);
    d->large_chunk = c && c->bin_small_max ? c->bin_small_max : env_u32("IPR_LARGE_CHUNK", 2048);
    if (d->large_chunk < 64) d->large_chunk = 64;
    if (!pull && d->dir_opt) {
        /* dense supersteps gather over the exact transpose; a symmetric graph
         * is its own transpose and keeps one copy */
        uint64_t* t_idx = dalloc((V + 1) * 8, "transpose idx");
        uint32_t* t_adj = dalloc(E * 4, "transpose adj");
        uint32_t* t_wgt = d->out_wgt ? dalloc(E * 4, "transpose wgt") : NULL;
        hip_check(ipr_hip_build_csc(V, E, d->out_idx, d->out_adj, d->out_wgt, t_idx, t_adj, t_wgt, st),
                  "build transpose");
        int same_idx = 0, same_adj = 0, same_w = 1;
        hip_check(ipr_hip_equal(t_idx, d->out_idx, (V + 1) * 8, &same_idx, st), "transpose check");
        if (same_idx) hip_check(ipr_hip_equal(t_adj, d->out_adj, E * 4, &same_adj, st), "transpose check");
        if (same_idx && same_adj && t_wgt)
            hip_check(ipr_hip_equal(t_wgt, d->out_wgt, E * 4, &same_w, st), "transpose check");
        if (same_idx && same_adj && same_w) {
            ipr_hip_free(t_idx);
            ipr_hip_free(t_adj);
            ipr_hip_free(t_wgt);
            d->t_idx = d->out_idx;
            d->t_adj = d->out_adj;
            d->t_wgt = d->out_wgt;
        } else {
            d->t_idx = t_idx;
            d->t_adj = t_adj;
            d->t_wgt = t_wgt;
        }
        const uint64_t K = IPR_GATHER_K, pass = 64 * K;
        /* split by record line, one class per XCD (IPR_GATHER_XCD=1): each
         * class's arcs start on a whole chunk, and its receivers get their own
         * g_dst entries */
        d->g_xcd = env_u32("IPR_GATHER_XCD", 0) == 1;
        /* receiver-blocked tiles (IPR_GATHER_TILED=1, the default where the
         * combiner folds by LDS atomics) replace the arc-balanced index */
        const int tiled = !d->g_xcd && a->can_tile && env_u32("IPR_GATHER_TILED", 1) == 1;
        const uint64_t ea = d->g_xcd ? E + IPR_GATHER_CLASSES * K : E;
        const uint64_t nd = d->g_xcd ? (E < IPR_GATHER_CLASSES * vv ? E : IPR_GATHER_CLASSES * vv) : vv;
        uint64_t nch = (ea + K - 1) / K, ndst = 0;
        uint32_t* g_arc = tiled ? NULL : dalloc(ea * 4, "gather sources");
        uint32_t* g_own = tiled ? NULL : dalloc(nch * 4, "gather chunk receivers");
        uint32_t* g_dst = tiled ? NULL : dalloc((nd ? nd : 1) * 4, "gather receivers");
        /* SPREAD reads records only for its frontier outside the gather, so
         * reordering them by out-degree costs little there (PUSH scans all V) */
        if (var == IPR_VARIANT_SPREAD && env_u32("IPR_RECORD_ORDER", 1) == 1) {
            uint32_t* bpos = dalloc(vv * 4, "record positions");
            uint32_t* bvert = dalloc(vv * 4, "record vertices");
            hip_check(ipr_hip_build_record_order(V, d->out_idx, bpos, bvert, st), "record order");
            d->bpos = bpos;
            d->bvert = bvert;
        }
        if (tiled) {
            /* the resident workgroups the LDS allows (one per CU with 128 KB
             * of receiver slots); tiles of at most tile_receivers receivers
             * and about one workgroup's share of the arcs: larger tiles share
             * more record lines between their arcs (measured with two 64 KB
             * workgroups per CU: CC R-MAT 22 gather 0.36 ms with a quarter
             * share, 0.28 ms with a whole one; CC R-MAT 24 1.52 -> 1.29 ms),
             * and the blocks that the receiver limit keeps smaller fill in
             * behind them, largest first */
            uint32_t cus = 1;
            hip_check(ipr_hip_cu_count(&cus), "compute units");
            d->tile_blocks = env_u32("IPR_TILE_BLOCKS", cus * (a->tile_wgs_per_cu ? a->tile_wgs_per_cu : 1));
            if (d->tile_blocks < 1) d->tile_blocks = 1;
            uint32_t max_recv = env_u32("IPR_TILE_RECV", a->tile_receivers);
            if (max_recv < 32 || max_recv > a->tile_receivers) max_recv = a->tile_receivers;
            uint64_t max_arcs = env_u32("IPR_TILE_ARCS", 0);
            if (max_arcs == 0) {
                max_arcs = E / (uint64_t)d->tile_blocks;
                if (max_arcs < 65536) max_arcs = 65536;
            }
            ipr_gather_tiles gt;
            hip_check(ipr_hip_build_gather_tiles(V, E, d->t_idx, d->t_adj, d->t_wgt, d->bpos, max_recv, max_arcs,
                                                 &gt, st),
                      "gather tiles");
            d->g_tq = gt.q;
            d->g_thead = gt.head;
            d->g_tesc = gt.esc;
            d->g_twgt = gt.wgt;
            d->g_twsz = gt.wgt_bytes;
            d->g_tblk = (const ipr_tblock*)gt.tiles;
            d->g_tn = gt.n_tiles;
        } else if (d->g_xcd) {
            uint32_t* g_wgt = d->t_wgt ? dalloc(ea * 4, "gather weights") : NULL;
            hip_check(ipr_hip_build_gather_classes(V, E, d->t_idx, d->t_adj, d->t_wgt, d->bpos, IPR_GATHER_K, g_arc,
                                                   g_wgt, g_own, g_dst, d->g_cbeg, d->g_cend, d->g_cdst, st),
                      "gather class index");
            d->g_wgt = g_wgt;
        }
        else
            hip_check(ipr_hip_build_gather(V, E, d->t_idx, d->t_adj, d->bpos, IPR_GATHER_K, g_arc, g_own, g_dst,
                                           &ndst, st),
                      "gather index");
        d->g_arc = g_arc;
        d->g_own = g_own;
        d->g_dst = g_dst;
        /* broadcasters' bitmap for gathers with few live arcs (below this
         * percentage of E) from few computed vertices (below live_vpct % of
         * V): unit SSSP R-MAT 24 superstep 1 (238K vertices, 52% of the arcs)
         * 0.54 -> 0.40 ms; with many computed vertices the skipped records
         * are cheap ones (CC R-MAT 22 superstep 2, 34% live: 2-4% slower) */
        d->live_pct = env_u32("IPR_GATHER_LIVE", 60);
        d->live_vpct = env_u32("IPR_GATHER_LIVE_VERTICES", 5);
        if (d->live_pct > 0) d->g_live = dalloc(((vv + 63) / 64) * 8, "gather live bitmap");
        /* 4-byte records for the gather (0 turns them off); padded to 16
         * entries for the window clear of the frontier compaction, cleared by
         * the init kernel */
        if (env_u32("IPR_GATHER_RECORD4", 1) == 1)
            d->brec4 = dalloc(((vv + 15) / 16) * 64, "gather records");
        /* the arcs split into ranges of about 2048 arcs (whole passes; four
         * passes keep the range counters' atomics rare at any E while the
         * ranges still balance); the wavefronts of one resident round of
         * workgroups (5 per CU: 32 KB of LDS and 96 VGPRs each) take them in
         * turn */
        uint32_t cus = 1;
        hip_check(ipr_hip_cu_count(&cus), "compute units");
        d->gather_blocks = env_u32("IPR_GATHER_BLOCKS", cus * 5);
        if (d->gather_blocks < 1) d->gather_blocks = 1;
        uint64_t per = env_u32("IPR_GATHER_RANGE_ARCS", 2048);
        per = (per + pass - 1) / pass * pass;
        d->g_per_wave = per ? per : pass;
    }
    if (pull) {
        /* ordered folds are sequential per vertex: keep the thread-per-vertex
         * bin short; reassociating folds use the lane-cooperative one */
        uint32_t ps = c && c->pull_small_max ? c->pull_small_max
                                             : env_u32("IPR_PULL_SMALL", d->pull_ordered ? 4 : 32);
        uint32_t pm = c && c->pull_medium_max ? c->pull_medium_max : env_u32("IPR_PULL_MEDIUM", 2048);
        for (int b = 0; b < IPR_NBINS; b++) d->pbins[b] = dalloc(vv * 4, "pull bins");
        hip_check(ipr_hip_build_pull_bins(d, ps, pm, st), "pull bins");
    }
    /* frontier compaction rounds per workgroup (0: picked from V) */
    d->compact_rounds = env_u32("IPR_COMPACT_ROUNDS", 0);
    /* spread: a sparse superstep is one launch (frontier scan, compute,
     * expansion; its last workgroup ends the superstep); a frontier vertex
     * above large_min out-arcs keeps the superstep on the general kernels,
     * which split it in chunks */
    d->fused = var == IPR_VARIANT_SPREAD && env_u32("IPR_FUSED", 1) == 1;
    d->fused_blocks = env_u32("IPR_FUSED_BLOCKS", 1536);
    if (d->fused_blocks < 1) d->fused_blocks = 1;
    d->fused_max_deg = env_u32("IPR_FUSED_MAX_DEG", d->large_min);
    {
        const char* fa = getenv("IPR_FUSED_MAX_ARCS");
        d->fused_max_arcs = fa && *fa ? strtoull(fa, NULL, 10) : ~0ull;
    }
    /* pull and spread: the last workgroup of the fused pull kernel or of the
     * frontier compaction ends the superstep (no finalize launch) */
    d->fin_in_kernel = (pull || var == IPR_VARIANT_SPREAD) && env_u32("IPR_FIN_IN_KERNEL", 1) == 1;
    /* the compaction's last workgroup must end the superstep to mark the
     * frontier as flags */
    d->dense_flags = var == IPR_VARIANT_SPREAD && d->fin_in_kernel && env_u32("IPR_DENSE_FLAGS", 1) == 1;
    uint32_t ug = c && c->use_graph ? (uint32_t)c->use_graph : env_u32("IPR_GRAPH", 1);
    D.use_graph = ug == 1;
    D.live_mode = env_u32("IPR_GATHER_LIVE_LAUNCH", 2);
    D.live_launch = D.live_mode != 0;
    D.plan_mode = env_u32("IPR_LAUNCH_PLAN", 1);
    D.plan_on = 0;
    D.max_chunk = c && c->max_chunk ? c->max_chunk : env_u32("IPR_MAX_CHUNK", 64);
    /* a power of two; the pull kernel's no-op launches are cheap */
    {
        uint32_t fc = env_u32("IPR_FIRST_CHUNK", pull ? 16 : 1), p2 = 1;
        while (p2 < fc && p2 < 64) p2 *= 2;
        D.first_chunk = p2;
    }
    if (D.max_chunk < 1) D.max_chunk = 1;
    D.adapt_chunk = env_u32("IPR_ADAPT_CHUNK", 1) == 1;
    D.prev_launches = 0;
    D.prev_fused = 0;
    hip_check(ipr_hip_host_alloc((void**)&D.h_ctl, sizeof(ipr_ctl)), "pinned ctl");
    hip_check(ipr_hip_event_create(&D.ev0), "event");
    hip_check(ipr_hip_event_create(&D.ev1), "event");
    if (D.max_chunk > 64) D.max_chunk = 64;
    D.kev_on = 0;
    if (c && c->kernel_events) ipr_harness_kernel_events(1);
    D.prof_on = a->thread_profiling || env_u32("IPR_PHASE_PROFILE", 0) == 1;
    if (D.prof_on) {
        for (int i = 0; i < 64; i++)
            for (int k = 0; k < 4; k++) hip_check(ipr_hip_event_create(&D.pev[k][i]), "profile event");
        for (int k = 0; k < 3; k++) {
            D.prof_ms[k] = (double*)safe_malloc(IPR_TRACE_CAP * sizeof(double));
            memset(D.prof_ms[k], 0, IPR_TRACE_CAP * sizeof(double));
        }
    }
    int rc = a->bind(d);
    if (rc) die("HIP failure in binding device state (%d)\n", rc);
}

/* ------------------------------------------------------------------------ */
/* loader                                                                    */
/* ------------------------------------------------------------------------ */

/* .idx holds V offsets of IP_NEIGHBOUR_COUNT_TYPE (src/iPregel_postamble.h:
 * 294-323 reads sizeof(IP_NEIGHBOUR_COUNT_TYPE) per entry; the reference's
 * applications use 8 bytes, benchmarks/template.c 4): widened to 64 bits */
static void load_from_files(const char* root, int want_weights, uint32_t id_size, uint32_t idx_size)
{
    char* p = with_ext(root, ".config");
    printf("\t- Loading configuration file from: \"%s\".\n", p);
    FILE* f = safe_fopen(p, "r");
    free(p);
    unsigned long long V = 0, E = 0;
    if (fscanf(f, "%llu %llu", &V, &E) != 2) {
        printf("\t- Failure in reading the number of vertices and edges. Abort...\n");
        fclose(f);
        fflush(stdout);
        exit(-1);
    }
    fclose(f);
    printf("\t\t- %llu vertices\n\t\t- %llu edges\n", V, E);
    D.V = V;
    D.E = E;
    D.h_idx = safe_malloc((V + 1) * sizeof(uint64_t));
    D.h_adj = safe_malloc((E ? E : 1) * sizeof(uint32_t));
    p = with_ext(root, ".idx");
    printf("\t- Loading offset file from: \"%s\".\n", p);
    if (idx_size == 4) {
        uint32_t* narrow = safe_malloc((V ? V : 1) * sizeof(uint32_t));
        read_elements(p, narrow, sizeof(uint32_t), V);
        #pragma omp parallel for schedule(static)
        for (uint64_t v = 0; v < V; v++) D.h_idx[v] = narrow[v];
        free(narrow);
    } else {
        read_elements(p, D.h_idx, sizeof(uint64_t), V);
    }
    free(p);
    D.h_idx[V] = E;
    /* .adj holds IP_VERTEX_ID_TYPE ids: 4 bytes in _32 builds, 8 in _64 ones;
     * the device keeps 32-bit ids (validate_graph bounds them) */
    p = with_ext(root, ".adj");
    printf("\t- Loading adjacency file from: \"%s\".\n", p);
    if (id_size == 8) {
        uint64_t* wide = safe_malloc((E ? E : 1) * sizeof(uint64_t));
        read_elements(p, wide, sizeof(uint64_t), E);
        int big = 0;
        #pragma omp parallel for schedule(static) reduction(|:big)
        for (uint64_t e = 0; e < E; e++) {
            big |= wide[e] > 0xffffffffull;
            D.h_adj[e] = (uint32_t)wide[e];
        }
        free(wide);
        if (big) die("Invalid graph: an arc points outside [0, V).\n");
    } else {
        read_elements(p, D.h_adj, sizeof(uint32_t), E);
    }
    free(p);
    if (want_weights) {
        p = with_ext(root, ".wgt");
        printf("\t- Loading weight file from: \"%s\".\n", p);
        D.h_wgt = safe_malloc((E ? E : 1) * sizeof(uint32_t));
        read_elements(p, D.h_wgt, sizeof(uint32_t), E);
        free(p);
    }
    D.own_host = 1;
}

/* The reference trusts its input; kernels index with these values, so they
 * are checked before anything reaches the device. */
static void validate_graph(void)
{
    uint64_t V = D.V, E = D.E;
    if (V > 0x7fffffffull) die("Graph has %llu vertices; this build supports ids below 2^31.\n",
                               (unsigned long long)V);
    if (V && D.h_idx[0] != 0) die("Invalid graph: first offset is %llu, expected 0.\n",
                                  (unsigned long long)D.h_idx[0]);
    int bad = 0;
    for (uint64_t v = 0; v < V; v++)
        if (D.h_idx[v + 1] < D.h_idx[v]) { bad = 1; break; }
    if (bad || (V && D.h_idx[V] != E)) die("Invalid graph: offsets are not non-decreasing up to E.\n");
    uint64_t badarc = 0;
    #pragma omp parallel for reduction(|:badarc) schedule(static)
    for (uint64_t e = 0; e < E; e++) badarc |= D.h_adj[e] >= V;
    if (badarc) die("Invalid graph: an arc points outside [0, V).\n");
}

static void host_free(void)
{
    if (D.own_host) {
        free(D.h_idx);
        free(D.h_adj);
        free(D.h_wgt);
    }
    D.h_idx = NULL;
    D.h_adj = NULL;
    D.h_wgt = NULL;
    D.own_host = 0;
}

/* ------------------------------------------------------------------------ */
/* ip_init                                                                   */
/* ------------------------------------------------------------------------ */

void ipr_drv_init(const ipr_app_desc* app, const char* file_path, int number_of_threads,
                  const char* schedule, int chunk_size, bool directed, bool weighted)
{
    /* src/iPregel_postamble.h:176-193: only "static" and "dynamic" exist */
    if (strcmp(schedule, "static") == 0)
        printf("Runtime schedule set to static(%d).\n", chunk_size);
    else if (strcmp(schedule, "dynamic") == 0)
        printf("Runtime schedule set to dynamic(%d).\n", chunk_size);
    else
        die("The schedule %s is unknown.\n", schedule);

    if (D.initialised) {
        dev_free_all();
        host_free();
        D.initialised = 0;
    }
    D.app = *app;
    printf("Version:%s\n", app->version);
    printf("Software:iPregel\n");
    printf("FileCommits:%s\n", app->commits);
    printf("Application:%s\n", app->application);
    printf("MpiProcessCount:0\n");
    printf("Machine:%s\n", app->machine);
    printf("CompilationFlags:%s\n", app->compilation_flags);
    time_t t = time(NULL);
    struct tm tm = *localtime(&t);
    printf("SubmissionDate:%02d/%02d/%d\n", tm.tm_mday, tm.tm_mon + 1, tm.tm_year + 1900);
    printf("SubmissionTime:%02d:%02d:%02d\n", tm.tm_hour, tm.tm_min, tm.tm_sec);
    double t0 = now_s();
    const char* name = file_path;
    for (const char* q = strchr(name, '/'); q; q = strchr(name, '/')) name = q + 1;
    printf("Graph:%s\n", name);
    printf("OpenmpThreadCount:%d\n", number_of_threads);

    D.directed = directed;
    D.weighted = weighted;
    double l0 = now_s();
    printf("[INFO] The application indicates that the graph must be %sdirected and %sweighted, "
           "so the graph passed is expected to be so.\n", directed ? "" : "un", weighted ? "" : "un");
    printf("[INFO] Starting graph loading.\n");
    int want_w = app->needs_weights;
    if (D.have_cfg && D.cfg.idx) {
        D.V = D.cfg.V;
        D.E = D.cfg.E;
        D.h_idx = (uint64_t*)D.cfg.idx;
        D.h_adj = (uint32_t*)D.cfg.adj;
        D.h_wgt = want_w ? (uint32_t*)D.cfg.wgt : NULL;
        D.own_host = 0;
        if (want_w && !D.h_wgt) die("The application needs edge weights but none were given.\n");
        printf("\t\t- %llu vertices\n\t\t- %llu edges\n", (unsigned long long)D.V, (unsigned long long)D.E);
    } else {
        load_from_files(file_path, want_w, D.app.id_size, D.app.idx_size ? D.app.idx_size : 8);
    }
    validate_graph();

    int device = D.have_cfg ? D.cfg.device : (int)env_u32("IPR_DEVICE", 0);
    char devname[128] = {0};
    uint64_t hbm = 0;
    hip_check(ipr_hip_init(device, devname, sizeof(devname), &hbm), "device init");
    printf("Device:%s\n", devname);
    snprintf(D.stats.device_name, sizeof(D.stats.device_name), "%s", devname);
    hip_check(ipr_hip_wallclock_hz(&D.wall_hz), "wall clock rate");
    if (D.have_cfg && D.cfg.stream) {
        D.stream = D.cfg.stream;
        D.own_stream = 0;
    } else {
        hip_check(ipr_hip_stream_create(&D.stream), "stream");
        D.own_stream = 1;
    }
    dev_setup();
    int cc = app->check_combiner();
    if (cc) die("The combiner declared on the compile line does not match ip_combine.\n");
    double l1 = now_s();
    D.stats.load_s = l1 - l0;
    printf("LoadingTime:%f\n", l1 - l0);
    D.initialised = 1;
    D.quiet = D.have_cfg ? D.cfg.quiet : 0;
    memset(&D.stats.supersteps, 0, sizeof(D.stats) - offsetof(ipr_run_stats, supersteps));
    D.stats.load_s = l1 - l0;
    snprintf(D.stats.device_name, sizeof(D.stats.device_name), "%s", devname);
    D.stats.V = D.V;
    D.stats.E = D.E;
    printf("InitialisationTime:%f\n", now_s() - t0);
    fflush(stdout);
}

uint64_t ipr_drv_vertices_count(void) { return D.V; }
uint64_t ipr_drv_edges_count(void) { return D.E; }
uint64_t ipr_drv_superstep(void) { return D.stats.supersteps; }
uint64_t ipr_drv_runs(void) { return D.stats.runs; }
int ipr_drv_directed(void) { return D.directed; }
const uint64_t* ipr_drv_host_out_idx(void) { return D.h_idx; }
const uint32_t* ipr_drv_host_out_adj(void) { return D.h_adj; }
const uint32_t* ipr_drv_host_out_wgt(void) { return D.h_wgt; }

int ipr_drv_vertex_state(void* val, uint8_t* act, uint8_t* hm, void* msg)
{
    if (!D.initialised) return -EINVAL;
    const uint64_t V = D.V;
    int rc = 0;
    if (val && ipr_hip_memcpy_d2h(val, D.dev.val, V * D.app.val_size, D.stream)) rc = -EIO;
    if (act && ipr_hip_memcpy_d2h(act, D.dev.act, V, D.stream)) rc = -EIO;
    if (hm && ipr_hip_memcpy_d2h(hm, D.dev.hm, V, D.stream)) rc = -EIO;
    if (msg && ipr_hip_memcpy_d2h(msg, D.dev.msg, V * D.app.msg_size, D.stream)) rc = -EIO;
    if (ipr_hip_stream_sync(D.stream)) rc = -EIO;
    return rc;
}

int ipr_drv_in_lists(uint64_t* idx, uint32_t* adj)
{
    if (!D.initialised) return -EINVAL;
    if (!D.dev.in_idx || !D.dev.in_adj) return -ENOENT;
    int rc = 0;
    if (idx && ipr_hip_memcpy_d2h(idx, D.dev.in_idx, (D.V + 1) * 8, D.stream)) rc = -EIO;
    if (adj && ipr_hip_memcpy_d2h(adj, D.dev.in_adj, D.E * 4, D.stream)) rc = -EIO;
    if (ipr_hip_stream_sync(D.stream)) rc = -EIO;
    return rc;
}

/* ------------------------------------------------------------------------ */
/* ip_run                                                                    */
/* ------------------------------------------------------------------------ */

static void launch_app(int k)
{
    int rc = D.app.launch(k, &D.dev, D.stream);
    if (rc) die("HIP failure launching kernel %d (%d)\n", k, rc);
}

/* HIP events bracket the delivery phase of each superstep when kernel timing
 * is on: mark 0 before it, 1 between the gather (or the fused pull kernel)
 * and the expansion, 2 after it */
static void kev_mark(int which)
{
    if (!D.kev_on || D.kev_slot >= 64) return;
    hip_check(ipr_hip_event_record(D.kev[which][D.kev_slot], D.stream), "kernel event");
}

/* phase profile marks: 0 before compute, 1 after compute, 2 after the
 * delivery (the reference's "Mailbox" phase), 3 after merging the frontier */
static void pmark(int which)
{
    if (!D.prof_on || D.kev_slot >= 64) return;
    hip_check(ipr_hip_event_record(D.pev[which][D.kev_slot], D.stream), "profile event");
}

/* one superstep of the variant's BSP loop, in phase order; `fused`: the
 * spread variant's sparse superstep as one kernel, which ends the superstep
 * itself (its time counts as the delivery phase of the kernel events and as
 * compute in the phase profile) */
static void enqueue_superstep(int fused, uint32_t step_in_chunk)
{
    pmark(0);
    if (fused) {
        /* the fused kernel takes its superstep from the chunk's base plus its
         * place in the chunk, passed in its copy of the device state; its last
         * workgroup ends the superstep and leaves the next frontier as flags */
        ipr_dev dk = D.dev;
        dk.step_in_chunk = step_in_chunk;
        kev_mark(0);
        kev_mark(1);
        int rc = D.app.launch(IPR_K_SPARSE_STEP, &dk, D.stream);
        if (rc) die("HIP failure launching kernel %d (%d)\n", IPR_K_SPARSE_STEP, rc);
        kev_mark(2);
        pmark(1);
        pmark(2);
        pmark(3);
        if ((D.kev_on || D.prof_on) && D.kev_slot < 64) D.kev_slot++;
        return;
    }
    switch (D.app.variant) {
    case IPR_VARIANT_PULL:
        /* the fused kernel is fetch(s-1) + compute(s): counted as compute */
        kev_mark(0);
        launch_app(IPR_K_PULL_STEP);
        kev_mark(1);
        kev_mark(2);
        pmark(1);
        pmark(2);
        break;
    case IPR_VARIANT_SPREAD_PULL:
        /* src/combiner_spread_single_broadcast_postamble.h ip_run: compute on
         * the targets (broadcasts mark out-neighbours), target filtering,
         * fetch for the targets; broadcast flags are epoch-stamped, so the
         * reset phase has nothing to do */
        launch_app(IPR_K_COMPUTE);
        pmark(1);
        kev_mark(0);
        kev_mark(1);
        launch_app(IPR_K_EXPAND);
        kev_mark(2);
        hip_check(ipr_hip_compact_frontier(&D.dev, -1, 0, D.stream), "compact");
        launch_app(IPR_K_FETCH);
        pmark(2);
        break;
    case IPR_VARIANT_SPREAD:
    case IPR_VARIANT_PUSH: {
        /* under a launch plan only the delivery kernel the superstep used */
        int gather = D.dev.dir_opt, expand = 1;
        if (D.planning && step_in_chunk < 64 && D.dev.dir_opt) {
            gather = (int)((D.plan_mask >> step_in_chunk) & 1u);
            expand = !gather;
        }
        launch_app(IPR_K_COMPUTE);
        pmark(1);
        if (gather && D.dev.g_live && D.live_launch) launch_app(IPR_K_GATHER_LIVE);
        kev_mark(0);
        if (gather) {
            D.dev.plan_alone = !expand;
            launch_app(IPR_K_GATHER);
            D.dev.plan_alone = 0;
        }
        kev_mark(1);
        if (expand) {
            D.dev.plan_alone = !gather && D.dev.dir_opt;
            launch_app(IPR_K_EXPAND);
            D.dev.plan_alone = 0;
        }
        kev_mark(2);
        pmark(2);
        /* spread: the compaction's last workgroup ends the superstep */
        if (D.app.variant == IPR_VARIANT_SPREAD)
            hip_check(ipr_hip_compact_frontier(&D.dev, D.dev.fin_in_kernel ? IPR_VARIANT_SPREAD : -1, 0, D.stream),
                      "compact");
        else
            launch_app(IPR_K_SWAP);
        break;
    }
    default:
        die("Engine variant %d is not available in this build.\n", D.app.variant);
    }
    pmark(3);
    /* the fused pull kernel's or the spread compaction's last workgroup ends
     * the superstep itself */
    if (!((D.app.variant == IPR_VARIANT_PULL || D.app.variant == IPR_VARIANT_SPREAD) && D.dev.fin_in_kernel))
        hip_check(ipr_hip_finalize(&D.dev, D.app.variant, D.stream), "finalize");
    /* the next superstep of the chunk records into the next event slot */
    if ((D.kev_on || D.prof_on) && D.kev_slot < 64) D.kev_slot++;
}

/* n supersteps (1 to 64): replayed from a cached HIP graph per (length,
 * kind of chunk) when graphs are on. A fused chunk starts at superstep
 * `base`, stored in the control block before the chunk. */
static void enqueue_chunk(uint32_t n, int fused, uint64_t base)
{
    const uint32_t slot = n;
    D.kev_slot = 0;
    if (fused) hip_check(ipr_hip_set_u64(&D.dev.ctl->chunk_base, base, D.stream), "chunk base");
    /* the first chunk of a run that repeats the planned one */
    D.planning = !fused && D.plan_on && D.chunk_no == 0 && n == D.plan_steps;
    void** cache = fused ? D.fgraphs : D.graphs;
    void** g = D.planning ? &D.plan_graph : (slot < IPR_MAX_GRAPHS ? &cache[slot] : NULL);
    if (D.use_graph && !D.kev_on && !D.prof_on && g) {
        if (!*g) {
            hip_check(ipr_hip_graph_begin(D.stream), "graph capture");
            for (uint32_t i = 0; i < n; i++) enqueue_superstep(fused, i);
            hip_check(ipr_hip_graph_end(D.stream, g), "graph capture");
        }
        hip_check(ipr_hip_graph_launch(*g, D.stream), "graph launch");
    } else {
        for (uint32_t i = 0; i < n; i++) enqueue_superstep(fused, i);
    }
    D.planning = 0;
    D.chunk_no++;
}

/* the superstep about to run can be a fused sparse one: its frontier's
 * out-arcs (summed while the frontier was built) would not gather, and none
 * of its vertices has more than fused_max_deg out-arcs */
static int fusable(const ipr_ctl* c)
{
    const ipr_dev* d = &D.dev;
    const uint64_t s = c->superstep, arcs = c->farcs[s & 1];
    if (!d->fused || s == 0) return 0;
    /* after a dense superstep the figures are left unknown (all ones) */
    if (arcs > d->E || arcs > d->fused_max_arcs || c->fmaxdeg[s & 1] > d->fused_max_deg) return 0;
    return !(d->dir_opt && arcs * (uint64_t)d->dir_alpha > d->E);
}

/* flags raised by the device-side API (IPR_ERR_* of ipregel_kernels.h) */
static void report_device_error(uint32_t err)
{
    /* src/combiner_single_broadcast_postamble.h:42-49 prints this and exits */
    if (err & 1u)
        die("The function send_message should not be used in the SINGLE_BROADCAST "
            "version; only broadcast() should be called, and once per superstep maximum.\n");
    if (err & 4u)
        die("The function ip_broadcast_weighted should not be used in the SINGLE_BROADCAST "
            "version; weighted broadcasts are delivered by the push variants only.\n");
    if (err & 2u)
        die("ip_broadcast_weighted cannot be combined with another broadcast in one superstep.\n");
    if (err & 8u)
        die("ip_send_message was given a vertex identifier outside [0, %llu).\n",
            (unsigned long long)D.V);
    die("Device-side error flags 0x%x.\n", err);
}

static void run_once(void)
{
    void* st = D.stream;
    D.chunk_no = 0;
    hip_check(ipr_hip_begin(&D.dev, st), "begin");
    launch_app(IPR_K_INIT);
    hip_check(ipr_hip_event_record(D.ev0, st), "event");
    /* supersteps of the first chunk: each chunk ends with a host check of the
     * control block (about 30 us), each superstep launched after the loop has
     * halted exits at once (one launch for pull, five for the others) */
    uint32_t chunk = D.first_chunk < D.max_chunk ? D.first_chunk : D.max_chunk;
    if (D.adapt_chunk && D.prev_launches && !D.prev_fused)
        chunk = D.prev_launches < D.max_chunk ? D.prev_launches : D.max_chunk;
    uint64_t launched = 0, prev_exec = 0;
    int fused = 0;
    for (;;) {
        /* a fused chunk starts at the superstep the last readback showed
         * (fused mode is only entered after one) */
        enqueue_chunk(chunk, fused, fused ? D.h_ctl->superstep : 0);
        launched += chunk;
        hip_check(ipr_hip_memcpy_d2h(D.h_ctl, D.dev.ctl, sizeof(ipr_ctl), st), "ctl readback");
        hip_check(ipr_hip_stream_sync(st), "superstep chunk");
        if (D.h_ctl->error) report_device_error(D.h_ctl->error);
        if (D.kev_on) {
            /* launches after the halting superstep exit at once: not counted */
            uint64_t live = D.h_ctl->executed - prev_exec;
            if (D.app.variant == IPR_VARIANT_PULL && prev_exec == 0 && live < chunk) live++;
            for (uint32_t i = 0; i < D.kev_slot && i < live; i++) {
                float k1 = 0, k2 = 0;
                hip_check(ipr_hip_event_elapsed_ms(D.kev[0][i], D.kev[1][i], &k1), "kernel event");
                hip_check(ipr_hip_event_elapsed_ms(D.kev[1][i], D.kev[2][i], &k2), "kernel event");
                D.stats.kernel_ms += k1 + k2;
                D.stats.kernel_launches++;
                uint64_t sidx = prev_exec + i;
                if (sidx < IPR_TRACE_CAP) {
                    D.kev_ms[0][sidx] = k1;
                    D.kev_ms[1][sidx] = k2;
                    if (sidx + 1 > D.kev_n) D.kev_n = sidx + 1;
                }
            }
        }
        if (D.prof_on) {
            uint64_t live = D.h_ctl->executed - prev_exec;
            if (D.app.variant == IPR_VARIANT_PULL && prev_exec == 0 && live < chunk) live++;
            for (uint32_t i = 0; i < D.kev_slot && i < live; i++) {
                const uint64_t sidx = prev_exec + i;
                if (sidx >= IPR_TRACE_CAP) break;
                for (int k = 0; k < 3; k++) {
                    float ms = 0;
                    hip_check(ipr_hip_event_elapsed_ms(D.pev[k][i], D.pev[k + 1][i], &ms), "profile event");
                    D.prof_ms[k][sidx] = ms;
                }
            }
        }
        prev_exec = D.h_ctl->executed;
        if (D.h_ctl->done) break;
        if (fused && D.h_ctl->bail) {
            /* back to the general kernels (the fused kernels kept the 4-byte
             * records' stamp-window clears); a frontier that a fused kernel
             * left as flags is compacted into its list first */
            hip_check(ipr_hip_fill(&D.dev.ctl->bail, 0, sizeof(uint32_t), st), "bail clear");
            if (D.h_ctl->front_flags) {
                const uint64_t s = D.h_ctl->superstep;
                hip_check(ipr_hip_set_u64(&D.dev.ctl->fsize[s & 1], 0, st), "relist counter");
                hip_check(ipr_hip_compact_frontier(&D.dev, -1, 1, st), "relist");
            }
            fused = 0;
            chunk = 1;
            continue;
        }
        if (!fused && fusable(D.h_ctl)) {
            fused = 1;
            chunk = D.first_chunk < D.max_chunk ? D.first_chunk : D.max_chunk;
            continue;
        }
        if (chunk < D.max_chunk) chunk *= 2;
        if (chunk > D.max_chunk) chunk = D.max_chunk;
    }
    if (D.h_ctl->plan_miss) {
        /* a superstep left its delivery kernel out: the run stopped there.
         * Drop the plan and run again (the run starts from its init). */
        D.plan_on = 0;
        D.plan_steps = 0;
        if (D.plan_graph) ipr_hip_graph_destroy(D.plan_graph);
        D.plan_graph = NULL;
        run_once();
        return;
    }
    hip_check(ipr_hip_event_record(D.ev1, st), "event");
    float ms = 0;
    hip_check(ipr_hip_event_elapsed_ms(D.ev0, D.ev1, &ms), "event");
    D.stats.last_ms = ms;
    D.stats.total_ms += ms;
    if (D.stats.runs == 0 || ms < D.stats.min_ms) D.stats.min_ms = ms;
    D.stats.runs++;
    D.stats.launches = launched;
    D.stats.supersteps = D.h_ctl->executed;
    D.stats.work_edges = D.h_ctl->work_edges;
    D.stats.work_vertices = D.h_ctl->work_vertices;
    D.stats.gathered = D.h_ctl->gathered;
    D.stats.fused_steps = D.h_ctl->fused_steps;
    D.stats.gather_tiles = D.dev.g_tn;
    /* pull: launch s finalizes superstep s-1, so one launch more than supersteps */
    D.prev_launches = (uint32_t)(D.h_ctl->executed + (D.app.variant == IPR_VARIANT_PULL ? 1 : 0));
    D.prev_fused = D.h_ctl->fused_steps > 0;
    /* the next run's launch plan: this run's gathered supersteps, for a run
     * of at most 64 supersteps without fused sparse ones (IPR_LAUNCH_PLAN=2
     * inverts it, which makes every planned run stop in its first superstep
     * and run again without the plan: a test of that path) */
    if (D.plan_mode >= 1 && D.dev.dir_opt && D.dev.g_tn &&
        (D.app.variant == IPR_VARIANT_SPREAD || D.app.variant == IPR_VARIANT_PUSH)) {
        const int ok = D.h_ctl->fused_steps == 0 && D.h_ctl->executed > 0 && D.h_ctl->executed <= 64;
        const uint64_t mask = D.plan_mode == 2 ? ~D.h_ctl->gathered_mask : D.h_ctl->gathered_mask;
        const uint32_t steps = ok ? (uint32_t)D.h_ctl->executed : 0u;
        if (ok != D.plan_on || mask != D.plan_mask || steps != D.plan_steps) {
            if (D.plan_graph) ipr_hip_graph_destroy(D.plan_graph);
            D.plan_graph = NULL;
        }
        D.plan_on = ok;
        D.plan_mask = mask;
        D.plan_steps = steps;
    }
    if (D.live_mode == 2 && D.dev.g_live) {
        const int want = D.h_ctl->live_built > 0 || D.h_ctl->live_missed != 0;
        if (want != D.live_launch) {
            /* the cached chunk graphs hold the old launch list */
            D.live_launch = want;
            for (int i = 0; i < IPR_MAX_GRAPHS; i++) {
                if (D.graphs[i]) ipr_hip_graph_destroy(D.graphs[i]);
                D.graphs[i] = NULL;
            }
            if (D.plan_graph) ipr_hip_graph_destroy(D.plan_graph);
            D.plan_graph = NULL;
        }
    }
}

static int run_and_report(int reps);

int ipr_drv_run(void)
{
    if (!D.initialised) die("ip_run called before ip_init.\n");
    if (D.have_cfg && D.cfg.defer_run) return 0;
    int reps = D.have_cfg && D.cfg.repeat > 0 ? D.cfg.repeat : 1;
    return run_and_report(reps);
}

int ipr_harness_run(int repeat)
{
    if (!D.initialised) return -EINVAL;
    int q = D.quiet;
    D.quiet = 1;
    int rc = run_and_report(repeat > 0 ? repeat : 1);
    D.quiet = q;
    return rc;
}

void* ipr_harness_stream(void) { return D.stream; }

int ipr_harness_kernel_events(int on)
{
    if (on && !D.kev[0][0]) {
        for (int i = 0; i < 64; i++)
            for (int k = 0; k < 3; k++)
                if (ipr_hip_event_create(&D.kev[k][i])) return -EIO;
    }
    if (on && !D.kev_ms[0]) {
        for (int k = 0; k < 2; k++) {
            D.kev_ms[k] = (double*)calloc(IPR_TRACE_CAP, sizeof(double));
            if (!D.kev_ms[k]) return -ENOMEM;
        }
    }
    D.kev_on = on != 0;
    return 0;
}

int ipr_harness_kernel_times(double* first, double* second, uint64_t cap)
{
    if (!D.initialised || !D.kev_ms[0]) return -EINVAL;
    uint64_t n = D.kev_n < cap ? D.kev_n : cap;
    for (uint64_t s = 0; s < n; s++) {
        if (first) first[s] = D.kev_ms[0][s];
        if (second) second[s] = D.kev_ms[1][s];
    }
    return (int)n;
}

/* the reference's IP_ENABLE_THREAD_PROFILING table
 * (src/combiner_spread_postamble.h:354-400), with the device as its one
 * column: phase times of the superstep in seconds, and its share of the
 * superstep's relaxed arcs (the device's: 100%) with their count */
static void print_phase_table(uint64_t s, uint64_t arcs)
{
    const double c = D.prof_ms[0][s] * 1e-3, mb = D.prof_ms[1][s] * 1e-3, mg = D.prof_ms[2][s] * 1e-3;
    printf("            +-----------+\n            |    Device |\n+-----------+-----------+\n");
    printf("|   Compute | %8.6fs |\n", c);
    printf("|   Merging | %8.6fs |\n", mg);
    printf("|   Mailbox | %8.6fs |\n", mb);
    printf("|     Total | %8.6fs |\n", c + mg + mb);
    printf("| EdgeCount | %8.3f%% |\n", arcs ? 100.0 : 0.0);
    printf("|      Arcs | %9llu |\n", (unsigned long long)arcs);
    printf("+-----------+-----------+\n");
}

static int run_and_report(int reps)
{
    if (D.V == 0) {
        printf("Total time of supersteps: %fs.\n", 0.0);
        D.stats.supersteps = 0;
        return 0;
    }
    for (int r = 0; r < reps; r++) run_once();
    uint64_t n = D.stats.supersteps;
    uint64_t m = n < IPR_TRACE_CAP ? n : IPR_TRACE_CAP;
    uint64_t* act = safe_malloc((m + 1) * 8);
    uint64_t* ts = safe_malloc((m + 2) * 8);
    hip_check(ipr_hip_memcpy_d2h(act, D.dev.trace_active, m * 8, D.stream), "trace");
    hip_check(ipr_hip_memcpy_d2h(ts, D.dev.trace_ts, (m + 1) * 8, D.stream), "trace");
    hip_check(ipr_hip_stream_sync(D.stream), "trace");
    double tick = D.wall_hz ? 1.0 / (double)D.wall_hz : 1e-8;
    uint64_t* sedges = NULL;
    if (!D.quiet && D.prof_on) {
        sedges = safe_malloc((m + 1) * 8);
        hip_check(ipr_hip_memcpy_d2h(sedges, D.dev.trace_edges, m * 8, D.stream), "trace");
        hip_check(ipr_hip_stream_sync(D.stream), "trace");
    }
    if (!D.quiet) {
        for (uint64_t s = 0; s < m; s++) {
            printf("Superstep%zuDuration:%f\n", (size_t)s, (double)(ts[s + 1] - ts[s]) * tick);
            printf("Superstep%zuActiveVertexCount:%zu\n", (size_t)s, (size_t)act[s]);
            if (sedges) print_phase_table(s, sedges[s]);
        }
    }
    free(sedges);
    D.stats.device_ms = m ? (double)(ts[m] - ts[0]) * tick * 1e3 : 0.0;
    double t = D.stats.last_ms * 1e-3;
    printf("Total time of supersteps: %fs.\n", t);
    printf("DeviceSuperstepTime:%f\n", D.stats.device_ms * 1e-3);
    printf("WorkEdges:%llu\n", (unsigned long long)D.stats.work_edges);
    printf("GTEPS:%f\n", t > 0 ? (double)D.stats.work_edges / t * 1e-9 : 0.0);
    fflush(stdout);
    free(act);
    free(ts);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ip_dump                                                                   */
/* ------------------------------------------------------------------------ */

void ipr_drv_dump(FILE* f)
{
    double t0 = now_s();
    if (D.have_cfg && D.cfg.skip_dump) {
        printf("DumpingTime:%f\n", now_s() - t0);
        return;
    }
    if (!D.initialised) die("ip_dump called before ip_init.\n");
    size_t bytes = (size_t)D.V * D.app.val_size;
    void* host = safe_malloc(bytes);
    hip_check(ipr_hip_memcpy_d2h(host, D.dev.val, bytes, D.stream), "value readback");
    hip_check(ipr_hip_stream_sync(D.stream), "value readback");
    D.app.serialise(f, host, D.h_idx, 0, D.V);
    fflush(f);
    free(host);
    printf("DumpingTime:%f\n", now_s() - t0);
}

/* ------------------------------------------------------------------------ */
/* harness                                                                   */
/* ------------------------------------------------------------------------ */

void ipr_harness_configure(const ipr_harness_cfg* cfg)
{
    if (cfg) {
        D.cfg = *cfg;
        D.have_cfg = 1;
    } else {
        memset(&D.cfg, 0, sizeof(D.cfg));
        D.have_cfg = 0;
    }
}

void ipr_harness_reset(void) { ipr_harness_configure(NULL); }

int ipr_harness_stats(ipr_run_stats* out)
{
    if (!out) return -EINVAL;
    *out = D.stats;
    return 0;
}

int ipr_harness_values(void* out, uint64_t bytes)
{
    if (!D.initialised) return -EINVAL;
    uint64_t need = D.V * D.app.val_size;
    if (bytes < need) return -EINVAL;
    if (ipr_hip_memcpy_d2h(out, D.dev.val, need, D.stream)) return -EIO;
    if (ipr_hip_stream_sync(D.stream)) return -EIO;
    return 0;
}

int ipr_harness_trace(uint64_t* active, uint64_t* edges, double* ns, uint64_t cap)
{
    if (!D.initialised) return -EINVAL;
    uint64_t n = D.stats.supersteps;
    if (n > IPR_TRACE_CAP) n = IPR_TRACE_CAP;
    if (n > cap) n = cap;
    uint64_t* ts = safe_malloc((n + 1) * 8);
    int rc = 0;
    if (active && ipr_hip_memcpy_d2h(active, D.dev.trace_active, n * 8, D.stream)) rc = -EIO;
    if (edges && ipr_hip_memcpy_d2h(edges, D.dev.trace_edges, n * 8, D.stream)) rc = -EIO;
    if (ipr_hip_memcpy_d2h(ts, D.dev.trace_ts, (n + 1) * 8, D.stream)) rc = -EIO;
    if (ipr_hip_stream_sync(D.stream)) rc = -EIO;
    if (!rc && ns) {
        double tick = D.wall_hz ? 1e9 / (double)D.wall_hz : 10.0;
        for (uint64_t s = 0; s < n; s++) ns[s] = (double)(ts[s + 1] - ts[s]) * tick;
    }
    free(ts);
    return rc ? rc : (int)n;
}

void ipr_harness_release(void)
{
    if (D.initialised) {
        ipr_hip_stream_sync(D.stream);
        dev_free_all();
        host_free();
        D.initialised = 0;
    }
}
