/*
 * ipregel_engine.h - definitions shared by the C host driver, the HIP C-ABI
 * library (libipregel_hip.so) and the kernels that iPregel.h instantiates in
 * the application's translation unit. Plain C, no HIP types.
 *
 * Device layout (structure of arrays in HBM, V vertices, E arcs):
 *   out_idx[V+1] u64, out_adj[E] u32, out_wgt[E] u32   out-edge CSR
 *   in_idx[V+1]  u64, in_adj[E]  u32, in_wgt[E]  u32   in-edge CSC (aliases the
 *                                                      CSR when directed=false,
 *                                                      src/iPregel_postamble.h:368-396)
 *   val[V]        application value                    (ip_vertex_t.value)
 *   msg[V]        delivered combined message            (ip_vertex_t.message)
 *   act[V], hm[V] active / has_message flags, one byte each
 *   bcm[2][V]     broadcast records {superstep+1, message} by superstep
 *                 parity (pull variants; a record counts only when its stamp
 *                 is the superstep that reads it, so nothing is cleared)
 *   bslot[V]      push variants: last broadcast {superstep+1 | weighted<<31, msg}
 *   mbox[2][V]    combiner mailboxes by parity (push variants)
 *   recv[2][V]    received-a-message flags (bytes) by parity (push variants);
 *                 the next frontier is compacted from them in ascending runs
 *   front[2][V]   frontier lists by parity (spread variants)
 *   lbins, lbase  broadcasters above large_min out-arcs and their first chunk
 */
#ifndef IPREGEL_ENGINE_H
#define IPREGEL_ENGINE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* engine variants: IP_USE_SPREAD x IP_USE_SINGLE_BROADCAST
 * (src/iPregel_preamble.h:319-331) */
enum {
    IPR_VARIANT_PUSH = 0,        /* combiner_{pre,post}amble.h */
    IPR_VARIANT_PULL = 1,        /* combiner_single_broadcast_{pre,post}amble.h */
    IPR_VARIANT_SPREAD = 2,      /* combiner_spread_{pre,post}amble.h */
    IPR_VARIANT_SPREAD_PULL = 3  /* combiner_spread_single_broadcast_{pre,post}amble.h */
};

/* combiner declared on the compile line (-DIP_COMBINER_MIN etc.) */
enum {
    IPR_COMBINER_GENERIC = 0,    /* user ip_combine in a CAS loop */
    IPR_COMBINER_MIN = 1,
    IPR_COMBINER_MAX = 2,
    IPR_COMBINER_SUM = 3
};

/* in-degree bins of the pull fetch and of the push gather */
enum { IPR_BIN_SMALL = 0, IPR_BIN_MEDIUM = 1, IPR_BIN_LARGE = 2, IPR_NBINS = 3 };

/* counters spread over IPR_CSLOTS slots (a workgroup adds its total to slot
 * blockIdx % IPR_CSLOTS) so that no single address takes every atomic; the
 * finalize kernel sums and clears them. Device atomics on one cache line are
 * performed one after another (about 12 ns each, whatever their address in
 * the line), so every slot has a 128-byte line of its own: IPR_CPAD words. */
#define IPR_CSLOTS 64
#define IPR_CPAD 16
#define IPR_CSLOT(id, slot) ((((uint64_t)(id)) * IPR_CSLOTS + (uint64_t)(slot)) * IPR_CPAD)
enum {
    IPR_C_ACT0 = 0,      /* ActiveVertexCount, superstep parity 0 */
    IPR_C_ACT1 = 1,      /* ... parity 1 */
    IPR_C_WORKV = 2,     /* vertices computed */
    IPR_C_EDGES = 3,     /* out-arcs of this superstep's broadcasters (W) */
    IPR_C_BCAST0 = 4,    /* broadcasts, superstep parity 0 */
    IPR_C_BCAST1 = 5,    /* ... parity 1 */
    IPR_C_GNEXT = 6,     /* gather range counters: IPR_GATHER_CTRS of them, one per slot
                          * (zeroed at every superstep's end); id 7 is unused */
    IPR_C_ARRIVE = 8,    /* workgroups of the running kernel that have finished, by
                          * blockIdx mod 64, for kernels whose last workgroup ends
                          * the superstep ... */
    IPR_C_ARRIVE_TOP = 9,/* ... slot 0: the slots whose workgroups have all finished */
    IPR_C_FARCS = 10,    /* spread: out-arcs of the next frontier */
    IPR_C_FMAX = 11,     /* spread: its largest out-degree (per-slot maxima) */
    IPR_C_FNEXT = 12,    /* spread: the next frontier's size, counted by the fused
                          * sparse kernel as its deliveries activate receivers */
    IPR_C_LPACK = 13,    /* large broadcasters registered in slot k: entries << IPR_LPACK_SHIFT
                          * | arc chunks; slot k's entries are lbins/lbase[k * lcap ...] */
    IPR_NCNT = 14
};

/* device-resident control block: every kernel reads the superstep and the
 * halting flag here, so a chunk of supersteps can be replayed as one HIP graph
 * without the host deciding anything in between */
typedef struct ipr_ctl {
    uint64_t superstep;          /* current superstep */
    uint32_t done;               /* set by the finalize kernel when the loop ends */
    uint32_t error;              /* device-side API misuse flags */
    uint64_t fsize[2];           /* frontier list sizes by parity */
    uint64_t pull_mode;          /* push apps: 1 when the last finalized superstep gathered */
    uint64_t step_edges;         /* W of the last finalized superstep */
    uint64_t work_edges;         /* W accumulated */
    uint64_t work_vertices;      /* vertices computed, accumulated */
    uint64_t trace_cap;          /* entries in the trace arrays */
    uint64_t executed;           /* supersteps executed so far */
    uint64_t last_active;        /* ActiveVertexCount of the last finalized superstep */
    uint64_t gathered;           /* supersteps delivered by gather */
    /* spread variant, by superstep parity: the out-arcs of the frontier's
     * vertices and the largest of their out-degrees, summed while the frontier
     * is built; they decide whether a superstep can run as one fused sparse
     * kernel, and are cleared when that frontier has been computed */
    uint64_t farcs[2];
    uint64_t fmaxdeg[2];
    uint32_t bail;               /* a fused sparse kernel found its superstep too dense and stopped */
    uint32_t fused_steps;        /* supersteps run by the fused sparse kernel */
    /* the superstep a chunk of fused sparse kernels starts at, set before the
     * chunk is launched; kernel k of the chunk runs superstep chunk_base + k
     * (never ctl->superstep, which its own last workgroup advances) */
    uint64_t chunk_base;
    /* spread: 1 when the current superstep's frontier exists only as raised
     * receive flags (a fused sparse kernel built it, no list), 0 when it is
     * the compacted list front[superstep & 1] */
    uint64_t front_flags;
    /* the broadcasters' bitmap: superstep + 1 it was last built for, bitmaps
     * built this run, and whether a gather wanted one that was not built
     * (the host launches the bitmap kernel only while runs use it) */
    uint64_t live_step;
    uint32_t live_built;
    uint32_t live_missed;
    /* push apps: bit s set when superstep s < 64 gathered; set when a
     * delivery kernel launched alone (a launch plan) found the superstep
     * needs the other one (the run stops; the host runs it again unplanned) */
    uint64_t gathered_mask;
    uint32_t plan_miss;
    uint32_t plan_pad;
} ipr_ctl;

#define IPR_LPACK_SHIFT 40
/* arc-balanced gather: arcs per lane per pass, and the end-of-list mark */
#ifndef IPR_GATHER_K
#define IPR_GATHER_K 8
#endif
#define IPR_GATHER_LAST 0x80000000u
/* gather split by record line: classes (one per XCD), positions per class
 * unit (the 4-byte records of one 128-byte line) */
#define IPR_GATHER_CLASSES 8
#define IPR_REC4_LINE 32
#define IPR_LPACK_MASK ((1ull << IPR_LPACK_SHIFT) - 1)

/* receiver-blocked gather: one block of consecutive receivers and its
 * in-arcs, sorted by the source's record position; a block whose arcs are
 * split between several tiles (a receiver with more in-arcs than a tile
 * holds) has `shared` set, and its tiles deliver with the combiner's atomic */
typedef struct ipr_tblock {
    uint64_t a0;                 /* first arc slot in g_tq / g_twgt (a multiple of 64) */
    uint32_t na;                 /* arc slots (a multiple of 64; unused slots hold IPR_TQ_NONE) */
    uint32_t v0;                 /* first receiver */
    uint32_t nr;                 /* receivers (at most 65536: 16-bit offsets) */
    uint32_t shared;
} ipr_tblock;

/* The receiver-blocked gather's index: each block's arcs, sorted by source
 * record position, fill whole groups of 64 slots. Slot i holds
 * (position - base) << 16 | receiver offset in g_tq[i]; g_thead[i / 64] is the
 * group's base position, or IPR_TQ_ESC | e when its positions span more than
 * 16 bits, and then the positions are g_tesc[64 e + i % 64] (the slot's high
 * half is unused). A slot whose receiver offset is IPR_TQ_NONE is empty. */
#define IPR_TQ_ESC 0x80000000u
#define IPR_TQ_NONE 0xFFFFu

typedef struct ipr_gather_tiles {
    void* tiles;                 /* device ipr_tblock[n_tiles], largest first */
    uint32_t n_tiles;
    uint32_t wgt_bytes;          /* width of wgt: 1, 2 or 4 (0: no weights) */
    uint64_t slots;              /* entries of q (a multiple of 64) */
    uint64_t escaped;            /* groups whose positions are in esc */
    uint32_t* q;                 /* device, slots */
    uint32_t* head;              /* device, slots / 64 */
    uint32_t* esc;               /* device, 64 per escaped group (NULL when none) */
    void* wgt;                   /* device, slots x wgt_bytes (NULL without weights) */
} ipr_gather_tiles;

typedef struct ipr_dev {
    ipr_ctl* ctl;
    uint64_t* cnt;               /* IPR_NCNT x IPR_CSLOTS counters, IPR_CPAD words apart */
    uint64_t V, E;
    const uint64_t* out_idx;
    const uint32_t* out_adj;
    const uint32_t* out_wgt;
    const uint64_t* in_idx;
    const uint32_t* in_adj;
    const uint32_t* in_wgt;
    /* the adjacency in the application's IP_VERTEX_ID_TYPE, for the
     * out_neighbours / in_neighbours pointers of ip_vertex_t: the 32-bit
     * arrays above in _32 builds, 64-bit copies in _64 builds that ask for
     * neighbour ids (IP_NEEDS_{OUT,IN}_NEIGHBOUR_IDS), NULL otherwise */
    const void* out_ids;
    const void* in_ids;
    /* push variants that gather dense supersteps: the exact transpose of the
     * out-edge CSR (aliases it when the graph is symmetric) */
    const uint64_t* t_idx;
    const uint32_t* t_adj;
    const uint32_t* t_wgt;
    /* arc-balanced gather over the transpose: sources with IPR_GATHER_LAST on
     * the last arc of each in-list, the receivers of the non-empty lists, and
     * per IPR_GATHER_K-arc chunk the receiver (position in g_dst) of its first
     * arc; each wavefront folds g_per_wave consecutive arcs */
    const uint32_t* g_arc;
    const uint32_t* g_own;
    const uint32_t* g_dst;
    /* SPREAD with gather: the broadcast record of vertex v lives at
     * bslot[bpos[v]], records ordered by descending out-degree class so that
     * the most-gathered ones share cache lines, ids ascending within a class
     * so that consecutive frontier vertices write nearby records (NULL:
     * bslot[v]) */
    const uint32_t* bpos;
    /* its inverse, bvert[bpos[v]] = v: a sweep over all records (the SPREAD
     * expansion of superstep 0) reads them in position order */
    const uint32_t* bvert;
    /* dense supersteps whose broadcasters' out-arcs are below live_pct % of
     * E: bit p of g_live is set when the record at position p was broadcast
     * this superstep, and the gather reads only those records */
    uint64_t* g_live;
    uint32_t live_pct;
    uint32_t live_vpct;          /* ... and the vertices computed are below this % of V */
    /* push apps with integral messages that gather: a 4-byte copy of each
     * broadcast record at the same position as in bslot (IPR_REC4_* layout),
     * read by the gather in place of the 8-byte record (NULL: off) */
    uint32_t* brec4;
    uint64_t g_per_wave;         /* arcs per range (whole passes) */
    /* g_xcd set: the gather arcs are split into IPR_GATHER_CLASSES classes by
     * the line of the source's record; class c's arcs are g_arc[g_cbeg[c],
     * g_cend[c]) (g_cbeg a multiple of IPR_GATHER_K), its g_own entries index
     * g_dst + g_cdst[c], and workgroup b starts with class b mod classes */
    uint32_t g_xcd;
    uint32_t g_xpad;
    const uint32_t* g_wgt;       /* g_xcd with weights: the arcs' weights in g_arc's order */
    uint64_t g_cbeg[IPR_GATHER_CLASSES];
    uint64_t g_cend[IPR_GATHER_CLASSES];
    uint64_t g_cdst[IPR_GATHER_CLASSES];
    uint32_t gather_blocks;      /* resident gather workgroups; their wavefronts take ranges in turn */
    /* receiver-blocked gather (g_tn > 0): the tiles, largest first, taken in
     * turn by the resident workgroups, each folding its block's messages in
     * LDS; the index g_tq / g_thead / g_tesc (layout at IPR_TQ_ESC), the
     * arcs' weights g_twgt in the same slots, g_twsz bytes each (weighted
     * apps: 1 or 2 when every weight fits) */
    const ipr_tblock* g_tblk;
    uint32_t g_tn;
    uint32_t tile_blocks;        /* resident workgroups of the receiver-blocked gather */
    const uint32_t* g_tq;
    const uint32_t* g_thead;
    const uint32_t* g_tesc;
    const void* g_twgt;
    uint32_t g_twsz;
    uint32_t g_tpad;
    uint32_t compact_rounds;     /* frontier compaction rounds per workgroup: 1, 2, 4 or 8 (0: by V) */
    void* val;
    void* msg;
    uint8_t* act;
    uint8_t* hm;
    void* bcm[2];
    void* mbox[2];
    uint8_t* recv[2];
    uint32_t* front[2];
    uint32_t* lbins;             /* large broadcasters of this superstep, IPR_CSLOTS regions of lcap */
    uint64_t* lbase;             /* their first arc chunk within their slot */
    uint64_t lcap;               /* vertices above large_min out-arcs */
    void* bslot;                 /* push variants: {superstep+1 | weighted<<31, message} */
    /* static pull / gather bins over all V (by in-degree), built at load */
    uint32_t* pbins[IPR_NBINS];
    uint64_t pbin_size[IPR_NBINS];
    uint64_t* trace_active;
    uint64_t* trace_ts;
    uint64_t* trace_edges;
    /* tuning */
    uint32_t large_min;          /* out-degree above which a broadcaster is split into chunks */
    uint32_t large_chunk;        /* arcs per workgroup task of a large broadcaster */
    uint32_t grid_blocks;        /* persistent-style grids */
    uint32_t pull_ordered;       /* 1: fold in in-list order (bit-exact) */
    uint32_t pull_wave_blocks;   /* pull: most workgroups of the wavefront bin (its waves loop beyond) */
    uint32_t dir_opt;            /* push apps: allow gathering dense supersteps */
    uint32_t dir_alpha;          /* gather when bcast out-arcs * alpha > E */
    /* spread variant: supersteps whose frontier has at most fused_max_arcs
     * out-arcs, none above fused_max_deg, and would not gather run as one
     * fused kernel (frontier scan, compute, expansion, end of superstep)
     * (0: off) */
    uint32_t fused;
    uint32_t fused_blocks;       /* workgroups of the fused kernel */
    uint64_t fused_max_arcs;
    uint64_t fused_max_deg;
    /* the last workgroup of the fused pull kernel (pull) or of the frontier
     * compaction (spread) ends the superstep: no separate finalize launch */
    uint32_t fin_in_kernel;
    /* position of a fused sparse kernel in its chunk (set per launch) */
    uint32_t step_in_chunk;
    /* set in the copy passed to a gather or expansion launched without the
     * other in its superstep (a launch plan): it flags ctl->plan_miss if the
     * superstep needs the other */
    uint32_t plan_alone;
    /* SPREAD: after a superstep that gathered, the compaction only counts
     * the raised flags and the next superstep reads its frontier from them */
    uint32_t dense_flags;
} ipr_dev;

/* 4-byte broadcast record: bits 31-26 the superstep's stamp (superstep mod
 * IPR_REC4_WINDOW, plus one; 0 never matches), bit 25 ip_broadcast_weighted,
 * bits 24-0 the message, or IPR_REC4_ESC when it does not fit (the reader
 * then takes the message from the 8-byte record). The stamps of a window of
 * IPR_REC4_WINDOW supersteps are distinct; every record is cleared to 0 at
 * the end of a window's last superstep, so a stale record never matches. */
/* gather range counters (workgroup b takes ranges from counter b mod 8) */
#define IPR_GATHER_CTRS 8
/* wavefronts per gather workgroup */
#ifndef IPR_GATHER_WAVES
#define IPR_GATHER_WAVES 4
#endif

#define IPR_REC4_SHIFT 26
#define IPR_REC4_WINDOW 63u
#define IPR_REC4_WEIGHTED 0x02000000u
#define IPR_REC4_ESC 0x01FFFFFFu

/* kernels the application TU provides (typed by the app's message/value) */
enum {
    IPR_K_INIT = 0,      /* identity-initialise mailboxes, flags */
    IPR_K_COMPUTE,       /* compute phase of the variant */
    IPR_K_EXPAND,        /* push delivery of recorded broadcasts */
    IPR_K_SWAP,          /* PUSH: mailbox -> message, reactivation */
    IPR_K_PULL_STEP,     /* PULL: fetch(s-1) fused with compute(s) */
    IPR_K_FETCH,         /* SPREAD_PULL: fetch for targets */
    IPR_K_GATHER,        /* push apps, dense superstep: receivers gather */
    IPR_K_GATHER_LIVE,   /* push apps, dense superstep with few broadcasters: their bitmap */
    IPR_K_SPARSE_STEP,   /* SPREAD, sparse superstep: compute, expansion, next frontier, end */
    IPR_K_COUNT
};

typedef struct ipr_app_desc {
    int variant;
    int combiner;
    uint32_t id_size, val_size, msg_size, mbox_size, bslot_size;
    uint32_t idx_size;           /* sizeof(IP_NEIGHBOUR_COUNT_TYPE): the width of the .idx offsets */
    int needs_in_lists;          /* pull variants, IP_NEEDS_IN_NEIGHBOUR_{IDS,COUNT} */
    int needs_out_ids;           /* IP_NEEDS_OUT_NEIGHBOUR_IDS */
    int needs_in_ids;            /* IP_NEEDS_IN_NEIGHBOUR_IDS (pull variants imply it) */
    int needs_weights;           /* IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS */
    int can_gather;              /* push app whose combiner makes gather exact */
    const char* application;     /* IP_APPLICATION */
    const char* version;         /* VERSION */
    const char* machine;         /* IP_MACHINE */
    const char* commits;         /* COMMITS */
    const char* compilation_flags;
    /* enqueue one app kernel on `stream` (a hipStream_t); 0 on success */
    int (*launch)(int kernel, const ipr_dev* dev, void* stream);
    /* make device-side globals point at this run's state */
    int (*bind)(const ipr_dev* dev);
    /* ip_dump: serialise vertices [first, first+count) from host copies */
    void (*serialise)(void* file, const void* val, const uint64_t* out_idx,
                      uint64_t first, uint64_t count);
    /* host-side self-check of the declared combiner (0 = consistent) */
    int (*check_combiner)(void);
    /* built with IP_ENABLE_THREAD_PROFILING: a per-superstep phase table */
    int thread_profiling;
    /* the receiver-blocked gather applies: an integral MIN/MAX/SUM combiner
     * on 4- or 8-byte slots (folded by LDS atomics); its LDS holds
     * tile_receivers receivers per tile */
    int can_tile;
    uint32_t tile_receivers;
    uint32_t tile_wgs_per_cu;    /* resident workgroups per CU its LDS allows */
} ipr_app_desc;

#ifdef __cplusplus
}
#endif
#endif
