/*
 * ipr_batch_int.h - internal interface of libipregel_batch.so, shared by the
 * kernels (ipr_batch.hip) and the serving loop (ipr_serve.cpp).
 *
 * Two state families run over a resident graph:
 *   PushState  frontier push with min-combined mailboxes. Scalar layout (one
 *              query: SSSP or CC) or batched layout (64 SSSP queries, one
 *              per wavefront lane: dist[v][64], mbox[v][64]).
 *   PprState   personalised PageRank pull over the transpose. Scalar layout
 *              (one seed) or batched layout (64 seeds: y[v][64]).
 * A state owns one HIP stream; each superstep (tick) is replayed as a HIP
 * graph captured once per frontier parity.
 */
#ifndef IPR_BATCH_INT_H
#define IPR_BATCH_INT_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ipregel_batch.h"

namespace iprb {

constexpr int BLOCK = 256;
constexpr uint32_t INF = 0xFFFFFFFFu;
constexpr int SPREAD = 16;        /* copies of each per-slot accumulator (fewer same-address atomics) */
constexpr int CHAINS = 256;       /* first-level chains of the canonical FP64 reductions */
constexpr uint32_t LDS_LARGE = 4096; /* chunk-prefix entries of the large list cached in LDS */
/* PPR: in-lists longer than HUB arcs are folded as HUB-arc chunks, each
 * compensated, then the chunk values folded in order. Part of the definition
 * of the FP32 result (scalar and batched kernels follow it alike), so it is a
 * constant and not a tuning knob. */
constexpr uint32_t HUB = 2048;
/* PPR: pull segments hold at most 64 consecutive vertices and at most SEG_ARCS
 * non-hub in-arcs (a single heavier vertex forms its own segment). The FP64
 * sums follow the segments, so this is part of the definition as well. */
constexpr uint32_t SEG_ARCS = 2048;

/* launch and binning parameters (chosen per graph by the autotuner or defaults) */
struct Tuning {
    uint32_t large;        /* push: out-degree above which a frontier vertex is relaxed in chunks */
    uint32_t chunk;        /* push: arcs per chunk of a large vertex */
    uint32_t grid;         /* workgroups of the grid-stride kernels */
    uint32_t delta;        /* SSSP near/far step of 64-slot states: a frontier vertex relaxes only
                              while its distance is below the query's threshold, which moves to
                              (smallest pending distance + delta) when nothing below it is left;
                              INF = plain BSP */
    uint32_t delta1;       /* the same for one-query states */
    uint32_t cc_mode;      /* CC: 0 label propagation (BSP supersteps), 1 union-find */
    uint32_t sssp_states;  /* serving: concurrent 64-slot SSSP states per graph */
};
Tuning default_tuning();

}  // namespace iprb

/* a graph resident in HBM (opaque in the C header) */
struct ipr_bgraph {
    int device;
    uint64_t V, E;
    uint64_t* oidx;       /* out-CSR offsets, V+1 */
    uint32_t* oadj;       /* out-CSR targets, E */
    uint32_t* owgt;       /* arc weights or NULL (weight 1) */
    uint64_t* iidx;       /* transpose offsets, V+1 */
    uint32_t* iadj;       /* transpose sources, ascending within each list */
    float* invdeg;        /* 1/out-degree in FP32, 0 for vertices without out-arcs */
    float* inw;           /* per transpose arc: invdeg of its source */
    uint32_t nhub, nchunk;
    uint32_t* hub_v;      /* vertices with in-degree > HUB, ascending */
    uint32_t* hub_c0;     /* first chunk of each hub, nhub+1 */
    uint32_t* chunk_h;    /* hub of each chunk */
    uint64_t nseg;        /* pull segments */
    uint32_t* seg_v0;     /* first vertex of each pull segment, nseg+1 */
    uint64_t nseg64;      /* fixed 64-vertex segments (result extraction) */
    unsigned long long inf_digest;   /* digest of the all-INF uint32 vector */
    uint64_t bytes;
    iprb::Tuning tune;
    hipStream_t stream;   /* stream of the one-shot entry points */
};

namespace iprb {

void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
double now_s();

/* host copy of the per-vertex digest term (the device one is identical) */
inline uint64_t digest_mix(uint64_t z)
{
    z ^= z >> 30; z *= 0xbf58476d1ce4e5b9ULL;
    z ^= z >> 27; z *= 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return z;
}

/* What one push superstep reports (read back to pinned host memory). */
struct TickOut {
    unsigned long long orm;       /* batched: OR of next-frontier masks (slots still running) */
    unsigned long long relax;     /* (query, arc) relaxations the next superstep issues */
    uint32_t nsmall, nlarge;      /* next frontier size */
    unsigned long long busy;      /* supersteps so far that left work for another */
    unsigned long long dig[64];   /* batched: per-slot digest of dist (telescoped) */
};

/* one accumulator on a 128-byte line of its own: device atomics on one line
 * are performed one after another (about 12 ns each on MI355X, whatever the
 * word), so the copies that spread an accumulator's atomics must not share
 * lines */
template <typename T> struct alignas(128) Lined {
    T v;
};

struct PushCtl {
    uint32_t nsmall[2], nlarge[2];
    Lined<unsigned long long> orm[64];
    Lined<unsigned long long> relax[64];
    Lined<unsigned long long> dig[SPREAD][64];
    uint32_t thr[64];                 /* near/far threshold per query */
    Lined<uint32_t> fmin[SPREAD][64]; /* smallest distance in the next frontier, per query */
    Lined<unsigned long long> pdef[SPREAD];  /* queries that parked (far) entries this superstep */
    Lined<unsigned long long> prem[SPREAD];  /* queries with parked entries left after a wake-up scan */
    unsigned long long pleft;         /* queries that may have parked entries */
    unsigned long long react;         /* queries whose threshold moved: wake parked entries below it */
    uint32_t pmin[64];                /* smallest parked distance per query (lower bound) */
    Lined<uint32_t> pdmin[SPREAD][64];  /* smallest distance parked this superstep */
    Lined<uint32_t> pscan[SPREAD][64];  /* smallest distance left parked after a wake-up scan */
    uint32_t step, pad;               /* superstep counter (frontier stamps) */
    unsigned long long relax_total;   /* cumulative over the run */
    unsigned long long busy_total;    /* supersteps of the run that left work for another */
    TickOut out;                      /* out.relax: relax_total after the superstep */
};

struct PushState {
    ipr_bgraph* g = nullptr;
    bool batched = false;
    uint32_t* val = nullptr;           /* scalar: V; batched: V*64 (dist) */
    uint32_t* mbox = nullptr;          /* same shape as val */
    unsigned long long* act = nullptr; /* batched: per-vertex mask of slots active this superstep */
    unsigned long long* vmask = nullptr; /* batched: per-vertex mask of slots whose dist is set */
    uint8_t* touched = nullptr;        /* V rounded up to 1024, zero outside supersteps */
    uint8_t* tgran = nullptr;          /* one byte per 256 vertices of the touched map */
    void* pend = nullptr;              /* parked (far) entries: byte per vertex, or slot mask */
    uint32_t* fst = nullptr;           /* superstep in which a vertex was last put in the next frontier */
    uint32_t* small[2] = {nullptr, nullptr};
    uint32_t* large[2] = {nullptr, nullptr};
    uint32_t* lpre = nullptr;          /* chunk prefix of the current large list */
    uint64_t large_cap = 0;
    PushCtl* ctl = nullptr;
    TickOut* hout = nullptr;           /* pinned, [2] (one per parity) */
    hipStream_t st = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipGraphExec_t gx[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  /* [weight-0][parity] */
    /* runs of 2^i supersteps in one graph (i = 1..6): [weight-0][first parity][i] */
    hipGraphExec_t gxn[2][2][7] = {};
    unsigned long long* dsum = nullptr;
    int cur = 0;                       /* frontier parity of the next superstep */
    uint32_t wzero = 0;                /* 1 while running CC (every arc weighs 0) */
    uint64_t bytes = 0;
    uint64_t supersteps = 0;
    int use_graph = 1;

    static int create(ipr_bgraph* g, bool batched, int use_graph, PushState** out);
    ~PushState();
    static uint64_t footprint(const ipr_bgraph* g, bool batched);
    /* scalar runs: reset values and seed the first frontier */
    int start_sssp(uint32_t source);
    int start_cc();                    /* label propagation (BSP supersteps) */
    int cc_union_find();               /* CC labels by union-find into val; ev[0] marks the end */
    int cc(uint64_t max_supersteps);   /* CC by the graph's cc_mode, to the end (blocking) */
    /* batched: inject distance 0 at sources for slots; clear released slots */
    int admit(int n, const int* slots, const uint32_t* sources);
    int release(unsigned long long mask);
    /* k supersteps, replayed as graphs of 64, 32, ... 1 supersteps; the last
     * one's TickOut lands in hout[*parity] when ev[*parity] completes */
    int steps(uint64_t k, int* parity);
    int wait(int parity);
    /* copy results to the host: scalar values (V) or one batched column */
    int values(uint32_t* host) ;
    int column(int slot, uint32_t* host);
    int columns(int n, uint32_t* host);  /* slots 0..n-1, query-major */
    int digest(unsigned long long* d);   /* scalar: digest of val */
    /* scalar: supersteps in doubling chunks until the frontier is empty */
    int run_to_end(uint64_t max_supersteps);
};

struct PprSlots {
    unsigned long long live, fresh, ext, pad;
    uint32_t seed[64];
    double invS[64], Dn[64], S[64], Sprev[64];
};

struct PprOut {
    double res2[128];              /* residual (L1 change of the last iteration) at [2*slot] */
    unsigned long long dig[64];    /* digest of the FP32 result */
};

struct PprState {
    ipr_bgraph* g = nullptr;
    int B = 1;                     /* 1 (scalar) or 64 (batched) */
    float* y[2] = {nullptr, nullptr};
    double* partS = nullptr;       /* [nseg][B] */
    double* partD = nullptr;
    float* hubpart = nullptr;      /* [nchunk][B] */
    double* hubS = nullptr;        /* [nhub][B] */
    double* hubD = nullptr;
    double* chain = nullptr;       /* [(CHAINS+1)][B][2] */
    double* res = nullptr;         /* [B][2] */
    unsigned long long* dig = nullptr; /* [SPREAD][B] */
    PprSlots* slots = nullptr;
    PprOut* out = nullptr;         /* device */
    PprOut* hout = nullptr;        /* pinned */
    float** outp = nullptr;        /* device array of B result pointers (extraction) */
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    hipGraphExec_t gx[2] = {nullptr, nullptr};
    uint64_t tick = 0;             /* iterations run; tick parity selects the y buffer read */
    uint64_t bytes = 0;
    int use_graph = 1;

    static int create(ipr_bgraph* g, int B, int use_graph, PprState** out);
    ~PprState();
    static uint64_t footprint(const ipr_bgraph* g, int B);
    /* admit seeds into slots (fresh: the next iteration starts from e_seed); release slots */
    int update(int nadmit, const int* slots, const uint32_t* seeds, unsigned long long release);
    int step();                     /* one iteration for every live slot */
    /* finish slots in mask after `first_mask` slots ran exactly one iteration:
     * residual and digest into hout (after wait()), results to device buffers */
    int extract(unsigned long long mask, unsigned long long first_mask, float* const* dev_out);
    int record();                   /* event after the enqueued work */
    int wait();
    bool done();
};

/* CC / SSSP digest of a device uint32 vector, and PPR digest of a host FP32 vector */
unsigned long long digest_u32_host(const uint32_t* v, uint64_t n);
unsigned long long digest_f32_host(const float* v, uint64_t n);

}  // namespace iprb

#endif
