/*
 * ipregel_kernels.h - the superstep kernels that call the application hooks.
 * Included by iPregel.h, i.e. instantiated once, in the application's
 * translation unit, with the application's message and value types.
 *
 * gfx950: 64-lane wavefronts, 256-thread workgroups (4 wavefronts).
 *
 * Push variants (PUSH, SPREAD), per superstep s:
 *   compute  every vertex of the work set (all V at superstep 0 and in PUSH,
 *            the frontier list in SPREAD) runs ip_compute; a broadcast is
 *            recorded in bslot[bpos[v]] = {s+1, message} (bpos: records by
 *            out-degree class for SPREAD, identity otherwise); broadcasters
 *            with more than large_min out-arcs are also listed, with their
 *            first arc chunk, in lbins/lbase (one atomic per wavefront).
 *   gather   (dense supersteps: broadcasters' out-arcs > E/dir_alpha) every
 *            vertex folds the records of its in-neighbours over the exact
 *            transpose, arc-balanced across wavefronts: same combined message
 *            and same receivers as the push; lists split between wavefronts
 *            are delivered with the combiner's atomic. With few live arcs a
 *            bitmap of this superstep's broadcasters is read first.
 *   expand   (sparse supersteps) large broadcasters one workgroup per chunk
 *            of large_chunk arcs; the rest of the work set by a workgroup
 *            prefix sum over the out-degrees of 256 vertices, so the
 *            workgroup's threads take consecutive arcs (load balance for the
 *            skewed R-MAT degrees). A delivery combines into mbox[(s+1)&1]
 *            with a hardware atomic (declared combiner) or a CAS loop around
 *            ip_combine and raises recv[(s+1)&1][u]
 *            (src/combiner_spread_postamble.h:46-98).
 *   The next frontier is compacted from recv (libipregel_hip.so) in
 *   ascending runs; PUSH's swap turns recv into has_message instead.
 * Pull variant (PULL): one kernel per superstep fetches superstep s-1's
 *   broadcasts over the in-lists and runs ip_compute for superstep s on the
 *   same vertex (broadcast slots double-buffered by superstep parity).
 * Counters are reduced per workgroup and added to one of IPR_CSLOTS slots.
 */
#ifndef IPREGEL_AMD_KERNELS_H
#define IPREGEL_AMD_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits>
#include <type_traits>

#include "ipregel_control.h"

#define IPR_BLOCK 256
#define IPR_WAVE 64
#define IPR_WAVES (IPR_BLOCK / IPR_WAVE)
#ifndef IPR_GATHER_HOT_BYTES
#define IPR_GATHER_HOT_BYTES 32768   /* LDS per gather workgroup for the hottest broadcast records */
#endif
#ifndef IPR_GATHER_UNROLL
#define IPR_GATHER_UNROLL 4   /* gathers in flight per lane (memory-level parallelism) */
#endif

/* device copy of the engine state, read by the device-side API functions
 * (constant memory: scalar loads, cached, never aliased by the kernels' stores) */
__constant__ ipr_dev ipr_d_dev;

/* the superstep, for ip_get_superstep() in application code: every kernel
 * that can run application code (ip_compute, ip_combine) stores it here first,
 * each thread the same value, so the call costs an LDS read instead of two
 * dependent global loads */
__shared__ uint64_t ipr_sm_superstep;

/* device-side API misuse, reported by the driver after the launch (bits of
 * ipr_ctl.error; the driver's report_device_error prints one message each) */
enum {
    IPR_ERR_SEND_IN_PULL = 1,        /* ip_send_message in a single-broadcast variant */
    IPR_ERR_MIXED_BROADCAST = 2,     /* ip_broadcast_weighted with another broadcast */
    IPR_ERR_WEIGHTED_IN_PULL = 4,    /* ip_broadcast_weighted in a single-broadcast variant */
    IPR_ERR_SEND_OUT_OF_RANGE = 8    /* ip_send_message to an id outside [0, V) */
};

__device__ inline void ipr_raise_error(uint32_t code)
{
    atomicOr(&ipr_d_dev.ctl->error, code);
}

/* ------------------------------------------------------------------------ */
/* wavefront and workgroup helpers                                           */
/* ------------------------------------------------------------------------ */

__device__ inline uint32_t ipr_lane(void) { return __lane_id(); }

__device__ inline uint64_t ipr_wave_sum(uint64_t x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ inline uint32_t ipr_wave_incl_scan(uint32_t x)
{
    uint32_t lane = ipr_lane();
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

/* workgroup exclusive scan of x (all threads); *total gets the sum.
 * `scratch` holds IPR_WAVES entries of LDS. Ends with a barrier. */
__device__ inline uint32_t ipr_block_excl_scan(uint32_t x, uint32_t* scratch, uint32_t* total)
{
    const uint32_t lane = ipr_lane(), wave = threadIdx.x / IPR_WAVE;
    uint32_t incl = ipr_wave_incl_scan(x);
    if (lane == IPR_WAVE - 1) scratch[wave] = incl;
    __syncthreads();
    uint32_t before = 0, sum = 0;
    for (uint32_t w = 0; w < IPR_WAVES; w++) {
        uint32_t t = scratch[w];
        if (w < wave) before += t;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return before + incl - x;
}

/* add per-thread counts a, b, c to counters ia, ib, ic: one atomic per
 * workgroup and counter, spread over IPR_CSLOTS slots. Every thread of the
 * workgroup must call it. */
__device__ inline void ipr_block_count3(const ipr_dev& d, int ia, uint64_t a, int ib, uint64_t b,
                                        int ic, uint64_t c)
{
    __shared__ uint64_t red[3][IPR_WAVES];
    const uint32_t lane = ipr_lane(), wave = threadIdx.x / IPR_WAVE;
    a = ipr_wave_sum(a);
    b = ipr_wave_sum(b);
    c = ipr_wave_sum(c);
    if (lane == 0) { red[0][wave] = a; red[1][wave] = b; red[2][wave] = c; }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint64_t t = 0;
        for (int w = 0; w < IPR_WAVES; w++) t += red[threadIdx.x][w];
        int id = threadIdx.x == 0 ? ia : (threadIdx.x == 1 ? ib : ic);
        if (t && id >= 0)
            atomicAdd((unsigned long long*)&d.cnt[IPR_CSLOT(id, blockIdx.x % IPR_CSLOTS)],
                      (unsigned long long)t);
    }
    __syncthreads();
}

template <typename T> __host__ __device__ inline bool ipr_same_bits(const T& a, const T& b)
{
    const unsigned char* pa = (const unsigned char*)&a;
    const unsigned char* pb = (const unsigned char*)&b;
    for (size_t i = 0; i < sizeof(T); i++)
        if (pa[i] != pb[i]) return false;
    return true;
}

/* ------------------------------------------------------------------------ */
/* combiner policies                                                         */
/* ------------------------------------------------------------------------ */

template <typename T, int KIND, typename Enable = void> struct ipr_comb;

/* atomic min/max on any 4- or 8-byte type by compare-and-swap on its bits */
template <typename T>
__device__ inline T ipr_atomic_minmax_cas(T* p, T m, bool is_min)
{
    typedef typename std::conditional<sizeof(T) == 4, unsigned int, unsigned long long>::type B;
    B* bp = (B*)p;
    B old = *bp;
    for (;;) {
        T cur;
        memcpy(&cur, &old, sizeof(T));
        bool better = is_min ? (m < cur) : (m > cur);
        if (!better) return cur;
        B nb;
        memcpy(&nb, &m, sizeof(T));
        B prev = atomicCAS(bp, old, nb);
        if (prev == old) return cur;
        old = prev;
    }
}

template <typename T> __device__ inline T ipr_atomic_min(T* p, T m)
{
    if constexpr (std::is_integral<T>::value && (sizeof(T) == 4 || sizeof(T) == 8))
        return atomicMin(p, m);
    else
        return ipr_atomic_minmax_cas(p, m, true);
}
template <typename T> __device__ inline T ipr_atomic_max(T* p, T m)
{
    if constexpr (std::is_integral<T>::value && (sizeof(T) == 4 || sizeof(T) == 8))
        return atomicMax(p, m);
    else
        return ipr_atomic_minmax_cas(p, m, false);
}

/* MIN: identity-initialised slot; the first delivery is the one that finds
 * the identity (or carries it). */
template <typename T> struct ipr_comb<T, IPR_COMBINER_MIN> {
    typedef T slot_t;
    static constexpr int exact_gather = std::is_integral<T>::value;
    /* deliver() answers true only for the first delivery to a slot (and for
     * a message equal to the empty value), so the receiver's flag is
     * raised without reading it first */
    static constexpr int first_once = 1;
    __host__ __device__ static T ident() { return std::numeric_limits<T>::has_infinity ? std::numeric_limits<T>::infinity() : std::numeric_limits<T>::max(); }
    __device__ static slot_t empty() { return ident(); }
    __device__ static T value(slot_t s) { return s; }
    __device__ static void fold(T* acc, T m) { if (m < *acc) *acc = m; }
    /* returns true when this delivery may be the first one (waiting for the
     * atomic's old value measured faster than raising the receiver's flag on
     * every lowering delivery: fewer flag reads) */
    __device__ static bool deliver(slot_t* p, T m)
    {
        T cur = *p;                       /* stale reads only overestimate */
        if (m < cur) return ipr_atomic_min(p, m) == ident();
        return cur == ident();            /* m == identity on an empty slot */
    }
    /* the same without reading the slot first: every delivery is an atomic,
     * one round trip less for those that lower the slot (measured faster in
     * the fused sparse kernel, where a receiver's flag is raised after) */
    __device__ static bool deliver_direct(slot_t* p, T m) { return ipr_atomic_min(p, m) == ident(); }
    /* single writer (gather): no atomics */
    __device__ static void merge(slot_t* p, T m) { if (m < *p) *p = m; }
};

template <typename T> struct ipr_comb<T, IPR_COMBINER_MAX> {
    typedef T slot_t;
    static constexpr int exact_gather = std::is_integral<T>::value;
    /* deliver() answers true only for the first delivery to a slot (and for
     * a message equal to the empty value), so the receiver's flag is
     * raised without reading it first */
    static constexpr int first_once = 1;
    __host__ __device__ static T ident() { return std::numeric_limits<T>::has_infinity ? -std::numeric_limits<T>::infinity() : std::numeric_limits<T>::lowest(); }
    __device__ static slot_t empty() { return ident(); }
    __device__ static T value(slot_t s) { return s; }
    __device__ static void fold(T* acc, T m) { if (m > *acc) *acc = m; }
    __device__ static bool deliver(slot_t* p, T m)
    {
        T cur = *p;
        if (m > cur) return ipr_atomic_max(p, m) == ident();
        return cur == ident();
    }
    __device__ static bool deliver_direct(slot_t* p, T m) { return ipr_atomic_max(p, m) == ident(); }
    __device__ static void merge(slot_t* p, T m) { if (m > *p) *p = m; }
};

/* SUM: zero-initialised slot, every delivery adds; firstness from the claim
 * bitmap alone. */
template <typename T> struct ipr_comb<T, IPR_COMBINER_SUM> {
    typedef T slot_t;
    static constexpr int exact_gather = std::is_integral<T>::value;
    /* deliver() answers true for every delivery */
    static constexpr int first_once = 0;
    __host__ __device__ static T ident() { return (T)0; }
    __device__ static slot_t empty() { return (T)0; }
    __device__ static T value(slot_t s) { return s; }
    __device__ static void fold(T* acc, T m) { *acc += m; }
    __device__ static bool deliver(slot_t* p, T m)
    {
        atomicAdd(p, m);
        return true;
    }
    __device__ static bool deliver_direct(slot_t* p, T m) { return deliver(p, m); }
    __device__ static void merge(slot_t* p, T m) { *p += m; }
};

/* GENERIC: 32-bit message packed with a presence flag in one 64-bit slot,
 * combined with the application's ip_combine in a CAS loop
 * (src/combiner_postamble.h:41-52 and :54-79 combine the same way, under a lock for
 * the first message). */
template <typename T> struct ipr_comb<T, IPR_COMBINER_GENERIC> {
    static_assert(sizeof(T) == 4,
                  "without -DIP_COMBINER_MIN/_MAX/_SUM, messages must be 4 bytes");
    typedef unsigned long long slot_t;
    static constexpr int exact_gather = 0;
    /* deliver() answers true only for the first delivery to a slot (and for
     * a message equal to the empty value), so the receiver's flag is
     * raised without reading it first */
    static constexpr int first_once = 1;
    __host__ __device__ static T ident() { return T(); }
    __device__ static slot_t empty() { return 0ull; }
    __device__ static T value(slot_t s) { uint32_t b = (uint32_t)s; T t; memcpy(&t, &b, 4); return t; }
    __device__ static void fold(T* acc, T m) { ip_combine(acc, m); }
    __device__ static bool deliver_direct(slot_t* p, T m) { return deliver(p, m); }
    __device__ static bool deliver(slot_t* p, T m)
    {
        slot_t old = *p;
        for (;;) {
            T a;
            uint32_t bits;
            if (old >> 32) { a = value(old); ip_combine(&a, m); }
            else a = m;
            memcpy(&bits, &a, 4);
            slot_t nw = (1ull << 32) | bits;
            if (nw == old) return false;
            slot_t prev = atomicCAS(p, old, nw);
            if (prev == old) return (old >> 32) == 0;
            old = prev;
        }
    }
    __device__ static void merge(slot_t* p, T m)
    {
        T a = m;
        if (*p >> 32) { a = value(*p); ip_combine(&a, m); }
        uint32_t bits;
        memcpy(&bits, &a, 4);
        *p = (1ull << 32) | bits;
    }
};

typedef ipr_comb<IP_MESSAGE_TYPE, IPR_COMBINER> ipr_policy;
typedef ipr_policy::slot_t ipr_slot_t;

/* broadcast record of a push-variant vertex: the superstep it broadcast in
 * (plus one; bit 31 set for ip_broadcast_weighted) and its message; a
 * gathering receiver gets both with one load */
struct ipr_bslot_t {
    uint32_t step;
    IP_MESSAGE_TYPE msg;
};

/* Push apps choose per superstep between pushing and gathering: gather when
 * the broadcasters' out-arcs (counted by the compute kernel) times dir_alpha
 * exceed E. Every wavefront of the delivery kernels derives the same answer
 * from the 64 counter slots; the finalize kernel counts gathered supersteps
 * by the same rule. */
__device__ inline uint64_t ipr_bcast_arcs(const ipr_dev& d)
{
    uint64_t x = d.cnt[IPR_CSLOT(IPR_C_EDGES, ipr_lane())];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ inline bool ipr_gathers(const ipr_dev& d)
{
    return d.dir_opt && ipr_bcast_arcs(d) * (uint64_t)d.dir_alpha > d.E;
}

/* a delivery kernel that finds its superstep is the other kernel's: when it
 * was launched alone (d.plan_alone), the superstep cannot be delivered, so
 * the run stops here and the host runs it again without a launch plan */
__device__ inline void ipr_plan_miss(const ipr_dev& d)
{
    if (d.plan_alone && blockIdx.x == 0 && threadIdx.x == 0) {
        d.ctl->plan_miss = 1;
        d.ctl->done = 1;
    }
}

/* a gathered superstep with few live arcs from few computed vertices reads
 * the broadcasters' bitmap before any record (wavefront-uniform): then most
 * of the records it skips are cold ones (few broadcasters holding many arcs
 * are hubs; the others' records are the ones that miss) */
__device__ inline bool ipr_gathers_live(const ipr_dev& d)
{
    if (!d.dir_opt || !d.g_live) return false;
    const uint64_t w = ipr_bcast_arcs(d);
    uint64_t n = d.cnt[IPR_CSLOT(IPR_C_WORKV, ipr_lane())];
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    return w * (uint64_t)d.dir_alpha > d.E && w * 100 < d.E * (uint64_t)d.live_pct &&
           n * 100 < d.V * (uint64_t)d.live_vpct;
}

/* a gather reads the broadcasters' bitmap when the superstep asks for it
 * and the bitmap kernel built it for this superstep (the host may leave that
 * launch out while runs do not use it); otherwise it reads every record, with
 * the same result, and tells the host (uniform across the launch) */
__device__ inline bool ipr_gather_bitmap(const ipr_dev& d)
{
    if (!ipr_gathers_live(d)) return false;
    ipr_ctl* ctl = d.ctl;
    if (ctl->live_step == ctl->superstep + 1) return true;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->live_missed = 1;
    return false;
}

/* position of vertex v's broadcast record */
__device__ __forceinline__ uint32_t ipr_brec(const ipr_dev& d, uint32_t v)
{
    return d.bpos ? d.bpos[v] : v;
}
#define IPR_BSTEP_WEIGHTED 0x80000000u

/* the 4-byte gather record of a broadcast in superstep s (integral
 * messages; a value that does not fit below IPR_REC4_ESC escapes) */
__device__ __forceinline__ uint32_t ipr_rec4(uint32_t stamp, IP_MESSAGE_TYPE m, bool weighted)
{
    uint32_t low = IPR_REC4_ESC;
    if constexpr (ipr_policy::exact_gather) {
        const uint64_t u = (uint64_t)m;
        if (u < IPR_REC4_ESC) low = (uint32_t)u;
    }
    return (stamp << IPR_REC4_SHIFT) | (weighted ? IPR_REC4_WEIGHTED : 0u) | low;
}

/* the record at position p: true when its vertex broadcast in superstep s
 * (want = s + 1, want4 its 4-byte stamp), with the message and whether it
 * is weighted. With 4-byte records on, the 8-byte record is written (and
 * read) only for messages that escape. */
__device__ __forceinline__ bool ipr_bcast_at(const ipr_dev& d, uint32_t p, uint32_t want, uint32_t want4,
                                             IP_MESSAGE_TYPE* m, bool* wf)
{
    const ipr_bslot_t* bs = (const ipr_bslot_t*)d.bslot;
    if (d.brec4) {
        const uint32_t q = d.brec4[p];
        const uint32_t low = q & IPR_REC4_ESC;
        *wf = (q & IPR_REC4_WEIGHTED) != 0;
        *m = (IP_MESSAGE_TYPE)low;
        if ((q >> IPR_REC4_SHIFT) != want4) return false;
        if (low == IPR_REC4_ESC) *m = bs[p].msg;
        return true;
    }
    const ipr_bslot_t b = bs[p];
    *wf = (b.step & IPR_BSTEP_WEIGHTED) != 0;
    *m = b.msg;
    return (b.step & ~IPR_BSTEP_WEIGHTED) == want;
}

/* raise the received flag of u: a plain byte store. Every writer stores the
 * same 1 and no kernel clears flags while deliveries run, so concurrent
 * stores need no atomic (the L2s write back dirty bytes only, so stores to
 * neighbouring flags from other XCDs are kept). A device atomic here was the
 * second of a delivery's two: random device atomics run at about 26G/s on
 * MI355X, plain random stores at about 84G/s. Callers that must learn
 * whether their delivery raised the flag use ipr_deliver_activates. */
__device__ inline void ipr_mark_recv(uint8_t* recv, uint32_t u)
{
    recv[u] = 1;
}

/* deliver one message to u in superstep s (push: concurrent senders) */
__device__ inline void ipr_deliver(const ipr_dev& d, uint64_t s, uint32_t u, IP_MESSAGE_TYPE m)
{
    ipr_slot_t* mb = (ipr_slot_t*)d.mbox[(s + 1) & 1];
    if (ipr_policy::deliver(mb + u, m)) ipr_mark_recv(d.recv[(s + 1) & 1], u);
}

/* what the expansion does for one arc: push variants deliver the message,
 * SPREAD_PULL only marks the out-neighbour as a broadcast target
 * (src/combiner_spread_single_broadcast_postamble.h ip_broadcast) */
__device__ inline void ipr_emit(const ipr_dev& d, uint64_t s, uint32_t u, IP_MESSAGE_TYPE m)
{
#if IPR_VARIANT == 3
    (void)m;
    ipr_mark_recv(d.recv[(s + 1) & 1], u);
#else
    ipr_deliver(d, s, u, m);
#endif
}

/* deliver one message to u in superstep s; true when this delivery is the
 * one that raised u's received flag (u joins the next frontier: the atomic
 * OR's old word tells, so exactly one delivery per receiver answers true) */
__device__ inline bool ipr_deliver_activates(const ipr_dev& d, uint64_t s, uint32_t u, IP_MESSAGE_TYPE m)
{
    ipr_slot_t* mb = (ipr_slot_t*)d.mbox[(s + 1) & 1];
    if (!ipr_policy::deliver_direct(mb + u, m)) return false;
    uint8_t* recv = d.recv[(s + 1) & 1];
    if (!ipr_policy::first_once && __atomic_load_n(recv + u, __ATOMIC_RELAXED)) return false;
    const uint32_t bit = 1u << (8 * (u & 3u));
    return (atomicOr((uint32_t*)(recv + (u & ~3u)), bit) & bit) == 0;
}

/* 1 in the fused sparse kernel, whose deliveries count the receivers they
 * activate (the next frontier's size and out-arcs, kept in the workgroup's
 * ipr_sm_fcount: count, arcs, largest out-degree); 0 in every other kernel
 * that runs application code. Each thread stores it before ip_compute. */
__shared__ uint32_t ipr_sm_counted;
__shared__ unsigned long long ipr_sm_fcount[3];

/* ip_send_message in push variants: immediate delivery */
__device__ inline void ipr_send_immediate(IP_VERTEX_ID_TYPE id, IP_MESSAGE_TYPE m)
{
    const ipr_dev& d = ipr_d_dev;
    if (!ipr_sm_counted) {
        ipr_deliver(d, ipr_sm_superstep, (uint32_t)id, m);
        return;
    }
    const uint32_t u = (uint32_t)id;
    if (ipr_deliver_activates(d, ipr_sm_superstep, u, m)) {
        const unsigned long long deg = d.out_idx[u + 1] - d.out_idx[u];
        atomicAdd(&ipr_sm_fcount[0], 1ull);
        atomicAdd(&ipr_sm_fcount[1], deg);
        atomicMax(&ipr_sm_fcount[2], deg);
    }
}

__device__ inline IP_MESSAGE_TYPE ipr_edge_msg(const uint32_t* wgt, IP_MESSAGE_TYPE m, bool weighted, uint64_t e)
{
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
    if (weighted) return (IP_MESSAGE_TYPE)(m + (IP_MESSAGE_TYPE)wgt[e]);
#else
    (void)wgt; (void)weighted; (void)e;
#endif
    return m;
}

/* ------------------------------------------------------------------------ */
/* vertex proxy                                                              */
/* ------------------------------------------------------------------------ */

__device__ inline void ipr_fill_vertex(const ipr_dev& d, struct ip_vertex_t* x, uint32_t v)
{
    x->id = (IP_VERTEX_ID_TYPE)v;
    x->value = ((const IP_VALUE_TYPE*)d.val)[v];
#ifdef IP_NEEDS_OUT_NEIGHBOUR_COUNT
    x->out_neighbour_count = (IP_NEIGHBOUR_COUNT_TYPE)(d.out_idx[v + 1] - d.out_idx[v]);
#endif
#ifdef IP_NEEDS_OUT_NEIGHBOUR_IDS
    /* out_ids holds IP_VERTEX_ID_TYPE entries: the 32-bit adjacency itself in
     * _32 builds, a 64-bit copy in _64 builds */
    x->out_neighbours = (IP_VERTEX_ID_TYPE*)d.out_ids + d.out_idx[v];
#endif
#ifdef IP_NEEDS_IN_NEIGHBOUR_COUNT
    x->in_neighbour_count = (IP_NEIGHBOUR_COUNT_TYPE)(d.in_idx[v + 1] - d.in_idx[v]);
#endif
#ifdef IP_NEEDS_IN_NEIGHBOUR_IDS
    x->in_neighbours = (IP_VERTEX_ID_TYPE*)d.in_ids + d.in_idx[v];
#endif
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
    x->out_neighbour_weights = (IP_EDGE_WEIGHT_TYPE*)(d.out_wgt + d.out_idx[v]);
#endif
    x->ipr_bcast = false;
    x->ipr_bcast_weighted = false;
}

__device__ inline void ipr_store_value(const ipr_dev& d, uint32_t v, const IP_VALUE_TYPE& before,
                                       const IP_VALUE_TYPE& after)
{
    if (!ipr_same_bits(before, after)) ((IP_VALUE_TYPE*)d.val)[v] = after;
}

/* ------------------------------------------------------------------------ */
/* init                                                                      */
/* ------------------------------------------------------------------------ */

__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_init(ipr_dev d)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < d.V; v += stride) {
        ((IP_VALUE_TYPE*)d.val)[v] = IP_VALUE_TYPE();
        d.act[v] = 1;  /* ip_init_vertex_range, e.g. src/combiner_postamble.h:94 */
        d.hm[v] = 0;
#if IPR_VARIANT == 1
        ((ipr_bslot_t*)d.bcm[0])[v].step = 0;
        ((ipr_bslot_t*)d.bcm[1])[v].step = 0;
#elif IPR_VARIANT == 3
        ((ipr_bslot_t*)d.bslot)[v].step = 0;
        d.recv[0][v] = 0;
        d.recv[1][v] = 0;
#else
        ((ipr_slot_t*)d.mbox[0])[v] = ipr_policy::empty();
        ((ipr_slot_t*)d.mbox[1])[v] = ipr_policy::empty();
        /* with 4-byte records the stamp lives there (cleared below): the
         * 8-byte record is read only for an escaping message, never for
         * its step */
        if (!d.brec4) ((ipr_bslot_t*)d.bslot)[v].step = 0;
        d.recv[0][v] = 0;
        d.recv[1][v] = 0;
#endif
        if (d.brec4) d.brec4[v] = 0;
    }
}

/* ------------------------------------------------------------------------ */
/* push variants: compute                                                    */
/* ------------------------------------------------------------------------ */

/* vertices per thread per round of the compute kernel. The sweep is bound by
 * load latency, not bytes: one vertex at a time, a wavefront waits for its
 * id, then its state, then its broadcast bookkeeping before the next vertex
 * starts. U at a time, each of those loads is issued for all U vertices
 * before the first is used. No branch may separate them: at a join the
 * compiler waits for every load in flight, so indices past the end are
 * clamped to the last entry (their results are dropped and their stores
 * skipped), and the list or full-sweep choice is a template parameter. */
#ifndef IPR_COMPUTE_U
#define IPR_COMPUTE_U 4
#endif
static_assert(IPR_COMPUTE_U >= 1 && IPR_COMPUTE_U <= 8, "IPR_COMPUTE_U out of range");

/* how the compute kernel finds its vertices: all V (superstep 0, PUSH), a
 * frontier list, or all V filtered by the frontier's raised receive flags
 * (SPREAD, after a superstep that gathered: the compaction only counted
 * them); in the last two every vertex run has a message */
enum { IPR_CR_SWEEP = 0, IPR_CR_LIST = 1, IPR_CR_FLAGS = 2 };

/* one round of U vertices per thread: entries base + u*IPR_BLOCK + thread
 * (coalesced per u) of the frontier list or of 0..V-1 */
template <int MODE>
__device__ __forceinline__ void ipr_compute_round(const ipr_dev& d, uint64_t s, const uint32_t* list, uint8_t* flags,
                                                  uint64_t n, uint64_t base, uint64_t& n_active, uint64_t& n_computed,
                                                  uint64_t& n_edges)
{
    constexpr int U = IPR_COMPUTE_U;
    constexpr bool LIST = MODE == IPR_CR_LIST, FLAGS = MODE == IPR_CR_FLAGS, FRONT = LIST || FLAGS;
    ipr_slot_t* mcur = (ipr_slot_t*)d.mbox[s & 1];
    ipr_bslot_t* bs = (ipr_bslot_t*)d.bslot;
    const uint32_t stamp4 = (uint32_t)(s % IPR_REC4_WINDOW) + 1u;
    /* the ids */
    uint32_t vs[U];
    bool run[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint64_t i = base + (uint64_t)u * IPR_BLOCK + threadIdx.x;
        run[u] = i < n;
        const uint64_t ic = run[u] ? i : n - 1;   /* base < n, so n >= 1 */
        vs[u] = LIST ? list[ic] : (uint32_t)ic;
    }
    if (FLAGS) {
        /* the frontier's flags, all U loaded before the first is used; the
         * others' state is not loaded */
        uint8_t fl[U];
#pragma unroll
        for (int u = 0; u < U; ++u) fl[u] = flags[vs[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) run[u] = run[u] && fl[u] != 0;
    }
    /* the state: message, value, and a listed vertex's broadcast bookkeeping
     * with it (most frontier vertices broadcast) */
    struct ip_vertex_t xs[U];
    bool hm_before[U];
    uint64_t o0[U], o1[U];
    uint32_t rp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t v = vs[u];
        struct ip_vertex_t& x = xs[u];
        o0[u] = 0;
        o1[u] = 0;
        rp[u] = 0;
        hm_before[u] = false;
        if (FLAGS && !run[u]) continue;
        if (IPR_VARIANT == IPR_VARIANT_SPREAD) {
            /* after the first superstep every frontier vertex has a message,
             * read straight from its mailbox (reset below) */
            x.active = true;
            x.has_message = FRONT;
            if (FRONT) x.message = ipr_policy::value(mcur[v]);
        } else if (IPR_VARIANT == IPR_VARIANT_SPREAD_PULL) {
            /* targets: message fetched at the end of the previous superstep */
            x.active = true;
            x.has_message = d.hm[v] != 0;
            if (x.has_message) x.message = ((const IP_MESSAGE_TYPE*)d.msg)[v];
        } else {
            /* PUSH: a full sweep that runs only the vertices with active or
             * has_message; the others' values are not loaded */
            x.active = d.act[v] != 0;
            x.has_message = d.hm[v] != 0;
            if (x.has_message) x.message = ((const IP_MESSAGE_TYPE*)d.msg)[v];
            run[u] = run[u] && (x.active || x.has_message);
            x.active = true;
            if (!run[u]) continue;
        }
        if (FRONT) {
            o0[u] = d.out_idx[v];
            o1[u] = d.out_idx[v + 1];
            rp[u] = ipr_brec(d, v);
        }
        ipr_fill_vertex(d, &x, v);
        hm_before[u] = x.has_message;
    }
    if (IPR_VARIANT == IPR_VARIANT_SPREAD && FRONT) {
        /* mailboxes are reset for reuse two supersteps on, and so are the
         * flags of a frontier left as flags */
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (run[u]) {
                mcur[vs[u]] = ipr_policy::empty();
                if (FLAGS) flags[vs[u]] = 0;
            }
    }
    /* the computes, then their value stores: a store between two computes
     * would make the compiler reload any application global the next one
     * reads (a store may alias it), and wait for the store with it */
    IP_VALUE_TYPE before[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!run[u]) continue;
        before[u] = xs[u].value;
        ip_compute(&xs[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!run[u]) continue;
        const uint32_t v = vs[u];
        struct ip_vertex_t& x = xs[u];
        ipr_store_value(d, v, before[u], x.value);
        n_computed++;
        if (IPR_VARIANT == IPR_VARIANT_PUSH) {
            if (!x.active) d.act[v] = 0;
            else if (!d.act[v]) d.act[v] = 1;
            if (hm_before[u] != x.has_message) d.hm[v] = x.has_message;
            n_active += x.active;
        }
    }
    /* a full sweep loads the bookkeeping only for the vertices that
     * broadcast, all of them before the first record is stored */
    if (!FRONT) {
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (run[u] && xs[u].ipr_bcast) {
                o0[u] = d.out_idx[vs[u]];
                o1[u] = d.out_idx[vs[u] + 1];
                rp[u] = ipr_brec(d, vs[u]);
            }
    }
    /* the records: every figure first, then the stores (a store before the
     * next vertex's bookkeeping is read would be waited for with it) */
    uint32_t large[U], chunks[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        large[u] = 0;
        chunks[u] = 0;
        if (!run[u] || !xs[u].ipr_bcast) continue;
        /* recorded even without out-arcs: with directed=false a SPREAD_PULL
         * in-list may still name this vertex */
        const uint64_t deg = o1[u] - o0[u];
        n_edges += deg;
        if (deg > d.large_min) {
            large[u] = 1;
            chunks[u] = (uint32_t)((deg + d.large_chunk - 1) / d.large_chunk);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!run[u] || !xs[u].ipr_bcast) continue;
        const struct ip_vertex_t& x = xs[u];
        ipr_bslot_t rec;
        rec.step = (uint32_t)(s + 1) | (x.ipr_bcast_weighted ? IPR_BSTEP_WEIGHTED : 0u);
        rec.msg = x.ipr_bcast_msg;
        if (d.brec4) {
            const uint32_t r4 = ipr_rec4(stamp4, rec.msg, x.ipr_bcast_weighted);
            d.brec4[rp[u]] = r4;
            if ((r4 & IPR_REC4_ESC) == IPR_REC4_ESC) bs[rp[u]] = rec;
        } else {
            bs[rp[u]] = rec;
        }
    }
    /* large broadcasters of this wavefront take consecutive entries and
     * consecutive chunk ranges of one slot (wavefront index mod IPR_CSLOTS)
     * with one atomic per u: slots on lines of their own keep the atomics
     * from queueing on one address (no workgroup barrier: a barrier would
     * wait for this round's stores before the next round) */
    const uint32_t lslot = (blockIdx.x * IPR_WAVES + threadIdx.x / IPR_WAVE) % IPR_CSLOTS;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint64_t lmask = __ballot(large[u]);
        if (!lmask) continue;
        const uint32_t lane = ipr_lane();
        const uint32_t el = (uint32_t)__popcll(lmask & ((1ull << lane) - 1ull));
        const uint32_t ci = ipr_wave_incl_scan(chunks[u]);
        const uint32_t ec = ci - chunks[u];
        const uint32_t tl = (uint32_t)__popcll(lmask), tc = __shfl(ci, IPR_WAVE - 1);
        unsigned long long lb = 0;
        if (lane == 0)
            lb = atomicAdd((unsigned long long*)&d.cnt[IPR_CSLOT(IPR_C_LPACK, lslot)],
                           ((unsigned long long)tl << IPR_LPACK_SHIFT) | tc);
        lb = __shfl(lb, 0);
        if (large[u]) {
            const uint64_t e = (uint64_t)lslot * d.lcap + (lb >> IPR_LPACK_SHIFT) + el;
            d.lbins[e] = vs[u];
            d.lbase[e] = (lb & IPR_LPACK_MASK) + ec;
        }
    }
}

/* SPREAD: src/combiner_spread_postamble.h:227-276 (superstep 0 over all V,
 * then the frontier) and :322-340 (the frontier's mailboxes become messages:
 * read straight from the mailbox, which is reset for reuse two supersteps
 * on). PUSH: src/combiner_postamble.h:151-164 (every vertex with active or
 * has_message). */
__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_compute_push(ipr_dev d)
{
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    const uint64_t s = ctl->superstep;
    ipr_sm_superstep = s;
    ipr_sm_counted = 0;
    uint64_t n = d.V;
    const uint32_t* list = nullptr;
    uint8_t* flags = nullptr;
    if ((IPR_VARIANT == IPR_VARIANT_SPREAD || IPR_VARIANT == IPR_VARIANT_SPREAD_PULL) && s > 0) {
        if (IPR_VARIANT == IPR_VARIANT_SPREAD && ctl->front_flags) {
            flags = d.recv[s & 1];
        } else {
            n = ctl->fsize[s & 1];
            list = d.front[s & 1];
        }
    }
    uint64_t n_active = 0, n_computed = 0, n_edges = 0;
    const uint64_t per_round = (uint64_t)IPR_BLOCK * IPR_COMPUTE_U;
    const uint64_t stride = (uint64_t)gridDim.x * per_round;
    if (list) {
        for (uint64_t base = (uint64_t)blockIdx.x * per_round; base < n; base += stride)
            ipr_compute_round<IPR_CR_LIST>(d, s, list, flags, n, base, n_active, n_computed, n_edges);
    } else if (flags) {
        for (uint64_t base = (uint64_t)blockIdx.x * per_round; base < n; base += stride)
            ipr_compute_round<IPR_CR_FLAGS>(d, s, list, flags, n, base, n_active, n_computed, n_edges);
    } else {
        for (uint64_t base = (uint64_t)blockIdx.x * per_round; base < n; base += stride)
            ipr_compute_round<IPR_CR_SWEEP>(d, s, list, flags, n, base, n_active, n_computed, n_edges);
    }
    ipr_block_count3(d, IPR_C_ACT0 + (int)(s & 1), n_active, IPR_C_WORKV, n_computed, IPR_C_EDGES, n_edges);
}

/* ------------------------------------------------------------------------ */
/* push variants: expansion (sparse supersteps)                              */
/* ------------------------------------------------------------------------ */

__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_expand(ipr_dev d)
{
    __shared__ uint32_t s_ex[IPR_BLOCK];
    __shared__ uint64_t s_off[IPR_BLOCK];
    __shared__ IP_MESSAGE_TYPE s_msg[IPR_BLOCK];
    __shared__ uint8_t s_wf[IPR_BLOCK];
    __shared__ uint32_t scan[IPR_WAVES];
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    if (ipr_gathers(d)) {
        ipr_plan_miss(d);
        return;
    }
    const uint64_t s = ctl->superstep;
    ipr_sm_superstep = s;
    ipr_sm_counted = 0;
    const uint32_t want = (uint32_t)(s + 1);
    const uint32_t want4 = (uint32_t)(s % IPR_REC4_WINDOW) + 1u;

    /* large broadcasters: one workgroup per chunk of large_chunk arcs. They
     * registered in IPR_CSLOTS slots; the chunks of slot k follow those of
     * slots 0..k-1, so every wavefront scans the 64 slot totals (lane k
     * holding slot k) and finds a chunk's slot by ballot */
    const uint32_t lane = ipr_lane();
    const uint64_t mine = d.lcap ? d.cnt[IPR_CSLOT(IPR_C_LPACK, lane)] : 0;
    const uint64_t my_chunks = mine & IPR_LPACK_MASK;
    uint64_t incl = my_chunks;
    for (uint32_t o = 1; o < IPR_WAVE; o <<= 1) {
        const uint64_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    const uint64_t nC = __shfl(incl, IPR_WAVE - 1);
    for (uint64_t c = blockIdx.x; c < nC; c += gridDim.x) {
        const uint32_t k = (uint32_t)__ffsll((unsigned long long)__ballot(incl > c)) - 1;
        const uint64_t ck = c - (__shfl(incl, k) - __shfl(my_chunks, k));   /* chunk within slot k */
        const uint64_t nL = __shfl(mine, k) >> IPR_LPACK_SHIFT;
        const uint32_t* lbins = d.lbins + (uint64_t)k * d.lcap;
        const uint64_t* lbase = d.lbase + (uint64_t)k * d.lcap;
        uint64_t lo = 0, hi = nL;            /* last j with lbase[j] <= ck */
        while (hi - lo > 1) {
            uint64_t mid = (lo + hi) / 2;
            if (lbase[mid] <= ck) lo = mid; else hi = mid;
        }
        uint32_t v = lbins[lo];
        IP_MESSAGE_TYPE bm;
        bool wf;
        ipr_bcast_at(d, ipr_brec(d, v), want, want4, &bm, &wf);   /* a listed broadcaster */
        uint64_t off = d.out_idx[v], deg = d.out_idx[v + 1] - off;
        uint64_t beg = (ck - lbase[lo]) * d.large_chunk;
        uint64_t end = beg + d.large_chunk < deg ? beg + d.large_chunk : deg;
        for (uint64_t k = beg + threadIdx.x; k < end; k += IPR_BLOCK)
            ipr_emit(d, s, d.out_adj[off + k], ipr_edge_msg(d.out_wgt, bm, wf, off + k));
    }

    /* the rest of the work set: 256 vertices per workgroup, their arcs laid
     * end to end by a prefix sum and taken 256 at a time */
    uint64_t n = d.V;
    const uint32_t* list = nullptr;
    /* a frontier left as flags (after a superstep that gathered) has no
     * list: the broadcasters are found by a sweep of the records, as in
     * superstep 0 */
    if ((IPR_VARIANT == IPR_VARIANT_SPREAD || IPR_VARIANT == IPR_VARIANT_SPREAD_PULL) && s > 0 &&
        !(IPR_VARIANT == IPR_VARIANT_SPREAD && ctl->front_flags)) {
        n = ctl->fsize[s & 1];
        list = d.front[s & 1];
    }
    /* a sweep over all V with records in out-degree order (SPREAD's superstep
     * 0, or a frontier left as flags) takes i as a record position: the
     * records are read in order, and only a broadcaster's vertex id (bvert)
     * and offsets after */
    const bool bypos = !list && d.bvert;
    /* one tile of IPR_BLOCK entries from `base`: every thread of the
     * workgroup calls it */
    auto tile = [&](uint64_t base) {
        const uint64_t i = base + threadIdx.x;
        uint32_t deg = 0;
        if (i < n) {
            uint32_t v = list ? list[i] : (uint32_t)i;
            /* a frontier vertex's offsets are read with its record (one round
             * trip less); the full sweep of PUSH reads them only for
             * broadcasters */
            uint64_t off = 0, end = 0;
            if (list) {
                off = d.out_idx[v];
                end = d.out_idx[v + 1];
            }
            IP_MESSAGE_TYPE bm;
            bool wf;
            if (ipr_bcast_at(d, bypos ? (uint32_t)i : ipr_brec(d, v), want, want4, &bm, &wf)) {
                if (bypos) v = d.bvert[i];
                if (!list) {
                    off = d.out_idx[v];
                    end = d.out_idx[v + 1];
                }
                uint64_t dg = end - off;
                if (dg <= d.large_min) {
                    deg = (uint32_t)dg;
                    s_off[threadIdx.x] = off;
                    s_msg[threadIdx.x] = bm;
                    s_wf[threadIdx.x] = wf;
                }
            }
        }
        uint32_t total = 0;
        uint32_t ex = ipr_block_excl_scan(deg, scan, &total);
        s_ex[threadIdx.x] = ex;
        __syncthreads();
        for (uint32_t t0 = 0; t0 < total; t0 += IPR_BLOCK) {
            uint32_t t = t0 + threadIdx.x;
            if (t < total) {
                uint32_t lo = 0;                 /* last owner with start <= t */
                for (uint32_t step = IPR_BLOCK / 2; step > 0; step >>= 1)
                    if (s_ex[lo + step] <= t) lo += step;
                uint64_t e = s_off[lo] + (t - s_ex[lo]);
                ipr_emit(d, s, d.out_adj[e], ipr_edge_msg(d.out_wgt, s_msg[lo], s_wf[lo] != 0, e));
            }
        }
        __syncthreads();
    };
    if (bypos && d.brec4) {
        /* the sweep by record position reads 4 records per thread with one
         * 16-byte load, a window of 4 tiles per workgroup, and runs the tile
         * code only on tiles that hold a broadcaster (wavefront w's four
         * records per lane are tile w's): with few broadcasters (SSSP's
         * superstep 0) a window costs one load and one barrier. The records
         * are padded to 16, so the load stays inside the array; records past
         * V are masked. */
        __shared__ uint32_t s_any[2][IPR_WAVES];
        const uint64_t span = (uint64_t)IPR_BLOCK * 4;
        uint32_t par = 0;
        for (uint64_t w0 = (uint64_t)blockIdx.x * span; w0 < n; w0 += (uint64_t)gridDim.x * span, par ^= 1u) {
            const uint64_t p = w0 + (uint64_t)threadIdx.x * 4;
            bool hit = false;
            if (p < n) {
                const uint4 q = *(const uint4*)(d.brec4 + p);
                const uint32_t r[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                for (int k = 0; k < 4; k++) hit |= p + k < n && (r[k] >> IPR_REC4_SHIFT) == want4;
            }
            const bool any = __ballot(hit) != 0;
            if (ipr_lane() == 0) s_any[par][threadIdx.x / IPR_WAVE] = any;
            __syncthreads();
            for (uint32_t w = 0; w < IPR_WAVES; w++)
                if (s_any[par][w]) tile(w0 + (uint64_t)w * IPR_BLOCK);
        }
    } else {
        const uint64_t tile_stride = (uint64_t)gridDim.x * IPR_BLOCK;
        for (uint64_t base = (uint64_t)blockIdx.x * IPR_BLOCK; base < n; base += tile_stride) tile(base);
    }
}

/* ------------------------------------------------------------------------ */
/* spread variant: one fused kernel per sparse superstep                     */
/* ------------------------------------------------------------------------ */

#if IPR_VARIANT == 2
/* vertices whose receive flags one workgroup scans per pass of the fused
 * kernel (one 16-byte load of flags per thread) */
#define IPR_SCAN_SPAN (IPR_BLOCK * 16)

/* One tile of the fused sparse kernel: thread t computes vertex v (when
 * `valid`), the tile's broadcasts are laid end to end by a workgroup prefix
 * sum and delivered 256 arcs at a time; receivers a delivery activates are
 * counted with their out-degrees in *n_new / *f_arcs / *f_max. Every thread of
 * the workgroup calls it. */
__device__ inline void ipr_sparse_tile(const ipr_dev& d, uint64_t s, ipr_slot_t* mcur, uint32_t v, bool valid,
                                       uint32_t* s_ex, uint64_t* s_off, IP_MESSAGE_TYPE* s_msg, uint8_t* s_wf,
                                       uint32_t* scan, uint64_t* n_computed, uint64_t* n_edges, uint64_t* n_new,
                                       uint64_t* f_arcs, uint64_t* f_max)
{
    uint32_t deg = 0;
    if (valid) {
        struct ip_vertex_t x;
        x.active = true;
        x.has_message = true;
        x.message = ipr_policy::value(mcur[v]);
        mcur[v] = ipr_policy::empty();
        const uint64_t o0 = d.out_idx[v], o1 = d.out_idx[v + 1];
        ipr_fill_vertex(d, &x, v);
        const IP_VALUE_TYPE before = x.value;
        ip_compute(&x);
        ipr_store_value(d, v, before, x.value);
        (*n_computed)++;
        if (x.ipr_bcast) {
            *n_edges += o1 - o0;
            deg = (uint32_t)(o1 - o0);
            s_off[threadIdx.x] = o0;
            s_msg[threadIdx.x] = x.ipr_bcast_msg;
            s_wf[threadIdx.x] = x.ipr_bcast_weighted;
        }
    }
    uint32_t total = 0;
    const uint32_t ex = ipr_block_excl_scan(deg, scan, &total);
    s_ex[threadIdx.x] = ex;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < total; t += IPR_BLOCK) {
        uint32_t lo = 0;                         /* last owner with start <= t */
        for (uint32_t step = IPR_BLOCK / 2; step > 0; step >>= 1)
            if (s_ex[lo + step] <= t) lo += step;
        const uint64_t e = s_off[lo] + (t - s_ex[lo]);
        const uint32_t u = d.out_adj[e];
        IP_MESSAGE_TYPE m = s_msg[lo];
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
        /* loaded with the arc's target whether the broadcast is weighted or
         * not: a load under the branch would wait for the target first */
        const uint32_t w = d.out_wgt[e];
        if (s_wf[lo]) m = (IP_MESSAGE_TYPE)(m + (IP_MESSAGE_TYPE)w);
#endif
        /* u's offsets are loaded with the delivery, not after it: the
         * out-degree is wanted only if this delivery activates u, but waiting
         * for the activation first would add a round trip */
        const uint64_t u0 = d.out_idx[u], u1 = d.out_idx[u + 1];
        if (ipr_deliver_activates(d, s, u, m)) {
            const uint64_t du = u1 - u0;
            (*n_new)++;
            *f_arcs += du;
            *f_max = du > *f_max ? du : *f_max;
        }
    }
    __syncthreads();
}

/* A whole sparse superstep of the spread variant in one launch
 * (src/combiner_spread_postamble.h:160-431 phases 1-3): every vertex of the
 * frontier takes its combined message from the mailbox and runs ip_compute;
 * the broadcasts of a tile of 256 vertices are laid end to end by a workgroup
 * prefix sum and delivered 256 arcs at a time, raising the receivers' flags;
 * the deliveries that activate a receiver count it, with its out-degree, so
 * the next frontier's size and out-arc figures are known without a
 * compaction; the last workgroup ends the superstep.
 *
 * The frontier is the list front[s & 1] after a superstep of the general
 * kernels (ctl->front_flags == 0), else the raised flags recv[s & 1] left by
 * the previous fused kernel: each workgroup scans IPR_SCAN_SPAN vertices at a
 * time, clears the flags it finds and computes those vertices in ascending
 * order from LDS. No broadcast record is written (nothing reads one in a
 * sparse superstep); the 4-byte records are cleared at the end of each stamp
 * window as the compaction does.
 *
 * Large broadcasters are not split into chunks here: a frontier with a vertex
 * above fused_max_deg out-arcs, or with enough arcs to gather (or above
 * fused_max_arcs), is left to the general kernels: the kernel raises
 * ctl->bail and returns before touching anything, and the host rebuilds the
 * list from the flags if the frontier has no list. */
__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_sparse_step(ipr_dev d)
{
    __shared__ uint32_t s_ex[IPR_BLOCK];
    __shared__ uint64_t s_off[IPR_BLOCK];
    __shared__ IP_MESSAGE_TYPE s_msg[IPR_BLOCK];
    __shared__ uint8_t s_wf[IPR_BLOCK];
    __shared__ uint32_t scan[IPR_WAVES];
    __shared__ uint16_t s_ids[IPR_SCAN_SPAN];  /* offsets within the pass */
    ipr_ctl* ctl = d.ctl;
    /* the superstep: the chunk's base plus this kernel's place in it (a
     * kernel after a bail in the same chunk finds ctl->superstep behind) */
    const uint64_t s = ctl->chunk_base + d.step_in_chunk;
    if (ctl->done || ctl->bail || ctl->superstep != s) return;
    const uint64_t arcs = ctl->farcs[s & 1], maxdeg = ctl->fmaxdeg[s & 1];
    if (s == 0 || arcs > d.E || arcs > d.fused_max_arcs || maxdeg > d.fused_max_deg ||
        (d.dir_opt && arcs * (uint64_t)d.dir_alpha > d.E)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl->bail = 1;
        return;
    }
    const bool from_flags = ctl->front_flags != 0;
    ipr_sm_superstep = s;
    ipr_sm_counted = 1;
    if (threadIdx.x < 3) ipr_sm_fcount[threadIdx.x] = 0;
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->fused_steps++;
    ipr_slot_t* mcur = (ipr_slot_t*)d.mbox[s & 1];
    uint64_t n_computed = 0, n_edges = 0, n_new = 0, f_arcs = 0, f_max = 0;
    const uint64_t groups = (d.V + 15) / 16;     /* flag arrays are padded to 16 */
    /* the last superstep of a stamp window clears the 4-byte gather records
     * (padded to 16 like the flags) */
    if (d.brec4 && (s + 1) % IPR_REC4_WINDOW == 0) {
        uint4* clr = (uint4*)d.brec4;
        const uint4 z = make_uint4(0, 0, 0, 0);
        for (uint64_t i = (uint64_t)blockIdx.x * IPR_BLOCK + threadIdx.x; i < groups * 4;
             i += (uint64_t)gridDim.x * IPR_BLOCK)
            clr[i] = z;
    }
    if (!from_flags) {
        const uint64_t n = ctl->fsize[s & 1];
        const uint32_t* list = d.front[s & 1];
        for (uint64_t base = (uint64_t)blockIdx.x * IPR_BLOCK; base < n; base += (uint64_t)gridDim.x * IPR_BLOCK) {
            const uint64_t i = base + threadIdx.x;
            ipr_sparse_tile(d, s, mcur, i < n ? list[i] : 0u, i < n, s_ex, s_off, s_msg, s_wf, scan,
                            &n_computed, &n_edges, &n_new, &f_arcs, &f_max);
        }
    } else {
        uint4* flags = (uint4*)d.recv[s & 1];
        const uint64_t stride = (uint64_t)gridDim.x * IPR_BLOCK;
        uint4 qn = make_uint4(0, 0, 0, 0);
        if ((uint64_t)blockIdx.x * IPR_BLOCK + threadIdx.x < groups)
            qn = flags[(uint64_t)blockIdx.x * IPR_BLOCK + threadIdx.x];
        for (uint64_t g0 = (uint64_t)blockIdx.x * IPR_BLOCK; g0 < groups; g0 += stride) {
            const uint64_t g = g0 + threadIdx.x;
            const uint4 q = qn;
            /* the next pass's flags are loaded while this pass runs */
            qn = make_uint4(0, 0, 0, 0);
            if (g + stride < groups) qn = flags[g + stride];
            /* flags are 0 or 1, so each raised flag is one bit (0, 8, 16 or 24) */
            const uint32_t c = __popc(q.x) + __popc(q.y) + __popc(q.z) + __popc(q.w);
            if (c) flags[g] = make_uint4(0, 0, 0, 0);
            uint32_t total = 0;
            uint32_t pos = ipr_block_excl_scan(c, scan, &total);
            if (!total) continue;                /* uniform across the workgroup */
            const uint32_t words[4] = {q.x, q.y, q.z, q.w};
            for (int k = 0; k < 4; k++) {
                uint32_t w = words[k];
                while (w) {
                    const uint32_t bit = __ffs(w) - 1;
                    s_ids[pos++] = (uint16_t)(threadIdx.x * 16 + k * 4 + bit / 8);
                    w &= w - 1;
                }
            }
            __syncthreads();
            for (uint32_t t0 = 0; t0 < total; t0 += IPR_BLOCK) {
                const uint32_t t = t0 + threadIdx.x;
                ipr_sparse_tile(d, s, mcur, t < total ? (uint32_t)(g0 * 16 + s_ids[t]) : 0u, t < total, s_ex, s_off,
                                s_msg, s_wf, scan, &n_computed, &n_edges, &n_new, &f_arcs, &f_max);
            }
        }
    }
    /* the workgroup's figures: one atomic per counter, spread over the slots
     * (ip_send_message deliveries counted theirs in LDS) */
    __syncthreads();
    if (threadIdx.x == 0) {
        n_new += ipr_sm_fcount[0];
        f_arcs += ipr_sm_fcount[1];
        f_max = ipr_sm_fcount[2] > f_max ? ipr_sm_fcount[2] : f_max;
    }
    ipr_block_count3(d, IPR_C_FNEXT, n_new, IPR_C_WORKV, n_computed, IPR_C_EDGES, n_edges);
    {
        __shared__ uint64_t r_arcs[IPR_WAVES], r_max[IPR_WAVES];
        for (int o = IPR_WAVE / 2; o > 0; o >>= 1) {
            f_arcs += __shfl_xor(f_arcs, o);
            const uint64_t m = __shfl_xor(f_max, o);
            f_max = m > f_max ? m : f_max;
        }
        if (ipr_lane() == 0) { r_arcs[threadIdx.x / IPR_WAVE] = f_arcs; r_max[threadIdx.x / IPR_WAVE] = f_max; }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t a = 0, m = 0;
            for (int w = 0; w < IPR_WAVES; w++) { a += r_arcs[w]; m = r_max[w] > m ? r_max[w] : m; }
            const uint32_t slot = blockIdx.x % IPR_CSLOTS;
            if (a) atomicAdd((unsigned long long*)&d.cnt[IPR_CSLOT(IPR_C_FARCS, slot)], (unsigned long long)a);
            if (m) atomicMax((unsigned long long*)&d.cnt[IPR_CSLOT(IPR_C_FMAX, slot)], (unsigned long long)m);
        }
    }
    /* the last workgroup to finish ends the superstep; the next frontier is
     * left as raised flags */
    if (ipr_last_workgroup(d, gridDim.x) && threadIdx.x < IPR_WAVE)
        ipr_finalize_step(d, IPR_VARIANT_SPREAD, 1);
}
#endif /* IPR_VARIANT == 2 (spread) */

/* ------------------------------------------------------------------------ */
/* push variants: gather (dense supersteps)                                  */
/* ------------------------------------------------------------------------ */

/* Lane-cooperative fold for vertices of small in-degree (reassociating
 * pull): lane L holds vertex v (valid); the in-lists of the wavefront's 64
 * vertices are laid end to end and taken 64*U arcs at a time. For the U
 * chunks in flight the source ids are loaded first, then their broadcast
 * records (stamp and message in one load), so no load waits on another of
 * its kind. A segmented scan across the lanes folds each vertex's
 * arcs and the segment's last lane hands the result to the vertex's lane.
 * Must be reached by all lanes of the wavefront. */
__device__ inline bool ipr_wave_pull_fold(const ipr_dev& d, uint32_t v, bool valid, const ipr_bslot_t* pr,
                                          uint32_t want, IP_MESSAGE_TYPE* out)
{
    const uint32_t lane = ipr_lane();
    uint64_t beg = 0;
    uint32_t deg = 0;
    if (valid) {
        beg = d.in_idx[v];
        deg = (uint32_t)(d.in_idx[v + 1] - beg);
    }
    const uint32_t incl = ipr_wave_incl_scan(deg);
    const uint32_t excl = incl - deg;
    const uint32_t total = __shfl(incl, IPR_WAVE - 1);
    IP_MESSAGE_TYPE acc = ipr_policy::ident();
    bool has = false;
    constexpr int U = IPR_GATHER_UNROLL;       /* chunks of 64 arcs in flight */
    for (uint32_t t00 = 0; t00 < total; t00 += U * IPR_WAVE) {
        uint64_t pos[U];
        uint32_t owner[U], u[U];
        ipr_bslot_t r[U];
        IP_MESSAGE_TYPE m[U];
        int h[U];
#pragma unroll
        for (int c = 0; c < U; c++) {
            const uint32_t t = t00 + c * IPR_WAVE + lane;
            uint32_t lo = 0;
#pragma unroll
            for (uint32_t step = 32; step > 0; step >>= 1) {
                uint32_t probe = __shfl(excl, (lo + step) & (IPR_WAVE - 1));
                if (lo + step < IPR_WAVE && probe <= t) lo += step;
            }
            const uint64_t o_beg = __shfl(beg, lo);
            const uint32_t o_ex = __shfl(excl, lo);
            /* lanes past the end read arc 0 (total > 0, so it exists) */
            owner[c] = t < total ? lo : IPR_WAVE + lane;   /* matches no other lane */
            pos[c] = t < total ? o_beg + (t - o_ex) : 0;
        }
#pragma unroll
        for (int c = 0; c < U; c++) u[c] = d.in_adj[pos[c]];
#pragma unroll
        for (int c = 0; c < U; c++) r[c] = pr[u[c]];
#pragma unroll
        for (int c = 0; c < U; c++) {
            h[c] = owner[c] < IPR_WAVE && r[c].step == want;
            m[c] = h[c] ? r[c].msg : ipr_policy::ident();
        }
#pragma unroll
        for (int c = 0; c < U; c++) {
            const uint32_t t0 = t00 + c * IPR_WAVE;
            if (t0 >= total) break;
            for (int o = 1; o < IPR_WAVE; o <<= 1) {
                IP_MESSAGE_TYPE y = __shfl_up(m[c], o);
                int yh = __shfl_up(h[c], o);
                uint32_t yo = __shfl_up(owner[c], o);
                if (lane >= (uint32_t)o && yo == owner[c]) {
                    ipr_policy::fold(&m[c], y);
                    h[c] |= yh;
                }
            }
            const uint32_t ss = excl > t0 ? excl : t0;
            const uint32_t se = excl + deg < t0 + IPR_WAVE ? excl + deg : t0 + IPR_WAVE;
            const uint32_t src = ss < se ? se - 1 - t0 : lane;
            IP_MESSAGE_TYPE y = __shfl(m[c], src);
            int yh = __shfl(h[c], src);
            if (ss < se) {
                ipr_policy::fold(&acc, y);
                has |= yh != 0;
            }
        }
    }
    *out = acc;
    return has;
}

typedef uint32_t ipr_u32x4 __attribute__((ext_vector_type(4)));

/* a receiver's folded messages of superstep s: the list may be split between
 * wavefronts, so the delivery is the combiner's atomic one */
__device__ inline void ipr_gather_flush(const ipr_dev& d, uint64_t s, uint32_t v, IP_MESSAGE_TYPE acc)
{
    ipr_policy::deliver((ipr_slot_t*)d.mbox[(s + 1) & 1] + v, acc);
    d.recv[(s + 1) & 1][v] = 1;
}

/* The bitmap of this superstep's broadcasters over record positions: one
 * wavefront per 64 positions, a ballot of their records' stamps, one word
 * written; no atomics, nothing to clear. Runs only when the gather will read
 * it. */
__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_gather_live(ipr_dev d)
{
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    if (!ipr_gathers_live(d)) return;
    /* the gather reads the bitmap only for the superstep stamped here (the
     * whole kernel ends before the gather starts) */
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->live_step = ctl->superstep + 1;
        ctl->live_built++;
    }
    const uint32_t want = (uint32_t)(ctl->superstep + 1);
    const uint32_t want4 = (uint32_t)(ctl->superstep % IPR_REC4_WINDOW) + 1u;
    const ipr_bslot_t* bs = (const ipr_bslot_t*)d.bslot;
    const uint64_t words = (d.V + 63) / 64;
    const uint64_t nw = (uint64_t)gridDim.x * IPR_WAVES;
    /* a wavefront covers many words: the stamps of U words are loaded
     * before the first ballot, U loads in flight per lane instead of one */
    constexpr int U = 8;
    for (uint64_t w0 = (uint64_t)blockIdx.x * IPR_WAVES + threadIdx.x / IPR_WAVE; w0 < words; w0 += nw * U) {
        bool on[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t p = (w0 + u * nw) * 64 + ipr_lane();
            on[u] = p < d.V && (d.brec4 ? (d.brec4[p] >> IPR_REC4_SHIFT) == want4
                                        : (bs[p].step & ~IPR_BSTEP_WEIGHTED) == want);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t wd = w0 + u * nw;
            const uint64_t m = __ballot(on[u]);
            if (ipr_lane() == 0 && wd < words) d.g_live[wd] = m;
        }
    }
}

/* Dense supersteps of push applications: every receiver folds the records of
 * the in-neighbours that broadcast in superstep s, over the exact transpose.
 * Arc-balanced, so hubs need no special case: the arcs split into ranges of
 * g_per_wave; one resident round of workgroups is launched, and each wavefront
 * takes ranges from a counter until none is left. A range is taken 64*K arcs
 * at a time; lane L folds
 * the K consecutive arcs base + L*K (sources by vector loads, then their K
 * records together), delivers the in-lists that end inside its K arcs, and one
 * segmented scan across the lanes joins the lists that span lanes. The list
 * still open after a pass carries into the next; the one open at the end of
 * the range is delivered and finished by whichever wavefront takes the next
 * range. Gathering is
 * only enabled for integral MIN/MAX/SUM combiners, whose result does not
 * depend on the order or grouping of the folds. */
/* at least 5 wavefronts per SIMD (at most 96 VGPRs): the occupancy its 32 KB
 * LDS copy allows, which the launch's 5 workgroups per CU assume. With arc
 * weights or messages wider than 4 bytes, 96 VGPRs would spill, and 4
 * wavefronts per SIMD are kept (apps whose messages are not integral never
 * launch it). */
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
constexpr int ipr_gather_min_waves = 4;
#else
constexpr int ipr_gather_min_waves = sizeof(IP_MESSAGE_TYPE) <= 4 ? 5 : 4;
#endif
__global__ void __launch_bounds__(IPR_GATHER_WAVES * IPR_WAVE, ipr_gather_min_waves) ipr_k_gather(ipr_dev d)
{
    constexpr int K = IPR_GATHER_K;
    static_assert(K % 4 == 0, "sources are read four at a time");
    /* with records in out-degree order the first ones are the most read:
     * each workgroup keeps a copy of IPR_GATHER_HOT_BYTES of them in LDS
     * (8192 records of 4 bytes or 4096 of 8; 32 KB still lets 5 workgroups
     * share a CU) */
    constexpr uint32_t HOT = IPR_GATHER_HOT_BYTES / sizeof(ipr_bslot_t);
    __shared__ ipr_bslot_t hot[HOT];
    uint32_t* hot4 = (uint32_t*)hot;
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    if (!ipr_gathers(d)) return;   /* never launched alone: plans need the tiled gather */
    const bool live = ipr_gather_bitmap(d);
    const uint64_t s = ctl->superstep;
    const uint32_t want = (uint32_t)(s + 1);
    const uint32_t want4 = (uint32_t)(s % IPR_REC4_WINDOW) + 1u;
    const uint32_t lane = ipr_lane();
    const ipr_bslot_t* bs = (const ipr_bslot_t*)d.bslot;
    const uint32_t* b4 = d.brec4;
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
    /* the arcs' weights in g_arc's order */
    const uint32_t* tw = d.g_wgt ? d.g_wgt : d.t_wgt;
#endif
    /* 4-byte records: twice as many of them in the same LDS */
    const uint32_t cap = b4 ? 2 * HOT : HOT;
    const uint32_t nhot = d.bpos ? (uint32_t)(d.V < cap ? d.V : cap) : 0u;
    constexpr uint32_t NT = IPR_GATHER_WAVES * IPR_WAVE;
    if (b4) {
        for (uint32_t h = threadIdx.x; h < nhot; h += NT) hot4[h] = b4[h];
    } else {
        for (uint32_t h = threadIdx.x; h < nhot; h += NT) hot[h] = bs[h];
    }
    __syncthreads();
    const uint64_t per = d.g_per_wave;
    const uint64_t nranges = (d.E + per - 1) / per;
    /* ranges gc, gc + G, gc + 2G ... come from range counter gc: G counters on
     * their own cache lines share the atomics (zeroed at every superstep's
     * end) */
    const uint32_t G = gridDim.x < IPR_GATHER_CTRS ? gridDim.x : IPR_GATHER_CTRS;
    const uint32_t gc = blockIdx.x % G;
    unsigned long long* ctr = (unsigned long long*)&d.cnt[IPR_CSLOT(IPR_C_GNEXT, gc)];
    /* arcs split by record line (g_xcd): workgroup b takes the ranges of class
     * b mod classes from that class's counter (workgroups reach the XCDs in
     * turn, so an XCD reads one share of the record lines), then helps with
     * the other classes once its own is done */
    uint32_t cls = d.g_xcd ? blockIdx.x % IPR_GATHER_CLASSES : 0u, ctried = 0;
    for (;;) {
        uint64_t w0 = 0, w1 = 0;
        const uint32_t* gd = d.g_dst;
        if (d.g_xcd) {
            bool got = false;
            while (ctried < IPR_GATHER_CLASSES) {
                unsigned long long k = 0;
                if (lane == 0)
                    k = atomicAdd((unsigned long long*)&d.cnt[IPR_CSLOT(IPR_C_GNEXT, cls)], 1ull);
                k = __shfl(k, 0);
                w0 = d.g_cbeg[cls] + (uint64_t)k * per;
                if (w0 < d.g_cend[cls]) {
                    w1 = w0 + per < d.g_cend[cls] ? w0 + per : d.g_cend[cls];
                    gd = d.g_dst + d.g_cdst[cls];
                    got = true;
                    break;
                }
                cls = (cls + 1) % IPR_GATHER_CLASSES;
                ctried++;
            }
            if (!got) break;
        } else {
            unsigned long long k = 0;
            if (lane == 0) k = atomicAdd(ctr, 1ull);
            k = __shfl(k, 0);
            const uint64_t rg = (uint64_t)k * G + gc;
            if (rg >= nranges) break;
            w0 = (uint64_t)rg * per;
            w1 = w0 + per < d.E ? w0 + per : d.E;
        }
        IP_MESSAGE_TYPE carry = ipr_policy::ident();
        int carry_has = 0;
        uint32_t carry_o = 0;
        for (uint64_t base = w0; base < w1; base += (uint64_t)IPR_WAVE * K) {
            const uint64_t t0 = base + (uint64_t)lane * K;
            uint32_t a[K];
            if (t0 + K <= w1) {
                /* the source stream is read once: keep it out of L2, which holds
                 * the randomly read broadcast records */
#pragma unroll
                for (int j = 0; j < K; j += 4) {
                    const ipr_u32x4 q = __builtin_nontemporal_load((const ipr_u32x4*)(d.g_arc + t0 + j));
                    a[j] = q[0]; a[j + 1] = q[1]; a[j + 2] = q[2]; a[j + 3] = q[3];
                }
            } else {
#pragma unroll
                for (int j = 0; j < K; j++) a[j] = t0 + j < w1 ? d.g_arc[t0 + j] : 0u;
            }
            const uint32_t o = t0 < w1 ? d.g_own[t0 / K] : 0u;   /* receiver of arc t0 */
            /* the K records: ok[j] bit 0 when source j broadcast in superstep s,
             * bit 1 when weighted; mv[j] its message */
            IP_MESSAGE_TYPE mv[K];
            uint32_t ok[K];
            uint64_t w[K];
            if (live) {
                /* the bitmap (V bits, cache-resident) first; records only of
                 * sources that broadcast */
#pragma unroll
                for (int j = 0; j < K; j++) w[j] = d.g_live[(a[j] & ~IPR_GATHER_LAST) >> 6];
            }
            if (b4) {
                uint32_t q[K];
#pragma unroll
                for (int j = 0; j < K; j++) {
                    const uint32_t p = a[j] & ~IPR_GATHER_LAST;
                    q[j] = 0;
                    if (!live || ((w[j] >> (a[j] & 63)) & 1ull)) q[j] = p < nhot ? hot4[p] : b4[p];
                }
#pragma unroll
                for (int j = 0; j < K; j++) {
                    const uint32_t low = q[j] & IPR_REC4_ESC;
                    ok[j] = (q[j] >> IPR_REC4_SHIFT) == want4 ? 1u | ((q[j] & IPR_REC4_WEIGHTED) ? 2u : 0u) : 0u;
                    mv[j] = (IP_MESSAGE_TYPE)low;
                    if (ok[j] && low == IPR_REC4_ESC) mv[j] = bs[a[j] & ~IPR_GATHER_LAST].msg;
                }
            } else {
                ipr_bslot_t r[K];
#pragma unroll
                for (int j = 0; j < K; j++) {
                    const uint32_t p = a[j] & ~IPR_GATHER_LAST;
                    r[j].step = 0;
                    if (!live || ((w[j] >> (a[j] & 63)) & 1ull)) r[j] = p < nhot ? hot[p] : bs[p];
                }
#pragma unroll
                for (int j = 0; j < K; j++) {
                    ok[j] = (r[j].step & ~IPR_BSTEP_WEIGHTED) == want
                                ? 1u | ((r[j].step & IPR_BSTEP_WEIGHTED) ? 2u : 0u) : 0u;
                    mv[j] = r[j].msg;
                }
            }
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
            uint32_t wt[K];
            if (t0 + K <= w1) {
#pragma unroll
                for (int j = 0; j < K; j += 4) {
                    const ipr_u32x4 q = __builtin_nontemporal_load((const ipr_u32x4*)(tw + t0 + j));
                    wt[j] = q[0]; wt[j + 1] = q[1]; wt[j + 2] = q[2]; wt[j + 3] = q[3];
                }
            } else {
#pragma unroll
                for (int j = 0; j < K; j++) wt[j] = t0 + j < w1 ? tw[t0 + j] : 0u;
            }
#endif
            /* head: the part of the list open on entry (it began in an earlier
             * lane); fv/fh: lists that begin and end in these K arcs */
            IP_MESSAGE_TYPE head = ipr_policy::ident(), cur = ipr_policy::ident();
            int head_has = 0, cur_has = 0;
            uint32_t nb = 0;
            IP_MESSAGE_TYPE fv[K];
            int fh[K];
#pragma unroll
            for (int j = 0; j < K; j++) {
                const bool live = t0 + j < w1;
                if (live && ok[j]) {
                    IP_MESSAGE_TYPE m = mv[j];
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
                    if (ok[j] & 2u) m = (IP_MESSAGE_TYPE)(m + (IP_MESSAGE_TYPE)wt[j]);
#endif
                    ipr_policy::fold(&cur, m);
                    cur_has = 1;
                }
                fv[j] = cur;
                fh[j] = 0;
                if (live && (a[j] & IPR_GATHER_LAST)) {
                    if (nb == 0) { head = cur; head_has = cur_has; }
                    else fh[j] = cur_has;
                    nb++;
                    cur = ipr_policy::ident();
                    cur_has = 0;
                }
            }
            uint32_t k = 0;
#pragma unroll
            for (int j = 0; j < K; j++) {
                if (t0 + j < w1 && (a[j] & IPR_GATHER_LAST)) {
                    if (fh[j]) ipr_gather_flush(d, s, gd[o + k], fv[j]);
                    k++;
                }
            }
            /* lanes without a list end extend the list open before them */
            const int f = nb > 0;
            if (lane == 0 && !f) { ipr_policy::fold(&cur, carry); cur_has |= carry_has; }
            IP_MESSAGE_TYPE S = cur;
            int Sh = cur_has, F = f;
            for (int dd = 1; dd < IPR_WAVE; dd <<= 1) {
                const IP_MESSAGE_TYPE y = __shfl_up(S, dd);
                const int yh = __shfl_up(Sh, dd), yf = __shfl_up(F, dd);
                if (lane >= (uint32_t)dd && !F) { ipr_policy::fold(&S, y); Sh |= yh; F |= yf; }
            }
            IP_MESSAGE_TYPE in = __shfl_up(S, 1);
            int inh = __shfl_up(Sh, 1);
            if (lane == 0) { in = carry; inh = carry_has; }
            if (f && (head_has | inh)) {
                ipr_policy::fold(&head, in);
                ipr_gather_flush(d, s, gd[o], head);
            }
            /* the last lane with arcs holds the list still open */
            const uint64_t left = w1 - base;
            const int last = left >= (uint64_t)IPR_WAVE * K ? IPR_WAVE - 1 : (int)((left - 1) / K);
            carry = __shfl(S, last);
            carry_has = __shfl(Sh, last);
            carry_o = __shfl(o + nb, last);
        }
        if (lane == 0 && carry_has) ipr_gather_flush(d, s, gd[carry_o], carry);
    }
}

/* Receiver-blocked gather: the same dense-superstep delivery as ipr_k_gather,
 * over tiles of a block of consecutive receivers whose in-arcs are sorted by
 * the source's record position. The lanes of a wavefront read consecutive
 * arcs, so their record loads walk the record array in ascending order and
 * share cache lines wherever a block's sources are dense (the high-degree
 * records, first in out-degree order): about 0.07-0.09 record-line requests
 * per arc at R-MAT scale 22, against 0.67 for the receiver-ordered gather.
 * The block's receivers fold in LDS with the combiner's atomic (integral
 * MIN/MAX/SUM: the result does not depend on the order); a receiver with a
 * message equal to the empty value also raises its bit in `th`. The tile's
 * folds are then delivered in receiver order: a plain read-modify-write (the
 * only writer of those mailboxes in this kernel; the compute kernel's
 * ip_send_message deliveries came before), or the combiner's atomic for the
 * tiles of one receiver whose in-arcs are split between tiles. Resident
 * workgroups take the tiles, largest first, from one counter. */
/* One workgroup of 16 wavefronts per CU, 128 KB of receiver slots (32768
 * receivers of 4-byte messages), 16 slots per lane per pass (the most that
 * stays within 128 VGPRs with weights). Measured on the dense gathers of CC
 * on R-MAT 24 per launch: 512 threads, 64 KB, 8 slots, two workgroups per
 * CU: 1.17 ms; 1024 threads, 128 KB, 8 slots: 0.98 ms; 16 slots: 0.93 ms
 * (4 slots 1.13 ms; 96 or 144 KB within 2%). */
#ifndef IPR_TILE_THREADS
#define IPR_TILE_THREADS 1024
#endif
#ifndef IPR_TILE_LDS_BYTES
#define IPR_TILE_LDS_BYTES 131072
#endif
/* slot groups per wavefront pass: as many as fit in the 128 VGPRs that four
 * wavefronts per SIMD leave (the LDS allows no more): 20 for 4-byte messages
 * without weights (125 VGPRs; 1-2% faster gathers than 16 on CC), 16 with
 * weights or 8-byte messages, where 20 spills */
#ifndef IPR_TILE_K
#if defined(IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS)
#define IPR_TILE_K 16
#else
#define IPR_TILE_K (sizeof(IP_MESSAGE_TYPE) > 4 ? 16 : 20)
#endif
#endif
constexpr bool ipr_tile_ok = ipr_policy::exact_gather && IPR_COMBINER != IPR_COMBINER_GENERIC &&
                             std::is_integral<IP_MESSAGE_TYPE>::value &&
                             (sizeof(IP_MESSAGE_TYPE) == 4 || sizeof(IP_MESSAGE_TYPE) == 8);
constexpr uint32_t ipr_tile_r = IPR_TILE_LDS_BYTES / sizeof(ipr_slot_t) < 65536u
                                    ? (uint32_t)(IPR_TILE_LDS_BYTES / sizeof(ipr_slot_t)) : 65536u;
static_assert(ipr_tile_r % 32 == 0, "tile receivers: whole words of the message bits");
static_assert(IPR_TILE_K >= 1 && IPR_TILE_K <= 64, "one lane loads each slot group's base");
/* resident workgroups per CU that its LDS allows (160 KB per CU on MI355X) */
constexpr uint32_t ipr_tile_wgs_per_cu =
    163840u / (uint32_t)(ipr_tile_r * sizeof(ipr_slot_t) + ipr_tile_r / 8 + 4) > 0
        ? 163840u / (uint32_t)(ipr_tile_r * sizeof(ipr_slot_t) + ipr_tile_r / 8 + 4) : 1u;

template <typename S>
__device__ __forceinline__ void ipr_tile_fold(S* p, IP_MESSAGE_TYPE m)
{
    if constexpr (ipr_tile_ok) {
        if constexpr (IPR_COMBINER == IPR_COMBINER_MIN) atomicMin(p, m);
        else if constexpr (IPR_COMBINER == IPR_COMBINER_MAX) atomicMax(p, m);
        else atomicAdd(p, m);
    }
}

/* The slot groups of a pass at base (whole groups: tiles start and end on
 * them; none past a1): lane j returns group j's base, and qv[j] gets each
 * lane's slot of group j (IPR_TQ_NONE past the end). The index is read once
 * per launch, so it is kept out of L2, which holds the records; the bases are
 * read back per group with readlane only after all the loads are issued.
 * (Issuing the next pass's slots before this pass's records measured the
 * same: the launch is bound by the record lines, not by the index latency.) */
__device__ __forceinline__ uint32_t ipr_tile_load(const ipr_dev& d, uint64_t base, uint64_t a1, uint32_t lane,
                                                  uint32_t (&qv)[IPR_TILE_K])
{
    const uint32_t ng = base < a1 ? (uint32_t)((a1 - base) >> 6 < IPR_TILE_K ? (a1 - base) >> 6 : IPR_TILE_K) : 0u;
#pragma unroll
    for (int j = 0; j < IPR_TILE_K; j++)
        qv[j] = (uint32_t)j < ng ? __builtin_nontemporal_load(d.g_tq + base + (uint64_t)j * IPR_WAVE + lane)
                                  : 0xFFFFFFFFu;
    return lane < ng ? d.g_thead[(base >> 6) + lane] : 0u;
}

__global__ void __launch_bounds__(IPR_TILE_THREADS) ipr_k_gather_tiled(ipr_dev d)
{
    if constexpr (ipr_tile_ok) {
        constexpr uint32_t R = ipr_tile_r, NT = IPR_TILE_THREADS, K = IPR_TILE_K;
        __shared__ ipr_slot_t ts[R];
        __shared__ uint32_t th[R / 32];
        __shared__ uint32_t tcur;
        ipr_ctl* ctl = d.ctl;
        if (ctl->done) return;
        if (!ipr_gathers(d)) {
            ipr_plan_miss(d);
            return;
        }
        /* few live arcs: the broadcasters' bitmap (built by
         * ipr_k_gather_live) is read first, and records only of sources
         * whose bit is set */
        const bool live = ipr_gather_bitmap(d);
        const uint64_t s = ctl->superstep;
        const uint32_t want = (uint32_t)(s + 1);
        const uint32_t want4 = (uint32_t)(s % IPR_REC4_WINDOW) + 1u;
        const ipr_bslot_t* bs = (const ipr_bslot_t*)d.bslot;
        const uint32_t* b4 = d.brec4;
        ipr_slot_t* mb = (ipr_slot_t*)d.mbox[(s + 1) & 1];
        uint8_t* rv = d.recv[(s + 1) & 1];
        unsigned long long* ctr = (unsigned long long*)&d.cnt[IPR_CSLOT(IPR_C_GNEXT, 0)];
        /* the wavefront index and the tile number are uniform: read them
         * through readfirstlane so that the tile's fields, the pass bounds and
         * the group bases live in scalar registers (scalar loads, scalar
         * branches on escape rows) */
        const uint32_t lane = ipr_lane(), wv = __builtin_amdgcn_readfirstlane(threadIdx.x / IPR_WAVE);
        const IP_MESSAGE_TYPE id = ipr_policy::ident();
        for (;;) {
            if (threadIdx.x == 0) tcur = (uint32_t)atomicAdd(ctr, 1ull);
            __syncthreads();
            const uint32_t t = __builtin_amdgcn_readfirstlane(tcur);
            if (t >= d.g_tn) break;
            const ipr_tblock tb = d.g_tblk[t];
            for (uint32_t r = threadIdx.x; r < tb.nr; r += NT) ts[r] = id;
            for (uint32_t w = threadIdx.x; w < (tb.nr + 31) / 32; w += NT) th[w] = 0;
            __syncthreads();
            const uint64_t a1 = tb.a0 + tb.na;
            for (uint64_t base = tb.a0 + (uint64_t)wv * IPR_WAVE * K; base < a1; base += (uint64_t)NT * K) {
                uint32_t p[K], r[K];
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
                uint32_t wt[K];
#endif
                uint32_t qv[K];
                const uint32_t hl = ipr_tile_load(d, base, a1, lane, qv);
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
                const uint32_t ng = (uint32_t)((a1 - base) >> 6 < K ? (a1 - base) >> 6 : K);
#pragma unroll
                for (int j = 0; j < (int)K; j++) {
                    const uint64_t i = base + (uint64_t)j * IPR_WAVE + lane;
                    wt[j] = 0;
                    if ((uint32_t)j < ng) {
                        if (d.g_twsz == 1) wt[j] = __builtin_nontemporal_load((const uint8_t*)d.g_twgt + i);
                        else if (d.g_twsz == 2) wt[j] = __builtin_nontemporal_load((const uint16_t*)d.g_twgt + i);
                        else wt[j] = __builtin_nontemporal_load((const uint32_t*)d.g_twgt + i);
                    }
                }
#endif
#pragma unroll
                for (int j = 0; j < (int)K; j++) {
                    const uint32_t h = __builtin_amdgcn_readlane(hl, j);
                    r[j] = qv[j] & 0xFFFFu;
                    if (h & IPR_TQ_ESC) p[j] = d.g_tesc[(uint64_t)(h & ~IPR_TQ_ESC) * IPR_WAVE + lane];
                    else p[j] = h + (qv[j] >> 16);
                    if (r[j] == IPR_TQ_NONE) p[j] = 0xFFFFFFFFu;
                }
                if (live) {
                    uint64_t lw[K];
#pragma unroll
                    for (int j = 0; j < (int)K; j++) lw[j] = p[j] != 0xFFFFFFFFu ? d.g_live[p[j] >> 6] : 0ull;
#pragma unroll
                    for (int j = 0; j < (int)K; j++)
                        if (!((lw[j] >> (p[j] & 63)) & 1ull)) p[j] = 0xFFFFFFFFu;
                }
                /* ok[j]: bit 0 when source j broadcast in superstep s, bit 1
                 * when weighted; mv[j] its message */
                uint32_t ok[K];
                IP_MESSAGE_TYPE mv[K];
                if (b4) {
                    uint32_t q[K];
#pragma unroll
                    for (int j = 0; j < (int)K; j++) q[j] = p[j] != 0xFFFFFFFFu ? b4[p[j]] : 0u;
#pragma unroll
                    for (int j = 0; j < (int)K; j++) {
                        const uint32_t low = q[j] & IPR_REC4_ESC;
                        ok[j] = (q[j] >> IPR_REC4_SHIFT) == want4 ? 1u | ((q[j] & IPR_REC4_WEIGHTED) ? 2u : 0u) : 0u;
                        mv[j] = (IP_MESSAGE_TYPE)low;
                        if (ok[j] && low == IPR_REC4_ESC) mv[j] = bs[p[j]].msg;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < (int)K; j++) {
                        ipr_bslot_t b;
                        b.step = 0;
                        b.msg = id;
                        if (p[j] != 0xFFFFFFFFu) b = bs[p[j]];
                        ok[j] = (b.step & ~IPR_BSTEP_WEIGHTED) == want
                                    ? 1u | ((b.step & IPR_BSTEP_WEIGHTED) ? 2u : 0u) : 0u;
                        mv[j] = b.msg;
                    }
                }
#pragma unroll
                for (int j = 0; j < (int)K; j++) {
                    if (ok[j]) {
                        IP_MESSAGE_TYPE m = mv[j];
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
                        if (ok[j] & 2u) m = (IP_MESSAGE_TYPE)(m + (IP_MESSAGE_TYPE)wt[j]);
#endif
                        ipr_tile_fold(&ts[r[j]], m);
                        if (m == id) atomicOr(&th[r[j] >> 5], 1u << (r[j] & 31));
                    }
                }
            }
            __syncthreads();
            for (uint32_t r = threadIdx.x; r < tb.nr; r += NT) {
                const ipr_slot_t x = ts[r];
                if (x != id || ((th[r >> 5] >> (r & 31)) & 1u)) {
                    const uint32_t v = tb.v0 + r;
                    if (tb.shared) ipr_policy::deliver(mb + v, x);
                    else ipr_policy::merge(mb + v, x);
                    rv[v] = 1;
                }
            }
            __syncthreads();
        }
    }
}

/* PUSH: src/combiner_postamble.h:166-183 - a delivered mailbox becomes the
 * vertex's message and reactivates it. */
__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_swap(ipr_dev d)
{
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    const uint64_t s = ctl->superstep;
    ipr_sm_superstep = s;
    ipr_sm_counted = 0;
    ipr_slot_t* mb = (ipr_slot_t*)d.mbox[(s + 1) & 1];
    uint8_t* rv = d.recv[(s + 1) & 1];
    uint64_t n_react = 0;
    /* the last superstep of a stamp window clears the 4-byte records */
    uint32_t* clr = d.brec4 && (s + 1) % IPR_REC4_WINDOW == 0 ? d.brec4 : nullptr;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < d.V; v += stride) {
        if (clr) clr[v] = 0;
        if (rv[v]) {
            rv[v] = 0;   /* clean for superstep s+2 */
            ((IP_MESSAGE_TYPE*)d.msg)[v] = ipr_policy::value(mb[v]);
            mb[v] = ipr_policy::empty();
            d.hm[v] = 1;
            if (!d.act[v]) { d.act[v] = 1; n_react++; }
        }
    }
    ipr_block_count3(d, IPR_C_ACT0 + (int)(s & 1), n_react, -1, 0, -1, 0);
}

/* ------------------------------------------------------------------------ */
/* pull variant                                                              */
/* ------------------------------------------------------------------------ */

/* fold of the broadcasting in-neighbours of v in in-list order
 * (src/combiner_single_broadcast_postamble.h:57-89); pr holds superstep
 * s-1's broadcast records, stamped want = s */
__device__ inline bool ipr_fetch_ordered(const ipr_dev& d, uint32_t v, const ipr_bslot_t* pr, uint32_t want,
                                         IP_MESSAGE_TYPE* out)
{
    uint64_t i = d.in_idx[v], end = d.in_idx[v + 1];
    bool has = false;
    IP_MESSAGE_TYPE acc = IP_MESSAGE_TYPE();
    for (; i < end; i++) {
        const ipr_bslot_t r = pr[d.in_adj[i]];
        if (r.step == want) {
            if (has) ip_combine(&acc, r.msg);
            else { acc = r.msg; has = true; }
        }
    }
    *out = acc;
    return has;
}

/* one vertex: fetch(s-1) result -> activation -> compute(s) */
__device__ inline void ipr_pull_vertex(const ipr_dev& d, uint64_t s, uint32_t v, bool hm,
                                       IP_MESSAGE_TYPE msg, uint64_t* n_react, uint64_t* n_active,
                                       uint64_t* n_bcast)
{
    bool act = d.act[v] != 0;
    const bool act0 = act;
    if (hm && !act) { act = true; (*n_react)++; }
    if (act) {
        struct ip_vertex_t x;
        ipr_fill_vertex(d, &x, v);
        x.active = true;
        x.has_message = hm;
        x.message = msg;
        const IP_VALUE_TYPE before = x.value;
        ip_compute(&x);
        ipr_store_value(d, v, before, x.value);
        act = x.active;
        if (act) (*n_active)++;
        if (x.ipr_bcast) {
            /* stamped s+1: read in superstep s+1, and no clearing needed */
            ipr_bslot_t rec;
            rec.step = (uint32_t)(s + 1);
            rec.msg = x.ipr_bcast_msg;
            ((ipr_bslot_t*)d.bcm[s & 1])[v] = rec;
            (*n_bcast)++;
        }
    }
    if (act != act0) d.act[v] = act;
}

/* Counters of the fused pull kernel: reactivations finish ActiveVertexCount
 * of superstep s-1, the active vertices after compute start that of s. */
__device__ inline void ipr_pull_counts(const ipr_dev& d, uint64_t s, uint64_t n_react,
                                       uint64_t n_active, uint64_t n_bcast)
{
    ipr_block_count3(d, s > 0 ? IPR_C_ACT0 + (int)((s + 1) & 1) : -1, n_react,
                     IPR_C_ACT0 + (int)(s & 1), n_active, IPR_C_BCAST0 + (int)(s & 1), n_bcast);
}

/* In-list order fold by one wavefront (vertices of larger in-degree): 64
 * in-neighbours are fetched at once, then folded lane by lane in list order
 * with readlane, so the result is bit-identical to the sequential fold. */
__device__ inline bool ipr_fetch_ordered_wave(const ipr_dev& d, uint32_t v, const ipr_bslot_t* pr,
                                              uint32_t want, IP_MESSAGE_TYPE* out)
{
    const uint32_t lane = ipr_lane();
    const uint64_t beg = d.in_idx[v], end = d.in_idx[v + 1];
    bool has = false;
    IP_MESSAGE_TYPE acc = IP_MESSAGE_TYPE();
    for (uint64_t base = beg; base < end; base += IPR_WAVE) {
        const uint64_t k = base + lane;
        bool f = false;
        IP_MESSAGE_TYPE m = IP_MESSAGE_TYPE();
        if (k < end) {
            const ipr_bslot_t r = pr[d.in_adj[k]];
            f = r.step == want;
            if (f) m = r.msg;
        }
        uint64_t mask = __ballot(f);
        while (mask) {
            const uint32_t j = __ffsll((unsigned long long)mask) - 1;
            IP_MESSAGE_TYPE mj = __shfl(m, j);
            if (has) ip_combine(&acc, mj);
            else { acc = mj; has = true; }
            mask &= mask - 1;
        }
    }
    *out = acc;
    return has;
}

/* Fused fetch(s-1) + compute(s) over vertices binned by in-degree (static
 * bins built at load). Blocks [0, nb0) serve bin 0, [nb0, nb0+nb1) bin 1,
 * the rest bin 2.
 * ORDERED (bit-exact with the reference's in-list order): bin 0 one thread
 * per vertex, bins 1-2 one wavefront per vertex with the ordered fold.
 * Otherwise (declared combiner, reassociating): bin 0 the lane-cooperative
 * segmented fold, bin 1 a wavefront and bin 2 a workgroup per vertex with
 * tree reductions. */
/* reassociating fold of one in-list slice: this thread takes the arcs
 * k0, k0 + STRIDE, ... below end, four at a time: their source ids first,
 * then the four records, so that no load waits on another of its kind */
template <uint32_t STRIDE>
__device__ inline void ipr_pull_chunks(const ipr_dev& d, const ipr_bslot_t* pr, uint32_t want, uint64_t k0,
                                       uint64_t end, IP_MESSAGE_TYPE* acc, bool* has)
{
    constexpr int U = 4;
    for (uint64_t k = k0; k < end; k += (uint64_t)STRIDE * U) {
        uint32_t src[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t ku = k + (uint64_t)u * STRIDE;
            src[u] = ku < end ? d.in_adj[ku] : 0u;
        }
        ipr_bslot_t r[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            r[u].step = 0;
            if (k + (uint64_t)u * STRIDE < end) r[u] = pr[src[u]];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (r[u].step == want) {
                ipr_policy::fold(acc, r[u].msg);
                *has = true;
            }
        }
    }
}

template <bool ORDERED>
__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_pull_step(ipr_dev d, uint32_t nb0, uint32_t nb1)
{
    __shared__ IP_MESSAGE_TYPE part[IPR_WAVES];
    __shared__ int parth[IPR_WAVES];
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    const uint64_t s = ctl->superstep;
    ipr_sm_superstep = s;
    ipr_sm_counted = 0;
    /* superstep s-1's broadcast records, stamped s */
    const ipr_bslot_t* pr = (const ipr_bslot_t*)d.bcm[(s + 1) & 1];
    const uint32_t want = (uint32_t)s;
    uint64_t n_react = 0, n_active = 0, n_bcast = 0;
    const uint32_t lane = ipr_lane();
    const uint32_t wave_in_block = threadIdx.x / IPR_WAVE;
    const uint32_t nb2 = gridDim.x - nb0 - nb1;
    const uint32_t b = blockIdx.x;
    if (b < nb0) {
        const uint64_t n = d.pbin_size[0];
        if (ORDERED) {
            for (uint64_t i = (uint64_t)b * blockDim.x + threadIdx.x; i < n; i += (uint64_t)nb0 * blockDim.x) {
                uint32_t v = d.pbins[0][i];
                IP_MESSAGE_TYPE msg = IP_MESSAGE_TYPE();
                bool hm = false;
                if (s > 0) hm = ipr_fetch_ordered(d, v, pr, want, &msg);
                ipr_pull_vertex(d, s, v, hm, msg, &n_react, &n_active, &n_bcast);
            }
        } else {
            const uint64_t stride = (uint64_t)nb0 * blockDim.x;
            for (uint64_t base = (uint64_t)b * blockDim.x + (threadIdx.x & ~(IPR_WAVE - 1)); base < n;
                 base += stride) {
                const uint64_t i = base + lane;
                const uint32_t v = i < n ? d.pbins[0][i] : 0;
                IP_MESSAGE_TYPE acc = ipr_policy::ident();
                bool hm = false;
                if (s > 0) hm = ipr_wave_pull_fold(d, v, i < n, pr, want, &acc);
                if (i < n) ipr_pull_vertex(d, s, v, hm, acc, &n_react, &n_active, &n_bcast);
            }
        }
    } else if (ORDERED) {
        /* bins 1 and 2: one wavefront per vertex, ordered */
        const bool one = b < nb0 + nb1;
        const uint64_t n = d.pbin_size[one ? 1 : 2];
        const uint32_t* list = d.pbins[one ? 1 : 2];
        const uint64_t first = (uint64_t)(b - (one ? nb0 : nb0 + nb1)) * IPR_WAVES + wave_in_block;
        const uint64_t nwaves = (uint64_t)(one ? nb1 : nb2) * IPR_WAVES;
        for (uint64_t i = first; i < n; i += nwaves) {
            uint32_t v = list[i];
            IP_MESSAGE_TYPE msg = IP_MESSAGE_TYPE();
            bool hm = false;
            if (s > 0) hm = ipr_fetch_ordered_wave(d, v, pr, want, &msg);
            if (lane == 0) ipr_pull_vertex(d, s, v, hm, msg, &n_react, &n_active, &n_bcast);
        }
    } else if (b < nb0 + nb1) {
        const uint64_t n = d.pbin_size[1];
        const uint64_t wave = (uint64_t)(b - nb0) * IPR_WAVES + wave_in_block;
        const uint64_t nwaves = (uint64_t)nb1 * IPR_WAVES;
        for (uint64_t i = wave; i < n; i += nwaves) {
            uint32_t v = d.pbins[1][i];
            IP_MESSAGE_TYPE acc = ipr_policy::ident();
            bool has = false;
            if (s > 0) {
                uint64_t beg = d.in_idx[v], end = d.in_idx[v + 1];
                ipr_pull_chunks<IPR_WAVE>(d, pr, want, beg + lane, end, &acc, &has);
                for (int o = 32; o > 0; o >>= 1) {
                    IP_MESSAGE_TYPE y = __shfl_xor(acc, o);
                    ipr_policy::fold(&acc, y);
                }
                has = __ballot(has) != 0;
            }
            if (lane == 0) ipr_pull_vertex(d, s, v, has, acc, &n_react, &n_active, &n_bcast);
        }
    } else {
        const uint64_t n = d.pbin_size[2];
        for (uint64_t i = b - nb0 - nb1; i < n; i += nb2) {
            uint32_t v = d.pbins[2][i];
            IP_MESSAGE_TYPE acc = ipr_policy::ident();
            bool has = false;
            if (s > 0) {
                uint64_t beg = d.in_idx[v], end = d.in_idx[v + 1];
                ipr_pull_chunks<IPR_BLOCK>(d, pr, want, beg + threadIdx.x, end, &acc, &has);
                for (int o = 32; o > 0; o >>= 1) {
                    IP_MESSAGE_TYPE y = __shfl_xor(acc, o);
                    ipr_policy::fold(&acc, y);
                }
                bool wh = __ballot(has) != 0;
                if (lane == 0) { part[wave_in_block] = acc; parth[wave_in_block] = wh; }
                __syncthreads();
                if (threadIdx.x == 0) {
                    acc = part[0];
                    has = parth[0];
                    for (int w = 1; w < IPR_WAVES; w++) {
                        ipr_policy::fold(&acc, part[w]);
                        has |= parth[w] != 0;
                    }
                }
                __syncthreads();
            }
            if (threadIdx.x == 0) ipr_pull_vertex(d, s, v, has, acc, &n_react, &n_active, &n_bcast);
        }
    }
    ipr_pull_counts(d, s, n_react, n_active, n_bcast);
    /* the last workgroup ends superstep s-1: no finalize launch */
    if (d.fin_in_kernel && ipr_last_workgroup(d, gridDim.x) && threadIdx.x < IPR_WAVE)
        ipr_finalize_step(d, IPR_VARIANT_PULL);
}

/* ------------------------------------------------------------------------ */
/* spread + single broadcast: fetch for the targets                          */
/* ------------------------------------------------------------------------ */

/* src/combiner_spread_single_broadcast_postamble.h ip_fetch_broadcast_messages
 * for every target of superstep s+1: the broadcasting in-neighbours (records
 * stamped with superstep s) folded in in-list order, no activation logic.
 * One thread per target; targets with more than 64 in-neighbours are taken by
 * their whole wavefront in turn (ordered wavefront fold). */
__device__ inline bool ipr_fetch_ordered_epoch_wave(const ipr_dev& d, uint32_t v, uint32_t want,
                                                    IP_MESSAGE_TYPE* out)
{
    const ipr_bslot_t* bs = (const ipr_bslot_t*)d.bslot;
    const uint32_t lane = ipr_lane();
    const uint64_t beg = d.in_idx[v], end = d.in_idx[v + 1];
    bool has = false;
    IP_MESSAGE_TYPE acc = IP_MESSAGE_TYPE();
    for (uint64_t base = beg; base < end; base += IPR_WAVE) {
        const uint64_t k = base + lane;
        bool f = false;
        IP_MESSAGE_TYPE m = IP_MESSAGE_TYPE();
        if (k < end) {
            ipr_bslot_t b = bs[ipr_brec(d, d.in_adj[k])];
            f = (b.step & ~IPR_BSTEP_WEIGHTED) == want;
            m = b.msg;
        }
        uint64_t mask = __ballot(f);
        while (mask) {
            const uint32_t j = __ffsll((unsigned long long)mask) - 1;
            IP_MESSAGE_TYPE mj = __shfl(m, j);
            if (has) ip_combine(&acc, mj);
            else { acc = mj; has = true; }
            mask &= mask - 1;
        }
    }
    *out = acc;
    return has;
}

__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_fetch_targets(ipr_dev d)
{
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    const uint64_t s = ctl->superstep;
    ipr_sm_superstep = s;
    ipr_sm_counted = 0;
    const uint32_t want = (uint32_t)(s + 1);
    const uint64_t n = ctl->fsize[(s + 1) & 1];
    const uint32_t* list = d.front[(s + 1) & 1];
    const ipr_bslot_t* bs = (const ipr_bslot_t*)d.bslot;
    const uint32_t lane = ipr_lane();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t nr = (n + IPR_WAVE - 1) & ~(uint64_t)(IPR_WAVE - 1);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += stride) {
        uint32_t v = 0;
        bool big = false;
        if (i < n) {
            v = list[i];
            uint64_t beg = d.in_idx[v], end = d.in_idx[v + 1];
            if (end - beg > IPR_WAVE) {
                big = true;
            } else {
                bool has = false;
                IP_MESSAGE_TYPE acc = IP_MESSAGE_TYPE();
                for (uint64_t k = beg; k < end; k++) {
                    ipr_bslot_t b = bs[ipr_brec(d, d.in_adj[k])];
                    if ((b.step & ~IPR_BSTEP_WEIGHTED) == want) {
                        if (has) ip_combine(&acc, b.msg);
                        else { acc = b.msg; has = true; }
                    }
                }
                d.hm[v] = has;
                if (has) ((IP_MESSAGE_TYPE*)d.msg)[v] = acc;
            }
        }
        uint64_t bigmask = __ballot(big);
        while (bigmask) {
            const uint32_t j = __ffsll((unsigned long long)bigmask) - 1;
            const uint32_t vj = __shfl(v, j);
            IP_MESSAGE_TYPE acc;
            bool has = ipr_fetch_ordered_epoch_wave(d, vj, want, &acc);
            if (lane == 0) {
                d.hm[vj] = has;
                if (has) ((IP_MESSAGE_TYPE*)d.msg)[vj] = acc;
            }
            bigmask &= bigmask - 1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* launch table                                                              */
/* ------------------------------------------------------------------------ */

/* workgroups per in-degree bin: threads, wavefronts, workgroups per vertex */
static void ipr_bin_blocks(const ipr_dev& d, uint32_t nb[3])
{
    uint64_t b0 = (d.pbin_size[0] + IPR_BLOCK - 1) / IPR_BLOCK;
    uint64_t b1 = (d.pbin_size[1] + IPR_WAVES - 1) / IPR_WAVES;
    uint64_t b2 = d.pbin_size[2];
    /* the wavefront bin loops when it has more vertices than pull_wave_blocks
     * workgroups hold (3 per CU by default): PageRank scale 16 launches 1850
     * workgroups without the cap, 1016 with it, and a run takes 0.235 ms
     * instead of 0.256 (a cap of 400: 0.26 ms, 1024: 0.245 ms) */
    const uint64_t cap1 = d.pull_wave_blocks ? d.pull_wave_blocks : 8192;
    nb[0] = (uint32_t)(b0 > 8192 ? 8192 : b0);
    nb[1] = (uint32_t)(b1 > cap1 ? cap1 : b1);
    nb[2] = (uint32_t)(b2 > 8192 ? 8192 : b2);
    if (nb[0] + nb[1] + nb[2] == 0) nb[0] = 1;
}

/* the compute kernel's resident workgroups on this device (0 = unknown): its
 * grid is capped there, since workgroups past one resident round would run
 * in a second round after the first has finished its loop */
static uint32_t ipr_compute_resident = 0;

static int ipr_app_launch(int kernel, const ipr_dev* dp, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    const ipr_dev& d = *dp;
    uint64_t vb = (d.V + IPR_BLOCK - 1) / IPR_BLOCK;
    uint32_t vblocks = (uint32_t)(vb == 0 ? 1 : (vb > 65536 ? 65536 : vb));
    uint32_t gblocks = d.grid_blocks ? d.grid_blocks : 2048;
    switch (kernel) {
    case IPR_K_INIT:
        hipLaunchKernelGGL(ipr_k_init, dim3(vblocks), dim3(IPR_BLOCK), 0, st, d);
        break;
    case IPR_K_COMPUTE: {
        uint32_t cb = vblocks < gblocks ? vblocks : gblocks;
        if (ipr_compute_resident && cb > ipr_compute_resident) cb = ipr_compute_resident;
        hipLaunchKernelGGL(ipr_k_compute_push, dim3(cb), dim3(IPR_BLOCK), 0, st, d);
        break;
    }
    case IPR_K_EXPAND:
        hipLaunchKernelGGL(ipr_k_expand, dim3(gblocks), dim3(IPR_BLOCK), 0, st, d);
        break;
    case IPR_K_FETCH:
        hipLaunchKernelGGL(ipr_k_fetch_targets, dim3(vblocks < gblocks ? vblocks : gblocks), dim3(IPR_BLOCK),
                           0, st, d);
        break;
    case IPR_K_SWAP:
        hipLaunchKernelGGL(ipr_k_swap, dim3(vblocks < gblocks ? vblocks : gblocks), dim3(IPR_BLOCK), 0, st, d);
        break;
    case IPR_K_GATHER:
        if (d.g_tn)
            hipLaunchKernelGGL(ipr_k_gather_tiled, dim3(d.tile_blocks ? d.tile_blocks : 1), dim3(IPR_TILE_THREADS), 0,
                               st, d);
        else
            hipLaunchKernelGGL(ipr_k_gather, dim3(d.gather_blocks ? d.gather_blocks : 1),
                               dim3(IPR_GATHER_WAVES * IPR_WAVE), 0, st, d);
        break;
    case IPR_K_GATHER_LIVE: {
        /* launched every superstep, it usually exits at once; when it runs,
         * its 4-byte loads need about eight wavefronts per SIMD in flight
         * (2048 workgroups: 512 left it at a third of the read rate) */
        uint64_t wb = ((d.V + 63) / 64 + IPR_WAVES - 1) / IPR_WAVES;
        hipLaunchKernelGGL(ipr_k_gather_live, dim3((uint32_t)(wb < 1 ? 1 : (wb > 2048 ? 2048 : wb))),
                           dim3(IPR_BLOCK), 0, st, d);
        break;
    }
    case IPR_K_SPARSE_STEP:
#if IPR_VARIANT == 2
        hipLaunchKernelGGL(ipr_k_sparse_step, dim3(d.fused_blocks ? d.fused_blocks : 1024), dim3(IPR_BLOCK), 0, st,
                           d);
        break;
#else
        return -1;
#endif
    case IPR_K_PULL_STEP: {
        uint32_t nb[3];
        ipr_bin_blocks(d, nb);
        if (d.pull_ordered || IPR_COMBINER == IPR_COMBINER_GENERIC)
            hipLaunchKernelGGL(ipr_k_pull_step<true>, dim3(nb[0] + nb[1] + nb[2]), dim3(IPR_BLOCK), 0, st,
                               d, nb[0], nb[1]);
        else
            hipLaunchKernelGGL(ipr_k_pull_step<false>, dim3(nb[0] + nb[1] + nb[2]), dim3(IPR_BLOCK), 0, st,
                               d, nb[0], nb[1]);
        break;
    }
    default:
        return -1;
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : -(int)e;
}

static int ipr_app_bind(const ipr_dev* d)
{
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(ipr_d_dev), d, sizeof(ipr_dev), 0, hipMemcpyHostToDevice);
    if (e != hipSuccess) return -(int)e;
    int dev = 0, cus = 0, occ = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(ipr_k_compute_push),
                                                         IPR_BLOCK, 0);
    if (e != hipSuccess) return -(int)e;
    ipr_compute_resident = occ > 0 && cus > 0 ? (uint32_t)occ * (uint32_t)cus : 0u;
    return 0;
}

/* the declared combiner must agree with ip_combine on sample pairs */
__global__ void ipr_k_check_combiner(const IP_MESSAGE_TYPE* a, const IP_MESSAGE_TYPE* b, int n, int* bad)
{
    ipr_sm_superstep = 0;   /* ip_combine may ask (it runs before the first superstep) */
    ipr_sm_counted = 0;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    IP_MESSAGE_TYPE x = a[i], y = x;
    ip_combine(&x, b[i]);
    ipr_policy::fold(&y, b[i]);
    if (!ipr_same_bits(x, y)) atomicAdd(bad, 1);
}

static int ipr_app_check_combiner(void)
{
    if (IPR_COMBINER == IPR_COMBINER_GENERIC) return 0;
    const int n = 256;
    IP_MESSAGE_TYPE ha[n], hb[n];
    uint64_t st = 0x1234567u;
    for (int i = 0; i < n; i++) {
        unsigned char* pa = (unsigned char*)&ha[i];
        unsigned char* pb = (unsigned char*)&hb[i];
        for (size_t k = 0; k < sizeof(IP_MESSAGE_TYPE); k++) {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            pa[k] = (unsigned char)(st >> 56);
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            pb[k] = (unsigned char)(st >> 56);
        }
        if (std::is_floating_point<IP_MESSAGE_TYPE>::value) {
            double fa = (double)(int64_t)(st >> 20) / 1048576.0, fb = (double)(i * 37 % 101) / 7.0;
            ha[i] = (IP_MESSAGE_TYPE)fa;
            hb[i] = (IP_MESSAGE_TYPE)fb;
        }
    }
    if (IPR_COMBINER == IPR_COMBINER_SUM && std::is_floating_point<IP_MESSAGE_TYPE>::value) {
        for (int i = 0; i < n; i++) { ha[i] = (IP_MESSAGE_TYPE)(i * 0.5); hb[i] = (IP_MESSAGE_TYPE)(i * 0.25); }
    }
    IP_MESSAGE_TYPE *da = nullptr, *db = nullptr;
    int* dbad = nullptr;
    int hbad = 0;
    if (hipMalloc(&da, sizeof(ha)) != hipSuccess) return -1;
    if (hipMalloc(&db, sizeof(hb)) != hipSuccess) { (void)hipFree(da); return -1; }
    if (hipMalloc(&dbad, sizeof(int)) != hipSuccess) { (void)hipFree(da); (void)hipFree(db); return -1; }
    int rc = 0;
    if (hipMemcpy(da, ha, sizeof(ha), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, hb, sizeof(hb), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dbad, &hbad, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) rc = -1;
    if (!rc) {
        hipLaunchKernelGGL(ipr_k_check_combiner, dim3(1), dim3(n), 0, 0, da, db, n, dbad);
        if (hipMemcpy(&hbad, dbad, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    }
    (void)hipFree(da); (void)hipFree(db); (void)hipFree(dbad);
    if (rc) return rc;
    return hbad ? 1 : 0;
}

#endif /* IPREGEL_AMD_KERNELS_H */
