/*
 * ipregel_kernels.h - the superstep kernels that call the application hooks.
 * Included by iPregel.h, i.e. instantiated once, in the application's
 * translation unit, with the application's message and value types.
 *
 * gfx950: 64-lane wavefronts, 256-thread workgroups (4 waves).
 *
 * Push variants (PUSH, SPREAD):
 *   compute  every vertex of the work set (all V at superstep 0 and in PUSH,
 *            the frontier list in SPREAD) runs ip_compute; a broadcast is
 *            recorded (message in bcm[0], vertex appended to one of three
 *            out-degree bins with one atomic per wavefront).
 *   expand   the recorded broadcasts are delivered: bin SMALL by
 *            lane-cooperative expansion (64 vertices per wavefront, edges
 *            spread over the lanes by a wavefront prefix sum), bin MEDIUM one
 *            wavefront per vertex, bin LARGE the whole grid per vertex.
 *            A delivery combines into mbox[(s+1)&1] with a hardware atomic
 *            (declared combiner) or a CAS loop (ip_combine); the first
 *            delivery to a vertex sets its claim bit and, in SPREAD, appends
 *            it to the next frontier (src/combiner_spread_postamble.h:202-227).
 *   gather   dense supersteps of push applications whose combiner makes the
 *            result order-independent: each vertex folds the broadcasts of
 *            its in-neighbours (CSC) without atomics; same messages, same
 *            frontier.
 * Pull variant (PULL): one kernel per superstep fetches superstep s-1's
 *   broadcasts over the in-lists and runs ip_compute for superstep s on the
 *   same vertex (broadcast slots double-buffered by superstep parity).
 */
#ifndef IPREGEL_AMD_KERNELS_H
#define IPREGEL_AMD_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits>
#include <type_traits>

#define IPR_BLOCK 256
#define IPR_WAVE 64

/* device copy of the engine state, read by the device-side API functions */
__device__ ipr_dev ipr_d_dev;

enum { IPR_ERR_SEND_IN_PULL = 1, IPR_ERR_MIXED_BROADCAST = 2 };

__device__ inline void ipr_raise_error(uint32_t code)
{
    atomicOr(&ipr_d_dev.ctl->error, code);
}

/* ------------------------------------------------------------------------ */
/* wavefront helpers                                                         */
/* ------------------------------------------------------------------------ */

__device__ inline uint32_t ipr_lane(void) { return __lane_id(); }

__device__ inline uint32_t ipr_lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

/* Append `item` to list[counter++] for every lane with pred; one atomic per
 * wavefront. Must be reached by all active lanes. */
__device__ inline void ipr_wave_append(unsigned long long* counter, uint32_t* list,
                                       bool pred, uint32_t item)
{
    uint64_t mask = __ballot(pred);
    if (mask == 0) return;
    uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    unsigned long long base = 0;
    if (ipr_lane() == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    if (pred) list[base + ipr_lanes_below(mask)] = item;
}

/* wavefront sum of a 64-bit counter, added once to *dst */
__device__ inline void ipr_wave_add(unsigned long long* dst, uint64_t x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    uint64_t mask = __ballot(1);
    if (ipr_lane() == (uint32_t)(__ffsll((unsigned long long)mask) - 1) && x)
        atomicAdd(dst, (unsigned long long)x);
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
    __host__ __device__ static T ident() { return std::numeric_limits<T>::has_infinity ? std::numeric_limits<T>::infinity() : std::numeric_limits<T>::max(); }
    __device__ static slot_t empty() { return ident(); }
    __device__ static T value(slot_t s) { return s; }
    __device__ static void fold(T* acc, T m) { if (m < *acc) *acc = m; }
    /* returns true when this delivery may be the first one */
    __device__ static bool deliver(slot_t* p, T m)
    {
        T cur = *p;                       /* stale reads only overestimate */
        if (m < cur) return ipr_atomic_min(p, m) == ident();
        return cur == ident();            /* m == identity on an empty slot */
    }
};

template <typename T> struct ipr_comb<T, IPR_COMBINER_MAX> {
    typedef T slot_t;
    static constexpr int exact_gather = std::is_integral<T>::value;
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
};

/* SUM: zero-initialised slot, every delivery adds; firstness from the claim
 * bitmap alone. */
template <typename T> struct ipr_comb<T, IPR_COMBINER_SUM> {
    typedef T slot_t;
    static constexpr int exact_gather = std::is_integral<T>::value;
    __host__ __device__ static T ident() { return (T)0; }
    __device__ static slot_t empty() { return (T)0; }
    __device__ static T value(slot_t s) { return s; }
    __device__ static void fold(T* acc, T m) { *acc += m; }
    __device__ static bool deliver(slot_t* p, T m)
    {
        atomicAdd(p, m);
        return true;
    }
};

/* GENERIC: 32-bit message packed with a presence flag in one 64-bit slot,
 * combined with the application's ip_combine in a CAS loop
 * (src/combiner_postamble.h:136-147 combines the same way, under a lock for
 * the first message). */
template <typename T> struct ipr_comb<T, IPR_COMBINER_GENERIC> {
    static_assert(sizeof(T) == 4,
                  "without -DIP_COMBINER_MIN/_MAX/_SUM, messages must be 4 bytes");
    typedef unsigned long long slot_t;
    static constexpr int exact_gather = 0;
    __host__ __device__ static T ident() { return T(); }
    __device__ static slot_t empty() { return 0ull; }
    __device__ static T value(slot_t s) { uint32_t b = (uint32_t)s; T t; memcpy(&t, &b, 4); return t; }
    __device__ static void fold(T* acc, T m) { ip_combine(acc, m); }
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
};

typedef ipr_comb<IP_MESSAGE_TYPE, IPR_COMBINER> ipr_policy;
typedef ipr_policy::slot_t ipr_slot_t;

/* broadcast record of a push-variant vertex: the superstep it broadcast in
 * (plus one, bit 31 set for ip_broadcast_weighted) and its message; one load
 * gives a gathering receiver both */
struct ipr_bslot_t {
    uint32_t step;
    IP_MESSAGE_TYPE msg;
};
#define IPR_BSTEP_WEIGHTED 0x80000000u

/* set the claim bit of u; true for the one caller that set it */
__device__ inline bool ipr_claim(uint32_t* bits, uint32_t u)
{
    uint32_t bit = 1u << (u & 31);
    uint32_t* w = bits + (u >> 5);
    if (__atomic_load_n(w, __ATOMIC_RELAXED) & bit) return false;
    return (atomicOr(w, bit) & bit) == 0;
}

/* deliver one message to u in superstep s: combine into the mailbox of
 * superstep s+1 and mark u as having received (the claim bitmap, from which
 * the next frontier is compacted and PUSH's swap reads has_message_next) */
__device__ inline void ipr_deliver(const ipr_dev& d, uint64_t s, uint32_t u, IP_MESSAGE_TYPE m)
{
    ipr_slot_t* mb = (ipr_slot_t*)d.mbox[(s + 1) & 1];
    if (ipr_policy::deliver(mb + u, m)) ipr_claim(d.claim[(s + 1) & 1], u);
}

/* ip_send_message in push variants: immediate delivery */
__device__ inline void ipr_send_immediate(IP_VERTEX_ID_TYPE id, IP_MESSAGE_TYPE m)
{
    const ipr_dev& d = ipr_d_dev;
    ipr_deliver(d, d.ctl->superstep, (uint32_t)id, m);
}

/* ------------------------------------------------------------------------ */
/* vertex proxy fill                                                         */
/* ------------------------------------------------------------------------ */

__device__ inline void ipr_fill_vertex(const ipr_dev& d, struct ip_vertex_t* x, uint32_t v)
{
    x->id = (IP_VERTEX_ID_TYPE)v;
    x->value = ((const IP_VALUE_TYPE*)d.val)[v];
#ifdef IP_NEEDS_OUT_NEIGHBOUR_COUNT
    x->out_neighbour_count = (IP_NEIGHBOUR_COUNT_TYPE)(d.out_idx[v + 1] - d.out_idx[v]);
#endif
#ifdef IP_NEEDS_OUT_NEIGHBOUR_IDS
    x->out_neighbours = (IP_VERTEX_ID_TYPE*)(d.out_adj + d.out_idx[v]);
#endif
#ifdef IP_NEEDS_IN_NEIGHBOUR_COUNT
    x->in_neighbour_count = (IP_NEIGHBOUR_COUNT_TYPE)(d.in_idx[v + 1] - d.in_idx[v]);
#endif
#ifdef IP_NEEDS_IN_NEIGHBOUR_IDS
    x->in_neighbours = (IP_VERTEX_ID_TYPE*)(d.in_adj + d.in_idx[v]);
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
        d.act[v] = 1;  /* ip_init_vertex_range, e.g. src/combiner_postamble.h:189 */
        d.hm[v] = 0;
#if IPR_IS_PULL
        ((uint8_t*)d.bcf[0])[v] = 0;
        ((uint8_t*)d.bcf[1])[v] = 0;
#else
        ((ipr_slot_t*)d.mbox[0])[v] = ipr_policy::empty();
        ((ipr_slot_t*)d.mbox[1])[v] = ipr_policy::empty();
        ((ipr_bslot_t*)d.bslot)[v].step = 0;
#endif
    }
#if !IPR_IS_PULL
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < d.words; w += stride) {
        d.claim[0][w] = 0;
        d.claim[1][w] = 0;
    }
#endif
}

/* ------------------------------------------------------------------------ */
/* push variants: compute                                                    */
/* ------------------------------------------------------------------------ */

/* SPREAD: src/combiner_spread_postamble.h:364-401 (superstep 0 over all V,
 * then the frontier) and :459-471 (the frontier's mailboxes become messages:
 * here read straight from the mailbox, which is reset for reuse two
 * supersteps on). PUSH: src/combiner_postamble.h:246-259 (every vertex with
 * active or has_message). A broadcast is recorded in bslot and the vertex is
 * appended to its out-degree bin. */
__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_compute_push(ipr_dev d)
{
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    const uint64_t s = ctl->superstep;
    uint64_t n = d.V;
    const uint32_t* list = nullptr;
    if (IPR_VARIANT == IPR_VARIANT_SPREAD && s > 0) {
        n = ctl->fsize[s & 1];
        list = d.front[s & 1];
    }
    ipr_slot_t* mcur = (ipr_slot_t*)d.mbox[s & 1];
    ipr_bslot_t* bs = (ipr_bslot_t*)d.bslot;
    uint64_t n_active = 0, n_computed = 0, n_edges = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t v = list ? list[i] : (uint32_t)i;
        struct ip_vertex_t x;
        bool run = true;
        if (IPR_VARIANT == IPR_VARIANT_SPREAD) {
            x.active = true;
            x.has_message = s > 0;
            if (s > 0) {
                x.message = ipr_policy::value(mcur[v]);
                mcur[v] = ipr_policy::empty();
            }
        } else {
            x.active = d.act[v] != 0;
            x.has_message = d.hm[v] != 0;
            if (x.has_message) x.message = ((const IP_MESSAGE_TYPE*)d.msg)[v];
            run = x.active || x.has_message;
            x.active = true;
        }
        bool small = false, medium = false;
        uint32_t entry = 0;
        if (run) {
            ipr_fill_vertex(d, &x, v);
            const IP_VALUE_TYPE before = x.value;
            const bool hm_before = x.has_message;
            ip_compute(&x);
            ipr_store_value(d, v, before, x.value);
            n_computed++;
            if (IPR_VARIANT == IPR_VARIANT_PUSH) {
                d.act[v] = x.active;
                if (hm_before != x.has_message) d.hm[v] = x.has_message;
                n_active += x.active;
            }
            if (x.ipr_bcast) {
                uint64_t deg = d.out_idx[v + 1] - d.out_idx[v];
                n_edges += deg;
                if (deg > 0) {
                    ipr_bslot_t rec;
                    rec.step = (uint32_t)(s + 1) | (x.ipr_bcast_weighted ? IPR_BSTEP_WEIGHTED : 0u);
                    rec.msg = x.ipr_bcast_msg;
                    bs[v] = rec;
                    entry = v | (x.ipr_bcast_weighted ? 0x80000000u : 0u);
                    if (deg <= d.bin_small_max) {
                        small = true;
                    } else if (deg <= d.bin_medium_max) {
                        medium = true;
                    } else {
                        /* one atomic hands out both the list slot and the
                         * vertex's first arc chunk, so lbase ascends with the
                         * slot and a chunk finds its vertex by binary search */
                        uint64_t chunks = (deg + d.large_chunk - 1) / d.large_chunk;
                        unsigned long long old = atomicAdd((unsigned long long*)&ctl->lpack,
                                                           (1ull << IPR_LPACK_SHIFT) | chunks);
                        uint64_t slot = old >> IPR_LPACK_SHIFT;
                        d.lbase[slot] = old & IPR_LPACK_MASK;
                        d.bins[IPR_BIN_LARGE][slot] = entry;
                    }
                }
            }
        }
        ipr_wave_append((unsigned long long*)&ctl->bsize[IPR_BIN_SMALL], d.bins[IPR_BIN_SMALL], small, entry);
        ipr_wave_append((unsigned long long*)&ctl->bsize[IPR_BIN_MEDIUM], d.bins[IPR_BIN_MEDIUM], medium, entry);
    }
    ipr_wave_add((unsigned long long*)&ctl->cnt[s & 1], n_active);
    ipr_wave_add((unsigned long long*)&ctl->work_vertices, n_computed);
    ipr_wave_add((unsigned long long*)&ctl->step_edges, n_edges);
}

/* ------------------------------------------------------------------------ */
/* push variants: expansion (sparse supersteps)                              */
/* ------------------------------------------------------------------------ */

__device__ inline IP_MESSAGE_TYPE ipr_edge_msg(const uint32_t* wgt, IP_MESSAGE_TYPE m, bool weighted, uint64_t e)
{
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
    if (weighted) return (IP_MESSAGE_TYPE)(m + (IP_MESSAGE_TYPE)wgt[e]);
#else
    (void)wgt; (void)weighted; (void)e;
#endif
    return m;
}

__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_expand(ipr_dev d)
{
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    if (ctl->pull_mode) return;
    const uint64_t s = ctl->superstep;
    const ipr_bslot_t* bs = (const ipr_bslot_t*)d.bslot;
    const uint32_t lane = ipr_lane();
    const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t gthreads = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t gwave = gtid / IPR_WAVE;
    const uint64_t nwaves = gthreads / IPR_WAVE;

    /* LARGE: one workgroup per chunk of large_chunk arcs */
    const uint64_t lp = ctl->lpack;
    const uint64_t nL = lp >> IPR_LPACK_SHIFT, nC = lp & IPR_LPACK_MASK;
    for (uint64_t c = blockIdx.x; c < nC; c += gridDim.x) {
        uint64_t lo = 0, hi = nL;            /* last j with lbase[j] <= c */
        while (hi - lo > 1) {
            uint64_t mid = (lo + hi) / 2;
            if (d.lbase[mid] <= c) lo = mid; else hi = mid;
        }
        uint32_t e0 = d.bins[IPR_BIN_LARGE][lo];
        uint32_t v = e0 & 0x7fffffffu;
        bool wf = (e0 >> 31) != 0;
        uint64_t off = d.out_idx[v], deg = d.out_idx[v + 1] - off;
        uint64_t beg = (c - d.lbase[lo]) * d.large_chunk;
        uint64_t end = beg + d.large_chunk < deg ? beg + d.large_chunk : deg;
        IP_MESSAGE_TYPE m = bs[v].msg;
        for (uint64_t k = beg + threadIdx.x; k < end; k += blockDim.x)
            ipr_deliver(d, s, d.out_adj[off + k], ipr_edge_msg(d.out_wgt, m, wf, off + k));
    }

    /* MEDIUM: one wavefront per vertex */
    const uint64_t nM = ctl->bsize[IPR_BIN_MEDIUM];
    for (uint64_t w = gwave; w < nM; w += nwaves) {
        uint32_t e0 = d.bins[IPR_BIN_MEDIUM][w];
        uint32_t v = e0 & 0x7fffffffu;
        bool wf = (e0 >> 31) != 0;
        uint64_t off = d.out_idx[v], deg = d.out_idx[v + 1] - off;
        IP_MESSAGE_TYPE m = bs[v].msg;
        for (uint64_t k = lane; k < deg; k += IPR_WAVE)
            ipr_deliver(d, s, d.out_adj[off + k], ipr_edge_msg(d.out_wgt, m, wf, off + k));
    }

    /* SMALL: 64 vertices per wavefront, their arcs spread over the lanes by a
     * wavefront prefix sum of the degrees */
    const uint64_t nS = ctl->bsize[IPR_BIN_SMALL];
    for (uint64_t base = gwave * IPR_WAVE; base < nS; base += nwaves * IPR_WAVE) {
        uint64_t i = base + lane;
        uint32_t deg = 0;
        uint64_t off = 0;
        bool wf = false;
        IP_MESSAGE_TYPE m = IP_MESSAGE_TYPE();
        if (i < nS) {
            uint32_t e0 = d.bins[IPR_BIN_SMALL][i];
            uint32_t v = e0 & 0x7fffffffu;
            wf = (e0 >> 31) != 0;
            off = d.out_idx[v];
            deg = (uint32_t)(d.out_idx[v + 1] - off);
            m = bs[v].msg;
        }
        uint32_t incl = ipr_wave_incl_scan(deg);
        uint32_t excl = incl - deg;
        uint32_t total = __shfl(incl, IPR_WAVE - 1);
        for (uint32_t t0 = 0; t0 < total; t0 += IPR_WAVE) {
            uint32_t t = t0 + lane;
            /* owner: the last lane whose exclusive start is <= t */
            uint32_t lo = 0;
            for (uint32_t step = 32; step > 0; step >>= 1) {
                uint32_t probe = __shfl(excl, (lo + step) & (IPR_WAVE - 1));
                if (lo + step < IPR_WAVE && probe <= t) lo += step;
            }
            uint64_t o_off = __shfl(off, lo);
            uint32_t o_ex = __shfl(excl, lo);
            IP_MESSAGE_TYPE o_m = __shfl(m, lo);
            bool o_wf = __shfl((int)wf, lo) != 0;
            if (t < total) {
                uint64_t e = o_off + (t - o_ex);
                ipr_deliver(d, s, d.out_adj[e], ipr_edge_msg(d.out_wgt, o_m, o_wf, e));
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* push variants: gather (dense supersteps)                                  */
/* ------------------------------------------------------------------------ */

/* fold of the messages broadcast in superstep s by the in-neighbours
 * [beg, end) of one vertex, strided by `step` from `first` */
__device__ inline bool ipr_gather_range(const ipr_dev& d, uint64_t s, uint64_t first, uint64_t end,
                                        uint64_t step, IP_MESSAGE_TYPE* acc)
{
    const ipr_bslot_t* bs = (const ipr_bslot_t*)d.bslot;
    const uint32_t want = (uint32_t)(s + 1);
    bool has = false;
    for (uint64_t k = first; k < end; k += step) {
        ipr_bslot_t b = bs[d.t_adj[k]];
        if ((b.step & ~IPR_BSTEP_WEIGHTED) == want) {
            IP_MESSAGE_TYPE m = b.msg;
#ifdef IP_NEEDS_OUT_NEIGHBOUR_WEIGHTS
            if (b.step & IPR_BSTEP_WEIGHTED) m = (IP_MESSAGE_TYPE)(m + (IP_MESSAGE_TYPE)d.t_wgt[k]);
#endif
            ipr_policy::fold(acc, m);
            has = true;
        }
    }
    return has;
}

/* Receivers fold their broadcasting in-neighbours over the CSC: the same
 * combined message and the same set of receivers as the push, without
 * atomics on the arcs (exact for integer min / max / sum combiners).
 * Vertices binned by in-degree; blocks [0, nb0) one thread per vertex,
 * [nb0, nb0+nb1) one wavefront per vertex, the rest one workgroup. */
__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_gather(ipr_dev d, uint32_t nb0, uint32_t nb1)
{
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    if (!ctl->pull_mode) return;
    const uint64_t s = ctl->superstep;
    const uint32_t lane = ipr_lane();
    const uint32_t b = blockIdx.x;
    if (b < nb0) {
        const uint64_t n = d.pbin_size[0];
        for (uint64_t i = (uint64_t)b * blockDim.x + threadIdx.x; i < n; i += (uint64_t)nb0 * blockDim.x) {
            uint32_t v = d.pbins[0][i];
            IP_MESSAGE_TYPE acc = ipr_policy::ident();
            if (ipr_gather_range(d, s, d.t_idx[v], d.t_idx[v + 1], 1, &acc)) ipr_deliver(d, s, v, acc);
        }
    } else if (b < nb0 + nb1) {
        const uint64_t n = d.pbin_size[1];
        const uint64_t wave = (uint64_t)(b - nb0) * (IPR_BLOCK / IPR_WAVE) + threadIdx.x / IPR_WAVE;
        const uint64_t nwaves = (uint64_t)nb1 * (IPR_BLOCK / IPR_WAVE);
        for (uint64_t i = wave; i < n; i += nwaves) {
            uint32_t v = d.pbins[1][i];
            IP_MESSAGE_TYPE acc = ipr_policy::ident();
            bool has = ipr_gather_range(d, s, d.t_idx[v] + lane, d.t_idx[v + 1], IPR_WAVE, &acc);
            for (int o = 32; o > 0; o >>= 1) {
                IP_MESSAGE_TYPE y = __shfl_xor(acc, o);
                ipr_policy::fold(&acc, y);
            }
            has = __ballot(has) != 0;
            if (lane == 0 && has) ipr_deliver(d, s, v, acc);
        }
    } else {
        __shared__ IP_MESSAGE_TYPE part[IPR_BLOCK / IPR_WAVE];
        __shared__ int parth[IPR_BLOCK / IPR_WAVE];
        const uint64_t n = d.pbin_size[2];
        const uint32_t nb2 = gridDim.x - nb0 - nb1;
        for (uint64_t i = b - nb0 - nb1; i < n; i += nb2) {
            uint32_t v = d.pbins[2][i];
            IP_MESSAGE_TYPE acc = ipr_policy::ident();
            bool has = ipr_gather_range(d, s, d.t_idx[v] + threadIdx.x, d.t_idx[v + 1], IPR_BLOCK, &acc);
            for (int o = 32; o > 0; o >>= 1) {
                IP_MESSAGE_TYPE y = __shfl_xor(acc, o);
                ipr_policy::fold(&acc, y);
            }
            bool wh = __ballot(has) != 0;
            if (lane == 0) { part[threadIdx.x / IPR_WAVE] = acc; parth[threadIdx.x / IPR_WAVE] = wh; }
            __syncthreads();
            if (threadIdx.x == 0) {
                bool any = false;
                for (int w = 0; w < IPR_BLOCK / IPR_WAVE; w++) {
                    ipr_policy::fold(&acc, part[w]);
                    any |= parth[w] != 0;
                }
                if (any) ipr_deliver(d, s, v, acc);
            }
            __syncthreads();
        }
    }
}

/* PUSH: src/combiner_postamble.h:263-278 - a delivered mailbox becomes the
 * vertex's message and reactivates it. */
__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_swap(ipr_dev d)
{
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    const uint64_t s = ctl->superstep;
    ipr_slot_t* mb = (ipr_slot_t*)d.mbox[(s + 1) & 1];
    uint32_t* cl = d.claim[(s + 1) & 1];
    uint64_t n_react = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < d.V; v += stride) {
        const uint32_t bit = 1u << (v & 31);
        if (cl[v >> 5] & bit) {
            ((IP_MESSAGE_TYPE*)d.msg)[v] = ipr_policy::value(mb[v]);
            mb[v] = ipr_policy::empty();
            atomicAnd(cl + (v >> 5), ~bit);   /* clean for superstep s+2 */
            d.hm[v] = 1;
            if (!d.act[v]) { d.act[v] = 1; n_react++; }
        }
    }
    ipr_wave_add((unsigned long long*)&ctl->cnt[s & 1], n_react);
}

/* ------------------------------------------------------------------------ */
/* pull variant                                                              */
/* ------------------------------------------------------------------------ */

/* fold of the broadcasting in-neighbours of v in in-list order
 * (src/combiner_single_broadcast_postamble.h:161-193) */
__device__ inline bool ipr_fetch_ordered(const ipr_dev& d, uint32_t v, const uint8_t* bf,
                                         const IP_MESSAGE_TYPE* bm, IP_MESSAGE_TYPE* out)
{
    uint64_t i = d.in_idx[v], end = d.in_idx[v + 1];
    bool has = false;
    IP_MESSAGE_TYPE acc = IP_MESSAGE_TYPE();
    for (; i < end; i++) {
        uint32_t u = d.in_adj[i];
        if (bf[u]) {
            IP_MESSAGE_TYPE m = bm[u];
            if (has) ip_combine(&acc, m);
            else { acc = m; has = true; }
        }
    }
    *out = acc;
    return has;
}

/* one vertex: fetch(s-1) result -> activation -> compute(s) */
__device__ inline void ipr_pull_vertex(const ipr_dev& d, uint64_t s, uint32_t v, bool hm,
                                       IP_MESSAGE_TYPE msg, uint64_t* n_react, uint64_t* n_active,
                                       uint64_t* n_computed, uint64_t* n_bcast)
{
    bool act = d.act[v] != 0;
    const bool act0 = act;
    if (hm && !act) { act = true; (*n_react)++; }
    uint8_t* bf_cur = (uint8_t*)d.bcf[s & 1];
    bf_cur[v] = 0;
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
        (*n_computed)++;
        if (act) (*n_active)++;
        if (x.ipr_bcast) {
            ((IP_MESSAGE_TYPE*)d.bcm[s & 1])[v] = x.ipr_bcast_msg;
            bf_cur[v] = 1;
            (*n_bcast)++;
        }
    }
    if (act != act0) d.act[v] = act;
}

/* ordered: one thread per vertex over all V */
__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_pull_step_ordered(ipr_dev d)
{
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    const uint64_t s = ctl->superstep;
    const uint8_t* bf = (const uint8_t*)d.bcf[(s + 1) & 1];
    const IP_MESSAGE_TYPE* bm = (const IP_MESSAGE_TYPE*)d.bcm[(s + 1) & 1];
    uint64_t n_react = 0, n_active = 0, n_computed = 0, n_bcast = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.V; i += stride) {
        uint32_t v = (uint32_t)i;
        IP_MESSAGE_TYPE msg = IP_MESSAGE_TYPE();
        bool hm = false;
        if (s > 0) hm = ipr_fetch_ordered(d, v, bf, bm, &msg);
        ipr_pull_vertex(d, s, v, hm, msg, &n_react, &n_active, &n_computed, &n_bcast);
    }
    if (s > 0) ipr_wave_add((unsigned long long*)&ctl->cnt[(s + 1) & 1], n_react);
    ipr_wave_add((unsigned long long*)&ctl->cnt[s & 1], n_active);
    ipr_wave_add((unsigned long long*)&ctl->work_vertices, n_computed);
    ipr_wave_add((unsigned long long*)&ctl->bcast[s & 1], n_bcast);
}

/* reassociating fold for declared combiners: vertices binned by in-degree
 * (static bins built at load); thread, wavefront or workgroup per vertex.
 * Blocks [0, nb0) serve bin 0, [nb0, nb0+nb1) bin 1, the rest bin 2. */
__global__ void __launch_bounds__(IPR_BLOCK) ipr_k_pull_step_binned(ipr_dev d, uint32_t nb0, uint32_t nb1)
{
    ipr_ctl* ctl = d.ctl;
    if (ctl->done) return;
    const uint64_t s = ctl->superstep;
    const uint8_t* bf = (const uint8_t*)d.bcf[(s + 1) & 1];
    const IP_MESSAGE_TYPE* bm = (const IP_MESSAGE_TYPE*)d.bcm[(s + 1) & 1];
    uint64_t n_react = 0, n_active = 0, n_computed = 0, n_bcast = 0;
    const uint32_t lane = ipr_lane();
    const uint32_t b = blockIdx.x;
    if (b < nb0) {
        const uint64_t n = d.pbin_size[0];
        for (uint64_t i = (uint64_t)b * blockDim.x + threadIdx.x; i < n; i += (uint64_t)nb0 * blockDim.x) {
            uint32_t v = d.pbins[0][i];
            IP_MESSAGE_TYPE msg = IP_MESSAGE_TYPE();
            bool hm = false;
            if (s > 0) hm = ipr_fetch_ordered(d, v, bf, bm, &msg);
            ipr_pull_vertex(d, s, v, hm, msg, &n_react, &n_active, &n_computed, &n_bcast);
        }
    } else if (b < nb0 + nb1) {
        const uint64_t n = d.pbin_size[1];
        const uint64_t wave = (uint64_t)(b - nb0) * (IPR_BLOCK / IPR_WAVE) + threadIdx.x / IPR_WAVE;
        const uint64_t nwaves = (uint64_t)nb1 * (IPR_BLOCK / IPR_WAVE);
        for (uint64_t i = wave; i < n; i += nwaves) {
            uint32_t v = d.pbins[1][i];
            IP_MESSAGE_TYPE acc = ipr_policy::ident();
            bool has = false;
            if (s > 0) {
                uint64_t beg = d.in_idx[v], end = d.in_idx[v + 1];
                for (uint64_t k = beg + lane; k < end; k += IPR_WAVE) {
                    uint32_t u = d.in_adj[k];
                    if (bf[u]) { ipr_policy::fold(&acc, bm[u]); has = true; }
                }
                for (int o = 32; o > 0; o >>= 1) {
                    IP_MESSAGE_TYPE y = __shfl_xor(acc, o);
                    ipr_policy::fold(&acc, y);
                }
                has = __ballot(has) != 0;
            }
            if (lane == 0)
                ipr_pull_vertex(d, s, v, has, acc, &n_react, &n_active, &n_computed, &n_bcast);
        }
    } else {
        __shared__ IP_MESSAGE_TYPE part[IPR_BLOCK / IPR_WAVE];
        __shared__ int parth[IPR_BLOCK / IPR_WAVE];
        const uint64_t n = d.pbin_size[2];
        const uint32_t nb2 = gridDim.x - nb0 - nb1;
        for (uint64_t i = b - nb0 - nb1; i < n; i += nb2) {
            uint32_t v = d.pbins[2][i];
            IP_MESSAGE_TYPE acc = ipr_policy::ident();
            bool has = false;
            if (s > 0) {
                uint64_t beg = d.in_idx[v], end = d.in_idx[v + 1];
                for (uint64_t k = beg + threadIdx.x; k < end; k += IPR_BLOCK) {
                    uint32_t u = d.in_adj[k];
                    if (bf[u]) { ipr_policy::fold(&acc, bm[u]); has = true; }
                }
                for (int o = 32; o > 0; o >>= 1) {
                    IP_MESSAGE_TYPE y = __shfl_xor(acc, o);
                    ipr_policy::fold(&acc, y);
                }
                bool wh = __ballot(has) != 0;
                if (lane == 0) { part[threadIdx.x / IPR_WAVE] = acc; parth[threadIdx.x / IPR_WAVE] = wh; }
                __syncthreads();
                if (threadIdx.x == 0) {
                    acc = part[0];
                    has = parth[0];
                    for (int w = 1; w < IPR_BLOCK / IPR_WAVE; w++) {
                        ipr_policy::fold(&acc, part[w]);
                        has |= parth[w] != 0;
                    }
                }
                __syncthreads();
            }
            if (threadIdx.x == 0)
                ipr_pull_vertex(d, s, v, has, acc, &n_react, &n_active, &n_computed, &n_bcast);
        }
    }
    if (s > 0) ipr_wave_add((unsigned long long*)&ctl->cnt[(s + 1) & 1], n_react);
    ipr_wave_add((unsigned long long*)&ctl->cnt[s & 1], n_active);
    ipr_wave_add((unsigned long long*)&ctl->work_vertices, n_computed);
    ipr_wave_add((unsigned long long*)&ctl->bcast[s & 1], n_bcast);
}

/* ------------------------------------------------------------------------ */
/* launch table                                                              */
/* ------------------------------------------------------------------------ */

/* workgroups per in-degree bin: threads, wavefronts, workgroups per vertex */
static void ipr_bin_blocks(const ipr_dev& d, uint32_t nb[3])
{
    uint64_t b0 = (d.pbin_size[0] + IPR_BLOCK - 1) / IPR_BLOCK;
    uint64_t b1 = (d.pbin_size[1] + (IPR_BLOCK / IPR_WAVE) - 1) / (IPR_BLOCK / IPR_WAVE);
    uint64_t b2 = d.pbin_size[2];
    nb[0] = (uint32_t)(b0 > 8192 ? 8192 : b0);
    nb[1] = (uint32_t)(b1 > 8192 ? 8192 : b1);
    nb[2] = (uint32_t)(b2 > 8192 ? 8192 : b2);
    if (nb[0] + nb[1] + nb[2] == 0) nb[0] = 1;
}

static int ipr_app_launch(int kernel, const ipr_dev* dp, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    const ipr_dev& d = *dp;
    uint32_t vblocks = (uint32_t)((d.V + IPR_BLOCK - 1) / IPR_BLOCK);
    if (vblocks == 0) vblocks = 1;
    if (vblocks > 65536) vblocks = 65536;
    switch (kernel) {
    case IPR_K_INIT:
        hipLaunchKernelGGL(ipr_k_init, dim3(vblocks), dim3(IPR_BLOCK), 0, st, d);
        break;
    case IPR_K_COMPUTE:
        hipLaunchKernelGGL(ipr_k_compute_push, dim3(vblocks), dim3(IPR_BLOCK), 0, st, d);
        break;
    case IPR_K_EXPAND:
        hipLaunchKernelGGL(ipr_k_expand, dim3(d.grid_blocks), dim3(IPR_BLOCK), 0, st, d);
        break;
    case IPR_K_SWAP:
        hipLaunchKernelGGL(ipr_k_swap, dim3(vblocks), dim3(IPR_BLOCK), 0, st, d);
        break;
    case IPR_K_GATHER: {
        uint32_t nb[3];
        ipr_bin_blocks(d, nb);
        hipLaunchKernelGGL(ipr_k_gather, dim3(nb[0] + nb[1] + nb[2]), dim3(IPR_BLOCK), 0, st, d, nb[0], nb[1]);
        break;
    }
    case IPR_K_PULL_STEP:
        if (d.pull_ordered || IPR_COMBINER == IPR_COMBINER_GENERIC) {
            hipLaunchKernelGGL(ipr_k_pull_step_ordered, dim3(vblocks), dim3(IPR_BLOCK), 0, st, d);
        } else {
            uint32_t nb[3];
            ipr_bin_blocks(d, nb);
            hipLaunchKernelGGL(ipr_k_pull_step_binned, dim3(nb[0] + nb[1] + nb[2]), dim3(IPR_BLOCK), 0, st,
                               d, nb[0], nb[1]);
        }
        break;
    default:
        return -1;
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : -(int)e;
}

static int ipr_app_bind(const ipr_dev* d)
{
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(ipr_d_dev), d, sizeof(ipr_dev), 0, hipMemcpyHostToDevice);
    return e == hipSuccess ? 0 : -(int)e;
}

/* the declared combiner must agree with ip_combine on sample pairs */
__global__ void ipr_k_check_combiner(const IP_MESSAGE_TYPE* a, const IP_MESSAGE_TYPE* b, int n, int* bad)
{
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
