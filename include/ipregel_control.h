/*
 * ipregel_control.h - the end of a superstep on the device: halting test,
 * ActiveVertexCount, W and the per-superstep trace, superstep increment.
 * HIP device code, shared by the finalize kernel of libipregel_hip.so and by
 * the application-TU kernels that end their superstep themselves: there, the
 * last workgroup to finish does it, so the superstep needs no extra launch.
 *
 * Reference loop tests: src/combiner_postamble.h:134 (full-sweep push,
 * `while active > 0`), src/combiner_single_broadcast_postamble.h:163 (pull),
 * src/combiner_spread_postamble.h:212 (spread: `first || active > 0` with
 * active = |next frontier|, :278-285).
 *
 * Coherence: the eight L2 caches of an MI355X are not coherent with one
 * another within a kernel. Counters are only ever changed by device-scope
 * atomics, which are performed where every XCD sees them, and read here with
 * device-scope atomic loads. A workgroup waits until its own memory
 * operations (its counter atomics among them) have completed before it counts
 * itself as finished, so the last one to arrive reads complete sums. No
 * device-scope release fence is used: on this GPU it writes back the XCD's
 * whole L2, once per workgroup.
 */
#ifndef IPREGEL_CONTROL_H
#define IPREGEL_CONTROL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ipregel_engine.h"

__device__ __forceinline__ uint64_t ipr_coherent_load(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void ipr_coherent_store(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* sum (and optionally clear) one spread counter: lanes 0..63 of one
 * wavefront, lane k owning slot k */
__device__ inline uint64_t ipr_take(const ipr_dev& d, int id, bool clear)
{
    uint64_t* p = &d.cnt[IPR_CSLOT(id, threadIdx.x & 63)];
    uint64_t x = ipr_coherent_load(p);
    if (clear) ipr_coherent_store(p, 0);
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

/* the largest of one spread counter's slots (optionally cleared): lanes
 * 0..63 of one wavefront */
__device__ inline uint64_t ipr_take_max(const ipr_dev& d, int id, bool clear)
{
    uint64_t* p = &d.cnt[IPR_CSLOT(id, threadIdx.x & 63)];
    uint64_t x = ipr_coherent_load(p);
    if (clear) ipr_coherent_store(p, 0);
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t y = __shfl_xor(x, o);
        x = y > x ? y : x;
    }
    return x;
}

/* Called by every thread of workgroups 0..n-1 after their last store and
 * atomic of the kernel; true in the workgroup that finished last. Arrivals are
 * counted in two levels, so that no address takes more than n/64 + 64 of
 * them: workgroup b counts in slot b mod 64, and the workgroup that
 * completes a slot counts the slot in the top counter. Counters are left at
 * 0 for the next kernel. */
__device__ inline bool ipr_last_workgroup(const ipr_dev& d, uint32_t n)
{
    __shared__ uint32_t s_last;
    /* this wavefront's memory operations have completed (s_waitcnt with
     * every counter at 0) before the workgroup counts itself */
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t slot = blockIdx.x % IPR_CSLOTS;
        const uint32_t share = n / IPR_CSLOTS + (slot < n % IPR_CSLOTS ? 1u : 0u);
        const uint32_t slots = n < IPR_CSLOTS ? n : IPR_CSLOTS;
        unsigned long long* sc = (unsigned long long*)&d.cnt[IPR_CSLOT(IPR_C_ARRIVE, slot)];
        unsigned long long* top = (unsigned long long*)&d.cnt[IPR_CSLOT(IPR_C_ARRIVE_TOP, 0)];
        uint32_t last = 0;
        if (atomicAdd(sc, 1ull) + 1 == (unsigned long long)share) {
            atomicExch(sc, 0ull);
            if (atomicAdd(top, 1ull) + 1 == (unsigned long long)slots) {
                atomicExch(top, 0ull);
                last = 1;
            }
        }
        s_last = last;
    }
    __syncthreads();
    return s_last != 0;
}

/* End of superstep s (PULL: of s-1, after the fused fetch(s-1)+compute(s)),
 * run by lanes 0..63 of one wavefront. SPREAD: `flags_front` is 1 when the
 * next frontier was left as raised receive flags (the fused sparse kernel),
 * 0 when it was compacted into a list. */
__device__ inline void ipr_finalize_step(const ipr_dev& d, int variant, int flags_front = 0)
{
    ipr_ctl* c = d.ctl;
    if (c->done) return;
    const uint64_t s = c->superstep;
    uint64_t fin, active, edges, computed;
    if (variant == IPR_VARIANT_PULL) {
        if (s == 0) {
            if ((threadIdx.x & 63) == 0) c->superstep = 1;
            return;
        }
        fin = s - 1;
        active = ipr_take(d, IPR_C_ACT0 + (int)(fin & 1), true);
        const uint64_t bc = ipr_take(d, IPR_C_BCAST0 + (int)(fin & 1), true);
        computed = ipr_take(d, IPR_C_WORKV, true);
        edges = bc ? d.in_idx[d.V] : 0;
    } else {
        fin = s;
        /* spread: the next frontier's size, from the compaction's list counter
         * or from the fused kernel's count of activated receivers */
        active = variant == IPR_VARIANT_PUSH ? ipr_take(d, IPR_C_ACT0 + (int)(s & 1), true)
                                             : ipr_coherent_load(&c->fsize[(s + 1) & 1]) +
                                                   ipr_take(d, IPR_C_FNEXT, true);
        edges = ipr_take(d, IPR_C_EDGES, true);
        computed = ipr_take(d, IPR_C_WORKV, true);
        if (variant != IPR_VARIANT_PUSH) {
            /* the out-arcs of the frontier just compacted, for the next
             * superstep's fused-or-general choice */
            const uint64_t fa = ipr_take(d, IPR_C_FARCS, true);
            const uint64_t fm = ipr_take_max(d, IPR_C_FMAX, true);
            /* after a superstep that gathered the compaction does not add
             * them up: unknown, so the next superstep is not fused */
            const bool dense = d.dir_opt && edges * (uint64_t)d.dir_alpha > d.E;
            if ((threadIdx.x & 63) == 0) {
                c->farcs[(s + 1) & 1] = dense ? ~0ull : fa;
                c->fmaxdeg[(s + 1) & 1] = dense ? ~0ull : fm;
            }
        }
        /* the gather's range counters start from 0 in the next superstep */
        ipr_coherent_store(&d.cnt[IPR_CSLOT(IPR_C_GNEXT, threadIdx.x & 63)], 0);
        /* and the large-broadcaster registrations of the next superstep too */
        ipr_coherent_store(&d.cnt[IPR_CSLOT(IPR_C_LPACK, threadIdx.x & 63)], 0);
    }
    if ((threadIdx.x & 63) != 0) return;
    if (variant == IPR_VARIANT_PULL) {
        c->done = active == 0;
        c->executed = fin + 1;
        c->superstep = c->done ? s : s + 1;
    } else {
        if (variant != IPR_VARIANT_PUSH) {
            /* superstep s's frontier is consumed: its list counter and its
             * out-arc figures are reused by superstep s+2 */
            c->fsize[s & 1] = 0;
            c->farcs[s & 1] = 0;
            c->fmaxdeg[s & 1] = 0;
            c->fsize[(s + 1) & 1] = active;
            c->front_flags = (uint64_t)flags_front;
        }
        c->done = active == 0;
        c->executed = s + 1;
        c->superstep = s + 1;
    }
    c->step_edges = edges;
    c->work_edges += edges;
    c->work_vertices += computed;
    /* the delivery kernels gathered this superstep by the same rule */
    c->pull_mode = variant != IPR_VARIANT_PULL && d.dir_opt && edges * (uint64_t)d.dir_alpha > d.E;
    if (c->pull_mode) c->gathered++;
    c->last_active = active;
    if (fin < c->trace_cap) {
        d.trace_active[fin] = active;
        d.trace_edges[fin] = edges;
        d.trace_ts[fin + 1] = wall_clock64();
    }
}

#endif /* IPREGEL_CONTROL_H */
