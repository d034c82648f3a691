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

/* Called by every thread of workgroups 0..n-1 after their last store and
 * atomic of the kernel; true in the workgroup that finished last. Arrivals are
 * counted in two levels, so that no address takes more than n/64 + 64 of
 * them: workgroup b counts in slot b mod 64, and the workgroup that
 * completes a slot counts the slot in the top counter. The last workgroup
 * (every slot complete by then) sets the counters back to 0 for the next
 * kernel, with stores issued after its two atomics rather than an exchange
 * after each. */
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
        if (atomicAdd(sc, 1ull) + 1 == (unsigned long long)share &&
            atomicAdd(top, 1ull) + 1 == (unsigned long long)slots)
            last = 1;
        s_last = last;
    }
    __syncthreads();
    const bool last = s_last != 0;
    if (last && threadIdx.x < IPR_CSLOTS) {
        ipr_coherent_store(&d.cnt[IPR_CSLOT(IPR_C_ARRIVE, threadIdx.x)], 0);
        if (threadIdx.x == 0) ipr_coherent_store(&d.cnt[IPR_CSLOT(IPR_C_ARRIVE_TOP, 0)], 0);
    }
    return last;
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
    const uint32_t lane = threadIdx.x & 63;
    if (variant == IPR_VARIANT_PULL) {
        if (s == 0) {
            if (lane == 0) c->superstep = 1;
            return;
        }
        fin = s - 1;
        /* every figure loaded before any is cleared or reduced: one round
         * trip (a load waited for after a store waits for the store too) */
        uint64_t* pa = &d.cnt[IPR_CSLOT(IPR_C_ACT0 + (int)(fin & 1), lane)];
        uint64_t* pb = &d.cnt[IPR_CSLOT(IPR_C_BCAST0 + (int)(fin & 1), lane)];
        uint64_t* pw = &d.cnt[IPR_CSLOT(IPR_C_WORKV, lane)];
        uint64_t xa = ipr_coherent_load(pa), xb = ipr_coherent_load(pb), xw = ipr_coherent_load(pw);
        const uint64_t in_arcs = d.in_idx[d.V];
        ipr_coherent_store(pa, 0);
        ipr_coherent_store(pb, 0);
        ipr_coherent_store(pw, 0);
        for (int o = 32; o > 0; o >>= 1) {
            xa += __shfl_xor(xa, o);
            xb += __shfl_xor(xb, o);
            xw += __shfl_xor(xw, o);
        }
        active = xa;
        computed = xw;
        edges = xb ? in_arcs : 0;
    } else {
        fin = s;
        const bool push = variant == IPR_VARIANT_PUSH;
        /* the figures, loaded together: ActiveVertexCount (push) or the
         * fused kernel's count of activated receivers (spread, added to the
         * compaction's list counter), W, computed vertices and, for spread,
         * the out-arcs and largest out-degree of the frontier just compacted */
        uint64_t* pa = &d.cnt[IPR_CSLOT(push ? IPR_C_ACT0 + (int)(s & 1) : IPR_C_FNEXT, lane)];
        uint64_t* pe = &d.cnt[IPR_CSLOT(IPR_C_EDGES, lane)];
        uint64_t* pw = &d.cnt[IPR_CSLOT(IPR_C_WORKV, lane)];
        uint64_t* pf = &d.cnt[IPR_CSLOT(IPR_C_FARCS, lane)];
        uint64_t* pm = &d.cnt[IPR_CSLOT(IPR_C_FMAX, lane)];
        uint64_t xa = ipr_coherent_load(pa), xe = ipr_coherent_load(pe), xw = ipr_coherent_load(pw);
        uint64_t xf = 0, xm = 0, fsz = 0;
        if (!push) {
            xf = ipr_coherent_load(pf);
            xm = ipr_coherent_load(pm);
            fsz = ipr_coherent_load(&c->fsize[(s + 1) & 1]);
        }
        ipr_coherent_store(pa, 0);
        ipr_coherent_store(pe, 0);
        ipr_coherent_store(pw, 0);
        if (!push) {
            ipr_coherent_store(pf, 0);
            ipr_coherent_store(pm, 0);
        }
        /* the gather's range counters and the large-broadcaster
         * registrations start from 0 in the next superstep */
        ipr_coherent_store(&d.cnt[IPR_CSLOT(IPR_C_GNEXT, lane)], 0);
        ipr_coherent_store(&d.cnt[IPR_CSLOT(IPR_C_LPACK, lane)], 0);
        for (int o = 32; o > 0; o >>= 1) {
            xa += __shfl_xor(xa, o);
            xe += __shfl_xor(xe, o);
            xw += __shfl_xor(xw, o);
            xf += __shfl_xor(xf, o);
            const uint64_t y = __shfl_xor(xm, o);
            xm = y > xm ? y : xm;
        }
        active = push ? xa : fsz + xa;
        edges = xe;
        computed = xw;
        if (!push) {
            /* after a superstep that gathered the compaction does not add up
             * the frontier's out-arcs: unknown, so the next superstep is not
             * fused */
            const bool dense = d.dir_opt && edges * (uint64_t)d.dir_alpha > d.E;
            if (lane == 0) {
                c->farcs[(s + 1) & 1] = dense ? ~0ull : xf;
                c->fmaxdeg[(s + 1) & 1] = dense ? ~0ull : xm;
            }
        }
    }
    if (lane != 0) return;
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
    if (c->pull_mode) {
        c->gathered++;
        if (fin < 64) c->gathered_mask |= 1ull << fin;
    }
    c->last_active = active;
    if (fin < c->trace_cap) {
        d.trace_active[fin] = active;
        d.trace_edges[fin] = edges;
        d.trace_ts[fin + 1] = wall_clock64();
    }
}

#endif /* IPREGEL_CONTROL_H */
