/*
 * ipr_serve.cpp - the request loop of libipregel_batch.so (no counterpart in
 * the reference, which answers one query per process).
 *
 * Queries arrive on a trace (arrival time, kind, graph, parameter). Arrival
 * times are honoured against the wall clock. Two modes:
 *
 *   sequential  the plain per-query loop: in arrival order, each query runs
 *               alone on the scalar kernels (a one-lane state per graph) to
 *               completion before the next one starts.
 *   batched     continuous batching. Per (graph, kind) there is one batcher
 *               with a 64-slot state on its own stream (for SSSP, the graph's
 *               tuning may ask for several, each arrival joining the least
 *               loaded: on long-diameter graphs a superstep is latency-bound
 *               and concurrent states overlap). At every tick boundary
 *               a batcher retires finished slots, admits waiting queries into
 *               free slots and launches the next superstep (SSSP) or iteration
 *               (PPR) as a captured HIP graph; batchers of different graphs run
 *               concurrently on their streams. CC queries of one graph that are
 *               waiting when a CC run starts share that run (the labels do not
 *               depend on the query). A batcher's state is allocated on first
 *               use (states for the trace's graphs and kinds are made before
 *               the clock starts while the budget allows) and counted against
 *               the HBM budget; a batcher that cannot fit waits, after idle
 *               batchers have been evicted (oldest first).
 *
 * A query's result is the same in both modes: the batched kernels compute a
 * slot only from its own column, in the same order as the scalar kernels.
 */
#include <hip/hip_runtime.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <deque>
#include <map>
#include <vector>

#include "ipr_batch_int.h"

using namespace iprb;

namespace {

void sleep_until(double t0, double target)
{
    const double left = target - (now_s() - t0);
    if (left > 300e-6) {
        timespec ts;
        const double s = left - 200e-6;
        ts.tv_sec = (time_t)s;
        ts.tv_nsec = (long)((s - (double)ts.tv_sec) * 1e9);
        nanosleep(&ts, nullptr);
    }
}

struct Ctx {
    const ipr_query* q;
    double* t_admit;
    double* t_done;
    uint64_t* digest;
    void** results;
    double t0;
    int ppr_iters;
    int use_graph;
    uint64_t completed = 0;
    std::vector<char> delayed;     /* query found no slot or no memory at its first tick boundary */

    double t() const { return now_s() - t0; }
    void finish(int qi, unsigned long long d)
    {
        t_done[qi] = t();
        digest[qi] = d;
        completed++;
    }
};

/* ---------------------------------------------------------------- sequential */

/* one-query states of every graph the trace uses, made before the clock starts */
struct SeqStates {
    std::vector<PushState*> push;
    std::vector<PprState*> ppr;
    ~SeqStates()
    {
        for (auto* s : push) delete s;
        for (auto* s : ppr) delete s;
    }
    int prepare(ipr_bgraph** graphs, int ng, const ipr_query* q, int n, int use_graph)
    {
        push.assign(ng, nullptr);
        ppr.assign(ng, nullptr);
        for (int i = 0; i < n; i++) {
            const int gi = q[i].graph;
            int rc = 0;
            if (q[i].kind == IPR_Q_PPR) {
                if (!ppr[gi]) rc = PprState::create(graphs[gi], 1, use_graph, &ppr[gi]);
            } else if (!push[gi]) {
                rc = PushState::create(graphs[gi], false, use_graph, &push[gi]);
            }
            if (rc) return rc;
        }
        return 0;
    }
};

int serve_sequential(ipr_bgraph** graphs, SeqStates& ss, const std::vector<int>& order, Ctx& c)
{
    std::vector<PushState*>& push = ss.push;
    std::vector<PprState*>& ppr = ss.ppr;
    int rc = 0;
    float* dbuf = nullptr;
    uint64_t dbuf_n = 0;
    for (int qi : order) {
        const ipr_query& q = c.q[qi];
        ipr_bgraph* g = graphs[q.graph];
        while (c.t() < q.t_arrival) sleep_until(c.t0, q.t_arrival);
        c.t_admit[qi] = c.t();
        if (q.kind == IPR_Q_SSSP || q.kind == IPR_Q_CC) {
            PushState* s = push[q.graph];
            if (q.kind == IPR_Q_SSSP) {
                rc = s->start_sssp(q.param);
                if (!rc) rc = s->run_to_end(64 * g->V + 64);
            } else {
                rc = s->cc(g->V + 8);
            }
            unsigned long long d = 0;
            if (!rc) rc = s->digest(&d);
            if (!rc && c.results && c.results[qi]) rc = s->values((uint32_t*)c.results[qi]);
            if (rc) break;
            c.finish(qi, d);
        } else {
            PprState* s = ppr[q.graph];
            const int slot = 0;
            rc = s->update(1, &slot, &q.param, 1ull);
            for (int it = 0; !rc && it < c.ppr_iters; it++) rc = s->step();
            float* ptr = nullptr;
            if (!rc && c.results && c.results[qi]) {
                if (dbuf_n < g->V) {
                    if (dbuf) (void)hipFree(dbuf);
                    dbuf = nullptr;
                    dbuf_n = 0;
                    const hipError_t e = hipMalloc((void**)&dbuf, g->V * 4);
                    if (e != hipSuccess) { rc = hip_fail(e, "result buffer"); break; }
                    dbuf_n = g->V;
                }
                ptr = dbuf;
            }
            if (!rc) rc = s->extract(1ull, c.ppr_iters == 1 ? 1ull : 0ull, &ptr);
            if (!rc) rc = s->record();
            if (!rc) rc = s->wait();
            if (!rc && ptr) {
                const hipError_t e = hipMemcpy(c.results[qi], ptr, g->V * 4, hipMemcpyDeviceToHost);
                if (e != hipSuccess) rc = hip_fail(e, "result readback");
            }
            if (rc) break;
            c.finish(qi, s->hout->dig[0]);
        }
    }
    if (dbuf) (void)hipFree(dbuf);
    return rc;
}

/* ---------------------------------------------------------------- batched */

struct Batcher {
    int kind = 0;
    int gi = 0;
    int rep = 0;                   /* SSSP: which of the graph's concurrent states */
    ipr_bgraph* g = nullptr;
    uint64_t footprint = 0;
    PushState* ps = nullptr;
    PprState* pp = nullptr;
    std::deque<int> waiting;
    int slot_q[64];
    uint64_t slot_tick[64];        /* SSSP: tick that carries the admission; PPR: iterations done */
    float* slot_buf[64];           /* PPR: device result buffers (only when results are asked for) */
    unsigned long long live = 0, release = 0;
    bool inflight = false;
    int parity = 0;                /* SSSP/CC: parity of the last superstep of the tick in flight */
    uint64_t run = 1;              /* SSSP: supersteps per tick (doubles while nobody waits) */
    uint64_t ticks = 0;
    unsigned long long ext_mask = 0;   /* PPR: slots extracted with the tick in flight */
    /* CC */
    std::vector<int> cc_attached;
    bool cc_running = false;
    uint64_t cc_chunk = 1;
    double last_used = 0;

    Batcher()
    {
        for (int i = 0; i < 64; i++) {
            slot_q[i] = -1;
            slot_tick[i] = 0;
            slot_buf[i] = nullptr;
        }
    }
    bool allocated() const { return ps || pp; }
    bool busy() const { return inflight || live || cc_running || !waiting.empty(); }
    void free_state()
    {
        delete ps;
        delete pp;
        ps = nullptr;
        pp = nullptr;
    }
};

struct Server {
    ipr_bgraph** graphs;
    int ng;
    Ctx& c;
    uint64_t budget;
    uint64_t used = 0;
    ipr_serve_stats* st;
    std::map<long, Batcher> batchers;

    Server(ipr_bgraph** g, int n, Ctx& ctx, uint64_t b, ipr_serve_stats* s)
        : graphs(g), ng(n), c(ctx), budget(b), st(s) {}

    Batcher& get(int kind, int gi, int rep = 0)
    {
        const long key = ((long)gi * 4 + kind) * 16 + rep;
        auto it = batchers.find(key);
        if (it != batchers.end()) return it->second;
        Batcher& b = batchers[key];
        b.kind = kind;
        b.gi = gi;
        b.rep = rep;
        b.g = graphs[gi];
        b.footprint = kind == IPR_Q_SSSP  ? PushState::footprint(b.g, true)
                      : kind == IPR_Q_PPR ? PprState::footprint(b.g, 64)
                                          : PushState::footprint(b.g, false);
        return b;
    }

    int states_of(int kind, int gi) const
    {
        return kind == IPR_Q_SSSP ? (int)std::max<uint32_t>(1u, graphs[gi]->tune.sssp_states) : 1;
    }

    /* where an arriving query waits: a graph's SSSP queries are spread over
     * its concurrent states, each arrival joining the least loaded one */
    Batcher& route(int kind, int gi)
    {
        Batcher* best = nullptr;
        size_t best_load = 0;
        for (int r = 0; r < states_of(kind, gi); r++) {
            Batcher& b = get(kind, gi, r);
            const size_t load = (size_t)__builtin_popcountll(b.live) + b.waiting.size();
            if (!best || load < best_load) {
                best = &b;
                best_load = load;
            }
        }
        return *best;
    }

    /* allocate b's state within the budget; false: it has to wait */
    int ensure_state(Batcher& b, bool* ok)
    {
        *ok = true;
        if (b.allocated()) return 0;
        if (budget && b.footprint > budget) {
            set_error("ipr_serve: the %s state of graph %d needs %llu bytes, above the HBM budget", 
                      b.kind == IPR_Q_SSSP ? "SSSP" : b.kind == IPR_Q_PPR ? "PPR" : "CC", b.gi,
                      (unsigned long long)b.footprint);
            return -4;
        }
        if (budget && used + b.footprint > budget) {
            std::vector<Batcher*> idle;
            for (auto& kv : batchers)
                if (&kv.second != &b && kv.second.allocated() && !kv.second.busy()) idle.push_back(&kv.second);
            std::sort(idle.begin(), idle.end(), [](Batcher* x, Batcher* y) { return x->last_used < y->last_used; });
            for (Batcher* x : idle) {
                if (used + b.footprint <= budget) break;
                used -= x->footprint;
                x->free_state();
            }
            if (used + b.footprint > budget) {
                *ok = false;
                return 0;
            }
        }
        int rc;
        if (b.kind == IPR_Q_SSSP) rc = PushState::create(b.g, true, c.use_graph, &b.ps);
        else if (b.kind == IPR_Q_PPR) rc = PprState::create(b.g, 64, c.use_graph, &b.pp);
        else rc = PushState::create(b.g, false, c.use_graph, &b.ps);
        if (rc) return rc;
        used += b.footprint;
        st->hbm_peak = std::max<uint64_t>(st->hbm_peak, used);
        return 0;
    }

    int free_slots(const Batcher& b) const { return 64 - __builtin_popcountll(b.live); }

    void mark_delayed(const Batcher& b)
    {
        for (int qi : b.waiting)
            if (!c.delayed[qi]) {
                c.delayed[qi] = 1;
                st->admitted_waits++;
            }
    }

    /* retire the tick in flight */
    int complete(Batcher& b)
    {
        int rc = 0;
        b.inflight = false;
        b.last_used = c.t();
        if (b.kind == IPR_Q_SSSP) {
            const TickOut& o = b.ps->hout[b.parity];
            for (int q = 0; q < 64; q++) {
                if (!((b.live >> q) & 1ull) || b.slot_tick[q] > b.ticks - 1) continue;
                if ((o.orm >> q) & 1ull) continue;
                const int qi = b.slot_q[q];
                if (c.results && c.results[qi] && (rc = b.ps->column(q, (uint32_t*)c.results[qi])) != 0) return rc;
                c.finish(qi, o.dig[q]);
                b.live &= ~(1ull << q);
                b.release |= 1ull << q;
                b.slot_q[q] = -1;
            }
        } else if (b.kind == IPR_Q_PPR) {
            for (int q = 0; q < 64; q++) {
                if (!((b.ext_mask >> q) & 1ull)) continue;
                const int qi = b.slot_q[q];
                if (b.slot_buf[q]) {
                    const hipError_t e = hipMemcpy(c.results[qi], b.slot_buf[q], b.g->V * 4, hipMemcpyDeviceToHost);
                    (void)hipFree(b.slot_buf[q]);
                    b.slot_buf[q] = nullptr;
                    if (e != hipSuccess) return hip_fail(e, "result readback");
                }
                c.finish(qi, b.pp->hout->dig[q]);
                b.live &= ~(1ull << q);
                b.release |= 1ull << q;
                b.slot_q[q] = -1;
            }
            b.ext_mask = 0;
        } else {
            if (b.g->tune.cc_mode == 0) {
                const TickOut& o = b.ps->hout[b.parity];
                if (o.nsmall + o.nlarge != 0) {
                    b.cc_chunk = std::min<uint64_t>(b.cc_chunk * 2, 32);
                    return 0;
                }
            }
            unsigned long long d = 0;
            if ((rc = b.ps->digest(&d)) != 0) return rc;
            for (int qi : b.cc_attached) {
                if (c.results && c.results[qi] && (rc = b.ps->values((uint32_t*)c.results[qi])) != 0) return rc;
                c.finish(qi, d);
            }
            b.cc_attached.clear();
            b.cc_running = false;
        }
        return 0;
    }

    /* admit what fits and launch the next tick */
    int advance(Batcher& b)
    {
        int rc = 0;
        if (!b.waiting.empty() && !b.allocated()) {
            bool ok;
            if ((rc = ensure_state(b, &ok)) != 0) return rc;
            if (!ok) {
                mark_delayed(b);
                return 0;
            }
        }
        if (!b.allocated()) return 0;
        if (b.kind == IPR_Q_SSSP) {
            int slots[64];
            uint32_t srcs[64];
            int n = 0;
            if (b.release) {
                if ((rc = b.ps->release(b.release)) != 0) return rc;
                b.release = 0;
            }
            while (!b.waiting.empty() && free_slots(b) > 0) {
                const int q = __builtin_ctzll(~b.live);
                const int qi = b.waiting.front();
                b.waiting.pop_front();
                b.live |= 1ull << q;
                b.slot_q[q] = qi;
                b.slot_tick[q] = b.ticks;
                c.t_admit[qi] = c.t();
                slots[n] = q;
                srcs[n] = c.q[qi].param;
                n++;
            }
            mark_delayed(b);
            if (n && (rc = b.ps->admit(n, slots, srcs)) != 0) return rc;
            if (!b.live) return 0;
            /* one superstep per tick while queries wait for a slot, else a
             * doubling run of supersteps between host checks */
            b.run = (n || !b.waiting.empty()) ? 1 : std::min<uint64_t>(b.run * 2, 32);
            if ((rc = b.ps->steps(b.run, &b.parity)) != 0) return rc;
        } else if (b.kind == IPR_Q_PPR) {
            int slots[64];
            uint32_t seeds[64];
            int n = 0;
            while (!b.waiting.empty() && free_slots(b) > 0) {
                const int q = __builtin_ctzll(~b.live);
                const int qi = b.waiting.front();
                b.waiting.pop_front();
                b.live |= 1ull << q;
                b.slot_q[q] = qi;
                b.slot_tick[q] = 0;
                c.t_admit[qi] = c.t();
                slots[n] = q;
                seeds[n] = c.q[qi].param;
                n++;
                if (c.results && c.results[qi]) {
                    const hipError_t e = hipMalloc((void**)&b.slot_buf[q], b.g->V * 4);
                    if (e != hipSuccess) return hip_fail(e, "result buffer");
                }
            }
            mark_delayed(b);
            if (n || b.release) {
                if ((rc = b.pp->update(n, slots, seeds, b.release)) != 0) return rc;
                b.release = 0;
            }
            if (!b.live) return 0;
            /* iterations this tick: up to the next slot completion; one while
             * queries wait for a free slot */
            uint64_t k = 8;
            for (int q = 0; q < 64; q++)
                if ((b.live >> q) & 1ull) k = std::min<uint64_t>(k, (uint64_t)c.ppr_iters - b.slot_tick[q]);
            if (!b.waiting.empty() && free_slots(b) > 0) k = 1;
            for (uint64_t i = 0; i < k; i++)
                if ((rc = b.pp->step()) != 0) return rc;
            unsigned long long fin = 0, first = 0;
            for (int q = 0; q < 64; q++) {
                if (!((b.live >> q) & 1ull)) continue;
                b.slot_tick[q] += k;
                if (b.slot_tick[q] == (uint64_t)c.ppr_iters) {
                    fin |= 1ull << q;
                    if (c.ppr_iters == 1) first |= 1ull << q;
                }
            }
            if (fin) {
                if ((rc = b.pp->extract(fin, first, b.slot_buf)) != 0) return rc;
                b.ext_mask = fin;
            }
            if ((rc = b.pp->record()) != 0) return rc;
        } else {
            if (!b.cc_running) {
                if (b.waiting.empty()) return 0;
                while (!b.waiting.empty()) {
                    const int qi = b.waiting.front();
                    b.waiting.pop_front();
                    c.t_admit[qi] = c.t();
                    b.cc_attached.push_back(qi);
                }
                if (b.g->tune.cc_mode == 1) {
                    if ((rc = b.ps->cc_union_find()) != 0) return rc;
                    b.parity = 0;
                } else if ((rc = b.ps->start_cc()) != 0) {
                    return rc;
                }
                b.cc_running = true;
                b.cc_chunk = 1;
            }
            /* label propagation: doubling runs of supersteps per tick */
            if (b.g->tune.cc_mode == 0)
                if ((rc = b.ps->steps(b.cc_chunk, &b.parity)) != 0) return rc;
        }
        b.inflight = true;
        b.ticks++;
        st->ticks++;
        return 0;
    }

    /* 64-slot states of the (graph, kind) pairs of the trace, in order of first
     * arrival, while the budget lasts; made before the clock starts */
    int prewarm(const std::vector<int>& order)
    {
        for (int qi : order) {
            for (int r = 0; r < states_of(c.q[qi].kind, c.q[qi].graph); r++) {
                Batcher& b = get(c.q[qi].kind, c.q[qi].graph, r);
                if (b.allocated()) continue;
                if (budget && used + b.footprint > budget) continue;
                bool ok;
                const int rc = ensure_state(b, &ok);
                if (rc) return rc;
            }
        }
        return 0;
    }

    bool tick_done(Batcher& b)
    {
        if (b.kind == IPR_Q_PPR) return b.pp->done();
        return hipEventQuery(b.ps->ev[b.parity]) == hipSuccess;
    }

    int run(const std::vector<int>& order, uint64_t n)
    {
        size_t next = 0;
        int rc = 0;
        while (c.completed < n) {
            const double now = c.t();
            while (next < order.size() && c.q[order[next]].t_arrival <= now) {
                const int qi = order[next++];
                route(c.q[qi].kind, c.q[qi].graph).waiting.push_back(qi);
            }
            bool any_inflight = false;
            uint64_t live_slots = 0;
            for (auto& kv : batchers) {
                Batcher& b = kv.second;
                if (b.inflight) {
                    if (!tick_done(b)) {
                        any_inflight = true;
                        live_slots += __builtin_popcountll(b.live) + b.cc_attached.size();
                        continue;
                    }
                    if ((rc = complete(b)) != 0) return rc;
                }
                if ((rc = advance(b)) != 0) return rc;
                any_inflight |= b.inflight;
                live_slots += __builtin_popcountll(b.live) + b.cc_attached.size();
            }
            st->peak_slots = std::max<uint64_t>(st->peak_slots, live_slots);
            if (!any_inflight && next < order.size()) {
                bool waiting = false;
                for (auto& kv : batchers) waiting |= !kv.second.waiting.empty();
                if (!waiting) sleep_until(c.t0, c.q[order[next]].t_arrival);
            }
        }
        return 0;
    }

    ~Server()
    {
        for (auto& kv : batchers) {
            for (int q = 0; q < 64; q++)
                if (kv.second.slot_buf[q]) (void)hipFree(kv.second.slot_buf[q]);
            kv.second.free_state();
        }
    }
};

}  // namespace

extern "C" int ipr_serve(ipr_bgraph** graphs, int ng, const ipr_query* q, int n, const ipr_serve_opts* opt,
                         double* t_admit, double* t_done, uint64_t* digest, void** results, ipr_serve_stats* st)
{
    if (!graphs || ng < 1 || (n > 0 && (!q || !t_admit || !t_done || !digest)) || n < 0) {
        set_error("ipr_serve: bad arguments");
        return -2;
    }
    for (int i = 0; i < ng; i++)
        if (!graphs[i]) {
            set_error("ipr_serve: graph %d is null", i);
            return -2;
        }
    for (int i = 0; i < n; i++) {
        if (q[i].graph < 0 || q[i].graph >= ng || q[i].kind < IPR_Q_SSSP || q[i].kind > IPR_Q_CC ||
            q[i].param >= graphs[q[i].graph]->V || !(q[i].t_arrival >= 0.0)) {
            set_error("ipr_serve: query %d is malformed", i);
            return -2;
        }
    }
    ipr_serve_opts o;
    memset(&o, 0, sizeof o);
    if (opt) o = *opt;
    if (o.ppr_iters <= 0) o.ppr_iters = 20;
    if (o.use_graph <= 0) o.use_graph = 1;
    ipr_serve_stats local;
    memset(&local, 0, sizeof local);
    const hipError_t e = hipSetDevice(graphs[0]->device);
    if (e != hipSuccess) return hip_fail(e, "set device");
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [q](int a, int b) { return q[a].t_arrival < q[b].t_arrival; });
    Ctx c{q, t_admit, t_done, digest, results, 0.0, o.ppr_iters, o.use_graph};
    c.delayed.assign((size_t)n, 0);
    int rc;
    if (o.sequential) {
        SeqStates ss;
        if ((rc = ss.prepare(graphs, ng, q, n, o.use_graph)) == 0) {
            c.t0 = now_s();
            rc = serve_sequential(graphs, ss, order, c);
        }
    } else {
        Server s(graphs, ng, c, o.hbm_budget, &local);
        if ((rc = s.prewarm(order)) == 0) {
            c.t0 = now_s();
            rc = s.run(order, (uint64_t)n);
        }
    }
    if (!rc) {
        double first = 1e300, last = 0.0;
        for (int i = 0; i < n; i++) {
            first = std::min(first, q[i].t_arrival);
            last = std::max(last, t_done[i]);
        }
        local.wall_s = n ? last - first : 0.0;
    }
    if (st) *st = local;
    return rc;
}
