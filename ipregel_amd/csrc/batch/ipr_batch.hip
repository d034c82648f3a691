/*
 * ipr_batch.hip - resident graphs, and the scalar and batched query kernels of
 * libipregel_batch.so (gfx950). No counterpart in the reference, which runs one
 * application per process; the single-query semantics are those of its SSSP
 * and CC benchmarks (benchmarks/sssp.c:38-79, benchmarks/cc.c:31-63) with the
 * min combiner (combiner_spread_postamble.h), and of personalised PageRank as
 * stated in include/ipregel_batch.h.
 *
 * Push (SSSP, CC). A superstep relaxes the out-arcs of the frontier into
 * min-combined mailboxes, marks the targets in a byte map, then an apply pass
 * over the byte map folds mailboxes into values and builds the next frontier
 * (vertices with out-degree above `large` go to a second list whose arcs are
 * relaxed in fixed-size chunks spread over all workgroups). The batched layout
 * gives each of 64 queries one wavefront lane: dist[v][64] and mbox[v][64] are
 * 256-byte rows, a frontier vertex carries the mask of queries active at it,
 * and the arcs of a vertex are read once for all of them. A query enters by a
 * distance-0 message at its source, so admission needs no frontier surgery.
 *
 * Pull (PPR). y_{t+1}(v) = 0.85 (sum_{u in in(v)} y_t(u)/deg(u) / S_t + [v=s] D_t/S_t)
 * + 0.15 [v=s], with S_t = sum y_t and D_t the mass of y_t on vertices without
 * out-arcs; the result after T iterations is x = y_T / S_T. The in-sum is an
 * FP32 Kahan sum in in-list order (lists longer than HUB in HUB-arc chunks,
 * then the chunk values in order); S and D are FP64 sums in a fixed order
 * (per 64-vertex segment in id order, 256 interleaved chains of segments, a
 * fixed tree over the chains, plus the hubs in order). The scalar and the
 * batched kernels perform the same operations in the same order for a query,
 * so their results are bit-identical, whatever else shares the batch.
 */
#include <cstring>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include <algorithm>
#include <vector>

#include "ipregel_hip.h"
#include "ipr_batch_int.h"

using namespace iprb;

namespace iprb {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char* what)
{
    set_error("HIP failure in %s: %s", what, hipGetErrorString(e));
    return -1;
}

double now_s()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

Tuning default_tuning() { return Tuning{1024u, 1024u, 2048u, INF, INF, 1u, 1u}; }

/* IPR_BATCH_GRAPHS=0 makes every state launch its kernels directly instead of
 * replaying captured HIP graphs (same kernels, same order) */
static int graph_mode(int requested)
{
    const char* e = getenv("IPR_BATCH_GRAPHS");
    if (e && e[0] == '0') return 2;
    return requested;
}

unsigned long long digest_u32_host(const uint32_t* v, uint64_t n)
{
    unsigned long long s = 0;
    for (uint64_t i = 0; i < n; i++) s += digest_mix((i << 32) | v[i]);
    return s;
}

unsigned long long digest_f32_host(const float* v, uint64_t n)
{
    unsigned long long s = 0;
    for (uint64_t i = 0; i < n; i++) {
        uint32_t b;
        memcpy(&b, &v[i], 4);
        s += digest_mix((i << 32) | b);
    }
    return s;
}

}  // namespace iprb

extern "C" const char* ipr_batch_last_error(void) { return iprb::g_err; }

#define HCK(x, what)                                   \
    do {                                               \
        hipError_t e_ = (x);                           \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

/* ------------------------------------------------------------------------ */
/* device helpers                                                            */
/* ------------------------------------------------------------------------ */

__device__ __forceinline__ uint64_t dmix(uint64_t z)
{
    z ^= z >> 30; z *= 0xbf58476d1ce4e5b9ULL;
    z ^= z >> 27; z *= 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return z;
}
__device__ __forceinline__ unsigned long long dterm(uint64_t v, uint32_t bits)
{
    return (unsigned long long)dmix((v << 32) | bits);
}

__device__ __forceinline__ uint32_t rl(uint32_t x, int j)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)x, j);
}
__device__ __forceinline__ float rlf(float x, int j)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
}
__device__ __forceinline__ uint64_t rl64(uint64_t x, int j)
{
    return ((uint64_t)rl((uint32_t)(x >> 32), j) << 32) | rl((uint32_t)x, j);
}
__device__ __forceinline__ double rld(double x, int j)
{
    return __longlong_as_double((long long)rl64((uint64_t)__double_as_longlong(x), j));
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x)
{
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (l >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)x, o, 64), hi = __shfl_xor((uint32_t)(x >> 32), o, 64);
        x += ((unsigned long long)hi << 32) | lo;
    }
    return x;
}

/* exclusive scan over the NT threads of a workgroup; s_w holds NT/64+1 words */
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t x, uint32_t* s_w, uint32_t& total)
{
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    uint32_t inc = wave_incl_scan_u32(x);
    if (l == 63) s_w[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < NT / 64; ++i) {
            uint32_t t = s_w[i];
            s_w[i] = acc;
            acc += t;
        }
        s_w[NT / 64] = acc;
    }
    __syncthreads();
    uint32_t r = s_w[w] + inc - x;
    total = s_w[NT / 64];
    __syncthreads();
    return r;
}

/* last index o in [0, n) with p[o] <= k (p nondecreasing, p[0] <= k) */
template <typename T>
__device__ __forceinline__ uint32_t owner_of(const T* p, uint32_t n, uint64_t k)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)p[mid] <= k) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t sadd(uint32_t d, uint32_t w)
{
    uint32_t r = d + w;
    return (r < d || r == INF) ? INF - 1u : r;
}

/* ------------------------------------------------------------------------ */
/* push kernels                                                              */
/* ------------------------------------------------------------------------ */

struct PushArgs {
    uint64_t V;
    const uint64_t* oidx;
    const uint32_t* oadj;
    const uint32_t* owgt;
    uint32_t* val;
    uint32_t* mbox;
    unsigned long long* act;
    unsigned long long* vmask;
    uint8_t* touched;
    uint8_t* tgran;     /* per 256 vertices: any touched byte set */
    uint32_t* small0;
    uint32_t* small1;
    uint32_t* large0;
    uint32_t* large1;
    uint32_t* lpre;
    PushCtl* ctl;
    void* pend;         /* parked frontier entries (byte per vertex, or slot mask when batched) */
    uint32_t* fst;      /* frontier stamps */
    uint32_t large, chunk;
    uint32_t wzero;     /* 1: every arc weighs 0 (CC) */
    uint32_t delta;     /* near/far step (INF: plain BSP) */
};

/* a vertex receives a message: its byte in the touched map, and the byte of
 * its 256-vertex granule, which lets the apply pass skip untouched granules
 * without reading their bytes */
__device__ __forceinline__ void mark(const PushArgs& a, uint32_t v)
{
    a.touched[v] = 1;
    a.tgran[v >> 8] = 1;
}

__device__ __forceinline__ uint32_t* small_of(const PushArgs& a, int p) { return p ? a.small1 : a.small0; }
__device__ __forceinline__ uint32_t* large_of(const PushArgs& a, int p) { return p ? a.large1 : a.large0; }
__device__ __forceinline__ uint32_t arc_w(const PushArgs& a, uint64_t e)
{
    return a.wzero ? 0u : (a.owgt ? a.owgt[e] : 1u);
}

__global__ void k_push_begin(PushArgs a, int cur)
{
    const int t = threadIdx.x, nxt = cur ^ 1;
    if (t == 0) {
        a.ctl->nsmall[nxt] = 0;
        a.ctl->nlarge[nxt] = 0;
    }
    if (t < 64) {
        a.ctl->orm[t].v = 0;
        a.ctl->relax[t].v = 0;
        for (int k = 0; k < SPREAD; ++k) {
            a.ctl->fmin[k][t].v = INF;
            a.ctl->pdmin[k][t].v = INF;
        }
    }
    if (t < SPREAD) a.ctl->pdef[t].v = 0;
}

/* a new state: no parked entries, no thresholds */
__global__ void k_ctl_init(PushArgs a)
{
    const int t = threadIdx.x;   /* 64 threads */
    a.ctl->thr[t] = INF;
    a.ctl->pmin[t] = INF;
    for (int k = 0; k < SPREAD; ++k) {
        a.ctl->pdmin[k][t].v = INF;
        a.ctl->pscan[k][t].v = INF;
    }
}

/* chunk prefix of the large list (one workgroup of 1024) */
__global__ __launch_bounds__(1024) void k_push_lscan(PushArgs a, int cur)
{
    __shared__ uint32_t s_w[17];
    __shared__ uint32_t s_carry;
    const uint32_t n = a.ctl->nlarge[cur];
    const uint32_t* L = large_of(a, cur);
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        uint32_t c = 0;
        if (i < n) {
            const uint32_t u = L[i];
            const uint64_t deg = a.oidx[u + 1] - a.oidx[u];
            c = (uint32_t)((deg + a.chunk - 1) / a.chunk);
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan_u32<1024>(c, s_w, tot);
        if (i < n) a.lpre[i] = s_carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) a.lpre[n] = s_carry;
}

/* scalar: tiles of 256 frontier vertices, one arc per thread per round */
__global__ __launch_bounds__(256) void k_relax_small_s(PushArgs a, int cur)
{
    __shared__ uint32_t s_pre[257];
    __shared__ uint64_t s_e0[256];
    __shared__ uint32_t s_val[256];
    __shared__ uint32_t s_w[5];
    const uint32_t n = a.ctl->nsmall[cur];
    const uint32_t thr = a.ctl->thr[0];
    uint8_t* pend = (uint8_t*)a.pend;
    const uint32_t* F = small_of(a, cur);
    const uint32_t ntiles = (n + 255) / 256;
    const int t = threadIdx.x;
    bool parked = false;
    uint32_t pmn = INF;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint32_t i = tile * 256 + t;
        uint32_t deg = 0;
        if (i < n) {
            const uint32_t u = F[i];
            const uint32_t du = a.val[u];
            if (du < thr) {
                const uint64_t e0 = a.oidx[u];
                deg = (uint32_t)(a.oidx[u + 1] - e0);
                s_e0[t] = e0;
                s_val[t] = du;
            } else {
                /* far: parked until the threshold passes it */
                pend[u] = 1;
                parked = true;
                pmn = min(pmn, du);
            }
        }
        uint32_t A;
        const uint32_t ex = block_excl_scan_u32<256>(deg, s_w, A);
        s_pre[t] = ex;
        __syncthreads();
        for (uint32_t r = 0; r < A; r += 256) {
            const uint32_t k = r + t;
            if (k < A) {
                const uint32_t o = owner_of(s_pre, 256, k);
                const uint64_t e = s_e0[o] + (k - s_pre[o]);
                const uint32_t v = a.oadj[e];
                const uint32_t c = sadd(s_val[o], arc_w(a, e));
                if (c < a.mbox[v]) {
                    atomicMin(&a.mbox[v], c);
                    mark(a, v);
                }
            }
        }
        __syncthreads();
    }
    if (__syncthreads_or(parked)) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) pmn = min(pmn, (uint32_t)__shfl_xor(pmn, o, 64));
        if ((t & 63) == 0 && pmn != INF) {
            atomicOr(&a.ctl->pdef[blockIdx.x % SPREAD].v, 1ull);
            atomicMin(&a.ctl->pdmin[blockIdx.x % SPREAD][0].v, pmn);
        }
    }
}

/* position of the r-th (from 0) set bit of m; r < popcount(m) */
__host__ __device__ __forceinline__ uint32_t nth_bit(unsigned long long m, uint32_t r)
{
    uint32_t pos = 0, c;
    c = (uint32_t)__builtin_popcountll(m & 0xFFFFFFFFull);
    if (r >= c) { r -= c; m >>= 32; pos += 32; }
    c = (uint32_t)__builtin_popcountll(m & 0xFFFFull);
    if (r >= c) { r -= c; m >>= 16; pos += 16; }
    c = (uint32_t)__builtin_popcountll(m & 0xFFull);
    if (r >= c) { r -= c; m >>= 8; pos += 8; }
    c = (uint32_t)__builtin_popcountll(m & 0xFull);
    if (r >= c) { r -= c; m >>= 4; pos += 4; }
    c = (uint32_t)__builtin_popcountll(m & 0x3ull);
    if (r >= c) { r -= c; m >>= 2; pos += 2; }
    if (r >= (uint32_t)(m & 1ull)) pos += 1;
    return pos;
}

/* a frontier vertex with at most PAIR_SLOTS active slots and at most PAIR_DEG
 * out-arcs is relaxed as (vertex, slot) pairs, one pair per lane; the others
 * as rows, one arc per wavefront step for all slots at once */
constexpr uint32_t PAIR_SLOTS = 8, PAIR_DEG = 32;

/* batched: tiles of 64 frontier vertices */
__global__ __launch_bounds__(256) void k_relax_small_b(PushArgs a, int cur)
{
    __shared__ uint32_t s_pre[65];     /* arcs of row-mode vertices, exclusive prefix */
    __shared__ uint32_t s_pp[65];      /* pairs of pair-mode vertices, exclusive prefix */
    __shared__ uint64_t s_e0[64];
    __shared__ uint32_t s_deg[64];
    __shared__ uint32_t s_u[64];
    __shared__ unsigned long long s_m[64];
    __shared__ uint32_t s_d[64 * 64];
    const uint32_t n = a.ctl->nsmall[cur];
    const uint32_t* F = small_of(a, cur);
    const uint32_t ntiles = (n + 63) / 64;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t thrl = a.ctl->thr[lane];
    unsigned long long* pend = (unsigned long long*)a.pend;
    unsigned long long parked = 0;
    uint32_t pmn = INF;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        if (t < 64) {
            const uint32_t i = tile * 64 + t;
            uint32_t deg = 0, u = 0;
            unsigned long long m = 0;
            if (i < n) {
                u = F[i];
                const uint64_t e0 = a.oidx[u];
                deg = (uint32_t)(a.oidx[u + 1] - e0);
                m = a.act[u];
                s_e0[t] = e0;
            }
            s_u[t] = u;
            s_m[t] = m;
            s_deg[t] = deg;
        }
        __syncthreads();
        const uint32_t nt = min(64u, n - tile * 64);
        for (uint32_t r = wave; r < nt; r += 4) {
            /* lanes whose distance is below their threshold relax now; the
             * others stay pending at the vertex */
            const uint32_t d = a.val[(uint64_t)s_u[r] * 64 + lane];
            s_d[r * 64 + lane] = d;
            const bool on = (s_m[r] >> lane) & 1ull;
            const bool near = on && d < thrl;
            if (on && !near) pmn = min(pmn, d);
            const unsigned long long nm = __ballot(near), dm = __ballot(on && !near);
            if (lane == 0) {
                s_m[r] = nm;
                if (dm) {
                    /* far slots: parked until their threshold passes them */
                    pend[s_u[r]] |= dm;
                    parked |= dm;
                }
            }
        }
        __syncthreads();
        if (t < 64) {
            const unsigned long long nm = s_m[t];
            const uint32_t deg = s_deg[t], k = (uint32_t)__popcll(nm);
            const bool pair = k && k <= PAIR_SLOTS && deg <= PAIR_DEG;
            const uint32_t rowdeg = (k && !pair) ? deg : 0u, pairs = pair ? k : 0u;
            const uint32_t ia = wave_incl_scan_u32(rowdeg), ib = wave_incl_scan_u32(pairs);
            s_pre[t] = ia - rowdeg;
            s_pp[t] = ib - pairs;
            if (t == 63) {
                s_pre[64] = ia;
                s_pp[64] = ib;
            }
        }
        __syncthreads();
        /* rows: 256 arcs per round, each wavefront relaxes the 64 it loaded */
        const uint32_t A = s_pre[64];
        for (uint32_t base = 0; base < A; base += 256) {
            const uint32_t k = base + t;
            uint32_t v = 0, w = 0, o = 0;
            unsigned long long vmk = 0;
            if (k < A) {
                o = owner_of(s_pre, 64, k);
                const uint64_t e = s_e0[o] + (k - s_pre[o]);
                v = a.oadj[e];
                w = arc_w(a, e);
                vmk = a.vmask[v];
            }
            const uint32_t wb = base + (uint32_t)wave * 64;
            const uint32_t cnt = A > wb ? min(64u, A - wb) : 0u;
            for (uint32_t j0 = 0; j0 < cnt; j0 += 8) {
                /* the targets' current distances, 8 rows in flight; a message
                 * that does not beat a set distance is dropped */
                uint32_t dv[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const uint32_t j = (j0 + i) & 63;
                    const unsigned long long vm = rl64(vmk, j);
                    const bool need = j0 + i < cnt && ((vm >> lane) & 1ull) && ((s_m[rl(o, j)] >> lane) & 1ull);
                    dv[i] = need ? a.val[(uint64_t)rl(v, j) * 64 + lane] : INF;
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const uint32_t j = (j0 + i) & 63;
                    if (j0 + i >= cnt) break;
                    const uint32_t vj = rl(v, j), oj = rl(o, j);
                    const uint32_t cd = sadd(s_d[oj * 64 + lane], rl(w, j));
                    const bool go = ((s_m[oj] >> lane) & 1ull) && cd < dv[i];
                    if (go) atomicMin(&a.mbox[(uint64_t)vj * 64 + lane], cd);
                    if (__ballot(go) && lane == 0) mark(a, vj);
                }
            }
        }
        /* pairs: lane = (vertex, slot), walking the vertex's short out-list */
        const uint32_t P = s_pp[64];
        for (uint32_t pb = (uint32_t)wave * 64; pb < P; pb += 256) {
            const uint32_t p = pb + lane;
            if (p < P) {
                const uint32_t o = owner_of(s_pp, 64, p);
                const uint32_t q = nth_bit(s_m[o], p - s_pp[o]);
                const uint32_t du = s_d[o * 64 + q];
                const uint64_t e0 = s_e0[o];
                const uint32_t deg = s_deg[o];
                for (uint32_t k0 = 0; k0 < deg; k0 += 4) {
                    /* four arcs: targets, weights, then their distances, then the messages */
                    uint32_t vv[4], cd[4], dv[4];
                    unsigned long long vm[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const bool ok = k0 + i < deg;
                        vv[i] = ok ? a.oadj[e0 + k0 + i] : 0u;
                        cd[i] = ok ? sadd(du, arc_w(a, e0 + k0 + i)) : INF;
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        vm[i] = k0 + i < deg ? a.vmask[vv[i]] : 0ull;
                        dv[i] = k0 + i < deg ? a.val[(uint64_t)vv[i] * 64 + q] : INF;
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const uint32_t cur = ((vm[i] >> q) & 1ull) ? dv[i] : INF;
                        if (k0 + i < deg && cd[i] < cur) {
                            atomicMin(&a.mbox[(uint64_t)vv[i] * 64 + q], cd[i]);
                            mark(a, vv[i]);
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    if (lane == 0 && parked) atomicOr(&a.ctl->pdef[blockIdx.x % SPREAD].v, parked);
    if (pmn != INF) atomicMin(&a.ctl->pdmin[blockIdx.x % SPREAD][lane].v, pmn);
}

/* large frontier vertices: chunk c of the large list goes to workgroup c mod grid */
template <bool BATCHED>
__global__ __launch_bounds__(256) void k_relax_large(PushArgs a, int cur)
{
    __shared__ uint32_t s_lp[LDS_LARGE + 1];
    const uint32_t n = a.ctl->nlarge[cur];
    if (n == 0) return;
    const uint32_t T = a.lpre[n];
    const uint32_t* L = large_of(a, cur);
    const bool inl = n + 1 <= LDS_LARGE;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (inl)
        for (uint32_t i = t; i <= n; i += 256) s_lp[i] = a.lpre[i];
    __syncthreads();
    const uint32_t* lp = inl ? s_lp : a.lpre;
    const uint32_t thrl = a.ctl->thr[BATCHED ? lane : 0];
    for (uint32_t c = blockIdx.x; c < T; c += gridDim.x) {
        const uint32_t i = owner_of(lp, n, c);
        const uint32_t u = L[i];
        const uint64_t e0 = a.oidx[u] + (uint64_t)(c - lp[i]) * a.chunk;
        const uint64_t e1 = min(e0 + a.chunk, a.oidx[u + 1]);
        if (!BATCHED) {
            const uint32_t du = a.val[u];
            if (du >= thrl) {
                if (c == lp[i] && t == 0) {
                    ((uint8_t*)a.pend)[u] = 1;
                    atomicOr(&a.ctl->pdef[blockIdx.x % SPREAD].v, 1ull);
                    atomicMin(&a.ctl->pdmin[blockIdx.x % SPREAD][0].v, du);
                }
                continue;
            }
            for (uint64_t e = e0 + t; e < e1; e += 256) {
                const uint32_t v = a.oadj[e];
                const uint32_t c = sadd(du, arc_w(a, e));
                if (c < a.mbox[v]) {
                    atomicMin(&a.mbox[v], c);
                    mark(a, v);
                }
            }
        } else {
            const unsigned long long m = a.act[u];
            const uint32_t du = a.val[(uint64_t)u * 64 + lane];
            const bool act_l = (m >> lane) & 1ull;
            const bool on = act_l && du < thrl;
            const unsigned long long dm = __ballot(act_l && !on);
            if (dm && c == lp[i] && t == 0) {
                ((unsigned long long*)a.pend)[u] |= dm;
                atomicOr(&a.ctl->pdef[blockIdx.x % SPREAD].v, dm);
            }
            if (act_l && !on && c == lp[i] && wave == 0) atomicMin(&a.ctl->pdmin[blockIdx.x % SPREAD][lane].v, du);
            if (!__ballot(on)) continue;
            for (uint64_t base = e0; base < e1; base += 256) {
                const uint64_t e = base + t;
                uint32_t v = 0, w = 0;
                unsigned long long vmk = 0;
                if (e < e1) {
                    v = a.oadj[e];
                    w = arc_w(a, e);
                    vmk = a.vmask[v];
                }
                const uint64_t wb = base + (uint64_t)wave * 64;
                const uint32_t cnt = e1 > wb ? (uint32_t)min<uint64_t>(64, e1 - wb) : 0u;
                for (uint32_t j0 = 0; j0 < cnt; j0 += 8) {
                    uint32_t dv[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const uint32_t j = (j0 + i) & 63;
                        const unsigned long long vm = rl64(vmk, j);
                        dv[i] = (on && j0 + i < cnt && ((vm >> lane) & 1ull)) ? a.val[(uint64_t)rl(v, j) * 64 + lane]
                                                                               : INF;
                    }
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const uint32_t j = (j0 + i) & 63;
                        if (j0 + i >= cnt) break;
                        const uint32_t vj = rl(v, j);
                        const uint32_t cd = sadd(du, rl(w, j));
                        const bool go = on && cd < dv[i];
                        if (go) atomicMin(&a.mbox[(uint64_t)vj * 64 + lane], cd);
                        if (__ballot(go) && lane == 0) mark(a, vj);
                    }
                }
            }
        }
    }
}

/* next-frontier entries a workgroup of an apply kernel buffers in LDS before
 * one reservation on the list counter (entries beyond the buffer are reserved
 * one at a time) */
#define APPLY_BUF_S 1024
#define APPLY_BUF_L 256

/* flush of an apply workgroup's buffered next-frontier entries */
__device__ inline void apply_flush(PushArgs& a, int nxt, uint32_t* S, uint32_t* Lg, const uint32_t* buf_s,
                                   const uint32_t* buf_l, uint32_t cnt_s, uint32_t cnt_l, uint32_t* base)
{
    const uint32_t ns = cnt_s < APPLY_BUF_S ? cnt_s : APPLY_BUF_S;
    const uint32_t nl = cnt_l < APPLY_BUF_L ? cnt_l : APPLY_BUF_L;
    if (threadIdx.x == 0) {
        base[0] = ns ? atomicAdd(&a.ctl->nsmall[nxt], ns) : 0u;
        base[1] = nl ? atomicAdd(&a.ctl->nlarge[nxt], nl) : 0u;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) S[base[0] + i] = buf_s[i];
    for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) Lg[base[1] + i] = buf_l[i];
}

/* one-query apply: a workgroup takes 1024-vertex spans (four granule bytes),
 * a thread 4 vertices; improved vertices go to the LDS buffers */
__global__ __launch_bounds__(256) void k_apply_s(PushArgs a, int cur)
{
    __shared__ uint32_t buf_s[APPLY_BUF_S], buf_l[APPLY_BUF_L];
    __shared__ uint32_t cnt_s, cnt_l, base[2];
    const int nxt = cur ^ 1, t = threadIdx.x;
    uint32_t* S = small_of(a, nxt);
    uint32_t* Lg = large_of(a, nxt);
    uint8_t* pend = (uint8_t*)a.pend;
    const uint32_t step = a.ctl->step;
    unsigned long long rel = 0;
    uint32_t fmn = INF;
    if (t == 0) {
        cnt_s = 0;
        cnt_l = 0;
    }
    for (uint64_t gb = (uint64_t)blockIdx.x * 1024; gb < a.V; gb += (uint64_t)gridDim.x * 1024) {
        /* the four granule bytes of these 1024 vertices (padded map) */
        if (*(const uint32_t*)(a.tgran + (gb >> 8)) == 0) continue;
        __syncthreads();   /* every thread has read the bytes before they are cleared */
        if (t == 0) *(uint32_t*)(a.tgran + (gb >> 8)) = 0;
        const uint64_t v0 = gb + (uint64_t)t * 4;
        uint32_t f4 = 0;
        if (v0 < a.V) f4 = *(const uint32_t*)(a.touched + v0);   /* padded to 1024 */
        if (!f4) continue;
        *(uint32_t*)(a.touched + v0) = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!((f4 >> (8 * k)) & 0xff)) continue;
            const uint64_t v = v0 + k;
            const uint32_t m = a.mbox[v], d = a.val[v];
            if (m < d) {
                a.val[v] = m;
                if (pend[v]) pend[v] = 0;     /* its parked entry is superseded */
                a.fst[v] = step;
                fmn = min(fmn, m);
                const uint64_t deg = a.oidx[v + 1] - a.oidx[v];
                rel += deg;
                if (deg > a.large) {
                    const uint32_t q = atomicAdd(&cnt_l, 1u);
                    if (q < APPLY_BUF_L) buf_l[q] = (uint32_t)v;
                    else Lg[atomicAdd(&a.ctl->nlarge[nxt], 1u)] = (uint32_t)v;
                } else {
                    const uint32_t q = atomicAdd(&cnt_s, 1u);
                    if (q < APPLY_BUF_S) buf_s[q] = (uint32_t)v;
                    else S[atomicAdd(&a.ctl->nsmall[nxt], 1u)] = (uint32_t)v;
                }
            }
        }
    }
    __syncthreads();
    apply_flush(a, nxt, S, Lg, buf_s, buf_l, cnt_s, cnt_l, base);
    rel = wave_sum_u64(rel);
    if ((t & 63) == 0 && rel) atomicAdd(&a.ctl->relax[blockIdx.x & 63].v, rel);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) fmn = min(fmn, (uint32_t)__shfl_xor(fmn, o, 64));
    if ((t & 63) == 0 && fmn != INF) atomicMin(&a.ctl->fmin[blockIdx.x % SPREAD][0].v, fmn);
}

/* batched apply: a workgroup takes 256-vertex granules; each wavefront scans
 * its 64 byte flags and folds the marked rows 8 at a time, all loads issued
 * first */
__global__ __launch_bounds__(256) void k_apply_b(PushArgs a, int cur)
{
    __shared__ uint32_t buf_s[APPLY_BUF_S], buf_l[APPLY_BUF_L];
    __shared__ uint32_t cnt_s, cnt_l, base[2];
    __shared__ unsigned long long s_gm;
    const int nxt = cur ^ 1, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint32_t* S = small_of(a, nxt);
    uint32_t* Lg = large_of(a, nxt);
    unsigned long long* pend = (unsigned long long*)a.pend;
    const uint32_t step = a.ctl->step;
    unsigned long long orm = 0, rel = 0, dig = 0;
    uint32_t fmn = INF;
    if (t == 0) {
        cnt_s = 0;
        cnt_l = 0;
    }
    /* granules g = blockIdx.x + k * gridDim.x; their touched bytes are read
     * 64 at a time by one wavefront, so untouched granules cost no round trip.
     * Inside a granule each wavefront owns 64 vertices and works alone: no
     * workgroup barrier, no global atomic per granule. */
    const uint64_t ngran = (a.V + 255) >> 8;
    for (uint64_t g0 = blockIdx.x; g0 < ngran; g0 += (uint64_t)gridDim.x * 64) {
        __syncthreads();
        if (wave == 0) {
            const uint64_t g = g0 + (uint64_t)lane * gridDim.x;
            const unsigned long long m = __ballot(g < ngran && a.tgran[g] != 0);
            if (lane == 0) s_gm = m;
        }
        __syncthreads();
        unsigned long long gm = s_gm;
        while (gm) {
            const uint64_t gb = (g0 + (uint64_t)(__ffsll(gm) - 1) * gridDim.x) * 256;
            gm &= gm - 1;
            if (t == 0) a.tgran[gb >> 8] = 0;
            const uint64_t vb = gb + (uint64_t)wave * 64, vl = vb + lane;
            const uint8_t fl = vl < a.V ? a.touched[vl] : 0;
            unsigned long long tm = __ballot(fl != 0);
            if (!tm) continue;
            if (fl) a.touched[vl] = 0;
            while (tm) {
                uint32_t js[8];
                uint32_t nb = 0;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    js[i] = tm ? (uint32_t)(__ffsll((unsigned long long)tm) - 1) : 0u;
                    if (tm) {
                        tm &= tm - 1;
                        nb = i + 1;
                    }
                }
                uint32_t mm[8], dd[8];
                unsigned long long vmv[8], pmv[8];
                uint64_t dg[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if ((uint32_t)i < nb) {
                        const uint64_t v = vb + js[i], ix = v * 64 + lane;
                        mm[i] = a.mbox[ix];
                        dd[i] = a.val[ix];
                        vmv[i] = a.vmask[v];
                        pmv[i] = pend[v];
                        dg[i] = a.oidx[v + 1] - a.oidx[v];
                    }
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if ((uint32_t)i >= nb) break;
                    const uint64_t v = vb + js[i], ix = v * 64 + lane;
                    const uint32_t m = mm[i];
                    const unsigned long long vm = vmv[i];
                    const uint32_t d = ((vm >> lane) & 1ull) ? dd[i] : INF;
                    const bool imp = m < d;
                    if (imp) {
                        a.val[ix] = m;
                        dig += dterm(v, m) - dterm(v, d);
                    }
                    if (m != INF) a.mbox[ix] = INF;
                    /* next frontier at v: the improved slots (parked entries of
                     * those slots at v are superseded) */
                    const unsigned long long im = __ballot(imp), pm = pmv[i];
                    if (im) {
                        const uint64_t deg = dg[i];
                        if (lane == 0) {
                            a.vmask[v] = vm | im;
                            a.act[v] = im;
                            a.fst[v] = step;
                            if (pm & im) pend[v] = pm & ~im;
                            if (deg > a.large) {
                                const uint32_t q = atomicAdd(&cnt_l, 1u);
                                if (q < APPLY_BUF_L) buf_l[q] = (uint32_t)v;
                                else Lg[atomicAdd(&a.ctl->nlarge[nxt], 1u)] = (uint32_t)v;
                            } else {
                                const uint32_t q = atomicAdd(&cnt_s, 1u);
                                if (q < APPLY_BUF_S) buf_s[q] = (uint32_t)v;
                                else S[atomicAdd(&a.ctl->nsmall[nxt], 1u)] = (uint32_t)v;
                            }
                        }
                        if (imp) fmn = min(fmn, m);
                        orm |= im;
                        rel += deg * (unsigned long long)__popcll(im);
                    }
                }
            }
        }
    }
    __syncthreads();
    apply_flush(a, nxt, S, Lg, buf_s, buf_l, cnt_s, cnt_l, base);
    /* the four wavefronts' figures are combined in LDS first: one wavefront
     * issues the workgroup's atomics (orm and rel are wavefront-uniform;
     * dig and fmn are per query, lane q holding query q) */
    __shared__ unsigned long long s_dig[4][64], s_orm[4], s_rel[4];
    __shared__ uint32_t s_fmn[4][64];
    s_dig[wave][lane] = dig;
    s_fmn[wave][lane] = fmn;
    if (lane == 0) {
        s_orm[wave] = orm;
        s_rel[wave] = rel;
    }
    __syncthreads();
    if (wave == 0) {
        const unsigned long long dsum = s_dig[0][lane] + s_dig[1][lane] + s_dig[2][lane] + s_dig[3][lane];
        const uint32_t fm = min(min(s_fmn[0][lane], s_fmn[1][lane]), min(s_fmn[2][lane], s_fmn[3][lane]));
        const unsigned long long om = s_orm[0] | s_orm[1] | s_orm[2] | s_orm[3];
        const unsigned long long rs = s_rel[0] + s_rel[1] + s_rel[2] + s_rel[3];
        if (lane == 0 && om) atomicOr(&a.ctl->orm[blockIdx.x & 63].v, om);
        if (lane == 0 && rs) atomicAdd(&a.ctl->relax[blockIdx.x & 63].v, rs);
        if (dsum) atomicAdd(&a.ctl->dig[blockIdx.x % SPREAD][lane].v, dsum);
        if (fm != INF) atomicMin(&a.ctl->fmin[blockIdx.x % SPREAD][lane].v, fm);
    }
}

/* after the apply pass: per query, if no entry of the next frontier lies below
 * the threshold, move the threshold (past the smallest next-frontier distance,
 * or by delta when entries are parked) and ask for a wake-up scan */
__global__ void k_push_thr(PushArgs a)
{
    const int t = threadIdx.x;   /* 64 threads, one per slot */
    unsigned long long pd = 0;
    for (int k = 0; k < SPREAD; ++k) pd |= a.ctl->pdef[k].v;
    const unsigned long long pleft = a.ctl->pleft | pd;
    uint32_t fm = INF;
    for (int k = 0; k < SPREAD; ++k) fm = min(fm, a.ctl->fmin[k][t].v);
    const uint32_t th = a.ctl->thr[t];
    uint32_t pm = a.ctl->pmin[t];
    for (int k = 0; k < SPREAD; ++k) pm = min(pm, a.ctl->pdmin[k][t].v);
    a.ctl->pmin[t] = pm;
    const bool parked = (pleft >> t) & 1ull;
    bool wake = false;
    if (fm >= th) {
        /* past the smallest distance waiting, in the frontier or parked */
        const uint32_t tg = parked ? min(fm, pm) : fm;
        uint32_t nt = th;
        if (tg != INF) nt = a.delta >= INF - tg ? INF : tg + a.delta;
        else if (parked) nt = a.delta >= INF - th ? INF : th + a.delta;
        a.ctl->thr[t] = nt;
        wake = parked && (pm < nt || pm == INF);
    }
    const unsigned long long wm = __ballot(wake);
    if (t == 0) {
        a.ctl->pleft = pleft;
        a.ctl->react = wm;
    }
}

/* wake-up scan: parked entries of the queries in react that now lie below their
 * threshold join the next frontier (once per vertex: frontier stamps); the set
 * of queries with entries still parked is recomputed exactly */
template <bool BATCHED>
__global__ __launch_bounds__(256) void k_push_react(PushArgs a, int nxt)
{
    __shared__ uint32_t s_w[5];
    __shared__ uint32_t s_base[2];
    __shared__ uint32_t s_pmin[64];
    const unsigned long long react = a.ctl->react;
    if (!react) return;
    if (threadIdx.x < 64) s_pmin[threadIdx.x] = INF;
    __syncthreads();
    const uint32_t step = a.ctl->step;
    const int t = threadIdx.x;
    uint32_t* S = small_of(a, nxt);
    uint32_t* Lg = large_of(a, nxt);
    unsigned long long rem = 0, orm = 0, rel = 0;
    for (uint64_t gb = (uint64_t)blockIdx.x * 256; gb < a.V; gb += (uint64_t)gridDim.x * 256) {
        const uint64_t v = gb + t;
        unsigned long long nb = 0;
        uint32_t isl = 0, iss = 0;
        if (v < a.V) {
            const unsigned long long pm =
                BATCHED ? ((unsigned long long*)a.pend)[v] : (unsigned long long)((uint8_t*)a.pend)[v];
            unsigned long long bits = pm & react;
            while (bits) {
                const int q = __ffsll(bits) - 1;
                bits &= bits - 1;
                const uint32_t d = BATCHED ? a.val[v * 64 + q] : a.val[v];
                if (d < a.ctl->thr[q]) nb |= 1ull << q;
                else atomicMin(&s_pmin[q], d);
            }
            rem |= pm & ~nb;
            if (nb) {
                if (BATCHED) {
                    ((unsigned long long*)a.pend)[v] = pm & ~nb;
                    if (a.fst[v] == step) {
                        a.act[v] |= nb;
                    } else {
                        a.act[v] = nb;
                        a.fst[v] = step;
                        isl = 1;
                    }
                } else {
                    ((uint8_t*)a.pend)[v] = 0;
                    if (a.fst[v] != step) {
                        a.fst[v] = step;
                        isl = 1;
                    }
                }
                const uint64_t deg = a.oidx[v + 1] - a.oidx[v];
                orm |= nb;
                rel += deg * (unsigned long long)__popcll(nb);
                if (isl && deg <= a.large) {
                    iss = 1;
                    isl = 0;
                }
            }
        }
        if (!__syncthreads_or(iss | isl)) continue;
        uint32_t ts, tl;
        const uint32_t xs = block_excl_scan_u32<256>(iss, s_w, ts);
        const uint32_t xl = block_excl_scan_u32<256>(isl, s_w, tl);
        if (t == 0) {
            s_base[0] = ts ? atomicAdd(&a.ctl->nsmall[nxt], ts) : 0;
            s_base[1] = tl ? atomicAdd(&a.ctl->nlarge[nxt], tl) : 0;
        }
        __syncthreads();
        if (iss) S[s_base[0] + xs] = (uint32_t)v;
        if (isl) Lg[s_base[1] + xl] = (uint32_t)v;
        __syncthreads();
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)rem, o, 64), hi = __shfl_xor((uint32_t)(rem >> 32), o, 64);
        rem |= ((unsigned long long)hi << 32) | lo;
        lo = __shfl_xor((uint32_t)orm, o, 64);
        hi = __shfl_xor((uint32_t)(orm >> 32), o, 64);
        orm |= ((unsigned long long)hi << 32) | lo;
    }
    rel = wave_sum_u64(rel);
    __syncthreads();
    if (t < 64 && s_pmin[t] != INF) atomicMin(&a.ctl->pscan[blockIdx.x % SPREAD][t].v, s_pmin[t]);
    if ((t & 63) == 0) {
        if (rem) atomicOr(&a.ctl->prem[blockIdx.x % SPREAD].v, rem);
        if (orm) atomicOr(&a.ctl->orm[blockIdx.x & 63].v, orm);
        if (rel) atomicAdd(&a.ctl->relax[blockIdx.x & 63].v, rel);
    }
}

/* end of the superstep: what the host reads back */
__global__ void k_push_out(PushArgs a, int nxt, int batched)
{
    const int t = threadIdx.x;   /* 64 threads */
    unsigned long long o = a.ctl->orm[t].v, r = a.ctl->relax[t].v, d = 0;
    for (int k = 0; k < SPREAD; ++k) d += a.ctl->dig[k][t].v;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)o, s, 64), hi = __shfl_xor((uint32_t)(o >> 32), s, 64);
        o |= ((unsigned long long)hi << 32) | lo;
    }
    r = wave_sum_u64(r);
    a.ctl->out.dig[t] = d;
    const unsigned long long scanned = a.ctl->react;
    if ((scanned >> t) & 1ull) {
        uint32_t pm = INF;
        for (int k = 0; k < SPREAD; ++k) {
            pm = min(pm, a.ctl->pscan[k][t].v);
            a.ctl->pscan[k][t].v = INF;
        }
        a.ctl->pmin[t] = pm;
    }
    __syncthreads();
    if (t == 0) {
        unsigned long long pleft = a.ctl->pleft;
        if (a.ctl->react) {
            pleft = 0;
            for (int k = 0; k < SPREAD; ++k) {
                pleft |= a.ctl->prem[k].v;
                a.ctl->prem[k].v = 0;
            }
            a.ctl->pleft = pleft;
            a.ctl->react = 0;
        }
        const uint32_t ns = a.ctl->nsmall[nxt], nl = a.ctl->nlarge[nxt];
        a.ctl->relax_total += r;
        /* running queries: in the next frontier, or with parked entries */
        a.ctl->out.orm = (batched ? o : (ns + nl ? 1ull : 0ull)) | pleft;
        if (a.ctl->out.orm) a.ctl->busy_total += 1;
        a.ctl->out.busy = a.ctl->busy_total;
        a.ctl->out.relax = a.ctl->relax_total;
        a.ctl->out.nsmall = ns;
        a.ctl->out.nlarge = nl;
        a.ctl->step += 1;
    }
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t x)
{
    const uint64_t st = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) p[i] = x;
}

/* scalar SSSP: the source receives distance 0 in the first superstep */
__global__ void k_start_sssp_s(PushArgs a, int cur, uint32_t src)
{
    a.ctl->thr[0] = a.delta;
    a.ctl->nsmall[cur] = 0;
    a.ctl->nlarge[cur] = 0;
    a.mbox[src] = 0;
    mark(a, src);
}

/* scalar CC: labels = ids, every vertex in the first frontier */
__global__ __launch_bounds__(256) void k_start_cc_s(PushArgs a, int cur)
{
    __shared__ uint32_t s_w[5];
    __shared__ uint32_t s_base[2];
    uint32_t* S = small_of(a, cur);
    uint32_t* Lg = large_of(a, cur);
    const int t = threadIdx.x;
    for (uint64_t gb = (uint64_t)blockIdx.x * 256; gb < a.V; gb += (uint64_t)gridDim.x * 256) {
        const uint64_t v = gb + t;
        uint32_t isl = 0, iss = 0;
        if (v < a.V) {
            a.val[v] = (uint32_t)v;
            a.mbox[v] = (uint32_t)v;
            const uint64_t deg = a.oidx[v + 1] - a.oidx[v];
            if (deg > a.large) isl = 1; else iss = 1;
        }
        uint32_t ts, tl;
        const uint32_t xs = block_excl_scan_u32<256>(iss, s_w, ts);
        const uint32_t xl = block_excl_scan_u32<256>(isl, s_w, tl);
        if (t == 0) {
            s_base[0] = ts ? atomicAdd(&a.ctl->nsmall[cur], ts) : 0;
            s_base[1] = tl ? atomicAdd(&a.ctl->nlarge[cur], tl) : 0;
        }
        __syncthreads();
        if (iss) S[s_base[0] + xs] = (uint32_t)v;
        if (isl) Lg[s_base[1] + xl] = (uint32_t)v;
        __syncthreads();
    }
}

__global__ void k_ctl_zero(PushArgs a, int cur)
{
    a.ctl->relax_total = 0;
    a.ctl->busy_total = 0;
    a.ctl->thr[0] = INF;
    a.ctl->pleft = 0;
    a.ctl->react = 0;
    a.ctl->pmin[0] = INF;
    a.ctl->nsmall[cur] = 0;
    a.ctl->nlarge[cur] = 0;
    a.ctl->nsmall[cur ^ 1] = 0;
    a.ctl->nlarge[cur ^ 1] = 0;
}

struct AdmitArgs {
    int n;
    int slot[64];
    uint32_t src[64];
};

/* batched admission: a distance-0 message at the source of each new slot */
__global__ void k_admit_b(PushArgs a, AdmitArgs ad, unsigned long long base)
{
    const int i = threadIdx.x;
    if (i >= ad.n) return;
    const int q = ad.slot[i];
    const uint32_t s = ad.src[i];
    a.mbox[(uint64_t)s * 64 + q] = 0;
    mark(a, s);
    a.ctl->thr[q] = a.delta;
    a.ctl->pmin[q] = INF;
    a.ctl->dig[0][q].v = base;
    for (int k = 1; k < SPREAD; ++k) a.ctl->dig[k][q].v = 0;
}

__global__ void k_release_b(PushArgs a, unsigned long long mask)
{
    const uint64_t st = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < a.V; v += st) {
        const unsigned long long m = a.vmask[v];
        if (m & mask) a.vmask[v] = m & ~mask;
    }
}

/* batched column(s) to query-major output: wavefront per vertex, lane per slot */
__global__ void k_columns_b(PushArgs a, int n, int first, uint32_t* out)
{
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t v = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; v < a.V; v += nw) {
        const unsigned long long vm = a.vmask[v];
        const int q = first + lane;
        if (lane < n) out[(uint64_t)lane * a.V + v] = ((vm >> q) & 1ull) ? a.val[v * 64 + q] : INF;
    }
}

__global__ void k_digest_u32(const uint32_t* x, uint64_t n, unsigned long long* out)
{
    const uint64_t st = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st)
        s += dterm(i, x ? x[i] : INF);
    s = wave_sum_u64(s);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

/* ------------------------------------------------------------------------ */
/* CC by union-find                                                          */
/* ------------------------------------------------------------------------ */
/* parent[v] <= v always; a root is linked only under a smaller root, so each
 * tree's root is the smallest id of its component: the labels equal the
 * label-propagation fixed point (component minimum) on a symmetric graph,
 * whatever the order of the links. Path halving writes only ancestors, so
 * concurrent finds and links stay consistent. */

__device__ __forceinline__ uint32_t uf_root(uint32_t* p, uint32_t x)
{
    volatile uint32_t* vp = p;
    for (;;) {
        const uint32_t px = vp[x];
        if (px == x) return x;
        const uint32_t ppx = vp[px];
        if (ppx == px) return px;
        vp[x] = ppx;
        x = ppx;
    }
}

__device__ __forceinline__ void uf_link(uint32_t* p, uint32_t u, uint32_t v)
{
    uint32_t ru = uf_root(p, u), rv = uf_root(p, v);
    while (ru != rv) {
        const uint32_t hi = max(ru, rv), lo = min(ru, rv);
        const uint32_t old = atomicCAS(&p[hi], hi, lo);
        if (old == hi) return;
        /* hi stopped being a root meanwhile: retry from the current roots */
        ru = uf_root(p, old);
        rv = uf_root(p, lo);
    }
}

__global__ void k_uf_init(uint32_t* p, uint64_t V)
{
    const uint64_t st = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += st) p[v] = (uint32_t)v;
}

/* wavefront per vertex, lanes over its arcs */
__global__ __launch_bounds__(256) void k_uf_link(const uint64_t* idx, const uint32_t* adj, uint32_t* p, uint64_t V)
{
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t u = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; u < V; u += nw) {
        const uint64_t b = idx[u], e = idx[u + 1];
        for (uint64_t k = b + lane; k < e; k += 64) {
            const uint32_t v = adj[k];
            if (v != (uint32_t)u) uf_link(p, (uint32_t)u, v);
        }
    }
}

__global__ void k_uf_flatten(uint32_t* p, uint64_t V)
{
    const uint64_t st = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += st) p[v] = uf_root(p, (uint32_t)v);
}

/* ------------------------------------------------------------------------ */
/* PPR kernels                                                               */
/* ------------------------------------------------------------------------ */

struct PprArgs {
    uint64_t V, nseg, nseg64;
    const uint32_t* seg_v0;
    uint32_t nhub, nchunk;
    const uint64_t* oidx;
    const uint64_t* iidx;
    const uint32_t* iadj;
    const float* inw;           /* per transpose arc: 1/out-degree of its source (FP32) */
    const float* invdeg;
    const uint32_t* hub_v;
    const uint32_t* hub_c0;
    const uint32_t* chunk_h;
    float* y0;
    float* y1;
    double* partS;
    double* partD;
    float* hubpart;
    double* hubS;
    double* hubD;
    double* chain;
    double* res;
    unsigned long long* dig;
    PprSlots* slots;
    PprOut* out;
};

__device__ __forceinline__ void kahan(float& s, float& c, float x)
{
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

__device__ __forceinline__ float ppr_y(float acc, double invS, double Dn, bool seed)
{
    double t = (double)acc * invS;
    if (seed) t += Dn;
    double y = 0.85 * t;
    if (seed) y += 0.15;
    return (float)y;
}

/* y_t(u) as the pull reads it: e_seed for a slot in its first iteration */
__device__ __forceinline__ float yval64(const float* yin, uint32_t u, int lane, uint32_t seed, bool fresh,
                                        bool live)
{
    if (!live) return 0.f;
    if (fresh) return u == seed ? 1.f : 0.f;
    return yin[(uint64_t)u * 64 + lane];
}

/* Both pulls walk a 64-vertex segment as runs of consecutive non-hub vertices
 * (their in-lists are contiguous in the transpose) in 64-arc chunks: arc ids
 * and weights are loaded coalesced, then folded in list order. Before arc a
 * is folded, every vertex whose list ends at or before a is closed (PPR_CLOSE),
 * so a vertex's Kahan sum covers exactly its own list, in order. */

/* batched: lane q folds slot q; 16 rows of 256 bytes in flight per wavefront */
#define PPR_STREAM64(RA, RB, EJ, CLOSE)                                                             \
    for (uint64_t e_ = (RA); e_ < (RB); e_ += 64) {                                               \
        const uint32_t cnt_ = (uint32_t)min<uint64_t>(64, (RB) - e_);                              \
        uint32_t ul_ = 0;                                                                           \
        float wl_ = 0.f;                                                                            \
        if ((uint32_t)lane < cnt_) {                                                                \
            ul_ = a.iadj[e_ + lane];                                                                \
            wl_ = a.inw[e_ + lane];                                                                 \
        }                                                                                           \
        for (uint32_t k_ = 0; k_ < cnt_; k_ += 16) {                                                \
            uint32_t uu_[16];                                                                       \
            float vv_[16];                                                                          \
            _Pragma("unroll") for (int i_ = 0; i_ < 16; ++i_) uu_[i_] = rl(ul_, (k_ + i_) & 63);    \
            _Pragma("unroll") for (int i_ = 0; i_ < 16; ++i_) vv_[i_] =                             \
                (k_ + i_ < cnt_) ? yval64(yin, uu_[i_], lane, seed, fresh, live) : 0.f;              \
            /* products first: the fold then reads registers only, so a row    */                   \
            /* store when a vertex closes does not stall the next fold step     */                   \
            _Pragma("unroll") for (int i_ = 0; i_ < 16; ++i_) vv_[i_] *= rlf(wl_, (k_ + i_) & 63);  \
            if (k_ + 16 <= cnt_ && e_ + k_ + 16 <= (EJ)) {                                          \
                _Pragma("unroll") for (int i_ = 0; i_ < 16; ++i_) kahan(s, c, vv_[i_]);             \
            } else {                                                                                \
                _Pragma("unroll") for (int i_ = 0; i_ < 16; ++i_) {                                 \
                    if (k_ + i_ < cnt_) {                                                           \
                        while (e_ + k_ + i_ >= (EJ)) { CLOSE; }                                     \
                        kahan(s, c, vv_[i_]);                                                       \
                    }                                                                               \
                }                                                                                   \
            }                                                                                       \
        }                                                                                           \
    }

/* batched pull. Work items: the hub chunks first, then the 64-vertex segments;
 * y and the segment sums follow id order. */
__global__ __launch_bounds__(256) void kp_pull_b(PprArgs a, int cur)
{
    const int lane = threadIdx.x & 63;
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const PprSlots* sl = a.slots;
    const bool live = (sl->live >> lane) & 1ull, fresh = (sl->fresh >> lane) & 1ull;
    const uint32_t seed = sl->seed[lane];
    const double invS = sl->invS[lane], Dn = sl->Dn[lane];
    const float* yin = cur ? a.y1 : a.y0;
    float* yout = cur ? a.y0 : a.y1;
    const uint64_t nitems = (uint64_t)a.nchunk + a.nseg;
    for (uint64_t it = wid; it < nitems; it += nw) {
        if (it < a.nchunk) {
            const uint32_t h = a.chunk_h[it];
            const uint32_t v = a.hub_v[h];
            const uint64_t e0 = a.iidx[v] + (uint64_t)(it - a.hub_c0[h]) * HUB;
            const uint64_t e1 = min<uint64_t>(e0 + HUB, a.iidx[v + 1]);
            float s = 0.f, c = 0.f;
            PPR_STREAM64(e0, e1, ~0ull, (void)0)
            if (live) a.hubpart[it * 64 + lane] = s - c;
            continue;
        }
        const uint64_t seg = it - a.nchunk;
        const uint64_t base = a.seg_v0[seg];
        const uint32_t nv = a.seg_v0[seg + 1] - (uint32_t)base;
        const bool in = (uint32_t)lane < nv;
        const uint64_t E0 = in ? a.iidx[base + lane] : 0, E1 = in ? a.iidx[base + lane + 1] : 0;
        const float dg = in ? a.invdeg[base + lane] : 1.f;
        const unsigned long long hubm = __ballot(in && E1 - E0 > HUB), dangm = __ballot(in && dg == 0.f);
        double sS = 0.0, sD = 0.0;
        uint32_t j = 0;
        while (j < nv) {
            if ((hubm >> j) & 1ull) {
                j++;
                continue;
            }
            const unsigned long long rest = hubm >> j;
            const uint32_t h = rest ? j + (uint32_t)__builtin_ctzll(rest) : nv;
            const uint64_t ra = rl64(E0, j), rb = rl64(E1, h - 1);
            float s = 0.f, c = 0.f;
            uint64_t ej = rl64(E1, j);
#define CLOSE_B                                                    \
    {                                                              \
        const float y_ = ppr_y(s - c, invS, Dn, base + j == seed); \
        if (live) yout[(base + j) * 64 + lane] = y_;               \
        sS += (double)y_;                                          \
        if ((dangm >> j) & 1ull) sD += (double)y_;                 \
        s = 0.f;                                                   \
        c = 0.f;                                                   \
        j++;                                                       \
        ej = j < h ? rl64(E1, j) : ~0ull;                          \
    }
            PPR_STREAM64(ra, rb, ej, CLOSE_B)
            while (j < h) CLOSE_B
#undef CLOSE_B
        }
        if (live) {
            a.partS[seg * 64 + lane] = sS;
            a.partD[seg * 64 + lane] = sD;
        }
    }
}

/* scalar pull: the batched kernel's work items, order and arithmetic for one
 * query. Arcs are taken 256 at a time: each lane loads four (source ids and
 * weights coalesced, then the four sources' y), so an item of 2048 arcs waits
 * on 16 round trips instead of 64. A hub chunk is one Kahan chain, run by
 * every lane over the 256 products in list order (as each lane of the batched
 * kernel runs it for its slot). In a segment, the 256 products go to LDS
 * and lane v runs vertex v's own chain over its products among them, in list
 * order; a window costs as many fold steps as its longest vertex run instead
 * of 256. Each vertex's sum is the same chain over the same products in the
 * same order, so the result is bit-identical to the batched kernel's lane. */
struct PprQuad {
    float p[4];
};

__device__ __forceinline__ PprQuad ppr_load4(const PprArgs& a, const float* yin, uint64_t e, uint64_t end, int lane,
                                             bool fresh, uint32_t seed)
{
    uint32_t u[4];
    float w[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint64_t x = e + (uint64_t)k * 64 + (uint64_t)lane;
        u[k] = x < end ? a.iadj[x] : 0u;
        w[k] = x < end ? a.inw[x] : 0.f;
    }
    PprQuad q;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint64_t x = e + (uint64_t)k * 64 + (uint64_t)lane;
        q.p[k] = x < end ? (fresh ? (u[k] == seed ? 1.f : 0.f) : yin[u[k]]) * w[k] : 0.f;
    }
    return q;
}

__global__ __launch_bounds__(256) void kp_pull_s(PprArgs a, int cur)
{
    const int lane = threadIdx.x & 63;
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const PprSlots* sl = a.slots;
    if (!(sl->live & 1ull)) return;
    const bool fresh = sl->fresh & 1ull;
    const uint32_t seed = sl->seed[0];
    const double invS = sl->invS[0], Dn = sl->Dn[0];
    const float* yin = cur ? a.y1 : a.y0;
    float* yout = cur ? a.y0 : a.y1;
    __shared__ float wins[4][256];
    float* win = wins[threadIdx.x >> 6];
    const uint64_t nitems = (uint64_t)a.nchunk + a.nseg;
    for (uint64_t it = wid; it < nitems; it += nw) {
        if (it < a.nchunk) {
            const uint32_t h = a.chunk_h[it];
            const uint32_t v = a.hub_v[h];
            const uint64_t e0 = a.iidx[v] + (uint64_t)(it - a.hub_c0[h]) * HUB;
            const uint64_t e1 = min<uint64_t>(e0 + HUB, a.iidx[v + 1]);
            float s = 0.f, c = 0.f;
            for (uint64_t e = e0; e < e1; e += 256) {
                const PprQuad q = ppr_load4(a, yin, e, e1, lane, fresh, seed);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint64_t ek = e + (uint64_t)k * 64;
                    if (ek >= e1) break;
                    const uint32_t cnt = (uint32_t)min<uint64_t>(64, e1 - ek);
                    for (uint32_t i = 0; i < cnt; ++i) kahan(s, c, rlf(q.p[k], (int)i));
                }
            }
            if (lane == 0) a.hubpart[it] = s - c;
            continue;
        }
        const uint64_t seg = it - a.nchunk;
        const uint64_t base = a.seg_v0[seg];
        const uint32_t nv = a.seg_v0[seg + 1] - (uint32_t)base;
        const bool in = (uint32_t)lane < nv;
        const uint64_t E0 = in ? a.iidx[base + lane] : 0, E1 = in ? a.iidx[base + lane + 1] : 0;
        const float dg = in ? a.invdeg[base + lane] : 1.f;
        const bool hub = in && E1 - E0 > HUB;
        const unsigned long long hubm = __ballot(hub), dangm = __ballot(in && dg == 0.f);
        /* the non-hub vertices' lists lie end to end between hubs: stream
         * each run of them once */
        float s = 0.f, c = 0.f;
        uint32_t j = 0;
        while (j < nv) {
            if ((hubm >> j) & 1ull) {
                j++;
                continue;
            }
            const unsigned long long rest = hubm >> j;
            const uint32_t h = rest ? j + (uint32_t)__builtin_ctzll(rest) : nv;
            const uint64_t ra = rl64(E0, j), rb = rl64(E1, h - 1);
            const bool mine = (uint32_t)lane >= j && (uint32_t)lane < h;
            for (uint64_t e = ra; e < rb; e += 256) {
                const PprQuad q = ppr_load4(a, yin, e, rb, lane, fresh, seed);
                /* the window's 256 products in LDS, in arc order */
#pragma unroll
                for (int k = 0; k < 4; k++) win[k * 64 + lane] = q.p[k];
                __builtin_amdgcn_wave_barrier();
                const uint64_t en = min<uint64_t>(e + 256, rb);
                /* this lane's vertex's products in the window: [lo, hi) */
                uint32_t lo = 0, hi = 0;
                if (mine) {
                    const uint64_t b0 = E0 > e ? E0 : e, b1 = E1 < en ? E1 : en;
                    if (b1 > b0) { lo = (uint32_t)(b0 - e); hi = (uint32_t)(b1 - e); }
                }
                for (uint32_t t = lo; t < hi; t++) kahan(s, c, win[t]);
                __builtin_amdgcn_wave_barrier();
            }
            j = h;
        }
        /* y of the lane's vertex, then the segment's sums in vertex order */
        float y = 0.f;
        if (in && !hub) y = ppr_y(s - c, invS, Dn, base + (uint32_t)lane == seed);
        double sS = 0.0, sD = 0.0;
        for (uint32_t k = 0; k < nv; k++) {
            if ((hubm >> k) & 1ull) continue;
            const float yk = rlf(y, (int)k);
            sS += (double)yk;
            if ((dangm >> k) & 1ull) sD += (double)yk;
        }
        if (in && !hub) yout[base + lane] = y;
        if (lane == 0) {
            a.partS[seg] = sS;
            a.partD[seg] = sD;
        }
    }
}

/* hubs: fold the chunk values in order; batched wavefront per hub, scalar thread per hub */
template <int B>
__global__ __launch_bounds__(256) void kp_hubfin(PprArgs a, int cur)
{
    const PprSlots* sl = a.slots;
    float* yout = cur ? a.y0 : a.y1;
    if (B == 64) {
        const int lane = threadIdx.x & 63;
        const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
        const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
        const bool live = (sl->live >> lane) & 1ull;
        const uint32_t seed = sl->seed[lane];
        const double invS = sl->invS[lane], Dn = sl->Dn[lane];
        for (uint64_t h = wid; h < a.nhub; h += nw) {
            float s = 0.f, c = 0.f;
            for (uint32_t k = a.hub_c0[h]; k < a.hub_c0[h + 1]; ++k) kahan(s, c, a.hubpart[(uint64_t)k * 64 + lane]);
            const uint32_t v = a.hub_v[h];
            const float y = ppr_y(s - c, invS, Dn, v == seed);
            if (live) {
                yout[(uint64_t)v * 64 + lane] = y;
                a.hubS[h * 64 + lane] = (double)y;
                a.hubD[h * 64 + lane] = a.invdeg[v] == 0.f ? (double)y : 0.0;
            }
        }
    } else {
        if (!(sl->live & 1ull)) return;
        const uint32_t seed = sl->seed[0];
        const double invS = sl->invS[0], Dn = sl->Dn[0];
        const uint64_t st = (uint64_t)gridDim.x * blockDim.x;
        for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < a.nhub; h += st) {
            float s = 0.f, c = 0.f;
            for (uint32_t k = a.hub_c0[h]; k < a.hub_c0[h + 1]; ++k) kahan(s, c, a.hubpart[k]);
            const uint32_t v = a.hub_v[h];
            const float y = ppr_y(s - c, invS, Dn, v == seed);
            yout[v] = y;
            a.hubS[h] = (double)y;
            a.hubD[h] = a.invdeg[v] == 0.f ? (double)y : 0.0;
        }
    }
}

/* canonical FP64 reduction, level 1: chain r < CHAINS sums p[i] for i = r mod
 * CHAINS in ascending i; chain CHAINS sums the hub terms in order. p1/h1 may be
 * NULL (one quantity). Thread (r, q) with q fastest; slots outside *maskp idle. */
__global__ __launch_bounds__(256) void k_chain(const double* p0, const double* p1, const double* h0,
                                               const double* h1, uint64_t n, uint32_t nh, int B,
                                               const unsigned long long* maskp, double* chain)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t r = g / (uint64_t)B;
    const int q = (int)(g % (uint64_t)B);
    if (r > (uint64_t)CHAINS) return;
    if (!((*maskp >> q) & 1ull)) return;
    double s0 = 0.0, s1 = 0.0;
    if (r < (uint64_t)CHAINS) {
        for (uint64_t i = r; i < n; i += CHAINS) {
            s0 += p0[i * B + q];
            if (p1) s1 += p1[i * B + q];
        }
    } else {
        for (uint32_t h = 0; h < nh; ++h) {
            s0 += h0[(uint64_t)h * B + q];
            if (h1) s1 += h1[(uint64_t)h * B + q];
        }
    }
    chain[(r * B + q) * 2] = s0;
    chain[(r * B + q) * 2 + 1] = s1;
}

/* level 2: fixed tree over the CHAINS chains, plus the hub chain; workgroup per slot */
__global__ __launch_bounds__(256) void k_tree(const double* chain, int B, const unsigned long long* maskp,
                                              double* res)
{
    __shared__ double s0[256], s1[256];
    const int q = blockIdx.x, t = threadIdx.x;
    if (!((*maskp >> q) & 1ull)) return;
    s0[t] = chain[((uint64_t)t * B + q) * 2];
    s1[t] = chain[((uint64_t)t * B + q) * 2 + 1];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s0[t] += s0[t + w];
            s1[t] += s1[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        res[q * 2] = s0[0] + chain[((uint64_t)CHAINS * B + q) * 2];
        res[q * 2 + 1] = s1[0] + chain[((uint64_t)CHAINS * B + q) * 2 + 1];
    }
}

/* after an iteration: S, 1/S, D/S of every live slot; no slot is fresh any more */
__global__ void kp_params(PprArgs a, int B)
{
    const int q = threadIdx.x;
    PprSlots* sl = a.slots;
    if (q < B && ((sl->live >> q) & 1ull)) {
        const double S = a.res[q * 2], D = a.res[q * 2 + 1];
        sl->Sprev[q] = sl->S[q];
        sl->S[q] = S;
        sl->invS[q] = 1.0 / S;
        sl->Dn[q] = D / S;
    }
    __syncthreads();
    if (q == 0) sl->fresh = 0;
}

struct SlotUpd {
    int n;
    int slot[64];
    uint32_t seed[64];
    unsigned long long release;
};

__global__ void kp_slots(PprArgs a, SlotUpd u)
{
    if (threadIdx.x != 0) return;
    PprSlots* sl = a.slots;
    sl->live &= ~u.release;
    for (int i = 0; i < u.n; ++i) {
        const int q = u.slot[i];
        const uint32_t s = u.seed[i];
        const unsigned long long b = 1ull << q;
        sl->live |= b;
        sl->fresh |= b;
        sl->seed[q] = s;
        sl->invS[q] = 1.0;
        sl->S[q] = 1.0;
        sl->Sprev[q] = 1.0;
        sl->Dn[q] = a.oidx[s + 1] == a.oidx[s] ? 1.0 : 0.0;
    }
}

/* result of the slots in mask: x = y_T / S_T in FP32, digest, and the L1
 * change of the last iteration per segment (id order) */
__global__ __launch_bounds__(256) void kx_ext_b(PprArgs a, int latest, unsigned long long mask,
                                                unsigned long long first, double* partR, float* const* outp)
{
    const int lane = threadIdx.x & 63;
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.slots->ext = mask;
    const bool on = (mask >> lane) & 1ull, fst = (first >> lane) & 1ull;
    const double S = a.slots->S[lane], Sp = a.slots->Sprev[lane];
    const uint32_t seed = a.slots->seed[lane];
    const float* yT = latest ? a.y1 : a.y0;
    const float* yP = latest ? a.y0 : a.y1;
    float* o = on && outp ? outp[lane] : nullptr;
    unsigned long long dig = 0;
    for (uint64_t seg = wid; seg < a.nseg64; seg += nw) {
        const uint64_t base = seg * 64;
        const uint32_t nv = (uint32_t)min<uint64_t>(64, a.V - base);
        double sr = 0.0;
        if (on) {
            for (uint32_t j = 0; j < nv; ++j) {
                const uint64_t v = base + j;
                const float x = (float)((double)yT[v * 64 + lane] / S);
                const float xp = fst ? (v == seed ? 1.f : 0.f) : (float)((double)yP[v * 64 + lane] / Sp);
                sr += fabs((double)x - (double)xp);
                dig += dterm(v, __float_as_uint(x));
                if (o) o[v] = x;
            }
            partR[seg * 64 + lane] = sr;
        }
    }
    if (on && dig) atomicAdd(&a.dig[(blockIdx.x % SPREAD) * 64 + lane], dig);
}

__global__ __launch_bounds__(256) void kx_ext_s(PprArgs a, int latest, unsigned long long first, double* partR,
                                                float* out)
{
    const int lane = threadIdx.x & 63;
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.slots->ext = 1ull;
    const bool fst = first & 1ull;
    const double S = a.slots->S[0], Sp = a.slots->Sprev[0];
    const uint32_t seed = a.slots->seed[0];
    const float* yT = latest ? a.y1 : a.y0;
    const float* yP = latest ? a.y0 : a.y1;
    unsigned long long dig = 0;
    for (uint64_t seg = wid; seg < a.nseg64; seg += nw) {
        const uint64_t base = seg * 64, v = base + lane;
        const uint32_t nv = (uint32_t)min<uint64_t>(64, a.V - base);
        double d = 0.0;
        if (v < a.V) {
            const float x = (float)((double)yT[v] / S);
            const float xp = fst ? (v == seed ? 1.f : 0.f) : (float)((double)yP[v] / Sp);
            d = fabs((double)x - (double)xp);
            dig += dterm(v, __float_as_uint(x));
            if (out) out[v] = x;
        }
        double sr = 0.0;
        for (uint32_t j = 0; j < nv; ++j) sr += rld(d, j);
        if (lane == 0) partR[seg] = sr;
    }
    dig = wave_sum_u64(dig);
    if (lane == 0 && dig) atomicAdd(&a.dig[blockIdx.x % SPREAD], dig);
}

/* digest spreads of the slots in mask to out, residuals alongside; spreads cleared */
__global__ void kx_finish(PprArgs a, int B, unsigned long long mask)
{
    const int q = threadIdx.x;
    if (q >= B || !((mask >> q) & 1ull)) return;
    unsigned long long d = 0;
    for (int k = 0; k < SPREAD; ++k) {
        d += a.dig[k * B + q];
        a.dig[k * B + q] = 0;
    }
    a.out->dig[q] = d;
    a.out->res2[2 * q] = a.res[2 * q];
}

/* ------------------------------------------------------------------------ */
/* graph                                                                     */
/* ------------------------------------------------------------------------ */

__global__ void k_inw(uint64_t E, const uint32_t* iadj, const float* inv, float* inw)
{
    const uint64_t st = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += st) inw[e] = inv[iadj[e]];
}

__global__ void k_invdeg(uint64_t V, const uint64_t* oidx, float* inv)
{
    const uint64_t st = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += st) {
        const uint64_t d = oidx[v + 1] - oidx[v];
        inv[v] = d ? (float)(1.0 / (double)d) : 0.f;
    }
}

static unsigned grid_for(uint64_t work, unsigned per_block, unsigned cap)
{
    uint64_t b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

template <typename T>
static int dmalloc(T** p, uint64_t n, uint64_t* acc, const char* what)
{
    *p = nullptr;
    if (n == 0) n = 1;
    HCK(hipMalloc((void**)p, n * sizeof(T)), what);
    if (acc) *acc += n * sizeof(T);
    return 0;
}

extern "C" int ipr_batch_graph_create(uint64_t V, uint64_t E, const uint64_t* idx, const uint32_t* adj,
                                      const uint32_t* wgt, int device, ipr_bgraph** out)
{
    if (!out || (V && !idx) || (E && !adj)) {
        set_error("ipr_batch_graph_create: null argument");
        return -2;
    }
    if (V == 0 || V > 0xFFFFFFFEull) {
        set_error("ipr_batch_graph_create: vertex count %llu outside [1, 2^32-2]", (unsigned long long)V);
        return -2;
    }
    if (idx[0] != 0 || idx[V] != E) {
        set_error("ipr_batch_graph_create: offsets do not span [0, E]");
        return -2;
    }
    for (uint64_t v = 0; v < V; v++)
        if (idx[v + 1] < idx[v]) {
            set_error("ipr_batch_graph_create: offsets decrease at vertex %llu", (unsigned long long)v);
            return -2;
        }
    for (uint64_t e = 0; e < E; e++)
        if (adj[e] >= V) {
            set_error("ipr_batch_graph_create: arc %llu points outside the graph", (unsigned long long)e);
            return -2;
        }
    HCK(hipSetDevice(device), "set device");
    ipr_bgraph* g = new ipr_bgraph();
    memset(g, 0, sizeof *g);
    g->device = device;
    g->V = V;
    g->E = E;
    g->tune = default_tuning();
    if (wgt && E) {
        /* near/far step: four mean arc weights */
        uint64_t sum = 0;
        for (uint64_t e = 0; e < E; e++) sum += wgt[e];
        const uint64_t d = 4 * ((sum + E - 1) / E);
        g->tune.delta = d >= INF ? INF : (uint32_t)(d ? d : 1);
        g->tune.delta1 = g->tune.delta;
    }
    g->nseg64 = (V + 63) / 64;
    int rc = 0;
    uint64_t* kidx = nullptr;
    std::vector<uint64_t> hin;
    std::vector<uint32_t> hv, hc0, ch, sv;
    unsigned long long* dsum = nullptr;
#define GO(x) do { if ((rc = (x)) != 0) goto fail; } while (0)
#define GH(x, w) do { hipError_t e_ = (x); if (e_ != hipSuccess) { rc = hip_fail(e_, w); goto fail; } } while (0)
    GH(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "graph stream");
    GO(dmalloc(&g->oidx, V + 1, &g->bytes, "graph offsets"));
    GO(dmalloc(&g->oadj, E, &g->bytes, "graph arcs"));
    if (wgt) GO(dmalloc(&g->owgt, E, &g->bytes, "graph weights"));
    GO(dmalloc(&g->iidx, V + 1, &g->bytes, "transpose offsets"));
    GO(dmalloc(&g->iadj, E, &g->bytes, "transpose arcs"));
    GO(dmalloc(&g->invdeg, V, &g->bytes, "inverse degrees"));
    GO(dmalloc(&g->inw, E, &g->bytes, "transpose arc weights"));
    GH(hipMemcpyAsync(g->oidx, idx, (V + 1) * 8, hipMemcpyHostToDevice, g->stream), "upload offsets");
    if (E) GH(hipMemcpyAsync(g->oadj, adj, E * 4, hipMemcpyHostToDevice, g->stream), "upload arcs");
    if (wgt && E) GH(hipMemcpyAsync(g->owgt, wgt, E * 4, hipMemcpyHostToDevice, g->stream), "upload weights");
    GH(hipStreamSynchronize(g->stream), "upload");
    if (ipr_hip_build_csc(V, E, g->oidx, g->oadj, nullptr, g->iidx, g->iadj, nullptr, g->stream) != 0) {
        set_error("transpose: %s", ipr_hip_last_error());
        rc = -1;
        goto fail;
    }
    hipLaunchKernelGGL(k_invdeg, dim3(grid_for(V, 256, 4096)), dim3(256), 0, g->stream, V, g->oidx, g->invdeg);
    hipLaunchKernelGGL(k_inw, dim3(grid_for(E, 256, 8192)), dim3(256), 0, g->stream, E, g->iadj, g->invdeg, g->inw);
    GH(hipGetLastError(), "inverse degrees");
    /* hubs and their chunks, from the transpose offsets */
    hin.resize(V + 1);
    GH(hipMemcpyAsync(hin.data(), g->iidx, (V + 1) * 8, hipMemcpyDeviceToHost, g->stream), "transpose offsets");
    GH(hipStreamSynchronize(g->stream), "transpose offsets");
    for (uint64_t v = 0; v < V; v++) {
        const uint64_t d = hin[v + 1] - hin[v];
        if (d > HUB) {
            hc0.push_back((uint32_t)ch.size());
            const uint32_t k = (uint32_t)((d + HUB - 1) / HUB);
            for (uint32_t i = 0; i < k; i++) ch.push_back((uint32_t)hv.size());
            hv.push_back((uint32_t)v);
        }
    }
    hc0.push_back((uint32_t)ch.size());
    /* pull segments: <= 64 vertices and <= SEG_ARCS non-hub in-arcs */
    {
        uint64_t v = 0;
        sv.push_back(0);
        while (v < V) {
            uint64_t arcs = 0;
            uint32_t n = 0;
            while (v < V && n < 64) {
                const uint64_t d = hin[v + 1] - hin[v];
                const uint64_t c = d > HUB ? 0 : d;
                if (n > 0 && arcs + c > SEG_ARCS) break;
                arcs += c;
                v++;
                n++;
            }
            sv.push_back((uint32_t)v);
        }
        g->nseg = sv.size() - 1;
    }
    g->nhub = (uint32_t)hv.size();
    g->nchunk = (uint32_t)ch.size();
    GO(dmalloc(&g->hub_v, hv.size(), &g->bytes, "hub list"));
    GO(dmalloc(&g->hub_c0, hc0.size(), &g->bytes, "hub chunks"));
    GO(dmalloc(&g->chunk_h, ch.size(), &g->bytes, "chunk hubs"));
    GO(dmalloc(&g->seg_v0, sv.size(), &g->bytes, "segments"));
    GH(hipMemcpy(g->seg_v0, sv.data(), sv.size() * 4, hipMemcpyHostToDevice), "segments");
    if (!hv.empty()) GH(hipMemcpy(g->hub_v, hv.data(), hv.size() * 4, hipMemcpyHostToDevice), "hub list");
    GH(hipMemcpy(g->hub_c0, hc0.data(), hc0.size() * 4, hipMemcpyHostToDevice), "hub chunks");
    if (!ch.empty()) GH(hipMemcpy(g->chunk_h, ch.data(), ch.size() * 4, hipMemcpyHostToDevice), "chunk hubs");
    GO(dmalloc(&dsum, 1, nullptr, "digest"));
    GH(hipMemsetAsync(dsum, 0, 8, g->stream), "digest");
    hipLaunchKernelGGL(k_digest_u32, dim3(grid_for(V, 256, 2048)), dim3(256), 0, g->stream,
                       (const uint32_t*)nullptr, V, dsum);
    GH(hipGetLastError(), "digest");
    GH(hipMemcpyAsync(&g->inf_digest, dsum, 8, hipMemcpyDeviceToHost, g->stream), "digest");
    GH(hipStreamSynchronize(g->stream), "graph create");
    (void)hipFree(dsum);
    (void)kidx;
    *out = g;
    return 0;
fail:
    if (dsum) (void)hipFree(dsum);
    ipr_batch_graph_destroy(g);
    return rc ? rc : -1;
#undef GO
#undef GH
}

extern "C" int ipr_batch_graph_destroy(ipr_bgraph* g)
{
    if (!g) return 0;
    (void)hipSetDevice(g->device);
    void* ps[] = {g->oidx, g->oadj, g->owgt, g->iidx, g->iadj, g->inw, g->invdeg, g->hub_v, g->hub_c0, g->chunk_h,
                  g->seg_v0};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return 0;
}

extern "C" uint64_t ipr_batch_graph_bytes(const ipr_bgraph* g) { return g ? g->bytes : 0; }

extern "C" int ipr_batch_graph_get_tuning(const ipr_bgraph* g, ipr_batch_tuning* t)
{
    if (!g || !t) {
        set_error("ipr_batch_graph_get_tuning: null argument");
        return -2;
    }
    t->large = g->tune.large;
    t->chunk = g->tune.chunk;
    t->grid = g->tune.grid;
    t->delta = g->tune.delta;
    t->delta1 = g->tune.delta1;
    t->cc_mode = g->tune.cc_mode;
    t->sssp_states = g->tune.sssp_states;
    return 0;
}

extern "C" int ipr_batch_graph_set_tuning(ipr_bgraph* g, const ipr_batch_tuning* t)
{
    if (!g || !t) {
        set_error("ipr_batch_graph_set_tuning: null argument");
        return -2;
    }
    if (t->large < 1 || t->chunk < 64 || t->chunk > (1u << 20) || t->grid < 1 || t->grid > 65535 ||
        t->delta < 1 || t->delta1 < 1 || t->cc_mode > 1 || t->sssp_states < 1 || t->sssp_states > 8) {
        set_error("ipr_batch_graph_set_tuning: large >= 1, chunk in [64, 2^20], grid in [1, 65535], "
                  "delta and delta1 >= 1, cc_mode 0 or 1, sssp_states in [1, 8]");
        return -2;
    }
    g->tune.large = t->large;
    g->tune.chunk = t->chunk;
    g->tune.grid = t->grid;
    g->tune.delta = t->delta;
    g->tune.delta1 = t->delta1;
    g->tune.cc_mode = t->cc_mode;
    g->tune.sssp_states = t->sssp_states;
    return 0;
}

extern "C" uint64_t ipr_batch_slot_bytes(const ipr_bgraph* g, int kind)
{
    if (!g) return 0;
    switch (kind) {
    case IPR_Q_SSSP: return PushState::footprint(g, true);
    case IPR_Q_PPR: return PprState::footprint(g, 64);
    case IPR_Q_CC: return PushState::footprint(g, false);
    default: return 0;
    }
}

/* ------------------------------------------------------------------------ */
/* PushState                                                                 */
/* ------------------------------------------------------------------------ */

static PushArgs push_args(const PushState* s)
{
    PushArgs a;
    const ipr_bgraph* g = s->g;
    a.V = g->V;
    a.oidx = g->oidx;
    a.oadj = g->oadj;
    a.owgt = g->owgt;
    a.val = s->val;
    a.mbox = s->mbox;
    a.act = s->act;
    a.vmask = s->vmask;
    a.touched = s->touched;
    a.tgran = s->tgran;
    a.small0 = s->small[0];
    a.small1 = s->small[1];
    a.large0 = s->large[0];
    a.large1 = s->large[1];
    a.lpre = s->lpre;
    a.ctl = s->ctl;
    a.pend = s->pend;
    a.fst = s->fst;
    a.large = g->tune.large;
    a.chunk = g->tune.chunk;
    a.wzero = 0;
    a.delta = s->batched ? g->tune.delta : g->tune.delta1;
    return a;
}

static uint64_t large_capacity(const ipr_bgraph* g)
{
    /* vertices with out-degree above `large` number at most E / large */
    uint64_t c = g->E / (g->tune.large + 1ull) + 1;
    return c < g->V ? c : g->V;
}

uint64_t PushState::footprint(const ipr_bgraph* g, bool batched)
{
    const uint64_t V = g->V, W = batched ? 64 : 1;
    const uint64_t lc = large_capacity(g);
    uint64_t b = 2 * V * W * 4 + ((V + 1023) / 1024) * 1028 + 2 * V * 4 + 2 * lc * 4 + (lc + 1) * 4 + sizeof(PushCtl);
    b += (batched ? 3 * V * 8 : V) + V * 4;
    return b;
}

int PushState::create(ipr_bgraph* g, bool batched, int use_graph, PushState** out)
{
    HCK(hipSetDevice(g->device), "set device");
    PushState* s = new PushState();
    s->g = g;
    s->batched = batched;
    s->use_graph = graph_mode(use_graph);
    const uint64_t V = g->V, W = batched ? 64 : 1;
    s->large_cap = large_capacity(g);
    int rc = 0;
#define GO(x) do { if ((rc = (x)) != 0) { delete s; return rc; } } while (0)
#define GH(x, w) do { hipError_t e_ = (x); if (e_ != hipSuccess) { rc = hip_fail(e_, w); delete s; return rc; } } while (0)
    GH(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking), "state stream");
    GH(hipEventCreateWithFlags(&s->ev[0], hipEventDisableTiming), "state event");
    GH(hipEventCreateWithFlags(&s->ev[1], hipEventDisableTiming), "state event");
    GO(dmalloc(&s->val, V * W, &s->bytes, "values"));
    GO(dmalloc(&s->mbox, V * W, &s->bytes, "mailboxes"));
    if (batched) {
        GO(dmalloc(&s->act, V, &s->bytes, "active masks"));
        GO(dmalloc(&s->vmask, V, &s->bytes, "valid masks"));
    }
    const uint64_t tb = ((V + 1023) / 1024) * 1024;
    GO(dmalloc(&s->touched, tb, &s->bytes, "touched map"));
    GO(dmalloc(&s->tgran, tb / 256, &s->bytes, "touched granules"));
    if (batched) GO(dmalloc((unsigned long long**)&s->pend, V, &s->bytes, "pending masks"));
    else GO(dmalloc((uint8_t**)&s->pend, V, &s->bytes, "pending flags"));
    GO(dmalloc(&s->fst, V, &s->bytes, "frontier stamps"));
    for (int p = 0; p < 2; p++) {
        GO(dmalloc(&s->small[p], V, &s->bytes, "frontier"));
        GO(dmalloc(&s->large[p], s->large_cap, &s->bytes, "large frontier"));
    }
    GO(dmalloc(&s->lpre, s->large_cap + 1, &s->bytes, "chunk prefix"));
    GO(dmalloc(&s->ctl, 1, &s->bytes, "control"));
    GH(hipHostMalloc((void**)&s->hout, 2 * sizeof(TickOut), hipHostMallocDefault), "pinned tick output");
    memset(s->hout, 0, 2 * sizeof(TickOut));
    GH(hipMemsetAsync(s->touched, 0, tb, s->st), "clear touched");
    GH(hipMemsetAsync(s->tgran, 0, tb / 256, s->st), "clear granules");
    GH(hipMemsetAsync(s->pend, 0, batched ? V * 8 : V, s->st), "clear pending");
    GH(hipMemsetAsync(s->fst, 0xFF, V * 4, s->st), "clear stamps");
    GH(hipMemsetAsync(s->ctl, 0, sizeof(PushCtl), s->st), "clear control");
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, s->st, push_args(s));
    GH(hipGetLastError(), "init control");
    if (batched) {
        GH(hipMemsetAsync(s->mbox, 0xFF, V * 64 * 4, s->st), "clear mailboxes");
        GH(hipMemsetAsync(s->vmask, 0, V * 8, s->st), "clear valid masks");
        GH(hipMemsetAsync(s->act, 0, V * 8, s->st), "clear active masks");
    }
    GH(hipStreamSynchronize(s->st), "state create");
#undef GO
#undef GH
    *out = s;
    return 0;
}

PushState::~PushState()
{
    if (g) (void)hipSetDevice(g->device);
    if (st) (void)hipStreamSynchronize(st);
    for (int w = 0; w < 2; w++)
        for (int p = 0; p < 2; p++)
            if (gx[w][p]) (void)hipGraphExecDestroy(gx[w][p]);
    for (int w = 0; w < 2; w++)
        for (int p = 0; p < 2; p++)
            for (int i = 0; i < 7; i++)
                if (gxn[w][p][i]) (void)hipGraphExecDestroy(gxn[w][p][i]);
    void* ps[] = {val, mbox, act, vmask, touched, tgran, pend, fst, small[0], small[1], large[0], large[1], lpre, ctl, dsum};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (hout) (void)hipHostFree(hout);
    for (int p = 0; p < 2; p++)
        if (ev[p]) (void)hipEventDestroy(ev[p]);
    if (st) (void)hipStreamDestroy(st);
}

static int enqueue_superstep(const PushState* s, int cur, uint32_t wzero)
{
    PushArgs a = push_args(s);
    a.wzero = wzero;
    const ipr_bgraph* g = s->g;
    const unsigned G = g->tune.grid;
    hipStream_t st = s->st;
    hipLaunchKernelGGL(k_push_begin, dim3(1), dim3(64), 0, st, a, cur);
    hipLaunchKernelGGL(k_push_lscan, dim3(1), dim3(1024), 0, st, a, cur);
    if (s->batched) {
        hipLaunchKernelGGL(k_relax_small_b, dim3(grid_for(g->V, 64, G)), dim3(256), 0, st, a, cur);
        hipLaunchKernelGGL(k_relax_large<true>, dim3(G), dim3(256), 0, st, a, cur);
        hipLaunchKernelGGL(k_apply_b, dim3(grid_for(g->V, 256, G)), dim3(256), 0, st, a, cur);
    } else {
        hipLaunchKernelGGL(k_relax_small_s, dim3(grid_for(g->V, 256, G)), dim3(256), 0, st, a, cur);
        hipLaunchKernelGGL(k_relax_large<false>, dim3(G), dim3(256), 0, st, a, cur);
        hipLaunchKernelGGL(k_apply_s, dim3(grid_for(g->V, 1024, G)), dim3(256), 0, st, a, cur);
    }
    hipLaunchKernelGGL(k_push_thr, dim3(1), dim3(64), 0, st, a);
    if (s->batched) hipLaunchKernelGGL(k_push_react<true>, dim3(grid_for(g->V, 256, G)), dim3(256), 0, st, a, cur ^ 1);
    else hipLaunchKernelGGL(k_push_react<false>, dim3(grid_for(g->V, 256, G)), dim3(256), 0, st, a, cur ^ 1);
    hipLaunchKernelGGL(k_push_out, dim3(1), dim3(64), 0, st, a, cur ^ 1, s->batched ? 1 : 0);
    HCK(hipGetLastError(), "superstep launch");
    return 0;
}

int PushState::start_sssp(uint32_t source)
{
    if (batched || source >= g->V) {
        set_error("start_sssp: %s", batched ? "batched state" : "source outside the graph");
        return -2;
    }
    PushArgs a = push_args(this);
    const unsigned G = g->tune.grid;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(g->V, 256, G)), dim3(256), 0, st, val, g->V, INF);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(g->V, 256, G)), dim3(256), 0, st, mbox, g->V, INF);
    hipLaunchKernelGGL(k_ctl_zero, dim3(1), dim3(1), 0, st, a, cur);
    hipLaunchKernelGGL(k_start_sssp_s, dim3(1), dim3(1), 0, st, a, cur, source);
    HCK(hipGetLastError(), "start sssp");
    wzero = 0;
    return 0;
}

int PushState::start_cc()
{
    if (batched) {
        set_error("start_cc: batched state");
        return -2;
    }
    PushArgs a = push_args(this);
    hipLaunchKernelGGL(k_ctl_zero, dim3(1), dim3(1), 0, st, a, cur);
    hipLaunchKernelGGL(k_start_cc_s, dim3(grid_for(g->V, 256, g->tune.grid)), dim3(256), 0, st, a, cur);
    HCK(hipGetLastError(), "start cc");
    wzero = 1;
    return 0;
}

int PushState::cc_union_find()
{
    if (batched) {
        set_error("cc_union_find: batched state");
        return -2;
    }
    const unsigned G = g->tune.grid;
    hipLaunchKernelGGL(k_uf_init, dim3(grid_for(g->V, 256, G)), dim3(256), 0, st, val, g->V);
    hipLaunchKernelGGL(k_uf_link, dim3(grid_for(g->V, 4, G)), dim3(256), 0, st, (const uint64_t*)g->oidx,
                       (const uint32_t*)g->oadj, val, g->V);
    hipLaunchKernelGGL(k_uf_flatten, dim3(grid_for(g->V, 256, G)), dim3(256), 0, st, val, g->V);
    HCK(hipGetLastError(), "union-find");
    HCK(hipEventRecord(ev[0], st), "union-find event");
    supersteps++;
    return 0;
}

int PushState::cc(uint64_t max_supersteps)
{
    int rc;
    if (g->tune.cc_mode == 1) {
        if ((rc = cc_union_find()) != 0) return rc;
        return wait(0);
    }
    if ((rc = start_cc()) != 0) return rc;
    return run_to_end(max_supersteps);
}

int PushState::admit(int n, const int* slots, const uint32_t* sources)
{
    if (!batched || n < 0 || n > 64) {
        set_error("admit: %s", batched ? "more than 64 admissions" : "scalar state");
        return -2;
    }
    if (n == 0) return 0;
    AdmitArgs ad;
    memset(&ad, 0, sizeof ad);
    ad.n = n;
    for (int i = 0; i < n; i++) {
        if (slots[i] < 0 || slots[i] >= 64 || sources[i] >= g->V) {
            set_error("admit: slot %d / source %u out of range", slots[i], sources[i]);
            return -2;
        }
        ad.slot[i] = slots[i];
        ad.src[i] = sources[i];
    }
    hipLaunchKernelGGL(k_admit_b, dim3(1), dim3(64), 0, st, push_args(this), ad, g->inf_digest);
    HCK(hipGetLastError(), "admit");
    return 0;
}

int PushState::release(unsigned long long mask)
{
    if (!batched || !mask) return 0;
    hipLaunchKernelGGL(k_release_b, dim3(grid_for(g->V, 256, g->tune.grid)), dim3(256), 0, st, push_args(this),
                       mask);
    HCK(hipGetLastError(), "release");
    return 0;
}

int PushState::steps(uint64_t k, int* parity)
{
    if (k == 0) return 0;
    while (k) {
        uint32_t i = 0;
        while (i < 6 && (2ull << i) <= k) i++;
        const uint32_t n = 1u << i;
        const int p0 = cur;
        if (use_graph == 1 && n > 1) {
            hipGraphExec_t& x = gxn[wzero][p0][i];
            if (!x) {
                hipGraph_t graph = nullptr;
                HCK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture supersteps");
                int rc = 0;
                for (uint32_t j = 0; j < n && !rc; j++) rc = enqueue_superstep(this, p0 ^ (int)(j & 1), wzero);
                const hipError_t e = hipStreamEndCapture(st, &graph);
                if (rc) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return rc;
                }
                HCK(e, "end capture");
                const hipError_t ei = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HCK(ei, "instantiate supersteps");
            }
            HCK(hipGraphLaunch(x, st), "superstep graph");
        } else if (use_graph == 1) {
            hipGraphExec_t& x = gx[wzero][p0];
            if (!x) {
                hipGraph_t graph = nullptr;
                HCK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture superstep");
                const int rc = enqueue_superstep(this, p0, wzero);
                const hipError_t e = hipStreamEndCapture(st, &graph);
                if (rc) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return rc;
                }
                HCK(e, "end capture");
                const hipError_t ei = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HCK(ei, "instantiate superstep");
            }
            HCK(hipGraphLaunch(x, st), "superstep graph");
        } else {
            for (uint32_t j = 0; j < n; j++) {
                const int rc = enqueue_superstep(this, p0 ^ (int)(j & 1), wzero);
                if (rc) return rc;
            }
        }
        cur = p0 ^ (int)(n & 1);
        supersteps += n;
        k -= n;
    }
    const int p = cur ^ 1;   /* parity of the last superstep */
    HCK(hipMemcpyAsync(&hout[p], &ctl->out, sizeof(TickOut), hipMemcpyDeviceToHost, st), "tick readback");
    HCK(hipEventRecord(ev[p], st), "tick event");
    if (parity) *parity = p;
    return 0;
}

int PushState::wait(int parity)
{
    HCK(hipEventSynchronize(ev[parity]), "tick wait");
    return 0;
}

int PushState::run_to_end(uint64_t max_supersteps)
{
    uint64_t k = 1, done = 0;
    for (;;) {
        int p = 0;
        const int rc0 = steps(k, &p);
        if (rc0) return rc0;
        done += k;
        const int rc = wait(p);
        if (rc) return rc;
        if (hout[p].orm == 0) return 0;   /* empty frontier, nothing parked */
        if (done >= max_supersteps) {
            set_error("run did not converge within %llu supersteps", (unsigned long long)max_supersteps);
            return -3;
        }
        k = std::min<uint64_t>(k * 2, 64);
    }
}

int PushState::values(uint32_t* host)
{
    HCK(hipMemcpyAsync(host, val, g->V * 4, hipMemcpyDeviceToHost, st), "values readback");
    HCK(hipStreamSynchronize(st), "values readback");
    return 0;
}

int PushState::columns(int n, uint32_t* host)
{
    if (!batched || n < 1 || n > 64) {
        set_error("columns: bad request");
        return -2;
    }
    uint32_t* d = nullptr;
    HCK(hipMallocAsync((void**)&d, (uint64_t)n * g->V * 4, st), "columns buffer");
    hipLaunchKernelGGL(k_columns_b, dim3(grid_for(g->V, 4, 8192)), dim3(256), 0, st, push_args(this), n, 0, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, d, (uint64_t)n * g->V * 4, hipMemcpyDeviceToHost, st);
    (void)hipFreeAsync(d, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    HCK(e, "columns");
    return 0;
}

int PushState::column(int slot, uint32_t* host)
{
    if (!batched || slot < 0 || slot >= 64) {
        set_error("column: bad slot");
        return -2;
    }
    uint32_t* d = nullptr;
    HCK(hipMallocAsync((void**)&d, g->V * 4, st), "column buffer");
    hipLaunchKernelGGL(k_columns_b, dim3(grid_for(g->V, 4, 8192)), dim3(256), 0, st, push_args(this), 1, slot, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, d, g->V * 4, hipMemcpyDeviceToHost, st);
    (void)hipFreeAsync(d, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    HCK(e, "column");
    return 0;
}

int PushState::digest(unsigned long long* d)
{
    if (!dsum) {
        HCK(hipMalloc((void**)&dsum, 8), "digest word");
        bytes += 8;
    }
    HCK(hipMemsetAsync(dsum, 0, 8, st), "digest clear");
    hipLaunchKernelGGL(k_digest_u32, dim3(grid_for(g->V, 256, 2048)), dim3(256), 0, st, (const uint32_t*)val, g->V,
                       dsum);
    HCK(hipGetLastError(), "digest");
    HCK(hipMemcpyAsync(&hout[0].dig[0], dsum, 8, hipMemcpyDeviceToHost, st), "digest readback");
    HCK(hipStreamSynchronize(st), "digest");
    *d = hout[0].dig[0];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* PprState                                                                  */
/* ------------------------------------------------------------------------ */

static PprArgs ppr_args(const PprState* s)
{
    PprArgs a;
    const ipr_bgraph* g = s->g;
    a.V = g->V;
    a.nseg = g->nseg;
    a.nseg64 = g->nseg64;
    a.seg_v0 = g->seg_v0;
    a.nhub = g->nhub;
    a.nchunk = g->nchunk;
    a.oidx = g->oidx;
    a.iidx = g->iidx;
    a.iadj = g->iadj;
    a.inw = g->inw;
    a.invdeg = g->invdeg;
    a.hub_v = g->hub_v;
    a.hub_c0 = g->hub_c0;
    a.chunk_h = g->chunk_h;
    a.y0 = s->y[0];
    a.y1 = s->y[1];
    a.partS = s->partS;
    a.partD = s->partD;
    a.hubpart = s->hubpart;
    a.hubS = s->hubS;
    a.hubD = s->hubD;
    a.chain = s->chain;
    a.res = s->res;
    a.dig = s->dig;
    a.slots = s->slots;
    a.out = s->out;
    return a;
}

uint64_t PprState::footprint(const ipr_bgraph* g, int B)
{
    const uint64_t b = (uint64_t)B;
    return 2 * g->V * b * 4 + 2 * std::max<uint64_t>(g->nseg, g->nseg64) * b * 8 + (uint64_t)g->nchunk * b * 4 + 2ull * g->nhub * b * 8 +
           (CHAINS + 1ull) * b * 16 + b * 16 + SPREAD * b * 8 + sizeof(PprSlots) + sizeof(PprOut) + b * 8;
}

int PprState::create(ipr_bgraph* g, int B, int use_graph, PprState** out)
{
    if (B != 1 && B != 64) {
        set_error("PprState: width must be 1 or 64");
        return -2;
    }
    HCK(hipSetDevice(g->device), "set device");
    PprState* s = new PprState();
    s->g = g;
    s->B = B;
    s->use_graph = graph_mode(use_graph);
    const uint64_t b = (uint64_t)B;
    int rc = 0;
#define GO(x) do { if ((rc = (x)) != 0) { delete s; return rc; } } while (0)
#define GH(x, w) do { hipError_t e_ = (x); if (e_ != hipSuccess) { rc = hip_fail(e_, w); delete s; return rc; } } while (0)
    GH(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking), "state stream");
    GH(hipEventCreateWithFlags(&s->ev, hipEventDisableTiming), "state event");
    GO(dmalloc(&s->y[0], g->V * b, &s->bytes, "ppr values"));
    GO(dmalloc(&s->y[1], g->V * b, &s->bytes, "ppr values"));
    const uint64_t np = std::max<uint64_t>(g->nseg, g->nseg64);
    GO(dmalloc(&s->partS, np * b, &s->bytes, "segment sums"));
    GO(dmalloc(&s->partD, np * b, &s->bytes, "segment sums"));
    GO(dmalloc(&s->hubpart, (uint64_t)g->nchunk * b, &s->bytes, "hub chunks"));
    GO(dmalloc(&s->hubS, (uint64_t)g->nhub * b, &s->bytes, "hub sums"));
    GO(dmalloc(&s->hubD, (uint64_t)g->nhub * b, &s->bytes, "hub sums"));
    GO(dmalloc(&s->chain, (CHAINS + 1ull) * b * 2, &s->bytes, "chains"));
    GO(dmalloc(&s->res, b * 2, &s->bytes, "sums"));
    GO(dmalloc(&s->dig, SPREAD * b, &s->bytes, "digests"));
    GO(dmalloc(&s->slots, 1, &s->bytes, "slots"));
    GO(dmalloc(&s->out, 1, &s->bytes, "output"));
    GO(dmalloc(&s->outp, b, &s->bytes, "output pointers"));
    GH(hipHostMalloc((void**)&s->hout, sizeof(PprOut), hipHostMallocDefault), "pinned output");
    memset(s->hout, 0, sizeof(PprOut));
    GH(hipMemsetAsync(s->slots, 0, sizeof(PprSlots), s->st), "clear slots");
    GH(hipMemsetAsync(s->dig, 0, SPREAD * b * 8, s->st), "clear digests");
    GH(hipMemsetAsync(s->out, 0, sizeof(PprOut), s->st), "clear output");
    GH(hipStreamSynchronize(s->st), "state create");
#undef GO
#undef GH
    *out = s;
    return 0;
}

PprState::~PprState()
{
    if (g) (void)hipSetDevice(g->device);
    if (st) (void)hipStreamSynchronize(st);
    for (int p = 0; p < 2; p++)
        if (gx[p]) (void)hipGraphExecDestroy(gx[p]);
    void* ps[] = {y[0], y[1], partS, partD, hubpart, hubS, hubD, chain, res, dig, slots, out, outp};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (hout) (void)hipHostFree(hout);
    if (ev) (void)hipEventDestroy(ev);
    if (st) (void)hipStreamDestroy(st);
}

int PprState::update(int nadmit, const int* sl, const uint32_t* seeds, unsigned long long release)
{
    if (nadmit < 0 || nadmit > B) {
        set_error("ppr update: too many admissions");
        return -2;
    }
    SlotUpd u;
    memset(&u, 0, sizeof u);
    u.n = nadmit;
    u.release = release;
    for (int i = 0; i < nadmit; i++) {
        if (sl[i] < 0 || sl[i] >= B || seeds[i] >= g->V) {
            set_error("ppr update: slot %d / seed %u out of range", sl[i], seeds[i]);
            return -2;
        }
        u.slot[i] = sl[i];
        u.seed[i] = seeds[i];
    }
    hipLaunchKernelGGL(kp_slots, dim3(1), dim3(1), 0, st, ppr_args(this), u);
    HCK(hipGetLastError(), "ppr slots");
    return 0;
}

static int enqueue_iteration(const PprState* s, int cur)
{
    const PprArgs a = ppr_args(s);
    const ipr_bgraph* g = s->g;
    const unsigned G = g->tune.grid;
    hipStream_t st = s->st;
    const uint64_t b = (uint64_t)s->B;
    const unsigned gp = grid_for(g->nseg + g->nchunk, 4, G);
    if (s->B == 64) hipLaunchKernelGGL(kp_pull_b, dim3(gp), dim3(256), 0, st, a, cur);
    else hipLaunchKernelGGL(kp_pull_s, dim3(gp), dim3(256), 0, st, a, cur);
    if (g->nhub) {
        if (s->B == 64) hipLaunchKernelGGL(kp_hubfin<64>, dim3(grid_for(g->nhub, 4, G)), dim3(256), 0, st, a, cur);
        else hipLaunchKernelGGL(kp_hubfin<1>, dim3(grid_for(g->nhub, 256, G)), dim3(256), 0, st, a, cur);
    }
    const unsigned gc = (unsigned)(((CHAINS + 1) * b + 255) / 256);
    hipLaunchKernelGGL(k_chain, dim3(gc), dim3(256), 0, st, (const double*)s->partS, (const double*)s->partD,
                       (const double*)s->hubS, (const double*)s->hubD, g->nseg, g->nhub, s->B,
                       (const unsigned long long*)&s->slots->live, s->chain);
    hipLaunchKernelGGL(k_tree, dim3(s->B), dim3(256), 0, st, (const double*)s->chain, s->B,
                       (const unsigned long long*)&s->slots->live, s->res);
    hipLaunchKernelGGL(kp_params, dim3(1), dim3(64), 0, st, a, s->B);
    HCK(hipGetLastError(), "ppr iteration launch");
    return 0;
}

int PprState::step()
{
    const int p = (int)(tick & 1);
    if (use_graph == 1) {
        hipGraphExec_t& x = gx[p];
        if (!x) {
            hipGraph_t graph = nullptr;
            HCK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture iteration");
            const int rc = enqueue_iteration(this, p);
            const hipError_t e = hipStreamEndCapture(st, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            HCK(e, "end capture");
            const hipError_t ei = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HCK(ei, "instantiate iteration");
        }
        HCK(hipGraphLaunch(x, st), "iteration graph");
    } else {
        const int rc = enqueue_iteration(this, p);
        if (rc) return rc;
    }
    tick++;
    return 0;
}

int PprState::extract(unsigned long long mask, unsigned long long first, float* const* dev_out)
{
    if (!mask) return 0;
    const PprArgs a = ppr_args(this);
    const int latest = (int)(tick & 1);
    const unsigned gs = grid_for(g->nseg64, 4, g->tune.grid);
    bool any_out = false;
    if (dev_out)
        for (int q = 0; q < B; q++) any_out |= dev_out[q] != nullptr;
    if (B == 64) {
        if (any_out) HCK(hipMemcpyAsync(outp, dev_out, 64 * sizeof(float*), hipMemcpyHostToDevice, st), "output pointers");
        hipLaunchKernelGGL(kx_ext_b, dim3(gs), dim3(256), 0, st, a, latest, mask, first, partS,
                           any_out ? (float* const*)outp : (float* const*)nullptr);
    } else {
        hipLaunchKernelGGL(kx_ext_s, dim3(gs), dim3(256), 0, st, a, latest, first, partS,
                           any_out ? dev_out[0] : (float*)nullptr);
    }
    const unsigned gc = (unsigned)(((CHAINS + 1) * (uint64_t)B + 255) / 256);
    hipLaunchKernelGGL(k_chain, dim3(gc), dim3(256), 0, st, (const double*)partS, (const double*)nullptr,
                       (const double*)nullptr, (const double*)nullptr, g->nseg64, 0u, B,
                       (const unsigned long long*)&slots->ext, chain);
    hipLaunchKernelGGL(k_tree, dim3(B), dim3(256), 0, st, (const double*)chain, B,
                       (const unsigned long long*)&slots->ext, res);
    hipLaunchKernelGGL(kx_finish, dim3(1), dim3(64), 0, st, a, B, mask);
    HCK(hipGetLastError(), "ppr extraction");
    HCK(hipMemcpyAsync(hout, out, sizeof(PprOut), hipMemcpyDeviceToHost, st), "ppr readback");
    return 0;
}

int PprState::record()
{
    HCK(hipEventRecord(ev, st), "ppr event");
    return 0;
}

int PprState::wait()
{
    HCK(hipEventSynchronize(ev), "ppr wait");
    return 0;
}

bool PprState::done() { return hipEventQuery(ev) == hipSuccess; }

/* ------------------------------------------------------------------------ */
/* one-shot entry points                                                     */
/* ------------------------------------------------------------------------ */

namespace {
struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t st;
    explicit Timer(hipStream_t s) : st(s)
    {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, st);
    }
    double stop()
    {
        float ms = 0.f;
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    }
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
}  // namespace

static int check_graph(const ipr_bgraph* g, const char* fn)
{
    if (!g) {
        set_error("%s: null graph", fn);
        return -2;
    }
    const hipError_t e = hipSetDevice(g->device);
    if (e != hipSuccess) return hip_fail(e, "set device");
    return 0;
}

extern "C" int ipr_batch_sssp(ipr_bgraph* g, const uint32_t* sources, int nq, uint32_t* out, ipr_batch_stats* st)
{
    int rc = check_graph(g, "ipr_batch_sssp");
    if (rc) return rc;
    if (nq < 1 || nq > 64 || !sources || !out) {
        set_error("ipr_batch_sssp: need 1..64 sources and an output");
        return -2;
    }
    PushState* s = nullptr;
    if ((rc = PushState::create(g, true, 1, &s)) != 0) return rc;
    int slots[64];
    for (int i = 0; i < nq; i++) slots[i] = i;
    const unsigned long long live = nq == 64 ? ~0ull : ((1ull << nq) - 1);
    Timer tm(s->st);
    uint64_t relax = 0;
    uint64_t ran = 0;
    if ((rc = s->admit(nq, slots, sources)) == 0) {
        /* doubling runs of supersteps between host checks; the supersteps
         * past the end find nothing to do and change nothing */
        for (uint64_t k = 1;; k = std::min<uint64_t>(k * 2, 64)) {
            int p = 0;
            if ((rc = s->steps(k, &p)) != 0 || (rc = s->wait(p)) != 0) break;
            relax = s->hout[p].relax;
            ran = s->hout[p].busy + 1;
            if (!(s->hout[p].orm & live)) break;
            if (s->supersteps > 64 * g->V + 64) {
                set_error("ipr_batch_sssp: no convergence");
                rc = -3;
                break;
            }
        }
    }
    const double ms = tm.stop();
    if (!rc) rc = s->columns(nq, out);
    if (!rc && st) {
        st->supersteps = ran;
        st->relaxations = relax;
        st->ms = ms;
    }
    delete s;
    return rc;
}

extern "C" int ipr_single_sssp(ipr_bgraph* g, uint32_t source, uint32_t* out, ipr_batch_stats* st)
{
    int rc = check_graph(g, "ipr_single_sssp");
    if (rc) return rc;
    if (!out) {
        set_error("ipr_single_sssp: null output");
        return -2;
    }
    PushState* s = nullptr;
    if ((rc = PushState::create(g, false, 1, &s)) != 0) return rc;
    Timer tm(s->st);
    if ((rc = s->start_sssp(source)) == 0) rc = s->run_to_end(64 * g->V + 64);
    const double ms = tm.stop();
    if (!rc) rc = s->values(out);
    if (!rc && st) {
        st->supersteps = s->hout[(s->cur ^ 1)].busy + 1;
        st->relaxations = s->hout[(s->cur ^ 1)].relax;
        st->ms = ms;
    }
    delete s;
    return rc;
}

extern "C" int ipr_batch_cc(ipr_bgraph* g, uint32_t* out, ipr_batch_stats* st)
{
    int rc = check_graph(g, "ipr_batch_cc");
    if (rc) return rc;
    if (!out) {
        set_error("ipr_batch_cc: null output");
        return -2;
    }
    PushState* s = nullptr;
    if ((rc = PushState::create(g, false, 1, &s)) != 0) return rc;
    Timer tm(s->st);
    rc = s->cc(g->V + 8);
    const double ms = tm.stop();
    if (!rc) rc = s->values(out);
    if (!rc && st) {
        st->supersteps = g->tune.cc_mode == 1 ? s->supersteps : s->hout[(s->cur ^ 1)].busy + 1;
        st->relaxations = g->tune.cc_mode == 1 ? g->E : s->hout[(s->cur ^ 1)].relax + g->E;
        st->ms = ms;
    }
    delete s;
    return rc;
}

static int ppr_run(ipr_bgraph* g, int B, const uint32_t* seeds, int nq, int iters, float* out, double* residual,
                   ipr_batch_stats* st)
{
    PprState* s = nullptr;
    int rc = PprState::create(g, B, 1, &s);
    if (rc) return rc;
    float* dbuf = nullptr;
    hipError_t e = hipMalloc((void**)&dbuf, (uint64_t)nq * g->V * 4);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "ppr result buffer");
    }
    float* ptrs[64] = {nullptr};
    for (int q = 0; q < nq; q++) ptrs[q] = dbuf + (uint64_t)q * g->V;
    int slots[64];
    for (int i = 0; i < nq; i++) slots[i] = i;
    const unsigned long long mask = nq == 64 ? ~0ull : ((1ull << nq) - 1);
    Timer tm(s->st);
    rc = s->update(nq, slots, seeds, 0);
    for (int it = 0; !rc && it < iters; it++) rc = s->step();
    const double ms = tm.stop();
    if (!rc) rc = s->extract(mask, iters == 1 ? mask : 0ull, ptrs);
    if (!rc) rc = s->record();
    if (!rc) rc = s->wait();
    if (!rc) {
        e = hipMemcpy(out, dbuf, (uint64_t)nq * g->V * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "ppr readback");
    }
    if (!rc && residual)
        for (int q = 0; q < nq; q++) residual[q] = s->hout->res2[2 * q];
    if (!rc && st) {
        st->supersteps = (uint64_t)iters;
        st->relaxations = (uint64_t)iters * g->E * (uint64_t)nq;
        st->ms = ms;
    }
    (void)hipFree(dbuf);
    delete s;
    return rc;
}

extern "C" int ipr_batch_ppr(ipr_bgraph* g, const uint32_t* seeds, int nq, int iters, float* out,
                             double* residual, ipr_batch_stats* st)
{
    int rc = check_graph(g, "ipr_batch_ppr");
    if (rc) return rc;
    if (nq < 1 || nq > 64 || iters < 1 || !seeds || !out) {
        set_error("ipr_batch_ppr: need 1..64 seeds, iters >= 1 and an output");
        return -2;
    }
    for (int i = 0; i < nq; i++)
        if (seeds[i] >= g->V) {
            set_error("ipr_batch_ppr: seed %u outside the graph", seeds[i]);
            return -2;
        }
    return ppr_run(g, 64, seeds, nq, iters, out, residual, st);
}

extern "C" int ipr_single_ppr(ipr_bgraph* g, uint32_t seed, int iters, float* out, double* residual,
                              ipr_batch_stats* st)
{
    int rc = check_graph(g, "ipr_single_ppr");
    if (rc) return rc;
    if (iters < 1 || !out || seed >= g->V) {
        set_error("ipr_single_ppr: bad arguments");
        return -2;
    }
    return ppr_run(g, 1, &seed, 1, iters, out, residual, st);
}
