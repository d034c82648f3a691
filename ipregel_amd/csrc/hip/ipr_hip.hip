/*
 * ipr_hip.hip - libipregel_hip.so: the thin C-ABI between the C host driver
 * and HIP on gfx950 (memory, streams, events, graphs), plus the kernels that
 * do not depend on the application's types: superstep control, CSC build,
 * in-degree binning.
 */
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <mutex>

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "ipregel_hip.h"
#include "ipregel_control.h"

static thread_local char g_err[512];

static int fail(hipError_t e, const char* what)
{
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
    return -(int)e;
}
#define CK(call)                                            \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return fail(e_, #call);       \
    } while (0)

extern "C" const char* ipr_hip_last_error(void) { return g_err; }

/* Scratch of the allocation-free entry point (ipr_hip_equal_noalloc): a
 * file-scope device array, one copy per GPU, held until the process ends.
 * Host code never passes its name to a kernel: the name is a host-side
 * stand-in, and the device address of each GPU's copy is looked up once per
 * device (hipGetSymbolAddress, with that device current) and kept here. */
#define IPR_NOALLOC_SCRATCH_BYTES 256
#define IPR_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) unsigned int scratch[IPR_NOALLOC_SCRATCH_BYTES / 4];
static unsigned int* g_scratch[IPR_NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

/* the current device's copy of `scratch` (looked up on first use) */
static int scratch_of_current_device(unsigned int** out)
{
    int dev = 0;
    CK(hipGetDevice(&dev));
    if (dev < 0 || dev >= IPR_NOALLOC_MAX_DEVICES) {
        snprintf(g_err, sizeof(g_err), "device %d is beyond the %d devices the scratch table holds", dev,
                 IPR_NOALLOC_MAX_DEVICES);
        return -(int)hipErrorInvalidDevice;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        CK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (unsigned int*)p;
    }
    *out = g_scratch[dev];
    return 0;
}

extern "C" int ipr_hip_init(int device, char* name, size_t name_len, uint64_t* hbm_bytes)
{
    int n = 0;
    CK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        snprintf(g_err, sizeof(g_err), "device %d not present (%d visible)", device, n);
        return -(int)hipErrorInvalidDevice;
    }
    CK(hipSetDevice(device));
    unsigned int* sp = nullptr;
    int rc = scratch_of_current_device(&sp);
    if (rc) return rc;
    hipDeviceProp_t p;
    CK(hipGetDeviceProperties(&p, device));
    if (name && name_len) snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    if (hbm_bytes) *hbm_bytes = (uint64_t)p.totalGlobalMem;
    return 0;
}

extern "C" int ipr_hip_cu_count(uint32_t* cus)
{
    int dev = 0, n = 0;
    CK(hipGetDevice(&dev));
    CK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    *cus = n > 0 ? (uint32_t)n : 1u;
    return 0;
}

extern "C" int ipr_hip_wallclock_hz(uint64_t* hz)
{
    int dev = 0, khz = 0;
    CK(hipGetDevice(&dev));
    CK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    *hz = (uint64_t)khz * 1000ull;
    return 0;
}

extern "C" int ipr_hip_malloc(void** p, size_t bytes)
{
    *p = nullptr;
    CK(hipMalloc(p, bytes ? bytes : 16));
    return 0;
}
extern "C" int ipr_hip_free(void* p)
{
    if (p) CK(hipFree(p));
    return 0;
}
extern "C" int ipr_hip_host_alloc(void** p, size_t bytes)
{
    *p = nullptr;
    CK(hipHostMalloc(p, bytes ? bytes : 16, hipHostMallocDefault));
    return 0;
}
extern "C" int ipr_hip_host_free(void* p)
{
    if (p) CK(hipHostFree(p));
    return 0;
}
extern "C" int ipr_hip_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream)
{
    if (!bytes) return 0;
    CK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return 0;
}
extern "C" int ipr_hip_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream)
{
    if (!bytes) return 0;
    CK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    return 0;
}
extern "C" int ipr_hip_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream)
{
    if (!bytes) return 0;
    CK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return 0;
}

__global__ void k_fill(uint8_t* p, uint8_t b, size_t n)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n16 = n / 16;
    uint4 v;
    v.x = v.y = v.z = v.w = 0x01010101u * b;
    if (((uintptr_t)p & 15) == 0) {
        for (size_t k = i; k < n16; k += stride) ((uint4*)p)[k] = v;
        for (size_t k = n16 * 16 + i; k < n; k += stride) p[k] = b;
    } else {
        for (size_t k = i; k < n; k += stride) p[k] = b;
    }
}

extern "C" int ipr_hip_fill(void* dst, int byte, size_t bytes, void* stream)
{
    if (!bytes) return 0;
    size_t blocks = (bytes / 16 + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (uint8_t*)dst, (uint8_t)byte, bytes);
    CK(hipGetLastError());
    return 0;
}

__global__ void k_set_u64(uint64_t* p, uint64_t v)
{
    *p = v;
}

extern "C" int ipr_hip_set_u64(void* dst, uint64_t value, void* stream)
{
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, (hipStream_t)stream, (uint64_t*)dst, value);
    CK(hipGetLastError());
    return 0;
}

extern "C" int ipr_hip_stream_create(void** stream)
{
    hipStream_t s;
    CK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return 0;
}
extern "C" int ipr_hip_stream_destroy(void* stream)
{
    if (stream) CK(hipStreamDestroy((hipStream_t)stream));
    return 0;
}
extern "C" int ipr_hip_stream_sync(void* stream)
{
    CK(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}
extern "C" int ipr_hip_event_create(void** ev)
{
    hipEvent_t e;
    CK(hipEventCreate(&e));
    *ev = (void*)e;
    return 0;
}
extern "C" int ipr_hip_event_destroy(void* ev)
{
    if (ev) CK(hipEventDestroy((hipEvent_t)ev));
    return 0;
}
extern "C" int ipr_hip_event_record(void* ev, void* stream)
{
    CK(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return 0;
}
extern "C" int ipr_hip_event_elapsed_ms(void* start, void* stop, float* ms)
{
    CK(hipEventSynchronize((hipEvent_t)stop));
    CK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return 0;
}

extern "C" int ipr_hip_graph_begin(void* stream)
{
    CK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return 0;
}
extern "C" int ipr_hip_graph_end(void* stream, void** exec)
{
    hipGraph_t g;
    hipGraphExec_t x;
    CK(hipStreamEndCapture((hipStream_t)stream, &g));
    hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(e, "hipGraphInstantiate");
    *exec = (void*)x;
    return 0;
}
extern "C" int ipr_hip_graph_launch(void* exec, void* stream)
{
    CK(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream));
    return 0;
}
extern "C" int ipr_hip_graph_destroy(void* exec)
{
    if (exec) CK(hipGraphExecDestroy((hipGraphExec_t)exec));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CSC                                                                       */
/* ------------------------------------------------------------------------ */

/* (dst << 32 | src) for every arc, one wavefront per source row */
__global__ void k_coo_keys(uint64_t V, const uint64_t* idx, const uint32_t* adj, uint64_t* keys)
{
    uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / 64;
    uint32_t lane = threadIdx.x & 63;
    for (uint64_t u = wave; u < V; u += nw) {
        uint64_t b = idx[u], e = idx[u + 1];
        for (uint64_t k = b + lane; k < e; k += 64) keys[k] = ((uint64_t)adj[k] << 32) | u;
    }
}

__global__ void k_iota_u32(uint32_t* p, uint64_t n)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = (uint32_t)i;
}

/* sorted keys -> in_adj, in_wgt and the row boundaries in_idx */
__global__ void k_csc_finish(uint64_t V, uint64_t E, const uint64_t* keys, const uint32_t* perm,
                             const uint32_t* out_wgt, uint64_t* in_idx, uint32_t* in_adj,
                             uint32_t* in_wgt)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= E; k += stride) {
        uint64_t dcur = k < E ? (keys[k] >> 32) : V;
        uint64_t dprev = k > 0 ? (keys[k - 1] >> 32) : (uint64_t)-1;
        if (k < E) {
            in_adj[k] = (uint32_t)keys[k];
            if (in_wgt) in_wgt[k] = out_wgt[perm[k]];
        }
        for (uint64_t v = dprev + 1; v <= dcur && v <= V; v++) in_idx[v] = k;
    }
}

static int nbits(uint64_t x)
{
    int b = 0;
    while (x) { b++; x >>= 1; }
    return b ? b : 1;
}

extern "C" int ipr_hip_build_csc(uint64_t V, uint64_t E, const uint64_t* out_idx,
                                 const uint32_t* out_adj, const uint32_t* out_wgt,
                                 uint64_t* in_idx, uint32_t* in_adj, uint32_t* in_wgt,
                                 void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    uint64_t *kin = nullptr, *kout = nullptr;
    uint32_t *pin = nullptr, *pout = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    int rc = 0;
    const unsigned end_bit = 32 + nbits(V ? V - 1 : 0);
    hipError_t e;
    if (E == 0) {
        hipLaunchKernelGGL(k_csc_finish, dim3(1), dim3(256), 0, st, V, E, (const uint64_t*)nullptr,
                           (const uint32_t*)nullptr, out_wgt, in_idx, in_adj, (uint32_t*)nullptr);
        CK(hipGetLastError());
        return 0;
    }
    if ((e = hipMalloc(&kin, E * 8)) != hipSuccess) { rc = fail(e, "csc keys"); goto out; }
    if ((e = hipMalloc(&kout, E * 8)) != hipSuccess) { rc = fail(e, "csc keys"); goto out; }
    hipLaunchKernelGGL(k_coo_keys, dim3(4096), dim3(256), 0, st, V, out_idx, out_adj, kin);
    if (out_wgt) {
        if ((e = hipMalloc(&pin, E * 4)) != hipSuccess) { rc = fail(e, "csc perm"); goto out; }
        if ((e = hipMalloc(&pout, E * 4)) != hipSuccess) { rc = fail(e, "csc perm"); goto out; }
        hipLaunchKernelGGL(k_iota_u32, dim3(4096), dim3(256), 0, st, pin, E);
        e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, pin, pout, (size_t)E, 0, end_bit, st);
    } else {
        e = rocprim::radix_sort_keys(nullptr, tmp_bytes, kin, kout, (size_t)E, 0, end_bit, st);
    }
    if (e != hipSuccess) { rc = fail(e, "radix sort size"); goto out; }
    if ((e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 16)) != hipSuccess) { rc = fail(e, "csc temp"); goto out; }
    if (out_wgt)
        e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, pin, pout, (size_t)E, 0, end_bit, st);
    else
        e = rocprim::radix_sort_keys(tmp, tmp_bytes, kin, kout, (size_t)E, 0, end_bit, st);
    if (e != hipSuccess) { rc = fail(e, "radix sort"); goto out; }
    {
        uint64_t blocks = (E + 1 + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(k_csc_finish, dim3((unsigned)blocks), dim3(256), 0, st, V, E, kout, pout,
                           out_wgt, in_idx, in_adj, out_wgt ? in_wgt : (uint32_t*)nullptr);
    }
    if ((e = hipGetLastError()) != hipSuccess) { rc = fail(e, "csc finish"); goto out; }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) { rc = fail(e, "csc sync"); goto out; }
out:
    (void)hipFree(kin); (void)hipFree(kout); (void)hipFree(pin); (void)hipFree(pout); (void)hipFree(tmp);
    return rc;
}

__global__ void k_widen_ids(const uint32_t* src, uint64_t* dst, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

extern "C" int ipr_hip_widen_ids(const uint32_t* src, uint64_t* dst, uint64_t n, void* stream)
{
    if (n == 0) return 0;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_widen_ids, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, src, dst, n);
    CK(hipGetLastError());
    return 0;
}

/* ------------------------------------------------------------------------ */
/* arc-balanced gather index                                                 */
/* ------------------------------------------------------------------------ */

/* sort key: out-degree classes [2^k, 2^(k+1)) in descending order, vertices
 * without out-arcs last; the radix sort is stable, so ids ascend within a
 * class. The hot records gather at the front, and a run of consecutive
 * frontier vertices of one class still writes consecutive records. */
__global__ void k_degree_keys(uint64_t V, const uint64_t* idx, uint32_t* key, uint32_t* id)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        const uint64_t dg = idx[v + 1] - idx[v];
        key[v] = dg == 0 ? 64u : 63u - (63u - (uint32_t)__clzll(dg));
        id[v] = (uint32_t)v;
    }
}

__global__ void k_scatter_pos(uint64_t V, const uint32_t* order, uint32_t* pos)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += stride)
        pos[order[i]] = (uint32_t)i;
}

extern "C" int ipr_hip_build_record_order(uint64_t V, const uint64_t* out_idx, uint32_t* bpos, uint32_t* bvert,
                                          void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (V == 0) return 0;
    if (V >= (1ull << 32)) {
        snprintf(g_err, sizeof(g_err), "record order: V=%llu needs 32-bit ids", (unsigned long long)V);
        return -1;
    }
    uint32_t *kin = nullptr, *kout = nullptr, *vin = nullptr, *vout = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    int rc = 0;
    hipError_t e;
    uint64_t vb = (V + 255) / 256;
    if (vb > 65536) vb = 65536;
    if ((e = hipMalloc(&kin, V * 4)) != hipSuccess) { rc = fail(e, "record order keys"); goto out; }
    if ((e = hipMalloc(&kout, V * 4)) != hipSuccess) { rc = fail(e, "record order keys"); goto out; }
    if ((e = hipMalloc(&vin, V * 4)) != hipSuccess) { rc = fail(e, "record order ids"); goto out; }
    if ((e = hipMalloc(&vout, V * 4)) != hipSuccess) { rc = fail(e, "record order ids"); goto out; }
    hipLaunchKernelGGL(k_degree_keys, dim3((unsigned)vb), dim3(256), 0, st, V, out_idx, kin, vin);
    e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, vin, vout, (size_t)V, 0, 7, st);
    if (e != hipSuccess) { rc = fail(e, "record order sort size"); goto out; }
    if ((e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 16)) != hipSuccess) { rc = fail(e, "record order temp"); goto out; }
    e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, vin, vout, (size_t)V, 0, 7, st);
    if (e != hipSuccess) { rc = fail(e, "record order sort"); goto out; }
    hipLaunchKernelGGL(k_scatter_pos, dim3((unsigned)vb), dim3(256), 0, st, V, vout, bpos);
    if (bvert && (e = hipMemcpyAsync(bvert, vout, V * 4, hipMemcpyDeviceToDevice, st)) != hipSuccess) {
        rc = fail(e, "record order inverse");
        goto out;
    }
    if ((e = hipGetLastError()) != hipSuccess) { rc = fail(e, "record order"); goto out; }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) { rc = fail(e, "record order sync"); goto out; }
out:
    (void)hipFree(kin); (void)hipFree(kout); (void)hipFree(vin); (void)hipFree(vout); (void)hipFree(tmp);
    return rc;
}

/* the transpose's sources, as record positions when records are reordered */
__global__ void k_gather_sources(uint64_t E, const uint32_t* t_adj, const uint32_t* bpos, uint32_t* g_arc)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < E; k += stride)
        g_arc[k] = bpos ? bpos[t_adj[k]] : t_adj[k];
}

__global__ void k_nonempty(uint64_t V, const uint64_t* idx, uint32_t* flag)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride)
        flag[v] = idx[v + 1] > idx[v] ? 1u : 0u;
}

/* list ranks -> g_dst; IPR_GATHER_LAST on the last arc of every non-empty list */
__global__ void k_gather_index(uint64_t V, const uint64_t* idx, const uint32_t* rank, uint32_t* g_dst,
                               uint32_t* g_arc)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        const uint64_t e = idx[v + 1];
        if (e > idx[v]) {
            g_dst[rank[v]] = (uint32_t)v;
            g_arc[e - 1] |= IPR_GATHER_LAST;
        }
    }
}

/* rank of the list holding arc c*K: the last v with idx[v] <= c*K (idx[V] = E
 * is larger, and that v's list is non-empty) */
__global__ void k_gather_owner(uint64_t V, uint64_t E, const uint64_t* idx, const uint32_t* rank, uint32_t K,
                               uint32_t* g_own)
{
    const uint64_t nch = (E + K - 1) / K;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nch; c += stride) {
        const uint64_t p = c * K;
        uint64_t lo = 0, hi = V;
        while (hi - lo > 1) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (idx[mid] <= p) lo = mid;
            else hi = mid;
        }
        g_own[c] = rank[lo];
    }
}

extern "C" int ipr_hip_build_gather(uint64_t V, uint64_t E, const uint64_t* t_idx, const uint32_t* t_adj,
                                    const uint32_t* bpos, uint32_t K, uint32_t* g_arc, uint32_t* g_own,
                                    uint32_t* g_dst, uint64_t* n_dst, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    *n_dst = 0;
    if (V >= (1ull << 31) || K == 0 || K % 4) {
        snprintf(g_err, sizeof(g_err), "gather index: needs V < 2^31 and K a multiple of 4 (V=%llu K=%u)",
                 (unsigned long long)V, K);
        return -1;
    }
    if (E == 0 || V == 0) return 0;
    uint32_t *flag = nullptr, *rank = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    uint32_t last[2] = {0, 0};
    int rc = 0;
    hipError_t e;
    uint64_t vb = (V + 255) / 256, cb = ((E + K - 1) / K + 255) / 256;
    if (vb > 65536) vb = 65536;
    if (cb > 65536) cb = 65536;
    {
        uint64_t eb = (E + 255) / 256;
        if (eb > 65536) eb = 65536;
        hipLaunchKernelGGL(k_gather_sources, dim3((unsigned)eb), dim3(256), 0, st, E, t_adj, bpos, g_arc);
    }
    if ((e = hipMalloc(&flag, V * 4)) != hipSuccess) { rc = fail(e, "gather flags"); goto out; }
    if ((e = hipMalloc(&rank, V * 4)) != hipSuccess) { rc = fail(e, "gather ranks"); goto out; }
    hipLaunchKernelGGL(k_nonempty, dim3((unsigned)vb), dim3(256), 0, st, V, t_idx, flag);
    e = rocprim::exclusive_scan(nullptr, tmp_bytes, flag, rank, 0u, (size_t)V, rocprim::plus<uint32_t>(), st);
    if (e != hipSuccess) { rc = fail(e, "gather scan size"); goto out; }
    if ((e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 16)) != hipSuccess) { rc = fail(e, "gather scan temp"); goto out; }
    e = rocprim::exclusive_scan(tmp, tmp_bytes, flag, rank, 0u, (size_t)V, rocprim::plus<uint32_t>(), st);
    if (e != hipSuccess) { rc = fail(e, "gather scan"); goto out; }
    hipLaunchKernelGGL(k_gather_index, dim3((unsigned)vb), dim3(256), 0, st, V, t_idx, rank, g_dst, g_arc);
    hipLaunchKernelGGL(k_gather_owner, dim3((unsigned)cb), dim3(256), 0, st, V, E, t_idx, rank, K, g_own);
    if ((e = hipGetLastError()) != hipSuccess) { rc = fail(e, "gather index"); goto out; }
    if ((e = hipMemcpyAsync(&last[0], rank + V - 1, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipMemcpyAsync(&last[1], flag + V - 1, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess) { rc = fail(e, "gather sync"); goto out; }
    *n_dst = (uint64_t)last[0] + last[1];
out:
    (void)hipFree(flag); (void)hipFree(rank); (void)hipFree(tmp);
    return rc;
}

/* gather split into IPR_GATHER_CLASSES classes by the cache line of the
 * source's record (line mod classes), so that the workgroups of one XCD read a
 * fixed share of the record lines: per class a transpose (counts, scans,
 * scatter; in-lists keep their order), then the index above per class */
__device__ __forceinline__ uint32_t ipr_arc_class(const uint32_t* bpos, uint32_t u)
{
    const uint32_t p = bpos ? bpos[u] : u;
    return (p / IPR_REC4_LINE) % IPR_GATHER_CLASSES;
}

__global__ void k_class_count(uint64_t V, const uint64_t* t_idx, const uint32_t* t_adj, const uint32_t* bpos,
                              uint64_t* cidx)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        uint64_t n[IPR_GATHER_CLASSES] = {};
        for (uint64_t e = t_idx[v]; e < t_idx[v + 1]; e++) n[ipr_arc_class(bpos, t_adj[e])]++;
        for (int c = 0; c < IPR_GATHER_CLASSES; c++) cidx[(uint64_t)c * (V + 1) + v] = n[c];
    }
}

__global__ void k_class_scatter(uint64_t V, const uint64_t* t_idx, const uint32_t* t_adj, const uint32_t* t_wgt,
                                const uint32_t* bpos, const uint64_t* cidx, const uint64_t* abase, uint32_t* cadj,
                                uint32_t* cwgt)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        uint64_t o[IPR_GATHER_CLASSES];
        for (int c = 0; c < IPR_GATHER_CLASSES; c++) o[c] = abase[c] + cidx[(uint64_t)c * (V + 1) + v];
        for (uint64_t e = t_idx[v]; e < t_idx[v + 1]; e++) {
            const uint32_t u = t_adj[e];
            const uint64_t q = o[ipr_arc_class(bpos, u)]++;
            cadj[q] = u;
            if (t_wgt) cwgt[q] = t_wgt[e];
        }
    }
}

extern "C" int ipr_hip_build_gather_classes(uint64_t V, uint64_t E, const uint64_t* t_idx, const uint32_t* t_adj,
                                            const uint32_t* t_wgt, const uint32_t* bpos, uint32_t K,
                                            uint32_t* g_arc, uint32_t* g_wgt, uint32_t* g_own, uint32_t* g_dst,
                                            uint64_t* cbeg, uint64_t* cend, uint64_t* cdst, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    for (int c = 0; c < IPR_GATHER_CLASSES; c++) cbeg[c] = cend[c] = cdst[c] = 0;
    if (E == 0 || V == 0) return 0;
    if (K == 0 || K % 4) {
        snprintf(g_err, sizeof(g_err), "gather classes: K=%u is not a multiple of 4", K);
        return -1;
    }
    const uint64_t nc = IPR_GATHER_CLASSES;
    uint64_t *ccnt = nullptr, *cidx = nullptr, *abase_d = nullptr;
    uint32_t *cadj = nullptr, *cwgt = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    uint64_t ec[IPR_GATHER_CLASSES], abase[IPR_GATHER_CLASSES];
    int rc = 0;
    hipError_t e;
    uint64_t vb = (V + 255) / 256;
    if (vb > 65536) vb = 65536;
    if ((e = hipMalloc(&ccnt, nc * (V + 1) * 8)) != hipSuccess) { rc = fail(e, "gather class counts"); goto out; }
    if ((e = hipMalloc(&cidx, nc * (V + 1) * 8)) != hipSuccess) { rc = fail(e, "gather class offsets"); goto out; }
    if ((e = hipMalloc(&abase_d, nc * 8)) != hipSuccess) { rc = fail(e, "gather class bases"); goto out; }
    if ((e = hipMalloc(&cadj, E * 4)) != hipSuccess) { rc = fail(e, "gather class arcs"); goto out; }
    if (t_wgt && g_wgt && (e = hipMalloc(&cwgt, E * 4)) != hipSuccess) { rc = fail(e, "gather class weights"); goto out; }
    /* the count slot V of every class is 0, so an exclusive scan over V+1
     * entries leaves the class's arc count at V */
    if ((e = hipMemsetAsync(ccnt, 0, nc * (V + 1) * 8, st)) != hipSuccess) { rc = fail(e, "gather class clear"); goto out; }
    hipLaunchKernelGGL(k_class_count, dim3((unsigned)vb), dim3(256), 0, st, V, t_idx, t_adj, bpos, ccnt);
    e = rocprim::exclusive_scan(nullptr, tmp_bytes, ccnt, cidx, (uint64_t)0, (size_t)(V + 1),
                                rocprim::plus<uint64_t>(), st);
    if (e != hipSuccess) { rc = fail(e, "gather class scan size"); goto out; }
    if ((e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 16)) != hipSuccess) { rc = fail(e, "gather class scan temp"); goto out; }
    for (uint64_t c = 0; c < nc; c++) {
        e = rocprim::exclusive_scan(tmp, tmp_bytes, ccnt + c * (V + 1), cidx + c * (V + 1), (uint64_t)0,
                                    (size_t)(V + 1), rocprim::plus<uint64_t>(), st);
        if (e != hipSuccess) { rc = fail(e, "gather class scan"); goto out; }
        if ((e = hipMemcpyAsync(&ec[c], cidx + c * (V + 1) + V, 8, hipMemcpyDeviceToHost, st)) != hipSuccess) {
            rc = fail(e, "gather class count"); goto out;
        }
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) { rc = fail(e, "gather class sync"); goto out; }
    {
        /* class c's arcs start at a multiple of K in g_arc (whole chunks for
         * g_own, 16-byte source loads) */
        uint64_t a = 0, s = 0;
        for (uint64_t c = 0; c < nc; c++) {
            abase[c] = s;
            s += ec[c];
            cbeg[c] = a;
            cend[c] = a + ec[c];
            a = (cend[c] + K - 1) / K * K;
        }
        if (s != E) {
            snprintf(g_err, sizeof(g_err), "gather classes: %llu of %llu arcs classed", (unsigned long long)s,
                     (unsigned long long)E);
            rc = -1;
            goto out;
        }
    }
    if ((e = hipMemcpyAsync(abase_d, abase, nc * 8, hipMemcpyHostToDevice, st)) != hipSuccess) { rc = fail(e, "gather class bases"); goto out; }
    hipLaunchKernelGGL(k_class_scatter, dim3((unsigned)vb), dim3(256), 0, st, V, t_idx, t_adj, cwgt ? t_wgt : nullptr,
                       bpos, cidx, abase_d, cadj, cwgt);
    if ((e = hipGetLastError()) != hipSuccess) { rc = fail(e, "gather class scatter"); goto out; }
    for (uint64_t c = 0, d = 0; c < nc; c++) {
        uint64_t nd = 0;
        cdst[c] = d;
        if (ec[c]) {
            rc = ipr_hip_build_gather(V, ec[c], cidx + c * (V + 1), cadj + abase[c], bpos, K, g_arc + cbeg[c],
                                      g_own + cbeg[c] / K, g_dst + d, &nd, stream);
            if (rc) goto out;
            /* the weights in the same order as the class's sources */
            if (cwgt && (e = hipMemcpyAsync(g_wgt + cbeg[c], cwgt + abase[c], ec[c] * 4, hipMemcpyDeviceToDevice,
                                            st)) != hipSuccess) {
                rc = fail(e, "gather class weights");
                goto out;
            }
        }
        d += nd;
    }
out:
    (void)hipStreamSynchronize(st);
    (void)hipFree(ccnt); (void)hipFree(cidx); (void)hipFree(abase_d); (void)hipFree(cadj); (void)hipFree(cwgt);
    (void)hipFree(tmp);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* receiver-blocked gather index                                             */
/* ------------------------------------------------------------------------ */

/* sort keys of block b's arcs, (b << kshift) | (source record position << 16)
 * | receiver offset in the block; workgroups take blocks in turn, a wavefront
 * per receiver */
__global__ void k_tile_keys(const uint64_t* t_idx, const uint32_t* t_adj, const uint32_t* bpos,
                            const uint32_t* bv0, const uint32_t* bnr, uint64_t nblk, unsigned kshift,
                            uint64_t* keys)
{
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (uint64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const uint32_t v0 = bv0[b], nr = bnr[b];
        for (uint32_t r = wv; r < nr; r += nw) {
            const uint64_t e0 = t_idx[v0 + r], e1 = t_idx[(uint64_t)v0 + r + 1];
            for (uint64_t k = e0 + lane; k < e1; k += 64) {
                const uint32_t u = t_adj[k];
                const uint64_t p = bpos ? bpos[u] : u;
                keys[k] = (b << kshift) | (p << 16) | r;
            }
        }
    }
}

/* group g's block: the last b with pofs[b] <= 64 g */
__device__ inline uint64_t tile_block_of(const uint64_t* pofs, uint64_t nblk, uint64_t slot)
{
    uint64_t lo = 0, hi = nblk;
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (pofs[mid] <= slot) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* per group of 64 slots: its base position (the first arc's), or a mark
 * (0xFFFFFFFF) when the group's positions span more than 16 bits, counted */
__global__ void k_tile_heads(uint64_t ngroups, const uint64_t* keys, const uint64_t* pofs, const uint64_t* aofs,
                             const uint32_t* narcs, uint64_t nblk, uint64_t pmask, uint32_t* head,
                             unsigned long long* nesc)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
        const uint64_t slot = g * 64, b = tile_block_of(pofs, nblk, slot);
        const uint64_t off = slot - pofs[b];
        if (off >= narcs[b]) { head[g] = 0; continue; }   /* a whole group of padding */
        const uint64_t k0 = aofs[b] + off;
        const uint64_t k1 = aofs[b] + (off + 63 < narcs[b] ? off + 63 : narcs[b] - 1);
        const uint64_t p0 = (keys[k0] >> 16) & pmask, p1 = (keys[k1] >> 16) & pmask;
        if (p1 - p0 > 0xFFFFu) {
            head[g] = 0xFFFFFFFFu;
            atomicAdd(nesc, 1ull);
        } else {
            head[g] = (uint32_t)p0;
        }
    }
}

/* the marked groups get their escape rows */
__global__ void k_tile_escapes(uint64_t ngroups, uint32_t* head, unsigned long long* next)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride)
        if (head[g] == 0xFFFFFFFFu) head[g] = IPR_TQ_ESC | (uint32_t)atomicAdd(next, 1ull);
}

/* each sorted arc into its slot: position delta and receiver offset (or the
 * position into the group's escape row), and its weight at the slot's width */
__global__ void k_tile_slots(uint64_t E, const uint64_t* keys, const uint32_t* perm, const uint32_t* t_wgt,
                             const uint64_t* pofs, const uint64_t* aofs, unsigned kshift, uint64_t pmask,
                             const uint32_t* head, uint32_t* q, uint32_t* esc, void* wgt, uint32_t wbytes)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < E; k += stride) {
        const uint64_t key = keys[k], b = key >> kshift;
        const uint64_t slot = pofs[b] + (k - aofs[b]);
        const uint32_t pos = (uint32_t)((key >> 16) & pmask), rcv = (uint32_t)(key & 0xFFFFu);
        const uint32_t h = head[slot >> 6];
        if (h & IPR_TQ_ESC) {
            esc[(uint64_t)(h & ~IPR_TQ_ESC) * 64 + (slot & 63)] = pos;
            q[slot] = rcv;
        } else {
            q[slot] = ((pos - h) << 16) | rcv;
        }
        if (wbytes) {
            const uint32_t w = t_wgt[perm[k]];
            if (wbytes == 1) ((uint8_t*)wgt)[slot] = (uint8_t)w;
            else if (wbytes == 2) ((uint16_t*)wgt)[slot] = (uint16_t)w;
            else ((uint32_t*)wgt)[slot] = w;
        }
    }
}

__global__ void k_max_u32(const uint32_t* a, uint64_t n, unsigned int* out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) m = a[i] > m ? a[i] : m;
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t y = __shfl_xor(m, o);
        m = y > m ? y : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

extern "C" int ipr_hip_build_gather_tiles(uint64_t V, uint64_t E, const uint64_t* t_idx, const uint32_t* t_adj,
                                          const uint32_t* t_wgt, const uint32_t* bpos, uint32_t max_recv,
                                          uint64_t max_arcs, ipr_gather_tiles* out, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    memset(out, 0, sizeof(*out));
    if (V >= (1ull << 31) || max_recv == 0 || max_recv >= IPR_TQ_NONE || max_arcs == 0) {
        snprintf(g_err, sizeof(g_err), "gather tiles: needs V < 2^31, 1 <= receivers per block < 65535 and "
                 "arcs per tile >= 1 (V=%llu receivers=%u arcs=%llu)", (unsigned long long)V, max_recv,
                 (unsigned long long)max_arcs);
        return -1;
    }
    if (E == 0 || V == 0) return 0;
    int rc = 0;
    hipError_t e;
    uint64_t* hidx = (uint64_t*)malloc((V + 1) * 8);
    uint32_t *hv0 = nullptr, *hnr = nullptr, *hna = nullptr, *dv0 = nullptr, *dnr = nullptr, *dna = nullptr;
    uint32_t *pin = nullptr, *pout = nullptr, *dq = nullptr, *dhead = nullptr, *desc = nullptr;
    uint64_t *ha0 = nullptr, *hpo = nullptr, *da0 = nullptr, *dpo = nullptr, *kin = nullptr, *kout = nullptr;
    unsigned long long* dctr = nullptr;
    unsigned long long hctr[2] = {0, 0};
    ipr_tblock* ht = nullptr;
    void *dt = nullptr, *dw = nullptr, *tmp = nullptr;
    size_t tmp_bytes = 0;
    uint64_t nb = 0, nt = 0, slots = 0, ngroups = 0;
    unsigned kshift = 0, end_bit = 0;
    uint32_t wbytes = 0;
    const uint64_t gpt = max_arcs / 64 ? max_arcs / 64 : 1;   /* groups per tile of a split receiver */
    if (!hidx) { snprintf(g_err, sizeof(g_err), "gather tiles: host memory"); return -1; }
    if ((e = hipMemcpyAsync(hidx, t_idx, (V + 1) * 8, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess) { rc = fail(e, "gather tiles offsets"); goto out; }
    /* blocks of consecutive receivers: at most max_recv of them and at most
     * max_arcs in-arcs, except a receiver with more in-arcs, which forms a
     * block alone; blocks without arcs are left out. Each block's arcs start
     * on a group of 64 slots. */
    hv0 = (uint32_t*)malloc(V * 4);
    hnr = (uint32_t*)malloc(V * 4);
    hna = (uint32_t*)malloc(V * 4);
    ha0 = (uint64_t*)malloc(V * 8);
    hpo = (uint64_t*)malloc(V * 8);
    if (!hv0 || !hnr || !hna || !ha0 || !hpo) {
        snprintf(g_err, sizeof(g_err), "gather tiles: host memory");
        rc = -1;
        goto out;
    }
    for (uint64_t v = 0; v < V;) {
        const uint64_t v0 = v, a0 = hidx[v];
        while (v < V && v - v0 < max_recv && (v == v0 || hidx[v + 1] - a0 <= max_arcs)) v++;
        const uint64_t na = hidx[v] - a0;
        if (na > 0) {
            if (na >= (1ull << 32)) {
                snprintf(g_err, sizeof(g_err), "gather tiles: a receiver with %llu in-arcs", (unsigned long long)na);
                rc = -1;
                goto out;
            }
            const uint64_t ng = (na + 63) / 64;
            hv0[nb] = (uint32_t)v0;
            hnr[nb] = (uint32_t)(v - v0);
            hna[nb] = (uint32_t)na;
            ha0[nb] = a0;
            hpo[nb] = slots;
            slots += 64 * ng;
            nt += na > max_arcs ? (ng + gpt - 1) / gpt : 1;
            nb++;
        }
    }
    ngroups = slots / 64;
    kshift = 16 + (unsigned)nbits(V - 1);
    end_bit = kshift + (unsigned)nbits(nb - 1);
    if (end_bit > 64 || nt >= (1ull << 31)) {
        snprintf(g_err, sizeof(g_err), "gather tiles: %llu blocks do not fit the sort key",
                 (unsigned long long)nb);
        rc = -1;
        goto out;
    }
    /* tiles, largest first: a block above max_arcs (one receiver) is split
     * into tiles of whole groups that deliver by atomics */
    ht = (ipr_tblock*)malloc(nt * sizeof(ipr_tblock));
    if (!ht) { snprintf(g_err, sizeof(g_err), "gather tiles: host memory"); rc = -1; goto out; }
    {
        uint64_t k = 0;
        for (uint64_t b = 0; b < nb; b++) {
            const uint64_t ns = 64 * ((hna[b] + 63ull) / 64);
            if (hna[b] <= max_arcs) {
                ht[k++] = ipr_tblock{hpo[b], (uint32_t)ns, hv0[b], hnr[b], 0u};
                continue;
            }
            for (uint64_t o = 0; o < ns; o += 64 * gpt) {
                const uint64_t n = ns - o < 64 * gpt ? ns - o : 64 * gpt;
                ht[k++] = ipr_tblock{hpo[b] + o, (uint32_t)n, hv0[b], hnr[b], 1u};
            }
        }
        std::stable_sort(ht, ht + nt, [](const ipr_tblock& x, const ipr_tblock& y) { return x.na > y.na; });
    }
    if ((e = hipMalloc(&dv0, nb * 4)) != hipSuccess || (e = hipMalloc(&dnr, nb * 4)) != hipSuccess ||
        (e = hipMalloc(&dna, nb * 4)) != hipSuccess || (e = hipMalloc(&da0, nb * 8)) != hipSuccess ||
        (e = hipMalloc(&dpo, nb * 8)) != hipSuccess || (e = hipMalloc(&dctr, 16)) != hipSuccess) {
        rc = fail(e, "gather tile blocks");
        goto out;
    }
    if ((e = hipMalloc(&dt, nt * sizeof(ipr_tblock))) != hipSuccess) { rc = fail(e, "gather tiles"); goto out; }
    if ((e = hipMalloc(&kin, E * 8)) != hipSuccess) { rc = fail(e, "gather tile keys"); goto out; }
    if ((e = hipMalloc(&kout, E * 8)) != hipSuccess) { rc = fail(e, "gather tile keys"); goto out; }
    if ((e = hipMemcpyAsync(dv0, hv0, nb * 4, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(dnr, hnr, nb * 4, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(dna, hna, nb * 4, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(da0, ha0, nb * 8, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(dpo, hpo, nb * 8, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemsetAsync(dctr, 0, 16, st)) != hipSuccess ||
        (e = hipMemcpyAsync(dt, ht, nt * sizeof(ipr_tblock), hipMemcpyHostToDevice, st)) != hipSuccess) {
        rc = fail(e, "gather tile upload");
        goto out;
    }
    hipLaunchKernelGGL(k_tile_keys, dim3((unsigned)(nb < 8192 ? nb : 8192)), dim3(256), 0, st, t_idx, t_adj, bpos,
                       dv0, dnr, nb, kshift, kin);
    if (t_wgt) {
        if ((e = hipMalloc(&pin, E * 4)) != hipSuccess) { rc = fail(e, "gather tile perm"); goto out; }
        if ((e = hipMalloc(&pout, E * 4)) != hipSuccess) { rc = fail(e, "gather tile perm"); goto out; }
        hipLaunchKernelGGL(k_iota_u32, dim3(4096), dim3(256), 0, st, pin, E);
        e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, pin, pout, (size_t)E, 0, end_bit, st);
    } else {
        e = rocprim::radix_sort_keys(nullptr, tmp_bytes, kin, kout, (size_t)E, 0, end_bit, st);
    }
    if (e != hipSuccess) { rc = fail(e, "gather tile sort size"); goto out; }
    if ((e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 16)) != hipSuccess) { rc = fail(e, "gather tile temp"); goto out; }
    if (t_wgt)
        e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, pin, pout, (size_t)E, 0, end_bit, st);
    else
        e = rocprim::radix_sort_keys(tmp, tmp_bytes, kin, kout, (size_t)E, 0, end_bit, st);
    if (e != hipSuccess) { rc = fail(e, "gather tile sort"); goto out; }
    (void)hipFree(kin); kin = nullptr;
    (void)hipFree(pin); pin = nullptr;
    (void)hipFree(tmp); tmp = nullptr;
    /* weights at the narrowest width that holds the largest */
    if (t_wgt) {
        unsigned int* dmax = (unsigned int*)(dctr + 1);
        uint32_t wmax = 0;
        hipLaunchKernelGGL(k_max_u32, dim3(1024), dim3(256), 0, st, t_wgt, E, dmax);
        if ((e = hipMemcpyAsync(&wmax, dmax, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess) { rc = fail(e, "gather tile weights"); goto out; }
        wbytes = wmax < 256 ? 1 : (wmax < 65536 ? 2 : 4);
        if ((e = hipMalloc(&dw, slots * wbytes)) != hipSuccess) { rc = fail(e, "gather tile weights"); goto out; }
    }
    if ((e = hipMalloc(&dq, slots * 4)) != hipSuccess) { rc = fail(e, "gather tile slots"); goto out; }
    if ((e = hipMalloc(&dhead, ngroups * 4)) != hipSuccess) { rc = fail(e, "gather tile groups"); goto out; }
    if ((e = hipMemsetAsync(dq, 0xFF, slots * 4, st)) != hipSuccess) { rc = fail(e, "gather tile slots"); goto out; }
    if ((e = hipMemsetAsync(dctr, 0, 8, st)) != hipSuccess) { rc = fail(e, "gather tile groups"); goto out; }
    {
        uint64_t gb = (ngroups + 255) / 256;
        if (gb > 65536) gb = 65536;
        hipLaunchKernelGGL(k_tile_heads, dim3((unsigned)gb), dim3(256), 0, st, ngroups, kout, dpo, da0, dna, nb,
                           (1ull << (kshift - 16)) - 1, dhead, dctr);
        if ((e = hipMemcpyAsync(&hctr[0], dctr, 8, hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess) { rc = fail(e, "gather tile groups"); goto out; }
        if (hctr[0]) {
            if ((e = hipMalloc(&desc, hctr[0] * 64 * 4)) != hipSuccess) { rc = fail(e, "gather tile escapes"); goto out; }
            if ((e = hipMemsetAsync(dctr, 0, 8, st)) != hipSuccess) { rc = fail(e, "gather tile escapes"); goto out; }
            hipLaunchKernelGGL(k_tile_escapes, dim3((unsigned)gb), dim3(256), 0, st, ngroups, dhead, dctr);
        }
        uint64_t eb = (E + 255) / 256;
        if (eb > 65536) eb = 65536;
        hipLaunchKernelGGL(k_tile_slots, dim3((unsigned)eb), dim3(256), 0, st, E, kout, pout, t_wgt, dpo, da0,
                           kshift, (1ull << (kshift - 16)) - 1, dhead, dq, desc, dw, wbytes);
    }
    if ((e = hipGetLastError()) != hipSuccess) { rc = fail(e, "gather tile index"); goto out; }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) { rc = fail(e, "gather tile sync"); goto out; }
    out->tiles = dt;
    out->n_tiles = (uint32_t)nt;
    out->wgt_bytes = wbytes;
    out->slots = slots;
    out->escaped = hctr[0];
    out->q = dq;
    out->head = dhead;
    out->esc = desc;
    out->wgt = dw;
    dt = nullptr; dq = nullptr; dhead = nullptr; desc = nullptr; dw = nullptr;
out:
    free(hidx); free(hv0); free(hnr); free(hna); free(ha0); free(hpo); free(ht);
    (void)hipFree(dv0); (void)hipFree(dnr); (void)hipFree(dna); (void)hipFree(da0); (void)hipFree(dpo);
    (void)hipFree(dctr); (void)hipFree(kin); (void)hipFree(kout); (void)hipFree(pin); (void)hipFree(pout);
    (void)hipFree(tmp); (void)hipFree(dt); (void)hipFree(dw); (void)hipFree(dq); (void)hipFree(dhead);
    (void)hipFree(desc);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* in-degree bins                                                            */
/* ------------------------------------------------------------------------ */

/* One wavefront per 64 consecutive vertices and one atomic per wavefront and
 * bin, so every bin is a sequence of ascending-id runs (coalesced writes in
 * the kernels that walk it). */
__global__ void k_pull_bins(uint64_t V, const uint64_t* in_idx, uint32_t small_max, uint32_t medium_max,
                            uint32_t* b0, uint32_t* b1, uint32_t* b2, unsigned long long* cnt)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t Vr = (V + 63) & ~63ull;
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < Vr; v += stride) {
        int b = -1;
        if (v < V) {
            uint64_t deg = in_idx[v + 1] - in_idx[v];
            b = deg <= small_max ? 0 : (deg <= medium_max ? 1 : 2);
        }
        for (int k = 0; k < 3; k++) {
            uint64_t mask = __ballot(b == k);
            if (!mask) continue;
            uint32_t leader = __ffsll((unsigned long long)mask) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&cnt[k], (unsigned long long)__popcll(mask));
            base = __shfl(base, leader);
            uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
            uint32_t* dst = k == 0 ? b0 : (k == 1 ? b1 : b2);
            if (b == k) dst[base + below] = (uint32_t)v;
        }
    }
}

extern "C" int ipr_hip_build_pull_bins(ipr_dev* dev, uint32_t small_max, uint32_t medium_max, void* stream)
{
    /* pull variants bin their in-lists, push variants their gather transpose */
    const uint64_t* in_idx = dev->t_idx ? dev->t_idx : dev->in_idx;
    hipStream_t st = (hipStream_t)stream;
    unsigned long long* cnt = nullptr;
    unsigned long long h[3] = {0, 0, 0};
    CK(hipMalloc(&cnt, sizeof(h)));
    hipError_t e = hipMemsetAsync(cnt, 0, sizeof(h), st);
    if (e == hipSuccess) {
        uint64_t blocks = (dev->V + 255) / 256;
        if (blocks < 1) blocks = 1;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(k_pull_bins, dim3((unsigned)blocks), dim3(256), 0, st, dev->V, in_idx,
                           small_max, medium_max, dev->pbins[0], dev->pbins[1], dev->pbins[2], cnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(cnt);
    if (e != hipSuccess) return fail(e, "pull bins");
    for (int b = 0; b < 3; b++) dev->pbin_size[b] = h[b];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* frontier compaction                                                       */
/* ------------------------------------------------------------------------ */

/* The received flags of superstep s+1 become the frontier list. A workgroup
 * takes T rounds of 256 consecutive 16-flag groups (coalesced loads; the
 * flags are cleared for reuse two supersteps on). Positions follow (round,
 * thread) order, which is ascending vertex order: one workgroup scan per
 * round and one atomic per workgroup on the list counter. Each round's
 * entries are staged in LDS and written out by the whole workgroup, 64
 * consecutive entries per wavefront store. T is 4, or less for graphs that
 * would otherwise get under 256 workgroups. */
template <int T>
__global__ void __launch_bounds__(256) k_compact(ipr_dev d, int fin_variant, int relist)
{
    __shared__ uint32_t wsum[T][4];
    __shared__ uint32_t stage[256 * 16];
    __shared__ unsigned long long blk_base;
    ipr_ctl* c = d.ctl;
    /* a fused sparse kernel that stopped (bail) leaves its superstep to the
     * general kernels: no compaction either */
    if (c->done || c->bail) return;
    const uint64_t s = c->superstep;
    /* relist: the current superstep's frontier, left as flags by a fused
     * sparse kernel, into its list (its size is already known; the host
     * zeroes the counter first) */
    const uint64_t par = relist ? (s & 1) : ((s + 1) & 1);
    uint4* flags = (uint4*)d.recv[par];
    uint32_t* out = d.front[par];
    unsigned long long* counter = (unsigned long long*)&c->fsize[par];
    const uint64_t groups = (d.V + 15) / 16;           /* flag arrays are padded to 16 */
    /* with the fused sparse kernel on, after a superstep that did not gather:
     * out-arcs of the new frontier (sum and largest out-degree). After a
     * dense superstep they are not read (the end of the superstep marks them
     * unknown), which spares a read per entry of a large frontier. */
    uint64_t f_arcs = 0, f_max = 0;
    bool want_f = d.fused && !relist;
    if (want_f && d.dir_opt) {
        uint64_t w = d.cnt[IPR_CSLOT(IPR_C_EDGES, threadIdx.x & 63)];
        for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
        want_f = w * (uint64_t)d.dir_alpha <= d.E;
    }
    const uint64_t span = 256ull * T;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    /* the last superstep of a stamp window clears the 4-byte gather records
     * (padded to 16 like the flags) */
    uint4* clr = !relist && d.brec4 && (s + 1) % IPR_REC4_WINDOW == 0 ? (uint4*)d.brec4 : nullptr;
    if (relist && blockIdx.x == 0 && threadIdx.x == 0) c->front_flags = 0;   /* the list is the frontier */
    /* after a superstep that gathered (SPREAD): count the raised flags and
     * leave them for the next superstep's kernels, which read the frontier
     * from them (no list, no staging, no list-counter atomic) */
    if (d.dense_flags && !relist) {
        uint64_t w = d.cnt[IPR_CSLOT(IPR_C_EDGES, threadIdx.x & 63)];
        for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
        if (d.dir_opt && w * (uint64_t)d.dir_alpha > d.E) {
            uint32_t cnt = 0;
            for (uint64_t b0 = (uint64_t)blockIdx.x * span; b0 < groups; b0 += (uint64_t)gridDim.x * span) {
                uint4 q[T];
#pragma unroll
                for (int t = 0; t < T; t++) {
                    const uint64_t g = b0 + (uint64_t)t * 256 + threadIdx.x;
                    q[t] = make_uint4(0, 0, 0, 0);
                    if (g < groups) q[t] = flags[g];
                    if (clr && g < groups) {
                        const uint4 z = make_uint4(0, 0, 0, 0);
                        clr[4 * g] = z;
                        clr[4 * g + 1] = z;
                        clr[4 * g + 2] = z;
                        clr[4 * g + 3] = z;
                    }
                }
#pragma unroll
                for (int t = 0; t < T; t++) cnt += __popc(q[t].x) + __popc(q[t].y) + __popc(q[t].z) + __popc(q[t].w);
            }
            __shared__ uint32_t s_cnt[4];
            for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
            if (lane == 0) s_cnt[wave] = cnt;
            __syncthreads();
            if (threadIdx.x == 0) {
                const uint64_t tot = (uint64_t)s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
                if (tot) atomicAdd((unsigned long long*)&d.cnt[IPR_CSLOT(IPR_C_FNEXT, blockIdx.x % IPR_CSLOTS)],
                                   (unsigned long long)tot);
            }
            if (fin_variant >= 0 && ipr_last_workgroup(d, gridDim.x) && threadIdx.x < 64)
                ipr_finalize_step(d, fin_variant, 1);
            return;
        }
    }
    for (uint64_t b0 = (uint64_t)blockIdx.x * span; b0 < groups; b0 += (uint64_t)gridDim.x * span) {
        uint4 q[T];
        uint32_t n[T], incl[T];
#pragma unroll
        for (int t = 0; t < T; t++) {
            const uint64_t g = b0 + (uint64_t)t * 256 + threadIdx.x;
            q[t] = make_uint4(0, 0, 0, 0);
            if (g < groups) q[t] = flags[g];
            if (clr && g < groups) {
                const uint4 z = make_uint4(0, 0, 0, 0);
                clr[4 * g] = z;
                clr[4 * g + 1] = z;
                clr[4 * g + 2] = z;
                clr[4 * g + 3] = z;
            }
        }
#pragma unroll
        for (int t = 0; t < T; t++) {
            /* flags are 0 or 1, so each set flag is one bit (0, 8, 16 or 24) */
            n[t] = __popc(q[t].x) + __popc(q[t].y) + __popc(q[t].z) + __popc(q[t].w);
            if (n[t]) flags[b0 + (uint64_t)t * 256 + threadIdx.x] = make_uint4(0, 0, 0, 0);
            uint32_t x = n[t];
            for (int o = 1; o < 64; o <<= 1) {
                uint32_t y = __shfl_up(x, o);
                if (lane >= (uint32_t)o) x += y;
            }
            incl[t] = x;
            if (lane == 63) wsum[t][wave] = x;
        }
        __syncthreads();
        uint32_t total = 0;
        uint32_t first[T], cnt[T], before[T];
#pragma unroll
        for (int t = 0; t < T; t++) {
            first[t] = total;
            uint32_t b = 0;
            for (uint32_t w = 0; w < 4; w++) {
                if (w < wave) b += wsum[t][w];
                total += wsum[t][w];
            }
            before[t] = b;
            cnt[t] = total - first[t];
        }
        if (threadIdx.x == 0 && total) blk_base = atomicAdd(counter, (unsigned long long)total);
        __syncthreads();
        if (total) {
            const uint64_t base = blk_base;
#pragma unroll
            for (int t = 0; t < T; t++) {
                if (!cnt[t]) continue;   /* uniform across the workgroup */
                uint32_t pos = before[t] + incl[t] - n[t];
                const uint64_t g = b0 + (uint64_t)t * 256 + threadIdx.x;
                const uint32_t words[4] = {q[t].x, q[t].y, q[t].z, q[t].w};
                for (int k = 0; k < 4; k++) {
                    uint32_t w = words[k];
                    while (w) {
                        uint32_t bit = __ffs(w) - 1;
                        const uint32_t id = (uint32_t)(g * 16 + k * 4 + bit / 8);
                        stage[pos++] = id;
                        w &= w - 1;
                    }
                }
                __syncthreads();
                /* coalesced write-out; the out-degrees of consecutive entries
                 * are read here, by consecutive threads */
                for (uint32_t i = threadIdx.x; i < cnt[t]; i += 256) {
                    const uint32_t id = stage[i];
                    out[base + first[t] + i] = id;
                    if (want_f) {
                        const uint64_t dg = d.out_idx[id + 1] - d.out_idx[id];
                        f_arcs += dg;
                        f_max = dg > f_max ? dg : f_max;
                    }
                }
                __syncthreads();
            }
        }
    }
    if (want_f) {
        /* one atomic per workgroup and figure */
        __shared__ uint64_t r_arcs[4], r_max[4];
        for (int o = 32; o > 0; o >>= 1) {
            f_arcs += __shfl_xor(f_arcs, o);
            const uint64_t m = __shfl_xor(f_max, o);
            f_max = m > f_max ? m : f_max;
        }
        if (lane == 0) { r_arcs[wave] = f_arcs; r_max[wave] = f_max; }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t a = 0, m = 0;
            for (int w = 0; w < 4; w++) { a += r_arcs[w]; m = r_max[w] > m ? r_max[w] : m; }
            /* spread over the 64 slots; the end of the superstep collapses
             * them into ctl->farcs / fmaxdeg */
            const uint32_t slot = blockIdx.x % IPR_CSLOTS;
            if (a) atomicAdd((unsigned long long*)&d.cnt[IPR_CSLOT(IPR_C_FARCS, slot)], (unsigned long long)a);
            if (m) atomicMax((unsigned long long*)&d.cnt[IPR_CSLOT(IPR_C_FMAX, slot)], (unsigned long long)m);
        }
    }
    /* after a fused sparse superstep: the last workgroup ends the superstep
     * (every workgroup counts, so none reads the control block after it) */
    if (fin_variant >= 0 && !relist && ipr_last_workgroup(d, gridDim.x) && threadIdx.x < 64)
        ipr_finalize_step(d, fin_variant);
}

extern "C" int ipr_hip_compact_frontier(const ipr_dev* dev, int finalize_variant, int relist, void* stream)
{
    const uint64_t groups = (dev->V + 15) / 16;
    /* rounds per workgroup: 4, or fewer where that leaves under 256
     * workgroups (fewer workgroups: fewer atomics on the one list counter,
     * longer serial work; 4 measured best at V = 4M and 16M) */
    int T = 4;
    while (T > 1 && groups < 256ull * 256 * T) T >>= 1;
    if (dev->compact_rounds == 1 || dev->compact_rounds == 2 || dev->compact_rounds == 4 || dev->compact_rounds == 8)
        T = (int)dev->compact_rounds;
    const uint64_t span = 256ull * T;
    uint64_t blocks = (groups + span - 1) / span;
    if (blocks < 1) blocks = 1;
    if (blocks > 16384) blocks = 16384;
    const dim3 grid((unsigned)blocks), block(256);
    hipStream_t st = (hipStream_t)stream;
    switch (T) {
    case 8: hipLaunchKernelGGL(k_compact<8>, grid, block, 0, st, *dev, finalize_variant, relist); break;
    case 4: hipLaunchKernelGGL(k_compact<4>, grid, block, 0, st, *dev, finalize_variant, relist); break;
    case 2: hipLaunchKernelGGL(k_compact<2>, grid, block, 0, st, *dev, finalize_variant, relist); break;
    default: hipLaunchKernelGGL(k_compact<1>, grid, block, 0, st, *dev, finalize_variant, relist); break;
    }
    CK(hipGetLastError());
    return 0;
}

__global__ void k_differ(const uint32_t* a, const uint32_t* b, uint64_t n, unsigned int* flag)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        if (a[i] != b[i]) { atomicOr(flag, 1u); return; }
}

extern "C" int ipr_hip_equal(const void* a, const void* b, uint64_t bytes, int* equal, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    unsigned int* flag = nullptr;
    unsigned int h = 0;
    CK(hipMalloc(&flag, sizeof(unsigned int)));
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(unsigned int), st);
    if (e == hipSuccess && bytes >= 4) {
        uint64_t n = bytes / 4;
        uint64_t blocks = (n + 255) / 256;
        if (blocks > 16384) blocks = 16384;
        hipLaunchKernelGGL(k_differ, dim3((unsigned)blocks), dim3(256), 0, st, (const uint32_t*)a,
                           (const uint32_t*)b, n, flag);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(flag);
    if (e != hipSuccess) return fail(e, "equality check");
    *equal = h == 0;
    return 0;
}

__global__ void k_clear_u32(unsigned int* p, uint32_t n)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

/* ipr_hip_equal without an allocation: the difference flag is the first word
 * of this device's `scratch`, cleared by a kernel (not a memset node) before
 * the comparison. Calls that overlap on one GPU share the flag and race. */
extern "C" int ipr_hip_equal_noalloc(const void* a, const void* b, uint64_t bytes, int* equal, void* stream)
{
    if (!equal || (bytes && (!a || !b))) {
        snprintf(g_err, sizeof(g_err), "ipr_hip_equal_noalloc: null argument");
        return -(int)hipErrorInvalidValue;
    }
    if (bytes % 4) {
        snprintf(g_err, sizeof(g_err), "ipr_hip_equal_noalloc: %llu bytes is not a multiple of 4",
                 (unsigned long long)bytes);
        return -(int)hipErrorInvalidValue;
    }
    unsigned int* flag = nullptr;
    int rc = scratch_of_current_device(&flag);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    unsigned int h = 0;
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, st, flag, 1u);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && bytes >= 4) {
        const uint64_t n = bytes / 4;
        uint64_t blocks = (n + 255) / 256;
        if (blocks > 16384) blocks = 16384;
        hipLaunchKernelGGL(k_differ, dim3((unsigned)blocks), dim3(256), 0, st, (const uint32_t*)a,
                           (const uint32_t*)b, n, flag);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(e, "equality check (no allocation)");
    *equal = h == 0;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* superstep control                                                         */
/* ------------------------------------------------------------------------ */

__global__ void k_begin(ipr_dev d)
{
    /* the counters' slot words (the rest of each slot's line is never used) */
    for (uint32_t i = threadIdx.x; i < IPR_NCNT * IPR_CSLOTS; i += blockDim.x) d.cnt[(uint64_t)i * IPR_CPAD] = 0;
    if (threadIdx.x != 0) return;
    ipr_ctl* c = d.ctl;
    uint64_t cap = c->trace_cap;
    memset(c, 0, sizeof(ipr_ctl));
    c->trace_cap = cap;
    if (cap) d.trace_ts[0] = wall_clock64();
}

/* End of one superstep (ipregel_control.h): one wavefront. Launched after
 * the superstep's kernels when none of them ends the superstep itself. */
__global__ void k_finalize(ipr_dev d, int variant)
{
    ipr_finalize_step(d, variant);
}

extern "C" int ipr_hip_begin(const ipr_dev* dev, void* stream)
{
    hipLaunchKernelGGL(k_begin, dim3(1), dim3(256), 0, (hipStream_t)stream, *dev);
    CK(hipGetLastError());
    return 0;
}

extern "C" int ipr_hip_finalize(const ipr_dev* dev, int variant, void* stream)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(64), 0, (hipStream_t)stream, *dev, variant);
    CK(hipGetLastError());
    return 0;
}

