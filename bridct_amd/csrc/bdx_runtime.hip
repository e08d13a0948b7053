// Thin C ABI of the HIP side (declared in bridct_device.h): device checks, memory, streams
// and the route choice of one batched transform.
//
// Route per shape (the GPU replacement of rect_policy.inc / the square-variant table of
// src/bridct.c:47-74):
//   * h * (w + 1) floats fit 144 KiB of LDS: one whole-array kernel (bdx_small.hip);
//   * otherwise: row pass (input -> output) then column pass in place on the output
//     (bdx_passes.hip); the fused round trip adds a third, inverse row pass.  Calls whose
//     intermediate is at least twice the Infinity Cache run these passes per chunk of
//     arrays (BDX_CHUNK_MIN_BYTES).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <atomic>

#include "bdx_kernels.h"
#include "bdx_lines.h"
#include "bridct_device.h"

namespace {

int small_dispatch(int h, int w, const float *in, float *out, size_t batch, int pack, int mode,
                   const float *tab, hipStream_t s, const size_t *offs = nullptr)
{
    using namespace bdx;
    pack = pack < 0 ? 0 : pack;   // the two-pass value on a shape that only fits this kernel
    switch (h) {
    case 8: return launch_small_h8(w, in, out, batch, pack, mode, tab, s, offs);
    case 16: return launch_small_h16(w, in, out, batch, pack, mode, tab, s, offs);
    case 32: return launch_small_h32(w, in, out, batch, pack, mode, tab, s, offs);
    case 64: return launch_small_h64(w, in, out, batch, pack, mode, tab, s, offs);
    case 128: return launch_small_h128(w, in, out, batch, pack, mode, tab, s, offs);
    case 256: return launch_small_h256(w, in, out, batch, pack, mode, tab, s, offs);
    case 512: return launch_small_h512(w, in, out, batch, pack, mode, tab, s, offs);
    case 1024: return launch_small_h1024(w, in, out, batch, pack, mode, tab, s, offs);
    default: return -1;
    }
}

// Device copy with streaming float4 accesses: the practical HBM ceiling that the transform
// kernels are compared with (a read and a write per element, like one pass).
__global__ __launch_bounds__(256) void k_copy_stream(const float *__restrict__ a, float *__restrict__ b, size_t n4)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256)
        bdx::st4<true>(b + 4 * i, bdx::ld4<true>(a + 4 * i));
}

__global__ void k_zero(unsigned *p, unsigned n)
{
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = 0u;
}

// Failure injection for the tests of the allocation-failure paths, and a count of the
// runtime resources the library holds (device and pinned buffers, streams, events, graph
// executables).
// fault_after < 0: off; k >= 0: k more acquisitions succeed, the next one fails, then off.
std::atomic<long> fault_after{-1};
std::atomic<long> live_resources{0};
// Workgroups per launch of the kernels that split large grids (0: the hardware limit).
std::atomic<unsigned> grid_limit{0};

// True when this acquisition must fail.
bool injected_fault()
{
    long n = fault_after.load(std::memory_order_relaxed);
    while (n >= 0)
        if (fault_after.compare_exchange_weak(n, n - 1, std::memory_order_relaxed))
            return n == 0;
    return false;
}

int acquired(bool ok)
{
    if (!ok)
        return -1;
    live_resources.fetch_add(1, std::memory_order_relaxed);
    return 0;
}

int released(void *handle, bool ok)
{
    if (handle && ok)
        live_resources.fetch_sub(1, std::memory_order_relaxed);
    return ok ? 0 : -1;
}

}  // namespace

extern "C" {

long bdx_debug_alloc_faults(long after)
{
    fault_after.store(after < 0 ? -1 : after, std::memory_order_relaxed);
    return live_resources.load(std::memory_order_relaxed);
}

long bdx_debug_grid_limit(long blocks)
{
    const long prev = (long)grid_limit.load(std::memory_order_relaxed);
    grid_limit.store(blocks > 0 && blocks <= 0x7fffffffL ? (unsigned)blocks : 0u, std::memory_order_relaxed);
    return prev;
}

int bdx_device_ready(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
        return -1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return -1;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return -1;
    return strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 0 : -2;
}

int bdx_current_device(int *device)
{
    return hipGetDevice(device) == hipSuccess ? 0 : -1;
}

int bdx_malloc(void **ptr, size_t bytes)
{
    return acquired(!injected_fault() && hipMalloc(ptr, bytes) == hipSuccess);
}

int bdx_free(void *ptr)
{
    return released(ptr, hipFree(ptr) == hipSuccess);
}

int bdx_host_malloc(void **ptr, size_t bytes)
{
    return acquired(!injected_fault() && hipHostMalloc(ptr, bytes, hipHostMallocDefault) == hipSuccess);
}

int bdx_host_free(void *ptr)
{
    return released(ptr, hipHostFree(ptr) == hipSuccess);
}

int bdx_event_create(void **event)
{
    hipEvent_t e;
    if (acquired(!injected_fault() && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess))
        return -1;
    *event = (void *)e;
    return 0;
}

int bdx_event_destroy(void *event)
{
    return released(event, hipEventDestroy((hipEvent_t)event) == hipSuccess);
}

int bdx_event_record(void *event, void *stream)
{
    return hipEventRecord((hipEvent_t)event, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int bdx_event_sync(void *event)
{
    return hipEventSynchronize((hipEvent_t)event) == hipSuccess ? 0 : -1;
}

int bdx_copy_stream(const float *d_src, float *d_dst, size_t floats, void *stream)
{
    if (floats % 4 || ((uintptr_t)d_src | (uintptr_t)d_dst) % 16)
        return -1;
    const size_t n4 = floats / 4, blocks = (n4 + 255) / 256;
    hipLaunchKernelGGL(k_copy_stream, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0,
                       (hipStream_t)stream, d_src, d_dst, n4);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int bdx_stream_wait_event(void *stream, void *event)
{
    return hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0) == hipSuccess ? 0 : -1;
}

int bdx_stream_create(void **stream)
{
    hipStream_t s;
    if (acquired(!injected_fault() && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess))
        return -1;
    *stream = (void *)s;
    (void)bdx::xcd_count();   // read here, before any launch a caller may capture into a graph
    return 0;
}

int bdx_stream_destroy(void *stream)
{
    return released(stream, hipStreamDestroy((hipStream_t)stream) == hipSuccess);
}

int bdx_stream_sync(void *stream)
{
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int bdx_copy_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int bdx_copy_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int bdx_copy_d2d(void *dst, const void *src, size_t bytes, void *stream)
{
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int bdx_device_cus(void)
{
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    return n;
}

int bdx_batch_launch(const float *d_in, float *d_out, const bdx_run *d_runs, const bdx_item *d_items,
                     unsigned nitems, const bdx_unit *d_units, unsigned nunits, unsigned qbase,
                     unsigned *d_counters, unsigned ncounters, int mode, const float *d_tables, unsigned blocks,
                     unsigned spin_limit, int flags, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    // a fill kernel, not a memset node, so that the sequence can be captured in a graph; the
    // tiny kernel uses no counters (the caller zeroed them once, for the error flag)
    if (!(flags & BDX_LAUNCH_TINY))
        hipLaunchKernelGGL(k_zero, dim3((ncounters + 255) / 256), dim3(256), 0, s, d_counters, ncounters);
    if (hipGetLastError() != hipSuccess || nitems == 0)
        return nitems == 0 ? 0 : -1;
    if (!(flags & BDX_LAUNCH_TINY) && (qbase % 2 || qbase + BDX_QWORDS(nunits) > ncounters))
        return -1;
    switch (mode) {
    case BDX_FORWARD:
        return bdx::launch_batch_m0(flags, d_in, d_out, d_runs, d_items, nitems, d_units, nunits, qbase, d_counters,
                                    d_tables, blocks, spin_limit, s);
    case BDX_INVERSE:
        return bdx::launch_batch_m1(flags, d_in, d_out, d_runs, d_items, nitems, d_units, nunits, qbase, d_counters,
                                    d_tables, blocks, spin_limit, s);
    case BDX_ROUNDTRIP:
        return bdx::launch_batch_m2(flags, d_in, d_out, d_runs, d_items, nitems, d_units, nunits, qbase, d_counters,
                                    d_tables, blocks, spin_limit, s);
    default:
        return -1;
    }
}

int bdx_graph_capture(void *stream, int (*fn)(void *), void *arg, void **graph_exec)
{
    hipStream_t s = (hipStream_t)stream;
    hipGraph_t graph = nullptr;
    if (injected_fault() || hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return -1;
    const int rc = fn(arg);
    const hipError_t end = hipStreamEndCapture(s, &graph);
    if (rc || end != hipSuccess || !graph) {
        if (graph)
            (void)hipGraphDestroy(graph);
        return -1;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t inst = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (acquired(inst == hipSuccess))
        return -1;
    *graph_exec = (void *)exec;
    return 0;
}

int bdx_graph_launch(void *graph_exec, void *stream)
{
    return hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int bdx_graph_destroy(void *graph_exec)
{
    return released(graph_exec, hipGraphExecDestroy((hipGraphExec_t)graph_exec) == hipSuccess);
}

int bdx_transform_pass(const float *d_in, float *d_out, int h, int w, size_t batch, int pack, int mode,
                       int pass, const float *d_tables, void *stream)
{
    if (batch == 0)
        return 0;
    if ((mode != BDX_FORWARD && mode != BDX_INVERSE) || (pass != 0 && pass != 1))
        return -1;
    hipStream_t s = (hipStream_t)stream;
    if (BDX_WHOLE_ARRAY(h, w, pack))
        return pass == 0 ? small_dispatch(h, w, d_in, d_out, batch, pack, mode, d_tables, s) : 0;
    const bool nt = 2 * batch * (size_t)h * w * sizeof(float) > BDX_STREAM_MIN_BYTES;
    if (pass == 0)
        return bdx::launch_rows(w, d_in, d_out, batch * (size_t)h, mode == BDX_FORWARD ? bdx::FA : bdx::IA,
                                d_tables, s, nullptr, 0, nt);
    return bdx::launch_cols(h, w, d_in, d_out, batch, mode, d_tables, s, nullptr, nt);
}

static int transform(const float *d_in, float *d_out, int h, int w, size_t batch, int pack, int mode,
                     const float *d_tables, hipStream_t s, const size_t *offs)
{
    if (batch == 0)
        return 0;
    if (mode < BDX_FORWARD || mode > BDX_ROUNDTRIP)
        return -1;
    if (BDX_WHOLE_ARRAY(h, w, pack))
        return small_dispatch(h, w, d_in, d_out, batch, pack, mode, d_tables, s, offs);
    const size_t hw = (size_t)h * w;
    const int rop = mode == BDX_FORWARD ? bdx::FA : mode == BDX_INVERSE ? bdx::IA : bdx::FA2;
    // chunks of whole arrays when the intermediate is far larger than the Infinity Cache;
    // the gather form advances its offset table, the contiguous form its array pointers
    const size_t per = batch * hw * sizeof(float) >= BDX_CHUNK_MIN_BYTES
                           ? std::max<size_t>(1, BDX_CHUNK_BYTES / (hw * sizeof(float)))
                           : batch;
    for (size_t a = 0; a < batch; a += per) {
        const size_t n = std::min(per, batch - a), nrows = n * h;
        const float *in = offs ? d_in : d_in + a * hw;
        float *out = offs ? d_out : d_out + a * hw;
        const size_t *o = offs ? offs + a : nullptr;
        const bool nt = 2 * nrows * (size_t)w * sizeof(float) > BDX_STREAM_MIN_BYTES;
        if (bdx::launch_rows(w, in, out, nrows, rop, d_tables, s, o, h, nt) ||
            bdx::launch_cols(h, w, out, out, n, mode, d_tables, s, o, nt))
            return -1;
        if (mode == BDX_ROUNDTRIP && bdx::launch_rows(w, out, out, nrows, bdx::I0, d_tables, s, o, h, nt))
            return -1;
    }
    return 0;
}

int bdx_transform(const float *d_in, float *d_out, int h, int w, size_t batch, int pack, int mode,
                  const float *d_tables, void *stream)
{
    return transform(d_in, d_out, h, w, batch, pack, mode, d_tables, (hipStream_t)stream, nullptr);
}

int bdx_transform_gather(const float *d_in, float *d_out, int h, int w, size_t count, const size_t *d_offsets,
                         int pack, int mode, const float *d_tables, void *stream)
{
    if (!d_offsets)
        return -1;
    return transform(d_in, d_out, h, w, count, pack, mode, d_tables, (hipStream_t)stream, d_offsets);
}

}  // extern "C"

namespace bdx {

size_t max_grid_blocks(int threads)
{
    // a launch covers at most 2^32 - 1 work-items per grid dimension
    const size_t hw = 0xffffffffu / (unsigned)threads, cap = grid_limit.load(std::memory_order_relaxed);
    return cap && cap < hw ? cap : hw;
}

}  // namespace bdx
