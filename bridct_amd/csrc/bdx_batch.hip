// Persistent batch kernel: ragged batches of arrays of mixed supported shapes in one
// launch.  Compiled once per transform mode BDX_MODE (0 forward, 1 inverse, 2 fused
// round trip).
//
// BDX_BATCH_THREADS-lane workgroups, as many as stay resident, take work items until the
// batch is done.  Items (built on the host, see bdx_item in bridct_device.h) are:
//   SMALL  several whole arrays of one run (small_unit: one HBM read and write per pixel);
//   ROW    a band of rows of a larger array, input -> output (row_unit);
//   COL    a strip of columns of that array, in place on the output (col_unit); it waits
//          until every ROW item of the array has completed;
//   ROW2   (round trip only) the inverse row band, in place, after every COL item.
// Units (bdx_unit): one two-pass array with all its items, or one SMALL item.
//
// Queues per XCD.  A workgroup reads the id of its XCD and draws tickets from that XCD's
// queue.  The tickets of a queue are cut into slots: slot j is the ROW items of the unit
// u_j the queue takes j-th, then the COL items of u_{j-1}, then (round trip) the ROW2
// items of u_{j-2}.  The workgroup that draws the first ticket of slot j takes u_j from the
// global unit counter and publishes the slot (unit and end ticket, one 64-bit word);
// workgroups with later tickets of the slot wait for that.  So every item of an array runs
// on one XCD: the row pass's output stays in that XCD's L2 for the column pass, and a
// dependency needs no cache write-back, only that the producer's stores have reached
// the L2 before its counter moves, and that the consumer's CU drops its L1 lines (a CU may
// hold a line of the array from before the producers wrote it) before reading: an
// agent-scope acquire, no whole-L2 write-back on the producer's side.  Every wait
// is for items of earlier tickets of the same queue, taken by running workgroups, which
// never wait on later ones: the schedule cannot deadlock.  A wait longer than spin_limit
// sleeps sets the error flag and proceeds instead of hanging.  The columns of u_{j-1} are
// drawn after the rows of u_j, so they seldom wait.
#include "bdx_units.h"

#ifndef BDX_MODE
#error "compile with -DBDX_MODE=<0|1|2>"
#endif

namespace bdx {
namespace {

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima): this wave's vector memory
// operations, cache write-backs and invalidations included, have completed.
__device__ __forceinline__ void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0F70); }

constexpr int TH = BDX_BATCH_THREADS;
constexpr int ROW_OP = BDX_MODE == 0 ? FA : BDX_MODE == 1 ? IA : FA2;

// Each unit instantiation is its own (non-inlined) function, so that it keeps its own
// register allocation instead of the union of all paths of the dispatch.
//
// A callee sees its arguments as per-lane values of unknown address space, so the shared
// unit code lost what the kernel knew: that arrays and tables are in global memory and
// the work area in LDS (it emitted FLAT accesses: a 64-bit address per lane, counted by
// both the memory and the LDS wait counters), and that every argument is the same on all
// lanes (it built every address with 64-bit vector arithmetic instead of a scalar base).
// The pointer parameters therefore carry their address space, and each function takes
// its arguments back through readfirstlane (all lanes of a workgroup pass the same values).
typedef const float __attribute__((address_space(1))) *gcptr;
typedef float __attribute__((address_space(1))) *gptr;
typedef float __attribute__((address_space(3))) *lptr;
__device__ __forceinline__ gcptr to_g(const float *p) { return (gcptr)p; }
__device__ __forceinline__ gptr to_g(float *p) { return (gptr)p; }
__device__ __forceinline__ lptr to_l(float *p) { return (lptr)p; }
__device__ __forceinline__ unsigned long long uni64(unsigned long long v)
{
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (unsigned long long)hi << 32 | lo;
}
__device__ __forceinline__ const float *uni(gcptr p) { return (const float *)(gcptr)uni64((unsigned long long)p); }
__device__ __forceinline__ float *uni(gptr p) { return (float *)(gptr)uni64((unsigned long long)p); }
__device__ __forceinline__ float *uni(lptr p)
{
    return (float *)(lptr)(unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((unsigned)(unsigned long long)p);
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ size_t uni(size_t v) { return (size_t)uni64(v); }

template <int H, int W>
__device__ __noinline__ void small_hw(gcptr in, gptr out, size_t off, int count, lptr smem, gcptr __restrict__ tab)
{
    if constexpr (BDX_BATCH_SMALL(H, W)) {
        static_assert(BDX_BATCH_PER(H, W) >= 1, "a whole array must fit the LDS area");
        small_unit<H, W, BDX_MODE, TH>(uni(in), uni(out), uni(off), uni(count), uni(smem), uni(tab));
    }
}

template <int H>
__device__ __forceinline__ void small_h(int w, const float *in, float *out, size_t off, int count, float *smem,
                                        const float *__restrict__ tab)
{
    switch (w) {
    case 8: small_hw<H, 8>(to_g(in), to_g(out), off, count, to_l(smem), to_g(tab)); break;
    case 16: small_hw<H, 16>(to_g(in), to_g(out), off, count, to_l(smem), to_g(tab)); break;
    case 32: small_hw<H, 32>(to_g(in), to_g(out), off, count, to_l(smem), to_g(tab)); break;
    case 64: small_hw<H, 64>(to_g(in), to_g(out), off, count, to_l(smem), to_g(tab)); break;
    case 128: small_hw<H, 128>(to_g(in), to_g(out), off, count, to_l(smem), to_g(tab)); break;
    case 256: small_hw<H, 256>(to_g(in), to_g(out), off, count, to_l(smem), to_g(tab)); break;
    case 512: small_hw<H, 512>(to_g(in), to_g(out), off, count, to_l(smem), to_g(tab)); break;
    default: small_hw<H, 1024>(to_g(in), to_g(out), off, count, to_l(smem), to_g(tab)); break;
    }
}

__device__ __forceinline__ void small_any(int h, int w, const float *in, float *out, size_t off, int count,
                                          float *smem, const float *__restrict__ tab)
{
    switch (h) {
    case 8: small_h<8>(w, in, out, off, count, smem, tab); break;
    case 16: small_h<16>(w, in, out, off, count, smem, tab); break;
    case 32: small_h<32>(w, in, out, off, count, smem, tab); break;
    case 64: small_h<64>(w, in, out, off, count, smem, tab); break;
    case 128: small_h<128>(w, in, out, off, count, smem, tab); break;
    case 256: small_h<256>(w, in, out, off, count, smem, tab); break;
    case 512: small_h<512>(w, in, out, off, count, smem, tab); break;
    default: small_h<1024>(w, in, out, off, count, smem, tab); break;
    }
}

template <int W, int OP>
__device__ __noinline__ void row_w(gcptr in, gptr out, size_t row0, int nrows, lptr smem, gcptr __restrict__ tab)
{
    static_assert(row_unit_lds<W, OP, TH>() <= BDX_BATCH_LDS && row_unit_rows<W, TH>() == BDX_ROW_BAND(W),
                  "row unit geometry");
    row_unit<W, OP, TH>(uni(in), uni(out), uni(row0), uni(nrows), uni(smem), uni(tab), false);
}

template <int OP>
__device__ __forceinline__ void row_any(int w, const float *in, float *out, size_t row0, int nrows, float *smem,
                                        const float *__restrict__ tab)
{
    switch (w) {
    case 8: row_w<8, OP>(to_g(in), to_g(out), row0, nrows, to_l(smem), to_g(tab)); break;
    case 16: row_w<16, OP>(to_g(in), to_g(out), row0, nrows, to_l(smem), to_g(tab)); break;
    case 32: row_w<32, OP>(to_g(in), to_g(out), row0, nrows, to_l(smem), to_g(tab)); break;
    case 64: row_w<64, OP>(to_g(in), to_g(out), row0, nrows, to_l(smem), to_g(tab)); break;
    case 128: row_w<128, OP>(to_g(in), to_g(out), row0, nrows, to_l(smem), to_g(tab)); break;
    case 256: row_w<256, OP>(to_g(in), to_g(out), row0, nrows, to_l(smem), to_g(tab)); break;
    case 512: row_w<512, OP>(to_g(in), to_g(out), row0, nrows, to_l(smem), to_g(tab)); break;
    default: row_w<1024, OP>(to_g(in), to_g(out), row0, nrows, to_l(smem), to_g(tab)); break;
    }
}

template <int H>
__device__ __noinline__ void col_h(gcptr in, gptr out, int w, size_t base, int c0, int ncols, lptr smem,
                                   gcptr __restrict__ tab)
{
    static_assert(col_unit_lds<H, TH>() <= BDX_BATCH_LDS && col_unit_cols<H, TH>() == BDX_COL_STRIP(H),
                  "column unit geometry");
    col_unit<H, BDX_MODE, TH>(uni(in), uni(out), uni(w), uni(base), uni(c0), uni(ncols), uni(smem), uni(tab), false);
}

__device__ __forceinline__ void col_any(int h, int w, const float *in, float *out, size_t base, int c0, int ncols,
                                        float *smem, const float *__restrict__ tab)
{
    switch (h) {
    case 8: col_h<8>(to_g(in), to_g(out), w, base, c0, ncols, to_l(smem), to_g(tab)); break;
    case 16: col_h<16>(to_g(in), to_g(out), w, base, c0, ncols, to_l(smem), to_g(tab)); break;
    case 32: col_h<32>(to_g(in), to_g(out), w, base, c0, ncols, to_l(smem), to_g(tab)); break;
    case 64: col_h<64>(to_g(in), to_g(out), w, base, c0, ncols, to_l(smem), to_g(tab)); break;
    case 128: col_h<128>(to_g(in), to_g(out), w, base, c0, ncols, to_l(smem), to_g(tab)); break;
    case 256: col_h<256>(to_g(in), to_g(out), w, base, c0, ncols, to_l(smem), to_g(tab)); break;
    case 512: col_h<512>(to_g(in), to_g(out), w, base, c0, ncols, to_l(smem), to_g(tab)); break;
    default: col_h<1024>(to_g(in), to_g(out), w, base, c0, ncols, to_l(smem), to_g(tab)); break;
    }
}

// Waits until *p >= target: thread 0 spins, then acquires at agent scope, which invalidates
// this CU's L1 (a CU that ran a column strip of the array holds whole lines of it, the
// neighbouring strips' columns included, from before their producers wrote them; an
// invalidation at workgroup scope left such lines in place), and the block follows through
// the barrier.  The producers are on this XCD, so its L2 already holds their stores.
__device__ __forceinline__ void wait_for(unsigned *p, unsigned target, unsigned *err, unsigned limit)
{
    if (threadIdx.x == 0) {
        unsigned n = 0;
        while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++n > limit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        wait_vm();
    }
    __syncthreads();
}

// Counts the item as done once the block's stores have reached the L2 of this XCD: every
// wave waits for its stores, then thread 0 increments the counter.
__device__ __forceinline__ void signal_done(unsigned *p)
{
    wait_vm();
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Id of the XCD that runs this workgroup (hardware register XCC_ID, bits 3:0).
__device__ __forceinline__ unsigned xcd_id()
{
    unsigned v;
    __asm__ volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return v & (BDX_XQ - 1);
}

// Items of slot j: the rows of u_j, the columns of u_{j-L} and (round trip) the ROW2 items
// of u_{j-2L}, L = BDX_LAG.
__device__ __forceinline__ unsigned slot_items(const bdx_unit *__restrict__ units, unsigned u0, unsigned u1,
                                               unsigned u2)
{
    return (u0 != BDX_UNIT_END ? units[u0].nrow : 0u) + (u1 != BDX_UNIT_END ? units[u1].ncol : 0u) +
           (BDX_MODE == 2 && u2 != BDX_UNIT_END ? units[u2].nrow2 : 0u);
}

// Units of the last K + 1 slots of a queue walk, newest first (K = 2 BDX_LAG).
struct SlotHistory {
    static constexpr int K = 2 * BDX_LAG;
    unsigned u[K + 1];
    __device__ __forceinline__ void init()
    {
#pragma unroll
        for (int i = 0; i <= K; i++)
            u[i] = BDX_UNIT_END;
    }
    __device__ __forceinline__ void push(unsigned v)
    {
#pragma unroll
        for (int i = K; i > 0; i--)
            u[i] = u[i - 1];
        u[0] = v;
    }
    // undoes push (the walk stays at this slot for the next ticket)
    __device__ __forceinline__ void pop()
    {
#pragma unroll
        for (int i = 0; i < K; i++)
            u[i] = u[i + 1];
        u[K] = BDX_UNIT_END;
    }
    // this slot and every later one are empty: the units of the slots it and they draw
    // columns or ROW2 items from are all exhausted
    __device__ __forceinline__ bool terminal() const
    {
        bool t = true;
#pragma unroll
        for (int i = 0; i <= (BDX_MODE == 2 ? K : BDX_LAG); i++)
            t = t && u[i] == BDX_UNIT_END;
        return t;
    }
};

__global__ __launch_bounds__(TH, 2) void k_batch(const float *in, float *out, const bdx_run *__restrict__ runs,
                                              const bdx_item *__restrict__ items, const bdx_unit *__restrict__ units,
                                              unsigned nunits, unsigned qbase, unsigned *ctr,
                                              const float *__restrict__ tab, unsigned spin_limit)
{
    __shared__ float smem[BDX_BATCH_LDS];
    __shared__ unsigned s_item;
    constexpr int F = BDX_MODE == 2 ? 1 : 0;   // item list family
    unsigned *err = ctr + BDX_CTR_ERROR;
    const unsigned q = xcd_id();
    unsigned *ticket = ctr + qbase + q, *unit_next = ctr + qbase + BDX_XQ;
    unsigned long long *slots = reinterpret_cast<unsigned long long *>(ctr + BDX_QSLOT(qbase, nunits, q, 0));
    // thread 0's walk over the slots of this queue: slot j starts at ticket `start`; hist.u[i]
    // is the unit of slot j - 1 - i
    unsigned j = 0, start = 0;
    SlotHistory hist;
    hist.init();
    for (;;) {
        if (threadIdx.x == 0) {
            const unsigned v = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            unsigned it = ~0u;   // item to run, or ~0u: no more work on this queue
            for (unsigned n = 0; j < BDX_SLOTS(nunits);) {
                unsigned long long w = __hip_atomic_load(slots + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (w == 0 && v == start) {   // first ticket of slot j: take its unit and publish it
                    unsigned u = __hip_atomic_fetch_add(unit_next, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    u = u < nunits ? u : BDX_UNIT_END;
                    w = (unsigned long long)(u == BDX_UNIT_END ? BDX_UNIT_END : u + 1u) << 32 |
                        (start + slot_items(units, u, hist.u[BDX_LAG - 1], hist.u[2 * BDX_LAG - 1]));
                    __hip_atomic_store(slots + j, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                while (w == 0) {   // published by the holder of its first ticket, drawn earlier
                    if (++n > spin_limit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                    w = __hip_atomic_load(slots + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                if (w == 0)
                    break;   // error flag set: give up
                const unsigned hi = (unsigned)(w >> 32), end = (unsigned)w;
                const unsigned u0 = hi == BDX_UNIT_END ? BDX_UNIT_END : hi - 1u;
                const unsigned u1 = hist.u[BDX_LAG - 1], u2 = hist.u[2 * BDX_LAG - 1];
                hist.push(u0);
                if (hist.terminal())
                    break;   // this and every later slot are empty
                if (v < end) {
                    unsigned o = v - start;
                    const unsigned r0 = u0 != BDX_UNIT_END ? units[u0].nrow : 0u;
                    const unsigned c1 = u1 != BDX_UNIT_END ? units[u1].ncol : 0u;
                    if (o < r0)
                        it = units[u0].item0[F] + o;
                    else if ((o -= r0) < c1)
                        it = units[u1].item0[F] + units[u1].nrow + o;
                    else
                        it = units[u2].item0[F] + units[u2].nrow + units[u2].ncol + (o - c1);
                    hist.pop();
                    break;
                }
                start = end;
                j++;
            }
            s_item = it;
        }
        __syncthreads();
        const unsigned i = s_item;
        __syncthreads();   // every lane has read s_item before thread 0 writes it again
        if (i == ~0u)
            break;
        const bdx_item it = items[i];
        const bdx_run rn = runs[it.run];
        const unsigned kind = it.kind_ctr >> 28, c = it.kind_ctr & 0x0FFFFFFFu;
        const int h = (int)rn.h, w = (int)rn.w;
        const size_t base = rn.off + (size_t)it.idx * h * w;
        unsigned *rows_done = ctr + BDX_CTR_BASE + 2 * c, *cols_done = rows_done + 1;
        if (kind == BDX_ITEM_SMALL) {
            small_any(h, w, in, out, base, (int)it.unit, smem, tab);
        } else if (kind == BDX_ITEM_ROW) {
            const int r0 = (int)it.unit * BDX_ROW_UNIT(w), r1 = min(r0 + BDX_ROW_UNIT(w), h);
            for (int r = r0; r < r1; r += BDX_ROW_BAND(w)) {
                if (r > r0)
                    __syncthreads();   // the LDS of the previous band is free
                row_any<ROW_OP>(w, in + base, out + base, (size_t)r, min(BDX_ROW_BAND(w), r1 - r), smem, tab);
            }
            signal_done(rows_done);
        } else if (kind == BDX_ITEM_COL) {
            wait_for(rows_done, (unsigned)BDX_CEIL(h, BDX_ROW_UNIT(w)), err, spin_limit);
            const int c0 = (int)it.unit * BDX_COL_UNIT(h), c1 = min(c0 + BDX_COL_UNIT(h), w);
            for (int cc = c0; cc < c1; cc += BDX_COL_STRIP(h)) {
                if (cc > c0)
                    __syncthreads();
                col_any(h, w, out, out, base, cc, min(BDX_COL_STRIP(h), c1 - cc), smem, tab);
            }
            if (BDX_MODE == 2)
                signal_done(cols_done);
        } else if constexpr (BDX_MODE == 2) {
            wait_for(cols_done, (unsigned)BDX_CEIL(w, BDX_COL_UNIT(h)), err, spin_limit);
            const int r0 = (int)it.unit * BDX_ROW_UNIT(w), r1 = min(r0 + BDX_ROW_UNIT(w), h);
            for (int r = r0; r < r1; r += BDX_ROW_BAND(w)) {
                if (r > r0)
                    __syncthreads();
                row_any<I0>(w, out + base, out + base, (size_t)r, min(BDX_ROW_BAND(w), r1 - r), smem, tab);
            }
        }
        __syncthreads();   // LDS is reused by the next item
    }
}

// ---- tiny kernel: whole arrays with both axes <= BDX_TINY_MAX only -----------------------

template <int H, bool NT>
__device__ __forceinline__ void tiny_h(int w, const float *in, float *out, size_t off, int count, float *smem,
                                       const float *__restrict__ tab)
{
    switch (w) {
    case 8: small_unit<H, 8, BDX_MODE, TH, NT>(in, out, off, count, smem, tab); break;
    case 16: small_unit<H, 16, BDX_MODE, TH, NT>(in, out, off, count, smem, tab); break;
    default: small_unit<H, 32, BDX_MODE, TH, NT>(in, out, off, count, smem, tab); break;
    }
}

// Every item of a tiny batch is a SMALL item (the host checked the shapes), and no item
// depends on another: one workgroup per item, in a plain grid the hardware schedules, with
// the units inlined and a small LDS area, so that many workgroups stay resident per CU.
template <bool NT>
__global__ __launch_bounds__(TH) void k_batch_tiny(const float *in, float *out, const bdx_run *__restrict__ runs,
                                                   const bdx_item *__restrict__ items, const float *__restrict__ tab)
{
    static_assert(BDX_TINY_PER(BDX_TINY_MAX, BDX_TINY_MAX) >= 1, "a tiny array must fit one item");
    static_assert(BDX_TINY_PER(8, 8) * 8 * 9 <= BDX_TINY_LDS && BDX_TINY_PER(32, 32) * 32 * 33 <= BDX_TINY_LDS,
                  "a full tiny item must fit the LDS area");
    __shared__ float smem[BDX_TINY_LDS];
    const bdx_item it = items[blockIdx.x];
    const bdx_run rn = runs[it.run];
    const int h = (int)rn.h, w = (int)rn.w;
    const size_t base = rn.off + (size_t)it.idx * h * w;
    switch (h) {
    case 8: tiny_h<8, NT>(w, in, out, base, (int)it.unit, smem, tab); break;
    case 16: tiny_h<16, NT>(w, in, out, base, (int)it.unit, smem, tab); break;
    default: tiny_h<32, NT>(w, in, out, base, (int)it.unit, smem, tab); break;
    }
}

}  // namespace

#define BDX_CAT2(a, b) a##b
#define BDX_CAT(a, b) BDX_CAT2(a, b)

int BDX_CAT(launch_batch_m, BDX_MODE)(int flags, const float *in, float *out, const bdx_run *runs,
                                      const bdx_item *items, unsigned nitems, const bdx_unit *units,
                                      unsigned nunits, unsigned qbase, unsigned *ctr, const float *tab,
                                      unsigned blocks, unsigned spin_limit, hipStream_t s)
{
    if (flags & BDX_LAUNCH_TINY) {
        // one workgroup per item; item lists longer than a launch allows run in several
        const size_t step = max_grid_blocks(TH);
        for (size_t i = 0; i < nitems; i += step) {
            const unsigned n = (unsigned)(nitems - i < step ? nitems - i : step);
            if (flags & BDX_LAUNCH_STREAM)
                hipLaunchKernelGGL(k_batch_tiny<true>, dim3(n), dim3(TH), 0, s, in, out, runs, items + i, tab);
            else
                hipLaunchKernelGGL(k_batch_tiny<false>, dim3(n), dim3(TH), 0, s, in, out, runs, items + i, tab);
            if (hipGetLastError() != hipSuccess)
                return -1;
        }
        return 0;
    }
    hipLaunchKernelGGL(k_batch, dim3(blocks), dim3(TH), 0, s, in, out, runs, items, units, nunits, qbase, ctr, tab,
                       spin_limit);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace bdx
