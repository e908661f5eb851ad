// Counting sort of particle indices by a small integer key ("bin"), into a
// caller-provided workspace, and the traversal that hands the sorted order to
// whole wavefronts 64 slots at a time.  Used by the adjoint kernels: with the
// particles of one bin (radial interval, spline cell) in every wavefront,
// per-bin table gradients are summed across the wavefront before they go to
// memory.
//
// Each bin's slot range is padded to a multiple of 64, so a 64-slot chunk
// never straddles two bins; padding slots are recognised by their position
// (slot >= off[b] + hist[b]).
#pragma once
#include "agx_common.h"

namespace agx {

// Threads per workgroup of the count and scatter kernels: large workgroups,
// so that the grid (held small: see g_sort_grid) still holds several
// wavefronts per SIMD to cover the latency of the position reads.
constexpr int kSortBlock = 1024;
constexpr int kSortMaxBins = 8192;

struct SortWork {
    int *hist;       // [nb] particles per bin
    int *off;        // [nb + 1] first slot of each bin
    int *cursor;     // [nb] next free slot of each bin while scattering
    int *queue;      // [1] next run of chunks (the dynamic traversal)
    uint16_t *bin;   // [n]
    int *idx;        // [n + 64 nb] particle of each slot
};

constexpr int64_t sort_header_bytes(int nb) { return ((int64_t)4 * (3 * nb + 2) + 255) & ~(int64_t)255; }

constexpr int64_t sort_work_size(int64_t n, int nb)
{
    return sort_header_bytes(nb) + ((2 * n + 255) & ~(int64_t)255) + 4 * (n + 64 * (int64_t)nb);
}

inline SortWork sort_work(void *base, int64_t n, int nb)
{
    SortWork w;
    int *h = (int *)base;
    w.hist = h;
    w.off = h + nb;
    w.cursor = h + 2 * nb + 1;
    w.queue = h + 3 * nb + 1;
    uint8_t *p = (uint8_t *)base + sort_header_bytes(nb);
    w.bin = (uint16_t *)p;
    w.idx = (int *)(p + ((2 * n + 255) & ~(int64_t)255));
    return w;
}

// The kernels below have internal linkage: every translation unit that sorts
// gets its own copy.
namespace {

__global__ __launch_bounds__(kSortBlock) void sort_init_kernel(SortWork W, int nb)
{
    for (int b = threadIdx.x; b < nb; b += kSortBlock) W.hist[b] = 0;
    if (threadIdx.x == 0) W.queue[0] = 0;
}

// Bin of every particle (key(i) in [0, nb)) and the histogram: LDS per
// workgroup (nb ints of dynamic LDS), then one global add per nonzero bin.
template <class Key>
__global__ __launch_bounds__(kSortBlock) void sort_count_kernel(Key key, int64_t n, SortWork W,
                                                               int nb)
{
    extern __shared__ int sort_cnt[];
    for (int b = threadIdx.x; b < nb; b += kSortBlock) sort_cnt[b] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * kSortBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kSortBlock) {
        const int b = key(i);
        W.bin[i] = (uint16_t)b;
        atomicAdd(&sort_cnt[b], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += kSortBlock)
        if (sort_cnt[b]) atomicAdd(&W.hist[b], sort_cnt[b]);
}

// Exclusive scan of the 64-padded bin sizes: one workgroup of 1024 threads,
// each summing a contiguous block of bins, then a scan over the 1024 sums.
__global__ __launch_bounds__(1024) void sort_scan_kernel(SortWork W, int nb)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = t * per, b1 = min(b0 + per, nb);
    int s = 0;
    for (int b = b0; b < b1; ++b) s += (W.hist[b] + 63) & ~63;
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {   // Hillis-Steele inclusive scan
        const int v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int o = part[t] - s;
    for (int b = b0; b < b1; ++b) {
        W.off[b] = o;
        W.cursor[b] = o;
        o += (W.hist[b] + 63) & ~63;
    }
    if (t == 1023) W.off[nb] = part[1023];
}

// Particle indices into their bins' slots: each workgroup reserves one range
// per bin (one global atomic), then places its particles with LDS counters
// (2 nb ints of dynamic LDS).  Any particle-to-workgroup mapping works.
__global__ __launch_bounds__(kSortBlock) void sort_scatter_kernel(int64_t n, int nb, SortWork W)
{
    extern __shared__ int sort_lds[];
    int *cnt = sort_lds, *base = sort_lds + nb;
    for (int b = threadIdx.x; b < nb; b += kSortBlock) cnt[b] = 0;
    __syncthreads();
    const int64_t i0 = (int64_t)blockIdx.x * kSortBlock + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * kSortBlock;
    for (int64_t i = i0; i < n; i += stride) atomicAdd(&cnt[W.bin[i]], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += kSortBlock) {
        base[b] = cnt[b] ? atomicAdd(&W.cursor[b], cnt[b]) : 0;
        cnt[b] = 0;
    }
    __syncthreads();
    for (int64_t i = i0; i < n; i += stride) {
        const int b = W.bin[i];
        W.idx[base[b] + atomicAdd(&cnt[b], 1)] = (int)i;
    }
}

__global__ void sort_requeue_kernel(SortWork W)
{
    if (threadIdx.x == 0) W.queue[0] = 0;
}

}  // namespace

// Rewind the run queue of a workspace that already holds a sort (to traverse
// the same sorted order again).
inline void sort_requeue(const SortWork &W, hipStream_t stream)
{
    hipLaunchKernelGGL(sort_requeue_kernel, dim3(1), dim3(64), 0, stream, W);
}

// Workgroups of the count and scatter launches: every workgroup adds its
// histogram to the global one with one atomic per nonzero bin, and those
// atomics serialise per bin, so the grid grows with N only slowly: about one
// workgroup per 4096 points, 64 to 1024 (measured on MI355X: 256 best at
// N = 1M, 1024 at N = 10M).  0 = that rule; agx_set_option AGX_OPT_SORT_GRID.
inline int g_sort_grid = 0;

// The four sort launches on `stream`; key is a device functor int(int64_t i).
template <class Key>
inline void sort_launch(Key key, int64_t n, int nb, const SortWork &W, hipStream_t stream)
{
    const int64_t n_blk = (n + kSortBlock - 1) / kSortBlock;
    int64_t want = g_sort_grid;
    if (want <= 0) want = n / 4096 < 64 ? 64 : (n / 4096 > 1024 ? 1024 : n / 4096);
    const unsigned grid = (unsigned)(n_blk < want ? n_blk : want);
    hipLaunchKernelGGL(sort_init_kernel, dim3(1), dim3(kSortBlock), 0, stream, W, nb);
    hipLaunchKernelGGL(sort_count_kernel<Key>, dim3(grid), dim3(kSortBlock), (size_t)nb * 4, stream,
                       key, n, W, nb);
    hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(1024), 0, stream, W, nb);
    hipLaunchKernelGGL(sort_scatter_kernel, dim3(grid), dim3(kSortBlock), (size_t)nb * 8, stream, n,
                       nb, W);
}

// The chunks of the sorted order, handed to whole wavefronts: for each chunk,
// body(bin, valid, i) is called by all 64 lanes (padding lanes with valid =
// false, i = 0).  Must be called by every lane of the wavefront.
//
// run_len > 0: runs of run_len consecutive chunks taken from the queue until
// none is left, which balances the work to within a run.  run_len = 0:
// static contiguous ranges, one per wavefront of the grid, with bin 0 (the
// extrapolation bin: chunks of a fraction of the interior chunks' work) and
// the interior chunks split separately, so every wavefront gets equal shares
// of both.  The queue's counter serialises at its one L2 address (about 16 ns
// per increment, measured), which costs more than the balance is worth when
// each wavefront has only a few chunks; the launcher picks (sorted_shape).
// Contiguous ranges also keep most of a wavefront's consecutive chunks in one
// bin, so per-bin state (an LDS slab) is reloaded or flushed only where a bin
// ends.
template <class Body>
__device__ __forceinline__ void for_each_sorted_chunk(const SortWork &W, int nb, int run_len,
                                                      Body body)
{
    const int lane = threadIdx.x & 63;
    const int total = cldi(W.off, nb) >> 6;
    // chunks [c0, c1) in order, from bin b0 on (the bin of c0 or an earlier one)
    auto range = [&](int c0, int c1, int b) {
        for (int c = c0; c < c1; ++c) {
            while (cldi(W.off, b + 1) <= c * 64) ++b;
            const int sl = c * 64 + lane;
            const bool valid = sl < cldi(W.off, b) + cldi(W.hist, b);
            const int64_t i = valid ? (int64_t)W.idx[sl] : 0;
            body(b, valid, i);
        }
    };
    if (run_len == 0) {
        const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
        const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) +
                          __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int64_t ne = cldi(W.off, 1) >> 6, ni = total - ne;   // bin 0, interior
        range((int)(w * ne / nw), (int)((w + 1) * ne / nw), 0);
        const int i0 = (int)(ne + w * ni / nw), i1 = (int)(ne + (w + 1) * ni / nw);
        if (i0 >= i1) return;
        // bin of chunk i0: the last b with off[b] <= 64 i0
        int b = 0, hi = nb;
        while (hi - b > 1) {
            const int mid = (b + hi) >> 1;
            if (cldi(W.off, mid) <= i0 * 64) b = mid; else hi = mid;
        }
        range(i0, i1, b);
        return;
    }
    for (;;) {
        int run = 0;
        if (lane == 0) run = atomicAdd(W.queue, 1);
        run = __builtin_amdgcn_readfirstlane(run);
        const int c0 = run * run_len;
        if (c0 >= total) break;
        int b = 0, hi = nb;
        while (hi - b > 1) {
            const int mid = (b + hi) >> 1;
            if (cldi(W.off, mid) <= c0 * 64) b = mid; else hi = mid;
        }
        range(c0, min(c0 + run_len, total), b);
    }
}

// Launch shape of a kernel that walks the sorted order with workgroups of
// `block` threads, resident_wgs of which fit on the device at once: no more
// workgroups than that (later ones would only start late), and with at least
// 32 chunks per wavefront the queue with runs of a quarter of that share
// (few increments per unit of work), otherwise static ranges (run_len 0).
struct SortedShape {
    unsigned grid;
    int run_len;
};

inline SortedShape sorted_shape(int64_t n, int nb, int64_t resident_wgs, int block)
{
    const int64_t per_wg = block / 64;
    const int64_t chunks = (n + 63) / 64 + nb;
    int64_t wg = (chunks + per_wg - 1) / per_wg;
    if (wg > resident_wgs) wg = resident_wgs;
    if (wg < 1) wg = 1;
    const int64_t share = chunks / (wg * per_wg);
    return SortedShape{(unsigned)wg, share >= 32 ? (int)(share / 4) : 0};
}

}  // namespace agx
