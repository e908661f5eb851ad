// Ragged ensembles of independent Multipole potentials in one persistent
// launch.
//
// The host plan (agx_batch_plan) cuts every potential's particle range into
// work items of at most 8192 particles and orders them by estimated cost,
// largest first.  One workgroup per resident slot of the chip loops: it takes
// the next item from a device counter (one atomic per item), loads that
// potential's descriptor with scalar loads, and evaluates the item by the
// path the plan chose for that potential:
//   lane   one particle per lane, coefficients gathered from L2
//   lds    the potential's whole table staged into LDS, then one particle
//          per lane reading coefficients from LDS
//   binned the item is one binned tile (counting sort by radial interval in
//          LDS, wavefront-uniform interval, scalar coefficient loads)
// The LDS of the tile sort and the staged table share one buffer, so the
// kernel keeps the occupancy of the single-potential kernels.  Items whose
// potential does not fit the chosen path (table too large for LDS, too many
// radial intervals or too many particles for a tile) fall back to the lane
// path on the device.  No reference counterpart: the reference evaluates one
// potential per launch (gpu_potential.py:757-793).
#include <algorithm>
#include <vector>

#include "multipole_dev.h"

namespace agx {

// The tile sort of the batch kernel takes at most 96 bins (n_int <= 95), so
// that its LDS (about 37 KB) lets four 256-thread workgroups share a CU.
using BatchBin = BinLDST<kMaxTile, 96>;
constexpr int kBatchTableDoubles = (int)(sizeof(BatchBin) / sizeof(double));

union BatchLDS {
    BatchBin bin;
    double tab[kBatchTableDoubles];
};

__global__ void batch_queue_reset(int *q) { if (threadIdx.x == 0) q[0] = 0; }

template <int ORD>
__device__ __forceinline__ void lane_point(const MPK &P, const double *tab_base, const double *xyz,
                                           int64_t i, double *phi, double *force, double *deriv)
{
    const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    PointOut o{};
    Where w = classify(P, x, y, z);
    if (w.kind == 2)
        interior<ORD, false>(P, tab_base + (int64_t)w.k * P.n_lm * 6, x, y, z, w.r2, w.r, w.lnr,
                             w.k, o, 0.0);
    else
        extrap<ORD>(P, w, x, y, z, o, 0.0);
    store<ORD>(o, i, phi, force, deriv);
}

template <int ORD>
__global__ __launch_bounds__(kBinThreads) void mp_batch_kernel(
    const char *__restrict__ descs, int64_t stride, int n_pot,
    const agx_batch_item *__restrict__ items, int64_t n_items, const double *__restrict__ xyz,
    int64_t n_rows, double *__restrict__ phi, double *__restrict__ force,
    double *__restrict__ deriv, int *__restrict__ queue)
{
    __shared__ BatchLDS U;
    __shared__ int item_s;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) item_s = atomicAdd(queue, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(item_s);
        __syncthreads();
        if (it >= n_items) break;
        const int *iw = (const int *)(items + it);
        const int64_t start = (int64_t)(uint32_t)cldi(iw, 0) | ((int64_t)cldi(iw, 1) << 32);
        const int count = cldi(iw, 2), pot = cldi(iw, 3), path = cldi(iw, 4);
        // an item outside the arrays or naming no potential is skipped
        if (start < 0 || count < 1 || start > n_rows - count || pot < 0 || pot >= n_pot) continue;
        const MPK P = load_desc<MPK>(descs + (int64_t)pot * stride);
        const int tab_len = P.n_int * P.n_lm * 6;
        if (path == AGX_BATCH_BINNED && P.n_int + 1 <= BatchBin::kBins && count <= BatchBin::kTile) {
            binned_tile(P, xyz, start, count, U.bin, [&](int b, bool valid, int64_t i) {
                if (!valid) return;
                const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
                PointOut o{};
                if (b > 0) {
                    const int k = b - 1;
                    const double r2 = fma(x, x, fma(y, y, z * z));
                    const double r = sqrt(r2);
                    interior<ORD, true>(P, P.poly + (int64_t)k * P.n_lm * 6, x, y, z, r2, r, agx_log(r),
                                        k, o, 0.0);
                } else {
                    Where w = classify(P, x, y, z);
                    extrap<ORD>(P, w, x, y, z, o, 0.0);
                }
                store<ORD>(o, i, phi, force, deriv);
            });
        } else if (path == AGX_BATCH_LDS && tab_len <= kBatchTableDoubles) {
            const dvec2 *src = (const dvec2 *)P.poly;
            dvec2 *dst = (dvec2 *)U.tab;
            for (int j = tid; j < tab_len / 2; j += kBinThreads) dst[j] = src[j];
            __syncthreads();
            for (int j = tid; j < count; j += kBinThreads)
                lane_point<ORD>(P, U.tab, xyz, start + j, phi, force, deriv);
            __syncthreads();
        } else {
            for (int j = tid; j < count; j += kBinThreads)
                lane_point<ORD>(P, P.poly, xyz, start + j, phi, force, deriv);
        }
    }
}

template <int ORD>
static unsigned batch_grid(int64_t n_items)
{
    static int per_cu[4] = {0, 0, 0, 0};
    static int n_cu = 0;
    if (!n_cu) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            n_cu = v;
        else
            n_cu = 256;
    }
    if (!per_cu[ORD]) {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, mp_batch_kernel<ORD>, kBinThreads, 0) !=
                hipSuccess || b < 1)
            b = 1;
        per_cu[ORD] = b;
    }
    int64_t g = (int64_t)n_cu * per_cu[ORD];
    if (g > n_items) g = n_items;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace agx

using namespace agx;

extern "C" int64_t agx_batch_lds_table_doubles(void) { return kBatchTableDoubles; }

extern "C" int32_t agx_batch_max_bins(void) { return BatchBin::kBins; }

extern "C" int agx_batch_plan(int32_t n_pot, const int64_t *counts, const int32_t *chunk,
                              const int32_t *path, const double *cost, agx_batch_item *items,
                              int64_t capacity, int64_t *n_items)
{
    if (n_pot < 0) return AGX_E_SIZE;
    if (!n_items || (n_pot > 0 && (!counts || !chunk || !path || !cost))) return AGX_E_NULL;
    int64_t total = 0;
    for (int32_t p = 0; p < n_pot; ++p) {
        if (counts[p] < 0 || chunk[p] < 1 || chunk[p] > AGX_BATCH_MAX_ITEM) return AGX_E_SIZE;
        if (path[p] < AGX_BATCH_LANE || path[p] > AGX_BATCH_BINNED) return AGX_E_WHAT;
        total += (counts[p] + chunk[p] - 1) / chunk[p];
    }
    *n_items = total;
    if (!items) return AGX_OK;
    if (capacity < total) return AGX_E_WORKSPACE;
    struct Work {
        double c;
        agx_batch_item it;
    };
    std::vector<Work> w;
    w.reserve((size_t)total);
    int64_t start = 0;
    for (int32_t p = 0; p < n_pot; ++p) {
        for (int64_t o = 0; o < counts[p]; o += chunk[p]) {
            agx_batch_item it;
            it.start = start + o;
            it.count = (int32_t)std::min<int64_t>(chunk[p], counts[p] - o);
            it.pot = p;
            it.path = path[p];
            it.pad = 0;
            w.push_back(Work{cost[p] * it.count, it});
        }
        start += counts[p];
    }
    // longest first: the queue then hands the short items to whichever
    // workgroups finish early
    std::stable_sort(w.begin(), w.end(), [](const Work &a, const Work &b) { return a.c > b.c; });
    for (size_t j = 0; j < w.size(); ++j) items[j] = w[j].it;
    return AGX_OK;
}

extern "C" int agx_multipole_batch_eval(const void *d_desc, int64_t desc_stride, int32_t n_pot,
                                        const agx_batch_item *d_items, int64_t n_items,
                                        const double *d_xyz, int64_t n_rows, double *d_phi,
                                        double *d_force,
                                        double *d_deriv, uint32_t what, int32_t *d_queue,
                                        void *stream)
{
    if (n_items < 0 || n_pot < 0 || n_rows < 0) return AGX_E_SIZE;
    if (desc_stride < (int64_t)sizeof(MPK) || (desc_stride & 7)) return AGX_E_SIZE;
    if (what == 0 || (what & ~(AGX_PHI | AGX_FORCE | AGX_DERIV))) return AGX_E_WHAT;
    if (n_items == 0) return AGX_OK;
    if (!d_desc || !d_items || !d_xyz || !d_queue) return AGX_E_NULL;
    if (((what & AGX_PHI) && !d_phi) || ((what & AGX_FORCE) && !d_force) ||
        ((what & AGX_DERIV) && !d_deriv))
        return AGX_E_NULL;
    hipStream_t s = (hipStream_t)stream;
    double *phi = (what & AGX_PHI) ? d_phi : nullptr;
    double *force = (what & AGX_FORCE) ? d_force : nullptr;
    double *deriv = (what & AGX_DERIV) ? d_deriv : nullptr;
    const char *descs = (const char *)d_desc;
    hipLaunchKernelGGL(batch_queue_reset, dim3(1), dim3(64), 0, s, (int *)d_queue);
    if (deriv)
        hipLaunchKernelGGL(mp_batch_kernel<ORD_HESS>, dim3(batch_grid<ORD_HESS>(n_items)),
                           dim3(kBinThreads), 0, s, descs, desc_stride, n_pot, d_items, n_items,
                           d_xyz, n_rows, phi, force, deriv, (int *)d_queue);
    else if (force)
        hipLaunchKernelGGL(mp_batch_kernel<ORD_GRAD>, dim3(batch_grid<ORD_GRAD>(n_items)),
                           dim3(kBinThreads), 0, s, descs, desc_stride, n_pot, d_items, n_items,
                           d_xyz, n_rows, phi, force, nullptr, (int *)d_queue);
    else
        hipLaunchKernelGGL(mp_batch_kernel<ORD_POT>, dim3(batch_grid<ORD_POT>(n_items)),
                           dim3(kBinThreads), 0, s, descs, desc_stride, n_pot, d_items, n_items,
                           d_xyz, n_rows, phi, nullptr, nullptr, (int *)d_queue);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}
