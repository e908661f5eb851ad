// Ragged ensembles of independent Multipole potentials in one persistent
// launch.
//
// The host plan (agx_batch_plan) cuts every potential's particle range into
// work items of at most AGX_BATCH_MAX_ITEM particles and orders them by
// estimated cost, largest first.  One 1024-thread workgroup per CU loops: it
// takes the next item from a device counter (one atomic per item), loads that
// potential's descriptor with scalar loads, and evaluates the item by the path
// the plan chose for that potential:
//   lane  one particle per lane, coefficients gathered from L2
//   lds   the potential's whole table staged into LDS (up to ~160 KB), then
//         one particle per lane reading coefficients from LDS
// One workgroup of 16 wavefronts per CU owns the CU's whole LDS, so any table
// of the single-potential LDS path fits; an item whose table does not fit
// falls back to the lane path on the device.  No reference counterpart: the
// reference evaluates one potential per launch (gpu_potential.py:757-793).
#include <algorithm>
#include <vector>

#include "multipole_dev.h"

namespace agx {

constexpr int kBatchBlock = 1024;
// dynamic LDS for the table: the CU's 160 KB less room for the static
// item index
constexpr int kBatchTableDoubles = (160 * 1024 - 256) / 8;

__global__ void batch_queue_reset(int *q) { if (threadIdx.x == 0) q[0] = 0; }

// rows [start, start + count) with the table at tab_base; each thread reads
// its next position one iteration ahead.  TL > 0: the potential's program is
// the triaxial program of degree TL and its table is in LDS (interior()).
template <int ORD, int STRIDE, bool SIN, int TL = 0>
__device__ __forceinline__ void batch_rows(const MPK &P, const double *tab_base,
                                           const double *__restrict__ xyz, int64_t start, int count,
                                           double *phi, double *force, double *deriv)
{
    int j = threadIdx.x;
    double xn = 0.0, yn = 0.0, zn = 0.0;
    if (j < count) {
        const int64_t i = start + j;
        xn = xyz[3 * i]; yn = xyz[3 * i + 1]; zn = xyz[3 * i + 2];
    }
    for (; j < count; j += kBatchBlock) {
        const double x = xn, y = yn, z = zn;
        const int jn = j + kBatchBlock;
        if (jn < count) {
            const int64_t in = start + jn;
            xn = xyz[3 * in]; yn = xyz[3 * in + 1]; zn = xyz[3 * in + 2];
        }
        PointOut o{};
        Where w = classify(P, x, y, z);
        if (w.kind == 2)
            interior<ORD, false, NoHook, TL ? 2 : STRIDE, TL ? false : SIN, TL>(
                P, tab_base + (int64_t)w.k * P.n_lm * 6, x, y, z, w.r2, w.r, w.ir, w.lnr, w.k, o,
                0.0);
        else
            extrap<ORD>(P, w, x, y, z, o, 0.0);
        store<ORD>(o, start + j, phi, force, deriv);
    }
}

// STRIDE and SIN as for interior(), from the flags every potential's program
// satisfies.
template <int ORD, int STRIDE, bool SIN>
__global__ __launch_bounds__(kBatchBlock) void mp_batch_kernel(
    const char *__restrict__ descs, int64_t stride, int n_pot,
    const agx_batch_item *__restrict__ items, int64_t n_items, const double *__restrict__ xyz,
    int64_t n_rows, double *__restrict__ phi, double *__restrict__ force,
    double *__restrict__ deriv, int *__restrict__ queue)
{
    extern __shared__ __align__(16) double tab[];
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
        const MPK P = with_law(load_desc<MPK>(descs + (int64_t)pot * stride));
        const int tab_len = P.n_int * P.n_lm * 6;
        if ((path & 0xFF) == AGX_BATCH_LDS && tab_len <= kBatchTableDoubles) {
            const dvec2 *src = (const dvec2 *)P.poly;
            dvec2 *dst = (dvec2 *)tab;
            for (int j = tid; j < tab_len / 2; j += kBatchBlock) dst[j] = src[j];
            __syncthreads();
            // the triaxial degree in bits 8-15 of the item's path, where it
            // agrees with the descriptor's sizes: the unrolled program (not
            // for the Hessian, which the 1024-thread workgroups hold to 128
            // VGPRs already)
            const int tl = (path >> 8) & 0xFF;
            const bool tri = ORD != ORD_HESS && tl >= 2 && tl <= 10 && !(tl & 1) &&
                             P.n_groups == tl / 2 + 1 && P.n_steps == tri_step0(tl, tl / 2 + 1) &&
                             P.n_lm == P.n_steps;
            if constexpr (ORD != ORD_HESS) {
                if (tri) {
                    switch (tl) {
                    case 2: batch_rows<ORD, STRIDE, SIN, 2>(P, tab, xyz, start, count, phi, force, deriv); break;
                    case 4: batch_rows<ORD, STRIDE, SIN, 4>(P, tab, xyz, start, count, phi, force, deriv); break;
                    case 6: batch_rows<ORD, STRIDE, SIN, 6>(P, tab, xyz, start, count, phi, force, deriv); break;
                    case 8: batch_rows<ORD, STRIDE, SIN, 8>(P, tab, xyz, start, count, phi, force, deriv); break;
                    default: batch_rows<ORD, STRIDE, SIN, 10>(P, tab, xyz, start, count, phi, force, deriv); break;
                    }
                }
            }
            if (!tri) batch_rows<ORD, STRIDE, SIN>(P, tab, xyz, start, count, phi, force, deriv);
            __syncthreads();
        } else {
            batch_rows<ORD, STRIDE, SIN>(P, P.poly, xyz, start, count, phi, force, deriv);
        }
    }
}

// Grid of the persistent kernel: the workgroups resident at once on the current
// device (cached per device by resident_workgroups), at most one per work item.
// The table's dynamic LDS exceeds 64 KB, which has to be allowed per kernel
// first; 0 if that fails, so the caller reports a launch error.
template <int ORD, int STRIDE, bool SIN>
static unsigned batch_grid(int64_t n_items)
{
    const size_t lds = (size_t)kBatchTableDoubles * sizeof(double);
    const void *k = (const void *)mp_batch_kernel<ORD, STRIDE, SIN>;
    if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return 0;
    int64_t g = resident_workgroups(k, kBatchBlock, lds, 256);
    if (g > n_items) g = n_items;
    if (g < 1) g = 1;
    return (unsigned)g;
}

template <int ORD, int STRIDE, bool SIN>
static bool launch_batch(const char *descs, int64_t stride, int n_pot, const agx_batch_item *items,
                         int64_t n_items, const double *xyz, int64_t n_rows, double *phi,
                         double *force, double *deriv, int *queue, hipStream_t s)
{
    const unsigned grid = batch_grid<ORD, STRIDE, SIN>(n_items);
    if (!grid) return false;
    hipLaunchKernelGGL((mp_batch_kernel<ORD, STRIDE, SIN>), dim3(grid), dim3(kBatchBlock),
                       kBatchTableDoubles * sizeof(double), s, descs, stride, n_pot, items, n_items,
                       xyz, n_rows, phi, force, deriv, queue);
    return true;
}

template <int ORD>
static bool launch_batch(uint32_t flags, const char *descs, int64_t stride, int n_pot,
                         const agx_batch_item *items, int64_t n_items, const double *xyz,
                         int64_t n_rows, double *phi, double *force, double *deriv, int *queue,
                         hipStream_t s)
{
    const int st = prog_stride(flags);
    const bool co = flags & AGX_PROG_COS_ONLY;
#define AGX_LAUNCH_BATCH(S, C)                                                                  \
    return launch_batch<ORD, S, C>(descs, stride, n_pot, items, n_items, xyz, n_rows, phi, force, \
                                   deriv, queue, s)
    if (st == 1 && co) AGX_LAUNCH_BATCH(1, false);
    if (st == 1) AGX_LAUNCH_BATCH(1, true);
    if (st == 2 && co) AGX_LAUNCH_BATCH(2, false);
    if (st == 2) AGX_LAUNCH_BATCH(2, true);
    if (co) AGX_LAUNCH_BATCH(0, false);
    AGX_LAUNCH_BATCH(0, true);
#undef AGX_LAUNCH_BATCH
}

}  // namespace agx

using namespace agx;

extern "C" int64_t agx_batch_lds_table_doubles(void) { return kBatchTableDoubles; }


extern "C" int agx_batch_plan(int32_t n_pot, const int64_t *counts, const int32_t *chunk,
                              const int32_t *path, const double *cost, agx_batch_item *items,
                              int64_t capacity, int64_t *n_items)
{
    if (n_pot < 0) return AGX_E_SIZE;
    if (!n_items || (n_pot > 0 && (!counts || !chunk || !path || !cost))) return AGX_E_NULL;
    int64_t total = 0;
    for (int32_t p = 0; p < n_pot; ++p) {
        if (counts[p] < 0 || chunk[p] < 1 || chunk[p] > AGX_BATCH_MAX_ITEM) return AGX_E_SIZE;
        // AGX_BATCH_LANE or AGX_BATCH_LDS, the latter with a triaxial degree
        // (even, <= 32) in bits 8-15
        const int base = path[p] & 0xFF, tl = path[p] >> 8;
        if (base < AGX_BATCH_LANE || base > AGX_BATCH_LDS) return AGX_E_WHAT;
        if (tl < 0 || tl > 32 || (tl & 1) || (tl && base != AGX_BATCH_LDS)) return AGX_E_WHAT;
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
                                        double *d_deriv, uint32_t what, uint32_t prog_flags,
                                        int32_t *d_queue, void *stream)
{
    if (n_items < 0 || n_pot < 0 || n_rows < 0) return AGX_E_SIZE;
    if (desc_stride < (int64_t)sizeof(MPK) || (desc_stride & 7)) return AGX_E_SIZE;
    if (what == 0 || (what & ~(AGX_PHI | AGX_FORCE | AGX_DERIV))) return AGX_E_WHAT;
    if (prog_flags & ~kProgFlags) return AGX_E_WHAT;
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
    int *q = (int *)d_queue;
    bool ok;
    if (deriv)
        ok = launch_batch<ORD_HESS>(prog_flags, descs, desc_stride, n_pot, d_items, n_items, d_xyz,
                                    n_rows, phi, force, deriv, q, s);
    else if (force)
        ok = launch_batch<ORD_GRAD>(prog_flags, descs, desc_stride, n_pot, d_items, n_items, d_xyz,
                                    n_rows, phi, force, nullptr, q, s);
    else
        ok = launch_batch<ORD_POT>(prog_flags, descs, desc_stride, n_pot, d_items, n_items, d_xyz,
                                   n_rows, phi, nullptr, nullptr, q, s);
    return (hipGetLastError() == hipSuccess && ok) ? AGX_OK : AGX_E_LAUNCH;
}
