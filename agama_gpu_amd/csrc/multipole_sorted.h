// The interval-sorted Multipole kernel and its launch (gfx950, FP64), shared by
// multipole.hip (the generic programs) and multipole_full.hip (the full
// programs unrolled, for the variant that stages slabs in LDS).
#pragma once
#include "multipole_lds.h"
#include "sort_dev.h"

namespace agx {

constexpr int64_t kSortedSlabLds = 40 * 1024;

// The interval-sorted order (sort_dev.h): 64 particles of one interval per
// wavefront and chunk, every wavefront full except one per interval.
// SLAB: the wavefront copies its interval's coefficients (n_lm x 6) into a
// private LDS slab whenever its chunks move to another interval and reads
// them from there (uniform addresses: broadcast reads); otherwise they come
// by scalar loads, as in the binned kernel.  Scalar loads suit tables whose
// slabs are long (the per-term arithmetic covers their latency, and LDS
// would hold few slabs); for slabs of a few KB the scalar cache, shared by
// many wavefronts each streaming a different interval, misses, and the
// wavefronts wait on L2 latency most of the time.  STRIDE and SIN as for
// interior(): the program flags the variant requires.  TL / FL > 0 (with
// SLAB): the triaxial / full program of that degree unrolled (interior()),
// its registers held as in mp_lds_kernel.
template <int ORD, int TL, int FL>
constexpr int sorted_min_waves()
{
    return TL == 0 && FL == 0 ? 1 : lds_min_waves<ORD, TL, FL>();
}
template <int ORD, bool SLAB, int STRIDE, bool SIN, int TL = 0, int FL = 0>
__global__ __launch_bounds__(kBlock, (sorted_min_waves<ORD, TL, FL>())) void mp_sorted_kernel(
    MPK P, const double *__restrict__ xyz, double *__restrict__ phi, double *__restrict__ force,
    double *__restrict__ deriv, double inv4piG, SortWork W, int run_len)
{
    static_assert(SLAB || (TL == 0 && FL == 0), "unrolled programs read LDS slabs");
    P = with_law(P);
    extern __shared__ __align__(16) double slab_lds[];   // SLAB: kBlock / 64 slabs
    const int slab_len = P.n_lm * 6;
    double *my_slab = slab_lds + (threadIdx.x >> 6) * slab_len;
    int staged = -1;                 // interval held in my_slab (wavefront-uniform)
    for_each_sorted_chunk(W, P.n_int + 1, run_len, [&](int b, bool valid, int64_t i) {
        if (SLAB && b > 0 && b - 1 != staged) {
            // every lane of the wavefront copies; LDS accesses of one
            // wavefront complete in order, so the reads below (and the
            // previous interval's reads before these writes) need no wait
            const dvec2 *src = (const dvec2 *)(P.poly + (int64_t)(b - 1) * slab_len);
            dvec2 *dst = (dvec2 *)my_slab;
            for (int q = threadIdx.x & 63; q < slab_len / 2; q += 64) dst[q] = src[q];
            staged = b - 1;
        }
        if (!valid) return;
        double x, y, z;
        load_pos(P, xyz, i, x, y, z);
        PointOut o{};
        if (b > 0) {
            const int k = b - 1;
            const double r2 = fma(x, x, fma(y, y, z * z));
            double r, ir;
            sqrt_rsqrt(r2, r, ir);      // r >= r_min: r2 is far above 2^-767
            interior<ORD, !SLAB, NoHook, STRIDE, SIN, TL, FL>(
                P, SLAB ? my_slab : P.poly + (int64_t)k * slab_len, x, y, z, r2, r, ir,
                agx_log(r), k, o, inv4piG);
        } else {
            const Where w = classify(P, x, y, z);
            extrap<ORD>(P, w, x, y, z, o, inv4piG);
        }
        store<ORD, false>(o, i, phi, force, deriv);
    });
}

struct SortedLaunch {
    int64_t n;
    const double *xyz;
    double *phi, *force, *deriv;
    SortWork W;
};

template <int ORD, bool SLAB, int STRIDE, bool SIN, int TL = 0, int FL = 0>
static void launch_sorted(const MPK &P, const SortedLaunch &L, size_t lds, hipStream_t s)
{
    auto kern = mp_sorted_kernel<ORD, SLAB, STRIDE, SIN, TL, FL>;
    const SortedShape sh = sorted_shape(
        L.n, P.n_int + 1, resident_workgroups((const void *)kern, kBlock, lds, 256 * 3), kBlock);
    hipLaunchKernelGGL(kern, dim3(sh.grid), dim3(kBlock), lds, s, P, L.xyz, L.phi, L.force, L.deriv,
                       0.0, L.W, sh.run_len);
}

// The LDS-slab sorted kernel of the full program of degree fl (even, 2..10);
// false (nothing launched) where no such variant is built.  multipole_full.hip.
bool launch_sorted_full(int ord, int fl, const MPK &P, const SortedLaunch &L, size_t lds,
                        hipStream_t s);

}  // namespace agx
