// The LDS-staged Multipole kernel and its launch (gfx950, FP64), shared by
// multipole.hip (the generic programs) and multipole_tri.hip (the triaxial
// programs, compile-time unrolled; a separate translation unit so that its
// instantiations compile in parallel with the rest).
#pragma once
#include "multipole_dev.h"

namespace agx {

constexpr int64_t kLdsPerCu = 160 * 1024;
constexpr int64_t kLdsTableBytes = kLdsPerCu;

// 16-byte words between consecutive slabs of the table in LDS: the slab's
// 3 n_lm words, plus one when that is even.
__host__ __device__ inline int lds_slab_words(int n_lm) { return 3 * n_lm + (n_lm & 1 ? 0 : 1); }

// One particle per lane with the whole coefficient table staged in LDS once
// per workgroup: the per-term gathers of the lane kernel become LDS reads.
// Used for tables of up to 160 KB (the LDS of one CU); the workgroup size is
// chosen so that the workgroups that fit in LDS hold 16 wavefronts per CU.
// STRIDE and SIN as for interior(): the program flags the variant requires.
// TL > 0 / FL > 0: the triaxial / full program of that degree, unrolled
// (interior()); its register use is held to the occupancy the LDS allows
// anyway (4 waves per SIMD, 3 for the Hessian), or the unrolled steps' loads
// would all be scheduled first.  The potential alone is held to 64 VGPRs (8
// waves: it runs in workgroups twice as large).
template <int ORD, int TL, int FL>
constexpr int lds_min_waves()
{
    return ORD == ORD_POT ? 8 : TL == 0 && FL == 0 ? 1 : ORD == ORD_HESS ? 3 : 4;
}
template <int ORD, int BLOCK, int STRIDE, bool SIN, int TL = 0, int FL = 0>
__global__ __launch_bounds__(BLOCK, (lds_min_waves<ORD, TL, FL>())) void mp_lds_kernel(
    MPK P, const double *__restrict__ xyz, int64_t n, double *__restrict__ phi,
    double *__restrict__ force, double *__restrict__ deriv, double inv4piG)
{
    P = with_law(P);
    extern __shared__ __align__(16) double lds_tab[];
    // Positions are read one grid-stride iteration ahead (the first before
    // the table staging), so the HBM latency of the next particle overlaps
    // the arithmetic of the current one.
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double xn = 0.0, yn = 0.0, zn = 0.0;
    if (i < n) { xn = xyz[3 * i]; yn = xyz[3 * i + 1]; zn = xyz[3 * i + 2]; }
    // The slabs in LDS are lds_slab_words(n_lm) 16-byte words apart: an odd
    // number, so that the lanes' slabs (one radial interval each) start in
    // every bank group (an even stride leaves half or a quarter of them
    // unused and the lanes' reads conflict).
    const int wn = P.n_lm * 3, wp = lds_slab_words(P.n_lm);  // 16-byte words
    const int nv = P.n_int * wn;
    const dvec2 *src = (const dvec2 *)P.poly;
    dvec2 *dst = (dvec2 *)lds_tab;
    if (wp == wn) {
        for (int j = threadIdx.x; j < nv; j += BLOCK) dst[j] = src[j];
    } else {
        // word j of the table goes to slab j / wn, word j % wn (both carried
        // along the loop: BLOCK = q wn + rem, rem < wn)
        const int q = BLOCK / wn, rem = BLOCK - q * wn;
        int k = (int)threadIdx.x / wn, r = (int)threadIdx.x - k * wn;
        for (int j = threadIdx.x; j < nv; j += BLOCK) {
            dst[k * wp + r] = src[j];
            k += q;
            r += rem;
            if (r >= wn) { r -= wn; ++k; }
        }
    }
    __syncthreads();
    const int slab_len = 2 * wp;                       // doubles
    for (; i < n; i += stride) {
        const double x = xn - P.ox, y = yn - P.oy, z = zn - P.oz;
        const int64_t inx = i + stride;
        if (inx < n) { xn = xyz[3 * inx]; yn = xyz[3 * inx + 1]; zn = xyz[3 * inx + 2]; }
        PointOut o{};
        Where w = classify(P, x, y, z);
        if (w.kind == 2)
            interior<ORD, false, NoHook, STRIDE, SIN, TL, FL>(P, lds_tab + w.k * slab_len, x, y,
                                                              z, w.r2, w.r, w.ir, w.lnr, w.k, o,
                                                              inv4piG);
        else
            extrap<ORD>(P, w, x, y, z, o, inv4piG);
        store<ORD>(o, i, phi, force, deriv);
    }
}

// Dynamic LDS beyond 64 KB has to be allowed per kernel.
template <int ORD, int BLOCK, int STRIDE, bool SIN, int TL, int FL>
static void lds_attr(int64_t bytes)
{
    static int64_t set = 0;
    if (bytes > set) {
        if (hipFuncSetAttribute((const void *)mp_lds_kernel<ORD, BLOCK, STRIDE, SIN, TL, FL>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsPerCu) ==
            hipSuccess)
            set = kLdsPerCu;
    }
}

template <int ORD, int BLOCK, int STRIDE, bool SIN, int TL = 0, int FL = 0>
static void launch_lds(const MPK &P, unsigned grid, int64_t tb, const double *xyz, int64_t n,
                       double *phi, double *force, double *deriv, double inv4piG, hipStream_t s)
{
    if (BLOCK > 256) lds_attr<ORD, BLOCK, STRIDE, SIN, TL, FL>(tb);
    hipLaunchKernelGGL((mp_lds_kernel<ORD, BLOCK, STRIDE, SIN, TL, FL>), dim3(grid), dim3(BLOCK),
                       (size_t)tb, s, P, xyz, n, phi, force, deriv, inv4piG);
}

// The unrolled kernel of the triaxial program of degree tl (even, 2..12) for
// this evaluation order and workgroup size (tables up to 80 KB); false
// (nothing launched) where no such variant is built: the caller then runs
// the generic one.  multipole_tri.hip.
bool launch_lds_triaxial(int ord, int block, int tl, const MPK &P, unsigned grid, int64_t tb,
                         const double *xyz, int64_t n, double *phi, double *force, double *deriv,
                         double inv4piG, hipStream_t s);
// The same for the full program of degree fl (even, 2..10; every workgroup
// size: tables up to 160 KB; multipole_full.hip).
bool launch_lds_full(int ord, int block, int fl, const MPK &P, unsigned grid, int64_t tb,
                     const double *xyz, int64_t n, double *phi, double *force, double *deriv,
                     double inv4piG, hipStream_t s);

}  // namespace agx
