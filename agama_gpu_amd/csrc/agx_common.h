// Shared device/host helpers for the gfx950 potential kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "agama_gpu.h"

namespace agx {

// Loads through the constant address space: a wave-uniform address becomes a
// scalar (SMEM) load into SGPRs, a divergent one a vector load.
typedef __attribute__((address_space(4))) const double cdouble_t;
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(4))) const dvec2 cdvec2_t;
typedef __attribute__((address_space(3))) const dvec2 ldvec2_t;   // in LDS
typedef __attribute__((address_space(4))) const int cint_t;

__device__ __forceinline__ double cld(const double *p, int i)
{
    return ((cdouble_t *)p)[i];
}
__device__ __forceinline__ int cldi(const int *p, int i)
{
    return ((cint_t *)p)[i];
}

// 32-byte program records, read with one scalar load each.
typedef int ivec8 __attribute__((ext_vector_type(8)));
typedef int ivec2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(4))) const ivec8 civec8_t;

__device__ __forceinline__ ivec8 cld_rec(const void *base, int idx)
{
    return ((civec8_t *)base)[idx];
}
// FP64 in dwords 4-5 (lo) or 6-7 (hi) of a record.  The pair is copied into
// its own vector first: clang's __builtin_bit_cast of a swizzle expression
// (v.s45) in this toolchain reads the vector's first elements instead of the
// swizzled ones.
__device__ __forceinline__ double rec_lo(const ivec8 &v)
{
    ivec2 t;
    t.x = v.s4;
    t.y = v.s5;
    return __builtin_bit_cast(double, t);
}
__device__ __forceinline__ double rec_hi(const ivec8 &v)
{
    ivec2 t;
    t.x = v.s6;
    t.y = v.s7;
    return __builtin_bit_cast(double, t);
}
// FP64 in dwords 2-3 of a record.
__device__ __forceinline__ double rec_mid(const ivec8 &v)
{
    ivec2 t;
    t.x = v.s2;
    t.y = v.s3;
    return __builtin_bit_cast(double, t);
}
// The two int16 halves of dword 1 (sign-extended): low and high.
__device__ __forceinline__ int rec_off_lo(const ivec8 &v) { return (int)(short)(v.s1 & 0xFFFF); }
__device__ __forceinline__ int rec_off_hi(const ivec8 &v) { return v.s1 >> 16; }

// 1/x from the hardware reciprocal estimate refined by two Newton steps
// (within 2 ulp of the rounded quotient for normal x; the IEEE division
// sequence costs about twice as many instructions).
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    r = fma(fma(-x, r, 1.0), r, r);
    return r;
}

// sqrt(x) and 1/sqrt(x) together, for x >= 2^-767: the hardware reciprocal
// square root estimate refined by two steps of Goldschmidt's coupled
// iteration, g -> sqrt(x) and h -> 1/(2 sqrt(x)) with e = 1/2 - g h, g += g e,
// h += h e (each step squares the relative error), within 1 ulp for sqrt and
// 2 ulp for the reciprocal.  About 11 instructions, against about 20 for the
// library sqrt (which also scales tiny inputs and passes zero and infinity
// through) plus 5 for a reciprocal.
__device__ __forceinline__ void sqrt_rsqrt(double x, double &s, double &rs)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    s = g;
    rs = h + h;
}

// Natural logarithm of a positive normal x, within a few ulp: x = m 2^e with
// m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(f), f = (m-1)/(m+1), |f| <= 0.1716,
// by its odd series to f^21 (the next term is below 2^-53 of the sum).  About
// 30 instructions against about 90 for the library log, whose extra work
// handles zero, subnormal, negative and non-finite inputs that do not occur
// where this is used.
// An FP64 constant pinned to an SGPR pair: the FMAs then take it as their
// scalar operand.  Left to itself the compiler materialises each literal FP64
// constant with two VALU moves into the accumulator register before the FMA.
__device__ __forceinline__ double sconst(double c)
{
    asm volatile("" : "+s"(c));
    return c;
}

// A copy of v that the compiler cannot identify with v: expressions formed
// from it are not merged with the same expressions of v formed earlier, so a
// value needed only at the end of a kernel is recomputed there instead of
// being held in registers across a loop.
__device__ __forceinline__ double vopaque(double v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// A pointer (wave-uniform, in scalar registers) that the compiler cannot
// identify with its value before this point: loads through it at constant
// offsets are not hoisted out of an enclosing loop, where their results
// would be held in registers for the whole loop.
template <class T>
__device__ __forceinline__ const T *sopaque(const T *p)
{
    asm volatile("" : "+s"(p));
    return p;
}

__device__ __forceinline__ double agx_log(double x)
{
    double m = __builtin_amdgcn_frexp_mant(x);     // [0.5, 1)
    int e = __builtin_amdgcn_frexp_exp(x);
    if (m < 0.70710678118654752440) { m += m; --e; }
    const double f = (m - 1.0) * rcp_nr(m + 1.0);
    const double f2 = f * f;
    double p = sconst(1.0 / 21.0);
    p = fma(p, f2, sconst(1.0 / 19.0));
    p = fma(p, f2, sconst(1.0 / 17.0));
    p = fma(p, f2, sconst(1.0 / 15.0));
    p = fma(p, f2, sconst(1.0 / 13.0));
    p = fma(p, f2, sconst(1.0 / 11.0));
    p = fma(p, f2, sconst(1.0 / 9.0));
    p = fma(p, f2, sconst(1.0 / 7.0));
    p = fma(p, f2, sconst(1.0 / 5.0));
    p = fma(p, f2, sconst(1.0 / 3.0));
    const double lnm = fma(2.0 * f * f2, p, 2.0 * f);
    const double de = (double)e;
    // ln 2 split so that e * LN2_HI is exact for |e| < 2^11
    return fma(de, sconst(6.93147180369123816490e-01), fma(de, sconst(1.90821492927058770002e-10), lnm));
}

// ln x for x > 0 given also xm1 = x - 1 to full relative precision: where x
// lies in [1/sqrt 2, sqrt 2) the series variable (x - 1) / (x + 1) is formed
// from xm1, so an argument near 1 keeps the relative precision of its
// distance from 1 (agx_log(x) would have only that of x itself there).
__device__ __forceinline__ double agx_log_xm1(double x, double xm1)
{
    double m = __builtin_amdgcn_frexp_mant(x);     // [0.5, 1)
    int e = __builtin_amdgcn_frexp_exp(x);
    if (m < 0.70710678118654752440) { m += m; --e; }
    const double f = (e == 0 ? xm1 : m - 1.0) * rcp_nr(m + 1.0);
    const double f2 = f * f;
    double p = sconst(1.0 / 21.0);
    p = fma(p, f2, sconst(1.0 / 19.0));
    p = fma(p, f2, sconst(1.0 / 17.0));
    p = fma(p, f2, sconst(1.0 / 15.0));
    p = fma(p, f2, sconst(1.0 / 13.0));
    p = fma(p, f2, sconst(1.0 / 11.0));
    p = fma(p, f2, sconst(1.0 / 9.0));
    p = fma(p, f2, sconst(1.0 / 7.0));
    p = fma(p, f2, sconst(1.0 / 5.0));
    p = fma(p, f2, sconst(1.0 / 3.0));
    const double lnm = fma(2.0 * f * f2, p, 2.0 * f);
    const double de = (double)e;
    return fma(de, sconst(6.93147180369123816490e-01), fma(de, sconst(1.90821492927058770002e-10), lnm));
}

// e^x for |x| < 700, within a few ulp: x = k ln2 + r with |r| <= ln2/2
// (Cody-Waite reduction, ln 2 in two parts), e^r by its Taylor series to r^13
// (the next term is below 2^-57), scaled by 2^k.  About 20 VALU instructions.
// Beyond that range (|x| up to about 1e9) the scaling by 2^k underflows to 0
// or overflows to infinity as e^x does.
__device__ __forceinline__ double agx_exp(double x)
{
    const double k = __builtin_rint(x * sconst(1.44269504088896338700e+00));
    double r = fma(-k, sconst(6.93147180369123816490e-01), x);
    r = fma(-k, sconst(1.90821492927058770002e-10), r);
    double p = sconst(1.0 / 6227020800.0);                 // 1/13!
    p = fma(p, r, sconst(1.0 / 479001600.0));              // 1/12!
    p = fma(p, r, sconst(1.0 / 39916800.0));
    p = fma(p, r, sconst(1.0 / 3628800.0));
    p = fma(p, r, sconst(1.0 / 362880.0));
    p = fma(p, r, sconst(1.0 / 40320.0));
    p = fma(p, r, sconst(1.0 / 5040.0));
    p = fma(p, r, sconst(1.0 / 720.0));
    p = fma(p, r, sconst(1.0 / 120.0));
    p = fma(p, r, sconst(1.0 / 24.0));
    p = fma(p, r, sconst(1.0 / 6.0));
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(p, (int)k);
}

constexpr int kBlock = 256;

// Grid-stride loop over points i < n of an AoS (n,3) position array, calling
// body(i, x, y, z); each position is read one iteration ahead, so its HBM
// latency overlaps the previous point's arithmetic.
template <int BLOCK, class Body>
__device__ __forceinline__ void for_each_point(const double *__restrict__ xyz, int64_t n, Body body)
{
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double xn = 0.0, yn = 0.0, zn = 0.0;
    if (i < n) { xn = xyz[3 * i]; yn = xyz[3 * i + 1]; zn = xyz[3 * i + 2]; }
    for (; i < n; i += stride) {
        const double x = xn, y = yn, z = zn;
        const int64_t inx = i + stride;
        if (inx < n) { xn = xyz[3 * inx]; yn = xyz[3 * inx + 1]; zn = xyz[3 * inx + 2]; }
        body(i, x, y, z);
    }
}

// ---- cross-lane sums without LDS -------------------------------------------
// DPP move of both 32-bit halves of a double (all rows and banks enabled).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}

// gfx950 lane swaps of a pair of doubles, both halves:
//   ROW16 = false (v_permlane32_swap): lanes 32-63 of a trade places with lanes 0-31 of b;
//   ROW16 = true  (v_permlane16_swap): rows 1 and 3 of a (16 lanes each) trade places
//                                      with rows 0 and 2 of b.
template <bool ROW16>
__device__ __forceinline__ void lane_swap(double &a, double &b)
{
    const unsigned long long ua = __builtin_bit_cast(unsigned long long, a);
    const unsigned long long ub = __builtin_bit_cast(unsigned long long, b);
    const unsigned alo = (unsigned)ua, ahi = (unsigned)(ua >> 32);
    const unsigned blo = (unsigned)ub, bhi = (unsigned)(ub >> 32);
    unsigned r0, r1, r2, r3;
    if constexpr (ROW16) {
        const auto lo = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
        const auto hi = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
        r0 = lo[0]; r1 = lo[1]; r2 = hi[0]; r3 = hi[1];
    } else {
        const auto lo = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
        const auto hi = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
        r0 = lo[0]; r1 = lo[1]; r2 = hi[0]; r3 = hi[1];
    }
    a = __builtin_bit_cast(double, ((unsigned long long)r2 << 32) | r0);
    b = __builtin_bit_cast(double, ((unsigned long long)r3 << 32) | r1);
}

// Sum over the 16 lanes of each row (every lane of the row gets its row's sum).
__device__ __forceinline__ double row_sum(double v)
{
    v += dpp_f64<0x128>(v);   // row_ror:8
    v += dpp_f64<0x124>(v);   // row_ror:4
    v += dpp_f64<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_f64<0xB1>(v);    // quad_perm [1,0,3,2]
    return v;
}

// Sum of v over the 64 lanes of the wavefront (every lane gets the total).
// Every lane must be active.
__device__ __forceinline__ double wave_sum(double v)
{
    double a = v, b = v;
    lane_swap<false>(a, b);
    v = a + b;                // lane i: v_i + v_(i^32)
    a = v; b = v;
    lane_swap<true>(a, b);
    return row_sum(a + b);    // + the paired row, then the row
}

// Wavefront sums of v0..v5 added to g[0..5] by FP64 atomics (every lane must
// be active).  The swaps transpose as they add: after the 32-lane swaps the
// low half of a register holds one quantity's pair sums and the high half
// another's; after the 16-lane swaps each row holds a different quantity, so
// six sums cost five swaps and two row reductions, and the adds are two
// atomic instructions (4 and 2 lanes).
// Transposed wavefront sums of v0..v5: on return, row r (lanes 16r..16r+15) of d
// holds the total of quantity ((r & 1) << 1) | (r >> 1), i.e. rows 0..3 hold
// quantities 0, 2, 1, 3, and rows 0-1 / 2-3 of e hold quantities 4 / 5.
// Every lane must be active.  The swaps transpose as they add: after the
// 32-lane swaps the low half of a register holds one quantity's pair sums and
// the high half another's; after the 16-lane swaps each row holds a different
// quantity, so six sums cost five swaps and two row reductions.
__device__ __forceinline__ void wave_sum6(double v0, double v1, double v2, double v3, double v4,
                                          double v5, double &d, double &e)
{
    lane_swap<false>(v0, v1);
    lane_swap<false>(v2, v3);
    lane_swap<false>(v4, v5);
    double a = v0 + v1;      // lanes 0-31: quantity 0, lanes 32-63: quantity 1
    double b = v2 + v3;      // 2 | 3
    double c = v4 + v5;      // 4 | 5
    lane_swap<true>(a, b);
    d = row_sum(a + b);
    double c2 = c;
    lane_swap<true>(c, c2);
    e = row_sum(c + c2);
}

// Transposed wavefront sums of v[0..15]: on return, row r of d[k] holds the
// total of v[4k + (((r & 1) << 1) | (r >> 1))].  Every lane must be active.
__device__ __forceinline__ void wave_sum16(double v[16], double d[4])
{
#pragma unroll
    for (int j = 0; j < 8; ++j) lane_swap<false>(v[2 * j], v[2 * j + 1]);
    double a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = v[2 * j] + v[2 * j + 1];   // lo: v[2j], hi: v[2j+1]
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        lane_swap<true>(a[2 * k], a[2 * k + 1]);
        d[k] = row_sum(a[2 * k] + a[2 * k + 1]);
    }
}

// wave_sum6, added to g[0..5] by FP64 atomics (two atomic instructions, 4 and 2 lanes).
__device__ __forceinline__ void wave_sum6_add(double *g, double v0, double v1, double v2,
                                              double v3, double v4, double v5, int lane)
{
    double d, e;
    wave_sum6(v0, v1, v2, v3, v4, v5, d, e);
    const int row = lane >> 4;
    if ((lane & 15) == 0) atomicAdd(g + (((row & 1) << 1) | (row >> 1)), d);
    if ((lane & 31) == 0) atomicAdd(g + 4 + (row >> 1), e);
}

// Derivative order of an evaluation: potential; +gradient; +Hessian; density.
enum { ORD_POT = 0, ORD_GRAD = 1, ORD_HESS = 2, ORD_DENS = 3 };

// One evaluated point (density is returned in phi for ORD_DENS).
struct PointOut {
    double phi;
    double g[3];   // +grad Phi
    double h[6];   // +Hessian [xx, yy, zz, xy, yz, xz]
};

// Agama's gravitational constant, kpc (km/s)^2 / Msol (reference gpu_potential.py:157)
constexpr double G_AGAMA_D = 4.30091727067736e-06;

// Workgroups per CU of the grid-stride launches (agx_set_option
// AGX_OPT_GRID_PER_CU; defined in capi.hip).
extern int g_grid_per_cu;
// AGX_OPT_CYL_PATH of agx_set_option (defined in cylspline.hip).
int cyl_set_path(int64_t value);
int adj_set_stage(int64_t value);
int cyl_path();

// Workgroups of `kernel` (block threads, lds bytes of dynamic LDS each) that
// are resident on the current device at once: the runtime's occupancy per CU
// times the CU count.  Cached per device, kernel, block and LDS size;
// `fallback` if the runtime cannot tell.  Defined in capi.hip.
int64_t resident_workgroups(const void *kernel, int block, size_t lds, int64_t fallback);

// Grid size for a grid-stride launch over n items: enough blocks to fill
// 256 CUs several times over (oversubscribed, so that workgroups finishing
// early free their slots for the remaining work), never more than needed.
inline unsigned grid_for(int64_t n, int per_cu_blocks = g_grid_per_cu)
{
    int64_t want = (n + kBlock - 1) / kBlock;
    int64_t cap = (int64_t)256 * per_cu_blocks;
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (unsigned)want;
}

// Read a device descriptor image with uniform (scalar) loads.
template <typename T>
__device__ __forceinline__ T load_desc(const void *src)
{
    static_assert(sizeof(T) % 4 == 0, "descriptor size");
    T d;
    int *dst = (int *)&d;
#pragma unroll
    for (int q = 0; q < (int)(sizeof(T) / 4); ++q) dst[q] = cldi((const int *)src, q);
    return d;
}

}  // namespace agx
