// CylSpline (azimuthal-harmonic bicubic spline) device code for gfx950, FP64.
//
// Same potential as the reference's cylspl_eval_device
// (_cylspl_potential_kernel.cu:396-863): per azimuthal harmonic m a cubic
// Hermite spline in the asinh-scaled coordinates lR = asinh(R/Rscale),
// lz = asinh(z/Rscale) from node values and derivatives (f, f_lR, f_lz,
// f_lR lz), log-scaling of m=0 (Phi0 = -exp(f0), others Phi_m/Phi0), the
// Fourier sum in phi, and outside the grid a PowerLaw multipole
// Phi = sum_lm W_lm (r0/r)^(l+1) mul Ptilde_lm T_m.  Differences from the
// reference, each a correction:
//  * the Cartesian Hessian includes the dPhi/dphi terms of the
//    cylindrical-to-Cartesian chain rule (the reference drops them, :820-846);
//  * outside the grid the gradient includes dPhi/dtheta and the Hessian all
//    harmonics (the reference keeps only dPhi/dr, dPhi/dphi and a monopole
//    Hessian, :354-382);
//  * the associated Legendre recurrence uses P_(l-2) (the reference's
//    sphHarm_device feeds 0 / P_(l-3) there, :205-223).
#pragma once
#include "sph_dev.h"

namespace agx {

struct CSK {
    const double *lR, *lz;        // [nR], [nz] asinh-scaled grids
    const double *nodes;          // [n_harm][nR][nz][4]
    const int *m_arr;             // [n_harm]
    const double *W;              // [(lmax_o+1)^2] outer PowerLaw, index l(l+1)+m
    int nR, nz, n_harm, log_scaling, lmax_o, mmax_o;
    double Rscale, r0;
    double iRs;                   // 1 / Rscale
};

// The point's cylindrical radius and scaled coordinates.  asinh(X / Rscale)
// = ln a, a = (X + q) / Rscale with q = sqrt(X^2 + Rscale^2), whose derivative
// in X, 1/q, comes from the same square root.  The logarithm also takes
// a - 1 = (X + X^2 / (q + Rscale)) / Rscale, formed without the cancellation
// of X + q - Rscale, so that a coordinate much smaller than Rscale keeps its
// relative precision (ln(X + q) - ln Rscale loses it: near the axis the
// Hessian's Phi_R / R - Phi_RR would then carry errors of 1e-7 relative to
// Phi_RR at R = 1e-9 Rscale).  lR and |lz| are clamped at 0 (rounding near
// the axis and the midplane must not move a point below a grid that starts at
// 0).  R^2 below 2^-767 is the axis: R = 1/R = 0.
struct CylCoords {
    double R, iR, lR, lz, jR, jz;
};

__device__ __forceinline__ CylCoords cyl_coords(const CSK &P, double x, double y, double z)
{
    CylCoords c;
    const double R2 = fma(x, x, y * y), Rs2 = P.Rscale * P.Rscale;
    c.R = 0.0;
    c.iR = 0.0;
    if (R2 >= 0x1p-767) sqrt_rsqrt(R2, c.R, c.iR);
    double qR, qz;
    sqrt_rsqrt(R2 + Rs2, qR, c.jR);
    sqrt_rsqrt(fma(z, z, Rs2), qz, c.jz);
    // (compared, not fmax: a NaN coordinate stays NaN, fails every bounds test
    // of the caller and gives a NaN result, as in the reference)
    const double az = fabs(z);
    const double tR = agx_log_xm1((c.R + qR) * P.iRs,
                                  fma(c.R, c.R * rcp_nr(qR + P.Rscale), c.R) * P.iRs);
    const double tz = agx_log_xm1((az + qz) * P.iRs,
                                  fma(az, az * rcp_nr(qz + P.Rscale), az) * P.iRs);
    c.lR = tR < 0.0 ? 0.0 : tR;
    c.lz = copysign(tz < 0.0 ? 0.0 : tz, z);
    return c;
}

// A grid node: from the table (LDS = false: the constant address space, a
// lane-divergent index is a vector load through the L1) or from a copy of the
// grid in LDS (LDS = true: g points into a __shared__ array, and the binary
// search's chain of dependent reads costs LDS latency instead of L1 latency).
template <bool LDS>
__device__ __forceinline__ double gnode(const double *g, int i)
{
    return LDS ? g[i] : cld(g, i);
}

// cell index i with grid[i] <= x < grid[i+1], clamped to [0, n-2]
template <bool LDS = false>
__device__ __forceinline__ int cell_search(double x, const double *g, int n)
{
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (gnode<LDS>(g, mid) > x) hi = mid; else lo = mid;
    }
    return lo;
}

// Most grid nodes (lR and lz together) a kernel stages in LDS.
constexpr int kGridLDS = 512;

// Copy the lR then the lz grid into grids[0 .. nR + nz) (nR + nz <= kGridLDS);
// every thread of the workgroup takes part, and a barrier follows.
__device__ __forceinline__ void stage_grids(const CSK &P, double *grids)
{
    for (int k = threadIdx.x; k < P.nR + P.nz; k += blockDim.x)
        grids[k] = k < P.nR ? cld(P.lR, k) : cld(P.lz, k - P.nR);
    __syncthreads();
}

// Cubic Hermite weights of one grid axis at a point of the cell [x0, x1]
// (the reference's evalCubic1D, :110-128, in weight form): h = x1 - x0,
// ih = 1/h, t = (x - x0) ih, T = 1 - t.  A quantity with values c0, c1 and
// slopes s0, s1 at the cell's two nodes is
//     c0 w0 + c1 w1 + s0 u0 + s1 u1,
//     w0 = T^2 (1 + 2t),  w1 = t^2 (1 + 2T),  u0 = h t T^2,  u1 = -h t^2 T,
// and its first and second derivatives in x are (c0 - c1) dw + s0 du0 + s1 du1
// and the same with d2w, d2u0, d2u1 (the value weights' derivatives are
// opposite).  The value keeps both weights rather than c1 + (c0 - c1) w0: near
// a node whose value is 0 (an m != 0 harmonic on the axis, where the value
// grows as R^2 and is then divided by R) w0 rounds to 1 and the difference
// form would lose every digit of the value.  The weights depend on the point
// only, so every harmonic of a cell shares them: per harmonic the bicubic
// patch costs 10 FMA per quantity and axis instead of forming the
// one-dimensional Hermite polynomial afresh for each.
template <bool D1, bool D2>
struct HermAxis {
    double w0, w1, u0, u1, dw, du0, du1, d2w, d2u0, d2u1;
    __device__ __forceinline__ HermAxis(double t, double h, double ih)
    {
        const double T = 1.0 - t, tT = t * T;
        w0 = T * T * fma(2.0, t, 1.0);
        w1 = t * t * fma(2.0, T, 1.0);
        u0 = h * tT * T;
        u1 = -h * tT * t;
        if (D1) {
            dw = -6.0 * tT * ih;
            du0 = T * fma(-3.0, t, 1.0);
            du1 = t * fma(3.0, t, -2.0);
        }
        if (D2) {
            d2w = 6.0 * (t - T) * (ih * ih);
            d2u0 = fma(6.0, t, -4.0) * ih;
            d2u1 = fma(6.0, t, -2.0) * ih;
        }
    }
    // value f and, with E1 / E2 (which need D1 / D2), the first and second
    // derivatives of the quantity (c0, c1, s0, s1)
    template <bool E1, bool E2>
    __device__ __forceinline__ void eval(double c0, double c1, double s0, double s1, double &f,
                                         double &df, double &d2f) const
    {
        static_assert((!E1 || D1) && (!E2 || D2), "weights not formed");
        f = fma(c0, w0, fma(c1, w1, fma(s0, u0, s1 * u1)));
        const double dif = c0 - c1;
        if (E1) df = fma(dif, dw, fma(s0, du0, s1 * du1));
        if (E2) d2f = fma(dif, d2w, fma(s0, d2u0, s1 * d2u1));
    }
};

// Constants of the outer multipole's normalised Legendre recurrences for
// l, m <= 8, evaluated at compile time (a square root by Newton's iteration,
// converged to the last bit): P~_l^m = a ct P~_(l-1)^m - b P~_(l-2)^m, and
// P~_m^m = f_m sin(theta) P~_(m-1)^(m-1).
constexpr double const_sqrt(double x)
{
    double r = x > 1.0 ? x : 1.0;
    for (int i = 0; i < 200; ++i) r = 0.5 * (r + x / r);
    return r;
}

struct OuterRecurrence {
    double a[9][9], b[9][9], f[9];   // [l][m], [m]
};

constexpr OuterRecurrence make_outer_recurrence()
{
    OuterRecurrence t{};
    for (int m = 0; m <= 8; ++m) {
        t.f[m] = (m == 0) ? 0.0 : -const_sqrt((2.0 * m + 1.0) / (2.0 * m));
        for (int l = m + 1; l <= 8; ++l) {
            if (l == m + 1) {
                t.a[l][m] = const_sqrt(2.0 * m + 3.0);
                t.b[l][m] = 0.0;
            } else {
                const double l2 = (double)l * l, m2 = (double)m * m;
                t.a[l][m] = const_sqrt((4.0 * l2 - 1.0) / (l2 - m2));
                t.b[l][m] = const_sqrt((2.0 * l + 1.0) * ((l - 1.0) * (l - 1.0) - m2) /
                                       ((2.0 * l - 3.0) * (l2 - m2)));
            }
        }
    }
    return t;
}

__constant__ const OuterRecurrence kOuterRec = make_outer_recurrence();

// Outer PowerLaw multipole (lmax_o <= 8): spherical derivatives.  R and 1/R
// as cyl_coords gives them.
template <bool G, bool H, bool RHO>
__device__ __forceinline__ void cyl_outer(const CSK &P, double x, double y, double z, double R,
                                          double iR, PointOut &o)
{
    if (RHO) { o.phi = 0.0; return; }   // a harmonic expansion: zero density
    const double r2 = fma(R, R, z * z);
    double r, ir;
    if (r2 >= 0x1p-767) {
        sqrt_rsqrt(r2, r, ir);
    } else {
        r = sqrt(r2);
        ir = 1.0 / r;
    }
    const double ct = z * ir, st = R * ir;
    const bool pole = st < 1.0e-10;
    double cph = 1.0, sph = 0.0;
    if (R > 0.0) { cph = x * iR; sph = y * iR; }
    const double ist = pole ? 0.0 : r * iR;
    const double q = P.r0 * ir;
    double Sv = 0, Sr = 0, Srr = 0, St = 0, Stt = 0, Srt = 0, Sf = 0, Sff = 0, Srf = 0, Stf = 0;
    double Sfn = 0, Sfp = 0;
    double cm = 1.0, sm = 0.0, pmm = 1.0 / sqrt(4.0 * M_PI), dpmm = 0.0;
    for (int m = 0; m <= P.mmax_o; ++m) {
        if (m > 0) {
            const double t = fma(cm, cph, -sm * sph);
            sm = fma(sm, cph, cm * sph); cm = t;
            const double f = kOuterRec.f[m];
            dpmm = f * fma(ct, pmm, st * dpmm);   // d/dtheta of f sin(theta) P_(m-1,m-1)
            pmm = f * st * pmm;
        }
        const double mul = (m == 0) ? 2.0 * sqrt(M_PI) : 2.0 * sqrt(2.0 * M_PI);
        double Q1 = pmm, dQ1 = dpmm, Q2 = 0.0, dQ2 = 0.0;
        double rl = q;
        for (int l = 0; l < m; ++l) rl *= q;      // (r0/r)^(m+1)
        double Vc = 0, Vs = 0, Rc = 0, Rs = 0, RRc = 0, RRs = 0, Tc = 0, Ts = 0;
        double TTc = 0, TTs = 0, RTc = 0, RTs = 0;
        for (int l = m; l <= P.lmax_o; ++l) {
            if (l > m) {
                const double a = kOuterRec.a[l][m], b = kOuterRec.b[l][m];
                const double Qn = fma(a * ct, Q1, -b * Q2);
                const double dQn = fma(a, fma(ct, dQ1, -st * Q1), -b * dQ2);
                Q2 = Q1; Q1 = Qn; dQ2 = dQ1; dQ1 = dQn;
                rl *= q;
            }
            const double Y = mul * Q1, dY = mul * dQ1;
            const double d2Y = pole ? 0.0
                : -ct * ist * dY - ((double)(l * (l + 1)) - (double)(m * m) * ist * ist) * Y;
            const double wc = cld(P.W, l * (l + 1) + m) * rl;
            const double ws = (m > 0) ? cld(P.W, l * (l + 1) - m) * rl : 0.0;
            const double k1 = -(double)(l + 1), k2 = k1 * k1;   // d/dlnr, d2/dlnr2 of (r0/r)^(l+1)
            Vc = fma(wc, Y, Vc); Vs = fma(ws, Y, Vs);
            Rc = fma(k1 * wc, Y, Rc); Rs = fma(k1 * ws, Y, Rs);
            Tc = fma(wc, dY, Tc); Ts = fma(ws, dY, Ts);
            if (H) {
                RRc = fma(k2 * wc, Y, RRc); RRs = fma(k2 * ws, Y, RRs);
                TTc = fma(wc, d2Y, TTc); TTs = fma(ws, d2Y, TTs);
                RTc = fma(k1 * wc, dY, RTc); RTs = fma(k1 * ws, dY, RTs);
            }
        }
        Sv = fma(cm, Vc, fma(sm, Vs, Sv));
        Sr = fma(cm, Rc, fma(sm, Rs, Sr));
        St = fma(cm, Tc, fma(sm, Ts, St));
        if (m > 0) {
            const double fm = (double)m, Fv = fm * fma(cm, Vs, -sm * Vc);
            Sfn += Fv;
            if (m == 1) Sfp += fma(cm, Ts, -sm * Tc);
            if (H) {
                Sf += Fv;
                Sff -= fm * fm * fma(cm, Vc, sm * Vs);
                Srf += fm * fma(cm, Rs, -sm * Rc);
                Stf += fm * fma(cm, Ts, -sm * Tc);
            }
        }
        if (H) {
            Srr = fma(cm, RRc, fma(sm, RRs, Srr));
            Stt = fma(cm, TTc, fma(sm, TTs, Stt));
            Srt = fma(cm, RTc, fma(sm, RTs, Srt));
        }
    }
    o.phi = Sv;
    if (!G && !H) return;
    if (H && pole) {
        // The Hessian on the z axis (sin(theta) < 1e-10): the regular limit,
        // as for the Multipole kernels (the chain rule's pole branch drops
        // d2Y/dtheta2 and the |m| = 1, 2 limits).  A term is
        // W (r0/r)^(l+1) mul Qh(z/r) h_m(x, y) / r^m with Qh = P~_lm / sin^m
        // (the recurrence from the seed without its sin factors, at
        // cos(theta) = s), V its value: m = 0 gives H_zz = (l+1)(l+2) V / r^2
        // and H_xx = H_yy = -(l+1)(l+2) V / (2 r^2), m = 1 gives H_xz (cos),
        // H_yz (sin) = -s (l+2) V / r^2, m = 2 gives H_xx = -H_yy = 2 V / r^2
        // (cos) and H_xy = 2 V / r^2 (sin).
        const double sg = ct < 0.0 ? -1.0 : 1.0;
        double h0 = 0.0, hzz = 0.0, h2c = 0.0, h2s = 0.0, hxz = 0.0, hyz = 0.0;
        double ph = 1.0 / sqrt(4.0 * M_PI);
        for (int m = 0; m <= P.mmax_o && m <= 2; ++m) {
            if (m > 0) ph *= kOuterRec.f[m];
            const double mul = (m == 0) ? 2.0 * sqrt(M_PI) : 2.0 * sqrt(2.0 * M_PI);
            double Q1 = ph, Q2 = 0.0, rl = q;
            for (int l = 0; l < m; ++l) rl *= q;
            for (int l = m; l <= P.lmax_o; ++l) {
                if (l > m) {
                    const double Qn = fma(kOuterRec.a[l][m] * sg, Q1, -kOuterRec.b[l][m] * Q2);
                    Q2 = Q1; Q1 = Qn;
                    rl *= q;
                }
                const double Y = mul * Q1 * rl, k = (double)(l + 2);
                const double vc = cld(P.W, l * (l + 1) + m) * Y;
                const double vs = (m > 0) ? cld(P.W, l * (l + 1) - m) * Y : 0.0;
                if (m == 0) { hzz = fma((double)(l + 1) * k, vc, hzz); h0 = fma(-0.5 * (l + 1) * k, vc, h0); }
                else if (m == 1) { hxz = fma(-sg * k, vc, hxz); hyz = fma(-sg * k, vs, hyz); }
                else { h2c = fma(2.0, vc, h2c); h2s = fma(2.0, vs, h2s); }
            }
        }
        const double ir2 = ir * ir;
        o.h[0] = (h0 + h2c) * ir2; o.h[1] = (h0 - h2c) * ir2; o.h[2] = hzz * ir2;
        o.h[3] = h2s * ir2; o.h[4] = hyz * ir2; o.h[5] = hxz * ir2;
    }
    SphDerivs d;
    d.Pr = Sr; d.Pt = St; d.Pfos = pole ? (ct < 0.0 ? -Sfp : Sfp) : Sfn * ist;   // south pole: -
    d.Prr = Srr; d.Prt = Srt; d.Prf = Srf; d.Ptt = Stt; d.Ptf = Stf; d.Pf = Sf; d.Pff = Sff;
    sph_to_cart<H>(x, y, z, ir, R, ct, st, cph, sph, d, o.g, o.h, H && pole);
}

// Outside the grid box in scaled coordinates (a NaN coordinate is inside:
// every comparison fails, and the spline then gives NaN).
__device__ __forceinline__ bool cyl_outside(const CSK &P, const CylCoords &cc)
{
    return cc.lR < cld(P.lR, 0) || cc.lR > cld(P.lR, P.nR - 1) || cc.lz < cld(P.lz, 0) ||
           cc.lz > cld(P.lz, P.nz - 1);
}

// The outer PowerLaw expansion of a point outside the box, for ORD.
template <int ORD>
__device__ __forceinline__ void cylspline_outside(const CSK &P, double x, double y, double z,
                                                  const CylCoords &cc, PointOut &o)
{
    constexpr bool G = ORD == ORD_GRAD || ORD == ORD_HESS;
    constexpr bool H = ORD == ORD_HESS || ORD == ORD_DENS;   // density from the trace
    cyl_outer<G || H, H, ORD == ORD_DENS>(P, x, y, z, cc.R, cc.iR, o);
}

template <int ORD, bool LDS = false>
__device__ __forceinline__ void cylspline_inside(const CSK &P, double x, double y, double z,
                                                 const CylCoords &cc, PointOut &o,
                                                 const double *gR = nullptr,
                                                 const double *gz = nullptr);

// One point of a CylSpline potential.  ORD as the Multipole kernels.
template <int ORD>
__device__ __forceinline__ void cylspline_point(const CSK &P, double x, double y, double z,
                                                PointOut &o)
{
    const CylCoords cc = cyl_coords(P, x, y, z);
    if (cyl_outside(P, cc))
        cylspline_outside<ORD>(P, x, y, z, cc, o);
    else
        cylspline_inside<ORD>(P, x, y, z, cc, o);
}

// One harmonic's bicubic spline in cell (iR, iz) with the point's axis
// weights aR, aZ: its value v, dv/dlR (G), d2v/dlR2 (H) and, with Z, dv/dlz
// (G), d2v/dlRdlz, d2v/dlz2 (H).  topR / topZ: the point is on the cell's
// upper lR / lz node.
template <bool G, bool H, bool Z, class AR, class AZ>
__device__ __forceinline__ void harm_cell(const CSK &P, int mm, int iR, int iz, const AR &aR,
                                          const AZ &aZ, bool topR, bool topZ, double &v,
                                          double &vR, double &vRR, double &vZ, double &vRZ,
                                          double &vZZ)
{
    const int64_t plane = (int64_t)P.nR * P.nz;
    // corners (iR,iz), (iR,iz+1) and (iR+1,iz), (iR+1,iz+1): two 64-byte runs
    const dvec2 *lo = (const dvec2 *)(P.nodes + ((int64_t)mm * plane + (int64_t)iR * P.nz + iz) * 4);
    const dvec2 *hi = lo + 2 * P.nz;
    const dvec2 a0 = lo[0], a1 = lo[1], b0 = lo[2], b1 = lo[3];   // (iR,iz): f,fx | fy,fxy ; (iR,iz+1)
    const dvec2 c0 = hi[0], c1 = hi[1], e0 = hi[2], e1 = hi[3];   // (iR+1,iz) ; (iR+1,iz+1)
    // offset by one corner value for accuracy (reference :551-554)
    const double off = topR ? (topZ ? e0.x : c0.x) : (topZ ? b0.x : a0.x);
    // step 1, along lz: f and f_lR on the lower (0) and upper (1) lR node,
    // with their lz derivatives
    constexpr bool ZG = Z && G, ZH = Z && H;
    double f0, f0z, f0zz, f1, f1z, f1zz, g0, g0z, g0zz, g1, g1z, g1zz;
    aZ.template eval<ZG, ZH>(a0.x - off, b0.x - off, a1.x, b1.x, f0, f0z, f0zz);
    aZ.template eval<ZG, ZH>(c0.x - off, e0.x - off, c1.x, e1.x, f1, f1z, f1zz);
    aZ.template eval<ZG, ZH>(a0.y, b0.y, a1.y, b1.y, g0, g0z, g0zz);
    aZ.template eval<ZG, ZH>(c0.y, e0.y, c1.y, e1.y, g1, g1z, g1zz);
    // step 2, along lR
    double dummy;
    aR.template eval<G, H>(f0, f1, g0, g1, v, vR, vRR);
    v += off;
    if (ZG) aR.template eval<H, false>(f0z, f1z, g0z, g1z, vZ, vRZ, dummy);
    if (ZH) aR.template eval<false, false>(f0zz, f1zz, g0zz, g1zz, vZZ, dummy, dummy);
}

// The bicubic spline of a point inside the box.  LDS: the grids are read from
// the LDS copies gR (nR nodes) and gz (nz nodes) that stage_grids made;
// otherwise from the table.
template <int ORD, bool LDS>
__device__ __forceinline__ void cylspline_inside(const CSK &P, double x, double y, double z,
                                                 const CylCoords &cc, PointOut &o,
                                                 const double *gR, const double *gz)
{
    constexpr bool G = ORD == ORD_GRAD || ORD == ORD_HESS;
    constexpr bool H = ORD == ORD_HESS || ORD == ORD_DENS;   // density from the trace
    if (!LDS) { gR = P.lR; gz = P.lz; }
    const double R = cc.R, lR = cc.lR, lz = cc.lz;
    const int iR = cell_search<LDS>(lR, gR, P.nR), iz = cell_search<LDS>(lz, gz, P.nz);
    const double lR0 = gnode<LDS>(gR, iR), lR1 = gnode<LDS>(gR, iR + 1);
    const double lz0 = gnode<LDS>(gz, iz), lz1 = gnode<LDS>(gz, iz + 1);
    const double hR = lR1 - lR0, hz = lz1 - lz0;
    const double ihR = rcp_nr(hR), ihz = rcp_nr(hz);
    const HermAxis<G || H, H> aR((lR - lR0) * ihR, hR, ihR);
    const HermAxis<G || H, H> aZ((lz - lz0) * ihz, hz, ihz);
    double cph = 1.0, sph = 0.0;
    if (R > 0.0) { cph = x * cc.iR; sph = y * cc.iR; }

    // Fourier sums of the (scaled) harmonics and their lR, lz, phi derivatives
    double S = P.log_scaling ? 1.0 : 0.0, SR = 0, SZ = 0, SRR = 0, SRZ = 0, SZZ = 0;
    double Sp = 0, Spp = 0, SRp = 0, SZp = 0;
    double F0 = 0, F0R = 0, F0Z = 0, F0RR = 0, F0RZ = 0, F0ZZ = 0;
    int cur_m = 0;
    double cm = 1.0, sm = 0.0;
    const bool topR = lR == lR1, topZ = lz == lz1;
    for (int mm = 0; mm < P.n_harm; ++mm) {
        const int mv = cldi(P.m_arr, mm), am = mv < 0 ? -mv : mv;
        double v, vR = 0, vRR = 0, vZ = 0, vRZ = 0, vZZ = 0;
        harm_cell<G || H, H, true>(P, mm, iR, iz, aR, aZ, topR, topZ, v, vR, vRR, vZ, vRZ, vZZ);
        if (mv == 0) {
            F0 = v; F0R = vR; F0Z = vZ; F0RR = vRR; F0RZ = vRZ; F0ZZ = vZZ;
            if (!P.log_scaling) {
                S += v; SR += vR; SZ += vZ; SRR += vRR; SRZ += vRZ; SZZ += vZZ;
            }
            continue;
        }
        while (cur_m < am) {                       // cos/sin(|m| phi) by rotation
            const double t = fma(cm, cph, -sm * sph);
            sm = fma(sm, cph, cm * sph); cm = t; ++cur_m;
        }
        if (cur_m > am) {                          // harmonics are sorted by m, |m| may drop
            cm = 1.0; sm = 0.0; cur_m = 0;
            while (cur_m < am) {
                const double t = fma(cm, cph, -sm * sph);
                sm = fma(sm, cph, cm * sph); cm = t; ++cur_m;
            }
        }
        const double T = (mv > 0) ? cm : sm;
        const double dT = (mv > 0) ? -(double)am * sm : (double)am * cm;
        const double d2T = -(double)(am * am) * T;
        S = fma(v, T, S); SR = fma(vR, T, SR); SZ = fma(vZ, T, SZ);
        Sp = fma(v, dT, Sp);
        if (H) {
            SRR = fma(vRR, T, SRR); SRZ = fma(vRZ, T, SRZ); SZZ = fma(vZZ, T, SZZ);
            SRp = fma(vR, dT, SRp); SZp = fma(vZ, dT, SZp); Spp = fma(v, d2T, Spp);
        }
    }
    // ---- on the z axis: the regular part of d2Phi/dlR2 -----------------------
    // Every harmonic's spline has zero slope in lR at the axis, so near it
    // Phi = Phi(axis) + lR^2 / 2 sum_m Q_m T_m(phi) + O(lR^3), lR ~ R / Rscale.
    // Of the lR^2 terms only m = 0 (x^2 + y^2), m = +2 (x^2 - y^2) and m = -2
    // (2 x y) are smooth in x, y; the others' second derivatives depend on the
    // direction of approach and average to zero over it, which is the limit
    // taken here.  Q_m is the phi-Fourier coefficient of d2Phi/dlR2 at the
    // point (with log-scaling the product rule of Phi = -exp(F0) S).
    double axQ0 = 0.0, axQc = 0.0, axQs = 0.0;
    if (H && !(R > 1.0e-30)) {
        double Pa = 1.0, PaR = 0.0, PaRR = 0.0;
        if (P.log_scaling) {
            Pa = -agx_exp(F0); PaR = Pa * F0R; PaRR = Pa * (F0RR + F0R * F0R);
            axQ0 = PaRR;
        } else {
            axQ0 = F0RR;
        }
        for (int mm = 0; mm < P.n_harm; ++mm) {
            const int mv = cldi(P.m_arr, mm);
            if (mv != 2 && mv != -2) continue;
            double v, vR, vRR, d0, d1, d2;
            harm_cell<H, H, false>(P, mm, iR, iz, aR, aZ, topR, topZ, v, vR, vRR, d0, d1, d2);
            const double q = fma(Pa, vRR, fma(2.0 * PaR, vR, PaRR * v));
            if (mv > 0) axQc = q; else axQs = q;
        }
    }
    // ---- physical derivatives in (lR, lz, phi) ------------------------------
    double Pv, PR, PZ, Pp, PRR = 0, PZZ = 0, PRZ = 0, PRp = 0, PZp = 0, Ppp = 0;
    if (P.log_scaling) {
        const double P0 = -agx_exp(F0);
        const double P0R = P0 * F0R, P0Z = P0 * F0Z;
        Pv = P0 * S;
        PR = fma(P0, SR, P0R * S);
        PZ = fma(P0, SZ, P0Z * S);
        Pp = P0 * Sp;
        if (H) {
            const double P0RR = P0 * (F0RR + F0R * F0R), P0ZZ = P0 * (F0ZZ + F0Z * F0Z);
            const double P0RZ = P0 * (F0RZ + F0R * F0Z);
            PRR = P0 * SRR + 2.0 * P0R * SR + P0RR * S;
            PZZ = P0 * SZZ + 2.0 * P0Z * SZ + P0ZZ * S;
            PRZ = P0 * SRZ + P0R * SZ + P0Z * SR + P0RZ * S;
            PRp = P0 * SRp + P0R * Sp;
            PZp = P0 * SZp + P0Z * Sp;
            Ppp = P0 * Spp;
        }
    } else {
        Pv = S; PR = SR; PZ = SZ; Pp = Sp;
        PRR = SRR; PZZ = SZZ; PRZ = SRZ; PRp = SRp; PZp = SZp; Ppp = Spp;
    }
    if (ORD == ORD_POT) { o.phi = Pv; return; }
    // ---- (lR, lz) -> (R, z): dlX/dX = 1/sqrt(X^2 + Rscale^2) -----------------
    // the Jacobian and 1/R formed again here (from opaque copies of R and z)
    // rather than held in registers across the harmonic loop
    const double Rq = vopaque(R), zq = vopaque(z), Rs2 = P.Rscale * P.Rscale;
    double qq, jR, jz;
    sqrt_rsqrt(fma(Rq, Rq, Rs2), qq, jR);
    sqrt_rsqrt(fma(zq, zq, Rs2), qq, jz);
    const double invR = R > 0.0 ? rcp_nr(Rq) : 0.0;
    const double PhR = PR * jR, PhZ = PZ * jz;
    double HRR = 0, HZZ = 0, HRZ = 0, HRp = 0, HZp = 0;
    if (H) {
        HRR = PRR * jR * jR - PR * R * jR * jR * jR;
        HZZ = PZZ * jz * jz - PZ * z * jz * jz * jz;
        HRZ = PRZ * jR * jz;
        HRp = PRp * jR;
        HZp = PZp * jz;
    }
    if (ORD == ORD_DENS) {
        // Laplacian in cylindrical coordinates: Phi_RR + Phi_R/R + Phi_pp/R^2 + Phi_zz
        const double lap = (R > 1.0e-30) ? HRR + (PhR + Ppp * invR) * invR + HZZ
                                         : 2.0 * axQ0 * jR * jR + HZZ;
        o.phi = lap * (1.0 / (4.0 * M_PI * G_AGAMA_D));
        return;
    }
    o.phi = Pv;
    // ---- (R, z, phi) -> Cartesian ---------------------------------------------
    if (R > 1.0e-30) {
        o.g[0] = PhR * cph - Pp * sph * invR;
        o.g[1] = PhR * sph + Pp * cph * invR;
    } else {
        o.g[0] = 0.0; o.g[1] = 0.0;
    }
    o.g[2] = PhZ;
    if (ORD != ORD_HESS) return;
    if (R > 1.0e-30) {
        const double iR2 = invR * invR;
        const double c2 = cph * cph, s2 = sph * sph, cs = cph * sph;
        // d/dx = cos d/dR - sin/R d/dphi, d/dy = sin d/dR + cos/R d/dphi
        o.h[0] = HRR * c2 - 2.0 * HRp * cs * invR + (PhR * invR) * s2 + (2.0 * Pp * iR2) * cs + Ppp * s2 * iR2;
        o.h[1] = HRR * s2 + 2.0 * HRp * cs * invR + (PhR * invR) * c2 - (2.0 * Pp * iR2) * cs + Ppp * c2 * iR2;
        o.h[3] = HRR * cs + HRp * (c2 - s2) * invR - (PhR * invR) * cs - Pp * (c2 - s2) * iR2 - Ppp * cs * iR2;
        o.h[4] = HRZ * sph + HZp * cph * invR;
        o.h[5] = HRZ * cph - HZp * sph * invR;
    } else {
        // the slopes in lR vanish on the axis, and with them the m = +-1 part of
        // d2Phi/dRdz: H_xz = H_yz = 0
        const double j2 = jR * jR;
        o.h[0] = j2 * (axQ0 + axQc); o.h[1] = j2 * (axQ0 - axQc); o.h[3] = j2 * axQs;
        o.h[4] = 0.0; o.h[5] = 0.0;
    }
    o.h[2] = HZZ;
}

inline CSK make_cyl_args(const agx_cylspline *cs)
{
    CSK P;
    P.lR = cs->lR_grid; P.lz = cs->lz_grid; P.nodes = cs->nodes; P.m_arr = cs->m_arr;
    P.W = cs->W_outer;
    P.nR = cs->nR; P.nz = cs->nz; P.n_harm = cs->n_harm; P.log_scaling = cs->log_scaling;
    P.lmax_o = cs->lmax_outer; P.mmax_o = cs->mmax_outer;
    P.Rscale = cs->Rscale; P.r0 = cs->r0_outer;
    P.iRs = 1.0 / cs->Rscale;
    return P;
}

}  // namespace agx
