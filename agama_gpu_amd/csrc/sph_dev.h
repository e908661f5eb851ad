// Spherical (ln r, theta, phi) derivatives -> Cartesian gradient and Hessian.
#pragma once
#include "agx_common.h"

namespace agx {

// Derivatives of Phi with respect to t = ln r, theta and phi.  Pfos is
// (dPhi/dphi)/sin(theta), finite on the z axis; the second-order entries are
// only read when the Hessian is requested.
struct SphDerivs {
    double Pr, Pt, Pfos;
    double Prr, Prt, Prf, Ptt, Ptf, Pf, Pff;
};

// Chain rule of the reference Multipole kernel (_multipole_potential_kernel.cu
// :390-409 gradient, :728-799 Hessian).  g: +grad Phi; h: [xx,yy,zz,xy,yz,xz].
// On the z axis the phi terms are dropped (their coefficients vanish there).
// keep_h: leave h as it is (the caller has formed it another way).
template <bool HESS>
__device__ __forceinline__ void sph_to_cart(double x, double y, double z, double ir, double R,
                                            double ct, double st, double cph, double sph,
                                            const SphDerivs &d, double g[3], double h[6],
                                            bool keep_h = false)
{
    const double ir2 = ir * ir;
    g[0] = fma(d.Pr * x, ir2, fma(d.Pt, ct * cph * ir, -d.Pfos * sph * ir));
    g[1] = fma(d.Pr * y, ir2, fma(d.Pt, ct * sph * ir, d.Pfos * cph * ir));
    g[2] = fma(d.Pr * z, ir2, -d.Pt * st * ir);
    if (!HESS || keep_h) return;
    const double iR = (R > 1.0e-300) ? 1.0 / R : 0.0;
    const double c3[3] = {x, y, z};
    const double dl[3] = {x * ir2, y * ir2, z * ir2};
    const double dt[3] = {ct * cph * ir, ct * sph * ir, -st * ir};
    const double dp[3] = {-sph * iR, cph * iR, 0.0};
    const double ir3 = ir2 * ir, ir4 = ir2 * ir2, ir5 = ir4 * ir;
    const double dct[3] = {-z * x * ir3, -z * y * ir3, ir - z * z * ir3};   // d cos(theta)
    const double isth = (st > 1.0e-14) ? 1.0 / st : 0.0;
    const double cos3 = ct * isth * isth * isth;
    double iR4 = 0.0;
    if (R > 1.0e-14) iR4 = iR * iR * iR * iR;
    const int ia[6] = {0, 1, 2, 0, 1, 0}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        const int a = ia[p], b = ib[p];
        const double dab = (a == b) ? 1.0 : 0.0;
        const double d2lr = dab * ir2 - 2.0 * c3[a] * c3[b] * ir4;
        const double d2ct = -(dab * z + (a == 2 ? c3[b] : 0.0) + (b == 2 ? c3[a] : 0.0)) * ir3
                            + 3.0 * z * c3[a] * c3[b] * ir5;
        const double d2th = -d2ct * isth - dct[a] * dct[b] * cos3;
        double d2ph = 0.0;
        if (a == 0 && b == 0) d2ph = 2.0 * x * y * iR4;
        else if (a == 1 && b == 1) d2ph = -2.0 * x * y * iR4;
        else if (a == 0 && b == 1) d2ph = (y * y - x * x) * iR4;
        h[p] = d.Prr * dl[a] * dl[b]
             + d.Prt * (dl[a] * dt[b] + dl[b] * dt[a])
             + d.Prf * (dl[a] * dp[b] + dl[b] * dp[a])
             + d.Ptt * dt[a] * dt[b]
             + d.Ptf * (dt[a] * dp[b] + dt[b] * dp[a])
             + d.Pff * dp[a] * dp[b]
             + d.Pr * d2lr + d.Pt * d2th + d.Pf * d2ph;
    }
}

}  // namespace agx
