// Closed-form analytic potentials (device code, FP64).
//
// Same potentials as the reference's CuPy elementwise kernels
// (_analytic_potentials.py:176-1165): NFW, Plummer, Hernquist, spherical
// Dehnen, Isochrone, Miyamoto-Nagai, logarithmic halo, DiskAnsatz and uniform
// acceleration.  Each type is one device function that returns Phi, +grad Phi
// and the Hessian together, so one pass over the particle serves every
// requested output (the reference launches one kernel per output and re-reads
// the positions each time, _analytic_potentials.py:118-130).
#pragma once
#include "agx_common.h"

namespace agx {

// Component types of the C ABI (agx_component.type).
enum {
    T_NFW = AGX_T_NFW, T_PLUMMER = AGX_T_PLUMMER, T_HERNQUIST = AGX_T_HERNQUIST,
    T_DEHNEN = AGX_T_DEHNEN, T_ISOCHRONE = AGX_T_ISOCHRONE, T_MN = AGX_T_MIYAMOTO_NAGAI,
    T_LOGHALO = AGX_T_LOGHALO, T_DISK = AGX_T_DISK_ANSATZ, T_UNIFORM = AGX_T_UNIFORM_ACCEL,
};

struct Eval3 {
    double phi, rho;
    double g[3];
    double h[6];   // [xx, yy, zz, xy, yz, xz]
};

// Spherical potential from Phi(r), Phi'(r), Phi''(r):
//   grad = Phi' x / r,  H_ij = (Phi'' - Phi'/r) x_i x_j / r^2 + Phi'/r delta_ij
// In terms of the unit vector u = x/r, so that no power of 1/r beyond the
// first is formed (at the clamped radius 1e-300 of the origin the gradient is
// 0), with the diagonal as Phi'' u_i^2 + Phi'/r (1 - u_i^2): the form
// (Phi'' - Phi'/r) u_i^2 + Phi'/r cancels where Phi'/r >> Phi'' (near a cusp).
template <bool G, bool H>
__device__ __forceinline__ void spherical_out(double x, double y, double z, double r, double P,
                                              double d1, double d2, Eval3 &o)
{
    o.phi = P;
    if (G || H) {
        const double ir = 1.0 / r, ux = x * ir, uy = y * ir, uz = z * ir;
        if (G) { o.g[0] = d1 * ux; o.g[1] = d1 * uy; o.g[2] = d1 * uz; }
        if (H) {
            const double d1r = d1 * ir, A = d2 - d1r;
            const double xx = ux * ux, yy = uy * uy, zz = uz * uz;
            o.h[0] = fma(d2, xx, d1r * (yy + zz));
            o.h[1] = fma(d2, yy, d1r * (xx + zz));
            o.h[2] = fma(d2, zz, d1r * (xx + yy));
            o.h[3] = A * ux * uy; o.h[4] = A * uy * uz; o.h[5] = A * ux * uz;
        }
    }
}

// r clamped below at 1e-300 as the reference's kernels do (`r < 1e-300 ? 1e-300
// : r`); a NaN radius passes through, so a bad position gives NaN.
__device__ __forceinline__ double clamp_r(double r) { return r < 1e-300 ? 1e-300 : r; }

// p[]: NFW {GM, rs, M}  (reference :176-243)
template <bool G, bool H, bool RHO>
__device__ __forceinline__ void nfw(const double *p, double x, double y, double z, Eval3 &o)
{
    const double GM = p[0], rs = p[1];
    const double r = clamp_r(sqrt(fma(x, x, fma(y, y, z * z))));
    const double u = r / rs, L = log1p(u), rr = r + rs, ir = 1.0 / r;
    const double P = -GM * L * ir;
    if (RHO) { o.rho = p[2] / (4.0 * M_PI * r * rr * rr); return; }
    double d1, d2;
    if (u < 0.0625) {
        // Phi' = GM/rs^2 sum_{n>=2} (-1)^n (n-1)/n u^(n-2) and
        // Phi'' = GM/rs^3 sum_{n>=3} (-1)^n (n-1)(n-2)/n u^(n-3), 17 terms each
        // (the first omitted term is below 1e-19 of the sum for u < 1/16):
        // the closed forms below lose digits as 1/u (Phi') and 1/u^2 (Phi'')
        // by cancellation, and give NaN at the origin
        double a = 17.0 / 18.0, b = -(18.0 * 17.0) / 19.0;
#pragma unroll
        for (int n = 17; n >= 2; --n) a = fma(a, -u, (double)(n - 1) / n);
#pragma unroll
        for (int n = 18; n >= 3; --n) b = fma(b, -u, -(double)((n - 1) * (n - 2)) / n);
        const double c = GM / (rs * rs);
        d1 = c * a;
        d2 = c / rs * b;
    } else {
        d1 = GM * (L * ir * ir - ir / rr);
        d2 = GM * (-2.0 * L * ir * ir * ir + ir * ir / rr + (2.0 * r + rs) * ir * ir / (rr * rr));
    }
    spherical_out<G, H>(x, y, z, r, P, d1, d2, o);
}

// p[]: Plummer {GM, b, M}  (reference :288-340)
template <bool G, bool H, bool RHO>
__device__ __forceinline__ void plummer(const double *p, double x, double y, double z, Eval3 &o)
{
    const double GM = p[0], b2 = p[1] * p[1];
    const double s2 = fma(x, x, fma(y, y, fma(z, z, b2)));
    const double is = 1.0 / sqrt(s2), is3 = is * is * is;
    if (RHO) { o.rho = 3.0 * p[2] * b2 * is3 * is * is / (4.0 * M_PI); return; }
    o.phi = -GM * is;
    const double c = GM * is3;
    if (G) { o.g[0] = c * x; o.g[1] = c * y; o.g[2] = c * z; }
    if (H) {
        const double A = -3.0 * c * is * is;
        o.h[0] = fma(A * x, x, c); o.h[1] = fma(A * y, y, c); o.h[2] = fma(A * z, z, c);
        o.h[3] = A * x * y; o.h[4] = A * y * z; o.h[5] = A * x * z;
    }
}

// p[]: Hernquist {GM, a, M}  (reference :383-444)
template <bool G, bool H, bool RHO>
__device__ __forceinline__ void hernquist(const double *p, double x, double y, double z, Eval3 &o)
{
    const double GM = p[0], a = p[1];
    const double r = clamp_r(sqrt(fma(x, x, fma(y, y, z * z))));
    const double ra = r + a, ira = 1.0 / ra;
    if (RHO) { o.rho = p[2] * a / (2.0 * M_PI * r * ra * ra * ra); return; }
    const double d1 = GM * ira * ira;
    spherical_out<G, H>(x, y, z, r, -GM * ira, d1, -2.0 * d1 * ira, o);
}

// p[]: spherical Dehnen {GM, a, gamma, M}, gamma in [0, 2]  (reference :486-549;
// gamma = 2 uses the log form -(GM/a) ln(1 + a/r))
template <bool G, bool H, bool RHO>
__device__ __forceinline__ void dehnen(const double *p, double x, double y, double z, Eval3 &o)
{
    const double GM = p[0], a = p[1], g = p[2];
    const double r = clamp_r(sqrt(fma(x, x, fma(y, y, z * z))));
    const double ra = r + a;
    if (RHO) {
        o.rho = (3.0 - g) * p[3] / (4.0 * M_PI * a * a * a) * pow(r / a, -g) * pow(1.0 + r / a, g - 4.0);
        return;
    }
    const double e = 2.0 - g;
    const double P = (e == 0.0) ? -(GM / a) * log1p(a / r)
                                : -(GM / a) * (1.0 - pow(r / ra, e)) / e;
    const double d1 = GM * pow(r, 1.0 - g) * pow(ra, g - 3.0);
    const double d2 = d1 * ((1.0 - g) / r + (g - 3.0) / ra);
    spherical_out<G, H>(x, y, z, r, P, d1, d2, o);
}

// p[]: Isochrone {GM, b, M}  (reference :603-676)
template <bool G, bool H, bool RHO>
__device__ __forceinline__ void isochrone(const double *p, double x, double y, double z, Eval3 &o)
{
    const double GM = p[0], b = p[1];
    const double r2 = fma(x, x, fma(y, y, z * z));
    const double s = sqrt(r2 + b * b), bs = b + s;
    const double d1r = GM / (s * bs * bs);                           // Phi'/r
    const double A = -GM * (b + 3.0 * s) / (s * s * s * bs * bs * bs); // (Phi'' - Phi'/r)/r^2
    if (RHO) { o.rho = (A * r2 + 3.0 * d1r) * (1.0 / (4.0 * M_PI * G_AGAMA_D)); return; }
    o.phi = -GM / bs;
    if (G) { o.g[0] = d1r * x; o.g[1] = d1r * y; o.g[2] = d1r * z; }
    if (H) {
        o.h[0] = fma(A * x, x, d1r); o.h[1] = fma(A * y, y, d1r); o.h[2] = fma(A * z, z, d1r);
        o.h[3] = A * x * y; o.h[4] = A * y * z; o.h[5] = A * x * z;
    }
}

// p[]: Miyamoto-Nagai {GM, a, b, M}  (reference :720-793)
template <bool G, bool H, bool RHO>
__device__ __forceinline__ void miyamoto_nagai(const double *p, double x, double y, double z,
                                               Eval3 &o)
{
    const double GM = p[0], a = p[1], b = p[2];
    const double R2 = fma(x, x, y * y);
    const double zeta = sqrt(fma(z, z, b * b)), az = a + zeta;
    const double S2 = fma(az, az, R2), iS = 1.0 / sqrt(S2), iS3 = iS * iS * iS, iS5 = iS3 * iS * iS;
    if (RHO) {
        o.rho = p[3] * b * b / (4.0 * M_PI) * fma(a, R2, (a + 3.0 * zeta) * az * az) * iS5
                / (zeta * zeta * zeta);
        return;
    }
    o.phi = -GM * iS;
    const double izeta = 1.0 / zeta, fz = az * z * izeta;            // d(S^2/2)/dz
    if (G) { o.g[0] = GM * x * iS3; o.g[1] = GM * y * iS3; o.g[2] = GM * fz * iS3; }
    if (H) {
        o.h[0] = GM * (iS3 - 3.0 * x * x * iS5);
        o.h[1] = GM * (iS3 - 3.0 * y * y * iS5);
        o.h[2] = GM * ((1.0 + a * b * b * izeta * izeta * izeta) * iS3 - 3.0 * fz * fz * iS5);
        o.h[3] = -3.0 * GM * x * y * iS5;
        o.h[4] = -3.0 * GM * y * fz * iS5;
        o.h[5] = -3.0 * GM * x * fz * iS5;
    }
}

// p[]: logarithmic halo {v0^2, rc^2, p^2, q^2}  (reference :842-892)
template <bool G, bool H, bool RHO>
__device__ __forceinline__ void loghalo(const double *p, double x, double y, double z, Eval3 &o)
{
    const double v2 = p[0], rc2 = p[1], ip2 = 1.0 / p[2], iq2 = 1.0 / p[3];
    const double m2 = rc2 + x * x + y * y * ip2 + z * z * iq2, im2 = 1.0 / m2, im4 = im2 * im2;
    const double gx = x, gy = y * ip2, gz = z * iq2;
    if (RHO) {
        const double lap = v2 * ((1.0 + ip2 + iq2) * im2 - 2.0 * (gx * gx + gy * gy + gz * gz) * im4);
        o.rho = lap * (1.0 / (4.0 * M_PI * G_AGAMA_D));
        return;
    }
    o.phi = 0.5 * v2 * log(m2);
    const double c = v2 * im2;
    if (G) { o.g[0] = c * gx; o.g[1] = c * gy; o.g[2] = c * gz; }
    if (H) {
        const double k = -2.0 * v2 * im4;
        o.h[0] = fma(k * gx, gx, c); o.h[1] = fma(k * gy, gy, c * ip2); o.h[2] = fma(k * gz, gz, c * iq2);
        o.h[3] = k * gx * gy; o.h[4] = k * gy * gz; o.h[5] = k * gx * gz;
    }
}

// p[]: DiskAnsatz {4 pi G Sigma, R_d, h, r_in}: Phi = f(r) H(z),
//   f = 4 pi G Sigma exp(-r/R_d - r_in/r)
//   H = h/2 (exp(-|z|/h) - 1 + |z|/h)  (h > 0, exponential),
//       |h|/2 ln cosh(z/|h|)            (h < 0, isothermal),
//       |z|/2                           (h = 0, thin)
// (reference :950-1052 uses the exponential H for every sign of h).  Near
// the midplane H ~ z^2: both profiles and their slopes are formed without the
// cancellation of the terms of order 1 (the reference's exp(-u) - 1 + u keeps
// only 1e-16 / u^2 of relative precision), from expm1 and a series.
__device__ __forceinline__ double exp_m1_plus(double u)     // e^-u - 1 + u, u >= 0
{
    if (u >= 0.5) return u + expm1(-u);
    // u^2 sum_k (-u)^k / (k+2)!, to k = 14 (0.5^15 / 17! ~ 1e-19)
    double p = 1.0 / 355687428096000.0;                    // 1/16!
    const double c[14] = {1.0 / 20922789888000.0, 1.0 / 1307674368000.0, 1.0 / 87178291200.0,
                          1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0,
                          1.0 / 3628800.0, 1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0,
                          1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0};
#pragma unroll
    for (int k = 0; k < 14; ++k) p = fma(-u, p, c[k]);
    return u * u * fma(-u, p, 0.5);
}

template <bool G, bool H, bool RHO>
__device__ __forceinline__ void disk_ansatz(const double *p, double x, double y, double z, Eval3 &o)
{
    const double A4 = p[0], Rd = p[1], hz = p[2], rin = p[3];
    const double r = clamp_r(sqrt(fma(x, x, fma(y, y, z * z)))), ir = 1.0 / r;
    const double f = A4 * exp(-r / Rd - rin * ir);
    const double fo = -1.0 / Rd + rin * ir * ir;
    const double f1 = f * fo, f2 = f * (fo * fo - 2.0 * rin * ir * ir * ir);
    double Hv, H1, H2;
    const double az = fabs(z), sg = (z > 0.0) ? 1.0 : ((z < 0.0) ? -1.0 : 0.0);
    if (hz > 0.0) {
        const double u = az / hz, em = expm1(-u);            // e^-u - 1
        Hv = 0.5 * hz * exp_m1_plus(u); H1 = -0.5 * sg * em; H2 = 0.5 * (1.0 + em) / hz;
    } else if (hz < 0.0) {
        const double h = -hz, u = az / h, em = expm1(-2.0 * u);
        // ln cosh(u): ln(1 + 2 sinh^2(u/2)) near 0, u + ln(1 + e^{-2u}) - ln 2
        // beyond (no overflow for large u)
        if (u < 1.0) {
            const double sh = 0.5 * (expm1(0.5 * u) - expm1(-0.5 * u));
            Hv = 0.5 * h * log1p(2.0 * sh * sh);
        } else {
            Hv = 0.5 * h * (u + log1p(1.0 + em) - M_LN2);
        }
        const double th = -em / (2.0 + em);                  // tanh(u)
        H1 = 0.5 * sg * th; H2 = 0.5 * (1.0 - th * th) / h;
    } else {
        Hv = 0.5 * az; H1 = 0.5 * sg; H2 = 0.0;
    }
    const double zr = z * ir;
    if (RHO) {
        const double lap = Hv * (f2 + 2.0 * f1 * ir) + 2.0 * f1 * H1 * zr + f * H2;
        o.rho = lap * (1.0 / (4.0 * M_PI * G_AGAMA_D));
        return;
    }
    o.phi = f * Hv;
    const double c = Hv * f1 * ir;
    if (G) { o.g[0] = c * x; o.g[1] = c * y; o.g[2] = fma(c, z, H1 * f); }
    if (H) {
        const double xr = x * ir, yr = y * ir;
        const double A = Hv * (f2 - f1 * ir);
        o.h[0] = fma(A * xr, xr, c); o.h[1] = fma(A * yr, yr, c);
        o.h[2] = Hv * (f2 * zr * zr + f1 * (1.0 - zr * zr) * ir) + 2.0 * f1 * H1 * zr + f * H2;
        o.h[3] = A * xr * yr;
        o.h[4] = fma(A * yr, zr, f1 * H1 * yr);
        o.h[5] = fma(A * xr, zr, f1 * H1 * xr);
    }
}

// p[]: uniform acceleration {ax, ay, az}: Phi = -(a . x)  (reference :1112-1165)
template <bool G, bool H, bool RHO>
__device__ __forceinline__ void uniform_accel(const double *p, double x, double y, double z, Eval3 &o)
{
    if (RHO) { o.rho = 0.0; return; }
    o.phi = -(p[0] * x + p[1] * y + p[2] * z);
    if (G) { o.g[0] = -p[0]; o.g[1] = -p[1]; o.g[2] = -p[2]; }
    if (H) { for (int k = 0; k < 6; ++k) o.h[k] = 0.0; }
}

template <bool G, bool H, bool RHO>
__device__ __forceinline__ bool eval_analytic(int type, const double *p, double x, double y,
                                              double z, Eval3 &o)
{
    switch (type) {
    case T_NFW: nfw<G, H, RHO>(p, x, y, z, o); return true;
    case T_PLUMMER: plummer<G, H, RHO>(p, x, y, z, o); return true;
    case T_HERNQUIST: hernquist<G, H, RHO>(p, x, y, z, o); return true;
    case T_DEHNEN: dehnen<G, H, RHO>(p, x, y, z, o); return true;
    case T_ISOCHRONE: isochrone<G, H, RHO>(p, x, y, z, o); return true;
    case T_MN: miyamoto_nagai<G, H, RHO>(p, x, y, z, o); return true;
    case T_LOGHALO: loghalo<G, H, RHO>(p, x, y, z, o); return true;
    case T_DISK: disk_ansatz<G, H, RHO>(p, x, y, z, o); return true;
    case T_UNIFORM: uniform_accel<G, H, RHO>(p, x, y, z, o); return true;
    default: return false;
    }
}

}  // namespace agx
