// Multipole (spherical-harmonic BFE) device code for gfx950, FP64 (shared by
// the Multipole kernels and the fused composite kernel).
//
// Same potential as the reference's multipole_*_kernel
// (_multipole_potential_kernel.cu:140-958): quintic splines in ln r per
// active (l,m) harmonic with Agama log-scaling, l=0 power/log-law
// extrapolation inside r_min and beyond r_max, Agama-normalised associated
// Legendre functions times cos/sin(|m| phi), Cartesian gradient and Hessian by
// the chain rule.  The arithmetic is arranged for CDNA4 instead of following
// the reference line by line:
//
//  * The Legendre recurrence runs in Agama's normalisation directly, with the
//    per-(l,m) constants alpha_l, beta_l and the P_mm seed precomputed on the
//    host (one FMA pair per step; no sqrt or division per step).
//  * Terms are grouped by |m|: inside a group cos(m phi) and sin(m phi) are
//    constants, so each term accumulates three products into per-group sums
//    (V = sum C Q, R = sum C' Q, T = sum C dQ/dtheta) and the trigonometric
//    factors are applied once per group.  The phi-derivative needs no extra
//    per-term work: it is m(-sin V_cos + cos V_sin) per group.
//  * With log-scaling every l>0 radial function is C_sc * Phi_0, so the
//    monopole factor and its derivatives multiply the group sums once at the
//    end instead of once per term.
//  * Powers s^j and the derivative weights j s^(j-1) are formed once per
//    particle; a term costs 5 FMA for C and 4 for C'.
//  * Table layout is interval-major [k][c][6]: particles in one radial
//    interval read one contiguous slab (scalar loads in the binned kernel).
#pragma once
#include <cmath>
#include "sph_dev.h"

namespace agx {

struct MPK {
    const double *poly;              // [n_int][n_lm][6]
    const void *prog;                // Legendre program records (agx_multipole.program)
    int n_int, n_lm, n_groups, n_steps, log_scaling, inner_log, outer_log;
    double logr_min, dlogr, inv_dlogr, invPhi0;
    double inner_s, inner_U, inner_W, outer_s, outer_U, outer_W;
    double inv_rmin, inv_rmax;       // 1 / r at the first and last node
    double rmax;                     // r at the last node
    const double *law;               // NULL or device [invPhi0, inner U, W, outer U, W]
    const double *origin;            // NULL or device [x0, y0, z0]: evaluated at x - origin
    double ox, oy, oz;               // the origin (0 without one), set by with_law
};

// The table with its five law scalars taken from device memory when it
// carries them there (agx_multipole.law), and its origin; called once per
// kernel or descriptor.
__device__ __forceinline__ MPK with_law(MPK P)
{
    if (P.law) {
        P.invPhi0 = cld(P.law, 0);
        P.inner_U = cld(P.law, 1); P.inner_W = cld(P.law, 2);
        P.outer_U = cld(P.law, 3); P.outer_W = cld(P.law, 4);
    }
    P.ox = P.oy = P.oz = 0.0;
    if (P.origin) { P.ox = cld(P.origin, 0); P.oy = cld(P.origin, 1); P.oz = cld(P.origin, 2); }
    return P;
}

// Position i of an AoS (n, 3) array relative to the table's origin (x - 0 is
// x exactly, so a table without one sees the positions unchanged).
__device__ __forceinline__ void load_pos(const MPK &P, const double *xyz, int64_t i, double &x,
                                         double &y, double &z)
{
    x = xyz[3 * i] - P.ox;
    y = xyz[3 * i + 1] - P.oy;
    z = xyz[3 * i + 2] - P.oz;
}


// l=0 extrapolation law and its first two ln r derivatives.
//   log mode:   Phi = (W + U t) e^{v t}
//   power mode: Phi = W e^{v t} + U e^{s t}
// t = ln(r / r_edge), er = r / r_edge = e^t and ier = 1 / er: the exponents
// that occur in practice (0, -1, 2) are powers of er, which avoids the
// exponential.
__device__ __forceinline__ double epow(double a, double t, double er, double ier)
{
    if (a == 0.0) return 1.0;
    if (a == -1.0) return ier;
    if (a == 2.0) return er * er;
    return agx_exp(a * t);
}

__device__ __forceinline__ void law(double v, double s, double U, double W, int logm, double t,
                                    double er, double ier, double &P, double &D, double &D2)
{
    double ev = epow(v, t, er, ier);
    if (logm) {
        double q = fma(U, t, W);
        P = q * ev;
        D = fma(v, q, U) * ev;
        D2 = fma(v * v, q, 2.0 * v * U) * ev;
    } else {
        double es = epow(s, t, er, ier);
        double wv = W * ev, us = U * es;
        P = wv + us;
        D = fma(v, wv, s * us);
        D2 = fma(v * v, wv, s * s * us);
    }
}

// Radial/extrapolated part: fills `o` for points with a purely radial l=0
// potential (inside r_min or beyond r_max).  dP, d2P are d/d ln r; ir = 1/r.
template <int ORD>
__device__ __forceinline__ void radial_out(double x, double y, double z, double ir,
                                           double P, double D, double D2, PointOut &o)
{
    o.phi = P;
    if (ORD == ORD_GRAD || ORD == ORD_HESS) {
        double ir2 = ir * ir, g = D * ir2;
        o.g[0] = g * x; o.g[1] = g * y; o.g[2] = g * z;
        if (ORD == ORD_HESS) {
            // H_ij = (D2 - 2D) x_i x_j / r^4 + D delta_ij / r^2 for a radial Phi
            double A = (D2 - 2.0 * D) * ir2 * ir2, B = g;
            o.h[0] = fma(A * x, x, B); o.h[1] = fma(A * y, y, B); o.h[2] = fma(A * z, z, B);
            o.h[3] = A * x * y; o.h[4] = A * y * z; o.h[5] = A * x * z;
        }
    }
}

// (dPhi/dphi) / sin(theta) on the z axis from the m = 1 terms' theta
// derivatives S: Y / sin(theta) -> dY/dtheta at the north pole (cos theta = 1)
// and -dY/dtheta at the south pole, where sin(theta) ~ pi - theta.  (The
// reference takes +dY/dtheta at both, _multipole_potential_kernel.cu:332-334,
// which reverses the m = 1 part of the force at points (0, 0, z < 0).)
__device__ __forceinline__ double pole_limit(double S, double ct) { return ct < 0.0 ? -S : S; }

// Per-term hook of interior(): term() sees every step of the Legendre
// recurrence (cos/sin slab offsets, -1 when absent, P_l^m-scaled value and
// theta derivative, m, cos/sin(m phi), 1/sin(theta), pole flag) and sums()
// the l > 0 sums before the log-scaling (value, d/ds, d/dtheta, d/dphi with
// its pole limit).  The adjoint kernel scatters table gradients through it.
struct NoHook {
    __device__ void term(int, int, double, double, int, double, double, double, bool) const {}
    __device__ void sums(double, double, double, double, double) const {}
};

// The triaxial program of even degree L (a table with every harmonic of even
// l <= L and even m in 0..l, cos terms only, columns in program order: the
// shape of every table with triaxial symmetry and no pruned term): group g
// has m = 2g and the steps l = 2g, 2g + 2, ..., L; step j of the program
// has its cos term in column j (step 0, l = m = 0, none: the monopole is
// evaluated apart).  tri_step0(L, g) is the first step of group g.
__host__ __device__ constexpr int tri_step0(int L, int g)
{
    return g * (L / 2 + 1) - g * (g - 1) / 2;
}

// The full program of degree L (a table with every harmonic up to l = L, cos
// and sin, columns in program order: the shape of every table without
// symmetry and no pruned term): group g has m = g and the steps l = g, g + 1,
// .., L; a step j <= L (group 0) has its cos term in column j (step 0, l = 0,
// none), a later step j its cos term in column 2j - L - 1 and its sin term in
// column 2j - L.  full_step0(L, g) is the first step of group g.
__host__ __device__ constexpr int full_step0(int L, int g)
{
    return g * (L + 1) - g * (g - 1) / 2;
}

// Interior evaluation on radial interval k (slab = poly + k*n_lm*6).
// UNI: the slab address is wave-uniform (scalar loads of coefficients).
// STRIDE: 1 or 2 when every group of the program that recurs steps by that
// stride (AGX_PROG_STRIDE1 / AGX_PROG_STRIDE2), 0 for any program; SIN false
// when the program has no sin terms (AGX_PROG_COS_ONLY).
// TL > 0: the program is the triaxial program of degree TL (tri_step0), FL > 0
// the full program of degree FL (full_step0): its groups, steps and column
// offsets are compile-time constants, so the group and step loops unroll
// completely; the records supply only the recurrence coefficients and q_mm.
// The slab must then be in LDS (the loads are LDS loads at immediate offsets).
template <int ORD, bool UNI, class Hook = NoHook, int STRIDE = 0, bool SIN = true, int TL = 0,
          int FL = 0>
__device__ __forceinline__ void interior(const MPK &P, const double *slab, double x, double y,
                                         double z, double r2, double r, double ir, double lnr,
                                         int k, PointOut &o, double inv4piG,
                                         const Hook &hook = Hook{})
{
    constexpr bool R1 = ORD != ORD_POT;                      // d/d ln r
    constexpr bool D1 = ORD == ORD_GRAD || ORD == ORD_HESS;  // angular first derivatives
    constexpr bool D2 = ORD == ORD_HESS || ORD == ORD_DENS;  // d^2/d ln r^2
    constexpr bool ANG2 = ORD == ORD_HESS;                   // second angular derivatives
    const double idl = P.inv_dlogr;
    double s = (lnr - fma((double)k, P.dlogr, P.logr_min)) * idl;
    s = fmin(fmax(s, 0.0), 1.0);

    auto load6 = [&](int off, double a[6]) {
        if (UNI) {
            cdvec2_t *p = (cdvec2_t *)(slab + off);
            dvec2 q0 = p[0], q1 = p[1], q2 = p[2];
            a[0] = q0.x; a[1] = q0.y; a[2] = q1.x; a[3] = q1.y; a[4] = q2.x; a[5] = q2.y;
        } else {
            const dvec2 *p = (const dvec2 *)(slab + off);
            dvec2 q0 = p[0], q1 = p[1], q2 = p[2];
            a[0] = q0.x; a[1] = q0.y; a[2] = q1.x; a[3] = q1.y; a[4] = q2.x; a[5] = q2.y;
        }
    };
    // C, C' and C'' of the quintic in s by Horner's scheme with synthetic
    // division: 5 + 4 (+ 3) FMA, and only s itself is held in registers
    // (powers of s and their derivative weights would take 16-22 VGPRs)
    auto horner = [&](const double a[6], double &C, double &dC, double &d2C) {
        double b = a[5], c = 0.0, d = 0.0;
#pragma unroll
        for (int q = 4; q >= 0; --q) {
            if (D2 && q <= 3) d = (q == 3) ? c : fma(d, s, c);
            if (R1 && q <= 4) c = (q == 4) ? b : fma(c, s, b);
            b = fma(b, s, a[q]);
        }
        C = b;
        if (R1) dC = c;
        if (D2) d2C = 2.0 * d;
    };

    // ---- angular basis ------------------------------------------------------
    // R^2 below 2^-767 (R < 4e-116, far below 1e-10 r for any interior point)
    // is taken as the axis: the pole branch, whose result does not depend on
    // cos/sin(phi)
    const double R2 = fma(x, x, y * y);
    double R = 0.0, iR = 0.0, cph = 1.0, sph = 0.0;
    if (R2 >= 0x1p-767) { sqrt_rsqrt(R2, R, iR); cph = x * iR; sph = y * iR; }
    const double ct = z * ir, st = R * ir;
    const bool pole = st < 1.0e-10;
    // cos^2 theta and its theta derivative, once per point when every
    // recurring group steps by 2
    // (unrolled, TL > 0: cos^2 kept opaque, or each step's cos^2 + beta
    // contracts into an FMA of cos * cos whose accumulator, beta, must first
    // be copied from scalar registers: two moves per step)
    double ct2 = STRIDE == 2 ? ct * ct : 0.0;
    if (TL > 0) ct2 = vopaque(ct2);
    const double dct2 = STRIDE == 2 ? -2.0 * ct * st : 0.0;
    // 1/sin(theta) = r/R, needed inside the term loop only for the Hessian
    const double ist_loop = (ANG2 && !pole) ? r * iR : 0.0;

    // scaled sums over all l>0 terms (trig applied)
    double Sv = 0.0, Sr = 0.0, St = 0.0, Sfn = 0.0, Sfp = 0.0;
    double Srr = 0.0, Stt = 0.0, Srt = 0.0, Sf = 0.0, Sff = 0.0, Srf = 0.0, Stf = 0.0, SL = 0.0;

    double cm = 1.0, sm = 0.0, sp = 1.0, sp1 = 0.0;   // cos/sin(m phi), sin^m, sin^(m-1)
    int cur_m = 0;
    // Program records are read one iteration ahead (the zero records after
    // the groups and after the steps keep the look-ahead in bounds), so the
    // scalar-load latency overlaps the current step's arithmetic.
    constexpr bool CT = TL > 0 || FL > 0;
    constexpr int kUnrollGroups = TL > 0 ? TL / 2 + 1 : FL > 0 ? FL + 1 : 1;
    constexpr int kUnrollPairs = TL > 0 ? TL / 4 + 1 : FL > 0 ? FL / 2 + 1 : 1;
    const int n_groups = CT ? kUnrollGroups : P.n_groups;
    const int S0 = n_groups + 1;
    // Unrolled (CT), the records' addresses are constants: the pointer is
    // re-derived per point, so that their loads stay here instead of being
    // hoisted out of the caller's particle loop, and again after every step
    // (in terms()), so that each record is read one step ahead and not all
    // of them at the start.
    const void *prog = CT ? sopaque(P.prog) : P.prog;
    static_assert(!(TL > 0 && FL > 0) && (!CT || !UNI), "one compile-time program, LDS slabs");
    static_assert(TL == 0 || (STRIDE == 2 && !SIN), "the triaxial program: stride 2, cos terms");
    static_assert(FL == 0 || (STRIDE == 1 && SIN), "the full program: stride 1, cos and sin terms");
    typedef __attribute__((address_space(3))) const double *lds_dptr;
    lds_dptr lslab = CT ? (lds_dptr)slab : nullptr;
    ivec8 gr = cld_rec(prog, 0);
#pragma unroll kUnrollGroups
    for (int g = 0; g < n_groups; ++g) {
        const ivec8 gn = cld_rec(prog, g + 1);
        const int m = TL > 0 ? 2 * g : FL > 0 ? g : gr.s0;
        if (m - cur_m == 2) {
            // two steps at once (tables with even-m symmetry): rotate by 2 phi
            const double c2 = fma(cph, cph, -sph * sph), s2 = 2.0 * cph * sph;
            sp1 = sp * st; sp = sp1 * st;
            const double t = fma(cm, c2, -sm * s2);
            sm = fma(sm, c2, cm * s2);
            cm = t;
            cur_m = m;
        }
        while (cur_m < m) {
            sp1 = sp; sp *= st;
            double t = fma(cm, cph, -sm * sph);
            sm = fma(sm, cph, cm * sph);
            cm = t;
            ++cur_m;
        }
        const double q = rec_lo(gr), fm = CT ? (double)m : rec_hi(gr);   // qmm and m
        // Two recurrence buffers A and B are used alternately: each step
        // writes the new P_l over the older of the two, so no register copies
        // are needed.  With scalar (UNI) coefficient loads the cos-term
        // coefficients of the next step are loaded into the other buffer (in
        // SGPRs) while this step computes; per-lane loads are issued at their
        // use, where prefetching costs more VGPRs than it saves in latency.
        double QA = q * sp, dQA = 0.0, QB = 0.0, dQB = 0.0;
        if (D1) dQA = fm * q * sp1 * ct;                     // 0 for m = 0

        double Vc = 0.0, Vs = 0.0, Rc = 0.0, Rs = 0.0, Tc = 0.0, Ts = 0.0;
        double RRc = 0.0, RRs = 0.0, TTc = 0.0, TTs = 0.0, RTc = 0.0, RTs = 0.0, Lc = 0.0, Ls = 0.0;
        auto add_term = [&](const double a[6], double Q, double dQ, double d2Q, double ll1, double &V,
                            double &Rr, double &Tt, double &RRr, double &TTt, double &RTt, double &LL) {
            double C, dC = 0.0, d2C = 0.0;
            horner(a, C, dC, d2C);
            V = fma(C, Q, V);
            if (R1) {
                Rr = fma(dC, Q, Rr);
                if (D1) Tt = fma(C, dQ, Tt);
                if (ANG2) RTt = fma(dC, dQ, RTt);
            }
            if (D2) RRr = fma(d2C, Q, RRr);
            if (ANG2) TTt = fma(C, d2Q, TTt);
            if (ORD == ORD_DENS) LL = fma(ll1 * C, Q, LL);
        };
        // the cos and sin terms of one step (degree rec.s0, current Q, dQ);
        // the cos coefficients were loaded into ca beforehand
        // Steps j0 .. j1-1 (the first has l = m); coff: the cos offset of step
        // j (record rec), -1 when it has no cos term
        const int j0 = TL > 0 ? tri_step0(TL, g) : FL > 0 ? full_step0(FL, g) : gr.s1;
        const int j1 = TL > 0 ? tri_step0(TL, g + 1) : FL > 0 ? full_step0(FL, g + 1) : gr.s2;
        auto coff = [&](const ivec8 &rec, int j) {
            if (TL > 0) return j > 0 ? 6 * j : -1;
            if (FL > 0) return j == 0 ? -1 : j <= FL ? 6 * j : 6 * (2 * j - FL - 1);
            return rec_off_lo(rec);
        };
        auto soff = [&](const ivec8 &rec, int j) {
            if (TL > 0) return -1;
            if (FL > 0) return j <= FL ? -1 : 6 * (2 * j - FL);
            return rec_off_hi(rec);
        };
        auto terms = [&](const ivec8 &rec, int j, double Q, double dQ, const double ca[6]) {
            const int l = CT ? m + (TL > 0 ? 2 : 1) * (j - j0) : rec.s0;
            const int oc = coff(rec, j), os = soff(rec, j);
            double d2Q = 0.0;
            if (ANG2 && !pole)
                d2Q = -ct * ist_loop * dQ -
                      ((double)(l * (l + 1)) - (double)(m * m) * ist_loop * ist_loop) * Q;
            const double ll1 = (double)(l * (l + 1));
            if (CT) {
                // the slab is in LDS; the next step's loads depend on this
                // step's sums, so that the unrolled steps do not all load
                // first (every step's coefficients live at once: spills)
                auto lds6 = [&](int off, double a[6]) {
                    ldvec2_t *p = (ldvec2_t *)(lslab + off);
                    const dvec2 q0 = p[0], q1 = p[1], q2 = p[2];
                    a[0] = q0.x; a[1] = q0.y; a[2] = q1.x; a[3] = q1.y; a[4] = q2.x; a[5] = q2.y;
                };
                auto tie = [&](double V, double Rr, double Tt, double RRr, double TTt, double RTt,
                               double LL) {
                    asm volatile("" : "+v"(lslab), "+s"(prog) : "v"(V));
                    if (R1) asm volatile("" : "+v"(lslab) : "v"(Rr));
                    if (D1) asm volatile("" : "+v"(lslab) : "v"(Tt));
                    if (D2) asm volatile("" : "+v"(lslab) : "v"(RRr));
                    if (ANG2) asm volatile("" : "+v"(lslab) : "v"(TTt), "v"(RTt));
                    if (ORD == ORD_DENS) asm volatile("" : "+v"(lslab) : "v"(LL));
                };
                double ac[6], as[6];
                if (oc >= 0) lds6(oc, ac);
                if (SIN && os >= 0) lds6(os, as);
                if (oc >= 0) add_term(ac, Q, dQ, d2Q, ll1, Vc, Rc, Tc, RRc, TTc, RTc, Lc);
                if (SIN && os >= 0) add_term(as, Q, dQ, d2Q, ll1, Vs, Rs, Ts, RRs, TTs, RTs, Ls);
                if (oc >= 0) tie(Vc, Rc, Tc, RRc, TTc, RTc, Lc);
                if (SIN && os >= 0) tie(Vs, Rs, Ts, RRs, TTs, RTs, Ls);

            } else if (oc >= 0) {
                if (UNI) {
                    add_term(ca, Q, dQ, d2Q, ll1, Vc, Rc, Tc, RRc, TTc, RTc, Lc);
                } else {
                    double ac[6];
                    load6(oc, ac);
                    add_term(ac, Q, dQ, d2Q, ll1, Vc, Rc, Tc, RRc, TTc, RTc, Lc);
                }
            }
            if (!CT && SIN && os >= 0) {
                double as[6];
                load6(os, as);
                add_term(as, Q, dQ, d2Q, ll1, Vs, Rs, Ts, RRs, TTs, RTs, Ls);
            }
            hook.term(oc, os, Q, dQ, m, cm, sm, ist_loop, pole);
        };
        // One step of stride s (1, or 2 when the group's terms all have l - m
        // even): Qo holds P_{l-2s} and becomes P_l, Qn is P_{l-s}, and with
        // c = cos (s = 1) or cos^2 (s = 2) and t = alpha (c + beta)
        //   P_l = t P_{l-s} + gamma P_{l-2s},
        //   dP_l/dtheta = t dP_{l-s} + alpha (dc/dtheta) P_{l-s} + gamma dP_{l-2s};
        // beta is 0 where s = 1.
        // for a program of mixed strides, c and dc/dtheta per group from
        // h = s - 1 by arithmetic rather than a select, so that no cos^2 or
        // cos sin stays live across the group loop
        double cc = ct, dcc = -st;
        if (STRIDE == 2) {
            cc = ct2;
            dcc = dct2;
        } else if (STRIDE == 0) {
            const double h = (double)(gr.s3 - 1);
            const double u = fma(ct, h, 1.0 - h);              // cos or 1
            cc = ct * u;
            dcc = -st * fma(h, ct, u);
        }
        auto recur = [&](const ivec8 &rec, double &Qo, double &dQo, double Qn, double dQn) {
            const double al = rec_lo(rec), be = rec_hi(rec), ga = rec_mid(rec);
            const double t = STRIDE != 1 ? al * (cc + be) : al * cc;
            if (D1) dQo = fma(t, dQn, fma(al * dcc, Qn, ga * dQo));
            Qo = fma(t, Qn, ga * Qo);
        };
        // Records past the group are read (the next group's, or the zero
        // record) but not used.
        double cA[6], cB[6];
        const ivec8 r0 = cld_rec(prog, S0 + j0);
        if (UNI && coff(r0, j0) >= 0) load6(coff(r0, j0), cA);
        ivec8 r1 = cld_rec(prog, S0 + j0 + 1);
        if (UNI && j0 + 1 < j1 && coff(r1, j0 + 1) >= 0) load6(coff(r1, j0 + 1), cB);
        terms(r0, j0, QA, dQA, cA);
        int j = j0 + 1;
#pragma unroll kUnrollPairs
        for (; j + 1 < j1; j += 2) {
            const ivec8 r2 = cld_rec(prog, S0 + j + 1);
            if (UNI && coff(r2, j + 1) >= 0) load6(coff(r2, j + 1), cA);
            recur(r1, QB, dQB, QA, dQA);
            terms(r1, j, QB, dQB, cB);
            const ivec8 r3 = cld_rec(prog, S0 + j + 2);
            if (UNI && j + 2 < j1 && coff(r3, j + 2) >= 0) load6(coff(r3, j + 2), cB);
            recur(r2, QA, dQA, QB, dQB);
            terms(r2, j + 1, QA, dQA, cA);
            r1 = r3;
        }
        if (j < j1) {
            recur(r1, QB, dQB, QA, dQA);
            terms(r1, j, QB, dQB, cB);
        }
        // ---- apply cos/sin(m phi) to the group sums ------------------------
        // acc + cos C + sin S, and cos S - sin C; without SIN (no sin terms in
        // the program) the S parts are zero and left out
        auto comb = [&](double c, double sv, double acc) {
            return SIN ? fma(cm, c, fma(sm, sv, acc)) : fma(cm, c, acc);
        };
        auto rot = [&](double sv, double c) { return SIN ? fma(cm, sv, -sm * c) : -sm * c; };
        Sv = comb(Vc, Vs, Sv);
        if (R1) Sr = comb(Rc, Rs, Sr);
        if (D1) {
            // every phi-derivative term carries the factor m (0 for m = 0,
            // where also sin(m phi) = 0), so no test of m is needed
            St = comb(Tc, Ts, St);
            const double Fv = fm * rot(Vs, Vc);                     // dPhi/dphi (group)
            Sfn += Fv;
            // pole limit of Y/sin: the m = 1 group only (a scalar 0/1 weight)
            Sfp = fma(m == 1 ? 1.0 : 0.0, rot(Ts, Tc), Sfp);
            if (ANG2) {
                Sf += Fv;
                Sff -= fm * fm * (SIN ? fma(cm, Vc, sm * Vs) : cm * Vc);
                Srf += fm * rot(Rs, Rc);
                Stf += fm * rot(Ts, Tc);
            }
        }
        if (D2) Srr = comb(RRc, RRs, Srr);
        if (ANG2) {
            Stt = comb(TTc, TTs, Stt);
            Srt = comb(RTc, RTs, Srt);
        }
        if (ORD == ORD_DENS) SL = comb(Lc, Ls, SL);
        gr = gn;
        // unrolled: the running sums formed here, at the group's end (left to
        // itself the compiler sinks every group's combination to the end of
        // the program, holding each group's sums and cos/sin(m phi) until
        // then: spills)
        if (CT) {
            asm volatile("" : "+v"(Sv));
            if (R1) asm volatile("" : "+v"(Sr));
            if (D1) asm volatile("" : "+v"(St), "+v"(Sfn), "+v"(Sfp));
            if (D2) asm volatile("" : "+v"(Srr));
            if (ANG2) asm volatile("" : "+v"(Stt), "+v"(Srt), "+v"(Sf), "+v"(Sff), "+v"(Srf), "+v"(Stf));
            if (ORD == ORD_DENS) asm volatile("" : "+v"(SL));
        }
    }

    hook.sums(Sv, Sr, St, Sfn, Sfp);
    // ---- monopole with the log-scaling inverse (after the term loop, so its
    // values are not held in registers across it) ----------------------------
    double a0[6];
    load6(0, a0);
    double C0, dC0 = 0.0, d2C0 = 0.0;
    horner(a0, C0, dC0, d2C0);
    if (R1) dC0 *= idl;
    if (D2) d2C0 *= idl * idl;
    double lam = 1.0, mu = 0.0, nu = 0.0;
    if (P.log_scaling) {
        double ex = agx_exp(C0);
        double P0 = rcp_nr(P.invPhi0 - ex);
        double dPdX = P0 * P0 * ex;
        if (D2) d2C0 = dPdX * fma(dC0 * dC0, P0 * (P.invPhi0 + ex), d2C0);
        if (R1) dC0 = dPdX * dC0;
        C0 = P0;
        lam = C0; mu = dC0; nu = d2C0;
    }

    // ---- spherical derivatives of the full potential -------------------------
    const double Pv = fma(lam, Sv, C0);
    if (ORD == ORD_POT) { o.phi = Pv; return; }
    const double Pr = fma(lam * idl, Sr, fma(mu, Sv, dC0));
    if (ORD == ORD_DENS) {
        const double Prr = fma(lam * idl * idl, Srr, fma(2.0 * mu * idl, Sr, fma(nu, Sv, d2C0)));
        o.phi = inv4piG * (Prr + Pr - lam * SL) / r2;
        return;
    }
    o.phi = Pv;
    const double ist = pole ? 0.0 : r * iR;
    if (ORD == ORD_GRAD) {
        // g = (Pr u_r + Pt u_theta + Pfos u_phi) / r with the unit vectors from
        // the angles, so x, y, z are not needed after the term loop
        const double Pt = lam * St, Pf = lam * (pole ? pole_limit(Sfp, ct) : Sfn * ist);
        const double a = fma(Pr, st, Pt * ct);
        o.g[0] = ir * fma(a, cph, -Pf * sph);
        o.g[1] = ir * fma(a, sph, Pf * cph);
        o.g[2] = ir * fma(Pr, ct, -Pt * st);
        return;
    }
    SphDerivs d;
    d.Pr = Pr;
    d.Pt = lam * St;
    d.Pfos = lam * (pole ? pole_limit(Sfp, ct) : Sfn * ist);
    if (ORD == ORD_HESS) {
        d.Prr = fma(lam * idl * idl, Srr, fma(2.0 * mu * idl, Sr, fma(nu, Sv, d2C0)));
        d.Ptt = lam * Stt;
        d.Prt = fma(lam * idl, Srt, mu * St);
        d.Pf = lam * Sf; d.Pff = lam * Sff; d.Ptf = lam * Stf;
        d.Prf = fma(lam * idl, Srf, mu * Sf);
    }
    if (ORD == ORD_HESS && pole) {
        // ---- the Hessian on the z axis (sin(theta) < 1e-10) ------------------
        // The spherical chain rule is singular there (and its pole branch
        // drops d2Y/dtheta2 and the |m| = 1, 2 limits, as the reference's
        // does).  On the axis only |m| <= 2 have second derivatives: with
        // Qh = P~_lm / sin^m(theta), regular on the axis, a term is
        // F(r) Qh(z/r) h_m(x, y) / r^m with h_0 = 1, h_1 = x or y, h_2 =
        // x^2 - y^2 or 2xy, and in t = ln r, s = sign(cos theta), V its value:
        //   m = 0: H_zz = (V_tt - V_t) / r^2, H_xx = H_yy = (V_t - l(l+1) V / 2) / r^2
        //   m = 1: H_xz (cos), H_yz (sin) = s (V_t - V) / r^2
        //   m = 2: H_xx = -H_yy = 2 V / r^2 (cos), H_xy = 2 V / r^2 (sin).
        // Qh is the program's recurrence run from its seed without the
        // sin^m factor, at cos(theta) = s.  The monopole and the log-scaling
        // factors are formed again here (rarely taken; nothing is held in
        // registers across the term loop for it); the chain rule below then
        // forms only the gradient for these points.
        double a0[6], c0, dc0, d2c0;
        load6(0, a0);
        horner(a0, c0, dc0, d2c0);
        dc0 *= idl;
        d2c0 *= idl * idl;
        double lm = 1.0, mu = 0.0, nu = 0.0;
        if (P.log_scaling) {
            const double ex = agx_exp(c0), p0 = rcp_nr(P.invPhi0 - ex), dpdx = p0 * p0 * ex;
            d2c0 = dpdx * fma(dc0 * dc0, p0 * (P.invPhi0 + ex), d2c0);
            dc0 = dpdx * dc0;
            c0 = p0;
            lm = c0; mu = dc0; nu = d2c0;
        }
        const double sg = ct < 0.0 ? -1.0 : 1.0;
        double v0 = 0.0, r0 = 0.0, rr0 = 0.0, l0 = 0.0;   // m = 0: sums of C, C', C'', l(l+1) C
        double v1c = 0.0, r1c = 0.0, v1s = 0.0, r1s = 0.0, v2c = 0.0, v2s = 0.0;
        const int S0p = P.n_groups + 1;
        for (int g = 0; g < P.n_groups; ++g) {
            const ivec8 gr = cld_rec(P.prog, g);
            const int m = gr.s0;
            if (m > 2) break;                       // groups come in increasing |m|
            const double cc = (gr.s3 == 2) ? 1.0 : sg;   // cos^2 or cos theta
            double Qm1 = 0.0, Qm2 = 0.0;
            for (int j = gr.s1; j < gr.s2; ++j) {
                const ivec8 rec = cld_rec(P.prog, S0p + j);
                const double Q = (j == gr.s1)
                    ? rec_lo(gr) : fma(rec_lo(rec) * (cc + rec_hi(rec)), Qm1, rec_mid(rec) * Qm2);
                Qm2 = Qm1;
                Qm1 = Q;
                const int oc = rec_off_lo(rec), os = rec_off_hi(rec);
                double a[6], c, dc, d2c;
                if (oc >= 0) {
                    load6(oc, a);
                    horner(a, c, dc, d2c);
                    if (m == 0) {
                        v0 = fma(c, Q, v0); r0 = fma(dc, Q, r0); rr0 = fma(d2c, Q, rr0);
                        l0 = fma((double)(rec.s0 * (rec.s0 + 1)) * c, Q, l0);
                    } else if (m == 1) {
                        v1c = fma(c, Q, v1c); r1c = fma(dc, Q, r1c);
                    } else {
                        v2c = fma(c, Q, v2c);
                    }
                }
                if (SIN && os >= 0) {
                    load6(os, a);
                    horner(a, c, dc, d2c);
                    if (m == 1) { v1s = fma(c, Q, v1s); r1s = fma(dc, Q, r1s); }
                    else if (m == 2) { v2s = fma(c, Q, v2s); }
                }
            }
        }
        const double ir2 = ir * ir, li = lm * idl;
        const double V0t = fma(li, r0, fma(mu, v0, dc0));
        const double V0tt = fma(li * idl, rr0, fma(2.0 * mu * idl, r0, fma(nu, v0, d2c0)));
        const double hxx0 = fma(-0.5 * lm, l0, V0t) * ir2;
        const double q2c = 2.0 * lm * v2c * ir2;
        o.h[0] = hxx0 + q2c;
        o.h[1] = hxx0 - q2c;
        o.h[2] = (V0tt - V0t) * ir2;
        o.h[3] = 2.0 * lm * v2s * ir2;
        o.h[4] = sg * (fma(li, r1s, mu * v1s) - lm * v1s) * ir2;
        o.h[5] = sg * (fma(li, r1c, mu * v1c) - lm * v1c) * ir2;
    }
    sph_to_cart<ORD == ORD_HESS>(x, y, z, ir, R, ct, st, cph, sph, d, o.g, o.h,
                                 ORD == ORD_HESS && pole);
}

// Classification of one point: 0 origin, 1 inner law, 2 interior (k), 3 outer law.
// Every comparison is written so that a NaN position fails it and falls through
// to the outer law, whose value is then NaN (as in the reference, where NaN
// reaches the spline and propagates); r^2 = +inf (|x| above about 1.3e154) is
// the outer law too, and extrap() returns its limit there (Phi, force and
// Hessian 0; the reference's outer law at r = inf gives the same).
struct Where {
    int kind, k;
    double r2, r, ir, lnr;
};

__device__ __forceinline__ bool beyond_fp64(const Where &w) { return w.r2 == __builtin_inf(); }

__device__ __forceinline__ Where classify(const MPK &P, double x, double y, double z)
{
    Where w;
    w.r2 = fma(x, x, fma(y, y, z * z));
    if (w.r2 >= 0x1p-767) {   // (+inf gives r = NaN; beyond_fp64 catches it)
        sqrt_rsqrt(w.r2, w.r, w.ir);
    } else {                  // r < 4e-116 (the origin or the inner law), or NaN
        w.r = sqrt(w.r2);
        w.ir = 1.0 / w.r;
    }
    w.k = 0;
    w.lnr = 0.0;
    if (w.r < 1.0e-300) { w.kind = 0; return w; }
    w.lnr = agx_log(w.r);
    if (w.lnr < P.logr_min) { w.kind = 1; return w; }
    const double kd = (w.lnr - P.logr_min) * P.inv_dlogr;    // >= 0 here, or NaN
    if (!(kd < (double)P.n_int)) { w.kind = 3; return w; }
    w.kind = 2;
    w.k = (int)kd;
    return w;
}

template <int ORD>
__device__ __forceinline__ void extrap(const MPK &P, const Where &w, double x, double y, double z,
                                       PointOut &o, double inv4piG)
{
    double Pv, D, D2;
    if (w.kind == 0) {      // origin: Phi = W, zero force and Hessian (reference :166-171)
        o.phi = (ORD == ORD_DENS) ? 0.0 : P.inner_W;
#pragma unroll
        for (int a = 0; a < 3; ++a) o.g[a] = 0.0;
#pragma unroll
        for (int a = 0; a < 6; ++a) o.h[a] = 0.0;
        return;
    }
    if (w.kind == 1) {
        const double er = w.r * P.inv_rmin;
        law(0.0, P.inner_s, P.inner_U, P.inner_W, P.inner_log, w.lnr - P.logr_min, er,
            rcp_nr(er), Pv, D, D2);
    } else {
        if (beyond_fp64(w)) {   // r = inf: every outer law decays to 0 with its derivatives
            o.phi = 0.0;
#pragma unroll
            for (int a = 0; a < 3; ++a) o.g[a] = 0.0;
#pragma unroll
            for (int a = 0; a < 6; ++a) o.h[a] = 0.0;
            return;
        }
        if (ORD == ORD_DENS) { o.phi = 0.0 * w.r2; return; }   // 0 beyond r_max; NaN stays NaN
        law(-1.0, P.outer_s, P.outer_U, P.outer_W, P.outer_log,
            w.lnr - fma((double)P.n_int, P.dlogr, P.logr_min), w.r * P.inv_rmax, P.rmax * w.ir,
            Pv, D, D2);
    }
    if (ORD == ORD_DENS) { o.phi = inv4piG * (D2 + D) * (w.ir * w.ir); return; }
    radial_out<ORD>(x, y, z, w.ir, Pv, D, D2, o);
}

// Outputs are written once and not read again by the kernel.  NT: streaming
// (non-temporal) stores, for kernels whose wavefronts write runs of
// consecutive particles; they leave fewer dirty L2 lines to write back when
// the kernel ends (the halo table at N=10M: 4.0e10 -> 4.3e10 particles/s
// over back-to-back launches).  Kernels that write particles in a scattered
// order (the interval-sorted one) use ordinary stores, so that the partial
// lines of neighbouring particles merge in L2 before they go to memory.
template <int ORD, bool NT = true>
__device__ __forceinline__ void store(const PointOut &o, int64_t i, double *phi, double *force,
                                      double *deriv)
{
    auto put = [](double v, double *p) {
        if (NT) __builtin_nontemporal_store(v, p);
        else *p = v;
    };
    if (phi) put(o.phi, phi + i);
    if ((ORD == ORD_GRAD || ORD == ORD_HESS) && force) {
        put(-o.g[0], force + 3 * i);
        put(-o.g[1], force + 3 * i + 1);
        put(-o.g[2], force + 3 * i + 2);
    }
    if (ORD == ORD_HESS && deriv) {
#pragma unroll
        for (int p = 0; p < 6; ++p) put(-o.h[p], deriv + 6 * i + p);
    }
}

// ---------------------------------------------------------------------------
// Radially binned traversal.  A workgroup takes a tile of particles, sorts the
// tile's indices by radial interval in LDS (counting sort, each interval
// padded to whole wavefronts), and then every wavefront handles 64 slots that
// share one interval k.  With k wave-uniform, the interval's coefficient slab
// is read with scalar loads into SGPRs and feeds the FP64 FMAs directly (no
// per-lane gathers of the table), and per-interval sums can be reduced across
// the wavefront.  Bin 0 collects the particles outside the grid (origin, inner
// and outer laws), which need no table.  body(bin, valid, particle) is called
// by every lane of every chunk (padding lanes with valid = false), so it may
// use wavefront-wide operations.
// ---------------------------------------------------------------------------
constexpr int kBinThreads = 256;
constexpr int kMaxBins = 128;          // n_int + 1 bins at most
constexpr int kMaxTile = 8192;         // particles per tile

// LDS of one binned tile: TILE particles at most, BINS bins (n_int + 1) at
// most.  BinLDS (8192, 128) is about 42 KB.
template <int TILE, int BINS>
struct BinLDST {
    static constexpr int kTile = TILE, kBins = BINS;
    static constexpr int kSlots = TILE + BINS * 63, kChunks = kSlots / 64;
    uint16_t slot_idx[kSlots];
    uint8_t pbin[TILE];
    int cnt[BINS], off[BINS + 1], cur[BINS];
    uint8_t chunk_bin[kChunks];
    int n_chunks;
};
using BinLDS = BinLDST<kMaxTile, kMaxBins>;

// One tile: particles [base, base + cnt_here), cnt_here <= kMaxTile.  Called by
// every thread of the workgroup; ends with a barrier, so L may be reused.
template <class LDS, class Body>
__device__ __forceinline__ void binned_tile(const MPK &P, const double *__restrict__ xyz,
                                            int64_t base, int cnt_here, LDS &L, Body body)
{
    const int nb = P.n_int + 1;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6, n_waves = kBinThreads / 64;
    for (int b = tid; b < nb; b += kBinThreads) { L.cnt[b] = 0; L.cur[b] = 0; }
    __syncthreads();
    // phase 1: classify and histogram
    for (int j = tid; j < cnt_here; j += kBinThreads) {
        const int64_t i = base + j;
        double x, y, z;
        load_pos(P, xyz, i, x, y, z);
        Where w = classify(P, x, y, z);
        const int b = (w.kind == 2) ? w.k + 1 : 0;
        L.pbin[j] = (uint8_t)b;
        atomicAdd(&L.cnt[b], 1);
    }
    __syncthreads();
    // phase 2: wave-padded offsets, chunk -> bin map
    if (tid == 0) {
        int o = 0;
        for (int b = 0; b < nb; ++b) { L.off[b] = o; o += (L.cnt[b] + 63) & ~63; }
        L.off[nb] = o;
        L.n_chunks = o >> 6;
    }
    __syncthreads();
    for (int b = tid; b < nb; b += kBinThreads)
        for (int c = L.off[b] >> 6; c < (L.off[b + 1] >> 6); ++c) L.chunk_bin[c] = (uint8_t)b;
    for (int q = tid; q < L.off[nb]; q += kBinThreads) L.slot_idx[q] = 0xFFFF;
    __syncthreads();
    // phase 3: scatter tile-local indices into their bin's slots
    for (int j = tid; j < cnt_here; j += kBinThreads) {
        const int b = L.pbin[j];
        const int pos = L.off[b] + atomicAdd(&L.cur[b], 1);
        L.slot_idx[pos] = (uint16_t)j;
    }
    __syncthreads();
    // phase 4: one interval per wavefront chunk
    const int nch = L.n_chunks;
    for (int c = wave; c < nch; c += n_waves) {
        const int b = __builtin_amdgcn_readfirstlane((int)L.chunk_bin[c]);
        const int j = L.slot_idx[c * 64 + lane];
        body(b, j != 0xFFFF, base + (j == 0xFFFF ? 0 : j));
    }
    __syncthreads();
}

// Tiles of `tile` particles, grid-strided over the workgroups.
template <class Body>
__device__ __forceinline__ void binned_tiles(const MPK &P, const double *__restrict__ xyz,
                                             int64_t n, int64_t tile, Body body)
{
    __shared__ BinLDS L;
    const int64_t n_tiles = (n + tile - 1) / tile;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const int64_t base = t * tile;
        const int cnt_here = (int)((n - base < tile) ? (n - base) : tile);
        binned_tile(P, xyz, base, cnt_here, L, body);
    }
}

// Sort key of the interval-sorted traversals (sort_dev.h): 0 outside the grid
// (origin, inner and outer laws), k + 1 in interval k.
struct MpBinKey {
    MPK P;                           // as built on the host (make_args)
    const double *xyz;
    __device__ int operator()(int64_t i) const
    {
        const MPK Q = with_law(P);   // the origin (uniform loads, hoisted out of the loop)
        double x, y, z;
        load_pos(Q, xyz, i, x, y, z);
        const Where w = classify(Q, x, y, z);
        return (w.kind == 2) ? w.k + 1 : 0;
    }
};

constexpr uint32_t kProgFlags = AGX_PROG_STRIDE1 | AGX_PROG_COS_ONLY | AGX_PROG_STRIDE2 |
                                AGX_PROG_TRIAXIAL_MASK | AGX_PROG_FULL_MASK;

// The stride a kernel variant may assume from the program flags: 1, 2, or 0
// (any).  A program with no recurring group satisfies both flags.
inline int prog_stride(uint32_t flags)
{
    return (flags & AGX_PROG_STRIDE1) ? 1 : (flags & AGX_PROG_STRIDE2) ? 2 : 0;
}

inline MPK make_args(const agx_multipole *mp)
{
    MPK P;
    P.poly = mp->poly_km;
    P.prog = mp->program;
    P.n_int = mp->n_int; P.n_lm = mp->n_lm; P.n_groups = mp->n_groups; P.n_steps = mp->n_steps;
    P.log_scaling = mp->log_scaling; P.inner_log = mp->inner_log; P.outer_log = mp->outer_log;
    P.logr_min = mp->logr_min; P.dlogr = mp->dlogr; P.inv_dlogr = mp->inv_dlogr;
    P.invPhi0 = mp->invPhi0;
    P.inner_s = mp->inner_s; P.inner_U = mp->inner_U; P.inner_W = mp->inner_W;
    P.outer_s = mp->outer_s; P.outer_U = mp->outer_U; P.outer_W = mp->outer_W;
    P.inv_rmin = std::exp(-mp->logr_min);
    P.inv_rmax = std::exp(-(mp->logr_min + (double)mp->n_int * mp->dlogr));
    P.rmax = std::exp(mp->logr_min + (double)mp->n_int * mp->dlogr);
    P.law = mp->law;
    P.origin = mp->origin;
    P.ox = P.oy = P.oz = 0.0;
    return P;
}

}  // namespace agx
