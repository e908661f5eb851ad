// Adjoint (reverse-mode) kernel of the Multipole potential+force evaluation.
//
// Given upstream gradients w_phi = dL/dPhi_i and w_F = dL/dF_i (F = -grad Phi)
// at N points, one launch produces
//   dL/dx_i   = w_phi grad Phi - H w_F            (Hessian-vector product)
//   dL/dpoly  [n_int][n_lm][6]  (Horner coefficients of every term, interval)
//   dL/dscal  [invPhi0, inner U, inner W, outer U, outer W]
// The table gradient is a scatter into the interval a particle falls in: the
// particles are sorted by interval per tile (binned_tiles), so a wavefront's
// 64 lanes share the interval; each coefficient's contribution is summed
// across the wavefront first and added to memory with one FP64 atomic per
// wavefront.  No reference counterpart: the reference has no derivatives with
// respect to its tables.
#include "multipole_dev.h"

namespace agx {

// s^j a + j s^(j-1) b for j = 0..5, summed over the wavefront and added to g[0..5]
__device__ __forceinline__ void scatter6(double *g, double s, double a, double b, int lane)
{
    const double s2 = s * s, s3 = s2 * s, s4 = s2 * s2;
    wave_sum6_add(g, a, fma(s, a, b), fma(s2, a, 2.0 * s * b), fma(s3, a, 3.0 * s2 * b),
                  fma(s4, a, 4.0 * s3 * b), fma(s4 * s, a, 5.0 * s4 * b), lane);
}

__device__ __forceinline__ void adjoint_interior(const MPK &P, int k, bool valid, double x, double y,
                                                 double z, double wphi, const double wf[3],
                                                 double *gx, double *gpoly, double *gscal,
                                                 int lane)
{
    const double *slab = P.poly + (int64_t)k * P.n_lm * 6;
    double *gslab = gpoly + (int64_t)k * P.n_lm * 6;
    if (!valid) {   // a finite point inside interval k; its seeds are zero
        x = exp(fma((double)k + 0.5, P.dlogr, P.logr_min)); y = 0.0; z = 0.0;
        wphi = 0.0;
    }
    const double wfx = valid ? wf[0] : 0.0, wfy = valid ? wf[1] : 0.0, wfz = valid ? wf[2] : 0.0;
    const double r2 = fma(x, x, fma(y, y, z * z)), r = sqrt(r2), lnr = agx_log(r);
    // Hessian-vector product for dL/dx
    {
        PointOut o{};
        interior<ORD_HESS, true>(P, slab, x, y, z, r2, r, lnr, k, o, 0.0);
        if (valid) {
            gx[0] = wphi * o.g[0] - (o.h[0] * wfx + o.h[3] * wfy + o.h[5] * wfz);
            gx[1] = wphi * o.g[1] - (o.h[3] * wfx + o.h[1] * wfy + o.h[4] * wfz);
            gx[2] = wphi * o.g[2] - (o.h[5] * wfx + o.h[4] * wfy + o.h[2] * wfz);
        }
    }
    // geometry and seeds: g = Pr u_r + Pt u_t + Pfos u_f
    const double idl = P.inv_dlogr;
    double s = (lnr - fma((double)k, P.dlogr, P.logr_min)) * idl;
    s = fmin(fmax(s, 0.0), 1.0);
    const double R = sqrt(fma(x, x, y * y)), ir = 1.0 / r, ir2 = ir * ir;
    const double ct = z * ir, st = R * ir;
    const bool pole = st < 1.0e-10;
    double cph = 1.0, sph = 0.0;
    if (R > 1.0e-300) { cph = x / R; sph = y / R; }
    const double ist = pole ? 0.0 : 1.0 / st;
    const double wPr = -(wfx * x + wfy * y + wfz * z) * ir2;
    const double wPt = -(wfx * ct * cph + wfy * ct * sph - wfz * st) * ir;
    const double wPf = -(-wfx * sph + wfy * cph) * ir;

    // monopole radial function and the log-scaling factors
    const double s2 = s * s, s3 = s2 * s, s4 = s2 * s2, s5 = s4 * s;
    typedef __attribute__((address_space(4))) const double cd_t;
    cd_t *a0 = (cd_t *)slab;
    const double c0 = fma(a0[5], s5, fma(a0[4], s4, fma(a0[3], s3, fma(a0[2], s2, fma(a0[1], s, a0[0])))));
    const double c0d = fma(5.0 * a0[5], s4, fma(4.0 * a0[4], s3, fma(3.0 * a0[3], s2, fma(2.0 * a0[2], s, a0[1]))));
    double lam = 1.0, mu = 0.0, P0 = 0.0, ex = 0.0, dPdX = 0.0;
    if (P.log_scaling) {
        ex = exp(c0);
        P0 = 1.0 / (P.invPhi0 - ex);
        dPdX = P0 * P0 * ex;
        lam = P0;
        mu = dPdX * c0d * idl;
    }

    // walk the active terms as the forward kernel does
    double Sv = 0.0, Sr = 0.0, St = 0.0, Sx = 0.0;
    double cm = 1.0, sm = 0.0, sp = 1.0, sp1 = 0.0;
    int cur_m = 0;
    const int S0 = P.n_groups + 1;
    for (int g = 0; g < P.n_groups; ++g) {
        const ivec8 gr = cld_rec(P.prog, g);
        const int m = gr.s0;
        while (cur_m < m) {
            sp1 = sp; sp *= st;
            const double t = fma(cm, cph, -sm * sph);
            sm = fma(sm, cph, cm * sph); cm = t;
            ++cur_m;
        }
        const double q = rec_lo(gr);
        double Q1 = q * sp, dQ1 = (m > 0) ? (double)m * q * sp1 * ct : 0.0, Q2 = 0.0, dQ2 = 0.0;
        const int j0 = gr.s1, j1 = gr.s2;
        for (int j = j0; j < j1; ++j) {
            const ivec8 st_rec = cld_rec(P.prog, S0 + j);
            const int l = st_rec.s0;
            if (l > m) {
                const double al = rec_lo(st_rec), be = rec_hi(st_rec);
                const double Qn = fma(al * ct, Q1, -be * Q2);
                const double dQn = fma(al, fma(ct, dQ1, -st * Q1), -be * dQ2);
                Q2 = Q1; Q1 = Qn; dQ2 = dQ1; dQ1 = dQn;
            }
            for (int side = 0; side < 2; ++side) {
                const int off = side == 0 ? st_rec.s1 : st_rec.s2;
                if (off < 0) continue;
                const double T = side == 0 ? cm : sm;
                const double dT = side == 0 ? -(double)m * sm : (double)m * cm;
                const double dTp = side == 0 ? -sm : cm;        // pole limit, m = 1
                const double X = pole ? ((m == 1) ? dQ1 * dTp : 0.0) : Q1 * dT * ist;
                cd_t *a = (cd_t *)(slab + off);
                const double C = fma(a[5], s5, fma(a[4], s4, fma(a[3], s3, fma(a[2], s2, fma(a[1], s, a[0])))));
                const double Cd = fma(5.0 * a[5], s4, fma(4.0 * a[4], s3, fma(3.0 * a[3], s2, fma(2.0 * a[2], s, a[1]))));
                const double QT = Q1 * T;
                Sv = fma(C, QT, Sv); Sr = fma(Cd, QT, Sr); St = fma(C, dQ1 * T, St); Sx = fma(C, X, Sx);
                const double alpha = lam * (wphi * QT + wPt * dQ1 * T + wPf * X) + wPr * mu * QT;
                const double beta = wPr * lam * idl * QT;
                scatter6(gslab + off, s, alpha, beta, lane);
            }
        }
    }
    // monopole coefficients and invPhi0
    double gc0, gc0d, ginv = 0.0;
    if (P.log_scaling) {
        const double gC0 = wphi * (1.0 + Sv) + wPr * idl * Sr + wPt * St + wPf * Sx;
        const double gdC0 = wPr * (1.0 + Sv);
        gc0 = gC0 * dPdX + gdC0 * dPdX * (1.0 + 2.0 * P0 * ex) * c0d * idl;
        gc0d = gdC0 * dPdX * idl;
        ginv = -gC0 * P0 * P0 - gdC0 * 2.0 * P0 * P0 * P0 * ex * c0d * idl;
    } else {
        gc0 = wphi;
        gc0d = wPr * idl;
    }
    scatter6(gslab, s, gc0, gc0d, lane);
    const double gi = wave_sum(ginv);
    if (lane == 0 && P.log_scaling) atomicAdd(gscal + 0, gi);
}

__device__ __forceinline__ void adjoint_extrap(const MPK &P, bool valid, double x, double y,
                                               double z, double wphi, const double wf[3],
                                               double *gx, double *gscal, int lane)
{
    double gU_in = 0.0, gW_in = 0.0, gU_out = 0.0, gW_out = 0.0;
    if (valid) {
        Where w = classify(P, x, y, z);
        if (w.kind == 0) {
            gW_in = wphi;
            gx[0] = gx[1] = gx[2] = 0.0;
        } else {
            const bool in = w.kind == 1;
            const double v = in ? 0.0 : -1.0;
            const double sx = in ? P.inner_s : P.outer_s, U = in ? P.inner_U : P.outer_U;
            const double W = in ? P.inner_W : P.outer_W;
            const int logm = in ? P.inner_log : P.outer_log;
            const double t = in ? w.lnr - P.logr_min : w.lnr - fma((double)P.n_int, P.dlogr, P.logr_min);
            PointOut o{};
            extrap<ORD_HESS>(P, w, x, y, z, o, 0.0);
            gx[0] = wphi * o.g[0] - (o.h[0] * wf[0] + o.h[3] * wf[1] + o.h[5] * wf[2]);
            gx[1] = wphi * o.g[1] - (o.h[3] * wf[0] + o.h[1] * wf[1] + o.h[4] * wf[2]);
            gx[2] = wphi * o.g[2] - (o.h[5] * wf[0] + o.h[4] * wf[1] + o.h[2] * wf[2]);
            const double wPr = -(wf[0] * x + wf[1] * y + wf[2] * z) / w.r2;
            const double ev = (v == 0.0) ? 1.0 : exp(v * t);
            double pU, dU;                       // d(Phi, dPhi/dln r)/dU
            if (logm) { pU = t * ev; dU = fma(v, t, 1.0) * ev; }
            else { const double es = (sx != 0.0) ? exp(sx * t) : 1.0; pU = es; dU = sx * es; }
            const double gU = wphi * pU + wPr * dU, gW = wphi * ev + wPr * v * ev;
            if (in) { gU_in = gU; gW_in = gW; } else { gU_out = gU; gW_out = gW; }
            (void)U; (void)W;
        }
    }
    gU_in = wave_sum(gU_in); gW_in = wave_sum(gW_in);
    gU_out = wave_sum(gU_out); gW_out = wave_sum(gW_out);
    if (lane == 0) {
        atomicAdd(gscal + 1, gU_in); atomicAdd(gscal + 2, gW_in);
        atomicAdd(gscal + 3, gU_out); atomicAdd(gscal + 4, gW_out);
    }
}

__global__ __launch_bounds__(kBinThreads) void mp_adjoint_kernel(
    MPK P, const double *__restrict__ xyz, int64_t n, int64_t tile, const double *__restrict__ gphi,
    const double *__restrict__ gforce, double *__restrict__ gxyz, double *__restrict__ gpoly,
    double *__restrict__ gscal)
{
    const int lane = threadIdx.x & 63;
    binned_tiles(P, xyz, n, tile, [&](int b, bool valid, int64_t i) {
        double x = 1.0, y = 0.0, z = 0.0, wphi = 0.0, wf[3] = {0.0, 0.0, 0.0};
        if (valid) {
            x = xyz[3 * i]; y = xyz[3 * i + 1]; z = xyz[3 * i + 2];
            if (gphi) wphi = gphi[i];
            if (gforce) { wf[0] = gforce[3 * i]; wf[1] = gforce[3 * i + 1]; wf[2] = gforce[3 * i + 2]; }
        }
        double gx[3] = {0.0, 0.0, 0.0};
        if (b > 0)
            adjoint_interior(P, b - 1, valid, x, y, z, wphi, wf, gx, gpoly, gscal, lane);
        else
            adjoint_extrap(P, valid, x, y, z, wphi, wf, gx, gscal, lane);
        if (valid && gxyz) { gxyz[3 * i] = gx[0]; gxyz[3 * i + 1] = gx[1]; gxyz[3 * i + 2] = gx[2]; }
    });
}

}  // namespace agx

using namespace agx;

extern "C" int agx_multipole_adjoint(const agx_multipole *mp, const double *d_xyz, int64_t n,
                                     const double *d_gphi, const double *d_gforce, double *d_gxyz,
                                     double *d_gpoly, double *d_gscal, void *stream)
{
    if (!mp) return AGX_E_NULL;
    if (n < 0) return AGX_E_SIZE;
    if (!mp->poly_km || !mp->program || !d_gpoly || !d_gscal) return AGX_E_NULL;
    if (mp->n_int + 1 > kMaxBins) return AGX_E_SIZE;
    if (mp->lmax > 32) return AGX_E_LMAX;
    if (n == 0) return AGX_OK;
    if (!d_xyz || (!d_gphi && !d_gforce)) return AGX_E_NULL;
    MPK P = make_args(mp);
    const int64_t max_wg = 256 * 3;
    int64_t tile = (n + max_wg - 1) / max_wg;
    if (tile < 1024) tile = 1024;
    if (tile > kMaxTile) tile = kMaxTile;
    int64_t n_tiles = (n + tile - 1) / tile;
    const unsigned grid = (unsigned)(n_tiles < max_wg ? n_tiles : max_wg);
    hipLaunchKernelGGL(mp_adjoint_kernel, dim3(grid), dim3(kBinThreads), 0, (hipStream_t)stream, P,
                       d_xyz, n, tile, d_gphi, d_gforce, d_gxyz, d_gpoly, d_gscal);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}
