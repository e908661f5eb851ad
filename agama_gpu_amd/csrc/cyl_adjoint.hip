// Adjoint (reverse-mode) kernel of the CylSpline potential+force evaluation.
//
// Given upstream gradients w_phi = dL/dPhi_i and w_F = dL/dF_i (F = -grad Phi)
// at N points, one call produces
//   dL/dx_i      = w_phi grad Phi - H w_F            (Hessian-vector product)
//   dL/dnodes    [n_harm][nR][nz][4]  (f, f_lR, f_lz, f_lRlz of every node)
//   dL/dW        [(lmax_outer+1)^2]   (the outer PowerLaw coefficients)
// Inside the grid the value of harmonic m at a point is linear in the 16 node
// quantities of its cell, with the bicubic Hermite weights A_a(tR) B_b(tz)
// (value nodes), A_a B'_b (lz slopes), A'_a B_b (lR slopes), A'_a B'_b
// (cross slopes).  The particles are first sorted by grid cell (a counting
// sort into the caller's workspace), so the 64 lanes of a wavefront share a
// cell: the 16 gradients of each harmonic are summed across the wavefront
// (lane swaps) and added to memory by 16 lanes, one FP64 atomic each.
// Outside the grid Phi is linear in W; each W entry's contribution is summed
// over the wavefront before one atomic.  Grids of more than 8191 cells fall
// back to one atomic per particle and node quantity.  No reference
// counterpart (the reference has no derivatives with respect to positions or
// tables).
#include "cylspline_dev.h"
#include "sort_dev.h"

namespace agx {

// Hermite basis on [0,1] for the value at the lower / upper node and the
// slope (times the step h) at the lower / upper node, and their t-derivatives
__device__ __forceinline__ void herm_basis(double t, double h, double w[4], double dw[4])
{
    const double T = 1.0 - t;
    w[0] = T * T * (1.0 + 2.0 * t);
    w[1] = t * t * (1.0 + 2.0 * T);
    w[2] = h * t * T * T;
    w[3] = -h * t * t * T;
    dw[0] = -6.0 * t * T;
    dw[1] = 6.0 * t * T;
    dw[2] = h * (1.0 - 4.0 * t + 3.0 * t * t);
    dw[3] = h * (3.0 * t * t - 2.0 * t);
}

__device__ __forceinline__ void cyl_adjoint_outer(const CSK &P, bool valid, double x, double y,
                                                  double z, double R, double wphi,
                                                  const double wf[3], double *gW)
{
    const double r2 = fma(R, R, z * z), r = sqrt(r2), ir = 1.0 / r;
    const double ct = z * ir, st = R * ir;
    const bool pole = st < 1.0e-10;
    double cph = 1.0, sph = 0.0;
    if (R > 1.0e-300) { cph = x / R; sph = y / R; }
    const double ist = pole ? 0.0 : 1.0 / st;
    const double psg = ct < 0.0 ? -1.0 : 1.0;   // pole limit of Y / sin(theta): +-dY/dtheta
    const double q = P.r0 * ir;
    // seeds: dL/d(d Phi / d ln r), dL/d(d Phi / d theta), dL/d((d Phi / d phi) / sin theta)
    const double wPr = valid ? -(wf[0] * x + wf[1] * y + wf[2] * z) * ir * ir : 0.0;
    const double wPt = valid ? -(wf[0] * ct * cph + wf[1] * ct * sph - wf[2] * st) * ir : 0.0;
    const double wPf = valid ? -(-wf[0] * sph + wf[1] * cph) * ir : 0.0;
    const double wv = valid ? wphi : 0.0;
    const int lane = threadIdx.x & 63;
    double cm = 1.0, sm = 0.0, pmm = 1.0 / sqrt(4.0 * M_PI), dpmm = 0.0;
    for (int m = 0; m <= P.mmax_o; ++m) {
        if (m > 0) {
            const double t = fma(cm, cph, -sm * sph);
            sm = fma(sm, cph, cm * sph); cm = t;
            const double f = kOuterRec.f[m];
            dpmm = f * fma(ct, pmm, st * dpmm);
            pmm = f * st * pmm;
        }
        const double mul = (m == 0) ? 2.0 * sqrt(M_PI) : 2.0 * sqrt(2.0 * M_PI);
        double Q1 = pmm, dQ1 = dpmm, Q2 = 0.0, dQ2 = 0.0;
        double rl = q;
        for (int l = 0; l < m; ++l) rl *= q;
        for (int l = m; l <= P.lmax_o; ++l) {
            if (l > m) {
                const double a = kOuterRec.a[l][m], b = kOuterRec.b[l][m];
                const double Qn = fma(a * ct, Q1, -b * Q2);
                const double dQn = fma(a, fma(ct, dQ1, -st * Q1), -b * dQ2);
                Q2 = Q1; Q1 = Qn; dQ2 = dQ1; dQ1 = dQn;
                rl *= q;
            }
            const double Y = mul * Q1, dY = mul * dQ1;
            const double k1 = -(double)(l + 1);
            // cos term W[l(l+1)+m]
            {
                const double fphi = pole ? ((m == 1) ? -psg * sm * dY : 0.0) : -(double)m * sm * Y * ist;
                const double g = rl * (cm * Y * fma(wPr, k1, wv) + wPt * cm * dY + wPf * fphi);
                const double gs = wave_sum(g);
                if (lane == 0) atomicAdd(gW + l * (l + 1) + m, gs);
            }
            if (m > 0) {   // sin term W[l(l+1)-m]
                const double fphi = pole ? ((m == 1) ? psg * cm * dY : 0.0) : (double)m * cm * Y * ist;
                const double g = rl * (sm * Y * fma(wPr, k1, wv) + wPt * sm * dY + wPf * fphi);
                const double gs = wave_sum(g);
                if (lane == 0) atomicAdd(gW + l * (l + 1) - m, gs);
            }
        }
    }
}

// One point inside the grid.  SORTED: every lane of the wavefront is in cell
// (iR, iz) (wavefront-uniform) and the node gradients are summed across the
// wavefront before the atomics; otherwise the cell is found per lane and each
// lane adds its own 16 gradients per harmonic.
template <bool SORTED>
__device__ __forceinline__ void cyl_adjoint_inner(const CSK &P, double x, double y, double z,
                                                  double R, double lR, double lz, double wphi,
                                                  const double wf[3], double *gnodes, int iR,
                                                  int iz)
{
    if (!SORTED) { iR = cell_search(lR, P.lR, P.nR); iz = cell_search(lz, P.lz, P.nz); }
    const double lR0 = cld(P.lR, iR), lR1 = cld(P.lR, iR + 1);
    const double lz0 = cld(P.lz, iz), lz1 = cld(P.lz, iz + 1);
    const double hR = lR1 - lR0, hz = lz1 - lz0;
    const double ihR = rcp_nr(hR), ihz = rcp_nr(hz);
    const double tR = (lR - lR0) * ihR, tz = (lz - lz0) * ihz;
    double A[4], dA[4], B[4], dB[4];
    herm_basis(tR, hR, A, dA);
    herm_basis(tz, hz, B, dB);
    for (int k = 0; k < 4; ++k) { dA[k] *= ihR; dB[k] *= ihz; }   // d/dlR, d/dlz
    double cph = 1.0, sph = 0.0;
    if (R > 1.0e-300) { cph = x / R; sph = y / R; }
    const double Rs2 = P.Rscale * P.Rscale;
    const double jR = 1.0 / sqrt(fma(R, R, Rs2)), jz = 1.0 / sqrt(fma(z, z, Rs2));
    // seeds: dL/dPR, dL/dPZ (PR = dPhi/dlR, PZ = dPhi/dlz), dL/dPp (Pp = dPhi/dphi)
    // on the axis the forward force has no (x, y) part
    const double aR = (R > 1.0e-30) ? -jR * fma(wf[0], cph, wf[1] * sph) : 0.0;
    const double aZ = -jz * wf[2];
    const double ap = (R > 1.0e-30) ? -fma(-wf[0], sph, wf[1] * cph) / R : 0.0;
    const int64_t plane = (int64_t)P.nR * P.nz;

    // value and lR, lz derivatives of harmonic mm at the point
    auto harmonic = [&](int mm, double &v, double &vR, double &vZ) {
        const double *nd = P.nodes + ((int64_t)mm * plane + (int64_t)iR * P.nz + iz) * 4;
        v = 0.0; vR = 0.0; vZ = 0.0;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const double *c = nd + (a * P.nz + b) * 4;   // f, f_lR, f_lz, f_lRlz
                const double f = c[0], fx = c[1], fy = c[2], fxy = c[3];
                const double Aa = A[a], Ada = A[2 + a], Bb = B[b], Bdb = B[2 + b];
                const double dAa = dA[a], dAda = dA[2 + a], dBb = dB[b], dBdb = dB[2 + b];
                v += Aa * Bb * f + Ada * Bb * fx + Aa * Bdb * fy + Ada * Bdb * fxy;
                vR += dAa * Bb * f + dAda * Bb * fx + dAa * Bdb * fy + dAda * Bdb * fxy;
                vZ += Aa * dBb * f + Ada * dBb * fx + Aa * dBdb * fy + Ada * dBdb * fxy;
            }
    };
    auto scatter = [&](int mm, double gv, double gvR, double gvZ) {
        double *gd = gnodes + ((int64_t)mm * plane + (int64_t)iR * P.nz + iz) * 4;
        if constexpr (SORTED) {
            double v[16], d[4];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const double Aa = A[a], Ada = A[2 + a], Bb = B[b], Bdb = B[2 + b];
                    const double dAa = dA[a], dAda = dA[2 + a], dBb = dB[b], dBdb = dB[2 + b];
                    double *q = v + (2 * a + b) * 4;
                    q[0] = gv * Aa * Bb + gvR * dAa * Bb + gvZ * Aa * dBb;
                    q[1] = gv * Ada * Bb + gvR * dAda * Bb + gvZ * Ada * dBb;
                    q[2] = gv * Aa * Bdb + gvR * dAa * Bdb + gvZ * Aa * dBdb;
                    q[3] = gv * Ada * Bdb + gvR * dAda * Bdb + gvZ * Ada * dBdb;
                }
            wave_sum16(v, d);
            const int lane = threadIdx.x & 63;
            if ((lane & 15) == 0) {
                const int r = lane >> 4, q = ((r & 1) << 1) | (r >> 1);
#pragma unroll
                for (int c = 0; c < 4; ++c) atomicAdd(gd + ((c >> 1) * P.nz + (c & 1)) * 4 + q, d[c]);
            }
            return;
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                double *c = gd + (a * P.nz + b) * 4;
                const double Aa = A[a], Ada = A[2 + a], Bb = B[b], Bdb = B[2 + b];
                const double dAa = dA[a], dAda = dA[2 + a], dBb = dB[b], dBdb = dB[2 + b];
                atomicAdd(c + 0, gv * Aa * Bb + gvR * dAa * Bb + gvZ * Aa * dBb);
                atomicAdd(c + 1, gv * Ada * Bb + gvR * dAda * Bb + gvZ * Ada * dBb);
                atomicAdd(c + 2, gv * Aa * Bdb + gvR * dAa * Bdb + gvZ * Aa * dBdb);
                atomicAdd(c + 3, gv * Ada * Bdb + gvR * dAda * Bdb + gvZ * Ada * dBdb);
            }
    };
    auto trig = [&](int am, double &c, double &s) {
        c = 1.0; s = 0.0;
        for (int k = 0; k < am; ++k) {
            const double t = fma(c, cph, -s * sph);
            s = fma(s, cph, c * sph); c = t;
        }
    };

    // pass 1: the Fourier sums (as the forward kernel)
    double S = P.log_scaling ? 1.0 : 0.0, SR = 0.0, SZ = 0.0, Sp = 0.0;
    double F0 = 0.0, F0R = 0.0, F0Z = 0.0;
    for (int mm = 0; mm < P.n_harm; ++mm) {
        const int mv = cldi(P.m_arr, mm), am = mv < 0 ? -mv : mv;
        double v, vR, vZ;
        harmonic(mm, v, vR, vZ);
        if (mv == 0) {
            F0 = v; F0R = vR; F0Z = vZ;
            if (!P.log_scaling) { S += v; SR += vR; SZ += vZ; }
            continue;
        }
        double c, s;
        trig(am, c, s);
        const double T = (mv > 0) ? c : s, dT = (mv > 0) ? -(double)am * s : (double)am * c;
        S = fma(v, T, S); SR = fma(vR, T, SR); SZ = fma(vZ, T, SZ); Sp = fma(v, dT, Sp);
    }
    // pass 2: gradients of every harmonic's (v, vR, vZ), spread onto its cell
    double P0 = 1.0, P0R = 0.0, P0Z = 0.0;
    if (P.log_scaling) { P0 = -exp(F0); P0R = P0 * F0R; P0Z = P0 * F0Z; }
    for (int mm = 0; mm < P.n_harm; ++mm) {
        const int mv = cldi(P.m_arr, mm), am = mv < 0 ? -mv : mv;
        double gv, gvR, gvZ;
        if (mv == 0 && P.log_scaling) {
            const double gP0 = wphi * S + aR * SR + aZ * SZ + ap * Sp;
            const double gP0R = aR * S, gP0Z = aZ * S;
            gv = P0 * (gP0 + gP0R * F0R + gP0Z * F0Z);
            gvR = gP0R * P0;
            gvZ = gP0Z * P0;
        } else {
            double c = 1.0, s = 0.0;
            if (mv != 0) trig(am, c, s);
            const double T = (mv >= 0) ? c : s;
            const double dT = (mv == 0) ? 0.0 : ((mv > 0) ? -(double)am * s : (double)am * c);
            gv = P0 * (wphi * T + ap * dT) + (aR * P0R + aZ * P0Z) * T;
            gvR = aR * P0 * T;
            gvZ = aZ * P0 * T;
        }
        scatter(mm, gv, gvR, gvZ);
    }
}

__global__ __launch_bounds__(kBlock) void cyl_adjoint_kernel(
    CSK P, const double *__restrict__ xyz, int64_t n, const double *__restrict__ gphi,
    const double *__restrict__ gforce, double *__restrict__ gxyz, double *__restrict__ gnodes,
    double *__restrict__ gW)
{
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    // whole wavefronts iterate together (the outer-law sums are wavefront-wide)
    const int64_t n_pad = (n + 63) & ~(int64_t)63;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n_pad; i += stride) {
        const bool valid = i < n;
        double x = 1.0e30, y = 0.0, z = 0.0, wphi = 0.0, wf[3] = {0.0, 0.0, 0.0};
        if (valid) {
            x = xyz[3 * i]; y = xyz[3 * i + 1]; z = xyz[3 * i + 2];
            if (gphi) wphi = gphi[i];
            if (gforce) { wf[0] = gforce[3 * i]; wf[1] = gforce[3 * i + 1]; wf[2] = gforce[3 * i + 2]; }
        }
        if (valid && gxyz) {
            PointOut o{};
            cylspline_point<ORD_HESS>(P, x, y, z, o);
            gxyz[3 * i] = wphi * o.g[0] - (o.h[0] * wf[0] + o.h[3] * wf[1] + o.h[5] * wf[2]);
            gxyz[3 * i + 1] = wphi * o.g[1] - (o.h[3] * wf[0] + o.h[1] * wf[1] + o.h[4] * wf[2]);
            gxyz[3 * i + 2] = wphi * o.g[2] - (o.h[5] * wf[0] + o.h[4] * wf[1] + o.h[2] * wf[2]);
        }
        const CylCoords cc = cyl_coords(P, x, y, z);
        const double R = cc.R, lR = cc.lR, lz = cc.lz;
        const bool outside = lR < cld(P.lR, 0) || lR > cld(P.lR, P.nR - 1) ||
                             lz < cld(P.lz, 0) || lz > cld(P.lz, P.nz - 1);
        if (valid && !outside) cyl_adjoint_inner<false>(P, x, y, z, R, lR, lz, wphi, wf, gnodes, 0, 0);
        // the outer sums run for every wavefront that has an outside point
        if (__ballot(valid && outside)) cyl_adjoint_outer(P, valid && outside, x, y, z, R, wphi, wf, gW);
    }
}

// Sort key: 0 outside the grid, 1 + iR (nz - 1) + iz in cell (iR, iz).
struct CylCellKey {
    CSK P;
    const double *xyz;
    __device__ int operator()(int64_t i) const
    {
        const CylCoords c = cyl_coords(P, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        const double lR = c.lR, lz = c.lz;
        if (lR < cld(P.lR, 0) || lR > cld(P.lR, P.nR - 1) || lz < cld(P.lz, 0) ||
            lz > cld(P.lz, P.nz - 1))
            return 0;
        return 1 + cell_search(lR, P.lR, P.nR) * (P.nz - 1) + cell_search(lz, P.lz, P.nz);
    }
};

// The sorted order: 64 particles of one cell (or of the outside) per chunk.
__global__ __launch_bounds__(kBlock) void cyl_adjoint_sorted_kernel(
    CSK P, const double *__restrict__ xyz, const double *__restrict__ gphi,
    const double *__restrict__ gforce, double *__restrict__ gxyz, double *__restrict__ gnodes,
    double *__restrict__ gW, SortWork W, int nb, int run_len)
{
    for_each_sorted_chunk(W, nb, run_len, [&](int b, bool valid, int64_t i) {
        double x = 1.0e30, y = 0.0, z = 0.0, wphi = 0.0, wf[3] = {0.0, 0.0, 0.0};
        int iR = 0, iz = 0;
        if (b > 0) {
            iR = (b - 1) / (P.nz - 1);
            iz = (b - 1) - iR * (P.nz - 1);
            // padding lanes: the cell's centre, with zero seeds
            const double lRc = 0.5 * (cld(P.lR, iR) + cld(P.lR, iR + 1));
            const double lzc = 0.5 * (cld(P.lz, iz) + cld(P.lz, iz + 1));
            x = P.Rscale * sinh(lRc); z = P.Rscale * sinh(lzc);
        }
        if (valid) {
            x = xyz[3 * i]; y = xyz[3 * i + 1]; z = xyz[3 * i + 2];
            if (gphi) wphi = gphi[i];
            if (gforce) { wf[0] = gforce[3 * i]; wf[1] = gforce[3 * i + 1]; wf[2] = gforce[3 * i + 2]; }
            if (gxyz) {
                PointOut o{};
                cylspline_point<ORD_HESS>(P, x, y, z, o);
                gxyz[3 * i] = wphi * o.g[0] - (o.h[0] * wf[0] + o.h[3] * wf[1] + o.h[5] * wf[2]);
                gxyz[3 * i + 1] = wphi * o.g[1] - (o.h[3] * wf[0] + o.h[1] * wf[1] + o.h[4] * wf[2]);
                gxyz[3 * i + 2] = wphi * o.g[2] - (o.h[5] * wf[0] + o.h[4] * wf[1] + o.h[2] * wf[2]);
            }
        }
        const CylCoords cc = cyl_coords(P, x, y, z);
        const double R = cc.R;
        if (b > 0) {
            cyl_adjoint_inner<true>(P, x, y, z, R, cc.lR, cc.lz, wphi, wf, gnodes, iR, iz);
        } else {
            cyl_adjoint_outer(P, valid, x, y, z, R, wphi, wf, gW);
        }
    });
}

inline int cyl_bins(const agx_cylspline *cs)
{
    return (cs->nR - 1) * (cs->nz - 1) + 1;
}

}  // namespace agx

using namespace agx;

extern "C" int64_t agx_cylspline_adjoint_work_size(const agx_cylspline *cs, int64_t n)
{
    if (!cs || n < 0 || cs->nR < 2 || cs->nz < 2) return 0;
    const int nb = cyl_bins(cs);
    return nb <= kSortMaxBins ? sort_work_size(n, nb) : 0;
}

extern "C" int agx_cylspline_adjoint(const agx_cylspline *cs, const double *d_xyz, int64_t n,
                                     const double *d_gphi, const double *d_gforce, double *d_gxyz,
                                     double *d_gnodes, double *d_gW, void *d_work,
                                     int64_t work_bytes, void *stream)
{
    if (!cs) return AGX_E_NULL;
    if (n < 0 || n > (int64_t)0x7FFFFFFF - 64 * (int64_t)kSortMaxBins) return AGX_E_SIZE;
    if (!cs->lR_grid || !cs->lz_grid || !cs->nodes || !cs->m_arr || !cs->W_outer || !d_gnodes ||
        !d_gW)
        return AGX_E_NULL;
    if (cs->nR < 2 || cs->nz < 2 || cs->n_harm < 1) return AGX_E_SIZE;
    if (n == 0) return AGX_OK;
    if (!d_xyz || (!d_gphi && !d_gforce)) return AGX_E_NULL;
    CSK P = make_cyl_args(cs);
    hipStream_t st = (hipStream_t)stream;
    const int nb = cyl_bins(cs);
    if (nb <= kSortMaxBins) {
        if (!d_work) return AGX_E_NULL;
        if (work_bytes < sort_work_size(n, nb)) return AGX_E_WORKSPACE;
        const SortWork W = sort_work(d_work, n, nb);
        sort_launch(CylCellKey{P, d_xyz}, n, nb, W, st);
        const SortedShape sh = sorted_shape(
            n, nb, resident_workgroups((const void *)cyl_adjoint_sorted_kernel, kBlock, 0, 256 * 2),
            kBlock);
        hipLaunchKernelGGL(cyl_adjoint_sorted_kernel, dim3(sh.grid), dim3(kBlock), 0, st, P, d_xyz,
                           d_gphi, d_gforce, d_gxyz, d_gnodes, d_gW, W, nb, sh.run_len);
    } else {
        const int64_t n_pad = (n + 63) & ~(int64_t)63;
        hipLaunchKernelGGL(cyl_adjoint_kernel, dim3(grid_for(n_pad)), dim3(kBlock), 0, st, P, d_xyz,
                           n, d_gphi, d_gforce, d_gxyz, d_gnodes, d_gW);
    }
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}
