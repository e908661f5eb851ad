// Adjoint (reverse-mode) kernel of the Multipole potential+force evaluation.
//
// Given upstream gradients w_phi = dL/dPhi_i and w_F = dL/dF_i (F = -grad Phi)
// at N points, one call produces
//   dL/dx_i   = w_phi grad Phi - H w_F            (Hessian-vector product)
//   dL/dpoly  [n_int][n_lm][6]  (Horner coefficients of every term, interval)
//   dL/dscal  [invPhi0, inner U, inner W, outer U, outer W]
// The table gradient is a scatter into the interval a particle falls in.  The
// particles are first sorted by interval over the whole array (a counting
// sort into the caller's workspace: bin, scan, scatter), so that every
// wavefront of the adjoint kernel takes 64 particles of one interval (only a
// bin's last chunk is partial).  Each term's six coefficient sums are reduced
// across the wavefront by lane swaps and added to a wavefront-private LDS
// copy of that interval's table slab; the slab goes to memory (FP64 atomics)
// only when the wavefront moves to another interval.  Wavefronts take
// contiguous ranges or runs of chunks (for_each_sorted_chunk).  No reference
// counterpart: the reference has no derivatives with respect to its tables.
#include <type_traits>

#include "multipole_dev.h"
#include "sort_dev.h"

namespace agx {

constexpr int kAdjBlock = 256;
// LDS slabs are used while 4 of them (one per wavefront) fit in this many bytes
constexpr int kAdjSlabLds = 53 * 1024;

// s^j a + j s^(j-1) b for j = 0..5, summed over the wavefront and added to
// g[0..5]: a wavefront-private LDS slab (SLAB) or the table gradient in memory.
template <bool SLAB>
__device__ __forceinline__ void scatter6(double *g, double s, double a, double b, int lane)
{
    const double s2 = s * s, s3 = s2 * s, s4 = s2 * s2;
    const double v1 = fma(s, a, b), v2 = fma(s2, a, 2.0 * s * b), v3 = fma(s3, a, 3.0 * s2 * b);
    const double v4 = fma(s4, a, 4.0 * s3 * b), v5 = fma(s4 * s, a, 5.0 * s4 * b);
    if constexpr (SLAB) {
        double d, e;
        wave_sum6(a, v1, v2, v3, v4, v5, d, e);
        const int row = lane >> 4;
        if ((lane & 15) == 0) g[((row & 1) << 1) | (row >> 1)] += d;
        if ((lane & 31) == 0) g[4 + (row >> 1)] += e;
    } else {
        wave_sum6_add(g, a, v1, v2, v3, v4, v5, lane);
    }
}

// Staged table gradient (the SLAB kernels).  Per term the wavefront's 64
// particles each contribute alpha s^j + j beta s^(j-1) to the six Horner
// coefficients; instead of six cross-lane sums per term, (alpha, beta) of
// kAdjBatch terms are staged in a wavefront-private LDS block and summed in
// one pass in which lane group t (8 lanes) takes term t and each lane eight
// particles: eight products per particle and coefficient pair, then three
// 8-lane butterfly steps per coefficient, and one LDS add per coefficient and
// term into the slab.  About 25 VALU instructions per term against 49 for
// the per-term cross-lane sums.
constexpr int kAdjBatch = 8;
static int g_adj_stage = 1;      // AGX_OPT_ADJ_STAGE
struct AdjStage {
    dvec2 ab[kAdjBatch][64];     // (alpha, beta) of the staged terms, per lane
    double s[64];                // each lane's position in its interval
    int off[kAdjBatch];          // slab offsets of the staged terms
};

// Sum over the 8 lanes of each half-row (every lane gets its group's sum).
__device__ __forceinline__ double sum8(double v)
{
    v += dpp_f64<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);   // row_half_mirror
    return v;
}

// Adds the staged terms' coefficient sums to the slab g; every lane of the
// wavefront takes part (the butterfly reads all of them).
__device__ __forceinline__ void flush_stage(AdjStage *st, double *g, int count, int lane)
{
    __builtin_amdgcn_wave_barrier();
    const int t = lane >> 3, q = lane & 7;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0, a5 = 0.0;
    if (t < count) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int pt = q + 8 * u;
            const dvec2 ab = st->ab[t][pt];
            const double s = st->s[pt], s2 = s * s, s3 = s2 * s, s4 = s2 * s2;
            const double b = ab.y;
            a0 += ab.x;
            a1 += fma(ab.x, s, b);
            a2 += fma(ab.x, s2, 2.0 * b * s);
            a3 += fma(ab.x, s3, 3.0 * b * s2);
            a4 += fma(ab.x, s4, 4.0 * b * s3);
            a5 += fma(ab.x, s4 * s, 5.0 * b * s4);
        }
    }
    a0 = sum8(a0); a1 = sum8(a1); a2 = sum8(a2);
    a3 = sum8(a3); a4 = sum8(a4); a5 = sum8(a5);
    if (t < count && q == 0) {
        double *gt = g + st->off[t];
        gt[0] += a0; gt[1] += a1; gt[2] += a2; gt[3] += a3; gt[4] += a4; gt[5] += a5;
    }
    __builtin_amdgcn_wave_barrier();
}

// Table gradient of one step of the interior evaluation (the hook of
// interior<ORD_HESS>): per cos/sin term, the gradient of the quintic's six
// coefficients is s^j alpha + j s^(j-1) beta with
//   alpha = lam (w_phi Q T + w_theta dQ T + w_phi' X) + w_lnr mu Q T,
//   beta  = w_lnr lam / dlogr Q T,
// X the term's d/dphi basis over sin(theta) (its pole limit for m = 1).
template <bool SLAB, bool STAGE>
struct AdjHook {
    double *gslab;
    double *sum;                 // [4]: l > 0 sums of value, d/ds, d/dtheta, d/dphi
    double s, lam, mu, idl, wphi, wPr, wPt, wPf;
    double psg;                  // sign of cos(theta): the pole limit's sign
    int lane;
    AdjStage *st;                // STAGE: the wavefront's staging block
    mutable int staged;          // STAGE: terms staged in st (wavefront-uniform)
    // (alpha, beta) of a term at slab offset off: staged (STAGE) or summed now
    __device__ void add(int off, double alpha, double beta) const
    {
        if constexpr (STAGE) {
            st->ab[staged][lane] = dvec2{alpha, beta};
            if (lane == 0) st->off[staged] = off;
            if (++staged == kAdjBatch) {
                flush_stage(st, gslab, kAdjBatch, lane);
                staged = 0;
            }
        } else {
            scatter6<SLAB>(gslab + off, s, alpha, beta, lane);
        }
    }
    __device__ void side(int off, double Q, double dQ, int m, double T, double dT, double dTp,
                         double ist, bool pole) const
    {
        if (off < 0) return;
        const double X = pole ? ((m == 1) ? psg * dQ * dTp : 0.0) : Q * dT * ist;
        const double QT = Q * T;
        const double alpha = lam * (wphi * QT + wPt * dQ * T + wPf * X) + wPr * mu * QT;
        const double beta = wPr * lam * idl * QT;
        add(off, alpha, beta);
    }
    __device__ void term(int oc, int os, double Q, double dQ, int m, double cm, double sm,
                         double ist, bool pole) const
    {
        side(oc, Q, dQ, m, cm, -(double)m * sm, -sm, ist, pole);
        side(os, Q, dQ, m, sm, (double)m * cm, cm, ist, pole);
    }
    __device__ void sums(double Sv, double Sr, double St, double Sfn, double Sfp) const
    {
        sum[0] = Sv; sum[1] = Sr; sum[2] = St; sum[3] = Sfn;
        sum[4] = Sfp;
    }
};

// One chunk of interval k.  gslab: where this interval's coefficient sums go
// (LDS slab or memory); the invPhi0 gradient is accumulated per lane in ginv.
// One pass: the Hessian-vector product's evaluation carries the table
// gradient's scatter as its per-term hook.
template <bool SLAB, int STRIDE, bool STAGE>
__device__ __forceinline__ void adjoint_interior(const MPK &P, int k, bool valid, double x, double y,
                                                 double z, double wphi, const double wf[3],
                                                 double *gx, double *gslab, double &ginv_acc,
                                                 int lane, AdjStage *stg)
{
    const double *slab = P.poly + (int64_t)k * P.n_lm * 6;
    if (!valid) {   // a finite point inside interval k; its seeds are zero
        x = exp(fma((double)k + 0.5, P.dlogr, P.logr_min)); y = 0.0; z = 0.0;
        wphi = 0.0;
    }
    const double wfx = valid ? wf[0] : 0.0, wfy = valid ? wf[1] : 0.0, wfz = valid ? wf[2] : 0.0;
    const double r2 = fma(x, x, fma(y, y, z * z)), r = sqrt(r2), lnr = agx_log(r);
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

    // evaluation with the Hessian (for dL/dx) and the per-term scatter
    double sum[5];
    const double psg = ct < 0.0 ? -1.0 : 1.0;
    if (STAGE) stg->s[lane] = s;
    const AdjHook<SLAB, STAGE> hook{gslab, sum, s, lam, mu, idl, wphi, wPr, wPt, wPf, psg, lane,
                                    stg, 0};
    {
        PointOut o{};
        interior<ORD_HESS, true, AdjHook<SLAB, STAGE>, STRIDE>(P, slab, x, y, z, r2, r, ir, lnr, k, o,
                                                          0.0, hook);
        if (valid) {
            gx[0] = wphi * o.g[0] - (o.h[0] * wfx + o.h[3] * wfy + o.h[5] * wfz);
            gx[1] = wphi * o.g[1] - (o.h[3] * wfx + o.h[1] * wfy + o.h[4] * wfz);
            gx[2] = wphi * o.g[2] - (o.h[5] * wfx + o.h[4] * wfy + o.h[2] * wfz);
        }
    }
    const double Sv = sum[0], Sr = sum[1], St = sum[2];
    const double Sx = pole ? psg * sum[4] : sum[3] * ist;
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
    hook.add(0, gc0, gc0d);
    if (STAGE && hook.staged > 0) flush_stage(stg, gslab, hook.staged, lane);
    ginv_acc += ginv;
}

// One chunk outside the grid; the law gradients are accumulated per lane in
// acc = [inner U, inner W, outer U, outer W].
__device__ __forceinline__ void adjoint_extrap(const MPK &P, bool valid, double x, double y,
                                               double z, double wphi, const double wf[3],
                                               double *gx, double acc[4])
{
    double gU_in = 0.0, gW_in = 0.0, gU_out = 0.0, gW_out = 0.0;
    if (valid) {
        Where w = classify(P, x, y, z);
        if (w.kind == 0) {
            gW_in = wphi;
            gx[0] = gx[1] = gx[2] = 0.0;
        } else if (beyond_fp64(w)) {     // r = inf: Phi and F are 0 for every table
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
    acc[0] += gU_in; acc[1] += gW_in; acc[2] += gU_out; acc[3] += gW_out;
}

// Adds a wavefront's LDS slab to memory and clears it.
__device__ __forceinline__ void flush_slab(double *slab, double *g, int cnt, int lane)
{
    __builtin_amdgcn_wave_barrier();
    for (int q = lane; q < cnt; q += 64) {
        const double v = slab[q];
        if (v != 0.0) atomicAdd(g + q, v);
        slab[q] = 0.0;
    }
    __builtin_amdgcn_wave_barrier();
}

// The sorted order, 64 particles of one bin per wavefront and chunk.  Three
// wavefronts per SIMD (at most 168 VGPRs: a few spill; measured faster than
// two wavefronts without spills, 0.757 against 0.79 ms per gradient step).
// STRIDE 1: the program recurs by 1 only (AGX_PROG_STRIDE1); 0: any program.
// GSUM: also add minus the sum over particles of dL/dx to gscal[5..7] (the
// gradient with respect to a centre subtracted from every position).
template <bool SLAB, int STRIDE, bool GSUM, bool STAGE>
__global__ __launch_bounds__(kAdjBlock, 3) void mp_adjoint_kernel(
    MPK P, const double *__restrict__ xyz, const double *__restrict__ gphi,
    const double *__restrict__ gforce, double *__restrict__ gxyz, double *__restrict__ gpoly,
    double *__restrict__ gscal, SortWork W, int run_len)
{
    P = with_law(P);
    extern __shared__ double adj_lds[];
    const int lane = threadIdx.x & 63;
    const int slab_n = P.n_lm * 6;
    double *slab = adj_lds + (SLAB ? (threadIdx.x >> 6) * slab_n : 0);
    // STAGE: the wavefronts' staging blocks follow their slabs
    AdjStage *st = (AdjStage *)(adj_lds + (kAdjBlock / 64) * slab_n) + (threadIdx.x >> 6);
    if (SLAB)
        for (int q = lane; q < slab_n; q += 64) slab[q] = 0.0;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};   // invPhi0, inner U, W, outer U, W
    // GSUM: the wavefront's sum of dL/dx, in LDS (kept in registers across
    // the evaluation it would cost spills)
    __shared__ double gsum_lds[kAdjBlock / 64][4];
    double *gsum = gsum_lds[threadIdx.x >> 6];
    if (GSUM && lane < 3) gsum[lane] = 0.0;
    int cur = -1;                                // interval + 1 held by the slab
    for_each_sorted_chunk(W, P.n_int + 1, run_len, [&](int b, bool valid, int64_t i) {
        if (SLAB && b != cur) {
            if (cur > 0) flush_slab(slab, gpoly + (int64_t)(cur - 1) * slab_n, slab_n, lane);
            cur = b;
        }
        double x = 1.0, y = 0.0, z = 0.0, wphi = 0.0, wf[3] = {0.0, 0.0, 0.0};
        if (valid) {
            load_pos(P, xyz, i, x, y, z);
            if (gphi) wphi = gphi[i];
            if (gforce) { wf[0] = gforce[3 * i]; wf[1] = gforce[3 * i + 1]; wf[2] = gforce[3 * i + 2]; }
        }
        double gx[3] = {0.0, 0.0, 0.0};
        if (b > 0)
            adjoint_interior<SLAB, STRIDE, STAGE>(P, b - 1, valid, x, y, z, wphi, wf, gx,
                                            SLAB ? slab : gpoly + (int64_t)(b - 1) * slab_n,
                                            acc[0], lane, st);
        else
            adjoint_extrap(P, valid, x, y, z, wphi, wf, gx, acc + 1);
        if (valid && gxyz) { gxyz[3 * i] = gx[0]; gxyz[3 * i + 1] = gx[1]; gxyz[3 * i + 2] = gx[2]; }
        if (GSUM) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double v = wave_sum(valid ? gx[q] : 0.0);
                if (lane == 0) gsum[q] += v;
            }
        }
    });
    if (SLAB && cur > 0) flush_slab(slab, gpoly + (int64_t)(cur - 1) * slab_n, slab_n, lane);
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == 0 && v != 0.0) atomicAdd(gscal + q, v);
    }
    if (GSUM) {
        __builtin_amdgcn_wave_barrier();
        if (lane < 3 && gsum[lane] != 0.0) atomicAdd(gscal + 5 + lane, -gsum[lane]);
    }
}

// The adjoint's option of agx_set_option (which lives in multipole.hip).
int adj_set_stage(int64_t value)
{
    if (value < 0 || value > 1) return AGX_E_SIZE;
    g_adj_stage = (int)value;
    return AGX_OK;
}

}  // namespace agx

using namespace agx;

extern "C" int agx_multipole_adjoint(const agx_multipole *mp, const double *d_xyz, int64_t n,
                                     const double *d_gphi, const double *d_gforce, double *d_gxyz,
                                     double *d_gpoly, double *d_gscal, void *d_work,
                                     int64_t work_bytes, uint32_t flags, void *stream)
{
    if (!mp) return AGX_E_NULL;
    if (n < 0 || n > (int64_t)0x7FFFFFFF - 64 * kMaxBins) return AGX_E_SIZE;
    if (!mp->poly_km || !mp->program || !d_gpoly || !d_gscal) return AGX_E_NULL;
    if (mp->n_int < 1 || mp->n_int + 1 > kMaxBins || mp->n_lm < 1) return AGX_E_SIZE;
    if (mp->lmax > 32) return AGX_E_LMAX;
    if (mp->prog_flags & ~kProgFlags) return AGX_E_WHAT;
    if (flags & ~(AGX_WORK_SORTED | AGX_ADJ_CENTER_GRAD)) return AGX_E_WHAT;
    if (n == 0) return AGX_OK;
    if (work_bytes < sort_work_size(n, kMaxBins)) return AGX_E_WORKSPACE;
    if (!d_xyz || (!d_gphi && !d_gforce) || !d_work) return AGX_E_NULL;
    MPK P = make_args(mp);
    const SortWork W = sort_work(d_work, n, kMaxBins);
    const int nb = mp->n_int + 1;
    hipStream_t st = (hipStream_t)stream;
    if (!(flags & AGX_WORK_SORTED)) sort_launch(MpBinKey{P, d_xyz}, n, nb, W, st);
    const size_t slab_bytes = (size_t)mp->n_lm * 6 * sizeof(double) * (kAdjBlock / 64);
    const size_t stage_bytes = slab_bytes + sizeof(AdjStage) * (kAdjBlock / 64);
    const bool slab = slab_bytes <= (size_t)kAdjSlabLds;
    const bool stage = g_adj_stage && stage_bytes <= (size_t)kAdjSlabLds;
    const bool s1 = mp->prog_flags & AGX_PROG_STRIDE1;
    const bool gs = flags & AGX_ADJ_CENTER_GRAD;
    auto pick = [&](auto gsum) {
        constexpr bool G = decltype(gsum)::value;
        if (stage)
            return s1 ? mp_adjoint_kernel<true, 1, G, true> : mp_adjoint_kernel<true, 0, G, true>;
        return slab ? (s1 ? mp_adjoint_kernel<true, 1, G, false>
                          : mp_adjoint_kernel<true, 0, G, false>)
                    : (s1 ? mp_adjoint_kernel<false, 1, G, false>
                          : mp_adjoint_kernel<false, 0, G, false>);
    };
    auto kern = gs ? pick(std::true_type{}) : pick(std::false_type{});
    const size_t lds = stage ? stage_bytes : (slab ? slab_bytes : 0);
    const SortedShape sh = sorted_shape(
        n, nb, resident_workgroups((const void *)kern, kAdjBlock, lds, 256 * 3), kAdjBlock);
    // a reused sort has had its queue consumed; only the dynamic traversal reads it
    if ((flags & AGX_WORK_SORTED) && sh.run_len > 0) sort_requeue(W, st);
    hipLaunchKernelGGL(kern, dim3(sh.grid), dim3(kAdjBlock), lds, st, P, d_xyz, d_gphi, d_gforce,
                       d_gxyz, d_gpoly, d_gscal, W, sh.run_len);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}
