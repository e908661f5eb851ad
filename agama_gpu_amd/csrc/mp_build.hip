// Multipole kernel table built on the device from the raw coefficient table,
// and its adjoint: the differentiable rebuild behind torch.autograd fitting.
//
// Forward (one workgroup per kernel term c, one thread per radial node):
//   P = phi[:, cols[c]], D = dphi[:, cols[c]], dl = D r  (and the same for the
//   (0,0) column: p0, d0, dl0)
//   law = [invPhi0, inner U, inner W, outer U, outer W] from p0, d0 at the ends
//   log-scaled tables: f0 = ln(invPhi0 - 1/p0), df0 = dl0 / (p0 (invPhi0 p0 - 1)),
//                      f_c = P/p0, df_c = (dl - f_c dl0) / p0;  else f = P, df = dl
//   d2 = Mf f + Md df   (the C2 quintic system, inverted on the host)
//   poly[k][c] = Horner coefficients of the quintic on [r_k, r_k+1] in s, with
//                slopes df h and second derivatives d2 h^2 (h = dlogr)
// The same steps as the host build (multipole_prep.build_multipole_data,
// restating gpu_potential.py:408-626) with the discrete choices (columns,
// law modes) frozen.  Adjoint: the reverse of each step; the (0,0) column
// collects contributions from every term (FP64 atomics) and a last one-thread
// launch applies the law scalars' dependence on the end nodes.  No reference
// counterpart (the reference has no derivatives with respect to its tables).
#include "agx_common.h"

namespace agx {

constexpr int kBuildThreads = 128;
constexpr int kBuildMaxNodes = 128;

struct BuildArgs {
    const double *phi, *dphi, *r;
    const int *cols;
    const double *Mf, *Md;
    int nR, n_cols, n_lm, log_scaling, inner_mode, inner_reset, outer_mode;
    double dlogr;
};

// law scalars; p0a/d0a at the first node, p0b/d0b at the last
struct Law {
    double invPhi0, Uin, Win, Uout, Wout;
};

__device__ __forceinline__ Law law_of(const BuildArgs &A, double p0a, double d0a, double p0b,
                                      double d0b)
{
    Law L;
    L.invPhi0 = 0.0; L.Uin = 0.0; L.Win = p0a;
    const double ra = A.r[0], rb = A.r[A.nR - 1];
    if (A.log_scaling) {
        if (A.inner_mode == 1) {
            L.Uin = ra * d0a;
        } else if (A.inner_mode == 2) {
            L.Uin = 0.5 * ra * d0a;
            L.Win = p0a - L.Uin;
            if (!A.inner_reset) L.invPhi0 = 1.0 / L.Win;
        }
    }
    if (A.outer_mode == 1) { L.Uout = rb * d0b + p0b; L.Wout = p0b; }
    else { L.Uout = 0.0; L.Wout = p0b; }
    return L;
}

// f, df of column c at node i (LDS arrays hold the column and column 0)
struct Scaled {
    double f, df, ratio;
};

__device__ __forceinline__ Scaled scaled(const BuildArgs &A, int c, double P, double dl, double p0,
                                         double dl0, double invPhi0)
{
    Scaled s;
    s.ratio = 0.0;
    if (!A.log_scaling) { s.f = P; s.df = dl; return s; }
    if (c == 0) {
        s.f = log(invPhi0 - 1.0 / p0);
        s.df = dl0 / (p0 * (invPhi0 * p0 - 1.0));
    } else {
        s.ratio = P / p0;
        s.f = s.ratio;
        s.df = (dl - s.ratio * dl0) / p0;
    }
    return s;
}

__global__ __launch_bounds__(kBuildThreads) void mp_build_kernel(BuildArgs A, double *__restrict__ poly,
                                                                 double *__restrict__ law)
{
    __shared__ double f[kBuildMaxNodes], df[kBuildMaxNodes], d2[kBuildMaxNodes];
    const int c = blockIdx.x, nR = A.nR, col = A.cols[c], col0 = A.cols[0];
    const Law L = law_of(A, A.phi[col0], A.dphi[col0], A.phi[(nR - 1) * A.n_cols + col0],
                         A.dphi[(nR - 1) * A.n_cols + col0]);
    for (int i = threadIdx.x; i < nR; i += kBuildThreads) {
        const double P = A.phi[i * A.n_cols + col], dl = A.dphi[i * A.n_cols + col] * A.r[i];
        const double p0 = A.phi[i * A.n_cols + col0], dl0 = A.dphi[i * A.n_cols + col0] * A.r[i];
        const Scaled s = scaled(A, c, P, dl, p0, dl0, L.invPhi0);
        f[i] = s.f; df[i] = s.df;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nR; i += kBuildThreads) {
        double acc = 0.0;
        for (int j = 0; j < nR; ++j) acc = fma(A.Mf[i * nR + j], f[j], fma(A.Md[i * nR + j], df[j], acc));
        d2[i] = acc;
    }
    __syncthreads();
    const double h = A.dlogr, h2 = h * h;
    for (int k = threadIdx.x; k < nR - 1; k += kBuildThreads) {
        const double f0 = f[k], f1 = f[k + 1], m0 = df[k] * h, m1 = df[k + 1] * h;
        const double q0 = d2[k] * h2, q1 = d2[k + 1] * h2, dlt = f1 - f0;
        double *o = poly + ((int64_t)k * A.n_lm + c) * 6;
        o[0] = f0;
        o[1] = m0;
        o[2] = 0.5 * q0;
        o[3] = 10.0 * dlt - 6.0 * m0 - 4.0 * m1 - 1.5 * q0 + 0.5 * q1;
        o[4] = -15.0 * dlt + 8.0 * m0 + 7.0 * m1 + 1.5 * q0 - q1;
        o[5] = 6.0 * dlt - 3.0 * m0 - 3.0 * m1 + 0.5 * (q1 - q0);
    }
    if (c == 0 && threadIdx.x == 0) {
        law[0] = L.invPhi0; law[1] = L.Uin; law[2] = L.Win; law[3] = L.Uout; law[4] = L.Wout;
    }
}

// The law scalars' dependence on the (0,0) column's end nodes, given
// dL/dinvPhi0 (ginv): one thread, with atomics (the term kernels add to the
// same entries concurrently).
__device__ void law_adjoint(const BuildArgs &A, const double *__restrict__ glaw,
                            double *__restrict__ gphi, double *__restrict__ gdphi, double ginv)
{
    const int nR = A.nR, col0 = A.cols[0];
    const Law L = law_of(A, A.phi[col0], A.dphi[col0], A.phi[(nR - 1) * A.n_cols + col0],
                         A.dphi[(nR - 1) * A.n_cols + col0]);
    const double ra = A.r[0], rb = A.r[nR - 1];
    double gWin = glaw[2];
    double gp0a = 0.0, gd0a = 0.0;
    if (A.log_scaling) {
        if (A.inner_mode == 1) {
            gp0a += gWin;
            gd0a += glaw[1] * ra;
        } else if (A.inner_mode == 2) {
            if (!A.inner_reset) gWin -= ginv / (L.Win * L.Win);
            gp0a += gWin;
            gd0a += (glaw[1] - gWin) * 0.5 * ra;
        } else {
            gp0a += gWin;
        }
    } else {
        gp0a += gWin;
    }
    atomicAdd(&gphi[col0], gp0a);
    if (gd0a != 0.0) atomicAdd(&gdphi[col0], gd0a);
    const int64_t e = (int64_t)(nR - 1) * A.n_cols + col0;
    if (A.outer_mode == 1) {
        atomicAdd(&gphi[e], glaw[3] + glaw[4]);
        atomicAdd(&gdphi[e], glaw[3] * rb);
    } else {
        atomicAdd(&gphi[e], glaw[4]);
    }
}

// Reverse of mp_build_kernel for term c: adds dL/dphi, dL/dphi' of column
// cols[c] (plain adds for c > 0: one writer) and of the (0,0) column
// (atomics, term 0's own included: every term's log-scaling depends on it).
// Term 0 also writes work[0] = dL/dinvPhi0 (its own uses plus the law
// output's) and adds the law scalars' gradients (law_adjoint).
__global__ __launch_bounds__(kBuildThreads) void mp_build_adjoint_kernel(
    BuildArgs A, const double *__restrict__ gpoly, const double *__restrict__ glaw,
    double *__restrict__ gphi, double *__restrict__ gdphi, double *__restrict__ work)
{
    __shared__ double gf[kBuildMaxNodes], gdf[kBuildMaxNodes], gd2[kBuildMaxNodes];
    __shared__ double red[kBuildThreads];
    const int c = blockIdx.x, nR = A.nR, col = A.cols[c], col0 = A.cols[0];
    const Law L = law_of(A, A.phi[col0], A.dphi[col0], A.phi[(nR - 1) * A.n_cols + col0],
                         A.dphi[(nR - 1) * A.n_cols + col0]);
    const double h = A.dlogr, h2 = h * h;
    // Horner coefficients -> (f, m = df h, q = d2 h^2) at both ends of each interval
    for (int i = threadIdx.x; i < nR; i += kBuildThreads) {
        double a = 0.0, b = 0.0, q = 0.0;
        if (i < nR - 1) {   // node i as the left end of interval i
            const double *g = gpoly + ((int64_t)i * A.n_lm + c) * 6;
            a += g[0] - 10.0 * g[3] + 15.0 * g[4] - 6.0 * g[5];
            b += g[1] - 6.0 * g[3] + 8.0 * g[4] - 3.0 * g[5];
            q += 0.5 * g[2] - 1.5 * g[3] + 1.5 * g[4] - 0.5 * g[5];
        }
        if (i > 0) {        // node i as the right end of interval i - 1
            const double *g = gpoly + ((int64_t)(i - 1) * A.n_lm + c) * 6;
            a += 10.0 * g[3] - 15.0 * g[4] + 6.0 * g[5];
            b += -4.0 * g[3] + 7.0 * g[4] - 3.0 * g[5];
            q += 0.5 * g[3] - g[4] + 0.5 * g[5];
        }
        gf[i] = a; gdf[i] = b * h; gd2[i] = q * h2;
    }
    __syncthreads();
    // d2 = Mf f + Md df
    for (int i = threadIdx.x; i < nR; i += kBuildThreads) {
        double af = 0.0, ad = 0.0;
        for (int j = 0; j < nR; ++j) {
            af = fma(A.Mf[j * nR + i], gd2[j], af);
            ad = fma(A.Md[j * nR + i], gd2[j], ad);
        }
        gf[i] += af; gdf[i] += ad;
    }
    __syncthreads();
    double ginv = 0.0;
    for (int i = threadIdx.x; i < nR; i += kBuildThreads) {
        const double ri = A.r[i];
        const double P = A.phi[i * A.n_cols + col], dl = A.dphi[i * A.n_cols + col] * ri;
        const double p0 = A.phi[i * A.n_cols + col0], dl0 = A.dphi[i * A.n_cols + col0] * ri;
        double gP = 0.0, gdl = 0.0, gp0 = 0.0, gdl0 = 0.0;
        if (!A.log_scaling) {
            gP = gf[i]; gdl = gdf[i];
        } else if (c == 0) {
            const double a = L.invPhi0 - 1.0 / p0, bq = p0 * (L.invPhi0 * p0 - 1.0);
            gp0 = gf[i] / (a * p0 * p0) - gdf[i] * dl0 * (2.0 * L.invPhi0 * p0 - 1.0) / (bq * bq);
            ginv += gf[i] / a - gdf[i] * dl0 * p0 * p0 / (bq * bq);
            gdl0 = gdf[i] / bq;
        } else {
            const double ratio = P / p0, dfc = (dl - ratio * dl0) / p0;
            const double gr = gf[i] - gdf[i] * dl0 / p0;
            gdl = gdf[i] / p0;
            gdl0 = -gdf[i] * ratio / p0;
            gp0 = -gdf[i] * dfc / p0 - gr * ratio / p0;
            gP = gr / p0;
        }
        if (c == 0) {   // the other terms add to the same entries concurrently
            atomicAdd(&gphi[i * A.n_cols + col], gP + gp0);
            atomicAdd(&gdphi[i * A.n_cols + col], (gdl + gdl0) * ri);
        } else {
            gphi[i * A.n_cols + col] += gP;
            gdphi[i * A.n_cols + col] += gdl * ri;
            if (gp0 != 0.0) atomicAdd(&gphi[i * A.n_cols + col0], gp0);
            if (gdl0 != 0.0) atomicAdd(&gdphi[i * A.n_cols + col0], gdl0 * ri);
        }
    }
    if (c == 0) {
        red[threadIdx.x] = ginv;
        __syncthreads();
        for (int s = kBuildThreads / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            work[0] = red[0] + glaw[0];
            law_adjoint(A, glaw, gphi, gdphi, red[0] + glaw[0]);
        }
    }
}

static int build_args(const agx_multipole_source *s, BuildArgs &A)
{
    if (!s) return AGX_E_NULL;
    if (!s->phi || !s->dphi || !s->r || !s->cols || !s->Mf || !s->Md) return AGX_E_NULL;
    if (s->nR < 2 || s->nR > kBuildMaxNodes || s->n_lm < 1 || s->n_cols < s->n_lm) return AGX_E_SIZE;
    if (s->inner_mode < 0 || s->inner_mode > 2 || s->outer_mode < 0 || s->outer_mode > 1)
        return AGX_E_WHAT;
    A.phi = s->phi; A.dphi = s->dphi; A.r = s->r; A.cols = s->cols; A.Mf = s->Mf; A.Md = s->Md;
    A.nR = s->nR; A.n_cols = s->n_cols; A.n_lm = s->n_lm;
    A.log_scaling = s->log_scaling; A.inner_mode = s->inner_mode; A.inner_reset = s->inner_reset;
    A.outer_mode = s->outer_mode; A.dlogr = s->dlogr;
    return AGX_OK;
}

}  // namespace agx

using namespace agx;

extern "C" int agx_multipole_build(const agx_multipole_source *src, double *d_poly, double *d_law,
                                   void *stream)
{
    BuildArgs A;
    const int rc = build_args(src, A);
    if (rc != AGX_OK) return rc;
    if (!d_poly || !d_law) return AGX_E_NULL;
    hipLaunchKernelGGL(mp_build_kernel, dim3(A.n_lm), dim3(kBuildThreads), 0, (hipStream_t)stream, A,
                       d_poly, d_law);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}

extern "C" int agx_multipole_build_adjoint(const agx_multipole_source *src, const double *d_gpoly,
                                           const double *d_glaw, double *d_gphi, double *d_gdphi,
                                           double *d_work, void *stream)
{
    BuildArgs A;
    const int rc = build_args(src, A);
    if (rc != AGX_OK) return rc;
    if (!d_gpoly || !d_glaw || !d_gphi || !d_gdphi || !d_work) return AGX_E_NULL;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(mp_build_adjoint_kernel, dim3(A.n_lm), dim3(kBuildThreads), 0, st, A, d_gpoly,
                       d_glaw, d_gphi, d_gdphi, d_work);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}
