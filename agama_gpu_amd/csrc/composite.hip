// Fused composite evaluation: any sum of analytic and expansion components,
// each under its own shift/scale/amplitude transform, in ONE launch.
//
// The reference evaluates a Composite as one launch per component plus an
// elementwise add per output (gpu_potential.py:1477-1510), a Shifted
// potential as a host spline evaluation plus a subtract kernel
// (gpu_potential.py:1720-1744), a Scaled one as two more elementwise kernels
// (gpu_potential.py:1878-1900) and an Evolving one as two full evaluations and
// a lerp (gpu_potential.py:1577-1644).  Here the wrapper tree is flattened on
// the host into leaves, each with
//     x_leaf = sigma (x - center),  Phi += A Phi_leaf(x_leaf),
//     grad  += A sigma grad_leaf,   Hessian += A sigma^2 H_leaf,
//     rho   += A sigma^2 rho_leaf
// (A carries Scaled amplitudes, 1/scale factors and Evolving lerp weights),
// and every particle runs through all leaves once: one read of the positions,
// one write of each output.
#include "leaf_dev.h"

namespace agx {

constexpr int kMaxComponents = AGX_MAX_COMPONENTS;

struct CompositeArgs {
    int n;
    int defer;          // the CylSpline leaf whose points outside its grid box are
                        // deferred (see composite_kernel), or -1
    agx_component c[kMaxComponents];
};

// STRIDE and SIN as for interior(), from the program flags every Multipole
// leaf's table satisfies.
//
// With A.defer >= 0 the points outside that CylSpline leaf's grid box are
// deferred: a lane adds every other leaf (and the spline part when its point
// is inside the box) and, for a point outside, appends the point's index and
// partial sums to a workgroup queue in LDS instead of evaluating the outer
// expansion; whenever the queue holds a workgroup's worth (kBlock) of points,
// every lane takes one, adds the outer expansion and writes the outputs, and
// the rest are taken at the end.  The outer expansion (lmax 8) then runs in
// full wavefronts instead of in nearly every wavefront (almost every one of
// them holds a point outside the box).  The deferred leaf's grids are read
// from a copy in LDS; a leaf whose grids do not fit there (nR + nz >
// kGridLDS) is evaluated point by point instead.  Every output is written
// once.
template <int ORD, int STRIDE, bool SIN>
__global__ __launch_bounds__(kBlock, ORD == ORD_HESS ? 3 : 1) void composite_kernel(CompositeArgs A,
                                                           const double *__restrict__ xyz,
                                                           int64_t n, double *__restrict__ phi,
                                                           double *__restrict__ force,
                                                           double *__restrict__ deriv)
{
    constexpr bool G = ORD == ORD_GRAD || ORD == ORD_HESS;
    constexpr bool H = ORD == ORD_HESS;
    constexpr int NQ = 2 * kBlock;            // the queue never holds 2 kBlock points
    __shared__ int64_t q_idx[NQ];
    __shared__ double q_sum[(1 + (G ? 3 : 0) + (H ? 6 : 0)) * NQ];
    __shared__ double grids[kGridLDS];         // the deferred leaf's lR and lz grids
    __shared__ int queued;
    const int t = threadIdx.x;
    // the deferred leaf (the outer expansion has no density), when its grids
    // fit in LDS; otherwise it is evaluated point by point like other leaves
    int dc = ORD == ORD_DENS ? -1 : A.defer;
    if (dc >= 0) {
        const CSK cs = load_desc<CSK>(A.c[dc].table);
        if (cs.nR + cs.nz <= kGridLDS) stage_grids(cs, grids);   // (uniform branch)
        else dc = -1;
    }
    auto store = [&](int64_t i, double P, const double g[3], const double h[6]) {
        if (phi) __builtin_nontemporal_store(P, phi + i);    // streaming: written once
        if (G && force)
            for (int k = 0; k < 3; ++k) __builtin_nontemporal_store(-g[k], force + 3 * i + k);
        if (H && deriv)
            for (int k = 0; k < 6; ++k) __builtin_nontemporal_store(-h[k], deriv + 6 * i + k);
    };
    // the deferred leaf's outer expansion for queue entry e, then the outputs
    auto finish = [&](int e) {
        const int64_t j = q_idx[e];
        double P = q_sum[e], g[3] = {0.0, 0.0, 0.0}, h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (G)
            for (int k = 0; k < 3; ++k) g[k] = q_sum[(1 + k) * NQ + e];
        if (H)
            for (int k = 0; k < 6; ++k) h[k] = q_sum[(4 + k) * NQ + e];
        cyl_leaf_outside<ORD>(A.c[dc], xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2], P, g, h);
        store(j, P, g, h);
    };
    if (t == 0) queued = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += stride) {
        const int64_t i = base + t;
        if (i < n) {
            const double x0 = xyz[3 * i], y0 = xyz[3 * i + 1], z0 = xyz[3 * i + 2];
            double P = 0.0, g[3] = {0.0, 0.0, 0.0}, h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            bool later = false;
            for (int c = 0; c < A.n; ++c) {
                const agx_component &C = A.c[c];
                if (c == dc)
                    later = !cyl_leaf_inside<ORD>(C, x0, y0, z0, P, g, h, grids);
                else
                    add_leaf<ORD, true, STRIDE, SIN>(C.type, C.p, C.table, C.ampl, C.scale,
                                                     C.center, x0, y0, z0, P, g, h);
            }
            if (later) {
                const int e = atomicAdd(&queued, 1);
                q_idx[e] = i;
                q_sum[e] = P;
                if (G)
                    for (int k = 0; k < 3; ++k) q_sum[(1 + k) * NQ + e] = g[k];
                if (H)
                    for (int k = 0; k < 6; ++k) q_sum[(4 + k) * NQ + e] = h[k];
            } else {
                store(i, P, g, h);
            }
        }
        if (dc < 0) continue;                    // nothing is ever queued: no barriers
        __syncthreads();
        const int qn = queued;
        // every lane holds the same qn before any lane can add to the queue
        // again (the branch below has barriers)
        __syncthreads();
        if (qn >= kBlock) {
            finish(qn - kBlock + t);
            __syncthreads();                     // every lane has read its entry
            if (t == 0) queued = qn - kBlock;
            __syncthreads();
        }
    }
    if (dc >= 0) {
        const int qn = queued;
        for (int e = t; e < qn; e += kBlock) finish(e);
    }
}

// Sums whose leaves are all analytic types, which are bound by HBM.  Each
// workgroup moves tiles of kBlock particles through LDS, so that the global
// reads of the positions and the writes of force and deriv are contiguous
// 8-byte accesses across the lanes (instead of one access per component at a
// 24- or 48-byte stride per lane), and the next tile's positions are read
// into registers while the current tile computes.  No expansion code is
// compiled in.
template <int ORD>
__global__ __launch_bounds__(kBlock) void analytic_kernel(CompositeArgs A,
                                                          const double *__restrict__ xyz,
                                                          int64_t n, double *__restrict__ phi,
                                                          double *__restrict__ force,
                                                          double *__restrict__ deriv)
{
    constexpr bool G = ORD == ORD_GRAD || ORD == ORD_HESS;
    constexpr bool H = ORD == ORD_HESS;
    __shared__ double L[(H ? 6 : 3) * kBlock];
    const int t = threadIdx.x;
    const int64_t tstride = (int64_t)gridDim.x * kBlock;
    int64_t base = (int64_t)blockIdx.x * kBlock;
    double pre[3];
    auto fetch = [&](int64_t b) {          // the 3 kBlock position doubles from particle b on
        const int64_t avail = 3 * (n - b);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int j = t + k * kBlock;
            pre[k] = (j < avail) ? xyz[3 * b + j] : 0.0;
        }
    };
    // a contiguous run of m doubles per particle from LDS to out + m * base
    auto flush = [&](double *out, int m, int64_t b) {
        const int64_t avail = (int64_t)m * (n - b);
        for (int j = t; j < m * kBlock; j += kBlock)
            if (j < avail) __builtin_nontemporal_store(L[j], out + (int64_t)m * b + j);
    };
    if (base < n) fetch(base);
    for (; base < n; base += tstride) {
#pragma unroll
        for (int k = 0; k < 3; ++k) L[t + k * kBlock] = pre[k];
        __syncthreads();
        const double x0 = L[3 * t], y0 = L[3 * t + 1], z0 = L[3 * t + 2];
        if (base + tstride < n) fetch(base + tstride);
        const int64_t i = base + t;
        double P = 0.0, g[3] = {0.0, 0.0, 0.0}, h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (i < n) {
            for (int c = 0; c < A.n; ++c) {
                const agx_component &C = A.c[c];
                add_leaf<ORD, false>(C.type, C.p, C.table, C.ampl, C.scale, C.center, x0, y0, z0,
                                     P, g, h);
            }
            if (phi) __builtin_nontemporal_store(P, phi + i);
        }
        __syncthreads();                   // the positions in L are read
        if (G && force) {
#pragma unroll
            for (int k = 0; k < 3; ++k) L[3 * t + k] = -g[k];
            __syncthreads();
            flush(force, 3, base);
            __syncthreads();
        }
        if (H && deriv) {
#pragma unroll
            for (int k = 0; k < 6; ++k) L[6 * t + k] = -h[k];
            __syncthreads();
            flush(deriv, 6, base);
            __syncthreads();
        }
    }
}

static int check_components(const agx_component *comps, int n_comps, CompositeArgs &A,
                            bool &any_expansion, uint32_t &mp_flags)
{
    if (n_comps < 1 || n_comps > kMaxComponents) return AGX_E_SIZE;
    if (!comps) return AGX_E_NULL;
    A.n = n_comps;
    A.defer = -1;
    any_expansion = false;
    // AND over the Multipole leaves (of the stride and sin flags: the
    // composite kernel does not use the triaxial or full degree)
    mp_flags = kProgFlags & ~(AGX_PROG_TRIAXIAL_MASK | AGX_PROG_FULL_MASK);
    for (int c = 0; c < n_comps; ++c) {
        const agx_component &C = comps[c];
        const bool expansion = C.type == AGX_T_MULTIPOLE || C.type == AGX_T_CYLSPLINE;
        const bool analytic = C.type >= AGX_T_NFW && C.type <= AGX_T_UNIFORM_ACCEL;
        if (!expansion && !analytic) return AGX_E_TYPE;
        if (expansion && !C.table) return AGX_E_NULL;
        any_expansion = any_expansion || expansion;
        if (C.type == AGX_T_MULTIPOLE) {
            if (C.prog_flags & ~kProgFlags) return AGX_E_WHAT;
            mp_flags &= C.prog_flags;
        }
        if (C.type == AGX_T_CYLSPLINE && A.defer < 0 && cyl_path() == 0) A.defer = c;
        A.c[c] = C;
    }
    return AGX_OK;
}

template <int ORD>
static void launch_composite(const CompositeArgs &A, bool expansion, uint32_t mp_flags,
                             const double *xyz, int64_t n, double *phi, double *force,
                             double *deriv, hipStream_t s)
{
    // three Multipole variants: dense programs (stride 1), the m >= 0 same-parity
    // programs of triaxial and axisymmetric tables (stride 2, no sin terms), and
    // the generic code for anything else
    const int st = prog_stride(mp_flags);
    if (expansion && st == 1)
        hipLaunchKernelGGL((composite_kernel<ORD, 1, true>), dim3(grid_for(n)), dim3(kBlock), 0,
                           s, A, xyz, n, phi, force, deriv);
    else if (expansion && st == 2 && (mp_flags & AGX_PROG_COS_ONLY))
        hipLaunchKernelGGL((composite_kernel<ORD, 2, false>), dim3(grid_for(n)), dim3(kBlock), 0,
                           s, A, xyz, n, phi, force, deriv);
    else if (expansion)
        hipLaunchKernelGGL((composite_kernel<ORD, 0, true>), dim3(grid_for(n)), dim3(kBlock), 0,
                           s, A, xyz, n, phi, force, deriv);
    else
        hipLaunchKernelGGL(analytic_kernel<ORD>, dim3(grid_for(n)), dim3(kBlock), 0, s, A, xyz, n,
                           phi, force, deriv);
}

}  // namespace agx

using namespace agx;

extern "C" int agx_composite_eval(const agx_component *comps, int n_comps, const double *d_xyz,
                                  int64_t n, double *d_phi, double *d_force, double *d_deriv,
                                  uint32_t what, void *stream)
{
    CompositeArgs A;
    bool exp_leaves;
    uint32_t mp_flags;
    int rc = check_components(comps, n_comps, A, exp_leaves, mp_flags);
    if (rc) return rc;
    if (n < 0) return AGX_E_SIZE;
    if (what == 0 || (what & ~(AGX_PHI | AGX_FORCE | AGX_DERIV))) return AGX_E_WHAT;
    if (n == 0) return AGX_OK;
    if (!d_xyz || ((what & AGX_PHI) && !d_phi) || ((what & AGX_FORCE) && !d_force) ||
        ((what & AGX_DERIV) && !d_deriv))
        return AGX_E_NULL;
    double *phi = (what & AGX_PHI) ? d_phi : nullptr;
    double *force = (what & AGX_FORCE) ? d_force : nullptr;
    double *deriv = (what & AGX_DERIV) ? d_deriv : nullptr;
    hipStream_t s = (hipStream_t)stream;
    if (deriv)
        launch_composite<ORD_HESS>(A, exp_leaves, mp_flags, d_xyz, n, phi, force, deriv, s);
    else if (force)
        launch_composite<ORD_GRAD>(A, exp_leaves, mp_flags, d_xyz, n, phi, force, nullptr, s);
    else
        launch_composite<ORD_POT>(A, exp_leaves, mp_flags, d_xyz, n, phi, nullptr, nullptr, s);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}

extern "C" int agx_composite_density(const agx_component *comps, int n_comps, const double *d_xyz,
                                     int64_t n, double *d_rho, void *stream)
{
    CompositeArgs A;
    bool exp_leaves;
    uint32_t mp_flags;
    int rc = check_components(comps, n_comps, A, exp_leaves, mp_flags);
    if (rc) return rc;
    if (n < 0) return AGX_E_SIZE;
    if (n == 0) return AGX_OK;
    if (!d_xyz || !d_rho) return AGX_E_NULL;
    launch_composite<ORD_DENS>(A, exp_leaves, mp_flags, d_xyz, n, d_rho, nullptr, nullptr,
                               (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}

extern "C" int64_t agx_multipole_pack_size(void) { return (int64_t)sizeof(MPK); }

extern "C" int64_t agx_cylspline_pack_size(void) { return (int64_t)sizeof(CSK); }

extern "C" int agx_cylspline_pack(const agx_cylspline *cs, void *dst, int64_t size)
{
    if (!cs || !dst) return AGX_E_NULL;
    if (size < (int64_t)sizeof(CSK)) return AGX_E_WORKSPACE;
    CSK P = make_cyl_args(cs);
    __builtin_memcpy(dst, &P, sizeof(CSK));
    return AGX_OK;
}

extern "C" int agx_multipole_pack(const agx_multipole *mp, void *dst, int64_t size)
{
    if (!mp || !dst) return AGX_E_NULL;
    if (size < (int64_t)sizeof(MPK)) return AGX_E_WORKSPACE;
    if (mp->origin) return AGX_E_WHAT;   // a leaf's centre is agx_component.center
    MPK P = make_args(mp);
    __builtin_memcpy(dst, &P, sizeof(MPK));
    return AGX_OK;
}
