// CylSpline kernels and C-ABI entry points (gfx950, FP64).
#include "cylspline_dev.h"

namespace agx {

template <int ORD>
__device__ __forceinline__ void cyl_store(const PointOut &o, int64_t i, double *phi, double *force,
                                          double *deriv)
{
    // streaming (non-temporal) stores: written once, not read again
    if (phi) __builtin_nontemporal_store(o.phi, phi + i);
    if ((ORD == ORD_GRAD || ORD == ORD_HESS) && force)
        for (int k = 0; k < 3; ++k) __builtin_nontemporal_store(-o.g[k], force + 3 * i + k);
    if (ORD == ORD_HESS && deriv)
        for (int k = 0; k < 6; ++k) __builtin_nontemporal_store(-o.h[k], deriv + 6 * i + k);
}

// One point per lane, whichever branch it takes (the lanes of a wavefront
// with points inside and outside the grid box run both).
template <int ORD>
__global__ __launch_bounds__(kBlock) void cyl_kernel(CSK P, const double *__restrict__ xyz,
                                                     int64_t n, double *__restrict__ phi,
                                                     double *__restrict__ force,
                                                     double *__restrict__ deriv)
{
    for_each_point<kBlock>(xyz, n, [&](int64_t i, double x, double y, double z) {
        PointOut o{};
        cylspline_point<ORD>(P, x, y, z, o);
        cyl_store<ORD>(o, i, phi, force, deriv);
    });
}

// The same evaluation with the points outside the grid box deferred: each
// lane evaluates its point if it lies inside the box and otherwise appends its
// index to a workgroup queue in LDS; whenever the queue holds a workgroup's
// worth (kBlock) of points, every lane takes one and evaluates the outer
// expansion, and the rest are taken at the end.  The outer expansion (lmax 8:
// about as many instructions as the whole spline part) then runs in full
// wavefronts; without the queue nearly every wavefront holds a point outside
// the box and runs both branches.  The queue never holds 2 kBlock points:
// before a block of points it holds fewer than kBlock.  The grids are read
// from a copy in LDS (the launch requires nR + nz <= kGridLDS).
template <int ORD>
__global__ __launch_bounds__(kBlock) void cyl_deferred_kernel(CSK P, const double *__restrict__ xyz,
                                                              int64_t n, double *__restrict__ phi,
                                                              double *__restrict__ force,
                                                              double *__restrict__ deriv)
{
    __shared__ int64_t queue[2 * kBlock];
    __shared__ double grids[kGridLDS];         // lR then lz (nR + nz <= kGridLDS)
    __shared__ int queued;
    const int t = threadIdx.x;
    if (t == 0) queued = 0;
    stage_grids(P, grids);
    auto outer = [&](int64_t j) {
        const double x = xyz[3 * j], y = xyz[3 * j + 1], z = xyz[3 * j + 2];
        const CylCoords cc = cyl_coords(P, x, y, z);
        PointOut o{};
        cylspline_outside<ORD>(P, x, y, z, cc, o);
        cyl_store<ORD>(o, j, phi, force, deriv);
    };
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += stride) {
        const int64_t i = base + t;
        if (i < n) {
            const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
            const CylCoords cc = cyl_coords(P, x, y, z);
            if (cyl_outside(P, cc)) {
                queue[atomicAdd(&queued, 1)] = i;
            } else {
                PointOut o{};
                cylspline_inside<ORD, true>(P, x, y, z, cc, o, grids, grids + P.nR);
                cyl_store<ORD>(o, i, phi, force, deriv);
            }
        }
        __syncthreads();
        const int q = queued;
        // every lane holds the same q before any lane can add to the queue
        // again (a lane running ahead into the next block must not change
        // what a slower lane reads here: the branch below has barriers)
        __syncthreads();
        if (q >= kBlock) {
            outer(queue[q - kBlock + t]);
            __syncthreads();                     // every lane has read its entry
            if (t == 0) queued = q - kBlock;
            __syncthreads();
        }
    }
    const int q = queued;
    for (int k = t; k < q; k += kBlock) outer(queue[k]);
}

// Tunable (agx_set_option AGX_OPT_CYL_PATH): 0 deferred outer points, 1 one
// point per lane.
static int g_cyl_path = 0;

static int check_cs(const agx_cylspline *cs, const double *xyz, int64_t n)
{
    if (!cs) return AGX_E_NULL;
    if (n < 0) return AGX_E_SIZE;
    if (n > 0 && !xyz) return AGX_E_NULL;
    if (!cs->lR_grid || !cs->lz_grid || !cs->nodes || !cs->m_arr || !cs->W_outer) return AGX_E_NULL;
    if (cs->nR < 2 || cs->nz < 2 || cs->n_harm < 1) return AGX_E_SIZE;
    if (cs->lmax_outer < 0 || cs->lmax_outer > 8 || cs->mmax_outer > cs->lmax_outer) return AGX_E_LMAX;
    return AGX_OK;
}

}  // namespace agx

using namespace agx;

extern "C" int agx_cylspline_eval(const agx_cylspline *cs, const double *d_xyz, int64_t n,
                                  double *d_phi, double *d_force, double *d_deriv, uint32_t what,
                                  void *stream)
{
    int rc = check_cs(cs, d_xyz, n);
    if (rc) return rc;
    if (what == 0 || (what & ~(AGX_PHI | AGX_FORCE | AGX_DERIV))) return AGX_E_WHAT;
    if (n == 0) return AGX_OK;
    if (((what & AGX_PHI) && !d_phi) || ((what & AGX_FORCE) && !d_force) ||
        ((what & AGX_DERIV) && !d_deriv))
        return AGX_E_NULL;
    CSK P = make_cyl_args(cs);
    hipStream_t s = (hipStream_t)stream;
    double *phi = (what & AGX_PHI) ? d_phi : nullptr;
    double *force = (what & AGX_FORCE) ? d_force : nullptr;
    double *deriv = (what & AGX_DERIV) ? d_deriv : nullptr;
    dim3 grid(grid_for(n)), block(kBlock);
    const bool defer = g_cyl_path == 0 && cs->nR + cs->nz <= kGridLDS;
    if (deriv)
        hipLaunchKernelGGL(defer ? cyl_deferred_kernel<ORD_HESS> : cyl_kernel<ORD_HESS>, grid, block,
                           0, s, P, d_xyz, n, phi, force, deriv);
    else if (force)
        hipLaunchKernelGGL(defer ? cyl_deferred_kernel<ORD_GRAD> : cyl_kernel<ORD_GRAD>, grid, block,
                           0, s, P, d_xyz, n, phi, force, nullptr);
    else
        hipLaunchKernelGGL(defer ? cyl_deferred_kernel<ORD_POT> : cyl_kernel<ORD_POT>, grid, block,
                           0, s, P, d_xyz, n, phi, nullptr, nullptr);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}

extern "C" int agx_cylspline_density(const agx_cylspline *cs, const double *d_xyz, int64_t n,
                                     double *d_rho, void *stream)
{
    int rc = check_cs(cs, d_xyz, n);
    if (rc) return rc;
    if (n == 0) return AGX_OK;
    if (!d_rho) return AGX_E_NULL;
    CSK P = make_cyl_args(cs);
    const bool defer = g_cyl_path == 0 && cs->nR + cs->nz <= kGridLDS;
    hipLaunchKernelGGL(defer ? cyl_deferred_kernel<ORD_DENS> : cyl_kernel<ORD_DENS>,
                       dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, P, d_xyz, n, d_rho,
                       nullptr, nullptr);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}

namespace agx {
// The CylSpline path option of agx_set_option (which lives in multipole.hip).
int cyl_set_path(int64_t value)
{
    if (value < 0 || value > 1) return AGX_E_SIZE;
    g_cyl_path = (int)value;
    return AGX_OK;
}
int cyl_path() { return g_cyl_path; }
}  // namespace agx
