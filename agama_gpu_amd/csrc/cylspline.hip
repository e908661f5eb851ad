// CylSpline kernels and C-ABI entry points (gfx950, FP64).
#include "cylspline_dev.h"

namespace agx {

template <int ORD>
__global__ __launch_bounds__(kBlock) void cyl_kernel(CSK P, const double *__restrict__ xyz,
                                                     int64_t n, double *__restrict__ phi,
                                                     double *__restrict__ force,
                                                     double *__restrict__ deriv)
{
    for_each_point<kBlock>(xyz, n, [&](int64_t i, double x, double y, double z) {
        PointOut o{};
        cylspline_point<ORD>(P, x, y, z, o);
        // streaming (non-temporal) stores: written once, not read again
        if (phi) __builtin_nontemporal_store(o.phi, phi + i);
        if ((ORD == ORD_GRAD || ORD == ORD_HESS) && force)
            for (int k = 0; k < 3; ++k) __builtin_nontemporal_store(-o.g[k], force + 3 * i + k);
        if (ORD == ORD_HESS && deriv)
            for (int k = 0; k < 6; ++k) __builtin_nontemporal_store(-o.h[k], deriv + 6 * i + k);
    });
}

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
    if (deriv)
        hipLaunchKernelGGL(cyl_kernel<ORD_HESS>, grid, block, 0, s, P, d_xyz, n, phi, force, deriv);
    else if (force)
        hipLaunchKernelGGL(cyl_kernel<ORD_GRAD>, grid, block, 0, s, P, d_xyz, n, phi, force, nullptr);
    else
        hipLaunchKernelGGL(cyl_kernel<ORD_POT>, grid, block, 0, s, P, d_xyz, n, phi, nullptr, nullptr);
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
    hipLaunchKernelGGL(cyl_kernel<ORD_DENS>, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                       P, d_xyz, n, d_rho, nullptr, nullptr);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}
