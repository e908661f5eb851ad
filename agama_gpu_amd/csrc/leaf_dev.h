// One leaf of a flattened potential tree: a native component (analytic type,
// Multipole or CylSpline descriptor image) evaluated at sigma (x - center)
// and accumulated with weight A:
//     Phi += A Phi_leaf,  grad += A sigma grad_leaf,  H += A sigma^2 H_leaf,
//     rho += A sigma^2 rho_leaf  (rho returned in P for ORD_DENS).
// Shared by the fused composite kernel and the leapfrog kernel.
#pragma once
#include "analytic_dev.h"
#include "cylspline_dev.h"
#include "multipole_dev.h"

namespace agx {

// Adds one leaf's value (rho for ORD_DENS), gradient and Hessian with the
// leaf's weights ampl, ampl sigma, ampl sigma^2 (rho: ampl sigma^2).
template <int ORD>
__device__ __forceinline__ void accumulate(double phi, double rho, const double eg[3],
                                           const double eh[6], double ampl, double sg, double &P,
                                           double g[3], double h[6])
{
    constexpr bool G = ORD == ORD_GRAD || ORD == ORD_HESS;
    constexpr bool H = ORD == ORD_HESS;
    if (ORD == ORD_DENS) {
        P = fma(ampl * sg * sg, rho, P);
    } else {
        P = fma(ampl, phi, P);
        if (G) {
            const double ag = ampl * sg;
            for (int k = 0; k < 3; ++k) g[k] = fma(ag, eg[k], g[k]);
        }
        if (H) {
            const double ah = ampl * sg * sg;
            for (int k = 0; k < 6; ++k) h[k] = fma(ah, eh[k], h[k]);
        }
    }
}

// EXP = false: the caller guarantees an analytic type (no expansion code is
// compiled in, which keeps analytic-only kernels small in registers).
template <int ORD, bool EXP = true, int STRIDE = 0, bool SIN = true>
__device__ __forceinline__ void add_leaf(int type, const double *p, const void *table, double ampl,
                                         double sg, const double *center, double x0, double y0,
                                         double z0, double &P, double g[3], double h[6])
{
    constexpr bool G = ORD == ORD_GRAD || ORD == ORD_HESS;
    constexpr bool H = ORD == ORD_HESS;
    constexpr bool RHO = ORD == ORD_DENS;
    const double x = sg * (x0 - center[0]), y = sg * (y0 - center[1]), z = sg * (z0 - center[2]);
    Eval3 e;
    if (EXP && type == AGX_T_MULTIPOLE) {
        const MPK mp = with_law(load_desc<MPK>(table));
        PointOut o{};
        Where w = classify(mp, x, y, z);
        if (w.kind == 2)
            interior<ORD, false, NoHook, STRIDE, SIN>(mp, mp.poly + (int64_t)w.k * mp.n_lm * 6, x,
                                                      y, z, w.r2, w.r, w.ir, w.lnr, w.k, o,
                                                      1.0 / (4.0 * M_PI * G_AGAMA_D));
        else
            extrap<ORD>(mp, w, x, y, z, o, 1.0 / (4.0 * M_PI * G_AGAMA_D));
        if (RHO) e.rho = o.phi;
        else e.phi = o.phi;
        for (int k = 0; k < 3; ++k) e.g[k] = o.g[k];
        for (int k = 0; k < 6; ++k) e.h[k] = o.h[k];
    } else if (EXP && type == AGX_T_CYLSPLINE) {
        const CSK cs = load_desc<CSK>(table);
        PointOut o{};
        cylspline_point<ORD>(cs, x, y, z, o);
        if (RHO) e.rho = o.phi;
        else e.phi = o.phi;
        for (int k = 0; k < 3; ++k) e.g[k] = o.g[k];
        for (int k = 0; k < 6; ++k) e.h[k] = o.h[k];
    } else {
        eval_analytic<G, H, RHO>(type, p, x, y, z, e);
    }
    accumulate<ORD>(e.phi, e.rho, e.g, e.h, ampl, sg, P, g, h);
}

// A CylSpline leaf in two halves, for callers that defer the points outside
// its grid box: cyl_leaf_inside adds the spline part and returns true, or
// adds nothing and returns false for a point outside the box;
// cyl_leaf_outside adds the outer expansion of such a point.  grids: the
// leaf's lR and lz grids as stage_grids copies them into LDS.
template <int ORD>
__device__ __forceinline__ bool cyl_leaf_inside(const agx_component &C, double x0, double y0,
                                                double z0, double &P, double g[3], double h[6],
                                                const double *grids)
{
    const double sg = C.scale;
    const double x = sg * (x0 - C.center[0]), y = sg * (y0 - C.center[1]), z = sg * (z0 - C.center[2]);
    const CSK cs = load_desc<CSK>(C.table);
    const CylCoords cc = cyl_coords(cs, x, y, z);
    if (cyl_outside(cs, cc)) return false;
    PointOut o{};
    cylspline_inside<ORD, true>(cs, x, y, z, cc, o, grids, grids + cs.nR);
    accumulate<ORD>(o.phi, o.phi, o.g, o.h, C.ampl, sg, P, g, h);
    return true;
}

template <int ORD>
__device__ __forceinline__ void cyl_leaf_outside(const agx_component &C, double x0, double y0,
                                                 double z0, double &P, double g[3], double h[6])
{
    const double sg = C.scale;
    const double x = sg * (x0 - C.center[0]), y = sg * (y0 - C.center[1]), z = sg * (z0 - C.center[2]);
    const CSK cs = load_desc<CSK>(C.table);
    const CylCoords cc = cyl_coords(cs, x, y, z);
    PointOut o{};
    cylspline_outside<ORD>(cs, x, y, z, cc, o);
    accumulate<ORD>(o.phi, o.phi, o.g, o.h, C.ampl, sg, P, g, h);
}

}  // namespace agx
