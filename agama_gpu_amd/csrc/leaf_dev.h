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
    if (RHO) {
        P = fma(ampl * sg * sg, e.rho, P);
    } else {
        P = fma(ampl, e.phi, P);
        if (G) {
            const double ag = ampl * sg;
            for (int k = 0; k < 3; ++k) g[k] = fma(ag, e.g[k], g[k]);
        }
        if (H) {
            const double ah = ampl * sg * sg;
            for (int k = 0; k < 6; ++k) h[k] = fma(ah, e.h[k], h[k]);
        }
    }
}

}  // namespace agx
