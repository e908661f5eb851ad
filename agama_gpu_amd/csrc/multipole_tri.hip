// The unrolled LDS-staged Multipole kernels of the triaxial programs
// (interior() with TL > 0): degrees 2..12 (the potential alone 2..10), every
// evaluation order, in the workgroup sizes of tables up to 80 KB (larger ones
// run the generic kernels).  A translation unit of its own: these
// instantiations compile in parallel with multipole.hip.
#include "multipole_lds.h"

namespace agx {

template <int ORD, int BLOCK>
static bool tri_launch(int tl, const MPK &P, unsigned grid, int64_t tb, const double *xyz,
                       int64_t n, double *phi, double *force, double *deriv, double inv4piG,
                       hipStream_t s)
{
    // (the potential alone at degree 12 ran 4 % slower unrolled)
    if (ORD == ORD_POT && tl == 12) return false;
    switch (tl) {
#define AGX_TRI_CASE(L)                                                                        \
    case L:                                                                                    \
        launch_lds<ORD, BLOCK, 2, false, L>(P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s); \
        return true;
        AGX_TRI_CASE(2)
        AGX_TRI_CASE(4)
        AGX_TRI_CASE(6)
        AGX_TRI_CASE(8)
        AGX_TRI_CASE(10)
        AGX_TRI_CASE(12)
#undef AGX_TRI_CASE
    default:
        return false;
    }
}

// The workgroup sizes the LDS path uses for tables up to 80 KB (the
// generic dispatch in multipole.hip): 512 and 1024 threads for the
// potential alone, 256 and 384 for the Hessian, 256 and 512 otherwise.
template <int ORD>
static bool tri_launch(int block, int tl, const MPK &P, unsigned grid, int64_t tb,
                       const double *xyz, int64_t n, double *phi, double *force, double *deriv,
                       double inv4piG, hipStream_t s)
{
    constexpr int B1 = ORD == ORD_POT ? 512 : 256;
    constexpr int B2 = ORD == ORD_POT ? 1024 : ORD == ORD_HESS ? 384 : 512;
    if (block == B1)
        return tri_launch<ORD, B1>(tl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    if (block == B2)
        return tri_launch<ORD, B2>(tl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    return false;
}

bool launch_lds_triaxial(int ord, int block, int tl, const MPK &P, unsigned grid, int64_t tb,
                         const double *xyz, int64_t n, double *phi, double *force, double *deriv,
                         double inv4piG, hipStream_t s)
{
    switch (ord) {
    case ORD_POT:
        return tri_launch<ORD_POT>(block, tl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    case ORD_GRAD:
        return tri_launch<ORD_GRAD>(block, tl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    case ORD_HESS:
        return tri_launch<ORD_HESS>(block, tl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    case ORD_DENS:
        return tri_launch<ORD_DENS>(block, tl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    default:
        return false;
    }
}

}  // namespace agx
