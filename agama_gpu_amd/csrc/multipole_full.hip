// The unrolled Multipole kernels of full programs (interior() with FL > 0:
// every harmonic up to l = FL, cos and sin, as in a table without symmetry),
// in every evaluation order: the LDS-staged kernel in the workgroup sizes the
// LDS path uses, and the interval-sorted kernel with LDS slabs.  A translation unit of its own: these instantiations
// compile in parallel with the rest.
#include "multipole_sorted.h"

namespace agx {

template <int ORD, int BLOCK>
static bool full_launch(int fl, const MPK &P, unsigned grid, int64_t tb, const double *xyz,
                        int64_t n, double *phi, double *force, double *deriv, double inv4piG,
                        hipStream_t s)
{
    switch (fl) {
#define AGX_FULL_CASE(L)                                                                          \
    case L:                                                                                       \
        launch_lds<ORD, BLOCK, 1, true, 0, L>(P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s); \
        return true;
        AGX_FULL_CASE(2)
        AGX_FULL_CASE(4)
        AGX_FULL_CASE(6)
        AGX_FULL_CASE(8)
        AGX_FULL_CASE(10)
#undef AGX_FULL_CASE
    default:
        return false;
    }
}

// the workgroup sizes of the generic dispatch (multipole.hip)
template <int ORD>
static bool full_launch(int block, int fl, const MPK &P, unsigned grid, int64_t tb,
                        const double *xyz, int64_t n, double *phi, double *force, double *deriv,
                        double inv4piG, hipStream_t s)
{
    constexpr int B1 = ORD == ORD_POT ? 512 : 256;
    constexpr int B2 = ORD == ORD_POT ? 1024 : ORD == ORD_HESS ? 384 : 512;
    constexpr int B3 = ORD == ORD_POT ? 1024 : ORD == ORD_HESS ? 768 : 1024;
    if (block == B1)
        return full_launch<ORD, B1>(fl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    if (block == B2)
        return full_launch<ORD, B2>(fl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    if (block == B3)
        return full_launch<ORD, B3>(fl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    return false;
}

template <int ORD>
static bool sorted_full(int fl, const MPK &P, const SortedLaunch &L, size_t lds, hipStream_t s)
{
    switch (fl) {
#define AGX_SORTED_FULL_CASE(F)                                  \
    case F:                                                      \
        launch_sorted<ORD, true, 1, true, 0, F>(P, L, lds, s);   \
        return true;
        AGX_SORTED_FULL_CASE(2)
        AGX_SORTED_FULL_CASE(4)
        AGX_SORTED_FULL_CASE(6)
        AGX_SORTED_FULL_CASE(8)
        AGX_SORTED_FULL_CASE(10)
#undef AGX_SORTED_FULL_CASE
    default:
        return false;
    }
}

bool launch_sorted_full(int ord, int fl, const MPK &P, const SortedLaunch &L, size_t lds,
                        hipStream_t s)
{
    switch (ord) {
    case ORD_POT: return sorted_full<ORD_POT>(fl, P, L, lds, s);
    case ORD_GRAD: return sorted_full<ORD_GRAD>(fl, P, L, lds, s);
    case ORD_HESS: return sorted_full<ORD_HESS>(fl, P, L, lds, s);
    case ORD_DENS: return sorted_full<ORD_DENS>(fl, P, L, lds, s);
    default: return false;
    }
}

bool launch_lds_full(int ord, int block, int fl, const MPK &P, unsigned grid, int64_t tb,
                     const double *xyz, int64_t n, double *phi, double *force, double *deriv,
                     double inv4piG, hipStream_t s)
{
    switch (ord) {
    case ORD_POT:
        return full_launch<ORD_POT>(block, fl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    case ORD_GRAD:
        return full_launch<ORD_GRAD>(block, fl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    case ORD_HESS:
        return full_launch<ORD_HESS>(block, fl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    case ORD_DENS:
        return full_launch<ORD_DENS>(block, fl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    default:
        return false;
    }
}

}  // namespace agx
