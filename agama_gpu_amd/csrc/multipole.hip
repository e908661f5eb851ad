// Multipole kernels and C-ABI entry points (gfx950, FP64).
#include <mutex>
#include <type_traits>
#include "multipole_dev.h"
#include "multipole_lds.h"
#include "multipole_sorted.h"
#include "sort_dev.h"

namespace agx {

// One particle per lane, grid-stride over N.
template <int ORD>
__global__ __launch_bounds__(kBlock) void mp_lane_kernel(MPK P, const double *__restrict__ xyz,
                                                         int64_t n, double *__restrict__ phi,
                                                         double *__restrict__ force,
                                                         double *__restrict__ deriv, double inv4piG)
{
    P = with_law(P);
    for_each_point<kBlock>(xyz, n, [&](int64_t i, double x, double y, double z) {
        x -= P.ox; y -= P.oy; z -= P.oz;
        PointOut o{};
        Where w = classify(P, x, y, z);
        if (w.kind == 2)
            interior<ORD, false>(P, P.poly + (int64_t)w.k * P.n_lm * 6, x, y, z, w.r2, w.r, w.ir,
                                 w.lnr, w.k, o, inv4piG);
        else
            extrap<ORD>(P, w, x, y, z, o, inv4piG);
        store<ORD>(o, i, phi, force, deriv);
    });
}

template <int ORD>
__global__ __launch_bounds__(kBinThreads) void mp_binned_kernel(
    MPK P, const double *__restrict__ xyz, int64_t n, int64_t tile, double *__restrict__ phi,
    double *__restrict__ force, double *__restrict__ deriv, double inv4piG)
{
    P = with_law(P);
    binned_tiles(P, xyz, n, tile, [&](int b, bool valid, int64_t i) {
        if (!valid) return;
        double x, y, z;
        load_pos(P, xyz, i, x, y, z);
        PointOut o{};
        if (b > 0) {
            const int k = b - 1;
            const double r2 = fma(x, x, fma(y, y, z * z));
            double r, ir;
            sqrt_rsqrt(r2, r, ir);      // r >= r_min: r2 is far above 2^-767
            interior<ORD, true>(P, P.poly + (int64_t)k * P.n_lm * 6, x, y, z, r2, r, ir, agx_log(r),
                                k, o, inv4piG);
        } else {
            Where w = classify(P, x, y, z);
            extrap<ORD>(P, w, x, y, z, o, inv4piG);
        }
        store<ORD>(o, i, phi, force, deriv);
    });
}

// The descriptor and N; the caller checks the position and output pointers
// after its own flags, so that no invalid argument can reach a launch.
static int check_mp(const agx_multipole *mp, int64_t n)
{
    if (!mp) return AGX_E_NULL;
    if (n < 0) return AGX_E_SIZE;
    if (!mp->poly_km || !mp->program) return AGX_E_NULL;
    if (mp->n_int < 1 || mp->n_lm < 1 || mp->n_groups < 0 || mp->n_steps < 0) return AGX_E_SIZE;
    if (mp->lmax > 32) return AGX_E_LMAX;
    if (mp->prog_flags & ~kProgFlags) return AGX_E_WHAT;
    // a triaxial degree must agree with the program's size (its column
    // offsets then stay inside the table's slabs)
    const int tl = (mp->prog_flags & AGX_PROG_TRIAXIAL_MASK) >> AGX_PROG_TRIAXIAL_SHIFT;
    if (tl && ((tl & 1) || mp->n_groups != tl / 2 + 1 || mp->n_steps != tri_step0(tl, tl / 2 + 1) ||
               mp->n_lm != mp->n_steps))
        return AGX_E_WHAT;
    // likewise a full degree (and not both)
    const int fl = (mp->prog_flags & AGX_PROG_FULL_MASK) >> AGX_PROG_FULL_SHIFT;
    if (fl && (tl || mp->n_groups != fl + 1 || mp->n_steps != full_step0(fl, fl + 1) ||
               mp->n_lm != (fl + 1) * (fl + 1)))
        return AGX_E_WHAT;
    return AGX_OK;
}

// Tunables (agx_set_option): evaluation path and binned tile size.
static int g_mp_path = 0;            // 0 auto, 1 lane, 2 binned, 3 LDS-staged lane
static int64_t g_mp_tile = 0;        // 0 = automatic
static int64_t g_mp_bin_min_n = 200000;
static int64_t g_mp_bin_min_terms = 24;
static int g_mp_lds_grid = 4;        // LDS path: grid = this x resident workgroups
static int g_mp_bin_grid = 2;        // binned path: grid = this x resident workgroups
static int g_mp_triaxial = 1;        // LDS path: unrolled kernels for triaxial programs

struct BinLaunch {
    unsigned grid;
    int64_t tile;
};

static BinLaunch bin_launch(int64_t n)
{
    // three 42 KB-LDS workgroups fit per CU; g_mp_bin_grid x that many
    // workgroups in the grid
    const int64_t max_wg = 256 * 3 * (int64_t)g_mp_bin_grid;
    int64_t tile = g_mp_tile > 0 ? g_mp_tile : (n + max_wg - 1) / max_wg;
    if (tile < 1024) tile = 1024;
    if (tile > kMaxTile) tile = kMaxTile;
    int64_t n_tiles = (n + tile - 1) / tile;
    int64_t grid = n_tiles < max_wg ? n_tiles : max_wg;
    // equal tile counts per workgroup: rebalance the tile size
    int64_t per = (n_tiles + grid - 1) / grid;
    tile = (n + grid * per - 1) / (grid * per);
    if (tile < 1) tile = 1;
    return BinLaunch{(unsigned)grid, tile};
}

// Binning pays when the per-term table reads dominate (many active terms) and
// N is large enough to fill the chip with tiles; with few terms its sort
// costs more than the reads it removes.
static bool binned_auto(const agx_multipole *mp, int64_t n)
{
    return mp->n_int + 1 <= kMaxBins && n >= g_mp_bin_min_n && mp->n_lm >= g_mp_bin_min_terms;
}

// LDS bytes of the staged table (slabs padded: lds_slab_words)
static int64_t table_bytes(const agx_multipole *mp)
{
    return (int64_t)mp->n_int * lds_slab_words(mp->n_lm) * 16;
}

static unsigned lds_grid(int64_t n, int block, int64_t per_cu)
{
    int64_t want = (n + block - 1) / block, cap = 256 * per_cu;
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (unsigned)want;
}

template <int ORD, int BLOCK>
static void launch_lds(const agx_multipole *mp, const MPK &P, unsigned grid, int64_t tb,
                       const double *xyz, int64_t n, double *phi, double *force, double *deriv,
                       double inv4piG, hipStream_t s)
{
    const int st = prog_stride(mp->prog_flags);
    const bool co = mp->prog_flags & AGX_PROG_COS_ONLY;
    const int tl = (mp->prog_flags & AGX_PROG_TRIAXIAL_MASK) >> AGX_PROG_TRIAXIAL_SHIFT;
    if (tl && g_mp_triaxial &&
        launch_lds_triaxial(ORD, BLOCK, tl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s))
        return;
    const int fl = (mp->prog_flags & AGX_PROG_FULL_MASK) >> AGX_PROG_FULL_SHIFT;
    if (fl && g_mp_triaxial &&
        launch_lds_full(ORD, BLOCK, fl, P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s))
        return;
    if (st == 1 && co)
        launch_lds<ORD, BLOCK, 1, false>(P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    else if (st == 1)
        launch_lds<ORD, BLOCK, 1, true>(P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    else if (st == 2 && co)
        launch_lds<ORD, BLOCK, 2, false>(P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    else if (st == 2)
        launch_lds<ORD, BLOCK, 2, true>(P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    else if (co)
        launch_lds<ORD, BLOCK, 0, false>(P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
    else
        launch_lds<ORD, BLOCK, 0, true>(P, grid, tb, xyz, n, phi, force, deriv, inv4piG, s);
}

template <int ORD>
static void launch_mp(const agx_multipole *mp, const MPK &P, const double *xyz, int64_t n,
                      double *phi, double *force, double *deriv, double inv4piG, hipStream_t s)
{
    const int64_t tb = table_bytes(mp);
    const bool lds_ok = tb <= kLdsTableBytes;
    int path = g_mp_path;
    // automatic: the LDS-staged table whenever it fits (and N is not tiny),
    // else binning for many terms and large N, else per-lane table reads
    if (path == 0 || path == 4) path = (lds_ok && n >= 4096) ? 3 : (binned_auto(mp, n) ? 2 : 1);
    if (path == 3 && !lds_ok) path = 1;
    if (path == 2 && mp->n_int + 1 > kMaxBins) path = 1;
    if (path == 3) {
        // Workgroups per CU that fit in LDS (at most 8) and a workgroup size
        // that gives them the wavefronts the kernel's registers allow
        // (lds_block): 4 per SIMD for pot+force and density, 8 for the
        // potential alone, 3 for the Hessian.
        int64_t fit = kLdsPerCu / tb;
        if (fit > 8) fit = 8;
        // The grid is g_mp_lds_grid times the workgroups that fit at once:
        // static partitions over exactly the resident workgroups end unevenly
        // (the waves sharing a SIMD do not progress at the same rate: the
        // first workgroups finish 35 % before the last), so the grid is
        // oversubscribed and the dispatcher hands freed slots to new
        // workgroups; each one re-stages the table from L2.
        const int64_t per_cu = fit * g_mp_lds_grid;
        auto go = [&](auto block) {
            constexpr int B = decltype(block)::value;
            launch_lds<ORD, B>(mp, P, lds_grid(n, B, per_cu), tb, xyz, n, phi, force, deriv,
                               inv4piG, s);
        };
        using std::integral_constant;
        if constexpr (ORD == ORD_POT) {
            // at most 64 VGPRs: 8 waves per SIMD, 32 wavefronts per CU
            if (fit >= 4) go(integral_constant<int, 512>{});
            else go(integral_constant<int, 1024>{});
        } else if constexpr (ORD == ORD_HESS) {
            // 143-168 VGPRs: 3 waves per SIMD, 12 wavefronts per CU (a
            // 1024-thread workgroup would cap it at 128 VGPRs: spills)
            if (fit >= 3) go(integral_constant<int, 256>{});
            else if (fit == 2) go(integral_constant<int, 384>{});
            else go(integral_constant<int, 768>{});
        } else {
            // 4 waves per SIMD, 16 wavefronts per CU
            if (fit >= 4) go(integral_constant<int, 256>{});
            else if (fit >= 2) go(integral_constant<int, 512>{});
            else go(integral_constant<int, 1024>{});
        }
    } else if (path == 2) {
        BinLaunch bl = bin_launch(n);
        hipLaunchKernelGGL(mp_binned_kernel<ORD>, dim3(bl.grid), dim3(kBinThreads), 0, s, P, xyz,
                           n, bl.tile, phi, force, deriv, inv4piG);
    } else {
        hipLaunchKernelGGL(mp_lane_kernel<ORD>, dim3(grid_for(n)), dim3(kBlock), 0, s, P, xyz, n,
                           phi, force, deriv, inv4piG);
    }
}

}  // namespace agx

using namespace agx;

extern "C" int agx_set_option(int key, int64_t value)
{
    switch (key) {
    case AGX_OPT_MP_PATH:
        if (value < 0 || value > 4) return AGX_E_SIZE;
        g_mp_path = (int)value;
        return AGX_OK;
    case AGX_OPT_MP_TILE:
        if (value < 0 || value > kMaxTile) return AGX_E_SIZE;
        g_mp_tile = value;
        return AGX_OK;
    case AGX_OPT_MP_BIN_MIN_N:
        if (value < 0) return AGX_E_SIZE;
        g_mp_bin_min_n = value;
        return AGX_OK;
    case AGX_OPT_MP_BIN_MIN_TERMS:
        if (value < 0) return AGX_E_SIZE;
        g_mp_bin_min_terms = value;
        return AGX_OK;
    case AGX_OPT_GRID_PER_CU:
        if (value < 1 || value > 64) return AGX_E_SIZE;
        g_grid_per_cu = (int)value;
        return AGX_OK;
    case AGX_OPT_MP_BIN_GRID:
        if (value < 1 || value > 16) return AGX_E_SIZE;
        g_mp_bin_grid = (int)value;
        return AGX_OK;
    case AGX_OPT_MP_LDS_GRID:
        if (value < 1 || value > 64) return AGX_E_SIZE;
        g_mp_lds_grid = (int)value;
        return AGX_OK;
    case AGX_OPT_SORT_GRID:
        if (value < 0 || value > 4096) return AGX_E_SIZE;
        g_sort_grid = (int)value;
        return AGX_OK;
    case AGX_OPT_CYL_PATH:
        return cyl_set_path(value);
    case AGX_OPT_ADJ_STAGE:
        return adj_set_stage(value);
    case AGX_OPT_MP_TRIAXIAL:
        if (value < 0 || value > 1) return AGX_E_SIZE;
        g_mp_triaxial = (int)value;
        return AGX_OK;
    default:
        return AGX_E_WHAT;
    }
}

extern "C" int agx_multipole_eval(const agx_multipole *mp, const double *d_xyz, int64_t n,
                                  double *d_phi, double *d_force, double *d_deriv, uint32_t what,
                                  void *stream)
{
    int rc = check_mp(mp, n);
    if (rc) return rc;
    if (what == 0 || (what & ~(AGX_PHI | AGX_FORCE | AGX_DERIV))) return AGX_E_WHAT;
    if (n == 0) return AGX_OK;   // empty input: nothing to write, outputs may be NULL
    if (!d_xyz || ((what & AGX_PHI) && !d_phi) || ((what & AGX_FORCE) && !d_force) ||
        ((what & AGX_DERIV) && !d_deriv))
        return AGX_E_NULL;
    MPK P = make_args(mp);
    hipStream_t s = (hipStream_t)stream;
    double *phi = (what & AGX_PHI) ? d_phi : nullptr;
    double *force = (what & AGX_FORCE) ? d_force : nullptr;
    double *deriv = (what & AGX_DERIV) ? d_deriv : nullptr;
    if (deriv)
        launch_mp<ORD_HESS>(mp, P, d_xyz, n, phi, force, deriv, 0.0, s);
    else if (force)
        launch_mp<ORD_GRAD>(mp, P, d_xyz, n, phi, force, nullptr, 0.0, s);
    else
        launch_mp<ORD_POT>(mp, P, d_xyz, n, phi, nullptr, nullptr, 0.0, s);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}

extern "C" int64_t agx_multipole_work_size(int64_t n)
{
    return n < 0 ? 0 : sort_work_size(n, kMaxBins);
}

// LDS coefficient slabs of the sorted kernel while a workgroup's four take at
// most this many bytes (slabs up to 10 KB: n_lm <= 213, lmax <= 13).
// The sorted kernel's variant: LDS slabs for short slabs, with the program
// flags' stride and sin variants; scalar coefficient loads and the general
// program otherwise.
template <int ORD>
static void launch_sorted(const agx_multipole *mp, const MPK &P, const SortedLaunch &L,
                          hipStream_t s)
{
    const int64_t lds = (int64_t)(kBlock / 64) * mp->n_lm * 6 * (int64_t)sizeof(double);
    if (lds > kSortedSlabLds) {
        launch_sorted<ORD, false, 0, true>(P, L, 0, s);
        return;
    }
    const int st = prog_stride(mp->prog_flags);
    const bool co = mp->prog_flags & AGX_PROG_COS_ONLY;
    const int fl = (mp->prog_flags & AGX_PROG_FULL_MASK) >> AGX_PROG_FULL_SHIFT;
    if (fl && g_mp_triaxial && launch_sorted_full(ORD, fl, P, L, (size_t)lds, s)) return;
    if (st == 1 && !co)
        launch_sorted<ORD, true, 1, true>(P, L, (size_t)lds, s);
    else if (st == 2 && co)
        launch_sorted<ORD, true, 2, false>(P, L, (size_t)lds, s);
    else
        launch_sorted<ORD, true, 0, true>(P, L, (size_t)lds, s);
}

// The interval-sorted evaluation with its workspace at d_work (at least
// sort_work_size(n, kMaxBins) bytes); arguments already validated.  Every
// workspace word that the kernels read is written by them first (the sort's
// init kernel clears the histogram), so no clearing is needed.
static int run_sorted(const agx_multipole *mp, const double *d_xyz, int64_t n, double *d_phi,
                      double *d_force, double *d_deriv, uint32_t what, void *d_work,
                      hipStream_t s)
{
    MPK P = make_args(mp);
    const SortWork W = sort_work(d_work, n, kMaxBins);
    const int nb = mp->n_int + 1;
    sort_launch(MpBinKey{P, d_xyz}, n, nb, W, s);
    double *phi = (what & AGX_PHI) ? d_phi : nullptr;
    double *force = (what & AGX_FORCE) ? d_force : nullptr;
    double *deriv = (what & AGX_DERIV) ? d_deriv : nullptr;
    const SortedLaunch L{n, d_xyz, phi, force, deriv, W};
    if (deriv)
        launch_sorted<ORD_HESS>(mp, P, L, s);
    else if (force)
        launch_sorted<ORD_GRAD>(mp, P, L, s);
    else
        launch_sorted<ORD_POT>(mp, P, L, s);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}

// Validation shared by the two sorted entry points: the descriptor, `what`
// and N, then (for N > 0) the pointers.
static int check_sorted(const agx_multipole *mp, int64_t n, uint32_t what)
{
    int rc = check_mp(mp, n);
    if (rc) return rc;
    if (what == 0 || (what & ~(AGX_PHI | AGX_FORCE | AGX_DERIV))) return AGX_E_WHAT;
    if (n > (int64_t)0x7FFFFFFF - 64 * kMaxBins || mp->n_int + 1 > kMaxBins) return AGX_E_SIZE;
    return AGX_OK;
}

static int check_sorted_ptrs(const double *d_xyz, double *d_phi, double *d_force,
                             double *d_deriv, uint32_t what)
{
    if (!d_xyz || ((what & AGX_PHI) && !d_phi) || ((what & AGX_FORCE) && !d_force) ||
        ((what & AGX_DERIV) && !d_deriv))
        return AGX_E_NULL;
    return AGX_OK;
}

// Device scratch of agx_multipole_eval_sorted_noalloc: one copy per GPU,
// resident for the life of the process.  Host code never uses this name as an
// address (on the host it is only a stand-in for the symbol); the device
// address of each GPU's copy comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[AGX_NOALLOC_SCRATCH_BYTES];
static_assert(sort_work_size(AGX_NOALLOC_MAX_N, kMaxBins) <= AGX_NOALLOC_SCRATCH_BYTES,
              "AGX_NOALLOC_MAX_N points must fit the static scratch");

constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mutex;

// This device's copy of `scratch`, or NULL if its address cannot be had.
static void *scratch_for_current_device()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

extern "C" int agx_multipole_eval_sorted(const agx_multipole *mp, const double *d_xyz, int64_t n,
                                         double *d_phi, double *d_force, double *d_deriv,
                                         uint32_t what, void *d_work, int64_t work_bytes,
                                         void *stream)
{
    int rc = check_sorted(mp, n, what);
    if (rc || n == 0) return rc;
    if (work_bytes < sort_work_size(n, kMaxBins)) return AGX_E_WORKSPACE;
    rc = check_sorted_ptrs(d_xyz, d_phi, d_force, d_deriv, what);
    if (rc) return rc;
    if (!d_work) return AGX_E_NULL;
    return run_sorted(mp, d_xyz, n, d_phi, d_force, d_deriv, what, d_work, (hipStream_t)stream);
}

extern "C" int agx_multipole_eval_sorted_noalloc(const agx_multipole *mp, const double *d_xyz,
                                                 int64_t n, double *d_phi, double *d_force,
                                                 double *d_deriv, uint32_t what, void *stream)
{
    int rc = check_sorted(mp, n, what);
    if (rc || n == 0) return rc;
    if (n > AGX_NOALLOC_MAX_N || sort_work_size(n, kMaxBins) > AGX_NOALLOC_SCRATCH_BYTES)
        return AGX_E_SCRATCH;
    rc = check_sorted_ptrs(d_xyz, d_phi, d_force, d_deriv, what);
    if (rc) return rc;
    void *work = scratch_for_current_device();
    if (!work) return AGX_E_LAUNCH;
    return run_sorted(mp, d_xyz, n, d_phi, d_force, d_deriv, what, work, (hipStream_t)stream);
}

extern "C" int agx_multipole_density(const agx_multipole *mp, const double *d_xyz, int64_t n,
                                     double inv_4piG, double *d_rho, void *stream)
{
    int rc = check_mp(mp, n);
    if (rc) return rc;
    if (n == 0) return AGX_OK;
    if (!d_xyz || !d_rho) return AGX_E_NULL;
    MPK P = make_args(mp);
    launch_mp<ORD_DENS>(mp, P, d_xyz, n, d_rho, nullptr, nullptr, inv_4piG, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}
