// Library-level entry points of the C ABI (version, error text).
#include "agx_common.h"

#include <mutex>

namespace agx {
int g_grid_per_cu = 32;

namespace {
struct Resident {
    int dev, block;
    const void *kernel;
    size_t lds;
    int64_t wgs;
};
constexpr int kResidentSlots = 128;
Resident g_resident[kResidentSlots];
int g_resident_n = 0;
std::mutex g_resident_mutex;
}  // namespace

int64_t resident_workgroups(const void *kernel, int block, size_t lds, int64_t fallback)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return fallback;
    std::lock_guard<std::mutex> lock(g_resident_mutex);
    for (int q = 0; q < g_resident_n; ++q) {
        const Resident &e = g_resident[q];
        if (e.dev == dev && e.kernel == kernel && e.block == block && e.lds == lds) return e.wgs;
    }
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        per_cu < 1 || cus < 1)
        return fallback;
    const int64_t wgs = (int64_t)per_cu * cus;
    if (g_resident_n < kResidentSlots) g_resident[g_resident_n++] = Resident{dev, block, kernel, lds, wgs};
    return wgs;
}
}  // namespace agx

// 1.1.0: agx_multipole.prog_flags, the program records' packed offsets and
// gamma, the prog_flags argument of agx_multipole_batch_eval
// 1.2.0: agx_component.prog_flags (formerly padding) selects the composite
// kernel's Multipole variant, so callers built against 1.1 must zero it or set
// it from the table's program; NaN positions give NaN results
// 1.3.0: agx_multipole.origin (new last field), AGX_ADJ_CENTER_GRAD
// 1.4.0: AGX_OPT_ADJ_STAGE; CylSpline near-axis / midplane and DiskAnsatz
// midplane precision
extern "C" int agx_version(void) { return 10500; }

extern "C" const char *agx_error_string(int code)
{
    switch (code) {
    case AGX_OK: return "success";
    case AGX_E_NULL: return "a required pointer argument is NULL";
    case AGX_E_SIZE: return "a size argument is out of range";
    case AGX_E_LMAX: return "lmax exceeds the supported maximum of 32";
    case AGX_E_LAUNCH: return "kernel launch failed";
    case AGX_E_WHAT: return "no output requested, or an unknown output flag";
    case AGX_E_TYPE: return "unknown potential component type";
    case AGX_E_WORKSPACE: return "workspace too small";
    case AGX_E_SCRATCH:
        return "N exceeds the static scratch of agx_multipole_eval_sorted_noalloc "
               "(AGX_NOALLOC_MAX_N points)";
    default: return "unknown error code";
    }
}
