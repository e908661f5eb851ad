// Library-level entry points of the C ABI (version, error text).
#include "agx_common.h"

namespace agx {
int g_grid_per_cu = 32;
}

// 1.1.0: agx_multipole.prog_flags, the program records' packed offsets and
// gamma, the prog_flags argument of agx_multipole_batch_eval
// 1.2.0: agx_component.prog_flags (formerly padding) selects the composite
// kernel's Multipole variant, so callers built against 1.1 must zero it or set
// it from the table's program; NaN positions give NaN results
extern "C" int agx_version(void) { return 10200; }

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
