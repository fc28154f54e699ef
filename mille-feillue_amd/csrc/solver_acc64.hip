// Persistent CG / BiCGSTAB kernels for 16-row tiles on sync-bound systems
// (one row per thread, or one workgroup) whose layouts hold low-precision
// tiles: their FP32, FP16 and E4M3 tiles accumulate in FP64 against the
// FP64 vector (instantiated here so they compile in parallel).
#include "persistent.h"

namespace mf {

KernelChoice pick_b16_acc64(int iw, int method, long long rows_per_wg, int G)
{
    return pick_tile_acc64<16>(iw, method, rows_per_wg, G);
}

}  // namespace mf
