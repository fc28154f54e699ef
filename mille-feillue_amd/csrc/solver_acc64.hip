// Persistent CG / BiCGSTAB kernels for 16-row tiles, 1 or 2 rows per thread,
// on layouts without FP64 tiles, whose FP32, FP16 and E4M3 tiles accumulate
// in FP64 against the FP64 vector (instantiated here so they compile in
// parallel).
#include "persistent.h"

namespace mf {

KernelChoice pick_b16_acc64(int iw, int method, long long rows_per_wg)
{
    return pick_tile_acc64<16>(iw, method, rows_per_wg);
}

}  // namespace mf
