// Persistent CG / BiCGSTAB kernels for 16-row tiles built for layouts without
// FP64 tiles (FP32, FP16 and E4M3 only; instantiated here so they compile in
// parallel).
#include "persistent.h"

namespace mf {

KernelChoice pick_b16_low(int iw, int method, long long rows_per_wg)
{
    return pick_tile_cm<16, CM_LOW>(iw, method, rows_per_wg);
}

}  // namespace mf
