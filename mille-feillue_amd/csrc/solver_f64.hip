// Persistent CG / BiCGSTAB kernels for 16-row tiles built for layouts of
// FP64 tiles only (instantiated here so they compile in parallel).
#include "persistent.h"

namespace mf {

KernelChoice pick_b16_f64(int iw, int method, long long rows_per_wg)
{
    return pick_tile_cm<16, CM_F64>(iw, method, rows_per_wg);
}

}  // namespace mf
