// Persistent CG / BiCGSTAB kernels for 16-row tiles built for layouts whose
// tiles are all E4M3 (instantiated here so they compile in parallel).
#include "persistent.h"

namespace mf {

KernelChoice pick_b16_e4m3(int iw, int method, long long rows_per_wg)
{
    return pick_tile_cm<16, CM_E4M3>(iw, method, rows_per_wg);
}

}  // namespace mf
