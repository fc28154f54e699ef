// Persistent CG / BiCGSTAB kernels for 32-row tiles (instantiated here so
// the tile sizes compile in parallel).
#include "persistent.h"

namespace mf {

KernelChoice pick_b32(int iw, int method, long long rows_per_wg, bool stream)
{
    return pick_tile<32>(iw, method, rows_per_wg, stream);
}

}  // namespace mf
