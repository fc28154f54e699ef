// Persistent CG / BiCGSTAB kernels for 8-row tiles (instantiated here so
// the tile sizes compile in parallel).
#include "persistent.h"

namespace mf {

KernelChoice pick_b8(int iw, int method, long long rows_per_wg, bool stream)
{
    return pick_tile<8>(iw, method, rows_per_wg, stream);
}

}  // namespace mf
