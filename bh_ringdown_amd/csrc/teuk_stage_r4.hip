// Stage kernels for 4-row theta chunks (teuk_stage.inc).
#include "teuk_stage.inc"

namespace teuk {
template hipError_t launch_stage_rows<4>(const Grid &, const DevState &, int, int, bool, int, int, dim3, dim3, int,
                                         hipStream_t);
template hipError_t launch_stage_batch_rows<4>(const BatchTile *, int, int, int, int, dim3, int, hipStream_t);
}  // namespace teuk
