// cg_tensor_veno.hip -- per-call tensor launches for LK_VENO line operators.
#include "cg_tensor_plain.hpp"

namespace cg {
template int launch_tensor_kind<LK_VENO>(int, int, int, const TensorArgs &, const TileCfg &,
                                          hipStream_t);
} // namespace cg
