// cg_tensor_vlag.hip -- per-call tensor launches for LK_VLAG line operators.
#include "cg_tensor_plain.hpp"

namespace cg {
template int launch_tensor_kind<LK_VLAG>(int, int, int, const TensorArgs &, const TileCfg &,
                                         hipStream_t);
} // namespace cg
