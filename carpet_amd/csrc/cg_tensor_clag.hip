// cg_tensor_clag.hip -- per-call tensor launches for LK_CLAG line operators.
#include "cg_tensor_plain.hpp"

namespace cg {
template int launch_tensor_kind<LK_CLAG>(int, int, int, const TensorArgs &, const TileCfg &,
                                         hipStream_t);
} // namespace cg
