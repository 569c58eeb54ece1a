// cg_batch_clag_t.hip -- batched tensor launches for LK_CLAG line operators,
// 4x16-base tiles (FP64).
#include "cg_batch_tensor.hpp"

namespace cg {
template int batch_tensor_kind<LK_CLAG, 3>(const BatchLaunch &, const TensorArgs *, const Item *,
                                           const int *, int *, hipStream_t);
template const void *batch_tensor_fn_kind<LK_CLAG, 3>(const BatchLaunch &, int *);
} // namespace cg
