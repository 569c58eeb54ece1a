// cg_batch_clag.hip -- batched tensor launches for LK_CLAG line operators.
#include "cg_batch_tensor.hpp"

namespace cg {
template int batch_tensor_kind<LK_CLAG>(const BatchLaunch &, const TensorArgs *, const Item *,
                                        const int *, int *, hipStream_t);
template const void *batch_tensor_fn_kind<LK_CLAG>(const BatchLaunch &, int *);
} // namespace cg
