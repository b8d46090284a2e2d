// mixed_k_dup_learned.hip -- instantiation of the persistent mixed-batch kernel, dup layout,
// learned weights where set (with the plain fallback) (device code in mixed_kernel.h).
#include "mixed_kernel.h"

namespace nrldpc {

const void* mixed_kernel_dup_learned() { return reinterpret_cast<const void*>(k_decode_mixed<true>); }

}  // namespace nrldpc
