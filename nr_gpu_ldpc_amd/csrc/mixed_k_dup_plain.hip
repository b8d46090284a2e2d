// mixed_k_dup_plain.hip -- instantiation of the persistent mixed-batch kernel, dup layout,
// plain min-sum (device code in mixed_kernel.h).
#include "mixed_kernel.h"

namespace nrldpc {

const void* mixed_kernel_dup_plain() { return reinterpret_cast<const void*>(k_decode_mixed<false>); }

}  // namespace nrldpc
