// mixed_k_slim_plain.hip -- instantiation of the persistent mixed-batch kernel, slim layout,
// plain min-sum (device code in mixed_kernel.h).
#include "mixed_kernel.h"

namespace nrldpc {

const void* mixed_kernel_slim_plain() { return reinterpret_cast<const void*>(k_decode_mixed_slim<false>); }

}  // namespace nrldpc
