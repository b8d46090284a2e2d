// decoder_slim_qc.hip -- instantiations of the slim-layout decoder (k_decode_slim) for one code kind:
// the lifting size with compiled-in offsets and the any-Z kernel, plain and learned.
#include "status.h"
#include "decoder_kernel.h"

namespace nrldpc {

template <int KIND, int ZC, bool LEARNED>
int launch_decode_slim(const DecodeArgs& a, int grid, int threads, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_decode_slim<KIND, ZC, LEARNED>), dim3(grid), dim3(threads), lds, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

template int launch_decode_slim<kWifi, 0, false>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_slim<kWifi, 0, true>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_slim<kWifi, 96, false>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_slim<kWifi, 96, true>(const DecodeArgs&, int, int, size_t, hipStream_t);

}  // namespace nrldpc
