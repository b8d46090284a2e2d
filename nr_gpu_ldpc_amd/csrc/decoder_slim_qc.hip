// decoder_slim_qc.hip -- instantiations of the slim-layout decoder (k_decode_slim) for one code kind:
// the lifting size with compiled-in offsets and the any-Z kernel, plain, learned and the
// binary16 (HALF_PRC) schedule.
#include "status.h"
#include "decoder_kernel.h"

namespace nrldpc {

template <int KIND, int ZC, bool LEARNED, bool HALF>
int launch_decode_slim(const DecodeArgs& a, int grid, int threads, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_decode_slim<KIND, ZC, LEARNED, HALF>), dim3(grid), dim3(threads), lds, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

template int launch_decode_slim<kWifi, 0, false, false>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_slim<kWifi, 0, true, false>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_slim<kWifi, 96, false, false>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_slim<kWifi, 96, true, false>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_slim<kWifi, 0, false, true>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_slim<kWifi, 96, false, true>(const DecodeArgs&, int, int, size_t, hipStream_t);

}  // namespace nrldpc
