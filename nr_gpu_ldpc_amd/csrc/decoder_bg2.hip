// decoder_bg2.hip -- instantiations of the fused decoder for one code kind.
#include "status.h"
#include "decoder_kernel.h"

namespace nrldpc {

template <int KIND, int ZC, bool LEARNED>
int launch_decode(const DecodeArgs& a, int blocks, int threads, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_decode<KIND, ZC, LEARNED>), dim3(blocks), dim3(threads), lds, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

template <int KIND, int ZC, bool LEARNED>
int occupancy_decode(int threads, size_t lds) {
    int n = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_decode<KIND, ZC, LEARNED>, threads, lds) == hipSuccess ? n : 0;
}

template <int KIND, int ZC>
int launch_decode_half(const DecodeArgs& a, int blocks, int threads, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_decode<KIND, ZC, false, false, true>), dim3(blocks), dim3(threads), lds, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

template <int KIND>
int launch_decode_train(const DecodeArgs& a, int blocks, int threads, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_decode<KIND, 0, true, true>), dim3(blocks), dim3(threads), lds, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

template int launch_decode<kBG2, 0, false>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode<kBG2, 0, true>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode<kBG2, 208, false>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode<kBG2, 208, true>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int occupancy_decode<kBG2, 0, false>(int, size_t);
template int occupancy_decode<kBG2, 0, true>(int, size_t);
template int occupancy_decode<kBG2, 208, false>(int, size_t);
template int occupancy_decode<kBG2, 208, true>(int, size_t);

template int launch_decode_train<kBG2>(const DecodeArgs&, int, int, size_t, hipStream_t);

template int launch_decode_half<kBG2, 0>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_half<kBG2, 208>(const DecodeArgs&, int, int, size_t, hipStream_t);

}  // namespace nrldpc
