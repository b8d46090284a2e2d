// decoder_pair.hip -- instantiations and launch of the two-codewords-per-lane CHAR_PRC decoder
// (decoder_pair.h).
#include "status.h"
#include "decoder_pair.h"

namespace nrldpc {

template <int KIND, int ZC>
int launch_decode_char_pair(const DecodeArgs& a, int blocks, int threads, size_t lds, hipStream_t s) {
    if (a.early_exit)
        hipLaunchKernelGGL((pairk::k_decode_char_pair<KIND, ZC, true>), dim3(blocks), dim3(threads), lds, s, a);
    else
        hipLaunchKernelGGL((pairk::k_decode_char_pair<KIND, ZC, false>), dim3(blocks), dim3(threads), lds, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

// Two row groups per workgroup (2 cpb Z threads): instantiated where cpb Z fills whole wavefronts.
template <int KIND, int ZC>
int launch_decode_char_split(const DecodeArgs& a, int blocks, int threads, size_t lds, hipStream_t s) {
    if (a.early_exit)
        hipLaunchKernelGGL((pairk::k_decode_char_split<KIND, ZC, true>), dim3(blocks), dim3(threads), lds, s, a);
    else
        hipLaunchKernelGGL((pairk::k_decode_char_split<KIND, ZC, false>), dim3(blocks), dim3(threads), lds, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

template int launch_decode_char_split<kBG1, 0>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_char_split<kBG1, 384>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_char_split<kBG2, 0>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_char_pair<kBG1, 0>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_char_pair<kBG1, 384>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_char_pair<kBG2, 0>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_char_pair<kBG2, 208>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_char_pair<kWifi, 0>(const DecodeArgs&, int, int, size_t, hipStream_t);
template int launch_decode_char_pair<kWifi, 96>(const DecodeArgs&, int, int, size_t, hipStream_t);

}  // namespace nrldpc
