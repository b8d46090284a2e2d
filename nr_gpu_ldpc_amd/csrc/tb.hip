// tb.hip -- device side of transport blocks: de-rate-matching and code-block CRC checks.
#include <hip/hip_runtime.h>

#include "status.h"
#include "tb.h"

namespace nrldpc {

// Known-zero filler bits enter the decoder with this LLR (any value above the check-node
// start value 1000 is never a minimum; finite so that binary16 storage also holds it).
constexpr float kFillerLlr = 10000.0f;

// One thread per decoder column of every code block: the sum, in bit-selection order, of the
// received LLRs that carry it (repetitions combine; punctured and unsent columns stay 0).
__global__ void __launch_bounds__(256) k_derate(DerateArgs a) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.n) return;
    const int c = n % a.n_vn;
    float v = 0.0f;
    if (c >= a.f0 && c < a.f1) {
        v = kFillerLlr;
    } else {
        for (int k = a.start[n]; k < a.start[n + 1]; k++) v += a.in[a.index[k]];
    }
    a.out[n] = v;
}

int derate_launch(const DerateArgs& a, hipStream_t s) {
    if (a.n <= 0) return LDPC_OK;
    hipLaunchKernelGGL(k_derate, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

// One workgroup per code block: lane 0 checks the CRC24B (register of TS 38.212 5.1, bit 0
// first); all lanes copy the block's data bits into the transport block.
__global__ void __launch_bounds__(256) k_tb_check(TbCheckArgs a) {
    const int r = blockIdx.x;
    const uint8_t* h = a.hd + (size_t)r * a.hd_bytes;
    const int data = a.K_prime - a.L_cb;
    if (a.tb_bits) {
        for (int k = threadIdx.x; k < data; k += blockDim.x) {
            const long long dst = (long long)r * data + k;
            if (dst < a.B) a.tb_bits[dst] = (h[k >> 3] >> (k & 7)) & 1u;
        }
    }
    if (threadIdx.x == 0) {
        int ok = 1;
        if (a.L_cb == 24) {
            uint32_t reg = 0;
            for (int k = 0; k < a.K_prime; k++) {
                reg ^= ((h[k >> 3] >> (k & 7)) & 1u) << 23;
                reg = (reg & 0x800000u) ? ((reg << 1) ^ 0x800063u) : (reg << 1);
                reg &= 0xffffffu;
            }
            ok = reg == 0u;
        }
        a.cb_ok[r] = ok;
    }
}

int tb_check_launch(const TbCheckArgs& a, hipStream_t s) {
    if (a.C <= 0) return LDPC_OK;
    hipLaunchKernelGGL(k_tb_check, dim3(a.C), dim3(256), 0, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

}  // namespace nrldpc
