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

// One workgroup per code block: all lanes stage the block's first K' decisions in LDS and copy
// its data bits into the transport block; lane 0 then runs the CRC24B register (TS 38.212 5.1,
// bit 0 first) a byte at a time through a 256-entry table in LDS: for the message bits
// m0..m7 of one byte, the register becomes (reg << 8) ^ T[(reg >> 16) ^ (m0 m1 ... m7 as an
// MSB-first byte)], T[i] being the register after 8 zero-input steps from i << 16.
constexpr int kTbMaxBytes = (8448 + 7) / 8;

__global__ void __launch_bounds__(256) k_tb_check(TbCheckArgs a) {
    __shared__ uint32_t tab[256];
    __shared__ uint8_t hb[kTbMaxBytes];
    const int r = blockIdx.x;
    const uint8_t* h = a.hd + (size_t)r * a.hd_bytes;
    const int data = a.K_prime - a.L_cb;
    if (a.tb_bits) {
        for (int k = threadIdx.x; k < data; k += blockDim.x) {
            const long long dst = (long long)r * data + k;
            if (dst < a.B) a.tb_bits[dst] = (h[k >> 3] >> (k & 7)) & 1u;
        }
    }
    if (a.L_cb != 24) {
        if (threadIdx.x == 0) a.cb_ok[r] = 1;
        return;
    }
    const int nb = (a.K_prime + 7) / 8;  // K' <= 8448: at most kTbMaxBytes
    for (int i = threadIdx.x; i < 256; i += blockDim.x) {
        uint32_t reg = (uint32_t)i << 16;
        for (int b = 0; b < 8; b++) reg = (reg & 0x800000u) ? ((reg << 1) ^ 0x800063u) : (reg << 1);
        tab[i] = reg & 0xffffffu;
    }
    for (int n = threadIdx.x; n < nb; n += blockDim.x) hb[n] = h[n];
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t reg = 0;
        const int full = a.K_prime / 8;
        for (int n = 0; n < full; n++) {
            const uint32_t v = __builtin_bitreverse32((uint32_t)hb[n]) >> 24;
            reg = ((reg << 8) ^ tab[((reg >> 16) ^ v) & 0xffu]) & 0xffffffu;
        }
        for (int k = full * 8; k < a.K_prime; k++) {
            reg ^= ((hb[k >> 3] >> (k & 7)) & 1u) << 23;
            reg = (reg & 0x800000u) ? ((reg << 1) ^ 0x800063u) : (reg << 1);
            reg &= 0xffffffu;
        }
        a.cb_ok[r] = reg == 0u;
    }
}

int tb_check_launch(const TbCheckArgs& a, hipStream_t s) {
    if (a.C <= 0) return LDPC_OK;
    if (a.K_prime > 8448 || a.K_prime < 0) {
        set_last_error("code block longer than 8448 bits");
        return LDPC_EINVAL;
    }
    hipLaunchKernelGGL(k_tb_check, dim3(a.C), dim3(256), 0, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

}  // namespace nrldpc
