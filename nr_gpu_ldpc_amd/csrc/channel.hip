// channel.hip -- on-device codeword generation, AWGN channel, syndrome and error counting.
//
// Random numbers come from Philox4x32-10 (Salmon et al., SC'11) keyed by a 64-bit seed and
// counted by (codeword stream, block index) with a domain tag that separates information
// bits from channel noise.  Gaussian samples use Box-Muller with a logarithm and sine/cosine
// built from basic IEEE operations in a fixed order (no contraction), so every sample is
// reproducible bit for bit on any IEEE-754 host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "channel.h"
#include "status.h"

namespace nrldpc {

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    uint32_t c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    c[0] = c0; c[1] = c1; c[2] = c2; c[3] = c3;
}

constexpr uint32_t kTagBits = 0x42495453u;
constexpr uint32_t kTagNoise = 0x4E4F4953u;

__device__ __forceinline__ void philox_block(uint64_t seed, uint64_t stream, uint64_t block, uint32_t tag,
                                             uint32_t o[4]) {
    o[0] = (uint32_t)block;
    o[1] = (uint32_t)(block >> 32);
    o[2] = (uint32_t)stream;
    o[3] = (uint32_t)(stream >> 32) ^ tag;
    philox4x32_10(o, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// log(u) for u in (0, 1]: exponent split, then 2*atanh(s) with s = (m-1)/(m+1), |s| <= 0.172.
__device__ __forceinline__ float nr_log(float u) {
    uint32_t b = __float_as_uint(u);
    int e = (int)((b >> 23) & 0xff) - 127;
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    const float s = __fdiv_rn(m - 1.0f, m + 1.0f);
    const float s2 = s * s;
    float p = 0.0909090909f;
    p = p * s2 + 0.111111111f;
    p = p * s2 + 0.142857143f;
    p = p * s2 + 0.2f;
    p = p * s2 + 0.333333333f;
    p = p * s2 + 1.0f;
    const float lm = 2.0f * s * p;
    return lm + (float)e * 0.693147181f;
}

// sin / cos of 2*pi*u, u in [0, 1): quadrant reduction, Taylor polynomials on [0, pi/2).
__device__ __forceinline__ void nr_sincos_2pi(float u, float* sn, float* cs) {
    const float v = u * 4.0f;
    const int q = (int)v;
    const float f = v - (float)q;
    const float x = f * 1.57079633f;
    const float x2 = x * x;
    float s = -2.50521084e-8f;
    s = s * x2 + 2.75573192e-6f;
    s = s * x2 - 1.98412698e-4f;
    s = s * x2 + 8.33333333e-3f;
    s = s * x2 - 1.66666667e-1f;
    s = s * x2 + 1.0f;
    s = s * x;
    float c = 2.08767570e-9f;
    c = c * x2 - 2.75573192e-7f;
    c = c * x2 + 2.48015873e-5f;
    c = c * x2 - 1.38888889e-3f;
    c = c * x2 + 4.16666667e-2f;
    c = c * x2 - 0.5f;
    c = c * x2 + 1.0f;
    switch (q & 3) {
        case 0: *sn = s; *cs = c; break;
        case 1: *sn = c; *cs = -s; break;
        case 2: *sn = -s; *cs = -c; break;
        default: *sn = -c; *cs = s; break;
    }
}

// Correctly rounded square root of a positive normal float: the hardware estimate is within
// one ulp, so test its two neighbours with exact fma residuals and step to the right one.
__device__ __forceinline__ float cr_sqrt(float x) {
    float s = __builtin_amdgcn_sqrtf(x);
    const uint32_t si = __float_as_uint(s);
    const float dn = __uint_as_float(si - 1u), up = __uint_as_float(si + 1u);
    const float rdn = fmaf(-dn, s, x);
    const float rup = fmaf(-up, s, x);
    s = (rdn <= 0.0f) ? dn : s;
    s = (rup > 0.0f) ? up : s;
    return s;
}

// Four standard normal samples 4*block .. 4*block+3 of a codeword's noise stream.
__device__ __forceinline__ void gauss4(uint64_t seed, uint64_t stream, uint64_t block, float g[4]) {
    uint32_t o[4];
    philox_block(seed, stream, block, kTagNoise, o);
#pragma unroll
    for (int pr = 0; pr < 2; pr++) {
        const float u1 = (float)(o[2 * pr] >> 8) * 5.9604644775390625e-8f + 2.98023223876953125e-8f;
        const float u2 = (float)(o[2 * pr + 1] >> 8) * 5.9604644775390625e-8f;
        const float rad = cr_sqrt(-2.0f * nr_log(u1));
        float sn, cs;
        nr_sincos_2pi(u2, &sn, &cs);
        g[2 * pr] = rad * cs;
        g[2 * pr + 1] = rad * sn;
    }
}

extern __shared__ __attribute__((aligned(16))) unsigned char c_smem[];

// One workgroup per codeword: info bits (Philox, or the caller's) -> encoding program ->
// codeword bits (optional) -> BPSK + AWGN -> LLR (skipped when a.llr is null).
__global__ void __launch_bounds__(256) k_channel(ChannelArgs a) {
    const int cw = blockIdx.x;
    const int Z = a.Z, nvn = a.n_vn, K = a.K;
    const uint64_t stream = a.stream_base + (uint64_t)cw;
    uint8_t* bits = c_smem;
    for (int n = threadIdx.x; n < nvn; n += blockDim.x) bits[n] = 0;
    __syncthreads();
    if (a.info_in) {
        const uint8_t* in = a.info_in + (size_t)cw * K;
        for (int n = threadIdx.x; n < K; n += blockDim.x) bits[n] = in[n] & 1u;
    }
    // information bits: word w of the stream gives bits 32w .. 32w+31 (LSB first)
    for (int w = threadIdx.x; !a.info_in && w < (K + 31) / 32; w += blockDim.x) {
        uint32_t o[4];
        philox_block(a.seed, stream, (uint64_t)(w >> 2), kTagBits, o);
        const uint32_t word = o[w & 3];
        for (int k = 0; k < 32 && 32 * w + k < K; k++) bits[32 * w + k] = (word >> k) & 1u;
    }
    __syncthreads();
    for (int s = 0; s < a.n_steps; s++) {
        const int tgt = a.step_target[s];
        const int t0 = a.step_start[s], t1 = a.step_start[s + 1];
        for (int x = threadIdx.x; x < Z; x += blockDim.x) {
            uint32_t v = 0;
            for (int t = t0; t < t1; t++) {
                int y = x + a.term_shift[t];
                if (y >= Z) y -= Z;
                v ^= bits[a.term_col[t] * Z + y];
            }
            bits[tgt * Z + x] = (uint8_t)v;
        }
        __syncthreads();
    }
    if (a.cw_bits)
        for (int n = threadIdx.x; n < nvn; n += blockDim.x) a.cw_bits[(size_t)cw * nvn + n] = bits[n];
    if (!a.llr) return;  // encode only (no barrier follows)
    const float sigma = a.sigma_dev ? a.sigma_dev[cw] : a.sigma_all;
    const float ss = sigma * sigma;
    const int p0 = a.punct * Z;
    float* llr = a.llr + (size_t)cw * nvn;
    for (int n = threadIdx.x; n < p0; n += blockDim.x) llr[n] = 0.0f;
    // transmitted symbol i = n - p0 uses noise sample i
    const int ntx = nvn - p0;
    for (int blk = threadIdx.x; blk < (ntx + 3) / 4; blk += blockDim.x) {
        float g[4];
        gauss4(a.seed, stream, (uint64_t)blk, g);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = 4 * blk + q;
            if (i < ntx) {
                const float x = bits[p0 + i] ? -1.0f : 1.0f;
                const float y = x + g[q] * sigma;
                llr[p0 + i] = (y * 2.0f) / ss;
            }
        }
    }
}

// Syndrome weight per codeword from packed hard decisions.  The codeword's decisions and the
// code's edge list (block column | V mod Z << 16, row-major) are staged in LDS.
//   Z >= 32: 32 parity checks per operation.  Block column j is laid out twice over, as the
//   bit string D_j[b] = HD[j*Z + (b mod Z)], b < 2Z; checks 32w..32w+31 of block row i are
//   then the XOR over the row's edges of the 32 bits of D_j starting at bit 32w + V (one
//   funnel shift each), and the weight is the popcount of the valid bits.
//   Z < 32: one check at a time.
__global__ void __launch_bounds__(256) k_syndrome(SyndromeArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint32_t smw[];
    const int cw = blockIdx.x;
    const int Z = a.Z, rows = a.rows, cols = a.cols;
    const int n_e = a.n_edges;
    int* rs = reinterpret_cast<int*>(smw);       // [rows + 1]
    uint32_t* et = smw + (rows + 1);              // [n_e]
    uint32_t* W = et + n_e;                       // stream words, zero padded: [nw + 2]
    const int nw = (a.hd_bytes + 3) / 4;
    const int WD = (2 * Z + 31) / 32 + 2;         // words of one doubled block column
    uint32_t* D = W + nw + 2;                     // [cols][WD] (Z >= 32)
    for (int i = threadIdx.x; i <= rows; i += blockDim.x) rs[i] = a.row_start[i];
    for (int e = threadIdx.x; e < n_e; e += blockDim.x)
        et[e] = (uint32_t)a.edge_col[e] | ((uint32_t)a.edge_vm[e] << 16);
    const uint8_t* hd = a.hd + (size_t)cw * a.hd_bytes;
    for (int k = threadIdx.x; k < nw + 2; k += blockDim.x) {
        uint32_t v = 0;
        for (int q = 0; q < 4; q++) {
            const int b = 4 * k + q;
            if (b < a.hd_bytes) v |= (uint32_t)hd[b] << (8 * q);
        }
        W[k] = v;
    }
    __syncthreads();
    auto ext = [](const uint32_t* w, int p) {  // 32 bits of w starting at bit p
        return __builtin_amdgcn_alignbit(w[(p >> 5) + 1], w[p >> 5], (uint32_t)(p & 31));
    };
    int wsum = 0;
    if (Z >= 32) {
        for (int idx = threadIdx.x; idx < cols * WD; idx += blockDim.x) {
            const int j = idx / WD, w = idx - j * WD;
            const int b0 = 32 * w;
            uint32_t v = 0;
            if (b0 < 2 * Z) {
                const int p = b0 >= Z ? b0 - Z : b0;
                const int n1 = Z - p;  // bits before the column wraps
                v = ext(W, j * Z + p);
                if (n1 < 32) v = (v & ((1u << n1) - 1u)) | (ext(W, j * Z) << n1);
            }
            D[idx] = v;
        }
        __syncthreads();
        const int nwr = (Z + 31) / 32;
        for (int item = threadIdx.x; item < rows * nwr; item += blockDim.x) {
            const int i = item / nwr, w = item - i * nwr;
            uint32_t par = 0;
            for (int e = rs[i]; e < rs[i + 1]; e++) {
                const uint32_t t = et[e];
                par ^= ext(D + (int)(t & 0xFFFFu) * WD, 32 * w + (int)(t >> 16));
            }
            const int valid = Z - 32 * w;
            if (valid < 32) par &= (1u << valid) - 1u;
            wsum += __popc(par);
        }
    } else {
        for (int chk = threadIdx.x; chk < rows * Z; chk += blockDim.x) {
            const int i = chk / Z, r = chk - i * Z;
            uint32_t p = 0;
            for (int e = rs[i]; e < rs[i + 1]; e++) {
                const uint32_t t = et[e];
                int c = r + (int)(t >> 16);
                if (c >= Z) c -= Z;
                const int n = (int)(t & 0xFFFFu) * Z + c;
                p ^= W[n >> 5] >> (n & 31);
            }
            wsum += (int)(p & 1u);
        }
    }
    __syncthreads();
    int* red = reinterpret_cast<int*>(smw);
    red[threadIdx.x] = wsum;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.weight[cw] = red[0];
}

// Indices of the codewords with a nonzero syndrome weight, in ascending order, and their
// count (one workgroup; block-wide prefix sums over chunks of 1024).
__global__ void __launch_bounds__(1024) k_compact_failures(const int* __restrict__ weight, int n_cw,
                                                            int* __restrict__ list, int* __restrict__ count) {
    __shared__ int warp_sums[16];
    __shared__ int base_s;
    if (threadIdx.x == 0) base_s = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c0 = 0; c0 < n_cw; c0 += 1024) {
        const int i = c0 + (int)threadIdx.x;
        const bool f = i < n_cw && weight[i] != 0;
        const uint64_t m = __ballot(f);
        const int below = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) warp_sums[wv] = __popcll(m);
        __syncthreads();
        int off = base_s;
        for (int k = 0; k < wv; k++) off += warp_sums[k];
        if (f) list[off + below] = i;
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int k = 0; k < 16; k++) t += warp_sums[k];
            base_s += t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = base_s;
}

// Bit and frame errors over the first K columns.
__global__ void __launch_bounds__(256) k_count_errors(ErrorArgs a) {
    const int cw = blockIdx.x;
    const uint8_t* hd = a.hd + (size_t)cw * a.hd_bytes;
    const uint8_t* ref = a.cw_bits + (size_t)cw * a.n_vn;
    int e = 0;
    for (int n = threadIdx.x; n < a.K; n += blockDim.x) e += (int)(((hd[n >> 3] >> (n & 7)) & 1u) != ref[n]);
    __shared__ int red[256];
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] > 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.counts[0]), (unsigned long long)red[0]);
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.counts[1]), 1ull);
    }
}

// One thread per output byte: bit k of byte b is column g % n_vn of codeword g / n_vn, g = 8b + k.
__global__ void __launch_bounds__(256) k_stream_pack(StreamPackArgs a) {
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (8 * b >= a.n_bits) return;
    uint32_t v = 0;
    for (int k = 0; k < 8; k++) {
        const long long g = 8 * b + k;
        if (g >= a.n_bits) break;
        const long long cw = g / a.n_vn;
        const int c = (int)(g - cw * a.n_vn);
        v |= ((a.hd[cw * a.hd_bytes + (c >> 3)] >> (c & 7)) & 1u) << k;
    }
    a.out[b] = (uint8_t)v;
}

int stream_pack_launch(const StreamPackArgs& a, hipStream_t s) {
    const long long bytes = (a.n_bits + 7) / 8;
    const long long blocks = (bytes + 255) / 256;
    if (blocks > 0x7fffffffLL) {
        set_last_error("stream pack: batch too large");
        return LDPC_EINVAL;
    }
    hipLaunchKernelGGL(k_stream_pack, dim3((unsigned)blocks), dim3(256), 0, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

__global__ void __launch_bounds__(256) k_to_half(const float* in, size_t n, uint16_t* out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = __builtin_bit_cast(uint16_t, (_Float16)in[i]);
}

__global__ void __launch_bounds__(256) k_from_int8(const int8_t* in, size_t n, float scale, float* out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i] * scale;
}

int from_int8_launch(const int8_t* in, size_t n, float scale, float* out, hipStream_t s) {
    if (n == 0) return LDPC_OK;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_from_int8, dim3((unsigned)blocks), dim3(256), 0, s, in, n, scale, out);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

int to_half_launch(const float* in, size_t n, uint16_t* out, hipStream_t s) {
    if (n == 0) return LDPC_OK;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_to_half, dim3((unsigned)blocks), dim3(256), 0, s, in, n, out);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

int channel_launch(const ChannelArgs& a, int n_cw, hipStream_t s) {
    const size_t lds = ((size_t)a.n_vn + 15) & ~(size_t)15;
    hipLaunchKernelGGL(k_channel, dim3(n_cw), dim3(256), lds, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

int syndrome_launch(const SyndromeArgs& a, int n_cw, hipStream_t s) {
    const size_t nw = (size_t)(a.hd_bytes + 3) / 4 + 2;
    const size_t wd = (size_t)(2 * a.Z + 31) / 32 + 2;
    const size_t words = (size_t)(a.rows + 1) + (size_t)a.n_edges + nw + (a.Z >= 32 ? (size_t)a.cols * wd : 0);
    const size_t lds = std::max<size_t>(256 * sizeof(int), words * sizeof(uint32_t));
    hipLaunchKernelGGL(k_syndrome, dim3(n_cw), dim3(256), lds, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

int compact_failures_launch(const int* weight, int n_cw, int* list, int* count, hipStream_t s) {
    hipLaunchKernelGGL(k_compact_failures, dim3(1), dim3(1024), 0, s, weight, n_cw, list, count);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

int count_errors_launch(const ErrorArgs& a, int n_cw, hipStream_t s) {
    hipLaunchKernelGGL(k_count_errors, dim3(n_cw), dim3(256), 0, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

}  // namespace nrldpc
