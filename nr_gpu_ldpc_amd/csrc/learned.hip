// learned.hip -- training kernels of the learned offset/normalized min-sum decoder.
//
// Decoder (forward, k_decode<..., LEARNED>):  R_t(e) = s * alpha_t(e) * max(m - beta_t(e), 0)
// with one (alpha, beta) per (iteration t, base-graph edge e); alpha = 0.75, beta = 0 is the
// plain decoder.  Loss: mean binary cross-entropy of the soft output sigmoid(-APP_T) against
// the transmitted bits, over every decoder column, each codeword's terms optionally weighted.
//
// Backward (this file), in exact-arithmetic form APP_t = L + sum_e R_t(e),
// Q_t(e) = APP_{t-1}[c(e)] - R_{t-1}(e):
//   gR_t(e)   = gAPP_t[c(e)] - gQ_{t+1}(e)
//   dalpha    += gR * s * relu(m - beta),   dbeta -= gR * s * alpha * [m > beta]
//   gQ_t(src) += gR * s * alpha * [m > beta] * sign(Q_t(src)),  src = arg-min excluding e
//   gAPP_{t-1}[c] = sum_e gQ_t(e)
// gQ_t is nonzero only at the first and second arg-min edge of every expanded row, so the
// carried state per row is two values and two indices.  One workgroup walks its codewords
// one after another; weight gradients are reduced across the wavefront, accumulated in LDS
// and added to the global gradient once per workgroup.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "code.h"
#include "learned.h"
#include "status.h"

namespace nrldpc {

template <int B, int E, class F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int KIND>
__global__ void __launch_bounds__(512) k_noms_backward(BackwardArgs a) {
    using S = Structure<KIND>;
    constexpr int ROWS = S::ROWS;
    extern __shared__ float gacc[];  // [n_iter][nnz][2] weight-gradient accumulators
    const int Z = a.Z, r = threadIdx.x, nvn = a.n_vn, T = a.n_iter;
    const bool act = r < Z;
    const int nacc = T * a.nnz * 2;
    for (int n = threadIdx.x; n < nacc; n += blockDim.x) gacc[n] = 0.f;
    float* gcur = a.gapp + (size_t)blockIdx.x * 2 * nvn;
    float* gnxt = gcur + nvn;
    float loss_part = 0.f;
    __syncthreads();

    for (int cw = blockIdx.x; cw < a.n_cw; cw += gridDim.x) {
        const float* app = a.app + (size_t)cw * nvn;
        const uint8_t* bits = a.bits + (size_t)cw * nvn;
        const float wcw = a.cw_weight ? a.cw_weight[cw] : 1.f;  // the codeword's loss weight
        const float gnorm = wcw * a.inv_norm;
        // loss and dLoss/dAPP_T
        for (int n = r; n < nvn; n += blockDim.x) {
            const float x = app[n];
            const float sp = x > 0.f ? x + log1pf(expf(-x)) : log1pf(expf(x));  // softplus(x)
            const float sg = 1.f / (1.f + expf(-x));
            loss_part += wcw * (bits[n] ? sp : sp - x);
            gcur[n] = (bits[n] ? sg : sg - 1.f) * gnorm;
        }
        float gq1[ROWS], gq2[ROWS];
        uint32_t pa1[ROWS], pa2[ROWS];
#pragma unroll
        for (int i = 0; i < ROWS; i++) { gq1[i] = 0.f; gq2[i] = 0.f; pa1[i] = 31u; pa2[i] = 31u; }
        __syncthreads();
        for (int t = T - 1; t >= 0; --t) {
            const float2* __restrict__ w = a.weights + (size_t)(t + 1) * a.nnz;
            float* gt = gacc + (size_t)t * a.nnz * 2;
            // phase 1: check-node backward of iteration t for every row
            sfor<0, ROWS>([&](auto IC) {
                constexpr int i = IC;
                constexpr int D = S::deg(i);
                constexpr int E0 = S::T.row_start[i];
                if (i >= a.m_rows) return;
                const size_t ri = (((size_t)cw * T + t) * ROWS + i) * (size_t)Z + (act ? r : 0);
                const float m1 = a.rec_m1[ri], m2 = a.rec_m2[ri];
                const uint32_t rw = a.rec_w[ri];
                const uint32_t neg = rw & 0xFFFFFu, a1 = (rw >> 20) & 31u, a2 = (rw >> 25) & 31u;
                const uint32_t sgn = __popc(neg) & 1u;
                const float gn1 = gq1[i], gn2 = gq2[i];
                const uint32_t p1 = pa1[i], p2 = pa2[i];
                float g1 = 0.f, g2 = 0.f;
                sfor<0, D>([&](auto KC) {
                    constexpr int k = KC;
                    constexpr int e = E0 + k;
                    constexpr int j = S::T.ecol[e];
                    int c = r + a.edge_vm[e];
                    if (c >= Z) c -= Z;
                    float gR = act ? gcur[j * Z + c] : 0.f;
                    gR -= (p1 == (uint32_t)k) ? gn1 : 0.f;
                    gR -= (p2 == (uint32_t)k) ? gn2 : 0.f;
                    const uint32_t nk = (neg >> (D - 1 - k)) & 1u;
                    const float s = (sgn ^ nk) ? -1.f : 1.f;
                    const bool isa1 = a1 == (uint32_t)k;
                    const float m = isa1 ? m2 : m1;
                    const float2 wt = w[e];
                    const float off = m - wt.y;
                    const bool pos = off > 0.f;
                    const float gs = gR * s;
                    const float da = pos ? gs * off : 0.f;
                    const float db = pos ? -gs * wt.x : 0.f;
                    const float gm = pos ? gs * wt.x : 0.f;
                    if (isa1) g2 += gm;
                    else g1 += gm;
                    const float sa = wave_sum(da), sb = wave_sum(db);
                    if ((threadIdx.x & 63) == 0) {
                        atomicAdd(&gt[2 * e], sa);
                        atomicAdd(&gt[2 * e + 1], sb);
                    }
                });
                // sign(Q) at the arg-mins; no gradient through the 1000.0f start value
                g1 = (a1 < 31u) ? (((neg >> (D - 1 - a1)) & 1u) ? -g1 : g1) : 0.f;
                g2 = (a2 < 31u) ? (((neg >> (D - 1 - a2)) & 1u) ? -g2 : g2) : 0.f;
                gq1[i] = g1;
                gq2[i] = g2;
                pa1[i] = a1;
                pa2[i] = a2;
            });
            __syncthreads();
            if (t == 0) break;
            // phase 2: gAPP_{t-1}[c] = sum of gQ_t over the column's edges
            for (int n = r; n < nvn; n += blockDim.x) gnxt[n] = 0.f;
            __syncthreads();
            if (act) {
                sfor<0, ROWS>([&](auto IC) {
                    constexpr int i = IC;
                    constexpr int E0 = S::T.row_start[i];
                    if (i >= a.m_rows) return;
                    if (pa1[i] < 31u) {
                        const int e = E0 + (int)pa1[i];
                        int c = r + a.edge_vm[e];
                        if (c >= Z) c -= Z;
                        atomicAdd(&gnxt[a.edge_col[e] * Z + c], gq1[i]);
                    }
                    if (pa2[i] < 31u) {
                        const int e = E0 + (int)pa2[i];
                        int c = r + a.edge_vm[e];
                        if (c >= Z) c -= Z;
                        atomicAdd(&gnxt[a.edge_col[e] * Z + c], gq2[i]);
                    }
                });
            }
            __syncthreads();
            float* tmp = gcur;
            gcur = gnxt;
            gnxt = tmp;
        }
        __syncthreads();
        // restore the buffer order for the next codeword
        if (gcur != a.gapp + (size_t)blockIdx.x * 2 * nvn) {
            gnxt = gcur;
            gcur = a.gapp + (size_t)blockIdx.x * 2 * nvn;
        }
    }
    __syncthreads();
    for (int n = threadIdx.x; n < nacc; n += blockDim.x)
        if (gacc[n] != 0.f) atomicAdd(&a.grad[n], gacc[n]);
    const float ls = wave_sum(loss_part);  // every lane (r >= Z too) summed loss terms
    if ((threadIdx.x & 63) == 0) atomicAdd(a.loss, ls * a.inv_norm);
}

__global__ void k_adam(AdamArgs a) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.n) return;
    const float g = a.grad[n];
    const float m = a.b1 * a.m[n] + (1.f - a.b1) * g;
    const float v = a.b2 * a.v[n] + (1.f - a.b2) * g * g;
    a.m[n] = m;
    a.v[n] = v;
    const float mh = m / (1.f - a.b1t), vh = v / (1.f - a.b2t);
    a.w[n] -= a.lr * mh / (sqrtf(vh) + a.eps);
}

// One workgroup per codeword; rows of codewords that keep their learned result are untouched.
__global__ void __launch_bounds__(256) k_fallback_merge(MergeArgs a) {
    const int cw = blockIdx.x;
    if (a.w_learned[cw] == 0 || a.w_plain[cw] != 0) return;
    const size_t hb = (size_t)cw * a.hd_bytes;
    for (int n = threadIdx.x; n < a.hd_bytes; n += blockDim.x) a.hd[hb + n] = a.hd_plain[hb + n];
    if (a.app) {
        const size_t ab = (size_t)cw * a.n_vn;
        for (int n = threadIdx.x; n < a.n_vn; n += blockDim.x) a.app[ab + n] = a.app_plain[ab + n];
    }
    if (a.iters && threadIdx.x == 0) a.iters[cw] = a.iters_plain[cw];
}

int fallback_merge_launch(const MergeArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_fallback_merge, dim3(a.n_cw), dim3(256), 0, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

int backward_launch(int kind, const BackwardArgs& a, int n_wg, hipStream_t s) {
    const int threads = ((a.Z + 63) / 64) * 64;
    const size_t lds = (size_t)a.n_iter * a.nnz * 2 * sizeof(float);
    if (threads > 512 || lds > 64 * 1024) {
        set_last_error("backward launch exceeds the workgroup size or LDS limit");
        return LDPC_EINVAL;
    }
    if (kind == kBG1) hipLaunchKernelGGL(k_noms_backward<kBG1>, dim3(n_wg), dim3(threads), lds, s, a);
    else if (kind == kBG2) hipLaunchKernelGGL(k_noms_backward<kBG2>, dim3(n_wg), dim3(threads), lds, s, a);
    else hipLaunchKernelGGL(k_noms_backward<kWifi>, dim3(n_wg), dim3(threads), lds, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

int adam_launch(const AdamArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

}  // namespace nrldpc
