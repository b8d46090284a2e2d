// decoder.hip -- fused flooding (learned) scaled min-sum decoder for gfx950.
//
// One workgroup decodes `cpb` codewords; lane (cw, r) owns expanded row r of every block row
// and sub-column r of every block column of codeword cw.  All iterations run inside one
// launch with the whole decoder state on chip:
//
//   LDS, per codeword   APP of the block columns of degree > 1, twice:
//                       * read copy  [NND][2Z]  (sub-columns 0..Z-1 stored twice so the
//                         circulant read at (r + V) never wraps: one VALU add per edge)
//                       * accumulator [NND][Z]  (APP_new, updated in place)
//   VGPRs, per lane     check-node state of every block row: the two scaled minima
//                       (|R| values), the sign bits of R and the arg-min index
//                       (12 bytes/row instead of the reference's R per edge), and the APP
//                       of every degree-1 (extension parity) column the lane's row feeds.
//
// Arithmetic follows the reference kernels exactly (cnp_kernel_1st_iter / cnp_kernel /
// vnp_kernel_normal / vnp_kernel_last_iter, src/ldpc_kernels.cu:72-468):
//   Q = APP_old - R_old;  two-min with the 1000.0f start value;  R = +-fl(0.75 m);
//   dt = fl(R_new - R_old);  APP_new = fl(...fl(APP_old + dt_i0) + dt_i1 ...) over incident
//   block rows in ascending order;  HD = APP > 0 ? 0 : 1.
// The arg-min index is replaced by the equality test |Q_k| == m1, which selects the same R
// (a tie at the minimum forces m2 == m1), and (m1, m2) are updated with min / med3, which
// reproduce the `<` / `else if <` rule exactly.  Per-column update order is kept by
// processing block rows in ascending order and placing a workgroup barrier before the VN
// update of every block row that shares an LDS column with the rows since the previous
// barrier (compile-time "epochs").
#include "code.h"
#include "decoder.h"
#include "status.h"

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#ifndef NRLDPC_LDS_ATOMIC_ACC
#define NRLDPC_LDS_ATOMIC_ACC 0
#endif

namespace nrldpc {

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

__device__ __forceinline__ float as_f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t as_u(float f) { return __float_as_uint(f); }

extern __shared__ __attribute__((aligned(16))) unsigned char g_smem[];

__device__ __forceinline__ float lds_ld(uint32_t a) { return *reinterpret_cast<float*>(g_smem + a); }
__device__ __forceinline__ void lds_st(uint32_t a, float v) { *reinterpret_cast<float*>(g_smem + a) = v; }
__device__ __forceinline__ uint32_t lds_ldu(uint32_t a) { return *reinterpret_cast<uint32_t*>(g_smem + a); }
__device__ __forceinline__ void lds_stu(uint32_t a, uint32_t v) { *reinterpret_cast<uint32_t*>(g_smem + a) = v; }

// VN update of one LDS-resident column: APP_new += dt.  Within one epoch every column is
// updated by at most one lane, so a plain read-modify-write is race free.
__device__ __forceinline__ void lds_acc(uint32_t a, float dt) {
    float* p = reinterpret_cast<float*>(g_smem + a);
    *p = *p + dt;
}

// Per-edge table entries are read through the constant address space so that the compiler
// issues scalar (s_load) reads of these wave-uniform values.
using CEdge = const EdgeParam __attribute__((address_space(4)));

template <int KIND>
constexpr bool nd_columns_leading() {
    using S = Structure<KIND>;
    for (int j = 0; j < S::NND; j++)
        if (S::T.nd_index[j] != j) return false;
    return true;
}

template <int KIND, bool LEARNED>
__global__ void __launch_bounds__(kMaxBlock) k_decode(DecodeArgs a) {
    using S = Structure<KIND>;
    static_assert(nd_columns_leading<KIND>(), "LDS-resident columns must be the leading ones");
    constexpr int ROWS = S::ROWS;
    constexpr int NND = S::NND;
    constexpr int ND1 = S::ND1 > 0 ? S::ND1 : 1;

    const int Z = a.Z;
    const int tid = threadIdx.x;
    const int cwl = tid / Z;
    const int r = tid - cwl * Z;
    const int cw = blockIdx.x * a.cpb + cwl;
    const bool live = cw < a.n_cw;
    const int cwc = live ? cw : a.n_cw - 1;
    const uint32_t Z4 = 4u * Z;
    const uint32_t base = cwl * a.lds_cw_bytes;
    uint32_t rRd = base + 4u * r;                  // read copy, block column 0
    uint32_t rAcc = base + 2u * NND * Z4 + 4u * r; // accumulator, block column 0
    uint32_t rAccW = rAcc - Z4;                    // same, for lanes whose (r+V) wraps
    const uint32_t rHi = ((uint32_t)r << 16) | 0xFFFFu;  // compared with EdgeParam::acc
    const uint32_t flagAddr = base + 3u * NND * Z4;
    CEdge* et = (CEdge*)(a.etab);
    const int m_rows = a.m_rows;

    constexpr int NW = S::NW;
    float A1[ROWS], A2[ROWS];  // plain: fl(0.75 m1), fl(0.75 m2); learned: m1, m2
    uint32_t W[NW];            // per row: sign of R per edge + arg-min index (see Structure)
    float app1[ND1];           // APP of the degree-1 columns fed by this lane's rows
#pragma unroll
    for (int i = 0; i < ROWS; i++) { A1[i] = 0.f; A2[i] = 0.f; }
#pragma unroll
    for (int i = 0; i < NW; i++) W[i] = 0u;

    // ---- load channel LLRs
    const float* __restrict__ L = a.llr + (size_t)cwc * a.n_vn;
#pragma unroll 2
    for (int j = 0; j < NND; j++) {
        const float v = L[j * Z + r];
        lds_st(rRd + j * 2 * Z4, v);
        lds_st(rRd + j * 2 * Z4 + Z4, v);
        lds_st(rAcc + j * Z4, v);
    }
    static_for<NND, S::COLS>([&](auto JC) {
        constexpr int j = JC;
        constexpr int e = S::T.d1_edge[j];
        constexpr int i = S::T.d1_row[j];
        if (i < m_rows) {
            int c = r + (int)et[e].off_rd;
            if (c >= Z) c -= Z;
            app1[S::T.d1_slot[j]] = L[j * Z + c];
        }
    });
    __syncthreads();

    bool emitted = false;
    const bool early = a.early_exit != 0;
    const float2* __restrict__ wt = a.weights;  // learned: [(n_iter+1)][nnz] (alpha, beta), row 0 zero

    // Writes hard decisions (and APP) of every lane's codeword for which `mine` holds.
    auto emit = [&](bool mine, int iters) {
        if (mine) {
            // Loop-variant copies: no address of this path is precomputed outside it.
            uint32_t eb = base;
            int er = r;
            float* ap_out = a.app_out;
            CEdge* ee = et;
            asm volatile("" : "+v"(eb), "+v"(er), "+s"(ap_out), "+s"(ee));
            if (ap_out) ap_out += (size_t)cw * a.n_vn;
#pragma unroll 1
            for (int j = 0; j < NND; j++) {
                const float x = lds_ld(eb + 2u * NND * Z4 + 4u * er + j * Z4);
                g_smem[eb + j * Z + er] = x > 0.f ? 0 : 1;
                if (ap_out) ap_out[j * Z + er] = x;
            }
            static_for<NND, S::COLS>([&](auto JC) {
                constexpr int j = JC;
                constexpr int e = S::T.d1_edge[j];
                constexpr int i = S::T.d1_row[j];
                if (i < m_rows) {
                    int c = er + (int)ee[e].off_rd;
                    if (c >= Z) c -= Z;
                    const float x = app1[S::T.d1_slot[j]];
                    g_smem[eb + j * Z + c] = x > 0.f ? 0 : 1;
                    if (ap_out) ap_out[j * Z + c] = x;
                }
            });
            if (er < 8) g_smem[eb + a.n_vn + er] = 0;
        }
        __syncthreads();
        if (mine) {
            for (int b = r; b < a.hd_bytes; b += Z) {
                const uint32_t v0 = lds_ldu(base + 8u * b), v1 = lds_ldu(base + 8u * b + 4u);
                const uint32_t lo = (v0 | (v0 >> 7) | (v0 >> 14) | (v0 >> 21)) & 0xFu;
                const uint32_t hi = (v1 | (v1 >> 7) | (v1 >> 14) | (v1 >> 21)) & 0xFu;
                a.hd[(size_t)cw * a.hd_bytes + b] = (uint8_t)(lo | (hi << 4));
            }
            if (r == 0 && a.iters_out) a.iters_out[cw] = iters;
        }
        __syncthreads();
    };

    int t = 0;
    for (; t < a.n_iter; ++t) {
        // Loop-variant copies of the lane bases keep address arithmetic from being hoisted
        // out of the iteration loop into long-lived VGPRs.
        asm volatile("" : "+v"(rRd), "+v"(rAcc), "+v"(rAccW));
        const float2* __restrict__ w_old = LEARNED ? wt + (size_t)t * a.nnz : nullptr;
        const float2* __restrict__ w_new = LEARNED ? wt + (size_t)(t + 1) * a.nnz : nullptr;
        static_for<0, ROWS>([&](auto IC) {
            constexpr int i = IC;
            constexpr int D = S::deg(i);
            constexpr int E0 = S::T.row_start[i];
            constexpr int WW = S::T.w_word[i];
            constexpr int WO = S::T.w_off[i];                 // sign bits start here
            constexpr int IO = S::half(i) ? WO + 11 : 24;     // arg-min index starts here
            int mr = m_rows;
            asm volatile("" : "+s"(mr));
            if (i >= mr) return;
            // Keep this row's table reads inside the row (bounded SGPR live ranges).
            asm volatile("" : "+s"(et));
            float q[D], ro[D];
            float m1 = 1000.0f, m2 = 1000.0f;
            uint32_t neg = 0u;
            const float a1 = A1[i], a2 = A2[i];
            const uint32_t wo = W[WW];
            uint32_t idxo = (wo >> IO) & 31u;
            asm volatile("" : "+v"(idxo));
            // pass 1: Q = APP_old - R_old, two-min, signs
            static_for<0, D>([&](auto KC) {
                constexpr int k = KC;
                constexpr int e = E0 + k;
                constexpr int j = S::T.ecol[e];
                float x;
                if constexpr (j < NND) x = lds_ld(rRd + et[e].off_rd);
                else x = app1[S::T.d1_slot[j]];
                const float mo = (idxo == (uint32_t)k) ? a2 : a1;
                float rv;
                if constexpr (LEARNED) {
                    const float2 w = w_old[e];
                    const float mag = w.x * fmaxf(mo - w.y, 0.0f);
                    rv = as_f(as_u(mag) ^ ((wo << (31 - WO - k)) & 0x80000000u));
                } else {
                    rv = as_f((as_u(mo) & 0x7fffffffu) | ((wo << (31 - WO - k)) & 0x80000000u));
                }
                const float qv = x - rv;
                q[k] = qv;
                ro[k] = rv;
                const float aq = fabsf(qv);
                m2 = __builtin_amdgcn_fmed3f(m1, m2, aq);
                m1 = fminf(m1, aq);
                neg |= (qv < 0.0f ? 1u : 0u) << k;
            });
            const uint32_t sgn = __popc(neg) & 1u;
            const uint32_t sbit = sgn << 31;
            float n1, n2;
            if constexpr (LEARNED) { n1 = m1; n2 = m2; }
            else { n1 = 0.75f * m1; n2 = 0.75f * m2; }
            const float n1s = as_f(as_u(n1) ^ sbit), n2s = as_f(as_u(n2) ^ sbit);
            // pass 2: R_new = (product of the other signs) * scaled min, dt = R_new - R_old
            uint32_t idxn = 0u;
            float dtv[D];
            static_for<0, D>([&](auto KC) {
                constexpr int k = KC;
                constexpr int e = E0 + k;
                const bool eq = fabsf(q[k]) == m1;
                idxn = eq ? (uint32_t)k : idxn;
                float ms;
                if constexpr (LEARNED) {
                    const float2 w = w_new[e];
                    const float mag = w.x * fmaxf((eq ? n2 : n1) - w.y, 0.0f);
                    ms = as_f(as_u(mag) ^ sbit);
                } else {
                    ms = eq ? n2s : n1s;
                }
                const float rn = (q[k] < 0.0f) ? -ms : ms;
                dtv[k] = rn - ro[k];
            });
            A1[i] = n1;
            A2[i] = n2;
            constexpr uint32_t dmask = (1u << D) - 1u;
            const uint32_t fresh_w = ((sgn ? ~neg : neg) & dmask) | (idxn << (IO - WO));
            if constexpr (!S::half(i)) W[WW] = fresh_w;
            else if constexpr (WO == 0) W[WW] = (W[WW] & 0xFFFF0000u) | fresh_w;
            else W[WW] = (W[WW] & 0x0000FFFFu) | (fresh_w << 16);
            if constexpr (S::T.epoch_start[i] != 0 && i != 0) __syncthreads();
            // VN update, ascending block row per column
#if NRLDPC_LDS_ATOMIC_ACC
            static_for<0, D>([&](auto KC) {
                constexpr int k = KC;
                constexpr int e = E0 + k;
                constexpr int j = S::T.ecol[e];
                if constexpr (j < NND) {
                    const uint32_t acc = et[e].acc;
                    const uint32_t addr = (rHi >= acc ? rAccW : rAcc) + (acc & 0xFFFFu);
                    __hip_atomic_fetch_add(reinterpret_cast<float*>(g_smem + addr), dtv[k],
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    float& ap = app1[S::T.d1_slot[j]];
                    ap = ap + dtv[k];
                }
            });
#else
            // read-modify-write in chunks of 8 edges: loads first, then stores
            static_for<0, (D + 7) / 8>([&](auto CC) {
                constexpr int C0 = CC * 8;
                constexpr int CN = (D - C0) < 8 ? (D - C0) : 8;
                uint32_t addr[CN];
                float cur[CN];
                static_for<0, CN>([&](auto KC) {
                    constexpr int k = C0 + KC;
                    constexpr int e = E0 + k;
                    constexpr int j = S::T.ecol[e];
                    if constexpr (j < NND) {
                        const uint32_t acc = et[e].acc;
                        addr[KC] = (rHi >= acc ? rAccW : rAcc) + (acc & 0xFFFFu);
                        cur[KC] = lds_ld(addr[KC]);
                    }
                });
                static_for<0, CN>([&](auto KC) {
                    constexpr int k = C0 + KC;
                    constexpr int e = E0 + k;
                    constexpr int j = S::T.ecol[e];
                    if constexpr (j < NND) {
                        lds_st(addr[KC], cur[KC] + dtv[k]);
                    } else {
                        float& ap = app1[S::T.d1_slot[j]];
                        ap = ap + dtv[k];
                    }
                });
            });
#endif
        });
        __syncthreads();
        if (t + 1 == a.n_iter) break;
        // APP_new becomes the read copy of the next iteration.
#pragma unroll 2
        for (int j = 0; j < NND; j++) {
            const float v = lds_ld(rAcc + j * Z4);
            lds_st(rRd + j * 2 * Z4, v);
            lds_st(rRd + j * 2 * Z4 + Z4, v);
        }
        if (early && r == 0) lds_stu(flagAddr, 0u);
        __syncthreads();
        if (early) {
            // syndrome of the current hard decision; any unsatisfied check keeps decoding
            uint32_t bad = 0u;
            static_for<0, ROWS>([&](auto IC) {
                constexpr int i = IC;
                constexpr int D = S::deg(i);
                constexpr int E0 = S::T.row_start[i];
                int mr = m_rows;
                asm volatile("" : "+s"(mr));
                if (i >= mr) return;
                asm volatile("" : "+s"(et));
                uint32_t p = 0u;
                static_for<0, D>([&](auto KC) {
                    constexpr int e = E0 + KC;
                    constexpr int j = S::T.ecol[e];
                    float x;
                    if constexpr (j < NND) x = lds_ld(rRd + et[e].off_rd);
                    else x = app1[S::T.d1_slot[j]];
                    p ^= x > 0.f ? 0u : 1u;
                });
                bad |= p;
            });
            if (bad) lds_stu(flagAddr, 1u);
            __syncthreads();
            const bool conv = lds_ldu(flagAddr) == 0u;
            const bool fresh = live && conv && !emitted;
            if (__syncthreads_or(fresh ? 1 : 0)) {
                // staging reuses the read copy of the converged codewords only
                emit(fresh, t + 1);
                if (fresh) emitted = true;
            }
            if (__syncthreads_and((emitted || !live) ? 1 : 0)) break;
        }
    }
    if (__syncthreads_or((live && !emitted) ? 1 : 0))
        emit(live && !emitted, t < a.n_iter ? t + 1 : a.n_iter);
}

// ---------------------------------------------------------------- host side

template <int KIND, bool LEARNED>
static int launch_one(const DecodeArgs& a, int blocks, int threads, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_decode<KIND, LEARNED>), dim3(blocks), dim3(threads), lds, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

// Dynamic LDS available to one workgroup on the current device (queried once per process).
int max_lds_bytes() {
    static int v = -1;
    if (v < 0) {
        int dev = 0, bytes = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            bytes = 64 * 1024;
        v = bytes - 1024;  // headroom for the kernels' static LDS
    }
    return v;
}

size_t decoder_lds_bytes_per_cw(int kind, int Z) {
    int nnd = 0;
    if (kind == kBG1) nnd = Structure<kBG1>::NND;
    else if (kind == kBG2) nnd = Structure<kBG2>::NND;
    else nnd = Structure<kWifi>::NND;
    return ((size_t)3 * nnd * 4 * Z + 16 + 15) & ~(size_t)15;
}

int decoder_launch(int kind, bool learned, const DecodeArgs& a, hipStream_t s) {
    const int blocks = (a.n_cw + a.cpb - 1) / a.cpb;
    const int threads = a.cpb * a.Z;
    const size_t lds = (size_t)a.cpb * a.lds_cw_bytes;
    if (threads > kMaxBlock || lds > (size_t)max_lds_bytes() || blocks < 1) {
        set_last_error("launch shape exceeds the workgroup size or LDS limit");
        return LDPC_EINVAL;
    }
    if (kind == kBG1) return learned ? launch_one<kBG1, true>(a, blocks, threads, lds, s)
                                     : launch_one<kBG1, false>(a, blocks, threads, lds, s);
    if (kind == kBG2) return learned ? launch_one<kBG2, true>(a, blocks, threads, lds, s)
                                     : launch_one<kBG2, false>(a, blocks, threads, lds, s);
    return learned ? launch_one<kWifi, true>(a, blocks, threads, lds, s)
                   : launch_one<kWifi, false>(a, blocks, threads, lds, s);
}

}  // namespace nrldpc
